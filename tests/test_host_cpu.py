"""Host-side pieces without a GPU: MPOGenerator, MPS initialisation and the
oracle's two-site sweep against the reference's golden energies.

Pins: (1) every generated MPO contracts to the same dense Hamiltonian as an
independent numpy Kronecker sum of the same terms; (2) exact diagonalisation
of that dense Hamiltonian reproduces the reference's golden energies; (3) the
oracle sweep, started from the product MPS initialisation, reproduces them.
"""
import os

import numpy as np
import pytest

from mps2_amd import hamiltonians as HM
from mps2_amd.flat import BTensor
from oracle import oracle_py as O


def mpo_dense(model):
    """H[out..., in...] from the MPO tensors (W_0=(p_in,b,p_out), W_i=(b_l,p_in,p_out,b_r), W_N-1=(p_in,b_l,p_out))."""
    N = model.N
    Ws = [model.mpo(i).dense() for i in range(N)]
    # running tensor R[b, out..., in...]
    W0 = Ws[0]                       # (i, b, j)
    R = np.transpose(W0, (1, 2, 0))  # (b, j, i)
    d = [model.site_dims[0]]
    for k in range(1, N):
        W = Ws[k]
        if k == N - 1:
            W = np.transpose(W, (1, 0, 2))[..., None]   # (b_l, p_in, p_out, 1)
        # R (b, J, I) ; W (b, i, j, b')
        R = np.einsum("bJI,bijc->cJjIi", R.reshape(R.shape[0], int(np.prod(d)), int(np.prod(d))), W)
        d.append(model.site_dims[k])
        R = R.reshape(R.shape[0], int(np.prod(d)), int(np.prod(d)))
    return R[0]


def kron_terms_dense(N, dims, terms, dtype=float):
    """terms: list of (coef, [(site, op_in_out_matrix)]) with identity elsewhere."""
    D = int(np.prod(dims))
    H = np.zeros((D, D), dtype=dtype)
    for coef, ops in terms:
        mats = [np.eye(d) for d in dims]
        for s, op in ops:
            mats[s] = mats[s] @ np.asarray(op).T   # M[out, in] = op[in, out]
        full = mats[0]
        for m in mats[1:]:
            full = np.kron(full, m)
        H += coef * full
    return H


def heis_terms(pairs, J=1.0):
    t = []
    for (i, j) in pairs:
        t += [(J, [(i, HM.SZ), (j, HM.SZ)]), (0.5 * J, [(i, HM.SP), (j, HM.SM)]), (0.5 * J, [(i, HM.SM), (j, HM.SP)])]
    return t


def test_heisenberg_mpo_matches_kron_sum():
    N = 6
    m = HM.heisenberg_chain(N)
    # 5 channels inside; the first bond needs no "finished" channel and the
    # last no "not yet started" one (the edge pattern of test_mpogen.cc:115-140)
    assert list(m.bond_dims) == [4] + [5] * (N - 3) + [4]
    H = kron_terms_dense(N, [2] * N, heis_terms([(i, i + 1) for i in range(N - 1)]))
    assert np.abs(mpo_dense(m) - H).max() < 1e-13


def test_j1j2_cylinder_mpo_matches_kron_sum():
    Lx, W = 3, 3
    m = HM.j1j2_cylinder(Lx, W, 1.0, 0.5)
    j1, j2 = HM.j1j2_cylinder_bonds(Lx, W)
    H = kron_terms_dense(Lx * W, [2] * (Lx * W), heis_terms(j1, 1.0) + heis_terms(j2, 0.5))
    assert np.abs(mpo_dense(m) - H).max() < 1e-12


def fermion_terms_with_strings(pairs_ops, N, f):
    """Expand (coef, opA, i, opB, j) with the insertion operator f on the sites strictly between."""
    t = []
    for coef, A, i, B, j in pairs_ops:
        ops = [(i, A)] + [(k, f) for k in range(i + 1, j)] + [(j, B)]
        t.append((coef, ops))
    return t


def test_tj_mpo_matches_kron_sum():
    o = HM.TJ
    pairs = [(0, 1), (0, 2), (2, 3), (1, 3)]
    m = HM.tj_model(4, pairs)
    spec = []
    for (i, j) in pairs:
        spec += [(-3.0, o["cdagup"], i, o["cup"], j), (-3.0, o["cdagdn"], i, o["cdn"], j),
                 (-3.0, o["cup"], i, o["cdagup"], j), (-3.0, o["cdn"], i, o["cdagdn"], j)]
    terms = fermion_terms_with_strings(spec, 4, o["f"])
    for (i, j) in pairs:
        terms += [(1.0, [(i, o["sz"]), (j, o["sz"])]), (0.5, [(i, o["sp"]), (j, o["sm"])]),
                  (0.5, [(i, o["sm"]), (j, o["sp"])])]
    H = kron_terms_dense(4, [3] * 4, terms)
    assert np.abs(mpo_dense(m) - H).max() < 1e-13


def sector_ground_energy(H, site_qns, total):
    """Lowest eigenvalue of H restricted to basis states with total QN `total`."""
    import itertools
    states = list(itertools.product(*[range(len(q)) for q in site_qns]))
    keep = [k for k, s in enumerate(states)
            if tuple(np.sum([site_qns[i][s[i]] for i in range(len(s))], axis=0)) == tuple(total)]
    Hs = H[np.ix_(keep, keep)]
    return np.linalg.eigvalsh(Hs)[0]


@pytest.mark.parametrize("name", ["heisenberg_1d_N6", "heisenberg_2d_6", "ising_1d_N6", "tj_1d_N4", "tj_2d_N4",
                                  "hubbard_2x2", "kondo_8", "kitaev_simple_8"])
def test_ed_reproduces_reference_golden_energy(name):
    case = HM.golden_cases()[name]
    m = case["build"]()
    H = mpo_dense(m)
    assert np.abs(H - H.T).max() < 1e-12
    sites = {"tj": HM.TJ_SITE, "hubbard": HM.HUB_SITE}
    if name.startswith("kondo"):
        e = np.linalg.eigvalsh(H)[0]
    else:
        key = "tj" if name.startswith("tj") else ("hubbard" if name.startswith("hubbard") else None)
        site = sites[key] if key else HM.SPIN_SITE
        qns = [[q for q, _ in site]] * m.N
        labs = case["init"]["labs"][0]
        total = tuple(np.sum([qns[i][labs[i]] for i in range(m.N)], axis=0))
        e = sector_ground_energy(H, qns, total)
    assert abs(e - case["e0"]) < max(case["tol"], 1e-11), (e, case["e0"])


def test_ed_reproduces_reference_complex_rashba_energy():
    """Target of the complex (GQTEN_Complex) path, which this build does not
    provide yet: the reference's Rashba t-J case
    (test_two_site_update_finite_vmps.cc:664-789) on a 3x2 lattice, periodic in
    x, with imaginary couplings i*lambda on the y bonds.  Its sweep keeps every
    state (D = 30 >= 27 at the centre bond), so its golden energy is the exact
    ground energy of the Hamiltonian its AddTerm calls define: operators as
    op[in, out], the insertion operator f on the sites strictly between the two
    operators of a term, N = 5 from the start state {0, 1, 0, 2, 0, 1}."""
    import itertools

    def op(entries):
        m = np.zeros((3, 3), complex)
        for (i, j), v in entries.items():
            m[i, j] = v
        return m

    f = op({(0, 0): -1, (1, 1): -1, (2, 2): 1})
    sz, sp, sm = op({(0, 0): 0.5, (1, 1): -0.5}), op({(0, 1): 1}), op({(1, 0): 1})
    cup, cdagup, cdn, cdagdn = op({(2, 0): 1}), op({(0, 2): 1}), op({(2, 1): 1}), op({(1, 2): 1})
    ntot = op({(0, 0): 1, (1, 1): 1})
    t, J, lamb = 3.0, 1.0, 0.03
    Nx, Ny = 3, 2
    N = Nx * Ny
    spec = []

    def bond(s0, s1, so):
        s0, s1 = min(s0, s1), max(s0, s1)
        spec.extend([(-t, cdagup, s0, cup, s1), (-t, cdagdn, s0, cdn, s1), (-t, cup, s0, cdagup, s1),
                     (-t, cdn, s0, cdagdn, s1)] + [(c, A, s0, B, s1) for c, A, B in so])
        for c, A, B in [(J, sz, sz), (J / 2, sp, sm), (J / 2, sm, sp), (-J / 4, ntot, ntot)]:
            terms.append((c, [(s0, A), (s1, B)]))

    terms = []
    for x in range(Nx):
        for y in range(Ny):
            s0 = x * Ny + y
            if x != Nx - 1:
                so = [(lamb, cdagup, cdn), (lamb, cup, cdagdn), (-lamb, cdagdn, cup), (-lamb, cdn, cdagup)]
            else:
                so = [(lamb, cdn, cdagup), (lamb, cdagdn, cup), (-lamb, cup, cdagdn), (-lamb, cdagup, cdn)]
            bond(s0, ((x + 1) % Nx) * Ny + y, so)
            if y != Ny - 1:
                il = 1j * lamb
                bond(s0, x * Ny + y + 1, [(il, cdagup, cdn), (il, cdagdn, cup), (-il, cup, cdagdn), (-il, cdn, cdagup)])
    terms += fermion_terms_with_strings(spec, N, f)
    H = kron_terms_dense(N, [3] * N, terms, dtype=complex)
    assert np.abs(H - H.conj().T).max() < 1e-14
    assert np.abs(H.imag).max() > 0.02          # the imaginary couplings are present
    qn = [1, 1, 0]
    keep = [k for k, s in enumerate(itertools.product(range(3), repeat=N)) if sum(qn[v] for v in s) == 5]
    e = np.linalg.eigvalsh(H[np.ix_(keep, keep)])[0]
    assert abs(e - (-11.018692166942165)) < 1e-12, e
    # without the imaginary y-bond couplings the energy changes: they are not a no-op
    Hr = H.real
    assert abs(np.linalg.eigvalsh(Hr[np.ix_(keep, keep)])[0] - e) > 1e-6


@pytest.mark.parametrize("name", ["heisenberg_1d_N6", "heisenberg_2d_6", "ising_1d_N6", "tj_1d_N4", "hubbard_2x2",
                                  "kitaev_simple_8", "kondo_8"])
def test_oracle_sweep_reproduces_golden_energy(name):
    case = HM.golden_cases()[name]
    m = case["build"]()
    init = case["init"]
    m.init_mps(init["labs"], enlarged_dim=init.get("enlarged", 0), seed=7, centralize=False)
    raw = [m.mps(i) for i in range(m.N)]
    mps = O.centralize0(raw)
    mpo = [m.mpo(i) for i in range(m.N)]
    if init.get("enlarged", 0):
        # ExtendDirectRandomInitMps normalises the head tensor after Centralize(0)
        h = mps[0]
        n = h.norm()
        mps[0] = BTensor(h.legs, h.qw, {c: b / n for c, b in h.blocks.items()})
    r = O.dmrg(mps, mpo, case["sweeps"], case["dmin"], case["dmax"], case["trunc"], case["lerr"], case["lmax"])
    assert abs(r["energy"] - case["e0"]) < case["tol"], (r["energy"], case["e0"])


def test_oracle_centralize_l100_random_start_is_right_canonical():
    """Centralize(0) of the seeded random 100-site start: the tensors it
    right-canonicalises carry the product of the site norms (~1e80 at
    site 1), past the range where sums of squares of their entries stay in
    FP64.  Every site i >= 1 must still come out right-orthonormal
    (sum_{s,r} A_i[l,s,r] A_i[l',s,r] = delta_{l l'}) and the state's norm
    must sit in the head tensor."""
    N = 100
    m = HM.heisenberg_chain(N)
    m.init_mps([[i % 2 for i in range(N)], [(i + 1) % 2 for i in range(N)]], enlarged_dim=8, seed=0x5EED,
               centralize=False)
    mps = O.centralize0([m.mps(i) for i in range(N)])
    for i in range(1, N):
        a = mps[i].dense()
        a2 = a.reshape(a.shape[0], -1)
        assert np.abs(a2 @ a2.T - np.eye(a2.shape[0])).max() < 1e-12, i
    assert 1e79 < mps[0].norm() < 1e82


def test_oracle_config1_heisenberg_l16_matches_ed():
    """Heisenberg OBC L=16, U(1) Sz=0, Dmax=64, 4 sweeps, Lanczos tol 1e-10:
    the oracle's final energy against exact diagonalisation."""
    import json
    import os
    ed = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))
    m = HM.heisenberg_chain(16)
    m.init_mps([[i % 2 for i in range(16)], [(i + 1) % 2 for i in range(16)]], enlarged_dim=4, seed=0x5EED,
               centralize=False)
    raw = [m.mps(i) for i in range(m.N)]
    mps = O.centralize0(raw)
    h = mps[0]
    n = h.norm()
    mps[0] = BTensor(h.legs, h.qw, {c: b / n for c, b in h.blocks.items()})
    r = O.dmrg(mps, [m.mpo(i) for i in range(m.N)], 4, 1, 64, 0.0, 1e-10, 200)
    assert abs(r["energy"] - ed["heisenberg_obc_L16_sz0"]) < 1e-8
    assert r["dims"].max() == 64


def test_cpp_host_containers(tmp_path):
    """The reference's DuoVector, TenVec file, SiteVec and MPO unit-test cases
    (test_duovector.cc, test_ten_vec.cc, test_site_vec.cc, test_mpo.cc) as a
    C++ program over the drop-in headers, compiled here and run on the host
    (no device call)."""
    import shutil
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib = os.path.join(root, "mps2_amd")
    if not os.path.exists(os.path.join(lib, "libmps2_amd.so")):
        pytest.fail("mps2_amd/libmps2_amd.so is not built (run __graft_entry__.build())")
    cxx = shutil.which("g++") or shutil.which("c++")
    exe = str(tmp_path / "test_containers")
    subprocess.run([cxx, "-O1", "-std=c++17", "-I", os.path.join(root, "include"), "-o", exe,
                    os.path.join(root, "tests", "cpp", "test_containers.cc"), "-L", lib, "-lmps2_amd",
                    "-Wl,-rpath," + lib], check=True, timeout=300)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout


@pytest.mark.parametrize("case", ["one_site", "zero_coef", "repeated", "chain4", "long_range5", "two_coupling4"])
def test_mpogen_fsm_cases_bond_dims_and_operator(case):
    """The term sets of test_mpogen.cc:79-207 (cases 1-5): the bond dimensions
    of the generated MPO equal the shapes of the reference's compressed FSM
    matrices (case 1: 1x1, 1x1; case 3: 1x2, 2x3, 3x2, 2x1; case 4: 1x2, 2x2,
    2x2, 2x2, 2x1; case 5: 1x2, 2x4, 4x2, 2x1), and the MPO contracts to the
    Kronecker sum of the same terms (a zero coefficient adds nothing, a
    repeated term doubles)."""
    from mps2_amd.gqmps2 import Model
    sz = np.diag([0.5, -0.5])
    site = [((0,), 1), ((1,), 1)]
    spec = {
        "one_site": (2, [(1.0, [0])], [1]),
        "zero_coef": (2, [(1.0, [0]), (0.0, [0])], [1]),
        "repeated": (2, [(1.0, [0, 1]), (1.0, [0, 1])], [1]),
        "chain4": (4, [(1.0, [0, 1]), (1.0, [1, 2]), (1.0, [2, 3])], [2, 3, 2]),
        "long_range5": (5, [(1.0, [0, 4]), (1.0, [1, 2, 4]), (1.0, [1, 2, 3])], [2, 2, 2, 2]),
        "two_coupling4": (4, [(0.5, [0, 1]), (0.5, [0, 2]), (0.5, [1, 3]), (0.5, [2, 3]), (2.0, [0, 3]),
                              (2.0, [1, 2])], [2, 4, 2]),
    }[case]
    N, terms, dims = spec
    m = Model([site] * N)
    for coef, sites in terms:
        m.term(coef, [sz] * len(sites), sites)
    assert list(m.gen_mpo()) == dims
    H = kron_terms_dense(N, [2] * N, [(c, [(s, sz) for s in st]) for c, st in terms])
    assert np.abs(mpo_dense(m) - H).max() < 1e-14
    m.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mpogen_random_terms_match_kron_sum(seed):
    """Random charge-conserving term sets on 6 spin-1/2 sites (one to four
    operators, insertion strings between them, repeated terms, terms that
    cancel, proportional remainders): the generated MPO contracts to the
    Kronecker sum of the same terms."""
    from mps2_amd.gqmps2 import Model
    rng = np.random.default_rng(seed)
    N = 6
    SZ, SP, SM = HM.SZ, HM.SP, HM.SM
    F = 2 * SZ
    patterns = [[SZ], [SZ, SZ], [SP, SM], [SM, SP], [SZ, SP, SM], [SP, SZ, SM], [SP, SP, SM, SM], [SZ, SZ, SZ]]
    m = Model([HM.SPIN_SITE] * N)
    dense_terms = []

    def add(coef, ops, sites, inst):
        m.term(coef, ops, sites, inst=[inst] * (len(ops) - 1) if inst is not None else None)
        full = list(zip(sites, ops))
        if inst is not None:
            for a, b in zip(sites[:-1], sites[1:]):
                full += [(k, inst) for k in range(a + 1, b)]
        dense_terms.append((coef, full))

    for _ in range(14):
        ops = patterns[rng.integers(len(patterns))]
        sites = sorted(rng.choice(N, size=len(ops), replace=False).tolist())
        coef = float(rng.choice([0.5, 1.0, 2.0, -1.5, rng.uniform(-1, 1)]))
        inst = F if (len(ops) > 1 and rng.random() < 0.4) else None
        add(coef, ops, sites, inst)
        if rng.random() < 0.2:
            add(coef, ops, sites, inst)            # repeated
        if rng.random() < 0.15:
            add(-coef, ops, sites, inst)           # cancels one copy
    m.gen_mpo()
    H = kron_terms_dense(N, [2] * N, dense_terms)
    assert np.abs(mpo_dense(m) - H).max() < 1e-13
    m.close()


def test_mpogen_tail_insertion_string():
    """AddTerm with as many insertion operators as physical ones
    (mpogen_impl.h:144-199): the last one is listed from the last physical
    site to the end of the chain; at the last physical site the physical
    operator takes precedence (the first listed operator of a site is used,
    mpogen_impl.h:99-113), so the string covers the sites after it."""
    from mps2_amd.gqmps2 import Model
    N = 6
    F = 2 * HM.SZ
    m = Model([HM.SPIN_SITE] * N)
    m.term(0.7, [HM.SP, HM.SM], [1, 3], inst=[F, HM.SZ])
    m.term(-0.3, [HM.SZ, HM.SZ], [0, 4])
    m.gen_mpo()
    # F between 1 and 3 (site 2), then SZ on every site after the last physical one
    t1 = [(1, HM.SP), (2, F), (3, HM.SM), (4, HM.SZ), (5, HM.SZ)]
    H = kron_terms_dense(N, [2] * N, [(0.7, t1), (-0.3, [(0, HM.SZ), (4, HM.SZ)])])
    assert np.abs(mpo_dense(m) - H).max() < 1e-13
    m.close()


def test_oracle_ground_state_correlations_match_ed():
    """Oracle pipeline on the host: two-site DMRG of the 10-site Heisenberg
    chain (exact at D=32), then the oracle's transfer-chain expectation values
    on its MPS against the exact ground state: <Sz_i>, <Sz_i Sz_j>,
    <S+_i S-_j> and the energy as the expectation of the MPO-free sum."""
    n = 10
    m = HM.heisenberg_chain(n)
    m.init_mps([[i % 2 for i in range(n)]], centralize=False)
    raw = [m.mps(i) for i in range(n)]
    mps = O.centralize0(raw)
    r = O.dmrg(mps, [m.mpo(i) for i in range(n)], 6, 8, 64, 0.0, 1e-12, 200)
    H = kron_terms_dense(n, [2] * n, heis_terms([(i, i + 1) for i in range(n - 1)]))
    w, v = np.linalg.eigh(H)
    psi = v[:, 0]
    assert abs(r["energy"] - w[0]) < 1e-10
    phys = HM.SPIN_SITE
    from mps2_amd.flat import Leg, OUT
    pl = Leg(OUT, phys)
    from tests import helpers as TH

    def mps_val(opmap):
        return O.expect(r["mps"], [TH.dense_op_tensor(pl, opmap[k]) if k in opmap else None for k in range(n)])

    def ed_val(opmap):
        return TH.dense_expect(psi, [2] * n, [opmap.get(k) for k in range(n)])

    norm = mps_val({})
    assert abs(norm - 1.0) < 1e-12
    worst = 0.0
    for i in range(n):
        worst = max(worst, abs(mps_val({i: HM.SZ}) - ed_val({i: HM.SZ})))
        for j in range(i + 1, n):
            for a, b in ((HM.SZ, HM.SZ), (HM.SP, HM.SM)):
                worst = max(worst, abs(mps_val({i: a, j: b}) - ed_val({i: a, j: b})))
    assert worst < 1e-9, worst
    e_corr = sum(mps_val({i: HM.SZ, i + 1: HM.SZ}) + 0.5 * mps_val({i: HM.SP, i + 1: HM.SM})
                 + 0.5 * mps_val({i: HM.SM, i + 1: HM.SP}) for i in range(n - 1))
    assert abs(e_corr - w[0]) < 1e-9


def test_model_create_rejects_bad_site_tables():
    """gqmps2_model_create returns an error status (raised here) for a
    zero-dimensional sector, a site without sectors, an unsupported QN width
    and a one-site chain, and a good table still builds afterwards."""
    from mps2_amd import gqmps2 as G
    up_dn = [((1,), 1), ((-1,), 1)]
    bad = [([up_dn, [((1,), 0), ((-1,), 1)]], 1, "positive"),
           ([up_dn, []], 1, "at least one sector"),
           ([[((1, 0, 0), 1)], [((1, 0, 0), 1)]], 3, "qn_width"),
           ([up_dn], 1, "two sites")]
    for sites, qw, msg in bad:
        with pytest.raises(RuntimeError, match=msg):
            G.Model(sites, qw=qw)
    m = G.Model([up_dn, up_dn])
    m.close()


def test_cpp_case_params_parser(tmp_path):
    """test_case_params_parser.cc:17-53 as a C++ program over the headers,
    compiled here and run on the CPU with the reference's fixture
    (tests/golden/test-params.json, a copy of its testdata file), plus edge
    cases of the JSON reader; a file without a CaseParams object exits with
    status 1 and the reference's message."""
    import shutil
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib = os.path.join(root, "mps2_amd")
    cxx = shutil.which("g++") or shutil.which("c++")
    exe = str(tmp_path / "test_case_params_parser")
    subprocess.run([cxx, "-O1", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"), "-o", exe,
                    os.path.join(root, "tests", "cpp", "test_case_params_parser.cc"), "-L", lib, "-lmps2_amd",
                    "-Wl,-rpath," + lib], check=True, timeout=300)
    fixture = os.path.join(root, "tests", "golden", "test-params.json")
    r = subprocess.run([exe, fixture], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout
    r = subprocess.run([exe, fixture, "missing"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "CaseParams object not found, exit!" in r.stdout


def test_model_add_term_rejects_bad_terms():
    """gqmps2_model_add_term returns an error status (raised here) where the
    generator would assert: sites out of order or out of range, an insertion
    count that is neither one per gap nor one per gap plus the tail, and an
    operator or insertion operator built for another local space than the
    sites it lands on; good terms are still accepted afterwards."""
    import ctypes as C
    from mps2_amd import gqmps2 as G
    spin = [((1,), 1), ((-1,), 1)]
    m = G.Model([spin, spin, [((0,), 4)], spin])
    SZ, F = np.diag([0.5, -0.5]), np.diag([-1.0, -1.0])
    bad = [(lambda: m.term(1.0, [SZ, SZ], [1, 0]), "ascending"),
           (lambda: m.term(1.0, [SZ, SZ], [0, 7]), "out of range"),
           (lambda: m.term(1.0, [SZ, SZ, SZ], [0, 1, 3], inst=[F]), "one per gap"),
           (lambda: m.term(1.0, [SZ, SZ], [0, 3], inst=[F]), "insertion operator built")]
    for f, msg in bad:
        with pytest.raises(RuntimeError, match=msg):
            f()
    big = m.op(2, np.eye(4))                       # registered for the 4-dimensional site
    ids = np.asarray([m.op(0, SZ), big], dtype=np.int32)
    st = np.asarray([0, 1], dtype=np.int64)
    with pytest.raises(RuntimeError, match="local space"):
        G.check(G.lib().gqmps2_model_add_term(m.h, 1.0, 2, ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                              st.ctypes.data_as(C.POINTER(C.c_int64)), 0, None))
    with pytest.raises(ValueError, match="must be 4x4"):
        m.op(2, np.eye(2))                         # would be read as 16 elements
    m.term(1.0, [SZ, SZ], [0, 1])
    m.term(1.0, [SZ, np.eye(4)], [1, 2])
    with pytest.raises(RuntimeError, match="insertion operator built"):
        m.term(1.0, [SZ, SZ], [0, 1], inst=[F, F])     # tail string over the 4-dimensional site 2
    m.term(1.0, [SZ, np.eye(4)], [0, 2], inst=[F, F])  # gap site 1 and tail site 3 are spins
    m.close()


def test_model_entry_points_check_handles_and_sites():
    """NULL model handles, MPO export before gen_mpo, and site numbers outside
    the chain give error statuses from the model-level C ABI."""
    import ctypes as C
    from mps2_amd import gqmps2 as G
    L = G.lib()
    i32p, i64p, dp = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double)
    n = C.c_int64()
    for rc in (L.gqmps2_mps_checkpoint(None), L.gqmps2_mps_restore(None),
               L.gqmps2_mps_export_sizes(None, 0, C.byref(n), C.byref(n)),
               L.gqmps2_model_mpo_export_sizes(None, 0, C.byref(n), C.byref(n)),
               L.gqmps2_mps_export(None, 0, None, None), L.gqmps2_model_mpo_export(None, 0, None, None)):
        assert rc != 0
        assert "NULL" in L.gqmps2_last_error().decode()
    spin = [((1,), 1), ((-1,), 1)]
    m = G.Model([spin] * 4)
    rc = L.gqmps2_model_mpo_export_sizes(m.h, 0, C.byref(n), C.byref(n))
    assert rc != 0 and "gen_mpo" in L.gqmps2_last_error().decode()
    m.term(1.0, [np.diag([0.5, -0.5])] * 2, [0, 1])
    m.gen_mpo()
    for site in (-1, 4):
        assert L.gqmps2_model_mpo_export_sizes(m.h, site, C.byref(n), C.byref(n)) != 0
        assert L.gqmps2_mps_export_sizes(m.h, site, C.byref(n), C.byref(n)) != 0
    m.init_mps([[0, 1, 0, 1]], centralize=False)
    meta = np.zeros(64, dtype=np.int32)
    data = np.zeros(64)
    assert L.gqmps2_mps_export(m.h, 4, meta.ctypes.data_as(i32p), data.ctypes.data_as(dp)) != 0
    assert L.gqmps2_mps_import(m.h, 9, meta.ctypes.data_as(i32p), data.ctypes.data_as(dp)) != 0
    labs = np.zeros(4, dtype=np.int64)
    assert L.gqmps2_mps_init(m.h, 1, labs.ctypes.data_as(i64p), -1, 0, 0) != 0
    # a good export/import round trip still works
    t = m.mps(1)
    m.set_mps(1, t)
    assert np.array_equal(m.mps(1).dense(), t.dense())
    m.close()


def test_model_checkpoint_and_sweeps_need_mps_and_mpo():
    """checkpoint before init_mps, and dmrg / bench_bond before gen_mpo, return
    error statuses (the reference asserts mps.size() == mpo.size(),
    two_site_update_finite_vmps_impl.h:83) instead of reading missing tensors."""
    from mps2_amd import gqmps2 as G
    L = G.lib()
    spin = [((1,), 1), ((-1,), 1)]
    m = G.Model([spin] * 4)
    assert L.gqmps2_mps_checkpoint(m.h) != 0
    assert "MPS not initialised" in L.gqmps2_last_error().decode()
    m.term(1.0, [np.diag([0.5, -0.5])] * 2, [0, 1])
    m.init_mps([[0, 1, 0, 1]], centralize=False)
    with pytest.raises(RuntimeError, match="gen_mpo"):
        m.dmrg(1, 1, 4, 0.0, 1e-10)
    assert L.gqmps2_bench_bond(m.h, 1, 1, 1) != 0
    assert "gen_mpo" in L.gqmps2_last_error().decode()
    m.close()
