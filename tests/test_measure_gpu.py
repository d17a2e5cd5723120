"""Device measurements (MeasureOneSiteOp / MeasureTwoSiteOp /
MeasureMultiSiteOp, finite_mps_measu.h:150-582) through the C ABI.

Checks: the reference's own product-state cases (test_finite_mps_measu.cc:43-170)
exactly, both as a C++ program over the drop-in headers and through the C
ABI; DMRG ground states against the oracle's gauge-free transfer chain on
the same MPS, against exact diagonalisation (Heisenberg chain) and against
the free-fermion closed form (Jordan-Wigner strings); bit-identical results
whether or not consecutive events reuse the open string; the result file.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from mps2_amd import engine as E
from mps2_amd import hamiltonians as Hm
from mps2_amd.flat import OUT, Leg
from mps2_amd.gqmps2 import Model
from oracle import oracle_py as O
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu

OCC_SITE = [((0,), 1), ((1,), 1)]     # basis 0 = empty (N=0), 1 = occupied (N=1)
NTOT = np.diag([0.0, 1.0])
ID2 = np.eye(2)
F = np.diag([1.0, -1.0])              # Jordan-Wigner string (1 - 2n)
ADAG = np.array([[0.0, 1.0], [0.0, 0.0]])
ANN = np.array([[0.0, 0.0], [1.0, 0.0]])


def _events(pairs, a, b, inst):
    return [([a, b], [i, j], [inst]) for (i, j) in pairs]


def _oracle_pair(mps, phys, a, b, i, j, inst):
    n = len(mps)
    ops = [None] * n
    ops[i] = H.dense_op_tensor(phys, a)
    ops[j] = H.dense_op_tensor(phys, b)
    if inst is not None:
        for k in range(i + 1, j):
            ops[k] = H.dense_op_tensor(phys, inst)
    return O.expect(mps, ops)


def test_cpp_reference_cases(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "test_finite_mps_measu")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASSED" in r.stdout
    assert (tmp_path / "op1.json").exists() and (tmp_path / "op1op2.json").exists()


def test_reference_product_states_capi(tmp_path):
    n = 6
    m = Model([OCC_SITE] * n)
    labs1 = [1] * n
    labs2 = [i % 2 for i in range(n)]
    sites_set = [(0, 1), (0, 2), (0, 5), (1, 2), (1, 3), (4, 5)]
    for labs in (labs1, labs2):
        m.init_mps([labs])
        got = m.measure_one_site([NTOT], [tmp_path / "op1"])[0]
        assert list(got) == labs
    m.init_mps([labs1])
    assert list(m.measure_multi_site(_events(sites_set, ID2, ID2, ID2), tmp_path / "a")) == [1.0] * 6
    assert list(m.measure_multi_site(_events(sites_set, NTOT, NTOT, ID2), tmp_path / "b")) == [1.0] * 6
    m.init_mps([labs2])
    assert list(m.measure_multi_site(_events(sites_set, ID2, ID2, ID2), tmp_path / "c")) == [1.0] * 6
    assert list(m.measure_multi_site(_events(sites_set, NTOT, NTOT, ID2), tmp_path / "d")) == [0, 0, 0, 0, 1, 0]
    m.close()


def _heisenberg_ed(n):
    sz, sp, sm = Hm.SZ, Hm.SP, Hm.SM
    dim = 2 ** n

    def site_op(o, k):
        return np.kron(np.kron(np.eye(2 ** k), o), np.eye(2 ** (n - k - 1)))

    Hd = np.zeros((dim, dim))
    for i in range(n - 1):
        Hd += site_op(sz, i) @ site_op(sz, i + 1)
        Hd += 0.5 * (site_op(sp, i) @ site_op(sm, i + 1) + site_op(sm, i) @ site_op(sp, i + 1))
    w, v = np.linalg.eigh(Hd)
    return w[0], v[:, 0]


def test_heisenberg_correlations_vs_oracle_and_ed(tmp_path):
    n = 10
    m = Hm.heisenberg_chain(n)
    m.init_mps([[i % 2 for i in range(n)]])
    r = m.dmrg(6, 8, 64, 0.0, 1e-12)
    e_ed, psi_ed = _heisenberg_ed(n)
    assert abs(r["energy"] - e_ed) < 1e-10, (r["energy"], e_ed)
    m.checkpoint()
    mps = [m.mps(i) for i in range(n)]
    phys = Leg(OUT, Hm.SPIN_SITE)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]

    # one-site values
    sz1 = m.measure_one_site([Hm.SZ, 2 * Hm.SZ], [tmp_path / "sz", tmp_path / "sz2"])
    for i in range(n):
        ops = [None] * n
        ops[i] = H.dense_op_tensor(phys, Hm.SZ)
        want = O.expect(mps, ops)
        assert abs(sz1[0, i] - want) < 1e-12
        assert abs(sz1[1, i] - 2 * want) < 2e-12
        assert abs(sz1[0, i] - H.dense_expect(psi_ed, [2] * n, ops_dense(n, {i: Hm.SZ}))) < 1e-9

    # two-site strings, identity insertion (-1 = the site's identity, the shortcut path)
    m.restore()
    zz = m.measure_multi_site(_events(pairs, Hm.SZ, Hm.SZ, None), tmp_path / "zz")
    m.restore()
    pm = m.measure_multi_site(_events(pairs, Hm.SP, Hm.SM, None), tmp_path / "pm")
    worst_o = worst_ed = 0.0
    for k, (i, j) in enumerate(pairs):
        for got, a, b in ((zz[k], Hm.SZ, Hm.SZ), (pm[k], Hm.SP, Hm.SM)):
            want = _oracle_pair(mps, phys, a, b, i, j, None)
            ed = H.dense_expect(psi_ed, [2] * n, ops_dense(n, {i: a, j: b}))
            worst_o = max(worst_o, abs(got - want))
            worst_ed = max(worst_ed, abs(got - ed))
    print(f"max |device - oracle| {worst_o:.2e}, max |device - ED| {worst_ed:.2e}")
    assert worst_o < 1e-12
    assert worst_ed < 1e-8

    # reusing the open string changes nothing: same events, each head's
    # partners in descending order (no reuse), from the same saved state
    desc = sorted(pairs, key=lambda p: (p[0], -p[1]))
    m.restore()
    zz_desc = m.measure_multi_site(_events(desc, Hm.SZ, Hm.SZ, None), tmp_path / "zz_desc")
    back = {p: v for p, v in zip(desc, zz_desc)}
    assert np.array_equal(zz, np.array([back[p] for p in pairs]))

    # result file: the reference's layout, 12 significant digits
    rows = json.loads((tmp_path / "zz.json").read_text())
    assert [tuple(r[0]) for r in rows] == pairs
    assert np.allclose([r[1] for r in rows], zz, rtol=1e-11, atol=1e-15)
    m.close()


def ops_dense(n, d):
    return [d.get(k) for k in range(n)]


def test_free_fermion_strings_closed_form(tmp_path):
    """Spinless fermions H = -t sum_i (c+_i c_{i+1} + h.c.) on an open chain
    of 10 sites at 5 particles: <c+_i c_j> = sum_{k occupied} phi_k(i) phi_k(j)
    exactly, which needs the Jordan-Wigner string F = 1 - 2n on the sites in
    between (a non-identity insertion operator)."""
    n, nf, t = 10, 5, 1.0
    m = Model([OCC_SITE] * n)
    for i in range(n - 1):
        m.term(-t, [ADAG, ANN], [i, i + 1])
        m.term(-t, [ANN, ADAG], [i, i + 1])
    m.gen_mpo()
    m.init_mps([[1, 0] * (n // 2)])
    r = m.dmrg(6, 8, 64, 0.0, 1e-12)
    h1 = -t * (np.eye(n, k=1) + np.eye(n, k=-1))
    eps, phi = np.linalg.eigh(h1)
    G = phi[:, :nf] @ phi[:, :nf].T
    assert abs(r["energy"] - eps[:nf].sum()) < 1e-10
    m.checkpoint()
    mps = [m.mps(i) for i in range(n)]
    phys = Leg(OUT, OCC_SITE)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    got = m.measure_multi_site(_events(pairs, ADAG, ANN, F), tmp_path / "cdc")
    dens = m.measure_one_site([NTOT], [tmp_path / "n"])[0]
    worst_o = 0.0
    for k, (i, j) in enumerate(pairs):
        worst_o = max(worst_o, abs(got[k] - _oracle_pair(mps, phys, ADAG, ANN, i, j, F)))
    want = np.array([G[i, j] for (i, j) in pairs])
    print(f"max |device - closed form| {np.abs(got - want).max():.2e}, |device - oracle| {worst_o:.2e}")
    assert np.abs(got - want).max() < 1e-8
    assert np.abs(dens - np.diag(G)).max() < 1e-8
    assert worst_o < 1e-12
    # without the string the values are those of hard-core bosons, not G
    m.restore()
    nostr = m.measure_multi_site(_events(pairs, ADAG, ANN, None), tmp_path / "bos")
    assert np.abs(nostr - want).max() > 1e-2
    m.close()


def test_measure_argument_errors(tmp_path):
    n = 6
    m = Model([OCC_SITE] * n)
    m.init_mps([[1] * n])
    with pytest.raises(RuntimeError, match="ascending"):
        m.measure_multi_site([([NTOT, NTOT], [3, 1], [None])], tmp_path / "x")
    with pytest.raises(RuntimeError, match="out of range"):
        m.measure_multi_site([([NTOT, NTOT], [1, 7], [None])], tmp_path / "x")
    # a failed call leaves the model usable
    assert list(m.measure_multi_site([([NTOT, NTOT], [1, 3], [None])], tmp_path / "y")) == [1.0]
    m.close()


def test_correlation_sweep_reuse_at_d256(tmp_path):
    """<Sz_0 Sz_j>, j = 1..39, on a 40-site Heisenberg ground state at D=256:
    ascending partners reuse the open string (39 middle steps in all),
    descending ones rebuild it (741); values agree bit for bit, and the
    ascending sweep agrees with the oracle's chain on the same MPS."""
    import time
    n = 40
    m = Hm.heisenberg_chain(n)
    m.init_mps([[i % 2 for i in range(n)]])
    m.dmrg(3, 8, 256, 0.0, 1e-10, dmax_schedule=[64, 256, 256])
    m.checkpoint()
    asc = [(0, j) for j in range(1, n)]
    times = {}
    vals = {}
    # one untimed pass first: it creates the contraction plans of every site,
    # which both timed passes then find in the plan cache
    m.restore()
    m.measure_multi_site(_events(asc[::-1], Hm.SZ, Hm.SZ, None), tmp_path / "warm")
    for name, pairs in (("asc", asc), ("desc", asc[::-1])):
        m.restore()
        # Centralize(0) of the restored MPS (39 untruncated SVDs) happens here, untimed
        m.measure_multi_site(_events([(0, 1)], Hm.SZ, Hm.SZ, None), tmp_path / "centre")
        t0 = time.perf_counter()
        v = m.measure_multi_site(_events(pairs, Hm.SZ, Hm.SZ, None), tmp_path / name)
        times[name] = time.perf_counter() - t0
        vals[name] = dict(zip(pairs, v))
    print(f"<Sz_0 Sz_j> sweep, L={n}, D=256: reuse {times['asc'] * 1e3:.1f} ms, "
          f"rebuild {times['desc'] * 1e3:.1f} ms")
    assert all(vals["asc"][p] == vals["desc"][p] for p in asc)
    mps = [m.mps(i) for i in range(n)]
    phys = Leg(OUT, Hm.SPIN_SITE)
    for j in (1, 2, 17, 39):
        want = _oracle_pair(mps, phys, Hm.SZ, Hm.SZ, 0, j, None)
        assert abs(vals["asc"][(0, j)] - want) < 1e-12, (j, vals["asc"][(0, j)], want)
    m.close()


def test_cpp_finite_mps_cases(tmp_path):
    """test_finite_mps.cc:31-260 as a C++ program over the drop-in headers:
    Centralize to every site (canonical tensors contract to the identity to
    1e-14, flags, centre), copy/move, const versus mutable access, Dump/Load,
    and TruncateMPS (bond dimensions, left-canonical tensors, normalised
    result, unchanged state when nothing is cut)."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_finite_mps")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout


def test_truncate_mps_matches_oracle():
    """TruncateMPS (finite_mps.h:245-303) on the device against the oracle's
    restatement, on a random 8-site U(1) MPS cut to Dmax = 3: the same kept
    dimension on every bond, the same truncated state (overlap 1 to 1e-10),
    and a normalised result; with no cut the state is unchanged."""
    rng = np.random.default_rng(92)
    n = 8
    src = H.random_spin_mps(rng, n, maxdim=3)
    psi = H.mps_dense_state(src)
    psi = psi / np.linalg.norm(psi)
    for dmax in (3, 1000):
        m = Model([Hm.SPIN_SITE] * n)
        for i, t in enumerate(src):
            m.set_mps(i, t)
        m.truncate_mps(0.0, 1, dmax)
        got = [m.mps(i) for i in range(n)]
        ref, dims = H.oracle_truncate_mps(src, 0.0, 1, dmax)
        assert [g.legs[-1].dim for g in got[:-1]] == dims, dmax
        a, b = H.mps_dense_state(got), H.mps_dense_state(ref)
        assert abs(np.linalg.norm(a) - 1.0) < 1e-12
        assert abs(abs(a @ b) - 1.0) < 1e-10, (dmax, a @ b)
        if dmax == 1000:
            assert abs(abs(a @ psi) - 1.0) < 1e-12
        m.close()
