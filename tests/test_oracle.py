"""Pins of the CPU FP64 oracle (no GPU).

Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) fails it: dense numpy einsum on small block-sparse
operands, closed forms, and dense eigensolvers.
"""
import numpy as np
import pytest

from mps2_amd.flat import BTensor, IN, OUT, Leg
from oracle import oracle_py as O
from tests import helpers as H


def contract_case(rng, axes, make):
    """Operands of one of the four axis patterns with charges that admit
    blocks on both sides (odd charges on l, even on r, spin-1/2 p)."""
    l0 = H.bond(IN, (3, 2, 4), (1, -1, 3))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (3, 2, 1), (0, 2, -2))
    A = make([l0, p, r], (0,), rng)
    if axes == [[2], [0]]:
        legs = [r.inverse(), p, H.bond(OUT, (2, 3), (1, -1))]
    elif axes == [[0], [0]]:
        legs = [l0.inverse(), p, H.bond(OUT, (2, 3, 1), (0, 2, -2))]
    elif axes == [[2, 1], [0, 1]]:
        legs = [r.inverse(), p.inverse(), H.bond(OUT, (2, 3, 2), (1, -1, 3))]
    else:
        legs = [r.inverse(), H.bond(OUT, (2, 3), (1, 3)), l0.inverse()]
    B = make(legs, (0,), rng)
    assert len(A.blocks) >= 4 and len(B.blocks) >= 2
    return A, B


@pytest.mark.parametrize("axes,spec", [
    ([[2], [0]], "abc,cde->abde"),
    ([[0], [0]], "abc,ade->bcde"),
    ([[2, 1], [0, 1]], "abc,cbe->ae"),
    ([[0, 2], [2, 0]], "abc,cda->bd"),
])
def test_contract_matches_dense_einsum(axes, spec):
    rng = np.random.default_rng(3)
    A, B = contract_case(rng, axes, BTensor.random)
    C = O.contract(A, B, axes)
    ref = np.einsum(spec, A.dense(), B.dense())
    assert len(C.blocks) >= 1 and np.abs(ref).max() > 0.1
    assert np.abs(C.dense() - ref).max() < 1e-13
    # output blocks: exactly the coordinates that receive a block-pair product
    for c, blk in C.blocks.items():
        assert blk.shape == C.block_shape(c)


def test_contract_rejects_non_inverse_legs():
    rng = np.random.default_rng(0)
    p = H.spin_phys(OUT)
    A = BTensor.random([p, p.inverse()], (0,), rng)
    with pytest.raises(RuntimeError):
        O.contract(A, A, [[0], [0]])


def test_tridiag_toeplitz_closed_form():
    # lambda_k = a + 2 b cos(k pi / (n + 1)); lowest is k = n for b > 0
    n, a, b = 37, 0.3, 0.7
    ev, v = O.tridiag_lowest(np.full(n, a), np.full(n - 1, b))
    lam = a + 2 * b * np.cos(n * np.pi / (n + 1))
    assert abs(ev - lam) < 1e-14
    k = np.arange(1, n + 1)
    ref = np.sin(k * n * np.pi / (n + 1))
    ref /= np.linalg.norm(ref)
    assert min(np.abs(v - ref).max(), np.abs(v + ref).max()) < 1e-12


def test_tridiag_random_vs_eigh():
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 60, 200):
        a, b = rng.normal(size=n), rng.normal(size=max(n - 1, 0))
        ev, v = O.tridiag_lowest(a, b)
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        w, V = np.linalg.eigh(T)
        assert abs(ev - w[0]) < 1e-12 * max(1, abs(w).max())
        assert np.linalg.norm(T @ v - ev * v) < 1e-11


def test_svd_reconstructs_and_truncates():
    rng = np.random.default_rng(5)
    l0 = H.bond(IN, (3, 2, 4), (0, 2, -2))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (3, 2, 1, 2), (0, 2, -2, 4))
    psi = BTensor.random([l0, p, p, r], (0,), rng)
    U, S, Vt, err, D = O.svd(psi, 2, (0,), 0.0, 1000, 1000)
    rec = np.einsum("abk,kl,lcd->abcd", U.dense(), S.dense(), Vt.dense())
    assert np.abs(rec - psi.dense()).max() < 1e-13
    u = U.dense().reshape(-1, U.dense().shape[-1])
    assert np.abs(u.T @ u - np.eye(u.shape[1])).max() < 1e-13
    v = Vt.dense().reshape(Vt.dense().shape[0], -1)
    assert np.abs(v @ v.T - np.eye(v.shape[0])).max() < 1e-13
    # singular values of the QN-blocked matrix equal those of the dense matrix
    dense_s = np.linalg.svd(psi.dense().reshape(psi.dense().shape[0] * 2, -1), compute_uv=False)
    ours = np.sort(np.diag(S.dense()))[::-1]
    assert np.abs(ours - dense_s[:len(ours)]).max() < 1e-13
    # truncation error = discarded weight
    U5, S5, V5, err5, D5 = O.svd(psi, 2, (0,), 0.0, 5, 5)
    assert D5 == 5
    assert abs(err5 - (dense_s[5:] ** 2).sum() / (dense_s ** 2).sum()) < 1e-13
    # trunc_err drives D when Dmax does not bind
    target = (dense_s[7:] ** 2).sum() / (dense_s ** 2).sum() * 1.0000001
    _, _, _, e7, D7 = O.svd(psi, 2, (0,), target, 1, 1000)
    assert D7 == 7 and e7 <= target


def test_svd_degenerate_cut_is_deterministic():
    # two sectors with identical spectra: the cut takes the lower-QN sector first
    q = H.bond(IN, (2, 2), (-1, 1))
    out = H.bond(OUT, (2, 2), (-1, 1))
    t = BTensor([q, out], 1)
    t.blocks[(0, 0)] = np.diag([3.0, 1.0])
    t.blocks[(1, 1)] = np.diag([3.0, 1.0])
    U, S, Vt, err, D = O.svd(t, 1, (0,), 0.0, 1, 3)
    assert D == 3
    dims = [d for _, d in U.legs[-1].sectors]
    qns = [qq for qq, _ in U.legs[-1].sectors]
    assert qns == [(-1,), (1,)] and dims == [2, 1]


def test_lanczos_matches_dense_eigh():
    rng = np.random.default_rng(11)
    L, W1, W2, R, psi = H.hermitian_cent_problem(rng)
    Hd = H.dense_heff_cent(L, W1, W2, R)
    assert np.abs(Hd - Hd.T).max() < 1e-12
    # restrict to the divergence-0 subspace spanned by psi's allowed blocks
    it, e, gs = O.lanczos(L, W1, W2, R, psi, "cent", 1e-12, 200)
    w = np.linalg.eigvalsh(Hd)
    x = gs.dense().ravel()
    assert abs(np.linalg.norm(x) - 1) < 1e-10
    rq = x @ Hd @ x
    assert abs(rq - e) < 1e-8
    assert e >= w[0] - 1e-8
    # max_iter termination mode (test_lanczos_solver.cc:150-168)
    it2, e2, _ = O.lanczos(L, W1, W2, R, psi, "cent", 1e-16, 20)
    assert it2 == 19


def test_heff_hermiticity_invariant():
    rng = np.random.default_rng(2)
    L, W1, W2, R, psi = H.hermitian_cent_problem(rng)
    x = psi
    y = BTensor.random(psi.legs, (0,), rng)
    hx = O.heff_mul(L, W1, W2, R, x, "cent")
    hy = O.heff_mul(L, W1, W2, R, y, "cent")
    lhs = float((y.dense() * hx.dense()).sum())
    rhs = float((hy.dense() * x.dense()).sum())
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))


def test_heff_cent_matches_dense():
    rng = np.random.default_rng(4)
    L, W1, W2, R, psi = H.random_cent_problem(rng)
    hp = O.heff_mul(L, W1, W2, R, psi, "cent")
    ref = np.einsum("abc,bdef,fghi,jik,adgj->cehk", L.dense(), W1.dense(), W2.dense(), R.dense(), psi.dense())
    assert np.abs(hp.dense() - ref).max() < 1e-12


def test_svd_scale_equivariance():
    """SVD(c A) = U (c S) V^T: scaling by 2^300 (exact in binary floating
    point) leaves U and V^T bit-identical and S scaled exactly; scales
    1e150 and 1e-150, where unscaled sums of squares would leave the FP64
    range, agree with the unscaled result to rounding and keep the same
    dimensions and truncation error."""
    rng = np.random.default_rng(77)
    l0 = H.bond(IN, (30, 50, 20), (0, 2, -2))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (31, 51, 21, 2), (0, 2, -2, 4))
    psi = BTensor.random([l0, p, p, r], (0,), rng)
    U0, S0, V0, e0, D0 = O.svd(psi, 2, (0,), 0.0, 1, 40)
    s0 = np.diag(S0.dense())
    for sc in (2.0 ** 300, 1e150, 1e-150):
        t = BTensor(psi.legs, psi.qw, {c: b * sc for c, b in psi.blocks.items()})
        U, S, V, e, D = O.svd(t, 2, (0,), 0.0, 1, 40)
        assert D == D0 and U.legs[-1].key() == U0.legs[-1].key()
        s = np.diag(S.dense())
        if sc == 2.0 ** 300:
            assert np.array_equal(s, s0 * sc) and e == e0
            assert all(np.array_equal(U.blocks[k], U0.blocks[k]) for k in U0.blocks)
            assert all(np.array_equal(V.blocks[k], V0.blocks[k]) for k in V0.blocks)
        else:
            assert np.abs(s / sc - s0).max() < 1e-13 * s0.max()
            assert abs(e - e0) < 1e-13
    bad = BTensor(psi.legs, psi.qw, {c: b.copy() for c, b in psi.blocks.items()})
    next(iter(bad.blocks.values())).flat[0] = np.nan
    with pytest.raises(RuntimeError, match="NaN"):
        O.svd(bad, 2, (0,), 0.0, 1, 40)


def test_svd_unresolved_values_rule():
    """Values below 1e-11 s_max have no reproducible order: with trunc_err = 0
    they are all kept (a direct SVD returns them as tiny positive numbers),
    and a cut among them is a tie decided by (sector, index), so one state
    beyond the rank goes to the lowest-QN sector.  A cut above them is
    decided by value (sector ranks 2, 4, 3, 3)."""
    psi = H.rank_deficient_state()
    U, S, V, e, D = O.svd(psi, 2, (0,), 0.0, 1, 10_000)
    s = np.diag(S.dense())
    assert D == 124 and (s < 1e-11 * s.max()).sum() == 112
    U, S, V, e, D = O.svd(psi, 2, (0,), 0.0, 1, 12)
    assert U.legs[-1].key() == (1, (((-3,), 2), ((-1,), 4), ((1,), 3), ((3,), 3)))
    U, S, V, e, D = O.svd(psi, 2, (0,), 0.0, 1, 13)
    assert U.legs[-1].key() == (1, (((-3,), 3), ((-1,), 4), ((1,), 3), ((3,), 3)))
    U, S, V, e, D = O.svd(psi, 2, (0,), 0.0, 1, 16)
    assert U.legs[-1].key() == (1, (((-3,), 6), ((-1,), 4), ((1,), 3), ((3,), 3)))
    # the discarded weight of the unresolved values is negligible
    assert e < 1e-20


# ---------------------------------------------------------------- measurements
SZ = np.diag([0.5, -0.5])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])
SM = np.array([[0.0, 0.0], [1.0, 0.0]])


def _expect_cases(n, rng):
    cases = [[None] * n]                                   # the squared norm
    for i in range(n):                                     # one-site
        ops = [None] * n
        ops[i] = SZ
        cases.append(ops)
    for (i, j) in [(0, 1), (0, n - 1), (1, 4), (2, 3)]:    # strings with and without insertions
        ops = [None] * n
        ops[i], ops[j] = SZ, SZ
        cases.append(ops)
        ops = [None] * n
        ops[i], ops[j] = SP, SM
        for k in range(i + 1, j):
            ops[k] = 2 * SZ
        cases.append(ops)
    cases.append([np.diag(rng.uniform(-1, 1, 2)) for _ in range(n)])   # every site
    return cases


def test_expect_matches_dense_state():
    """The oracle's transfer chain against the dense state vector of a random
    (non-canonical, unnormalised) U(1) MPS: one-site values, strings with
    identity and non-identity insertions, QN-changing pairs (S+ S-) and an
    operator on every site.  A transposed operator, a skipped site or a
    wrong contraction axis fails one of them."""
    rng = np.random.default_rng(77)
    n = 6
    p = H.spin_phys(OUT)
    for div in (0, 2):
        mps = H.random_spin_mps(rng, n, div=div)
        psi = H.mps_dense_state(mps)
        norm2 = float(psi @ psi)
        for ops in _expect_cases(n, rng):
            want = H.dense_expect(psi, [2] * n, ops)
            got = O.expect(mps, [H.dense_op_tensor(p, o) if o is not None else None for o in ops])
            assert abs(got - want) <= 1e-13 * norm2, (ops, got, want)
        # S+ alone changes the charge: zero, with no block left in the scalar
        ops = [None] * n
        ops[2] = SP
        assert O.expect(mps, [H.dense_op_tensor(p, o) if o is not None else None for o in ops]) == 0.0


def test_expect_gauge_invariant():
    """Moving the gauge (Centralize(0) by untruncated SVDs) leaves every value
    unchanged, the property the reference's canonical-form shortcut uses."""
    rng = np.random.default_rng(78)
    n = 6
    p = H.spin_phys(OUT)
    mps = H.random_spin_mps(rng, n)
    cen = O.centralize0(mps)
    for ops in _expect_cases(n, rng):
        t = [H.dense_op_tensor(p, o) if o is not None else None for o in ops]
        a, b = O.expect(mps, t), O.expect(cen, t)
        assert abs(a - b) <= 1e-12 * abs(O.expect(mps, [None] * n)), (a, b)


def test_expect_reference_product_states():
    """test_finite_mps_measu.cc:43-170 (double cases): a 6-site chain of
    (N=0, N=1) sites in the product states DirectStateInitMps builds; <n_i>,
    <1 1> and <n_i n_j> with identity insertions, compared exactly."""
    phys = H.Leg(OUT, [((0,), 1), ((1,), 1)])
    ntot = H.dense_op_tensor(phys, np.diag([0.0, 1.0]))
    idt = H.dense_op_tensor(phys, np.eye(2))
    n = 6
    labs1 = [1] * n
    labs2 = [i % 2 for i in range(n)]
    sites_set = [(0, 1), (0, 2), (0, 5), (1, 2), (1, 3), (4, 5)]

    def two(mps, a_op, b_op, s):
        ops = [None] * n
        ops[s[0]], ops[s[1]] = a_op, b_op
        for k in range(s[0] + 1, s[1]):
            ops[k] = idt
        return O.expect(mps, ops)

    for labs in (labs1, labs2):
        mps = H.product_mps(phys, labs)
        for i in range(n):
            ops = [None] * n
            ops[i] = ntot
            assert O.expect(mps, ops) == labs[i]
    m1, m2 = H.product_mps(phys, labs1), H.product_mps(phys, labs2)
    assert [two(m1, idt, idt, s) for s in sites_set] == [1.0] * 6
    assert [two(m1, ntot, ntot, s) for s in sites_set] == [1.0] * 6
    assert [two(m2, idt, idt, s) for s in sites_set] == [1.0] * 6
    assert [two(m2, ntot, ntot, s) for s in sites_set] == [0, 0, 0, 0, 1, 0]


def test_oracle_truncate_mps_against_schmidt_values():
    """The oracle's TruncateMPS restatement on four spin-1/2 sites, where only
    the middle bond can exceed Dmax = 2: that cut keeps the two largest
    Schmidt values of the state across it (normalised), the result is
    normalised, and nothing changes when nothing is cut."""
    rng = np.random.default_rng(91)
    n = 4   # bonds 0 and 2 have rank <= 2, bond 1 up to 4; even: total charge 0
    mps = H.random_spin_mps(rng, n, maxdim=3)
    psi = H.mps_dense_state(mps)
    psi = psi / np.linalg.norm(psi)
    full, dims_full = H.oracle_truncate_mps(mps, 0.0, 1, 1000)
    phi = H.mps_dense_state(full)
    assert abs(abs(phi @ psi) - 1.0) < 1e-12
    cut, _ = H.oracle_truncate_mps(mps, 0.0, 1, 2)
    phi = H.mps_dense_state(cut)
    assert abs(np.linalg.norm(phi) - 1.0) < 1e-12
    s_psi = H.schmidt_values(psi, [2] * n, 1)
    s_phi = H.schmidt_values(phi, [2] * n, 1)
    top = s_psi[:2] / np.linalg.norm(s_psi[:2])
    assert np.allclose(s_phi[:2], top, atol=1e-12) and np.all(s_phi[2:] < 1e-12)


def test_u1xu1_contract_and_svd_against_dense():
    """U(1)xU(1) charges (qw = 2, the t-J / Hubbard sites): the oracle's
    Contract against numpy on the dense arrays, and its SVD against the dense
    SVD (reconstruction, orthonormal factors, singular values)."""
    rng = np.random.default_rng(13)
    site = Leg(OUT, [((0, 0), 1), ((1, 0), 1), ((0, 1), 1), ((1, 1), 1)])   # Hubbard basis
    l0 = Leg(IN, [((0, 0), 2), ((1, 0), 3), ((0, 1), 3), ((1, 1), 2), ((2, 1), 1)])
    r = Leg(OUT, [((-1, -1), 2), ((0, -1), 3), ((-1, 0), 3), ((0, 0), 4), ((-2, -1), 1), ((1, 0), 2), ((0, 1), 2)])
    psi = BTensor.random([l0, site, site, r], (0, 0), rng, qw=2)
    assert len(psi.blocks) > 20
    B = BTensor.random([r.inverse(), site, Leg(OUT, [((0, 0), 2), ((1, 1), 3), ((1, 0), 2)])], (0, 0), rng, qw=2)
    C = O.contract(psi, B, [[3], [0]])
    ref = np.einsum("abcd,def->abcef", psi.dense(), B.dense())
    assert np.abs(C.dense() - ref).max() < 1e-13
    U, S, Vt, err, D = O.svd(psi, 2, (0, 0), 0.0, 1000, 1000)
    rec = np.einsum("abk,kl,lcd->abcd", U.dense(), S.dense(), Vt.dense())
    assert np.abs(rec - psi.dense()).max() < 1e-13
    u = U.dense().reshape(-1, U.dense().shape[-1])
    assert np.abs(u.T @ u - np.eye(u.shape[1])).max() < 1e-13
    dense_s = np.linalg.svd(psi.dense().reshape(l0.dim * 4, -1), compute_uv=False)
    ours = np.sort(np.diag(S.dense()))[::-1]
    assert np.abs(ours - dense_s[:len(ours)]).max() < 1e-13


# ------------------------------------------------------------------ complex (GQTEN_Complex)

@pytest.mark.parametrize("axes,spec", [
    ([[2], [0]], "abc,cde->abde"),
    ([[0], [0]], "abc,ade->bcde"),
    ([[2, 1], [0, 1]], "abc,cbe->ae"),
    ([[0, 2], [2, 0]], "abc,cda->bd"),
])
def test_complex_contract_matches_dense_einsum(axes, spec):
    """Complex Contract against numpy einsum on the dense complex operands
    (the same four axis patterns as the real test), and a real operand with
    a complex one (promoted, as the MPS update's real S times a complex Vt)."""
    rng = np.random.default_rng(13)
    A, B = contract_case(rng, axes, BTensor.random_complex)
    legs = B.legs
    C = O.contract(A, B, axes)
    ref = np.einsum(spec, A.dense(), B.dense())
    assert len(C.blocks) >= 1 and np.abs(ref.imag).max() > 0.1
    assert C.is_complex and np.abs(C.dense() - ref).max() < 1e-13
    Br = BTensor.random(legs, (0,), rng)
    Cr = O.contract(A, Br, axes)
    assert np.abs(Cr.dense() - np.einsum(spec, A.dense(), Br.dense())).max() < 1e-13


def test_complex_svd_reconstructs_and_truncates():
    """Complex truncated SVD: U S Vt = psi, U^H U = 1, Vt Vt^H = 1, singular
    values of the QN-blocked matrix equal numpy's of the dense matrix, the
    discarded weight is the truncation error, trunc_err drives D."""
    rng = np.random.default_rng(15)
    l0 = H.bond(IN, (3, 2, 4), (0, 2, -2))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (3, 2, 1, 2), (0, 2, -2, 4))
    psi = BTensor.random_complex([l0, p, p, r], (0,), rng)
    U, S, Vt, err, D = O.svd(psi, 2, (0,), 0.0, 1000, 1000)
    assert U.is_complex and Vt.is_complex and not S.is_complex
    rec = np.einsum("abk,kl,lcd->abcd", U.dense(), S.dense(), Vt.dense())
    assert np.abs(rec - psi.dense()).max() < 1e-13
    u = U.dense().reshape(-1, U.dense().shape[-1])
    assert np.abs(u.conj().T @ u - np.eye(u.shape[1])).max() < 1e-13
    v = Vt.dense().reshape(Vt.dense().shape[0], -1)
    assert np.abs(v @ v.conj().T - np.eye(v.shape[0])).max() < 1e-13
    dense_s = np.linalg.svd(psi.dense().reshape(psi.dense().shape[0] * 2, -1), compute_uv=False)
    ours = np.sort(np.diag(S.dense()))[::-1]
    assert np.abs(ours - dense_s[:len(ours)]).max() < 1e-13
    U5, S5, V5, err5, D5 = O.svd(psi, 2, (0,), 0.0, 5, 5)
    assert D5 == 5
    assert abs(err5 - (dense_s[5:] ** 2).sum() / (dense_s ** 2).sum()) < 1e-13
    target = (dense_s[7:] ** 2).sum() / (dense_s ** 2).sum() * 1.0000001
    _, _, _, e7, D7 = O.svd(psi, 2, (0,), target, 1, 1000)
    assert D7 == 7 and e7 <= target
    # a wide sector (fewer rows than columns) takes the conjugate-transpose route
    Uw, Sw, Vw, _, _ = O.svd(psi, 1, (0,), 0.0, 1000, 1000)
    recw = np.einsum("ak,kl,lbcd->abcd", Uw.dense(), Sw.dense(), Vw.dense())
    assert np.abs(recw - psi.dense()).max() < 1e-13


def test_complex_lanczos_matches_dense_eigh():
    """Complex Hermitian H_eff (every factor Hermitian): H is Hermitian, the
    oracle Lanczos ground energy is real, not below the dense lowest
    eigenvalue, and equals the Rayleigh quotient of its ground state; the
    Hermiticity invariant <y|Hx> = conj <x|Hy> holds for the matvec."""
    rng = np.random.default_rng(17)
    L, W1, W2, R, psi = H.hermitian_cent_problem_c(rng)
    Hd = H.dense_heff_cent(L, W1, W2, R)
    assert np.abs(Hd - Hd.conj().T).max() < 1e-12 and np.abs(Hd.imag).max() > 0.1
    it, e, gs = O.lanczos(L, W1, W2, R, psi, "cent", 1e-12, 200)
    x = gs.dense().ravel()
    assert gs.is_complex and abs(np.linalg.norm(x) - 1) < 1e-10
    assert abs((x.conj() @ Hd @ x).real - e) < 1e-8
    # the ground state of the symmetry sector the start vector lies in
    w = np.linalg.eigvalsh(Hd)
    assert e >= w[0] - 1e-8
    y = BTensor.random_complex(psi.legs, (0,), rng)
    hx = O.heff_mul(L, W1, W2, R, psi, "cent").dense()
    hy = O.heff_mul(L, W1, W2, R, y, "cent").dense()
    lhs = np.vdot(y.dense(), hx)
    rhs = np.conj(np.vdot(psi.dense(), hy))
    assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs))


def test_complex_dag_conjugates_and_dot_is_real_part():
    rng = np.random.default_rng(19)
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (2, 3), (1, -1))
    a = BTensor.random_complex([p, r], (0,), rng)
    b = BTensor.random_complex([p, r], (0,), rng)
    A, B = O.OTen.from_bt(a), O.OTen.from_bt(b)
    d = O.OTen(O.lib().oracle_dag(A.h)).to_bt()
    assert np.abs(d.dense() - np.conj(a.dense())).max() == 0.0
    assert [l.dir for l in d.legs] == [1 - l.dir for l in a.legs]
    assert abs(O.lib().oracle_dot(A.h, B.h) - np.vdot(a.dense(), b.dense()).real) < 1e-14
