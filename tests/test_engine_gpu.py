"""Parity of the HIP engine (through the C ABI) with the CPU FP64 oracle.

Bar: identical block sets (bit-exact bookkeeping) and FP64 values within
1e-12 relative (contractions reorder sums; no other difference is allowed).
"""
import numpy as np
import pytest

from mps2_amd import engine as E
from mps2_amd.flat import BTensor, IN, OUT
from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-12


def same_blocks(a: BTensor, b: BTensor):
    return sorted(a.blocks) == sorted(b.blocks) and [l.key() for l in a.legs] == [l.key() for l in b.legs]


@pytest.mark.parametrize("case", ["trail_lead", "lead_lead", "mid_perm", "trail_trail", "a_perm", "scalar"])
def test_contract_parity(ctx, case):
    rng = np.random.default_rng(sum(map(ord, case)))
    l0 = H.bond(IN, (17, 40, 9), (0, 2, -2))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (33, 70, 5), (1, -1, 3))
    A = BTensor.random([l0, p, r], (1,), rng)
    if case == "trail_lead":
        B = BTensor.random([r.inverse(), p, H.bond(OUT, (66, 130), (0, 2))], (0,), rng)
        axes = [[2], [0]]
    elif case == "lead_lead":
        B = BTensor.random([l0.inverse(), p, H.bond(OUT, (20, 31), (0, 2))], (0,), rng)
        axes = [[0], [0]]
    elif case == "mid_perm":
        B = BTensor.random([r.inverse(), H.bond(OUT, (2, 3), (1, 3)), p.inverse()], (0,), rng)
        axes = [[1, 2], [2, 0]]
    elif case == "trail_trail":
        B = BTensor.random([H.bond(OUT, (7, 19), (0, 2)), p.inverse(), r.inverse()], (0,), rng)
        axes = [[1, 2], [1, 2]]
    elif case == "a_perm":
        B = BTensor.random([p.inverse(), H.bond(OUT, (3, 8), (1, -1))], (0,), rng)
        axes = [[1], [0]]
    else:
        B = A.dag()
        axes = [[0, 1, 2], [0, 1, 2]]
    ref = O.contract(A, B, axes)
    dA, dB = ctx.put(A), ctx.put(B)
    got = E.contract(dA, dB, axes).get()
    assert same_blocks(got, ref)
    assert H.max_rel_diff(got, ref) < TOL


def test_contract_large_blocks(ctx):
    # blocks above the 128x128 tile threshold, uneven edges
    rng = np.random.default_rng(9)
    l0 = H.bond(IN, (300, 157), (0, 2))
    r = H.bond(OUT, (211, 129), (0, 2))
    A = BTensor.random([l0, r], (0,), rng)
    B = BTensor.random([r.inverse(), H.bond(OUT, (190, 77), (0, 2))], (0,), rng)
    ref = O.contract(A, B, [[1], [0]])
    got = E.contract(ctx.put(A), ctx.put(B), [[1], [0]]).get()
    assert same_blocks(got, ref)
    assert H.max_rel_diff(got, ref) < TOL


@pytest.mark.parametrize("where", ["cent", "lend", "rend"])
def test_heff_parity(ctx, where):
    rng = np.random.default_rng(21)
    prob = {"cent": H.random_cent_problem, "lend": H.random_lend_problem, "rend": H.random_rend_problem}[where](rng)
    L, W1, W2, R, psi = prob
    ref = O.heff_mul(L, W1, W2, R, psi, where)
    d = [ctx.put(x) if x is not None else None for x in prob]
    got = E.heff_mul(*d, where).get()
    # the device result lives on the full block layout; compare densely
    assert np.abs(got.dense() - ref.dense()).max() <= TOL * np.abs(ref.dense()).max()


def test_lanczos_parity(ctx):
    rng = np.random.default_rng(31)
    L, W1, W2, R, psi = H.hermitian_cent_problem(rng, dl=(6, 9, 5), dr=(4, 10, 6))
    it_o, e_o, gs_o = O.lanczos(L, W1, W2, R, psi, "cent", 1e-10, 200)
    d = [ctx.put(x) for x in (L, W1, W2, R, psi)]
    it_g, e_g, gs = E.lanczos(*d, "cent", 1e-10, 200, E.PREC_FP64)
    gs_g = gs.get()
    assert it_g == it_o
    assert abs(e_g - e_o) < 1e-12 * max(1, abs(e_o))
    ov = float((gs_g.dense() * gs_o.dense()).sum())
    assert abs(abs(ov) - 1) < 1e-10
    Hd = H.dense_heff_cent(L, W1, W2, R)
    x = gs_g.dense().ravel()
    assert abs(x @ Hd @ x - e_g) < 1e-8


def test_lanczos_max_iter_mode(ctx):
    rng = np.random.default_rng(33)
    L, W1, W2, R, psi = H.hermitian_cent_problem(rng)
    it_o, e_o, _ = O.lanczos(L, W1, W2, R, psi, "cent", 1e-16, 20)
    d = [ctx.put(x) for x in (L, W1, W2, R, psi)]
    it_g, e_g, _ = E.lanczos(*d, "cent", 1e-16, 20, E.PREC_FP64)
    assert it_g == it_o == 19
    assert abs(e_g - e_o) < 1e-11


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 200])
def test_tridiag_parity(ctx, n):
    rng = np.random.default_rng(n)
    a, b = rng.normal(size=n), rng.normal(size=n - 1)
    e_o, v_o = O.tridiag_lowest(a, b)
    e_g, v_g = ctx.tridiag_lowest(a, b)
    assert abs(e_g - e_o) < 1e-13 * max(1, abs(e_o))
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    assert np.linalg.norm(T @ v_g - e_g * v_g) < 1e-11
    assert abs(abs(float(v_g @ v_o)) - 1) < 1e-10


def graded_block_matrix(rng, dims, qns, decades, multiplet=True):
    """Block-diagonal 2-leg tensor whose sector matrices are U diag(s) V^T with
    s_j = 10^(-decades j / n): singular values spread over `decades` orders
    of magnitude, the sectors +q and -q carrying identical spectra (exact
    multiplets across sectors, as SU(2) symmetry produces)."""
    l = H.bond(IN, dims, qns)
    r = H.bond(OUT, dims, qns)
    t = BTensor([l, r])
    spectra = {}
    for i, (q, n) in enumerate(zip(qns, dims)):
        key = abs(q) if multiplet else q
        if key not in spectra:
            spectra[key] = 10.0 ** (-decades * np.arange(n) / n) * (1 + 0.3 * rng.random())
        s = spectra[key][:n]
        u, _ = np.linalg.qr(rng.standard_normal((n, n)))
        v, _ = np.linalg.qr(rng.standard_normal((n, n)))
        t.blocks[(i, i)] = (u * s) @ v.T
    return t


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
def test_svd_deep_cut_parity(ctx, orth):
    """Cut deep in a graded spectrum (s ~ 1e-8 s_max) with exact multiplets:
    the kept counts per sector must match the oracle's direct SVD, which
    needs singular values accurate far below the Gram matrix's resolution."""
    rng = np.random.default_rng(7 + orth)
    psi = graded_block_matrix(rng, (300, 420, 300), (-2, 0, 2), 12.0)
    dpsi = ctx.put(psi)
    for dmax in (520, 700):
        Uo, So, Vo, eo, Do = O.svd(psi, 1, (0,), 0.0, 1, dmax)
        u, s, vt, eg, Dg = E.svd(dpsi, 1, (0,), 0.0, 1, dmax, orth)
        Sg = s.get()
        assert Dg == Do
        assert u.get().legs[-1].key() == Uo.legs[-1].key()
        so, sg = np.diag(So.dense()), np.diag(Sg.dense())
        assert so.min() < 1e-6 * so.max()  # the cut lies deep in the spectrum
        assert np.abs(so - sg).max() < 1e-13 * so.max()
        np.testing.assert_allclose(sg, so, rtol=1e-8)


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
def test_svd_settled_cut_parity(ctx, orth):
    """Cut at the resolution floor (Dmax above the number of values over
    1e-11 s_max; spectra over 16 decades with exact +-q multiplets), a cut
    that needed a second refinement round before: the refinement now stops
    after one round once every value's place relative to the cut's tie window
    is settled within its error bound (2 eps (scale + 2 s s_max)), and the
    kept counts per sector, D and
    the kept values above the floor still match the oracle's direct SVD to
    within that bound, at most 12 eps s_max at the floor (asserted with the
    oracle's own error added)."""
    rng = np.random.default_rng(11 + orth)
    psi = graded_block_matrix(rng, (300, 420, 300), (-2, 0, 2), 16.0)
    dpsi = ctx.put(psi)
    before = E.ctx_counters(ctx.h, ("svd_refine_settled", "svd_refine_rounds"))
    Uo, So, Vo, eo, Do = O.svd(psi, 1, (0,), 0.0, 1, 900)
    u, s, vt, eg, Dg = E.svd(dpsi, 1, (0,), 0.0, 1, 900, orth)
    after = E.ctx_counters(ctx.h, ("svd_refine_settled", "svd_refine_rounds"))
    assert after["svd_refine_settled"] - before["svd_refine_settled"] == 1
    assert after["svd_refine_rounds"] - before["svd_refine_rounds"] == 1
    assert Dg == Do == 900
    assert u.get().legs[-1].key() == Uo.legs[-1].key()
    so, sg = np.diag(So.dense()), np.diag(s.get().dense())
    smax = so.max()
    assert np.sum(so > 1e-11 * smax) < 900  # the cut lies at the floor
    assert np.abs(so - sg).max() < 1e-13 * smax, np.abs(so - sg).max() / smax
    big = so > 1e-8 * smax
    np.testing.assert_allclose(sg[big], so[big], rtol=1e-8)
    # refined values between the floor and 1e-8 s_max: within 12 eps s_max
    # plus the oracle's own error
    res = so > 1e-11 * smax
    mid = res & ~big
    assert np.abs(so - sg)[mid].max() <= 1e-14 * smax, np.abs(so - sg)[mid].max() / smax
    # the kept values below it are tied by the truncation rule; after one
    # round they are resolved only to about sqrt(eps 1e-10) s_max = 2e-13 s_max
    assert np.all(sg[~res] < 2e-11 * smax)


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
@pytest.mark.parametrize("dims", [(3, 2, 4), (40, 70, 25), (150, 260, 120)])
def test_svd_parity(ctx, orth, dims):
    rng = np.random.default_rng(sum(dims) + orth)
    l0 = H.bond(IN, dims, (0, 2, -2))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, tuple(d + 1 for d in dims) + (2,), (0, 2, -2, 4))
    psi = BTensor.random([l0, p, p, r], (0,), rng)
    dpsi = ctx.put(psi)
    for dmax in (10_000, 7, max(3, sum(dims) // 2)):
        Uo, So, Vo, eo, Do = O.svd(psi, 2, (0,), 0.0, 1, dmax)
        u, s, vt, eg, Dg = E.svd(dpsi, 2, (0,), 0.0, 1, dmax, orth)
        Ug, Sg, Vg = u.get(), s.get(), vt.get()
        assert Dg == Do
        # bit-exact bookkeeping: identical new-bond sectors and kept counts
        assert Ug.legs[-1].key() == Uo.legs[-1].key()
        assert sorted(Sg.blocks) == sorted(So.blocks)
        so, sg = np.diag(So.dense()), np.diag(Sg.dense())
        assert np.abs(so - sg).max() < 1e-10 * so.max()
        assert abs(eg - eo) < 1e-10
        # the product the sweep uses is exact; the orthonormal side is orthonormal
        if orth == E.ORTH_LEFT:
            u2 = Ug.dense().reshape(-1, Ug.dense().shape[-1])
            assert np.abs(u2.T @ u2 - np.eye(u2.shape[1])).max() < 1e-12
            svt = np.einsum("kl,lcd->kcd", Sg.dense(), Vg.dense(), optimize=True)
            proj = np.einsum("abk,abcd->kcd", Ug.dense(), psi.dense(), optimize=True)
            assert np.abs(svt - proj).max() < 1e-11
        else:
            v2 = Vg.dense().reshape(Vg.dense().shape[0], -1)
            assert np.abs(v2 @ v2.T - np.eye(v2.shape[0])).max() < 1e-12
            us = np.einsum("abk,kl->abl", Ug.dense(), Sg.dense(), optimize=True)
            proj = np.einsum("abcd,lcd->abl", psi.dense(), Vg.dense(), optimize=True)
            assert np.abs(us - proj).max() < 1e-11
        if dmax == 10_000:
            rec = np.einsum("abk,kl,lcd->abcd", Ug.dense(), Sg.dense(), Vg.dense(), optimize=True)
            assert np.abs(rec - psi.dense()).max() < 1e-11


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
def test_svd_large_sector_known_spectrum(ctx, orth):
    """A 3300-wide sector beside sectors of 700 and 64, each with a known
    spectrum graded over 10 decades.  The eigensolver's first 228 columns of
    the large sector take its through-memory row path, the rest its one-,
    two- and three-rows-per-thread paths, across 412 deferred panel updates.
    The singular values match the constructed ones (1e-9 s_max above
    1e-2 s_max, 1e-6 s_max everywhere), U S Vt reproduces the matrix to
    2e-9 s_max (a direction whose Gram eigenvalue comes out <= 0, at
    rounding level, gets s = 0 and drops its component, of size s < 1e-8
    s_max here), and a cut at 2000 (singular value 1.3e-5 s_max, so the
    tail is refined) keeps per sector what the exact values keep (the gap at
    the cut is 0.2%)."""
    rng = np.random.default_rng(31 + orth)
    dims, qns = (3300, 700, 64), (-2, 0, 2)
    l, r = H.bond(IN, dims, qns), H.bond(OUT, dims, qns)
    psi = BTensor([l, r])
    exact = []
    for i, n in enumerate(dims):
        s = 10.0 ** (-10.0 * np.arange(n) / n) * (1.0 + 0.37 * i)
        u, _ = np.linalg.qr(rng.standard_normal((n, n)))
        v, _ = np.linalg.qr(rng.standard_normal((n, n)))
        psi.blocks[(i, i)] = (u * s) @ v.T
        exact += [(x, i) for x in s]
    exact.sort(key=lambda t: -t[0])
    smax = exact[0][0]
    dpsi = ctx.put(psi)
    u, s, vt, eg, Dg = E.svd(dpsi, 1, (0,), 0.0, 1, sum(dims), orth)
    Ug, Sg, Vg = u.get(), s.get(), vt.get()
    sg = np.sort(np.concatenate([np.diag(b) for b in Sg.blocks.values()]))[::-1]
    se = np.array([x for x, _ in exact])
    assert Dg == sum(dims) and sg.shape == se.shape
    big = se >= 1e-2 * smax
    assert np.abs(sg - se)[big].max() <= 1e-9 * smax, np.abs(sg - se)[big].max() / smax
    assert np.abs(sg - se).max() <= 1e-6 * smax, np.abs(sg - se).max() / smax
    rec = np.einsum("ak,kl,lb->ab", Ug.dense(), Sg.dense(), Vg.dense(), optimize=True)
    assert np.abs(rec - psi.dense()).max() < 2e-9 * smax, np.abs(rec - psi.dense()).max() / smax
    u, s, vt, eg, Dg = E.svd(dpsi, 1, (0,), 0.0, 1, 2000, orth)
    kept = {q: 0 for q in range(len(dims))}
    for _, q in exact[:2000]:
        kept[q] += 1
    assert Dg == 2000
    assert u.get().legs[-1].key() == H.bond(OUT, tuple(kept[q] for q in range(3)), qns).key()


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
def test_svd_rank_deficient_parity(ctx, orth):
    """psi = A B through a bond of total dimension 12: every sector matrix is
    numerically rank-deficient (112 of 124 singular values at rounding level,
    1e-17 relative).  Values below 1e-11 s_max have no reproducible order in
    any SVD; the rule keeps them all with trunc_err = 0 (a direct SVD returns
    them as tiny positive numbers) and breaks a cut among them by (sector,
    index).  The Gram path's +-noise eigenvalues there must give the
    oracle's kept counts per sector: all kept (Dmax 10000), a cut at the
    rank (12), and cuts inside the unresolved values (13, 16).  The product
    used by the sweep reconstructs psi and the orthonormal side is
    orthonormal."""
    psi = H.rank_deficient_state()
    dpsi = ctx.put(psi)
    for dmax in (10_000, 9, 12, 13, 16):
        Uo, So, Vo, eo, Do = O.svd(psi, 2, (0,), 0.0, 1, dmax)
        u, s, vt, eg, Dg = E.svd(dpsi, 2, (0,), 0.0, 1, dmax, orth)
        Ug, Sg, Vg = u.get(), s.get(), vt.get()
        assert Dg == Do, (dmax, Dg, Do)
        assert Ug.legs[-1].key() == Uo.legs[-1].key(), (dmax, Ug.legs[-1].key(), Uo.legs[-1].key())
        so, sg = np.sort(np.diag(So.dense())), np.sort(np.diag(Sg.dense()))
        big = so > 1e-6 * so.max()
        assert np.abs(so[big] - sg[big]).max() < 1e-10 * so.max()
        assert np.all(sg[~big] < 1e-6 * so.max())
        if orth == E.ORTH_LEFT:
            u2 = Ug.dense().reshape(-1, Ug.dense().shape[-1])
            assert np.abs(u2.T @ u2 - np.eye(u2.shape[1])).max() < 1e-12
        else:
            v2 = Vg.dense().reshape(Vg.dense().shape[0], -1)
            assert np.abs(v2 @ v2.T - np.eye(v2.shape[0])).max() < 1e-12
        if dmax >= 12:
            rec = np.einsum("abk,kl,lcd->abcd", Ug.dense(), Sg.dense(), Vg.dense(), optimize=True)
            assert np.abs(rec - psi.dense()).max() < 1e-11 * np.abs(psi.dense()).max()


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
def test_svd_sector_scales(ctx, orth):
    """Sectors of very different weight, as at the ends of a long chain where
    an Sz sector's Gram matrix peaks at 1e-126 of the largest: sector scales
    1e-63, 1 and 1e-30, each spectrum spread over 14 decades.  Each sector is
    solved at its own exact power-of-two scale, so the small sectors neither
    underflow nor lose relative accuracy: all values finite, kept counts per
    sector equal to the oracle's (everything kept, and a cut at 100), values
    above 1e-4 of their sector's largest within 1e-10 of it, and all values
    within 1e-7 of it (the Gram eigensolver's floor for values far below any
    cut, which no refinement round targets)."""
    rng = np.random.default_rng(91 + orth)
    dims, qns, scales = (60, 120, 80), (-2, 0, 2), (1e-63, 1.0, 1e-30)
    l, r = H.bond(IN, dims, qns), H.bond(OUT, dims, qns)
    psi = BTensor([l, r])
    for i, (n, sc) in enumerate(zip(dims, scales)):
        s = sc * 10.0 ** (-14.0 * np.arange(n) / n)
        u, _ = np.linalg.qr(rng.standard_normal((n, n)))
        v, _ = np.linalg.qr(rng.standard_normal((n, n)))
        psi.blocks[(i, i)] = (u * s) @ v.T
    dpsi = ctx.put(psi)
    for dmax in (sum(dims), 100):
        Uo, So, Vo, eo, Do = O.svd(psi, 1, (0,), 0.0, 1, dmax)
        u, s, vt, eg, Dg = E.svd(dpsi, 1, (0,), 0.0, 1, dmax, orth)
        Ug, Sg = u.get(), s.get()
        assert Dg == Do and Ug.legs[-1].key() == Uo.legs[-1].key(), (dmax, Ug.legs[-1].key(), Uo.legs[-1].key())
        for k, blk in So.blocks.items():
            so, sg = np.diag(blk), np.diag(Sg.blocks[k])
            assert np.all(np.isfinite(sg))
            big = so >= 1e-4 * so.max()
            assert np.abs(so - sg)[big].max() <= 1e-10 * so.max(), (k, np.abs(so - sg)[big].max() / so.max())
            assert np.abs(so - sg).max() <= 1e-7 * so.max(), (k, np.abs(so - sg).max() / so.max())


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
def test_svd_scale_invariance(ctx, orth):
    """Unnormalised inputs: a right-canonicalisation of a random 100-site
    start carries the product of the site norms (~1e80), whose Gram entries
    (1e160) and their squares overflow FP64 unless the SVD rescales.  The
    power-of-two scaling is exact: scaling the input by 2^300 leaves U and
    V^T bit-identical and multiplies S by exactly 2^300.  Inputs of scale
    1e150 and 1e-150 (not powers of two) agree with the oracle's SVD of the
    same tensor to FP64 accuracy; a NaN input raises instead of running."""
    rng = np.random.default_rng(77 + orth)
    l0 = H.bond(IN, (30, 50, 20), (0, 2, -2))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (31, 51, 21, 2), (0, 2, -2, 4))
    psi = BTensor.random([l0, p, p, r], (0,), rng)
    dmax = 40
    u1, s1, v1, e1, D1 = E.svd(ctx.put(psi), 2, (0,), 0.0, 1, dmax, orth)
    big = BTensor(psi.legs, psi.qw, {c: b * 2.0 ** 300 for c, b in psi.blocks.items()})
    u2, s2, v2, e2, D2 = E.svd(ctx.put(big), 2, (0,), 0.0, 1, dmax, orth)
    assert D2 == D1 and e2 == e1
    for a, b in ((u1.get(), u2.get()), (v1.get(), v2.get())):
        assert sorted(a.blocks) == sorted(b.blocks)
        for k in a.blocks:
            assert np.array_equal(a.blocks[k], b.blocks[k])
    S1, S2 = s1.get(), s2.get()
    for k in S1.blocks:
        assert np.array_equal(S1.blocks[k] * 2.0 ** 300, S2.blocks[k])
    for scale in (1e150, 1e-150):
        t = BTensor(psi.legs, psi.qw, {c: b * scale for c, b in psi.blocks.items()})
        Uo, So, Vo, eo, Do = O.svd(t, 2, (0,), 0.0, 1, dmax)
        u, s, vt, eg, Dg = E.svd(ctx.put(t), 2, (0,), 0.0, 1, dmax, orth)
        assert Dg == Do and u.get().legs[-1].key() == Uo.legs[-1].key()
        so, sg = np.diag(So.dense()), np.diag(s.get().dense())
        assert np.all(np.isfinite(sg))
        assert np.abs(so - sg).max() < 1e-10 * so.max()
        assert abs(eg - eo) < 1e-10
    bad = BTensor(psi.legs, psi.qw, {c: b.copy() for c, b in psi.blocks.items()})
    next(iter(bad.blocks.values())).flat[0] = np.nan
    with pytest.raises(E.EngineError, match="NaN"):
        E.svd(ctx.put(bad), 2, (0,), 0.0, 1, dmax, orth)


def test_linear_combine_and_normalize(ctx):
    rng = np.random.default_rng(41)
    l0 = H.bond(IN, (5, 6), (0, 2))
    p = H.spin_phys(OUT)
    x = BTensor.random([l0, p, l0.inverse()], (1,), rng)
    y = BTensor.random([l0, p, l0.inverse()], (1,), rng)
    dx, dy = ctx.put(x), ctx.put(y)
    out = dx.copy()
    E.linear_combine([0.5, -2.0], [dx, dy], 3.0, out)
    ref = 3.0 * x.dense() + 0.5 * x.dense() - 2.0 * y.dense()
    assert np.abs(out.get().dense() - ref).max() < 1e-13
    n = out.normalize()
    assert abs(n - np.linalg.norm(ref)) < 1e-12 * n
    assert abs(E.dot(out, out) - 1) < 1e-14


def low_prec_bound(prec, K):
    """Elementwise error bound of a low-precision product, as a multiple of
    sum_k |a_k||b_k|.  FP32 MFMA: 1e-5 (operands rounded to FP32, FP32
    accumulation).  Split-BF16x3: x = hi + lo with hi = bf16(x), lo =
    bf16(x - hi) leaves |x - hi - lo| <= 2^-16 |x|, and dropping lo_a lo_b
    (<= 2^-16 |ab|) gives at most 3 * 2^-16 per product (plus 2^-23 for the
    FP32 rounding of the operands), then at most (K + 2) 2^-24 from FP32
    accumulation."""
    if prec == E.PREC_MIXED:
        return 1e-5
    return 3.1 * 2.0 ** -16 + (K + 2) * 2.0 ** -24


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_MIXED, E.PREC_MIXED_BF16X3])
def test_grouped_gemm_shapes_vs_numpy(ctx, prec):
    """Single-block contractions over edge shapes and all four operand layouts
    against numpy (FP64: 1e-12 relative; FP32 MFMA and split-BF16x3: the
    elementwise bound of low_prec_bound relative to sum_k |a||b|)."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(5)

    def lg(d, dr):
        return Leg(dr, [((0,), d)])

    for (M, K, N) in [(1, 1, 1), (2, 3, 1), (3, 2, 5), (16, 4, 16), (17, 5, 3), (8, 2, 64), (64, 16, 64),
                      (65, 17, 65), (130, 40, 97), (300, 211, 190)]:
        for case in ("NN", "TN", "NT", "TT"):
            a = rng.uniform(-1, 1, (M, K))
            b = rng.uniform(-1, 1, (K, N))
            if case[0] == "N":
                A, ax = BTensor([lg(M, IN), lg(K, OUT)], 1, {(0, 0): a}), [1]
            else:
                A, ax = BTensor([lg(K, OUT), lg(M, IN)], 1, {(0, 0): a.T.copy()}), [0]
            if case[1] == "N":
                B, bx = BTensor([lg(K, IN), lg(N, OUT)], 1, {(0, 0): b}), [0]
            else:
                B, bx = BTensor([lg(N, OUT), lg(K, IN)], 1, {(0, 0): b.T.copy()}), [1]
            got = E.contract_ex(ctx.put(A), ctx.put(B), [ax, bx], prec).get().dense()
            ref = a @ b
            if prec == E.PREC_FP64:
                assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (M, K, N, case)
            else:
                bound = np.abs(a) @ np.abs(b)
                assert (np.abs(got - ref) <= low_prec_bound(prec, K) * bound + 1e-30).all(), (M, K, N, case)


@pytest.mark.parametrize("prec", [E.PREC_MIXED, E.PREC_MIXED_BF16X3])
def test_lanczos_mixed_matches_fp64(ctx, prec):
    """FP32-MFMA (or split-BF16x3) matvec + FP64 residual restarts reach the
    FP64 ground state: |dE|/|E| <= 1e-10 and state overlap >= 1 - 1e-9."""
    rng = np.random.default_rng(77)
    L, W1, W2, R, psi = H.hermitian_cent_problem(rng, dl=(9, 14, 7), dr=(8, 15, 6))
    d = [ctx.put(x) for x in (L, W1, W2, R, psi)]
    _, e64, g64 = E.lanczos(*d, "cent", 1e-12, 200, E.PREC_FP64)
    _, emx, gmx = E.lanczos(*d, "cent", 1e-12, 200, prec)
    assert abs(emx - e64) <= 1e-10 * abs(e64)
    ov = abs(float((g64.get().dense() * gmx.get().dense()).sum()))
    assert ov >= 1 - 1e-9


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_MIXED, E.PREC_MIXED_BF16X3])
def test_long_k_walks_over_several_pairs(ctx, prec):
    """Output blocks fed by long K walks over two block pairs (K = 800 and 750,
    two contracted legs, 97 K slabs per tile): values against numpy for
    leading and trailing contracted legs on both operands, and bit-identical
    on a second run (owner-computes, fixed order)."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(11)
    k1 = Leg(OUT, [((0,), 20), ((1,), 30)])
    k2 = Leg(OUT, [((0,), 25), ((1,), 40)])
    for (M, N) in [(200, 150), (64, 64), (129, 257)]:
        m = Leg(IN, [((0,), M)])
        n = Leg(OUT, [((0,), N)])
        A0 = BTensor.random([m, k1, k2], (1,), rng)
        B0 = BTensor.random([k1.inverse(), k2.inverse(), n], (-1,), rng)
        ref = np.tensordot(A0.dense(), B0.dense(), axes=([1, 2], [0, 1]))
        bound = np.tensordot(np.abs(A0.dense()), np.abs(B0.dense()), axes=([1, 2], [0, 1]))
        for case in ("NN", "TN", "NT", "TT"):
            if case[0] == "N":
                A, ax = A0, [1, 2]
            else:
                A = BTensor([k1, k2, m], 1, {(c[1], c[2], c[0]): b.transpose(1, 2, 0).copy() for c, b in A0.blocks.items()})
                ax = [0, 1]
            if case[1] == "N":
                B, bx = B0, [0, 1]
            else:
                B = BTensor([n, k1.inverse(), k2.inverse()], 1,
                            {(c[2], c[0], c[1]): b.transpose(2, 0, 1).copy() for c, b in B0.blocks.items()})
                bx = [1, 2]
            dA, dB = ctx.put(A), ctx.put(B)
            got = E.contract_ex(dA, dB, [ax, bx], prec).get()
            again = E.contract_ex(dA, dB, [ax, bx], prec).get()
            assert sorted(got.blocks) == [(0, 0)]
            assert np.array_equal(got.blocks[(0, 0)], again.blocks[(0, 0)]), (M, N, case)
            g = got.dense()
            if prec == E.PREC_FP64:
                assert np.abs(g - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (M, N, case)
            else:
                assert (np.abs(g - ref) <= low_prec_bound(prec, 1550) * bound + 1e-30).all(), (M, N, case)


def test_u1xu1_contract_and_svd_parity(ctx):
    """U(1)xU(1) charges (qw = 2): device Contract and SVD against the oracle
    on the same tensors (identical block sets, 1e-12 values, identical kept
    dimension and new-bond sectors)."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(13)
    site = Leg(OUT, [((0, 0), 1), ((1, 0), 1), ((0, 1), 1), ((1, 1), 1)])
    l0 = Leg(IN, [((0, 0), 12), ((1, 0), 23), ((0, 1), 23), ((1, 1), 12), ((2, 1), 7)])
    r = Leg(OUT, [((-1, -1), 12), ((0, -1), 23), ((-1, 0), 23), ((0, 0), 24), ((-2, -1), 7), ((1, 0), 9), ((0, 1), 8)])
    psi = BTensor.random([l0, site, site, r], (0, 0), rng, qw=2)
    B = BTensor.random([r.inverse(), site, Leg(OUT, [((0, 0), 9), ((1, 1), 13), ((1, 0), 5)])], (0, 0), rng, qw=2)
    assert len(psi.blocks) > 20 and B.blocks
    ref = O.contract(psi, B, [[3], [0]])
    got = E.contract(ctx.put(psi), ctx.put(B), [[3], [0]]).get()
    assert same_blocks(got, ref)
    assert H.max_rel_diff(got, ref) < TOL
    for dmax in (1000, 40):
        u, s, vt, err, D = O.svd(psi, 2, (0, 0), 0.0, 1, dmax)
        gu, gs, gvt, gerr, gD = E.svd(ctx.put(psi), 2, (0, 0), 0.0, 1, dmax, E.ORTH_LEFT)
        gs = gs.get()
        assert gD == D and gs.legs[0].key() == s.legs[0].key()
        assert dmax == 1000 or err > 0.0   # the second case cuts
        assert np.abs(np.sort(np.diag(gs.dense())) - np.sort(np.diag(s.dense()))).max() < 1e-12 * np.abs(s.dense()).max()
        assert abs(gerr - err) < 1e-12


def test_tensor_outlives_its_context():
    """A tensor freed after its context was destroyed returns its memory to
    HIP; the context is deleted with its last tensor (no use after free)."""
    rng = np.random.default_rng(5)
    A = BTensor.random([H.bond(IN, (5, 7), (0, 2)), H.bond(OUT, (5, 7), (0, 2))], (0,), rng)
    c = E.Context(0)
    d1, d2 = c.put(A), c.put(A)
    s = E.dot(d1, d2)
    c.close()
    assert abs(s - float((A.dense() ** 2).sum())) < 1e-12 * abs(s)
    got = d1.get()   # still readable: the context lives until its last tensor goes
    assert H.max_rel_diff(got, A) == 0.0
    d1.free()
    d2.free()
    c2 = E.Context(0)   # a fresh context works afterwards
    d3 = c2.put(A)
    assert H.max_rel_diff(d3.get(), A) == 0.0
    d3.free()
    c2.close()


@pytest.mark.parametrize("prec", [E.PREC_MIXED, E.PREC_MIXED_BF16X3])
def test_fp32_large_tile_launch_vs_numpy(ctx, prec):
    """A contraction large enough for the FP32 128x128 tile launch (16 output
    blocks of 500..512 square, at least one tile per compute unit), with
    one to two block pairs per output block and K extents that are not multiples of
    the slab depth (5 x 3 = 15, 6 x 5 = 30, ...): every block against the oracle to the
    low-precision bound (low_prec_bound) of sum_k |a||b|, and bit-identical on a second run."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(17)
    qs = list(range(16))
    m = Leg(IN, [((q,), 500 + (q * 7) % 13) for q in qs])
    n = Leg(OUT, [((q,), 512 - (q * 5) % 11) for q in qs])
    k1 = Leg(OUT, [((q,), 5 + q % 3) for q in range(0, 14, 2)])
    k2 = Leg(OUT, [((q,), 3 + 2 * q) for q in range(4)])
    A = BTensor.random([m, k1, k2], (0,), rng)
    B = BTensor.random([k1.inverse(), k2.inverse(), n], (0,), rng)
    dA, dB = ctx.put(A), ctx.put(B)
    got = E.contract_ex(dA, dB, [[1, 2], [0, 1]], prec).get()
    again = E.contract_ex(dA, dB, [[1, 2], [0, 1]], prec).get()
    ref = O.contract(A, B, [[1, 2], [0, 1]])
    assert sorted(got.blocks) == sorted(ref.blocks) and len(ref.blocks) == 16
    absA = BTensor(A.legs, A.qw, {c: np.abs(b) for c, b in A.blocks.items()})
    absB = BTensor(B.legs, B.qw, {c: np.abs(b) for c, b in B.blocks.items()})
    bound = O.contract(absA, absB, [[1, 2], [0, 1]])
    for c in ref.blocks:
        assert np.array_equal(got.blocks[c], again.blocks[c]), c
        assert (np.abs(got.blocks[c] - ref.blocks[c]) <= low_prec_bound(prec, 60) * bound.blocks[c] + 1e-30).all(), c


def test_split_k_launch_vs_oracle(ctx):
    """A contraction with few 128x128 tiles (4 output blocks of about 300
    square: 36 tiles) and long K (3 block pairs of about 1500 per output
    block) takes the split-K launches (four parts per tile, added in part
    order by the reduction) in all three arithmetic forms.  FP32: every
    element within the worst-case bound of FP32 accumulation, (K + 2) 2^-24
    sum_k |a||b| (operands rounded to FP32, K additions); split-BF16x3 within
    its bound (low_prec_bound); FP64 to the engine's FP64 bar; each
    bit-identical on a second run."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(41)
    m = Leg(IN, [((q,), 290 + 7 * q) for q in range(4)])
    n = Leg(OUT, [((q,), 301 - 5 * q) for q in range(4)])
    k1 = Leg(OUT, [((q,), 50 + q) for q in range(-3, 4)])
    k2 = Leg(OUT, [((q,), 28 + 2 * q) for q in range(3)])
    A = BTensor.random([m, k1, k2], (0,), rng)
    B = BTensor.random([k1.inverse(), k2.inverse(), n], (0,), rng)
    dA, dB = ctx.put(A), ctx.put(B)
    got = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_MIXED).get()
    again = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_MIXED).get()
    ref = O.contract(A, B, [[1, 2], [0, 1]])
    assert sorted(got.blocks) == sorted(ref.blocks) and len(ref.blocks) >= 3
    absA = BTensor(A.legs, A.qw, {c: np.abs(b) for c, b in A.blocks.items()})
    absB = BTensor(B.legs, B.qw, {c: np.abs(b) for c, b in B.blocks.items()})
    bound = O.contract(absA, absB, [[1, 2], [0, 1]])
    kmax = sum(d for _, d in k1.sectors) * sum(d for _, d in k2.sectors)
    for c in ref.blocks:
        assert np.array_equal(got.blocks[c], again.blocks[c]), c
        tol = (kmax + 2) * 2.0 ** -24
        assert (np.abs(got.blocks[c] - ref.blocks[c]) <= tol * bound.blocks[c] + 1e-30).all(), c
    bf3 = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_MIXED_BF16X3).get()
    bf3_again = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_MIXED_BF16X3).get()
    for c in ref.blocks:
        assert np.array_equal(bf3.blocks[c], bf3_again.blocks[c]), c
        tol = low_prec_bound(E.PREC_MIXED_BF16X3, kmax)
        assert (np.abs(bf3.blocks[c] - ref.blocks[c]) <= tol * bound.blocks[c] + 1e-30).all(), c
    # FP64: the same shape runs split-K on 64x64 tiles (100 tiles, fewer than
    # one round of the compute units); the engine's FP64 bar, bit-identical
    # on a second run
    got64 = E.contract(dA, dB, [[1, 2], [0, 1]]).get()
    again64 = E.contract(dA, dB, [[1, 2], [0, 1]]).get()
    assert same_blocks(got64, ref)
    assert H.max_rel_diff(got64, ref) < TOL
    for c in ref.blocks:
        assert np.array_equal(got64.blocks[c], again64.blocks[c]), c


def test_autotuned_launches_are_bit_identical():
    """The GEMM autotuner times its candidate kernels on the first launches of
    a large plan (>= 1e9 flops) and keeps the fastest; every candidate has
    the same per-element accumulation order, so all launches, before, during
    and after the decision, and with tuning off, give bit-identical results,
    and the decision is counted."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(23)
    qs = list(range(8))
    m = Leg(IN, [((q,), 500 + (q * 7) % 13) for q in qs])
    n = Leg(OUT, [((q,), 512 - (q * 5) % 11) for q in qs])
    k1 = Leg(OUT, [((q,), 60 + 5 * (q % 3)) for q in range(0, 8, 2)])
    k2 = Leg(OUT, [((q,), 20 + 3 * q) for q in range(3)])
    A = BTensor.random([m, k1, k2], (0,), rng)
    B = BTensor.random([k1.inverse(), k2.inverse(), n], (0,), rng)
    c = E.Context(0)
    try:
        dA, dB = c.put(A), c.put(B)
        for prec in (E.PREC_MIXED, E.PREC_MIXED_BF16X3):
            E.check(E.lib().mps2_ctx_set_tuning(c.h, 0))
            ref = E.contract_ex(dA, dB, [[1, 2], [0, 1]], prec).get()
            E.check(E.lib().mps2_ctx_set_tuning(c.h, 1))
            for it in range(12):
                got = E.contract_ex(dA, dB, [[1, 2], [0, 1]], prec).get()
                for cc in ref.blocks:
                    assert np.array_equal(got.blocks[cc], ref.blocks[cc]), (prec, it, cc)
        cnt = E.ctx_counters(c.h)
        assert cnt["tune_decisions"] >= 1 and cnt["tune_probes"] >= 4, cnt
        E.check(E.lib().mps2_ctx_set_tuning(c.h, 0))
        assert E.lib().mps2_ctx_set_tuning(c.h, 7) != 0
    finally:
        c.close()


def test_plan_found_by_layout_content_is_exact():
    """Contraction plans are cached by layout object and by layout content: a
    tensor uploaded anew with a layout seen before (new layout objects, new
    data) finds the earlier plan (counted as a content hit, no plan built),
    and the contraction through it equals the oracle and is bit-identical to
    the same data contracted through a freshly built plan.  The blocks are
    large enough for tiles of several blocks, edges and K steps, so the
    device-decoded tile order is exercised."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(31)
    m = Leg(IN, [((q,), 150 + 37 * q) for q in range(4)])
    k = Leg(OUT, [((q,), 90 + 11 * q) for q in range(4)])
    n = Leg(OUT, [((q,), 70 + 29 * q) for q in range(5)])
    A1 = BTensor.random([m, k], (0,), rng)
    B1 = BTensor.random([k.inverse(), n], (0,), rng)
    A2 = BTensor.random([m, k], (0,), rng)
    B2 = BTensor.random([k.inverse(), n], (0,), rng)
    c = E.Context(0)
    try:
        got1 = E.contract(c.put(A1), c.put(B1), [[1], [0]]).get()
        before = E.ctx_counters(c.h, ("plan_builds", "plan_content_hits"))
        got2 = E.contract(c.put(A2), c.put(B2), [[1], [0]]).get()
        after = E.ctx_counters(c.h, ("plan_builds", "plan_content_hits"))
        assert after["plan_builds"] == before["plan_builds"], (before, after)
        assert after["plan_content_hits"] == before["plan_content_hits"] + 1, (before, after)
        for got, (A, B) in ((got1, (A1, B1)), (got2, (A2, B2))):
            ref = O.contract(A, B, [[1], [0]])
            assert same_blocks(got, ref)
            assert H.max_rel_diff(got, ref) < TOL
    finally:
        c.close()
    c2 = E.Context(0)   # a new context starts with no plans: this one is built
    try:
        fresh = E.contract(c2.put(A2), c2.put(B2), [[1], [0]]).get()
        assert E.ctx_counters(c2.h, ("plan_builds",))["plan_builds"] >= 1
        for cc in fresh.blocks:
            assert np.array_equal(fresh.blocks[cc], got2.blocks[cc]), cc
    finally:
        c2.close()


def test_bf16x3_error_is_fp32_class(ctx):
    """Split-BF16x3 against FP32 MFMA on one large grouped launch (8 output
    blocks of about 500 x 500, K = 330): the relative Frobenius error against
    the FP64 product stays below 1e-5 (each product carries at most 3 * 2^-16
    and about 5e-6 root-mean-square, independent from product to product; a
    dropped cross term hi*lo alone would leave about 2e-3), FP32 MFMA below
    1e-6, and the kernel is deterministic."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(29)
    qs = list(range(8))
    m = Leg(IN, [((q,), 490 + (q * 7) % 13) for q in qs])
    n = Leg(OUT, [((q,), 500 - (q * 5) % 11) for q in qs])
    k1 = Leg(OUT, [((q,), 60 + 5 * (q % 3)) for q in range(0, 8, 2)])
    k2 = Leg(OUT, [((q,), 20 + 3 * q) for q in range(2)])
    A = BTensor.random([m, k1, k2], (0,), rng)
    B = BTensor.random([k1.inverse(), k2.inverse(), n], (0,), rng)
    dA, dB = ctx.put(A), ctx.put(B)
    ref = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_FP64).get()
    f32 = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_MIXED).get()
    bf3 = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_MIXED_BF16X3).get()
    bf3b = E.contract_ex(dA, dB, [[1, 2], [0, 1]], E.PREC_MIXED_BF16X3).get()
    num_f, num_b, den = 0.0, 0.0, 0.0
    for c, r in ref.blocks.items():
        assert np.array_equal(bf3.blocks[c], bf3b.blocks[c]), c
        num_f += float(((f32.blocks[c] - r) ** 2).sum())
        num_b += float(((bf3.blocks[c] - r) ** 2).sum())
        den += float((r ** 2).sum())
    rel_f, rel_b = np.sqrt(num_f / den), np.sqrt(num_b / den)
    print(f"relative Frobenius error: fp32 {rel_f:.3e}, split-bf16x3 {rel_b:.3e}")
    assert rel_b < 1e-5, (rel_f, rel_b)
    assert rel_f < 1e-6, (rel_f, rel_b)


def test_precision_code_checked(ctx):
    """An unknown precision code is an argument error, not a silent FP64 run."""
    rng = np.random.default_rng(31)
    L, W1, W2, R, psi = H.hermitian_cent_problem(rng, dl=(3, 4, 2), dr=(2, 4, 3))
    d = [ctx.put(x) for x in (L, W1, W2, R, psi)]
    with pytest.raises(E.EngineError):
        E.lanczos(*d, "cent", 1e-10, 50, 3)


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_MIXED, E.PREC_MIXED_BF16X3])
def test_narrow_outputs_vs_numpy(ctx, prec):
    """Contractions whose output blocks are at most 16 columns wide run on the
    vector-ALU kernel (one thread per output row): output blocks of 1..16
    columns and 1..1000 rows (across the 256-row workgroup edge), each fed by
    one or two block pairs over two contracted legs (K from 1 to 120), all four
    operand layouts, against numpy (FP64 1e-12; the FP32 forms to
    low_prec_bound), and bit-identical on a second run."""
    from mps2_amd.flat import Leg
    rng = np.random.default_rng(37)
    m = Leg(IN, [((0,), 1), ((1,), 255), ((2,), 257), ((3,), 1000)])
    k1 = Leg(OUT, [((0,), 1), ((1,), 7), ((2,), 40)])
    k2 = Leg(OUT, [((0,), 1), ((1,), 3)])
    for nd in (1, 3, 8, 13, 16):
        n = Leg(OUT, [((q,), max(1, nd - q)) for q in range(4)])
        A0 = BTensor.random([m, k1, k2], (0,), rng)
        B0 = BTensor.random([k1.inverse(), k2.inverse(), n], (0,), rng)
        ref = np.tensordot(A0.dense(), B0.dense(), axes=([1, 2], [0, 1]))
        bound = np.tensordot(np.abs(A0.dense()), np.abs(B0.dense()), axes=([1, 2], [0, 1]))
        for case in ("NN", "TN", "NT", "TT"):
            if case[0] == "N":
                A, ax = A0, [1, 2]
            else:
                A = BTensor([k1, k2, m], 1, {(c[1], c[2], c[0]): b.transpose(1, 2, 0).copy()
                                             for c, b in A0.blocks.items()})
                ax = [0, 1]
            if case[1] == "N":
                B, bx = B0, [0, 1]
            else:
                B = BTensor([n, k1.inverse(), k2.inverse()], 1,
                            {(c[2], c[0], c[1]): b.transpose(2, 0, 1).copy() for c, b in B0.blocks.items()})
                bx = [1, 2]
            dA, dB = ctx.put(A), ctx.put(B)
            got = E.contract_ex(dA, dB, [ax, bx], prec).get()
            again = E.contract_ex(dA, dB, [ax, bx], prec).get()
            for c in got.blocks:
                assert np.array_equal(got.blocks[c], again.blocks[c]), (nd, case, c)
            g = got.dense()
            if prec == E.PREC_FP64:
                assert np.abs(g - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (nd, case)
            else:
                assert (np.abs(g - ref) <= low_prec_bound(prec, 120) * bound + 1e-30).all(), (nd, case)


# ------------------------------------------------------------------ complex (GQTEN_Complex)

def _complex_cases(rng):
    """Complex contractions covering every operand layout of the device
    plan: contracted legs trailing or leading on A (A read in place or
    permuted to rows), on B (B expanded from either layout), in the middle
    (permutations), output blocks narrow (vector-ALU kernel) and wide
    (matrix cores), and a real operand with a complex one."""
    from tests.test_oracle import contract_case
    out = []
    for axes in ([[2], [0]], [[0], [0]], [[2, 1], [0, 1]], [[0, 2], [2, 0]]):
        A, B = contract_case(rng, axes, BTensor.random_complex)
        out.append((A, B, axes))
    # wide outputs (matrix-core tiles) with several pairs per block
    from mps2_amd.flat import Leg
    m = Leg(IN, [((0,), 70), ((1,), 130)])
    k1 = Leg(OUT, [((0,), 9), ((1,), 20)])
    k2 = Leg(OUT, [((0,), 5), ((1,), 7)])
    n = Leg(OUT, [((q,), 40 + 30 * q) for q in range(3)])
    A = BTensor.random_complex([m, k1, k2], (1,), rng)
    B = BTensor.random_complex([k1.inverse(), k2.inverse(), n], (0,), rng)
    out.append((A, B, [[1, 2], [0, 1]]))
    At = BTensor([k1, k2, m], 1, {(c[1], c[2], c[0]): b.transpose(1, 2, 0).copy() for c, b in A.blocks.items()})
    out.append((At, B, [[0, 1], [0, 1]]))
    Bt = BTensor([n, k1.inverse(), k2.inverse()], 1,
                 {(c[2], c[0], c[1]): b.transpose(2, 0, 1).copy() for c, b in B.blocks.items()})
    out.append((A, Bt, [[1, 2], [1, 2]]))
    # real times complex (the MPS update's S Vt) and complex times real
    out.append((BTensor.random(A.legs, (1,), rng), B, [[1, 2], [0, 1]]))
    out.append((A, BTensor.random(B.legs, (0,), rng), [[1, 2], [0, 1]]))
    return out


def test_complex_contract_parity(ctx):
    """Complex Contract on the device against the oracle's complex Contract
    (itself pinned by numpy einsum): identical output block sets, values to
    1e-12 relative, for every operand layout of the device plan."""
    rng = np.random.default_rng(41)
    for i, (A, B, axes) in enumerate(_complex_cases(rng)):
        ref = O.contract(A, B, axes)
        got = E.contract(ctx.put(A), ctx.put(B), axes).get()
        assert got.is_complex, i
        assert sorted(got.blocks) == sorted(ref.blocks), i
        assert H.max_rel_diff(got, ref) < 1e-12, (i, H.max_rel_diff(got, ref))


def test_complex_heff_and_lanczos_parity(ctx):
    """Complex Hermitian effective Hamiltonian (every factor Hermitian): the
    device matvec against the oracle's to 1e-12, and the device Lanczos
    (MPS2_PREC_FP64) against the oracle's: same iteration count, energy to
    1e-12 relative, ground states overlapping to 1 - 1e-10 in modulus; the
    mixed precisions reject complex problems."""
    rng = np.random.default_rng(43)
    L, W1, W2, R, psi = H.hermitian_cent_problem_c(rng, dl=(5, 7, 4), dr=(4, 8, 5))
    d = [ctx.put(x) for x in (L, W1, W2, R, psi)]
    hp = E.heff_mul(*d, "cent").get()
    ho = O.heff_mul(L, W1, W2, R, psi, "cent")
    assert hp.is_complex and H.max_rel_diff(hp, ho) < 1e-12
    it_o, e_o, gs_o = O.lanczos(L, W1, W2, R, psi, "cent", 1e-12, 200)
    it_g, e_g, gs_g = E.lanczos(*d, "cent", 1e-12, 200, E.PREC_FP64)
    assert it_g == it_o, (it_g, it_o)
    assert abs(e_g - e_o) <= 1e-12 * abs(e_o), (e_g, e_o)
    ov = np.vdot(gs_o.dense(), gs_g.get().dense())
    assert abs(abs(ov) - 1.0) < 1e-10, ov
    for prec in (E.PREC_MIXED, E.PREC_MIXED_BF16X3):
        with pytest.raises(E.EngineError):
            E.lanczos(*d, "cent", 1e-12, 200, prec)


@pytest.mark.parametrize("orth", [E.ORTH_LEFT, E.ORTH_RIGHT])
def test_complex_svd_parity(ctx, orth):
    """Complex truncated SVD on the device (one-sided Jacobi per sector)
    against the oracle's complex SVD: identical new-bond sectors and kept
    counts, singular values to 1e-12 of the largest, truncation error to
    1e-12; the factors reconstruct the kept part, the orthonormal side is
    unitary to 1e-12; untruncated, U S Vt = psi."""
    rng = np.random.default_rng(47)
    l0 = H.bond(IN, (7, 12, 9), (0, 2, -2))
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (8, 11, 6, 5), (0, 2, -2, 4))
    psi = BTensor.random_complex([l0, p, p, r], (0,), rng)
    dpsi = ctx.put(psi)
    for dmax in (1000, 25, 9):
        Uo, So, Vo, eo, Do = O.svd(psi, 2, (0,), 0.0, 1, dmax)
        u, s, vt, eg, Dg = E.svd(dpsi, 2, (0,), 0.0, 1, dmax, orth)
        Ug, Sg, Vg = u.get(), s.get(), vt.get()
        assert Dg == Do and Ug.legs[-1].key() == Uo.legs[-1].key(), (dmax, Dg, Do)
        so, sg = np.diag(So.dense()), np.diag(Sg.dense())
        assert np.abs(so - sg).max() < 1e-12 * so.max(), dmax
        assert abs(eg - eo) < 1e-12, dmax
        ud = Ug.dense().reshape(-1, Dg)
        vd = Vg.dense().reshape(Dg, -1)
        if orth == E.ORTH_LEFT:
            assert np.abs(ud.conj().T @ ud - np.eye(Dg)).max() < 1e-12
        else:
            assert np.abs(vd @ vd.conj().T - np.eye(Dg)).max() < 1e-12
        rec = np.einsum("abk,kl,lcd->abcd", Ug.dense(), Sg.dense(), Vg.dense())
        ref = np.einsum("abk,kl,lcd->abcd", Uo.dense(), So.dense(), Vo.dense())
        assert np.abs(rec - ref).max() < 1e-11 * np.abs(ref).max(), dmax
        if dmax == 1000:
            assert np.abs(rec - psi.dense()).max() < 1e-12 * np.abs(psi.dense()).max()


def test_complex_dag_and_dot(ctx):
    """Dag of a complex tensor conjugates its data and inverts its indexes;
    mps2_dot returns Re <a|b>; Normalize uses the complex 2-norm."""
    rng = np.random.default_rng(53)
    p = H.spin_phys(OUT)
    r = H.bond(OUT, (3, 4), (1, -1))
    a = BTensor.random_complex([p, r], (0,), rng)
    b = BTensor.random_complex([p, r], (0,), rng)
    da, db = ctx.put(a), ctx.put(b)
    g = da.dag().get()
    assert np.array_equal(g.dense(), np.conj(a.dense()))
    assert [l.dir for l in g.legs] == [1 - l.dir for l in a.legs]
    assert abs(E.dot(da, db) - np.vdot(a.dense(), b.dense()).real) < 1e-13
    n = da.normalize()
    assert abs(n - np.linalg.norm(a.dense())) < 1e-13
    assert abs(np.linalg.norm(da.get().dense()) - 1.0) < 1e-14
