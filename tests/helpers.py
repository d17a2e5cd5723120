"""Random block-sparse test problems for the parity tests.

The leg conventions follow the reference: MPS A=(l IN, p OUT, r OUT),
L=(ket OUT, mpo OUT, bra IN), R=(ket IN, mpo IN, bra OUT),
W=(b_l IN, p_in IN, p_out OUT, b_r OUT), W_0=(p_in, b_r, p_out),
W_{N-1}=(p_in, b_l, p_out)  (two_site_update_finite_vmps_impl.h:117-129,
mpogen_impl.h:359,377,398; test_lanczos_solver.cc:121-168).
"""
import numpy as np

from mps2_amd.flat import BTensor, Leg, IN, OUT
from oracle import oracle_py as O


def bond(dir, dims, qns, qw=1):
    secs = []
    for q, d in zip(qns, dims):
        q = tuple(q) if isinstance(q, (tuple, list)) else (q,) + (0,) * (qw - 1)
        secs.append((q, d))
    return Leg(dir, secs)


def spin_phys(dir=OUT):
    return Leg(dir, [((1,), 1), ((-1,), 1)])


def sym_env(rng, ket: Leg, mpo: Leg, bra: Leg, scale=1.0):
    """Environment with zero-QN MPO sectors, symmetric in (ket, bra) per MPO index."""
    t = BTensor([ket, mpo, bra], 1)
    for c in t.allowed_coords((0,)):
        if c[0] != c[2]:
            continue
        b = rng.uniform(-1, 1, size=t.block_shape(c)) * scale
        t.blocks[c] = 0.5 * (b + b.transpose(2, 1, 0))
    return t


def sym_w(rng, bl: Leg, pin: Leg, pout: Leg, br: Leg):
    t = BTensor([bl, pin, pout, br], 1)
    for c in t.allowed_coords((0,)):
        if c[1] != c[2]:
            continue
        b = rng.uniform(-1, 1, size=t.block_shape(c))
        t.blocks[c] = 0.5 * (b + b.transpose(0, 2, 1, 3))
    return t


def hermitian_cent_problem(rng, dl=(3, 4, 2), dr=(2, 5, 3), qns=(-2, 0, 2), wdims=(2, 3)):
    """Hermitian H_eff = sum_b L_b (x) W1 (x) W2 (x) R_b with U(1) sectors on
    the virtual bonds and zero-QN MPO channels; psi has divergence 0."""
    l_leg = bond(IN, dl, qns)
    r_leg = bond(OUT, dr, qns)
    p = spin_phys(OUT)
    m1 = bond(OUT, wdims, (0,) * len(wdims))
    m2 = bond(OUT, wdims, (0,) * len(wdims))
    m3 = bond(OUT, wdims, (0,) * len(wdims))
    L = sym_env(rng, l_leg.inverse(), m1, l_leg)
    W1 = sym_w(rng, m1.inverse(), p.inverse(), p, m2)
    W2 = sym_w(rng, m2.inverse(), p.inverse(), p, m3)
    R = sym_env(rng, r_leg.inverse(), m3.inverse(), r_leg)
    psi = BTensor.random([l_leg, p, p, r_leg], (0,), rng)
    return L, W1, W2, R, psi


def random_cent_problem(rng, dl=(3, 4, 2), dr=(2, 5, 3), qns=(-2, 0, 2), wq=(0, 2, -2), wd=(2, 1, 1)):
    """Non-Hermitian random H_eff with QN-changing MPO channels (S+/S- like)."""
    l_leg = bond(IN, dl, qns)
    r_leg = bond(OUT, dr, qns)
    p = spin_phys(OUT)
    m1 = bond(OUT, wd, wq)
    m2 = bond(OUT, wd, wq)
    m3 = bond(OUT, wd, wq)
    L = BTensor.random([l_leg.inverse(), m1, l_leg], (0,), rng)
    W1 = BTensor.random([m1.inverse(), p.inverse(), p, m2], (0,), rng)
    W2 = BTensor.random([m2.inverse(), p.inverse(), p, m3], (0,), rng)
    R = BTensor.random([r_leg.inverse(), m3.inverse(), r_leg], (0,), rng)
    psi = BTensor.random([l_leg, p, p, r_leg], (0,), rng)
    return L, W1, W2, R, psi


def random_lend_problem(rng, dr=(2, 5, 3), qns=(-1, 1, 3), wq=(0, 2, -2), wd=(2, 1, 1)):
    r_leg = bond(OUT, dr, qns)
    p = spin_phys(OUT)
    m1 = bond(OUT, wd, wq)
    m2 = bond(OUT, wd, wq)
    W1 = BTensor.random([p.inverse(), m1, p], (0,), rng)
    W2 = BTensor.random([m1.inverse(), p.inverse(), p, m2], (0,), rng)
    R = BTensor.random([r_leg.inverse(), m2.inverse(), r_leg], (0,), rng)
    psi = BTensor.random([p, p, r_leg], (1,), rng)
    return None, W1, W2, R, psi


def random_rend_problem(rng, dl=(2, 5, 3), qns=(-1, 1, 3), wq=(0, 2, -2), wd=(2, 1, 1)):
    l_leg = bond(IN, dl, qns)
    p = spin_phys(OUT)
    m1 = bond(OUT, wd, wq)
    m2 = bond(OUT, wd, wq)
    L = BTensor.random([l_leg.inverse(), m1, l_leg], (0,), rng)
    W1 = BTensor.random([m1.inverse(), p.inverse(), p, m2], (0,), rng)
    W2 = BTensor.random([p.inverse(), m2.inverse(), p], (0,), rng)
    psi = BTensor.random([l_leg, p, p], (-1,), rng)
    return L, W1, W2, None, psi


def dense_heff_cent(L, W1, W2, R):
    """Dense H_eff matrix (rows: output (l', p1', p2', r'), cols: input)."""
    Ld, W1d, W2d, Rd = L.dense(), W1.dense(), W2.dense(), R.dense()
    # H[(lb,t1,t2,rb),(l,p1,p2,r)] = L[l,b,lb] W1[b,p1,t1,b'] W2[b',p2,t2,b''] R[r,b'',rb]
    H = np.einsum("abc,bdef,fghi,jik->cehkadgj", Ld, W1d, W2d, Rd)
    n = Ld.shape[2] * W1d.shape[2] * W2d.shape[2] * Rd.shape[2]
    m = Ld.shape[0] * W1d.shape[1] * W2d.shape[1] * Rd.shape[0]
    return H.reshape(n, m)


def max_rel_diff(a: BTensor, b: BTensor):
    da, db = a.dense(), b.dense()
    s = max(np.abs(db).max(), 1e-300)
    return float(np.abs(da - db).max() / s)


def rank_deficient_state():
    """psi = A B through a bond of total dimension 12 (sectors -3, -1, 1, 3 of
    dimensions 2, 4, 3, 3): 12 resolved singular values, the other 112 at
    rounding level."""
    rng = np.random.default_rng(5)
    l0 = bond(IN, (20, 30, 12), (0, 2, -2))
    p = spin_phys(OUT)
    k = bond(OUT, (3, 4, 3, 2), (1, -1, 3, -3))
    r = bond(OUT, (21, 31, 13, 2), (0, 2, -2, 4))
    A = BTensor.random([l0, p, k], (0,), rng)
    B = BTensor.random([k.inverse(), p, r], (0,), rng)
    return O.contract(A, B, [[2], [0]])


def mps_dense_state(mps):
    """Dense state vector of an MPS (site 0 = slowest index)."""
    psi = mps[0].dense()
    for t in mps[1:]:
        psi = np.tensordot(psi, t.dense(), axes=([-1], [0]))
    return psi.reshape(-1)


def dense_expect(psi, dims, ops):
    """<psi| prod_i ops[i] |psi> for dense ops in the op(i = IN, j = OUT)
    convention, i.e. (O psi)_j = sum_i op[i, j] psi_i; None = identity."""
    v = psi.reshape(dims)
    w = v
    for k, o in enumerate(ops):
        if o is not None:
            w = np.moveaxis(np.tensordot(o, w, axes=([0], [k])), 0, k)
    return float((v * w).sum())


def random_spin_mps(rng, n, maxdim=3, div=0):
    """Random (non-canonical, unnormalised) U(1) MPS of spin-1/2 sites; bond k
    carries every charge of the right parity within |q| <= k + 1."""
    p = spin_phys(OUT)
    bonds = []
    for k in range(n - 1):
        qs = list(range(-(k + 1), k + 2, 2))
        bonds.append(bond(OUT, [int(rng.integers(1, maxdim + 1)) for _ in qs], qs))
    mps = [BTensor.random([p, bonds[0]], (0,), rng)]
    for k in range(1, n - 1):
        mps.append(BTensor.random([bonds[k - 1].inverse(), p, bonds[k]], (0,), rng))
    mps.append(BTensor.random([bonds[-1].inverse(), p], (div,), rng))
    return mps


def product_mps(phys: Leg, labs):
    """DirectStateInitMps (finite_mps_init.h:216-262) in numpy: one-dimensional
    bonds whose charges follow the occupied sectors, every tensor = 1 at the
    labelled state."""
    n = len(labs)
    offs = phys.offsets()

    def sect(c):
        for s, o in enumerate(offs):
            if o <= c < o + phys.sectors[s][1]:
                return s
        raise ValueError(c)

    q = [phys.sectors[sect(c)][0] for c in labs]
    qw = len(q[0])
    tot = tuple(sum(x[k] for x in q) for k in range(qw))
    rq = tuple(tot[k] - q[0][k] for k in range(qw))
    legs_r = Leg(OUT, [(rq, 1)])
    mps = []
    t = BTensor([phys, legs_r], qw)
    t.blocks[(sect(labs[0]), 0)] = np.ones((1, 1))
    mps.append(t)
    for i in range(1, n - 1):
        l = legs_r.inverse()
        rq = tuple(l.sectors[0][0][k] - q[i][k] for k in range(qw))
        legs_r = Leg(OUT, [(rq, 1)])
        t = BTensor([l, phys, legs_r], qw)
        t.blocks[(0, sect(labs[i]), 0)] = np.ones((1, 1, 1))
        mps.append(t)
    t = BTensor([legs_r.inverse(), phys], qw)
    t.blocks[(0, sect(labs[-1]))] = np.ones((1, 1))
    mps.append(t)
    return mps


def dense_op_tensor(phys: Leg, dense):
    """Block tensor (phys IN, phys OUT) of a dense operator (zeros dropped)."""
    t = BTensor([phys.inverse(), phys], len(phys.sectors[0][0]))
    offs = phys.offsets()
    for a, (qa, da) in enumerate(phys.sectors):
        for b, (qb, db) in enumerate(phys.sectors):
            blk = np.asarray(dense)[offs[a]:offs[a] + da, offs[b]:offs[b] + db]
            if np.any(blk != 0):
                t.blocks[(a, b)] = np.array(blk, dtype=np.float64)
    return t


def oracle_truncate_mps(mps, trunc_err, dmin, dmax):
    """TruncateMPS (finite_mps.h:245-303) on the oracle: Centralize(0),
    normalise the first tensor, then per bond a truncated SVD (U carries the
    tensor's divergence), normalised S, S Vt into the next tensor.  Returns
    (mps, kept dimension per bond)."""
    mps = O.centralize0(list(mps))
    h = mps[0]
    n = h.norm()
    mps[0] = BTensor(h.legs, h.qw, {c: b / n for c, b in h.blocks.items()})
    dims = []
    for i in range(len(mps) - 1):
        A = mps[i]
        ldiv = A.block_div(next(iter(A.blocks)))
        u, s, vt, _, D = O.svd(A, 1 if i == 0 else 2, ldiv, trunc_err, dmin, dmax)
        sn = s.norm()
        s = BTensor(s.legs, s.qw, {c: b / sn for c, b in s.blocks.items()})
        mps[i] = u
        mps[i + 1] = O.contract(O.contract(s, vt, [[1], [0]]), mps[i + 1], [[1], [0]])
        dims.append(D)
    return mps, dims


def schmidt_values(psi, dims, cut):
    """Schmidt values of a dense state across the cut after site `cut`."""
    left = int(np.prod(dims[:cut + 1]))
    return np.linalg.svd(psi.reshape(left, -1), compute_uv=False)


def run_pair(model, sweeps, dmin, dmax, trunc, lerr, lmax=200, labs=None, enlarged=8, seed=11):
    """The GPU sweep (gqmps2 host layer, FP64) and the oracle sweep from the
    same right-canonical start: per-update records of both."""
    N = model.N
    labs = labs or [[i % 2 for i in range(N)], [(i + 1) % 2 for i in range(N)]]
    model.init_mps(labs, enlarged_dim=enlarged, seed=seed, centralize=True)
    mps0 = [model.mps(i) for i in range(N)]
    mpo = [model.mpo(i) for i in range(N)]
    g = model.dmrg(sweeps, dmin, dmax, trunc, lerr, lmax)
    o = O.dmrg(mps0, mpo, sweeps, dmin, dmax, trunc, lerr, lmax)
    return g, o


# ------------------------------------------------------------------ complex (GQTEN_Complex) problems

def herm_env_c(rng, ket: Leg, mpo: Leg, bra: Leg, scale=1.0):
    """Complex environment, Hermitian in (ket, bra) for every MPO index."""
    t = BTensor([ket, mpo, bra], 1)
    for c in t.allowed_coords((0,)):
        if c[0] != c[2]:
            continue
        shp = t.block_shape(c)
        b = (rng.uniform(-1, 1, size=shp) + 1j * rng.uniform(-1, 1, size=shp)) * scale
        t.blocks[c] = 0.5 * (b + np.conj(b.transpose(2, 1, 0)))
    return t


def herm_w_c(rng, bl: Leg, pin: Leg, pout: Leg, br: Leg):
    """Complex MPO tensor, Hermitian in (p_in, p_out) for every bond pair."""
    t = BTensor([bl, pin, pout, br], 1)
    for c in t.allowed_coords((0,)):
        if c[1] != c[2]:
            continue
        shp = t.block_shape(c)
        b = rng.uniform(-1, 1, size=shp) + 1j * rng.uniform(-1, 1, size=shp)
        t.blocks[c] = 0.5 * (b + np.conj(b.transpose(0, 2, 1, 3)))
    return t


def hermitian_cent_problem_c(rng, dl=(3, 4, 2), dr=(2, 5, 3), qns=(-2, 0, 2), wdims=(2, 3)):
    """The complex twin of hermitian_cent_problem: H_eff = sum L_b (x) W1 (x)
    W2 (x) R_b with every factor Hermitian in its (in, out) pair, so H_eff is
    Hermitian; psi is complex with divergence 0."""
    l_leg = bond(IN, dl, qns)
    r_leg = bond(OUT, dr, qns)
    p = spin_phys(OUT)
    m1 = bond(OUT, wdims, (0,) * len(wdims))
    m2 = bond(OUT, wdims, (0,) * len(wdims))
    m3 = bond(OUT, wdims, (0,) * len(wdims))
    L = herm_env_c(rng, l_leg.inverse(), m1, l_leg)
    W1 = herm_w_c(rng, m1.inverse(), p.inverse(), p, m2)
    W2 = herm_w_c(rng, m2.inverse(), p.inverse(), p, m3)
    R = herm_env_c(rng, r_leg.inverse(), m3.inverse(), r_leg)
    psi = BTensor.random_complex([l_leg, p, p, r_leg], (0,), rng)
    return L, W1, W2, R, psi
