"""Lattice Hamiltonians used by the tests and the benchmark, built through
the MPOGenerator of the host layer (mps2_amd.gqmps2.Model).

Operator matrices use the gqten element convention op(i, j) (i = IN / ket
coordinate, j = OUT coordinate), and the local bases, operators and terms of
the reference's own tests (tests/test_algorithm/test_two_site_update_finite_vmps.cc:92-132
spin-1/2, :491-551 t-J, :792-885 Hubbard, :1017-1072 Kondo), so the golden
energies there apply unchanged.
"""
from __future__ import annotations

import numpy as np

from .gqmps2 import Model

# ------------------------------------------------------------------ spin-1/2
# basis 0 = Sz +1/2 (qn +1), 1 = Sz -1/2 (qn -1); test_two_site_update_finite_vmps.cc:94-124
SPIN_SITE = [((1,), 1), ((-1,), 1)]
SZ = np.diag([0.5, -0.5])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])   # sp({0,1}) = 1
SM = np.array([[0.0, 0.0], [1.0, 0.0]])   # sm({1,0}) = 1
ID2 = np.eye(2)


def heisenberg_terms(model: Model, pairs, J=1.0):
    for (i, j) in pairs:
        i, j = min(i, j), max(i, j)
        model.term(J, [SZ, SZ], [i, j])
        model.term(0.5 * J, [SP, SM], [i, j])
        model.term(0.5 * J, [SM, SP], [i, j])


def heisenberg_chain(N, J=1.0) -> Model:
    m = Model([SPIN_SITE] * N, qw=1)
    heisenberg_terms(m, [(i, i + 1) for i in range(N - 1)], J)
    m.bond_dims = m.gen_mpo()
    return m


def ising_chain(N) -> Model:
    m = Model([SPIN_SITE] * N, qw=1)
    for i in range(N - 1):
        m.term(1.0, [SZ, SZ], [i, i + 1])
    m.bond_dims = m.gen_mpo()
    return m


def heisenberg_2d_6() -> Model:
    """test_two_site_update_finite_vmps.cc:249-281 (6-site ladder)."""
    m = Model([SPIN_SITE] * 6, qw=1)
    heisenberg_terms(m, [(0, 1), (0, 2), (1, 3), (2, 3), (2, 4), (3, 5), (4, 5)])
    m.bond_dims = m.gen_mpo()
    return m


def kitaev_simple(Nx=4, Ny=2) -> Model:
    """test_two_site_update_finite_vmps.cc:310-345 (ZZ bonds of a honeycomb strip)."""
    N = Nx * Ny
    m = Model([SPIN_SITE] * N, qw=1)
    for x in range(Nx):
        for y in range(Ny):
            if x % 2 == 1:
                s0 = Ny * (x % Nx) + y % Ny
                s1 = Ny * ((x - 1) % Nx) + (y + 1) % Ny
                s0, s1 = min(s0, s1), max(s0, s1)
                m.term(1.0, [SZ, SZ], [s0, s1])
    m.bond_dims = m.gen_mpo()
    return m


def snake_index(x, y, W):
    return x * W + (y if x % 2 == 0 else W - 1 - y)


def j1j2_cylinder_bonds(Lx, W):
    """Nearest (J1) and next-nearest (J2) bonds of a width-W cylinder
    (periodic in y), sites in snake order."""
    j1, j2 = set(), set()
    for x in range(Lx):
        for y in range(W):
            s = snake_index(x, y, W)
            if W > 1:
                t = snake_index(x, (y + 1) % W, W)
                if t != s:
                    j1.add((min(s, t), max(s, t)))
            if x + 1 < Lx:
                t = snake_index(x + 1, y, W)
                j1.add((min(s, t), max(s, t)))
                if W > 1:
                    for dy in (1, -1):
                        t = snake_index(x + 1, (y + dy) % W, W)
                        j2.add((min(s, t), max(s, t)))
    return sorted(j1), sorted(j2)


def j1j2_cylinder(Lx, W, J1=1.0, J2=0.5) -> Model:
    j1, j2 = j1j2_cylinder_bonds(Lx, W)
    m = Model([SPIN_SITE] * (Lx * W), qw=1)
    heisenberg_terms(m, j1, J1)
    heisenberg_terms(m, j2, J2)
    m.bond_dims = m.gen_mpo()
    return m


# ------------------------------------------------------------------ t-J (U1 x U1: N, Sz)
# test_two_site_update_finite_vmps.cc:491-551; basis 0 = up, 1 = down, 2 = hole
TJ_SITE = [((1, 1), 1), ((1, -1), 1), ((0, 0), 1)]


def _m3(entries):
    a = np.zeros((3, 3))
    for (i, j), v in entries.items():
        a[i, j] = v
    return a


TJ = dict(
    f=_m3({(0, 0): -1, (1, 1): -1, (2, 2): 1}),
    sz=_m3({(0, 0): 0.5, (1, 1): -0.5}),
    sp=_m3({(1, 0): 1}),
    sm=_m3({(0, 1): 1}),
    cup=_m3({(0, 2): 1}),
    cdagup=_m3({(2, 0): 1}),
    cdn=_m3({(1, 2): 1}),
    cdagdn=_m3({(2, 1): 1}),
)


def tj_model(N, pairs, t=3.0, J=1.0) -> Model:
    m = Model([TJ_SITE] * N, qw=2)
    o = TJ
    for (i, j) in pairs:
        m.term(-t, [o["cdagup"], o["cup"]], [i, j], inst=[o["f"]])
        m.term(-t, [o["cdagdn"], o["cdn"]], [i, j], inst=[o["f"]])
        m.term(-t, [o["cup"], o["cdagup"]], [i, j], inst=[o["f"]])
        m.term(-t, [o["cdn"], o["cdagdn"]], [i, j], inst=[o["f"]])
        m.term(J, [o["sz"], o["sz"]], [i, j])
        m.term(J / 2, [o["sp"], o["sm"]], [i, j])
        m.term(J / 2, [o["sm"], o["sp"]], [i, j])
    m.bond_dims = m.gen_mpo()
    return m


# ------------------------------------------------------------------ Hubbard (U1 x U1: Nup, Ndn)
# test_two_site_update_finite_vmps.cc:792-885; basis 0 = empty, 1 = up, 2 = down, 3 = double
HUB_SITE = [((0, 0), 1), ((1, 0), 1), ((0, 1), 1), ((1, 1), 1)]


def _m4(entries):
    a = np.zeros((4, 4))
    for (i, j), v in entries.items():
        a[i, j] = v
    return a


HUB = dict(
    f=_m4({(0, 0): 1, (1, 1): -1, (2, 2): -1, (3, 3): 1}),
    nupdn=_m4({(3, 3): 1}),
    adagupf=_m4({(1, 0): 1, (3, 2): -1}),
    aup=_m4({(0, 1): 1, (2, 3): 1}),
    adagdn=_m4({(2, 0): 1, (3, 1): 1}),
    fadn=_m4({(0, 2): 1, (1, 3): -1}),
    naupf=_m4({(0, 1): 1, (2, 3): -1}),
    adagup=_m4({(1, 0): 1, (3, 2): 1}),
    nadn=_m4({(0, 2): -1, (1, 3): -1}),
    fadagdn=_m4({(2, 0): -1, (3, 1): 1}),
)


def hubbard_hop_terms(m: Model, s0, s1, tt):
    o = HUB
    m.term(1.0, [-tt * o["adagupf"], o["aup"]], [s0, s1], inst=[o["f"]])
    m.term(1.0, [-tt * o["adagdn"], o["fadn"]], [s0, s1], inst=[o["f"]])
    m.term(1.0, [o["naupf"], -tt * o["adagup"]], [s0, s1], inst=[o["f"]])
    m.term(1.0, [o["nadn"], -tt * o["fadagdn"]], [s0, s1], inst=[o["f"]])


def hubbard_2x2(t0=1.0, t1=0.5, U=2.0) -> Model:
    """test_two_site_update_finite_vmps.cc:888-935 (2x2 plaquette, t, t', U)."""
    Nx = Ny = 2
    m = Model([HUB_SITE] * (Nx * Ny), qw=2)
    idx = lambda x, y: x * Ny + y   # noqa: E731
    for i in range(Nx):
        for j in range(Ny):
            s0 = idx(i, j)
            m.term(U, [HUB["nupdn"]], [s0])
            if i != Nx - 1:
                hubbard_hop_terms(m, s0, idx(i + 1, j), t0)
            if j != Ny - 1:
                hubbard_hop_terms(m, s0, idx(i, j + 1), t0)
            if j != Ny - 1:
                if i != 0:
                    a, b = sorted((s0, idx(i - 1, j + 1)))
                    hubbard_hop_terms(m, a, b, t1)
                if i != Nx - 1:
                    a, b = sorted((s0, idx(i + 1, j + 1)))
                    hubbard_hop_terms(m, a, b, t1)
    m.bond_dims = m.gen_mpo()
    return m


def hubbard_cylinder(Lx, W, t=1.0, tp=0.0, U=8.0) -> Model:
    """Square-lattice Hubbard cylinder (periodic in y), snake order."""
    N = Lx * W
    m = Model([HUB_SITE] * N, qw=2)
    for s in range(N):
        m.term(U, [HUB["nupdn"]], [s])
    j1, j2 = j1j2_cylinder_bonds(Lx, W)
    for (a, b) in j1:
        hubbard_hop_terms(m, a, b, t)
    if tp != 0.0:
        for (a, b) in j2:
            hubbard_hop_terms(m, a, b, tp)
    m.bond_dims = m.gen_mpo()
    return m


# ------------------------------------------------------------------ Kondo (non-uniform sites)
# test_two_site_update_finite_vmps.cc:1017-1113: even sites extended (dim 4),
# odd sites localized (dim 2), trivial Sz sector
KONDO_E = [((0,), 4)]
KONDO_L = [((0,), 2)]


def _m(d, entries):
    a = np.zeros((d, d))
    for (i, j), v in entries.items():
        a[i, j] = v
    return a


def kondo_chain(Nx=4, t=0.25, Jk=1.0, Jz=0.5) -> Model:
    N = 2 * Nx
    m = Model([KONDO_E if i % 2 == 0 else KONDO_L for i in range(N)], qw=1)
    sz = _m(4, {(0, 0): 0.5, (1, 1): -0.5})
    sp = _m(4, {(0, 1): 1.0})
    sm = _m(4, {(1, 0): 1.0})
    bupcF = _m(4, {(2, 1): -1, (0, 3): 1})
    Fbdnc = _m(4, {(2, 0): 1, (1, 3): -1})
    bupaF = _m(4, {(1, 2): 1, (3, 0): -1})
    Fbdna = _m(4, {(0, 2): -1, (3, 1): 1})
    bupc = _m(4, {(2, 1): 1, (0, 3): 1})
    bdnc = _m(4, {(2, 0): 1, (1, 3): 1})
    bupa = _m(4, {(1, 2): 1, (3, 0): 1})
    bdna = _m(4, {(0, 2): 1, (3, 1): 1})
    Sz = _m(2, {(0, 0): 0.5, (1, 1): -0.5})
    Sp = _m(2, {(0, 1): 1.0})
    Sm = _m(2, {(1, 0): 1.0})
    for i in range(0, N - 2, 2):
        m.term(-t, [bupcF, bupa], [i, i + 2])
        m.term(-t, [bdnc, Fbdna], [i, i + 2])
        m.term(t, [bupaF, bupc], [i, i + 2])
        m.term(t, [bdna, Fbdnc], [i, i + 2])
        m.term(Jz, [Sz, Sz], [i + 1, i + 3])
    for i in range(0, N, 2):
        m.term(Jk, [sz, Sz], [i, i + 1])
        m.term(Jk / 2, [sp, Sm], [i, i + 1])
        m.term(Jk / 2, [sm, Sp], [i, i + 1])
    m.bond_dims = m.gen_mpo()
    return m


# ------------------------------------------------------------------ golden cases
# (builder, init dict, sweep params, E0, tolerance): the real-valued cases of
# test_two_site_update_finite_vmps.cc with their lines.
def golden_cases():
    N6 = 6
    return {
        "ising_1d_N6": dict(  # :135-151
            build=lambda: ising_chain(N6), init=dict(labs=[[i % 2 for i in range(N6)]]),
            sweeps=4, dmin=1, dmax=10, trunc=1e-5, lerr=1e-7, lmax=200, e0=-0.25 * (N6 - 1), tol=1e-10),
        "heisenberg_1d_N6": dict(  # :192-213
            build=lambda: heisenberg_chain(N6), init=dict(labs=[[i % 2 for i in range(N6)]]),
            sweeps=4, dmin=8, dmax=8, trunc=1e-9, lerr=1e-7, lmax=200, e0=-2.493577133888, tol=1e-12),
        "heisenberg_2d_6": dict(  # :249-281
            build=heisenberg_2d_6, init=dict(labs=[[i % 2 for i in range(N6)]]),
            sweeps=4, dmin=8, dmax=8, trunc=1e-9, lerr=1e-7, lmax=200, e0=-3.129385241572, tol=1e-12),
        "kitaev_simple_8": dict(  # :310-345
            build=kitaev_simple,
            init=dict(labs=[[i % 2 for i in range(8)], [(i + 1) % 2 for i in range(8)]], enlarged=2),
            sweeps=4, dmin=8, dmax=8, trunc=1e-4, lerr=1e-10, lmax=200, e0=-1.0, tol=1e-12),
        "tj_1d_N4": dict(  # :554-577
            build=lambda: tj_model(4, [(0, 1), (1, 2), (2, 3)]), init=dict(labs=[[2, 1, 2, 0]]),
            sweeps=11, dmin=8, dmax=8, trunc=1e-9, lerr=1e-8, lmax=20, e0=-6.947478526233, tol=1e-10),
        "tj_2d_N4": dict(  # :604-637
            build=lambda: tj_model(4, [(0, 1), (0, 2), (2, 3), (1, 3)]), init=dict(labs=[[2, 0, 1, 2]]),
            sweeps=10, dmin=8, dmax=8, trunc=1e-9, lerr=1e-8, lmax=20, e0=-8.868563739680, tol=1e-10),
        "hubbard_2x2": dict(  # :888-951
            build=hubbard_2x2, init=dict(labs=[[1 if i % 2 == 0 else 2 for i in range(4)]]),
            sweeps=10, dmin=16, dmax=16, trunc=1e-9, lerr=1e-8, lmax=20, e0=-2.828427124746, tol=1e-10),
        "kondo_8": dict(  # :1080-1113
            build=kondo_chain, init=dict(labs=[[i % 2 for i in range(8)]]),
            sweeps=5, dmin=64, dmax=64, trunc=1e-9, lerr=1e-8, lmax=20, e0=-3.180025784229132, tol=1e-10),
    }
