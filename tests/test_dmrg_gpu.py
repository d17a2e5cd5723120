"""Two-site DMRG on the GPU (gqmps2 host layer -> C ABI -> HIP engine).

Pins: the reference's golden ground energies (test_two_site_update_finite_vmps.cc),
and per-update parity with the CPU FP64 oracle sweep started from the same
right-canonical MPS (energies to 1e-10 relative, identical kept bond
dimensions, identical Lanczos iteration counts).
"""
import os

import numpy as np
import pytest

from mps2_amd import hamiltonians as HM
from tests.helpers import run_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["heisenberg_1d_N6", "heisenberg_2d_6", "ising_1d_N6", "tj_1d_N4", "tj_2d_N4",
                                  "hubbard_2x2", "kitaev_simple_8", "kondo_8"])
def test_golden_energy_on_gpu(name):
    case = HM.golden_cases()[name]
    m = case["build"]()
    init = case["init"]
    m.init_mps(init["labs"], enlarged_dim=init.get("enlarged", 0), seed=7, centralize=True)
    r = m.dmrg(case["sweeps"], case["dmin"], case["dmax"], case["trunc"], case["lerr"], case["lmax"])
    assert abs(r["energy"] - case["e0"]) < case["tol"], (r["energy"], case["e0"])


def test_heisenberg_chain_per_update_parity_exact_bond():
    """Dmax above every bond's rank: the GPU sweep and the oracle sweep take
    the same path, update by update.  trunc_err = 1e-10 drops the numerically
    zero and near-zero singular values (s^2 < 1e-10), whose vectors are
    arbitrary (or, through a Gram matrix, resolved only to ~eps/s^2)."""
    m = HM.heisenberg_chain(12)
    g, o = run_pair(m, 2, 1, 64, 1e-10, 1e-10)
    e_g, e_o = g["energies"], o["energies"]
    assert np.abs(e_g - e_o).max() < 1e-10 * np.abs(e_o).max()
    assert (g["dims"] == o["dims"]).all()
    assert (g["iters"] == o["iters"]).all()


def test_heisenberg_chain_truncated_sweep_parity():
    """Binding Dmax (24 on L=16): kept dimensions identical at every update.
    In the first sweep the state is far from converged and the two Lanczos
    solvers (1e-10 energy tolerance) return vectors that differ at about the
    square root of that tolerance; the truncation carries the difference into
    the next environments, so mid-sweep energies agree to 1.4e-7 relative
    (measured; bar 1e-6), and the converged energies to 1.3e-13 (bar 1e-11)."""
    m = HM.heisenberg_chain(16)
    g, o = run_pair(m, 2, 1, 24, 0.0, 1e-10)
    e_g, e_o = g["energies"], o["energies"]
    assert (g["dims"] == o["dims"]).all()
    rel = np.abs(e_g - e_o) / np.abs(e_o)
    assert rel.max() < 1e-6, rel.max()
    assert abs(g["energy"] - o["energy"]) < 1e-11 * abs(o["energy"])


def test_j1j2_small_cylinder_per_update_parity():
    m = HM.j1j2_cylinder(4, 2, 1.0, 0.5)
    g, o = run_pair(m, 2, 1, 32, 1e-12, 1e-10)
    assert np.abs(g["energies"] - o["energies"]).max() < 1e-10 * np.abs(o["energies"]).max()
    assert (g["dims"] == o["dims"]).all()


def test_heisenberg_l16_config1_matches_ed():
    """Heisenberg OBC L=16, Dmax=64, 4 sweeps, Lanczos tol 1e-10 against the
    exact ground energy -6.911737145574 (ED, Sz=0, tests/golden/ed_energies.json)."""
    import json
    import os
    ed = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))
    m = HM.heisenberg_chain(16)
    m.init_mps([[i % 2 for i in range(16)], [(i + 1) % 2 for i in range(16)]], enlarged_dim=4, seed=0x5EED)
    r = m.dmrg(4, 1, 64, 0.0, 1e-10, 200)
    assert abs(r["energy"] - ed["heisenberg_obc_L16_sz0"]) < 1e-8


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_sweep_matches_fp64(precision):
    """Mixed-precision sweeps (1: FP32-MFMA matvec, 2: split-BF16x3) against
    FP64 sweeps from one checkpoint (L=20, Dmax 48, Lanczos tolerance 1e-11):
    every update's energy within 1e-10 relative (measured 1.2e-13 for both),
    the final within 1e-12, kept dimensions identical at every update."""
    m = HM.heisenberg_chain(20)
    m.init_mps([[i % 2 for i in range(20)], [(i + 1) % 2 for i in range(20)]], enlarged_dim=8, seed=3)
    m.dmrg(2, 1, 32, 0.0, 1e-10, 200)
    m.checkpoint()
    g64 = m.dmrg(2, 1, 48, 0.0, 1e-11, 200)
    m.restore()
    gmx = m.dmrg(2, 1, 48, 0.0, 1e-11, 200, precision=precision)
    rel = np.abs(gmx["energies"] - g64["energies"]) / np.abs(g64["energies"])
    assert rel.max() < 1e-10, rel.max()
    assert abs(gmx["energy"] - g64["energy"]) <= 1e-12 * abs(g64["energy"])
    assert np.array_equal(gmx["dims"], g64["dims"])


def test_heisenberg_l100_config2_ramp():
    """Heisenberg OBC L=100, U(1) Sz=0, D ramped 256,256,512,512,768,768,1024,1024
    over 8 sweeps (FP64 on the GPU), from a seeded random MPS.  No reference
    output exists at this size, so the checks are what the physics fixes:
    the energy is variational (non-increasing over the last sweeps, within the
    Lanczos tolerance), D reaches 1024, and the converged energy agrees with
    the DMRG value of this chain, E0 = -44.12773924, to 1e-5.

    Then mixed precision against FP64, one sweep each from the same state.  At
    Dmax 1024 this chain's cuts fall among Schmidt values of 1e-11 .. 1e-20
    s_max, below what a state converged to a 1e-10 energy tolerance fixes, so
    the per-sector split of those bonds (and the count of the next bond, which
    follows from it) may differ between the two runs, as between any two
    solvers; energies must agree per update to 1e-9 and finally to 1e-10, and
    kept counts may differ only at updates discarding less than 1e-16 (values
    below 1e-11 s_max enter the reported weight as 1e-22 each).  At Dmax 256
    the cuts are resolved and the kept counts must be identical."""
    N = 100
    m = HM.heisenberg_chain(N)
    m.init_mps([[i % 2 for i in range(N)], [(i + 1) % 2 for i in range(N)]], enlarged_dim=8, seed=0x5EED)
    sched = [256, 256, 512, 512, 768, 768, 1024, 1024]
    r = m.dmrg(len(sched), 1, 1024, 0.0, 1e-10, 200, dmax_schedule=sched, verbose=True)
    per = 2 * (N - 1)
    e_sweep = r["energies"].reshape(len(sched), per)[:, -1]
    dmax_sweep = r["dims"].reshape(len(sched), per).max(axis=1)
    info = f"energies per sweep {e_sweep.tolist()}, max D per sweep {dmax_sweep.tolist()}"
    assert np.all(np.diff(e_sweep[-4:]) <= 1e-9), info
    assert dmax_sweep[-1] == 1024, info
    assert abs(r["energy"] - (-44.12773924)) < 1e-5, info
    m.checkpoint()
    for dmax in (1024, 256):
        m.restore()
        g64 = m.dmrg(1, 1, dmax, 0.0, 1e-10, 200, verbose=True)
        m.restore()
        gmx = m.dmrg(1, 1, dmax, 0.0, 1e-10, 200, precision=1, verbose=True)
        rel = np.abs(gmx["energies"] - g64["energies"]) / np.abs(g64["energies"])
        assert rel.max() < 1e-9, (dmax, rel.max())
        assert abs(gmx["energy"] - g64["energy"]) <= 1e-10 * abs(g64["energy"])
        diff = np.nonzero(gmx["dims"] != g64["dims"])[0]
        msg = (dmax, diff.tolist(), g64["dims"][diff].tolist(), gmx["dims"][diff].tolist(),
               g64["trunc_errs"][diff].tolist())
        if dmax == 256:
            assert diff.size == 0, msg
        else:
            # an update whose count differs keeps everything or cuts among
            # values the converged state does not fix
            assert np.all(np.maximum(g64["trunc_errs"][diff], gmx["trunc_errs"][diff]) < 1e-16), msg


@pytest.mark.parametrize("tp,U", [(0.0, 8.0), (-0.2, 10.0)])
def test_hubbard_cylinder_mixed_matches_fp64(tp, U):
    """Fermi-Hubbard on a 4x4 cylinder (periodic in y, snake order), t = 1,
    t' = tp, U, 14 electrons (filling 0.875), U(1)xU(1) (N_up, N_dn) blocks
    of local dimension 4: the reduced form of the Hubbard cylinder workloads.
    Even at D = 512 this system is strongly truncated (discarded weight ~1e-4
    per cut, the energy varying by ~2e-3 within a sweep), a regime where the
    reported energies (taken before each truncation) need not fall sweep by
    sweep, and where a sweep's trajectory amplifies the two solvers'
    eigenvector-level differences: at D = 256 the t' case differs by 3.5e-10,
    at D = 512 by < 1e-10.  So: an FP64 ramp 64 -> 512 lowers the energy,
    then one FP64 and one mixed sweep at D = 512 from the same state (Lanczos
    tolerance 1e-12) agree at every update to 1e-9 relative, finally to
    1e-10, with identical kept dimensions."""
    N = 16
    m = HM.hubbard_cylinder(4, 4, t=1.0, tp=tp, U=U)
    up_dn = [1 if i % 2 == 0 else 2 for i in range(14)]
    m.init_mps([up_dn + [0, 0], [0, 0] + up_dn[::-1]], enlarged_dim=4, seed=11)
    sched = [64, 128, 256, 512, 512, 512, 512]
    r = m.dmrg(len(sched), 1, 512, 0.0, 1e-10, 200, dmax_schedule=sched)
    e_sweep = r["energies"].reshape(len(sched), 2 * (N - 1))[:, -1]
    assert e_sweep[-1] < e_sweep[0] - 1.0 and e_sweep[-1] < e_sweep[2], e_sweep
    m.checkpoint()
    g64 = m.dmrg(1, 1, 512, 0.0, 1e-12, 200)
    m.restore()
    gmx = m.dmrg(1, 1, 512, 0.0, 1e-12, 200, precision=1)
    rel = np.abs(gmx["energies"] - g64["energies"]) / np.abs(g64["energies"])
    assert rel.max() < 1e-9, rel.max()
    assert abs(gmx["energy"] - g64["energy"]) <= 1e-10 * abs(g64["energy"])
    assert np.array_equal(gmx["dims"], g64["dims"]), (g64["dims"], gmx["dims"])
    assert g64["dims"].max() == 512


@pytest.mark.gpu
def test_cpp_spin_chain_cases_with_resume(tmp_path):
    """test_two_site_update_finite_vmps.cc:91-222 as a C++ program over the
    drop-in headers: Ising E0 and measurements of the loaded result, Heisenberg
    E0 = -2.493577133888 to 1e-12, and the same energy when a second run
    continues from the MPS files."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                       "test_two_site_update_finite_vmps")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout


@pytest.mark.gpu
def test_cpp_lanczos_solver_cases(tmp_path):
    """test_lanczos_solver.cc:42-340 as a C++ program over the drop-in headers:
    centre-bond LanczosSolver to an error target and in the maximum-iteration
    mode, and the left end with a default-constructed left environment; each
    energy against the lowest eigenvalue of the dense effective Hamiltonian
    (built with the same device contractions, upper triangle as the
    reference's LAPACK call) to the reference's 1e-8."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                       "test_lanczos_solver")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "DONE" in r.stdout
    cases = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("CASE ")]
    assert [c[1] for c in cases] == ["cent_err", "cent_maxiter", "lend_err"]
    for c in cases:
        kv = dict(x.split("=") for x in c[2:])
        dim, e = int(kv["dim"]), float(kv["E"])
        H = np.fromfile(str(tmp_path / (kv["file"] + ".bin")), dtype=np.float64).reshape(dim, dim)
        w0 = np.linalg.eigvalsh(H, UPLO="U")[0]
        if c[1] != "cent_maxiter":
            assert abs(e - w0) < 1e-8, (c[1], e, w0)
            continue
        # Stopped by the iteration cap (20): the energy is the Ritz value of a
        # Krylov space K_m(H, psi0) with m <= 21, an upper bound of w0.  (The
        # reference asserts w0 to 1e-8 here because its random draw converges
        # within 20 steps; this draw does not.)
        Hs = np.triu(H) + np.triu(H, 1).T
        v = np.fromfile(str(tmp_path / "cent_maxiter_init.bin"), dtype=np.float64)
        assert int(kv["iters"]) <= 20 and e >= w0 - 1e-9
        Q = (v / np.linalg.norm(v))[:, None]
        ritz = []
        for m in range(1, 22):
            ritz.append(np.linalg.eigvalsh(Q.T @ Hs @ Q)[0])
            w = Hs @ Q[:, -1]
            for _ in range(2):
                w -= Q @ (Q.T @ w)
            Q = np.hstack([Q, (w / np.linalg.norm(w))[:, None]])
        assert min(abs(e - r) for r in ritz) < 1e-8 * abs(w0), (e, ritz[-3:])


@pytest.mark.gpu
def test_cpp_tj_cases(tmp_path):
    """test_two_site_update_finite_vmps.cc:491-640 (real cases) as a C++
    program over the drop-in headers: U(1)xU(1) t-J model (t = 3, J = 1) with
    the two-operator AddTerm and a Jordan-Wigner insertion operator; 4-site
    chain E0 = -6.947478526233 and 2x2 plaquette E0 = -8.868563739680, each to
    the reference's 1e-10."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                       "test_tj_sweeps")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout


@pytest.mark.gpu
def test_cpp_hubbard_kondo_cases(tmp_path):
    """test_two_site_update_finite_vmps.cc:790-951 and :1013-1113 (real cases)
    as a C++ program over the drop-in headers: the Hubbard 2x2 cluster with
    hopping operators built as scalar multiples of tensors, E0 =
    -2.828427124746, and the Kondo double chain on a SiteVec of two different
    local spaces, E0 = -3.180025784229132, each to the reference's 1e-10."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                       "test_hubbard_kondo_sweeps")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout


@pytest.mark.gpu
def test_cpp_spin_2d_cases(tmp_path):
    """test_two_site_update_finite_vmps.cc:249-345 (real cases) as a C++
    program over the drop-in headers: 2D Heisenberg on six sites with the
    vector form of AddTerm, E0 = -3.129385241572, and the Kitaev simple case
    started from ExtendDirectRandomInitMps, E0 = -1.0, each to the reference's
    1e-12."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                       "test_spin_2d_sweeps")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout


def test_cpp_complex_cases(tmp_path):
    """The reference's GQTEN_Complex sweep cases (test_two_site_update_finite_vmps.cc:168-182,
    225-247, 285-303, 347-369, 374-484, 640-659, 704-786) as a C++ program over
    the drop-in headers: complex halves of 1D Ising and Heisenberg, 2D
    Heisenberg, Kitaev simple case, the complex t-J plaquette, the 2D Kitaev
    complex case (3x4, D = 60, E0 = -4.57509167674 to 2e-10) and the Rashba
    t-J case with imaginary couplings (E0 = -11.018692166942165 to 1e-10), each
    at the reference's tolerance."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                       "test_complex_sweeps")
    assert os.path.exists(exe), "tests/cpp is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "PASSED" in r.stdout
