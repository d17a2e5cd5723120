// Complex (GQTEN_Complex) cases of tests/test_algorithm/test_two_site_update_finite_vmps.cc
// (GraceQ/MPS2 0.2-alpha.0), written against the gqmps2-compatible headers:
//  * the complex halves of the spin cases: 1D Ising (:168-182, E0 = -0.25 (N-1)
//    to 1e-10), 1D Heisenberg (:225-247, -2.493577133888 to 1e-12), 2D
//    Heisenberg (:285-303, -3.129385241572 to 1e-12), Kitaev simple case with
//    the ExtendDirectRandomInitMps start (:347-369, -1.0 to 1e-12);
//  * the 2D Kitaev complex case on a 3x4 cylinder (:374-484): sy with
//    imaginary entries, Gamma terms, D = 60, trunc_err 1e-4, E0 =
//    -4.57509167674 to 2e-10;
//  * the complex half of the U(1)xU(1) t-J plaquette (:640-659, -8.868563739680
//    to 1e-10);
//  * the t-J Rashba case (:704-786): spin-orbit hopping with imaginary
//    couplings i*lambda, U(1) charge, E0 = -11.018692166942165 to 1e-10.
// The measurements the reference's complex Ising half runs after its sweep are
// not repeated here (the complex path covers the sweep).
// Needs a GPU.  Exit status 0 = all cases pass.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <utility>
#include <vector>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;
using namespace gqten;

using U1QN = QN<U1QNVal>;
using U1U1QN = QN<U1QNVal, U1QNVal>;
using ZTensor = GQTensor<GQTEN_Complex, U1QN>;
using ZTensor2 = GQTensor<GQTEN_Complex, U1U1QN>;
using ZMPS = FiniteMPS<GQTEN_Complex, U1QN>;
using ZMPS2 = FiniteMPS<GQTEN_Complex, U1U1QN>;

static int g_fail = 0, g_checks = 0;

template <typename QNT>
static void RunCase(FiniteMPS<GQTEN_Complex, QNT> &mps, const MPO<GQTensor<GQTEN_Complex, QNT>> &mpo,
                    const SweepParams &sp, double e_ref, double tol, const char *what) {
  ++g_checks;
  double e0 = TwoSiteFiniteVMPS(mps, mpo, sp);
  if (!(std::fabs(e0 - e_ref) <= tol)) {
    ++g_fail;
    std::printf("FAIL %s: E0 = %.15f, reference %.15f (tolerance %.1e)\n", what, e0, e_ref, tol);
  } else {
    std::printf("ok %s: E0 = %.15f (reference %.15f)\n", what, e0, e_ref);
  }
  ++g_checks;
  if (!mps.empty()) {
    ++g_fail;
    std::printf("FAIL %s: container not empty after the run\n", what);
  }
  std::filesystem::remove_all(sp.mps_path);
  std::filesystem::remove_all(sp.temp_path);
}

static void KeepOrder(size_t &x, size_t &y) {
  if (x > y) std::swap(x, y);
}
static size_t Coors2Idx(size_t x, size_t y, size_t, size_t Ny) { return x * Ny + y; }
static size_t Coors2IdxSquare(int x, int y, size_t, size_t Ny) { return x * Ny + y; }
static size_t Coors2IdxHoneycomb(int x, int y, size_t Nx, size_t Ny) { return Ny * (x % Nx) + y % Ny; }

static void SpinCases() {
  const size_t N = 6;
  U1QN qn0({QNCard("Sz", U1QNVal(0))});
  Index<U1QN> pb_out({QNSector<U1QN>(U1QN({QNCard("Sz", U1QNVal(1))}), 1),
                      QNSector<U1QN>(U1QN({QNCard("Sz", U1QNVal(-1))}), 1)},
                     GQTenIndexDirType::OUT);
  auto pb_in = InverseIndex(pb_out);
  ZTensor zsz({pb_in, pb_out}), zsp({pb_in, pb_out}), zsm({pb_in, pb_out});
  zsz({0, 0}) = 0.5;
  zsz({1, 1}) = -0.5;
  zsp({0, 1}) = 1;
  zsm({1, 0}) = 1;
  SiteVec<GQTEN_Complex, U1QN> sites6(N, pb_out);
  std::vector<size_t> labs;
  for (size_t i = 0; i < N; ++i) labs.push_back(i % 2);
  {  // 1D Ising
    MPOGenerator<GQTEN_Complex, U1QN> gen(sites6, qn0);
    for (size_t i = 0; i + 1 < N; ++i) gen.AddTerm(1, {zsz, zsz}, {i, i + 1});
    auto mpo = gen.Gen();
    SweepParams sweep(4, 1, 10, 1.0E-5, LanczosParams(1.0E-7));
    ZMPS mps(sites6);
    DirectStateInitMps(mps, labs, qn0);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -0.25 * (N - 1), 1.0E-10, "complex 1D Ising");
  }
  {  // 1D Heisenberg
    MPOGenerator<GQTEN_Complex, U1QN> gen(sites6, qn0);
    for (size_t i = 0; i + 1 < N; ++i) {
      gen.AddTerm(1, {zsz, zsz}, {i, i + 1});
      gen.AddTerm(0.5, {zsp, zsm}, {i, i + 1});
      gen.AddTerm(0.5, {zsm, zsp}, {i, i + 1});
    }
    auto mpo = gen.Gen();
    SweepParams sweep(4, 8, 8, 1.0E-9, LanczosParams(1.0E-7));
    ZMPS mps(sites6);
    DirectStateInitMps(mps, labs, qn0);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -2.493577133888, 1.0E-12, "complex 1D Heisenberg");
  }
  {  // 2D Heisenberg
    MPOGenerator<GQTEN_Complex, U1QN> gen(sites6, qn0);
    std::vector<std::pair<size_t, size_t>> nn = {{0, 1}, {0, 2}, {1, 3}, {2, 3}, {2, 4}, {3, 5}, {4, 5}};
    for (auto &p : nn) {
      gen.AddTerm(1, {zsz, zsz}, {p.first, p.second});
      gen.AddTerm(0.5, {zsp, zsm}, {p.first, p.second});
      gen.AddTerm(0.5, {zsm, zsp}, {p.first, p.second});
    }
    auto mpo = gen.Gen();
    SweepParams sweep(4, 8, 8, 1.0E-9, LanczosParams(1.0E-7));
    ZMPS mps(sites6);
    DirectStateInitMps(mps, labs, qn0);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -3.129385241572, 1.0E-12, "complex 2D Heisenberg");
  }
  {  // Kitaev simple case
    const size_t Nx = 4, Ny = 2, N8 = Nx * Ny;
    SiteVec<GQTEN_Complex, U1QN> sites(N8, pb_out);
    MPOGenerator<GQTEN_Complex, U1QN> gen(sites, qn0);
    for (int x = 0; x < (int)Nx; ++x)
      for (int y = 0; y < (int)Ny; ++y)
        if (x % 2 == 1) {
          size_t s0 = Coors2IdxHoneycomb(x, y, Nx, Ny), s1 = Coors2IdxHoneycomb(x - 1, y + 1, Nx, Ny);
          KeepOrder(s0, s1);
          gen.AddTerm(1, {zsz, zsz}, {s0, s1});
        }
    auto mpo = gen.Gen();
    SweepParams sweep(4, 8, 8, 1.0E-4, LanczosParams(1.0E-10));
    std::vector<size_t> labs1, labs2;
    for (size_t i = 0; i < N8; ++i) {
      labs1.push_back(i % 2);
      labs2.push_back((i + 1) % 2);
    }
    ZMPS mps(sites);
    ExtendDirectRandomInitMps(mps, {labs1, labs2}, qn0, 2);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -1.0, 1.0E-12, "complex Kitaev simple");
  }
}

static void KitaevComplexCase() {
  auto zero_div = U1QN({QNCard("N", U1QNVal(0))});
  auto idx_out = Index<U1QN>({QNSector<U1QN>(U1QN({QNCard("N", U1QNVal(1))}), 2)}, GQTenIndexDirType::OUT);
  auto idx_in = InverseIndex(idx_out);
  ZTensor sz({idx_in, idx_out}), sx({idx_in, idx_out}), sy({idx_in, idx_out});
  sz({0, 0}) = GQTEN_Complex(0.5, 0);
  sz({1, 1}) = GQTEN_Complex(-0.5, 0);
  sx({0, 1}) = GQTEN_Complex(0.5, 0);
  sx({1, 0}) = GQTEN_Complex(0.5, 0);
  sy({0, 1}) = GQTEN_Complex(0, -0.5);
  sy({1, 0}) = GQTEN_Complex(0, 0.5);
  const double J = -1.0, K = 1.0, Gm = 0.1, h = 0.1;
  const size_t Nx = 3, Ny = 4, N = Nx * Ny;
  SiteVec<GQTEN_Complex, U1QN> site_vec(N, idx_out);
  MPOGenerator<GQTEN_Complex, U1QN> gen(site_vec, zero_div);
  for (size_t x = 0; x < Nx; ++x)
    for (size_t y = 0; y < Ny; ++y) {
      const size_t s = Coors2Idx(x, y, Nx, Ny);
      gen.AddTerm(-h, sz, s);
      gen.AddTerm(-h, sx, s);
      gen.AddTerm(-h, sy, s);
      if (y % 2 == 0) {  // x-link
        size_t s0 = Coors2Idx(x, y, Nx, Ny), s1 = Coors2Idx(x, y + 1, Nx, Ny);
        gen.AddTerm(J, sz, s0, sz, s1);
        gen.AddTerm(J, sx, s0, sx, s1);
        gen.AddTerm(J, sy, s0, sy, s1);
        gen.AddTerm(K, sx, s0, sx, s1);
        gen.AddTerm(Gm, sz, s0, sy, s1);
        gen.AddTerm(Gm, sy, s0, sz, s1);
      }
      if (y % 2 == 1) {  // z-link
        size_t s0 = Coors2Idx(x, y, Nx, Ny), s1 = Coors2Idx(x, (y + 1) % Ny, Nx, Ny);
        KeepOrder(s0, s1);
        gen.AddTerm(J, sz, s0, sz, s1);
        gen.AddTerm(J, sx, s0, sx, s1);
        gen.AddTerm(J, sy, s0, sy, s1);
        gen.AddTerm(K, sz, s0, sz, s1);
        gen.AddTerm(Gm, sx, s0, sy, s1);
        gen.AddTerm(Gm, sy, s0, sx, s1);
      }
      if ((y % 2 == 1) && (x != Nx - 1)) {  // y-link, cylinder
        size_t s0 = Coors2Idx(x, y, Nx, Ny), s1 = Coors2Idx(x + 1, y - 1, Nx, Ny);
        KeepOrder(s0, s1);
        gen.AddTerm(J, sz, s0, sz, s1);
        gen.AddTerm(J, sx, s0, sx, s1);
        gen.AddTerm(J, sy, s0, sy, s1);
        gen.AddTerm(K, sy, s0, sy, s1);
        gen.AddTerm(Gm, sz, s0, sx, s1);
        gen.AddTerm(Gm, sx, s0, sz, s1);
      }
    }
  auto mpo = gen.Gen();
  ZMPS mps(site_vec);
  std::vector<size_t> labs(N);
  bool was_up = false;
  for (size_t i = 0; i < N; ++i) {
    labs[i] = was_up ? 1 : 0;
    was_up = !was_up;
  }
  SweepParams sweep(4, 60, 60, 1.0E-4, LanczosParams(1.0E-10));
  DirectStateInitMps(mps, labs, zero_div);
  mps.Dump(sweep.mps_path, true);
  RunCase(mps, mpo, sweep, -4.57509167674, 2.0E-10, "complex 2D Kitaev (3x4, D = 60)");
}

static void TjComplexPlaquette() {
  const size_t N = 4;
  const double t = 3.0, J = 1.0;
  U1U1QN qn0({QNCard("N", U1QNVal(0)), QNCard("Sz", U1QNVal(0))});
  Index<U1U1QN> pb_out({QNSector<U1U1QN>(U1U1QN({QNCard("N", U1QNVal(1)), QNCard("Sz", U1QNVal(1))}), 1),
                        QNSector<U1U1QN>(U1U1QN({QNCard("N", U1QNVal(1)), QNCard("Sz", U1QNVal(-1))}), 1),
                        QNSector<U1U1QN>(U1U1QN({QNCard("N", U1QNVal(0)), QNCard("Sz", U1QNVal(0))}), 1)},
                       GQTenIndexDirType::OUT);
  auto pb_in = InverseIndex(pb_out);
  ZTensor2 zf({pb_in, pb_out}), zsz({pb_in, pb_out}), zsp({pb_in, pb_out}), zsm({pb_in, pb_out}),
      zcup({pb_in, pb_out}), zcdagup({pb_in, pb_out}), zcdn({pb_in, pb_out}), zcdagdn({pb_in, pb_out});
  zf({0, 0}) = -1;
  zf({1, 1}) = -1;
  zf({2, 2}) = 1;
  zsz({0, 0}) = 0.5;
  zsz({1, 1}) = -0.5;
  zsp({1, 0}) = 1;
  zsm({0, 1}) = 1;
  zcup({0, 2}) = 1;
  zcdagup({2, 0}) = 1;
  zcdn({1, 2}) = 1;
  zcdagdn({2, 1}) = 1;
  SiteVec<GQTEN_Complex, U1U1QN> sites(N, pb_out);
  MPOGenerator<GQTEN_Complex, U1U1QN> gen(sites, qn0);
  std::vector<std::pair<size_t, size_t>> nn = {{0, 1}, {0, 2}, {2, 3}, {1, 3}};
  for (auto &p : nn) {
    gen.AddTerm(-t, zcdagup, p.first, zcup, p.second, zf);
    gen.AddTerm(-t, zcdagdn, p.first, zcdn, p.second, zf);
    gen.AddTerm(-t, zcup, p.first, zcdagup, p.second, zf);
    gen.AddTerm(-t, zcdn, p.first, zcdagdn, p.second, zf);
    gen.AddTerm(J, zsz, p.first, zsz, p.second);
    gen.AddTerm(J / 2, zsp, p.first, zsm, p.second);
    gen.AddTerm(J / 2, zsm, p.first, zsp, p.second);
  }
  auto mpo = gen.Gen();
  SweepParams sweep(10, 8, 8, 1.0E-9, LanczosParams(1.0E-8, 20));
  ZMPS2 mps(sites);
  DirectStateInitMps(mps, {2, 0, 1, 2}, qn0);
  mps.Dump(sweep.mps_path, true);
  RunCase(mps, mpo, sweep, -8.868563739680, 1.0E-10, "complex t-J plaquette");
}

static void RashbaCase() {
  U1QN qn0({QNCard("N", U1QNVal(0))});
  Index<U1QN> pb_out({QNSector<U1QN>(U1QN({QNCard("N", U1QNVal(1))}), 2),
                      QNSector<U1QN>(U1QN({QNCard("N", U1QNVal(0))}), 1)},
                     GQTenIndexDirType::OUT);
  auto pb_in = InverseIndex(pb_out);
  ZTensor zf({pb_in, pb_out}), zsz({pb_in, pb_out}), zsp({pb_in, pb_out}), zsm({pb_in, pb_out}),
      zcup({pb_in, pb_out}), zcdagup({pb_in, pb_out}), zcdn({pb_in, pb_out}), zcdagdn({pb_in, pb_out}),
      zntot({pb_in, pb_out});
  zf({0, 0}) = -1;
  zf({1, 1}) = -1;
  zf({2, 2}) = 1;
  zsz({0, 0}) = 0.5;
  zsz({1, 1}) = -0.5;
  zsp({0, 1}) = 1;
  zsm({1, 0}) = 1;
  zcup({2, 0}) = 1;
  zcdagup({0, 2}) = 1;
  zcdn({2, 1}) = 1;
  zcdagdn({1, 2}) = 1;
  zntot({0, 0}) = 1;
  zntot({1, 1}) = 1;
  const double t = 3.0, J = 1.0, lamb = 0.03;
  const GQTEN_Complex ilamb(0, lamb);
  const size_t Nx = 3, Ny = 2, Ntot = Nx * Ny;
  const char BCx = 'p', BCy = 'o';
  SiteVec<GQTEN_Complex, U1QN> site_vec(Ntot, pb_out);
  MPOGenerator<GQTEN_Complex, U1QN> gen(site_vec, qn0);
  for (int x = 0; x < (int)Nx; ++x)
    for (int y = 0; y < (int)Ny; ++y) {
      if (!((BCx == 'o') && (x == (int)Nx - 1))) {
        size_t s0 = Coors2IdxSquare(x, y, Nx, Ny), s1 = Coors2IdxSquare((x + 1) % Nx, y, Nx, Ny);
        KeepOrder(s0, s1);
        gen.AddTerm(-t, zcdagup, s0, zcup, s1, zf);
        gen.AddTerm(-t, zcdagdn, s0, zcdn, s1, zf);
        gen.AddTerm(-t, zcup, s0, zcdagup, s1, zf);
        gen.AddTerm(-t, zcdn, s0, zcdagdn, s1, zf);
        gen.AddTerm(J, zsz, s0, zsz, s1);
        gen.AddTerm(J / 2, zsp, s0, zsm, s1);
        gen.AddTerm(J / 2, zsm, s0, zsp, s1);
        gen.AddTerm(-J / 4, zntot, s0, zntot, s1);
        if (x != (int)Nx - 1) {
          gen.AddTerm(lamb, zcdagup, s0, zcdn, s1, zf);
          gen.AddTerm(lamb, zcup, s0, zcdagdn, s1, zf);
          gen.AddTerm(-lamb, zcdagdn, s0, zcup, s1, zf);
          gen.AddTerm(-lamb, zcdn, s0, zcdagup, s1, zf);
        } else {  // at the boundary
          gen.AddTerm(lamb, zcdn, s0, zcdagup, s1, zf);
          gen.AddTerm(lamb, zcdagdn, s0, zcup, s1, zf);
          gen.AddTerm(-lamb, zcup, s0, zcdagdn, s1, zf);
          gen.AddTerm(-lamb, zcdagup, s0, zcdn, s1, zf);
        }
      }
      if (!((BCy == 'o') && (y == (int)Ny - 1))) {
        size_t s0 = Coors2IdxSquare(x, y, Nx, Ny), s2 = Coors2IdxSquare(x, (y + 1) % Ny, Nx, Ny);
        KeepOrder(s0, s2);
        gen.AddTerm(-t, zcdagup, s0, zcup, s2, zf);
        gen.AddTerm(-t, zcdagdn, s0, zcdn, s2, zf);
        gen.AddTerm(-t, zcup, s0, zcdagup, s2, zf);
        gen.AddTerm(-t, zcdn, s0, zcdagdn, s2, zf);
        gen.AddTerm(J, zsz, s0, zsz, s2);
        gen.AddTerm(J / 2, zsp, s0, zsm, s2);
        gen.AddTerm(J / 2, zsm, s0, zsp, s2);
        gen.AddTerm(-J / 4, zntot, s0, zntot, s2);
        if (y != (int)Ny - 1) {
          gen.AddTerm(ilamb, zcdagup, s0, zcdn, s2, zf);
          gen.AddTerm(ilamb, zcdagdn, s0, zcup, s2, zf);
          gen.AddTerm(-ilamb, zcup, s0, zcdagdn, s2, zf);
          gen.AddTerm(-ilamb, zcdn, s0, zcdagup, s2, zf);
        } else {  // at the boundary
          gen.AddTerm(ilamb, zcup, s0, zcdagdn, s2, zf);
          gen.AddTerm(ilamb, zcdn, s0, zcdagup, s2, zf);
          gen.AddTerm(-ilamb, zcdagup, s0, zcdn, s2, zf);
          gen.AddTerm(-ilamb, zcdagdn, s0, zcup, s2, zf);
        }
      }
    }
  auto mpo = gen.Gen();
  SweepParams sweep(8, 30, 30, 1.0E-4, LanczosParams(1.0E-14, 100));
  ZMPS mps(site_vec);
  DirectStateInitMps(mps, {0, 1, 0, 2, 0, 1}, qn0);
  mps.Dump(sweep.mps_path, true);
  RunCase(mps, mpo, sweep, -11.018692166942165, 1.0E-10, "t-J Rashba (complex couplings)");
}

int main() {
  Verbose() = false;
  SpinCases();
  KitaevComplexCase();
  TjComplexPlaquette();
  RashbaCase();
  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
