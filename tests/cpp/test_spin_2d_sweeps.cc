// 2D spin cases of tests/test_algorithm/test_two_site_update_finite_vmps.cc
// (GraceQ/MPS2 0.2-alpha.0), real cases, written against the
// gqmps2-compatible headers:
//  * 2D Heisenberg on 6 sites (:249-281): the vector form of AddTerm over
//    seven bonds, E0 = -3.129385241572 to 1e-12;
//  * Kitaev simple case on a 4x2 honeycomb (:310-345): z-bonds only, the start
//    state from ExtendDirectRandomInitMps (two product states, enlarged bond
//    dimension 2), E0 = -1.0 to 1e-12.
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
using Tensor = GQTensor<GQTEN_Double, U1QN>;
using MPST = FiniteMPS<GQTEN_Double, U1QN>;

static int g_fail = 0, g_checks = 0;

static void RunCase(MPST &mps, const MPO<Tensor> &mpo, const SweepParams &sp, double e_ref, const char *what) {
  ++g_checks;
  double e0 = TwoSiteFiniteVMPS(mps, mpo, sp);
  if (!(std::fabs(e0 - e_ref) <= 1.0E-12)) {
    ++g_fail;
    std::printf("FAIL %s: E0 = %.15f, reference %.12f\n", what, e0, e_ref);
  }
  ++g_checks;
  if (!mps.empty()) {
    ++g_fail;
    std::printf("FAIL %s: container not empty after the run\n", what);
  }
  std::filesystem::remove_all(sp.mps_path);
  std::filesystem::remove_all(sp.temp_path);
}

static size_t Coors2IdxHoneycomb(int x, int y, size_t Nx, size_t Ny) { return Ny * (x % Nx) + y % Ny; }

int main() {
  Verbose() = false;
  U1QN qn0({QNCard("Sz", U1QNVal(0))});
  Index<U1QN> pb_out({QNSector<U1QN>(U1QN({QNCard("Sz", U1QNVal(1))}), 1),
                      QNSector<U1QN>(U1QN({QNCard("Sz", U1QNVal(-1))}), 1)},
                     GQTenIndexDirType::OUT);
  auto pb_in = InverseIndex(pb_out);
  Tensor sz({pb_in, pb_out}), sp({pb_in, pb_out}), sm({pb_in, pb_out});
  sz({0, 0}) = 0.5;
  sz({1, 1}) = -0.5;
  sp({0, 1}) = 1;
  sm({1, 0}) = 1;

  {  // 2D Heisenberg
    const size_t N = 6;
    SiteVec<GQTEN_Double, U1QN> sites(N, pb_out);
    MPOGenerator<GQTEN_Double, U1QN> gen(sites, qn0);
    std::vector<std::pair<size_t, size_t>> nn = {{0, 1}, {0, 2}, {1, 3}, {2, 3}, {2, 4}, {3, 5}, {4, 5}};
    for (auto &p : nn) {
      gen.AddTerm(1, {sz, sz}, {p.first, p.second});
      gen.AddTerm(0.5, {sp, sm}, {p.first, p.second});
      gen.AddTerm(0.5, {sm, sp}, {p.first, p.second});
    }
    auto mpo = gen.Gen();
    SweepParams sweep(4, 8, 8, 1.0E-9, LanczosParams(1.0E-7));
    std::vector<size_t> labs;
    for (size_t i = 0; i < N; ++i) labs.push_back(i % 2);
    MPST mps(sites);
    DirectStateInitMps(mps, labs, qn0);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -3.129385241572, "2D Heisenberg");
  }
  {  // Kitaev simple case
    const size_t Nx = 4, Ny = 2, N = Nx * Ny;
    SiteVec<GQTEN_Double, U1QN> sites(N, pb_out);
    MPOGenerator<GQTEN_Double, U1QN> gen(sites, qn0);
    for (int x = 0; x < (int)Nx; ++x)
      for (int y = 0; y < (int)Ny; ++y)
        if (x % 2 == 1) {
          size_t s0 = Coors2IdxHoneycomb(x, y, Nx, Ny), s1 = Coors2IdxHoneycomb(x - 1, y + 1, Nx, Ny);
          if (s0 > s1) std::swap(s0, s1);
          gen.AddTerm(1, {sz, sz}, {s0, s1});
        }
    auto mpo = gen.Gen();
    SweepParams sweep(4, 8, 8, 1.0E-4, LanczosParams(1.0E-10));
    std::vector<size_t> labs1, labs2;
    for (size_t i = 0; i < N; ++i) {
      labs1.push_back(i % 2);
      labs2.push_back((i + 1) % 2);
    }
    MPST mps(sites);
    ExtendDirectRandomInitMps(mps, {labs1, labs2}, qn0, 2);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -1.0, "Kitaev simple");
  }
  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
