// Hubbard and Kondo cases of tests/test_algorithm/test_two_site_update_finite_vmps.cc
// (GraceQ/MPS2 0.2-alpha.0), real cases, written against the
// gqmps2-compatible headers:
//  * Hubbard 2x2 (:790-951): (Nup, Ndn) sites, hopping operators built as
//    scalar multiples (-t0 * a^+_up f), next-nearest bonds ordered by
//    KeepOrder, E0 = -2.828427124746 to 1e-10;
//  * Kondo insulator double chain (:1013-1113): a SiteVec of two different
//    local spaces (4-dim conduction, 2-dim local sites), terms two sites
//    apart with identities between, E0 = -3.180025784229132 to 1e-10.
// Needs a GPU.  Exit status 0 = all cases pass.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <vector>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;
using namespace gqten;

using U1QN = QN<U1QNVal>;
using U1U1QN = QN<U1QNVal, U1QNVal>;

static int g_fail = 0, g_checks = 0;

template <typename MPST, typename MPOT>
static void RunCase(MPST &mps, const MPOT &mpo, const SweepParams &sp, double e_ref, const char *what) {
  ++g_checks;
  double e0 = TwoSiteFiniteVMPS(mps, mpo, sp);
  if (!(std::fabs(e0 - e_ref) <= 1.0E-10)) {
    ++g_fail;
    std::printf("FAIL %s: E0 = %.15f, reference %.15f\n", what, e0, e_ref);
  }
  std::filesystem::remove_all(sp.mps_path);
  std::filesystem::remove_all(sp.temp_path);
}

static void KeepOrder(size_t &x, size_t &y) {
  if (x > y) std::swap(x, y);
}

static void Hubbard2x2() {
  using Tensor = GQTensor<GQTEN_Double, U1U1QN>;
  const size_t Nx = 2, Ny = 2, N = Nx * Ny;
  const double t0 = 1.0, t1 = 0.5, U = 2.0;
  auto qn = [](int up, int dn) { return U1U1QN({QNCard("Nup", U1QNVal(up)), QNCard("Ndn", U1QNVal(dn))}); };
  Index<U1U1QN> pb_out({QNSector<U1U1QN>(qn(0, 0), 1), QNSector<U1U1QN>(qn(1, 0), 1),
                        QNSector<U1U1QN>(qn(0, 1), 1), QNSector<U1U1QN>(qn(1, 1), 1)},
                       GQTenIndexDirType::OUT);
  auto pb_in = InverseIndex(pb_out);
  SiteVec<GQTEN_Double, U1U1QN> sites(N, pb_out);
  Tensor f({pb_in, pb_out}), nupdn({pb_in, pb_out}), adagupf({pb_in, pb_out}), aup({pb_in, pb_out}),
      adagdn({pb_in, pb_out}), fadn({pb_in, pb_out}), naupf({pb_in, pb_out}), adagup({pb_in, pb_out}),
      nadn({pb_in, pb_out}), fadagdn({pb_in, pb_out});
  f({0, 0}) = 1;
  f({1, 1}) = -1;
  f({2, 2}) = -1;
  f({3, 3}) = 1;
  nupdn({3, 3}) = 1;
  adagupf({1, 0}) = 1;
  adagupf({3, 2}) = -1;
  aup({0, 1}) = 1;
  aup({2, 3}) = 1;
  adagdn({2, 0}) = 1;
  adagdn({3, 1}) = 1;
  fadn({0, 2}) = 1;
  fadn({1, 3}) = -1;
  naupf({0, 1}) = 1;
  naupf({2, 3}) = -1;
  adagup({1, 0}) = 1;
  adagup({3, 2}) = 1;
  nadn({0, 2}) = -1;
  nadn({1, 3}) = -1;
  fadagdn({2, 0}) = -1;
  fadagdn({3, 1}) = 1;

  MPOGenerator<GQTEN_Double, U1U1QN> gen(sites, qn(0, 0));
  auto hop = [&](size_t a, size_t b, double t) {
    gen.AddTerm(1, -t * adagupf, a, aup, b, f);
    gen.AddTerm(1, -t * adagdn, a, fadn, b, f);
    gen.AddTerm(1, naupf, a, -t * adagup, b, f);
    gen.AddTerm(1, nadn, a, -t * fadagdn, b, f);
  };
  for (size_t i = 0; i < Nx; ++i) {
    for (size_t j = 0; j < Ny; ++j) {
      size_t s0 = i * Ny + j;
      gen.AddTerm(U, nupdn, s0);
      if (i != Nx - 1) hop(s0, (i + 1) * Ny + j, t0);
      if (j != Ny - 1) hop(s0, i * Ny + j + 1, t0);
      if (j != Ny - 1) {
        if (i != 0) {
          size_t a = s0, b = (i - 1) * Ny + j + 1;
          KeepOrder(a, b);
          hop(a, b, t1);
        }
        if (i != Nx - 1) {
          size_t a = s0, b = (i + 1) * Ny + j + 1;
          KeepOrder(a, b);
          hop(a, b, t1);
        }
      }
    }
  }
  auto mpo = gen.Gen();
  SweepParams sweep(10, 16, 16, 1.0E-9, LanczosParams(1.0E-8, 20));
  FiniteMPS<GQTEN_Double, U1U1QN> mps(sites);
  std::vector<size_t> labs(N);
  for (size_t i = 0; i < N; ++i) labs[i] = (i % 2 == 0 ? 1 : 2);
  DirectStateInitMps(mps, labs, qn(0, 0));
  mps.Dump(sweep.mps_path, true);
  RunCase(mps, mpo, sweep, -2.828427124746, "Hubbard 2x2");
}

static void KondoDoubleChain() {
  using Tensor = GQTensor<GQTEN_Double, U1QN>;
  const size_t Nx = 4, N = 2 * Nx;
  const double t = 0.25, Jk = 1.0, Jz = 0.5;
  U1QN qn0({QNCard("Sz", U1QNVal(0))});
  Index<U1QN> pb_outE({QNSector<U1QN>(qn0, 4)}, GQTenIndexDirType::OUT);
  Index<U1QN> pb_outL({QNSector<U1QN>(qn0, 2)}, GQTenIndexDirType::OUT);
  auto pb_inE = InverseIndex(pb_outE), pb_inL = InverseIndex(pb_outL);
  Tensor sz({pb_inE, pb_outE}), sp({pb_inE, pb_outE}), sm({pb_inE, pb_outE}), bupcF({pb_inE, pb_outE}),
      bupaF({pb_inE, pb_outE}), Fbdnc({pb_inE, pb_outE}), Fbdna({pb_inE, pb_outE}), bupc({pb_inE, pb_outE}),
      bupa({pb_inE, pb_outE}), bdnc({pb_inE, pb_outE}), bdna({pb_inE, pb_outE});
  Tensor Sz({pb_inL, pb_outL}), Sp({pb_inL, pb_outL}), Sm({pb_inL, pb_outL});
  sz({0, 0}) = 0.5;
  sz({1, 1}) = -0.5;
  sp({0, 1}) = 1.0;
  sm({1, 0}) = 1.0;
  bupcF({2, 1}) = -1;
  bupcF({0, 3}) = 1;
  Fbdnc({2, 0}) = 1;
  Fbdnc({1, 3}) = -1;
  bupaF({1, 2}) = 1;
  bupaF({3, 0}) = -1;
  Fbdna({0, 2}) = -1;
  Fbdna({3, 1}) = 1;
  bupc({2, 1}) = 1;
  bupc({0, 3}) = 1;
  bdnc({2, 0}) = 1;
  bdnc({1, 3}) = 1;
  bupa({1, 2}) = 1;
  bupa({3, 0}) = 1;
  bdna({0, 2}) = 1;
  bdna({3, 1}) = 1;
  Sz({0, 0}) = 0.5;
  Sz({1, 1}) = -0.5;
  Sp({0, 1}) = 1.0;
  Sm({1, 0}) = 1.0;
  std::vector<Index<U1QN>> pb_set(N);
  for (size_t i = 0; i < N; ++i) pb_set[i] = (i % 2 == 0 ? pb_outE : pb_outL);

  SiteVec<GQTEN_Double, U1QN> sites(pb_set);
  FiniteMPS<GQTEN_Double, U1QN> mps(sites);
  MPOGenerator<GQTEN_Double, U1QN> gen(sites, qn0);
  for (size_t i = 0; i < N - 2; i += 2) {
    gen.AddTerm(-t, bupcF, i, bupa, i + 2);
    gen.AddTerm(-t, bdnc, i, Fbdna, i + 2);
    gen.AddTerm(t, bupaF, i, bupc, i + 2);
    gen.AddTerm(t, bdna, i, Fbdnc, i + 2);
    gen.AddTerm(Jz, Sz, i + 1, Sz, i + 3);
  }
  for (size_t i = 0; i < N; i += 2) {
    gen.AddTerm(Jk, sz, i, Sz, i + 1);
    gen.AddTerm(Jk / 2, sp, i, Sm, i + 1);
    gen.AddTerm(Jk / 2, sm, i, Sp, i + 1);
  }
  auto mpo = gen.Gen();
  SweepParams sweep(5, 64, 64, 1.0E-9, LanczosParams(1.0E-8, 20));
  std::vector<size_t> labs;
  for (size_t i = 0; i < N; ++i) labs.push_back(i % 2);
  DirectStateInitMps(mps, labs, qn0);
  mps.Dump(sweep.mps_path, true);
  RunCase(mps, mpo, sweep, -3.180025784229132, "Kondo double chain");
}

int main() {
  Verbose() = false;
  Hubbard2x2();
  KondoDoubleChain();
  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
