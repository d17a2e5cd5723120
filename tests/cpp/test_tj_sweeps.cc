// t-J cases of tests/test_algorithm/test_two_site_update_finite_vmps.cc:490-640
// (GraceQ/MPS2 0.2-alpha.0), real cases, written against the
// gqmps2-compatible headers: U(1)xU(1) (N, Sz) sites, the two-operator
// AddTerm form with a Jordan-Wigner insertion operator
// (AddTerm(coef, op1, i, op2, j, inst_op), mpogen_impl.h:214-245), 4-site
// chain E0 = -6.947478526233 and 2x2 plaquette E0 = -8.868563739680, each to
// the reference's 1e-10.  Needs a GPU.  Exit status 0 = all cases pass.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <utility>
#include <vector>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;
using namespace gqten;

using U1U1QN = QN<U1QNVal, U1QNVal>;
using IndexT2 = Index<U1U1QN>;
using QNSctT2 = QNSector<U1U1QN>;
using Tensor = GQTensor<GQTEN_Double, U1U1QN>;
using SiteVecT = SiteVec<GQTEN_Double, U1U1QN>;
using MPST = FiniteMPS<GQTEN_Double, U1U1QN>;

static int g_fail = 0, g_checks = 0;

static void RunCase(MPST &mps, const MPO<Tensor> &mpo, const SweepParams &sp, double e_ref, const char *what) {
  ++g_checks;
  double e0 = TwoSiteFiniteVMPS(mps, mpo, sp);
  if (!(std::fabs(e0 - e_ref) <= 1.0E-10)) {
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

int main() {
  Verbose() = false;
  const size_t N = 4;
  const double t = 3.0, J = 1.0;
  U1U1QN qn0({QNCard("N", U1QNVal(0)), QNCard("Sz", U1QNVal(0))});
  IndexT2 pb_out({QNSctT2(U1U1QN({QNCard("N", U1QNVal(1)), QNCard("Sz", U1QNVal(1))}), 1),
                  QNSctT2(U1U1QN({QNCard("N", U1QNVal(1)), QNCard("Sz", U1QNVal(-1))}), 1),
                  QNSctT2(U1U1QN({QNCard("N", U1QNVal(0)), QNCard("Sz", U1QNVal(0))}), 1)},
                 GQTenIndexDirType::OUT);
  IndexT2 pb_in = InverseIndex(pb_out);
  SiteVecT sites(N, pb_out);
  Tensor f({pb_in, pb_out}), sz({pb_in, pb_out}), sp({pb_in, pb_out}), sm({pb_in, pb_out}), cup({pb_in, pb_out}),
      cdagup({pb_in, pb_out}), cdn({pb_in, pb_out}), cdagdn({pb_in, pb_out});
  f({0, 0}) = -1;
  f({1, 1}) = -1;
  f({2, 2}) = 1;
  sz({0, 0}) = 0.5;
  sz({1, 1}) = -0.5;
  sp({1, 0}) = 1;
  sm({0, 1}) = 1;
  cup({0, 2}) = 1;
  cdagup({2, 0}) = 1;
  cdn({1, 2}) = 1;
  cdagdn({2, 1}) = 1;

  auto add_bond = [&](MPOGenerator<GQTEN_Double, U1U1QN> &gen, size_t a, size_t b) {
    gen.AddTerm(-t, cdagup, a, cup, b, f);
    gen.AddTerm(-t, cdagdn, a, cdn, b, f);
    gen.AddTerm(-t, cup, a, cdagup, b, f);
    gen.AddTerm(-t, cdn, a, cdagdn, b, f);
    gen.AddTerm(J, sz, a, sz, b);
    gen.AddTerm(J / 2, sp, a, sm, b);
    gen.AddTerm(J / 2, sm, a, sp, b);
  };

  {  // 1D chain (:557-588)
    MPOGenerator<GQTEN_Double, U1U1QN> gen(sites, qn0);
    for (size_t i = 0; i + 1 < N; ++i) add_bond(gen, i, i + 1);
    auto mpo = gen.Gen();
    SweepParams sweep(11, 8, 8, 1.0E-9, LanczosParams(1.0E-8, 20));
    MPST mps(sites);
    DirectStateInitMps(mps, {2, 1, 2, 0}, qn0);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -6.947478526233, "t-J 1D");
  }
  {  // 2x2 plaquette (:612-640)
    MPOGenerator<GQTEN_Double, U1U1QN> gen(sites, qn0);
    for (auto p : std::vector<std::pair<size_t, size_t>>{{0, 1}, {0, 2}, {2, 3}, {1, 3}}) add_bond(gen, p.first, p.second);
    auto mpo = gen.Gen();
    SweepParams sweep(10, 8, 8, 1.0E-9, LanczosParams(1.0E-8, 20));
    MPST mps(sites);
    DirectStateInitMps(mps, {2, 0, 1, 2}, qn0);
    mps.Dump(sweep.mps_path, true);
    RunCase(mps, mpo, sweep, -8.868563739680, "t-J 2x2");
  }
  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
