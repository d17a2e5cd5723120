// Measurement cases of tests/test_one_dim_tn/test_finite_mps/test_finite_mps_measu.cc:43-170
// (GraceQ/MPS2 0.2-alpha.0), double-precision cases, written against the
// gqmps2-compatible headers of this repository exactly as a user of the
// reference writes them: a 6-site chain of (N=0, N=1) sites, product states
// from DirectStateInitMps, <n_i> and <n_i n_j> / <1 1> strings.  Needs a GPU
// (every contraction runs on the device).  Exit status 0 = all cases pass.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;
using namespace gqten;

using U1QN = QN<U1QNVal>;
using IndexT = Index<U1QN>;
using QNSctT = QNSector<U1QN>;
using DGQTensor = GQTensor<GQTEN_Double, U1QN>;
using DSiteVec = SiteVec<GQTEN_Double, U1QN>;
using DMPS = FiniteMPS<GQTEN_Double, U1QN>;

static int g_fail = 0, g_checks = 0;

// gtest's EXPECT_DOUBLE_EQ: equal within 4 units in the last place.
static void ExpectDoubleEq(double got, double want, const char *what, size_t k) {
  ++g_checks;
  double tol = 4 * std::ldexp(1.0, -52) * std::max(std::fabs(got), std::fabs(want));
  if (!(std::fabs(got - want) <= tol)) {
    ++g_fail;
    std::printf("FAIL %s[%zu]: got %.17g want %.17g\n", what, k, got, want);
  }
}

int main() {
  const size_t N = 6;
  U1QN qn0({QNCard("N", U1QNVal(0))});
  IndexT pb_out({QNSctT(U1QN({QNCard("N", U1QNVal(0))}), 1), QNSctT(U1QN({QNCard("N", U1QNVal(1))}), 1)},
                GQTenIndexDirType::OUT);
  IndexT pb_in = InverseIndex(pb_out);
  DGQTensor dntot({pb_in, pb_out}), did({pb_in, pb_out});
  dntot({0, 0}) = 0;
  dntot({1, 1}) = 1;
  did({0, 0}) = 1;
  did({1, 1}) = 1;
  DMPS dmps(DSiteVec(N, pb_out));
  std::vector<size_t> labs1(N, 1), labs2;
  for (size_t i = 0; i < N; ++i) labs2.push_back(i % 2);

  // MeasureOneSiteOp: <n_i> equals the occupation of the product state.
  for (auto *labs : {&labs1, &labs2}) {
    DirectStateInitMps(dmps, *labs, qn0);
    auto res = MeasureOneSiteOp(dmps, dntot, "op1");
    if (res.size() != N) ++g_fail;
    for (size_t i = 0; i < N; ++i) ExpectDoubleEq(res[i].avg, (double)(*labs)[i], "one-site", i);
  }

  // MeasureTwoSiteOp with an identity insertion operator.
  std::vector<std::vector<size_t>> sites_set = {{0, 1}, {0, 2}, {0, 5}, {1, 2}, {1, 3}, {4, 5}};
  auto two = [&](const std::vector<DGQTensor> &phys, const std::vector<double> &want, const char *what) {
    auto res = MeasureTwoSiteOp(dmps, phys, did, sites_set, "op1op2");
    if (res.size() != want.size()) ++g_fail;
    for (size_t k = 0; k < want.size(); ++k) {
      ExpectDoubleEq(res[k].avg, want[k], what, k);
      if (res[k].sites != sites_set[k]) ++g_fail;
    }
  };
  std::vector<double> ones(sites_set.size(), 1.0), dres2 = {0, 0, 0, 0, 1, 0};
  DirectStateInitMps(dmps, labs1, qn0);
  two({did, did}, ones, "id-id all-occupied");
  two({dntot, dntot}, ones, "n-n all-occupied");
  DirectStateInitMps(dmps, labs2, qn0);
  two({did, did}, ones, "id-id alternating");
  two({dntot, dntot}, dres2, "n-n alternating");

  // The result file of the last call: 6 events in the reference's layout.
  std::ifstream f("op1op2.json");
  std::stringstream ss;
  ss << f.rdbuf();
  std::string txt = ss.str();
  if (txt.rfind("[\n  [[0, 1], ", 0) != 0 || txt.back() != ']') {
    ++g_fail;
    std::printf("FAIL dump layout:\n%s\n", txt.c_str());
  }

  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
