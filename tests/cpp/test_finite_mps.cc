// FiniteMPS cases of tests/test_one_dim_tn/test_finite_mps/test_finite_mps.cc:31-260
// (GraceQ/MPS2 0.2-alpha.0), written against the gqmps2-compatible headers
// as a user of the reference writes them: a 5-site MPS of random block
// tensors; Centralize to every site (canonical tensors contract to the
// identity, canonical-type flags, centre), copy and move, const versus
// mutable element access, Dump/Load, TruncateMPS.  TruncateMPS is checked
// beyond the reference test (which only runs it): bond dimensions within
// Dmax, left-canonical tensors, a normalised result, and the unchanged
// state when nothing is cut.  Needs a GPU.  Exit status 0 = all pass.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>
#include <utility>
#include <vector>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;
using namespace gqten;

using U1QN = QN<U1QNVal>;
using IndexT = Index<U1QN>;
using QNSctT = QNSector<U1QN>;
using Tensor = GQTensor<GQTEN_Double, U1QN>;
using SiteVecT = SiteVec<GQTEN_Double, U1QN>;
using MPST = FiniteMPS<GQTEN_Double, U1QN>;

static int g_fail = 0, g_checks = 0;
static void Check(bool ok, const std::string &what) {
  ++g_checks;
  if (!ok) {
    ++g_fail;
    std::printf("FAIL %s\n", what.c_str());
  }
}

struct Fixture {
  U1QN qn0{{QNCard("N", U1QNVal(0))}}, qn1{{QNCard("N", U1QNVal(1))}}, qn2{{QNCard("N", U1QNVal(2))}};
  IndexT pb_out{{QNSctT(qn0, 1), QNSctT(qn1, 1)}, GQTenIndexDirType::OUT};
  IndexT vb01_out{{QNSctT(qn0, 1), QNSctT(qn1, 1)}, GQTenIndexDirType::OUT};
  IndexT vb01_in = InverseIndex(vb01_out);
  IndexT vb012_out{{QNSctT(qn0, 1), QNSctT(qn1, 2), QNSctT(qn2, 1)}, GQTenIndexDirType::OUT};
  IndexT vb012_in = InverseIndex(vb012_out);
  Tensor t0{{pb_out, vb01_out}}, t1{{vb01_in, pb_out, vb012_out}}, t2{{vb012_in, pb_out, vb012_out}},
      t3{{vb012_in, pb_out, vb01_out}}, t4{{vb01_in, pb_out}};
  SiteVecT site_vec{5, pb_out};
  MPST mps{site_vec};
  Fixture() {
    t0.Random(qn1);
    t1.Random(qn1);
    t2.Random(qn0);
    t3.Random(qn0);
    t4.Random(qn0);
    mps[0] = t0;
    mps[1] = t1;
    mps[2] = t2;
    mps[3] = t3;
    mps[4] = t4;
  }
};

// A^dag A (left) or A A^dag (right) is the identity (test_finite_mps.cc:67-112).
static void CheckCanonical(const MPST &mps, size_t i, bool left, const std::string &what) {
  const size_t N = mps.size();
  std::vector<std::vector<size_t>> ax;
  if (left) ax = i == 0 ? std::vector<std::vector<size_t>>{{0}, {0}} : std::vector<std::vector<size_t>>{{0, 1}, {0, 1}};
  else ax = i == N - 1 ? std::vector<std::vector<size_t>>{{1}, {1}} : std::vector<std::vector<size_t>>{{1, 2}, {1, 2}};
  Tensor res;
  Tensor ten = mps[i];
  auto dag = Dag(ten);
  Contract(&ten, &dag, ax, &res);
  auto sh = res.GetShape();
  bool ok = sh.size() == 2 && sh[0] == sh[1];
  double worst = 0;
  for (size_t a = 0; ok && a < sh[0]; ++a)
    for (size_t b = 0; b < sh[1]; ++b) worst = std::max(worst, std::fabs(res(a, b) - (a == b ? 1.0 : 0.0)));
  Check(ok && worst <= 1e-14, what + " site " + std::to_string(i) + (left ? " left" : " right") +
                                  "-canonical (max |A^dag A - 1| = " + std::to_string(worst) + ")");
}

static void CheckCenter(const MPST &mps, int center, const std::string &what) {
  Check(mps.GetCenter() == center, what + ": centre");
  auto types = mps.GetTensCanoType();
  for (size_t i = 0; i < mps.size(); ++i) {
    if ((int)i < center) {
      Check(types[i] == MPSTenCanoType::LEFT, what + ": LEFT flag");
      CheckCanonical(mps, i, true, what);
    } else if ((int)i > center) {
      Check(types[i] == MPSTenCanoType::RIGHT, what + ": RIGHT flag");
      CheckCanonical(mps, i, false, what);
    } else {
      Check(types[i] == MPSTenCanoType::NONE, what + ": NONE flag at the centre");
    }
  }
}

static void CentralizeAll(MPST &mps, const std::string &what) {
  for (int c : {0, 2, 4}) {
    mps.Centralize(c);
    CheckCenter(mps, c, what + " Centralize(" + std::to_string(c) + ")");
  }
  for (size_t c = 0; c < mps.size(); ++c) {
    mps.Centralize((int)c);
    CheckCenter(mps, (int)c, what + " sweep Centralize(" + std::to_string(c) + ")");
  }
}

// The whole state as one tensor (legs: the physical legs in site order).
static Tensor FullState(const MPST &mps) {
  Tensor acc = mps[0];
  for (size_t i = 1; i < mps.size(); ++i) {
    Tensor next;
    Contract(&acc, &mps[i], {{acc.Rank() - 1}, {0}}, &next);
    acc = std::move(next);
  }
  return acc;
}
static double Overlap(const Tensor &a, const Tensor &b) {
  std::vector<size_t> ax(a.Rank());
  for (size_t k = 0; k < ax.size(); ++k) ax[k] = k;
  auto bd = Dag(b);
  Tensor r;
  Contract(&a, &bd, {ax, ax}, &r);
  return r();
}

int main() {
  Verbose() = false;
  {  // TestCentralize (test_finite_mps.cc:133-158)
    Fixture f;
    CentralizeAll(f.mps, "random");
    f.mps[0].Random(f.qn0);
    CentralizeAll(f.mps, "site 0 redrawn");
    f.mps[1].Random(f.qn0);
    f.mps[2].Random(f.qn0);
    f.mps[4].Random(f.qn0);
    CentralizeAll(f.mps, "sites 1, 2, 4 redrawn");
  }
  {  // TestCopyAndMove (test_finite_mps.cc:161-205)
    Fixture f;
    f.mps.Centralize(2);
    const MPST &cr = f.mps;
    MPST copy(f.mps);
    const MPST &crc = copy;
    Check(copy.GetCenter() == f.mps.GetCenter(), "copy: centre");
    for (size_t i = 0; i < 5; ++i) {
      Check(copy.GetTenCanoType(i) == f.mps.GetTenCanoType(i), "copy: flags");
      Check(crc[i] == cr[i], "copy: tensors equal");
      Check(crc(i) != cr(i), "copy: tensors are distinct objects");
    }
    MPST copy2 = f.mps;
    auto raw = copy.cdata();
    MPST moved(std::move(copy));
    const MPST &crm = moved;
    Check(moved.GetCenter() == f.mps.GetCenter(), "move: centre");
    for (size_t i = 0; i < 5; ++i) {
      Check(moved.GetTenCanoType(i) == f.mps.GetTenCanoType(i), "move: flags");
      Check(crm[i] == cr[i], "move: tensors equal");
      Check(crm(i) == raw[i], "move: the same tensor objects");
    }
    auto raw2 = copy2.cdata();
    MPST moved2 = std::move(copy2);
    const MPST &crm2 = moved2;
    for (size_t i = 0; i < 5; ++i) Check(crm2(i) == raw2[i], "move-assign: the same tensor objects");
  }
  {  // TestElemAccess (test_finite_mps.cc:208-225)
    Fixture f;
    f.mps.Centralize(2);
    const MPST &cr = f.mps;
    Tensor ten = cr[1];
    Check(f.mps.GetTenCanoType(1) == MPSTenCanoType::LEFT, "const access keeps LEFT");
    ten = f.mps[1];
    Check(f.mps.GetTenCanoType(1) == MPSTenCanoType::NONE, "mutable access resets the flag");
    const MPST *cp = &f.mps;
    ten = (*cp)[3];
    Check(f.mps.GetTenCanoType(3) == MPSTenCanoType::RIGHT, "const access keeps RIGHT");
    MPST *p = &f.mps;
    ten = (*p)[3];
    Check(f.mps.GetTenCanoType(3) == MPSTenCanoType::NONE, "mutable access resets RIGHT");
  }
  {  // TestIO (test_finite_mps.cc:228-245)
    Fixture f;
    MPST mps2(SiteVecT(5, f.pb_out));
    f.mps.Dump();
    mps2.Load();
    for (size_t i = 0; i < 5; ++i) Check(mps2[i] == f.mps[i], "Dump/Load default path");
    f.mps.Dump("mps2");
    mps2.Load("mps2");
    for (size_t i = 0; i < 5; ++i) Check(mps2[i] == f.mps[i], "Dump/Load named path");
    f.mps.Dump("mps3", true);
    Check(f.mps.empty(), "Dump with release empties the container");
    for (auto d : {"mps", "mps2", "mps3"}) std::filesystem::remove_all(d);
  }
  {  // TestTruncate (test_finite_mps.cc:248-254) with checks of the result
    for (size_t dmax : {size_t(3), size_t(2), size_t(1000)}) {
      Fixture f;
      Tensor before = FullState(f.mps);
      const double n0 = std::sqrt(Overlap(before, before));
      TruncateMPS(f.mps, 0.0, 1, dmax);
      const std::string what = "TruncateMPS Dmax=" + std::to_string(dmax);
      for (size_t i = 0; i + 1 < 5; ++i) {
        CheckCanonical(f.mps, i, true, what);
        auto sh = f.mps[i].GetShape();
        Check(sh.back() <= dmax, what + ": bond dimension within Dmax");
      }
      Tensor after = FullState(f.mps);
      const double n1 = std::sqrt(Overlap(after, after));
      Check(std::fabs(n1 - 1.0) <= 1e-13, what + ": normalised result (" + std::to_string(n1) + ")");
      const double fid = Overlap(before, after) / n0;
      Check(fid <= 1.0 + 1e-13 && fid > 0.0, what + ": fidelity in (0, 1]");
      if (dmax == 1000) Check(std::fabs(fid - 1.0) <= 1e-13, what + ": nothing cut, state unchanged");
    }
  }
  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
