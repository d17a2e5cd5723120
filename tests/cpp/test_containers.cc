// Host-side container cases of the reference's unit tests, written against
// the gqmps2-compatible headers: DuoVector (tests/test_one_dim_tn/test_duovector.cc:20-110),
// TenVec tensor files (test_ten_vec.cc:27-60), SiteVec (tests/test_site_vec.cc:20-102)
// and MPO (test_mpo.cc:20-71), GraceQ/MPS2 0.2-alpha.0.  No device call is
// made: tensors are filled, compared, written and read on the host.
// Exit status 0 = all cases pass.
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <stdexcept>
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

static int g_fail = 0, g_checks = 0;
static void Check(bool ok, const char *what) {
  ++g_checks;
  if (!ok) {
    ++g_fail;
    std::printf("FAIL %s\n", what);
  }
}

template <typename ElemT>
static void DuoVectorConstructors(size_t size) {
  DuoVector<ElemT> dv(size);
  Check(dv.size() == size, "DuoVector size");
  for (auto p : dv.cdata()) Check(p == nullptr, "DuoVector starts empty");
  for (size_t i = 0; i < size; ++i) dv[i] = ElemT(i + 5);
  DuoVector<ElemT> copy(dv);
  for (size_t i = 0; i < size; ++i) {
    Check(copy[i] == dv[i], "copy equal");
    Check(copy(i) != dv(i), "copy distinct");
  }
  auto raw = copy.cdata();
  DuoVector<ElemT> moved(std::move(copy));
  for (size_t i = 0; i < size; ++i) {
    Check(moved(i) == raw[i], "move keeps elements");
    Check(moved[i] == dv[i], "move equal");
  }
  auto copy2 = dv;
  auto raw2 = copy2.cdata();
  auto moved2 = std::move(copy2);
  for (size_t i = 0; i < size; ++i) Check(moved2(i) == raw2[i] && moved2[i] == dv[i], "move-assign");
}

static bool IsIdOp(const Tensor &t) {
  auto sh = t.GetShape();
  if (t.GetIndexes().size() != 2 || sh[0] != sh[1]) return false;
  if (t.GetIndexes()[0].GetDir() != IN || InverseIndex(t.GetIndexes()[0]) != t.GetIndexes()[1]) return false;
  for (size_t i = 0; i < sh[0]; ++i)
    for (size_t j = 0; j < sh[1]; ++j)
      if (t(i, j) != (i == j ? 1.0 : 0.0)) return false;
  return true;
}

int main() {
  // ---- DuoVector
  {
    DuoVector<int> d0;
    Check(d0.size() == 0, "default DuoVector is empty");
    DuoVectorConstructors<int>(1);
    DuoVectorConstructors<int>(3);
    DuoVectorConstructors<double>(1);
    DuoVectorConstructors<double>(3);
    DuoVector<int> iv(1);
    iv[0] = 3;
    Check(iv[0] == 3 && iv.front() == 3 && iv.back() == 3, "element access");
    iv.front() = 6;
    Check(iv.front() == 6 && iv.back() == 6, "front/back write");
    auto pe = iv.cdata()[0];
    iv[0] = 5;
    Check(iv.cdata()[0] == pe && iv[0] == 5, "assignment keeps the element object");
    auto pe2 = new int(4);
    delete iv(0);
    iv(0) = pe2;
    Check(iv[0] == 4 && iv.cdata()[0] == pe2, "raw pointer replacement");
    DuoVector<int> iv2(2);
    iv2.alloc(0);
    Check(iv2.cdata()[0] != nullptr && iv2.cdata()[1] == nullptr, "alloc");
    iv2[0] = 3;
    iv2.dealloc(0);
    iv2.dealloc(1);
    Check(iv2.cdata()[0] == nullptr && iv2.cdata()[1] == nullptr, "dealloc");
  }
  // ---- TenVec files
  {
    U1QN qn0({QNCard("N", U1QNVal(0))}), qn1({QNCard("N", U1QNVal(1))}), qnm1({QNCard("N", U1QNVal(-1))});
    IndexT out({QNSctT(qn0, 2), QNSctT(qn1, 2)}, GQTenIndexDirType::OUT);
    auto in = InverseIndex(out);
    Tensor t0({in, out}), t1({in, out}), t2({in, out});
    t0.Random(qn0);
    t1.Random(qn1);
    t2.Random(qnm1);
    TenVec<Tensor> tv(3);
    tv[0] = t0;
    tv[1] = t1;
    tv[2] = t2;
    tv.DumpTen(0, "ten0." + kGQTenFileSuffix);
    tv.DumpTen(1, "ten1." + kGQTenFileSuffix, true);
    tv.DumpTen(2, "ten2." + kGQTenFileSuffix, false);
    tv.dealloc(0);
    tv.dealloc(2);
    Check(tv.empty(), "TenVec empty after release");
    tv.LoadTen(0, "ten2." + kGQTenFileSuffix);
    tv.LoadTen(1, "ten0." + kGQTenFileSuffix);
    tv.LoadTen(2, "ten1." + kGQTenFileSuffix);
    Check(tv[0] == t2 && tv[1] == t0 && tv[2] == t1, "TenVec files round trip");
    for (auto f : {"ten0.", "ten1.", "ten2."}) std::filesystem::remove(std::string(f) + kGQTenFileSuffix);
  }
  // ---- SiteVec
  {
    IndexT pb_out1({QNSctT(U1QN({QNCard("N", U1QNVal(0))}), 1), QNSctT(U1QN({QNCard("N", U1QNVal(1))}), 1)}, OUT);
    IndexT pb_in1 = InverseIndex(pb_out1);
    IndexT pb_out2({QNSctT(U1QN({QNCard("N", U1QNVal(1))}), 3)}, OUT);
    IndexT pb_in2 = InverseIndex(pb_out2);
    for (int N : {1, 3})
      for (auto *lh : {&pb_out1, &pb_in1}) {
        SiteVec<GQTEN_Double, U1QN> sv(N, *lh);
        Check(sv.size == (size_t)N, "SiteVec size");
        Check(sv.sites == IndexVec<U1QN>(N, pb_out1), "SiteVec sites are OUT");
        for (int i = 0; i < N; ++i) Check(IsIdOp(sv.id_ops[i]), "SiteVec identity operators");
      }
    std::vector<IndexVec<U1QN>> lists = {{pb_out1}, {pb_out2, pb_out2}, {pb_out1, pb_out2, pb_out1},
                                         {pb_in2, pb_out1, pb_out1, pb_in1, pb_out2}};
    for (auto &l : lists) {
      SiteVec<GQTEN_Double, U1QN> sv(l);
      Check(sv.size == l.size(), "SiteVec size (list)");
      for (size_t i = 0; i < l.size(); ++i) {
        Check(sv.sites[i] == (l[i].GetDir() == OUT ? l[i] : InverseIndex(l[i])), "SiteVec site (list)");
        Check(IsIdOp(sv.id_ops[i]), "SiteVec identity operator (list)");
      }
    }
  }
  // ---- environment file names (two_site_update_finite_vmps_impl.h:55-69)
  {
    Check(GenEnvTenName("r", 7, ".temp") == ".temp/renv7.gqten", "GenEnvTenName right");
    Check(GenEnvTenName("l", 12, "tmp/x") == "tmp/x/lenv12.gqten", "GenEnvTenName left");
    std::ofstream("env_probe.gqten") << "x";
    RemoveFile("env_probe.gqten");
    Check(!IsPathExist("env_probe.gqten"), "RemoveFile deletes the file");
  }
  // ---- corrupt flat data (tensor files, FFI input) is rejected before parsing
  {
    U1QN q0({QNCard("N", U1QNVal(0))}), q1({QNCard("N", U1QNVal(1))});
    IndexT o({QNSctT(q0, 2), QNSctT(q1, 3)}, OUT);
    Tensor t({InverseIndex(o), o});
    t.Random(q0);
    std::vector<int32_t> meta;
    std::vector<double> data;
    t.ToFlat(meta, data);
    Tensor back;
    back.FromFlat(meta, data);
    Check(back == t, "flat round trip");
    const size_t nblk = (size_t)meta[2], coords = meta.size() - 2 * nblk;
    Check(nblk == 2 && data.size() == 13, "flat layout of the probe tensor");
    auto rejects = [&](std::vector<int32_t> m, std::vector<double> d) {
      Tensor r;
      try {
        r.FromFlat(m, d);
      } catch (const std::runtime_error &) {
        return true;
      }
      return false;
    };
    auto d_short = data;
    d_short.pop_back();
    auto d_long = data;
    d_long.push_back(0.0);
    Check(rejects(meta, d_short) && rejects(meta, d_long), "element count differs from the blocks");
    auto m = meta;
    m[coords] = 2;  // sector 2 of a two-sector leg
    Check(rejects(m, data), "block coordinate out of range");
    m = meta;
    m[coords + 2] = m[coords];
    m[coords + 3] = m[coords + 1];
    Check(rejects(m, data), "repeated block");
    m = meta;
    m.pop_back();
    Check(rejects(m, data), "truncated block list");
    m = meta;
    m[3] = 7;
    Check(rejects(m, data), "bad direction");
    m = meta;
    m[1] = 2;
    Check(rejects(m, data), "QN width");
    m = meta;
    m[0] = 3;
    Check(rejects(m, data), "rank beyond the legs present");
    Check(rejects(std::vector<int32_t>{1, 1}, data), "header too short");
  }
  // ---- MPO
  {
    U1QN qn({QNCard("N", U1QNVal(0))});
    IndexT o({QNSctT(qn, 3)}, OUT), i({QNSctT(qn, 4)}, IN);
    Tensor a({o}), b({i, o}), c({i, o, o});
    a.Random(qn);
    b.Random(qn);
    c.Random(qn);
    for (int N : {0, 5}) {
      MPO<Tensor> mpo(N);
      Check(mpo.size() == (size_t)N, "MPO size");
      for (int k = 0; k < N; ++k) Check(mpo[k] == Tensor(), "MPO elements default-constructed");
    }
    MPO<Tensor> mpo(3);
    mpo[0] = a;
    mpo[1] = b;
    mpo[2] = c;
    Check(mpo[0] == a && mpo[1] == b && mpo[2] == c, "MPO element assignment");
    mpo[1] = c;
    Check(mpo[1] == c, "MPO element reassignment");
  }
  // ---- tensor arithmetic (scalar multiples and sums, as site operators are built)
  {
    U1QN q0({QNCard("N", U1QNVal(0))}), q1({QNCard("N", U1QNVal(1))});
    IndexT o({QNSctT(q1, 2), QNSctT(q0, 1)}, OUT);
    IndexT i = InverseIndex(o);
    Tensor a({i, o}), b({i, o});
    a({0, 0}) = 1.5;
    a({1, 1}) = -2.0;
    b({1, 1}) = 0.5;
    b({2, 2}) = 4.0;   // a block that a does not hold
    Tensor m = -3.0 * a, n = a * 2.0, neg = -a, s = a + b;
    Check(m(0, 0) == -4.5 && m(1, 1) == 6.0 && m(2, 2) == 0.0, "scalar * tensor");
    Check(n(0, 0) == 3.0 && n(1, 1) == -4.0, "tensor * scalar");
    Check(neg(0, 0) == -1.5 && neg(1, 1) == 2.0, "unary minus");
    Check(s(0, 0) == 1.5 && s(1, 1) == -1.5 && s(2, 2) == 4.0 && s(0, 1) == 0.0, "sum, including a new block");
    Check(a(0, 0) == 1.5 && a(2, 2) == 0.0 && b(0, 0) == 0.0, "operands unchanged");
    Tensor c = a;
    c += c;
    c *= 0.5;
    Check(c == a, "c += c; c *= 0.5 gives c back");
    Tensor w({o, o});
    bool threw = false;
    try {
      a += w;
    } catch (const std::invalid_argument &) {
      threw = true;
    }
    Check(threw, "sum of tensors with different indexes is rejected");
  }
  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
