// MPS / FiniteMPS with canonicalisation and the two initialisations of the
// reference.
//
// Follows include/gqmps2/one_dim_tn/mps/mps.h:31-107,
// one_dim_tn/mps/finite_mps/finite_mps.h:50-302 and finite_mps_init.h:216-330
// of GraceQ/MPS2 0.2-alpha.0.  Canonicalisation SVDs run on the device and
// build the canonical factor exactly orthonormal (U for left, Vt for right).
#ifndef MPS2_AMD_GQMPS2_FINITE_MPS_H
#define MPS2_AMD_GQMPS2_FINITE_MPS_H

#include <iomanip>

#include "gqmps2/framework.h"

namespace gqmps2 {

inline std::string GenMPSTenName(const std::string &mps_path, const size_t idx) {
  return mps_path + "/" + kMpsTenBaseName + std::to_string(idx) + "." + kGQTenFileSuffix;
}

template <typename TenElemT, typename QNT>
class MPS : public TenVec<GQTensor<TenElemT, QNT>> {
 public:
  using LocalTenT = GQTensor<TenElemT, QNT>;
  MPS(const SiteVec<TenElemT, QNT> &site_vec) : TenVec<LocalTenT>(site_vec.size), site_vec_(site_vec) {}
  const SiteVec<TenElemT, QNT> &GetSitesInfo() const { return site_vec_; }

  void Dump(const std::string &mps_path = kMpsPath) const {
    if (!IsPathExist(mps_path)) CreatPath(mps_path);
    for (size_t i = 0; i < this->size(); ++i) this->DumpTen(i, GenMPSTenName(mps_path, i));
  }
  void Dump(const std::string &mps_path = kMpsPath, const bool release_mem = false) {
    if (!IsPathExist(mps_path)) CreatPath(mps_path);
    for (size_t i = 0; i < this->size(); ++i) this->DumpTen(i, GenMPSTenName(mps_path, i), release_mem);
  }
  void Load(const std::string &mps_path = kMpsPath) {
    for (size_t i = 0; i < this->size(); ++i) this->LoadTen(i, GenMPSTenName(mps_path, i));
  }

 private:
  SiteVec<TenElemT, QNT> site_vec_;
};

enum class MPSTenCanoType { NONE, LEFT, RIGHT };
const int kUncentralizedCenterIdx = -1;

template <typename TenElemT, typename QNT>
class FiniteMPS : public MPS<TenElemT, QNT> {
 public:
  using LocalTenT = typename MPS<TenElemT, QNT>::LocalTenT;
  FiniteMPS(const SiteVec<TenElemT, QNT> &site_vec)
      : MPS<TenElemT, QNT>(site_vec), center_(kUncentralizedCenterIdx), cano_(site_vec.size, MPSTenCanoType::NONE) {}

  LocalTenT &operator[](const size_t idx) {
    cano_[idx] = MPSTenCanoType::NONE;
    center_ = kUncentralizedCenterIdx;
    return DuoVector<LocalTenT>::operator[](idx);
  }
  const LocalTenT &operator[](const size_t idx) const { return DuoVector<LocalTenT>::operator[](idx); }
  LocalTenT *&operator()(const size_t idx) {
    cano_[idx] = MPSTenCanoType::NONE;
    center_ = kUncentralizedCenterIdx;
    return DuoVector<LocalTenT>::operator()(idx);
  }
  const LocalTenT *operator()(const size_t idx) const { return DuoVector<LocalTenT>::operator()(idx); }

  void Centralize(const int target_center) {
    assert(target_center >= 0);
    int tail = (int)this->size() - 1;
    if (target_center != 0) LeftCanonicalize_(target_center - 1);
    if (target_center != tail) RightCanonicalize_(target_center + 1);
    center_ = target_center;
  }
  int GetCenter() const { return center_; }
  std::vector<MPSTenCanoType> GetTensCanoType() const { return cano_; }
  MPSTenCanoType GetTenCanoType(const size_t idx) const { return cano_[idx]; }

 private:
  int center_;
  std::vector<MPSTenCanoType> cano_;

  void LeftCanonicalize_(const size_t stop) {
    size_t start = 0;
    for (size_t i = 0; i <= stop; ++i) {
      start = i;
      if (cano_[i] != MPSTenCanoType::LEFT) break;
      if (i == stop) return;
    }
    for (size_t i = start; i <= stop; ++i) LeftCanonicalizeTen_(i);
  }
  // finite_mps.h:172-197
  void LeftCanonicalizeTen_(const size_t i) {
    size_t ldims = i == 0 ? 1 : 2;
    GQTensor<double, QNT> s;
    LocalTenT u, vt;
    const LocalTenT &A = DuoVector<LocalTenT>::operator[](i);
    size_t D = FullDim_(A, ldims);
    SVD(&A, ldims, Div(A), 0.0, D, D, &u, &s, &vt, nullptr, nullptr, SvdOrth::Left);
    DuoVector<LocalTenT>::operator[](i) = std::move(u);
    LocalTenT sv, next;
    Contract(&s, &vt, {{1}, {0}}, &sv);
    Contract(&sv, &DuoVector<LocalTenT>::operator[](i + 1), {{1}, {0}}, &next);
    DuoVector<LocalTenT>::operator[](i + 1) = std::move(next);
    cano_[i] = MPSTenCanoType::LEFT;
    cano_[i + 1] = MPSTenCanoType::NONE;
  }
  void RightCanonicalize_(const size_t stop) {
    size_t tail = this->size() - 1, start = tail;
    for (size_t i = tail; i >= stop; --i) {
      start = i;
      if (cano_[i] != MPSTenCanoType::RIGHT) break;
      if (i == stop) return;
    }
    for (size_t i = start; i >= stop; --i) RightCanonicalizeTen_(i);
  }
  // finite_mps.h:213-243
  void RightCanonicalizeTen_(const size_t i) {
    assert(i > 0);
    LocalTenT u, vt;
    GQTensor<double, QNT> s;
    const LocalTenT &A = DuoVector<LocalTenT>::operator[](i);
    auto qndiv = Div(A);
    size_t D = FullDim_(A, 1);
    SVD(&A, 1, qndiv - qndiv, 0.0, D, D, &u, &s, &vt, nullptr, nullptr, SvdOrth::Right);
    DuoVector<LocalTenT>::operator[](i) = std::move(vt);
    LocalTenT us, prev;
    Contract(&u, &s, {{1}, {0}}, &us);
    std::vector<size_t> ax = {(i - 1) == 0 ? (size_t)1 : (size_t)2};
    Contract(&DuoVector<LocalTenT>::operator[](i - 1), &us, {ax, {0}}, &prev);
    DuoVector<LocalTenT>::operator[](i - 1) = std::move(prev);
    cano_[i] = MPSTenCanoType::RIGHT;
    cano_[i - 1] = MPSTenCanoType::NONE;
  }
  // mock_gqten::SVD keeps D = max(lsize, rsize) (utilities.h:29-55)
  static size_t FullDim_(const LocalTenT &A, size_t ldims) {
    size_t l = 1, r = 1;
    auto sh = A.GetShape();
    for (size_t k = 0; k < sh.size(); ++k) (k < ldims ? l : r) *= sh[k];
    return l >= r ? l : r;
  }
};

// TruncateMPS (finite_mps.h:245-303): Centralize(0), normalise the first
// tensor, then one truncated device SVD per bond from left to right, each
// with a normalised S carried into the next tensor.  Leaves every tensor but
// the last left-canonical; the MPS counts as not centralised (the tensors are
// replaced through operator()).
template <typename TenElemT, typename QNT>
void TruncateMPS(FiniteMPS<TenElemT, QNT> &mps, const GQTEN_Double trunc_err, const size_t Dmin, const size_t Dmax) {
  using LocalTenT = GQTensor<TenElemT, QNT>;
  const size_t N = mps.size();
  if (N < 2) throw std::invalid_argument("TruncateMPS: at least two sites");
  mps.Centralize(0);
  mps[0].Normalize();
  for (size_t i = 0; i + 1 < N; ++i) {
    const size_t ldims = i == 0 ? 1 : 2;
    GQTensor<GQTEN_Double, QNT> s;
    LocalTenT vt;
    auto *u = new LocalTenT;
    double err = 0;
    size_t D = 0;
    const LocalTenT &A = mps[i];
    SVD(&A, ldims, Div(A), trunc_err, Dmin, Dmax, u, &s, &vt, &err, &D);
    if (Verbose())
      std::cout << "Truncate MPS bond " << std::setw(4) << i << " TruncErr = " << std::setprecision(2)
                << std::scientific << err << std::fixed << " D = " << std::setw(5) << D << std::scientific
                << std::endl;
    s.Normalize();
    delete mps(i);
    mps(i) = u;
    LocalTenT sv;
    Contract(&s, &vt, {{1}, {0}}, &sv);
    auto *next = new LocalTenT;
    Contract(&sv, mps(i + 1), {{1}, {0}}, next);
    delete mps(i + 1);
    mps(i + 1) = next;
  }
}

// finite_mps_init.h:216-262
template <typename TenElemT, typename QNT>
void DirectStateInitMps(FiniteMPS<TenElemT, QNT> &mps, const std::vector<size_t> &stat_labs, const QNT &zero_div) {
  using TenT = GQTensor<TenElemT, QNT>;
  using IndexT = Index<QNT>;
  using QNSctT = QNSector<QNT>;
  auto N = mps.size();
  assert(N == stat_labs.size());
  for (size_t i = 0; i < N; ++i) mps.dealloc(i);
  auto pb = mps.GetSitesInfo().sites;
  auto div = pb[0].GetQNSctFromActualCoor(stat_labs[0]).GetQn();
  for (size_t i = 1; i < N; ++i) div += pb[i].GetQNSctFromActualCoor(stat_labs[i]).GetQn();
  auto rvb_qn = div - pb[0].GetQNSctFromActualCoor(stat_labs[0]).GetQn();
  IndexT rvb({QNSctT(rvb_qn, 1)}, GQTenIndexDirType::OUT), lvb;
  mps(0) = new TenT({pb[0], rvb});
  (mps[0])({stat_labs[0], 0}) = 1;
  for (size_t i = 1; i < N - 1; ++i) {
    lvb = InverseIndex(rvb);
    rvb_qn = zero_div - pb[i].GetQNSctFromActualCoor(stat_labs[i]).GetQn() + lvb.GetQNSctFromActualCoor(0).GetQn();
    rvb = IndexT({QNSctT(rvb_qn, 1)}, GQTenIndexDirType::OUT);
    mps(i) = new TenT({lvb, pb[i], rvb});
    (mps[i])({0, stat_labs[i], 0}) = 1;
  }
  lvb = InverseIndex(rvb);
  mps(N - 1) = new TenT({lvb, pb[N - 1]});
  (mps[N - 1])({0, stat_labs[N - 1]}) = 1;
  mps.Centralize(0);
}

// finite_mps_init.h:269-330
template <typename TenElemT, typename QNT>
void ExtendDirectRandomInitMps(FiniteMPS<TenElemT, QNT> &mps, const std::vector<std::vector<size_t>> &stat_labs_set,
                               const QNT &zero_div, const size_t enlarged_dim) {
  using TenT = GQTensor<TenElemT, QNT>;
  using IndexT = Index<QNT>;
  using QNSctT = QNSector<QNT>;
  auto nfus = stat_labs_set.size();
  assert(nfus >= 1);
  auto N = mps.size();
  for (size_t i = 0; i < N; ++i) mps.dealloc(i);
  auto pb = mps.GetSitesInfo().sites;
  auto div = pb[0].GetQNSctFromActualCoor(stat_labs_set[0][0]).GetQn();
  for (size_t i = 1; i < N; ++i) div += pb[i].GetQNSctFromActualCoor(stat_labs_set[0][i]).GetQn();
  std::vector<QNSctT> secs;
  for (size_t j = 0; j < nfus; ++j)
    secs.push_back(QNSctT(div - pb[0].GetQNSctFromActualCoor(stat_labs_set[j][0]).GetQn(), enlarged_dim));
  IndexT rvb(secs, GQTenIndexDirType::OUT), lvb;
  secs.clear();
  mps(0) = new TenT({pb[0], rvb});
  mps[0].Random(div);
  for (size_t i = 1; i < N - 1; ++i) {
    lvb = InverseIndex(rvb);
    for (size_t j = 0; j < nfus; ++j) {
      auto q = zero_div - pb[i].GetQNSctFromActualCoor(stat_labs_set[j][i]).GetQn() +
               lvb.GetQNSctFromActualCoor(j * enlarged_dim).GetQn();
      secs.push_back(QNSctT(q, enlarged_dim));
    }
    rvb = IndexT(secs, GQTenIndexDirType::OUT);
    secs.clear();
    mps(i) = new TenT({lvb, pb[i], rvb});
    mps[i].Random(zero_div);
  }
  lvb = InverseIndex(rvb);
  mps(N - 1) = new TenT({lvb, pb[N - 1]});
  mps[N - 1].Random(zero_div);
  mps.Centralize(0);
  mps[0].Normalize();
}

}  // namespace gqmps2

#endif
