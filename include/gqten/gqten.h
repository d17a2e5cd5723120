// gqten-compatible tensor API over the mps2_amd device engine.
//
// Mirrors the subset of GraceQ/tensor used by GraceQ/MPS2's hot path
// (QN / QNCard / U1QNVal, QNSector, Index, GQTensor<GQTEN_Double, QNT>,
// Contract, SVD, LinearCombine, Dag, Div, Timer; call sites listed in
// include/mps2_amd.h) so that code written against the reference compiles
// unchanged.  Small tensors built element by element (operators, MPO tensors,
// initial MPS) live on the host until an operation needs them on the device;
// contractions, SVD, linear combinations, Dag and norms run on the GPU through
// the C ABI, while the set-up edits of user code (element writes, Random,
// Transpose, scalar multiples and sums) act on the host copy.  Element types
// GQTEN_Double and GQTEN_Complex (std::complex<double>); a contraction of a
// real tensor with a complex one (the real singular values times a complex
// factor) gives a complex tensor.
#ifndef MPS2_AMD_GQTEN_H
#define MPS2_AMD_GQTEN_H

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../mps2_amd.h"

namespace gqten {

using GQTEN_Double = double;
using GQTEN_Complex = std::complex<double>;

// Real part of a scalar (lanczos_solver_impl.h:64-67).
inline double Real(const double x) { return x; }
inline double Real(const GQTEN_Complex &x) { return x.real(); }

enum class GQTenIndexDirType { IN = 0, OUT = 1, NDIR = 2 };
// gqten's short names for the two directions (IndexT({...}, OUT) in user code)
constexpr GQTenIndexDirType IN = GQTenIndexDirType::IN;
constexpr GQTenIndexDirType OUT = GQTenIndexDirType::OUT;

// ---------------------------------------------------------------- errors / context
inline void CheckMps2(int rc, const char *what) {
  if (rc != MPS2_OK) {
    throw std::runtime_error(std::string(what) + ": " + mps2_last_error());
  }
}

// One engine context per process (device 0 unless MPS2_DEVICE is set).
inline mps2_ctx DefaultCtx() {
  static std::shared_ptr<mps2_ctx_s> ctx = [] {
    int dev = 0;
    if (const char *e = std::getenv("MPS2_DEVICE")) dev = std::atoi(e);
    mps2_ctx c = nullptr;
    CheckMps2(mps2_ctx_create(dev, &c), "mps2_ctx_create");
    return std::shared_ptr<mps2_ctx_s>(c, [](mps2_ctx p) { mps2_ctx_destroy(p); });
  }();
  return ctx.get();
}

// ---------------------------------------------------------------- QN
struct U1QNVal {
  int val;
  explicit U1QNVal(int v) : val(v) {}
};

struct QNCard {
  std::string name;
  int val;
  QNCard(const std::string &n, const U1QNVal &v) : name(n), val(v.val) {}
};

template <typename... QNVals>
class QN {
 public:
  static constexpr int kWidth = (int)sizeof...(QNVals);
  static_assert(kWidth >= 1 && kWidth <= 4, "1 to 4 U(1) components");
  QN() { vals_.fill(0); }
  explicit QN(const std::vector<QNCard> &cards) {
    vals_.fill(0);
    if ((int)cards.size() != kWidth) throw std::invalid_argument("QN: wrong number of QNCards");
    for (int i = 0; i < kWidth; ++i) {
      vals_[i] = cards[i].val;
      names_[i] = cards[i].name;
    }
  }
  int operator[](int i) const { return vals_[i]; }
  int &operator[](int i) { return vals_[i]; }
  QN operator+(const QN &o) const {
    QN r(*this);
    for (int i = 0; i < kWidth; ++i) r.vals_[i] += o.vals_[i];
    return r;
  }
  QN operator-(const QN &o) const {
    QN r(*this);
    for (int i = 0; i < kWidth; ++i) r.vals_[i] -= o.vals_[i];
    return r;
  }
  QN operator-() const {
    QN r(*this);
    for (int i = 0; i < kWidth; ++i) r.vals_[i] = -r.vals_[i];
    return r;
  }
  QN &operator+=(const QN &o) { return *this = *this + o; }
  QN &operator-=(const QN &o) { return *this = *this - o; }
  bool operator==(const QN &o) const { return vals_ == o.vals_; }
  bool operator!=(const QN &o) const { return !(*this == o); }
  bool operator<(const QN &o) const { return vals_ < o.vals_; }

 private:
  std::array<int, kWidth> vals_;
  std::array<std::string, kWidth> names_;
};

template <typename QNT>
class QNSector {
 public:
  QNSector() : dim_(0) {}
  QNSector(const QNT &qn, size_t dim) : qn_(qn), dim_(dim) {}
  const QNT &GetQn() const { return qn_; }
  size_t dim() const { return dim_; }
  size_t GetDegeneracy() const { return dim_; }
  bool operator==(const QNSector &o) const { return qn_ == o.qn_ && dim_ == o.dim_; }
  bool operator!=(const QNSector &o) const { return !(*this == o); }

 private:
  QNT qn_;
  size_t dim_;
};

template <typename QNT>
using QNSectorVec = std::vector<QNSector<QNT>>;

template <typename QNT>
class Index {
 public:
  Index() : dir_(GQTenIndexDirType::NDIR), dim_(0) {}
  Index(const QNSectorVec<QNT> &s, GQTenIndexDirType d) : qnscts_(s), dir_(d), dim_(0) {
    for (auto &x : qnscts_) dim_ += x.dim();
  }
  size_t dim() const { return dim_; }
  GQTenIndexDirType GetDir() const { return dir_; }
  const QNSectorVec<QNT> &GetQNScts() const { return qnscts_; }
  size_t GetQNSctNum() const { return qnscts_.size(); }
  const QNSector<QNT> &GetQNSct(size_t i) const { return qnscts_[i]; }
  // Sector holding the actual (dense) coordinate.
  const QNSector<QNT> &GetQNSctFromActualCoor(size_t coor) const {
    size_t s, off;
    Locate(coor, s, off);
    return qnscts_[s];
  }
  void Locate(size_t coor, size_t &sct, size_t &off) const {
    size_t acc = 0;
    for (size_t i = 0; i < qnscts_.size(); ++i) {
      if (coor < acc + qnscts_[i].dim()) {
        sct = i;
        off = coor - acc;
        return;
      }
      acc += qnscts_[i].dim();
    }
    throw std::out_of_range("Index: coordinate out of range");
  }
  void Inverse() {
    if (dir_ == GQTenIndexDirType::IN) dir_ = GQTenIndexDirType::OUT;
    else if (dir_ == GQTenIndexDirType::OUT) dir_ = GQTenIndexDirType::IN;
  }
  bool operator==(const Index &o) const { return dir_ == o.dir_ && qnscts_ == o.qnscts_; }
  bool operator!=(const Index &o) const { return !(*this == o); }

 private:
  QNSectorVec<QNT> qnscts_;
  GQTenIndexDirType dir_;
  size_t dim_;
};

template <typename QNT>
using IndexVec = std::vector<Index<QNT>>;

template <typename QNT>
Index<QNT> InverseIndex(const Index<QNT> &idx) {
  Index<QNT> r(idx);
  r.Inverse();
  return r;
}

// ---------------------------------------------------------------- device handle
struct DevTen {
  mps2_ten h = nullptr;
  explicit DevTen(mps2_ten t) : h(t) {}
  ~DevTen() {
    if (h) mps2_ten_free(h);
  }
  DevTen(const DevTen &) = delete;
  DevTen &operator=(const DevTen &) = delete;
};

// ---------------------------------------------------------------- GQTensor
template <typename ElemT, typename QNT>
class GQTensor {
  static_assert(std::is_same<ElemT, GQTEN_Double>::value || std::is_same<ElemT, GQTEN_Complex>::value,
                "GQTensor element type: GQTEN_Double or GQTEN_Complex");

 public:
  using BlockKey = std::vector<int32_t>;
  static constexpr bool kComplex = std::is_same<ElemT, GQTEN_Complex>::value;
  static constexpr int kDoubles = kComplex ? 2 : 1;   // doubles per element in the flat format

  GQTensor() : is_default_(true) {}
  explicit GQTensor(const IndexVec<QNT> &idxs) : indexes_(idxs), is_default_(false) {}

  size_t Rank() const { return indexes_.size(); }
  const IndexVec<QNT> &GetIndexes() const { return indexes_; }
  const Index<QNT> &GetIndex(size_t i) const { return indexes_[i]; }
  std::vector<size_t> GetShape() const {
    std::vector<size_t> s;
    for (auto &i : indexes_) s.push_back(i.dim());
    return s;
  }
  bool IsDefault() const { return is_default_; }

  // ---- element access (host side)
  ElemT &operator()(const std::vector<size_t> &coors) {
    ToHost();
    dev_.reset();
    return ElemRef(coors);
  }
  template <typename... Ints>
  ElemT &operator()(size_t c0, Ints... cs) {
    return (*this)(std::vector<size_t>{c0, (size_t)cs...});
  }
  // read-only element access on a const tensor (as gqten's const operator())
  ElemT operator()(const std::vector<size_t> &coors) const { return GetElem(coors); }
  template <typename... Ints>
  ElemT operator()(size_t c0, Ints... cs) const {
    return GetElem(std::vector<size_t>{c0, (size_t)cs...});
  }
  ElemT GetElem(const std::vector<size_t> &coors) const {
    ToHost();
    BlockKey k;
    size_t off;
    Addr(coors, k, off);
    auto it = hblk_.find(k);
    return it == hblk_.end() ? ElemT(0) : it->second[off];
  }
  // rank-0 scalar
  ElemT operator()() const {
    ToHost();
    auto it = hblk_.find(BlockKey());
    return it == hblk_.end() ? ElemT(0) : it->second[0];
  }

  double Normalize() {
    ToDevice();
    double n = 0;
    CheckMps2(mps2_normalize(dev_->h, &n), "Normalize");
    hblk_.clear();
    host_valid_ = false;
    return n;
  }

  void Random(const QNT &div);
  void Transpose(const std::vector<size_t> &perm);

  // gqten's tensor arithmetic: scalar multiples and sums of tensors with equal
  // indexes (used to build site operators, e.g. -t * c_up in
  // test_two_site_update_finite_vmps.cc:898-901).  Computed on the host copy.
  GQTensor &operator*=(const ElemT s) {
    ToHost();
    dev_.reset();
    for (auto &kv : hblk_)
      for (auto &x : kv.second) x *= s;
    return *this;
  }
  GQTensor operator*(const ElemT s) const {
    GQTensor r(*this);
    r *= s;
    return r;
  }
  GQTensor operator-() const { return (*this) * ElemT(-1.0); }
  GQTensor &operator+=(const GQTensor &o) {
    if (is_default_ || o.is_default_ || indexes_ != o.indexes_)
      throw std::invalid_argument("GQTensor sum: the tensors must have the same indexes");
    ToHost();
    o.ToHost();
    dev_.reset();
    for (auto &kv : o.hblk_) {
      auto it = hblk_.find(kv.first);
      if (it == hblk_.end()) {
        hblk_.emplace(kv.first, kv.second);
        continue;
      }
      for (size_t e = 0; e < kv.second.size(); ++e) it->second[e] += kv.second[e];
    }
    return *this;
  }
  GQTensor operator+(const GQTensor &o) const {
    GQTensor r(*this);
    r += o;
    return r;
  }

  bool operator==(const GQTensor &o) const {
    if (is_default_ || o.is_default_) return is_default_ == o.is_default_;
    if (indexes_ != o.indexes_) return false;
    ToHost();
    o.ToHost();
    return hblk_ == o.hblk_;
  }
  bool operator!=(const GQTensor &o) const { return !(*this == o); }

  // ---- engine interop
  mps2_ten Handle() const {
    ToDevice();
    return dev_->h;
  }
  // Adopt a device tensor produced by the engine; indexes are rebuilt from it.
  void AdoptDevice(mps2_ten h) {
    int cz = 0;
    CheckMps2(mps2_ten_is_complex(h, &cz), "is_complex");
    if ((cz != 0) != kComplex) {
      mps2_ten_free(h);
      throw std::runtime_error(kComplex ? "GQTensor<GQTEN_Complex>: the engine returned a real tensor"
                                        : "GQTensor<GQTEN_Double>: the engine returned a complex tensor "
                                          "(a complex operand needs a GQTEN_Complex result)");
    }
    dev_ = std::make_shared<DevTen>(h);
    is_default_ = false;
    hblk_.clear();
    host_valid_ = false;
    int64_t nm = 0;
    CheckMps2(mps2_ten_export_sizes(h, &nm, nullptr), "export_sizes");
    std::vector<int32_t> meta(nm);
    CheckMps2(mps2_ten_export(h, meta.data(), nullptr), "export meta");
    ParseLegs(meta);
  }
  bool OnDevice() const { return (bool)dev_; }
  void ReleaseHost() const {
    if (dev_) {
      hblk_.clear();
      host_valid_ = false;
    }
  }
  // Flat exchange format (include/mps2_amd.h).
  void ToFlat(std::vector<int32_t> &meta, std::vector<double> &data) const;
  void FromFlat(const std::vector<int32_t> &meta, const std::vector<double> &data);

  void ToDevice() const {
    if (dev_) return;
    if (is_default_) throw std::runtime_error("operation on a default-constructed GQTensor");
    std::vector<int32_t> meta;
    std::vector<double> data;
    ToFlat(meta, data);
    if (data.empty()) data.push_back(0.0);
    if (data.size() < 2) data.resize(2, 0.0);
    mps2_ten h = nullptr;
    if (kComplex) CheckMps2(mps2_ten_import_complex(DefaultCtx(), meta.data(), data.data(), &h), "upload");
    else CheckMps2(mps2_ten_import(DefaultCtx(), meta.data(), data.data(), &h), "upload");
    dev_ = std::make_shared<DevTen>(h);
  }
  void ToHost() const {
    if (host_valid_ || !dev_) return;
    int64_t nm = 0, nd = 0;
    CheckMps2(mps2_ten_export_sizes(dev_->h, &nm, &nd), "export_sizes");
    std::vector<int32_t> meta(nm);
    std::vector<double> data(std::max<int64_t>(nd, 1));
    CheckMps2(mps2_ten_export(dev_->h, meta.data(), data.data()), "download");
    LoadBlocks(meta, data);
    host_valid_ = true;
  }

 private:
  IndexVec<QNT> indexes_;
  bool is_default_ = true;
  mutable std::map<BlockKey, std::vector<ElemT>> hblk_;
  mutable bool host_valid_ = true;
  mutable std::shared_ptr<DevTen> dev_;

  size_t BlockSize(const BlockKey &k) const {
    size_t s = 1;
    for (size_t i = 0; i < k.size(); ++i) s *= indexes_[i].GetQNSct(k[i]).dim();
    return s;
  }
  void Addr(const std::vector<size_t> &coors, BlockKey &k, size_t &off) const {
    if (coors.size() != indexes_.size()) throw std::invalid_argument("GQTensor: wrong number of coordinates");
    k.resize(coors.size());
    off = 0;
    for (size_t i = 0; i < coors.size(); ++i) {
      size_t s, o;
      indexes_[i].Locate(coors[i], s, o);
      k[i] = (int32_t)s;
      off = off * indexes_[i].GetQNSct(s).dim() + o;
    }
  }
  ElemT &ElemRef(const std::vector<size_t> &coors) {
    BlockKey k;
    size_t off;
    Addr(coors, k, off);
    auto &b = hblk_[k];
    if (b.empty()) b.assign(BlockSize(k), ElemT(0));
    return b[off];
  }
  void ParseLegs(const std::vector<int32_t> &m) {
    int p = 0;
    int rank = m[p++];
    int qw = m[p++];
    p++;
    if (qw != QNT::kWidth) throw std::runtime_error("QN width mismatch");
    indexes_.clear();
    for (int l = 0; l < rank; ++l) {
      auto dir = m[p++] == 1 ? GQTenIndexDirType::OUT : GQTenIndexDirType::IN;
      int ns = m[p++];
      QNSectorVec<QNT> secs;
      for (int s = 0; s < ns; ++s) {
        QNT q;
        for (int k = 0; k < qw; ++k) q[k] = m[p++];
        secs.emplace_back(q, (size_t)m[p++]);
      }
      indexes_.emplace_back(secs, dir);
    }
  }
  void LoadBlocks(const std::vector<int32_t> &m, const std::vector<double> &data) const {
    int p = 0;
    int rank = m[p++];
    int qw = m[p++];
    int nblk = m[p++];
    for (int l = 0; l < rank; ++l) {
      p++;
      int ns = m[p++];
      p += ns * (qw + 1);
    }
    hblk_.clear();
    size_t off = 0;
    for (int b = 0; b < nblk; ++b) {
      BlockKey k(m.begin() + p, m.begin() + p + rank);
      p += rank;
      size_t sz = BlockSize(k);
      std::vector<ElemT> v(sz);
      for (size_t e = 0; e < sz; ++e) v[e] = ElemFrom(&data[off + kDoubles * e]);
      hblk_[k] = std::move(v);
      off += kDoubles * sz;
    }
  }
  static ElemT ElemFrom(const double *x) {
    if constexpr (kComplex) return ElemT(x[0], x[1]);
    else return x[0];
  }
};

// Seeded generator behind GQTensor::Random (uniform in [0, 1)).
inline std::mt19937_64 &RandomEngine() {
  static std::mt19937_64 g(0x5EEDull);
  return g;
}
inline void SetRandomSeed(uint64_t s) { RandomEngine().seed(s); }

template <typename QNT>
QNT BlockDiv(const IndexVec<QNT> &idx, const std::vector<int32_t> &k) {
  QNT d;
  for (size_t i = 0; i < idx.size(); ++i) {
    const QNT &q = idx[i].GetQNSct(k[i]).GetQn();
    if (idx[i].GetDir() == GQTenIndexDirType::OUT) d += q;
    else d -= q;
  }
  return d;
}

template <typename ElemT, typename QNT>
void GQTensor<ElemT, QNT>::Random(const QNT &div) {
  ToHost();
  dev_.reset();
  hblk_.clear();
  host_valid_ = true;
  size_t r = indexes_.size();
  BlockKey k(r, 0);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  if (r == 0) return;
  while (true) {
    if (BlockDiv(indexes_, k) == div) {
      auto &b = hblk_[k];
      b.resize(BlockSize(k));
      for (auto &x : b) {
        if constexpr (kComplex) {
          const double re = u(RandomEngine());
          x = ElemT(re, u(RandomEngine()));
        } else {
          x = u(RandomEngine());
        }
      }
    }
    int i = (int)r - 1;
    for (; i >= 0; --i) {
      if (++k[i] < (int32_t)indexes_[i].GetQNSctNum()) break;
      k[i] = 0;
    }
    if (i < 0) break;
  }
}

template <typename ElemT, typename QNT>
void GQTensor<ElemT, QNT>::Transpose(const std::vector<size_t> &perm) {
  ToHost();
  dev_.reset();
  IndexVec<QNT> ni;
  for (auto p : perm) ni.push_back(indexes_[p]);
  std::map<BlockKey, std::vector<ElemT>> nb;
  size_t r = perm.size();
  for (auto &kv : hblk_) {
    BlockKey k(r);
    std::vector<size_t> od(r), nd(r);
    for (size_t i = 0; i < r; ++i) od[i] = indexes_[i].GetQNSct(kv.first[i]).dim();
    for (size_t i = 0; i < r; ++i) {
      k[i] = kv.first[perm[i]];
      nd[i] = od[perm[i]];
    }
    std::vector<ElemT> out(kv.second.size());
    std::vector<size_t> ostr(r, 1);
    for (int i = (int)r - 2; i >= 0; --i) ostr[i] = ostr[i + 1] * od[i + 1];
    std::vector<size_t> idx(r, 0);
    for (size_t e = 0; e < out.size(); ++e) {
      size_t so = 0;
      for (size_t i = 0; i < r; ++i) so += idx[i] * ostr[perm[i]];
      out[e] = kv.second[so];
      for (int i = (int)r - 1; i >= 0; --i) {
        if (++idx[i] < nd[i]) break;
        idx[i] = 0;
      }
    }
    nb[k] = std::move(out);
  }
  indexes_ = ni;
  hblk_ = std::move(nb);
}

template <typename ElemT, typename QNT>
void GQTensor<ElemT, QNT>::ToFlat(std::vector<int32_t> &meta, std::vector<double> &data) const {
  ToHost();
  meta.clear();
  data.clear();
  meta.push_back((int32_t)indexes_.size());
  meta.push_back(QNT::kWidth);
  meta.push_back((int32_t)hblk_.size());
  for (auto &idx : indexes_) {
    meta.push_back(idx.GetDir() == GQTenIndexDirType::OUT ? 1 : 0);
    meta.push_back((int32_t)idx.GetQNSctNum());
    for (auto &s : idx.GetQNScts()) {
      for (int k = 0; k < QNT::kWidth; ++k) meta.push_back(s.GetQn()[k]);
      meta.push_back((int32_t)s.dim());
    }
  }
  for (auto &kv : hblk_)
    for (auto c : kv.first) meta.push_back(c);
  for (auto &kv : hblk_)
    for (const ElemT &x : kv.second) {
      if constexpr (kComplex) {
        data.push_back(x.real());
        data.push_back(x.imag());
      } else {
        data.push_back(x);
      }
    }
}

// Structure check of flat input from outside the engine (tensor files, FFI
// callers) before it is parsed: header, legs, sectors, block coordinates in
// range and not repeated, and exactly as many elements as the blocks hold.
template <typename QNT>
inline void CheckFlat(const std::vector<int32_t> &m, const size_t ndata) {
  auto bad = [](const char *what) { throw std::runtime_error(std::string("malformed tensor data: ") + what); };
  if (m.size() < 3) bad("header");
  const int64_t rank = m[0], qw = m[1], nblk = m[2];
  if (rank < 0 || rank > 64) bad("rank");
  if (qw != QNT::kWidth) bad("QN width");
  if (nblk < 0) bad("block count");
  size_t p = 3;
  std::vector<std::vector<int64_t>> dims((size_t)rank);
  for (int64_t l = 0; l < rank; ++l) {
    if (p + 2 > m.size()) bad("legs");
    if (m[p] != 0 && m[p] != 1) bad("direction");
    const int64_t ns = m[p + 1];
    p += 2;
    if (ns < 0 || (int64_t)((m.size() - p) / (size_t)(qw + 1)) < ns) bad("sectors");
    for (int64_t s = 0; s < ns; ++s) {
      p += (size_t)qw;
      if (m[p] < 0) bad("sector dimension");
      dims[l].push_back(m[p++]);
    }
  }
  if ((int64_t)(m.size() - p) != nblk * rank) bad("block list");
  std::set<std::vector<int32_t>> seen;
  size_t total = 0;
  for (int64_t b = 0; b < nblk; ++b) {
    std::vector<int32_t> k(m.begin() + p, m.begin() + p + rank);
    p += (size_t)rank;
    size_t sz = 1;
    for (int64_t l = 0; l < rank; ++l) {
      if (k[l] < 0 || k[l] >= (int64_t)dims[l].size()) bad("block coordinate");
      sz *= (size_t)dims[l][k[l]];
      if (sz > ndata) bad("data size");
    }
    if (!seen.insert(k).second) bad("repeated block");
    total += sz;
    if (total > ndata) bad("data size");
  }
  if (total != ndata) bad("data size");
}

template <typename ElemT, typename QNT>
void GQTensor<ElemT, QNT>::FromFlat(const std::vector<int32_t> &meta, const std::vector<double> &data) {
  if (data.size() % kDoubles) throw std::runtime_error("malformed tensor data: data size");
  CheckFlat<QNT>(meta, data.size() / kDoubles);
  is_default_ = false;
  dev_.reset();
  ParseLegs(meta);
  LoadBlocks(meta, data);
  host_valid_ = true;
}

// ---------------------------------------------------------------- free functions
template <typename ElemT, typename QNT>
GQTensor<ElemT, QNT> operator*(const double s, const GQTensor<ElemT, QNT> &t) {
  return t * ElemT(s);
}
template <typename QNT>
GQTensor<GQTEN_Complex, QNT> operator*(const GQTEN_Complex s, const GQTensor<GQTEN_Complex, QNT> &t) {
  return t * s;
}

// Contract of two tensors of either element type (a real operand with a
// complex one gives a complex result, which `out` must hold).
template <typename TA, typename TB, typename TO>
void Contract(const TA *a, const TB *b, const std::vector<std::vector<size_t>> &axes, TO *out) {
  std::vector<int32_t> xa(axes[0].begin(), axes[0].end()), xb(axes[1].begin(), axes[1].end());
  mps2_ten h = nullptr;
  CheckMps2(mps2_contract(a->Handle(), b->Handle(), (int)xa.size(), xa.data(), xb.data(), &h), "Contract");
  out->AdoptDevice(h);
}

template <typename ElemT, typename QNT>
GQTensor<ElemT, QNT> Dag(const GQTensor<ElemT, QNT> &t) {
  mps2_ten h = nullptr;
  CheckMps2(mps2_ten_dag(t.Handle(), &h), "Dag");
  GQTensor<ElemT, QNT> r;
  r.AdoptDevice(h);
  return r;
}

template <typename ElemT, typename QNT>
QNT Div(const GQTensor<ElemT, QNT> &t) {
  if (!t.OnDevice()) {
    std::vector<int32_t> meta;
    std::vector<double> data;
    t.ToFlat(meta, data);
    int r = meta[0], qw = meta[1], nb = meta[2];
    QNT d;
    if (nb == 0) return d;
    int p = 3;
    for (int l = 0; l < r; ++l) {
      p++;
      int ns = meta[p++];
      p += ns * (qw + 1);
    }
    std::vector<int32_t> k(meta.begin() + p, meta.begin() + p + r);
    return BlockDiv(t.GetIndexes(), k);
  }
  int32_t q[4] = {0, 0, 0, 0};
  CheckMps2(mps2_ten_div(t.Handle(), q), "Div");
  QNT d;
  for (int i = 0; i < QNT::kWidth; ++i) d[i] = q[i];
  return d;
}

// out = beta*out + sum_i coefs[i]*tens[i]
template <typename ElemT, typename QNT>
void LinearCombine(const std::vector<ElemT> &coefs, const std::vector<GQTensor<ElemT, QNT> *> &tens,
                   const ElemT beta, GQTensor<ElemT, QNT> *out) {
  std::vector<mps2_ten> hs;
  for (auto *t : tens) hs.push_back(t->Handle());
  if (out->IsDefault() || out->GetIndexes().empty()) {
    // start from a zero tensor with the indexes of the first operand
    mps2_ten z = nullptr;
    CheckMps2(mps2_ten_copy(hs[0], &z), "LinearCombine");
    GQTensor<ElemT, QNT> tmp;
    tmp.AdoptDevice(z);
    std::vector<double> zc{0.0};
    mps2_ten h0 = tmp.Handle();
    CheckMps2(mps2_linear_combine(0, zc.data(), nullptr, 0.0, h0), "LinearCombine");
    *out = tmp;
  }
  // the engine takes real coefficients (those of the Lanczos recurrence,
  // lanczos_solver_impl.h:151-158,173,215, are real for either element type)
  std::vector<double> c;
  for (const ElemT &x : coefs) {
    if (std::imag(x) != 0.0) throw std::invalid_argument("LinearCombine: complex coefficients are not supported");
    c.push_back(std::real(x));
  }
  if (std::imag(beta) != 0.0) throw std::invalid_argument("LinearCombine: complex coefficients are not supported");
  CheckMps2(mps2_linear_combine((int)hs.size(), c.data(), hs.data(), std::real(beta), out->Handle()), "LinearCombine");
  out->ReleaseHost();
}

template <typename ElemT, typename QNT>
void LinearCombine(const size_t n, const ElemT *coefs, const std::vector<GQTensor<ElemT, QNT> *> &tens,
                   const ElemT beta, GQTensor<ElemT, QNT> *out) {
  std::vector<ElemT> c(coefs, coefs + n);
  std::vector<GQTensor<ElemT, QNT> *> t(tens.begin(), tens.begin() + n);
  LinearCombine(c, t, beta, out);
}

enum class SvdOrth { Left = MPS2_ORTH_LEFT, Right = MPS2_ORTH_RIGHT };

// Truncated SVD with an explicit choice of the exactly-orthonormal factor.
template <typename ElemT, typename QNT>
void SVD(const GQTensor<ElemT, QNT> *t, const size_t ldims, const QNT &lqndiv, const double trunc_err,
         const size_t Dmin, const size_t Dmax, GQTensor<ElemT, QNT> *u, GQTensor<double, QNT> *s,
         GQTensor<ElemT, QNT> *vt, double *actual_trunc_err, size_t *D, SvdOrth orth) {
  int32_t q[4] = {0, 0, 0, 0};
  for (int i = 0; i < QNT::kWidth; ++i) q[i] = lqndiv[i];
  mps2_ten hu = nullptr, hs = nullptr, hv = nullptr;
  double err = 0;
  int64_t d = 0;
  CheckMps2(mps2_svd(t->Handle(), (int)ldims, q, trunc_err, (int64_t)Dmin, (int64_t)Dmax, (int)orth, &hu, &hs, &hv,
                     &err, &d),
            "SVD");
  u->AdoptDevice(hu);
  s->AdoptDevice(hs);
  vt->AdoptDevice(hv);
  if (actual_trunc_err) *actual_trunc_err = err;
  if (D) *D = (size_t)d;
}

// gqten signature (U exactly orthonormal).
template <typename ElemT, typename QNT>
void SVD(const GQTensor<ElemT, QNT> *t, const size_t ldims, const QNT &lqndiv, const double trunc_err,
         const size_t Dmin, const size_t Dmax, GQTensor<ElemT, QNT> *u, GQTensor<double, QNT> *s,
         GQTensor<ElemT, QNT> *vt, double *actual_trunc_err, size_t *D) {
  SVD(t, ldims, lqndiv, trunc_err, Dmin, Dmax, u, s, vt, actual_trunc_err, D, SvdOrth::Left);
}

// ---------------------------------------------------------------- Timer
class Timer {
 public:
  explicit Timer(const std::string &notes) : notes_(notes), start_(Clock::now()) {}
  void Restart() { start_ = Clock::now(); }
  double Elapsed() const { return std::chrono::duration<double>(Clock::now() - start_).count(); }
  double PrintElapsed(size_t precision = 5) const {
    double e = Elapsed();
    std::cout << "[timing] " << notes_ << "\t" << std::setw(precision + 3) << std::setprecision(precision)
              << std::fixed << e << std::endl;
    return e;
  }

 private:
  using Clock = std::chrono::steady_clock;
  std::string notes_;
  Clock::time_point start_;
};

}  // namespace gqten

#endif  // MPS2_AMD_GQTEN_H
