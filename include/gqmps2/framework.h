// gqmps2 framework pieces: constants, file helpers, SiteVec, DuoVector /
// TenVec containers, MPO alias.
//
// Follows include/gqmps2/{consts.h:26-33, utilities.h:76-91, site_vec.h:35-132,
// one_dim_tn/framework/duovector.h:33-222, one_dim_tn/framework/ten_vec.h:35-86,
// one_dim_tn/mpo/mpo.h:25} of GraceQ/MPS2 0.2-alpha.0.  Tensor files use this
// build's own versioned binary format (the gqten stream format is not
// reproducible here); the directory layout and file names are the reference's.
#ifndef MPS2_AMD_GQMPS2_FRAMEWORK_H
#define MPS2_AMD_GQMPS2_FRAMEWORK_H

#include <sys/stat.h>
#include <sys/types.h>

#include <cassert>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "gqten/gqten.h"

namespace gqmps2 {
using namespace gqten;

const std::string kMpsPath = "mps";
const std::string kRuntimeTempPath = ".temp";
const std::string kEnvFileBaseName = "env";
const std::string kMpsTenBaseName = "mps_ten";
const std::string kGQTenFileSuffix = "gqten";
const int kLanczEnergyOutputPrecision = 16;
const std::vector<size_t> kNullUintVec;
const std::vector<std::vector<size_t>> kNullUintVecVec;

// Console output of the library (MPO bond dimensions, per-update lines) as in
// the reference; FFI drivers that print machine-readable results switch it off.
inline bool &Verbose() {
  static bool v = true;
  return v;
}

inline bool IsPathExist(const std::string &path) {
  struct stat buffer;
  return stat(path.c_str(), &buffer) == 0;
}

inline void CreatPath(const std::string &path) {
  if (mkdir(path.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH) == -1) {
    std::cout << "error creating directory!" << std::endl;
    exit(1);
  }
}

// ---------------------------------------------------------------- tensor files
// Layout: "MPS2TEN1", int64 nmeta, int32 meta[nmeta], int64 ndata, double data[ndata]
// (GQTEN_Complex tensors: data holds (real, imaginary) pairs, interleaved).
template <typename TenT>
inline void WriteGQTensorTOFile(const TenT &t, const std::string &file) {
  std::vector<int32_t> meta;
  std::vector<double> data;
  t.ToFlat(meta, data);
  std::ofstream ofs(file, std::ofstream::binary);
  const char magic[8] = {'M', 'P', 'S', '2', 'T', 'E', 'N', '1'};
  ofs.write(magic, 8);
  int64_t nm = (int64_t)meta.size(), nd = (int64_t)data.size();
  ofs.write(reinterpret_cast<const char *>(&nm), sizeof(nm));
  ofs.write(reinterpret_cast<const char *>(meta.data()), nm * sizeof(int32_t));
  ofs.write(reinterpret_cast<const char *>(&nd), sizeof(nd));
  ofs.write(reinterpret_cast<const char *>(data.data()), nd * sizeof(double));
  if (!ofs) throw std::runtime_error("failed to write " + file);
}

template <typename TenT>
inline void ReadGQTensorFromFile(TenT &t, const std::string &file) {
  std::ifstream ifs(file, std::ifstream::binary);
  if (!ifs) throw std::runtime_error("cannot open " + file);
  char magic[8];
  ifs.read(magic, 8);
  if (!ifs || std::string(magic, 8) != "MPS2TEN1") throw std::runtime_error("not a tensor file: " + file);
  int64_t nm = 0, nd = 0;
  ifs.read(reinterpret_cast<char *>(&nm), sizeof(nm));
  if (!ifs || nm < 3 || nm > (1ll << 40)) throw std::runtime_error("corrupt tensor file: " + file);
  std::vector<int32_t> meta(nm);
  ifs.read(reinterpret_cast<char *>(meta.data()), nm * sizeof(int32_t));
  ifs.read(reinterpret_cast<char *>(&nd), sizeof(nd));
  if (!ifs || nd < 0 || nd > (1ll << 40)) throw std::runtime_error("corrupt tensor file: " + file);
  std::vector<double> data(nd);
  ifs.read(reinterpret_cast<char *>(data.data()), nd * sizeof(double));
  if (!ifs) throw std::runtime_error("truncated tensor file: " + file);
  t.FromFlat(meta, data);
}

// ---------------------------------------------------------------- SiteVec
template <typename QNT>
inline Index<QNT> SetIndexDirOut(const Index<QNT> &idx) {
  return idx.GetDir() == GQTenIndexDirType::OUT ? idx : InverseIndex(idx);
}

template <typename TenElemT, typename QNT>
GQTensor<TenElemT, QNT> GenIdOp(const Index<QNT> &idx_out) {
  auto idx_in = InverseIndex(idx_out);
  GQTensor<TenElemT, QNT> id_op({idx_in, idx_out});
  for (size_t i = 0; i < idx_out.dim(); ++i) id_op(i, i) = 1.0;
  return id_op;
}

template <typename TenElemT, typename QNT>
class SiteVec {
 public:
  using TenT = GQTensor<TenElemT, QNT>;
  SiteVec(const size_t N, const Index<QNT> &local_hilbert_space) {
    assert(N > 0);
    size = N;
    sites = IndexVec<QNT>(N, SetIndexDirOut(local_hilbert_space));
    GenIdOpsVec_();
  }
  SiteVec(const IndexVec<QNT> &local_hilbert_spaces) {
    size = local_hilbert_spaces.size();
    assert(size > 0);
    for (auto &i : local_hilbert_spaces) sites.emplace_back(SetIndexDirOut(i));
    GenIdOpsVec_();
  }
  size_t size;
  IndexVec<QNT> sites;
  std::vector<TenT> id_ops;

 private:
  void GenIdOpsVec_() {
    for (size_t i = 0; i < size; ++i) id_ops.emplace_back(GenIdOp<TenElemT, QNT>(sites[i]));
  }
};

// ---------------------------------------------------------------- DuoVector / TenVec
// Fixed-size vector of owned element pointers; nullptr = not resident.
template <typename ElemT>
class DuoVector {
 public:
  DuoVector() {}
  explicit DuoVector(const size_t size) : raw_(size, nullptr) {}
  DuoVector(const DuoVector &o) : raw_(o.size(), nullptr) {
    for (size_t i = 0; i < o.size(); ++i)
      if (o.raw_[i]) raw_[i] = new ElemT(*o.raw_[i]);
  }
  DuoVector &operator=(const DuoVector &o) {
    if (this == &o) return *this;
    clear();
    raw_.assign(o.size(), nullptr);
    for (size_t i = 0; i < o.size(); ++i)
      if (o.raw_[i]) raw_[i] = new ElemT(*o.raw_[i]);
    return *this;
  }
  DuoVector(DuoVector &&o) noexcept : raw_(std::move(o.raw_)) {}
  virtual ~DuoVector() { clear(); }

  // mutable access default-constructs a missing element (duovector.h:121-131)
  ElemT &operator[](const size_t i) {
    if (!raw_[i]) raw_[i] = new ElemT;
    return *raw_[i];
  }
  const ElemT &operator[](const size_t i) const { return *raw_[i]; }
  ElemT *&operator()(const size_t i) { return raw_[i]; }
  const ElemT *operator()(const size_t i) const { return raw_[i]; }
  ElemT &front() { return *raw_.front(); }
  const ElemT &front() const { return *raw_.front(); }
  ElemT &back() { return *raw_.back(); }
  const ElemT &back() const { return *raw_.back(); }
  // element pointers (duovector.h:161-170)
  std::vector<const ElemT *> cdata() const { return std::vector<const ElemT *>(raw_.begin(), raw_.end()); }
  ElemT *alloc(const size_t i) {
    delete raw_[i];
    raw_[i] = new ElemT;
    return raw_[i];
  }
  void dealloc(const size_t i) {
    delete raw_[i];
    raw_[i] = nullptr;
  }
  void clear() {
    for (auto &p : raw_) {
      delete p;
      p = nullptr;
    }
  }
  size_t size() const { return raw_.size(); }
  bool empty() const {
    for (auto p : raw_)
      if (p) return false;
    return true;
  }

 protected:
  std::vector<ElemT *> raw_;
};

template <typename TenT>
class TenVec : public DuoVector<TenT> {
 public:
  TenVec() {}
  explicit TenVec(const size_t size) : DuoVector<TenT>(size) {}
  void LoadTen(const size_t idx, const std::string &file) {
    this->alloc(idx);
    ReadGQTensorFromFile((*this)[idx], file);
  }
  void DumpTen(const size_t idx, const std::string &file) const { WriteGQTensorTOFile((*this)[idx], file); }
  void DumpTen(const size_t idx, const std::string &file, const bool release_mem = false) {
    WriteGQTensorTOFile((*this)[idx], file);
    if (release_mem) this->dealloc(idx);
  }
};

template <typename TenT>
using MPO = TenVec<TenT>;

}  // namespace gqmps2

#endif
