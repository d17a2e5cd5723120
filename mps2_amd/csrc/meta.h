// Host-side quantum-number bookkeeping of block-sparse tensors.
//
// The integer side of GraceQ/tensor's GQTensor (indexes of QN sectors with an
// IN/OUT direction, blocks keyed by one sector per leg, divergence
// sum(OUT qn) - sum(IN qn)) as used by the reference hot path
// (include/gqmps2/algorithm/vmps/two_site_update_finite_vmps_impl.h:117-317,
// lanczos_solver_impl.h:228-256).  Block matching stays on the host and is
// exact; only dense block data lives on the device.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace mps2 {

constexpr int kMaxQnWidth = 4;
constexpr int kMaxRank = 8;

struct QN {
  int32_t v[kMaxQnWidth] = {0, 0, 0, 0};
  bool operator==(const QN &o) const {
    for (int i = 0; i < kMaxQnWidth; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
  bool operator!=(const QN &o) const { return !(*this == o); }
  bool operator<(const QN &o) const {
    for (int i = 0; i < kMaxQnWidth; ++i)
      if (v[i] != o.v[i]) return v[i] < o.v[i];
    return false;
  }
  QN operator+(const QN &o) const {
    QN r;
    for (int i = 0; i < kMaxQnWidth; ++i) r.v[i] = v[i] + o.v[i];
    return r;
  }
  QN operator-(const QN &o) const {
    QN r;
    for (int i = 0; i < kMaxQnWidth; ++i) r.v[i] = v[i] - o.v[i];
    return r;
  }
};

enum Dir : int32_t { kIn = 0, kOut = 1 };

struct Leg {
  int32_t dir = kIn;
  std::vector<QN> qn;
  std::vector<int64_t> dim;
  bool operator==(const Leg &o) const { return dir == o.dir && qn == o.qn && dim == o.dim; }
  int64_t total() const {
    int64_t t = 0;
    for (auto d : dim) t += d;
    return t;
  }
  Leg inverse() const {
    Leg r = *this;
    r.dir = dir == kIn ? kOut : kIn;
    return r;
  }
  int sign() const { return dir == kOut ? 1 : -1; }
};

using Coord = std::vector<int32_t>;

struct CoordHash {
  size_t operator()(const Coord &c) const {
    size_t h = 1469598103934665603ull;
    for (auto x : c) h = (h ^ (uint32_t)x) * 1099511628211ull;
    return h;
  }
};

// Immutable layout of a block-sparse tensor: legs + sorted block list with
// element offsets into one contiguous buffer.
struct Meta;

// The layout with every leg direction flipped (Dag), remembered by the
// layout it was made from so that repeated Dag of tensors sharing a layout
// yields one layout object (and the contraction-plan cache, keyed by layout
// objects, hits).  A copied Meta starts unlinked: a copy may be modified.
struct DagLink {
  std::weak_ptr<const Meta> p;
  DagLink() = default;
  DagLink(const DagLink &) {}
  DagLink &operator=(const DagLink &) {
    p.reset();
    return *this;
  }
};

struct Meta {
  int qw = 1;
  std::vector<Leg> legs;
  std::vector<Coord> coords;            // sorted lexicographically
  std::vector<int64_t> offset;          // element offset of each block
  std::vector<int64_t> bsize;           // elements in each block
  int64_t total = 0;
  uint64_t digest = 0;                  // hash of (qw, legs, coords), set by finalize()
  std::unordered_map<Coord, int32_t, CoordHash> lookup;
  mutable DagLink dag;                  // see DagLink

  int rank() const { return (int)legs.size(); }
  int64_t dim_of(int blk, int leg) const { return legs[leg].dim[coords[blk][leg]]; }
  QN block_div(const Coord &c) const;
  QN div() const;                       // divergence of the first block (zero QN if empty)
  int32_t find(const Coord &c) const {
    auto it = lookup.find(c);
    return it == lookup.end() ? -1 : it->second;
  }
  // Sort coords, fill offsets / sizes / lookup.
  void finalize();
  bool same_layout(const Meta &o) const;
};

using MetaP = std::shared_ptr<const Meta>;

// All divergence-allowed blocks of a leg set (the complete block structure of
// a vector space of fixed divergence: the Lanczos vectors share it).
MetaP full_meta(int qw, const std::vector<Leg> &legs, const QN &div);

// Parse / emit the flat exchange format (see include/mps2_amd.h).
MetaP meta_from_flat(const int32_t *m, int64_t *consumed);
int64_t meta_flat_size(const Meta &m);
void meta_to_flat(const Meta &m, int32_t *out);

// ------------------------------------------------------------------ Contract plan
// One GEMM operand view: how element (row, col) of a block matrix is addressed
// after the optional permutation of the operand.
enum OpLayout : int32_t { kNormal = 0, kTrans = 1 };

struct PairH {
  int64_t a_off, b_off;   // element offsets in the (possibly permuted) operand buffers
  int32_t K;
};

struct BlkH {
  int64_t c_off;          // element offset in the output buffer
  int32_t M, N;
  int32_t pair_begin, pair_end;
};

// Contractions whose output blocks are all at most kNarrowMaxN columns wide
// run on the vector ALUs (narrow_gemm_kernel), the others on the matrix cores.
constexpr int kNarrowMaxN = 16;

struct ContractPlan {
  MetaP a, b, out;
  std::vector<int> ax, bx;
  bool permute_a = false, permute_b = false;
  std::vector<int> aperm, bperm;        // permutation applied when permute_* is set
  OpLayout la = kNormal, lb = kNormal;  // layout seen by the GEMM
  std::vector<PairH> pairs;
  std::vector<BlkH> blks;
  double flops = 0;                     // 2*M*N*K summed over pairs
};

// Plan C = Contract(A, B, {ax, bx}) with gqten semantics: out legs = free legs
// of A then free legs of B, contracted legs must be mutually inverse, output
// blocks are those receiving at least one block-pair product.
std::shared_ptr<ContractPlan> plan_contract(const MetaP &a, const MetaP &b, const std::vector<int> &ax,
                                            const std::vector<int> &bx);

// Permuted layout of a meta: legs reordered by perm (new leg i = old perm[i]).
MetaP permuted_meta(const Meta &m, const std::vector<int> &perm);

std::string qn_str(const QN &q, int qw);

}  // namespace mps2
