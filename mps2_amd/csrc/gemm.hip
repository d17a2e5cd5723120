// Grouped block-sparse GEMM on CDNA4 matrix cores.
//
// One launch covers every output block of one contraction stage.  Each
// workgroup owns one BMxBN tile of one output block and accumulates, in a
// fixed order, the products of all block pairs that feed that block (owner
// computes: no atomics, bit-reproducible).  Operands are addressed through
// per-pair offsets; the A operand is read either row-major (M x K) or
// transposed (stored K x M), likewise B, so contraction legs that lead or
// trail a block need no permutation pass.
//
// Matrix core: v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32),
// 4 wavefronts of 64 lanes in a 2x2 arrangement, LDS-staged BK=16 slabs.
// FP32: the instruction's k index is the lane group g = lane / 16; within a
// slab, MFMA step s feeds k = 4g + s, so each lane's four k values of a
// fragment are contiguous in k-major LDS rows and come in one 16-byte read.
// FP64 (one wavefront per SIMD at 128x128) keeps m/n-major rows and one
// scalar read per fragment and step, which measured faster there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "runtime.h"

namespace mps2 {
namespace {

template <typename T>
struct Mfma;

// Operand layout (both types): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16].  Accumulator element r of lane l is D[row(l, r)][l % 16],
// where the row mapping differs between the FP64 and FP32 instructions
// (measured on gfx950): FP64 rows l/16 + 4r, FP32 rows 4(l/16) + r.
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int kThreads = 256;

// Per-thread share of one BK-deep slab of A (BM x BK) and B (BK x BN),
// loaded from global memory into registers, then stored to LDS: FP32 as
// As[m][k], Bs[n][k] with row stride SK (k contiguous, rows 16-byte
// aligned), FP64 as As[k][m], Bs[k][n] with strides PA, PB.  Out-of-range
// elements read as zero.
template <typename T, int BM, int BN, bool TA, bool TB, int kBK>
struct Slab {
  static constexpr bool kQ = sizeof(T) == 4;
  static constexpr int NA = BM * kBK / kThreads;
  static constexpr int NB = BN * kBK / kThreads;
  static constexpr int SK = kBK + 4;
  static constexpr int PA = BM + 1, PB = BN + 1;
  static constexpr int LA = kQ ? BM * SK : kBK * PA, LB = kQ ? BN * SK : kBK * PB;  // LDS elements
  T ra[NA], rb[NB];

  __device__ __forceinline__ void load(const T *__restrict__ Ab, const T *__restrict__ Bb, int M, int N, int K,
                                       int m0, int n0, int k0, int tid) {
    if (!TA) {
      constexpr int RSTEP = kThreads / kBK;
      const int kk = tid % kBK, gk = k0 + kk;
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int gm = m0 + tid / kBK + i * RSTEP;
        ra[i] = (gm < M && gk < K) ? Ab[(int64_t)gm * K + gk] : T(0);
      }
    } else {
      constexpr int KSTEP = kThreads / BM;
      const int gm = m0 + tid % BM;
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int gk = k0 + tid / BM + i * KSTEP;
        ra[i] = (gm < M && gk < K) ? Ab[(int64_t)gk * M + gm] : T(0);
      }
    }
    if (!TB) {
      constexpr int KSTEP = kThreads / BN;
      const int gn = n0 + tid % BN;
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int gk = k0 + tid / BN + i * KSTEP;
        rb[i] = (gn < N && gk < K) ? Bb[(int64_t)gk * N + gn] : T(0);
      }
    } else {
      constexpr int RSTEP = kThreads / kBK;
      const int kk = tid % kBK, gk = k0 + kk;
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int gn = n0 + tid / kBK + i * RSTEP;
        rb[i] = (gn < N && gk < K) ? Bb[(int64_t)gn * K + gk] : T(0);
      }
    }
  }

  // LDS index of A(m, k) and B(k, n) within a slab
  static __device__ __forceinline__ int ia(int m, int k) { return kQ ? m * SK + k : k * PA + m; }
  static __device__ __forceinline__ int ib(int k, int n) { return kQ ? n * SK + k : k * PB + n; }

  __device__ __forceinline__ void store(T *As, T *Bs, int tid) const {
    if (!TA) {
      constexpr int RSTEP = kThreads / kBK;
      const int kk = tid % kBK;
#pragma unroll
      for (int i = 0; i < NA; ++i) As[ia(tid / kBK + i * RSTEP, kk)] = ra[i];
    } else {
      constexpr int KSTEP = kThreads / BM;
      const int mm = tid % BM;
#pragma unroll
      for (int i = 0; i < NA; ++i) As[ia(mm, tid / BM + i * KSTEP)] = ra[i];
    }
    if (!TB) {
      constexpr int KSTEP = kThreads / BN;
      const int nn = tid % BN;
#pragma unroll
      for (int i = 0; i < NB; ++i) Bs[ib(tid / BN + i * KSTEP, nn)] = rb[i];
    } else {
      constexpr int RSTEP = kThreads / kBK;
      const int kk = tid % kBK;
#pragma unroll
      for (int i = 0; i < NB; ++i) Bs[ib(kk, tid / kBK + i * RSTEP)] = rb[i];
    }
  }
};

// Four consecutive k values of one FP32 LDS row (16-byte aligned).
struct Quad {
  float v[4];
  __device__ __forceinline__ void load(const float *p) {
    const float4 q = *reinterpret_cast<const float4 *>(p);
    v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
  }
};

// The (block pair, K slab) sequence of a tile is walked as one flat loop so
// the global loads of the next slab (possibly of the next pair) overlap the
// matrix-core work on the current one.  NBUF = 2: registers -> LDS double
// buffer, one barrier per slab; NBUF = 1: one LDS buffer (half the LDS, so
// more workgroups per compute unit), two barriers per slab.  BK is the slab
// depth; for FP32 lane group g covers k in [g BK/4, (g+1) BK/4) and the
// MFMA steps walk that range four k at a time (one 16-byte LDS read per
// fragment per four steps).
// TAG only names the instance (profiling: isolates one stage's launches).
template <typename T, int BM, int BN, bool TA, bool TB, int TAG, int BK, int NBUF>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(const T *__restrict__ A, const T *__restrict__ B,
                                                                 T *__restrict__ C, const TileD *__restrict__ tiles,
                                                                 const BlkD *__restrict__ blks,
                                                                 const PairD *__restrict__ pairs, int ntiles, int sub) {
  using V4 = typename Mfma<T>::v4;
  using SlabT = Slab<T, BM, BN, TA, TB, BK>;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  __shared__ __align__(16) T As[NBUF][SlabT::LA];
  __shared__ __align__(16) T Bs[NBUF][SlabT::LB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tix = blockIdx.x;
  if (tix >= ntiles) return;
  const TileD tile = tiles[tix];
  if (tile.blk < 0) return;  // padding of a shorter per-XCD queue
  const BlkD blk = blks[tile.blk];
  const int M = blk.M, N = blk.N;
  const int m0 = tile.m0, n0 = tile.n0;
  // Tiles listed for an upper bound of a block whose size is set on the device.
  if (m0 >= M || n0 >= N || blk.pb >= blk.pe) return;

  V4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = V4{0, 0, 0, 0};

  SlabT slab;
  int p = blk.pb, k0 = 0;
  PairD pr = pairs[p];
  slab.load(A + pr.a_off, B + pr.b_off, M, N, pr.K, m0, n0, 0, tid);
  slab.store(As[0], Bs[0], tid);
  __syncthreads();
  int cur = 0;
  while (true) {
    int np = p, nk0 = k0 + BK;
    PairD npr = pr;
    if (nk0 >= pr.K) {
      np = p + 1;
      nk0 = 0;
      if (np < blk.pe) npr = pairs[np];
    }
    const bool has_next = np < blk.pe;
    if (has_next) slab.load(A + npr.a_off, B + npr.b_off, M, N, npr.K, m0, n0, nk0, tid);
    if constexpr (SlabT::kQ) {
      constexpr int SK = SlabT::SK;
      constexpr int KPL = BK / 4;   // k values per lane group and slab
      const float *as = As[cur] + (wm * WM + (lane & 15)) * SK + KPL * (lane >> 4);
      const float *bs = Bs[cur] + (wn * WN + (lane & 15)) * SK + KPL * (lane >> 4);
#pragma unroll
      for (int h = 0; h < KPL / 4; ++h) {
        Quad af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i].load(as + i * 16 * SK + 4 * h);
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j].load(bs + j * 16 * SK + 4 * h);
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::op(af[i].v[st], bf[j].v[st], acc[i][j]);
      }
    } else {
      constexpr int PA = SlabT::PA, PB = SlabT::PB;
      const T *as = As[cur];
      const T *bs = Bs[cur];
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        const int kr = ks + (lane >> 4);
        T af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = as[kr * PA + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = bs[kr * PB + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::op(af[i], bf[j], acc[i][j]);
      }
    }
    if constexpr (NBUF == 2) {
      if (has_next) slab.store(As[cur ^ 1], Bs[cur ^ 1], tid);
      __syncthreads();
      if (!has_next) break;
      cur ^= 1;
    } else {
      if (!has_next) break;
      __syncthreads();
      slab.store(As[0], Bs[0], tid);
      __syncthreads();
    }
    p = np;
    k0 = nk0;
    pr = npr;
  }
  // ---- epilogue: element r of each 16x16 fragment is row Mfma<T>::row(lane, r), column lane%16
  T *Cb = C + blk.c_off;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * WM + i * 16 + Mfma<T>::row(lane, r);
        if (gm < M && gn < N) {
          if (sub) Cb[(int64_t)gm * N + gn] -= acc[i][j][r];
          else Cb[(int64_t)gm * N + gn] = acc[i][j][r];
        }
      }
    }
}

template <typename T, int BM, int BN, int TAG, int BK, int NBUF>
void launch_tiles_v(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileD *tiles, int ntiles,
                    const BlkD *blks, const PairD *pairs, int sub) {
  dim3 grid(ntiles), block(kThreads);
  if (la == 0 && lb == 0)
    grouped_gemm_kernel<T, BM, BN, false, false, TAG, BK, NBUF>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
  else if (la == 0 && lb == 1)
    grouped_gemm_kernel<T, BM, BN, false, true, TAG, BK, NBUF>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
  else if (la == 1 && lb == 0)
    grouped_gemm_kernel<T, BM, BN, true, false, TAG, BK, NBUF>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
  else
    grouped_gemm_kernel<T, BM, BN, true, true, TAG, BK, NBUF>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
}

// FP32 128x128 slab variant (MPS2_GEMM_VARIANT): 0 = BK 16, double LDS
// buffer; 1 = BK 32, double buffer; 2 = BK 32, one buffer.
int f32_variant() {
  static const int v = [] {
    const char *e = std::getenv("MPS2_GEMM_VARIANT");
    int x = e ? std::atoi(e) : 0;
    return x < 0 || x > 2 ? 0 : x;
  }();
  return v;
}

template <typename T, int BM, int BN, int TAG>
void launch_tiles(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileD *tiles, int ntiles,
                  const BlkD *blks, const PairD *pairs, int sub) {
  if constexpr (sizeof(T) == 4 && BM == 128) {
    const int v = f32_variant();
    if (v == 1) return launch_tiles_v<T, BM, BN, TAG, 32, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
    if (v == 2) return launch_tiles_v<T, BM, BN, TAG, 32, 1>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
  }
  launch_tiles_v<T, BM, BN, TAG, 16, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
}

}  // namespace

// XCD-aware tile order.  Workgroup i of a launch runs on XCD i % 8, and each
// XCD has its own L2.  The tiles of one output block that share operand
// panels (the same rows of A or columns of B) are gathered into groups of up
// to kGroup x kGroup tiles, each group is assigned whole to one XCD
// (longest-processing-time on the groups' MFMA work, so the XCDs finish
// together), and each XCD's queue runs heaviest group first with a group's
// tiles consecutive, so they are resident together and read each panel
// through that XCD's L2 once per group rather than once per tile.  The list
// interleaves the queues (entry 8k + x is the k-th tile of XCD x); shorter
// queues are padded with blk = -1 entries, which exit at once.
// MPS2_XCDS=8 with MPS2_TILE_GROUP (1..8) selects this order.  The default
// (MPS2_XCDS=1, group 1) is one queue of single tiles, heaviest first: on the
// D=4096 J1-J2 matvec the grouped order (groups of 2, 4 or 8) made the FP32
// stage-4 launch 15-19% slower (4.67 -> 5.38..5.58 ms; FP64 unchanged).
std::vector<TileD> schedule_tiles(const std::vector<BlkD> &blks, const std::vector<int64_t> &kc, int ts) {
  static const int kXcds = [] {
    const char *e = std::getenv("MPS2_XCDS");
    int x = e ? std::atoi(e) : 1;
    return x == 8 ? 8 : 1;
  }();
  static const int kGroup = [] {
    const char *e = std::getenv("MPS2_TILE_GROUP");
    int g = e ? std::atoi(e) : (kXcds == 8 ? 4 : 1);
    return g < 1 ? 1 : (g > 8 ? 8 : g);
  }();
  struct Group {
    int blk, tm0, tm1, tn0, tn1;
    double cost;
  };
  std::vector<Group> groups;
  for (size_t b = 0; b < blks.size(); ++b) {
    const int tm = (blks[b].M + ts - 1) / ts, tn = (blks[b].N + ts - 1) / ts;
    for (int i = 0; i < tm; i += kGroup)
      for (int j = 0; j < tn; j += kGroup) {
        Group g{(int)b, i, std::min(tm, i + kGroup), j, std::min(tn, j + kGroup), 0.0};
        g.cost = (double)(g.tm1 - g.tm0) * (g.tn1 - g.tn0) * (double)std::max<int64_t>(kc[b], 1);
        groups.push_back(g);
      }
  }
  std::stable_sort(groups.begin(), groups.end(), [](const Group &x, const Group &y) { return x.cost > y.cost; });
  std::vector<std::vector<TileD>> q(kXcds);
  std::vector<double> load(kXcds, 0.0);
  for (const Group &g : groups) {
    int x = 0;
    for (int k = 1; k < kXcds; ++k)
      if (load[k] < load[x]) x = k;
    load[x] += g.cost;
    for (int i = g.tm0; i < g.tm1; ++i)
      for (int j = g.tn0; j < g.tn1; ++j) q[x].push_back(TileD{g.blk, i * ts, j * ts, 0});
  }
  size_t len = 0;
  for (auto &v : q) len = std::max(len, v.size());
  std::vector<TileD> out;
  out.reserve(len * kXcds);
  for (size_t k = 0; k < len; ++k)
    for (int x = 0; x < kXcds; ++x) out.push_back(k < q[x].size() ? q[x][k] : TileD{-1, 0, 0, 0});
  return out;
}

template <typename T>
void launch_grouped_gemm(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileD *tiles, int ntiles,
                         const BlkD *blks, const PairD *pairs, int tile_m, int tile_n, int sub, int tag) {
  if (ntiles <= 0) return;
  const bool big = tile_m == 128 && tile_n == 128;
  if (tag == 1) {
    if (big) launch_tiles<T, 128, 128, 1>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
    else launch_tiles<T, 64, 64, 1>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
  } else if (tag == 2) {
    if (big) launch_tiles<T, 128, 128, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
    else launch_tiles<T, 64, 64, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
  } else {
    if (big) launch_tiles<T, 128, 128, 0>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
    else launch_tiles<T, 64, 64, 0>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
  }
  c->gemm_launches++;
}

template void launch_grouped_gemm<double>(Ctx *, int, int, const double *, const double *, double *, const TileD *,
                                          int, const BlkD *, const PairD *, int, int, int, int);
template void launch_grouped_gemm<float>(Ctx *, int, int, const float *, const float *, float *, const TileD *, int,
                                         const BlkD *, const PairD *, int, int, int, int);

}  // namespace mps2
