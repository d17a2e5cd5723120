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

// Tile tix of a launch (TileMap, runtime.h): from the list, or by binary
// search of the launch-order prefix sums for the block (the largest position
// whose start is <= tix; blocks without tiles are never found) and its
// row-major tile (i, j).
__device__ __forceinline__ TileD tile_at(const TileMap &T, const BlkD *__restrict__ blks, int tix) {
  if (T.list) return T.list[tix];
  int lo = 0, hi = T.nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (T.start[mid] <= tix) lo = mid;
    else hi = mid;
  }
  const int b = T.order[lo], local = tix - T.start[lo];
  const int tn = (blks[b].N + T.tn - 1) / T.tn;
  return TileD{b, (local / tn) * T.tm, (local % tn) * T.tn, 0};
}

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

// Number of a wavefront's 16-row (or 16-column) fragments that hold any row
// (column) of the block, out of F, given the rows left from the wavefront's
// first one (<= 0: none); read through readfirstlane so the count is a scalar
// and the branches on it are uniform.  Fragments past it only multiply the
// zero padding of an edge tile and are skipped.
template <int F>
__device__ __forceinline__ int live_frags(int rows_left) {
  const int f = rows_left <= 0 ? 0 : min(F, (rows_left + 15) / 16);
  return __builtin_amdgcn_readfirstlane(f);
}

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
// SPLIT > 1 (split-K): workgroup w runs part w % SPLIT of tile w / SPLIT, the
// part-th of SPLIT equal runs of the tile's (block pair, slab) sequence;
// part 0 writes C, part s > 0 the scratch copy s - 1 (stride sstride, same
// offsets), and splitk_reduce_kernel adds them in part order.
template <typename T, int BM, int BN, bool TA, bool TB, int TAG, int BK, int NBUF, int SPLIT>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(const T *__restrict__ A, const T *__restrict__ B,
                                                                 T *__restrict__ C, TileMap tiles,
                                                                 const BlkD *__restrict__ blks,
                                                                 const PairD *__restrict__ pairs, int ntiles, int sub,
                                                                 T *__restrict__ scratch, int64_t sstride) {
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
  int tix = blockIdx.x, part = 0;
  if constexpr (SPLIT > 1) {
    part = tix % SPLIT;
    tix /= SPLIT;
  }
  if (tix >= ntiles) return;
  const TileD tile = tile_at(tiles, blks, tix);
  if (tile.blk < 0) return;  // padding of a shorter per-XCD queue
  const BlkD blk = blks[tile.blk];
  const int M = blk.M, N = blk.N;
  const int m0 = tile.m0, n0 = tile.n0;
  // Tiles listed for an upper bound of a block whose size is set on the device.
  if (m0 >= M || n0 >= N || blk.pb >= blk.pe) return;

  // The slabs this workgroup multiplies: all of the tile's, or one part's
  // (first slab at pair p, depth k0; `left` of them).
  int p = blk.pb, k0 = 0, left = 0x7fffffff;
  if constexpr (SPLIT > 1) {
    int total = 0;
    for (int q = blk.pb; q < blk.pe; ++q) total += (pairs[q].K + BK - 1) / BK;
    const int j0 = (int)((int64_t)total * part / SPLIT);
    left = (int)((int64_t)total * (part + 1) / SPLIT) - j0;
    int j = j0;
    while (p < blk.pe) {
      const int ns = (pairs[p].K + BK - 1) / BK;
      if (j < ns) break;
      j -= ns;
      ++p;
    }
    k0 = j * BK;
  }

  V4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = V4{0, 0, 0, 0};
  // Fragments beyond the block are skipped (wave-uniform scalar branches):
  // FP64 stage 4 of the D=4096 matvec 7.8 -> 7.4 ms; FP32 with split-K
  // 4.08 -> 3.86 ms (without split-K it measured even).
  constexpr bool kSkip = true;
  const int fmv = kSkip ? live_frags<FM>(M - m0 - wm * WM) : FM;
  const int fnv = kSkip ? live_frags<FN>(N - n0 - wn * WN) : FN;

  SlabT slab;
  int cur = 0;
  // One slab's products; mid() runs halfway through them (the next slab's
  // LDS store, so that it and the wait for its global loads overlap the
  // matrix-core work).
  auto multiply = [&](auto &&mid) {
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
        for (int st = 0; st < 4; ++st) {
          if (4 * h + st == KPL / 2) mid();
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              if (!kSkip || (i < fmv && j < fnv)) acc[i][j] = Mfma<T>::op(af[i].v[st], bf[j].v[st], acc[i][j]);
        }
      }
    } else {
      constexpr int PA = SlabT::PA, PB = SlabT::PB;
      const T *as = As[cur];
      const T *bs = Bs[cur];
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        if (ks == BK / 2) mid();
        const int kr = ks + (lane >> 4);
        T af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = as[kr * PA + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = bs[kr * PB + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            if (!kSkip || (i < fmv && j < fnv)) acc[i][j] = Mfma<T>::op(af[i], bf[j], acc[i][j]);
      }
    }
  };
  if (left > 0) {  // uniform: a split part without slabs writes zeros
    PairD pr = pairs[p];
    slab.load(A + pr.a_off, B + pr.b_off, M, N, pr.K, m0, n0, k0, tid);
    slab.store(As[0], Bs[0], tid);
    __syncthreads();
    while (true) {
      int np = p, nk0 = k0 + BK;
      PairD npr = pr;
      if (nk0 >= pr.K) {
        np = p + 1;
        nk0 = 0;
        if (np < blk.pe) npr = pairs[np];
      }
      const bool has_next = np < blk.pe && (SPLIT == 1 || left > 1);
      if (has_next) slab.load(A + npr.a_off, B + npr.b_off, M, N, npr.K, m0, n0, nk0, tid);
      if constexpr (NBUF == 2) {
        multiply([&] {
          if (has_next) slab.store(As[cur ^ 1], Bs[cur ^ 1], tid);
        });
        __syncthreads();
        if (!has_next) break;
        cur ^= 1;
      } else {
        multiply([] {});
        if (!has_next) break;
        __syncthreads();
        slab.store(As[0], Bs[0], tid);
        __syncthreads();
      }
      p = np;
      k0 = nk0;
      pr = npr;
      if constexpr (SPLIT > 1) --left;
    }
  }
  // ---- epilogue: element r of each 16x16 fragment is row Mfma<T>::row(lane, r), column lane%16
  T *Cb = C + blk.c_off;
  if constexpr (SPLIT > 1) {
    if (part > 0) {
      Cb = scratch + (int64_t)(part - 1) * sstride + blk.c_off;
      sub = 0;
    }
  }
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

// ---------------------------------------------------------------- FP32, direct-to-LDS loads
// FP32 grouped GEMM for row-major A (M x K) and row-major B (K x N) whose
// slabs go from global memory straight into LDS (global_load_lds_dword: lane
// i of a wavefront writes the dword at LDS base + 4 i), so no registers stage
// them and NST - 1 slabs stay in flight while one is multiplied.  Slab depth
// 16.  LDS per stage:
//   A: 128 rows x 16 k, row m at 16 m, k stored at (k ^ ((m >> 2 & 3) << 2))
//      (quads permuted per row group: the 16-byte fragment reads of 16
//      consecutive rows fall on distinct banks);
//   B: 16 k rows x 128 n, row k at kLdB * k (kLdB = 132: the four lane groups'
//      rows of a fragment read fall on distinct banks).
// Out-of-range rows and columns read clamped, valid addresses (their products
// land only in outputs the epilogue does not store); k beyond the pair's K
// reads B from a zero array, so those terms vanish.
__device__ float g_dma_zero[64];

// s_waitcnt vmcnt(N) lgkmcnt(0) (expcnt not waited for), gfx9 encoding:
// vmcnt bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8].
template <int N>
__device__ __forceinline__ void dma_wait() {
  static_assert(N >= 0 && N < 64, "vmcnt holds 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4));
}

// Workgroup barrier without the release/acquire fences of __syncthreads (a
// workgroup-scope release would wait for every load-to-LDS still in flight,
// the ones the pipeline keeps in flight on purpose).  Each thread first
// waits for its own loads of the slab about to be read (dma_wait) and for its
// LDS reads (lgkmcnt 0), so after the barrier that slab is complete in LDS
// and no wavefront still reads the stage the next loads overwrite; the signal
// fences keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void dma_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

constexpr int kDmaBK = 16;
constexpr int kLdB = 132;
constexpr int kDmaA = 128 * kDmaBK;   // floats per A stage
constexpr int kDmaB = kDmaBK * kLdB;  // floats per B stage

// One step's MFMA operands from LDS: four A quads at a + 1024 i bytes and
// sixteen B values at b + 4 kLdB s + 64 j bytes, in one asm block that waits
// for its own reads (lgkmcnt 0) before it ends, so every output is complete
// when the block returns.  The compiler does not see these reads as LDS
// accesses and so adds no wait for the loads-to-LDS still in flight; the
// explicit waits and barriers of the loop order them.
struct Frags {
  float4 a[4];
  float b[4][4];  // [s][j]
};
__device__ __forceinline__ void read_frags(Frags &f, uint32_t a, uint32_t b) {
  asm volatile(
      "ds_read_b128 %0, %20\n\t"
      "ds_read_b128 %1, %20 offset:1024\n\t"
      "ds_read_b128 %2, %20 offset:2048\n\t"
      "ds_read_b128 %3, %20 offset:3072\n\t"
      "ds_read_b32 %4, %21\n\t"
      "ds_read_b32 %5, %21 offset:64\n\t"
      "ds_read_b32 %6, %21 offset:128\n\t"
      "ds_read_b32 %7, %21 offset:192\n\t"
      "ds_read_b32 %8, %21 offset:528\n\t"
      "ds_read_b32 %9, %21 offset:592\n\t"
      "ds_read_b32 %10, %21 offset:656\n\t"
      "ds_read_b32 %11, %21 offset:720\n\t"
      "ds_read_b32 %12, %21 offset:1056\n\t"
      "ds_read_b32 %13, %21 offset:1120\n\t"
      "ds_read_b32 %14, %21 offset:1184\n\t"
      "ds_read_b32 %15, %21 offset:1248\n\t"
      "ds_read_b32 %16, %21 offset:1584\n\t"
      "ds_read_b32 %17, %21 offset:1648\n\t"
      "ds_read_b32 %18, %21 offset:1712\n\t"
      "ds_read_b32 %19, %21 offset:1776\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(f.a[0]), "=&v"(f.a[1]), "=&v"(f.a[2]), "=&v"(f.a[3]), "=&v"(f.b[0][0]), "=&v"(f.b[0][1]),
        "=&v"(f.b[0][2]), "=&v"(f.b[0][3]), "=&v"(f.b[1][0]), "=&v"(f.b[1][1]), "=&v"(f.b[1][2]),
        "=&v"(f.b[1][3]), "=&v"(f.b[2][0]), "=&v"(f.b[2][1]), "=&v"(f.b[2][2]), "=&v"(f.b[2][3]),
        "=&v"(f.b[3][0]), "=&v"(f.b[3][1]), "=&v"(f.b[3][2]), "=&v"(f.b[3][3])
      : "v"(a), "v"(b)
      : "memory");
}
static_assert(kLdB * 4 == 528, "B row offsets in read_frags assume kLdB = 132");

template <int NST, int TAG>
__global__ __launch_bounds__(kThreads) void grouped_gemm_dma_kernel(const float *__restrict__ A,
                                                                     const float *__restrict__ B, float *__restrict__ C,
                                                                     TileMap tiles,
                                                                     const BlkD *__restrict__ blks,
                                                                     const PairD *__restrict__ pairs, int ntiles,
                                                                     int sub) {
  static_assert(NST == 3 || NST == 4, "three or four stages");
  using V4 = Mfma<float>::v4;
  constexpr int WM = 64, WN = 64, FM = 4, FN = 4;
  __shared__ __align__(16) float As[NST][kDmaA];
  __shared__ __align__(16) float Bs[NST][kDmaB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tix = blockIdx.x;
  if (tix >= ntiles) return;
  const TileD tile = tile_at(tiles, blks, tix);
  if (tile.blk < 0) return;
  const BlkD blk = blks[tile.blk];
  const int M = blk.M, N = blk.N, m0 = tile.m0, n0 = tile.n0;
  if (m0 >= M || n0 >= N || blk.pb >= blk.pe) return;
  int nslab = 0;
  for (int p = blk.pb; p < blk.pe; ++p) nslab += (max(pairs[p].K, 0) + kDmaBK - 1) / kDmaBK;

  // Issue state: the slab (pair ip, depth ik0) whose loads go out next.
  // A: 8 loads per thread and slab; load j of wave w covers rows
  // 32 w + 4 j .. +3 (lane i: row + i / 16, LDS column i % 16, holding k
  // column (i % 16) ^ 4 (j % 4): the quad order of row group j).
  // B: 8 loads per thread; load j covers k row 4 w + j / 2, columns
  // 64 (j % 2) .. +63 (lane i: column + i).
  int ip = blk.pb, ik0 = 0;
  while (ip < blk.pe && pairs[ip].K <= 0) ++ip;  // pairs without K contribute nothing
  if (ip >= blk.pe) ip = blk.pb;                 // (then nslab == 0 and nothing is issued)
  PairD ipr = pairs[ip];
  auto issue = [&](int stage) {
    const float *Ab = A + ipr.a_off;
    const float *Bb = B + ipr.b_off;
    const int K = ipr.K;
    float *as = As[stage];
    float *bs = Bs[stage];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int r = 32 * wave + 4 * j + (lane >> 4);
      const int k = ik0 + ((lane & 15) ^ ((j & 3) << 2));
      const uint32_t off = (uint32_t)min(m0 + r, M - 1) * (uint32_t)K + (uint32_t)min(k, K - 1);
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(Ab + off),
                                       (__attribute__((address_space(3))) void *)(as + (32 * wave + 4 * j) * 16), 4,
                                       0, 0);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kr = 4 * wave + (j >> 1), c0 = 64 * (j & 1), k = ik0 + kr;
      const uint32_t off = (uint32_t)k * (uint32_t)N + (uint32_t)min(n0 + c0 + lane, N - 1);
      const float *src = k < K ? Bb + off : g_dma_zero + lane;
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                       (__attribute__((address_space(3))) void *)(bs + kr * kLdB + c0), 4, 0, 0);
    }
    ik0 += kDmaBK;
    if (ik0 >= K) {
      ik0 = 0;
      do {
        ++ip;
      } while (ip < blk.pe && pairs[ip].K <= 0);
      if (ip < blk.pe) ipr = pairs[ip];
    }
  };

  V4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = V4{0, 0, 0, 0};

  // LDS byte addresses of this lane's fragments within a stage: A row
  // m = 64 wm + lane % 16 (+16 i), quad lane / 16 permuted by the row group
  // ((m >> 2) & 3 = (lane % 16) >> 2); B row k = 4 (lane / 16) (+ s), column
  // 64 wn + lane % 16 (+16 j).
  const int ma = wm * WM + (lane & 15);
  const uint32_t a_lane = (uint32_t)(ma * 16 + ((4 * (lane >> 4)) ^ (((lane & 15) >> 2) << 2))) * 4u;
  const uint32_t b_lane = (uint32_t)((4 * (lane >> 4)) * kLdB + wn * WN + (lane & 15)) * 4u;
  const uint32_t a_base = (uint32_t)(uintptr_t)&As[0][0], b_base = (uint32_t)(uintptr_t)&Bs[0][0];

  int issued = 0;
  for (; issued < NST - 1 && issued < nslab; ++issued) issue(issued);
  int st = 0;  // stage of slab t
  for (int t = 0; t < nslab; ++t) {
    // loads still allowed in flight: those of slabs t+1 .. t+NST-2
    const int ahead = min(NST - 2, nslab - 1 - t);
    if (ahead >= 2) dma_wait<32>();
    else if (ahead == 1) dma_wait<16>();
    else dma_wait<0>();
    dma_barrier();
    if (issued < nslab) {
      issue(issued % NST);
      ++issued;
    }
    Frags f;
    read_frags(f, a_base + (uint32_t)st * (kDmaA * 4u) + a_lane, b_base + (uint32_t)st * (kDmaB * 4u) + b_lane);
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const float av = s4 == 0 ? f.a[i].x : s4 == 1 ? f.a[i].y : s4 == 2 ? f.a[i].z : f.a[i].w;
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<float>::op(av, f.b[s4][j], acc[i][j]);
      }
    st = st + 1 == NST ? 0 : st + 1;
  }

  float *Cb = C + blk.c_off;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * WM + i * 16 + Mfma<float>::row(lane, r);
        if (gm < M && gn < N) {
          if (sub) Cb[(int64_t)gm * N + gn] -= acc[i][j][r];
          else Cb[(int64_t)gm * N + gn] = acc[i][j][r];
        }
      }
    }
}

// ---------------------------------------------------------------- FP32 operands, split-BF16x3 products
// FP32 operands x are split as x = hi + lo with hi = bf16(x) and lo =
// bf16(x - hi) (round to nearest; x - hi is exact in FP32), and each product
// a b is formed as hi_a hi_b + hi_a lo_b + lo_a hi_b on the BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.  The representation
// error |x - hi - lo| <= 2^-16 |x| and the dropped lo_a lo_b term leave each
// product accurate to a few units of 2^-16, independent from product to
// product, so a K-term sum carries about that error times its root-sum-square
// rather than its sum: the same order as FP32 accumulation over the K depths
// of the DMRG stages.
//
// Tiles BM x BN, 4 wavefronts in 2 x 2, 32-deep slabs loaded as FP32 into
// registers (the FP32 kernel's Slab), split when stored to LDS.  LDS row r of
// A (row m) and of B (column n): hi for k = 0..31, lo for k = 0..31, 8 bf16
// of padding (144 bytes per row).  Lane l of a wavefront feeds the MFMA rows
// l % 16 with k = 8 (l / 16) .. +7 of both operands (one 16-byte read per
// fragment per half); accumulator element r of lane l is row 4 (l / 16) + r,
// column l % 16 (measured on gfx950 with an in-bounds probe).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
constexpr int kBf3BK = 32;
constexpr int kBf3Row = 2 * kBf3BK + 8;   // bf16 per LDS row

typedef float f4v_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf4v_t __attribute__((ext_vector_type(4)));

// Per-thread share of one 32-deep slab, as quads of four consecutive k of
// one row of A (BM x 32) or one column of B (32 x BN): QA and QB quads per
// thread, loaded from global memory as FP32 (out-of-range elements read as
// zero) and stored to LDS split into BF16 hi and lo parts, one 8-byte write
// per part and quad.  Operands whose k is contiguous in memory (A row-major,
// B stored N x K) give 8 threads per row (a row's 32 k in one 128-byte run);
// operands stored k-major (A as K x M, B as K x N) give consecutive threads
// consecutive rows (each of the four loads coalesced along the row index).
template <int BM, int BN, bool TA, bool TB>
struct Bf3Slab {
  static constexpr int BK = kBf3BK;
  static constexpr int QA = BM * BK / 4 / kThreads, QB = BN * BK / 4 / kThreads;
  f4v_t ra[QA], rb[QB];

  // (row, quad) of this thread's j-th quad
  template <int BR, bool KCONTIG>
  static __device__ __forceinline__ void pos(int tid, int j, int &row, int &quad) {
    if (KCONTIG) {
      row = tid / 8 + j * (kThreads / 8);
      quad = tid % 8;
    } else {
      row = tid % BR;
      quad = tid / BR + j * (kThreads / BR);
    }
  }

  __device__ __forceinline__ void load(const float *__restrict__ Ab, const float *__restrict__ Bb, int M, int N,
                                       int K, int m0, int n0, int k0, int tid) {
#pragma unroll
    for (int j = 0; j < QA; ++j) {
      int r, q;
      pos<BM, !TA>(tid, j, r, q);
      const int gm = m0 + r, gk = k0 + 4 * q;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        ra[j][e] = (gm < M && gk + e < K) ? (TA ? Ab[(int64_t)(gk + e) * M + gm] : Ab[(int64_t)gm * K + gk + e])
                                          : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < QB; ++j) {
      int r, q;
      pos<BN, TB>(tid, j, r, q);
      const int gn = n0 + r, gk = k0 + 4 * q;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        rb[j][e] = (gn < N && gk + e < K) ? (TB ? Bb[(int64_t)gn * K + gk + e] : Bb[(int64_t)(gk + e) * N + gn])
                                          : 0.0f;
    }
  }

  static __device__ __forceinline__ void put(__bf16 *row, int q, f4v_t x) {
    const bf4v_t hi = __builtin_convertvector(x, bf4v_t);
    const bf4v_t lo = __builtin_convertvector(x - __builtin_convertvector(hi, f4v_t), bf4v_t);
    *reinterpret_cast<bf4v_t *>(row + 4 * q) = hi;
    *reinterpret_cast<bf4v_t *>(row + BK + 4 * q) = lo;
  }

  __device__ __forceinline__ void store(__bf16 *as, __bf16 *bs, int tid) const {
#pragma unroll
    for (int j = 0; j < QA; ++j) {
      int r, q;
      pos<BM, !TA>(tid, j, r, q);
      put(as + r * kBf3Row, q, ra[j]);
    }
#pragma unroll
    for (int j = 0; j < QB; ++j) {
      int r, q;
      pos<BN, TB>(tid, j, r, q);
      put(bs + r * kBf3Row, q, rb[j]);
    }
  }
};

// SPLIT > 1: split-K as in grouped_gemm_kernel.
template <int BM, int BN, bool TA, bool TB, int TAG, int SPLIT>
__global__ __launch_bounds__(kThreads) void grouped_gemm_bf3_kernel(const float *__restrict__ A,
                                                                     const float *__restrict__ B,
                                                                     float *__restrict__ C,
                                                                     TileMap tiles,
                                                                     const BlkD *__restrict__ blks,
                                                                     const PairD *__restrict__ pairs, int ntiles,
                                                                     int sub, float *__restrict__ scratch,
                                                                     int64_t sstride) {
  using V4 = Mfma<float>::v4;
  using SlabT = Bf3Slab<BM, BN, TA, TB>;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int RS = kBf3Row, BK = kBf3BK;
  __shared__ __align__(16) __bf16 As[2][BM * RS];
  __shared__ __align__(16) __bf16 Bs[2][BN * RS];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  int tix = blockIdx.x, part = 0;
  if constexpr (SPLIT > 1) {
    part = tix % SPLIT;
    tix /= SPLIT;
  }
  if (tix >= ntiles) return;
  const TileD tile = tile_at(tiles, blks, tix);
  if (tile.blk < 0) return;
  const BlkD blk = blks[tile.blk];
  const int M = blk.M, N = blk.N;
  const int m0 = tile.m0, n0 = tile.n0;
  if (m0 >= M || n0 >= N || blk.pb >= blk.pe) return;
  int p = blk.pb, k0 = 0, left = 0x7fffffff;
  if constexpr (SPLIT > 1) {
    int total = 0;
    for (int q = blk.pb; q < blk.pe; ++q) total += (pairs[q].K + BK - 1) / BK;
    const int j0 = (int)((int64_t)total * part / SPLIT);
    left = (int)((int64_t)total * (part + 1) / SPLIT) - j0;
    int j = j0;
    while (p < blk.pe) {
      const int ns = (pairs[p].K + BK - 1) / BK;
      if (j < ns) break;
      j -= ns;
      ++p;
    }
    k0 = j * BK;
  }

  V4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = V4{0, 0, 0, 0};

  SlabT slab;
  int cur = 0;
  const int frag_row = (lane & 15) * RS + 8 * (lane >> 4);
  if (left > 0) {  // uniform: a split part without slabs writes zeros
  PairD pr = pairs[p];
  slab.load(A + pr.a_off, B + pr.b_off, M, N, pr.K, m0, n0, k0, tid);
  slab.store(As[0], Bs[0], tid);
  __syncthreads();
  while (true) {
    int np = p, nk0 = k0 + BK;
    PairD npr = pr;
    if (nk0 >= pr.K) {
      np = p + 1;
      nk0 = 0;
      if (np < blk.pe) npr = pairs[np];
    }
    const bool has_next = np < blk.pe && (SPLIT == 1 || left > 1);
    if (has_next) slab.load(A + npr.a_off, B + npr.b_off, M, N, npr.K, m0, n0, nk0, tid);
    {
      const __bf16 *as = As[cur] + wm * WM * RS + frag_row;
      const __bf16 *bs = Bs[cur] + wn * WN * RS + frag_row;
      bf16x8_t ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        ah[i] = *reinterpret_cast<const bf16x8_t *>(as + i * 16 * RS);
        al[i] = *reinterpret_cast<const bf16x8_t *>(as + i * 16 * RS + BK);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        bh[j] = *reinterpret_cast<const bf16x8_t *>(bs + j * 16 * RS);
        bl[j] = *reinterpret_cast<const bf16x8_t *>(bs + j * 16 * RS + BK);
      }
      // small cross terms first, then hi x hi; every accumulator is touched
      // once per pass, so consecutive MFMAs are independent
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    }
    if (has_next) slab.store(As[cur ^ 1], Bs[cur ^ 1], tid);
    __syncthreads();
    if (!has_next) break;
    cur ^= 1;
    p = np;
    k0 = nk0;
    pr = npr;
    if constexpr (SPLIT > 1) --left;
  }
  }
  float *Cb = C + blk.c_off;
  if constexpr (SPLIT > 1) {
    if (part > 0) {
      Cb = scratch + (int64_t)(part - 1) * sstride + blk.c_off;
      sub = 0;
    }
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * WM + i * 16 + 4 * (lane >> 4) + r;
        if (gm < M && gn < N) {
          if (sub) Cb[(int64_t)gm * N + gn] -= acc[i][j][r];
          else Cb[(int64_t)gm * N + gn] = acc[i][j][r];
        }
      }
    }
}

template <int BM, int BN, int TAG, int SPLIT = 1>
void launch_tiles_bf3(Ctx *c, int la, int lb, const float *A, const float *B, float *C, const TileMap &tiles,
                      int ntiles, const BlkD *blks, const PairD *pairs, int sub, float *scratch = nullptr,
                      int64_t sstride = 0) {
  dim3 grid((unsigned)((int64_t)ntiles * SPLIT)), block(kThreads);
  if (la == 0 && lb == 0)
    grouped_gemm_bf3_kernel<BM, BN, false, false, TAG, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
  else if (la == 0 && lb == 1)
    grouped_gemm_bf3_kernel<BM, BN, false, true, TAG, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
  else if (la == 1 && lb == 0)
    grouped_gemm_bf3_kernel<BM, BN, true, false, TAG, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
  else
    grouped_gemm_bf3_kernel<BM, BN, true, true, TAG, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
}

// ---------------------------------------------------------------- narrow outputs (vector ALU)
// Contractions whose output blocks are at most 16 columns wide (the two-site
// MPO stage of the matvec: K and N are an MPO bond sector times the physical
// dimensions, a few to a few tens) do about 2N flops per element of A read,
// so they are HBM-bound, and a matrix-core tile, which pads N to 64 or 128
// and K to its slab depth, spends most of its time on padding.  Here one
// thread owns one row of one output block and NC of its columns: it walks the
// block's pairs in their fixed order, reads its row of A (coalesced across
// the wavefront when A is stored K x M, the layout of that stage) and
// accumulates the NC outputs in registers with fused multiply-adds.  B's
// values are the same for the whole wavefront.  Fixed order, so the result
// is bit-reproducible.  A tile is (block, 256 rows from m0, NC columns from n0).
template <typename T, bool TA, bool TB, int NC, int TAG>
__global__ __launch_bounds__(256) void narrow_gemm_kernel(const T *__restrict__ A, const T *__restrict__ B,
                                                          T *__restrict__ C, TileMap tiles,
                                                          const BlkD *__restrict__ blks,
                                                          const PairD *__restrict__ pairs, int ntiles, int sub) {
  const int tix = blockIdx.x;
  if (tix >= ntiles) return;
  const TileD tile = tile_at(tiles, blks, tix);
  if (tile.blk < 0) return;
  const BlkD blk = blks[tile.blk];
  const int M = blk.M, N = blk.N;
  const int m = tile.m0 + (int)threadIdx.x, n0 = tile.n0;
  if (m >= M || n0 >= N) return;
  const int nn = min(NC, N - n0);
  T acc[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) acc[j] = T(0);
  for (int p = blk.pb; p < blk.pe; ++p) {
    const PairD pr = pairs[p];
    const int K = pr.K;
    const T *Ab = A + pr.a_off;
    const T *Bb = B + pr.b_off;
    for (int k = 0; k < K; ++k) {
      const T a = TA ? Ab[(int64_t)k * M + m] : Ab[(int64_t)m * K + k];
#pragma unroll
      for (int j = 0; j < NC; ++j)
        if (j < nn) {
          const T b = TB ? Bb[(int64_t)(n0 + j) * K + k] : Bb[(int64_t)k * N + n0 + j];
          acc[j] = fma(a, b, acc[j]);
        }
    }
  }
  T *Cr = C + blk.c_off + (int64_t)m * N + n0;
#pragma unroll
  for (int j = 0; j < NC; ++j)
    if (j < nn) Cr[j] = sub ? Cr[j] - acc[j] : acc[j];
}

template <typename T, int NC, int TAG>
void launch_narrow_nc(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                      const BlkD *blks, const PairD *pairs, int sub) {
  dim3 grid(ntiles), block(256);
  if (la == 0 && lb == 0)
    narrow_gemm_kernel<T, false, false, NC, TAG><<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
  else if (la == 0 && lb == 1)
    narrow_gemm_kernel<T, false, true, NC, TAG><<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
  else if (la == 1 && lb == 0)
    narrow_gemm_kernel<T, true, false, NC, TAG><<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
  else
    narrow_gemm_kernel<T, true, true, NC, TAG><<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub);
}

template <typename T, int TAG>
void launch_narrow_tag(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                       const BlkD *blks, const PairD *pairs, int nc, int sub) {
  if (nc == 4) launch_narrow_nc<T, 4, TAG>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
  else if (nc == 8) launch_narrow_nc<T, 8, TAG>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
  else launch_narrow_nc<T, 16, TAG>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
}

template <typename T, int BM, int BN, int TAG, int BK, int NBUF, int SPLIT = 1>
void launch_tiles_v(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                    const BlkD *blks, const PairD *pairs, int sub, T *scratch = nullptr, int64_t sstride = 0) {
  dim3 grid((unsigned)((int64_t)ntiles * SPLIT)), block(kThreads);
  if (la == 0 && lb == 0)
    grouped_gemm_kernel<T, BM, BN, false, false, TAG, BK, NBUF, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
  else if (la == 0 && lb == 1)
    grouped_gemm_kernel<T, BM, BN, false, true, TAG, BK, NBUF, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
  else if (la == 1 && lb == 0)
    grouped_gemm_kernel<T, BM, BN, true, false, TAG, BK, NBUF, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
  else
    grouped_gemm_kernel<T, BM, BN, true, true, TAG, BK, NBUF, SPLIT>
        <<<grid, block, 0, c->stream>>>(A, B, C, tiles, blks, pairs, ntiles, sub, scratch, sstride);
}

// C += scratch part 1 + ... + part SPLIT - 1 on every output block, in part
// order (fixed, so split launches are reproducible).
template <typename T, int SPLIT>
__global__ void splitk_reduce_kernel(T *__restrict__ C, const T *__restrict__ scratch, int64_t sstride,
                                     const BlkD *__restrict__ blks) {
  const BlkD b = blks[blockIdx.y];
  const int64_t n = (int64_t)b.M * b.N;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T v = C[b.c_off + i];
#pragma unroll
    for (int s = 0; s < SPLIT - 1; ++s) v += scratch[(int64_t)s * sstride + b.c_off + i];
    C[b.c_off + i] = v;
  }
}

// Split-K launch (FP32 128x128 tiles, plain writes): SPLIT parts per tile,
// then the reduction.  out_span: one past the largest output element the
// launch writes (c_off + M N over its blocks); nblk: its output blocks.
template <typename T, int BM, int BN, int TAG, int SPLIT>
void launch_split(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                  const BlkD *blks, const PairD *pairs, int64_t out_span, int nblk, int64_t max_mn) {
  BufP scratch = c->alloc((size_t)(SPLIT - 1) * (size_t)out_span * sizeof(T));
  T *sp = static_cast<T *>(scratch->p);
  launch_tiles_v<T, BM, BN, TAG, 16, 2, SPLIT>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, 0, sp, out_span);
  const unsigned gx = (unsigned)std::min<int64_t>((max_mn + 255) / 256, 1024);
  splitk_reduce_kernel<T, SPLIT><<<dim3(std::max(gx, 1u), (unsigned)nblk), 256, 0, c->stream>>>(C, sp, out_span, blks);
}

// Variants: 0 = registers -> LDS double buffer; 4 = direct-to-LDS with four
// stages (FP32 128x128 tiles of row-major A and B); 8 / 9 = split-K in 2 / 4
// parts (FP32 128x128 and FP64 64x64 tiles, launches that pass their output
// extent).  MPS2_GEMM_VARIANT (0, 4, 8 or 9) forces one for every such
// launch; otherwise the caller's choice runs.
int forced_variant() {
  static const int v = [] {
    const char *e = std::getenv("MPS2_GEMM_VARIANT");
    if (!e) return -1;
    const int x = std::atoi(e);
    return x == 0 || x == 4 || x == 8 || x == 9 ? x : -1;
  }();
  return v;
}

template <typename T, int BM, int BN, int TAG>
void launch_tiles(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                  const BlkD *blks, const PairD *pairs, int sub, int variant, const SplitInfo &si) {
  const int f = forced_variant();
  const int v = f >= 0 ? f : variant;
  if constexpr (sizeof(T) == 4) {
    if (c->f32_split) {
      if (v == 9 && sub == 0 && si.out_span > 0 && si.nblk > 0 && si.nblk <= 65535) {
        BufP scratch = c->alloc((size_t)3 * (size_t)si.out_span * sizeof(float));
        float *sp = static_cast<float *>(scratch->p);
        launch_tiles_bf3<BM, BN, TAG, 4>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, 0, sp, si.out_span);
        const unsigned gx = (unsigned)std::min<int64_t>((si.max_mn + 255) / 256, 1024);
        splitk_reduce_kernel<float, 4><<<dim3(std::max(gx, 1u), (unsigned)si.nblk), 256, 0, c->stream>>>(
            C, sp, si.out_span, blks);
      } else {
        launch_tiles_bf3<BM, BN, TAG>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub);
      }
      return;
    }
  }
  if constexpr (sizeof(T) == 4 && BM == 128) {
    if (v == 4 && la == 0 && lb == 0) {
      grouped_gemm_dma_kernel<4, TAG><<<dim3(ntiles), dim3(kThreads), 0, c->stream>>>(A, B, C, tiles, blks, pairs,
                                                                                     ntiles, sub);
      return;
    }
  }
  // split-K: FP32 128x128 and FP64 64x64 tiles
  if constexpr ((sizeof(T) == 4 && BM == 128) || (sizeof(T) == 8 && BM == 64)) {
    if ((v == 8 || v == 9) && sub == 0 && si.out_span > 0 && si.nblk > 0 && si.nblk <= 65535) {
      if (v == 8) launch_split<T, BM, BN, TAG, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, si.out_span, si.nblk, si.max_mn);
      else launch_split<T, BM, BN, TAG, 4>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, si.out_span, si.nblk, si.max_mn);
      return;
    }
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
namespace {
int xcd_queues() {
  static const int v = [] {
    const char *e = std::getenv("MPS2_XCDS");
    int x = e ? std::atoi(e) : 1;
    return x == 8 ? 8 : 1;
  }();
  return v;
}
int tile_group() {
  static const int v = [] {
    const char *e = std::getenv("MPS2_TILE_GROUP");
    int g = e ? std::atoi(e) : (xcd_queues() == 8 ? 4 : 1);
    return g < 1 ? 1 : (g > 8 ? 8 : g);
  }();
  return v;
}
}  // namespace

std::vector<int32_t> tile_block_order(const std::vector<BlkD> &blks, const std::vector<int64_t> &kc) {
  if (xcd_queues() != 1 || tile_group() != 1) return {};
  std::vector<int32_t> order(blks.size());
  for (size_t b = 0; b < blks.size(); ++b) order[b] = (int32_t)b;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    return std::max<int64_t>(kc[x], 1) > std::max<int64_t>(kc[y], 1);
  });
  return order;
}

std::vector<TileD> schedule_tiles(const std::vector<BlkD> &blks, const std::vector<int64_t> &kc, int ts) {
  const int kXcds = xcd_queues(), kGroup = tile_group();
  if (kXcds == 1 && kGroup == 1) {
    // one queue of single tiles, heaviest first: the tiles of a block share
    // its cost, so order the blocks (stably) and emit each block's tiles
    const std::vector<int32_t> order = tile_block_order(blks, kc);
    std::vector<TileD> out;
    for (int b : order) {
      const int tm = (blks[b].M + ts - 1) / ts, tn = (blks[b].N + ts - 1) / ts;
      for (int i = 0; i < tm; ++i)
        for (int j = 0; j < tn; ++j) out.push_back(TileD{b, i * ts, j * ts, 0});
    }
    return out;
  }
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
void launch_grouped_gemm(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                         const BlkD *blks, const PairD *pairs, int tile_m, int tile_n, int sub, int tag,
                         int variant, const SplitInfo &si) {
  if (ntiles <= 0) return;
  const bool big = tile_m == 128 && tile_n == 128;
  if (tag == 1) {
    if (big) launch_tiles<T, 128, 128, 1>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub, variant, si);
    else launch_tiles<T, 64, 64, 1>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub, variant, si);
  } else if (tag == 2) {
    if (big) launch_tiles<T, 128, 128, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub, variant, si);
    else launch_tiles<T, 64, 64, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub, variant, si);
  } else {
    if (big) launch_tiles<T, 128, 128, 0>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub, variant, si);
    else launch_tiles<T, 64, 64, 0>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, sub, variant, si);
  }
  c->gemm_launches++;
}

template <typename T>
void launch_narrow_gemm(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                        const BlkD *blks, const PairD *pairs, int nc, int sub, int tag) {
  if (ntiles <= 0) return;
  if (tag == 1) launch_narrow_tag<T, 1>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, nc, sub);
  else if (tag == 2) launch_narrow_tag<T, 2>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, nc, sub);
  else launch_narrow_tag<T, 0>(c, la, lb, A, B, C, tiles, ntiles, blks, pairs, nc, sub);
  c->gemm_launches++;
}

std::vector<TileD> narrow_tiles(const std::vector<BlkD> &blks, int nc) {
  std::vector<TileD> t;
  for (size_t b = 0; b < blks.size(); ++b)
    for (int m0 = 0; m0 < blks[b].M; m0 += kNarrowRows)
      for (int n0 = 0; n0 < blks[b].N; n0 += nc) t.push_back(TileD{(int32_t)b, m0, n0, 0});
  return t;
}

template void launch_narrow_gemm<double>(Ctx *, int, int, const double *, const double *, double *, const TileMap &, int,
                                         const BlkD *, const PairD *, int, int, int);
template void launch_narrow_gemm<float>(Ctx *, int, int, const float *, const float *, float *, const TileMap &, int,
                                        const BlkD *, const PairD *, int, int, int);

template void launch_grouped_gemm<double>(Ctx *, int, int, const double *, const double *, double *, const TileMap &,
                                          int, const BlkD *, const PairD *, int, int, int, int, int, const SplitInfo &);
template void launch_grouped_gemm<float>(Ctx *, int, int, const float *, const float *, float *, const TileMap &, int,
                                         const BlkD *, const PairD *, int, int, int, int, int, const SplitInfo &);

}  // namespace mps2
