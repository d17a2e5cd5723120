// Truncated SVD of a block-sparse tensor on the device, the counterpart of
// gqten's SVD as called by TwoSiteFiniteVMPSUpdate
// (include/gqmps2/algorithm/vmps/two_site_update_finite_vmps_impl.h:258-263)
// and FiniteMPS::{Left,Right}CanonicalizeTen_ (finite_mps.h:172-243).
//
// Per new-bond QN sector q the state is a dense matrix M_q (rows: left-leg
// tuples, columns: right-leg tuples, taken from blocks holding a nonzero
// element, both in lexicographic tuple order).  The side that must come out
// orthonormal (left for a left-canonical U, right for a right-canonical Vt)
// is found from the eigen-decomposition of its Gram matrix (M M^T or M^T M,
// computed with the grouped MFMA GEMM); singular values are sqrt(eigenvalue).
// The other factor is M projected on that basis (S Vt = U^T M, or U S = M V),
// so both factors of the product used by the sweep are exact even for
// (near-)zero singular values.
//
// Eigen-decomposition (eigh.hip): Householder tridiagonalisation of every
// sector's Gram matrix, tridiagonal divide and conquer, and back-
// transformation of the kept eigenvectors only, all sectors batched per
// launch.  The batched block Jacobi method below (each Gram padded to nb*b,
// b = 32; per round of a round-robin tournament every block pair (I, J)
// diagonalised in LDS, then applied to G and the accumulated eigenvectors
// on MFMA; sweeps until no rotation exceeds 1e-15 * trace(G_q)) remains
// selectable with MPS2_EIGH=jacobi.
//
// Truncation (host, on the singular values only; every nonzero value below
// kResolve s_max = 1e-11 s_max, which no FP64 SVD orders reproducibly,
// counts as kResolve s_max):
//  1. all singular values, sorted descending, ties by (sector, index);
//  2. D_err = smallest D >= 1 with sum_{i>=D} s_i^2 <= trunc_err * sum s_i^2;
//  3. D = min(max(D_err, Dmin), Dmax, count);
//  4. values within relative 1e-6 of the last kept value form the cut cluster;
//     the cut keeps as many of them as step 3 does, chosen by (sector, index).
// New-bond sectors are ordered by ascending QN; states in a sector by
// descending singular value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <chrono>
#include <map>

#include "eigh.h"
#include "runtime.h"

namespace mps2 {

namespace {

constexpr int kB = 32;        // Jacobi block width
constexpr int kP = 2 * kB;    // subproblem size
constexpr double kDegTol = 1e-6;
// Singular values below kResolve * s_max (s_max over all sectors, weights
// below 1e-22) carry no reproducible order: a direct SVD's absolute error is
// p(n) eps s_max, p(n) up to a few hundred, and the refined Gram tail below
// keeps artifacts of up to eps s_max^2 / s_r ~ 1e-12 s_max (leakage of the
// values s_r >= 3e-4 s_max just above the refinement threshold into the tail
// basis).  The truncation treats them as tied.
constexpr double kResolve = kSvdResolve;
constexpr double kEps = 2.220446049250313e-16;

// Multiple of eps in the error bound of a value (MPS2_SVD_SETTLE_BAND,
// default 2, twelve times the error measured on graded spectra, 0.16 eps
// times the problem's largest eigenvalue; 0 keeps refining until the
// uniform accuracy rule is met).
double settle_band() {
  static const double v = [] {
    const char *e = std::getenv("MPS2_SVD_SETTLE_BAND");
    const double x = e ? std::atof(e) : 2.0;
    return x > 0.0 ? x : 0.0;
  }();
  return v;
}

// Jacobi sweeps per 2b x 2b subproblem and per outer round (MPS2_JACOBI_INNER,
// default 1), and the rotation threshold relative to trace(G_q)
// (MPS2_JACOBI_TOL, default 1e-15).
int jacobi_inner_sweeps() {
  static int v = [] {
    const char *e = std::getenv("MPS2_JACOBI_INNER");
    int x = e ? std::atoi(e) : 1;
    return x < 1 ? 1 : (x > 15 ? 15 : x);
  }();
  return v;
}
// Eigensolver of the Gram matrices: tridiagonalisation + divide and conquer
// (default), or the block Jacobi method (MPS2_EIGH=jacobi).
bool use_jacobi() {
  const char *e = std::getenv("MPS2_EIGH");
  return e && std::string(e) == "jacobi";
}
double jacobi_rel_tol() {
  static double v = [] {
    const char *e = std::getenv("MPS2_JACOBI_TOL");
    double x = e ? std::atof(e) : 1e-15;
    return x > 0 && x < 1e-6 ? x : 1e-15;
  }();
  return v;
}

BufP upload(Ctx *c, const void *h, size_t bytes) { return c->upload(h, bytes); }

// Grouped GEMM from host-side block / pair lists (not cached).
void raw_gemm(Ctx *c, int la, int lb, const double *A, const double *B, double *C, const std::vector<BlkD> &blks,
              const std::vector<PairD> &pairs, int sub = 0) {
  std::vector<int64_t> kc(blks.size(), 0);
  for (size_t i = 0; i < blks.size(); ++i)
    for (int p = blks[i].pb; p < blks[i].pe; ++p) kc[i] += pairs[p].K;
  std::vector<TileD> tiles = schedule_tiles(blks, kc, 64);
  if (tiles.empty()) return;
  auto dt = upload(c, tiles.data(), tiles.size() * sizeof(TileD));
  auto db = upload(c, blks.data(), blks.size() * sizeof(BlkD));
  auto dp = upload(c, pairs.data(), pairs.size() * sizeof(PairD));
  launch_grouped_gemm<double>(c, la, lb, A, B, C, static_cast<const TileD *>(dt->p), (int)tiles.size(),
                              static_cast<const BlkD *>(db->p), static_cast<const PairD *>(dp->p), 64, 64, sub);
}

// Largest |x| over n device doubles (NaN propagates); host value.
double dev_maxabs(Ctx *c, const double *x, int64_t n) {
  if (n <= 0) return 0.0;
  const int64_t off = 0;
  auto doff = upload(c, &off, sizeof(int64_t));
  auto dsz = upload(c, &n, sizeof(int64_t));
  auto dout = c->alloc(sizeof(double));
  k_block_maxabs(c, x, static_cast<const int64_t *>(doff->p), static_cast<const int64_t *>(dsz->p), 1,
                 static_cast<double *>(dout->p));
  double h = 0;
  MPS2_HIP(hipMemcpyAsync(&h, dout->p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  return h;
}

// Eigen-decomposition with the stage named in any numeric error.
EighResult eigh_named(Ctx *c, double *G, const std::vector<int64_t> &goff, const std::vector<int> &ns,
                      const std::string &stage) {
  try {
    return eigh_batched(c, G, goff, ns);
  } catch (const Error &e) {
    int nmax = 0;
    for (int x : ns) nmax = std::max(nmax, x);
    throw Error(e.code, "SVD " + stage + " (" + std::to_string(ns.size()) + " sectors, largest " +
                            std::to_string(nmax) + "): " + e.what());
  }
}

struct SecDesc {     // per sector, device
  int64_t g_off;     // offset of padded Gram / eigenvector matrices
  int32_t np;        // padded size nb * kB
  int32_t nb;        // number of blocks (even)
  double tau;        // rotation threshold
};

struct PairWork {    // one (sector, I, J) subproblem
  int32_t sec, I, J, slot;   // slot: Q storage index
};

// Tournament (circle method) pairing of nb players in round r (nb even).
void tournament(int nb, int r, std::vector<std::pair<int, int>> &out) {
  out.clear();
  int m = nb - 1;
  for (int k = 0; k < nb / 2; ++k) {
    int p, q;
    if (k == 0) {
      p = nb - 1;
      q = r % m;
    } else {
      p = (r + k) % m;
      q = (r - k + m) % m;
    }
    out.push_back({std::min(p, q), std::max(p, q)});
  }
}

__device__ inline void jacobi_rot(double app, double aqq, double apq, double &c, double &s) {
  double zeta = (aqq - app) / (2.0 * apq);
  double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  c = 1.0 / sqrt(1.0 + t * t);
  s = t * c;
}

// Diagonalise the 64x64 subproblem [[G_II, G_IJ], [G_IJ^T, G_JJ]] by cyclic
// Jacobi in LDS; write the accumulated rotation Q (64x64, row-major).
// Each of the 63 rounds of a sweep rotates 32 disjoint index pairs (p_k, q_k).
// The 32 x 32 blocks {p_k, q_k} x {p_k', q_k'} of S are disjoint, so thread
// (k, k') applies both rotations to its own 2x2 block in one step
// (S <- G S G^T), and 64 x 32 thread-pairs rotate the columns of Q.
constexpr int kEigThreads = 1024;

__global__ __launch_bounds__(kEigThreads) void pair_eig_kernel(const double *G, const SecDesc *secs,
                                                               const PairWork *work, double *Qout, int *rotated,
                                                               const int *active, int max_inner) {
  __shared__ double S[kP][kP + 1];
  __shared__ double Q[kP][kP + 1];
  __shared__ double cs[kP / 2], sn[kP / 2];
  __shared__ int pp[kP / 2], qq[kP / 2];
  __shared__ int any_rot, any_sweep;
  const PairWork w = work[blockIdx.x];
  if (!active[w.sec]) return;
  const SecDesc sd = secs[w.sec];
  const double *g = G + sd.g_off;
  const int np = sd.np;
  const int tid = threadIdx.x;
  for (int e = tid; e < kP * kP; e += kEigThreads) {
    int i = e / kP, j = e % kP;
    int gi = (i < kB ? w.I * kB + i : w.J * kB + (i - kB));
    int gj = (j < kB ? w.I * kB + j : w.J * kB + (j - kB));
    double v1 = g[(int64_t)gi * np + gj], v2 = g[(int64_t)gj * np + gi];
    S[i][j] = 0.5 * (v1 + v2);
    Q[i][j] = (i == j) ? 1.0 : 0.0;
  }
  if (tid == 0) any_rot = 0;
  __syncthreads();
  const double tau = sd.tau;
  const int bk = tid >> 5, bk2 = tid & 31;   // this thread's 2x2 block of S
  for (int sweep = 0; sweep < max_inner; ++sweep) {
    if (tid == 0) any_sweep = 0;
    __syncthreads();
    for (int r = 0; r < kP - 1; ++r) {
      if (tid < kP / 2) {
        int k = tid, m = kP - 1, p, q;
        if (k == 0) {
          p = kP - 1;
          q = r;
        } else {
          p = (r + k) % m;
          q = (r - k + m) % m;
        }
        if (p > q) {
          int t = p;
          p = q;
          q = t;
        }
        pp[k] = p;
        qq[k] = q;
        double apq = S[p][q];
        if (fabs(apq) > tau) {
          double c, s;
          jacobi_rot(S[p][p], S[q][q], apq, c, s);
          cs[k] = c;
          sn[k] = s;
          any_sweep = 1;
        } else {
          cs[k] = 1.0;
          sn[k] = 0.0;
        }
      }
      __syncthreads();
      {
        const double c1 = cs[bk], s1 = sn[bk], c2 = cs[bk2], s2 = sn[bk2];
        if (s1 != 0.0 || s2 != 0.0) {
          const int p = pp[bk], q = qq[bk], p2 = pp[bk2], q2 = qq[bk2];
          const double a = S[p][p2], b = S[p][q2], cc = S[q][p2], d = S[q][q2];
          const double r0 = c1 * a - s1 * cc, r1 = c1 * b - s1 * d;   // rotated rows
          const double r2 = s1 * a + c1 * cc, r3 = s1 * b + c1 * d;
          double npp = c2 * r0 - s2 * r1, npq = s2 * r0 + c2 * r1;   // rotated columns
          double nqp = c2 * r2 - s2 * r3, nqq = s2 * r2 + c2 * r3;
          if (bk == bk2) {   // the pivot block: its off-diagonal is annihilated
            npq = 0.0;
            nqp = 0.0;
          }
          S[p][p2] = npp;
          S[p][q2] = npq;
          S[q][p2] = nqp;
          S[q][q2] = nqq;
        }
      }
#pragma unroll
      for (int u = 0; u < (kP / 2) * kP / kEigThreads; ++u) {
        const int e = tid + u * kEigThreads, k = e >> 6, i = e & 63;
        const double c = cs[k], s = sn[k];
        if (s != 0.0) {
          const int p = pp[k], q = qq[k];
          const double x = Q[i][p], y = Q[i][q];
          Q[i][p] = c * x - s * y;
          Q[i][q] = s * x + c * y;
        }
      }
      __syncthreads();
    }
    int done = !any_sweep;
    if (!done && tid == 0) any_rot = 1;
    __syncthreads();
    if (done) break;
  }
  double *qo = Qout + (int64_t)w.slot * kP * kP;
  for (int e = tid; e < kP * kP; e += kEigThreads) qo[e] = Q[e / kP][e % kP];
  if (tid == 0 && any_rot) atomicOr(rotated + w.sec, 1);
}

// One 64x64 tile of the round's update.  The round's pairs partition the
// column blocks, so tiles (pair P rows, pair P2 columns) partition the matrix
// and every tile is read and written by its own workgroup only:
//   mat 0:  G[P, P2] <- Q_P^T G[P, P2] Q_P2
//   mat 1:  V[P, P2] <- V[P, P2] Q_P2
struct TileWork {
  int32_t P, P2;   // absolute indices into the PairWork list (same sector)
  int32_t mat, pad;
};

__device__ inline int pair_row(const PairWork &w, int i) { return i < kB ? w.I * kB + i : w.J * kB + (i - kB); }

__global__ __launch_bounds__(256) void apply_tiles_kernel(double *G, double *V, const SecDesc *secs,
                                                          const PairWork *pw, const TileWork *tw, const double *Qall,
                                                          const int *active) {
  __shared__ double X[kP][kP + 1];
  __shared__ double Q1[kP][kP + 1];
  __shared__ double Q2[kP][kP + 1];
  const TileWork t = tw[blockIdx.x];
  const PairWork w1 = pw[t.P], w2 = pw[t.P2];
  if (!active[w1.sec]) return;
  const SecDesc sd = secs[w1.sec];
  const int np = sd.np, tid = threadIdx.x;
  double *g = (t.mat == 0 ? G : V) + sd.g_off;
  const double *q1 = Qall + (int64_t)w1.slot * kP * kP;
  const double *q2 = Qall + (int64_t)w2.slot * kP * kP;
  for (int e = tid; e < kP * kP; e += 256) {
    int i = e / kP, j = e % kP;
    X[i][j] = g[(int64_t)pair_row(w1, i) * np + pair_row(w2, j)];
    Q2[i][j] = q2[e];
    if (t.mat == 0) Q1[i][j] = q1[e];
  }
  __syncthreads();
  // 64x64x64 products on FP64 MFMA (16x16x4): wave w owns output rows
  // 16*w .. 16*w+15 (four 16x16 fragments along the columns).  Lane l feeds
  // A[l%16][l/16] and B[l/16][l%16]; accumulator element r of lane l is row
  // l/16 + 4r, column l%16 of the fragment.
  typedef double d4 __attribute__((ext_vector_type(4)));
  const int lane = tid & 63, wv = tid >> 6;
  const int i0 = 16 * wv;
  d4 acc[4];
  if (t.mat == 0) {
    // T = Q1^T X
#pragma unroll
    for (int f = 0; f < 4; ++f) acc[f] = d4{0, 0, 0, 0};
    for (int k0 = 0; k0 < kP; k0 += 4) {
      const double a = Q1[k0 + (lane >> 4)][i0 + (lane & 15)];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const double b = X[k0 + (lane >> 4)][16 * f + (lane & 15)];
        acc[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[f], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int r = 0; r < 4; ++r) X[i0 + (lane >> 4) + 4 * r][16 * f + (lane & 15)] = acc[f][r];
    __syncthreads();
  }
  // Y = X Q2
#pragma unroll
  for (int f = 0; f < 4; ++f) acc[f] = d4{0, 0, 0, 0};
  for (int k0 = 0; k0 < kP; k0 += 4) {
    const double a = X[i0 + (lane & 15)][k0 + (lane >> 4)];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const double b = Q2[k0 + (lane >> 4)][16 * f + (lane & 15)];
      acc[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[f], 0, 0, 0);
    }
  }
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) + 4 * r, j = 16 * f + (lane & 15);
      g[(int64_t)pair_row(w1, i) * np + pair_row(w2, j)] = acc[f][r];
    }
}

// Copy the n x n Gram into the padded np x np layout; set V = I.
__global__ void pad_init_kernel(const double *Gsrc, int64_t src_off, int n, double *G, double *V, int64_t g_off,
                                int np) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)np * np;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(e / np), j = (int)(e % np);
    G[g_off + e] = (i < n && j < n) ? Gsrc[src_off + (int64_t)i * n + j] : 0.0;
    V[g_off + e] = (i == j) ? 1.0 : 0.0;
  }
}

// out[k] = diag entry, and gather selected eigenvector columns:
// E[i * kq + t] = V[i * np + col[t]] for i < n.
__global__ void gather_cols_kernel(const double *V, int64_t g_off, int np, int n, const int32_t *cols, int kq,
                                   double *E) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * kq;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(e / kq), t = (int)(e % kq);
    E[e] = V[g_off + (int64_t)i * np + cols[t]];
  }
}

__global__ void diag_kernel(const double *G, int64_t g_off, int np, int n, double *out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = G[g_off + (int64_t)i * np + i];
}



// Row (mode 0) or column (mode 1) scaling of one matrix per blockIdx.y.
struct ScaleD {
  int64_t off, sc_off;
  int32_t rows, cols;
  int32_t mode, pad;
};
__global__ void scale_batch_kernel(double *X, const double *sc, const ScaleD *desc) {
  const ScaleD d = desc[blockIdx.y];
  const int64_t n = (int64_t)d.rows * d.cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / d.cols, j = e - i * d.cols;
    X[d.off + e] *= sc[d.sc_off + (d.mode == 0 ? i : j)];
  }
}

// Sums of squares of the rows (mode 0: one wavefront per row, lanes along
// the row) or columns (mode 1: one thread per column) of each descriptor's
// block (rows x cols, row-major at off); results at out[sc_off + i].
__global__ __launch_bounds__(256) void sumsq_batch_kernel(const double *__restrict__ X, double *out,
                                                          const ScaleD *__restrict__ desc) {
  const ScaleD d = desc[blockIdx.y];
  if (d.mode == 0) {
    const int lane = threadIdx.x & 63;
    for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < d.rows; i += gridDim.x * 4) {
      const double *row = X + d.off + (int64_t)i * d.cols;
      double acc = 0;
      for (int j = lane; j < d.cols; j += 64) acc += row[j] * row[j];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
      if (lane == 0) out[d.sc_off + i] = acc;
    }
  } else {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < d.cols; j += gridDim.x * blockDim.x) {
      double acc = 0;
      for (int i = 0; i < d.rows; ++i) {
        const double x = X[d.off + (int64_t)i * d.cols + j];
        acc += x * x;
      }
      out[d.sc_off + j] = acc;
    }
  }
}

// Kept eigenvectors of a refined sector from the level-1 vectors computed
// for its refinement: E[e_off + i*kq + t] = map[t] >= 0 ? ES[i][map[t]]
// (tail basis, n x r) : EL[i][-1 - map[t]] (large directions, n x nl).
struct KeepD {
  int64_t e_off, s_off, l_off, map_off;
  int32_t n, kq, r, nl;
};
__global__ void gather_kept_kernel(const double *__restrict__ ES, const double *__restrict__ EL,
                                   const int32_t *__restrict__ map, const KeepD *__restrict__ desc, double *E) {
  const KeepD d = desc[blockIdx.y];
  const int64_t tot = (int64_t)d.n * d.kq;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < tot; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = g / d.kq;
    const int t = (int)(g - i * d.kq);
    const int m = map[d.map_off + t];
    E[d.e_off + g] = m >= 0 ? ES[d.s_off + i * d.r + m] : EL[d.l_off + i * d.nl + (-1 - m)];
  }
}

// Transposed block copies, one descriptor per blockIdx.x:
// dst[dst_off + c * ldd + r] = src[src_off + r * lds + c], r < rows, c < cols.
__global__ void copy2d_t_kernel(const double *src, double *dst, const Copy2D *desc) {
  const Copy2D d = desc[blockIdx.x];
  const int64_t n = (int64_t)d.rows * d.cols;
  for (int64_t e = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.y * blockDim.x) {
    const int64_t r = e / d.cols, cc = e - r * d.cols;
    dst[d.dst_off + cc * d.ldd + r] = d.scale * src[d.src_off + r * d.lds + cc];
  }
}

inline int gridn(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096); }

// Batched symmetric eigen-decomposition of Gram matrices (contiguous n_q x n_q
// row-major at goffs[q] in Gsrc).  Returns eigenvalues (unsorted, per sector)
// and leaves padded eigenvectors in V (np_q x np_q at the returned offsets).
struct EigOut {
  BufP Gp, V;
  std::vector<SecDesc> secs;
  std::vector<std::vector<double>> evals;
};

EigOut batched_eigh(Ctx *c, const double *Gsrc, const std::vector<int64_t> &goffs, const std::vector<int> &ns,
                    const std::vector<double> &traces) {
  int nsec = (int)ns.size();
  EigOut eo;
  eo.secs.resize(nsec);
  int64_t tot = 0;
  int maxnb = 0;
  for (int q = 0; q < nsec; ++q) {
    int nb = (ns[q] + kB - 1) / kB;
    if (nb < 2) nb = 2;
    if (nb & 1) nb++;
    eo.secs[q].g_off = tot;
    eo.secs[q].nb = nb;
    eo.secs[q].np = nb * kB;
    eo.secs[q].tau = std::max(jacobi_rel_tol() * traces[q], 1e-300);
    tot += (int64_t)nb * kB * nb * kB;
    maxnb = std::max(maxnb, nb);
  }
  eo.Gp = c->alloc(std::max<int64_t>(tot, 1) * sizeof(double));
  eo.V = c->alloc(std::max<int64_t>(tot, 1) * sizeof(double));
  double *G = static_cast<double *>(eo.Gp->p), *V = static_cast<double *>(eo.V->p);
  for (int q = 0; q < nsec; ++q)
    pad_init_kernel<<<gridn((int64_t)eo.secs[q].np * eo.secs[q].np), 256, 0, c->stream>>>(
        Gsrc, goffs[q], ns[q], G, V, eo.secs[q].g_off, eo.secs[q].np);
  auto dsecs = upload(c, eo.secs.data(), nsec * sizeof(SecDesc));
  // Work lists for every round of one sweep: pair subproblems, then the
  // (pair rows x pair columns) tiles of G and V.
  std::vector<PairWork> pw;
  std::vector<TileWork> tw;
  std::vector<int> round_p0(maxnb, 0), round_p1(maxnb, 0), round_t0(maxnb, 0), round_t1(maxnb, 0);
  std::vector<std::pair<int, int>> tp;
  int slots = 0;
  for (int r = 0; r < maxnb - 1; ++r) {
    round_p0[r] = (int)pw.size();
    round_t0[r] = (int)tw.size();
    int slot = 0;
    for (int q = 0; q < nsec; ++q) {
      int nb = eo.secs[q].nb;
      if (r >= nb - 1) continue;
      tournament(nb, r, tp);
      int first = (int)pw.size();
      for (auto &pq : tp) pw.push_back(PairWork{q, pq.first, pq.second, slot++});
      int last = (int)pw.size();
      for (int a = first; a < last; ++a)
        for (int b = first; b < last; ++b) {
          tw.push_back(TileWork{a, b, 0, 0});
          tw.push_back(TileWork{a, b, 1, 0});
        }
    }
    slots = std::max(slots, slot);
    round_p1[r] = (int)pw.size();
    round_t1[r] = (int)tw.size();
  }
  auto dpw = upload(c, pw.data(), pw.size() * sizeof(PairWork));
  auto dtw = upload(c, tw.data(), tw.size() * sizeof(TileWork));
  auto dq = c->alloc((int64_t)std::max(slots, 1) * kP * kP * sizeof(double));
  auto drot = c->alloc(std::max(nsec, 1) * sizeof(int));
  auto dact = c->alloc(std::max(nsec, 1) * sizeof(int));
  const PairWork *PW = static_cast<const PairWork *>(dpw->p);
  const TileWork *TW = static_cast<const TileWork *>(dtw->p);
  const SecDesc *SD = static_cast<const SecDesc *>(dsecs->p);
  double *Q = static_cast<double *>(dq->p);
  int *ROT = static_cast<int *>(drot->p);
  int *ACT = static_cast<int *>(dact->p);
  std::vector<int> hrot(nsec), hact(nsec, 1);
  MPS2_HIP(hipMemcpyAsync(ACT, hact.data(), nsec * sizeof(int), hipMemcpyHostToDevice, c->stream));
  for (int sweep = 0; sweep < 40; ++sweep) {
    MPS2_HIP(hipMemsetAsync(ROT, 0, nsec * sizeof(int), c->stream));
    for (int r = 0; r < maxnb - 1; ++r) {
      int np_ = round_p1[r] - round_p0[r], nt = round_t1[r] - round_t0[r];
      if (np_ == 0) continue;
      pair_eig_kernel<<<np_, kEigThreads, 0, c->stream>>>(G, SD, PW + round_p0[r], Q, ROT, ACT,
                                                          jacobi_inner_sweeps());
      apply_tiles_kernel<<<nt, 256, 0, c->stream>>>(G, V, SD, PW, TW + round_t0[r], Q, ACT);
    }
    MPS2_HIP(hipMemcpyAsync(hrot.data(), ROT, nsec * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    bool any = false;
    for (int q = 0; q < nsec; ++q) {
      hact[q] = hact[q] && hrot[q] != 0;
      any |= hact[q] != 0;
    }
    c->count("svd_sweeps");
    c->count("svd_rounds", maxnb - 1);
    if (!any) break;
    MPS2_HIP(hipMemcpyAsync(ACT, hact.data(), nsec * sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  // eigenvalues
  eo.evals.resize(nsec);
  auto dd = c->alloc(std::max<int64_t>(tot, 1) * sizeof(double));
  for (int q = 0; q < nsec; ++q)
    diag_kernel<<<gridn(ns[q]), 256, 0, c->stream>>>(G, eo.secs[q].g_off, eo.secs[q].np, ns[q],
                                                    static_cast<double *>(dd->p) + eo.secs[q].g_off);
  for (int q = 0; q < nsec; ++q) {
    eo.evals[q].resize(ns[q]);
    if (ns[q])
      MPS2_HIP(hipMemcpyAsync(eo.evals[q].data(), static_cast<double *>(dd->p) + eo.secs[q].g_off,
                              ns[q] * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  c->sync();
  return eo;
}

struct SvInfo {
  double v;
  int sec;
  int j;
};

void truncate_rule(const std::vector<std::vector<double>> &sv, double trunc_err, int64_t dmin, int64_t dmax,
                   std::vector<int64_t> &kept, double *err, int64_t *D, int64_t *derr_out = nullptr) {
  std::vector<SvInfo> all;
  for (int q = 0; q < (int)sv.size(); ++q)
    for (int j = 0; j < (int)sv[q].size(); ++j) all.push_back({sv[q][j], q, j});
  std::sort(all.begin(), all.end(), [](const SvInfo &x, const SvInfo &y) {
    if (x.v != y.v) return x.v > y.v;
    if (x.sec != y.sec) return x.sec < y.sec;
    return x.j < y.j;
  });
  int64_t cnt = (int64_t)all.size();
  double total = 0;
  for (auto &x : all) total += x.v * x.v;
  int64_t derr = cnt;
  {
    double tail = 0;
    for (int64_t i = cnt - 1; i >= 1; --i) {
      tail += all[i].v * all[i].v;
      if (tail <= trunc_err * total) derr = i;
      else break;
    }
  }
  if (derr_out) *derr_out = derr;
  int64_t d = std::max(derr, dmin);
  d = std::min(d, dmax);
  d = std::min(d, cnt);
  if (d < 1 && cnt > 0) d = 1;
  std::vector<char> keep(cnt, 0);
  for (int64_t i = 0; i < d; ++i) keep[i] = 1;
  if (d > 0 && d < cnt) {
    double v = all[d - 1].v;
    std::vector<int64_t> cl;
    int64_t kin = 0;
    for (int64_t i = 0; i < cnt; ++i)
      if (std::fabs(all[i].v - v) <= kDegTol * v) {
        cl.push_back(i);
        if (i < d) ++kin;
      }
    std::sort(cl.begin(), cl.end(), [&](int64_t x, int64_t y) {
      if (all[x].sec != all[y].sec) return all[x].sec < all[y].sec;
      return all[x].j < all[y].j;
    });
    for (int64_t t = 0; t < (int64_t)cl.size(); ++t) keep[cl[t]] = t < kin ? 1 : 0;
  }
  kept.assign(sv.size(), 0);
  double disc = 0;
  for (int64_t i = 0; i < cnt; ++i) {
    if (keep[i]) kept[all[i].sec]++;
    else disc += all[i].v * all[i].v;
  }
  *err = total > 0 ? disc / total : 0.0;
  *D = d;
}

}  // namespace

void svd_truncate_rule(const std::vector<std::vector<double>> &sv, double trunc_err, int64_t dmin, int64_t dmax,
                       std::vector<int64_t> &kept, double *err, int64_t *D) {
  truncate_rule(sv, trunc_err, dmin, dmax, kept, err, D);
}

SvdOut svd_impl(const TenP &t, int ldims, const QN &ldiv, double trunc_err, int64_t dmin, int64_t dmax,
                OrthSide side);

SvdOut svd(const TenP &t, int ldims, const QN &ldiv, double trunc_err, int64_t dmin, int64_t dmax, OrthSide side) {
  auto t0 = std::chrono::steady_clock::now();
  SvdOut o = t->cplx ? svd_complex(t, ldims, ldiv, trunc_err, dmin, dmax, side)
                     : svd_impl(t, ldims, ldiv, trunc_err, dmin, dmax, side);
  t->ctx->count("svd_calls");
  t->ctx->count("svd_ms", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  return o;
}

SvdOut svd_impl(const TenP &t, int ldims, const QN &ldiv, double trunc_err, int64_t dmin, int64_t dmax,
                OrthSide side) {
  MPS2_REQUIRE(t->prec == kF64, "SVD: FP64 tensors only");
  Ctx *c = t->ctx;
  const Meta &m = *t->meta;
  int r = m.rank();
  MPS2_REQUIRE(ldims >= 1 && ldims < r, "SVD: ldims out of range");
  int nblk = (int)m.coords.size();
  // 1. nonzero blocks
  std::vector<double> bmax(nblk, 0.0);
  if (nblk) {
    auto doff = upload(c, m.offset.data(), nblk * sizeof(int64_t));
    auto dsz = upload(c, m.bsize.data(), nblk * sizeof(int64_t));
    auto dout = c->alloc(nblk * sizeof(double));
    k_block_maxabs(c, t->d(), static_cast<const int64_t *>(doff->p), static_cast<const int64_t *>(dsz->p), nblk,
                   static_cast<double *>(dout->p));
    MPS2_HIP(hipMemcpyAsync(bmax.data(), dout->p, nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->sync();
  }
  // The Gram matrices square the entries and the eigensolver squares and
  // inverts them again, so unnormalised inputs (a canonicalisation sweep
  // over a random start carries the product of all site norms, 1e80 at 100
  // sites) would overflow, and sectors of negligible weight (1e-63 of the
  // largest entry at the ends of a long chain) underflow.  Each sector matrix
  // is assembled scaled by 2^-e_q with 2^e_q <= max|M_q| < 2^(e_q+1): exact
  // in binary floating point, so U and V^T are unchanged.  Host-side values
  // are kept in one unit, the tensor scaled by 2^-e with e from the largest
  // entry overall (S is scaled back by 2^e); a sector's device values differ
  // from it by the exact factor 2^(e - e_q) (eigenvalues 4^(e - e_q)).
  double gmax = 0.0;
  for (int b = 0; b < nblk; ++b) {
    if (!std::isfinite(bmax[b])) throw Error(kErrNumeric, "SVD: input tensor holds NaN or Inf");
    gmax = std::max(gmax, bmax[b]);
  }
  const int sexp = gmax > 0.0 ? std::max(-1000, std::min(1000, std::ilogb(gmax))) : 0;
  const double pre = std::ldexp(1.0, -sexp), post = std::ldexp(1.0, sexp);
  // 2. sector structure
  struct Sec {
    QN qn;
    std::map<Coord, int64_t> rows, cols;
    int64_t m = 0, n = 0;
  };
  std::map<QN, Sec> smap;
  auto kqn = [&](const Coord &cc) {
    QN q = ldiv;
    for (int i = 0; i < ldims; ++i) q = m.legs[i].dir == kOut ? q - m.legs[i].qn[cc[i]] : q + m.legs[i].qn[cc[i]];
    return q;
  };
  for (int b = 0; b < nblk; ++b) {
    if (!(bmax[b] > 0.0)) continue;
    const Coord &cc = m.coords[b];
    Sec &s = smap[kqn(cc)];
    s.qn = kqn(cc);
    s.rows[Coord(cc.begin(), cc.begin() + ldims)] = 0;
    s.cols[Coord(cc.begin() + ldims, cc.end())] = 0;
  }
  std::vector<Sec> secs;
  for (auto &kv : smap) secs.push_back(kv.second);
  int ns = (int)secs.size();
  int64_t moff_tot = 0;
  std::vector<int64_t> moff(ns);
  for (int q = 0; q < ns; ++q) {
    auto &s = secs[q];
    for (auto &kv : s.rows) {
      kv.second = s.m;
      int64_t d = 1;
      for (int i = 0; i < ldims; ++i) d *= m.legs[i].dim[kv.first[i]];
      s.m += d;
    }
    for (auto &kv : s.cols) {
      kv.second = s.n;
      int64_t d = 1;
      for (int i = ldims; i < r; ++i) d *= m.legs[i].dim[kv.first[i - ldims]];
      s.n += d;
    }
    moff[q] = moff_tot;
    moff_tot += s.m * s.n;
  }
  // per-sector exponents dq = e_q - e (<= 0): device value = host value * 2^-dq
  std::vector<int> dq(ns, 0);
  {
    std::vector<double> smax(ns, 0.0);
    for (int b = 0; b < nblk; ++b) {
      if (!(bmax[b] > 0.0)) continue;
      const int q = (int)std::distance(smap.begin(), smap.find(kqn(m.coords[b])));
      smax[q] = std::max(smax[q], bmax[b]);
    }
    for (int q = 0; q < ns; ++q)
      dq[q] = smax[q] > 0.0 ? std::max(-1000, std::min(1000, std::ilogb(smax[q]))) - sexp : 0;
  }
  // 3. assemble sector matrices
  auto Mb = c->alloc(std::max<int64_t>(moff_tot, 1) * sizeof(double));
  double *Md = static_cast<double *>(Mb->p);
  k_fill(c, Md, std::max<int64_t>(moff_tot, 1), 0.0);
  {
    std::vector<Copy2D> cp;
    for (int b = 0; b < nblk; ++b) {
      if (!(bmax[b] > 0.0)) continue;
      const Coord &cc = m.coords[b];
      int q = (int)std::distance(smap.begin(), smap.find(kqn(cc)));
      auto &s = secs[q];
      int64_t ro = s.rows.at(Coord(cc.begin(), cc.begin() + ldims));
      int64_t co = s.cols.at(Coord(cc.begin() + ldims, cc.end()));
      int64_t bm = 1, bn = 1;
      for (int i = 0; i < ldims; ++i) bm *= m.legs[i].dim[cc[i]];
      for (int i = ldims; i < r; ++i) bn *= m.legs[i].dim[cc[i]];
      cp.push_back(Copy2D{m.offset[b], moff[q] + ro * s.n + co, (int32_t)bm, (int32_t)bn, (int32_t)bn,
                          (int32_t)s.n, std::ldexp(pre, -dq[q])});
    }
    if (!cp.empty()) {
      auto dcp = upload(c, cp.data(), cp.size() * sizeof(Copy2D));
      k_copy2d(c, t->d(), Md, static_cast<const Copy2D *>(dcp->p), (int)cp.size());
    }
  }
  // 4. Gram matrices of the orthonormal side
  std::vector<int> gn(ns);
  std::vector<int64_t> goff(ns);
  int64_t gtot = 0;
  std::vector<BlkD> gb;
  std::vector<PairD> gp;
  for (int q = 0; q < ns; ++q) {
    int64_t n_ = side == kOrthLeft ? secs[q].m : secs[q].n;
    int64_t k_ = side == kOrthLeft ? secs[q].n : secs[q].m;
    gn[q] = (int)n_;
    goff[q] = gtot;
    gb.push_back(BlkD{gtot, (int32_t)n_, (int32_t)n_, (int32_t)gp.size(), (int32_t)gp.size() + 1});
    gp.push_back(PairD{moff[q], moff[q], (int32_t)k_, 0});
    gtot += n_ * n_;
  }
  auto Gb = c->alloc(std::max<int64_t>(gtot, 1) * sizeof(double));
  double *Gd = static_cast<double *>(Gb->p);
  // left: G = M M^T (A = M normal, B = M^T stored as M -> trans)
  // right: G = M^T M (A = M^T stored as M -> trans, B = M normal)
  if (side == kOrthLeft) raw_gemm(c, kNormal, kTrans, Md, Md, Gd, gb, gp);
  else raw_gemm(c, kTrans, kNormal, Md, Md, Gd, gb, gp);
  // traces for the rotation threshold
  std::vector<double> traces(ns, 0.0);
  if (use_jacobi()) {
    auto dd = c->alloc(std::max<int64_t>(gtot, 1) * sizeof(double));
    std::vector<std::vector<double>> dg(ns);
    for (int q = 0; q < ns; ++q) {
      diag_kernel<<<gridn(gn[q]), 256, 0, c->stream>>>(Gd, goff[q], gn[q], gn[q], static_cast<double *>(dd->p) + goff[q]);
      dg[q].resize(gn[q]);
    }
    for (int q = 0; q < ns; ++q)
      if (gn[q])
        MPS2_HIP(hipMemcpyAsync(dg[q].data(), static_cast<double *>(dd->p) + goff[q], gn[q] * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
    c->sync();
    for (int q = 0; q < ns; ++q)
      for (double x : dg[q]) traces[q] += x;
  }
  auto te0 = std::chrono::steady_clock::now();
  const bool jac = use_jacobi();
  EigOut eo;
  EighResult er;
  if (jac) eo = batched_eigh(c, Gd, goff, gn, traces);
  else er = eigh_named(c, Gd, goff, gn, "level-1 eigensolve");
  std::vector<std::vector<double>> evals = jac ? eo.evals : er.evals;
  for (int q = 0; q < ns; ++q)
    for (double &x : evals[q]) x = std::ldexp(x, 2 * dq[q]);
  c->count("svd_eigh_ms", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - te0).count());
  {
    int maxn = 0;
    for (int x : gn) maxn = std::max(maxn, x);
    c->count("svd_sectors", ns);
    c->count("svd_max_sector_sum", maxn);
  }
  // 5. singular values: top min(m, n) eigenvalues per sector, descending.
  // sv holds sqrt(max(lambda, 0)) (S and the factors); svr holds what the
  // truncation rule sees.  Below kResolve * s_max a direct SVD returns
  // rounding noise, tiny but positive, and the Gram eigenvalues there are
  // +-noise.  svr counts every nonzero value below kResolve * s_max
  // (negative lambda included) as exactly kResolve * s_max: with trunc_err = 0
  // all of them are kept, as a direct SVD keeps its positive noise, and a cut
  // among them is a tie, decided by (sector, index) instead of by rounding.
  // An exactly zero lambda (a structural zero, which deflation produces
  // exactly) stays zero and is cut.
  std::vector<std::vector<double>> sv(ns), svr(ns);
  std::vector<std::vector<int32_t>> order(ns);
  auto singular_values = [&]() {
    double top = 0.0;
    for (int q = 0; q < ns; ++q)
      for (double x : evals[q]) top = std::max(top, x);
    const double thr = kResolve * std::sqrt(top);
    for (int q = 0; q < ns; ++q) {
      std::vector<int32_t> idx(gn[q]);
      for (int i = 0; i < gn[q]; ++i) idx[i] = i;
      const auto &ev = evals[q];
      auto rule = [&](int32_t i) {
        return ev[i] > 0.0 ? std::max(std::sqrt(ev[i]), thr) : (ev[i] < 0.0 ? thr : 0.0);
      };
      std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) {
        const double rx = rule(x), ry = rule(y);
        return rx != ry ? rx > ry : ev[x] > ev[y];
      });
      int64_t rq = std::min(secs[q].m, secs[q].n);
      idx.resize(rq);
      order[q] = idx;
      sv[q].resize(rq);
      svr[q].resize(rq);
      for (int64_t j = 0; j < rq; ++j) {
        sv[q][j] = std::sqrt(std::max(ev[idx[j]], 0.0));
        svr[q][j] = rule(idx[j]);
      }
    }
  };
  singular_values();
  SvdOut out;
  truncate_rule(svr, trunc_err, dmin, dmax, out.kept, &out.trunc_err, &out.D);
  // Tail refinement (tridiagonal path).  An eigenvalue of G is accurate to
  // about n eps lambda_max absolutely, i.e. a singular value s only to a
  // relative (s_max / s)^2 eps, while the truncation (and its 1e-6 tie
  // window) needs the values near the cut to a relative error like a direct
  // SVD's, eps s_max / s.  Eigenvalues above kAccept * lambda_max are
  // resolved to ~1e-9 relative.  When the cut lies below that in a sector,
  // the eigenpairs of that sector below the tail threshold (the tail S) are
  // recomputed from the Gram of the projected tail, W_S = E_S^T M (left) or
  // M E_S (right): its rows carry the true small singular values up to
  // eps s_max absolute, and its eigensolver's rounding now scales with the
  // tail's own maximum.  Repeated while the cut still lies below kAccept
  // times the refined tail's maximum.
  std::vector<int> nsub(ns, 0);
  std::vector<int64_t> esoff(ns, 0);
  BufP Esub;
  // Large directions of the refined sectors (level-1 values above the tail):
  // F_L = E_L^T M (left) or M E_L (right), and 1/s^2 per direction.
  std::vector<int> nL(ns, 0);
  std::vector<int64_t> fLoff(ns, 0), iLoff(ns, 0);
  BufP FLb, ILb;
  // level-1 vectors of the sectors refined in round 0: tail (r0 columns) and
  // large directions (nL columns), reused for the kept vectors
  BufP ES0, EL0;
  std::vector<int> r0(ns, -1);
  std::vector<int64_t> es0off(ns, 0), el0off(ns, 0);
  if (!jac) {
    const auto tr0 = std::chrono::steady_clock::now();
    // A cut below kAccept times the current problem's largest eigenvalue is
    // not resolved to the tie window and triggers a round; the round's tail
    // holds the eigenvalues below max(kTail * top, kMargin * cut^2).  The
    // level-1 eigenvalues are accurate to ~0.2 eps * top absolutely (graded
    // spectra of known values, n = 1000), so a tail boundary at 1e-10 top
    // mixes the tail with the large directions by theta < 1e-6, a relative
    // error theta^2 in the refined values; the tail's own maximum, ~1e-10 top,
    // then resolves a cut at 1e-8 s_max in one round.
    constexpr double kAccept = 1e-7, kTail = 1e-10, kMargin = 1e3;
    // phase clocks of the refinement (engine counters svd_refine_{vec,eigh}_ms),
    // device-timeline events
    hipEvent_t tv0 = nullptr;
    std::vector<double> lam_top(ns, 0.0);  // scale of each sector's current problem
    for (int q = 0; q < ns; ++q)
      for (double x : evals[q]) lam_top[q] = std::max(lam_top[q], x);
    // Error scale of each value: the largest eigenvalue of the problem it
    // was last computed from (its eigensolver's error is a few eps times that).
    std::vector<std::vector<double>> esc(ns);
    for (int q = 0; q < ns; ++q) esc[q].assign(evals[q].size(), lam_top[q]);
    // True when no value's place relative to the cut's tie window (relative
    // kDegTol around the D-th largest value) can change within its error
    // bound, band eps (scale + 2 s s_max) (the eigensolver's error and that of
    // the projected tail), and D does not hang on the trunc_err tail sum: a
    // further round would then leave the kept counts as they are.  The kept
    // values near the cut then carry that bound, in s eps (scale / s + 2 s_max)
    // at band 2: at most 12 eps s_max at kResolve s_max, within a direct SVD's
    // p(n) eps s_max.
    auto settled = [&](int64_t derr_now) -> bool {
      const double band = settle_band();
      if (band <= 0.0) return false;
      if (!(trunc_err <= 0.0 || derr_now >= out.D + 2 || derr_now + 2 <= out.D)) return false;
      double top = 0.0;
      for (int q = 0; q < ns; ++q)
        for (double x : evals[q]) top = std::max(top, x);
      const double smax = std::sqrt(top), thr = kResolve * smax;
      auto rule = [&](double x) { return x > 0.0 ? std::max(std::sqrt(x), thr) : thr; };
      auto err = [&](int q, int j) {
        return band * kEps * (esc[q][j] + 2.0 * std::sqrt(std::max(evals[q][j], 0.0)) * smax);
      };
      struct Ent {
        double v;
        int q, j, pos;
      };
      std::vector<Ent> es;
      for (int q = 0; q < ns; ++q)
        for (int t = 0; t < (int)order[q].size(); ++t) es.push_back({svr[q][t], q, order[q][t], t});
      if (out.D <= 0 || out.D >= (int64_t)es.size()) return false;
      std::nth_element(es.begin(), es.begin() + (out.D - 1), es.end(), [](const Ent &x, const Ent &y) {
        if (x.v != y.v) return x.v > y.v;
        if (x.q != y.q) return x.q < y.q;
        return x.pos < y.pos;
      });
      const Ent cut = es[out.D - 1];
      const double lc = evals[cut.q][cut.j], ec = err(cut.q, cut.j);
      const double vlo = rule(lc - ec), vhi = rule(lc + ec);
      for (const Ent &x : es) {
        const double l = evals[x.q][x.j];
        if (l == 0.0) continue;  // a structural zero is always cut
        const double e = err(x.q, x.j), lo = rule(l - e), hi = rule(l + e);
        const bool below = hi < vlo * (1.0 - kDegTol);
        const bool above = lo > vhi * (1.0 + kDegTol);
        const bool inside = lo > vhi * (1.0 - kDegTol) && hi < vlo * (1.0 + kDegTol);
        if (!(below || above || inside)) return false;
      }
      return true;
    };
    for (int round = 0; round < 3; ++round) {
      std::vector<double> all;
      for (auto &v : svr) all.insert(all.end(), v.begin(), v.end());
      if (out.D <= 0 || out.D >= (int64_t)all.size()) break;
      std::nth_element(all.begin(), all.begin() + (out.D - 1), all.end(), std::greater<double>());
      // A cut below kResolve * s_max is a tie among the values counted as
      // kResolve * s_max; what must be resolved then is which values lie
      // above that level, so the threshold takes the cut's place.
      const double cut =
          std::max(all[out.D - 1], kResolve * *std::max_element(all.begin(), all.end()));
      const double cut2 = cut * cut;
      std::vector<int> rs, rn;
      for (int q = 0; q < ns; ++q) {
        if (!(cut2 < kAccept * lam_top[q])) continue;
        const double lim = std::max(kTail * lam_top[q], kMargin * cut2);
        if (round > 0) {
          // the tail basis prepared by the previous round
          if (nsub[q] > 0) {
            rs.push_back(q);
            rn.push_back(nsub[q]);
          }
          continue;
        }
        int r = 0;
        while (r < gn[q] && evals[q][r] < lim) ++r;  // ascending
        if (r < 1 || evals[q][r - 1] < 1e-4 * cut2) continue;  // no value near the cut
        rs.push_back(q);
        rn.push_back(r);
      }
      if (rs.empty()) break;
      c->count("svd_refine_sectors", (double)rs.size());
      c->count("svd_refine_rounds");
      const int nr = (int)rs.size();
      if (round == 0) {
        // explicit tail bases from the tridiagonal solution
        std::vector<std::vector<int32_t>> cols(ns);
        int64_t tot = 0;
        std::fill(nsub.begin(), nsub.end(), 0);
        for (int i = 0; i < nr; ++i) {
          const int q = rs[i];
          cols[q].resize(rn[i]);
          for (int j = 0; j < rn[i]; ++j) cols[q][j] = j;
          esoff[q] = tot;
          nsub[q] = rn[i];
          tot += (int64_t)gn[q] * rn[i];
        }
        Esub = c->alloc(std::max<int64_t>(tot, 1) * sizeof(double));
        tv0 = c->mark();
        eigh_vectors_all(c, er, cols, static_cast<double *>(Esub->p), esoff);
        ES0 = Esub;
        es0off = esoff;
        for (int i = 0; i < nr; ++i) r0[rs[i]] = rn[i];
        // The computed tail basis leans into each large direction u_r by
        // theta ~ eps lambda_max / lambda_r, which puts rows theta s_r v_r^T
        // into W_S and lifts true tail values to about eps s_max^2 / s_r
        // (1e-12 s_max for s_r near the tail threshold).  The v_r = M^T u_r / s_r
        // of the large values are accurate (their tail admixture is scaled by
        // s_tail / s_r), so W_S is projected off them below, every round.
        std::vector<std::vector<int32_t>> colsL(ns);
        std::vector<int64_t> eLoff(ns, 0);
        int64_t totE = 0, totF = 0, totI = 0;
        std::vector<double> is2;
        for (int i = 0; i < nr; ++i) {
          const int q = rs[i], kl = gn[q] - rn[i];
          nL[q] = kl;
          if (kl == 0) continue;
          colsL[q].resize(kl);
          for (int j = 0; j < kl; ++j) colsL[q][j] = rn[i] + j;
          eLoff[q] = totE;
          totE += (int64_t)gn[q] * kl;
          fLoff[q] = totF;
          totF += (int64_t)kl * (side == kOrthLeft ? secs[q].n : secs[q].m);
          iLoff[q] = totI;
          totI += kl;
          for (int j = 0; j < kl; ++j) is2.push_back(std::ldexp(1.0 / evals[q][rn[i] + j], 2 * dq[q]));
        }
        if (totE > 0) {
          auto ELb = c->alloc(totE * sizeof(double));
          eigh_vectors_all(c, er, colsL, static_cast<double *>(ELb->p), eLoff);
          EL0 = ELb;
          el0off = eLoff;
          FLb = c->alloc(totF * sizeof(double));
          std::vector<BlkD> fb;
          std::vector<PairD> fp;
          for (int q = 0; q < ns; ++q) {
            if (nL[q] == 0) continue;
            const int32_t kl = nL[q], sm = (int32_t)secs[q].m, sn = (int32_t)secs[q].n, idx = (int32_t)fb.size();
            if (side == kOrthLeft) {
              fb.push_back(BlkD{fLoff[q], kl, sn, idx, idx + 1});  // F_L = E_L^T M
              fp.push_back(PairD{eLoff[q], moff[q], sm, 0});
            } else {
              fb.push_back(BlkD{fLoff[q], sm, kl, idx, idx + 1});  // F_L = M E_L
              fp.push_back(PairD{moff[q], eLoff[q], sn, 0});
            }
          }
          if (side == kOrthLeft) raw_gemm(c, kTrans, kNormal, static_cast<double *>(ELb->p), Md,
                                          static_cast<double *>(FLb->p), fb, fp);
          else raw_gemm(c, kNormal, kNormal, Md, static_cast<double *>(ELb->p), static_cast<double *>(FLb->p), fb, fp);
          ILb = upload(c, is2.data(), is2.size() * sizeof(double));
          // The large values themselves: |u_r^T M|^2 (|M v_r|^2), the Rayleigh
          // quotient of the exact M M^T, whose error is second order in the
          // eigenvector's (~theta^2 times the gap, 1e-14 relative), where the
          // eigenvalue carries the absolute ~eps * top of G.
          std::vector<ScaleD> nd;
          for (int q = 0; q < ns; ++q) {
            if (nL[q] == 0) continue;
            if (side == kOrthLeft) nd.push_back(ScaleD{fLoff[q], iLoff[q], nL[q], (int32_t)secs[q].n, 0, 0});
            else nd.push_back(ScaleD{fLoff[q], iLoff[q], (int32_t)secs[q].m, nL[q], 1, 0});
          }
          auto dnd = upload(c, nd.data(), nd.size() * sizeof(ScaleD));
          auto bn = c->alloc(std::max<int64_t>(totI, 1) * sizeof(double));
          sumsq_batch_kernel<<<dim3(64, (unsigned)nd.size()), 256, 0, c->stream>>>(
              static_cast<const double *>(FLb->p), static_cast<double *>(bn->p), static_cast<const ScaleD *>(dnd->p));
          std::vector<double> lv(totI);
          MPS2_HIP(hipMemcpyAsync(lv.data(), bn->p, totI * sizeof(double), hipMemcpyDeviceToHost, c->stream));
          c->sync();
          for (double x : lv)
            if (!std::isfinite(x)) throw Error(kErrNumeric, "SVD refinement: non-finite large-direction values");
          for (int q = 0; q < ns; ++q)
            for (int j = 0; j < nL[q]; ++j) evals[q][gn[q] - nL[q] + j] = std::ldexp(lv[iLoff[q] + j], 2 * dq[q]);
        }
      }
      if (round == 0) c->phase("svd_refine_vec_ms", tv0, c->mark());
      const double *Es = static_cast<const double *>(Esub->p);
      // W_S and the Gram of the projected tail
      std::vector<int64_t> woff(nr), g2off(nr);
      int64_t wtot = 0, gtot2 = 0;
      for (int i = 0; i < nr; ++i) {
        const int q = rs[i], r = rn[i];
        woff[i] = wtot;
        wtot += (int64_t)r * (side == kOrthLeft ? secs[q].n : secs[q].m);
        g2off[i] = gtot2;
        gtot2 += (int64_t)r * r;
      }
      auto bW = c->alloc(std::max<int64_t>(wtot, 1) * sizeof(double));
      auto bG = c->alloc(std::max<int64_t>(gtot2, 1) * sizeof(double));
      double *W_ = static_cast<double *>(bW->p), *G_ = static_cast<double *>(bG->p);
      std::vector<BlkD> wb, gb2;
      std::vector<PairD> wp, gp2;
      for (int i = 0; i < nr; ++i) {
        const int q = rs[i], r = rn[i];
        const int32_t sm = (int32_t)secs[q].m, sn = (int32_t)secs[q].n;
        if (side == kOrthLeft) {
          wb.push_back(BlkD{woff[i], r, sn, i, i + 1});  // W = ES^T M  (r x n_cols)
          wp.push_back(PairD{esoff[q], moff[q], sm, 0});
          gb2.push_back(BlkD{g2off[i], r, r, i, i + 1});  // G2 = W W^T
          gp2.push_back(PairD{woff[i], woff[i], sn, 0});
        } else {
          wb.push_back(BlkD{woff[i], sm, r, i, i + 1});  // W = M ES  (m_rows x r)
          wp.push_back(PairD{moff[q], esoff[q], sn, 0});
          gb2.push_back(BlkD{g2off[i], r, r, i, i + 1});  // G2 = W^T W
          gp2.push_back(PairD{woff[i], woff[i], sm, 0});
        }
      }
      if (side == kOrthLeft) raw_gemm(c, kTrans, kNormal, Es, Md, W_, wb, wp);
      else raw_gemm(c, kNormal, kNormal, Md, Es, W_, wb, wp);
      if (FLb) {
        // W_S -= (W_S F_L^T) diag(1/s^2) F_L   (left; right: F_L diag(1/s^2) F_L^T W_S)
        const double *FL = static_cast<const double *>(FLb->p);
        std::vector<int64_t> poff(nr, 0);
        int64_t ptot = 0;
        for (int i = 0; i < nr; ++i) {
          poff[i] = ptot;
          ptot += (int64_t)rn[i] * nL[rs[i]];
        }
        auto bP = c->alloc(std::max<int64_t>(ptot, 1) * sizeof(double));
        double *P = static_cast<double *>(bP->p);
        std::vector<BlkD> pb, ub;
        std::vector<PairD> pp, up;
        std::vector<ScaleD> sd;
        for (int i = 0; i < nr; ++i) {
          const int q = rs[i];
          const int32_t r = rn[i], kl = nL[q], sm = (int32_t)secs[q].m, sn = (int32_t)secs[q].n;
          if (kl == 0) continue;
          const int32_t idx = (int32_t)pb.size();
          if (side == kOrthLeft) {
            pb.push_back(BlkD{poff[i], r, kl, idx, idx + 1});  // P = W F_L^T
            pp.push_back(PairD{woff[i], fLoff[q], sn, 0});
            sd.push_back(ScaleD{poff[i], iLoff[q], r, kl, 1, 0});
            ub.push_back(BlkD{woff[i], r, sn, idx, idx + 1});  // W -= P F_L
            up.push_back(PairD{poff[i], fLoff[q], kl, 0});
          } else {
            pb.push_back(BlkD{poff[i], kl, r, idx, idx + 1});  // P = F_L^T W
            pp.push_back(PairD{fLoff[q], woff[i], sm, 0});
            sd.push_back(ScaleD{poff[i], iLoff[q], kl, r, 0, 0});
            ub.push_back(BlkD{woff[i], sm, r, idx, idx + 1});  // W -= F_L P
            up.push_back(PairD{fLoff[q], poff[i], kl, 0});
          }
        }
        if (!pb.empty()) {
          if (side == kOrthLeft) raw_gemm(c, kNormal, kTrans, W_, FL, P, pb, pp);
          else raw_gemm(c, kTrans, kNormal, FL, W_, P, pb, pp);
          auto dsd = upload(c, sd.data(), sd.size() * sizeof(ScaleD));
          scale_batch_kernel<<<dim3(256, (unsigned)sd.size()), 256, 0, c->stream>>>(
              P, static_cast<const double *>(ILb->p), static_cast<const ScaleD *>(dsd->p));
          if (side == kOrthLeft) raw_gemm(c, kNormal, kNormal, P, FL, W_, ub, up, 1);
          else raw_gemm(c, kNormal, kNormal, FL, P, W_, ub, up, 1);
        }
      }
      if (side == kOrthLeft) raw_gemm(c, kNormal, kTrans, W_, W_, G_, gb2, gp2);
      else raw_gemm(c, kTrans, kNormal, W_, W_, G_, gb2, gp2);
      if (!std::isfinite(dev_maxabs(c, G_, gtot2)))
        throw Error(kErrNumeric, "SVD refinement round " + std::to_string(round) + ": non-finite Gram of the projected tail");
      const hipEvent_t te0 = c->mark();
      EighResult er2 = eigh_named(c, G_, g2off, rn, "refinement round " + std::to_string(round));
      c->phase("svd_refine_eigh_ms", te0, c->mark());
      for (int i = 0; i < nr; ++i) {
        const int q = rs[i], r = rn[i];
        for (double &x : er2.evals[i]) x = std::ldexp(x, 2 * dq[q]);
        for (int j = 0; j < r; ++j) evals[q][j] = er2.evals[i][j];
        lam_top[q] = 0.0;
        for (int j = 0; j < r; ++j) lam_top[q] = std::max(lam_top[q], er2.evals[i][j]);
        for (int j = 0; j < r; ++j) esc[q][j] = lam_top[q];
      }
      singular_values();
      int64_t derr_now = 0;
      truncate_rule(svr, trunc_err, dmin, dmax, out.kept, &out.trunc_err, &out.D, &derr_now);
      // The kept vectors stay the tridiagonal solution's (the products S Vt =
      // U^T M, or U S = M V, are exact projections for any basis of the kept
      // span); refined vectors are formed only for a tail that a further
      // round must resolve.
      if (round == 2) break;
      if (settled(derr_now)) {
        c->count("svd_refine_settled");
        break;
      }
      all.clear();
      for (auto &v : svr) all.insert(all.end(), v.begin(), v.end());
      if (out.D <= 0 || out.D >= (int64_t)all.size()) break;
      std::nth_element(all.begin(), all.begin() + (out.D - 1), all.end(), std::greater<double>());
      const double cutn =
          std::max(all[out.D - 1], kResolve * *std::max_element(all.begin(), all.end()));
      const double cut2n = cutn * cutn;
      std::vector<std::vector<int32_t>> cols2(nr);
      std::vector<int> r2(nr, 0);
      bool more = false;
      for (int i = 0; i < nr; ++i) {
        const int q = rs[i];
        const double lim = std::max(kTail * lam_top[q], kMargin * cut2n);
        int r = 0;
        if (cut2n < kAccept * lam_top[q])
          while (r < rn[i] && er2.evals[i][r] < lim) ++r;
        if (r >= 1 && er2.evals[i][r - 1] >= 1e-4 * cut2n) {
          r2[i] = r;
          cols2[i].resize(r);
          for (int j = 0; j < r; ++j) cols2[i][j] = j;
          more = true;
        }
      }
      if (!more) break;
      // next tail bases: ES2 = ES B2[:, 0:r2]
      std::vector<int64_t> b2off(nr), e2off(ns, 0);
      int64_t btot = 0, etot2 = 0;
      for (int i = 0; i < nr; ++i) {
        b2off[i] = btot;
        btot += (int64_t)rn[i] * r2[i];
        e2off[rs[i]] = etot2;
        etot2 += (int64_t)gn[rs[i]] * r2[i];
      }
      auto bB = c->alloc(std::max<int64_t>(btot, 1) * sizeof(double));
      eigh_vectors_all(c, er2, cols2, static_cast<double *>(bB->p), b2off);
      auto En = c->alloc(std::max<int64_t>(etot2, 1) * sizeof(double));
      std::vector<BlkD> eb;
      std::vector<PairD> ep;
      std::vector<int> nsub2(ns, 0);
      for (int i = 0; i < nr; ++i) {
        if (r2[i] == 0) continue;
        const int q = rs[i], idx = (int)eb.size();
        eb.push_back(BlkD{e2off[q], gn[q], r2[i], idx, idx + 1});
        ep.push_back(PairD{esoff[q], b2off[i], rn[i], 0});
        nsub2[q] = r2[i];
      }
      raw_gemm(c, kNormal, kNormal, Es, static_cast<double *>(bB->p), static_cast<double *>(En->p), eb, ep);
      Esub = En;
      esoff = e2off;
      nsub = nsub2;
    }
    c->count("svd_refine_ms",
             std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr0).count());
  }
  // 6. new bond index
  Leg knew;
  knew.dir = kOut;
  std::vector<int> secpos(ns, -1);
  for (int q = 0; q < ns; ++q)
    if (out.kept[q] > 0) {
      secpos[q] = (int)knew.qn.size();
      knew.qn.push_back(secs[q].qn);
      knew.dim.push_back(out.kept[q]);
    }
  auto um = std::make_shared<Meta>(), sm = std::make_shared<Meta>(), vm = std::make_shared<Meta>();
  um->qw = sm->qw = vm->qw = m.qw;
  for (int i = 0; i < ldims; ++i) um->legs.push_back(m.legs[i]);
  um->legs.push_back(knew);
  sm->legs = {knew.inverse(), knew};
  vm->legs.push_back(knew.inverse());
  for (int i = ldims; i < r; ++i) vm->legs.push_back(m.legs[i]);
  for (int q = 0; q < ns; ++q) {
    if (secpos[q] < 0) continue;
    for (auto &kv : secs[q].rows) {
      Coord cc = kv.first;
      cc.push_back(secpos[q]);
      um->coords.push_back(cc);
    }
    for (auto &kv : secs[q].cols) {
      Coord cc{secpos[q]};
      cc.insert(cc.end(), kv.first.begin(), kv.first.end());
      vm->coords.push_back(cc);
    }
    sm->coords.push_back(Coord{secpos[q], secpos[q]});
  }
  um->finalize();
  sm->finalize();
  vm->finalize();
  out.u = ten_zeros(c, um);
  out.vt = ten_zeros(c, vm);
  // S on the host
  {
    std::vector<double> sh(sm->total, 0.0);
    for (int q = 0; q < ns; ++q) {
      if (secpos[q] < 0) continue;
      int32_t b = sm->find(Coord{secpos[q], secpos[q]});
      int64_t kq = out.kept[q];
      for (int64_t k = 0; k < kq; ++k) sh[sm->offset[b] + k * kq + k] = sv[q][k] * post;
    }
    out.s = ten_upload(c, sm, sh.data());
  }
  // 7. factors, batched over the kept sectors: E (orthonormal side, n x kq
  //    row-major) and F (the other side, projected and divided by s)
  std::vector<std::vector<int32_t>> kcols(ns);
  std::vector<int64_t> eoff(ns, 0), foff(ns, 0), ioff(ns, 0);
  int64_t etot = 0, ftot = 0, itot = 0;
  std::vector<double> inv;
  for (int q = 0; q < ns; ++q) {
    if (secpos[q] < 0) continue;
    const int64_t kq = out.kept[q];
    kcols[q].assign(order[q].begin(), order[q].begin() + kq);
    eoff[q] = etot;
    etot += gn[q] * kq;
    foff[q] = ftot;
    ftot += kq * (side == kOrthLeft ? secs[q].n : secs[q].m);
    ioff[q] = itot;
    itot += kq;
    for (int64_t k = 0; k < kq; ++k) inv.push_back(sv[q][k] > 0 ? std::ldexp(1.0 / sv[q][k], dq[q]) : 0.0);
  }
  auto Eb = c->alloc(std::max<int64_t>(etot, 1) * sizeof(double));
  auto Fb = c->alloc(std::max<int64_t>(ftot, 1) * sizeof(double));
  double *Ed = static_cast<double *>(Eb->p), *Fd = static_cast<double *>(Fb->p);
  std::vector<BufP> keep;
  if (jac) {
    for (int q = 0; q < ns; ++q) {
      if (secpos[q] < 0) continue;
      const int64_t kq = out.kept[q];
      keep.push_back(upload(c, kcols[q].data(), kq * sizeof(int32_t)));
      gather_cols_kernel<<<gridn(gn[q] * kq), 256, 0, c->stream>>>(
          static_cast<double *>(eo.V->p), eo.secs[q].g_off, eo.secs[q].np, gn[q],
          static_cast<int32_t *>(keep.back()->p), (int)kq, Ed + eoff[q]);
    }
  } else {
    // refined sectors: copy from the level-1 tail / large-direction vectors;
    // the others are back-transformed
    std::vector<std::vector<int32_t>> bcols(kcols);
    std::vector<KeepD> kd;
    std::vector<int32_t> kmap;
    for (int q = 0; q < ns; ++q) {
      if (secpos[q] < 0 || r0[q] < 0) continue;
      const int32_t kq = (int32_t)out.kept[q];
      kd.push_back(KeepD{eoff[q], es0off[q], el0off[q], (int64_t)kmap.size(), gn[q], kq, r0[q], nL[q]});
      for (int32_t idx : kcols[q]) kmap.push_back(idx < r0[q] ? idx : -1 - (idx - r0[q]));
      bcols[q].clear();
    }
    eigh_vectors_all(c, er, bcols, Ed, eoff);
    if (!kd.empty()) {
      keep.push_back(upload(c, kmap.data(), kmap.size() * sizeof(int32_t)));
      const int32_t *dmap = static_cast<const int32_t *>(keep.back()->p);
      keep.push_back(upload(c, kd.data(), kd.size() * sizeof(KeepD)));
      gather_kept_kernel<<<dim3(256, (unsigned)kd.size()), 256, 0, c->stream>>>(
          ES0 ? static_cast<const double *>(ES0->p) : nullptr, EL0 ? static_cast<const double *>(EL0->p) : nullptr,
          dmap, static_cast<const KeepD *>(keep.back()->p), Ed);
    }
  }
  {
    std::vector<BlkD> fb;
    std::vector<PairD> fp;
    std::vector<ScaleD> sd;
    for (int q = 0; q < ns; ++q) {
      if (secpos[q] < 0) continue;
      const int32_t kq = (int32_t)out.kept[q], idx = (int32_t)fb.size();
      if (side == kOrthLeft) {
        // F = E^T M (kq x n), rows scaled by 1/s
        fb.push_back(BlkD{foff[q], kq, (int32_t)secs[q].n, idx, idx + 1});
        fp.push_back(PairD{eoff[q], moff[q], (int32_t)secs[q].m, 0});
        sd.push_back(ScaleD{foff[q], ioff[q], kq, (int32_t)secs[q].n, 0, 0});
      } else {
        // F = M E (m x kq), columns scaled by 1/s
        fb.push_back(BlkD{foff[q], (int32_t)secs[q].m, kq, idx, idx + 1});
        fp.push_back(PairD{moff[q], eoff[q], (int32_t)secs[q].n, 0});
        sd.push_back(ScaleD{foff[q], ioff[q], (int32_t)secs[q].m, kq, 1, 0});
      }
    }
    if (!fb.empty()) {
      if (side == kOrthLeft) raw_gemm(c, kTrans, kNormal, Ed, Md, Fd, fb, fp);
      else raw_gemm(c, kNormal, kNormal, Md, Ed, Fd, fb, fp);
      keep.push_back(upload(c, inv.data(), inv.size() * sizeof(double)));
      const double *dinv = static_cast<const double *>(keep.back()->p);
      keep.push_back(upload(c, sd.data(), sd.size() * sizeof(ScaleD)));
      scale_batch_kernel<<<dim3(256, (unsigned)sd.size()), 256, 0, c->stream>>>(
          Fd, dinv, static_cast<const ScaleD *>(keep.back()->p));
    }
  }
  // scatter into the U and Vt blocks
  std::vector<Copy2D> cpu, cpv, cpvt;
  for (int q = 0; q < ns; ++q) {
    if (secpos[q] < 0) continue;
    auto &s = secs[q];
    const int64_t kq = out.kept[q];
    for (auto &kv : s.rows) {
      Coord cc = kv.first;
      cc.push_back(secpos[q]);
      int32_t b = um->find(cc);
      int64_t bm = um->bsize[b] / kq;
      const int64_t src = (side == kOrthLeft ? eoff[q] : foff[q]) + kv.second * kq;
      cpu.push_back(Copy2D{src, um->offset[b], (int32_t)bm, (int32_t)kq, (int32_t)kq, (int32_t)kq, 1.0});
    }
    for (auto &kv : s.cols) {
      Coord cc{secpos[q]};
      cc.insert(cc.end(), kv.first.begin(), kv.first.end());
      int32_t b = vm->find(cc);
      int64_t bn = vm->bsize[b] / kq;
      if (side == kOrthLeft)
        cpv.push_back(Copy2D{foff[q] + kv.second, vm->offset[b], (int32_t)kq, (int32_t)bn, (int32_t)s.n, (int32_t)bn,
                             1.0});
      else  // Vt block (kq x bn) = E[rows kv.second .. +bn, :]^T
        cpvt.push_back(Copy2D{eoff[q] + kv.second * kq, vm->offset[b], (int32_t)bn, (int32_t)kq, (int32_t)kq,
                              (int32_t)bn, 1.0});
    }
  }
  if (!cpu.empty()) {
    keep.push_back(upload(c, cpu.data(), cpu.size() * sizeof(Copy2D)));
    k_copy2d(c, side == kOrthLeft ? Ed : Fd, out.u->d(), static_cast<const Copy2D *>(keep.back()->p), (int)cpu.size());
  }
  if (!cpv.empty()) {
    keep.push_back(upload(c, cpv.data(), cpv.size() * sizeof(Copy2D)));
    k_copy2d(c, Fd, out.vt->d(), static_cast<const Copy2D *>(keep.back()->p), (int)cpv.size());
  }
  if (!cpvt.empty()) {
    keep.push_back(upload(c, cpvt.data(), cpvt.size() * sizeof(Copy2D)));
    copy2d_t_kernel<<<dim3((unsigned)cpvt.size(), 64), 256, 0, c->stream>>>(
        Ed, out.vt->d(), static_cast<const Copy2D *>(keep.back()->p));
  }
  c->sync();
  return out;
}

}  // namespace mps2
