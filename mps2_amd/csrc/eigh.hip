// Batched symmetric eigen-decomposition of the sector Gram matrices of the
// truncated SVD, on the device, FP64:
//
//  1. Householder tridiagonalisation, lower form (A = Q T Q^T,
//     Q = H_0 H_1 ... H_{n-2}, H_c = I - tau_c v_c v_c^T, v_c[c+1] = 1), one
//     column per step for every sector at once: one workgroup per sector
//     finishes w_{c-1} = tau y - tau^2/2 (y^T v) v, forms row c of the updated
//     matrix and builds the reflector v_c; then many workgroups apply the
//     rank-2 update of column c-1 to the trailing square and form
//     y = A_trail v_c in the same pass.
//  2. Tridiagonal divide and conquer (Cuppen), bottom-up from 1x1 blocks,
//     every merge of one level in the same launches.  A merge of adjacent
//     blocks diag(T1', T2') + rho u u^T (u = e_last + sign(beta) e_first,
//     rho = |beta|) works on D + rho z z^T with D the sub-block eigenvalues:
//     deflation of negligible z_i and of close pairs of D (Givens rotation),
//     secular-equation roots by Newton steps safeguarded by bisection in
//     coordinates shifted to the nearer pole, Gu-Eisenstat recomputation of
//     z from the roots (orthogonal vectors), and the product of the
//     sub-block eigenvectors with the merge's vectors on MFMA.
//  3. Back-transformation of selected eigenvectors: E <- Q E panel by panel
//     in reverse, Q_p = I - V_p T_p V_p^T (compact WY over kNB = 128
//     reflectors), two MFMA GEMMs per panel.
//
// Layouts (per sector of size n; offsets in the sector descriptors):
//   A      n x n row-major (the Gram matrix, overwritten),
//   VH, WT n x n column-major (column c = v_c; WT = V_p T_p per panel),
//   Z      n x n column-major eigenvectors of T (block-diagonal during the
//          merges), lam ascending eigenvalues,
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "eigh.h"

namespace mps2 {
namespace {

constexpr int kNB = 128;  // compact-WY panel width of the back-transformation
constexpr int kT = 256;   // threads of the per-sector kernels
constexpr int kRowsPerWave = 1;
constexpr int kSymvRows = 4 * kRowsPerWave;  // 4 wavefronts per workgroup
constexpr double kEps = 2.220446049250313e-16;

// Order key of eigenvalue estimates in the merges: NaN sorts as +inf, so the
// merge positions (dc_setup) and ranks (dc_rank) form permutations for every
// input and the gathers indexed through them stay inside their blocks even
// when the data is not finite (the host then rejects the eigenvalues).
__device__ __forceinline__ double okey(double v) { return v == v ? v : INFINITY; }

struct ESec {
  int64_t nn;   // offset in n*n buffers
  int64_t nv;   // offset in n-vectors
  int64_t np;   // unused
  int64_t tb;   // offset in T blocks (npanels * kNB * kNB)
  int n;
  int pad;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum over the workgroup (up to 1024 threads), fixed order; red holds
// blockDim.x / 64 doubles.
__device__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;
}

__device__ double block_max(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s = fmax(s, red[i]);
  return s;
}

// Column step c of one sector, run by one workgroup: finish w_{c-1} from
// y = A' v_{c-1} (w = tau y - tau^2/2 (y.v) v), form row c of the matrix
// updated by (v_{c-1}, w_{c-1}), take d[c], and build the reflector v_c of
// row c (alpha = a[c+1], x = a[c+2..]) into VH column c, e[c], tau[c].
__device__ void row_step(const ESec &S, int c, const double *__restrict__ A, double *VH, double *wbuf,
                         const double *ybuf, double *abuf, double *dv, double *ev, double *tauv, double *red,
                         double *sh) {
  const int n = S.n, tid = threadIdx.x, nt = blockDim.x;
  const double *arow = A + S.nn + (int64_t)c * n;
  constexpr int P = 8;
  if (n - c <= P * nt) {
    // Rows r = c + tid + k nt (k < P) stay in this thread's registers: their
    // loads are issued before the first reduction, and the reflector is
    // written from the same registers (no round trip through memory).
    double ar[P], yr[P], vr[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int r = c + tid + k * nt;
      ar[k] = r < n ? arow[r] : 0.0;
    }
    double ss = 0;
    if (c > 0) {
      const double tau = tauv[S.nv + c - 1];
      const double *vp = VH + S.nn + (int64_t)(c - 1) * n;
      const double *y = ybuf + S.nv;
      double *w = wbuf + S.nv;
      const double vc = vp[c], yc = y[c];
      double part = 0;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int r = c + tid + k * nt;
        yr[k] = r < n ? y[r] : 0.0;
        vr[k] = r < n ? vp[r] : 0.0;
        part += yr[k] * vr[k];
      }
      const double kcoef = -0.5 * tau * tau * block_sum(part, red);
      const double wc = tau * yc + kcoef * vc;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int r = c + tid + k * nt;
        if (r < n) {
          const double x = tau * yr[k] + kcoef * vr[k];
          w[r] = x;
          ar[k] = ar[k] - vc * x - wc * vr[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int r = c + tid + k * nt;
      if (r < n && r >= c + 2) ss += ar[k] * ar[k];
    }
    if (tid == 0) dv[S.nv + c] = ar[0];  // row c is thread 0's first
    if (c >= n - 1) return;
    if (tid == 1) sh[0] = ar[0];  // a[c+1], read by all after the barrier below
    const double xn = sqrt(block_sum(ss, red));
    const double alpha = sh[0];
    double beta, tau, sc;
    if (xn == 0.0) {
      beta = alpha;
      tau = 0.0;
      sc = 0.0;
    } else {
      beta = -copysign(hypot(alpha, xn), alpha);
      tau = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    double *vh = VH + S.nn + (int64_t)c * n;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int r = c + tid + k * nt;
      if (r > c && r < n) vh[r] = r == c + 1 ? 1.0 : ar[k] * sc;
    }
    if (tid == 0) {
      ev[S.nv + c] = beta;
      tauv[S.nv + c] = tau;
    }
    return;
  }
  double *a = abuf + S.nv;
  // One pass forms w_{c-1}, the corrected row c and the squared norm of its
  // tail; w_{c-1}[c] is recomputed by every thread instead of broadcast.
  double ss = 0;
  if (c > 0) {
    const double tau = tauv[S.nv + c - 1];
    const double *vp = VH + S.nn + (int64_t)(c - 1) * n;
    const double *y = ybuf + S.nv;
    double *w = wbuf + S.nv;
    double part = 0;
    for (int r = c + tid; r < n; r += nt) part += y[r] * vp[r];
    const double kcoef = -0.5 * tau * tau * block_sum(part, red);
    const double vc = vp[c], wc = tau * y[c] + kcoef * vc;
    for (int r = c + tid; r < n; r += nt) {
      const double x = tau * y[r] + kcoef * vp[r];
      w[r] = x;
      const double av = arow[r] - vc * x - wc * vp[r];
      a[r] = av;
      if (r >= c + 2) ss += av * av;
    }
  } else {
    for (int r = tid; r < n; r += nt) {
      const double av = arow[r];
      a[r] = av;
      if (r >= 2) ss += av * av;
    }
  }
  if (tid == 0) dv[S.nv + c] = a[c];  // a[c] was written by this thread
  if (c >= n - 1) return;
  const double xn = sqrt(block_sum(ss, red));
  const double alpha = a[c + 1];
  double beta, tau, sc;
  if (xn == 0.0) {
    beta = alpha;
    tau = 0.0;
    sc = 0.0;
  } else {
    beta = -copysign(hypot(alpha, xn), alpha);
    tau = (beta - alpha) / beta;
    sc = 1.0 / (alpha - beta);
  }
  double *vh = VH + S.nn + (int64_t)c * n;
  for (int r = c + 1 + tid; r < n; r += nt) vh[r] = r == c + 1 ? 1.0 : a[r] * sc;
  if (tid == 0) {
    ev[S.nv + c] = beta;
    tauv[S.nv + c] = tau;
  }
}

// Column step c of every sector (one workgroup of kTR threads per sector).
constexpr int kTR = 512;
__global__ __launch_bounds__(kTR) void trd_row_kernel(const ESec *__restrict__ secs, int c,
                                                      const double *__restrict__ A, double *VH, double *wbuf,
                                                      const double *__restrict__ ybuf, double *abuf, double *dv,
                                                      double *ev, double *tauv) {
  __shared__ double red[kTR / 64];
  __shared__ double sh[2];
  const ESec S = secs[blockIdx.x];
  if (c >= S.n) return;
  row_step(S, c, A, VH, wbuf, ybuf, abuf, dv, ev, tauv, red, sh);
}

// Rows r > c of the trailing square, one pass: apply the rank-2 update of
// column c-1, A[r][q] -= v[r] w[q] + w[r] v[q] (q > c), store, and form
// y[r] = sum_{q > c} A[r][q] v_c[q]; kSymvRows rows per workgroup, kRowsPerWave
// per wavefront, lanes along the row.  (A device-scope fence per workgroup,
// to hand the column step to the last workgroup, would write back the whole
// L2 this kernel dirties; the column step is a launch of its own instead.)
__global__ __launch_bounds__(kT) void trd_update_symv_kernel(const ESec *__restrict__ secs, int nsec, int c, double *A,
                                                             const double *__restrict__ VH,
                                                             const double *__restrict__ wbuf, double *ybuf) {
  int b = blockIdx.x;
  int q = 0;
  for (; q < nsec; ++q) {
    const int n = secs[q].n;
    if (c < n - 1) {
      const int nwg = (n - c - 1 + kSymvRows - 1) / kSymvRows;
      if (b < nwg) break;
      b -= nwg;
    }
  }
  if (q >= nsec) return;
  const ESec S = secs[q];
  const int n = S.n, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double *v = VH + S.nn + (int64_t)c * n;
  const double *vp = VH + S.nn + (int64_t)(c > 0 ? c - 1 : 0) * n;
  const double *w = wbuf + S.nv;
  const int r0 = c + 1 + b * kSymvRows + wv * kRowsPerWave;
  double *ar[kRowsPerWave];
  double vr[kRowsPerWave], wr[kRowsPerWave], acc[kRowsPerWave];
  bool ok[kRowsPerWave];
#pragma unroll
  for (int i = 0; i < kRowsPerWave; ++i) {
    const int r = r0 + i;
    ok[i] = r < n;
    ar[i] = A + S.nn + (int64_t)(ok[i] ? r : c + 1) * n;
    vr[i] = (ok[i] && c > 0) ? vp[r] : 0.0;
    wr[i] = (ok[i] && c > 0) ? w[r] : 0.0;
    acc[i] = 0.0;
  }
  // Columns in chunks of four 64-wide strides: the loads of a chunk are all
  // issued before its arithmetic (memory-level parallelism for a wave that
  // is otherwise alone on its SIMD); each lane still sums its columns in
  // increasing order.  Rows that are not ok point at row c + 1 (< n), so the
  // chunk loads stay in bounds; only ok rows are written.
  constexpr int U = 4;
  int col = c + 1 + lane;
  if (c > 0) {
    for (; col + (U - 1) * 64 < n; col += U * 64) {
      double x[U], vq[U], wq[U], t[U][kRowsPerWave];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        x[u] = v[col + 64 * u];
        vq[u] = vp[col + 64 * u];
        wq[u] = w[col + 64 * u];
#pragma unroll
        for (int i = 0; i < kRowsPerWave; ++i) t[u][i] = ar[i][col + 64 * u];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < kRowsPerWave; ++i)
          if (ok[i]) {
            const double tt = t[u][i] - vr[i] * wq[u] - wr[i] * vq[u];
            ar[i][col + 64 * u] = tt;
            acc[i] += tt * x[u];
          }
    }
    for (; col < n; col += 64) {
      const double x = v[col], vq = vp[col], wq = w[col];
#pragma unroll
      for (int i = 0; i < kRowsPerWave; ++i)
        if (ok[i]) {
          const double t = ar[i][col] - vr[i] * wq - wr[i] * vq;
          ar[i][col] = t;
          acc[i] += t * x;
        }
    }
  } else {
    for (; col + (U - 1) * 64 < n; col += U * 64) {
      double x[U], t[U][kRowsPerWave];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        x[u] = v[col + 64 * u];
#pragma unroll
        for (int i = 0; i < kRowsPerWave; ++i) t[u][i] = ar[i][col + 64 * u];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < kRowsPerWave; ++i)
          if (ok[i]) acc[i] += t[u][i] * x[u];
    }
    for (; col < n; col += 64) {
      const double x = v[col];
#pragma unroll
      for (int i = 0; i < kRowsPerWave; ++i)
        if (ok[i]) acc[i] += ar[i][col] * x;
    }
  }
#pragma unroll
  for (int i = 0; i < kRowsPerWave; ++i) {
    const double sum = wave_sum(acc[i]);
    if (lane == 0 && ok[i]) ybuf[S.nv + r0 + i] = sum;
  }
}

// T block of panel p (compact WY, forward columnwise) from the Gram of its
// reflectors G = V_p^T V_p:  T[j][j] = tau_j,  T[0:j, j] = -tau_j T[0:j, 0:j] G[0:j, j].
// One workgroup of kNB threads per (sector, panel); T is built in LDS
// (column-major, T[col * kNB + row]) and then stored.
struct PanelRef {
  int q, p;
  int64_t goff;  // G block (jn x jn row-major) in the panel-Gram buffer
};
__global__ __launch_bounds__(kNB) void trd_tblock_kernel(const ESec *__restrict__ secs, const PanelRef *__restrict__ pr,
                                                         const double *__restrict__ Gp, const double *__restrict__ tauv,
                                                         double *Tb) {
  __shared__ double Ts[kNB * kNB];
  const PanelRef P = pr[blockIdx.x];
  const ESec S = secs[P.q];
  const int n = S.n, lane = threadIdx.x;
  const int k = P.p * kNB, jn = min(kNB, n - k);
  double *T = Tb + S.tb + (int64_t)P.p * kNB * kNB;
  const double *G = Gp + P.goff;
  for (int i = lane; i < kNB * kNB; i += kNB) Ts[i] = 0.0;
  __syncthreads();
  for (int j = 0; j < jn; ++j) {
    const double tj = (k + j < n - 1) ? tauv[S.nv + k + j] : 0.0;
    double s = 0;
    if (lane < j)
      for (int l = lane; l < j; ++l) s += Ts[l * kNB + lane] * G[l * jn + j];
    if (lane < j) Ts[j * kNB + lane] = -tj * s;
    if (lane == j) Ts[j * kNB + j] = tj;
    __syncthreads();
  }
  for (int i = lane; i < kNB * kNB; i += kNB) T[i] = Ts[i];
}

// WT[:, k + j] = sum_{l <= j} VH[:, k + l] T_p[l][j]  (column-major, per panel).
__global__ void trd_wt_kernel(const ESec *__restrict__ secs, int nsec, const double *__restrict__ VH,
                              const double *__restrict__ Tb, double *WT, int64_t total) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int q = 0;
  for (; q < nsec; ++q) {
    const int64_t nn = (int64_t)secs[q].n * secs[q].n;
    if (g < nn) break;
    g -= nn;
  }
  const ESec S = secs[q];
  const int n = S.n;
  const int col = (int)(g / n), r = (int)(g % n);
  const int p = col / kNB, j = col % kNB, k = p * kNB;
  const double *T = Tb + S.tb + (int64_t)p * kNB * kNB;
  double s = 0;
  for (int l = 0; l <= j; ++l) s += VH[S.nn + (int64_t)(k + l) * n + r] * T[j * kNB + l];
  WT[S.nn + (int64_t)col * n + r] = s;
}

// ------------------------------------------------------------ divide & conquer

struct DcMerge {
  int q;          // sector
  int s, n1, n2;  // block start and the two sizes
  int64_t roff;   // k*k region: sector nn offset + s * n
};

struct DcInfo {
  int kk;      // non-deflated count
  int nrot;    // deflation rotations
  int nd;      // deflated count
  int pad;
  double rho;  // rho * |z|^2
};

struct DcBufs {
  double *lam, *Z, *dv, *ev;
  double *Dp, *zp, *dK, *zK, *zh, *tau, *rc, *rs;
  int *perm, *flag, *Kidx, *Didx, *org, *rpj, *ri;
  double *Scr, *P, *U, *C;
  DcInfo *info;
  BlkD *gblk;
  PairD *gpair;
};

// Leaves: lam[i] = d[i] - |e[i-1]| - |e[i]|, Z = I.
__global__ void dc_init_kernel(const ESec *__restrict__ secs, int nsec, DcBufs B, int64_t total) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int q = 0;
  for (; q < nsec; ++q) {
    const int64_t nn = (int64_t)secs[q].n * secs[q].n;
    if (g < nn) break;
    g -= nn;
  }
  const ESec S = secs[q];
  const int n = S.n;
  const int col = (int)(g / n), r = (int)(g % n);
  B.Z[S.nn + g] = col == r ? 1.0 : 0.0;
  if (col == 0) {
    double x = B.dv[S.nv + r];
    if (r > 0) x -= fabs(B.ev[S.nv + r - 1]);
    if (r < n - 1) x -= fabs(B.ev[S.nv + r]);
    B.lam[S.nv + r] = x;
  }
}

// One workgroup per merge: z, D in merged sorted order, deflation (one
// thread, sequential over the sorted order), compaction.  With lds_cap > 0
// (every merge of the launch has k <= lds_cap) D, z and the flags live in
// dynamic LDS during the sequential scan and D, z are written back after it;
// otherwise they are used in place in global memory.
constexpr int kDcLdsBytes = 128 * 1024;
__global__ __launch_bounds__(kT) void dc_setup_kernel(const ESec *__restrict__ secs, const DcMerge *__restrict__ ms,
                                                      DcBufs B, int lds_cap) {
  __shared__ double red[kT / 64];
  extern __shared__ double dc_lds[];
  const DcMerge M = ms[blockIdx.x];
  const ESec S = secs[M.q];
  const int n = S.n, s = M.s, n1 = M.n1, n2 = M.n2, k = n1 + n2, tid = threadIdx.x;
  const int64_t vo = S.nv + s;
  const bool in_lds = lds_cap > 0;
  double *Dp = in_lds ? dc_lds : B.Dp + vo;
  double *zp = in_lds ? dc_lds + lds_cap : B.zp + vo;
  int *flag = in_lds ? reinterpret_cast<int *>(dc_lds + 2 * (int64_t)lds_cap) : B.flag + vo;
  const double beta = B.ev[S.nv + s + n1 - 1];
  const double sg = beta >= 0 ? 1.0 : -1.0;
  const double *l1 = B.lam + vo, *l2 = B.lam + vo + n1;
  double zz = 0, dm = 0;
  for (int i = tid; i < k; i += kT) {
    double D, z;
    int pos;
    if (i < n1) {
      D = l1[i];
      z = B.Z[S.nn + (int64_t)(s + i) * n + (s + n1 - 1)];
      int lo = 0, hi = n2;  // count of l2 < D
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (okey(l2[mid]) < okey(D)) lo = mid + 1;
        else hi = mid;
      }
      pos = i + lo;
    } else {
      const int t = i - n1;
      D = l2[t];
      z = sg * B.Z[S.nn + (int64_t)(s + i) * n + (s + n1)];
      int lo = 0, hi = n1;  // count of l1 <= D
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (okey(l1[mid]) <= okey(D)) lo = mid + 1;
        else hi = mid;
      }
      pos = t + lo;
    }
    Dp[pos] = D;
    zp[pos] = z;
    B.perm[vo + pos] = i;
    zz += z * z;
    dm = fmax(dm, fabs(D));
  }
  const double zn2 = block_sum(zz, red);
  const double dmax = block_max(dm, red);
  const double zn = sqrt(zn2);
  const double rho = fabs(beta) * zn2;
  const double tol = 8.0 * kEps * fmax(dmax, rho);
  __syncthreads();
  if (tid == 0) {
    int pj = -1, nrot = 0;
    for (int i = 0; i < k; ++i) {
      const double zi = zp[i] / zn;
      zp[i] = zi;
      flag[i] = 0;
      if (rho * fabs(zi) <= tol) {
        flag[i] = 1;
        continue;
      }
      if (pj < 0) {
        pj = i;
        continue;
      }
      const double t = hypot(zi, zp[pj]);
      const double cc = zi / t, ss = -zp[pj] / t;
      if (fabs((Dp[i] - Dp[pj]) * cc * ss) <= tol) {
        B.rpj[vo + nrot] = pj;
        B.ri[vo + nrot] = i;
        B.rc[vo + nrot] = cc;
        B.rs[vo + nrot] = ss;
        ++nrot;
        zp[i] = t;
        zp[pj] = 0.0;
        const double dpj = Dp[pj] * cc * cc + Dp[i] * ss * ss;
        Dp[i] = Dp[pj] * ss * ss + Dp[i] * cc * cc;
        Dp[pj] = dpj;
        flag[pj] = 1;
      }
      pj = i;
    }
    int kk = 0, nd = 0;
    for (int i = 0; i < k; ++i) {
      if (flag[i]) {
        B.Didx[vo + nd++] = i;
      } else {
        B.Kidx[vo + kk] = i;
        B.dK[vo + kk] = Dp[i];
        B.zK[vo + kk] = zp[i];
        ++kk;
      }
    }
    DcInfo inf;
    inf.kk = kk;
    inf.nrot = nrot;
    inf.nd = nd;
    inf.pad = 0;
    inf.rho = rho;
    B.info[blockIdx.x] = inf;
    B.gblk[blockIdx.x].M = kk;
    B.gpair[blockIdx.x].K = kk;
  }
  if (in_lds) {
    __syncthreads();
    for (int i = tid; i < k; i += kT) {
      B.Dp[vo + i] = Dp[i];
      B.zp[vo + i] = zp[i];
    }
  }
}

// Per merge row r: gather the sorted columns of diag(Q1, Q2), apply the
// deflation rotations in order, write the non-deflated columns to P.
__global__ __launch_bounds__(256) void dc_rotate_kernel(const ESec *__restrict__ secs, const DcMerge *__restrict__ ms,
                                                        DcBufs B) {
  const DcMerge M = ms[blockIdx.y];
  const ESec S = secs[M.q];
  const int n = S.n, s = M.s, k = M.n1 + M.n2;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= k) return;
  const int64_t vo = S.nv + s;
  const DcInfo inf = B.info[blockIdx.y];
  double *X = B.Scr + M.roff;
  for (int pos = 0; pos < k; ++pos) X[(int64_t)pos * k + r] = B.Z[S.nn + (int64_t)(s + B.perm[vo + pos]) * n + s + r];
  for (int t = 0; t < inf.nrot; ++t) {
    const int pj = B.rpj[vo + t], i = B.ri[vo + t];
    const double cc = B.rc[vo + t], ss = B.rs[vo + t];
    const double x = X[(int64_t)pj * k + r], y = X[(int64_t)i * k + r];
    X[(int64_t)pj * k + r] = cc * x + ss * y;
    X[(int64_t)i * k + r] = -ss * x + cc * y;
  }
  double *P = B.P + M.roff;
  for (int t = 0; t < inf.kk; ++t) P[(int64_t)t * k + r] = X[(int64_t)B.Kidx[vo + t] * k + r];
}

// Secular equation f(lambda) = 1 + rho sum_i z_i^2 / (d_i - lambda): root j
// lies in (d_j, d_{j+1}) (the last in (d_last, d_last + rho |z|^2)).  Solved
// in coordinates shifted to the nearer pole o (lambda = d_o + tau).  Each
// step fits the two-pole model A + b1/(L - x) + b2/(R - x) (L, R the
// interval ends) to the value and slope of the pole sums left and right of
// the interval (the fixed-weight method of Bunch, Nielsen and Sorensen) and
// takes its root; a step leaving the bisection bracket is replaced by
// bisection.
__global__ __launch_bounds__(256) void dc_roots_kernel(const ESec *__restrict__ secs, const DcMerge *__restrict__ ms,
                                                       DcBufs B) {
  // one wavefront per root, lanes split the pole sums
  const DcMerge M = ms[blockIdx.y];
  const ESec S = secs[M.q];
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const DcInfo inf = B.info[blockIdx.y];
  const int kk = inf.kk;
  if (j >= kk) return;
  const int64_t vo = S.nv + M.s;
  const double *dK = B.dK + vo, *zK = B.zK + vo;
  const double rho = inf.rho;
  const bool last = j + 1 >= kk;
  int o;
  double lo, hi;
  if (!last) {
    const double gap = dK[j + 1] - dK[j];
    const double mid = 0.5 * gap;
    double f = 0.0;
    for (int i = lane; i < kk; i += 64) f += rho * zK[i] * zK[i] / ((dK[i] - dK[j]) - mid);
    f = 1.0 + wave_sum(f);
    if (f >= 0) {
      o = j;
      lo = 0.0;
      hi = mid;
    } else {
      o = j + 1;
      lo = mid - gap;
      hi = 0.0;
    }
  } else {
    double zz = 0;
    for (int i = lane; i < kk; i += 64) zz += zK[i] * zK[i];
    zz = wave_sum(zz);
    o = j;
    lo = 0.0;
    hi = rho * zz;
  }
  const double dO = dK[o];
  const double L = dK[j] - dO, R = last ? 0.0 : dK[j + 1] - dO;
  double tau = 0.5 * (lo + hi);
  for (int it = 0; it < 100; ++it) {
    double psi = 0, dpsi = 0, phi = 0, dphi = 0;
    for (int i = lane; i < kk; i += 64) {
      const double inv = 1.0 / ((dK[i] - dO) - tau);
      const double w = rho * zK[i] * zK[i] * inv;
      const bool left = i <= j;
      psi += left ? w : 0.0;
      dpsi += left ? w * inv : 0.0;
      phi += left ? 0.0 : w;
      dphi += left ? 0.0 : w * inv;
    }
    psi = wave_sum(psi);
    dpsi = wave_sum(dpsi);
    phi = wave_sum(phi);
    dphi = wave_sum(dphi);
    const double f = 1.0 + psi + phi;
    if (f == 0.0) break;
    if (f > 0.0) hi = tau;
    else lo = tau;
    const double b1 = dpsi * (L - tau) * (L - tau);
    const double a1 = psi - b1 / (L - tau);
    double tn;
    if (last) {
      const double A = 1.0 + a1;
      tn = A > 0.0 ? L + b1 / A : 0.5 * (lo + hi);
    } else {
      const double b2 = dphi * (R - tau) * (R - tau);
      const double a2 = phi - b2 / (R - tau);
      const double A = 1.0 + a1 + a2, D = R - L;
      // A u^2 + (A D + b1 + b2) u + b1 D = 0 with u = L - x in (-D, 0)
      const double qb = A * D + b1 + b2, qc = b1 * D;
      double u;
      if (A == 0.0) {
        u = -qc / qb;
      } else {
        const double disc = fmax(qb * qb - 4.0 * A * qc, 0.0);
        const double qq = -0.5 * (qb + copysign(sqrt(disc), qb));
        const double u1 = qq / A, u2 = qq != 0.0 ? qc / qq : u1;
        u = (u1 > -D && u1 < 0.0) ? u1 : u2;
      }
      tn = L - u;
    }
    if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
    const double step = fabs(tn - tau);
    tau = tn;
    if (step <= 2.0 * kEps * fabs(tau) || hi - lo <= 2.0 * kEps * fmax(fabs(lo), fabs(hi))) break;
  }
  if (lane == 0) {
    B.org[vo + j] = o;
    B.tau[vo + j] = tau;
  }
}

// Gu-Eisenstat: zhat_i^2 = (lambda_i - d_i) / rho * prod_{j != i} (lambda_j - d_i) / (d_j - d_i).
__global__ __launch_bounds__(256) void dc_zhat_kernel(const ESec *__restrict__ secs, const DcMerge *__restrict__ ms,
                                                      DcBufs B) {
  const DcMerge M = ms[blockIdx.y];
  const ESec S = secs[M.q];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const DcInfo inf = B.info[blockIdx.y];
  const int kk = inf.kk;
  if (i >= kk) return;
  const int64_t vo = S.nv + M.s;
  const double *dK = B.dK + vo;
  const int *org = B.org + vo;
  const double *tau = B.tau + vo;
  const double di = dK[i];
  double w = ((dK[org[i]] - di) + tau[i]) / inf.rho;
  for (int j = 0; j < kk; ++j) {
    if (j == i) continue;
    w *= ((dK[org[j]] - di) + tau[j]) / (dK[j] - di);
  }
  B.zh[vo + i] = copysign(sqrt(fmax(w, 0.0)), B.zK[vo + i]);
}

// Vector j of the merge (one wavefront): u_i = zhat_i / (d_i - lambda_j),
// normalised, written as row j of U (kk x kk row-major).
__global__ __launch_bounds__(256) void dc_vec_kernel(const ESec *__restrict__ secs, const DcMerge *__restrict__ ms,
                                                     DcBufs B) {
  const DcMerge M = ms[blockIdx.y];
  const ESec S = secs[M.q];
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const DcInfo inf = B.info[blockIdx.y];
  const int kk = inf.kk;
  if (j >= kk) return;
  const int64_t vo = S.nv + M.s;
  const double *dK = B.dK + vo, *zh = B.zh + vo;
  const double dO = dK[B.org[vo + j]], tj = B.tau[vo + j];
  double ss = 0;
  for (int i = lane; i < kk; i += 64) {
    const double u = zh[i] / ((dK[i] - dO) - tj);
    ss += u * u;
  }
  const double inv = 1.0 / sqrt(wave_sum(ss));
  double *U = B.U + M.roff + (int64_t)j * kk;
  for (int i = lane; i < kk; i += 64) U[i] = zh[i] / ((dK[i] - dO) - tj) * inv;
}

// Eigenvalues of the merge in ascending order (stable): rank of each value.
__global__ __launch_bounds__(256) void dc_rank_kernel(const ESec *__restrict__ secs, const DcMerge *__restrict__ ms,
                                                      DcBufs B) {
  const DcMerge M = ms[blockIdx.y];
  const ESec S = secs[M.q];
  const int k = M.n1 + M.n2;
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u >= k) return;
  const int64_t vo = S.nv + M.s;
  const DcInfo inf = B.info[blockIdx.y];
  const int kk = inf.kk;
  auto val = [&](int w) {
    return w < kk ? B.dK[vo + B.org[vo + w]] + B.tau[vo + w] : B.Dp[vo + B.Didx[vo + w - kk]];
  };
  const double x = val(u), kx = okey(x);
  int rank = 0;
  for (int w = 0; w < k; ++w) {
    const double ky = okey(val(w));
    rank += (ky < kx) || (ky == kx && w < u);
  }
  // lam of this block was consumed by the setup kernel; write the merged values
  B.zp[vo + rank] = x;    // staged: copied into lam by dc_write_kernel
  B.flag[vo + rank] = u;  // source of output column `rank`
}

// Output column `rank` of the merge (one wavefront): the merged eigenvector
// into Z's block, its eigenvalue into lam.
__global__ __launch_bounds__(256) void dc_write_kernel(const ESec *__restrict__ secs, const DcMerge *__restrict__ ms,
                                                       DcBufs B) {
  const DcMerge M = ms[blockIdx.y];
  const ESec S = secs[M.q];
  const int n = S.n, s = M.s, k = M.n1 + M.n2;
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (col >= k) return;
  const int64_t vo = S.nv + s;
  const DcInfo inf = B.info[blockIdx.y];
  const int u = B.flag[vo + col];
  const double *src = u < inf.kk ? B.C + M.roff + (int64_t)u * k : B.Scr + M.roff + (int64_t)B.Didx[vo + u - inf.kk] * k;
  double *dst = B.Z + S.nn + (int64_t)(s + col) * n + s;
  for (int r = lane; r < k; r += 64) dst[r] = src[r];
  if (lane == 0) B.lam[vo + col] = B.zp[vo + col];
}

// E (n x kq row-major) = Z[:, cols] (Z column-major).
__global__ void gather_vec_kernel(const double *__restrict__ Z, int64_t zoff, int n, const int32_t *__restrict__ cols,
                                  int kq, double *E) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n * kq) return;
  const int i = (int)(g / kq), t = (int)(g % kq);
  E[g] = Z[zoff + (int64_t)cols[t] * n + i];
}

inline int nblk(int64_t n, int t = 256) { return (int)std::max<int64_t>((n + t - 1) / t, 1); }

BufP upload_(Ctx *c, const void *h, size_t bytes) {
  auto b = c->alloc(std::max<size_t>(bytes, 1));
  if (bytes) MPS2_HIP(hipMemcpyAsync(b->p, h, bytes, hipMemcpyHostToDevice, c->stream));
  return b;
}

template <typename T>
T *dp(const BufP &b) {
  return static_cast<T *>(b->p);
}

}  // namespace

struct EighState {
  int nsec = 0;
  std::vector<int> n;
  std::vector<ESec> secs;
  BufP dsecs, VH, WT, Tb, Z;
  int64_t total_nn = 0;
};

EighResult eigh_batched(Ctx *c, double *G, const std::vector<int64_t> &goffs, const std::vector<int> &ns) {
  EighResult res;
  auto st = std::make_shared<EighState>();
  res.st = st;
  const int nsec = (int)ns.size();
  st->nsec = nsec;
  st->n = ns;
  st->secs.resize(nsec);
  int64_t nn = 0, nv = 0, tb = 0;
  int maxn = 0;
  for (int q = 0; q < nsec; ++q) {
    MPS2_REQUIRE(goffs[q] == nn, "eigh: Gram matrices must be packed back to back");
    ESec &S = st->secs[q];
    S.n = ns[q];
    S.nn = nn;
    S.nv = nv;
    S.np = 0;
    S.tb = tb;
    S.pad = 0;
    nn += (int64_t)ns[q] * ns[q];
    nv += ns[q];
    tb += (int64_t)((ns[q] + kNB - 1) / kNB) * kNB * kNB;
    maxn = std::max(maxn, ns[q]);
  }
  st->total_nn = nn;
  res.evals.resize(nsec);
  if (nsec == 0 || maxn == 0) return res;
  st->dsecs = upload_(c, st->secs.data(), nsec * sizeof(ESec));
  const ESec *SD = dp<ESec>(st->dsecs);
  st->VH = c->alloc(nn * sizeof(double));
  st->WT = c->alloc(nn * sizeof(double));
  st->Tb = c->alloc(std::max<int64_t>(tb, 1) * sizeof(double));
  st->Z = c->alloc(nn * sizeof(double));
  auto bd = c->alloc(nv * sizeof(double)), be = c->alloc(nv * sizeof(double)), btau = c->alloc(nv * sizeof(double)),
       by = c->alloc(nv * sizeof(double)), bw = c->alloc(nv * sizeof(double)), ba = c->alloc(nv * sizeof(double));
  double *VH = dp<double>(st->VH), *Tb = dp<double>(st->Tb);
  double *dv = dp<double>(bd), *ev = dp<double>(be), *tauv = dp<double>(btau), *yb = dp<double>(by);
  k_fill(c, VH, nn, 0.0);
  k_fill(c, ev, nv, 0.0);
  k_fill(c, tauv, nv, 0.0);
  // ---- 1. tridiagonalisation, one column per step for every sector
  for (int col = 0; col < maxn; ++col) {
    trd_row_kernel<<<nsec, kTR, 0, c->stream>>>(SD, col, G, VH, dp<double>(bw), yb, dp<double>(ba), dv, ev, tauv);
    int64_t nwg = 0;
    for (int q = 0; q < nsec; ++q)
      if (col < ns[q] - 1) nwg += (ns[q] - col - 1 + kSymvRows - 1) / kSymvRows;
    if (nwg > 0) trd_update_symv_kernel<<<(int)nwg, kT, 0, c->stream>>>(SD, nsec, col, G, VH, dp<double>(bw), yb);
  }
  // WY blocks: Gram of each panel's reflectors (MFMA), then the T recurrence
  {
    std::vector<PanelRef> prs;
    std::vector<BlkD> gb;
    std::vector<PairD> gp;
    std::vector<TileD> gt;
    int64_t goff = 0;
    for (int q = 0; q < nsec; ++q)
      for (int p = 0; p * kNB < ns[q]; ++p) {
        const int k = p * kNB, jn = std::min(kNB, ns[q] - k);
        const int idx = (int)prs.size();
        prs.push_back(PanelRef{q, p, goff});
        gb.push_back(BlkD{goff, jn, jn, idx, idx + 1});
        gp.push_back(PairD{st->secs[q].nn + (int64_t)k * ns[q], st->secs[q].nn + (int64_t)k * ns[q], ns[q], 0});
        for (int m0 = 0; m0 < jn; m0 += 64)
          for (int n0 = 0; n0 < jn; n0 += 64) gt.push_back(TileD{idx, m0, n0, 0});
        goff += (int64_t)jn * jn;
      }
    auto dpr = upload_(c, prs.data(), prs.size() * sizeof(PanelRef));
    auto dgb = upload_(c, gb.data(), gb.size() * sizeof(BlkD));
    auto dgp = upload_(c, gp.data(), gp.size() * sizeof(PairD));
    auto dgt = upload_(c, gt.data(), gt.size() * sizeof(TileD));
    auto bg = c->alloc(std::max<int64_t>(goff, 1) * sizeof(double));
    launch_grouped_gemm<double>(c, 0, 1, VH, VH, dp<double>(bg), dp<TileD>(dgt), (int)gt.size(), dp<BlkD>(dgb),
                                dp<PairD>(dgp), 64, 64);
    trd_tblock_kernel<<<(int)prs.size(), kNB, 0, c->stream>>>(SD, dp<PanelRef>(dpr), dp<double>(bg), tauv, Tb);
    trd_wt_kernel<<<nblk(nn), 256, 0, c->stream>>>(SD, nsec, VH, Tb, dp<double>(st->WT), nn);
    c->sync();
  }
  c->count("eigh_trd_columns", maxn);
  // ---- 2. divide and conquer on (d, e)
  auto bl = c->alloc(nv * sizeof(double));
  DcBufs B{};
  B.lam = dp<double>(bl);
  B.Z = dp<double>(st->Z);
  B.dv = dv;
  B.ev = ev;
  std::vector<BufP> keep;
  auto vd = [&](double *&p) {
    keep.push_back(c->alloc(nv * sizeof(double)));
    p = dp<double>(keep.back());
  };
  auto vi = [&](int *&p) {
    keep.push_back(c->alloc(nv * sizeof(int)));
    p = dp<int>(keep.back());
  };
  auto vn = [&](double *&p) {
    keep.push_back(c->alloc(nn * sizeof(double)));
    p = dp<double>(keep.back());
  };
  vd(B.Dp), vd(B.zp), vd(B.dK), vd(B.zK), vd(B.zh), vd(B.tau), vd(B.rc), vd(B.rs);
  vi(B.perm), vi(B.flag), vi(B.Kidx), vi(B.Didx), vi(B.org), vi(B.rpj), vi(B.ri);
  vn(B.Scr), vn(B.P), vn(B.U), vn(B.C);
  dc_init_kernel<<<nblk(nn), 256, 0, c->stream>>>(SD, nsec, B, nn);
  // level structure: blocks per sector, merged pairwise
  std::vector<std::vector<std::pair<int, int>>> blocks(nsec);
  for (int q = 0; q < nsec; ++q)
    for (int i = 0; i < ns[q]; ++i) blocks[q].push_back({i, 1});
  int levels = 0;
  while (true) {
    std::vector<DcMerge> ms;
    std::vector<std::vector<std::pair<int, int>>> nxt(nsec);
    for (int q = 0; q < nsec; ++q) {
      auto &bq = blocks[q];
      size_t b = 0;
      for (; b + 1 < bq.size(); b += 2) {
        DcMerge m;
        m.q = q;
        m.s = bq[b].first;
        m.n1 = bq[b].second;
        m.n2 = bq[b + 1].second;
        m.roff = st->secs[q].nn + (int64_t)m.s * ns[q];
        ms.push_back(m);
        nxt[q].push_back({m.s, m.n1 + m.n2});
      }
      if (b < bq.size()) nxt[q].push_back(bq[b]);
    }
    if (ms.empty()) break;
    blocks.swap(nxt);
    ++levels;
    const int nm = (int)ms.size();
    int kmax = 0;
    std::vector<BlkD> gb(nm);
    std::vector<PairD> gp(nm);
    std::vector<TileD> tiles;
    for (int i = 0; i < nm; ++i) {
      const int k = ms[i].n1 + ms[i].n2;
      kmax = std::max(kmax, k);
      gb[i] = BlkD{ms[i].roff, k, k, i, i + 1};  // M set on the device (non-deflated count)
      gp[i] = PairD{ms[i].roff, ms[i].roff, k, 0};
      for (int m0 = 0; m0 < k; m0 += 64)
        for (int n0 = 0; n0 < k; n0 += 64) tiles.push_back(TileD{i, m0, n0, 0});
    }
    auto dms = upload_(c, ms.data(), nm * sizeof(DcMerge));
    auto dgb = upload_(c, gb.data(), nm * sizeof(BlkD));
    auto dgp = upload_(c, gp.data(), nm * sizeof(PairD));
    auto dti = upload_(c, tiles.data(), tiles.size() * sizeof(TileD));
    auto dinfo = c->alloc(nm * sizeof(DcInfo));
    B.info = dp<DcInfo>(dinfo);
    B.gblk = dp<BlkD>(dgb);
    B.gpair = dp<PairD>(dgp);
    const DcMerge *MS = dp<DcMerge>(dms);
    // D, z, flags of the largest merge in LDS when they fit (20 bytes per entry)
    const int lds_cap = (int64_t)kmax * 20 <= kDcLdsBytes ? kmax : 0;
    static const bool lds_ok = hipFuncSetAttribute((const void *)dc_setup_kernel,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   kDcLdsBytes) == hipSuccess;
    const int cap = lds_ok ? lds_cap : 0;
    dc_setup_kernel<<<nm, kT, (size_t)cap * 20, c->stream>>>(SD, MS, B, cap);
    const dim3 g256(nblk(kmax), nm), g4(nblk(kmax, 4), nm);
    dc_rotate_kernel<<<g256, 256, 0, c->stream>>>(SD, MS, B);
    dc_roots_kernel<<<g4, 256, 0, c->stream>>>(SD, MS, B);
    dc_zhat_kernel<<<g256, 256, 0, c->stream>>>(SD, MS, B);
    dc_vec_kernel<<<g4, 256, 0, c->stream>>>(SD, MS, B);
    // C (kk x k row-major) = U (kk x kk) P (kk x k): the merged vectors, column-major k x kk
    launch_grouped_gemm<double>(c, 0, 0, B.U, B.P, B.C, dp<TileD>(dti), (int)tiles.size(), dp<BlkD>(dgb),
                                dp<PairD>(dgp), 64, 64);
    dc_rank_kernel<<<g256, 256, 0, c->stream>>>(SD, MS, B);
    dc_write_kernel<<<g4, 256, 0, c->stream>>>(SD, MS, B);
    c->sync();  // keeps this level's descriptors alive until its kernels ran
  }
  c->count("eigh_dc_levels", levels);
  for (int q = 0; q < nsec; ++q) {
    res.evals[q].resize(ns[q]);
    MPS2_HIP(hipMemcpyAsync(res.evals[q].data(), B.lam + st->secs[q].nv, ns[q] * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  }
  c->sync();
  for (int q = 0; q < nsec; ++q)
    for (double x : res.evals[q])
      if (!std::isfinite(x)) throw Error(kErrNumeric, "eigh: non-finite eigenvalue (input holds NaN or Inf)");
  return res;
}

void eigh_vectors_all(Ctx *c, const EighResult &r, const std::vector<std::vector<int32_t>> &cols, double *E,
                      const std::vector<int64_t> &eoff) {
  const EighState &st = *r.st;
  const int nsec = st.nsec;
  // E_q = Z_q[:, cols_q]
  std::vector<BufP> keep;
  std::vector<int64_t> xoff(nsec, 0);
  int64_t xtot = 0;
  int maxpan = 0;
  for (int q = 0; q < nsec; ++q) {
    const ESec &S = st.secs[q];
    const int n = S.n, kq = (int)cols[q].size();
    if (kq == 0 || n == 0) continue;
    keep.push_back(upload_(c, cols[q].data(), kq * sizeof(int32_t)));
    gather_vec_kernel<<<nblk((int64_t)n * kq), 256, 0, c->stream>>>(dp<double>(st.Z), S.nn, n,
                                                                    dp<int32_t>(keep.back()), kq, E + eoff[q]);
    xoff[q] = xtot;
    xtot += (int64_t)kNB * kq;
    maxpan = std::max(maxpan, (n + kNB - 1) / kNB);
  }
  if (maxpan == 0) {
    c->sync();
    return;
  }
  // E <- Q E, Q = prod_p (I - V_p T_p V_p^T), panels from the last: per panel
  // index p, X_q = VH_p^T E_q and E_q -= WT_p X_q for every sector at once.
  std::vector<BlkD> blks;
  std::vector<PairD> pairs;
  std::vector<TileD> tiles;
  std::vector<int> t1b(maxpan), t1e(maxpan), t2b(maxpan), t2e(maxpan);
  for (int p = maxpan - 1; p >= 0; --p) {
    t1b[p] = (int)tiles.size();
    for (int q = 0; q < nsec; ++q) {
      const ESec &S = st.secs[q];
      const int n = S.n, kq = (int)cols[q].size(), k = p * kNB;
      if (kq == 0 || k >= n) continue;
      const int jn = std::min(kNB, n - k), bi = (int)blks.size();
      blks.push_back(BlkD{xoff[q], jn, kq, (int32_t)pairs.size(), (int32_t)pairs.size() + 1});
      pairs.push_back(PairD{S.nn + (int64_t)k * n, eoff[q], n, 0});
      for (int m0 = 0; m0 < jn; m0 += 64)
        for (int n0 = 0; n0 < kq; n0 += 64) tiles.push_back(TileD{bi, m0, n0, 0});
    }
    t1e[p] = (int)tiles.size();
    t2b[p] = (int)tiles.size();
    for (int q = 0; q < nsec; ++q) {
      const ESec &S = st.secs[q];
      const int n = S.n, kq = (int)cols[q].size(), k = p * kNB;
      if (kq == 0 || k >= n) continue;
      const int jn = std::min(kNB, n - k), bi = (int)blks.size();
      blks.push_back(BlkD{eoff[q], n, kq, (int32_t)pairs.size(), (int32_t)pairs.size() + 1});
      pairs.push_back(PairD{S.nn + (int64_t)k * n, xoff[q], jn, 0});
      for (int m0 = 0; m0 < n; m0 += 64)
        for (int n0 = 0; n0 < kq; n0 += 64) tiles.push_back(TileD{bi, m0, n0, 0});
    }
    t2e[p] = (int)tiles.size();
  }
  auto bx = c->alloc(std::max<int64_t>(xtot, 1) * sizeof(double));
  auto db = upload_(c, blks.data(), blks.size() * sizeof(BlkD));
  auto dpp = upload_(c, pairs.data(), pairs.size() * sizeof(PairD));
  auto dt = upload_(c, tiles.data(), tiles.size() * sizeof(TileD));
  double *X = dp<double>(bx);
  for (int p = maxpan - 1; p >= 0; --p) {
    launch_grouped_gemm<double>(c, 0, 0, dp<double>(st.VH), E, X, dp<TileD>(dt) + t1b[p], t1e[p] - t1b[p],
                                dp<BlkD>(db), dp<PairD>(dpp), 64, 64);
    launch_grouped_gemm<double>(c, 1, 0, dp<double>(st.WT), X, E, dp<TileD>(dt) + t2b[p], t2e[p] - t2b[p],
                                dp<BlkD>(db), dp<PairD>(dpp), 64, 64, 1);
  }
  c->sync();
}

}  // namespace mps2
