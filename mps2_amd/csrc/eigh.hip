// Batched symmetric eigen-decomposition of the sector Gram matrices of the
// truncated SVD, on the device, FP64:
//
//  1. Householder tridiagonalisation, lower form (A = Q T Q^T,
//     Q = H_0 H_1 ... H_{n-2}, H_c = I - tau_c v_c v_c^T, v_c[c+1] = 1), one
//     column per step for every sector at once, with the trailing update
//     deferred over panels of kPB reflectors (see the tridiagonalisation
//     section): per column a row-parallel kernel finishes w_{c-1} and forms
//     column c, a second one builds the reflector v_c and forms y = S v_c
//     from the stored trailing square without writing it; every kPB columns
//     a tiled pass applies the panel's rank-2kPB update.
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
// strides of 64 columns whose loads the product kernel keeps in flight at
// once (16 against 4: 0.3% faster on a 1700-column sector batch)
constexpr int kSymvChunk = 16;
constexpr double kEps = 2.220446049250313e-16;

// Order key of eigenvalue estimates in the merges: NaN sorts as +inf, so the
// merge positions (dc_setup) and ranks (dc_rank) form permutations for every
// input and the gathers indexed through them stay inside their blocks even
// when the data is not finite (the host then rejects the eigenvalues).
__device__ __forceinline__ double okey(double v) { return v == v ? v : INFINITY; }

struct ESec {
  int64_t nn;   // offset in n*n buffers
  int64_t nv;   // offset in n-vectors
  int64_t ps;   // first product-partial workgroup of the tridiagonalisation
  int64_t tb;   // offset in T blocks (npanels * kNB * kNB)
  int n;
  int pe;       // first reflector-partial workgroup of the tridiagonalisation
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

// ------------------------------------------------------ tridiagonalisation
// Blocked (deferred-update) Householder reduction, the scheme of LAPACK's
// dsytrd/dlatrd: the trailing matrix S receives the rank-2 terms of kPB
// reflectors at once, A_true = S - sum_{i pending} (v_i w_i^T + w_i v_i^T), so
// the matrix-vector product of each column reads S without writing it (half
// the HBM traffic of a rank-2 update per column) and the deferred update is
// one tiled read-modify-write pass per kPB columns.
//
// Per column c, every sector of the batch in the same phase, no phase on a
// single workgroup:
//   reflector phase (one thread per row r >= c): sums the product phase's
//     per-workgroup partials of y.v_{c-1} and of the panel projections
//     p_i = w_i.v_{c-1}, s_i = v_i.v_{c-1} (i in the panel of c - 1, before
//     it), finishes w_{c-1} = tau (y - sum_i (v_i p_i + w_i s_i)) + alpha v_{c-1}
//     with alpha = -tau^2/2 (y.v_{c-1} - 2 p.s), forms column c of A_true,
//     a = row c of S - sum_{i pending} (v_i w_i[c] + w_i v_i[c]), sets
//     d[c] = a[c] and writes per-workgroup partial sums of a[r]^2 (r >= c + 2);
//   update phase (first column c = k > 0 of a panel): S -= the previous
//     panel's terms on rows and columns > c;
//   product phase (one wavefront per row r > c): every workgroup forms the
//     same reflector of column c from a[c + 1] and the partials of |a|^2
//     (e[c], tau[c]; v_c = a / (alpha - beta) below row c + 1, formed on the
//     fly from a and written by the wavefront of each row), y = S v_c on its
//     rows (S only read), and per-workgroup partials of the next column's
//     y.v_c, w_i.v_c, v_i.v_c.
// Partials are summed in a fixed order (within a workgroup, then over the
// workgroups in index order), so the results do not depend on scheduling.
// S is kept as a full square (both triangles) on rows and columns > c.
// One launch per phase (two per column, three at a panel start); per-launch
// latency, not the work of a column, bounds the column loop at these sizes
// (about 15 us per column).
constexpr int kPB = 8;          // reflectors per deferred update
constexpr int kNS = 2 * kPB;    // partials per product item: y.v, p_i (kPB - 1), s_i (kPB - 1), one unused
constexpr int kTB = 512;        // threads of every tridiagonalisation workgroup
constexpr int kER = kTB;        // rows per reflector item
constexpr int kPR = kTB / 64;   // rows per product item (one per wavefront)
constexpr int kUT = 64;         // update tile
constexpr int kUR = kUT * 64 / kTB;  // rows per thread of an update tile

struct TrdArgs {
  double *S;        // Gram matrices, n x n row-major per sector (overwritten)
  double *VH;       // reflectors, column-major per sector (column c = v_c)
  double *W0, *W1;  // w_i of even / odd panels, n x kPB column-major per sector (offset nv * kPB)
  double *y;        // result of the last product phase (n per sector)
  double *a;        // column c of A_true (n per sector)
  double *d, *e, *tau;
  double *ps;       // product partials, kNS per item (sector base ESec::ps items)
  double *pe;       // reflector partials of |a|^2, one per item (sector base ESec::pe)
};

__device__ __forceinline__ double *wcol(const TrdArgs &A, const ESec &S, int i) {
  return ((i / kPB) & 1 ? A.W1 : A.W0) + S.nv * kPB + (int64_t)(i % kPB) * S.n;
}

// Reflector item: thread tid of item b of a sector owns row
// r = c + b * kER + tid; the items of the sectors with c < n follow each
// other.
__device__ void reflect_item(const ESec *__restrict__ secs, int nsec, int c, int item, const TrdArgs &A) {
  constexpr int kG = kTB / kNS;  // thread groups summing the product partials
  __shared__ double red[kTB / 64];
  __shared__ double grp[kG][kNS];
  __shared__ double fin[kNS];
  __shared__ double un[2][kPB];  // v_i[c], w_i[c] of the pending slots
  int b = item, q = 0;
  for (; q < nsec; ++q) {
    const int n = secs[q].n;
    if (c < n) {
      const int nb = (n - c + kER - 1) / kER;
      if (b < nb) break;
      b -= nb;
    }
  }
  if (q >= nsec) return;
  __syncthreads();  // the shared arrays may still be read for a previous item
  const ESec S = secs[q];
  const int n = S.n, tid = threadIdx.x;
  const int r = c + b * kER + tid;
  const bool act = r < n;
  const int rc = act ? r : n - 1;
  const int k = c - c % kPB;                             // first column of c's panel
  const int pb = c == k ? (c >= kPB ? c - kPB : c) : k;  // pending reflectors [pb, c)
  const int npend = c - pb;
  const int ncorr = npend > 0 ? npend - 1 : 0;           // [pb, c - 1): panel of c - 1, before it
  const int slast = c - 1 - pb;                          // slot of reflector c - 1 (w_{c-1} is formed here)
  const double *V = A.VH + S.nn;
  // slot i: reflector pb + i if i < npend, else reflector 0 with weight 0
  auto vcol = [&](int i) { return V + (int64_t)(i < npend ? pb + i : 0) * n; };
  auto wcl = [&](int i) { return wcol(A, S, i < npend ? pb + i : 0); };
  if (tid < 2 * kPB) {
    const int i = tid % kPB;
    un[tid / kPB][i] = (i < npend ? 1.0 : 0.0) * (tid < kPB ? vcol(i) : wcl(i))[c];
  }
  // ---- loads: this row of S's row c, y, v_{c-1}, the pending reflectors;
  // the product partials of column c - 1
  const double sr = A.S[S.nn + (int64_t)c * n + rc];
  const double yr = A.y[S.nv + rc];
  const double vr = (c > 0 && act) ? V[(int64_t)(c - 1) * n + rc] : 0.0;
  double vv[kPB], ww[kPB];
#pragma unroll
  for (int i = 0; i < kPB; ++i) {
    vv[i] = vcol(i)[rc];
    ww[i] = wcl(i)[rc];
  }
  double part = 0.0;
  if (c > 0) {
    const int nps = (n - c + kPR - 1) / kPR;  // product items of column c - 1
    const double *P = A.ps + S.ps * kNS;
    for (int g = tid / kNS; g < nps; g += kG) part += P[(int64_t)g * kNS + tid % kNS];
  }
  grp[tid / kNS][tid % kNS] = part;
  __syncthreads();  // grp, un
  if (tid < kNS) {
    double t = 0.0;
    for (int g = 0; g < kG; ++g) t += grp[g][tid];
    fin[tid] = t;
  }
  __syncthreads();
  double wr = 0.0, wc = 0.0;  // w_{c-1}[r], w_{c-1}[c]
  if (c > 0) {
    double pv[kPB - 1], sv[kPB - 1], ps = 0.0;
#pragma unroll
    for (int i = 0; i < kPB - 1; ++i) {
      const double m = i < ncorr ? 1.0 : 0.0;
      pv[i] = m * fin[1 + i];
      sv[i] = m * fin[kPB + i];
      ps = fma(pv[i], sv[i], ps);
    }
    const double tp = A.tau[S.nv + c - 1];
    const double al = -0.5 * tp * tp * (fin[0] - 2.0 * ps);
    double tc = A.y[S.nv + c], t = yr;
#pragma unroll
    for (int i = 0; i < kPB - 1; ++i) {
      tc = fma(-un[0][i], pv[i], tc);
      tc = fma(-un[1][i], sv[i], tc);
      t = fma(-vv[i], pv[i], t);
      t = fma(-ww[i], sv[i], t);
    }
    wc = fma(al, 1.0, tp * tc);  // v_{c-1}[c] = 1
    wr = fma(al, vr, tp * t);
    if (act) wcol(A, S, c - 1)[r] = wr;
  }
  // ---- column c of A_true on this row
  double a = sr;
#pragma unroll
  for (int i = 0; i < kPB; ++i) {
    a = fma(-vv[i], i == slast ? wc : un[1][i], a);
    a = fma(-(i == slast ? wr : ww[i]), un[0][i], a);
  }
  if (act) A.a[S.nv + r] = a;
  if (r == c) A.d[S.nv + c] = a;
  const double am = (act && r >= c + 2) ? a : 0.0;
  const double ss = block_sum(am * am, red);
  if (tid == 0) A.pe[S.pe + b] = ss;
}

// Update item (first column c of a panel):
//   S[r][q] -= sum_{i in [c - kPB, c)} (v_i[r] w_i[q] + w_i[r] v_i[q])   for r, q > c,
// one 64 x 64 tile (the full square), the panel entries of the tile's rows
// and columns staged in LDS; thread t updates column t % 64 of kUR rows.
__device__ void update_item(const ESec *__restrict__ secs, int nsec, int c, int item, const TrdArgs &A) {
  __shared__ double pv[4][kPB][kUT];  // v_i, w_i on the tile's rows; v_i, w_i on its columns
  int b = item, q = 0, nb = 0;
  for (; q < nsec; ++q) {
    const int n = secs[q].n;
    if (c < n - 1) {
      nb = (n - c - 1 + kUT - 1) / kUT;
      if (b < nb * nb) break;
      b -= nb * nb;
    }
  }
  if (q >= nsec) return;
  __syncthreads();  // pv may still be read for a previous item
  const ESec S = secs[q];
  const int n = S.n, t = threadIdx.x;
  const int r0 = c + 1 + (b / nb) * kUT, q0 = c + 1 + (b % nb) * kUT;
  const int i0 = c - kPB;
  for (int idx = t; idx < 4 * kPB * kUT; idx += kTB) {
    const int which = idx / (kPB * kUT), i = (idx / kUT) % kPB, x = idx % kUT;
    const int rr = (which < 2 ? r0 : q0) + x;
    double val = 0.0;
    if (rr < n) val = (which & 1) ? wcol(A, S, i0 + i)[rr] : A.VH[S.nn + (int64_t)(i0 + i) * n + rr];
    pv[which][i][x] = val;
  }
  __syncthreads();
  const int x = t & 63, ry = (t >> 6) * kUR;
  const int qq = q0 + x;
  if (qq >= n) return;
  double vq[kPB], wq[kPB];
#pragma unroll
  for (int i = 0; i < kPB; ++i) {
    vq[i] = pv[2][i][x];
    wq[i] = pv[3][i][x];
  }
  double *Sm = A.S + S.nn;
  // all kUR loads in flight first, then the rows one at a time (the panel
  // entries of a row read from LDS at one address per wavefront: broadcast)
  double sv[kUR];
#pragma unroll
  for (int j = 0; j < kUR; ++j) {
    const int r = r0 + ry + j;
    sv[j] = r < n ? Sm[(int64_t)r * n + qq] : 0.0;
  }
#pragma unroll 2
  for (int j = 0; j < kUR; ++j) {
    double s = sv[j];
#pragma unroll
    for (int i = 0; i < kPB; ++i) {
      s = fma(-pv[0][i][ry + j], wq[i], s);
      s = fma(-pv[1][i][ry + j], vq[i], s);
    }
    sv[j] = s;
  }
#pragma unroll
  for (int j = 0; j < kUR; ++j) {
    const int r = r0 + ry + j;
    if (r < n) Sm[(int64_t)r * n + qq] = sv[j];
  }
}

// Product item: wavefront wv of item b of a sector owns row
// r = c + 1 + b * kPR + wv (lanes along the row); the items of the sectors
// with c < n - 1 follow each other.
__device__ void product_item(const ESec *__restrict__ secs, int nsec, int c, int item, const TrdArgs &A) {
  __shared__ double pt[kPR][kNS];
  int b = item, q = 0;
  for (; q < nsec; ++q) {
    const int n = secs[q].n;
    if (c < n - 1) {
      const int nwg = (n - c - 1 + kPR - 1) / kPR;
      if (b < nwg) break;
      b -= nwg;
    }
  }
  if (q >= nsec) return;
  const ESec S = secs[q];
  const int n = S.n, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = c + 1 + b * kPR + wv;
  const bool act = r < n;
  const double *a = A.a + S.nv;
  double *V = A.VH + S.nn;
  // this lane's value for the next column's partials (lane 1 + i: w_i, lane
  // kPB + i: v_i of slot i of column c + 1's step; slots below c), loaded now
  const int c1 = c + 1, k1 = c1 - c1 % kPB;
  const int pb1 = c1 == k1 ? (c1 >= kPB ? c1 - kPB : c1) : k1;
  const int ncorr1 = c1 - pb1 > 0 ? c1 - pb1 - 1 : 0;
  double sl = 0.0;
  if (act && lane >= 1 && lane < 2 * kPB - 1) {
    const int i = lane < kPB ? lane - 1 : lane - kPB;
    if (i < ncorr1) sl = lane < kPB ? wcol(A, S, pb1 + i)[r] : V[(int64_t)(pb1 + i) * n + r];
  }
  // reflector of column c: alpha = a[c + 1], |x|^2 = the reflector phase's
  // partials summed in item order
  const int ne = (n - c + kER - 1) / kER;
  double x2 = 0.0;
  for (int g = 0; g < ne; ++g) x2 += A.pe[S.pe + g];
  const double alpha = a[c + 1];
  // y[r] = sum_{q > c} S[r][q] v_c[q], v_c[q] = 1 (q = c + 1) or a[q] sc.
  // Columns in chunks of U 64-wide strides, every load of a chunk issued
  // before its arithmetic (the last chunk reads zeros past the row end, which
  // leave the sum unchanged); each lane sums its columns in increasing order.
  // The first chunk is loaded before the reflector's scalars are formed, so
  // its round trip overlaps theirs.
  constexpr int U = kSymvChunk / 2;
  const double *row = A.S + S.nn + (int64_t)r * n;
  double x[U], t[U];
  auto load = [&](int col) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int cc = col + 64 * u;
      const bool in = act && cc < n;
      x[u] = in ? a[cc] : 0.0;
      t[u] = in ? row[cc] : 0.0;
    }
  };
  int col = c + 1 + lane;
  load(col);
  const double xn = sqrt(x2);
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
  if (b == 0 && threadIdx.x == 0) {
    A.e[S.nv + c] = beta;
    A.tau[S.nv + c] = tau;
  }
  double acc = 0.0, vr = 0.0;
  if (act) {
    while (col < n) {
#pragma unroll
      for (int u = 0; u < U; ++u) acc = fma(t[u], col + 64 * u == c + 1 ? 1.0 : x[u] * sc, acc);
      col += U * 64;
      if (col < n) load(col);
    }
    vr = r == c + 1 ? 1.0 : a[r] * sc;
    if (lane == 0) V[(int64_t)c * n + r] = vr;
  }
  const double y = wave_sum(acc);
  if (act && lane == 0) A.y[S.nv + r] = y;
  // partials of column c + 1's step over this item's rows: y.v_c, then
  // w_i.v_c and v_i.v_c of its slots
  __syncthreads();  // pt may still be read for a previous item
  if (lane < kNS) pt[wv][lane] = act ? (lane == 0 ? y : sl) * vr : 0.0;
  __syncthreads();
  if (threadIdx.x < kNS) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kPR; ++w) t += pt[w][threadIdx.x];
    A.ps[(S.ps + b) * kNS + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kTB) void trd_reflect_kernel(const ESec *__restrict__ secs, int nsec, int c, TrdArgs A) {
  reflect_item(secs, nsec, c, blockIdx.x, A);
}
__global__ __launch_bounds__(kTB) void trd_update_kernel(const ESec *__restrict__ secs, int nsec, int c, TrdArgs A) {
  update_item(secs, nsec, c, blockIdx.x, A);
}
__global__ __launch_bounds__(kTB) void trd_product_kernel(const ESec *__restrict__ secs, int nsec, int c, TrdArgs A) {
  product_item(secs, nsec, c, blockIdx.x, A);
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

BufP upload_(Ctx *c, const void *h, size_t bytes) { return c->upload(h, bytes); }

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
  int64_t nn = 0, nv = 0, tb = 0, nps = 0;
  int maxn = 0, npe = 0;
  for (int q = 0; q < nsec; ++q) {
    MPS2_REQUIRE(goffs[q] == nn, "eigh: Gram matrices must be packed back to back");
    ESec &S = st->secs[q];
    S.n = ns[q];
    S.nn = nn;
    S.nv = nv;
    S.ps = nps;
    S.tb = tb;
    S.pe = npe;
    nps += (ns[q] + kPR - 1) / kPR;
    npe += (ns[q] + kER - 1) / kER;
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
       by = c->alloc(nv * sizeof(double)), ba = c->alloc(nv * sizeof(double)),
       bw0 = c->alloc(nv * kPB * sizeof(double)), bw1 = c->alloc(nv * kPB * sizeof(double)),
       bps = c->alloc(std::max<int64_t>(nps, 1) * kNS * sizeof(double)),
       bpe = c->alloc(std::max<int64_t>(npe, 1) * sizeof(double));
  double *VH = dp<double>(st->VH), *Tb = dp<double>(st->Tb);
  double *dv = dp<double>(bd), *ev = dp<double>(be), *tauv = dp<double>(btau);
  k_fill(c, VH, nn, 0.0);
  k_fill(c, ev, nv, 0.0);
  k_fill(c, tauv, nv, 0.0);
  k_fill(c, dp<double>(bw0), nv * kPB, 0.0);  // read with zero weight before written
  k_fill(c, dp<double>(bw1), nv * kPB, 0.0);
  const hipEvent_t t_trd = c->mark();
  {
    double n3 = 0;
    for (int q = 0; q < nsec; ++q) n3 += (double)ns[q] * ns[q] * ns[q];
    c->count("eigh_n3_sum_g", n3 * 1e-9);   // sum of n^3 over sectors, 1e9 units
  }
  // ---- 1. tridiagonalisation, one column per step for every sector
  const TrdArgs TA{G,  VH,   dp<double>(bw0), dp<double>(bw1), dp<double>(by), dp<double>(ba),
                   dv, ev, tauv, dp<double>(bps), dp<double>(bpe)};
  for (int col = 0; col < maxn; ++col) {
    int64_t ner = 0, nwg = 0, nut = 0;
    for (int q = 0; q < nsec; ++q) {
      if (col < ns[q]) ner += (ns[q] - col + kER - 1) / kER;
      if (col < ns[q] - 1) {
        nwg += (ns[q] - col - 1 + kPR - 1) / kPR;
        const int64_t nb = (ns[q] - col - 1 + kUT - 1) / kUT;
        nut += nb * nb;
      }
    }
    if (ner > 0) trd_reflect_kernel<<<(unsigned)ner, kTB, 0, c->stream>>>(SD, nsec, col, TA);
    if (col % kPB == 0 && col >= kPB && nut > 0)
      trd_update_kernel<<<(unsigned)nut, kTB, 0, c->stream>>>(SD, nsec, col, TA);
    if (nwg > 0) trd_product_kernel<<<(unsigned)nwg, kTB, 0, c->stream>>>(SD, nsec, col, TA);
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
  }
  c->count("eigh_trd_columns", maxn);
  const hipEvent_t t_dc = c->mark();
  c->phase("eigh_trd_ms", t_trd, t_dc);
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
    // This level's descriptor buffers go back to the pool here; a later
    // allocation reusing them is written in stream order, after these kernels.
  }
  c->count("eigh_dc_levels", levels);
  c->phase("eigh_dc_ms", t_dc, c->mark());
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
  const hipEvent_t t_bt = c->mark();
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
    c->phase("eigh_back_ms", t_bt, c->mark());
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
  c->phase("eigh_back_ms", t_bt, c->mark());
}

}  // namespace mps2
