// Lowest eigenpair of the Lanczos tridiagonal matrix on the device, the
// counterpart of TridiagGsSolver / LAPACKE_dstev
// (include/gqmps2/algorithm/lanczos_solver_impl.h:259-312).
//
// Eigenvalue: Sturm-sequence multisection (one wavefront evaluates 64 points
// per round; the count of negative pivots of T - x I equals the number of
// eigenvalues below x).  Eigenvector: inverse iteration with a tridiagonal
// Gaussian elimination with partial pivoting, three sweeps from a flat start.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "runtime.h"

namespace mps2 {

// Work space of the allocation-free entry point (mps2_tridiag_lowest_noalloc):
// one array per GPU, held until the process ends, sized for matrices of up
// to kNoAllocMaxN rows (the inverse iteration's 5n + 8 doubles).  Host code
// never uses the name: set-up (tridiag_scratch_init, from the context
// constructor with the device current) takes each device's address once and
// kernels receive that address.  Calls that overlap on one GPU race on it.
constexpr int kNoAllocMaxN = 4096;
static __device__ __align__(256) double scratch[5 * kNoAllocMaxN + 8];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void tridiag_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(kErrArg, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  MPS2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(kErrHip, "hipGetSymbolAddress returned no address for the tridiagonal scratch array");
  g_scratch_ptr[device] = static_cast<double *>(p);
}

namespace {

__device__ int sturm_count(int n, const double *a, const double *b, double x) {
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) ++cnt;
  for (int i = 1; i < n; ++i) {
    if (q == 0.0) q = 1e-300;
    q = (a[i] - x) - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

// One workgroup of 64 lanes.  out[0] = eigenvalue, out[1..n] = eigenvector.
__global__ __launch_bounds__(64) void tridiag_lowest_kernel(int n, const double *a, const double *b, int want_vec,
                                                            double *out, double *work) {
  __shared__ double lo_s, hi_s;
  __shared__ int cnt_s[64];
  const int lane = threadIdx.x;
  if (n == 1) {
    if (lane == 0) {
      out[0] = a[0];
      if (want_vec) out[1] = 1.0;
    }
    return;
  }
  if (lane == 0) {
    double lo = a[0], hi = a[0];
    for (int i = 0; i < n; ++i) {
      double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i + 1 < n ? fabs(b[i]) : 0.0);
      lo = fmin(lo, a[i] - r);
      hi = fmax(hi, a[i] + r);
    }
    double w = fmax(hi - lo, 1e-300);
    lo_s = lo - 1e-12 * w - 1e-300;
    hi_s = hi + 1e-12 * w + 1e-300;
  }
  __syncthreads();
  for (int it = 0; it < 40; ++it) {
    double lo = lo_s, hi = hi_s;
    double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    cnt_s[lane] = sturm_count(n, a, b, x);
    __syncthreads();
    if (lane == 0) {
      // smallest point with at least one eigenvalue below it
      int k = 64;
      for (int i = 0; i < 64; ++i)
        if (cnt_s[i] >= 1) {
          k = i;
          break;
        }
      double nlo = k == 0 ? lo : lo + (hi - lo) * (double)k / 65.0;
      double nhi = k == 64 ? hi : lo + (hi - lo) * (double)(k + 1) / 65.0;
      lo_s = nlo;
      hi_s = nhi;
    }
    __syncthreads();
    if (hi_s - lo_s <= 2e-16 * fmax(fabs(lo_s), fabs(hi_s)) + 1e-300) break;
  }
  if (lane != 0) return;
  double lam = 0.5 * (lo_s + hi_s);
  out[0] = lam;
  if (!want_vec) return;
  // inverse iteration: (T - lam I) x = y, Gaussian elimination with partial pivoting
  double *x = out + 1;
  double *d = work, *du = work + n, *du2 = work + 2 * n, *dl = work + 3 * n;
  int *piv = reinterpret_cast<int *>(work + 4 * n);
  double scale = 0;
  for (int i = 0; i < n; ++i) scale = fmax(scale, fabs(a[i]) + (i + 1 < n ? fabs(b[i]) : 0.0));
  double eps = 1e-15 * fmax(scale, 1e-300);
  for (int i = 0; i < n; ++i) {
    d[i] = a[i] - lam;
    du[i] = i + 1 < n ? b[i] : 0.0;
    dl[i] = i + 1 < n ? b[i] : 0.0;
    du2[i] = 0.0;
  }
  // LU factorisation with row interchanges (dgttrf-like)
  for (int i = 0; i + 1 < n; ++i) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      piv[i] = 0;
      double di = d[i] == 0.0 ? eps : d[i];
      d[i] = di;
      double f = dl[i] / di;
      dl[i] = f;
      d[i + 1] -= f * du[i];
    } else {
      piv[i] = 1;
      double f = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = f;
      double t = du[i];
      du[i] = d[i + 1];
      d[i + 1] = t - f * d[i + 1];
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
    }
  }
  if (d[n - 1] == 0.0) d[n - 1] = eps;
  for (int i = 0; i < n; ++i) x[i] = 1.0;
  for (int sweep = 0; sweep < 3; ++sweep) {
    // forward: apply L^{-1} with the recorded interchanges
    for (int i = 0; i + 1 < n; ++i) {
      if (piv[i] == 0) {
        x[i + 1] -= dl[i] * x[i];
      } else {
        double t = x[i];
        x[i] = x[i + 1];
        x[i + 1] = t - dl[i] * x[i];
      }
    }
    // backward: U x = y
    x[n - 1] /= d[n - 1];
    if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
    double nr = 0;
    for (int i = 0; i < n; ++i) nr += x[i] * x[i];
    nr = sqrt(nr);
    for (int i = 0; i < n; ++i) x[i] /= nr;
  }
}

}  // namespace

int tridiag_lowest_dev(Ctx *c, int n, const double *a, const double *b, double *eval, double *vec) {
  if (n <= 0) return 1;
  if (n == 1) {
    *eval = a[0];
    if (vec) vec[0] = 1.0;
    return 0;
  }
  // pack a, b into pinned scratch, copy to device
  auto in = c->alloc((2 * n) * sizeof(double));
  auto out = c->alloc((n + 1) * sizeof(double));
  auto work = c->alloc((5 * n + 8) * sizeof(double));
  std::vector<double> h(2 * n, 0.0);
  for (int i = 0; i < n; ++i) h[i] = a[i];
  for (int i = 0; i + 1 < n; ++i) h[n + i] = b[i];
  MPS2_HIP(hipMemcpyAsync(in->p, h.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const double *da = static_cast<const double *>(in->p);
  tridiag_lowest_kernel<<<1, 64, 0, c->stream>>>(n, da, da + n, vec ? 1 : 0, static_cast<double *>(out->p),
                                                 static_cast<double *>(work->p));
  std::vector<double> o(n + 1);
  MPS2_HIP(hipMemcpyAsync(o.data(), out->p, (vec ? n + 1 : 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  *eval = o[0];
  if (vec)
    for (int i = 0; i < n; ++i) vec[i] = o[1 + i];
  return 0;
}

void tridiag_lowest_async(Ctx *c, int n, const double *d_a, const double *d_b, int want_vec, double *d_out,
                          double *d_work) {
  if (n < 1) throw Error(kErrArg, "tridiag_lowest_async: n must be >= 1");
  tridiag_lowest_kernel<<<1, 64, 0, c->stream>>>(n, d_a, n > 1 ? d_b : d_a, want_vec ? 1 : 0, d_out, d_work);
}

void tridiag_lowest_noalloc(Ctx *c, int n, const double *d_a, const double *d_b, int want_vec, double *d_out) {
  if (n < 1 || n > kNoAllocMaxN)
    throw Error(kErrArg, "tridiag_lowest_noalloc: n = " + std::to_string(n) + " outside 1.." +
                             std::to_string(kNoAllocMaxN) + " (the size of its static work space)");
  if (!d_a || !d_out || (n > 1 && !d_b)) throw Error(kErrArg, "tridiag_lowest_noalloc: NULL device pointer");
  double *work = g_scratch_ptr[c->device];
  if (!work) throw Error(kErrState, "tridiag_lowest_noalloc: scratch not set up on this device");
  tridiag_lowest_kernel<<<1, 64, 0, c->stream>>>(n, d_a, n > 1 ? d_b : d_a, want_vec ? 1 : 0, d_out, work);
  MPS2_HIP(hipGetLastError());
}

}  // namespace mps2
