// Runtime (context, caching allocator, device tensors) and the HBM-bound
// helper kernels of the engine: fills, axpby, casts, deterministic
// reductions, block permutations and strided block copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cstring>

#include "runtime.h"

namespace mps2 {

// ------------------------------------------------------------------ runtime
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

Buf::~Buf() {
  if (ctx && p) ctx->release(p, bytes);
}

Ctx::Ctx(int dev) : device(dev) {
  int n = 0;
  MPS2_HIP(hipGetDeviceCount(&n));
  if (n <= 0) throw Error(kErrNoDevice, "no HIP device visible");
  if (dev < 0 || dev >= n) throw Error(kErrArg, "device index out of range");
  MPS2_HIP(hipSetDevice(dev));
  tridiag_scratch_init(dev);
  MPS2_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  MPS2_HIP(hipGetDeviceProperties(&prop, dev));
  num_cu = prop.multiProcessorCount;
  MPS2_HIP(hipHostMalloc((void **)&host_scalars, 256 * sizeof(double), hipHostMallocDefault));
  red_partial = alloc(kRedBlocks * 8 * sizeof(double));
  dev_scalars = alloc(256 * sizeof(double));
}

namespace {
constexpr size_t kPinMax = 256u << 20;   // pinned staging cap; larger uploads sync
}

BufP Ctx::upload(const void *h, size_t bytes) {
  auto b = alloc(std::max<size_t>(bytes, 16));
  if (!bytes) return b;
  PinBlk *blk = nullptr;
  for (auto &pb : pins) {
    if (pb.cap < bytes) continue;
    if (pb.pending) {
      const hipError_t e = hipEventQuery(pb.ev);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) MPS2_HIP(e);
      pb.pending = false;
    }
    if (!blk || pb.cap < blk->cap) blk = &pb;
  }
  if (!blk) {
    size_t cap = 64u << 10;
    while (cap < bytes) cap <<= 1;
    if (pinned_bytes + cap > kPinMax) {
      MPS2_HIP(hipMemcpyAsync(b->p, h, bytes, hipMemcpyHostToDevice, stream));
      sync();
      return b;
    }
    PinBlk nb;
    nb.cap = cap;
    MPS2_HIP(hipHostMalloc((void **)&nb.p, cap, hipHostMallocDefault));
    MPS2_HIP(hipEventCreateWithFlags(&nb.ev, hipEventDisableTiming));
    pinned_bytes += cap;
    count("pinned_blocks");
    pins.push_back(nb);
    blk = &pins.back();
  }
  std::memcpy(blk->p, h, bytes);
  MPS2_HIP(hipMemcpyAsync(b->p, blk->p, bytes, hipMemcpyHostToDevice, stream));
  MPS2_HIP(hipEventRecord(blk->ev, stream));
  blk->pending = true;
  return b;
}

Ctx::~Ctx() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &ph : phases_pending)
    for (hipEvent_t e : {ph.a, ph.b})
      if (std::find(ev_free.begin(), ev_free.end(), e) == ev_free.end()) ev_free.push_back(e);
  phases_pending.clear();
  for (hipEvent_t e : ev_free) (void)hipEventDestroy(e);
  ev_free.clear();
  for (auto &pb : pins) {
    if (pb.ev) (void)hipEventDestroy(pb.ev);
    if (pb.p) (void)hipHostFree(pb.p);
  }
  pins.clear();
  red_partial.reset();
  dev_scalars.reset();
  for (auto &kv : free_) (void)hipFree(kv.second);
  free_.clear();
  if (host_scalars) (void)hipHostFree(host_scalars);
  if (stream) (void)hipStreamDestroy(stream);
}

size_t Ctx::round_bytes(size_t b) {
  if (b == 0) b = 1;
  if (b <= (1u << 20)) {
    size_t r = 256;
    while (r < b) r <<= 1;
    return r;
  }
  const size_t g = 2u << 20;
  return (b + g - 1) / g * g;
}

// Stream-ordered reuse: a released block may be handed out again at once,
// since every later use is queued on the same stream.  Best fit among cached
// blocks of at most 1.25x the request (+ one granule), so the varying tensor
// sizes of a sweep reuse memory instead of growing the cache.
BufP Ctx::alloc(size_t bytes) {
  size_t rb = round_bytes(bytes);
  void *p = nullptr;
  auto it = free_.lower_bound(rb);
  if (it != free_.end() && it->first <= rb + rb / 4 + (2u << 20)) {
    p = it->second;
    rb = it->first;
    free_.erase(it);
    cached_bytes -= rb;
  } else {
    count("hip_malloc");
    hipError_t e = hipMalloc(&p, rb);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, rb);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(kErrOom, "hipMalloc of " + std::to_string(rb) + " bytes failed");
      }
    }
  }
  live_bytes += rb;
  peak_bytes = std::max(peak_bytes, live_bytes);
  ++nbufs;
  auto b = std::make_shared<Buf>();
  b->ctx = this;
  b->p = p;
  b->bytes = rb;
  return b;
}

void Ctx::release(void *p, size_t bytes) {
  --nbufs;
  live_bytes -= bytes;
  if (closing) {
    // The owner has destroyed this context: hand the block back to HIP and
    // delete the context with its last buffer (nothing touches it after).
    (void)hipStreamSynchronize(stream);
    (void)hipFree(p);
    if (nbufs == 0) delete this;
    return;
  }
  free_.emplace(bytes, p);
  cached_bytes += bytes;
}

void Ctx::close() {
  if (stream) (void)hipStreamSynchronize(stream);
  plan_cache_clear(this);
  red_partial.reset();
  dev_scalars.reset();
  for (auto &kv : free_) (void)hipFree(kv.second);
  free_.clear();
  cached_bytes = 0;
  closing = true;
  if (nbufs == 0) delete this;
}

void Ctx::trim() {
  count("alloc_trim");
  (void)hipStreamSynchronize(stream);
  for (auto &kv : free_) (void)hipFree(kv.second);
  free_.clear();
  cached_bytes = 0;
}

void Ctx::sync() {
  counters["syncs"] += 1;
  MPS2_HIP(hipStreamSynchronize(stream));
  resolve_phases();
}

hipEvent_t Ctx::mark() {
  hipEvent_t e;
  if (!ev_free.empty()) {
    e = ev_free.back();
    ev_free.pop_back();
  } else {
    MPS2_HIP(hipEventCreate(&e));
  }
  MPS2_HIP(hipEventRecord(e, stream));
  return e;
}

void Ctx::phase(const char *name, hipEvent_t a, hipEvent_t b) { phases_pending.push_back(Phase{name, a, b}); }

void Ctx::resolve_phases() {
  // every event recorded before the stream sync that precedes this call has
  // completed
  for (auto &ph : phases_pending) {
    float ms = 0;
    MPS2_HIP(hipEventElapsedTime(&ms, ph.a, ph.b));
    counters[ph.name] += ms;
    for (hipEvent_t e : {ph.a, ph.b})
      if (std::find(ev_free.begin(), ev_free.end(), e) == ev_free.end()) ev_free.push_back(e);
  }
  phases_pending.clear();
}

void Ctx::flush_phases() {
  if (!phases_pending.empty()) sync();
}

TenP ten_alloc(Ctx *c, MetaP m, Prec p, bool cplx) {
  if (cplx && p != kF64) throw Error(kErrArg, "complex tensors are FP64");
  auto t = std::make_shared<Ten>();
  t->ctx = c;
  t->meta = std::move(m);
  t->prec = p;
  t->cplx = cplx;
  size_t es = p == kF64 ? sizeof(double) : sizeof(float);
  t->buf = c->alloc(std::max<int64_t>(t->nr(), 1) * es);
  return t;
}

TenP ten_zeros(Ctx *c, MetaP m, Prec p, bool cplx) {
  auto t = ten_alloc(c, std::move(m), p, cplx);
  if (p == kF64) k_fill(c, t->d(), std::max<int64_t>(t->nr(), 1), 0.0);
  else k_fillf(c, t->f(), std::max<int64_t>(t->nr(), 1), 0.0f);
  return t;
}

TenP ten_upload(Ctx *c, MetaP m, const double *host, bool cplx) {
  auto t = ten_alloc(c, std::move(m), kF64, cplx);
  if (t->nr() > 0) {
    MPS2_HIP(hipMemcpyAsync(t->d(), host, t->nr() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->sync();
  }
  return t;
}

void ten_download(const Ten &t, double *host) {
  if (t.n() == 0) return;
  if (t.prec != kF64) throw Error(kErrState, "download of a non-FP64 tensor");
  MPS2_HIP(hipMemcpyAsync(host, t.d(), t.nr() * sizeof(double), hipMemcpyDeviceToHost, t.ctx->stream));
  t.ctx->sync();
}

TenP ten_copy(const Ten &t) {
  auto r = ten_alloc(t.ctx, t.meta, t.prec, t.cplx);
  size_t es = t.prec == kF64 ? 8 : 4;
  if (t.n() > 0)
    MPS2_HIP(hipMemcpyAsync(r->buf->p, t.buf->p, t.nr() * es, hipMemcpyDeviceToDevice, t.ctx->stream));
  return r;
}

TenP ten_promote(const TenP &t) {
  if (t->cplx) return t;
  MPS2_REQUIRE(t->prec == kF64, "complex promotion of a non-FP64 tensor");
  auto r = ten_alloc(t->ctx, t->meta, kF64, true);
  k_promote(t->ctx, t->n(), t->d(), r->d());
  return r;
}

TenP ten_dag(const TenP &t) {
  // The flipped layout is shared by every Dag of tensors with t's layout
  // (and Dag of it returns t's layout); the data buffer is shared as well.
  static std::mutex mu;
  MetaP m;
  {
    std::lock_guard<std::mutex> lock(mu);
    m = t->meta->dag.p.lock();
    if (!m) {
      auto nm = std::make_shared<Meta>(*t->meta);
      for (auto &l : nm->legs) l.dir = l.dir == kIn ? kOut : kIn;
      nm->dag.p = t->meta;
      t->meta->dag.p = nm;
      m = nm;
    }
  }
  if (t->cplx) {
    auto r = ten_copy(*t);
    k_conj(t->ctx, t->n(), r->d());
    r->meta = m;
    return r;
  }
  auto r = std::make_shared<Ten>(*t);
  r->meta = m;
  return r;
}

// ------------------------------------------------------------------ elementwise
namespace {

inline int grid_for(int64_t n, int threads, int cap = 65536) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

__global__ void fill_kernel(double *p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fillf_kernel(float *p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void axpby_kernel(int64_t n, double a, const double *x, double b, double *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
__global__ void scale_kernel(int64_t n, double a, double *x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}
__global__ void d2f_kernel(int64_t n, const double *x, float *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
__global__ void f2d_kernel(int64_t n, const float *x, double *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}

__device__ inline double block_sum(double v, double *sh) {
  // 256 threads = 4 wavefronts of 64
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  }
  return r;
}

__global__ __launch_bounds__(kRedThreads) void dot_partial_kernel(int64_t n, const double *x, const double *y,
                                                                  double *partial) {
  __shared__ double sh[8];
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  double r = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(kRedThreads) void reduce_final_kernel(const double *partial, int np, double *out) {
  __shared__ double sh[8];
  double s = 0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += partial[i];
  double r = block_sum(s, sh);
  if (threadIdx.x == 0) *out = r;
}

// Max |x| per block; blockIdx.y splits a block into chunks, combined by a
// 64-bit atomic max on the bit pattern (monotone for non-negative doubles,
// order-independent).  out must start at zero.
__global__ __launch_bounds__(kRedThreads) void block_maxabs_kernel(const double *x, const int64_t *off,
                                                                   const int64_t *size, double *out) {
  __shared__ double sh[8];
  const int b = blockIdx.x;
  const double *p = x + off[b];
  const int64_t n = size[b];
  // NaN-propagating maximum (fmax would drop NaN): a block holding NaN
  // reports NaN, whose bit pattern (sign cleared) orders above +inf.
  auto nmax = [](double a, double v) { return (v != v || v > a) ? v : a; };
  double m = 0;
  for (int64_t i = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.y * blockDim.x)
    m = nmax(m, fabs(p[i]));
  for (int o = 32; o > 0; o >>= 1) m = nmax(m, __shfl_down(m, o, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = nmax(r, sh[i]);
    if (!(r <= 0)) atomicMax(reinterpret_cast<unsigned long long *>(out) + b, (unsigned long long)__double_as_longlong(r));
  }
}

__global__ void lincomb_kernel(int64_t n, int nv, const double *const *vecs, const double *coef, double *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0;
    for (int v = 0; v < nv; ++v) s += coef[v] * vecs[v][i];
    out[i] = s;
  }
}

// Each thread writes one destination element; its block is found by binary
// search over the destination start offsets.
template <typename T>
__global__ void permute_kernel(const T *src, T *dst, const PermDesc *desc, const int64_t *dst_start, int nblk,
                               int64_t total) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = nblk - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (dst_start[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    const PermDesc &d = desc[lo];
    int64_t loc = g - d.dst_off;
    if (loc < 0 || loc >= d.size) continue;
    int64_t so = d.src_off;
    for (int i = d.rank - 1; i >= 0; --i) {
      int64_t di = d.dims[i];
      int64_t q = loc / di;
      so += (loc - q * di) * d.sstride[i];
      loc = q;
    }
    dst[g] = src[so];
  }
}

// blockIdx.x selects the descriptor, blockIdx.y strides over its elements.
__global__ void copy2d_kernel(const double *src, double *dst, const Copy2D *desc) {
  const Copy2D &d = desc[blockIdx.x];
  int64_t n = (int64_t)d.rows * d.cols;
  for (int64_t e = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.y * blockDim.x) {
    int64_t r = e / d.cols, cc = e - r * d.cols;
    dst[d.dst_off + r * d.ldd + cc] = d.scale * src[d.src_off + r * d.lds + cc];
  }
}

// Plain streaming copy (16 bytes per thread per step), for bandwidth
// calibration.
__global__ void calib_copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// y = y + (sign * *a) * x with the coefficient read from device memory.
__global__ void axpy_dev_kernel(int64_t n, double sign, const double *a, const double *x, double *y) {
  const double c = sign * *a;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = c * x[i] + y[i];
}
__global__ void scale_dev_kernel(int64_t n, const double *a, double *x) {
  const double c = *a;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= c;
}

// Partial sums of <vecs[v], w> for v in [v0, v0 + kMultiV): block b covers a
// strided share of the elements; partial[v * nblk + b].
constexpr int kMultiV = 8;
__global__ __launch_bounds__(kRedThreads) void multidot_partial_kernel(int64_t n, int nv, int v0,
                                                                       const double *const *vecs, const double *w,
                                                                       double *partial) {
  __shared__ double sh[8];
  double s[kMultiV];
#pragma unroll
  for (int j = 0; j < kMultiV; ++j) s[j] = 0;
  const int nloc = min(kMultiV, nv - v0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double x = w[i];
#pragma unroll
    for (int j = 0; j < kMultiV; ++j)
      if (j < nloc) s[j] += vecs[v0 + j][i] * x;
  }
  for (int j = 0; j < nloc; ++j) {
    const double r = block_sum(s[j], sh);
    if (threadIdx.x == 0) partial[(int64_t)(v0 + j) * gridDim.x + blockIdx.x] = r;
    __syncthreads();
  }
}
// out[v] = sum of the partials of vector v (one block per vector, fixed order).
__global__ __launch_bounds__(kRedThreads) void multidot_final_kernel(const double *partial, int np, double *out) {
  __shared__ double sh[8];
  double s = 0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += partial[(int64_t)blockIdx.x * np + i];
  const double r = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}
// w = w - sum_v coef[v] * vecs[v]  (coefficients in device memory)
__global__ void multiaxpy_kernel(int64_t n, int nv, const double *const *vecs, const double *coef, double *w) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0;
    for (int v = 0; v < nv; ++v) s += coef[v] * vecs[v][i];
    w[i] -= s;
  }
}

}  // namespace

void k_axpy_dev(Ctx *c, int64_t n, double sign, const double *d_a, const double *x, double *y) {
  if (n > 0) axpy_dev_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, sign, d_a, x, y);
}
void k_scale_dev(Ctx *c, int64_t n, const double *d_a, double *x) {
  if (n > 0) scale_dev_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, d_a, x);
}
void k_dot_to(Ctx *c, int64_t n, const double *x, const double *y, double *d_out) {
  double *part = static_cast<double *>(c->red_partial->p);
  dot_partial_kernel<<<kRedBlocks, kRedThreads, 0, c->stream>>>(n, x, y, part);
  reduce_final_kernel<<<1, kRedThreads, 0, c->stream>>>(part, kRedBlocks, d_out);
}
void k_multidot(Ctx *c, int64_t n, int nv, const double *const *d_vecs, const double *w, double *d_out,
                double *d_partial) {
  if (nv <= 0) return;
  constexpr int nblk = 256;
  for (int v0 = 0; v0 < nv; v0 += kMultiV)
    multidot_partial_kernel<<<nblk, kRedThreads, 0, c->stream>>>(n, nv, v0, d_vecs, w, d_partial);
  multidot_final_kernel<<<nv, kRedThreads, 0, c->stream>>>(d_partial, nblk, d_out);
}
void k_multiaxpy(Ctx *c, int64_t n, int nv, const double *const *d_vecs, const double *d_coef, double *w) {
  if (n > 0 && nv > 0) multiaxpy_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, nv, d_vecs, d_coef, w);
}

void k_calib_copy(Ctx *c, const void *src, void *dst, int64_t bytes) {
  const int64_t n = bytes / 16;
  if (n > 0)
    calib_copy_kernel<<<grid_for(n, 256, 65536), 256, 0, c->stream>>>(static_cast<const double2 *>(src),
                                                                      static_cast<double2 *>(dst), n);
}

void k_fill(Ctx *c, double *p, int64_t n, double v) {
  fill_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(p, n, v);
}
void k_fillf(Ctx *c, float *p, int64_t n, float v) {
  fillf_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(p, n, v);
}
void k_axpby(Ctx *c, int64_t n, double a, const double *x, double b, double *y) {
  if (n > 0) axpby_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, a, x, b, y);
}
void k_scale(Ctx *c, int64_t n, double a, double *x) {
  if (n > 0) scale_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, a, x);
}
void k_cast_d2f(Ctx *c, int64_t n, const double *x, float *y) {
  if (n > 0) d2f_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, x, y);
}
void k_cast_f2d(Ctx *c, int64_t n, const float *x, double *y) {
  if (n > 0) f2d_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, x, y);
}
void k_dot(Ctx *c, int64_t n, const double *x, const double *y, int slot) {
  k_dot_to(c, n, x, y, static_cast<double *>(c->dev_scalars->p) + slot);
}
void k_block_maxabs(Ctx *c, const double *x, const int64_t *d_off, const int64_t *d_size, int nblk, double *d_out) {
  if (nblk <= 0) return;
  fill_kernel<<<grid_for(nblk, 256), 256, 0, c->stream>>>(d_out, nblk, 0.0);
  block_maxabs_kernel<<<dim3(nblk, 64), 256, 0, c->stream>>>(x, d_off, d_size, d_out);
}
void k_lincomb(Ctx *c, int64_t n, int nv, const double *const *d_vecs, const double *d_coef, double *out) {
  if (n > 0) lincomb_kernel<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, nv, d_vecs, d_coef, out);
}
template <typename T>
void k_permute(Ctx *c, const T *src, T *dst, const PermDesc *d_desc, const int64_t *d_dst_start, int nblk,
               int64_t total) {
  if (total > 0 && nblk > 0)
    permute_kernel<T><<<grid_for(total, 256, 16384), 256, 0, c->stream>>>(src, dst, d_desc, d_dst_start, nblk, total);
}
template void k_permute<double>(Ctx *, const double *, double *, const PermDesc *, const int64_t *, int, int64_t);
template void k_permute<Cplx>(Ctx *, const Cplx *, Cplx *, const PermDesc *, const int64_t *, int, int64_t);
template void k_permute<float>(Ctx *, const float *, float *, const PermDesc *, const int64_t *, int, int64_t);

namespace {
// x[2i+1] = -x[2i+1]: complex conjugate in place (interleaved pairs).
__global__ void conj_kernel(double *x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[2 * i + 1] = -x[2 * i + 1];
}
// out[2i] = re[i], out[2i+1] = 0.
__global__ void promote_kernel(const double *re, double *out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    out[2 * i] = re[i];
    out[2 * i + 1] = 0.0;
  }
}
// One descriptor per blockIdx.y; grid-stride over its K x N complex entries.
__global__ void expand_cplx_kernel(const double *src, double *dst, const CplxExpand *desc) {
  const CplxExpand d = desc[blockIdx.y];
  const int64_t kn = (int64_t)d.K * d.N;
  const int64_t ld = 2 * (int64_t)d.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < kn; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / d.N, n = e - k * d.N;
    const int64_t so = d.src_off + (d.trans ? n * d.K + k : e);
    const double re = src[2 * so], im = src[2 * so + 1];
    double *row0 = dst + d.dst_off + 2 * k * ld + 2 * n;
    row0[0] = re;
    row0[1] = im;
    row0[ld] = -im;
    row0[ld + 1] = re;
  }
}
}  // namespace

void k_conj(Ctx *c, int64_t n, double *x) {
  if (n > 0) conj_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(x, n);
}

void k_promote(Ctx *c, int64_t n, const double *re, double *out) {
  if (n > 0) promote_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(re, out, n);
}

void k_expand_cplx(Ctx *c, const double *src, double *dst, const CplxExpand *d_desc, int n, int64_t max_kn) {
  if (n <= 0) return;
  dim3 g(grid_for(max_kn, 256, 1024), n);
  expand_cplx_kernel<<<g, 256, 0, c->stream>>>(src, dst, d_desc);
}

void k_copy2d(Ctx *c, const double *src, double *dst, const Copy2D *d_desc, int n) {
  if (n <= 0) return;
  dim3 g(n, 32);
  copy2d_kernel<<<g, 256, 0, c->stream>>>(src, dst, d_desc);
}

}  // namespace mps2
