// Device runtime of the block-sparse engine: one context per GPU holding one
// HIP stream, a stream-ordered caching allocator and a plan cache.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "meta.h"

namespace mps2 {

enum Status : int {
  kOk = 0,
  kErrArg = 1,
  kErrHip = 2,
  kErrOom = 3,
  kErrNumeric = 4,
  kErrState = 5,
  kErrNoDevice = 6,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define MPS2_HIP(expr)                                                                              \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      throw ::mps2::Error(e_ == hipErrorOutOfMemory ? ::mps2::kErrOom : ::mps2::kErrHip,           \
                          std::string(#expr) + ": " + hipGetErrorString(e_));                       \
  } while (0)

#define MPS2_REQUIRE(cond, msg)                                       \
  do {                                                                \
    if (!(cond)) throw ::mps2::Error(::mps2::kErrArg, (msg));         \
  } while (0)

struct Ctx;

// Device allocation owned through shared_ptr; returned to the context cache
// on release (stream order makes reuse safe: all work runs on ctx->stream).
struct Buf {
  Ctx *ctx = nullptr;
  void *p = nullptr;
  size_t bytes = 0;
  ~Buf();
};
using BufP = std::shared_ptr<Buf>;

enum Prec : int { kF64 = 0, kF32 = 1 };

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  // caching allocator
  std::multimap<size_t, void *> free_;  // cached blocks by size (best fit)
  size_t cached_bytes = 0, live_bytes = 0, peak_bytes = 0;
  // pinned host staging for scalars
  double *host_scalars = nullptr;   // 256 doubles
  BufP red_partial;                 // reduction partials (device)
  BufP dev_scalars;                 // 256 doubles on device
  // statistics
  double gemm_flops = 0;
  int64_t gemm_launches = 0;
  std::string last_kernel;
  // per-launch timing of grouped GEMMs (benchmarking only)
  bool time_gemms = false;
  struct GemmTiming {
    hipEvent_t start, stop;
    double flops;
    double bytes;   // operand and result bytes of the launch (algorithmic)
  };
  std::vector<GemmTiming> gemm_timings;
  // Profiling aid: within each effective-Hamiltonian matvec, the grouped GEMM
  // of stage tag_stage (0-based; -1 = none) runs as a separately named kernel
  // instance, so a profiler's per-kernel statistics isolate its launches.
  int tag_stage = -1;
  int stage_idx = 0;
  // Benchmark of one GPU's share: with shard_g > 1 the tagged stage computes
  // only the output blocks owned by rank shard_r of shard_g (blocks assigned
  // by longest-processing-time on their flops); shard_flops records the
  // share's algorithmic flops of the last such launch.
  int shard_g = 1, shard_r = 0;
  double shard_flops = 0;
  // named event counters (svd calls, Jacobi sweeps, host syncs, ...)
  std::map<std::string, double> counters;
  void count(const char *name, double v = 1.0) { counters[name] += v; }
  // Buffers handed out and not yet released.  A closed context (close())
  // stays alive until the last of them is released: every Buf keeps a raw
  // pointer to its context.
  int64_t nbufs = 0;
  bool closing = false;
  // Grouped-GEMM autotuning (tune.hip): 1 = tune large plans, 0 = heuristic.
  int tune_mode = 1;
  // Arithmetic of FP32 grouped GEMMs: 0 = FP32 matrix cores, 1 = split-BF16x3
  // (operands split into two BF16 parts, three BF16 products per multiply,
  // FP32 accumulation).  Set for the duration of a call by SplitScope.
  int f32_split = 0;

  // Host-to-device uploads of small host arrays (descriptor lists, tile
  // lists, scalars) go through pinned staging blocks, each reused once the
  // copy recorded after its last use has completed: no stream sync per
  // upload, so the device keeps running queued work while the host plans.
  struct PinBlk {
    char *p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  std::vector<PinBlk> pins;
  size_t pinned_bytes = 0;

  // Phase clocks on the device timeline: mark() records a timing event on
  // the stream, phase(name, a, b) adds the time between two marks to
  // counters[name] once both have completed (resolved at the next sync(), so
  // timing a phase costs no sync of its own; flush_phases() forces it).
  hipEvent_t mark();
  void phase(const char *name, hipEvent_t a, hipEvent_t b);
  void flush_phases();
  struct Phase {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Phase> phases_pending;
  std::vector<hipEvent_t> ev_free;
  void resolve_phases();
  // A device buffer holding `bytes` of host data h (copied to staging at once;
  // the caller may free h on return).
  BufP upload(const void *h, size_t bytes);

  explicit Ctx(int dev);
  ~Ctx();
  // Destroy on behalf of the owner: frees the plans and the cache now, and
  // deletes the context at once or, while tensors allocated from it are
  // alive, when the last of them is released.
  void close();
  BufP alloc(size_t bytes);
  void release(void *p, size_t bytes);
  void trim();
  void sync();
  static size_t round_bytes(size_t b);
};

// A device-resident block-sparse tensor.  Data is FP64 unless prec == kF32
// (mixed-precision working copies).  Tensors sharing a Buf (Dag views) never
// write through it.
// Complex tensors (gqten's GQTEN_Complex) hold FP64 (re, im) pairs,
// interleaved: element i of the layout at doubles 2i, 2i + 1.
struct Ten {
  Ctx *ctx = nullptr;
  MetaP meta;
  BufP buf;
  Prec prec = kF64;
  bool cplx = false;
  double *d() const { return static_cast<double *>(buf->p); }
  float *f() const { return static_cast<float *>(buf->p); }
  int64_t n() const { return meta->total; }                  // elements
  int64_t nr() const { return meta->total * (cplx ? 2 : 1); }  // stored reals
};
using TenP = std::shared_ptr<Ten>;

TenP ten_alloc(Ctx *c, MetaP m, Prec p = kF64, bool cplx = false);
TenP ten_zeros(Ctx *c, MetaP m, Prec p = kF64, bool cplx = false);
// host: n (real) or 2 n (complex, interleaved) doubles
TenP ten_upload(Ctx *c, MetaP m, const double *host, bool cplx = false);
void ten_download(const Ten &t, double *host);
TenP ten_copy(const Ten &t);
// Flipped directions; real tensors share the data, complex ones get the
// complex conjugate.
TenP ten_dag(const TenP &t);
// The complex form of a real FP64 tensor (imaginary parts zero); t itself if
// it is complex.
TenP ten_promote(const TenP &t);

// Arithmetic of a matvec, contraction or Lanczos call (the C ABI's precision
// argument): FP64, FP32 on the FP32 matrix cores, or split-BF16x3 on the BF16
// matrix cores.  The last two keep their working copies in FP32.
enum Arith : int { kArithF64 = 0, kArithF32 = 1, kArithBf16x3 = 2 };
inline Prec storage_of(Arith a) { return a == kArithF64 ? kF64 : kF32; }
// The C ABI's precision code (0, 1, 2) as an Arith; anything else is an
// argument error.
Arith arith_of(int precision);
// Sets ctx->f32_split (split-BF16x3 FP32 GEMMs) for the lifetime of the scope.
struct SplitScope {
  Ctx *c;
  int old;
  SplitScope(Ctx *c_, bool on) : c(c_), old(c_->f32_split) { c->f32_split = on ? 1 : 0; }
  ~SplitScope() { c->f32_split = old; }
  SplitScope(const SplitScope &) = delete;
  SplitScope &operator=(const SplitScope &) = delete;
};

// ------------------------------------------------------------------ kernels (kernels.hip)
void k_fill(Ctx *c, double *p, int64_t n, double v);
void k_fillf(Ctx *c, float *p, int64_t n, float v);
void k_axpby(Ctx *c, int64_t n, double a, const double *x, double b, double *y);     // y = a x + b y
void k_scale(Ctx *c, int64_t n, double a, double *x);
void k_cast_d2f(Ctx *c, int64_t n, const double *x, float *y);
void k_cast_f2d(Ctx *c, int64_t n, const float *x, double *y);
// Deterministic reductions; result left in device scalar slot `slot` of ctx->dev_scalars.
void k_dot(Ctx *c, int64_t n, const double *x, const double *y, int slot);
void k_block_maxabs(Ctx *c, const double *x, const int64_t *d_off, const int64_t *d_size, int nblk, double *d_out);
// out = sum_i coef[i] * vec[i] (device pointer table, device coefficients)
void k_lincomb(Ctx *c, int64_t n, int nv, const double *const *d_vecs, const double *d_coef, double *out);
// Device-scalar forms (no host round trip): y += sign * (*d_a) * x; x *= *d_a;
// *d_out = <x, y> (deterministic two-pass reduction).
void k_axpy_dev(Ctx *c, int64_t n, double sign, const double *d_a, const double *x, double *y);
void k_scale_dev(Ctx *c, int64_t n, const double *d_a, double *x);
void k_dot_to(Ctx *c, int64_t n, const double *x, const double *y, double *d_out);
// d_out[v] = <vecs[v], w> for v < nv (d_partial: nv * 256 doubles of work
// space); w -= sum_v coef[v] vecs[v].  Fixed summation order.
void k_multidot(Ctx *c, int64_t n, int nv, const double *const *d_vecs, const double *w, double *d_out,
                double *d_partial);
void k_multiaxpy(Ctx *c, int64_t n, int nv, const double *const *d_vecs, const double *d_coef, double *w);
// Block-wise permutation / copy (descriptors on device).
struct PermDesc {
  int64_t src_off, dst_off, size;
  int32_t rank, pad;
  int64_t dims[kMaxRank];      // dst dims
  int64_t sstride[kMaxRank];   // src stride for each dst axis
};
template <typename T>
void k_permute(Ctx *c, const T *src, T *dst, const PermDesc *d_desc, const int64_t *d_dst_start, int nblk,
               int64_t total);
// Strided 2D copies: dst[dr*ldd + dc] = src[sr*lds + sc] over (rows x cols).
struct Copy2D {
  int64_t src_off, dst_off;
  int32_t rows, cols;
  int32_t lds, ldd;
  double scale;   // multiply by scale (1 for plain copies)
};
void k_copy2d(Ctx *c, const double *src, double *dst, const Copy2D *d_desc, int n);
// Complex element type of the permutation kernel.
struct alignas(16) Cplx {
  double re, im;
};
// Conjugate n interleaved complex values in place; promote n reals to complex.
void k_conj(Ctx *c, int64_t n, double *x);
void k_promote(Ctx *c, int64_t n, const double *re, double *out);
// Real form of a complex GEMM operand B (K x N complex, row-major, or N x K
// with trans): the 2K x 2N real matrix [[Re, Im], [-Im, Re]] per element,
// row-major at dst_off (doubles), so that A (M x K complex read as M x 2K
// reals) times it is C (M x N complex read as M x 2N reals).  src_off in
// complex elements.
struct CplxExpand {
  int64_t src_off, dst_off;
  int32_t K, N;
  int32_t trans, pad;
};
void k_expand_cplx(Ctx *c, const double *src, double *dst, const CplxExpand *d_desc, int n, int64_t max_kn);
// bytes (a multiple of 16) copied by a plain streaming kernel (calibration)
void k_calib_copy(Ctx *c, const void *src, void *dst, int64_t bytes);

// Grouped GEMM over a contraction plan (gemm.hip).
struct TileD {
  int32_t blk;
  int32_t m0, n0;
  int32_t pad;
};
struct BlkD {
  int64_t c_off;
  int32_t M, N;
  int32_t pb, pe;
};
struct PairD {
  int64_t a_off, b_off;
  int32_t K, pad;
};
// The tiles of one grouped launch: an explicit list (list != nullptr), or
// decoded on the device from the blocks in launch order (order[0..nb)) and
// the prefix sums of their tile counts (start[0..nb], start[0] = 0,
// start[nb] = the launch's tile count) for tm x tn tiles, each block's tiles
// row-major.  The decoded form keeps a plan's device footprint to a few
// integers per block.
struct TileMap {
  const TileD *list = nullptr;
  const int32_t *order = nullptr;
  const int32_t *start = nullptr;
  int nb = 0, tm = 0, tn = 0;
};
// Output extent of a launch, for split-K: one past the largest element it
// writes, its output blocks, the largest block's M N.
struct SplitInfo {
  int64_t out_span = 0;
  int nblk = 0;
  int64_t max_mn = 0;
  double avg_k = 0;   // flops / (2 sum M N): the mean K of an output element
};
// variant: 0 = register-staged slabs; 4 = direct-to-LDS slabs with four
// stages (FP32 128x128 tiles, row-major A and B only); 8 / 9 = split-K in
// 2 / 4 parts per tile (FP32 128x128 or FP64 64x64 tiles, plain writes;
// needs si).
template <typename T>
void launch_grouped_gemm(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                         const BlkD *blks, const PairD *pairs, int tile_m, int tile_n, int sub, int tag,
                         int variant, const SplitInfo &si);
template <typename T>
void launch_grouped_gemm(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileD *tiles, int ntiles,
                         const BlkD *blks, const PairD *pairs, int tile_m, int tile_n, int sub = 0, int tag = 0,
                         int variant = 0) {
  TileMap m;
  m.list = tiles;
  launch_grouped_gemm<T>(c, la, lb, A, B, C, m, ntiles, blks, pairs, tile_m, tile_n, sub, tag, variant, SplitInfo{});
}
// Narrow contractions (kNarrowMaxN, meta.h) run on the vector ALUs
// (narrow_gemm_kernel, gemm.hip): one thread per output row and nc (4, 8 or
// 16) columns, kNarrowRows rows per workgroup.
constexpr int kNarrowRows = 256;
template <typename T>
void launch_narrow_gemm(Ctx *c, int la, int lb, const T *A, const T *B, T *C, const TileMap &tiles, int ntiles,
                        const BlkD *blks, const PairD *pairs, int nc, int sub = 0, int tag = 0);
std::vector<TileD> narrow_tiles(const std::vector<BlkD> &blks, int nc);
// Tile list of one grouped-GEMM launch (tile edge ts) ordered for the XCDs.
// kc[i] is block i's summed K extent (its per-tile cost).  See gemm.hip.
std::vector<TileD> schedule_tiles(const std::vector<BlkD> &blks, const std::vector<int64_t> &kc, int ts);
// The single-queue order of schedule_tiles (blocks heaviest first, stable),
// or an empty vector when an XCD-grouped order is selected (explicit lists).
std::vector<int32_t> tile_block_order(const std::vector<BlkD> &blks, const std::vector<int64_t> &kc);

// ------------------------------------------------------------------ autotuning (tune.hip)
// A grouped-GEMM kernel choice: tile edge (64 or 128) and, for FP32 128x128
// tiles, the slab-loading variant of launch_grouped_gemm.
struct GemmChoice {
  int tile = 64;
  int variant = 0;
};
// Tuning state of one plan and precision: candidates, the timed probes and
// the decision (index into cand, -1 while undecided).
struct PlanTune {
  std::vector<GemmChoice> cand;
  int decided = -1;
  int next = 0;                 // probes issued
  bool keyed = false;
  uint64_t key = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::vector<int> ev_cand;
  double best_ms = 0, heuristic_ms = 0;
  struct Probe {
    bool active = false;
    int cand = 0;
  };
  PlanTune() = default;
  PlanTune(const PlanTune &) = delete;
  PlanTune &operator=(const PlanTune &) = delete;
  ~PlanTune();
};

// ------------------------------------------------------------------ ops (ops.hip)
// Device-side contraction plan (descriptors uploaded once, cached by layout).
struct DevPlan {
  std::shared_ptr<ContractPlan> hp;
  BufP blks, pairs;
  int tile_m = 64, tile_n = 64;   // heuristic FP32 tile
  // the output tiled 64x64 and 128x128: decoded on the device from the block
  // order and prefix sums in tilemap (or explicit lists in tiles64 /
  // tiles128 for the XCD-grouped order)
  BufP tilemap, tiles64, tiles128;
  TileMap map64, map128;
  int ntiles64 = 0, ntiles128 = 0;
  SplitInfo split;                // output extent (split-K launches)
  // complex plans (hp in real terms, complex_view in ops.hip): the expansion
  // descriptors of B's blocks and the largest K x N among them
  bool cplx = false;
  BufP expand;
  int nexpand = 0;
  int64_t expand_max_kn = 0;
  // narrow plans (every output block at most kNarrowMaxN wide): the vector-ALU
  // kernel's column chunk (0 = matrix-core kernels) and its tile map
  int narrow_nc = 0;
  TileMap map_narrow;
  int ntiles_narrow = 0;
  mutable PlanTune tune[3];       // [0] FP64, [1] FP32, [2] split-BF16x3
  BufP perm_a, perm_b, perm_a_start, perm_b_start;   // permutation descriptors
  int nperm_a = 0, nperm_b = 0;
  // tiles of one owner's output blocks, by (owners, rank): tiles, count, flops
  struct Share {
    BufP tiles;
    int ntiles = 0;
    double flops = 0;
  };
  mutable std::map<std::pair<int, int>, Share> shares;
};

// Kernel for one launch of plan dp; with probe->active the caller brackets
// the launch with tune_probe_begin / tune_probe_end.
GemmChoice tune_select(Ctx *c, const DevPlan &dp, bool fp64, PlanTune::Probe *probe);
void tune_probe_begin(Ctx *c, const DevPlan &dp, bool fp64, PlanTune::Probe *probe);
void tune_probe_end(Ctx *c, const DevPlan &dp, bool fp64, PlanTune::Probe *probe);

std::shared_ptr<DevPlan> get_plan(Ctx *c, const MetaP &a, const MetaP &b, const std::vector<int> &ax,
                                  const std::vector<int> &bx);
// Drops every cached plan of context c (thread-safe).
void plan_cache_clear(Ctx *c);
TenP contract(const TenP &a, const TenP &b, const std::vector<int> &ax, const std::vector<int> &bx);
// Contract writing into `out` whose meta must contain every output block of the
// plan; blocks of `out` not produced are zero-filled.
void contract_into(const TenP &a, const TenP &b, const std::vector<int> &ax, const std::vector<int> &bx,
                   const TenP &out);
double dot(const Ten &a, const Ten &b);
double norm(const Ten &a);
double normalize(Ten &a);
// out = beta*out + sum c_i t_i; all tensors must share out's layout or be
// embeddable into it (missing blocks read as zero).
void lincomb(const std::vector<double> &coefs, const std::vector<TenP> &ts, double beta, const TenP &out);
// Copy t into the layout `m` (blocks of m absent from t are zero; blocks of t
// absent from m are an error).
TenP embed(const TenP &t, const MetaP &m);

// Effective-Hamiltonian matvec (where = lend | cent | rend) with the
// reference's contraction sequence; result embedded into `out` (full layout).
void heff_mul(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &psi,
              const std::string &where, const TenP &out, Prec prec);

// Truncated SVD (svd.hip).
struct SvdOut {
  TenP u, s, vt;
  double trunc_err = 0;
  int64_t D = 0;
  std::vector<int64_t> kept;   // per new-bond sector
};
enum OrthSide : int { kOrthLeft = 0, kOrthRight = 1 };
SvdOut svd(const TenP &t, int ldims, const QN &ldiv, double trunc_err, int64_t dmin, int64_t dmax,
           OrthSide side);
// Complex tensors (svd_complex.hip; svd() dispatches there).
SvdOut svd_complex(const TenP &t, int ldims, const QN &ldiv, double trunc_err, int64_t dmin, int64_t dmax,
                   OrthSide side);
// Singular values below kSvdResolve * s_max have no reproducible order in any
// FP64 SVD; the truncation treats every nonzero one of them as exactly that.
constexpr double kSvdResolve = 1e-11;
// The truncation rule (svd.hip) on per-sector singular values (descending
// within each sector): kept count per sector, discarded weight, total kept.
void svd_truncate_rule(const std::vector<std::vector<double>> &sv, double trunc_err, int64_t dmin, int64_t dmax,
                       std::vector<int64_t> &kept, double *err, int64_t *D);

// Symmetric tridiagonal lowest eigenpair on the device (tridiag.hip).
int tridiag_lowest_dev(Ctx *c, int n, const double *a, const double *b, double *eval, double *vec);
// Allocation-free form on device data: d_out[0] = eigenvalue, d_out[1..n] =
// eigenvector (want_vec); work space in a static per-device array.
void tridiag_lowest_noalloc(Ctx *c, int n, const double *d_a, const double *d_b, int want_vec, double *d_out);
// The same on device data with the caller's work space (5 n + 8 doubles),
// stream-ordered, no sync.
void tridiag_lowest_async(Ctx *c, int n, const double *d_a, const double *d_b, int want_vec, double *d_out,
                          double *d_work);
// Takes the device address of that array; called once per device, with the
// device current (context construction).
void tridiag_scratch_init(int device);

// Lanczos (lanczos.hip).
struct LanczosOut {
  int64_t iters = 0;
  double energy = 0;
  TenP gs;
  int64_t matvecs_lo = 0, matvecs_hi = 0;   // low-precision / FP64 matvecs
};
LanczosOut lanczos(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &init,
                   const std::string &where, double error, int64_t max_iter, Arith arith);

}  // namespace mps2
