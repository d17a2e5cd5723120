/*
 * mps2_amd: C ABI of the MI355X block-sparse tensor engine behind the
 * two-site finite vMPS (DMRG) hot path of GraceQ/MPS2.
 *
 * This is the boundary the host C++ (gqmps2-compatible) layer and any FFI
 * caller bind to.  Each entry point replaces one GraceQ/tensor (gqten) call,
 * or one gqmps2 routine, made on the hot path; the reference call sites are
 * cited per function (paths relative to the GraceQ/MPS2 0.2-alpha.0 tree).
 *
 * Conventions
 *  - Every function returns 0 on success or a nonzero status
 *    (MPS2_ERR_*); mps2_last_error() gives the message of the calling
 *    thread's last failure.  The reference exits the process on solver
 *    failure (lanczos_solver_impl.h:288-291); here the caller decides.
 *  - Tensors are opaque device-resident handles owned by the caller
 *    (mps2_ten_free).  One context = one GPU + one HIP stream; a context and
 *    its tensors must be used from one host thread at a time.
 *  - Flat tensor format used for import/export:
 *      int32 meta[]: rank, qn_width, nblocks,
 *                    per leg: dir (0 = IN, 1 = OUT), nsect,
 *                             per sector: qn[qn_width], dim
 *                    per block: sector coordinate of each leg
 *                    (blocks unique, sorted lexicographically)
 *      double data[]: the blocks in meta order, each dense row-major.
 *    Block divergence = sum(qn of OUT legs) - sum(qn of IN legs).
 *  - Precision for mps2_lanczos: MPS2_PREC_FP64 is the plain FP64 form (all
 *    contractions FP64 on FP64 MFMA); MPS2_PREC_MIXED runs the matvec
 *    contractions in FP32 on MFMA with FP64 residual correction;
 *    MPS2_PREC_MIXED_BF16X3 is the same restarted form with the matvec
 *    products split-BF16x3 (each FP32 operand as two BF16 parts, three BF16
 *    products per multiply on the BF16 matrix cores, FP32 accumulation).
 *    Any other precision code is an argument error.
 */
#ifndef MPS2_AMD_H
#define MPS2_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPS2_OK 0
#define MPS2_ERR_ARG 1
#define MPS2_ERR_HIP 2
#define MPS2_ERR_OOM 3
#define MPS2_ERR_NUMERIC 4
#define MPS2_ERR_STATE 5
#define MPS2_ERR_NO_DEVICE 6

#define MPS2_PREC_FP64 0
#define MPS2_PREC_MIXED 1
#define MPS2_PREC_MIXED_BF16X3 2

#define MPS2_ORTH_LEFT 0
#define MPS2_ORTH_RIGHT 1

typedef struct mps2_ctx_s *mps2_ctx;
typedef struct mps2_ten_s *mps2_ten;

/* Library version (major*10000 + minor*100 + patch). */
int mps2_version(void);
/* Message of this thread's last failure ("" if none). */
const char *mps2_last_error(void);
/* Number of visible HIP devices (0 on a host without a GPU). */
int mps2_device_count(int *n);

/* Context: one device, one stream, caching allocator, plan cache. */
int mps2_ctx_create(int device, mps2_ctx *out);
/* Invalidates the handle.  Tensors created from it may still be exported and
 * must still be freed; the device context goes away with the last of them. */
int mps2_ctx_destroy(mps2_ctx ctx);
int mps2_ctx_sync(mps2_ctx ctx);
/* Counters: algorithmic GEMM flops issued, grouped-GEMM launches, live and
 * peak device bytes held by the allocator. */
int mps2_ctx_stats(mps2_ctx ctx, double *gemm_flops, int64_t *gemm_launches, int64_t *live_bytes,
                   int64_t *peak_bytes);
int mps2_ctx_reset_stats(mps2_ctx ctx);
/* Named event counters of the context ("svd_calls", "svd_sweeps",
 * "svd_rounds", "syncs", ...); a name never counted reads 0.  Reset with
 * mps2_ctx_reset_stats. */
int mps2_ctx_counter(mps2_ctx ctx, const char *name, double *value);

/* Grouped-GEMM autotuning of the context's contraction plans (no reference
 * counterpart: the reference's contractions run in MKL).  mode 1 (default):
 * the first launches of an FP32 plan of at least 1e9 flops time the candidate
 * kernels, whose results are identical, and keep the fastest, shared by
 * plans of the same block-shape histogram and stored in a cache file
 * ($MPS2_TUNE_CACHE, else ~/.cache/mps2_amd/gemm_tune_gfx950.txt); mode 0:
 * the heuristic kernel only.  Counters "tune_probes", "tune_decisions",
 * "tune_changed", "tune_cache_hits". */
int mps2_ctx_set_tuning(mps2_ctx ctx, int mode);

/* Tensors.  Replace GQTensor construction, element fill and the stream I/O
 * used by TenVec::LoadTen/DumpTen (one_dim_tn/framework/ten_vec.h:50-85). */
int mps2_ten_import(mps2_ctx ctx, const int32_t *meta, const double *data, mps2_ten *out);
/* A GQTensor<GQTEN_Complex> (gqten's std::complex<double> element type,
 * test_two_site_update_finite_vmps.cc:374-484,704-786): the same meta, data
 * holding two doubles (real, imaginary) per element, interleaved.  Complex
 * tensors are exported in that form (mps2_ten_export_sizes counts doubles);
 * Contract, Dag (conjugating), LinearCombine and Normalize (real
 * coefficients), the effective Hamiltonian, Lanczos (MPS2_PREC_FP64) and SVD
 * accept them; a real operand of a contraction with a complex one enters as
 * complex; mps2_dot returns Re <a|b>. */
int mps2_ten_import_complex(mps2_ctx ctx, const int32_t *meta, const double *data, mps2_ten *out);
/* 1 for a complex tensor, 0 for a real one. */
int mps2_ten_is_complex(mps2_ten t, int *is_complex);
int mps2_ten_export_sizes(mps2_ten t, int64_t *nmeta, int64_t *ndata);
int mps2_ten_export(mps2_ten t, int32_t *meta, double *data);
int mps2_ten_free(mps2_ten t);
int mps2_ten_copy(mps2_ten t, mps2_ten *out);
/* gqten Dag for real tensors: flip every index direction (data shared).
 * Reference call sites: lanczos_solver_impl.h:135,197;
 * two_site_update_finite_vmps_impl.h:119,128,291,298,309,316. */
int mps2_ten_dag(mps2_ten t, mps2_ten *out);
/* gqten Div: divergence QN (qn_width ints).
 * Reference call site: two_site_update_finite_vmps_impl.h:260. */
int mps2_ten_div(mps2_ten t, int32_t *qn);

/* gqten Contract(a, b, {axes_a, axes_b}, out): output legs are the free legs
 * of a then of b; contracted legs must be mutually inverse indexes.
 * Reference call sites: lanczos_solver_impl.h:136,198,231-234,242-244,252-254;
 * two_site_update_finite_vmps_impl.h:118,120,124,127,129,243,272,276,290-317. */
int mps2_contract(mps2_ten a, mps2_ten b, int nax, const int32_t *axes_a, const int32_t *axes_b, mps2_ten *out);
/* Contract with the arithmetic precision selectable (MPS2_PREC_MIXED: operands
 * rounded to FP32, FP32 MFMA with FP32 accumulation, result widened to FP64;
 * MPS2_PREC_MIXED_BF16X3: the same FP32 operands multiplied split-BF16x3).
 * Same semantics as mps2_contract otherwise; used by the mixed matvec. */
int mps2_contract_ex(mps2_ten a, mps2_ten b, int nax, const int32_t *axes_a, const int32_t *axes_b, int precision,
                     mps2_ten *out);

/* gqten LinearCombine: out = beta*out + sum_i coefs[i]*ts[i] (block sets
 * are united).  Reference call sites: lanczos_solver_impl.h:151,153,173,215. */
int mps2_linear_combine(int n, const double *coefs, const mps2_ten *ts, double beta, mps2_ten out);

/* GQTensor::Normalize: scale to unit 2-norm, return the previous norm.
 * Reference call sites: lanczos_solver_impl.h:120,160. */
int mps2_normalize(mps2_ten t, double *norm_before);

/* Full contraction <a|b> of two real tensors (Contract(a, Dag(b)) over all
 * legs), as in lanczos_solver_impl.h:134-141,196-204. */
int mps2_dot(mps2_ten a, mps2_ten b, double *out);

/* gqten truncated SVD(t, ldims, ldiv, trunc_err, Dmin, Dmax, &u, &s, &vt,
 * &actual_trunc_err, &D): legs [0, ldims) go to u (plus a new OUT leg, u
 * carries divergence ldiv), s is the real diagonal (IN, OUT) on the new bond,
 * vt has the new IN leg then the remaining legs.  orth_side selects which
 * factor is built exactly orthonormal (MPS2_ORTH_LEFT: u; MPS2_ORTH_RIGHT:
 * vt); the other is t projected on it and divided by s.
 * Reference call sites: two_site_update_finite_vmps_impl.h:258-263;
 * finite_mps.h:184,221 (through mock_gqten::SVD, utilities.h:29-55). */
int mps2_svd(mps2_ten t, int ldims, const int32_t *ldiv, double trunc_err, int64_t dmin, int64_t dmax,
             int orth_side, mps2_ten *u, mps2_ten *s, mps2_ten *vt, double *actual_trunc_err, int64_t *d_kept);

/* Effective Hamiltonian times state: eff_ham_mul_state_{lend,cent,rend}
 * (lanczos_solver_impl.h:228-256).  L (for lend) and R (for rend) may be
 * NULL, as in the reference where they are unused. */
int mps2_heff_mul(mps2_ten L, mps2_ten W1, mps2_ten W2, mps2_ten R, mps2_ten psi, const char *where, mps2_ten *out);

/* The effective Hamiltonian bound once to its operands: L, W1, W2, R and the
 * block layout of `psi`, at one precision (MPS2_PREC_MIXED and
 * MPS2_PREC_MIXED_BF16X3 make the FP32 operand copies once).  It is the operator that eff_ham_mul_state_{lend,
 * cent,rend} (lanczos_solver_impl.h:228-256) applies at every Lanczos step
 * (:149,171), with its setup hoisted out of the loop.  FP64 operands are
 * shared with the caller's tensors (kept alive by the operator; they must not
 * be modified while it is in use).  mps2_heff_run applies it n times to the
 * bound state (psi embedded into the full two-site layout, at the
 * operator's precision), stream-ordered, without a host sync: the matvec
 * alone, as a benchmark step or a warm-up. */
typedef struct mps2_heff_s *mps2_heff;
int mps2_heff_create(mps2_ten L, mps2_ten W1, mps2_ten W2, mps2_ten R, mps2_ten psi, const char *where,
                     int precision, mps2_heff *out);
int mps2_heff_run(mps2_heff h, int n);
/* Algorithmic flops of one application (sum over QN-matched block pairs of
 * 2*M*N*K) and the FP64 bytes of its operands plus two states. */
int mps2_heff_info(mps2_heff h, double *flops, double *bytes);
int mps2_heff_destroy(mps2_heff h);

/* LanczosSolver({L, W1, W2, R}, init, {error, max_iterations}, where)
 * (lanczos_solver_impl.h:79-225).  init is not consumed (the caller frees it).
 * Returns iterations, ground energy and a new ground-state tensor. */
int mps2_lanczos(mps2_ten L, mps2_ten W1, mps2_ten W2, mps2_ten R, mps2_ten init, const char *where, double error,
                 int64_t max_iter, int precision, int64_t *iters, double *energy, mps2_ten *gs);

/* Benchmark of one effective-Hamiltonian matvec (no reference counterpart:
 * measurement only).  Runs `warmup` untimed then `reps` timed matvecs on the
 * context stream, bracketed by HIP events, and times every grouped-GEMM
 * launch with its own events.  out[0] ms per matvec, out[1] algorithmic
 * flops per matvec (sum over QN-matched block pairs of 2*M*N*K), out[2] ms
 * per matvec inside grouped GEMMs, out[3] index of the stage with the longest
 * GEMM time, out[4] that stage's flops per launch, out[5] its ms per launch
 * measured in a second pass of `reps` matvecs in which that stage runs as a
 * separately named kernel instance (grouped_gemm_kernel<..., 1>), so a
 * profiler's per-kernel statistics of that instance cover these launches,
 * out[6] FP64 bytes of L, W1, W2, R and two states (the fused-ideal traffic),
 * out[7..10] ms per launch of stages 1..4, out[11] the dominant stage's
 * algorithmic bytes per launch (its operands read once and its result
 * written once, at the working precision). */
int mps2_heff_bench(mps2_ten L, mps2_ten W1, mps2_ten W2, mps2_ten R, mps2_ten psi, const char *where,
                    int precision, int warmup, int reps, double *out);

/* One GPU's share of a sharded matvec stage (no reference counterpart:
 * measurement of the owner-computes design).  The output blocks of matvec
 * stage `stage` (0-based) are assigned to `nshard` owners by
 * longest-processing-time on their flops; `reps` matvecs run with that stage
 * computing only rank `shard`'s blocks (the output is then incomplete).
 * out[0] ms per launch of the share, out[1] the share's algorithmic flops,
 * out[2] the whole stage's flops. */
int mps2_heff_bench_shard(mps2_ten L, mps2_ten W1, mps2_ten W2, mps2_ten R, mps2_ten psi, const char *where,
                          int precision, int stage, int nshard, int shard, int reps, double *out);

/* Plain device-to-device copy of `bytes` (a multiple of 16) repeated `reps`
 * times by a streaming kernel (calib_copy_kernel); ms per copy from HIP
 * events.  Measurement only (no reference counterpart): attainable HBM
 * bandwidth, and the known-size kernel that scales profiler byte counters. */
int mps2_copy_bench(mps2_ctx ctx, int64_t bytes, int reps, double *ms_per_copy);

/* TridiagGsSolver (lanczos_solver_impl.h:259-312): lowest eigenpair of the
 * symmetric tridiagonal (a[0..n-1], b[0..n-2]), solved on the device.  vec may
 * be NULL (jobz 'N'). */
int mps2_tridiag_lowest(mps2_ctx ctx, int n, const double *a, const double *b, double *eval, double *vec);

/* Allocation-free form of mps2_tridiag_lowest.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  d_a (n values), d_b (n - 1 values) and d_out are device
 * pointers the caller owns; d_out[0] receives the lowest eigenvalue and,
 * when want_vec is nonzero, d_out[1..n] its unit eigenvector.  One kernel
 * launch on the context's stream, with no allocation and no host sync (the
 * caller synchronises, e.g. mps2_ctx_sync).  The kernel's work space is a
 * static array per GPU, held by the library until the process ends, sized
 * for n <= MPS2_NOALLOC_MAX_N; larger n returns MPS2_ERR_ARG.  Every call
 * on one GPU uses that GPU's array, so calls that overlap on one GPU (other
 * streams or threads) race. */
#define MPS2_NOALLOC_MAX_N 4096
int mps2_tridiag_lowest_noalloc(mps2_ctx ctx, int n, const double *d_a, const double *d_b, int want_vec,
                                double *d_out);

#ifdef __cplusplus
}
#endif
#endif /* MPS2_AMD_H */
