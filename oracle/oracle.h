/*
 * CPU FP64 oracle for the two-site finite vMPS (DMRG) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product (mps2_amd/) never
 * links, loads or calls it.
 *
 * It restates, in plain C++ with its own GEMM / Jacobi SVD / implicit-QL
 * tridiagonal solver, the algorithm of GraceQ/MPS2 0.2-alpha.0:
 *   - Contract / Dag / LinearCombine / Normalize semantics of the gqten calls
 *     made by include/gqmps2/algorithm/lanczos_solver_impl.h
 *   - LanczosSolver            lanczos_solver_impl.h:79-225
 *   - eff_ham_mul_state_*      lanczos_solver_impl.h:228-256
 *   - TridiagGsSolver (dstev)  lanczos_solver_impl.h:259-312
 *   - TwoSiteFiniteVMPS sweep  two_site_update_finite_vmps_impl.h:77-340
 *   - InitEnvs                 two_site_update_finite_vmps_impl.h:104-135
 *   - FiniteMPS::Centralize    finite_mps.h:148-243
 *
 * Tensors cross this C ABI in the "flat tensor" format shared with the
 * product library (include/mps2_amd.h):
 *   int32 meta[]: rank, qn_width, nblocks,
 *                 per leg: dir (0 IN, 1 OUT), nsect, per sector: qn[qn_width], dim
 *                 per block: sector coordinate per leg (blocks sorted lexicographically)
 *   double data[]: blocks in meta order, each dense row-major over its sector dims.
 */
#ifndef MPS2_ORACLE_H
#define MPS2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *oracle_ten;

const char *oracle_last_error(void);

/* Tensor import / export / free. */
oracle_ten oracle_ten_import(const int32_t *meta, const double *data);
/* Complex tensor: data holds (real, imaginary) pairs, interleaved; the
 * operations below accept complex tensors (dot returns Re <a|b>). */
oracle_ten oracle_ten_import_complex(const int32_t *meta, const double *data);
int oracle_ten_is_complex(oracle_ten t);
int oracle_ten_export_sizes(oracle_ten t, int64_t *nmeta, int64_t *ndata);
int oracle_ten_export(oracle_ten t, int32_t *meta, double *data);
void oracle_ten_free(oracle_ten t);

/* gqten-equivalent operations. */
oracle_ten oracle_contract(oracle_ten a, oracle_ten b, int nax,
                           const int32_t *axes_a, const int32_t *axes_b);
oracle_ten oracle_dag(oracle_ten a);
double oracle_normalize(oracle_ten a); /* returns pre-normalisation norm */
double oracle_dot(oracle_ten a, oracle_ten b);

/* Truncated SVD, ψ split after `ldims` legs; U carries divergence ldiv. */
int oracle_svd(oracle_ten t, int ldims, const int32_t *ldiv, double trunc_err,
               int64_t dmin, int64_t dmax, oracle_ten *u, oracle_ten *s,
               oracle_ten *vt, double *actual_trunc_err, int64_t *d_kept);

/* Lowest eigenpair of the symmetric tridiagonal (a[0..n-1], b[0..n-2]).
 * vec may be NULL.  Returns 0 on success. */
int oracle_tridiag_lowest(int n, const double *a, const double *b,
                          double *eval, double *vec);

/* Lanczos on the effective Hamiltonian {L, W1, W2, R}; where = "lend",
 * "cent" or "rend".  init is not modified (the reference consumes it). */
int oracle_lanczos(oracle_ten L, oracle_ten W1, oracle_ten W2, oracle_ten R,
                   oracle_ten init, const char *where, double error,
                   int64_t max_iter, int64_t *iters, double *energy,
                   oracle_ten *gs);

/* H_eff * psi for one bond (where = "lend" | "cent" | "rend"). */
oracle_ten oracle_heff_mul(oracle_ten L, oracle_ten W1, oracle_ten W2,
                           oracle_ten R, oracle_ten psi, const char *where);

/* Right-canonicalise an MPS in place (Centralize(0)). mps[i] are replaced. */
int oracle_centralize0(int n, oracle_ten *mps);

/* Full two-site DMRG: InitEnvs + `sweeps` sweeps.  mps[] replaced by the
 * final MPS.  Per-update logs (2(N-1) per sweep) written when non-NULL:
 * energies, kept D, Lanczos iterations, truncation errors. */
int oracle_dmrg(int n, oracle_ten *mps, oracle_ten *mpo, int64_t sweeps,
                int64_t dmin, int64_t dmax, double trunc_err,
                double lanczos_error, int64_t lanczos_max_iter,
                double *energies, int64_t *dims, int64_t *iters,
                double *trunc_errs, double *final_energy);

/* <psi| prod_i ops[i] |psi> of the MPS as given (any gauge, unnormalised):
 * the full transfer chain; ops[i] NULL (or ops NULL) = identity at site i. */
int oracle_expect(int n, oracle_ten *mps, oracle_ten *ops, double *out);

/* Threads used by the oracle's GEMM (OpenMP); returns the count. */
int oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
