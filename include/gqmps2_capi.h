/*
 * gqmps2_capi: C ABI over the gqmps2-compatible host layer (include/gqmps2/),
 * for FFI callers (Python ctypes in tests/ and bench.py).
 *
 * A model object owns a SiteVec, an MPOGenerator, the generated MPO and an
 * MPS.  The calls mirror the reference's user-level flow
 * (doc/index.md:338-406; test_two_site_update_finite_vmps.cc:92-220):
 * SiteVec -> MPOGenerator::AddTerm* -> Gen() -> {Direct,ExtendDirectRandom}InitMps
 * -> TwoSiteFiniteVMPS.  Status codes and mps2_last_error() as in mps2_amd.h.
 *
 * Operators are given as dense d x d row-major matrices in the gqten element
 * convention op(i, j) with i the IN (ket) coordinate and j the OUT coordinate
 * of the site's physical index.
 */
#ifndef GQMPS2_CAPI_H
#define GQMPS2_CAPI_H

#include <stdint.h>

#include "mps2_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gqmps2_model_s *gqmps2_model;

/* qn_width 1 = U(1), 2 = U(1)xU(1).  Site i has nsect[i] sectors; qns and
 * dims hold all sites' sectors back to back (qn_width ints per sector). */
int gqmps2_model_create(int qn_width, int nsites, const int32_t *nsect, const int32_t *qns, const int32_t *dims,
                        gqmps2_model *out);
int gqmps2_model_destroy(gqmps2_model m);

/* Register an operator acting on `site`'s local space; returns its id. */
int gqmps2_model_add_op(gqmps2_model m, int site, const double *dense, int32_t *op_id);

/* MPOGenerator::AddTerm (mpogen_impl.h:87-245): coef * prod_k op[k] at
 * sites[k] (ascending).  With ninst > 0, insertion operators are placed
 * between consecutive physical operators (ninst == nops - 1) and, when
 * ninst == nops, also as a tail string to the end of the chain. */
int gqmps2_model_add_term(gqmps2_model m, double coef, int nops, const int32_t *op_ids, const int64_t *sites,
                          int ninst, const int32_t *inst_op_ids);

/* MPOGenerator::Gen (mpogen_impl.h:248-285); bond_dims gets N-1 values. */
int gqmps2_model_gen_mpo(gqmps2_model m, int64_t *bond_dims);
int gqmps2_model_mpo_export_sizes(gqmps2_model m, int site, int64_t *nmeta, int64_t *ndata);
int gqmps2_model_mpo_export(gqmps2_model m, int site, int32_t *meta, double *data);

/* DirectStateInitMps (enlarged_dim == 0, nsets == 1) or
 * ExtendDirectRandomInitMps (finite_mps_init.h:216-330), seeded.  With
 * centralize == 0 the raw tensors are kept (no Centralize(0)). */
int gqmps2_mps_init(gqmps2_model m, int nsets, const int64_t *stat_labs, int64_t enlarged_dim, uint64_t seed,
                    int centralize);
int gqmps2_mps_export_sizes(gqmps2_model m, int site, int64_t *nmeta, int64_t *ndata);
int gqmps2_mps_export(gqmps2_model m, int site, int32_t *meta, double *data);
int gqmps2_mps_import(gqmps2_model m, int site, const int32_t *meta, const double *data);
/* Save / restore the MPS inside device memory (for runs from one state). */
int gqmps2_mps_checkpoint(gqmps2_model m);
int gqmps2_mps_restore(gqmps2_model m);

/* Two-site DMRG: InitEnvs + `sweeps` sweeps (two_site_update_finite_vmps_impl.h:77-161)
 * on the resident MPS.  dmax_schedule (length sweeps) may be NULL (constant
 * dmax).  Per-update outputs (2(N-1) per sweep) may be NULL. */
int gqmps2_dmrg(gqmps2_model m, int64_t sweeps, int64_t dmin, int64_t dmax, const int64_t *dmax_schedule,
                double trunc_err, double lanczos_error, int64_t lanczos_max_iter, int precision, int verbose,
                double *energies, int64_t *dims, int64_t *iters, double *trunc_errs, double *update_times,
                double *sweep_times, double *final_energy);

/* Measurement hook (no reference counterpart): during later gqmps2_dmrg
 * calls, right before the Lanczos solve of the right-moving update at `site`
 * of the first sweep, run mps2_heff_bench on that bond's effective
 * Hamiltonian for FP64 and mixed precision and keep copies of L, W1, W2, R
 * and the initial state.  site < 0 disables it. */
int gqmps2_set_bench(gqmps2_model m, int64_t site, int warmup, int reps);
/* Results: out[0..11] FP64, out[12..23] mixed (layout of mps2_heff_bench). */
int gqmps2_get_bench(gqmps2_model m, double *out);
/* New handles on the kept tensors (L, W1, W2, R, psi); NULL where absent. */
int gqmps2_bench_tensors(gqmps2_model m, mps2_ten *out5);
/* Move the MPS centre to `site` (FiniteMPS::Centralize), build the left and
 * right environments of the bond (site, site+1) and benchmark its matvec as
 * gqmps2_set_bench does; results through gqmps2_get_bench and
 * gqmps2_bench_tensors.  Leaves the MPS centred at `site`. */
int gqmps2_bench_bond(gqmps2_model m, int64_t site, int warmup, int reps);

/* Host wall-clock seconds (Lanczos, SVD, environment growth) of each update
 * of the last gqmps2_dmrg call, 3 values per update, n values at most. */
int gqmps2_phase_times(gqmps2_model m, double *out, int64_t n);
/* Console output of the host layer on (1, the reference's behaviour) or off. */
int gqmps2_set_verbose(int verbose);

/* TruncateMPS (finite_mps.h:245-303) of the resident MPS: Centralize(0),
 * then one truncated SVD per bond, left to right, with normalised S. */
int gqmps2_mps_truncate(gqmps2_model m, double trunc_err, int64_t dmin, int64_t dmax);

/* MeasureOneSiteOp (finite_mps_measu.h:150-199): each of the nops operators
 * at every site, the MPS centred there; avgs[j * N + i] = <op_j at site i>.
 * Writes basenames[j] + ".json" (DumpMeasuRes, finite_mps_measu.h:556-582).
 * Leaves the MPS centred at site N-1. */
int gqmps2_measure_one_site(gqmps2_model m, int nops, const int32_t *op_ids, const char *const *basenames,
                            double *avgs);
/* MeasureMultiSiteOp (finite_mps_measu.h:290-361): event e has nphys[e] >= 2
 * physical operators phys_op_ids[] at ascending sites[] (both back to back
 * over the events) and nphys[e] - 1 entries of inst_op_ids[], the insertion
 * operator repeated over each gap between consecutive physical operators
 * (-1: the site's identity).  avgs[e] = <event e>; writes basename + ".json".
 * MeasureTwoSiteOp is the nphys == 2 case. */
int gqmps2_measure_multi_site(gqmps2_model m, int64_t nevents, const int32_t *nphys, const int64_t *sites,
                              const int32_t *phys_op_ids, const int32_t *inst_op_ids, const char *basename,
                              double *avgs);

/* Engine context used by this library (for mps2_ctx_stats). */
mps2_ctx gqmps2_engine_ctx(void);
/* Message of this thread's last gqmps2_* failure. */
const char *gqmps2_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GQMPS2_CAPI_H */
