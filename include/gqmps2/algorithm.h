// LanczosSolver and the two-site finite variational MPS (DMRG) sweep.
//
// Follows include/gqmps2/algorithm/lanczos_solver.h:27-43,
// lanczos_solver_impl.h:71-225 and algorithm/vmps/two_site_update_finite_vmps.h:27-55,
// two_site_update_finite_vmps_impl.h:43-340 of GraceQ/MPS2 0.2-alpha.0.
//
// Differences in mechanism, not in results: every tensor (MPS, MPO,
// environments, Lanczos vectors) stays resident in GPU memory for the whole
// run instead of being streamed to disk around each update; the MPS is read
// from `mps_path` when the container is empty and written back (and released)
// at the end, as the reference leaves it.  The Lanczos solve, the SVD and all
// contractions run on the device through the mps2_amd C ABI.
#ifndef MPS2_AMD_GQMPS2_ALGORITHM_H
#define MPS2_AMD_GQMPS2_ALGORITHM_H

#include <cmath>
#include <cstdio>
#include <functional>
#include <iomanip>

#include "gqmps2/finite_mps.h"

namespace gqmps2 {

// Precision of the Lanczos matvec contractions (extension; FP64 is the
// reference behaviour): Mixed runs them on the FP32 matrix cores, MixedBf16x3
// split-BF16x3 on the BF16 matrix cores, both with FP64 residual correction.
enum class LanczosPrecision {
  FP64 = MPS2_PREC_FP64,
  Mixed = MPS2_PREC_MIXED,
  MixedBf16x3 = MPS2_PREC_MIXED_BF16X3
};

struct LanczosParams {
  LanczosParams(double err, size_t max_iter, LanczosPrecision prec = LanczosPrecision::FP64)
      : error(err), max_iterations(max_iter), precision(prec) {}
  LanczosParams(double err) : LanczosParams(err, 200) {}
  LanczosParams() : LanczosParams(1.0E-7, 200) {}
  LanczosParams(const LanczosParams &o) = default;
  double error;
  size_t max_iterations;
  LanczosPrecision precision;
};

template <typename TenT>
struct LanczosRes {
  size_t iters;
  double gs_eng;
  TenT *gs_vec;
};

// lanczos_solver_impl.h:79-225.  As in the reference, the solver takes
// ownership of `pinit_state` (deleted on return) and returns a new ground
// state the caller deletes.
template <typename TenT>
LanczosRes<TenT> LanczosSolver(const std::vector<TenT *> &rpeff_ham, TenT *pinit_state, const LanczosParams &params,
                               const std::string &where) {
  // the absent environment of an end bond may be a null pointer or, as in the
  // reference's callers (test_lanczos_solver.cc:270-300), a default-constructed tensor
  auto h = [](TenT *t) -> mps2_ten { return t && !t->IsDefault() ? t->Handle() : nullptr; };
  int64_t iters = 0;
  double e = 0;
  mps2_ten gs = nullptr;
  int rc = mps2_lanczos(h(rpeff_ham[0]), rpeff_ham[1]->Handle(), rpeff_ham[2]->Handle(), h(rpeff_ham[3]),
                        pinit_state->Handle(), where.c_str(), params.error, (int64_t)params.max_iterations,
                        (int)params.precision, &iters, &e, &gs);
  delete pinit_state;
  if (rc != MPS2_OK) {
    std::cout << "Lanczos solver failed: " << mps2_last_error() << std::endl;
    exit(1);
  }
  LanczosRes<TenT> r;
  r.iters = (size_t)iters;
  r.gs_eng = e;
  r.gs_vec = new TenT;
  r.gs_vec->AdoptDevice(gs);
  return r;
}

struct SweepParams {
  SweepParams(const size_t sweeps, const size_t dmin, const size_t dmax, const double trunc_err,
              const LanczosParams &lancz_params, const std::string mps_path = kMpsPath,
              const std::string temp_path = kRuntimeTempPath)
      : sweeps(sweeps), Dmin(dmin), Dmax(dmax), trunc_err(trunc_err), lancz_params(lancz_params), mps_path(mps_path),
        temp_path(temp_path) {}
  size_t sweeps;
  size_t Dmin;
  size_t Dmax;
  double trunc_err;
  LanczosParams lancz_params;
  std::string mps_path;
  std::string temp_path;
  // Extension: write the whole MPS to mps_path at the end of every sweep
  // (the reference keeps the MPS on disk by streaming every tensor around
  // each update, two_site_update_finite_vmps_impl.h:343-494, so a run that
  // stops can be continued from the files).  0 writes only when the call
  // returns.
  size_t checkpoint_every = 1;
};

// Per-update record (extension: the reference only prints it).
struct UpdateRecord {
  size_t sweep, site;
  char dir;
  double energy, trunc_err, ee, lancz_time, update_time;
  size_t D, iters;
  double svd_time, env_time;
};

// Environment file names and removal (two_site_update_finite_vmps_impl.h:55-69):
// "<temp_path>/<dir>env<blk_len>.gqten", dir "l" or "r".  The sweep here keeps
// its environments in device memory and writes none of these files; the helpers
// remain for callers that look for or clean up such files under temp_path.
inline std::string GenEnvTenName(const std::string &dir, const long blk_len, const std::string temp_path) {
  return temp_path + "/" + dir + kEnvFileBaseName + std::to_string(blk_len) + "." + kGQTenFileSuffix;
}

inline void RemoveFile(const std::string &file) {
  if (std::remove(file.c_str())) perror(("Unable to delete " + file).c_str());
}

// two_site_update_finite_vmps_impl.h:43-52
template <typename DTenT>
inline double MeasureEE(const DTenT &s, const size_t sdim) {
  double ee = 0;
  for (size_t i = 0; i < sdim; ++i) {
    double p = std::pow(s.GetElem({i, i}), 2.0);
    if (p > 0) ee += -p * std::log(p);
  }
  return ee;
}

template <typename TenElemT, typename QNT>
class TwoSiteVMPSEngine {
 public:
  using TenT = GQTensor<TenElemT, QNT>;
  TwoSiteVMPSEngine(FiniteMPS<TenElemT, QNT> &mps, const MPO<TenT> &mpo, const SweepParams &sp)
      : mps_(mps), mpo_(mpo), sp_(sp), N_(mps.size()), lenvs_(N_), renvs_(N_) {}

  std::function<void(const UpdateRecord &)> on_update;
  // Called with the effective Hamiltonian and initial state right before each
  // Lanczos solve (measurement hook; extension).
  std::function<void(const std::vector<TenT *> &, const TenT &, const std::string &, char, size_t)> before_lanczos;
  bool quiet = false;

  // two_site_update_finite_vmps_impl.h:104-135
  void InitEnvs(size_t max_len = (size_t)-1) {
    for (size_t i = 1; i + 2 <= N_ && i <= max_len; ++i) {
      const TenT &A = mps_[N_ - i];
      if (i == 1) {
        TenT temp;
        Contract(&A, &mpo_.back(), {{1}, {0}}, &temp);
        auto ad = Dag(A);
        Contract(&temp, &ad, {{2}, {1}}, &renvs_[i]);
      } else {
        TenT t1, t2;
        Contract(&A, &renvs_[i - 1], {{2}, {0}}, &t1);
        Contract(&t1, &mpo_[N_ - i], {{1, 2}, {1, 3}}, &t2);
        auto ad = Dag(A);
        Contract(&t2, &ad, {{3, 1}, {1, 2}}, &renvs_[i]);
      }
    }
    // the environments are complete before the first update's phase clocks start
    CheckMps2(mps2_ctx_sync(DefaultCtx()), "sync");
  }

  // two_site_update_finite_vmps_impl.h:143-161
  double Sweep(size_t sweep_no) {
    double e0 = 0;
    for (size_t i = 0; i + 1 < N_; ++i) e0 = Update('r', i, sweep_no);
    for (size_t i = N_ - 1; i > 0; --i) e0 = Update('l', i, sweep_no);
    return e0;
  }

  // two_site_update_finite_vmps_impl.h:164-340
  double Update(const char dir, const size_t site, size_t sweep_no = 0) {
    Timer update_timer("update");
    std::vector<std::vector<size_t>> init_axes, us_axes;
    std::string where;
    size_t svd_ldims, ls, rs, llen, rlen;
    if (dir == 'r') {
      ls = site;
      rs = site + 1;
      llen = site;
      rlen = N_ - (site + 2);
      if (site == 0) {
        init_axes = {{1}, {0}};
        where = "lend";
        svd_ldims = 1;
      } else if (site == N_ - 2) {
        init_axes = {{2}, {0}};
        where = "rend";
        svd_ldims = 2;
      } else {
        init_axes = {{2}, {0}};
        where = "cent";
        svd_ldims = 2;
      }
    } else if (dir == 'l') {
      ls = site - 1;
      rs = site;
      llen = site - 1;
      rlen = N_ - site - 1;
      if (site == N_ - 1) {
        init_axes = {{2}, {0}};
        where = "rend";
        svd_ldims = 2;
        us_axes = {{2}, {0}};
      } else if (site == 1) {
        init_axes = {{1}, {0}};
        where = "lend";
        svd_ldims = 1;
        us_axes = {{1}, {0}};
      } else {
        init_axes = {{2}, {0}};
        where = "cent";
        svd_ldims = 2;
        us_axes = {{2}, {0}};
      }
    } else {
      std::cout << "dir must be 'r' or 'l', but " << dir << std::endl;
      exit(1);
    }
    std::vector<TenT *> eff_ham(4);
    eff_ham[0] = where == "lend" ? nullptr : &lenvs_[llen];
    eff_ham[1] = const_cast<TenT *>(&mpo_[ls]);
    eff_ham[2] = const_cast<TenT *>(&mpo_[rs]);
    eff_ham[3] = where == "rend" ? nullptr : &renvs_[rlen];
    auto init_state = new TenT;
    Contract(&mps_[ls], &mps_[rs], init_axes, init_state);
    if (before_lanczos) before_lanczos(eff_ham, *init_state, where, dir, site);
    Timer lancz_timer("Lancz");
    auto lr = LanczosSolver(eff_ham, init_state, sp_.lancz_params, where);
    double lancz_t = lancz_timer.Elapsed();

    Timer svd_timer("SVD");
    TenT u, vt;
    GQTensor<double, QNT> s;
    double err = 0;
    size_t D = 0;
    SVD(lr.gs_vec, svd_ldims, Div(mps_[ls]), sp_.trunc_err, sp_.Dmin, sp_.Dmax, &u, &s, &vt, &err, &D,
        dir == 'r' ? SvdOrth::Left : SvdOrth::Right);
    delete lr.gs_vec;
    double ee = MeasureEE(s, D);
    double svd_t = svd_timer.Elapsed();
    Timer env_timer("env");

    if (dir == 'r') {
      mps_[ls] = std::move(u);
      TenT other;
      Contract(&s, &vt, {{1}, {0}}, &other);
      mps_[rs] = std::move(other);
    } else {
      TenT other;
      Contract(&u, &s, us_axes, &other);
      mps_[ls] = std::move(other);
      mps_[rs] = std::move(vt);
    }

    if (dir == 'r') {
      if (site != N_ - 2) {
        if (site == 0) {
          TenT temp, env;
          Contract(&mps_[site], &mpo_[site], {{0}, {0}}, &temp);
          auto ad = Dag(mps_[site]);
          Contract(&temp, &ad, {{2}, {0}}, &env);
          lenvs_[llen + 1] = std::move(env);
        } else {
          TenT t1, t2, env;
          Contract(&lenvs_[llen], &mps_[site], {{0}, {0}}, &t1);
          Contract(&t1, &mpo_[site], {{0, 2}, {0, 1}}, &t2);
          auto ad = Dag(mps_[site]);
          Contract(&t2, &ad, {{0, 2}, {0, 1}}, &env);
          lenvs_[llen + 1] = std::move(env);
        }
      }
    } else {
      if (site != 1) {
        if (site == N_ - 1) {
          TenT temp, env;
          Contract(&mps_[site], &mpo_[site], {{1}, {0}}, &temp);
          auto ad = Dag(mps_[site]);
          Contract(&temp, &ad, {{2}, {1}}, &env);
          renvs_[rlen + 1] = std::move(env);
        } else {
          TenT t1, t2, env;
          Contract(&mps_[site], &renvs_[rlen], {{2}, {0}}, &t1);
          Contract(&t1, &mpo_[site], {{1, 2}, {1, 3}}, &t2);
          auto ad = Dag(mps_[site]);
          Contract(&t2, &ad, {{3, 1}, {1, 2}}, &env);
          renvs_[rlen + 1] = std::move(env);
        }
      }
    }
    CheckMps2(mps2_ctx_sync(DefaultCtx()), "sync");
    double env_t = env_timer.Elapsed();
    double upd_t = update_timer.Elapsed();
    if (!quiet) {
      std::cout << "Site " << std::setw(4) << site << " E0 = " << std::setw(20)
                << std::setprecision(kLanczEnergyOutputPrecision) << std::fixed << lr.gs_eng
                << " TruncErr = " << std::setprecision(2) << std::scientific << err << std::fixed
                << " D = " << std::setw(5) << D << " Iter = " << std::setw(3) << lr.iters
                << " LanczT = " << std::setw(8) << lancz_t << " TotT = " << std::setw(8) << upd_t
                << " S = " << std::setw(10) << std::setprecision(7) << ee;
      std::cout << std::scientific << std::endl;
    }
    if (on_update)
      on_update(UpdateRecord{sweep_no, site, dir, lr.gs_eng, err, ee, lancz_t, upd_t, D, lr.iters, svd_t, env_t});
    return lr.gs_eng;
  }

  std::vector<TenT> &lenvs() { return lenvs_; }
  std::vector<TenT> &renvs() { return renvs_; }

 private:
  FiniteMPS<TenElemT, QNT> &mps_;
  const MPO<TenT> &mpo_;
  SweepParams sp_;
  size_t N_;
  std::vector<TenT> lenvs_, renvs_;
};

// two_site_update_finite_vmps_impl.h:77-101
template <typename TenElemT, typename QNT>
double TwoSiteFiniteVMPS(FiniteMPS<TenElemT, QNT> &mps, const MPO<GQTensor<TenElemT, QNT>> &mpo,
                         const SweepParams &sweep_params,
                         std::function<void(const UpdateRecord &)> on_update = nullptr) {
  assert(mps.size() == mpo.size());
  bool from_disk = false;
  for (size_t i = 0; i < mps.size(); ++i)
    if (mps(i) == nullptr) {
      mps.LoadTen(i, GenMPSTenName(sweep_params.mps_path, i));
      from_disk = true;
    }
  if (!IsPathExist(sweep_params.temp_path)) CreatPath(sweep_params.temp_path);
  TwoSiteVMPSEngine<TenElemT, QNT> eng(mps, mpo, sweep_params);
  eng.on_update = on_update;
  eng.InitEnvs();
  std::cout << "\n";
  double e0 = 0;
  for (size_t sweep = 1; sweep <= sweep_params.sweeps; ++sweep) {
    std::cout << "sweep " << sweep << std::endl;
    Timer sweep_timer("sweep");
    e0 = eng.Sweep(sweep);
    sweep_timer.PrintElapsed();
    std::cout << "\n";
    // Sweep-level checkpoint: a right-canonical MPS centred at site 0, the
    // state TwoSiteFiniteVMPS starts from when it reads mps_path.
    if (sweep_params.checkpoint_every && sweep % sweep_params.checkpoint_every == 0 && sweep < sweep_params.sweeps) {
      const FiniteMPS<TenElemT, QNT> &cmps = mps;
      cmps.Dump(sweep_params.mps_path);
    }
  }
  (void)from_disk;
  mps.Dump(sweep_params.mps_path, true);
  return e0;
}

}  // namespace gqmps2

#endif
