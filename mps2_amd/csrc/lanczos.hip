// Lanczos ground-state solver on the effective Hamiltonian, device-resident.
//
// Plain FP64 form: the recurrence, stopping rule and final Ritz combination
// of LanczosSolver (include/gqmps2/algorithm/lanczos_solver_impl.h:79-225):
// no reorthogonalisation, every basis vector kept (in HBM), stop when the
// tridiagonal ground energy decreases by less than `error`, or m reaches the
// effective dimension, or m == max_iter - 1.
//
// All Lanczos vectors share one block layout (every divergence-allowed block
// of the two-site state), so the BLAS-1 steps are flat vector kernels.
#include <hip/hip_runtime.h>

#include <cmath>

#include "runtime.h"

namespace mps2 {

namespace {

int64_t eff_dim_of(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const std::string &where) {
  auto d = [](const TenP &t, int leg) { return t->meta->legs[leg].total(); };
  if (where == "cent") return d(L, 0) * d(W1, 1) * d(W2, 1) * d(R, 0);
  if (where == "lend") return d(W1, 0) * d(W2, 1) * d(R, 0);
  if (where == "rend") return d(L, 0) * d(W1, 1) * d(W2, 0);
  throw Error(kErrArg, "where must be lend, cent or rend");
}

TenP final_combination(Ctx *c, const MetaP &m, const std::vector<TenP> &bases, int n, const std::vector<double> &z) {
  auto gs = ten_alloc(c, m);
  std::vector<const double *> ptrs(n);
  for (int i = 0; i < n; ++i) ptrs[i] = bases[i]->d();
  auto dp = c->alloc(n * sizeof(void *));
  auto dc = c->alloc(n * sizeof(double));
  MPS2_HIP(hipMemcpyAsync(dp->p, ptrs.data(), n * sizeof(void *), hipMemcpyHostToDevice, c->stream));
  MPS2_HIP(hipMemcpyAsync(dc->p, z.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_lincomb(c, m->total, n, static_cast<const double *const *>(dp->p), static_cast<const double *>(dc->p), gs->d());
  c->sync();
  return gs;
}

}  // namespace

LanczosOut lanczos_plain(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &init,
                         const std::string &where, double error, int64_t max_iter) {
  Ctx *c = init->ctx;
  MPS2_REQUIRE(max_iter >= 2, "Lanczos: max_iterations must be >= 2");
  int64_t eff_dim = eff_dim_of(L, W1, W2, R, where);
  MetaP full = full_meta(init->meta->qw, init->meta->legs, init->meta->div());
  std::vector<TenP> bases;
  std::vector<double> a(max_iter, 0.0), b(max_iter, 0.0), N(max_iter, 0.0);
  bases.push_back(embed(init, full));
  normalize(*bases[0]);
  TenP last = ten_alloc(c, full);
  heff_mul(L, W1, W2, R, bases[0], where, last, kF64);
  LanczosOut res;
  res.matvecs_hi = 1;
  a[0] = dot(*last, *bases[0]);
  double energy0 = a[0];
  int64_t m = 0;
  while (true) {
    m += 1;
    TenP gamma = last;
    if (m == 1) {
      k_axpby(c, gamma->n(), -a[m - 1], bases[m - 1]->d(), 1.0, gamma->d());
    } else {
      k_axpby(c, gamma->n(), -a[m - 1], bases[m - 1]->d(), 1.0, gamma->d());
      k_axpby(c, gamma->n(), -std::sqrt(N[m - 1]), bases[m - 2]->d(), 1.0, gamma->d());
    }
    double ng = normalize(*gamma);
    if (ng == 0.0) {
      res.iters = m;
      res.energy = energy0;
      if (m == 1) {
        res.gs = ten_copy(*bases[0]);
      } else {
        double ev;
        std::vector<double> z(m);
        tridiag_lowest_dev(c, (int)m, a.data(), b.data(), &ev, z.data());
        res.gs = final_combination(c, full, bases, (int)m, z);
      }
      return res;
    }
    N[m] = ng * ng;
    b[m - 1] = ng;
    bases.push_back(gamma);
    last = ten_alloc(c, full);
    heff_mul(L, W1, W2, R, bases[m], where, last, kF64);
    res.matvecs_hi++;
    a[m] = dot(*last, *bases[m]);
    double e_new;
    tridiag_lowest_dev(c, (int)m + 1, a.data(), b.data(), &e_new, nullptr);
    if ((energy0 - e_new) < error || m == eff_dim || m == max_iter - 1) {
      std::vector<double> z(m + 1);
      double ev;
      tridiag_lowest_dev(c, (int)m + 1, a.data(), b.data(), &ev, z.data());
      res.iters = m;
      res.energy = e_new;
      res.gs = final_combination(c, full, bases, (int)m + 1, z);
      return res;
    }
    energy0 = e_new;
  }
}

LanczosOut lanczos(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &init,
                   const std::string &where, double error, int64_t max_iter, Prec prec) {
  if (prec == kF64) return lanczos_plain(L, W1, W2, R, init, where, error, max_iter);
  throw Error(kErrArg, "Lanczos: precision mode not available");
}

}  // namespace mps2
