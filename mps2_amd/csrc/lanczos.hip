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

#include <algorithm>
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

// ---- device-side scalar steps of the recurrences (one thread each)

__global__ void lz_set_ptr_kernel(const double **table, int i, const double *p) { table[i] = p; }

// From dot = <g, g>: b = |g|, c2 = sqrt(b * b) (the coefficient the
// reference forms from N = b^2 one step later), inv = 1 / b (0 when b = 0).
__global__ void lz_norm_kernel(const double *dot, double *b, double *c2, double *inv) {
  const double ng = sqrt(*dot);
  *b = ng;
  *c2 = sqrt(ng * ng);
  *inv = ng > 0.0 ? 1.0 / ng : 0.0;
}
__global__ void lz_rsqrt_kernel(const double *dot, double *norm, double *inv) {
  const double n = sqrt(*dot);
  *norm = n;
  *inv = n > 0.0 ? 1.0 / n : 0.0;
}
// Small coefficient vectors travel as kernel arguments (no host staging).
struct CoefArg {
  int n;
  double v[64];
};
__global__ void lz_set_coef_kernel(CoefArg a, double *dst) {
  const int i = threadIdx.x;
  if (i < a.n) dst[i] = a.v[i];
}
__global__ void lz_pack_kernel(const double *src, int n, double *dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Per-solve device state: the recurrence scalars and the basis pointer table.
struct LzDev {
  Ctx *c;
  BufP buf;        // scalars
  BufP table;      // const double * per basis vector
  BufP work;       // tridiagonal work space
  BufP partial;    // multi-dot partial sums
  int cap;
  double *a, *b, *c2, *scal, *out;
  LzDev(Ctx *ctx, int max_vec) : c(ctx), cap(max_vec) {
    const int m = max_vec + 2;
    buf = c->alloc((size_t)(4 * m + 8 + (m + 1)) * sizeof(double));
    double *p = static_cast<double *>(buf->p);
    a = p, b = p + m, c2 = p + 2 * m, scal = p + 3 * m, out = p + 3 * m + 8;
    table = c->alloc((size_t)m * sizeof(void *));
    work = c->alloc((size_t)(5 * m + 8) * sizeof(double));
    partial = c->alloc((size_t)m * 256 * sizeof(double));
  }
  const double **ptrs() const { return static_cast<const double **>(table->p); }
  void set_ptr(int i, const double *p) {
    if (i >= cap + 2) throw Error(kErrState, "Lanczos: basis table overflow");
    lz_set_ptr_kernel<<<1, 1, 0, c->stream>>>(const_cast<const double **>(ptrs()), i, p);
  }
  // n doubles from device to the context's pinned host scalars, then sync
  const double *read(const double *src, int n) {
    MPS2_HIP(hipMemcpyAsync(c->host_scalars, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    return c->host_scalars;
  }
};

TenP combine_dev(Ctx *c, const MetaP &m, const LzDev &dv, int n, const double *d_coef, bool cplx = false) {
  auto gs = ten_alloc(c, m, kF64, cplx);
  k_lincomb(c, m->total * (cplx ? 2 : 1), n, dv.ptrs(), d_coef, gs->d());
  return gs;
}

bool any_complex(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &init) {
  for (const TenP *t : {&L, &W1, &W2, &R, &init})
    if (*t && (*t)->cplx) return true;
  return false;
}

}  // namespace

// Plain FP64 form (lanczos_solver_impl.h:79-225) with the recurrence scalars
// kept on the device: alpha, beta and the coefficients are reduced into and
// read from device memory, the tridiagonal ground energy is solved on the
// device from those arrays, and the host reads two numbers per iteration
// (energy, |gamma|) for the stopping rule.  Complex (GQTEN_Complex) problems
// run the same recurrence on the interleaved real form of the vectors: the
// Hermitian H_eff makes alpha = <v|Hv> real (the reference takes Real() of
// it, lanczos_solver_impl.h:64-67,136-141), every coefficient is real, and
// Re <a|b> is the real dot product of the two real forms.
LanczosOut lanczos_plain(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &init,
                         const std::string &where, double error, int64_t max_iter) {
  Ctx *c = init->ctx;
  MPS2_REQUIRE(max_iter >= 2, "Lanczos: max_iterations must be >= 2");
  int64_t eff_dim = eff_dim_of(L, W1, W2, R, where);
  MetaP full = full_meta(init->meta->qw, init->meta->legs, init->meta->div());
  const bool cz = any_complex(L, W1, W2, R, init);
  const int64_t n = full->total * (cz ? 2 : 1);
  LzDev dv(c, (int)max_iter);
  std::vector<TenP> bases;
  bases.push_back(embed(cz ? ten_promote(init) : init, full));
  k_dot_to(c, n, bases[0]->d(), bases[0]->d(), dv.scal);
  lz_rsqrt_kernel<<<1, 1, 0, c->stream>>>(dv.scal, dv.scal + 1, dv.scal + 2);
  k_scale_dev(c, n, dv.scal + 2, bases[0]->d());
  dv.set_ptr(0, bases[0]->d());
  TenP last = ten_alloc(c, full, kF64, cz);
  heff_mul(L, W1, W2, R, bases[0], where, last, kF64);
  LanczosOut res;
  res.matvecs_hi = 1;
  k_dot_to(c, n, last->d(), bases[0]->d(), dv.a);
  double energy0 = dv.read(dv.a, 1)[0];
  int64_t m = 0;
  while (true) {
    m += 1;
    TenP gamma = last;
    k_axpy_dev(c, n, -1.0, dv.a + (m - 1), bases[m - 1]->d(), gamma->d());
    if (m > 1) k_axpy_dev(c, n, -1.0, dv.c2 + (m - 1), bases[m - 2]->d(), gamma->d());
    k_dot_to(c, n, gamma->d(), gamma->d(), dv.scal);
    lz_norm_kernel<<<1, 1, 0, c->stream>>>(dv.scal, dv.b + (m - 1), dv.c2 + m, dv.scal + 2);
    k_scale_dev(c, n, dv.scal + 2, gamma->d());
    bases.push_back(gamma);
    dv.set_ptr((int)m, gamma->d());
    last = ten_alloc(c, full, kF64, cz);
    heff_mul(L, W1, W2, R, bases[m], where, last, kF64);
    res.matvecs_hi++;
    k_dot_to(c, n, last->d(), bases[m]->d(), dv.a + m);
    tridiag_lowest_async(c, (int)m + 1, dv.a, dv.b, 0, dv.out, static_cast<double *>(dv.work->p));
    lz_pack_kernel<<<1, 64, 0, c->stream>>>(dv.b + (m - 1), 1, dv.scal + 3);
    lz_pack_kernel<<<1, 64, 0, c->stream>>>(dv.out, 1, dv.scal + 4);
    const double *h = dv.read(dv.scal + 3, 2);
    const double ng = h[0], e_new = h[1];
    if (ng == 0.0) {
      // gamma vanished: the Krylov space is invariant (the reference returns
      // the ground state of the m x m tridiagonal, lanczos_solver_impl.h:161-170)
      res.iters = m;
      res.energy = energy0;
      if (m == 1) {
        res.gs = ten_copy(*bases[0]);
      } else {
        tridiag_lowest_async(c, (int)m, dv.a, dv.b, 1, dv.out, static_cast<double *>(dv.work->p));
        res.gs = combine_dev(c, full, dv, (int)m, dv.out + 1, cz);
      }
      return res;
    }
    if ((energy0 - e_new) < error || m == eff_dim || m == max_iter - 1) {
      tridiag_lowest_async(c, (int)m + 1, dv.a, dv.b, 1, dv.out, static_cast<double *>(dv.work->p));
      res.iters = m;
      res.energy = e_new;
      res.gs = combine_dev(c, full, dv, (int)m + 1, dv.out + 1, cz);
      c->sync();
      return res;
    }
    energy0 = e_new;
  }
}

namespace {

TenP cast_to(const TenP &t, Prec p) {
  if (!t) return nullptr;
  if (t->prec == p) return t;
  auto r = ten_alloc(t->ctx, t->meta, p);
  if (p == kF32) k_cast_d2f(t->ctx, t->n(), t->d(), r->f());
  else k_cast_f2d(t->ctx, t->n(), t->f(), r->d());
  return r;
}

// Lowest eigenpair of a small dense symmetric matrix (row-major n x n, n of
// the order of the restart length): Householder reduction to tridiagonal
// form (A = Q T Q^T), the lowest eigenvalue of T by bisection on Sturm
// counts, its vector by two steps of inverse iteration (Gaussian elimination
// with partial pivoting on T - lambda I), then vec = Q z.  O(n^3) once
// instead of per Jacobi sweep; the eigenvalue is accurate to a few eps ||A||.
double small_sym_lowest(int n, std::vector<double> A, std::vector<double> &vec) {
  vec.assign(n, 0.0);
  if (n == 1) {
    vec[0] = 1.0;
    return A[0];
  }
  auto at = [&](int r, int c) -> double & { return A[(size_t)r * n + c]; };
  std::vector<std::vector<double>> refl;  // v_i (length n - i - 1, v[0] = 1), reflector i
  std::vector<double> taus;
  std::vector<double> p(n), w(n);
  for (int i = 0; i + 2 < n; ++i) {
    const int m = n - i - 1;
    const double alpha = at(i + 1, i);
    double xn2 = 0;
    for (int r = i + 2; r < n; ++r) xn2 += at(r, i) * at(r, i);
    std::vector<double> v(m, 0.0);
    v[0] = 1.0;
    if (xn2 == 0.0) {
      refl.push_back(v);
      taus.push_back(0.0);
      continue;
    }
    const double beta = -std::copysign(std::sqrt(alpha * alpha + xn2), alpha);
    const double tau = (beta - alpha) / beta, sc = 1.0 / (alpha - beta);
    for (int r = 1; r < m; ++r) v[r] = at(i + 1 + r, i) * sc;
    // A22 <- H A22 H, H = I - tau v v^T:  p = tau A22 v, w = p - (tau/2)(p.v) v
    double pv = 0;
    for (int r = 0; r < m; ++r) {
      double t = 0;
      for (int c = 0; c < m; ++c) t += at(i + 1 + r, i + 1 + c) * v[c];
      p[r] = tau * t;
      pv += p[r] * v[r];
    }
    const double K = 0.5 * tau * pv;
    for (int r = 0; r < m; ++r) w[r] = p[r] - K * v[r];
    for (int r = 0; r < m; ++r)
      for (int c = 0; c < m; ++c) at(i + 1 + r, i + 1 + c) -= v[r] * w[c] + w[r] * v[c];
    at(i + 1, i) = at(i, i + 1) = beta;
    for (int r = i + 2; r < n; ++r) at(r, i) = at(i, r) = 0.0;
    refl.push_back(v);
    taus.push_back(tau);
  }
  std::vector<double> d(n), e(n, 0.0);
  double tnorm = 0;
  for (int i = 0; i < n; ++i) {
    d[i] = at(i, i);
    if (i + 1 < n) e[i] = at(i + 1, i);
  }
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
    tnorm = std::max(tnorm, std::fabs(d[i]) + r);
  }
  const double eps = 2.220446049250313e-16;
  const double pivmin = std::max(tnorm * eps * eps, 1e-300);
  // number of eigenvalues of T below x
  auto below = [&](double x) {
    int cnt = 0;
    double q = d[0] - x;
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++cnt;
    for (int i = 1; i < n; ++i) {
      q = d[i] - x - e[i - 1] * e[i - 1] / q;
      if (std::fabs(q) < pivmin) q = -pivmin;
      if (q < 0) ++cnt;
    }
    return cnt;
  };
  lo -= 2 * eps * tnorm + pivmin;
  hi += 2 * eps * tnorm + pivmin;
  for (int it = 0; it < 200 && hi - lo > 2 * eps * std::max(std::fabs(lo), std::fabs(hi)) + pivmin; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (below(mid) >= 1) hi = mid;
    else lo = mid;
  }
  const double lam = 0.5 * (lo + hi);
  // inverse iteration on the tridiagonal T - lam I (dense, partial pivoting)
  std::vector<double> z(n, 1.0 / std::sqrt((double)n)), M((size_t)n * n);
  for (int step = 0; step < 2; ++step) {
    std::fill(M.begin(), M.end(), 0.0);
    for (int i = 0; i < n; ++i) {
      M[(size_t)i * n + i] = d[i] - lam;
      if (i + 1 < n) M[(size_t)i * n + i + 1] = M[(size_t)(i + 1) * n + i] = e[i];
    }
    std::vector<double> b = z;
    for (int col = 0; col < n; ++col) {
      int piv = col;
      for (int r = col + 1; r < n; ++r)
        if (std::fabs(M[(size_t)r * n + col]) > std::fabs(M[(size_t)piv * n + col])) piv = r;
      if (piv != col) {
        for (int c2 = 0; c2 < n; ++c2) std::swap(M[(size_t)col * n + c2], M[(size_t)piv * n + c2]);
        std::swap(b[col], b[piv]);
      }
      double &pv = M[(size_t)col * n + col];
      if (std::fabs(pv) < eps * tnorm) pv = pv < 0 ? -eps * tnorm : eps * tnorm;
      if (pv == 0.0) pv = pivmin;
      for (int r = col + 1; r < n; ++r) {
        const double f = M[(size_t)r * n + col] / pv;
        if (f == 0.0) continue;
        for (int c2 = col; c2 < n; ++c2) M[(size_t)r * n + c2] -= f * M[(size_t)col * n + c2];
        b[r] -= f * b[col];
      }
    }
    for (int r = n - 1; r >= 0; --r) {
      double t = b[r];
      for (int c2 = r + 1; c2 < n; ++c2) t -= M[(size_t)r * n + c2] * b[c2];
      b[r] = t / M[(size_t)r * n + r];
    }
    double nb = 0;
    for (double x : b) nb += x * x;
    nb = std::sqrt(nb);
    if (!(nb > 0) || !std::isfinite(nb)) break;
    for (int i = 0; i < n; ++i) z[i] = b[i] / nb;
  }
  // vec = H_0 H_1 ... H_{n-3} z
  for (int i = (int)refl.size() - 1; i >= 0; --i) {
    const std::vector<double> &v = refl[i];
    const int m = (int)v.size();
    double t = 0;
    for (int r = 0; r < m; ++r) t += v[r] * z[i + 1 + r];
    t *= taus[i];
    for (int r = 0; r < m; ++r) z[i + 1 + r] -= t * v[r];
  }
  vec = z;
  return lam;
}

}  // namespace

// Mixed-precision form: restarted Lanczos with FP64 residual correction.
// Each restart k computes, in FP64, h = H v_k, E_k = <v_k|h> and the residual
// r = h - E_k v_k.  The Krylov space {v_k, r, H~ r, ...} is then extended with
// the low-precision operator H~ (FP32 operands and accumulation; products on
// the FP32 matrix cores, or split-BF16x3 on the BF16 ones), every new
// vector orthogonalised in FP64 by the three-term step and then once more
// against the whole basis.  The projected
// matrix uses the exact FP64 row of v_k (E_k, |r|, 0, ...) and FP64 dot
// products of the FP32 images, so the error left by H~ enters the Ritz vector
// only at second order and each restart contracts it further: the result
// converges to the FP64 answer.  Stop when the projected solve lowers the
// energy by less than `error` (the reference's energy criterion, returning
// that Ritz pair), when the FP64 Rayleigh quotient decreases by less than
// `error` between restarts, or when the residual vanishes (returning v and
// its FP64 Rayleigh quotient).  A converged state thus costs one FP64 and one
// FP32 matvec, where the plain form spends two FP64 matvecs.
LanczosOut lanczos_mixed(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &init,
                         const std::string &where, double error, int64_t max_iter, bool split) {
  Ctx *c = init->ctx;
  MPS2_REQUIRE(max_iter >= 2, "Lanczos: max_iterations must be >= 2");
  MPS2_REQUIRE(!any_complex(L, W1, W2, R, init),
               "mixed-precision Lanczos: real tensors only (complex problems run MPS2_PREC_FP64)");
  int64_t eff_dim = eff_dim_of(L, W1, W2, R, where);
  MetaP full = full_meta(init->meta->qw, init->meta->legs, init->meta->div());
  TenP L32 = cast_to(L, kF32), W132 = cast_to(W1, kF32), W232 = cast_to(W2, kF32), R32 = cast_to(R, kF32);
  const int64_t n = full->total;
  const int kInnerMax = (int)std::min<int64_t>(std::max<int64_t>(max_iter, 2), 40);
  LzDev dv(c, kInnerMax + 2);
  double *d_col = dv.a;   // orthogonalisation coefficients, norm and alpha (device)
  TenP v = embed(init, full);
  k_dot_to(c, n, v->d(), v->d(), dv.scal);
  lz_rsqrt_kernel<<<1, 1, 0, c->stream>>>(dv.scal, dv.scal + 1, dv.scal + 2);
  k_scale_dev(c, n, dv.scal + 2, v->d());
  TenP h = ten_alloc(c, full);
  TenP x32 = ten_alloc(c, full, kF32), y32 = ten_alloc(c, full, kF32);
  LanczosOut res;
  double E_prev = 0;
  int64_t total_inner = 0;
  for (int restart = 0; restart < 64; ++restart) {
    heff_mul(L, W1, W2, R, v, where, h, kF64);
    res.matvecs_hi++;
    // E = <h, v>, r = h - E v (in h), beta = |r|: one read for both
    k_dot_to(c, n, h->d(), v->d(), dv.scal + 3);
    k_axpy_dev(c, n, -1.0, dv.scal + 3, v->d(), h->d());
    k_dot_to(c, n, h->d(), h->d(), dv.scal + 4);
    const double *hs = dv.read(dv.scal + 3, 2);
    const double E = hs[0], beta = std::sqrt(hs[1]);
    bool done = (restart > 0 && (E_prev - E) < error) || beta == 0.0 || total_inner >= max_iter - 1 ||
                (eff_dim <= 1);
    if (done) {
      res.iters = std::max<int64_t>(total_inner, 1);
      res.energy = E;
      res.gs = v;
      return res;
    }
    E_prev = E;
    std::vector<TenP> Q{v};
    dv.set_ptr(0, v->d());
    auto q1 = ten_alloc(c, full);
    MPS2_HIP(hipMemcpyAsync(q1->d(), h->d(), n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    k_scale(c, n, 1.0 / beta, q1->d());
    Q.push_back(q1);
    dv.set_ptr(1, q1->d());
    // projected matrix, grown column by column (kept as a full symmetric matrix)
    std::vector<std::vector<double>> T(2, std::vector<double>(2, 0.0));
    T[0][0] = E;
    T[0][1] = T[1][0] = beta;
    std::vector<double> y;
    double theta_prev = E;
    double wprev = beta;  // norm that made q_j: beta for q_1, then the previous step's
    int m_in = 0;
    for (int j = 1; j <= kInnerMax; ++j) {
      // w = H~ q_j
      k_cast_d2f(c, n, Q[j]->d(), x32->f());
      {
        SplitScope scope(c, split);
        heff_mul(L32, W132, W232, R32, x32, where, y32, kF32);
      }
      res.matvecs_lo++;
      auto w = ten_alloc(c, full);
      k_cast_f2d(c, n, y32->f(), w->d());
      ++total_inner;
      const int k = (int)Q.size();  // j + 1
      // Orthogonalisation in FP64, two passes: the three-term step removes
      // the large components (alpha_j = <q_j, w> along q_j, the previous
      // norm beta_j along q_{j-1}), then one classical Gram-Schmidt pass
      // against all of Q[0..j] removes what the low-precision operator and
      // the rounding of that step leave along every q_i.  The projected
      // matrix's new column is <Q_i, w> = the second pass's coefficient plus
      // alpha_j (i = j) or beta_j (i = j - 1).
      const double bprev = wprev;
      double *d_alpha = d_col + k + 1;
      k_dot_to(c, n, Q[j]->d(), w->d(), d_alpha);
      k_axpy_dev(c, n, -1.0, d_alpha, Q[j]->d(), w->d());
      k_axpby(c, n, -bprev, Q[j - 1]->d(), 1.0, w->d());
      k_multidot(c, n, k, dv.ptrs(), w->d(), d_col, static_cast<double *>(dv.partial->p));
      k_multiaxpy(c, n, k, dv.ptrs(), d_col, w->d());
      k_dot_to(c, n, w->d(), w->d(), d_col + k);
      const double *hc = dv.read(d_col, k + 2);
      std::vector<double> col(hc, hc + k);
      const double wn = std::sqrt(hc[k]);
      col[j] += hc[k + 1];
      col[j - 1] += bprev;
      col[0] = (j == 1) ? beta : 0.0;
      for (int i = 1; i <= j; ++i) {
        T[i][j] = col[i];
        T[j][i] = col[i];
      }
      int kk = j + 1;
      std::vector<double> A(kk * kk);
      for (int a = 0; a < kk; ++a)
        for (int b = 0; b < kk; ++b) A[a * kk + b] = T[a][b];
      double theta = small_sym_lowest(kk, A, y);
      m_in = j;
      // Stagnation test of the plain form: the Ritz value moved by less than `error`.
      bool inner_done = std::fabs(theta_prev - theta) < error || wn <= 1e-14 * std::fabs(theta) ||
                        (int64_t)Q.size() >= eff_dim || j == kInnerMax || total_inner >= max_iter - 1;
      theta_prev = theta;
      if (inner_done) break;
      k_scale(c, n, 1.0 / wn, w->d());
      wprev = wn;
      Q.push_back(w);
      dv.set_ptr(k, w->d());
      for (auto &row : T) row.push_back(0.0);
      T.push_back(std::vector<double>(T[0].size(), 0.0));
    }
    // v <- sum_i y_i Q_i  (y has m_in + 1 entries: Q[0..m_in]), normalised
    const int nv = m_in + 1;
    CoefArg ca{};
    ca.n = nv;
    for (int i = 0; i < nv; ++i) ca.v[i] = y[i];
    lz_set_coef_kernel<<<1, 64, 0, c->stream>>>(ca, dv.c2);
    v = combine_dev(c, full, dv, nv, dv.c2);
    k_dot_to(c, n, v->d(), v->d(), dv.scal);
    lz_rsqrt_kernel<<<1, 1, 0, c->stream>>>(dv.scal, dv.scal + 1, dv.scal + 2);
    k_scale_dev(c, n, dv.scal + 2, v->d());
    // The projected solve lowers the energy by less than `error`: the
    // reference's stopping rule (lanczos_solver_impl.h:208-221) holds for this
    // Krylov space, so return its Ritz pair.  Row and column 0 of the
    // projected matrix are exact FP64 (E, |r|); the FP32 entries enter theta
    // weighted by the squared Ritz coefficients of Q[1..], each below
    // (E - theta) / gap, so theta carries their error only at second order.
    // theta <= E by Cauchy interlacing.
    if ((E - theta_prev) < error) {
      res.iters = std::max<int64_t>(total_inner, 1);
      res.energy = theta_prev;
      res.gs = v;
      return res;
    }
  }
  res.iters = total_inner;
  res.energy = E_prev;
  res.gs = v;
  return res;
}

LanczosOut lanczos(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &init,
                   const std::string &where, double error, int64_t max_iter, Arith arith) {
  Ctx *c = init->ctx;
  LanczosOut r = arith == kArithF64 ? lanczos_plain(L, W1, W2, R, init, where, error, max_iter)
                                    : lanczos_mixed(L, W1, W2, R, init, where, error, max_iter,
                                                    arith == kArithBf16x3);
  c->count("lanczos_calls");
  c->count("matvec_fp64", (double)r.matvecs_hi);
  c->count("matvec_fp32", (double)r.matvecs_lo);
  if (!std::isfinite(r.energy)) throw Error(kErrNumeric, "Lanczos: non-finite energy (environment or state holds NaN or Inf)");
  return r;
}

}  // namespace mps2
