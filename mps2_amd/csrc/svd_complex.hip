// Truncated SVD of a complex (GQTEN_Complex) block-sparse tensor on the
// device, the complex counterpart of svd.hip (gqten's SVD as called by
// TwoSiteFiniteVMPSUpdate, two_site_update_finite_vmps_impl.h:258-263, and
// by FiniteMPS canonicalisation, finite_mps.h:172-243).
//
// Per new-bond QN sector q the state is a complex matrix M_q (rows: left-leg
// tuples, columns: right-leg tuples, from blocks holding a nonzero element,
// in lexicographic tuple order), assembled scaled by a power of two per
// sector exactly as the real path does.  Each sector is diagonalised by the
// one-sided (Hestenes) Jacobi method on the rows of a matrix X: X = M_q when
// the left factor must come out orthonormal, X = M_q^H when the right one
// must.  Unitary plane rotations Q make the rows of X' = Q X mutually
// orthogonal; their norms are the singular values, Q is accumulated from the
// identity.  Left: U = Q^H, S Vt = X' (so S Vt = U^H M exactly).  Right:
// Vt = Q, U S = X'^H (= M V).  One-sided Jacobi resolves every singular value
// to a relative accuracy set by its own size (no Gram matrix is formed), so
// no tail refinement is needed; the work is O(sweeps r^2 c) per sector with
// r - 1 dependent rounds per sweep, which suits the small complex models the
// reference tests (bond dimensions up to 60) rather than D in the thousands.
//
// Truncation, new-bond order and the treatment of values below
// kSvdResolve * s_max are the real path's rule (svd_truncate_rule).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>

#include "runtime.h"

namespace mps2 {

namespace {

struct CJacD {
  int64_t x_off, q_off, s_off;   // complex elements (X, Q), doubles (S)
  int32_t r, cc;                 // rows of X (and of Q), columns of X
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One workgroup (kJacThreads threads) per sector.  Round-robin (circle
// method) ordering: in each of the P - 1 rounds of a sweep (P = r rounded up
// to even) the P / 2 disjoint row pairs are rotated concurrently, one
// wavefront per pair; a barrier separates the rounds.  A pair is rotated when
// |<x_p, x_q>| > tol sqrt(|x_p|^2 |x_q|^2).  Sweeps repeat until a sweep
// rotates nothing, at most max_sweeps.  Then S[s_off + i] = |x'_i|.
constexpr int kJacThreads = 1024;

__global__ __launch_bounds__(kJacThreads) void cjacobi_kernel(double *X, double *Q, double *S, const CJacD *desc,
                                                              double tol, int max_sweeps, int *sweeps) {
  const CJacD d = desc[blockIdx.x];
  const int r = d.r, cc = d.cc;
  double *x = X + 2 * d.x_off;
  double *q = Q + 2 * d.q_off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int64_t e = threadIdx.x; e < (int64_t)r * r; e += blockDim.x) {
    const int64_t i = e / r, j = e - i * r;
    q[2 * e] = i == j ? 1.0 : 0.0;
    q[2 * e + 1] = 0.0;
  }
  __shared__ int rotated;
  const int P = r + (r & 1);
  int sweep = 0;
  __syncthreads();
  if (r >= 2) {
    for (; sweep < max_sweeps; ++sweep) {
      if (threadIdx.x == 0) rotated = 0;
      __syncthreads();
      for (int t = 0; t < P - 1; ++t) {
        for (int k = wave; k < P / 2; k += nw) {
          int a, b;
          if (k == 0) {
            a = P - 1;
            b = t;
          } else {
            a = (t + k) % (P - 1);
            b = (t - k + P - 1) % (P - 1);
          }
          if (a >= r || b >= r) continue;
          const int ip = min(a, b), iq = max(a, b);
          double *xp = x + 2 * (int64_t)ip * cc, *xq = x + 2 * (int64_t)iq * cc;
          double al = 0, be = 0, gr = 0, gi = 0;
          for (int j = lane; j < cc; j += 64) {
            const double pr = xp[2 * j], pi = xp[2 * j + 1], qr = xq[2 * j], qi = xq[2 * j + 1];
            al += pr * pr + pi * pi;
            be += qr * qr + qi * qi;
            gr += pr * qr + pi * qi;   // gamma = sum x_p conj(x_q)
            gi += pi * qr - pr * qi;
          }
          al = wave_sum(al);
          be = wave_sum(be);
          gr = wave_sum(gr);
          gi = wave_sum(gi);
          const double g = sqrt(gr * gr + gi * gi);
          if (!(g > tol * sqrt(al * be))) continue;
          // real Jacobi rotation of (x_p, e x_q), e = gamma / |gamma|, whose
          // Gram matrix [[al, g], [g, be]] is real; x_q' takes back the phase
          const double zeta = (be - al) / (2.0 * g);
          const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
          const double er = gr / g, ei = gi / g;
          auto rot = [&](double *u, double *v, int len) {
            for (int j = lane; j < len; j += 64) {
              const double ur = u[2 * j], ui = u[2 * j + 1], vr = v[2 * j], vi = v[2 * j + 1];
              const double evr = er * vr - ei * vi, evi = er * vi + ei * vr;       // e v
              const double cur = er * ur + ei * ui, cui = er * ui - ei * ur;       // conj(e) u
              u[2 * j] = c * ur - s * evr;
              u[2 * j + 1] = c * ui - s * evi;
              v[2 * j] = s * cur + c * vr;
              v[2 * j + 1] = s * cui + c * vi;
            }
          };
          rot(xp, xq, cc);
          rot(q + 2 * (int64_t)ip * r, q + 2 * (int64_t)iq * r, r);
          if (lane == 0) rotated = 1;
        }
        __syncthreads();
      }
      const int any = rotated;
      __syncthreads();
      if (!any) break;
    }
  }
  for (int i = wave; i < r; i += nw) {
    const double *xi = x + 2 * (int64_t)i * cc;
    double ss = 0;
    for (int j = lane; j < cc; j += 64) ss += xi[2 * j] * xi[2 * j] + xi[2 * j + 1] * xi[2 * j + 1];
    ss = wave_sum(ss);
    if (lane == 0) S[d.s_off + i] = sqrt(ss);
  }
  if (threadIdx.x == 0 && sweeps) sweeps[blockIdx.x] = sweep;
}

// Factor assembly, one descriptor per blockIdx.y (complex element units):
//   mode 0: dst[i][j] = src[sel(i)][j] * scl(i)          (rows x cols)
//   mode 1: dst[i][j] = conj(src[sel(j)][i]) * scl(j)    (rows x cols)
// sel(x) = sel[sel_off + x] (x when sel_off < 0), scl likewise (1 when < 0).
struct CFacD {
  int64_t src_off, dst_off;
  int32_t rows, cols, lds, ldd;
  int32_t mode, pad;
  int64_t sel_off, scl_off;
};

__global__ void cfac_kernel(const double *src, double *dst, const CFacD *desc, const int32_t *sel,
                            const double *scl) {
  const CFacD d = desc[blockIdx.y];
  const int64_t tot = (int64_t)d.rows * d.cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / d.cols, j = e - i * d.cols;
    const int64_t x = d.mode == 0 ? i : j, y = d.mode == 0 ? j : i;
    const int64_t sr = d.sel_off >= 0 ? sel[d.sel_off + x] : x;
    const double f = d.scl_off >= 0 ? scl[d.scl_off + x] : 1.0;
    const int64_t so = 2 * (d.src_off + sr * d.lds + y);
    const double re = src[so], im = src[so + 1];
    const int64_t o = 2 * (d.dst_off + i * d.ldd + j);
    dst[o] = re * f;
    dst[o + 1] = (d.mode == 0 ? im : -im) * f;
  }
}

BufP upload_c(Ctx *c, const void *h, size_t bytes) { return c->upload(h, bytes); }

void run_cfac(Ctx *c, const double *src, double *dst, const std::vector<CFacD> &d, const std::vector<int32_t> &sel,
              const std::vector<double> &scl, int64_t max_elems) {
  if (d.empty()) return;
  auto dd = upload_c(c, d.data(), d.size() * sizeof(CFacD));
  // (descriptors with sel_off / scl_off < 0 read neither table)
  auto ds = upload_c(c, sel.data(), sel.size() * sizeof(int32_t));
  auto dscl = upload_c(c, scl.data(), scl.size() * sizeof(double));
  const int gx = (int)std::min<int64_t>(std::max<int64_t>((max_elems + 255) / 256, 1), 1024);
  cfac_kernel<<<dim3(gx, (unsigned)d.size()), 256, 0, c->stream>>>(src, dst, static_cast<const CFacD *>(dd->p),
                                                                    static_cast<const int32_t *>(ds->p),
                                                                    static_cast<const double *>(dscl->p));
}

}  // namespace

SvdOut svd_complex(const TenP &t, int ldims, const QN &ldiv, double trunc_err, int64_t dmin, int64_t dmax,
                   OrthSide side) {
  MPS2_REQUIRE(t->prec == kF64 && t->cplx, "complex SVD: complex FP64 tensor expected");
  Ctx *c = t->ctx;
  const Meta &m = *t->meta;
  const int r = m.rank();
  MPS2_REQUIRE(ldims >= 1 && ldims < r, "SVD: ldims out of range");
  const int nblk = (int)m.coords.size();
  // 1. nonzero blocks (largest |re|, |im| per block, from the real form)
  std::vector<double> bmax(nblk, 0.0);
  if (nblk) {
    std::vector<int64_t> off2(nblk), sz2(nblk);
    for (int b = 0; b < nblk; ++b) off2[b] = 2 * m.offset[b], sz2[b] = 2 * m.bsize[b];
    auto doff = upload_c(c, off2.data(), nblk * sizeof(int64_t));
    auto dsz = upload_c(c, sz2.data(), nblk * sizeof(int64_t));
    auto dout = c->alloc(nblk * sizeof(double));
    k_block_maxabs(c, t->d(), static_cast<const int64_t *>(doff->p), static_cast<const int64_t *>(dsz->p), nblk,
                   static_cast<double *>(dout->p));
    MPS2_HIP(hipMemcpyAsync(bmax.data(), dout->p, nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->sync();
  }
  double gmax = 0.0;
  for (int b = 0; b < nblk; ++b) {
    if (!std::isfinite(bmax[b])) throw Error(kErrNumeric, "SVD: input tensor holds NaN or Inf");
    gmax = std::max(gmax, bmax[b]);
  }
  const int sexp = gmax > 0.0 ? std::max(-1000, std::min(1000, std::ilogb(gmax))) : 0;
  const double pre = std::ldexp(1.0, -sexp), post = std::ldexp(1.0, sexp);
  // 2. sector structure (as the real path)
  struct Sec {
    QN qn;
    std::map<Coord, int64_t> rows, cols;
    int64_t m = 0, n = 0;
  };
  std::map<QN, Sec> smap;
  auto kqn = [&](const Coord &cc) {
    QN q = ldiv;
    for (int i = 0; i < ldims; ++i) q = m.legs[i].dir == kOut ? q - m.legs[i].qn[cc[i]] : q + m.legs[i].qn[cc[i]];
    return q;
  };
  for (int b = 0; b < nblk; ++b) {
    if (!(bmax[b] > 0.0)) continue;
    const Coord &cc = m.coords[b];
    Sec &s = smap[kqn(cc)];
    s.qn = kqn(cc);
    s.rows[Coord(cc.begin(), cc.begin() + ldims)] = 0;
    s.cols[Coord(cc.begin() + ldims, cc.end())] = 0;
  }
  std::vector<Sec> secs;
  for (auto &kv : smap) secs.push_back(kv.second);
  const int ns = (int)secs.size();
  std::vector<int64_t> moff(ns);
  int64_t mtot = 0;
  for (int q = 0; q < ns; ++q) {
    auto &s = secs[q];
    for (auto &kv : s.rows) {
      kv.second = s.m;
      int64_t d = 1;
      for (int i = 0; i < ldims; ++i) d *= m.legs[i].dim[kv.first[i]];
      s.m += d;
    }
    for (auto &kv : s.cols) {
      kv.second = s.n;
      int64_t d = 1;
      for (int i = ldims; i < r; ++i) d *= m.legs[i].dim[kv.first[i - ldims]];
      s.n += d;
    }
    moff[q] = mtot;
    mtot += s.m * s.n;
  }
  std::vector<int> dq(ns, 0);
  {
    std::vector<double> smax(ns, 0.0);
    for (int b = 0; b < nblk; ++b) {
      if (!(bmax[b] > 0.0)) continue;
      const int q = (int)std::distance(smap.begin(), smap.find(kqn(m.coords[b])));
      smax[q] = std::max(smax[q], bmax[b]);
    }
    for (int q = 0; q < ns; ++q)
      dq[q] = smax[q] > 0.0 ? std::max(-1000, std::min(1000, std::ilogb(smax[q]))) - sexp : 0;
  }
  // 3. sector matrices (complex, m x n row-major), scaled by 2^-(sexp + dq)
  auto Mb = c->alloc(std::max<int64_t>(2 * mtot, 1) * sizeof(double));
  double *Md = static_cast<double *>(Mb->p);
  k_fill(c, Md, std::max<int64_t>(2 * mtot, 1), 0.0);
  {
    std::vector<Copy2D> cp;
    for (int b = 0; b < nblk; ++b) {
      if (!(bmax[b] > 0.0)) continue;
      const Coord &cc = m.coords[b];
      const int q = (int)std::distance(smap.begin(), smap.find(kqn(cc)));
      auto &s = secs[q];
      const int64_t ro = s.rows.at(Coord(cc.begin(), cc.begin() + ldims));
      const int64_t co = s.cols.at(Coord(cc.begin() + ldims, cc.end()));
      int64_t bm = 1, bn = 1;
      for (int i = 0; i < ldims; ++i) bm *= m.legs[i].dim[cc[i]];
      for (int i = ldims; i < r; ++i) bn *= m.legs[i].dim[cc[i]];
      cp.push_back(Copy2D{2 * m.offset[b], 2 * (moff[q] + ro * s.n + co), (int32_t)bm, (int32_t)(2 * bn),
                          (int32_t)(2 * bn), (int32_t)(2 * s.n), std::ldexp(pre, -dq[q])});
    }
    if (!cp.empty()) {
      auto dcp = upload_c(c, cp.data(), cp.size() * sizeof(Copy2D));
      k_copy2d(c, t->d(), Md, static_cast<const Copy2D *>(dcp->p), (int)cp.size());
      c->sync();
    }
  }
  // 4. X per sector: M (left) or M^H (right); Q and the row norms
  const bool left = side == kOrthLeft;
  std::vector<int> xr(ns), xc(ns);
  std::vector<int64_t> xoff(ns), qoff(ns), soff(ns);
  int64_t xtot = 0, qtot = 0, stot = 0, xmax = 0;
  for (int q = 0; q < ns; ++q) {
    xr[q] = (int)(left ? secs[q].m : secs[q].n);
    xc[q] = (int)(left ? secs[q].n : secs[q].m);
    xoff[q] = xtot;
    xtot += (int64_t)xr[q] * xc[q];
    qoff[q] = qtot;
    qtot += (int64_t)xr[q] * xr[q];
    soff[q] = stot;
    stot += xr[q];
    xmax = std::max<int64_t>(xmax, (int64_t)xr[q] * xc[q]);
  }
  BufP Xb;
  double *Xd = Md;
  if (!left) {
    Xb = c->alloc(std::max<int64_t>(2 * xtot, 1) * sizeof(double));
    Xd = static_cast<double *>(Xb->p);
    std::vector<CFacD> fd;
    for (int q = 0; q < ns; ++q)
      fd.push_back(CFacD{moff[q], xoff[q], xr[q], xc[q], (int32_t)secs[q].n, xc[q], 1, 0, -1, -1});
    run_cfac(c, Md, Xd, fd, {}, {}, xmax);
  }
  auto Qb = c->alloc(std::max<int64_t>(2 * qtot, 1) * sizeof(double));
  auto Sb = c->alloc(std::max<int64_t>(stot, 1) * sizeof(double));
  double *Qd = static_cast<double *>(Qb->p), *Sd = static_cast<double *>(Sb->p);
  auto te0 = std::chrono::steady_clock::now();
  {
    std::vector<CJacD> jd;
    for (int q = 0; q < ns; ++q) jd.push_back(CJacD{xoff[q], qoff[q], soff[q], xr[q], xc[q]});
    if (!jd.empty()) {
      auto djd = upload_c(c, jd.data(), jd.size() * sizeof(CJacD));
      auto dsw = c->alloc(jd.size() * sizeof(int));
      const int maxc = *std::max_element(xc.begin(), xc.end());
      const double tol = 2.3e-16 * std::sqrt((double)std::max(maxc, 1)) * 4.0;
      cjacobi_kernel<<<(unsigned)jd.size(), kJacThreads, 0, c->stream>>>(Xd, Qd, Sd, static_cast<const CJacD *>(djd->p),
                                                                          tol, 60, static_cast<int *>(dsw->p));
      std::vector<int> sw(jd.size());
      MPS2_HIP(hipMemcpyAsync(sw.data(), dsw->p, sw.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      for (int x : sw) c->count("svd_cplx_sweeps", x);
    }
  }
  c->count("svd_eigh_ms", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - te0).count());
  std::vector<double> sh(std::max<int64_t>(stot, 1));
  MPS2_HIP(hipMemcpyAsync(sh.data(), Sd, stot * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  // 5. singular values per sector, descending (host units: 2^dq), and the
  //    values the truncation rule sees (below kSvdResolve s_max: tied)
  std::vector<std::vector<int32_t>> order(ns);
  std::vector<std::vector<double>> sv(ns), svr(ns);
  double top = 0;
  for (int q = 0; q < ns; ++q)
    for (int i = 0; i < xr[q]; ++i) top = std::max(top, std::ldexp(sh[soff[q] + i], dq[q]));
  const double thr = kSvdResolve * top;
  for (int q = 0; q < ns; ++q) {
    std::vector<int32_t> idx(xr[q]);
    for (int i = 0; i < xr[q]; ++i) idx[i] = i;
    auto val = [&](int32_t i) { return std::ldexp(sh[soff[q] + i], dq[q]); };
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return val(a) > val(b); });
    const int64_t rq = std::min(secs[q].m, secs[q].n);
    idx.resize(rq);
    order[q] = idx;
    for (int64_t j = 0; j < rq; ++j) {
      const double v = val(idx[j]);
      sv[q].push_back(v);
      svr[q].push_back(v > 0.0 ? std::max(v, thr) : 0.0);
    }
  }
  SvdOut out;
  svd_truncate_rule(svr, trunc_err, dmin, dmax, out.kept, &out.trunc_err, &out.D);
  // 6. new bond index and the output layouts (as the real path)
  Leg knew;
  knew.dir = kOut;
  std::vector<int> secpos(ns, -1);
  for (int q = 0; q < ns; ++q)
    if (out.kept[q] > 0) {
      secpos[q] = (int)knew.qn.size();
      knew.qn.push_back(secs[q].qn);
      knew.dim.push_back(out.kept[q]);
    }
  auto um = std::make_shared<Meta>(), sm = std::make_shared<Meta>(), vm = std::make_shared<Meta>();
  um->qw = sm->qw = vm->qw = m.qw;
  for (int i = 0; i < ldims; ++i) um->legs.push_back(m.legs[i]);
  um->legs.push_back(knew);
  sm->legs = {knew.inverse(), knew};
  vm->legs.push_back(knew.inverse());
  for (int i = ldims; i < r; ++i) vm->legs.push_back(m.legs[i]);
  for (int q = 0; q < ns; ++q) {
    if (secpos[q] < 0) continue;
    for (auto &kv : secs[q].rows) {
      Coord cc = kv.first;
      cc.push_back(secpos[q]);
      um->coords.push_back(cc);
    }
    for (auto &kv : secs[q].cols) {
      Coord cc{secpos[q]};
      cc.insert(cc.end(), kv.first.begin(), kv.first.end());
      vm->coords.push_back(cc);
    }
    sm->coords.push_back(Coord{secpos[q], secpos[q]});
  }
  um->finalize();
  sm->finalize();
  vm->finalize();
  out.u = ten_zeros(c, um, kF64, true);
  out.vt = ten_zeros(c, vm, kF64, true);
  {
    std::vector<double> s(sm->total, 0.0);
    for (int q = 0; q < ns; ++q) {
      if (secpos[q] < 0) continue;
      const int32_t b = sm->find(Coord{secpos[q], secpos[q]});
      const int64_t kq = out.kept[q];
      for (int64_t k = 0; k < kq; ++k) s[sm->offset[b] + k * kq + k] = sv[q][k] * post;
    }
    out.s = ten_upload(c, sm, s.data());
  }
  // 7. dense factors per kept sector: Ud (m x kq), Vd (kq x n)
  std::vector<int64_t> uoff(ns, 0), voff(ns, 0);
  int64_t utot = 0, vtot = 0, fmax = 0;
  std::vector<int32_t> sel;
  std::vector<double> inv;
  std::vector<int64_t> seloff(ns, 0);
  for (int q = 0; q < ns; ++q) {
    if (secpos[q] < 0) continue;
    const int64_t kq = out.kept[q];
    uoff[q] = utot;
    utot += secs[q].m * kq;
    voff[q] = vtot;
    vtot += kq * secs[q].n;
    fmax = std::max<int64_t>(fmax, std::max(secs[q].m, secs[q].n) * kq);
    seloff[q] = (int64_t)sel.size();
    for (int64_t k = 0; k < kq; ++k) {
      sel.push_back(order[q][k]);
      inv.push_back(sv[q][k] > 0 ? std::ldexp(1.0 / sv[q][k], dq[q]) : 0.0);
    }
  }
  auto Ub = c->alloc(std::max<int64_t>(2 * utot, 1) * sizeof(double));
  auto Vb = c->alloc(std::max<int64_t>(2 * vtot, 1) * sizeof(double));
  double *Ud = static_cast<double *>(Ub->p), *Vd = static_cast<double *>(Vb->p);
  {
    std::vector<CFacD> fu, fv;   // sources: Q for the orthonormal side, X' for the other
    for (int q = 0; q < ns; ++q) {
      if (secpos[q] < 0) continue;
      const int32_t kq = (int32_t)out.kept[q], mq = (int32_t)secs[q].m, nq = (int32_t)secs[q].n;
      if (left) {
        // U = Q^H: U[i][j] = conj(Q[sel j][i]); Vt = X'[sel j] / s_j
        fu.push_back(CFacD{qoff[q], uoff[q], mq, kq, xr[q], kq, 1, 0, seloff[q], -1});
        fv.push_back(CFacD{xoff[q], voff[q], kq, nq, xc[q], nq, 0, 0, seloff[q], seloff[q]});
      } else {
        // Vt = Q[sel j]; U[i][j] = conj(X'[sel j][i]) / s_j
        fv.push_back(CFacD{qoff[q], voff[q], kq, nq, xr[q], nq, 0, 0, seloff[q], -1});
        fu.push_back(CFacD{xoff[q], uoff[q], mq, kq, xc[q], kq, 1, 0, seloff[q], seloff[q]});
      }
    }
    run_cfac(c, left ? Qd : Xd, Ud, fu, sel, inv, fmax);
    run_cfac(c, left ? Xd : Qd, Vd, fv, sel, inv, fmax);
  }
  // 8. scatter into the U and Vt blocks (real form: two doubles per element)
  std::vector<Copy2D> cpu, cpv;
  for (int q = 0; q < ns; ++q) {
    if (secpos[q] < 0) continue;
    auto &s = secs[q];
    const int64_t kq = out.kept[q];
    for (auto &kv : s.rows) {
      Coord cc = kv.first;
      cc.push_back(secpos[q]);
      const int32_t b = um->find(cc);
      const int64_t bm = um->bsize[b] / kq;
      cpu.push_back(Copy2D{2 * (uoff[q] + kv.second * kq), 2 * um->offset[b], (int32_t)bm, (int32_t)(2 * kq),
                           (int32_t)(2 * kq), (int32_t)(2 * kq), 1.0});
    }
    for (auto &kv : s.cols) {
      Coord cc{secpos[q]};
      cc.insert(cc.end(), kv.first.begin(), kv.first.end());
      const int32_t b = vm->find(cc);
      const int64_t bn = vm->bsize[b] / kq;
      cpv.push_back(Copy2D{2 * (voff[q] + kv.second), 2 * vm->offset[b], (int32_t)kq, (int32_t)(2 * bn),
                           (int32_t)(2 * s.n), (int32_t)(2 * bn), 1.0});
    }
  }
  if (!cpu.empty()) {
    auto d = upload_c(c, cpu.data(), cpu.size() * sizeof(Copy2D));
    k_copy2d(c, Ud, out.u->d(), static_cast<const Copy2D *>(d->p), (int)cpu.size());
    c->sync();
  }
  if (!cpv.empty()) {
    auto d = upload_c(c, cpv.data(), cpv.size() * sizeof(Copy2D));
    k_copy2d(c, Vd, out.vt->d(), static_cast<const Copy2D *>(d->p), (int)cpv.size());
    c->sync();
  }
  return out;
}

}  // namespace mps2
