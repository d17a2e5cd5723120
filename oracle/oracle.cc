// CPU FP64 oracle for the two-site finite vMPS hot path.
//
// TEST INFRASTRUCTURE ONLY: this is the checker the HIP product path is
// compared against.  Nothing in mps2_amd/ links or calls it.  Each function
// names the reference file:line whose behaviour it restates (paths relative to
// the GraceQ/MPS2 0.2-alpha.0 tree, headers under include/gqmps2/).
//
// Own numerics (no BLAS/LAPACK on the box): blocked OpenMP DGEMM, one-sided
// (Hestenes) Jacobi SVD, implicit-QL symmetric tridiagonal eigensolver.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

using Coord = std::vector<int>;
using QNv = std::vector<int>;

struct Leg {
  int dir = 0;                  // 0 IN, 1 OUT
  std::vector<QNv> qn;          // per sector
  std::vector<int64_t> dim;     // per sector
  bool operator==(const Leg &o) const { return dir == o.dir && qn == o.qn && dim == o.dim; }
  int64_t total() const {
    int64_t t = 0;
    for (auto d : dim) t += d;
    return t;
  }
};

// A complex tensor (gqten's GQTEN_Complex) stores every element as two
// doubles (real, imaginary), interleaved: its block vectors are twice the
// block size.
struct Ten {
  int qw = 1;
  bool cplx = false;
  std::vector<Leg> legs;
  std::map<Coord, std::vector<double>> blk;   // sorted lexicographically
  int rank() const { return (int)legs.size(); }
  int64_t blk_size(const Coord &c) const {
    int64_t s = 1;
    for (int i = 0; i < rank(); ++i) s *= legs[i].dim[c[i]];
    return s;
  }
};

Ten *T(oracle_ten h) { return reinterpret_cast<Ten *>(h); }

QNv qn_zero(int qw) { return QNv(qw, 0); }
QNv qn_add(const QNv &a, const QNv &b, int sb = 1) {
  QNv r(a);
  for (size_t i = 0; i < r.size(); ++i) r[i] += sb * b[i];
  return r;
}

// ---------------------------------------------------------------- GEMM
// C[M,N] += A[M,K] * B[K,N], all row-major, contiguous.
void gemm_acc(int64_t M, int64_t N, int64_t K, const double *A, const double *B, double *C) {
  if (M == 0 || N == 0 || K == 0) return;
  const int64_t MB = 32, KB = 256;
  int64_t nmb = (M + MB - 1) / MB;
  int64_t work = M * N * K;
#pragma omp parallel for schedule(dynamic) if (work > 2000000)
  for (int64_t ib = 0; ib < nmb; ++ib) {
    int64_t i0 = ib * MB, i1 = std::min(M, i0 + MB);
    for (int64_t k0 = 0; k0 < K; k0 += KB) {
      int64_t k1 = std::min(K, k0 + KB);
      for (int64_t i = i0; i < i1; ++i) {
        double *c = C + i * N;
        const double *a = A + i * K;
        int64_t k = k0;
        for (; k + 4 <= k1; k += 4) {
          const double a0 = a[k], a1 = a[k + 1], a2 = a[k + 2], a3 = a[k + 3];
          const double *b0 = B + k * N, *b1 = b0 + N, *b2 = b1 + N, *b3 = b2 + N;
#pragma omp simd
          for (int64_t j = 0; j < N; ++j) c[j] += a0 * b0[j] + a1 * b1[j] + a2 * b2[j] + a3 * b3[j];
        }
        for (; k < k1; ++k) {
          const double a0 = a[k];
          const double *b0 = B + k * N;
#pragma omp simd
          for (int64_t j = 0; j < N; ++j) c[j] += a0 * b0[j];
        }
      }
    }
  }
}

// Permute a dense row-major block: dst[perm-ordered] = src.
// dst axis i is src axis perm[i].
void permute(const double *src, const std::vector<int64_t> &dims, const std::vector<int> &perm,
             double *dst) {
  int r = (int)dims.size();
  int64_t n = 1;
  for (auto d : dims) n *= d;
  if (n == 0) return;
  bool ident = true;
  for (int i = 0; i < r; ++i) ident &= (perm[i] == i);
  if (ident) {
    std::memcpy(dst, src, n * sizeof(double));
    return;
  }
  std::vector<int64_t> sstride(r, 1);
  for (int i = r - 2; i >= 0; --i) sstride[i] = sstride[i + 1] * dims[i + 1];
  std::vector<int64_t> ddims(r), dstr(r);
  for (int i = 0; i < r; ++i) {
    ddims[i] = dims[perm[i]];
    dstr[i] = sstride[perm[i]];
  }
  std::vector<int64_t> idx(r, 0);
  int64_t soff = 0;
  for (int64_t t = 0; t < n; ++t) {
    dst[t] = src[soff];
    for (int i = r - 1; i >= 0; --i) {
      idx[i]++;
      soff += dstr[i];
      if (idx[i] < ddims[i]) break;
      soff -= dstr[i] * ddims[i];
      idx[i] = 0;
    }
  }
}

// ---------------------------------------------------------------- I/O
Ten *import_ten(const int32_t *m, const double *data, bool cplx = false) {
  auto t = std::make_unique<Ten>();
  t->cplx = cplx;
  int p = 0;
  int rank = m[p++];
  t->qw = m[p++];
  int nblk = m[p++];
  t->legs.resize(rank);
  for (int l = 0; l < rank; ++l) {
    t->legs[l].dir = m[p++];
    int ns = m[p++];
    for (int s = 0; s < ns; ++s) {
      QNv q(m + p, m + p + t->qw);
      p += t->qw;
      t->legs[l].qn.push_back(q);
      t->legs[l].dim.push_back(m[p++]);
    }
  }
  std::vector<Coord> coords(nblk);
  for (int b = 0; b < nblk; ++b) {
    coords[b] = Coord(m + p, m + p + rank);
    p += rank;
  }
  int64_t off = 0;
  for (int b = 0; b < nblk; ++b) {
    int64_t sz = t->blk_size(coords[b]) * (cplx ? 2 : 1);
    t->blk[coords[b]] = std::vector<double>(data + off, data + off + sz);
    off += sz;
  }
  return t.release();
}

void export_sizes(const Ten &t, int64_t *nmeta, int64_t *ndata) {
  int64_t nm = 3;
  for (auto &l : t.legs) nm += 2 + (int64_t)l.qn.size() * (t.qw + 1);
  nm += (int64_t)t.blk.size() * t.rank();
  int64_t nd = 0;
  for (auto &kv : t.blk) nd += (int64_t)kv.second.size();
  *nmeta = nm;
  *ndata = nd;
}

void export_ten(const Ten &t, int32_t *m, double *data) {
  int p = 0;
  m[p++] = t.rank();
  m[p++] = t.qw;
  m[p++] = (int32_t)t.blk.size();
  for (auto &l : t.legs) {
    m[p++] = l.dir;
    m[p++] = (int32_t)l.qn.size();
    for (size_t s = 0; s < l.qn.size(); ++s) {
      for (int k = 0; k < t.qw; ++k) m[p++] = l.qn[s][k];
      m[p++] = (int32_t)l.dim[s];
    }
  }
  for (auto &kv : t.blk)
    for (int c : kv.first) m[p++] = c;
  int64_t off = 0;
  for (auto &kv : t.blk) {
    std::memcpy(data + off, kv.second.data(), kv.second.size() * sizeof(double));
    off += (int64_t)kv.second.size();
  }
}

Leg inverse(const Leg &l) {
  Leg r = l;
  r.dir = 1 - l.dir;
  return r;
}

// ---------------------------------------------------------------- Contract
// gqten Contract(a, b, axes, out): out legs = free legs of a (in order) then
// free legs of b; contracted legs must be mutually inverse indexes.  Output
// blocks are exactly those that receive a block-pair contribution.  Called at
// lanczos_solver_impl.h:231-234 and two_site_update_finite_vmps_impl.h:118-317.
// Complex form of a tensor (a real one gets zero imaginary parts).
Ten promote(const Ten &a) {
  if (a.cplx) return a;
  Ten r;
  r.qw = a.qw;
  r.cplx = true;
  r.legs = a.legs;
  for (auto &kv : a.blk) {
    std::vector<double> v(2 * kv.second.size(), 0.0);
    for (size_t e = 0; e < kv.second.size(); ++e) v[2 * e] = kv.second[e];
    r.blk[kv.first] = std::move(v);
  }
  return r;
}

// C[M,N] += A[M,K] * B[K,N] in complex arithmetic (interleaved doubles).
void zgemm_acc(int64_t M, int64_t N, int64_t K, const double *A, const double *B, double *C) {
#pragma omp parallel for schedule(static) if (M * N * K > 200000)
  for (int64_t i = 0; i < M; ++i)
    for (int64_t k = 0; k < K; ++k) {
      const double ar = A[2 * (i * K + k)], ai = A[2 * (i * K + k) + 1];
      const double *b = B + 2 * k * N;
      double *c = C + 2 * i * N;
      for (int64_t j = 0; j < N; ++j) {
        c[2 * j] += ar * b[2 * j] - ai * b[2 * j + 1];
        c[2 * j + 1] += ar * b[2 * j + 1] + ai * b[2 * j];
      }
    }
}

// Permute a dense row-major complex block (the real permute on (re, im)
// pairs: a trailing axis of extent 2 that stays last).
void zpermute(const double *src, const std::vector<int64_t> &dims, const std::vector<int> &perm, double *dst) {
  std::vector<int64_t> d2(dims);
  d2.push_back(2);
  std::vector<int> p2(perm);
  p2.push_back((int)dims.size());
  permute(src, d2, p2, dst);
}

Ten contract(const Ten &a, const Ten &b, const std::vector<int> &ax, const std::vector<int> &bx);

// gqten Contract of complex tensors (a real operand enters as complex).
Ten contract_complex(const Ten &a0, const Ten &b0, const std::vector<int> &ax, const std::vector<int> &bx) {
  const Ten a = promote(a0), b = promote(b0);
  int ra = a.rank(), rb = b.rank();
  std::vector<char> a_c(ra, 0), b_c(rb, 0);
  for (size_t i = 0; i < ax.size(); ++i) {
    if (!(a.legs[ax[i]] == inverse(b.legs[bx[i]])))
      throw std::runtime_error("contract: contracted indexes are not inverse of each other");
    a_c[ax[i]] = 1;
    b_c[bx[i]] = 1;
  }
  std::vector<int> afree, bfree;
  for (int i = 0; i < ra; ++i)
    if (!a_c[i]) afree.push_back(i);
  for (int i = 0; i < rb; ++i)
    if (!b_c[i]) bfree.push_back(i);
  Ten out;
  out.qw = a.qw;
  out.cplx = true;
  for (int i : afree) out.legs.push_back(a.legs[i]);
  for (int i : bfree) out.legs.push_back(b.legs[i]);
  std::vector<int> aperm(afree), bperm(bx);
  aperm.insert(aperm.end(), ax.begin(), ax.end());
  bperm.insert(bperm.end(), bfree.begin(), bfree.end());
  std::map<Coord, std::vector<const std::pair<const Coord, std::vector<double>> *>> bgrp;
  for (auto &kv : b.blk) {
    Coord k;
    for (int i : bx) k.push_back(kv.first[i]);
    bgrp[k].push_back(&kv);
  }
  std::vector<double> at, bt;
  for (auto &ka : a.blk) {
    Coord k;
    for (int i : ax) k.push_back(ka.first[i]);
    auto it = bgrp.find(k);
    if (it == bgrp.end()) continue;
    std::vector<int64_t> adims(ra);
    for (int i = 0; i < ra; ++i) adims[i] = a.legs[i].dim[ka.first[i]];
    int64_t M = 1, K = 1;
    for (int i : afree) M *= adims[i];
    for (int i : ax) K *= adims[i];
    at.resize(2 * M * K);
    zpermute(ka.second.data(), adims, aperm, at.data());
    for (auto *pb : it->second) {
      std::vector<int64_t> bdims(rb);
      for (int i = 0; i < rb; ++i) bdims[i] = b.legs[i].dim[pb->first[i]];
      int64_t N = 1;
      for (int i : bfree) N *= bdims[i];
      bt.resize(2 * K * N);
      zpermute(pb->second.data(), bdims, bperm, bt.data());
      Coord oc;
      for (int i : afree) oc.push_back(ka.first[i]);
      for (int i : bfree) oc.push_back(pb->first[i]);
      auto &ob = out.blk[oc];
      if (ob.empty()) ob.assign(2 * M * N, 0.0);
      zgemm_acc(M, N, K, at.data(), bt.data(), ob.data());
    }
  }
  return out;
}

Ten contract(const Ten &a, const Ten &b, const std::vector<int> &ax, const std::vector<int> &bx) {
  if (ax.size() != bx.size()) throw std::runtime_error("contract: axes size mismatch");
  if (a.cplx || b.cplx) return contract_complex(a, b, ax, bx);
  int ra = a.rank(), rb = b.rank();
  std::vector<char> a_c(ra, 0), b_c(rb, 0);
  for (size_t i = 0; i < ax.size(); ++i) {
    if (!(a.legs[ax[i]] == inverse(b.legs[bx[i]])))
      throw std::runtime_error("contract: contracted indexes are not inverse of each other");
    a_c[ax[i]] = 1;
    b_c[bx[i]] = 1;
  }
  std::vector<int> afree, bfree;
  for (int i = 0; i < ra; ++i)
    if (!a_c[i]) afree.push_back(i);
  for (int i = 0; i < rb; ++i)
    if (!b_c[i]) bfree.push_back(i);
  Ten out;
  out.qw = a.qw;
  for (int i : afree) out.legs.push_back(a.legs[i]);
  for (int i : bfree) out.legs.push_back(b.legs[i]);

  std::vector<int> aperm(afree), bperm(bx);
  aperm.insert(aperm.end(), ax.begin(), ax.end());
  bperm.insert(bperm.end(), bfree.begin(), bfree.end());

  // Group b blocks by contracted coordinate (in ax order).
  std::map<Coord, std::vector<const std::pair<const Coord, std::vector<double>> *>> bgrp;
  for (auto &kv : b.blk) {
    Coord k;
    for (int i : bx) k.push_back(kv.first[i]);
    bgrp[k].push_back(&kv);
  }
  std::vector<double> at, bt;
  for (auto &ka : a.blk) {
    Coord k;
    for (int i : ax) k.push_back(ka.first[i]);
    auto it = bgrp.find(k);
    if (it == bgrp.end()) continue;
    std::vector<int64_t> adims(ra);
    for (int i = 0; i < ra; ++i) adims[i] = a.legs[i].dim[ka.first[i]];
    int64_t M = 1, K = 1;
    for (int i : afree) M *= adims[i];
    for (int i : ax) K *= adims[i];
    at.resize(M * K);
    permute(ka.second.data(), adims, aperm, at.data());
    for (auto *pb : it->second) {
      std::vector<int64_t> bdims(rb);
      for (int i = 0; i < rb; ++i) bdims[i] = b.legs[i].dim[pb->first[i]];
      int64_t N = 1;
      for (int i : bfree) N *= bdims[i];
      bt.resize(K * N);
      permute(pb->second.data(), bdims, bperm, bt.data());
      Coord oc;
      for (int i : afree) oc.push_back(ka.first[i]);
      for (int i : bfree) oc.push_back(pb->first[i]);
      auto &ob = out.blk[oc];
      if (ob.empty()) ob.assign(M * N, 0.0);
      gemm_acc(M, N, K, at.data(), bt.data(), ob.data());
    }
  }
  return out;
}

// gqten Dag: inverse indexes and complex conjugate data (the identity on
// real data), lanczos_solver_impl.h:135.
Ten dag(const Ten &a) {
  Ten r = a;
  for (auto &l : r.legs) l.dir = 1 - l.dir;
  if (r.cplx)
    for (auto &kv : r.blk)
      for (size_t e = 1; e < kv.second.size(); e += 2) kv.second[e] = -kv.second[e];
  return r;
}

double norm2(const Ten &a) {
  double s = 0;
  for (auto &kv : a.blk)
    for (double v : kv.second) s += v * v;
  return s;
}

// gqten Normalize(): scale to unit 2-norm, return the old norm
// (lanczos_solver_impl.h:120,160).
double normalize(Ten &a) {
  double n = std::sqrt(norm2(a));
  if (n > 0)
    for (auto &kv : a.blk)
      for (double &v : kv.second) v /= n;
  return n;
}

// out = beta*out + sum_i c_i t_i over the union of block sets
// (gqten LinearCombine, lanczos_solver_impl.h:151-158,173,215).
void lincomb(const std::vector<double> &c, const std::vector<const Ten *> &ts, double beta, Ten &out) {
  for (auto *t : ts)
    if (t->cplx != out.cplx) throw std::runtime_error("lincomb: mixed real and complex tensors");
  for (auto &kv : out.blk)
    for (double &v : kv.second) v *= beta;
  for (size_t i = 0; i < ts.size(); ++i) {
    for (auto &kv : ts[i]->blk) {
      auto &ob = out.blk[kv.first];
      if (ob.empty()) ob.assign(kv.second.size(), 0.0);
      for (size_t e = 0; e < kv.second.size(); ++e) ob[e] += c[i] * kv.second[e];
    }
  }
}

// Re <a|b> = Re sum conj(a) b: for complex tensors the dot product of the
// interleaved doubles.
double dot(const Ten &a, const Ten &b) {
  if (a.cplx != b.cplx) return dot(promote(a), promote(b));
  double s = 0;
  for (auto &kv : a.blk) {
    auto it = b.blk.find(kv.first);
    if (it == b.blk.end()) continue;
    for (size_t e = 0; e < kv.second.size(); ++e) s += kv.second[e] * it->second[e];
  }
  return s;
}

// ---------------------------------------------------------------- tridiagonal
// Lowest eigenpair of a symmetric tridiagonal matrix by implicit QL with
// Wilkinson shifts (the textbook tql2 iteration), the stand-in for LAPACK
// dstev used by TridiagGsSolver (lanczos_solver_impl.h:259-312).
int tridiag_lowest(int n, const double *a, const double *b, double *eval, std::vector<double> *vec) {
  std::vector<double> d(a, a + n), e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = b[i];
  std::vector<double> z;
  bool wantv = vec != nullptr;
  if (wantv) {
    z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;   // z[row*n + col]
  }
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m != l) {
        if (++iter > 200) return 1;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], bb = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * bb;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - bb;
          if (wantv) {
            for (int k = 0; k < n; ++k) {
              double zf = z[(size_t)k * n + i + 1];
              z[(size_t)k * n + i + 1] = s * z[(size_t)k * n + i] + c * zf;
              z[(size_t)k * n + i] = c * z[(size_t)k * n + i] - s * zf;
            }
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  int imin = 0;
  for (int i = 1; i < n; ++i)
    if (d[i] < d[imin]) imin = i;
  *eval = d[imin];
  if (wantv) {
    vec->resize(n);
    for (int k = 0; k < n; ++k) (*vec)[k] = z[(size_t)k * n + imin];
  }
  return 0;
}

// ---------------------------------------------------------------- SVD
// Thin SVD A (m x n, row-major) = U diag(s) V^T by one-sided cyclic Jacobi
// (Hestenes 1958; de Rijk 1989 convergence test), r = min(m, n) values sorted
// descending.  Exactly-zero columns of U are completed by Gram-Schmidt so U
// and V always have orthonormal columns.
void svd_jacobi(int64_t m, int64_t n, const double *A, std::vector<double> &U, std::vector<double> &s,
                std::vector<double> &V) {
  bool tr = m < n;
  int64_t rows = tr ? n : m, cols = tr ? m : n;   // work on a (rows x cols) with rows >= cols
  // W: cols vectors of length rows (column-major of the tall matrix)
  std::vector<double> W((size_t)cols * rows), X((size_t)cols * cols, 0.0);
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) {
      double v = A[i * n + j];
      if (!tr) W[(size_t)j * rows + i] = v;
      else W[(size_t)i * rows + j] = v;
    }
  for (int64_t j = 0; j < cols; ++j) X[(size_t)j * cols + j] = 1.0;
  const double tol = 1e-15;
  for (int sweep = 0; sweep < 80; ++sweep) {
    int64_t rot = 0;
    for (int64_t p = 0; p < cols - 1; ++p) {
      for (int64_t q = p + 1; q < cols; ++q) {
        double *wp = &W[(size_t)p * rows], *wq = &W[(size_t)q * rows];
        double al = 0, be = 0, ga = 0;
        for (int64_t i = 0; i < rows; ++i) {
          al += wp[i] * wp[i];
          be += wq[i] * wq[i];
          ga += wp[i] * wq[i];
        }
        if (al == 0.0 || be == 0.0) continue;
        if (std::fabs(ga) <= tol * std::sqrt(al * be)) continue;
        ++rot;
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int64_t i = 0; i < rows; ++i) {
          double x = wp[i], y = wq[i];
          wp[i] = c * x - sn * y;
          wq[i] = sn * x + c * y;
        }
        double *xp = &X[(size_t)p * cols], *xq = &X[(size_t)q * cols];
        for (int64_t i = 0; i < cols; ++i) {
          double x = xp[i], y = xq[i];
          xp[i] = c * x - sn * y;
          xq[i] = sn * x + c * y;
        }
      }
    }
    if (rot == 0) break;
  }
  std::vector<double> nrm(cols);
  for (int64_t j = 0; j < cols; ++j) {
    double a = 0;
    for (int64_t i = 0; i < rows; ++i) a += W[(size_t)j * rows + i] * W[(size_t)j * rows + i];
    nrm[j] = std::sqrt(a);
  }
  std::vector<int64_t> ord(cols);
  for (int64_t j = 0; j < cols; ++j) ord[j] = j;
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return nrm[x] > nrm[y]; });
  // tall-side singular vectors Y (rows x cols, column-major), other side X.
  std::vector<double> Y((size_t)cols * rows, 0.0), Xs((size_t)cols * cols);
  s.assign(cols, 0.0);
  std::vector<char> have(cols, 0);
  for (int64_t jj = 0; jj < cols; ++jj) {
    int64_t j = ord[jj];
    s[jj] = nrm[j];
    std::memcpy(&Xs[(size_t)jj * cols], &X[(size_t)j * cols], cols * sizeof(double));
    if (nrm[j] > 1e-280) {
      for (int64_t i = 0; i < rows; ++i) Y[(size_t)jj * rows + i] = W[(size_t)j * rows + i] / nrm[j];
      have[jj] = 1;
    }
  }
  for (int64_t jj = 0; jj < cols; ++jj) {   // completion for exact zeros
    if (have[jj]) continue;
    for (int64_t e = 0; e < rows; ++e) {
      std::vector<double> v(rows, 0.0);
      v[e] = 1.0;
      for (int pass = 0; pass < 2; ++pass)
        for (int64_t k = 0; k < cols; ++k) {
          if (!have[k]) continue;
          double d = 0;
          for (int64_t i = 0; i < rows; ++i) d += Y[(size_t)k * rows + i] * v[i];
          for (int64_t i = 0; i < rows; ++i) v[i] -= d * Y[(size_t)k * rows + i];
        }
      double nv = 0;
      for (double x : v) nv += x * x;
      nv = std::sqrt(nv);
      if (nv > 0.5) {
        for (int64_t i = 0; i < rows; ++i) Y[(size_t)jj * rows + i] = v[i] / nv;
        have[jj] = 1;
        break;
      }
    }
  }
  // U (m x r) row-major, V (n x r) row-major
  int64_t r = cols;
  U.assign((size_t)m * r, 0.0);
  V.assign((size_t)n * r, 0.0);
  const std::vector<double> &UL = tr ? Xs : Y;   // column-major, length m each
  const std::vector<double> &VL = tr ? Y : Xs;   // column-major, length n each
  for (int64_t k = 0; k < r; ++k) {
    for (int64_t i = 0; i < m; ++i) U[(size_t)i * r + k] = UL[(size_t)k * m + i];
    for (int64_t i = 0; i < n; ++i) V[(size_t)i * r + k] = VL[(size_t)k * n + i];
  }
}

// Complex thin SVD A (m x n, row-major, interleaved) = U diag(s) V^H by
// one-sided cyclic Jacobi on the columns of the tall orientation W (A, or A^H
// when m < n): for a column pair with g = sum conj(w_p) w_q = |g| e^{i phi},
// the real rotation of (w_p, e^{-i phi} w_q) orthogonalises them (the
// Hermitian Jacobi step); the rotations accumulate X with W X = Y diag(s).
// U = Y, V = X for A; U = X, V = Y for A^H.  r = min(m, n) values, descending.
void svd_jacobi_c(int64_t m, int64_t n, const double *A, std::vector<double> &U, std::vector<double> &s,
                  std::vector<double> &V) {
  bool tr = m < n;
  int64_t rows = tr ? n : m, cols = tr ? m : n;
  std::vector<double> W(2 * (size_t)cols * rows), X(2 * (size_t)cols * cols, 0.0);
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) {
      const double re = A[2 * (i * n + j)], im = A[2 * (i * n + j) + 1];
      if (!tr) {
        W[2 * ((size_t)j * rows + i)] = re;
        W[2 * ((size_t)j * rows + i) + 1] = im;
      } else {   // W = A^H: column i holds conj(row i of A)
        W[2 * ((size_t)i * rows + j)] = re;
        W[2 * ((size_t)i * rows + j) + 1] = -im;
      }
    }
  for (int64_t j = 0; j < cols; ++j) X[2 * ((size_t)j * cols + j)] = 1.0;
  const double tol = 1e-15;
  auto rot = [](double *u, double *v, int64_t len, double c, double sn, double er, double ei) {
    // u' = c u - sn e^{-i phi} v ;  v' = sn e^{i phi} u + c v   (e = er + i ei)
    for (int64_t i = 0; i < len; ++i) {
      const double ur = u[2 * i], ui = u[2 * i + 1], vr = v[2 * i], vi = v[2 * i + 1];
      const double evr = er * vr + ei * vi, evi = er * vi - ei * vr;   // conj(e) v
      const double eur = er * ur - ei * ui, eui = er * ui + ei * ur;   // e u
      u[2 * i] = c * ur - sn * evr;
      u[2 * i + 1] = c * ui - sn * evi;
      v[2 * i] = sn * eur + c * vr;
      v[2 * i + 1] = sn * eui + c * vi;
    }
  };
  for (int sweep = 0; sweep < 80; ++sweep) {
    int64_t nrot = 0;
    for (int64_t p = 0; p < cols - 1; ++p) {
      for (int64_t q = p + 1; q < cols; ++q) {
        double *wp = &W[2 * (size_t)p * rows], *wq = &W[2 * (size_t)q * rows];
        double al = 0, be = 0, gr = 0, gi = 0;
        for (int64_t i = 0; i < rows; ++i) {
          const double pr = wp[2 * i], pi = wp[2 * i + 1], qr = wq[2 * i], qi = wq[2 * i + 1];
          al += pr * pr + pi * pi;
          be += qr * qr + qi * qi;
          gr += pr * qr + pi * qi;   // g = sum conj(w_p) w_q
          gi += pr * qi - pi * qr;
        }
        const double g = std::sqrt(gr * gr + gi * gi);
        if (al == 0.0 || be == 0.0) continue;
        if (g <= tol * std::sqrt(al * be)) continue;
        ++nrot;
        const double zeta = (be - al) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        rot(wp, wq, rows, c, sn, gr / g, gi / g);
        rot(&X[2 * (size_t)p * cols], &X[2 * (size_t)q * cols], cols, c, sn, gr / g, gi / g);
      }
    }
    if (nrot == 0) break;
  }
  std::vector<double> nrm(cols);
  for (int64_t j = 0; j < cols; ++j) {
    double a = 0;
    for (int64_t i = 0; i < rows; ++i) {
      const double re = W[2 * ((size_t)j * rows + i)], im = W[2 * ((size_t)j * rows + i) + 1];
      a += re * re + im * im;
    }
    nrm[j] = std::sqrt(a);
  }
  std::vector<int64_t> ord(cols);
  for (int64_t j = 0; j < cols; ++j) ord[j] = j;
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return nrm[x] > nrm[y]; });
  const int64_t r = cols;
  s.assign(r, 0.0);
  U.assign(2 * (size_t)m * r, 0.0);
  V.assign(2 * (size_t)n * r, 0.0);
  for (int64_t k = 0; k < r; ++k) {
    const int64_t j = ord[k];
    s[k] = nrm[j];
    const double inv = nrm[j] > 1e-280 ? 1.0 / nrm[j] : 0.0;
    // Y column = W column / s (tall side), X column (other side)
    std::vector<double> &T = tr ? V : U;   // tall side: length rows
    std::vector<double> &O = tr ? U : V;   // other side: length cols
    for (int64_t i = 0; i < rows; ++i) {
      T[2 * ((size_t)i * r + k)] = W[2 * ((size_t)j * rows + i)] * inv;
      T[2 * ((size_t)i * r + k) + 1] = W[2 * ((size_t)j * rows + i) + 1] * inv;
    }
    for (int64_t i = 0; i < cols; ++i) {
      O[2 * ((size_t)i * r + k)] = X[2 * ((size_t)j * cols + i)];
      O[2 * ((size_t)i * r + k) + 1] = X[2 * ((size_t)j * cols + i) + 1];
    }
  }
}

// Truncation rule shared (as a specification, not as code) with the product:
//  1. all singular values of all sectors, sorted by value descending, ties by
//     (sector position, index in sector) ascending;
//  2. D_err = smallest D >= 1 with sum_{i>=D} s_i^2 <= trunc_err * sum s_i^2;
//  3. D = min(max(D_err, Dmin), Dmax, count);
//  4. values within relative kDegTol of the last kept value form the cut
//     cluster; the cut keeps as many of them as step 3 keeps, chosen by
//     (sector position, index) ascending, so near-ties do not depend on
//     rounding.
const double kDegTol = 1e-6;

struct SvInfo {
  double v;
  int sec;
  int64_t j;
};

void truncate_rule(const std::vector<std::vector<double>> &sv, double trunc_err, int64_t dmin, int64_t dmax,
                   std::vector<int64_t> &kept, double *err, int64_t *D) {
  std::vector<SvInfo> all;
  for (int q = 0; q < (int)sv.size(); ++q)
    for (int64_t j = 0; j < (int64_t)sv[q].size(); ++j) all.push_back({sv[q][j], q, j});
  std::sort(all.begin(), all.end(), [](const SvInfo &x, const SvInfo &y) {
    if (x.v != y.v) return x.v > y.v;
    if (x.sec != y.sec) return x.sec < y.sec;
    return x.j < y.j;
  });
  int64_t cnt = (int64_t)all.size();
  double total = 0;
  for (auto &x : all) total += x.v * x.v;
  int64_t Derr = cnt;
  {
    double tail = 0;
    // smallest D with tail(D) <= trunc_err*total; tail(cnt) = 0
    int64_t Dm = cnt;
    for (int64_t i = cnt - 1; i >= 1; --i) {
      tail += all[i].v * all[i].v;
      if (tail <= trunc_err * total) Dm = i;
      else break;
    }
    Derr = Dm;
  }
  int64_t d = std::max(Derr, dmin);
  d = std::min(d, dmax);
  d = std::min(d, cnt);
  if (d < 1 && cnt > 0) d = 1;
  std::vector<char> keep(cnt, 0);
  for (int64_t i = 0; i < d; ++i) keep[i] = 1;
  if (d < cnt && d > 0) {
    double v = all[d - 1].v;
    std::vector<int64_t> cl;
    int64_t kin = 0;
    for (int64_t i = 0; i < cnt; ++i)
      if (std::fabs(all[i].v - v) <= kDegTol * v) {
        cl.push_back(i);
        if (i < d) ++kin;
      }
    std::sort(cl.begin(), cl.end(), [&](int64_t x, int64_t y) {
      if (all[x].sec != all[y].sec) return all[x].sec < all[y].sec;
      return all[x].j < all[y].j;
    });
    for (int64_t t = 0; t < (int64_t)cl.size(); ++t) keep[cl[t]] = t < kin ? 1 : 0;
  }
  kept.assign(sv.size(), 0);
  double disc = 0;
  for (int64_t i = 0; i < cnt; ++i) {
    if (keep[i]) kept[all[i].sec]++;
    else disc += all[i].v * all[i].v;
  }
  *err = total > 0 ? disc / total : 0.0;
  *D = d;
}

int qn_cmp(const QNv &a, const QNv &b) {
  for (size_t i = 0; i < a.size(); ++i) {
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  }
  return 0;
}

// gqten truncated SVD as called at two_site_update_finite_vmps_impl.h:258-263:
// legs [0, ldims) become U's legs plus a new OUT leg k; S is the (IN, OUT)
// diagonal on the new bond; Vt has the new IN leg then the remaining legs.
// New-bond sectors are ordered by ascending quantum number; inside a sector
// states are ordered by descending singular value.  Rows/columns of a sector
// matrix come from the blocks that hold a nonzero element.
void svd(const Ten &t, int ldims, const QNv &ldiv, double trunc_err, int64_t dmin, int64_t dmax, Ten &U, Ten &S,
         Ten &Vt, double *err, int64_t *D) {
  int r = t.rank();
  auto sgn = [&](int leg) { return t.legs[leg].dir == 1 ? 1 : -1; };
  // sector qn of the new bond for a block
  auto kqn = [&](const Coord &c) {
    QNv q = ldiv;
    for (int i = 0; i < ldims; ++i) q = qn_add(q, t.legs[i].qn[c[i]], -sgn(i));
    return q;
  };
  struct Sec {
    QNv qn;
    std::map<Coord, int64_t> rows, cols;   // tuple -> offset
    int64_t m = 0, n = 0;
  };
  std::vector<Sec> secs;
  auto find_sec = [&](const QNv &q) -> Sec & {
    for (auto &s : secs)
      if (s.qn == q) return s;
    secs.push_back(Sec());
    secs.back().qn = q;
    return secs.back();
  };
  for (auto &kv : t.blk) {
    bool nz = false;
    for (double v : kv.second) nz |= (v != 0.0);
    if (!nz) continue;
    Sec &s = find_sec(kqn(kv.first));
    s.rows[Coord(kv.first.begin(), kv.first.begin() + ldims)] = 0;
    s.cols[Coord(kv.first.begin() + ldims, kv.first.end())] = 0;
  }
  std::sort(secs.begin(), secs.end(), [](const Sec &a, const Sec &b) { return qn_cmp(a.qn, b.qn) < 0; });
  for (auto &s : secs) {
    for (auto &kv : s.rows) {
      kv.second = s.m;
      int64_t d = 1;
      for (int i = 0; i < ldims; ++i) d *= t.legs[i].dim[kv.first[i]];
      s.m += d;
    }
    for (auto &kv : s.cols) {
      kv.second = s.n;
      int64_t d = 1;
      for (int i = ldims; i < r; ++i) d *= t.legs[i].dim[kv.first[i - ldims]];
      s.n += d;
    }
  }
  int ns = (int)secs.size();
  // LAPACK's gesdd, behind gqten's SVD, is accurate at any finite scale (it
  // rescales out-of-range matrices); the Jacobi sums of squares below are
  // not, so the sector matrices are taken scaled by 2^-e, 2^e <= max|t| <
  // 2^(e+1) (exact, so U and V are unchanged), and S is scaled back.
  double gmax = 0.0;
  for (auto &kv : t.blk)
    for (double v : kv.second) {
      if (!std::isfinite(v)) throw std::runtime_error("svd: input holds NaN or Inf");
      gmax = std::max(gmax, std::fabs(v));
    }
  const int sexp = gmax > 0.0 ? std::max(-1000, std::min(1000, std::ilogb(gmax))) : 0;
  const double pre = std::ldexp(1.0, -sexp), post = std::ldexp(1.0, sexp);
  const int w = t.cplx ? 2 : 1;   // doubles per element
  std::vector<std::vector<double>> Us(ns), Ss(ns), Vs(ns);
  for (int q = 0; q < ns; ++q) {
    auto &s = secs[q];
    std::vector<double> M((size_t)w * s.m * s.n, 0.0);
    for (auto &kv : t.blk) {
      Coord rc(kv.first.begin(), kv.first.begin() + ldims), cc(kv.first.begin() + ldims, kv.first.end());
      auto ir = s.rows.find(rc);
      auto ic = s.cols.find(cc);
      if (ir == s.rows.end() || ic == s.cols.end()) continue;
      if (kqn(kv.first) != s.qn) continue;
      int64_t bm = 1, bn = 1;
      for (int i = 0; i < ldims; ++i) bm *= t.legs[i].dim[kv.first[i]];
      for (int i = ldims; i < r; ++i) bn *= t.legs[i].dim[kv.first[i]];
      for (int64_t i = 0; i < bm; ++i)
        for (int64_t j = 0; j < bn; ++j)
          for (int h = 0; h < w; ++h)
            M[w * ((size_t)(ir->second + i) * s.n + ic->second + j) + h] = kv.second[w * (i * bn + j) + h] * pre;
    }
    // each sector additionally at its own scale 2^-e_q (exact): sectors of
    // negligible weight would otherwise underflow in the sums of squares
    double mq = 0.0;
    for (double v : M) mq = std::max(mq, std::fabs(v));
    const int eq = mq > 0.0 ? std::max(-1000, std::min(1000, std::ilogb(mq))) : 0;
    for (double &v : M) v = std::ldexp(v, -eq);
    if (t.cplx) svd_jacobi_c(s.m, s.n, M.data(), Us[q], Ss[q], Vs[q]);
    else svd_jacobi(s.m, s.n, M.data(), Us[q], Ss[q], Vs[q]);
    for (double &v : Ss[q]) v = std::ldexp(v, eq);
  }
  // Below 1e-11 s_max (s_max over all sectors; weights below 1e-22) FP64
  // SVD values carry no reproducible order (a direct SVD's absolute error is
  // p(n) eps s_max; a Gram-based one keeps larger artifacts).  The rule
  // counts every nonzero value below 1e-11 s_max as exactly 1e-11 s_max, so
  // a cut among them is a tie decided by (sector, index), not by rounding;
  // exact zeros stay zero and are cut.
  double smax_all = 0.0;
  for (auto &v : Ss)
    for (double x : v) smax_all = std::max(smax_all, x);
  const double thr = 1e-11 * smax_all;
  std::vector<std::vector<double>> Sr(Ss);
  for (auto &v : Sr)
    for (double &x : v)
      if (x > 0.0 && x < thr) x = thr;
  std::vector<int64_t> kept;
  truncate_rule(Sr, trunc_err, dmin, dmax, kept, err, D);
  Leg knew;
  knew.dir = 1;
  std::vector<int> secpos(ns, -1);
  for (int q = 0; q < ns; ++q)
    if (kept[q] > 0) {
      secpos[q] = (int)knew.qn.size();
      knew.qn.push_back(secs[q].qn);
      knew.dim.push_back(kept[q]);
    }
  U = Ten();
  S = Ten();
  Vt = Ten();
  U.qw = S.qw = Vt.qw = t.qw;
  U.cplx = Vt.cplx = t.cplx;
  for (int i = 0; i < ldims; ++i) U.legs.push_back(t.legs[i]);
  U.legs.push_back(knew);
  S.legs.push_back(inverse(knew));
  S.legs.push_back(knew);
  Vt.legs.push_back(inverse(knew));
  for (int i = ldims; i < r; ++i) Vt.legs.push_back(t.legs[i]);
  for (int q = 0; q < ns; ++q) {
    if (secpos[q] < 0) continue;
    auto &s = secs[q];
    int64_t kq = kept[q], rq = (int64_t)Ss[q].size();
    for (auto &kv : s.rows) {
      Coord c = kv.first;
      c.push_back(secpos[q]);
      int64_t bm = 1;
      for (int i = 0; i < ldims; ++i) bm *= t.legs[i].dim[kv.first[i]];
      std::vector<double> b((size_t)w * bm * kq);
      for (int64_t i = 0; i < bm; ++i)
        for (int64_t k = 0; k < kq; ++k)
          for (int h = 0; h < w; ++h) b[w * (i * kq + k) + h] = Us[q][w * ((size_t)(kv.second + i) * rq + k) + h];
      U.blk[c] = std::move(b);
    }
    for (auto &kv : s.cols) {
      Coord c;
      c.push_back(secpos[q]);
      c.insert(c.end(), kv.first.begin(), kv.first.end());
      int64_t bn = 1;
      for (int i = ldims; i < r; ++i) bn *= t.legs[i].dim[kv.first[i - ldims]];
      std::vector<double> b((size_t)w * kq * bn);
      for (int64_t k = 0; k < kq; ++k)
        for (int64_t j = 0; j < bn; ++j)
          for (int h = 0; h < w; ++h)   // Vt = V^H: conjugate
            b[w * (k * bn + j) + h] = (h ? -1.0 : 1.0) * Vs[q][w * ((size_t)(kv.second + j) * rq + k) + h];
      Vt.blk[c] = std::move(b);
    }
    std::vector<double> sd((size_t)kq * kq, 0.0);
    for (int64_t k = 0; k < kq; ++k) sd[k * kq + k] = Ss[q][k] * post;
    S.blk[Coord{secpos[q], secpos[q]}] = std::move(sd);
  }
}

// ---------------------------------------------------------------- matvec
// eff_ham_mul_state_{cent,lend,rend}: lanczos_solver_impl.h:228-256.
Ten heff_mul(const Ten *L, const Ten *W1, const Ten *W2, const Ten *R, const Ten &psi, const std::string &where) {
  if (where == "cent") {
    Ten t = contract(*L, psi, {0}, {0});
    t = contract(t, *W1, {0, 2}, {0, 1});
    t = contract(t, *W2, {4, 1}, {0, 1});
    return contract(t, *R, {4, 1}, {1, 0});
  } else if (where == "lend") {
    Ten t = contract(psi, *W1, {0}, {0});
    t = contract(t, *W2, {0, 2}, {1, 0});
    return contract(t, *R, {0, 3}, {0, 1});
  } else if (where == "rend") {
    Ten t = contract(psi, *L, {0}, {0});
    t = contract(t, *W1, {2, 0}, {0, 1});
    return contract(t, *W2, {3, 0}, {1, 0});
  }
  throw std::runtime_error("heff_mul: where must be lend, cent or rend");
}

struct LanczosOut {
  int64_t iters;
  double energy;
  Ten gs;
};

// LanczosSolver: lanczos_solver_impl.h:79-225 (plain three-term recurrence,
// every basis vector kept, no reorthogonalisation, stop on energy decrease).
LanczosOut lanczos(const Ten *L, const Ten *W1, const Ten *W2, const Ten *R, const Ten &init,
                   const std::string &where, double error, int64_t max_iter) {
  int64_t eff_dim = 1;
  auto d0 = [](const Ten *t, int leg) { return t->legs[leg].total(); };
  if (where == "cent") eff_dim = d0(L, 0) * d0(W1, 1) * d0(W2, 1) * d0(R, 0);
  else if (where == "lend") eff_dim = d0(W1, 0) * d0(W2, 1) * d0(R, 0);
  else eff_dim = d0(L, 0) * d0(W1, 1) * d0(W2, 0);
  std::vector<Ten> bases;
  bases.reserve(max_iter);
  std::vector<double> a(max_iter, 0.0), b(max_iter, 0.0), N(max_iter, 0.0);
  const bool cz = init.cplx || (L && L->cplx) || W1->cplx || W2->cplx || (R && R->cplx);
  bases.push_back(cz ? promote(init) : init);
  normalize(bases[0]);
  Ten last = heff_mul(L, W1, W2, R, bases[0], where);
  a[0] = dot(last, bases[0]);
  double energy0 = a[0];
  int64_t m = 0;
  LanczosOut res;
  while (true) {
    m += 1;
    Ten gamma = last;
    if (m == 1) lincomb({-a[m - 1]}, {&bases[m - 1]}, 1.0, gamma);
    else lincomb({-a[m - 1], -std::sqrt(N[m - 1])}, {&bases[m - 1], &bases[m - 2]}, 1.0, gamma);
    double ng = normalize(gamma);
    if (ng == 0.0) {
      if (m == 1) {
        res.iters = m;
        res.energy = energy0;
        res.gs = bases[0];
        return res;
      }
      double ev;
      std::vector<double> z;
      if (tridiag_lowest((int)m, a.data(), b.data(), &ev, &z)) throw std::runtime_error("?stev error.");
      Ten gs;
      gs.qw = init.qw;
      gs.cplx = bases[0].cplx;
      gs.legs = bases[0].legs;
      std::vector<const Ten *> ps;
      for (int64_t i = 0; i < m; ++i) ps.push_back(&bases[i]);
      lincomb(z, ps, 0.0, gs);
      res.iters = m;
      res.energy = energy0;
      res.gs = gs;
      return res;
    }
    N[m] = ng * ng;
    b[m - 1] = ng;
    bases.push_back(std::move(gamma));
    last = heff_mul(L, W1, W2, R, bases[m], where);
    a[m] = dot(last, bases[m]);
    double ev;
    if (tridiag_lowest((int)m + 1, a.data(), b.data(), &ev, nullptr)) throw std::runtime_error("?stev error.");
    double e_new = ev;
    if ((energy0 - e_new) < error || m == eff_dim || m == max_iter - 1) {
      std::vector<double> z;
      tridiag_lowest((int)m + 1, a.data(), b.data(), &ev, &z);
      Ten gs;
      gs.qw = init.qw;
      gs.cplx = bases[0].cplx;
      gs.legs = bases[0].legs;
      std::vector<const Ten *> ps;
      for (int64_t i = 0; i <= m; ++i) ps.push_back(&bases[i]);
      lincomb(z, ps, 0.0, gs);
      res.iters = m;
      res.energy = e_new;
      res.gs = std::move(gs);
      return res;
    }
    energy0 = e_new;
  }
}

QNv div_of(const Ten &t) {
  QNv d = qn_zero(t.qw);
  if (t.blk.empty()) return d;
  const Coord &c = t.blk.begin()->first;
  for (int i = 0; i < t.rank(); ++i) d = qn_add(d, t.legs[i].qn[c[i]], t.legs[i].dir == 1 ? 1 : -1);
  return d;
}

// FiniteMPS::RightCanonicalizeTen_ (finite_mps.h:213-243) for i = N-1..1,
// i.e. Centralize(0) (finite_mps.h:148-157) on a fresh MPS; the untruncated
// SVD is mock_gqten::SVD (utilities.h:29-55).
void centralize0(std::vector<Ten> &mps) {
  int n = (int)mps.size();
  for (int i = n - 1; i >= 1; --i) {
    Ten &A = mps[i];
    int64_t lsize = A.legs[0].total(), rsize = 1;
    for (int l = 1; l < A.rank(); ++l) rsize *= A.legs[l].total();
    int64_t Dm = std::max(lsize, rsize);
    Ten u, s, vt;
    double err;
    int64_t D;
    svd(A, 1, qn_zero(A.qw), 0.0, Dm, Dm, u, s, vt, &err, &D);
    mps[i] = std::move(vt);
    Ten us = contract(u, s, {1}, {0});
    int ax = (i - 1 == 0) ? 1 : 2;
    mps[i - 1] = contract(mps[i - 1], us, {ax}, {0});
  }
}

// TwoSiteFiniteVMPS + InitEnvs + Sweep + Update (two_site_update_finite_vmps_impl.h:77-340),
// all tensors resident in memory instead of streamed to disk.
struct Dmrg {
  int N;
  std::vector<Ten> &mps;
  const std::vector<Ten> &mpo;
  std::vector<Ten> lenvs, renvs;
  int64_t dmin, dmax;
  double trunc_err, lerr;
  int64_t lmax;
  Dmrg(std::vector<Ten> &m, const std::vector<Ten> &w) : N((int)m.size()), mps(m), mpo(w) {}

  void init_envs() {   // :104-135
    lenvs.assign(N, Ten());
    renvs.assign(N, Ten());
    for (int i = 1; i <= N - 2; ++i) {
      const Ten &A = mps[N - i];
      if (i == 1) {
        Ten temp = contract(A, mpo.back(), {1}, {0});
        renvs[i] = contract(temp, dag(A), {2}, {1});
      } else {
        Ten t1 = contract(A, renvs[i - 1], {2}, {0});
        Ten t2 = contract(t1, mpo[N - i], {1, 2}, {1, 3});
        renvs[i] = contract(t2, dag(A), {3, 1}, {1, 2});
      }
    }
  }

  double update(char dir, int site, double *e_out, int64_t *d_out, int64_t *it_out, double *te_out) {   // :164-340
    int ls, rs, llen, rlen;
    std::vector<int> ia, ib, usa, usb;
    std::string where;
    int svd_ldims;
    if (dir == 'r') {
      ls = site;
      rs = site + 1;
      llen = site;
      rlen = N - (site + 2);
      if (site == 0) {
        ia = {1}; ib = {0}; where = "lend"; svd_ldims = 1;
      } else if (site == N - 2) {
        ia = {2}; ib = {0}; where = "rend"; svd_ldims = 2;
      } else {
        ia = {2}; ib = {0}; where = "cent"; svd_ldims = 2;
      }
    } else {
      ls = site - 1;
      rs = site;
      llen = site - 1;
      rlen = N - site - 1;
      if (site == N - 1) {
        ia = {2}; ib = {0}; where = "rend"; svd_ldims = 2; usa = {2}; usb = {0};
      } else if (site == 1) {
        ia = {1}; ib = {0}; where = "lend"; svd_ldims = 1; usa = {1}; usb = {0};
      } else {
        ia = {2}; ib = {0}; where = "cent"; svd_ldims = 2; usa = {2}; usb = {0};
      }
    }
    Ten init = contract(mps[ls], mps[rs], ia, ib);
    LanczosOut lr = lanczos(&lenvs[llen], &mpo[ls], &mpo[rs], &renvs[rlen], init, where, lerr, lmax);
    Ten u, s, vt;
    double terr;
    int64_t D;
    svd(lr.gs, svd_ldims, div_of(mps[ls]), trunc_err, dmin, dmax, u, s, vt, &terr, &D);
    if (dir == 'r') {
      mps[ls] = std::move(u);
      mps[rs] = contract(s, vt, {1}, {0});
    } else {
      mps[ls] = contract(u, s, usa, usb);
      mps[rs] = std::move(vt);
    }
    if (dir == 'r') {
      if (site != N - 2) {
        if (site == 0) {
          Ten temp = contract(mps[site], mpo[site], {0}, {0});
          lenvs[llen + 1] = contract(temp, dag(mps[site]), {2}, {0});
        } else {
          Ten t1 = contract(lenvs[llen], mps[site], {0}, {0});
          Ten t2 = contract(t1, mpo[site], {0, 2}, {0, 1});
          lenvs[llen + 1] = contract(t2, dag(mps[site]), {0, 2}, {0, 1});
        }
      }
    } else {
      if (site != 1) {
        if (site == N - 1) {
          Ten temp = contract(mps[site], mpo[site], {1}, {0});
          renvs[rlen + 1] = contract(temp, dag(mps[site]), {2}, {1});
        } else {
          Ten t1 = contract(mps[site], renvs[rlen], {2}, {0});
          Ten t2 = contract(t1, mpo[site], {1, 2}, {1, 3});
          renvs[rlen + 1] = contract(t2, dag(mps[site]), {3, 1}, {1, 2});
        }
      }
    }
    if (e_out) *e_out = lr.energy;
    if (d_out) *d_out = D;
    if (it_out) *it_out = lr.iters;
    if (te_out) *te_out = terr;
    return lr.energy;
  }
};

// ---------------------------------------------------------------- measurements
// <psi| prod_i op_i |psi> of the MPS as it stands (any gauge, no
// normalisation): the transfer chain from site 0 to site N-1 with the identity
// where ops[i] is NULL.  This is what MeasureOneSiteOp / MeasureMultiSiteOp
// return (finite_mps_measu.h:150-554) without their canonical-form shortcut:
// the reference centres the MPS at the first operator so that the chain left
// of it and right of the last operator is the identity; here every site is
// contracted, so the value does not depend on the gauge.
// MPS legs (l IN, p OUT, r OUT), ends (p, r) and (l, p); op (p IN, p OUT).
double expect(const std::vector<Ten> &mps, const std::vector<const Ten *> &ops) {
  const int n = (int)mps.size();
  if (n < 2 || (int)ops.size() != n) throw std::invalid_argument("expect: need N >= 2 tensors and N operator slots");
  Ten E;
  {
    const Ten &A = mps[0];
    Ten bra = dag(A);
    if (ops[0]) E = contract(contract(A, *ops[0], {0}, {0}), bra, {1}, {0});
    else E = contract(A, bra, {0}, {0});
  }
  for (int i = 1; i < n; ++i) {
    const Ten &A = mps[i];
    Ten bra = dag(A);
    Ten t = contract(E, A, {0}, {0});   // (bra bond, p[, r])
    if (ops[i]) t = contract(t, *ops[i], {1}, {0});   // (bra bond[, r], p')
    const bool last = i == n - 1;
    if (ops[i] && !last) E = contract(t, bra, {0, 2}, {0, 1});
    else E = contract(t, bra, {0, 1}, {0, 1});
  }
  if (E.rank() != 0) throw std::logic_error("expect: chain did not close");
  auto it = E.blk.find(Coord());
  return it == E.blk.end() ? 0.0 : it->second[0];
}

template <class F>
int guarded(F f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

oracle_ten oracle_ten_import(const int32_t *meta, const double *data) {
  Ten *t = nullptr;
  guarded([&] { t = import_ten(meta, data); });
  return t;
}

oracle_ten oracle_ten_import_complex(const int32_t *meta, const double *data) {
  Ten *t = nullptr;
  guarded([&] { t = import_ten(meta, data, true); });
  return t;
}

int oracle_ten_is_complex(oracle_ten t) { return T(t)->cplx ? 1 : 0; }

int oracle_ten_export_sizes(oracle_ten t, int64_t *nmeta, int64_t *ndata) {
  export_sizes(*T(t), nmeta, ndata);
  return 0;
}

int oracle_ten_export(oracle_ten t, int32_t *meta, double *data) {
  export_ten(*T(t), meta, data);
  return 0;
}

void oracle_ten_free(oracle_ten t) { delete T(t); }

oracle_ten oracle_contract(oracle_ten a, oracle_ten b, int nax, const int32_t *axes_a, const int32_t *axes_b) {
  Ten *r = nullptr;
  guarded([&] {
    std::vector<int> x(axes_a, axes_a + nax), y(axes_b, axes_b + nax);
    r = new Ten(contract(*T(a), *T(b), x, y));
  });
  return r;
}

oracle_ten oracle_dag(oracle_ten a) { return new Ten(dag(*T(a))); }

double oracle_normalize(oracle_ten a) { return normalize(*T(a)); }

double oracle_dot(oracle_ten a, oracle_ten b) { return dot(*T(a), *T(b)); }

int oracle_svd(oracle_ten t, int ldims, const int32_t *ldiv, double trunc_err, int64_t dmin, int64_t dmax,
               oracle_ten *u, oracle_ten *s, oracle_ten *vt, double *actual_trunc_err, int64_t *d_kept) {
  return guarded([&] {
    Ten U, S, V;
    QNv q(ldiv, ldiv + T(t)->qw);
    svd(*T(t), ldims, q, trunc_err, dmin, dmax, U, S, V, actual_trunc_err, d_kept);
    *u = new Ten(std::move(U));
    *s = new Ten(std::move(S));
    *vt = new Ten(std::move(V));
  });
}

int oracle_tridiag_lowest(int n, const double *a, const double *b, double *eval, double *vec) {
  std::vector<double> z;
  int rc = tridiag_lowest(n, a, b, eval, vec ? &z : nullptr);
  if (rc == 0 && vec) std::memcpy(vec, z.data(), n * sizeof(double));
  return rc;
}

int oracle_lanczos(oracle_ten L, oracle_ten W1, oracle_ten W2, oracle_ten R, oracle_ten init, const char *where,
                   double error, int64_t max_iter, int64_t *iters, double *energy, oracle_ten *gs) {
  return guarded([&] {
    LanczosOut r = lanczos(T(L), T(W1), T(W2), T(R), *T(init), where, error, max_iter);
    *iters = r.iters;
    *energy = r.energy;
    *gs = new Ten(std::move(r.gs));
  });
}

oracle_ten oracle_heff_mul(oracle_ten L, oracle_ten W1, oracle_ten W2, oracle_ten R, oracle_ten psi,
                           const char *where) {
  Ten *r = nullptr;
  guarded([&] { r = new Ten(heff_mul(T(L), T(W1), T(W2), T(R), *T(psi), where)); });
  return r;
}

int oracle_centralize0(int n, oracle_ten *mps) {
  return guarded([&] {
    std::vector<Ten> v;
    for (int i = 0; i < n; ++i) v.push_back(*T(mps[i]));
    centralize0(v);
    for (int i = 0; i < n; ++i) {
      delete T(mps[i]);
      mps[i] = new Ten(std::move(v[i]));
    }
  });
}

int oracle_dmrg(int n, oracle_ten *mps, oracle_ten *mpo, int64_t sweeps, int64_t dmin, int64_t dmax,
                double trunc_err, double lanczos_error, int64_t lanczos_max_iter, double *energies, int64_t *dims,
                int64_t *iters, double *trunc_errs, double *final_energy) {
  return guarded([&] {
    std::vector<Ten> m, w;
    for (int i = 0; i < n; ++i) {
      m.push_back(*T(mps[i]));
      w.push_back(*T(mpo[i]));
    }
    Dmrg d(m, w);
    d.dmin = dmin;
    d.dmax = dmax;
    d.trunc_err = trunc_err;
    d.lerr = lanczos_error;
    d.lmax = lanczos_max_iter;
    d.init_envs();
    double e0 = 0;
    int64_t k = 0;
    for (int64_t sw = 0; sw < sweeps; ++sw) {
      for (int i = 0; i < n - 1; ++i, ++k)
        e0 = d.update('r', i, energies ? energies + k : nullptr, dims ? dims + k : nullptr,
                      iters ? iters + k : nullptr, trunc_errs ? trunc_errs + k : nullptr);
      for (int i = n - 1; i > 0; --i, ++k)
        e0 = d.update('l', i, energies ? energies + k : nullptr, dims ? dims + k : nullptr,
                      iters ? iters + k : nullptr, trunc_errs ? trunc_errs + k : nullptr);
    }
    if (final_energy) *final_energy = e0;
    for (int i = 0; i < n; ++i) {
      delete T(mps[i]);
      mps[i] = new Ten(std::move(m[i]));
    }
  });
}

int oracle_expect(int n, oracle_ten *mps, oracle_ten *ops, double *out) {
  return guarded([&] {
    std::vector<Ten> m;
    std::vector<const Ten *> o;
    for (int i = 0; i < n; ++i) {
      m.push_back(*T(mps[i]));
      o.push_back(ops && ops[i] ? T(ops[i]) : nullptr);
    }
    *out = expect(m, o);
  });
}

int oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

}  // extern "C"
