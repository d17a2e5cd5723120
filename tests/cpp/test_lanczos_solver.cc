// LanczosSolver cases of tests/test_algorithm/test_lanczos_solver.cc:42-340
// (GraceQ/MPS2 0.2-alpha.0), double-precision cases, written against the
// gqmps2-compatible headers: a centre-bond effective Hamiltonian (random
// symmetric block and site matrices, D = 20, d = 2, MPO bond 2) solved to
// an error target and in the maximum-iteration mode, and a left-end one.
// The reference checks the Lanczos energy against LAPACK on the dense
// effective Hamiltonian; this program writes that dense matrix (built with
// the same device contractions) next to the energy, and the calling test
// diagonalises it.  Output lines: "CASE <name> E=<energy> iters=<n> dim=<n>",
// files <name>.bin (dim x dim doubles, row-major) and, for the centre-bond
// cases, the initial state <case>_init.bin.  Needs a GPU.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;
using namespace gqten;

using U1QN = QN<U1QNVal>;
using IndexT = Index<U1QN>;
using QNSctT = QNSector<U1QN>;
using Tensor = GQTensor<GQTEN_Double, U1QN>;

static const size_t d = 2, D = 20, dh = 2;
static std::mt19937_64 g_rng(2019);

static std::vector<double> RandSym(size_t n) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<double> m(n * n);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = i; j < n; ++j) m[i * n + j] = m[j * n + i] = u(g_rng);
  return m;
}

// All elements of t in row-major order of its legs (a dim x dim matrix, or a
// vector when dim is 1).
static void WriteDense(const Tensor &t, const std::string &name, size_t dim) {
  auto sh = t.GetShape();
  std::vector<size_t> c(sh.size(), 0);
  std::vector<double> v;
  v.reserve(dim > 1 ? dim * dim : 0);
  while (true) {
    v.push_back(t.GetElem(c));
    int k = (int)sh.size() - 1;
    for (; k >= 0; --k) {
      if (++c[k] < sh[k]) break;
      c[k] = 0;
    }
    if (k < 0) break;
  }
  FILE *f = std::fopen((name + ".bin").c_str(), "wb");
  std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
}

int main() {
  Verbose() = false;
  U1QN qn0({QNCard("Sz", U1QNVal(0))});
  IndexT idx_din({QNSctT(qn0, d)}, GQTenIndexDirType::IN), idx_dout = InverseIndex(idx_din);
  IndexT idx_Din({QNSctT(qn0, D)}, GQTenIndexDirType::IN), idx_Dout = InverseIndex(idx_Din);
  IndexT idx_vin({QNSctT(qn0, dh)}, GQTenIndexDirType::IN), idx_vout = InverseIndex(idx_vin);

  // ---- centre bond (test_lanczos_solver.cc:121-168)
  {
    Tensor lblock({idx_Dout, idx_vout, idx_Din}), lsite({idx_vin, idx_din, idx_dout, idx_vout}),
        rblock({idx_Din, idx_vin, idx_Dout});
    auto bm = RandSym(D);
    for (size_t i = 0; i < D; ++i)
      for (size_t j = 0; j < D; ++j)
        for (size_t k = 0; k < dh; ++k) {
          lblock({i, k, j}) = bm[i * D + j];
          rblock({j, k, i}) = bm[i * D + j];
        }
    auto sm = RandSym(d);
    for (size_t i = 0; i < d; ++i)
      for (size_t j = 0; j < d; ++j)
        for (size_t k = 0; k < dh; ++k) lsite({k, i, j, k}) = sm[i * d + j];
    Tensor rsite(lsite);
    std::vector<Tensor *> H = {&lblock, &lsite, &rsite, &rblock};
    Tensor t1, t2, heff;
    Contract(H[0], H[1], {{1}, {0}}, &t1);
    Contract(&t1, H[2], {{4}, {0}}, &t2);
    Contract(&t2, H[3], {{6}, {1}}, &heff);
    heff.Transpose({0, 2, 4, 6, 1, 3, 5, 7});
    const size_t dim = D * d * d * D;
    WriteDense(heff, "cent", dim);
    const char *names[2] = {"cent_err", "cent_maxiter"};
    LanczosParams params[2] = {LanczosParams(1.0E-9), LanczosParams(1.0E-16, 20)};
    for (int c = 0; c < 2; ++c) {
      auto *init = new Tensor({idx_Din, idx_dout, idx_dout, idx_Dout});
      init->Random(qn0);
      WriteDense(*init, std::string(names[c]) + "_init", 1);
      auto res = LanczosSolver(H, init, params[c], "cent");
      std::printf("CASE %s E=%.17g iters=%zu dim=%zu file=cent\n", names[c], res.gs_eng, res.iters, dim);
      delete res.gs_vec;
    }
  }
  // ---- left end (test_lanczos_solver.cc:270-340)
  {
    Tensor lsite({idx_din, idx_vout, idx_dout}), rsite({idx_vin, idx_din, idx_dout, idx_vout}),
        rblock({idx_Din, idx_vin, idx_Dout});
    auto bm = RandSym(D);
    for (size_t i = 0; i < D; ++i)
      for (size_t j = 0; j < D; ++j)
        for (size_t k = 0; k < dh; ++k) rblock({j, k, i}) = bm[i * D + j];
    auto sm = RandSym(d);
    for (size_t i = 0; i < d; ++i)
      for (size_t j = 0; j < d; ++j)
        for (size_t k = 0; k < dh; ++k) {
          lsite({i, k, j}) = sm[i * d + j];
          rsite({k, i, j, k}) = sm[i * d + j];
        }
    Tensor null_ten;
    std::vector<Tensor *> H = {&null_ten, &lsite, &rsite, &rblock};
    Tensor t1, heff;
    Contract(H[1], H[2], {{1}, {0}}, &t1);
    Contract(&t1, H[3], {{4}, {1}}, &heff);
    heff.Transpose({0, 2, 4, 1, 3, 5});
    const size_t dim = d * d * D;
    WriteDense(heff, "lend", dim);
    auto *init = new Tensor({idx_dout, idx_dout, idx_Dout});
    init->Random(qn0);
    auto res = LanczosSolver(H, init, LanczosParams(1.0E-9), "lend");
    std::printf("CASE lend_err E=%.17g iters=%zu dim=%zu file=lend\n", res.gs_eng, res.iters, dim);
    delete res.gs_vec;
  }
  std::printf("DONE\n");
  return 0;
}
