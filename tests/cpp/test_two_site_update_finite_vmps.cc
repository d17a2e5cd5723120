// Spin-chain cases of tests/test_algorithm/test_two_site_update_finite_vmps.cc:91-222
// (GraceQ/MPS2 0.2-alpha.0), double-precision cases, written against the
// gqmps2-compatible headers as a user of the reference writes them:
//   * 1D Ising chain (6 sites): MPS dumped to disk, TwoSiteFiniteVMPS reads it,
//     E0 = -0.25 (N-1); then the MPS is loaded again and measured with
//     MeasureOneSiteOp / MeasureTwoSiteOp;
//   * 1D Heisenberg chain (6 sites, D = 8): E0 = -2.493577133888 to 1e-12,
//     then the "continue simulation" case: the container is cleared and a
//     second TwoSiteFiniteVMPS call resumes from the MPS files;
//   * sweep-level checkpoint (extension): a 10-site run stopped inside sweep
//     3 resumes from the files of sweep 2 and reproduces sweeps 3 and 4.
// Needs a GPU.  Exit status 0 = all cases pass.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;
using namespace gqten;

using U1QN = QN<U1QNVal>;
using IndexT = Index<U1QN>;
using QNSctT = QNSector<U1QN>;
using DGQTensor = GQTensor<GQTEN_Double, U1QN>;
using DSiteVec = SiteVec<GQTEN_Double, U1QN>;
using DMPS = FiniteMPS<GQTEN_Double, U1QN>;

static int g_fail = 0, g_checks = 0;

static void ExpectNear(double got, double want, double tol, const char *what) {
  ++g_checks;
  if (!(std::fabs(got - want) <= tol)) {
    ++g_fail;
    std::printf("FAIL %s: got %.15g want %.15g (tol %.1e)\n", what, got, want, tol);
  }
}
static void ExpectTrue(bool ok, const char *what) {
  ++g_checks;
  if (!ok) {
    ++g_fail;
    std::printf("FAIL %s\n", what);
  }
}

static void RemoveFolder(const std::string &p) { std::filesystem::remove_all(p); }

// test_two_site_update_finite_vmps.cc:79-88
static void RunCase(DMPS &mps, const MPO<DGQTensor> &mpo, const SweepParams &sp, double e_ref, double tol,
                    const char *what) {
  double e0 = TwoSiteFiniteVMPS(mps, mpo, sp);
  ExpectNear(e0, e_ref, tol, what);
  ExpectTrue(mps.empty(), "MPS container empty after the run (tensors left on disk)");
}

int main() {
  Verbose() = false;
  const size_t N = 6;
  U1QN qn0({QNCard("Sz", U1QNVal(0))});
  IndexT pb_out({QNSctT(U1QN({QNCard("Sz", U1QNVal(1))}), 1), QNSctT(U1QN({QNCard("Sz", U1QNVal(-1))}), 1)},
                GQTenIndexDirType::OUT);
  IndexT pb_in = InverseIndex(pb_out);
  DSiteVec sites(N, pb_out);
  DGQTensor did({pb_in, pb_out}), dsz({pb_in, pb_out}), dsp({pb_in, pb_out}), dsm({pb_in, pb_out});
  did({0, 0}) = 1;
  did({1, 1}) = 1;
  dsz({0, 0}) = 0.5;
  dsz({1, 1}) = -0.5;
  dsp({0, 1}) = 1;
  dsm({1, 0}) = 1;
  DMPS dmps(sites);
  std::vector<size_t> stat_labs;
  for (size_t i = 0; i < N; ++i) stat_labs.push_back(i % 2);

  // ---- 1D Ising (test_two_site_update_finite_vmps.cc:135-165)
  {
    MPOGenerator<GQTEN_Double, U1QN> gen(sites, qn0);
    for (size_t i = 0; i + 1 < N; ++i) gen.AddTerm(1, {dsz, dsz}, {i, i + 1});
    auto mpo = gen.Gen();
    SweepParams sp(4, 1, 10, 1.0E-5, LanczosParams(1.0E-7));
    DirectStateInitMps(dmps, stat_labs, qn0);
    dmps.Dump(sp.mps_path, true);
    RunCase(dmps, mpo, sp, -0.25 * (N - 1), 1.0E-10, "Ising E0");

    dmps.Load(sp.mps_path);
    auto sz = MeasureOneSiteOp(dmps, dsz, "dsz");
    double tot = 0;
    for (auto &r : sz) tot += r.avg;
    ExpectNear(tot, 0.0, 1e-10, "Ising sum_i <Sz_i> (Sz = 0 sector)");
    std::vector<std::vector<size_t>> sites_set;
    for (size_t i = 0; i < N; ++i)
      for (size_t j = i + 1; j < N; ++j) sites_set.push_back({i, j});
    auto zz = MeasureTwoSiteOp(dmps, {dsz, dsz}, did, sites_set, "dszdsz");
    // every Neel ground state (and superposition of the two) has <Sz_i Sz_j> = (-1)^(j-i) / 4
    for (auto &r : zz) {
      double want = ((r.sites[1] - r.sites[0]) % 2 ? -0.25 : 0.25);
      ExpectNear(r.avg, want, 1e-9, "Ising <Sz_i Sz_j>");
    }
    dmps.clear();
    RemoveFolder(sp.mps_path);
    RemoveFolder(sp.temp_path);
  }

  // ---- 1D Heisenberg + continue simulation (test_two_site_update_finite_vmps.cc:185-222)
  {
    MPOGenerator<GQTEN_Double, U1QN> gen(sites, qn0);
    for (size_t i = 0; i + 1 < N; ++i) {
      gen.AddTerm(1, {dsz, dsz}, {i, i + 1});
      gen.AddTerm(0.5, {dsp, dsm}, {i, i + 1});
      gen.AddTerm(0.5, {dsm, dsp}, {i, i + 1});
    }
    auto mpo = gen.Gen();
    SweepParams sp(4, 8, 8, 1.0E-9, LanczosParams(1.0E-7));
    DirectStateInitMps(dmps, stat_labs, qn0);
    dmps.Dump(sp.mps_path, true);
    RunCase(dmps, mpo, sp, -2.493577133888, 1.0E-12, "Heisenberg E0");
    ExpectTrue(IsPathExist(sp.mps_path + "/mps_ten0.gqten") && IsPathExist(sp.temp_path),
               "MPS files and runtime directory exist after the run");
    dmps.clear();
    RunCase(dmps, mpo, sp, -2.493577133888, 1.0E-12, "Heisenberg E0, continued from the MPS files");
    RemoveFolder(sp.mps_path);
    RemoveFolder(sp.temp_path);
  }

  // ---- sweep-level checkpoint: a run stopped inside sweep 3 continues from
  // the MPS files written at the end of sweep 2 and reproduces sweeps 3 and 4
  // of the uninterrupted run (an own extension; the reference streams the MPS
  // to disk around every update)
  {
    const size_t M = 10;
    DSiteVec sites10(M, pb_out);
    MPOGenerator<GQTEN_Double, U1QN> gen(sites10, qn0);
    for (size_t i = 0; i + 1 < M; ++i) {
      gen.AddTerm(1, {dsz, dsz}, {i, i + 1});
      gen.AddTerm(0.5, {dsp, dsm}, {i, i + 1});
      gen.AddTerm(0.5, {dsm, dsp}, {i, i + 1});
    }
    auto mpo = gen.Gen();
    std::vector<size_t> labs10;
    for (size_t i = 0; i < M; ++i) labs10.push_back(i % 2);
    const size_t per_sweep = 2 * (M - 1);
    SweepParams sp(4, 1, 6, 0.0, LanczosParams(1.0E-10), "ckpt_mps", "ckpt_tmp");
    std::vector<double> ref, got;
    DMPS a(sites10);
    std::vector<size_t> labs10b;
    for (size_t i = 0; i < M; ++i) labs10b.push_back((i + 1) % 2);
    SetRandomSeed(99);
    ExtendDirectRandomInitMps(a, {labs10, labs10b}, qn0, 3);
    a.Dump(sp.mps_path, true);
    TwoSiteFiniteVMPS(a, mpo, sp, [&](const UpdateRecord &r) { ref.push_back(r.energy); });
    ExpectTrue(ref.size() == 4 * per_sweep, "uninterrupted run: four sweeps of updates");
    RemoveFolder(sp.mps_path);
    DMPS b(sites10);
    SetRandomSeed(99);
    ExtendDirectRandomInitMps(b, {labs10, labs10b}, qn0, 3);
    b.Dump(sp.mps_path, true);
    struct Stop {};
    bool stopped = false;
    try {
      TwoSiteFiniteVMPS(b, mpo, sp, [&](const UpdateRecord &r) {
        got.push_back(r.energy);
        if (got.size() == 2 * per_sweep + 3) throw Stop{};  // inside sweep 3
      });
    } catch (const Stop &) {
      stopped = true;
    }
    ExpectTrue(stopped, "the run stopped inside sweep 3");
    DMPS c(sites10);
    SweepParams sp2 = sp;
    sp2.sweeps = 2;
    std::vector<double> resumed;
    TwoSiteFiniteVMPS(c, mpo, sp2, [&](const UpdateRecord &r) { resumed.push_back(r.energy); });
    ExpectTrue(resumed.size() == 2 * per_sweep, "resumed run: two sweeps of updates");
    double worst = 0;
    for (size_t k = 0; k < resumed.size() && 2 * per_sweep + k < ref.size(); ++k)
      worst = std::max(worst, std::fabs(resumed[k] - ref[2 * per_sweep + k]));
    ExpectNear(worst, 0.0, 1e-12, "resumed sweeps 3 and 4 reproduce the uninterrupted run, every update");
    for (size_t k = 0; k + 1 < got.size() && k < ref.size(); ++k)
      worst = std::max(worst, std::fabs(got[k] - ref[k]));
    ExpectNear(worst, 0.0, 1e-12, "the stopped run agrees with the uninterrupted one before the stop");
    RemoveFolder(sp.mps_path);
    RemoveFolder(sp.temp_path);
  }

  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
