// Autotuner of the grouped-GEMM launches of contraction plans.
//
// Candidates are kernels whose per-element accumulation order is the same
// (same slab depth and MFMA step order), so the choice changes the time of a
// launch, never its result: for FP32, 128x128 tiles with register-staged
// slabs, 128x128 tiles with direct-to-LDS slabs (row-major operands only) and
// 64x64 tiles; for split-BF16x3, 128x128 and 64x64 tiles.  FP64 launches keep
// 64x64 tiles.  Decisions are kept per arithmetic class (FP64, FP32,
// split-BF16x3).
// The first launches of a large plan run each candidate twice in turn, timed
// with HIP events on the context stream (no host sync: the events are read
// once the last one has completed, until then the heuristic kernel runs).
// The fastest candidate is kept for the plan and, under the plan's shape key
// (a hash of the histogram of its output blocks' M, N and summed K in
// quarter-octave bins, blocks below 1e-3 of the launch's flops left out, with
// the operand layouts), shared by every later plan of the same shape in
// the process and appended to a cache file ($MPS2_TUNE_CACHE, else
// $XDG_CACHE_HOME or ~/.cache, mps2_amd/gemm_tune_gfx950.txt) that later
// processes read.  MPS2_TUNE=0, or mps2_ctx_set_tuning(ctx, 0), keeps the
// heuristic choice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>

#include "runtime.h"

namespace mps2 {

namespace {

constexpr double kTuneMinFlops = 1e9;   // smaller launches keep the heuristic
constexpr int kTuneRounds = 2;          // timed runs per candidate

struct CacheEntry {
  GemmChoice choice;
  double best_ms = 0, heuristic_ms = 0;
};

std::mutex g_tune_mu;
// (shape key, class code) -> choice; class codes as in the cache file:
// 1 = FP64, 0 = FP32, 2 = split-BF16x3.
std::map<std::pair<uint64_t, int>, CacheEntry> g_tune;
bool g_tune_loaded = false;

std::string cache_path() {
  if (const char *e = std::getenv("MPS2_TUNE_CACHE")) return e;
  std::string base;
  if (const char *x = std::getenv("XDG_CACHE_HOME")) base = x;
  else if (const char *h = std::getenv("HOME")) base = std::string(h) + "/.cache";
  else return "";
  return base + "/mps2_amd/gemm_tune_gfx950.txt";
}

void load_cache_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const std::string p = cache_path();
  if (p.empty()) return;
  std::ifstream f(p);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream s(line);
    std::string key;
    int code = 0, tile = 0, variant = 0;
    double best = 0, heur = 0;
    if (!(s >> key >> code >> tile >> variant >> best >> heur)) continue;
    if ((tile != 64 && tile != 128) || (variant != 0 && variant != 4 && variant != 9) || code < 0 || code > 2) continue;
    CacheEntry ce;
    ce.choice = GemmChoice{tile, variant};
    ce.best_ms = best;
    ce.heuristic_ms = heur;
    g_tune[{std::strtoull(key.c_str(), nullptr, 16), code}] = ce;
  }
}

void append_cache(uint64_t key, int code, const CacheEntry &ce) {
  const std::string p = cache_path();
  if (p.empty()) return;
  const auto slash = p.find_last_of('/');
  if (slash != std::string::npos) {
    // create the directory chain (errors leave the cache in memory only)
    std::string dir = p.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
  }
  std::ofstream f(p, std::ios::app);
  if (!f) return;
  char buf[160];
  std::snprintf(buf, sizeof buf, "%016llx %d %d %d %.6f %.6f\n", (unsigned long long)key, code, ce.choice.tile,
                ce.choice.variant, ce.best_ms, ce.heuristic_ms);
  f << buf;
}

bool tuning_enabled_env() {
  static const bool on = [] {
    const char *e = std::getenv("MPS2_TUNE");
    return !(e && std::string(e) == "0");
  }();
  return on;
}

// Log bin of a size, four bins per octave: neighbouring bonds of a sweep,
// whose sector dimensions differ by a few percent, share a key.
int64_t lbin(int64_t x) { return x <= 0 ? -1 : (int64_t)std::lround(4.0 * std::log2((double)x)); }

uint64_t shape_key(const ContractPlan &hp, int code) {
  std::vector<std::array<int64_t, 3>> s;
  s.reserve(hp.blks.size());
  for (const BlkH &b : hp.blks) {
    int64_t kc = 0;
    for (int p = b.pair_begin; p < b.pair_end; ++p) kc += hp.pairs[p].K;
    // blocks too small to matter for the launch time stay out of the key
    if (2.0 * b.M * (double)b.N * (double)kc < 1e-3 * hp.flops) continue;
    s.push_back({lbin(b.M), lbin(b.N), lbin(kc)});
  }
  std::sort(s.begin(), s.end());
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (v >> (8 * i)) & 0xff;
      h *= 1099511628211ull;
    }
  };
  mix(code == 1 ? 1 : code == 0 ? 2 : 3);
  mix((uint64_t)hp.la);
  mix((uint64_t)hp.lb);
  mix(s.size());
  for (auto &x : s)
    for (int64_t v : x) mix((uint64_t)v);
  return h;
}

// Arithmetic class of an FP64 / FP32 launch on context c: index into
// DevPlan::tune (0 FP64, 1 FP32, 2 split-BF16x3) and the cache file's code.
int tune_index(const Ctx *c, bool fp64) { return fp64 ? 0 : (c->f32_split ? 2 : 1); }
int class_code(int idx) { return idx == 0 ? 1 : idx == 1 ? 0 : 2; }

}  // namespace

PlanTune::~PlanTune() {
  for (auto &e : ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
}

std::vector<GemmChoice> tune_candidates(const DevPlan &dp, bool fp64, bool split, int num_cu) {
  std::vector<GemmChoice> c;
  // Split-K in four parts, untuned (a fixed rule, so the result does not
  // depend on timing), for launches with long K (mean 2048 or more) whose
  // last round of workgroups would leave the compute units idle: four parts
  // per tile when that shortens the rounds of resident workgroups (slots:
  // three FP32 128x128 or four FP64 64x64 workgroups per compute unit) by
  // 10% or more.  The centre-bond stage 4 at D=4096 (978 FP32 tiles, 3318
  // FP64 tiles, K ~ 16 thousand): FP32 4.64 -> 4.14 ms (2 parts: 4.50 ms).
  // (the three partial copies stay under 2 GiB of scratch)
  const bool split_fits = dp.split.nblk <= 65535 && 3 * dp.split.out_span * (fp64 ? 8 : 4) <= (int64_t(1) << 31);
  auto split_pays = [&](int64_t tiles, int64_t slots) {
    if (tiles <= 0 || slots <= 0 || dp.split.avg_k < 2048 || !split_fits) return false;
    const double r1 = std::ceil((double)tiles / slots), r4 = std::ceil(4.0 * tiles / slots) / 4.0;
    return r4 < 0.9 * r1;
  };
  if (!fp64 && !split && dp.tile_m == 128 && split_pays(dp.ntiles128, 3 * (int64_t)num_cu)) {
    c.push_back(GemmChoice{128, 9});
    return c;
  }
  if (fp64 && split_pays(dp.ntiles64, 4 * (int64_t)num_cu)) {
    c.push_back(GemmChoice{64, 9});
    return c;
  }
  // Split-BF16x3 (two workgroups per compute unit) is bound by the operand
  // panels' L2 misses; four parts per tile put the same part of neighbouring
  // tiles on one XCD at once, so very long K pays there by locality alone
  // (the centre-bond stage 4: 2.46 -> 2.21 ms).
  if (split && dp.tile_m == 128 && dp.split.avg_k >= 4096 && dp.ntiles128 < 8 * (int64_t)num_cu && split_fits) {
    c.push_back(GemmChoice{128, 9});
    return c;
  }
  const GemmChoice h = fp64 ? GemmChoice{64, 0} : GemmChoice{dp.tile_m, 0};
  c.push_back(h);
  if (split) {
    c.push_back(GemmChoice{h.tile == 128 ? 64 : 128, 0});
    return c;
  }
  // FP64 keeps its 64x64 tiles: 128x128 tiles allow one workgroup per compute
  // unit and measured 19% slower on the D=4096 matvec and slower in every
  // probe of a sweep, so timing them only cost time.
  if (!fp64) {
    if (h.tile != 128) c.push_back(GemmChoice{128, 0});
    if (dp.hp->la == kNormal && dp.hp->lb == kNormal) c.push_back(GemmChoice{128, 4});
    if (h.tile != 64) c.push_back(GemmChoice{64, 0});
  }
  return c;
}

GemmChoice tune_select(Ctx *c, const DevPlan &dp, bool fp64, PlanTune::Probe *probe) {
  const int idx = tune_index(c, fp64), code = class_code(idx);
  PlanTune &t = dp.tune[idx];
  probe->active = false;
  if (t.cand.empty()) t.cand = tune_candidates(dp, fp64, idx == 2, c->num_cu);
  if (!tuning_enabled_env() || !c->tune_mode) return t.cand[0];
  if (t.decided >= 0) return t.cand[t.decided];
  if (dp.hp->flops < kTuneMinFlops || t.cand.size() < 2) return t.cand[0];
  if (!t.keyed) {
    t.key = shape_key(*dp.hp, code);
    t.keyed = true;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    load_cache_locked();
    auto it = g_tune.find({t.key, code});
    if (it != g_tune.end())
      for (size_t i = 0; i < t.cand.size(); ++i)
        if (t.cand[i].tile == it->second.choice.tile && t.cand[i].variant == it->second.choice.variant) {
          t.decided = (int)i;
          c->count("tune_cache_hits");
          return t.cand[i];
        }
  }
  const int total = (int)t.cand.size() * kTuneRounds;
  if (t.next < total) {
    const int k = t.next++ % (int)t.cand.size();
    probe->active = true;
    probe->cand = k;
    c->count("tune_probes");
    return t.cand[k];
  }
  // every probe issued: decide once the last one has completed
  if (hipEventQuery(t.ev.back().second) != hipSuccess) {
    (void)hipGetLastError();
    return t.cand[0];
  }
  std::vector<double> best(t.cand.size(), 1e300);
  for (size_t i = 0; i < t.ev.size(); ++i) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.ev[i].first, t.ev[i].second) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    best[t.ev_cand[i]] = std::min(best[t.ev_cand[i]], (double)ms);
  }
  int w = 0;
  for (size_t i = 1; i < best.size(); ++i)
    if (best[i] < best[w]) w = (int)i;
  t.decided = w;
  t.best_ms = best[w];
  t.heuristic_ms = best[0];
  c->count("tune_decisions");
  if (w != 0) c->count("tune_changed");
  CacheEntry ce;
  ce.choice = t.cand[w];
  ce.best_ms = best[w];
  ce.heuristic_ms = best[0];
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[{t.key, code}] = ce;
    append_cache(t.key, code, ce);
  }
  return t.cand[w];
}

void tune_probe_begin(Ctx *c, const DevPlan &dp, bool fp64, PlanTune::Probe *probe) {
  if (!probe->active) return;
  PlanTune &t = dp.tune[tune_index(c, fp64)];
  hipEvent_t a, b;
  MPS2_HIP(hipEventCreate(&a));
  MPS2_HIP(hipEventCreate(&b));
  t.ev.emplace_back(a, b);
  t.ev_cand.push_back(probe->cand);
  MPS2_HIP(hipEventRecord(a, c->stream));
}

void tune_probe_end(Ctx *c, const DevPlan &dp, bool fp64, PlanTune::Probe *probe) {
  if (!probe->active) return;
  MPS2_HIP(hipEventRecord(dp.tune[tune_index(c, fp64)].ev.back().second, c->stream));
}

}  // namespace mps2
