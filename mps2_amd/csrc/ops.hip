// Block-sparse tensor operations on the device: Contract (permute + grouped
// GEMM over a cached plan), LinearCombine, dot / norm / Normalize, layout
// embedding, and the effective-Hamiltonian matvec.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <tuple>

#include "runtime.h"

namespace mps2 {

namespace {

BufP upload_bytes(Ctx *c, const void *host, size_t bytes) { return c->upload(host, bytes); }

using PlanKey = std::tuple<Ctx *, const Meta *, const Meta *, std::vector<int>, std::vector<int>, const Meta *, bool>;

struct PlanEntry {
  std::shared_ptr<DevPlan> plan;
  MetaP keep_out;   // retarget layout kept alive while cached
  MetaP ka, kb;     // the key's operand layouts, kept alive (their addresses are the key)
};

// Process-wide plan cache (keys carry the context), FIFO eviction; every
// access holds g_plan_mu, so contexts may be driven from different threads.
std::mutex g_plan_mu;
std::map<PlanKey, PlanEntry> g_plans;
std::deque<PlanKey> g_plan_order;
constexpr size_t kPlanCacheCap = 512;

// Second level, keyed by layout content: a tensor made anew with a layout
// seen before (the same bond in a later sweep, once the kept sector
// dimensions have settled) finds its plan, tile map and autotuner state
// again.  Hits are confirmed by comparing the layouts themselves
// (Meta::same_layout); the digests only index.  MPS2_PLAN_CONTENT=0 turns
// this level off (every new layout object then builds its plan).
using ContentKey = std::tuple<Ctx *, uint64_t, uint64_t, std::vector<int>, std::vector<int>, uint64_t, bool>;
struct ContentEntry {
  std::shared_ptr<DevPlan> plan;
  MetaP target;
};
std::map<ContentKey, ContentEntry> g_cplans;
std::deque<ContentKey> g_cplan_order;
constexpr size_t kContentCacheCap = 4096;

void perm_descs(const Meta &m, const std::vector<int> &perm, std::vector<PermDesc> &out, std::vector<int64_t> &starts) {
  int r = m.rank();
  out.resize(m.coords.size());
  starts.resize(m.coords.size());
  for (size_t b = 0; b < m.coords.size(); ++b) {
    PermDesc d;
    std::memset(&d, 0, sizeof(d));
    d.src_off = m.offset[b];
    d.dst_off = m.offset[b];
    d.size = m.bsize[b];
    d.rank = r;
    int64_t st[kMaxRank];
    int64_t s = 1;
    for (int i = r - 1; i >= 0; --i) {
      st[i] = s;
      s *= m.legs[i].dim[m.coords[b][i]];
    }
    for (int i = 0; i < r; ++i) {
      d.dims[i] = m.legs[perm[i]].dim[m.coords[b][perm[i]]];
      d.sstride[i] = st[perm[i]];
    }
    out[b] = d;
    starts[b] = d.dst_off;
  }
}

// The plan of a complex contraction in real terms.  A is read as an M x 2K
// real matrix (row-major complex, contracted legs trailing: an A stored with
// them leading is permuted first), B as its 2K x 2N real form built per call
// (k_expand_cplx, four reals per element), C written as M x 2N reals; pair and
// block offsets, K and N are rescaled accordingly, and the flops are the real
// ones (four times 2 M N K).  exp lists the expansion of each B block used.
std::shared_ptr<ContractPlan> complex_view(const std::shared_ptr<ContractPlan> &hp, std::vector<CplxExpand> &exp,
                                           int64_t &max_kn) {
  auto v = std::make_shared<ContractPlan>(*hp);
  if (v->la == kTrans) {
    std::vector<int> afree;
    for (int i = 0; i < v->a->rank(); ++i)
      if (std::find(v->ax.begin(), v->ax.end(), i) == v->ax.end()) afree.push_back(i);
    v->permute_a = true;
    v->aperm = afree;
    v->aperm.insert(v->aperm.end(), v->ax.begin(), v->ax.end());
  }
  v->la = kNormal;
  const int btrans = (!v->permute_b && v->lb == kTrans) ? 1 : 0;
  std::map<int64_t, int> seen;
  max_kn = 0;
  for (const BlkH &b : hp->blks)
    for (int p = b.pair_begin; p < b.pair_end; ++p) {
      const PairH &ph = hp->pairs[p];
      if (seen.count(ph.b_off)) continue;
      seen[ph.b_off] = (int)exp.size();
      exp.push_back(CplxExpand{ph.b_off, 4 * ph.b_off, ph.K, b.N, btrans, 0});
      max_kn = std::max<int64_t>(max_kn, (int64_t)ph.K * b.N);
    }
  v->lb = kNormal;
  for (auto &ph : v->pairs) {
    ph.a_off *= 2;
    ph.b_off *= 4;
    ph.K *= 2;
  }
  for (auto &b : v->blks) {
    b.c_off *= 2;
    b.N *= 2;
  }
  v->flops *= 4;
  return v;
}

std::shared_ptr<DevPlan> build_dev_plan(Ctx *c, const std::shared_ptr<ContractPlan> &hp0, const Meta *target,
                                        bool cplx) {
  auto dp = std::make_shared<DevPlan>();
  std::shared_ptr<ContractPlan> hp = hp0;
  if (cplx) {
    std::vector<CplxExpand> exp;
    hp = complex_view(hp0, exp, dp->expand_max_kn);
    dp->cplx = true;
    dp->nexpand = (int)exp.size();
    if (!exp.empty()) dp->expand = upload_bytes(c, exp.data(), exp.size() * sizeof(CplxExpand));
  }
  dp->hp = hp;
  // Tile shape: 128x128 tiles move half the operand bytes per flop of 64x64
  // tiles; FP32 uses them unless edge padding inflates the work by more than
  // 30% or there are too few of them to occupy every compute unit.  FP64
  // always runs 64x64 tiles (a second list): at 128x128 its registers allow
  // one workgroup per compute unit, and 64x64 measured 19% faster on the
  // D=4096 matvec.
  double w64 = 0, w128 = 0;
  int64_t t128 = 0;
  for (auto &b : hp->blks) {
    double kc = 0;
    for (int p = b.pair_begin; p < b.pair_end; ++p) kc += hp->pairs[p].K;
    int64_t n64 = (int64_t)((b.M + 63) / 64) * ((b.N + 63) / 64);
    int64_t n128 = (int64_t)((b.M + 127) / 128) * ((b.N + 127) / 128);
    w64 += (double)n64 * 4096.0 * kc;
    w128 += (double)n128 * 16384.0 * kc;
    t128 += n128;
  }
  int ts = (w128 <= 1.3 * w64 && t128 >= c->num_cu) ? 128 : 64;
  dp->tile_m = dp->tile_n = ts;
  std::vector<BlkD> blks(hp->blks.size());
  std::vector<PairD> pairs(hp->pairs.size());
  for (size_t i = 0; i < hp->pairs.size(); ++i) {
    pairs[i].a_off = hp->pairs[i].a_off;
    pairs[i].b_off = hp->pairs[i].b_off;
    pairs[i].K = hp->pairs[i].K;
    pairs[i].pad = 0;
  }
  std::vector<int64_t> kcs(hp->blks.size(), 0);
  for (size_t i = 0; i < hp->blks.size(); ++i) {
    const BlkH &b = hp->blks[i];
    blks[i].c_off = b.c_off;
    if (target) {
      const Coord &oc = hp->out->coords[i];
      int32_t tb = target->find(oc);
      if (tb < 0) throw Error(kErrArg, "contract_into: target layout lacks an output block");
      blks[i].c_off = target->offset[tb] * (cplx ? 2 : 1);
    }
    blks[i].M = b.M;
    blks[i].N = b.N;
    blks[i].pb = b.pair_begin;
    blks[i].pe = b.pair_end;
    for (int p = b.pair_begin; p < b.pair_end; ++p) kcs[i] += hp->pairs[p].K;
  }
  int max_n = 0;
  for (auto &b : blks) max_n = std::max(max_n, (int)b.N);
  // MPS2_NARROW=0 keeps narrow contractions on the matrix-core kernels.
  static const bool narrow_on = [] {
    const char *e = std::getenv("MPS2_NARROW");
    return !(e && std::string(e) == "0");
  }();
  if (narrow_on && max_n <= kNarrowMaxN && !hp->blks.empty())
    dp->narrow_nc = max_n <= 4 ? 4 : (max_n <= 8 ? 8 : 16);
  // Tiles: the blocks in launch order (heaviest first) and, per tile shape,
  // the prefix sums of their tile counts, decoded on the device; explicit
  // lists only for an XCD-grouped order (schedule_tiles, gemm.hip).
  const std::vector<int32_t> order = tile_block_order(blks, kcs);
  if (!order.empty() || blks.empty()) {
    const int nb = (int)blks.size();
    std::vector<int32_t> h((size_t)nb + 3 * ((size_t)nb + 1), 0);
    for (int i = 0; i < nb; ++i) h[i] = order[i];
    int32_t *s64 = h.data() + nb, *s128 = s64 + nb + 1, *sn = s128 + nb + 1;
    const int nc = dp->narrow_nc ? dp->narrow_nc : 1;
    for (int i = 0; i < nb; ++i) {
      const BlkD &bk = blks[order[i]];
      auto cnt = [&](int tm, int tn) {
        return (int64_t)((bk.M + tm - 1) / tm) * ((bk.N + tn - 1) / tn);
      };
      const int64_t n64 = s64[i] + cnt(64, 64), n128 = s128[i] + cnt(128, 128), nn = sn[i] + cnt(kNarrowRows, nc);
      if (n64 > INT32_MAX || n128 > INT32_MAX || nn > INT32_MAX) throw Error(kErrArg, "contraction: too many tiles");
      s64[i + 1] = (int32_t)n64;
      s128[i + 1] = (int32_t)n128;
      sn[i + 1] = (int32_t)nn;
    }
    dp->tilemap = upload_bytes(c, h.data(), h.size() * sizeof(int32_t));
    const int32_t *d = static_cast<const int32_t *>(dp->tilemap->p);
    auto mk = [&](const int32_t *start, int tm, int tn) {
      TileMap m;
      m.order = d;
      m.start = start;
      m.nb = nb;
      m.tm = tm;
      m.tn = tn;
      return m;
    };
    dp->map64 = mk(d + nb, 64, 64);
    dp->map128 = mk(d + 2 * nb + 1, 128, 128);
    dp->map_narrow = mk(d + 3 * nb + 2, kNarrowRows, nc);
    dp->ntiles64 = s64[nb];
    dp->ntiles128 = s128[nb];
    dp->ntiles_narrow = sn[nb];
  } else {
    auto upload_tiles = [&](int edge, int &count, TileMap &m) {
      std::vector<TileD> t = schedule_tiles(blks, kcs, edge);
      count = (int)t.size();
      BufP buf = upload_bytes(c, t.data(), t.size() * sizeof(TileD));
      m.list = static_cast<const TileD *>(buf->p);
      return buf;
    };
    dp->tiles64 = upload_tiles(64, dp->ntiles64, dp->map64);
    dp->tiles128 = upload_tiles(128, dp->ntiles128, dp->map128);
    if (dp->narrow_nc) {
      std::vector<TileD> t = narrow_tiles(blks, dp->narrow_nc);
      dp->ntiles_narrow = (int)t.size();
      dp->tilemap = upload_bytes(c, t.data(), t.size() * sizeof(TileD));
      dp->map_narrow.list = static_cast<const TileD *>(dp->tilemap->p);
    }
  }
  for (const BlkD &bk : blks) {
    const int64_t mn = (int64_t)bk.M * bk.N;
    dp->split.out_span = std::max(dp->split.out_span, bk.c_off + mn);
    dp->split.max_mn = std::max(dp->split.max_mn, mn);
  }
  dp->split.nblk = (int)blks.size();
  {
    double mn = 0;
    for (const BlkD &bk : blks) mn += (double)bk.M * bk.N;
    dp->split.avg_k = mn > 0 ? hp->flops / (2.0 * mn) : 0.0;
  }
  dp->blks = upload_bytes(c, blks.data(), blks.size() * sizeof(BlkD));
  dp->pairs = upload_bytes(c, pairs.data(), pairs.size() * sizeof(PairD));
  if (hp->permute_a) {
    std::vector<PermDesc> d;
    std::vector<int64_t> s;
    perm_descs(*hp->a, hp->aperm, d, s);
    dp->perm_a = upload_bytes(c, d.data(), d.size() * sizeof(PermDesc));
    dp->perm_a_start = upload_bytes(c, s.data(), s.size() * sizeof(int64_t));
    dp->nperm_a = (int)d.size();
  }
  if (hp->permute_b) {
    std::vector<PermDesc> d;
    std::vector<int64_t> s;
    perm_descs(*hp->b, hp->bperm, d, s);
    dp->perm_b = upload_bytes(c, d.data(), d.size() * sizeof(PermDesc));
    dp->perm_b_start = upload_bytes(c, s.data(), s.size() * sizeof(int64_t));
    dp->nperm_b = (int)d.size();
  }
  return dp;
}

std::shared_ptr<DevPlan> get_plan_t(Ctx *c, const MetaP &a, const MetaP &b, const std::vector<int> &ax,
                                    const std::vector<int> &bx, const MetaP &target, bool cplx = false) {
  PlanKey key{c, a.get(), b.get(), ax, bx, target.get(), cplx};
  std::lock_guard<std::mutex> lock(g_plan_mu);
  auto it = g_plans.find(key);
  if (it != g_plans.end() && it->second.ka == a && it->second.kb == b) return it->second.plan;
  const ContentKey ck{c, a->digest, b->digest, ax, bx, target ? target->digest : 0, cplx};
  std::shared_ptr<DevPlan> dp;
  static const bool content_on = [] {
    const char *e = std::getenv("MPS2_PLAN_CONTENT");
    return !(e && std::string(e) == "0");
  }();
  auto ci = content_on ? g_cplans.find(ck) : g_cplans.end();
  if (ci != g_cplans.end()) {
    const ContractPlan &h = *ci->second.plan->hp;
    const bool same_target = target ? (ci->second.target && ci->second.target->same_layout(*target))
                                    : !ci->second.target;
    if (h.a->same_layout(*a) && h.b->same_layout(*b) && same_target) {
      dp = ci->second.plan;
      c->count("plan_content_hits");
    }
  }
  if (!dp) {
    const auto t0 = std::chrono::steady_clock::now();
    auto hp = plan_contract(a, b, ax, bx);
    dp = build_dev_plan(c, hp, target.get(), cplx);
    c->count("plan_builds");
    c->count("plan_build_ms",
             std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    if (!content_on) {
    } else if (ci != g_cplans.end()) {
      ci->second = ContentEntry{dp, target};
    } else {
      if (g_cplans.size() >= kContentCacheCap) {
        g_cplans.erase(g_cplan_order.front());
        g_cplan_order.pop_front();
      }
      g_cplans[ck] = ContentEntry{dp, target};
      g_cplan_order.push_back(ck);
    }
  }
  if (it != g_plans.end()) {
    // same addresses, new layouts (the old ones were freed): replace in place
    it->second = PlanEntry{dp, target, a, b};
    return dp;
  }
  if (g_plans.size() >= kPlanCacheCap) {
    g_plans.erase(g_plan_order.front());
    g_plan_order.pop_front();
  }
  g_plans[key] = PlanEntry{dp, target, a, b};
  g_plan_order.push_back(key);
  return dp;
}

template <typename T>
T *ptr(const TenP &t) {
  return static_cast<T *>(t->buf->p);
}

// Output blocks owned by rank r of g (longest-processing-time assignment on
// their flops, ties by block index) and their tiles.
const DevPlan::Share &plan_share(Ctx *c, const DevPlan &dp, int g, int r) {
  auto key = std::make_pair(g, r);
  auto it = dp.shares.find(key);
  if (it != dp.shares.end()) return it->second;
  const auto &blks = dp.hp->blks;
  std::vector<double> cost(blks.size(), 0.0);
  for (size_t i = 0; i < blks.size(); ++i) {
    double kc = 0;
    for (int p = blks[i].pair_begin; p < blks[i].pair_end; ++p) kc += dp.hp->pairs[p].K;
    cost[i] = 2.0 * blks[i].M * (double)blks[i].N * kc;
  }
  std::vector<int> idx(blks.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<double> load(g, 0.0);
  std::vector<int> owner(blks.size(), 0);
  for (int i : idx) {
    int best = 0;
    for (int k = 1; k < g; ++k)
      if (load[k] < load[best]) best = k;
    owner[i] = best;
    load[best] += cost[i];
  }
  DevPlan::Share sh;
  std::vector<TileD> tl;
  const int ts = dp.tile_m;
  for (size_t i = 0; i < blks.size(); ++i) {
    if (owner[i] != r) continue;
    sh.flops += cost[i];
    for (int m0 = 0; m0 < blks[i].M; m0 += ts)
      for (int n0 = 0; n0 < blks[i].N; n0 += ts) tl.push_back(TileD{(int32_t)i, m0, n0, 0});
  }
  sh.ntiles = (int)tl.size();
  if (!tl.empty()) sh.tiles = upload_bytes(c, tl.data(), tl.size() * sizeof(TileD));
  return dp.shares.emplace(key, std::move(sh)).first->second;
}

template <typename T>
void run_plan(Ctx *c, const DevPlan &dp, const TenP &a, const TenP &b, T *out) {
  const T *A = ptr<T>(a);
  const T *B = ptr<T>(b);
  BufP pa, pb, pe;
  if (dp.cplx) {
    // permute in complex elements, then B's real form
    if (dp.hp->permute_a) {
      pa = c->alloc(std::max<int64_t>(a->nr(), 2) * sizeof(double));
      k_permute<Cplx>(c, reinterpret_cast<const Cplx *>(A), static_cast<Cplx *>(pa->p),
                      static_cast<const PermDesc *>(dp.perm_a->p), static_cast<const int64_t *>(dp.perm_a_start->p),
                      dp.nperm_a, a->n());
      A = static_cast<const T *>(pa->p);
    }
    if (dp.hp->permute_b) {
      pb = c->alloc(std::max<int64_t>(b->nr(), 2) * sizeof(double));
      k_permute<Cplx>(c, reinterpret_cast<const Cplx *>(B), static_cast<Cplx *>(pb->p),
                      static_cast<const PermDesc *>(dp.perm_b->p), static_cast<const int64_t *>(dp.perm_b_start->p),
                      dp.nperm_b, b->n());
      B = static_cast<const T *>(pb->p);
    }
    pe = c->alloc(std::max<int64_t>(4 * b->n(), 1) * sizeof(double));
    k_expand_cplx(c, reinterpret_cast<const double *>(B), static_cast<double *>(pe->p),
                  static_cast<const CplxExpand *>(dp.expand->p), dp.nexpand, dp.expand_max_kn);
    B = static_cast<const T *>(pe->p);
  } else if (dp.hp->permute_a) {
    pa = c->alloc(std::max<int64_t>(a->n(), 1) * sizeof(T));
    k_permute<T>(c, A, static_cast<T *>(pa->p), static_cast<const PermDesc *>(dp.perm_a->p),
                 static_cast<const int64_t *>(dp.perm_a_start->p), dp.nperm_a, a->n());
    A = static_cast<const T *>(pa->p);
  }
  if (!dp.cplx && dp.hp->permute_b) {
    pb = c->alloc(std::max<int64_t>(b->n(), 1) * sizeof(T));
    k_permute<T>(c, B, static_cast<T *>(pb->p), static_cast<const PermDesc *>(dp.perm_b->p),
                 static_cast<const int64_t *>(dp.perm_b_start->p), dp.nperm_b, b->n());
    B = static_cast<const T *>(pb->p);
  }
  Ctx::GemmTiming tm{};
  if (c->time_gemms) {
    MPS2_HIP(hipEventCreate(&tm.start));
    MPS2_HIP(hipEventCreate(&tm.stop));
    MPS2_HIP(hipEventRecord(tm.start, c->stream));
  }
  // tag 1: the benchmarked stage; tag 2: one owner's share of it
  const int tag = (c->tag_stage >= 0 && c->stage_idx == c->tag_stage) ? (c->shard_g > 1 ? 2 : 1) : 0;
  c->stage_idx++;
  const bool fp64 = sizeof(T) == sizeof(double);
  if (dp.narrow_nc && tag != 2) {
    c->count(dp.hp->la == kTrans ? "gemm_narrow" : "gemm_narrow_rowmajor_a");
    launch_narrow_gemm<T>(c, dp.hp->la, dp.hp->lb, A, B, out, dp.map_narrow, dp.ntiles_narrow,
                          static_cast<const BlkD *>(dp.blks->p),
                          static_cast<const PairD *>(dp.pairs->p), dp.narrow_nc, 0, tag);
    if (c->time_gemms) {
      MPS2_HIP(hipEventRecord(tm.stop, c->stream));
      tm.flops = dp.hp->flops;
      tm.bytes = (double)sizeof(T) * ((double)a->nr() + (double)b->nr() + (double)dp.hp->out->total * (dp.cplx ? 2 : 1));
      c->gemm_timings.push_back(tm);
    }
    c->gemm_flops += dp.hp->flops;
    return;
  }
  PlanTune::Probe probe;
  const GemmChoice ch = tag == 2 ? GemmChoice{dp.tile_m, 0} : tune_select(c, dp, fp64, &probe);
  TileMap tiles = ch.tile == 128 ? dp.map128 : dp.map64;
  int ntiles = ch.tile == 128 ? dp.ntiles128 : dp.ntiles64;
  int tm_ = ch.tile, tn_ = ch.tile;
  if (tag == 2) {
    const DevPlan::Share &sh = plan_share(c, dp, c->shard_g, c->shard_r);
    tiles = TileMap{};
    tiles.list = sh.ntiles ? static_cast<const TileD *>(sh.tiles->p) : nullptr;
    ntiles = sh.ntiles;
    tm_ = dp.tile_m, tn_ = dp.tile_n;  // the share's list is built with the plan's tile size
    c->shard_flops = sh.flops;
  }
  tune_probe_begin(c, dp, fp64, &probe);
  launch_grouped_gemm<T>(c, dp.hp->la, dp.hp->lb, A, B, out, tiles, ntiles, static_cast<const BlkD *>(dp.blks->p),
                         static_cast<const PairD *>(dp.pairs->p), tm_, tn_, 0, tag, tag == 2 ? 0 : ch.variant,
                         dp.split);
  tune_probe_end(c, dp, fp64, &probe);
  if (c->time_gemms) {
    MPS2_HIP(hipEventRecord(tm.stop, c->stream));
    tm.flops = dp.hp->flops;
    tm.bytes = (double)sizeof(T) * ((double)a->nr() + (double)b->nr() + (double)dp.hp->out->total * (dp.cplx ? 2 : 1));
    c->gemm_timings.push_back(tm);
  }
  c->gemm_flops += dp.hp->flops;
}

}  // namespace

Arith arith_of(int precision) {
  if (precision < kArithF64 || precision > kArithBf16x3)
    throw Error(kErrArg, "precision must be MPS2_PREC_FP64, MPS2_PREC_MIXED or MPS2_PREC_MIXED_BF16X3");
  return static_cast<Arith>(precision);
}

void plan_cache_clear(Ctx *c) {
  std::vector<PlanEntry> dropped;   // released after the lock (their buffers go back to c)
  std::lock_guard<std::mutex> lock(g_plan_mu);
  for (auto it = g_plans.begin(); it != g_plans.end();) {
    if (std::get<0>(it->first) == c) {
      dropped.push_back(std::move(it->second));
      it = g_plans.erase(it);
    } else {
      ++it;
    }
  }
  std::deque<PlanKey> keep;
  for (auto &k : g_plan_order)
    if (std::get<0>(k) != c) keep.push_back(k);
  g_plan_order.swap(keep);
  std::vector<ContentEntry> cdropped;
  for (auto it = g_cplans.begin(); it != g_cplans.end();) {
    if (std::get<0>(it->first) == c) {
      cdropped.push_back(std::move(it->second));
      it = g_cplans.erase(it);
    } else {
      ++it;
    }
  }
  std::deque<ContentKey> ckeep;
  for (auto &k : g_cplan_order)
    if (std::get<0>(k) != c) ckeep.push_back(k);
  g_cplan_order.swap(ckeep);
}

std::shared_ptr<DevPlan> get_plan(Ctx *c, const MetaP &a, const MetaP &b, const std::vector<int> &ax,
                                  const std::vector<int> &bx) {
  return get_plan_t(c, a, b, ax, bx, nullptr);
}

TenP contract(const TenP &a0, const TenP &b0, const std::vector<int> &ax, const std::vector<int> &bx) {
  MPS2_REQUIRE(a0->ctx == b0->ctx, "Contract: tensors live on different contexts");
  MPS2_REQUIRE(a0->prec == b0->prec, "Contract: operand precisions differ");
  Ctx *c = a0->ctx;
  if (a0->cplx || b0->cplx) {
    // a real operand enters in complex form (gqten promotes mixed products,
    // e.g. the real S times a complex Vt of the MPS update)
    TenP a = ten_promote(a0), b = ten_promote(b0);
    auto dp = get_plan_t(c, a->meta, b->meta, ax, bx, nullptr, true);
    auto out = ten_alloc(c, dp->hp->out, kF64, true);
    if (dp->hp->out->total == 0) return out;
    run_plan<double>(c, *dp, a, b, out->d());
    return out;
  }
  const TenP &a = a0, &b = b0;
  auto dp = get_plan(c, a->meta, b->meta, ax, bx);
  auto out = ten_alloc(c, dp->hp->out, a->prec);
  if (dp->hp->out->total == 0) return out;
  if (a->prec == kF64) run_plan<double>(c, *dp, a, b, out->d());
  else run_plan<float>(c, *dp, a, b, out->f());
  return out;
}

void contract_into(const TenP &a0, const TenP &b0, const std::vector<int> &ax, const std::vector<int> &bx,
                   const TenP &out) {
  MPS2_REQUIRE(a0->prec == b0->prec && a0->prec == out->prec, "Contract: operand precisions differ");
  const bool cplx = a0->cplx || b0->cplx;
  MPS2_REQUIRE(out->cplx == cplx, "Contract: a complex product needs a complex output (and a real one a real output)");
  TenP a = cplx ? ten_promote(a0) : a0, b = cplx ? ten_promote(b0) : b0;
  Ctx *c = a->ctx;
  auto dp = get_plan_t(c, a->meta, b->meta, ax, bx, out->meta, cplx);
  if (!(dp->hp->out->coords == out->meta->coords)) {
    if (out->prec == kF64) k_fill(c, out->d(), std::max<int64_t>(out->nr(), 1), 0.0);
    else k_fillf(c, out->f(), std::max<int64_t>(out->n(), 1), 0.0f);
  }
  if (dp->hp->out->total == 0) return;
  if (a->prec == kF64) run_plan<double>(c, *dp, a, b, out->d());
  else run_plan<float>(c, *dp, a, b, out->f());
}

namespace {
MetaP union_meta(const Meta &a, const Meta &b) {
  MPS2_REQUIRE(a.legs == b.legs, "layouts with different indexes cannot be combined");
  auto m = std::make_shared<Meta>();
  m->qw = a.qw;
  m->legs = a.legs;
  m->coords = a.coords;
  m->coords.insert(m->coords.end(), b.coords.begin(), b.coords.end());
  m->finalize();
  return m;
}
}  // namespace

TenP embed(const TenP &t, const MetaP &m) {
  MPS2_REQUIRE(t->prec == kF64, "embed: FP64 tensors only");
  Ctx *c = t->ctx;
  if (t->meta->same_layout(*m)) {
    auto r = ten_copy(*t);
    r->meta = m;
    return r;
  }
  MPS2_REQUIRE(t->meta->legs == m->legs, "embed: indexes differ");
  auto r = ten_zeros(c, m, kF64, t->cplx);
  const int64_t w = t->cplx ? 2 : 1;   // doubles per element
  std::vector<Copy2D> d;
  for (size_t b = 0; b < t->meta->coords.size(); ++b) {
    int32_t tb = m->find(t->meta->coords[b]);
    MPS2_REQUIRE(tb >= 0, "embed: target layout lacks a block");
    int64_t sz = t->meta->bsize[b] * w;
    if (sz == 0) continue;
    d.push_back(Copy2D{t->meta->offset[b] * w, m->offset[tb] * w, 1, (int32_t)sz, (int32_t)sz, (int32_t)sz, 1.0});
  }
  if (!d.empty()) {
    auto dd = upload_bytes(c, d.data(), d.size() * sizeof(Copy2D));
    k_copy2d(c, t->d(), r->d(), static_cast<const Copy2D *>(dd->p), (int)d.size());
  }
  return r;
}

// Re <a|b> (for complex tensors the real part of sum conj(a) b, which is the
// dot product of their interleaved real forms).
double dot(const Ten &a, const Ten &b) {
  MPS2_REQUIRE(a.prec == kF64 && b.prec == kF64, "dot: FP64 tensors only");
  Ctx *c = a.ctx;
  if (a.cplx != b.cplx) {
    auto ta = ten_promote(std::make_shared<Ten>(a)), tb = ten_promote(std::make_shared<Ten>(b));
    return dot(*ta, *tb);
  }
  if (!a.meta->same_layout(*b.meta)) {
    auto m = union_meta(*a.meta, *b.meta);
    auto ta = std::make_shared<Ten>(a), tb = std::make_shared<Ten>(b);
    auto ea = embed(ta, m), eb = embed(tb, m);
    return dot(*ea, *eb);
  }
  if (a.n() == 0) return 0.0;
  k_dot(c, a.nr(), a.d(), b.d(), 0);
  MPS2_HIP(hipMemcpyAsync(c->host_scalars, c->dev_scalars->p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  return c->host_scalars[0];
}

double norm(const Ten &a) { return std::sqrt(dot(a, a)); }

// In-place writers call this first: a buffer shared with a Dag view or a
// copied handle is duplicated before it is modified.
void make_unique_buf(Ten &a) {
  if (a.buf.use_count() > 1) {
    auto r = ten_copy(a);
    a.buf = r->buf;
  }
}

double normalize(Ten &a) {
  double n = norm(a);
  if (n > 0) {
    make_unique_buf(a);
    k_scale(a.ctx, a.nr(), 1.0 / n, a.d());
  }
  return n;
}

void lincomb(const std::vector<double> &coefs, const std::vector<TenP> &ts, double beta, const TenP &out) {
  MPS2_REQUIRE(coefs.size() == ts.size(), "LinearCombine: coefficient count differs from tensor count");
  Ctx *c = out->ctx;
  for (auto &t : ts)
    MPS2_REQUIRE(t->cplx == out->cplx, "LinearCombine: tensors and output must all be real or all complex");
  // Grow `out` to the union of all block sets (gqten LinearCombine semantics).
  MetaP m = out->meta;
  for (auto &t : ts)
    if (!t->meta->same_layout(*m)) {
      bool sub = t->meta->legs == m->legs;
      if (sub)
        for (auto &cc : t->meta->coords)
          if (m->find(cc) < 0) {
            sub = false;
            break;
          }
      if (!sub) m = union_meta(*m, *t->meta);
    }
  if (m != out->meta) {
    auto grown = embed(out, m);
    out->meta = grown->meta;
    out->buf = grown->buf;
  } else {
    make_unique_buf(*out);
  }
  if (beta != 1.0) k_scale(c, out->nr(), beta, out->d());
  for (size_t i = 0; i < ts.size(); ++i) {
    TenP t = ts[i];
    if (!t->meta->same_layout(*out->meta)) t = embed(t, out->meta);
    k_axpby(c, out->nr(), coefs[i], t->d(), 1.0, out->d());
  }
}

void heff_mul(const TenP &L, const TenP &W1, const TenP &W2, const TenP &R, const TenP &psi,
              const std::string &where, const TenP &out, Prec prec) {
  (void)prec;
  psi->ctx->stage_idx = 0;
  if (where == "cent") {
    // Stages 2 and 3 of lanczos_solver_impl.h:228-236 (t1 W1 over {0,2}x{0,1},
    // then W2 over {4,1}x{0,1}) taken as one contraction with the two-site
    // MPO W12 = W1 W2 (W1 leg 3 with W2 leg 0; legs b_l, p1_in, p1_out,
    // p2_in, p2_out, b_r): t1 legs (m, p1, p2) with W12 legs (b_l, p1_in,
    // p2_in) give the same t3 legs (b, r, p1_out, p2_out, b_r) with one
    // memory-bound intermediate fewer.
    auto w12 = contract(W1, W2, {3}, {0});
    auto t1 = contract(L, psi, {0}, {0});
    auto t3 = contract(t1, w12, {0, 2, 3}, {0, 1, 3});
    t1.reset();
    // (t3 leg 1 (r) with R leg 0 (ket), t3 leg 4 (mpo) with R leg 1): the
    // reference's pairing {4, 1} x {1, 0}, listed in R's storage order so R,
    // which is the same for every Lanczos step, is read in place.
    contract_into(t3, R, {1, 4}, {0, 1}, out);
  } else if (where == "lend") {
    auto t1 = contract(psi, W1, {0}, {0});
    auto t2 = contract(t1, W2, {0, 2}, {1, 0});
    t1.reset();
    contract_into(t2, R, {0, 3}, {0, 1}, out);
  } else if (where == "rend") {
    auto t1 = contract(psi, L, {0}, {0});
    auto t2 = contract(t1, W1, {2, 0}, {0, 1});
    t1.reset();
    contract_into(t2, W2, {3, 0}, {1, 0}, out);
  } else {
    throw Error(kErrArg, "where must be lend, cent or rend");
  }
}

}  // namespace mps2
