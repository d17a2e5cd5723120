// extern "C" driver over the gqmps2-compatible host layer
// (declared in include/gqmps2_capi.h).
#include "gqmps2_capi.h"

#include <memory>
#include <string>
#include <vector>

#include "gqmps2/gqmps2.h"

namespace {

thread_local std::string g_err;

template <class F>
int guard(F f) {
  try {
    f();
    return MPS2_OK;
  } catch (const std::exception &e) {
    g_err = e.what();
    return MPS2_ERR_ARG;
  }
}

struct ModelBase {
  virtual ~ModelBase() = default;
  virtual int32_t add_op(int site, const double *dense) = 0;
  virtual void add_term(double coef, int nops, const int32_t *ops, const int64_t *sites, int ninst,
                        const int32_t *inst) = 0;
  virtual std::vector<int64_t> gen_mpo() = 0;
  virtual void export_mpo(int site, std::vector<int32_t> &meta, std::vector<double> &data) = 0;
  virtual void mps_init(int nsets, const int64_t *labs, int64_t enlarged, uint64_t seed, int centralize) = 0;
  virtual void export_mps(int site, std::vector<int32_t> &meta, std::vector<double> &data) = 0;
  virtual void import_mps(int site, const int32_t *meta, const double *data) = 0;
  virtual void checkpoint() = 0;
  virtual void restore() = 0;
  int64_t bench_site = -1;
  int bench_warmup = 2, bench_reps = 10;
  bool bench_done = false;
  double bench_out[24] = {0};
  std::vector<double> phase_times;   // (lanczos, svd, env) seconds per update of the last dmrg call
  std::vector<mps2_ten> bench_tens = std::vector<mps2_ten>(5, nullptr);
  void clear_bench_tens() {
    for (auto &h : bench_tens) {
      if (h) mps2_ten_free(h);
      h = nullptr;
    }
  }
  virtual double dmrg(int64_t sweeps, int64_t dmin, int64_t dmax, const int64_t *sched, double trunc_err, double lerr,
                      int64_t lmax, int prec, int verbose, double *en, int64_t *dims, int64_t *its, double *te,
                      double *ut, double *st) = 0;
  virtual void bench_bond(int64_t site, int warmup, int reps) = 0;
  virtual void truncate_mps(double trunc_err, int64_t dmin, int64_t dmax) = 0;
  virtual void measure_one_site(int nops, const int32_t *op_ids, const char *const *basenames, double *avgs) = 0;
  virtual void measure_multi_site(int64_t nevents, const int32_t *nphys, const int64_t *sites, const int32_t *phys_ids,
                                  const int32_t *inst_ids, const char *basename, double *avgs) = 0;
};

template <typename QNT>
struct Model : ModelBase {
  using TenT = gqten::GQTensor<double, QNT>;
  size_t N;
  std::unique_ptr<gqmps2::SiteVec<double, QNT>> sites;
  std::unique_ptr<gqmps2::MPOGenerator<double, QNT>> gen;
  std::vector<TenT> ops;
  std::vector<int> op_site;
  gqmps2::MPO<TenT> mpo;
  std::unique_ptr<gqmps2::FiniteMPS<double, QNT>> mps;
  std::vector<TenT> saved;

  // Read-only view of the MPS: FiniteMPS's mutable accessors mark a tensor
  // as no longer canonical (as in the reference), so checks and reads go
  // through this view and a centred MPS stays centred between calls.
  const gqmps2::FiniteMPS<double, QNT> &cm() const { return *mps; }

  Model(int nsites, const int32_t *nsect, const int32_t *qns, const int32_t *dims) : N(nsites) {
    gqten::IndexVec<QNT> idx;
    int p = 0;
    for (int i = 0; i < nsites; ++i) {
      if (nsect[i] < 1) throw std::invalid_argument("every site needs at least one sector");
      gqten::QNSectorVec<QNT> secs;
      for (int s = 0; s < nsect[i]; ++s) {
        QNT q;
        for (int k = 0; k < QNT::kWidth; ++k) q[k] = qns[p * QNT::kWidth + k];
        if (dims[p] < 1) throw std::invalid_argument("sector dimensions must be positive");
        secs.emplace_back(q, (size_t)dims[p]);
        ++p;
      }
      idx.emplace_back(secs, gqten::GQTenIndexDirType::OUT);
    }
    sites = std::make_unique<gqmps2::SiteVec<double, QNT>>(idx);
    gen = std::make_unique<gqmps2::MPOGenerator<double, QNT>>(*sites, QNT());
    mps = std::make_unique<gqmps2::FiniteMPS<double, QNT>>(*sites);
  }

  int32_t add_op(int site, const double *dense) override {
    if (site < 0 || (size_t)site >= N) throw std::out_of_range("site out of range");
    auto pout = sites->sites[site];
    auto pin = gqten::InverseIndex(pout);
    TenT op({pin, pout});
    size_t d = pout.dim();
    for (size_t i = 0; i < d; ++i)
      for (size_t j = 0; j < d; ++j)
        if (dense[i * d + j] != 0.0) op({i, j}) = dense[i * d + j];
    ops.push_back(op);
    op_site.push_back(site);
    return (int32_t)ops.size() - 1;
  }

  void add_term(double coef, int nops, const int32_t *oid, const int64_t *st, int ninst, const int32_t *inst) override {
    // The generator asserts these; a C caller gets an error status instead.
    if (nops < 1) throw std::invalid_argument("a term needs at least one operator");
    if (ninst != 0 && (nops < 2 || (ninst != nops - 1 && ninst != nops)))
      throw std::invalid_argument("insertion operators: one per gap, or one per gap plus one for the tail");
    if (ninst != 0 && !inst) throw std::invalid_argument("NULL insertion operator list");
    auto space_of = [&](int32_t id) -> const gqten::Index<QNT> & { return sites->sites[op_site.at(id)]; };
    for (int k = 0; k < nops; ++k) {
      if (st[k] < 0 || (size_t)st[k] >= N) throw std::out_of_range("term site out of range");
      if (k > 0 && st[k] < st[k - 1]) throw std::invalid_argument("term sites must be in ascending order");
      if (space_of(oid[k]) != sites->sites[st[k]])
        throw std::invalid_argument("operator built for another local space than its site's");
    }
    for (int k = 0; k < ninst; ++k) {
      const int64_t from = st[k] + 1, to = k + 1 < nops ? st[k + 1] : (int64_t)N;
      for (int64_t j = from; j < to; ++j)
        if (space_of(inst[k]) != sites->sites[j])
          throw std::invalid_argument("insertion operator built for another local space than its site's");
    }
    std::vector<TenT> o;
    std::vector<size_t> s;
    for (int k = 0; k < nops; ++k) {
      o.push_back(ops.at(oid[k]));
      s.push_back((size_t)st[k]);
    }
    if (ninst == 0) {
      gen->AddTerm(coef, o, s);
    } else {
      std::vector<TenT> io;
      for (int k = 0; k < ninst; ++k) io.push_back(ops.at(inst[k]));
      gen->AddTerm(coef, o, s, io);
    }
  }

  std::vector<int64_t> gen_mpo() override {
    mpo = gen->Gen();
    std::vector<int64_t> bd;
    for (size_t i = 0; i + 1 < N; ++i) {
      size_t leg = i == 0 ? 1 : 3;
      bd.push_back((int64_t)mpo[i].GetIndexes()[leg].dim());
    }
    return bd;
  }

  void export_mpo(int site, std::vector<int32_t> &meta, std::vector<double> &data) override {
    if (mpo.size() == 0) throw std::runtime_error("no MPO yet (call gqmps2_model_gen_mpo)");
    if (site < 0 || (size_t)site >= mpo.size()) throw std::out_of_range("site out of range");
    mpo[site].ToFlat(meta, data);
  }

  void mps_init(int nsets, const int64_t *labs, int64_t enlarged, uint64_t seed, int centralize) override {
    gqten::SetRandomSeed(seed);
    std::vector<std::vector<size_t>> sets(nsets, std::vector<size_t>(N));
    for (int k = 0; k < nsets; ++k)
      for (size_t i = 0; i < N; ++i) sets[k][i] = (size_t)labs[k * N + i];
    mps = std::make_unique<gqmps2::FiniteMPS<double, QNT>>(*sites);
    if (centralize) {
      if (enlarged == 0) gqmps2::DirectStateInitMps(*mps, sets[0], QNT());
      else gqmps2::ExtendDirectRandomInitMps(*mps, sets, QNT(), (size_t)enlarged);
    } else {
      // same tensors, no Centralize(0): build through a site-local copy
      InitRaw(sets, enlarged);
    }
  }

  void InitRaw(const std::vector<std::vector<size_t>> &sets, int64_t enlarged) {
    using IndexT = gqten::Index<QNT>;
    using QNSctT = gqten::QNSector<QNT>;
    auto pb = sites->sites;
    size_t nf = sets.size();
    size_t ed = enlarged == 0 ? 1 : (size_t)enlarged;
    auto div = pb[0].GetQNSctFromActualCoor(sets[0][0]).GetQn();
    for (size_t i = 1; i < N; ++i) div += pb[i].GetQNSctFromActualCoor(sets[0][i]).GetQn();
    std::vector<QNSctT> secs;
    for (size_t j = 0; j < nf; ++j) secs.push_back(QNSctT(div - pb[0].GetQNSctFromActualCoor(sets[j][0]).GetQn(), ed));
    IndexT rvb(secs, gqten::GQTenIndexDirType::OUT), lvb;
    secs.clear();
    (*mps)(0) = new TenT({pb[0], rvb});
    if (enlarged == 0) (*mps)[0]({sets[0][0], 0}) = 1;
    else (*mps)[0].Random(div);
    for (size_t i = 1; i + 1 < N; ++i) {
      lvb = gqten::InverseIndex(rvb);
      for (size_t j = 0; j < nf; ++j)
        secs.push_back(QNSctT(QNT() - pb[i].GetQNSctFromActualCoor(sets[j][i]).GetQn() +
                                  lvb.GetQNSctFromActualCoor(j * ed).GetQn(),
                              ed));
      rvb = IndexT(secs, gqten::GQTenIndexDirType::OUT);
      secs.clear();
      (*mps)(i) = new TenT({lvb, pb[i], rvb});
      if (enlarged == 0) (*mps)[i]({0, sets[0][i], 0}) = 1;
      else (*mps)[i].Random(QNT());
    }
    lvb = gqten::InverseIndex(rvb);
    (*mps)(N - 1) = new TenT({lvb, pb[N - 1]});
    if (enlarged == 0) (*mps)[N - 1]({0, sets[0][N - 1]}) = 1;
    else (*mps)[N - 1].Random(QNT());
  }

  void export_mps(int site, std::vector<int32_t> &meta, std::vector<double> &data) override {
    if (site < 0 || (size_t)site >= N) throw std::out_of_range("site out of range");
    if (!cm()(site)) throw std::runtime_error("MPS tensor not resident");
    cm()[site].ToFlat(meta, data);
  }

  void import_mps(int site, const int32_t *meta, const double *data) override {
    if (site < 0 || (size_t)site >= N) throw std::out_of_range("site out of range");
    // size the meta: rank, qw, nblk then legs then coords
    int rank = meta[0], qw = meta[1], nb = meta[2];
    if (rank < 0 || rank > 64 || qw != QNT::kWidth || nb < 0) throw std::invalid_argument("malformed tensor meta");
    int p = 3;
    for (int l = 0; l < rank; ++l) {
      int ns = meta[p + 1];
      if (ns < 0) throw std::invalid_argument("malformed tensor meta");
      p += 2 + ns * (qw + 1);
    }
    p += nb * rank;
    std::vector<int32_t> mm(meta, meta + p);
    int64_t total = 0;
    {
      // count elements from the block list
      int q = 3;
      std::vector<std::vector<int64_t>> dims(rank);
      for (int l = 0; l < rank; ++l) {
        int ns = meta[q + 1];
        q += 2;
        for (int s = 0; s < ns; ++s) {
          q += qw;
          dims[l].push_back(meta[q++]);
        }
      }
      for (int b = 0; b < nb; ++b) {
        int64_t sz = 1;
        for (int l = 0; l < rank; ++l) {
          const int32_t c = meta[q++];
          if (c < 0 || c >= (int32_t)dims[l].size()) throw std::invalid_argument("tensor meta: block coordinate out of range");
          if (dims[l][c] < 0) throw std::invalid_argument("tensor meta: negative sector dimension");
          sz *= dims[l][c];
        }
        total += sz;
      }
    }
    if (total > 0 && !data) throw std::invalid_argument("NULL data for a tensor with elements");
    std::vector<double> d(data, data + total);
    TenT t;
    t.FromFlat(mm, d);
    if ((*mps)(site)) mps->dealloc(site);
    (*mps)(site) = new TenT(t);
  }

  void checkpoint() override {
    require_mps();
    saved.clear();
    for (size_t i = 0; i < N; ++i) {
      TenT c;
      mps2_ten h = nullptr;
      gqten::CheckMps2(mps2_ten_copy(cm()[i].Handle(), &h), "checkpoint");
      c.AdoptDevice(h);
      saved.push_back(std::move(c));
    }
  }

  void restore() override {
    if (saved.size() != N) throw std::runtime_error("no checkpoint");
    for (size_t i = 0; i < N; ++i) {
      mps2_ten h = nullptr;
      gqten::CheckMps2(mps2_ten_copy(saved[i].Handle(), &h), "restore");
      TenT c;
      c.AdoptDevice(h);
      (*mps)[i] = std::move(c);
    }
  }

  void bench_bond(int64_t site, int warmup, int reps) override {
    if (site < 1 || (size_t)site + 2 >= N) throw std::invalid_argument("bench_bond: site must be an interior bond");
    require_mps();
    require_mpo();
    mps->Centralize((int)site);
    gqmps2::SweepParams sp(1, 1, 1, 0.0, gqmps2::LanczosParams());
    gqmps2::TwoSiteVMPSEngine<double, QNT> eng(*mps, mpo, sp);
    auto &L = eng.lenvs();
    auto &R = eng.renvs();
    // left environments of sites 0 .. site-1 (two_site_update_finite_vmps_impl.h:288-299)
    for (int64_t s = 0; s < site; ++s) {
      if (s == 0) {
        TenT temp;
        Contract(&cm()[0], &mpo[0], {{0}, {0}}, &temp);
        auto ad = gqten::Dag(cm()[0]);
        Contract(&temp, &ad, {{2}, {0}}, &L[1]);
      } else {
        TenT t1, t2;
        Contract(&L[s], &cm()[s], {{0}, {0}}, &t1);
        Contract(&t1, &mpo[s], {{0, 2}, {0, 1}}, &t2);
        auto ad = gqten::Dag(cm()[s]);
        Contract(&t2, &ad, {{0, 2}, {0, 1}}, &L[s + 1]);
      }
    }
    // right environments up to length N - site - 2
    size_t rlen = N - (size_t)site - 2;
    eng.InitEnvs(rlen);
    std::vector<TenT *> H{&L[site], &mpo[site], &mpo[site + 1], &R[rlen]};
    TenT init;
    Contract(&cm()[site], &cm()[site + 1], {{2}, {0}}, &init);
    for (int p = 0; p < 2; ++p)
      gqten::CheckMps2(mps2_heff_bench(H[0]->Handle(), H[1]->Handle(), H[2]->Handle(), H[3]->Handle(), init.Handle(),
                                       "cent", p == 0 ? MPS2_PREC_FP64 : MPS2_PREC_MIXED, warmup, reps,
                                       bench_out + 12 * p),
                       "heff_bench");
    clear_bench_tens();
    mps2_ten src[5] = {H[0]->Handle(), H[1]->Handle(), H[2]->Handle(), H[3]->Handle(), init.Handle()};
    for (int k = 0; k < 5; ++k) gqten::CheckMps2(mps2_ten_copy(src[k], &bench_tens[k]), "bench copy");
    bench_done = true;
  }

  void require_mps() const {
    for (size_t i = 0; i < N; ++i)
      if (!cm()(i)) throw std::runtime_error("MPS not initialised");
  }

  // The sweep indexes one MPO tensor per site (the reference asserts
  // mps.size() == mpo.size(), two_site_update_finite_vmps_impl.h:83).
  void require_mpo() const {
    if (mpo.size() != N) throw std::runtime_error("no MPO yet (call gqmps2_model_gen_mpo)");
  }

  void truncate_mps(double trunc_err, int64_t dmin, int64_t dmax) override {
    require_mps();
    if (dmin < 1 || dmax < dmin) throw std::invalid_argument("truncate_mps: need 1 <= dmin <= dmax");
    gqmps2::TruncateMPS(*mps, trunc_err, (size_t)dmin, (size_t)dmax);
  }

  void measure_one_site(int nops, const int32_t *op_ids, const char *const *basenames, double *avgs) override {
    require_mps();
    if (nops < 1) throw std::invalid_argument("measure_one_site: no operator");
    std::vector<TenT> o;
    std::vector<std::string> names;
    for (int j = 0; j < nops; ++j) {
      o.push_back(ops.at(op_ids[j]));
      if (!basenames[j]) throw std::invalid_argument("measure_one_site: NULL file basename");
      names.emplace_back(basenames[j]);
    }
    auto set = gqmps2::MeasureOneSiteOp(*mps, o, names);
    for (int j = 0; j < nops; ++j)
      for (size_t i = 0; i < N; ++i) avgs[j * N + i] = set[j][i].avg;
  }

  void measure_multi_site(int64_t nevents, const int32_t *nphys, const int64_t *st, const int32_t *phys_ids,
                          const int32_t *inst_ids, const char *basename, double *avgs) override {
    require_mps();
    if (!basename) throw std::invalid_argument("measure_multi_site: NULL file basename");
    const auto &id_ops = mps->GetSitesInfo().id_ops;
    // one device copy per operator, shared by the per-event copies below
    auto op_at = [&](int32_t id) -> const TenT & {
      const TenT &t = ops.at(id);
      t.ToDevice();
      return t;
    };
    gqmps2::TenVV<double, QNT> phys_set(nevents);
    gqmps2::TenVVV<double, QNT> inst_set(nevents);
    std::vector<std::vector<size_t>> sites_set(nevents);
    size_t p = 0, g = 0;
    for (int64_t e = 0; e < nevents; ++e) {
      if (nphys[e] < 2) throw std::invalid_argument("measure_multi_site: an event needs two operators or more");
      for (int k = 0; k < nphys[e]; ++k, ++p) {
        if (st[p] < 0 || (size_t)st[p] >= N) throw std::out_of_range("measure_multi_site: site out of range");
        phys_set[e].push_back(op_at(phys_ids[p]));
        sites_set[e].push_back((size_t)st[p]);
      }
      if (!gqmps2::IsOrderKept(sites_set[e])) throw std::invalid_argument("measure_multi_site: sites not ascending");
      for (int k = 0; k + 1 < nphys[e]; ++k, ++g) {
        std::vector<TenT> gap;
        for (size_t s = sites_set[e][k] + 1; s < sites_set[e][k + 1]; ++s)
          gap.push_back(inst_ids[g] < 0 ? id_ops[s] : op_at(inst_ids[g]));
        inst_set[e].push_back(std::move(gap));
      }
    }
    auto res = gqmps2::MeasureMultiSiteOp(*mps, phys_set, inst_set, sites_set, basename);
    for (int64_t e = 0; e < nevents; ++e) avgs[e] = res[e].avg;
  }

  double dmrg(int64_t sweeps, int64_t dmin, int64_t dmax, const int64_t *sched, double trunc_err, double lerr,
              int64_t lmax, int prec, int verbose, double *en, int64_t *dims, int64_t *its, double *te, double *ut,
              double *st) override {
    require_mps();
    require_mpo();
    if (sweeps < 0 || dmin < 1 || dmax < dmin) throw std::invalid_argument("dmrg: need sweeps >= 0 and 1 <= dmin <= dmax");
    if (prec != MPS2_PREC_FP64 && prec != MPS2_PREC_MIXED && prec != MPS2_PREC_MIXED_BF16X3)
      throw std::invalid_argument("dmrg: precision must be MPS2_PREC_FP64, MPS2_PREC_MIXED or MPS2_PREC_MIXED_BF16X3");
    gqmps2::LanczosParams lp(lerr, (size_t)lmax, static_cast<gqmps2::LanczosPrecision>(prec));
    double e0 = 0;
    int64_t k = 0;
    phase_times.clear();
    auto record = [&](const gqmps2::UpdateRecord &r) {
      if (en) en[k] = r.energy;
      if (dims) dims[k] = (int64_t)r.D;
      if (its) its[k] = (int64_t)r.iters;
      if (te) te[k] = r.trunc_err;
      if (ut) ut[k] = r.update_time;
      phase_times.push_back(r.lancz_time);
      phase_times.push_back(r.svd_time);
      phase_times.push_back(r.env_time);
      ++k;
    };
    std::unique_ptr<gqmps2::TwoSiteVMPSEngine<double, QNT>> eng;
    for (int64_t s = 0; s < sweeps; ++s) {
      int64_t dm = sched ? sched[s] : dmax;
      gqmps2::SweepParams sp((size_t)1, (size_t)dmin, (size_t)dm, trunc_err, lp);
      auto *old = eng.release();
      eng = std::make_unique<gqmps2::TwoSiteVMPSEngine<double, QNT>>(*mps, mpo, sp);
      if (old) {
        eng->renvs() = std::move(old->renvs());
        eng->lenvs() = std::move(old->lenvs());
        delete old;
      } else {
        eng->InitEnvs();
      }
      eng->quiet = !verbose;
      eng->on_update = record;
      if (bench_site >= 0 && !bench_done) {
        eng->before_lanczos = [this](const std::vector<TenT *> &H, const TenT &init, const std::string &where,
                                     char dir, size_t site) {
          if (bench_done || dir != 'r' || (int64_t)site != bench_site) return;
          auto h = [](TenT *t) -> mps2_ten { return t ? t->Handle() : nullptr; };
          for (int p = 0; p < 2; ++p)
            gqten::CheckMps2(mps2_heff_bench(h(H[0]), H[1]->Handle(), H[2]->Handle(), h(H[3]), init.Handle(),
                                             where.c_str(), p == 0 ? MPS2_PREC_FP64 : MPS2_PREC_MIXED, bench_warmup,
                                             bench_reps, bench_out + 12 * p),
                             "heff_bench");
          clear_bench_tens();
          mps2_ten src[5] = {h(H[0]), H[1]->Handle(), H[2]->Handle(), h(H[3]), init.Handle()};
          for (int k = 0; k < 5; ++k)
            if (src[k]) gqten::CheckMps2(mps2_ten_copy(src[k], &bench_tens[k]), "bench copy");
          bench_done = true;
        };
      }
      gqten::Timer t("sweep");
      e0 = eng->Sweep((size_t)s + 1);
      gqten::CheckMps2(mps2_ctx_sync(gqten::DefaultCtx()), "sync");
      if (st) st[s] = t.Elapsed();
      if (verbose) t.PrintElapsed();
    }
    return e0;
  }
};

}  // namespace

struct gqmps2_model_s {
  std::unique_ptr<ModelBase> m;
};

extern "C" {

int gqmps2_model_create(int qn_width, int nsites, const int32_t *nsect, const int32_t *qns, const int32_t *dims,
                        gqmps2_model *out) {
  return guard([&] {
    if (!out || !nsect || !qns || !dims) throw std::invalid_argument("NULL argument");
    if (nsites < 2) throw std::invalid_argument("at least two sites are required");
    auto h = std::make_unique<gqmps2_model_s>();
    if (qn_width == 1) h->m = std::make_unique<Model<gqten::QN<gqten::U1QNVal>>>(nsites, nsect, qns, dims);
    else if (qn_width == 2)
      h->m = std::make_unique<Model<gqten::QN<gqten::U1QNVal, gqten::U1QNVal>>>(nsites, nsect, qns, dims);
    else
      throw std::invalid_argument("qn_width must be 1 or 2");
    *out = h.release();
  });
}

int gqmps2_model_destroy(gqmps2_model m) {
  return guard([&] {
    if (m) m->m->clear_bench_tens();
    delete m;
  });
}

int gqmps2_model_add_op(gqmps2_model m, int site, const double *dense, int32_t *op_id) {
  return guard([&] {
    if (!m || !dense || !op_id) throw std::invalid_argument("NULL argument");
    *op_id = m->m->add_op(site, dense);
  });
}

int gqmps2_model_add_term(gqmps2_model m, double coef, int nops, const int32_t *op_ids, const int64_t *sites,
                          int ninst, const int32_t *inst_op_ids) {
  return guard([&] {
    if (!m || !op_ids || !sites) throw std::invalid_argument("NULL argument");
    m->m->add_term(coef, nops, op_ids, sites, ninst, inst_op_ids);
  });
}

int gqmps2_model_gen_mpo(gqmps2_model m, int64_t *bond_dims) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    auto bd = m->m->gen_mpo();
    if (bond_dims)
      for (size_t i = 0; i < bd.size(); ++i) bond_dims[i] = bd[i];
  });
}

int gqmps2_model_mpo_export_sizes(gqmps2_model m, int site, int64_t *nmeta, int64_t *ndata) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    std::vector<int32_t> meta;
    std::vector<double> data;
    m->m->export_mpo(site, meta, data);
    if (nmeta) *nmeta = (int64_t)meta.size();
    if (ndata) *ndata = (int64_t)data.size();
  });
}

int gqmps2_model_mpo_export(gqmps2_model m, int site, int32_t *meta, double *data) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    std::vector<int32_t> mt;
    std::vector<double> dt;
    m->m->export_mpo(site, mt, dt);
    if (meta) std::copy(mt.begin(), mt.end(), meta);
    if (data) std::copy(dt.begin(), dt.end(), data);
  });
}

int gqmps2_mps_init(gqmps2_model m, int nsets, const int64_t *stat_labs, int64_t enlarged_dim, uint64_t seed,
                    int centralize) {
  return guard([&] {
    if (!m || !stat_labs || nsets < 1 || enlarged_dim < 0) throw std::invalid_argument("bad arguments");
    m->m->mps_init(nsets, stat_labs, enlarged_dim, seed, centralize);
  });
}

int gqmps2_mps_export_sizes(gqmps2_model m, int site, int64_t *nmeta, int64_t *ndata) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    std::vector<int32_t> meta;
    std::vector<double> data;
    m->m->export_mps(site, meta, data);
    if (nmeta) *nmeta = (int64_t)meta.size();
    if (ndata) *ndata = (int64_t)data.size();
  });
}

int gqmps2_mps_export(gqmps2_model m, int site, int32_t *meta, double *data) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    std::vector<int32_t> mt;
    std::vector<double> dt;
    m->m->export_mps(site, mt, dt);
    if (meta) std::copy(mt.begin(), mt.end(), meta);
    if (data) std::copy(dt.begin(), dt.end(), data);
  });
}

int gqmps2_mps_import(gqmps2_model m, int site, const int32_t *meta, const double *data) {
  return guard([&] {
    if (!m || !meta) throw std::invalid_argument("NULL argument");
    m->m->import_mps(site, meta, data);
  });
}

int gqmps2_mps_checkpoint(gqmps2_model m) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    m->m->checkpoint();
  });
}

int gqmps2_mps_restore(gqmps2_model m) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    m->m->restore();
  });
}

int gqmps2_dmrg(gqmps2_model m, int64_t sweeps, int64_t dmin, int64_t dmax, const int64_t *dmax_schedule,
                double trunc_err, double lanczos_error, int64_t lanczos_max_iter, int precision, int verbose,
                double *energies, int64_t *dims, int64_t *iters, double *trunc_errs, double *update_times,
                double *sweep_times, double *final_energy) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    double e = m->m->dmrg(sweeps, dmin, dmax, dmax_schedule, trunc_err, lanczos_error, lanczos_max_iter, precision,
                          verbose, energies, dims, iters, trunc_errs, update_times, sweep_times);
    if (final_energy) *final_energy = e;
  });
}

int gqmps2_set_bench(gqmps2_model m, int64_t site, int warmup, int reps) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    m->m->bench_site = site;
    m->m->bench_warmup = warmup;
    m->m->bench_reps = reps;
    m->m->bench_done = false;
  });
}

int gqmps2_phase_times(gqmps2_model m, double *out, int64_t n) {
  return guard([&] {
    if (!m || !out) throw std::invalid_argument("NULL argument");
    auto &v = m->m->phase_times;
    for (int64_t i = 0; i < n && i < (int64_t)v.size(); ++i) out[i] = v[i];
  });
}

int gqmps2_set_verbose(int verbose) {
  return guard([&] { gqmps2::Verbose() = verbose != 0; });
}

int gqmps2_bench_bond(gqmps2_model m, int64_t site, int warmup, int reps) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    m->m->bench_bond(site, warmup, reps);
  });
}

int gqmps2_get_bench(gqmps2_model m, double *out) {
  return guard([&] {
    if (!m || !out) throw std::invalid_argument("NULL argument");
    if (!m->m->bench_done) throw std::runtime_error("no benchmark recorded");
    for (int i = 0; i < 24; ++i) out[i] = m->m->bench_out[i];
  });
}

int gqmps2_bench_tensors(gqmps2_model m, mps2_ten *out5) {
  return guard([&] {
    if (!m || !out5) throw std::invalid_argument("NULL argument");
    for (int k = 0; k < 5; ++k) {
      out5[k] = nullptr;
      if (m->m->bench_tens[k]) gqten::CheckMps2(mps2_ten_copy(m->m->bench_tens[k], &out5[k]), "bench copy");
    }
  });
}

int gqmps2_mps_truncate(gqmps2_model m, double trunc_err, int64_t dmin, int64_t dmax) {
  return guard([&] {
    if (!m) throw std::invalid_argument("NULL model");
    m->m->truncate_mps(trunc_err, dmin, dmax);
  });
}

int gqmps2_measure_one_site(gqmps2_model m, int nops, const int32_t *op_ids, const char *const *basenames,
                            double *avgs) {
  return guard([&] {
    if (!m || !op_ids || !basenames || !avgs) throw std::invalid_argument("NULL argument");
    m->m->measure_one_site(nops, op_ids, basenames, avgs);
  });
}

int gqmps2_measure_multi_site(gqmps2_model m, int64_t nevents, const int32_t *nphys, const int64_t *sites,
                              const int32_t *phys_op_ids, const int32_t *inst_op_ids, const char *basename,
                              double *avgs) {
  return guard([&] {
    if (!m || !nphys || !sites || !phys_op_ids || !inst_op_ids || !avgs) throw std::invalid_argument("NULL argument");
    m->m->measure_multi_site(nevents, nphys, sites, phys_op_ids, inst_op_ids, basename, avgs);
  });
}

mps2_ctx gqmps2_engine_ctx(void) {
  try {
    return gqten::DefaultCtx();
  } catch (...) {
    return nullptr;
  }
}

const char *gqmps2_last_error(void) { return g_err.c_str(); }

}  // extern "C"
