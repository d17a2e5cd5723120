// Host-side quantum-number bookkeeping (see meta.h).
#include "meta.h"

#include <algorithm>
#include <sstream>
#include <stdexcept>

namespace mps2 {

QN Meta::block_div(const Coord &c) const {
  QN d;
  for (int i = 0; i < rank(); ++i) {
    const QN &q = legs[i].qn[c[i]];
    d = legs[i].dir == kOut ? d + q : d - q;
  }
  return d;
}

QN Meta::div() const { return coords.empty() ? QN() : block_div(coords[0]); }

void Meta::finalize() {
  std::sort(coords.begin(), coords.end());
  coords.erase(std::unique(coords.begin(), coords.end()), coords.end());
  offset.resize(coords.size());
  bsize.resize(coords.size());
  lookup.clear();
  lookup.reserve(coords.size() * 2);
  int64_t off = 0;
  for (size_t b = 0; b < coords.size(); ++b) {
    int64_t s = 1;
    for (int l = 0; l < rank(); ++l) s *= legs[l].dim[coords[b][l]];
    offset[b] = off;
    bsize[b] = s;
    off += s;
    lookup[coords[b]] = (int32_t)b;
  }
  total = off;
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t x) { h = (h ^ x) * 1099511628211ull; };
  mix((uint64_t)qw);
  for (const Leg &l : legs) {
    mix((uint64_t)(uint32_t)l.dir);
    for (const QN &q : l.qn)
      for (int i = 0; i < kMaxQnWidth; ++i) mix((uint32_t)q.v[i]);
    for (int64_t d : l.dim) mix((uint64_t)d);
    mix(0x9e3779b97f4a7c15ull);
  }
  for (const Coord &c : coords)
    for (int32_t x : c) mix((uint32_t)x);
  digest = h;
}

bool Meta::same_layout(const Meta &o) const {
  return qw == o.qw && legs == o.legs && coords == o.coords;
}

MetaP full_meta(int qw, const std::vector<Leg> &legs, const QN &div) {
  auto m = std::make_shared<Meta>();
  m->qw = qw;
  m->legs = legs;
  int r = (int)legs.size();
  if (r == 0) {
    m->coords.push_back(Coord());
    m->finalize();
    return m;
  }
  // The last leg's sector is fixed by the divergence of the others.
  std::map<QN, std::vector<int32_t>> last;
  for (int32_t s = 0; s < (int32_t)legs[r - 1].qn.size(); ++s) last[legs[r - 1].qn[s]].push_back(s);
  Coord c(r, 0);
  std::vector<int32_t> ns(r);
  for (int i = 0; i < r; ++i) ns[i] = (int32_t)legs[i].qn.size();
  for (int i = 0; i < r - 1; ++i)
    if (ns[i] == 0) {
      m->finalize();
      return m;
    }
  while (true) {
    QN partial;
    for (int i = 0; i < r - 1; ++i) {
      const QN &q = legs[i].qn[c[i]];
      partial = legs[i].dir == kOut ? partial + q : partial - q;
    }
    // div = partial + sign_last * q_last  ->  q_last = sign_last * (div - partial)
    QN need = div - partial;
    if (legs[r - 1].dir == kIn) need = QN() - need;
    auto it = last.find(need);
    if (it != last.end())
      for (int32_t s : it->second) {
        c[r - 1] = s;
        m->coords.push_back(c);
      }
    int i = r - 2;
    for (; i >= 0; --i) {
      if (++c[i] < ns[i]) break;
      c[i] = 0;
    }
    if (i < 0) break;
  }
  m->finalize();
  return m;
}

MetaP meta_from_flat(const int32_t *m, int64_t *consumed) {
  auto t = std::make_shared<Meta>();
  int64_t p = 0;
  int rank = m[p++];
  t->qw = m[p++];
  int nblk = m[p++];
  if (rank < 0 || rank > kMaxRank) throw std::invalid_argument("flat tensor: rank out of range");
  if (t->qw < 1 || t->qw > kMaxQnWidth) throw std::invalid_argument("flat tensor: qn width out of range");
  if (nblk < 0) throw std::invalid_argument("flat tensor: negative block count");
  t->legs.resize(rank);
  for (int l = 0; l < rank; ++l) {
    t->legs[l].dir = m[p++];
    int ns = m[p++];
    if (ns < 0) throw std::invalid_argument("flat tensor: negative sector count");
    for (int s = 0; s < ns; ++s) {
      QN q;
      for (int k = 0; k < t->qw; ++k) q.v[k] = m[p++];
      int32_t d = m[p++];
      if (d < 0) throw std::invalid_argument("flat tensor: negative sector dim");
      t->legs[l].qn.push_back(q);
      t->legs[l].dim.push_back(d);
    }
  }
  for (int b = 0; b < nblk; ++b) {
    Coord c(m + p, m + p + rank);
    p += rank;
    for (int l = 0; l < rank; ++l)
      if (c[l] < 0 || c[l] >= (int32_t)t->legs[l].qn.size())
        throw std::invalid_argument("flat tensor: block coordinate out of range");
    t->coords.push_back(c);
  }
  std::vector<Coord> given = t->coords;
  t->finalize();
  if (given != t->coords) throw std::invalid_argument("flat tensor: blocks must be unique and sorted");
  if (consumed) *consumed = p;
  return t;
}

int64_t meta_flat_size(const Meta &m) {
  int64_t n = 3;
  for (auto &l : m.legs) n += 2 + (int64_t)l.qn.size() * (m.qw + 1);
  n += (int64_t)m.coords.size() * m.rank();
  return n;
}

void meta_to_flat(const Meta &m, int32_t *o) {
  int64_t p = 0;
  o[p++] = m.rank();
  o[p++] = m.qw;
  o[p++] = (int32_t)m.coords.size();
  for (auto &l : m.legs) {
    o[p++] = l.dir;
    o[p++] = (int32_t)l.qn.size();
    for (size_t s = 0; s < l.qn.size(); ++s) {
      for (int k = 0; k < m.qw; ++k) o[p++] = l.qn[s].v[k];
      o[p++] = (int32_t)l.dim[s];
    }
  }
  for (auto &c : m.coords)
    for (auto x : c) o[p++] = x;
}

MetaP permuted_meta(const Meta &m, const std::vector<int> &perm) {
  auto t = std::make_shared<Meta>();
  t->qw = m.qw;
  for (int i : perm) t->legs.push_back(m.legs[i]);
  for (auto &c : m.coords) {
    Coord n;
    for (int i : perm) n.push_back(c[i]);
    t->coords.push_back(n);
  }
  t->finalize();
  return t;
}

std::shared_ptr<ContractPlan> plan_contract(const MetaP &a, const MetaP &b, const std::vector<int> &ax,
                                            const std::vector<int> &bx) {
  if (ax.size() != bx.size()) throw std::invalid_argument("Contract: axes lists differ in length");
  if (a->qw != b->qw) throw std::invalid_argument("Contract: QN widths differ");
  int ra = a->rank(), rb = b->rank(), k = (int)ax.size();
  std::vector<char> ac(ra, 0), bc(rb, 0);
  for (int i = 0; i < k; ++i) {
    if (ax[i] < 0 || ax[i] >= ra || bx[i] < 0 || bx[i] >= rb) throw std::invalid_argument("Contract: axis out of range");
    if (ac[ax[i]] || bc[bx[i]]) throw std::invalid_argument("Contract: repeated axis");
    ac[ax[i]] = bc[bx[i]] = 1;
    if (!(a->legs[ax[i]] == b->legs[bx[i]].inverse()))
      throw std::invalid_argument("Contract: contracted indexes are not inverse of each other");
  }
  auto P = std::make_shared<ContractPlan>();
  P->a = a;
  P->b = b;
  P->ax = ax;
  P->bx = bx;
  std::vector<int> afree, bfree;
  for (int i = 0; i < ra; ++i)
    if (!ac[i]) afree.push_back(i);
  for (int i = 0; i < rb; ++i)
    if (!bc[i]) bfree.push_back(i);
  // Operand layouts.  A leg whose sectors all have dimension 1 (a spin-1/2 or
  // fermion site, for instance) leaves the memory order of every block
  // unchanged wherever it sits, so it is left out of the test: a contraction
  // whose other contracted legs lead or trail the block reads it in place.
  auto unit = [](const Leg &l) {
    for (auto d : l.dim)
      if (d != 1) return false;
    return true;
  };
  std::vector<int> apos(ra, -1), bpos(rb, -1);
  int na = 0, nb = 0;
  for (int i = 0; i < ra; ++i)
    if (!unit(a->legs[i])) apos[i] = na++;
  for (int i = 0; i < rb; ++i)
    if (!unit(b->legs[i])) bpos[i] = nb++;
  std::vector<int> axn, bxn;
  for (int i = 0; i < k; ++i)
    if (apos[ax[i]] >= 0) {
      axn.push_back(apos[ax[i]]);
      bxn.push_back(bpos[bx[i]]);
    }
  const int kn = (int)axn.size();
  bool a_trail = true, a_lead = true, b_lead = true, b_trail = true;
  for (int i = 0; i < kn; ++i) {
    a_trail &= axn[i] == na - kn + i;
    a_lead &= axn[i] == i;
    b_lead &= bxn[i] == i;
    b_trail &= bxn[i] == nb - kn + i;
  }
  // Widest output block (the product of B's free dimensions over B's blocks):
  // a narrow contraction runs on the vector-ALU kernel, one thread per output
  // row, which reads A coalesced when A is stored K x M, so a permuted A is
  // then laid out with the contracted legs leading.
  int64_t max_n = 0;
  for (int32_t j = 0; j < (int32_t)b->coords.size(); ++j) {
    int64_t n = 1;
    for (int i : bfree) n *= b->dim_of(j, i);
    max_n = std::max(max_n, n);
  }
  const bool narrow = max_n <= kNarrowMaxN;
  if (a_trail) P->la = kNormal;
  else if (a_lead) P->la = kTrans;
  else if (narrow) {
    P->permute_a = true;
    P->la = kTrans;
    P->aperm = ax;
    P->aperm.insert(P->aperm.end(), afree.begin(), afree.end());
  } else {
    P->permute_a = true;
    P->la = kNormal;
    P->aperm = afree;
    P->aperm.insert(P->aperm.end(), ax.begin(), ax.end());
  }
  if (b_lead) P->lb = kNormal;
  else if (b_trail) P->lb = kTrans;
  else {
    P->permute_b = true;
    P->lb = kNormal;
    P->bperm = bx;
    P->bperm.insert(P->bperm.end(), bfree.begin(), bfree.end());
  }
  // Output legs.
  auto out = std::make_shared<Meta>();
  out->qw = a->qw;
  for (int i : afree) out->legs.push_back(a->legs[i]);
  for (int i : bfree) out->legs.push_back(b->legs[i]);
  // Match blocks on contracted coordinates.
  std::unordered_map<Coord, std::vector<int32_t>, CoordHash> bgrp;
  for (int32_t j = 0; j < (int32_t)b->coords.size(); ++j) {
    Coord key(k);
    for (int i = 0; i < k; ++i) key[i] = b->coords[j][bx[i]];
    bgrp[key].push_back(j);
  }
  struct Raw {
    int32_t ia, ib;
  };
  std::map<Coord, std::vector<Raw>> grp;
  for (int32_t ia = 0; ia < (int32_t)a->coords.size(); ++ia) {
    const Coord &ca = a->coords[ia];
    Coord key(k);
    for (int i = 0; i < k; ++i) key[i] = ca[ax[i]];
    auto it = bgrp.find(key);
    if (it == bgrp.end()) continue;
    for (int32_t ib : it->second) {
      Coord oc;
      oc.reserve(afree.size() + bfree.size());
      for (int i : afree) oc.push_back(ca[i]);
      for (int i : bfree) oc.push_back(b->coords[ib][i]);
      grp[oc].push_back({ia, ib});
    }
  }
  for (auto &kv : grp) out->coords.push_back(kv.first);
  out->finalize();
  P->out = out;
  P->blks.reserve(out->coords.size());
  for (auto &kv : grp) {
    int32_t ob = out->find(kv.first);
    BlkH bh;
    bh.c_off = out->offset[ob];
    int64_t M = 1, N = 1;
    const Raw &r0 = kv.second[0];
    for (int i : afree) M *= a->dim_of(r0.ia, i);
    for (int i : bfree) N *= b->dim_of(r0.ib, i);
    bh.M = (int32_t)M;
    bh.N = (int32_t)N;
    bh.pair_begin = (int32_t)P->pairs.size();
    for (auto &r : kv.second) {
      int64_t K = 1;
      for (int i : ax) K *= a->dim_of(r.ia, i);
      PairH ph;
      ph.a_off = a->offset[r.ia];
      ph.b_off = b->offset[r.ib];
      ph.K = (int32_t)K;
      P->pairs.push_back(ph);
      P->flops += 2.0 * (double)M * (double)N * (double)K;
    }
    bh.pair_end = (int32_t)P->pairs.size();
    P->blks.push_back(bh);
  }
  return P;
}

std::string qn_str(const QN &q, int qw) {
  std::ostringstream s;
  s << "(";
  for (int i = 0; i < qw; ++i) s << (i ? "," : "") << q.v[i];
  s << ")";
  return s.str();
}

}  // namespace mps2
