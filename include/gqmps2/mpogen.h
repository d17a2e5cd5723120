// MPOGenerator: exact MPO from a sum of n-body terms.
//
// Public API of include/gqmps2/one_dim_tn/mpo/mpogen/mpogen.h:40-121 and the
// term expansion of mpogen_impl.h:87-245 (identity / insertion operators
// between the physical operators, optional tail string), the leg orders of the
// generated tensors (W_0 = (p_in, b_r, p_out), W_i = (b_l, p_in, p_out, b_r),
// W_{N-1} = (p_in, b_l, p_out): mpogen_impl.h:359,377,398) and the virtual
// bond quantum numbers (target QN = zero_div - Div(op) + left QN,
// mpogen_impl.h:290-299; columns grouped by QN in order of first appearance,
// mpogen_impl.h:301-349).
//
// Construction (own design, replacing the reference's FSM + symbolic
// compression): a finite-state automaton whose intermediate states are the
// distinct unfinished term prefixes (head site, operator sequence so far).
// Coefficients are applied on the last operator, so terms that share a prefix
// share a state; the ready (identity) and done states close the automaton.
#ifndef MPS2_AMD_GQMPS2_MPOGEN_H
#define MPS2_AMD_GQMPS2_MPOGEN_H

#include <algorithm>
#include <iomanip>
#include <map>

#include "gqmps2/framework.h"

namespace gqmps2 {

template <typename TenElemT, typename QNT>
class MPOGenerator {
 public:
  using IndexT = Index<QNT>;
  using QNSctT = QNSector<QNT>;
  using GQTensorT = GQTensor<TenElemT, QNT>;
  using GQTensorVec = std::vector<GQTensorT>;

  MPOGenerator(const SiteVec<TenElemT, QNT> &site_vec, const QNT &zero_div)
      : N_(site_vec.size), site_vec_(site_vec), zero_div_(zero_div) {
    for (size_t i = 0; i < N_; ++i) {
      pb_out_.push_back(site_vec.sites[i]);
      pb_in_.push_back(InverseIndex(site_vec.sites[i]));
    }
  }

  // mpogen_impl.h:87-121 (ascending site indexes; identity on skipped sites)
  void AddTerm(const TenElemT coef, const GQTensorVec &local_ops, const std::vector<size_t> &local_ops_idxs) {
    assert(local_ops.size() == local_ops_idxs.size());
    assert(std::is_sorted(local_ops_idxs.begin(), local_ops_idxs.end()));
    if (local_ops.empty()) return;
    if (local_ops_idxs.back() >= N_) throw std::out_of_range("MPOGenerator::AddTerm: site index out of range");
    if (coef == TenElemT(0)) return;
    Term t;
    t.coef = coef;
    t.head = local_ops_idxs.front();
    size_t tail = local_ops_idxs.back();
    for (size_t i = t.head; i <= tail; ++i) {
      auto it = std::find(local_ops_idxs.begin(), local_ops_idxs.end(), i);
      if (it != local_ops_idxs.end()) t.ops.push_back(Intern_(local_ops[it - local_ops_idxs.begin()]));
      else t.ops.push_back(Intern_(site_vec_.id_ops[i]));
    }
    terms_.push_back(std::move(t));
  }

  // mpogen_impl.h:144-199 (insertion operators between / after physical ops)
  void AddTerm(const TenElemT coef, const GQTensorVec &phys_ops, const std::vector<size_t> &phys_ops_idxs,
               const GQTensorVec &inst_ops, const std::vector<std::vector<size_t>> &inst_ops_idxs_set = kNullUintVecVec) {
    assert(phys_ops.size() >= 2 && phys_ops.size() == phys_ops_idxs.size());
    assert(inst_ops.size() == phys_ops.size() - 1 || inst_ops.size() == phys_ops.size());
    GQTensorVec ops;
    std::vector<size_t> idxs;
    for (size_t i = 0; i + 1 < phys_ops.size(); ++i) {
      ops.push_back(phys_ops[i]);
      idxs.push_back(phys_ops_idxs[i]);
      if (inst_ops_idxs_set == kNullUintVecVec) {
        for (size_t j = phys_ops_idxs[i] + 1; j < phys_ops_idxs[i + 1]; ++j) {
          ops.push_back(inst_ops[i]);
          idxs.push_back(j);
        }
      } else {
        for (auto j : inst_ops_idxs_set[i]) {
          ops.push_back(inst_ops[i]);
          idxs.push_back(j);
        }
      }
    }
    ops.push_back(phys_ops.back());
    idxs.push_back(phys_ops_idxs.back());
    if (inst_ops.size() == phys_ops.size()) {
      if (inst_ops_idxs_set == kNullUintVecVec) {
        for (size_t j = phys_ops_idxs.back(); j < N_; ++j) {
          ops.push_back(inst_ops.back());
          idxs.push_back(j);
        }
      } else {
        for (auto j : inst_ops_idxs_set.back()) {
          ops.push_back(inst_ops.back());
          idxs.push_back(j);
        }
      }
    }
    AddTerm(coef, ops, idxs);
  }

  // mpogen_impl.h:214-245
  void AddTerm(const TenElemT coef, const GQTensorT &op1, const size_t op1_idx, const GQTensorT &op2 = GQTensorT(),
               const size_t op2_idx = 0, const GQTensorT &inst_op = GQTensorT(),
               const std::vector<size_t> &inst_op_idxs = kNullUintVec) {
    if (op2 == GQTensorT()) {
      AddTerm(coef, GQTensorVec{op1}, std::vector<size_t>{op1_idx});
    } else if (inst_op == GQTensorT()) {
      AddTerm(coef, GQTensorVec{op1, op2}, std::vector<size_t>{op1_idx, op2_idx});
    } else if (inst_op_idxs == kNullUintVec) {
      AddTerm(coef, GQTensorVec{op1, op2}, std::vector<size_t>{op1_idx, op2_idx}, GQTensorVec{inst_op});
    } else {
      AddTerm(coef, GQTensorVec{op1, op2}, std::vector<size_t>{op1_idx, op2_idx}, GQTensorVec{inst_op},
              std::vector<std::vector<size_t>>{inst_op_idxs});
    }
  }

  MPO<GQTensorT> Gen() {
    // ---- automaton states per bond b (between site b and b+1)
    // state keys: ready = {-1}, done = {-2}, prefix = {head, op_0, ..., op_b}
    std::vector<std::map<std::vector<long>, size_t>> sid(N_ > 0 ? N_ - 1 : 0);
    std::vector<std::vector<std::vector<long>>> order(N_ > 0 ? N_ - 1 : 0);
    auto state = [&](size_t b, const std::vector<long> &key) {
      auto it = sid[b].find(key);
      if (it != sid[b].end()) return it->second;
      size_t id = order[b].size();
      sid[b][key] = id;
      order[b].push_back(key);
      return id;
    };
    const std::vector<long> kReady{-1}, kDone{-2};
    for (size_t b = 0; b + 1 < N_; ++b) state(b, kReady);
    // transitions on site i: (from state key at bond i-1, to key at bond i, coef, op)
    struct Tr {
      std::vector<long> from, to;
      TenElemT coef;
      size_t op;
    };
    std::vector<std::vector<Tr>> tr(N_);
    // An edge into a shared prefix state is added once; edges into the done
    // state carry one term's coefficient each and are summed.
    std::vector<std::map<std::pair<std::vector<long>, std::vector<long>>, bool>> seen(N_);
    for (auto &t : terms_) {
      size_t tail = t.head + t.ops.size() - 1;
      std::vector<long> prev = kReady;
      for (size_t i = t.head; i <= tail; ++i) {
        std::vector<long> cur;
        TenElemT c = 1;
        if (i == tail) {
          cur = kDone;
          c = t.coef;
          tr[i].push_back({prev, cur, c, t.ops[i - t.head]});
        } else {
          cur = {(long)t.head};
          for (size_t k = 0; k <= i - t.head; ++k) cur.push_back((long)t.ops[k]);
          state(i, cur);
          if (!seen[i].count({prev, cur})) {
            seen[i][{prev, cur}] = true;
            tr[i].push_back({prev, cur, c, t.ops[i - t.head]});
          }
        }
        prev = cur;
      }
    }
    for (size_t b = 0; b + 1 < N_; ++b) state(b, kDone);
    // ---- QN of each state and bond indexes (grouped by QN, first appearance)
    std::vector<std::map<std::vector<long>, QNT>> sqn(N_ > 0 ? N_ - 1 : 0);
    for (size_t b = 0; b + 1 < N_; ++b) {
      sqn[b][kReady] = zero_div_;
      sqn[b][kDone] = zero_div_;
    }
    for (size_t i = 0; i < N_; ++i)
      for (auto &x : tr[i]) {
        if (i + 1 >= N_ || x.to == kDone) continue;
        QNT lq = (i == 0 || x.from == kReady) ? zero_div_ : sqn[i - 1].at(x.from);
        sqn[i][x.to] = zero_div_ - Div(ops_[x.op]) + lq;
      }
    std::vector<IndexT> bonds(N_ > 0 ? N_ - 1 : 0);
    std::vector<std::map<std::vector<long>, size_t>> coor(N_ > 0 ? N_ - 1 : 0);
    for (size_t b = 0; b + 1 < N_; ++b) {
      std::vector<QNT> qorder;
      std::map<QNT, std::vector<std::vector<long>>> groups;
      for (auto &key : order[b]) {
        const QNT &q = sqn[b].at(key);
        if (!groups.count(q)) qorder.push_back(q);
        groups[q].push_back(key);
      }
      std::vector<QNSctT> secs;
      size_t c = 0;
      for (auto &q : qorder) {
        secs.emplace_back(q, groups[q].size());
        for (auto &key : groups[q]) coor[b][key] = c++;
      }
      bonds[b] = IndexT(secs, GQTenIndexDirType::OUT);
      if (Verbose()) std::cout << std::setw(3) << c << std::endl;   // bond dimension, as mpogen_impl.h:254-256
    }
    // ---- realise the tensors
    MPO<GQTensorT> mpo(N_);
    for (size_t i = 0; i < N_; ++i) {
      if (N_ == 1) throw std::invalid_argument("MPOGenerator: at least two sites are required");
      if (i == 0) mpo(i) = new GQTensorT({pb_in_[0], bonds[0], pb_out_[0]});
      else if (i == N_ - 1) mpo(i) = new GQTensorT({pb_in_[i], InverseIndex(bonds[i - 1]), pb_out_[i]});
      else mpo(i) = new GQTensorT({InverseIndex(bonds[i - 1]), pb_in_[i], pb_out_[i], bonds[i]});
      std::vector<Tr> all = tr[i];
      const size_t id = Intern_(site_vec_.id_ops[i]);
      if (i > 0 && i + 1 < N_) {
        all.push_back({kReady, kReady, TenElemT(1), id});
        all.push_back({kDone, kDone, TenElemT(1), id});
      } else if (i == 0) {
        all.push_back({kReady, kReady, TenElemT(1), id});
      } else {
        all.push_back({kDone, kDone, TenElemT(1), id});
      }
      for (auto &x : all) {
        const GQTensorT &op = ops_[x.op];
        size_t d = op.GetIndexes()[0].dim();
        for (size_t a = 0; a < d; ++a)
          for (size_t bb = 0; bb < d; ++bb) {
            double v = op.GetElem({a, bb});
            if (v == 0.0) continue;
            v *= x.coef;
            if (i == 0) {
              if (N_ > 1 && !coor[0].count(x.to)) continue;
              (*mpo(i))({a, coor[0].at(x.to), bb}) += v;
            } else if (i == N_ - 1) {
              (*mpo(i))({a, coor[i - 1].at(x.from), bb}) += v;
            } else {
              (*mpo(i))({coor[i - 1].at(x.from), a, bb, coor[i].at(x.to)}) += v;
            }
          }
      }
    }
    return mpo;
  }

 private:
  struct Term {
    TenElemT coef;
    size_t head;
    std::vector<size_t> ops;
  };
  size_t N_;
  SiteVec<TenElemT, QNT> site_vec_;
  QNT zero_div_;
  std::vector<IndexT> pb_in_, pb_out_;
  std::vector<GQTensorT> ops_;
  std::vector<Term> terms_;

  size_t Intern_(const GQTensorT &op) {
    for (size_t i = 0; i < ops_.size(); ++i)
      if (ops_[i] == op) return i;
    ops_.push_back(op);
    return ops_.size() - 1;
  }
};

}  // namespace gqmps2

#endif
