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
// compression): the channels of each bond are the distinct remainders of the
// sum of terms to the right of that bond, up to a scale (see Gen()).
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
    if (N_ < 2) throw std::invalid_argument("MPOGenerator: at least two sites are required");
    // A channel of bond b (between sites b and b+1) stands for a weighted set
    // of operator strings on sites b+1 .. N-1 (its "future"), normalised so
    // that its first string has weight 1.  Starting from the whole sum of
    // terms before site 0, each site splits every future by its first
    // operator; futures equal up to a scale share one channel (the scale goes
    // on the edge into it).  Terms with a common prefix therefore share
    // channels on the left and terms with proportional remainders share them
    // on the right, which reproduces the compressed FSM dimensions of
    // test_mpogen.cc:79-207.
    using Str = std::vector<size_t>;
    using Fut = std::map<Str, TenElemT>;
    // total order of futures (complex weights ordered by real, then imaginary part)
    struct FutLess {
      bool operator()(const Fut &x, const Fut &y) const {
        auto ex = x.begin(), ey = y.begin();
        for (; ex != x.end() && ey != y.end(); ++ex, ++ey) {
          if (ex->first != ey->first) return ex->first < ey->first;
          const double xr = std::real(ex->second), yr = std::real(ey->second);
          if (xr != yr) return xr < yr;
          const double xi = std::imag(ex->second), yi = std::imag(ey->second);
          if (xi != yi) return xi < yi;
        }
        return ex == x.end() && ey != y.end();
      }
    };
    std::vector<size_t> id_op(N_);
    for (size_t i = 0; i < N_; ++i) id_op[i] = Intern_(site_vec_.id_ops[i]);
    Fut root;
    for (auto &t : terms_) {
      Str str(N_);
      for (size_t i = 0; i < N_; ++i) str[i] = id_op[i];
      for (size_t k = 0; k < t.ops.size(); ++k) str[t.head + k] = t.ops[k];
      root[str] += t.coef;
    }
    for (auto it = root.begin(); it != root.end();) it = it->second == TenElemT(0) ? root.erase(it) : std::next(it);
    struct Edge {
      size_t from, op, to;
      TenElemT w;
    };
    std::vector<std::vector<Edge>> edges(N_);
    std::vector<std::vector<QNT>> chan_qn(N_);  // channels of bond i (i < N-1)
    std::vector<Fut> cur{root};
    std::vector<QNT> cur_qn{zero_div_};
    for (size_t i = 0; i < N_; ++i) {
      std::vector<Fut> next;
      std::map<Fut, size_t, FutLess> nidx;
      std::vector<QNT> next_qn;
      for (size_t j = 0; j < cur.size(); ++j) {
        std::map<size_t, Fut> by_op;
        for (auto &kv : cur[j]) by_op[kv.first[0]][Str(kv.first.begin() + 1, kv.first.end())] += kv.second;
        for (auto &ob : by_op) {
          Fut &f = ob.second;
          for (auto it = f.begin(); it != f.end();) it = it->second == TenElemT(0) ? f.erase(it) : std::next(it);
          if (f.empty()) continue;
          if (i + 1 == N_) {
            edges[i].push_back({j, ob.first, 0, f.begin()->second});  // the empty string's weight
            continue;
          }
          const TenElemT scale = f.begin()->second;
          for (auto &kv : f) kv.second /= scale;
          const QNT q = zero_div_ - Div(ops_[ob.first]) + cur_qn[j];
          auto found = nidx.find(f);
          size_t k;
          if (found == nidx.end()) {
            k = next.size();
            nidx.emplace(f, k);
            next.push_back(std::move(f));
            next_qn.push_back(q);
          } else {
            k = found->second;
            if (!(next_qn[k] == q)) throw std::runtime_error("MPOGenerator: terms with different quantum-number divergence");
          }
          edges[i].push_back({j, ob.first, k, scale});
        }
      }
      if (i + 1 < N_) {
        if (next.empty()) throw std::runtime_error("MPOGenerator: no terms");
        chan_qn[i] = next_qn;
        cur.swap(next);
        cur_qn.swap(next_qn);
      }
    }
    // ---- bond indexes: channels grouped by QN in order of first appearance (mpogen_impl.h:301-349)
    std::vector<IndexT> bonds(N_ - 1);
    std::vector<std::vector<size_t>> coor(N_ - 1);
    for (size_t b = 0; b + 1 < N_; ++b) {
      std::vector<QNT> qorder;
      std::vector<std::vector<size_t>> groups;
      for (size_t k = 0; k < chan_qn[b].size(); ++k) {
        size_t g = 0;
        while (g < qorder.size() && !(qorder[g] == chan_qn[b][k])) ++g;
        if (g == qorder.size()) {
          qorder.push_back(chan_qn[b][k]);
          groups.emplace_back();
        }
        groups[g].push_back(k);
      }
      std::vector<QNSctT> secs;
      coor[b].assign(chan_qn[b].size(), 0);
      size_t c = 0;
      for (size_t g = 0; g < qorder.size(); ++g) {
        secs.emplace_back(qorder[g], groups[g].size());
        for (auto k : groups[g]) coor[b][k] = c++;
      }
      bonds[b] = IndexT(secs, GQTenIndexDirType::OUT);
      if (Verbose()) std::cout << std::setw(3) << c << std::endl;   // bond dimension, as mpogen_impl.h:254-256
    }
    // ---- realise the tensors
    MPO<GQTensorT> mpo(N_);
    for (size_t i = 0; i < N_; ++i) {
      if (i == 0) mpo(i) = new GQTensorT({pb_in_[0], bonds[0], pb_out_[0]});
      else if (i == N_ - 1) mpo(i) = new GQTensorT({pb_in_[i], InverseIndex(bonds[i - 1]), pb_out_[i]});
      else mpo(i) = new GQTensorT({InverseIndex(bonds[i - 1]), pb_in_[i], pb_out_[i], bonds[i]});
      for (auto &e : edges[i]) {
        const GQTensorT &op = ops_[e.op];
        size_t d = op.GetIndexes()[0].dim();
        for (size_t a = 0; a < d; ++a)
          for (size_t bb = 0; bb < d; ++bb) {
            TenElemT v = op.GetElem({a, bb});
            if (v == TenElemT(0)) continue;
            v *= e.w;
            if (i == 0) (*mpo(i))({a, coor[0][e.to], bb}) += v;
            else if (i == N_ - 1) (*mpo(i))({a, coor[i - 1][e.from], bb}) += v;
            else (*mpo(i))({coor[i - 1][e.from], a, bb, coor[i][e.to]}) += v;
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
