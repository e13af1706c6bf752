#include "plan.hpp"

#include <algorithm>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <utility>

namespace mpir {

namespace {

// Best-fit allocator over one linear pool; free blocks are coalesced.
class PoolAlloc {
 public:
  int64_t alloc(int64_t size) {
    size = (size + 63) & ~int64_t(63);
    auto it = by_size_.lower_bound({size, -1});
    if (it != by_size_.end()) {
      const int64_t bsize = it->first, off = it->second;
      by_size_.erase(it);
      by_off_.erase(off);
      if (bsize > size) insert(off + size, bsize - size);
      return off;
    }
    const int64_t off = top_;
    top_ += size;
    return off;
  }
  void release(int64_t off, int64_t size) {
    size = (size + 63) & ~int64_t(63);
    auto nx = by_off_.lower_bound(off);
    if (nx != by_off_.end() && off + size == nx->first) {
      size += nx->second;
      by_size_.erase({nx->second, nx->first});
      nx = by_off_.erase(nx);
    }
    if (nx != by_off_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        off = pv->first;
        size += pv->second;
        by_size_.erase({pv->second, pv->first});
        by_off_.erase(pv);
      }
    }
    insert(off, size);
  }
  int64_t peak() const { return top_; }

 private:
  void insert(int64_t off, int64_t size) {
    by_off_[off] = size;
    by_size_.insert({size, off});
  }
  std::map<int64_t, int64_t> by_off_;
  std::set<std::pair<int64_t, int64_t>> by_size_;
  int64_t top_ = 0;
};

// Task lists of the dependency-driven solves from the per-level lists. A
// forward task of a front waits for its children (small fronts, gathers),
// for its gathers (column blocks) or for its column blocks (R blocks); a
// backward task waits for its parent (R partials, small fronts) or for its
// R partials (column blocks). Children sit on lower levels and a front's
// own task kinds follow each other within its level, so every wait is on
// tasks earlier in the list.
void build_dag_schedule(FactorPlan& P) {
  const Supernodes& S = P.sn;
  const int32_t ns = S.ns;
  P.dag_meta.assign(static_cast<size_t>(ns) * kMetaStride, 0);
  for (int32_t s = 0; s < ns; ++s) {
    int32_t* m = &P.dag_meta[static_cast<size_t>(s) * kMetaStride];
    m[kMetaParent] = S.sparent[s];
    m[kMetaNChild] = P.child_ptr[s + 1] - P.child_ptr[s];
    m[kMetaNBFinal] = 1;  // small fronts: one backward task
  }
  auto count = [&](const std::vector<int32_t>& tasks, int stride, int field) {
    for (size_t q = 0; q < tasks.size(); q += stride) P.dag_meta[static_cast<size_t>(tasks[q]) * kMetaStride + field]++;
  };
  for (int32_t q = 0; q < static_cast<int32_t>(P.big_list.size()); ++q)
    P.dag_meta[static_cast<size_t>(P.big_list[q]) * kMetaStride + kMetaNBFinal] = 0;
  count(P.sfwd_task, 2, kMetaNGather);
  count(P.sxc_task, 3, kMetaNXc);
  count(P.su_task, 2, kMetaNU);
  count(P.rpart_task, 3, kMetaNRpart);
  count(P.sbwd_task, 2, kMetaNBFinal);
  auto grouped = [](std::vector<Task4>& out, int32_t kind, int32_t b, int32_t e) {
    for (int32_t q = b; q < e; q += kDagGroup) out.push_back({kind, q, std::min(kDagGroup, e - q), 0});
  };
  auto single = [](std::vector<Task4>& out, int32_t kind, int32_t b, int32_t e) {
    for (int32_t q = b; q < e; ++q) out.push_back({kind, q, 1, 0});
  };
  P.dag_fwd.clear();
  P.dag_bwd.clear();
  for (int32_t l = 0; l < S.nlevels; ++l) {
    grouped(P.dag_fwd, kDagSmall, P.small_ptr[l], P.small_ptr[l + 1]);
    grouped(P.dag_fwd, kDagGather, P.sfwd_ptr[l], P.sfwd_ptr[l + 1]);
    single(P.dag_fwd, kDagXc, P.sxc_ptr[l], P.sxc_ptr[l + 1]);
    single(P.dag_fwd, kDagU, P.su_ptr[l], P.su_ptr[l + 1]);
  }
  for (int32_t l = S.nlevels - 1; l >= 0; --l) {
    single(P.dag_bwd, kDagRpart, P.rpart_ptr[l], P.rpart_ptr[l + 1]);
    single(P.dag_bwd, kDagBxc, P.sbwd_ptr[l], P.sbwd_ptr[l + 1]);
    grouped(P.dag_bwd, kDagSmall, P.small_ptr[l], P.small_ptr[l + 1]);
  }
}

}  // namespace

FactorPlan build_plan_from_supernodes(Supernodes sn) {
  FactorPlan P;
  P.sn = std::move(sn);
  Supernodes& S = P.sn;
  const int32_t ns = S.ns;
  P.panel_entries = S.panel_off[ns];

  // children lists
  P.child_ptr.assign(static_cast<size_t>(ns) + 1, 0);
  for (int32_t s = 0; s < ns; ++s)
    if (S.sparent[s] >= 0) P.child_ptr[S.sparent[s] + 1]++;
  for (int32_t s = 0; s < ns; ++s) P.child_ptr[s + 1] += P.child_ptr[s];
  P.child_list.resize(P.child_ptr[ns]);
  {
    std::vector<int32_t> next(P.child_ptr.begin(), P.child_ptr.end() - 1);
    for (int32_t s = 0; s < ns; ++s)
      if (S.sparent[s] >= 0) P.child_list[next[S.sparent[s]]++] = s;
  }

  // update-matrix pool with level lifetimes
  P.upd_off.assign(ns, -1);
  PoolAlloc pool;
  for (int32_t l = 0; l < S.nlevels; ++l) {
    for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int32_t s = S.level_sn[q];
      const int64_t nr = S.nrow[s];
      if (nr > 0) P.upd_off[s] = pool.alloc(nr * nr);
    }
    for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int32_t s = S.level_sn[q];
      for (int32_t t = P.child_ptr[s]; t < P.child_ptr[s + 1]; ++t) {
        const int32_t c = P.child_list[t];
        const int64_t nr = S.nrow[c];
        if (nr > 0) pool.release(P.upd_off[c], nr * nr);
      }
    }
  }
  P.upd_pool = pool.peak();

  // front classes, assembly tasks, blocked-factor tasks
  P.small_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.big_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.asm_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.level_step_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.pan_ptr.assign(1, 0);
  P.upd_ptr.assign(1, 0);
  P.pan_aug_ptr.clear();
  P.upd_aug_ptr.clear();
  for (int32_t l = 0; l < S.nlevels; ++l) {
    int32_t max_steps = 0;
    for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int32_t s = S.level_sn[q];
      const int32_t nc = S.first[s + 1] - S.first[s];
      const int32_t m = nc + S.nrow[s];
      if (m <= kSmallFront) {
        P.small_list.push_back(s);
        continue;
      }
      P.big_list.push_back(s);
      max_steps = std::max(max_steps, (nc + kPanelNB - 1) / kPanelNB);
      // 64 x 64 tiles of the lower part; tiles made only of panel columns
      // are kept below only if a child's update lands in them
      for (int32_t c0 = 0; c0 < m; c0 += kAsmTile) {
        const int32_t c1 = std::min(m, c0 + kAsmTile);
        for (int32_t t0 = c0; t0 < m; t0 += kAsmTile) P.asm_task.push_back({s, c0, c1, t0});
      }
    }
    P.small_ptr[l + 1] = static_cast<int32_t>(P.small_list.size());
    P.big_ptr[l + 1] = static_cast<int32_t>(P.big_list.size());
    P.asm_ptr[l + 1] = static_cast<int32_t>(P.asm_task.size());
    for (int32_t k = 0; k < max_steps; ++k) {
      // real rows first, then the identity rows below R_s (aug rows), so a
      // run that only rebuilds the inverses can launch the tail of a step
      for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
          P.pan_aug_ptr.push_back(static_cast<int32_t>(P.pan_task.size()));
          P.upd_aug_ptr.push_back(static_cast<int32_t>(P.upd_task.size()));
        }
        for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) {
          const int32_t s = P.big_list[q];
          const int32_t nc = S.first[s + 1] - S.first[s];
          const int32_t m = nc + S.nrow[s];
          const int32_t k0 = k * kPanelNB;
          if (k0 >= nc) continue;
          const int32_t nbk = std::min(kPanelNB, nc - k0);
          const int32_t t0 = k0 + nbk;
          if (pass == 0) {
            // the diagonal block's own task, then rows below it in chunks of 64
            P.pan_task.push_back({s, k0, t0, k});
            for (int32_t r0 = t0; r0 < m; r0 += kPanelRows) P.pan_task.push_back({s, r0, std::min(m, r0 + kPanelRows), k});
            for (int32_t c0 = t0; c0 < m; c0 += kTile)
              for (int32_t r0 = c0; r0 < m; r0 += kTile) P.upd_task.push_back({s, r0, c0, k});
          } else {
            // aug rows i < t0 are the only ones nonzero in this step's columns
            for (int32_t r0 = 0; r0 < t0; r0 += kPanelRows)
              P.pan_task.push_back({s, m + r0, m + std::min(t0, r0 + kPanelRows), k});
            for (int32_t c0 = t0; c0 < nc; c0 += kTile)
              for (int32_t r0 = 0; r0 < t0; r0 += kTile) P.upd_task.push_back({s, m + r0, c0, k});
          }
        }
      }
      P.pan_ptr.push_back(static_cast<int32_t>(P.pan_task.size()));
      P.upd_ptr.push_back(static_cast<int32_t>(P.upd_task.size()));
    }
    P.level_step_ptr[l + 1] = P.level_step_ptr[l] + max_steps;
  }
  P.nsmall = static_cast<int64_t>(P.small_list.size());
  P.nbig = static_cast<int64_t>(P.big_list.size());

  auto lower = [&](int32_t c, int32_t v) {
    const int32_t* rel = S.relidx.data() + S.rows_off[c];
    return static_cast<int32_t>(std::lower_bound(rel, rel + S.nrow[c], v) - rel);
  };
  // child ranges of every assembly tile; panel-only tiles without any
  // child contribution are dropped (their entries are final after the
  // scatter of the matrix values)
  {
    std::vector<Task4> kept;
    std::vector<int32_t> kept_level(static_cast<size_t>(S.nlevels) + 1, 0);
    P.asm_child_ptr.assign(1, 0);
    for (int32_t l = 0; l < S.nlevels; ++l) {
      for (int32_t t = P.asm_ptr[l]; t < P.asm_ptr[l + 1]; ++t) {
        const Task4 tk = P.asm_task[t];
        const int32_t s = tk.a, c0 = tk.b, c1 = tk.c, t0 = tk.d;
        const int32_t nc = S.first[s + 1] - S.first[s];
        const int32_t m = nc + S.nrow[s];
        const int32_t t1 = std::min(m, t0 + kAsmTile);
        const size_t before = P.asm_child.size();
        for (int32_t q = P.child_ptr[s]; q < P.child_ptr[s + 1]; ++q) {
          const int32_t c = P.child_list[q];
          if (S.nrow[c] == 0) continue;
          const int32_t b0 = lower(c, c0), b1 = lower(c, c1), a0 = lower(c, t0), a1 = lower(c, t1);
          if (b0 >= b1 || a1 <= std::max(b0, a0)) continue;
          P.asm_child.insert(P.asm_child.end(), {c, b0, b1, a0, a1});
        }
        if (c1 <= nc && P.asm_child.size() == before) continue;
        kept.push_back(tk);
        P.asm_child_ptr.push_back(static_cast<int32_t>(P.asm_child.size() / 5));
      }
      kept_level[l + 1] = static_cast<int32_t>(kept.size());
    }
    P.asm_task.swap(kept);
    P.asm_ptr.swap(kept_level);
  }

  // solve tasks of the big fronts, per level
  P.sfwd_child_ptr.assign(1, 0);
  P.part_off.assign(S.ns, -1);
  P.xfp_off.assign(S.ns, -1);
  P.rpart_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.sfwd_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.sxc_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.su_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.sbwd_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  for (int32_t l = 0; l < S.nlevels; ++l) {
    for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) {
      const int32_t s = P.big_list[q];
      const int32_t nc = S.first[s + 1] - S.first[s];
      const int32_t nr = S.nrow[s];
      const int32_t m = nc + nr;
      const int32_t ncb = (nc + kSolveRB - 1) / kSolveRB;
      const int32_t nrb = (m + kSolveRB - 1) / kSolveRB;
      for (int32_t i = 0; i < nrb; ++i) {
        P.sfwd_task.push_back(s);
        P.sfwd_task.push_back(i);
        const int32_t r0 = i * kSolveRB, r1 = std::min(m, r0 + kSolveRB);
        for (int32_t q2 = P.child_ptr[s]; q2 < P.child_ptr[s + 1]; ++q2) {
          const int32_t c = P.child_list[q2];
          const int32_t q0 = lower(c, r0), q1 = lower(c, r1);
          if (q1 > q0) P.sfwd_child.insert(P.sfwd_child.end(), {c, q0, q1});
        }
        P.sfwd_child_ptr.push_back(static_cast<int32_t>(P.sfwd_child.size() / 3));
      }
      // forward column blocks by row chunk; backward row blocks, longest first
      const int32_t nhx = (nc + kXcRows - 1) / kXcRows;
      P.xfp_off[s] = P.nxfp;
      P.nxfp += static_cast<int64_t>(ncb) * nhx * kSolveRB;
      for (int32_t j = 0; j < ncb; ++j) {
        const int32_t rend = std::min(nc, (j + 1) * kSolveRB);
        for (int32_t h = 0; h * kXcRows < rend; ++h) {
          P.sxc_task.push_back(s);
          P.sxc_task.push_back(j);
          P.sxc_task.push_back(h);
        }
      }
      for (int32_t b = 0; b * kSolveRB < nr; ++b) {
        P.su_task.push_back(s);
        P.su_task.push_back(b);
      }
      for (int32_t i = 0; i < ncb; ++i) {
        P.sbwd_task.push_back(s);
        P.sbwd_task.push_back(i);
      }
      const int32_t nch = (nr + kRpartRows - 1) / kRpartRows;
      P.part_off[s] = P.npart;
      P.npart += static_cast<int64_t>(ncb) * nch * kSolveRB;
      for (int32_t j = 0; j < ncb; ++j)
        for (int32_t c = 0; c < nch; ++c) {
          P.rpart_task.push_back(s);
          P.rpart_task.push_back(j);
          P.rpart_task.push_back(c);
        }
    }
    P.rpart_ptr[l + 1] = static_cast<int32_t>(P.rpart_task.size() / 3);
    P.sfwd_ptr[l + 1] = static_cast<int32_t>(P.sfwd_task.size() / 2);
    P.sxc_ptr[l + 1] = static_cast<int32_t>(P.sxc_task.size() / 3);
    P.su_ptr[l + 1] = static_cast<int32_t>(P.su_task.size() / 2);
    P.sbwd_ptr[l + 1] = static_cast<int32_t>(P.sbwd_task.size() / 2);
  }
  build_dag_schedule(P);
  return P;
}

namespace {
void fill_value_map(FactorPlan& P, const CscUpper& C, const std::vector<int32_t>& val_map);
}

FactorPlan build_factor_plan(const CscUpper& C, const std::vector<int32_t>& val_map,
                             const int32_t* parent, const int32_t* Lnz) {
  FactorPlan P = build_plan_from_supernodes(build_supernodes(C, parent, Lnz));
  fill_value_map(P, C, val_map);
  return P;
}

FactorPlan build_factor_plan_forced(const CscUpper& C, const std::vector<int32_t>& val_map,
                                    const int32_t* parent, const std::vector<int32_t>& first) {
  FactorPlan P = build_plan_from_supernodes(build_supernodes_forced(C, parent, first));
  fill_value_map(P, C, val_map);
  return P;
}

Analysis analyze_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, bool relax) {
  Analysis R;
  CscUpper A;
  A.n = n;
  A.p.assign(Ap, Ap + n + 1);
  A.i.assign(Ai, Ai + Ap[n]);
  A.x.assign(A.i.size(), 0.0);
  std::vector<int32_t> work(n), Lnz(n), parent(n);
  if (ordering == 1 || ordering == 2) {
    std::vector<int32_t> p0 = ordering == 1 ? amd_order(A) : nd_order(A);
    std::vector<int32_t> vm;
    CscUpper C0 = permute_upper(A, p0, &vm);
    if (etree_counts(n, C0.p.data(), C0.i.data(), work.data(), Lnz.data(), parent.data()) < 0)
      throw std::invalid_argument("etree failed on the ordered matrix");
    std::vector<int32_t> post = tree_postorder(n, parent.data());
    R.perm.resize(n);
    for (int32_t k = 0; k < n; ++k) R.perm[k] = p0[post[k]];
  } else {
    R.perm.resize(n);
    std::iota(R.perm.begin(), R.perm.end(), 0);
  }
  R.C = permute_upper(A, R.perm, &R.val_map);
  int64_t lnz = etree_counts(n, R.C.p.data(), R.C.i.data(), work.data(), Lnz.data(), parent.data());
  if (lnz == -2) throw std::overflow_error("nonzeros of L overflow int32");
  if (lnz < 0) throw std::invalid_argument("etree failed");
  if (ordering != 0 && relax) {
    RelaxedPartition rp = relax_supernodes(R.C, parent.data(), Lnz.data());
    std::vector<int32_t> perm2(n);
    for (int32_t k = 0; k < n; ++k) perm2[k] = R.perm[rp.order[k]];
    R.perm.swap(perm2);
    R.C = permute_upper(A, R.perm, &R.val_map);
    lnz = etree_counts(n, R.C.p.data(), R.C.i.data(), work.data(), Lnz.data(), parent.data());
    if (lnz < 0) throw std::logic_error("etree failed after relaxation");
    R.plan = build_factor_plan_forced(R.C, R.val_map, parent.data(), rp.first);
    R.relaxed = true;
  } else {
    R.plan = build_factor_plan(R.C, R.val_map, parent.data(), Lnz.data());
  }
  R.nnz_l = lnz;
  R.flops = static_cast<double>(n);
  for (int32_t j = 0; j < n; ++j) R.flops += static_cast<double>(Lnz[j]) * (static_cast<double>(Lnz[j]) + 2.0);
  R.parent = std::move(parent);
  R.Lnz = std::move(Lnz);
  R.C.x.clear();
  return R;
}

namespace {
void fill_value_map(FactorPlan& P, const CscUpper& C, const std::vector<int32_t>& val_map) {
  const Supernodes& S = P.sn;
  const int32_t n = C.n;
  // value destinations: C entry (i, j), i <= j, is L-view entry (j, i) of
  // the panel of the supernode that owns column i
  P.val_dst.assign(C.nnz(), -1);
  for (int32_t j = 0; j < n; ++j) {
    for (int32_t q = C.p[j]; q < C.p[j + 1]; ++q) {
      const int32_t i = C.i[q];
      const int32_t s = S.col2sn[i];
      const int32_t f = S.first[s], l = S.first[s + 1];
      const int64_t ld = panel_ld(l - f, S.nrow[s]);
      int64_t rp;
      if (j < l) {
        rp = j - f;
      } else {
        const int32_t* R = S.rows.data() + S.rows_off[s];
        const int32_t* it = std::lower_bound(R, R + S.nrow[s], j);
        if (it == R + S.nrow[s] || *it != j) throw std::logic_error("entry outside the L pattern");
        rp = (l - f) + (it - R);
      }
      P.val_dst[val_map[q]] = S.panel_off[s] + static_cast<int64_t>(i - f) * ld + rp;
    }
  }
}
}  // namespace

}  // namespace mpir
