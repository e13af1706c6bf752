#include "plan.hpp"

#include <algorithm>
#include <cstdint>
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

// Trailing-update tiles of front s: columns c0 in [c_lo, c_hi) and rows r0
// in [c0, r_hi) (lower part, `lower`) or in [0, r_hi) (the identity rows
// below the front, listed at r_off + r0), 64 x 64 each. They are listed in
// super-blocks of st x st tiles (8 by default), column block by column
// block: an XCD takes a contiguous run of a launch's list (see xcd_task in
// the factor kernels), and tiles of one super-block share their row
// operands as well as their column operands, about 1 MB for a rank-256
// update at st = 8, which stays in the XCD's L2 while the super-block runs.
void push_tiles(std::vector<Task4>& out, int32_t s, int32_t c_lo, int32_t c_hi, int32_t r_hi, int32_t r_off,
                bool lower, int32_t code, int32_t st) {
  const int32_t W = st * kTile;
  for (int32_t cb = c_lo; cb < c_hi; cb += W) {
    const int32_t ce = std::min(c_hi, cb + W);
    for (int32_t rb = lower ? cb : 0; rb < r_hi; rb += W) {
      const int32_t re = std::min(r_hi, rb + W);
      for (int32_t c0 = cb; c0 < ce; c0 += kTile) {
        const int32_t base = lower ? c0 : 0;  // rows base, base + 64, ...
        int32_t r0 = base;
        if (r0 < rb) r0 += (rb - base + kTile - 1) / kTile * kTile;
        for (; r0 < re; r0 += kTile) out.push_back({s, r_off + r0, c0, code});
      }
    }
  }
}

}  // namespace

// Task lists of the dependency-driven solves from the per-level lists. A
// forward task of a front waits for its children (small fronts, fx, fm) or
// for all its y values (fu); a backward task waits for its parent (small
// fronts, br, bm) or for its br tasks (bx). Children sit on lower levels and a
// front's own kinds follow each other within its level, so every wait is on
// tasks earlier in the list.
void build_dag_schedule(FactorPlan& P) {
  const Supernodes& S = P.sn;
  const int32_t ns = S.ns;
  P.dag_meta.assign(static_cast<size_t>(ns) * kMetaStride, 0);
  for (int32_t s = 0; s < ns; ++s) {
    int32_t* m = &P.dag_meta[static_cast<size_t>(s) * kMetaStride];
    const int32_t nc = S.first[s + 1] - S.first[s], nr = S.nrow[s];
    m[kMetaParent] = S.sparent[s];
    m[kMetaNChild] = P.child_ptr[s + 1] - P.child_ptr[s];
    m[kMetaNBFinal] = 1;  // small fronts: one backward task
    if (nc + nr > kSmallFront) {
      m[kMetaNH] = (nc + kGemvCols - 1) / kGemvCols;
      m[kMetaNHR] = (nr + kXcRows - 1) / kXcRows;
      m[kMetaNBFinal] = (nc + kSolveRB - 1) / kSolveRB;
    }
  }
  auto count = [&](const std::vector<int32_t>& tasks, int field) {
    for (size_t q = 0; q < tasks.size(); q += 3) P.dag_meta[static_cast<size_t>(tasks[q]) * kMetaStride + field]++;
  };
  count(P.fx_task, kMetaNFx);
  count(P.fu_task, kMetaNFu);
  count(P.fm_task, kMetaNFu);
  count(P.br_task, kMetaNBr);
  auto grouped = [](std::vector<Task4>& out, int32_t b, int32_t e) {
    for (int32_t q = b; q < e; q += kDagGroup) out.push_back({kDagSmall, q, std::min(kDagGroup, e - q), 0});
  };
  auto tiles = [](std::vector<Task4>& out, int32_t kind, const std::vector<int32_t>& t, int32_t b, int32_t e) {
    for (int32_t q = b; q < e; ++q) out.push_back({kind, t[3 * q], t[3 * q + 1], t[3 * q + 2]});
  };
  P.dag_fwd.clear();
  P.dag_bwd.clear();
  for (int32_t l = 0; l < S.nlevels; ++l) {
    grouped(P.dag_fwd, P.small_ptr[l], P.small_ptr[l + 1]);
    tiles(P.dag_fwd, kDagFx, P.fx_task, P.fx_ptr[l], P.fx_ptr[l + 1]);
    tiles(P.dag_fwd, kDagFu, P.fu_task, P.fu_ptr[l], P.fu_ptr[l + 1]);
    tiles(P.dag_fwd, kDagFm, P.fm_task, P.fm_ptr[l], P.fm_ptr[l + 1]);
  }
  for (int32_t l = S.nlevels - 1; l >= 0; --l) {
    tiles(P.dag_bwd, kDagBr, P.br_task, P.br_ptr[l], P.br_ptr[l + 1]);
    tiles(P.dag_bwd, kDagBx, P.bx_task, P.bx_ptr[l], P.bx_ptr[l + 1]);
    tiles(P.dag_bwd, kDagBm, P.bm_task, P.bm_ptr[l], P.bm_ptr[l + 1]);
    grouped(P.dag_bwd, P.small_ptr[l], P.small_ptr[l + 1]);
  }
}

void build_solve_schedule(FactorPlan& P, const SolveTuning& t) {
  const Supernodes& S = P.sn;
  const int32_t ns = S.ns;
  P.tuning = t;
  for (auto* v : {&P.fx_task, &P.fu_task, &P.br_task, &P.bx_task, &P.fm_task, &P.bm_task, &P.inv_task}) v->clear();
  P.nypart = P.nupart = P.nrpart = P.nxpart = 0;
  P.ncb_total = 0;
  auto narrow = [&](int32_t s) {
    const int32_t nc = S.first[s + 1] - S.first[s];
    return nc + S.nrow[s] > kSmallFront && nc <= kSolveRB && S.nrow[s] <= t.narrow_rows;
  };
  // solve tasks of the big fronts, per level
  P.fx_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.fu_ptr = P.br_ptr = P.bx_ptr = P.fm_ptr = P.bm_ptr = P.fx_ptr;
  P.cb_base.assign(ns, 0);
  P.ypart_off.assign(ns, 0);
  P.upart_off.assign(ns, 0);
  P.rpart_off.assign(ns, 0);
  P.xpart_off.assign(ns, 0);
  for (int32_t l = 0; l < S.nlevels; ++l) {
    for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) {
      const int32_t s = P.big_list[q];
      const int32_t nc = S.first[s + 1] - S.first[s];
      const int32_t nr = S.nrow[s];
      const int32_t ncb = (nc + kSolveRB - 1) / kSolveRB;
      const int32_t nrb = (nr + kSolveRB - 1) / kSolveRB;
      const int32_t nh = (nc + kGemvCols - 1) / kGemvCols;
      const int32_t nhx = (nc + kXcRows - 1) / kXcRows;
      const int32_t nhr = (nr + kXcRows - 1) / kXcRows;
      // the factor's W pass: rows up to the block's last column (W is upper
      // triangular)
      for (int32_t j = 0; j < ncb; ++j) {
        const int32_t rend = std::min(nc, (j + 1) * kSolveRB);
        for (int32_t h = 0; h * kXcRows < rend; ++h) P.inv_task.insert(P.inv_task.end(), {s, j, h});
      }
      P.cb_base[s] = P.ncb_total;
      P.ncb_total += ncb;
      P.ypart_off[s] = P.nypart;
      P.nypart += static_cast<int64_t>(ncb) * nhx * kSolveRB;
      if (narrow(s)) {
        const int32_t nch = std::max<int32_t>(1, (nr + kNarrowChunk - 1) / kNarrowChunk);
        for (int32_t h = 0; h < nch; ++h) P.fm_task.insert(P.fm_task.end(), {s, h, 0});
        P.bm_task.insert(P.bm_task.end(), {s, 0, 0});
        continue;
      }
      P.upart_off[s] = P.nupart;
      P.nupart += static_cast<int64_t>(nrb) * nh * kSolveRB;
      P.rpart_off[s] = P.nrpart;
      P.nrpart += static_cast<int64_t>(ncb) * nhr * kSolveRB;
      P.xpart_off[s] = P.nxpart;
      P.nxpart += static_cast<int64_t>(ncb) * nh * kSolveRB;
      for (int32_t j = 0; j < ncb; ++j) {
        const int32_t rend = std::min(nc, (j + 1) * kSolveRB);
        for (int32_t h = 0; h * kXcRows < rend; ++h) P.fx_task.insert(P.fx_task.end(), {s, j, h});
      }
      for (int32_t b = 0; b < nrb; ++b)
        for (int32_t h = 0; h < nh; ++h) P.fu_task.insert(P.fu_task.end(), {s, b, h});
      for (int32_t j = 0; j < ncb; ++j)
        for (int32_t h = 0; h < nhr; ++h) P.br_task.insert(P.br_task.end(), {s, j, h});
      // bx: columns from the block's first row on
      for (int32_t i = 0; i < ncb; ++i)
        for (int32_t h = i * kSolveRB / kGemvCols; h < nh; ++h) P.bx_task.insert(P.bx_task.end(), {s, i, h});
    }
    P.fx_ptr[l + 1] = static_cast<int32_t>(P.fx_task.size() / 3);
    P.fu_ptr[l + 1] = static_cast<int32_t>(P.fu_task.size() / 3);
    P.br_ptr[l + 1] = static_cast<int32_t>(P.br_task.size() / 3);
    P.bx_ptr[l + 1] = static_cast<int32_t>(P.bx_task.size() / 3);
    P.fm_ptr[l + 1] = static_cast<int32_t>(P.fm_task.size() / 3);
    P.bm_ptr[l + 1] = static_cast<int32_t>(P.bm_task.size() / 3);
  }
  // row gathers: children in order, each child's rows ascending
  {
    P.frow_off.assign(static_cast<size_t>(ns) + 1, 0);
    for (int32_t s = 0; s < ns; ++s) {
      const int64_t m = static_cast<int64_t>(S.first[s + 1] - S.first[s]) + S.nrow[s];
      const int64_t next = P.frow_off[s] + m;
      if (next > INT32_MAX) throw std::length_error("front rows exceed the 32-bit gather index");
      P.frow_off[s + 1] = static_cast<int32_t>(next);
    }
    std::vector<int32_t> cnt(static_cast<size_t>(P.frow_off[ns]) + 1, 0);
    for (int32_t s = 0; s < ns; ++s)
      for (int32_t q = P.child_ptr[s]; q < P.child_ptr[s + 1]; ++q) {
        const int32_t c = P.child_list[q];
        for (int32_t r = 0; r < S.nrow[c]; ++r) cnt[P.frow_off[s] + S.relidx[S.rows_off[c] + r] + 1]++;
      }
    for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
    P.grow_ptr = cnt;
    P.gent.assign(static_cast<size_t>(cnt.back()), Task4{0, 0, 0, 0});
    std::vector<int32_t> next(cnt.begin(), cnt.end() - 1);
    for (int32_t s = 0; s < ns; ++s)
      for (int32_t q = P.child_ptr[s]; q < P.child_ptr[s + 1]; ++q) {
        const int32_t c = P.child_list[q];
        const int32_t ncc = S.first[c + 1] - S.first[c];
        // small and narrow children leave u final in the update vector
        const bool partial = ncc + S.nrow[c] > kSmallFront && !narrow(c);
        const int32_t nhc = partial ? (ncc + kGemvCols - 1) / kGemvCols : 0;
        for (int32_t r = 0; r < S.nrow[c]; ++r) {
          const int64_t src = S.rows_off[c] + r;
          const int64_t pb = partial ? P.upart_off[c] + static_cast<int64_t>(r / kSolveRB) * nhc * kSolveRB + r % kSolveRB : 0;
          if (src > INT32_MAX || pb > INT32_MAX) throw std::length_error("update vectors exceed the 32-bit gather index");
          P.gent[next[P.frow_off[s] + S.relidx[S.rows_off[c] + r]]++] =
              Task4{static_cast<int32_t>(src), static_cast<int32_t>(pb), nhc, 0};
        }
      }
    // two unused entries past the end: a row's first two entries are read
    // before its count is known
    P.gent.push_back(Task4{0, 0, 0, 0});
    P.gent.push_back(Task4{0, 0, 0, 0});
  }
  build_dag_schedule(P);
}

void build_factor_steps(FactorPlan& P, const FactorTuning& t) {
  if (t.upd_group < 1 || t.upd_group > 64 || t.super_tiles < 1 || t.super_tiles > 64)
    throw std::invalid_argument("factor tuning out of range");
  const Supernodes& S = P.sn;
  P.ftuning = t;
  P.level_step_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.pan_ptr.assign(1, 0);
  P.upd_ptr.assign(1, 0);
  P.pan_aug_ptr.clear();
  P.upd_aug_ptr.clear();
  P.pan_task.clear();
  P.upd_task.clear();
  for (int32_t l = 0; l < S.nlevels; ++l) {
    int32_t max_steps = 0;
    for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) {
      const int32_t s = P.big_list[q];
      max_steps = std::max(max_steps, (S.first[s + 1] - S.first[s] + kPanelNB - 1) / kPanelNB);
    }
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
          // the step's group of t.upd_group steps ends at column gend; its
          // trailing columns up to gend take this step's rank-64 update,
          // those right of it the whole group's update at its last step
          const int32_t nk = (nc + kPanelNB - 1) / kPanelNB;
          const int32_t g0 = k / t.upd_group * t.upd_group;
          const int32_t glast = std::min(nk - 1, g0 + t.upd_group - 1);
          const int32_t gend = std::min(nc, (glast + 1) * kPanelNB);
          // In a front's last group gend = nc: the last near tile may reach
          // past nc, into the columns of the update matrix that the group's
          // far tiles cover, so its near tiles keep to the panel columns.
          const bool last_group = glast == nk - 1;
          const int32_t near = k | (1 << kUpdStepsShift) | ((last_group ? 1 : 0) << kUpdPanelBit);
          const int32_t far = k | ((k - g0 + 1) << kUpdStepsShift);
          const bool group_end = k == glast;
          if (pass == 0) {
            // the diagonal block's own task, then rows below it in chunks of 64
            P.pan_task.push_back({s, k0, t0, k});
            for (int32_t r0 = t0; r0 < m; r0 += kPanelRows) P.pan_task.push_back({s, r0, std::min(m, r0 + kPanelRows), k});
            push_tiles(P.upd_task, s, t0, gend, m, 0, true, near, t.super_tiles);
            if (group_end) push_tiles(P.upd_task, s, gend, m, m, 0, true, far, t.super_tiles);
          } else {
            // aug rows i < t0 are the only ones nonzero in this step's columns
            for (int32_t r0 = 0; r0 < t0; r0 += kPanelRows)
              P.pan_task.push_back({s, m + r0, m + std::min(t0, r0 + kPanelRows), k});
            push_tiles(P.upd_task, s, t0, gend, t0, m, false, near, t.super_tiles);
            if (group_end) push_tiles(P.upd_task, s, gend, nc, t0, m, false, far, t.super_tiles);
          }
        }
      }
      P.pan_ptr.push_back(static_cast<int32_t>(P.pan_task.size()));
      P.upd_ptr.push_back(static_cast<int32_t>(P.upd_task.size()));
    }
    P.level_step_ptr[l + 1] = P.level_step_ptr[l] + max_steps;
  }
  // the dependency-driven list follows the steps when it is in use (its
  // pool layout is); the assembly lists are final once the plan is built
  if (P.dag_pool && P.asm_child_ptr.size() == P.asm_task.size() + 1) build_factor_dag(P);
}

int32_t factor_dag_level(const FactorPlan& P) {
  int32_t split = 0;
  for (int32_t l = 0; l < P.sn.nlevels; ++l)
    if (P.small_ptr[l + 1] > P.small_ptr[l]) split = l + 1;
  return split;
}

// The list runs level by level. Within a level: the assembly tiles, the
// fronts' first diagonal blocks and panel rows, then per step k: the update
// tiles of column block k + 1, the diagonal blocks and panel rows of step
// k + 1 (they need only those tiles), the step's other update tiles. So a
// front's next panel step is handed out ahead of the bulk of the current
// step's trailing update (one step of lookahead), and every wait is on
// tasks earlier in the list.
void build_factor_dag(FactorPlan& P) {
  const Supernodes& S = P.sn;
  const int32_t ns = S.ns, L = S.nlevels;
  const int32_t split = factor_dag_level(P);
  P.fdag_level = split;
  P.fdag.clear();
  P.fdag_meta.assign(static_cast<size_t>(ns) * kFMetaStride, 0);
  for (int32_t s = 0; s < ns; ++s) P.fdag_meta[static_cast<size_t>(s) * kFMetaStride + kFMetaParent] = -1;
  P.fdag_nsteps = P.fdag_ncb = 0;
  if (split >= L) return;
  auto meta = [&](int32_t s) { return &P.fdag_meta[static_cast<size_t>(s) * kFMetaStride]; };
  std::vector<char> in_dag(ns, 0);
  std::vector<int32_t> nk(ns, 0);
  int32_t nsteps = 0, ncb = 0;
  for (int32_t l = split; l < L; ++l)
    for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) {
      const int32_t s = P.big_list[q];
      const int32_t nc = S.first[s + 1] - S.first[s], m = nc + S.nrow[s];
      in_dag[s] = 1;
      nk[s] = (nc + kPanelNB - 1) / kPanelNB;
      meta(s)[kFMetaStep] = nsteps;
      meta(s)[kFMetaCb] = ncb;
      nsteps += nk[s];
      ncb += (m + kTile - 1) / kTile;
    }
  P.fdag_nsteps = nsteps;
  P.fdag_ncb = ncb;
  const FactorDagIndex X{ns, nsteps};
  std::vector<int32_t> n_asm(ns, 0), n_child(ns, 0), n_pan(nsteps, 0);
  for (int32_t s = 0; s < ns; ++s) {
    if (!in_dag[s]) continue;
    const int32_t p = S.sparent[s];
    if (p < 0) continue;
    if (!in_dag[p]) throw std::logic_error("factor DAG: parent of a front in the launch is not in it");
    meta(s)[kFMetaParent] = p;
    n_child[p]++;
  }
  for (int32_t t = P.asm_ptr[split]; t < P.asm_ptr[L]; ++t) n_asm[P.asm_task[t].a]++;
  const int32_t step_lo = P.level_step_ptr[split], step_hi = P.level_step_ptr[L];
  // update tiles per (front, column block, last step): prior counts
  std::vector<int64_t> cs_off(static_cast<size_t>(ncb) + 1, 0);
  for (int32_t s = 0; s < ns; ++s) {
    if (!in_dag[s]) continue;
    const int32_t m = S.first[s + 1] - S.first[s] + S.nrow[s];
    for (int32_t j = 0; j < (m + kTile - 1) / kTile; ++j) cs_off[meta(s)[kFMetaCb] + j + 1] = nk[s];
  }
  for (int32_t b = 0; b < ncb; ++b) cs_off[b + 1] += cs_off[b];
  std::vector<int32_t> cs(static_cast<size_t>(cs_off[ncb]), 0);
  auto upd_fields = [&](const Task4& tk, int32_t& kfirst, int32_t& klast) {
    klast = tk.d & ((1 << kUpdStepsShift) - 1);
    const int32_t steps = (tk.d >> kUpdStepsShift) & ((1 << (kUpdPanelBit - kUpdStepsShift)) - 1);
    kfirst = klast - steps + 1;
  };
  for (int32_t t = P.upd_ptr[step_lo]; t < P.upd_ptr[step_hi]; ++t) {
    const Task4 tk = P.upd_task[t];
    int32_t kf, kl;
    upd_fields(tk, kf, kl);
    meta(tk.a)[kFMetaNUpd]++;
    if (tk.c % kTile == 0) cs[cs_off[meta(tk.a)[kFMetaCb] + tk.c / kTile] + kl]++;
  }
  // exclusive prefix over the steps: cs[b, k] = tiles of block b with last
  // step < k; col_total[b] = all of them
  std::vector<int32_t> col_total(ncb, 0);
  for (int32_t b = 0; b < ncb; ++b) {
    int32_t run = 0;
    for (int64_t q = cs_off[b]; q < cs_off[b + 1]; ++q) {
      const int32_t c = cs[q];
      cs[q] = run;
      run += c;
    }
    col_total[b] = run;
  }
  auto prior = [&](int32_t s, int32_t b, int32_t k) -> int32_t {
    const int32_t cb = meta(s)[kFMetaCb] + b;
    return k >= nk[s] ? col_total[cb] : cs[cs_off[cb] + k];
  };
  for (int32_t s = 0; s < ns; ++s)
    if (in_dag[s] && meta(s)[kFMetaParent] >= 0 && meta(s)[kFMetaNUpd] == 0)
      throw std::logic_error("factor DAG: a front with a parent has no update tasks");
  // panel tasks: the diagonal block (row0 = k0), other rows; the block's own
  // identity rows are written by the diagonal-block task
  auto pan_kind = [&](const Task4& tk) {
    const int32_t s = tk.a, k0 = tk.d * kPanelNB;
    const int32_t m = S.first[s + 1] - S.first[s] + S.nrow[s];
    return tk.b == k0 ? kFDiag : tk.b == m + k0 ? -1 : kFPan;
  };
  for (int32_t t = P.pan_ptr[step_lo]; t < P.pan_ptr[step_hi]; ++t) {
    const Task4 tk = P.pan_task[t];
    if (pan_kind(tk) >= 0) n_pan[meta(tk.a)[kFMetaStep] + tk.d]++;
  }
  auto panel_wait = [&](int32_t s, int32_t k, int32_t& w, int32_t& tgt) {
    if (k == 0) {
      w = X.asm_done(s);
      tgt = n_asm[s];
    } else {
      w = X.col_done(meta(s)[kFMetaCb] + k);
      tgt = prior(s, k, k);
    }
  };
  auto emit_panels = [&](int32_t g) {
    for (int32_t t = P.pan_ptr[g]; t < P.pan_ptr[g + 1]; ++t) {
      const Task4 tk = P.pan_task[t];
      const int32_t kind = pan_kind(tk);
      if (kind < 0) continue;
      const int32_t s = tk.a, k = tk.d;
      FTask f{s | (kind << kFKindShift), 0, 0, 0, 0, 0, -1, 0, -1, 0};
      panel_wait(s, k, f.wa, f.ta);
      if (kind == kFDiag) {
        f.x = k;
      } else {
        f.x = tk.b;
        f.y = tk.c;
        f.z = k;
        f.wb = X.diag_done(meta(s)[kFMetaStep] + k);
        f.tb = 1;
      }
      P.fdag.push_back(f);
    }
  };
  // next = true: the tiles of column block k + 1 only; false: the others
  auto emit_updates = [&](int32_t g, bool next) {
    for (int32_t t = P.upd_ptr[g]; t < P.upd_ptr[g + 1]; ++t) {
      const Task4 tk = P.upd_task[t];
      int32_t kf, kl;
      upd_fields(tk, kf, kl);
      const bool is_next = tk.c == (kl + 1) * kTile;
      if (is_next != next) continue;
      const int32_t s = tk.a;
      const int32_t m = S.first[s + 1] - S.first[s] + S.nrow[s];
      const int32_t st = meta(s)[kFMetaStep] + kl;
      FTask f{s | (kFUpd << kFKindShift), tk.b, tk.c, tk.d, X.pan_done(st), n_pan[st], -1, 0, -1, 0};
      const int32_t b0 = tk.c / kTile, b1 = (std::min(m, tk.c + kTile) - 1) / kTile;
      f.wb = X.col_done(meta(s)[kFMetaCb] + b0);
      f.tb = prior(s, b0, kf);
      if (tk.c % kTile != 0 && b1 > b0) {
        f.wc = X.col_done(meta(s)[kFMetaCb] + b1);
        f.tc = prior(s, b1, kf);
      }
      P.fdag.push_back(f);
    }
  };
  for (int32_t l = split; l < L; ++l) {
    for (int32_t t = P.asm_ptr[l]; t < P.asm_ptr[l + 1]; ++t) {
      const int32_t s = P.asm_task[t].a;
      P.fdag.push_back(FTask{s | (kFAsm << kFKindShift), t, 0, 0, X.child_done(s), n_child[s], -1, 0, -1, 0});
    }
    const int32_t g0 = P.level_step_ptr[l], g1 = P.level_step_ptr[l + 1];
    if (g1 > g0) emit_panels(g0);
    for (int32_t g = g0; g < g1; ++g) {
      emit_updates(g, true);
      if (g + 1 < g1) emit_panels(g + 1);
      emit_updates(g, false);
    }
  }
}

// Update-matrix pool: U_s at upd_off[s], allocated level by level. A
// child's block is released once its parent's level has been assembled,
// except from the first level of the dependency-driven factorization on
// when `dag`: there fronts of different levels run at once, so a block is
// kept (a front of a higher level that is not its ancestor could otherwise
// reuse it before its parent has read it).
void build_update_pool(FactorPlan& P, bool dag) {
  const Supernodes& S = P.sn;
  const int32_t ns = S.ns;
  int32_t dag_level = 0;
  for (int32_t s = 0; s < ns; ++s)
    if (S.first[s + 1] - S.first[s] + S.nrow[s] <= kSmallFront) dag_level = std::max(dag_level, S.level[s] + 1);
  P.fdag_level = dag_level;
  if (!dag) dag_level = S.nlevels;
  P.upd_off.assign(ns, -1);
  PoolAlloc pool;
  for (int32_t l = 0; l < S.nlevels; ++l) {
    for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int32_t s = S.level_sn[q];
      const int64_t nr = S.nrow[s];
      if (nr > 0) P.upd_off[s] = pool.alloc(nr * nr);
    }
    if (l >= dag_level) continue;
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
  P.dag_pool = dag;
  // the dependency-driven task list exists only for a pool laid out for it
  // (built here once the assembly lists are final)
  if (dag && P.asm_child_ptr.size() == P.asm_task.size() + 1) {
    build_factor_dag(P);
  } else if (!dag) {
    P.fdag.clear();
    P.fdag_meta.clear();
    P.fdag_nsteps = P.fdag_ncb = 0;
  }
}

FactorPlan build_plan_from_supernodes(Supernodes sn, const FactorTuning& ft, const SolveTuning& st) {
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

  build_update_pool(P, false);

  // front classes, assembly tasks, blocked-factor tasks
  P.small_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.big_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  P.asm_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  for (int32_t l = 0; l < S.nlevels; ++l) {
    for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int32_t s = S.level_sn[q];
      const int32_t nc = S.first[s + 1] - S.first[s];
      const int32_t m = nc + S.nrow[s];
      if (m <= kSmallFront) {
        P.small_list.push_back(s);
        continue;
      }
      P.big_list.push_back(s);
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
  }
  build_factor_steps(P, ft);
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

  build_solve_schedule(P, st);
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

Analysis analyze_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, int relax) {
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
  if (ordering != 0 && relax > 0) {
    RelaxedPartition rp = relax_supernodes(R.C, parent.data(), Lnz.data(), relax);
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
