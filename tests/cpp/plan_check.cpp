// Host-only checks of the factor and solve schedules of mpir_amd/csrc/plan.cpp
// for several tunings (FactorTuning, SolveTuning), on quasi-definite KKT
// patterns from the generator. For every big front:
//   * each trailing-update entry of each panel step is applied exactly once,
//     by a task launched after that step's panel task and, for a panel
//     column, before the panel step of its column block; real rows take the
//     lower part of every trailing column, identity rows the panel columns
//     (an identity row still zero in the step's columns may take it once);
//   * each panel step's tasks cover its rows exactly once;
//   * the solve tasks cover the front once: narrow fronts (nc <= 64, nrow <=
//     narrow_rows) with one fm task per 256 rows and one bm task, the others
//     with the fx / fu / br / bx grids;
//   * gather entries read partial sums only from split (not narrow or small)
//     children, and every task of the dependency-driven lists comes after
//     the tasks it waits on.
// The dependency-driven factorization list is checked against the per-level
// lists (check_factor_dag). Mistakes planted in a copy of a plan (an update
// task dropped or doubled, a narrow-front task dropped; a diagonal block moved
// after its dependents, a short wait, a dropped panel task) must be caught. Prints one line per case and
// "plan_check ok" at the end; exits 1 on the first violation.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>
#include <string>
#include <vector>

#include "plan.hpp"

using namespace mpir;

namespace {

struct Violation {
  std::string what;
};
[[noreturn]] void fail(const std::string& what) { throw Violation{what}; }

void check_factor(const FactorPlan& P) {
  const Supernodes& S = P.sn;
  // per big front: its position in the step schedule (level, first step)
  for (int32_t l = 0; l < S.nlevels; ++l) {
    for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) {
      const int32_t s = P.big_list[q];
      const int32_t nc = S.first[s + 1] - S.first[s];
      const int32_t m = nc + S.nrow[s];
      const int32_t nk = (nc + kPanelNB - 1) / kPanelNB;
      const int64_t rows = static_cast<int64_t>(m) + nc;  // real rows, then identity rows
      // applied[(r * m + c) * nk + kk] counts applications of step kk
      std::vector<uint8_t> applied(static_cast<size_t>(rows) * m * nk, 0);
      for (int32_t k = 0; k < nk; ++k) {
        const int32_t g = P.level_step_ptr[l] + k;
        if (g >= P.level_step_ptr[l + 1]) fail("front has more steps than its level");
        // panel rows of step k
        const int32_t k0 = k * kPanelNB, t0 = std::min(nc, k0 + kPanelNB);
        std::vector<uint8_t> prow(static_cast<size_t>(rows), 0);
        for (int32_t t = P.pan_ptr[g]; t < P.pan_ptr[g + 1]; ++t) {
          const Task4 tk = P.pan_task[t];
          if (tk.a != s) continue;
          if (tk.d != k) fail("panel task of another step");
          for (int32_t r = tk.b; r < tk.c; ++r) prow[r]++;
        }
        for (int64_t r = 0; r < rows; ++r) {
          const bool want = (r >= k0 && r < m) || (r >= m && r < m + t0);
          if (prow[r] != (want ? 1 : 0)) fail("panel rows of front " + std::to_string(s) + " step " + std::to_string(k));
        }
        // update tasks launched after panel step k
        for (int32_t t = P.upd_ptr[g]; t < P.upd_ptr[g + 1]; ++t) {
          const Task4 tk = P.upd_task[t];
          if (tk.a != s) continue;
          const int32_t klast = tk.d & ((1 << kUpdStepsShift) - 1);
          const int32_t steps = (tk.d >> kUpdStepsShift) & ((1 << (kUpdPanelBit - kUpdStepsShift)) - 1);
          const bool panel_only = (tk.d >> kUpdPanelBit) & 1;
          if (klast != k || steps < 1 || steps > k + 1) fail("update task steps");
          const int32_t r0 = tk.b, c0 = tk.c;
          const int32_t rlim = r0 >= m ? m + nc : m;  // as the update kernel
          for (int32_t r = r0; r < std::min(r0 + kTile, rlim); ++r)
            for (int32_t c = c0; c < std::min(c0 + kTile, m); ++c) {
              if (r < c || !(c < nc || (r < m && !panel_only))) continue;
              for (int32_t kk = k - steps + 1; kk <= k; ++kk) {
                // a panel column must be final before its own panel step
                if (c < nc && c / kPanelNB <= k) fail("update after the panel step of its column");
                applied[(static_cast<size_t>(r) * m + c) * nk + kk]++;
              }
            }
        }
      }
      for (int64_t r = 0; r < rows; ++r)
        for (int32_t c = 0; c < m; ++c)
          for (int32_t kk = 0; kk < nk; ++kk) {
            const int32_t t0 = std::min(nc, (kk + 1) * kPanelNB);
            const bool want = c >= t0 && ((r < m && r >= c) || (r >= m && r - m < t0 && c < nc));
            // identity rows i >= t0 are still zero in step kk's columns: a
            // group's update may apply the step to them (adding zero), once
            const bool zero = c >= t0 && r >= m && r - m >= t0 && c < nc;
            const int got = applied[(static_cast<size_t>(r) * m + c) * nk + kk];
            if (zero ? got > 1 : got != (want ? 1 : 0))
              fail("update of front " + std::to_string(s) + " entry (" + std::to_string(r) + ", " + std::to_string(c) +
                   ") step " + std::to_string(kk) + ": applied " +
                   std::to_string(applied[(static_cast<size_t>(r) * m + c) * nk + kk]) + " times");
          }
    }
  }
}

void check_solve(const FactorPlan& P) {
  const Supernodes& S = P.sn;
  const int32_t ns = S.ns;
  std::vector<int> kind(ns, 0);  // 0 small, 1 split, 2 narrow
  std::vector<int32_t> nfm(ns, 0), nbm(ns, 0), nfx(ns, 0), nfu(ns, 0), nbr(ns, 0), nbx(ns, 0);
  for (int32_t s : P.big_list) kind[s] = 1;
  for (size_t t = 0; t < P.fm_task.size(); t += 3) {
    const int32_t s = P.fm_task[t], h = P.fm_task[t + 1];
    if (h != nfm[s]++) fail("fm chunks out of order");
    kind[s] = 2;
  }
  for (size_t t = 0; t < P.bm_task.size(); t += 3) nbm[P.bm_task[t]]++;
  for (size_t t = 0; t < P.fx_task.size(); t += 3) nfx[P.fx_task[t]]++;
  for (size_t t = 0; t < P.fu_task.size(); t += 3) nfu[P.fu_task[t]]++;
  for (size_t t = 0; t < P.br_task.size(); t += 3) nbr[P.br_task[t]]++;
  for (size_t t = 0; t < P.bx_task.size(); t += 3) nbx[P.bx_task[t]]++;
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t nc = S.first[s + 1] - S.first[s], nr = S.nrow[s];
    const bool big = nc + nr > kSmallFront;
    const bool narrow = big && nc <= kSolveRB && nr <= P.tuning.narrow_rows;
    if ((kind[s] == 2) != narrow) fail("narrow front classification of " + std::to_string(s));
    if (narrow) {
      const int32_t nch = std::max<int32_t>(1, (nr + kNarrowChunk - 1) / kNarrowChunk);
      if (nfm[s] != nch || nbm[s] != 1 || nfx[s] || nfu[s] || nbr[s] || nbx[s]) fail("narrow front tasks");
    } else if (big) {
      const int32_t ncb = (nc + kSolveRB - 1) / kSolveRB, nrb = (nr + kSolveRB - 1) / kSolveRB;
      const int32_t nh = (nc + kGemvCols - 1) / kGemvCols, nhr = (nr + kXcRows - 1) / kXcRows;
      int32_t want_fx = 0, want_bx = 0;
      for (int32_t j = 0; j < ncb; ++j) want_fx += (std::min(nc, (j + 1) * kSolveRB) + kXcRows - 1) / kXcRows;
      for (int32_t i = 0; i < ncb; ++i) want_bx += nh - i * kSolveRB / kGemvCols;
      if (nfx[s] != want_fx || nfu[s] != nrb * nh || nbr[s] != ncb * nhr || nbx[s] != want_bx || nfm[s] || nbm[s])
        fail("split front tasks of " + std::to_string(s));
    } else if (nfm[s] || nbm[s] || nfx[s] || nfu[s] || nbr[s] || nbx[s]) {
      fail("small front with big-front tasks");
    }
    const int32_t* mt = &P.dag_meta[static_cast<size_t>(s) * kMetaStride];
    if (big && mt[kMetaNFu] != (narrow ? nfm[s] : nfu[s])) fail("NFu of front " + std::to_string(s));
  }
  // gather entries: partial sums only from split children
  for (int32_t s = 0; s < ns; ++s)
    for (int32_t q = P.child_ptr[s]; q < P.child_ptr[s + 1]; ++q) {
      const int32_t c = P.child_list[q];
      for (int32_t r = 0; r < S.nrow[c]; ++r) {
        const int32_t row = P.frow_off[s] + S.relidx[S.rows_off[c] + r];
        bool found = false;
        for (int32_t e = P.grow_ptr[row]; e < P.grow_ptr[row + 1]; ++e) {
          const Task4 g = P.gent[e];
          if (g.a != S.rows_off[c] + r) continue;
          found = true;
          const int32_t want = kind[c] == 1 ? (S.first[c + 1] - S.first[c] + kGemvCols - 1) / kGemvCols : 0;
          if (g.c != want) fail("gather partial sums of child " + std::to_string(c));
        }
        if (!found) fail("missing gather entry");
      }
    }
  // dependency order of the task lists
  auto front_of = [&](const Task4& t, std::vector<int32_t>& out) {
    out.clear();
    if (t.a == kDagSmall)
      for (int32_t i = 0; i < t.c; ++i) out.push_back(P.small_list[t.b + i]);
    else
      out.push_back(t.b);
  };
  std::vector<int32_t> fr;
  std::vector<int64_t> last(ns, -1), first(ns, -1);
  for (size_t i = 0; i < P.dag_fwd.size(); ++i) {
    front_of(P.dag_fwd[i], fr);
    for (int32_t s : fr) {
      for (int32_t q = P.child_ptr[s]; q < P.child_ptr[s + 1]; ++q)
        if (last[P.child_list[q]] < 0 || last[P.child_list[q]] >= static_cast<int64_t>(i))
          fail("forward task before its children");
      last[s] = static_cast<int64_t>(i);
    }
  }
  std::fill(last.begin(), last.end(), -1);
  for (size_t i = 0; i < P.dag_bwd.size(); ++i) {
    front_of(P.dag_bwd[i], fr);
    for (int32_t s : fr) {
      const int32_t p = S.sparent[s];
      if (p >= 0 && (last[p] < 0 || last[p] >= static_cast<int64_t>(i))) fail("backward task before its parent");
      last[s] = static_cast<int64_t>(i);
    }
  }
}

// The dependency-driven factorization list (plan.hpp, FTask) against the
// per-level lists it replaces:
//   * it holds exactly the levels' assembly, panel and update tasks from
//     fdag_level on (a step's own identity rows excepted: the diagonal-block
//     task writes them), each once;
//   * every wait names the counter and the count that the task's
//     predecessors reach, and all those predecessors come earlier in the
//     list: a front's assembly after all of its children's update tiles; a
//     panel step after every update of its column block (after the
//     assembly at step 0), its rows after its diagonal block; an update
//     tile after every panel task of its last step and after every earlier
//     tile of the column blocks it overlaps;
//   * tiles that do not start on a 64-column block touch only the update
//     matrix (columns >= nc) at the front's last step;
//   * the update-matrix blocks read or written inside the launch are
//     disjoint;
//   * a run of the list by a few workgroups that take tasks in list order
//     and wait as the kernel does, finishing tasks in random order,
//     completes every task.
void check_factor_dag(const FactorPlan& P) {
  const Supernodes& S = P.sn;
  const int32_t ns = S.ns, L = S.nlevels, split = P.fdag_level;
  if (split != factor_dag_level(P)) fail("fdag_level");
  if (split >= L) {
    if (!P.fdag.empty()) fail("DAG tasks without DAG levels");
    return;
  }
  for (int32_t l = split; l < L; ++l)
    if (P.small_ptr[l + 1] > P.small_ptr[l]) fail("small front in a DAG level");
  const FactorDagIndex X{ns, P.fdag_nsteps};
  auto meta = [&](int32_t s) { return &P.fdag_meta[static_cast<size_t>(s) * kFMetaStride]; };
  auto kind_of = [](const FTask& f) { return f.sk >> kFKindShift; };
  auto sn_of = [](const FTask& f) { return f.sk & ((1 << kFKindShift) - 1); };
  auto m_of = [&](int32_t s) { return S.first[s + 1] - S.first[s] + S.nrow[s]; };
  auto nc_of = [&](int32_t s) { return S.first[s + 1] - S.first[s]; };
  auto steps_of = [](int32_t code, int32_t& kf, int32_t& kl) {
    kl = code & ((1 << kUpdStepsShift) - 1);
    kf = kl - ((code >> kUpdStepsShift) & ((1 << (kUpdPanelBit - kUpdStepsShift)) - 1)) + 1;
  };
  // coverage
  {
    std::map<std::tuple<int, int, int, int, int>, int> want, got;
    for (int32_t t = P.asm_ptr[split]; t < P.asm_ptr[L]; ++t) want[{kFAsm, t, 0, 0, 0}]++;
    const int32_t g0 = P.level_step_ptr[split], g1 = P.level_step_ptr[L];
    for (int32_t t = P.pan_ptr[g0]; t < P.pan_ptr[g1]; ++t) {
      const Task4 tk = P.pan_task[t];
      const int32_t k0 = tk.d * kPanelNB;
      if (tk.b == k0) want[{kFDiag, tk.a, tk.d, 0, 0}]++;
      else if (tk.b != m_of(tk.a) + k0) want[{kFPan, tk.a, tk.b, tk.c, tk.d}]++;
    }
    for (int32_t t = P.upd_ptr[g0]; t < P.upd_ptr[g1]; ++t) {
      const Task4 tk = P.upd_task[t];
      want[{kFUpd, tk.a, tk.b, tk.c, tk.d}]++;
    }
    for (const FTask& f : P.fdag) {
      const int k = kind_of(f);
      if (k == kFAsm) got[{kFAsm, f.x, 0, 0, 0}]++;
      else if (k == kFDiag) got[{kFDiag, sn_of(f), f.x, 0, 0}]++;
      else got[{k, sn_of(f), f.x, f.y, f.z}]++;
    }
    if (want != got) fail("DAG tasks differ from the levels' tasks");
  }
  // predecessors per counter, for the position and count checks
  const int64_t nt = static_cast<int64_t>(P.fdag.size());
  std::vector<int32_t> n_asm(ns, 0), n_child(ns, 0), n_upd(ns, 0);
  std::map<int32_t, int32_t> n_pan;  // step counter -> panel tasks incl. the diagonal block
  for (const FTask& f : P.fdag) {
    const int32_t s = sn_of(f);
    switch (kind_of(f)) {
      case kFAsm: n_asm[s]++; break;
      case kFDiag: n_pan[meta(s)[kFMetaStep] + f.x]++; break;
      case kFPan: n_pan[meta(s)[kFMetaStep] + f.z]++; break;
      default: n_upd[s]++;
    }
  }
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t p = meta(s)[kFMetaParent];
    if (p >= 0) {
      if (p != S.sparent[s]) fail("DAG parent");
      n_child[p]++;
    }
    if (n_upd[s] != meta(s)[kFMetaNUpd]) fail("DAG update count of a front");
  }
  // aligned update tiles per column-block counter: (last step, position)
  std::map<int32_t, std::vector<std::pair<int32_t, int64_t>>> col_tiles;
  std::map<int32_t, int64_t> last_upd_pos, diag_pos;
  std::map<int32_t, int64_t> last_pan_pos, last_asm_pos;
  for (int64_t q = 0; q < nt; ++q) {
    const FTask& f = P.fdag[q];
    const int32_t s = sn_of(f);
    if (kind_of(f) == kFUpd) {
      int32_t kf, kl;
      steps_of(f.z, kf, kl);
      if (f.y % kTile == 0) col_tiles[X.col_done(meta(s)[kFMetaCb] + f.y / kTile)].push_back({kl, q});
      last_upd_pos[s] = q;
    } else if (kind_of(f) == kFDiag) {
      diag_pos[meta(s)[kFMetaStep] + f.x] = q;
      last_pan_pos[meta(s)[kFMetaStep] + f.x] = std::max(last_pan_pos[meta(s)[kFMetaStep] + f.x], q);
    } else if (kind_of(f) == kFPan) {
      last_pan_pos[meta(s)[kFMetaStep] + f.z] = std::max(last_pan_pos[meta(s)[kFMetaStep] + f.z], q);
    } else {
      last_asm_pos[s] = q;
    }
  }
  // tiles of counter c with last step < k, and whether they all precede q
  auto col_before = [&](int32_t c, int32_t k, int64_t q) {
    int32_t cnt = 0;
    for (const auto& e : col_tiles[c])
      if (e.first < k) {
        if (e.second >= q) fail("an earlier update tile comes later in the list");
        ++cnt;
      }
    return cnt;
  };
  for (int64_t q = 0; q < nt; ++q) {
    const FTask& f = P.fdag[q];
    const int32_t s = sn_of(f);
    const int32_t cb = meta(s)[kFMetaCb], st = meta(s)[kFMetaStep];
    switch (kind_of(f)) {
      case kFAsm:
        if (f.wa != X.child_done(s) || f.ta != n_child[s] || f.wb >= 0 || f.wc >= 0) fail("asm waits");
        for (int32_t t = P.child_ptr[s]; t < P.child_ptr[s + 1]; ++t) {
          const int32_t c = P.child_list[t];
          if (meta(c)[kFMetaParent] == s && last_upd_pos[c] >= q) fail("assembly before a child's update tile");
        }
        break;
      case kFDiag:
      case kFPan: {
        const int32_t k = kind_of(f) == kFDiag ? f.x : f.z;
        if (k == 0) {
          if (f.wa != X.asm_done(s) || f.ta != n_asm[s]) fail("step-0 panel waits");
          if (n_asm[s] > 0 && last_asm_pos[s] >= q) fail("panel before its assembly");
        } else {
          if (f.wa != X.col_done(cb + k) || f.ta != col_before(f.wa, k, q)) fail("panel waits on its column block");
          for (const auto& e : col_tiles[f.wa])
            if (e.first >= k) fail("a tile of a factored column block is updated after its panel step");
        }
        if (kind_of(f) == kFPan) {
          if (f.wb != X.diag_done(st + k) || f.tb != 1 || diag_pos.count(st + k) == 0 || diag_pos[st + k] >= q)
            fail("panel rows before their diagonal block");
        } else if (f.wb >= 0) {
          fail("diagonal block waits");
        }
        if (f.wc >= 0) fail("panel third wait");
        break;
      }
      default: {
        int32_t kf, kl;
        steps_of(f.z, kf, kl);
        if (f.wa != X.pan_done(st + kl) || f.ta != n_pan[st + kl] || last_pan_pos[st + kl] >= q)
          fail("update tile before its panel step");
        const int32_t m = m_of(s), nc = nc_of(s);
        const int32_t b0 = f.y / kTile, b1 = (std::min(m, f.y + kTile) - 1) / kTile;
        if (f.wb != X.col_done(cb + b0) || f.tb != col_before(f.wb, kf, q)) fail("update tile waits on its block");
        if (f.y % kTile != 0) {
          const int32_t nk = (nc + kPanelNB - 1) / kPanelNB;
          if (f.y < nc || kl != nk - 1) fail("unaligned update tile before the last step or on panel columns");
          if (b1 > b0 && (f.wc != X.col_done(cb + b1) || f.tc != col_before(f.wc, kf, q)))
            fail("unaligned tile waits on its second block");
          if (b1 == b0 && f.wc >= 0) fail("unaligned tile third wait");
        } else if (f.wc >= 0) {
          fail("aligned tile third wait");
        }
      }
    }
  }
  // update-matrix blocks live in the launch: the DAG fronts' and their children's
  if (!P.dag_pool) fail("the update pool is not laid out for the DAG");
  {
    std::vector<std::pair<int64_t, int64_t>> iv;
    std::vector<char> seen(ns, 0);
    for (int32_t l = split; l < L; ++l)
      for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) {
        const int32_t s = P.big_list[q];
        for (int32_t t = P.child_ptr[s]; t <= P.child_ptr[s + 1]; ++t) {
          const int32_t c = t < P.child_ptr[s + 1] ? P.child_list[t] : s;
          if (seen[c] || S.nrow[c] == 0) continue;
          seen[c] = 1;
          const int64_t nr = S.nrow[c];
          iv.push_back({P.upd_off[c], P.upd_off[c] + nr * nr});
        }
      }
    std::sort(iv.begin(), iv.end());
    for (size_t i = 1; i < iv.size(); ++i)
      if (iv[i].first < iv[i - 1].second) fail("update-matrix blocks of the DAG levels overlap");
  }
  // a run by a few workgroups in list order with random finishing order
  {
    std::vector<int32_t> cnt(static_cast<size_t>(X.col_done(P.fdag_ncb)), 0);
    std::vector<int32_t> ucnt(ns, 0);
    uint64_t rng = 0x9e3779b97f4a7c15ull ^ static_cast<uint64_t>(nt);
    auto rnd = [&]() {
      rng ^= rng << 13;
      rng ^= rng >> 7;
      rng ^= rng << 17;
      return rng;
    };
    const int G = 5;
    std::vector<int64_t> held;
    int64_t next = 0, done = 0;
    auto ready = [&](const FTask& f) {
      return cnt[f.wa] >= f.ta && (f.wb < 0 || cnt[f.wb] >= f.tb) && (f.wc < 0 || cnt[f.wc] >= f.tc);
    };
    while (done < nt) {
      while (static_cast<int>(held.size()) < G && next < nt) held.push_back(next++);
      std::vector<size_t> can;
      for (size_t i = 0; i < held.size(); ++i)
        if (ready(P.fdag[held[i]])) can.push_back(i);
      if (can.empty()) fail("the list deadlocks with a few workgroups");
      const size_t pick = can[rnd() % can.size()];
      const FTask f = P.fdag[held[pick]];
      held.erase(held.begin() + static_cast<int64_t>(pick));
      ++done;
      const int32_t s = sn_of(f);
      switch (kind_of(f)) {
        case kFAsm: cnt[X.asm_done(s)]++; break;
        case kFDiag:
          cnt[X.diag_done(meta(s)[kFMetaStep] + f.x)]++;
          cnt[X.pan_done(meta(s)[kFMetaStep] + f.x)]++;
          break;
        case kFPan: cnt[X.pan_done(meta(s)[kFMetaStep] + f.z)]++; break;
        default:
          if (f.y % kTile == 0) cnt[X.col_done(meta(s)[kFMetaCb] + f.y / kTile)]++;
          if (++ucnt[s] == meta(s)[kFMetaNUpd] && meta(s)[kFMetaParent] >= 0)
            cnt[X.child_done(meta(s)[kFMetaParent])]++;
      }
    }
  }
}

// true if check() finds a violation in the plan
template <class F>
bool caught(F&& check, const FactorPlan& P) {
  try {
    check(P);
  } catch (const Violation&) {
    return true;
  }
  return false;
}

int run() {
  struct Case {
    int dim, k, m, a_mode, cnt;
    int ordering;
  };
  const Case cases[] = {{2, 24, 144, 0, 3, 1}, {3, 8, 128, 1, 4, 2}, {3, 10, 250, 1, 4, 1}, {2, 40, 400, 0, 3, 2},
                        {3, 14, 686, 1, 4, 2}, {3, 22, 2662, 1, 4, 2}};
  const FactorTuning ft[] = {{4, 8}, {1, 1}, {2, 4}, {3, 16}, {8, 2}, {64, 64}};
  const SolveTuning st[] = {{1024}, {0}, {64}, {300}, {1 << 30}};
  for (int amalg : {1, kDefaultAmalgamation})
  for (const Case& cs : cases) {
    KktSpec sp;
    sp.dim = cs.dim;
    sp.k = cs.k;
    sp.m = cs.m;
    sp.a_mode = cs.a_mode;
    sp.cnt_lo = sp.cnt_hi = cs.cnt;
    const KktProblem kp = generate_kkt(sp);
    Analysis an = analyze_pattern(kp.K.n, kp.K.p.data(), kp.K.i.data(), cs.ordering, amalg);
    FactorPlan& P = an.plan;
    int32_t max_steps = 0;
    for (int32_t s : P.big_list) max_steps = std::max(max_steps, (P.sn.first[s + 1] - P.sn.first[s] + 63) / 64);
    build_update_pool(P, true);
    for (const FactorTuning& f : ft) {
      build_factor_steps(P, f);
      check_factor(P);
      check_factor_dag(P);
    }
    for (const SolveTuning& s : st) {
      build_solve_schedule(P, s);
      check_solve(P);
    }
    std::printf("n %d, amalgamation %d: %d supernodes, %lld big fronts, up to %d panel steps per front, %zu DAG tasks from level %d of %d: ok\n",
                kp.K.n, amalg, P.sn.ns, static_cast<long long>(P.nbig), max_steps, P.fdag.size(), P.fdag_level,
                P.sn.nlevels);
  }
  // planted mistakes, in copies of a plan with many steps per front
  KktSpec sp;
  sp.dim = 2;
  sp.k = 40;
  sp.m = 400;
  const KktProblem kp = generate_kkt(sp);
  Analysis an = analyze_pattern(kp.K.n, kp.K.p.data(), kp.K.i.data(), 2, 1);
  build_update_pool(an.plan, true);
  build_factor_steps(an.plan, FactorTuning{4, 8});
  build_solve_schedule(an.plan, SolveTuning{1024});
  const FactorPlan& base = an.plan;
  if (base.upd_task.size() < 10 || base.fm_task.empty()) fail("planted-mistake plan too small");
  {
    FactorPlan bad = base;
    bad.upd_task.erase(bad.upd_task.begin() + static_cast<int64_t>(bad.upd_task.size() / 2));
    for (size_t i = 0; i < bad.upd_ptr.size(); ++i)
      if (bad.upd_ptr[i] > static_cast<int32_t>(base.upd_task.size() / 2)) bad.upd_ptr[i]--;
    if (!caught(check_factor, bad)) fail("a dropped update task was not caught");
  }
  {
    FactorPlan bad = base;
    const size_t at = bad.upd_task.size() / 3;
    bad.upd_task.insert(bad.upd_task.begin() + static_cast<int64_t>(at), bad.upd_task[at]);
    for (size_t i = 0; i < bad.upd_ptr.size(); ++i)
      if (bad.upd_ptr[i] > static_cast<int32_t>(at)) bad.upd_ptr[i]++;
    if (!caught(check_factor, bad)) fail("a doubled update task was not caught");
  }
  {
    FactorPlan bad = base;
    bad.bm_task.erase(bad.bm_task.begin(), bad.bm_task.begin() + 3);
    if (!caught(check_solve, bad)) fail("a dropped narrow-front task was not caught");
  }
  if (base.fdag.size() < 10) fail("planted-mistake plan has no DAG levels");
  {
    // the level-lifetime pool reuses blocks inside the DAG levels
    FactorPlan lev = base;
    build_update_pool(lev, false);
    if (!lev.fdag.empty()) fail("a level-lifetime pool kept the DAG list");
    FactorPlan bad = base;
    bad.upd_off = lev.upd_off;
    if (!caught(check_factor_dag, bad)) fail("a level-lifetime update pool was not caught");
  }
  {
    // a diagonal block moved to the end of the list
    FactorPlan bad = base;
    auto it = std::find_if(bad.fdag.begin(), bad.fdag.end(), [](const FTask& f) { return (f.sk >> kFKindShift) == kFDiag; });
    const FTask f = *it;
    bad.fdag.erase(it);
    bad.fdag.push_back(f);
    if (!caught(check_factor_dag, bad)) fail("a diagonal block after its dependents was not caught");
  }
  {
    // an update tile that waits for one earlier tile too few
    FactorPlan bad = base;
    for (FTask& f : bad.fdag)
      if ((f.sk >> kFKindShift) == kFUpd && f.tb > 0) {
        f.tb--;
        break;
      }
    if (!caught(check_factor_dag, bad)) fail("a short update-tile wait was not caught");
  }
  {
    // a dropped panel task
    FactorPlan bad = base;
    auto it = std::find_if(bad.fdag.begin(), bad.fdag.end(), [](const FTask& f) { return (f.sk >> kFKindShift) == kFPan; });
    bad.fdag.erase(it);
    if (!caught(check_factor_dag, bad)) fail("a dropped panel task was not caught");
  }
  std::printf("planted mistakes caught\n");
  std::printf("plan_check ok\n");
  return 0;
}

}  // namespace

int main() {
  try {
    return run();
  } catch (const Violation& v) {
    std::printf("FAIL: %s\n", v.what.c_str());
    return 1;
  }
}
