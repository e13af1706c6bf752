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
// Mistakes planted in a copy of a plan (an update task dropped or doubled, a
// narrow-front task dropped) must be caught. Prints one line per case and
// "plan_check ok" at the end; exits 1 on the first violation.
#include <cstdio>
#include <cstdlib>
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
  for (const Case& cs : cases) {
    KktSpec sp;
    sp.dim = cs.dim;
    sp.k = cs.k;
    sp.m = cs.m;
    sp.a_mode = cs.a_mode;
    sp.cnt_lo = sp.cnt_hi = cs.cnt;
    const KktProblem kp = generate_kkt(sp);
    Analysis an = analyze_pattern(kp.K.n, kp.K.p.data(), kp.K.i.data(), cs.ordering, true);
    FactorPlan& P = an.plan;
    int32_t max_steps = 0;
    for (int32_t s : P.big_list) max_steps = std::max(max_steps, (P.sn.first[s + 1] - P.sn.first[s] + 63) / 64);
    for (const FactorTuning& f : ft) {
      build_factor_steps(P, f);
      check_factor(P);
    }
    for (const SolveTuning& s : st) {
      build_solve_schedule(P, s);
      check_solve(P);
    }
    std::printf("n %d: %d supernodes, %lld big fronts, up to %d panel steps per front: ok\n", kp.K.n, P.sn.ns,
                static_cast<long long>(P.nbig), max_steps);
  }
  // planted mistakes, in copies of a plan with many steps per front
  KktSpec sp;
  sp.dim = 2;
  sp.k = 40;
  sp.m = 400;
  const KktProblem kp = generate_kkt(sp);
  Analysis an = analyze_pattern(kp.K.n, kp.K.p.data(), kp.K.i.data(), 2, true);
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
