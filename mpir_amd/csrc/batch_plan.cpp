#include "batch_plan.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "host/sparse.hpp"

namespace mpir {

namespace {
// f(i) for i in [0, n) on up to `threads` host threads
template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
  const int nt = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(threads, n)));
  if (nt == 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (int64_t i; (i = next.fetch_add(1)) < n;) f(i);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
}
int join_threads() { return static_cast<int>(std::max(1u, std::min(8u, std::thread::hardware_concurrency()))); }
}  // namespace

uint64_t pattern_key(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, int relax) {
  // FNV-1a over the dimension, the options and the index arrays
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* data, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) {
      h ^= p[i];
      h *= 1099511628211ull;
    }
  };
  const int32_t opts[3] = {n, ordering, relax};
  mix(opts, sizeof opts);
  mix(Ap, sizeof(int32_t) * (static_cast<size_t>(n) + 1));
  mix(Ai, sizeof(int32_t) * static_cast<size_t>(Ap[n]));
  return h;
}

PatternEntry analyze_entry(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, int relax) {
  if (n < 1) throw std::invalid_argument("matrix dimension must be positive");
  if (Ap[0] != 0) throw std::invalid_argument("column pointers must start at 0");
  for (int32_t j = 0; j < n; ++j) {
    if (Ap[j + 1] <= Ap[j]) throw std::invalid_argument("empty column in the upper-triangular input");
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q) {
      if (Ai[q] < 0 || Ai[q] > j) throw std::invalid_argument("entry below the diagonal or out of range");
      if (q > Ap[j] && Ai[q] <= Ai[q - 1]) throw std::invalid_argument("row indices must be strictly ascending");
    }
  }
  PatternEntry e;
  e.key = pattern_key(n, Ap, Ai, ordering, relax);
  e.ordering = ordering;
  e.relax = relax;
  e.n = n;
  e.nnz = Ap[n];
  e.Ap.assign(Ap, Ap + n + 1);
  e.Ai.assign(Ai, Ai + e.nnz);
  e.an = analyze_pattern(n, Ap, Ai, ordering, relax);
  CsrFull F = csr_from_upper(e.an.C);
  e.csr_p = std::move(F.p);
  e.csr_j = std::move(F.j);
  e.csr_src.resize(F.src.size());
  for (size_t q = 0; q < F.src.size(); ++q) e.csr_src[q] = e.an.val_map[F.src[q]];
  e.ecol.resize(e.nnz);
  e.diag_pos.assign(n, -1);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q) {
      e.ecol[q] = j;
      if (Ai[q] == j) e.diag_pos[j] = q;
    }
  e.est_bytes = pattern_footprint(e, 4, 1, 10);
  return e;
}

int64_t pattern_footprint(const PatternEntry& e, int factor_bytes, int nrhs, int max_gmres) {
  const FactorPlan& P = e.an.plan;
  const int64_t n = e.n, nnz = e.nnz, ncsr = static_cast<int64_t>(e.csr_j.size());
  int64_t b = 0;
  b += (P.panel_entries + P.upd_pool + 2 * n) * factor_bytes;            // panels, update pool, D, D^{-1}
  b += 2 * nnz * 8 + ncsr * (8 + 4 + 4) + (n + 1) * 4;                    // values, CSR
  b += static_cast<int64_t>(P.sn.rows.size()) * (8 + 4 + 4);              // update vectors, rows, relidx
  b += (P.nypart + P.nupart + P.nrpart + P.nxpart) * 8;                   // solve partial sums
  b += static_cast<int64_t>(nrhs) * n * 8 * (max_gmres + 1 + 6);          // basis and work vectors
  b += static_cast<int64_t>(P.pan_task.size() + P.upd_task.size() + P.asm_task.size()) * 16 +
       static_cast<int64_t>(P.dag_fwd.size() + P.dag_bwd.size() + P.gent.size()) * 16 +  // schedules
       static_cast<int64_t>(P.fdag.size()) * static_cast<int64_t>(sizeof(FTask)) +
       (3 * static_cast<int64_t>(P.sn.ns) + 2 * P.fdag_nsteps + P.fdag_ncb) * 4;  // factor task list, counters
  return b;
}

std::shared_ptr<const PatternEntry> PatternCache::find_locked(uint64_t key, int32_t n, const int32_t* Ap,
                                                              const int32_t* Ai, int ordering, int relax) const {
  auto range = map_.equal_range(key);
  for (auto it = range.first; it != range.second; ++it) {
    const PatternEntry& e = *it->second;
    if (e.n != n || e.ordering != ordering || e.relax != relax || e.nnz != Ap[n]) continue;
    if (std::memcmp(e.Ap.data(), Ap, sizeof(int32_t) * (n + 1)) != 0) continue;
    if (std::memcmp(e.Ai.data(), Ai, sizeof(int32_t) * e.nnz) != 0) continue;
    return it->second;
  }
  return nullptr;
}

std::shared_ptr<const PatternEntry> PatternCache::get(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering,
                                                      int relax) {
  return get_many({n}, {Ap}, {Ai}, ordering, relax, 1)[0];
}

std::vector<std::shared_ptr<const PatternEntry>> PatternCache::get_many(const std::vector<int32_t>& n,
                                                                        const std::vector<const int32_t*>& Ap,
                                                                        const std::vector<const int32_t*>& Ai,
                                                                        int ordering, int relax, int threads) {
  const size_t ns = n.size();
  std::vector<std::shared_ptr<const PatternEntry>> out(ns);
  std::vector<uint64_t> key(ns);
  // first system of every pattern missing from the cache, and the systems
  // that share a missing pattern with an earlier one
  std::vector<size_t> todo;
  std::vector<size_t> same_as(ns, SIZE_MAX);
  {
    for (size_t i = 0; i < ns; ++i) {
      if (n[i] < 1 || !Ap[i] || !Ai[i]) throw std::invalid_argument("bad system pattern");
      key[i] = pattern_key(n[i], Ap[i], Ai[i], ordering, relax);
    }
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t i = 0; i < ns; ++i) {
      out[i] = find_locked(key[i], n[i], Ap[i], Ai[i], ordering, relax);
      if (out[i]) {
        ++hits_;
        continue;
      }
      for (size_t t : todo)
        if (key[t] == key[i] && n[t] == n[i] && Ap[t][n[i]] == Ap[i][n[i]] &&
            std::memcmp(Ap[t], Ap[i], sizeof(int32_t) * (n[i] + 1)) == 0 &&
            std::memcmp(Ai[t], Ai[i], sizeof(int32_t) * Ap[i][n[i]]) == 0) {
          same_as[i] = t;
          break;
        }
      if (same_as[i] == SIZE_MAX) todo.push_back(i);
    }
  }
  // analyse the missing patterns, largest first, on a pool of threads
  std::sort(todo.begin(), todo.end(), [&](size_t a, size_t b) { return Ap[a][n[a]] > Ap[b][n[b]]; });
  std::vector<std::shared_ptr<const PatternEntry>> made(ns);
  std::vector<std::string> errors(ns);
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t t; (t = next.fetch_add(1)) < todo.size();) {
      const size_t i = todo[t];
      try {
        made[i] = std::make_shared<const PatternEntry>(analyze_entry(n[i], Ap[i], Ai[i], ordering, relax));
      } catch (const std::exception& ex) {
        errors[i] = ex.what();
      }
    }
  };
  const int nt = std::max(1, std::min<int>(threads, static_cast<int>(todo.size())));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  for (size_t i : todo)
    if (!made[i]) throw std::invalid_argument("system " + std::to_string(i) + ": " + errors[i]);
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t i : todo) {
      // another thread may have added the same pattern meanwhile
      auto have = find_locked(key[i], n[i], Ap[i], Ai[i], ordering, relax);
      if (have) {
        made[i] = have;
      } else {
        map_.emplace(key[i], made[i]);
        ++misses_;
      }
      out[i] = made[i];
    }
    for (size_t i = 0; i < ns; ++i)
      if (same_as[i] != SIZE_MAX) {
        out[i] = out[same_as[i]];
        ++hits_;
      }
  }
  return out;
}

size_t PatternCache::size() const {
  std::lock_guard<std::mutex> lk(mu_);
  return map_.size();
}

void merge_plans(const std::vector<const FactorPlan*>& P, const std::vector<int32_t>& col_base, FactorPlan& M) {
  const size_t S = P.size();
  // these lists start empty (the vectors keep their memory)
  for (auto* v : {&M.child_ptr, &M.child_list, &M.small_ptr, &M.small_list, &M.big_ptr, &M.big_list, &M.asm_ptr,
                  &M.level_step_ptr, &M.fx_ptr, &M.fu_ptr, &M.br_ptr, &M.bx_ptr, &M.fx_task, &M.fu_task, &M.br_task,
                  &M.bx_task, &M.fm_ptr, &M.bm_ptr, &M.fm_task, &M.bm_task, &M.inv_task, &M.cb_base, &M.dag_meta, &M.sn.level_ptr, &M.sn.level_sn})
    v->clear();
  for (auto* v : {&M.upd_off, &M.ypart_off, &M.upart_off, &M.rpart_off, &M.xpart_off}) v->clear();
  for (auto* v : {&M.dag_fwd, &M.dag_bwd}) v->clear();
  // asm_task, pan_task, upd_task, gent, frow_off, grow_ptr (and val_dst, in
  // combine_patterns) are resized and then written entry by entry: not
  // clearing them first means that only growth is zero-filled
  Supernodes& T = M.sn;
  // bases of every system in the joined numbering
  struct Base {
    int32_t sn = 0, col = 0, frow = 0, cb = 0;
    int64_t rows = 0, panel = 0, upd = 0, gent = 0, yp = 0, up = 0, rp = 0, xp = 0;
  };
  std::vector<Base> b(S + 1);
  int32_t nlev = 0;
  for (size_t s = 0; s < S; ++s) {
    const FactorPlan& p = *P[s];
    const Supernodes& q = p.sn;
    Base& n = b[s + 1];
    n.sn = b[s].sn + q.ns;
    n.col = col_base[s + 1];
    n.rows = b[s].rows + static_cast<int64_t>(q.rows.size());
    n.panel = b[s].panel + p.panel_entries;
    n.upd = b[s].upd + p.upd_pool;
    n.frow = b[s].frow + p.frow_off[q.ns];
    n.gent = b[s].gent + static_cast<int64_t>(p.gent.size()) - 2;
    n.cb = b[s].cb + p.ncb_total;
    n.yp = b[s].yp + p.nypart;
    n.up = b[s].up + p.nupart;
    n.rp = b[s].rp + p.nrpart;
    n.xp = b[s].xp + p.nxpart;
    nlev = std::max(nlev, q.nlevels);
    if (q.first[0] != 0 || q.n != col_base[s + 1] - col_base[s]) throw std::logic_error("plan does not match its system");
  }
  const Base& E = b[S];
  if (E.rows >= INT32_MAX || E.gent >= INT32_MAX || E.up >= INT32_MAX)
    throw std::length_error("batch exceeds the 32-bit solve gather indices");
  // supernodes
  T.n = E.col;
  T.ns = E.sn;
  T.nlevels = nlev;
  T.first.resize(static_cast<size_t>(T.ns) + 1);
  T.nrow.resize(T.ns);
  T.rows_off.resize(static_cast<size_t>(T.ns) + 1);
  T.sparent.resize(T.ns);
  T.level.resize(T.ns);
  T.panel_off.resize(static_cast<size_t>(T.ns) + 1);
  T.rows.resize(E.rows);
  T.relidx.resize(E.rows);
  T.col2sn.resize(T.n);
  T.nnz_l = 0;
  T.flops = 0.0;
  T.max_front = 0;
  for (size_t s = 0; s < S; ++s) {
    const Supernodes& q = P[s]->sn;
    T.nnz_l += q.nnz_l;
    T.flops += q.flops;
    T.max_front = std::max(T.max_front, q.max_front);
  }
  parallel_for(static_cast<int64_t>(S), join_threads(), [&](int64_t s) {
    const Supernodes& q = P[s]->sn;
    const Base& o = b[s];
    for (int32_t j = 0; j < q.ns; ++j) {
      const int64_t d = o.sn + j;
      T.first[d] = o.col + q.first[j];
      T.nrow[d] = q.nrow[j];
      T.rows_off[d] = o.rows + q.rows_off[j];
      T.sparent[d] = q.sparent[j] >= 0 ? o.sn + q.sparent[j] : -1;
      T.level[d] = q.level[j];
      T.panel_off[d] = o.panel + q.panel_off[j];
    }
    for (size_t r = 0; r < q.rows.size(); ++r) {
      T.rows[o.rows + r] = o.col + q.rows[r];
      T.relidx[o.rows + r] = q.relidx[r];
    }
    for (int32_t c = 0; c < q.n; ++c) T.col2sn[o.col + c] = o.sn + q.col2sn[c];
  });
  T.first[T.ns] = T.n;
  T.rows_off[T.ns] = E.rows;
  T.panel_off[T.ns] = E.panel;
  T.level_ptr.assign(static_cast<size_t>(nlev) + 1, 0);
  for (int32_t l = 0; l < nlev; ++l) {
    for (size_t s = 0; s < S; ++s) {
      const Supernodes& q = P[s]->sn;
      if (l >= q.nlevels) continue;
      for (int32_t t = q.level_ptr[l]; t < q.level_ptr[l + 1]; ++t) T.level_sn.push_back(b[s].sn + q.level_sn[t]);
    }
    T.level_ptr[l + 1] = static_cast<int32_t>(T.level_sn.size());
  }
  // children, update pool (each system keeps its own region)
  M.child_ptr.assign(1, 0);
  for (size_t s = 0; s < S; ++s) {
    const FactorPlan& p = *P[s];
    const int32_t c0 = M.child_ptr.back();
    for (int32_t j = 0; j < p.sn.ns; ++j) M.child_ptr.push_back(c0 + p.child_ptr[j + 1]);
    for (int32_t c : p.child_list) M.child_list.push_back(b[s].sn + c);
    for (int64_t u : p.upd_off) M.upd_off.push_back(u >= 0 ? b[s].upd + u : -1);
    M.cb_base.insert(M.cb_base.end(), p.cb_base.begin(), p.cb_base.end());
    for (size_t j = M.cb_base.size() - p.cb_base.size(); j < M.cb_base.size(); ++j) M.cb_base[j] += b[s].cb;
    for (int64_t v : p.ypart_off) M.ypart_off.push_back(b[s].yp + v);
    for (int64_t v : p.upart_off) M.upart_off.push_back(b[s].up + v);
    for (int64_t v : p.rpart_off) M.rpart_off.push_back(b[s].rp + v);
    for (int64_t v : p.xpart_off) M.xpart_off.push_back(b[s].xp + v);
  }
  M.upd_pool = E.upd;
  M.panel_entries = E.panel;
  M.ncb_total = E.cb;
  M.nypart = E.yp;
  M.nupart = E.up;
  M.nrpart = E.rp;
  M.nxpart = E.xp;
  // per-level lists
  auto lvl_list = [&](std::vector<int32_t>& ptr, std::vector<int32_t>& list,
                      const std::vector<int32_t> FactorPlan::*pp, const std::vector<int32_t> FactorPlan::*pl) {
    ptr.assign(static_cast<size_t>(nlev) + 1, 0);
    for (int32_t l = 0; l < nlev; ++l) {
      for (size_t s = 0; s < S; ++s) {
        const FactorPlan& p = *P[s];
        if (l >= p.sn.nlevels) continue;
        const std::vector<int32_t>& sp = p.*pp;
        const std::vector<int32_t>& sl = p.*pl;
        for (int32_t t = sp[l]; t < sp[l + 1]; ++t) list.push_back(b[s].sn + sl[t]);
      }
      ptr[l + 1] = static_cast<int32_t>(list.size());
    }
  };
  lvl_list(M.small_ptr, M.small_list, &FactorPlan::small_ptr, &FactorPlan::small_list);
  lvl_list(M.big_ptr, M.big_list, &FactorPlan::big_ptr, &FactorPlan::big_list);
  auto lvl_triples = [&](std::vector<int32_t>& ptr, std::vector<int32_t>& list,
                         const std::vector<int32_t> FactorPlan::*pp, const std::vector<int32_t> FactorPlan::*pl) {
    ptr.assign(static_cast<size_t>(nlev) + 1, 0);
    for (int32_t l = 0; l < nlev; ++l) {
      for (size_t s = 0; s < S; ++s) {
        const FactorPlan& p = *P[s];
        if (l >= p.sn.nlevels) continue;
        const std::vector<int32_t>& sp = p.*pp;
        const std::vector<int32_t>& sl = p.*pl;
        for (int32_t t = sp[l]; t < sp[l + 1]; ++t)
          list.insert(list.end(), {b[s].sn + sl[3 * t], sl[3 * t + 1], sl[3 * t + 2]});
      }
      ptr[l + 1] = static_cast<int32_t>(list.size() / 3);
    }
  };
  lvl_triples(M.fx_ptr, M.fx_task, &FactorPlan::fx_ptr, &FactorPlan::fx_task);
  lvl_triples(M.fu_ptr, M.fu_task, &FactorPlan::fu_ptr, &FactorPlan::fu_task);
  lvl_triples(M.br_ptr, M.br_task, &FactorPlan::br_ptr, &FactorPlan::br_task);
  lvl_triples(M.bx_ptr, M.bx_task, &FactorPlan::bx_ptr, &FactorPlan::bx_task);
  lvl_triples(M.fm_ptr, M.fm_task, &FactorPlan::fm_ptr, &FactorPlan::fm_task);
  lvl_triples(M.bm_ptr, M.bm_task, &FactorPlan::bm_ptr, &FactorPlan::bm_task);
  for (size_t s = 0; s < S; ++s) {
    const std::vector<int32_t>& v = P[s]->inv_task;
    for (size_t t = 0; t < v.size(); t += 3) M.inv_task.insert(M.inv_task.end(), {b[s].sn + v[t], v[t + 1], v[t + 2]});
  }
  // the systems' plans come from the pattern cache, all with the default
  // tunings; the handle re-plans the joined forest for its own
  M.tuning = S > 0 ? P[0]->tuning : SolveTuning{};
  M.ftuning = S > 0 ? P[0]->ftuning : FactorTuning{};
  // assembly tasks with their child ranges, and the panel and update tasks
  // per (level, step): the real rows of every system, then their identity
  // rows. Positions come from counts per (level, system) and (level, step,
  // pass, system); the copies then run in parallel per system.
  {
    const size_t LS = static_cast<size_t>(nlev) * S;
    std::vector<int64_t> at(LS + 1, 0), ct(LS + 1, 0);
    for (int32_t l = 0; l < nlev; ++l)
      for (size_t s = 0; s < S; ++s) {
        const FactorPlan& p = *P[s];
        const size_t i = static_cast<size_t>(l) * S + s;
        int64_t na = 0, nc = 0;
        if (l < p.sn.nlevels) {
          na = p.asm_ptr[l + 1] - p.asm_ptr[l];
          nc = p.asm_child_ptr[p.asm_ptr[l + 1]] - p.asm_child_ptr[p.asm_ptr[l]];
        }
        at[i + 1] = at[i] + na;
        ct[i + 1] = ct[i] + nc;
      }
    if (at[LS] >= INT32_MAX || 5 * ct[LS] >= INT32_MAX) throw std::length_error("batch exceeds the 32-bit task indices");
    M.asm_task.resize(static_cast<size_t>(at[LS]));
    M.asm_child.resize(static_cast<size_t>(5 * ct[LS]));
    M.asm_child_ptr.resize(static_cast<size_t>(at[LS]) + 1);
    M.asm_child_ptr[at[LS]] = static_cast<int32_t>(ct[LS]);
    M.asm_ptr.assign(static_cast<size_t>(nlev) + 1, 0);
    for (int32_t l = 0; l < nlev; ++l) M.asm_ptr[l + 1] = static_cast<int32_t>(at[static_cast<size_t>(l + 1) * S]);
    // steps: per level the most of any system; (level, step) pairs in order
    M.level_step_ptr.assign(static_cast<size_t>(nlev) + 1, 0);
    for (int32_t l = 0; l < nlev; ++l) {
      int32_t steps = 0;
      for (size_t s = 0; s < S; ++s) {
        const FactorPlan& p = *P[s];
        if (l < p.sn.nlevels) steps = std::max(steps, p.level_step_ptr[l + 1] - p.level_step_ptr[l]);
      }
      M.level_step_ptr[l + 1] = M.level_step_ptr[l] + steps;
    }
    const int32_t G = M.level_step_ptr[nlev];
    const size_t GS = static_cast<size_t>(G) * 2 * S;
    std::vector<int64_t> pt(GS + 1, 0), ut(GS + 1, 0);
    // source range of (level l, step k, pass) in system s, or empty
    auto src = [&](const FactorPlan& p, int32_t l, int32_t k, int pass, int32_t& pa, int32_t& pe, int32_t& ua,
                   int32_t& ue) {
      pa = pe = ua = ue = 0;
      if (l >= p.sn.nlevels || k >= p.level_step_ptr[l + 1] - p.level_step_ptr[l]) return;
      const int32_t g = p.level_step_ptr[l] + k;
      pa = pass == 0 ? p.pan_ptr[g] : p.pan_aug_ptr[g];
      pe = pass == 0 ? p.pan_aug_ptr[g] : p.pan_ptr[g + 1];
      ua = pass == 0 ? p.upd_ptr[g] : p.upd_aug_ptr[g];
      ue = pass == 0 ? p.upd_aug_ptr[g] : p.upd_ptr[g + 1];
    };
    for (int32_t l = 0; l < nlev; ++l)
      for (int32_t g = M.level_step_ptr[l]; g < M.level_step_ptr[l + 1]; ++g)
        for (int pass = 0; pass < 2; ++pass)
          for (size_t s = 0; s < S; ++s) {
            int32_t pa, pe, ua, ue;
            src(*P[s], l, g - M.level_step_ptr[l], pass, pa, pe, ua, ue);
            const size_t i = (static_cast<size_t>(g) * 2 + pass) * S + s;
            pt[i + 1] = pt[i] + (pe - pa);
            ut[i + 1] = ut[i] + (ue - ua);
          }
    if (pt[GS] >= INT32_MAX || ut[GS] >= INT32_MAX) throw std::length_error("batch exceeds the 32-bit task indices");
    M.pan_task.resize(static_cast<size_t>(pt[GS]));
    M.upd_task.resize(static_cast<size_t>(ut[GS]));
    M.pan_ptr.resize(static_cast<size_t>(G) + 1);
    M.upd_ptr.resize(static_cast<size_t>(G) + 1);
    M.pan_aug_ptr.resize(G);
    M.upd_aug_ptr.resize(G);
    for (int32_t g = 0; g <= G; ++g) {
      M.pan_ptr[g] = static_cast<int32_t>(pt[static_cast<size_t>(g) * 2 * S]);
      M.upd_ptr[g] = static_cast<int32_t>(ut[static_cast<size_t>(g) * 2 * S]);
    }
    for (int32_t g = 0; g < G; ++g) {
      M.pan_aug_ptr[g] = static_cast<int32_t>(pt[(static_cast<size_t>(g) * 2 + 1) * S]);
      M.upd_aug_ptr[g] = static_cast<int32_t>(ut[(static_cast<size_t>(g) * 2 + 1) * S]);
    }
    parallel_for(static_cast<int64_t>(S), join_threads(), [&](int64_t s) {
      const FactorPlan& p = *P[s];
      const int32_t sn = b[s].sn;
      for (int32_t l = 0; l < nlev && l < p.sn.nlevels; ++l) {
        const size_t i = static_cast<size_t>(l) * S + s;
        const int32_t t0 = p.asm_ptr[l], c0 = p.asm_child_ptr[t0];
        for (int32_t t = t0; t < p.asm_ptr[l + 1]; ++t) {
          Task4 tk = p.asm_task[t];
          tk.a += sn;
          M.asm_task[at[i] + (t - t0)] = tk;
          M.asm_child_ptr[at[i] + (t - t0)] = static_cast<int32_t>(ct[i] + (p.asm_child_ptr[t] - c0));
        }
        for (int32_t c = c0; c < p.asm_child_ptr[p.asm_ptr[l + 1]]; ++c) {
          const int32_t* e = &p.asm_child[5 * static_cast<size_t>(c)];
          int32_t* d = &M.asm_child[5 * static_cast<size_t>(ct[i] + (c - c0))];
          d[0] = sn + e[0];
          d[1] = e[1];
          d[2] = e[2];
          d[3] = e[3];
          d[4] = e[4];
        }
        for (int32_t g = M.level_step_ptr[l]; g < M.level_step_ptr[l + 1]; ++g)
          for (int pass = 0; pass < 2; ++pass) {
            int32_t pa, pe, ua, ue;
            src(p, l, g - M.level_step_ptr[l], pass, pa, pe, ua, ue);
            const size_t j = (static_cast<size_t>(g) * 2 + pass) * S + s;
            for (int32_t t = pa; t < pe; ++t) {
              Task4 tk = p.pan_task[t];
              tk.a += sn;
              M.pan_task[pt[j] + (t - pa)] = tk;
            }
            for (int32_t t = ua; t < ue; ++t) {
              Task4 tk = p.upd_task[t];
              tk.a += sn;
              M.upd_task[ut[j] + (t - ua)] = tk;
            }
          }
      }
    });
  }
  // row gathers of the solves
  M.frow_off.resize(static_cast<size_t>(T.ns) + 1);
  M.grow_ptr.resize(static_cast<size_t>(E.frow) + 1);
  M.gent.resize(static_cast<size_t>(E.gent) + 2);
  parallel_for(static_cast<int64_t>(S), join_threads(), [&](int64_t s) {
    const FactorPlan& p = *P[s];
    const Base& o = b[s];
    for (int32_t j = 0; j < p.sn.ns; ++j) M.frow_off[o.sn + j] = o.frow + p.frow_off[j];
    for (int32_t r = 0; r < p.frow_off[p.sn.ns]; ++r)
      M.grow_ptr[o.frow + r] = static_cast<int32_t>(o.gent + p.grow_ptr[r]);
    for (size_t t = 0; t + 2 < p.gent.size(); ++t) {
      Task4 g = p.gent[t];
      g.a += static_cast<int32_t>(o.rows);
      if (g.c > 0) g.b += static_cast<int32_t>(o.up);
      M.gent[o.gent + t] = g;
    }
  });
  M.frow_off[T.ns] = E.frow;
  M.grow_ptr[E.frow] = static_cast<int32_t>(E.gent);
  M.gent[E.gent] = Task4{0, 0, 0, 0};
  M.gent[E.gent + 1] = Task4{0, 0, 0, 0};
  M.nsmall = static_cast<int64_t>(M.small_list.size());
  M.nbig = static_cast<int64_t>(M.big_list.size());
  build_dag_schedule(M);
}

void combine_patterns(const std::vector<const PatternEntry*>& parts, BatchSystem& B, FactorPlan& plan) {
  const int32_t S = static_cast<int32_t>(parts.size());
  if (S < 1) throw std::invalid_argument("a batch needs at least one system");
  B.nsys = S;
  B.col_base.assign(S + 1, 0);
  B.ent_base.assign(S + 1, 0);
  std::vector<int64_t> csr_base(S + 1, 0), panel_base(S + 1, 0);
  B.nnz_l = 0;
  B.flops = 0.0;
  B.relaxed = true;
  for (int32_t s = 0; s < S; ++s) {
    const PatternEntry& e = *parts[s];
    const int64_t nb = static_cast<int64_t>(B.col_base[s]) + e.n;
    if (nb >= (int64_t(1) << 31)) throw std::invalid_argument("batch exceeds 2^31 rows");
    B.col_base[s + 1] = static_cast<int32_t>(nb);
    B.ent_base[s + 1] = B.ent_base[s] + e.nnz;
    csr_base[s + 1] = csr_base[s] + static_cast<int64_t>(e.csr_j.size());
    panel_base[s + 1] = panel_base[s] + e.an.plan.panel_entries;
    B.nnz_l += e.an.nnz_l;
    B.flops += e.an.flops;
    B.relaxed = B.relaxed && e.an.relaxed;
  }
  const int32_t n = B.col_base[S];
  const int64_t nnz = B.ent_base[S], ncsr = csr_base[S];
  if (nnz >= (int64_t(1) << 31) || ncsr >= (int64_t(1) << 31))
    throw std::invalid_argument("batch exceeds 2^31 matrix entries");
  std::vector<const FactorPlan*> plans;
  for (int32_t s = 0; s < S; ++s) plans.push_back(&parts[s]->an.plan);
  merge_plans(plans, B.col_base, plan);
  B.perm.resize(n);
  B.local_perm.resize(n);
  B.sys_of_row.resize(n);
  B.diag_pos.resize(n);
  B.erow.resize(nnz);
  B.ecol.resize(nnz);
  B.csr_p.resize(static_cast<size_t>(n) + 1);
  B.csr_j.resize(ncsr);
  B.csr_src.resize(ncsr);
  plan.val_dst.resize(nnz);
  B.csr_p[0] = 0;
  parallel_for(S, join_threads(), [&](int64_t s) {
    const PatternEntry& e = *parts[s];
    const Analysis& a = e.an;
    const int32_t cb = B.col_base[s];
    const int64_t eb = B.ent_base[s];
    const int64_t csr_at = csr_base[s];
    for (int32_t k = 0; k < e.n; ++k) {
      B.perm[cb + k] = cb + a.perm[k];
      B.local_perm[cb + k] = a.perm[k];
      B.sys_of_row[cb + k] = static_cast<int32_t>(s);
      B.csr_p[cb + k + 1] = static_cast<int32_t>(csr_at + e.csr_p[k + 1]);
      B.diag_pos[cb + k] = static_cast<int32_t>(eb + e.diag_pos[k]);
    }
    for (int64_t q = 0; q < e.nnz; ++q) {
      B.erow[eb + q] = cb + e.Ai[q];
      B.ecol[eb + q] = cb + e.ecol[q];
      plan.val_dst[eb + q] = panel_base[s] + a.plan.val_dst[q];
    }
    for (size_t q = 0; q < e.csr_j.size(); ++q) {
      B.csr_j[csr_at + q] = cb + e.csr_j[q];
      B.csr_src[csr_at + q] = static_cast<int32_t>(eb + e.csr_src[q]);
    }
  });
}

}  // namespace mpir
