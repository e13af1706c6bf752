#include "batch_plan.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "host/sparse.hpp"

namespace mpir {

uint64_t pattern_key(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, bool relax) {
  // FNV-1a over the dimension, the options and the index arrays
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* data, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) {
      h ^= p[i];
      h *= 1099511628211ull;
    }
  };
  const int32_t opts[3] = {n, ordering, relax ? 1 : 0};
  mix(opts, sizeof opts);
  mix(Ap, sizeof(int32_t) * (static_cast<size_t>(n) + 1));
  mix(Ai, sizeof(int32_t) * static_cast<size_t>(Ap[n]));
  return h;
}

PatternEntry analyze_entry(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, bool relax) {
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
       static_cast<int64_t>(P.dag_fwd.size() + P.dag_bwd.size() + P.gent.size()) * 16;  // schedules
  return b;
}

std::shared_ptr<const PatternEntry> PatternCache::find_locked(uint64_t key, int32_t n, const int32_t* Ap,
                                                              const int32_t* Ai, int ordering, bool relax) const {
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
                                                      bool relax) {
  return get_many({n}, {Ap}, {Ai}, ordering, relax, 1)[0];
}

std::vector<std::shared_ptr<const PatternEntry>> PatternCache::get_many(const std::vector<int32_t>& n,
                                                                        const std::vector<const int32_t*>& Ap,
                                                                        const std::vector<const int32_t*>& Ai,
                                                                        int ordering, bool relax, int threads) {
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

BatchSystem combine_patterns(const std::vector<const PatternEntry*>& parts) {
  BatchSystem B;
  const int32_t S = static_cast<int32_t>(parts.size());
  if (S < 1) throw std::invalid_argument("a batch needs at least one system");
  B.nsys = S;
  B.col_base.assign(S + 1, 0);
  B.ent_base.assign(S + 1, 0);
  int64_t ncsr = 0, nrows = 0, ns_tot = 0;
  for (int32_t s = 0; s < S; ++s) {
    const int64_t nb = static_cast<int64_t>(B.col_base[s]) + parts[s]->n;
    if (nb >= (int64_t(1) << 31)) throw std::invalid_argument("batch exceeds 2^31 rows");
    B.col_base[s + 1] = static_cast<int32_t>(nb);
    B.ent_base[s + 1] = B.ent_base[s] + parts[s]->nnz;
    ncsr += static_cast<int64_t>(parts[s]->csr_j.size());
    nrows += static_cast<int64_t>(parts[s]->an.plan.sn.rows.size());
    ns_tot += parts[s]->an.plan.sn.ns;
  }
  const int32_t n = B.col_base[S];
  const int64_t nnz = B.ent_base[S];
  if (nnz >= (int64_t(1) << 31) || ncsr >= (int64_t(1) << 31))
    throw std::invalid_argument("batch exceeds 2^31 matrix entries");

  Analysis& A = B.an;
  A.perm.resize(n);
  B.local_perm.resize(n);
  B.sys_of_row.resize(n);
  A.C.n = n;
  A.C.p.assign(static_cast<size_t>(n) + 1, 0);
  A.C.i.resize(nnz);
  A.val_map.resize(nnz);
  A.parent.resize(n);
  A.Lnz.resize(n);
  B.Ap.assign(static_cast<size_t>(n) + 1, 0);
  B.Ai.resize(nnz);
  B.csr_p.assign(static_cast<size_t>(n) + 1, 0);
  B.csr_j.resize(ncsr);
  B.csr_src.resize(ncsr);
  B.ecol.resize(nnz);
  B.diag_pos.resize(n);
  A.nnz_l = 0;
  A.flops = 0.0;
  A.relaxed = true;

  Supernodes& T = A.plan.sn;
  T.n = n;
  T.ns = static_cast<int32_t>(ns_tot);
  T.first.reserve(ns_tot + 1);
  T.nrow.reserve(ns_tot);
  T.rows_off.reserve(ns_tot + 1);
  T.rows.reserve(nrows);
  T.sparent.reserve(ns_tot);
  T.col2sn.resize(n);
  std::vector<int64_t> val_dst_sys(nnz);  // per original entry: panel index within its system
  std::vector<int64_t> panel_base_sn;     // first supernode of each system
  int64_t csr_at = 0;
  int32_t sn_base = 0;
  for (int32_t s = 0; s < S; ++s) {
    const PatternEntry& e = *parts[s];
    const Analysis& a = e.an;
    const Supernodes& P = a.plan.sn;
    const int32_t cb = B.col_base[s];
    const int64_t eb = B.ent_base[s];
    for (int32_t k = 0; k < e.n; ++k) {
      A.perm[cb + k] = cb + a.perm[k];
      B.local_perm[cb + k] = a.perm[k];
      B.sys_of_row[cb + k] = s;
      A.parent[cb + k] = a.parent[k] >= 0 ? cb + a.parent[k] : -1;
      A.Lnz[cb + k] = a.Lnz[k];
      A.C.p[cb + k + 1] = static_cast<int32_t>(eb + a.C.p[k + 1]);
      B.Ap[cb + k + 1] = static_cast<int32_t>(eb + e.Ap[k + 1]);
      B.csr_p[cb + k + 1] = static_cast<int32_t>(csr_at + e.csr_p[k + 1]);
      B.diag_pos[cb + k] = static_cast<int32_t>(eb + e.diag_pos[k]);
      T.col2sn[cb + k] = sn_base + P.col2sn[k];
    }
    for (int64_t q = 0; q < e.nnz; ++q) {
      A.C.i[eb + q] = cb + a.C.i[q];
      A.val_map[eb + q] = static_cast<int32_t>(eb + a.val_map[q]);
      B.Ai[eb + q] = cb + e.Ai[q];
      B.ecol[eb + q] = cb + e.ecol[q];
      val_dst_sys[eb + q] = a.plan.val_dst[q];
    }
    for (size_t q = 0; q < e.csr_j.size(); ++q) {
      B.csr_j[csr_at + q] = cb + e.csr_j[q];
      B.csr_src[csr_at + q] = static_cast<int32_t>(eb + e.csr_src[q]);
    }
    csr_at += static_cast<int64_t>(e.csr_j.size());
    const int64_t rb = static_cast<int64_t>(T.rows.size());
    for (int32_t j = 0; j < P.ns; ++j) {
      T.first.push_back(cb + P.first[j]);
      T.nrow.push_back(P.nrow[j]);
      T.rows_off.push_back(rb + P.rows_off[j]);
      T.sparent.push_back(P.sparent[j] >= 0 ? sn_base + P.sparent[j] : -1);
    }
    for (int32_t r : P.rows) T.rows.push_back(cb + r);
    panel_base_sn.push_back(sn_base);
    sn_base += P.ns;
    A.nnz_l += a.nnz_l;
    A.flops += a.flops;
    A.relaxed = A.relaxed && a.relaxed;
  }
  T.first.push_back(n);
  T.rows_off.push_back(static_cast<int64_t>(T.rows.size()));
  // levels, panels and relative indices of the forest; per system they
  // equal the system's own
  finish_supernodes(T);
  FactorPlan plan = build_plan_from_supernodes(std::move(T));
  plan.val_dst.resize(nnz);
  for (int32_t s = 0; s < S; ++s) {
    const int64_t pb = plan.sn.panel_off[panel_base_sn[s]];
    for (int64_t q = B.ent_base[s]; q < B.ent_base[s + 1]; ++q) plan.val_dst[q] = pb + val_dst_sys[q];
  }
  A.plan = std::move(plan);
  return B;
}

}  // namespace mpir
