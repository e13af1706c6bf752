// Fill-reducing ordering: approximate minimum degree (AMD).
//
// The reference has no ordering code at all (yex33/MPIR applies QDLDL to the
// matrix as given: include/qdldl-luir/luir.hpp:83, examples/example.cpp:87-138).
// This is written from the published method -- Amestoy, Davis & Duff, "An
// approximate minimum degree ordering algorithm", SIAM J. Matrix Anal. Appl.
// 17(4), 1996 -- on a quotient graph of variables and elements:
//   * pivot = variable of least approximate external degree (LIFO buckets);
//   * the new element absorbs every element adjacent to the pivot;
//   * |Le \ Lp| is computed for every element touching the new element and
//     gives the approximate degree bound
//       d_i <= min(n - k, d_i + |Lp \ i|, |Ai \ i| + |Lp \ i| + sum |Le \ Lp|);
//   * aggressive absorption of elements contained in the new element;
//   * indistinguishable variables are merged into supervariables (hash of
//     adjacency, then exact comparison), and variables adjacent only to the
//     new element are eliminated with it (mass elimination);
//   * rows denser than max(16, 10 sqrt(n)) are removed and ordered last.
// Storage is one growable list per node rather than one compacted workspace.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <utility>

#include "sparse.hpp"

namespace mpir {

namespace {

enum : uint8_t { kVar = 0, kElem = 1, kDeadVar = 2, kDeadElem = 3, kDense = 4 };

struct DegreeBuckets {
  std::vector<int32_t> head, next, prev;
  explicit DegreeBuckets(int32_t n) : head(static_cast<size_t>(n) + 1, -1), next(n, -1), prev(n, -1) {}
  void insert(int32_t i, int32_t d) {
    next[i] = head[d];
    prev[i] = -1;
    if (head[d] >= 0) prev[head[d]] = i;
    head[d] = i;
  }
  void remove(int32_t i, int32_t d) {
    if (prev[i] >= 0) next[prev[i]] = next[i]; else head[d] = next[i];
    if (next[i] >= 0) prev[next[i]] = prev[i];
    next[i] = prev[i] = -1;
  }
};

}  // namespace

std::vector<int32_t> amd_order(const CscUpper& A, int64_t* nnz_est) {
  const int32_t n = A.n;
  std::vector<int32_t> perm;
  if (n == 0) return perm;

  // symmetric adjacency without the diagonal
  std::vector<int32_t> cnt(n, 0);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int32_t i = A.i[q];
      if (i == j) continue;
      cnt[i]++;
      cnt[j]++;
    }
  std::vector<std::vector<int32_t>> Av(n), E(n), Le(n);
  for (int32_t i = 0; i < n; ++i) Av[i].reserve(cnt[i]);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int32_t i = A.i[q];
      if (i == j) continue;
      Av[i].push_back(j);
      Av[j].push_back(i);
    }

  std::vector<uint8_t> kind(n, kVar);
  const double dense_thr = std::max(16.0, 10.0 * std::sqrt(static_cast<double>(n)));
  int32_t ndense = 0;
  for (int32_t i = 0; i < n; ++i)
    if (static_cast<double>(Av[i].size()) > dense_thr && n > 16) {
      kind[i] = kDense;
      ndense++;
    }
  if (ndense > 0) {
    for (int32_t i = 0; i < n; ++i) {
      if (kind[i] == kDense) { Av[i].clear(); continue; }
      auto& L = Av[i];
      L.erase(std::remove_if(L.begin(), L.end(), [&](int32_t j) { return kind[j] == kDense; }), L.end());
    }
  }

  std::vector<int32_t> nv(n, 1), deg(n, 0), mark(n, 0), wstamp(n, 0), merged_into(n, -1), elim_step(n, -1);
  std::vector<int64_t> esize(n, 0), wval(n, 0);
  std::vector<uint64_t> hashv(n, 0);
  DegreeBuckets buckets(n);
  int32_t mindeg = n;
  for (int32_t i = n - 1; i >= 0; --i) {
    if (kind[i] != kVar) continue;
    deg[i] = static_cast<int32_t>(Av[i].size());
    buckets.insert(i, deg[i]);
    mindeg = std::min(mindeg, deg[i]);
  }

  int32_t stamp = 0, wcur = 0;
  int64_t nel = 0;
  const int64_t nlive = n - ndense;
  int32_t step = 0;
  int64_t lnz = 0;
  std::vector<int32_t> Lme;
  std::vector<std::pair<uint64_t, int32_t>> hashed;
  std::vector<uint8_t> gone;  // per Lme slot: removed by mass elimination / merge

  while (nel < nlive) {
    while (mindeg <= n && buckets.head[mindeg] < 0) ++mindeg;
    if (mindeg > n) throw std::logic_error("amd: degree buckets exhausted early");
    const int32_t p = buckets.head[mindeg];
    buckets.remove(p, mindeg);
    elim_step[p] = step++;

    // ---- new element p: union of absorbed elements and p's variables
    ++stamp;
    mark[p] = stamp;
    Lme.clear();
    int64_t degme = 0;
    for (int32_t e : E[p]) {
      if (kind[e] != kElem) continue;
      for (int32_t i : Le[e]) {
        if (kind[i] != kVar || mark[i] == stamp) continue;
        mark[i] = stamp;
        Lme.push_back(i);
        degme += nv[i];
      }
      kind[e] = kDeadElem;
      merged_into[e] = p;
      std::vector<int32_t>().swap(Le[e]);
    }
    for (int32_t i : Av[p]) {
      if (kind[i] != kVar || mark[i] == stamp) continue;
      mark[i] = stamp;
      Lme.push_back(i);
      degme += nv[i];
    }
    std::vector<int32_t>().swap(E[p]);
    std::vector<int32_t>().swap(Av[p]);
    kind[p] = kElem;
    nel += nv[p];
    // entries of L contributed by this pivot block (estimate for statistics)
    lnz += static_cast<int64_t>(nv[p]) * (nv[p] - 1) / 2 + static_cast<int64_t>(nv[p]) * degme;
    for (int32_t i : Lme) buckets.remove(i, deg[i]);

    // ---- |Le \ Lme| for every element adjacent to a variable of Lme
    ++wcur;
    for (int32_t i : Lme)
      for (int32_t e : E[i]) {
        if (kind[e] != kElem || e == p) continue;
        if (wstamp[e] != wcur) { wstamp[e] = wcur; wval[e] = esize[e]; }
        wval[e] -= nv[i];
      }

    // ---- degree update, aggressive absorption, mass elimination
    gone.assign(Lme.size(), 0);
    hashed.clear();
    const int64_t remaining = nlive - nel;
    for (size_t t = 0; t < Lme.size(); ++t) {
      const int32_t i = Lme[t];
      int64_t dext = 0;
      uint64_t h = static_cast<uint64_t>(p);
      std::vector<int32_t>& Ei = E[i];
      // rebuilt element list, p first
      std::vector<int32_t> kept;
      kept.reserve(Ei.size() + 1);
      kept.push_back(p);
      for (int32_t e : Ei) {
        if (kind[e] != kElem || e == p) continue;
        const int64_t we = (wstamp[e] == wcur) ? wval[e] : esize[e];
        if (we <= 0) {  // e is a subset of the new element: absorb it
          kind[e] = kDeadElem;
          merged_into[e] = p;
          std::vector<int32_t>().swap(Le[e]);
          continue;
        }
        dext += we;
        kept.push_back(e);
        h += static_cast<uint64_t>(e);
      }
      Ei.swap(kept);
      std::vector<int32_t>& Ai = Av[i];
      size_t out = 0;
      for (int32_t j : Ai) {
        if (kind[j] != kVar || mark[j] == stamp) continue;
        Ai[out++] = j;
        dext += nv[j];
        h += static_cast<uint64_t>(j);
      }
      Ai.resize(out);
      if (Ei.size() == 1 && Ai.empty()) {  // adjacent only to p: mass elimination
        kind[i] = kDeadVar;
        merged_into[i] = p;
        nel += nv[i];
        degme -= nv[i];
        gone[t] = 1;
        std::vector<int32_t>().swap(Ei);
        continue;
      }
      int64_t d = std::min<int64_t>(static_cast<int64_t>(deg[i]) + degme - nv[i], dext + degme - nv[i]);
      d = std::min<int64_t>(d, remaining - nv[i]);
      deg[i] = static_cast<int32_t>(std::max<int64_t>(d, 0));
      hashv[i] = h;
      hashed.emplace_back(h, static_cast<int32_t>(t));
    }

    // ---- supervariable detection among the survivors
    std::sort(hashed.begin(), hashed.end());
    for (size_t a = 0; a < hashed.size();) {
      size_t b = a + 1;
      while (b < hashed.size() && hashed[b].first == hashed[a].first) ++b;
      if (b - a > 1) {
        for (size_t x = a; x < b; ++x) {
          const int32_t tx = hashed[x].second;
          if (gone[tx]) continue;
          const int32_t ix = Lme[tx];
          ++stamp;
          for (int32_t e : E[ix]) mark[e] = stamp;
          for (int32_t j : Av[ix]) mark[j] = stamp;
          for (size_t y = x + 1; y < b; ++y) {
            const int32_t ty = hashed[y].second;
            if (gone[ty]) continue;
            const int32_t iy = Lme[ty];
            if (E[iy].size() != E[ix].size() || Av[iy].size() != Av[ix].size()) continue;
            bool same = true;
            for (int32_t e : E[iy]) if (mark[e] != stamp) { same = false; break; }
            if (same) for (int32_t j : Av[iy]) if (mark[j] != stamp) { same = false; break; }
            if (!same) continue;
            nv[ix] += nv[iy];
            deg[ix] = std::max(0, deg[ix] - nv[iy]);
            nv[iy] = 0;
            kind[iy] = kDeadVar;
            merged_into[iy] = ix;
            gone[ty] = 1;
            std::vector<int32_t>().swap(E[iy]);
            std::vector<int32_t>().swap(Av[iy]);
          }
        }
      }
      a = b;
    }
    // stamps of the comparison pass must not collide with the Lme mark
    ++stamp;

    // ---- finalise element p and reinsert its variables
    size_t out = 0;
    int64_t esz = 0;
    const int64_t remaining2 = nlive - nel;
    for (size_t t = 0; t < Lme.size(); ++t) {
      if (gone[t]) continue;
      const int32_t i = Lme[t];
      Lme[out++] = i;
      esz += nv[i];
      int64_t d = std::min<int64_t>(deg[i], remaining2 - nv[i]);
      deg[i] = static_cast<int32_t>(std::max<int64_t>(d, 0));
      buckets.insert(i, deg[i]);
      mindeg = std::min(mindeg, deg[i]);
    }
    Lme.resize(out);
    Le[p].assign(Lme.begin(), Lme.end());
    esize[p] = esz;
  }

  // ---- expand supervariables into the final order
  std::vector<int32_t> root_step(n, -1);
  for (int32_t i = 0; i < n; ++i) {
    if (kind[i] == kDense) continue;
    int32_t r = i;
    int guard = 0;
    while (elim_step[r] < 0) {
      r = merged_into[r];
      if (r < 0 || ++guard > n) throw std::logic_error("amd: broken supervariable chain");
    }
    root_step[i] = elim_step[r];
  }
  std::vector<int32_t> idx;
  idx.reserve(n);
  for (int32_t i = 0; i < n; ++i)
    if (kind[i] != kDense) idx.push_back(i);
  // within a pivot group the principal variable first, then the rest by index
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
    if (root_step[a] != root_step[b]) return root_step[a] < root_step[b];
    const bool pa = elim_step[a] >= 0, pb = elim_step[b] >= 0;
    if (pa != pb) return pa;
    return a < b;
  });
  perm = std::move(idx);
  for (int32_t i = 0; i < n; ++i)
    if (kind[i] == kDense) perm.push_back(i);
  if (nnz_est) *nnz_est = lnz;
  return perm;
}

CscUpper permute_upper(const CscUpper& A, const std::vector<int32_t>& perm,
                       std::vector<int32_t>* val_map) {
  const int32_t n = A.n;
  if (static_cast<int32_t>(perm.size()) != n) throw std::invalid_argument("permutation size mismatch");
  std::vector<int32_t> pinv(n, -1);
  for (int32_t k = 0; k < n; ++k) {
    const int32_t o = perm[k];
    if (o < 0 || o >= n || pinv[o] >= 0) throw std::invalid_argument("not a permutation");
    pinv[o] = k;
  }
  CscUpper C;
  C.n = n;
  C.p.assign(static_cast<size_t>(n) + 1, 0);
  const int64_t nnz = A.nnz();
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int32_t a = pinv[A.i[q]], b = pinv[j];
      C.p[std::max(a, b) + 1]++;
    }
  for (int32_t j = 0; j < n; ++j) C.p[j + 1] += C.p[j];
  std::vector<int32_t> next(C.p.begin(), C.p.end() - 1);
  std::vector<int32_t> src(nnz);
  C.i.resize(nnz);
  C.x.resize(nnz);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int32_t a = pinv[A.i[q]], b = pinv[j];
      const int32_t col = std::max(a, b), row = std::min(a, b);
      const int32_t pos = next[col]++;
      C.i[pos] = row;
      src[pos] = q;
    }
  // sort rows within each column, carrying the source index
  std::vector<std::pair<int32_t, int32_t>> tmp;
  for (int32_t j = 0; j < n; ++j) {
    const int32_t b = C.p[j], e = C.p[j + 1];
    bool sorted = true;
    for (int32_t q = b + 1; q < e; ++q)
      if (C.i[q - 1] > C.i[q]) { sorted = false; break; }
    if (!sorted) {
      tmp.clear();
      for (int32_t q = b; q < e; ++q) tmp.emplace_back(C.i[q], src[q]);
      std::sort(tmp.begin(), tmp.end());
      for (int32_t q = b; q < e; ++q) { C.i[q] = tmp[q - b].first; src[q] = tmp[q - b].second; }
    }
  }
  for (int64_t q = 0; q < nnz; ++q) C.x[q] = A.x[src[q]];
  if (val_map) *val_map = std::move(src);
  return C;
}

}  // namespace mpir
