// Symbolic analysis: elimination tree, column counts, row pattern of L,
// supernodes and the level schedule of the supernodal (assembly) tree.
#include <algorithm>
#include <limits>
#include <stdexcept>

#include "sparse.hpp"

namespace mpir {

// Same contract and same tree as QDLDL_etree (include/MPIR/qdldl.hpp:48-105):
// for each column j and each upper entry (i, j), walk from i towards the root
// until a node already visited for j; every node on the way gains a row-j
// entry in L and the first unknown parent met becomes j.
int64_t etree_counts(int32_t n, const int32_t* Ap, const int32_t* Ai,
                     int32_t* work, int32_t* Lnz, int32_t* parent, bool check_i32) {
  for (int32_t c = 0; c < n; ++c) {
    work[c] = 0;
    Lnz[c] = 0;
    parent[c] = -1;
    if (Ap[c] == Ap[c + 1]) return -1;
  }
  for (int32_t j = 0; j < n; ++j) {
    work[j] = j;
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q) {
      int32_t i = Ai[q];
      if (i > j) return -1;
      while (work[i] != j) {
        if (parent[i] == -1) parent[i] = j;
        Lnz[i]++;
        work[i] = j;
        i = parent[i];
      }
    }
  }
  int64_t total = 0;
  for (int32_t c = 0; c < n; ++c) {
    if (check_i32 && total > std::numeric_limits<int32_t>::max() - Lnz[c]) return -2;
    total += Lnz[c];
  }
  return total;
}

std::vector<int32_t> tree_postorder(int32_t n, const int32_t* parent) {
  std::vector<int32_t> head(n, -1), next(n, -1), post;
  post.reserve(n);
  // children lists in ascending order: insert in descending index order
  for (int32_t j = n - 1; j >= 0; --j) {
    const int32_t p = parent[j];
    if (p < 0) continue;
    next[j] = head[p];
    head[p] = j;
  }
  std::vector<int32_t> stack;
  for (int32_t r = 0; r < n; ++r) {
    if (parent[r] >= 0) continue;
    stack.push_back(r);
    while (!stack.empty()) {
      const int32_t v = stack.back();
      const int32_t c = head[v];
      if (c < 0) {
        post.push_back(v);
        stack.pop_back();
      } else {
        head[v] = next[c];
        stack.push_back(c);
      }
    }
  }
  return post;
}

void l_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, const int32_t* parent,
               const int32_t* Lnz, std::vector<int64_t>& Lp, std::vector<int32_t>& Li) {
  Lp.assign(static_cast<size_t>(n) + 1, 0);
  for (int32_t c = 0; c < n; ++c) Lp[c + 1] = Lp[c] + Lnz[c];
  Li.resize(Lp[n]);
  std::vector<int64_t> fill(Lp.begin(), Lp.end() - 1);
  std::vector<int32_t> seen(n, -1);
  // row k of L = nodes reached from the entries of column k of the upper
  // part by walking up the tree (the row subtree of k), appended in k order
  for (int32_t k = 0; k < n; ++k) {
    seen[k] = k;
    for (int32_t q = Ap[k]; q < Ap[k + 1]; ++q) {
      int32_t i = Ai[q];
      while (i >= 0 && i < k && seen[i] != k) {
        seen[i] = k;
        Li[fill[i]++] = k;
        i = parent[i];
      }
    }
  }
}

// Levels, panel offsets, statistics and relative indices of a supernodal
// partition whose first/nrow/rows_off/rows/sparent are filled in.
void finish_supernodes(Supernodes& S) {
  const int32_t n = S.n;
  const int32_t ns = S.ns;
  // levels: leaves 0, parent above its highest child
  S.level.assign(ns, 0);
  for (int32_t s = 0; s < ns; ++s)
    if (S.sparent[s] >= 0) S.level[S.sparent[s]] = std::max(S.level[S.sparent[s]], S.level[s] + 1);
  S.nlevels = 0;
  for (int32_t s = 0; s < ns; ++s) S.nlevels = std::max(S.nlevels, S.level[s] + 1);
  S.level_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  for (int32_t s = 0; s < ns; ++s) S.level_ptr[S.level[s] + 1]++;
  for (int32_t l = 0; l < S.nlevels; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
  S.level_sn.resize(ns);
  {
    std::vector<int32_t> next(S.level_ptr.begin(), S.level_ptr.end() - 1);
    for (int32_t s = 0; s < ns; ++s) S.level_sn[next[S.level[s]]++] = s;
  }
  // panels, statistics
  S.panel_off.assign(static_cast<size_t>(ns) + 1, 0);
  S.nnz_l = 0;
  S.flops = static_cast<double>(n);
  S.max_front = 0;
  for (int32_t s = 0; s < ns; ++s) {
    const int64_t nc = S.first[s + 1] - S.first[s];
    const int64_t m = nc + S.nrow[s];
    S.panel_off[s + 1] = S.panel_off[s] + m * nc;
    S.max_front = std::max<int32_t>(S.max_front, static_cast<int32_t>(m));
  }
  for (int32_t s = 0; s < ns; ++s) {
    const int64_t nc = S.first[s + 1] - S.first[s];
    for (int64_t c = 0; c < nc; ++c) {
      const double lnz = static_cast<double>(nc - c - 1 + S.nrow[s]);
      S.nnz_l += static_cast<int64_t>(lnz);
      S.flops += lnz * (lnz + 2.0);
    }
  }
  // relative indices of each R_s inside its parent's front [cols(p), R_p]
  S.relidx.resize(S.rows.size());
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t p = S.sparent[s];
    if (p < 0) {
      if (S.nrow[s] != 0) throw std::logic_error("root supernode with rows below");
      continue;
    }
    const int32_t fp = S.first[p], lp = S.first[p + 1];
    const int32_t* Rp = S.rows.data() + S.rows_off[p];
    const int32_t nrp = S.nrow[p];
    int32_t cursor = 0;
    for (int64_t q = S.rows_off[s]; q < S.rows_off[s + 1]; ++q) {
      const int32_t r = S.rows[q];
      int32_t pos;
      if (r < lp) {
        if (r < fp) throw std::logic_error("row of child below parent's first column");
        pos = r - fp;
      } else {
        while (cursor < nrp && Rp[cursor] < r) ++cursor;
        if (cursor >= nrp || Rp[cursor] != r) throw std::logic_error("child rows not contained in parent front");
        pos = (lp - fp) + cursor;
      }
      S.relidx[q] = pos;
    }
  }
}

Supernodes build_supernodes(const CscUpper& C, const int32_t* parent, const int32_t* Lnz) {
  Supernodes S;
  const int32_t n = C.n;
  S.n = n;
  // columns j and j+1 share a supernode when j+1 is j's parent and column j's
  // pattern is column j+1's pattern plus row j+1 (counts differ by one);
  // then every column's pattern is {later columns of the block} + R_s.
  S.first.push_back(0);
  for (int32_t j = 0; j + 1 < n; ++j) {
    const bool chain = parent[j] == j + 1 && Lnz[j] == Lnz[j + 1] + 1;
    if (!chain) S.first.push_back(j + 1);
  }
  if (n > 0) S.first.push_back(n);
  S.ns = n > 0 ? static_cast<int32_t>(S.first.size()) - 1 : 0;
  const int32_t ns = S.ns;
  S.col2sn.assign(n, -1);
  S.nrow.assign(ns, 0);
  S.rows_off.assign(static_cast<size_t>(ns) + 1, 0);
  S.sparent.assign(ns, -1);
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t f = S.first[s], l = S.first[s + 1];
    for (int32_t j = f; j < l; ++j) S.col2sn[j] = s;
    S.nrow[s] = Lnz[l - 1];
    S.rows_off[s + 1] = S.rows_off[s] + S.nrow[s];
  }
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t l = S.first[s + 1];
    const int32_t pc = parent[l - 1];
    S.sparent[s] = pc >= 0 ? S.col2sn[pc] : -1;
    if (S.sparent[s] >= 0 && S.sparent[s] <= s) throw std::logic_error("supernodal tree not topological");
  }
  // R_s = rows >= l_s of (the strictly-lower entries of the columns of s)
  // united with the children's R_c: the supernodal form of
  // struct(L_j) = struct(A_j) + union of struct(L_c) over tree children c.
  std::vector<int32_t> tp(static_cast<size_t>(n) + 1, 0), ti(C.nnz());
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = C.p[j]; q < C.p[j + 1]; ++q)
      if (C.i[q] != j) tp[C.i[q] + 1]++;
  for (int32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
  {
    std::vector<int32_t> next(tp.begin(), tp.end() - 1);
    for (int32_t j = 0; j < n; ++j)
      for (int32_t q = C.p[j]; q < C.p[j + 1]; ++q)
        if (C.i[q] != j) ti[next[C.i[q]]++] = j;  // row C.i[q] of the upper part
  }
  std::vector<int32_t> cptr(static_cast<size_t>(ns) + 1, 0), clist(ns);
  for (int32_t s = 0; s < ns; ++s)
    if (S.sparent[s] >= 0) cptr[S.sparent[s] + 1]++;
  for (int32_t s = 0; s < ns; ++s) cptr[s + 1] += cptr[s];
  {
    std::vector<int32_t> next(cptr.begin(), cptr.end() - 1);
    for (int32_t s = 0; s < ns; ++s)
      if (S.sparent[s] >= 0) clist[next[S.sparent[s]]++] = s;
  }
  S.rows.resize(S.rows_off[ns]);
  std::vector<int32_t> mark(n, -1);
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t f = S.first[s], l = S.first[s + 1];
    int32_t* R = S.rows.data() + S.rows_off[s];
    int64_t cnt = 0;
    auto add = [&](int32_t r) {
      if (r >= l && mark[r] != s) {
        if (cnt >= S.nrow[s]) throw std::logic_error("supernode rows exceed the column count");
        mark[r] = s;
        R[cnt++] = r;
      }
    };
    for (int32_t j = f; j < l; ++j)
      for (int32_t q = tp[j]; q < tp[j + 1]; ++q) add(ti[q]);
    for (int32_t t = cptr[s]; t < cptr[s + 1]; ++t) {
      const int32_t c = clist[t];
      for (int64_t q = S.rows_off[c]; q < S.rows_off[c + 1]; ++q) add(S.rows[q]);
    }
    if (cnt != S.nrow[s]) throw std::logic_error("supernode rows disagree with the column count");
    std::sort(R, R + cnt);
  }
  finish_supernodes(S);
  return S;
}

Supernodes supernodes_from_l(int32_t n, const int64_t* Lp, const int32_t* Li) {
  Supernodes S;
  S.n = n;
  std::vector<int32_t> parent(n, -1), Lnz(n, 0);
  for (int32_t j = 0; j < n; ++j) {
    const int64_t b = Lp[j], e = Lp[j + 1];
    if (e < b) throw std::invalid_argument("column pointers must be non-decreasing");
    Lnz[j] = static_cast<int32_t>(e - b);
    for (int64_t q = b; q < e; ++q) {
      if (Li[q] <= j || Li[q] >= n) throw std::invalid_argument("L entries must lie strictly below the diagonal");
      if (q > b && Li[q] <= Li[q - 1]) throw std::invalid_argument("L row indices must be ascending");
    }
    if (e > b) parent[j] = Li[b];
  }
  S.first.push_back(0);
  for (int32_t j = 0; j + 1 < n; ++j) {
    bool chain = parent[j] == j + 1 && Lnz[j] == Lnz[j + 1] + 1;
    if (chain) {  // column j's rows must be {j+1} + column j+1's rows
      for (int64_t t = 0; t < Lnz[j + 1]; ++t)
        if (Li[Lp[j] + 1 + t] != Li[Lp[j + 1] + t]) { chain = false; break; }
    }
    if (!chain) S.first.push_back(j + 1);
  }
  if (n > 0) S.first.push_back(n);
  S.ns = n > 0 ? static_cast<int32_t>(S.first.size()) - 1 : 0;
  const int32_t ns = S.ns;
  S.col2sn.assign(n, -1);
  S.nrow.assign(ns, 0);
  S.rows_off.assign(static_cast<size_t>(ns) + 1, 0);
  S.sparent.assign(ns, -1);
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t f = S.first[s], l = S.first[s + 1];
    for (int32_t j = f; j < l; ++j) S.col2sn[j] = s;
    S.nrow[s] = Lnz[l - 1];
    S.rows_off[s + 1] = S.rows_off[s] + S.nrow[s];
  }
  S.rows.resize(S.rows_off[ns]);
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t l = S.first[s + 1];
    std::copy(Li + Lp[l - 1], Li + Lp[l], S.rows.begin() + S.rows_off[s]);
    const int32_t pc = parent[l - 1];
    S.sparent[s] = pc >= 0 ? S.col2sn[pc] : -1;
  }
  finish_supernodes(S);  // throws when a child's rows do not nest in its parent
  return S;
}

}  // namespace mpir
