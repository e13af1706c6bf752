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
    S.panel_off[s + 1] = S.panel_off[s] + panel_ld(nc, S.nrow[s]) * nc;
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

namespace mpir {

// Relaxed supernodes. Starting from the fundamental supernodes of the
// (postordered) factor pattern, a child is merged into its parent when the
// merged block stays small or gains few explicit zeros (thresholds after
// Ashcraft & Grimes, "The influence of relaxed supernode partitions on the
// multifrontal method", ACM TOMS 15(4), 1989). Merging moves the child's
// columns to just before the parent's, which is a topological reordering of
// the elimination tree: the fill (and every L value) is unchanged.
// Returns new -> old column order (of C's columns) and the block starts.
RelaxedPartition relax_supernodes(const CscUpper& C, const int32_t* parent, const int32_t* Lnz, int level) {
  const int32_t n = C.n;
  // fundamental partition
  std::vector<int32_t> first{0};
  for (int32_t j = 0; j + 1 < n; ++j)
    if (!(parent[j] == j + 1 && Lnz[j] == Lnz[j + 1] + 1)) first.push_back(j + 1);
  first.push_back(n);
  const int32_t ns = static_cast<int32_t>(first.size()) - 1;
  std::vector<int32_t> col2sn(n);
  for (int32_t s = 0; s < ns; ++s)
    for (int32_t j = first[s]; j < first[s + 1]; ++j) col2sn[j] = s;
  std::vector<int64_t> nc(ns), nr(ns);
  std::vector<double> zeros(ns, 0.0);
  std::vector<int32_t> sp(ns, -1);
  for (int32_t s = 0; s < ns; ++s) {
    nc[s] = first[s + 1] - first[s];
    nr[s] = Lnz[first[s + 1] - 1];
    const int32_t pc = parent[first[s + 1] - 1];
    sp[s] = pc >= 0 ? col2sn[pc] : -1;
  }
  // children lists (ascending)
  std::vector<int32_t> cptr(static_cast<size_t>(ns) + 1, 0), cl(ns);
  for (int32_t s = 0; s < ns; ++s)
    if (sp[s] >= 0) cptr[sp[s] + 1]++;
  for (int32_t s = 0; s < ns; ++s) cptr[s + 1] += cptr[s];
  {
    std::vector<int32_t> next(cptr.begin(), cptr.end() - 1);
    for (int32_t s = 0; s < ns; ++s)
      if (sp[s] >= 0) cl[next[sp[s]]++] = s;
  }
  // bottom-up merging; merged[c] = true when c joined its parent's group
  std::vector<uint8_t> merged(ns, 0);
  auto entries = [](double c, double r) { return c * (c - 1) / 2 + c * r; };
  for (int32_t p = 0; p < ns; ++p) {
    // children by descending column count: big children first
    std::vector<int32_t> kids(cl.begin() + cptr[p], cl.begin() + cptr[p + 1]);
    std::sort(kids.begin(), kids.end(), [&](int32_t a, int32_t b) { return nc[a] > nc[b] || (nc[a] == nc[b] && a < b); });
    for (int32_t c : kids) {
      const double ncm = static_cast<double>(nc[p] + nc[c]);
      const double z = zeros[p] + zeros[c] + static_cast<double>(nc[c]) * static_cast<double>(nc[p] + nr[p] - nr[c]);
      const double frac = z / std::max(1.0, entries(ncm, static_cast<double>(nr[p])));
      // merged front sizes up to which a zero fraction is accepted, per level
      const double small = level >= 3 ? 64 : level == 2 ? 32 : 16;
      const double mid = level >= 3 ? 128 : level == 2 ? 64 : 48;
      const double mid_frac = level >= 2 ? 0.2 : 0.1;
      const bool take = ncm <= 4 || (ncm <= small && frac < 0.8) || (ncm <= mid && frac < mid_frac) || frac < 0.05;
      if (!take) continue;
      merged[c] = 1;
      zeros[p] = z;
      nc[p] += nc[c];
    }
  }
  // new order: postorder of the merged tree; a group lists the columns of its
  // merged members (recursively, each member after its own merged children)
  // after all non-merged subtrees hanging below the group
  RelaxedPartition out;
  out.order.reserve(n);
  out.first.push_back(0);
  std::vector<int32_t> gmem;  // members of the current group
  // iterative DFS over groups: a group is identified by its top supernode
  struct Frame { int32_t top; int stage; std::vector<int32_t> subs; size_t next; };
  std::vector<Frame> stack;
  auto collect = [&](int32_t top, std::vector<int32_t>& members, std::vector<int32_t>& subs) {
    // members in column order (merged children first, recursively), subs =
    // non-merged children of any member (roots of other groups)
    std::vector<std::pair<int32_t, int>> st{{top, 0}};
    while (!st.empty()) {
      const int32_t v = st.back().first;
      if (st.back().second == 0) {
        st.back().second = 1;
        for (int32_t t = cptr[v]; t < cptr[v + 1]; ++t) {
          const int32_t c = cl[t];
          if (merged[c]) st.push_back({c, 0});
          else subs.push_back(c);
        }
      } else {
        members.push_back(v);
        st.pop_back();
      }
    }
  };
  for (int32_t root = 0; root < ns; ++root) {
    if (sp[root] >= 0) continue;
    stack.push_back({root, 0, {}, 0});
    while (!stack.empty()) {
      Frame& fr = stack.back();
      if (fr.stage == 0) {
        std::vector<int32_t> members;
        collect(fr.top, members, fr.subs);
        std::sort(fr.subs.begin(), fr.subs.end());
        fr.stage = 1;
        gmem = std::move(members);
        stack.back().next = 0;
        // remember members in the frame via a parallel vector
        out.pending.push_back(std::move(gmem));
        continue;
      }
      if (fr.next < fr.subs.size()) {
        const int32_t child = fr.subs[fr.next++];
        stack.push_back({child, 0, {}, 0});
        continue;
      }
      // all sub-groups emitted: emit this group's columns
      std::vector<int32_t> members = std::move(out.pending.back());
      out.pending.pop_back();
      for (int32_t v : members)
        for (int32_t j = first[v]; j < first[v + 1]; ++j) out.order.push_back(j);
      out.first.push_back(static_cast<int32_t>(out.order.size()));
      stack.pop_back();
    }
  }
  if (static_cast<int32_t>(out.order.size()) != n) throw std::logic_error("relaxed supernodes lost columns");
  return out;
}

Supernodes build_supernodes_forced(const CscUpper& C, const int32_t* parent, const std::vector<int32_t>& first_in) {
  Supernodes S;
  const int32_t n = C.n;
  S.n = n;
  S.first = first_in;
  S.ns = static_cast<int32_t>(S.first.size()) - 1;
  const int32_t ns = S.ns;
  S.col2sn.assign(n, -1);
  S.sparent.assign(ns, -1);
  for (int32_t s = 0; s < ns; ++s)
    for (int32_t j = S.first[s]; j < S.first[s + 1]; ++j) S.col2sn[j] = s;
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t pc = parent[S.first[s + 1] - 1];
    S.sparent[s] = pc >= 0 ? S.col2sn[pc] : -1;
    if (S.sparent[s] >= 0 && S.sparent[s] <= s) throw std::logic_error("relaxed supernodal tree not topological");
  }
  std::vector<int32_t> tp(static_cast<size_t>(n) + 1, 0), ti(C.nnz());
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = C.p[j]; q < C.p[j + 1]; ++q)
      if (C.i[q] != j) tp[C.i[q] + 1]++;
  for (int32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
  {
    std::vector<int32_t> next(tp.begin(), tp.end() - 1);
    for (int32_t j = 0; j < n; ++j)
      for (int32_t q = C.p[j]; q < C.p[j + 1]; ++q)
        if (C.i[q] != j) ti[next[C.i[q]]++] = j;
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
  // R_s by union (no count check: relaxed blocks hold explicit zeros)
  S.nrow.assign(ns, 0);
  S.rows_off.assign(static_cast<size_t>(ns) + 1, 0);
  std::vector<int32_t> mark(n, -1), buf;
  std::vector<std::vector<int32_t>> R(ns);
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t f = S.first[s], l = S.first[s + 1];
    buf.clear();
    auto add = [&](int32_t r) {
      if (r >= l && mark[r] != s) {
        mark[r] = s;
        buf.push_back(r);
      }
    };
    for (int32_t j = f; j < l; ++j)
      for (int32_t q = tp[j]; q < tp[j + 1]; ++q) add(ti[q]);
    for (int32_t t = cptr[s]; t < cptr[s + 1]; ++t) {
      for (int32_t r : R[clist[t]]) add(r);
    }
    std::sort(buf.begin(), buf.end());
    R[s] = buf;
    S.nrow[s] = static_cast<int32_t>(buf.size());
    S.rows_off[s + 1] = S.rows_off[s] + S.nrow[s];
    for (int32_t t = cptr[s]; t < cptr[s + 1]; ++t) std::vector<int32_t>().swap(R[clist[t]]);
  }
  // rows of the remaining (root) lists were kept; rebuild the flat array
  S.rows.resize(S.rows_off[ns]);
  {
    // recompute R in a second pass to fill the flat array (children lists
    // were released as soon as their parent absorbed them)
    std::fill(mark.begin(), mark.end(), -1);
    for (int32_t s = 0; s < ns; ++s) {
      const int32_t f = S.first[s], l = S.first[s + 1];
      int32_t* Rs = S.rows.data() + S.rows_off[s];
      int64_t cnt = 0;
      auto add = [&](int32_t r) {
        if (r >= l && mark[r] != s) {
          mark[r] = s;
          Rs[cnt++] = r;
        }
      };
      for (int32_t j = f; j < l; ++j)
        for (int32_t q = tp[j]; q < tp[j + 1]; ++q) add(ti[q]);
      for (int32_t t = cptr[s]; t < cptr[s + 1]; ++t) {
        const int32_t c = clist[t];
        for (int64_t q = S.rows_off[c]; q < S.rows_off[c + 1]; ++q) add(S.rows[q]);
      }
      std::sort(Rs, Rs + cnt);
    }
  }
  finish_supernodes(S);
  return S;
}

}  // namespace mpir
