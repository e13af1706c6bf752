// Fill-reducing ordering: nested dissection with multilevel bisection.
//
// The reference has no ordering code (yex33/MPIR factors the matrix as given,
// include/qdldl-luir/luir.hpp:83). This ordering follows the published
// multilevel scheme (Karypis & Kumar, "A fast and high quality multilevel
// scheme for partitioning irregular graphs", SIAM J. Sci. Comput. 20(1),
// 1998; George, "Nested dissection of a regular finite element mesh", SIAM
// J. Numer. Anal. 10, 1973):
//   * coarsen by heavy-edge matching until the graph is small;
//   * bisect the coarsest graph by greedy graph growing from several seeds;
//   * project back level by level, refining the edge cut with
//     Fiduccia-Mattheyses passes (best prefix kept, balance bounded);
//   * turn the edge cut into a minimum vertex separator (a minimum vertex
//     cover of the cut's bipartite graph, Konig's theorem);
//   * order part A, then part B, then the separator, recursively; subgraphs
//     below a leaf size are ordered by AMD.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <utility>
#include <vector>

#include "sparse.hpp"

namespace mpir {

namespace {

struct Graph {
  int32_t n = 0;
  std::vector<int32_t> xadj, adj;  // CSR, no self loops
  std::vector<int32_t> ewgt;       // edge weights, parallel to adj
  std::vector<int32_t> vwgt;       // vertex weights
  int64_t total_vwgt() const { return std::accumulate(vwgt.begin(), vwgt.end(), int64_t(0)); }
};

// Small deterministic generator (SplitMix64) for visiting orders and seeds.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  int32_t below(int32_t n) { return static_cast<int32_t>(next() % static_cast<uint64_t>(n)); }
};

// Heavy-edge matching; returns the coarse graph and cmap (fine -> coarse).
Graph coarsen(const Graph& G, std::vector<int32_t>& cmap, Rng& rng) {
  const int32_t n = G.n;
  std::vector<int32_t> match(n, -1), order(n);
  std::iota(order.begin(), order.end(), 0);
  for (int32_t i = n - 1; i > 0; --i) std::swap(order[i], order[rng.below(i + 1)]);
  for (int32_t v : order) {
    if (match[v] >= 0) continue;
    int32_t best = -1, bw = -1;
    for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
      const int32_t u = G.adj[q];
      if (match[u] < 0 && u != v && G.ewgt[q] > bw) {
        bw = G.ewgt[q];
        best = u;
      }
    }
    if (best >= 0) {
      match[v] = best;
      match[best] = v;
    } else {
      match[v] = v;
    }
  }
  cmap.assign(n, -1);
  int32_t nc = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (cmap[v] >= 0) continue;
    cmap[v] = nc;
    cmap[match[v]] = nc;
    ++nc;
  }
  Graph C;
  C.n = nc;
  C.vwgt.assign(nc, 0);
  for (int32_t v = 0; v < n; ++v) C.vwgt[cmap[v]] += G.vwgt[v];
  C.xadj.assign(static_cast<size_t>(nc) + 1, 0);
  std::vector<int32_t> pos(nc, -1);
  std::vector<int32_t> members(static_cast<size_t>(nc) * 2, -1);
  for (int32_t v = 0; v < n; ++v) {
    const int32_t c = cmap[v];
    if (members[2 * c] < 0) members[2 * c] = v; else members[2 * c + 1] = v;
  }
  C.adj.reserve(G.adj.size() / 2 + 16);
  C.ewgt.reserve(G.adj.size() / 2 + 16);
  for (int32_t c = 0; c < nc; ++c) {
    const int32_t start = static_cast<int32_t>(C.adj.size());
    for (int k = 0; k < 2; ++k) {
      const int32_t v = members[2 * c + k];
      if (v < 0) continue;
      for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
        const int32_t cu = cmap[G.adj[q]];
        if (cu == c) continue;
        if (pos[cu] < start) {
          pos[cu] = static_cast<int32_t>(C.adj.size());
          C.adj.push_back(cu);
          C.ewgt.push_back(G.ewgt[q]);
        } else {
          C.ewgt[pos[cu]] += G.ewgt[q];
        }
      }
    }
    C.xadj[c + 1] = static_cast<int32_t>(C.adj.size());
  }
  return C;
}

int64_t edge_cut(const Graph& G, const std::vector<uint8_t>& part) {
  int64_t cut = 0;
  for (int32_t v = 0; v < G.n; ++v)
    for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q)
      if (part[G.adj[q]] != part[v]) cut += G.ewgt[q];
  return cut / 2;
}

// Greedy graph growing from `seed`: part 0 grows by the vertex with the
// largest (edges into part 0) - (edges outside) until it holds half the weight.
std::vector<uint8_t> grow_partition(const Graph& G, int32_t seed, int64_t target) {
  const int32_t n = G.n;
  std::vector<uint8_t> part(n, 1);
  std::vector<int64_t> gain(n, 0);
  std::vector<uint8_t> inq(n, 0);
  using Item = std::pair<int64_t, int32_t>;
  std::priority_queue<Item> pq;
  int64_t w0 = 0;
  auto add = [&](int32_t v) {
    part[v] = 0;
    w0 += G.vwgt[v];
    for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
      const int32_t u = G.adj[q];
      if (part[u] == 0) continue;
      gain[u] += 2 * G.ewgt[q];
      inq[u] = 1;
      pq.push({gain[u], u});
    }
  };
  for (int32_t v = 0; v < n; ++v) {
    int64_t s = 0;
    for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q) s += G.ewgt[q];
    gain[v] = -s;
  }
  add(seed);
  int32_t next_unreached = 0;
  while (w0 < target) {
    int32_t pick = -1;
    while (!pq.empty()) {
      const Item it = pq.top();
      pq.pop();
      if (part[it.second] == 0 || it.first != gain[it.second]) continue;
      pick = it.second;
      break;
    }
    if (pick < 0) {  // disconnected: continue from an unreached vertex
      while (next_unreached < n && part[next_unreached] == 0) ++next_unreached;
      if (next_unreached >= n) break;
      pick = next_unreached;
    }
    add(pick);
  }
  return part;
}

// Fiduccia-Mattheyses refinement of a bisection (edge cut), balance bounded by
// max_w for each side. A few passes; each pass keeps its best prefix.
void fm_refine(const Graph& G, std::vector<uint8_t>& part, int64_t max_w, int passes) {
  const int32_t n = G.n;
  std::vector<int64_t> gain(n);
  std::vector<uint8_t> locked(n);
  int64_t w[2] = {0, 0};
  for (int32_t v = 0; v < n; ++v) w[part[v]] += G.vwgt[v];
  for (int pass = 0; pass < passes; ++pass) {
    std::fill(locked.begin(), locked.end(), 0);
    using Item = std::pair<int64_t, int32_t>;
    std::priority_queue<Item> pq[2];
    for (int32_t v = 0; v < n; ++v) {
      int64_t ext = 0, in = 0;
      bool boundary = false;
      for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
        if (part[G.adj[q]] != part[v]) { ext += G.ewgt[q]; boundary = true; }
        else in += G.ewgt[q];
      }
      gain[v] = ext - in;
      if (boundary) pq[part[v]].push({gain[v], v});
    }
    std::vector<int32_t> moved;
    int64_t cur = 0, best = 0;
    size_t best_len = 0;
    int64_t best_imb = std::max(w[0], w[1]);
    const size_t limit = std::max<size_t>(64, static_cast<size_t>(n) / 20);
    size_t since_best = 0;
    while (since_best < limit) {
      // move from the side whose move keeps the balance; prefer the heavier
      int32_t v = -1;
      int from = -1;
      for (int attempt = 0; attempt < 2 && v < 0; ++attempt) {
        const int side = attempt == 0 ? (w[0] >= w[1] ? 0 : 1) : (w[0] >= w[1] ? 1 : 0);
        while (!pq[side].empty()) {
          const Item it = pq[side].top();
          const int32_t u = it.second;
          if (locked[u] || part[u] != side || it.first != gain[u]) { pq[side].pop(); continue; }
          if (w[1 - side] + G.vwgt[u] > max_w) break;
          pq[side].pop();
          v = u;
          from = side;
          break;
        }
      }
      if (v < 0) break;
      locked[v] = 1;
      cur -= gain[v];  // cut decreases by the gain
      part[v] = static_cast<uint8_t>(1 - from);
      w[from] -= G.vwgt[v];
      w[1 - from] += G.vwgt[v];
      moved.push_back(v);
      for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
        const int32_t u = G.adj[q];
        if (locked[u]) continue;
        gain[u] += (part[u] == from) ? 2 * G.ewgt[q] : -2 * G.ewgt[q];
        pq[part[u]].push({gain[u], u});
      }
      const int64_t imb = std::max(w[0], w[1]);
      if (cur < best || (cur == best && imb < best_imb)) {
        best = cur;
        best_len = moved.size();
        best_imb = imb;
        since_best = 0;
      } else {
        ++since_best;
      }
    }
    for (size_t t = moved.size(); t > best_len; --t) {  // roll back past the best prefix
      const int32_t v = moved[t - 1];
      const int to = 1 - part[v];
      w[part[v]] -= G.vwgt[v];
      w[to] += G.vwgt[v];
      part[v] = static_cast<uint8_t>(to);
    }
    if (best_len == 0) break;
  }
}

// Multilevel bisection of G; part[v] in {0, 1}.
std::vector<uint8_t> bisect(const Graph& G, Rng& rng) {
  std::vector<Graph> levels;
  std::vector<std::vector<int32_t>> cmaps;
  const Graph* cur = &G;
  while (cur->n > 160) {
    std::vector<int32_t> cmap;
    Graph C = coarsen(*cur, cmap, rng);
    if (C.n > 0.92 * cur->n) break;
    cmaps.push_back(std::move(cmap));
    levels.push_back(std::move(C));
    cur = &levels.back();
  }
  const int64_t total = cur->total_vwgt();
  const int64_t max_w = static_cast<int64_t>(0.55 * static_cast<double>(total)) + 1;
  std::vector<uint8_t> best;
  int64_t best_cut = -1;
  const int tries = std::min<int32_t>(8, cur->n);
  for (int t = 0; t < tries; ++t) {
    std::vector<uint8_t> p = grow_partition(*cur, rng.below(cur->n), total / 2);
    fm_refine(*cur, p, max_w, 4);
    const int64_t c = edge_cut(*cur, p);
    if (best_cut < 0 || c < best_cut) { best_cut = c; best = std::move(p); }
  }
  for (int l = static_cast<int>(levels.size()) - 1; l >= 0; --l) {
    const Graph& fine = l == 0 ? G : levels[l - 1];
    std::vector<uint8_t> p(fine.n);
    for (int32_t v = 0; v < fine.n; ++v) p[v] = best[cmaps[l][v]];
    const int64_t tw = fine.total_vwgt();
    fm_refine(fine, p, static_cast<int64_t>(0.55 * static_cast<double>(tw)) + 1, 3);
    best = std::move(p);
  }
  return best;
}

// Minimum vertex cover of the bipartite graph of cut edges (Konig): maximum
// matching by augmenting paths, then the alternating-reachability cover.
std::vector<uint8_t> vertex_separator(const Graph& G, const std::vector<uint8_t>& part) {
  const int32_t n = G.n;
  std::vector<int32_t> mate(n, -1);
  std::vector<int32_t> left;  // boundary vertices of part 0
  for (int32_t v = 0; v < n; ++v) {
    if (part[v] != 0) continue;
    for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q)
      if (part[G.adj[q]] == 1) { left.push_back(v); break; }
  }
  std::vector<int32_t> seen(n, -1);
  int32_t stamp = 0;
  // iterative augmenting path search from each left vertex
  std::vector<std::pair<int32_t, int32_t>> stack;  // (left vertex, next edge)
  std::vector<int32_t> via(n, -1);
  for (int32_t root : left) {
    ++stamp;
    stack.clear();
    stack.push_back({root, G.xadj[root]});
    bool found = false;
    int32_t end_right = -1;
    while (!stack.empty() && !found) {
      auto& top = stack.back();
      const int32_t a = top.first;
      if (top.second >= G.xadj[a + 1]) { stack.pop_back(); continue; }
      const int32_t b = G.adj[top.second++];
      if (part[b] != 1 || seen[b] == stamp) continue;
      seen[b] = stamp;
      via[b] = a;
      if (mate[b] < 0) { found = true; end_right = b; break; }
      stack.push_back({mate[b], G.xadj[mate[b]]});
    }
    if (found) {  // flip the path
      int32_t b = end_right;
      while (b >= 0) {
        const int32_t a = via[b];
        const int32_t prev = mate[a];
        mate[a] = b;
        mate[b] = a;
        b = (a == root) ? -1 : prev;
      }
    }
  }
  // Z = vertices reachable from unmatched left vertices by alternating paths
  std::vector<uint8_t> inZ(n, 0);
  std::vector<int32_t> queue;
  for (int32_t a : left)
    if (mate[a] < 0) { inZ[a] = 1; queue.push_back(a); }
  for (size_t h = 0; h < queue.size(); ++h) {
    const int32_t a = queue[h];
    for (int32_t q = G.xadj[a]; q < G.xadj[a + 1]; ++q) {
      const int32_t b = G.adj[q];
      if (part[b] != 1 || inZ[b]) continue;
      inZ[b] = 1;
      const int32_t a2 = mate[b];
      if (a2 >= 0 && !inZ[a2]) { inZ[a2] = 1; queue.push_back(a2); }
    }
  }
  std::vector<uint8_t> sep(n, 0);
  for (int32_t a : left)
    if (!inZ[a]) sep[a] = 1;
  for (int32_t v = 0; v < n; ++v)
    if (part[v] == 1 && inZ[v]) {
      bool cut = false;
      for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q)
        if (part[G.adj[q]] == 0) { cut = true; break; }
      if (cut) sep[v] = 1;
    }
  return sep;
}

Graph induced(const Graph& G, const std::vector<int32_t>& verts, std::vector<int32_t>& local) {
  Graph S;
  S.n = static_cast<int32_t>(verts.size());
  for (int32_t i = 0; i < S.n; ++i) local[verts[i]] = i;
  S.xadj.assign(static_cast<size_t>(S.n) + 1, 0);
  S.vwgt.assign(S.n, 1);
  for (int32_t i = 0; i < S.n; ++i) {
    const int32_t v = verts[i];
    for (int32_t q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
      const int32_t u = local[G.adj[q]];
      if (u >= 0) {
        S.adj.push_back(u);
        S.ewgt.push_back(1);
      }
    }
    S.xadj[i + 1] = static_cast<int32_t>(S.adj.size());
  }
  for (int32_t v : verts) local[v] = -1;
  return S;
}

void order_leaf(const Graph& S, const std::vector<int32_t>& verts, std::vector<int32_t>& out) {
  CscUpper A;
  A.n = S.n;
  A.p.assign(static_cast<size_t>(S.n) + 1, 0);
  for (int32_t j = 0; j < S.n; ++j) {
    for (int32_t q = S.xadj[j]; q < S.xadj[j + 1]; ++q)
      if (S.adj[q] < j) A.i.push_back(S.adj[q]);
    std::sort(A.i.begin() + A.p[j], A.i.end());
    A.i.push_back(j);
    A.p[j + 1] = static_cast<int32_t>(A.i.size());
  }
  A.x.assign(A.i.size(), 0.0);
  const std::vector<int32_t> p = amd_order(A);
  for (int32_t k : p) out.push_back(verts[k]);
}

void dissect(const Graph& G, std::vector<int32_t> verts, std::vector<int32_t>& local, std::vector<int32_t>& out,
             int32_t leaf, Rng& rng, int depth) {
  Graph S = induced(G, verts, local);
  if (S.n <= leaf || depth > 60) {
    order_leaf(S, verts, out);
    return;
  }
  std::vector<uint8_t> part = bisect(S, rng);
  std::vector<uint8_t> sep = vertex_separator(S, part);
  std::vector<int32_t> A, B, C;
  for (int32_t i = 0; i < S.n; ++i) {
    if (sep[i]) C.push_back(verts[i]);
    else if (part[i] == 0) A.push_back(verts[i]);
    else B.push_back(verts[i]);
  }
  if (A.empty() || B.empty()) {  // no useful split
    order_leaf(S, verts, out);
    return;
  }
  std::vector<int32_t>().swap(verts);
  dissect(G, std::move(A), local, out, leaf, rng, depth + 1);
  dissect(G, std::move(B), local, out, leaf, rng, depth + 1);
  for (int32_t v : C) out.push_back(v);
}

}  // namespace

std::vector<int32_t> nd_order(const CscUpper& A, int32_t leaf_size) {
  const int32_t n = A.n;
  Graph G;
  G.n = n;
  std::vector<int32_t> deg(n, 0);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = A.p[j]; q < A.p[j + 1]; ++q)
      if (A.i[q] != j) { deg[A.i[q]]++; deg[j]++; }
  G.xadj.assign(static_cast<size_t>(n) + 1, 0);
  for (int32_t v = 0; v < n; ++v) G.xadj[v + 1] = G.xadj[v] + deg[v];
  G.adj.resize(G.xadj[n]);
  std::vector<int32_t> fill(G.xadj.begin(), G.xadj.end() - 1);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int32_t i = A.i[q];
      if (i == j) continue;
      G.adj[fill[i]++] = j;
      G.adj[fill[j]++] = i;
    }
  G.ewgt.assign(G.adj.size(), 1);
  G.vwgt.assign(n, 1);
  std::vector<int32_t> verts(n), local(n, -1), out;
  std::iota(verts.begin(), verts.end(), 0);
  out.reserve(n);
  Rng rng(0x5eed);
  dissect(G, std::move(verts), local, out, std::max(leaf_size, 8), rng, 0);
  if (static_cast<int32_t>(out.size()) != n) throw std::logic_error("nested dissection lost vertices");
  return out;
}

}  // namespace mpir
