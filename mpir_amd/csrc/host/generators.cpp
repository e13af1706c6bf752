// Deterministic synthetic quasi-definite KKT systems K = [H A^T; A -D].
//
// The reference ships no generator (it reads SuiteSparse .mtx files,
// examples/example.cpp:87-138, and builds b from x_ref = 1 at :153-158); the
// workloads here are defined from scratch:
//   H  = grid Laplacian (2D 5-point or 3D 7-point, Dirichlet, lexicographic,
//        x fastest; diagonal 2*dim, off-diagonals -1) + h_shift * I
//   A  = m x n1, a_mode 0: cnt distinct uniform columns per row, values N(0,1)
//               a_mode 1: one uniform column c plus its +x, +y, +z grid
//                         neighbours (-x/-y/-z at the upper face), N(0,1)
//               a_mode 2: cnt in cnt_lo..cnt_hi (uniform); one uniform
//                         column c plus the first cnt-1 cells of the offset
//                         list +x, +y, (+z), -x, -y, (-z), +x+y, -x-y, +x-y,
//                         -x+y, +2x, -2x that lie inside the grid, N(0,1)
//   D  = d_value * I
//   x_true ~ U(-1, 1),  b = K x_true (FP64, row-wise ascending-column sums)
// Random numbers: SplitMix64 (Steele, Lea & Flood 2014), 53-bit uniforms,
// normals by the Marsaglia polar method. All draws come from one stream in
// this order: A rows (count, columns, values), then x_true.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "sparse.hpp"

namespace mpir {

namespace {

struct SplitMix64 {
  uint64_t s;
  bool have_spare = false;
  double spare = 0.0;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
  int64_t below(int64_t n) {
    int64_t v = static_cast<int64_t>(uniform() * static_cast<double>(n));
    return v < n ? v : n - 1;
  }
  double normal() {
    if (have_spare) {
      have_spare = false;
      return spare;
    }
    double u, v, q;
    do {
      u = 2.0 * uniform() - 1.0;
      v = 2.0 * uniform() - 1.0;
      q = u * u + v * v;
    } while (q >= 1.0 || q == 0.0);
    const double f = std::sqrt(-2.0 * std::log(q) / q);
    spare = v * f;
    have_spare = true;
    return u * f;
  }
};

}  // namespace

KktProblem generate_kkt(const KktSpec& spec) {
  if (spec.dim != 2 && spec.dim != 3) throw std::invalid_argument("dim must be 2 or 3");
  if (spec.k < 2) throw std::invalid_argument("grid side must be >= 2");
  const int64_t k = spec.k;
  const int64_t n1 = spec.dim == 2 ? k * k : k * k * k;
  const int64_t m = spec.m;
  const int64_t n = n1 + m;
  if (n >= (int64_t(1) << 31)) throw std::invalid_argument("system too large for int32 indices");
  if (spec.a_mode == 0 && (spec.cnt_lo < 1 || spec.cnt_hi < spec.cnt_lo || spec.cnt_hi > n1))
    throw std::invalid_argument("bad constraint-row count range");
  if (spec.a_mode == 1 && spec.dim != 3) throw std::invalid_argument("a_mode 1 needs a 3D grid");
  if (spec.a_mode == 2 && (spec.cnt_lo < 1 || spec.cnt_hi < spec.cnt_lo || spec.cnt_hi > 5 || k < 3))
    throw std::invalid_argument("a_mode 2 needs 1 <= cnt_lo <= cnt_hi <= 5 and a grid side >= 3");
  if (spec.a_mode < 0 || spec.a_mode > 2) throw std::invalid_argument("a_mode must be 0, 1 or 2");

  SplitMix64 rng(spec.seed);

  // constraint rows: (sorted columns, values)
  std::vector<int64_t> arow_ptr(m + 1, 0);
  std::vector<int32_t> acol;
  std::vector<double> aval;
  acol.reserve(static_cast<size_t>(m) * 5);
  aval.reserve(static_cast<size_t>(m) * 5);
  std::vector<int64_t> cols;
  std::vector<double> vals;
  for (int64_t r = 0; r < m; ++r) {
    cols.clear();
    vals.clear();
    if (spec.a_mode == 0) {
      int cnt = spec.cnt_lo;
      if (spec.cnt_hi > spec.cnt_lo) cnt += static_cast<int>(rng.below(spec.cnt_hi - spec.cnt_lo + 1));
      while (static_cast<int>(cols.size()) < cnt) {
        int64_t c = rng.below(n1);
        if (std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
      }
    } else if (spec.a_mode == 2) {
      int cnt = spec.cnt_lo;
      if (spec.cnt_hi > spec.cnt_lo) cnt += static_cast<int>(rng.below(spec.cnt_hi - spec.cnt_lo + 1));
      const int64_t c = rng.below(n1);
      const int64_t x = c % k, y = (c / k) % k, z = spec.dim == 3 ? c / (k * k) : 0;
      static const int off2[][3] = {{1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0}, {1, 1, 0}, {-1, -1, 0},
                                    {1, -1, 0}, {-1, 1, 0}, {2, 0, 0}, {-2, 0, 0}};
      static const int off3[][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
                                    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {2, 0, 0}, {-2, 0, 0}};
      const int (*off)[3] = spec.dim == 3 ? off3 : off2;
      const int noff = spec.dim == 3 ? 12 : 10;
      cols.push_back(c);
      for (int t = 0; t < noff && static_cast<int>(cols.size()) < cnt; ++t) {
        const int64_t xx = x + off[t][0], yy = y + off[t][1], zz = z + off[t][2];
        if (xx < 0 || xx >= k || yy < 0 || yy >= k || zz < 0 || zz >= (spec.dim == 3 ? k : 1)) continue;
        cols.push_back(xx + k * (yy + k * zz));
      }
      if (static_cast<int>(cols.size()) < cnt) throw std::logic_error("a_mode 2: too few grid neighbours");
    } else {
      const int64_t c = rng.below(n1);
      const int64_t x = c % k, y = (c / k) % k, z = c / (k * k);
      cols.push_back(c);
      cols.push_back(x + 1 < k ? c + 1 : c - 1);
      cols.push_back(y + 1 < k ? c + k : c - k);
      cols.push_back(z + 1 < k ? c + k * k : c - k * k);
    }
    for (size_t t = 0; t < cols.size(); ++t) vals.push_back(rng.normal());
    // sort columns, carrying values
    std::vector<size_t> ord(cols.size());
    for (size_t t = 0; t < ord.size(); ++t) ord[t] = t;
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return cols[a] < cols[b]; });
    for (size_t t : ord) {
      acol.push_back(static_cast<int32_t>(cols[t]));
      aval.push_back(vals[t]);
    }
    arow_ptr[r + 1] = static_cast<int64_t>(acol.size());
  }

  KktProblem P;
  P.n1 = static_cast<int32_t>(n1);
  CscUpper& K = P.K;
  K.n = static_cast<int32_t>(n);
  K.p.assign(n + 1, 0);
  K.i.reserve(static_cast<size_t>(n1) * (spec.dim + 1) + acol.size() + m);
  K.x.reserve(K.i.capacity());
  const double diag = 2.0 * spec.dim + spec.h_shift;
  for (int64_t j = 0; j < n1; ++j) {
    const int64_t x = j % k, y = (j / k) % k, z = j / (k * k);
    if (spec.dim == 3 && z > 0) { K.i.push_back(static_cast<int32_t>(j - k * k)); K.x.push_back(-1.0); }
    if (y > 0) { K.i.push_back(static_cast<int32_t>(j - k)); K.x.push_back(-1.0); }
    if (x > 0) { K.i.push_back(static_cast<int32_t>(j - 1)); K.x.push_back(-1.0); }
    K.i.push_back(static_cast<int32_t>(j));
    K.x.push_back(diag);
    K.p[j + 1] = static_cast<int32_t>(K.i.size());
  }
  for (int64_t r = 0; r < m; ++r) {
    for (int64_t q = arow_ptr[r]; q < arow_ptr[r + 1]; ++q) {
      K.i.push_back(acol[q]);
      K.x.push_back(aval[q]);
    }
    K.i.push_back(static_cast<int32_t>(n1 + r));
    K.x.push_back(-spec.d_value);
    K.p[n1 + r + 1] = static_cast<int32_t>(K.i.size());
  }

  P.x_true.resize(n);
  for (int64_t i = 0; i < n; ++i) P.x_true[i] = 2.0 * rng.uniform() - 1.0;
  P.b.assign(n, 0.0);
  CsrFull F = csr_from_upper(K);
  spmv_full(F, P.x_true.data(), P.b.data());
  return P;
}

CsrFull csr_from_upper(const CscUpper& A) {
  const int32_t n = A.n;
  CsrFull F;
  F.n = n;
  F.p.assign(static_cast<size_t>(n) + 1, 0);
  // row i of the full matrix = column i of the upper part (rows <= i) plus
  // row i of the upper part (columns > i).
  for (int32_t j = 0; j < n; ++j) {
    for (int32_t q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int32_t i = A.i[q];
      F.p[j + 1]++;
      if (i != j) F.p[i + 1]++;
    }
  }
  for (int32_t i = 0; i < n; ++i) F.p[i + 1] += F.p[i];
  const int64_t nnz = F.p[n];
  // a pattern-only input (empty x) gives a pattern-only CSR
  const bool values = !A.x.empty();
  F.j.resize(nnz);
  F.x.resize(values ? nnz : 0);
  F.src.resize(nnz);
  std::vector<int32_t> next(F.p.begin(), F.p.end() - 1);
  // Row r of the full matrix = column r of the upper part (columns <= r, in
  // ascending order, diagonal last) followed by row r of the upper part
  // (columns > r). Pass 1 writes every lower part, pass 2 appends the upper
  // parts while sweeping columns in ascending order, so rows stay sorted.
  for (int32_t r = 0; r < n; ++r) {
    for (int32_t q = A.p[r]; q < A.p[r + 1]; ++q) {  // column r: rows <= r
      const int32_t pos = next[r]++;
      F.j[pos] = A.i[q];
      if (values) F.x[pos] = A.x[q];
      F.src[pos] = q;
    }
  }
  for (int32_t c = 0; c < n; ++c) {  // entries (r, c), r < c, go to row r
    for (int32_t q = A.p[c]; q < A.p[c + 1]; ++q) {
      const int32_t r = A.i[q];
      if (r == c) continue;
      const int32_t pos = next[r]++;
      F.j[pos] = c;
      if (values) F.x[pos] = A.x[q];
      F.src[pos] = q;
    }
  }
  return F;
}

void spmv_full(const CsrFull& A, const double* x, double* y) {
  for (int32_t r = 0; r < A.n; ++r) {
    double s = 0.0;
    for (int32_t q = A.p[r]; q < A.p[r + 1]; ++q) s += A.x[q] * x[A.j[q]];
    y[r] = s;
  }
}

}  // namespace mpir
