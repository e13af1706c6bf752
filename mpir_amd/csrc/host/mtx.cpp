// Matrix Market input as the reference's driver reads it
// (examples/example.cpp:87-138 of yex33/MPIR): the coordinate triplets of a
// symmetric matrix stored by its lower triangle are read transposed (the
// file's row becomes the column), without expanding the symmetry, sorted by
// (column, row) and compressed into an upper-triangular CSC.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sparse.hpp"

namespace mpir {

MtxUpper read_mtx_upper(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument("cannot open " + path);
  std::string line;
  if (!std::getline(in, line)) throw std::invalid_argument(path + ": empty file");
  std::string banner, object, format, field, symmetry;
  {
    std::istringstream hs(line);
    hs >> banner >> object >> format >> field >> symmetry;
  }
  auto lower = [](std::string v) {
    std::transform(v.begin(), v.end(), v.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
    return v;
  };
  if (banner != "%%MatrixMarket" || lower(object) != "matrix" || lower(format) != "coordinate")
    throw std::invalid_argument(path + ": not a Matrix Market coordinate matrix");
  field = lower(field);
  if (field != "real" && field != "integer" && field != "double")
    throw std::invalid_argument(path + ": only real / integer values are supported");
  MtxUpper M;
  M.symmetric = lower(symmetry) == "symmetric";
  while (std::getline(in, line))
    if (!line.empty() && line[0] != '%') break;
  int64_t nrows = 0, ncols = 0, nnz = 0;
  {
    std::istringstream ss(line);
    if (!(ss >> nrows >> ncols >> nnz) || nrows < 1 || nrows != ncols || nnz < 0)
      throw std::invalid_argument(path + ": bad size line (square matrix expected)");
  }
  if (nrows > INT32_MAX || nnz > INT32_MAX) throw std::invalid_argument(path + ": too large for int32 indices");
  // transposed read: the file's (i, j) becomes row j, column i
  std::vector<int32_t> rs(nnz), cs(nnz);
  std::vector<double> vs(nnz);
  for (int64_t q = 0; q < nnz; ++q) {
    int64_t i, j;
    double v;
    if (!(in >> i >> j >> v)) throw std::invalid_argument(path + ": truncated entry list");
    if (i < 1 || i > nrows || j < 1 || j > ncols) throw std::invalid_argument(path + ": index out of range");
    cs[q] = static_cast<int32_t>(i - 1);
    rs[q] = static_cast<int32_t>(j - 1);
    vs[q] = v;
  }
  std::vector<int64_t> order(nnz);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return cs[a] != cs[b] ? cs[a] < cs[b] : rs[a] < rs[b];
  });
  M.n = static_cast<int32_t>(nrows);
  M.p.assign(static_cast<size_t>(nrows) + 1, 0);
  M.i.resize(nnz);
  M.x.resize(nnz);
  for (int64_t q = 0; q < nnz; ++q) {
    const int64_t s = order[q];
    M.i[q] = rs[s];
    M.x[q] = vs[s];
    M.p[cs[s] + 1]++;
  }
  for (int32_t c = 0; c < M.n; ++c) M.p[c + 1] += M.p[c];
  return M;
}

}  // namespace mpir
