// Symbolic phase of the FGPILU preconditioner (the ILU path of
// GmresLDLIR::Compute, include/MPIR/gmres_ir.hpp:365-445 of yex33/MPIR):
// the ILU_k pattern, the value map into the matrix, and the level sets of
// the two triangular solves.
#include "ilu.hpp"

#include <algorithm>
#include <stdexcept>

namespace mpir {

IluStructure ilu_structure(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t ilu_k) {
  if (n < 1) throw std::invalid_argument("matrix dimension must be positive");
  if (ilu_k < 0) throw std::invalid_argument("ILU level must be non-negative");
  IluStructure S;
  S.n = n;
  // the upper pattern of A, then ilu_k times: column col collects the rows of
  // A's columns row1 for every row1 of its current pattern (:369-408)
  S.Up.assign(Ap, Ap + n + 1);
  S.Ui.assign(Ai, Ai + Ap[n]);
  std::vector<char> marker(n, 0);
  std::vector<int32_t> rows;
  for (int32_t it = 0; it < ilu_k; ++it) {
    std::vector<int32_t> Np(static_cast<size_t>(n) + 1, 0), Ni;
    Ni.reserve(S.Ui.size() * 2);
    for (int32_t col = 0; col < n; ++col) {
      rows.clear();
      for (int32_t i1 = S.Up[col]; i1 < S.Up[col + 1]; ++i1) {
        const int32_t row1 = S.Ui[i1];
        if (row1 > col) continue;
        for (int32_t i2 = Ap[row1]; i2 < Ap[row1 + 1]; ++i2) {
          const int32_t row2 = Ai[i2];
          if (row2 > col) continue;
          if (!marker[row2]) {
            marker[row2] = 1;
            rows.push_back(row2);
          }
        }
      }
      std::sort(rows.begin(), rows.end());
      for (int32_t r : rows) {
        marker[r] = 0;
        Ni.push_back(r);
      }
      if (Ni.size() >= (size_t(1) << 31)) throw std::overflow_error("ILU pattern exceeds 2^31 entries");
      Np[col + 1] = static_cast<int32_t>(Ni.size());
    }
    S.Up.swap(Np);
    S.Ui.swap(Ni);
  }
  const int64_t nnz = S.Up[n];
  // per entry: its column, the matrix entry it starts from (ValueAt,
  // :826-836), the diagonal of each column
  S.col.resize(nnz);
  S.src.assign(nnz, -1);
  S.diag.assign(n, -1);
  for (int32_t col = 0; col < n; ++col) {
    int32_t q = Ap[col];
    for (int32_t idx = S.Up[col]; idx < S.Up[col + 1]; ++idx) {
      const int32_t row = S.Ui[idx];
      S.col[idx] = col;
      while (q < Ap[col + 1] && Ai[q] < row) ++q;
      if (q < Ap[col + 1] && Ai[q] == row) S.src[idx] = q;
      if (row == col) S.diag[col] = idx;
    }
    if (S.diag[col] < 0) throw std::invalid_argument("ILU pattern without a diagonal entry");
  }
  // U by rows: the entries (col > row, idx) of row `row`, columns ascending
  S.ur_ptr.assign(static_cast<size_t>(n) + 1, 0);
  for (int64_t idx = 0; idx < nnz; ++idx)
    if (S.Ui[idx] < S.col[idx]) S.ur_ptr[S.Ui[idx] + 1]++;
  for (int32_t r = 0; r < n; ++r) S.ur_ptr[r + 1] += S.ur_ptr[r];
  S.ur_col.resize(S.ur_ptr[n]);
  S.ur_idx.resize(S.ur_ptr[n]);
  {
    std::vector<int32_t> next(S.ur_ptr.begin(), S.ur_ptr.end() - 1);
    for (int32_t col = 0; col < n; ++col)
      for (int32_t idx = S.Up[col]; idx < S.Up[col + 1]; ++idx) {
        const int32_t row = S.Ui[idx];
        if (row >= col) continue;
        S.ur_col[next[row]] = col;
        S.ur_idx[next[row]++] = idx;
      }
  }
  // level sets: row r of L depends on the rows c < r of its entries (column
  // r of U); row r of U on the rows c > r of its entries
  std::vector<int32_t> lev(n, 0);
  int32_t nl = 0;
  for (int32_t r = 0; r < n; ++r) {
    int32_t l = 0;
    for (int32_t idx = S.Up[r]; idx < S.Up[r + 1]; ++idx)
      if (S.Ui[idx] < r) l = std::max(l, lev[S.Ui[idx]] + 1);
    lev[r] = l;
    nl = std::max(nl, l + 1);
  }
  auto bucket = [n](const std::vector<int32_t>& level, int32_t nlev, std::vector<int32_t>& ptr,
                    std::vector<int32_t>& rows_out) {
    ptr.assign(static_cast<size_t>(nlev) + 1, 0);
    for (int32_t r = 0; r < n; ++r) ptr[level[r] + 1]++;
    for (int32_t l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
    rows_out.resize(n);
    std::vector<int32_t> next(ptr.begin(), ptr.end() - 1);
    for (int32_t r = 0; r < n; ++r) rows_out[next[level[r]]++] = r;
  };
  bucket(lev, nl, S.lo_ptr, S.lo_rows);
  std::fill(lev.begin(), lev.end(), 0);
  nl = 0;
  for (int32_t r = n; r-- > 0;) {
    int32_t l = 0;
    for (int32_t t = S.ur_ptr[r]; t < S.ur_ptr[r + 1]; ++t) l = std::max(l, lev[S.ur_col[t]] + 1);
    lev[r] = l;
    nl = std::max(nl, l + 1);
  }
  bucket(lev, nl, S.up_ptr, S.up_rows);
  return S;
}

}  // namespace mpir
