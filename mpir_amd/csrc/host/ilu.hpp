// Symbolic phase of the FGPILU preconditioner (the ILU path of the
// reference's GmresLDLIR::Compute): ILU_k pattern, value map, level sets.
#pragma once

#include <cstdint>
#include <vector>

namespace mpir {

// U is held by columns (Up, Ui: the upper triangle, rows ascending); L by
// rows with the same index arrays: row r of L has its entries at the
// positions of column r of U, mirrored (L[r, c] at the index of U[c, r]).
struct IluStructure {
  int32_t n = 0;
  std::vector<int32_t> Up, Ui;        // n+1, nnz
  std::vector<int32_t> col;           // column of each entry
  std::vector<int32_t> src;           // the matrix entry of each pattern entry (upper CSC index), or -1
  std::vector<int32_t> diag;          // index of U[c, c]
  std::vector<int32_t> ur_ptr, ur_col, ur_idx;  // U by rows: strictly-upper entries (column, index)
  std::vector<int32_t> lo_ptr, lo_rows;  // level sets of the L solve (rows ascending within a level)
  std::vector<int32_t> up_ptr, up_rows;  // level sets of the U solve
};

// pattern of ILU(ilu_k) as the reference's Compute builds it, from the upper
// CSC pattern of A (rows ascending, diagonal present)
IluStructure ilu_structure(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t ilu_k);

}  // namespace mpir
