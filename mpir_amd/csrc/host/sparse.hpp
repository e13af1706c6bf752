// Host-side sparse containers and symbolic analysis for quasi-definite
// K = [H A^T; A -D] systems stored as the upper triangle in CSC form, the
// input convention of QDLDL (include/MPIR/qdldl.hpp:25-31 in yex33/MPIR).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mpir {

// Upper-triangular CSC: column j holds rows i <= j, rows ascending, no
// duplicates, diagonal present (the QDLDL input contract).
struct CscUpper {
  int32_t n = 0;
  std::vector<int32_t> p;  // n+1
  std::vector<int32_t> i;  // nnz
  std::vector<double> x;   // nnz
  int64_t nnz() const { return p.empty() ? 0 : p[n]; }
};

// Full symmetric CSR (both triangles), columns ascending per row.
struct CsrFull {
  int32_t n = 0;
  std::vector<int32_t> p;
  std::vector<int32_t> j;
  std::vector<double> x;
  // For each CSR entry, the index of the triu CSC entry it mirrors; lets
  // fresh values be scattered without rebuilding the pattern.
  std::vector<int32_t> src;
};

// Parameters of the synthetic KKT generator (grid Laplacian H plus random
// constraint block A). Exact definitions live in generators.cpp.
struct KktSpec {
  int dim = 2;           // 2 or 3 (grid dimension of H)
  int k = 32;            // grid points per side
  double h_shift = 1e-2; // H = Laplacian + h_shift * I
  int64_t m = 256;       // constraint rows
  int a_mode = 0;        // 0: cnt_lo..cnt_hi distinct uniform columns per row
                         // 1: one uniform column plus its +x,+y,+z grid
                         //    neighbours (reflected at the upper face)
  int cnt_lo = 3, cnt_hi = 3;
  double d_value = 1e-2;  // D = d_value * I
  uint64_t seed = 0;
};

struct KktProblem {
  CscUpper K;
  int32_t n1 = 0;  // size of the H block
  std::vector<double> x_true;
  std::vector<double> b;
};

KktProblem generate_kkt(const KktSpec& spec);

// Symmetric CSR of the full matrix from its upper CSC.
CsrFull csr_from_upper(const CscUpper& A);

// b = K x with the accumulation order of a row-wise ascending-column sweep.
void spmv_full(const CsrFull& A, const double* x, double* y);

// ---------------------------------------------------------------- ordering
// Approximate minimum degree ordering of the symmetric pattern of A (upper
// CSC, diagonal ignored). perm[k] = original index of the k-th pivot.
std::vector<int32_t> amd_order(const CscUpper& A, int64_t* stats_nnz_est = nullptr);

// Nested dissection (multilevel bisection, vertex separators, AMD on
// subgraphs of at most leaf_size vertices). perm[k] = original index.
std::vector<int32_t> nd_order(const CscUpper& A, int32_t leaf_size = 200);

// C = triu(P A P^T) with perm[k] = old index of new index k. val_map[q] = the
// index into A's arrays that entry q of C takes its value from.
CscUpper permute_upper(const CscUpper& A, const std::vector<int32_t>& perm,
                       std::vector<int32_t>* val_map);

// ---------------------------------------------------------------- symbolic
// Elimination tree and column counts with QDLDL_etree's contract
// (include/MPIR/qdldl.hpp:48-105): returns sum(Lnz), -1 for bad input
// (empty column or entry below the diagonal), -2 on int32 overflow.
// The int32 overflow check applies when check_i32 is set.
int64_t etree_counts(int32_t n, const int32_t* Ap, const int32_t* Ai,
                     int32_t* work, int32_t* Lnz, int32_t* parent, bool check_i32 = true);

// Postorder of a forest given parent pointers (children visited in
// ascending index order). post[k] = node placed at position k.
std::vector<int32_t> tree_postorder(int32_t n, const int32_t* parent);

// Row pattern of L (CSC, rows ascending per column) from the etree.
void l_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai,
               const int32_t* parent, const int32_t* Lnz,
               std::vector<int64_t>& Lp, std::vector<int32_t>& Li);

// Supernodal partition and the multifrontal schedule built on it.
// A Matrix Market file read as the reference's driver reads it
// (examples/example.cpp:87-138): triplets transposed, sorted by (column,
// row), compressed to CSC; for a symmetric file stored by its lower
// triangle this is the upper-triangular CSC.
struct MtxUpper {
  int32_t n = 0;
  bool symmetric = false;
  std::vector<int32_t> p, i;
  std::vector<double> x;
};
MtxUpper read_mtx_upper(const std::string& path);

// Fronts with at most this many rows are factored and solved by one
// wavefront in LDS; larger ("big") fronts use the blocked kernels.
constexpr int kSmallFront = 64;
// Column-block width of the blocked factorization of big fronts.
constexpr int kPanelNB = 64;

// Leading dimension of a supernode's panel. A big front (m = nc + nr rows)
// carries nc extra rows below R_s, initialised to the identity, which the
// blocked factorization turns into L_11^{-T} D^{-1} of its diagonal block:
// the triangular solves then become plain matrix-vector products.
constexpr int64_t panel_ld(int64_t nc, int64_t nr) { return nc + nr > kSmallFront ? 2 * nc + nr : nc + nr; }

// Position of L[r, c] (front-local row r > column c) in a panel. Big fronts
// keep the strictly-lower part of each kPanelNB-wide diagonal block transposed in
// the block's upper triangle (their factorization writes it there while the
// lower triangle is still being read).
constexpr int64_t panel_l_index(int64_t nc, int64_t nr, int64_t r, int64_t c) {
  return (nc + nr > kSmallFront && r < nc && r / kPanelNB == c / kPanelNB) ? c + r * panel_ld(nc, nr)
                                                                  : r + c * panel_ld(nc, nr);
}

struct Supernodes {
  int32_t n = 0;
  int32_t ns = 0;                   // number of supernodes
  std::vector<int32_t> first;       // ns+1, column ranges [first[s], first[s+1])
  std::vector<int32_t> nrow;        // rows strictly below the diagonal block
  std::vector<int64_t> rows_off;    // ns+1, offsets into rows
  std::vector<int32_t> rows;        // R_s, ascending, all >= first[s+1]
  std::vector<int32_t> sparent;     // parent supernode or -1
  std::vector<int32_t> col2sn;      // n
  std::vector<int32_t> level;       // height above the leaves
  int32_t nlevels = 0;
  std::vector<int32_t> level_ptr;   // nlevels+1 into level_sn
  std::vector<int32_t> level_sn;    // supernodes grouped by level
  std::vector<int64_t> panel_off;   // ns+1, panel storage (panel_ld * ncol_s)
  std::vector<int32_t> relidx;      // per supernode, position of each R_s row
                                    // inside the parent's front index list
  // statistics
  int64_t nnz_l = 0;      // strictly-lower entries of L
  double flops = 0.0;     // factor flops, sum Lnz_j (Lnz_j + 2) + n
  int32_t max_front = 0;
};

Supernodes build_supernodes(const CscUpper& C, const int32_t* parent, const int32_t* Lnz);
// Supernodes of a given factor pattern (CSC, rows ascending, strictly lower),
// e.g. the output of QDLDL_factor. Throws std::invalid_argument /
// std::logic_error if the pattern is not closed under its elimination tree.
Supernodes supernodes_from_l(int32_t n, const int64_t* Lp, const int32_t* Li);
void finish_supernodes(Supernodes& S);

// Relaxed supernode partition of a postordered pattern: order[k] = column of
// C placed at position k (a topological reordering of the elimination tree),
// first = block starts in the new order.
struct RelaxedPartition {
  std::vector<int32_t> order;
  std::vector<int32_t> first;
  std::vector<std::vector<int32_t>> pending;  // work space of the traversal
};
// level 1 .. 3: how freely a child is merged into its parent
// (larger fronts and a shallower tree for more explicit zeros).
RelaxedPartition relax_supernodes(const CscUpper& C, const int32_t* parent, const int32_t* Lnz, int level = 1);
// Supernodes for a given partition; R_s is the union structure.
Supernodes build_supernodes_forced(const CscUpper& C, const int32_t* parent, const std::vector<int32_t>& first);

}  // namespace mpir
