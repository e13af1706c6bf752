// Many independent systems as one block-diagonal system.
//
// The symbolic phase of a sparsity pattern (ordering, elimination tree,
// supernodes, value map, SpMV pattern) is computed once and kept in a
// PatternCache. A batch of systems is the block-diagonal matrix
// diag(K_1, ..., K_S): its elimination tree is the forest of the systems'
// trees, so one factorization plan over that forest factors every system
// with one launch per level and step, and one dependency-driven launch per
// direction solves them all. Every front is computed exactly as it is when
// its system is alone, so per-system results do not depend on the batch.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "plan.hpp"

namespace mpir {

// The symbolic phase of one pattern.
struct PatternEntry {
  uint64_t key = 0;
  int ordering = 0;
  int relax = kDefaultAmalgamation;
  int32_t n = 0;
  int64_t nnz = 0;
  std::vector<int32_t> Ap, Ai;  // the pattern itself (upper CSC), for equality
  Analysis an;                  // perm (new -> old), C, val_map, etree, plan
  // full CSR of P K P^T: row pointers, columns, source entry of the
  // original upper CSC
  std::vector<int32_t> csr_p, csr_j, csr_src;
  std::vector<int32_t> ecol;      // column of each original entry
  std::vector<int32_t> diag_pos;  // original entry of each diagonal
  int64_t est_bytes = 0;          // device bytes of one system (see footprint)
};

PatternEntry analyze_entry(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, int relax);

// Device bytes a solver needs for one system of this pattern with the given
// factor element size, right-hand sides and GMRES cycle length.
int64_t pattern_footprint(const PatternEntry& e, int factor_bytes, int nrhs, int max_gmres);

// Thread-safe map pattern -> symbolic phase (keyed by a hash of n, Ap, Ai,
// ordering and relaxation; equal keys are checked entry by entry).
class PatternCache {
 public:
  std::shared_ptr<const PatternEntry> get(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, int relax);
  // entries for many patterns, analysing the missing ones on `threads`
  // host threads
  std::vector<std::shared_ptr<const PatternEntry>> get_many(const std::vector<int32_t>& n,
                                                            const std::vector<const int32_t*>& Ap,
                                                            const std::vector<const int32_t*>& Ai, int ordering,
                                                            int relax, int threads);
  int64_t hits() const { return hits_; }
  int64_t misses() const { return misses_; }
  size_t size() const;

 private:
  std::shared_ptr<const PatternEntry> find_locked(uint64_t key, int32_t n, const int32_t* Ap, const int32_t* Ai,
                                                  int ordering, int relax) const;
  mutable std::mutex mu_;
  std::unordered_multimap<uint64_t, std::shared_ptr<const PatternEntry>> map_;
  int64_t hits_ = 0, misses_ = 0;
};

uint64_t pattern_key(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, int relax);

// The block-diagonal system of a batch.
struct BatchSystem {
  int32_t nsys = 0;
  std::vector<int32_t> col_base;    // nsys+1: first row of each system
  std::vector<int64_t> ent_base;    // nsys+1: first original entry
  std::vector<int32_t> sys_of_row;  // system of each joined row
  std::vector<int32_t> local_perm;  // joined new row -> old row within its system
  std::vector<int32_t> perm;        // joined new row -> joined old row
  std::vector<int32_t> erow, ecol;  // row and column of every original entry (joined numbering)
  std::vector<int32_t> diag_pos;    // original entry of every diagonal
  std::vector<int32_t> csr_p, csr_j, csr_src;  // full CSR of P K P^T, source entries
  int64_t nnz_l = 0;  // nonzeros of L (without relaxation zeros)
  double flops = 0.0;
  bool relaxed = true;
};

// Joins the systems into B and their plans into plan (the forest of their
// elimination trees, see merge_plans); both keep their vectors' memory
// from one batch to the next.
void combine_patterns(const std::vector<const PatternEntry*>& parts, BatchSystem& B, FactorPlan& plan);

// The plan of the forest of the systems' elimination trees from the
// systems' own plans: supernodes, schedules and storage offsets shifted by
// the systems before each one; per level and step the systems' tasks follow
// each other (for the panel steps: every system's real rows, then every
// system's identity rows). Each system's update matrices keep their own
// region of the pool. No value map.
void merge_plans(const std::vector<const FactorPlan*>& plans, const std::vector<int32_t>& col_base, FactorPlan& M);

}  // namespace mpir
