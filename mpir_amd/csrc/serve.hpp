// Serving: a stream of solve requests on one device with continuous batching.
//
// Each request names a registered sparsity pattern and carries fresh values
// and 1..R right-hand sides. Requests are admitted in arrival order into
// batches by their device-memory footprint: a batch takes queued requests
// while the bytes of every batch in flight stay within the budget. Several
// lanes (a host thread and a HIP stream each, with their own solver handle)
// run batches concurrently, so requests that arrive while one batch is on
// the device are admitted by the next free lane and overlap with it on the
// device. The symbolic phase of a pattern is computed once (PatternCache)
// and reused by every request of that pattern. A request's result does not
// depend on the batch it lands in (see batch.hip).
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <vector>

#include "batch_plan.hpp"

namespace mpir {

struct ServeOptions {
  int uf = 1;                 // factor precision (kF32)
  int lanes = 2;              // concurrent batches
  int64_t budget_bytes = int64_t(64) << 30;  // device bytes of all batches in flight
  int max_batch = 64;         // requests per batch
  double tol = 1e-10;
  int ordering = 2;           // nested dissection
  int max_ir = 10, max_gmres = 10;
  int keep_x = 0;             // keep every solution on the host
  int threads = 8;            // host threads preparing request data
};

// A registered pattern with its base values. Request values are derived
// from the base values entry by entry (see materialize).
struct ServePattern {
  std::shared_ptr<const PatternEntry> entry;
  int32_t n = 0, n1 = 0;
  int64_t nnz = 0;
  std::vector<double> base;
  std::vector<uint8_t> cls;  // 0: diagonal of H, 1: off-diagonal of H, 2: A, 3: diagonal of -D
  std::vector<int32_t> col;  // column of each entry
};

// Request data from (pattern, seed), deterministic: values, nrhs
// right-hand sides b = K x_true (n x nrhs, column-major), x_true ~ U(-1,1).
// Values: off-diagonals of H and entries of A times (1 + 0.05 u), u ~ U(-1,1);
// -D times (1 + 0.05 u); the diagonal of H times (1.05 + 0.05 v), v ~ U(0,1),
// which keeps H strictly diagonally dominant (Sum |h_ij| <= 1.05 Sum |base|
// < 1.05 h_ii) and K quasi-definite.
void materialize(const ServePattern& p, uint64_t seed, int nrhs, double* Ax, double* b, double* x_true);

struct ServeRecord {
  int32_t pattern = 0;
  uint64_t seed = 0;
  int32_t nrhs = 1;
  double t_arrive = 0.0;  // ms after the start of the run (input)
  // outputs (ms after the start of the run)
  double t_admit = 0.0, t_done = 0.0;
  int32_t batch = -1, lane = -1, status = 0;
  int32_t ir_max = 0, gmres_max = 0;
  double rel_residual_max = 0.0;
  int64_t bytes = 0;  // footprint charged at admission
};

struct ServeSummary {
  int64_t completed = 0, failed = 0, batches = 0;
  int64_t max_backlog = 0;          // arrived, not yet admitted
  int64_t peak_inflight_bytes = 0;
  int64_t max_concurrent = 0;       // batches on the device at once
  double wall_ms = 0.0;
  std::vector<double> lane_busy_ms;
  // time of every batch phase, summed over batches: request data, joined
  // symbolic plan and its upload, numeric factorization, GMRES-IR
  double prepare_ms = 0.0, plan_ms = 0.0, factor_ms = 0.0, solve_ms = 0.0;
};

class GmresLdl;

class Server {
 public:
  explicit Server(const ServeOptions& opt);
  ~Server();
  // symbolic phase now (from the cache when the pattern is known); returns
  // the pattern id
  int32_t add_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* base, int32_t n1);
  // many patterns, analysed on `threads` host threads; returns the first id
  int32_t add_patterns(int32_t count, const int32_t* n, const int32_t* const* Ap, const int32_t* const* Ai,
                       const double* const* base, const int32_t* n1, int threads);
  const ServePattern& pattern(int32_t id) const;
  int32_t npatterns() const { return static_cast<int32_t>(patterns_.size()); }
  // Runs a stream of requests (t_arrive ascending) to completion.
  void run(std::vector<ServeRecord>& reqs);
  const ServeSummary& summary() const { return sum_; }
  // solution of request i of the last run (keep_x)
  const std::vector<double>& solution(int64_t i) const;
  // bytes a batch of these requests needs on the device
  int64_t batch_bytes(const std::vector<const ServeRecord*>& rs) const;

 private:
  void add_entry(std::shared_ptr<const PatternEntry> entry, int32_t n, const int32_t* Ap, const int32_t* Ai,
                 const double* base, int32_t n1);
  void lane_main(int lane, std::vector<ServeRecord>* reqs);
  void process(int lane, std::vector<ServeRecord>* reqs, const std::vector<int64_t>& ids, int32_t batch_id);

  ServeOptions opt_;
  PatternCache cache_;
  std::vector<ServePattern> patterns_;
  std::vector<std::unique_ptr<GmresLdl>> solvers_;
  std::vector<std::vector<double>> x_;
  // run state
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<int64_t> queue_;
  int64_t next_ = 0, outstanding_ = 0;
  int64_t inflight_bytes_ = 0, concurrent_ = 0;
  int32_t batch_seq_ = 0;
  double t0_ = 0.0;
  ServeSummary sum_;
};

}  // namespace mpir
