// Serving: solve requests on one device with continuous batching.
//
// Each request names a registered sparsity pattern and carries fresh values
// and 1..R right-hand sides, supplied by the caller (submit) or derived from
// a seed (submit_seeded, for benchmark streams). Requests are admitted in
// arrival order into batches by their device-memory footprint: a batch takes
// queued requests while the bytes of every batch in flight stay within the
// budget. Several lanes (a host thread and a HIP stream each, with their own
// solver handle) run for the server's lifetime and take batches whenever
// they are free, so requests that arrive while one batch is on the device
// are admitted by the next free lane and overlap with it on the device. The
// symbolic phase of a pattern is computed once (PatternCache) and reused by
// every request of that pattern. A request's result does not depend on the
// batch it lands in (see batch.hip). Callers poll or wait for a request by
// its id; the reference splits the same work into Compute (values) and
// Solve (right-hand sides), gmres_ir.hpp:271-285.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
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

// One request inside the server (times in ms on the server's clock).
struct ServeRequest {
  int32_t pattern = 0, nrhs = 1;
  bool seeded = false;
  uint64_t seed = 0;
  std::vector<double> ax, b;  // the caller's values and right-hand sides (freed once solved)
  std::vector<double> x;      // n x nrhs, column-major, once solved
  double t_arrive = 0.0, t_admit = 0.0, t_done = 0.0;
  int32_t batch = -1, lane = -1, status = 0;
  int32_t ir_max = 0, gmres_max = 0;
  double rel_residual_max = 0.0;
  int64_t bytes = 0;  // footprint charged at admission
  bool done = false;
};

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
  double plan_join_ms = 0.0, plan_upload_ms = 0.0;  // the two parts of plan_ms
  // device buffers of the process that had to grow during the run (each
  // such reallocation waits for the whole device)
  int64_t regrowths = 0;
};

class GmresLdl;

class Server {
 public:
  explicit Server(const ServeOptions& opt);
  ~Server();  // finishes the requests in flight, then stops the lanes
  // symbolic phase now (from the cache when the pattern is known); returns
  // the pattern id
  int32_t add_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* base, int32_t n1);
  // many patterns, analysed on `threads` host threads; returns the first id
  int32_t add_patterns(int32_t count, const int32_t* n, const int32_t* const* Ap, const int32_t* const* Ai,
                       const double* const* base, const int32_t* n1, int threads);
  const ServePattern& pattern(int32_t id) const;
  int32_t npatterns() const;
  // A request with the caller's values (nnz of the pattern, upper CSC order)
  // and nrhs right-hand sides (n x nrhs, column-major), copied; returns its id.
  int64_t submit(int32_t pattern, const double* Ax, const double* b, int32_t nrhs);
  // a request whose data materialize() derives from the seed in the lane
  int64_t submit_seeded(int32_t pattern, uint64_t seed, int32_t nrhs);
  // Waits up to timeout_ms (< 0: no limit) for request id. Returns false if
  // it is still pending; else copies its record (and x when non-null) out
  // and, with release, forgets the request.
  bool wait(int64_t id, double timeout_ms, ServeRequest* info, double* x, bool release);
  // Runs a stream of requests (t_arrive ascending, ms after the call) to
  // completion: each is submitted (seeded) at its arrival time.
  void run(std::vector<ServeRecord>& reqs);
  ServeSummary summary() const;  // a copy, taken under the server's lock
  // solution of request i of the last run (keep_x)
  const std::vector<double>& solution(int64_t i) const;
  // bytes a batch of these requests needs on the device
  int64_t batch_bytes(const std::vector<const ServeRequest*>& rs) const;

 private:
  void add_entry(std::shared_ptr<const PatternEntry> entry, int32_t n, const int32_t* Ap, const int32_t* Ai,
                 const double* base, int32_t n1);
  int64_t enqueue(std::unique_ptr<ServeRequest> r);
  void lane_main(int lane);
  void process(int lane, const std::vector<ServeRequest*>& batch, int32_t batch_id);

  ServeOptions opt_;
  PatternCache cache_;
  std::deque<ServePattern> patterns_;  // stable references; growth guarded by mu_
  std::vector<std::unique_ptr<GmresLdl>> solvers_;
  std::vector<std::vector<double>> x_;  // solutions of the last run (keep_x)
  // server state, guarded by mu_
  mutable std::mutex mu_;
  std::condition_variable work_cv_;  // lanes: a request arrived, memory freed, or stop
  std::condition_variable done_cv_;  // waiters: a batch completed
  std::vector<std::unique_ptr<ServeRequest>> reqs_;  // by id; null once released
  std::deque<int64_t> queue_;
  int64_t inflight_bytes_ = 0, concurrent_ = 0;
  int32_t batch_seq_ = 0;
  bool stop_ = false;
  double t0_ = 0.0;
  ServeSummary sum_;
  std::vector<std::thread> lanes_;
};

}  // namespace mpir
