#include "serve.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <thread>

#include "common.hpp"
#include "solver.hpp"

namespace mpir {

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// SplitMix64 (Steele, Lea & Flood 2014), 53-bit uniforms in [0, 1)
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
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
  const int nt = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(threads, n)));
  if (nt == 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (int64_t i; (i = next.fetch_add(1)) < n;) f(i);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
}

}  // namespace

void materialize(const ServePattern& p, uint64_t seed, int nrhs, double* Ax, double* b, double* x_true) {
  Rng rng(seed * 0x2545F4914F6CDD1Dull + 0x632BE59BD9B4E019ull);
  for (int64_t q = 0; q < p.nnz; ++q) {
    const double u = rng.uniform();
    const double f = p.cls[q] == 0 ? 1.05 + 0.05 * u : 1.0 + 0.05 * (2.0 * u - 1.0);
    Ax[q] = p.base[q] * f;
  }
  const int32_t n = p.n;
  const std::vector<int32_t>& Ai = p.entry->Ai;
  for (int r = 0; r < nrhs; ++r) {
    double* xt = x_true + static_cast<int64_t>(r) * n;
    double* br = b + static_cast<int64_t>(r) * n;
    for (int32_t i = 0; i < n; ++i) xt[i] = 2.0 * rng.uniform() - 1.0;
    // b = K x_true from the upper CSC (column sweep, both triangles)
    std::fill(br, br + n, 0.0);
    for (int64_t q = 0; q < p.nnz; ++q) {
      const int32_t i = Ai[q], j = p.col[q];
      br[j] += Ax[q] * xt[i];
      if (i != j) br[i] += Ax[q] * xt[j];
    }
  }
}

Server::Server(const ServeOptions& opt) : opt_(opt) {
  MPIR_CHECK(opt.lanes >= 1 && opt.lanes <= 8, "lanes must be 1..8");
  MPIR_CHECK(opt.max_batch >= 1, "max_batch must be positive");
  MPIR_CHECK(opt.budget_bytes > 0, "memory budget must be positive");
  MPIR_CHECK(opt.uf == kF64 || opt.uf == kF32, "serving factor precision must be f64 or f32");
  for (int l = 0; l < opt.lanes; ++l) {
    auto s = std::make_unique<GmresLdl>(opt.uf, kF64, kF64);
    s->set_tol(opt.tol);
    s->set_ordering(opt.ordering);
    // Lanes run concurrently on their own streams. The level schedule of the
    // triangular solves never waits across workgroups, so a lane's solve
    // cannot stall behind another lane's kernels; on joined batches it is
    // also the faster schedule.
    s->set_schedule(kScheduleLevels);
    s->set_max_ir(opt.max_ir);
    s->set_max_gmres(opt.max_gmres);
    solvers_.push_back(std::move(s));
  }
}

Server::~Server() = default;

int32_t Server::add_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* base, int32_t n1) {
  return add_patterns(1, &n, &Ap, &Ai, &base, &n1, 1);
}

int32_t Server::add_patterns(int32_t count, const int32_t* n, const int32_t* const* Ap, const int32_t* const* Ai,
                             const double* const* base, const int32_t* n1, int threads) {
  MPIR_CHECK(count >= 1 && n && Ap && Ai && base && n1, "bad pattern list");
  for (int32_t k = 0; k < count; ++k) {
    MPIR_CHECK(n[k] >= 1 && Ap[k] && Ai[k] && base[k], "bad pattern");
    MPIR_CHECK(n1[k] >= 0 && n1[k] <= n[k], "n1 must lie in [0, n]");
  }
  std::vector<std::shared_ptr<const PatternEntry>> entries;
  try {
    entries = cache_.get_many(std::vector<int32_t>(n, n + count), std::vector<const int32_t*>(Ap, Ap + count),
                              std::vector<const int32_t*>(Ai, Ai + count), opt_.ordering, true, threads);
  } catch (const std::invalid_argument& e) {
    throw Error(kErrInvalid, e.what());
  }
  const int32_t first = npatterns();
  for (int32_t k = 0; k < count; ++k) add_entry(entries[k], n[k], Ap[k], Ai[k], base[k], n1[k]);
  return first;
}

void Server::add_entry(std::shared_ptr<const PatternEntry> entry, int32_t n, const int32_t* Ap, const int32_t* Ai,
                       const double* base, int32_t n1) {
  ServePattern p;
  p.entry = std::move(entry);
  p.n = n;
  p.n1 = n1;
  p.nnz = Ap[n];
  p.base.assign(base, base + p.nnz);
  p.cls.resize(p.nnz);
  p.col.resize(p.nnz);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q) {
      const int32_t i = Ai[q];
      p.col[q] = j;
      p.cls[q] = i == j ? (j < n1 ? 0 : 3) : (j < n1 ? 1 : 2);
    }
  patterns_.push_back(std::move(p));
}

const ServePattern& Server::pattern(int32_t id) const {
  MPIR_CHECK(id >= 0 && id < npatterns(), "pattern id out of range");
  return patterns_[id];
}

const std::vector<double>& Server::solution(int64_t i) const {
  MPIR_CHECK(i >= 0 && i < static_cast<int64_t>(x_.size()) && !x_[i].empty(), "no kept solution for that request");
  return x_[i];
}

int64_t Server::batch_bytes(const std::vector<const ServeRecord*>& rs) const {
  const int sf = opt_.uf == kF64 ? 8 : 4;
  int64_t fixed = 0, n = 0;
  int r = 0;
  for (const ServeRecord* q : rs) {
    const PatternEntry& e = *patterns_[q->pattern].entry;
    fixed += pattern_footprint(e, sf, 0, opt_.max_gmres);
    n += e.n;
    r = std::max(r, q->nrhs);
  }
  // every system carries max-nrhs vector slots (basis, work vectors, I/O)
  return fixed + static_cast<int64_t>(r) * n * 8 * (opt_.max_gmres + 9);
}

void Server::run(std::vector<ServeRecord>& reqs) {
  for (size_t i = 0; i < reqs.size(); ++i) {
    const ServeRecord& r = reqs[i];
    MPIR_CHECK(r.pattern >= 0 && r.pattern < npatterns(), "request names an unknown pattern");
    MPIR_CHECK(r.nrhs >= 1 && r.nrhs <= 64, "requests carry 1..64 right-hand sides");
    MPIR_CHECK(i == 0 || r.t_arrive >= reqs[i - 1].t_arrive, "requests must be sorted by arrival time");
  }
  sum_ = ServeSummary();
  sum_.lane_busy_ms.assign(opt_.lanes, 0.0);
  x_.assign(opt_.keep_x ? reqs.size() : 0, std::vector<double>());
  queue_.clear();
  next_ = 0;
  outstanding_ = static_cast<int64_t>(reqs.size());
  inflight_bytes_ = concurrent_ = 0;
  batch_seq_ = 0;
  t0_ = now_ms();
  std::vector<std::thread> lanes;
  for (int l = 0; l < opt_.lanes; ++l) lanes.emplace_back(&Server::lane_main, this, l, &reqs);
  for (auto& t : lanes) t.join();
  sum_.wall_ms = now_ms() - t0_;
  for (const ServeRecord& r : reqs) {
    if (r.status == 0) sum_.completed++;
    else sum_.failed++;
  }
}

void Server::lane_main(int lane, std::vector<ServeRecord>* reqs) {
  const int64_t total = static_cast<int64_t>(reqs->size());
  for (;;) {
    std::vector<int64_t> ids;
    int64_t bytes = 0;
    int32_t bid = 0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      for (;;) {
        // release the requests that have arrived
        const double t = now_ms() - t0_;
        while (next_ < total && (*reqs)[next_].t_arrive <= t) queue_.push_back(next_++);
        sum_.max_backlog = std::max<int64_t>(sum_.max_backlog, static_cast<int64_t>(queue_.size()));
        if (outstanding_ == 0) return;
        if (queue_.empty()) {
          if (next_ >= total) {
            cv_.wait(lk);  // others still running; nothing left to admit
            if (outstanding_ == 0) return;
            continue;
          }
          const double wait = (*reqs)[next_].t_arrive - t;
          cv_.wait_for(lk, std::chrono::duration<double, std::milli>(std::max(0.0, wait)));
          continue;
        }
        // admission in arrival order by device footprint
        std::vector<const ServeRecord*> cand;
        while (!queue_.empty() && static_cast<int>(ids.size()) < opt_.max_batch) {
          const int64_t id = queue_.front();
          cand.push_back(&(*reqs)[id]);
          const int64_t b = batch_bytes(cand);
          if (inflight_bytes_ + b > opt_.budget_bytes && !(ids.empty() && inflight_bytes_ == 0)) {
            cand.pop_back();
            break;
          }
          ids.push_back(id);
          bytes = b;
          queue_.pop_front();
        }
        if (ids.empty()) {
          cv_.wait(lk);  // memory frees when a batch in flight completes
          continue;
        }
        inflight_bytes_ += bytes;
        concurrent_++;
        sum_.peak_inflight_bytes = std::max(sum_.peak_inflight_bytes, inflight_bytes_);
        sum_.max_concurrent = std::max(sum_.max_concurrent, concurrent_);
        bid = batch_seq_++;
        sum_.batches++;
        const double ta = now_ms() - t0_;
        for (int64_t id : ids) {
          (*reqs)[id].t_admit = ta;
          (*reqs)[id].bytes = bytes;
        }
        break;
      }
    }
    const double t_begin = now_ms();
    process(lane, reqs, ids, bid);
    {
      std::lock_guard<std::mutex> lk(mu_);
      sum_.lane_busy_ms[lane] += now_ms() - t_begin;
      inflight_bytes_ -= bytes;
      concurrent_--;
      outstanding_ -= static_cast<int64_t>(ids.size());
    }
    cv_.notify_all();
  }
}

void Server::process(int lane, std::vector<ServeRecord>* reqs, const std::vector<int64_t>& ids, int32_t batch_id) {
  GmresLdl& s = *solvers_[lane];
  const size_t B = ids.size();
  const double t_prep = now_ms();
  std::vector<std::vector<double>> ax(B), bv(B), xv(B);
  std::vector<std::shared_ptr<const PatternEntry>> parts(B);
  std::vector<int32_t> nrhs(B);
  // request data (deterministic from the pattern and the seed)
  parallel_for(static_cast<int64_t>(B), opt_.threads, [&](int64_t k) {
    const ServeRecord& r = (*reqs)[ids[k]];
    const ServePattern& p = patterns_[r.pattern];
    ax[k].resize(p.nnz);
    bv[k].resize(static_cast<size_t>(p.n) * r.nrhs);
    xv[k].resize(static_cast<size_t>(p.n) * r.nrhs);
    std::vector<double> xt(static_cast<size_t>(p.n) * r.nrhs);
    materialize(p, r.seed, r.nrhs, ax[k].data(), bv[k].data(), xt.data());
  });
  std::vector<const double*> axp(B), bp(B);
  std::vector<double*> xp(B);
  for (size_t k = 0; k < B; ++k) {
    const ServeRecord& r = (*reqs)[ids[k]];
    parts[k] = patterns_[r.pattern].entry;
    nrhs[k] = r.nrhs;
    axp[k] = ax[k].data();
    bp[k] = bv[k].data();
    xp[k] = xv[k].data();
  }
  int status = 0;
  const double t1 = now_ms();
  double t2 = t1, t3 = t1;
  try {
    s.analyze_batch(parts);
    t2 = now_ms();
    s.factor_batch(axp.data(), false);
    t3 = now_ms();
    s.solve_batch(nrhs.data(), bp.data(), xp.data(), false);
  } catch (const Error& e) {
    status = e.code;
  } catch (const std::exception&) {
    status = kErrInternal;
  }
  const double t4 = now_ms();
  {
    std::lock_guard<std::mutex> lk(mu_);
    sum_.prepare_ms += t1 - t_prep;
    sum_.plan_ms += t2 - t1;
    sum_.factor_ms += t3 - t2;
    sum_.solve_ms += t4 - t3;
  }
  const double td = t4 - t0_;
  const BatchStats& st = s.batch_stats();
  for (size_t k = 0; k < B; ++k) {
    ServeRecord& r = (*reqs)[ids[k]];
    r.t_done = td;
    r.batch = batch_id;
    r.lane = lane;
    r.status = status;
    if (status == 0) {
      const int R = st.max_rhs;
      for (int c = 0; c < r.nrhs; ++c) {
        const size_t o = k * R + c;
        r.ir_max = std::max(r.ir_max, st.ir_iterations[o]);
        r.gmres_max = std::max(r.gmres_max, st.gmres_iterations[o]);
        r.rel_residual_max = std::max(r.rel_residual_max, st.final_rel_residual[o]);
      }
      if (opt_.keep_x) x_[ids[k]] = std::move(xv[k]);
    }
  }
}

}  // namespace mpir
