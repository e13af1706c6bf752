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
    // the dependency-driven triangular solves: lanes on other streams
    // cannot starve them (a workgroup takes the task class of its arrival
    // order, solve.hip)
    s->set_schedule(kScheduleDag);
    s->set_max_ir(opt.max_ir);
    s->set_max_gmres(opt.max_gmres);
    solvers_.push_back(std::move(s));
  }
  sum_.lane_busy_ms.assign(opt_.lanes, 0.0);
  t0_ = now_ms();
  for (int l = 0; l < opt_.lanes; ++l) lanes_.emplace_back(&Server::lane_main, this, l);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  work_cv_.notify_all();
  for (auto& t : lanes_) t.join();
}

ServeSummary Server::summary() const {
  std::lock_guard<std::mutex> lk(mu_);
  return sum_;
}

int32_t Server::npatterns() const {
  std::lock_guard<std::mutex> lk(mu_);
  return static_cast<int32_t>(patterns_.size());
}

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
                              std::vector<const int32_t*>(Ai, Ai + count), opt_.ordering, kDefaultAmalgamation,
                              threads);
  } catch (const std::invalid_argument& e) {
    throw Error(kErrInvalid, e.what());
  }
  std::lock_guard<std::mutex> lk(mu_);
  const int32_t first = static_cast<int32_t>(patterns_.size());
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
  std::lock_guard<std::mutex> lk(mu_);
  MPIR_CHECK(id >= 0 && id < static_cast<int32_t>(patterns_.size()), "pattern id out of range");
  return patterns_[id];
}

const std::vector<double>& Server::solution(int64_t i) const {
  MPIR_CHECK(i >= 0 && i < static_cast<int64_t>(x_.size()) && !x_[i].empty(), "no kept solution for that request");
  return x_[i];
}

int64_t Server::batch_bytes(const std::vector<const ServeRequest*>& rs) const {
  const int sf = opt_.uf == kF64 ? 8 : 4;
  int64_t fixed = 0, n = 0;
  int r = 0;
  for (const ServeRequest* q : rs) {
    const PatternEntry& e = *patterns_[q->pattern].entry;
    fixed += pattern_footprint(e, sf, 0, opt_.max_gmres);
    n += e.n;
    r = std::max(r, q->nrhs);
  }
  // every system carries max-nrhs vector slots (basis, work vectors, I/O)
  return fixed + static_cast<int64_t>(r) * n * 8 * (opt_.max_gmres + 9);
}

int64_t Server::enqueue(std::unique_ptr<ServeRequest> r) {
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    MPIR_CHECK(!stop_, "the server is stopping");
    r->t_arrive = now_ms() - t0_;
    id = static_cast<int64_t>(reqs_.size());
    reqs_.push_back(std::move(r));
    queue_.push_back(id);
    sum_.max_backlog = std::max<int64_t>(sum_.max_backlog, static_cast<int64_t>(queue_.size()));
  }
  work_cv_.notify_one();
  return id;
}

int64_t Server::submit(int32_t pattern, const double* Ax, const double* b, int32_t nrhs) {
  MPIR_CHECK(Ax && b, "null request data");
  MPIR_CHECK(nrhs >= 1 && nrhs <= 64, "requests carry 1..64 right-hand sides");
  const ServePattern& p = this->pattern(pattern);
  auto r = std::make_unique<ServeRequest>();
  r->pattern = pattern;
  r->nrhs = nrhs;
  r->ax.assign(Ax, Ax + p.nnz);
  r->b.assign(b, b + static_cast<int64_t>(p.n) * nrhs);
  return enqueue(std::move(r));
}

int64_t Server::submit_seeded(int32_t pattern, uint64_t seed, int32_t nrhs) {
  MPIR_CHECK(nrhs >= 1 && nrhs <= 64, "requests carry 1..64 right-hand sides");
  (void)this->pattern(pattern);  // range check
  auto r = std::make_unique<ServeRequest>();
  r->pattern = pattern;
  r->nrhs = nrhs;
  r->seeded = true;
  r->seed = seed;
  return enqueue(std::move(r));
}

bool Server::wait(int64_t id, double timeout_ms, ServeRequest* info, double* x, bool release) {
  std::unique_lock<std::mutex> lk(mu_);
  MPIR_CHECK(id >= 0 && id < static_cast<int64_t>(reqs_.size()) && reqs_[id], "unknown or released request id");
  ServeRequest* r = reqs_[id].get();
  auto ready = [&] { return r->done; };
  if (timeout_ms < 0) done_cv_.wait(lk, ready);
  else if (timeout_ms > 0) done_cv_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready);
  if (!r->done) return false;
  if (x && r->status == 0) std::copy(r->x.begin(), r->x.end(), x);
  if (info) {
    *info = ServeRequest();
    info->pattern = r->pattern;
    info->nrhs = r->nrhs;
    info->seeded = r->seeded;
    info->seed = r->seed;
    info->t_arrive = r->t_arrive;
    info->t_admit = r->t_admit;
    info->t_done = r->t_done;
    info->batch = r->batch;
    info->lane = r->lane;
    info->status = r->status;
    info->ir_max = r->ir_max;
    info->gmres_max = r->gmres_max;
    info->rel_residual_max = r->rel_residual_max;
    info->bytes = r->bytes;
    info->done = true;
    if (!x) info->x = r->x;
  }
  if (release) reqs_[id].reset();
  return true;
}

void Server::run(std::vector<ServeRecord>& reqs) {
  const int32_t np = npatterns();
  for (size_t i = 0; i < reqs.size(); ++i) {
    const ServeRecord& r = reqs[i];
    MPIR_CHECK(r.pattern >= 0 && r.pattern < np, "request names an unknown pattern");
    MPIR_CHECK(r.nrhs >= 1 && r.nrhs <= 64, "requests carry 1..64 right-hand sides");
    MPIR_CHECK(i == 0 || r.t_arrive >= reqs[i - 1].t_arrive, "requests must be sorted by arrival time");
  }
  double t_run;
  {
    std::lock_guard<std::mutex> lk(mu_);
    MPIR_CHECK(queue_.empty() && concurrent_ == 0, "run needs an idle server");
    sum_ = ServeSummary();
    sum_.lane_busy_ms.assign(opt_.lanes, 0.0);
    t_run = now_ms();
  }
  const int64_t grown = DevBuf::regrowths().load();
  x_.assign(opt_.keep_x ? reqs.size() : 0, std::vector<double>());
  // submit each request at its arrival time (a late submitter catches up)
  std::vector<int64_t> ids(reqs.size());
  for (size_t i = 0; i < reqs.size(); ++i) {
    const double wait = t_run + reqs[i].t_arrive - now_ms();
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
    ids[i] = submit_seeded(reqs[i].pattern, reqs[i].seed, reqs[i].nrhs);
  }
  const double off = t_run - t0_;  // run clock = server clock - off
  for (size_t i = 0; i < reqs.size(); ++i) {
    ServeRequest info;
    wait(ids[i], -1.0, &info, nullptr, true);
    ServeRecord& r = reqs[i];
    r.t_admit = info.t_admit - off;
    r.t_done = info.t_done - off;
    r.batch = info.batch;
    r.lane = info.lane;
    r.status = info.status;
    r.ir_max = info.ir_max;
    r.gmres_max = info.gmres_max;
    r.rel_residual_max = info.rel_residual_max;
    r.bytes = info.bytes;
    if (opt_.keep_x && info.status == 0) x_[i] = std::move(info.x);
  }
  std::lock_guard<std::mutex> lk(mu_);
  sum_.wall_ms = now_ms() - t_run;
  sum_.regrowths = DevBuf::regrowths().load() - grown;
}

void Server::lane_main(int lane) {
  for (;;) {
    std::vector<ServeRequest*> batch;
    int64_t bytes = 0;
    int32_t bid = 0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      for (;;) {
        if (queue_.empty()) {
          if (stop_) return;
          work_cv_.wait(lk);
          continue;
        }
        // admission in arrival order by device footprint
        std::vector<const ServeRequest*> cand;
        while (!queue_.empty() && static_cast<int>(batch.size()) < opt_.max_batch) {
          ServeRequest* r = reqs_[queue_.front()].get();
          cand.push_back(r);
          const int64_t b = batch_bytes(cand);
          if (inflight_bytes_ + b > opt_.budget_bytes && !(batch.empty() && inflight_bytes_ == 0)) {
            cand.pop_back();
            break;
          }
          batch.push_back(r);
          bytes = b;
          queue_.pop_front();
        }
        if (batch.empty()) {
          work_cv_.wait(lk);  // memory frees when a batch in flight completes
          continue;
        }
        inflight_bytes_ += bytes;
        concurrent_++;
        sum_.peak_inflight_bytes = std::max(sum_.peak_inflight_bytes, inflight_bytes_);
        sum_.max_concurrent = std::max(sum_.max_concurrent, concurrent_);
        bid = batch_seq_++;
        sum_.batches++;
        const double ta = now_ms() - t0_;
        for (ServeRequest* r : batch) {
          r->t_admit = ta;
          r->bytes = bytes;
        }
        break;
      }
    }
    const double t_begin = now_ms();
    process(lane, batch, bid);
    {
      std::lock_guard<std::mutex> lk(mu_);
      sum_.lane_busy_ms[lane] += now_ms() - t_begin;
      inflight_bytes_ -= bytes;
      concurrent_--;
      for (ServeRequest* r : batch) {
        r->done = true;
        if (r->status == 0) sum_.completed++;
        else sum_.failed++;
      }
    }
    done_cv_.notify_all();
    work_cv_.notify_all();
  }
}

void Server::process(int lane, const std::vector<ServeRequest*>& batch, int32_t batch_id) {
  GmresLdl& s = *solvers_[lane];
  const size_t B = batch.size();
  const double t_prep = now_ms();
  std::vector<const ServePattern*> pats(B);
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t k = 0; k < B; ++k) pats[k] = &patterns_[batch[k]->pattern];
  }
  std::vector<std::vector<double>> xv(B);
  std::vector<std::shared_ptr<const PatternEntry>> parts(B);
  std::vector<int32_t> nrhs(B);
  // request data: the caller's, or derived from the pattern and the seed
  parallel_for(static_cast<int64_t>(B), opt_.threads, [&](int64_t k) {
    ServeRequest& r = *batch[k];
    const ServePattern& p = *pats[k];
    xv[k].resize(static_cast<size_t>(p.n) * r.nrhs);
    if (r.seeded) {
      r.ax.resize(p.nnz);
      r.b.resize(static_cast<size_t>(p.n) * r.nrhs);
      std::vector<double> xt(static_cast<size_t>(p.n) * r.nrhs);
      materialize(p, r.seed, r.nrhs, r.ax.data(), r.b.data(), xt.data());
    }
  });
  std::vector<const double*> axp(B), bp(B);
  std::vector<double*> xp(B);
  for (size_t k = 0; k < B; ++k) {
    parts[k] = pats[k]->entry;
    nrhs[k] = batch[k]->nrhs;
    axp[k] = batch[k]->ax.data();
    bp[k] = batch[k]->b.data();
    xp[k] = xv[k].data();
  }
  int status = 0;
  const double t1 = now_ms();
  double t2 = t1, t3 = t1;
  double join = 0.0, upload = 0.0;
  try {
    s.analyze_batch(parts);
    t2 = now_ms();
    join = s.batch_stats().join_ms;
    upload = s.batch_stats().upload_ms;
    s.factor_batch(axp.data(), false);
    t3 = now_ms();
    s.solve_batch(nrhs.data(), bp.data(), xp.data(), false);
  } catch (const Error& e) {
    status = e.code;
  } catch (const std::exception&) {
    status = kErrInternal;
  }
  const double t4 = now_ms();
  const BatchStats& st = s.batch_stats();
  std::lock_guard<std::mutex> lk(mu_);
  sum_.prepare_ms += t1 - t_prep;
  sum_.plan_ms += t2 - t1;
  sum_.plan_join_ms += join;
  sum_.plan_upload_ms += upload;
  sum_.factor_ms += t3 - t2;
  sum_.solve_ms += t4 - t3;
  const double td = t4 - t0_;
  for (size_t k = 0; k < B; ++k) {
    ServeRequest& r = *batch[k];
    r.t_done = td;
    r.batch = batch_id;
    r.lane = lane;
    r.status = status;
    std::vector<double>().swap(r.ax);
    std::vector<double>().swap(r.b);
    if (status == 0) {
      const int R = st.max_rhs;
      for (int c = 0; c < r.nrhs; ++c) {
        const size_t o = k * R + c;
        // an instance whose data are not finite stops with reason 4
        if (st.stop_reason[o] == 4) r.status = kErrNumeric;
        r.ir_max = std::max(r.ir_max, st.ir_iterations[o]);
        r.gmres_max = std::max(r.gmres_max, st.gmres_iterations[o]);
        r.rel_residual_max = std::max(r.rel_residual_max, st.final_rel_residual[o]);
      }
      r.x = std::move(xv[k]);
    }
  }
}

}  // namespace mpir
