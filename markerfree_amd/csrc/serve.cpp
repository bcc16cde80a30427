// Batched and served forms over the per-series path.
//
// Batched: mf_batch_run aligns many series on one context concurrently --
// each series has its own HIP stream and a host worker drives it, so the
// kernels of different series (different sizes, tilt counts, stages)
// overlap on the device and one series' host-side steps hide behind the
// others' kernels.
//
// Served: an mf_server accepts series as they arrive (mf_submit), admits
// them in arrival order against an HBM budget (the context's budget, or the
// device memory free at creation) using the per-series byte bound, keeps
// admitted series resident on the device until they finish, and runs them
// on a fixed pool of workers; results are fetched by id.  Requests that
// arrive while others run start as soon as a worker and memory free up.
//
// Every series computes exactly what it computes alone: the kernels are
// deterministic (fixed reduction orders, no atomics on results) and a
// series shares no mutable device state with another.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "context.h"
#include "markerfree/mf_api.h"
#include "series.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Free page-locked blocks kept for reuse, at most 'cap' of them.  The first
// host write to newly page-locked memory costs about 0.3 ms per MB (page
// population), several times a plain copy, so a request's stack copy and
// its aligned stack go into recycled blocks when one of a fitting size is
// free (no more than twice the size asked for).  The server sets the cap to
// the number of blocks its admitted and queued requests can hold at once
// (two per running request, one per queued request): the pool keeps no more
// blocks than the requests themselves can hold.
class PinnedPool {
 public:
  explicit PinnedPool(size_t cap) : cap_(cap) {}
  PinnedPool(const PinnedPool&) = delete;
  PinnedPool& operator=(const PinnedPool&) = delete;
  ~PinnedPool() {
    for (auto& b : free_) mf_host_free(b.second);
  }
  // A block of at least n floats (its size in *got), or nullptr.
  float* take(size_t n, size_t* got) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = free_.lower_bound(n);
      if (it != free_.end() && it->first <= 2 * n) {
        float* p = it->second;
        *got = it->first;
        free_.erase(it);
        return p;
      }
    }
    void* q = nullptr;
    if (mf_host_alloc(n * sizeof(float), &q) != MF_OK) return nullptr;
    *got = n;
    return (float*)q;
  }
  void give(float* p, size_t n) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (free_.size() < cap_) {
        free_.emplace(n, p);
        return;
      }
    }
    mf_host_free(p);
  }

 private:
  std::mutex mu_;
  std::multimap<size_t, float*> free_;  // size in floats -> block
  size_t cap_;
};

// Page-locked float buffer from the pool: the request's stack copy and its
// aligned stack move to and from the device by DMA at the PCIe rate.
class PinnedFloats {
 public:
  PinnedFloats() = default;
  PinnedFloats(const PinnedFloats&) = delete;
  PinnedFloats& operator=(const PinnedFloats&) = delete;
  ~PinnedFloats() { clear(); }
  // Returns false when page-locked memory cannot be had.
  bool resize(PinnedPool* pool, size_t n) {
    clear();
    if (!n) return true;
    p_ = pool->take(n, &cap_);
    if (!p_) return false;
    pool_ = pool;
    n_ = n;
    return true;
  }
  void clear() {
    if (p_) pool_->give(p_, cap_);
    p_ = nullptr;
    pool_ = nullptr;
    n_ = cap_ = 0;
  }
  float* data() { return p_; }
  const float* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  PinnedPool* pool_ = nullptr;
  float* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

struct Request {
  long long id = 0;
  int nx = 0, ny = 0, nz = 0;
  PinnedFloats stack;               // host copy (file order)
  std::vector<float> tilts;
  mf_options opt{};
  size_t bytes = 0;                 // admission charge
  int state = MF_REQ_QUEUED;
  std::string error;
  // Results.
  std::vector<double> sx, sy;
  std::vector<float> rot;
  PinnedFloats aligned;
  float offset = 0.0f;
  int bx = 0, by = 0;
  double t_submit = 0, t_start = 0, t_done = 0;  // seconds since server start
};

}  // namespace

struct mf_server {
  mf_ctx* ctx = nullptr;
  // Declared first, destroyed last: every request returns its blocks here.
  std::unique_ptr<PinnedPool> pool;
  size_t budget = 0, reserved = 0;
  int queue_cap = 0;
  long long next_id = 1;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;  // waiting, arrival order
  std::map<long long, std::shared_ptr<Request>> all;
  std::vector<std::thread> workers;
  int running = 0;
  bool stopping = false;
  long long completed = 0;
};

namespace {

// Runs one series end to end and fills the request's results.  Returns
// false with the message in *err on failure.  It does not touch r.state or
// the times: the worker publishes those under the server lock, so a reader
// that sees MF_REQ_DONE also sees the results and t_done.
bool process(mf_ctx* ctx, PinnedPool* pool, Request& r, std::string* err) {
  mf_series* s = nullptr;
  if (mf_series_create(ctx, r.nx, r.ny, r.nz, r.stack.data(), r.tilts.data(),
                       &s) != MF_OK) {
    *err = mf_last_error(ctx);
    return false;
  }
  auto fail = [&]() {
    *err = mf_series_last_error(s);
    mf_series_destroy(s);
    return false;
  };
  if (mf_run(s, &r.opt) != MF_OK) return fail();
  if (mf_correct(s, r.opt.out_bin, nullptr, &r.bx, &r.by) != MF_OK)
    return fail();
  if (!r.aligned.resize(pool, (size_t)r.bx * r.by * r.nz)) {
    mf_series_destroy(s);
    *err = "no page-locked host memory for the aligned stack";
    return false;
  }
  if (mf_correct(s, r.opt.out_bin, r.aligned.data(), &r.bx, &r.by) != MF_OK)
    return fail();
  r.sx.resize(r.nz);
  r.sy.resize(r.nz);
  r.rot.resize(r.nz);
  if (mf_get_params(s, r.sx.data(), r.sy.data(), r.rot.data(), &r.offset) !=
      MF_OK)
    return fail();
  mf_series_destroy(s);
  return true;
}

void worker_loop(mf_server* sv) {
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      // Head-of-line admission: the oldest waiting request starts when its
      // bytes fit next to the running ones.
      sv->cv_work.wait(lk, [&] {
        return sv->stopping ||
               (!sv->queue.empty() &&
                sv->reserved + sv->queue.front()->bytes <= sv->budget);
      });
      if (sv->stopping && sv->queue.empty()) return;
      if (sv->stopping) {
        // Drain without starting new work.
        for (auto& q : sv->queue) {
          q->state = MF_REQ_FAILED;
          q->error = "server stopped";
        }
        sv->queue.clear();
        sv->cv_done.notify_all();
        return;
      }
      r = sv->queue.front();
      sv->queue.pop_front();
      sv->reserved += r->bytes;
      sv->running++;
      r->state = MF_REQ_RUNNING;
      r->t_start = seconds_since(sv->t0);
    }
    std::string err;
    const bool ok = process(sv->ctx, sv->pool.get(), *r, &err);
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      r->state = ok ? MF_REQ_DONE : MF_REQ_FAILED;
      if (!ok) r->error = err;
      sv->reserved -= r->bytes;
      sv->running--;
      sv->completed++;
      r->t_done = seconds_since(sv->t0);
      r->stack.clear();
    }
    sv->cv_work.notify_all();
    sv->cv_done.notify_all();
  }
}

template <typename F>
mf_status guarded(std::string* err, F&& f) {
  try {
    f();
    return MF_OK;
  } catch (const mf::Error& e) {
    if (err) *err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (err) *err = "host allocation failed";
    return MF_ENOMEM;
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    return MF_EDEVICE;
  }
}

}  // namespace

extern "C" {

mf_status mf_batch_run(mf_series** series, int n, const mf_options* o,
                       int concurrency) {
  if (!series || n < 0 || !o) return MF_EINVAL;
  for (int i = 0; i < n; i++)
    if (!series[i]) return MF_EINVAL;
  const int workers = std::max(1, std::min(concurrency, n));
  std::vector<mf_status> st(n, MF_OK);
  std::mutex mu;
  int next = 0;
  auto loop = [&] {
    for (;;) {
      int i;
      {
        std::lock_guard<std::mutex> g(mu);
        if (next >= n) return;
        i = next++;
      }
      st[i] = mf_run(series[i], o);
      // Only the stack and the parameters are needed from here on.
      if (st[i] == MF_OK) st[i] = mf_series_trim(series[i]);
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < workers; w++) pool.emplace_back(loop);
  loop();
  for (auto& t : pool) t.join();
  for (int i = 0; i < n; i++)
    if (st[i] != MF_OK) return st[i];
  return MF_OK;
}

size_t mf_series_hbm_bytes(int nx, int ny, int nz, const mf_options* o) {
  if (!o || nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return mf::series_hbm_bytes(nx, ny, nz, o);
}

mf_status mf_server_create(mf_ctx* ctx, int workers, size_t hbm_budget,
                           int queue_cap, mf_server** out) {
  if (!ctx || !out || workers < 1 || queue_cap < 1) return MF_EINVAL;
  *out = nullptr;
  mf_server* sv = nullptr;
  std::string err;
  mf_status st = guarded(&err, [&] {
    sv = new mf_server;
    sv->ctx = ctx;
    // Recycled page-locked blocks: as many as the requests can hold.
    sv->pool.reset(new PinnedPool((size_t)2 * workers + queue_cap));
    size_t budget = hbm_budget ? hbm_budget : ctx->budget;
    if (!budget) {
      size_t free_b = 0, total_b = 0;
      MF_HIP(hipSetDevice(ctx->device));
      MF_HIP(hipMemGetInfo(&free_b, &total_b));
      budget = free_b - free_b / 16;  // leave headroom for the runtime
    }
    sv->budget = budget;
    sv->queue_cap = queue_cap;
    sv->t0 = Clock::now();
    for (int w = 0; w < workers; w++) sv->workers.emplace_back(worker_loop, sv);
  });
  if (st != MF_OK) {
    ctx->last_error = err;
    delete sv;
    return st;
  }
  *out = sv;
  return MF_OK;
}

void mf_server_destroy(mf_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stopping = true;
  }
  sv->cv_work.notify_all();
  for (auto& t : sv->workers) t.join();
  delete sv;
}

mf_status mf_submit(mf_server* sv, int nx, int ny, int nz,
                    const float* stack_zyx, const float* tilts,
                    const mf_options* o, long long* id) {
  if (!sv || !stack_zyx || !tilts || !o || !id || nx < 8 || ny < 8 || nz < 3)
    return MF_EINVAL;
  auto r = std::make_shared<Request>();
  r->nx = nx;
  r->ny = ny;
  r->nz = nz;
  r->opt = *o;
  r->opt.verbose = 0;
  r->bytes = mf::series_hbm_bytes(nx, ny, nz, o);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stopping) return MF_ESTATE;
    if (r->bytes > sv->budget) return MF_ENOMEM;  // could never be admitted
    if ((int)sv->queue.size() >= sv->queue_cap) return MF_EBUSY;
  }
  const size_t n = (size_t)nx * ny * nz;
  if (!r->stack.resize(sv->pool.get(), n)) return MF_ENOMEM;
  std::memcpy(r->stack.data(), stack_zyx, n * sizeof(float));
  r->tilts.assign(tilts, tilts + nz);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    r->id = sv->next_id++;
    r->t_submit = seconds_since(sv->t0);
    sv->queue.push_back(r);
    sv->all[r->id] = r;
    *id = r->id;
  }
  sv->cv_work.notify_all();
  return MF_OK;
}

mf_status mf_server_clock(mf_server* sv, double* seconds) {
  if (!sv || !seconds) return MF_EINVAL;
  *seconds = seconds_since(sv->t0);
  return MF_OK;
}

mf_status mf_poll(mf_server* sv, long long id, int* state) {
  if (!sv || !state) return MF_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->all.find(id);
  if (it == sv->all.end()) return MF_EINVAL;
  *state = it->second->state;
  return MF_OK;
}

mf_status mf_wait(mf_server* sv, long long id, int* state) {
  if (!sv || !state) return MF_EINVAL;
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->all.find(id);
  if (it == sv->all.end()) return MF_EINVAL;
  std::shared_ptr<Request> r = it->second;
  sv->cv_done.wait(lk, [&] {
    return r->state == MF_REQ_DONE || r->state == MF_REQ_FAILED;
  });
  *state = r->state;
  return MF_OK;
}

mf_status mf_fetch(mf_server* sv, long long id, double* shift_x,
                   double* shift_y, float* rotate, float* angle_offset,
                   float* aligned, int* bx, int* by, double* times3,
                   char* err, size_t err_size) {
  if (!sv) return MF_EINVAL;
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return MF_EINVAL;
    r = it->second;
    if (r->state == MF_REQ_QUEUED || r->state == MF_REQ_RUNNING)
      return MF_ESTATE;
    // State and times are published together under the lock; the results
    // below were written before the state changed and are not modified
    // after it, so they are read without the lock.
    if (times3) {
      times3[0] = r->t_submit;
      times3[1] = r->t_start;
      times3[2] = r->t_done;
    }
  }
  if (r->state == MF_REQ_FAILED) {
    if (err && err_size) {
      std::strncpy(err, r->error.c_str(), err_size - 1);
      err[err_size - 1] = '\0';
    }
    return MF_EDEVICE;
  }
  const size_t n = r->nz;
  if (shift_x) std::memcpy(shift_x, r->sx.data(), n * sizeof(double));
  if (shift_y) std::memcpy(shift_y, r->sy.data(), n * sizeof(double));
  if (rotate) std::memcpy(rotate, r->rot.data(), n * sizeof(float));
  if (angle_offset) *angle_offset = r->offset;
  if (bx) *bx = r->bx;
  if (by) *by = r->by;
  if (aligned)
    std::memcpy(aligned, r->aligned.data(), r->aligned.size() * sizeof(float));
  return MF_OK;
}

mf_status mf_release(mf_server* sv, long long id) {
  if (!sv) return MF_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->all.find(id);
  if (it == sv->all.end()) return MF_EINVAL;
  if (it->second->state == MF_REQ_QUEUED || it->second->state == MF_REQ_RUNNING)
    return MF_ESTATE;
  sv->all.erase(it);
  return MF_OK;
}

mf_status mf_server_stats(mf_server* sv, size_t* budget, size_t* reserved,
                          int* queued, int* running, long long* completed) {
  if (!sv) return MF_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (budget) *budget = sv->budget;
  if (reserved) *reserved = sv->reserved;
  if (queued) *queued = (int)sv->queue.size();
  if (running) *running = sv->running;
  if (completed) *completed = sv->completed;
  return MF_OK;
}

}  // extern "C"
