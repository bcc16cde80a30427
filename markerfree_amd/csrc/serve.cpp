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
#include <new>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "context.h"
#include "markerfree/mf_api.h"
#include "pinned.h"
#include "series.h"

namespace {

using Clock = std::chrono::steady_clock;
using mf::PinnedPool;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Host float buffer, page-locked from the pool where the system grants it
// (the request's stack copy and its aligned stack then move to and from the
// device by DMA at the PCIe rate), else ordinary memory (staged copies).
class PinnedFloats {
 public:
  PinnedFloats() = default;
  PinnedFloats(const PinnedFloats&) = delete;
  PinnedFloats& operator=(const PinnedFloats&) = delete;
  ~PinnedFloats() { clear(); }
  // Returns false only when no host memory at all can be had.
  bool resize(PinnedPool* pool, size_t n) {
    clear();
    if (!n) return true;
    p_ = pool->take(n, &cap_);
    if (p_) {
      pool_ = pool;
    } else {
      heap_.reset(new (std::nothrow) float[n]);
      if (!heap_) return false;
      p_ = heap_.get();
      cap_ = n;
    }
    n_ = n;
    return true;
  }
  void clear() {
    if (p_ && pool_) pool_->give(p_, cap_);
    heap_.reset();
    p_ = nullptr;
    pool_ = nullptr;
    n_ = cap_ = 0;
  }
  float* data() { return p_; }
  const float* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  PinnedPool* pool_ = nullptr;
  std::unique_ptr<float[]> heap_;
  float* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

struct Request {
  long long id = 0;
  int nx = 0, ny = 0, nz = 0;
  PinnedFloats stack;               // host copy (file order)
  const float* borrowed = nullptr;  // the caller's stack (mf_submit_borrowed)
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
  const float* raw = r.borrowed ? r.borrowed : r.stack.data();
  if (mf_series_create(ctx, r.nx, r.ny, r.nz, raw, r.tilts.data(), &s) !=
      MF_OK) {
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
    *err = "no host memory for the aligned stack";
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
      r->borrowed = nullptr;  // the caller may reuse its stack now
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

mf_status mf_batch_run_each(mf_series** series, int n, const mf_options* opts,
                            int concurrency, int correct, mf_batch_done done,
                            void* user) {
  if (n == 0) return MF_OK;
  if (!series || n < 0 || !opts) return MF_EINVAL;
  for (int i = 0; i < n; i++)
    if (!series[i] || opts[i].out_bin < 1) return MF_EINVAL;
  mf_ctx* ctx = series[0]->ctx;
  for (int i = 0; i < n; i++)
    if (series[i]->ctx != ctx) return MF_EINVAL;
  const int workers = std::max(1, std::min(concurrency, n));
  // Largest series first: the longest ones do not start last and leave the
  // other workers idle at the end.
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const mf_series *x = series[a], *y = series[b];
    return (long long)x->nx * x->ny * x->nz > (long long)y->nx * y->ny * y->nz;
  });
  PinnedPool* pool = nullptr;
  if (correct) {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!ctx->out_pool) ctx->out_pool.reset(new PinnedPool(0));
    ctx->out_pool->keep_at_least((size_t)workers);
    pool = ctx->out_pool.get();
  }
  std::vector<mf_status> st(n, MF_OK);
  std::mutex mu;
  int next = 0;
  auto loop = [&] {
    PinnedFloats out;       // this worker's output block
    for (;;) {
      int i;
      {
        std::lock_guard<std::mutex> g(mu);
        if (next >= n) break;
        i = order[next++];
      }
      mf_series* s = series[i];
      mf_status r = mf_run(s, &opts[i]);
      // Only the stack and the parameters are needed from here on.
      if (r == MF_OK) r = mf_series_trim(s);
      int bx = 0, by = 0;
      if (r == MF_OK && pool) {
        r = mf_correct(s, opts[i].out_bin, nullptr, &bx, &by);
        const size_t need = (size_t)bx * by * s->nz;
        if (r == MF_OK && need > out.size() && !out.resize(pool, need)) {
          s->last_error = "no host memory for the aligned stack";
          r = MF_ENOMEM;
        }
        if (r == MF_OK)
          r = mf_correct(s, opts[i].out_bin, out.data(), &bx, &by);
        if (r == MF_OK) r = mf_series_trim(s);  // the correction's buffer
      }
      st[i] = r;
      if (done)
        done(i, r, r == MF_OK && pool ? out.data() : nullptr, bx, by, s->nz,
             user);
    }
  };
  std::vector<std::thread> pool_threads;
  try {
    for (int w = 1; w < workers; w++) pool_threads.emplace_back(loop);
  } catch (const std::system_error&) {
    // Fewer host threads than asked for: the ones started and this thread
    // share the series.
  }
  loop();
  for (auto& t : pool_threads) t.join();
  for (int i = 0; i < n; i++)
    if (st[i] != MF_OK) return st[i];
  return MF_OK;
}

mf_status mf_batch_run(mf_series** series, int n, const mf_options* o,
                       int concurrency) {
  if (n == 0) return MF_OK;
  if (!series || n < 0 || !o) return MF_EINVAL;
  std::vector<mf_options> opts(std::max(n, 0), *o);
  return mf_batch_run_each(series, n, opts.data(), concurrency, 0, nullptr,
                           nullptr);
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
    {
      std::lock_guard<std::mutex> g(ctx->err_mu);
      ctx->last_error = err;
    }
    if (sv) mf_server_destroy(sv);  // joins any worker already started
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

namespace {

mf_status submit(mf_server* sv, int nx, int ny, int nz, const float* stack_zyx,
                 const float* tilts, const mf_options* o, bool borrow,
                 long long* id) {
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
    if (r->bytes > sv->budget) {
      std::lock_guard<std::mutex> g(sv->ctx->err_mu);
      sv->ctx->last_error = "the series' device bytes exceed the server's budget";
      return MF_ENOMEM;  // could never be admitted
    }
    if ((int)sv->queue.size() >= sv->queue_cap) return MF_EBUSY;
  }
  const size_t n = (size_t)nx * ny * nz;
  if (borrow) {
    r->borrowed = stack_zyx;
  } else {
    if (!r->stack.resize(sv->pool.get(), n)) {
      std::lock_guard<std::mutex> g(sv->ctx->err_mu);
      sv->ctx->last_error = "no host memory for the stack's copy";
      return MF_ENOMEM;
    }
    std::memcpy(r->stack.data(), stack_zyx, n * sizeof(float));
  }
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

}  // namespace

mf_status mf_submit(mf_server* sv, int nx, int ny, int nz,
                    const float* stack_zyx, const float* tilts,
                    const mf_options* o, long long* id) {
  return submit(sv, nx, ny, nz, stack_zyx, tilts, o, false, id);
}

mf_status mf_submit_borrowed(mf_server* sv, int nx, int ny, int nz,
                             const float* stack_zyx, const float* tilts,
                             const mf_options* o, long long* id) {
  return submit(sv, nx, ny, nz, stack_zyx, tilts, o, true, id);
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
