// Library context: device, stream, HBM accounting, kernel timing, workspaces.
#pragma once
#include <atomic>
#include <functional>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "pinned.h"

struct mf_ctx;

namespace mf {

class FftCache;

// Sums HIP-event-timed durations of named kernels (bench/roofline support).
class KernelTimer {
 public:
  ~KernelTimer();
  void enable(bool on) { enabled_ = on; }
  // Times only launches named 'name' (empty: all).
  void select(const std::string& name) { only_ = name; }
  // Times every k-th launch of each selected name (k >= 1); the others are
  // counted only, so that a sampled timing adds fewer stream markers.
  void every(int k) { every_ = k < 1 ? 1 : k; }
  bool enabled() const { return enabled_; }
  // Returns an event pair to record around one launch, or null when off.
  hipEvent_t* begin(const char* name, hipStream_t s);
  void end(hipEvent_t* ev, hipStream_t s);
  void reset();
  void query(const std::string& name, double* total_ms, long long* n);
  // Selected launches of 'name' since the last reset, timed or not.
  long long count(const std::string& name);

 private:
  struct Rec {
    std::string name;
    hipEvent_t ev[2];
  };
  bool enabled_ = false;
  int every_ = 1;
  std::string only_;
  std::map<std::string, long long> seen_;
  std::vector<Rec*> pending_;
  std::vector<Rec*> free_;  // event pairs reused across launches
  std::map<std::string, std::pair<double, long long>> done_;
  std::mutex mu_;
  void drain();
};

// Scope helper: times one launch when timing is enabled.
struct TimedLaunch {
  KernelTimer& t;
  hipStream_t s;
  hipEvent_t* ev;
  TimedLaunch(KernelTimer& t_, const char* name, hipStream_t s_)
      : t(t_), s(s_), ev(t_.begin(name, s_)) {}
  ~TimedLaunch() { t.end(ev, s); }
};

// A grow-only device buffer owned by the context or a series.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  // Grows the buffer to at least 'bytes'.  With the stream 'st' that last
  // used it, a smaller old block waits for that stream only and goes to the
  // context's cache; without, it is freed (hipFree waits for the device).
  void* ensure(mf_ctx* ctx, size_t bytes, hipStream_t st = nullptr);
  // 'cache': the caller has synchronised every stream that used the buffer,
  // so its block may go to the context's cache for reuse by any stream.
  void release(bool cache = false);
  void* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  mf_ctx* ctx_ = nullptr;
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Device memory through the context: budget and peak accounting in
// requested bytes, blocks in rounded sizes.  dev_free with 'cache' keeps the
// block (up to the context's cache_cap bytes in all) for a later dev_alloc
// of nearly its size, instead of returning it; the caller guarantees that
// no pending device work still uses it.  dev_flush returns
// every cached block (dev_alloc does so once when hipMalloc runs out).
void* dev_alloc(mf_ctx* ctx, size_t bytes);
void dev_free(mf_ctx* ctx, void* p, size_t bytes, bool cache = false);
void dev_flush(mf_ctx* ctx);

// Whether host address p lies in page-locked memory (hipHostMalloc), which
// a device-to-host copy can fill while the host goes on.
bool host_pinned(const void* p);

// Device copy of a constant host table, uploaded once per key and owned by
// the context (filters, tapers, trigonometric tables).  'build' runs only the
// first time a key is requested.
const void* ctx_table(mf_ctx* ctx, const std::string& key,
                      const std::function<std::vector<float>()>& build);

// Launch-shape tunable 'key' (for example "cc_rows/1024x1024/threads") as set
// through mf_tune_set, else 'dflt' (the built-in heuristic).
int tune_get(mf_ctx* ctx, const std::string& key, int dflt);

// Raises kernel k's dynamic-LDS limit on the context's device to at least
// 'bytes' (hipFuncSetAttribute once per kernel and size; kernels launched
// with more than 64 KB of dynamic LDS need it).
void ensure_dyn_lds(mf_ctx* ctx, const void* k, size_t bytes);

// Resident workgroups per CU of kernel k for a launch shape, queried once
// per (kernel, threads, dynamic LDS) and remembered.
int occupancy(mf_ctx* ctx, const void* k, int threads, size_t lds);

}  // namespace mf

struct mf_ctx {
  int device = 0;
  int num_cu = 0;         // compute units of the device (persistent grids)
  hipStream_t stream = nullptr;
  size_t budget = 0;      // bytes the context may hold (0: unlimited)
  size_t in_use = 0;      // bytes currently allocated through the context
  size_t peak = 0;        // largest in_use since creation or the last reset
  // Blocks of released series buffers kept for reuse (mf::dev_free), by
  // size; 'cached' bytes in all, at most cache_cap.
  std::multimap<size_t, void*> free_blocks;
  std::map<void*, size_t> block_size;  // live blocks' rounded sizes
  size_t cached = 0;
  size_t cache_cap = (size_t)16 << 30;
  std::string last_error;
  mf::KernelTimer timer;
  std::map<void*, size_t> user_allocs;  // mf_dev_malloc bookkeeping
  mf::DevBuf slab_ws;                   // materialised-slab workspace
  std::unique_ptr<mf::FftCache> fft;
  // Page-locked output blocks of the batched form (mf_batch_run_each),
  // kept between calls.
  std::unique_ptr<mf::PinnedPool> out_pool;
  std::map<std::string, std::unique_ptr<mf::DevBuf>> tables;
  std::map<std::string, int> tune;      // launch-shape tunables
  std::atomic<unsigned long long> tune_gen{0};  // changes to 'tune'
  // Device address of the allocation-free peak search's scratch array on
  // this context's device (cc.hip).
  void* peak_scratch = nullptr;
  // Per-kernel launch facts (ensure_dyn_lds, occupancy): host calls that
  // are not repeated per launch.
  std::map<const void*, size_t> dyn_lds;
  std::map<std::tuple<const void*, int, size_t>, int> occ;
  std::mutex mu, table_mu, tune_mu, launch_mu, err_mu;
  ~mf_ctx();
};
