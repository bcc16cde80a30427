#include <algorithm>
#include "context.h"
#include "pinned.h"

#include "fft.h"

namespace mf {

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e,
           hipGetErrorString(e), what, file, line);
  mf_status code = (e == hipErrorOutOfMemory) ? MF_ENOMEM : MF_EDEVICE;
  throw Error(code, buf);
}

KernelTimer::~KernelTimer() {
  for (Rec* r : pending_) free_.push_back(r);
  for (Rec* r : free_) {
    (void)hipEventDestroy(r->ev[0]);
    (void)hipEventDestroy(r->ev[1]);
    delete r;
  }
}

hipEvent_t* KernelTimer::begin(const char* name, hipStream_t s) {
  if (!enabled_ || (!only_.empty() && only_ != name)) return nullptr;
  Rec* r = nullptr;
  {
    std::lock_guard<std::mutex> g(mu_);
    if (seen_[name]++ % every_ != 0) return nullptr;
    if (!free_.empty()) {
      r = free_.back();
      free_.pop_back();
    }
  }
  if (!r) {
    r = new Rec;
    MF_HIP(hipEventCreate(&r->ev[0]));
    MF_HIP(hipEventCreate(&r->ev[1]));
  }
  r->name = name;
  MF_HIP(hipEventRecord(r->ev[0], s));
  std::lock_guard<std::mutex> g(mu_);
  pending_.push_back(r);
  return r->ev;
}

void KernelTimer::end(hipEvent_t* ev, hipStream_t s) {
  if (!ev) return;
  (void)hipEventRecord(ev[1], s);
}

void KernelTimer::drain() {
  for (Rec* r : pending_) {
    (void)hipEventSynchronize(r->ev[1]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r->ev[0], r->ev[1]);
    auto& d = done_[r->name];
    d.first += ms;
    d.second += 1;
    free_.push_back(r);
  }
  pending_.clear();
}

void KernelTimer::reset() {
  std::lock_guard<std::mutex> g(mu_);
  drain();
  done_.clear();
  seen_.clear();
}

long long KernelTimer::count(const std::string& name) {
  std::lock_guard<std::mutex> g(mu_);
  const auto it = seen_.find(name);
  return it == seen_.end() ? 0 : it->second;
}

void KernelTimer::query(const std::string& name, double* total_ms,
                        long long* n) {
  std::lock_guard<std::mutex> g(mu_);
  drain();
  auto it = done_.find(name);
  *total_ms = it == done_.end() ? 0.0 : it->second.first;
  *n = it == done_.end() ? 0 : it->second.second;
}

namespace {
// Block sizes: 2 MiB multiples from 1 MiB up, so that buffers of nearby
// sizes share cached blocks; 4 KiB multiples below.
size_t block_bytes(size_t b) {
  const size_t g = b >= ((size_t)1 << 20) ? ((size_t)2 << 20) : 4096;
  return (b + g - 1) / g * g;
}
}  // namespace

void* dev_alloc(mf_ctx* ctx, size_t bytes) {
  if (bytes == 0) bytes = 256;
  const size_t rb = block_bytes(bytes);
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
      throw Error(MF_ENOMEM, "HBM budget of the context exceeded");
    ctx->in_use += bytes;
    ctx->peak = std::max(ctx->peak, ctx->in_use);
    // A cached block of at least rb and at most 5/4 rb bytes.
    auto it = ctx->free_blocks.lower_bound(rb);
    if (it != ctx->free_blocks.end() && it->first <= rb + rb / 4) {
      void* p = it->second;
      ctx->block_size[p] = it->first;
      ctx->cached -= it->first;
      ctx->free_blocks.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, rb);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    dev_flush(ctx);  // the cache's blocks back to the device, then retry
    e = hipMalloc(&p, rb);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();  // thrown below; later checks must not see it
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->in_use -= bytes;
    throw_hip(e, "hipMalloc", __FILE__, __LINE__);
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->block_size[p] = rb;
  return p;
}

void dev_free(mf_ctx* ctx, void* p, size_t bytes, bool cache) {
  if (!p) return;
  if (bytes == 0) bytes = 256;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->in_use -= bytes;
    size_t sz = 0;
    auto it = ctx->block_size.find(p);
    if (it != ctx->block_size.end()) {
      sz = it->second;
      ctx->block_size.erase(it);
    }
    if (cache && sz > 0 && ctx->cached + sz <= ctx->cache_cap) {
      ctx->free_blocks.emplace(sz, p);
      ctx->cached += sz;
      return;
    }
  }
  (void)hipFree(p);
}

void dev_flush(mf_ctx* ctx) {
  std::multimap<size_t, void*> blocks;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    blocks.swap(ctx->free_blocks);
    ctx->cached = 0;
  }
  for (auto& kv : blocks) (void)hipFree(kv.second);
}

bool host_pinned(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // ordinary memory: no error for later calls
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

const void* ctx_table(mf_ctx* ctx, const std::string& key,
                      const std::function<std::vector<float>()>& build) {
  std::lock_guard<std::mutex> g(ctx->table_mu);
  auto it = ctx->tables.find(key);
  if (it != ctx->tables.end()) return it->second->get();
  const std::vector<float> host = build();
  const size_t bytes = host.size() * sizeof(float);
  std::unique_ptr<DevBuf> b(new DevBuf);
  void* d = b->ensure(ctx, bytes);
  MF_HIP(hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice));
  ctx->tables[key] = std::move(b);
  return d;
}

int tune_get(mf_ctx* ctx, const std::string& key, int dflt) {
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  auto it = ctx->tune.find(key);
  return it == ctx->tune.end() ? dflt : it->second;
}

void ensure_dyn_lds(mf_ctx* ctx, const void* k, size_t bytes) {
  if (bytes <= 64 * 1024) return;
  std::lock_guard<std::mutex> g(ctx->launch_mu);
  size_t& have = ctx->dyn_lds[k];
  if (have >= bytes) return;
  MF_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes));
  have = bytes;
}

int occupancy(mf_ctx* ctx, const void* k, int threads, size_t lds) {
  std::lock_guard<std::mutex> g(ctx->launch_mu);
  const auto key = std::make_tuple(k, threads, lds);
  auto it = ctx->occ.find(key);
  if (it != ctx->occ.end()) return it->second;
  int occ = 0;
  MF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, threads, lds));
  ctx->occ[key] = occ;
  return occ;
}

DevBuf::~DevBuf() { release(); }

void* DevBuf::ensure(mf_ctx* ctx, size_t bytes, hipStream_t st) {
  if (p_ && n_ >= bytes) return p_;
  if (p_ && st) {
    // Only this buffer's own stream is waited for: a hipFree here would wait
    // for every stream of the device, another series' downloads included.
    MF_HIP(hipStreamSynchronize(st));
    dev_free(ctx_, p_, n_, true);
    p_ = nullptr;
    n_ = 0;
  }
  release();
  ctx_ = ctx;
  p_ = dev_alloc(ctx, bytes);
  n_ = bytes;
  return p_;
}

void DevBuf::release(bool cache) {
  if (p_) {
    if (ctx_) (void)hipStreamSynchronize(ctx_->stream);
    dev_free(ctx_, p_, n_, cache);
  }
  p_ = nullptr;
  n_ = 0;
}

}  // namespace mf

mf_ctx::~mf_ctx() {
  if (stream) (void)hipStreamSynchronize(stream);
  slab_ws.release();
  tables.clear();
  fft.reset();
  out_pool.reset();
  for (auto& kv : user_allocs) (void)hipFree(kv.first);
  user_allocs.clear();
  mf::dev_flush(this);
  timer.reset();
  if (stream) (void)hipStreamDestroy(stream);
}
