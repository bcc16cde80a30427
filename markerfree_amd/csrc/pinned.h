// Page-locked host blocks recycled between requests (served form) and
// between batched series (the batched form's output buffers).
#pragma once
#include <cstddef>
#include <map>
#include <mutex>

#include "markerfree/mf_api.h"

namespace mf {

// Free page-locked blocks kept for reuse, at most 'cap' of them and at most
// kMaxHeldBytes in all (page-locked memory is a limited system resource).  The first
// host write to newly page-locked memory costs about 0.3 ms per MB (page
// population), several times a plain copy, so a request's stack copy and
// its aligned stack go into recycled blocks when one of a fitting size is
// free (no more than twice the size asked for).  The server sets the cap to
// the number of blocks its admitted and queued requests can hold at once
// (two per running request, one per queued request): the pool keeps no more
// blocks than the requests themselves can hold.
class PinnedPool {
 public:
  static constexpr size_t kMaxHeldBytes = (size_t)32 << 30;
  explicit PinnedPool(size_t cap) : cap_(cap) {}
  // Raises the number of free blocks kept to at least 'cap'.
  void keep_at_least(size_t cap) {
    std::lock_guard<std::mutex> lk(mu_);
    if (cap > cap_) cap_ = cap;
  }
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
        held_ -= it->first * sizeof(float);
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
      if (free_.size() < cap_ && held_ + n * sizeof(float) <= kMaxHeldBytes) {
        free_.emplace(n, p);
        held_ += n * sizeof(float);
        return;
      }
    }
    mf_host_free(p);
  }

 private:
  std::mutex mu_;
  std::multimap<size_t, float*> free_;  // size in floats -> block
  size_t cap_, held_ = 0;
};

}  // namespace mf
