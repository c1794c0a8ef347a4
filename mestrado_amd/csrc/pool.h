// Host-side allocators of device memory regions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <vector>

namespace mst {

// First-fit allocator over offsets [0, cap) in units of `align`.
struct Pool {
  int64_t cap = 0, align = 1, used = 0;
  std::map<int64_t, int64_t> free_;  // offset -> size
  void init(int64_t c, int64_t a) {
    cap = c;
    align = a;
    used = 0;
    free_.clear();
    if (c > 0) free_[0] = c;
  }
  int64_t round(int64_t n) const { return (n + align - 1) / align * align; }
  // offset of a block of n, or -1
  int64_t alloc(int64_t n) {
    n = round(std::max<int64_t>(n, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t off = it->first, rest = it->second - n;
        free_.erase(it);
        if (rest > 0) free_[off + n] = rest;
        used += n;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    n = round(std::max<int64_t>(n, 1));
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  // gives back the tail of a block of n beyond its first keep bytes
  void shrink(int64_t off, int64_t n, int64_t keep) {
    const int64_t a = round(std::max<int64_t>(n, 1)), b = round(std::max<int64_t>(keep, 1));
    if (b < a) release(off + b, a - b);
  }
};

// Device blocks carved from regions allocated on demand (hipMalloc of at
// least `chunk` bytes each); everything is freed with the object.
class GrowPool {
 public:
  explicit GrowPool(int64_t chunk) : chunk_(chunk) {}
  GrowPool(const GrowPool&) = delete;
  GrowPool& operator=(const GrowPool&) = delete;
  ~GrowPool() {
    for (auto& r : regions_) hipFree(r.base);
  }
  uint8_t* alloc(int64_t n) {
    for (auto& r : regions_) {
      const int64_t off = r.pool.alloc(n);
      if (off >= 0) return r.base + off;
    }
    Region r;
    const int64_t cap = std::max<int64_t>(chunk_, (n + 255) / 256 * 256);
    if (hipMalloc(&r.base, cap) != hipSuccess) return nullptr;
    r.pool.init(cap, 256);
    regions_.push_back(r);
    return regions_.back().base + regions_.back().pool.alloc(n);
  }
  void release(uint8_t* p, int64_t n) {
    for (auto& r : regions_)
      if (p >= r.base && p < r.base + r.pool.cap) return r.pool.release(p - r.base, n);
  }
  void shrink(uint8_t* p, int64_t n, int64_t keep) {
    for (auto& r : regions_)
      if (p >= r.base && p < r.base + r.pool.cap) return r.pool.shrink(p - r.base, n, keep);
  }

 private:
  struct Region {
    uint8_t* base = nullptr;
    Pool pool;
  };
  int64_t chunk_;
  std::vector<Region> regions_;
};

}  // namespace mst
