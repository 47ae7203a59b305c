// pool.cpp -- error reporting and the pool allocator (host bookkeeping).
#include <cstdlib>
#include <cstring>

#include "mbx_internal.hpp"

namespace mbx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int status, const std::string &msg) {
  set_error(msg);
  return status;
}

int hip_fail(hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  set_error(m);
  return e == hipErrorOutOfMemory ? MBX_ERR_OOM : MBX_ERR_HIP;
}

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

Pool::Pool(bool device, size_t fixed_capacity)
    : device_(device), fixed_(fixed_capacity != 0) {
  if (fixed_) init_status_ = add_arena(round_up(fixed_capacity, kAlign));
}

Pool::~Pool() {
  for (auto &a : arenas_) raw_free(a.base);
}

void *Pool::raw_alloc(size_t size) {
  if (device_) {
    void *p = nullptr;
    if (hipMalloc(&p, size) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    return p;
  }
  return std::aligned_alloc(kAlign, size);
}

void Pool::raw_free(void *p) {
  if (!p) return;
  if (device_) (void)hipFree(p);
  else std::free(p);
}

int Pool::add_arena(size_t size) {
  char *p = static_cast<char *>(raw_alloc(size));
  if (!p) return fail(MBX_ERR_OOM, "pool: arena allocation of " +
                                       std::to_string(size) + " bytes failed");
  arenas_.push_back({p, size});
  free_[p] = size;
  capacity_ += size;
  return MBX_OK;
}

int Pool::alloc(size_t bytes, void **out) {
  if (!out) return fail(MBX_ERR_INVALID, "pool alloc: null out");
  *out = nullptr;
  std::lock_guard<std::mutex> g(mu_);
  const size_t need = round_up(bytes ? bytes : 1, kAlign);
  for (int attempt = 0; attempt < 2; ++attempt) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      char *p = it->first;
      const size_t rest = it->second - need;
      free_.erase(it);
      if (rest) free_[p + need] = rest;
      live_[p] = need;
      used_ += need;
      if (used_ > peak_) peak_ = used_;
      *out = p;
      return MBX_OK;
    }
    if (fixed_ || attempt) break;
    int st = add_arena(need > kArena ? round_up(need, kAlign) : kArena);
    if (st != MBX_OK) return st;
  }
  return fail(MBX_ERR_OOM, "pool exhausted: need " + std::to_string(need) +
                               " bytes, " + std::to_string(capacity_ - used_) +
                               " free of " + std::to_string(capacity_));
}

int Pool::free(void *vp) {
  if (!vp) return MBX_OK;
  std::lock_guard<std::mutex> g(mu_);
  char *p = static_cast<char *>(vp);
  auto it = live_.find(p);
  if (it == live_.end())
    return fail(MBX_ERR_INVALID, "pool free: pointer not allocated here");
  size_t size = it->second;
  live_.erase(it);
  used_ -= size;
  // coalesce with neighbours inside the same arena
  auto next = free_.lower_bound(p);
  if (next != free_.end() && p + size == next->first) {
    bool same_arena = false;
    for (auto &a : arenas_)
      if (p >= a.base && next->first < a.base + a.size) same_arena = true;
    if (same_arena) {
      size += next->second;
      next = free_.erase(next);
    }
  }
  if (next != free_.begin()) {
    auto prev = std::prev(next);
    bool same_arena = false;
    for (auto &a : arenas_)
      if (prev->first >= a.base && p < a.base + a.size) same_arena = true;
    if (same_arena && prev->first + prev->second == p) {
      prev->second += size;
      return MBX_OK;
    }
  }
  free_[p] = size;
  return MBX_OK;
}

void Pool::stats(size_t *used, size_t *capacity, size_t *peak, size_t *nlive,
                 size_t *largest_free) const {
  std::lock_guard<std::mutex> g(mu_);
  if (used) *used = used_;
  if (capacity) *capacity = capacity_;
  if (peak) *peak = peak_;
  if (nlive) *nlive = live_.size();
  if (largest_free) {
    size_t m = 0;
    for (auto &f : free_) m = f.second > m ? f.second : m;
    *largest_free = m;
  }
}

}  // namespace mbx

struct mbx_hostpool {
  mbx::Pool pool;
  explicit mbx_hostpool(size_t cap) : pool(false, cap) {}
};

extern "C" {

const char *mbx_last_error(void) { return mbx::g_last_error.c_str(); }

const char *mbx_strerror(int s) {
  switch (s) {
  case MBX_OK: return "ok";
  case MBX_ERR_INVALID: return "invalid argument";
  case MBX_ERR_OOM: return "out of memory";
  case MBX_ERR_HIP: return "HIP runtime error";
  case MBX_ERR_ALIGN: return "bitmap pointer not 8-byte aligned";
  case MBX_ERR_CAPACITY: return "capacity exceeded";
  case MBX_ERR_NODEVICE: return "no usable GPU";
  case MBX_ERR_ADMISSION: return "refused by memory admission";
  case MBX_ERR_UNSUPPORTED: return "no dedicated kernel for the program's shape";
  default: return "unknown status";
  }
}

int mbx_hostpool_create(size_t capacity, mbx_hostpool **out) {
  if (!out || capacity == 0)
    return mbx::fail(MBX_ERR_INVALID, "hostpool: capacity must be > 0");
  auto *p = new mbx_hostpool(capacity);
  if (p->pool.init_status() != MBX_OK) {
    delete p;
    return MBX_ERR_OOM;
  }
  *out = p;
  return MBX_OK;
}

int mbx_hostpool_alloc(mbx_hostpool *p, size_t bytes, void **ptr) {
  if (!p) return mbx::fail(MBX_ERR_INVALID, "hostpool: null");
  return p->pool.alloc(bytes, ptr);
}

int mbx_hostpool_free(mbx_hostpool *p, void *ptr) {
  if (!p) return mbx::fail(MBX_ERR_INVALID, "hostpool: null");
  return p->pool.free(ptr);
}

int mbx_hostpool_stats(mbx_hostpool *p, size_t *used, size_t *capacity,
                       size_t *peak, size_t *nlive, size_t *largest_free) {
  if (!p) return mbx::fail(MBX_ERR_INVALID, "hostpool: null");
  p->pool.stats(used, capacity, peak, nlive, largest_free);
  return MBX_OK;
}

int mbx_hostpool_destroy(mbx_hostpool *p) {
  delete p;
  return MBX_OK;
}

}  // extern "C"
