/*
 * device_arena.cpp — first-fit carving of driver chunks (device_arena.h).
 */
#include "device_arena.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace blz {

namespace {
constexpr size_t kAlign = 256;
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
}  // namespace

DeviceArena::DeviceArena(size_t chunk_bytes) : chunk_bytes_(align_up(std::max<size_t>(chunk_bytes, kAlign), kAlign)) {}

DeviceArena::~DeviceArena() {
  for (Chunk &c : chunks_) (void)hipFree(c.base);
}

void *DeviceArena::alloc(size_t bytes) {
  const size_t need = align_up(std::max<size_t>(bytes, 1), kAlign);
  std::lock_guard<std::mutex> g(mu_);
  for (size_t ci = 0; ci < chunks_.size(); ci++) {
    Chunk &c = chunks_[ci];
    for (auto it = c.free_blocks.begin(); it != c.free_blocks.end(); ++it) {
      if (it->second < need) continue;
      const size_t off = it->first, left = it->second - need;
      c.free_blocks.erase(it);
      if (left) c.free_blocks.emplace(off + need, left);
      uint8_t *p = c.base + off;
      live_.emplace(p, std::make_pair(ci, need));
      in_use_ += need;
      return p;
    }
  }
  Chunk c;
  c.size = std::max(chunk_bytes_, align_up(need, 2u << 20));
  if (hipMalloc((void **)&c.base, c.size) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (c.size > need) c.free_blocks.emplace(need, c.size - need);
  chunks_.push_back(c);
  reserved_ += c.size;
  live_.emplace(c.base, std::make_pair(chunks_.size() - 1, need));
  in_use_ += need;
  return c.base;
}

void DeviceArena::free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  Chunk &c = chunks_[it->second.first];
  size_t off = (size_t)((uint8_t *)p - c.base), len = it->second.second;
  in_use_ -= len;
  live_.erase(it);
  // merge with the free neighbours on either side
  auto next = c.free_blocks.lower_bound(off);
  if (next != c.free_blocks.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      off = prev->first;
      len += prev->second;
      c.free_blocks.erase(prev);
    }
  }
  if (next != c.free_blocks.end() && off + len == next->first) {
    len += next->second;
    c.free_blocks.erase(next);
  }
  c.free_blocks.emplace(off, len);
}

size_t DeviceArena::reserved() const {
  std::lock_guard<std::mutex> g(mu_);
  return reserved_;
}

size_t DeviceArena::in_use() const {
  std::lock_guard<std::mutex> g(mu_);
  return in_use_;
}

}  // namespace blz
