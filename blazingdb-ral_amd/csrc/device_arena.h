/*
 * device_arena.h — device memory for the serving loop: large chunks taken from the
 * driver once, carved first-fit with coalescing of freed neighbours.
 *
 * The serving loop allocates every admitted query's outputs and workspace and
 * frees them when the query is released; through the stream-ordered pool each of
 * those calls cost 0.1-0.2 ms (5000 queries per second spent 2.3 s per second in
 * the allocator). Here an allocation is a map lookup under a lock. A free is
 * immediate: the loop frees a query's memory only after the query's batch has
 * completed on the device and the caller has released it, so no kernel still uses
 * it. Chunks go back to the driver when the arena is destroyed.
 */
#ifndef BLZ_DEVICE_ARENA_H_
#define BLZ_DEVICE_ARENA_H_

#include <stddef.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace blz {

class DeviceArena {
 public:
  // chunk_bytes: the driver allocation unit (a larger request gets a chunk of its own size).
  explicit DeviceArena(size_t chunk_bytes);
  ~DeviceArena();
  DeviceArena(const DeviceArena &) = delete;
  DeviceArena &operator=(const DeviceArena &) = delete;

  // 256-byte aligned; nullptr when the device has no memory for a new chunk.
  void *alloc(size_t bytes);
  // `p` must come from alloc and must not be in use by any pending device work.
  void free(void *p);
  size_t reserved() const;  // bytes taken from the driver
  size_t in_use() const;    // bytes handed out

 private:
  struct Chunk {
    uint8_t *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_blocks;  // offset -> bytes, disjoint, not adjacent
  };
  mutable std::mutex mu_;
  size_t chunk_bytes_;
  std::vector<Chunk> chunks_;
  std::unordered_map<const void *, std::pair<size_t, size_t>> live_;  // pointer -> (chunk, bytes)
  size_t reserved_ = 0, in_use_ = 0;
};

}  // namespace blz
#endif
