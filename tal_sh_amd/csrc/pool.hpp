// Buffer pool: a best-fit, coalescing sub-allocator over one large region.
//
// Replaces the reference's multi-level argument-buffer tree
// (mem_manager.cpp:175-307 arg_buf_allocate, :544-692 get/free_buf_entry),
// which hands out whole power-of-two-ish blocks (25 % average / 50 % worst
// internal fragmentation, DOC/MemManager_manual.txt:13-21) and caps a tensor
// at 1/6 of the buffer.  Here a request takes exactly its size rounded up to
// the alignment, free neighbours are merged on release, and the only size
// limit is the region itself.  The TAL-SH contract is kept:
//   TRY_LATER      the request fits the region but no free range is large
//                  enough right now (retry after other entries are freed);
//   DEVICE_UNABLE  the request can never fit this region.
// Handles are small non-negative integers (TAL-SH's buf_entry).  Not
// thread-safe by itself; the runtime serialises calls.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <vector>

namespace tal {

class Arena {
 public:
  void init(void *base, size_t size, size_t align);
  void reset();
  // 0 on success (ptr, handle set), TRY_LATER or DEVICE_UNABLE otherwise.
  int alloc(size_t bytes, void **ptr, int *handle);
  // 0 on success, -1 for an unknown handle.
  int release(int handle);
  // Bytes held by a live handle (its request rounded up to the alignment), 0 otherwise.
  size_t length(int handle) const;
  bool owns(const void *p) const;
  void *base() const { return base_; }
  size_t size() const { return size_; }
  size_t free_bytes() const { return free_total_; }
  size_t largest_free() const;
  size_t live_entries() const { return live_; }

 private:
  struct Slot {
    size_t off = 0, len = 0;
    bool used = false;
  };
  void insert_free(size_t off, size_t len);
  void erase_free(std::map<size_t, size_t>::iterator it);

  char *base_ = nullptr;
  size_t size_ = 0, align_ = 256, free_total_ = 0, live_ = 0;
  std::map<size_t, size_t> by_off_;            // free ranges: offset -> length
  std::multimap<size_t, size_t> by_len_;       // free ranges: length -> offset
  std::vector<Slot> slots_;
  std::vector<int> free_slots_;
};

}  // namespace tal
