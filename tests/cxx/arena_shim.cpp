// C entry points over tal::Arena (tal_sh_amd/csrc/pool.cpp) for the CPU test
// tests/test_pool.py, which builds this file together with pool.cpp using g++.
// The arena never dereferences its base, so a fake base address is used.
#include <cstddef>
#include <cstdint>

#include "pool.hpp"

extern "C" {

void *arena_new(size_t size, size_t align) {
  auto *a = new tal::Arena();
  a->init(reinterpret_cast<void *>(uintptr_t{1} << 40), size, align);
  return a;
}

void arena_delete(void *a) { delete static_cast<tal::Arena *>(a); }

// 0 / TRY_LATER / DEVICE_UNABLE; on success the offset from the base and the handle.
int arena_alloc(void *a, size_t bytes, size_t *offset, int *handle) {
  auto *ar = static_cast<tal::Arena *>(a);
  void *p = nullptr;
  const int e = ar->alloc(bytes, &p, handle);
  if (e == 0) *offset = static_cast<size_t>(static_cast<char *>(p) - static_cast<char *>(ar->base()));
  return e;
}

int arena_release(void *a, int handle) { return static_cast<tal::Arena *>(a)->release(handle); }
size_t arena_free_bytes(void *a) { return static_cast<tal::Arena *>(a)->free_bytes(); }
size_t arena_largest_free(void *a) { return static_cast<tal::Arena *>(a)->largest_free(); }
size_t arena_length(void *a, int handle) { return static_cast<tal::Arena *>(a)->length(handle); }
size_t arena_live(void *a) { return static_cast<tal::Arena *>(a)->live_entries(); }
size_t arena_size(void *a) { return static_cast<tal::Arena *>(a)->size(); }
}
