// hostpool.hpp - pinned host memory for result pages, recycled across results.
#pragma once
#include <cstdint>

namespace hj {
// a block of at least `bytes` bytes of pinned host memory (the copy engines write result pages
// straight into it), from the pool's free blocks or newly allocated; plain malloc'd memory when
// pinned memory cannot be had
uint8_t* host_pages_alloc(uint64_t bytes);
// a block from host_pages_alloc back to the pool (any other pointer is free'd)
void host_pages_free(void* p);
}  // namespace hj
