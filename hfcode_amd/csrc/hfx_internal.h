// Library-internal services shared by the translation units of libhfx.
#pragma once
#include <cstddef>

namespace hfx {
// Device memory from the per-process cache of freed blocks (hfx_host.hip).
// cached_free: the caller guarantees that no queued work still uses the block.
void* cached_alloc(size_t bytes);
void cached_free(void* p);
}  // namespace hfx
