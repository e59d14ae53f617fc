// Library-internal services shared by the translation units of libhfx.
#pragma once
#include <cstddef>

#include "../../include/hfx.h"

namespace hfx {
// F += G(P) on device arrays (as hfx_fock_nd_device) with the FP32 split of
// this call (0: every quartet in FP64) instead of the context's setting.
int fock_nd_device_split(hfx_ctx* ctx, const double* dP, double* dF, double thr, int mode, void* stream,
                         double split);
// The context's HFX_OPT_MIXED_SPLIT.
double ctx_split(const hfx_ctx* ctx);
// Device memory from the per-process cache of freed blocks (hfx_host.hip).
// cached_free: the caller guarantees that no queued work still uses the block.
void* cached_alloc(size_t bytes);
void cached_free(void* p);
}  // namespace hfx
