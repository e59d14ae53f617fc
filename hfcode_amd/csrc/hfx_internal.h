// Library-internal services shared by the translation units of libhfx.
#pragma once
#include <cstddef>
#include <mutex>
#include <shared_mutex>
#include <vector>

#include "../../include/hfx.h"

namespace hfx {
// Stream capture and device-wide operations.  While a thread captures a HIP
// graph from a stream, another thread's synchronous operation on the null
// stream or implicit device synchronisation (hipMemcpy, hipFree, a library's
// workspace reallocation) would break that capture.  The library issues its
// copies on its own streams, and takes this lock around the rest: shared by
// capture sections (captures on different streams may overlap), exclusive
// for the device-wide operations.  A thread inside a capture section does
// not allocate (cached_alloc returns null there).
std::shared_mutex& capture_mutex();
extern thread_local bool t_capturing;
struct CaptureSection {
  std::shared_lock<std::shared_mutex> lk;
  CaptureSection() : lk(capture_mutex()) { t_capturing = true; }
  ~CaptureSection() { t_capturing = false; }
};
struct DeviceSection {
  std::unique_lock<std::shared_mutex> lk;
  DeviceSection() : lk(capture_mutex()) {}
};
// F += G(P) on device arrays (as hfx_fock_nd_device) with the FP32 split of
// this call (0: every quartet in FP64) instead of the context's setting.
int fock_nd_device_split(hfx_ctx* ctx, const double* dP, double* dF, double thr, int mode, void* stream,
                         double split);
// The context's HFX_OPT_MIXED_SPLIT.
double ctx_split(const hfx_ctx* ctx);
// rocBLAS handles (rocblas_handle) from a per-process pool: creating and
// destroying one allocates and frees its device workspace, and hipFree waits
// for the whole device.  A handle is used by one caller at a time.
void* take_blas_handle();
void give_blas_handle(void* h);
// Molecules, matrix stride and each molecule's basis size.
int ctx_dims(const hfx_ctx* ctx, int* nmol, int* nb_stride, std::vector<int>* nbs);
// Device memory from the per-process cache of freed blocks (hfx_host.hip).
// cached_free: the caller guarantees that no queued work still uses the block.
void* cached_alloc(size_t bytes);
void cached_free(void* p);
}  // namespace hfx
