// TEST INFRASTRUCTURE: a serial CPU stand-in for the parts of the HIP runtime
// that libhfx uses, so that the unmodified library sources (host code, support
// kernels and generated class kernels) can be compiled with g++ into
// tests/cpuemu/libhfx_cpuemu.so and checked on a machine without a GPU.
//
// Every launch runs as one workgroup of one thread (HFX_WAVE = 1): the
// kernels are grid-stride loops, so one thread visits every work item.
// This library is loaded only by tests; the product (hfcode_amd/fock.py)
// loads the gfx950 build libhfx.so and nothing else.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifndef HFX_WAVE
#define HFX_WAVE 1
#define HFX_PIN(x) asm volatile("" : "+m"(x))
#endif

#define __global__
#define __device__
#define __host__
#define __shared__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __restrict__

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct int2 {
  int x, y;
};
inline int2 make_int2(int a, int b) { return int2{a, b}; }
struct int4 {
  int x, y, z, w;
};
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }

struct hfx_emu_idx {
  unsigned x, y, z;
};
static const hfx_emu_idx threadIdx = {0, 0, 0};
static const hfx_emu_idx blockIdx = {0, 0, 0};
static const hfx_emu_idx blockDim = {1, 1, 1};
static const hfx_emu_idx gridDim = {1, 1, 1};

inline void __syncthreads() {}
inline int __syncthreads_count(int p) { return p ? 1 : 0; }
#define __builtin_amdgcn_fence(order, scope) do { } while (0)
#define __builtin_amdgcn_wave_barrier() do { } while (0)
inline unsigned long long __ballot(bool p) { return p ? 1ull : 0ull; }
inline int __lane_id() { return 0; }
inline int __ffsll(long long m) { return m ? __builtin_ffsll(m) : 0; }
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline int __clzll(long long m) { return m ? __builtin_clzll((unsigned long long)m) : 64; }
template <typename T> inline T __shfl(T v, int) { return v; }
template <typename T> inline T __shfl_xor(T v, int) { return v; }
template <typename T> inline T __shfl_up(T v, unsigned) { return v; }
template <typename T> inline T __shfl_down(T v, unsigned) { return v; }
template <typename T> inline T atomicMax(T* p, T v) {
  T old = *p;
  if (v > old) *p = v;
  return old;
}
inline long long __double_as_longlong(double x) { long long r; std::memcpy(&r, &x, 8); return r; }
inline double __longlong_as_double(long long x) { double r; std::memcpy(&r, &x, 8); return r; }
template <typename T> inline T atomicAdd(T* p, T v) {
  T old = *p;
  *p = old + v;
  return old;
}
using std::fabs;
using std::fma;
using std::fmax;
using std::sqrt;
using std::min;
using std::max;

// dynamic LDS of the class kernels
extern char hfx_smem[];
#define HFX_EMU_SMEM_BYTES (1 << 20)

// ---- runtime API ------------------------------------------------------------
typedef int hipError_t;
#define hipSuccess 0
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
#define hipStreamNonBlocking 1
#define hipEventDisableTiming 2
struct hipDeviceProp_t {
  int multiProcessorCount;
  size_t sharedMemPerBlock;
};
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipMalloc(void** p, size_t n) {
  *p = std::calloc(1, n);
  return *p ? hipSuccess : 2;
}
template <typename T>
inline hipError_t hipMalloc(T** p, size_t n) {
  return hipMalloc((void**)p, n);
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  std::memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  p->multiProcessorCount = 1;
  p->sharedMemPerBlock = HFX_EMU_SMEM_BYTES;
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = (void*)1;
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = (void*)1;
  return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}
#define hipLaunchKernelGGL(k, g, b, shm, st, ...) k(__VA_ARGS__)
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) {
  *n = 1;
  return hipSuccess;
}

namespace hfx {
struct ClassArgs;
}
// Class kernels are launched through a type-erased pointer; all of them take
// one ClassArgs by value.
template <typename A = hfx::ClassArgs>
inline hipError_t hipLaunchKernel(const void* fn, dim3, dim3, void** args, size_t shm, hipStream_t) {
  if (shm > HFX_EMU_SMEM_BYTES) return 3;
  typedef void (*K)(const A);
  ((K)fn)(*(const A*)args[0]);
  return hipSuccess;
}
