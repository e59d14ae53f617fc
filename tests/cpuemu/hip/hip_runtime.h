// TEST INFRASTRUCTURE: a serial CPU stand-in for the parts of the HIP runtime
// that libhfx uses, so that the unmodified library sources (host code, support
// kernels and generated class kernels) can be compiled with g++ into
// tests/cpuemu/libhfx_cpuemu.so and checked on a machine without a GPU.
//
// Every launch runs as one workgroup of one thread (HFX_WAVE = 1) per
// blockIdx.y (the kernels are grid-stride loops in x, so one thread visits
// every work item); built with HFX_EMU_LANES, as one workgroup of the
// launch's blockDim.x lanes in 64-lane waves (below).  Work launched while a stream is being captured is
// recorded instead and run again by hipGraphLaunch, as a HIP graph would.
// This library is loaded only by tests; the product (hfcode_amd/fock.py)
// loads the gfx950 build libhfx.so and nothing else.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#ifdef HFX_EMU_LANES
// Lanes mode (tests/cpuemu/libhfx_lanes.so): every workgroup runs as
// blockDim.x lanes, 64 to a wave, each lane a fiber on the calling host
// thread.  A lane runs until it reaches a wave collective (ballot, shuffle,
// wave barrier) or a block barrier; the collective completes once every
// lane of the wave (block) that has not exited waits at the same call site,
// with the GPU's results (exited lanes are inactive: no ballot bit).  Lanes
// of a wave at different collectives, or a shuffle reading an exited lane,
// abort with a message.  So code in which lanes see each other (ballot
// compaction, __syncthreads_count votes, LDS written by one lane and read by
// another between wave barriers) runs with its 64-lane semantics.
#define HFX_WAVE 64
#else
#define HFX_WAVE 1
#endif
#define HFX_PIN(x) asm volatile("" : "+m"(x))
#define HFX_PIN_GPTR(x) HFX_PIN(x)
#define HFX_RCPF(x) (1.0f / (x))
#define HFX_RSQF(x) (1.0f / std::sqrt(x))
#define HFX_EXP2F(x) std::exp2(x)

#define __global__
#define __device__
#define __host__
// LDS: one copy per host thread (the sources declare it static or extern)
#define __shared__ __thread
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __restrict__
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

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
inline int __ffsll(long long m) { return m ? __builtin_ffsll(m) : 0; }
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline int __clzll(long long m) { return m ? __builtin_clzll((unsigned long long)m) : 64; }
inline long long __double_as_longlong(double x) { long long r; std::memcpy(&r, &x, 8); return r; }
inline double __longlong_as_double(long long x) { double r; std::memcpy(&r, &x, 8); return r; }
#ifndef HFX_EMU_LANES
static const hfx_emu_idx threadIdx = {0, 0, 0};
static const hfx_emu_idx blockDim = {1, 1, 1};
extern thread_local hfx_emu_idx blockIdx;   // x = 0; y runs over the launch's grid.y
extern thread_local hfx_emu_idx gridDim;    // x = 1

inline void __syncthreads() {}
inline int __syncthreads_count(int p) { return p ? 1 : 0; }
#define __builtin_amdgcn_fence(order, scope) do { } while (0)
#define __builtin_amdgcn_wave_barrier() do { } while (0)
inline unsigned long long __ballot(bool p) { return p ? 1ull : 0ull; }
inline int __lane_id() { return 0; }
template <typename T> inline T __shfl(T v, int) { return v; }
template <typename T> inline T __shfl_xor(T v, int) { return v; }
inline int __builtin_amdgcn_update_dpp(int old, int, int, int, int, bool) { return old; }
inline int __builtin_amdgcn_readlane(int v, int) { return v; }
template <typename T> inline T __shfl_up(T v, unsigned) { return v; }
template <typename T> inline T __shfl_down(T v, unsigned) { return v; }
template <typename T> inline T atomicMax(T* p, T v) {
  T old = *p;
  if (v > old) *p = v;
  return old;
}
template <typename T> inline T atomicAdd(T* p, T v) {
  T old = *p;
  *p = old + v;
  return old;
}
#else
extern thread_local hfx_emu_idx threadIdx;
extern thread_local hfx_emu_idx blockDim;
extern thread_local hfx_emu_idx blockIdx;   // x = 0; y runs over the launch's grid.y
extern thread_local hfx_emu_idx gridDim;    // x = 1
enum { HFX_LC_BALLOT = 1, HFX_LC_SHFL, HFX_LC_WAVE_BARRIER, HFX_LC_SYNC, HFX_LC_SYNC_COUNT };
// the calling lane's part in a collective at source line `site`; returns
// its result once the collective has completed
unsigned long long hfx_lane_collective(int kind, int site, unsigned long long v, int src);
// runs f as one workgroup of nt lanes (blockIdx.y = y of gy)
void hfx_lane_run_block(unsigned nt, unsigned y, unsigned gy, const std::function<void()>& f);
inline int __lane_id() { return (int)(threadIdx.x & 63u); }
template <typename T>
inline T hfx_lane_shfl(T v, int src, int site) {
  static_assert(sizeof(T) <= 8, "shuffle of more than 8 bytes");
  unsigned long long b = 0;
  std::memcpy(&b, &v, sizeof(T));
  b = hfx_lane_collective(HFX_LC_SHFL, site, b, src & 63);
  T r;
  std::memcpy(&r, &b, sizeof(T));
  return r;
}
// DPP row_shr:k (ctrl 0x110 + k): lane l reads lane l - k of its row of 16
// lanes, lanes with no such source keep `old`; readlane: lane l's value
inline int hfx_lane_dpp(int old, int src, int ctrl, int site) {
  if (ctrl <= 0x110 || ctrl > 0x11F) {
    std::fprintf(stderr, "emulated DPP control 0x%x not supported\n", ctrl);
    std::abort();
  }
  const int k = ctrl - 0x110, l = __lane_id();
  const bool ok = (l & 15) >= k;
  const int v = hfx_lane_shfl(src, ok ? l - k : l, site);
  return ok ? v : old;
}
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) hfx_lane_dpp((old), (src), (ctrl), __LINE__)
#define __builtin_amdgcn_readlane(v, l) hfx_lane_shfl((int)(v), (int)(l), __LINE__)
#define __ballot(p) hfx_lane_collective(HFX_LC_BALLOT, __LINE__, (p) ? 1ull : 0ull, 0)
#define __shfl(v, s) hfx_lane_shfl((v), (int)(s), __LINE__)
#define __shfl_xor(v, m) hfx_lane_shfl((v), __lane_id() ^ (int)(m), __LINE__)
#define __syncthreads() ((void)hfx_lane_collective(HFX_LC_SYNC, __LINE__, 0, 0))
#define __syncthreads_count(p) ((int)hfx_lane_collective(HFX_LC_SYNC_COUNT, __LINE__, (p) ? 1ull : 0ull, 0))
#define __builtin_amdgcn_fence(order, scope) do { } while (0)
#define __builtin_amdgcn_wave_barrier() ((void)hfx_lane_collective(HFX_LC_WAVE_BARRIER, __LINE__, 0, 0))
// lanes run one at a time on one host thread: plain read-modify-writes
template <typename T> inline T atomicMax(T* p, T v) {
  T old = *p;
  if (v > old) *p = v;
  return old;
}
template <typename T> inline T atomicAdd(T* p, T v) {
  T old = *p;
  *p = old + v;
  return old;
}
#endif
using std::fabs;
using std::fma;
using std::fmax;
using std::sqrt;
using std::min;
using std::max;

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
  int maxThreadsPerMultiProcessor;
};
// no scratch memory on the CPU: kernels report none and the limits are unsupported
struct hipFuncAttributes {
  size_t localSizeBytes;
};
enum hipLimit_t { hipLimitStackSize = 0, hipExtLimitScratchMin = 0x1000, hipExtLimitScratchMax = 0x1001,
                  hipExtLimitScratchCurrent = 0x1002 };
inline hipError_t hipFuncGetAttributes(hipFuncAttributes* a, const void*) {
  a->localSizeBytes = 0;
  return 0;
}
inline hipError_t hipDeviceGetLimit(size_t*, hipLimit_t) { return 1; }
inline hipError_t hipDeviceSetLimit(hipLimit_t, size_t) { return 1; }
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
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
  *p = const_cast<void*>(symbol);
  return hipSuccess;
}
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
// stream capture: the recorded operations of the graph being captured (per thread)
typedef std::vector<std::function<void()>> hfx_emu_ops;
inline hfx_emu_ops*& hfx_emu_capture() {
  static thread_local hfx_emu_ops* cur = nullptr;
  return cur;
}
// runs f once per blockIdx.y of the grid (lanes mode: a kernel as `threads`
// lanes per block; threads = 1 is the stand-in libraries' host work), or
// records that for a graph
inline void hfx_emu_launch(unsigned gy, std::function<void()> f, unsigned threads = 1) {
  auto run = [gy, f, threads]() {
#ifdef HFX_EMU_LANES
    if (threads > 1) {
      for (unsigned y = 0; y < gy; ++y) hfx_lane_run_block(threads, y, gy, f);
      return;
    }
#else
    (void)threads;
#endif
    const hfx_emu_idx b = blockIdx, g = gridDim;
    gridDim = hfx_emu_idx{1, gy, 1};
    for (unsigned y = 0; y < gy; ++y) {
      blockIdx = hfx_emu_idx{0, y, 0};
      f();
    }
    blockIdx = b;
    gridDim = g;
  };
  if (hfx_emu_ops* c = hfx_emu_capture()) c->push_back(run);
  else run();
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  if (hfx_emu_ops* c = hfx_emu_capture()) {
    c->push_back([=]() { std::memmove(d, s, n); });
    return hipSuccess;
  }
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  if (hfx_emu_ops* c = hfx_emu_capture()) {
    c->push_back([=]() { std::memset(d, v, n); });
    return hipSuccess;
  }
  std::memset(d, v, n);
  return hipSuccess;
}
// graphs
struct hfx_emu_graph {
  hfx_emu_ops ops;
};
typedef hfx_emu_graph* hipGraph_t;
typedef hfx_emu_graph* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  if (hfx_emu_capture()) return 4;
  hfx_emu_capture() = new hfx_emu_ops;
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  hfx_emu_ops* c = hfx_emu_capture();
  if (!c) return 4;
  *g = new hfx_emu_graph{std::move(*c)};
  delete c;
  hfx_emu_capture() = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t) {
  *e = new hfx_emu_graph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& f : e->ops) f();
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  p->multiProcessorCount = 1;
  p->sharedMemPerBlock = HFX_EMU_SMEM_BYTES;
  p->maxThreadsPerMultiProcessor = 64;
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
// the arguments are evaluated (and copied) at the launch, as HIP does
template <typename K, typename... A>
inline void hfx_emu_kernel(dim3 g, dim3 b, K k, A... a) {
  hfx_emu_launch(g.y, [=]() { k(a...); }, b.x);
}
#define hipLaunchKernelGGL(k, g, b, shm, st, ...) hfx_emu_kernel(dim3(g), dim3(b), k, __VA_ARGS__)
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) {
  *n = 1;
  return hipSuccess;
}

// Class and digestion kernels are launched through a type-erased pointer with
// one argument struct by value (ClassArgs or DigestArgs).  The stand-in copies
// that struct (its size from hfx_emu_arg_size, which knows the digestion
// kernels) into a larger zeroed buffer and passes the buffer by value: both
// are aggregates passed in memory, so the kernel reads its own struct.
struct hfx_emu_argbuf {
  alignas(16) char b[512];
};
size_t hfx_emu_arg_size(const void* fn);
inline hipError_t hipLaunchKernel(const void* fn, dim3 g, dim3 b, void** args, size_t shm, hipStream_t) {
  if (shm > HFX_EMU_SMEM_BYTES) return 3;
  hfx_emu_argbuf buf;
  std::memset(buf.b, 0, sizeof(buf.b));
  std::memcpy(buf.b, args[0], hfx_emu_arg_size(fn));
  typedef void (*K)(hfx_emu_argbuf);
  hfx_emu_launch(g.y, [fn, buf]() { ((K)fn)(buf); }, b.x);
  return hipSuccess;
}
