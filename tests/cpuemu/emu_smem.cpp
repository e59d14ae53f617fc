// TEST INFRASTRUCTURE: backing store of the emulated LDS, the emulated block
// index, the argument size of type-erased launches and, in lanes mode, the
// lane scheduler (see hip/hip_runtime.h).
#include "hip/hip_runtime.h"

#include <unordered_set>

#include "eri_dispatch.h"
#include "hfx_kernels.h"

thread_local hfx_emu_idx blockIdx = {0, 0, 0};
thread_local hfx_emu_idx gridDim = {1, 1, 1};
#ifdef HFX_EMU_LANES
#include <execinfo.h>
#include <signal.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <cstdio>

thread_local hfx_emu_idx threadIdx = {0, 0, 0};
thread_local hfx_emu_idx blockDim = {1, 1, 1};

// Switches fibers: pushes the callee-saved registers and the SSE/x87 control
// words on the current stack, stores the stack pointer in *from, loads `to`
// (a pointer stored the same way) and pops its registers.  glibc's
// swapcontext would do the same plus a signal-mask system call per switch.
extern "C" void hfx_fiber_switch(void** from, void* to);
asm(R"(
  .text
  .globl hfx_fiber_switch
  .hidden hfx_fiber_switch
  .type hfx_fiber_switch, @function
hfx_fiber_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  subq $8, %rsp
  stmxcsr (%rsp)
  fnstcw 4(%rsp)
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  ldmxcsr (%rsp)
  fldcw 4(%rsp)
  addq $8, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
  .size hfx_fiber_switch, .-hfx_fiber_switch
)");

namespace {
constexpr size_t kFiberStack = (size_t)4 << 20;
enum LaneState { kRun, kAtWave, kAtBlock, kDone };
struct Lane {
  void* sp;                        // saved stack pointer while switched out
  LaneState state;
  int kind, site, src;
  unsigned long long val, res;
};
// one workgroup in flight on this host thread
struct Block {
  void* main = nullptr;             // the scheduler's saved stack pointer
  std::vector<Lane> lanes;
  std::vector<char*> stacks;       // kept for the thread's later launches
  const std::function<void()>* f = nullptr;
  unsigned cur = 0;
  bool active = false;
};
thread_local Block t_block;

[[noreturn]] void lane_fail(const char* what, unsigned lane, int site) {
  std::fprintf(stderr, "lanes emulation: %s (lane %u, source line %d, blockIdx.y %u)\n", what, lane, site,
               blockIdx.y);
  std::fflush(stderr);
  std::abort();
}

// a fault inside a lane: name the lane and the call stack, then die
void lane_fault(int sig, siginfo_t* si, void* uc) {
  const Block& b = t_block;
  void* pc = (void*)static_cast<ucontext_t*>(uc)->uc_mcontext.gregs[REG_RIP];
  char msg[200];
  const int n = std::snprintf(msg, sizeof msg,
                              "lanes emulation: signal %d in lane %u of blockIdx.y %u (active %d), address %p, pc:\n",
                              sig, b.cur, blockIdx.y, (int)b.active, si->si_addr);
  if (n > 0) (void)!write(2, msg, (size_t)n);
  backtrace_symbols_fd(&pc, 1, 2);
  void* ret = *reinterpret_cast<void**>(static_cast<ucontext_t*>(uc)->uc_mcontext.gregs[REG_RSP]);
  backtrace_symbols_fd(&ret, 1, 2);
  void* frames[48];
  backtrace_symbols_fd(frames, backtrace(frames, 48), 2);
  signal(sig, SIG_DFL);
  raise(sig);
}

void install_fault_handler() {
  static thread_local bool done = false;
  if (done) return;
  done = true;
  static thread_local char alt[1 << 16];
  stack_t ss{};
  ss.ss_sp = alt;
  ss.ss_size = sizeof alt;
  sigaltstack(&ss, nullptr);
  struct sigaction sa {};
  sa.sa_sigaction = lane_fault;
  sa.sa_flags = SA_ONSTACK | SA_SIGINFO;
  sigaction(SIGSEGV, &sa, nullptr);
  sigaction(SIGBUS, &sa, nullptr);
}

[[noreturn]] void lane_entry() {
  Block& b = t_block;
  (*b.f)();
  b.lanes[b.cur].state = kDone;
  void* dead;
  hfx_fiber_switch(&dead, b.main);   // never resumed
  __builtin_unreachable();
}

// a fresh lane's stack: the frame hfx_fiber_switch pops, returning into lane_entry
void* lane_stack(char* base, size_t size) {
  char* top = base + size;
  top -= (uintptr_t)top % 16;
  void** sp = reinterpret_cast<void**>(top - 72);   // lane_entry starts with rsp = 8 (mod 16)
  const uint32_t mxcsr = 0x1F80;
  const uint16_t fpcw = 0x037F;
  std::memcpy(sp, &mxcsr, 4);
  std::memcpy(reinterpret_cast<char*>(sp) + 4, &fpcw, 2);
  for (int k = 1; k <= 6; ++k) sp[k] = nullptr;     // r15 r14 r13 r12 rbx rbp
  sp[7] = reinterpret_cast<void*>(&lane_entry);
  return sp;
}

// completes the collective at which the lanes [w0, w0 + 64) of one wave wait;
// false when the wave has no lane waiting at a wave collective
bool resolve_wave(Block& b, unsigned w0) {
  int lead = -1;
  for (unsigned i = w0; i < w0 + 64; ++i)
    if (b.lanes[i].state == kAtWave) {
      lead = (int)i;
      break;
    }
  if (lead < 0) return false;
  const Lane& L = b.lanes[lead];
  unsigned long long mask = 0;
  for (unsigned i = w0; i < w0 + 64; ++i) {
    const Lane& x = b.lanes[i];
    if (x.state == kDone) continue;
    if (x.state != kAtWave || x.kind != L.kind || x.site != L.site)
      lane_fail("lanes of one wave wait at different collectives", i, x.site);
    if (x.val) mask |= 1ull << (i - w0);
  }
  for (unsigned i = w0; i < w0 + 64; ++i) {
    Lane& x = b.lanes[i];
    if (x.state != kAtWave) continue;
    if (L.kind == HFX_LC_BALLOT) {
      x.res = mask;
    } else if (L.kind == HFX_LC_SHFL) {
      const Lane& from = b.lanes[w0 + (unsigned)x.src];
      if (from.state != kAtWave) lane_fail("shuffle reads a lane that has exited", i, x.site);
      x.res = from.val;
    } else {
      x.res = 0;
    }
  }
  for (unsigned i = w0; i < w0 + 64; ++i)
    if (b.lanes[i].state == kAtWave) b.lanes[i].state = kRun;
  return true;
}
}  // namespace

unsigned long long hfx_lane_collective(int kind, int site, unsigned long long v, int src) {
  Block& b = t_block;
  if (!b.active) lane_fail("collective outside a kernel", 0, site);
  const unsigned me = b.cur;
  Lane& L = b.lanes[me];
  L.kind = kind;
  L.site = site;
  L.val = v;
  L.src = src;
  L.state = (kind == HFX_LC_SYNC || kind == HFX_LC_SYNC_COUNT) ? kAtBlock : kAtWave;
  hfx_fiber_switch(&L.sp, b.main);
  return b.lanes[me].res;
}

void hfx_lane_run_block(unsigned nt, unsigned y, unsigned gy, const std::function<void()>& f) {
  Block& b = t_block;
  if (b.active) lane_fail("launch from inside a kernel", b.cur, 0);
  install_fault_handler();
  if (nt == 0 || nt % 64 != 0 || nt > 1024) lane_fail("block size is not a multiple of 64 up to 1024", nt, 0);
  while (b.stacks.size() < nt) {
    void* p = mmap(nullptr, kFiberStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (p == MAP_FAILED) lane_fail("no memory for lane stacks", (unsigned)b.stacks.size(), 0);
    mprotect(p, 4096, PROT_NONE);       // guard page: a stack overflow faults
    b.stacks.push_back(static_cast<char*>(p));
  }
  b.lanes.assign(nt, Lane{});
  b.f = &f;
  blockIdx = hfx_emu_idx{0, y, 0};
  gridDim = hfx_emu_idx{1, gy, 1};
  blockDim = hfx_emu_idx{nt, 1, 1};
  for (unsigned i = 0; i < nt; ++i) {
    Lane& L = b.lanes[i];
    L.sp = lane_stack(b.stacks[i] + 4096, kFiberStack - 4096);
    L.state = kRun;
  }
  b.active = true;
  for (;;) {
    for (unsigned i = 0; i < nt; ++i)
      if (b.lanes[i].state == kRun) {
        b.cur = i;
        threadIdx = hfx_emu_idx{i, 0, 0};
        hfx_fiber_switch(&b.main, b.lanes[i].sp);
      }
    bool progress = false;
    for (unsigned w0 = 0; w0 < nt; w0 += 64) progress |= resolve_wave(b, w0);
    if (progress) continue;
    // every lane has exited or waits at a block barrier
    int lead = -1;
    unsigned long long count = 0;
    for (unsigned i = 0; i < nt; ++i) {
      const Lane& x = b.lanes[i];
      if (x.state == kDone) continue;
      if (lead < 0) lead = (int)i;
      const Lane& L = b.lanes[lead];
      if (x.kind != L.kind || x.site != L.site) lane_fail("lanes of one block wait at different barriers", i, x.site);
      count += x.val;
    }
    if (lead < 0) break;
    for (unsigned i = 0; i < nt; ++i)
      if (b.lanes[i].state == kAtBlock) {
        b.lanes[i].res = count;
        b.lanes[i].state = kRun;
      }
  }
  b.active = false;
  threadIdx = hfx_emu_idx{0, 0, 0};
}
#endif

namespace hfx {
__thread char hfx_smem[HFX_EMU_SMEM_BYTES] __attribute__((aligned(64)));
}

size_t hfx_emu_arg_size(const void* fn) {
  static const std::unordered_set<const void*>* dig = [] {
    auto* s = new std::unordered_set<const void*>;
    for (const auto& k : hfx::kClassKernels)
      for (const void* f : {k.dig, k.digj, k.digf, k.digm, k.digjm, k.digfm, k.dig1, k.dig1f})
        if (f) s->insert(f);
    return s;
  }();
  static_assert(sizeof(hfx::DigestArgs) <= sizeof(hfx_emu_argbuf) && sizeof(hfx::ClassArgs) <= sizeof(hfx_emu_argbuf),
                "argument buffer too small");
  return dig->count(fn) ? sizeof(hfx::DigestArgs) : sizeof(hfx::ClassArgs);
}
