// Host side of the hfx C ABI (include/hfx.h): context set-up, primitive-pair
// data, Schwarz lists, and the launch sequence of one Fock build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hfx.h"
#include "eri_dispatch.h"   // from the generated-kernel directory (-I)
#include "hfx_internal.h"
#include "hfx_kernels.h"

using namespace hfx;

namespace {

const char* kPairNames[HFX_NPAIRTYPES] = {"ss", "sp", "sd", "sl", "pp", "pd", "pl", "dd", "dl", "ll"};
const char* kClassNames[HFX_NCLASSES] = {
    "ssss", "sssp", "sssd", "sssl", "sspp", "sspd", "sspl", "ssdd", "ssdl", "ssll", "spsp", "spsd",
    "spsl", "sppp", "sppd", "sppl", "spdd", "spll", "sdsd", "sdsl", "sdpp", "sdpd", "sddd", "sddl",
    "sdll", "slsl", "slpp", "slpl", "sldd", "sldl", "slll", "pppp", "pppd", "pppl", "ppdd", "ppll",
    "pdpd", "pddd", "plpl", "plll", "dddd", "dddl", "ddll", "dldl", "dlll", "llll"};
// same-type class per pair type, used by the diagonal builds
const int kDiagClass[HFX_NPAIRTYPES] = {0, 10, 18, 25, 31, 36, 38, 40, 43, 45};

const double kShthresh = 1.0e-11;  // shthresh = intthresh/10 (molprops.f90:163)
const double kNegab = 1.0e-12;     // extrathresh = intthresh*0.01 (precalc_kabs.f90:13)
// Primitives per shell, at most: the reference's compile-time maxprim
// (molprops.f90:9-18).  k_screen keeps the set of its survivors' ket
// primitive-pair counts + 1 in two 64-bit words and counts them in
// 128-entry LDS tables, which needs kMaxShellPrims^2 + 1 < 128.
const int kMaxShellPrims = 10;
static_assert(kMaxShellPrims * kMaxShellPrims + 1 < 128, "k_screen's primitive-count set holds keys below 128");
// Screening statistics per class and list: surviving quartets, primitive
// quartets, bra primitive iterations, bra pairs with a survivor (runs).
const int kStatN = 4;

int letter(int shtyp) {  // canonical pair-name letter order s < p < d < L
  switch (shtyp) {
    case 1: return 0;
    case 2: return 1;
    case 3: return 2;
    case 12: return 3;
  }
  return -1;
}
const char kLetters[4] = {'s', 'p', 'd', 'l'};

int pair_type_index(char a, char b) {
  for (int t = 0; t < HFX_NPAIRTYPES; ++t)
    if (kPairNames[t][0] == a && kPairNames[t][1] == b) return t;
  return -1;
}

int ncomp_of(int shtyp) { return shtyp == 1 ? 1 : shtyp == 2 ? 3 : shtyp == 3 ? 5 : 4; }

struct Shell {
  int type, nprim, ao;
  double x[3];
  double zeta[16], djk[16], djkl[16];
};

// Boys F_j(x) by the series F_j(x) = e^-x sum_k (2x)^k / ((2j+1)(2j+3)...(2j+2k+1)),
// in long double; converges for all x on the grid (x <= 28).
void boys_table(std::vector<double>& t64, std::vector<float>& t32) {
  t64.assign(HFX_BOYS_ROWS * HFX_BOYS_COLS, 0.0);
  t32.assign(HFX_BOYS_ROWS * HFX_BOYS_COLS, 0.0f);
  for (int i = 0; i < HFX_BOYS_ROWS; ++i) {
    const long double x = (long double)i / (long double)HFX_BOYS_STEP;
    for (int j = 0; j < HFX_BOYS_COLS; ++j) {
      long double term = 1.0L / (2 * j + 1), sum = term;
      for (int k = 1; k < 400; ++k) {
        term *= 2.0L * x / (long double)(2 * j + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
      }
      const long double v = expl(-x) * sum;
      t64[i * HFX_BOYS_COLS + j] = (double)v;
      t32[i * HFX_BOYS_COLS + j] = (float)v;
    }
  }
}

// Device memory of the contexts comes from a per-process cache of freed
// blocks (size classes, four per power of two, per device): hipFree waits
// for the whole device, so a job that ends while others are running would
// stall its thread until all their queued work had drained.  A block is
// returned to the cache only when its context's stream no longer uses it;
// hfx_release_cached_memory hands the cached blocks back to HIP.
struct BlockCache {
  std::mutex mu;
  std::multimap<std::pair<int, size_t>, void*> free_blocks;   // (device, bytes) -> block
  std::unordered_map<void*, std::pair<int, size_t>> live;
};
BlockCache& block_cache() {
  static BlockCache* bc = new BlockCache;   // lives until the process ends
  return *bc;
}

size_t size_class(size_t n) {
  if (n <= 4096) return 4096;
  const int lg = 63 - __builtin_clzll((unsigned long long)(n - 1));
  const size_t step = (size_t)1 << (lg - 2);
  return (n + step - 1) / step * step;
}

void release_cached(int device) {   // device < 0: every device
  BlockCache& bc = block_cache();
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> lk(bc.mu);
    for (auto it = bc.free_blocks.begin(); it != bc.free_blocks.end();) {
      if (device < 0 || it->first.first == device) {
        drop.push_back(it->second);
        it = bc.free_blocks.erase(it);
      } else {
        ++it;
      }
    }
  }
  hfx::DeviceSection ds;
  for (void* p : drop) hipFree(p);
}

void* dev_alloc(size_t bytes) {
  if (hfx::t_capturing) return nullptr;   // never inside a capture section
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const std::pair<int, size_t> key(dev, size_class(bytes));
  BlockCache& bc = block_cache();
  {
    std::lock_guard<std::mutex> lk(bc.mu);
    auto it = bc.free_blocks.find(key);
    if (it != bc.free_blocks.end()) {
      void* p = it->second;
      bc.free_blocks.erase(it);
      bc.live[p] = key;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e;
  {
    hfx::DeviceSection ds;
    e = hipMalloc(&p, key.second);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release_cached(dev);            // memory held by the cache, then try once more
    hfx::DeviceSection ds;
    if (hipMalloc(&p, key.second) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  std::lock_guard<std::mutex> lk(bc.mu);
  bc.live[p] = key;
  return p;
}

// The caller guarantees that no queued work still uses p.
void dev_free(void* p) {
  if (!p) return;
  BlockCache& bc = block_cache();
  {
    std::lock_guard<std::mutex> lk(bc.mu);
    auto it = bc.live.find(p);
    if (it != bc.live.end()) {
      bc.free_blocks.emplace(it->second, p);
      bc.live.erase(it);
      return;
    }
  }
  hfx::DeviceSection ds;
  hipFree(p);
}

}  // namespace

// Device buffers of hfx_ctx_onee_noalloc: a fixed array in every GPU's copy
// of the code object, held until the process ends.  Its host-side name is
// not a device address; each device's address is taken once
// (scratch_address) and used everywhere host code passes the array on.
constexpr size_t kScratchBytes = (size_t)48 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;   // per device, null until taken
std::vector<char> g_scratch_done;
}  // namespace

// The scratch array's address on the current device (taken on the first
// call per device, with that device current); null if HIP refuses it.
static void* scratch_address(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0) return nullptr;
  if ((size_t)dev >= g_scratch_ptr.size()) {
    g_scratch_ptr.resize(dev + 1, nullptr);
    g_scratch_done.resize(dev + 1, 0);
  }
  if (!g_scratch_done[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fprintf(stderr, "hfx: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      p = nullptr;
    }
    g_scratch_ptr[dev] = p;
    g_scratch_done[dev] = 1;
  }
  return g_scratch_ptr[dev];
}

// Dispatches that need more scratch memory (spilled registers) than the
// device's scratch-limit threshold get a use-once scratch allocation per
// dispatch; below it the runtime keeps the allocation on the queue.  The
// largest combination kernels spill kilobytes per lane, so the threshold is
// raised, once per device, to what the library's kernels need at full
// occupancy (at most the device's maximum).
static void raise_scratch_limit(int dev, const hipDeviceProp_t& prop) {
  static std::mutex mu;
  static std::vector<char> done;
  std::lock_guard<std::mutex> lk(mu);
  if (dev < 0) return;
  if ((size_t)dev >= done.size()) done.resize(dev + 1, 0);
  if (done[dev]) return;
  done[dev] = 1;
  size_t lane = 0;
  auto take = [&](const void* fn) {
    hipFuncAttributes at;
    if (fn && hipFuncGetAttributes(&at, fn) == hipSuccess) lane = std::max(lane, (size_t)at.localSizeBytes);
    else (void)hipGetLastError();
  };
  for (const ClassKernels& k : kClassKernels) {
    for (const void* fn : {k.f64, k.f64j, k.f32, k.dig, k.digj, k.digf, k.digm, k.digjm, k.digfm, k.dig1, k.dig1f})
      take(fn);
    for (int i = 0; i < k.ncombo; ++i) {
      take(k.c64[i]);
      for (int j = 0; j < k.npre[i]; ++j) take(k.pre64[i][j]);
    }
    for (int i = 0; i < k.ncombo32; ++i) {
      take(k.c32[i]);
      for (int j = 0; j < k.npre32[i]; ++j) take(k.pre32[i][j]);
    }
  }
  size_t mx = 0, cur = 0;
  if (lane == 0 || hipDeviceGetLimit(&mx, hipExtLimitScratchMax) != hipSuccess ||
      hipDeviceGetLimit(&cur, hipExtLimitScratchCurrent) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  const size_t want = std::min(mx, lane * HFX_WAVE * (size_t)prop.multiProcessorCount *
                                       (size_t)std::max(1, prop.maxThreadsPerMultiProcessor / HFX_WAVE));
  if (want > cur && hipDeviceSetLimit(hipExtLimitScratchCurrent, want) != hipSuccess) (void)hipGetLastError();
}

std::shared_mutex& hfx::capture_mutex() {
  static std::shared_mutex* m = new std::shared_mutex;   // lives until the process ends
  return *m;
}
thread_local bool hfx::t_capturing = false;

void* hfx::cached_alloc(size_t bytes) { return dev_alloc(bytes); }
void hfx::cached_free(void* p) { dev_free(p); }

namespace {

// A copy in order on `st` (never the null stream: see hfx::capture_mutex),
// waited for.
hipError_t copy_sync(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
  return e != hipSuccess ? e : hipStreamSynchronize(st);
}

template <typename T>
int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  *p = static_cast<T*>(dev_alloc(n * sizeof(T)));
  return *p ? HFX_OK : HFX_E_NOMEM;
}

template <typename T>
void dfree(T*& p) {
  dev_free((void*)p);
  p = nullptr;
}

}  // namespace

struct ClassPlan {
  int cls = 0, t1 = 0, t2 = 0, same = 0;
  long long cap = 0;           // maximum number of quartets
  int2* list64 = nullptr;
  int2* list32 = nullptr;
  double* blocks64 = nullptr;  // debug
  double* blocks32 = nullptr;
};

// One molecule of a context.  A context holds one molecule (hfx_ctx_create)
// or several whose builds share every launch (hfx_batch_create): shells are
// numbered across the molecules, pairs of a type are grouped by molecule,
// and molecule m's nb x nb matrices start at m * nb * nb (nb = the context's
// stride, the largest molecule's basis size).
struct MolPart {
  int nb = 0, nsh = 0, sh0 = 0;                 // basis functions, shells, first shell
  int npair[HFX_NPAIRTYPES] = {0};
  int pbeg[HFX_NPAIRTYPES] = {0};               // first global pair of the molecule's type-t list
  std::vector<double> schw[HFX_NPAIRTYPES];     // per pair (the molecule's list order)
  std::vector<int> order[HFX_NPAIRTYPES];       // 0-based positions in that list
  int last[HFX_NPAIRTYPES] = {0};
};

// A build captured in a HIP graph, replayed by later calls with the same arguments.
struct GraphKey {
  const double* P;
  double* F;
  const double* thr_diag;
  const double* thr_nd;
  const double* split_mol;
  int mode, zero_f;
  double split;
  bool operator==(const GraphKey& o) const {
    return P == o.P && F == o.F && thr_diag == o.thr_diag && thr_nd == o.thr_nd && split_mol == o.split_mol &&
           mode == o.mode && zero_f == o.zero_f && split == o.split;
  }
};
struct CachedGraph {
  GraphKey key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct hfx_ctx {
  int nb = 0, nsh = 0, dev = 0, ncu = 0;        // nb: matrix stride; nsh: shells of all molecules
  int nmol = 1, nsp = 0;                        // molecules; nsp: largest shell count (maxdm block)
  std::vector<MolPart> mols;
  std::vector<Shell> shells;
  int npair[HFX_NPAIRTYPES] = {0};              // pairs of the type over all molecules
  int offset[HFX_NPAIRTYPES] = {0};             // first global pair of the type
  int act_len[HFX_NPAIRTYPES] = {0};            // active (truncated) list length over all molecules
  int nchunk_max[HFX_NPAIRTYPES] = {0};         // largest per-molecule active length, in waves
  bool have_schwarz = false;
  bool device_chunks = false;     // block classes chunked by capacity: no count read-back (graph capture)
  bool capturing = false;         // a HIP graph is being captured from the context's stream
  std::vector<CachedGraph> graphs;  // captured builds; dropped whenever a buffer they use is replaced
  bool graphs_off = false;          // a capture failed: later builds run uncaptured
  void* scratch = nullptr;          // address of the scratch array on the context's device
  int cls_per_cu[HFX_NCLASSES] = {};          // launch width per class (workgroups per CU; 0: default)
  long long cls_chunk[HFX_NCLASSES] = {};     // quartets per chunk per block class (0: default)
  std::vector<PairRec> pairs;
  std::vector<PrimRec> prims;
  std::vector<PrimRec32> prims32;   // prims in FP32, P relative to the pair's centres
  std::vector<ClassPlan> classes;
  int totpairs = 0;
  double split = 0.0;
  int debug_blocks = 0;
  int profile = 0;
  double* d_eblk = nullptr;     // integral blocks of the block class kernels
  double* d_vmax = nullptr;     // their partial max|e|, kMaxCombo rows of vmax_rows
  double* d_cbuf = nullptr;     // contracted values of split combinations (pre-kernels to the last kernel)
  long long cbuf_doubles = 0;   // capacity of d_cbuf
  long long eblk_doubles = 0;   // capacity of d_eblk
  long long vmax_rows = 0;      // quartets per row of d_vmax
  size_t max_shm = 0;           // LDS bytes a workgroup may use
  long long chunk_opt = 0;      // HFX_OPT_CHUNK (0: as large as the buffer allows)
  int class_ran[2][HFX_NCLASSES] = {};
  int dig_occ[3][HFX_NCLASSES] = {};   // resident digestion workgroups per CU [fp64, fp64 J-only, fp32]
  int digest1 = 1;                     // HFX_OPT_DIGEST: one-pass digestion where its LDS rows fit
  int dig1_occ[2][HFX_NCLASSES] = {};  // resident one-pass digestion workgroups per CU [fp64, fp32]
  std::vector<double> diag_cost;   // diagonal builds: estimated work per pair type (stream balancing)
  // class start, end, and a second start (the other side stream when a
  // class's first kernels start on both): HFX_OPT_PROFILE spans
  hipEvent_t cev[2][HFX_NCLASSES][3] = {};
  int last_kind = 0;  // 0 none, 1 non-diag, 2 diag
  // device
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // side streams for the combination kernels of a large chunk (independent
  // writers of disjoint block elements), created on first use; fork/join events.
  // Used only while this is the process's one live context: with several,
  // their own streams already keep the GPU busy.
  static constexpr int kSide = 2;   // with the context's stream and the null stream: four hardware queues
  bool counted = false;           // holds one count of live_contexts
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr, join[kSide] = {};
  hipEvent_t done[2] = {};        // digestion of the last chunk in each half of the integral buffer
  // a build's block-class lists pipelined across lists (launch_split): on,
  // chunks so far, halves used, and the point after which the first chunks'
  // combination kernels may start
  bool overlap = true;   // HFX_OPT_OVERLAP
  bool xp = false;
  long long xp_j = 0;
  bool xp_used[2] = {false, false};
  hipEvent_t xp_start = nullptr;
  PairRec* d_pairs = nullptr;
  PrimRec* d_prims = nullptr;
  PrimRec32* d_prims32 = nullptr;
  ShellRec* d_shells = nullptr;
  double* d_pairbuf = nullptr;    // X_cd of the digestion per ket pair (HFX_PAIRBUF_STRIDE per pair)
  double* d_boys64 = nullptr;
  float* d_boys32 = nullptr;
  double* d_P = nullptr;
  float* d_P32 = nullptr;         // FP32 copy of the density of a mixed-precision build
  double* d_F = nullptr;
  double* d_X = nullptr;
  double* d_maxdm = nullptr;
  unsigned long long* d_pmax = nullptr;        // per molecule: max over maxdm (bits of the double)
  int2* d_molsh = nullptr;                     // per molecule: first shell, shell count
  int* d_bmol[HFX_NPAIRTYPES] = {nullptr};       // molecule per active position
  int* d_aseg[HFX_NPAIRTYPES] = {nullptr};       // first active position per molecule (nmol + 1)
  unsigned long long* d_masks = nullptr;       // screening survivor bits of one class
  size_t masks_words = 0;
  unsigned long long* d_masks_side = nullptr;  // the same for the screens on the side streams (kSide x masks_words)
  size_t masks_side_words = 0;
  int act_sorted[HFX_NPAIRTYPES] = {0};         // active Schwarz values non-increasing
  double* d_schw_out = nullptr;   // per global pair
  int* d_act[HFX_NPAIRTYPES] = {nullptr};        // active global pair indices (rank order)
  double* d_act_schw[HFX_NPAIRTYPES] = {nullptr};
  int4* d_act_info[HFX_NPAIRTYPES] = {nullptr};   // shells and primitive counts per active position
  double* d_pair_schw[HFX_NPAIRTYPES] = {nullptr};  // Schwarz per pair (list order)
  int2* d_diag_list[HFX_NPAIRTYPES] = {nullptr};
  // one-electron set-up on the device (allocated on first use)
  int4* d_onee_items = nullptr;
  std::vector<int> onee_count;                 // work items per (molecule, la, lb), consecutive in d_onee_items
  int *d_sh_type = nullptr, *d_sh_np = nullptr, *d_sh_ao = nullptr;
  double *d_sh_xyz = nullptr, *d_sh_z = nullptr, *d_sh_c = nullptr, *d_sh_cl = nullptr;
  int* d_counts = nullptr;                     // [2][HFX_NCLASSES]
  unsigned long long* d_stats = nullptr;       // [2][HFX_NCLASSES][kStatN]
};

#define HIPOK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "hfx: %s failed: %s\n", #x, hipGetErrorString(e_));    \
      return HFX_E_HIP;                                                       \
    }                                                                         \
  } while (0)

namespace {

int grid_for(long long n, int nt, int maxblocks) {
  long long g = (n + nt - 1) / nt;
  if (g < 1) g = 1;
  if (g > maxblocks) g = maxblocks;
  return (int)g;
}

// Waits for the context's stream and its side streams.  A build joins its
// side streams into the context's stream, but one that failed between
// launching a chunk's kernels and joining them leaves work only there.
hipError_t sync_all_streams(hfx_ctx* c) {
  hipError_t e = c->stream ? hipStreamSynchronize(c->stream) : hipSuccess;
  for (int s = 0; s < hfx_ctx::kSide; ++s)
    if (c->side[s]) {
      const hipError_t f = hipStreamSynchronize(c->side[s]);
      if (e == hipSuccess) e = f;
    }
  return e;
}

// Captured builds hold the addresses of the context's buffers: they are
// dropped (and captured again on their next call) whenever one is replaced.
void drop_graphs(hfx_ctx* c) {
  if (c->graphs.empty()) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  hfx::DeviceSection ds;
  for (CachedGraph& g : c->graphs) {
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.graph) hipGraphDestroy(g.graph);
  }
  c->graphs.clear();
}

// Active (Schwarz-truncated) lists on the device, from each molecule's
// schw/order/last: per pair type, the molecules' truncated lists one after
// the other (each in its own Schwarz order), with the molecule of every
// position and where each molecule's segment starts.
int upload_active(hfx_ctx* c) {
  HIPOK(sync_all_streams(c));   // the lists below are replaced; no build may still read them
  drop_graphs(c);
  for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
    if (c->npair[t] == 0) continue;
    std::vector<int> act, bmol, aseg(c->nmol + 1, 0);
    std::vector<double> asc;
    std::vector<int4> info;
    int sorted = 1, maxlen = 0;
    for (int m = 0; m < c->nmol; ++m) {
      const MolPart& mp = c->mols[m];
      aseg[m] = (int)act.size();
      maxlen = std::max(maxlen, mp.last[t]);
      for (int k = 0; k < mp.last[t]; ++k) {
        const int pi = mp.order[t][k];
        const int g = mp.pbeg[t] + pi;
        act.push_back(g);
        asc.push_back(mp.schw[t][pi]);
        if (k > 0 && asc.back() > asc[asc.size() - 2]) sorted = 0;
        const PairRec& pr = c->pairs[g];
        info.push_back(make_int4(pr.lsa, pr.lsb, pr.np, pr.npall));
        bmol.push_back(m);
      }
    }
    aseg[c->nmol] = (int)act.size();
    c->act_len[t] = (int)act.size();
    c->nchunk_max[t] = (maxlen + HFX_WAVE - 1) / HFX_WAVE;
    c->act_sorted[t] = sorted;
    dfree(c->d_act[t]);
    dfree(c->d_act_schw[t]);
    dfree(c->d_act_info[t]);
    dfree(c->d_bmol[t]);
    dfree(c->d_aseg[t]);
    if (dalloc(&c->d_act[t], act.size()) || dalloc(&c->d_act_schw[t], asc.size()) ||
        dalloc(&c->d_act_info[t], info.size()) || dalloc(&c->d_bmol[t], bmol.size()) ||
        dalloc(&c->d_aseg[t], aseg.size()))
      return HFX_E_NOMEM;
    if (!act.empty()) {
      HIPOK(copy_sync(c->stream, c->d_act_info[t], info.data(), info.size() * sizeof(int4), hipMemcpyHostToDevice));
      HIPOK(copy_sync(c->stream, c->d_act[t], act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice));
      HIPOK(copy_sync(c->stream, c->d_act_schw[t], asc.data(), asc.size() * sizeof(double), hipMemcpyHostToDevice));
      HIPOK(copy_sync(c->stream, c->d_bmol[t], bmol.data(), bmol.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    HIPOK(copy_sync(c->stream, c->d_aseg[t], aseg.data(), aseg.size() * sizeof(int), hipMemcpyHostToDevice));
    std::vector<double> ps(c->npair[t]);
    for (int m = 0; m < c->nmol; ++m) {
      const MolPart& mp = c->mols[m];
      for (int k = 0; k < mp.npair[t]; ++k) ps[mp.pbeg[t] - c->offset[t] + k] = mp.schw[t][k];
    }
    HIPOK(copy_sync(c->stream, c->d_pair_schw[t], ps.data(), ps.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  // class plans and their work lists: capacity = candidate quartets of every molecule
  size_t need_masks = 0;
  for (auto& cp : c->classes) {
    dfree(cp.list64);
    dfree(cp.list32);
    long long cap = 0;
    for (int m = 0; m < c->nmol; ++m) {
      const long long nb = c->mols[m].last[cp.t1], nk = c->mols[m].last[cp.t2];
      cap += cp.same ? nb * (nb - 1) / 2 : nb * nk;
    }
    cp.cap = cap;
    // quartet positions and counts are 32-bit on the device
    if (cp.cap > (long long)INT32_MAX - 4 * HFX_NT) return HFX_E_UNSUPPORTED;
    if (cp.cap > 0 && dalloc(&cp.list64, (size_t)cp.cap)) return HFX_E_NOMEM;
    need_masks = std::max(need_masks, (size_t)c->act_len[cp.t1] * (size_t)c->nchunk_max[cp.t2] * 2);
  }
  if (need_masks > c->masks_words) {
    dfree(c->d_masks);
    if (dalloc(&c->d_masks, need_masks)) return HFX_E_NOMEM;
    c->masks_words = need_masks;
  }
  c->have_schwarz = true;
  return HFX_OK;
}

int ensure_mixed_lists(hfx_ctx* c) {
  if (!c->d_P32 && dalloc(&c->d_P32, (size_t)c->nb * c->nb * c->nmol)) return HFX_E_NOMEM;
  for (auto& cp : c->classes)
    if (!cp.list32 && cp.cap > 0 && dalloc(&cp.list32, (size_t)cp.cap)) return HFX_E_NOMEM;
  return HFX_OK;
}

ClassArgs base_args(hfx_ctx* c, const double* dP, double thr) {
  ClassArgs a;
  memset(&a, 0, sizeof(a));
  a.pairs = c->d_pairs;
  a.prims = c->d_prims;
  a.prims32 = c->d_prims32;
  a.P = dP;
  a.X = c->d_X;
  a.boys64 = c->d_boys64;
  a.boys32 = c->d_boys32;
  a.pairbuf = c->d_pairbuf;
  a.thr = thr;
  a.nb = c->nb;
  return a;
}

int launch_class(hfx_ctx* c, int cls, int prec, bool jonly, ClassArgs a, long long cap, hipStream_t st = nullptr) {
  const ClassKernels& k = kClassKernels[cls];
  if (!st) st = c->stream;
  const bool prof = c->profile && !c->device_chunks;
  if (prof) {
    HIPOK(hipEventRecord(c->cev[prec][cls][0], st));
    HIPOK(hipEventRecord(c->cev[prec][cls][2], st));
  }
  const void* fn = prec == 1 ? k.f32 : (jonly ? k.f64j : k.f64);
  const int K = prec == 1 ? 5 : 8;
  const size_t boys_shm = (size_t)HFX_BOYS_ROWS * (k.M + K) * (prec == 1 ? sizeof(float) : sizeof(double));
  // rows of X of the resident bra pair (shells A and B) after the Boys table,
  // where at least two workgroups per CU still fit (off in the diagonal builds)
  const size_t rows_off = (boys_shm + 15) / 16 * 16;
  const size_t rows_shm = (size_t)(k.dims[0] + k.dims[1]) * c->nb * sizeof(double);
  const bool rows = !a.diag && rows_off + rows_shm <= c->max_shm / 2;
  const size_t shm = rows ? rows_off + rows_shm : boys_shm;
  a.lds_rows = rows ? 1 : 0;
  a.rows_offset = (int)rows_off;
  int per_cu = rows ? (int)std::max<size_t>(2, std::min<size_t>(8, c->max_shm / shm)) : 8;
  if (c->cls_per_cu[cls] > 0) per_cu = c->cls_per_cu[cls];   // tuned launch width
  const int grid = grid_for(cap, HFX_NT, c->ncu * per_cu);
  void* args[] = {&a};
  HIPOK(hipLaunchKernel(fn, dim3(grid), dim3(HFX_NT), args, shm, st));
  if (prof) HIPOK(hipEventRecord(c->cev[prec][cls][1], st));
  c->class_ran[prec][cls] = 1;
  return HFX_OK;
}

const long long kEblkDoubles = (1ll << 20) * 625;       // integral buffer of the block classes, at most (5.2 GB) ...
const long long kEblkDoublesAlone = (1ll << 21) * 625;  // ... or this when it grows in the one live context (10.5 GB)
const long long kMaxChunk = 1 << 22;                      // quartets per chunk, at most
const int kMaxCombo = 16;                            // combination kernels per class, at most

const long long kConcurrentChunk = 1 << 17;   // quartets of a chunk from which its combinations run side by side
std::atomic<int> live_contexts{0};             // contexts created and not yet destroyed, process-wide

// Creates what is still missing (a failed call leaves the handles it made,
// which the next call keeps and hfx_ctx_destroy frees); the fork event last.
int ensure_side_streams(hfx_ctx* c) {
  if (c->fork) return HFX_OK;
  for (int s = 0; s < hfx_ctx::kSide; ++s) {
    if (!c->side[s]) HIPOK(hipStreamCreateWithFlags(&c->side[s], hipStreamNonBlocking));
    if (!c->join[s]) HIPOK(hipEventCreateWithFlags(&c->join[s], hipEventDisableTiming));
  }
  for (int h = 0; h < 2; ++h)
    if (!c->done[h]) HIPOK(hipEventCreateWithFlags(&c->done[h], hipEventDisableTiming));
  if (!c->xp_start) HIPOK(hipEventCreateWithFlags(&c->xp_start, hipEventDisableTiming));
  HIPOK(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  return HFX_OK;
}

// Heavy classes: combination kernels (integrals to HBM, tiles of 64
// quartets) + the class's digestion kernel, chunk by chunk over quartets
// [0, n).  n is the list's length read back from the device, or its
// capacity when the context sizes its launches on the device
// (device_chunks): chunks past the list's end then exit at once, and the
// whole build is one fixed launch sequence on the context's stream (no side
// streams, no buffer growth, no host synchronisation), which a HIP graph can
// capture.  The digestion walks contiguous ranges of the (bra-contiguous)
// list and, where the workgroup's LDS holds the rows of X that a bra pair's
// quartets update (rows of shells A and B: J_ab, K_ac, K_ad, K_bc, K_bd),
// accumulates them there and flushes them with row-coalesced atomics.
// The integral buffer and vmax grow to `need` doubles (capped by the
// context's limit; vmax to a sixteenth of that in rows).  Growing frees the
// old buffers (an implicit device synchronisation), so they grow at least
// twofold.
int grow_eblk(hfx_ctx* c, long long need) {
  const long long limit = live_contexts.load() == 1 ? kEblkDoublesAlone : kEblkDoubles;
  need = std::min(limit, need);
  if (need <= c->eblk_doubles) return HFX_OK;
  if (c->capturing) return HFX_E_STATE;   // buffers are sized by an uncaptured build first
  const long long cap = std::min(limit, std::max(need, 2 * c->eblk_doubles));
  const long long rows = std::min(kMaxChunk, (cap / 16 + HFX_NT - 1) / HFX_NT * HFX_NT);
  HIPOK(sync_all_streams(c));   // earlier chunks may still use the old buffers
  drop_graphs(c);
  dfree(c->d_eblk);
  dfree(c->d_vmax);
  c->eblk_doubles = c->vmax_rows = 0;
  if (dalloc(&c->d_eblk, (size_t)cap) || dalloc(&c->d_vmax, (size_t)rows * kMaxCombo)) return HFX_E_NOMEM;
  c->eblk_doubles = cap;
  c->vmax_rows = rows;
  return HFX_OK;
}

// After a chunk's digestion (on the context's stream): the event that frees
// its half of the integral buffer; across lists also the half's use and the
// chunk count, and a list that used the whole buffer frees both halves.
int after_digestion(hfx_ctx* c, bool pipelined, bool xp, int h) {
  if (pipelined) HIPOK(hipEventRecord(c->done[h], c->stream));
  if (xp) {
    c->xp_used[h] = true;
    ++c->xp_j;
  } else if (c->xp && c->done[0]) {
    for (int k = 0; k < 2; ++k) {
      HIPOK(hipEventRecord(c->done[k], c->stream));
      c->xp_used[k] = true;
    }
  }
  return HFX_OK;
}

// A block class's own stream and share of the integral buffer and vmax, for
// block classes that run side by side (the diagonal builds): the class runs
// in one chunk of `rows` quartets, all its kernels on `st`.
struct BlockSlot {
  hipStream_t st;
  long long eoff, voff, rows;   // doubles into d_eblk, doubles into d_vmax, quartets
};

int launch_split(hfx_ctx* c, int cls, int prec, bool jonly, ClassArgs a, long long n, double* dump,
                 long long runs = 0, const BlockSlot* slot = nullptr) {
  const ClassKernels& k = kClassKernels[cls];
  if (n <= 0) return HFX_OK;
  hipStream_t ms = slot ? slot->st : c->stream;   // the class's stream (digestion, profile events)
  if (prec == 1 && a.diag) {   // the diagonal builds' element cut (k_cut_small) is on FP64 blocks
    fprintf(stderr, "hfx: FP32 blocks in a diagonal build\n");
    return HFX_E_STATE;
  }
  const long long ne = (long long)k.dims[0] * k.dims[1] * k.dims[2] * k.dims[3];
  // The integral buffer grows to what this class needs for all its quartets
  // at once (small molecules take megabytes), up to kEblkDoubles, or
  // kEblkDoublesAlone for the process's one live context (its two halves
  // then hold chunks as large as one buffer would with several contexts);
  // larger lists are chunked.  Growing frees the old buffer (an implicit
  // device synchronisation), so it grows at least twofold.
  const long long limit = live_contexts.load() == 1 ? kEblkDoublesAlone : kEblkDoubles;
  if (slot) {   // the caller sized the buffers for every slot
    const int nc = prec == 1 ? k.ncombo32 : k.ncombo;
    if (slot->rows < n || slot->rows % HFX_NT || slot->eoff + slot->rows * ne > c->eblk_doubles ||
        slot->voff + nc * slot->rows > c->vmax_rows * kMaxCombo || (prec == 1 ? k.cbuf32 : k.cbuf64) > 0) {
      fprintf(stderr, "hfx: block slot does not fit the integral buffer\n");
      return HFX_E_STATE;
    }
  } else {
    int r = grow_eblk(c, ((n + HFX_NT - 1) / HFX_NT) * HFX_NT * ne);
    if (r) return r;
  }
  long long chunk = slot ? slot->rows : std::min(c->vmax_rows, (c->eblk_doubles / ne) / HFX_NT * HFX_NT);
  if (!slot && c->chunk_opt > 0) chunk = std::min(chunk, c->chunk_opt);
  if (!slot && c->cls_chunk[cls] > 0) chunk = std::min(chunk, c->cls_chunk[cls]);   // tuned quartets per chunk
  // Several chunks, one live context: the buffers split into two halves and
  // the chunks alternate between them; the combination kernels run on the
  // side streams, the digestion on the context's stream, so chunk j's
  // digestion overlaps chunk j+1's integrals.  Chunk j+2 reuses chunk j's
  // half once that digestion is done (event done[j % 2]).
  const long long half_e = c->eblk_doubles / 2 / 64 * 64;   // doubles per half, 512-byte aligned
  const long long half_v = c->vmax_rows / 2 * kMaxCombo;    // vmax doubles per half
  const long long half_chunk = std::min(c->vmax_rows / 2, half_e / ne) / HFX_NT * HFX_NT;
  // Across lists (c->xp, one build's block classes with one live context):
  // every list runs in half-buffer chunks, the chunk count continuing from
  // the previous list, so a list's first combination kernels wait only for
  // the digestion that last read their half, not for the previous list's
  // last digestion.
  const bool xp = c->xp && !slot && !a.diag && dump == nullptr && half_chunk >= HFX_NT;
  const bool pipelined = xp || (!slot && n > chunk && half_chunk >= HFX_NT && live_contexts.load() == 1 &&
                                !c->device_chunks);
  if (pipelined) {
    chunk = std::min(chunk, half_chunk);
    int r = ensure_side_streams(c);
    if (r) return r;
  }
  // Split combinations: their pre-kernels store cbn doubles per quartet of
  // the chunk (floats: half), per half of the pipeline; the buffer has the
  // integral buffer's budget, and chunks shrink to fit it.
  const long long cbn = prec == 1 ? (k.cbuf32 + 1) / 2 : k.cbuf64;
  if (cbn > 0) {
    const long long halves = pipelined ? 2 : 1;
    const long long cmax = limit / (halves * cbn) / HFX_NT * HFX_NT;
    if (cmax < HFX_NT) return HFX_E_NOMEM;
    chunk = std::min(chunk, cmax);
    const long long need = halves * cbn * chunk;
    if (need > c->cbuf_doubles) {
      if (c->capturing) return HFX_E_STATE;   // sized by an uncaptured build first
      const long long cap = std::min(limit, std::max(need, 2 * c->cbuf_doubles));
      HIPOK(sync_all_streams(c));
      drop_graphs(c);
      dfree(c->d_cbuf);
      c->cbuf_doubles = 0;
      if (dalloc(&c->d_cbuf, (size_t)cap)) return HFX_E_NOMEM;
      c->cbuf_doubles = cap;
    }
  }
  // every chunk's blocks and partial maxima stay inside its half (or the
  // whole buffer) of the integral buffer and of vmax
  const bool fits = slot ? chunk >= HFX_NT
                         : chunk >= HFX_NT && chunk * ne <= (pipelined ? half_e : c->eblk_doubles) &&
                               chunk <= (pipelined ? c->vmax_rows / 2 : c->vmax_rows);
  if (!fits || k.ncombo > kMaxCombo || k.ncombo32 > kMaxCombo) {
    fprintf(stderr, "hfx: chunk of %lld quartets does not fit the integral buffer\n", chunk);
    return HFX_E_STATE;
  }
  const int K = prec == 1 ? 5 : 8;
  // LDS rows of X for the digestion (off for the diagonal builds: every bra once)
  const size_t rows_shm = (size_t)(k.dims[0] + k.dims[1]) * c->nb * sizeof(double);
  // staging of X_cd: 16 quartets x nC nD doubles per wave
  const size_t tb_shm = (size_t)(HFX_NT / HFX_WAVE) * std::min(HFX_WAVE, 16) * k.dims[2] * k.dims[3] * sizeof(double);
  const bool rows = !a.diag && rows_shm + tb_shm <= c->max_shm;
  const size_t dig_shm = (rows ? rows_shm : 0) + tb_shm;
  // one-pass digestion (one molecule, J and K, off-diagonal): X rows, then the
  // same rows of P in the blocks' precision, the X_cd staging and one int per wave
  const size_t one_prow_off = rows_shm;
  const size_t one_tb_off = (one_prow_off + (size_t)(k.dims[0] + k.dims[1]) * c->nb *
                             (prec == 1 ? sizeof(float) : sizeof(double)) + 15) / 16 * 16;
  const int one_nt = prec == 1 ? k.dig1_nt32 : DIG1_NT;   // threads per one-pass workgroup
  const size_t one_run_off = one_tb_off + (size_t)(one_nt / HFX_WAVE) * std::min(HFX_WAVE, 16) * k.dims[2] *
                                              k.dims[3] * sizeof(double);
  const size_t one_shm = one_run_off + (size_t)(one_nt / HFX_WAVE) * sizeof(int);
  // where it measured faster (the class's table entry for the list's
  // precision) and the list's bra runs are long (a step never crosses a bra
  // pair, so short runs leave most of the workgroup idle); runs = 0: unknown
  // (launches sized on the device), the table alone decides; HFX_OPT_DIGEST
  // 2 takes it wherever it fits
  const bool one_ok = c->digest1 == 2 ||
                      (c->digest1 == 1 && (k.dig1_use >> (prec == 1 ? 1 : 0) & 1) &&
                       (runs <= 0 || n >= (long long)k.dig1_min_run * runs));
  const bool one_pass = one_ok && !a.diag && !jonly && c->nmol == 1 && dump == nullptr &&
                        a.schw_out == nullptr && k.dig1 != nullptr && one_shm <= c->max_shm;
  const int ncombo = prec == 1 ? k.ncombo32 : k.ncombo;
  const bool prof = c->profile && !c->device_chunks;
  if (prof && !xp) {
    HIPOK(hipEventRecord(c->cev[prec][cls][0], ms));
    HIPOK(hipEventRecord(c->cev[prec][cls][2], ms));
  }
  if (pipelined && !xp) {
    // the side streams start after everything already on the context's stream
    HIPOK(hipEventRecord(c->fork, c->stream));
    for (int s = 0; s < hfx_ctx::kSide; ++s) HIPOK(hipStreamWaitEvent(c->side[s], c->fork, 0));
  }
  for (long long q0 = 0, j = 0; q0 < n; q0 += chunk, ++j) {
    const long long q1 = std::min(n, q0 + chunk);
    const int h = xp ? (int)(c->xp_j & 1) : pipelined ? (int)(j & 1) : 0;
    double* eblk = slot ? c->d_eblk + slot->eoff : c->d_eblk + h * half_e;
    double* vmax = slot ? c->d_vmax + slot->voff : c->d_vmax + h * half_v;
    ClassArgs b = a;
    b.eblk = eblk;
    b.vmax = vmax;
    // (halves at a fixed offset: lists of other classes alternate halves too)
    b.cbuf = cbn > 0 ? c->d_cbuf + (size_t)h * (size_t)(c->cbuf_doubles / 2) : nullptr;
    b.vstride = (int)chunk;
    b.q0 = (int)q0;
    b.q1 = (int)q1;
    void* args[] = {&b};
    // A chunk's combination kernels write disjoint elements of the blocks and
    // their own rows of vmax, so they run side by side (pipelined: all on the
    // side streams; else, for a large chunk with one live context, on the
    // context's stream and the side streams); the digestion waits for all of
    // them.  Kernels go longest first (generator's cost estimate) to the
    // stream with the least work so far.
    int nside = 0;
    if (pipelined) {
      nside = std::min(ncombo, (int)hfx_ctx::kSide);
      if (xp) {   // the half's last reader, or the build's start
        for (int s = 0; s < nside; ++s)
          HIPOK(hipStreamWaitEvent(c->side[s], c->xp_used[h] ? c->done[h] : c->xp_start, 0));
        if (prof && j == 0) {   // before the class's first kernels: the span runs from the earlier of the two
          HIPOK(hipEventRecord(c->cev[prec][cls][0], c->side[0]));
          HIPOK(hipEventRecord(c->cev[prec][cls][2], c->side[nside > 1 ? 1 : 0]));
        }
      } else if (j >= 2)
        for (int s = 0; s < nside; ++s) HIPOK(hipStreamWaitEvent(c->side[s], c->done[h], 0));
    } else if (!slot && ncombo > 1 && (q1 - q0 >= kConcurrentChunk || a.diag) && live_contexts.load() == 1 &&
               !c->device_chunks) {
      // (the diagonal builds' chunks are small but their kernels few and latency-bound)
      nside = std::min(ncombo - 1, (int)hfx_ctx::kSide);
      int r = ensure_side_streams(c);
      if (r) return r;
      HIPOK(hipEventRecord(c->fork, c->stream));
      for (int s = 0; s < nside; ++s) HIPOK(hipStreamWaitEvent(c->side[s], c->fork, 0));
    }
    const float* cost = prec == 1 ? k.combo_cost32 : k.combo_cost;
    int order[kMaxCombo];
    for (int ic = 0; ic < ncombo; ++ic) order[ic] = ic;
    std::stable_sort(order, order + ncombo, [&](int x, int y) { return cost[x] > cost[y]; });
    const int nstream = pipelined ? nside : nside + 1;
    double load[hfx_ctx::kSide + 1] = {};
    for (int io = 0; io < ncombo; ++io) {
      const int ic = order[io];
      const void* fn = prec == 1 ? k.c32[ic] : k.c64[ic];
      const int cM = prec == 1 ? k.combo_M32[ic] : k.combo_M[ic];
      const size_t shm = (size_t)HFX_BOYS_ROWS * (cM + K) * (prec == 1 ? sizeof(float) : sizeof(double));
      int sb = 0;
      for (int s = 1; s < nstream; ++s)
        if (load[s] < load[sb]) sb = s;
      load[sb] += cost[ic];
      hipStream_t st = pipelined ? c->side[sb] : (sb == 0 ? ms : c->side[sb - 1]);
      const int per_cu = c->cls_per_cu[cls] > 0 ? c->cls_per_cu[cls] : 8;   // tuned launch width
      const dim3 grid(grid_for(q1 - q0, HFX_NT, c->ncu * per_cu));
      // a split combination's pre-kernels first, on the same stream
      const int npre = prec == 1 ? k.npre32[ic] : k.npre[ic];
      for (int jp = 0; jp < npre; ++jp)
        HIPOK(hipLaunchKernel(prec == 1 ? k.pre32[ic][jp] : k.pre64[ic][jp], grid, dim3(HFX_NT), args, shm, st));
      HIPOK(hipLaunchKernel(fn, grid, dim3(HFX_NT), args, shm, st));
    }
    for (int s = 0; s < nside; ++s) {
      HIPOK(hipEventRecord(c->join[s], c->side[s]));
      HIPOK(hipStreamWaitEvent(ms, c->join[s], 0));
    }
    if (a.diag) {   // the diagonal builds' element cut, on the chunk's whole tiles
      const long long nel = (q1 - q0 + 63) / 64 * 64 * ne;
      hipLaunchKernelGGL(k_cut_small, dim3(grid_for(nel, 256, 2048)), dim3(256), 0, ms, eblk, nel);
    }
    DigestArgs g;
    memset(&g, 0, sizeof(g));
    g.list = a.list;
    g.count = a.count;
    g.eblk = eblk;
    g.vmax = vmax;
    g.vstride = (int)chunk;
    g.pairs = c->d_pairs;
    g.P = a.P;
    g.P32 = c->d_P32;
    g.X = a.X;
    g.pairbuf = c->d_pairbuf;
    g.dump = dump;
    g.schw_out = a.schw_out;
    g.thr = a.thr;
    g.nb = c->nb;
    g.diag = a.diag;
    g.jonly = jonly;
    g.q0 = (int)q0;
    g.q1 = (int)q1;
    g.lds_rows = rows ? 1 : 0;
    g.ncombo = ncombo;
    g.tb_offset = (int)(dig_shm - tb_shm);
    for (int i = 0; i < 4; ++i) g.dims[i] = k.dims[i];
    g.swap_bra = k.swap_bra;
    g.swap_ket = k.swap_ket;
    g.thr_mol = a.thr_mol;
    void* gargs[] = {&g};
    // FP32 quartets: blocks stored as floats and contracted with the FP32 copy
    // of P; every quartet's sums are added to X in FP64
    const bool multi = c->nmol > 1;   // molecule-aware form only for batch contexts
    if (one_pass) {
      // one pass over each block, rows A and B of X and of P in LDS; a
      // workgroup per resident slot, each a contiguous range of the chunk
      g.prow_offset = (int)one_prow_off;
      g.tb_offset = (int)one_tb_off;
      g.run_offset = (int)one_run_off;
      const void* fn1 = prec == 1 ? k.dig1f : k.dig1;
      int& occ1 = c->dig1_occ[prec == 1 ? 1 : 0][cls];
      if (occ1 == 0) {
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, fn1, one_nt, one_shm) != hipSuccess || nblk < 1) {
          (void)hipGetLastError();
          nblk = 1;
        }
        occ1 = std::min(nblk, 4);
      }
      HIPOK(hipLaunchKernel(fn1, dim3(c->ncu * occ1), dim3(one_nt), gargs, one_shm, ms));
      if (int r2 = after_digestion(c, pipelined, xp, h)) return r2;
      continue;
    }
    const void* dfn = prec == 1 ? (multi ? k.digfm : k.digf) : (jonly ? (multi ? k.digjm : k.digj) : (multi ? k.digm : k.dig));
    // with LDS rows: one workgroup per resident slot (registers and LDS, as
    // the runtime reports them; at most four per CU), each a contiguous range
    // of steps
    int& occ = c->dig_occ[prec == 1 ? 2 : (jonly ? 1 : 0)][cls];   // one context: one of the two forms
    if (rows && occ == 0) {
      int nblk = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, dfn, HFX_NT, dig_shm) != hipSuccess || nblk < 1) {
        (void)hipGetLastError();
        nblk = (int)std::max<size_t>(1, std::min<size_t>(2, c->max_shm / std::max<size_t>(dig_shm, 1)));
      }
      occ = std::min(nblk, 4);
    }
    const int grid = rows ? grid_for(q1 - q0, HFX_NT, c->ncu * occ) : grid_for(q1 - q0, HFX_NT, c->ncu * 8);
    HIPOK(hipLaunchKernel(dfn, dim3(grid), dim3(HFX_NT), gargs, dig_shm, ms));
    if (int r2 = after_digestion(c, pipelined, xp, h)) return r2;
  }
  if (prof) HIPOK(hipEventRecord(c->cev[prec][cls][1], ms));
  c->class_ran[prec][cls] = 1;
  return HFX_OK;
}

int zero_build_state(hfx_ctx* c) {
  memset(c->class_ran, 0, sizeof(c->class_ran));
  const size_t nx = (size_t)c->nb * c->nb * c->nmol;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for((long long)nx, 256, 2048)), dim3(256), 0, c->stream, c->d_X, nx, 0.0);
  const size_t npb = (size_t)c->totpairs * HFX_PAIRBUF_STRIDE;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for((long long)npb, 256, 2048)), dim3(256), 0, c->stream, c->d_pairbuf, npb,
                     0.0);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, c->stream, c->d_counts, 2 * HFX_NCLASSES, 0);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, c->stream, c->d_stats, 2 * HFX_NCLASSES * kStatN);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c->nmol, 256, 64)), dim3(256), 0, c->stream, c->d_pmax, c->nmol);
  HIPOK(hipGetLastError());
  return HFX_OK;
}

int launch_maxdm(hfx_ctx* c, const double* dP) {
  const long long n = (long long)c->nsp * c->nsp;
  const dim3 grid(grid_for(n, 256, std::max(1, 4096 / c->nmol)), c->nmol);
  hipLaunchKernelGGL(k_maxdm, grid, dim3(256), 0, c->stream, dP, c->nb, c->d_shells, c->d_molsh, c->nsp, c->d_maxdm,
                     c->d_pmax);
  HIPOK(hipGetLastError());
  return HFX_OK;
}

int alloc_debug_blocks(hfx_ctx* c, ClassPlan& cp, long long cap, int prec) {
  double** b = prec ? &cp.blocks32 : &cp.blocks64;
  if (*b) {
    HIPOK(hipStreamSynchronize(c->stream));
    dfree(*b);
  }
  if (cap > 0 && dalloc(b, (size_t)cap * 625)) return HFX_E_NOMEM;
  if (cap > 0) {
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(cap * 625, 256, 4096)), dim3(256), 0, c->stream, *b,
                       (size_t)cap * 625, 0.0);
    HIPOK(hipGetLastError());
  }
  return HFX_OK;
}

// Off-diagonal build on device arrays: F += G(P); quartets whose bound is
// below `split` (> 0) in FP32.  thr_mol (device, one per molecule) replaces
// thr when given.  P and F hold the context's nmol stacked matrices.
int build_nd(hfx_ctx* c, const double* dP, double* dF, double thr, int mode, double split,
             const double* thr_mol = nullptr, const double* split_mol = nullptr) {
  if (!c->have_schwarz) return HFX_E_STATE;
  const bool jonly = mode == HFX_J_ONLY;
  const bool mixed = (split > 0.0 || split_mol) && !jonly;   // split_mol: the FP32 split per molecule
  if (mixed) {
    int r = ensure_mixed_lists(c);
    if (r) return r;
  }
  int r;
  if (!c->capturing) HIPOK(hipEventRecord(c->ev0, c->stream));
  if ((r = zero_build_state(c))) return r;
  if ((r = launch_maxdm(c, dP))) return r;
  if (mixed) {   // the FP32 quartets are contracted with an FP32 copy of P
    const size_t nb2 = (size_t)c->nb * c->nb * c->nmol;
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for((long long)nb2, 256, 2048)), dim3(256), 0, c->stream, dP, c->d_P32, nb2);
    HIPOK(hipGetLastError());
  }
  // The classes' screens are independent but each leaves most of the GPU idle
  // (one wave per bra position): with one live context and launches sized
  // on the host, they run side by side on the context's stream and the side
  // streams (their own survivor-mask buffers), costliest first (bra x ket
  // candidates) to the least-loaded stream, and join before the counts are read.
  const bool conc_screen = live_contexts.load() == 1 && !c->device_chunks && !c->capturing;
  hipStream_t sst[hfx_ctx::kSide + 1] = {c->stream};
  unsigned long long* smask[hfx_ctx::kSide + 1] = {c->d_masks};
  int nsst = 1;
  if (conc_screen) {
    int r = ensure_side_streams(c);
    if (r) return r;
    if (c->masks_side_words < c->masks_words * hfx_ctx::kSide) {
      HIPOK(sync_all_streams(c));   // an earlier build's screens may still read the old buffer
      dfree(c->d_masks_side);
      c->masks_side_words = 0;
      if (dalloc(&c->d_masks_side, c->masks_words * hfx_ctx::kSide)) return HFX_E_NOMEM;
      c->masks_side_words = c->masks_words * hfx_ctx::kSide;
    }
    HIPOK(hipEventRecord(c->fork, c->stream));
    for (int k = 0; k < hfx_ctx::kSide; ++k) {
      HIPOK(hipStreamWaitEvent(c->side[k], c->fork, 0));
      sst[k + 1] = c->side[k];
      smask[k + 1] = c->d_masks_side + (size_t)k * c->masks_words;
    }
    nsst = hfx_ctx::kSide + 1;
  }
  std::vector<int> sorder;
  for (int i = 0; i < (int)c->classes.size(); ++i)
    if (c->classes[i].cap > 0) sorder.push_back(i);
  std::stable_sort(sorder.begin(), sorder.end(), [&](int x, int y) {
    return (double)c->act_len[c->classes[x].t1] * c->act_len[c->classes[x].t2] >
           (double)c->act_len[c->classes[y].t1] * c->act_len[c->classes[y].t2];
  });
  double sload[hfx_ctx::kSide + 1] = {};
  for (int ci : sorder) {
    auto& cp = c->classes[ci];
    int si = 0;
    for (int k = 1; k < nsst; ++k)
      if (sload[k] < sload[si]) si = k;
    sload[si] += (double)c->act_len[cp.t1] * c->act_len[cp.t2];
    ScreenArgs s;
    memset(&s, 0, sizeof(s));
    s.bra = c->d_act[cp.t1];
    s.ket = c->d_act[cp.t2];
    s.binfo = c->d_act_info[cp.t1];
    s.kinfo = c->d_act_info[cp.t2];
    s.schw_bra = c->d_act_schw[cp.t1];
    s.schw_ket = c->d_act_schw[cp.t2];
    s.pairs = c->d_pairs;
    s.maxdm = c->d_maxdm;
    s.md_stride = (long long)c->nsp * c->nsp;
    s.bmol = c->nmol > 1 ? c->d_bmol[cp.t1] : nullptr;
    s.kseg = c->nmol > 1 ? c->d_aseg[cp.t2] : nullptr;
    s.thr_mol = thr_mol;
    s.split_mol = mixed ? split_mol : nullptr;
    s.list64 = cp.list64;
    s.list32 = mixed ? cp.list32 : nullptr;
    s.count64 = c->d_counts + cp.cls;
    s.count32 = c->d_counts + HFX_NCLASSES + cp.cls;
    s.stats64 = c->d_stats + kStatN * cp.cls;
    s.stats32 = c->d_stats + kStatN * (HFX_NCLASSES + cp.cls);
    s.thr = thr;
    s.split = mixed ? split : 0.0;
    s.nbra = c->act_len[cp.t1];
    s.nket = c->act_len[cp.t2];
    s.same = cp.same;
    s.jonly = jonly;
    s.nsh = c->nsp;
    s.all_prims = kClassKernels[cp.cls].all_prims;
    s.pmax = c->d_pmax;
    s.masks = smask[si];
    s.nchunk = c->nchunk_max[cp.t2];
    s.ket_sorted = c->act_sorted[cp.t2];
    // one wave per bra position
    hipLaunchKernelGGL(k_screen, dim3(grid_for(s.nbra, 256 / HFX_WAVE, c->ncu * 16)), dim3(256), 0, sst[si], s);
    HIPOK(hipGetLastError());
  }
  for (int k = 1; k < nsst; ++k) {
    HIPOK(hipEventRecord(c->join[k - 1], sst[k]));
    HIPOK(hipStreamWaitEvent(c->stream, c->join[k - 1], 0));
  }
  // surviving-quartet counts of the block classes decide their chunking
  int hcounts[2 * HFX_NCLASSES] = {0};
  unsigned long long hstats[2 * HFX_NCLASSES * kStatN] = {0};   // bra runs pick the digestion form
  bool any_split = false;
  for (auto& cp : c->classes)
    if (cp.cap > 0 && kClassKernels[cp.cls].split) any_split = true;
  if (any_split && !c->device_chunks) {
    HIPOK(hipMemcpyAsync(hcounts, c->d_counts, sizeof(hcounts), hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipMemcpyAsync(hstats, c->d_stats, sizeof(hstats), hipMemcpyDeviceToHost, c->stream));
    HIPOK(hipStreamSynchronize(c->stream));
  }
  // block-class lists pipelined across lists (launch_split) while this
  // build launches them: one live context, launches sized on the host
  struct XpScope {
    hfx_ctx* c;
    ~XpScope() { c->xp = false; }
  } xp_scope{c};
  if (c->overlap && any_split && !c->device_chunks && !c->capturing && !c->debug_blocks && c->nmol == 1 &&
      live_contexts.load() == 1) {
    if ((r = ensure_side_streams(c))) return r;
    c->xp = true;
    c->xp_j = 0;
    c->xp_used[0] = c->xp_used[1] = false;
    HIPOK(hipEventRecord(c->xp_start, c->stream));   // lists, counts and P32 are ready
  }
  // Class order: with lists overlapping, the fused classes (on the
  // context's stream, like the digestion) go between the block classes,
  // costliest first, so that each runs beside the next block class's
  // integral kernels on the side streams rather than alone.
  std::vector<int> corder;
  for (int i = 0; i < (int)c->classes.size(); ++i) corder.push_back(i);
  if (c->xp) {
    std::vector<int> blk, fus;
    std::vector<double> cost(c->classes.size(), 0.0);
    for (int i = 0; i < (int)c->classes.size(); ++i) {
      const int cl = c->classes[i].cls;
      const ClassKernels& k = kClassKernels[cl];
      for (int p = 0; p < 2; ++p) {
        const unsigned long long* st = hstats + kStatN * (p * HFX_NCLASSES + cl);
        cost[i] += (double)st[1] * k.flops_prim + (double)st[0] * k.flops_quartet;
      }
      (k.split ? blk : fus).push_back(i);
    }
    std::stable_sort(fus.begin(), fus.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    corder.clear();
    size_t f = 0;
    for (int i : blk) {
      corder.push_back(i);
      if (f < fus.size()) corder.push_back(fus[f++]);
    }
    while (f < fus.size()) corder.push_back(fus[f++]);
  }
  // with lists overlapping, every other fused class on a side stream (the
  // context's stream carries every digestion)
  int nfused = 0, side_fused = 0;
  for (int ci : corder) {
    ClassPlan& cp = c->classes[ci];
    if (cp.cap <= 0) continue;
    hipStream_t fst = nullptr;
    if (c->xp && !kClassKernels[cp.cls].split && (nfused++ & 1)) {
      if (side_fused == 0)   // (after the build's set-up, whatever the block lists launched)
        for (int k = 0; k < hfx_ctx::kSide; ++k) HIPOK(hipStreamWaitEvent(c->side[k], c->xp_start, 0));
      fst = c->side[side_fused++ % hfx_ctx::kSide];
    }
    for (int prec = 0; prec < (mixed ? 2 : 1); ++prec) {
      ClassArgs a = base_args(c, dP, thr);
      a.thr_mol = thr_mol;
      a.list = prec ? cp.list32 : cp.list64;
      a.count = c->d_counts + prec * HFX_NCLASSES + cp.cls;
      a.diag = 0;
      if (c->debug_blocks) {
        if ((r = alloc_debug_blocks(c, cp, cp.cap, prec))) return r;
        a.blocks = prec ? cp.blocks32 : cp.blocks64;
      }
      if (kClassKernels[cp.cls].split)
        r = launch_split(c, cp.cls, prec, jonly, a,
                         c->device_chunks ? cp.cap : hcounts[prec * HFX_NCLASSES + cp.cls], a.blocks,
                         c->device_chunks ? 0 : (long long)hstats[kStatN * (prec * HFX_NCLASSES + cp.cls) + 3]);
      else
        r = launch_class(c, cp.cls, prec, jonly, a, cp.cap, fst);
      if (r) return r;
    }
  }
  if (side_fused > 0)   // the side streams' fused classes before X is folded into F
    for (int k = 0; k < hfx_ctx::kSide; ++k) {
      HIPOK(hipEventRecord(c->join[k], c->side[k]));
      HIPOK(hipStreamWaitEvent(c->stream, c->join[k], 0));
    }
  hipLaunchKernelGGL(k_pairbuf_to_x, dim3(grid_for((long long)c->totpairs * HFX_PAIRBUF_STRIDE, 256, 2048)), dim3(256), 0,
                     c->stream, c->d_pairbuf, c->d_pairs, c->d_shells, c->totpairs, c->d_X, c->nb);
  hipLaunchKernelGGL(k_add_sym, dim3(grid_for((long long)c->nb * c->nb * c->nmol, 256, 4096)), dim3(256), 0, c->stream,
                     c->d_X, dF, c->nb, c->nmol);
  HIPOK(hipGetLastError());
  if (!c->capturing) HIPOK(hipEventRecord(c->ev1, c->stream));
  c->last_kind = 1;
  return HFX_OK;
}

// Diagonal-pair build (fock_build_diag, or fock_build_initial_diag when
// initial != 0: no density test, Schwarz values written out); thr_mol as in
// build_nd.
int build_diag(hfx_ctx* c, const double* dP, double* dF, double thr, bool initial, const double* thr_mol = nullptr) {
  int r;
  if (!c->capturing) HIPOK(hipEventRecord(c->ev0, c->stream));
  if ((r = zero_build_state(c))) return r;
  if ((r = launch_maxdm(c, dP))) return r;
  // The pair types' screens and classes are small, latency-bound launches
  // (a few thousand quartets, one thread each): with one live context and
  // launches sized on the host, each pair type's screen and class go to the
  // least-loaded of the context's stream and the side streams, longest
  // first.  A class's time is taken as its longest thread's operations (the
  // pair with the most primitive pairs: np^2 primitive quartets, np bra
  // primitives, one quartet), x1.5 for block classes (combination kernels
  // one after another, then the digestion).  The block classes each get
  // their own share of the integral buffer, and all join before X is folded
  // into F.
  const bool conc = live_contexts.load() == 1 && !c->device_chunks && !c->capturing && !c->debug_blocks;
  hipStream_t streams[1 + hfx_ctx::kSide] = {c->stream};
  int nstream = 1;
  if (conc) {
    if ((r = ensure_side_streams(c))) return r;
    HIPOK(hipEventRecord(c->fork, c->stream));
    for (int k = 0; k < hfx_ctx::kSide; ++k) {
      HIPOK(hipStreamWaitEvent(c->side[k], c->fork, 0));
      streams[nstream++] = c->side[k];
    }
  }
  if (c->diag_cost.empty()) {   // per pair type, from the pairs' primitive counts
    c->diag_cost.assign(HFX_NPAIRTYPES, 0.0);
    for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
      const ClassKernels& k = kClassKernels[kDiagClass[t]];
      for (int g = c->offset[t]; g < c->offset[t] + c->npair[t]; ++g) {
        const double np = k.all_prims ? c->pairs[g].npall : c->pairs[g].np;
        const double ops = (np * np * k.flops_prim + np * k.flops_braprim + k.flops_quartet) * (k.split ? 1.5 : 1.0);
        c->diag_cost[t] = std::max(c->diag_cost[t], ops);
      }
    }
  }
  int order[HFX_NPAIRTYPES];
  for (int t = 0; t < HFX_NPAIRTYPES; ++t) order[t] = t;
  if (conc)   // (else the pair types' order, as every context's build keeps it)
    std::stable_sort(order, order + HFX_NPAIRTYPES, [&](int x, int y) { return c->diag_cost[x] > c->diag_cost[y]; });
  // block classes side by side: one chunk each, in consecutive shares of the buffers
  BlockSlot slots[HFX_NPAIRTYPES];
  bool slotted = conc;
  if (conc) {
    long long eoff = 0, voff = 0;
    for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
      const ClassKernels& k = kClassKernels[kDiagClass[t]];
      if (c->npair[t] == 0 || !k.split) continue;
      const long long ne = (long long)k.dims[0] * k.dims[1] * k.dims[2] * k.dims[3];
      const long long rows = (c->npair[t] + HFX_NT - 1) / HFX_NT * HFX_NT;
      slots[t] = BlockSlot{c->stream, eoff, voff, rows};
      eoff += rows * ne;
      voff += (long long)k.ncombo * rows;
    }
    if ((r = grow_eblk(c, eoff))) return r;
    // (the integral buffer's cap: one class at a time in chunks instead)
    slotted = eoff <= c->eblk_doubles && voff <= c->vmax_rows * kMaxCombo;
  }
  double load[1 + hfx_ctx::kSide] = {};
  for (int io = 0; io < HFX_NPAIRTYPES; ++io) {
    const int t = order[io];
    const int n = c->npair[t];
    if (n == 0) continue;
    const int cls = kDiagClass[t];
    hipStream_t st = c->stream;
    if (conc && (slotted || !kClassKernels[cls].split)) {
      int sb = 0;
      for (int k = 1; k < nstream; ++k)
        if (load[k] < load[sb]) sb = k;
      load[sb] += c->diag_cost[t];
      st = streams[sb];
    }
    DiagScreenArgs s;
    memset(&s, 0, sizeof(s));
    s.pairs = c->d_pairs;
    s.schw = c->d_pair_schw[t];
    s.maxdm = c->d_maxdm;
    s.md_stride = (long long)c->nsp * c->nsp;
    s.thr_mol = thr_mol;
    s.list = c->d_diag_list[t];
    s.count = c->d_counts + cls;
    s.stats = c->d_stats + kStatN * cls;
    s.thr = thr;
    s.first = c->offset[t];
    s.npair = n;
    s.nsh = c->nsp;
    s.check = initial ? 0 : 1;
    s.all_prims = kClassKernels[cls].all_prims;
    hipLaunchKernelGGL(k_screen_diag, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, s);
    HIPOK(hipGetLastError());
    ClassArgs a = base_args(c, dP, thr);
    a.thr_mol = thr_mol;
    a.list = c->d_diag_list[t];
    a.count = c->d_counts + cls;
    a.diag = 1;
    a.schw_out = initial ? c->d_schw_out : nullptr;
    if (c->debug_blocks) {
      ClassPlan* cp = nullptr;
      for (auto& x : c->classes)
        if (x.cls == cls) cp = &x;
      if (cp) {
        if ((r = alloc_debug_blocks(c, *cp, n, 0))) return r;
        a.blocks = cp->blocks64;
      }
    }
    if (kClassKernels[cls].split && conc && slotted) {
      slots[t].st = st;
      r = launch_split(c, cls, 0, false, a, n, a.blocks, 0, &slots[t]);
    } else if (kClassKernels[cls].split) {
      r = launch_split(c, cls, 0, false, a, n, a.blocks);
    } else {
      r = launch_class(c, cls, 0, false, a, n, st);
    }
    if (r) return r;
  }
  if (conc) {
    for (int k = 0; k < hfx_ctx::kSide; ++k) {
      HIPOK(hipEventRecord(c->join[k], c->side[k]));
      HIPOK(hipStreamWaitEvent(c->stream, c->join[k], 0));
    }
  }
  hipLaunchKernelGGL(k_pairbuf_to_x, dim3(grid_for((long long)c->totpairs * HFX_PAIRBUF_STRIDE, 256, 2048)), dim3(256), 0,
                     c->stream, c->d_pairbuf, c->d_pairs, c->d_shells, c->totpairs, c->d_X, c->nb);
  hipLaunchKernelGGL(k_add_sym, dim3(grid_for((long long)c->nb * c->nb * c->nmol, 256, 4096)), dim3(256), 0, c->stream,
                     c->d_X, dF, c->nb, c->nmol);
  HIPOK(hipGetLastError());
  if (!c->capturing) HIPOK(hipEventRecord(c->ev1, c->stream));
  c->last_kind = 2;
  return HFX_OK;
}

// diag + non-diag build of every molecule with per-molecule thresholds
// (device arrays), F optionally cleared first: one SCF iteration's build.
int build_full(hfx_ctx* c, const GraphKey& k) {
  int r;
  if (k.zero_f) {
    const size_t n = (size_t)c->nb * c->nb * c->nmol;
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for((long long)n, 256, 4096)), dim3(256), 0, c->stream, k.F, n, 0.0);
    HIPOK(hipGetLastError());
  }
  if ((r = build_diag(c, k.P, k.F, 0.0, false, k.thr_diag))) return r;
  return build_nd(c, k.P, k.F, 0.0, k.mode, k.split, k.thr_nd, k.split_mol);
}

// The same build as a HIP graph: replayed when one with these arguments was
// captured before; otherwise run once in the fixed launch sequence of
// device-sized chunks (which sizes every buffer it needs), then captured
// from the context's stream for the next call.
int graph_build(hfx_ctx* c, const GraphKey& k) {
  if (c->graphs_off) return build_full(c, k);
  for (const CachedGraph& g : c->graphs)
    if (g.key == k) {
      HIPOK(hipGraphLaunch(g.exec, c->stream));
      return HFX_OK;
    }
  const bool prev = c->device_chunks;
  c->device_chunks = true;
  int r = build_full(c, k);
  if (r) {
    c->device_chunks = prev;
    return r;
  }
  CachedGraph g;
  g.key = k;
  // relaxed: other threads of the process (other contexts' jobs) keep
  // allocating and copying while this stream is captured
  hipError_t e;
  {
    hfx::CaptureSection cs;
    e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      c->capturing = true;
      r = build_full(c, k);
      c->capturing = false;
      const hipError_t e2 = hipStreamEndCapture(c->stream, &g.graph);
      if (e2 != hipSuccess) e = e2;
    }
  }
  if (r == HFX_OK && e == hipSuccess) {
    hfx::DeviceSection ds;
    e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
  }
  c->device_chunks = prev;
  if (r != HFX_OK || e != hipSuccess) {
    // nothing captured ran; F holds the uncaptured build made above
    hfx::DeviceSection ds;
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.graph) hipGraphDestroy(g.graph);
    (void)hipGetLastError();
    fprintf(stderr, "hfx: build not captured (%s); this context's builds run uncaptured\n",
            r != HFX_OK ? "build error" : hipGetErrorString(e));
    c->graphs_off = true;
    return HFX_OK;
  }
  c->graphs.push_back(g);
  return HFX_OK;
}

// Descending Schwarz order and truncation index, as index_sort / find_threshold
// (index_sort.f90:1-119) define them: `last` is the position of the first
// pair below shthresh (that pair included), or n when none is below.
void sort_schwarz(MolPart& mp, int t) {
  const int n = mp.npair[t];
  std::vector<int>& o = mp.order[t];
  o.resize(n);
  for (int i = 0; i < n; ++i) o[i] = i;
  const std::vector<double>& s = mp.schw[t];
  std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return s[a] > s[b]; });
  int last = n;
  if (n > 0 && s[o[0]] < kShthresh) {
    last = 1;
  } else if (n > 0 && s[o[n - 1]] > kShthresh) {
    last = n;
  } else {
    for (int i = 0; i < n; ++i)
      if (s[o[i]] < kShthresh) { last = i + 1; break; }
  }
  mp.last[t] = last;
}

int copy_fock_host(hfx_ctx* c, const double* P, double* F, int nb) {
  const size_t bytes = (size_t)nb * nb * c->nmol * sizeof(double);
  HIPOK(hipMemcpyAsync(c->d_P, P, bytes, hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(c->d_F, F, bytes, hipMemcpyHostToDevice, c->stream));
  return HFX_OK;
}

int fetch_fock_host(hfx_ctx* c, double* F, int nb) {
  const size_t bytes = (size_t)nb * nb * c->nmol * sizeof(double);
  HIPOK(hipMemcpyAsync(F, c->d_F, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPOK(hipStreamSynchronize(c->stream));
  return HFX_OK;
}

// After an initial-diagonal build without caller tables: the Schwarz values
// the build wrote out, sorted per pair type, and the screening lists.
int finish_schwarz(hfx_ctx* c) {
  int r;
  if (!c->have_schwarz) {
    std::vector<double> s(c->totpairs);
    HIPOK(copy_sync(c->stream, s.data(), c->d_schw_out, s.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (MolPart& mp : c->mols)
      for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
        for (int k = 0; k < mp.npair[t]; ++k) mp.schw[t][k] = s[mp.pbeg[t] + k];
        sort_schwarz(mp, t);
      }
    if ((r = upload_active(c))) return r;
  }
  return HFX_OK;
}

}  // namespace

extern "C" {

const char* hfx_version(void) { return "hfx-mi355x 0.1 (gfx950)"; }

int hfx_ctx_create(const hfx_mol_desc* d, hfx_ctx** out) { return hfx_batch_create(d, 1, out); }

int hfx_batch_create(const hfx_mol_desc* descs, int nmol, hfx_ctx** out) {
  if (!descs || !out || nmol < 1) return HFX_E_ARG;
  *out = nullptr;
  for (int m = 0; m < nmol; ++m) {
    const hfx_mol_desc* d = descs + m;
    if (d->nb <= 0 || d->nsh <= 0 || !d->shtyp || !d->nzet || !d->aosh || !d->xsh || !d->zeta || !d->djk ||
        d->maxprim <= 0 || d->maxprim > 16)
      return HFX_E_ARG;
  }
  hfx_ctx* c = new hfx_ctx();
  c->nmol = nmol;
  c->mols.resize(nmol);
  for (int m = 0; m < nmol; ++m) {
    const hfx_mol_desc* d = descs + m;
    MolPart& mp = c->mols[m];
    mp.nb = d->nb;
    mp.nsh = d->nsh;
    mp.sh0 = c->nsh;
    c->nb = std::max(c->nb, d->nb);
    c->nsp = std::max(c->nsp, d->nsh);
    c->nsh += d->nsh;
    bool hasP = false, hasD = false, hasL = false;
    for (int i = 0; i < d->nsh; ++i) {
      Shell s;
      s.type = d->shtyp[i];
      if (letter(s.type) < 0) { delete c; return HFX_E_ARG; }
      hasP |= s.type == 2; hasD |= s.type == 3; hasL |= s.type == 12;
      s.nprim = d->nzet[i];
      if (s.nprim <= 0 || s.nprim > d->maxprim || s.nprim > kMaxShellPrims) { delete c; return HFX_E_UNSUPPORTED; }
      s.ao = d->aosh[i] - 1;
      if (s.ao < 0 || s.ao + ncomp_of(s.type) > d->nb) { delete c; return HFX_E_ARG; }
      for (int k = 0; k < 3; ++k) s.x[k] = d->xsh[3 * i + k];
      for (int k = 0; k < s.nprim; ++k) {
        s.zeta[k] = d->zeta[(size_t)d->maxprim * i + k];
        s.djk[k] = d->djk[(size_t)d->maxprim * i + k];
        s.djkl[k] = (s.type == 12 && d->djkl) ? d->djkl[(size_t)d->maxprim * i + k] : 1.0;
      }
      c->shells.push_back(s);
    }
    if (hasP && hasD && hasL) { delete c; return HFX_E_UNSUPPORTED; }  // spread_basis.f90:174-177
  }

  // pair lists: type-major, then molecule by molecule in each molecule's list order
  const double rt2pi54 = std::sqrt(2.0 * M_PI * M_PI * std::sqrt(M_PI));  // nrtype.f90:12-13
  bool schwarz_given = true;
  int tot = 0;
  std::vector<int> gprs;   // 2 global shells per global pair
  std::vector<int> gmol;   // molecule per global pair
  for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
    c->offset[t] = tot;
    for (int m = 0; m < nmol; ++m) {
      const hfx_mol_desc* d = descs + m;
      MolPart& mp = c->mols[m];
      const int n = d->npair[t];
      mp.npair[t] = n;
      mp.pbeg[t] = tot;
      tot += n;
      if (n == 0) continue;
      if (!d->prs[t] || n < 0) { delete c; return HFX_E_ARG; }
      for (int k = 0; k < n; ++k) {
        const int s1 = d->prs[t][2 * k] - 1, s2 = d->prs[t][2 * k + 1] - 1;
        if (s1 < 0 || s2 < 0 || s1 >= d->nsh || s2 >= d->nsh) { delete c; return HFX_E_ARG; }
        if (kLetters[letter(c->shells[mp.sh0 + s1].type)] != kPairNames[t][0] ||
            kLetters[letter(c->shells[mp.sh0 + s2].type)] != kPairNames[t][1]) { delete c; return HFX_E_ARG; }
        gprs.push_back(mp.sh0 + s1);
        gprs.push_back(mp.sh0 + s2);
        gmol.push_back(m);
      }
      if (d->schw[t] && d->order[t]) {
        mp.schw[t].assign(d->schw[t], d->schw[t] + n);
        mp.order[t].resize(n);
        for (int k = 0; k < n; ++k) {
          mp.order[t][k] = d->order[t][k] - 1;
          if (mp.order[t][k] < 0 || mp.order[t][k] >= n) { delete c; return HFX_E_ARG; }
        }
        mp.last[t] = d->last[t];
        if (mp.last[t] < 0 || mp.last[t] > n) { delete c; return HFX_E_ARG; }
      } else {
        schwarz_given = false;
        mp.schw[t].assign(n, 0.0);
        mp.order[t].resize(n);
        for (int k = 0; k < n; ++k) mp.order[t][k] = k;
        mp.last[t] = 0;
      }
    }
    c->npair[t] = tot - c->offset[t];
  }
  c->totpairs = tot;
  const size_t nb2 = (size_t)c->nb * c->nb;
  // primitive pairs, precalc_kabs.f90:94-115, with the negab ones moved last
  c->pairs.resize(tot);
  for (int g = 0; g < tot; ++g) {
    const int m = gmol[g];
    const Shell& A = c->shells[gprs[2 * g]];
    const Shell& B = c->shells[gprs[2 * g + 1]];
    PairRec& pr = c->pairs[g];
    pr.ax = A.x[0]; pr.ay = A.x[1]; pr.az = A.x[2];
    pr.bx = B.x[0]; pr.by = B.x[1]; pr.bz = B.x[2];
    pr.sha = gprs[2 * g];
    pr.shb = gprs[2 * g + 1];
    pr.aoa = A.ao;
    pr.aob = B.ao;
    pr.mol = m;
    pr.lsa = pr.sha - c->mols[m].sh0;
    pr.lsb = pr.shb - c->mols[m].sh0;
    pr.mbase = (long long)m * (long long)nb2;
    const double dx = A.x[0] - B.x[0], dy = A.x[1] - B.x[1], dz = A.x[2] - B.x[2];
    const double rab2 = dx * dx + dy * dy + dz * dz;
    std::vector<PrimRec> keep, neg;
    std::vector<PrimRec32> keep32, neg32;
    for (int za = 0; za < A.nprim; ++za) {
      for (int zb = 0; zb < B.nprim; ++zb) {
        PrimRec r;
        const double a = A.zeta[za], b = B.zeta[zb];
        r.p = a + b;
        r.pinv = 1.0 / r.p;
        const double q = a * b * r.pinv;
        r.Px = (a * A.x[0] + b * B.x[0]) * r.pinv;
        r.Py = (a * A.x[1] + b * B.x[1]) * r.pinv;
        r.Pz = (a * A.x[2] + b * B.x[2]) * r.pinv;
        const double dd = A.djk[za] * B.djk[zb];
        r.K = rt2pi54 * std::exp(-q * rab2) * r.pinv * dd;
        r.wa = A.type == 12 ? A.djkl[za] : 1.0;
        r.wb = B.type == 12 ? B.djkl[zb] : 1.0;
        const PrimRec32 r32{(float)r.p, (float)r.pinv, (float)r.K, (float)r.wa, (float)r.wb,
                            (float)(r.Px - A.x[0]), (float)(r.Py - A.x[1]), (float)(r.Pz - A.x[2]),
                            (float)(r.Px - B.x[0]), (float)(r.Py - B.x[1]), (float)(r.Pz - B.x[2]), 0.0f};
        if (std::fabs(r.K) * std::sqrt(r.pinv) < kNegab) {
          neg.push_back(r);
          neg32.push_back(r32);
        } else {
          keep.push_back(r);
          keep32.push_back(r32);
        }
      }
    }
    pr.p0 = (int)c->prims.size();
    pr.np = (int)keep.size();
    pr.npall = (int)(keep.size() + neg.size());
    c->prims.insert(c->prims.end(), keep.begin(), keep.end());
    c->prims.insert(c->prims.end(), neg.begin(), neg.end());
    c->prims32.insert(c->prims32.end(), keep32.begin(), keep32.end());
    c->prims32.insert(c->prims32.end(), neg32.begin(), neg32.end());
  }
  // classes
  for (int cl = 0; cl < HFX_NCLASSES; ++cl) {
    ClassPlan cp;
    cp.cls = cl;
    cp.t1 = pair_type_index(kClassNames[cl][0], kClassNames[cl][1]);
    cp.t2 = pair_type_index(kClassNames[cl][2], kClassNames[cl][3]);
    cp.same = cp.t1 == cp.t2;
    if (c->npair[cp.t1] == 0 || c->npair[cp.t2] == 0) continue;
    c->classes.push_back(cp);
  }

  // device
  int r = HFX_OK;
  auto fail = [&](int code) { hfx_ctx_destroy(c); return code; };
  if (hipGetDevice(&c->dev) != hipSuccess) return fail(HFX_E_HIP);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->dev) != hipSuccess) return fail(HFX_E_HIP);
  c->ncu = prop.multiProcessorCount;
  c->max_shm = prop.sharedMemPerBlock;
  c->scratch = scratch_address(c->dev);   // set-up of hfx_ctx_onee_noalloc, once per device
  raise_scratch_limit(c->dev, prop);      // spilling kernels keep their scratch on the queue
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(HFX_E_HIP);
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(HFX_E_HIP);
  std::vector<ShellRec> sr(c->nsh);
  for (int i = 0; i < c->nsh; ++i) sr[i] = ShellRec{c->shells[i].type, c->shells[i].ao, ncomp_of(c->shells[i].type), 0};
  std::vector<int2> molsh(nmol);
  for (int m = 0; m < nmol; ++m) molsh[m] = make_int2(c->mols[m].sh0, c->mols[m].nsh);
  std::vector<double> b64;
  std::vector<float> b32;
  boys_table(b64, b32);
  const size_t nmat = nb2 * nmol;
  if ((r = dalloc(&c->d_pairs, c->pairs.size())) || (r = dalloc(&c->d_prims, c->prims.size())) ||
      (r = dalloc(&c->d_prims32, c->prims32.size())) ||
      (r = dalloc(&c->d_shells, sr.size())) || (r = dalloc(&c->d_molsh, molsh.size())) ||
      (r = dalloc(&c->d_pairbuf, (size_t)c->totpairs * HFX_PAIRBUF_STRIDE)) || (r = dalloc(&c->d_boys64, b64.size())) ||
      (r = dalloc(&c->d_boys32, b32.size())) || (r = dalloc(&c->d_P, nmat)) || (r = dalloc(&c->d_F, nmat)) ||
      (r = dalloc(&c->d_X, nmat)) || (r = dalloc(&c->d_maxdm, (size_t)c->nsp * c->nsp * nmol)) ||
      (r = dalloc(&c->d_pmax, nmol)) || (r = dalloc(&c->d_schw_out, (size_t)tot)) ||
      (r = dalloc(&c->d_counts, 2 * HFX_NCLASSES)) || (r = dalloc(&c->d_stats, 2 * HFX_NCLASSES * kStatN)))
    return fail(r);
  for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
    if (c->npair[t] == 0) continue;
    if ((r = dalloc(&c->d_diag_list[t], c->npair[t])) || (r = dalloc(&c->d_pair_schw[t], c->npair[t])))
      return fail(r);
  }
  if (copy_sync(c->stream, c->d_prims, c->prims.data(), c->prims.size() * sizeof(PrimRec), hipMemcpyHostToDevice) ||
      copy_sync(c->stream, c->d_prims32, c->prims32.data(), c->prims32.size() * sizeof(PrimRec32),
                hipMemcpyHostToDevice) ||
      copy_sync(c->stream, c->d_shells, sr.data(), sr.size() * sizeof(ShellRec), hipMemcpyHostToDevice) ||
      copy_sync(c->stream, c->d_molsh, molsh.data(), molsh.size() * sizeof(int2), hipMemcpyHostToDevice) ||
      copy_sync(c->stream, c->d_boys64, b64.data(), b64.size() * sizeof(double), hipMemcpyHostToDevice) ||
      copy_sync(c->stream, c->d_boys32, b32.data(), b32.size() * sizeof(float), hipMemcpyHostToDevice) ||
      copy_sync(c->stream, c->d_pairs, c->pairs.data(), c->pairs.size() * sizeof(PairRec), hipMemcpyHostToDevice))
    return fail(HFX_E_HIP);
  if (schwarz_given) {
    if ((r = upload_active(c))) return fail(r);
  } else {
    for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
      if (!c->npair[t]) continue;
      std::vector<double> ps(c->npair[t], 0.0);
      if (copy_sync(c->stream, c->d_pair_schw[t], ps.data(), ps.size() * sizeof(double), hipMemcpyHostToDevice))
        return fail(HFX_E_HIP);
    }
  }
  c->counted = true;
  live_contexts.fetch_add(1);
  *out = c;
  return HFX_OK;
}

void hfx_ctx_destroy(hfx_ctx* c) {
  if (!c) return;
  sync_all_streams(c);
  drop_graphs(c);
  for (auto& cp : c->classes) {
    dfree(cp.list64); dfree(cp.list32); dfree(cp.blocks64); dfree(cp.blocks32);
  }
  for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
    dfree(c->d_act[t]); dfree(c->d_act_schw[t]); dfree(c->d_act_info[t]); dfree(c->d_pair_schw[t]);
    dfree(c->d_diag_list[t]); dfree(c->d_bmol[t]); dfree(c->d_aseg[t]);
  }
  dfree(c->d_molsh);
  dfree(c->d_pairs); dfree(c->d_prims); dfree(c->d_prims32); dfree(c->d_shells); dfree(c->d_pairbuf); dfree(c->d_boys64); dfree(c->d_boys32);
  dfree(c->d_P); dfree(c->d_P32); dfree(c->d_F); dfree(c->d_X); dfree(c->d_maxdm); dfree(c->d_schw_out);
  dfree(c->d_onee_items); dfree(c->d_sh_type); dfree(c->d_sh_np); dfree(c->d_sh_ao);
  dfree(c->d_sh_xyz); dfree(c->d_sh_z); dfree(c->d_sh_c); dfree(c->d_sh_cl);
  dfree(c->d_counts); dfree(c->d_stats); dfree(c->d_eblk); dfree(c->d_vmax); dfree(c->d_pmax); dfree(c->d_masks);
  dfree(c->d_masks_side); dfree(c->d_cbuf);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < HFX_NCLASSES; ++k)
      for (int e = 0; e < 3; ++e)
        if (c->cev[p][k][e]) hipEventDestroy(c->cev[p][k][e]);
  for (int s = 0; s < hfx_ctx::kSide; ++s) {
    if (c->side[s]) hipStreamDestroy(c->side[s]);
    if (c->join[s]) hipEventDestroy(c->join[s]);
  }
  for (int h = 0; h < 2; ++h)
    if (c->done[h]) hipEventDestroy(c->done[h]);
  if (c->fork) hipEventDestroy(c->fork);
  if (c->xp_start) hipEventDestroy(c->xp_start);
  if (c->stream) hipStreamDestroy(c->stream);
  if (c->counted) live_contexts.fetch_sub(1);
  delete c;
}

int hfx_set_class_launch(hfx_ctx* c, int cls, int per_cu, long long chunk) {
  if (!c || cls < 0 || cls >= HFX_NCLASSES || per_cu < 0 || per_cu > 64 || chunk < 0 || chunk > (1ll << 30))
    return HFX_E_ARG;
  if (c->cls_per_cu[cls] != per_cu || c->cls_chunk[cls] != chunk / HFX_NT * HFX_NT) {
    drop_graphs(c);   // captured launches carry the old shape
    c->cls_per_cu[cls] = per_cu;
    c->cls_chunk[cls] = chunk > 0 ? std::max<long long>(HFX_NT, chunk / HFX_NT * HFX_NT) : 0;
  }
  return HFX_OK;
}

int hfx_set_option(hfx_ctx* c, int key, double value) {
  if (!c) return HFX_E_ARG;
  if (key == HFX_OPT_MIXED_SPLIT) {
    if (!(value >= 0.0)) return HFX_E_ARG;
    c->split = value;
    return HFX_OK;
  }
  if (key == HFX_OPT_DEBUG_BLOCKS) {
    c->debug_blocks = value != 0.0;
    return HFX_OK;
  }
  if (key == HFX_OPT_CHUNK) {
    if (!(value >= 0.0) || value > 1e12) return HFX_E_ARG;
    c->chunk_opt = (long long)value / HFX_NT * HFX_NT;
    if (value > 0.0 && c->chunk_opt == 0) c->chunk_opt = HFX_NT;
    return HFX_OK;
  }
  if (key == HFX_OPT_OVERLAP) {
    if (value != 0.0 && value != 1.0) return HFX_E_ARG;
    c->overlap = value != 0.0;
    return HFX_OK;
  }
  if (key == HFX_OPT_DIGEST) {
    if (value != 0.0 && value != 1.0 && value != 2.0) return HFX_E_ARG;
    if (c->digest1 != (int)value) drop_graphs(c);   // captured launches carry the old kernels
    c->digest1 = (int)value;
    return HFX_OK;
  }
  if (key == HFX_OPT_PROFILE) {
    if (value != 0.0 && !c->cev[0][0][0]) {
      for (int p = 0; p < 2; ++p)
        for (int k = 0; k < HFX_NCLASSES; ++k)
          for (int e = 0; e < 3; ++e) HIPOK(hipEventCreate(&c->cev[p][k][e]));
    }
    c->profile = value != 0.0;
    return HFX_OK;
  }
  return HFX_E_ARG;
}

int hfx_fock_nd(hfx_ctx* c, const double* P, double* F, int nb, double thr, int mode) {
  if (!c || !P || !F || nb != c->nb || (mode != HFX_JK && mode != HFX_J_ONLY)) return HFX_E_ARG;
  if (!c->have_schwarz) return HFX_E_STATE;
  int r;
  if ((r = copy_fock_host(c, P, F, nb))) return r;
  if ((r = build_nd(c, c->d_P, c->d_F, thr, mode, c->split))) return r;
  return fetch_fock_host(c, F, nb);
}

int hfx_fock_diag(hfx_ctx* c, const double* P, double* F, int nb, double thr) {
  if (!c || !P || !F || nb != c->nb) return HFX_E_ARG;
  if (!c->have_schwarz) return HFX_E_STATE;
  int r;
  if ((r = copy_fock_host(c, P, F, nb))) return r;
  if ((r = build_diag(c, c->d_P, c->d_F, thr, false))) return r;
  return fetch_fock_host(c, F, nb);
}

}  // extern "C"

namespace {

// One-electron set-up of a context: the work items per (molecule, la, lb)
// of the shell-pair parts and the shell tables the kernels read, built on
// the host; where they and the temporaries live is the caller's (library
// allocations, or the fixed scratch array of hfx_ctx_onee_noalloc).
struct OneeTables {
  std::vector<int4> items;
  std::vector<int> count;                    // per (molecule, la, lb)
  std::vector<int> ty, np, ao;
  std::vector<double> xyz, z, cc, cl;
};
struct OneeDev {
  int4* items;
  int *ty, *np, *ao;
  double *xyz, *z, *cc, *cl;
  double *q, *cxyz;                          // charges, coordinates of every molecule's atoms
  double *S, *T, *V;                         // nmol * nb * nb each
};
constexpr int kOneeMaxPrim = 16;

void onee_tables(const hfx_ctx* c, OneeTables& t) {
  auto nparts = [](int ty) { return ty == 12 ? 2 : 1; };
  auto part_l = [](int ty, int p) { return ty == 1 ? 0 : ty == 2 ? 1 : ty == 3 ? 2 : p; };
  t.count.assign((size_t)c->nmol * 9, 0);
  t.items.clear();
  for (int m = 0; m < c->nmol; ++m) {
    const MolPart& mp = c->mols[m];
    std::vector<int4> items[3][3];
    for (int i = mp.sh0; i < mp.sh0 + mp.nsh; ++i)
      for (int j = mp.sh0; j <= i; ++j)
        for (int pa = 0; pa < nparts(c->shells[i].type); ++pa)
          for (int pb = 0; pb < nparts(c->shells[j].type); ++pb)
            items[part_l(c->shells[i].type, pa)][part_l(c->shells[j].type, pb)].push_back(make_int4(i, j, pa, pb));
    for (int la = 0; la < 3; ++la)
      for (int lb = 0; lb < 3; ++lb) {
        t.count[(size_t)m * 9 + la * 3 + lb] = (int)items[la][lb].size();
        t.items.insert(t.items.end(), items[la][lb].begin(), items[la][lb].end());
      }
  }
  const int nsh = c->nsh, mp = kOneeMaxPrim;
  t.ty.resize(nsh); t.np.resize(nsh); t.ao.resize(nsh);
  t.xyz.resize(3 * (size_t)nsh);
  t.z.assign((size_t)mp * nsh, 0.0); t.cc.assign((size_t)mp * nsh, 0.0); t.cl.assign((size_t)mp * nsh, 0.0);
  for (int i = 0; i < nsh; ++i) {
    const Shell& sh = c->shells[i];
    t.ty[i] = sh.type;
    t.np[i] = sh.nprim;
    t.ao[i] = sh.ao;
    for (int k = 0; k < 3; ++k) t.xyz[3 * i + k] = sh.x[k];
    for (int k = 0; k < mp; ++k) {
      t.z[(size_t)mp * i + k] = sh.zeta[k];
      t.cc[(size_t)mp * i + k] = sh.djk[k];
      t.cl[(size_t)mp * i + k] = sh.djkl[k];
    }
  }
}

int onee_upload_tables(hfx_ctx* c, const OneeTables& t, const OneeDev& d) {
  const int nsh = c->nsh;
  HIPOK(hipMemcpyAsync(d.items, t.items.data(), t.items.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(d.ty, t.ty.data(), nsh * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(d.np, t.np.data(), nsh * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(d.ao, t.ao.data(), nsh * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(d.xyz, t.xyz.data(), t.xyz.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(d.z, t.z.data(), t.z.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(d.cc, t.cc.data(), t.cc.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPOK(hipMemcpyAsync(d.cl, t.cl.data(), t.cl.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return HFX_OK;
}

// The launches: charges and coordinates of ntot atoms uploaded, S, T, V
// cleared (a fill kernel), one kernel pair per (molecule, la, lb), the
// matrices copied back to the host; waits for all of it.
int onee_run(hfx_ctx* c, const OneeDev& d, const std::vector<int>& count, const int* natom, int ntot,
             const double* charges, const double* coords, double* S, double* T, double* V) {
  const int nb = c->nb;
  const size_t nmat = (size_t)nb * nb * c->nmol;
  if (ntot > 0) {
    HIPOK(hipMemcpyAsync(d.q, charges, ntot * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPOK(hipMemcpyAsync(d.cxyz, coords, 3 * (size_t)ntot * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  for (double* x : {d.S, d.T, d.V})   // the padding of smaller molecules' blocks stays zero
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for((long long)nmat, 256, 4096)), dim3(256), 0, c->stream, x, nmat, 0.0);
  OneeArgs a;
  memset(&a, 0, sizeof(a));
  a.sh_type = d.ty; a.sh_np = d.np; a.sh_ao = d.ao; a.sh_xyz = d.xyz;
  a.sh_z = d.z; a.sh_c = d.cc; a.sh_cl = d.cl; a.maxprim = kOneeMaxPrim;
  a.boys64 = c->d_boys64;
  a.nb = nb;
  int off = 0, atom0 = 0;
  for (int m = 0; m < c->nmol; ++m) {
    const size_t base = (size_t)m * nb * nb;
    a.charges = d.q + atom0; a.coords = d.cxyz + 3 * atom0; a.natom = natom[m];
    a.S = d.S + base; a.T = d.T + base; a.V = d.V + base;
    atom0 += natom[m];
    for (int la = 0; la < 3; ++la)
      for (int lb = 0; lb < 3; ++lb) {
        const int n = count[(size_t)m * 9 + la * 3 + lb];
        if (n == 0) continue;
        a.items = d.items + off;
        a.nitems = n;
        off += n;
        const int grid = grid_for(n, HFX_NT, c->ncu * 8);
        const size_t shm = (size_t)HFX_BOYS_ROWS * (la + lb + 8) * sizeof(double);
        hipLaunchKernelGGL(kOneeST[la][lb], dim3(grid), dim3(HFX_NT), 0, c->stream, a);
        hipLaunchKernelGGL(kOneeV[la][lb], dim3(grid), dim3(HFX_NT), shm, c->stream, a);
      }
  }
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(S, d.S, nmat * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPOK(hipMemcpyAsync(T, d.T, nmat * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPOK(hipMemcpyAsync(V, d.V, nmat * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPOK(hipStreamSynchronize(c->stream));
  return HFX_OK;
}

int onee_natoms(const hfx_ctx* c, const int* natom, const double* charges, const double* coords, int* ntot) {
  *ntot = 0;
  for (int m = 0; m < c->nmol; ++m) {
    if (natom[m] < 0) return HFX_E_ARG;
    *ntot += natom[m];
  }
  if (*ntot > 0 && (!charges || !coords)) return HFX_E_ARG;
  return HFX_OK;
}

}  // namespace

extern "C" {

// S, T, V on the device: work items per (molecule, la, lb) of the shell-pair
// parts, one generated kernel pair per (la, lb); each molecule's matrices in
// its block of the stacked outputs, its own nuclei for V.  natom[m] atoms of
// molecule m, their charges and coordinates concatenated.  The tables stay
// in the context's memory for later calls; the temporaries are allocated
// per call.
int hfx_batch_onee(hfx_ctx* c, const int* natom, const double* charges, const double* coords, double* S, double* T,
                   double* V) {
  if (!c || !natom || !S || !T || !V) return HFX_E_ARG;
  int ntot = 0, r;
  if ((r = onee_natoms(c, natom, charges, coords, &ntot))) return r;
  const int nsh = c->nsh;
  if (!c->d_onee_items) {
    OneeTables t;
    onee_tables(c, t);
    const int mp = kOneeMaxPrim;
    if (dalloc(&c->d_onee_items, t.items.size()) || dalloc(&c->d_sh_type, nsh) || dalloc(&c->d_sh_np, nsh) ||
        dalloc(&c->d_sh_ao, nsh) || dalloc(&c->d_sh_xyz, 3 * (size_t)nsh) || dalloc(&c->d_sh_z, (size_t)mp * nsh) ||
        dalloc(&c->d_sh_c, (size_t)mp * nsh) || dalloc(&c->d_sh_cl, (size_t)mp * nsh))
      return HFX_E_NOMEM;
    OneeDev d = {c->d_onee_items, c->d_sh_type, c->d_sh_np, c->d_sh_ao, c->d_sh_xyz, c->d_sh_z, c->d_sh_c, c->d_sh_cl,
                 nullptr, nullptr, nullptr, nullptr, nullptr};
    if ((r = onee_upload_tables(c, t, d))) return r;
    c->onee_count = t.count;
  }
  double *d_q = nullptr, *d_xyz = nullptr, *d_S = nullptr, *d_T = nullptr, *d_V = nullptr;
  const size_t nmat = (size_t)c->nb * c->nb * c->nmol;
  r = HFX_OK;
  if (dalloc(&d_q, std::max(ntot, 1)) || dalloc(&d_xyz, 3 * std::max(ntot, 1)) || dalloc(&d_S, nmat) ||
      dalloc(&d_T, nmat) || dalloc(&d_V, nmat))
    r = HFX_E_NOMEM;
  if (r == HFX_OK) {
    OneeDev d = {c->d_onee_items, c->d_sh_type, c->d_sh_np, c->d_sh_ao, c->d_sh_xyz, c->d_sh_z, c->d_sh_c, c->d_sh_cl,
                 d_q, d_xyz, d_S, d_T, d_V};
    r = onee_run(c, d, c->onee_count, natom, ntot, charges, coords, S, T, V);
  }
  hipStreamSynchronize(c->stream);   // nothing queued may still use the temporaries
  dfree(d_q); dfree(d_xyz); dfree(d_S); dfree(d_T); dfree(d_V);
  return r;
}

int hfx_ctx_onee(hfx_ctx* c, int natom, const double* charges, const double* coords, double* S, double* T,
                 double* V) {
  if (!c || natom < 0 || c->nmol != 1) return HFX_E_ARG;
  return hfx_batch_onee(c, &natom, charges, coords, S, T, V);
}

// The same matrices with every device buffer in `scratch` (the tables,
// nuclei and S, T, V are laid out there on each call): no allocation.
int hfx_ctx_onee_noalloc(hfx_ctx* c, int natom, const double* charges, const double* coords, double* S, double* T,
                         double* V) {
  if (!c || natom < 0 || c->nmol != 1 || !S || !T || !V) return HFX_E_ARG;
  int ntot = 0, r;
  if ((r = onee_natoms(c, &natom, charges, coords, &ntot))) return r;
  if (!c->scratch) return HFX_E_HIP;   // its address could not be taken on this device
  OneeTables t;
  onee_tables(c, t);
  const size_t nsh = (size_t)c->nsh, nmat = (size_t)c->nb * c->nb, mp = kOneeMaxPrim;
  size_t off = 0;
  auto take = [&](size_t bytes) {      // 256-byte aligned pieces of the scratch array
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  const size_t o_items = take(t.items.size() * sizeof(int4)), o_ty = take(nsh * sizeof(int)),
               o_np = take(nsh * sizeof(int)), o_ao = take(nsh * sizeof(int)), o_xyz = take(3 * nsh * sizeof(double)),
               o_z = take(mp * nsh * sizeof(double)), o_cc = take(mp * nsh * sizeof(double)),
               o_cl = take(mp * nsh * sizeof(double)), o_q = take(std::max(ntot, 1) * sizeof(double)),
               o_cxyz = take(3 * (size_t)std::max(ntot, 1) * sizeof(double)), o_S = take(nmat * sizeof(double)),
               o_T = take(nmat * sizeof(double)), o_V = take(nmat * sizeof(double));
  if (off > kScratchBytes) {
    fprintf(stderr, "hfx_ctx_onee_noalloc: %zu bytes of device buffers needed (nb = %d), the scratch array holds %zu\n",
            off, c->nb, (size_t)kScratchBytes);
    return HFX_E_UNSUPPORTED;
  }
  unsigned char* base = static_cast<unsigned char*>(c->scratch);
  OneeDev d = {reinterpret_cast<int4*>(base + o_items), reinterpret_cast<int*>(base + o_ty),
               reinterpret_cast<int*>(base + o_np), reinterpret_cast<int*>(base + o_ao),
               reinterpret_cast<double*>(base + o_xyz), reinterpret_cast<double*>(base + o_z),
               reinterpret_cast<double*>(base + o_cc), reinterpret_cast<double*>(base + o_cl),
               reinterpret_cast<double*>(base + o_q), reinterpret_cast<double*>(base + o_cxyz),
               reinterpret_cast<double*>(base + o_S), reinterpret_cast<double*>(base + o_T),
               reinterpret_cast<double*>(base + o_V)};
  if ((r = onee_upload_tables(c, t, d))) return r;
  return onee_run(c, d, t.count, &natom, ntot, charges, coords, S, T, V);
}

int hfx_fock_initial_diag(hfx_ctx* c, const double* P, double* F, int nb, double thr) {
  if (!c || !P || !F || nb != c->nb) return HFX_E_ARG;
  int r;
  if ((r = copy_fock_host(c, P, F, nb))) return r;
  if ((r = build_diag(c, c->d_P, c->d_F, thr, true))) return r;
  if ((r = fetch_fock_host(c, F, nb))) return r;
  return finish_schwarz(c);
}

int hfx_fock_initial_diag_device(hfx_ctx* c, const double* dP, double* dF, double thr, void* stream) {
  if (!c || !dP || !dF) return HFX_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (s && s != c->stream) {
    hipEvent_t e;
    HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPOK(hipEventRecord(e, s));
    HIPOK(hipStreamWaitEvent(c->stream, e, 0));
    hipEventDestroy(e);
  }
  int r = build_diag(c, dP, dF, thr, true);
  if (r) return r;
  if (!c->have_schwarz) HIPOK(hipStreamSynchronize(c->stream));
  if ((r = finish_schwarz(c))) return r;
  if (s && s != c->stream) HIPOK(hipStreamWaitEvent(s, c->ev1, 0));
  return HFX_OK;
}

void hfx_release_cached_memory(void) { release_cached(-1); }

void* hfx_ctx_stream(hfx_ctx* c) { return c ? (void*)c->stream : nullptr; }

int hfx_ctx_nb(hfx_ctx* c) { return c ? c->nb : HFX_E_ARG; }

int hfx_ctx_nmol(hfx_ctx* c) { return c ? c->nmol : HFX_E_ARG; }

int hfx_fock_nd_device(hfx_ctx* c, const double* dP, double* dF, double thr, int mode, void* stream) {
  return c ? hfx::fock_nd_device_split(c, dP, dF, thr, mode, stream, c->split) : HFX_E_ARG;
}

}  // extern "C"

int hfx::fock_nd_device_split(hfx_ctx* c, const double* dP, double* dF, double thr, int mode, void* stream,
                              double split) {
  if (!c || !dP || !dF || (mode != HFX_JK && mode != HFX_J_ONLY) || !(split >= 0.0)) return HFX_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (s && s != c->stream) {
    hipEvent_t e;
    HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPOK(hipEventRecord(e, s));
    HIPOK(hipStreamWaitEvent(c->stream, e, 0));
    hipEventDestroy(e);
  }
  int r = build_nd(c, dP, dF, thr, mode, split);
  if (r) return r;
  if (s && s != c->stream) HIPOK(hipStreamWaitEvent(s, c->ev1, 0));
  return HFX_OK;
}

double hfx::ctx_split(const hfx_ctx* c) { return c ? c->split : 0.0; }

int hfx::ctx_dims(const hfx_ctx* c, int* nmol, int* nb_stride, std::vector<int>* nbs) {
  if (!c) return HFX_E_ARG;
  if (nmol) *nmol = c->nmol;
  if (nb_stride) *nb_stride = c->nb;
  if (nbs) {
    nbs->clear();
    for (const MolPart& mp : c->mols) nbs->push_back(mp.nb);
  }
  return HFX_OK;
}

extern "C" int hfx_batch_fock_device(hfx_ctx* c, const double* dP, double* dF, const double* d_thr_diag,
                                     const double* d_thr_nd, const double* d_split, int mode, int flags,
                                     void* stream) {
  if (!c || !dP || !dF || !d_thr_diag || !d_thr_nd || (mode != HFX_JK && mode != HFX_J_ONLY) ||
      (flags & ~(HFX_BUILD_ZERO_F | HFX_BUILD_GRAPH)))
    return HFX_E_ARG;
  if (!c->have_schwarz) return HFX_E_STATE;
  hipStream_t s = (hipStream_t)stream;
  if (s && s != c->stream) {
    hipEvent_t e;
    HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPOK(hipEventRecord(e, s));
    HIPOK(hipStreamWaitEvent(c->stream, e, 0));
    hipEventDestroy(e);
  }
  GraphKey k;
  k.P = dP;
  k.F = dF;
  k.thr_diag = d_thr_diag;
  k.thr_nd = d_thr_nd;
  k.split_mol = d_split;
  k.split = d_split ? 0.0 : c->split;
  k.mode = mode;
  k.zero_f = (flags & HFX_BUILD_ZERO_F) ? 1 : 0;
  const bool graph = (flags & HFX_BUILD_GRAPH) && !c->debug_blocks && !c->profile;
  int r = graph ? graph_build(c, k) : build_full(c, k);
  if (r) return r;
  c->last_kind = 1;
  if (s && s != c->stream) {
    hipEvent_t e;
    HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPOK(hipEventRecord(e, c->stream));
    HIPOK(hipStreamWaitEvent(s, e, 0));
    hipEventDestroy(e);
  }
  return HFX_OK;
}

extern "C" {

int hfx_fock_diag_device(hfx_ctx* c, const double* dP, double* dF, double thr, void* stream) {
  if (!c || !dP || !dF) return HFX_E_ARG;
  if (!c->have_schwarz) return HFX_E_STATE;
  hipStream_t s = (hipStream_t)stream;
  if (s && s != c->stream) {
    hipEvent_t e;
    HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPOK(hipEventRecord(e, s));
    HIPOK(hipStreamWaitEvent(c->stream, e, 0));
    hipEventDestroy(e);
  }
  int r = build_diag(c, dP, dF, thr, false);
  if (r) return r;
  if (s && s != c->stream) HIPOK(hipStreamWaitEvent(s, c->ev1, 0));
  return HFX_OK;
}

int hfx_get_stats(hfx_ctx* c, hfx_stats* out) {
  if (!c || !out) return HFX_E_ARG;
  memset(out, 0, sizeof(*out));
  if (!c->last_kind) return HFX_OK;
  HIPOK(hipStreamSynchronize(c->stream));
  std::vector<unsigned long long> st(2 * HFX_NCLASSES * kStatN);
  HIPOK(copy_sync(c->stream, st.data(), c->d_stats, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int cl = 0; cl < HFX_NCLASSES; ++cl) {
    const ClassKernels& k = kClassKernels[cl];
    for (int p = 0; p < 2; ++p) {
      const unsigned long long* s = &st[kStatN * (p * HFX_NCLASSES + cl)];
      out->quartets[cl][p] = (long long)s[0];
      out->prim_quartets[cl][p] = (long long)s[1];
      out->bra_prims[cl][p] = (long long)s[2];
      out->bra_runs[cl][p] = (long long)s[3];
      const double f = (double)s[1] * (p ? k.flops_prim_fp32 : k.flops_prim) + (double)s[2] * k.flops_braprim;
      out->class_flops[cl][p] = f + (double)s[0] * k.flops_quartet;
      if (c->profile && c->class_ran[p][cl]) {
        float cms = 0.f, cms2 = 0.f;
        HIPOK(hipEventElapsedTime(&cms, c->cev[p][cl][0], c->cev[p][cl][1]));
        HIPOK(hipEventElapsedTime(&cms2, c->cev[p][cl][2], c->cev[p][cl][1]));
        out->class_ms[cl][p] = std::max(cms, cms2);
      }
      if (p) {
        out->flops_fp32 += f;
        out->flops_fp64 += (double)s[0] * k.flops_quartet;  // digestion is FP64
      } else {
        out->flops_fp64 += f + (double)s[0] * k.flops_quartet;
      }
    }
  }
  float ms = 0.f;
  HIPOK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  out->ms_total = ms;
  return HFX_OK;
}

int hfx_get_schwarz(hfx_ctx* c, int t, double* schw, int* order, int* last) {
  return hfx_batch_get_schwarz(c, 0, t, schw, order, last);
}

int hfx_batch_get_schwarz(hfx_ctx* c, int m, int t, double* schw, int* order, int* last) {
  if (!c || t < 0 || t >= HFX_NPAIRTYPES || m < 0 || m >= c->nmol) return HFX_E_ARG;
  if (!c->have_schwarz) return HFX_E_STATE;
  const MolPart& mp = c->mols[m];
  for (int k = 0; k < mp.npair[t]; ++k) {
    if (schw) schw[k] = mp.schw[t][k];
    if (order) order[k] = mp.order[t][k] + 1;
  }
  if (last) *last = mp.last[t];
  return HFX_OK;
}

long long hfx_debug_blocks(hfx_ctx* c, int cls, int prec, int* quartets, double* blocks) {
  if (!c || cls < 0 || cls >= HFX_NCLASSES || prec < 0 || prec > 1) return HFX_E_ARG;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return HFX_E_HIP;
  ClassPlan* cp = nullptr;
  for (auto& x : c->classes)
    if (x.cls == cls) cp = &x;
  if (!cp) return 0;
  int n = 0;
  if (copy_sync(c->stream, &n, c->d_counts + prec * HFX_NCLASSES + cls, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return HFX_E_HIP;
  if (n <= 0) return 0;
  const int2* list = nullptr;
  if (c->last_kind == 2) {
    for (int t = 0; t < HFX_NPAIRTYPES; ++t)
      if (kDiagClass[t] == cls) list = c->d_diag_list[t];
  } else {
    list = prec ? cp->list32 : cp->list64;
  }
  double* b = prec ? cp->blocks32 : cp->blocks64;
  if (quartets && list && copy_sync(c->stream, quartets, list, (size_t)n * sizeof(int2), hipMemcpyDeviceToHost) != hipSuccess)
    return HFX_E_HIP;
  if (blocks && b && copy_sync(c->stream, blocks, b, (size_t)n * 625 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return HFX_E_HIP;
  return n;
}

int hfx_pair_info(hfx_ctx* c, int g, int* type, int* index) {
  if (!c || g < 0 || g >= c->totpairs) return HFX_E_ARG;
  for (int t = HFX_NPAIRTYPES - 1; t >= 0; --t) {
    if (c->npair[t] && g >= c->offset[t]) {
      if (type) *type = t;
      if (index) *index = g - c->offset[t] + 1;
      return HFX_OK;
    }
  }
  return HFX_E_ARG;
}

}  // extern "C"
