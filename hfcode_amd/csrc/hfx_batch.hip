// Direct RHF for a batch of molecules on the GPU (hfx_batch_scf_*): the
// iteration of direct_hartreefock.f90:98-223 run for every molecule of a
// batch context at once.  Every molecule's S, H, S^-1/2, P, P_old, F, F_old,
// DIIS error and Fock rings and orbitals stay in device memory, stacked like
// the context's matrices (molecule m's nb x nb block at m * nb * nb; a
// smaller molecule fills the top-left nb_m x nb_m part).  Per iteration:
//
//   device  D = P (full build) or P - P_old (incremental build, FP64)
//           one shared two-electron build of all molecules (a HIP graph),
//           F = H + G or F_old + G, energy, F' = X F X^T, DIIS error
//           X^T (FPS - SPF) X, the rings, max|error|, max|dP|, the new
//           error's inner products with the stored ones (a second graph);
//   host    reads those scalars for all molecules in one copy and makes each
//           molecule's decisions (DIIS start, threshold switch, rebuild
//           cadence, DIIS coefficients, convergence) as hfx_scf_run does;
//   device  DIIS combination, batched eigensolve of F' (rocSOLVER), C = X C',
//           new densities of the molecules still iterating.
//
// Padding: a molecule's unused rows and columns hold the identity in S (so
// S^-1/2 is the identity there) and kPadLevel on the diagonal of F' (its
// padding eigenvalues sort after every real orbital energy, and the padding
// eigenvectors have no component on the molecule's own basis functions), so
// the padded blocks never mix into the molecule's orbitals.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/hfx.h"
#include "hfx_internal.h"

namespace {

constexpr int kMaxDiis = 16;
constexpr int kNT = 256;
constexpr int kRedBlocks = 64;          // partial sums per molecule and reduction
constexpr double kPadLevel = 1.0e3;     // F' diagonal of padding rows (hartree)
// per-molecule scalar slots read by the host each iteration
constexpr int kSlotEnergy = 0;   // sum P (H + F)
constexpr int kSlotResmax = 1;   // max |error|
constexpr int kSlotDeltap = 2;   // max |P - P_old|
constexpr int kSlotDots = 3;     // <error_new, ring slot k>, k < ndiis
constexpr int kSlots = kSlotDots + kMaxDiis;
// per-molecule flags
constexpr int kFull = 1;         // full build (D = P, F = H + G) this iteration
constexpr int kActive = 2;       // still iterating: its density is updated
constexpr int kCombine = 4;      // F' = DIIS combination of the ring

__device__ inline long long grid_stride() { return (long long)gridDim.x * blockDim.x; }

// D_m = P_m (full) or P_m - Pold_m
__global__ void kb_density_input(const double* P, const double* Pold, const int* flags, double* D, long long n2,
                                 int nmol) {
  const long long n = n2 * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const int m = (int)(i / n2);
    D[i] = (flags[m] & kFull) ? P[i] : P[i] - Pold[i];
  }
}

// F_m = (full ? H_m : Fold_m) + G_m
__global__ void kb_assemble(const double* H, const double* Fold, const double* G, const int* flags, double* F,
                            long long n2, int nmol) {
  const long long n = n2 * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const int m = (int)(i / n2);
    F[i] = ((flags[m] & kFull) ? H[i] : Fold[i]) + G[i];
  }
}

// diag of the padding rows of each molecule's F' = kPadLevel (and of S: 1)
__global__ void kb_pad_diag(double* A, const int* nbm, int nb, int nmol, double v) {
  const long long n = (long long)nb * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const int m = (int)(i / nb), r = (int)(i - (long long)m * nb);
    if (r >= nbm[m]) A[(long long)m * nb * nb + (long long)r * nb + r] = v;
  }
}

// out = A - A^T per matrix
__global__ void kb_antisym(const double* A, double* out, int nb, int nmol) {
  const long long n2 = (long long)nb * nb, n = n2 * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const long long m0 = i / n2 * n2, e = i - m0;
    const int r = (int)(e / nb), c = (int)(e - (long long)r * nb);
    out[i] = A[i] - A[m0 + (long long)c * nb + r];
  }
}

// ring slot *head <- R, F' (every molecule)
__global__ void kb_ring_put(const double* R, const double* fob, double* ring_r, double* ring_f, const int* head,
                            long long n) {
  const long long off = (long long)(*head) * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    ring_r[off + i] = R[i];
    ring_f[off + i] = fob[i];
  }
}

__global__ void kb_zero(double* a, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) a[i] = 0.0;
}

__global__ void kb_copy(const double* a, double* b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) b[i] = a[i];
}

// Per-molecule reductions, stage 1: molecule m = blockIdx.y, second operand
// z (ring slot for kDotRing) from blockIdx.z; nvb virtual blocks per (m, z),
// each writing its partial to part[(z * nmol + m) * nvb + vb].
enum { kDot3 = 0, kMaxAbs = 1, kMaxAbsDiff = 2, kDotRing = 3 };
template <int OP>
__global__ void kb_reduce1(const double* a, const double* b, const double* c, long long n2, long long ring_stride,
                           int nmol, int nz, int nvb, double* part) {
  static __shared__ double sm[kNT];
  for (int z = blockIdx.z; z < nz; z += gridDim.z)
    for (int m = blockIdx.y; m < nmol; m += gridDim.y) {
      const double* am = a + (long long)m * n2;
      const double* bm = OP == kDotRing ? b + (long long)z * ring_stride + (long long)m * n2
                                        : (b ? b + (long long)m * n2 : nullptr);
      const double* cm = c ? c + (long long)m * n2 : nullptr;
      for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        double acc = 0.0;
        for (long long i = (long long)vb * blockDim.x + threadIdx.x; i < n2; i += (long long)nvb * blockDim.x) {
          if (OP == kDot3) acc += am[i] * (bm[i] + cm[i]);
          else if (OP == kDotRing) acc += am[i] * bm[i];
          else if (OP == kMaxAbs) acc = fmax(acc, fabs(am[i]));
          else acc = fmax(acc, fabs(am[i] - bm[i]));
        }
        sm[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
          if ((int)threadIdx.x < s)
            sm[threadIdx.x] = (OP == kDot3 || OP == kDotRing) ? sm[threadIdx.x] + sm[threadIdx.x + s]
                                                               : fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
          __syncthreads();
        }
        if (threadIdx.x == 0) part[((long long)z * nmol + m) * nvb + vb] = sm[0];
        __syncthreads();
      }
    }
}

// stage 2: scal[m * kSlots + slot0 + z] = reduce over the partials of (m, z)
template <int OP>
__global__ void kb_reduce2(const double* part, int nper, int nmol, int nz, double* scal, int slot0) {
  static __shared__ double sm[kNT];
  for (int y = blockIdx.x; y < nmol * nz; y += gridDim.x) {
    const int z = y / nmol, m = y - z * nmol;
    const double* row = part + (long long)y * nper;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nper; i += blockDim.x)
      acc = (OP == kDot3 || OP == kDotRing) ? acc + row[i] : fmax(acc, row[i]);
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s)
        sm[threadIdx.x] = (OP == kDot3 || OP == kDotRing) ? sm[threadIdx.x] + sm[threadIdx.x + s]
                                                           : fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) scal[(long long)m * kSlots + slot0 + z] = sm[0];
    __syncthreads();
  }
}

// F'_m = sum_j coef[m][j] ring_f[j]_m for molecules with kCombine (physical slots)
__global__ void kb_combine(const double* ring_f, const double* coef, const int* flags, int ndiis, double* fob,
                           long long n2, int nmol) {
  const long long n = n2 * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const int m = (int)(i / n2);
    if (!(flags[m] & kCombine)) continue;
    double acc = 0.0;
    for (int j = 0; j < ndiis; ++j) {
      const double w = coef[m * kMaxDiis + j];
      if (w != 0.0) acc += w * ring_f[(long long)j * n + i];
    }
    fob[i] = acc;
  }
}

// occupied columns of C (column-major: column j of molecule m at m n2 + j nb),
// the rest zero: C_occ
__global__ void kb_occ_cols(const double* C, const int* nocc, double* Co, int nb, int nmol) {
  const long long n2 = (long long)nb * nb, n = n2 * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const int m = (int)(i / n2);
    const int col = (int)((i - (long long)m * n2) / nb);
    Co[i] = col < nocc[m] ? C[i] : 0.0;
  }
}

// molecules still iterating: Pold = P, P = Pnew
__global__ void kb_advance(double* P, double* Pold, const double* Pnew, const int* flags, long long n2, int nmol) {
  const long long n = n2 * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const int m = (int)(i / n2);
    if (!(flags[m] & kActive)) continue;
    Pold[i] = P[i];
    P[i] = Pnew[i];
  }
}

// S^-1/2 from the eigenpairs of S: W(:, j) = V(:, j) / sqrt(w_j) per molecule
__global__ void kb_scale_cols(const double* V, const double* w, double* W, int nb, int nmol) {
  const long long n2 = (long long)nb * nb, n = n2 * nmol;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const int m = (int)(i / n2);
    const int j = (int)((i - (long long)m * n2) / nb);
    W[i] = V[i] / sqrt(w[(long long)m * nb + j]);
  }
}

bool solve_small(std::vector<double>& A, std::vector<double>& b, int n) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[piv * n + k])) piv = i;
    if (A[piv * n + k] == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * b[j];
    b[k] = s / A[k * n + k];
  }
  return true;
}

// host-side state of one molecule (the driver variables of direct_hartreefock.f90)
struct MolState {
  int nb = 0, nocc = 0;
  double unuc = 0.0;
  int status = 0;            // 0 iterating, 1 converged, 2 maxiter reached, < 0 failed
  int it = 0, iterations = 0;
  bool diis = false, fullfock = true;
  double thresh = 0.0, eelec = 0.0, etot = 0.0, deltae = 0.0, deltap = 0.0, resmax = 0.0;
  int nusediis = 0, nfull = 0;
  std::vector<double> eo, D;
  bool full_now = true;      // kind of the current iteration's build
};

}  // namespace

struct hfx_batch_scf {
  hfx_ctx* ctx = nullptr;
  hipStream_t st = nullptr;
  rocblas_handle hb = nullptr;
  hfx_scf_options o;
  int nmol = 0, nb = 0, ndiis = 0, maxnocc = 0, use_graphs = 1, it = 0;
  long long n2 = 0, nall = 0;
  std::vector<MolState> mol;
  std::vector<double*> bufs;
  double *S = nullptr, *H = nullptr, *X = nullptr, *P = nullptr, *Pold = nullptr, *F = nullptr, *Fold = nullptr;
  double *D = nullptr, *G = nullptr, *fob = nullptr, *T1 = nullptr, *A = nullptr, *R = nullptr;
  double *cmatob = nullptr, *cmat = nullptr, *ring_r = nullptr, *ring_f = nullptr, *eig = nullptr, *E = nullptr;
  double *part = nullptr, *scal = nullptr, *thr_d = nullptr, *thr_n = nullptr, *split_d = nullptr, *coef = nullptr;
  int *flags = nullptr, *head = nullptr, *nbm = nullptr, *nocc_d = nullptr;
  rocblas_int* info = nullptr;
  // pinned host staging of what the host writes and reads every iteration
  double* h_scal = nullptr;      // nmol * kSlots
  double* h_par = nullptr;       // thr_d, thr_n, split (3 nmol), coef (nmol kMaxDiis)
  int* h_iflags = nullptr;       // flags (nmol), head, info (nmol)
  hipGraphExec_t g_post = nullptr;
  hipGraph_t g_post_graph = nullptr;
  int headv = 0;

  ~hfx_batch_scf() {
    if (st) hipStreamSynchronize(st);
    {
      hfx::DeviceSection ds;
      if (g_post) hipGraphExecDestroy(g_post);
      if (g_post_graph) hipGraphDestroy(g_post_graph);
    }
    for (double* b : bufs) hfx::cached_free(b);
    for (void* p : {(void*)flags, (void*)head, (void*)nbm, (void*)nocc_d, (void*)info})
      if (p) hfx::cached_free(p);
    hfx::DeviceSection ds;
    if (h_scal) hipHostFree(h_scal);
    if (h_par) hipHostFree(h_par);
    if (h_iflags) hipHostFree(h_iflags);
    if (hb) hfx::give_blas_handle(hb);
  }
  double* alloc(long long n) {
    double* p = static_cast<double*>(hfx::cached_alloc((size_t)std::max<long long>(n, 1) * sizeof(double)));
    if (p) bufs.push_back(p);
    return p;
  }
  int grid() const { return (int)std::min<long long>(4096, (nall + kNT - 1) / kNT); }
  // rocBLAS / rocSOLVER may resize their workspace (allocation and free, a
  // device-wide operation) in a call: outside capture sections such calls
  // hold the device lock
  static std::unique_lock<std::shared_mutex> device_lock() {
    return hfx::t_capturing ? std::unique_lock<std::shared_mutex>()
                            : std::unique_lock<std::shared_mutex>(hfx::capture_mutex());
  }
  // C_b = op(A_b) op(B_b) for every molecule b
  bool gemm(const double* a, bool ta, const double* b, bool tb, double* c, double alpha = 1.0, int k = -1) {
    const double beta = 0.0;
    auto lk = device_lock();
    return rocblas_dgemm_strided_batched(hb, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                         tb ? rocblas_operation_transpose : rocblas_operation_none, nb, nb,
                                         k < 0 ? nb : k, &alpha, a, nb, n2, b, nb, n2, &beta, c, nb, n2,
                                         nmol) == rocblas_status_success;
  }
  bool eigh(double* a, double* w) {
    auto lk = device_lock();
    return rocsolver_dsyevd_strided_batched(hb, rocblas_evect_original, rocblas_fill_upper, nb, a, nb, n2, w, nb, E,
                                            nb, info, nmol) == rocblas_status_success;
  }
  template <int OP>
  void reduce(const double* a, const double* b, const double* c, int slot0, int nz = 1) {
    hipLaunchKernelGGL(kb_reduce1<OP>, dim3(kRedBlocks, std::min(nmol, 65535), nz), dim3(kNT), 0, st, a, b, c, n2, nall,
                       nmol, nz, kRedBlocks, part);
    hipLaunchKernelGGL(kb_reduce2<OP>, dim3(std::min(4096, nmol * nz)), dim3(kNT), 0, st, part, kRedBlocks, nmol, nz,
                       scal, slot0);
  }
};

namespace {

#define BOK(x)                                             \
  do {                                                     \
    if (!(x)) {                                            \
      fprintf(stderr, "hfx_batch_scf: %s failed\n", #x);   \
      return HFX_E_HIP;                                    \
    }                                                      \
  } while (0)
#define BHIP(x) BOK((x) == hipSuccess)

// density of every molecule from the orbitals in cmat: Pnew = 2 C_occ C_occ^T
int new_density(hfx_batch_scf* b, double* Pnew) {
  hipLaunchKernelGGL(kb_occ_cols, dim3(b->grid()), dim3(kNT), 0, b->st, b->cmat, b->nocc_d, b->T1, b->nb, b->nmol);
  BOK(b->gemm(b->T1, false, b->T1, true, Pnew, 2.0, b->maxnocc));
  return HFX_OK;
}

// orbitals of F' (in fob): eigensolve, C = X C'
int orbitals(hfx_batch_scf* b) {
  hipLaunchKernelGGL(kb_copy, dim3(b->grid()), dim3(kNT), 0, b->st, b->fob, b->cmatob, b->nall);
  BOK(b->eigh(b->cmatob, b->eig));
  BOK(b->gemm(b->X, false, b->cmatob, false, b->cmat));
  return HFX_OK;
}

// Everything between the two-electron build and the host's read-back: F, the
// energy, F' = X F X^T, the DIIS error X^T (F P S - S P F) X, the rings, the
// reductions; F_old = F.  One launch sequence on the stream (captured once).
int post_build(hfx_batch_scf* b) {
  const int g = b->grid();
  hipLaunchKernelGGL(kb_assemble, dim3(g), dim3(kNT), 0, b->st, b->H, b->Fold, b->G, b->flags, b->F, b->n2, b->nmol);
  b->reduce<kDot3>(b->P, b->H, b->F, kSlotEnergy);
  BOK(b->gemm(b->F, false, b->X, true, b->T1));
  BOK(b->gemm(b->X, false, b->T1, false, b->fob));
  hipLaunchKernelGGL(kb_pad_diag, dim3(std::max(1, std::min(4096, (b->nb * b->nmol + kNT - 1) / kNT))), dim3(kNT), 0,
                     b->st, b->fob, b->nbm, b->nb, b->nmol, kPadLevel);
  BOK(b->gemm(b->P, false, b->S, false, b->T1));
  BOK(b->gemm(b->F, false, b->T1, false, b->A));
  hipLaunchKernelGGL(kb_antisym, dim3(g), dim3(kNT), 0, b->st, b->A, b->T1, b->nb, b->nmol);
  BOK(b->gemm(b->T1, false, b->X, false, b->A));
  BOK(b->gemm(b->X, true, b->A, false, b->R));
  hipLaunchKernelGGL(kb_ring_put, dim3(g), dim3(kNT), 0, b->st, b->R, b->fob, b->ring_r, b->ring_f, b->head, b->nall);
  b->reduce<kMaxAbs>(b->R, nullptr, nullptr, kSlotResmax);
  b->reduce<kMaxAbsDiff>(b->P, b->Pold, nullptr, kSlotDeltap);
  b->reduce<kDotRing>(b->R, b->ring_r, nullptr, kSlotDots, b->ndiis);
  hipLaunchKernelGGL(kb_copy, dim3(g), dim3(kNT), 0, b->st, b->F, b->Fold, b->nall);
  BHIP(hipGetLastError());
  return HFX_OK;
}

int run_post_build(hfx_batch_scf* b) {
  if (!b->use_graphs) return post_build(b);
  if (b->g_post) {
    BHIP(hipGraphLaunch(b->g_post, b->st));
    return HFX_OK;
  }
  int r = post_build(b);   // first call: eager (rocBLAS sets up its workspace), then capture
  if (r) return r;
  hipError_t e;
  {
    hfx::CaptureSection cs;
    e = hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      r = post_build(b);
      const hipError_t e2 = hipStreamEndCapture(b->st, &b->g_post_graph);
      if (e2 != hipSuccess) e = e2;
    }
  }
  if (r == HFX_OK && e == hipSuccess) {
    hfx::DeviceSection ds;
    e = hipGraphInstantiate(&b->g_post, b->g_post_graph, nullptr, nullptr, 0);
  }
  if (r != HFX_OK || e != hipSuccess) {   // the sequence stays eager
    hfx::DeviceSection ds;
    (void)hipGetLastError();
    if (b->g_post) hipGraphExecDestroy(b->g_post);
    if (b->g_post_graph) hipGraphDestroy(b->g_post_graph);
    b->g_post = nullptr;
    b->g_post_graph = nullptr;
    fprintf(stderr, "hfx_batch_scf: the post-build sequence is not captured (%s); it runs uncaptured\n",
            hipGetErrorString(e));
    b->use_graphs = 0;
  }
  return HFX_OK;
}

}  // namespace

extern "C" {

int hfx_batch_scf_create(hfx_ctx* ctx, const double* S, const double* H, const int* nocc, const double* unuc,
                         const double* P0, const hfx_scf_options* opt, int use_graphs, hfx_batch_scf** out) {
  if (!ctx || !S || !H || !nocc || !unuc || !out) return HFX_E_ARG;
  *out = nullptr;
  hfx_scf_options o;
  hfx_scf_defaults(&o);
  if (opt) o = *opt;
  if (o.maxiter < 1 || o.ndiis < 1 || o.ndiis > kMaxDiis || (o.mode != HFX_JK && o.mode != HFX_J_ONLY) ||
      o.levelshift != 0.0)
    return HFX_E_ARG;
  int nmol = 0, nb = 0;
  std::vector<int> nbs;
  if (hfx::ctx_dims(ctx, &nmol, &nb, &nbs)) return HFX_E_ARG;
  for (int m = 0; m < nmol; ++m)
    if (nocc[m] <= 0 || nocc[m] > nbs[m]) return HFX_E_ARG;
  hfx_batch_scf* b = new hfx_batch_scf();
  auto fail = [&](int r) {
    delete b;
    return r;
  };
  b->ctx = ctx;
  b->st = (hipStream_t)hfx_ctx_stream(ctx);
  b->o = o;
  b->nmol = nmol;
  b->nb = nb;
  b->ndiis = o.ndiis;
  b->use_graphs = use_graphs ? 1 : 0;
  b->n2 = (long long)nb * nb;
  b->nall = b->n2 * nmol;
  b->mol.resize(nmol);
  for (int m = 0; m < nmol; ++m) {
    MolState& s = b->mol[m];
    s.nb = nbs[m];
    s.nocc = nocc[m];
    s.unuc = unuc[m];
    s.thresh = o.lowthresh;
    s.eo.assign(o.ndiis, 0.0);
    s.D.assign((size_t)o.ndiis * o.ndiis, 0.0);
    b->maxnocc = std::max(b->maxnocc, nocc[m]);
  }
  if (!(b->hb = (rocblas_handle)hfx::take_blas_handle())) return fail(HFX_E_HIP);
  if (rocblas_set_stream(b->hb, b->st) != rocblas_status_success) return fail(HFX_E_HIP);
  double** mats[] = {&b->S, &b->H, &b->X, &b->P, &b->Pold, &b->F, &b->Fold, &b->D, &b->G, &b->fob, &b->T1, &b->A,
                     &b->R, &b->cmatob, &b->cmat};
  for (double** p : mats)
    if (!(*p = b->alloc(b->nall))) return fail(HFX_E_NOMEM);
  if (!(b->ring_r = b->alloc(b->nall * o.ndiis)) || !(b->ring_f = b->alloc(b->nall * o.ndiis)) ||
      !(b->eig = b->alloc((long long)nb * nmol)) || !(b->E = b->alloc((long long)nb * nmol)) ||
      !(b->part = b->alloc((long long)kRedBlocks * nmol * kMaxDiis)) || !(b->scal = b->alloc((long long)nmol * kSlots)) ||
      !(b->thr_d = b->alloc(3LL * nmol)) || !(b->coef = b->alloc((long long)nmol * kMaxDiis)))
    return fail(HFX_E_NOMEM);
  b->thr_n = b->thr_d + nmol;     // thr_d, thr_n, split: one array, uploaded in one copy
  b->split_d = b->thr_d + 2 * nmol;
  b->flags = static_cast<int*>(hfx::cached_alloc(sizeof(int) * nmol));
  b->head = static_cast<int*>(hfx::cached_alloc(sizeof(int)));
  b->nbm = static_cast<int*>(hfx::cached_alloc(sizeof(int) * nmol));
  b->nocc_d = static_cast<int*>(hfx::cached_alloc(sizeof(int) * nmol));
  b->info = static_cast<rocblas_int*>(hfx::cached_alloc(sizeof(rocblas_int) * nmol));
  if (!b->flags || !b->head || !b->nbm || !b->nocc_d || !b->info) return fail(HFX_E_NOMEM);
  {
    hfx::DeviceSection ds;
    if (hipHostMalloc((void**)&b->h_scal, sizeof(double) * nmol * kSlots, 0) != hipSuccess ||
        hipHostMalloc((void**)&b->h_par, sizeof(double) * nmol * (3 + kMaxDiis), 0) != hipSuccess ||
        hipHostMalloc((void**)&b->h_iflags, sizeof(int) * (2 * nmol + 1), 0) != hipSuccess)
      return fail(HFX_E_NOMEM);
  }
  const size_t bytes = (size_t)b->nall * sizeof(double);
  hipStream_t st = b->st;
  if (hipMemcpyAsync(b->S, S, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(b->H, H, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(b->nbm, nbs.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(b->nocc_d, nocc, sizeof(int) * nmol, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(HFX_E_HIP);
  const int g = b->grid();
  // S^-1/2 (direct_hartreefock.f90:23-29), identity padding
  hipLaunchKernelGGL(kb_copy, dim3(g), dim3(kNT), 0, st, b->S, b->T1, b->nall);
  hipLaunchKernelGGL(kb_pad_diag, dim3(std::max(1, std::min(4096, (nb * nmol + kNT - 1) / kNT))), dim3(kNT), 0, st, b->T1,
                     b->nbm, nb, nmol, 1.0);
  hipLaunchKernelGGL(kb_copy, dim3(g), dim3(kNT), 0, st, b->T1, b->S, b->nall);
  if (!b->eigh(b->T1, b->eig)) return fail(HFX_E_HIP);
  hipLaunchKernelGGL(kb_scale_cols, dim3(g), dim3(kNT), 0, st, b->T1, b->eig, b->A, nb, nmol);
  if (!b->gemm(b->A, false, b->T1, true, b->X)) return fail(HFX_E_HIP);
  if (P0) {   // superposition of atomic densities (direct_hartreefock.f90:50-60)
    if (hipMemcpyAsync(b->P, P0, bytes, hipMemcpyHostToDevice, st) != hipSuccess) return fail(HFX_E_HIP);
  } else {    // core Hamiltonian (:62-93): F' = X H X^T
    if (!b->gemm(b->H, false, b->X, true, b->T1) || !b->gemm(b->X, false, b->T1, false, b->fob)) return fail(HFX_E_HIP);
    hipLaunchKernelGGL(kb_pad_diag, dim3(std::max(1, std::min(4096, (nb * nmol + kNT - 1) / kNT))), dim3(kNT), 0, st,
                       b->fob, b->nbm, nb, nmol, kPadLevel);
    int r = orbitals(b);
    if (!r) r = new_density(b, b->P);
    if (r) return fail(r);
  }
  hipLaunchKernelGGL(kb_copy, dim3(g), dim3(kNT), 0, st, b->P, b->Pold, b->nall);
  for (double* z : {b->ring_r, b->ring_f})   // entries not written yet stay finite
    hipLaunchKernelGGL(kb_zero, dim3(g), dim3(kNT), 0, st, z, b->nall * o.ndiis);
  if (hipGetLastError() != hipSuccess) return fail(HFX_E_HIP);
  if (!P0) {   // energy of the core guess: sum P H (direct_hartreefock.f90:83-88)
    hipLaunchKernelGGL(kb_zero, dim3(g), dim3(kNT), 0, st, b->G, b->nall);
    b->reduce<kDot3>(b->P, b->H, b->G, kSlotEnergy);   // sum P (H + 0)
    if (hipMemcpyAsync(b->h_scal, b->scal, sizeof(double) * nmol * kSlots, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(HFX_E_HIP);
    for (int m = 0; m < nmol; ++m) b->mol[m].eelec = b->h_scal[(size_t)m * kSlots + kSlotEnergy];
  }
  rocblas_int ok = 0;
  std::vector<rocblas_int> inf(nmol);
  if (hipMemcpyAsync(inf.data(), b->info, sizeof(rocblas_int) * nmol, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(HFX_E_HIP);
  for (rocblas_int v : inf) ok |= v;
  if (ok) return fail(HFX_E_STATE);   // an eigensolve of S did not converge
  *out = b;
  return HFX_OK;
}

void hfx_batch_scf_destroy(hfx_batch_scf* b) { delete b; }

int hfx_batch_scf_step(hfx_batch_scf* b, int* status, double* energy, int* iterations) {
  if (!b) return HFX_E_ARG;
  const int nmol = b->nmol, ndiis = b->ndiis;
  const hfx_scf_options& o = b->o;
  hipStream_t st = b->st;
  int live = 0;
  for (const MolState& s : b->mol) live += s.status == 0;
  if (live == 0) return 0;
  // the previous step's last copies read the pinned staging buffers, which
  // are rewritten below
  BHIP(hipStreamSynchronize(st));
  b->it += 1;
  // this iteration's build per molecule (direct_hartreefock.f90:100-121)
  double* thr_d = b->h_par;
  double* thr_n = b->h_par + nmol;
  double* split = b->h_par + 2 * nmol;
  double* coef = b->h_par + 3 * nmol;
  int* fl = b->h_iflags;
  const double csplit = hfx::ctx_split(b->ctx);
  for (int m = 0; m < nmol; ++m) {
    MolState& s = b->mol[m];
    fl[m] = 0;
    if (s.status != 0) {   // finished: no work in the shared build, density frozen
      thr_d[m] = thr_n[m] = INFINITY;
      split[m] = 0.0;
      continue;
    }
    s.it += 1;
    s.full_now = s.fullfock;
    if (s.fullfock) {
      thr_d[m] = thr_n[m] = s.thresh;
      split[m] = csplit;
      fl[m] = kFull | kActive;
      if (s.it > 1) s.fullfock = false;
      s.nfull = 0;
    } else {   // incremental: dP, non-diagonal part at thresh * fdifffac, FP64
      thr_d[m] = s.thresh;
      thr_n[m] = s.thresh * o.fdifffac;
      split[m] = 0.0;
      fl[m] = kActive;
      if (++s.nfull == 15) s.fullfock = true;
    }
  }
  BHIP(hipMemcpyAsync(b->thr_d, thr_d, sizeof(double) * 3 * nmol, hipMemcpyHostToDevice, st));   // thr_d, thr_n, split
  BHIP(hipMemcpyAsync(b->flags, fl, sizeof(int) * nmol, hipMemcpyHostToDevice, st));
  b->headv = (b->headv + ndiis - 1) % ndiis;   // shift the history: the new entry is logical 0
  fl[nmol] = b->headv;
  BHIP(hipMemcpyAsync(b->head, fl + nmol, sizeof(int), hipMemcpyHostToDevice, st));
  int r;
  if (b->it == 1) {
    // iteration 1 of a new batch (every molecule starts together): the
    // initial diagonal build makes the Schwarz tables (fock_build_initial_diag)
    hipLaunchKernelGGL(kb_zero, dim3(b->grid()), dim3(kNT), 0, st, b->G, b->nall);
    BHIP(hipGetLastError());
    if ((r = hfx_fock_initial_diag_device(b->ctx, b->P, b->G, o.lowthresh, nullptr))) return r;
    if ((r = hfx::fock_nd_device_split(b->ctx, b->P, b->G, o.lowthresh, o.mode, nullptr, csplit))) return r;
  } else {
    hipLaunchKernelGGL(kb_density_input, dim3(b->grid()), dim3(kNT), 0, st, b->P, b->Pold, b->flags, b->D, b->n2,
                       nmol);
    BHIP(hipGetLastError());
    // per-molecule FP32 split only when the context runs mixed precision at all
    if ((r = hfx_batch_fock_device(b->ctx, b->D, b->G, b->thr_d, b->thr_n, csplit > 0.0 ? b->split_d : nullptr, o.mode,
                                   HFX_BUILD_ZERO_F | (b->use_graphs ? HFX_BUILD_GRAPH : 0), nullptr)))
      return r;
  }
  if ((r = run_post_build(b))) return r;
  BHIP(hipMemcpyAsync(b->h_scal, b->scal, sizeof(double) * nmol * kSlots, hipMemcpyDeviceToHost, st));
  BHIP(hipMemcpyAsync(b->h_iflags + nmol + 1, b->info, sizeof(rocblas_int) * (nmol - 0), hipMemcpyDeviceToHost, st));
  BHIP(hipStreamSynchronize(st));
  // the driver's decisions per molecule (direct_hartreefock.f90:123-216)
  const int h = b->headv;
  for (int m = 0; m < nmol; ++m) {
    MolState& s = b->mol[m];
    for (int j = 0; j < kMaxDiis; ++j) coef[m * kMaxDiis + j] = 0.0;
    if (s.status != 0) continue;
    if (b->h_iflags[nmol + 1 + m] != 0) {   // the previous eigensolve did not converge
      s.status = HFX_E_STATE;
      fl[m] = 0;
      continue;
    }
    const double* sc = b->h_scal + (size_t)m * kSlots;
    const double eeold = s.eelec;
    s.eelec = 0.5 * sc[kSlotEnergy];
    s.deltae = s.eelec - eeold;
    s.etot = s.eelec + s.unuc;
    for (int k = 0; k < ndiis; ++k) s.D[(size_t)h * ndiis + k] = s.D[(size_t)k * ndiis + h] = sc[kSlotDots + k];
    s.resmax = sc[kSlotResmax];
    for (int j = ndiis - 1; j > 0; --j) s.eo[j] = s.eo[j - 1];
    s.eo[0] = s.eelec;
    if (s.it >= 2 && s.resmax < o.diis_start && !s.diis) s.diis = true;
    if (s.it >= 7 && s.thresh > o.intthresh && s.resmax < 0.005) {
      for (double& e : s.eo) e = 2000.0;
      s.nusediis = 0;
      s.fullfock = true;
      s.thresh = o.intthresh;
    }
    if (s.diis) {   // (:161-199)
      s.nusediis = std::min(s.nusediis + 1, ndiis);
      if (s.nusediis > 1) {
        const int n = s.nusediis, mm = n + 1;
        std::vector<double> B((size_t)mm * mm, 0.0), rhs(mm, 0.0);
        for (int j = 0; j < mm; ++j) B[(size_t)n * mm + j] = B[(size_t)j * mm + n] = -1.0;
        B[(size_t)n * mm + n] = 0.0;
        rhs[n] = -1.0;
        for (int j = 0; j < n; ++j)
          for (int k = 0; k < n; ++k) B[(size_t)j * mm + k] = s.D[(size_t)((h + j) % ndiis) * ndiis + (h + k) % ndiis];
        int ii = 0;
        for (int j = 1; j < ndiis; ++j)
          if (s.eo[j] < s.eo[ii]) ii = j;
        for (int j = 0; j < n; ++j)
          if (j != ii) B[(size_t)j * mm + j] *= 1.05;
        if (!solve_small(B, rhs, mm)) {
          s.status = HFX_E_STATE;
          fl[m] = 0;
          continue;
        }
        double cmax = 0.0;
        for (int j = 0; j < n; ++j) cmax = std::max(cmax, std::fabs(rhs[j]));
        if (cmax < 1.5) {
          for (int j = 0; j < n; ++j) coef[m * kMaxDiis + (h + j) % ndiis] = rhs[j];
          fl[m] |= kCombine;
        } else {
          s.nusediis = 0;
        }
      }
    }
    s.deltap = sc[kSlotDeltap];
    s.iterations = s.it;
    if (std::fabs(s.deltae) < o.conv_e && s.deltap < o.conv_p && s.it > 1) {
      s.status = 1;
      fl[m] &= ~kActive;   // converged: its density stays the one of this build
    } else if (s.it >= o.maxiter) {
      s.status = 2;
      fl[m] &= ~kActive;
    }
  }
  BHIP(hipMemcpyAsync(b->coef, coef, sizeof(double) * nmol * kMaxDiis, hipMemcpyHostToDevice, st));
  BHIP(hipMemcpyAsync(b->flags, fl, sizeof(int) * nmol, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(kb_combine, dim3(b->grid()), dim3(kNT), 0, st, b->ring_f, b->coef, b->flags, ndiis, b->fob, b->n2,
                     nmol);
  BHIP(hipGetLastError());
  if ((r = orbitals(b))) return r;
  if ((r = new_density(b, b->A))) return r;
  hipLaunchKernelGGL(kb_advance, dim3(b->grid()), dim3(kNT), 0, st, b->P, b->Pold, b->A, b->flags, b->n2, nmol);
  BHIP(hipGetLastError());
  live = 0;
  for (int m = 0; m < nmol; ++m) {
    const MolState& s = b->mol[m];
    live += s.status == 0;
    if (status) status[m] = s.status;
    if (energy) energy[m] = s.etot;
    if (iterations) iterations[m] = s.iterations;
  }
  return live;
}

int hfx_batch_scf_result(hfx_batch_scf* b, int m, double* P, double* F, double* eigs) {
  if (!b || m < 0 || m >= b->nmol) return HFX_E_ARG;
  const int nbm = b->mol[m].nb, nb = b->nb;
  std::vector<double> tmp((size_t)b->n2);
  hipStream_t st = b->st;
  const double* src[3] = {b->P, b->Fold, nullptr};
  double* dst[3] = {P, F, nullptr};
  for (int k = 0; k < 2; ++k) {
    if (!dst[k]) continue;
    BHIP(hipMemcpyAsync(tmp.data(), src[k] + (size_t)m * b->n2, sizeof(double) * b->n2, hipMemcpyDeviceToHost, st));
    BHIP(hipStreamSynchronize(st));
    for (int i = 0; i < nbm; ++i)
      for (int j = 0; j < nbm; ++j) dst[k][(size_t)i * nbm + j] = tmp[(size_t)i * nb + j];
  }
  if (eigs) {
    BHIP(hipMemcpyAsync(tmp.data(), b->eig + (size_t)m * nb, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    BHIP(hipStreamSynchronize(st));
    for (int i = 0; i < nbm; ++i) eigs[i] = tmp[i];
  }
  return HFX_OK;
}

}  // extern "C"
