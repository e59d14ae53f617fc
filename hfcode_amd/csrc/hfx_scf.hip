// Direct RHF driver on the GPU (hfx_scf_run): the iteration of
// direct_hartreefock.f90:98-223 with every nb x nb matrix resident in device
// memory.  Products are rocBLAS dgemm, eigensolves rocSOLVER dsyevd (in place
// of wrapper_dsyev), element-wise work and reductions are the kernels below,
// all on the Fock-build context's stream.  Per iteration the host reads one
// small block of scalars (energy, max|resid|, max|P - Pold|, inner products
// of the new DIIS error with the stored ones) and makes the driver's
// decisions; the DIIS system ((n+1) x (n+1), n <= ndiis) is solved on the
// host as wrapper_dsysv does.
//
// Storage: all matrices are symmetric except the eigenvectors, the
// orbitals and the commutator; rocBLAS is column-major, so eigenvectors are
// the columns of cmatob / cmat as in the reference's Fortran arrays.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/hfx.h"
#include "hfx_internal.h"

namespace {

constexpr int kMaxDiis = 16;
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 256;
// scalar block read by the host once per iteration
constexpr int kSlotEnergy = 0;    // sum P (H + F)
constexpr int kSlotResmax = 1;    // max |resid|
constexpr int kSlotDeltap = 2;    // max |P - Pold|
constexpr int kSlotDots = 8;      // <resid_new, resid_slot k>, k < ndiis
constexpr int kNumSlots = kSlotDots + kMaxDiis;

enum RedOp { kDot3 = 0, kDot = 1, kMaxAbs = 2, kMaxAbsDiff = 3 };

struct Ptrs {
  const double* p[kMaxDiis];
  double c[kMaxDiis];
};

// Stage 1 of a deterministic reduction over n elements, for ny second
// operands b.p[y] (kDot; the other ops use y = 0 only): nvb virtual blocks
// per operand, each a partial (sum or max) into part[y * nvb + vb].
template <int OP>
__global__ void k_reduce1(const double* a, Ptrs b, const double* c, long long n, int ny, int nvb, double* part) {
  static __shared__ double sm[kRedThreads];
  for (int y = blockIdx.y; y < ny; y += gridDim.y) {
    const double* bb = b.p[y];
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
      double acc = 0.0;
      for (long long i = (long long)vb * blockDim.x + threadIdx.x; i < n; i += (long long)nvb * blockDim.x) {
        if (OP == kDot3) acc += a[i] * (bb[i] + c[i]);
        else if (OP == kDot) acc += a[i] * bb[i];
        else if (OP == kMaxAbs) acc = fmax(acc, fabs(a[i]));
        else acc = fmax(acc, fabs(a[i] - bb[i]));
      }
      sm[threadIdx.x] = acc;
      __syncthreads();
      for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
          sm[threadIdx.x] = (OP == kDot3 || OP == kDot) ? sm[threadIdx.x] + sm[threadIdx.x + s]
                                                        : fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
      }
      if (threadIdx.x == 0) part[(size_t)y * nvb + vb] = sm[0];
      __syncthreads();
    }
  }
}

// Stage 2: out[slot0 + y] = reduce(part[y * nper .. y * nper + nper)), y < ny.
template <int OP>
__global__ void k_reduce2(const double* part, int nper, int ny, double* out, int slot0) {
  static __shared__ double sm[kRedThreads];
  for (int y = blockIdx.x; y < ny; y += gridDim.x) {
    const double* row = part + (size_t)y * nper;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nper; i += blockDim.x)
      acc = (OP == kDot3 || OP == kDot) ? acc + row[i] : fmax(acc, row[i]);
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s)
        sm[threadIdx.x] = (OP == kDot3 || OP == kDot) ? sm[threadIdx.x] + sm[threadIdx.x + s]
                                                      : fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[slot0 + y] = sm[0];
    __syncthreads();
  }
}

// out = a + s * b
__global__ void k_axpy(const double* a, const double* b, double s, double* out, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = a[i] + s * b[i];
}

// out = A - A^T (nb x nb)
__global__ void k_antisym(const double* A, double* out, int nb) {
  const long long n = (long long)nb * nb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / nb), c = (int)(i - (long long)r * nb);
    out[i] = A[i] - A[(long long)c * nb + r];
  }
}

// out = sum_k w.c[k] * w.p[k], k < m
__global__ void k_combine(Ptrs w, int m, double* out, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < m; ++k) acc += w.c[k] * w.p[k][i];
    out[i] = acc;
  }
}

// Column scaling of a column-major nb x nb matrix: W(:, j) = V(:, j) * f(d[j]),
// f = 1/sqrt (S^-1/2) or the level shift (0 for j < nocc, shift after).
__global__ void k_scale_cols(const double* V, const double* d, double* W, int nb, int mode, int nocc, double shift) {
  const long long n = (long long)nb * nb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(i / nb);
    const double f = mode == 0 ? 1.0 / sqrt(d[j]) : (j < nocc ? 0.0 : shift);
    W[i] = V[i] * f;
  }
}

// rocBLAS handles are reused across solves (creating and destroying one
// allocates and frees its device workspace, and hipFree waits for the whole
// device); a handle is used by one solve at a time.
std::mutex g_handles_mu;
std::vector<rocblas_handle>* g_handles = new std::vector<rocblas_handle>;   // lives until the process ends

rocblas_handle take_handle_impl() {
  {
    std::lock_guard<std::mutex> lk(g_handles_mu);
    if (!g_handles->empty()) {
      rocblas_handle h = g_handles->back();
      g_handles->pop_back();
      return h;
    }
  }
  rocblas_handle h = nullptr;
  hfx::DeviceSection ds;   // allocates the handle's workspace
  return rocblas_create_handle(&h) == rocblas_status_success ? h : nullptr;
}

void give_handle_impl(rocblas_handle h) {
  std::lock_guard<std::mutex> lk(g_handles_mu);
  g_handles->push_back(h);
}

}  // namespace

void* hfx::take_blas_handle() { return take_handle_impl(); }
void hfx::give_blas_handle(void* h) { give_handle_impl((rocblas_handle)h); }

namespace {

rocblas_handle take_handle() { return take_handle_impl(); }
void give_handle(rocblas_handle h) { give_handle_impl(h); }

struct Solver {
  hfx_ctx* ctx = nullptr;
  hipStream_t st = nullptr;
  rocblas_handle hb = nullptr;
  int nb = 0;
  long long n2 = 0;
  std::vector<double*> bufs;
  double *S = nullptr, *H = nullptr, *X = nullptr, *P = nullptr, *Pold = nullptr, *F = nullptr, *Fold = nullptr;
  double *dP = nullptr, *dF = nullptr, *fob = nullptr, *T1 = nullptr, *A = nullptr;
  double *cmatob = nullptr, *cmat = nullptr, *W = nullptr, *eig = nullptr, *E = nullptr;
  double* ring_r[kMaxDiis] = {};
  double* ring_f[kMaxDiis] = {};
  double *part = nullptr, *scal = nullptr;
  double* par = nullptr;          // build parameters in device memory: [thr_full, thr_diag, thr_nd, split, 0]
  double* hpar = nullptr;         // their pinned host copy
  rocblas_int* info = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};

  ~Solver() {
    if (st) hipStreamSynchronize(st);   // nothing queued may still use the buffers
    if (hpar) {
      hfx::DeviceSection ds;
      hipHostFree(hpar);
    }
    for (double* b : bufs) hfx::cached_free(b);
    if (info) hfx::cached_free(info);
    if (hb) give_handle(hb);
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
  }
  int grid() const { return (int)std::min<long long>(2048, (n2 + 255) / 256); }
  double* alloc(long long n) {
    double* p = static_cast<double*>(hfx::cached_alloc((size_t)std::max<long long>(n, 1) * sizeof(double)));
    if (p) bufs.push_back(p);
    return p;
  }
  // C = op(A) * op(B), all nb x nb
  bool gemm(const double* a, bool ta, const double* b, bool tb, double* c, double alpha = 1.0, double beta = 0.0,
            int k = -1) {
    const int kk = k < 0 ? nb : k;
    hfx::DeviceSection ds;   // a call may resize the handle's workspace
    return rocblas_dgemm(hb, ta ? rocblas_operation_transpose : rocblas_operation_none,
                         tb ? rocblas_operation_transpose : rocblas_operation_none, nb, nb, kk, &alpha, a, nb, b,
                         nb, &beta, c, nb) == rocblas_status_success;
  }
  bool eigh(double* a_inout, double* w) {
    hfx::DeviceSection ds;
    return rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, nb, a_inout, nb, w, E, info) ==
           rocblas_status_success;
  }
  template <int OP>
  bool reduce(const double* a, const double* b, const double* c, int slot) {
    Ptrs p;
    memset(&p, 0, sizeof(p));
    p.p[0] = b;
    hipLaunchKernelGGL(k_reduce1<OP>, dim3(kRedBlocks, 1), dim3(kRedThreads), 0, st, a, p, c, n2, 1, kRedBlocks,
                       part);
    hipLaunchKernelGGL(k_reduce2<OP>, dim3(1), dim3(kRedThreads), 0, st, part, kRedBlocks, 1, scal, slot);
    return hipGetLastError() == hipSuccess;
  }
  bool dots(const double* a, double* const* slots, int m, int slot0) {
    Ptrs p;
    memset(&p, 0, sizeof(p));
    for (int k = 0; k < m; ++k) p.p[k] = slots[k];
    hipLaunchKernelGGL(k_reduce1<kDot>, dim3(kRedBlocks, m), dim3(kRedThreads), 0, st, a, p, (const double*)nullptr,
                       n2, m, kRedBlocks, part);
    hipLaunchKernelGGL(k_reduce2<kDot>, dim3(m), dim3(kRedThreads), 0, st, part, kRedBlocks, m, scal, slot0);
    return hipGetLastError() == hipSuccess;
  }
  // P = 2 C(:, 1:nocc) C(:, 1:nocc)^T
  bool density(double* p, int nocc) { return gemm(cmat, false, cmat, true, p, 2.0, 0.0, nocc); }
};

// (n x n) A x = b by Gaussian elimination with partial pivoting (A, b overwritten).
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

#define SCFOK(x)                                                   \
  do {                                                             \
    if (!(x)) {                                                    \
      fprintf(stderr, "hfx_scf: %s failed\n", #x);                 \
      return HFX_E_HIP;                                            \
    }                                                              \
  } while (0)
#define SCFHIP(x) SCFOK((x) == hipSuccess)
#define SCFBUILD(x)      \
  do {                   \
    int r_ = (x);        \
    if (r_) return r_;   \
  } while (0)

int run(Solver& s, const double* S, const double* H, int nocc, double unuc, const double* P0,
        const hfx_scf_options& o, hfx_scf_result* res, double* Pout, double* Fout, double* eigs_out,
        double* history) {
  const auto t_start = std::chrono::steady_clock::now();
  const int nb = s.nb;
  const size_t bytes = (size_t)s.n2 * sizeof(double);
  const int ndiis = o.ndiis;
  if (!(s.hb = take_handle())) return HFX_E_HIP;
  if (rocblas_set_stream(s.hb, s.st) != rocblas_status_success) return HFX_E_HIP;
  double** mats[] = {&s.S, &s.H, &s.X, &s.P, &s.Pold, &s.F, &s.Fold, &s.dP, &s.dF,
                     &s.fob, &s.T1, &s.A, &s.cmatob, &s.cmat, &s.W};
  for (double** m : mats)
    if (!(*m = s.alloc(s.n2))) return HFX_E_NOMEM;
  for (int k = 0; k < ndiis; ++k) {
    if (!(s.ring_r[k] = s.alloc(s.n2)) || !(s.ring_f[k] = s.alloc(s.n2))) return HFX_E_NOMEM;
    SCFHIP(hipMemsetAsync(s.ring_r[k], 0, bytes, s.st));   // entries not written yet stay finite
  }
  if (!(s.eig = s.alloc(nb)) || !(s.E = s.alloc(nb)) || !(s.part = s.alloc((long long)kRedBlocks * kMaxDiis)) ||
      !(s.scal = s.alloc(kNumSlots)))
    return HFX_E_NOMEM;
  if (!(s.info = static_cast<rocblas_int*>(hfx::cached_alloc(sizeof(rocblas_int))))) return HFX_E_NOMEM;
  if (!(s.par = s.alloc(8))) return HFX_E_NOMEM;
  {
    hfx::DeviceSection ds;
    if (hipHostMalloc((void**)&s.hpar, 8 * sizeof(double), 0) != hipSuccess) return HFX_E_NOMEM;
  }
  const double split = hfx::ctx_split(s.ctx);
  SCFHIP(hipEventCreate(&s.ev[0]));
  SCFHIP(hipEventCreate(&s.ev[1]));
  SCFHIP(hipMemcpyAsync(s.S, S, bytes, hipMemcpyHostToDevice, s.st));
  SCFHIP(hipMemcpyAsync(s.H, H, bytes, hipMemcpyHostToDevice, s.st));

  // S^-1/2 = U diag(w^-1/2) U^T (direct_hartreefock.f90:23-29)
  SCFHIP(hipMemcpyAsync(s.T1, s.S, bytes, hipMemcpyDeviceToDevice, s.st));
  SCFOK(s.eigh(s.T1, s.eig));
  {
    rocblas_int info_s = 0;
    SCFHIP(hipMemcpyAsync(&info_s, s.info, sizeof(info_s), hipMemcpyDeviceToHost, s.st));
    SCFHIP(hipStreamSynchronize(s.st));
    if (info_s != 0) return HFX_E_STATE;
  }
  hipLaunchKernelGGL(k_scale_cols, dim3(s.grid()), dim3(256), 0, s.st, s.T1, s.eig, s.W, nb, 0, 0, 0.0);
  SCFOK(s.gemm(s.W, false, s.T1, true, s.X));

  bool diis = false, fullfock = true, uselevshft, have_cmatob = false;
  double thresh = o.lowthresh, eelec = 0.0;
  std::vector<double> eo(ndiis, 0.0);
  // inner products between ring slots (physical indices); slot h holds the newest entry
  std::vector<double> D((size_t)ndiis * ndiis, 0.0);
  int head = 0, nusediis = 0, nfull = 0;
  if (P0) {  // superposition of atomic densities (direct_hartreefock.f90:50-60)
    SCFHIP(hipMemcpyAsync(s.P, P0, bytes, hipMemcpyHostToDevice, s.st));
    uselevshft = false;
  } else {   // core Hamiltonian (:62-93)
    SCFOK(s.gemm(s.H, false, s.X, true, s.T1));
    SCFOK(s.gemm(s.X, false, s.T1, false, s.cmatob));
    SCFOK(s.eigh(s.cmatob, s.eig));
    have_cmatob = true;
    SCFOK(s.gemm(s.X, false, s.cmatob, false, s.cmat));
    SCFOK(s.density(s.P, nocc));
    SCFOK(s.reduce<kDot>(s.P, s.H, nullptr, kSlotEnergy));
    double e0;
    SCFHIP(hipMemcpyAsync(&e0, s.scal + kSlotEnergy, sizeof(double), hipMemcpyDeviceToHost, s.st));
    SCFHIP(hipStreamSynchronize(s.st));
    eelec = e0;
    uselevshft = true;
  }
  SCFHIP(hipMemcpyAsync(s.Pold, s.P, bytes, hipMemcpyDeviceToDevice, s.st));
  const bool levshft = o.levelshift != 0.0;
  double etot = 0.0, fock_ms = 0.0;
  int it = 0;
  bool converged = false;
  std::vector<double> sc(kNumSlots);
  for (it = 1; it <= o.maxiter; ++it) {
    const double eeold = eelec;
    SCFHIP(hipEventRecord(s.ev[0], s.st));
    // Builds after the first iteration replay HIP graphs (hfx_batch_fock_device
    // with HFX_BUILD_GRAPH): thresholds and split are read from device memory,
    // so one graph serves every full build and one every incremental build.
    // The pinned parameters are rewritten only after the previous iteration's
    // copy of them has run (the stream is synchronised once per iteration).
    s.hpar[0] = thresh;                  // full build: diag and non-diag
    s.hpar[1] = thresh;                  // incremental: diag ...
    s.hpar[2] = thresh * o.fdifffac;     // ... and non-diag
    s.hpar[3] = split;                   // full builds: the context's FP32 split
    s.hpar[4] = 0.0;                     // incremental builds: FP64
    SCFHIP(hipMemcpyAsync(s.par, s.hpar, 5 * sizeof(double), hipMemcpyHostToDevice, s.st));
    if (fullfock) {  // F = H + G(P)
      SCFHIP(hipMemcpyAsync(s.F, s.H, bytes, hipMemcpyDeviceToDevice, s.st));
      if (it == 1) {
        SCFBUILD(hfx_fock_initial_diag_device(s.ctx, s.P, s.F, thresh, nullptr));
        SCFBUILD(hfx_fock_nd_device(s.ctx, s.P, s.F, thresh, o.mode, nullptr));
      } else {
        SCFBUILD(hfx_batch_fock_device(s.ctx, s.P, s.F, s.par, s.par, split > 0.0 ? s.par + 3 : nullptr, o.mode,
                                       o.use_graphs ? HFX_BUILD_GRAPH : 0, nullptr));
      }
      if (it > 1) fullfock = false;
      nfull = 0;
    } else {         // F = F_old + G(P - P_old), non-diagonal part at thresh * fdifffac
      // The incremental build stays FP64 whatever the context's mixed split:
      // the split is an absolute bound, and near convergence almost every
      // quartet of dP would fall below it.
      hipLaunchKernelGGL(k_axpy, dim3(s.grid()), dim3(256), 0, s.st, s.P, s.Pold, -1.0, s.dP, s.n2);
      SCFBUILD(hfx_batch_fock_device(s.ctx, s.dP, s.dF, s.par + 1, s.par + 2, split > 0.0 ? s.par + 4 : nullptr,
                                     o.mode, HFX_BUILD_ZERO_F | (o.use_graphs ? HFX_BUILD_GRAPH : 0), nullptr));
      hipLaunchKernelGGL(k_axpy, dim3(s.grid()), dim3(256), 0, s.st, s.Fold, s.dF, 1.0, s.F, s.n2);
      if (++nfull == 15) fullfock = true;
    }
    SCFHIP(hipEventRecord(s.ev[1], s.st));
    // energy, DIIS error, orthogonal-basis Fock matrix (:123-144)
    SCFOK(s.reduce<kDot3>(s.P, s.H, s.F, kSlotEnergy));
    SCFOK(s.gemm(s.F, false, s.X, true, s.T1));
    SCFOK(s.gemm(s.X, false, s.T1, false, s.fob));
    SCFOK(s.gemm(s.P, false, s.S, false, s.T1));
    SCFOK(s.gemm(s.F, false, s.T1, false, s.A));
    hipLaunchKernelGGL(k_antisym, dim3(s.grid()), dim3(256), 0, s.st, s.A, s.T1, nb);
    SCFOK(s.gemm(s.T1, false, s.X, false, s.A));
    head = (head + ndiis - 1) % ndiis;           // shift the history: the new entry is logical 0
    SCFOK(s.gemm(s.X, true, s.A, false, s.ring_r[head]));
    SCFHIP(hipMemcpyAsync(s.ring_f[head], s.fob, bytes, hipMemcpyDeviceToDevice, s.st));
    for (int j = ndiis - 1; j > 0; --j) eo[j] = eo[j - 1];
    SCFOK(s.reduce<kMaxAbs>(s.ring_r[head], nullptr, nullptr, kSlotResmax));
    SCFOK(s.reduce<kMaxAbsDiff>(s.P, s.Pold, nullptr, kSlotDeltap));
    SCFOK(s.dots(s.ring_r[head], s.ring_r, ndiis, kSlotDots));
    SCFHIP(hipMemcpyAsync(sc.data(), s.scal, kNumSlots * sizeof(double), hipMemcpyDeviceToHost, s.st));
    rocblas_int info_prev = 0;   // the last eigensolve so far (S, the guess or the previous iteration)
    SCFHIP(hipMemcpyAsync(&info_prev, s.info, sizeof(info_prev), hipMemcpyDeviceToHost, s.st));
    SCFHIP(hipStreamSynchronize(s.st));
    if (info_prev != 0) return HFX_E_STATE;   // dsyevd did not converge
    float ms = 0.f;
    SCFHIP(hipEventElapsedTime(&ms, s.ev[0], s.ev[1]));
    fock_ms += ms;
    eelec = 0.5 * sc[kSlotEnergy];
    const double deltae = eelec - eeold;
    etot = eelec + unuc;
    std::swap(s.F, s.Fold);                       // faobold = f
    for (int k = 0; k < ndiis; ++k) D[(size_t)head * ndiis + k] = D[(size_t)k * ndiis + head] = sc[kSlotDots + k];
    const double resmax = sc[kSlotResmax];
    eo[0] = eelec;
    if (it >= 2 && resmax < o.diis_start && !diis) {
      diis = true;
      uselevshft = false;
    }
    if (it >= 7 && thresh > o.intthresh && resmax < 0.005) {
      for (double& e : eo) e = 2000.0;
      nusediis = 0;
      fullfock = true;
      uselevshft = false;
      thresh = o.intthresh;
    }
    if (diis) {  // (:161-199)
      nusediis = std::min(nusediis + 1, ndiis);
      if (nusediis > 1) {
        const int n = nusediis, m = n + 1;
        std::vector<double> B((size_t)m * m, 0.0), rhs(m, 0.0);
        for (int j = 0; j < m; ++j) B[(size_t)n * m + j] = B[(size_t)j * m + n] = -1.0;
        B[(size_t)n * m + n] = 0.0;
        rhs[n] = -1.0;
        for (int j = 0; j < n; ++j)
          for (int k = 0; k < n; ++k) B[(size_t)j * m + k] = D[(size_t)((head + j) % ndiis) * ndiis + (head + k) % ndiis];
        int ii = 0;
        for (int j = 1; j < ndiis; ++j)
          if (eo[j] < eo[ii]) ii = j;
        for (int j = 0; j < n; ++j)
          if (j != ii) B[(size_t)j * m + j] *= 1.05;
        if (!solve_small(B, rhs, m)) return HFX_E_STATE;
        double cmax = 0.0;
        for (int j = 0; j < n; ++j) cmax = std::max(cmax, std::fabs(rhs[j]));
        if (cmax < 1.5) {
          Ptrs w;
          memset(&w, 0, sizeof(w));
          for (int j = 0; j < n; ++j) {
            w.p[j] = s.ring_f[(head + j) % ndiis];
            w.c[j] = rhs[j];
          }
          hipLaunchKernelGGL(k_combine, dim3(s.grid()), dim3(256), 0, s.st, w, n, s.fob, s.n2);
        } else {
          nusediis = 0;
        }
      }
    }
    if (uselevshft && levshft && have_cmatob) {  // fob += cmatob shft cmatob^T
      hipLaunchKernelGGL(k_scale_cols, dim3(s.grid()), dim3(256), 0, s.st, s.cmatob, (const double*)nullptr, s.W, nb,
                         1, nocc, o.levelshift);
      SCFOK(s.gemm(s.W, false, s.cmatob, true, s.fob, 1.0, 1.0));
    }
    SCFHIP(hipMemcpyAsync(s.cmatob, s.fob, bytes, hipMemcpyDeviceToDevice, s.st));
    SCFOK(s.eigh(s.cmatob, s.eig));
    have_cmatob = true;
    SCFOK(s.gemm(s.X, false, s.cmatob, false, s.cmat));
    const double deltap = sc[kSlotDeltap];
    if (history) {
      double* h = history + 4 * (size_t)(it - 1);
      h[0] = etot; h[1] = deltae; h[2] = deltap; h[3] = resmax;
    }
    if (std::fabs(deltae) < o.conv_e && deltap < o.conv_p && it > 1) {
      converged = true;
      break;
    }
    std::swap(s.P, s.Pold);                       // pold = p
    SCFOK(s.density(s.P, nocc));
  }
  rocblas_int info = 0;
  SCFHIP(hipMemcpyAsync(&info, s.info, sizeof(info), hipMemcpyDeviceToHost, s.st));
  if (Pout) SCFHIP(hipMemcpyAsync(Pout, s.P, bytes, hipMemcpyDeviceToHost, s.st));
  if (Fout) SCFHIP(hipMemcpyAsync(Fout, s.Fold, bytes, hipMemcpyDeviceToHost, s.st));
  if (eigs_out) SCFHIP(hipMemcpyAsync(eigs_out, s.eig, nb * sizeof(double), hipMemcpyDeviceToHost, s.st));
  SCFHIP(hipStreamSynchronize(s.st));
  if (info != 0) return HFX_E_STATE;   // dsyevd did not converge
  res->energy = etot;
  res->iterations = std::min(it, o.maxiter);
  res->converged = converged ? 1 : 0;
  res->fock_ms = fock_ms;
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return HFX_OK;
}

}  // namespace

extern "C" {

void hfx_scf_defaults(hfx_scf_options* o) {
  if (!o) return;
  o->maxiter = 25;
  o->ndiis = 8;
  o->conv_e = 5e-8;
  o->conv_p = 1e-5;
  o->lowthresh = 5e-7;
  o->intthresh = 1e-10;
  o->fdifffac = 0.1;
  o->diis_start = 0.3;
  o->levelshift = 0.0;
  o->mode = HFX_JK;
  o->use_graphs = 1;
}

int hfx_scf_run(hfx_ctx* ctx, const double* S, const double* H, int nocc, double unuc, const double* P0,
                const hfx_scf_options* opt, hfx_scf_result* res, double* P, double* F, double* eigs,
                double* history) {
  if (!ctx || !S || !H || !res) return HFX_E_ARG;
  hfx_scf_options o;
  hfx_scf_defaults(&o);
  if (opt) o = *opt;
  const int nb = hfx_ctx_nb(ctx);
  if (nb <= 0 || nocc <= 0 || nocc > nb || o.maxiter < 1 || o.ndiis < 1 || o.ndiis > kMaxDiis ||
      (o.mode != HFX_JK && o.mode != HFX_J_ONLY))
    return HFX_E_ARG;
  memset(res, 0, sizeof(*res));
  Solver s;
  s.ctx = ctx;
  s.st = (hipStream_t)hfx_ctx_stream(ctx);
  s.nb = nb;
  s.n2 = (long long)nb * nb;
  return run(s, S, H, nocc, unuc, P0, o, res, P, F, eigs, history);
}

}  // extern "C"
