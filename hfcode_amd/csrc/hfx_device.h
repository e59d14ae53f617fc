// Device-side records and helpers shared by every hfx kernel (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HFX_NT 256                 // threads per workgroup of the class kernels
#define DIG1_NT 512                // threads per workgroup of the one-pass digestion (two waves per SIMD)
#ifndef HFX_WAVE
#define HFX_WAVE 64                // wavefront width of gfx950
#endif
#define HFX_INTTHRESH 1.0e-10      // element cut of the diagonal builds (molprops.f90:162)
#define HFX_BOYS_ROWS 449          // Boys grid x = 0, 1/16, ..., 28
#define HFX_BOYS_COLS 16           // F_0 .. F_15 per grid point
#define HFX_BOYS_STEP 16.0         // grid points per unit x
#define HFX_BOYS_XMAX 28.0         // above: asymptotic form (boys.f90:22-24)
#define HFX_PAIRBUF_STRIDE 25      // doubles per shell pair in the X_cd pair buffer (d x d)

namespace hfx {

// One shell pair, in the orientation of the reference pair lists
// (precalc_kabs.f90:38-90): shell a is the first listed shell.  A context
// may hold several molecules (a batch): their shells are numbered one after
// the other, AO offsets are the molecule's own, and each molecule's nb x nb
// matrices (P, X, F) are stacked, molecule m's block at m * nb * nb with nb
// the context's row stride.
struct PairRec {
  double ax, ay, az, bx, by, bz;   // centres (bohr)
  int sha, shb;                    // shells, 0-based, numbered across the context's molecules
  int aoa, aob;                    // first AO in the molecule, 0-based
  int p0;                          // first primitive pair record
  int np;                          // primitive pairs that pass the negab test
  int npall;                       // all primitive pairs (negab ones follow)
  int mol;                         // molecule of the pair (0 with one molecule)
  int lsa, lsb;                    // shells a, b numbered within the molecule
  long long mbase;                 // first element of the molecule's block of P, X and F
};

// One primitive pair (precalc_kabs.f90:94-115): p = za + zb, 1/p, the
// product centre P, K_ab = sqrt(2) pi^(5/4) exp(-q R_AB^2) / p * d_a d_b, and
// the L-shell p/s coefficient ratios of the two primitives (1 otherwise).
struct PrimRec {
  double p, pinv, Px, Py, Pz, K, wa, wb;
};
// The same in FP32 for the FP32 kernels, with P relative to the pair's two
// centres (P - A, P - B, rounded from FP64): differences of centres are
// formed per quartet in FP64, so no FP32 difference of absolute positions
// loses digits.
struct PrimRec32 {
  float p, pinv, K, wa, wb, PAx, PAy, PAz, PBx, PBy, PBz, pad;
};

// Per-shell data the density-maximum kernel needs.
struct ShellRec {
  int type;      // 1 s, 2 p, 3 d, 12 L
  int ao;        // first AO, 0-based
  int ncomp;     // 1, 3, 5, 4
  int pad;
};

struct ClassArgs {
  const int2* list;        // surviving quartets (bra pair, ket pair)
  const int* count;        // number of entries in list
  const PairRec* pairs;
  const PrimRec* prims;
  const PrimRec32* prims32;   // the same records for the FP32 kernels
  const double* P;         // density, nb x nb
  double* X;               // accumulator; F += X + X^T at the end
  const double* boys64;    // HFX_BOYS_ROWS x HFX_BOYS_COLS
  const float* boys32;
  double* blocks;          // optional: dump of every block (625 per quartet)
  double* schw_out;        // optional: sqrt(max|(ab|ab)|) per bra pair
  double* eblk;            // block kernels: integrals of quartets [q0, q1), tiles of 64 quartets, element-major
                           // (stored as R: doubles from the FP64 kernels, floats from the FP32 ones)
  double* vmax;            // block kernels: max|e| of each combination's part, [combination][quartet - q0]
  double thr;
  int nb;
  int diag;                // 0 off-diagonal build, 1 bra==ket build
  int q0, q1;              // block kernels: chunk of the quartet list
  int vstride;             // block kernels: row length of vmax
  int lds_rows;            // fused kernels: 1 = rows of X of the resident bra pair in LDS
  int rows_offset;         // fused kernels: byte offset of those rows in the dynamic LDS
  double* pairbuf;         // fused kernels: J_cd per ket pair (HFX_PAIRBUF_STRIDE doubles per pair)
  const double* thr_mol;   // block test threshold per molecule (device), or null: thr for every quartet
  double* cbuf;            // block kernels: contracted values of split combinations, [value][quartet - q0]
                           // (stored as R, vstride apart)
};

// Resident bra pair for a step [lo, hi) of a bra-contiguous quartet list
// (one quartet per thread): of the current resident bra and the bras of the
// step's first, middle and last quartets, the one that owns the most
// quartets of the step, the current one on ties (its rows then stay in LDS
// without a flush).  Called by every thread of the block.
__device__ inline int pick_resident(const int2* list, int lo, int hi, int cur) {
  const int t = lo + (int)threadIdx.x;
  const int mine = t < hi ? list[t].x : -1;
  const int cand[3] = {list[lo].x, list[lo + (hi - lo - 1) / 2].x, list[hi - 1].x};
  int best = cur, nbest = cur >= 0 ? __syncthreads_count(mine == cur) : -1;
  for (int k = 0; k < 3; ++k) {
    if (cand[k] == best) continue;
    const int n = __syncthreads_count(mine == cand[k]);
    if (n > nbest) {
      best = cand[k];
      nbest = n;
    }
  }
  return best;
}

// Compiler barrier between slices of an unrolled load sequence: keeps the
// scheduler from hoisting every load of a block into registers at once.
#define HFX_LOAD_FENCE() asm volatile("" ::: "memory")
// Materialises a running accumulator at this point: its updates cannot be
// sunk past it (into the conditional flush at the end of a pass), which
// would keep every loaded element alive until then.
#ifndef HFX_PIN
#define HFX_PIN(x) asm volatile("" : "+v"(x))
#endif
// The same for a pointer into global memory, keeping its address space: an
// opaque generic pointer would turn the loads through it into flat loads,
// which also count against the LDS wait counter, so that every LDS wait
// (atomics, row reads) would first drain them.
#ifndef HFX_PIN_GPTR
template <typename T>
__device__ __forceinline__ const T* hfx_pin_global(const T* p) {
  using GT = const __attribute__((address_space(1))) T*;
  GT g = (GT)p;
  asm volatile("" : "+v"(g));
  return (const T*)g;
}
#define HFX_PIN_GPTR(p) ((p) = hfx_pin_global(p))
#endif

// Wave-level barrier: LDS written by one lane is visible to the others.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// 1/x and sqrt(1/x) of the primitive loop and the Boys function: the IEEE
// forms in FP64; in FP32 the hardware's approximations (v_rcp_f32,
// v_rsq_f32, 1 ulp), as the correctly rounded FP32 forms cost about as many
// instructions as the FP64 ones and the FP32 path's results carry ~1e-7
// relative error anyway.  hfx_sqrt_rcp(x, r) = sqrt(r) with r = 1/x.
#ifndef HFX_RCPF
#define HFX_RCPF(x) __builtin_amdgcn_rcpf(x)
#define HFX_RSQF(x) __builtin_amdgcn_rsqf(x)
#define HFX_EXP2F(x) __builtin_amdgcn_exp2f(x)
#endif
__device__ __forceinline__ double hfx_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float hfx_rcp(float x) { return HFX_RCPF(x); }
__device__ __forceinline__ double hfx_sqrt_rcp(double, double r) { return sqrt(r); }
__device__ __forceinline__ float hfx_sqrt_rcp(float x, float) { return HFX_RSQF(x); }

// Sum of x over the lanes of a wave, the same value in every lane (every
// lane of the wave active).  A scan by DPP row shifts 1, 2, 4, 8 (zero fill)
// leaves each row's total of 16 lanes in its lane 15; the four row totals
// are read with readlane.  Cross-lane moves in the VALU: no LDS traffic and
// none of the LDS latency of a ds_bpermute (__shfl_xor) tree.
template <int CTRL>
__device__ __forceinline__ double hfx_dpp(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double hfx_readlane(double x, int l) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double wave_sum(double x) {
#if HFX_WAVE == 1
  return x;
#else
  static_assert(HFX_WAVE == 64, "four rows of 16 lanes");
  x += hfx_dpp<0x111>(x);   // row_shr:1
  x += hfx_dpp<0x112>(x);   // row_shr:2
  x += hfx_dpp<0x114>(x);   // row_shr:4
  x += hfx_dpp<0x118>(x);   // row_shr:8
  return (hfx_readlane(x, 15) + hfx_readlane(x, 31)) + (hfx_readlane(x, 47) + hfx_readlane(x, 63));
#endif
}

template <typename R> struct BoysTerms;
template <> struct BoysTerms<double> { static constexpr int K = 8; };
template <> struct BoysTerms<float> { static constexpr int K = 5; };

template <typename R> __device__ inline const R* boys_src(const ClassArgs& a);
template <> __device__ inline const double* boys_src<double>(const ClassArgs& a) { return a.boys64; }
template <> __device__ inline const float* boys_src<float>(const ClassArgs& a) { return a.boys32; }

// Copy columns 0..ncol-1 of the Boys table into LDS (row-major, ncol wide).
template <typename R>
__device__ inline void stage_boys(R* tab, const ClassArgs& a, int ncol) {
  const R* src = boys_src<R>(a);
  const int n = HFX_BOYS_ROWS * ncol;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int r = i / ncol, c = i - r * ncol;
    tab[i] = src[r * HFX_BOYS_COLS + c];
  }
  __syncthreads();
}

// Boys F_m(T), m = 0..M.
//  T > 28: F_0 = sqrt(pi/T)/2 and upward recursion without the exp term, the
//          long-range branch of the reference (boys.f90:22-24, 117-123);
//  else:   Taylor series about the nearest grid point x_i,
//          F_m(T) = sum_k F_{m+k}(x_i) (x_i - T)^k / k!, using dF_m/dx = -F_{m+1};
//          K terms give < 1e-16 (double) / 1e-10 (float) relative error for
//          |x_i - T| <= 1/32 (FP32: F_M only, then downward recursion).
template <typename R, int M>
__device__ __forceinline__ void boys(const R* __restrict__ tab, int ncol, R T, R* F) {
  constexpr int K = BoysTerms<R>::K;
  if (T > (R)HFX_BOYS_XMAX) {
    const R t = hfx_rcp(T);
    F[0] = (R)0.88622692545275801365 * hfx_sqrt_rcp(T, t);
#pragma unroll
    for (int m = 1; m <= M; ++m) F[m] = F[m - 1] * ((R)m - (R)0.5) * t;
  } else {
    int i = (int)(T * (R)HFX_BOYS_STEP + (R)0.5);
    i = i < 0 ? 0 : (i > HFX_BOYS_ROWS - 1 ? HFX_BOYS_ROWS - 1 : i);
    const R d = (R)i * (R)(1.0 / HFX_BOYS_STEP) - T;
    R c[K];
    c[0] = 1;
#pragma unroll
    for (int k = 1; k < K; ++k) c[k] = c[k - 1] * d * (R)(1.0 / k);
    const R* row = tab + i * ncol;
    if constexpr (sizeof(R) == sizeof(float) && M > 0) {
      // FP32: F_M by the series, the lower orders by the downward recursion
      // F_m = (2T F_{m+1} + exp(-T)) / (2m + 1) (positive terms: no
      // cancellation, relative error <= (M + 1) ulp)
      R s = row[M + K - 1] * c[K - 1];
#pragma unroll
      for (int k = K - 2; k >= 0; --k) s = fma(row[M + k], c[k], s);
      F[M] = s;
      const R e = HFX_EXP2F(T * (R)-1.44269504088896340736);
      const R t2 = T + T;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], e) * (R)(1.0 / (2 * m + 1));
    } else {
#pragma unroll
      for (int m = 0; m <= M; ++m) {
        R s = row[m + K - 1] * c[K - 1];
#pragma unroll
        for (int k = K - 2; k >= 0; --k) s = fma(row[m + k], c[k], s);
        F[m] = s;
      }
    }
  }
}

}  // namespace hfx
