// Support kernels of the Fock build: fills, density maxima, screening and
// work-list compaction, and the final symmetrised add into F.
#include "hfx_kernels.h"

namespace hfx {

__global__ void k_fill_f64(double* x, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = v;
}

__global__ void k_fill_i32(int* x, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

__global__ void k_fill_u64(unsigned long long* x, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = 0ull;
}

// maxdm(a,b) = max |P| over the AO block of shell pair (a,b), stored as a
// symmetric nsp x nsp block per molecule (blockIdx.y = molecule; molsh[m] =
// its first shell and shell count; P block of molecule m at m * nb * nb).
// The block is the one evaluate_maxdm reads (evaluate_maxdm.f90:1-101):
// rows = AOs of the first shell of the pair in list order (s before p before
// d before L), columns = AOs of the second.  For a (d, L) pair the reference
// reads rows ao_d..ao_d+3 and columns ao_L..ao_L+4 (evaluate_maxdm.f90:82-90):
// 4 of the 5 d functions and the AO after the L shell.  That block is
// reproduced so that screening decisions match the reference exactly; a
// column past nb is clamped.
//
// pmax[m] receives the maximum over molecule m's entries (the bits of a
// non-negative double, whose integer order is the numeric order); it bounds
// every screening factor, which lets the screening stop early along a
// Schwarz-sorted ket list.
__global__ void k_maxdm(const double* __restrict__ P, int nb, const ShellRec* __restrict__ sh,
                        const int2* __restrict__ molsh, int nsp, double* __restrict__ maxdm,
                        unsigned long long* __restrict__ pmax) {
  const int mol = blockIdx.y;
  const int2 ms = molsh[mol];
  const int nsh = ms.y;
  const double* Pm = P + (size_t)mol * nb * nb;
  double* md = maxdm + (size_t)mol * nsp * nsp;
  const long long n = (long long)nsh * nsh;
  double tmax = 0.0;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / nsh), j = (int)(idx - (long long)i * nsh);
    if (j > i) continue;
    ShellRec s1 = sh[ms.x + i], s2 = sh[ms.x + j];
    // list orientation: canonical letter order s < p < d < L
    auto rank = [](int t) { return t == 1 ? 0 : t == 2 ? 1 : t == 3 ? 2 : 3; };
    if (rank(s2.type) < rank(s1.type)) { ShellRec t = s1; s1 = s2; s2 = t; }
    int nr = s1.ncomp, nc = s2.ncomp;
    if (s1.type == 3 && s2.type == 12) { nr = 4; nc = 5; }
    double m = 0.0;
    for (int r = 0; r < nr; ++r) {
      const int ar = s1.ao + r;
      for (int c = 0; c < nc; ++c) {
        int ac = s2.ao + c;
        if (ac >= nb) ac = nb - 1;
        m = fmax(m, fabs(Pm[(size_t)ar * nb + ac]));
      }
    }
    md[(size_t)i * nsp + j] = m;
    md[(size_t)j * nsp + i] = m;
    tmax = fmax(tmax, m);
  }
  for (int o = HFX_WAVE / 2; o > 0; o >>= 1) tmax = fmax(tmax, __shfl_xor(tmax, o));
  if (__lane_id() == 0) atomicMax(pmax + mol, (unsigned long long)__double_as_longlong(tmax));
}

// Wave-aggregated append of (bra, ket) to a list: one atomic per wave.
__device__ inline void wave_append(bool pred, int2 v, int2* list, int* count, unsigned long long* stats,
                                   unsigned long long primq, unsigned long long brap) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  const int n = __popcll(mask);
  int base = 0;
  if (lane == leader) base = atomicAdd(count, n);
  base = __shfl(base, leader);
  if (pred) {
    const int off = __popcll(mask & ((1ull << lane) - 1ull));
    list[base + off] = v;
  }
  // statistics: wave sums of primitive quartets and bra primitive iterations
  unsigned long long a = pred ? primq : 0ull, b = pred ? brap : 0ull;
  for (int o = HFX_WAVE / 2; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  if (lane == leader) {
    atomicAdd(&stats[0], (unsigned long long)n);
    atomicAdd(&stats[1], a);
    atomicAdd(&stats[2], b);
  }
}

__device__ inline unsigned long long wave_sum(unsigned long long a) {
  for (int o = HFX_WAVE / 2; o > 0; o >>= 1) a += __shfl_xor(a, o);
  return a;
}

// Screening of one class (fock_build_non_diag.f90:63-67, e.g. :58-67):
// bound = schw_bra * schw_ket * max(maxdm(ab), maxdm(cd),
//                                   1/4 max(maxdm(ac), maxdm(ad), maxdm(bc), maxdm(bd)))
// (J only: max(maxdm(ab), maxdm(cd)), fock_build_non_diag_j.f90:61-62);
// survivors have bound >= thr.  Same-type classes take unordered pairs of
// distinct pairs (ket position < bra position, fock_build_non_diag.f90:52-57).
// With mixed precision, survivors with bound < split go to the FP32 list.
//
// One wave per bra position.  Kets are in descending Schwarz order, so
// (schw_bra * schw_ket) * max|P| < thr holds from some ket position on, and
// no quartet past it can survive (the product is monotone in each factor);
// a 64-ary search finds that position.  The survivors of the bra are then
// marked in bit masks (one word per 64 kets), the bra's range in each list
// is reserved with one atomic, and the survivors are written there grouped
// by ket primitive count: every bra's quartets are contiguous in its list,
// which the digestion of the block classes relies on to accumulate rows of
// X in LDS.
__global__ void k_screen(const ScreenArgs s) {
  // per wave (four per workgroup): survivors per ket key (FP64, FP32 list),
  // then each key's next list position (FP64, FP32); lane 0 writes them
  __shared__ int kcount[4][4][128];   // launched with 256 threads
  int (*kw)[128] = kcount[(threadIdx.x / HFX_WAVE) & 3];
  const int lane = __lane_id();
  const int nw = gridDim.x * (blockDim.x / HFX_WAVE);
  const int ns = s.nsh;
  for (int i = blockIdx.x * (blockDim.x / HFX_WAVE) + threadIdx.x / HFX_WAVE; i < s.nbra; i += nw) {
    // the bra's molecule: its kets, maxdm block, density maximum and threshold
    const int mol = s.bmol ? s.bmol[i] : 0;
    const int kb = s.kseg ? s.kseg[mol] : 0;
    const int ke = s.kseg ? s.kseg[mol + 1] : s.nket;
    const double* md = s.maxdm + mol * s.md_stride;
    const double pm = __longlong_as_double((long long)s.pmax[mol]);
    const double thr = s.thr_mol ? s.thr_mol[mol] : s.thr;
    const double split = s.split_mol ? s.split_mol[mol] : s.split;
    const int jend = s.same ? min(i, ke) : ke;
    const double sb = s.schw_bra[i];
    // first ket position whose bound cannot reach thr
    int lo = kb, hi = jend;
    if (s.ket_sorted) {
      while (hi - lo > HFX_WAVE) {
        const int step = (hi - lo + HFX_WAVE - 1) / HFX_WAVE;
        const int p = lo + lane * step;
        const bool keep = p < hi && !(sb * s.schw_ket[p] * pm < thr);
        const int c = __popcll(__ballot(keep));          // the kept probes are a prefix
        const int pc = lo + c * step;                    // first failing probe, if any
        const int nlo = c == 0 ? lo : lo + (c - 1) * step + 1;
        hi = pc < hi ? pc : hi;
        lo = nlo;
      }
      const int p = lo + lane;
      const bool keep = p < hi && !(sb * s.schw_ket[p] * pm < thr);
      hi = lo + __popcll(__ballot(keep));
    }
    const int jcut = hi;
    const int4 b = s.binfo[i];
    const int gb = s.bra[i];
    unsigned long long* mrow = s.masks + (size_t)i * s.nchunk * 2;
    int n64 = 0, n32 = 0;
    unsigned long long pq64 = 0, bp64 = 0, pq32 = 0, bp32 = 0;
    // a survivor's key: its ket's primitive-pair count + 1 (1..101); the
    // keys present (uniform), and each key's survivors counted in kw[0..1]
    unsigned long long klo = 0, khi = 0;
    for (int e = lane; e < 256; e += HFX_WAVE) kw[e >> 7][e & 127] = 0;
    wave_sync();
    for (int j0 = kb; j0 < jcut; j0 += HFX_WAVE) {
      const int j = j0 + lane;
      bool k64 = false, k32 = false;
      int key = 0;
      if (j < jcut) {
        const int4 k = s.kinfo[j];
        double mv = fmax(md[(size_t)b.x * ns + b.y], md[(size_t)k.x * ns + k.y]);
        if (!s.jonly) {
          const double ex = fmax(fmax(md[(size_t)b.x * ns + k.x], md[(size_t)b.x * ns + k.y]),
                                 fmax(md[(size_t)b.y * ns + k.x], md[(size_t)b.y * ns + k.y]));
          mv = fmax(mv, 0.25 * ex);
        }
        const double bound = sb * s.schw_ket[j] * mv;
        if (!(bound < thr)) {
          if (split > 0.0 && bound < split) k32 = true; else k64 = true;
          const unsigned long long nb_ = s.all_prims ? b.w : b.z;
          const unsigned long long nk_ = s.all_prims ? k.w : k.z;
          if (k64) { pq64 += nb_ * nk_; bp64 += nb_; } else { pq32 += nb_ * nk_; bp32 += nb_; }
          key = (int)nk_ + 1;
        }
      }
      const unsigned long long m64 = __ballot(k64), m32 = __ballot(k32);
      if (lane == 0) {
        mrow[((j0 - kb) / HFX_WAVE) * 2] = m64;
        mrow[((j0 - kb) / HFX_WAVE) * 2 + 1] = m32;
      }
      n64 += __popcll(m64);
      n32 += __popcll(m32);
      // the chunk's survivors counted key by key (each pass takes every lane
      // with the first remaining lane's key)
      for (unsigned long long rem = m64 | m32; rem;) {
        const int kk = __builtin_amdgcn_readlane(key, __ffsll((long long)rem) - 1);
        const unsigned long long same = __ballot(key == kk) & rem;
        if (lane == 0) {
          kw[0][kk] += __popcll(same & m64);
          kw[1][kk] += __popcll(same & m32);
        }
        if (kk < 64) klo |= 1ull << kk; else khi |= 1ull << (kk - 64);
        rem &= ~same;
      }
    }
    pq64 = wave_sum(pq64); bp64 = wave_sum(bp64); pq32 = wave_sum(pq32); bp32 = wave_sum(bp32);
    int base64 = 0, base32 = 0;
    if (lane == 0) {
      if (n64) {
        base64 = atomicAdd(s.count64, n64);
        atomicAdd(&s.stats64[0], (unsigned long long)n64);
        atomicAdd(&s.stats64[1], pq64);
        atomicAdd(&s.stats64[2], bp64);
        atomicAdd(&s.stats64[3], 1ull);   // one more bra run in the list
      }
      if (n32) {
        base32 = atomicAdd(s.count32, n32);
        atomicAdd(&s.stats32[0], (unsigned long long)n32);
        atomicAdd(&s.stats32[1], pq32);
        atomicAdd(&s.stats32[2], bp32);
        atomicAdd(&s.stats32[3], 1ull);
      }
    }
    base64 = __shfl(base64, 0);
    base32 = __shfl(base32, 0);
    // Survivors grouped by ket primitive count, largest first (ket order
    // inside a group): the class kernels loop over a quartet's primitive
    // pairs, and a wave then runs quartets of equal trip count together.
    // Each present key's first list position, keys in descending order ...
    if (lane == 0) {
      for (unsigned long long hl = khi, ll = klo; hl | ll;) {
        int kk;
        if (hl) { kk = 127 - __clzll((long long)hl); hl &= ~(1ull << (kk - 64)); }
        else { kk = 63 - __clzll((long long)ll); ll &= ~(1ull << kk); }
        kw[2][kk] = base64;
        kw[3][kk] = base32;
        base64 += kw[0][kk];
        base32 += kw[1][kk];
      }
    }
    wave_sync();   // positions and the masks lane 0 stored are read by every lane
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // ... then one scan of the bra's masks, each chunk's survivors placed key
    // by key (a chunk's keys are distinct passes, so its position reads
    // precede lane 0's updates of the same key)
    for (int j0 = kb; j0 < jcut; j0 += HFX_WAVE) {
      const unsigned long long m64 = mrow[((j0 - kb) / HFX_WAVE) * 2];
      const unsigned long long m32 = mrow[((j0 - kb) / HFX_WAVE) * 2 + 1];
      if ((m64 | m32) == 0ull) continue;
      const bool surv = (m64 | m32) >> lane & 1ull;
      int key = 0;
      int2 v = make_int2(0, 0);
      if (surv) {
        const int4 k = s.kinfo[j0 + lane];
        key = (s.all_prims ? k.w : k.z) + 1;
        v = make_int2(gb, s.ket[j0 + lane]);
      }
      for (unsigned long long rem = m64 | m32; rem;) {
        const int kk = __builtin_amdgcn_readlane(key, __ffsll((long long)rem) - 1);
        const unsigned long long same = __ballot(key == kk) & rem;
        const unsigned long long s64 = same & m64, s32 = same & m32;
        const int p64 = kw[2][kk], p32 = kw[3][kk];
        if (same >> lane & 1ull) {
          if (m64 >> lane & 1ull) s.list64[p64 + __popcll(s64 & below)] = v;
          else s.list32[p32 + __popcll(s32 & below)] = v;
        }
        wave_sync();   // every lane has read the positions before lane 0 moves them
        if (lane == 0) {
          kw[2][kk] = p64 + __popcll(s64);
          kw[3][kk] = p32 + __popcll(s32);
        }
        rem &= ~same;
      }
      wave_sync();   // the chunk's updates are seen by the next chunk's reads
    }
  }
}

// Diagonal-pair screening (fock_build_diag.f90:17-21): every pair of the type
// (not truncated by `last`), bound = schw^2 * max(maxdm(ab), maxdm(bb)/4, maxdm(aa)/4).
// With check == 0 every pair is taken (fock_build_initial_diag.f90:31-35).
__global__ void k_screen_diag(const DiagScreenArgs s) {
  const int nround = ((s.npair + HFX_WAVE - 1) / HFX_WAVE) * HFX_WAVE;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nround; i += gridDim.x * blockDim.x) {
    bool ok = i < s.npair;
    int g = 0;
    unsigned long long primq = 0, brap = 0;
    if (ok) {
      g = s.first + i;
      const PairRec b = s.pairs[g];
      if (s.check) {
        const int ns = s.nsh;
        const double* md = s.maxdm + b.mol * s.md_stride;
        const double mv = fmax(md[(size_t)b.lsa * ns + b.lsb],
                               fmax(0.25 * md[(size_t)b.lsb * ns + b.lsb], 0.25 * md[(size_t)b.lsa * ns + b.lsa]));
        const double sc = s.schw[i];
        if (sc * sc * mv < (s.thr_mol ? s.thr_mol[b.mol] : s.thr)) ok = false;
      }
      if (ok) {
        const int np_ = s.all_prims ? b.npall : b.np;
        primq = (unsigned long long)np_ * np_;
        brap = np_;
      }
    }
    wave_append(ok, make_int2(g, g), s.list, s.count, s.stats, primq, brap);
  }
}

// X += pair buffer: one thread per (pair, element); the pairs are distinct
// shell pairs, so every element of X is written by at most one thread.
__global__ void k_pairbuf_to_x(const double* __restrict__ pairbuf, const PairRec* __restrict__ pairs,
                               const ShellRec* __restrict__ shells, int npairs, double* __restrict__ X, int nb) {
  const long long n = (long long)npairs * HFX_PAIRBUF_STRIDE;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double v = pairbuf[i];
    if (v == 0.0) continue;
    const int p = (int)(i / HFX_PAIRBUF_STRIDE), e = (int)(i - (long long)p * HFX_PAIRBUF_STRIDE);
    const PairRec& pr = pairs[p];
    const int nbc = shells[pr.shb].ncomp;
    const int ia = e / nbc, ib = e - ia * nbc;
    if (ia < shells[pr.sha].ncomp) X[pr.mbase + (size_t)(pr.aoa + ia) * nb + pr.aob + ib] += v;
  }
}

// y = (float)x: the density for the FP32 quartets' contractions
__global__ void k_to_f32(const double* __restrict__ x, float* __restrict__ y, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

// Diagonal builds: integral elements with |e| <= 1e-10 are dropped
// (fock_build_diag.f90:22) between a chunk's combination kernels and its
// digestion; the block maxima were taken from the raw values.
__global__ void k_cut_small(double* __restrict__ e, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double v = e[i];
    if (v != 0.0 && fabs(v) <= HFX_INTTHRESH) e[i] = 0.0;
  }
}

// F += X + X^T per nb x nb matrix (the accumulator holds each symmetric
// contribution once); nmat stacked matrices.
__global__ void k_add_sym(const double* __restrict__ X, double* __restrict__ F, int nb, int nmat) {
  const long long n2 = (long long)nb * nb, n = n2 * nmat;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long m0 = idx / n2 * n2, r = idx - m0;   // matrix, element within it
    const int i = (int)(r / nb), j = (int)(r - (long long)i * nb);
    F[idx] += X[idx] + X[m0 + (size_t)j * nb + i];
  }
}

}  // namespace hfx
