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
// symmetric nsh x nsh matrix.  The block is the one evaluate_maxdm reads
// (evaluate_maxdm.f90:1-101): rows = AOs of the first shell of the pair in
// list order (s before p before d before L), columns = AOs of the second.
// For a (d, L) pair the reference reads rows ao_d..ao_d+3 and columns
// ao_L..ao_L+4 (evaluate_maxdm.f90:82-90): 4 of the 5 d functions and the AO
// after the L shell.  That block is reproduced so that screening decisions
// match the reference exactly; a column past nb is clamped.
__global__ void k_maxdm(const double* __restrict__ P, int nb, const ShellRec* __restrict__ sh, int nsh,
                        double* __restrict__ maxdm) {
  const long long n = (long long)nsh * nsh;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / nsh), j = (int)(idx - (long long)i * nsh);
    if (j > i) continue;
    ShellRec s1 = sh[i], s2 = sh[j];
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
        m = fmax(m, fabs(P[(size_t)ar * nb + ac]));
      }
    }
    maxdm[(size_t)i * nsh + j] = m;
    maxdm[(size_t)j * nsh + i] = m;
  }
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

// Screening of one class (fock_build_non_diag.f90:63-67, e.g. :58-67):
// bound = schw_bra * schw_ket * max(maxdm(ab), maxdm(cd),
//                                   1/4 max(maxdm(ac), maxdm(ad), maxdm(bc), maxdm(bd)))
// (J only: max(maxdm(ab), maxdm(cd)), fock_build_non_diag_j.f90:61-62);
// survivors have bound >= thr.  Same-type classes take unordered pairs of
// distinct pairs (ket position < bra position, fock_build_non_diag.f90:52-57).
// With mixed precision, survivors with bound < split go to the FP32 list.
__global__ void k_screen(const ScreenArgs s) {
  const long long n = (long long)s.nbra * s.nket;
  // Every lane of a wave must reach wave_append, so the loop bound is
  // rounded up to whole waves.
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long nround = ((n + HFX_WAVE - 1) / HFX_WAVE) * HFX_WAVE;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nround; idx += stride) {
    bool ok = idx < n;
    int i = 0, j = 0;
    if (ok) {
      i = (int)(idx / s.nket);
      j = (int)(idx - (long long)i * s.nket);
      if (s.same && j >= i) ok = false;
    }
    bool keep64 = false, keep32 = false;
    int gb = 0, gk = 0;
    unsigned long long primq = 0, brap = 0;
    if (ok) {
      gb = s.bra[i];
      gk = s.ket[j];
      const int4 b = s.binfo[i];
      const int4 k = s.kinfo[j];
      const int ns = s.nsh;
      const double* md = s.maxdm;
      double mv = fmax(md[(size_t)b.x * ns + b.y], md[(size_t)k.x * ns + k.y]);
      if (!s.jonly) {
        const double ex = fmax(fmax(md[(size_t)b.x * ns + k.x], md[(size_t)b.x * ns + k.y]),
                               fmax(md[(size_t)b.y * ns + k.x], md[(size_t)b.y * ns + k.y]));
        mv = fmax(mv, 0.25 * ex);
      }
      const double bound = s.schw_bra[i] * s.schw_ket[j] * mv;
      if (!(bound < s.thr)) {
        if (s.split > 0.0 && bound < s.split) keep32 = true; else keep64 = true;
        const int nb_ = s.all_prims ? b.w : b.z;
        const int nk_ = s.all_prims ? k.w : k.z;
        primq = (unsigned long long)nb_ * (unsigned long long)nk_;
        brap = (unsigned long long)nb_;
      }
    }
    const int2 v = make_int2(gb, gk);
    wave_append(keep64, v, s.list64, s.count64, s.stats64, primq, brap);
    if (s.split > 0.0) wave_append(keep32, v, s.list32, s.count32, s.stats32, primq, brap);
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
        const double mv = fmax(s.maxdm[(size_t)b.sha * ns + b.shb],
                               fmax(0.25 * s.maxdm[(size_t)b.shb * ns + b.shb],
                                    0.25 * s.maxdm[(size_t)b.sha * ns + b.sha]));
        const double sc = s.schw[i];
        if (sc * sc * mv < s.thr) ok = false;
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

// Digestion of integral blocks produced by the block class kernels: one wave
// per quartet.  The block is staged in LDS, its max |e| decides the block test
// (fock_build_non_diag.f90:69: skip when max|e| < thr), and the six J/K
// contractions of the symmetric-case digestion (fock_build_non_diag.f90:
// 2871-2955 in closed form) are spread over the lanes:
//   X_ab += 2w sum_cd e P_cd,  X_cd += 2w sum_ab e P_ab,
//   X_ac -= w/2 sum_bd e P_bd, X_ad -= w/2 sum_bc e P_bc,
//   X_bc -= w/2 sum_ad e P_ad, X_bd -= w/2 sum_ac e P_ac,
// w = 1/2 per pair made of one shell twice (1/2 more for bra == ket).
// In the diagonal builds elements with |e| <= 1e-10 are dropped first
// (fock_build_diag.f90:22).
#define HFX_DIG_WAVES 4
__global__ void __launch_bounds__(64 * HFX_DIG_WAVES) k_digest(const DigestArgs d) {
  __shared__ double s_e[HFX_DIG_WAVES][625];
  __shared__ double s_p[HFX_DIG_WAVES][150];
  const int lane = threadIdx.x % HFX_WAVE;
  const int wv = threadIdx.x / HFX_WAVE;
  const int nA = d.dims[0], nB = d.dims[1], nC = d.dims[2], nD = d.dims[3];
  const int ne = nA * nB * nC * nD;
  const int qend = min(d.q1, *d.count);
  double* se = s_e[wv];
  double* sp = s_p[wv];
  const int nwaves = gridDim.x * (blockDim.x / HFX_WAVE);
  for (int iq = d.q0 + blockIdx.x * (blockDim.x / HFX_WAVE) + wv; iq < qend; iq += nwaves) {
    const int2 bk = d.list[iq];
    const PairRec pb = d.pairs[bk.x];
    const PairRec pk = d.pairs[bk.y];
    const double* blk = d.eblk + (size_t)(iq - d.q0) * ne;
    double vmax = 0.0;
    for (int i = lane; i < ne; i += HFX_WAVE) {
      double v = blk[i];
      vmax = fmax(vmax, fabs(v));
      if (d.diag && fabs(v) <= HFX_INTTHRESH) v = 0.0;
      se[i] = v;
      if (d.dump) {
        const int id = i % nD, ic = (i / nD) % nC, ib = (i / (nD * nC)) % nB, ia = i / (nD * nC * nB);
        const int i1 = d.swap_bra ? ib : ia, i2 = d.swap_bra ? ia : ib;
        const int i3 = d.swap_ket ? id : ic, i4 = d.swap_ket ? ic : id;
        const int n1 = d.swap_bra ? nB : nA, n2 = d.swap_bra ? nA : nB, n3 = d.swap_ket ? nD : nC;
        d.dump[(size_t)iq * 625 + i1 + n1 * (i2 + n2 * (i3 + n3 * i4))] = v;
      }
    }
    for (int o = HFX_WAVE / 2; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o));
    if (d.schw_out && lane == 0) d.schw_out[bk.x] = sqrt(vmax);
    if (!d.diag && vmax < d.thr) continue;
    const int shA = d.swap_bra ? pb.shb : pb.sha, shB = d.swap_bra ? pb.sha : pb.shb;
    const int shC = d.swap_ket ? pk.shb : pk.sha, shD = d.swap_ket ? pk.sha : pk.shb;
    const int aoA = d.swap_bra ? pb.aob : pb.aoa, aoB = d.swap_bra ? pb.aoa : pb.aob;
    const int aoC = d.swap_ket ? pk.aob : pk.aoa, aoD = d.swap_ket ? pk.aoa : pk.aob;
    // density blocks: ab, cd, ac, ad, bc, bd
    const int o_ab = 0, o_cd = o_ab + nA * nB, o_ac = o_cd + nC * nD, o_ad = o_ac + nA * nC;
    const int o_bc = o_ad + nA * nD, o_bd = o_bc + nB * nC, np = o_bd + nB * nD;
    for (int i = lane; i < np; i += HFX_WAVE) {
      int r, c;
      if (i < o_cd) { r = aoA + (i - o_ab) / nB; c = aoB + (i - o_ab) % nB; }
      else if (i < o_ac) { r = aoC + (i - o_cd) / nD; c = aoD + (i - o_cd) % nD; }
      else if (i < o_ad) { r = aoA + (i - o_ac) / nC; c = aoC + (i - o_ac) % nC; }
      else if (i < o_bc) { r = aoA + (i - o_ad) / nD; c = aoD + (i - o_ad) % nD; }
      else if (i < o_bd) { r = aoB + (i - o_bc) / nC; c = aoC + (i - o_bc) % nC; }
      else { r = aoB + (i - o_bd) / nD; c = aoD + (i - o_bd) % nD; }
      sp[i] = d.P[(size_t)r * d.nb + c];
    }
    wave_sync();
    const double w = ((shA == shB) ? 0.5 : 1.0) * ((shC == shD) ? 0.5 : 1.0) * (d.diag ? 0.5 : 1.0);
    const int nout = d.jonly ? o_ac : np;
    for (int o = lane; o < nout; o += HFX_WAVE) {
      double acc = 0.0;
      int row, col;
      double wt;
      if (o < o_cd) {          // J_ab
        const int a = (o - o_ab) / nB, b = (o - o_ab) % nB;
        for (int c = 0; c < nC; ++c)
          for (int e = 0; e < nD; ++e) acc += se[((a * nB + b) * nC + c) * nD + e] * sp[o_cd + c * nD + e];
        row = aoA + a; col = aoB + b; wt = 2.0 * w;
      } else if (o < o_ac) {   // J_cd
        const int c = (o - o_cd) / nD, e = (o - o_cd) % nD;
        for (int a = 0; a < nA; ++a)
          for (int b = 0; b < nB; ++b) acc += se[((a * nB + b) * nC + c) * nD + e] * sp[o_ab + a * nB + b];
        row = aoC + c; col = aoD + e; wt = 2.0 * w;
      } else if (o < o_ad) {   // K_ac
        const int a = (o - o_ac) / nC, c = (o - o_ac) % nC;
        for (int b = 0; b < nB; ++b)
          for (int e = 0; e < nD; ++e) acc += se[((a * nB + b) * nC + c) * nD + e] * sp[o_bd + b * nD + e];
        row = aoA + a; col = aoC + c; wt = -0.5 * w;
      } else if (o < o_bc) {   // K_ad
        const int a = (o - o_ad) / nD, e = (o - o_ad) % nD;
        for (int b = 0; b < nB; ++b)
          for (int c = 0; c < nC; ++c) acc += se[((a * nB + b) * nC + c) * nD + e] * sp[o_bc + b * nC + c];
        row = aoA + a; col = aoD + e; wt = -0.5 * w;
      } else if (o < o_bd) {   // K_bc
        const int b = (o - o_bc) / nC, c = (o - o_bc) % nC;
        for (int a = 0; a < nA; ++a)
          for (int e = 0; e < nD; ++e) acc += se[((a * nB + b) * nC + c) * nD + e] * sp[o_ad + a * nD + e];
        row = aoB + b; col = aoC + c; wt = -0.5 * w;
      } else {                 // K_bd
        const int b = (o - o_bd) / nD, e = (o - o_bd) % nD;
        for (int a = 0; a < nA; ++a)
          for (int c = 0; c < nC; ++c) acc += se[((a * nB + b) * nC + c) * nD + e] * sp[o_ac + a * nC + c];
        row = aoB + b; col = aoD + e; wt = -0.5 * w;
      }
      atomicAdd(&d.X[(size_t)row * d.nb + col], wt * acc);
    }
    wave_sync();
  }
}

// F += X + X^T (the accumulator holds each symmetric contribution once).
__global__ void k_add_sym(const double* __restrict__ X, double* __restrict__ F, int nb) {
  const long long n = (long long)nb * nb;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / nb), j = (int)(idx - (long long)i * nb);
    F[idx] += X[idx] + X[(size_t)j * nb + i];
  }
}

}  // namespace hfx
