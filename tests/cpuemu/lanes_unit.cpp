// TEST INFRASTRUCTURE: targeted checks of libhfx's cross-lane device code with
// 64-lane waves, on the CPU.  Built by tests/test_cpu_lanes.py with g++ against
// the lanes-mode HIP stand-in (-DHFX_EMU_LANES) and linked with
// tests/cpuemu/libhfx_lanes.so, whose k_screen and k_screen_diag are the
// library's own kernels compiled from hfcode_amd/csrc/hfx_kernels.hip.
//
//   lanes_unit all        every check below; prints one line per check, exit 0 if all pass
//   lanes_unit divergent  a ballot inside a branch taken by some lanes of a wave: the
//                         stand-in must stop with "wait at different collectives"
//
// Checks:
// * the stand-in's wave collectives (ballot masks, shuffles, block counts);
// * k_screen: on random pair lists (one and several molecules, same-type and
//   distinct-type classes, J only, mixed precision, Schwarz-sorted and
//   unsorted kets, lengths that are not multiples of 64, empty molecules,
//   an infinite threshold) the FP64 and FP32 lists hold exactly the
//   quartets a direct loop over every (bra, ket) finds with bound >= thr,
//   each bra's quartets contiguous and ordered by ket primitive count
//   (largest first) then ket position, counts and statistics equal;
// * k_screen_diag (ballot compaction through wave_append): the list holds
//   exactly the diagonal pairs a direct loop keeps, once each;
// * wave_sum (DPP row scans + readlane): every lane of a wave gets the sum of
//   the wave's 64 values, for each of four waves of a workgroup;
// * pick_resident: the pair chosen is one of its candidates (the current
//   pair, the step's first, middle and last quartets' bras) and none of them
//   owns more of the step's quartets; the current pair is kept on a tie.
#include "hip/hip_runtime.h"

#include <algorithm>
#include <cstdio>
#include <map>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "hfx_kernels.h"

using hfx::DiagScreenArgs;
using hfx::PairRec;
using hfx::ScreenArgs;

static int g_fail = 0;
#define CHECK(cond, ...)                        \
  do {                                          \
    if (!(cond)) {                              \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                 \
      std::printf("\n");                        \
      ++g_fail;                                 \
      return false;                             \
    }                                           \
  } while (0)

// ---- the stand-in's collectives ----------------------------------------------
__global__ void k_selfcheck(unsigned long long* out) {
  const int t = threadIdx.x, lane = __lane_id();
  const unsigned long long m = __ballot(t % 3 == 0);
  const int x = __shfl_xor(t, 5);
  const int y = __shfl(t * 2, 63 - lane);
  const int c = __syncthreads_count(t < 77);
  out[4 * t + 0] = m;
  out[4 * t + 1] = (unsigned long long)x;
  out[4 * t + 2] = (unsigned long long)y;
  out[4 * t + 3] = (unsigned long long)c;
}

static bool check_collectives() {
  std::vector<unsigned long long> out(4 * 256);
  hipLaunchKernelGGL(k_selfcheck, dim3(1), dim3(256), 0, 0, out.data());
  for (int t = 0; t < 256; ++t) {
    const int w0 = t & ~63, lane = t & 63;
    unsigned long long m = 0;
    for (int l = 0; l < 64; ++l)
      if ((w0 + l) % 3 == 0) m |= 1ull << l;
    CHECK(out[4 * t] == m, "ballot of lane %d", t);
    CHECK(out[4 * t + 1] == (unsigned long long)(w0 + (lane ^ 5)), "shfl_xor of lane %d", t);
    CHECK(out[4 * t + 2] == (unsigned long long)(2 * (w0 + 63 - lane)), "shfl of lane %d", t);
    CHECK(out[4 * t + 3] == 77, "syncthreads_count of lane %d", t);
  }
  return true;
}

// ---- wave_sum -----------------------------------------------------------------
__global__ void k_wave_sum(const double* in, double* out) { out[threadIdx.x] = hfx::wave_sum(in[threadIdx.x]); }

static bool check_wave_sum(unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (int rep = 0; rep < 4; ++rep) {
    std::vector<double> in(256), out(256, -7.0);
    for (int t = 0; t < 256; ++t)   // one wave of zeros but one lane, the others mixed magnitudes
      in[t] = rep == 1 && t < 64 ? (t == 37 ? 3.0 : 0.0) : u(rng) * std::pow(10.0, (double)(rng() % 7) - 3.0);
    hipLaunchKernelGGL(k_wave_sum, dim3(1), dim3(256), 0, 0, in.data(), out.data());
    for (int w = 0; w < 4; ++w) {
      double ref = 0.0, mag = 0.0;
      for (int l = 0; l < 64; ++l) {
        ref += in[64 * w + l];
        mag += std::fabs(in[64 * w + l]);
      }
      for (int l = 0; l < 64; ++l)
        CHECK(std::fabs(out[64 * w + l] - ref) <= 1e-14 * mag, "wave_sum rep %d wave %d lane %d: %.17g vs %.17g", rep,
              w, l, out[64 * w + l], ref);
    }
  }
  std::printf("ok wave_sum\n");
  return true;
}

// a ballot reached by the even lanes only: not a valid wave collective here
__global__ void k_divergent(int* out) {
  if (threadIdx.x % 2 == 0) out[threadIdx.x] = (int)__popcll(__ballot(true));
  out[threadIdx.x] += (int)__popcll(__ballot(threadIdx.x < 3));
}

// ---- k_screen ----------------------------------------------------------------
struct ScreenCase {
  int nmol, same, jonly, mixed, all_prims, sorted;
  std::vector<int> nbra_mol, nket_mol;   // per molecule (same: nket_mol = nbra_mol)
  bool inf_thr_mol1;
};

struct Quartet {
  int bra, ket, key;
  bool fp32;
};

static double log_uniform(std::mt19937_64& r, double lo_exp, double hi_exp) {
  std::uniform_real_distribution<double> u(lo_exp, hi_exp);
  return std::pow(10.0, u(r));
}

static bool run_screen_case(const ScreenCase& sc, unsigned seed, const char* name) {
  std::mt19937_64 rng(seed);
  const int nsh = 9;
  const int nmol = sc.nmol;
  // bra and ket lists: molecules one after another, as a batch context lays them out
  std::vector<int> bmol, kseg(nmol + 1, 0), bseg(nmol + 1, 0);
  for (int m = 0; m < nmol; ++m) {
    bseg[m + 1] = bseg[m] + sc.nbra_mol[m];
    kseg[m + 1] = kseg[m] + (sc.same ? sc.nbra_mol[m] : sc.nket_mol[m]);
    for (int i = 0; i < sc.nbra_mol[m]; ++i) bmol.push_back(m);
  }
  const int nbra = bseg[nmol], nket = kseg[nmol];
  auto make_side = [&](int n, const std::vector<int>& seg, std::vector<int4>& info, std::vector<double>& schw,
                       std::vector<int>& gidx, int gbase) {
    info.resize(n);
    schw.resize(n);
    gidx.resize(n);
    std::uniform_int_distribution<int> sh(0, nsh - 1), np(1, 9);
    for (int m = 0; m < nmol; ++m) {
      std::vector<double> v;
      for (int i = seg[m]; i < seg[m + 1]; ++i) v.push_back(log_uniform(rng, -7.0, 0.0));
      if (sc.sorted) std::sort(v.begin(), v.end(), std::greater<double>());
      for (int i = seg[m]; i < seg[m + 1]; ++i) {
        const int a = sh(rng), b = sh(rng), p = np(rng);
        info[i] = make_int4(std::max(a, b), std::min(a, b), p, p + (int)(rng() % 4));
        schw[i] = v[i - seg[m]];
        gidx[i] = gbase + 7 * i + 3;
      }
    }
  };
  std::vector<int4> binfo, kinfo;
  std::vector<double> sbra, sket;
  std::vector<int> gbra, gket;
  make_side(nbra, bseg, binfo, sbra, gbra, 1000);
  if (sc.same) {
    kinfo = binfo;
    sket = sbra;
    gket = gbra;
  } else {
    make_side(nket, kseg, kinfo, sket, gket, 500000);
  }
  // per molecule: symmetric nsh x nsh density maxima, their maximum, threshold and split
  std::vector<double> maxdm((size_t)nmol * nsh * nsh), thr(nmol), split(nmol);
  std::vector<unsigned long long> pmax(nmol);
  for (int m = 0; m < nmol; ++m) {
    double mx = 0.0;
    for (int i = 0; i < nsh; ++i)
      for (int j = 0; j <= i; ++j) {
        const double v = log_uniform(rng, -5.0, 0.3);
        maxdm[(size_t)m * nsh * nsh + i * nsh + j] = maxdm[(size_t)m * nsh * nsh + j * nsh + i] = v;
        mx = std::max(mx, v);
      }
    std::memcpy(&pmax[m], &mx, 8);
    thr[m] = log_uniform(rng, -9.0, -7.0);
    split[m] = sc.mixed ? log_uniform(rng, -6.0, -4.5) : 0.0;
  }
  if (sc.inf_thr_mol1 && nmol > 1) thr[1] = INFINITY;

  // direct loop over every (bra, ket) of the same molecule
  std::map<int, std::vector<Quartet>> want;   // per bra position, in the expected list order
  unsigned long long st64[3] = {0, 0, 0}, st32[3] = {0, 0, 0};
  int n64 = 0, n32 = 0;
  const int nsq = nsh * nsh;
  for (int i = 0; i < nbra; ++i) {
    const int m = bmol[i];
    const double* md = maxdm.data() + (size_t)m * nsq;
    const int jend = sc.same ? std::min(i, kseg[m + 1]) : kseg[m + 1];
    for (int j = kseg[m]; j < jend; ++j) {
      const int4 b = binfo[i], k = kinfo[j];
      double mv = std::max(md[b.x * nsh + b.y], md[k.x * nsh + k.y]);
      if (!sc.jonly)
        mv = std::max(mv, 0.25 * std::max(std::max(md[b.x * nsh + k.x], md[b.x * nsh + k.y]),
                                          std::max(md[b.y * nsh + k.x], md[b.y * nsh + k.y])));
      const double bound = sbra[i] * sket[j] * mv;
      if (bound < thr[m]) continue;
      const bool f32 = sc.mixed && bound < split[m];
      const unsigned long long nb_ = sc.all_prims ? b.w : b.z, nk_ = sc.all_prims ? k.w : k.z;
      unsigned long long* st = f32 ? st32 : st64;
      st[0] += 1;
      st[1] += nb_ * nk_;
      st[2] += nb_;
      (f32 ? n32 : n64) += 1;
      want[i].push_back(Quartet{i, j, (int)nk_, f32});
    }
    auto& v = want[i];
    std::stable_sort(v.begin(), v.end(), [](const Quartet& x, const Quartet& y) { return x.key > y.key; });
  }

  int maxseg = 0;
  for (int m = 0; m < nmol; ++m) maxseg = std::max(maxseg, kseg[m + 1] - kseg[m]);
  const int nchunk = std::max(1, (maxseg + 63) / 64);
  const int cap = nbra * std::max(1, maxseg) + 1;
  std::vector<int2> list64(cap, make_int2(-1, -1)), list32(cap, make_int2(-1, -1));
  std::vector<unsigned long long> masks((size_t)nbra * nchunk * 2 + 2, 0ull), stats(8, 0ull);
  std::vector<int> counts(2, 0);
  ScreenArgs s;
  std::memset(&s, 0, sizeof(s));
  s.bra = gbra.data();
  s.ket = gket.data();
  s.binfo = binfo.data();
  s.kinfo = kinfo.data();
  s.schw_bra = sbra.data();
  s.schw_ket = sket.data();
  s.maxdm = maxdm.data();
  s.md_stride = nsq;
  s.bmol = nmol > 1 ? bmol.data() : nullptr;
  s.kseg = nmol > 1 ? kseg.data() : nullptr;
  s.thr_mol = thr.data();
  s.split_mol = sc.mixed ? split.data() : nullptr;
  s.list64 = list64.data();
  s.list32 = sc.mixed ? list32.data() : nullptr;
  s.count64 = &counts[0];
  s.count32 = &counts[1];
  s.stats64 = &stats[0];
  s.stats32 = &stats[4];
  s.pmax = pmax.data();
  s.masks = masks.data();
  s.thr = thr[0];
  s.split = sc.mixed ? split[0] : 0.0;
  s.nbra = nbra;
  s.nket = nket;
  s.same = sc.same;
  s.jonly = sc.jonly;
  s.nsh = nsh;
  s.all_prims = sc.all_prims;
  s.nchunk = nchunk;
  s.ket_sorted = sc.sorted;
  hipLaunchKernelGGL(hfx::k_screen, dim3(1), dim3(256), 0, 0, s);

  CHECK(counts[0] == n64 && counts[1] == n32, "%s: counts %d %d, want %d %d", name, counts[0], counts[1], n64, n32);
  for (int k = 0; k < 3; ++k)
    CHECK(stats[k] == st64[k] && stats[4 + k] == st32[k], "%s: statistics entry %d", name, k);
  // bra runs: bras with at least one survivor in the list
  unsigned long long runs64 = 0, runs32 = 0;
  for (int i = 0; i < nbra; ++i) {
    bool h64 = false, h32 = false;
    for (const Quartet& q : want[i]) (q.fp32 ? h32 : h64) = true;
    runs64 += h64;
    runs32 += h32;
  }
  CHECK(stats[3] == runs64 && stats[7] == runs32, "%s: bra runs %llu %llu, want %llu %llu", name, stats[3], stats[7],
        runs64, runs32);
  // every bra's quartets: one contiguous run per list, in the expected order
  for (int prec = 0; prec < 2; ++prec) {
    const std::vector<int2>& L = prec ? list32 : list64;
    const int n = prec ? n32 : n64;
    std::map<int, std::vector<int>> got;   // bra position -> list positions
    std::map<int, int> bra_of;
    for (int i = 0; i < nbra; ++i) bra_of[gbra[i]] = i;
    std::map<int, int> ket_of;
    for (int j = 0; j < nket; ++j) ket_of[gket[j]] = j;
    for (int p = 0; p < n; ++p) {
      CHECK(bra_of.count(L[p].x) && ket_of.count(L[p].y), "%s: list entry %d (%d, %d) unknown", name, p, L[p].x,
            L[p].y);
      got[bra_of[L[p].x]].push_back(p);
    }
    CHECK(L[n].x == -1, "%s: entry past the count written", name);
    for (int i = 0; i < nbra; ++i) {
      std::vector<int> exp;
      for (const Quartet& q : want[i])
        if (q.fp32 == (prec == 1)) exp.push_back(gket[q.ket]);
      const std::vector<int>& pos = got[i];
      CHECK(pos.size() == exp.size(), "%s: bra %d has %zu quartets in list %d, want %zu", name, i, pos.size(), prec,
            exp.size());
      for (size_t k = 0; k < pos.size(); ++k) {
        CHECK(pos[k] == pos[0] + (int)k, "%s: bra %d not contiguous in list %d", name, i, prec);
        CHECK(L[pos[k]].y == exp[k], "%s: bra %d, quartet %zu of list %d: ket %d, want %d", name, i, k, prec,
              L[pos[k]].y, exp[k]);
      }
    }
  }
  std::printf("ok k_screen %s: %d bras, %d kets, %d + %d quartets\n", name, nbra, nket, n64, n32);
  return true;
}

// ---- k_screen_diag -----------------------------------------------------------
static bool run_diag_case(int npair, int check, int nmol, unsigned seed) {
  std::mt19937_64 rng(seed);
  const int nsh = 7, first = 11;
  std::vector<PairRec> pairs(first + npair);
  std::vector<double> schw(npair), maxdm((size_t)nmol * nsh * nsh), thr(nmol);
  for (auto& v : maxdm) v = log_uniform(rng, -5.0, 0.0);
  for (int m = 0; m < nmol; ++m) thr[m] = log_uniform(rng, -9.0, -6.0);
  for (int i = 0; i < npair; ++i) {
    PairRec& p = pairs[first + i];
    std::memset(&p, 0, sizeof(p));
    p.mol = (int)(rng() % nmol);
    p.lsa = (int)(rng() % nsh);
    p.lsb = (int)(rng() % nsh);
    p.np = 1 + (int)(rng() % 9);
    p.npall = p.np + (int)(rng() % 3);
    schw[i] = log_uniform(rng, -6.0, 0.0);
  }
  std::vector<int> want;
  unsigned long long st[3] = {0, 0, 0};
  for (int i = 0; i < npair; ++i) {
    const PairRec& b = pairs[first + i];
    const double* md = maxdm.data() + (size_t)b.mol * nsh * nsh;
    const double mv = std::max(md[b.lsa * nsh + b.lsb],
                               std::max(0.25 * md[b.lsb * nsh + b.lsb], 0.25 * md[b.lsa * nsh + b.lsa]));
    if (check && schw[i] * schw[i] * mv < thr[b.mol]) continue;
    want.push_back(first + i);
    st[0] += 1;
    st[1] += (unsigned long long)b.np * b.np;
    st[2] += b.np;
  }
  std::vector<int2> list(npair + 64, make_int2(-1, -1));
  std::vector<unsigned long long> stats(3, 0);
  int count = 0;
  DiagScreenArgs s;
  std::memset(&s, 0, sizeof(s));
  s.pairs = pairs.data();
  s.schw = schw.data();
  s.maxdm = maxdm.data();
  s.md_stride = nsh * nsh;
  s.thr_mol = thr.data();
  s.list = list.data();
  s.count = &count;
  s.stats = stats.data();
  s.thr = thr[0];
  s.first = first;
  s.npair = npair;
  s.nsh = nsh;
  s.check = check;
  s.all_prims = 0;
  hipLaunchKernelGGL(hfx::k_screen_diag, dim3(1), dim3(256), 0, 0, s);
  CHECK(count == (int)want.size(), "k_screen_diag: count %d, want %zu", count, want.size());
  std::vector<int> got;
  for (int p = 0; p < count; ++p) {
    CHECK(list[p].x == list[p].y, "k_screen_diag: entry %d is not a diagonal pair", p);
    got.push_back(list[p].x);
  }
  std::sort(got.begin(), got.end());
  CHECK(got == want, "k_screen_diag: the listed pairs differ from the direct loop's");
  for (int k = 0; k < 3; ++k) CHECK(stats[k] == st[k], "k_screen_diag: statistics entry %d", k);
  std::printf("ok k_screen_diag: %d pairs, check %d, %d molecules, %d kept\n", npair, check, nmol, count);
  return true;
}

// ---- pick_resident -----------------------------------------------------------
struct PickCase {
  int lo, hi, cur;
};

// out[c * blockDim.x + lane]: every lane's choice for step c
__global__ void k_pick(const int2* list, const PickCase* cases, int ncase, int* out) {
  for (int c = 0; c < ncase; ++c)
    out[c * blockDim.x + threadIdx.x] = hfx::pick_resident(list, cases[c].lo, cases[c].hi, cases[c].cur);
}

static bool check_pick_resident(unsigned seed) {
  std::mt19937_64 rng(seed);
  const int n = 4096;
  // bras in runs of random length, as in a bra-contiguous quartet list
  std::vector<int2> list(n);
  int b = 0;
  for (int i = 0; i < n;) {
    const int run = 1 + (int)(rng() % 300);
    for (int k = 0; k < run && i < n; ++k, ++i) list[i] = make_int2(b, (int)(rng() % 1000));
    b += 1 + (int)(rng() % 3);
  }
  std::vector<PickCase> cases;
  for (int c = 0; c < 60; ++c) {
    const int lo = (int)(rng() % (n - 256));
    const int len = c % 3 == 0 ? 256 : 1 + (int)(rng() % 256);   // full and partial steps
    const int kind = c % 4;
    const int cur = kind == 0 ? -1 : kind == 1 ? list[lo].x : kind == 2 ? list[lo + len - 1].x : 100000;
    cases.push_back(PickCase{lo, lo + len, cur});
  }
  std::vector<int> out(cases.size() * 256, -7);
  hipLaunchKernelGGL(k_pick, dim3(1), dim3(256), 0, 0, list.data(), cases.data(), (int)cases.size(), out.data());
  for (size_t c = 0; c < cases.size(); ++c) {
    const PickCase& pc = cases[c];
    auto owned = [&](int bra) {
      int k = 0;
      for (int i = pc.lo; i < pc.hi; ++i) k += list[i].x == bra;
      return k;
    };
    std::vector<int> cand = {list[pc.lo].x, list[pc.lo + (pc.hi - pc.lo - 1) / 2].x, list[pc.hi - 1].x};
    if (pc.cur >= 0) cand.push_back(pc.cur);
    const int r = out[c * 256];
    for (int t = 1; t < 256; ++t)   // the kernels branch on the choice: every lane must agree
      CHECK(out[c * 256 + t] == r, "pick_resident case %zu: lane %d chose %d, lane 0 %d", c, t, out[c * 256 + t], r);
    CHECK(std::find(cand.begin(), cand.end(), r) != cand.end(), "pick_resident case %zu: %d is no candidate", c, r);
    for (int x : cand) CHECK(owned(r) >= owned(x), "pick_resident case %zu: %d owns more than %d", c, x, r);
    if (pc.cur >= 0 && owned(pc.cur) == owned(r))
      CHECK(r == pc.cur, "pick_resident case %zu: tie not kept on the current pair", c);
  }
  std::printf("ok pick_resident: %zu steps\n", cases.size());
  return true;
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "all";
  if (mode == "divergent") {
    std::vector<int> out(64, 0);
    hipLaunchKernelGGL(k_divergent, dim3(1), dim3(64), 0, 0, out.data());
    std::printf("divergent ballot was not detected\n");
    return 0;
  }
  if (check_collectives()) std::printf("ok collectives\n");
  const std::vector<std::pair<std::string, ScreenCase>> cases = {
      {"one molecule, distinct types, sorted", {1, 0, 0, 0, 0, 1, {150}, {333}, false}},
      {"one molecule, distinct types, unsorted", {1, 0, 0, 0, 1, 0, {97}, {200}, false}},
      {"one molecule, same type", {1, 1, 0, 0, 0, 1, {300}, {}, false}},
      {"one molecule, J only, mixed", {1, 0, 1, 1, 0, 1, {120}, {257}, false}},
      {"one molecule, same type, mixed, all primitives", {1, 1, 0, 1, 1, 1, {200}, {}, false}},
      {"short ket list", {1, 0, 0, 1, 0, 1, {70}, {5}, false}},
      {"four molecules, distinct types, mixed", {4, 0, 0, 1, 0, 1, {40, 0, 77, 65}, {130, 12, 0, 64}, false}},
      {"four molecules, same type, mixed", {4, 1, 0, 1, 0, 1, {64, 129, 1, 90}, {}, false}},
      {"three molecules, one at infinite threshold", {3, 0, 0, 1, 0, 1, {50, 60, 70}, {100, 150, 65}, true}},
      {"three molecules, unsorted, J only", {3, 0, 1, 0, 0, 0, {33, 44, 55}, {66, 77, 88}, false}},
  };
  unsigned seed = 17;
  for (auto& [name, sc] : cases) run_screen_case(sc, seed++, name.c_str());
  run_diag_case(1000, 1, 1, 5);
  run_diag_case(777, 1, 3, 6);
  run_diag_case(130, 0, 2, 7);
  run_diag_case(0, 1, 1, 8);
  check_pick_resident(9);
  check_wave_sum(10);
  std::printf(g_fail ? "%d checks FAILED\n" : "all checks passed\n", g_fail);
  return g_fail ? 1 : 0;
}
