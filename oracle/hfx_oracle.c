/* ORACLE-SIDE TEST INFRASTRUCTURE -- the checker, never the product.
 *
 * CPU restatement (plain C, FP64, single thread) of the reference's direct-SCF
 * Fock build (brichjnh/hfcode), used by tests/ and by the smoke check to
 * check libhfx.  Each function cites the reference lines it restates.  The
 * two-electron integrals themselves are computed by McMurchie--Davidson
 * (Hermite expansion), an algorithm independent of both the reference's
 * Obara--Saika class routines and the GPU's generated HGP kernels, so an
 * error in either recursion shows up here.
 *
 * Built by oracle/Makefile into oracle/_ref/libhfx_oracle.so.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/hfx.h"

#define MAXP 16
#define PI 3.141592653589793238462643383279502884197

/* ------------------------------------------------------------------------ */
/* Boys function exactly as the reference evaluates it.                      */
/* Grid: eval_boys_func_on_grid (index_functions.f90:195-226) with the       */
/* Gamma-ratio series of buildgammarats (index_functions.f90:171-193);       */
/* evaluation: boys_0/1/2/n (boys.f90:7-158).                                */
/* ------------------------------------------------------------------------ */
#define NGAMMA 10000
#define MAXBOYS 9
#define MAXGAMMA 50
#define TWODELTA 1.0e-3
static double g_boysvals[NGAMMA + 1][MAXBOYS + 1][4];
static int g_boys_ready = 0;

static void boys_grid(void) {
  double gr[MAXGAMMA + 1][MAXBOYS + 4];
  for (int i = 0; i <= MAXGAMMA; ++i)
    for (int j = 0; j <= MAXBOYS + 3; ++j) gr[i][j] = tgamma(j + 0.5) / tgamma(i + j + 1.5);
  for (int i = 0; i <= NGAMMA; ++i) {
    const double x = (i + 0.5) * TWODELTA;
    double bf[MAXBOYS + 4] = {0};
    double xpow = 1.0;
    const double efac = exp(-x) / 2.0;
    for (int jn = 0; jn <= MAXGAMMA; ++jn) {
      for (int j = 0; j <= MAXBOYS + 3; ++j) bf[j] += xpow * gr[jn][j];
      xpow *= x;
    }
    for (int j = 0; j <= MAXBOYS + 3; ++j) bf[j] *= efac;
    for (int j = 0; j <= MAXBOYS; ++j) {
      g_boysvals[i][j][0] = bf[j] + x * bf[j + 1] + .5 * x * x * bf[j + 2] + x * x * x * bf[j + 3] / 6.0;
      g_boysvals[i][j][1] = bf[j + 1] + x * bf[j + 2] + .5 * x * x * bf[j + 3];
      g_boysvals[i][j][2] = .5 * bf[j + 2] + .5 * x * bf[j + 3];
      g_boysvals[i][j][3] = bf[j + 3] / 6.0;
    }
  }
  g_boys_ready = 1;
}

static double grid_eval(int j, int m, double x) {
  const double* c = g_boysvals[j][m];
  return c[0] - x * (c[1] - x * (c[2] - x * c[3]));
}

/* boys.f90:7-158: medrange 2 / 4 / 8 / 8 for n = 0 / 1 / 2 / >2 */
void oracle_boys(int n, double x, double* fn) {
  if (!g_boys_ready) boys_grid();
  const double hsqrtpi = sqrt(PI) / 2.0;
  const double medrange = n == 0 ? 2.0 : n == 1 ? 4.0 : 8.0;
  if (x < 1.0e-15) {
    for (int i = 0; i <= n; ++i) fn[i] = 1.0 / (2 * i + 1);
  } else if (x > 28.0) {
    fn[0] = hsqrtpi / sqrt(x);
    for (int i = 1; i <= n; ++i) fn[i] = fn[i - 1] * (i - 0.5) / x;
  } else if (x >= medrange) {
    const double sx = sqrt(x);
    fn[0] = hsqrtpi * erf(sx) / sx;
    const double efac = .5 * exp(-x);
    for (int i = 0; i < n; ++i) fn[i + 1] = ((i + 0.5) * fn[i] - efac) / x;
  } else {
    const int j = (int)(x / TWODELTA);
    if (n <= 1) {
      for (int i = 0; i <= n; ++i) fn[i] = grid_eval(j, i, x);
    } else {
      const double efac = .5 * exp(-x);
      fn[n] = grid_eval(j, n, x);
      for (int i = n; i >= 1; --i) fn[i - 1] = (x * fn[i] + efac) * (2.0 / (2 * (i - 1) + 1));
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Shells, primitive pairs (precalc_kabs.f90:94-115)                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int type, nprim, ao, ncomp;
  double x[3];
  double z[MAXP], c[MAXP], cl[MAXP];
} osh;

static int ncomp_of(int t) { return t == 1 ? 1 : t == 2 ? 3 : t == 3 ? 5 : 4; }

static int load_shells(const hfx_mol_desc* d, osh* s) {
  for (int i = 0; i < d->nsh; ++i) {
    s[i].type = d->shtyp[i];
    s[i].nprim = d->nzet[i];
    s[i].ao = d->aosh[i] - 1;
    s[i].ncomp = ncomp_of(s[i].type);
    if (s[i].nprim > MAXP) return -1;
    for (int k = 0; k < 3; ++k) s[i].x[k] = d->xsh[3 * i + k];
    for (int k = 0; k < s[i].nprim; ++k) {
      s[i].z[k] = d->zeta[(size_t)d->maxprim * i + k];
      s[i].c[k] = d->djk[(size_t)d->maxprim * i + k];
      s[i].cl[k] = (s[i].type == 12 && d->djkl) ? d->djkl[(size_t)d->maxprim * i + k] : 1.0;
    }
  }
  return 0;
}

/* negab: |K_ab| / sqrt(p) < 1e-12 (precalc_kabs.f90:13, 112-114) */
static int negab(const osh* A, int ia, const osh* B, int ib) {
  const double a = A->z[ia], b = B->z[ib], p = a + b, pm1 = 1.0 / p;
  double r2 = 0;
  for (int k = 0; k < 3; ++k) r2 += (A->x[k] - B->x[k]) * (A->x[k] - B->x[k]);
  const double rt2pi54 = sqrt(2.0 * PI * PI * sqrt(PI));
  const double kab = rt2pi54 * exp(-a * b * pm1 * r2) * pm1 * A->c[ia] * B->c[ib];
  return fabs(kab) * sqrt(pm1) < 1.0e-12;
}

/* ------------------------------------------------------------------------ */
/* McMurchie--Davidson primitive ERIs over Cartesian components              */
/* ------------------------------------------------------------------------ */
static void cart_list(int l, int out[][3]) { /* reference order, index_functions.f90:57-68 */
  int n = 0;
  for (int i = 0; i <= l; ++i)
    for (int k = 0; k <= i; ++k) {
      out[n][0] = l - i;
      out[n][1] = i - k;
      out[n][2] = k;
      ++n;
    }
}

/* Hermite coefficients E^{ij}_t, i,j <= 2, t <= 4, for one dimension */
static void hermite_E(double a, double b, double XA, double XB, double E[3][3][5]) {
  const double p = a + b, q = a * b / p, XAB = XA - XB;
  const double XPA = -b / p * XAB, XPB = a / p * XAB;
  memset(E, 0, sizeof(double) * 45);
  E[0][0][0] = exp(-q * XAB * XAB);
  for (int i = 0; i <= 2; ++i)
    for (int j = 0; j <= 2; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
          v = XPA * E[i - 1][j][t] + (t + 1 <= 4 ? (t + 1) * E[i - 1][j][t + 1] : 0.0);
          if (t > 0) v += E[i - 1][j][t - 1] / (2 * p);
        } else {
          v = XPB * E[i][j - 1][t] + (t + 1 <= 4 ? (t + 1) * E[i][j - 1][t + 1] : 0.0);
          if (t > 0) v += E[i][j - 1][t - 1] / (2 * p);
        }
        E[i][j][t] = v;
      }
    }
}

/* R_{tuv}^{(0)}(alpha, PQ), t+u+v <= L, via the auxiliary recursion */
static double Rtab[9][9][9][9];
static void hermite_R(int L, double alpha, const double* PQ, const double* F) {
  for (int n = 0; n <= L; ++n) Rtab[n][0][0][0] = pow(-2.0 * alpha, n) * F[n];
  for (int n = L - 1; n >= 0; --n) {
    for (int t = 0; t <= L - n; ++t)
      for (int u = 0; u <= L - n - t; ++u)
        for (int v = 0; v <= L - n - t - u; ++v) {
          if (t + u + v == 0) continue;
          double r;
          if (t > 0) {
            r = PQ[0] * Rtab[n + 1][t - 1][u][v];
            if (t > 1) r += (t - 1) * Rtab[n + 1][t - 2][u][v];
          } else if (u > 0) {
            r = PQ[1] * Rtab[n + 1][t][u - 1][v];
            if (u > 1) r += (u - 1) * Rtab[n + 1][t][u - 2][v];
          } else {
            r = PQ[2] * Rtab[n + 1][t][u][v - 1];
            if (v > 1) r += (v - 1) * Rtab[n + 1][t][u][v - 2];
          }
          Rtab[n][t][u][v] = r;
        }
  }
}

/* spherical d from Cartesian (xx, xy, xz, yy, yz, zz): matharmond,
 * index_functions.f90:147-168 */
static const double kSph[5][6] = {{-0.5, 0, 0, -0.5, 0, 1},
                                  {0, 0, 1.7320508075688772, 0, 0, 0},
                                  {0, 0, 0, 0, 1.7320508075688772, 0},
                                  {0, 1.7320508075688772, 0, 0, 0, 0},
                                  {0.8660254037844386, 0, 0, -0.8660254037844386, 0, 0}};

/* AO component -> (l, Cartesian coefficient vector over cart(l), weighted-by-L-ratio) */
typedef struct {
  int l, weighted, ncart;
  double coef[6];
} comp_t;

static void comp_of(int type, int k, comp_t* c) {
  memset(c, 0, sizeof(*c));
  if (type == 1) { c->l = 0; c->ncart = 1; c->coef[0] = 1; }
  else if (type == 2) { c->l = 1; c->ncart = 3; c->coef[k] = 1; }
  else if (type == 3) { c->l = 2; c->ncart = 6; for (int u = 0; u < 6; ++u) c->coef[u] = kSph[k][u]; }
  else if (k == 0) { c->l = 0; c->ncart = 1; c->coef[0] = 1; }
  else { c->l = 1; c->ncart = 3; c->weighted = 1; c->coef[k - 1] = 1; }
}

/* Contracted block (ab|cd) over the AO components of four shells, AO order,
 * block[((ia*nb+ib)*nc+ic)*nd+id].  Primitive pairs flagged negab are skipped
 * except for the (sl|sl) class (i2e_slsl, s_d_l_terms.f90:1629-1693, has no
 * negab test). */
static void eri_block(const osh* S, int s1, int s2, int s3, int s4, int skip_negab, double* block) {
  const osh *A = &S[s1], *B = &S[s2], *C = &S[s3], *D = &S[s4];
  const int na = A->ncomp, nb = B->ncomp, nc = C->ncomp, nd = D->ncomp;
  memset(block, 0, sizeof(double) * na * nb * nc * nd);
  comp_t ca[5], cb[5], cc[5], cd[5];
  for (int i = 0; i < na; ++i) comp_of(A->type, i, &ca[i]);
  for (int i = 0; i < nb; ++i) comp_of(B->type, i, &cb[i]);
  for (int i = 0; i < nc; ++i) comp_of(C->type, i, &cc[i]);
  for (int i = 0; i < nd; ++i) comp_of(D->type, i, &cd[i]);
  int cartl[3][6][3];
  for (int l = 0; l <= 2; ++l) cart_list(l, cartl[l]);
  const int lmaxA = A->type == 3 ? 2 : (A->type == 1 ? 0 : 1), lmaxB = B->type == 3 ? 2 : (B->type == 1 ? 0 : 1);
  const int lmaxC = C->type == 3 ? 2 : (C->type == 1 ? 0 : 1), lmaxD = D->type == 3 ? 2 : (D->type == 1 ? 0 : 1);
  const int L = lmaxA + lmaxB + lmaxC + lmaxD;
  for (int pa = 0; pa < A->nprim; ++pa)
    for (int pb = 0; pb < B->nprim; ++pb) {
      if (skip_negab && negab(A, pa, B, pb)) continue;
      const double a = A->z[pa], b = B->z[pb], p = a + b;
      double P[3], Eab[3][3][3][5];
      for (int k = 0; k < 3; ++k) {
        P[k] = (a * A->x[k] + b * B->x[k]) / p;
        hermite_E(a, b, A->x[k], B->x[k], Eab[k]);
      }
      for (int pc = 0; pc < C->nprim; ++pc)
        for (int pd = 0; pd < D->nprim; ++pd) {
          if (skip_negab && negab(C, pc, D, pd)) continue;
          const double c = C->z[pc], d = D->z[pd], q = c + d;
          double Q[3], Ecd[3][3][3][5], PQ[3], F[9];
          for (int k = 0; k < 3; ++k) {
            Q[k] = (c * C->x[k] + d * D->x[k]) / q;
            hermite_E(c, d, C->x[k], D->x[k], Ecd[k]);
            PQ[k] = P[k] - Q[k];
          }
          const double alpha = p * q / (p + q);
          const double T = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
          oracle_boys(L, T, F);
          hermite_R(L, alpha, PQ, F);
          const double pref = 2.0 * pow(PI, 2.5) / (p * q * sqrt(p + q));
          for (int ia = 0; ia < na; ++ia)
            for (int ib = 0; ib < nb; ++ib)
              for (int ic = 0; ic < nc; ++ic)
                for (int id = 0; id < nd; ++id) {
                  double w = A->c[pa] * B->c[pb] * C->c[pc] * D->c[pd];
                  if (ca[ia].weighted) w *= A->cl[pa];
                  if (cb[ib].weighted) w *= B->cl[pb];
                  if (cc[ic].weighted) w *= C->cl[pc];
                  if (cd[id].weighted) w *= D->cl[pd];
                  double sum = 0.0;
                  for (int xa = 0; xa < ca[ia].ncart; ++xa) {
                    if (ca[ia].coef[xa] == 0) continue;
                    const int* ea = cartl[ca[ia].l][xa];
                    for (int xb = 0; xb < cb[ib].ncart; ++xb) {
                      if (cb[ib].coef[xb] == 0) continue;
                      const int* eb = cartl[cb[ib].l][xb];
                      for (int xc = 0; xc < cc[ic].ncart; ++xc) {
                        if (cc[ic].coef[xc] == 0) continue;
                        const int* ec = cartl[cc[ic].l][xc];
                        for (int xd = 0; xd < cd[id].ncart; ++xd) {
                          if (cd[id].coef[xd] == 0) continue;
                          const int* ed = cartl[cd[id].l][xd];
                          double v = 0.0;
                          for (int t = 0; t <= ea[0] + eb[0]; ++t)
                            for (int u = 0; u <= ea[1] + eb[1]; ++u)
                              for (int s = 0; s <= ea[2] + eb[2]; ++s) {
                                const double e1 = Eab[0][ea[0]][eb[0]][t] * Eab[1][ea[1]][eb[1]][u] *
                                                  Eab[2][ea[2]][eb[2]][s];
                                for (int t2 = 0; t2 <= ec[0] + ed[0]; ++t2)
                                  for (int u2 = 0; u2 <= ec[1] + ed[1]; ++u2)
                                    for (int s2 = 0; s2 <= ec[2] + ed[2]; ++s2) {
                                      const double e2 = Ecd[0][ec[0]][ed[0]][t2] * Ecd[1][ec[1]][ed[1]][u2] *
                                                        Ecd[2][ec[2]][ed[2]][s2];
                                      const double sg = ((t2 + u2 + s2) & 1) ? -1.0 : 1.0;
                                      v += e1 * e2 * sg * Rtab[0][t + t2][u + u2][s + s2];
                                    }
                              }
                          sum += ca[ia].coef[xa] * cb[ib].coef[xb] * cc[ic].coef[xc] * cd[id].coef[xd] * v;
                        }
                      }
                    }
                  }
                  block[((ia * nb + ib) * nc + ic) * nd + id] += w * pref * sum;
                }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Density maxima per shell pair (evaluate_maxdm.f90:1-101)                  */
/* ------------------------------------------------------------------------ */
static int letter(int t) { return t == 1 ? 0 : t == 2 ? 1 : t == 3 ? 2 : 3; }

static double maxdm_pair(const osh* S, int i, int j, const double* P, int nbf) {
  const osh *s1 = &S[i], *s2 = &S[j];
  if (letter(s2->type) < letter(s1->type)) { const osh* t = s1; s1 = s2; s2 = t; }
  int nr = s1->ncomp, nc = s2->ncomp;
  if (s1->type == 3 && s2->type == 12) { nr = 4; nc = 5; }   /* evaluate_maxdm.f90:82-90 */
  double m = 0.0;
  for (int r = 0; r < nr; ++r)
    for (int c = 0; c < nc; ++c) {
      int ac = s2->ao + c;
      if (ac >= nbf) ac = nbf - 1;
      const double v = fabs(P[(size_t)(s1->ao + r) * nbf + ac]);
      if (v > m) m = v;
    }
  return m;
}

/* ------------------------------------------------------------------------ */
/* Digestion of one block into G (closed form of the symmetric-case updates  */
/* of fock_build_non_diag.f90:70-117, 2871-2955 and fock_build_diag.f90)    */
/* ------------------------------------------------------------------------ */
static void digest(const osh* S, int s1, int s2, int s3, int s4, const double* blk, const double* P, double* G,
                   int nbf, double w, int jonly) {
  const osh *A = &S[s1], *B = &S[s2], *C = &S[s3], *D = &S[s4];
  const int na = A->ncomp, nb = B->ncomp, nc = C->ncomp, nd = D->ncomp;
#define PP(i, j) P[(size_t)(i) * nbf + (j)]
#define GG(i, j) G[(size_t)(i) * nbf + (j)]
  for (int ia = 0; ia < na; ++ia)
    for (int ib = 0; ib < nb; ++ib)
      for (int ic = 0; ic < nc; ++ic)
        for (int id = 0; id < nd; ++id) {
          const double e = w * blk[((ia * nb + ib) * nc + ic) * nd + id];
          if (e == 0.0) continue;
          const int a = A->ao + ia, b = B->ao + ib, c = C->ao + ic, d = D->ao + id;
          GG(a, b) += 2 * PP(c, d) * e; GG(b, a) += 2 * PP(c, d) * e;
          GG(c, d) += 2 * PP(a, b) * e; GG(d, c) += 2 * PP(a, b) * e;
          if (jonly) continue;
          GG(a, c) -= .5 * PP(b, d) * e; GG(c, a) -= .5 * PP(b, d) * e;
          GG(a, d) -= .5 * PP(b, c) * e; GG(d, a) -= .5 * PP(b, c) * e;
          GG(b, c) -= .5 * PP(a, d) * e; GG(c, b) -= .5 * PP(a, d) * e;
          GG(b, d) -= .5 * PP(a, c) * e; GG(d, b) -= .5 * PP(a, c) * e;
        }
#undef PP
#undef GG
}

static const char* kPairNames[10] = {"ss", "sp", "sd", "sl", "pp", "pd", "pl", "dd", "dl", "ll"};
/* class order of fock_build_non_diag.f90: (bra type, ket type) */
static void class_types(int cls, int* t1, int* t2) {
  static const char* names[46] = {
      "ssss", "sssp", "sssd", "sssl", "sspp", "sspd", "sspl", "ssdd", "ssdl", "ssll", "spsp", "spsd",
      "spsl", "sppp", "sppd", "sppl", "spdd", "spll", "sdsd", "sdsl", "sdpp", "sdpd", "sddd", "sddl",
      "sdll", "slsl", "slpp", "slpl", "sldd", "sldl", "slll", "pppp", "pppd", "pppl", "ppdd", "ppll",
      "pdpd", "pddd", "plpl", "plll", "dddd", "dddl", "ddll", "dldl", "dlll", "llll"};
  for (int t = 0; t < 10; ++t) {
    if (names[cls][0] == kPairNames[t][0] && names[cls][1] == kPairNames[t][1]) *t1 = t;
    if (names[cls][2] == kPairNames[t][0] && names[cls][3] == kPairNames[t][1]) *t2 = t;
  }
}

/* kind 0: fock_build_non_diag (J - K/2), 1: J only (fock_build_non_diag_j),
 * 2: fock_build_diag, 3: fock_build_initial_diag (schw_out[g] per global pair,
 * pairs numbered type by type).  F += G. */
int oracle_fock(const hfx_mol_desc* d, const double* P, double* F, double thr, int kind, double* schw_out) {
  if (!g_boys_ready) boys_grid();
  const int nbf = d->nb;
  osh* S = (osh*)calloc(d->nsh, sizeof(osh));
  if (load_shells(d, S)) { free(S); return -1; }
  double* md = (double*)malloc(sizeof(double) * d->nsh * d->nsh);
  for (int i = 0; i < d->nsh; ++i)
    for (int j = 0; j <= i; ++j) md[i * d->nsh + j] = md[j * d->nsh + i] = maxdm_pair(S, i, j, P, nbf);
  double* G = (double*)calloc((size_t)nbf * nbf, sizeof(double));
  double* blk = (double*)malloc(sizeof(double) * 625);
  const int slsl_type = 3;
  if (kind >= 2) {
    int g = 0;
    for (int t = 0; t < 10; ++t) {
      for (int k = 0; k < d->npair[t]; ++k, ++g) {
        const int s1 = d->prs[t][2 * k] - 1, s2 = d->prs[t][2 * k + 1] - 1;
        if (kind == 2) {   /* fock_build_diag.f90:17-21 */
          const double mv = fmax(md[s1 * d->nsh + s2], fmax(.25 * md[s2 * d->nsh + s2], .25 * md[s1 * d->nsh + s1]));
          const double sc = d->schw[t][k];
          if (sc * sc * mv < thr) continue;
        }
        eri_block(S, s1, s2, s1, s2, t != slsl_type, blk);
        const int n = S[s1].ncomp * S[s2].ncomp * S[s1].ncomp * S[s2].ncomp;
        double vmax = 0;
        for (int i = 0; i < n; ++i) {
          vmax = fmax(vmax, fabs(blk[i]));
          if (fabs(blk[i]) <= 1.0e-10) blk[i] = 0.0;    /* abs(e) > intthresh */
        }
        if (kind == 3 && schw_out) schw_out[g] = sqrt(vmax);
        const double w = 0.5 * ((s1 == s2) ? 0.25 : 1.0);
        digest(S, s1, s2, s1, s2, blk, P, G, nbf, w, 0);
      }
    }
  } else {
    for (int cls = 0; cls < 46; ++cls) {
      int t1 = 0, t2 = 0;
      class_types(cls, &t1, &t2);
      if (!d->npair[t1] || !d->npair[t2]) continue;
      const int same = t1 == t2;
      for (int ii = 0; ii < d->last[t1]; ++ii) {
        const int pi = d->order[t1][ii] - 1;
        const int s1 = d->prs[t1][2 * pi] - 1, s2 = d->prs[t1][2 * pi + 1] - 1;
        const int jend = same ? ii : d->last[t2];
        for (int jj = 0; jj < jend; ++jj) {
          const int pj = d->order[t2][jj] - 1;
          const int s3 = d->prs[t2][2 * pj] - 1, s4 = d->prs[t2][2 * pj + 1] - 1;
          /* fock_build_non_diag.f90:61-67 */
          double mv = fmax(md[s1 * d->nsh + s2], md[s3 * d->nsh + s4]);
          if (kind == 0) {
            const double ex = fmax(fmax(md[s1 * d->nsh + s3], md[s1 * d->nsh + s4]),
                                   fmax(md[s2 * d->nsh + s3], md[s2 * d->nsh + s4]));
            mv = fmax(mv, .25 * ex);
          }
          if (d->schw[t1][pi] * d->schw[t2][pj] * mv < thr) continue;
          eri_block(S, s1, s2, s3, s4, cls != 25, blk);
          const int n = S[s1].ncomp * S[s2].ncomp * S[s3].ncomp * S[s4].ncomp;
          double vmax = 0;
          for (int i = 0; i < n; ++i) vmax = fmax(vmax, fabs(blk[i]));
          if (vmax < thr) continue;  /* fock_build_non_diag.f90:69 */
          const double w = ((s1 == s2) ? .5 : 1.0) * ((s3 == s4) ? .5 : 1.0);
          digest(S, s1, s2, s3, s4, blk, P, G, nbf, w, kind == 1);
        }
      }
    }
  }
  for (size_t i = 0; i < (size_t)nbf * nbf; ++i) F[i] += G[i];
  free(S); free(md); free(G); free(blk);
  return 0;
}

/* One contracted block (AO order) for 1-based shells; negab rule of the class. */
int oracle_eri_block(const hfx_mol_desc* d, int s1, int s2, int s3, int s4, int skip_negab, double* out) {
  if (!g_boys_ready) boys_grid();
  osh* S = (osh*)calloc(d->nsh, sizeof(osh));
  if (load_shells(d, S)) { free(S); return -1; }
  eri_block(S, s1 - 1, s2 - 1, s3 - 1, s4 - 1, skip_negab, out);
  free(S);
  return 0;
}
