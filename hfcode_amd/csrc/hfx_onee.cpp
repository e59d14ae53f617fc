// One-electron matrices S, T, V of a molecule in the basis of include/hfx.h
// (host set-up, once per molecule; the reference computes them in
// cmpt1e.f90:1-258 with one_electron_terms.f90:7-1058).
//
// Obara--Saika for Cartesian Gaussians:
//   overlap     s(i+1,j) = PA s(i,j) + (i s(i-1,j) + j s(i,j-1)) / 2p
//   kinetic     t(i,j)   = b(2j+1) s(i,j) - 2b^2 s(i,j+2) - j(j-1)/2 s(i,j-2)
//   nuclear     [a+1_k|b]^(m) = PA_k [a|b]^(m) - PC_k [a|b]^(m+1)
//                 + a_k/2p ([a-1_k|b]^(m) - [a-1_k|b]^(m+1))
//                 + b_k/2p ([a|b-1_k]^(m) - [a|b-1_k]^(m+1)),
//               [0|0]^(m) = 2 pi / p exp(-ab/p AB^2) F_m(p PC^2),
// then d shells to the five real spherical functions of matharmond
// (index_functions.f90:147-168) and L shells as s + p with the p/s
// coefficient ratio (spread_basis.f90:149-151).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/hfx.h"

namespace {

struct Sh {
  int type, nprim, ao, ncomp;
  double x[3];
  const double *z, *c, *cl;
};

// Boys F_m(x), m = 0..M, to ~1e-16: series for x < 30, asymptotic above.
void boys_host(int M, double x, double* F) {
  if (x < 30.0) {
    long double term = 1.0L / (2 * M + 1), sum = term;
    for (int k = 1; k < 500; ++k) {
      term *= 2.0L * x / (long double)(2 * M + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-20L) break;
    }
    const long double ex = expl(-(long double)x);
    long double f = ex * sum;
    F[M] = (double)f;
    for (int m = M; m > 0; --m) {
      f = (2.0L * x * f + ex) / (long double)(2 * m - 1);
      F[m - 1] = (double)f;
    }
  } else {
    F[0] = 0.5 * std::sqrt(M_PI / x);
    for (int m = 1; m <= M; ++m) F[m] = F[m - 1] * (m - 0.5) / x;
  }
}

std::vector<int> cart_list(int l) {  // (lx, ly, lz) triples, reference order
  std::vector<int> out;
  for (int i = 0; i <= l; ++i)
    for (int n = 0; n <= i; ++n) {
      out.push_back(l - i);
      out.push_back(i - n);
      out.push_back(n);
    }
  return out;
}

// 1D overlap table s[i][j], i <= imax, j <= jmax, for one primitive pair.
void overlap1d(double PA, double PB, double oo2p, int imax, int jmax, double s[8][8]) {
  s[0][0] = 1.0;
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      if (i == 0 && j == 0) continue;
      double v;
      if (i > 0) {
        v = PA * s[i - 1][j];
        if (i > 1) v += (i - 1) * oo2p * s[i - 2][j];
        if (j > 0) v += j * oo2p * s[i - 1][j - 1];
      } else {
        v = PB * s[i][j - 1];
        if (j > 1) v += (j - 1) * oo2p * s[i][j - 2];
      }
      s[i][j] = v;
    }
}

// Nuclear-attraction recursion over (a, b) Cartesian exponents with
// auxiliary index m; memo keyed by (ax,ay,az,bx,by,bz,m), l <= 2 per shell.
struct NucRec {
  double PA[3], PB[3], PC[3], oo2p;
  double F[8];
  double pref;
  double memo[3][3][3][3][3][3][6];
  bool have[3][3][3][3][3][3][6];
  double get(int a0, int a1, int a2, int b0, int b1, int b2, int m) {
    if (a0 < 0 || a1 < 0 || a2 < 0 || b0 < 0 || b1 < 0 || b2 < 0) return 0.0;
    if (have[a0][a1][a2][b0][b1][b2][m]) return memo[a0][a1][a2][b0][b1][b2][m];
    double v;
    int a[3] = {a0, a1, a2}, b[3] = {b0, b1, b2};
    if (a0 + a1 + a2 + b0 + b1 + b2 == 0) {
      v = pref * F[m];
    } else {
      int k = -1;
      bool onA = (a0 + a1 + a2) > 0;
      const int* t = onA ? a : b;
      for (int i = 0; i < 3; ++i)
        if (t[i] > 0) { k = i; break; }
      int a1v[3] = {a[0], a[1], a[2]}, b1v[3] = {b[0], b[1], b[2]};
      const double* P_ = onA ? PA : PB;
      if (onA) a1v[k] -= 1; else b1v[k] -= 1;
      v = P_[k] * get(a1v[0], a1v[1], a1v[2], b1v[0], b1v[1], b1v[2], m) -
          PC[k] * get(a1v[0], a1v[1], a1v[2], b1v[0], b1v[1], b1v[2], m + 1);
      // a_k (reduced) and b_k terms
      const int na = onA ? a1v[k] : a[k];
      const int nbk = onA ? b[k] : b1v[k];
      if (na > 0) {
        int a2v[3] = {a1v[0], a1v[1], a1v[2]};
        a2v[k] -= 1;
        v += na * oo2p * (get(a2v[0], a2v[1], a2v[2], b1v[0], b1v[1], b1v[2], m) -
                          get(a2v[0], a2v[1], a2v[2], b1v[0], b1v[1], b1v[2], m + 1));
      }
      if (nbk > 0) {
        int b2v[3] = {b1v[0], b1v[1], b1v[2]};
        b2v[k] -= 1;
        v += nbk * oo2p * (get(a1v[0], a1v[1], a1v[2], b2v[0], b2v[1], b2v[2], m) -
                           get(a1v[0], a1v[1], a1v[2], b2v[0], b2v[1], b2v[2], m + 1));
      }
    }
    have[a0][a1][a2][b0][b1][b2][m] = true;
    memo[a0][a1][a2][b0][b1][b2][m] = v;
    return v;
  }
};

// spherical d from Cartesian (xx, xy, xz, yy, yz, zz)
const double SQ3 = 1.7320508075688772;
const double kSph[5][6] = {{-0.5, 0, 0, -0.5, 0, 1}, {0, 0, SQ3, 0, 0, 0}, {0, 0, 0, 0, SQ3, 0},
                           {0, SQ3, 0, 0, 0, 0},     {SQ3 / 2, 0, 0, -SQ3 / 2, 0, 0}};

// Parts of a shell: (l, weighted by the L ratio)
int nparts(int type) { return type == 12 ? 2 : 1; }
int part_l(int type, int p) { return type == 1 ? 0 : type == 2 ? 1 : type == 3 ? 2 : p; }

}  // namespace

extern "C" int hfx_onee(const hfx_mol_desc* d, int natom, const double* charges, const double* coords, double* S,
                        double* T, double* V) {
  if (!d || !S || !T || !V || natom < 0 || (natom > 0 && (!charges || !coords))) return HFX_E_ARG;
  const int nb = d->nb;
  std::vector<Sh> sh(d->nsh);
  for (int i = 0; i < d->nsh; ++i) {
    Sh& s = sh[i];
    s.type = d->shtyp[i];
    s.nprim = d->nzet[i];
    s.ao = d->aosh[i] - 1;
    s.ncomp = s.type == 1 ? 1 : s.type == 2 ? 3 : s.type == 3 ? 5 : 4;
    for (int k = 0; k < 3; ++k) s.x[k] = d->xsh[3 * i + k];
    s.z = d->zeta + (size_t)d->maxprim * i;
    s.c = d->djk + (size_t)d->maxprim * i;
    s.cl = d->djkl ? d->djkl + (size_t)d->maxprim * i : nullptr;
    if (s.nprim <= 0 || s.nprim > d->maxprim || s.ao < 0 || s.ao + s.ncomp > nb) return HFX_E_ARG;
  }
  std::memset(S, 0, sizeof(double) * nb * nb);
  std::memset(T, 0, sizeof(double) * nb * nb);
  std::memset(V, 0, sizeof(double) * nb * nb);
  NucRec* nr = new NucRec();
  for (int i = 0; i < d->nsh; ++i) {
    for (int j = 0; j <= i; ++j) {
      const Sh &A = sh[i], &B = sh[j];
      const double AB2 = (A.x[0] - B.x[0]) * (A.x[0] - B.x[0]) + (A.x[1] - B.x[1]) * (A.x[1] - B.x[1]) +
                         (A.x[2] - B.x[2]) * (A.x[2] - B.x[2]);
      for (int pa = 0; pa < nparts(A.type); ++pa) {
        for (int pb = 0; pb < nparts(B.type); ++pb) {
          const int la = part_l(A.type, pa), lb = part_l(B.type, pb);
          const std::vector<int> ca = cart_list(la), cb = cart_list(lb);
          const int na = (int)ca.size() / 3, nbb = (int)cb.size() / 3;
          double s6[6][6] = {{0}}, t6[6][6] = {{0}}, v6[6][6] = {{0}};
          for (int za = 0; za < A.nprim; ++za) {
            for (int zb = 0; zb < B.nprim; ++zb) {
              const double a = A.z[za], b = B.z[zb], p = a + b, oo2p = 0.5 / p;
              double w = A.c[za] * B.c[zb];
              if (A.type == 12 && pa == 1) w *= A.cl[za];
              if (B.type == 12 && pb == 1) w *= B.cl[zb];
              const double Kx = std::exp(-a * b / p * AB2);
              double P[3], PA[3], PB[3];
              for (int k = 0; k < 3; ++k) {
                P[k] = (a * A.x[k] + b * B.x[k]) / p;
                PA[k] = P[k] - A.x[k];
                PB[k] = P[k] - B.x[k];
              }
              double s1[3][8][8];
              for (int k = 0; k < 3; ++k) overlap1d(PA[k], PB[k], oo2p, la + 1, lb + 2, s1[k]);
              const double s00 = std::pow(M_PI / p, 1.5) * Kx;
              for (int x = 0; x < na; ++x)
                for (int y = 0; y < nbb; ++y) {
                  const int* ea = &ca[3 * x];
                  const int* eb = &cb[3 * y];
                  double sx[3], tx[3];
                  for (int k = 0; k < 3; ++k) {
                    const int ii = ea[k], jj = eb[k];
                    sx[k] = s1[k][ii][jj];
                    tx[k] = b * (2 * jj + 1) * s1[k][ii][jj] - 2 * b * b * s1[k][ii][jj + 2] -
                            (jj >= 2 ? 0.5 * jj * (jj - 1) * s1[k][ii][jj - 2] : 0.0);
                  }
                  s6[x][y] += w * s00 * sx[0] * sx[1] * sx[2];
                  t6[x][y] += w * s00 * (tx[0] * sx[1] * sx[2] + sx[0] * tx[1] * sx[2] + sx[0] * sx[1] * tx[2]);
                }
              for (int n = 0; n < natom; ++n) {
                const double* C = coords + 3 * n;
                double PC2 = 0;
                for (int k = 0; k < 3; ++k) {
                  nr->PA[k] = PA[k];
                  nr->PB[k] = PB[k];
                  nr->PC[k] = P[k] - C[k];
                  PC2 += nr->PC[k] * nr->PC[k];
                }
                nr->oo2p = oo2p;
                boys_host(la + lb, p * PC2, nr->F);
                nr->pref = 2.0 * M_PI / p * Kx;
                std::memset(nr->have, 0, sizeof(nr->have));
                for (int x = 0; x < na; ++x)
                  for (int y = 0; y < nbb; ++y) {
                    const int* ea = &ca[3 * x];
                    const int* eb = &cb[3 * y];
                    v6[x][y] -= w * charges[n] * nr->get(ea[0], ea[1], ea[2], eb[0], eb[1], eb[2], 0);
                  }
              }
            }
          }
          // to spherical / AO components and into the matrices
          auto ncomp_part = [](int type, int l) { return type == 3 ? 5 : (l == 0 ? 1 : 3); };
          auto off_part = [](int type, int p) { return (type == 12 && p == 1) ? 1 : 0; };
          const int nca = ncomp_part(A.type, la), ncb = ncomp_part(B.type, lb);
          for (int x = 0; x < nca; ++x)
            for (int y = 0; y < ncb; ++y) {
              double sv = 0, tv = 0, vv = 0;
              for (int u = 0; u < na; ++u) {
                const double fa = A.type == 3 ? kSph[x][u] : (u == x ? 1.0 : 0.0);
                if (fa == 0.0) continue;
                for (int v = 0; v < nbb; ++v) {
                  const double fb = B.type == 3 ? kSph[y][v] : (v == y ? 1.0 : 0.0);
                  if (fb == 0.0) continue;
                  sv += fa * fb * s6[u][v];
                  tv += fa * fb * t6[u][v];
                  vv += fa * fb * v6[u][v];
                }
              }
              const int mu = A.ao + off_part(A.type, pa) + x, nu = B.ao + off_part(B.type, pb) + y;
              S[(size_t)mu * nb + nu] = S[(size_t)nu * nb + mu] = sv;
              T[(size_t)mu * nb + nu] = T[(size_t)nu * nb + mu] = tv;
              V[(size_t)mu * nb + nu] = V[(size_t)nu * nb + mu] = vv;
            }
        }
      }
    }
  }
  delete nr;
  return HFX_OK;
}
