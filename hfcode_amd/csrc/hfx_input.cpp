// Host-side set-up of a molecule from the reference's own text formats, so
// that jobs can be handed over without the Fortran front end:
//   input file   -- read_infile.f90:12-73 (list-directed records: natom,
//                   title, six keyword lines, prtlevl, ncore, one line per
//                   atom "label x y z" in Angstrom, bohr = 1.889725989);
//   basis file   -- read_basfile.f90:12-91 ("ELEMENTS n dmat"; per element
//                   the shell count, the label, per shell "T nprim [scale]"
//                   and nprim lines "zeta c [c_p]", then "****"; with dmat = 1
//                   a DENSITY section, one nbf x nbf block per element in
//                   records of at most 7 values);
//   normalisation -- read_basfile.f90:100-179;
//   shell table  -- spread_basis.f90:53-150 (atoms in input order, shells in
//                   basis-file order, AOs numbered shell by shell, L shells
//                   keep c_p / c_s in djkl), P + D + L rejected
//                   (spread_basis.f90:174-177);
//   pair lists   -- precalc_kabs.f90:38-90 (sh1 = 1..nsh, sh2 = 1..sh1, the
//                   shell of the earlier letter s, p, d, L first);
//   E_nuc        -- nuclear_repulsion.f90:1-20.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "hfx.h"

namespace {

const double kBohr = 1.889725989;
const double kPi = 3.141592653589793238462643383279502884197;
const int kMaxPrim = 10;  // molprops.f90:18
const char* kElements[] = {"h",  "he", "li", "be", "b",  "c",  "n",  "o",  "f",  "ne", "na", "mg",
                           "al", "si", "p",  "s",  "cl", "ar", "k",  "ca", "sc", "ti", "v",  "cr",
                           "mn", "fe", "co", "ni", "cu", "zn", "ga", "ge", "as", "se", "br", "kr"};

std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

// Records (lines) split into whitespace/comma separated tokens; empty lines
// are skipped, as a list-directed read skips them.
struct Records {
  std::vector<std::vector<std::string>> rec;
  size_t pos = 0;
  explicit Records(const char* text) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
      for (auto& ch : line)
        if (ch == ',') ch = ' ';
      std::istringstream ls(line);
      std::vector<std::string> t;
      std::string w;
      while (ls >> w) t.push_back(w);
      if (!t.empty()) rec.push_back(t);
    }
  }
  const std::vector<std::string>* next(size_t need) {
    if (pos >= rec.size() || rec[pos].size() < need) return nullptr;
    return &rec[pos++];
  }
};

bool to_double(const std::string& s, double* v) {
  std::string t = s;
  for (auto& ch : t)
    if (ch == 'd' || ch == 'D') ch = 'e';   // Fortran exponent letter
  char* end = nullptr;
  *v = std::strtod(t.c_str(), &end);
  return end && *end == '\0' && end != t.c_str();
}

bool to_int(const std::string& s, int* v) {
  char* end = nullptr;
  const long x = std::strtol(s.c_str(), &end, 10);
  if (!end || *end != '\0' || end == s.c_str()) return false;
  *v = (int)x;
  return true;
}

// Contraction coefficients times the primitive normalisation, scaled so that
// the contracted function has unit self-overlap (read_basfile.f90:100-179;
// for d the x^2 component is the one normalised).
void normalize(const std::vector<double>& z, int l, std::vector<double>& c) {
  const double pi32 = std::sqrt(kPi * kPi * kPi);
  const size_t n = z.size();
  std::vector<double> g(n);
  for (size_t i = 0; i < n; ++i) {
    double ofac;
    if (l == 0) ofac = std::pow(2.0 * z[i] / kPi, 0.75);
    else if (l == 1) ofac = std::pow(2.0 / kPi, 0.75) * 2.0 * std::pow(z[i], 1.25);
    else ofac = std::pow(2.0 / kPi, 0.75) * 4.0 * std::pow(z[i], 1.75) / std::sqrt(3.0);
    g[i] = c[i] * ofac;
  }
  double nfac = 0.0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      const double p = z[i] + z[j];
      if (l == 0) nfac += pi32 / std::sqrt(p * p * p) * g[i] * g[j];
      else if (l == 1) nfac += pi32 / (2.0 * p * p * std::sqrt(p)) * g[i] * g[j];
      else nfac += pi32 * 3.0 / (4.0 * p * p * p * std::sqrt(p)) * g[i] * g[j];
    }
  const double s = std::sqrt(nfac);
  for (size_t i = 0; i < n; ++i) c[i] = g[i] / s;
}

struct BasShell {
  int type = 0;  // 1 s, 2 p, 3 d, 12 L
  std::vector<double> z, c, cp;
};

struct BasElement {
  std::string label;
  std::vector<BasShell> shells;
  std::vector<double> dens;  // nbf x nbf atomic density (dmat = 1)
  int nbf = 0;
};

int ncomp(int t) { return t == 1 ? 1 : t == 2 ? 3 : t == 3 ? 5 : 4; }
int letter_rank(int t) { return t == 1 ? 0 : t == 2 ? 1 : t == 3 ? 2 : 3; }

}  // namespace

struct hfx_molecule {
  int natom = 0, nocc = 0, nb = 0, nsh = 0;
  int densup = 0, levshft = 0, prtlevl = 0;
  double lshft = 0.0, unuc = 0.0;
  std::vector<double> charges, coords;  // coords: natom x 3, bohr
  std::vector<int> shtyp, nzet, atsh, aosh;
  std::vector<double> xsh, zeta, djk, djkl;  // xsh nsh x 3; the others nsh x kMaxPrim
  std::vector<int> prs[HFX_NPAIRTYPES];      // 2 per pair, 1-based
  std::vector<int> atom_element;              // basis element of each atom
  std::vector<BasElement> basis;
  bool has_density = false;
};

extern "C" int hfx_molecule_read(const char* inp, const char* dat, hfx_molecule** out) {
  if (!inp || !dat || !out) return HFX_E_ARG;
  *out = nullptr;
  hfx_molecule* m = new hfx_molecule();
  auto fail = [&](int code) { delete m; return code; };

  // ---- input file (read_infile.f90:12-73)
  Records ri(inp);
  const std::vector<std::string>* r;
  if (!(r = ri.next(1)) || !to_int((*r)[0], &m->natom) || m->natom < 1) return fail(HFX_E_ARG);
  if (!(r = ri.next(1))) return fail(HFX_E_ARG);                    // title
  if (!(r = ri.next(2))) return fail(HFX_E_ARG);                    // USEDENSITYSUPGUESS
  m->densup = std::tolower((unsigned char)(*r)[1][0]) == 'y';
  if (!(r = ri.next(2))) return fail(HFX_E_ARG);                    // DIRECTSCF
  if (!(r = ri.next(3))) return fail(HFX_E_ARG);                    // LEVELSHIFT flag value
  m->levshft = std::tolower((unsigned char)(*r)[1][0]) == 'y';
  if (!to_double((*r)[2], &m->lshft)) return fail(HFX_E_ARG);
  if (!(r = ri.next(3))) return fail(HFX_E_ARG);                    // UHF flag sz
  if (std::tolower((unsigned char)(*r)[1][0]) == 'y') return fail(HFX_E_UNSUPPORTED);
  if (!(r = ri.next(2))) return fail(HFX_E_ARG);                    // MP2
  if (!(r = ri.next(2))) return fail(HFX_E_ARG);                    // CCSD
  if (!(r = ri.next(2)) || !to_int((*r)[1], &m->prtlevl)) return fail(HFX_E_ARG);   // PRINT
  if (!(r = ri.next(2))) return fail(HFX_E_ARG);                    // ncore
  std::vector<std::string> labels;
  for (int i = 0; i < m->natom; ++i) {
    if (!(r = ri.next(4))) return fail(HFX_E_ARG);
    const std::string lab = lower((*r)[0]);
    int z = 0;
    for (int k = 0; k < (int)(sizeof(kElements) / sizeof(kElements[0])); ++k)
      if (lab == kElements[k]) z = k + 1;
    if (z == 0) return fail(HFX_E_ARG);
    labels.push_back(lab);
    m->charges.push_back((double)z);
    for (int d = 1; d <= 3; ++d) {
      double x;
      if (!to_double((*r)[d], &x)) return fail(HFX_E_ARG);
      m->coords.push_back(x * kBohr);
    }
  }
  int zsum = 0;
  for (double q : m->charges) zsum += (int)q;
  m->nocc = zsum / 2;  // read_infile.f90:78
  for (int i = 0; i < m->natom; ++i)
    for (int j = i + 1; j < m->natom; ++j) {
      double r2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double dx = m->coords[3 * i + d] - m->coords[3 * j + d];
        r2 += dx * dx;
      }
      m->unuc += m->charges[i] * m->charges[j] / std::sqrt(r2);
    }

  // ---- basis file (read_basfile.f90:12-91)
  Records rb(dat);
  int nel = 0, dmat = 0;
  if (!(r = rb.next(3)) || !to_int((*r)[1], &nel) || !to_int((*r)[2], &dmat) || nel < 1) return fail(HFX_E_ARG);
  for (int e = 0; e < nel; ++e) {
    BasElement el;
    int nshell = 0;
    if (!(r = rb.next(1)) || !to_int((*r)[0], &nshell) || nshell < 1) return fail(HFX_E_ARG);
    if (!(r = rb.next(1))) return fail(HFX_E_ARG);
    el.label = lower((*r)[0]);
    for (int s = 0; s < nshell; ++s) {
      BasShell sh;
      int np = 0;
      if (!(r = rb.next(2)) || !to_int((*r)[1], &np) || np < 1) return fail(HFX_E_ARG);
      const char t = (char)std::tolower((unsigned char)(*r)[0][0]);
      sh.type = t == 's' ? 1 : t == 'p' ? 2 : t == 'd' ? 3 : t == 'l' ? 12 : 0;
      if (sh.type == 0) return fail(HFX_E_ARG);
      if (np > kMaxPrim) return fail(HFX_E_UNSUPPORTED);
      for (int k = 0; k < np; ++k) {
        if (!(r = rb.next(sh.type == 12 ? 3 : 2))) return fail(HFX_E_ARG);
        double z, c, cp = 0.0;
        if (!to_double((*r)[0], &z) || !to_double((*r)[1], &c)) return fail(HFX_E_ARG);
        if (sh.type == 12 && !to_double((*r)[2], &cp)) return fail(HFX_E_ARG);
        sh.z.push_back(z);
        sh.c.push_back(c);
        sh.cp.push_back(cp);
      }
      if (sh.type == 12) {
        normalize(sh.z, 0, sh.c);
        normalize(sh.z, 1, sh.cp);
      } else {
        normalize(sh.z, sh.type == 1 ? 0 : sh.type == 2 ? 1 : 2, sh.c);
      }
      el.nbf += ncomp(sh.type);
      el.shells.push_back(sh);
    }
    if (!(r = rb.next(1))) return fail(HFX_E_ARG);  // "****"
    m->basis.push_back(el);
  }
  if (dmat == 1) {
    if (!(r = rb.next(1))) return fail(HFX_E_ARG);  // DENSITY
    for (auto& el : m->basis) {
      if (!(r = rb.next(1))) return fail(HFX_E_ARG);  // element label
      const int n = el.nbf;
      el.dens.assign((size_t)n * n, 0.0);
      for (int ii = 0; ii < n; ++ii)
        for (int k0 = 0; k0 < n; k0 += 7) {
          const int cnt = std::min(7, n - k0);
          if (!(r = rb.next(cnt))) return fail(HFX_E_ARG);
          for (int k = 0; k < cnt; ++k)
            if (!to_double((*r)[k], &el.dens[(size_t)ii * n + k0 + k])) return fail(HFX_E_ARG);
        }
      if (!(r = rb.next(1))) return fail(HFX_E_ARG);  // "****"
    }
    m->has_density = true;
  }

  // ---- shell table (spread_basis.f90:53-150)
  bool has_p = false, has_d = false, has_l = false;
  int iao = 0;
  for (int i = 0; i < m->natom; ++i) {
    int map = -1;
    for (int e = 0; e < (int)m->basis.size(); ++e)
      if (m->basis[e].label == labels[i]) { map = e; break; }
    if (map < 0) return fail(HFX_E_ARG);  // element not in the basis
    m->atom_element.push_back(map);
    for (const auto& sh : m->basis[map].shells) {
      m->shtyp.push_back(sh.type);
      m->nzet.push_back((int)sh.z.size());
      m->atsh.push_back(i + 1);
      m->aosh.push_back(iao + 1);
      for (int d = 0; d < 3; ++d) m->xsh.push_back(m->coords[3 * i + d]);
      for (int k = 0; k < kMaxPrim; ++k) {
        const bool in = k < (int)sh.z.size();
        m->zeta.push_back(in ? sh.z[k] : 0.0);
        m->djk.push_back(in ? sh.c[k] : 0.0);
        m->djkl.push_back(in && sh.type == 12 ? sh.cp[k] / sh.c[k] : 0.0);
      }
      iao += ncomp(sh.type);
      has_p |= sh.type == 2;
      has_d |= sh.type == 3;
      has_l |= sh.type == 12;
    }
  }
  if (has_p && has_d && has_l) return fail(HFX_E_UNSUPPORTED);  // spread_basis.f90:174-177
  m->nb = iao;
  m->nsh = (int)m->shtyp.size();

  // ---- pair lists (precalc_kabs.f90:38-90)
  auto type_index = [](int ta, int tb) {
    static const int idx[4][4] = {{0, 1, 2, 3}, {-1, 4, 5, 6}, {-1, -1, 7, 8}, {-1, -1, -1, 9}};
    return idx[letter_rank(ta)][letter_rank(tb)];
  };
  for (int s1 = 1; s1 <= m->nsh; ++s1)
    for (int s2 = 1; s2 <= s1; ++s2) {
      const int t1 = m->shtyp[s1 - 1], t2 = m->shtyp[s2 - 1];
      const bool keep = letter_rank(t1) <= letter_rank(t2);
      const int a = keep ? s1 : s2, b = keep ? s2 : s1;
      const int t = type_index(keep ? t1 : t2, keep ? t2 : t1);
      m->prs[t].push_back(a);
      m->prs[t].push_back(b);
    }
  *out = m;
  return HFX_OK;
}

extern "C" int hfx_molecule_dims(const hfx_molecule* m, int* natom, int* nb, int* nsh, int* nocc, double* unuc) {
  if (!m) return HFX_E_ARG;
  if (natom) *natom = m->natom;
  if (nb) *nb = m->nb;
  if (nsh) *nsh = m->nsh;
  if (nocc) *nocc = m->nocc;
  if (unuc) *unuc = m->unuc;
  return HFX_OK;
}

extern "C" int hfx_molecule_options(const hfx_molecule* m, int* densup, int* levelshift, double* shift, int* prtlevl) {
  if (!m) return HFX_E_ARG;
  if (densup) *densup = m->densup;
  if (levelshift) *levelshift = m->levshft;
  if (shift) *shift = m->lshft;
  if (prtlevl) *prtlevl = m->prtlevl;
  return HFX_OK;
}

extern "C" int hfx_molecule_desc(const hfx_molecule* m, hfx_mol_desc* d) {
  if (!m || !d) return HFX_E_ARG;
  std::memset(d, 0, sizeof(*d));
  d->nb = m->nb;
  d->nsh = m->nsh;
  d->maxprim = kMaxPrim;
  d->shtyp = m->shtyp.data();
  d->nzet = m->nzet.data();
  d->aosh = m->aosh.data();
  d->xsh = m->xsh.data();
  d->zeta = m->zeta.data();
  d->djk = m->djk.data();
  d->djkl = m->djkl.data();
  for (int t = 0; t < HFX_NPAIRTYPES; ++t) {
    d->npair[t] = (int)m->prs[t].size() / 2;
    d->prs[t] = d->npair[t] ? m->prs[t].data() : nullptr;
  }
  return HFX_OK;
}

extern "C" int hfx_molecule_shells(const hfx_molecule* m, int* atsh) {
  if (!m || !atsh) return HFX_E_ARG;
  std::memcpy(atsh, m->atsh.data(), m->atsh.size() * sizeof(int));
  return HFX_OK;
}

extern "C" int hfx_molecule_atoms(const hfx_molecule* m, double* charges, double* coords) {
  if (!m) return HFX_E_ARG;
  if (charges) std::memcpy(charges, m->charges.data(), m->charges.size() * sizeof(double));
  if (coords) std::memcpy(coords, m->coords.data(), m->coords.size() * sizeof(double));
  return HFX_OK;
}

// Superposition of the atomic density blocks (direct_hartreefock.f90:50-64).
extern "C" int hfx_molecule_density_guess(const hfx_molecule* m, double* P) {
  if (!m || !P) return HFX_E_ARG;
  if (!m->has_density) return HFX_E_STATE;
  std::memset(P, 0, (size_t)m->nb * m->nb * sizeof(double));
  int off = 0;
  for (int i = 0; i < m->natom; ++i) {
    const BasElement& el = m->basis[m->atom_element[i]];
    for (int a = 0; a < el.nbf; ++a)
      for (int b = 0; b < el.nbf; ++b) P[(size_t)(off + a) * m->nb + off + b] = el.dens[(size_t)a * el.nbf + b];
    off += el.nbf;
  }
  return HFX_OK;
}

extern "C" void hfx_molecule_free(hfx_molecule* m) { delete m; }
