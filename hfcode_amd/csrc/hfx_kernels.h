// Declarations of the support kernels (hfx_kernels.hip) and their arguments.
#pragma once
#include "hfx_device.h"

namespace hfx {

struct ScreenArgs {
  const int* bra;              // global pair index per bra list position
  const int* ket;
  const int4* binfo;           // per list position: shells a, b (numbered in the molecule), np, npall
  const int4* kinfo;
  const double* schw_bra;      // Schwarz value per list position
  const double* schw_ket;
  const PairRec* pairs;
  const double* maxdm;         // per molecule an ns x ns block, md_stride doubles apart
  long long md_stride;
  const int* bmol;             // molecule per bra position (null: one molecule)
  const int* kseg;             // first ket position of each molecule, nmol + 1 entries (null: [0, nket))
  const double* thr_mol;       // threshold per molecule (null: thr)
  const double* split_mol;     // FP32 split per molecule (null: split)
  int2* list64;
  int2* list32;
  int* count64;
  int* count32;
  unsigned long long* stats64; // [quartets, primitive quartets, bra primitives, bra pairs with a survivor]
  unsigned long long* stats32;
  const unsigned long long* pmax;  // per molecule: max over its maxdm, bits of a non-negative double
  unsigned long long* masks;   // survivor bits: [bra][chunk of HFX_WAVE kets][list 64, list 32]
  double thr;
  double split;                // > 0: bound < split -> FP32 list
  int nbra, nket, same, jonly, nsh, all_prims;   // nsh: row length of a molecule's maxdm block
  int nchunk;                  // chunks per bra row of masks
  int ket_sorted;              // schw_ket is non-increasing within each molecule (the early stop applies)
};

struct DiagScreenArgs {
  const PairRec* pairs;
  const double* schw;          // Schwarz value per pair of the type (list order)
  const double* maxdm;         // per molecule an nsh x nsh block, md_stride doubles apart
  long long md_stride;
  const double* thr_mol;       // threshold per molecule (null: thr)
  int2* list;
  int* count;
  unsigned long long* stats;
  double thr;
  int first, npair, nsh, check, all_prims;
};

struct DigestArgs {
  const int2* list;            // quartet list of the class
  const int* count;            // its length
  const double* eblk;          // integrals of quartets [q0, q1): tiles of 64 quartets, element-major, oriented
                               // layout; floats for the FP32 list (read through the kernel's element type)
  const double* vmax;          // partial max|e| per combination kernel, [combination][quartet - q0]
  int vstride;
  const PairRec* pairs;
  const double* P;
  const float* P32;            // P in FP32: the FP32 list's quartets are contracted in FP32
  double* X;
  double* pairbuf;             // X_cd per ket pair: HFX_PAIRBUF_STRIDE doubles per global pair,
                               // element (i, j) of the pair's shells a, b at i * n_b + j
  double* dump;                // optional: blocks in the reference routine's layout, 625 per quartet
  double* schw_out;            // optional: sqrt(max|e|) per bra pair (initial diagonal build)
  double thr;
  int nb, diag, jonly, q0, q1;
  int lds_rows;                // accumulate rows A and B of X in LDS ((nA + nB) x nb doubles)
  int ncombo;                  // rows of vmax (combination kernels of the list's precision)
  int tb_offset;               // byte offset in LDS of the per-wave staging of X_cd (after the rows)
  int dims[4];                 // components of the oriented shells A, B, C, D
  int swap_bra, swap_ket;      // oriented A is the pair's second shell (bra / ket)
  const double* thr_mol;       // block test threshold per molecule (device), or null: thr
  int prow_offset;             // one-pass digestion: byte offset in LDS of the rows A and B of P
  int run_offset;              // one-pass digestion: byte offset in LDS of one int per wave
};

// One-electron kernels (generated, gen/onee.hip): one thread per work item
// (shell a, shell b, part a, part b; 0-based shells, a >= b), S/T or V block
// of the item written to both triangles of the nb x nb matrices.
struct OneeArgs {
  const int4* items;
  int nitems;
  const int* sh_type;          // per shell: 1 s, 2 p, 3 d, 12 L
  const int* sh_np;            // primitives
  const int* sh_ao;            // first AO, 0-based
  const double* sh_xyz;        // centre, bohr (nsh x 3)
  const double* sh_z;          // exponents (nsh x maxprim)
  const double* sh_c;          // normalised coefficients
  const double* sh_cl;         // L shells: p / s coefficient ratio
  int maxprim;
  const double* charges;       // nuclei
  const double* coords;        // natom x 3, bohr
  int natom;
  const double* boys64;        // HFX_BOYS_ROWS x HFX_BOYS_COLS
  double* S;
  double* T;
  double* V;
  int nb;
};
typedef void (*OneeKernel)(OneeArgs);
extern const OneeKernel kOneeST[3][3];   // [la][lb]
extern const OneeKernel kOneeV[3][3];

__global__ void k_fill_f64(double* x, size_t n, double v);
__global__ void k_fill_i32(int* x, int n, int v);
__global__ void k_fill_u64(unsigned long long* x, int n);
__global__ void k_maxdm(const double* P, int nb, const ShellRec* sh, const int2* molsh, int nsp, double* maxdm,
                        unsigned long long* pmax);
__global__ void k_screen(const ScreenArgs s);
__global__ void k_screen_diag(const DiagScreenArgs s);
__global__ void k_add_sym(const double* X, double* F, int nb, int nmat);
__global__ void k_cut_small(double* e, long long n);
__global__ void k_to_f32(const double* x, float* y, size_t n);
// X += the pair buffer, block by block (every global pair once)
__global__ void k_pairbuf_to_x(const double* pairbuf, const PairRec* pairs, const ShellRec* shells, int npairs,
                               double* X, int nb);

}  // namespace hfx
