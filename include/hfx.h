/* hfx -- MI355X (gfx950) direct-SCF Fock build, C ABI.
 *
 * Drop-in for the two-electron Fock build of brichjnh/hfcode:
 *   fock_build_non_diag(f,p,thr)  fock_build_non_diag.f90:1-10  (J - K/2)
 *   fock_build_non_diag_k(...)    fock_build_non_diag_k.f90     (same routine)
 *   fock_build_non_diag_j_only    fock_build_non_diag_j.f90:1   (J only)
 *   fock_build_diag(f,p,thr)      fock_build_diag.f90:1
 *   fock_build_initial_diag       fock_build_initial_diag.f90:1 (Schwarz tables)
 * All matrices are nb x nb FP64; P must be symmetric.  F is ACCUMULATED into
 * (F += G), never overwritten, as the reference routines do.  Plain pointers
 * and sizes only; every call returns 0 or a negative HFX_E_* code.
 */
#ifndef HFX_H
#define HFX_H

#ifdef __cplusplus
extern "C" {
#endif

#define HFX_OK 0
#define HFX_E_ARG (-1)          /* bad argument (null pointer, nb mismatch, ...) */
#define HFX_E_UNSUPPORTED (-2)  /* basis the reference rejects (P+D+L mix, > 10 primitives) */
#define HFX_E_HIP (-3)          /* HIP runtime error (no device, launch failure) */
#define HFX_E_NOMEM (-4)        /* device allocation failed */
#define HFX_E_STATE (-5)        /* Schwarz tables missing: run hfx_fock_initial_diag first */

/* build modes */
#define HFX_JK 0                /* F += J - K/2        (fock_build_non_diag / _k) */
#define HFX_J_ONLY 1            /* F += J, Coulomb-only screening (fock_build_non_diag_j) */

/* options for hfx_set_option */
#define HFX_OPT_MIXED_SPLIT 1   /* > 0: quartets with Schwarz x density bound below this value
                                   are evaluated in FP32 (accumulation stays FP64); 0 = all FP64 */
#define HFX_OPT_DEBUG_BLOCKS 2  /* 1: keep every integral block of the next build (tests) */
#define HFX_OPT_PROFILE 3       /* 1: time every class kernel with HIP events (hfx_stats.class_ms) */
#define HFX_OPT_CHUNK 4         /* quartets per chunk of the block classes (integrals stored, then
                                   digested), rounded down to a multiple of 256; 0 = largest the
                                   integral buffer holds (default).  A tuning knob: results do not
                                   depend on it. */
#define HFX_OPT_DIGEST 5        /* block classes of a one-molecule context: 1 = digest the integral
                                   blocks in one pass with the resident bra's rows of X and P in LDS
                                   where that measured faster and the bra pairs' runs of quartets are
                                   long (default); 2 = wherever its LDS fits; 0 = never (multi-pass
                                   digestion).  Same sums, in another order. */
#define HFX_OPT_OVERLAP 6       /* 1 (default): with one live context, a build's block-class lists run
                                   in half-buffer chunks, each list's first integral kernels beside the
                                   previous list's last digestion; 0: lists one after another.  A
                                   scheduling knob: same results. */

/* pair types, in the order of the reference's pair lists (molprops.f90:122-131) */
#define HFX_NPAIRTYPES 10       /* ss sp sd sl pp pd pl dd dl ll */
#define HFX_NCLASSES 46         /* (ss|ss) ... (ll|ll), order of fock_build_non_diag.f90 */

typedef struct hfx_ctx hfx_ctx;

/* Molecule + basis, as held by module molprops (molprops.f90:37-168).
 * Shell arrays are indexed by shell; all indices are 1-based, as in Fortran. */
typedef struct hfx_mol_desc {
  int nb;                      /* basis functions (5 spherical d, L = s+p) */
  int nsh;                     /* shells */
  int maxprim;                 /* leading dimension of zeta/djk/djkl (molprops maxprim = 10) */
  const int* shtyp;            /* [nsh] 1 = s, 2 = p, 3 = d, 12 = L  (shtyp) */
  const int* nzet;             /* [nsh] primitives per shell          (nzet) */
  const int* aosh;             /* [nsh] first AO of the shell, 1-based (aosh) */
  const double* xsh;           /* [3*nsh] centres in bohr, (3,nsh)    (xsh) */
  const double* zeta;          /* [maxprim*nsh] exponents             (zeta) */
  const double* djk;           /* [maxprim*nsh] normalised coefficients, s part for L (djk) */
  const double* djkl;          /* [maxprim*nsh] L-shell p/s ratios; ignored for s, p, d (djkl) */
  /* shell-pair lists and Schwarz tables per pair type (ssprs/schwss/orderss/lastss, ...).
   * npair[t] = 0 skips the type.  schw/order may be NULL: the library then builds the
   * Schwarz tables itself on the first hfx_fock_initial_diag call. */
  int npair[HFX_NPAIRTYPES];
  const int* prs[HFX_NPAIRTYPES];       /* [2*npair] shells of each pair, 1-based */
  const double* schw[HFX_NPAIRTYPES];   /* [npair] sqrt(max|(ab|ab)|) */
  const int* order[HFX_NPAIRTYPES];     /* [npair] descending Schwarz order, 1-based */
  int last[HFX_NPAIRTYPES];             /* truncation index (index_sort.f90:95-119) */
} hfx_mol_desc;

/* Per-build statistics (of the most recent build on the context). */
typedef struct hfx_stats {
  long long quartets[HFX_NCLASSES][2];       /* surviving contracted quartets [fp64, fp32] */
  long long prim_quartets[HFX_NCLASSES][2];  /* primitive quartets evaluated [fp64, fp32] */
  long long bra_prims[HFX_NCLASSES][2];      /* bra primitive iterations [fp64, fp32] */
  double flops_fp64, flops_fp32;             /* algorithmic FLOPs of the build */
  double ms_total;                           /* device time of the build, ms */
  double class_flops[HFX_NCLASSES][2];       /* algorithmic FLOPs per class kernel [fp64, fp32] */
  float class_ms[HFX_NCLASSES][2];           /* class span, ms (HFX_OPT_PROFILE = 1): first to last of the
                                                class's launches; with overlapping lists (HFX_OPT_OVERLAP)
                                                it includes some of its neighbours' work */
  long long bra_runs[HFX_NCLASSES][2];       /* bra pairs with a surviving quartet [fp64, fp32] */
} hfx_stats;

/* Create a context on the current HIP device: copies the shells, builds the
 * primitive-pair data (precalc_kabs.f90:94-295) and the Boys table, and
 * allocates the device buffers.  Returns HFX_E_UNSUPPORTED for bases the
 * reference rejects. */
int hfx_ctx_create(const hfx_mol_desc* desc, hfx_ctx** out);
void hfx_ctx_destroy(hfx_ctx* ctx);
/* Device memory released by destroyed contexts stays cached in the process
 * for later contexts (no hipFree, which waits for the whole device, while
 * other jobs run); this hands the cached blocks back to HIP. */
void hfx_release_cached_memory(void);
int hfx_set_option(hfx_ctx* ctx, int key, double value);
/* Launch shape of class cls (0..45) on this context: per_cu workgroups per
 * compute unit at most for its kernels (0: the library's default), and for
 * the block classes at most `chunk` quartets per chunk of stored integrals
 * (0: as many as the integral buffer holds).  Tuning only: results do not
 * depend on it (hfcode_amd/gen/autotune_batch.py picks them per class). */
int hfx_set_class_launch(hfx_ctx* ctx, int cls, int per_cu, long long chunk);

/* F += off-diagonal two-electron part of the Fock matrix.
 * Replaces fock_build_non_diag(f,p,thr) (mode HFX_JK) and
 * fock_build_non_diag_j_only(f,p,thr) (mode HFX_J_ONLY).
 * P and F are host arrays of nb*nb doubles. */
int hfx_fock_nd(hfx_ctx* ctx, const double* P, double* F, int nb, double thr, int mode);

/* F += diagonal-pair part (ab|ab): replaces fock_build_diag(f,p,thr). */
int hfx_fock_diag(hfx_ctx* ctx, const double* P, double* F, int nb, double thr);

/* Builds the Schwarz tables from the diagonal quartets and adds their
 * contribution to F: replaces fock_build_initial_diag(f,p,thr).  Uses the
 * tables passed in hfx_mol_desc when present (they are then not rebuilt). */
int hfx_fock_initial_diag(hfx_ctx* ctx, const double* P, double* F, int nb, double thr);

/* Same builds on device-resident nb*nb arrays (hipMalloc'd), enqueued on the
 * context's stream; `stream` (a hipStream_t, may be NULL for the context
 * stream) is waited on before and made to wait after.  No host copies. */
int hfx_fock_nd_device(hfx_ctx* ctx, const double* dP, double* dF, double thr, int mode, void* stream);
int hfx_fock_diag_device(hfx_ctx* ctx, const double* dP, double* dF, double thr, void* stream);
/* fock_build_initial_diag on device arrays; when the context has no Schwarz
 * tables yet it waits for the build and sorts them on the host. */
int hfx_fock_initial_diag_device(hfx_ctx* ctx, const double* dP, double* dF, double thr, void* stream);

/* The context's HIP stream (a hipStream_t), for callers that enqueue their own
 * work in order with the builds, its number of basis functions (for a batch
 * context: the row stride of its stacked matrices, the largest molecule's
 * basis size) and its number of molecules. */
void* hfx_ctx_stream(hfx_ctx* ctx);
int hfx_ctx_nb(hfx_ctx* ctx);
int hfx_ctx_nmol(hfx_ctx* ctx);

/* ---- Batches: many molecules, one launch sequence ----------------------
 * No counterpart in the reference, which builds one molecule per process
 * (fock_build_non_diag.f90:33-57 distributes one molecule's quartets over
 * OpenMP threads).  A batch context holds nmol molecules: each class is
 * screened and evaluated for all of them in one launch of its kernels, over
 * one quartet list that never pairs shells of different molecules.  Its
 * matrices are stacked: molecule m's nb x nb block (nb = hfx_ctx_nb(ctx),
 * the largest molecule's basis size; a smaller molecule uses the top-left
 * nb_m x nb_m part and leaves the rest zero) starts at element m * nb * nb.
 * Every entry point above takes such stacked arrays on a batch context (the
 * host-array builds copy nmol * nb * nb doubles each way); hfx_ctx_onee needs
 * a single-molecule context. */
int hfx_batch_create(const hfx_mol_desc* descs, int nmol, hfx_ctx** out);
/* Schwarz tables of molecule m (0-based), as hfx_get_schwarz. */
int hfx_batch_get_schwarz(hfx_ctx* ctx, int m, int t, double* schw, int* order, int* last);

#define HFX_BUILD_ZERO_F 1      /* hfx_batch_fock_device: F = 0 before the build */
#define HFX_BUILD_GRAPH 2       /* ... : capture the build in a HIP graph on its first call with these
                                   arguments and replay it on later calls (no host synchronisation, no
                                   count read-back; the pointed-to values may change between calls) */
/* One SCF iteration's two-electron build for every molecule: F += diagonal-
 * pair part at thr_diag[m] (fock_build_diag) + off-diagonal part at thr_nd[m]
 * (fock_build_non_diag), mode HFX_JK or HFX_J_ONLY.  thr_diag, thr_nd and
 * split (FP32 split per molecule, or NULL for the context's
 * HFX_OPT_MIXED_SPLIT) are device arrays of nmol doubles; a molecule with an
 * infinite threshold adds nothing.  Needs the Schwarz tables
 * (hfx_fock_initial_diag*). */
int hfx_batch_fock_device(hfx_ctx* ctx, const double* dP, double* dF, const double* thr_diag, const double* thr_nd,
                          const double* split, int mode, int flags, void* stream);
/* One-electron matrices S, T, V of every molecule (stacked host arrays,
 * nmol * nb * nb each), molecule m with natom[m] point charges; charges and
 * coordinates (bohr) of all molecules concatenated.  hfx_ctx_onee for a batch. */
int hfx_batch_onee(hfx_ctx* ctx, const int* natom, const double* charges, const double* coords, double* S, double* T,
                   double* V);

/* Statistics of the most recent build (blocks until it has finished). */
int hfx_get_stats(hfx_ctx* ctx, hfx_stats* out);

/* Schwarz tables as used by the context (after initial_diag when built by the
 * library).  schw/order have npair[t] entries; order is 1-based. */
int hfx_get_schwarz(hfx_ctx* ctx, int t, double* schw, int* order, int* last);

/* Debug: integral blocks of the most recent build (HFX_OPT_DEBUG_BLOCKS = 1).
 * For class c (0..45) and precision prec (0 fp64, 1 fp32): number of
 * quartets; when quartets/blocks are non-NULL they receive (bra pair, ket
 * pair) 0-based global pair indices and 625 doubles per quartet, each block
 * laid out as the reference routine's output argument (column-major, s shells
 * without a dimension). */
long long hfx_debug_blocks(hfx_ctx* ctx, int c, int prec, int* quartets, double* blocks);

/* Global pair index -> (type, 1-based position in the type's list). */
int hfx_pair_info(hfx_ctx* ctx, int gpair, int* type, int* index);

/* One-electron matrices S, T, V (nb*nb each, symmetric) of the molecule in
 * `desc` with point charges `charges[natom]` at `coords[3*natom]` (bohr).
 * Host set-up, once per molecule; replaces cmpt1e (cmpt1e.f90:1-258). */
int hfx_onee(const hfx_mol_desc* desc, int natom, const double* charges, const double* coords, double* S,
             double* T, double* V);
/* The same matrices computed on the context's GPU (generated Obara--Saika
 * kernels, one thread per shell pair and L-shell part pair); host arrays in
 * and out. */
int hfx_ctx_onee(hfx_ctx* ctx, int natom, const double* charges, const double* coords, double* S, double* T,
                 double* V);

/* hfx_ctx_onee without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Every device buffer
 * (S, T, V 24 nb^2 bytes, work items 8 to 32 nsh^2 bytes, shell tables
 * 420 nsh bytes, nuclei) lies in the library's fixed 48 MiB scratch array, one per
 * GPU for the life of the process; a molecule needing more is refused with
 * HFX_E_UNSUPPORTED (nb up to about 1300).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (from different threads or
 * contexts) race.  Single-molecule contexts only; host arrays in and out. */
int hfx_ctx_onee_noalloc(hfx_ctx* ctx, int natom, const double* charges, const double* coords, double* S, double* T,
                         double* V);

/* Molecules in the reference's own text formats (host set-up, no HIP calls).
 * hfx_molecule_read parses an input file (read_infile.f90:12-73) and a basis
 * file (read_basfile.f90:12-91, with the normalisation of :100-179), both
 * given as text, and builds the shell table and pair lists as
 * spread_basis.f90:53-150 and precalc_kabs.f90:38-90 do.  Errors:
 * HFX_E_ARG malformed text or an element missing from the basis,
 * HFX_E_UNSUPPORTED UHF, more than 10 primitives or a P + D + L basis
 * (spread_basis.f90:174-177). */
typedef struct hfx_molecule hfx_molecule;
int hfx_molecule_read(const char* input_text, const char* basis_text, hfx_molecule** out);
int hfx_molecule_dims(const hfx_molecule* m, int* natom, int* nb, int* nsh, int* nocc, double* unuc);
/* USEDENSITYSUPGUESS, LEVELSHIFT flag and value, PRINT level of the input. */
int hfx_molecule_options(const hfx_molecule* m, int* densup, int* levelshift, double* shift, int* prtlevl);
/* Descriptor for hfx_ctx_create / hfx_onee; its arrays belong to m (valid until hfx_molecule_free). */
int hfx_molecule_desc(const hfx_molecule* m, hfx_mol_desc* desc);
/* 1-based atom of each shell (nsh). */
int hfx_molecule_shells(const hfx_molecule* m, int* atsh);
/* Nuclear charges (natom) and coordinates (natom x 3, bohr). */
int hfx_molecule_atoms(const hfx_molecule* m, double* charges, double* coords);
/* Superposition of the basis file's atomic densities (nb x nb), the guess of
 * direct_hartreefock.f90:50-64; HFX_E_STATE when the basis file has none (dmat = 0). */
int hfx_molecule_density_guess(const hfx_molecule* m, double* P);
void hfx_molecule_free(hfx_molecule* m);

/* Direct RHF on the GPU: the driver loop of direct_hartreefock.f90:98-223
 * (full, rebuild and incremental dP builds, energy, DIIS error
 * S^-1/2 (FPS - SPF) S^-1/2, DIIS over up to ndiis Fock matrices with 1.05
 * damping and the 1.5 coefficient cap, threshold switch, level shift,
 * convergence test) with S^-1/2 from wrapper_dsyev (direct_hartreefock.f90:23-29)
 * and the eigensolves of wrapper_dsyev as rocSOLVER dsyevd.  Every matrix
 * stays on the device on the context's stream; per iteration the host reads
 * a few scalars (energy, max|resid|, max|dP|, the new DIIS inner products)
 * and solves the (n+1) x (n+1) DIIS system (wrapper_dsysv).  No counterpart
 * entry point in the reference: it replaces the driver, not a routine. */
typedef struct hfx_scf_options {
  int maxiter;          /* maxiter (25) */
  int ndiis;            /* ndiis (8), at most 16 */
  double conv_e;        /* conve (5e-8) */
  double conv_p;        /* conv (1e-5) */
  double lowthresh;     /* lowthresh (5e-7) */
  double intthresh;     /* intthresh (1e-10) */
  double fdifffac;      /* fdifffac (0.1) */
  double diis_start;    /* DIIS starts when max|resid| < this (0.3) */
  double levelshift;    /* lshft; 0 = no level shift (levshft false) */
  int mode;             /* HFX_JK */
  int use_graphs;       /* 1: builds after the first iteration replay HIP graphs (hfx_batch_fock_device) */
} hfx_scf_options;
typedef struct hfx_scf_result {
  double energy;        /* etot of the last iteration */
  int iterations;
  int converged;        /* 1 when |dE| < conv_e and max|dP| < conv_p */
  double fock_ms;       /* device time of the Fock builds */
  double total_ms;      /* wall time of hfx_scf_run */
} hfx_scf_result;
void hfx_scf_defaults(hfx_scf_options* opt);
/* S, H: nb*nb host arrays (H = T + V).  P0: starting density (nb*nb, the
 * atomic-density guess) or NULL for the core-Hamiltonian guess.  Outputs
 * (each may be NULL): P the density of the last build, F its Fock matrix,
 * eigs the final orbital energies (nb), history 4 doubles per iteration
 * (etot, deltae, deltap, max|resid|; maxiter rows). */
int hfx_scf_run(hfx_ctx* ctx, const double* S, const double* H, int nocc, double unuc, const double* P0,
                const hfx_scf_options* opt, hfx_scf_result* res, double* P, double* F, double* eigs,
                double* history);

/* Direct RHF for every molecule of a batch context at once (the driver loop
 * of direct_hartreefock.f90:98-223, as hfx_scf_run, for all of them): every
 * molecule's S, H, S^-1/2, P, F, DIIS history and orbitals stay in device
 * memory; per iteration one shared two-electron build (hfx_batch_fock_device,
 * captured in a HIP graph when use_graphs != 0, as is the sequence of
 * products and reductions after it), batched rocBLAS products, one batched
 * rocSOLVER eigensolve, and one small read-back of every molecule's energy,
 * max|error|, max|dP| and DIIS inner products for the host's per-molecule
 * decisions.  Full builds use the context's HFX_OPT_MIXED_SPLIT, incremental
 * builds run in FP64.  A molecule that has converged (or reached maxiter)
 * adds no work to later shared builds and keeps its density.
 * S, H, P0 (or NULL: core-Hamiltonian guess) are stacked host arrays as the
 * context's; nocc and unuc one per molecule; opt->levelshift must be 0.
 * No counterpart in the reference (one molecule per process). */
typedef struct hfx_batch_scf hfx_batch_scf;
int hfx_batch_scf_create(hfx_ctx* ctx, const double* S, const double* H, const int* nocc, const double* unuc,
                         const double* P0, const hfx_scf_options* opt, int use_graphs, hfx_batch_scf** out);
/* One iteration of every molecule still iterating; per molecule status
 * (0 iterating, 1 converged, 2 maxiter reached without converging, < 0 an
 * HFX_E_* failure of that molecule), energy (etot of its last iteration) and
 * iterations (each may be NULL).  Returns the number still iterating, or a
 * negative HFX_E_* code. */
int hfx_batch_scf_step(hfx_batch_scf* b, int* status, double* energy, int* iterations);
/* Molecule m's density of its last build, that build's Fock matrix and its
 * orbital energies (nb_m * nb_m, nb_m * nb_m, nb_m; each may be NULL). */
int hfx_batch_scf_result(hfx_batch_scf* b, int m, double* P, double* F, double* eigs);
void hfx_batch_scf_destroy(hfx_batch_scf* b);

const char* hfx_version(void);

#ifdef __cplusplus
}
#endif
#endif
