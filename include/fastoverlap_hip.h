/*
 * fastoverlap_hip.h -- C ABI of libfastoverlap_hip.so, the MI355X (gfx950)
 * implementation of fastoverlap's overlap-maximisation hot path.
 *
 * The entry points replace the numerical core behind the reference's
 * drop-in classes and Fortran modules:
 *
 *   fastoverlap/sphericalAlignment.py:160-188, :237-273  SphericalAlign
 *   fastoverlap/soft.py:115-130                          SOFT.iSOFT
 *   fastoverlap/periodicAlignment.py:27-110, :362-460    PeriodicAlign
 *   f90/fastclusters.f90:129-269                         CLUSTERFASTOVERLAP::ALIGN
 *   f90/fastbulk.f90:279-334                             BULKFASTOVERLAP::ALIGN
 *   f90/fastutils.f90:117-167                            SETPERM (group layout)
 *
 * Conventions
 *  - Every function returns FO_OK (0) or a positive FO_ERR_* code and never
 *    terminates the caller; fo_last_error() returns a thread-local message.
 *  - Positions are row-major (natoms, 3) FP64 -- identical to Fortran's flat
 *    3*NATOMS arrays.  Several pairs are concatenated atom-wise.
 *  - Permutation groups ("species") use the SETPERM layout but 0-based:
 *    ngroups[p] groups per pair, group_sizes concatenated over pairs,
 *    group_atoms concatenated (atom indices local to the pair).
 *  - perm[i] (output) = atom of structure 2 matched to atom i of structure 1,
 *    so X2[perm] aligns with X1 (0-based).
 *  - C^l coefficients use the compact l-major order: for l = 0..L,
 *    m1 = -l..l, m2 = -l..l, complex FP64 interleaved (re, im).
 *  - Host buffers are copied in and out; nothing is retained after a
 *    synchronous call returns.  Distinct contexts may be used from
 *    different threads; one context serialises its own calls.
 */
#ifndef FASTOVERLAP_HIP_H
#define FASTOVERLAP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FO_OK 0
#define FO_ERR_INVALID 1      /* bad argument / shape */
#define FO_ERR_HIP 2          /* HIP runtime error */
#define FO_ERR_NOMEM 3        /* exceeds device memory or the context budget */
#define FO_ERR_UNSUPPORTED 4  /* size outside the supported range */
#define FO_ERR_STATE 5        /* call sequence error (e.g. missing stage) */
#define FO_ERR_NOTFOUND 6     /* unknown request id */
#define FO_ERR_CANCELLED 7    /* request cancelled / server stopped */

typedef struct fo_context fo_context;
typedef struct fo_sph_batch fo_sph_batch;
typedef struct fo_per_batch fo_per_batch;
typedef struct fo_server fo_server;

/* ---- library / context ------------------------------------------------ */
int fo_version(void);
const char* fo_last_error(void);
int fo_device_count(int* count);
/* hbm_budget: bytes this context may allocate (0 = no limit). */
int fo_context_create(int device, unsigned long long hbm_budget, fo_context** out);
int fo_context_destroy(fo_context* ctx);
int fo_context_memory(fo_context* ctx, unsigned long long* used, unsigned long long* budget);

/* ---- spherical batches --------------------------------------------------
 * A batch holds npairs cluster pairs sharing one SO(3) grid side grid_n
 * (= 2B; the reference uses 2*(Jmax+1), fastoverlap/sphericalAlignment.py:50).
 * Per pair: natoms, Jmax L (grid_n >= 2(L+1), L <= 63) and kernel width
 * scale (<= 0: mean nearest-neighbour distance / 3, as the reference's
 * calcScale path, :145-147).
 */
#define FO_SPH_NO_INVERT 1     /* only the given orientation (invert=False) */
#define FO_SPH_KEEP_GRID 2     /* keep the FP32 overlap grid(s) */
#define FO_SPH_NO_CENTER 4     /* positions are used as given */
#define FO_SPH_GENERAL 8       /* two independent Hermitian coefficient sets */
#define FO_SPH_DFT 16          /* direct DFT planes even when grid_n is a power of two */
#define FO_SPH_AUTOTUNE 32     /* choose launch shapes by timing on the first full run */
#define FO_SPH_BETA_GEMM 64    /* beta contraction as a split-BF16x3 GEMM on the matrix
                                  cores (FP32 accumulation) instead of FP32 in the plane
                                  kernel; alignment mode, power-of-two grid sides */
#define FO_SPH_EXACT_PLANES 128 /* every beta plane of the grid evaluated in FP64 for the
                                  argmax (exhaustive check of the bounded FP32 selection;
                                  about grid_n times the selection's work) */

/* stages for fo_sph_batch_run (bit mask) */
#define FO_STAGE_PREP 1        /* centroids, kernel width */
#define FO_STAGE_COEFF 2       /* Y_lm, radial kernel, C^l (FP64) */
#define FO_STAGE_GRID 4        /* inverse SOFT (FP32) + exact FP64 argmax */
#define FO_STAGE_ROTOPT 8      /* FP64 rotation refinement from the grid peak */
#define FO_STAGE_FINISH 16     /* orientation choice, LAP, quaternion fit */
#define FO_STAGE_ALL 31

typedef struct {
  int npairs;
  int grid_n;
  const int* natoms;      /* [npairs] */
  const int* L;           /* [npairs] */
  const double* scale;    /* [npairs] */
  const int* ngroups;     /* [npairs] */
  const int* group_sizes; /* [sum ngroups] */
  const int* group_atoms; /* [sum group_sizes] */
  int flags;
} fo_sph_batch_desc;

int fo_sph_batch_create(fo_context* ctx, const fo_sph_batch_desc* desc, fo_sph_batch** out);
/* SphericalHarmonicAlign (fastoverlap/sphericalAlignment.py:276-391): the
 * coefficient stage projects each structure's Gaussian density (width =
 * the pair's scale) on the harmonic-oscillator basis of radial degree
 * <= nmax and length scale harmscale and forms
 * C^l = sum_groups sum_n conj(c1) c2 (:363-368); every other stage is the
 * batch's own.  basis 0: the Python class's basis (:340-356,
 * utils.py:408-427); basis 1: the Fortran module's HARMONICCOEFFS
 * (f90/fastclusters.f90:492-540, :654-687), normalised oscillator functions
 * of r / harmscale. */
int fo_sph_batch_create_harmonic(fo_context* ctx, const fo_sph_batch_desc* desc, int nmax,
                                 double harmscale, int basis, fo_sph_batch** out);
/* per pair, sum over (l, m1, m2) of |C^l_{m1 m2}|^2 of the current
 * coefficients (CALCSIMILARITY's NORM, f90/fastclusters.f90:805-812) */
int fo_sph_batch_coef_norm2(fo_sph_batch* b, double* norm2);
/* harmonic coefficients of one pair after FO_STAGE_COEFF (single-chunk
 * batches): c1, c2 [ngroups][q(l,m)][nmax+1] complex, q(l,m) = l(l+1)/2 + m,
 * 0 <= m <= l <= L (c[l,-m] = (-1)^m conj(c[l,m])) */
int fo_sph_batch_harmonic_coeffs(fo_sph_batch* b, int pair, double* c1, double* c2);
int fo_sph_batch_destroy(fo_sph_batch* b);
/* bytes of device memory the batch holds */
int fo_sph_batch_footprint(const fo_sph_batch* b, unsigned long long* bytes);
/* positions: host arrays of 3*sum(natoms) doubles */
int fo_sph_batch_set_positions(fo_sph_batch* b, const double* pos1, const double* pos2);
/* positions already in device memory (device-to-device copy, async on the batch stream) */
int fo_sph_batch_set_positions_device(fo_sph_batch* b, const void* dpos1, const void* dpos2);
/* compact FP64 coefficients, set A (and B in FO_SPH_GENERAL mode; may be NULL) */
int fo_sph_batch_set_coeffs(fo_sph_batch* b, const double* coefA, const double* coefB);
/* grid peaks / start angles for FO_STAGE_ROTOPT without FO_STAGE_GRID:
 * R0[npairs*norient*3] */
int fo_sph_batch_set_start(fo_sph_batch* b, const double* R0);
/* chosen rotation for FO_STAGE_FINISH alone: R[npairs*3]; the objective of
 * orientation 0 is set to 0 and, with inversion, orientation 1 to +inf */
int fo_sph_batch_set_rotation(fo_sph_batch* b, const double* R);
/* enqueue the stages on the batch stream; use_graph=1 captures/replays a
 * hipGraph of the launches */
int fo_sph_batch_run(fo_sph_batch* b, int stages, int use_graph);
int fo_sph_batch_sync(fo_sph_batch* b);
/* outputs (any pointer may be NULL):
 *  dist[npairs], X1[3N], X2[3N], perm[N], R[npairs*3], rmat[npairs*9],
 *  inverted[npairs] */
int fo_sph_batch_results(fo_sph_batch* b, double* dist, double* X1, double* X2, int* perm,
                         double* R, double* rmat, int* inverted);
/* per (pair, orientation): ind[3], fun, grid peak value, margin, error bound,
 * FP64 candidates, fully re-evaluated planes */
int fo_sph_batch_peaks(fo_sph_batch* b, int* ind, double* fun, double* peak_value,
                       double* margin, double* delta, int* ncand, int* noverflow);
/* rotation refinement per (pair, orientation): R[3], objective, gradient
 * max-norm at the last evaluated point (a final Newton step shorter than
 * 1e-5 rad is taken without evaluating, with the model objective), Newton
 * iterations, objective evaluations (incl. the line search) */
int fo_sph_batch_opt(fo_sph_batch* b, double* R, double* fun, double* gnorm, int* iters,
                     int* evals);
/* kernel width used per pair (resolved on the device when <= 0 was given) */
int fo_sph_batch_scales(fo_sph_batch* b, double* scale);
int fo_sph_batch_coeffs(fo_sph_batch* b, double* coef);
/* FP32 grid(s) of one pair: norient * n^3 floats, axes (alpha, beta, gamma) */
int fo_sph_batch_grid(fo_sph_batch* b, int pair, float* grid);
int fo_sph_batch_norient(const fo_sph_batch* b);
/* kernel times (ms) of the last run: [prep, ylm, radial, contract, plane,
 * select, rotopt, finish]; requires timing enabled */
int fo_sph_batch_timing(fo_sph_batch* b, int enable, float* ms8);
/* Launch shapes of the batch (no counterpart in the reference: the
 * project's autotuner): params[0] = pairs per XCD chunk of the plane kernel,
 * params[1] = reserved (accepted in [-1, 15], no effect),
 * params[2] = radial workgroups per pair (0: default).  set = 0 reads them
 * (after an FO_SPH_AUTOTUNE batch's first full run: the tuned ones), set = 1
 * imposes them.  Every shape gives identical results. */
int fo_sph_batch_tuning(fo_sph_batch* b, int* params, int set);
/* Precision of the grid stage per (pair, orientation): val32 = largest value
 * of the FP32 (or split-BF16x3) grid, val64 = FP64 value at the exact
 * argmax, delta = the bound on |grid - FP64 grid| used to collect the
 * candidates re-evaluated in FP64 (no counterpart in the reference, whose
 * grid is FP64). */
int fo_sph_batch_precision(fo_sph_batch* b, double* val32, double* val64, double* delta);
/* Grid peak start angles per (pair, orientation): the interpolated maximum
 * as Euler angles, R0[npairs*norient*3] (fastoverlap/utils.py:319-338 +
 * fastoverlap/soft.py:127-130). */
int fo_sph_batch_start(fo_sph_batch* b, double* R0);
/* Multi-peak search on the kept grids (FO_SPH_KEEP_GRID, after
 * FO_STAGE_GRID): findPeaks of fastoverlap/utils.py:367-396 as used by
 * findRotations / malign (fastoverlap/sphericalAlignment.py:196-235).  Per
 * (pair, orientation) up to npeaks (<= 64) subtractive Gaussian fits on
 * (2*width+1)^3 windows (width <= 3).  fit[npairs*norient*npeaks*11]: A,
 * mu, s00, s01, s02, s11, s12, s22, centre (3 fractional grid indices);
 * count[npairs*norient] successful fits (0: fall back to the interpolated
 * maximum); residual (may be NULL): the grids after the subtractions. */
int fo_sph_batch_find_peaks(fo_sph_batch* b, int npeaks, int width, double* fit, int* count,
                            float* residual);

/* Build and keep resident the tables of grid side grid_n (Wigner-d table,
 * twiddles, recurrence coefficients) -- the tables fastoverlap/soft.py:48-96
 * builds per SOFT instance.  Allocates; a set-up call for
 * fo_sph_align_noalloc (every other entry point builds tables on demand). */
int fo_sph_prepare(fo_context* ctx, int grid_n);

/* Allocation-free alignment (no counterpart in the reference): the same
 * pipeline as fo_sph_batch_create + _set_positions_device + _run +
 * _results for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * calls no hipMalloc/hipFree: every intermediate lives in a fixed 64 MiB
 * device array per GPU, held until the process ends; a batch needing more
 * returns FO_ERR_NOMEM.  The grid's tables must be resident
 * (fo_sph_prepare, else FO_ERR_STATE).  flags: FO_SPH_NO_INVERT,
 * FO_SPH_NO_CENTER, FO_SPH_DFT.  Inputs d_pos1/d_pos2 and the outputs
 * (d_dist[npairs], d_X2[3N], d_perm[N], d_rmat[9*npairs],
 * d_inverted[npairs]; any output may be NULL) are DEVICE pointers.
 * Synchronous; calls on one GPU share its array and are serialised inside
 * the process (overlapping use of one GPU's array would race). */
int fo_sph_align_noalloc(fo_context* ctx, const fo_sph_batch_desc* desc, const double* d_pos1,
                         const double* d_pos2, double* d_dist, double* d_X2, int* d_perm,
                         double* d_rmat, int* d_inverted);

/* ---- single-call spherical helpers (drop-in SphericalAlign) ------------ */
/* sphHarm: Y[l, m mod (2L+1), j], complex interleaved, (L+1)*(2L+1)*npts */
int fo_sph_harm(fo_context* ctx, int L, int npts, const double* theta, const double* phi,
                double* Y);
/* getEnergyGradient: f = -Re sum X * D(rot), grad[3]; X dense complex
 * (L+1, 2L+1, 2L+1) with m wrapped mod 2L+1 */
int fo_sph_energy(fo_context* ctx, int L, const double* Xdense, const double* rot, double* f,
                  double* grad);
/* calcWignerMatrices(rot): D and grad[3] dense (L+1, 2L+1, 2L+1) complex */
int fo_sph_wigner_D(fo_context* ctx, int L, const double* rot, double* D, double* grad);
/* Wigner table of SOFT(bw) in the dense layout Ds[l, m1, m2, k] (FP64) */
int fo_sph_wigner_table(fo_context* ctx, int bw, double* Ds);

/* SOFT.SOFT(data): forward SO(3) transform of a complex (2bw)^3 grid
 * (axes alpha, beta, gamma; interleaved re, im) into compact coefficients
 * l <= bw-1 (fastoverlap/soft.py:98-113), FP64. */
int fo_soft_forward(fo_context* ctx, int bw, const double* data, double* flmm);

/* ---- linear assignment -------------------------------------------------
 * metric 0: squared Euclidean; 1: Euclidean with minimum image in box[3].
 * perm[i] = row i's column (per group); dist = |X1 - X2[perm]| (periodic for
 * metric 1) */
int fo_lap_align(fo_context* ctx, int natoms, const double* X1, const double* X2, int ngroups,
                 const int* group_sizes, const int* group_atoms, int metric, const double* box,
                 int* perm, double* dist);

/* ---- periodic batches ---------------------------------------------------
 * All pairs share the k-cube half width nk (2nk+1 vectors per axis) and the
 * FFT side F; per pair natoms, box[3] and scale (> 0).
 */
#define FO_PER_KEEP_GRID 2      /* keep |f| (FP32) */
#define FO_PER_INPUT_COEFFS 4   /* coefficients supplied (quickAlign path) */
#define FO_PER_KEEP_COEFFS 8    /* keep per-group structure factors */
#define FO_PER_KEEP_FFT 16      /* keep the complex transform f (FP32) */

#define FO_PSTAGE_SF 1          /* structure factors + weighted product */
#define FO_PSTAGE_GRID 2        /* 3D DFT planes + exact FP64 argmax */
#define FO_PSTAGE_REFINE 4      /* LAP + mean-shift iteration */
#define FO_PSTAGE_ALL 7

typedef struct {
  int npairs;
  int nk;
  int F;
  const int* natoms;
  const double* box;      /* [npairs*3] */
  const double* scale;    /* [npairs] */
  const int* ngroups;
  const int* group_sizes;
  const int* group_atoms;
  int niter;              /* refinement iterations (reference default 10) */
  int flags;
} fo_per_batch_desc;

int fo_per_batch_create(fo_context* ctx, const fo_per_batch_desc* desc, fo_per_batch** out);
int fo_per_batch_destroy(fo_per_batch* b);
int fo_per_batch_footprint(const fo_per_batch* b, unsigned long long* bytes);
int fo_per_batch_set_positions(fo_per_batch* b, const double* pos1, const double* pos2);
int fo_per_batch_set_positions_device(fo_per_batch* b, const void* dpos1, const void* dpos2);
/* per-group coefficients (sum ngroups, 2nk+1, 2nk+1, 2nk+1) complex, for
 * FO_PER_INPUT_COEFFS */
int fo_per_batch_set_coeffs(fo_per_batch* b, const double* C1, const double* C2);
/* initial displacement for FO_PSTAGE_REFINE alone: disp[npairs*3] */
int fo_per_batch_set_disp(fo_per_batch* b, const double* disp);
int fo_per_batch_run(fo_per_batch* b, int stages, int use_graph);
int fo_per_batch_sync(fo_per_batch* b);
int fo_per_batch_results(fo_per_batch* b, double* dist, double* X1, double* X2, int* perm,
                         double* disp);
int fo_per_batch_peaks(fo_per_batch* b, int* ind, double* disp0, double* peak_value,
                       double* margin, double* delta, int* ncand, int* noverflow);
int fo_per_batch_coeffs(fo_per_batch* b, double* C1, double* C2);
int fo_per_batch_product(fo_per_batch* b, double* C);
int fo_per_batch_grid(fo_per_batch* b, int pair, float* fabs_grid);
int fo_per_batch_timing(fo_per_batch* b, int enable, float* ms4);
/* complex f = fftn(C, (F,F,F)) of one pair, FP32 interleaved (re, im)
 * (FO_PER_KEEP_FFT; fastoverlap/periodicAlignment.py:439) */
int fo_per_batch_fft(fo_per_batch* b, int pair, float* f);
/* Per-structure coefficients shared by the pairs (FO_PER_INPUT_COEFFS):
 * Cs[nstruct][ngroups][2nk+1]^3 complex; pair p uses structure idx1[p] as
 * its first and idx2[p] as its second structure (every pair has the same
 * groups) -- the coefficient reuse of alignGroup,
 * fastoverlap/periodicAlignment.py:462-473. */
int fo_per_batch_set_coeffs_indexed(fo_per_batch* b, const double* Cs, int nstruct,
                                    const int* idx1, const int* idx2);
/* Multi-peak search on the kept |f| grids (FO_PER_KEEP_GRID, after
 * FO_PSTAGE_GRID): findDisps with npeaks > 1
 * (fastoverlap/periodicAlignment.py:442-456); layout as
 * fo_sph_batch_find_peaks with one grid per pair. */
int fo_per_batch_find_peaks(fo_per_batch* b, int npeaks, int width, double* fit, int* count,
                            float* residual);

/* ---- serving --------------------------------------------------------------
 * A continuous-batching server: requests are queued by shape class, admitted
 * while their device footprint fits the budget, batched and run on several
 * streams so that batches overlap on the device.
 */
typedef struct {
  int kind;               /* 0 spherical, 1 periodic */
  int natoms;
  const double* pos1;
  const double* pos2;
  int ngroups;
  const int* group_sizes;
  const int* group_atoms;
  /* spherical */
  int L;
  int grid_n;             /* 0: next power of two >= 2(L+1) */
  int invert;
  /* periodic */
  double box[3];
  int nk;
  int F;
  double scale;           /* <= 0: reference default */
} fo_request;

typedef struct {
  int status;
  double dist;
  double R[3];            /* spherical Euler angles */
  double rmat[9];
  int inverted;
  double disp[3];         /* periodic displacement */
  int ind[6];             /* argmax per orientation (spherical) or [0:3] */
  double latency_ms;      /* submit -> result ready */
  double queue_ms;        /* submit -> admitted to a batch */
  int batch_size;         /* pairs in the batch that served it */
} fo_result;

typedef struct {
  int max_batch;          /* pairs per batch */
  int streams;            /* concurrent batches */
  double max_wait_ms;     /* how long a partial batch waits for more work */
  unsigned long long hbm_budget; /* admission limit (bytes) */
} fo_server_config;

int fo_server_create(fo_context* ctx, const fo_server_config* cfg, fo_server** out);
int fo_server_destroy(fo_server* s);
/* copies the request; returns an id */
int fo_server_submit(fo_server* s, const fo_request* req, long long* id);
/* wait (timeout_ms < 0: forever) for a result; X2/perm may be NULL */
int fo_server_wait(fo_server* s, long long id, double timeout_ms, fo_result* res, double* X2,
                   int* perm);
int fo_server_stats(fo_server* s, long long* completed, long long* batches,
                    unsigned long long* peak_bytes);

/* ---- Fortran callers ------------------------------------------------------
 * Bound by the iso_c_binding modules CLUSTERFASTOVERLAP_HIP,
 * BULKFASTOVERLAP_HIP and FASTOVERLAPUTILS_HIP (fastoverlap_amd/fortran).
 * Argument meaning follows CLUSTERFASTOVERLAP::ALIGN
 * (f90/fastclusters.f90:129-155) and BULKFASTOVERLAP::ALIGN
 * (f90/fastbulk.f90:279-289); permgroup is 1-based as in SETPERM
 * (f90/fastutils.f90:117-167); ngroups = 0 means one group of all atoms.
 * RMATBEST is column-major and acts on column vectors of the (inverted,
 * when *inverted) input; bestperm is 1-based.
 */
int fo_f90_cluster_align(double* coordsb, double* coordsa, int natoms, int L, double* kwidth,
                         int invert, int ngroups, const int* npermsize, const int* permgroup,
                         double* distance, double* dist2, double* rmatbest, int* bestperm,
                         int* inverted);
/* ALIGNHARM (f90/fastclusters.f90:271-391): the cluster ALIGN with the
 * coefficients of the harmonic-oscillator basis (HARMONICCOEFFS, :654-687)
 * of radial degree <= n and length scale hwidth; arguments otherwise as
 * fo_f90_cluster_align. */
int fo_f90_cluster_align_harm(double* coordsb, double* coordsa, int natoms, int n, int L,
                              double hwidth, double* kwidth, int invert, int ngroups,
                              const int* npermsize, const int* permgroup, double* distance,
                              double* dist2, double* rmatbest, int* bestperm, int* inverted);
/* CALCOVERLAPMATRICES (f90/fastclusters.f90:853-866): coords column-major
 * [3 natoms][nlist] used as given; norms (sum |C|^2) and maxovers (maximum
 * of the real overlap grid at bandwidth L + 1) column-major [nlist][nlist],
 * symmetric. */
int fo_f90_harm_overlap_matrices(const double* coords, int natoms, int nlist, int n, int L,
                                 double hwidth, double kwidth, int ngroups, const int* npermsize,
                                 const int* permgroup, double* norms, double* maxovers);
/* ndisps = NDISPLACEMENTS (<= 1: the interpolated overlap maximum; above:
 * that many Gaussian-fitted peaks, each refined, as ALIGNCOEFFS,
 * f90/fastbulk.f90:501-565); ohcell 1 = the COMMONS flag OHCELLT, which also
 * tries the 48 octahedral images of coordsa, each from the displacements of
 * the untransformed pair as f90/fastbulk.f90:458-481 does; ohcell 2 takes
 * each image's displacements from its own overlap. */
int fo_f90_bulk_align(double* coordsb, double* coordsa, int natoms, double boxlx, double boxly,
                      double boxlz, double kernelwidth, int ndisps, int ohcell, int ngroups,
                      const int* npermsize, const int* permgroup, double* distance,
                      double* dist2, int* bestperm, double* disp);
/* CALCDEFAULTS (f90/fastbulk.f90:253-277): default kernel width, NWAVE and
 * NFSPACE of the group alignment. */
int fo_f90_bulk_defaults(int natoms, double boxlx, double boxly, double boxlz, double* kwidth,
                         int* nwave, int* nfspace);
/* ALIGNGROUP (f90/fastbulk.f90:336-412): every coords2 structure j aligned
 * onto every coords1 structure i (i >= j only when sym; coords2 may then be
 * NULL).  distmat is column-major [n1][n2] (distmat[i + n1 j]); aligned is
 * [n2][n1][3 natoms] (ALIGNEDCOORDS2(3*NATOMS, N1LIST, N2LIST)); cells sym
 * leaves out are zero.  ndisps = 0 means 10. */
int fo_f90_bulk_align_group(const double* coords1, int n1, const double* coords2, int n2,
                            int natoms, double boxlx, double boxly, double boxlz, double kwidth,
                            int ndisps, int nwave, int nfspace, int sym, int ngroups,
                            const int* npermsize, const int* permgroup, double* distmat,
                            double* aligned);
/* Row-major 3x3 matrix of octahedral operation op (0-based) in the order of
 * OHOPS (f90/alignutils.f90:830-987): y = rmat x for each atom. */
int fo_oh_operation(int op, double* rmat);

#ifdef __cplusplus
}
#endif
#endif /* FASTOVERLAP_HIP_H */
