/*
 * skkr_gpu.h -- C-ABI of the MI355X (gfx950) per-(E,k) layer kernel.
 *
 * Drop-in for the k-loop body of the reference's cpacoord
 * (skkr_kpar_v6/nrsp-scf/cpacoord.f:147-317,
 *  skkr_kpar_v6/rsp-scf/cpacoord_ksave.f:188-400,
 *  skkr_dbogo_bsf/rsp-scf/cpacoord_ksave.f:289-867): per k-point the layer
 * structure constants G(k), the two lead self-energies by decimation (dugo ->
 * surfgfd), the Godfrin block-tridiagonal sweeps (tau2d) and the BZ-weighted
 * sum of the site-diagonal blocks of tau.
 *
 * Conventions (those of the reference):
 *  - complex arrays are complex*16: interleaved (re, im) doubles, Fortran
 *    column-major per matrix, no padding (leading dimension = matrix width);
 *  - site-major ordering inside a principal layer (PL), row index
 *    kk = site*lk + orbital (skkr_kpar_v6/main/packer.f:47-95);
 *  - the caller owns every host array; the context owns device memory;
 *  - functions return 0 (SKKR_OK) or an SKKR_ERR_* code; where the reference
 *    stops (decimation non-convergence, cpacoord.f:233-238, surfgf.f:100-103)
 *    the per-k status array names the failing k and the host decides.
 *
 * Every entry point is callable from C, C++ and Fortran (ISO_C_BINDING,
 * module skkr_gpu in skkr_budapest_amd/fortran/skkr_gpu_mod.f90).
 */
#ifndef SKKR_GPU_H
#define SKKR_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SKKR_OK = 0,
  SKKR_ERR_ARG = 1,          /* bad argument or size */
  SKKR_ERR_HIP = 2,          /* HIP runtime error (no device, launch failure) */
  SKKR_ERR_NOMEM = 3,        /* device allocation failed */
  SKKR_ERR_SINGULAR = 4,     /* zero pivot in an inversion */
  SKKR_ERR_SGF_DIVERGED = 5, /* decimation |alpha|_F > 1e10 (surfgf.f:89-94) */
  SKKR_ERR_SGF_ITERMAX = 6,  /* decimation iter > itermax (surfgf.f:100-103) */
  SKKR_ERR_UNSUPPORTED = 7,  /* size/mode not built */
  SKKR_ERR_STR_NOCONV = 8,   /* structure-constant shell sum not converged
                                (kamstr.f:358-359,388-389, directg.f:128-131) */
  SKKR_ERR_TOO_LARGE = 9     /* skkr_surfgfd_noalloc: more than its static work space */
};

/* precision modes (no reference counterpart; the reference is FP64 only) */
enum {
  SKKR_FP64 = 0,        /* every product and inversion in FP64 */
  SKKR_MIXED_FP32 = 1,  /* not built: FP32 products lose ~5 digits in the
                           decimation (SKKR_ERR_UNSUPPORTED) */
  SKKR_MIXED_BF16 = 2,  /* split-BF16 products: every complex GEMM of the
                           decimation and the PL sweeps of skkr_tau_bz* as
                           exact BF16 slice products (Ozaki scheme) summed in
                           FP32 and recombined in FP64 -- S = 5 slices per
                           operand (2^-44 relative per product) in the
                           decimation and the sweeps; inversions and the BZ
                           sum in FP64; for Im E < fp64_dec_below_imE the
                           decimation runs in FP64.  The per-routine entry
                           points stay FP64. */
  SKKR_MIXED_INT8 = 3   /* INT8 digit products: the same GEMMs as exact
                           products of S balanced base-256 digits per
                           operand value (S = 6: 2^-47 of its row / column
                           maximum per K-chunk of 32; S = 5 in the sweeps
                           above fp64_dec_below_imE and in the decimation
                           for Im E >= 0.1: 2^-39; S = 4 in a decimation
                           chain's late iterations, once |alpha'| fell to
                           2^-6 of its value after iteration 1, for n >= 64;
                           from 2^-20 on the chain keeps eps^-1 of the
                           iteration before) on the INT8
                           MFMA, summed in int32 and recombined in FP64 per
                           chunk; the rest as SKKR_MIXED_BF16 */
};

typedef struct skkr_ctx skkr_ctx;

typedef struct {
  int lk;              /* site block width (kmymax / lmmax of the reference) */
  int sites_per_pl;    /* ninprc (equal for every PL and both leads) */
  int n_pl;            /* nprc: interface principal layers */
  double errmax;       /* decimation stop |alpha|_F < errmax (/iterpar/) */
  int itermax_l;       /* itermaxl */
  int itermax_r;       /* itermaxr */
  int precision;       /* SKKR_FP64 | SKKR_MIXED_BF16 | SKKR_MIXED_INT8
                          (SKKR_MIXED_FP32: SKKR_ERR_UNSUPPORTED) */
  double fp64_dec_below_imE; /* mixed modes: FP64 decimation when Im E < this */
  int device;          /* HIP device ordinal */
  int batch;           /* k-points per launch group; 0 = automatic */
} skkr_config;

typedef struct {
  double ms_total;        /* wall time of the call, host clock */
  double ms_fourier;      /* device time per phase, HIP events on ctx stream */
  double ms_decimation;
  double ms_sweeps;
  double ms_accumulate;
  double ms_refine;
  double mean_iter_l, mean_iter_r;
  int max_iter;
  int n_batches;
  long long n_launches;   /* kernel launches and graph replays issued by the host */
  /* dominant kernel: launches and summed event time (HIP events bracketing
   * each launch on the ctx stream; filled only when timing_detail != 0) */
  int timing_detail;
  long long dom_launches;
  double dom_ms;
  double dom_flops;       /* algorithmic flops of those launches */
  char dom_name[32];
  double dom_bytes;       /* algorithmic HBM bytes of those launches (A, B read, C written) */
  double dom_flops_exec;  /* SKKR_MIXED_BF16: BF16 MFMA flops the split products
                             executed (slice pairs x 4 real products, padded
                             tiles); SKKR_MIXED_INT8: INT8 MFMA ops (digit
                             pairs x 4 real products, padded tiles) */
  /* inversion launches (gjinv kernels), timed like the GEMMs; in skkr_tau_bz*
   * the flops and bytes of both classes count only the decimation chains that
   * were still iterating (n <= 128) */
  long long inv_launches;
  double inv_ms, inv_flops, inv_bytes;
  /* batches whose blocked inversions flagged multiplier growth and ran again
   * with the row-wise kernel that pivots over all rows (gjinv.f:13-22) */
  int redo_batches;
} skkr_stats;

/* Context: one per host thread / rank; calls on a context are serialised on
 * its HIP stream. */
int skkr_create(skkr_ctx **ctx, const skkr_config *cfg);
void skkr_destroy(skkr_ctx *ctx);
const char *skkr_error_string(int code);

/* Real-space structure constants of the bulk-like stack (replaces the G(k)
 * producer gstore -> kamstr/directg -> scrstr -> getg2d,
 * skkr_kpar_v6/main/gstore.f:6-283).  R: 2*n_R doubles (x,y per vector);
 * GR_intra, GR_inter: n_R matrices n x n (n = lk*sites_per_pl) of the
 * coupling inside a PL and from PL p to PL p+1; G(k) = sum_R G_r(R) e^{-ik.R},
 * the coupling from PL p+1 back to PL p is G_inter(-R)^dagger. */
int skkr_set_structure(skkr_ctx *ctx, int n_R, const double *R,
                       const double *GR_intra, const double *GR_inter);

/* Inverse single-site t-matrices at the current energy (tminv, tminvl,
 * tminvr of cpacoord.f:31-35): lk x lk per site, n_pl*sites_per_pl interface
 * sites then sites_per_pl sites for each lead. */
int skkr_set_tinv(skkr_ctx *ctx, const double *tinv_int, const double *tinv_l,
                  const double *tinv_r);

/* The k-loop body for nk k-points at energy E (cpacoord.f:147-317):
 * tau_bz (lk x lk x n_pl*sites_per_pl) += sum_k w_k tau_site(k).
 * Optional outputs (NULL to skip): tr_tau_k (nk x sites, complex) traces of the
 * unweighted site-diagonal tau per k (taukdiag of cpacoord.f:300-309 summed
 * over l), iters_k (2*nk: left, right decimation iterations), status_k (nk). */
int skkr_tau_bz(skkr_ctx *ctx, double e_re, double e_im, int nk,
                const double *k, const double *w, double *tau_bz,
                double *tr_tau_k, int *iters_k, int *status_k,
                skkr_stats *stats);

/* Screening parameters of the interface sites for the physical (unscreened)
 * representation of tau, phystau.f (skkr_kpar_v6/nrsp-scf/phystau.f:2-58,
 * rsp-scf/phystau.f, skkr_dbogo_bsf/rsp-scf/phystau_d.f):
 *   phys = (1 + A t^-1) tau (1 + t^-1 A) - (A + A t^-1 A),  A = diag(a).
 * a: sites x lk complex (site-major), a[site][i] = alphaintkkr(l(i), site)
 * with the reference's index map (lml for lm, ldex for kappa-mu, the hole
 * alphakkrh for the hole half of a BdG block).  l_of_index: lk ints in
 * [0, nl), the trace group (l) of each basis index.  lk <= 64. */
int skkr_set_screening(skkr_ctx *ctx, int nl, const int *l_of_index, const double *a);

/* skkr_tau_bz plus the physical representation of the k-loop outputs of
 * cpacoord (nrsp-scf/cpacoord.f:294-311, then :375):
 *   tau_bz_phys (sites x lk x lk, accumulated like tau_bz) += L S R - wsum C,
 *     S = this call's BZ sum, wsum = sum of its weights (the reference's
 *     phystau of the whole BZ sum when the weights of all calls add to 1);
 *   taukdiag (nk x nl x sites complex, k fastest: taukdiag(ik,l,li)) = per-l
 *     traces of phystau(tau(k)) of each site (optional, NULL to skip).
 * Needs skkr_set_screening. */
int skkr_tau_bz_phys(skkr_ctx *ctx, double e_re, double e_im, int nk,
                     const double *k, const double *w, double *tau_bz,
                     double *tau_bz_phys, double *taukdiag, int *iters_k,
                     int *status_k, skkr_stats *stats);

/* Options of a context (no reference counterpart; execution choices that
 * leave FP64 results bitwise unchanged):
 *   "graphs"    1/0  capture the PL sweeps of a batch in a HIP graph and replay
 *               it (default 1; bitwise equal to the plain launches)
 *   "gemm_tile" 0 (by shape, or the autotuned choice) | 32 | 64 | 96 | 128
 *   "inv_kernel" 0 blocked MFMA Gauss-Jordan | 1 row-wise Gauss-Jordan
 *                (different rounding; not bitwise equal to 0) | 2 blocked
 *                without row exchanges inside the 16x16 diagonal blocks: for
 *                accretive matrices only (positive-definite anti-Hermitian
 *                part), larger rounding error (also SKKR_INV_KERNEL=rowwise
 *                or nopivot in the environment)
 *   "gemm_3m"   0 four-product | 1 three-product (Gauss) complex GEMM on the
 *               128 tile: 3/4 of the MFMA work, normwise FP64 accuracy,
 *               different rounding (also SKKR_GEMM_3M=1 in the environment)
 *   "split_dec", "split_sw", "split_sw_low"  SKKR_MIXED_BF16: slices per
 *               operand (2..5, or 0 for FP64 products) of the decimation, of
 *               the sweeps, and of the sweeps below fp64_dec_below_imE
 *               (defaults 5, 5, 5); SKKR_MIXED_INT8: digits per operand
 *               (2..6, or 0; defaults 6, 5, 6)
 *   "split_dec_high" the decimation's slices / digits for Im E >= 0.1 (0: as
 *               split_dec; defaults: BF16 0, INT8 5)
 *   "split_dec_late", "dec_late_log2"  slices / digits (2..6, or 0: off) of a
 *               decimation chain's products from the iteration after which
 *               |alpha'|_F is at most 2^-dec_late_log2 (0..60; 0: off) of its
 *               value after iteration 1, when fewer than the decimation's
 *               and n >= 64 (defaults: BF16 0, INT8 4 and 6)
 *   "dec_keep_y_log2" 0..60: a decimation chain's iterations keep eps^-1 of
 *               the iteration before once |alpha'|_F is at most 2^-value of
 *               its value after iteration 1 (eps then changes by far less
 *               than the products' error; mixed modes only, 0: off;
 *               defaults: BF16 0, INT8 20)
 *   "zgemm_split" 0 | 2..5: skkr_zgemm_batch runs the split-BF16 product with
 *               that many slices (0: FP64)
 *   "zgemm_int8" 0 | 2..6: skkr_zgemm_batch runs the INT8 digit product with
 *               that many digits (0: FP64)
 *   "int8_tile" 64 | 128: work-group tile of the INT8 digit product (64 x 64,
 *               two work-groups per CU, the default; or 128 x 64, one per
 *               CU); the same bits (also SKKR_INT8_TILE in the environment) */
int skkr_set_option(skkr_ctx *ctx, const char *key, int value);

typedef struct {
  int tile_nn;          /* chosen work-group tile for n x n x n products */
  int tile_2nn;         /* chosen tile for the 2n x n x n decimation products */
  double us_nn[3];      /* mean launch time, tiles 32 / 64 / 128 (0 = not run) */
  double us_2nn[3];
} skkr_tuning;

/* Time the GEMM tile variants on this context's batch shapes (device buffers
 * of a full batch, contents irrelevant) and keep the fastest per shape. */
int skkr_autotune(skkr_ctx *ctx, skkr_tuning *out);

/* ---- per-routine batched entry points (host arrays in and out) ---------- */

/* G(k) -> PL blocks for nk k-points (the gstore+packer step,
 * skkr_kpar_v6/main/packer.f:6-100): m00 = -G_intra(k), m01 = -G_inter(k),
 * m10 = m01^dagger, each nk x (n x n). */
int skkr_pl_blocks(skkr_ctx *ctx, int nk, const double *k, double *m00,
                   double *m01, double *m10);

/* In-place inverse of nb matrices n x n (gjinv, skkr_kpar_v6/main/gjinv.f:2);
 * n <= 128, or 128 < n <= 256 when n equals the context's lk*sites_per_pl.
 * status (nb, optional) gets SKKR_ERR_SINGULAR per matrix. */
int skkr_gjinv_batch(skkr_ctx *ctx, int n, int nb, double *a, int *status);

/* c = a*b for nb pairs of n x n matrices (doubmt2, matops.f:204-228). */
int skkr_zgemm_batch(skkr_ctx *ctx, int n, int nb, const double *a,
                     const double *b, double *c);

/* Decimation for nb independent semi-infinite chains (surfgfd,
 * skkr_kpar_v6/main/surfgf.f:2): x = surface Green's function, iters and
 * status per chain. */
int skkr_surfgfd_batch(skkr_ctx *ctx, int nb, const double *ml,
                       const double *m0, const double *mr, double *x,
                       int *iters, int *status);

/* Lead self-energy for nb k-points (dugo bulk branch,
 * skkr_kpar_v6/main/dugo.f:9): side 'L' -> Y_1 = m10 g m01, 'R' -> X_P =
 * m01 g m10, with g the surface Green's function of m00 + tinv_lead. */
int skkr_dugo_batch(skkr_ctx *ctx, char side, int nb, const double *m00,
                    const double *m01, const double *m10, double *xy,
                    int *iters, int *status);

/* PL-diagonal tau for nb k-points (replaces tau2d, skkr_kpar_v6/main/tau2d.f:2)
 * with the interface t^-1 set by skkr_set_tinv and the lead self-energies
 * x = X_P, y = Y_1: tau (nb x sites x lk x lk) = w * site-diagonal blocks of
 * (t^-1 - G)^-1's PL-diagonal blocks (the same blocks as the Godfrin sweeps,
 * computed in recursive Green's-function form). */
int skkr_tau2d_batch(skkr_ctx *ctx, int nb, const double *m00,
                     const double *m01, const double *m10, const double *x,
                     const double *y, double w, double *tau, int *status);

/* tau2d for a layered stack (bulkgeo = .false., skkr_kpar_v6/main/tau2d.f:55-56,
 * 90-91): every principal layer p has its own packer blocks of its own g2d
 * plane, m00/m01/m10 (nb x P x n x n, k-major: block (k, p) at
 * (k * P + p) * n * n), as tau2d's packer calls produce them for
 * ipl = p + 1.  The matrix is block-tridiagonal with A_pp = m00_p + t^-1_p,
 * A_p,p+1 = m01_p and A_p+1,p = m10_p+1; m10 of the first and m01 of the last
 * PL are not read (the lead terms are y and x).  Equal sites per PL, as for
 * skkr_tau2d_batch; the other arguments and tau are those of
 * skkr_tau2d_batch. */
int skkr_tau2d_layered_batch(skkr_ctx *ctx, int nb, const double *m00,
                             const double *m01, const double *m10,
                             const double *x, const double *y, double w,
                             double *tau, int *status);

/* tau2d with the reference's own arguments, for one k-point
 * (skkr_kpar_v6/main/tau2d.f:2-3):
 *   tau2d(irel, wgeff, bulkgeo, lmax, lkmax, lkmaxp, nintfc, tminv, g2d, x, y, tau, mdimp)
 * followed by what tau2d takes from param.h and its common blocks: the
 * dimensions lmmaxp, minprc, mprc1, mintfc of g2d(lmmaxp, lmmaxp,
 * -minprc:minprc, minprc, 0:mprc1+1), tminv and tau(lkmaxp, lkmaxp, mintfc);
 * nprc and ninprc(0:nprc+1) of common /lay2d/; fac of common /relfac/ (read
 * when irel = 1).  Arrays are the reference's complex*16 arrays as declared
 * (column-major, padded); x and y are (mdimp, mdimp) with the leading
 * lkmax*ninprc(1) square used.  The packer (packer.f:6-100, relmtrx.f for
 * irel = 1) runs on the GPU from the g2d planes (plane 1 for every PL with
 * bulkgeo, plane p for PL p otherwise), then the sweeps of skkr_tau2d_batch
 * or skkr_tau2d_layered_batch.  tau(:, :, 1:nintfc) is written, the rest of
 * tau is left as it is.  The context must have lk = lkmax, n_pl = nprc and
 * sites_per_pl = ninprc(p) for p = 1..nprc (SKKR_ERR_UNSUPPORTED for
 * unequal principal layers); tminv replaces its interface t^-1, as
 * skkr_set_tinv would.  Returns SKKR_ERR_SINGULAR as gjinv would stop. */
int skkr_tau2d_g2d(skkr_ctx *ctx, int irel, double wgeff, int bulkgeo,
                   int lmax, int lkmax, int lkmaxp, int nintfc,
                   const double *tminv, const double *g2d, const double *x,
                   const double *y, double *tau, int mdimp, int lmmaxp,
                   int minprc, int mprc1, int mintfc, int nprc,
                   const int *ninprc, double fac_re, double fac_im);

/* dugo with the reference's own arguments, for one k-point
 * (skkr_kpar_v6/main/dugo.f:9-10):
 *   dugo(lmax, lkmax, tminvb, lkmaxp, g2d, g2di, xy, mdimp, side, bulk, irel, sgfconv)
 * followed by lmmaxp, minprc (param.h), nprc and ninprc(0:nprc+1) (common
 * /lay2d/), fac (common /relfac/) and iters (out, optional: the decimation
 * iterations).  g2d and g2di are (lmmaxp, lmmaxp, -minprc:minprc, minprc)
 * windows; tminvb is (lkmaxp, lkmaxp, minprc).  Both branches of dugo.f:44-72:
 * bulk (the lead's blocks from g2di, the outer factor from g2d) and not bulk
 * (the lead's blocks from g2d, the interface PL's coupling from g2di).  The
 * decimation stops on the context's errmax and itermax_l / itermax_r (the
 * values of common /iterpar/).  xy(1:n, 1:n), n = lkmax*ninprc(1 or nprc),
 * is written; sgfconv = 0 when the decimation diverged (surfgf.f:89-94),
 * which is also the return code; SKKR_ERR_SGF_ITERMAX where surfgf.f stops.
 * tminvb replaces the context's lead t^-1 of that side. */
int skkr_dugo_g2d(skkr_ctx *ctx, int lmax, int lkmax, const double *tminvb,
                  int lkmaxp, const double *g2d, const double *g2di,
                  double *xy, int mdimp, char side, int bulk, int irel,
                  int *sgfconv, int lmmaxp, int minprc, int nprc,
                  const int *ninprc, double fac_re, double fac_im,
                  int *iters);

/* Point group of an irreducible-wedge k-mesh (intbz = 1,
 * skkr_kpar_v6/nrsp-scf/cpacoord.f:98-101,281-299): the BZ sums of
 * skkr_tau_bz, skkr_tau_bz_phys and skkr_cpa_bz become
 * sum_k (w_k / ng) sum_g dg_g tau_k dg1_g per site, with the ng group
 * matrices dg and their inverses dg1 (ng x lk x lk each, the sphbas output in
 * the context's site basis).  The per-k traces are left as they are (the
 * group does not mix l).  ng = 0 switches it off; ng <= 48.  Pair sums
 * (skkr_tau_bz_pairs) need the full BZ and return SKKR_ERR_UNSUPPORTED while
 * a group with ng > 1 is set. */
int skkr_set_point_group(skkr_ctx *ctx, int ng, const double *dg,
                         const double *dg1);

/* ------------------------------------------------------------------------
 * Layer structure constants (the G(k) producer of the reference):
 * skkr_kpar_v6/main/gstore.f:6-7
 *   gstore(park, ce, factor, bulkgeo, side, lmax, nintfc_true, g2d)
 * for the bulk-like geometry (bulkgeo, gstore.f:99-160), called per k at
 * nrsp-scf/cpacoord.f:205,217,250.  Kambe's Ewald sum (kamstr.f) for close
 * layers, the direct reciprocal sum (directg.f) for layers with
 * |dz| * min|b| >= 5, screening G' = G (1 - alpha G)^-1 over (2 nextra + 3)
 * principal layers (scrstr.f), and the layer-tridiagonal blocks g2d.
 * Non-relativistic lm basis, nl2 = (lmax+1)^2 per layer; lmax <= 3 and
 * nl2 (2 nextra + 3) s <= 512 (above 256 rows 1 - alpha G is inverted by
 * 2 x 2 blocks around the leading 256).
 * ------------------------------------------------------------------------ */
typedef struct skkr_str skkr_str;

typedef struct {
  int lmax;          /* l cutoff of the structure constants (LMAX) */
  int s;             /* layers per principal layer (ninprc) */
  int nextra;        /* extra principal layers each side for the screening
                        (NEXTRA, readgeom.f) */
  double eta;        /* Ewald parameter (ETA, readini.f:161; 'factor' of
                        kamstr.f) */
  double a1[2], a2[2]; /* 2D Bravais basis in a.u. (common/brav2d/,
                          primkr2d.f) */
  int device;        /* HIP device */
  int batch;         /* k-points per batch (0: 256) */
} skkr_str_config;

/* cvec: 3 x (s+1) layer positions in a.u. (x, y, z per layer): the layers of
 * one principal layer and the first layer of the next one (gstore.f:105-122
 * builds the stack from these).  The direct and reciprocal lattices are
 * those of vector2d.f / genlatt2d.f with mdir = mrec = 20000 vectors. */
int skkr_str_create(skkr_str **out, const skkr_str_config *cfg,
                    const double *cvec);
void skkr_str_destroy(skkr_str *sc);

/* g2d for nk k-points at energy E:
 * alpha  complex (lmax+1) x s, l fastest: the screening parameter of each
 *        layer of the principal layer (alphalkkr(l, layer) for side 'L',
 *        alpharkkr for 'R'; alphamat.f);
 * k      2 x nk (park per k);
 * g2d    complex, per k nl2 x nl2 x (2s+1) x s x 3 column-major: the
 *        reference's g2d(m, n, jj = -s..s, icount = 1..s, icpl = 0..2);
 * status per k (optional): SKKR_OK, SKKR_ERR_STR_NOCONV or
 *        SKKR_ERR_SINGULAR. */
int skkr_gstore_batch(skkr_str *sc, double e_re, double e_im, const double *alpha,
                      int nk, const double *k, double *g2d, int *status);

/* Use the layer producer for the structure constants of the k-loop instead
 * of skkr_set_structure's Fourier sum (sc = NULL switches back): every
 * skkr_tau_bz / skkr_tau_bz_phys call then computes, per k at its energy,
 * the screened constants of gstore (side 'L', bulkgeo: cpacoord.f:205) and
 * builds the blocks as packer does (packer.f:36-97): M00, M01, M10 of the
 * middle principal layer for the decimation and the sweeps (tau2d.f:55-56),
 * and the self-energies Y_1 = M10 g_L M01(plane 0), X_P = M01 g_R M10(plane 2)
 * (dugo.f:71-75, 98, 106).  alpha: complex (lmax+1) x s as for
 * skkr_gstore_batch, at the energy of the next tau_bz calls (alphamat.f is
 * energy dependent).  Requires lk = (lmax+1)^2 (the non-relativistic lm
 * basis, one spin channel per context as nrsp-scf runs them) and
 * sites_per_pl = s.  A producer serves one context at a time. */
int skkr_set_structure_kkr(skkr_ctx *ctx, skkr_str *sc, const double *alpha);

/* The same for the relativistic k-loop (rsp-scf): rel = 1 evaluates the
 * structure constants at psq = E + E^2/c^2 (c = 274.072) and rel = 2 at E
 * (rsp-scf/cpacoord_ksave.f:93-99, wrel), then transforms every lm block to
 * the kappa-mu basis with the factor psq/E (packer.f with irel = 1,
 * relmtrx.f:2-45, Clebsch-Gordan tables of clebsch.f:2-70); requires
 * lk = 2 (lmax+1)^2.  rel = 0 is skkr_set_structure_kkr. */
int skkr_set_structure_kkr_rel(skkr_ctx *ctx, skkr_str *sc, const double *alpha,
                               int rel);

/* ------------------------------------------------------------------------
 * CPA loop of one energy: nrsp-scf/cpacoord.f:107-345 around the k-loop,
 * with the solver of nrsp-scf/fullcpa.f (rsp-scf/fullcpa.f is the same in
 * the kappa-mu basis):
 *   repeat: interface t^-1 = coherent mc; k-loop -> tau_c (BZ sum);
 *           cpacor: per site (skipped at concentration 1 or mc = ma or mb)
 *           xa = ((mc - ma)^-1 - tau_c)^-1, xb likewise,
 *           mc -= ((c xa + (1-c) xb)^-1 + tau_c)^-1,
 *           tau_a = tau_c + tau_c xa tau_c, tau_b likewise,
 *           error = max over sites of |dmc|_F / |mc|_F;
 *   until error <= cpatol or itcpam passes (cpacor's concpa / cpacon).
 * All of it runs on the device.  When the device holds every k-point of the
 * mesh (PL blocks, lead self-energies, Y_1: ~(32 + 2P) n^2 x 16 B per k), the
 * passes after the first rerun only the PL sweeps, as the reference reuses
 * its saved g2d and SSPOs (cpacoord.f:168-199); otherwise each pass is a full
 * k-loop.  The leads' t^-1 are those of skkr_set_tinv; the interface t^-1 of
 * the context is left at the final mc.
 *   conc            sites = n_pl * sites_per_pl concentrations of species A
 *   tinv_a, tinv_b  complex lk x lk x sites, species A and B (tminva, tminvb)
 *   tinv_c          complex lk x lk x sites: in, the starting coherent t^-1
 *                   unless init_ata (then tata, fullcpa.f:2-24); out, mc
 *   tau_c           the last pass's BZ sum (overwritten, not accumulated)
 *   tau_a, tau_b    species-projected tau of the last cpacor
 *   itcpa           passes run; cpaerr the last error
 * Returns a k-loop error as skkr_tau_bz does, SKKR_ERR_SINGULAR for a zero
 * pivot in the solver; lk <= 128. */
int skkr_cpa_bz(skkr_ctx *ctx, double e_re, double e_im, int nk, const double *k,
                const double *w, const double *conc, const double *tinv_a,
                const double *tinv_b, double *tinv_c, int init_ata, int itcpam,
                double cpatol, double *tau_c, double *tau_a, double *tau_b,
                int *itcpa, double *cpaerr, int *iters_k, int *status_k,
                skkr_stats *stats);

/* ------------------------------------------------------------------------
 * Off-diagonal tau of interface-site pairs, summed over the BZ with the
 * phase of a lateral lattice vector: the input of the cluster embedding,
 * skkr_dbogo_bsf/clu/cpacoord_save.f:432-478 (identity point operation) with
 * the full tau of clu/tau2d_god.f:2-270:
 *   tau_ij(pair) += sum_k w_k e^{+i k.R} tau_{li,lj}(k),
 *   tau_ji(pair) += sum_k w_k e^{-i k.R} tau_{lj,li}(k),  R = n1 a1 + n2 a2.
 * pairs: 4 x npair ints (li, lj: 1-based interface sites as kpair(3:4),
 * n1, n2 as kpair(1:2)); a1, a2: the 2D Bravais vectors (common/brav2d/).
 * Blocks between principal layers p < q come from the sweeps' Z_t = g_t M01
 * and G_q, p > q from W_t = M10 g_t and G_p (the recursion of tau2dgod).
 * skkr_tau_bz_pairs is skkr_tau_bz plus these sums (tau_ij, tau_ji complex
 * lk x lk x npair, accumulated). */
int skkr_set_pairs(skkr_ctx *ctx, int npair, const int *pairs, const double *a1,
                   const double *a2);
int skkr_tau_bz_pairs(skkr_ctx *ctx, double e_re, double e_im, int nk,
                      const double *k, const double *w, double *tau_bz,
                      double *tau_ij, double *tau_ji, int *iters_k,
                      int *status_k, skkr_stats *stats);

/* ------------------------------------------------------------------------
 * Decimation that allocates nothing.  No counterpart in the reference: for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  The same decimation as
 * skkr_surfgfd_batch (surfgfd, skkr_kpar_v6/main/surfgf.f:2-108, for nb
 * independent chains), but every argument is a DEVICE pointer the caller owns:
 *   ml, m0, mr  nb consecutive n x n complex matrices each (column-major)
 *   x           out: nb surface Green's functions
 *   iters       out: nb decimation iteration counts
 *   status      out: nb SKKR_* codes per chain
 * errmax, itermax as in surfgfd (itermax <= 64); stream: a hipStream_t (NULL:
 * the default stream); the call returns when the results are in place.
 * The chains' work buffers (9 n^2 complex per chain, rounded up to an even
 * chain count, plus bookkeeping) live in one 64 MiB array in device memory
 * that the library holds on every GPU it runs on until the process ends:
 * n <= 128 and nb <= 1024, and larger problems (e.g. more than 28 chains of
 * n = 128, 358 of n = 36) return SKKR_ERR_TOO_LARGE.  Calls that overlap on
 * one GPU share that array and race: serialise them. */
int skkr_surfgfd_noalloc(int device, int n, int nb, const double *ml, const double *m0,
                         const double *mr, double *x, int *iters, int *status,
                         double errmax, int itermax, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SKKR_GPU_H */
