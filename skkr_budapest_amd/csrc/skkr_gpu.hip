// skkr_gpu.hip -- host orchestration and C-ABI of the per-(E,k) layer kernel.
//
// One context owns a HIP stream and the device buffers of a k batch.  Per
// batch of B k-points:
//   1. lattice Fourier sum -> M00' = -G_intra(k), M01 = -G_inter(k),
//      M10 = M01^dagger                                   (gstore+packer)
//   2. decimation of both leads as 2B independent chains   (dugo/surfgfd)
//   3. lead self-energies Y_1 = M10 g_L M01, X_P = M01 g_R M10
//   4. Godfrin sweeps; the PL-diagonal inversion is fused into the X sweep
//      and writes only the site-diagonal blocks              (tau2d)
//   5. BZ accumulation in FP64, k summed in order            (cpacoord BZ sum)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <vector>

#include "../../include/skkr_gpu.h"
#include "kernels_f64.h"
#include "gjinv_blk.h"
#include "zgemm_split.h"
#include "zgemm_i8.h"
#include "kkrlat.h"
#include "kkrstr.h"
#include "cpa.h"

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      fprintf(stderr, "skkr_gpu: %s failed: %s (%s:%d)\n", #x,             \
              hipGetErrorString(e_), __FILE__, __LINE__);                  \
      return SKKR_ERR_HIP;                                                 \
    }                                                                      \
  } while (0)

// Im E from which the decimation takes split_dec_high slices / digits
constexpr double kDecHighImE = 0.1;

// work space of skkr_surfgfd_noalloc, per GPU (64 MiB)
#define SKKR_NOALLOC_SCRATCH_BYTES (64 << 20)

#define CHK(x)                  \
  do {                          \
    int r_ = (x);               \
    if (r_ != SKKR_OK) return r_; \
  } while (0)

typedef double2 z;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct PtrArr {  // device array of matrix pointers
  z** d = nullptr;
  int n = 0;
};

}  // namespace

struct skkr_ctx {
  skkr_config cfg;
  int n = 0, nn = 0, ns = 0, nR = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;        // freed by skkr_destroy
  std::vector<void*> batch_allocs;  // sized by the batch: freed when the batch grows
  // structure and t^-1
  double* dR = nullptr;
  z* dGRi = nullptr;  // intra
  z* dGRo = nullptr;  // inter (p -> p+1)
  z* dGRb = nullptr;  // inter^dagger per R (p+1 -> p), phase e^{+ik.R}
  z* dTint = nullptr;
  z* dTl = nullptr;
  z* dTr = nullptr;
  bool have_structure = false, have_tinv = false;
  // batch buffers
  int B = 0;
  double* dk = nullptr;
  double* dw = nullptr;
  z *M00 = nullptr, *M01 = nullptr, *M10 = nullptr;
  z *eps = nullptr, *xs = nullptr, *AB = nullptr, *Yinv = nullptr, *AYBY = nullptr;
  double* dnorm = nullptr;  // |alpha'|^2 partials per (chain, work-group)
  // 2x2 block inversion scratch (n > 128): assembled matrix, C = A21 A11^-1,
  // D = A11^-1 A12, and views of sub-blocks
  z *Asc = nullptr, *Csc = nullptr, *Dsc = nullptr;
  PtrArr p_Asc, p_A21, p_A12, p_A22, p_Csc, p_Dsc;
  int invB = 0;  // batch capacity of the 2x2 block inversion scratch
  z** dv_out[4] = {nullptr, nullptr, nullptr, nullptr};  // out + offsets (B11, B21, B12, B22)
  int *active = nullptr, *iters = nullptr, *status = nullptr, *itermax = nullptr,
      *nactive = nullptr, *status_sw = nullptr, *active0 = nullptr;
  // per-chain digit counts of the next decimation iteration (full / late) and
  // the iterations a chain ran at the late count; |alpha'| after iteration 1
  int *act_hi = nullptr, *act_lo = nullptr, *nlate = nullptr;
  double* nref = nullptr;
  int* act_inv = nullptr;  // chains whose next decimation iteration inverts eps
  int* nkeep = nullptr;    // and the iterations that kept the previous inverse
  int dec_hint = 8;  // iterations launched before the first convergence read-back
  z *Gs = nullptr, *Zs = nullptr, *Ycur = nullptr, *W = nullptr, *Zb = nullptr,
    *Xc = nullptr, *tauk = nullptr, *trk = nullptr;
  z* taubz = nullptr;
  // pointer arrays
  PtrArr p_ml, p_mr, p_m0, p_tl, p_eps, p_x, p_Yinv, p_AYBY, p_sigma;
  PtrArr p_AB[2], p_alpha[2], p_beta[2];  // ping-pong [alpha; beta]
  PtrArr p_M00, p_M01, p_M10, p_W, p_Z, p_X, p_Y;
  std::vector<PtrArr> p_G, p_Zs, p_tau, p_tr;
  // per-PL blocks of a layered (non-bulk) stack, device layout [p][k]
  z *M00pl = nullptr, *M01pl = nullptr, *M10pl = nullptr;
  int pl_cap = 0;
  std::vector<PtrArr> p_M00pl, p_M01pl, p_M10pl;
  bool layered = false;  // the sweeps read the per-PL blocks
  // point group of an irreducible-wedge mesh: tau_BZ <- (1/ng) sum_g dg tau_BZ dg1
  int ng_sym = 0, sym_cap = 0;
  z *dSym = nullptr, *dSymOut = nullptr, *dSymTmp = nullptr;  // dg (ng), dg1 (ng); sites each
  PtrArr p_symT, p_symOut, p_symTmp;
  std::vector<PtrArr> p_symD, p_symD1;  // dg_g, dg1_g repeated for every site
  // stats
  long long launches = 0;
  int timing_detail = 0;
  std::vector<hipEvent_t> evpool;  // pairs bracketing each dominant-kernel launch
  hipEvent_t phase_ev[5] = {};     // phase boundaries of a tau_bz call, tile timing
  std::vector<double> evflops, evbytes;
  std::vector<int> evclass;  // 0 GEMM, 1 inversion
  size_t evused = 0;
  double dom_ms = 0.0, dom_flops = 0.0, dom_bytes = 0.0;  // GEMM launches
  long long dom_launches = 0;
  double inv_ms = 0.0, inv_flops = 0.0, inv_bytes = 0.0;  // inversion launches
  long long inv_launches = 0;
  int dec_rounds = 0;  // iteration rounds of the last run_decimation
  int redo_batches = 0;  // batches of the last tau_bz call run again with full pivoting
  int inv_kernel = 0;  // 0 blocked MFMA Gauss-Jordan, 1 row-wise register Gauss-Jordan
  int gemm_3m = 0;     // 128-wide tiles: 1 = three-product (Gauss) complex GEMM
  int gemm_tile_force = 0;  // 0 = by shape, else 32 | 64 | 96 | 128
  int tile_nn = 0, tile_2nn = 0;  // autotuned tiles (0 = by shape)
  int graphs = 1;  // sweeps of a batch replayed from a captured graph (bitwise equal to launches)
  bool capturing = false;
  // split products (SKKR_MIXED_BF16 / SKKR_MIXED_INT8): slices (BF16) or
  // digits (INT8) of the GEMMs being launched (0 = FP64 MFMA kernel), the
  // kind, and the choices per phase
  int split = 0;
  int split_i8 = 0;       // 1: INT8 digit products (zgemm_i8.h), 0: BF16 slices
  int i8_tile = 64;       // INT8 kernel: 64 (64 x 64 tiles, two per CU) or 128 (128 x 64, one per CU)
  int split_dec = 5;      // decimation and lead self-energies, Im E >= fp64_dec_below_imE
  int split_dec_high = 0; // the same for Im E >= kDecHighImE (0: as split_dec)
  int split_dec_late = 0; // a chain's decimation iterations once |alpha'| fell below
  int dec_late_log2 = 6;  //   2^-dec_late_log2 of its value after iteration 1 (0: off)
  int dec_keep_y_log2 = 0; // a chain's iterations keep eps^-1 once |alpha'| fell below
                           //   2^-dec_keep_y_log2 of its value after iteration 1 (0: off)
  bool dom_nominal_off = false;  // the launch adds no nominal work to the roofline sums
  int split_sw = 5;       // PL sweeps, Im E >= fp64_dec_below_imE
  int split_sw_low = 5;   // PL sweeps below it (the decimation is FP64 there)
  int zgemm_split = 0;    // skkr_zgemm_batch: 0 FP64, else slices or digits
  int zgemm_i8 = 0;       // skkr_zgemm_batch: the product kind (1 INT8 digits)
  double dom_flops_exec = 0.0;  // executed BF16 MFMA flops (INT8 MFMA ops) of the split launches
  // screening of the interface sites (physical tau, phystau.f)
  bool have_screening = false;
  int nl = 0;
  z* dA = nullptr;       // sites x lk screening parameters per basis index
  int* dLidx = nullptr;  // lk: l of each basis index
  z* physx = nullptr;    // sites x lk x lk scratch
  z* taubz_phys = nullptr;
  z* phk = nullptr;      // per-k per-site per-l traces of a batch
  int phk_cap = 0;
  std::map<int, hipGraphExec_t> sweep_graphs;  // keyed by 8 nb + split
  std::map<long long, hipGraphExec_t> dec_graphs;  // decimation rounds, keyed by iterations and products
  // structure constants from the layer producer instead of the Fourier sum
  skkr_str* str = nullptr;
  z* strAlpha = nullptr;          // (lmax+1) x s screening parameters
  z *MtL = nullptr, *MtR = nullptr;  // outer factors of the lead self-energies
  PtrArr p_mt;                    // per chain: MtL (left), MtR (right)
  int str_rel = 0;                // 0 lm basis; 1, 2 kappa-mu (psq = E + E^2/c^2, or E)
  double *dU1 = nullptr, *dU2 = nullptr;
  int *dI1 = nullptr, *dI2 = nullptr;
  // g2d windows of the reference-shaped tau2d / dugo entries
  z* dG2d = nullptr;
  size_t g2d_cap = 0;
  z* dMt = nullptr;  // dugo's extra coupling block (one chain)
  PtrArr p_dmt;      // every chain -> dMt
  int dmt_cap = 0;   // chains of p_dmt
  int mt_cap = 0;
  double e_re = 0.0, e_im = 0.0;  // energy of the current tau_bz call
  // CPA loop: 1 = keep the batch's blocks and lead self-energies, 2 = reuse
  // them (sweeps only; the k-points are those of the previous pass)
  int cpa_mode = 0;
  z* Y1s = nullptr;  // Y_1 of every k (the sweeps overwrite Ycur)
  int y1_cap = 0;
  int cpa_S = 0, cpa_lk = 0;  // sites and block size of the CPA buffers
  double *dConc = nullptr, *dErr3 = nullptr;
  int *dFlag = nullptr, *dInvSt = nullptr;
  z *dMa = nullptr, *dMb = nullptr, *dDa = nullptr, *dDb = nullptr, *dXab = nullptr, *dH = nullptr,
    *dT1 = nullptr, *dTa = nullptr, *dTb = nullptr;
  PtrArr p_ma, p_mb, p_mc, p_da, p_db, p_xab, p_xa, p_xb, p_h, p_t1, p_tauc, p_ta, p_tb;
  // site pairs of the off-diagonal BZ sums: (li, lj) 0-based, R = n1 a1 + n2 a2
  std::vector<int> pair_li, pair_lj;
  std::vector<double> pair_rx, pair_ry;
  bool pairs_on = false;  // the sweeps keep W_p = M10 g_p of every PL
  z* Ws = nullptr;
  int ws_cap = 0;
  std::vector<PtrArr> p_Ws;
  z *dTij = nullptr, *dTji = nullptr;  // npair x lk x lk accumulators
  int pair_cap = 0;
};

// ------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------
static int dalloc(skkr_ctx* c, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (hipMalloc(p, bytes) != hipSuccess) return SKKR_ERR_NOMEM;
  c->allocs.push_back(*p);
  return SKKR_OK;
}
template <class T>
static int dalloc_t(skkr_ctx* c, T** p, size_t count) {
  return dalloc(c, (void**)p, count * sizeof(T));
}

// buffers sized by the batch (ensure_batch and the per-batch extras)
static int dalloc_b(skkr_ctx* c, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (hipMalloc(p, bytes) != hipSuccess) return SKKR_ERR_NOMEM;
  c->batch_allocs.push_back(*p);
  return SKKR_OK;
}
template <class T>
static int dalloc_bt(skkr_ctx* c, T** p, size_t count) {
  return dalloc_b(c, (void**)p, count * sizeof(T));
}

static int make_ptrs(skkr_ctx* c, PtrArr* pa, const std::vector<z*>& v, bool batch = false) {
  if (batch)
    CHK(dalloc_bt(c, &pa->d, v.size()));
  else
    CHK(dalloc_t(c, &pa->d, v.size()));
  pa->n = (int)v.size();
  HIPCHK(hipMemcpy(pa->d, v.data(), v.size() * sizeof(z*), hipMemcpyHostToDevice));
  return SKKR_OK;
}

// Frees every buffer sized by the batch (before a larger batch is allocated);
// the capacities of the per-batch extras go back to zero, so they are
// allocated again at the new size when next used.
static int free_batch(skkr_ctx* c) {
  if (c->batch_allocs.empty()) return SKKR_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (void* p : c->batch_allocs) hipFree(p);
  c->batch_allocs.clear();
  c->B = 0;
  c->invB = c->ws_cap = c->pl_cap = c->mt_cap = c->phk_cap = c->y1_cap = 0;
  c->Asc = c->Csc = c->Dsc = nullptr;
  return SKKR_OK;
}

static void drop_graphs(skkr_ctx* c) {
  for (auto& kv : c->sweep_graphs) hipGraphExecDestroy(kv.second);
  c->sweep_graphs.clear();
  for (auto& kv : c->dec_graphs) hipGraphExecDestroy(kv.second);
  c->dec_graphs.clear();
}

static int ensure_batch(skkr_ctx* c, int B) {
  if (c->B >= B) return SKKR_OK;
  drop_graphs(c);
  free_batch(c);
  const int n = c->n;
  const size_t nn = c->nn, P = c->cfg.n_pl, lk = c->cfg.lk, ns = c->ns;
  const size_t B2 = 2 * (size_t)B;
  CHK(dalloc_bt(c, &c->dk, 2 * (size_t)B));
  CHK(dalloc_bt(c, &c->dw, (size_t)B));
  CHK(dalloc_bt(c, &c->M00, B * nn));
  CHK(dalloc_bt(c, &c->M01, B * nn));
  CHK(dalloc_bt(c, &c->M10, B * nn));
  CHK(dalloc_bt(c, &c->eps, B2 * nn));
  CHK(dalloc_bt(c, &c->xs, B2 * nn));
  CHK(dalloc_bt(c, &c->AB, 2 * B2 * 2 * nn));
  // |alpha'|^2 partials: one per work-group of the 2n x n product, at most
  // the count of the smallest (32 x 32) tile
  CHK(dalloc_bt(c, &c->dnorm, (size_t)B2 * ((2 * n + 31) / 32) * ((n + 31) / 32)));
  CHK(dalloc_bt(c, &c->Yinv, B2 * nn));
  CHK(dalloc_bt(c, &c->AYBY, B2 * 2 * nn));
  CHK(dalloc_bt(c, &c->active, B2));
  CHK(dalloc_bt(c, &c->iters, B2));
  CHK(dalloc_bt(c, &c->status, B2));
  CHK(dalloc_bt(c, &c->itermax, B2));
  // still-active chain count after each decimation iteration (index = iteration)
  CHK(dalloc_bt(c, &c->nactive, (size_t)std::max(c->cfg.itermax_l, c->cfg.itermax_r) + 2));
  CHK(dalloc_bt(c, &c->active0, B2));
  CHK(dalloc_bt(c, &c->act_hi, B2));
  CHK(dalloc_bt(c, &c->act_lo, B2));
  CHK(dalloc_bt(c, &c->nlate, B2));
  CHK(dalloc_bt(c, &c->nref, B2));
  CHK(dalloc_bt(c, &c->act_inv, B2));
  CHK(dalloc_bt(c, &c->nkeep, B2));
  CHK(dalloc_bt(c, &c->status_sw, (size_t)B));
  CHK(dalloc_bt(c, &c->Gs, P * B * nn));
  CHK(dalloc_bt(c, &c->Zs, (P > 1 ? P - 1 : 1) * B * nn));
  CHK(dalloc_bt(c, &c->Ycur, B * nn));
  CHK(dalloc_bt(c, &c->W, B * nn));
  CHK(dalloc_bt(c, &c->Zb, B * nn));
  CHK(dalloc_bt(c, &c->Xc, B * nn));
  CHK(dalloc_bt(c, &c->tauk, (size_t)B * P * ns * lk * lk / (ns ? 1 : 1)));
  CHK(dalloc_bt(c, &c->trk, (size_t)B * P * ns));
  if (n > 128 && c->invB < B) {
    const size_t h1 = 128, h2 = n - 128;
    CHK(dalloc_bt(c, &c->Asc, B2 * nn));
    CHK(dalloc_bt(c, &c->Csc, B2 * h1 * h2));
    CHK(dalloc_bt(c, &c->Dsc, B2 * h1 * h2));
    for (int v = 0; v < 4; ++v) CHK(dalloc_bt(c, &c->dv_out[v], B2));
  }
  std::vector<z*> v(B2);
  auto fill = [&](PtrArr* pa, auto fn, size_t cnt) -> int {
    std::vector<z*> t(cnt);
    for (size_t j = 0; j < cnt; ++j) t[j] = fn(j);
    return make_ptrs(c, pa, t, true);
  };
  const size_t Bs = B;
  // decimation chain j: left lead for j < B (ml = M10, mr = M01, tinv = left),
  // right lead for j >= B (ml = M01, mr = M10, tinv = right)
  CHK(fill(&c->p_ml, [&](size_t j) { return j < Bs ? c->M10 + j * nn : c->M01 + (j - Bs) * nn; }, B2));
  CHK(fill(&c->p_mr, [&](size_t j) { return j < Bs ? c->M01 + j * nn : c->M10 + (j - Bs) * nn; }, B2));
  CHK(fill(&c->p_m0, [&](size_t j) { return c->M00 + (j % Bs) * nn; }, B2));
  CHK(fill(&c->p_tl, [&](size_t j) { return j < Bs ? c->dTl : c->dTr; }, B2));
  for (int h = 0; h < 2; ++h) {
    z* base = c->AB + (size_t)h * B2 * 2 * nn;
    CHK(fill(&c->p_AB[h], [&](size_t j) { return base + j * 2 * nn; }, B2));
    CHK(fill(&c->p_alpha[h], [&](size_t j) { return base + j * 2 * nn; }, B2));
    CHK(fill(&c->p_beta[h], [&](size_t j) { return base + j * 2 * nn + n; }, B2));
  }
  CHK(fill(&c->p_eps, [&](size_t j) { return c->eps + j * nn; }, B2));
  CHK(fill(&c->p_x, [&](size_t j) { return c->xs + j * nn; }, B2));
  CHK(fill(&c->p_Yinv, [&](size_t j) { return c->Yinv + j * nn; }, B2));
  CHK(fill(&c->p_AYBY, [&](size_t j) { return c->AYBY + j * 2 * nn; }, B2));
  // self-energy outputs: Y_1 into Ycur, X_P into Xc
  CHK(fill(&c->p_sigma, [&](size_t j) { return j < Bs ? c->Ycur + j * nn : c->Xc + (j - Bs) * nn; }, B2));
  if (n > 128 && c->invB < B) {
    const size_t h1 = 128, h2 = n - 128;
    CHK(fill(&c->p_Asc, [&](size_t j) { return c->Asc + j * nn; }, B2));
    CHK(fill(&c->p_A21, [&](size_t j) { return c->Asc + j * nn + h1; }, B2));
    CHK(fill(&c->p_A12, [&](size_t j) { return c->Asc + j * nn + h1 * n; }, B2));
    CHK(fill(&c->p_A22, [&](size_t j) { return c->Asc + j * nn + h1 + h1 * n; }, B2));
    CHK(fill(&c->p_Csc, [&](size_t j) { return c->Csc + j * h1 * h2; }, B2));
    CHK(fill(&c->p_Dsc, [&](size_t j) { return c->Dsc + j * h1 * h2; }, B2));
    c->invB = B;
  }
  CHK(fill(&c->p_M00, [&](size_t j) { return c->M00 + j * nn; }, Bs));
  CHK(fill(&c->p_M01, [&](size_t j) { return c->M01 + j * nn; }, Bs));
  CHK(fill(&c->p_M10, [&](size_t j) { return c->M10 + j * nn; }, Bs));
  CHK(fill(&c->p_W, [&](size_t j) { return c->W + j * nn; }, Bs));
  CHK(fill(&c->p_Z, [&](size_t j) { return c->Zb + j * nn; }, Bs));
  CHK(fill(&c->p_X, [&](size_t j) { return c->Xc + j * nn; }, Bs));
  CHK(fill(&c->p_Y, [&](size_t j) { return c->Ycur + j * nn; }, Bs));
  c->p_G.assign(P, PtrArr());
  c->p_Zs.assign(P, PtrArr());
  c->p_tau.assign(P, PtrArr());
  c->p_tr.assign(P, PtrArr());
  for (size_t p = 0; p < P; ++p) {
    CHK(fill(&c->p_G[p], [&](size_t j) { return c->Gs + (p * Bs + j) * nn; }, Bs));
    if (p + 1 < P) CHK(fill(&c->p_Zs[p], [&](size_t j) { return c->Zs + (p * Bs + j) * nn; }, Bs));
    CHK(fill(&c->p_tau[p], [&](size_t j) { return c->tauk + (j * P * ns + p * ns) * lk * lk; }, Bs));
    CHK(fill(&c->p_tr[p], [&](size_t j) { return c->trk + j * P * ns + p * ns; }, Bs));
  }
  c->B = B;
  return SKKR_OK;
}

// scratch of the 2x2 block inversion only (n > 128), for contexts that
// invert matrices outside the k-loop
static int ensure_inverse(skkr_ctx* c, int B) {
  const int n = c->n;
  if (n <= 128 || c->invB >= B) return SKKR_OK;
  drop_graphs(c);  // captured launches hold the previous buffers
  const size_t nn = c->nn, B2 = 2 * (size_t)B, h1 = 128, h2 = n - 128;
  CHK(dalloc_bt(c, &c->Asc, B2 * nn));
  CHK(dalloc_bt(c, &c->Csc, B2 * h1 * h2));
  CHK(dalloc_bt(c, &c->Dsc, B2 * h1 * h2));
  for (int v = 0; v < 4; ++v) CHK(dalloc_bt(c, &c->dv_out[v], B2));
  auto fill = [&](PtrArr* pa, auto fn, size_t cnt) -> int {
    std::vector<z*> t(cnt);
    for (size_t j = 0; j < cnt; ++j) t[j] = fn(j);
    return make_ptrs(c, pa, t, true);
  };
  CHK(fill(&c->p_Asc, [&](size_t j) { return c->Asc + j * nn; }, B2));
  CHK(fill(&c->p_A21, [&](size_t j) { return c->Asc + j * nn + h1; }, B2));
  CHK(fill(&c->p_A12, [&](size_t j) { return c->Asc + j * nn + h1 * n; }, B2));
  CHK(fill(&c->p_A22, [&](size_t j) { return c->Asc + j * nn + h1 + h1 * n; }, B2));
  CHK(fill(&c->p_Csc, [&](size_t j) { return c->Csc + j * h1 * h2; }, B2));
  CHK(fill(&c->p_Dsc, [&](size_t j) { return c->Dsc + j * h1 * h2; }, B2));
  c->invB = B;
  return SKKR_OK;
}

// The context's phase events, created on first use and kept until
// skkr_destroy, so that no return path leaks them.
static int ensure_phase_events(skkr_ctx* c) {
  for (hipEvent_t& e : c->phase_ev)
    if (!e) HIPCHK(hipEventCreate(&e));
  return SKKR_OK;
}

// ------------------------------------------------------------------------
// kernel launchers
// ------------------------------------------------------------------------
// Dominant-kernel timing: HIP events recorded on the context stream around
// each launch (no synchronisation); read back when the stream has drained.
static void dom_begin(skkr_ctx* c) {
  if (!c->timing_detail || c->capturing) return;
  while (c->evpool.size() < c->evused + 2) {
    hipEvent_t e;
    hipEventCreate(&e);
    c->evpool.push_back(e);
  }
  hipEventRecord(c->evpool[c->evused], c->stream);
}
static void dom_end(skkr_ctx* c, double flops, double bytes, int cls = 0) {
  if (!c->timing_detail || c->capturing) return;
  hipEventRecord(c->evpool[c->evused + 1], c->stream);
  c->evflops.push_back(flops);
  c->evbytes.push_back(bytes);
  c->evclass.push_back(cls);
  c->evused += 2;
}
// call after the stream was synchronised
static void dom_collect(skkr_ctx* c) {
  for (size_t i = 0; i + 1 < c->evused; i += 2) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->evpool[i], c->evpool[i + 1]);
    if (c->evclass[i / 2] == 0) {
      c->dom_ms += ms;
      c->dom_flops += c->evflops[i / 2];
      c->dom_bytes += c->evbytes[i / 2];
      c->dom_launches++;
    } else {
      c->inv_ms += ms;
      c->inv_flops += c->evflops[i / 2];
      c->inv_bytes += c->evbytes[i / 2];
      c->inv_launches++;
    }
  }
  c->evused = 0;
  c->evflops.clear();
  c->evbytes.clear();
  c->evclass.clear();
}

// Work-group tile of a GEMM shape: the context's tuned choice (skkr_autotune
// or SKKR_GEMM_TILE), else the largest tile that the shape fills.
static void gemm_tile(const skkr_ctx* c, int m, int n, int* tm, int* tiles) {
  int t = (m >= 128 && n >= 128) ? 128 : ((m > 32 || n > 32) ? 64 : 32);
  if (t == 64) {
    // the 96 x 96 tile where it pads much less than 64 x 64 (n = 72: 96^2 against 128^2)
    auto area = [&](int q) { return (double)((m + q - 1) / q * q) * ((n + q - 1) / q * q); };
    if (area(96) < 0.85 * area(64)) t = 96;
  }
  if (c && c->tile_nn && m == c->n && n == c->n) t = c->tile_nn;
  if (c && c->tile_2nn && m == 2 * c->n && n == c->n) t = c->tile_2nn;
  if (c && c->gemm_tile_force > 0) t = c->gemm_tile_force;
  *tm = t;
  // the kernels of the 128 tile cover 128 x 64
  const int tn = (t == 128) ? 64 : t;
  *tiles = ((m + t - 1) / t) * ((n + tn - 1) / tn);
}

// Split-BF16 product (zgemm_split.h) with S = c->split slices per operand,
// or with c->split_i8 the INT8 digit product (zgemm_i8.h) with S digits.
static int split_tiles(int m, int n) { return ((m + 63) / 64) * ((n + 63) / 64); }
static int i8_tiles(const skkr_ctx* c, int m, int n) {
  const int t = c->i8_tile;
  return ((m + t - 1) / t) * ((n + 63) / 64);
}
static int split_tiles(const skkr_ctx* c, int m, int n) { return c->split_i8 ? i8_tiles(c, m, n) : split_tiles(m, n); }

static int launch_zgemm_i8(skkr_ctx* c, int m, int n, int k, z** A, int lda, z** B, int ldb,
                           z** C, int ldc, z alpha, z beta, int nb, const int* active,
                           const EpiArgs& epi) {
  dim3 grid(i8_tiles(c, m, n), nb);
  dom_begin(c);
  const z* const* a = (const z* const*)A;
  const z* const* b = (const z* const*)B;
  if (c->i8_tile == 64) {
    switch (c->split) {
      case 2: zgemm_i8_64_kernel<2><<<grid, 256, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 3: zgemm_i8_64_kernel<3><<<grid, 256, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 4: zgemm_i8_64_kernel<4><<<grid, 256, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 5: zgemm_i8_64_kernel<5><<<grid, 256, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 6: zgemm_i8_64_kernel<6><<<grid, 256, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      default: return SKKR_ERR_UNSUPPORTED;
    }
  } else {
    switch (c->split) {
      case 2: zgemm_i8_kernel<2><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 3: zgemm_i8_kernel<3><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 4: zgemm_i8_kernel<4><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 5: zgemm_i8_kernel<5><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      case 6: zgemm_i8_kernel<6><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
      default: return SKKR_ERR_UNSUPPORTED;
    }
  }
  HIPCHK(hipGetLastError());
  // executed INT8 MFMA work: S (S + 1) / 2 digit pairs x 4 real products on
  // the padded tiles (i8_tile x 64 x 32)
  const int ti = c->i8_tile;
  const double pm = (double)((m + ti - 1) / ti * ti), pn = (n + 63) / 64 * 64.0, pk = (k + 31) / 32 * 32.0;
  const double pairs = c->split * (c->split + 1) / 2.0;
  const double nom = c->dom_nominal_off ? 0.0 : 1.0;
  if (c->timing_detail && !c->capturing) c->dom_flops_exec += nom * pairs * 4.0 * 2.0 * pm * pn * pk * nb;
  dom_end(c, nom * 8.0 * m * n * (double)k * nb,
          nom * 16.0 * ((double)m * k + (double)k * n + (double)m * n) * nb);
  c->launches++;
  return SKKR_OK;
}

static int launch_zgemm_split(skkr_ctx* c, int m, int n, int k, z** A, int lda, z** B, int ldb,
                              z** C, int ldc, z alpha, z beta, int nb, const int* active,
                              const EpiArgs& epi) {
  if (c->split_i8) return launch_zgemm_i8(c, m, n, k, A, lda, B, ldb, C, ldc, alpha, beta, nb, active, epi);
  dim3 grid(split_tiles(m, n), nb);
  dom_begin(c);
  const z* const* a = (const z* const*)A;
  const z* const* b = (const z* const*)B;
  switch (c->split) {
    case 2: zgemm_split_kernel<2><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
    case 3: zgemm_split_kernel<3><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
    case 4: zgemm_split_kernel<4><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
    case 5: zgemm_split_kernel<5><<<grid, 512, 0, c->stream>>>(m, n, k, a, lda, b, ldb, C, ldc, alpha, beta, active, epi); break;
    default: return SKKR_ERR_UNSUPPORTED;
  }
  HIPCHK(hipGetLastError());
  // executed BF16 MFMA work: S (S + 1) / 2 slice pairs x 4 real products on
  // the padded 64 x 64 x 32 tiles
  const double pm = (m + 63) / 64 * 64.0, pn = (n + 63) / 64 * 64.0, pk = (k + 31) / 32 * 32.0;
  const double pairs = c->split * (c->split + 1) / 2.0;
  const double nom = c->dom_nominal_off ? 0.0 : 1.0;
  if (c->timing_detail && !c->capturing) c->dom_flops_exec += nom * pairs * 4.0 * 2.0 * pm * pn * pk * nb;
  dom_end(c, nom * 8.0 * m * n * (double)k * nb,
          nom * 16.0 * ((double)m * k + (double)k * n + (double)m * n) * nb);
  c->launches++;
  return SKKR_OK;
}

static int launch_zgemm(skkr_ctx* c, int m, int n, int k, z** A, int lda, z** B, int ldb,
                        z** C, int ldc, z alpha, z beta, int nb, const int* active,
                        const EpiArgs* epi_in = nullptr) {
  if (nb <= 0) return SKKR_OK;
  EpiArgs epi;
  memset(&epi, 0, sizeof(epi));
  if (epi_in) epi = *epi_in;
  if (c && c->split > 0) return launch_zgemm_split(c, m, n, k, A, lda, B, ldb, C, ldc, alpha, beta, nb,
                                                   active, epi);
  int tm, tiles;
  gemm_tile(c, m, n, &tm, &tiles);
  dim3 grid(tiles, nb);
  dom_begin(c);
  const bool m3 = c && c->gemm_3m;
  if (tm == 128 && m3)
    zgemm_f64_kernel<128, 64, 16, 2, true><<<grid, 512, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  else if (tm == 128)  // 8-deep stages: 54 KB LDS, two work-groups per CU
    zgemm_f64_kernel<128, 64, 8><<<grid, 512, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  else if (tm == 96 && m3)
    zgemm_f64_kernel<96, 96, 8, 2, true><<<grid, 576, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  else if (tm == 96)  // 9 waves, 8-deep stages (52 KB LDS)
    zgemm_f64_kernel<96, 96, 8><<<grid, 576, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  else if (tm == 64 && m3)
    zgemm_f64_kernel<64, 64, 16, 2, true><<<grid, 256, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  else if (tm == 64)
    zgemm_f64_kernel<64, 64, 16><<<grid, 256, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  else if (m3)
    zgemm_f64_kernel<32, 32, 16, 2, true><<<grid, 64, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  else
    zgemm_f64_kernel<32, 32, 16><<<grid, 64, 0, c->stream>>>(
        m, n, k, (const z* const*)A, lda, (const z* const*)B, ldb, C, ldc, alpha, beta, active, epi);
  HIPCHK(hipGetLastError());
  dom_end(c, 8.0 * m * n * (double)k * nb,
          16.0 * ((double)m * k + (double)k * n + (double)m * n) * nb);
  c->launches++;
  return SKKR_OK;
}

// The blocked kernel for T = ceil(n / 16) block columns (the matrix padded
// with the identity to 16 T rows): T x T tiles of 16 x 16 over
// (T / TI) x (T / TJ) waves, each holding TI x TJ tiles in registers.  The
// work goes as T^3, so n = 72 runs as 80 rather than 128 rows.
template <bool PIV>
static int launch_gjinv_blk(const InvArgs& a, int nb, hipStream_t st) {
  switch ((a.n + 15) / 16) {
    case 1:
    case 2: gjinv_blk_f64_kernel<2, 2, 2, PIV><<<nb, 64, 0, st>>>(a); break;
    case 3:  // <3,1,3> spills; <3,1,1> (9 waves) measured no faster than 64 rows at n = 36
    case 4: gjinv_blk_f64_kernel<4, 2, 2, PIV><<<nb, 256, 0, st>>>(a); break;
    case 5: gjinv_blk_f64_kernel<5, 1, 5, PIV><<<nb, 320, 0, st>>>(a); break;
    case 6: gjinv_blk_f64_kernel<6, 2, 3, PIV><<<nb, 384, 0, st>>>(a); break;
    case 7: gjinv_blk_f64_kernel<7, 1, 7, PIV><<<nb, 448, 0, st>>>(a); break;
    case 8: gjinv_blk_f64_kernel<8, 2, 4, PIV><<<nb, 512, 0, st>>>(a); break;
    default: return SKKR_ERR_UNSUPPORTED;
  }
  HIPCHK(hipGetLastError());
  return SKKR_OK;
}

// algorithmic work of an inversion launch: 8 n^3 flops per matrix; the
// matrix (and each subtracted term) read, the inverse written
static void inv_end(skkr_ctx* c, const InvArgs& a, int nb) {
  const double n = a.n;
  const double terms = 2.0 + (a.sub1 ? 1.0 : 0.0) + (a.sub2 ? 1.0 : 0.0);
  dom_end(c, 8.0 * n * n * n * nb, 16.0 * terms * n * n * nb, 1);
}

static int launch_gjinv(skkr_ctx* c, const InvArgs& a, int nb) {
  if (nb <= 0) return SKKR_OK;
  if (a.n < 1) return SKKR_ERR_ARG;
  if (c->inv_kernel == 0 || c->inv_kernel == 2) {  // blocked MFMA Gauss-Jordan (2: no pivoting in the 16 x 16 blocks)
    dom_begin(c);
    CHK(c->inv_kernel == 0 ? launch_gjinv_blk<true>(a, nb, c->stream) : launch_gjinv_blk<false>(a, nb, c->stream));
    inv_end(c, a, nb);
    c->launches++;
    return SKKR_OK;
  }
  dom_begin(c);
  if (a.n <= 32) {
    gjinv_f64_kernel<8, 8, 4, 4><<<nb, 64, 0, c->stream>>>(a);
  } else if (a.n <= 64) {
    gjinv_f64_kernel<16, 16, 4, 4><<<nb, 256, 0, c->stream>>>(a);
  } else if (a.n <= 128) {
    gjinv_f64_kernel<16, 32, 8, 4><<<nb, 512, 0, c->stream>>>(a);
  } else {
    return SKKR_ERR_UNSUPPORTED;
  }
  HIPCHK(hipGetLastError());
  inv_end(c, a, nb);
  c->launches++;
  return SKKR_OK;
}

// Inverse for any supported n: the register-resident kernel for n <= 128;
// for 128 < n <= 256 a 2x2 block elimination around A11 (128 x 128):
//   B11 = A11^-1 + D S^-1 C, B12 = -D S^-1, B21 = -S^-1 C, B22 = S^-1,
//   C = A21 A11^-1, D = A11^-1 A12, S = A22 - C A12.
// No rows move between the halves; the matrices inverted here have a
// positive-definite anti-Hermitian part, so A11 and S stay invertible.
// Temporarily sets the slices of the context's GEMM launches (restored on
// every return path).
struct SplitScope {
  skkr_ctx* c;
  int saved;
  SplitScope(skkr_ctx* c_, int s) : c(c_), saved(c_->split) { c->split = s; }
  ~SplitScope() { c->split = saved; }
};

static int run_inverse(skkr_ctx* c, const InvArgs& a, int nb) {
  if (nb <= 0) return SKKR_OK;
  SplitScope fp64(c, 0);  // inversions (and their block GEMMs) stay FP64
  if (a.n <= 128) return launch_gjinv(c, a, nb);
  if (a.n > 256 || a.mode != 0 || !c->Asc || nb > 2 * c->invB || a.n != c->n) return SKKR_ERR_UNSUPPORTED;
  const int n = a.n, h1 = 128, h2 = n - 128;
  const z one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  {
    dim3 grid((unsigned)(((long long)n * n + 255) / 256), nb);
    assemble_kernel<<<grid, 256, 0, c->stream>>>(a, c->p_Asc.d);
    HIPCHK(hipGetLastError());
    const long long off[4] = {0, h1, (long long)h1 * n, h1 + (long long)h1 * n};
    for (int v = 0; v < 4; ++v)
      ptr_offset_kernel<<<(nb + 255) / 256, 256, 0, c->stream>>>(nb, a.out, off[v], c->dv_out[v]);
    HIPCHK(hipGetLastError());
    c->launches += 5;
  }
  z** B11 = c->dv_out[0];
  z** B21 = c->dv_out[1];
  z** B12 = c->dv_out[2];
  z** B22 = c->dv_out[3];
  InvArgs i1;
  memset(&i1, 0, sizeof(i1));
  i1.n = h1;
  i1.lk = h1;
  i1.ld = n;
  i1.ldo = n;
  i1.base = c->p_Asc.d;
  i1.out = B11;
  i1.active = a.active;
  i1.status = a.status;
  CHK(launch_gjinv(c, i1, nb));
  CHK(launch_zgemm(c, h2, h1, h1, c->p_A21.d, n, B11, n, c->p_Csc.d, h2, one, zero, nb, a.active, nullptr));
  CHK(launch_zgemm(c, h2, h2, h1, c->p_Csc.d, h2, c->p_A12.d, n, c->p_A22.d, n, mone, one, nb, a.active, nullptr));
  InvArgs i2 = i1;
  i2.n = h2;
  i2.lk = h2;
  i2.base = c->p_A22.d;
  i2.out = B22;
  CHK(launch_gjinv(c, i2, nb));
  CHK(launch_zgemm(c, h1, h2, h1, B11, n, c->p_A12.d, n, c->p_Dsc.d, h1, one, zero, nb, a.active, nullptr));
  CHK(launch_zgemm(c, h1, h2, h2, c->p_Dsc.d, h1, B22, n, B12, n, mone, zero, nb, a.active, nullptr));
  CHK(launch_zgemm(c, h2, h1, h2, B22, n, c->p_Csc.d, h2, B21, n, mone, zero, nb, a.active, nullptr));
  CHK(launch_zgemm(c, h1, h1, h2, B12, n, c->p_Csc.d, h2, B11, n, mone, one, nb, a.active, nullptr));
  return SKKR_OK;
}

static InvArgs inv_args(skkr_ctx* c) {
  InvArgs a;
  memset(&a, 0, sizeof(a));
  a.n = c->n;
  a.lk = c->cfg.lk;
  return a;
}

static int launch_fourier(skkr_ctx* c, const z* GR, double sgn, z* out, int nk) {
  constexpr int KB = 32;
  dim3 grid((c->nn + 255) / 256, (nk + KB - 1) / KB);
  if (c->nR <= 16)
    fourier_kernel<16, KB><<<grid, 256, 0, c->stream>>>(c->nn, c->nR, GR, c->dR, c->dk, nk, sgn, -1.0, out, c->nn);
  else if (c->nR <= 32)
    fourier_kernel<32, KB><<<grid, 256, 0, c->stream>>>(c->nn, c->nR, GR, c->dR, c->dk, nk, sgn, -1.0, out, c->nn);
  else
    return SKKR_ERR_UNSUPPORTED;
  HIPCHK(hipGetLastError());
  c->launches++;
  return SKKR_OK;
}

// pl blocks of nk (<= B) k-points already in c->dk
static int run_fourier(skkr_ctx* c, int nk) {
  CHK(launch_fourier(c, c->dGRi, -1.0, c->M00, nk));
  CHK(launch_fourier(c, c->dGRo, -1.0, c->M01, nk));
  CHK(launch_fourier(c, c->dGRb, +1.0, c->M10, nk));
  return SKKR_OK;
}

// Decimation iterations first .. last of all 2B chains (surfgf.f:47-96): per
// iteration y = eps^-1, [ay; by] = [alpha; beta] y, then [ay; by] beta and
// [ay; by] alpha with the updates of eps, x and the next [alpha; beta] fused
// into the GEMM epilogues, and the per-chain stop test.  cur: the buffer of
// [alpha; beta] that iteration `first` reads.
// Whether the decimation products of the current digit count switch, per
// chain, to split_dec_late digits late in the iteration (the products then
// add a correction of relative size ~|alpha|^2 |y| / |eps| to eps and x).
// Each product is then two launches from iteration 3 on; below n = 64 the
// second launch costs more than the fewer digits save (n = 36, 512 chains:
// the decimation 1.88 -> 1.96 ms per energy).
constexpr int kDecLateMinN = 64;
static bool dec_late_on(const skkr_ctx* c, int s) {
  return s > 0 && c->split_dec_late > 0 && c->split_dec_late < s && c->dec_late_log2 > 0 && c->act_hi &&
         c->n >= kDecLateMinN;
}
// Whether a chain's late decimation iterations keep eps^-1 of the iteration
// before: once |alpha'| <= 2^-dec_keep_y_log2 of its first value, eps changed
// by a relative ~u in the last iteration and changes by ~u^2 in the next, so
// the stale inverse puts an error ~u^3 into eps and x (mixed modes only).
static bool dec_keep_y_on(const skkr_ctx* c, int s) {
  return s > 0 && c->dec_keep_y_log2 > 0 && c->act_inv;
}

static int dec_iterations(skkr_ctx* c, int first, int last, int cur, int tiles) {
  const int n = c->n, B2 = 2 * c->B;
  const z one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  const bool late = dec_late_on(c, c->split);
  const bool keep_y = dec_keep_y_on(c, c->split);
  const int* act = late ? c->act_hi : c->active;
  // A product for the chains at the full digit count, then, from iteration 3
  // (the first that a chain can run at the late count), for those at the
  // late count; each chain is in exactly one of the two launches.
  auto product = [&](int iter, int m, z** A, int lda, z** B, int ldb, z** C, int ldc,
                     const EpiArgs* e) -> int {
    CHK(launch_zgemm(c, m, n, n, A, lda, B, ldb, C, ldc, one, zero, B2, act, e));
    if (late && iter >= 3) {
      SplitScope sl(c, c->split_dec_late);
      c->dom_nominal_off = true;
      const int r = launch_zgemm(c, m, n, n, A, lda, B, ldb, C, ldc, one, zero, B2, c->act_lo, e);
      c->dom_nominal_off = false;
      CHK(r);
    }
    return SKKR_OK;
  };
  for (int iter = first; iter <= last; ++iter) {
    InvArgs ia = inv_args(c);
    ia.base = c->p_eps.d;
    ia.out = c->p_Yinv.d;
    ia.active = keep_y ? c->act_inv : c->active;
    ia.status = c->status;
    CHK(run_inverse(c, ia, B2));
    CHK(product(iter, 2 * n, c->p_AB[cur].d, 2 * n, c->p_Yinv.d, n, c->p_AYBY.d, 2 * n, nullptr));
    EpiArgs e;
    memset(&e, 0, sizeof(e));
    e.n = n;
    e.eps = c->p_eps.d;
    e.x = c->p_x.d;
    e.abn = c->p_AB[cur ^ 1].d;
    e.norm = c->dnorm;
    e.tiles = tiles;
    e.mode = EPI_DEC_BETA;
    CHK(product(iter, 2 * n, c->p_AYBY.d, 2 * n, c->p_beta[cur].d, 2 * n, nullptr, 2 * n, &e));
    e.mode = EPI_DEC_ALPHA;
    CHK(product(iter, 2 * n, c->p_AYBY.d, 2 * n, c->p_alpha[cur].d, 2 * n, nullptr, 2 * n, &e));
    dec_check_kernel<<<(B2 + 255) / 256, 256, 0, c->stream>>>(
        B2, tiles, iter, c->cfg.itermax_l, c->cfg.itermax_r, c->B, c->cfg.errmax, c->dnorm,
        c->active, c->iters, c->status, c->nactive + iter,
        late ? std::ldexp(1.0, -c->dec_late_log2) : 0.0, c->nref, late ? c->act_hi : nullptr,
        c->act_lo, c->nlate, keep_y ? std::ldexp(1.0, -c->dec_keep_y_log2) : 0.0,
        keep_y ? c->act_inv : nullptr, c->nkeep);
    HIPCHK(hipGetLastError());
    c->launches++;
    cur ^= 1;
  }
  return SKKR_OK;
}

// One round of decimation iterations: replayed from a captured HIP graph
// (one per first / last iteration and product kind; the kernels and their
// arguments are those of the launches) unless per-launch timing is asked for.
static int dec_round(skkr_ctx* c, int first, int last, int cur, int tiles) {
  if (!c->graphs || c->capturing || c->timing_detail) return dec_iterations(c, first, last, cur, tiles);
  const long long key = (((long long)first * 256 + last) * 16 + c->split) * 2 + c->split_i8;
  auto it = c->dec_graphs.find(key);
  if (it == c->dec_graphs.end()) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    const long long l0 = c->launches;
    const int r = dec_iterations(c, first, last, cur, tiles);
    c->launches = l0;
    c->capturing = false;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != SKKR_OK) {
      if (g) hipGraphDestroy(g);
      return r;
    }
    HIPCHK(e);
    hipGraphExec_t ge = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIPCHK(ei);
    it = c->dec_graphs.emplace(key, ge).first;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
  c->launches++;  // one replay counts as one launch
  return SKKR_OK;
}

// Decimation of chains j in [0, 2B) whose active flag the caller set; chains
// use p_ml/p_mr/p_m0/p_tl.  Per iteration (surfgf.f:47-96): y = eps^-1,
// [ay; by] = [alpha; beta] y, then [ay; by] beta and [ay; by] alpha with the
// updates of eps, x and the next [alpha; beta] fused into the GEMM epilogues.
// Leaves g_s = x^-1 in Yinv and the self-energy ml g_s mr in p_sigma when
// want_sigma (ml g_s mt with the outer factors mt when given).
static int run_decimation(skkr_ctx* c, bool add_tinv, bool want_sigma, const PtrArr* mt = nullptr,
                          const PtrArr* lt = nullptr) {
  const int n = c->n, B2 = 2 * c->B;
  dec_init_kernel<<<B2, 256, 0, c->stream>>>(n, c->cfg.lk, c->p_ml.d, c->p_mr.d, c->p_m0.d,
                                             add_tinv ? c->p_tl.d : nullptr, c->p_AB[0].d,
                                             c->p_eps.d, c->p_x.d, c->active,
                                             dec_late_on(c, c->split) ? c->act_hi : nullptr,
                                             c->act_lo, c->nlate,
                                             dec_keep_y_on(c, c->split) ? c->act_inv : nullptr,
                                             c->nkeep);
  HIPCHK(hipGetLastError());
  // the chains started (restored for the final inversion and products)
  HIPCHK(hipMemcpyAsync(c->active0, c->active, B2 * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
  const int itermax = std::max(c->cfg.itermax_l, c->cfg.itermax_r);
  HIPCHK(hipMemsetAsync(c->nactive, 0, sizeof(int) * (itermax + 2), c->stream));
  const z one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  int tm, tiles;
  gemm_tile(c, 2 * n, n, &tm, &tiles);
  if (c->split > 0) tiles = split_tiles(c, 2 * n, n);  // work-groups of the split kernel
  int cur = 0;
  c->dec_rounds = 0;
  // Iterations are launched in rounds without waiting: the first round as many
  // as the previous call needed, then two at a time; the host reads the
  // active count once per round.  Iterations after a chain has stopped cost
  // only the early exit of its work-groups (every kernel tests the active
  // flag), so the results do not depend on the round lengths.
  int iter = 1, round = std::max(2, std::min(c->dec_hint, itermax + 1));
  while (iter <= itermax + 1) {
    const int last = std::min(itermax + 1, iter + round - 1);
    CHK(dec_round(c, iter, last, cur, tiles));
    cur ^= (last - iter + 1) & 1;
    iter = last + 1;
    c->dec_rounds = last;
    int nact = 0;
    HIPCHK(hipMemcpyAsync(&nact, c->nactive + (iter - 1), sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (nact == 0) break;
    round = 2;
  }
  // the next call starts with one round of the iterations this one needed:
  // the first iteration after which no chain was left
  {
    int cnt[66] = {0};
    const int nread = std::min(iter, 66);
    HIPCHK(hipMemcpyAsync(cnt, c->nactive, sizeof(int) * nread, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int used = iter - 1;
    for (int j = 1; j < nread; ++j)
      if (cnt[j] == 0) {
        used = j;
        break;
      }
    c->dec_hint = used;
  }
  HIPCHK(hipMemcpyAsync(c->active, c->active0, B2 * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
  // g_s = x^-1 for every chain that was started
  InvArgs ia = inv_args(c);
  ia.base = c->p_x.d;
  ia.out = c->p_Yinv.d;
  ia.active = c->active;
  ia.status = c->status;
  CHK(run_inverse(c, ia, B2));
  if (want_sigma) {
    // T = ml g_s (into AYBY as n x n), sigma = T mr (lt / mt in place of ml / mr)
    CHK(launch_zgemm(c, n, n, n, lt ? lt->d : c->p_ml.d, n, c->p_Yinv.d, n, c->p_AYBY.d, n, one, zero,
                     B2, c->active));
    CHK(launch_zgemm(c, n, n, n, c->p_AYBY.d, n, mt ? mt->d : c->p_mr.d, n, c->p_sigma.d, n, one, zero, B2,
                     c->active));
  }
  return SKKR_OK;
}

// PL-diagonal blocks of tau for nb k-points: Y_1 in Ycur, X_P in Xc, blocks
// in M00/M01/M10.  Recursive Green's-function form of the block-tridiagonal
// inverse (same tau_pp as the Godfrin sweeps of tau2d.f:53-149):
//   forward  g_p = (T_p - Y_p)^-1, Z_p = g_p M01, Y_{p+1} = M10 Z_p;
//   last     G_P = (T_P - X_P - Y_P)^-1;
//   backward G_p = g_p + Z_p (G_{p+1} (M10 g_p)).
// One inversion per PL instead of three; the rest are GEMMs.
static int launch_extract(skkr_ctx* c, int p, int nb) {
  extract_sites_kernel<<<nb, 256, 0, c->stream>>>(c->n, c->cfg.lk, c->p_G[p].d, c->p_tau[p].d,
                                                  c->p_tr[p].d);
  HIPCHK(hipGetLastError());
  c->launches++;
  return SKKR_OK;
}

static int sweep_launches(skkr_ctx* c, int nb) {
  const int n = c->n, P = c->cfg.n_pl, s = c->cfg.sites_per_pl, lk = c->cfg.lk;
  const size_t tblk = (size_t)s * lk * lk;
  const z one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  // blocks of PL p: one set for a bulk-like stack, per PL for a layered one
  // (A_pp = M00_p + t^-1_p, A_p,p+1 = M01_p, A_p+1,p = M10_p+1)
  auto m00 = [&](int p) { return c->layered ? c->p_M00pl[p].d : c->p_M00.d; };
  auto m01 = [&](int p) { return c->layered ? c->p_M01pl[p].d : c->p_M01.d; };
  auto m10 = [&](int p) { return c->layered ? c->p_M10pl[p].d : c->p_M10.d; };
  for (int p = 0; p < P - 1; ++p) {
    InvArgs ia = inv_args(c);
    ia.base = m00(p);
    ia.tinv = c->dTint + p * tblk;
    ia.sub1 = c->p_Y.d;
    ia.out = c->p_G[p].d;
    ia.status = c->status_sw;
    CHK(run_inverse(c, ia, nb));
    CHK(launch_zgemm(c, n, n, n, c->p_G[p].d, n, m01(p), n, c->p_Zs[p].d, n, one, zero, nb,
                     nullptr));
    CHK(launch_zgemm(c, n, n, n, m10(p + 1), n, c->p_Zs[p].d, n, c->p_Y.d, n, one, zero, nb,
                     nullptr));
  }
  {
    InvArgs ia = inv_args(c);
    ia.base = m00(P - 1);
    ia.tinv = c->dTint + (P - 1) * tblk;
    ia.sub1 = c->p_X.d;
    ia.sub2 = c->p_Y.d;
    ia.out = c->p_G[P - 1].d;
    ia.status = c->status_sw;
    CHK(run_inverse(c, ia, nb));
    CHK(launch_extract(c, P - 1, nb));
  }
  for (int p = P - 2; p >= 0; --p) {
    z** W = c->pairs_on ? c->p_Ws[p].d : c->p_W.d;
    CHK(launch_zgemm(c, n, n, n, m10(p + 1), n, c->p_G[p].d, n, W, n, one, zero, nb,
                     nullptr));
    CHK(launch_zgemm(c, n, n, n, c->p_G[p + 1].d, n, W, n, c->p_Z.d, n, one, zero, nb,
                     nullptr));
    CHK(launch_zgemm(c, n, n, n, c->p_Zs[p].d, n, c->p_Z.d, n, c->p_G[p].d, n, one, one, nb,
                     nullptr));
    CHK(launch_extract(c, p, nb));
  }
  return SKKR_OK;
}

// The sweeps of a batch, optionally as a captured HIP graph (one per nb; all
// buffers and pointer arrays are fixed for the context's batch size).
static int run_sweeps(skkr_ctx* c, int nb) {
  if (!c->graphs || c->pairs_on || c->layered) return sweep_launches(c, nb);
  const int key = 8 * nb + c->split + 1000000 * c->split_i8;
  auto it = c->sweep_graphs.find(key);
  if (it == c->sweep_graphs.end()) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    const int r = sweep_launches(c, nb);
    c->capturing = false;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != SKKR_OK) {
      if (g) hipGraphDestroy(g);
      return r;
    }
    HIPCHK(e);
    hipGraphExec_t ge = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIPCHK(ei);
    it = c->sweep_graphs.emplace(key, ge).first;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
  c->launches++;  // one replay counts as one launch
  return SKKR_OK;
}

static int status_code(int bits) {
  if (bits & 8) return SKKR_ERR_STR_NOCONV;
  if (bits & 4) return SKKR_ERR_SGF_ITERMAX;
  if (bits & 2) return SKKR_ERR_SGF_DIVERGED;
  if (bits & 1) return SKKR_ERR_SINGULAR;
  return SKKR_OK;
}

// ------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------
extern "C" {

const char* skkr_error_string(int code) {
  switch (code) {
    case SKKR_OK: return "ok";
    case SKKR_ERR_ARG: return "invalid argument";
    case SKKR_ERR_HIP: return "HIP runtime error";
    case SKKR_ERR_NOMEM: return "device out of memory";
    case SKKR_ERR_SINGULAR: return "singular pivot";
    case SKKR_ERR_SGF_DIVERGED: return "decimation diverged";
    case SKKR_ERR_SGF_ITERMAX: return "decimation exceeded itermax";
    case SKKR_ERR_UNSUPPORTED: return "unsupported size or mode";
    case SKKR_ERR_STR_NOCONV: return "structure-constant shell sum not converged";
    case SKKR_ERR_TOO_LARGE: return "problem larger than the static work space";
  }
  return "unknown error";
}

int skkr_create(skkr_ctx** out, const skkr_config* cfg) {
  if (!out || !cfg) return SKKR_ERR_ARG;
  *out = nullptr;
  if (cfg->lk <= 0 || cfg->sites_per_pl <= 0 || cfg->n_pl <= 0 || cfg->errmax <= 0 ||
      cfg->itermax_l < 1 || cfg->itermax_r < 1)
    return SKKR_ERR_ARG;
  if (cfg->precision != SKKR_FP64 && cfg->precision != SKKR_MIXED_BF16 && cfg->precision != SKKR_MIXED_INT8)
    return SKKR_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0)
    return SKKR_ERR_HIP;
  HIPCHK(hipSetDevice(cfg->device));
  skkr_ctx* c = new skkr_ctx();
  c->cfg = *cfg;
  if (cfg->precision == SKKR_MIXED_INT8) {
    // 6 digits (2^-47 of the line maximum per operand value) in the
    // decimation and in the sweeps near the real axis, 5 (2^-39) in the
    // sweeps above fp64_dec_below_imE: the sweeps need ~36-40 bits, the
    // decimation ~45 (config 1: tau within 3.7e-10 of FP64)
    c->split_i8 = 1;
    c->split_dec = c->split_sw_low = 6;
    c->split_sw = 5;
    // far from the real axis the decimation converges in fewer iterations and
    // amplifies less: 5 digits there (config-1 system: tau within 7.4e-10 of
    // FP64 at Im E = 0.1, 1.5e-10 at 0.2)
    c->split_dec_high = 5;
    // a chain's decimation products once |alpha'| is below 2^-6 of its value
    // after iteration 1 add corrections ~|alpha|^2 |y| small against eps and
    // x: 4 digits there (config-1 system at Im E 0.02 / 0.04 / 0.15: tau
    // unchanged at 3.7e-10 / 1.8e-10 / 1.0e-10 of FP64, 1-3% less time)
    c->split_dec_late = 4;
    c->dec_late_log2 = 6;
    // and from 2^-20 on it keeps eps^-1 of the iteration before (the same
    // three energies: tau within 3.7e-10 / 1.8e-10 / 1.0e-10, 0.4-1.5% less
    // time)
    c->dec_keep_y_log2 = 20;
  }
  if (const char* e = getenv("SKKR_INV_KERNEL"))
    c->inv_kernel = (strcmp(e, "rowwise") == 0) ? 1 : (strcmp(e, "nopivot") == 0) ? 2 : 0;
  if (const char* e = getenv("SKKR_GEMM_3M")) c->gemm_3m = (atoi(e) != 0) ? 1 : 0;
  if (const char* e = getenv("SKKR_GEMM_TILE")) {
    const int t = atoi(e);
    if (t == 32 || t == 64 || t == 96 || t == 128) c->gemm_tile_force = t;
  }
  if (const char* e = getenv("SKKR_INT8_TILE")) {
    const int t = atoi(e);
    if (t == 64 || t == 128) c->i8_tile = t;
  }
  c->n = cfg->lk * cfg->sites_per_pl;
  c->nn = c->n * c->n;
  c->ns = cfg->sites_per_pl;
  if (c->n > 256) {
    delete c;
    return SKKR_ERR_UNSUPPORTED;
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return SKKR_ERR_HIP;
  }
  const size_t tsz = (size_t)cfg->lk * cfg->lk;
  int r = SKKR_OK;
  if ((r = dalloc_t(c, &c->dTint, tsz * cfg->n_pl * cfg->sites_per_pl)) != SKKR_OK ||
      (r = dalloc_t(c, &c->dTl, tsz * cfg->sites_per_pl)) != SKKR_OK ||
      (r = dalloc_t(c, &c->dTr, tsz * cfg->sites_per_pl)) != SKKR_OK ||
      (r = dalloc_t(c, &c->taubz, tsz * cfg->n_pl * cfg->sites_per_pl)) != SKKR_OK) {
    skkr_destroy(c);
    return r;
  }
  *out = c;
  return SKKR_OK;
}

void skkr_destroy(skkr_ctx* c) {
  if (!c) return;
  hipSetDevice(c->cfg.device);
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (void* p : c->allocs) hipFree(p);
  for (void* p : c->batch_allocs) hipFree(p);
  if (c->stream) hipStreamDestroy(c->stream);
  for (hipEvent_t e : c->evpool) hipEventDestroy(e);
  for (hipEvent_t e : c->phase_ev)
    if (e) hipEventDestroy(e);
  delete c;
}

int skkr_set_structure(skkr_ctx* c, int nR, const double* R, const double* GRi,
                       const double* GRo) {
  if (!c || nR <= 0 || nR > 32 || !R || !GRi || !GRo) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  if (c->have_structure && nR != c->nR) return SKKR_ERR_ARG;
  const int n = c->n;
  const size_t nn = c->nn;
  if (!c->have_structure) {
    c->nR = nR;
    CHK(dalloc_t(c, &c->dR, 2 * (size_t)nR));
    CHK(dalloc_t(c, &c->dGRi, nR * nn));
    CHK(dalloc_t(c, &c->dGRo, nR * nn));
    CHK(dalloc_t(c, &c->dGRb, nR * nn));
  }
  // back coupling: G_inter(R)^dagger for every R, summed with e^{+ik.R}
  std::vector<double> back(2 * nR * nn);
  for (int r = 0; r < nR; ++r)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        size_t src = (size_t)r * nn + i + (size_t)j * n;
        size_t dst = (size_t)r * nn + j + (size_t)i * n;
        back[2 * dst] = GRo[2 * src];
        back[2 * dst + 1] = -GRo[2 * src + 1];
      }
  HIPCHK(hipMemcpy(c->dR, R, 2 * sizeof(double) * nR, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dGRi, GRi, sizeof(z) * nR * nn, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dGRo, GRo, sizeof(z) * nR * nn, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dGRb, back.data(), sizeof(z) * nR * nn, hipMemcpyHostToDevice));
  c->have_structure = true;
  return SKKR_OK;
}

int skkr_set_tinv(skkr_ctx* c, const double* ti, const double* tl, const double* tr) {
  if (!c || !ti || !tl || !tr) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  const size_t tsz = (size_t)c->cfg.lk * c->cfg.lk;
  HIPCHK(hipMemcpy(c->dTint, ti, sizeof(z) * tsz * c->cfg.n_pl * c->ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dTl, tl, sizeof(z) * tsz * c->ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dTr, tr, sizeof(z) * tsz * c->ns, hipMemcpyHostToDevice));
  c->have_tinv = true;
  return SKKR_OK;
}

static int auto_batch(skkr_ctx* c, int nk) {
  int B = c->cfg.batch;
  if (B <= 0) {
    // ~ 28 n^2 + 2P n^2 complex per k (+ 3 n^2 block-inverse scratch for
    // n > 128); keep a batch under ~24 GB
    double per_k = (31.0 + 2.0 * c->cfg.n_pl) * c->nn * 16.0;
    B = (int)std::min(1024.0, std::max(1.0, 24e9 / per_k));
  }
  return std::max(1, std::min(B, nk));
}

// The k-loop of one energy; with phys, also the physical representation:
// tau_bz_phys += phystau of this call's BZ sum, taukdiag = per-k per-l traces.
static int str_blocks(skkr_ctx* c, int nb);

// W_p of every PL for the pair sums
static int ensure_ws(skkr_ctx* c) {
  const int P = c->cfg.n_pl;
  if (c->ws_cap >= c->B || P < 2) return SKKR_OK;
  const size_t nn = c->nn, Bs = c->B;
  CHK(dalloc_bt(c, &c->Ws, (size_t)(P - 1) * Bs * nn));
  c->p_Ws.assign(P - 1, PtrArr());
  for (int p = 0; p < P - 1; ++p) {
    std::vector<z*> v(Bs);
    for (size_t j = 0; j < Bs; ++j) v[j] = c->Ws + ((size_t)p * Bs + j) * nn;
    CHK(make_ptrs(c, &c->p_Ws[p], v, true));
  }
  c->ws_cap = c->B;
  return SKKR_OK;
}

// Device buffers of the per-PL blocks of a layered stack (P x B matrices each).
static int ensure_layered(skkr_ctx* c) {
  const int P = c->cfg.n_pl;
  if (c->pl_cap >= c->B) return SKKR_OK;
  const size_t nn = c->nn, Bs = c->B;
  CHK(dalloc_bt(c, &c->M00pl, (size_t)P * Bs * nn));
  CHK(dalloc_bt(c, &c->M01pl, (size_t)P * Bs * nn));
  CHK(dalloc_bt(c, &c->M10pl, (size_t)P * Bs * nn));
  c->p_M00pl.assign(P, PtrArr());
  c->p_M01pl.assign(P, PtrArr());
  c->p_M10pl.assign(P, PtrArr());
  for (int p = 0; p < P; ++p) {
    std::vector<z*> v(Bs);
    for (size_t j = 0; j < Bs; ++j) v[j] = c->M00pl + ((size_t)p * Bs + j) * nn;
    CHK(make_ptrs(c, &c->p_M00pl[p], v, true));
    for (size_t j = 0; j < Bs; ++j) v[j] = c->M01pl + ((size_t)p * Bs + j) * nn;
    CHK(make_ptrs(c, &c->p_M01pl[p], v, true));
    for (size_t j = 0; j < Bs; ++j) v[j] = c->M10pl + ((size_t)p * Bs + j) * nn;
    CHK(make_ptrs(c, &c->p_M10pl[p], v, true));
  }
  c->pl_cap = c->B;
  return SKKR_OK;
}

// Off-diagonal BZ sums of the site pairs for the nb k-points of a batch
// (after the sweeps).  Blocks of tau between PLs p and q of every k:
//   p < q: tau_pq = (-Z_p)(-Z_{p+1}) ... (-Z_{q-1}) G_q,
//   p > q: tau_pq = G_p (-W_{p-1}) (-W_{p-2}) ... (-W_q),
// with Z_t = g_t M01 and W_t = M10 g_t of the sweeps -- the same blocks as
// tau2dgod's recursion (clu/tau2d_god.f) -- built right-to-left (p < q) or
// left-to-right (p > q) in the free W / Z scratch; then each pair's lk x lk
// block is summed over k with its phase.
static int pair_sums(skkr_ctx* c, int nb) {
  const int n = c->n, lk = c->cfg.lk, s = c->ns, np = (int)c->pair_li.size();
  const z mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  std::vector<std::pair<int, int>> done;
  for (int q2 = 0; q2 < np; ++q2)
    for (int dir = 0; dir < 2; ++dir) {
      const int la = dir ? c->pair_lj[q2] : c->pair_li[q2], lb = dir ? c->pair_li[q2] : c->pair_lj[q2];
      const int p = la / s, q = lb / s;
      std::pair<int, int> key(p, q);
      bool seen = false;
      for (auto& d : done) seen = seen || (d == key);
      if (seen) continue;
      done.push_back(key);
      z** T = c->p_G[p].d;
      if (p < q) {
        z** buf[2] = {c->p_W.d, c->p_Z.d};
        z** cur = c->p_G[q].d;
        for (int t = q - 1, h = 0; t >= p; --t, h ^= 1) {
          CHK(launch_zgemm(c, n, n, n, c->p_Zs[t].d, n, cur, n, buf[h], n, mone, zero, nb, nullptr));
          cur = buf[h];
        }
        T = cur;
      } else if (p > q) {
        z** buf[2] = {c->p_W.d, c->p_Z.d};
        z** cur = c->p_G[p].d;
        for (int t = p - 1, h = 0; t >= q; --t, h ^= 1) {
          CHK(launch_zgemm(c, n, n, n, cur, n, c->p_Ws[t].d, n, buf[h], n, mone, zero, nb, nullptr));
          cur = buf[h];
        }
        T = cur;
      }
      // every pair (either direction) whose blocks lie between p and q
      for (int r = 0; r < np; ++r)
        for (int d2 = 0; d2 < 2; ++d2) {
          const int xa = d2 ? c->pair_lj[r] : c->pair_li[r], xb = d2 ? c->pair_li[r] : c->pair_lj[r];
          if (xa / s != p || xb / s != q) continue;
          z* acc = (d2 ? c->dTji : c->dTij) + (size_t)r * lk * lk;
          pair_accumulate_kernel<<<(lk * lk + 255) / 256, 256, 0, c->stream>>>(
              lk, n, nb, (const double2* const*)T, xa % s, xb % s, c->pair_rx[r], c->pair_ry[r], d2 ? -1.0 : 1.0,
              c->dk, c->dw, acc);
          HIPCHK(hipGetLastError());
          c->launches++;
        }
    }
  return SKKR_OK;
}

// outer self-energy factors of the chains (layer producer only)
static int ensure_mt(skkr_ctx* c) {
  if (c->mt_cap >= c->B) return SKKR_OK;
  const size_t nn = c->nn, Bs = c->B;
  CHK(dalloc_bt(c, &c->MtL, Bs * nn));
  CHK(dalloc_bt(c, &c->MtR, Bs * nn));
  std::vector<z*> v(2 * Bs);
  for (size_t j = 0; j < 2 * Bs; ++j) v[j] = j < Bs ? c->MtL + j * nn : c->MtR + (j - Bs) * nn;
  CHK(make_ptrs(c, &c->p_mt, v, true));
  c->mt_cap = c->B;
  return SKKR_OK;
}

// The point-group sum of an irreducible-wedge BZ integration,
// nrsp-scf/cpacoord.f:294-299: every k adds (w_k/ng) dg_g tau_k dg1_g for
// every group element.  Applied once to the BZ sum on the device (equal by
// linearity): per element two batched products over the sites,
// tmp = tau dg1_g and out += (1/ng) dg_g tmp; then tau <- out.
static int apply_point_group(skkr_ctx* c) {
  const int lk = c->cfg.lk, S = (int)(c->cfg.n_pl * c->ns);
  const z zero = make_double2(0.0, 0.0), one = make_double2(1.0, 0.0);
  const z scale = make_double2(1.0 / c->ng_sym, 0.0);
  const int td = c->timing_detail;
  c->timing_detail = 0;  // not part of the dominant-kernel statistics
  int r = SKKR_OK;
  for (int g = 0; g < c->ng_sym && r == SKKR_OK; ++g) {
    r = launch_zgemm(c, lk, lk, lk, c->p_symT.d, lk, c->p_symD1[g].d, lk, c->p_symTmp.d, lk, one, zero, S,
                     nullptr);
    if (r == SKKR_OK)
      r = launch_zgemm(c, lk, lk, lk, c->p_symD[g].d, lk, c->p_symTmp.d, lk, c->p_symOut.d, lk, scale,
                       g == 0 ? zero : one, S, nullptr);
  }
  c->timing_detail = td;
  CHK(r);
  HIPCHK(hipMemcpyAsync(c->taubz, c->dSymOut, sizeof(z) * (size_t)S * lk * lk, hipMemcpyDeviceToDevice,
                        c->stream));
  return SKKR_OK;
}

static int tau_bz_run(skkr_ctx* c, int nk, const double* k, const double* w, double* tau_bz,
                      double* tr_tau_k, int* iters_k, int* status_k, skkr_stats* st,
                      bool phys, double* tau_bz_phys, double* taukdiag) {
  if (!c || nk < 0 || (nk > 0 && (!k || !w)) || !tau_bz) return SKKR_ERR_ARG;
  if (!(c->have_structure || c->str) || !c->have_tinv) return SKKR_ERR_ARG;
  if (phys && (!c->have_screening || !tau_bz_phys)) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  auto t0 = std::chrono::steady_clock::now();
  CHK(ensure_batch(c, auto_batch(c, std::max(nk, 1))));
  if (c->str) CHK(ensure_mt(c));
  if (c->pairs_on) CHK(ensure_ws(c));
  // the chain layout on the device (left chains 0..B-1, right chains B..2B-1)
  // follows the allocated batch, which an earlier call may have made larger
  const int B = c->B;
  const int P = c->cfg.n_pl, lk = c->cfg.lk, ns = c->ns;
  const size_t sites = (size_t)P * ns, tsz = sites * lk * lk;
  c->timing_detail = st ? st->timing_detail : 0;
  c->dom_ms = 0;
  c->dom_flops = 0;
  c->dom_launches = 0;
  c->dom_bytes = 0;
  c->inv_ms = 0;
  c->inv_flops = 0;
  c->inv_launches = 0;
  c->inv_bytes = 0;
  c->evused = 0;
  c->evflops.clear();
  c->evbytes.clear();
  c->evclass.clear();
  c->launches = 0;
  c->dom_flops_exec = 0;
  c->redo_batches = 0;
  // mixed precision (SKKR_MIXED_BF16): split-BF16 products in the decimation
  // and the sweeps; the decimation stays FP64 for Im E below the threshold
  int s_dec = 0, s_sw = 0;
  if (c->cfg.precision == SKKR_MIXED_BF16 || c->cfg.precision == SKKR_MIXED_INT8) {
    const bool near_axis = c->e_im < c->cfg.fp64_dec_below_imE;
    const bool far = c->e_im >= kDecHighImE && c->split_dec_high > 0;
    s_dec = near_axis ? 0 : far ? c->split_dec_high : c->split_dec;
    s_sw = near_axis ? c->split_sw_low : c->split_sw;
  }
  HIPCHK(hipMemsetAsync(c->taubz, 0, sizeof(z) * tsz, c->stream));
  CHK(ensure_phase_events(c));
  hipEvent_t* ev = c->phase_ev;
  double ms_f = 0, ms_d = 0, ms_s = 0, ms_a = 0;
  double sum_l = 0, sum_r = 0;
  int maxit = 0, nbatches = 0;
  int rc = SKKR_OK;
  std::vector<int> h_it(2 * B), h_st(2 * B), h_sw(B), act(2 * B), h_nl(2 * B, 0), h_nk(2 * B, 0);
  std::vector<double> h_tr(2 * sites * B);
  std::vector<double> h_ph;
  // the CPA passes keep every k resident in one batch
  if (c->cpa_mode && nk > B) return SKKR_ERR_ARG;
  const bool reuse = c->cpa_mode == 2;
  for (int k0 = 0; k0 < nk; k0 += B) {
    const int nb = std::min(B, nk - k0);
    ++nbatches;
    if (!reuse) {
      HIPCHK(hipMemcpyAsync(c->dk, k + 2 * (size_t)k0, 2 * sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->dw, w + k0, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
      for (int j = 0; j < 2 * B; ++j) act[j] = (j % B) < nb ? 1 : 0;
      HIPCHK(hipMemcpyAsync(c->active, act.data(), 2 * B * sizeof(int), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemsetAsync(c->iters, 0, 2 * B * sizeof(int), c->stream));
      HIPCHK(hipMemsetAsync(c->status, 0, 2 * B * sizeof(int), c->stream));
    }
    HIPCHK(hipMemsetAsync(c->status_sw, 0, B * sizeof(int), c->stream));
    hipEventRecord(ev[0], c->stream);
    if (!reuse) {
      if (c->str)
        CHK(str_blocks(c, nb));
      else
        CHK(run_fourier(c, nb));
    }
    hipEventRecord(ev[1], c->stream);
    if (!reuse) {
      SplitScope sd(c, s_dec);
      CHK(run_decimation(c, true, true, c->str ? &c->p_mt : nullptr));
      if (c->cpa_mode == 1)
        HIPCHK(hipMemcpyAsync(c->Y1s, c->Ycur, sizeof(z) * c->nn * nb, hipMemcpyDeviceToDevice, c->stream));
    } else {
      HIPCHK(hipMemcpyAsync(c->Ycur, c->Y1s, sizeof(z) * c->nn * nb, hipMemcpyDeviceToDevice, c->stream));
    }
    hipEventRecord(ev[2], c->stream);
    {
      SplitScope ss(c, s_sw);
      CHK(run_sweeps(c, nb));
    }
    // A blocked inversion whose multipliers grew (GJ_STATUS_REDO: a diagonal
    // block singular or small against its column) runs the batch's
    // decimation and sweeps again with the row-wise kernel, which pivots
    // over all rows.
    {
      HIPCHK(hipMemcpyAsync(h_st.data(), c->status, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(h_sw.data(), c->status_sw, B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      bool redo = false;
      for (int b = 0; b < nb; ++b)
        redo = redo || ((h_st[b] | h_st[B + b] | h_sw[b]) & GJ_STATUS_REDO);
      if (redo) {
        ++c->redo_batches;
        struct InvKernel {
          skkr_ctx* c;
          int saved;
          ~InvKernel() { c->inv_kernel = saved; }
        } keep{c, c->inv_kernel};
        c->inv_kernel = 1;
        HIPCHK(hipMemsetAsync(c->status_sw, 0, B * sizeof(int), c->stream));
        if (!reuse) {
          HIPCHK(hipMemcpyAsync(c->active, act.data(), 2 * B * sizeof(int), hipMemcpyHostToDevice, c->stream));
          HIPCHK(hipMemsetAsync(c->iters, 0, 2 * B * sizeof(int), c->stream));
          HIPCHK(hipMemsetAsync(c->status, 0, 2 * B * sizeof(int), c->stream));
          SplitScope sd(c, s_dec);
          CHK(run_decimation(c, true, true, c->str ? &c->p_mt : nullptr));
          if (c->cpa_mode == 1)
            HIPCHK(hipMemcpyAsync(c->Y1s, c->Ycur, sizeof(z) * c->nn * nb, hipMemcpyDeviceToDevice, c->stream));
        } else {
          HIPCHK(hipMemcpyAsync(c->Ycur, c->Y1s, sizeof(z) * c->nn * nb, hipMemcpyDeviceToDevice, c->stream));
        }
        SplitScope ss(c, s_sw);
        CHK(sweep_launches(c, nb));  // eager: the captured graph holds the blocked kernel
      }
    }
    if (c->pairs_on) CHK(pair_sums(c, nb));
    hipEventRecord(ev[3], c->stream);
    {
      int ne = (int)tsz;
      bz_accumulate_kernel<<<(ne + 255) / 256, 256, 0, c->stream>>>(ne, nb, c->tauk, (long long)tsz, c->dw, c->taubz);
      HIPCHK(hipGetLastError());
      c->launches++;
    }
    if (phys && taukdiag) {
      const int nl = c->nl;
      if (c->phk_cap < B) {
        CHK(dalloc_bt(c, &c->phk, (size_t)B * sites * nl));
        c->phk_cap = B;
      }
      dim3 grid((unsigned)sites, (unsigned)nb);
      if (lk <= 32)
        phys_trace_kernel<32><<<grid, 64, 0, c->stream>>>(lk, nl, (int)sites, c->tauk, (long long)tsz,
                                                          c->dTint, c->dA, c->dLidx, c->phk);
      else
        phys_trace_kernel<64><<<grid, 64, 0, c->stream>>>(lk, nl, (int)sites, c->tauk, (long long)tsz,
                                                          c->dTint, c->dA, c->dLidx, c->phk);
      HIPCHK(hipGetLastError());
      c->launches++;
      h_ph.resize((size_t)2 * nb * sites * nl);
      HIPCHK(hipMemcpyAsync(h_ph.data(), c->phk, sizeof(z) * nb * sites * nl, hipMemcpyDeviceToHost,
                            c->stream));
    }
    hipEventRecord(ev[4], c->stream);
    HIPCHK(hipMemcpyAsync(h_it.data(), c->iters, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(h_st.data(), c->status, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    const bool late_dec = !reuse && dec_late_on(c, s_dec);
    if (c->timing_detail && late_dec)
      HIPCHK(hipMemcpyAsync(h_nl.data(), c->nlate, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    const bool keep_dec = !reuse && dec_keep_y_on(c, s_dec);
    if (c->timing_detail && keep_dec)
      HIPCHK(hipMemcpyAsync(h_nk.data(), c->nkeep, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(h_sw.data(), c->status_sw, B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (tr_tau_k)
      HIPCHK(hipMemcpyAsync(tr_tau_k + 2 * sites * (size_t)k0, c->trk, sizeof(z) * sites * nb,
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dom_collect(c);
    float t;
    hipEventElapsedTime(&t, ev[0], ev[1]); ms_f += t;
    hipEventElapsedTime(&t, ev[1], ev[2]); ms_d += t;
    hipEventElapsedTime(&t, ev[2], ev[3]); ms_s += t;
    hipEventElapsedTime(&t, ev[3], ev[4]); ms_a += t;
    if (c->timing_detail && !reuse && c->n <= 128) {
      // The decimation launches cover all 2B chains; only the active ones do
      // work.  Replace the nominal counts of those launches by the work of
      // the chains that ran: chain j took part in iterations 1..iters_j
      // (three 2n x n x n products and one inversion each), every started
      // chain in the final inversion and the two n^3 self-energy products.
      const double n3 = (double)c->n * c->n * c->n, n2 = (double)c->n * c->n;
      const double B2d = 2.0 * B, rounds = c->dec_rounds;
      double it_sum = 0.0;
      for (int b = 0; b < nb; ++b) it_sum += h_it[b] + h_it[B + b];
      const double started = 2.0 * nb;
      const double corr = (it_sum - rounds * B2d) * 3.0 * 16.0 * n3 + (started - B2d) * 2.0 * 8.0 * n3;
      c->dom_flops += corr;
      if (s_dec > 0) {  // executed split work of the same launches (padded tiles)
        const int tm = c->split_i8 ? c->i8_tile : 64;
        const double pad = (double)((2 * c->n + tm - 1) / tm * tm) * ((c->n + 63) / 64 * 64) *
                           ((c->n + 31) / 32 * 32) / (2.0 * n3);
        c->dom_flops_exec += corr * s_dec * (s_dec + 1) / 2.0 * pad;
        if (late_dec) {  // chain iterations at the late digit count (launches counted no nominal work)
          double late_sum = 0.0;
          for (int b = 0; b < nb; ++b) late_sum += h_nl[b] + h_nl[B + b];
          const int sl = c->split_dec_late;
          c->dom_flops_exec -= late_sum * 3.0 * 16.0 * n3 * (s_dec * (s_dec + 1) - sl * (sl + 1)) / 2.0 * pad;
        }
      }
      c->dom_bytes += (it_sum - rounds * B2d) * 3.0 * 80.0 * n2 + (started - B2d) * 2.0 * 48.0 * n2;
      double kept = 0.0;  // iterations that kept the previous inverse
      for (int b = 0; b < nb; ++b) kept += h_nk[b] + h_nk[B + b];
      c->inv_flops += (it_sum - kept - rounds * B2d) * 8.0 * n3 + (started - B2d) * 8.0 * n3;
      c->inv_bytes += (it_sum - kept - rounds * B2d) * 32.0 * n2 + (started - B2d) * 32.0 * n2;
    }
    for (int b = 0; b < nb; ++b) {
      int il = h_it[b], ir = h_it[B + b];
      sum_l += il;
      sum_r += ir;
      maxit = std::max(maxit, std::max(il, ir));
      int code = status_code(h_st[b] | h_st[B + b] | h_sw[b]);
      if (iters_k) {
        iters_k[2 * (size_t)(k0 + b)] = il;
        iters_k[2 * (size_t)(k0 + b) + 1] = ir;
      }
      if (status_k) status_k[k0 + b] = code;
      if (code != SKKR_OK && rc == SKKR_OK) rc = code;
    }
    if (phys && taukdiag) {
      // taukdiag(ik, l, site), ik fastest (cpacoord.f taukdiag(mkpar,0:lmaxp,mintfc))
      const int nl = c->nl;
      for (int b = 0; b < nb; ++b)
        for (size_t site = 0; site < sites; ++site)
          for (int l = 0; l < nl; ++l) {
            const size_t src = 2 * (((size_t)b * sites + site) * nl + l);
            const size_t dst = 2 * ((size_t)(k0 + b) + (size_t)nk * ((size_t)l + (size_t)nl * site));
            taukdiag[dst] = h_ph[src];
            taukdiag[dst + 1] = h_ph[src + 1];
          }
    }
  }
  if (c->ng_sym > 0) CHK(apply_point_group(c));
  std::vector<double> hbz(2 * tsz);
  HIPCHK(hipMemcpyAsync(hbz.data(), c->taubz, sizeof(z) * tsz, hipMemcpyDeviceToHost, c->stream));
  if (phys) {
    double wsum = 0.0;
    for (int i = 0; i < nk; ++i) wsum += w[i];
    if (lk <= 32)
      phys_full_kernel<32><<<(unsigned)sites, 256, 0, c->stream>>>(lk, c->taubz, c->dTint, c->dA, wsum,
                                                                  c->physx, c->taubz_phys);
    else
      phys_full_kernel<64><<<(unsigned)sites, 256, 0, c->stream>>>(lk, c->taubz, c->dTint, c->dA, wsum,
                                                                  c->physx, c->taubz_phys);
    HIPCHK(hipGetLastError());
    c->launches++;
  }
  std::vector<double> hph(phys ? 2 * tsz : 0);
  if (phys)
    HIPCHK(hipMemcpyAsync(hph.data(), c->taubz_phys, sizeof(z) * tsz, hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (size_t e = 0; e < 2 * tsz; ++e) tau_bz[e] += hbz[e];
  if (phys)
    for (size_t e = 0; e < 2 * tsz; ++e) tau_bz_phys[e] += hph[e];
  if (st) {
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    st->ms_fourier = ms_f;
    st->ms_decimation = ms_d;
    st->ms_sweeps = ms_s;
    st->ms_accumulate = ms_a;
    st->ms_refine = 0;
    st->mean_iter_l = nk ? sum_l / nk : 0;
    st->mean_iter_r = nk ? sum_r / nk : 0;
    st->max_iter = maxit;
    st->n_batches = nbatches;
    st->n_launches = c->launches;
    st->dom_launches = c->dom_launches;
    st->dom_ms = c->dom_ms;
    st->dom_flops = c->dom_flops;
    st->dom_bytes = c->dom_bytes;
    st->dom_flops_exec = c->dom_flops_exec;
    st->inv_launches = c->inv_launches;
    st->inv_ms = c->inv_ms;
    st->inv_flops = c->inv_flops;
    st->inv_bytes = c->inv_bytes;
    st->redo_batches = c->redo_batches;
    snprintf(st->dom_name, sizeof(st->dom_name), "%s",
             (s_dec > 0 && s_sw > 0) ? (c->split_i8 ? (c->i8_tile == 64 ? "zgemm_i8_64_kernel" : "zgemm_i8_kernel") : "zgemm_split_kernel")
                                     : (s_sw > 0 ? "zgemm_f64+split" : "zgemm_f64_kernel"));
  }
  c->timing_detail = 0;
  return rc;
}

int skkr_tau_bz(skkr_ctx* c, double e_re, double e_im, int nk, const double* k,
                const double* w, double* tau_bz, double* tr_tau_k, int* iters_k,
                int* status_k, skkr_stats* st) {
  if (!c) return SKKR_ERR_ARG;
  c->e_re = e_re;
  c->e_im = e_im;
  return tau_bz_run(c, nk, k, w, tau_bz, tr_tau_k, iters_k, status_k, st, false, nullptr,
                    nullptr);
}

int skkr_tau_bz_phys(skkr_ctx* c, double e_re, double e_im, int nk, const double* k,
                     const double* w, double* tau_bz, double* tau_bz_phys, double* taukdiag,
                     int* iters_k, int* status_k, skkr_stats* st) {
  if (!c) return SKKR_ERR_ARG;
  c->e_re = e_re;
  c->e_im = e_im;
  return tau_bz_run(c, nk, k, w, tau_bz, nullptr, iters_k, status_k, st, true, tau_bz_phys,
                    taukdiag);
}

int skkr_set_screening(skkr_ctx* c, int nl, const int* l_of_index, const double* a) {
  if (!c || nl < 1 || !l_of_index || !a) return SKKR_ERR_ARG;
  const int lk = c->cfg.lk;
  if (lk > 64) return SKKR_ERR_UNSUPPORTED;
  for (int i = 0; i < lk; ++i)
    if (l_of_index[i] < 0 || l_of_index[i] >= nl) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  const size_t sites = (size_t)c->cfg.n_pl * c->ns, tsz = sites * lk * lk;
  if (!c->dA) {
    CHK(dalloc_t(c, &c->dA, sites * lk));
    CHK(dalloc_t(c, &c->dLidx, (size_t)lk));
    CHK(dalloc_t(c, &c->physx, tsz));
    CHK(dalloc_t(c, &c->taubz_phys, tsz));
  }
  if (nl > c->nl) c->phk_cap = 0;  // traces buffer sized by nl: reallocate on next use
  HIPCHK(hipMemcpy(c->dA, a, sizeof(z) * sites * lk, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dLidx, l_of_index, sizeof(int) * lk, hipMemcpyHostToDevice));
  c->nl = nl;
  c->have_screening = true;
  return SKKR_OK;
}

int skkr_set_option(skkr_ctx* c, const char* key, int value) {
  if (!c || !key) return SKKR_ERR_ARG;
  if (strcmp(key, "graphs") == 0) {
    if (value != 0 && value != 1) return SKKR_ERR_ARG;
    c->graphs = value;
    return SKKR_OK;
  }
  if (strcmp(key, "gemm_tile") == 0) {
    if (value != 0 && value != 32 && value != 64 && value != 96 && value != 128) return SKKR_ERR_ARG;
    c->gemm_tile_force = value;
    drop_graphs(c);
    return SKKR_OK;
  }
  if (strcmp(key, "inv_kernel") == 0) {
    if (value < 0 || value > 2) return SKKR_ERR_ARG;
    c->inv_kernel = value;
    drop_graphs(c);
    return SKKR_OK;
  }
  if (strcmp(key, "split_dec") == 0 || strcmp(key, "split_sw") == 0 || strcmp(key, "split_sw_low") == 0 ||
      strcmp(key, "split_dec_high") == 0 || strcmp(key, "split_dec_late") == 0) {
    // BF16 slices 2..5, INT8 digits 2..6 (by the context's precision)
    const int smax = c->split_i8 ? 6 : 5;
    if (value != 0 && (value < 2 || value > smax)) return SKKR_ERR_ARG;
    if (strcmp(key, "split_dec") == 0) c->split_dec = value;
    else if (strcmp(key, "split_dec_high") == 0) c->split_dec_high = value;
    else if (strcmp(key, "split_dec_late") == 0) c->split_dec_late = value;
    else if (strcmp(key, "split_sw") == 0) c->split_sw = value;
    else c->split_sw_low = value;
    drop_graphs(c);
    return SKKR_OK;
  }
  if (strcmp(key, "zgemm_split") == 0 || strcmp(key, "zgemm_int8") == 0) {
    const bool i8 = strcmp(key, "zgemm_int8") == 0;
    if (value != 0 && (value < 2 || value > (i8 ? 6 : 5))) return SKKR_ERR_ARG;
    c->zgemm_split = value;
    c->zgemm_i8 = i8 ? 1 : 0;
    return SKKR_OK;
  }
  if (strcmp(key, "dec_late_log2") == 0 || strcmp(key, "dec_keep_y_log2") == 0) {
    if (value < 0 || value > 60) return SKKR_ERR_ARG;
    if (strcmp(key, "dec_late_log2") == 0) c->dec_late_log2 = value;
    else c->dec_keep_y_log2 = value;
    drop_graphs(c);
    return SKKR_OK;
  }
  if (strcmp(key, "int8_tile") == 0) {
    if (value != 64 && value != 128) return SKKR_ERR_ARG;
    c->i8_tile = value;
    drop_graphs(c);
    return SKKR_OK;
  }
  if (strcmp(key, "gemm_3m") == 0) {
    if (value != 0 && value != 1) return SKKR_ERR_ARG;
    c->gemm_3m = value;
    drop_graphs(c);
    return SKKR_OK;
  }
  return SKKR_ERR_ARG;
}

// mean launch time (us) of the plain GEMM m x n x k over nb batch entries
static int time_gemm(skkr_ctx* c, int tile, int m, int n, int k, z** A, int lda, z** B, int ldb,
                     z** C, int ldc, int nb, double* us) {
  CHK(ensure_phase_events(c));
  // the forced tile is restored on every return path
  struct Restore {
    skkr_ctx* c;
    int saved;
    ~Restore() { c->gemm_tile_force = saved; }
  } restore{c, c->gemm_tile_force};
  c->gemm_tile_force = tile;
  const z one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  hipEvent_t e0 = c->phase_ev[0], e1 = c->phase_ev[1];
  int r = launch_zgemm(c, m, n, k, A, lda, B, ldb, C, ldc, one, zero, nb, nullptr);
  const int reps = 3;
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps && r == SKKR_OK; ++i)
    r = launch_zgemm(c, m, n, k, A, lda, B, ldb, C, ldc, one, zero, nb, nullptr);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *us = 1e3 * ms / reps;
  return r;
}

int skkr_autotune(skkr_ctx* c, skkr_tuning* out) {
  if (!c) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  const int B = c->cfg.batch > 0 ? c->cfg.batch : 1024;
  CHK(ensure_batch(c, std::max(c->B, std::min(B, 1024))));
  const int n = c->n, nb = c->B;
  const int tiles[3] = {32, 64, 128};
  skkr_tuning t;
  memset(&t, 0, sizeof(t));
  double best_nn = 1e30, best_2nn = 1e30;
  for (int v = 0; v < 3; ++v) {
    if (tiles[v] == 32 && n > 64) continue;  // far too many work-groups to matter
    double us = 0.0;
    CHK(time_gemm(c, tiles[v], n, n, n, c->p_M10.d, n, c->p_M01.d, n, c->p_W.d, n, nb, &us));
    t.us_nn[v] = us;
    if (us < best_nn) { best_nn = us; t.tile_nn = tiles[v]; }
    CHK(time_gemm(c, tiles[v], 2 * n, n, n, c->p_AB[0].d, 2 * n, c->p_Yinv.d, n, c->p_AYBY.d,
                  2 * n, 2 * nb, &us));
    t.us_2nn[v] = us;
    if (us < best_2nn) { best_2nn = us; t.tile_2nn = tiles[v]; }
  }
  c->tile_nn = t.tile_nn;
  c->tile_2nn = t.tile_2nn;
  drop_graphs(c);
  if (out) *out = t;
  return SKKR_OK;
}

// ---------------- per-routine batched entry points ----------------------
int skkr_pl_blocks(skkr_ctx* c, int nk, const double* k, double* m00, double* m01,
                   double* m10) {
  if (!c || nk <= 0 || !k || !m00 || !m01 || !m10 || !c->have_structure) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  const size_t nn = c->nn;
  for (int k0 = 0; k0 < nk; k0 += 256) {
    int nb = std::min(256, nk - k0);
    CHK(ensure_batch(c, std::max(c->B, nb)));
    HIPCHK(hipMemcpyAsync(c->dk, k + 2 * (size_t)k0, 2 * sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
    CHK(run_fourier(c, nb));
    HIPCHK(hipMemcpyAsync(m00 + 2 * nn * k0, c->M00, sizeof(z) * nn * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(m01 + 2 * nn * k0, c->M01, sizeof(z) * nn * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(m10 + 2 * nn * k0, c->M10, sizeof(z) * nn * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return SKKR_OK;
}

// temporary device copy of nb host matrices + pointer array
struct TmpMats {
  z* d = nullptr;
  z** p = nullptr;
  ~TmpMats() {
    if (d) hipFree(d);
    if (p) hipFree(p);
  }
  int init(int nb, size_t nn, const double* host, hipStream_t s) {
    if (hipMalloc(&d, sizeof(z) * nn * nb) != hipSuccess) return SKKR_ERR_NOMEM;
    if (hipMalloc(&p, sizeof(z*) * nb) != hipSuccess) return SKKR_ERR_NOMEM;
    std::vector<z*> v(nb);
    for (int j = 0; j < nb; ++j) v[j] = d + j * nn;
    HIPCHK(hipMemcpy(p, v.data(), sizeof(z*) * nb, hipMemcpyHostToDevice));
    if (host) HIPCHK(hipMemcpyAsync(d, host, sizeof(z) * nn * nb, hipMemcpyHostToDevice, s));
    return SKKR_OK;
  }
};

// temporary device buffer, freed on every return path
struct TmpBuf {
  void* d = nullptr;
  ~TmpBuf() {
    if (d) hipFree(d);
  }
};

int skkr_gjinv_batch(skkr_ctx* c, int n, int nb, double* a, int* status) {
  if (!c || n <= 0 || n > 256 || nb <= 0 || !a) return SKKR_ERR_ARG;
  if (n > 128 && n != c->n) return SKKR_ERR_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->cfg.device));
  TmpMats m, orig;
  CHK(m.init(nb, (size_t)n * n, a, c->stream));
  CHK(orig.init(nb, (size_t)n * n, a, c->stream));  // input kept for a pivoted second pass
  TmpBuf sbuf;
  if (hipMalloc(&sbuf.d, sizeof(int) * nb) != hipSuccess) return SKKR_ERR_NOMEM;
  int* dst = (int*)sbuf.d;
  HIPCHK(hipMemsetAsync(dst, 0, sizeof(int) * nb, c->stream));
  InvArgs ia;
  memset(&ia, 0, sizeof(ia));
  ia.n = n;
  ia.lk = n;
  ia.base = m.p;
  ia.out = m.p;
  ia.status = dst;
  if (n > 128) CHK(ensure_inverse(c, std::max(c->invB, (nb + 1) / 2)));
  int r = run_inverse(c, ia, nb);
  std::vector<int> hs(nb);
  if (r == SKKR_OK && (c->inv_kernel == 0 || c->inv_kernel == 2)) {
    // matrices whose blocked elimination flagged multiplier growth: inverted
    // again from the input with the row-wise kernel (pivots over all rows)
    HIPCHK(hipMemcpyAsync(hs.data(), dst, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int> again(nb);
    bool any = false;
    for (int j = 0; j < nb; ++j) {
      again[j] = (hs[j] & GJ_STATUS_REDO) ? 1 : 0;
      any = any || again[j];
    }
    if (any) {
      TmpBuf abuf;
      if (hipMalloc(&abuf.d, sizeof(int) * nb) != hipSuccess) return SKKR_ERR_NOMEM;
      HIPCHK(hipMemcpy(abuf.d, again.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
      for (int j = 0; j < nb; ++j) hs[j] = again[j] ? 0 : hs[j];
      HIPCHK(hipMemcpy(dst, hs.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
      InvArgs ib = ia;
      ib.base = orig.p;
      ib.active = (int*)abuf.d;
      const int saved = c->inv_kernel;
      c->inv_kernel = 1;
      r = run_inverse(c, ib, nb);
      c->inv_kernel = saved;
    }
  }
  if (r == SKKR_OK) {
    HIPCHK(hipMemcpyAsync(a, m.d, sizeof(z) * n * n * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hs.data(), dst, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < nb; ++j) {
      int code = status_code(hs[j]);
      if (status) status[j] = code;
      if (code != SKKR_OK && r == SKKR_OK) r = code;
    }
  }
  return r;
}

int skkr_zgemm_batch(skkr_ctx* c, int n, int nb, const double* a, const double* b, double* cc) {
  if (!c || n <= 0 || nb <= 0 || !a || !b || !cc) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  TmpMats A, Bm, C;
  const size_t nn = (size_t)n * n;
  CHK(A.init(nb, nn, a, c->stream));
  CHK(Bm.init(nb, nn, b, c->stream));
  CHK(C.init(nb, nn, nullptr, c->stream));
  SplitScope sp(c, c->zgemm_split);
  struct KindScope {
    skkr_ctx* c;
    int saved;
    ~KindScope() { c->split_i8 = saved; }
  } kind{c, c->split_i8};
  c->split_i8 = c->zgemm_i8;
  CHK(launch_zgemm(c, n, n, n, A.p, n, Bm.p, n, C.p, n, make_double2(1, 0), make_double2(0, 0), nb, nullptr));
  HIPCHK(hipMemcpyAsync(cc, C.d, sizeof(z) * nn * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SKKR_OK;
}

// decimation of nb chains given by host matrices; lead t^-1 added for dugo
// Decimation of nb chains whose blocks are already in the device buffers
// (left lead: M10/M00/M01 entries 0..nb-1; right lead: M01/M00/M10, chains
// B..B+nb-1), again with the fully pivoting inversion when a blocked one
// flagged growth.  Leaves g_s in Yinv and, with want_sigma, the self-energy
// in Ycur (left) or Xc (right).
static int dec_device(skkr_ctx* c, int nb, bool add_tinv, char side, bool want_sigma,
                      const PtrArr* mt = nullptr, const PtrArr* lt = nullptr) {
  const int B = c->B;
  const int off = (side == 'R') ? B : 0;
  std::vector<int> act(2 * B, 0);
  for (int j = 0; j < nb; ++j) act[off + j] = 1;
  HIPCHK(hipMemcpyAsync(c->active, act.data(), 2 * B * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->iters, 0, 2 * B * sizeof(int), c->stream));
  HIPCHK(hipMemsetAsync(c->status, 0, 2 * B * sizeof(int), c->stream));
  CHK(run_decimation(c, add_tinv, want_sigma, mt, lt));
  // growth flagged in a blocked inversion: the chains again, pivoting over all rows
  std::vector<int> hs0(2 * B);
  HIPCHK(hipMemcpyAsync(hs0.data(), c->status, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  bool redo = false;
  for (int j = 0; j < nb; ++j) redo = redo || (hs0[off + j] & GJ_STATUS_REDO);
  if (redo) {
    const int saved = c->inv_kernel;
    c->inv_kernel = 1;
    HIPCHK(hipMemcpyAsync(c->active, act.data(), 2 * B * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->iters, 0, 2 * B * sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync(c->status, 0, 2 * B * sizeof(int), c->stream));
    const int r = run_decimation(c, add_tinv, want_sigma, mt, lt);
    c->inv_kernel = saved;
    CHK(r);
  }
  return SKKR_OK;
}

// iteration counts and status codes of the nb chains of one side
static int dec_results(skkr_ctx* c, int nb, char side, int* iters, int* status) {
  const int B = c->B;
  const int off = (side == 'R') ? B : 0;
  std::vector<int> hi(2 * B), hs(2 * B);
  HIPCHK(hipMemcpyAsync(hi.data(), c->iters, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(hs.data(), c->status, 2 * B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = SKKR_OK;
  for (int j = 0; j < nb; ++j) {
    int code = status_code(hs[off + j]);
    if (iters) iters[j] = hi[off + j];
    if (status) status[j] = code;
    if (code != SKKR_OK && rc == SKKR_OK) rc = code;
  }
  return rc;
}

static int dec_host(skkr_ctx* c, int nb, const double* ml, const double* m0, const double* mr,
                    bool add_tinv, char side, double* x_out, double* sigma_out, int* iters,
                    int* status) {
  const size_t nn = c->nn;
  CHK(ensure_batch(c, std::max(c->B, nb)));
  const int B = c->B;
  // chains occupy the left half [0, nb) (or the right half for side 'R' so that
  // the right lead t^-1 is used)
  const int off = (side == 'R') ? B : 0;
  z* mlbuf = (side == 'R') ? c->M01 : c->M10;
  z* mrbuf = (side == 'R') ? c->M10 : c->M01;
  HIPCHK(hipMemcpyAsync(mlbuf, ml, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(mrbuf, mr, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->M00, m0, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  CHK(dec_device(c, nb, add_tinv, side, sigma_out != nullptr));
  if (x_out)
    HIPCHK(hipMemcpyAsync(x_out, c->Yinv + off * nn, sizeof(z) * nn * nb, hipMemcpyDeviceToHost, c->stream));
  if (sigma_out) {
    z* src = (side == 'R') ? c->Xc : c->Ycur;
    HIPCHK(hipMemcpyAsync(sigma_out, src, sizeof(z) * nn * nb, hipMemcpyDeviceToHost, c->stream));
  }
  return dec_results(c, nb, side, iters, status);
}

int skkr_surfgfd_batch(skkr_ctx* c, int nb, const double* ml, const double* m0,
                       const double* mr, double* x, int* iters, int* status) {
  if (!c || nb <= 0 || !ml || !m0 || !mr || !x) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  return dec_host(c, nb, ml, m0, mr, false, 'L', x, nullptr, iters, status);
}

}  // extern "C"

// ------------------------------------------------------------------------
// Decimation without allocation (skkr_surfgfd_noalloc): the work buffers of
// the chains live in one statically allocated device array per GPU.
// ------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SKKR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDev = 64;
void* g_scratch[kMaxDev] = {};  // device address of scratch on each GPU
}  // namespace

// bytes of scratch that nb chains of n x n take (256-byte aligned pieces)
static size_t noalloc_bytes(int n, int nb, int itermax) {
  const size_t B2 = 2 * (size_t)((nb + 1) / 2), nn = (size_t)n * n;
  const size_t tiles = (size_t)((2 * n + 31) / 32) * ((n + 31) / 32);
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  return al(B2 * nn * sizeof(double2)) * 2          // eps, x
         + al(2 * B2 * 2 * nn * sizeof(double2))    // two [alpha; beta]
         + al(B2 * nn * sizeof(double2))            // y
         + al(B2 * 2 * nn * sizeof(double2))        // [ay; by]
         + al(B2 * tiles * sizeof(double))          // |alpha'|^2 partials
         + 2 * al(B2 * sizeof(int))                 // active, active0
         + al(((size_t)itermax + 2) * sizeof(int))  // active count per iteration
         + al(2 * B2 * sizeof(int))                 // iterations, status
         + 13 * al(B2 * sizeof(double2*));          // pointer arrays
}

// device address of this GPU's scratch (looked up on the first call per device)
static int scratch_ptr(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDev) return SKKR_ERR_ARG;
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = p;
  }
  *out = (unsigned char*)g_scratch[dev];
  return SKKR_OK;
}

// out[j] = base + j * stride (j < count), and base for count <= j < total
__global__ void fill_ptrs_kernel(double2** out, double2* base, long long stride, int count, int total) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < total) out[j] = base + (j < count ? j : 0) * stride;
}
__global__ void fill_int_kernel(int* out, int value, int count, int total) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < total) out[j] = (j < count) ? value : 0;
}

extern "C" {

int skkr_surfgfd_noalloc(int device, int n, int nb, const double* ml, const double* m0, const double* mr,
                         double* x, int* iters, int* status, double errmax, int itermax, void* stream) {
  if (n <= 0 || n > 128 || nb <= 0 || !ml || !m0 || !mr || !x || !iters || !status || errmax <= 0 ||
      itermax < 1 || itermax > 64)
    return SKKR_ERR_ARG;
  const size_t need = noalloc_bytes(n, nb, itermax);
  if (need > (size_t)SKKR_NOALLOC_SCRATCH_BYTES || nb > 1024) {
    fprintf(stderr,
            "skkr_surfgfd_noalloc: %d chains of %d x %d need %zu bytes of work space; the static scratch "
            "holds %d bytes and at most 1024 chains\n",
            nb, n, n, need, (int)SKKR_NOALLOC_SCRATCH_BYTES);
    return SKKR_ERR_TOO_LARGE;
  }
  HIPCHK(hipSetDevice(device));
  unsigned char* base = nullptr;
  CHK(scratch_ptr(device, &base));
  // a context on the host stack whose buffers are views of the scratch and
  // of the caller's arrays; it owns no device memory (allocs stay empty)
  skkr_ctx v;
  v.cfg.lk = n;
  v.cfg.sites_per_pl = 1;
  v.cfg.n_pl = 1;
  v.cfg.errmax = errmax;
  v.cfg.itermax_l = v.cfg.itermax_r = itermax;
  v.cfg.precision = SKKR_FP64;
  v.cfg.device = device;
  v.n = n;
  v.nn = n * n;
  v.ns = 1;
  v.stream = (hipStream_t)stream;
  v.graphs = 0;
  const int B = (nb + 1) / 2, B2 = 2 * B;
  v.B = B;
  const size_t nn = (size_t)n * n;
  const size_t tiles = (size_t)((2 * n + 31) / 32) * ((n + 31) / 32);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  z* eps = (z*)take(B2 * nn * sizeof(z));
  z* xs = (z*)take(B2 * nn * sizeof(z));
  z* AB = (z*)take(2 * (size_t)B2 * 2 * nn * sizeof(z));
  z* Yinv = (z*)take(B2 * nn * sizeof(z));
  z* AYBY = (z*)take(B2 * 2 * nn * sizeof(z));
  v.dnorm = (double*)take(B2 * tiles * sizeof(double));
  v.active = (int*)take(B2 * sizeof(int));
  v.active0 = (int*)take(B2 * sizeof(int));
  v.nactive = (int*)take(((size_t)itermax + 2) * sizeof(int));
  // iterations and status of the chains (those beyond nb are padding)
  int* itst = (int*)take(2 * (size_t)B2 * sizeof(int));
  v.iters = itst;
  v.status = itst + B2;
  z** pa[13];
  for (int i = 0; i < 13; ++i) pa[i] = (z**)take(B2 * sizeof(z*));
  hipStream_t st = v.stream;
  const unsigned g = (unsigned)((B2 + 255) / 256);
  auto fill = [&](z** out, z* b, long long stride, int count) {
    fill_ptrs_kernel<<<g, 256, 0, st>>>(out, b, stride, count, B2);
  };
  fill(pa[0], (z*)ml, (long long)nn, nb);
  fill(pa[1], (z*)mr, (long long)nn, nb);
  fill(pa[2], (z*)m0, (long long)nn, nb);
  fill(pa[3], eps, (long long)nn, B2);
  fill(pa[4], xs, (long long)nn, B2);
  fill(pa[5], Yinv, (long long)nn, B2);
  fill(pa[6], AYBY, 2 * (long long)nn, B2);
  for (int h = 0; h < 2; ++h) {
    z* abh = AB + (size_t)h * B2 * 2 * nn;
    fill(pa[7 + h], abh, 2 * (long long)nn, B2);       // [alpha; beta] (and alpha)
    fill(pa[9 + h], abh + n, 2 * (long long)nn, B2);   // beta
  }
  fill_int_kernel<<<g, 256, 0, st>>>(v.active, 1, nb, B2);
  fill_int_kernel<<<g, 256, 0, st>>>(v.iters, 0, 0, 2 * B2);  // iters and status zeroed
  HIPCHK(hipGetLastError());
  v.p_ml.d = pa[0];
  v.p_mr.d = pa[1];
  v.p_m0.d = pa[2];
  v.p_eps.d = pa[3];
  v.p_x.d = pa[4];
  v.p_Yinv.d = pa[5];
  v.p_AYBY.d = pa[6];
  v.p_AB[0].d = v.p_alpha[0].d = pa[7];
  v.p_AB[1].d = v.p_alpha[1].d = pa[8];
  v.p_beta[0].d = pa[9];
  v.p_beta[1].d = pa[10];
  int r = run_decimation(&v, false, false);
  // growth flagged in a blocked inversion: the chains again with full pivoting
  int hs[1024];
  if (r == SKKR_OK) {
    HIPCHK(hipMemcpyAsync(hs, v.status, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    bool redo = false;
    for (int j = 0; j < nb; ++j) redo = redo || (hs[j] & GJ_STATUS_REDO);
    if (redo) {
      fill_int_kernel<<<g, 256, 0, st>>>(v.active, 1, nb, B2);
      fill_int_kernel<<<g, 256, 0, st>>>(v.iters, 0, 0, 2 * B2);
      v.inv_kernel = 1;
      r = run_decimation(&v, false, false);
    }
  }
  if (r == SKKR_OK) {
    HIPCHK(hipMemcpyAsync(x, Yinv, sizeof(z) * nn * nb, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(iters, v.iters, sizeof(int) * nb, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(hs, v.status, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int j = 0; j < nb; ++j) {
      hs[j] = status_code(hs[j]);
      if (hs[j] != SKKR_OK && r == SKKR_OK) r = hs[j];
    }
    HIPCHK(hipMemcpyAsync(status, hs, sizeof(int) * nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  v.stream = nullptr;  // the caller's stream: not destroyed with the view
  return r;
}

int skkr_dugo_batch(skkr_ctx* c, char side, int nb, const double* m00, const double* m01,
                    const double* m10, double* xy, int* iters, int* status) {
  if (!c || nb <= 0 || !m00 || !m01 || !m10 || !xy || (side != 'L' && side != 'R') || !c->have_tinv)
    return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  if (side == 'L') return dec_host(c, nb, m10, m00, m01, true, 'L', nullptr, xy, iters, status);
  return dec_host(c, nb, m01, m00, m10, true, 'R', nullptr, xy, iters, status);
}

// The sweeps of nb k-points, again with the row-wise (fully pivoting)
// inversion when a blocked inversion flagged multiplier growth.  The forward
// sweep overwrites Y_1 (Ycur) with the later Y_p: the host's y is copied in
// again before the second pass.
static int sweeps_checked(skkr_ctx* c, int nb, const double* y_host) {
  CHK(run_sweeps(c, nb));
  std::vector<int> hs(nb);
  HIPCHK(hipMemcpyAsync(hs.data(), c->status_sw, nb * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  bool redo = false;
  for (int j = 0; j < nb; ++j) redo = redo || (hs[j] & GJ_STATUS_REDO);
  if (!redo) return SKKR_OK;
  const int saved = c->inv_kernel;
  c->inv_kernel = 1;
  HIPCHK(hipMemsetAsync(c->status_sw, 0, nb * sizeof(int), c->stream));
  HIPCHK(hipMemcpyAsync(c->Ycur, y_host, sizeof(z) * c->nn * nb, hipMemcpyHostToDevice, c->stream));
  const int r = sweep_launches(c, nb);
  c->inv_kernel = saved;
  return r;
}

int skkr_tau2d_batch(skkr_ctx* c, int nb, const double* m00, const double* m01,
                     const double* m10, const double* x, const double* y, double w,
                     double* tau, int* status) {
  if (!c || nb <= 0 || !m00 || !m01 || !m10 || !x || !y || !tau || !c->have_tinv)
    return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  CHK(ensure_batch(c, std::max(c->B, nb)));
  const size_t nn = c->nn;
  const int P = c->cfg.n_pl, lk = c->cfg.lk;
  const size_t tsz = (size_t)P * c->ns * lk * lk;
  HIPCHK(hipMemcpyAsync(c->M00, m00, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->M01, m01, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->M10, m10, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Xc, x, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Ycur, y, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->status_sw, 0, nb * sizeof(int), c->stream));
  CHK(sweeps_checked(c, nb, y));
  long long ne = (long long)tsz * nb;
  scale_copy_kernel<<<(unsigned)((ne + 255) / 256), 256, 0, c->stream>>>(ne, c->tauk, w, c->tauk);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(tau, c->tauk, sizeof(z) * tsz * nb, hipMemcpyDeviceToHost, c->stream));
  std::vector<int> hs(nb);
  HIPCHK(hipMemcpyAsync(hs.data(), c->status_sw, nb * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = SKKR_OK;
  for (int j = 0; j < nb; ++j) {
    int code = status_code(hs[j]);
    if (status) status[j] = code;
    if (code != SKKR_OK && rc == SKKR_OK) rc = code;
  }
  return rc;
}

int skkr_tau2d_layered_batch(skkr_ctx* c, int nb, const double* m00, const double* m01,
                             const double* m10, const double* x, const double* y, double w,
                             double* tau, int* status) {
  if (!c || nb <= 0 || !m00 || !m01 || !m10 || !x || !y || !tau || !c->have_tinv)
    return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  CHK(ensure_batch(c, std::max(c->B, nb)));
  CHK(ensure_layered(c));
  const size_t nn = c->nn, Bs = c->B;
  const int P = c->cfg.n_pl, lk = c->cfg.lk;
  const size_t tsz = (size_t)P * c->ns * lk * lk;
  // host [k][p] blocks -> device [p][k]
  const size_t blk = sizeof(z) * nn;
  for (int p = 0; p < P; ++p) {
    HIPCHK(hipMemcpy2DAsync(c->M00pl + (size_t)p * Bs * nn, blk, m00 + 2 * (size_t)p * nn, blk * P, blk, nb,
                            hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpy2DAsync(c->M01pl + (size_t)p * Bs * nn, blk, m01 + 2 * (size_t)p * nn, blk * P, blk, nb,
                            hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpy2DAsync(c->M10pl + (size_t)p * Bs * nn, blk, m10 + 2 * (size_t)p * nn, blk * P, blk, nb,
                            hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipMemcpyAsync(c->Xc, x, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Ycur, y, sizeof(z) * nn * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->status_sw, 0, nb * sizeof(int), c->stream));
  c->layered = true;
  const int rs = sweeps_checked(c, nb, y);
  c->layered = false;
  CHK(rs);
  long long ne = (long long)tsz * nb;
  scale_copy_kernel<<<(unsigned)((ne + 255) / 256), 256, 0, c->stream>>>(ne, c->tauk, w, c->tauk);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(tau, c->tauk, sizeof(z) * tsz * nb, hipMemcpyDeviceToHost, c->stream));
  std::vector<int> hs(nb);
  HIPCHK(hipMemcpyAsync(hs.data(), c->status_sw, nb * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = SKKR_OK;
  for (int j = 0; j < nb; ++j) {
    int code = status_code(hs[j]);
    if (status) status[j] = code;
    if (code != SKKR_OK && rc == SKKR_OK) rc = code;
  }
  return rc;
}

// ---------------------------------------------------------------------------
// tau2d and dugo with the reference's own arrays (g2d windows, padded leading
// dimensions), one k-point per call as the reference calls them
// ---------------------------------------------------------------------------

// Shape of the reference's g2d(lmmaxp, lmmaxp, -minprc:minprc, minprc) window
struct G2dShape {
  int lmmaxp, minprc;
  size_t window() const { return (size_t)lmmaxp * lmmaxp * (2 * minprc + 1) * minprc; }
};

// packer.f:6-100 for a principal layer of s layers between neighbours of s
// layers (ni = nim = nip = s): element (r, c) of the n x n blocks, n = lk s,
//   M00 = -G(i, i+j),   j = jp - i          (every layer pair of the PL),
//   M10 = -G(i, j),     j = jp - i - s     (jp >= i: the PL before),
//   M01 = -G(i, j),     j = jp - i + s     (jp <= i: the PL after),
// with G(i, j) the lm block g2d(:, :, j, i) (i 1-based).  irel = 1: every
// block through relmtrx.f:26-42, b(r, c) = u1(r) (f a)(ind1 r, ind1 c) u1(c)
// + u2(r) (f a)(ind2 r, ind2 c) u2(c), indices past the lm block reading 0.
// Null outputs are skipped.
__global__ __launch_bounds__(256) void g2d_pack_kernel(int s, int lk, int nl2, G2dShape sh,
                                                       const double2* __restrict__ win, int rel, double fr,
                                                       double fi, const double* __restrict__ u1,
                                                       const double* __restrict__ u2, const int* __restrict__ i1,
                                                       const int* __restrict__ i2,
                                                       double2* M00, double2* M10, double2* M01) {
  const int n = lk * s;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  const int i0 = r / lk, kr = r % lk, jp0 = c / lk, kc = c % lk;
  const int ld = sh.lmmaxp, nj = 2 * sh.minprc + 1;
  auto elem = [&](int a, int b, int j) {  // g2d(a, b, j, i0 + 1), a, b 0-based
    return win[a + (size_t)ld * (b + (size_t)ld * ((j + sh.minprc) + (size_t)nj * i0))];
  };
  auto blk = [&](int j) {
    if (!rel) {
      const double2 v = elem(kr, kc, j);
      return make_double2(-v.x, -v.y);
    }
    double2 acc = make_double2(0.0, 0.0);
    const int p1 = i1[kr], q1 = i1[kc], p2 = i2[kr], q2 = i2[kc];
    if (p1 < nl2 && q1 < nl2) {
      const double2 v = elem(p1, q1, j);
      const double cf = u1[kr] * u1[kc];
      acc.x += cf * (fr * v.x - fi * v.y);
      acc.y += cf * (fr * v.y + fi * v.x);
    }
    if (p2 < nl2 && q2 < nl2) {
      const double2 v = elem(p2, q2, j);
      const double cf = u2[kr] * u2[kc];
      acc.x += cf * (fr * v.x - fi * v.y);
      acc.y += cf * (fr * v.y + fi * v.x);
    }
    return make_double2(-acc.x, -acc.y);
  };
  const double2 zero = make_double2(0.0, 0.0);
  if (M00) M00[idx] = blk(jp0 - i0);
  if (M10) M10[idx] = (jp0 >= i0) ? blk(jp0 - i0 - s) : zero;
  if (M01) M01[idx] = (jp0 <= i0) ? blk(jp0 - i0 + s) : zero;
}

static int ensure_clebsch(skkr_ctx* c) {
  if (c->dU1) return SKKR_OK;
  std::vector<double> u1, u2;
  std::vector<int> i1, i2;
  kkrh::clebsch_tables(u1, u2, i1, i2);
  CHK(dalloc_t(c, &c->dU1, u1.size()));
  CHK(dalloc_t(c, &c->dU2, u2.size()));
  CHK(dalloc_t(c, &c->dI1, i1.size()));
  CHK(dalloc_t(c, &c->dI2, i2.size()));
  HIPCHK(hipMemcpy(c->dU1, u1.data(), sizeof(double) * u1.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dU2, u2.data(), sizeof(double) * u2.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dI1, i1.data(), sizeof(int) * i1.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dI2, i2.data(), sizeof(int) * i2.size(), hipMemcpyHostToDevice));
  return SKKR_OK;
}

static int ensure_g2d(skkr_ctx* c, size_t count) {
  if (count <= c->g2d_cap) return SKKR_OK;
  if (c->dG2d) {
    HIPCHK(hipStreamSynchronize(c->stream));
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->dG2d);
    if (it != c->allocs.end()) c->allocs.erase(it);
    HIPCHK(hipFree(c->dG2d));
    c->dG2d = nullptr;
  }
  CHK(dalloc_t(c, &c->dG2d, count));
  c->g2d_cap = count;
  return SKKR_OK;
}

static int launch_g2d_pack(skkr_ctx* c, int irel, int nl2, const G2dShape& sh, const z* win, double fr,
                           double fi, z* m00, z* m10, z* m01) {
  const int s = c->cfg.sites_per_pl, lk = c->cfg.lk;
  const long long ne = (long long)c->nn;
  g2d_pack_kernel<<<(unsigned)((ne + 255) / 256), 256, 0, c->stream>>>(
      s, lk, nl2, sh, win, irel, fr, fi, c->dU1, c->dU2, c->dI1, c->dI2, m00, m10, m01);
  HIPCHK(hipGetLastError());
  c->launches++;
  return SKKR_OK;
}

// common checks of the reference-shaped entries: the basis, the window and
// the padded dimensions against the context.  irel = 0 reads lkmax x lkmax
// blocks of g2d and never lmax (packer.f passes it to relmtrx only); irel = 1
// builds kappa-mu blocks of 2 (lmax+1)^2 from (lmax+1)^2 lm blocks.
static int check_ref_shape(skkr_ctx* c, int irel, int lmax, int lkmax, int lkmaxp, int mdimp, const G2dShape& sh,
                           int* nl2) {
  if (irel != 0 && irel != 1) return SKKR_ERR_ARG;
  if (lkmax != c->cfg.lk || lkmaxp < lkmax || mdimp < c->n || sh.minprc < c->cfg.sites_per_pl) return SKKR_ERR_ARG;
  if (irel) {
    if (lmax < 0 || lmax > 5 || lkmax != 2 * (lmax + 1) * (lmax + 1)) return SKKR_ERR_ARG;
    *nl2 = (lmax + 1) * (lmax + 1);
  } else {
    *nl2 = lkmax;
  }
  if (sh.lmmaxp < *nl2) return SKKR_ERR_ARG;
  return SKKR_OK;
}

// lk x lk blocks of a padded (lkmaxp, lkmaxp, *) array -> contiguous
static void unpad_blocks(const double* src, int lkmaxp, int lk, int nblk, std::vector<double>& dst) {
  dst.assign(2 * (size_t)lk * lk * nblk, 0.0);
  for (int b = 0; b < nblk; ++b)
    for (int cc = 0; cc < lk; ++cc)
      memcpy(&dst[2 * ((size_t)b * lk * lk + (size_t)cc * lk)],
             src + 2 * ((size_t)b * lkmaxp * lkmaxp + (size_t)cc * lkmaxp), sizeof(z) * lk);
}

int skkr_tau2d_g2d(skkr_ctx* c, int irel, double wgeff, int bulkgeo, int lmax, int lkmax, int lkmaxp,
                   int nintfc, const double* tminv, const double* g2d, const double* x, const double* y,
                   double* tau, int mdimp, int lmmaxp, int minprc, int mprc1, int mintfc, int nprc,
                   const int* ninprc, double fac_re, double fac_im) {
  if (!c || !tminv || !g2d || !x || !y || !tau || !ninprc) return SKKR_ERR_ARG;
  const G2dShape sh{lmmaxp, minprc};
  int nl2 = 0;
  CHK(check_ref_shape(c, irel, lmax, lkmax, lkmaxp, mdimp, sh, &nl2));
  const int P = c->cfg.n_pl, s = c->cfg.sites_per_pl, lk = c->cfg.lk, n = c->n;
  if (nprc != P || nintfc != P * s || mintfc < nintfc || mprc1 < 0) return SKKR_ERR_ARG;
  if (!bulkgeo && nprc > mprc1 + 1) return SKKR_ERR_ARG;  // planes 1..nprc of 0:mprc1+1
  for (int p = 1; p <= nprc; ++p)
    if (ninprc[p] != s) return SKKR_ERR_UNSUPPORTED;  // equal principal layers only
  HIPCHK(hipSetDevice(c->cfg.device));
  CHK(ensure_batch(c, std::max(c->B, 1)));
  if (!bulkgeo) CHK(ensure_layered(c));
  if (irel) CHK(ensure_clebsch(c));
  const size_t W = sh.window(), nn = c->nn, Bs = c->B;
  const int nplanes = bulkgeo ? 1 : nprc;
  CHK(ensure_g2d(c, W * nplanes));
  // planes 1..nplanes of g2d(.., 0:mprc1+1)
  HIPCHK(hipMemcpyAsync(c->dG2d, g2d + 2 * W, sizeof(z) * W * nplanes, hipMemcpyHostToDevice, c->stream));
  if (bulkgeo) {
    CHK(launch_g2d_pack(c, irel, nl2, sh, c->dG2d, fac_re, fac_im, c->M00, c->M10, c->M01));
  } else {
    for (int p = 0; p < nprc; ++p)
      CHK(launch_g2d_pack(c, irel, nl2, sh, c->dG2d + (size_t)p * W, fac_re, fac_im,
                          c->M00pl + (size_t)p * Bs * nn, c->M10pl + (size_t)p * Bs * nn,
                          c->M01pl + (size_t)p * Bs * nn));
  }
  // the interface t^-1 (replaces the context's, as skkr_set_tinv would)
  std::vector<double> ht;
  unpad_blocks(tminv, lkmaxp, lk, nintfc, ht);
  HIPCHK(hipMemcpyAsync(c->dTint, ht.data(), sizeof(double) * ht.size(), hipMemcpyHostToDevice, c->stream));
  // x, y: leading n x n of mdimp x mdimp
  std::vector<double> hy;
  unpad_blocks(y, mdimp, n, 1, hy);
  HIPCHK(hipMemcpy2DAsync(c->Xc, sizeof(z) * n, x, sizeof(z) * mdimp, sizeof(z) * n, n, hipMemcpyHostToDevice,
                          c->stream));
  HIPCHK(hipMemcpyAsync(c->Ycur, hy.data(), sizeof(z) * nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->status_sw, 0, sizeof(int), c->stream));
  c->layered = !bulkgeo;
  const int rs = sweeps_checked(c, 1, hy.data());
  c->layered = false;
  CHK(rs);
  const size_t tsz = (size_t)nintfc * lk * lk;
  scale_copy_kernel<<<(unsigned)((tsz + 255) / 256), 256, 0, c->stream>>>((long long)tsz, c->tauk, wgeff, c->tauk);
  HIPCHK(hipGetLastError());
  std::vector<double> ho(2 * tsz);
  int hs = 0;
  HIPCHK(hipMemcpyAsync(ho.data(), c->tauk, sizeof(z) * tsz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&hs, c->status_sw, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // tau(l, lp, ilay) of the padded (lkmaxp, lkmaxp, mintfc) array
  for (int b = 0; b < nintfc; ++b)
    for (int cc = 0; cc < lk; ++cc)
      memcpy(tau + 2 * ((size_t)b * lkmaxp * lkmaxp + (size_t)cc * lkmaxp),
             &ho[2 * ((size_t)b * lk * lk + (size_t)cc * lk)], sizeof(z) * lk);
  return status_code(hs);
}

int skkr_dugo_g2d(skkr_ctx* c, int lmax, int lkmax, const double* tminvb, int lkmaxp, const double* g2d,
                  const double* g2di, double* xy, int mdimp, char side, int bulk, int irel, int* sgfconv,
                  int lmmaxp, int minprc, int nprc, const int* ninprc, double fac_re, double fac_im,
                  int* iters) {
  if (!c || !tminvb || !g2d || !g2di || !xy || !ninprc || (side != 'L' && side != 'R')) return SKKR_ERR_ARG;
  const G2dShape sh{lmmaxp, minprc};
  int nl2 = 0;
  CHK(check_ref_shape(c, irel, lmax, lkmax, lkmaxp, mdimp, sh, &nl2));
  const int s = c->cfg.sites_per_pl, lk = c->cfg.lk, n = c->n;
  if (nprc < 1) return SKKR_ERR_ARG;
  // the lead PL and the interface PL next to it (dugo.f:44-72)
  const int ncpl = bulk ? ninprc[1] : (side == 'L' ? ninprc[0] : ninprc[nprc + 1]);
  const int ncpli = bulk ? ncpl : (side == 'L' ? ninprc[1] : ninprc[nprc]);
  if (ncpl != s || ncpli != s) return SKKR_ERR_UNSUPPORTED;  // equal principal layers only
  HIPCHK(hipSetDevice(c->cfg.device));
  CHK(ensure_batch(c, std::max(c->B, 1)));
  if (irel) CHK(ensure_clebsch(c));
  const size_t W = sh.window(), nn = c->nn;
  const int B = c->B;
  CHK(ensure_g2d(c, 2 * W));
  if (!c->dMt) CHK(dalloc_t(c, &c->dMt, nn));
  if (c->dmt_cap != 2 * B) {
    CHK(make_ptrs(c, &c->p_dmt, std::vector<z*>(2 * B, c->dMt)));
    c->dmt_cap = 2 * B;
  }
  HIPCHK(hipMemcpyAsync(c->dG2d, g2d, sizeof(z) * W, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dG2d + W, g2di, sizeof(z) * W, hipMemcpyHostToDevice, c->stream));
  const z* wg = c->dG2d;
  const z* wgi = c->dG2d + W;
  // bulk: the lead's blocks from g2di, the extra factor mt from g2d; otherwise
  // the lead's blocks from g2d and mt, the interface PL's coupling to the lead,
  // from g2di.  Chain 0 (left, ml = M10, mr = M01) or B (right, ml = M01, mr = M10).
  const z* wdec = bulk ? wgi : wg;
  const z* wmt = bulk ? wg : wgi;
  CHK(launch_g2d_pack(c, irel, nl2, sh, wdec, fac_re, fac_im, c->M00, c->M10, c->M01));
  const bool mt_is_m01 = (side == 'L') == (bool)bulk;  // bulk L, non-bulk R: mt = packer's m01
  CHK(launch_g2d_pack(c, irel, nl2, sh, wmt, fac_re, fac_im, nullptr, mt_is_m01 ? nullptr : c->dMt,
                      mt_is_m01 ? c->dMt : nullptr));
  // the lead t^-1 of this side
  std::vector<double> ht;
  unpad_blocks(tminvb, lkmaxp, lk, s, ht);
  HIPCHK(hipMemcpyAsync(side == 'L' ? c->dTl : c->dTr, ht.data(), sizeof(double) * ht.size(),
                        hipMemcpyHostToDevice, c->stream));
  // xy = m10 g mt (bulk L), mt g m01 (L), m01 g mt (bulk R), mt g m10 (R)
  CHK(dec_device(c, 1, true, side, true, bulk ? &c->p_dmt : nullptr, bulk ? nullptr : &c->p_dmt));
  std::vector<double> hxy(2 * nn);
  HIPCHK(hipMemcpyAsync(hxy.data(), side == 'R' ? c->Xc : c->Ycur, sizeof(z) * nn, hipMemcpyDeviceToHost,
                        c->stream));
  int it = 0, st = 0;
  const int rc = dec_results(c, 1, side, &it, &st);
  for (int cc = 0; cc < n; ++cc)
    memcpy(xy + 2 * (size_t)cc * mdimp, &hxy[2 * (size_t)cc * n], sizeof(z) * n);
  if (iters) *iters = it;
  if (sgfconv) *sgfconv = (st == SKKR_ERR_SGF_DIVERGED) ? 0 : 1;
  return rc;
}

int skkr_set_point_group(skkr_ctx* c, int ng, const double* dg, const double* dg1) {
  if (!c || ng < 0 || ng > 48) return SKKR_ERR_ARG;
  if (ng == 0) {
    c->ng_sym = 0;
    return SKKR_OK;
  }
  if (!dg || !dg1) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  const int lk = c->cfg.lk;
  const size_t S = (size_t)c->cfg.n_pl * c->ns, bl = (size_t)lk * lk;
  if (!c->dSymOut) {
    CHK(dalloc_t(c, &c->dSymOut, S * bl));
    CHK(dalloc_t(c, &c->dSymTmp, S * bl));
    std::vector<z*> v(S);
    for (size_t s = 0; s < S; ++s) v[s] = c->taubz + s * bl;
    CHK(make_ptrs(c, &c->p_symT, v));
    for (size_t s = 0; s < S; ++s) v[s] = c->dSymOut + s * bl;
    CHK(make_ptrs(c, &c->p_symOut, v));
    for (size_t s = 0; s < S; ++s) v[s] = c->dSymTmp + s * bl;
    CHK(make_ptrs(c, &c->p_symTmp, v));
  }
  if (ng > c->sym_cap) {
    CHK(dalloc_t(c, &c->dSym, 2 * (size_t)ng * bl));
    c->p_symD.assign(ng, PtrArr());
    c->p_symD1.assign(ng, PtrArr());
    for (int g = 0; g < ng; ++g) {
      CHK(make_ptrs(c, &c->p_symD[g], std::vector<z*>(S, c->dSym + g * bl)));
      CHK(make_ptrs(c, &c->p_symD1[g], std::vector<z*>(S, c->dSym + (ng + g) * bl)));
    }
    c->sym_cap = ng;
  } else {
    // the arrays of a larger earlier group: re-point them for this ng
    for (int g = 0; g < ng; ++g) {
      std::vector<z*> d(S, c->dSym + g * bl), d1(S, c->dSym + (ng + g) * bl);
      HIPCHK(hipMemcpy(c->p_symD[g].d, d.data(), sizeof(z*) * S, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->p_symD1[g].d, d1.data(), sizeof(z*) * S, hipMemcpyHostToDevice));
    }
  }
  HIPCHK(hipMemcpy(c->dSym, dg, sizeof(z) * ng * bl, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dSym + (size_t)ng * bl, dg1, sizeof(z) * ng * bl, hipMemcpyHostToDevice));
  c->ng_sym = ng;
  return SKKR_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------
// Layer structure constants (gstore.f with bulkgeo)
// ------------------------------------------------------------------------
struct skkr_str {
  skkr_str_config cfg;
  int nl2 = 0, nlay = 0, N = 0, B = 0, nkam = 0, ndir = 0, nblk = 0;
  double rlmin = 0.0;
  kkrh::Lattice lat;
  skkr_ctx* inv = nullptr;   // batched inversion and products of size min(N, 256)
  skkr_ctx* invS = nullptr;  // N > 384: the (N - 256)-row Schur complement
  // N > 256: 2 x 2 blocks around the leading 256 rows (views into d_g1)
  z *d_X11 = nullptr, *d_C = nullptr, *d_D = nullptr, *d_Si = nullptr;
  PtrArr p_X11, p_C, p_D, p_Si, p_A11, p_A21, p_A12, p_A22;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // tables
  double *d_kam_sep = nullptr, *d_dir_sep = nullptr, *d_gau = nullptr, *d_flm = nullptr,
         *d_yb = nullptr;
  kkr::cx* d_ycf = nullptr;
  int *d_kam_blk = nullptr, *d_dir_blk = nullptr, *d_r_off = nullptr, *d_rsh_off = nullptr,
      *d_rsh_n = nullptr, *d_rsh = nullptr, *d_gsh = nullptr, *d_bmap = nullptr, *d_lof = nullptr,
      *d_r_nvec = nullptr;
  double2 *d_rsorted = nullptr, *d_gsorted = nullptr;
  int ngsh = 0, ngvec = 0;
  // batch buffers
  double* d_k = nullptr;
  z *d_alpha = nullptr, *d_blk = nullptr, *d_g = nullptr, *d_g1 = nullptr, *d_gs = nullptr,
    *d_g2d = nullptr;
  int* d_status = nullptr;
  PtrArr p_g, p_g1, p_gs;
  // the rows R of the middle three PLs (planes nextra .. nextra + 2) of g, the
  // columns R of (1 - alpha g)^-1 and the blocks (R, R) and (R, all) of gs
  PtrArr p_gR, p_g1C, p_gsRC, p_gsR;
};

namespace {
template <class T>
int str_alloc(skkr_str* sc, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (count ? count : 1)) != hipSuccess) return SKKR_ERR_NOMEM;
  sc->allocs.push_back(q);
  *p = (T*)q;
  return SKKR_OK;
}
template <class T>
int str_upload(skkr_str* sc, T** p, const std::vector<T>& v) {
  CHK(str_alloc(sc, p, v.size()));
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return SKKR_OK;
}
int str_ptrs(skkr_str* sc, PtrArr* pa, z* base, size_t stride, int n) {
  std::vector<z*> v(n);
  for (int j = 0; j < n; ++j) v[j] = base + (size_t)j * stride;
  CHK(str_alloc(sc, &pa->d, (size_t)n));
  HIPCHK(hipMemcpy(pa->d, v.data(), sizeof(z*) * n, hipMemcpyHostToDevice));
  return SKKR_OK;
}
constexpr int kStrMaxVec = 4096;  // vectors kept per shell list (cut at a shell end)
}  // namespace

extern "C" {

int skkr_str_create(skkr_str** out, const skkr_str_config* cfg, const double* cvec) {
  if (!out || !cfg || !cvec) return SKKR_ERR_ARG;
  *out = nullptr;
  // nextra >= 1: the blocks of the outer principal layers reach s layers
  // beyond them (gstore.f:243-276 reads layers i - s .. i + s)
  if (cfg->lmax < 1 || cfg->s < 1 || cfg->nextra < 1 || !(cfg->eta > 0.0) || cfg->batch < 0)
    return SKKR_ERR_ARG;
  if (cfg->lmax > 3) return SKKR_ERR_UNSUPPORTED;
  const int nl2 = (cfg->lmax + 1) * (cfg->lmax + 1), nlay = (2 * cfg->nextra + 3) * cfg->s;
  if (nl2 * nlay > 512) return SKKR_ERR_UNSUPPORTED;
  const double cell = cfg->a1[0] * cfg->a2[1] - cfg->a1[1] * cfg->a2[0];
  if (!(std::fabs(cell) > 0.0)) return SKKR_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) return SKKR_ERR_HIP;
  HIPCHK(hipSetDevice(cfg->device));
  skkr_str* sc = new skkr_str();
  sc->cfg = *cfg;
  sc->nl2 = nl2;
  sc->nlay = nlay;
  sc->N = nl2 * nlay;
  sc->B = cfg->batch ? cfg->batch : 256;
  auto fail = [&](int r) {
    skkr_str_destroy(sc);
    return r;
  };
  // inversion / product engine of size N
  skkr_config ic;
  memset(&ic, 0, sizeof(ic));
  ic.lk = std::min(sc->N, 256);
  ic.sites_per_pl = 1;
  ic.n_pl = 1;
  ic.errmax = 1e-8;
  ic.itermax_l = ic.itermax_r = 1;
  ic.precision = SKKR_FP64;
  ic.device = cfg->device;
  int r = skkr_create(&sc->inv, &ic);
  if (r != SKKR_OK) return fail(r);
  if (sc->N - 256 > 128) {
    ic.lk = sc->N - 256;
    r = skkr_create(&sc->invS, &ic);
    if (r != SKKR_OK) return fail(r);
  }
  // 1 - alpha G has no definite anti-Hermitian part: its inversions pivot over
  // all rows (the row-wise kernel), within the halves of a 2 x 2 split
  sc->inv->inv_kernel = 1;
  if (sc->invS) sc->invS->inv_kernel = 1;
  sc->stream = sc->inv->stream;
  // lattice and the stack of layers (gstore.f:99-122)
  sc->lat = kkrh::make_lattice(cfg->a1, cfg->a2, 20000);
  const auto& L = sc->lat;
  sc->rlmin = std::sqrt(std::fmin(L.b1[0] * L.b1[0] + L.b1[1] * L.b1[1], L.b2[0] * L.b2[0] + L.b2[1] * L.b2[1]));
  const int s = cfg->s;
  std::vector<double> cv(3 * (size_t)nlay);
  double cb[3];
  for (int d = 0; d < 3; ++d) cb[d] = cvec[3 * s + d] - cvec[d];
  for (int i = 0; i < s; ++i)
    for (int d = 0; d < 3; ++d) cv[3 * i + d] = cvec[3 * i + d] - cvec[d];
  for (int ip = 2; ip <= 2 * cfg->nextra + 3; ++ip)
    for (int i = 0; i < s; ++i)
      for (int d = 0; d < 3; ++d) cv[3 * ((ip - 1) * s + i) + d] = cv[3 * i + d] + (ip - 1) * cb[d];
  // distinct separations of all layer pairs (gstore.f:164-202): block (j, i)
  // of G holds the pair (i, j) with sep = cvec(j) - cvec(i); Kambe for close
  // layers, the direct sum (j > i gives blocks (j, i) and (i, j)) otherwise
  std::vector<double> kam_sep, dir_sep;
  std::vector<int> kam_blk, dir_blk, bmap((size_t)nlay * nlay, -1);
  auto same = [](const double* a, const double* b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; };
  for (int i = 0; i < nlay; ++i)
    for (int j = 0; j < nlay; ++j) {
      const double sp[3] = {cv[3 * j] - cv[3 * i], cv[3 * j + 1] - cv[3 * i + 1], cv[3 * j + 2] - cv[3 * i + 2]};
      if (std::fabs(sp[2]) * sc->rlmin >= 5.0) {
        if (j <= i) continue;
        int u = -1;
        for (int q = 0; q < (int)dir_blk.size(); ++q)
          if (same(&dir_sep[3 * q], sp)) u = q;
        if (u < 0) {
          u = (int)dir_blk.size();
          for (int d = 0; d < 3; ++d) dir_sep.push_back(sp[d]);
          dir_blk.push_back(sc->nblk);
          sc->nblk += 2;
        }
        bmap[(size_t)j * nlay + i] = dir_blk[u];
        bmap[(size_t)i * nlay + j] = dir_blk[u] + 1;
      } else {
        int u = -1;
        for (int q = 0; q < (int)kam_blk.size(); ++q)
          if (same(&kam_sep[4 * q + 1], sp)) u = q;
        if (u < 0) {
          u = (int)kam_blk.size();
          kam_sep.push_back(std::sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]));
          for (int d = 0; d < 3; ++d) kam_sep.push_back(sp[d]);
          kam_blk.push_back(sc->nblk);
          sc->nblk += 1;
        }
        bmap[(size_t)j * nlay + i] = kam_blk[u];
      }
    }
  for (int v : bmap)
    if (v < 0) return fail(SKKR_ERR_ARG);
  sc->nkam = (int)kam_blk.size();
  sc->ndir = (int)dir_blk.size();
  // shells: reciprocal vectors around the origin, real vectors around each
  // Kambe pair's in-plane offset (vecsrt2d.f, gstore.f:184-188)
  std::vector<double> gs, rs_all;
  std::vector<int> gsh, rsh_all, r_off, rsh_off, rsh_n, r_nvec;
  const double zero2[2] = {0.0, 0.0};
  kkrh::shells(L.veck, zero2, L.distk, kStrMaxVec, gs, gsh);
  for (int u = 0; u < sc->nkam; ++u) {
    std::vector<double> rs;
    std::vector<int> rsh;
    const double rho[2] = {kam_sep[4 * u + 1], kam_sep[4 * u + 2]};
    kkrh::shells(L.rvec, rho, L.distr, kStrMaxVec, rs, rsh);
    r_off.push_back((int)(rs_all.size() / 2));
    rsh_off.push_back((int)rsh_all.size());
    rsh_n.push_back((int)rsh.size());
    r_nvec.push_back((int)(rs.size() / 2));
    rs_all.insert(rs_all.end(), rs.begin(), rs.end());
    rsh_all.insert(rsh_all.end(), rsh.begin(), rsh.end());
  }
  sc->ngsh = (int)gsh.size();
  sc->ngvec = (int)(gs.size() / 2);
  if (sc->ngsh == 0) return fail(SKKR_ERR_ARG);
  std::vector<double2> g2(gs.size() / 2), r2(rs_all.size() / 2);
  for (size_t i = 0; i < g2.size(); ++i) g2[i] = make_double2(gs[2 * i], gs[2 * i + 1]);
  for (size_t i = 0; i < r2.size(); ++i) r2[i] = make_double2(rs_all[2 * i], rs_all[2 * i + 1]);
  // tables
  const int l2 = 2 * cfg->lmax;
  std::vector<double> yb;
  std::vector<std::complex<double>> ycf;
  kkrh::ylm_tables(l2, yb, ycf);
  std::vector<kkr::cx> ycf2(ycf.size());
  for (size_t i = 0; i < ycf.size(); ++i) ycf2[i] = kkr::cx{ycf[i].real(), ycf[i].imag()};
  std::vector<int> lof(nl2);
  for (int l = 0, i = 0; l <= cfg->lmax; ++l)
    for (int m = -l; m <= l; ++m) lof[i++] = l;
  if ((r = str_upload(sc, &sc->d_kam_sep, kam_sep)) || (r = str_upload(sc, &sc->d_dir_sep, dir_sep)) ||
      (r = str_upload(sc, &sc->d_kam_blk, kam_blk)) || (r = str_upload(sc, &sc->d_dir_blk, dir_blk)) ||
      (r = str_upload(sc, &sc->d_bmap, bmap)) || (r = str_upload(sc, &sc->d_gsorted, g2)) ||
      (r = str_upload(sc, &sc->d_gsh, gsh)) || (r = str_upload(sc, &sc->d_rsorted, r2)) ||
      (r = str_upload(sc, &sc->d_r_off, r_off)) || (r = str_upload(sc, &sc->d_rsh_off, rsh_off)) ||
      (r = str_upload(sc, &sc->d_rsh_n, rsh_n)) || (r = str_upload(sc, &sc->d_rsh, rsh_all)) ||
      (r = str_upload(sc, &sc->d_r_nvec, r_nvec)) ||
      (r = str_upload(sc, &sc->d_gau, kkrh::gaunt_table(cfg->lmax))) ||
      (r = str_upload(sc, &sc->d_flm, kkrh::kambe_flm(l2))) || (r = str_upload(sc, &sc->d_yb, yb)) ||
      (r = str_upload(sc, &sc->d_ycf, ycf2)) || (r = str_upload(sc, &sc->d_lof, lof)))
    return fail(r);
  // batch buffers
  const int B = sc->B, N = sc->N;
  const size_t NN = (size_t)N * N, per = (size_t)nl2 * nl2 * (2 * s + 1) * s * 3;
  if ((r = str_alloc(sc, &sc->d_k, 2 * (size_t)B)) || (r = str_alloc(sc, &sc->d_alpha, (size_t)(cfg->lmax + 1) * s)) ||
      (r = str_alloc(sc, &sc->d_blk, (size_t)B * sc->nblk * nl2 * nl2)) ||
      (r = str_alloc(sc, &sc->d_g, B * NN)) || (r = str_alloc(sc, &sc->d_g1, B * NN)) ||
      (r = str_alloc(sc, &sc->d_gs, B * NN)) || (r = str_alloc(sc, &sc->d_g2d, B * per)) ||
      (r = str_alloc(sc, &sc->d_status, (size_t)B)) || (r = str_ptrs(sc, &sc->p_g, sc->d_g, NN, B)) ||
      (r = str_ptrs(sc, &sc->p_g1, sc->d_g1, NN, B)) || (r = str_ptrs(sc, &sc->p_gs, sc->d_gs, NN, B)) ||
      (r = ensure_inverse(sc->inv, B)))
    return fail(r);
  {
    const size_t r0 = (size_t)cfg->nextra * s * nl2;
    if ((r = str_ptrs(sc, &sc->p_gR, sc->d_g + r0, NN, B)) ||
        (r = str_ptrs(sc, &sc->p_g1C, sc->d_g1 + r0 * N, NN, B)) ||
        (r = str_ptrs(sc, &sc->p_gsRC, sc->d_gs + r0 + r0 * N, NN, B)) ||
        (r = str_ptrs(sc, &sc->p_gsR, sc->d_gs + r0, NN, B)))
      return fail(r);
  }
  if (N > 256) {
    const size_t h1 = 256, h2 = N - 256;
    if ((r = str_alloc(sc, &sc->d_X11, B * h1 * h1)) || (r = str_alloc(sc, &sc->d_C, B * h2 * h1)) ||
        (r = str_alloc(sc, &sc->d_D, B * h1 * h2)) || (r = str_alloc(sc, &sc->d_Si, B * h2 * h2)) ||
        (r = str_ptrs(sc, &sc->p_X11, sc->d_X11, h1 * h1, B)) || (r = str_ptrs(sc, &sc->p_C, sc->d_C, h2 * h1, B)) ||
        (r = str_ptrs(sc, &sc->p_D, sc->d_D, h1 * h2, B)) || (r = str_ptrs(sc, &sc->p_Si, sc->d_Si, h2 * h2, B)) ||
        (r = str_ptrs(sc, &sc->p_A11, sc->d_g1, NN, B)) || (r = str_ptrs(sc, &sc->p_A21, sc->d_g1 + h1, NN, B)) ||
        (r = str_ptrs(sc, &sc->p_A12, sc->d_g1 + h1 * N, NN, B)) ||
        (r = str_ptrs(sc, &sc->p_A22, sc->d_g1 + h1 + h1 * N, NN, B)) ||
        (sc->invS && (r = ensure_inverse(sc->invS, B))))
      return fail(r);
  }
  *out = sc;
  return SKKR_OK;
}

void skkr_str_destroy(skkr_str* sc) {
  if (!sc) return;
  hipSetDevice(sc->cfg.device);
  if (sc->stream) hipStreamSynchronize(sc->stream);
  for (void* p : sc->allocs) hipFree(p);
  if (sc->inv) skkr_destroy(sc->inv);
  if (sc->invS) skkr_destroy(sc->invS);
  delete sc;
}

}  // extern "C"

// In-place inverse of 1 - alpha G for 256 < N <= 512 (on sc->inv's stream):
// 2 x 2 blocks around A11 = the leading 256 rows,
//   X11 = A11^-1, C = A21 X11, S = A22 - C A12, D = X11 A12,
//   B12 = -D S^-1, B21 = -S^-1 C, B11 = X11 - B12 C, B22 = S^-1,
// each block written over the one it replaces once that one is consumed.
// No rows move between the halves: the matrices are those the screening
// transformation inverts, pivoted within the halves.
static int str_inverse_large(skkr_str* sc, int nb) {
  const int N = sc->N, h1 = 256, h2 = N - 256;
  skkr_ctx* e = sc->inv;
  const z one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  InvArgs ia;
  memset(&ia, 0, sizeof(ia));
  ia.n = h1;
  ia.lk = h1;
  ia.ld = N;
  ia.base = sc->p_A11.d;
  ia.out = sc->p_X11.d;
  ia.status = sc->d_status;
  CHK(run_inverse(e, ia, nb));
  CHK(launch_zgemm(e, h2, h1, h1, sc->p_A21.d, N, sc->p_X11.d, h1, sc->p_C.d, h2, one, zero, nb, nullptr));
  CHK(launch_zgemm(e, h2, h2, h1, sc->p_C.d, h2, sc->p_A12.d, N, sc->p_A22.d, N, mone, one, nb, nullptr));
  InvArgs is;
  memset(&is, 0, sizeof(is));
  is.n = h2;
  is.lk = h2;
  is.ld = N;
  is.base = sc->p_A22.d;
  is.out = sc->p_Si.d;
  is.status = sc->d_status;
  if (h2 <= 128) {
    is.ldo = h2;
    CHK(launch_gjinv(e, is, nb));
  } else {
    CHK(run_inverse(sc->invS, is, nb));
  }
  CHK(launch_zgemm(e, h1, h2, h1, sc->p_X11.d, h1, sc->p_A12.d, N, sc->p_D.d, h1, one, zero, nb, nullptr));
  CHK(launch_zgemm(e, h1, h2, h2, sc->p_D.d, h1, sc->p_Si.d, h2, sc->p_A12.d, N, mone, zero, nb, nullptr));
  CHK(launch_zgemm(e, h2, h1, h2, sc->p_Si.d, h2, sc->p_C.d, h2, sc->p_A21.d, N, mone, zero, nb, nullptr));
  {
    dim3 grid((unsigned)(((long long)h1 * h1 + 255) / 256), nb);
    kkr::kkr_copy_block_kernel<<<grid, 256, 0, e->stream>>>(h1, h1, (const double2* const*)sc->p_X11.d, h1,
                                                            sc->p_A11.d, N);
    HIPCHK(hipGetLastError());
    dim3 grid2((unsigned)(((long long)h2 * h2 + 255) / 256), nb);
    kkr::kkr_copy_block_kernel<<<grid2, 256, 0, e->stream>>>(h2, h2, (const double2* const*)sc->p_Si.d, h2,
                                                             sc->p_A22.d, N);
    HIPCHK(hipGetLastError());
  }
  CHK(launch_zgemm(e, h1, h1, h2, sc->p_A12.d, N, sc->p_C.d, h2, sc->p_A11.d, N, mone, one, nb, nullptr));
  return SKKR_OK;
}

// screened constants of nb <= sc->B k-points (device k) into sc->d_gs, on
// stream; status per k in sc->d_status.  Only the rows R of the middle three
// PLs are formed (gs[R, C] = g[R, :] (1 - alpha g)^-1[:, C]): against every
// column for gstore's g2d, against R alone when blocks_only (the k-loop's PL
// blocks).  The other entries of d_gs are left as they were.
static int str_run(skkr_str* sc, hipStream_t stream, double e_re, double e_im, const z* alpha, const double* dk,
                   int nb, bool blocks_only) {
  const int lmax = sc->cfg.lmax, s = sc->cfg.s, nl2 = sc->nl2, N = sc->N;
  kkr::StrArgs a;
  memset(&a, 0, sizeof(a));
  a.nk = nb;
  a.k = dk;
  a.e = kkr::cx{e_re, e_im};
  a.eta = sc->cfg.eta;
  a.vol = sc->lat.vol;
  a.nkam = sc->nkam;
  a.kam_sep = sc->d_kam_sep;
  a.kam_blk = sc->d_kam_blk;
  a.rsorted = sc->d_rsorted;
  a.r_off = sc->d_r_off;
  a.rsh_off = sc->d_rsh_off;
  a.rsh_n = sc->d_rsh_n;
  a.rsh = sc->d_rsh;
  a.r_nvec = sc->d_r_nvec;
  a.gsorted = sc->d_gsorted;
  a.gsh = sc->d_gsh;
  a.ngsh = sc->ngsh;
  a.ngvec = sc->ngvec;
  a.gau = sc->d_gau;
  a.flm = sc->d_flm;
  a.yb = sc->d_yb;
  a.ycf = sc->d_ycf;
  a.ndir = sc->ndir;
  a.dir_sep = sc->d_dir_sep;
  a.dir_blk = sc->d_dir_blk;
  a.blk = sc->d_blk;
  a.nblk = sc->nblk;
  a.status = sc->d_status;
  HIPCHK(hipMemsetAsync(sc->d_status, 0, sizeof(int) * nb, stream));
  if (sc->nkam) {
    dim3 grid(nb, sc->nkam);
    if (lmax == 1) kkr::kkr_kamstr_kernel<1><<<grid, 64, 0, stream>>>(a);
    else if (lmax == 2) kkr::kkr_kamstr_kernel<2><<<grid, 64, 0, stream>>>(a);
    else kkr::kkr_kamstr_kernel<3><<<grid, 64, 0, stream>>>(a);
    HIPCHK(hipGetLastError());
  }
  if (sc->ndir) {
    dim3 grid(nb, sc->ndir);
    if (lmax == 1) kkr::kkr_directg_kernel<1><<<grid, 64, 0, stream>>>(a);
    else if (lmax == 2) kkr::kkr_directg_kernel<2><<<grid, 64, 0, stream>>>(a);
    else kkr::kkr_directg_kernel<3><<<grid, 64, 0, stream>>>(a);
    HIPCHK(hipGetLastError());
  }
  {
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), nb);
    kkr::kkr_assemble_kernel<<<grid, 256, 0, stream>>>(N, nl2, sc->nlay, s, lmax + 1, sc->d_bmap, sc->d_blk,
                                                       sc->nblk, alpha, sc->d_lof, sc->d_g, sc->d_g1, nb);
    HIPCHK(hipGetLastError());
  }
  // the inversion and the product run on the engines' launch paths, on this stream
  hipStream_t own = sc->inv->stream, ownS = sc->invS ? sc->invS->stream : nullptr;
  sc->inv->stream = stream;
  if (sc->invS) sc->invS->stream = stream;
  int r = (N <= 256) ? [&] {
    InvArgs ia;
    memset(&ia, 0, sizeof(ia));
    ia.n = N;
    ia.lk = N;
    ia.base = sc->p_g1.d;
    ia.out = sc->p_g1.d;
    ia.status = sc->d_status;
    return run_inverse(sc->inv, ia, nb);
  }()
                     : str_inverse_large(sc, nb);
  if (r == SKKR_OK) {
    const int mr = 3 * s * nl2;
    r = blocks_only ? launch_zgemm(sc->inv, mr, mr, N, sc->p_gR.d, N, sc->p_g1C.d, N, sc->p_gsRC.d, N,
                                   make_double2(1, 0), make_double2(0, 0), nb, nullptr)
                    : launch_zgemm(sc->inv, mr, N, N, sc->p_gR.d, N, sc->p_g1.d, N, sc->p_gsR.d, N,
                                   make_double2(1, 0), make_double2(0, 0), nb, nullptr);
  }
  sc->inv->stream = own;
  if (sc->invS) sc->invS->stream = ownS;
  return r;
}

// M00, M01, M10 and the self-energy factors of nb k-points of the k-loop
// from the layer producer (replaces the Fourier sum)
static int str_blocks(skkr_ctx* c, int nb) {
  skkr_str* sc = c->str;
  for (int k0 = 0; k0 < nb; k0 += sc->B) {
    const int nbb = std::min(sc->B, nb - k0);
    // the relativistic k-loop evaluates the structure constants at
    // psq = E + E^2/c^2 and scales them by psq / E (rsp-scf/cpacoord_ksave.f:93-99)
    std::complex<double> E(c->e_re, c->e_im), psq = E;
    if (c->str_rel == 1) psq = E + E * E / (274.072 * 274.072);
    const std::complex<double> fac = psq / E;
    CHK(str_run(sc, c->stream, psq.real(), psq.imag(), c->strAlpha, c->dk + 2 * (size_t)k0, nbb, true));
    const size_t nn = c->nn, o = (size_t)k0 * nn;
    dim3 grid((unsigned)((nn + 255) / 256), nbb);
    if (c->str_rel)
      kkr::kkr_pack_rel_kernel<<<grid, 256, 0, c->stream>>>(
          sc->N, sc->nl2, sc->cfg.s, sc->cfg.nextra, sc->d_gs, nbb, (long long)nn, fac.real(), fac.imag(), c->dU1,
          c->dU2, c->dI1, c->dI2, c->M00 + o, c->M01 + o, c->M10 + o, c->MtL + o, c->MtR + o);
    else
      kkr::kkr_pack_kernel<<<grid, 256, 0, c->stream>>>(sc->N, sc->nl2, sc->cfg.s, sc->cfg.nextra, sc->d_gs, nbb,
                                                         (long long)nn, c->M00 + o, c->M01 + o, c->M10 + o,
                                                         c->MtL + o, c->MtR + o);
    HIPCHK(hipGetLastError());
    kkr::kkr_status_kernel<<<(nbb + 255) / 256, 256, 0, c->stream>>>(nbb, sc->d_status, c->status_sw + k0);
    HIPCHK(hipGetLastError());
    c->launches += 2;
  }
  return SKKR_OK;
}

extern "C" {

int skkr_set_structure_kkr_rel(skkr_ctx* c, skkr_str* sc, const double* alpha, int rel) {
  if (!c || rel < 0 || rel > 2) return SKKR_ERR_ARG;
  if (!sc) {
    c->str = nullptr;
    return SKKR_OK;
  }
  const int lk_need = rel ? 2 * sc->nl2 : sc->nl2;
  if (!alpha || c->cfg.lk != lk_need || c->cfg.sites_per_pl != sc->cfg.s || c->cfg.device != sc->cfg.device)
    return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  if (rel && !c->dU1) {
    std::vector<double> u1, u2;
    std::vector<int> i1, i2;
    kkrh::clebsch_tables(u1, u2, i1, i2);
    CHK(dalloc_t(c, &c->dU1, u1.size()));
    CHK(dalloc_t(c, &c->dU2, u2.size()));
    CHK(dalloc_t(c, &c->dI1, i1.size()));
    CHK(dalloc_t(c, &c->dI2, i2.size()));
    HIPCHK(hipMemcpy(c->dU1, u1.data(), sizeof(double) * u1.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->dU2, u2.data(), sizeof(double) * u2.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->dI1, i1.data(), sizeof(int) * i1.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->dI2, i2.data(), sizeof(int) * i2.size(), hipMemcpyHostToDevice));
  }
  c->str_rel = rel;
  return skkr_set_structure_kkr(c, sc, alpha);
}

int skkr_set_structure_kkr(skkr_ctx* c, skkr_str* sc, const double* alpha) {
  if (!c) return SKKR_ERR_ARG;
  if (!sc) {
    c->str = nullptr;
    return SKKR_OK;
  }
  const int lk_need = c->str_rel ? 2 * sc->nl2 : sc->nl2;
  if (!alpha || c->cfg.lk != lk_need || c->cfg.sites_per_pl != sc->cfg.s || c->cfg.device != sc->cfg.device)
    return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  const size_t na = (size_t)(sc->cfg.lmax + 1) * sc->cfg.s;
  if (!c->strAlpha) CHK(dalloc_t(c, &c->strAlpha, na));
  HIPCHK(hipMemcpyAsync(c->strAlpha, alpha, sizeof(z) * na, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->str = sc;
  return SKKR_OK;
}

int skkr_gstore_batch(skkr_str* sc, double e_re, double e_im, const double* alpha, int nk, const double* k,
                      double* g2d, int* status) {
  if (!sc || !alpha || !k || !g2d || nk < 0) return SKKR_ERR_ARG;
  if (nk == 0) return SKKR_OK;
  HIPCHK(hipSetDevice(sc->cfg.device));
  const int lmax = sc->cfg.lmax, s = sc->cfg.s, nl2 = sc->nl2, N = sc->N;
  const size_t per = (size_t)nl2 * nl2 * (2 * s + 1) * s * 3;
  HIPCHK(hipMemcpyAsync(sc->d_alpha, alpha, sizeof(z) * (lmax + 1) * s, hipMemcpyHostToDevice, sc->stream));
  int rc = SKKR_OK;
  std::vector<int> hs;
  for (int k0 = 0; k0 < nk; k0 += sc->B) {
    const int nb = std::min(sc->B, nk - k0);
    HIPCHK(hipMemcpyAsync(sc->d_k, k + 2 * (size_t)k0, sizeof(double) * 2 * nb, hipMemcpyHostToDevice, sc->stream));
    CHK(str_run(sc, sc->stream, e_re, e_im, sc->d_alpha, sc->d_k, nb, false));
    {
      dim3 grid((unsigned)((per + 255) / 256), nb);
      kkr::kkr_extract_kernel<<<grid, 256, 0, sc->stream>>>(N, nl2, s, sc->cfg.nextra, sc->d_gs, sc->d_g2d, nb);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(g2d + 2 * per * (size_t)k0, sc->d_g2d, sizeof(z) * per * nb, hipMemcpyDeviceToHost,
                          sc->stream));
    hs.resize(nb);
    HIPCHK(hipMemcpyAsync(hs.data(), sc->d_status, sizeof(int) * nb, hipMemcpyDeviceToHost, sc->stream));
    HIPCHK(hipStreamSynchronize(sc->stream));
    for (int j = 0; j < nb; ++j) {
      int code = SKKR_OK;
      if (hs[j] & (2 | 4 | 8)) code = SKKR_ERR_STR_NOCONV;  // kamstr G / R, directg
      else if (hs[j] & 1) code = SKKR_ERR_SINGULAR;
      if (status) status[k0 + j] = code;
      if (code != SKKR_OK && rc == SKKR_OK) rc = code;
    }
  }
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------------------
// CPA loop of one energy (nrsp-scf/cpacoord.f:107-345 with fullcpa.f)
// ------------------------------------------------------------------------
// small inversions with full row pivoting (the CPA matrices have no
// definite anti-Hermitian part): the register kernel of the row-wise path
static int launch_gjinv_pivot(skkr_ctx* c, const InvArgs& a, int nb) {
  if (nb <= 0) return SKKR_OK;
  if (a.n <= 32)
    gjinv_f64_kernel<8, 8, 4, 4><<<nb, 64, 0, c->stream>>>(a);
  else if (a.n <= 64)
    gjinv_f64_kernel<16, 16, 4, 4><<<nb, 256, 0, c->stream>>>(a);
  else if (a.n <= 128)
    gjinv_f64_kernel<16, 32, 8, 4><<<nb, 512, 0, c->stream>>>(a);
  else
    return SKKR_ERR_UNSUPPORTED;
  HIPCHK(hipGetLastError());
  c->launches++;
  return SKKR_OK;
}

static int cpa_buffers(skkr_ctx* c) {
  const int S = c->cfg.n_pl * c->ns, lk = c->cfg.lk;
  if (c->cpa_S == S && c->cpa_lk == lk) return SKKR_OK;
  const size_t nn = (size_t)lk * lk, tot = nn * S;
  CHK(dalloc_t(c, &c->dConc, (size_t)S));
  CHK(dalloc_t(c, &c->dErr3, (size_t)S));
  CHK(dalloc_t(c, &c->dFlag, (size_t)2 * S));
  CHK(dalloc_t(c, &c->dInvSt, (size_t)2 * S));
  CHK(dalloc_t(c, &c->dMa, tot));
  CHK(dalloc_t(c, &c->dMb, tot));
  CHK(dalloc_t(c, &c->dDa, tot));
  CHK(dalloc_t(c, &c->dDb, tot));
  CHK(dalloc_t(c, &c->dXab, 2 * tot));
  CHK(dalloc_t(c, &c->dH, tot));
  CHK(dalloc_t(c, &c->dT1, 2 * tot));
  CHK(dalloc_t(c, &c->dTa, tot));
  CHK(dalloc_t(c, &c->dTb, tot));
  auto arr = [&](PtrArr* pa, z* base, int cnt) {
    std::vector<z*> v(cnt);
    for (int j = 0; j < cnt; ++j) v[j] = base + (size_t)j * nn;
    return make_ptrs(c, pa, v);
  };
  CHK(arr(&c->p_ma, c->dMa, S));
  CHK(arr(&c->p_mb, c->dMb, S));
  CHK(arr(&c->p_mc, c->dTint, S));
  CHK(arr(&c->p_da, c->dDa, S));
  CHK(arr(&c->p_db, c->dDb, S));
  CHK(arr(&c->p_xab, c->dXab, 2 * S));
  CHK(arr(&c->p_xa, c->dXab, S));
  CHK(arr(&c->p_xb, c->dXab + tot, S));
  CHK(arr(&c->p_h, c->dH, S));
  CHK(arr(&c->p_t1, c->dT1, 2 * S));
  CHK(arr(&c->p_tauc, c->taubz, S));
  CHK(arr(&c->p_ta, c->dTa, S));
  CHK(arr(&c->p_tb, c->dTb, S));
  c->cpa_S = S;
  c->cpa_lk = lk;
  return SKKR_OK;
}

// tata (fullcpa.f:2-24) into the interface t^-1: mc = mb (c mb + (1-c) ma)^-1 ma
static int cpa_ata(skkr_ctx* c) {
  const int S = c->cpa_S, lk = c->cpa_lk;
  const size_t nn = (size_t)lk * lk;
  const long long tot = (long long)nn * S;
  const z one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  cpa::cpa_ata_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>((int)nn, tot, c->dConc, c->dMa, c->dMb,
                                                                           c->dH);
  HIPCHK(hipGetLastError());
  InvArgs ia;
  memset(&ia, 0, sizeof(ia));
  ia.n = lk;
  ia.lk = lk;
  ia.base = c->p_h.d;
  ia.out = c->p_h.d;
  ia.status = c->dInvSt;  // |= 1 on a zero pivot
  HIPCHK(hipMemsetAsync(c->dInvSt, 0, sizeof(int) * S, c->stream));
  CHK(launch_gjinv_pivot(c, ia, S));
  CHK(launch_zgemm(c, lk, lk, lk, c->p_mb.d, lk, c->p_h.d, lk, c->p_t1.d, lk, one, zero, S, nullptr));
  CHK(launch_zgemm(c, lk, lk, lk, c->p_t1.d, lk, c->p_ma.d, lk, c->p_mc.d, lk, one, zero, S, nullptr));
  std::vector<int> hs(S);
  HIPCHK(hipMemcpyAsync(hs.data(), c->dInvSt, sizeof(int) * S, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int v : hs)
    if (v & 1) return SKKR_ERR_SINGULAR;
  return SKKR_OK;
}

// cpacor / cpasol (fullcpa.f:25-153) on the device; returns the largest
// error3 of the updated sites in *err and SKKR_ERR_SINGULAR on a zero pivot
static int cpa_step(skkr_ctx* c, double* err) {
  const int S = c->cpa_S, lk = c->cpa_lk;
  const size_t nn = (size_t)lk * lk;
  const long long tot = (long long)nn * S;
  const unsigned g = (unsigned)((tot + 255) / 256);
  const z one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  int* flag2 = c->dFlag;  // [flags of the sites | the same again for the xb half]
  cpa::cpa_prep_kernel<<<S, 256, 0, c->stream>>>(lk, c->dConc, c->dMa, c->dMb, c->dTint, flag2, c->dDa, c->dDb,
                                                c->dXab, c->dXab + tot);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(flag2 + S, flag2, sizeof(int) * S, hipMemcpyDeviceToDevice, c->stream));
  int* st = c->dInvSt;
  HIPCHK(hipMemsetAsync(st, 0, sizeof(int) * 2 * S, c->stream));
  InvArgs ia;
  memset(&ia, 0, sizeof(ia));
  ia.n = lk;
  ia.lk = lk;
  ia.base = c->p_xab.d;
  ia.out = c->p_xab.d;
  ia.active = flag2;
  ia.status = st;
  // xa = ((mc - ma)^-1 - tauc)^-1, xb likewise (fullcpa.f:97-106)
  CHK(launch_gjinv_pivot(c, ia, 2 * S));
  cpa::cpa_sub_tau_kernel<<<g, 256, 0, c->stream>>>(tot, c->taubz, c->dXab, c->dXab + tot);
  HIPCHK(hipGetLastError());
  CHK(launch_gjinv_pivot(c, ia, 2 * S));
  // h = ((c xa + (1-c) xb)^-1 + tauc)^-1, mc -= h (fullcpa.f:108-122)
  cpa::cpa_mix_kernel<<<g, 256, 0, c->stream>>>((int)nn, tot, c->dConc, c->dXab, c->dXab + tot, c->dH);
  HIPCHK(hipGetLastError());
  ia.base = c->p_h.d;
  ia.out = c->p_h.d;
  CHK(launch_gjinv_pivot(c, ia, S));
  cpa::cpa_add_tau_kernel<<<g, 256, 0, c->stream>>>(tot, c->taubz, c->dH);
  HIPCHK(hipGetLastError());
  CHK(launch_gjinv_pivot(c, ia, S));
  cpa::cpa_update_kernel<<<S, 256, 0, c->stream>>>(lk, flag2, c->dH, c->taubz, c->dTint, c->dTa, c->dTb, c->dErr3);
  HIPCHK(hipGetLastError());
  // taua = tauc xa tauc + tauc, taub likewise (fullcpa.f:141-148)
  CHK(launch_zgemm(c, lk, lk, lk, c->p_tauc.d, lk, c->p_xa.d, lk, c->p_t1.d, lk, one, zero, S, flag2));
  CHK(launch_zgemm(c, lk, lk, lk, c->p_t1.d, lk, c->p_tauc.d, lk, c->p_ta.d, lk, one, one, S, flag2));
  CHK(launch_zgemm(c, lk, lk, lk, c->p_tauc.d, lk, c->p_xb.d, lk, c->p_t1.d, lk, one, zero, S, flag2));
  CHK(launch_zgemm(c, lk, lk, lk, c->p_t1.d, lk, c->p_tauc.d, lk, c->p_tb.d, lk, one, one, S, flag2));
  std::vector<double> e3(S);
  std::vector<int> hs(2 * S);
  HIPCHK(hipMemcpyAsync(e3.data(), c->dErr3, sizeof(double) * S, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(hs.data(), st, sizeof(int) * 2 * S, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double e = 0.0;
  for (double v : e3) e = std::max(e, v);
  *err = e;
  for (int v : hs)
    if (v & 1) return SKKR_ERR_SINGULAR;
  return SKKR_OK;
}

extern "C" {

int skkr_cpa_bz(skkr_ctx* c, double e_re, double e_im, int nk, const double* k, const double* w,
                const double* conc, const double* tinv_a, const double* tinv_b, double* tinv_c, int init_ata,
                int itcpam, double cpatol, double* tau_c, double* tau_a, double* tau_b, int* itcpa_out,
                double* cpaerr, int* iters_k, int* status_k, skkr_stats* st) {
  if (!c || nk <= 0 || !k || !w || !conc || !tinv_a || !tinv_b || !tinv_c || !tau_c || !tau_a || !tau_b ||
      itcpam < 1 || !(cpatol > 0.0))
    return SKKR_ERR_ARG;
  if (!(c->have_structure || c->str) || !c->have_tinv) return SKKR_ERR_ARG;
  if (c->cfg.lk > 128) return SKKR_ERR_UNSUPPORTED;
  const int S = c->cfg.n_pl * c->ns, lk = c->cfg.lk;
  for (int i = 0; i < S; ++i)
    if (!(conc[i] > 0.0 && conc[i] <= 1.0)) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  // every k resident when the device holds them (blocks, lead self-energies
  // and Y_1 are kept across the CPA passes); else each pass recomputes them
  const double per_k = (32.0 + 2.0 * c->cfg.n_pl) * c->nn * 16.0;
  size_t freeb = 0, totb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totb));
  const double have = (double)freeb + (c->B >= nk ? per_k * c->B : 0.0);
  // (an explicit skkr_config.batch below nk asks for batches: no residency)
  const bool resident = per_k * nk < 0.85 * have && (c->cfg.batch <= 0 || c->cfg.batch >= nk);
  if (resident) {
    CHK(ensure_batch(c, std::max(c->B, nk)));
    if (c->y1_cap < c->B) {
      CHK(dalloc_bt(c, &c->Y1s, (size_t)c->B * c->nn));
      c->y1_cap = c->B;
    }
  }
  CHK(cpa_buffers(c));
  const size_t nn = (size_t)lk * lk, tot = nn * S;
  HIPCHK(hipMemcpyAsync(c->dConc, conc, sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dMa, tinv_a, sizeof(z) * tot, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dMb, tinv_b, sizeof(z) * tot, hipMemcpyHostToDevice, c->stream));
  if (init_ata)
    CHK(cpa_ata(c));
  else
    HIPCHK(hipMemcpyAsync(c->dTint, tinv_c, sizeof(z) * tot, hipMemcpyHostToDevice, c->stream));
  c->e_re = e_re;
  c->e_im = e_im;
  std::vector<double> hbz(2 * tot);
  int itcpa = 1, rc = SKKR_OK;
  double err = 0.0;
  for (;;) {
    c->cpa_mode = resident ? (itcpa == 1 ? 1 : 2) : 0;
    std::fill(hbz.begin(), hbz.end(), 0.0);
    rc = tau_bz_run(c, nk, k, w, hbz.data(), nullptr, iters_k, status_k, st, false, nullptr, nullptr);
    if (rc != SKKR_OK) break;
    rc = cpa_step(c, &err);
    if (rc != SKKR_OK) break;
    const bool concpa = err > cpatol, cpacon = itcpa < itcpam;
    ++itcpa;
    if (!(concpa && cpacon)) break;
  }
  c->cpa_mode = 0;
  if (itcpa_out) *itcpa_out = itcpa - 1;
  if (cpaerr) *cpaerr = err;
  memcpy(tau_c, hbz.data(), sizeof(z) * tot);
  HIPCHK(hipMemcpyAsync(tinv_c, c->dTint, sizeof(z) * tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(tau_a, c->dTa, sizeof(z) * tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(tau_b, c->dTb, sizeof(z) * tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return rc;
}


int skkr_set_pairs(skkr_ctx* c, int npair, const int* pairs, const double* a1, const double* a2) {
  if (!c || npair < 0 || (npair > 0 && (!pairs || !a1 || !a2))) return SKKR_ERR_ARG;
  const int S = c->cfg.n_pl * c->ns, lk = c->cfg.lk;
  for (int q = 0; q < npair; ++q)
    if (pairs[4 * q] < 1 || pairs[4 * q] > S || pairs[4 * q + 1] < 1 || pairs[4 * q + 1] > S) return SKKR_ERR_ARG;
  HIPCHK(hipSetDevice(c->cfg.device));
  c->pair_li.resize(npair);
  c->pair_lj.resize(npair);
  c->pair_rx.resize(npair);
  c->pair_ry.resize(npair);
  for (int q = 0; q < npair; ++q) {
    c->pair_li[q] = pairs[4 * q] - 1;
    c->pair_lj[q] = pairs[4 * q + 1] - 1;
    const double n1 = pairs[4 * q + 2], n2 = pairs[4 * q + 3];
    c->pair_rx[q] = n1 * a1[0] + n2 * a2[0];
    c->pair_ry[q] = n1 * a1[1] + n2 * a2[1];
  }
  if (npair > c->pair_cap) {
    CHK(dalloc_t(c, &c->dTij, (size_t)npair * lk * lk));
    CHK(dalloc_t(c, &c->dTji, (size_t)npair * lk * lk));
    c->pair_cap = npair;
  }
  return SKKR_OK;
}

int skkr_tau_bz_pairs(skkr_ctx* c, double e_re, double e_im, int nk, const double* k, const double* w,
                      double* tau_bz, double* tau_ij, double* tau_ji, int* iters_k, int* status_k,
                      skkr_stats* st) {
  if (!c || !tau_ij || !tau_ji) return SKKR_ERR_ARG;
  const int np = (int)c->pair_li.size(), lk = c->cfg.lk;
  if (np == 0) return SKKR_ERR_ARG;
  if (c->ng_sym > 1) return SKKR_ERR_UNSUPPORTED;  // pair sums need the full BZ
  HIPCHK(hipSetDevice(c->cfg.device));
  const size_t ne = (size_t)np * lk * lk;
  HIPCHK(hipMemsetAsync(c->dTij, 0, sizeof(z) * ne, c->stream));
  HIPCHK(hipMemsetAsync(c->dTji, 0, sizeof(z) * ne, c->stream));
  c->e_re = e_re;
  c->e_im = e_im;
  c->pairs_on = true;
  const int rc = tau_bz_run(c, nk, k, w, tau_bz, nullptr, iters_k, status_k, st, false, nullptr, nullptr);
  c->pairs_on = false;
  std::vector<double> hij(2 * ne), hji(2 * ne);
  HIPCHK(hipMemcpyAsync(hij.data(), c->dTij, sizeof(z) * ne, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(hji.data(), c->dTji, sizeof(z) * ne, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (size_t e = 0; e < 2 * ne; ++e) {
    tau_ij[e] += hij[e];
    tau_ji[e] += hji[e];
  }
  return rc;
}

}  // extern "C"
