// FP64 device kernels of the per-(E,k) layer kernel (gfx950).
//
// Matrices are complex*16 column-major, one n x n block per k-point (or per
// decimation problem), addressed through per-batch pointers or strides.
#pragma once
#include <hip/hip_runtime.h>

typedef __attribute__((ext_vector_type(4))) double dx4;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
// 1/q with the scaling of Smith's method (no overflow for large |q|)
__device__ __forceinline__ double2 crecip(double2 q) {
  if (fabs(q.x) >= fabs(q.y)) {
    double r = q.y / q.x, d = q.x + q.y * r;
    return make_double2(1.0 / d, -r / d);
  } else {
    double r = q.x / q.y, d = q.x * r + q.y;
    return make_double2(r / d, -1.0 / d);
  }
}

// ---------------------------------------------------------------------------
// Lattice Fourier sum: out[b][e] = scale * sum_r GR[r][e] * exp(i*sgn*(k_b.R_r))
// One thread per matrix element e; GR[.][e] held in registers; the phases of
// KB k-points staged in LDS; consecutive threads write consecutive elements.
// ---------------------------------------------------------------------------
template <int MAXR, int KB>
__global__ __launch_bounds__(256) void fourier_kernel(
    int nn, int nR, const double2* __restrict__ GR, const double* __restrict__ R,
    const double* __restrict__ kpts, int nk, double sgn, double scale,
    double2* __restrict__ out, long long out_stride) {
  __shared__ double2 ph[KB * MAXR];
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int k0 = blockIdx.y * KB;
  for (int t = threadIdx.x; t < KB * nR; t += 256) {
    int kb = t / nR, r = t % nR;
    int kk = k0 + kb;
    double phi = 0.0;
    if (kk < nk) phi = kpts[2 * kk] * R[2 * r] + kpts[2 * kk + 1] * R[2 * r + 1];
    double sn, cs;
    sincos(phi, &sn, &cs);
    ph[kb * MAXR + r] = make_double2(cs, sgn * sn);
  }
  __syncthreads();
  if (e >= nn) return;
  double2 g[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r)
    g[r] = (r < nR) ? GR[(long long)r * nn + e] : make_double2(0.0, 0.0);
  const int kend = min(KB, nk - k0);
  for (int kb = 0; kb < kend; ++kb) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      if (r < nR) {
        double2 p = ph[kb * MAXR + r];
        acc.x += g[r].x * p.x - g[r].y * p.y;
        acc.y += g[r].x * p.y + g[r].y * p.x;
      }
    }
    out[(long long)(k0 + kb) * out_stride + e] = make_double2(scale * acc.x, scale * acc.y);
  }
}

// Epilogues of the decimation products (surfgf.f:62-81), applied per element
// of the 2n x n result [ay; by] * beta (DEC_BETA) or [ay; by] * alpha
// (DEC_ALPHA):
//   DEC_BETA : top rows  y3 = ay*beta : eps -= y3, x -= y3
//              bottom    beta' = by*beta -> ABn bottom
//   DEC_ALPHA: top rows  alpha' = ay*alpha -> ABn top, |alpha'|^2 partial
//              bottom    y4 = by*alpha : eps -= y4
// DEC_BETA runs first, so eps is updated as (eps - y3) - y4.
enum { EPI_PLAIN = 0, EPI_DEC_BETA = 1, EPI_DEC_ALPHA = 2 };
struct EpiArgs {
  int mode;
  int n;                       // chain block size (rows [0,n) are the top half)
  double2* const* eps;
  double2* const* x;
  double2* const* abn;         // next [alpha; beta], 2n x n, ld 2n
  double* norm;                // per (chain, work-group) partial |alpha'|^2
  int tiles;                   // work-groups per chain
};

// ---------------------------------------------------------------------------
// Batched complex GEMM on FP64 MFMA:
//   C[b] = alpha * A[b](m x k) * B[b](k x n) + beta * C[b]
// Work-group tile TM x TN, KC-deep LDS stages, each wave a 32 x 16*WT
// sub-tile of 2 x WT v_mfma_f64_16x16x4 blocks.  The MFMA computes the
// transposed block (B^T A^T) so that lanes 0..15 own 16 consecutive rows of C
// and the stores of a column are contiguous.
// Complex product, M3 = false: four real products into two accumulators,
//   re += ar*br - ai*bi, im += ar*bi + ai*br;
// M3 = true: three real products (Gauss) into three accumulators,
//   p1 += ar*br, p2 += ai*bi, p3 += (ar+ai)*(br+bi);  re = p1 - p2,
//   im = p3 - p1 - p2  (3/4 of the MFMA work; normwise FP64 accuracy).
// ---------------------------------------------------------------------------
// Global (not flat) loads and stores of a complex value: they count on the
// vector-memory counter only, so a store waits for nothing that LDS does
typedef double __attribute__((address_space(1))) gdbl_t;
__device__ __forceinline__ double2 gload2(const double2* p) {
  const gdbl_t* q = (const gdbl_t*)(const double*)p;
  return make_double2(q[0], q[1]);
}
__device__ __forceinline__ void gstore2(double2* p, double2 v) {
  gdbl_t* q = (gdbl_t*)(double*)p;
  q[0] = v.x;
  q[1] = v.y;
}

template <int TM, int TN, int KC, int WT = 2, bool M3 = false>
__global__ __launch_bounds__((TM / 32) * (TN / (16 * WT)) * 64) void zgemm_f64_kernel(
    int m, int n, int k, const double2* const* __restrict__ Ap, int lda,
    const double2* const* __restrict__ Bp, int ldb, double2* const* __restrict__ Cp,
    int ldc, double2 alpha, double2 beta, const int* __restrict__ active, EpiArgs epi) {
  constexpr int WTN = 16 * WT;               // wave tile width
  constexpr int NW = (TM / 32) * (TN / WTN);
  constexpr int NT = NW * 64;
  // double-buffered LDS stages; the next stage is fetched into registers while
  // the MFMAs of the current one run
  __shared__ double2 As[2][KC][TM];
  __shared__ double2 Bs[2][TN][KC + 1];
  constexpr int LA = (TM * KC + NT - 1) / NT;  // A elements per thread per stage
  constexpr int LB = (TN * KC + NT - 1) / NT;
  // Work-groups go to the 8 XCDs round-robin by linear id.  Renumber them so
  // that the tiles of one product land on one XCD and share its L2 (A is
  // read by every column tile): linear id L = 8*tiles*g + 8*t + r is tile t
  // of product 8g + r.  The last nb % 8 products keep the plain numbering.
  int b = blockIdx.y, tile = blockIdx.x;
  {
    const int tiles = gridDim.x;
    const int L = blockIdx.x + tiles * blockIdx.y;
    if (tiles > 1 && L < 8 * tiles * (int)(gridDim.y / 8)) {
      const int g = L / (8 * tiles), rem = L % (8 * tiles);
      tile = rem / 8;
      b = 8 * g + rem % 8;
    }
  }
  if (active && !active[b]) return;
  const int tiles_m = (m + TM - 1) / TM;
  const int row0 = (tile % tiles_m) * TM;
  const int col0 = (tile / tiles_m) * TN;
  const double2* __restrict__ A = Ap[b];
  const double2* __restrict__ B = Bp[b];
  double2* __restrict__ C = Cp ? Cp[b] : nullptr;  // epilogue modes write elsewhere
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w % (TM / 32), wn = w / (TM / 32);
  constexpr int P2 = M3 ? WT : 1;
  dx4 acc_re[2][WT], acc_im[2][WT], acc_p2[2][P2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < WT; ++j) {
      acc_re[i][j] = (dx4){0.0, 0.0, 0.0, 0.0};
      acc_im[i][j] = (dx4){0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int j = 0; j < P2; ++j) acc_p2[i][j] = (dx4){0.0, 0.0, 0.0, 0.0};
  }
  const int l16 = lane & 15, l4 = lane >> 4;
  double2 ra[LA], rb[LB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < LA; ++u) {
      const int idx = tid + NT * u;
      const int r = idx % TM, kk = idx / TM;
      const int gr = row0 + r, gk = k0 + kk;
      ra[u] = (idx < TM * KC && gr < m && gk < k) ? A[gr + (long long)gk * lda]
                                                   : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int idx = tid + NT * u;
      const int kk = idx % KC, c = idx / KC;
      const int gk = k0 + kk, gc = col0 + c;
      rb[u] = (idx < TN * KC && gk < k && gc < n) ? B[gk + (long long)gc * ldb]
                                                   : make_double2(0.0, 0.0);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < LA; ++u) {
      const int idx = tid + NT * u;
      if (idx < TM * KC) As[buf][idx / TM][idx % TM] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int idx = tid + NT * u;
      if (idx < TN * KC) Bs[buf][idx / KC][idx % KC] = rb[u];
    }
  };
  fetch(0);
  stash(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < k; k0 += KC) {
    const bool more = (k0 + KC < k);
    if (more) fetch(k0 + KC);
#pragma unroll
    for (int ks = 0; ks < KC; ks += 4) {
      double2 af[2], bf[WT];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = As[buf][ks + l4][wm * 32 + i * 16 + l16];
#pragma unroll
      for (int j = 0; j < WT; ++j) bf[j] = Bs[buf][wn * WTN + j * 16 + l16][ks + l4];
      // transposed block: MFMA A-operand = B^T (lane -> col of C),
      // B-operand = A^T (lane -> row of C)
      if constexpr (M3) {
        double as[2], bs[WT];
#pragma unroll
        for (int i = 0; i < 2; ++i) as[i] = af[i].x + af[i].y;
#pragma unroll
        for (int j = 0; j < WT; ++j) bs[j] = bf[j].x + bf[j].y;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < WT; ++j) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j].x, af[i].x, acc_re[i][j], 0, 0, 0);
            acc_p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j].y, af[i].y, acc_p2[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bs[j], as[i], acc_im[i][j], 0, 0, 0);
          }
      } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < WT; ++j) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j].x, af[i].x, acc_re[i][j], 0, 0, 0);
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bf[j].y, af[i].y, acc_re[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j].x, af[i].y, acc_im[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j].y, af[i].x, acc_im[i][j], 0, 0, 0);
          }
      }
    }
    if (more) stash(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  if constexpr (M3) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < WT; ++j) {
        const dx4 p1 = acc_re[i][j], p2 = acc_p2[i][j];
        acc_re[i][j] = p1 - p2;
        acc_im[i][j] = acc_im[i][j] - p1 - p2;
      }
  }
  // v_mfma_f64_16x16x4 result layout: lane l, register q holds D[l/16 + 4q][l%16];
  // D is the transposed block, so C row (l % 16), C column (l/16 + 4q)
  if (epi.mode == EPI_PLAIN) {
    const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < WT; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int gr = row0 + wm * 32 + i * 16 + l16;
          int gc = col0 + wn * WTN + j * 16 + l4 + 4 * q;
          if (gr < m && gc < n) {
            double2 v = make_double2(acc_re[i][j][q], acc_im[i][j][q]);
            double2 r = cmul(alpha, v);
            long long o = gr + (long long)gc * ldc;
            if (use_beta) r = cadd(r, cmul(beta, C[o]));
            C[o] = r;
          }
        }
    return;
  }
  const int hn = epi.n;
  double2* __restrict__ eps = epi.eps[b];
  double2* __restrict__ xx = epi.x[b];
  double2* __restrict__ abn = epi.abn[b];
  double part = 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < WT; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gr = row0 + wm * 32 + i * 16 + l16;
        const int gc = col0 + wn * WTN + j * 16 + l4 + 4 * q;
        if (gr < m && gc < n) {
          const double2 v = make_double2(acc_re[i][j][q], acc_im[i][j][q]);
          if (gr < hn) {
            const long long o = gr + (long long)gc * hn;
            if (epi.mode == EPI_DEC_BETA) {
              eps[o] = csub(eps[o], v);
              xx[o] = csub(xx[o], v);
            } else {
              abn[gr + (long long)gc * 2 * hn] = v;
              part += v.x * v.x + v.y * v.y;
            }
          } else {
            const long long o = (gr - hn) + (long long)gc * hn;
            if (epi.mode == EPI_DEC_BETA)
              abn[gr + (long long)gc * 2 * hn] = v;
            else
              eps[o] = csub(eps[o], v);
          }
        }
      }
  if (epi.mode == EPI_DEC_ALPHA) {
    // deterministic work-group sum of |alpha'|^2 (tree in LDS, fixed order);
    // a tail above the largest power of two P <= NT (9 waves) is folded first
    constexpr int P = (NT >= 512) ? 512 : (NT >= 256) ? 256 : (NT >= 128) ? 128 : 64;
    static_assert(NT >= 64 && NT - P <= P, "work-group size");
    __shared__ double red[NT];
    red[tid] = part;
    __syncthreads();
    if (P < NT) {
      if (tid < NT - P) red[tid] += red[tid + P];
      __syncthreads();
    }
    for (int st = P / 2; st > 0; st >>= 1) {
      if (tid < st) red[tid] += red[tid + st];
      __syncthreads();
    }
    if (tid == 0) epi.norm[(long long)b * epi.tiles + tile] = red[0];
  }
}

// ---------------------------------------------------------------------------
// Batched in-place Gauss-Jordan inversion, register resident, n <= TR*RPT and
// n <= TC*CPT.  Thread (tr, tc) owns rows tr + TR*i and columns tc + TC*j.
// Partial pivoting over unused rows of the current column with implicit row
// exchange: with pivot row pi(i) for column i the stored S gives
// A^-1[pinv(rho)][pi(gamma)] = S[rho][gamma].
//
// The matrix is assembled on load:
//   a = base + blockdiag(tinv) - sub1 - sub2      (any term optional)
// Epilogue: mode 0 writes the full inverse to out; mode 1 writes only the
// site-diagonal lk x lk blocks to tau (lk*lk per site, sites consecutive) and
// the traces of those blocks to tr.
// ---------------------------------------------------------------------------
struct InvArgs {
  int n, lk;
  int ld, ldo;                // leading dimensions of the inputs / of out (0 = n)
  const double2* const* base; // per batch (may repeat)
  const double2* tinv;        // site blocks lk*lk, shared by the batch (or null)
  const double2* const* sub1; // or null
  const double2* const* sub2; // or null
  double2* const* out;        // mode 0
  double2* const* tau;        // mode 1
  double2* const* tr;         // mode 1 (n/lk per batch)
  const int* active;          // or null
  int* status;                // or null; |= 1 on zero pivot
  int mode;
};

template <int TR, int TC, int RPT, int CPT>
__global__ __launch_bounds__(TR * TC) void gjinv_f64_kernel(InvArgs args) {
  constexpr int NT = TR * TC;
  constexpr int NMAX = (TR * RPT < TC * CPT) ? TR * RPT : TC * CPT;
  __shared__ double2 colbuf[2][NMAX];
  __shared__ double2 rowbuf[2][NMAX];
  __shared__ int rowused[NMAX];
  __shared__ int perm[NMAX];
  __shared__ int pinv[NMAX];
  __shared__ double2 diag[NMAX];
  const int b = blockIdx.x;
  if (args.active && !args.active[b]) return;
  const int n = args.n;
  const int tid = threadIdx.x;
  const int tr = tid % TR, tc = tid / TR;
  const int lane = tid & 63;
  double2 a[RPT][CPT];
  {
    const double2* __restrict__ base = args.base[b];
    const double2* __restrict__ s1 = args.sub1 ? args.sub1[b] : nullptr;
    const double2* __restrict__ s2 = args.sub2 ? args.sub2[b] : nullptr;
    const int lk = args.lk;
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        int r = tr + TR * i, c = tc + TC * j;
        double2 v = make_double2(0.0, 0.0);
        if (r < n && c < n) {
          const int ld = args.ld ? args.ld : n;
          long long o = r + (long long)c * ld;
          v = base[o];
          if (args.tinv && (r / lk) == (c / lk)) {
            int st = r / lk;
            v = cadd(args.tinv[(long long)st * lk * lk + (r % lk) + (c % lk) * lk], v);
          }
          if (s1) v = csub(v, s1[o]);
          if (s2) v = csub(v, s2[o]);
        }
        a[i][j] = v;
      }
  }
  for (int t = tid; t < NMAX; t += NT) rowused[t] = (t < n) ? 0 : 1;
  int singular = 0;
  // column 0 into colbuf[0]
#pragma unroll
  for (int j = 0; j < CPT; ++j)
    if (tc + TC * j == 0)
#pragma unroll
      for (int i = 0; i < RPT; ++i)
        if (tr + TR * i < n) colbuf[0][tr + TR * i] = a[i][j];
  __syncthreads();
  for (int it = 0; it < n; ++it) {
    const int cur = it & 1;
    // pivot: every wave computes argmax over unused rows (same result)
    double best = -1.0;
    int bi = NMAX;
    for (int r = lane; r < n; r += 64) {
      if (!rowused[r]) {
        double2 v = colbuf[cur][r];
        double mg = v.x * v.x + v.y * v.y;
        if (mg > best) { best = mg; bi = r; }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double ob = __shfl_xor(best, off);
      int oi = __shfl_xor(bi, off);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    int p = bi;
    if (p >= n) {  // only when every candidate is NaN: take the first unused row
      p = 0;
      while (p < n - 1 && rowused[p]) ++p;
      singular = 1;
    }
    const double2 piv = colbuf[cur][p];
    if (piv.x == 0.0 && piv.y == 0.0) singular = 1;
    const double2 q = crecip(piv);
    // scale pivot row, publish it
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      if (tr + TR * i == p) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
          int c = tc + TC * j;
          if (c < n) {
            double2 v = (c == it) ? q : cmul(a[i][j], q);
            a[i][j] = v;
            rowbuf[cur][c] = v;
          }
        }
      }
    }
    __syncthreads();
    // eliminate every other row; owners of column it+1 publish it
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = tr + TR * i;
      if (r < n && r != p) {
        const double2 f = colbuf[cur][r];
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
          const int c = tc + TC * j;
          if (c < n) {
            const double2 rb = rowbuf[cur][c];
            double2 v = (c == it) ? make_double2(0.0, 0.0) : a[i][j];
            v.x -= f.x * rb.x - f.y * rb.y;
            v.y -= f.x * rb.y + f.y * rb.x;
            a[i][j] = v;
          }
        }
      }
    }
    if (it + 1 < n) {
#pragma unroll
      for (int j = 0; j < CPT; ++j)
        if (tc + TC * j == it + 1)
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (tr + TR * i < n) colbuf[cur ^ 1][tr + TR * i] = a[i][j];
    }
    if (tid == 0) {
      rowused[p] = 1;
      perm[it] = p;
      pinv[p] = it;
    }
    __syncthreads();
  }
  if (singular && args.status && tid == 0) atomicOr(&args.status[b], 1);
  if (args.mode == 0) {
    double2* __restrict__ out = args.out[b];
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        int rho = tr + TR * i, gam = tc + TC * j;
        if (rho < n && gam < n) out[pinv[rho] + (long long)perm[gam] * (args.ldo ? args.ldo : n)] = a[i][j];
      }
  } else {
    const int lk = args.lk;
    double2* __restrict__ tau = args.tau[b];
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        int rho = tr + TR * i, gam = tc + TC * j;
        if (rho < n && gam < n) {
          int rr = pinv[rho], cc = perm[gam];
          if (rr / lk == cc / lk) {
            int st = rr / lk;
            tau[(long long)st * lk * lk + (rr % lk) + (cc % lk) * lk] = a[i][j];
            if (rr == cc) diag[rr] = a[i][j];
          }
        }
      }
    __syncthreads();
    const int ns = n / lk;
    for (int st = tid; st < ns; st += NT) {
      double2 s = make_double2(0.0, 0.0);
      for (int l = 0; l < lk; ++l) s = cadd(s, diag[st * lk + l]);
      args.tr[b][st] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Decimation helpers (surfgf.f): one work-group per chain.
// ---------------------------------------------------------------------------
// init: AB = [-ml; -mr] (2n x n, ld 2n), eps = x = m0 + blockdiag(tinv)
__global__ __launch_bounds__(256) void dec_init_kernel(
    int n, int lk, const double2* const* ml, const double2* const* mr,
    const double2* const* m0, const double2* const* tinv, double2* const* AB,
    double2* const* eps, double2* const* x, const int* active, int* act_hi, int* act_lo,
    int* nlate, int* act_inv, int* nkeep) {
  const int b = blockIdx.x;
  if (act_hi && threadIdx.x == 0) {  // every started chain begins at the full digit count
    act_hi[b] = active[b];
    act_lo[b] = 0;
    nlate[b] = 0;
  }
  if (act_inv && threadIdx.x == 0) {
    act_inv[b] = active[b];
    nkeep[b] = 0;
  }
  const double2* __restrict__ L = ml[b];
  const double2* __restrict__ Rm = mr[b];
  const double2* __restrict__ M0 = m0[b];
  const double2* __restrict__ T = tinv ? tinv[b] : nullptr;
  for (int e = threadIdx.x; e < n * n; e += 256) {
    int r = e % n, c = e / n;
    double2 v = M0[e];
    if (T && r / lk == c / lk)
      v = cadd(T[(long long)(r / lk) * lk * lk + (r % lk) + (c % lk) * lk], v);
    eps[b][e] = v;
    x[b][e] = v;
    AB[b][r + (long long)c * 2 * n] = make_double2(-L[e].x, -L[e].y);
    AB[b][n + r + (long long)c * 2 * n] = make_double2(-Rm[e].x, -Rm[e].y);
  }
}

// Convergence bookkeeping of surfgf.f:76-103 per chain after an iteration:
// err = |alpha'|_F from the work-group partials (summed in order); chains
// b < nhalf are left leads (itermax_l), the others right leads (itermax_r).
// With act_hi, the digit count of the chain's next iteration: the reduced
// one (act_lo) once err <= late_ratio times its err after iteration 1 (nref).
// With act_inv, whether the next iteration inverts eps: not once err <=
// inv_ratio times nref (the iteration then keeps the previous eps^-1).
__global__ __launch_bounds__(256) void dec_check_kernel(
    int nchains, int tiles, int iter, int itermax_l, int itermax_r, int nhalf, double errmax,
    const double* __restrict__ norm, int* active, int* iters, int* status, int* n_active,
    double late_ratio, double* nref, int* act_hi, int* act_lo, int* nlate, double inv_ratio,
    int* act_inv, int* nkeep) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nchains || !active[b]) return;
  double s = 0.0;
  for (int t = 0; t < tiles; ++t) s += norm[(long long)b * tiles + t];
  const double err = sqrt(s);
  const int itermax = (b < nhalf) ? itermax_l : itermax_r;
  if (err < errmax || iter > itermax) {
    active[b] = 0;
    iters[b] = iter;
    if (iter > itermax) status[b] |= 4;
  } else if (err > 1.0e10) {
    active[b] = 0;
    iters[b] = iter;
    status[b] |= 2;
  } else {
    atomicAdd(n_active, 1);
    if (iter == 1 && nref) nref[b] = err;
    if (act_hi) {
      const int lo = err <= late_ratio * nref[b];
      act_hi[b] = !lo;
      act_lo[b] = lo;
      nlate[b] += lo;
    }
    if (act_inv) {
      const int keep = err <= inv_ratio * nref[b];
      act_inv[b] = !keep;
      nkeep[b] += keep;
    }
    return;
  }
  if (act_hi) act_hi[b] = act_lo[b] = 0;
  if (act_inv) act_inv[b] = 0;
}

// Assembly a = base + blockdiag(tinv) - sub1 - sub2 into a dense n x n buffer
// (the load stage of the inversion kernels, for sizes inverted by blocks).
__global__ __launch_bounds__(256) void assemble_kernel(InvArgs args, double2* const* dst) {
  const int b = blockIdx.y;
  if (args.active && !args.active[b]) return;
  const int n = args.n, lk = args.lk, ld = args.ld ? args.ld : n;
  const long long e = blockIdx.x * 256ll + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const long long o = r + (long long)c * ld;
  double2 v = args.base[b][o];
  if (args.tinv && r / lk == c / lk)
    v = cadd(args.tinv[(long long)(r / lk) * lk * lk + (r % lk) + (c % lk) * lk], v);
  if (args.sub1) v = csub(v, args.sub1[b][o]);
  if (args.sub2) v = csub(v, args.sub2[b][o]);
  dst[b][e] = v;
}

// out[b] = in[b] + off (device-side views of sub-blocks)
__global__ void ptr_offset_kernel(int nb, double2* const* in, long long off, double2** out) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b < nb) out[b] = in[b] + off;
}

// ---------------------------------------------------------------------------
// Site-diagonal lk x lk blocks of an n x n matrix and their traces (the part
// of the PL-diagonal tau that the BZ sum keeps, tau2d.f:127-141).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void extract_sites_kernel(
    int n, int lk, const double2* const* G, double2* const* tau, double2* const* tr) {
  const int b = blockIdx.x;
  const double2* __restrict__ g = G[b];
  double2* __restrict__ t = tau[b];
  const int ns = n / lk;
  for (int e = threadIdx.x; e < ns * lk * lk; e += 256) {
    const int st = e / (lk * lk), rem = e % (lk * lk);
    const int l = rem % lk, lp = rem / lk;
    t[e] = g[(st * lk + l) + (long long)(st * lk + lp) * n];
  }
  for (int st = threadIdx.x; st < ns; st += 256) {
    double2 s = make_double2(0.0, 0.0);
    for (int l = 0; l < lk; ++l) s = cadd(s, g[(st * lk + l) + (long long)(st * lk + l) * n]);
    tr[b][st] = s;
  }
}

// ---------------------------------------------------------------------------
// BZ accumulation: acc[e] += sum_b w[b] * tau[b][e], b in order (deterministic)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bz_accumulate_kernel(
    int ne, int nb, const double2* __restrict__ tau, long long stride,
    const double* __restrict__ w, double2* __restrict__ acc) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  double2 s = acc[e];
  // ne is small (sites x lk^2, a few work-groups): issue 16 loads before the
  // adds, which stay in k order, so the sum is the same as one k at a time
  constexpr int U = 16;
  int b = 0;
  for (; b + U <= nb; b += U) {
    double2 t[U];
    double wb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      t[u] = tau[(long long)(b + u) * stride + e];
      wb[u] = w[b + u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      s.x += wb[u] * t[u].x;
      s.y += wb[u] * t[u].y;
    }
  }
  for (; b < nb; ++b) {
    double2 t = tau[(long long)b * stride + e];
    s.x += w[b] * t.x;
    s.y += w[b] * t.y;
  }
  acc[e] = s;
}

// Off-diagonal BZ sums of site pairs: acc[e] += sum_b w[b] e^{sgn i k_b.R}
// T_b[i lk + a, j lk + c], e = a + lk c, b in order (deterministic); T_b are
// the batch's n x n blocks of tau between the two principal layers.
__global__ __launch_bounds__(256) void pair_accumulate_kernel(
    int lk, int n, int nb, const double2* const* __restrict__ T, int i, int j, double rx, double ry,
    double sgn, const double* __restrict__ kpts, const double* __restrict__ w, double2* __restrict__ acc) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= lk * lk) return;
  const int a = e % lk, c = e / lk;
  const long long off = (long long)(i * lk + a) + (long long)(j * lk + c) * n;
  double2 s = acc[e];
  auto add = [&](int b, double2 t) {
    double sn, cs;
    sincos(kpts[2 * b] * rx + kpts[2 * b + 1] * ry, &sn, &cs);
    const double pr = w[b] * cs, pi = w[b] * sgn * sn;
    s.x += pr * t.x - pi * t.y;
    s.y += pr * t.y + pi * t.x;
  };
  // as in bz_accumulate_kernel: 8 loads (through the pointer array) ahead of
  // their adds, which stay in k order
  constexpr int U = 8;
  int b = 0;
  for (; b + U <= nb; b += U) {
    double2 t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = T[b + u][off];
#pragma unroll
    for (int u = 0; u < U; ++u) add(b + u, t[u]);
  }
  for (; b < nb; ++b) add(b, T[b][off]);
  acc[e] = s;
}

__global__ __launch_bounds__(256) void scale_copy_kernel(
    long long ne, const double2* __restrict__ src, double scale, double2* __restrict__ dst) {
  long long e = blockIdx.x * 256ll + threadIdx.x;
  if (e < ne) dst[e] = make_double2(scale * src[e].x, scale * src[e].y);
}

// ---------------------------------------------------------------------------
// Screened -> physical tau of a site block (phystau.f:30-58).  With
// A = diag(a), a_i the screening parameter of basis index i, and T = t^-1 of
// the site:
//   phys = (1 + A T) tau (1 + T A) - (A + A T A).
// phys_trace_kernel: per (site, k) the sums over basis indices of one l of
// the diagonal of phys (taukdiag of cpacoord.f:300-309), from
//   phys_ii = tau_ii + a_i ((T tau)_ii + (tau T)_ii) + a_i^2 (T tau T)_ii
//             - a_i - a_i^2 T_ii.
// One wave per (site, k); thread i < lk owns diagonal element i; the per-l
// sums are taken in index order.
// ---------------------------------------------------------------------------
template <int MAXLK>
__global__ __launch_bounds__(64) void phys_trace_kernel(
    int lk, int nl, int sites, const double2* __restrict__ tauk, long long kstride,
    const double2* __restrict__ tint, const double2* __restrict__ a,
    const int* __restrict__ lidx, double2* __restrict__ out) {
  __shared__ double2 st[MAXLK * MAXLK];  // tau, column-major
  __shared__ double2 sT[MAXLK * MAXLK];  // T, column-major
  __shared__ double2 sd[MAXLK];
  const int site = blockIdx.x, b = blockIdx.y;
  const long long nn = (long long)lk * lk;
  const double2* __restrict__ tau = tauk + (long long)b * kstride + site * nn;
  const double2* __restrict__ T = tint + site * nn;
  for (int e = threadIdx.x; e < lk * lk; e += 64) {
    st[e] = tau[e];
    sT[e] = T[e];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < lk; i += 64) {
    double2 tt = make_double2(0.0, 0.0), ttr = tt, ttt = tt;
    for (int c = 0; c < lk; ++c) {
      tt = cadd(tt, cmul(sT[i + c * lk], st[c + i * lk]));
      ttr = cadd(ttr, cmul(st[i + c * lk], sT[c + i * lk]));
    }
    for (int c2 = 0; c2 < lk; ++c2) {
      double2 row = make_double2(0.0, 0.0);  // (T tau)_{i,c2}
      for (int c = 0; c < lk; ++c) row = cadd(row, cmul(sT[i + c * lk], st[c + c2 * lk]));
      ttt = cadd(ttt, cmul(row, sT[c2 + i * lk]));
    }
    const double2 ai = a[(long long)site * lk + i], a2 = cmul(ai, ai);
    double2 d = cadd(st[i + i * lk], cmul(ai, cadd(tt, ttr)));
    d = cadd(d, cmul(a2, ttt));
    d = csub(d, cadd(ai, cmul(a2, sT[i + i * lk])));
    sd[i] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double2* __restrict__ o = out + ((long long)b * sites + site) * nl;
    for (int l = 0; l < nl; ++l) o[l] = make_double2(0.0, 0.0);
    for (int i = 0; i < lk; ++i) o[lidx[i]] = cadd(o[lidx[i]], sd[i]);
  }
}

// Full phys of one site block of the BZ sum (one work-group per site):
//   X = tau + A (T tau),  phys = X + (X T) A - wsum (A + A T A),
// wsum = sum of the BZ weights of the call (the constant term is affine in
// the weights, so the result adds up over calls like the BZ sum).  X goes
// through the scratch block x (site-major like tau).
template <int MAXLK>
__global__ __launch_bounds__(256) void phys_full_kernel(
    int lk, const double2* __restrict__ taubz, const double2* __restrict__ tint,
    const double2* __restrict__ a, double wsum, double2* __restrict__ x,
    double2* __restrict__ out) {
  __shared__ double2 sT[MAXLK * MAXLK];
  const int site = blockIdx.x;
  const long long nn = (long long)lk * lk;
  const double2* __restrict__ tau = taubz + site * nn;
  const double2* __restrict__ av = a + (long long)site * lk;
  double2* __restrict__ xs = x + site * nn;
  double2* __restrict__ os = out + site * nn;
  for (int e = threadIdx.x; e < lk * lk; e += 256) sT[e] = tint[site * nn + e];
  __syncthreads();
  for (int e = threadIdx.x; e < lk * lk; e += 256) {
    const int i = e % lk, j = e / lk;
    double2 y = make_double2(0.0, 0.0);
    for (int c = 0; c < lk; ++c) y = cadd(y, cmul(sT[i + c * lk], tau[c + j * lk]));
    xs[e] = cadd(tau[e], cmul(av[i], y));
  }
  __syncthreads();
  for (int e = threadIdx.x; e < lk * lk; e += 256) {
    const int i = e % lk, j = e / lk;
    double2 y = make_double2(0.0, 0.0);
    for (int c = 0; c < lk; ++c) y = cadd(y, cmul(xs[i + c * lk], sT[c + j * lk]));
    double2 cm = cmul(cmul(av[i], sT[i + j * lk]), av[j]);
    if (i == j) cm = cadd(cm, av[i]);
    const double2 v = csub(cadd(xs[e], cmul(y, av[j])), make_double2(wsum * cm.x, wsum * cm.y));
    os[e] = v;
  }
}
