// Split-BF16 complex GEMM on the BF16 MFMA of gfx950 (Ozaki scheme).
//
//   C[b] = alpha * A[b](m x k) * B[b](k x n) + beta * C[b]   (or the decimation
//   epilogues of zgemm_f64_kernel), FP64 complex in and out.
//
// Every K-chunk of 32 is split on the fly.  Each row of the A chunk and each
// column of the B chunk gets a power-of-two scale 2^e above its largest |re|,
// |im|; the scaled value y in (-1, 1) is cut into S signed integer slices
//   y = s_1 2^-8 + s_2 2^-17 + ... + s_S 2^-(8+9(S-1)) + r,
//   s_i = rint of the remainder times 2^8 (first) or 2^9, |s_i| <= 256,
//   |r| <= 2^-9S,
// and each s_i is stored exactly as a BF16 value (integers up to 256 have at
// most 8 significant bits).  The slice products s_i t_j run on
// v_mfma_f32_16x16x32_bf16 and are summed per group g = i + j - 2 (slices with
// i + j > S + 1 are dropped, the triangular truncation of the Ozaki scheme):
// a 32-term sum of products of integers <= 256 is at most 2^21, a group holds
// at most S pairs of two MFMAs each, so for S <= 4 every FP32 accumulator
// stays an integer below 2^24 and the sum is exact.  (S = 5 allows 5 * 2^22 in
// the largest group: exact unless nearly every term has its maximal value.)
// After each chunk the groups are recombined in FP64 in a fixed order,
//   v = acc_0 + 2^-9 (acc_1 + 2^-9 (acc_2 + ...)),  C += v 2^(eA + eB - 16),
// and the chunks are added in K order, so the result is deterministic and a
// CPU emulation of the same steps reproduces it bit for bit.  The error per
// product is bounded by about 2^(-9S+1) K max|A_row| max|B_col| per chunk:
// 2^-35 for S = 4, 2^-44 for S = 5.
//
// Work-group tile 64 x 64 of C, eight waves of 32 x 16 (two MFMA blocks of
// 16 x 16); all 512 threads split the chunk into LDS (4 complex values of A and
// 4 of B each), then every wave reads the slices it needs.  The next chunk's
// values are loaded while the current chunk's products run, and for S <= 4
// the slice planes are double-buffered (one barrier per chunk; LDS
// 2 x S x 16 KB, S x 16 KB for S = 5).  Slice s of a value:
//   y = x 2^(8-e);  s_1 = rint(y);  y <- 512 (y - s_i);  s_{i+1} = rint(y)
// (y - rint(y) is exact by Sterbenz, so every remainder is exact).  As in
// zgemm_f64_kernel the MFMA computes the transposed block
// (B^T A^T), so lane l owns row l % 16 of a block and its stores along a
// column are contiguous.  v_mfma_f32_16x16x32_bf16: lane l supplies
// X[l%16][8(l/16) .. +7] and Y[8(l/16) .. +7][l%16] and gets
// D[4(l/16) + q][l%16] in register q.
#pragma once
#include "kernels_f64.h"

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float fx4;
typedef __attribute__((ext_vector_type(4))) unsigned int ux4;

// exponent e with max < 2^e (0 for max = 0); max is finite and >= 0
__device__ __forceinline__ int scale_exp(double mx) {
  if (!(mx > 0.0)) return 0;
  return ilogb(mx) + 1;
}

// LDS slice planes of a 64 x 32 operand chunk: [slice][re|im][row or col]
// [4 quarters of 8 K values]; quarter kq of row r sits at position
// kq ^ ((r >> 1) & 3) so that the 16 rows read by a lane group fall into
// different banks.
__device__ __forceinline__ int split_quarter(int r, int kq) { return kq ^ ((r >> 1) & 3); }

// The operand values a thread splits per chunk: 4 consecutive K values (eighth
// t & 7 of the chunk) of row / column t >> 3; the eight eighths of a row are
// neighbouring lanes.  ld_row / ld_k: element strides of a row and of K.
// FULL: the whole 64 x 32 chunk lies inside the matrix (no bounds tests).
typedef const double __attribute__((address_space(1))) * gdptr_t;
template <bool FULL>
__device__ __forceinline__ void split_fetch(const double2* __restrict__ M, long long ld_row, long long ld_k,
                                            int rmax, int k0, int kmax, int r0, int tid, double2* v) {
  const int rc = tid >> 3, k8 = tid & 7;
  const int gr = r0 + rc;
  const gdptr_t g = (gdptr_t)(const double*)(M + gr * ld_row + (long long)(k0 + 4 * k8) * ld_k);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const long long o = 2 * u * ld_k;
    if (FULL) {
      v[u] = make_double2(g[o], g[o + 1]);
    } else {
      const int gk = k0 + 4 * k8 + u;
      v[u] = (gr < rmax && gk < kmax) ? make_double2(g[o], g[o + 1]) : make_double2(0.0, 0.0);
    }
  }
}

// max of an int over the 8 lanes of an aligned group (DPP: quad swaps, then
// the half-row mirror, which pairs each quad with the other quad of its 8)
__device__ __forceinline__ int max8_dpp(int x) {
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false));  // row_half_mirror
  return x;
}

// two exact FP32 values -> their BF16 upper halves in one 32-bit word
__device__ __forceinline__ unsigned pack_bf16(float lo, float hi) {
  return __builtin_amdgcn_perm(__float_as_uint(hi), __float_as_uint(lo), 0x07060302u);
}

// Scale and slice the fetched values into the LDS planes [S][re|im][64][8
// halves of quarters] and the row exponent.
template <int S>
__device__ __forceinline__ void split_store(const double2* v, unsigned long long* __restrict__ plane,
                                            int* __restrict__ eplane, int tid) {
  const int rc = tid >> 3, k8 = tid & 7;
  double mx = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) mx = fmax(mx, fmax(fabs(v[u].x), fabs(v[u].y)));
  // exponent of the row's largest value: the lanes' own exponents (zero rows
  // count as -inf) reduced over the 8 lanes of the row
  constexpr int NONE = -(1 << 20);
  int e = max8_dpp(mx > 0.0 ? ilogb(mx) + 1 : NONE);
  if (e == NONE) e = 0;
  if (k8 == 0) eplane[rc] = e;
  double yr[4], yi[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    yr[u] = ldexp(v[u].x, 8 - e);
    yi[u] = ldexp(v[u].y, 8 - e);
  }
  // 8-byte half of quarter k8 / 2
  const int pos = split_quarter(rc, k8 >> 1) * 2 + (k8 & 1);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    float fr[4], fi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double sr = rint(yr[u]), si = rint(yi[u]);
      fr[u] = (float)sr;
      fi[u] = (float)si;
      if (s + 1 < S) {
        yr[u] = (yr[u] - sr) * 512.0;
        yi[u] = (yi[u] - si) * 512.0;
      }
    }
    const unsigned long long pr = (unsigned long long)pack_bf16(fr[0], fr[1]) |
                                  ((unsigned long long)pack_bf16(fr[2], fr[3]) << 32);
    const unsigned long long pi = (unsigned long long)pack_bf16(fi[0], fi[1]) |
                                  ((unsigned long long)pack_bf16(fi[2], fi[3]) << 32);
    plane[((s * 2 + 0) * 64 + rc) * 8 + pos] = pr;
    plane[((s * 2 + 1) * 64 + rc) * 8 + pos] = pi;
  }
}

template <int S>
__global__ __launch_bounds__(512) void zgemm_split_kernel(
    int m, int n, int k, const double2* const* __restrict__ Ap, int lda,
    const double2* const* __restrict__ Bp, int ldb, double2* const* __restrict__ Cp, int ldc,
    double2 alpha, double2 beta, const int* __restrict__ active, EpiArgs epi) {
  constexpr int TM = 64, TN = 64, KC = 32;
  // double-buffered planes where they fit (S <= 4: 2 x 2 x S x 8 KB)
  constexpr int NBUF = (S <= 4) ? 2 : 1;
  constexpr int PL = S * 2 * 64 * 4;  // ux4 (8 BF16 values) per operand plane set
  __shared__ ux4 sA[NBUF * PL];
  __shared__ ux4 sB[NBUF * PL];
  __shared__ int eA[NBUF * TM], eB[NBUF * TN];

  // XCD-aware numbering as in zgemm_f64_kernel: the tiles of one product on
  // one XCD
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
  double2* __restrict__ C = Cp ? Cp[b] : nullptr;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // wave w: rows 32 (w & 1) .. +31 (two 16-row blocks), columns 16 (w >> 1) .. +15
  const int wm = w & 1, wn = w >> 1;
  const int l16 = lane & 15, l4 = lane >> 4;
  const int c = wn * 16 + l16;
  const int qc = split_quarter(c, l4);
  int qa[2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) qa[bi] = split_quarter(wm * 32 + bi * 16 + l16, l4);

  double cre[2][4], cim[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) cre[i][q] = cim[i][q] = 0.0;

  // A chunk: element (r, kk) at r + kk lda; B chunk: (kk, col) at kk + col ldb
  const bool full = (row0 + TM <= m) && (col0 + TN <= n) && (k % KC == 0);
  auto fetch = [&](int k0, double2* va_, double2* vb_) {
    if (full) {
      split_fetch<true>(A, 1, lda, m, k0, k, row0, tid, va_);
      split_fetch<true>(B, ldb, 1, n, k0, k, col0, tid, vb_);
    } else {
      split_fetch<false>(A, 1, lda, m, k0, k, row0, tid, va_);
      split_fetch<false>(B, ldb, 1, n, k0, k, col0, tid, vb_);
    }
  };
  double2 va[4], vb[4];
  fetch(0, va, vb);
  split_store<S>(va, reinterpret_cast<unsigned long long*>(sA), eA, tid);
  split_store<S>(vb, reinterpret_cast<unsigned long long*>(sB), eB, tid);
  __syncthreads();
  const int nchunk = (k + KC - 1) / KC;
  for (int ch = 0; ch < nchunk; ++ch) {
    const int buf = (NBUF == 2) ? (ch & 1) : 0;
    const bool more = ch + 1 < nchunk;
    // next chunk's operands in flight during this chunk's products
    if (more) fetch((ch + 1) * KC, va, vb);
    const ux4* __restrict__ pa = sA + buf * PL;
    const ux4* __restrict__ pb = sB + buf * PL;
    // X operand: the B slices of this wave's column block (lane: column l16, quarter l4)
    bf16x8 xr[S], xi[S];
#pragma unroll
    for (int sj = 0; sj < S; ++sj) {
      xr[sj] = __builtin_bit_cast(bf16x8, pb[((sj * 2 + 0) * TN + c) * 4 + qc]);
      xi[sj] = __builtin_bit_cast(bf16x8, pb[((sj * 2 + 1) * TN + c) * 4 + qc]);
    }
    // slice products, group g = si + sj (0-based slices), si + sj <= S - 1
    fx4 are[2][S], aim[2][S];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int g = 0; g < S; ++g) {
        are[i][g] = (fx4){0.f, 0.f, 0.f, 0.f};
        aim[i][g] = (fx4){0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
    for (int si = 0; si < S; ++si) {
      // Y operand: A slice si of the two row blocks (lane: row l16, quarter l4),
      // the imaginary part also negated for the real part of the product
      bf16x8 yr[2], yi[2], yn[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) {
        const int r = wm * 32 + bi * 16 + l16;
        const ux4 ur = pa[((si * 2 + 0) * TM + r) * 4 + qa[bi]];
        const ux4 ui = pa[((si * 2 + 1) * TM + r) * 4 + qa[bi]];
        yr[bi] = __builtin_bit_cast(bf16x8, ur);
        yi[bi] = __builtin_bit_cast(bf16x8, ui);
        yn[bi] = __builtin_bit_cast(bf16x8, ui ^ (ux4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u});
      }
#pragma unroll
      for (int sj = 0; sj + si < S; ++sj) {
        const int g = si + sj;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          are[bi][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[sj], yr[bi], are[bi][g], 0, 0, 0);
          aim[bi][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[sj], yi[bi], aim[bi][g], 0, 0, 0);
          are[bi][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xi[sj], yn[bi], are[bi][g], 0, 0, 0);
          aim[bi][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xi[sj], yr[bi], aim[bi][g], 0, 0, 0);
        }
      }
    }
    // recombine the groups in FP64 and add this chunk
    const int* __restrict__ ea_ = eA + buf * TM;
    const int* __restrict__ eb_ = eB + buf * TN;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
      const int ea = ea_[wm * 32 + bi * 16 + l16];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int eb = eb_[wn * 16 + 4 * l4 + q];
        double vr = (double)are[bi][S - 1][q], vi = (double)aim[bi][S - 1][q];
#pragma unroll
        for (int g = S - 2; g >= 0; --g) {
          vr = vr * (1.0 / 512.0) + (double)are[bi][g][q];
          vi = vi * (1.0 / 512.0) + (double)aim[bi][g][q];
        }
        cre[bi][q] += ldexp(vr, ea + eb - 16);
        cim[bi][q] += ldexp(vi, ea + eb - 16);
      }
    }
    if (more) {
      const int nb = (NBUF == 2) ? (buf ^ 1) : 0;
      if (NBUF == 1) __syncthreads();  // this chunk's planes are consumed
      split_store<S>(va, reinterpret_cast<unsigned long long*>(sA + nb * PL), eA + nb * TM, tid);
      split_store<S>(vb, reinterpret_cast<unsigned long long*>(sB + nb * PL), eB + nb * TN, tid);
      __syncthreads();
    }
  }
  __syncthreads();  // planes free for the reduction below

  // D is the transposed block: C row (l % 16), C column 4 (l / 16) + q
  auto row_of = [&](int bi) { return row0 + wm * 32 + bi * 16 + l16; };
  auto col_of = [&](int q) { return col0 + wn * 16 + 4 * l4 + q; };
  if (epi.mode == EPI_PLAIN) {
    const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
    // with beta, the lane's loads of C before its first store (an element
    // outside m x n reads element (0, 0) and stores nothing)
    double2 c0[2][4];
    if (use_beta) {
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int gr = row_of(bi), gc = col_of(q);
          c0[bi][q] = gload2(C + ((gr < m && gc < n) ? gr + (long long)gc * ldc : 0ll));
        }
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gr = row_of(bi), gc = col_of(q);
        if (gr < m && gc < n) {
          double2 r = cmul(alpha, make_double2(cre[bi][q], cim[bi][q]));
          if (use_beta) r = cadd(r, cmul(beta, c0[bi][q]));
          gstore2(C + gr + (long long)gc * ldc, r);
        }
      }
    return;
  }
  const int hn = epi.n;
  double2* __restrict__ eps = epi.eps[b];
  double2* __restrict__ xx = epi.x[b];
  double2* __restrict__ abn = epi.abn[b];
  double part = 0.0;
  // Top rows (gr < hn): beta mode eps -= v and x -= v, alpha mode alpha' = v;
  // bottom rows: beta mode beta' = v, alpha mode eps -= v.  The four values of
  // a row issue all their read-modify-write loads before the first store (a
  // column beyond n reads column n - 1 and stores nothing).
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
    const int gr = row_of(bi);
    if (gr < m) {
      const bool top = gr < hn;
      const bool rmw_eps = top ? epi.mode == EPI_DEC_BETA : epi.mode == EPI_DEC_ALPHA;
      const bool rmw_x = top && epi.mode == EPI_DEC_BETA;
      const int er = top ? gr : gr - hn;  // row in eps and x
      if (rmw_eps) {
        double2 e[4], xv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long long o = er + (long long)min(col_of(q), n - 1) * hn;
          e[q] = gload2(eps + o);
          if (rmw_x) xv[q] = gload2(xx + o);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int gc = col_of(q);
          if (gc < n) {
            const double2 v = make_double2(cre[bi][q], cim[bi][q]);
            const long long o = er + (long long)gc * hn;
            gstore2(eps + o, csub(e[q], v));
            if (rmw_x) gstore2(xx + o, csub(xv[q], v));
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int gc = col_of(q);
          if (gc < n) {
            const double2 v = make_double2(cre[bi][q], cim[bi][q]);
            gstore2(abn + gr + (long long)gc * 2 * hn, v);
            if (top) part += v.x * v.x + v.y * v.y;  // the top rows of alpha'
          }
        }
      }
    }
  }
  if (epi.mode == EPI_DEC_ALPHA) {
    // deterministic work-group sum of |alpha'|^2 (fixed-order tree in LDS,
    // in the slice planes, which are free now)
    double* red = reinterpret_cast<double*>(sA);
    red[tid] = part;
    __syncthreads();
    for (int st = 256; st > 0; st >>= 1) {
      if (tid < st) red[tid] += red[tid + st];
      __syncthreads();
    }
    if (tid == 0) epi.norm[(long long)b * epi.tiles + tile] = red[0];
  }
}
