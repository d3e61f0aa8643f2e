// Blocked Gauss-Jordan inversion on FP64 MFMA, one work-group per matrix.
//
// The matrix (padded to N = 16*T with an identity block) lives in registers
// as 16x16 tiles in the v_mfma_f64_16x16x4 result layout: lane l, register q
// of a tile hold row (l/16 + 4q), column (l%16).  Wave w owns the TI x TJ
// tiles (TI*wr+i, TJ*wc+j).
//
// Block Gauss-Jordan over the 16-wide block columns K = 0..T-1, with partial
// pivoting restricted to the rows of the diagonal block (rows are never
// exchanged between blocks).  Per block column K with diagonal block D:
//   1. block column K -> LDS;
//   2. the wave holding D inverts it in registers (Gauss-Jordan with partial
//      pivoting within the 16 rows, pivot search and broadcasts by DPP and
//      permlane swaps, then unscramble) -> Dinv; from the second block
//      column on this overlaps the other waves' rank update of step K-1;
//   3. W = -M[O,K] Dinv for the rows O outside block K, W[K] = Dinv, and
//      X = Dinv R for R the rows of block K in the other block columns, as
//      16x16 MFMA tile products, one tile per wave;
//   5. rank-16 MFMA update of every other tile: M[O, J] += W[O, :] R;
//   6. rows of block K <- X; block column K <- W.
// The matrices inverted here (t^-1 - G(k), the Schur complements and the
// decimation blocks) have a positive-definite anti-Hermitian part >= Im E,
// so every diagonal block stays invertible without row exchanges between
// blocks; on the configurations this matches LAPACK's pivoted inverse to
// ~1e-13 relative.  For any other matrix the kernel checks its multipliers:
// an element of W = -M[O,K] D^-1 above GJ_GROWTH_LIMIT (a diagonal block that
// is nearly singular or small against the rest of its column) or a zero pivot
// of a diagonal block sets GJ_STATUS_REDO in the matrix's status, and the host
// inverts that matrix again with the row-wise kernel, which pivots over all
// rows as gjinv.f does (and alone decides that a matrix is singular).
#pragma once
#include "kernels_f64.h"

// status bit of a matrix whose block elimination grew too large (see above)
constexpr int GJ_STATUS_REDO = 16;
constexpr double GJ_GROWTH_LIMIT = 1048576.0;  // 2^20

// Shuffle of a complex value: lane `src` of the wave (ds_bpermute)
__device__ __forceinline__ double2 shfl2(double2 v, int src) {
  return make_double2(__shfl(v.x, src), __shfl(v.y, src));
}

// Lane IT of each row of 16 lanes, in every lane of that row (DPP
// row_newbcast: a VALU move, no LDS round trip)
template <int IT>
__device__ __forceinline__ double2 row_bcast(double2 v) {
  return make_double2(__builtin_amdgcn_update_dpp(0.0, v.x, 0x150 + IT, 0xF, 0xF, true),
                      __builtin_amdgcn_update_dpp(0.0, v.y, 0x150 + IT, 0xF, 0xF, true));
}

// 1/z for a pivot z: conj(z) / |z|^2 with the reciprocal of |z|^2 from
// v_rcp_f64 refined by two Newton-Raphson steps r <- r (2 - s r).  A zero or
// NaN pivot gives non-finite values and is flagged by the caller.
__device__ __forceinline__ double2 pivot_recip(double2 z) {
  const double s2 = z.x * z.x + z.y * z.y;
  double r = __builtin_amdgcn_rcp(s2);
  r = fma(r, fma(-s2, r, 1.0), r);
  r = fma(r, fma(-s2, r, 1.0), r);
  return make_double2(z.x * r, -z.y * r);
}

// Pivot key of a candidate z in row `row`: the high word of |z|^2 (sign,
// exponent, 20 mantissa bits) with its four lowest bits replaced by
// 15 - row, compared as a signed integer -- the larger magnitude wins, the
// smaller row on a tie to 2^-16 relative; -1 for a row already pivoted.
// Zero magnitudes give keys 0..15, infinities and NaN keys >= 0x7ff00000.
__device__ __forceinline__ int pivot_key(double2 z, int row, bool used) {
  const double m = z.x * z.x + z.y * z.y;
  const int k = (__double2hiint(m) & ~15) | (15 - row);
  return used ? -1 : k;
}

// The larger of this lane's key and that of lane l ^ 32 (l ^ 16): the swap
// returns the two halves exchanged between two copies of the key, one of
// which is this lane's own value and the other its partner's.
__device__ __forceinline__ int max_xor32(int k) {
  const auto r = __builtin_amdgcn_permlane32_swap(k, k, false, false);
  return max((int)r[0], (int)r[1]);
}
__device__ __forceinline__ int max_xor16(int k) {
  const auto r = __builtin_amdgcn_permlane16_swap(k, k, false, false);
  return max((int)r[0], (int)r[1]);
}

// One elimination step on element q of a lane (row r of column l16), as
// selects rather than branches: the pivot row p takes rn = prow / piv, the
// pivot column starts from zero, every other element subtracts f_r rn.
__device__ __forceinline__ void gj_update(double2& dq, int r, int p, bool pivcol, double2 fq,
                                          double2 rn) {
  double2 o = pivcol ? make_double2(0.0, 0.0) : dq;
  o.x = fma(-fq.x, rn.x, fma(fq.y, rn.y, o.x));
  o.y = fma(-fq.x, rn.y, fma(-fq.y, rn.x, o.y));
  dq = (r == p) ? rn : o;
}

// Step IT of the 16x16 inversion below.  The column IT of the lane's four
// rows comes from lane IT of its row of lanes (DPP); the pivot is the
// largest key over the four rows of the lane and the four rows of lanes
// (permlane swaps), read into a scalar register; the pivot row p moves to
// every row of lanes by one shuffle and its column-IT element, the pivot,
// by DPP again.  sperm[IT] = p and spinv[p] = IT go to LDS (every lane
// stores the same value).
template <int IT, bool PIV>
__device__ __forceinline__ void gj16_step(double2& d0, double2& d1, double2& d2, double2& d3, int l16,
                                          int l4, unsigned& used, bool& bad, int* sperm, int* spinv) {
  const double2 f0 = row_bcast<IT>(d0), f1 = row_bcast<IT>(d1);
  const double2 f2 = row_bcast<IT>(d2), f3 = row_bcast<IT>(d3);
  int p = IT;
  if constexpr (PIV) {
    const int k0 = pivot_key(f0, l4, (used >> l4) & 1u);
    const int k1 = pivot_key(f1, l4 + 4, (used >> (l4 + 4)) & 1u);
    const int k2 = pivot_key(f2, l4 + 8, (used >> (l4 + 8)) & 1u);
    const int k3 = pivot_key(f3, l4 + 12, (used >> (l4 + 12)) & 1u);
    const int kb =
        __builtin_amdgcn_readfirstlane(max_xor16(max_xor32(max(max(k0, k1), max(k2, k3)))));
    if (kb < 0) {  // no unused row (not reachable before step 16)
      p = __builtin_ctz(~used & 0xFFFFu);
      bad = true;
    } else {
      p = 15 - (kb & 15);
      if (kb < 16 || kb >= 0x7ff00000) bad = true;  // zero, infinite or NaN pivot
    }
    used |= 1u << p;
    sperm[IT] = p;
    spinv[p] = IT;
  }
  // pivot row p of this lane's column: lane (p&3)*16 + l16, element p>>2
  const int qp = p >> 2;
  const double2 mine = qp == 0 ? d0 : qp == 1 ? d1 : qp == 2 ? d2 : d3;
  const double2 prow = shfl2(mine, (p & 3) * 16 + l16);
  const double2 piv = row_bcast<IT>(prow);
  if constexpr (!PIV) {
    if (!(piv.x * piv.x + piv.y * piv.y > 0.0)) bad = true;  // zero or NaN pivot
  }
  const double2 qv = pivot_recip(piv);
  const bool pivcol = (l16 == IT);
  const double2 rn = pivcol ? qv : cmul(prow, qv);
  gj_update(d0, l4, p, pivcol, f0, rn);
  gj_update(d1, l4 + 4, p, pivcol, f1, rn);
  gj_update(d2, l4 + 8, p, pivcol, f2, rn);
  gj_update(d3, l4 + 12, p, pivcol, f3, rn);
}

template <bool PIV, int IT = 0>
__device__ __forceinline__ void gj16_steps(double2& d0, double2& d1, double2& d2, double2& d3, int l16,
                                           int l4, unsigned& used, bool& bad, int* sperm, int* spinv) {
  if constexpr (IT < 16) {
    gj16_step<IT, PIV>(d0, d1, d2, d3, l16, l4, used, bad, sperm, spinv);
    gj16_steps<PIV, IT + 1>(d0, d1, d2, d3, l16, l4, used, bad, sperm, spinv);
  }
}

// 16x16 Gauss-Jordan inversion of a diagonal block held by one wave in the
// v_mfma_f64_16x16x4 result layout (lane l, element q: row l/16 + 4q, column
// l%16), with partial pivoting within the 16 rows (pivot rows recorded, no
// exchanges).  The elements are four separate variables (an index into an
// array by the wave-uniform pivot row would be lowered to scratch memory);
// the sixteen steps are unrolled so that the DPP lane selects are
// immediates.  Writes the unscrambled inverse to dinv[col * DP + row]; sets
// *sing on a zero, infinite or NaN pivot.
template <int DP, bool PIV>
__device__ __forceinline__ void diag_inverse16(double2& d0, double2& d1, double2& d2, double2& d3,
                                               int lane, double2* __restrict__ dinv, int* sperm,
                                               int* spinv, int* sing) {
  const int l16 = lane & 15, l4 = lane >> 4;
  unsigned used = 0;  // rows already pivoted (identical in every lane)
  bool bad = false;
  gj16_steps<PIV>(d0, d1, d2, d3, l16, l4, used, bad, sperm, spinv);
  if (bad && lane == 0) *sing = 1;
  if constexpr (PIV) {
    // unscramble: Dinv[pinv(rho)][perm(gam)] = S[rho][gam]
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int cc = sperm[l16];
    dinv[cc * DP + spinv[l4]] = d0;
    dinv[cc * DP + spinv[l4 + 4]] = d1;
    dinv[cc * DP + spinv[l4 + 8]] = d2;
    dinv[cc * DP + spinv[l4 + 12]] = d3;
  } else {
    dinv[l16 * DP + l4] = d0;
    dinv[l16 * DP + l4 + 4] = d1;
    dinv[l16 * DP + l4 + 8] = d2;
    dinv[l16 * DP + l4 + 12] = d3;
  }
}

template <int T, int TI, int TJ, bool PIV = true>
__global__ __launch_bounds__((T / TI) * (T / TJ) * 64, 2) void gjinv_blk_f64_kernel(InvArgs args) {
  constexpr int N = 16 * T;
  constexpr int WC = T / TJ;
  constexpr int NW = (T / TI) * WC;
  constexpr int NT = NW * 64;
  // [t][r] arrays padded to a row stride of N+1 (and 16x16 blocks to 17) so
  // that the 16 lanes of a column-strided access hit different LDS banks
  constexpr int NP = N + 1, DP = 17;
  __shared__ double2 pan[16 * NP];  // block column K, [t][r]
  __shared__ double2 Wb[16 * NP];   // W, [t][r]
  __shared__ double2 Rb[16 * N];    // R, [t][c]
  __shared__ double2 Xb[16 * N];    // X, [t][c]
  __shared__ double2 Dinv[16 * DP]; // [col][row]
  __shared__ int sperm[16], spinv[16];  // pivot rows of the diagonal block
  __shared__ int s_sing;
  __shared__ int s_redo;
  __shared__ double2 diag[N];

  const int b = blockIdx.x;
  if (args.active && !args.active[b]) return;
  const int n = args.n;
  const int lk = args.lk;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w / WC, wc = w % WC;
  const int l16 = lane & 15, l4 = lane >> 4;

  dx4 mre[TI][TJ], mim[TI][TJ];
  {
    const double2* __restrict__ base = args.base[b];
    const double2* __restrict__ s1 = args.sub1 ? args.sub1[b] : nullptr;
    const double2* __restrict__ s2 = args.sub2 ? args.sub2[b] : nullptr;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = (TI * wr + i) * 16 + l4 + 4 * q;
          const int c = (TJ * wc + j) * 16 + l16;
          double2 v;
          if (r < n && c < n) {
            const long long o = r + (long long)c * (args.ld ? args.ld : n);
            v = base[o];
            if (args.tinv && (r / lk) == (c / lk))
              v = cadd(args.tinv[(long long)(r / lk) * lk * lk + (r % lk) + (c % lk) * lk], v);
            if (s1) v = csub(v, s1[o]);
            if (s2) v = csub(v, s2[o]);
          } else {
            v = make_double2(r == c ? 1.0 : 0.0, 0.0);
          }
          mre[i][j][q] = v.x;
          mim[i][j][q] = v.y;
        }
  }
  if (tid == 0) {
    s_sing = 0;
    s_redo = 0;
  }
  bool grown = false;  // a multiplier this lane formed exceeded the limit (or was NaN)

  // prologue: block column 0 -> pan; the owner of tile (0,0) inverts D_0
  {
    double2 d[4];
    bool own = false;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        if (TJ * wc + j == 0) {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            pan[l16 * NP + (TI * wr + i) * 16 + l4 + 4 * q] = make_double2(mre[i][j][q], mim[i][j][q]);
          if (TI * wr + i == 0) {
            own = true;
#pragma unroll
            for (int q = 0; q < 4; ++q) d[q] = make_double2(mre[i][j][q], mim[i][j][q]);
          }
        }
    if (own) diag_inverse16<DP, PIV>(d[0], d[1], d[2], d[3], lane, Dinv, sperm, spinv, &s_sing);
  }

#pragma unroll 1
  for (int k = 0; k < T; ++k) {
    // (1) R = rows of block K in the other block columns -> Rb[t][c]
#pragma unroll
    for (int i = 0; i < TI; ++i)
      if (TI * wr + i == k)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          if (TJ * wc + j != k)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              Rb[(l4 + 4 * q) * N + (TJ * wc + j) * 16 + l16] = make_double2(mre[i][j][q], mim[i][j][q]);
    __syncthreads();  // pan (column K), Dinv and Rb are complete
    // (2) on MFMA, one 16x16 tile per wave and step:
    //     W[rho] = -pan[rho] Dinv for row tiles rho != K, W[K] = Dinv;
    //     X[gam] = Dinv R[gam] for column tiles gam != K.
#pragma unroll 1
    for (int rho = w; rho < T; rho += NW) {
      if (rho == k) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int u = l4 + 4 * q;
          Wb[l16 * NP + 16 * k + u] = Dinv[l16 * DP + u];
        }
        continue;
      }
      dx4 are = {0.0, 0.0, 0.0, 0.0}, aim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const double2 a = pan[(4 * ks + l4) * NP + 16 * rho + l16];
        const double2 d = Dinv[l16 * DP + 4 * ks + l4];
        are = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, d.x, are, 0, 0, 0);
        are = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, d.y, are, 0, 0, 0);
        aim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, d.y, aim, 0, 0, 0);
        aim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, d.x, aim, 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Wb[l16 * NP + 16 * rho + l4 + 4 * q] = make_double2(-are[q], -aim[q]);
        grown = grown || !(fabs(are[q]) <= GJ_GROWTH_LIMIT && fabs(aim[q]) <= GJ_GROWTH_LIMIT);
      }
    }
#pragma unroll 1
    for (int gam = w; gam < T; gam += NW) {
      if (gam == k) continue;
      dx4 are = {0.0, 0.0, 0.0, 0.0}, aim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const double2 d = Dinv[(4 * ks + l4) * DP + l16];
        const double2 r = Rb[(4 * ks + l4) * N + 16 * gam + l16];
        are = __builtin_amdgcn_mfma_f64_16x16x4f64(d.x, r.x, are, 0, 0, 0);
        are = __builtin_amdgcn_mfma_f64_16x16x4f64(-d.y, r.y, are, 0, 0, 0);
        aim = __builtin_amdgcn_mfma_f64_16x16x4f64(d.x, r.y, aim, 0, 0, 0);
        aim = __builtin_amdgcn_mfma_f64_16x16x4f64(d.y, r.x, aim, 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        Xb[(l4 + 4 * q) * N + 16 * gam + l16] = make_double2(are[q], aim[q]);
    }
    __syncthreads();  // W and X are complete; pan (column K) and Dinv are free
    // (3) rank-16 update M[O,J] += W[O] R[J] of the tiles outside block row K
    // and block column K, in two passes: first block column K+1 (look ahead:
    // it goes to pan, and the owner of tile (K+1, K+1) inverts D_{K+1} while
    // the other waves run the second pass), then the rest
#pragma unroll 1
    for (int ks = 0; ks < (k + 1 < T ? 4 : 0); ++ks) {
      const int t = 4 * ks + l4;
      const double2 bk = Rb[t * N + (k + 1) * 16 + l16];
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        const double2 af = Wb[t * NP + (TI * wr + i) * 16 + l16];
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          const int bi = TI * wr + i, bj = TJ * wc + j;
          if (bj == k + 1 && bi != k) {
            mre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af.x, bk.x, mre[i][j], 0, 0, 0);
            mre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-af.y, bk.y, mre[i][j], 0, 0, 0);
            mim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af.x, bk.y, mim[i][j], 0, 0, 0);
            mim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af.y, bk.x, mim[i][j], 0, 0, 0);
          }
        }
      }
    }
    {
      // block column K+1 (row block K from X) -> pan; D_{K+1} -> Dinv
      double2 d[4];
      bool own = false;
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          const int bi = TI * wr + i, bj = TJ * wc + j;
          if (bj == k + 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if (bi == k) {
                const double2 v = Xb[(l4 + 4 * q) * N + bj * 16 + l16];
                mre[i][j][q] = v.x;
                mim[i][j][q] = v.y;
              }
              pan[l16 * NP + bi * 16 + l4 + 4 * q] = make_double2(mre[i][j][q], mim[i][j][q]);
            }
            if (bi == k + 1) {
              own = true;
#pragma unroll
              for (int q = 0; q < 4; ++q) d[q] = make_double2(mre[i][j][q], mim[i][j][q]);
            }
          }
        }
      if (own) diag_inverse16<DP, PIV>(d[0], d[1], d[2], d[3], lane, Dinv, sperm, spinv, &s_sing);
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int t = 4 * ks + l4;
      double2 af[TI], bfr[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) af[i] = Wb[t * NP + (TI * wr + i) * 16 + l16];
#pragma unroll
      for (int j = 0; j < TJ; ++j) bfr[j] = Rb[t * N + (TJ * wc + j) * 16 + l16];
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          const int bi = TI * wr + i, bj = TJ * wc + j;
          if (bj != k && bi != k && bj != k + 1) {
            mre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].x, bfr[j].x, mre[i][j], 0, 0, 0);
            mre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-af[i].y, bfr[j].y, mre[i][j], 0, 0, 0);
            mim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].x, bfr[j].y, mim[i][j], 0, 0, 0);
            mim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].y, bfr[j].x, mim[i][j], 0, 0, 0);
          }
        }
    }
    // (4) rows of block K <- X (block column K+1 done above), block column K <- W
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        const int bi = TI * wr + i, bj = TJ * wc + j;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (bj == k) {
            const double2 v = Wb[l16 * NP + bi * 16 + l4 + 4 * q];
            mre[i][j][q] = v.x;
            mim[i][j][q] = v.y;
          } else if (bi == k && bj != k + 1) {
            const double2 v = Xb[(l4 + 4 * q) * N + bj * 16 + l16];
            mre[i][j][q] = v.x;
            mim[i][j][q] = v.y;
          }
        }
      }
    __syncthreads();  // the next pan, Dinv and this step's W, X, R reads are done
  }

  if (grown) s_redo = 1;
  __syncthreads();
  // a zero pivot of a diagonal block says nothing about the matrix (its rows
  // never moved between blocks): flagged for the pivoting second pass too
  if (s_sing && args.status && tid == 0) atomicOr(&args.status[b], 1 | GJ_STATUS_REDO);
  if (s_redo && args.status && tid == 0) atomicOr(&args.status[b], GJ_STATUS_REDO);
  if (args.mode == 0) {
    double2* __restrict__ out = args.out[b];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = (TI * wr + i) * 16 + l4 + 4 * q;
          const int c = (TJ * wc + j) * 16 + l16;
          if (r < n && c < n)
            out[r + (long long)c * (args.ldo ? args.ldo : n)] = make_double2(mre[i][j][q], mim[i][j][q]);
        }
  } else {
    double2* __restrict__ tau = args.tau[b];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = (TI * wr + i) * 16 + l4 + 4 * q;
          const int c = (TJ * wc + j) * 16 + l16;
          if (r < n && c < n && r / lk == c / lk) {
            const double2 v = make_double2(mre[i][j][q], mim[i][j][q]);
            tau[(long long)(r / lk) * lk * lk + (r % lk) + (c % lk) * lk] = v;
            if (r == c) diag[r] = v;
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
