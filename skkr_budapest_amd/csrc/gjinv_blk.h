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
//   2. wave 0 inverts D in LDS (in-place Gauss-Jordan, partial pivoting within
//      the 16 rows, then unscramble) -> Dinv;
//   3. W = -M[O,K] Dinv for the rows O outside block K, W[K] = Dinv, and
//      X = Dinv R for R the rows of block K in the other block columns, as
//      16x16 MFMA tile products, one tile per wave;
//   5. rank-16 MFMA update of every other tile: M[O, J] += W[O, :] R;
//   6. rows of block K <- X; block column K <- W.
// The matrices inverted here (t^-1 - G(k), the Schur complements and the
// decimation blocks) have a positive-definite anti-Hermitian part >= Im E,
// so every diagonal block stays invertible without row exchanges between
// blocks; on the configurations this matches LAPACK's pivoted inverse to
// ~1e-13 relative.
#pragma once
#include "kernels_f64.h"

template <int T, int TI, int TJ>
__global__ __launch_bounds__((T / TI) * (T / TJ) * 64) void gjinv_blk_f64_kernel(InvArgs args) {
  constexpr int N = 16 * T;
  constexpr int WC = T / TJ;
  constexpr int NW = (T / TI) * WC;
  constexpr int NT = NW * 64;
  __shared__ double2 pan[16 * N];   // block column K, [t][r]
  __shared__ double2 Wb[16 * N];    // W, [t][r]
  __shared__ double2 Rb[16 * N];    // R, [t][c]
  __shared__ double2 Xb[16 * N];    // X, [t][c]
  __shared__ double2 Db[16 * 16];   // D work area, [col][row]
  __shared__ double2 Dinv[16 * 16]; // [col][row]
  __shared__ int s_sing;
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
  if (tid == 0) s_sing = 0;

#pragma unroll 1
  for (int k = 0; k < T; ++k) {
    // (1) block column K -> pan[t][r]
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        if (TJ * wc + j == k) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = (TI * wr + i) * 16 + l4 + 4 * q;
            pan[l16 * N + r] = make_double2(mre[i][j][q], mim[i][j][q]);
          }
        }
    // R = rows of block K in the other block columns -> Rb[t][c]
#pragma unroll
    for (int i = 0; i < TI; ++i)
      if (TI * wr + i == k)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          if (TJ * wc + j != k)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              Rb[(l4 + 4 * q) * N + (TJ * wc + j) * 16 + l16] = make_double2(mre[i][j][q], mim[i][j][q]);
    __syncthreads();
    // (2) wave 0: Dinv = D^-1, in-place Gauss-Jordan with partial pivoting
    // inside the block.  Lane l holds column l%16, rows l/16 + 4q.
    if (w == 0) {
      double2 d[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) d[q] = pan[l16 * N + 16 * k + l4 + 4 * q];
#pragma unroll
      for (int q = 0; q < 4; ++q) Db[l16 * 16 + l4 + 4 * q] = d[q];
      __builtin_amdgcn_wave_barrier();
      unsigned used = 0;  // rows already pivoted (identical in every lane)
      int perm[16];
#pragma unroll
      for (int it = 0; it < 16; ++it) {
        int p = 0;
        double best = -1.0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const double2 v = Db[it * 16 + r];
          const double mg = v.x * v.x + v.y * v.y;
          if (!((used >> r) & 1u) && mg > best) { best = mg; p = r; }
        }
        if ((used >> p) & 1u) {  // only with NaN entries: take the first unused row
          p = __builtin_ctz(~used & 0xFFFFu);
          if (lane == 0) s_sing = 1;
        }
        used |= 1u << p;
        perm[it] = p;
        const double2 piv = Db[it * 16 + p];
        if (best <= 0.0 && lane == 0) s_sing = 1;
        const double2 qv = crecip(piv);
        const double2 prow = Db[l16 * 16 + p];
        const double2 rn = (l16 == it) ? qv : cmul(prow, qv);
        double2 f[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) f[q] = Db[it * 16 + l4 + 4 * q];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = l4 + 4 * q;
          if (r == p) {
            d[q] = rn;
          } else {
            double2 o = (l16 == it) ? make_double2(0.0, 0.0) : d[q];
            o.x -= f[q].x * rn.x - f[q].y * rn.y;
            o.y -= f[q].x * rn.y + f[q].y * rn.x;
            d[q] = o;
          }
        }
        // every lane has read this step's values before any lane writes
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < 4; ++q) Db[l16 * 16 + l4 + 4 * q] = d[q];
        __builtin_amdgcn_wave_barrier();
      }
      // unscramble: Dinv[pinv(rho)][perm(gam)] = S[rho][gam]
      int pinv[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) pinv[t] = 0;
#pragma unroll
      for (int t = 0; t < 16; ++t) pinv[perm[t] & 15] = t;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rho = l4 + 4 * q, gam = l16;
        int rr = 0, cc = 0;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          if (t == rho) rr = pinv[t];
          if (t == gam) cc = perm[t];
        }
        Dinv[cc * 16 + rr] = d[q];
      }
    }
    __syncthreads();
    // (3) on MFMA, one 16x16 tile per wave and step:
    //     W[rho] = -pan[rho] Dinv for row tiles rho != K, W[K] = Dinv;
    //     X[gam] = Dinv R[gam] for column tiles gam != K.
#pragma unroll 1
    for (int rho = w; rho < T; rho += NW) {
      if (rho == k) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int u = l4 + 4 * q;
          Wb[l16 * N + 16 * k + u] = Dinv[l16 * 16 + u];
        }
        continue;
      }
      dx4 are = {0.0, 0.0, 0.0, 0.0}, aim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const double2 a = pan[(4 * ks + l4) * N + 16 * rho + l16];
        const double2 d = Dinv[l16 * 16 + 4 * ks + l4];
        are = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, d.x, are, 0, 0, 0);
        are = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, d.y, are, 0, 0, 0);
        aim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, d.y, aim, 0, 0, 0);
        aim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, d.x, aim, 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        Wb[l16 * N + 16 * rho + l4 + 4 * q] = make_double2(-are[q], -aim[q]);
    }
#pragma unroll 1
    for (int gam = w; gam < T; gam += NW) {
      if (gam == k) continue;
      dx4 are = {0.0, 0.0, 0.0, 0.0}, aim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const double2 d = Dinv[(4 * ks + l4) * 16 + l16];
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
    __syncthreads();
    // (5) rank-16 update of the tiles outside block row K and block column K
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int t = 4 * ks + l4;
      double2 af[TI], bfr[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) af[i] = Wb[t * N + (TI * wr + i) * 16 + l16];
#pragma unroll
      for (int j = 0; j < TJ; ++j) bfr[j] = Rb[t * N + (TJ * wc + j) * 16 + l16];
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          if (TJ * wc + j != k && TI * wr + i != k) {
            mre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].x, bfr[j].x, mre[i][j], 0, 0, 0);
            mre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-af[i].y, bfr[j].y, mre[i][j], 0, 0, 0);
            mim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].x, bfr[j].y, mim[i][j], 0, 0, 0);
            mim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].y, bfr[j].x, mim[i][j], 0, 0, 0);
          }
    }
    // (6) rows of block K <- X, block column K <- W (LDS reads only: no barrier
    // needed after (5))
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = (TI * wr + i) * 16 + l4 + 4 * q;
          const int c = (TJ * wc + j) * 16 + l16;
          if (TJ * wc + j == k) {
            const double2 v = Wb[l16 * N + r];
            mre[i][j][q] = v.x;
            mim[i][j][q] = v.y;
          } else if (TI * wr + i == k) {
            const double2 v = Xb[(l4 + 4 * q) * N + c];
            mre[i][j][q] = v.x;
            mim[i][j][q] = v.y;
          }
        }
    __syncthreads();
  }

  if (s_sing && args.status && tid == 0) atomicOr(&args.status[b], 1);
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
