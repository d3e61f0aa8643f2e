// Complex GEMM from exact INT8 digit products on the INT8 MFMA of gfx950
// (Ozaki scheme with integer accumulation).
//
//   C[b] = alpha * A[b](m x k) * B[b](k x n) + beta * C[b]   (or the decimation
//   epilogues of zgemm_f64_kernel), FP64 complex in and out.
//
// Digits.  Every K-chunk of 32 gives each row of A and each column of B an
// exponent e with max(|re|, |im|) < 2^e.  A value x of that line becomes the
// integer V = rint(x 2^(8S-2-e)), |V| <= 2^(8S-2), written in S balanced
// base-256 digits,
//   V = sum_t d_t 256^t,  d_t in [-128, 127]   (t = 0 least significant),
// so x = V 2^(e-8S+2) to 2^-(8S-1) of the line maximum (S = 6: 2^-47).
// One FMA yields the digits: D = fma(x, 2^(8S-2-e), 1.5 2^52 + c) with
// c = sum_t 128 256^t rounds x 2^(8S-2-e) to the integer V (the product is
// exact, the sum has unit spacing) and leaves c + V in the low 8S bits of D;
// byte t of c + V is d_t + 128, so d_t = byte t XOR 0x80.
//
// Products.  Digit products d_t d'_u are summed on v_mfma_i32_32x32x32_i8 per
// group g = (S-1-t) + (S-1-u), the digit pairs with g <= S-1 (the triangular
// truncation of the Ozaki scheme; g >= S is dropped).  Group sums are exact
// int32: |d d'| <= 2^14, a group holds at most S pairs of 32-term sums.  Two
// groups share one accumulator, P = 256 G_2h + G_2h+1 (< 2^30, exact), and
//   C += sum_h P_h 2^(eA + eB - 20 - 16h)        (h ascending, in FP64)
// after each chunk; the chunks are added in K order.  The real part takes
// the products of the real digits and of the imaginary digits in two
// accumulators whose int32 difference is exact.  The result is deterministic
// and a CPU emulation of the same steps reproduces it bit for bit.  The error
// per product is about 2^-(8S-3) K max|A_row| max|B_col| per chunk.
//
// zgemm_i8_kernel: work-group tile 128 x 64 of C, eight waves of 32 x 32 (one
// MFMA block each); one work-group per CU (LDS).  (zgemm_i8_64_kernel at the
// end: 64 x 64 tiles, two work-groups per CU, the same bits.)  All 512 threads
// cut the next chunk into LDS digit
// planes [digit][re|im][row or column][32 bytes of K] between the current
// chunk's digit pairs; the planes are double-buffered (S = 6: 2 x 72 KB), one
// barrier per chunk.  Each wave reads the A and B digits of its 32 rows and
// columns one digit pair ahead of the products.  The MFMA computes the
// transposed block (B^T A^T): lane l holds row l % 32 of C and columns
// 8 (r / 4) + 4 (l / 32) + r % 4 in register r, so the stores along a column
// are contiguous.  Both operands give lane l the K values 16 (l / 32) .. +15
// of the chunk, so any fixed K order inside the instruction gives the same
// sum.
#pragma once
#include "kernels_f64.h"

typedef __attribute__((ext_vector_type(4))) int i8x4w;   // 16 int8 values
typedef __attribute__((ext_vector_type(16))) int i32x16;

// 1.5 2^52 + sum_{t < S} 128 256^t: adding it rounds to an integer and puts
// c + V in the low mantissa bits
template <int S>
__device__ __forceinline__ double i8_magic() {
  unsigned long long c = 0;
#pragma unroll
  for (int t = 0; t < S; ++t) c |= 0x80ull << (8 * t);
  return __longlong_as_double((long long)(0x4338000000000000ull | c));
}

// max of an int over the 8 lanes of an aligned group (DPP quad swaps and the
// half-row mirror)
__device__ __forceinline__ int i8_max8(int x) {
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false));  // row_half_mirror
  return x;
}

// 4 values x 8 bytes -> per byte t the dword [b_t(v0), b_t(v1), b_t(v2), b_t(v3)]
// (v_perm_b32: bytes of the second operand are 0-3, of the first 4-7)
template <int S>
__device__ __forceinline__ void i8_transpose(const unsigned* lo, const unsigned* hi, unsigned* d) {
  const unsigned t01 = __builtin_amdgcn_perm(lo[1], lo[0], 0x05010400u);
  const unsigned t23 = __builtin_amdgcn_perm(lo[3], lo[2], 0x05010400u);
  d[0] = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
  if (S > 1) d[1] = __builtin_amdgcn_perm(t23, t01, 0x07060302u);
  if (S > 2) {
    const unsigned u01 = __builtin_amdgcn_perm(lo[1], lo[0], 0x07030602u);
    const unsigned u23 = __builtin_amdgcn_perm(lo[3], lo[2], 0x07030602u);
    d[2] = __builtin_amdgcn_perm(u23, u01, 0x05040100u);
    if (S > 3) d[3] = __builtin_amdgcn_perm(u23, u01, 0x07060302u);
  }
  if (S > 4) {
    const unsigned h01 = __builtin_amdgcn_perm(hi[1], hi[0], 0x05010400u);
    const unsigned h23 = __builtin_amdgcn_perm(hi[3], hi[2], 0x05010400u);
    d[4] = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
    if (S > 5) d[5] = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
  }
}

// physical 16-byte half of logical half hf in plane row rc: the halves swap
// places by bit 2 xor bit 3 of the row, so that the 16-byte reads of 8 (or
// 16) consecutive rows by neighbouring lanes fall into distinct banks with 32
// (or 64) banks of 4 bytes
__device__ __forceinline__ int i8_half(int rc, int hf) { return hf ^ (((rc >> 2) ^ (rc >> 3)) & 1); }
// physical dword of K-quarter q (4 K values) in plane row rc
__device__ __forceinline__ int i8_dword(int rc, int q) { return (q & 3) | (i8_half(rc, q >> 2) << 2); }

// Cut 4 consecutive K values of line rc (K-quarter q = lane & 7; the 8
// quarters of a line are 8 neighbouring lanes) into the planes
// [S][re|im][rows][8 dwords] and store the line exponent.
template <int S>
__device__ __forceinline__ void i8_split_store(const double2* v, unsigned* __restrict__ plane, int rows,
                                               int* __restrict__ eplane, int rc, int q) {
  double mx = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) mx = fmax(mx, fmax(fabs(v[u].x), fabs(v[u].y)));
  constexpr int NONE = -(1 << 20);
  int e = i8_max8(mx > 0.0 ? ilogb(mx) + 1 : NONE);
  if (e == NONE) e = 0;      // all-zero line: every digit is 0
  e = max(e, -960);          // keeps the scale finite; lines below 2^-960 round to 0
  if (q == 0) eplane[rc] = e;
  const double f = ldexp(1.0, 8 * S - 2 - e);
  const double M = i8_magic<S>();
  unsigned lr[4], hr[4], li[4], hi[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const double dr = fma(v[u].x, f, M), di = fma(v[u].y, f, M);
    const unsigned long long br = (unsigned long long)__double_as_longlong(dr);
    const unsigned long long bi = (unsigned long long)__double_as_longlong(di);
    lr[u] = (unsigned)br;
    hr[u] = (unsigned)(br >> 32);
    li[u] = (unsigned)bi;
    hi[u] = (unsigned)(bi >> 32);
  }
  unsigned dr[S], di[S];
  i8_transpose<S>(lr, hr, dr);
  i8_transpose<S>(li, hi, di);
  const int pq = i8_dword(rc, q);
#pragma unroll
  for (int t = 0; t < S; ++t) {
    plane[((t * 2 + 0) * rows + rc) * 8 + pq] = dr[t] ^ 0x80808080u;
    plane[((t * 2 + 1) * rows + rc) * 8 + pq] = di[t] ^ 0x80808080u;
  }
}

// 4 consecutive K values (quarter q) of row / column rc of a chunk; ld_row /
// ld_k: element strides of a line and of K.  Outside the matrix (rc >= rmax
// or K >= kmax) the value is 0; the load then reads an element inside it, so
// the code has no branches.
template <bool FULL>
__device__ __forceinline__ void i8_fetch(const double2* __restrict__ M, long long ld_row, long long ld_k,
                                         int rmax, int k0, int kmax, int rc, int q, double2* v) {
  // global (not flat) loads: they wait on vmcnt only, not on the LDS counter
  typedef const double __attribute__((address_space(1)))* gptr;
  if (FULL) {
    const gptr g = (gptr)(const double*)(M + (long long)rc * ld_row + (long long)(k0 + 4 * q) * ld_k);
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = make_double2(g[2 * u * ld_k], g[2 * u * ld_k + 1]);
    return;
  }
  const bool rok = rc < rmax;
  const long long rbase = (long long)(rok ? rc : 0) * ld_row;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int gk = k0 + 4 * q + u;
    const bool ok = rok && gk < kmax;
    const gptr g = (gptr)(const double*)(M + rbase + (long long)(ok ? gk : 0) * ld_k);
    const double xr = g[0], xi = g[1];
    v[u] = make_double2(ok ? xr : 0.0, ok ? xi : 0.0);
  }
}

// The digit products of one K-chunk for a wave's 32 x 32 block of C and
// their fold into (cre, cim): planes pa (A, TM rows) and pb (B, TN columns),
// exponents ea_ / eb_, the lane's plane offsets oA / oB, its C row rA (in the
// tile), wave column wn and K-half hf.  after_h(h) runs at the end of group
// pair h, before its fold (the 128 x 64 kernel cuts the next chunk there).
template <int S, int TM, int TN, class F>
__device__ __forceinline__ void i8_chunk_products(const unsigned* __restrict__ pa, const unsigned* __restrict__ pb,
                                                  const int* __restrict__ ea_, const int* __restrict__ eb_,
                                                  int oA, int oB, int rA, int wn, int hf, double* cre,
                                                  double* cim, F&& after_h) {
  // the digit pairs of the chunk in order: group pair h, group g, A digit
  // S-1-i with B digit S-1-(g-i); the operands of pair p + 1 are read before
  // the products of pair p are issued
  constexpr int NP = S * (S + 1) / 2;
  // group of pair p: pairs g (g + 1) / 2 .. (g + 1) (g + 2) / 2 - 1 (no loop,
  // so that the pair loop unrolls completely)
  auto pair_g = [](int p) { return (p >= 1) + (p >= 3) + (p >= 6) + (p >= 10) + (p >= 15); };
  auto pair_i = [&](int p) { return p - pair_g(p) * (pair_g(p) + 1) / 2; };
  i8x4w op[2][4];
  auto load = [&](int p, i8x4w* o) {
    const int g = pair_g(p), i = pair_i(p);
    const int ta = S - 1 - i, tb = S - 1 - (g - i);
    o[0] = *reinterpret_cast<const i8x4w*>(pa + (ta * 2 + 0) * TM * 8 + oA);
    o[1] = *reinterpret_cast<const i8x4w*>(pa + (ta * 2 + 1) * TM * 8 + oA);
    o[2] = *reinterpret_cast<const i8x4w*>(pb + (tb * 2 + 0) * TN * 8 + oB);
    o[3] = *reinterpret_cast<const i8x4w*>(pb + (tb * 2 + 1) * TN * 8 + oB);
  };
  i32x16 ap, am, ai;
  load(0, op[0]);
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int g = pair_g(p), i = pair_i(p), h = g / 2;
    if (p + 1 < NP) load(p + 1, op[(p + 1) & 1]);
    const i8x4w* o = op[p & 1];  // A re, A im, B re, B im
    if (g == 2 * h + 1 && i == 0) {
      // group 2h is complete: P = 256 G_2h, and group 2h + 1 adds to it
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        ap[r] <<= 8;
        am[r] <<= 8;
        ai[r] <<= 8;
      }
    }
    if (g == 2 * h && i == 0) {
      const i32x16 z = {0};
      ap = __builtin_amdgcn_mfma_i32_32x32x32_i8(o[2], o[0], z, 0, 0, 0);
      am = __builtin_amdgcn_mfma_i32_32x32x32_i8(o[3], o[1], z, 0, 0, 0);
      ai = __builtin_amdgcn_mfma_i32_32x32x32_i8(o[2], o[1], z, 0, 0, 0);
    } else {
      ap = __builtin_amdgcn_mfma_i32_32x32x32_i8(o[2], o[0], ap, 0, 0, 0);
      am = __builtin_amdgcn_mfma_i32_32x32x32_i8(o[3], o[1], am, 0, 0, 0);
      ai = __builtin_amdgcn_mfma_i32_32x32x32_i8(o[2], o[1], ai, 0, 0, 0);
    }
    ai = __builtin_amdgcn_mfma_i32_32x32x32_i8(o[3], o[0], ai, 0, 0, 0);
    // two pairs' operands in flight at a time
    __builtin_amdgcn_sched_barrier(0);
    const bool last_of_h = (p + 1 == NP) || (pair_g(p + 1) / 2 != h);
    if (last_of_h) {
      if (g == 2 * h) {  // odd S: the last group pair has one group
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          ap[r] <<= 8;
          am[r] <<= 8;
          ai[r] <<= 8;
        }
      }
      after_h(h);
      // fold: C += P_h 2^(eA + eB - 20 - 16 h)
      const int ea = ea_[rA] - 20 - 16 * h;
#pragma unroll
      for (int a4 = 0; a4 < 4; ++a4) {
        const i8x4w eb = *reinterpret_cast<const i8x4w*>(eb_ + wn * 32 + 8 * a4 + 4 * hf);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 4 * a4 + q;
          cre[r] += ldexp((double)(ap[r] - am[r]), ea + eb[q]);
          cim[r] += ldexp((double)ai[r], ea + eb[q]);
        }
      }
      // the fold happens here (not sunk past the next pairs), so that only
      // one set of accumulators is live
#pragma unroll
      for (int r = 0; r < 16; ++r) asm volatile("" : "+v"(cre[r]), "+v"(cim[r]));
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Write the wave's 32 x 32 block (lane: C row row0 + rA; register r: column
// col0 + 32 wn + 8 (r / 4) + 4 hf + r % 4) with the plain or the decimation
// epilogue.  red: NT doubles of free LDS for the work-group's |alpha'|^2 sum.
template <int NT>
__device__ __forceinline__ void i8_epilogue(int m, int n, int row0, int col0, int rA, int wn, int hf,
                                            const double* cre, const double* cim, double2* __restrict__ C,
                                            int ldc, double2 alpha, double2 beta, const EpiArgs& epi, int b,
                                            int tile, double* red, int tid) {
  const int gr = row0 + rA;
  auto col_of = [&](int r) { return col0 + wn * 32 + 8 * (r >> 2) + 4 * hf + (r & 3); };
  if (epi.mode == EPI_PLAIN) {
    const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
    if (gr < m) {
      // with beta, all 16 loads of C before the first store (a column beyond
      // n reads column n - 1 and stores nothing)
      double2 c0[16];
      if (use_beta) {
#pragma unroll
        for (int r = 0; r < 16; ++r) c0[r] = gload2(C + gr + (long long)min(col_of(r), n - 1) * ldc);
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gc = col_of(r);
        if (gc < n) {
          double2 v = cmul(alpha, make_double2(cre[r], cim[r]));
          if (use_beta) v = cadd(v, cmul(beta, c0[r]));
          gstore2(C + gr + (long long)gc * ldc, v);
        }
      }
    }
    return;
  }
  const int hn = epi.n;
  double2* __restrict__ eps = epi.eps[b];
  double2* __restrict__ xx = epi.x[b];
  double2* __restrict__ abn = epi.abn[b];
  double part = 0.0;
  if (gr < m) {
    // Top rows (gr < hn): beta mode eps -= v and x -= v, alpha mode alpha' =
    // v; bottom rows: beta mode beta' = v, alpha mode eps -= v.  A lane's 16
    // values share its row, so each read-modify-write issues all its loads
    // before the first store (a column beyond n reads column n - 1 and stores
    // nothing): their latencies overlap instead of adding up.
    const bool top = gr < hn;
    const bool rmw_eps = top ? epi.mode == EPI_DEC_BETA : epi.mode == EPI_DEC_ALPHA;
    const bool rmw_x = top && epi.mode == EPI_DEC_BETA;
    const int er = top ? gr : gr - hn;  // row in eps and x
    if (rmw_eps) {
      double2 e[16], xv[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long long o = er + (long long)min(col_of(r), n - 1) * hn;
        e[r] = gload2(eps + o);
        if (rmw_x) xv[r] = gload2(xx + o);
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gc = col_of(r);
        if (gc < n) {
          const double2 v = make_double2(cre[r], cim[r]);
          const long long o = er + (long long)gc * hn;
          gstore2(eps + o, csub(e[r], v));
          if (rmw_x) gstore2(xx + o, csub(xv[r], v));
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gc = col_of(r);
        if (gc < n) {
          const double2 v = make_double2(cre[r], cim[r]);
          gstore2(abn + gr + (long long)gc * 2 * hn, v);
          if (top) part += v.x * v.x + v.y * v.y;  // the top rows of alpha'
        }
      }
    }
  }
  if (epi.mode == EPI_DEC_ALPHA) {
    // deterministic work-group sum of |alpha'|^2 (fixed-order tree in LDS)
    red[tid] = part;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
      if (tid < st) red[tid] += red[tid + st];
      __syncthreads();
    }
    if (tid == 0) epi.norm[(long long)b * epi.tiles + tile] = red[0];
  }
}

// XCD-aware numbering as in zgemm_f64_kernel: the tiles of one product on
// one XCD (work-groups go round-robin over the eight XCDs)
__device__ __forceinline__ void i8_tile_of(int& b, int& tile) {
  b = blockIdx.y;
  tile = blockIdx.x;
  const int tiles = gridDim.x;
  const int L = blockIdx.x + tiles * blockIdx.y;
  if (tiles > 1 && L < 8 * tiles * (int)(gridDim.y / 8)) {
    const int g = L / (8 * tiles), rem = L % (8 * tiles);
    tile = rem / 8;
    b = 8 * g + rem % 8;
  }
}

template <int S>
__global__ __launch_bounds__(512) void zgemm_i8_kernel(
    int m, int n, int k, const double2* const* __restrict__ Ap, int lda,
    const double2* const* __restrict__ Bp, int ldb, double2* const* __restrict__ Cp, int ldc,
    double2 alpha, double2 beta, const int* __restrict__ active, EpiArgs epi) {
  static_assert(S >= 2 && S <= 6, "the FMA digit cut holds c + V below 2^51: S <= 6 (and pair_g covers 21 pairs)");
  constexpr int TM = 128, TN = 64, KC = 32;
  constexpr int PA = S * 2 * TM * 8;  // dwords of one A plane set
  constexpr int PB = S * 2 * TN * 8;
  __shared__ __attribute__((aligned(16))) unsigned sA[2 * PA];
  __shared__ __attribute__((aligned(16))) unsigned sB[2 * PB];
  __shared__ __attribute__((aligned(16))) int eA[2 * TM];
  __shared__ __attribute__((aligned(16))) int eB[2 * TN];

  int b, tile;
  i8_tile_of(b, tile);
  if (active && !active[b]) return;
  const int tiles_m = (m + TM - 1) / TM;
  const int row0 = (tile % tiles_m) * TM;
  const int col0 = (tile / tiles_m) * TN;
  const double2* __restrict__ A = Ap[b];
  const double2* __restrict__ B = Bp[b];
  double2* __restrict__ C = Cp ? Cp[b] : nullptr;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 3, wn = w >> 2;  // wave: rows 32 wm .. +31, columns 32 wn .. +31
  const int l32 = lane & 31, hf = lane >> 5;
  // split items: A rows tid / 8 and tid / 8 + 64, B column tid / 8; quarter tid % 8
  const int irc = tid >> 3, iq = tid & 7;

  double cre[16], cim[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) cre[r] = cim[r] = 0.0;

  // A chunk: element (r, kk) at r + kk lda; B chunk: (kk, col) at kk + col ldb
  const bool full = (row0 + TM <= m) && (col0 + TN <= n) && (k % KC == 0);
  const double2* Ar = A + row0;
  const double2* Bc = B + (long long)col0 * ldb;
  double2 va0[4], va1[4], vb[4];
  auto fetch = [&](int k0) {
    if (full) {
      i8_fetch<true>(Ar, 1, lda, m - row0, k0, k, irc, iq, va0);
      i8_fetch<true>(Ar, 1, lda, m - row0, k0, k, irc + 64, iq, va1);
      i8_fetch<true>(Bc, ldb, 1, n - col0, k0, k, irc, iq, vb);
    } else {
      i8_fetch<false>(Ar, 1, lda, m - row0, k0, k, irc, iq, va0);
      i8_fetch<false>(Ar, 1, lda, m - row0, k0, k, irc + 64, iq, va1);
      i8_fetch<false>(Bc, ldb, 1, n - col0, k0, k, irc, iq, vb);
    }
  };
  // split item it (0, 1: A rows irc, irc + 64; 2: B column irc) into buffer buf
  auto store = [&](int it, int buf) {
    if (it == 0) i8_split_store<S>(va0, sA + buf * PA, TM, eA + buf * TM, irc, iq);
    if (it == 1) i8_split_store<S>(va1, sA + buf * PA, TM, eA + buf * TM, irc + 64, iq);
    if (it == 2) i8_split_store<S>(vb, sB + buf * PB, TN, eB + buf * TN, irc, iq);
  };
  fetch(0);
  store(0, 0);
  store(1, 0);
  store(2, 0);
  __syncthreads();

  // this lane's operand rows: B column cB and A row rA, half hf of the K chunk
  const int cB = wn * 32 + l32, rA = wm * 32 + l32;
  const int oB = cB * 8 + 4 * i8_half(cB, hf);
  const int oA = rA * 8 + 4 * i8_half(rA, hf);
  constexpr int H = (S + 1) / 2;  // group pairs
  const int nchunk = (k + KC - 1) / KC;
  for (int ch = 0; ch < nchunk; ++ch) {
    const int buf = ch & 1;
    const bool more = ch + 1 < nchunk;
    if (more) fetch((ch + 1) * KC);
    // the next chunk's split work between the group pairs (its loads have
    // had the first pair's products to arrive)
    auto after_h = [&](int h) {
      if (!more) return;
      if (H >= 3) {
        if (h == 1) store(0, buf ^ 1);
        if (h == 2) {
          store(1, buf ^ 1);
          store(2, buf ^ 1);
        }
      } else if (h == H - 1) {
        store(0, buf ^ 1);
        store(1, buf ^ 1);
        store(2, buf ^ 1);
      }
    };
    i8_chunk_products<S, TM, TN>(sA + buf * PA, sB + buf * PB, eA + buf * TM, eB + buf * TN, oA, oB, rA, wn,
                                 hf, cre, cim, after_h);
    if (more) __syncthreads();
  }
  __syncthreads();  // planes free for the reduction in the epilogue
  i8_epilogue<512>(m, n, row0, col0, rA, wn, hf, cre, cim, C, ldc, alpha, beta, epi, b, tile,
                   reinterpret_cast<double*>(sA), tid);
}

// The same product on 64 x 64 tiles: four waves of 32 x 32, one set of digit
// planes (S = 6: 48 KB), so that two work-groups share a CU and one's cut,
// loads and epilogue run beside the other's MFMAs.  Per chunk: cut the
// chunk's values into the planes, start the loads of the next chunk, barrier,
// the digit products and folds, barrier.  The digits, the pair order and the
// folds are those of zgemm_i8_kernel, so both give the same bits.
template <int S>
__global__ __launch_bounds__(256, 2) void zgemm_i8_64_kernel(
    int m, int n, int k, const double2* const* __restrict__ Ap, int lda,
    const double2* const* __restrict__ Bp, int ldb, double2* const* __restrict__ Cp, int ldc,
    double2 alpha, double2 beta, const int* __restrict__ active, EpiArgs epi) {
  static_assert(S >= 2 && S <= 6, "the FMA digit cut holds c + V below 2^51: S <= 6 (and pair_g covers 21 pairs)");
  constexpr int TM = 64, TN = 64, KC = 32;
  constexpr int PA = S * 2 * TM * 8;
  constexpr int PB = S * 2 * TN * 8;
  __shared__ __attribute__((aligned(16))) unsigned sA[PA];
  __shared__ __attribute__((aligned(16))) unsigned sB[PB];
  __shared__ __attribute__((aligned(16))) int eA[TM];
  __shared__ __attribute__((aligned(16))) int eB[TN];

  int b, tile;
  i8_tile_of(b, tile);
  if (active && !active[b]) return;
  const int tiles_m = (m + TM - 1) / TM;
  const int row0 = (tile % tiles_m) * TM;
  const int col0 = (tile / tiles_m) * TN;
  const double2* __restrict__ A = Ap[b];
  const double2* __restrict__ B = Bp[b];
  double2* __restrict__ C = Cp ? Cp[b] : nullptr;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;  // wave: rows 32 wm .. +31, columns 32 wn .. +31
  const int l32 = lane & 31, hf = lane >> 5;
  // split items: lines tid / 8 and tid / 8 + 32 of A and of B; quarter tid % 8
  const int irc = tid >> 3, iq = tid & 7;

  double cre[16], cim[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) cre[r] = cim[r] = 0.0;

  const bool full = (row0 + TM <= m) && (col0 + TN <= n) && (k % KC == 0);
  const double2* Ar = A + row0;
  const double2* Bc = B + (long long)col0 * ldb;
  double2 v[4][4];
  auto fetch = [&](int k0) {
    if (full) {
      i8_fetch<true>(Ar, 1, lda, m - row0, k0, k, irc, iq, v[0]);
      i8_fetch<true>(Ar, 1, lda, m - row0, k0, k, irc + 32, iq, v[1]);
      i8_fetch<true>(Bc, ldb, 1, n - col0, k0, k, irc, iq, v[2]);
      i8_fetch<true>(Bc, ldb, 1, n - col0, k0, k, irc + 32, iq, v[3]);
    } else {
      i8_fetch<false>(Ar, 1, lda, m - row0, k0, k, irc, iq, v[0]);
      i8_fetch<false>(Ar, 1, lda, m - row0, k0, k, irc + 32, iq, v[1]);
      i8_fetch<false>(Bc, ldb, 1, n - col0, k0, k, irc, iq, v[2]);
      i8_fetch<false>(Bc, ldb, 1, n - col0, k0, k, irc + 32, iq, v[3]);
    }
  };
  const int cB = wn * 32 + l32, rA = wm * 32 + l32;
  const int oB = cB * 8 + 4 * i8_half(cB, hf);
  const int oA = rA * 8 + 4 * i8_half(rA, hf);
  const int nchunk = (k + KC - 1) / KC;
  fetch(0);
  for (int ch = 0; ch < nchunk; ++ch) {
    i8_split_store<S>(v[0], sA, TM, eA, irc, iq);
    i8_split_store<S>(v[1], sA, TM, eA, irc + 32, iq);
    i8_split_store<S>(v[2], sB, TN, eB, irc, iq);
    i8_split_store<S>(v[3], sB, TN, eB, irc + 32, iq);
    if (ch + 1 < nchunk) fetch((ch + 1) * KC);
    __syncthreads();
    i8_chunk_products<S, TM, TN>(sA, sB, eA, eB, oA, oB, rA, wn, hf, cre, cim, [](int) {});
    __syncthreads();  // planes read; free for the next cut (or the reduction)
  }
  i8_epilogue<256>(m, n, row0, col0, rA, wn, hf, cre, cim, C, ldc, alpha, beta, epi, b, tile,
                   reinterpret_cast<double*>(sA), tid);
}
