// Layer structure constants of screened KKR on gfx950: Kambe's Ewald sum
// (skkr_kpar_v6/main/kamstr.f:20-392), the direct reciprocal sum for distant
// layers (directg.f:2-151), their assembly into the matrix of all layer pairs
// of a bulk-like stack and the screening G' = G (1 - alpha G)^-1
// (gstore.f:164-280, scrstr.f:2-138).  The inversion and the product run on
// the batched FP64 MFMA kernels of the k-loop.
//
// Work decomposition: one wavefront per (k-point, distinct layer
// separation).  The shell sums are sequential in the reference and their
// stop rule is per shell, so they stay sequential here; the lanes of the
// wavefront share the per-vector special functions (each lane evaluates
// them, no divergence) and split the output coefficients: Kambe's D_lm over
// lm, the direct sum's (lm, lm') entries over lanes.  Every shell's stop test
// reads the same LDS partials in every lane, so the loop exit is uniform.
#pragma once
#include <hip/hip_runtime.h>

namespace kkr {

struct cx {
  double r, i;
};
__device__ __forceinline__ cx mk(double r, double i = 0.0) { return cx{r, i}; }
__device__ __forceinline__ cx operator+(cx a, cx b) { return cx{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cx operator-(cx a, cx b) { return cx{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cx operator-(cx a) { return cx{-a.r, -a.i}; }
__device__ __forceinline__ cx operator*(cx a, cx b) {
  return cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ __forceinline__ cx operator*(double s, cx a) { return cx{s * a.r, s * a.i}; }
__device__ __forceinline__ cx operator*(cx a, double s) { return cx{a.r * s, a.i * s}; }
__device__ __forceinline__ cx operator/(cx a, double s) { return cx{a.r / s, a.i / s}; }
// Smith's division
__device__ __forceinline__ cx operator/(cx a, cx b) {
  if (fabs(b.r) >= fabs(b.i)) {
    double t = b.i / b.r, d = b.r + b.i * t;
    return cx{(a.r + a.i * t) / d, (a.i - a.r * t) / d};
  }
  double t = b.r / b.i, d = b.r * t + b.i;
  return cx{(a.r * t + a.i) / d, (a.i * t - a.r) / d};
}
__device__ __forceinline__ cx operator/(double s, cx b) { return mk(s) / b; }
__device__ __forceinline__ cx conj(cx a) { return cx{a.r, -a.i}; }
__device__ __forceinline__ double cabs(cx a) { return hypot(a.r, a.i); }
__device__ __forceinline__ cx cexp(cx a) {
  double s, c, e = exp(a.r);
  sincos(a.i, &s, &c);
  return cx{e * c, e * s};
}
// principal square root, branch cut on the negative real axis (signed zero)
__device__ __forceinline__ cx csqrt(cx a) {
  if (a.r == 0.0 && a.i == 0.0) return cx{0.0, a.i};
  double t = sqrt(0.5 * (fabs(a.r) + hypot(a.r, a.i)));
  if (a.r >= 0.0) return cx{t, a.i / (2.0 * t)};
  return cx{fabs(a.i) / (2.0 * t), copysign(t, a.i)};
}
__device__ __forceinline__ cx mul_i(cx a) { return cx{-a.i, a.r}; }
// i^n, exact
__device__ __forceinline__ cx ipow_i(int n) {
  switch (((n % 4) + 4) % 4) {
    case 0: return cx{1.0, 0.0};
    case 1: return cx{0.0, 1.0};
    case 2: return cx{-1.0, 0.0};
    default: return cx{0.0, -1.0};
  }
}
__device__ __forceinline__ cx cipow(cx z, int n) {
  cx r = mk(1.0);
  for (int i = 0; i < (n < 0 ? -n : n); ++i) r = r * z;
  return n < 0 ? 1.0 / r : r;
}

// ---------------------------------------------------------------- w(z)
// Faddeeva function with the branches, series lengths and constants of
// cerf.f:2-148: Taylor series of exp(z^2) int_0^z exp(t^2) for |z| < 4,
// Im z < 1; else Abramowitz-Stegun 7.1.29 (24 terms) whose real part is
// exp(x^2) erfc(x) from SLATEC's Chebyshev series.  The reference writes the
// Chebyshev coefficients, p and erfc's emach as default-real literals, so
// they carry single precision there; the same values are used here.
__constant__ double kERFCS[13] = {
    (double)-.049046121234691808f, (double)-.14226120510371364f, (double).010035582187599796f,
    (double)-.000576876469976748f, (double).000027419931252196f, (double)-.000001104317550734f,
    (double).000000038488755420f,  (double)-.000000001180858253f, (double).000000000032334215f,
    (double)-.000000000000799101f, (double).000000000000017990f,  (double)-.000000000000000371f,
    (double).000000000000000007f};
__constant__ double kERC2CS[23] = {
    (double)-.069601346602309501f, (double)-.041101339362620893f, (double).003914495866689626f,
    (double)-.000490639565054897f, (double).000071574790013770f,  (double)-.000011530716341312f,
    (double).000001994670590201f,  (double)-.000000364266647159f, (double).000000069443726100f,
    (double)-.000000013712209021f, (double).000000002788389661f,  (double)-.000000000581416472f,
    (double).000000000123892049f,  (double)-.000000000026906391f, (double).000000000005942614f,
    (double)-.000000000001332386f, (double).000000000000302804f,  (double)-.000000000000069666f,
    (double).000000000000016208f,  (double)-.000000000000003809f, (double).000000000000000904f,
    (double)-.000000000000000216f, (double).000000000000000052f};
__constant__ double kERFCCS[24] = {
    (double).0715179310202925f,    (double)-.026532434337606719f, (double).001711153977920853f,
    (double)-.000163751663458512f, (double).000019871293500549f,  (double)-.000002843712412769f,
    (double).000000460616130901f,  (double)-.000000082277530261f, (double).000000015921418724f,
    (double)-.000000003295071356f, (double).000000000722343973f,  (double)-.000000000166485584f,
    (double).000000000040103931f,  (double)-.000000000010048164f, (double).000000000002608272f,
    (double)-.000000000000699105f, (double).000000000000192946f,  (double)-.000000000000054704f,
    (double).000000000000015901f,  (double)-.000000000000004729f, (double).000000000000001432f,
    (double)-.000000000000000439f, (double).000000000000000138f,  (double)-.000000000000000048f};
// terms kept for a tail below 0.1 * 1e-12 (single precision): 10, 17, 17
constexpr int kNERFCS = 10, kNERC2CS = 17, kNERFCCS = 17;

__device__ __forceinline__ double csevl(double x, const double* cs, int n) {
  double b1 = 0.0, b0 = 0.0, b2 = 0.0, twox = 2.0 * x;
  for (int i = n - 1; i >= 0; --i) {
    b2 = b1;
    b1 = b0;
    b0 = twox * b1 - b2 + cs[i];
  }
  return 0.5 * (b0 - b2);
}

// exp(x^2) erfc(x) for x >= 0
__device__ double erfce(double x) {
  const double emach = (double)1.0e-12f;
  const double sqeps = sqrt(2.0 * emach);
  const double sqrtpi = (double)1.7724538509055160f;
  if (x <= emach) return 1.0;
  if (x > 1.0 / emach) return 0.0;
  double y = fabs(x);
  if (y <= 1.0) {
    if (y < sqeps) return 1.0 - 2.0 * x / sqrtpi;
    return (1.0 - x * (1.0 + csevl(2.0 * x * x - 1.0, kERFCS, kNERFCS))) * exp(x * x);
  }
  y = y * y;
  if (y <= 4.0) return 1.0 / fabs(x) * (0.5 + csevl((8.0 / y - 5.0) / 3.0, kERC2CS, kNERC2CS));
  return 1.0 / fabs(x) * (0.5 + csevl(8.0 / y - 1.0, kERFCCS, kNERFCCS));
}

__device__ cx cerf(cx z) {
  const double emach = 1.0e-10, eps = 5.0 * emach;
  const double pi = 3.14159265358979323846, api = 1.0 / pi;
  const double p = (double)0.3275911f;
  const double absz = hypot(z.r, z.i);
  if (absz < emach) return mk(1.0);
  const double x = z.r, y = z.i, ax = fabs(x), ay = fabs(y);
  const cx zz = cx{ax, ay}, zzs = zz * zz;
  cx xzzs = mk(0.0), w = mk(0.0);
  int nn = 1;
  if (x != ax) nn = 2;
  if (y != ay) nn = 5 - nn;
  if (nn == 3 || nn == 4) xzzs = (cabs(zzs) < 300.0) ? cexp(-zzs) : mk(0.0);
  bool series = !(ay >= 1.0 || absz >= 4.0);
  if (!series) {
    const double emx2 = exp(-ax * ax), epx = exp(ax), twoy = 2.0 * ay;
    double ex = epx, sn, cs;
    sincos(ax * twoy, &sn, &cs);
    const cx phase = cx{cs, -sn};
    const double t = 1.0 / (1.0 + p * ay);
    const double xe = (1.0 / t - 1.0) / 0.3275911;  // funerr: p in double there
    w = phase * erfce(xe) - api * (phase - mk(1.0)) / twoy;
    cx sum = w;
    bool conv = false;
    for (int n = 1; n <= 24; ++n) {
      const double an = (double)n, ah = an / 2.0;
      const double e2 = 2.0 * api * exp(-ah * ah);
      const cx cian = cx{0.0, an};
      const double denom = an * an + twoy * twoy;
      const cx term1 = e2 * (twoy * phase + 0.5 * ((cian - mk(twoy)) / ex - (cian + mk(twoy)) * ex)) / denom;
      sum = sum - term1;
      if (cabs(term1 / sum) < eps) {
        conv = true;
        break;
      }
      ex = ex * epx;
    }
    if (conv)
      w = sum * emx2;
    else
      series = true;
  }
  if (series) {
    double q = 1.0, factn = -1.0, factd = 1.0;
    xzzs = cexp(-zzs);
    cx term1 = zz, sum = zz;
    for (;;) {
      for (int n = 1; n <= 5; ++n) {
        factn += 2.0;
        factd += 2.0;
        const double fact = factn / (q * factd);
        term1 = fact * zzs * term1;
        sum = sum + term1;
        q += 1.0;
      }
      if (cabs(term1) < eps || q >= 1000.0) break;
    }
    sum = mul_i(2.0 * sqrt(api) * sum);
    w = xzzs + xzzs * sum;
  }
  if (nn == 4) w = 2.0 * xzzs - w;
  if (nn == 2 || nn == 4) w = conj(w);
  if (nn == 3) w = 2.0 * xzzs - w;
  return w;
}

// Gamma(1/2 - n, z), n = 0..N (incgamma.f:6-36)
template <int N>
__device__ __forceinline__ void incgamma(cx z, cx* gam) {
  const double pi = 3.14159265358979323846;
  const cx ex = cexp(-z), sz = csqrt(z);
  const cx a = sz * cexp(cx{0.0, pi / 2.0});
  cx s = sqrt(pi) * ex * cerf(a);
  gam[0] = s;
  double b = 1.0 / 2.0 - 1.0;
  cx zb = sz;
#pragma unroll
  for (int i = 1; i <= N; ++i) {
    zb = zb / z;
    s = (s - zb * ex) / b;
    gam[i] = s;
    b -= 1.0;
  }
}

// Kambe's Delta_n(x, z^2), n = 0..N (delkam.f:4-53)
template <int N>
__device__ __forceinline__ void delkam(cx x, cx zz, cx* delta) {
  const double pi = 3.14159265358979323846;
  const cx z = csqrt(zz), sx = csqrt(x);
  const double dc = sqrt(pi) / 2.0;
  cx b = 2.0 / z;
  const cx a = z / 2.0 / sx;
  const cx ex = cexp(a * a - x);
  const cx c0 = mul_i(sx);
  const cx wm = cerf(-a + c0), wp = cerf(a + c0);
  const cx c = ex * dc;
  delta[0] = (wm + wp) * c;
  if (N >= 1) delta[1] = mul_i((wm - wp) * c * b);
  cx fc = sx;
  b = b * b;
  double d = 0.5;
#pragma unroll
  for (int i = 0; i <= N - 2; ++i) {
    fc = fc / x;
    delta[i + 2] = b * (-d * delta[i + 1] - delta[i] + ex * fc);
    d += 1.0;
  }
}

// real-space integrals I_l, l = 0..N (ilkam.f:4-54)
template <int N>
__device__ __forceinline__ void ilkam(cx alpha, double t, cx kappa, cx* delta) {
  const double pi = 3.14159265358979323846;
  const cx z = kappa * t, x = alpha, sx = csqrt(x);
  cx b = 2.0 / z;
  const double dc = sqrt(pi) / 2.0;
  const cx a = z / 2.0 / sx;
  const cx ex = cexp(x - a * a);
  const cx c0 = mul_i(a);
  const cx wp = cerf(sx + c0), wm = cerf(-sx + c0);
  const cx c = ex * dc;
  cx del = (wp + wm) * c * b;
  cx delm1 = (wp - wm) * c / cx{0.0, 1.0};
  delta[0] = del;
  b = b * b;
  cx cx_ = sx;
  double d = 0.5;
#pragma unroll
  for (int i = 1; i <= N; ++i) {
    cx_ = cx_ / x;
    const cx del2 = (d * del - delm1 + cx_ * ex) * b;
    delta[i] = del2;
    delm1 = del;
    del = del2;
    d += 1.0;
  }
}

// ---------------------------------------------------------------- kernels
// n! for n <= 16
__constant__ double kFac[17] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0,
                                3628800.0, 39916800.0, 479001600.0, 6227020800.0, 87178291200.0,
                                1307674368000.0, 20922789888000.0};

struct StrArgs {
  int nk;                  // k-points in this batch
  const double* k;         // 2 per k
  cx e;                    // energy
  double eta, vol;         // Ewald parameter, 2D cell area
  // Kambe pairs
  int nkam;
  const double* kam_sep;   // 4 per pair: s, sx, sy, sz
  const int* kam_blk;      // output block of the pair
  const double2* rsorted;  // real-space vectors sorted around each pair's (sx, sy)
  const int* r_off;        // first vector of the pair in rsorted
  const int* rsh_off;      // first shell of the pair in rsh
  const int* rsh_n;        // shells of the pair
  const int* rsh;          // vectors per shell
  const int* r_nvec;       // vectors of the pair in rsorted
  const double2* gsorted;  // reciprocal vectors sorted around the origin
  const int* gsh;          // vectors per shell
  int ngsh, ngvec;         // shells, vectors
  const double* gau;       // Gaunt table (lm1, lm2, lm)
  const double* flm;       // F(lm) of kamstr.f:79-91
  const double* yb;        // ylmc.f b(l, m, k), (L2+1)^3
  const cx* ycf;           // ylmc.f prefactors per lm
  // direct pairs
  int ndir;
  const double* dir_sep;   // 3 per pair
  const int* dir_blk;      // output block of gp; gm goes to the next block
  // outputs
  double2* blk;            // per k: nblk blocks of nl2 x nl2 (column-major)
  int nblk;
  int* status;             // per k: |= 2 G sum, 4 R sum, 8 direct sum not converged
                           // (1 is the inversion's zero pivot)
};

// sum over the 64 lanes of a wavefront, returned to every lane
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wavefront per (k, Kambe pair): D_lm of kamstr.f:202-361 (G sum,
// formula 35 or 50), :363-373 (R sum, formula 47) and the Gaunt contraction
// :374-392.  Vectors are taken 64 at a time: first lane t evaluates the
// special functions of vector t of the chunk into LDS, then every lane (one
// per lm of l <= 2 LMAX) adds the chunk's vectors in order to its D_lm and
// the shells' stop tests run as they complete.
template <int LMAX>
__global__ __launch_bounds__(64) void kkr_kamstr_kernel(StrArgs a) {
  constexpr int L2 = 2 * LMAX, NLM2 = (L2 + 1) * (L2 + 1), NL2 = (LMAX + 1) * (LMAX + 1);
  static_assert(NLM2 <= 64, "lanes cover the l <= 2 lmax coefficients");
  const double pi = 3.14159265358979323846, ctol = 1.0e-4, tin = 1.0e-10;
  __shared__ cx s_f0[64][L2 + 1], s_f1[64][L2 + 1], s_cp[64][L2 + 1], s_cm[64][L2 + 1];
  __shared__ cx s_e1[64], s_e2[64];
  __shared__ double s_d0[64], s_d1[64];
  __shared__ cx s_cs[L2 + 1];
  __shared__ cx sd[64];
  const int lane = threadIdx.x;
  const int kb = blockIdx.x, pr = blockIdx.y;
  if (kb >= a.nk || pr >= a.nkam) return;  // uniform per work-group
  const double kx = a.k[2 * kb], ky = a.k[2 * kb + 1];
  const bool act = lane < NLM2;
  const int ln = act ? lane : 0;
  int l = 0;
  while ((l + 1) * (l + 1) <= ln) ++l;
  const int m = ln - l * (l + 1), mm = m < 0 ? -m : m;
  const cx e = a.e, kappa = csqrt(e);
  const cx cs1 = -1.0 / (a.vol * kappa);
  const cx cs2 = -kappa / 2.0 / sqrt(pi);
  const cx alpha = a.eta * e * a.vol / 4.0 / pi;
  const cx w = alpha * 2.0 / e;
  const double fl = a.flm[ln];
  double sx = a.kam_sep[4 * pr + 1], sy = a.kam_sep[4 * pr + 2], sz = a.kam_sep[4 * pr + 3];
  double s = a.kam_sep[4 * pr + 0];
  cx dlm1 = mk(0.0), dlm2 = mk(0.0);
  if (s < ctol && ln == 0 && act) {
    const cx sa = csqrt(alpha);
    const cx c = mul_i(cexp(-alpha) - cerf(sa)) - 1.0 / sqrt(pi) / sa;
    dlm1 = -kappa * (cexp(alpha) * c - cx{0.0, 1.0}) / 2.0 / sqrt(pi);
  }
  if (s > ctol && fabs(sz) < ctol) {
    sz = 1.1 * ctol;
    s = sqrt(sx * sx + sy * sy + sz * sz);
  }
  const bool f35 = (s > ctol && fabs(sz) > ctol), f50 = (s < ctol);
  if (lane == 0) {
    const cx c = kappa * sz;
    cx t = mk(1.0);
    s_cs[0] = t;
#pragma unroll
    for (int i = 1; i <= L2; ++i) {
      t = t * c;
      s_cs[i] = t;
    }
  }
  int status = 0;
  // ---- reciprocal-space sum (s_f1 = Delta_n for formula 35, s_f0 = Gamma for 50)
  {
    int ish = 0, left = a.gsh[0];
    double tst = 1.0e30;
    bool conv = false, done = false;
    for (int v0 = 0; v0 < a.ngvec && !done; v0 += 64) {
      const int nv = min(64, a.ngvec - v0);
      __syncthreads();
      if (lane < nv) {
        const double2 g = a.gsorted[v0 + lane];
        const double gkx = g.x + kx, gky = g.y + ky;
        double km = gkx * gkx + gky * gky;
        const cx kp = csqrt(e - mk(km));
        km = sqrt(km);
        cx ex2 = mk(1.0);
        if (km > 1.0e-10) ex2 = cx{gkx / km, gky / km};
        const cx xc = cx{-1.0, -1.2246467991473532e-16} * (e - mk(km * km)) * w / 2.0;
        cx f[L2 + 1];
        if (f35) {
          const cx zq = kp * sz;
          delkam<L2>(xc, zq * zq, f);
#pragma unroll
          for (int i = 0; i <= L2; ++i) s_f1[lane][i] = f[i];
          double sn, csn;
          sincos(gkx * sx + gky * sy, &sn, &csn);
          s_e1[lane] = cx{csn, sn} * cs1;
        } else {
          incgamma<L2>(xc, f);
#pragma unroll
          for (int i = 0; i <= L2; ++i) s_f0[lane][i] = f[i];
          s_e1[lane] = cs1;
        }
        const cx r = kp / kappa, c = r * r, q = km / kappa;
        cx t = 1.0 / (r * r * r), u = mk(1.0);
#pragma unroll
        for (int i = 0; i <= L2; ++i) {
          t = t * c;
          s_cp[lane][i] = t;
          s_cm[lane][i] = u;
          u = u * q;
        }
        s_e2[lane] = ex2;
      }
      __syncthreads();
      for (int v = 0; v < nv; ++v) {
        const cx ex2 = s_e2[v];
        cx ex = cipow(ex2, l + 1) * s_e1[v];
        for (int mi = -l; mi <= m; ++mi) ex = ex / ex2;
        cx csum = mk(0.0);
        if (f35) {
          const int ipar = (l - mm) % 2;
          for (int n = 0; n <= l - mm; ++n) {
            const cx an = s_cp[v][n] * s_f1[v][n];
            cx b = mk(0.0);
            const int kend = (n + n < l - mm) ? n + n : l - mm;
            for (int k = n + (n + ipar) % 2; k <= kend; k += 2) {
              int i = n + n - k, j = l - k;
              const cx cc = s_cs[i] * s_cm[v][j] / kFac[i] / kFac[k - n];
              i = (l - mm - k) / 2;
              j = (l + mm - k) / 2;
              b = b + cc / kFac[i] / kFac[j];
            }
            csum = csum + an * b;
          }
          if (act) dlm1 = dlm1 + ex * fl * csum * ipow_i(1 - m);
        } else if (f50) {
          if ((l - mm) % 2 == 0) {
            for (int n = 0; n <= (l - mm) / 2; ++n) {
              const cx an = s_cp[v][n] * s_cm[v][l - 2 * n] / kFac[n] * s_f0[v][n];
              csum = csum + an / kFac[(l - mm - 2 * n) / 2] / kFac[(l + mm - 2 * n) / 2];
            }
          }
          if (act) dlm1 = dlm1 + ex * csum * fl * ipow_i(1 - m);
        }
        if (--left == 0) {  // shell ish complete: stop test of kamstr.f:355-357
          const double d = wave_sum(act ? cabs(dlm1) : 0.0);
          if (fabs(tst - d) < tin && ish + 1 >= 3) {
            conv = done = true;
            break;
          }
          tst = d;
          if (++ish >= a.ngsh) {
            done = true;
            break;
          }
          left = a.gsh[ish];
        }
      }
    }
    if (!conv) status |= 2;
  }
  // ---- real-space sum around (sx, sy): s_f0 = I_l, s_e1 = ex1 cs2 / a,
  // s_e2 = a = -kappa rs / 2, s_cm[.][0] = e^{i phi}, s_d0 = cos theta,
  // s_d1 = sin theta (negative: vector skipped)
  if (!status) {
    const double2* rv = a.rsorted + a.r_off[pr];
    const int* rsh = a.rsh + a.rsh_off[pr];
    const int nsh = a.rsh_n[pr], nvec = a.r_nvec[pr];
    double yb[L2 + 1];
#pragma unroll
    for (int k = 0; k <= L2; ++k) yb[k] = a.yb[(l * (L2 + 1) + mm) * (L2 + 1) + k];
    const cx ycf = a.ycf[ln];
    int ish = 0, left = nsh > 0 ? rsh[0] : 0;
    double tst = 1.0e30;
    bool conv = false, done = (nsh == 0);
    for (int v0 = 0; v0 < nvec && !done; v0 += 64) {
      const int nv = min(64, nvec - v0);
      __syncthreads();
      if (lane < nv) {
        const double2 r = rv[v0 + lane];
        const double rk = r.x * kx + r.y * ky;
        double sn, csn;
        sincos(rk, &sn, &csn);
        const cx ex1 = cx{csn, -sn};
        const double rsx = r.x + sx, rsy = r.y + sy, rsz = sz;
        const double rs = sqrt(rsx * rsx + rsy * rsy + rsz * rsz);
        if (rs < 1.0e-5) {
          s_d1[lane] = -1.0;
        } else {
          const double cth = rsz / rs, dxy = sqrt(rsx * rsx + rsy * rsy);
          s_cm[lane][0] = (dxy < 1.0e-6) ? mk(1.0) : cx{rsx / dxy, rsy / dxy};
          s_d0[lane] = cth;
          s_d1[lane] = sqrt(1.0 - cth * cth) + 1.0e-30;
          cx f[L2 + 1];
          ilkam<L2>(alpha, rs, kappa, f);
#pragma unroll
          for (int i = 0; i <= L2; ++i) s_f0[lane][i] = f[i];
          const cx av = -kappa * rs / 2.0;
          s_e2[lane] = av;
          s_e1[lane] = ex1 * cs2 / av;
        }
      }
      __syncthreads();
      for (int v = 0; v < nv; ++v) {
        const double px = s_d1[v];
        if (px >= 0.0) {
          const double cth = s_d0[v];
          double t = 0.0, pk = 1.0;
#pragma unroll
          for (int k = 0; k <= L2; ++k) {
            if (k > l) break;
            t += yb[k] * pk;
            pk *= cth;
          }
          const cx ylm = ycf * pow(px, (double)mm) * t * cipow(s_cm[v][0], m);
          const cx av = s_e2[v];
          cx b = s_e1[v];
          for (int i = 0; i <= l; ++i) b = b * av;
          const cx c = s_f0[v][l] * b;
          if (act) dlm2 = dlm2 + c * conj(ylm);
        }
        if (--left == 0) {
          const double d = wave_sum(act ? cabs(dlm2) : 0.0);
          if (fabs(tst - d) < tin && ish + 1 >= 3) {
            conv = done = true;
            break;
          }
          tst = d;
          if (++ish >= nsh) {
            done = true;
            break;
          }
          left = rsh[ish];
        }
      }
    }
    if (!conv) status |= 4;
  }
  if (status) {
    if (lane == 0) atomicOr(a.status + kb, status);
    return;
  }
  // ---- Gaunt contraction alm(lm1, lm2) = sum_lm gau (D1 + D2)
  __syncthreads();
  sd[lane] = dlm1 + dlm2;
  __syncthreads();
  double2* out = a.blk + ((size_t)kb * a.nblk + a.kam_blk[pr]) * (NL2 * NL2);
  for (int e2 = lane; e2 < NL2 * NL2; e2 += 64) {
    const int lm1 = e2 / NL2, lm2 = e2 % NL2;  // o-order of gauntcalc.f
    const double* g = a.gau + (size_t)e2 * NLM2;
    cx acc = mk(0.0);
    for (int lm = 0; lm < NLM2; ++lm) acc = acc + g[lm] * sd[lm];
    out[lm1 + lm2 * NL2] = make_double2(acc.r, acc.i);
  }
}

// Y_lm of complex angles (sphrm.f:2-68), all lm <= (LMAX+1)^2
template <int LMAX>
__device__ __forceinline__ void sphrm(cx* ylm, cx ct, cx st, cx cf) {
  const double pi = 3.14159265358979323846, pii4 = 0.25 / pi;
  ylm[0] = mk(sqrt(pii4));
  double a = 1.0, b = 1.0, asg = 1.0;
  cx sf = mk(1.0), sa = mk(1.0);
  int lp = 1;
#pragma unroll
  for (int l = 1; l <= LMAX; ++l) {
    const double fl = (double)l;
    a = 0.5 * a * fl * (2.0 * fl - 1.0);
    b = fl * b;
    asg = -asg;
    const int lm = lp + 1;
    lp = lp + l + l + 1;
    sf = sf * cf;
    cx fac = sqrt((2.0 * fl + 1.0) * a / (4.0 * pi * b * b)) * sa;
    ylm[lm - 1] = fac * st / sf;
    ylm[lp - 1] = asg * fac * st * sf;
    fac = sqrt(2.0 * fl) * fac * ct;
    ylm[lm] = fac * cf / sf;
    ylm[lp - 2] = -asg * fac * sf / cf;
    sa = sa * st;
  }
#pragma unroll
  for (int m = 2; m <= LMAX; ++m) {
    const int mm = m + m - 4;
    const double fm = (double)(m - 2);
    double a2 = sqrt(1.0 / (fm + fm + 3.0));
    int ln = m * m - 1, lm = ln - m - m + 2;
#pragma unroll
    for (int l = m; l <= LMAX; ++l) {
      const double fl = (double)l;
      const double b2 = sqrt((fl + fm) * (fl - fm) / ((fl + fl + 1.0) * (fl + fl - 1.0)));
      const int lpp = ln + l + l;
      ylm[lpp - 1] = (ct * ylm[ln - 1] - a2 * ylm[lm - 1]) / b2;
      ylm[lpp - mm - 1] = (ct * ylm[ln - mm - 1] - a2 * ylm[lm - mm - 1]) / b2;
      a2 = b2;
      lm = ln;
      ln = lpp;
    }
  }
}

// One wavefront per (k, distant pair): gp, gm of directg.f:52-130.  Lane t
// first evaluates Y_lm(theta_k, phi_k) and the two phase factors of vector t
// of a 64-vector chunk into LDS; then lanes own entries e = lmj + NL2 * lmi
// of gp(lmj, lmi), gm(lmj, lmi) and add the chunk's vectors in order.  The
// stop test of directg.f:125 reads, at entry (lmi, lmj), gp(lmi, lmj) as it
// stands inside the last vector's double loop: updated already when
// lmj <= lmi.
template <int LMAX>
__global__ __launch_bounds__(64) void kkr_directg_kernel(StrArgs a) {
  constexpr int NL2 = (LMAX + 1) * (LMAX + 1), NE = NL2 * NL2, PER = (NE + 63) / 64;
  const double pi = 3.14159265358979323846, dthr = 1.0e-12;
  __shared__ cx s_ylm[64][NL2];
  __shared__ cx s_fp[64], s_fm[64];
  const int lane = threadIdx.x;
  const int kb = blockIdx.x, pr = blockIdx.y;
  if (kb >= a.nk || pr >= a.ndir) return;
  const double akx = a.k[2 * kb], aky = a.k[2 * kb + 1];
  const double s1 = a.dir_sep[3 * pr], s2 = a.dir_sep[3 * pr + 1], s3 = a.dir_sep[3 * pr + 2];
  const cx kappa2 = a.e, kappa = csqrt(kappa2);
  const double f1 = 8.0 * pi * pi / a.vol;
  // per entry: Y indices and the exact unit factors of directg.f:88-104
  // (ylm1 = sl il Y_lmi, ylm2 = il sm Y_lmmm, ylm3 = il sm Y_lmi, ylm4 = sl il Y_lmmm)
  int yi[PER], yj[PER];
  cx p1[PER], p2[PER], p3[PER], p4[PER];
  bool ok[PER], upd[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e2 = lane + 64 * q;
    ok[q] = e2 < NE;
    const int lmj = ok[q] ? e2 % NL2 : 0, lmi = ok[q] ? e2 / NL2 : 0;
    int li = 0, lj = 0;
    while ((li + 1) * (li + 1) <= lmi) ++li;
    while ((lj + 1) * (lj + 1) <= lmj) ++lj;
    const int mi = lmi - li * (li + 1), mj = lmj - lj * (lj + 1);
    const double sli = (li % 2) ? -1.0 : 1.0, slj = (lj % 2) ? -1.0 : 1.0;
    const double smi = (mi % 2) ? -1.0 : 1.0, smj = (mj % 2) ? -1.0 : 1.0;
    yi[q] = lmi;
    yj[q] = lmj - mj - mj;
    p1[q] = sli * ipow_i(li);
    p2[q] = ipow_i(lj) * smj;
    p3[q] = ipow_i(li) * smi;
    p4[q] = slj * ipow_i(lj);
    upd[q] = (e2 / NL2) <= (e2 % NL2);  // the stop test's read of entry (row e%NL2, col e/NL2)
  }
  cx gp[PER], gm[PER], gpo[PER], gmo[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) gp[q] = gm[q] = gpo[q] = gmo[q] = mk(0.0);
  int ish = 0, left = a.gsh[0];
  double term = 0.0;
  bool conv = false, done = false;
  for (int v0 = 0; v0 < a.ngvec && !done; v0 += 64) {
    const int nv = min(64, a.ngvec - v0);
    __syncthreads();
    if (lane < nv) {
      const double2 g = a.gsorted[v0 + lane];
      const double gkx = akx + g.x, gky = aky + g.y;
      const double gk2 = gkx * gkx + gky * gky, gk = sqrt(gk2);
      const cx gkz = csqrt(kappa2 - mk(gk2));
      cx phi = mk(1.0);
      if (gk > (double)1.0e-6f) phi = cx{gkx / gk, gky / gk};
      const cx cth = gkz / kappa, sth = mk(gk) / kappa;
      cx ylm[NL2];
      sphrm<LMAX>(ylm, cth, sth, phi);
#pragma unroll
      for (int j = 0; j < NL2; ++j) s_ylm[lane][j] = ylm[j];
      const cx argp = mul_i(mk(gkx * s1 + gky * s2) + gkz * s3);
      const cx argm = mul_i(mk(-gkx * s1 - gky * s2) + gkz * s3);
      s_fp[lane] = cexp(argp) * f1 / gkz;
      s_fm[lane] = cexp(argm) * f1 / gkz;
    }
    __syncthreads();
    for (int v = 0; v < nv; ++v) {
      const cx fp = s_fp[v], fm = s_fm[v];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (ok[q]) {
          const cx ya = s_ylm[v][yi[q]], yb = s_ylm[v][yj[q]];
          const cx ctp = (p1[q] * ya) * fp * (p2[q] * yb), ctm = (p3[q] * ya) * fm * (p4[q] * yb);
          gpo[q] = gp[q];
          gmo[q] = gm[q];
          gp[q] = gp[q] + cx{ctp.i, -ctp.r};  // ctp * (-i)
          gm[q] = gm[q] + cx{ctm.i, -ctm.r};
          term += cabs(ctp) + cabs(ctm);
        }
      }
      if (--left == 0) {  // shell ish complete
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < PER; ++q)
          if (ok[q]) sum += upd[q] ? cabs(gp[q]) + cabs(gm[q]) : cabs(gpo[q]) + cabs(gmo[q]);
        const double tt = wave_sum(term), ss = wave_sum(sum);
        if (tt / ss < dthr && ish >= 1) {
          conv = done = true;
          break;
        }
        term = 0.0;
        if (++ish >= a.ngsh) {
          done = true;
          break;
        }
        left = a.gsh[ish];
      }
    }
  }
  if (!conv) {
    if (lane == 0) atomicOr(a.status + kb, 8);
    return;
  }
  double2* outp = a.blk + ((size_t)kb * a.nblk + a.dir_blk[pr]) * NE;
  double2* outm = outp + NE;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e2 = lane + 64 * q;
    if (e2 < NE) {
      outp[e2] = make_double2(gp[q].r, gp[q].i);
      outm[e2] = make_double2(gm[q].r, gm[q].i);
    }
  }
}

// g (all layer pairs) and 1 - alpha g of one k: element (r, c) of block
// (r / nl2, c / nl2) comes from block bmap of the pair table
__global__ __launch_bounds__(256) void kkr_assemble_kernel(int N, int nl2, int nlay, int s, int lmax1,
                                                           const int* __restrict__ bmap,
                                                           const double2* __restrict__ blk, int nblk,
                                                           const double2* __restrict__ alpha,
                                                           const int* __restrict__ lof, double2* g,
                                                           double2* g1, int nk) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int kb = blockIdx.y;
  if (idx >= (long long)N * N || kb >= nk) return;
  const int r = (int)(idx % N), c = (int)(idx / N);
  const int la = r / nl2, m = r % nl2, lb = c / nl2, n = c % nl2;
  const int bi = bmap[la * nlay + lb];
  const double2 v = blk[((size_t)kb * nblk + bi) * nl2 * nl2 + m + n * nl2];
  const double2 al = alpha[lof[m] + lmax1 * (la % s)];
  const size_t o = (size_t)kb * N * N + idx;
  g[o] = v;
  double2 t = make_double2(-(al.x * v.x - al.y * v.y), -(al.x * v.y + al.y * v.x));
  if (r == c) t.x += 1.0;
  g1[o] = t;
}

// g2d(m, n, jj, icount, icpl) = G'(layer i, layer i + jj) of the 3 s layers
// of the middle PLs (gstore.f:205-280)
__global__ __launch_bounds__(256) void kkr_extract_kernel(int N, int nl2, int s, int nextra,
                                                          const double2* __restrict__ gs,
                                                          double2* __restrict__ g2d, int nk) {
  const int per = nl2 * nl2 * (2 * s + 1) * s * 3;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int kb = blockIdx.y;
  if (idx >= per || kb >= nk) return;
  int t = (int)idx;
  const int m = t % nl2;
  t /= nl2;
  const int n = t % nl2;
  t /= nl2;
  const int jj = t % (2 * s + 1) - s;
  t /= (2 * s + 1);
  const int icount = t % s, icpl = t / s;
  const int i = nextra * s + icpl * s + icount, j = i + jj;
  g2d[(size_t)kb * per + idx] = gs[(size_t)kb * N * N + (size_t)(i * nl2 + m) + (size_t)(j * nl2 + n) * N];
}

// PL blocks of the k-loop from the screened constants of one k (packer.f:36-97
// with equal ninprc, irel = 0): site i, j of the middle principal layer
// (plane 1) give M00, M01 (j <= i) and M10 (j >= i); M01 of plane 0 and M10
// of plane 2 are the outer factors of the lead self-energies (dugo.f:71-75,
// 98, 106).  Row (i, k1) of an n x n block is i * nl2 + k1.
__global__ __launch_bounds__(256) void kkr_pack_kernel(int N, int nl2, int s, int nextra,
                                                       const double2* __restrict__ gs, int nb, long long nn,
                                                       double2* M00, double2* M01, double2* M10,
                                                       double2* MtL, double2* MtR) {
  const int n = nl2 * s;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int kb = blockIdx.y;
  if (idx >= (long long)n * n || kb >= nb) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  const int i = r / nl2, k1 = r % nl2, jp = c / nl2, k2 = c % nl2;
  const double2* g = gs + (size_t)kb * N * N;
  auto at = [&](int la, int lb) {
    const double2 v = g[(size_t)(la * nl2 + k1) + (size_t)(lb * nl2 + k2) * N];
    return make_double2(-v.x, -v.y);
  };
  const double2 zero = make_double2(0.0, 0.0);
  const int p0 = nextra * s, p1 = p0 + s, p2 = p1 + s;
  const size_t o = (size_t)kb * nn + idx;
  M00[o] = at(p1 + i, p1 + jp);
  M01[o] = (jp <= i) ? at(p1 + i, p2 + jp) : zero;
  M10[o] = (jp >= i) ? at(p1 + i, p0 + jp) : zero;
  MtL[o] = (jp <= i) ? at(p0 + i, p1 + jp) : zero;
  MtR[o] = (jp >= i) ? at(p2 + i, p1 + jp) : zero;
}

// The same blocks in the kappa-mu basis (lk = 2 nl2), packer.f with irel = 1:
// each nl2 x nl2 block a goes through relmtrx.f:26-42,
//   b(i, j) = u1(i) (f a)(ind1 i, ind1 j) u1(j) + u2(i) (f a)(ind2 i, ind2 j) u2(j),
// f = psq / E; indices past the lm block carry coefficient 0.
__global__ __launch_bounds__(256) void kkr_pack_rel_kernel(int N, int nl2, int s, int nextra,
                                                           const double2* __restrict__ gs, int nb, long long nn,
                                                           double fr, double fi, const double* __restrict__ u1,
                                                           const double* __restrict__ u2, const int* __restrict__ i1,
                                                           const int* __restrict__ i2, double2* M00, double2* M01,
                                                           double2* M10, double2* MtL, double2* MtR) {
  const int lk = 2 * nl2, n = lk * s;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int kb = blockIdx.y;
  if (idx >= (long long)n * n || kb >= nb) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  const int i = r / lk, kr = r % lk, jp = c / lk, kc = c % lk;
  const double2* g = gs + (size_t)kb * N * N;
  const cx f = cx{fr, fi};
  const double a1 = u1[kr], b1 = u1[kc], a2 = u2[kr], b2 = u2[kc];
  const int p1 = i1[kr], q1 = i1[kc], p2 = i2[kr], q2 = i2[kc];
  auto at = [&](int la, int lb) {
    cx v1 = mk(0.0), v2 = mk(0.0);
    if (p1 < nl2 && q1 < nl2) {
      const double2 t = g[(size_t)(la * nl2 + p1) + (size_t)(lb * nl2 + q1) * N];
      v1 = f * cx{-t.x, -t.y};
    }
    if (p2 < nl2 && q2 < nl2) {
      const double2 t = g[(size_t)(la * nl2 + p2) + (size_t)(lb * nl2 + q2) * N];
      v2 = f * cx{-t.x, -t.y};
    }
    const cx b = (a1 * v1) * b1 + (a2 * v2) * b2;
    return make_double2(b.r, b.i);
  };
  const double2 zero = make_double2(0.0, 0.0);
  const int p0 = nextra * s, pl1 = p0 + s, pl2 = pl1 + s;
  const size_t o = (size_t)kb * nn + idx;
  M00[o] = at(pl1 + i, pl1 + jp);
  M01[o] = (jp <= i) ? at(pl1 + i, pl2 + jp) : zero;
  M10[o] = (jp >= i) ? at(pl1 + i, p0 + jp) : zero;
  MtL[o] = (jp <= i) ? at(p0 + i, pl1 + jp) : zero;
  MtR[o] = (jp >= i) ? at(pl2 + i, pl1 + jp) : zero;
}

// dst[b] (ldd) = src[b] (lds), an m x n block per batch entry
__global__ __launch_bounds__(256) void kkr_copy_block_kernel(int m, int n, const double2* const* __restrict__ src,
                                                             int lds, double2* const* __restrict__ dst, int ldd) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= (long long)m * n) return;
  const int r = (int)(e % m), c = (int)(e / m);
  dst[b][r + (long long)c * ldd] = src[b][r + (long long)c * lds];
}

// producer status of a k -> k-loop status bits (8: not converged, 1: pivot)
__global__ void kkr_status_kernel(int nb, const int* __restrict__ st, int* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const int v = st[j];
  if (v & (2 | 4 | 8)) out[j] |= 8;
  if (v & 1) out[j] |= 1;
}

}  // namespace kkr
