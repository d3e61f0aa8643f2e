/*
 * CPU FP64 oracle of the layer structure constants (screened KKR, 2D
 * periodic layers).  TEST INFRASTRUCTURE ONLY: the checker of the HIP
 * producer, never part of the product.
 *
 * Each routine restates one reference subroutine of skkr_kpar_v6/main,
 * keeping its branch logic, series, recursions and summation order:
 *
 *   orc_cerf      cerf.f:2-148     Faddeeva w(z): power series |z| < 4 and
 *                                  Im z < 1, else Abramowitz-Stegun 7.1.29
 *                                  with exp(x^2) erfc(x) from SLATEC's
 *                                  Chebyshev series (erfce, csevl, inits)
 *   orc_incgamma  incgamma.f:6-36  Gamma(1/2 - n, z), n = 0..N
 *   orc_delkam    delkam.f:4-53    Kambe's Delta_n(x, z^2)
 *   orc_ilkam     ilkam.f:4-54     Kambe's real-space integrals I_l
 *   ylmc          ylmc.f:5-123     complex Y_lm of a real vector
 *   sphrm         sphrm.f:2-68     Y_lm of complex (cos, sin, e^{i phi})
 *   blm, gaunt    blm.f:2-59, gauntcalc.f:2-28 (fac.f:2-10)
 *   lattice       vector2d.f:2-150, latgen2d.f:2-33, genlatt2d.f:2-63,
 *                 primkr2d.f:236-246 (reciprocal basis)
 *   vecsrt        vecsrt2d.f:2-95 with the heapsort of indekx.f:2-47
 *   orc_kamstr    kamstr.f:20-392  Ewald (Kambe) sum for one layer pair
 *   orc_directg   directg.f:2-151  reciprocal-space sum for distant layers
 *   orc_gstore    gstore.f:6-283 (bulkgeo branches, :99-160 and :164-280),
 *                 scrstr.f:2-138 (screening), scrinv.f:2-63 (the same
 *                 column-pivot Gauss-Jordan as gjinv.f: orc_gjinv)
 *
 * Matrices are column-major double complex, as complex*16 arrays are in the
 * reference.  Constants the reference writes as default-real literals
 * (cerf.f:11-12, the SLATEC coefficient tables, funerr's emach) are rounded
 * to single precision here too, because that is the value they carry there.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;
int orc_gjinv(cplx *a, int n, int lda);

#define PI_F 3.1415926535897932384626
#define LMX 12
#define SNG(x) ((double)(float)(x))

/* ------------------------------------------------------------------ cerf */

/* SLATEC erfc Chebyshev tables (public domain, W. Fullerton, LANL); values
 * as single-precision literals like the reference's DATA statements. */
static const float ERFCS[13] = {
    -.049046121234691808f, -.14226120510371364f,   .010035582187599796f, -.000576876469976748f,
    .000027419931252196f,  -.000001104317550734f,  .000000038488755420f, -.000000001180858253f,
    .000000000032334215f,  -.000000000000799101f,  .000000000000017990f, -.000000000000000371f,
    .000000000000000007f};
static const float ERC2CS[23] = {
    -.069601346602309501f, -.041101339362620893f, .003914495866689626f,  -.000490639565054897f,
    .000071574790013770f,  -.000011530716341312f, .000001994670590201f,  -.000000364266647159f,
    .000000069443726100f,  -.000000013712209021f, .000000002788389661f,  -.000000000581416472f,
    .000000000123892049f,  -.000000000026906391f, .000000000005942614f,  -.000000000001332386f,
    .000000000000302804f,  -.000000000000069666f, .000000000000016208f,  -.000000000000003809f,
    .000000000000000904f,  -.000000000000000216f, .000000000000000052f};
static const float ERFCCS[24] = {
    .0715179310202925f,    -.026532434337606719f, .001711153977920853f,  -.000163751663458512f,
    .000019871293500549f,  -.000002843712412769f, .000000460616130901f,  -.000000082277530261f,
    .000000015921418724f,  -.000000003295071356f, .000000000722343973f,  -.000000000166485584f,
    .000000000040103931f,  -.000000000010048164f, .000000000002608272f,  -.000000000000699105f,
    .000000000000192946f,  -.000000000000054704f, .000000000000015901f,  -.000000000000004729f,
    .000000000000001432f,  -.000000000000000439f, .000000000000000138f,  -.000000000000000048f};

/* number of terms whose tail stays within eta (cerf.f inits) */
static int inits(const float *os, int nos, double eta) {
  double err = 0.0;
  int i = nos;
  for (int ii = 1; ii <= nos; ++ii) {
    i = nos + 1 - ii;
    err += fabs((double)os[i - 1]);
    if (err > eta) break;
  }
  return i;
}

/* n-term Chebyshev series at x (cerf.f csevl) */
static double csevl(double x, const float *cs, int n) {
  double b1 = 0.0, b0 = 0.0, b2 = 0.0, twox = 2.0 * x;
  for (int i = 1; i <= n; ++i) {
    b2 = b1;
    b1 = b0;
    b0 = twox * b1 - b2 + (double)cs[n - i];
  }
  return 0.5 * (b0 - b2);
}

/* exp(x^2) erfc(x), x >= 0 (cerf.f erfce with emach = 1.0e-12 from funerr) */
static double erfce(double x) {
  const double emach = SNG(1.0e-12);
  const double eta = SNG(0.1) * emach;
  const double xmax = 1.0 / emach;
  const double sqeps = sqrt(2.0 * emach);
  const double sqrtpi = SNG(1.7724538509055160);
  const int nterf = inits(ERFCS, 13, eta), nterfc = inits(ERFCCS, 24, eta),
            nterc2 = inits(ERC2CS, 23, eta);
  if (x <= emach) return 1.0;
  if (x > xmax) return 0.0;
  double y = fabs(x);
  if (y <= 1.0) {
    if (y < sqeps) return 1.0 - 2.0 * x / sqrtpi;
    return (1.0 - x * (1.0 + csevl(2.0 * x * x - 1.0, ERFCS, nterf))) * exp(x * x);
  }
  y = y * y;
  if (y <= 4.0) return 1.0 / fabs(x) * (0.5 + csevl((8.0 / y - 5.0) / 3.0, ERC2CS, nterc2));
  return 1.0 / fabs(x) * (0.5 + csevl(8.0 / y - 1.0, ERFCCS, nterfc));
}

static double funerr(double t) {
  const double p = 0.3275911;
  double x = (1.0 / t - 1.0) / p;
  return erfce(x);
}

cplx orc_cerf(cplx z) {
  const double emach = 1.0e-10, eps = 5.0 * emach;
  const double pi = 4.0 * atan(1.0), api = 1.0 / pi;
  const double p = SNG(0.3275911);
  double absz = cabs(z);
  if (absz < emach) return 1.0;
  double x = creal(z), y = cimag(z), ax = fabs(x), ay = fabs(y);
  cplx zz = CMPLX(ax, ay), zzs = zz * zz, xzzs = 0.0, w;
  int nn = 1;
  if (x != ax) nn = 2;
  if (y != ay) nn = 5 - nn;
  if (nn == 3 || nn == 4) xzzs = (cabs(zzs) < 300.0) ? cexp(-zzs) : 0.0;
  int series = !(ay >= 1.0 || absz >= 4.0);
  if (!series) {
    /* Abramowitz-Stegun 7.1.29, at most 24 terms */
    double expn2[25];
    for (int n = 1; n <= 24; ++n) {
      double an = (double)n / 2.0;
      expn2[n] = 2.0 * api * exp(-an * an);
    }
    double emx2 = exp(-ax * ax), epx = exp(ax), ex = epx, twoy = 2.0 * ay;
    cplx phase = cos(ax * twoy) - I * sin(ax * twoy);
    double t = 1.0 / (1.0 + p * ay);
    w = phase * funerr(t) - api * (phase - 1.0) / twoy;
    cplx sum = w;
    int conv = 0;
    for (int n = 1; n <= 24; ++n) {
      double an = (double)n;
      cplx cian = I * an;
      double denom = an * an + twoy * twoy;
      cplx term1 = expn2[n] * (twoy * phase + 0.5 * ((cian - twoy) / ex - (cian + twoy) * ex)) / denom;
      sum = sum - term1;
      if (cabs(term1 / sum) < eps) {
        conv = 1;
        break;
      }
      ex = ex * epx;
    }
    if (conv)
      w = sum * emx2;
    else
      series = 1;
  }
  if (series) {
    double q = 1.0, factn = -1.0, factd = 1.0;
    xzzs = cexp(-zzs);
    cplx term1 = zz, sum = zz;
    for (;;) {
      for (int n = 1; n <= 5; ++n) {
        factn += 2.0;
        factd += 2.0;
        double fact = factn / (q * factd);
        term1 = fact * zzs * term1;
        sum = sum + term1;
        q += 1.0;
      }
      if (cabs(term1) < eps || q >= 1000.0) break;
    }
    double fact = 2.0 * sqrt(api);
    sum = fact * I * sum;
    w = xzzs + xzzs * sum;
  }
  switch (nn) {
    case 4: w = 2.0 * xzzs - w; /* fall through */
    case 2: w = conj(w); break;
    case 3: w = 2.0 * xzzs - w; break;
    default: break;
  }
  return w;
}

void orc_cerf_v(int n, const cplx *z, cplx *w) {
  for (int i = 0; i < n; ++i) w[i] = orc_cerf(z[i]);
}

/* ------------------------------------------------- Kambe's special functions */

void orc_incgamma(int n, cplx z, cplx *gam) {
  const double pi = PI_F;
  cplx ex = cexp(-z), sz = csqrt(z);
  cplx a = sz * cexp(I * pi / 2.0);
  cplx s = sqrt(pi) * ex * orc_cerf(a);
  gam[0] = s;
  if (n < 0) {
    for (int i = 1; i <= -n; ++i) gam[i] = cpow(z, (double)i - 0.5) * ex + ((double)i - 0.5) * gam[i - 1];
    return;
  }
  double b = 1.0 / 2.0 - 1.0;
  cplx zb = sz;
  for (int i = 1; i <= n; ++i) {
    zb = zb / z;
    s = (s - zb * ex) / b;
    gam[i] = s;
    b -= 1.0;
  }
}

void orc_delkam(int n, cplx x, cplx zz, cplx *delta) {
  const double pi = PI_F;
  cplx z = csqrt(zz), sx = csqrt(x);
  double dc = sqrt(pi) / 2.0;
  cplx b = 2.0 / z, a = z / 2.0 / sx;
  cplx ex = cexp(a * a - x);
  cplx c = I * sx;
  cplx am = -a + c, ap = a + c;
  cplx wm = orc_cerf(am), wp = orc_cerf(ap);
  c = ex * dc;
  cplx del0 = (wm + wp) * c;
  cplx del1 = (wm - wp) * c * b * I;
  delta[0] = del0;
  if (n >= 1) delta[1] = del1;
  cplx fc = sx;
  b = b * b;
  double d = 0.5;
  for (int i = 0; i <= n - 2; ++i) {
    fc = fc / x;
    delta[i + 2] = b * (-d * delta[i + 1] - delta[i] + ex * fc);
    d += 1.0;
  }
}

void orc_ilkam(int l, cplx alpha, double t, cplx kappa, cplx *delta) {
  const double pi = PI_F;
  cplx z = kappa * t, x = alpha, sx = csqrt(x);
  cplx b = 2.0 / z;
  double dc = sqrt(pi) / 2.0;
  cplx a = z / 2.0 / sx;
  cplx ex = cexp(x - a * a);
  cplx c = I * a;
  cplx ap = sx + c, am = -sx + c;
  cplx wp = orc_cerf(ap), wm = orc_cerf(am);
  c = ex * dc;
  cplx del = (wp + wm) * c * b;
  cplx delm1 = (wp - wm) * c / I;
  delta[0] = del;
  b = b * b;
  cplx cx = sx;
  double d = 0.5;
  for (int i = 1; i <= l; ++i) {
    cx = cx / x;
    cplx del2 = (d * del - delm1 + cx * ex) * b;
    delta[i] = del2;
    delm1 = del;
    del = del2;
    d += 1.0;
  }
}

/* complex integer power by repeated multiplication */
static cplx cipow(cplx z, int n) {
  if (n < 0) return 1.0 / cipow(z, -n);
  cplx r = 1.0;
  for (int i = 0; i < n; ++i) r *= z;
  return r;
}

static cplx ipow_i(int n) { /* i^n, exact */
  static const cplx t[4] = {1.0, I, -1.0, -I};
  return t[((n % 4) + 4) % 4];
}

/* complex spherical harmonics of a real vector, l <= lmax (ylmc.f) */
static void ylmc(double x, double y, double z, int lmax, cplx *ylm) {
  static double b[LMX + 1][LMX + 1][LMX + 1];
  static cplx cf[(LMX + 1) * (LMX + 1)];
  static int init = 0;
  if (!init) {
    double a[LMX + 1][LMX + 1], f[LMX + 1][LMX + 1], fact[2 * LMX + 1];
    memset(a, 0, sizeof a);
    memset(f, 0, sizeof f);
    memset(b, 0, sizeof b);
    a[0][0] = 1.0;
    a[1][1] = 1.0;
    a[2][0] = -0.5;
    a[2][2] = 1.5;
    for (int l = 3; l <= LMX; ++l) {
      /* float(2l-1)/float(l): a single-precision quotient in ylmc.f:48-49 */
      double fa1 = (double)((float)(2 * l - 1) / (float)l), fa2 = -(double)((float)(l - 1) / (float)l);
      a[l][0] = fa2 * a[l - 2][0];
      for (int n = 1; n <= l - 2; ++n) a[l][n] = fa1 * a[l - 1][n - 1] + fa2 * a[l - 2][n];
      a[l][l - 1] = fa1 * a[l - 1][l - 2];
      a[l][l] = fa1 * a[l - 1][l - 1];
    }
    for (int n = 0; n <= LMX; ++n) {
      f[n][0] = 1.0;
      for (int m = 0; m <= n - 1; ++m) f[n][m + 1] = f[n][m] * (double)(n - m);
    }
    for (int l = 0; l <= LMX; ++l)
      for (int m = 0; m <= l; ++m)
        for (int k = 0; k <= l - m; ++k) b[l][m][k] = a[l][k + m] * f[k + m][m];
    fact[0] = 1.0;
    for (int n = 1; n <= 2 * LMX; ++n) fact[n] = fact[n - 1] * (double)n;
    const double pi4 = 4.0 * acos(-1.0);
    int i = 0;
    for (int l = 0; l <= LMX; ++l)
      for (int m = -l; m <= l; ++m) {
        int mb = abs(m);
        cf[i++] = sqrt((double)(2 * l + 1) * fact[l - mb] / (pi4 * fact[l + mb])) * ipow_i(m + mb);
      }
    init = 1;
  }
  double r = sqrt(x * x + y * y + z * z), cth;
  cplx ephi;
  if (r < 1.0e-16) {
    cth = 0.0;
    ephi = 1.0;
  } else {
    cth = z / r;
    double d = sqrt(x * x + y * y);
    ephi = (d < 1.0e-6) ? 1.0 : (x + I * y) / d;
  }
  double px = sqrt(1.0 - cth * cth) + 1.0e-30;
  double p[LMX + 1];
  p[0] = 1.0;
  for (int k = 1; k <= lmax; ++k) p[k] = p[k - 1] * cth;
  int i = 0;
  for (int l = 0; l <= lmax; ++l)
    for (int m = -l; m <= l; ++m) {
      int mb = abs(m);
      double t = 0.0;
      for (int k = 0; k <= l; ++k) t += b[l][mb][k] * p[k];
      ylm[i] = cf[i] * pow(px, mb) * t * cipow(ephi, m);
      ++i;
    }
}

/* spherical harmonics of complex angles (sphrm.f, Pendry appendix A) */
static void sphrm(cplx *ylm, cplx ct, cplx st, cplx cf, int lmax) {
  const double pi = 4.0 * atan(1.0), pii4 = 0.25 / pi;
  ylm[0] = sqrt(pii4);
  if (lmax == 0) return;
  double a = 1.0, b = 1.0, asg = 1.0;
  cplx sf = 1.0, sa = 1.0;
  int lp = 1;
  for (int l = 1; l <= lmax; ++l) {
    double fl = (double)l;
    a = 0.5 * a * fl * (2.0 * fl - 1.0);
    b = fl * b;
    asg = -asg;
    int lm = lp + 1;
    lp = lp + l + l + 1;
    sf = sf * cf;
    cplx fac = sqrt((2.0 * fl + 1.0) * a / (4.0 * pi * b * b)) * sa;
    ylm[lm - 1] = fac * st / sf;
    ylm[lp - 1] = asg * fac * st * sf;
    fac = sqrt(2.0 * fl) * fac * ct;
    ylm[lm] = fac * cf / sf;
    ylm[lp - 2] = -asg * fac * sf / cf;
    sa = sa * st;
  }
  if (lmax == 1) return;
  for (int m = 2; m <= lmax; ++m) {
    int mm = m + m - 4;
    double fm = (double)(m - 2);
    double a2 = sqrt(1.0 / (fm + fm + 3.0));
    int ln = m * m - 1, lm = ln - m - m + 2;
    for (int l = m; l <= lmax; ++l) {
      double fl = (double)l;
      double b2 = sqrt((fl + fm) * (fl - fm) / ((fl + fl + 1.0) * (fl + fl - 1.0)));
      int lpp = ln + l + l;
      ylm[lpp - 1] = (ct * ylm[ln - 1] - a2 * ylm[lm - 1]) / b2;
      ylm[lpp - mm - 1] = (ct * ylm[ln - mm - 1] - a2 * ylm[lm - mm - 1]) / b2;
      a2 = b2;
      lm = ln;
      ln = lpp;
    }
  }
}

/* ----------------------------------------------------------------- Gaunt */

static double facnm(int n, int m) { /* n!/(n-m)! (fac.f) */
  double f = 1.0;
  if (m == 0) return f;
  for (int i = n - m + 1; i <= n; ++i) f *= (double)i;
  return f;
}

static double blm(int il1, int im1, int il2, int im2, int il3, int im3) {
  const double pi = PI_F;
  double onm = -1.0;
  int l1 = il1, l2 = il2, l3 = il3, m1 = abs(im1), m2 = abs(im2), m3 = abs(im3);
  int m = abs(im1 + im2 + im3) + (l1 + l2 + l3) % 2;
  if (m != 0 || l1 > l2 + l3 || l1 < abs(l2 - l3)) return 0.0;
  int l = l1;
  m = m1;
  m1 = m1 > m2 ? m1 : m2;
  m1 = m1 > m3 ? m1 : m3;
  if (m1 != m) {
    if (m1 == m2) {
      l1 = l2;
      l2 = l;
      m2 = m;
    } else if (m1 == m3) {
      l1 = l3;
      l3 = l;
      m3 = m;
    }
  }
  if (l2 < l3) {
    l = l2;
    m = m2;
    l2 = l3;
    m2 = m3;
    l3 = l;
    m3 = m;
  }
  int is = (l1 + l2 + l3) / 2;
  int e = is - l2 - m3 + m1;
  double sgn = (e % 2 == 0) ? 1.0 : -1.0;
  (void)onm;
  int a1 = l2 + l3 - m1, a2 = l2 - l3 + m1;
  double fl = facnm(a1 > a2 ? a1 : a2, abs(2 * (l3 - m1)));
  double v = sgn * sqrt((double)((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1)) / pi) / (double)(2 * (2 * is + 1)) *
             sqrt(facnm(l1 + m1, m1 + m1)) / facnm(is - l1, is - l1) * sqrt(facnm(l2 + m2, m2 + m2)) /
             facnm(is - l2, is - l2) * sqrt(facnm(l3 + m3, m3 + m3)) * ((l3 - m1) >= 0 ? fl : 1.0 / fl);
  if (l3 == 0) return v;
  for (int i = 1; i <= l3; ++i) v = v / (double)(2 * (2 * (is - i) + 1));
  double a = 1.0, aa = 1.0;
  int i1 = l1 + m1, i2 = (l2 + l3 - m1) > 1 ? (l2 + l3 - m1) : 1, i3 = 0, i4 = l1 - m1, i5 = l2 - l3 + m1,
      i6 = l3 - m3;
  int it = i4 < i6 ? i4 : i6;
  if (it == 0) return v;
  for (int i = 1; i <= it; ++i) {
    i1++;
    i3++;
    i5++;
    aa = -aa * (double)(i1 * i4 * i6) / (double)(i2 * i3 * i5);
    a += aa;
    i2--;
    i4--;
    i6--;
  }
  return a * v;
}

/* gau(o), o over (l1,m1,l2,m2,l,m), l <= 2 lmax (gauntcalc.f) */
static void gauntcalc(int lmax, double *gau) {
  const double pi = PI_F;
  int o = 0;
  for (int l1 = 0; l1 <= lmax; ++l1)
    for (int m1 = -l1; m1 <= l1; ++m1)
      for (int l2 = 0; l2 <= lmax; ++l2)
        for (int m2 = -l2; m2 <= l2; ++m2)
          for (int l = 0; l <= 2 * lmax; ++l)
            for (int m = -l; m <= l; ++m) {
              int e = (-l2 - l + l1 + 2 * m1) / 2;
              double s = (e % 2 == 0) ? 1.0 : -1.0;
              gau[o++] = blm(l2, m2, l1, -m1, l, m) * 4.0 * pi * s;
            }
}

/* --------------------------------------------------------------- lattice */

typedef struct {
  double a1[2], a2[2], b1[2], b2[2], vol, volbz, distr, distk;
  int maxr, maxk, mdir;
  double *rvec, *veck; /* 2 x maxr / maxk */
  int *indk, *nshk, maxshk;
} orc_lattice;

static int latgen2d(int n1, int n2, int m, double *r, int *mmax, double dist, const double *a1,
                    const double *a2) {
  int nr = 0;
  double dist2 = dist * dist;
  for (int i = -n1; i <= n1; ++i)
    for (int j = -n2; j <= n2; ++j) {
      double x = a1[0] * i + a2[0] * j, y = a1[1] * i + a2[1] * j;
      double r2 = x * x + y * y;
      if (r2 <= dist2) {
        nr++;
        if (nr > m) return 1;
        r[2 * (nr - 1)] = x;
        r[2 * (nr - 1) + 1] = y;
      }
    }
  *mmax = nr;
  return 0;
}

/* heapsort index: a[ind[i+1]] >= a[ind[i]] (indekx.f, 1-based like it) */
static void indekx(int n, const double *a, int *ind) {
  for (int i = 1; i <= n; ++i) ind[i - 1] = i;
  if (n < 2) return;
  int l = n / 2 + 1, ir = n, indxt, i, j;
  double q;
  for (;;) {
    if (l > 1) {
      l = l - 1;
      indxt = ind[l - 1];
      q = a[indxt - 1];
    } else {
      indxt = ind[ir - 1];
      q = a[indxt - 1];
      ind[ir - 1] = ind[0];
      ir = ir - 1;
      if (ir == 1) {
        ind[0] = indxt;
        return;
      }
    }
    i = l;
    j = l + l;
    while (j <= ir) {
      if (j < ir && a[ind[j - 1] - 1] < a[ind[j] - 1]) j = j + 1;
      if (q < a[ind[j - 1] - 1]) {
        ind[i - 1] = ind[j - 1];
        i = j;
        j = j + j;
      } else {
        j = ir + 1;
      }
    }
    ind[i - 1] = indxt;
  }
}

/* index vectors r + rho by length within dist, shells of equal length
 * (vecsrt2d.f); ind is 0-based on return, nsh[k] counts shell k. */
static int vecsrt(int maxx, const double *r, const double *rho, double dist, int *ind, int *nsh) {
  const double tol = 1.0e-6;
  double dist2 = dist * dist;
  double *rmod = (double *)malloc(sizeof(double) * (size_t)maxx);
  int *ix = (int *)malloc(sizeof(int) * (size_t)maxx);
  for (int i = 0; i < maxx; ++i) {
    double x = r[2 * i] + rho[0], y = r[2 * i + 1] + rho[1];
    rmod[i] = x * x + y * y;
    ind[i] = 0;
    nsh[i] = 0;
  }
  indekx(maxx, rmod, ix);
  double r1 = -1.0;
  int k = 0;
  for (int i = 1; i <= maxx; ++i) {
    double r2 = rmod[ix[i - 1] - 1];
    if (r2 > dist2) break;
    if (i == maxx) {
      if (fabs(r2 - r1) <= tol) k = k - 1;
      break;
    }
    ind[i - 1] = ix[i - 1] - 1;
    if (fabs(r2 - r1) > tol) k = k + 1;
    nsh[k - 1]++;
    r1 = r2;
  }
  free(rmod);
  free(ix);
  return k;
}

orc_lattice *orc_lattice_new(const double *a1, const double *a2, int mdir) {
  const double pi = 3.1415926535897932, twopi = pi + pi;
  orc_lattice *L = (orc_lattice *)calloc(1, sizeof(orc_lattice));
  L->mdir = mdir;
  memcpy(L->a1, a1, sizeof L->a1);
  memcpy(L->a2, a2, sizeof L->a2);
  /* reciprocal basis (primkr2d.f:236-246) */
  double v1[2] = {a2[1], -a2[0]}, v2[2] = {-a1[1], a1[0]};
  double dnt1 = a1[0] * v1[0] + a1[1] * v1[1], dnt2 = a2[0] * v2[0] + a2[1] * v2[1];
  for (int i = 0; i < 2; ++i) {
    L->b1[i] = 2.0 * PI_F * v1[i] / dnt1;
    L->b2[i] = 2.0 * PI_F * v2[i] / dnt2;
  }
  const double *b1 = L->b1, *b2 = L->b2;
  L->vol = fabs(a1[0] * a2[1] - a1[1] * a2[0]);
  L->volbz = fabs(b1[0] * b2[1] - b1[1] * b2[0]);
  double dr1 = 0, dr2 = 0, dk1 = 0, dk2 = 0;
  for (int i = 0; i < 2; ++i) {
    dr1 += (a1[i] + a2[i]) * (a1[i] + a2[i]);
    dr2 += (a1[i] - a2[i]) * (a1[i] - a2[i]);
    dk1 += (b1[i] + b2[i]) * (b1[i] + b2[i]);
    dk2 += (b1[i] - b2[i]) * (b1[i] - b2[i]);
  }
  double diagr = sqrt(fmax(dr1, dr2)), diagk = sqrt(fmax(dk1, dk2));
  double diagrm = sqrt(fmin(dr1, dr2)), diagkm = sqrt(fmin(dk1, dk2));
  double distr1 = twopi / sqrt(b1[0] * b1[0] + b1[1] * b1[1]);
  double distr2 = twopi / sqrt(b2[0] * b2[0] + b2[1] * b2[1]);
  double distk1 = twopi / sqrt(a1[0] * a1[0] + a1[1] * a1[1]);
  double distk2 = twopi / sqrt(a2[0] * a2[0] + a2[1] * a2[1]);
  L->rvec = (double *)malloc(sizeof(double) * 2 * (size_t)mdir);
  L->veck = (double *)malloc(sizeof(double) * 2 * (size_t)mdir);
  int md;
  for (md = 1; md <= mdir; ++md) {
    double distr = md * diagrm, totr = distr + diagr;
    int nr1 = (int)(totr / distr1) + 1, nr2 = (int)(totr / distr2) + 1;
    if (latgen2d(nr1, nr2, mdir, L->rvec, &L->maxr, totr, a1, a2)) break;
  }
  L->distr = (md - 1) * diagrm;
  {
    double totr = L->distr + diagr;
    int nr1 = (int)(totr / distr1) + 1, nr2 = (int)(totr / distr2) + 1;
    latgen2d(nr1, nr2, mdir, L->rvec, &L->maxr, totr, a1, a2);
  }
  for (md = 1; md <= mdir; ++md) {
    double distk = md * diagkm, totk = distk + diagk;
    int nk1 = (int)(totk / distk1) + 1, nk2 = (int)(totk / distk2) + 1;
    if (latgen2d(nk1, nk2, mdir, L->veck, &L->maxk, totk, b1, b2)) break;
  }
  L->distk = (md - 1) * diagkm;
  {
    double totk = L->distk + diagk;
    int nk1 = (int)(totk / distk1) + 1, nk2 = (int)(totk / distk2) + 1;
    latgen2d(nk1, nk2, mdir, L->veck, &L->maxk, totk, b1, b2);
  }
  L->indk = (int *)malloc(sizeof(int) * (size_t)L->maxk);
  L->nshk = (int *)malloc(sizeof(int) * (size_t)L->maxk);
  double ak[2] = {0.0, 0.0};
  L->maxshk = vecsrt(L->maxk, L->veck, ak, L->distk, L->indk, L->nshk);
  return L;
}

void orc_lattice_info(const orc_lattice *L, double *d, int *n) {
  d[0] = L->distr;
  d[1] = L->distk;
  d[2] = L->vol;
  d[3] = L->volbz;
  d[4] = L->b1[0];
  d[5] = L->b1[1];
  d[6] = L->b2[0];
  d[7] = L->b2[1];
  n[0] = L->maxr;
  n[1] = L->maxk;
  n[2] = L->maxshk;
}

/* vectors of the reciprocal lattice (in sorted order, 0-based) and real
 * lattice, for checks and for building the same lists elsewhere */
void orc_lattice_vectors(const orc_lattice *L, double *rvec, double *veck) {
  memcpy(rvec, L->rvec, sizeof(double) * 2 * (size_t)L->maxr);
  memcpy(veck, L->veck, sizeof(double) * 2 * (size_t)L->maxk);
}

void orc_lattice_free(orc_lattice *L) {
  if (!L) return;
  free(L->rvec);
  free(L->veck);
  free(L->indk);
  free(L->nshk);
  free(L);
}

/* --------------------------------------------------------------- kamstr */

/* Ewald (Kambe) structure constant of one layer pair, sep = (s, sx, sy, sz);
 * alm is lmmax x lmmax column-major with lmmax = (lmax+1)^2.  Returns 0, or
 * 1 / 2 when the G / R shell sum did not converge (the reference stops).
 * nsh_out[0..1] = shells used in G and R. */
int orc_kamstr(const orc_lattice *L, int lmax, cplx e, double kx, double ky, const double *ts,
               double factor, const double *gau, cplx *alm, int *nsh_out) {
  const double pi = PI_F, ctol = 1.0e-4, tin = 1.0e-10;
  const int lmax2 = 2 * lmax, lmmax = (lmax + 1) * (lmax + 1), lmmax2 = (lmax2 + 1) * (lmax2 + 1);
  const double v2 = L->vol;
  cplx kappa = csqrt(e);
  cplx cs1 = -1.0 / v2 / kappa;
  cplx cs2 = -kappa / 2.0 / sqrt(pi);
  cplx alpha = factor * e * v2 / 4.0 / pi;
  cplx w = alpha * 2.0 / e;
  double fac[2 * 2 * LMX + 1];
  fac[0] = 1.0;
  for (int i = 1; i <= lmax2 + lmax2; ++i) fac[i] = fac[i - 1] * (double)i;
  double flm[(2 * LMX + 1) * (2 * LMX + 1)];
  {
    double d = 2.0;
    for (int l = 0; l <= lmax2; ++l) {
      d = d / 2.0;
      double dl = 2.0 * l + 1.0;
      for (int m = 0; m <= l; ++m) {
        double a = d * sqrt(dl * fac[l + m] * fac[l - m]);
        flm[l * (l + 1) + m] = a;
        flm[l * (l + 1) - m] = a;
      }
    }
  }
  for (int i = 0; i < lmmax * lmmax; ++i) alm[i] = 0.0;
  cplx a = csqrt(alpha);
  cplx c = I * (cexp(-alpha) - orc_cerf(a)) - 1.0 / sqrt(pi) / a;
  cplx d003 = -kappa * (cexp(alpha) * c - I) / 2.0 / sqrt(pi);
  double sx = ts[1], sy = ts[2], sz = ts[3], s = ts[0];
  cplx dlm1[(2 * LMX + 1) * (2 * LMX + 1)], dlm2[(2 * LMX + 1) * (2 * LMX + 1)];
  for (int lm = 0; lm < lmmax2; ++lm) dlm1[lm] = dlm2[lm] = 0.0;
  if (s < ctol) dlm1[0] += d003;
  if (s > ctol && fabs(sz) < ctol) {
    sz = 1.1 * ctol;
    s = sqrt(sx * sx + sy * sy + sz * sz);
  }
  int f35 = (s > ctol && fabs(sz) > ctol), f50 = (s < ctol);
  cplx cs[2 * LMX + 1], cp[2 * LMX + 1], cm[2 * LMX + 1], del[2 * LMX + 2], gam[2 * LMX + 2];
  c = kappa * sz;
  a = 1.0;
  cs[0] = a;
  for (int l = 1; l <= lmax2; ++l) {
    a = a * c;
    cs[l] = a;
  }
  /* reciprocal-space (G) sum, shell by shell */
  int no = 0, conv = 0, is;
  double tst = 1.0e30;
  for (is = 1; is <= L->maxshk; ++is) {
    for (int iis = 0; iis < L->nshk[is - 1]; ++iis) {
      int ig = L->indk[no++];
      double gx = L->veck[2 * ig], gy = L->veck[2 * ig + 1];
      double gkx = gx + kx, gky = gy + ky;
      double km = gkx * gkx + gky * gky;
      cplx kp = csqrt(e - km);
      km = sqrt(km);
      cplx ex2 = 1.0;
      if (km > 1.0e-10) ex2 = CMPLX(gkx, gky) / km;
      cplx xc = cexp(-I * pi) * (e - km * km) * w / 2.0;
      orc_incgamma(lmax2, xc, gam);
      c = kp / kappa;
      c = c * c;
      a = cipow(kp / kappa, -3);
      for (int l = 0; l <= lmax2; ++l) {
        a = a * c;
        cp[l] = a;
      }
      c = km / kappa;
      for (int l = 0; l <= lmax2; ++l) cm[l] = cipow(c, l);
      if (f35) {
        cplx zc = kp * sz;
        zc = zc * zc;
        orc_delkam(lmax2, xc, zc, del);
        cplx ex1 = cexp(I * (gkx * sx + gky * sy)) * cs1;
        for (int l = 0; l <= lmax2; ++l) {
          cplx ex = cipow(ex2, l + 1) * ex1;
          for (int m = -l; m <= l; ++m) {
            ex = ex / ex2;
            int lm = l * (l + 1) + m, mm = abs(m);
            cplx csum = 0.0;
            for (int n = 0; n <= l - mm; ++n) {
              cplx an = cp[n] * del[n], b = 0.0;
              int ipar = (l - mm) % 2;
              int kend = (n + n < l - mm) ? n + n : l - mm;
              for (int k = n + (n + ipar) % 2; k <= kend; k += 2) {
                int i = n + n - k, j = l - k;
                cplx cc = cs[i] * cm[j] / fac[i] / fac[k - n];
                i = (l - mm - k) / 2;
                j = (l + mm - k) / 2;
                b = b + cc / fac[i] / fac[j];
              }
              csum = csum + an * b;
            }
            dlm1[lm] += ex * flm[lm] * csum * ipow_i(1 - m);
          }
        }
      } else if (f50) {
        for (int l = 0; l <= lmax2; ++l) {
          cplx ex = cipow(ex2, l + 1) * cs1;
          for (int m = -l; m <= l; ++m) {
            ex = ex / ex2;
            int lm = l * (l + 1) + m, mm = abs(m);
            cplx csum = 0.0;
            if ((l - mm) % 2 == 0) {
              for (int n = 0; n <= (l - mm) / 2; ++n) {
                int i = l - 2 * n;
                cplx an = cp[n] * cm[i] / fac[n] * gam[n];
                i = (l - mm - 2 * n) / 2;
                int j = (l + mm - 2 * n) / 2;
                csum = csum + an / fac[i] / fac[j];
              }
            }
            dlm1[lm] += ex * csum * flm[lm] * ipow_i(1 - m);
          }
        }
      }
    }
    double d = 0.0;
    for (int lm = 0; lm < lmmax2; ++lm) d += cabs(dlm1[lm]);
    if (fabs(tst - d) < tin && is >= 3) {
      conv = 1;
      break;
    }
    tst = d;
  }
  if (nsh_out) nsh_out[0] = is;
  if (!conv) return 1;
  /* real-space (R) sum around rho = (sx, sy), shell by shell */
  int *indr = (int *)malloc(sizeof(int) * (size_t)L->maxr);
  int *nshr = (int *)malloc(sizeof(int) * (size_t)L->maxr);
  double rho[2] = {ts[1], ts[2]};
  int maxshr = vecsrt(L->maxr, L->rvec, rho, L->distr, indr, nshr);
  cplx ylm[(2 * LMX + 1) * (2 * LMX + 1)], intil[2 * LMX + 2];
  no = 0;
  conv = 0;
  tst = 1.0e30;
  for (is = 1; is <= maxshr; ++is) {
    for (int iis = 0; iis < nshr[is - 1]; ++iis) {
      int ir = indr[no++];
      double rx = L->rvec[2 * ir], ry = L->rvec[2 * ir + 1];
      double rk = rx * kx + ry * ky;
      cplx ex1 = cexp(-I * rk);
      double rsx = rx + sx, rsy = ry + sy, rsz = sz;
      double rs = sqrt(rsx * rsx + rsy * rsy + rsz * rsz);
      if (rs < 1.0e-5) continue;
      ylmc(rsx, rsy, rsz, lmax2, ylm);
      orc_ilkam(lmax2, alpha, rs, kappa, intil);
      a = -kappa * rs / 2.0;
      cplx b = ex1 * cs2 / a;
      for (int l = 0; l <= lmax2; ++l) {
        b = b * a;
        c = intil[l] * b;
        for (int m = -l; m <= l; ++m) {
          int lm = l * (l + 1) + m;
          dlm2[lm] += c * conj(ylm[lm]);
        }
      }
    }
    double d = 0.0;
    for (int lm = 0; lm < lmmax2; ++lm) d += cabs(dlm2[lm]);
    if (fabs(tst - d) < tin && is >= 3) {
      conv = 1;
      break;
    }
    tst = d;
  }
  if (nsh_out) nsh_out[1] = is;
  free(indr);
  free(nshr);
  if (!conv) return 2;
  int o = 0;
  for (int l1 = 0; l1 <= lmax; ++l1)
    for (int m1 = -l1; m1 <= l1; ++m1) {
      int lm1 = l1 * (l1 + 1) + m1;
      for (int l2 = 0; l2 <= lmax; ++l2)
        for (int m2 = -l2; m2 <= l2; ++m2) {
          int lm2 = l2 * (l2 + 1) + m2;
          cplx csum = 0.0;
          for (int l = 0; l <= lmax2; ++l)
            for (int m = -l; m <= l; ++m) {
              int lm = l * (l + 1) + m;
              csum = csum + gau[o++] * (dlm1[lm] + dlm2[lm]);
            }
          alm[lm1 + lm2 * lmmax] = csum;
        }
    }
  return 0;
}

/* -------------------------------------------------------------- directg */

/* reciprocal-space sum for layers with sz > 0 (directg.f); gp and gm are
 * lmmax x lmmax column-major.  Returns 0, or 1 if not converged (the
 * reference stops).  *nsh_out = shells used. */
int orc_directg(const orc_lattice *L, int lmax, double akx, double aky, const double *sep, cplx ecomp,
                cplx *gp, cplx *gm, int *nsh_out) {
  const double dthr = 1.0e-12, pi = 4.0 * atan(1.0);
  const int lmx = lmax + 1, lmx2 = lmx * lmx;
  cplx kappa2 = ecomp, kappa = csqrt(kappa2);
  double f1 = 8.0 * pi * pi / L->vol;
  for (int i = 0; i < lmx2 * lmx2; ++i) gp[i] = gm[i] = 0.0;
  cplx ylm[(LMX + 1) * (LMX + 1)], y1[(LMX + 1) * (LMX + 1)], y2[(LMX + 1) * (LMX + 1)],
      y3[(LMX + 1) * (LMX + 1)], y4[(LMX + 1) * (LMX + 1)];
  int num = 0, i1;
  for (i1 = 1; i1 <= L->maxshk; ++i1) {
    double term = 0.0, sum = 0.0;
    for (int j1 = 0; j1 < L->nshk[i1 - 1]; ++j1) {
      int ig = L->indk[num++];
      double gkx = akx + L->veck[2 * ig], gky = aky + L->veck[2 * ig + 1];
      double gk2 = gkx * gkx + gky * gky, gk = sqrt(gk2);
      cplx gkz2 = kappa2 - gk2, gkz = csqrt(gkz2);
      cplx phi = 1.0;
      if (gk > SNG(1.0e-6)) phi = CMPLX(gkx, gky) / gk;
      cplx cth = gkz / kappa, sth = gk / kappa;
      sphrm(ylm, cth, sth, phi, lmax);
      int lm = 0;
      cplx il = -I;
      double sl = -1.0, sm = -1.0;
      for (int l = 0; l <= lmax; ++l) {
        sl = -sl;
        il = il * I;
        for (int m = -l; m <= l; ++m) {
          sm = -sm;
          lm = lm + 1;
          int lmmm = lm - m - m;
          y1[lm - 1] = sl * il * ylm[lm - 1];
          y2[lm - 1] = il * sm * ylm[lmmm - 1];
          y3[lm - 1] = il * sm * ylm[lm - 1];
          y4[lm - 1] = sl * il * ylm[lmmm - 1];
        }
      }
      sum = 0.0;
      cplx fp = cexp(I * (gkx * sep[0] + gky * sep[1] + gkz * sep[2]));
      cplx fm = cexp(I * (-gkx * sep[0] - gky * sep[1] + gkz * sep[2]));
      fp = fp * f1 / gkz;
      fm = fm * f1 / gkz;
      const cplx fac = -I;
      for (int lmi = 0; lmi < lmx2; ++lmi)
        for (int lmj = 0; lmj < lmx2; ++lmj) {
          cplx ctp = y1[lmi] * fp * y2[lmj], ctm = y3[lmi] * fm * y4[lmj];
          gp[lmj + lmi * lmx2] += ctp * fac;
          gm[lmj + lmi * lmx2] += ctm * fac;
          term += cabs(ctp) + cabs(ctm);
          sum += cabs(gp[lmi + lmj * lmx2]) + cabs(gm[lmi + lmj * lmx2]);
        }
    }
    if (term / sum < dthr && i1 > 1) break;
  }
  if (nsh_out) *nsh_out = i1;
  return (i1 > L->maxshk) ? 1 : 0;
}

/* --------------------------------------------------------------- gstore */

/* g2d of one k for the bulk-like geometry (gstore.f with bulkgeo, side 'L'
 * or 'R'): one PL of s layers at cvec[3*(i)] (i = 0..s, the last being the
 * first layer of the next PL), nextra extra PLs each side for screening,
 * alpha[l + (lmax+1)*i] the screening parameter of layer i.  g2d is
 * (nl2, nl2, 2s+1, s, 3) column-major: g2d(m, n, jj+s, icount, icpl).
 * Returns 0 or the first nonzero status of kamstr / directg / inversion
 * (10 + status for the inversion). */
int orc_gstore(const orc_lattice *L, int lmax, int s, int nextra, const double *cvec_pl, const cplx *alpha,
               double kx, double ky, cplx ce, double eta, cplx *g2d) {
  const int nl2 = (lmax + 1) * (lmax + 1), ntotal = (2 * nextra + 3) * s, num = 3 * s, N = nl2 * ntotal;
  double rlmin = fmin(L->b1[0] * L->b1[0] + L->b1[1] * L->b1[1], L->b2[0] * L->b2[0] + L->b2[1] * L->b2[1]);
  rlmin = sqrt(rlmin);
  int lg = (2 * lmax + 1) * (2 * lmax + 1);
  double *gau = (double *)malloc(sizeof(double) * (size_t)(nl2 * nl2 * lg));
  gauntcalc(lmax, gau);
  double *cv = (double *)malloc(sizeof(double) * 3 * (size_t)ntotal);
  double cb[3];
  for (int d = 0; d < 3; ++d) cb[d] = cvec_pl[3 * s + d] - cvec_pl[d];
  for (int i = 0; i < s; ++i)
    for (int d = 0; d < 3; ++d) cv[3 * i + d] = cvec_pl[3 * i + d] - cvec_pl[d];
  for (int iprc = 2; iprc <= 2 * nextra + 3; ++iprc)
    for (int inprc = 0; inprc < s; ++inprc)
      for (int d = 0; d < 3; ++d) cv[3 * ((iprc - 1) * s + inprc) + d] = cv[3 * inprc + d] + (iprc - 1) * cb[d];
  /* gtotal as the N x N matrix g: block (a, b) at rows a*nl2, cols b*nl2 */
  cplx *g = (cplx *)calloc((size_t)N * N, sizeof(cplx));
  cplx *blk = (cplx *)malloc(sizeof(cplx) * (size_t)(2 * nl2 * nl2));
  int status = 0;
  for (int i = 0; i < ntotal && !status; ++i)
    for (int j = 0; j < ntotal && !status; ++j) {
      double sep[4];
      sep[1] = cv[3 * j] - cv[3 * i];
      sep[2] = cv[3 * j + 1] - cv[3 * i + 1];
      sep[3] = cv[3 * j + 2] - cv[3 * i + 2];
      sep[0] = sqrt(sep[1] * sep[1] + sep[2] * sep[2] + sep[3] * sep[3]);
      if (fabs(sep[3]) * rlmin >= 5.0) {
        if (j > i) {
          int st = orc_directg(L, lmax, kx, ky, sep + 1, ce, blk, blk + nl2 * nl2, NULL);
          if (st) status = 3;
          for (int n = 0; n < nl2; ++n)
            for (int m = 0; m < nl2; ++m) {
              g[(size_t)(j * nl2 + m) + (size_t)(i * nl2 + n) * N] = blk[m + n * nl2];
              g[(size_t)(i * nl2 + m) + (size_t)(j * nl2 + n) * N] = blk[nl2 * nl2 + m + n * nl2];
            }
        }
      } else {
        int st = orc_kamstr(L, lmax, ce, kx, ky, sep, eta, gau, blk, NULL);
        if (st) status = st;
        for (int n = 0; n < nl2; ++n)
          for (int m = 0; m < nl2; ++m) g[(size_t)(j * nl2 + m) + (size_t)(i * nl2 + n) * N] = blk[m + n * nl2];
      }
    }
  if (!status) {
    /* screening: G' = G (1 - alpha G)^-1 (scrstr.f:56-135) */
    cplx *g1 = (cplx *)malloc(sizeof(cplx) * (size_t)N * N);
    for (int r = 0; r < N; ++r) {
      int il = r / nl2, ilm = r % nl2;
      int l = (int)sqrt((double)ilm + 0.5);
      while ((l + 1) * (l + 1) <= ilm) ++l;
      while (l * l > ilm) --l;
      int ill = il % s; /* mod(il, ninprc) with 0 -> ninprc, 0-based here */
      cplx al = alpha[l + (lmax + 1) * ill];
      for (int c2 = 0; c2 < N; ++c2) g1[r + (size_t)c2 * N] = -al * g[r + (size_t)c2 * N];
      g1[r + (size_t)r * N] = 1.0 + g1[r + (size_t)r * N];
    }
    int st = orc_gjinv(g1, N, N);
    if (st) status = 10 + st;
    /* gs = g g1 row by row, k summed in order (scrstr.f:106-133) */
    cplx *gs = (cplx *)malloc(sizeof(cplx) * (size_t)N * N);
    for (int r = 0; r < N && !status; ++r)
      for (int c2 = 0; c2 < N; ++c2) {
        cplx sum = 0.0;
        for (int k = 0; k < N; ++k) sum = sum + g[r + (size_t)k * N] * g1[k + (size_t)c2 * N];
        gs[r + (size_t)c2 * N] = sum;
      }
    /* extract the layer-tridiagonal blocks (gstore.f:205-280) */
    const int imin = nextra * s;
    for (int il = 1; il <= num && !status; ++il) {
      int i = imin + il; /* 1-based layer */
      int icpl = (il <= s) ? 0 : (il > num - s) ? 2 : 1;
      int icount = (il - 1) % s; /* 0-based */
      for (int jj = -s; jj <= s; ++jj) {
        int j = i + jj;
        for (int n = 0; n < nl2; ++n)
          for (int m = 0; m < nl2; ++m)
            g2d[m + (size_t)nl2 * (n + (size_t)nl2 * ((jj + s) + (size_t)(2 * s + 1) * (icount + (size_t)s * icpl)))] =
                gs[(size_t)((i - 1) * nl2 + m) + (size_t)((j - 1) * nl2 + n) * N];
      }
    }
    free(gs);
    free(g1);
  }
  free(g);
  free(blk);
  free(cv);
  free(gau);
  return status;
}

/* the Gaunt table of gauntcalc.f for checks */
void orc_gaunt(int lmax, double *gau) { gauntcalc(lmax, gau); }

void orc_ylmc(double x, double y, double z, int lmax, cplx *ylm) { ylmc(x, y, z, lmax, ylm); }

/* ----------------------------------------- pointer entry points (ctypes) */

void orc_incgamma_p(int n, const cplx *z, cplx *gam) { orc_incgamma(n, *z, gam); }
void orc_delkam_p(int n, const cplx *x, const cplx *zz, cplx *d) { orc_delkam(n, *x, *zz, d); }
void orc_ilkam_p(int l, const cplx *alpha, double t, const cplx *kappa, cplx *d) {
  orc_ilkam(l, *alpha, t, *kappa, d);
}
int orc_kamstr_p(const orc_lattice *L, int lmax, const cplx *e, double kx, double ky, const double *ts,
                 double factor, cplx *alm, int *nsh) {
  int lg = (2 * lmax + 1) * (2 * lmax + 1), nl2 = (lmax + 1) * (lmax + 1);
  double *gau = (double *)malloc(sizeof(double) * (size_t)(nl2 * nl2 * lg));
  gauntcalc(lmax, gau);
  int st = orc_kamstr(L, lmax, *e, kx, ky, ts, factor, gau, alm, nsh);
  free(gau);
  return st;
}
int orc_directg_p(const orc_lattice *L, int lmax, double kx, double ky, const double *sep, const cplx *e,
                  cplx *gp, cplx *gm, int *nsh) {
  return orc_directg(L, lmax, kx, ky, sep, *e, gp, gm, nsh);
}
int orc_gstore_p(const orc_lattice *L, int lmax, int s, int nextra, const double *cvec_pl, const cplx *alpha,
                 double kx, double ky, const cplx *e, double eta, cplx *g2d) {
  return orc_gstore(L, lmax, s, nextra, cvec_pl, alpha, kx, ky, *e, eta, g2d);
}
