// Host-side set-up of the layer structure constants: the 2D lattices of
// vector2d.f:2-150 / latgen2d.f:2-33 (skkr_kpar_v6/main), shells of vectors
// sorted by length around a point (vecsrt2d.f:2-95 with the heapsort of
// indekx.f:2-47, whose order within a shell sets the summation order), the
// Gaunt table of gauntcalc.f:2-28 / blm.f:2-59, Kambe's F(lm) of
// kamstr.f:79-91 and the spherical-harmonic coefficients of ylmc.f:26-89.
#pragma once
#include <cmath>
#include <complex>
#include <cstdlib>
#include <vector>

namespace kkrh {

constexpr double kPi = 3.1415926535897932384626;

struct Lattice {
  double a1[2], a2[2], b1[2], b2[2], vol = 0, volbz = 0, distr = 0, distk = 0;
  std::vector<double> rvec, veck;  // 2 per vector, generation order
};

// lattice vectors i a1 + j a2 within dist, i over [-n1, n1] outer, j inner;
// false when more than m vectors would be needed
static bool latgen(int n1, int n2, int m, std::vector<double>& r, double dist, const double* a1,
                   const double* a2) {
  r.clear();
  const double d2 = dist * dist;
  for (int i = -n1; i <= n1; ++i)
    for (int j = -n2; j <= n2; ++j) {
      const double x = a1[0] * i + a2[0] * j, y = a1[1] * i + a2[1] * j;
      if (x * x + y * y <= d2) {
        if ((int)(r.size() / 2) + 1 > m) return false;
        r.push_back(x);
        r.push_back(y);
      }
    }
  return true;
}

// direct and reciprocal lattices: the largest radius (a multiple of the
// short cell diagonal) whose vectors fit in mdir, padded by the long one
static Lattice make_lattice(const double* a1, const double* a2, int mdir) {
  Lattice L;
  for (int i = 0; i < 2; ++i) {
    L.a1[i] = a1[i];
    L.a2[i] = a2[i];
  }
  const double v1[2] = {a2[1], -a2[0]}, v2[2] = {-a1[1], a1[0]};
  const double dnt1 = a1[0] * v1[0] + a1[1] * v1[1], dnt2 = a2[0] * v2[0] + a2[1] * v2[1];
  for (int i = 0; i < 2; ++i) {
    L.b1[i] = 2.0 * kPi * v1[i] / dnt1;
    L.b2[i] = 2.0 * kPi * v2[i] / dnt2;
  }
  const double *b1 = L.b1, *b2 = L.b2;
  const double twopi = 3.1415926535897932 * 2.0;
  L.vol = std::fabs(a1[0] * a2[1] - a1[1] * a2[0]);
  L.volbz = std::fabs(b1[0] * b2[1] - b1[1] * b2[0]);
  double dr1 = 0, dr2 = 0, dk1 = 0, dk2 = 0;
  for (int i = 0; i < 2; ++i) {
    dr1 += (a1[i] + a2[i]) * (a1[i] + a2[i]);
    dr2 += (a1[i] - a2[i]) * (a1[i] - a2[i]);
    dk1 += (b1[i] + b2[i]) * (b1[i] + b2[i]);
    dk2 += (b1[i] - b2[i]) * (b1[i] - b2[i]);
  }
  const double diagr = std::sqrt(std::fmax(dr1, dr2)), diagk = std::sqrt(std::fmax(dk1, dk2));
  const double diagrm = std::sqrt(std::fmin(dr1, dr2)), diagkm = std::sqrt(std::fmin(dk1, dk2));
  const double distr1 = twopi / std::sqrt(b1[0] * b1[0] + b1[1] * b1[1]);
  const double distr2 = twopi / std::sqrt(b2[0] * b2[0] + b2[1] * b2[1]);
  const double distk1 = twopi / std::sqrt(a1[0] * a1[0] + a1[1] * a1[1]);
  const double distk2 = twopi / std::sqrt(a2[0] * a2[0] + a2[1] * a2[1]);
  auto grow = [&](double diagm, double diag, double dd1, double dd2, const double* p1, const double* p2,
                  std::vector<double>& out) {
    int md;
    for (md = 1; md <= mdir; ++md) {
      const double tot = md * diagm + diag;
      if (!latgen((int)(tot / dd1) + 1, (int)(tot / dd2) + 1, mdir, out, tot, p1, p2)) break;
    }
    const double dist = (md - 1) * diagm, tot = dist + diag;
    latgen((int)(tot / dd1) + 1, (int)(tot / dd2) + 1, mdir, out, tot, p1, p2);
    return dist;
  };
  L.distr = grow(diagrm, diagr, distr1, distr2, a1, a2, L.rvec);
  L.distk = grow(diagkm, diagk, distk1, distk2, b1, b2, L.veck);
  return L;
}

// heapsort index with a[ind[i+1]] >= a[ind[i]] (0-based)
static void heap_index(const std::vector<double>& a, std::vector<int>& ind) {
  const int n = (int)a.size();
  ind.resize(n);
  for (int i = 0; i < n; ++i) ind[i] = i;
  if (n < 2) return;
  int l = n / 2 + 1, ir = n;
  for (;;) {
    int it;
    double q;
    if (l > 1) {
      --l;
      it = ind[l - 1];
      q = a[it];
    } else {
      it = ind[ir - 1];
      q = a[it];
      ind[ir - 1] = ind[0];
      if (--ir == 1) {
        ind[0] = it;
        return;
      }
    }
    int i = l, j = l + l;
    while (j <= ir) {
      if (j < ir && a[ind[j - 1]] < a[ind[j]]) ++j;
      if (q < a[ind[j - 1]]) {
        ind[i - 1] = ind[j - 1];
        i = j;
        j += j;
      } else {
        j = ir + 1;
      }
    }
    ind[i - 1] = it;
  }
}

// vectors r + rho within dist sorted by length, grouped into shells of equal
// length (|r|^2 within 1e-6 of the previous); at most maxv vectors, cut at a
// shell boundary
static void shells(const std::vector<double>& r, const double* rho, double dist, int maxv,
                   std::vector<double>& sorted, std::vector<int>& nsh) {
  const int n = (int)(r.size() / 2);
  std::vector<double> rm(n);
  for (int i = 0; i < n; ++i) {
    const double x = r[2 * i] + rho[0], y = r[2 * i + 1] + rho[1];
    rm[i] = x * x + y * y;
  }
  std::vector<int> ix;
  heap_index(rm, ix);
  sorted.clear();
  nsh.clear();
  const double d2 = dist * dist;
  double r1 = -1.0;
  for (int i = 0; i < n - 1; ++i) {  // the last vector never starts or ends a kept shell
    const double r2 = rm[ix[i]];
    if (r2 > d2) break;
    if (std::fabs(r2 - r1) > 1.0e-6) {
      if ((int)(sorted.size() / 2) >= maxv) break;
      nsh.push_back(0);
    }
    nsh.back()++;
    sorted.push_back(r[2 * ix[i]]);
    sorted.push_back(r[2 * ix[i] + 1]);
    r1 = r2;
  }
}

// n!/(n-m)!
static double facnm(int n, int m) {
  double f = 1.0;
  for (int i = n - m + 1; i <= n; ++i) f *= (double)i;
  return f;
}

// integral of three spherical harmonics by Gaunt's closed form
static double gaunt3(int il1, int im1, int il2, int im2, int il3, int im3) {
  int l1 = il1, l2 = il2, l3 = il3, m1 = std::abs(im1), m2 = std::abs(im2), m3 = std::abs(im3);
  if (std::abs(im1 + im2 + im3) + (l1 + l2 + l3) % 2 != 0 || l1 > l2 + l3 || l1 < std::abs(l2 - l3))
    return 0.0;
  // the largest |m| first, then l2 >= l3
  const int mx = std::max(m1, std::max(m2, m3));
  if (mx != m1) {
    if (mx == m2) {
      std::swap(l1, l2);
      std::swap(m1, m2);
    } else {
      std::swap(l1, l3);
      std::swap(m1, m3);
    }
  }
  if (l2 < l3) {
    std::swap(l2, l3);
    std::swap(m2, m3);
  }
  const int is = (l1 + l2 + l3) / 2;
  const double sgn = ((is - l2 - m3 + m1) % 2 == 0) ? 1.0 : -1.0;
  const double fl = facnm(std::max(l2 + l3 - m1, l2 - l3 + m1), std::abs(2 * (l3 - m1)));
  double v = sgn * std::sqrt((double)((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1)) / kPi) /
             (double)(2 * (2 * is + 1)) * std::sqrt(facnm(l1 + m1, m1 + m1)) / facnm(is - l1, is - l1) *
             std::sqrt(facnm(l2 + m2, m2 + m2)) / facnm(is - l2, is - l2) * std::sqrt(facnm(l3 + m3, m3 + m3)) *
             ((l3 - m1) >= 0 ? fl : 1.0 / fl);
  if (l3 == 0) return v;
  for (int i = 1; i <= l3; ++i) v /= (double)(2 * (2 * (is - i) + 1));
  double a = 1.0, aa = 1.0;
  int i1 = l1 + m1, i2 = std::max(1, l2 + l3 - m1), i3 = 0, i4 = l1 - m1, i5 = l2 - l3 + m1, i6 = l3 - m3;
  const int it = std::min(i4, i6);
  for (int i = 1; i <= it; ++i) {
    ++i1;
    ++i3;
    ++i5;
    aa = -aa * (double)(i1 * i4 * i6) / (double)(i2 * i3 * i5);
    a += aa;
    --i2;
    --i4;
    --i6;
  }
  return it == 0 ? v : a * v;
}

// gau(lm1, lm2, lm), lm of l <= 2 lmax, as kamstr's contraction reads it
static std::vector<double> gaunt_table(int lmax) {
  std::vector<double> g;
  for (int l1 = 0; l1 <= lmax; ++l1)
    for (int m1 = -l1; m1 <= l1; ++m1)
      for (int l2 = 0; l2 <= lmax; ++l2)
        for (int m2 = -l2; m2 <= l2; ++m2)
          for (int l = 0; l <= 2 * lmax; ++l)
            for (int m = -l; m <= l; ++m) {
              const int e = (-l2 - l + l1 + 2 * m1) / 2;
              g.push_back(gaunt3(l2, m2, l1, -m1, l, m) * 4.0 * kPi * ((e % 2 == 0) ? 1.0 : -1.0));
            }
  return g;
}

// F(lm) = 2^-l sqrt((2l+1) (l+m)! (l-m)!)
static std::vector<double> kambe_flm(int l2max) {
  std::vector<double> fac(2 * l2max + 1), f((l2max + 1) * (l2max + 1));
  fac[0] = 1.0;
  for (int i = 1; i <= 2 * l2max; ++i) fac[i] = fac[i - 1] * (double)i;
  double d = 2.0;
  for (int l = 0; l <= l2max; ++l) {
    d /= 2.0;
    const double dl = 2.0 * l + 1.0;
    for (int m = 0; m <= l; ++m) {
      const double a = d * std::sqrt(dl * fac[l + m] * fac[l - m]);
      f[l * (l + 1) + m] = f[l * (l + 1) - m] = a;
    }
  }
  return f;
}

// associated-Legendre polynomial coefficients b(l, m, k) (cos^k) and the
// prefactors of Y_lm; the l-recursion factors (2l-1)/l and -(l-1)/l are
// single-precision quotients in the reference (ylmc.f:48-49) and are kept so
// b(l >= 3) carries the same values
static void ylm_tables(int l2max, std::vector<double>& b, std::vector<std::complex<double>>& cf) {
  const int n = l2max + 1;
  std::vector<double> a(n * n, 0.0), f(n * n, 0.0);
  auto A = [&](int l, int k) -> double& { return a[l * n + k]; };
  A(0, 0) = 1.0;
  if (n > 1) A(1, 1) = 1.0;
  if (n > 2) {
    A(2, 0) = -0.5;
    A(2, 2) = 1.5;
  }
  for (int l = 3; l < n; ++l) {
    const double fa1 = (double)((float)(2 * l - 1) / (float)l), fa2 = -(double)((float)(l - 1) / (float)l);
    A(l, 0) = fa2 * A(l - 2, 0);
    for (int k = 1; k <= l - 2; ++k) A(l, k) = fa1 * A(l - 1, k - 1) + fa2 * A(l - 2, k);
    A(l, l - 1) = fa1 * A(l - 1, l - 2);
    A(l, l) = fa1 * A(l - 1, l - 1);
  }
  for (int q = 0; q < n; ++q) {
    f[q * n] = 1.0;
    for (int m = 0; m < q; ++m) f[q * n + m + 1] = f[q * n + m] * (double)(q - m);
  }
  b.assign(n * n * n, 0.0);
  for (int l = 0; l < n; ++l)
    for (int m = 0; m <= l; ++m)
      for (int k = 0; k <= l - m; ++k) b[(l * n + m) * n + k] = A(l, k + m) * f[(k + m) * n + m];
  std::vector<double> fact(2 * n + 1);
  fact[0] = 1.0;
  for (int i = 1; i <= 2 * n; ++i) fact[i] = fact[i - 1] * (double)i;
  const double pi4 = 4.0 * std::acos(-1.0);
  cf.clear();
  const std::complex<double> ip[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
  for (int l = 0; l < n; ++l)
    for (int m = -l; m <= l; ++m) {
      const int mb = std::abs(m);
      cf.push_back(std::sqrt((double)(2 * l + 1) * fact[l - mb] / (pi4 * fact[l + mb])) * ip[(m + mb) % 4]);
    }
}

// lm -> kappa-mu rectangular Clebsch-Gordan matrices for l <= 5: u1, u2
// and the lm index (0-based) of each kappa-mu index ir = 2 kap^2 + kap + m
// (ms = -1/2) and ir + 1 (ms = +1/2), j = l -+ 1/2 (clebsch.f:21-67); an
// unset index points at lm 0 with coefficient 0
static void clebsch_tables(std::vector<double>& u1, std::vector<double>& u2, std::vector<int>& i1,
                           std::vector<int>& i2) {
  u1.assign(72, 0.0);
  u2.assign(72, 0.0);
  i1.assign(72, 0);
  i2.assign(72, 0);
  int inr = -1;
  for (int l = 0; l <= 5; ++l) {
    const double t = (double)(2 * l + 1);
    for (int m = -l; m <= l; ++m) {
      ++inr;
      if (l != 0) {
        const int kap = l, ir = 2 * kap * kap + kap + m;
        u1[ir - 1] = std::sqrt((l + m) / t);
        i1[ir - 1] = inr;
        u2[ir] = -std::sqrt((l - m) / t);
        i2[ir] = inr;
      }
      const int kap = -l - 1, ir = 2 * kap * kap + kap + m;
      u1[ir - 1] = std::sqrt((l - m + 1) / t);
      i1[ir - 1] = inr;
      u2[ir] = std::sqrt((l + m + 1) / t);
      i2[ir] = inr;
    }
  }
}

}  // namespace kkrh
