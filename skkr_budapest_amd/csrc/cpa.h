// Coherent-potential update on the device (skkr_kpar_v6/nrsp-scf/fullcpa.f):
// elementwise and per-site reduction kernels around the batched inversions
// and products of the CPA step.  Matrices are lk x lk column-major, one per
// interface site, site after site.
#pragma once
#include <hip/hip_runtime.h>

namespace cpa {

// compmat (matops.f:51-80) of ma and mb against mc, the site flag of cpacor
// (fullcpa.f:44-51: update unless concentration 1 or mc equals ma or mb),
// and da = mc - ma, db = mc - mb with copies xa = da, xb = db.
__global__ __launch_bounds__(256) void cpa_prep_kernel(int lk, const double* __restrict__ conc,
                                                       const double2* __restrict__ ma,
                                                       const double2* __restrict__ mb,
                                                       const double2* __restrict__ mc, int* __restrict__ flag,
                                                       double2* da, double2* db, double2* xa, double2* xb) {
  __shared__ double red[4][256];
  const int site = blockIdx.x, t = threadIdx.x, nn = lk * lk;
  const size_t o = (size_t)site * nn;
  double xa2 = 0.0, dxa2 = 0.0, xb2 = 0.0, dxb2 = 0.0;
  for (int e = t; e < nn; e += 256) {
    const double2 a = ma[o + e], b = mb[o + e], c = mc[o + e];
    const double2 dA = make_double2(a.x - c.x, a.y - c.y), dB = make_double2(b.x - c.x, b.y - c.y);
    xa2 += a.x * a.x + a.y * a.y;
    dxa2 += dA.x * dA.x + dA.y * dA.y;
    xb2 += b.x * b.x + b.y * b.y;
    dxb2 += dB.x * dB.x + dB.y * dB.y;
    const double2 u = make_double2(c.x - a.x, c.y - a.y), v = make_double2(c.x - b.x, c.y - b.y);
    da[o + e] = u;
    xa[o + e] = u;
    db[o + e] = v;
    xb[o + e] = v;
  }
  red[0][t] = xa2;
  red[1][t] = dxa2;
  red[2][t] = xb2;
  red[3][t] = dxb2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h)
      for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  if (t == 0) {
    const double tol = 1.0e-10;
    const bool ica = sqrt(red[1][0]) < tol * sqrt(red[0][0]);
    const bool icb = sqrt(red[3][0]) < tol * sqrt(red[2][0]);
    const bool cpalay = 1.0 - conc[site] > 1.0e-6;
    flag[site] = (cpalay && !ica && !icb) ? 1 : 0;
  }
}

// x -= tauc for the two species halves (x = [xa sites | xb sites])
__global__ __launch_bounds__(256) void cpa_sub_tau_kernel(long long nsn, const double2* __restrict__ tauc,
                                                          double2* xa, double2* xb) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= nsn) return;
  const double2 t = tauc[e];
  xa[e] = make_double2(xa[e].x - t.x, xa[e].y - t.y);
  xb[e] = make_double2(xb[e].x - t.x, xb[e].y - t.y);
}

// h = c xa + (1 - c) xb per site
__global__ __launch_bounds__(256) void cpa_mix_kernel(int nn, long long nsn, const double* __restrict__ conc,
                                                      const double2* __restrict__ xa,
                                                      const double2* __restrict__ xb, double2* h) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= nsn) return;
  const double ca = conc[e / nn], cb = 1.0 - ca;
  h[e] = make_double2(ca * xa[e].x + cb * xb[e].x, ca * xa[e].y + cb * xb[e].y);
}

// h += tauc
__global__ __launch_bounds__(256) void cpa_add_tau_kernel(long long nsn, const double2* __restrict__ tauc,
                                                          double2* h) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= nsn) return;
  h[e] = make_double2(h[e].x + tauc[e].x, h[e].y + tauc[e].y);
}

// flagged sites: mc -= h and error3 = |h|_F / |mc|_F (fullcpa.f:114-137);
// every site: taua = taub = tauc (the products add tauc xa tauc on
// flagged sites; unflagged keep tauc, fullcpa.f:46-49)
__global__ __launch_bounds__(256) void cpa_update_kernel(int lk, const int* __restrict__ flag,
                                                         const double2* __restrict__ h,
                                                         const double2* __restrict__ tauc, double2* mc,
                                                         double2* taua, double2* taub, double* err3) {
  __shared__ double red[2][256];
  const int site = blockIdx.x, t = threadIdx.x, nn = lk * lk;
  const size_t o = (size_t)site * nn;
  const bool upd = flag[site] != 0;
  double d1 = 0.0, d2 = 0.0;
  for (int e = t; e < nn; e += 256) {
    const double2 tc = tauc[o + e];
    taua[o + e] = tc;
    taub[o + e] = tc;
    if (upd) {
      const double2 hv = h[o + e];
      double2 m = mc[o + e];
      m = make_double2(m.x - hv.x, m.y - hv.y);
      mc[o + e] = m;
      d1 += hv.x * hv.x + hv.y * hv.y;
      d2 += m.x * m.x + m.y * m.y;
    }
  }
  red[0][t] = d1;
  red[1][t] = d2;
  __syncthreads();
  for (int hh = 128; hh > 0; hh >>= 1) {
    if (t < hh) {
      red[0][t] += red[0][t + hh];
      red[1][t] += red[1][t + hh];
    }
    __syncthreads();
  }
  if (t == 0) err3[site] = upd ? sqrt(red[0][0] / red[1][0]) : 0.0;
}

// t = c mb + (1 - c) ma per site (tata, fullcpa.f:15-20)
__global__ __launch_bounds__(256) void cpa_ata_kernel(int nn, long long nsn, const double* __restrict__ conc,
                                                      const double2* __restrict__ ma,
                                                      const double2* __restrict__ mb, double2* t) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= nsn) return;
  const double ca = conc[e / nn], cb = 1.0 - ca;
  t[e] = make_double2(ca * mb[e].x + cb * ma[e].x, ca * mb[e].y + cb * ma[e].y);
}

}  // namespace cpa
