/*
 * CPU restatement of the BIEF prism (P1 6-node) element matrices.
 * TEST INFRASTRUCTURE ONLY (see bief_oracle.c).
 *
 * Storage: t is T(NELMAX,6) (diagonal terms), xm is XM(NELMAX,15) in the
 * prism pair order (12,13,14,15,16,23,24,25,26,34,35,36,45,46,56).
 */
#include <math.h>

#include "bief_oracle.h"

void orc_mt01pp(double* t, double* xm, double xmul, const double* surfac,
                const double* z, const int* ikle, int nelem, int nelmax) {
  /* mt01pp.f (prism mass matrix with a 1e-3 floor on the column heights) */
  const double chouia = 1e-3;
  double sur360 = xmul / 360.0;
#define XM(c) xm[(long)((c)-1) * nelmax + e]
#define T(c) t[(long)((c)-1) * nelmax + e]
#define IK(k) ikle[(long)((k)-1) * nelmax + e]
  for (int e = 0; e < nelem; ++e) {
    double coef = surfac[e] * sur360;
    double h1 = fmax(z[IK(4)] - z[IK(1)], chouia) * coef;
    double h2 = fmax(z[IK(5)] - z[IK(2)], chouia) * coef;
    double h3 = fmax(z[IK(6)] - z[IK(3)], chouia) * coef;
    double ht = h1 + h2 + h3;
    XM(4) = h1 + h2 + ht;
    XM(5) = h1 + h3 + ht;
    XM(9) = h2 + h3 + ht;
    XM(7) = XM(4);
    XM(10) = XM(5);
    XM(11) = XM(9);
    XM(3) = 4 * h1 + ht + ht;
    XM(8) = 4 * h2 + ht + ht;
    XM(12) = 4 * h3 + ht + ht;
    XM(1) = XM(4) + XM(4);
    XM(2) = XM(5) + XM(5);
    XM(6) = XM(9) + XM(9);
    XM(13) = XM(1);
    XM(14) = XM(2);
    XM(15) = XM(6);
    T(1) = XM(3) + XM(3);
    T(2) = XM(8) + XM(8);
    T(3) = XM(12) + XM(12);
    T(4) = T(1);
    T(5) = T(2);
    T(6) = T(3);
  }
#undef XM
#undef T
#undef IK
}

/* One horizontal direction of the prism diffusion matrix (mt02pp.f:
 * 229-356 for x with F, 358-445 for y with G; the same sequence of
 * operations, kept so that results agree to the last bit). d1, d2, d3 are the
 * in-plane derivative factors of the base triangle, n[0..5] the nodal
 * coefficient divided by the column heights, hh the heights h1..h3, zz the
 * node elevations, dd = XMUL/2880/SURFAC. Writes the 15 contributions. */
static void mt02pp_dir(double d1, double d2, double d3, const double* n, const double* hh,
                       const double* zz, double dd, double* c) {
  const double h1 = hh[0], h2 = hh[1], h3 = hh[2];
  const double sh1 = h1 + h2 + h3;
  const double shh = h1 * h1 + h2 * h2 + h3 * h3;
  const double ri = -(zz[0] * d1 + zz[1] * d2 + zz[2] * d3);
  const double rs = -(zz[3] * d1 + zz[4] * d2 + zz[5] * d3);
  const double r = ri + rs;
  const double rri = r + ri + ri, rrs = r + rs + rs;
  const double rrri = r * r + 2 * ri * ri, rrrs = r * r + 2 * rs * rs;
  const double d12 = d1 * d2, d13 = d1 * d3, d23 = d2 * d3;
  const double sni = n[0] + n[1] + n[2], sns = n[3] + n[4] + n[5];
  const double snhi = n[0] * h1 + n[1] * h2 + n[2] * h3;
  const double snhs = n[3] * h1 + n[4] * h2 + n[5] * h3;
  double snhhi = (sni * sh1 + snhi + snhi) * sh1 + sni * shh +
                 2 * (n[0] * h1 * h1 + n[1] * h2 * h2 + n[2] * h3 * h3);
  double snhhs = (sns * sh1 + snhs + snhs) * sh1 + sns * shh +
                 2 * (n[3] * h1 * h1 + n[4] * h2 * h2 + n[5] * h3 * h3);
  const double snhh = snhhi + snhhs;
  snhhi = snhh + snhhi + snhhi;
  snhhs = snhh + snhhs + snhhs;
  const double hhi12 = snhhi * d12, hhi13 = snhhi * d13, hhi23 = snhhi * d23;
  const double hhs12 = snhhs * d12, hhs13 = snhhs * d13, hhs23 = snhhs * d23;
  const double hh12 = snhh * d12, hh13 = snhh * d13, hh23 = snhh * d23;
  const double snhi1 = (sni + n[0]) * (sh1 + h1) + snhi + n[0] * h1;
  const double snhs1 = (sns + n[3]) * (sh1 + h1) + snhs + n[3] * h1;
  const double snhi2 = (sni + n[1]) * (sh1 + h2) + snhi + n[1] * h2;
  const double snhs2 = (sns + n[4]) * (sh1 + h2) + snhs + n[4] * h2;
  const double snhi3 = (sni + n[2]) * (sh1 + h3) + snhi + n[2] * h3;
  const double snhs3 = (sns + n[5]) * (sh1 + h3) + snhs + n[5] * h3;
  const double hri1 = rri * snhi1 + r * snhs1, hrs1 = rrs * snhs1 + r * snhi1;
  const double hri2 = rri * snhi2 + r * snhs2, hrs2 = rrs * snhs2 + r * snhi2;
  const double hri3 = rri * snhi3 + r * snhs3, hrs3 = rrs * snhs3 + r * snhi3;
  const double rr11 = 2 * (rrri * (sni + n[0] + n[0]) + rrrs * (sns + n[3] + n[3]));
  const double rr22 = 2 * (rrri * (sni + n[1] + n[1]) + rrrs * (sns + n[4] + n[4]));
  const double rr33 = 2 * (rrri * (sni + n[2] + n[2]) + rrrs * (sns + n[5] + n[5]));
  const double rr12 = rrri * (sni + n[0] + n[1]) + rrrs * (sns + n[3] + n[4]);
  const double rr13 = rrri * (sni + n[0] + n[2]) + rrrs * (sns + n[3] + n[5]);
  const double rr23 = rrri * (sni + n[1] + n[2]) + rrrs * (sns + n[4] + n[5]);
  c[0] = dd * (hhi12 - d1 * hri2 - d2 * hri1 + rr12);
  c[1] = dd * (hhi13 - d1 * hri3 - d3 * hri1 + rr13);
  c[2] = dd * (-hh12 - hh13 + d1 * (hri1 - hrs1) - rr11);
  c[3] = dd * (hh12 + d1 * hri2 - d2 * hrs1 - rr12);
  c[4] = dd * (hh13 + d1 * hri3 - d3 * hrs1 - rr13);
  c[5] = dd * (hhi23 - d2 * hri3 - d3 * hri2 + rr23);
  c[6] = dd * (hh12 + d2 * hri1 - d1 * hrs2 - rr12);
  c[7] = dd * (-hh12 - hh23 + d2 * (hri2 - hrs2) - rr22);
  c[8] = dd * (hh23 + d2 * hri3 - d3 * hrs2 - rr23);
  c[9] = dd * (hh13 + d3 * hri1 - d1 * hrs3 - rr13);
  c[10] = dd * (hh23 + d3 * hri2 - d2 * hrs3 - rr23);
  c[11] = dd * (-hh13 - hh23 + d3 * (hri3 - hrs3) - rr33);
  c[12] = dd * (hhs12 + d1 * hrs2 + d2 * hrs1 + rr12);
  c[13] = dd * (hhs13 + d1 * hrs3 + d3 * hrs1 + rr13);
  c[14] = dd * (hhs23 + d2 * hrs3 + d3 * hrs2 + rr23);
}

void orc_mt02pp(double* t, double* xm, double xmul, const double* f, const double* g,
                const double* h, int h_zero, int npou0, const double* xel,
                const double* yel, const double* z, const double* surfac, const int* ikle,
                int nelem, int nelmax, int inchyd, int mon) {
  /* mt02pp.f:155-168 (XS06 cancelled when H is the zero function), 200-215
   * (npou0: NPOIN2 for a P1 vertical viscosity, 0 for P0 on the vertical),
   * 229-445 / 520-707 (with / without the hydrostatic inconsistency test),
   * 713-785 ('MON'), 791-831 (diagonal = minus the row sums). */
  const double chouia = 1e-4;
  const double xs06 = h_zero ? 0.0 : xmul / 6.0;
  const double xs2880 = xmul / 2880.0;
#define XM(c) xm[(long)((c)-1) * nelmax + e]
#define T(c) t[(long)((c)-1) * nelmax + e]
#define IK(k) ikle[(long)((k)-1) * nelmax + e]
  for (int e = 0; e < nelem; ++e) {
    int ik[6];
    double zz[6];
    for (int k = 0; k < 6; ++k) {
      ik[k] = IK(k + 1);
      zz[k] = z[ik[k]];
    }
    const double hh[3] = {zz[3] - zz[0], zz[4] - zz[1], zz[5] - zz[2]};
    int dry = hh[0] < chouia || hh[1] < chouia || hh[2] < chouia;
    if (inchyd) {
      double zb = fmax(fmax(zz[0], zz[1]), zz[2]);
      double zt = fmin(fmin(zz[3], zz[4]), zz[5]);
      dry = dry || zb > zt;
    }
    double nf[6], ng[6], nh[3];
    for (int k = 0; k < 6; ++k) {
      nf[k] = dry ? 0.0 : f[ik[k]] / hh[k % 3];
      ng[k] = dry ? 0.0 : g[ik[k]] / hh[k % 3];
    }
    for (int k = 0; k < 3; ++k) nh[k] = dry ? 0.0 : (h[ik[k]] + h[ik[k] + npou0]) / hh[k];
    const double dd = xs2880 / surfac[e];
    const double x2 = xel[(long)1 * nelmax + e], x3 = xel[(long)2 * nelmax + e];
    const double y2 = yel[(long)1 * nelmax + e], y3 = yel[(long)2 * nelmax + e];
    double cx[15], cy[15];
    mt02pp_dir(y2 - y3, y3, -y2, nf, hh, zz, dd, cx);
    mt02pp_dir(x3 - x2, -x3, x2, ng, hh, zz, dd, cy);
    for (int q = 0; q < 15; ++q) XM(q + 1) = cx[q] + cy[q];
    const double dz = surfac[e] * xs06;
    XM(3) = XM(3) - dz * nh[0];
    XM(8) = XM(8) - dz * nh[1];
    XM(12) = XM(12) - dz * nh[2];
  }
  if (mon) {
    for (int e = 0; e < nelem; ++e)
      for (int q = 1; q <= 15; ++q)
        XM(q) = xmul > 0.0 ? fmin(XM(q), 0.0) : fmax(XM(q), 0.0);
  }
  for (int e = 0; e < nelem; ++e) {
    /* row r of the pair table: the pairs that hold node r */
    T(1) = -XM(1) - XM(2) - XM(3) - XM(4) - XM(5);
    T(2) = -XM(1) - XM(6) - XM(7) - XM(8) - XM(9);
    T(3) = -XM(2) - XM(6) - XM(10) - XM(11) - XM(12);
    T(4) = -XM(3) - XM(7) - XM(10) - XM(13) - XM(14);
    T(5) = -XM(4) - XM(8) - XM(11) - XM(13) - XM(15);
    T(6) = -XM(5) - XM(9) - XM(12) - XM(14) - XM(15);
  }
#undef XM
#undef T
#undef IK
}
