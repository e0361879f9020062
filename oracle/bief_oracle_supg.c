/*
 * CPU restatement of the SUPG element routines of BIEF (P1 triangles):
 * KSPG11 (SUPG element vectors), MT03AA (MASUPG), MT04AA (MAUGUG) and MT05AA
 * (MATVGR). TEST INFRASTRUCTURE ONLY (see bief_oracle.h).
 *
 * Element terms are written in the order the reference evaluates them, so
 * that the results are the same to the last bit (built with
 * -ffp-contract=off). Storage: t[k*nelmax + e] diagonal terms; xm in the
 * non-symmetric 'Q' order (12,13,23,21,31,32) except MT04AA, which is
 * symmetric (12,13,23).
 */
#include <math.h>

#include "bief_oracle.h"

static double max3(double a, double b, double c) {
  double m = a > b ? a : b;
  return m > c ? m : c;
}

/* ksupg.f -> kspg11.f: the SUPG vector h/3 * U/|U| per element, the
 * element length h the largest height, reduced to the height opposite a
 * node whose gradient direction the velocity follows too closely. */
void orc_kspg11(double* kx, double* ky, const double* xel, const double* yel,
                const double* u, const double* v, const int* ikle, int nelem, int nelmax,
                double xmul) {
  for (int e = 0; e < nelem; ++e) {
    const int i1 = ikle[e], i2 = ikle[nelmax + e], i3 = ikle[2 * nelmax + e];
    const double x2 = xel[nelmax + e], x3 = xel[2 * nelmax + e];
    const double y2 = yel[nelmax + e], y3 = yel[2 * nelmax + e];
    /* unit inward normals of the three sides, from the basis gradients */
    double g1x = y2 - y3, g1y = x3 - x2;
    double r = 1.0 / sqrt(g1x * g1x + g1y * g1y);
    g1x = g1x * r;
    g1y = g1y * r;
    double g2x = y3, g2y = -x3;
    r = 1.0 / sqrt(g2x * g2x + g2y * g2y);
    g2x = g2x * r;
    g2y = g2y * r;
    double g3x = -y2, g3y = x2;
    r = 1.0 / sqrt(g3x * g3x + g3y * g3y);
    g3x = g3x * r;
    g3y = g3y * r;
    /* side lengths and heights */
    const double c3 = sqrt(x2 * x2 + y2 * y2);
    const double c1 = sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
    const double c2 = sqrt(x3 * x3 + y3 * y3);
    const double s = 0.5 * (x2 * y3 - x3 * y2);
    const double h1 = 2 * s / c1, h2 = 2 * s / c2, h3 = 2 * s / c3;
    double h = max3(h1, h2, h3);
    const double um = u[i1] + u[i2] + u[i3];
    const double vm = v[i1] + v[i2] + v[i3];
    double rn = sqrt(um * um + vm * vm);
    rn = 1.0 / (rn > 1e-10 ? rn : 1e-10);
    const double un = um * rn, vn = vm * rn;
    const double a1 = g1x * un + g1y * vn;
    const double a2 = g2x * un + g2y * vn;
    const double a3 = g3x * un + g3y * vn;
    if (a1 * h > h1) h = h1;
    if (a2 * h > h2) h = h2;
    if (a3 * h > h3) h = h3;
    /* the reference's literal, not 1/3 */
    kx[e] = 0.33333333 * xmul * h * un;
    ky[e] = 0.33333333 * xmul * h * vn;
  }
}

/* mt03aa.f (F, G per element; U, V per node):
 * A(i,j) = XMUL int (F,G).grad(phi_i) (U.grad(phi_j)) */
void orc_mt03aa(double* t, double* xm, double xmul, const double* f, const double* g,
                const double* u, const double* v, const double* xel, const double* yel,
                const double* surfac, const int* ikle, int nelem, int nelmax) {
  const double sur12 = xmul / 12.0;
  double* a12 = xm;
  double* a13 = xm + nelmax;
  double* a23 = xm + 2 * (long)nelmax;
  double* a21 = xm + 3 * (long)nelmax;
  double* a31 = xm + 4 * (long)nelmax;
  double* a32 = xm + 5 * (long)nelmax;
  for (int e = 0; e < nelem; ++e) {
    const int i1 = ikle[e], i2 = ikle[nelmax + e], i3 = ikle[2 * nelmax + e];
    const double x2 = xel[nelmax + e], x3 = xel[2 * nelmax + e];
    const double y2 = yel[nelmax + e], y3 = yel[2 * nelmax + e];
    const double u123 = u[i1] + u[i2] + u[i3];
    const double v123 = v[i1] + v[i2] + v[i3];
    const double den = sur12 / surfac[e];
    const double kkx = f[e] * den, kky = g[e] * den;
    const double p1 = x2 * v123 - y2 * u123;
    const double p2 = x3 * v123 - y3 * u123;
    const double q1 = x2 * kky - y2 * kkx;
    const double q2 = x3 * kky - y3 * kkx;
    const double b11 = (p1 - p2) * (q1 - q2);
    const double b22 = p2 * q2;
    const double b12 = p2 * (q1 - q2);
    const double b21 = (p1 - p2) * q2;
    const double b13 = -b11 - b12;
    const double b23 = -b22 - b21;
    t[e] = b11;
    t[nelmax + e] = b22;
    a12[e] = b12;
    a21[e] = b21;
    a13[e] = b13;
    a23[e] = b23;
    a31[e] = -b11 - b21;
    a32[e] = -b22 - b12;
    t[2 * (long)nelmax + e] = -b13 - b23;
  }
}

/* mt04aa.f (U, V per node): the symmetric streamline term
 * A(i,j) = XMUL int (U.grad(phi_i)) (U.grad(phi_j)), 'S' storage */
void orc_mt04aa(double* t, double* xm, double xmul, const double* u, const double* v,
                const double* xel, const double* yel, const double* surfac, const int* ikle,
                int nelem, int nelmax) {
  const double sur48 = xmul / 48.0;
  for (int e = 0; e < nelem; ++e) {
    const int i1 = ikle[e], i2 = ikle[nelmax + e], i3 = ikle[2 * nelmax + e];
    const double x2 = xel[nelmax + e], x3 = xel[2 * nelmax + e];
    const double y2 = yel[nelmax + e], y3 = yel[2 * nelmax + e];
    const double u1 = u[i1], u2 = u[i2], u3 = u[i3];
    const double v1 = v[i1], v2 = v[i2], v3 = v[i3];
    const double u123 = u1 + u2 + u3, v123 = v1 + v2 + v3;
    const double c = sur48 / surfac[e];
    const double uu = u1 * u123 + u2 * u2 + u2 * u3 + u3 * u3;
    const double uv = u1 * (v123 + v1) + u2 * (v123 + v2) + u3 * (v123 + v3);
    const double vv = v1 * v123 + v2 * v2 + v2 * v3 + v3 * v3;
    const double b12 = (2 * y3 * (y2 - y3) * uu + (x3 * (y3 - y2) + (x3 - x2) * y3) * uv +
                        2 * x3 * (x2 - x3) * vv) * c;
    const double b13 = (2 * y2 * (y3 - y2) * uu + (2 * x2 * y2 - x2 * y3 - x3 * y2) * uv +
                        2 * x2 * (x3 - x2) * vv) * c;
    const double b23 = (-(2 * y2 * y3 * uu) + (x2 * y3 + x3 * y2) * uv - 2 * x2 * x3 * vv) * c;
    xm[e] = b12;
    xm[nelmax + e] = b13;
    xm[2 * (long)nelmax + e] = b23;
    t[e] = -b12 - b13;
    t[nelmax + e] = -b23 - b12;
    t[2 * (long)nelmax + e] = -b13 - b23;
  }
}

/* mt05aa.f: A(i,j) = XMUL int phi_i U.grad(phi_j). u_p0: U, V per element
 * (else per node); upwind: the FORMUL(16:16) = 'N' variant (P1 U, V), the
 * positive part of the element's node-to-node fluxes. */
void orc_mt05aa(double* t, double* xm, double xmul, const double* u, const double* v,
                int u_p0, int upwind, const double* xel, const double* yel, const int* ikle,
                int nelem, int nelmax) {
  const double sur6 = xmul / 6.0, sur24 = xmul / 24.0;
  double* a12 = xm;
  double* a13 = xm + nelmax;
  double* a23 = xm + 2 * (long)nelmax;
  double* a21 = xm + 3 * (long)nelmax;
  double* a31 = xm + 4 * (long)nelmax;
  double* a32 = xm + 5 * (long)nelmax;
  double* a11 = t;
  double* a22 = t + nelmax;
  double* a33 = t + 2 * (long)nelmax;
  for (int e = 0; e < nelem; ++e) {
    const int i1 = ikle[e], i2 = ikle[nelmax + e], i3 = ikle[2 * nelmax + e];
    if (u_p0) {
      const double x2 = xel[nelmax + e] * sur24, x3 = xel[2 * nelmax + e] * sur24;
      const double y2 = yel[nelmax + e] * sur24, y3 = yel[2 * nelmax + e] * sur24;
      const double qu = 4 * u[e], qv = 4 * v[e];
      const double r1 = (x3 - x2) * qv + (y2 - y3) * qu;
      const double r2 = -(x3 * qv) + y3 * qu;
      const double r3 = x2 * qv - y2 * qu;
      a11[e] = r1;
      a22[e] = r2;
      a33[e] = r3;
      a12[e] = r2;
      a13[e] = r3;
      a23[e] = r3;
      a21[e] = r1;
      a31[e] = r1;
      a32[e] = r2;
    } else if (upwind) {
      const double x2 = xel[nelmax + e], x3 = xel[2 * nelmax + e];
      const double y2 = yel[nelmax + e], y3 = yel[2 * nelmax + e];
      const double us = (u[i1] + u[i2] + u[i3]) * sur6;
      const double vs = (v[i1] + v[i2] + v[i3]) * sur6;
      const double k1 = us * (y2 - y3) - vs * (x2 - x3);
      const double k2 = us * y3 - vs * x3;
      const double k3 = us * (-y2) - vs * (-x2);
#define POSMIN(a, b) fmax(fmin((a), (b)), 0.0)
      const double l12 = POSMIN(k1, -k2), l13 = POSMIN(k1, -k3);
      const double l21 = POSMIN(k2, -k1), l23 = POSMIN(k2, -k3);
      const double l31 = POSMIN(k3, -k1), l32 = POSMIN(k3, -k2);
#undef POSMIN
      a11[e] = l12 + l13;
      a22[e] = l21 + l23;
      a33[e] = l31 + l32;
      a12[e] = -l12;
      a13[e] = -l13;
      a23[e] = -l23;
      a21[e] = -l21;
      a31[e] = -l31;
      a32[e] = -l32;
    } else {
      const double x2 = xel[nelmax + e] * sur24, x3 = xel[2 * nelmax + e] * sur24;
      const double y2 = yel[nelmax + e] * sur24, y3 = yel[2 * nelmax + e] * sur24;
      const double u1 = u[i1], u2 = u[i2], u3 = u[i3];
      const double v1 = v[i1], v2 = v[i2], v3 = v[i3];
      const double u123 = u1 + u2 + u3, v123 = v1 + v2 + v3;
      /* row i of the element: phi_i tested against U.grad, U weighted toward node i */
      const double w1u = u123 + u1, w1v = v123 + v1;
      const double w2u = u123 + u2, w2v = v123 + v2;
      const double w3u = u123 + u3, w3v = v123 + v3;
      a11[e] = (x3 - x2) * w1v + (y2 - y3) * w1u;
      a22[e] = -(x3 * w2v) + y3 * w2u;
      a33[e] = x2 * w3v - y2 * w3u;
      a12[e] = -(x3 * w1v) + y3 * w1u;
      a13[e] = x2 * w1v - y2 * w1u;
      a23[e] = x2 * w2v - y2 * w2u;
      a21[e] = (x3 - x2) * w2v + (y2 - y3) * w2u;
      a31[e] = (x3 - x2) * w3v + (y2 - y3) * w3u;
      a32[e] = -(x3 * w3v) + y3 * w3u;
    }
  }
}
