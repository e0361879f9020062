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
