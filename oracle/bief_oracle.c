/*
 * CPU restatement (FP64, sequential) of TELEMAC's BIEF linear-system path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation and the
 * "port" CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product never does.
 *
 * Built with -ffp-contract=off so that every a*b+c is two roundings, as the
 * reference's generic x86-64 flang build computes it.
 *
 * Each function cites the reference routine (sources/utils/bief/...) whose
 * arithmetic and evaluation order it restates.
 */
#include "bief_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* coefficient maps                                                        */
/* ---------------------------------------------------------------------- */

/* index of the pair (i<j) in BIEF's row-major upper-triangle order:
 * triangles (12,13,23); prisms (12,13,14,15,16,23,...,56)
 * (matvct.f:152-169, mv0606.f:117-200) */
static int pair_index(int ndp, int i, int j) {
  int idx = 0;
  for (int a = 0; a < ndp; ++a)
    for (int b = a + 1; b < ndp; ++b) {
      if (a == i && b == j) return idx;
      ++idx;
    }
  return -1;
}

/* column of XA holding A(i,j), i != j (symmetric: A(j,i) = A(i,j);
 * non-symmetric: lower entries follow the npair upper ones in the same
 * pair order, e.g. triangles 21->4, 31->5, 32->6) */
static int coef_col(int ndp, char typext, int i, int j) {
  int npair = ndp * (ndp - 1) / 2;
  if (i < j) return pair_index(ndp, i, j);
  int p = pair_index(ndp, j, i);
  return (typext == 'Q') ? npair + p : p;
}

double orc_dot(int n, const double* x, const double* y) {
  /* dot.f:52-53: left-to-right accumulation */
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + x[i] * y[i];
  return s;
}

/* ---------------------------------------------------------------------- */
/* MATVCT + MV0303/MV0606 + ASSVEC                                         */
/* ---------------------------------------------------------------------- */

void orc_assvec(double* x, const int* ikle, int nelem, int nelmax, int ndp,
                const double* w) {
  /* assvec.f:89-93 -> assve1.f:100-102: k outer, elements in order */
  for (int k = 0; k < ndp; ++k) {
    const int* ik = ikle + (long)k * nelmax;
    const double* wk = w + (long)k * nelmax;
    for (int e = 0; e < nelem; ++e) x[ik[e]] = x[ik[e]] + wk[e];
  }
}

void orc_matvct(int transposed, double* x, const double* da, char typdia,
                const double* xa, char typext, const double* y,
                const int* ikle, int npoin, int nelem, int nelmax, int ndp,
                double* w) {
  /* off-diagonal contributions per element (mv0303.f:123-130 X=AY,
   * :405-420 X=TAY; mv0606.f:117-200): W_k = sum_{j != k} A(k,j) Y(i_j),
   * j increasing, first term not added to zero. */
  if (typext == 'Q' || typext == 'S') {
    int col[6][6];
    for (int k = 0; k < ndp; ++k)
      for (int j = 0; j < ndp; ++j)
        col[k][j] = (k == j) ? -1
                    : (transposed ? coef_col(ndp, typext, j, k)
                                  : coef_col(ndp, typext, k, j));
    for (int e = 0; e < nelem; ++e) {
      for (int k = 0; k < ndp; ++k) {
        double s = 0.0;
        int first = 1;
        for (int j = 0; j < ndp; ++j) {
          if (j == k) continue;
          double term = xa[(long)col[k][j] * nelmax + e] *
                        y[ikle[(long)j * nelmax + e]];
          if (first) {
            s = term;
            first = 0;
          } else {
            s = s + term;
          }
        }
        w[(long)k * nelmax + e] = s;
      }
    }
  } else {
    for (int k = 0; k < ndp; ++k)
      for (int e = 0; e < nelem; ++e) w[(long)k * nelmax + e] = 0.0;
  }
  /* diagonal contribution (mv0303.f:150-160): 'Q' X=YZ, 'I' X=Y, '0' X=0 */
  if (typdia == 'Q') {
    for (int i = 0; i < npoin; ++i) x[i] = y[i] * da[i];
  } else if (typdia == 'I') {
    for (int i = 0; i < npoin; ++i) x[i] = y[i];
  } else {
    for (int i = 0; i < npoin; ++i) x[i] = 0.0;
  }
  /* matvct.f:552-561: ASSVEC with INIT=.FALSE. */
  orc_assvec(x, ikle, nelem, nelmax, ndp, w);
}

/* MVSEG (mvseg.f:110-728) for a square matrix in segment storage
 * (MATRIX STORAGE = 3): gloseg holds the two nodes of each segment
 * (GLOSEG(NSEG,2)), xa the off-diagonal terms XA1 = A(g1,g2) in column 1
 * and, for TYPEXT 'Q', XA2 = A(g2,g1) in column 2. The diagonal part of
 * each operation first (the OV calls), then the segments in order, each
 * adding to its two nodes. */
static void orc_mvseg(int op, double* x, const double* da, char typdia, const double* xa,
                      char typext, const double* y, double c, const int* gloseg, int npoin,
                      int nseg, int nsegmax) {
  const int trans = op >= ORC_TAY;
  const double* xa1 = xa;
  const double* xa2 = typext == 'Q' ? xa + nsegmax : xa;
  if (trans) { /* X=TAY family: XA2 towards GLOSEG1, XA1 towards GLOSEG2 */
    const double* tmp = xa1;
    xa1 = xa2;
    xa2 = tmp;
  }
  for (int i = 0; i < npoin; ++i) {
    const double d = typdia == 'Q' ? da[i] : 1.0;
    const int diag = typdia == 'Q' || typdia == 'I';
    switch (op) {
      case ORC_AY: case ORC_TAY:
        x[i] = !diag ? 0.0 : (typdia == 'Q' ? y[i] * d : y[i]);
        break;
      case ORC_CAY:
        x[i] = !diag ? 0.0 : (typdia == 'Q' ? c * y[i] * d : c * y[i]);
        break;
      case ORC_NAY: case ORC_NTAY:
        x[i] = !diag ? 0.0 : (typdia == 'Q' ? -(y[i] * d) : -y[i]);
        break;
      case ORC_XPAY: case ORC_XPTAY:
        if (diag) x[i] = typdia == 'Q' ? x[i] + y[i] * d : x[i] + y[i];
        break;
      case ORC_XMAY: case ORC_XMTAY:
        if (diag) x[i] = typdia == 'Q' ? x[i] - y[i] * d : x[i] - y[i];
        break;
      default: /* X=X+CAY, X=X+CTAY */
        if (diag) x[i] = typdia == 'Q' ? x[i] + c * y[i] * d : x[i] + c * y[i];
        break;
    }
  }
  if (typext != 'Q' && typext != 'S') return;
  const int* g1 = gloseg;
  const int* g2 = gloseg + nsegmax;
  for (int s = 0; s < nseg; ++s) {
    const int a = g1[s], b = g2[s];
    switch (op) {
      case ORC_AY: case ORC_TAY: case ORC_XPAY: case ORC_XPTAY:
        x[a] = x[a] + xa1[s] * y[b];
        x[b] = x[b] + xa2[s] * y[a];
        break;
      case ORC_NAY: case ORC_NTAY: case ORC_XMAY: case ORC_XMTAY:
        x[a] = x[a] - xa1[s] * y[b];
        x[b] = x[b] - xa2[s] * y[a];
        break;
      default: /* X=CAY, X=X+CAY, X=X+CTAY */
        x[a] = x[a] + c * xa1[s] * y[b];
        x[b] = x[b] + c * xa2[s] * y[a];
        break;
    }
  }
}

void orc_matvec(int op, double* x, const double* da, char typdia,
                const double* xa, char typext, const double* y, double c,
                const int* ikle, int npoin, int nelem, int nelmax, int ndp,
                double* w, double* t) {
  if (ndp == 2) {
    /* segment storage: MATVCT calls MVSEG with the operation itself when X
     * and Y are different vectors (matvec.f:289-306) */
    orc_mvseg(op, x, da, typdia, xa, typext, y, c, ikle, npoin, nelem, nelmax);
    return;
  }
  /* matvec.f:349-380: T = A Y or A^T Y, then X combined with T by OS */
  int trans = (op >= ORC_TAY);
  orc_matvct(trans, t, da, typdia, xa, typext, y, ikle, npoin, nelem, nelmax,
             ndp, w);
  for (int i = 0; i < npoin; ++i) {
    switch (op) {
      case ORC_AY:
      case ORC_TAY: x[i] = t[i]; break;
      case ORC_CAY: x[i] = c * t[i]; break;
      case ORC_NAY:
      case ORC_NTAY: x[i] = -t[i]; break;
      case ORC_XPAY:
      case ORC_XPTAY: x[i] = x[i] + t[i]; break;
      case ORC_XMAY:
      case ORC_XMTAY: x[i] = x[i] - t[i]; break;
      case ORC_XPCAY:
      case ORC_XPCTAY: x[i] = x[i] + c * t[i]; break;
      default: break;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* element matrices                                                        */
/* ---------------------------------------------------------------------- */

void orc_mt01aa(double* t, double* xm, double xmul, const double* surfac,
                int nelem, int nelmax) {
  /* mt01aa.f:83-106 */
  double sur12 = xmul / 12.0;
  for (int e = 0; e < nelem; ++e) {
    double det = surfac[e] * sur12;
    xm[e] = det;
    xm[nelmax + e] = det;
    xm[2L * nelmax + e] = det;
    t[e] = det + det;
    t[nelmax + e] = det + det;
    t[2L * nelmax + e] = det + det;
  }
}

static void mt02_diag(double* t, const double* xm, int nelmax, int e) {
  double a12 = xm[e], a13 = xm[nelmax + e], a23 = xm[2L * nelmax + e];
  t[e] = -a12 - a13;
  t[nelmax + e] = -a12 - a23;
  t[2L * nelmax + e] = -a13 - a23;
}

void orc_mt02aa(int kind, double* t, double* xm, double xmul,
                const double* u, const double* v, int ldu,
                const double* xel, const double* yel, const double* surfac,
                const int* ikle, int nelem, int nelmax, int mon) {
  double xsur12 = xmul / 12.0, xsur48 = xmul / 48.0;
  double *a12 = xm, *a13 = xm + nelmax, *a23 = xm + 2L * nelmax;
  for (int e = 0; e < nelem; ++e) {
    double x2 = xel[nelmax + e], x3 = xel[2L * nelmax + e];
    double y2 = yel[nelmax + e], y3 = yel[2L * nelmax + e];
    int i1 = ikle[e], i2 = ikle[nelmax + e], i3 = ikle[2L * nelmax + e];
    if (kind == 0) {
      /* mt02aa.f:106-136, P0 isotropic viscosity */
      double s2d = xmul * 0.25 / surfac[e];
      double visc1 = u[e] * s2d;
      double bpe = visc1 * (y3 * y3 + x3 * x3);
      double cpf = visc1 * (y2 * y3 + x2 * x3);
      double apd = visc1 * (y2 * y2 + x2 * x2);
      t[e] = apd + bpe - 2 * cpf;
      t[nelmax + e] = bpe;
      t[2L * nelmax + e] = apd;
      a12[e] = cpf - bpe;
      a13[e] = cpf - apd;
      a23[e] = -cpf;
      continue;
    } else if (kind == 3) {
      /* mt02aa.f:142-176, P0 tensor viscosity */
      double s2d = xmul * 0.25 / surfac[e];
      double visc1 = u[e] * s2d, visc2 = u[(long)nelmax + e] * s2d;
      double visc3 = u[2L * nelmax + e] * s2d;
      double bpe = visc1 * y3 * y3 + visc2 * x3 * x3;
      double cpf = visc1 * y2 * y3 + visc2 * x2 * x3;
      double apd = visc1 * y2 * y2 + visc2 * x2 * x2;
      double x2mx3 = x2 - x3, y2my3 = y2 - y3;
      a12[e] = cpf - bpe - (y2my3 * x3 + x2mx3 * y3) * visc3;
      a13[e] = cpf - apd + (y2my3 * x2 + x2mx3 * y2) * visc3;
      a23[e] = -cpf + (x2 * y3 + x3 * y2) * visc3;
    } else if (kind == 2) {
      /* mt02aa.f:200-240, 'UV': coefficient U*V integrated exactly */
      double g1 = v[i1], g2 = v[i2], g3 = v[i3];
      double coef1 = xsur48 / surfac[e];
      double g123 = g1 + g2 + g3;
      double coef2 = u[i1] * (g1 + g123) + u[i2] * (g2 + g123) +
                     u[i3] * (g3 + g123);
      a12[e] = (y3 * (y2 - y3) + x3 * (x2 - x3)) * coef2 * coef1;
      a13[e] = -(y2 * (y2 - y3) + x2 * (x2 - x3)) * coef2 * coef1;
      a23[e] = -(y2 * y3 + x2 * x3) * coef2 * coef1;
    } else if (kind == 1) {
      /* mt02aa.f:247-290, P1 isotropic viscosity */
      double somvx = (u[i1] + u[i2] + u[i3]) * xsur12 / surfac[e];
      double x2x3 = x2 * x3 + y2 * y3;
      double x2aux = x2 * (-x2 + x3) + y2 * (-y2 + y3);
      double x3aux = x3 * (-x2 + x3) + y3 * (-y2 + y3);
      a12[e] = -somvx * x3aux;
      a13[e] = somvx * x2aux;
      a23[e] = -somvx * x2x3;
    } else {
      /* mt02aa.f:296-365, P1 tensor viscosity, columns at stride ldu */
      double somvx = u[i1] + u[i2] + u[i3];
      double somvy = u[i1 + ldu] + u[i2 + ldu] + u[i3 + ldu];
      double somvz = u[i1 + 2L * ldu] + u[i2 + 2L * ldu] + u[i3 + 2L * ldu];
      double aux = xsur12 / surfac[e];
      double x2x3 = x2 * x3, y2y3 = y2 * y3;
      double x2aux = x2 * (-x2 + x3), y2aux = y2 * (-y2 + y3);
      double x3aux = x3 * (-x2 + x3), y3aux = y3 * (-y2 + y3);
      double x2mx3 = x2 - x3, y2my3 = y2 - y3;
      a12[e] = (-somvx * y3aux - somvy * x3aux - y2my3 * x3 * somvz -
                x2mx3 * y3 * somvz) * aux;
      a13[e] = (somvx * y2aux + somvy * x2aux + y2my3 * x2 * somvz +
                x2mx3 * y2 * somvz) * aux;
      a23[e] = (-somvx * y2y3 - somvy * x2x3 + somvz * x3 * y2 +
                somvz * x2 * y3) * aux;
    }
    mt02_diag(t, xm, nelmax, e);
  }
  if (mon) {
    /* mt02aa.f:371-395 */
    for (int e = 0; e < nelem; ++e) {
      if (xmul > 0.0) {
        a12[e] = fmin(a12[e], 0.0);
        a13[e] = fmin(a13[e], 0.0);
        a23[e] = fmin(a23[e], 0.0);
      } else {
        a12[e] = fmax(a12[e], 0.0);
        a13[e] = fmax(a13[e], 0.0);
        a23[e] = fmax(a23[e], 0.0);
      }
      mt02_diag(t, xm, nelmax, e);
    }
  }
}

void orc_assemble_diag(double* da, const double* t, const int* ikle,
                       int npoin, int nelem, int nelmax, int ndp) {
  /* matrix.f:218-227: ASSVEC of the element diagonal with INIT=.TRUE. */
  for (int i = 0; i < npoin; ++i) da[i] = 0.0;
  orc_assvec(da, ikle, nelem, nelmax, ndp, t);
}

/* ---------------------------------------------------------------------- */
/* OM 'M=DMD' and PRECD1                                                   */
/* ---------------------------------------------------------------------- */

void orc_om_dmd(double* da, char* typdia, double* xa, char typext,
                const double* d, const int* ikle, int npoin, int nelem,
                int nelmax, int ndp) {
  /* om1111.f:559-600 (triangles), om4141.f (prisms) */
  if (*typdia == 'Q') {
    for (int i = 0; i < npoin; ++i) da[i] = da[i] * d[i];
    for (int i = 0; i < npoin; ++i) da[i] = da[i] * d[i];
  } else if (*typdia == 'I') {
    for (int i = 0; i < npoin; ++i) da[i] = d[i] * d[i];
    *typdia = 'Q';
  }
  if (typext == 'S' || typext == 'Q') {
    int npair = ndp * (ndp - 1) / 2;
    int ncoef = (typext == 'Q') ? 2 * npair : npair;
    for (int c = 0; c < ncoef; ++c) {
      int p = c % npair, i = 0, j = 0, idx = 0;
      for (int a = 0; a < ndp; ++a)
        for (int b = a + 1; b < ndp; ++b) {
          if (idx == p) {
            i = a;
            j = b;
          }
          ++idx;
        }
      /* XM(e,c) = XM(e,c) * D(IKLE(e,j)) * D(IKLE(e,i)); segments:
       * XM(s) = XM(s) * D(GLOSEG(s,1)) * D(GLOSEG(s,2)) (omseg.f:524-536) */
      double* col = xa + (long)c * nelmax;
      const int* ii = ikle + (long)i * nelmax;
      const int* jj = ikle + (long)j * nelmax;
      if (ndp == 2)
        for (int e = 0; e < nelem; ++e) col[e] = col[e] * d[ii[e]] * d[jj[e]];
      else
        for (int e = 0; e < nelem; ++e) col[e] = col[e] * d[jj[e]] * d[ii[e]];
    }
  }
}

void orc_precd1(double* x, double* da, char* typdia, double* xa, char typext,
                double* b, double* d, const int* ikle, int npoin, int nelem,
                int nelmax, int ndp, int precon) {
  /* precd1.f:90-149 (DIADON false, PREXSM true, one domain) */
  for (int i = 0; i < npoin; ++i) d[i] = da[i];
  if (precon == 5)
    for (int i = 0; i < npoin; ++i) d[i] = fabs(d[i]);
  for (int i = 0; i < npoin; ++i) d[i] = sqrt(d[i]);
  for (int i = 0; i < npoin; ++i) x[i] = x[i] * d[i];
  /* OS 'X=1/Y' with IOPT=2, D=1, EPS=1e-10 (ovd.f:92-102) */
  for (int i = 0; i < npoin; ++i) d[i] = (fabs(d[i]) > 1e-10) ? 1.0 / d[i] : 1.0;
  orc_om_dmd(da, typdia, xa, typext, d, ikle, npoin, nelem, nelmax, ndp);
  if (precon % 2 == 0 || precon % 3 == 0) *typdia = 'I';
  for (int i = 0; i < npoin; ++i) b[i] = b[i] * d[i];
}

/* ---------------------------------------------------------------------- */
/* Krylov loops                                                            */
/* ---------------------------------------------------------------------- */

typedef struct {
  const double* da;
  char typdia;
  const double* xa;
  char typext;
  const int* ikle;
  int npoin, nelem, nelmax, ndp;
  double *w, *t;
} mvctx;

static void mv_ay(mvctx* m, double* x, const double* y) {
  /* MATRBL 'X=AY' -> MATVEC (matrbl.f:194-202, matvec.f:349-380) */
  orc_matvec(ORC_AY, x, m->da, m->typdia, m->xa, m->typext, y, 0.0, m->ikle,
             m->npoin, m->nelem, m->nelmax, m->ndp, m->w, m->t);
}

void orc_gracjg(double* x, const double* da, char typdia, const double* xa,
                char typext, const double* b, const int* ikle, int npoin,
                int nelem, int nelmax, int ndp, double eps, int nitmax,
                double* work, orc_info* info) {
  /* gracjg.F:184-420, without Crout/GS/3D preconditioning, so G == R
   * (solve.f:474) */
  mvctx m = {da, typdia, xa, typext, ikle, npoin, nelem, nelmax, ndp,
             work, work + (long)ndp * nelmax};
  double* d = m.t + npoin;
  double* ad = d + npoin;
  double* r = ad + npoin;
  int n = npoin;
  const double rmin = 1e-15;
  memset(info, 0, sizeof(*info));
  int mit = 0;
  double xl = orc_dot(n, b, b);
  double testl = sqrt(xl);
  if (testl < rmin) {
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    info->zero_rhs = 1;
    info->converged = 1;
    info->relat = 1;
    return;
  }
  int relat = 1;
  if (xl < 1.0) {
    xl = 1.0;
    relat = 0;
  }
  info->relat = relat;
  mv_ay(&m, r, x);
  for (int i = 0; i < n; ++i) r[i] = r[i] - b[i];
  double rmrm = orc_dot(n, r, r);
  double rmgm = rmrm;
  if (rmrm < eps * eps * xl) {
    info->niter = 0;
    info->relres = sqrt(rmrm / xl);
    info->converged = 1;
    return;
  }
  for (int i = 0; i < n; ++i) d[i] = r[i];
  mv_ay(&m, ad, d);
  double dad = orc_dot(n, d, ad);
  double ro = rmgm / dad;
  for (int i = 0; i < n; ++i) x[i] = x[i] + (-ro) * d[i];
  for (;;) {
    mit = mit + 1;
    for (int i = 0; i < n; ++i) r[i] = r[i] + (-ro) * ad[i];
    double rm1gm1 = rmgm;
    rmrm = orc_dot(n, r, r);
    double rmdm = rmrm;
    rmgm = rmrm;
    if (rmrm <= xl * eps * eps) {
      info->converged = 1;
      break;
    }
    double beta = rmgm / rm1gm1;
    for (int i = 0; i < n; ++i) d[i] = r[i] + beta * d[i];
    mv_ay(&m, ad, d);
    dad = orc_dot(n, d, ad);
    ro = rmdm / dad;
    for (int i = 0; i < n; ++i) x[i] = x[i] + (-ro) * d[i];
    if (!(mit < nitmax)) break;
  }
  info->niter = mit;
  info->relres = sqrt(rmrm / xl);
}

void orc_cgstab(double* x, const double* da, char typdia, const double* xa,
                char typext, const double* b, const int* ikle, int npoin,
                int nelem, int nelmax, int ndp, double eps, int nitmax,
                double* work, orc_info* info) {
  /* cgstab.f:113-236, without Crout preconditioning */
  mvctx m = {da, typdia, xa, typext, ikle, npoin, nelem, nelmax, ndp,
             work, work + (long)ndp * nelmax};
  int n = npoin;
  double* p = m.t + npoin;
  double* q = p + n;
  double* r = q + n;
  double* s = r + n;
  double* t = s + n;
  double* v = t + n;
  const double rmin = 1e-15;
  memset(info, 0, sizeof(*info));
  int mit = 0;
  double xl = orc_dot(n, b, b);
  int relat = 1;
  if (xl < 1.0) {
    xl = 1.0;
    relat = 0;
  }
  info->relat = relat;
  double rmrm;
  if (sqrt(xl) < rmin) {
    /* unreachable after the clamp above, kept as in the reference */
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    info->zero_rhs = 1;
    info->converged = 1;
    return;
  }
  mv_ay(&m, v, x);
  for (int i = 0; i < n; ++i) r[i] = b[i] - v[i];
  rmrm = orc_dot(n, r, r);
  if (rmrm < eps * eps * xl) {
    info->converged = 1;
    info->relres = sqrt(rmrm / xl);
    return;
  }
  for (int i = 0; i < n; ++i) r[i] = b[i];
  for (int i = 0; i < n; ++i) r[i] = r[i] - v[i];
  for (int i = 0; i < n; ++i) p[i] = r[i];
  for (int i = 0; i < n; ++i) v[i] = 0.0;
  for (int i = 0; i < n; ++i) q[i] = 0.0;
  double alfa = 1.0, beta = 1.0, omeg1 = 1.0;
  for (;;) {
    mit = mit + 1;
    double beta1 = orc_dot(n, r, p);
    double omeg2 = omeg1 * beta1 / beta;
    double omeg = omeg2 / alfa;
    beta = beta1;
    for (int i = 0; i < n; ++i) q[i] = r[i] + omeg * q[i];
    for (int i = 0; i < n; ++i) q[i] = q[i] + (-omeg2) * v[i];
    mv_ay(&m, v, q);
    omeg1 = orc_dot(n, p, v);
    omeg1 = beta1 / omeg1;
    for (int i = 0; i < n; ++i) s[i] = r[i] + (-omeg1) * v[i];
    mv_ay(&m, t, s);
    alfa = orc_dot(n, t, s);
    double alfa1 = orc_dot(n, t, t);
    alfa = alfa / alfa1;
    for (int i = 0; i < n; ++i) x[i] = x[i] + omeg1 * q[i];
    for (int i = 0; i < n; ++i) x[i] = x[i] + alfa * s[i];
    for (int i = 0; i < n; ++i) r[i] = s[i] + (-alfa) * t[i];
    rmrm = orc_dot(n, r, r);
    if (rmrm <= xl * eps * eps) {
      info->converged = 1;
      break;
    }
    if (!(mit < nitmax)) break;
  }
  info->niter = mit;
  info->relres = sqrt(rmrm / xl);
}

void orc_gmres(double* x, const double* da, char typdia, const double* xa,
               char typext, const double* b, const int* ikle, int npoin,
               int nelem, int nelmax, int ndp, double eps, int nitmax,
               int krylov, double* work, orc_info* info) {
  /* gmres.f:171-372, without Crout/GS preconditioning; H(21,20) */
  mvctx m = {da, typdia, xa, typext, ikle, npoin, nelem, nelmax, ndp,
             work, work + (long)ndp * nelmax};
  int n = npoin;
  double* r0 = m.t + npoin;
  double* vv = r0 + n;              /* krylov vectors V(1..K) */
  double* av = vv + (long)krylov * n; /* krylov vectors AV(1..K) */
  double h[21][20], cc[20], ss[20], e[21];
  memset(info, 0, sizeof(*info));
  memset(h, 0, sizeof(h));
  double nb = sqrt(orc_dot(n, b, b));
  int relat = 1;
  if (nb < 1.0) {
    nb = 1.0;
    relat = 0;
  }
  info->relat = relat;
  int mit = 0;
  mv_ay(&m, r0, x);
  for (int i = 0; i < n; ++i) r0[i] = r0[i] - b[i];
  double nr0 = sqrt(orc_dot(n, r0, r0));
  double prec = nr0 / nb;
  if (prec <= eps) goto done;
  for (;;) {
    mit = mit + 1;
    int k = krylov;
#define V(j) (vv + (long)(j) * n)
#define AV(j) (av + (long)(j) * n)
    {
      double cst = -1.0 / nr0;
      for (int i = 0; i < n; ++i) V(0)[i] = cst * r0[i];
    }
    for (int j = 0; j < k - 1; ++j) {
      mv_ay(&m, AV(j), V(j));
      for (int i = 0; i < n; ++i) V(j + 1)[i] = AV(j)[i];
      for (int l = 0; l <= j; ++l) {
        h[l][j] = orc_dot(n, V(j + 1), V(l));
        double cst = -h[l][j];
        for (int i = 0; i < n; ++i) V(j + 1)[i] = V(j + 1)[i] + cst * V(l)[i];
      }
      h[j + 1][j] = sqrt(orc_dot(n, V(j + 1), V(j + 1)));
      if (h[j + 1][j] < 1e-20) {
        k = j + 1;
        goto lastcol;
      }
      {
        double cst = 1.0 / h[j + 1][j];
        for (int i = 0; i < n; ++i) V(j + 1)[i] = cst * V(j + 1)[i];
      }
    }
  lastcol:
    mv_ay(&m, AV(k - 1), V(k - 1));
    h[k][k - 1] = orc_dot(n, AV(k - 1), AV(k - 1));
    for (int l = 0; l < k; ++l) {
      h[l][k - 1] = orc_dot(n, AV(k - 1), V(l));
      h[k][k - 1] = h[k][k - 1] - h[l][k - 1] * h[l][k - 1];
    }
    h[k][k - 1] = sqrt(fmax(h[k][k - 1], 0.0));
    e[0] = nr0;
    for (int l = 0; l < k; ++l) {
      for (int j = 0; j < l; ++j) {
        double zz = cc[j] * h[j][l] + ss[j] * h[j + 1][l];
        h[j + 1][l] = -ss[j] * h[j][l] + cc[j] * h[j + 1][l];
        h[j][l] = zz;
      }
      double rr = sqrt(h[l][l] * h[l][l] + h[l + 1][l] * h[l + 1][l]);
      if (fabs(rr) < 1e-6) goto done;
      cc[l] = h[l][l] / rr;
      ss[l] = h[l + 1][l] / rr;
      h[l][l] = rr;
      e[l + 1] = -ss[l] * e[l];
      e[l] = cc[l] * e[l];
    }
    e[k - 1] = e[k - 1] / h[k - 1][k - 1];
    for (int j = k - 2; j >= 0; --j) {
      for (int l = j + 1; l < k; ++l) e[j] = e[j] - h[j][l] * e[l];
      e[j] = e[j] / h[j][j];
    }
    for (int l = 0; l < k; ++l) {
      for (int i = 0; i < n; ++i) x[i] = x[i] + e[l] * V(l)[i];
      for (int i = 0; i < n; ++i) r0[i] = r0[i] + e[l] * AV(l)[i];
    }
    nr0 = fabs(e[k]);
    prec = nr0 / nb;
    if (!(prec >= eps && mit < nitmax)) break;
#undef V
#undef AV
  }
done:
  info->niter = mit;
  info->relres = prec;
  info->converged = (prec < eps) ? 1 : 0;
}

long orc_solve_work_size(int slv, int krylov, int npoin, int nelmax, int ndp) {
  long nvec = 3;
  if (slv == 6) nvec = 6;
  if (slv == 7) nvec = 1 + 2L * krylov;
  return (long)ndp * nelmax + npoin + nvec * npoin + npoin /* d */;
}

int orc_solve(int slv, int precon, int krylov, int nitmax, double eps,
              double* x, double* da, char* typdia, double* xa, char typext,
              double* b, const int* ikle, int npoin, int nelem, int nelmax,
              int ndp, orc_info* info) {
  /* solve.f:387-649 for a single matrix (S = 0) */
  if (slv != 1 && slv != 6 && slv != 7) return 2;
  if (slv == 7 && (krylov < 1 || krylov > 20)) return 2;
  if (precon == 0) precon = 1;
  if (precon % 7 == 0 || precon % 11 == 0 || precon % 13 == 0 ||
      precon % 17 == 0 || precon % 3 == 0)
    return 2;
  long nw = orc_solve_work_size(slv, krylov, npoin, nelmax, ndp);
  double* work = (double*)malloc(sizeof(double) * nw);
  if (!work) return 4;
  double* d = work + nw - npoin;
  int diag = (precon % 2 == 0 || precon % 5 == 0);
  if (diag)
    orc_precd1(x, da, typdia, xa, typext, b, d, ikle, npoin, nelem, nelmax,
               ndp, precon);
  if (slv == 1)
    orc_gracjg(x, da, *typdia, xa, typext, b, ikle, npoin, nelem, nelmax, ndp,
               eps, nitmax, work, info);
  else if (slv == 6)
    orc_cgstab(x, da, *typdia, xa, typext, b, ikle, npoin, nelem, nelmax, ndp,
               eps, nitmax, work, info);
  else
    orc_gmres(x, da, *typdia, xa, typext, b, ikle, npoin, nelem, nelmax, ndp,
              eps, nitmax, krylov, work, info);
  if (diag)
    for (int i = 0; i < npoin; ++i) x[i] = x[i] * d[i];
  free(work);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* DIRICH and LUMP                                                         */
/* ---------------------------------------------------------------------- */

void orc_dirich(double* f, double* da, char typdia, double* xa, char typext,
                double* sm, const double* fbor, const int* limdir, int kdir,
                const int* ikle, int npoin, int nelem, int nelmax, int ndp) {
  /* diri01.f:94-139 with FBOR of the same discretisation as SM */
  double* w = (double*)malloc(sizeof(double) * ((long)ndp * nelmax + 2L * npoin));
  double* t = w + (long)ndp * nelmax;
  double* work2 = t + npoin;
  orc_matvec(ORC_AY, work2, da, typdia, xa, typext, fbor, 0.0, ikle, npoin,
             nelem, nelmax, ndp, w, t);
  for (int i = 0; i < npoin; ++i) sm[i] = sm[i] - work2[i];
  /* DIRAUX (diraux.f): W=1; at prescribed nodes SM=DA*FBOR, W=0, F=FBOR.
   * The diagonal Y passed is S%D: with TYPDIA 'I' its stored values are
   * still what DIRAUX reads. */
  for (int i = 0; i < npoin; ++i) work2[i] = 1.0;
  for (int i = 0; i < npoin; ++i) {
    if (limdir[i] == kdir) {
      sm[i] = da[i] * fbor[i];
      work2[i] = 0.0;
      f[i] = fbor[i];
    }
  }
  /* OM 'M=DMD' with the diagonal declared '0' so it is left untouched */
  char td = '0';
  orc_om_dmd(da, &td, xa, typext, work2, ikle, npoin, nelem, nelmax, ndp);
  free(w);
}

void orc_lump(double* diag, const double* da, char typdia, const double* xa,
              char typext, double xmul, const int* ikle, int npoin, int nelem,
              int nelmax, int ndp) {
  /* lump.f: DIAG = XMUL everywhere, then DIAG = A * DIAG */
  double* w = (double*)malloc(sizeof(double) * ((long)ndp * nelmax + 2L * npoin));
  double* t = w + (long)ndp * nelmax;
  double* ones = t + npoin;
  for (int i = 0; i < npoin; ++i) ones[i] = xmul;
  orc_matvec(ORC_AY, diag, da, typdia, xa, typext, ones, 0.0, ikle, npoin,
             nelem, nelmax, ndp, w, t);
  free(w);
}

/* ---------------------------------------------------------------------- */
/* prism element matrices (restated in a later part of this file)          */
/* ---------------------------------------------------------------------- */
