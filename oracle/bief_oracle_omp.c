/*
 * Multi-threaded (OpenMP) restatement of BIEF's SOLVE for the conjugate
 * gradient with diagonal preconditioning (SLV 1, PRECON 2 or 1), symmetric
 * or non-symmetric EBE storage of linear elements.
 *
 * TEST INFRASTRUCTURE ONLY: the multi-core CPU baseline beside the
 * sequential reference build. Only bench.py's CPU baseline leg and tests/
 * load it; the product never does.
 *
 * What the reference does on several cores is domain decomposition over MPI
 * (each rank a sub-mesh, the dot products summed by P_DSUM,
 * sources/utils/parallel/p_dsum.F:57-87, the interface nodes by PARCOM).
 * This build shares memory instead: the element loop of MATVEC and the
 * node loops run over threads, and the steps are those of the sequential
 * restatement (oracle/bief_oracle.c), with these properties:
 *  - the mat-vec is node-centric: each node gathers the element terms of
 *    the elements that contain it, in the order (local slot k, element e)
 *    that ASSVEC adds them (assvec.f:89-93 -> assve1.f:100-102), each term
 *    W_k(e) summed over j != k, j increasing, as MV0303/MV0606 form it
 *    (mv0303.f:123-130). So X = A Y equals the sequential product bit for
 *    bit;
 *  - dot products are summed per thread over fixed contiguous chunks and
 *    the chunk sums added in thread order: deterministic for a given thread
 *    count, but not the left-to-right order of dot.f:52-53 (as P_DSUM's sum
 *    over ranks is not either), so iterates differ from the sequential ones
 *    at rounding level and the iteration count by at most one or two.
 * Built with -ffp-contract=off, like the sequential restatement.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "bief_oracle.h"

typedef struct {
  int npoin, nelem, ndp;
  long *start;  /* npoin + 1 */
  int *slot;    /* per incidence: local slot k */
  int *elem;    /* per incidence: element e */
} orc_omp_plan;

static int pair_idx(int ndp, int i, int j) {
  int idx = 0;
  for (int a = 0; a < ndp; ++a)
    for (int b = a + 1; b < ndp; ++b) {
      if (a == i && b == j) return idx;
      ++idx;
    }
  return -1;
}

/* XA column of A(i,j), i != j: upper pairs first, then (for 'Q') the lower
 * ones in the same pair order (matvct.f:152-169) */
static int col_of(int ndp, char typext, int i, int j) {
  int npair = ndp * (ndp - 1) / 2;
  if (i < j) return pair_idx(ndp, i, j);
  int p = pair_idx(ndp, j, i);
  return typext == 'Q' ? npair + p : p;
}

void* orc_omp_plan_create(const int* ikle, int npoin, int nelem, int ndp) {
  if (npoin <= 0 || nelem <= 0 || ndp < 2 || ndp > 6) return NULL;
  orc_omp_plan* p = (orc_omp_plan*)calloc(1, sizeof(*p));
  if (!p) return NULL;
  long ninc = (long)nelem * ndp;
  p->npoin = npoin;
  p->nelem = nelem;
  p->ndp = ndp;
  p->start = (long*)calloc((size_t)npoin + 1, sizeof(long));
  p->slot = (int*)malloc(sizeof(int) * ninc);
  p->elem = (int*)malloc(sizeof(int) * ninc);
  if (!p->start || !p->slot || !p->elem) {
    free(p->start); free(p->slot); free(p->elem); free(p);
    return NULL;
  }
  for (long i = 0; i < ninc; ++i) {
    int n = ikle[i];
    if (n < 0 || n >= npoin) {
      free(p->start); free(p->slot); free(p->elem); free(p);
      return NULL;
    }
    p->start[n + 1]++;
  }
  for (int i = 0; i < npoin; ++i) p->start[i + 1] += p->start[i];
  long* fill = (long*)malloc(sizeof(long) * npoin);
  memcpy(fill, p->start, sizeof(long) * npoin);
  /* (k, e) ascending: ASSVEC's order of addition at every node */
  for (int k = 0; k < ndp; ++k)
    for (int e = 0; e < nelem; ++e) {
      int n = ikle[(long)k * nelem + e];
      long at = fill[n]++;
      p->slot[at] = k;
      p->elem[at] = e;
    }
  free(fill);
  return p;
}

void orc_omp_plan_free(void* plan) {
  orc_omp_plan* p = (orc_omp_plan*)plan;
  if (!p) return;
  free(p->start); free(p->slot); free(p->elem); free(p);
}

/* X = A Y with TYPDIA 'Q' or 'I' (matvct.f:552-561 + mv0303.f:150-160) */
static void omp_ay(const orc_omp_plan* p, double* x, const double* da, char typdia,
                   const double* xa, char typext, const int* ikle, const double* y,
                   int col[6][6]) {
  const int ndp = p->ndp, ne = p->nelem;
#pragma omp for schedule(static)
  for (int i = 0; i < p->npoin; ++i) {
    double s = typdia == 'Q' ? y[i] * da[i] : y[i];
    for (long a = p->start[i]; a < p->start[i + 1]; ++a) {
      const int k = p->slot[a], e = p->elem[a];
      double w = 0.0;
      int first = 1;
      for (int j = 0; j < ndp; ++j) {
        if (j == k) continue;
        double term = xa[(long)col[k][j] * ne + e] * y[ikle[(long)j * ne + e]];
        if (first) {
          w = term;
          first = 0;
        } else {
          w = w + term;
        }
      }
      s = s + w;
    }
    x[i] = s;
  }
}

/* chunk sums in thread order; partial has one slot per thread */
static double omp_dot(int n, const double* a, const double* b, double* partial) {
  int t = omp_get_thread_num(), nt = omp_get_num_threads();
  long lo = (long)n * t / nt, hi = (long)n * (t + 1) / nt;
  double s = 0.0;
  for (long i = lo; i < hi; ++i) s = s + a[i] * b[i];
  partial[t] = s;
#pragma omp barrier
  double tot = 0.0;
  for (int u = 0; u < nt; ++u) tot = tot + partial[u];
#pragma omp barrier
  return tot;
}

/* SOLVE (solve.f:387-649) with SLV 1 and PRECON 1 or 2: PRECD1
 * (precd1.f:90-149), GRACJG (gracjg.F:184-420, G == R), X = D X. da and xa
 * are overwritten with D A D as the reference does; b with D B. Returns 0,
 * 2 for an unsupported option, 4 for an allocation failure. */
int orc_omp_solve_cg(void* plan, const int* ikle, double* da, char typdia, double* xa,
                     char typext, double* b, double* x, double eps, int nitmax, int precon,
                     int nthreads, orc_info* info) {
  orc_omp_plan* p = (orc_omp_plan*)plan;
  if (!p || (typdia != 'Q' && typdia != 'I') || (typext != 'S' && typext != 'Q')) return 2;
  if (precon == 0) precon = 1;
  if (precon != 1 && precon != 2) return 2;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const int n = p->npoin, ndp = p->ndp, ne = p->nelem;
  double* buf = (double*)malloc(sizeof(double) * (4L * n + nthreads));
  if (!buf) return 4;
  double *dd = buf, *ad = dd + n, *r = ad + n, *dg = r + n;
  double* partial = dg + n;
  int col[6][6];
  for (int k = 0; k < ndp; ++k)
    for (int j = 0; j < ndp; ++j) col[k][j] = k == j ? -1 : col_of(ndp, typext, k, j);
  const int npair = ndp * (ndp - 1) / 2, ncoef = typext == 'Q' ? 2 * npair : npair;
  int pi[15], pj[15];
  for (int a = 0, c = 0; a < ndp; ++a)
    for (int bb = a + 1; bb < ndp; ++bb, ++c) {
      pi[c] = a;
      pj[c] = bb;
    }
  memset(info, 0, sizeof(*info));
  const int diag = precon == 2;
  char td = typdia;
  int mit = 0, conv = 0, zero = 0, relat = 1;
  double rmrm = 0.0, xl = 0.0;
#pragma omp parallel num_threads(nthreads)
  {
    if (diag) {
      /* PRECD1: D = sqrt(DA); X = X D; D = 1/D (ovd.f:92-102); A = D A D */
#pragma omp for schedule(static)
      for (int i = 0; i < n; ++i) {
        double s = sqrt(da[i]);
        x[i] = x[i] * s;
        dg[i] = fabs(s) > 1e-10 ? 1.0 / s : 1.0;
        da[i] = da[i] * dg[i];
        da[i] = da[i] * dg[i];
        b[i] = b[i] * dg[i];
      }
      for (int c = 0; c < ncoef; ++c) {
        double* cl = xa + (long)c * ne;
        const int* ii = ikle + (long)pi[c % npair] * ne;
        const int* jj = ikle + (long)pj[c % npair] * ne;
        /* om1111.f:559-600: XM = XM * D(j) * D(i); segments D(1) * D(2) */
        if (ndp == 2) {
#pragma omp for schedule(static)
          for (int e = 0; e < ne; ++e) cl[e] = cl[e] * dg[ii[e]] * dg[jj[e]];
        } else {
#pragma omp for schedule(static)
          for (int e = 0; e < ne; ++e) cl[e] = cl[e] * dg[jj[e]] * dg[ii[e]];
        }
      }
#pragma omp single
      td = 'I';
    }
    /* GRACJG */
    const double xb = omp_dot(n, b, b, partial);
    double xll = xb;
    int rel = 1;
    if (sqrt(xb) < 1e-15) {
#pragma omp for schedule(static)
      for (int i = 0; i < n; ++i) x[i] = 0.0;
#pragma omp single
      zero = 1;
    } else {
      if (xll < 1.0) {
        xll = 1.0;
        rel = 0;
      }
      omp_ay(p, r, da, td, xa, typext, ikle, x, col);
#pragma omp for schedule(static)
      for (int i = 0; i < n; ++i) r[i] = r[i] - b[i];
      double rr = omp_dot(n, r, r, partial);
      double rmgm = rr;
      int m = 0, ok = 0;
      if (rr < eps * eps * xll) {
        ok = 1;
      } else {
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) dd[i] = r[i];
        omp_ay(p, ad, da, td, xa, typext, ikle, dd, col);
        double dad = omp_dot(n, dd, ad, partial);
        double ro = rmgm / dad;
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) x[i] = x[i] + (-ro) * dd[i];
        for (;;) {
          m = m + 1;
#pragma omp for schedule(static)
          for (int i = 0; i < n; ++i) r[i] = r[i] + (-ro) * ad[i];
          double rm1gm1 = rmgm;
          rr = omp_dot(n, r, r, partial);
          rmgm = rr;
          if (rr <= xll * eps * eps) {
            ok = 1;
            break;
          }
          double beta = rmgm / rm1gm1;
#pragma omp for schedule(static)
          for (int i = 0; i < n; ++i) dd[i] = r[i] + beta * dd[i];
          omp_ay(p, ad, da, td, xa, typext, ikle, dd, col);
          dad = omp_dot(n, dd, ad, partial);
          ro = rr / dad;
#pragma omp for schedule(static)
          for (int i = 0; i < n; ++i) x[i] = x[i] + (-ro) * dd[i];
          if (!(m < nitmax)) break;
        }
      }
#pragma omp single
      {
        mit = m;
        conv = ok;
        rmrm = rr;
        xl = xll;
        relat = rel;
      }
    }
    if (diag) {
#pragma omp for schedule(static)
      for (int i = 0; i < n; ++i) x[i] = x[i] * dg[i];
    }
  }
  info->niter = mit;
  info->converged = zero ? 1 : conv;
  info->zero_rhs = zero;
  info->relat = relat;
  info->relres = zero ? 0.0 : sqrt(rmrm / xl);
  free(buf);
  return 0;
}
