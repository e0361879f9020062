/*
 * CPU restatement (FP64, sequential) of TELEMAC's BIEF linear-system path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it. The product (libbiefhip) never links or calls it.
 *
 * Conventions (shared with the reference's BIEF storage):
 *  - connectivity ikle is structure-of-arrays, 0-based: ikle[k*nelmax + e]
 *    is local node k of element e (BIEF's IKLE(NELMAX,NDP), minus one);
 *  - off-diagonal coefficients xa are SoA: xa[c*nelmax + e];
 *    triangles, symmetric 'S': c = (12,13,23);
 *    triangles, non-symmetric 'Q': c = (12,13,23,21,31,32)
 *    (sources/utils/bief/matvct.f:152-169);
 *    prisms, symmetric: 15 upper-triangle pairs in row-major order,
 *    non-symmetric: those 15 then the 15 transposed ones
 *    (sources/utils/bief/mv0606.f, A1-matrix_storage.tex Prism).
 *  - typdia in {'Q','I','0'}, typext in {'Q','S','0'}.
 */
#ifndef BIEF_ORACLE_H
#define BIEF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* MATVEC operation codes (sources/utils/bief/matvec.f:349-380, mv0303.f:25-35) */
enum {
  ORC_AY = 0,     /* X = A Y       */
  ORC_CAY = 1,    /* X = C A Y     */
  ORC_NAY = 2,    /* X = -A Y      */
  ORC_XPAY = 3,   /* X = X + A Y   */
  ORC_XMAY = 4,   /* X = X - A Y   */
  ORC_XPCAY = 5,  /* X = X + C A Y */
  ORC_TAY = 6,    /* X = A^T Y     */
  ORC_NTAY = 7,   /* X = -A^T Y    */
  ORC_XPTAY = 8,  /* X = X + A^T Y */
  ORC_XMTAY = 9,  /* X = X - A^T Y */
  ORC_XPCTAY = 10 /* X = X + C A^T Y */
};

double orc_dot(int n, const double* x, const double* y);

/* element-wise products W (ndp columns of nelmax) plus the diagonal term in x,
 * then the assembly of W onto x in BIEF's order. Elements 11 (triangle) or
 * 41 (prism). transposed != 0 selects A^T. */
void orc_matvct(int transposed, double* x, const double* da, char typdia,
                const double* xa, char typext, const double* y,
                const int* ikle, int npoin, int nelem, int nelmax, int ndp,
                double* w);

/* BIEF MATVEC: T = A Y (or A^T Y) into t, then X combined with T per op. */
void orc_matvec(int op, double* x, const double* da, char typdia,
                const double* xa, char typext, const double* y, double c,
                const int* ikle, int npoin, int nelem, int nelmax, int ndp,
                double* w, double* t);

/* ASSVEC (INIT=.FALSE.): x[ikle(e,k)] += w(e,k), k outer, e inner. */
void orc_assvec(double* x, const int* ikle, int nelem, int nelmax, int ndp,
                const double* w);

/* element matrices: diagonal terms in t (ndp columns), off-diagonal in xm */
void orc_mt01aa(double* t, double* xm, double xmul, const double* surfac,
                int nelem, int nelmax);
/* kind: 0 = P0 isotropic (U per element), 1 = P1 isotropic (U per node),
 *       2 = 'UV' product (U,V per node), 3 = P0 tensor (3 columns per element),
 *       4 = P1 tensor (3 columns per node, column stride ldu).
 * mon != 0 applies the 'MON' clipping. */
void orc_mt02aa(int kind, double* t, double* xm, double xmul,
                const double* u, const double* v, int ldu,
                const double* xel, const double* yel, const double* surfac,
                const int* ikle, int nelem, int nelmax, int mon);
/* prism mass matrix (MT01PP): t is T(NELMAX,6), xm the 15 symmetric pairs */
void orc_mt01pp(double* t, double* xm, double xmul, const double* surfac,
                const double* z, const int* ikle, int nelem, int nelmax);

/* prism diffusion matrix (MT02PP): diagonal tensor diffusion with P1
 * coefficients f (x), g (y), h (z) on the prism nodes; h_zero cancels the
 * vertical term (H of type '0'); npou0 = NPOIN2 (P1 vertical viscosity) or 0
 * (P0 on the vertical); xel, yel: the base triangle's local coordinates
 * (3 columns of nelmax, node 1 at the origin); z: node elevations; inchyd:
 * hydrostatic inconsistency test; mon: 'MON' clipping. */
void orc_mt02pp(double* t, double* xm, double xmul, const double* f, const double* g,
                const double* h, int h_zero, int npou0, const double* xel,
                const double* yel, const double* z, const double* surfac, const int* ikle,
                int nelem, int nelmax, int inchyd, int mon);

/* SUPG element routines (bief_oracle_supg.c), P1 triangles; xm in the 'Q'
 * order (12,13,23,21,31,32) except MT04AA ('S', 12,13,23).
 *  KSPG11 (ksupg.f -> kspg11.f): SUPG element vectors kx, ky (per element)
 *  MT03AA (MASUPG): f, g per element, u, v per node
 *  MT04AA (MAUGUG): u, v per node, symmetric
 *  MT05AA (MATVGR): u, v per element (u_p0) or per node; upwind: the 'N'
 *  variant of FORMUL(16:16) */
void orc_kspg11(double* kx, double* ky, const double* xel, const double* yel,
                const double* u, const double* v, const int* ikle, int nelem, int nelmax,
                double xmul);
void orc_mt03aa(double* t, double* xm, double xmul, const double* f, const double* g,
                const double* u, const double* v, const double* xel, const double* yel,
                const double* surfac, const int* ikle, int nelem, int nelmax);
void orc_mt04aa(double* t, double* xm, double xmul, const double* u, const double* v,
                const double* xel, const double* yel, const double* surfac, const int* ikle,
                int nelem, int nelmax);
void orc_mt05aa(double* t, double* xm, double xmul, const double* u, const double* v,
                int u_p0, int upwind, const double* xel, const double* yel, const int* ikle,
                int nelem, int nelmax);

/* assembles element diagonal terms t (ndp columns) into da (init to zero) */
void orc_assemble_diag(double* da, const double* t, const int* ikle,
                       int npoin, int nelem, int nelmax, int ndp);

/* OM 'M=DMD': A <- D A D. *typdia updated as in om1111.f. */
void orc_om_dmd(double* da, char* typdia, double* xa, char typext,
                const double* d, const int* ikle, int npoin, int nelem,
                int nelmax, int ndp);

/* PRECD1: diagonal preconditioning with PRECON 2 or 5 (precd1.f:90-149).
 * d receives the preconditioning diagonal (1/sqrt|diag|). */
void orc_precd1(double* x, double* da, char* typdia, double* xa, char typext,
                double* b, double* d, const int* ikle, int npoin, int nelem,
                int nelmax, int ndp, int precon);

typedef struct {
  int niter;      /* GRACJG/CGSTAB: iterations M; GMRES: restart cycles M */
  double relres;  /* sqrt(RMRM/XL) (CG, CGSTAB) or ||r||/NB (GMRES) */
  int relat;      /* 1 relative precision, 0 absolute (norm of B < 1) */
  int converged;  /* 1 if the stop test was met */
  int zero_rhs;   /* 1 if X=0 because the RHS norm was < 1e-15 */
} orc_info;

/* Krylov loops on the (already preconditioned) system. work must hold
 * ndp*nelmax (w) + npoin (t) + nvec*npoin doubles:
 * CG 3 vectors, CGSTAB 6, GMRES 2*krylov + 1. */
void orc_gracjg(double* x, const double* da, char typdia, const double* xa,
                char typext, const double* b, const int* ikle, int npoin,
                int nelem, int nelmax, int ndp, double eps, int nitmax,
                double* work, orc_info* info);
void orc_cgstab(double* x, const double* da, char typdia, const double* xa,
                char typext, const double* b, const int* ikle, int npoin,
                int nelem, int nelmax, int ndp, double eps, int nitmax,
                double* work, orc_info* info);
void orc_gmres(double* x, const double* da, char typdia, const double* xa,
               char typext, const double* b, const int* ikle, int npoin,
               int nelem, int nelmax, int ndp, double eps, int nitmax,
               int krylov, double* work, orc_info* info);

/* SOLVE for a single matrix (S=0): preconditioning (PRECON 0/1, 2, 5),
 * Krylov loop (slv 1 CG, 6 CGSTAB, 7 GMRES), undo of the variable change.
 * Overwrites A (da, xa, *typdia) and b exactly as the reference does.
 * Returns 0 on success, 2 on unsupported arguments. */
int orc_solve(int slv, int precon, int krylov, int nitmax, double eps,
              double* x, double* da, char* typdia, double* xa, char typext,
              double* b, const int* ikle, int npoin, int nelem, int nelmax,
              int ndp, orc_info* info);
/* doubles of work orc_solve allocates internally (for callers that size) */
long orc_solve_work_size(int slv, int krylov, int npoin, int nelmax, int ndp);

/* DIRICH -> DIRI01 -> DIRAUX with domain-sized FBOR/LIMDIR
 * (dirich.f:83, diri01.f:94-139, diraux.f): f receives the prescribed
 * values, sm is corrected, A's rows/columns at prescribed nodes zeroed. */
void orc_dirich(double* f, double* da, char typdia, double* xa, char typext,
                double* sm, const double* fbor, const int* limdir, int kdir,
                const int* ikle, int npoin, int nelem, int nelmax, int ndp);

/* LUMP: diag = A * (xmul everywhere) (lump.f) */
void orc_lump(double* diag, const double* da, char typdia, const double* xa,
              char typext, double xmul, const int* ikle, int npoin, int nelem,
              int nelmax, int ndp);

/* Multi-threaded SOLVE (CG, PRECON 1 or 2) for the multi-core CPU baseline
 * (bief_oracle_omp.c, built into liboracle_omp.so): a plan holds the
 * node-to-element incidences in ASSVEC's order of addition. */
void* orc_omp_plan_create(const int* ikle, int npoin, int nelem, int ndp);
void orc_omp_plan_free(void* plan);
int orc_omp_solve_cg(void* plan, const int* ikle, double* da, char typdia, double* xa,
                     char typext, double* b, double* x, double eps, int nitmax, int precon,
                     int nthreads, orc_info* info);

#ifdef __cplusplus
}
#endif
#endif
