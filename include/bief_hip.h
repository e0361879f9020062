/*
 * libbiefhip -- C ABI of the MI355X (gfx950) implementation of TELEMAC's
 * BIEF linear-system path: element-by-element (EBE) matrix-vector products,
 * diagonal preconditioning and the CG / BiCGSTAB / GMRES loops behind SOLVE.
 *
 * Plain C types only (no torch, no HIP types in the signatures). Arrays are
 * BIEF's own layouts:
 *   ikle  : int32, structure of arrays, 0-based: ikle[k*nelem + e] is local
 *           node k of element e (BIEF IKLE(NELMAX,NDP) minus one);
 *   da    : float64[npoin], the assembled diagonal (BIEF A%D%R);
 *   xa    : float64[ncoef*nelem], off-diagonal element terms, column c at
 *           xa[c*nelem] (BIEF A%X%R with NELMAX = nelem). Triangles:
 *           symmetric 'S' (12,13,23), non-symmetric 'Q' (12,13,23,21,31,32);
 *           prisms: 15 upper pairs row-major, 'Q' adds the 15 transposed
 *           (sources/utils/bief/matvct.f:152-169, mv0606.f).
 * Vectors are in the caller's node numbering; the library renumbers
 * internally and translates at its boundary.
 *
 * Entry points with a `_dev` suffix take device pointers (HBM-resident data,
 * the caller's numbering) and run on the context's stream; the others take
 * host pointers and are synchronous, as BIEF's Fortran routines are.
 *
 * Every function returns a bhip_status. The library never aborts the process
 * (the reference's PLANTE/STOP, sources/utils/special/plante.F, becomes
 * BHIP_EBADARG / BHIP_EDEVICE, which the Fortran shim turns back into PLANTE).
 */
#ifndef BIEF_HIP_H
#define BIEF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  BHIP_OK = 0,
  BHIP_MAXITER = 1,   /* NITMAX reached (the reference only prints a message) */
  BHIP_EBADARG = 2,   /* invalid argument / unsupported configuration         */
  BHIP_ZERO_RHS = 3,  /* ||B|| < 1e-15: X = 0 (gracjg.F:189-202)              */
  BHIP_EDEVICE = 4    /* HIP runtime error                                    */
} bhip_status;

/* MATVEC operation codes: the OP strings of sources/utils/bief/matvec.f:98-103
 * ('X=AY    ', 'X=CAY   ', ... 'X=X+CTAY'), in this order. */
typedef enum {
  BHIP_OP_AY = 0, BHIP_OP_CAY = 1, BHIP_OP_NAY = 2, BHIP_OP_XPAY = 3,
  BHIP_OP_XMAY = 4, BHIP_OP_XPCAY = 5, BHIP_OP_TAY = 6, BHIP_OP_NTAY = 7,
  BHIP_OP_XPTAY = 8, BHIP_OP_XMTAY = 9, BHIP_OP_XPCTAY = 10
} bhip_op;

/* precision of a solve */
enum {
  BHIP_PREC_FP64 = 0,   /* FP64 everywhere: the reference's arithmetic       */
  BHIP_PREC_MIXED = 1   /* FP32 EBE mat-vec + FP32 inner CG inside FP64
                           iterative refinement; FP64 residual, dots, update */
};

typedef struct bhip_ctx bhip_ctx;

/* Solver configuration: BIEF's SLVCFG (sources/utils/bief/bief_def.f:506-532)
 * plus the precision mode. */
typedef struct {
  int slv;        /* 1 CG (GRACJG), 6 BiCGSTAB (CGSTAB), 7 GMRES            */
  int precon;     /* 0/1 none, 2 diagonal, 5 |diagonal| (products of primes
                     3, 7, 11, 13, 17 are rejected)                          */
  int krylov;     /* GMRES restart dimension, 1..20                          */
  int nitmax;     /* CG/CGSTAB: iterations; GMRES: restart cycles            */
  double eps;     /* stop: ||r'||^2 <= max(||b'||^2,1) eps^2 (CG, CGSTAB);
                     ||r'|| <= max(||b'||,1) eps (GMRES)                    */
  int precision;  /* BHIP_PREC_FP64 or BHIP_PREC_MIXED (CG and GMRES)      */
  double inner_eps;  /* mixed CG: relative reduction per inner FP32 solve    */
  int inner_nitmax;  /* mixed CG: iteration cap of one inner solve (NITMAX
                        caps the inner iterations of the whole solve)        */
} bhip_solver_cfg;

/* Outcome of a solve (BIEF SLVCFG%NIT / OK are never set by the reference;
 * its listing line carries the same numbers, gracjg.F:426-435). */
typedef struct {
  int niter;        /* iterations M (GMRES: restart cycles)                  */
  int inner_iters;  /* mixed CG: total FP32 inner CG iterations; mixed
                       GMRES: FP32 Arnoldi products                          */
  int outer_iters;  /* mixed: FP64 refinement sweeps (FP64 residuals after
                       the first)                                            */
  int relat;        /* 1: relative precision, 0: absolute (||b'|| < 1)       */
  double relres;    /* sqrt(RMRM/XL) on the preconditioned system            */
  int converged;
} bhip_solve_info;

/* ---- mesh context ------------------------------------------------------ */

/* Creates the per-mesh context (replaces what SOLVE/MATVEC read from
 * BIEF_MESH, sources/utils/bief/bief_def.f:205-498: NPOIN, NELEM, IKLE).
 * ndp = 3 (P1 triangles, IELM 11) or 6 (P1 prisms, IELM 41).
 * x, y (may be NULL) are node coordinates used only to order nodes for
 * locality; nplan > 1 marks a prism mesh built by plane-major extrusion of a
 * 2D mesh of npoin/nplan nodes (node p*npoin2 + i), for column-wise blocks.
 * block_nodes: nodes owned per workgroup (0 = default 1024). */
int bhip_ctx_create(int npoin, int nelem, int ndp, const int32_t* ikle,
                    const double* x, const double* y, int nplan,
                    int block_nodes, bhip_ctx** out);
int bhip_ctx_destroy(bhip_ctx* ctx);
/* the HIP stream (hipStream_t) all work of the context runs on */
void* bhip_ctx_stream(bhip_ctx* ctx);
/* layout statistics: nblocks, max colours per block, element instances
 * (duplicated halo elements included), halo node count */
int bhip_ctx_stats(const bhip_ctx* ctx, int* nblocks, int* max_colors,
                   int64_t* elem_instances, int64_t* halo_nodes);

/* Host-only: builds and validates the block layout bhip_ctx_create would use
 * (node order, blocks, halo, per-block colouring with no two elements of a
 * colour sharing an owned node, 16-bit local ids), without touching a GPU.
 * stats[7] (may be NULL): nblocks, max colours, element instances, halo
 * nodes, max local nodes, max owned nodes, LDS bytes per FP64 block.
 * No counterpart in the reference (its vector-mode renumbering check is
 * VECLE3, sources/utils/bief/vecle3.f). */
int bhip_layout_check(int npoin, int nelem, int ndp, const int32_t* ikle,
                      const double* x, const double* y, int nplan,
                      int block_nodes, int64_t* stats);

/* Ensemble of nsys independent meshes in one context (no counterpart in the
 * reference, which solves one mesh per run): every entry point then works on
 * the union of the meshes -- vectors are the meshes' vectors one after the
 * other, xa the meshes' element terms one after the other (nelem = the sum)
 * -- and one launch covers all meshes. bhip_solve_dev runs CG (SLV 1,
 * PRECON 0/2/5), FP64 or mixed, with the scalars, stop tests, refinement
 * sweeps and iteration counts of each mesh its own; bhip_batch_info returns
 * each mesh's outcome. ikle holds the
 * meshes' connectivities one after the other (mesh s: ndp x nelem[s], SoA,
 * its own 0-based node ids); x, y (may be NULL) the coordinates. ndp 3
 * (triangles) or 2 (segment storage). */
int bhip_batch_create(int nsys, const int* npoin, const int* nelem, int ndp,
                      const int32_t* ikle, const double* x, const double* y,
                      int block_nodes, bhip_ctx** out);
int bhip_batch_info(const bhip_ctx* ctx, int nsys, bhip_solve_info* infos);
/* Per-mesh completion time of the last CG solve of an ensemble (ms, device
 * clock): from the solve's set-up kernel to the kernel that stopped that
 * mesh (converged, NITMAX, or a zero right-hand side: 0). */
int bhip_batch_latency(const bhip_ctx* ctx, int nsys, double* ms);

/* ---- matrix ------------------------------------------------------------ */

/* Loads an EBE matrix (replaces the A%D / A%X / A%TYPDIA / A%TYPEXT fields
 * BIEF_OBJ hands to MATVEC/SOLVE, bief_def.f:74-196). typdia 'Q','I','0';
 * typext 'S','Q','0'. The host variant copies from host memory and returns
 * when the copy is done; the _dev variant queues the copy on the context
 * stream, so the caller keeps da_dev / xa_dev valid until that stream has
 * run it (any later synchronising call on the context guarantees it). */
int bhip_matrix_set(bhip_ctx* ctx, char typdia, char typext, const double* da,
                    const double* xa);
int bhip_matrix_set_dev(bhip_ctx* ctx, char typdia, char typext,
                        const double* da_dev, const double* xa_dev);
/* OM 'M=N' between the context's matrix and one saved copy in the same
 * device layout (sources/utils/bief/om.f; telemac2d copies AM1 before SOLVE
 * overwrites it, propag.F:1643): save copies the current matrix aside,
 * restore copies it back (device to device, queued on the context stream). */
int bhip_matrix_save(bhip_ctx* ctx);
int bhip_matrix_restore(bhip_ctx* ctx);
/* Reads the context's (possibly preconditioned) matrix back in BIEF layout;
 * *typdia receives the current TYPDIA. xa must hold ncoef*nelem values, ncoef
 * as bhip_matrix_info reports it. */
int bhip_matrix_get(bhip_ctx* ctx, char* typdia, double* da, double* xa);
/* The context matrix's current TYPDIA, TYPEXT and number of off-diagonal
 * columns (the A%TYPDIA/A%TYPEXT a BIEF caller reads after MATRIX or OM; the
 * element-matrix and OM entry points change them). Any pointer may be NULL.
 * No single counterpart in the reference: BIEF keeps these in BIEF_OBJ. */
int bhip_matrix_info(const bhip_ctx* ctx, char* typdia, char* typext, int* ncoef);

/* ---- MATVEC ------------------------------------------------------------ */

/* X = op(A, Y, C): sources/utils/bief/matvec.f:116-401 (EBE storage, LEGO
 * assembly) -> matvct.f -> mv0303.f / mv0606.f -> assvec.f. */
int bhip_matvec(bhip_ctx* ctx, int op, double c, const double* y, double* x);
int bhip_matvec_dev(bhip_ctx* ctx, int op, double c, const double* y_dev,
                    double* x_dev);

/* ---- SOLVE ------------------------------------------------------------- */

/* SOLVE(X, A, B, TB, CFG, INFOGR, MESH, AUX) for one matrix
 * (sources/utils/bief/solve.f:2-654): diagonal preconditioning PRECDT/PRECD1
 * (precd1.f:90-149), then GRACJG (gracjg.F) / CGSTAB (cgstab.f) /
 * GMRES (gmres.f), then X <- X*D. As in the reference, the context's matrix
 * is left preconditioned (D*A*D, TYPDIA 'I') and b is overwritten by D*b.
 * Returns BHIP_OK, BHIP_MAXITER, BHIP_ZERO_RHS, or an error. */
int bhip_solve(bhip_ctx* ctx, const bhip_solver_cfg* cfg, double* x,
               double* b, bhip_solve_info* info);
int bhip_solve_dev(bhip_ctx* ctx, const bhip_solver_cfg* cfg, double* x_dev,
                   double* b_dev, bhip_solve_info* info);
/* The same SOLVE with device vectors in the context's internal node order
 * (bhip_ctx_perm), for callers that keep their fields in that order: no
 * permutation at the boundary. */
int bhip_solve_dev_internal(bhip_ctx* ctx, const bhip_solver_cfg* cfg,
                            double* x_dev, double* b_dev, bhip_solve_info* info);
/* bhip_matrix_restore followed by bhip_solve_dev, as telemac2d copies the
 * assembled matrix before each SOLVE (OM 'M=N', propag.F:1643) since SOLVE
 * leaves it preconditioned. With diagonal preconditioning the copy is not
 * made: PRECD1 and D*A*D read the saved matrix and write the context's. The
 * results, and the matrix left behind, are those of the two calls. */
int bhip_solve_dev_restored(bhip_ctx* ctx, const bhip_solver_cfg* cfg,
                            double* x_dev, double* b_dev, bhip_solve_info* info);
/* bhip_solve_dev for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. It calls no hipMalloc/hipFree: the partial
 * sums and solver states live in the library's static device scratch (about
 * 2.1 MB per GPU, held until the process ends) instead of the context's
 * buffers; the context's own work vectors (allocated by bhip_ctx_create) are
 * used as they are, and no graph is captured. Limits: FP64 CG (slv 1) only,
 * at most 65,600 partial sums per array (node blocks) and 64 systems;
 * anything beyond returns BHIP_EBADARG with a message. All calls on one GPU
 * share the scratch: calls that overlap in time on one GPU race. Results
 * equal bhip_solve_dev's to the bit. */
int bhip_solve_dev_noalloc(bhip_ctx* ctx, const bhip_solver_cfg* cfg,
                           double* x_dev, double* b_dev, bhip_solve_info* info);
/* The context's node order: perm[i] is the caller's node of internal node
 * i (npoin entries, host memory). */
int bhip_ctx_perm(const bhip_ctx* ctx, int32_t* perm);

/* Measurement: average duration (ms, HIP events on the context stream) of
 * the kernels of one CG iteration on the context's current (preconditioned)
 * matrix. FP64: the fused direction update + EBE mat-vec + D.AD kernel
 * (ms_dir) and the X/R update + R.R kernel (ms_upd). Mixed (after a mixed
 * solve): the same two kernels of the FP32 inner CG. Scratch vectors are
 * reused, so
 * call it after the solves whose results are needed. No counterpart in the
 * reference (it has no timers below the whole run, homere_telemac2d.F). */
int bhip_time_cg_kernels(bhip_ctx* ctx, int precision, int reps,
                         double* ms_dir, double* ms_upd);
/* Measurement: average duration (ms, HIP events on the context stream) of
 * the plain EBE product X = A Y (MV0303/MV0606 + ASSVEC, the kernel of
 * GMRES's Arnoldi steps) on the context's current matrix, FP64 or (after a
 * mixed solve) its FP32 copy. Scratch vectors are overwritten. No
 * counterpart in the reference. */
int bhip_time_matvec(bhip_ctx* ctx, int precision, int reps, double* ms);
/* Measurement: average duration (ms, HIP events on the context stream) of
 * the CG direction + EBE mat-vec kernel (as ms_dir of bhip_time_cg_kernels)
 * when every launch follows a read pass over flush_dev (flush_n
 * doubles of device memory, larger than the last-level cache): the kernel
 * then finds none of its data in any cache, so the figure is its
 * DRAM-bound time. Only the kernel is inside the events. No counterpart in
 * the reference. */
int bhip_time_cg_dir_cold(bhip_ctx* ctx, int precision, int reps, double* flush_dev,
                          long flush_n, double* ms_dir);

/* ---- operand preparation before SOLVE ----------------------------------- */

/* DIRICH with domain-sized FBOR / LIMDIR (sources/utils/bief/dirich.f:83 ->
 * diri01.f:94-139 -> diraux.f): SM -= A*FBOR; at nodes with
 * limdir == kdir: SM = DA*FBOR, F = FBOR; rows and columns of A at those
 * nodes zeroed (diagonal kept). */
int bhip_dirich(bhip_ctx* ctx, double* f, double* sm, const double* fbor,
                const int32_t* limdir, int kdir);

/* LUMP: DIAG = A * (XMUL everywhere) (sources/utils/bief/lump.f) */
int bhip_lump(bhip_ctx* ctx, double xmul, double* diag);
/* The same on device arrays in the caller's order (f, sm updated in place;
 * limdir int32), queued on the context stream: no host copy, no wait. */
int bhip_dirich_dev(bhip_ctx* ctx, double* f_dev, double* sm_dev,
                    const double* fbor_dev, const int32_t* limdir_dev, int kdir);
int bhip_lump_dev(bhip_ctx* ctx, double xmul, double* diag_dev);

/* The mesh's element geometry, kept on the device in the context's element
 * order: XEL, YEL [3*nelem] (node 1 at the origin) and SURFAC [nelem], as
 * INBIEF computes them once per mesh (inbief.f:384-397, for MESH%XEL,
 * MESH%YEL, MESH%SURFAC). Afterwards the element-matrix entries below take
 * xel = yel = surfac = NULL to use it: no per-call upload, and slot-ordered
 * reads in their kernels. Triangles and prisms (the base triangle's
 * values per prism). A later call replaces it. */
int bhip_ctx_set_geometry(bhip_ctx* ctx, const double* xel, const double* yel,
                          const double* surfac);

/* Element matrices on the device, assembled into the context's matrix as
 * BIEF's MATRIX does (matrix.f: element matrix, diagonal by ASSVEC, then
 * OM 'M=N' or 'M=M+N'). accumulate = 0: 'M=N'; 1: 'M=M+N'. xel, yel and
 * surfac all NULL: the geometry of bhip_ctx_set_geometry (here and in the
 * prism and SUPG entries).
 *  - MATMAS, MT01AA (mt01aa.f:83-106): mass, xmul*S/12 (off-diag) and
 *    xmul*S/6 (diagonal); surfac[nelem]
 *  - MATDIFUV, MT02AA 'UV' (mt02aa.f:200-240): diffusion with the P1
 *    product coefficient U*V; xel, yel [3*nelem] local coordinates,
 *    u, v [npoin]
 *  - MT02AA P1 isotropic (mt02aa.f:247-290): u [npoin], v ignored. */
int bhip_matrix_mt01aa(bhip_ctx* ctx, double xmul, const double* surfac,
                       int accumulate);
/* (accumulate = 1 into a non-symmetric 'Q' matrix adds the symmetric
 * element terms to both halves, as OM 'M=M+N' does) */
int bhip_matrix_mt02aa(bhip_ctx* ctx, int kind, double xmul,
                       const double* xel, const double* yel,
                       const double* surfac, const double* u, const double* v,
                       int accumulate);

/* Prism (IELM 41) element matrices on the device, assembled as above
 * (TYPDIA 'Q', TYPEXT 'S', 15 coefficients in the prism pair order).
 *  - MT01PP (mt01pp.f): mass, XMUL*SURFAC/360 combinations of the column
 *    heights z(top) - z(bottom) floored at 1e-3; surfac[nelem] is the area of
 *    each prism's base triangle, z[npoin] the node elevations.
 *  - MT02PP (mt02pp.f): diffusion with the diagonal tensor (f, g, h) [npoin]
 *    of P1 coefficients; xel, yel [3*nelem] the base triangle's local
 *    coordinates (node 1 at the origin); h_zero = 1 cancels the vertical
 *    term (H of type '0', mt02pp.f:165-167); npou0 = NPOIN2 for a P1
 *    vertical viscosity or 0 for P0 on the vertical (mt02pp.f:200-215);
 *    inchyd: the hydrostatic inconsistency test; mon: 'MON' clipping. */
int bhip_matrix_mt01pp(bhip_ctx* ctx, double xmul, const double* surfac,
                       const double* z, int accumulate);
int bhip_matrix_mt02pp(bhip_ctx* ctx, double xmul, const double* xel,
                       const double* yel, const double* surfac, const double* z,
                       const double* f, const double* g, const double* h,
                       int h_zero, int npou0, int inchyd, int mon,
                       int accumulate);

/* The element matrices above from device node fields in the caller's order
 * and the context's geometry (bhip_ctx_set_geometry, required), queued on
 * the context stream: the per-time-step assembly of MATRIX without host
 * copies or waits (matrix.f, mt01aa.f, mt02aa.f, mt01pp.f, mt02pp.f). */
int bhip_matrix_mt01aa_dev(bhip_ctx* ctx, double xmul, int accumulate);
int bhip_matrix_mt02aa_dev(bhip_ctx* ctx, int kind, double xmul,
                           const double* u_dev, const double* v_dev,
                           int accumulate);
int bhip_matrix_mt01pp_dev(bhip_ctx* ctx, double xmul, const double* z_dev,
                           int accumulate);
int bhip_matrix_mt02pp_dev(bhip_ctx* ctx, double xmul, const double* z_dev,
                           const double* f_dev, const double* g_dev,
                           const double* h_dev, int h_zero, int npou0,
                           int inchyd, int mon, int accumulate);

/* SUPG element matrices of CVDFTR (cvdftr.F:398-560) on the device, P1
 * triangles, assembled into the context's matrix as MATRIX does (host
 * inputs in the caller's numbering; xel, yel [3*nelem], surfac [nelem]):
 *  kind 3: MASUPG, MT03AA (mt03aa.f), f, g per element, u, v per node; 'Q'
 *  kind 4: MAUGUG, MT04AA (mt04aa.f), u, v per node; symmetric
 *  kind 5: MATVGR, MT05AA (mt05aa.f), u, v per node, or per element with
 *          BHIP_SUPG_U_P0; BHIP_SUPG_UPWIND selects FORMUL(16:16) = 'N'; 'Q'
 * BHIP_SUPG_KSUPG: the element vectors (f, g of kind 3, the per-element
 * u, v of kind 5) are KSUPG(1, u, v) of the nodal u, v given (ksupg.f ->
 * kspg11.f), computed on the device. op: 0 'M=N', 1 'M=M+N', 2 'M=M+TN'
 * (OM, om1111.f); a symmetric target becomes 'Q' first when needed. */
#define BHIP_SUPG_U_P0 1
#define BHIP_SUPG_UPWIND 2
#define BHIP_SUPG_KSUPG 4
int bhip_matrix_supg(bhip_ctx* ctx, int kind, double xmul, const double* xel,
                     const double* yel, const double* surfac, const double* f,
                     const double* g, const double* u, const double* v, int flags,
                     int op);
/* The same from device arrays (f, g per element in the caller's element
 * order; u, v per node in the caller's order, or per element with
 * BHIP_SUPG_U_P0) and the context's geometry, queued on the context stream. */
int bhip_matrix_supg_dev(bhip_ctx* ctx, int kind, double xmul,
                         const double* f_dev, const double* g_dev,
                         const double* u_dev, const double* v_dev, int flags,
                         int op);
/* OM on the context's matrix M, with N its saved copy (bhip_matrix_save)
 * (sources/utils/bief/om.f -> om1111.f): "M=X(M)" (the non-symmetric form
 * of M), "M=CN" (M = c M), "M=M+CN", "M=M+N", "M=M+TN" (N transposed). */
int bhip_matrix_om(bhip_ctx* ctx, const char* op, double c);

/* ---- the time-stepped wave equation and its adjoint ---------------------- *
 * One time step of telemac2d's wave-equation treatment (PROPAG with
 * TREATMENT OF THE LINEAR SYSTEM = 2, propag.F:664-716, 1003-1012,
 * 1150-1170, 1369-1465, 1634-1710) for implicitation 1, no advection or
 * diffusion of velocity, no sources, no tidal flats, Strickler friction
 * (friction_unif.f:124-127, friction_calc.f:130-134, fricti.f:123-135, with
 * the bottom-slope factor of slop10.f), the free surface prescribed on sea
 * nodes (free velocity there) and walls elsewhere, and the transposed kernels
 * of its hand-written adjoint. No single routine of the reference
 * corresponds: PROPAG builds the same operands from MATMAS (MT01AA),
 * MATDIFUV (MT02AA 'UV'), MATFGR (MT08AA), VECTOR 'GRADF'/'MASBAS', OS and
 * DIRICH calls. All node vectors are device pointers in the context's
 * internal order (bhip_ctx_perm); triangles only. */
typedef struct {
  double dt;     /* time step (s) */
  double grav;   /* gravity */
  double eps_u;  /* speed floor in |u| = sqrt(u^2 + v^2 + eps_u^2); the
                    reference's |u| is eps_u = 0 */
  double tide;   /* free surface prescribed on sea nodes at the new time */
} bhip_wave_params;

/* Slot geometry (xel, yel [3*nelem] local coordinates, surfac [nelem], caller
 * element order), basis integrals, and the sea nodes (sea [npoin], caller
 * order, nonzero = prescribed). */
int bhip_wave_setup(bhip_ctx* ctx, const double* xel, const double* yel,
                    const double* surfac, const int32_t* sea);
/* The system of the step from the state (eta, u, v), Strickler kstr and
 * bottom zb: AM1 = M/dt + MATDIFUV(-1, DM1, max(h, 0)) into the context's
 * matrix (TYPDIA 'Q', TYPEXT 'S', rows and columns of sea nodes cleared off
 * the diagonal) and its right-hand side rhs = CV1 = -BM1 T2 - BM2 T3 with
 * DIRICH's terms (D fb on sea nodes, fb = tide - eta). */
int bhip_wave_assemble(bhip_ctx* ctx, const bhip_wave_params* p,
                       const double* eta, const double* u, const double* v,
                       const double* kstr, const double* zb, double* rhs);
/* The new state from the increment deta = A^-1 rhs: eta' = eta + deta and
 * the momentum update u' = (GRADF(-g eta') + VOL/dt u) / AM2D. Uses the
 * friction of the last bhip_wave_assemble. */
int bhip_wave_momentum(bhip_ctx* ctx, const bhip_wave_params* p,
                       const double* eta, const double* deta, const double* u,
                       const double* v, double* eta_n, double* u_n,
                       double* v_n);
/* Adjoint of the momentum update with respect to eta' (after
 * bhip_wave_assemble of the same step): etot = ebar_n + grad^T(...), the
 * right-hand side of the adjoint solve. */
int bhip_wave_momentum_adj(bhip_ctx* ctx, const bhip_wave_params* p,
                           const double* ubar_n, const double* vbar_n,
                           const double* ebar_n, double* etot);
/* Adjoint of the system, its right-hand side and the node quantities, given
 * lam = A^-1 etot: writes the adjoint state of the step's start (ebar, ubar,
 * vbar) and adds to the adjoints of kstr and zb (kbar, zbbar). */
int bhip_wave_assemble_adj(bhip_ctx* ctx, const bhip_wave_params* p,
                           const double* lam, const double* deta,
                           const double* eta, const double* u, const double* v,
                           const double* u_n, const double* v_n,
                           const double* ubar_n, const double* vbar_n,
                           const double* etot, const double* kstr, double* ebar,
                           double* ubar, double* vbar, double* kbar,
                           double* zbbar);

#ifdef __cplusplus
}
#endif
#endif /* BIEF_HIP_H */
