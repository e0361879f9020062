// BiCGSTAB (CGSTAB) and restarted GMRES on the device-resident,
// preconditioned system (FP64), included by bief_hip.hip.
//
// Both follow the reference loops operation by operation (same scalar
// recurrences, same tests and exits). Reductions use the same fixed-order
// per-workgroup partials as CG, reduced again by every consumer workgroup,
// so no kernel synchronises workgroups internally.

namespace bhip {

// ---------------------------------------------------------------------------
// CGSTAB (sources/utils/bief/cgstab.f:113-236)
// ---------------------------------------------------------------------------
struct BicgScal {
  double beta[2];  // BETA of iteration m, by parity
  double beta1;    // BETA1 of the current iteration (written by kernel A)
  double omeg1;    // OMEG1 (kernel B), read by the next iteration's kernel A
  double alfa;     // ALFA (kernel C), read by the next iteration's kernel A
  double omeg, omeg2;
};

// elementwise pieces, each also the consumer of the previous reduction
// A: termination tests of the previous iteration, then
//    BETA1 = R.P ; OMEG2 = OMEG1 BETA1/BETA ; OMEG = OMEG2/ALFA ; BETA = BETA1
//    Q = R + OMEG Q ; Q = Q - OMEG2 V        (cgstab.f:172-181)
__global__ void k_bicg_a(double* __restrict__ q, const double* __restrict__ r,
                         const double* __restrict__ v, int n,
                         const double* __restrict__ part_rr, const double* __restrict__ part_rp,
                         int npart, SolveState* st, BicgScal* sc, int parity) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  if (load_relaxed(&st->done)) return;
  const int p = parity, q1 = 1 - p;
  const double rr = reduce_partials(part_rr, npart, sh_red, &sh_b);
  const double beta1 = reduce_partials(part_rp, npart, sh_red, &sh_b);
  const int mprev = st->m[q1];
  if (mprev >= 1) {
    // tests after iteration mprev (cgstab.f:229-232)
    if (rr <= st->eps2xl || mprev >= st->nitmax) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->final_m = mprev;
        st->final_rr = rr;
        store_relaxed(&st->done, rr <= st->eps2xl ? ST_CONVERGED : ST_MAXITER);
      }
      return;
    }
  }
  const double omeg2 = sc->omeg1 * beta1 / sc->beta[q1];
  const double omeg = omeg2 / sc->alfa;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double t = r[i] + omeg * q[i];
    q[i] = t + (-omeg2) * v[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->beta[p] = beta1;
    sc->beta1 = beta1;
    sc->omeg = omeg;
    sc->omeg2 = omeg2;
    st->m[p] = mprev + 1;
    st->rr[p] = rr;
  }
}

// B: OMEG1 = BETA1 / P.V ; S = R - OMEG1 V  (cgstab.f:190-197)
__global__ void k_bicg_b(double* __restrict__ s, const double* __restrict__ r,
                         const double* __restrict__ v, int n, const double* __restrict__ part_pv,
                         int npart, SolveState* st, BicgScal* sc) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  if (load_relaxed(&st->done)) return;
  const double pv = reduce_partials(part_pv, npart, sh_red, &sh_b);
  const double omeg1 = sc->beta1 / pv;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s[i] = r[i] + (-omeg1) * v[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->omeg1 = omeg1;
}

// C: ALFA = T.S / T.T ; X = X + OMEG1 Q ; X = X + ALFA S ; R = S - ALFA T ;
//    partial sums of R.R and R.P for the next iteration (cgstab.f:207-227)
__global__ void k_bicg_c(double* __restrict__ x, double* __restrict__ r,
                         const double* __restrict__ q, const double* __restrict__ s,
                         const double* __restrict__ t, const double* __restrict__ pv, int n,
                         const double* __restrict__ part_ts, const double* __restrict__ part_tt,
                         int npart_in, double* __restrict__ out_rr, double* __restrict__ out_rp,
                         SolveState* st, BicgScal* sc) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  if (load_relaxed(&st->done)) return;
  const double ts = reduce_partials(part_ts, npart_in, sh_red, &sh_b);
  const double tt = reduce_partials(part_tt, npart_in, sh_red, &sh_b);
  const double alfa = ts / tt;
  const double omeg1 = sc->omeg1;
  double prr = 0.0, prp = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double xv = x[i] + omeg1 * q[i];
    x[i] = xv + alfa * s[i];
    const double rv = s[i] + (-alfa) * t[i];
    r[i] = rv;
    prr += rv * rv;
    prp += rv * pv[i];
  }
  const double a1 = block_sum(prr, sh_red);
  const double a2 = block_sum(prp, sh_red);
  if (threadIdx.x == 0) {
    out_rr[blockIdx.x] = a1;
    out_rp[blockIdx.x] = a2;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alfa = alfa;
}

// set-up after V = A X: R = B - V ; RMRM ; P = R ; Q = V = 0 (cgstab.f:132-160)
__global__ void k_bicg_init(double* __restrict__ r, double* __restrict__ p,
                            double* __restrict__ q, double* __restrict__ v,
                            const double* __restrict__ b, int n, double* __restrict__ out_rr,
                            double* __restrict__ out_rp) {
  __shared__ double sh_red[kWaves];
  double prr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double rv = b[i] - v[i];
    r[i] = rv;
    p[i] = rv;
    q[i] = 0.0;
    v[i] = 0.0;
    prr += rv * rv;
  }
  const double a1 = block_sum(prr, sh_red);
  if (threadIdx.x == 0) {
    out_rr[blockIdx.x] = a1;
    out_rp[blockIdx.x] = a1;  // P = R: R.P = R.R
  }
}

// initial test RMRM < EPS^2 XL (cgstab.f:137); state of the loop
__global__ void k_bicg_start(const double* __restrict__ part_rr, int npart, double eps,
                             int nitmax, SolveState* st, BicgScal* sc) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  const double rr = reduce_partials(part_rr, npart, sh_red, &sh_b);
  if (threadIdx.x != 0) return;
  const double xl = st->xl;
  st->eps2xl = eps * eps * xl;
  st->nitmax = nitmax < 1 ? 1 : nitmax;
  st->m[0] = st->m[1] = 0;
  st->final_flag = 0;
  sc->alfa = 1.0;
  sc->beta[0] = sc->beta[1] = 1.0;
  sc->omeg1 = 1.0;
  if (rr < eps * eps * xl) {
    st->final_m = 0;
    st->final_rr = rr;
    st->done = ST_CONVERGED;
  } else {
    st->done = 0;
  }
}

static int cgstab_fp64(bhip_ctx* c, double* x, const double* b, double eps, int nitmax,
                       bhip_solve_info* info) {
  const Layout& L = c->L;
  const int n = L.npoin;
  const int g = c->grid_nodes;
  SolveState* sh = c->st_host;
  if (c->bicg_sc.alloc(16)) return BHIP_EDEVICE;
  BicgScal* sc = reinterpret_cast<BicgScal*>(c->bicg_sc.p);
  double *R = c->vr.p, *P = c->vd0.p, *Q = c->vd1.p, *S = c->vad.p, *T = c->vio0.p,
         *V = c->vio1.p;
  // XL = B.B, clamped to 1 (cgstab.f:113-120); the reference tests the zero
  // right-hand side only after the clamp, so it never triggers
  HIP_OK(enqueue_bb_setup(c, b, eps, nitmax, 0, 0.0, 0));
  {
    EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
    a.y = x;
    a.out = V;
    a.op = BHIP_OP_AY;
    HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_OP, FIN_NONE>(c, a, false, c->typdia,
                                                                       c->typext)));
  }
  double* prr[2] = {part_u(c, 0), part_u(c, 1)};
  double* prp[2] = {part_e(c, 0), part_e(c, 1)};
  hipLaunchKernelGGL(k_bicg_init, dim3(g), dim3(kThreads), 0, c->stream, R, P, Q, V, b, n,
                     prr[0], prp[0]);
  hipLaunchKernelGGL(k_bicg_start, dim3(1), dim3(kThreads), 0, c->stream, prr[0], g, eps,
                     nitmax, c->st_p, sc);
  HIP_OK(hipGetLastError());
  // the two EBE kernels of an iteration write their partials past the
  // four node-kernel arrays
  DevBuf<double>& ex = c->bicg_part;
  if (ex.alloc((size_t)3 * c->max_part)) return BHIP_EDEVICE;
  double* ppv = ex.p;
  double* pts = ex.p + c->max_part;
  double* ptt = ex.p + 2 * (size_t)c->max_part;
  int it = 0;
  auto enqueue_iter = [&]() -> int {
    const int p = (it & 1) ? 0 : 1;  // iteration it+1
    hipLaunchKernelGGL(k_bicg_a, dim3(g), dim3(kThreads), 0, c->stream, Q, R, V, n, prr[1 - p],
                       prp[1 - p], g, c->st_p, sc, p);
    HIP_OK(hipGetLastError());
    {
      EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
      a.y = Q;
      a.out = V;
      a.w = P;
      a.partial = ppv;
      a.skip_if_done = 1;
      HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_DOTW, FIN_PARTIAL>(
          c, a, false, c->typdia, c->typext)));
    }
    hipLaunchKernelGGL(k_bicg_b, dim3(g), dim3(kThreads), 0, c->stream, S, R, V, n, ppv,
                       L.nblocks, c->st_p, sc);
    HIP_OK(hipGetLastError());
    {
      EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
      a.y = S;
      a.out = T;
      a.partial = pts;
      a.partial2 = ptt;
      a.skip_if_done = 1;
      HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_DOT2, FIN_PARTIAL>(
          c, a, false, c->typdia, c->typext)));
    }
    hipLaunchKernelGGL(k_bicg_c, dim3(g), dim3(kThreads), 0, c->stream, x, R, Q, S, T, P, n,
                       pts, ptt, L.nblocks, prr[p], prp[p], c->st_p, sc);
    HIP_OK(hipGetLastError());
    ++it;
    return BHIP_OK;
  };
  int& hint = c->iter_hint[3];
  int nq = std::max(hint + 1, 4);
  for (;;) {
    for (int k = 0; k < nq; ++k)
      if (enqueue_iter()) return BHIP_EDEVICE;
    if (read_state(c)) return BHIP_EDEVICE;
    if (sh->done) break;
    nq = std::min(std::max(nq, 8), 64);
  }
  hint = sh->final_m;
  info->relat = sh->xl_raw >= 1.0;
  info->niter = sh->final_m;
  info->relres = std::sqrt(sh->final_rr / sh->xl);
  info->converged = sh->done == ST_CONVERGED;
  return sh->done == ST_CONVERGED ? BHIP_OK : BHIP_MAXITER;
}

// ---------------------------------------------------------------------------
// GMRES (sources/utils/bief/gmres.f:171-372), restarted with K <= 20
// ---------------------------------------------------------------------------

// x = c * y
__global__ void k_scale_into(double* __restrict__ x, const double* __restrict__ y, double cst,
                             int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] = cst * y[i];
}

// x = c * y in the basis precision (the mixed GMRES's V1 from the FP64 residual)
__global__ void k_scale_into_f32(float* __restrict__ x, const double* __restrict__ y, double cst,
                                 int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] = (float)(cst * y[i]);
}

// V <- V / sqrt(nrm2[0]) unless the norm is below the Krylov-reduction
// threshold of gmres.f:274 (the host then sees the breakdown in the
// Hessenberg column and ignores what follows)
template <typename T>
__global__ void k_scale_inv_norm(T* __restrict__ v, const double* __restrict__ nrm2, int n) {
  const double r = sqrt(nrm2[0]);
  if (!(r >= 1e-20)) return;
  const double s = 1.0 / r;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v[i] = (T)(s * (double)v[i]);
}

// One modified Gram-Schmidt step: H(i,j) = reduced partials of w.V_i;
// w = src - H(i,j) V_i (gmres.f:262-265); partial sums of w.next (next = the
// next basis vector, or w itself for the norm). w and src may be the same
// buffer (in-place steps after the first): each element is read before it is
// written at the same index, so neither is declared __restrict__.
// The basis precision T is FP64, or FP32 in the mixed GMRES (H and the
// dots stay FP64).
template <typename T>
__global__ void k_mgs(T* w, const T* src, const T* __restrict__ vi, const T* __restrict__ next,
                      int n, const double* __restrict__ part_in, int npart_in,
                      double* __restrict__ part_out, double* __restrict__ hij) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  const double h = reduce_partials(part_in, npart_in, sh_red, &sh_b);
  const T ht = (T)h;
  double acc = 0.0;
  const bool self = (next == nullptr);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T wv = src[i] + (-ht) * vi[i];
    w[i] = wv;
    acc += (double)wv * (double)(self ? wv : next[i]);
  }
  const double s = block_sum(acc, sh_red);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) *hij = h;
}

// partial sums of av.v_i for i < k and of av.av (last column, gmres.f:285-291)
template <int KMAX, typename T>
__global__ void k_multidot(const T* __restrict__ av, const T* __restrict__ vbase, int k, int n,
                           double* __restrict__ part, int stride) {
  __shared__ double sh_red[kWaves];
  double acc[KMAX + 1];
#pragma unroll
  for (int q = 0; q <= KMAX; ++q) acc[q] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double a = (double)av[i];
    acc[0] += a * a;
#pragma unroll
    for (int q = 0; q < KMAX; ++q)
      if (q < k) acc[q + 1] += a * (double)vbase[(size_t)q * n + i];
  }
#pragma unroll
  for (int q = 0; q <= KMAX; ++q) {
    if (q > k) break;
    const double s = block_sum(acc[q], sh_red);
    if (threadIdx.x == 0) part[(size_t)q * stride + blockIdx.x] = s;
  }
}

// reduce m partial arrays (stride apart) into dst[0..m)
__global__ void k_reduce_many(const double* __restrict__ part, int npart, int stride, int m,
                              double* __restrict__ dst) {
  __shared__ double sh_red[kWaves];
  __shared__ double sh_b;
  for (int q = 0; q < m; ++q) {
    const double s = reduce_partials(part + (size_t)q * stride, npart, sh_red, &sh_b);
    if (threadIdx.x == 0) dst[q] = s;
  }
}

// X = X + E(L) V(L) ; R0 = R0 + E(L) AV(L), L = 1..K in order (gmres.f:354-359)
struct Coef20 {
  double e[20];
};
__global__ void k_gmres_update(double* __restrict__ x, double* __restrict__ r0,
                               const double* __restrict__ vbase,
                               const double* __restrict__ avbase, Coef20 e, int k, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double xv = x[i], rv = r0[i];
    for (int l = 0; l < k; ++l) {
      xv = xv + e.e[l] * vbase[(size_t)l * n + i];
      rv = rv + e.e[l] * avbase[(size_t)l * n + i];
    }
    x[i] = xv;
    r0[i] = rv;
  }
}
// the mixed GMRES's correction: X = X + E(L) V(L) in FP64 with the FP32
// basis (the residual is then recomputed in FP64, not updated)
__global__ void k_gmres_xupd(double* __restrict__ x, const float* __restrict__ vbase, Coef20 e,
                             int k, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double xv = x[i];
    for (int l = 0; l < k; ++l) xv = xv + e.e[l] * (double)vbase[(size_t)l * n + i];
    x[i] = xv;
  }
}

// Restarted GMRES(K), gmres.f:171-372. T is the precision of the Krylov
// basis and of the Arnoldi products:
//  * double: the reference's arithmetic (R0 updated with the AV vectors,
//    gmres.f:354-359, the stop on |E(K+1)|);
//  * float (mixed): the Arnoldi products use the FP32 copy of the matrix
//    and the basis V, AV is FP32 (Hessenberg entries and dots FP64); every
//    restart is an FP64 refinement step -- X += V y in FP64, then the true
//    residual R0 = A X - B by the FP64 product -- and the stop rule
//    ||R0|| / max(||B||, 1) < EPS is tested on that FP64 residual.
template <typename T>
static int gmres_impl(bhip_ctx* c, double* x, const double* b, double eps, int nitmax,
                      int krylov, bool xa32_ready, bhip_solve_info* info) {
  constexpr bool kMixed = sizeof(T) == 4;
  const Layout& L = c->L;
  const int n = L.npoin;
  const int g = c->grid_nodes;
  const int K0 = krylov;
  T* Vb;
  if constexpr (kMixed) {
    if (ensure_f32(c)) return BHIP_EDEVICE;
    if (c->gm32.alloc((size_t)2 * K0 * n)) return BHIP_EDEVICE;
    Vb = c->gm32.p;
    const int npair = L.ndp * (L.ndp - 1) / 2;
    const long ncf = (long)(c->typext == 'Q' ? 2 * npair : npair) * L.nslot_pad;
    if (!xa32_ready)
      hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(kThreads), 0, c->stream, c->xa32.p, c->xa.p,
                         ncf);
    hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kThreads), 0, c->stream, c->da32.p, c->da.p,
                       (long)n);
    HIP_OK(hipGetLastError());
  } else {
    (void)xa32_ready;
    if (c->gm.alloc((size_t)2 * K0 * n)) return BHIP_EDEVICE;
    Vb = c->gm.p;
  }
  // device scalars: ||B||^2, ||R0||^2, then one Hessenberg column per
  // Arnoldi step (H(0..j, j) and ||w||^2), then the last column
  const int colw = K0 + 1;
  auto col = [&](int j) { return 2 + j * colw; };
  const int lastoff = col(K0 - 1);
  if (c->gm_scal.alloc((size_t)lastoff + K0 + 2)) return BHIP_EDEVICE;
  if (c->bicg_part.alloc((size_t)(K0 + 2) * c->max_part)) return BHIP_EDEVICE;
  T* AVb = Vb + (size_t)K0 * n;
  auto Vp = [&](int j) { return Vb + (size_t)j * n; };
  auto AVp = [&](int j) { return AVb + (size_t)j * n; };
  double* R0 = c->vr.p;
  double* ds = c->gm_scal.p;  // device scalars
  double* mp = c->bicg_part.p;
  const int stride = c->max_part;
  auto fetch = [&](double* h, int cnt) -> int {
    HIP_OK(hipMemcpyAsync(h, ds, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return BHIP_OK;
  };
  std::vector<double> hsv((size_t)lastoff + K0 + 2);
  double* hs = hsv.data();
  // NB = ||B|| clamped to 1 (gmres.f:188-194)
  hipLaunchKernelGGL(k_dot_partials<double>, dim3(g), dim3(kThreads), 0, c->stream, b, b, n,
                     part_u(c, 0));
  hipLaunchKernelGGL(k_reduce_into, dim3(1), dim3(kThreads), 0, c->stream, part_u(c, 0), g,
                     ds);
  // R0 = A X - B ; ||R0||^2 (gmres.f:198-207), FP64
  auto residual = [&]() -> int {
    EbeArgs<double, double> a = base_args<double, double>(c, c->xa.p, c->da.p);
    a.y = x;
    a.b = b;
    a.out = R0;
    a.partial = part_u(c, 1);
    HIP_OK((launch_ebe<double, double, STAGE_PLAIN, EPI_RESID, FIN_PARTIAL>(
        c, a, false, c->typdia, c->typext)));
    hipLaunchKernelGGL(k_reduce_into, dim3(1), dim3(kThreads), 0, c->stream, part_u(c, 1),
                       L.nblocks, ds + 1);
    HIP_OK(hipGetLastError());
    return BHIP_OK;
  };
  if (residual()) return BHIP_EDEVICE;
  if (fetch(hs, 2)) return BHIP_EDEVICE;
  double nb = std::sqrt(hs[0]);
  info->relat = 1;
  if (nb < 1.0) {
    nb = 1.0;
    info->relat = 0;
  }
  double nr0 = std::sqrt(hs[1]);
  double prec = nr0 / nb;
  int m = 0, products = 0;
  double H[21][20];
  double cg[20], sg[20], e[21];
  std::memset(H, 0, sizeof(H));
  // EBE product of the Arnoldi steps: FP64 matrix, or its FP32 copy
  auto ebe_args = [&]() {
    if constexpr (kMixed)
      return base_args<T, T>(c, c->xa32.p, c->da32.p);
    else
      return base_args<T, T>(c, c->xa.p, c->da.p);
  };
  if (!(prec <= eps)) {
    for (;;) {
      m = m + 1;
      int k = K0;
      // V1 = -R0 / ||R0|| (gmres.f:243)
      if constexpr (kMixed)
        hipLaunchKernelGGL(k_scale_into_f32, dim3(g), dim3(kThreads), 0, c->stream, Vp(0), R0,
                           -1.0 / nr0, n);
      else
        hipLaunchKernelGGL(k_scale_into, dim3(g), dim3(kThreads), 0, c->stream, Vp(0), R0,
                           -1.0 / nr0, n);
      HIP_OK(hipGetLastError());
      auto matvec_w = [&](int j, const T* w0) -> int {
        // AV_j = A V_j with partial sums of AV_j . w0
        EbeArgs<T, T> a = ebe_args();
        a.y = Vp(j);
        a.out = AVp(j);
        a.w = w0;
        a.partial = mp;
        HIP_OK((launch_ebe<T, T, STAGE_PLAIN, EPI_DOTW, FIN_PARTIAL>(c, a, false, c->typdia,
                                                                      c->typext)));
        ++products;
        return BHIP_OK;
      };
      // the Arnoldi steps are queued back to back: each normalises its new
      // basis vector on the device, and the host reads all the Hessenberg
      // columns once per cycle
      for (int j = 0; j < K0 - 1; ++j) {
        if (matvec_w(j, Vp(0))) return BHIP_EDEVICE;
        // V_{j+1} from AV_j by modified Gram-Schmidt over V_0..V_j
        int npin = L.nblocks;
        for (int i = 0; i <= j; ++i) {
          const T* src = (i == 0) ? AVp(j) : Vp(j + 1);
          const T* next = (i < j) ? Vp(i + 1) : nullptr;
          double* pin = mp + (size_t)(i & 1) * stride;
          double* pout = mp + (size_t)((i + 1) & 1) * stride;
          hipLaunchKernelGGL(k_mgs<T>, dim3(g), dim3(kThreads), 0, c->stream, Vp(j + 1), src,
                             Vp(i), next, n, pin, npin, pout, ds + col(j) + i);
          npin = g;
        }
        hipLaunchKernelGGL(k_reduce_into, dim3(1), dim3(kThreads), 0, c->stream,
                           mp + (size_t)((j + 1) & 1) * stride, g, ds + col(j) + j + 1);
        hipLaunchKernelGGL(k_scale_inv_norm<T>, dim3(g), dim3(kThreads), 0, c->stream,
                           Vp(j + 1), ds + col(j) + j + 1, n);
        HIP_OK(hipGetLastError());
      }
      // K-th column (gmres.f:285-296): AV_K, H(I,K) = AV_K.V_I,
      // H(K+1,K) = sqrt(max(||AV_K||^2 - sum H(I,K)^2, 0))
      auto last_column = [&](int kk) -> int {
        EbeArgs<T, T> a = ebe_args();
        a.y = Vp(kk - 1);
        a.out = AVp(kk - 1);
        a.op = BHIP_OP_AY;
        HIP_OK((launch_ebe<T, T, STAGE_PLAIN, EPI_OP, FIN_NONE>(c, a, false, c->typdia,
                                                                 c->typext)));
        ++products;
        hipLaunchKernelGGL((k_multidot<20, T>), dim3(g), dim3(kThreads), 0, c->stream,
                           AVp(kk - 1), Vb, kk, n, mp, stride);
        hipLaunchKernelGGL(k_reduce_many, dim3(1), dim3(kThreads), 0, c->stream, mp, g, stride,
                           kk + 1, ds + lastoff);
        HIP_OK(hipGetLastError());
        return BHIP_OK;
      };
      if (last_column(K0)) return BHIP_EDEVICE;
      if (fetch(hs, lastoff + K0 + 1)) return BHIP_EDEVICE;
      for (int j = 0; j < K0 - 1; ++j) {
        for (int i = 0; i <= j; ++i) H[i][j] = hs[col(j) + i];
        H[j + 1][j] = std::sqrt(hs[col(j) + j + 1]);
        if (H[j + 1][j] < 1e-20) {
          // reduction of the Krylov space (gmres.f:274-283): the last
          // column is that of V_j, computed again
          k = j + 1;
          if (last_column(k)) return BHIP_EDEVICE;
          if (fetch(hs, lastoff + k + 1)) return BHIP_EDEVICE;
          break;
        }
      }
      double hk = hs[lastoff];
      for (int i = 0; i < k; ++i) {
        H[i][k - 1] = hs[lastoff + 1 + i];
        hk = hk - H[i][k - 1] * H[i][k - 1];
      }
      H[k][k - 1] = std::sqrt(std::max(hk, 0.0));
      // Givens rotations (gmres.f:304-335)
      e[0] = nr0;
      bool exitnow = false;
      for (int i = 0; i < k; ++i) {
        for (int j = 0; j < i; ++j) {
          const double zz = cg[j] * H[j][i] + sg[j] * H[j + 1][i];
          H[j + 1][i] = -sg[j] * H[j][i] + cg[j] * H[j + 1][i];
          H[j][i] = zz;
        }
        const double rr = std::sqrt(H[i][i] * H[i][i] + H[i + 1][i] * H[i + 1][i]);
        if (std::fabs(rr) < 1e-6) {
          exitnow = true;
          break;
        }
        cg[i] = H[i][i] / rr;
        sg[i] = H[i + 1][i] / rr;
        H[i][i] = rr;
        e[i + 1] = -sg[i] * e[i];
        e[i] = cg[i] * e[i];
      }
      if (exitnow) break;
      // H y = e (gmres.f:343-350)
      e[k - 1] = e[k - 1] / H[k - 1][k - 1];
      for (int j = k - 2; j >= 0; --j) {
        for (int l = j + 1; l < k; ++l) e[j] = e[j] - H[j][l] * e[l];
        e[j] = e[j] / H[j][j];
      }
      Coef20 ce;
      for (int l = 0; l < 20; ++l) ce.e[l] = l < k ? e[l] : 0.0;
      if constexpr (kMixed) {
        // refinement: X += V y (FP64), then the true FP64 residual
        hipLaunchKernelGGL(k_gmres_xupd, dim3(g), dim3(kThreads), 0, c->stream, x, Vb, ce, k,
                           n);
        HIP_OK(hipGetLastError());
        if (residual()) return BHIP_EDEVICE;
        if (fetch(hs, 2)) return BHIP_EDEVICE;
        nr0 = std::sqrt(hs[1]);
      } else {
        hipLaunchKernelGGL(k_gmres_update, dim3(g), dim3(kThreads), 0, c->stream, x, R0, Vb,
                           AVb, ce, k, n);
        HIP_OK(hipGetLastError());
        nr0 = std::fabs(e[k]);
      }
      prec = nr0 / nb;
      if (!(prec >= eps && m < nitmax)) break;
    }
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  info->niter = m;
  if (kMixed) {
    info->inner_iters = products;
    info->outer_iters = m;
  }
  info->relres = prec;
  info->converged = prec < eps;
  return prec < eps ? BHIP_OK : BHIP_MAXITER;
}

static int gmres_fp64(bhip_ctx* c, double* x, const double* b, double eps, int nitmax,
                      int krylov, bhip_solve_info* info) {
  return gmres_impl<double>(c, x, b, eps, nitmax, krylov, false, info);
}

static int gmres_mixed(bhip_ctx* c, double* x, const double* b, double eps, int nitmax,
                       int krylov, bool xa32_ready, bhip_solve_info* info) {
  return gmres_impl<float>(c, x, b, eps, nitmax, krylov, xa32_ready, info);
}

}  // namespace bhip
