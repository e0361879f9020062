// The time-stepped wave equation of the adjoint harness, on the device, with
// its hand-written adjoint. Included at the end of bief_hip.hip.
//
// One step (eta, u, v) -> (eta', u', v') at the new time, with Strickler
// friction K and bottom zb, restates the wave-equation treatment of telemac2d
// PROPAG (propag.F:664-716, 1369-1408, 1634-1756) for a theta = 1 scheme
// without advection:
//   node      h = eta - zb, CF = 2 g / (K^2 h^(1/3)) (friction_calc.f:130-134),
//             f = CF |u|_e / (2 h), c = 1 / (1/dt + f), q = h c u / dt,
//             y = eta on inner nodes, the tide on sea nodes
//   element   A = M/dt + MATDIFUV(U = g c, V = h) (mt01aa.f:83-106,
//             mt02aa.f:200-240), rhs = int grad(phi).q_mean - A y + M eta/dt,
//             sea rows: rhs = D fb, rows and columns of sea nodes cleared in the
//             off-diagonal terms (DIRICH, diri01.f:94-139)
//   solve     A deta = rhs (SOLVE, CG + PRECON 2 on the device)
//   element   u' = c (u/dt - g grad(eta')), grad lumped: sum over the elements
//             of S/3 grad(eta') divided by the basis integral
// All node vectors are in the context's internal order; the matrix is written
// straight into the context (slot order), so the solve needs no upload.
//
// The adjoint kernels are the transposes of these, derived by hand: the
// element terms of each kernel are differentiated with respect to the node
// values they gather, and the results are assembled onto owned nodes with the
// same colour-by-colour LDS scheme (deterministic, no atomics).

namespace bhip {

struct WaveArgs {
  // layout
  const int32_t* node_start;
  const int32_t* halo_start;
  const int32_t* halo;
  const int32_t* color_off;
  const uint16_t* lid;
  int nslot_pad, nblocks;
  // slot geometry [5][nslot_pad]
  const double* geo;
  // parameters
  double dt, grav, eps_u, tide;
  // node fields (internal order)
  const double* w;    // 1 inner node, 0 sea node
  const double* vol;  // basis integrals
  double* h;
  double* c;
  double* qx;
  double* qy;
  double* y;
  double* da_keep;
};

__device__ inline int wave_block(const WaveArgs& a) { return xcd_block(blockIdx.x, a.nblocks); }

// the node global ids of slot s of block b
__device__ inline void slot_nodes(const WaveArgs& a, int s, int n0, int nown, int hb, int (&l)[3],
                                  int (&g)[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    l[k] = a.lid[(size_t)k * a.nslot_pad + s];
    g[k] = l[k] < nown ? n0 + l[k] : a.halo[hb + l[k] - nown];
  }
}

// walks the colours of block b, calling f(s) for each slot, a barrier after
// each colour
template <typename F>
__device__ inline void colour_walk(const WaveArgs& a, int b, F&& f) {
  const int32_t* co = a.color_off + (size_t)b * (kMaxColors + 1);
  const int send = co[kMaxColors];
  for (int col = 0; col < kMaxColors; ++col) {
    const int s0 = co[col];
    if (s0 >= send) break;
    const int s1 = co[col + 1];
    for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) f(s);
    __syncthreads();
  }
}

// the MT02AA 'UV' geometric pair factors and the P1 gradient factors of the
// slot's element (node 1 at the origin)
struct ElemGeo {
  double s, p12, p13, p23, bx[3], by[3];
};
__device__ inline ElemGeo elem_geo(const WaveArgs& a, int s) {
  const size_t np = a.nslot_pad;
  const double x2 = a.geo[s], x3 = a.geo[np + s], y2 = a.geo[2 * np + s], y3 = a.geo[3 * np + s];
  ElemGeo e;
  e.s = a.geo[4 * np + s];
  e.p12 = y3 * (y2 - y3) + x3 * (x2 - x3);
  e.p13 = -(y2 * (y2 - y3) + x2 * (x2 - x3));
  e.p23 = -(y2 * y3 + x2 * x3);
  e.bx[0] = y2 - y3;
  e.bx[1] = y3;
  e.bx[2] = -y2;
  e.by[0] = x3 - x2;
  e.by[1] = -x3;
  e.by[2] = x2;
  return e;
}

// node quantities of the step
__global__ void k_wave_nodes(WaveArgs a, int n, const double* __restrict__ eta,
                             const double* __restrict__ u, const double* __restrict__ v,
                             const double* __restrict__ kstr, const double* __restrict__ zb) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double h = eta[i] - zb[i];
    const double k = kstr[i];
    const double cf = 2.0 * a.grav / (k * k * cbrt(h));
    const double sp = sqrt(u[i] * u[i] + v[i] * v[i] + a.eps_u * a.eps_u);
    const double f = 0.5 * cf * sp / h;
    const double c = 1.0 / (1.0 / a.dt + f);
    a.h[i] = h;
    a.c[i] = c;
    a.qx[i] = h * c * u[i] / a.dt;
    a.qy[i] = h * c * v[i] / a.dt;
    a.y[i] = a.w[i] != 0.0 ? eta[i] : a.tide;
  }
}

// the system of the step: A (diagonal into da, internal order; off-diagonal
// terms with the sea rows and columns cleared into xa, slot order) and its
// right-hand side
__global__ void __launch_bounds__(kThreads)
k_wave_assemble(WaveArgs a, const double* __restrict__ eta, double* __restrict__ da,
                double* __restrict__ xa, double* __restrict__ rhs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = wave_block(a);
  const int n0 = a.node_start[b], nown = a.node_start[b + 1] - n0;
  const int hb = a.halo_start[b];
  double* acc_d = reinterpret_cast<double*>(smem_raw);
  double* acc_r = acc_d + nown;
  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc_d[l] = acc_r[l] = 0.0;
  __syncthreads();
  const double mdt = 1.0 / (12.0 * a.dt);
  colour_walk(a, b, [&](int s) {
    int l[3], g[3];
    slot_nodes(a, s, n0, nown, hb, l, g);
    const ElemGeo e = elem_geo(a, s);
    double hh[3], uu[3], qxk[3], qyk[3], yk[3], ek[3], wk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      hh[k] = a.h[g[k]];
      uu[k] = a.grav * a.c[g[k]];
      qxk[k] = a.qx[g[k]];
      qyk[k] = a.qy[g[k]];
      yk[k] = a.y[g[k]];
      ek[k] = eta[g[k]];
      wk[k] = a.w[g[k]];
    }
    const double m_off = e.s * mdt;
    const double g123 = hh[0] + hh[1] + hh[2];
    const double coef = (uu[0] * (hh[0] + g123) + uu[1] * (hh[1] + g123) +
                         uu[2] * (hh[2] + g123)) / (48.0 * e.s);
    const double a12 = e.p12 * coef, a13 = e.p13 * coef, a23 = e.p23 * coef;
    const double x12 = a12 + m_off, x13 = a13 + m_off, x23 = a23 + m_off;
    const double t[3] = {-a12 - a13 + 2 * m_off, -a12 - a23 + 2 * m_off,
                         -a13 - a23 + 2 * m_off};
    const double qxm = (qxk[0] + qxk[1] + qxk[2]) / 3.0;
    const double qym = (qyk[0] + qyk[1] + qyk[2]) / 3.0;
    const double ay[3] = {t[0] * yk[0] + x12 * yk[1] + x13 * yk[2],
                          x12 * yk[0] + t[1] * yk[1] + x23 * yk[2],
                          x13 * yk[0] + x23 * yk[1] + t[2] * yk[2]};
    const double me[3] = {m_off * (2 * ek[0] + ek[1] + ek[2]), m_off * (ek[0] + 2 * ek[1] + ek[2]),
                          m_off * (ek[0] + ek[1] + 2 * ek[2])};
    const size_t np = a.nslot_pad;
    xa[s] = x12 * (wk[0] * wk[1]);
    xa[np + s] = x13 * (wk[0] * wk[2]);
    xa[2 * np + s] = x23 * (wk[1] * wk[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (l[k] < nown) {
        acc_d[l[k]] += t[k];
        acc_r[l[k]] += 0.5 * (e.bx[k] * qxm + e.by[k] * qym) - ay[k] + me[k];
      }
    }
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    const double d = acc_d[l];
    da[i] = d;
    a.da_keep[i] = d;
    rhs[i] = a.w[i] != 0.0 ? acc_r[l] : d * (a.tide - eta[i]);
  }
}

// the new state from the increment: eta' = eta + deta, the lumped gradient of
// eta' and the momentum update
__global__ void __launch_bounds__(kThreads)
k_wave_momentum(WaveArgs a, const double* __restrict__ eta, const double* __restrict__ deta,
                const double* __restrict__ u, const double* __restrict__ v,
                double* __restrict__ eta_n, double* __restrict__ u_n, double* __restrict__ v_n) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = wave_block(a);
  const int n0 = a.node_start[b], nown = a.node_start[b + 1] - n0;
  const int hb = a.halo_start[b];
  double* acc_x = reinterpret_cast<double*>(smem_raw);
  double* acc_y = acc_x + nown;
  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc_x[l] = acc_y[l] = 0.0;
  __syncthreads();
  colour_walk(a, b, [&](int s) {
    int l[3], g[3];
    slot_nodes(a, s, n0, nown, hb, l, g);
    const ElemGeo e = elem_geo(a, s);
    double en[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) en[k] = eta[g[k]] + deta[g[k]];
    const double gx = (e.bx[0] * en[0] + e.bx[1] * en[1] + e.bx[2] * en[2]) / 6.0;
    const double gy = (e.by[0] * en[0] + e.by[1] * en[1] + e.by[2] * en[2]) / 6.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (l[k] < nown) {
        acc_x[l[k]] += gx;
        acc_y[l[k]] += gy;
      }
    }
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    const double c = a.c[i];
    eta_n[i] = eta[i] + deta[i];
    u_n[i] = c * (u[i] / a.dt - a.grav * (acc_x[l] / a.vol[i]));
    v_n[i] = c * (v[i] / a.dt - a.grav * (acc_y[l] / a.vol[i]));
  }
}

// adjoint of the momentum update with respect to eta' (the c and u parts are
// node-local and taken in k_wave_assemble_adj): etot = ebar + grad^T(...)
__global__ void __launch_bounds__(kThreads)
k_wave_momentum_adj(WaveArgs a, const double* __restrict__ ubar_n,
                    const double* __restrict__ vbar_n, const double* __restrict__ ebar_n,
                    double* __restrict__ etot) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = wave_block(a);
  const int n0 = a.node_start[b], nown = a.node_start[b + 1] - n0;
  const int hb = a.halo_start[b];
  double* acc = reinterpret_cast<double*>(smem_raw);
  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc[l] = 0.0;
  __syncthreads();
  colour_walk(a, b, [&](int s) {
    int l[3], g[3];
    slot_nodes(a, s, n0, nown, hb, l, g);
    const ElemGeo e = elem_geo(a, s);
    // adjoint of the lumped gradient: -g c ubar / vol at each node, summed
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double f = -a.grav * a.c[g[k]] / a.vol[g[k]];
      sx += f * ubar_n[g[k]];
      sy += f * vbar_n[g[k]];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (l[k] < nown) acc[l[k]] += (e.bx[k] * sx + e.by[k] * sy) / 6.0;
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    etot[i] = ebar_n[i] + acc[l];
  }
}

// adjoint of the system, its right-hand side and the node quantities, given
// lam = A^-1 etot (the adjoint solve) and the increment deta. Writes the
// adjoint state of step n (ebar, ubar, vbar) and adds to those of K and zb.
struct WaveAdjIn {
  const double* lam;
  const double* deta;
  const double* eta;
  const double* u;
  const double* v;
  const double* u_n;
  const double* v_n;
  const double* ubar_n;
  const double* vbar_n;
  const double* etot;
  const double* kstr;
};
struct WaveAdjOut {
  double* ebar;
  double* ubar;
  double* vbar;
  double* kbar;   // accumulated
  double* zbbar;  // accumulated
};

__global__ void __launch_bounds__(kThreads)
k_wave_assemble_adj(WaveArgs a, WaveAdjIn in, WaveAdjOut out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = wave_block(a);
  const int n0 = a.node_start[b], nown = a.node_start[b + 1] - n0;
  const int hb = a.halo_start[b];
  double* acc_c = reinterpret_cast<double*>(smem_raw);
  double* acc_h = acc_c + nown;
  double* acc_qx = acc_h + nown;
  double* acc_qy = acc_qx + nown;
  double* acc_y = acc_qy + nown;
  double* acc_m = acc_y + nown;
  for (int l = threadIdx.x; l < nown; l += blockDim.x)
    acc_c[l] = acc_h[l] = acc_qx[l] = acc_qy[l] = acc_y[l] = acc_m[l] = 0.0;
  __syncthreads();
  const double mdt = 1.0 / (12.0 * a.dt);
  colour_walk(a, b, [&](int s) {
    int l[3], g[3];
    slot_nodes(a, s, n0, nown, hb, l, g);
    const ElemGeo e = elem_geo(a, s);
    double hh[3], uu[3], yk[3], wk[3], lam[3], de[3], rho[3], dabar[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = g[k];
      hh[k] = a.h[i];
      uu[k] = a.grav * a.c[i];
      yk[k] = a.y[i];
      wk[k] = a.w[i];
      lam[k] = in.lam[i];
      de[k] = in.deta[i];
      rho[k] = wk[k] * lam[k];
      // adjoint of the assembled diagonal: the solve (-lam deta), the sea
      // rows D fb (lam fb), the product -A y (-rho y)
      const double fb = wk[k] != 0.0 ? 0.0 : a.tide - in.eta[i];
      dabar[k] = -lam[k] * de[k] + (wk[k] != 0.0 ? 0.0 : lam[k] * fb) - rho[k] * yk[k];
    }
    const double m_off = e.s * mdt;
    const double g123 = hh[0] + hh[1] + hh[2];
    const double u123 = uu[0] + uu[1] + uu[2];
    const double coef = (uu[0] * (hh[0] + g123) + uu[1] * (hh[1] + g123) +
                         uu[2] * (hh[2] + g123)) / (48.0 * e.s);
    const double a12 = e.p12 * coef, a13 = e.p13 * coef, a23 = e.p23 * coef;
    const double x12 = a12 + m_off, x13 = a13 + m_off, x23 = a23 + m_off;
    const double t[3] = {-a12 - a13 + 2 * m_off, -a12 - a23 + 2 * m_off,
                         -a13 - a23 + 2 * m_off};
    // adjoints of the off-diagonal terms: the solve's -(lam_k de_l + lam_l
    // de_k) through the cleared rows/columns, the product's -(rho_k y_l + rho_l y_k)
    const double xb12 = -(lam[0] * de[1] + lam[1] * de[0]) * (wk[0] * wk[1]) -
                        (rho[0] * yk[1] + rho[1] * yk[0]);
    const double xb13 = -(lam[0] * de[2] + lam[2] * de[0]) * (wk[0] * wk[2]) -
                        (rho[0] * yk[2] + rho[2] * yk[0]);
    const double xb23 = -(lam[1] * de[2] + lam[2] * de[1]) * (wk[1] * wk[2]) -
                        (rho[1] * yk[2] + rho[2] * yk[1]);
    // through X = a + m_off and T_k = -(sum of a_kl) + 2 m_off
    const double ab12 = xb12 - dabar[0] - dabar[1];
    const double ab13 = xb13 - dabar[0] - dabar[2];
    const double ab23 = xb23 - dabar[1] - dabar[2];
    const double cb2 = (ab12 * e.p12 + ab13 * e.p13 + ab23 * e.p23) / (48.0 * e.s);
    // -A rho (A symmetric) and M rho / dt
    const double arho[3] = {t[0] * rho[0] + x12 * rho[1] + x13 * rho[2],
                            x12 * rho[0] + t[1] * rho[1] + x23 * rho[2],
                            x13 * rho[0] + x23 * rho[1] + t[2] * rho[2]};
    const double mr[3] = {m_off * (2 * rho[0] + rho[1] + rho[2]),
                          m_off * (rho[0] + 2 * rho[1] + rho[2]),
                          m_off * (rho[0] + rho[1] + 2 * rho[2])};
    const double qxb = 0.5 * (e.bx[0] * rho[0] + e.bx[1] * rho[1] + e.bx[2] * rho[2]) / 3.0;
    const double qyb = 0.5 * (e.by[0] * rho[0] + e.by[1] * rho[1] + e.by[2] * rho[2]) / 3.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (l[k] < nown) {
        const int q = l[k];
        acc_c[q] += a.grav * cb2 * (hh[k] + g123);
        acc_h[q] += cb2 * (uu[k] + u123);
        acc_qx[q] += qxb;
        acc_qy[q] += qyb;
        acc_y[q] -= arho[k];
        acc_m[q] += mr[k];
      }
    }
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    const double dt = a.dt;
    const double w = a.w[i];
    const double lam = in.lam[i];
    const double ybar = acc_y[l];
    const double fbbar = ybar + (w != 0.0 ? 0.0 : lam * a.da_keep[i]);
    double ebar = in.etot[i] + ybar + acc_m[l] - (w != 0.0 ? 0.0 : fbbar);
    // the node quantities again
    const double h = a.h[i], c = a.c[i];
    const double uu = in.u[i], vv = in.v[i];
    const double k = in.kstr[i];
    const double cf = 2.0 * a.grav / (k * k * cbrt(h));
    const double sp = sqrt(uu * uu + vv * vv + a.eps_u * a.eps_u);
    const double f = 0.5 * cf * sp / h;
    const double qxbar = acc_qx[l], qybar = acc_qy[l];
    const double ubn = in.ubar_n[i], vbn = in.vbar_n[i];
    // u' = c (u/dt - g gx): du'/dc = u'/c
    double cbar = acc_c[l] + (ubn * in.u_n[i] + vbn * in.v_n[i]) / c +
                  (qxbar * uu + qybar * vv) * h / dt;
    double hbar = acc_h[l] + (qxbar * uu + qybar * vv) * c / dt;
    const double fbar = -c * c * cbar;
    const double cfbar = fbar * 0.5 * sp / h;
    const double spbar = fbar * 0.5 * cf / h;
    hbar += -fbar * f / h - cfbar * cf / (3.0 * h);
    out.ubar[i] = ubn * c / dt + qxbar * h * c / dt + spbar * uu / sp;
    out.vbar[i] = vbn * c / dt + qybar * h * c / dt + spbar * vv / sp;
    out.kbar[i] += cfbar * (-2.0 * cf / k);
    ebar += hbar;
    out.zbbar[i] -= hbar;
    out.ebar[i] = ebar;
  }
}

}  // namespace bhip

namespace {

WaveArgs wave_args(bhip_ctx* c, const bhip_wave_params* p) {
  const Layout& L = c->L;
  WaveArgs a;
  a.node_start = c->node_start.p;
  a.halo_start = c->halo_start.p;
  a.halo = c->halo.p;
  a.color_off = c->color_off.p;
  a.lid = c->lid.p;
  a.nslot_pad = L.nslot_pad;
  a.nblocks = L.nblocks;
  a.geo = c->wv_geo.p;
  a.dt = p ? p->dt : 1.0;
  a.grav = p ? p->grav : 9.81;
  a.eps_u = p ? p->eps_u : 0.0;
  a.tide = p ? p->tide : 0.0;
  a.w = c->wv_w.p;
  a.vol = c->wv_vol.p;
  a.h = c->wv_h.p;
  a.c = c->wv_c.p;
  a.qx = c->wv_qx.p;
  a.qy = c->wv_qy.p;
  a.y = c->wv_y.p;
  a.da_keep = c->wv_da.p;
  return a;
}

bool wave_params_ok(const bhip_wave_params* p) {
  return p && p->dt > 0.0 && p->grav > 0.0 && p->eps_u >= 0.0;
}

}  // namespace

extern "C" {

int bhip_wave_setup(bhip_ctx* c, const double* xel, const double* yel, const double* surfac,
                    const int32_t* sea) {
  if (!c || !xel || !yel || !surfac || !sea || c->L.ndp != 3 || c->nsys != 1)
    return BHIP_EBADARG;
  const Layout& L = c->L;
  const int n = L.npoin, ne = L.nelem;
  // k_wave_assemble_adj holds six accumulators per owned node in LDS
  if ((size_t)6 * L.max_owned * sizeof(double) > 64 * 1024) {
    fprintf(stderr, "libbiefhip: bhip_wave_setup: blocks of %d nodes need more than 64 KiB "
            "of LDS; use blocks of at most 1365 nodes\n", L.max_owned);
    return BHIP_EBADARG;
  }
  for (int e = 0; e < ne; ++e)
    if (!(surfac[e] > 0.0)) return BHIP_EBADARG;
  for (auto* v : {&c->wv_w, &c->wv_vol, &c->wv_h, &c->wv_c, &c->wv_qx, &c->wv_qy, &c->wv_y,
                  &c->wv_da})
    HIP_OK(v->alloc(n));
  HIP_OK(c->wv_geo.alloc((size_t)5 * L.nslot_pad));
  // geometry per slot, from the caller's element arrays
  std::vector<double> geo((size_t)5 * L.nslot_pad, 0.0);
  for (int s = 0; s < L.nslot; ++s) {
    const int e = L.slot_elem[s];
    geo[s] = xel[(size_t)ne + e];
    geo[(size_t)L.nslot_pad + s] = xel[2 * (size_t)ne + e];
    geo[2 * (size_t)L.nslot_pad + s] = yel[(size_t)ne + e];
    geo[3 * (size_t)L.nslot_pad + s] = yel[2 * (size_t)ne + e];
    geo[4 * (size_t)L.nslot_pad + s] = surfac[e];
  }
  // basis integrals (VOLU2D: assembled S/3) and the inner-node mask, internal order
  // (each block lists every element touching its owned nodes once)
  std::vector<double> vol(n, 0.0), w(n);
  for (int b = 0; b < L.nblocks; ++b) {
    const int n0 = L.node_start[b], nown = L.node_start[b + 1] - n0;
    const int32_t* co = L.color_off.data() + (size_t)b * (kMaxColors + 1);
    for (int s = co[0]; s < co[kMaxColors]; ++s)
      for (int k = 0; k < 3; ++k) {
        const int l = L.lid[(size_t)k * L.nslot_pad + s];
        if (l < nown) vol[n0 + l] += surfac[L.slot_elem[s]] / 3.0;
      }
  }
  for (int i = 0; i < n; ++i) w[i] = sea[L.perm[i]] ? 0.0 : 1.0;
  HIP_OK(hipMemcpyAsync(c->wv_geo.p, geo.data(), geo.size() * sizeof(double),
                        hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->wv_vol.p, vol.data(), n * sizeof(double), hipMemcpyHostToDevice,
                        c->stream));
  HIP_OK(hipMemcpyAsync(c->wv_w.p, w.data(), n * sizeof(double), hipMemcpyHostToDevice,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (!c->have_matrix || c->xa.n < (size_t)6 * L.nslot_pad) {
    HIP_OK(c->da.alloc(n));
    HIP_OK(c->xa.alloc((size_t)6 * L.nslot_pad));
  }
  c->wv_ready = true;
  return BHIP_OK;
}

int bhip_wave_assemble(bhip_ctx* c, const bhip_wave_params* p, const double* eta,
                       const double* u, const double* v, const double* kstr, const double* zb,
                       double* rhs) {
  if (!c || !c->wv_ready || !wave_params_ok(p) || !eta || !u || !v || !kstr || !zb || !rhs)
    return BHIP_EBADARG;
  const Layout& L = c->L;
  WaveArgs a = wave_args(c, p);
  hipLaunchKernelGGL(k_wave_nodes, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream, a,
                     L.npoin, eta, u, v, kstr, zb);
  hipLaunchKernelGGL(k_wave_assemble, dim3(L.nblocks), dim3(kThreads),
                     2 * (size_t)L.max_owned * sizeof(double), c->stream, a, eta, c->da.p,
                     c->xa.p, rhs);
  HIP_OK(hipGetLastError());
  // MATRIY's storage of the assembled system (matriy.f:375-376)
  c->typdia = 'Q';
  c->typext = 'S';
  c->ncoef = 3;
  c->have_matrix = true;
  return BHIP_OK;
}

int bhip_wave_momentum(bhip_ctx* c, const bhip_wave_params* p, const double* eta,
                       const double* deta, const double* u, const double* v, double* eta_n,
                       double* u_n, double* v_n) {
  if (!c || !c->wv_ready || !wave_params_ok(p) || !eta || !deta || !u || !v || !eta_n || !u_n ||
      !v_n)
    return BHIP_EBADARG;
  const Layout& L = c->L;
  hipLaunchKernelGGL(k_wave_momentum, dim3(L.nblocks), dim3(kThreads),
                     2 * (size_t)L.max_owned * sizeof(double), c->stream, wave_args(c, p), eta,
                     deta, u, v, eta_n, u_n, v_n);
  HIP_OK(hipGetLastError());
  return BHIP_OK;
}

int bhip_wave_momentum_adj(bhip_ctx* c, const bhip_wave_params* p, const double* ubar_n,
                           const double* vbar_n, const double* ebar_n, double* etot) {
  if (!c || !c->wv_ready || !wave_params_ok(p) || !ubar_n || !vbar_n || !ebar_n || !etot)
    return BHIP_EBADARG;
  const Layout& L = c->L;
  hipLaunchKernelGGL(k_wave_momentum_adj, dim3(L.nblocks), dim3(kThreads),
                     (size_t)L.max_owned * sizeof(double), c->stream, wave_args(c, p), ubar_n,
                     vbar_n, ebar_n, etot);
  HIP_OK(hipGetLastError());
  return BHIP_OK;
}

int bhip_wave_assemble_adj(bhip_ctx* c, const bhip_wave_params* p, const double* lam,
                           const double* deta, const double* eta, const double* u,
                           const double* v, const double* u_n, const double* v_n,
                           const double* ubar_n, const double* vbar_n, const double* etot,
                           const double* kstr, double* ebar, double* ubar, double* vbar,
                           double* kbar, double* zbbar) {
  if (!c || !c->wv_ready || !wave_params_ok(p) || !lam || !deta || !eta || !u || !v || !u_n ||
      !v_n || !ubar_n || !vbar_n || !etot || !kstr || !ebar || !ubar || !vbar || !kbar ||
      !zbbar)
    return BHIP_EBADARG;
  const Layout& L = c->L;
  WaveAdjIn in = {lam, deta, eta, u, v, u_n, v_n, ubar_n, vbar_n, etot, kstr};
  WaveAdjOut out = {ebar, ubar, vbar, kbar, zbbar};
  hipLaunchKernelGGL(k_wave_assemble_adj, dim3(L.nblocks), dim3(kThreads),
                     6 * (size_t)L.max_owned * sizeof(double), c->stream, wave_args(c, p), in,
                     out);
  HIP_OK(hipGetLastError());
  return BHIP_OK;
}

}  // extern "C"
