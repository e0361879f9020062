// The time-stepped wave equation of the adjoint harness, on the device, with
// its hand-written adjoint. Included at the end of bief_hip.hip.
//
// One step (eta, u, v) -> (eta', u', v') at the new time, with Strickler
// friction K and bottom zb, is telemac2d PROPAG's wave-equation treatment
// (TREATMENT OF THE LINEAR SYSTEM = 2; propag.F:664-716, 1003-1012,
// 1150-1170, 1369-1465, 1634-1710) for these options: P1 depth and velocity,
// implicitation 1 (TETAH = TETAU = TETAZCOMP = 1), no advection or diffusion
// of velocity, no sources, no tidal flats, one sub-iteration, Strickler law
// (KFROT 3, HFROT 1), the free surface prescribed on sea nodes with free
// velocity there, walls elsewhere (no boundary flux term then):
//   node      h = eta - zb; CF = 2 g / K^2 / max(h, 1e-4)^(1/3)
//             (friction_unif.f:124-127, friction_calc.f:130-134);
//             f = CF |u| / (2 max(h, 1e-9)), |u| raised to the shallow-water
//             speed below 3 cm (fricti.f:123-135); the sea nodes' prescribed
//             increment fb = tide - eta (HBOR - HN)
//   element   GX = GRADF(eta) (vc13aa.f), VS = MASBAS weighted by the bottom
//             slope factor (slop10.f; propag.F:1155-1161)
//   node      AM2D = VOL/dt + f VS, c = VOL / AM2D (so DM1 = -g c),
//             u* = c (u/dt - g GX/VOL)  (T2 = CV2 / AM2D, propag.F:1413)
//   element   A = M/dt + MATDIFUV(-1, DM1, max(h,0)) (mt01aa.f:83-106,
//             mt02aa.f:200-240); rhs = -MATFGR(h)u* (mt08aa.f; CV1 =
//             -BM1 T2 - BM2 T3, propag.F:1464-1465) - A fb; sea rows
//             rhs = D fb and their rows and columns cleared off the diagonal
//             (DIRICH, diri01.f:94-139)
//   solve     A deta = rhs (SOLVE, CG + PRECON 2 on the device)
//   element   u' = c (u/dt - g GRADF(eta + deta)/VOL)  (propag.F:1685-1707)
// All node vectors are in the context's internal order; the matrix is written
// straight into the context (slot order), so the solve needs no upload.
//
// The adjoint kernels are the transposes of these, derived by hand: the
// element terms of each kernel are differentiated with respect to the node
// values they gather, and the results are assembled onto owned nodes with the
// same colour-by-colour LDS scheme (deterministic, no atomics). The floors
// max(h, .) pass the derivative where h is above them and nothing below.

namespace bhip {

constexpr double kWaveHO = 3e-2;      // fricti.f: HO, the shallow-water threshold
constexpr double kWaveGHO = 9.81;     // fricti.f: the gravity of its shallow speed
constexpr double kWaveMinHF = 1e-9;   // fricti.f: MAX(HHN, 1.D-9)
constexpr double kWaveMinHC = 1e-4;   // friction_unif.f: MINH

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
  const double* vol;  // basis integrals (VOLU2D)
  double* h;          // eta - zb
  double* zb;         // the bottom of the step
  double* f;          // friction factor 0.5 CF |u| / h
  double* vs;         // slope-weighted basis integrals
  double* c;          // VOL / AM2D
  double* usx;        // u* (T2), v* (T3)
  double* usy;
  double* fb;         // prescribed increment on sea nodes, 0 elsewhere
  double* da_keep;    // assembled diagonal of the step's system
  // adjoint scratch: adjoints of u*, v*, VS and the partial adjoint of eta
  double* sbx;
  double* sby;
  double* svs;
  double* seb;
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
// slot's element (node 1 at the origin): dphi_k/dx = bx_k / (2S)
struct ElemGeo {
  double s, x2, x3, y2, y3, p12, p13, p23, bx[3], by[3];
};
__device__ inline ElemGeo elem_geo(const WaveArgs& a, int s) {
  const size_t np = a.nslot_pad;
  ElemGeo e;
  e.x2 = a.geo[s];
  e.x3 = a.geo[np + s];
  e.y2 = a.geo[2 * np + s];
  e.y3 = a.geo[3 * np + s];
  e.s = a.geo[4 * np + s];
  const double x2 = e.x2, x3 = e.x3, y2 = e.y2, y3 = e.y3;
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

// GRADF of a P1 field on the element, the same value for its three nodes
// (vc13aa.f, XMUL = 1)
__device__ inline double gradf_x(const ElemGeo& e, const double (&q)[3]) {
  return (e.y2 * (q[0] - q[2]) + e.y3 * (q[1] - q[0])) / 6.0;
}
__device__ inline double gradf_y(const ElemGeo& e, const double (&q)[3]) {
  return (e.x2 * (q[2] - q[0]) + e.x3 * (q[0] - q[1])) / 6.0;
}

// SLOP10 (slop10.f): sqrt((A + B + C) / A), the area factor of the inclined
// bottom, with its partial derivatives along the bottom differences z2 - z1,
// z3 - z1 when asked
__device__ inline double elem_slope(const ElemGeo& e, double z1, double z2, double z3,
                                    double* d2 = nullptr, double* d3 = nullptr) {
  const double Z2 = z2 - z1, Z3 = z3 - z1;
  const double ta = e.x2 * e.y3 - e.x3 * e.y2;
  const double tb = e.y2 * Z3 - Z2 * e.y3;
  const double tc = e.x3 * Z2 - Z3 * e.x2;
  const double A = ta * ta, B = tb * tb, C = tc * tc;
  const double sl = sqrt((A + B + C) / A);
  if (d2) {
    const double k = 1.0 / (2.0 * sl * A);
    *d2 = k * (-2.0 * e.y3 * tb + 2.0 * e.x3 * tc);
    *d3 = k * (2.0 * e.y2 * tb - 2.0 * e.x2 * tc);
  }
  return sl;
}

// node quantities of the step: depth, friction factor, prescribed increment
__global__ void k_wave_nodes(WaveArgs a, int n, const double* __restrict__ eta,
                             const double* __restrict__ u, const double* __restrict__ v,
                             const double* __restrict__ kstr, const double* __restrict__ zb) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double h = eta[i] - zb[i];
    const double k = kstr[i];
    const double hc = fmax(h, kWaveMinHC);
    const double cf = 2.0 * a.grav / (k * k) / pow(hc, 1.0 / 3.0);
    const double hf = fmax(h, kWaveMinHF);
    double sp = sqrt(u[i] * u[i] + v[i] * v[i] + a.eps_u * a.eps_u);
    if (hf < kWaveHO) sp = fmax(sp, sqrt(kWaveGHO * (kWaveHO - hf) * hf / kWaveHO));
    a.h[i] = h;
    a.zb[i] = zb[i];
    a.f[i] = 0.5 * cf * sp / hf;
    a.fb[i] = a.w[i] != 0.0 ? 0.0 : a.tide - eta[i];
  }
}

// GRADF(eta) and the slope-weighted basis integrals at owned nodes, then
// c = VOL / (VOL/dt + f VS) and the velocity predictor u* = c (u/dt - g GX/VOL)
__global__ void __launch_bounds__(kThreads)
k_wave_grad(WaveArgs a, const double* __restrict__ eta, const double* __restrict__ u,
            const double* __restrict__ v) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = wave_block(a);
  const int n0 = a.node_start[b], nown = a.node_start[b + 1] - n0;
  const int hb = a.halo_start[b];
  double* acc_x = reinterpret_cast<double*>(smem_raw);
  double* acc_y = acc_x + nown;
  double* acc_v = acc_y + nown;
  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc_x[l] = acc_y[l] = acc_v[l] = 0.0;
  __syncthreads();
  colour_walk(a, b, [&](int s) {
    int l[3], g[3];
    slot_nodes(a, s, n0, nown, hb, l, g);
    const ElemGeo e = elem_geo(a, s);
    double ek[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) ek[k] = eta[g[k]];
    const double gx = gradf_x(e, ek), gy = gradf_y(e, ek);
    const double vs = e.s / 3.0 * elem_slope(e, a.zb[g[0]], a.zb[g[1]], a.zb[g[2]]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (l[k] < nown) {
        acc_x[l[k]] += gx;
        acc_y[l[k]] += gy;
        acc_v[l[k]] += vs;
      }
    }
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    const double vol = a.vol[i];
    const double am = vol / a.dt + a.f[i] * acc_v[l];
    // OS 'X=1/Y' guarded at 1e-6 (propag.F:1384)
    const double c = fabs(am) > 1e-6 ? vol / am : 0.0;
    a.vs[i] = acc_v[l];
    a.c[i] = c;
    a.usx[i] = c * (u[i] / a.dt - a.grav * (acc_x[l] / vol));
    a.usy[i] = c * (v[i] / a.dt - a.grav * (acc_y[l] / vol));
  }
}

// the system of the step: A (diagonal into da, internal order; off-diagonal
// terms with the sea rows and columns cleared into xa, slot order) and its
// right-hand side
__global__ void __launch_bounds__(kThreads)
k_wave_assemble(WaveArgs a, double* __restrict__ da, double* __restrict__ xa,
                double* __restrict__ rhs) {
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
    double hp[3], uu[3], sx[3], sy[3], fb[3], wk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      hp[k] = fmax(a.h[g[k]], 0.0);
      uu[k] = a.grav * a.c[g[k]];
      sx[k] = a.usx[g[k]];
      sy[k] = a.usy[g[k]];
      fb[k] = a.fb[g[k]];
      wk[k] = a.w[g[k]];
    }
    const double m_off = e.s * mdt;
    const double g123 = hp[0] + hp[1] + hp[2];
    const double coef = (uu[0] * (hp[0] + g123) + uu[1] * (hp[1] + g123) +
                         uu[2] * (hp[2] + g123)) / (48.0 * e.s);
    const double a12 = e.p12 * coef, a13 = e.p13 * coef, a23 = e.p23 * coef;
    const double x12 = a12 + m_off, x13 = a13 + m_off, x23 = a23 + m_off;
    const double t[3] = {-a12 - a13 + 2 * m_off, -a12 - a23 + 2 * m_off,
                         -a13 - a23 + 2 * m_off};
    // -MATFGR(h) u*: r_k = bx_k Qx + by_k Qy with F = h/24 (mt08aa.f)
    const double F[3] = {hp[0] / 24.0, hp[1] / 24.0, hp[2] / 24.0};
    const double F123 = F[0] + F[1] + F[2];
    const double qx = F123 * (sx[0] + sx[1] + sx[2]) + F[0] * sx[0] + F[1] * sx[1] + F[2] * sx[2];
    const double qy = F123 * (sy[0] + sy[1] + sy[2]) + F[0] * sy[0] + F[1] * sy[1] + F[2] * sy[2];
    // DIRICH: b -= A fb
    const double af[3] = {t[0] * fb[0] + x12 * fb[1] + x13 * fb[2],
                          x12 * fb[0] + t[1] * fb[1] + x23 * fb[2],
                          x13 * fb[0] + x23 * fb[1] + t[2] * fb[2]};
    const size_t np = a.nslot_pad;
    xa[s] = x12 * (wk[0] * wk[1]);
    xa[np + s] = x13 * (wk[0] * wk[2]);
    xa[2 * np + s] = x23 * (wk[1] * wk[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (l[k] < nown) {
        acc_d[l[k]] += t[k];
        acc_r[l[k]] += e.bx[k] * qx + e.by[k] * qy - af[k];
      }
    }
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    const double d = acc_d[l];
    da[i] = d;
    a.da_keep[i] = d;
    rhs[i] = a.w[i] != 0.0 ? acc_r[l] : d * a.fb[i];
  }
}

// the new state from the increment: eta' = eta + deta, GRADF(eta') and the
// momentum update u' = c (u/dt - g GRADF(eta')/VOL)
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
    const double gx = gradf_x(e, en), gy = gradf_y(e, en);
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

// transpose of GRADF: etot = ebar + GRADF^T(-g c ubar / VOL, -g c vbar / VOL)
// (the momentum update's dependence on eta'; the c and u parts are
// node-local and taken in k_wave_assemble_adj)
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
// adjoint state of step n (ebar without the GRADF and slope transposes, which
// k_wave_grad_adj adds; ubar, vbar), adds to those of K and zb, and leaves
// the adjoints of u*, v* and VS in the scratch vectors.
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
  double* acc_sx = acc_h + nown;
  double* acc_sy = acc_sx + nown;
  double* acc_fb = acc_sy + nown;
  for (int l = threadIdx.x; l < nown; l += blockDim.x)
    acc_c[l] = acc_h[l] = acc_sx[l] = acc_sy[l] = acc_fb[l] = 0.0;
  __syncthreads();
  const double mdt = 1.0 / (12.0 * a.dt);
  colour_walk(a, b, [&](int s) {
    int l[3], g[3];
    slot_nodes(a, s, n0, nown, hb, l, g);
    const ElemGeo e = elem_geo(a, s);
    double hp[3], uu[3], sx[3], sy[3], fb[3], wk[3], lam[3], de[3], rho[3], dabar[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = g[k];
      hp[k] = fmax(a.h[i], 0.0);
      uu[k] = a.grav * a.c[i];
      sx[k] = a.usx[i];
      sy[k] = a.usy[i];
      fb[k] = a.fb[i];
      wk[k] = a.w[i];
      lam[k] = in.lam[i];
      de[k] = in.deta[i];
      rho[k] = wk[k] * lam[k];
      // adjoint of the element's diagonal term: the solve (-lam deta) and
      // the sea rows D fb (lam fb)
      dabar[k] = -lam[k] * de[k] + (1.0 - wk[k]) * lam[k] * fb[k];
    }
    const double m_off = e.s * mdt;
    const double g123 = hp[0] + hp[1] + hp[2];
    const double u123 = uu[0] + uu[1] + uu[2];
    const double coef = (uu[0] * (hp[0] + g123) + uu[1] * (hp[1] + g123) +
                         uu[2] * (hp[2] + g123)) / (48.0 * e.s);
    const double a12 = e.p12 * coef, a13 = e.p13 * coef, a23 = e.p23 * coef;
    const double x12 = a12 + m_off, x13 = a13 + m_off, x23 = a23 + m_off;
    const double t[3] = {-a12 - a13 + 2 * m_off, -a12 - a23 + 2 * m_off,
                         -a13 - a23 + 2 * m_off};
    // adjoints of the off-diagonal terms: the solve's -(lam_k de_l + lam_l
    // de_k) through the cleared rows/columns, DIRICH's -(rho_k fb_l + rho_l fb_k)
    const double xb12 = -(lam[0] * de[1] + lam[1] * de[0]) * (wk[0] * wk[1]) -
                        (rho[0] * fb[1] + rho[1] * fb[0]);
    const double xb13 = -(lam[0] * de[2] + lam[2] * de[0]) * (wk[0] * wk[2]) -
                        (rho[0] * fb[2] + rho[2] * fb[0]);
    const double xb23 = -(lam[1] * de[2] + lam[2] * de[1]) * (wk[1] * wk[2]) -
                        (rho[1] * fb[2] + rho[2] * fb[1]);
    // through X = a + m_off and T_k = -(sum of a_kl) + 2 m_off
    const double ab12 = xb12 - dabar[0] - dabar[1];
    const double ab13 = xb13 - dabar[0] - dabar[2];
    const double ab23 = xb23 - dabar[1] - dabar[2];
    const double cb2 = (ab12 * e.p12 + ab13 * e.p13 + ab23 * e.p23) / (48.0 * e.s);
    // DIRICH's -A fb: its transpose -A rho onto fb (A symmetric)
    const double arho[3] = {t[0] * rho[0] + x12 * rho[1] + x13 * rho[2],
                            x12 * rho[0] + t[1] * rho[1] + x23 * rho[2],
                            x13 * rho[0] + x23 * rho[1] + t[2] * rho[2]};
    // -MATFGR(h) u*: r_k = bx_k Qx + by_k Qy, Qx = F123 sum(u*) + sum(F u*)
    const double rx = e.bx[0] * rho[0] + e.bx[1] * rho[1] + e.bx[2] * rho[2];
    const double ry = e.by[0] * rho[0] + e.by[1] * rho[1] + e.by[2] * rho[2];
    const double F[3] = {hp[0] / 24.0, hp[1] / 24.0, hp[2] / 24.0};
    const double F123 = F[0] + F[1] + F[2];
    const double Sx = sx[0] + sx[1] + sx[2], Sy = sy[0] + sy[1] + sy[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (l[k] < nown) {
        const int q = l[k];
        acc_c[q] += a.grav * cb2 * (hp[k] + g123);
        acc_h[q] += cb2 * (uu[k] + u123) + (rx * (Sx + sx[k]) + ry * (Sy + sy[k])) / 24.0;
        acc_sx[q] += rx * (F123 + F[k]);
        acc_sy[q] += ry * (F123 + F[k]);
        acc_fb[q] -= arho[k];
      }
    }
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    const double dt = a.dt;
    const double w = a.w[i];
    const double lam = in.lam[i];
    const double fbbar = acc_fb[l] + (w != 0.0 ? 0.0 : lam * a.da_keep[i]);
    // eta' = eta + deta; fb = tide - eta on sea nodes
    double ebar = in.etot[i] - (w != 0.0 ? 0.0 : fbbar);
    // the node quantities again
    const double h = a.h[i], c = a.c[i], f = a.f[i], vs = a.vs[i], vol = a.vol[i];
    const double uu = in.u[i], vv = in.v[i];
    const double k = in.kstr[i];
    const double hc = fmax(h, kWaveMinHC);
    const double cf = 2.0 * a.grav / (k * k) / pow(hc, 1.0 / 3.0);
    const double hf = fmax(h, kWaveMinHF);
    const double sr = sqrt(uu * uu + vv * vv + a.eps_u * a.eps_u);
    double sh = 0.0;
    if (hf < kWaveHO) sh = sqrt(kWaveGHO * (kWaveHO - hf) * hf / kWaveHO);
    const bool shallow = hf < kWaveHO && sh > sr;
    const double sp = shallow ? sh : sr;
    const double sbx = acc_sx[l], sby = acc_sy[l];
    const double ubn = in.ubar_n[i], vbn = in.vbar_n[i];
    // u' = c (u/dt - g GX'/VOL) and u* = c (u/dt - g GX/VOL): du/dc = u/c
    double cbar = acc_c[l];
    if (c != 0.0)
      cbar += (ubn * in.u_n[i] + vbn * in.v_n[i] + sbx * a.usx[i] + sby * a.usy[i]) / c;
    // c = VOL / (VOL/dt + f VS)
    const double fbar = -c * c * vs / vol * cbar;
    const double vsbar = -c * c * f / vol * cbar;
    // f = 0.5 CF sp / hf
    const double cfbar = fbar * 0.5 * sp / hf;
    const double spbar = fbar * 0.5 * cf / hf;
    double hfbar = -fbar * f / hf;
    double ub = (ubn + sbx) * c / dt, vb = (vbn + sby) * c / dt;
    if (shallow) {
      hfbar += spbar * kWaveGHO * (kWaveHO - 2.0 * hf) / (2.0 * kWaveHO * sh);
    } else if (sr > 0.0) {
      ub += spbar * uu / sr;
      vb += spbar * vv / sr;
    }
    // CF = 2 g / K^2 / hc^(1/3)
    const double hcbar = -cfbar * cf / (3.0 * hc);
    out.kbar[i] += cfbar * (-2.0 * cf / k);
    const double hbar = (h > 0.0 ? acc_h[l] : 0.0) + (h > kWaveMinHC ? hcbar : 0.0) +
                        (h > kWaveMinHF ? hfbar : 0.0);
    ebar += hbar;
    out.zbbar[i] -= hbar;
    out.ubar[i] = ub;
    out.vbar[i] = vb;
    a.seb[i] = ebar;
    a.sbx[i] = sbx;
    a.sby[i] = sby;
    a.svs[i] = vsbar;
  }
}

// the transposes of GRADF (u* = c (u/dt - g GX/VOL): eta gets GRADF^T of
// -g c u*bar / VOL) and of the slope factor of VS (zb), from the scratch
// vectors k_wave_assemble_adj left; completes ebar and zbbar
__global__ void __launch_bounds__(kThreads)
k_wave_grad_adj(WaveArgs a, double* __restrict__ ebar, double* __restrict__ zbbar) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = wave_block(a);
  const int n0 = a.node_start[b], nown = a.node_start[b + 1] - n0;
  const int hb = a.halo_start[b];
  double* acc_e = reinterpret_cast<double*>(smem_raw);
  double* acc_z = acc_e + nown;
  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc_e[l] = acc_z[l] = 0.0;
  __syncthreads();
  colour_walk(a, b, [&](int s) {
    int l[3], g[3];
    slot_nodes(a, s, n0, nown, hb, l, g);
    const ElemGeo e = elem_geo(a, s);
    double sx = 0.0, sy = 0.0, sv = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = g[k];
      const double f = -a.grav * a.c[i] / a.vol[i];
      sx += f * a.sbx[i];
      sy += f * a.sby[i];
      sv += a.svs[i];
    }
    double d2, d3;
    (void)elem_slope(e, a.zb[g[0]], a.zb[g[1]], a.zb[g[2]], &d2, &d3);
    const double slb = sv * e.s / 3.0;
    const double zb3[3] = {-(d2 + d3) * slb, d2 * slb, d3 * slb};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (l[k] < nown) {
        acc_e[l[k]] += (e.bx[k] * sx + e.by[k] * sy) / 6.0;
        acc_z[l[k]] += zb3[k];
      }
    }
  });
  for (int l = threadIdx.x; l < nown; l += blockDim.x) {
    const int i = n0 + l;
    ebar[i] = a.seb[i] + acc_e[l];
    zbbar[i] += acc_z[l];
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
  a.zb = c->wv_zb.p;
  a.f = c->wv_f.p;
  a.vs = c->wv_vs.p;
  a.c = c->wv_c.p;
  a.usx = c->wv_usx.p;
  a.usy = c->wv_usy.p;
  a.fb = c->wv_fb.p;
  a.da_keep = c->wv_da.p;
  a.sbx = c->wv_sbx.p;
  a.sby = c->wv_sby.p;
  a.svs = c->wv_svs.p;
  a.seb = c->wv_seb.p;
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
  // k_wave_assemble_adj holds five accumulators per owned node in LDS
  if ((size_t)5 * L.max_owned * sizeof(double) > 64 * 1024) {
    fprintf(stderr, "libbiefhip: bhip_wave_setup: blocks of %d nodes need more than 64 KiB "
            "of LDS; use blocks of at most 1638 nodes\n", L.max_owned);
    return BHIP_EBADARG;
  }
  for (int e = 0; e < ne; ++e)
    if (!(surfac[e] > 0.0)) return BHIP_EBADARG;
  for (auto* v : {&c->wv_w, &c->wv_vol, &c->wv_h, &c->wv_zb, &c->wv_f, &c->wv_vs, &c->wv_c,
                  &c->wv_usx, &c->wv_usy, &c->wv_fb, &c->wv_da, &c->wv_sbx, &c->wv_sby,
                  &c->wv_svs, &c->wv_seb})
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
  hipLaunchKernelGGL(k_wave_grad, dim3(L.nblocks), dim3(kThreads),
                     3 * (size_t)L.max_owned * sizeof(double), c->stream, a, eta, u, v);
  hipLaunchKernelGGL(k_wave_assemble, dim3(L.nblocks), dim3(kThreads),
                     2 * (size_t)L.max_owned * sizeof(double), c->stream, a, c->da.p, c->xa.p,
                     rhs);
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
  const WaveArgs a = wave_args(c, p);
  hipLaunchKernelGGL(k_wave_assemble_adj, dim3(L.nblocks), dim3(kThreads),
                     5 * (size_t)L.max_owned * sizeof(double), c->stream, a, in, out);
  hipLaunchKernelGGL(k_wave_grad_adj, dim3(L.nblocks), dim3(kThreads),
                     2 * (size_t)L.max_owned * sizeof(double), c->stream, a, ebar, zbbar);
  HIP_OK(hipGetLastError());
  return BHIP_OK;
}

}  // extern "C"
