// SUPG element matrices (P1 triangles) on the device and the OM operations
// that CVDFTR combines them with. Included at the end of bief_hip.hip.
//
//   MASUPG, MT03AA (mt03aa.f): XMUL int (F,G).grad(phi_i) (U.grad(phi_j)),
//           F, G per element, U, V per node                         'Q'
//   MAUGUG, MT04AA (mt04aa.f): XMUL int (U.grad(phi_i)) (U.grad(phi_j)) 'S'
//   MATVGR, MT05AA (mt05aa.f): XMUL int phi_i U.grad(phi_j), U, V per node
//           or per element; the FORMUL(16:16) = 'N' upwind variant  'Q'
//   KSUPG  (ksupg.f -> kspg11.f): the SUPG element vectors, computed in
//           place from the nodal velocities when the caller asks for them
//
// Each slot computes its element's 3x3 matrix (every instance of an element
// computes the same values), writes the off-diagonal terms into the slot
// matrix and assembles the diagonal onto owned nodes in LDS, colour by
// colour (the deterministic scheme of the mat-vec). Element formulas keep
// the reference's order of operations.

namespace bhip {

enum { SUPG_U_P0 = 1, SUPG_UPWIND = 2, SUPG_KSUPG = 4 };

// kspg11.f for one element: the SUPG vector (kx, ky)
__device__ inline void kspg11_elem(double x2, double x3, double y2, double y3, double u1,
                                   double u2, double u3, double v1, double v2, double v3,
                                   double xmul, double& kx, double& ky) {
  double g1x = y2 - y3, g1y = x3 - x2;
  double r = 1.0 / sqrt(g1x * g1x + g1y * g1y);
  g1x *= r;
  g1y *= r;
  double g2x = y3, g2y = -x3;
  r = 1.0 / sqrt(g2x * g2x + g2y * g2y);
  g2x *= r;
  g2y *= r;
  double g3x = -y2, g3y = x2;
  r = 1.0 / sqrt(g3x * g3x + g3y * g3y);
  g3x *= r;
  g3y *= r;
  const double c3 = sqrt(x2 * x2 + y2 * y2);
  const double c1 = sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
  const double c2 = sqrt(x3 * x3 + y3 * y3);
  const double s = 0.5 * (x2 * y3 - x3 * y2);
  const double h1 = 2 * s / c1, h2 = 2 * s / c2, h3 = 2 * s / c3;
  double h = fmax(fmax(h1, h2), h3);
  const double um = u1 + u2 + u3, vm = v1 + v2 + v3;
  const double rn = 1.0 / fmax(sqrt(um * um + vm * vm), 1e-10);
  const double un = um * rn, vn = vm * rn;
  if ((g1x * un + g1y * vn) * h > h1) h = h1;
  if ((g2x * un + g2y * vn) * h > h2) h = h2;
  if ((g3x * un + g3y * vn) * h > h3) h = h3;
  // the reference's literal
  kx = 0.33333333 * xmul * h * un;
  ky = 0.33333333 * xmul * h * vn;
}

struct SupgArgs {
  int kind;  // 3 MT03AA, 4 MT04AA, 5 MT05AA
  int flags;
  double xmul;
  const double* surfac;  // per element (caller order)
  const double* xel;     // [3][nelem]
  const double* yel;
  const double* geo;     // or the context's slot-ordered geometry (then the three above unused)
  int nelem;
  const double* f;  // per element, MT03AA (unless SUPG_KSUPG)
  const double* g;
  const double* u;  // per node (internal order) or per element
  const double* v;
  const int32_t* slot_elem;
  const uint16_t* lid;
  int nslot_pad;
  const int32_t* node_start;
  const int32_t* halo_start;
  const int32_t* halo;
  const int32_t* color_off;
  int nblocks;
  double* xa;  // slot matrix, 6 columns ('Q') or 3 ('S')
  double* da;
  int op;        // 0 'M=N', 1 'M=M+N', 2 'M=M+TN'
  int target_q;  // the target matrix stores 6 columns
};

__global__ void __launch_bounds__(kThreads) k_elem_supg(SupgArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* acc = reinterpret_cast<double*>(smem_raw);
  const int b = xcd_block(blockIdx.x, a.nblocks);
  const int n0 = a.node_start[b], nown = a.node_start[b + 1] - n0;
  const int hb = a.halo_start[b];
  for (int l = threadIdx.x; l < nown; l += blockDim.x) acc[l] = 0.0;
  __syncthreads();
  const int32_t* co = a.color_off + (size_t)b * (kMaxColors + 1);
  const int send = co[kMaxColors];
  const size_t np = a.nslot_pad, ne = a.nelem;
  for (int col = 0; col < kMaxColors; ++col) {
    const int s0 = co[col];
    if (s0 >= send) break;
    const int s1 = co[col + 1];
    for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
      const int e = a.slot_elem[s];
      int l[3], g[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        l[k] = a.lid[(size_t)k * np + s];
        g[k] = l[k] < nown ? n0 + l[k] : a.halo[hb + l[k] - nown];
      }
      const SlotGeo G = load_geo(a.geo, np, s, a.xel, a.yel, a.surfac, ne, e, true);
      const double x2e = G.x2, x3e = G.x3;
      const double y2e = G.y2, y3e = G.y3;
      const bool p0 = (a.flags & SUPG_U_P0) && !(a.flags & SUPG_KSUPG);
      double u1 = 0, u2 = 0, u3 = 0, v1 = 0, v2 = 0, v3 = 0;
      if (!p0) {
        u1 = a.u[g[0]];
        u2 = a.u[g[1]];
        u3 = a.u[g[2]];
        v1 = a.v[g[0]];
        v2 = a.v[g[1]];
        v3 = a.v[g[2]];
      }
      // m[i][j] = A(i, j) of the element
      double m[3][3];
      if (a.kind == 3) {
        double ff, gg;
        if (a.flags & SUPG_KSUPG)
          kspg11_elem(x2e, x3e, y2e, y3e, u1, u2, u3, v1, v2, v3, 1.0, ff, gg);
        else {
          ff = a.f[e];
          gg = a.g[e];
        }
        const double den = (a.xmul / 12.0) / G.s;
        const double kkx = ff * den, kky = gg * den;
        const double u123 = u1 + u2 + u3, v123 = v1 + v2 + v3;
        const double p1 = x2e * v123 - y2e * u123;
        const double p2 = x3e * v123 - y3e * u123;
        const double q1 = x2e * kky - y2e * kkx;
        const double q2 = x3e * kky - y3e * kkx;
        m[0][0] = (p1 - p2) * (q1 - q2);
        m[1][1] = p2 * q2;
        m[0][1] = p2 * (q1 - q2);
        m[1][0] = (p1 - p2) * q2;
        m[0][2] = -m[0][0] - m[0][1];
        m[1][2] = -m[1][1] - m[1][0];
        m[2][0] = -m[0][0] - m[1][0];
        m[2][1] = -m[1][1] - m[0][1];
        m[2][2] = -m[0][2] - m[1][2];
      } else if (a.kind == 4) {
        const double c = (a.xmul / 48.0) / G.s;
        const double u123 = u1 + u2 + u3, v123 = v1 + v2 + v3;
        const double uu = u1 * u123 + u2 * u2 + u2 * u3 + u3 * u3;
        const double uv = u1 * (v123 + v1) + u2 * (v123 + v2) + u3 * (v123 + v3);
        const double vv = v1 * v123 + v2 * v2 + v2 * v3 + v3 * v3;
        const double b12 = (2 * y3e * (y2e - y3e) * uu +
                            (x3e * (y3e - y2e) + (x3e - x2e) * y3e) * uv +
                            2 * x3e * (x2e - x3e) * vv) * c;
        const double b13 = (2 * y2e * (y3e - y2e) * uu +
                            (2 * x2e * y2e - x2e * y3e - x3e * y2e) * uv +
                            2 * x2e * (x3e - x2e) * vv) * c;
        const double b23 =
            (-(2 * y2e * y3e * uu) + (x2e * y3e + x3e * y2e) * uv - 2 * x2e * x3e * vv) * c;
        m[0][1] = m[1][0] = b12;
        m[0][2] = m[2][0] = b13;
        m[1][2] = m[2][1] = b23;
        m[0][0] = -b12 - b13;
        m[1][1] = -b23 - b12;
        m[2][2] = -b13 - b23;
      } else if (a.flags & (SUPG_U_P0 | SUPG_KSUPG)) {
        // MT05AA with U, V per element (given, or the SUPG vectors)
        double ue, ve;
        if (a.flags & SUPG_KSUPG)
          kspg11_elem(x2e, x3e, y2e, y3e, u1, u2, u3, v1, v2, v3, 1.0, ue, ve);
        else {
          ue = a.u[e];
          ve = a.v[e];
        }
        const double sur24 = a.xmul / 24.0;
        const double x2 = x2e * sur24, x3 = x3e * sur24, y2 = y2e * sur24, y3 = y3e * sur24;
        const double qu = 4 * ue, qv = 4 * ve;
        const double r1 = (x3 - x2) * qv + (y2 - y3) * qu;
        const double r2 = -(x3 * qv) + y3 * qu;
        const double r3 = x2 * qv - y2 * qu;
        m[0][0] = m[1][0] = m[2][0] = r1;
        m[1][1] = m[0][1] = m[2][1] = r2;
        m[2][2] = m[0][2] = m[1][2] = r3;
      } else if (a.flags & SUPG_UPWIND) {
        const double sur6 = a.xmul / 6.0;
        const double us = (u1 + u2 + u3) * sur6, vs = (v1 + v2 + v3) * sur6;
        const double k1 = us * (y2e - y3e) - vs * (x2e - x3e);
        const double k2 = us * y3e - vs * x3e;
        const double k3 = us * (-y2e) - vs * (-x2e);
        const double l12 = fmax(fmin(k1, -k2), 0.0), l13 = fmax(fmin(k1, -k3), 0.0);
        const double l21 = fmax(fmin(k2, -k1), 0.0), l23 = fmax(fmin(k2, -k3), 0.0);
        const double l31 = fmax(fmin(k3, -k1), 0.0), l32 = fmax(fmin(k3, -k2), 0.0);
        m[0][0] = l12 + l13;
        m[1][1] = l21 + l23;
        m[2][2] = l31 + l32;
        m[0][1] = -l12;
        m[0][2] = -l13;
        m[1][2] = -l23;
        m[1][0] = -l21;
        m[2][0] = -l31;
        m[2][1] = -l32;
      } else {
        const double sur24 = a.xmul / 24.0;
        const double x2 = x2e * sur24, x3 = x3e * sur24, y2 = y2e * sur24, y3 = y3e * sur24;
        const double u123 = u1 + u2 + u3, v123 = v1 + v2 + v3;
        const double w1u = u123 + u1, w1v = v123 + v1;
        const double w2u = u123 + u2, w2v = v123 + v2;
        const double w3u = u123 + u3, w3v = v123 + v3;
        m[0][0] = (x3 - x2) * w1v + (y2 - y3) * w1u;
        m[1][1] = -(x3 * w2v) + y3 * w2u;
        m[2][2] = x2 * w3v - y2 * w3u;
        m[0][1] = -(x3 * w1v) + y3 * w1u;
        m[0][2] = x2 * w1v - y2 * w1u;
        m[1][2] = x2 * w2v - y2 * w2u;
        m[1][0] = (x3 - x2) * w2v + (y2 - y3) * w2u;
        m[2][0] = (x3 - x2) * w3v + (y2 - y3) * w3u;
        m[2][1] = -(x3 * w3v) + y3 * w3u;
      }
      // 'M=M+TN' adds the transpose
      const bool tr = a.op == 2;
      const double c12 = tr ? m[1][0] : m[0][1], c21 = tr ? m[0][1] : m[1][0];
      const double c13 = tr ? m[2][0] : m[0][2], c31 = tr ? m[0][2] : m[2][0];
      const double c23 = tr ? m[2][1] : m[1][2], c32 = tr ? m[1][2] : m[2][1];
      const double cv[6] = {c12, c13, c23, c21, c31, c32};
      const int ncol = a.target_q ? 6 : 3;
      for (int c = 0; c < ncol; ++c) {
        double* p = a.xa + (size_t)c * np + s;
        *p = a.op == 0 ? cv[c] : *p + cv[c];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (l[k] < nown) acc[l[k]] += m[k][k];
    }
    __syncthreads();
  }
  for (int l = threadIdx.x; l < nown; l += blockDim.x)
    a.da[n0 + l] = a.op == 0 ? acc[l] : a.da[n0 + l] + acc[l];
}

// OM on the context's matrix M and its saved copy N (slot layout):
//   0 'M=X(M)'  : the symmetric terms copied into the transposed half
//   1 'M=CN'    : M = c M (N = M, as CVDFTR calls it)
//   2 'M=M+CN'  : M = M + c N
//   3 'M=M+TN'  : M = M + N^T
__global__ void k_om_slots(int op, double c, double* __restrict__ xa,
                           const double* __restrict__ xn, int m_q, int n_q, int nslot_pad,
                           long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int col = (int)(i / nslot_pad);  // 0..5 of the 'Q' target
    const long s = i - (long)col * nslot_pad;
    if (op == 0) {
      if (col >= 3) xa[i] = xa[(long)(col - 3) * nslot_pad + s];
    } else if (op == 1) {
      xa[i] = c * xa[i];
    } else {
      // the N term at this position (symmetric N stores 3 columns)
      int src = col;
      if (op == 3) src = col < 3 ? col + 3 : col - 3;
      if (!n_q) src = src % 3;
      const double nv = xn[(long)src * nslot_pad + s];
      xa[i] = op == 2 ? xa[i] + c * nv : xa[i] + nv;
    }
  }
}

__global__ void k_om_diag(int op, double c, double* __restrict__ da,
                          const double* __restrict__ dn, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (op == 1)
      da[i] = c * da[i];
    else if (op == 2)
      da[i] = da[i] + c * dn[i];
    else if (op == 3)
      da[i] = da[i] + dn[i];
  }
}

}  // namespace bhip

extern "C" {

static int supg_impl(bhip_ctx* c, int kind, double xmul, const double* xel, const double* yel,
                     const double* surfac, const double* f, const double* g, const double* u,
                     const double* v, int flags, int op) {
  if (!c || c->L.ndp != 3 || !u || !v) return BHIP_EBADARG;
  // XEL, YEL, SURFAC all NULL: the context's geometry (bhip_ctx_set_geometry)
  const bool use_geo = !surfac && !xel && !yel;
  if (use_geo ? !c->have_geo : (!surfac || !xel || !yel)) return BHIP_EBADARG;
  if (kind != 3 && kind != 4 && kind != 5) return BHIP_EBADARG;
  if (op < 0 || op > 2 || (flags & ~(SUPG_U_P0 | SUPG_UPWIND | SUPG_KSUPG))) return BHIP_EBADARG;
  if (kind == 3 && !(flags & SUPG_KSUPG) && (!f || !g)) return BHIP_EBADARG;
  if (kind != 5 && (flags & (SUPG_U_P0 | SUPG_UPWIND))) return BHIP_EBADARG;
  if (kind == 4 && (flags & SUPG_KSUPG)) return BHIP_EBADARG;
  if ((flags & SUPG_UPWIND) && (flags & (SUPG_U_P0 | SUPG_KSUPG))) return BHIP_EBADARG;
  const Layout& L = c->L;
  const int n = L.npoin, ne = L.nelem;
  if (op != 0 && (!c->have_matrix || c->typdia != 'Q' || c->typext == '0')) return BHIP_EBADARG;
  // the result's storage: MT03AA and MT05AA are not symmetric; MT04AA is
  const bool sym_src = kind == 4;
  bool target_q;
  if (op == 0)
    target_q = !sym_src;
  else
    target_q = c->typext == 'Q' || !sym_src || op == 2;
  if (!c->have_matrix) {
    HIP_OK(c->da.alloc(n));
    HIP_OK(c->xa.alloc((size_t)6 * L.nslot_pad));
  }
  if (op != 0 && target_q && c->typext == 'S') {
    // OM 'M=X(M)' first: the target becomes non-symmetric
    hipLaunchKernelGGL(k_om_slots, dim3(grid_for((int)(6 * (long)L.nslot_pad))), dim3(kThreads), 0,
                       c->stream, 0, 0.0, c->xa.p, (const double*)nullptr, 1, 1, L.nslot_pad,
                       (long)6 * L.nslot_pad);
    HIP_OK(hipGetLastError());
  }
  const bool u_p0 = (flags & SUPG_U_P0) && !(flags & SUPG_KSUPG);
  // the caller's geometry arrays (host calls only): temporaries
  DevBuf<double> dS, dX, dY;
  if (!use_geo) {
    HIP_OK(dS.alloc(ne));
    HIP_OK(dX.alloc((size_t)3 * ne));
    HIP_OK(dY.alloc((size_t)3 * ne));
    HIP_OK(hipMemcpyAsync(dS.p, surfac, sizeof(double) * ne, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(dX.p, xel, sizeof(double) * 3 * ne, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(dY.p, yel, sizeof(double) * 3 * ne, hipMemcpyHostToDevice, c->stream));
  }
  const double *dF = nullptr, *dG = nullptr, *dU = nullptr, *dV = nullptr;
  if (kind == 3 && !(flags & SUPG_KSUPG)) {
    if (!(dF = elem_in(c, 0, f)) || !(dG = elem_in(c, 1, g))) return BHIP_EDEVICE;
  }
  if (u_p0) {
    if (!(dU = elem_in(c, 2, u)) || !(dV = elem_in(c, 3, v))) return BHIP_EDEVICE;
  } else {
    if (!(dU = node_in(c, 0, u)) || !(dV = node_in(c, 1, v))) return BHIP_EDEVICE;
  }
  SupgArgs a;
  a.kind = kind;
  a.flags = flags;
  a.xmul = xmul;
  a.surfac = dS.p;
  a.xel = dX.p;
  a.yel = dY.p;
  a.geo = use_geo ? c->geo.p : nullptr;
  a.nelem = ne;
  a.f = dF;
  a.g = dG;
  a.u = dU;
  a.v = dV;
  a.slot_elem = c->slot_elem.p;
  a.lid = c->lid.p;
  a.nslot_pad = L.nslot_pad;
  a.node_start = c->node_start.p;
  a.halo_start = c->halo_start.p;
  a.halo = c->halo.p;
  a.color_off = c->color_off.p;
  a.nblocks = L.nblocks;
  a.xa = c->xa.p;
  a.da = c->da.p;
  a.op = op;
  a.target_q = target_q ? 1 : 0;
  hipLaunchKernelGGL(k_elem_supg, dim3(L.nblocks), dim3(kThreads),
                     (size_t)L.max_owned * sizeof(double), c->stream, a);
  HIP_OK(hipGetLastError());
  if (!use_geo) HIP_OK(hipStreamSynchronize(c->stream));
  if (input_done(c)) return BHIP_EDEVICE;
  c->typdia = 'Q';
  c->typext = target_q ? 'Q' : 'S';
  c->ncoef = target_q ? 6 : 3;
  c->have_matrix = true;
  return BHIP_OK;
}

int bhip_matrix_supg(bhip_ctx* c, int kind, double xmul, const double* xel, const double* yel,
                     const double* surfac, const double* f, const double* g, const double* u,
                     const double* v, int flags, int op) {
  if (!c) return BHIP_EBADARG;
  InputMode mode(c, false);
  return supg_impl(c, kind, xmul, xel, yel, surfac, f, g, u, v, flags, op);
}

int bhip_matrix_supg_dev(bhip_ctx* c, int kind, double xmul, const double* f_dev,
                         const double* g_dev, const double* u_dev, const double* v_dev, int flags,
                         int op) {
  if (!c) return BHIP_EBADARG;
  InputMode mode(c, true);
  return supg_impl(c, kind, xmul, nullptr, nullptr, nullptr, f_dev, g_dev, u_dev, v_dev, flags,
                   op);
}

int bhip_matrix_om(bhip_ctx* c, const char* op, double cc) {
  if (!c || !op || !c->have_matrix || c->L.ndp != 3) return BHIP_EBADARG;
  const Layout& L = c->L;
  const std::string o(op, strnlen(op, 8));
  int code;
  if (o.rfind("M=X(M)", 0) == 0)
    code = 0;
  else if (o.rfind("M=CN", 0) == 0)
    code = 1;
  else if (o.rfind("M=M+CN", 0) == 0)
    code = 2;
  else if (o.rfind("M=M+TN", 0) == 0)
    code = 3;
  else if (o.rfind("M=M+N", 0) == 0) {
    code = 2;
    cc = 1.0;
  } else
    return BHIP_EBADARG;
  if (code == 0) {
    if (c->typext == 'Q') return BHIP_OK;
    if (c->typext != 'S') return BHIP_EBADARG;
  } else if (c->typdia != 'Q' || c->typext == '0') {
    return BHIP_EBADARG;
  }
  if (code >= 2 && (!c->have_saved || c->typdia_saved != 'Q' || c->typext_saved == '0'))
    return BHIP_EBADARG;
  const bool m_q = c->typext == 'Q';
  const bool n_q = code >= 2 && c->typext_saved == 'Q';
  // the result is non-symmetric when either operand is, or for M + N^T of a 'Q' N
  const bool out_q = code == 0 || m_q || n_q;
  if (out_q && !m_q && code != 0) {
    hipLaunchKernelGGL(k_om_slots, dim3(grid_for((int)(6 * (long)L.nslot_pad))), dim3(kThreads), 0,
                       c->stream, 0, 0.0, c->xa.p, (const double*)nullptr, 1, 1, L.nslot_pad,
                       (long)6 * L.nslot_pad);
  }
  const long ncol = out_q ? 6 : 3;
  hipLaunchKernelGGL(k_om_slots, dim3(grid_for((int)(ncol * L.nslot_pad))), dim3(kThreads), 0,
                     c->stream, code, cc, c->xa.p, c->xa_saved.p, 1, n_q ? 1 : 0, L.nslot_pad,
                     ncol * L.nslot_pad);
  if (code != 0)
    hipLaunchKernelGGL(k_om_diag, dim3(c->grid_nodes), dim3(kThreads), 0, c->stream, code, cc,
                       c->da.p, c->da_saved.p, L.npoin);
  HIP_OK(hipGetLastError());
  c->typext = out_q ? 'Q' : 'S';
  c->ncoef = out_q ? 6 : 3;
  return BHIP_OK;
}

}  // extern "C"
