"""Time-stepped telemac2d wave equation on the GPU, with its adjoint.

The forward step is the wave-equation treatment of PROPAG (sources/telemac2d/
propag.F, TREATMENT OF THE LINEAR SYSTEM = 2) for implicitation 1
(TETAH = TETAU = TETAZCOMP = 1), no advection or diffusion of velocity, no
sources, no tidal flats, one sub-iteration, Strickler friction (KFROT 3,
HFROT 1), the free surface prescribed on sea nodes (free velocity there) and
walls elsewhere:

  friction  CF = 2 g / K^2 / max(h, 1e-4)^(1/3)   (friction_unif.f:124-127,
                                                    friction_calc.f:130-134)
            f  = CF |u| / (2 max(h, 1e-9)), |u| raised to the shallow-water
                 speed below 3 cm                  (fricti.f:123-135)
            AM2D = VOL/dt + f VS, VS the basis integrals weighted by the
                 bottom-slope factor               (slop10.f, propag.F:1155-1170)
            c = VOL / AM2D                         (DM1 = -g c, propag.F:1397)
  momentum  u* = c (u/dt - g GRADF(eta)/VOL)       (T2 = CV2/AM2D, propag.F:1413)
            u' = c (u/dt - g GRADF(eta')/VOL)      (propag.F:1685-1707)
  continuity, for the increment deta = eta' - eta:
            (M/dt + MATDIFUV(-1, DM1, max(h,0))) deta = -MATFGR(h) u*
            (mt01aa.f, mt02aa.f:200-240, mt08aa.f; propag.F:679, 1407, 1464)
  Dirichlet values of eta' on the sea boundary enter through DIRICH
  (dirich.f: b -= A f, rows/columns of prescribed nodes cleared).

with h = eta - zb. The linear system is solved by the HIP CG behind SOLVE
(bief.Context, diagonal preconditioning), its matrix-vector products by the
HIP EBE mat-vec. The step's operands are pinned against the reference's own
routines run in that sequence (the PROPAG golden vectors); the reference
differentiates telemac2d with the external dco/fortran tool (COMPAD,
ad_solve.F), which is absent, so the gradients are checked against central
finite differences and a tangent-linear run.

Differentiation: the linear solve and the EBE mat-vec are torch.autograd
Functions with hand-written adjoints (adjoint CG solve with the same
symmetric matrix; transposed EBE mat-vec; the element outer products for the
matrix entries) and tangent-linear rules. The time loop is a Function of its
own that stores the state every `checkpoint` steps and, going backward,
recomputes each segment from its checkpoint before pulling the adjoint state
through it.
"""
from __future__ import annotations

import numpy as np
import torch

from . import bief
from .mesh import Mesh

GRAV = 9.81


# ---------------------------------------------------------------------------
# EBE operators as autograd Functions
# ---------------------------------------------------------------------------
def _pair_outer(ikle, a, b):
    """Per-element symmetric pair products a_i b_j + a_j b_i for the three
    pairs (12, 13, 23) of each triangle: the gradient of x^T A y with respect
    to the off-diagonal coefficients of a symmetric EBE matrix."""
    i1, i2, i3 = ikle
    return torch.stack([a[i1] * b[i2] + a[i2] * b[i1],
                        a[i1] * b[i3] + a[i3] * b[i1],
                        a[i2] * b[i3] + a[i3] * b[i2]])


class _Ops:
    """Holds the context and connectivity the Functions need."""

    def __init__(self, ctx: bief.Context, ikle: torch.Tensor, solve_kw: dict):
        self.ctx = ctx
        self.ikle = ikle
        self.solve_kw = solve_kw
        self.n_solves = 0
        self.n_iters = 0
        # the last solution of each kind of solve (forward, adjoint,
        # tangent): the initial guess of the next one of that kind, since
        # successive time steps give close solutions
        self.guess = {}

    def reset_guesses(self):
        self.guess = {}

    def set(self, da, xa):
        self.ctx.set_matrix("Q", "S", da.contiguous(), xa.contiguous())

    def solve(self, da, xa, b, kind="forward"):
        """A^-1 b. The right-hand side is normalised first: GRACJG's stop
        test turns absolute when ||b'|| < 1 (XL = max(||b'||^2, 1),
        gracjg.F:189-209), and adjoint and tangent right-hand sides are
        small, so without it they would barely be solved. The previous
        solution of the same kind is the initial guess."""
        # on the device (no host round trip); a zero b stays zero and takes
        # the solver's zero right-hand-side exit
        scale = torch.clamp(torch.linalg.vector_norm(b), min=1e-300)
        self.set(da, xa)
        g = self.guess.get(kind)
        x = g / scale if g is not None and g.shape == b.shape else torch.zeros_like(b)
        _, info = self.ctx.solve(x, b / scale, slv=1, precon=2, **self.solve_kw)
        if not info["converged"]:
            raise bief.BiefError(f"wave-equation solve did not converge: {info}")
        self.n_solves += 1
        self.n_iters += info["niter"]
        x = x * scale
        self.guess[kind] = x.detach()
        return x

    def matvec(self, da, xa, y, trans=False):
        self.set(da, xa)
        return self.ctx.matvec("X=TAY" if trans else "X=AY", y.contiguous())


def _tangent_product(ops, dda, dxa, x):
    """(dA) x for a tangent matrix (dda, dxa), either part possibly absent."""
    if dda is None and dxa is None:
        return torch.zeros_like(x)
    dda = torch.zeros_like(x) if dda is None else dda
    dxa = torch.zeros((3, ops.ikle.shape[1]), dtype=x.dtype, device=x.device) if dxa is None \
        else dxa
    return ops.matvec(dda, dxa, x)


class EbeSolve(torch.autograd.Function):
    """x = A^-1 b for the symmetric EBE matrix (da, xa): BIEF SOLVE (CG,
    diagonal preconditioning) on the GPU.
    Adjoint: A lam = gx (A symmetric), db = lam, dA = -lam x^T on the
    element pattern. Tangent: A dx = db - (dA) x."""

    @staticmethod
    def forward(ctx, ops, da, xa, b):
        x = ops.solve(da, xa, b)
        ctx.ops = ops
        ctx.save_for_backward(da, xa, x)
        ctx.save_for_forward(da, xa, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        ops = ctx.ops
        da, xa, x = ctx.saved_tensors
        lam = ops.solve(da, xa, gx.contiguous(), kind="adjoint")
        return None, -lam * x, -_pair_outer(ops.ikle, lam, x), lam

    @staticmethod
    def jvp(ctx, _ops_t, dda, dxa, db):
        ops = ctx.ops
        da, xa, x = ctx.saved_tensors
        rhs = -_tangent_product(ops, dda, dxa, x)
        if db is not None:
            rhs = rhs + db
        return ops.solve(da, xa, rhs.contiguous(), kind="tangent")


class EbeMatvec(torch.autograd.Function):
    """y = A x for the symmetric EBE matrix (da, xa) (MATVEC 'X=AY');
    its adjoint is the transposed product 'X=TAY'."""

    @staticmethod
    def forward(ctx, ops, da, xa, x):
        ctx.ops = ops
        ctx.save_for_backward(da, xa, x)
        ctx.save_for_forward(da, xa, x)
        return ops.matvec(da, xa, x)

    @staticmethod
    def backward(ctx, gy):
        ops = ctx.ops
        da, xa, x = ctx.saved_tensors
        gx = ops.matvec(da, xa, gy.contiguous(), trans=True)
        return None, gy * x, _pair_outer(ops.ikle, gy, x), gx

    @staticmethod
    def jvp(ctx, _ops_t, dda, dxa, dx):
        ops = ctx.ops
        da, xa, x = ctx.saved_tensors
        out = _tangent_product(ops, dda, dxa, x)
        if dx is not None:
            out = out + ops.matvec(da, xa, dx.contiguous())
        return out


# ---------------------------------------------------------------------------
# the model
# ---------------------------------------------------------------------------
class WaveModel:
    """Wave-equation time stepping on a triangle mesh, on the GPU.

    sea: node indices with the prescribed tide eta = amp sin(2 pi t / period).
    """

    def __init__(self, mesh: Mesh, sea, dt=2.0, amp=0.2, period=600.0, hmin_speed=1e-3,
                 precision=bief.PREC_FP64, eps=1e-10, device="cuda"):
        if not torch.cuda.is_available():
            raise bief.BiefError("the wave model runs on the GPU")
        self.mesh = mesh
        self.dev = torch.device(device)
        self.dt, self.amp, self.period = float(dt), float(amp), float(period)
        self.eps_u = float(hmin_speed)
        td = lambda a, dt_=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt_,
                                                           device=self.dev)
        self.ikle = td(mesh.ikle, torch.long)
        xe, ye = mesh.xel_yel()
        self.s = td(mesh.surfac())
        self.x2, self.x3 = td(xe[1]), td(xe[2])
        self.y2, self.y3 = td(ye[1]), td(ye[2])
        self.n = mesh.npoin
        # basis-function integrals (VOLU2D)
        vol = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        for k in range(3):
            vol.index_add_(0, self.ikle[k], self.s / 3.0)
        self.vol = vol
        mask = torch.zeros(self.n, dtype=torch.bool, device=self.dev)
        mask[torch.as_tensor(np.asarray(sea), device=self.dev)] = True
        self.sea = mask
        self.ctx = bief.Context(mesh)
        self.ops = _Ops(self.ctx, self.ikle, {"eps": eps, "nitmax": 5000, "precision": precision})
        # P1 gradient factors per element: dphi_k/dx = bx_k / (2S), dphi_k/dy = by_k / (2S)
        self.bx = torch.stack([self.y2 - self.y3, self.y3, -self.y2])
        self.by = torch.stack([self.x3 - self.x2, -self.x3, self.x2])

    # -- assembly helpers (torch, differentiable) ---------------------------
    def scatter(self, w):
        """ASSVEC of element terms w (3, ne) into a node vector."""
        out = torch.zeros(self.n, dtype=w.dtype, device=self.dev)
        for k in range(3):
            out = out.index_add(0, self.ikle[k], w[k])
        return out

    def system(self, h, cnode):
        """(da, xa) of AM1 = M/dt + MATDIFUV(-1, DM1 = -g c, h) (h = max(eta - zb, 0))."""
        i1, i2, i3 = self.ikle
        x2, x3, y2, y3, s = self.x2, self.x3, self.y2, self.y3, self.s
        # MT01AA (mt01aa.f:83-106)
        m_off = s * (1.0 / (12.0 * self.dt))
        # MT02AA 'UV' (mt02aa.f:200-240), U = g c, V = h, XMUL = 1
        u = GRAV * cnode
        g1, g2, g3 = h[i1], h[i2], h[i3]
        g123 = g1 + g2 + g3
        coef2 = u[i1] * (g1 + g123) + u[i2] * (g2 + g123) + u[i3] * (g3 + g123)
        coef = coef2 / (48.0 * s)
        a12 = (y3 * (y2 - y3) + x3 * (x2 - x3)) * coef
        a13 = -(y2 * (y2 - y3) + x2 * (x2 - x3)) * coef
        a23 = -(y2 * y3 + x2 * x3) * coef
        xa = torch.stack([a12 + m_off, a13 + m_off, a23 + m_off])
        t = torch.stack([-a12 - a13 + 2 * m_off, -a12 - a23 + 2 * m_off,
                         -a13 - a23 + 2 * m_off])
        return self.scatter(t), xa

    def tide(self, t):
        return self.amp * np.sin(2.0 * np.pi * t / self.period)

    def gradf(self, q):
        """GRADF of a node field, assembled (VECTOR 'GRADF', vc13aa.f, XMUL 1)."""
        i1, i2, i3 = self.ikle
        q1, q2, q3 = q[i1], q[i2], q[i3]
        gx = (self.y2 * (q1 - q3) + self.y3 * (q2 - q1)) / 6.0
        gy = (self.x2 * (q3 - q1) + self.x3 * (q1 - q2)) / 6.0
        return self.scatter(torch.stack([gx, gx, gx])), self.scatter(torch.stack([gy, gy, gy]))

    def slope_mass(self, zb):
        """Basis integrals weighted by the bottom-slope factor of each element
        (SLOP10, slop10.f; the masked VECTOR 'MASBAS' of propag.F:1155-1161)."""
        i1, i2, i3 = self.ikle
        z2, z3 = zb[i2] - zb[i1], zb[i3] - zb[i1]
        a = (self.x2 * self.y3 - self.x3 * self.y2) ** 2
        b = (self.y2 * z3 - z2 * self.y3) ** 2
        c = (self.x3 * z2 - z3 * self.x2) ** 2
        w = self.s / 3.0 * torch.sqrt((a + b + c) / a)
        return self.scatter(torch.stack([w, w, w]))

    def step(self, eta, u, v, k_str, zb, t_new):
        """One time step (eta, u, v) -> (eta', u', v') at time t_new."""
        dt = self.dt
        h = eta - zb
        hp = torch.clamp(h, min=0.0)
        hc = torch.clamp(h, min=1e-4)
        cf = 2.0 * GRAV / (k_str * k_str) / hc.pow(1.0 / 3.0)
        hf = torch.clamp(h, min=1e-9)
        sp = torch.sqrt(u * u + v * v + self.eps_u * self.eps_u)
        shallow = hf < 3e-2
        arg = torch.where(shallow, 9.81 * (3e-2 - hf) * hf / 3e-2, torch.ones_like(hf))
        sp = torch.where(shallow, torch.maximum(sp, torch.sqrt(arg)), sp)
        f = 0.5 * cf * sp / hf
        c = self.vol / (self.vol / dt + f * self.slope_mass(zb))
        da, xa = self.system(hp, c)
        gx, gy = self.gradf(eta)
        usx = c * (u / dt - GRAV * gx / self.vol)
        usy = c * (v / dt - GRAV * gy / self.vol)
        # -MATFGR(h) u*: r_k = bx_k Qx + by_k Qy, Qx = F123 sum(u*) + sum(F u*), F = h/24
        i1, i2, i3 = self.ikle
        f1, f2, f3 = hp[i1] / 24.0, hp[i2] / 24.0, hp[i3] / 24.0
        f123 = f1 + f2 + f3
        sx1, sx2, sx3 = usx[i1], usx[i2], usx[i3]
        sy1, sy2, sy3 = usy[i1], usy[i2], usy[i3]
        qx = f123 * (sx1 + sx2 + sx3) + f1 * sx1 + f2 * sx2 + f3 * sx3
        qy = f123 * (sy1 + sy2 + sy3) + f1 * sy1 + f2 * sy2 + f3 * sy3
        r_el = self.bx * qx + self.by * qy
        # DIRICH (diri01.f:94-139): prescribed increment fb on the sea
        # boundary, b -= A fb, sea rows D fb, their rows and columns cleared
        w = (~self.sea).to(eta.dtype)
        fb = torch.where(self.sea, self.tide(t_new) - eta, torch.zeros_like(eta))
        rhs = self.scatter(r_el) - EbeMatvec.apply(self.ops, da, xa, fb)
        rhs = torch.where(self.sea, da * fb, rhs)
        xa_d = xa * torch.stack([w[i1] * w[i2], w[i1] * w[i3], w[i2] * w[i3]])
        deta = EbeSolve.apply(self.ops, da, xa_d, rhs)
        eta_n = eta + deta
        gxn, gyn = self.gradf(eta_n)
        u_n = c * (u / dt - GRAV * gxn / self.vol)
        v_n = c * (v / dt - GRAV * gyn / self.vol)
        return eta_n, u_n, v_n

    def initial_state(self, zb):
        z = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        return z.clone(), z.clone(), z.clone()

    def run(self, k_str, zb, nsteps, gauges=None, obs=None):
        """Forward run from rest; returns the gauge series (nsteps, ngauge) or,
        with obs, the cost sum (eta_g - obs)^2."""
        self.ops.reset_guesses()
        eta, u, v = self.initial_state(zb)
        series = []
        cost = torch.zeros((), dtype=torch.float64, device=self.dev)
        for n in range(nsteps):
            eta, u, v = self.step(eta, u, v, k_str, zb, (n + 1) * self.dt)
            if gauges is not None:
                ge = eta[gauges]
                series.append(ge)
                if obs is not None:
                    cost = cost + ((ge - obs[n]) ** 2).sum()
        if obs is not None:
            return cost
        return torch.stack(series) if series else eta


class WaveCost(torch.autograd.Function):
    """J(K, zb) = sum over steps and gauges of (eta - obs)^2, with a
    hand-written reverse sweep: states stored every `checkpoint` steps, each
    segment recomputed from its checkpoint and differentiated step by step
    (adjoint solves and transposed mat-vecs inside), last segment first."""

    @staticmethod
    def forward(ctx, model, k_str, zb, nsteps, gauges, obs, checkpoint):
        model.ops.reset_guesses()
        last = ((nsteps - 1) // checkpoint) * checkpoint  # start of the last segment
        with torch.no_grad():
            eta, u, v = model.initial_state(zb)
            checkpoints = {0: (eta, u, v)}
            # the forward solve's initial guess at each checkpoint, so that a
            # recomputed segment repeats the forward sweep's solves exactly
            guesses = {0: None}
            cost = torch.zeros((), dtype=torch.float64, device=model.dev)
            for n in range(last):
                eta, u, v = model.step(eta, u, v, k_str, zb, (n + 1) * model.dt)
                cost = cost + ((eta[gauges] - obs[n]) ** 2).sum()
                if (n + 1) % checkpoint == 0:
                    checkpoints[n + 1] = (eta, u, v)
                    guesses[n + 1] = model.ops.guess.get("forward")
        # the last segment keeps its graph: the reverse sweep starts there
        # without recomputing it
        with torch.enable_grad():
            seg = _Segment(model, checkpoints[last], k_str, zb)
            jlast = seg.run(last, nsteps, gauges, obs)
        cost = cost + jlast.detach()
        ctx.model, ctx.nsteps, ctx.gauges, ctx.obs = model, nsteps, gauges, obs
        ctx.checkpoints, ctx.checkpoint, ctx.last_segment = checkpoints, checkpoint, seg
        ctx.guesses = guesses
        ctx.save_for_backward(k_str, zb)
        return cost

    @staticmethod
    def backward(ctx, gcost):
        model, nsteps, gauges, obs = ctx.model, ctx.nsteps, ctx.gauges, ctx.obs
        k_str, zb = ctx.saved_tensors
        gk = torch.zeros_like(k_str)
        gz = torch.zeros_like(zb)
        # adjoint of the state at the end of the segment being processed
        lam = [torch.zeros_like(zb) for _ in range(3)]
        model.ops.guess.pop("adjoint", None)
        for s0 in sorted(ctx.checkpoints, reverse=True):
            s1 = min(s0 + ctx.checkpoint, nsteps)
            with torch.enable_grad():
                if ctx.last_segment is not None:
                    seg, ctx.last_segment = ctx.last_segment, None
                else:
                    # the segment recomputed from its checkpoint, with its
                    # graph, its first solve from the forward initial guess
                    if ctx.guesses[s0] is None:
                        model.ops.guess.pop("forward", None)
                    else:
                        model.ops.guess["forward"] = ctx.guesses[s0]
                    seg = _Segment(model, ctx.checkpoints[s0], k_str, zb)
                    seg.run(s0, s1, gauges, obs)
                g = seg.grad(lam, gcost)
            lam = [gi if gi is not None else torch.zeros_like(zb) for gi in g[:3]]
            if g[3] is not None:
                gk += g[3]
            if g[4] is not None:
                gz += g[4]
        return None, gk, gz, None, None, None, None


class _Segment:
    """Steps s0..s1-1 from a stored state, with their autograd graph."""

    def __init__(self, model, state, k_str, zb):
        self.model = model
        self.e0, self.u0, self.v0 = (t.detach().requires_grad_(True) for t in state)
        self.kk = k_str.detach().requires_grad_(True)
        self.zz = zb.detach().requires_grad_(True)

    def run(self, s0, s1, gauges, obs):
        m = self.model
        e, u, v = self.e0, self.u0, self.v0
        j = torch.zeros((), dtype=torch.float64, device=m.dev)
        for n in range(s0, s1):
            e, u, v = m.step(e, u, v, self.kk, self.zz, (n + 1) * m.dt)
            j = j + ((e[gauges] - obs[n]) ** 2).sum()
        self.outs = (e, u, v, j)
        return j

    def grad(self, lam, gcost):
        """Adjoints of the segment's start state and of K, zb, given those
        of its end state (lam) and of the cost."""
        return torch.autograd.grad(self.outs, (self.e0, self.u0, self.v0, self.kk, self.zz),
                                   (lam[0], lam[1], lam[2], gcost), allow_unused=True)


def cost_and_gradient(model: WaveModel, k_str, zb, nsteps, gauges, obs, checkpoint=10):
    """(J, dJ/dK, dJ/dzb) through the checkpointed adjoint."""
    k = k_str.detach().requires_grad_(True)
    z = zb.detach().requires_grad_(True)
    j = WaveCost.apply(model, k, z, nsteps, gauges, obs, checkpoint)
    gk, gz = torch.autograd.grad(j, (k, z))
    return j.detach(), gk, gz


# ---------------------------------------------------------------------------
# the same model on the HIP step kernels, with the hand-written adjoint sweep
# ---------------------------------------------------------------------------
class DeviceWaveModel:
    """The wave-equation model of WaveModel with each step in three HIP
    launches (bhip_wave_assemble: node quantities + system + right-hand side
    written straight into the context; SOLVE on internal-order vectors;
    bhip_wave_momentum) and its adjoint in the transposed kernels
    (bhip_wave_momentum_adj, adjoint SOLVE, bhip_wave_assemble_adj). State
    vectors stay in the context's internal node order for the whole run;
    inputs and gradients are converted at the ends.

    cost_and_gradient keeps the state every `checkpoint` steps; going
    backward it recomputes each segment from its checkpoint, keeping that
    segment's states (the last segment's come from the forward run), and
    pulls the adjoint state through it step by step."""

    def __init__(self, mesh: Mesh, sea, dt=2.0, amp=0.2, period=600.0, hmin_speed=1e-3,
                 precision=bief.PREC_FP64, eps=1e-10, device="cuda", block_nodes=0):
        if not torch.cuda.is_available():
            raise bief.BiefError("the wave model runs on the GPU")
        self.mesh = mesh
        self.dev = torch.device(device)
        self.dt, self.amp, self.period = float(dt), float(amp), float(period)
        self.eps_u = float(hmin_speed)
        self.n = mesh.npoin
        # block_nodes: nodes per workgroup of the context (0: the library's
        # default); a mesh of a few hundred thousand nodes fills the GPU
        # better with smaller blocks (tune.block_nodes measures it)
        self.ctx = bief.Context(mesh, block_nodes=block_nodes)
        xe, ye = mesh.xel_yel()
        sea_mask = np.zeros(mesh.npoin, dtype=np.int32)
        sea_mask[np.asarray(sea)] = 1
        self.kern = bief.WaveStep(self.ctx, xe, ye, mesh.surfac(), sea_mask)
        perm = self.ctx.perm()
        iperm = np.empty_like(perm)
        iperm[perm] = np.arange(perm.size, dtype=np.int32)
        self.perm = torch.as_tensor(perm.astype(np.int64), device=self.dev)
        self.iperm = torch.as_tensor(iperm.astype(np.int64), device=self.dev)
        self.solve_kw = {"eps": eps, "nitmax": 5000, "precision": precision}
        self.guess = {}
        self.n_solves = 0
        self.n_iters = 0
        # torch's element-wise work runs on the context's own stream, so the
        # library calls need no event handshakes (inputs and results are
        # ordered with the caller's stream at the public methods)
        self.stream = self.ctx.torch_stream()

    # -- node order ----------------------------------------------------------
    def to_internal(self, t):
        return t[self.perm].contiguous()

    def to_caller(self, t):
        out = torch.empty_like(t)
        out[self.perm] = t
        return out

    def _zeros(self):
        return torch.zeros(self.n, dtype=torch.float64, device=self.dev)

    def _out(self):
        """A node vector that a kernel writes in full (no fill needed)."""
        return torch.empty(self.n, dtype=torch.float64, device=self.dev)

    def tide(self, t):
        return self.amp * np.sin(2.0 * np.pi * t / self.period)

    def _params(self, t_new):
        return bief.WaveStep.params(self.dt, GRAV, self.eps_u, self.tide(t_new))

    def _solve(self, rhs, kind):
        """A^-1 rhs on the context's current matrix; right-hand side
        normalised (GRACJG's stop test is absolute below ||b'|| = 1,
        gracjg.F:189-209), previous solution of the same kind as initial
        guess."""
        scale = torch.clamp(torch.linalg.vector_norm(rhs), min=1e-300)
        g = self.guess.get(kind)
        x = g / scale if g is not None else self._zeros()
        b = rhs / scale
        _, info = self.ctx.solve_internal(x, b, slv=1, precon=2, **self.solve_kw)
        if not info["converged"]:
            raise bief.BiefError(f"wave-equation solve did not converge: {info}")
        self.n_solves += 1
        self.n_iters += info["niter"]
        x = x * scale
        self.guess[kind] = x
        return x

    # -- forward -------------------------------------------------------------
    def step(self, eta, u, v, k_str, zb, t_new):
        """(eta, u, v) -> (eta', u', v', deta), internal order."""
        p = self._params(t_new)
        rhs = self._out()
        self.kern.assemble(p, eta, u, v, k_str, zb, rhs)
        deta = self._solve(rhs, "forward")
        eta_n, u_n, v_n = self._out(), self._out(), self._out()
        self.kern.momentum(p, eta, deta, u, v, eta_n, u_n, v_n)
        return eta_n, u_n, v_n, deta

    def _on_stream(self, fn, *args):
        """fn(*args) on the context's stream, ordered after the caller's
        stream's work and before what the caller queues next; returned
        tensors are marked as used by the caller's stream."""
        caller = torch.cuda.current_stream()
        self.stream.wait_stream(caller)
        with torch.cuda.stream(self.stream):
            out = fn(*args)
        caller.wait_stream(self.stream)
        stack = [out]
        while stack:
            o = stack.pop()
            if torch.is_tensor(o):
                o.record_stream(caller)
            elif isinstance(o, (tuple, list)):
                stack.extend(o)
            elif isinstance(o, dict):
                stack.extend(o.values())
        return out

    def run(self, k_str, zb, nsteps, gauges=None, obs=None):
        """As WaveModel.run (caller-order inputs, gauges as caller node ids)."""
        return self._on_stream(self._run, k_str, zb, nsteps, gauges, obs)

    def _run(self, k_str, zb, nsteps, gauges, obs):
        self.guess = {}
        k, z = self.to_internal(k_str), self.to_internal(zb)
        g = self.iperm[gauges] if gauges is not None else None
        eta, u, v = self._zeros(), self._zeros(), self._zeros()
        series = []
        cost = torch.zeros((), dtype=torch.float64, device=self.dev)
        for n in range(nsteps):
            eta, u, v, _ = self.step(eta, u, v, k, z, (n + 1) * self.dt)
            if g is not None:
                ge = eta[g]
                series.append(ge)
                if obs is not None:
                    cost = cost + ((ge - obs[n]) ** 2).sum()
        if obs is not None:
            return cost
        return torch.stack(series) if series else self.to_caller(eta)

    # -- adjoint -------------------------------------------------------------
    def forward_sweep(self, k_str, zb, nsteps, gauges, obs, checkpoint=10):
        """J = sum over steps and gauges of (eta - obs)^2, and the tape of the
        reverse sweep: the state every `checkpoint` steps and the states and
        increments of the last segment."""
        return self._on_stream(self._forward_sweep, k_str, zb, nsteps, gauges, obs, checkpoint)

    def _forward_sweep(self, k_str, zb, nsteps, gauges, obs, checkpoint):
        self.guess = {}
        k, z = self.to_internal(k_str), self.to_internal(zb)
        g = self.iperm[gauges]
        last = ((nsteps - 1) // checkpoint) * checkpoint
        eta, u, v = self._zeros(), self._zeros(), self._zeros()
        checkpoints = {0: (eta, u, v)}
        # the forward solve's initial guess at each checkpoint (the previous
        # step's solution), so that a recomputed segment repeats the forward
        # sweep's solves exactly
        guesses = {0: None}
        seg_last = [(eta, u, v)] if last == 0 else None
        deta_last = []
        cost = torch.zeros((), dtype=torch.float64, device=self.dev)
        for n in range(nsteps):
            eta, u, v, deta = self.step(eta, u, v, k, z, (n + 1) * self.dt)
            cost = cost + ((eta[g] - obs[n]) ** 2).sum()
            if n + 1 == last:
                seg_last = [(eta, u, v)]
            elif n + 1 > last:
                seg_last.append((eta, u, v))
            if n >= last:
                deta_last.append(deta)
            if (n + 1) % checkpoint == 0 and n + 1 <= last:
                checkpoints[n + 1] = (eta, u, v)
                guesses[n + 1] = self.guess.get("forward")
        tape = {"k": k, "z": z, "g": g, "obs": obs, "nsteps": nsteps, "checkpoint": checkpoint,
                "checkpoints": checkpoints, "guesses": guesses, "last": last,
                "seg_last": (seg_last, deta_last)}
        return cost, tape

    def reverse_sweep(self, tape, gcost=1.0):
        """(dJ/dK, dJ/dzb) times gcost, caller order, from forward_sweep's tape."""
        return self._on_stream(self._reverse_sweep, tape, gcost)

    def _reverse_sweep(self, tape, gcost):
        k, z, g, obs = tape["k"], tape["z"], tape["g"], tape["obs"]
        nsteps, checkpoint = tape["nsteps"], tape["checkpoint"]
        kbar, zbar = self._zeros(), self._zeros()
        lam_e, lam_u, lam_v = self._zeros(), self._zeros(), self._zeros()
        self.guess.pop("adjoint", None)
        for s0 in sorted(tape["checkpoints"], reverse=True):
            s1 = min(s0 + checkpoint, nsteps)
            if s0 == tape["last"]:
                states, detas = tape["seg_last"]
            else:
                # the segment recomputed from its checkpoint, its first solve
                # from the forward sweep's initial guess there
                states, detas = [tape["checkpoints"][s0]], []
                if tape["guesses"][s0] is None:
                    self.guess.pop("forward", None)
                else:
                    self.guess["forward"] = tape["guesses"][s0]
                for n in range(s0, s1):
                    e, uu, vv, d = self.step(*states[-1], k, z, (n + 1) * self.dt)
                    states.append((e, uu, vv))
                    detas.append(d)
            for n in range(s1 - 1, s0 - 1, -1):
                (e0, u0, v0), (e1, u1, v1) = states[n - s0], states[n - s0 + 1]
                p = self._params((n + 1) * self.dt)
                ebar_n = lam_e.clone()
                ebar_n.index_add_(0, g, (2.0 * gcost) * (e1[g] - obs[n]))
                # the step's system again, then the transposed step
                self.kern.assemble(p, e0, u0, v0, k, z, self._out())
                etot = self._out()
                self.kern.momentum_adj(p, lam_u, lam_v, ebar_n, etot)
                lam = self._solve(etot, "adjoint")
                eb, ub, vb = self._out(), self._out(), self._out()
                self.kern.assemble_adj(p, lam, detas[n - s0], e0, u0, v0, u1, v1, lam_u, lam_v,
                                       etot, k, eb, ub, vb, kbar, zbar)
                lam_e, lam_u, lam_v = eb, ub, vb
        return self.to_caller(kbar), self.to_caller(zbar)

    def cost_and_gradient(self, k_str, zb, nsteps, gauges, obs, checkpoint=10):
        """(J, dJ/dK, dJ/dzb) through the checkpointed adjoint sweep."""
        cost, tape = self.forward_sweep(k_str, zb, nsteps, gauges, obs, checkpoint)
        gk, gz = self.reverse_sweep(tape)
        return cost, gk, gz


class DeviceWaveCost(torch.autograd.Function):
    """J(K, zb) of DeviceWaveModel behind autograd: forward is the
    checkpointed forward sweep, backward the hand-written reverse sweep."""

    @staticmethod
    def forward(ctx, model, k_str, zb, nsteps, gauges, obs, checkpoint):
        cost, tape = model.forward_sweep(k_str.detach(), zb.detach(), nsteps, gauges, obs,
                                         checkpoint)
        ctx.model, ctx.tape = model, tape
        return cost

    @staticmethod
    def backward(ctx, gcost):
        gk, gz = ctx.model.reverse_sweep(ctx.tape, gcost)
        ctx.tape = None
        return None, gk, gz, None, None, None, None
