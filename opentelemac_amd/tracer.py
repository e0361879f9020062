"""The SUPG tracer system of CVDFTR on the device.

sources/utils/bief/cvdftr.F:398-560 builds, for an advected and diffused
tracer with classical SUPG (OPTADV = 1):

  TE  = KSUPG(1, U)                                  SUPG element vectors
  AM2 = MATVGR(1, U) + MASUPG(1, TE, U)               advection + SUPG
  AM1 = X(MATMAS(1/dt)) + T(MATVGR(1/dt, TE as P0))   mass + SUPG test term
  SM  = AM1 (dt SM0 + FN) + (TETAT - 1) AM2 FN
  AM1 = AM1 + TETAT AM2

then adds the diffusion (MATDIF, cvdftr.F:694-733; here implicit with an
isotropic P1 coefficient), applies DIRICH and calls SOLVE with a
non-symmetric solver (GMRES). Every matrix here is built on the device:
the element kernels of MT01AA, MT02AA, MT03AA, MT05AA (with KSUPG computed
in place) and the OM operations work on the context's slot matrix and its
saved copy, in the reference's order. The mesh geometry (XEL, YEL, SURFAC)
is put on the device once per context and read from there.
"""
from __future__ import annotations

import numpy as np

from . import bief


def supg_system(ctx: bief.Context, mesh, u, v, nu, dt, fn, tetat=1.0, sm0=None):
    """Assembles AM1 into the context (non-symmetric, 'Q') and returns the
    right-hand side SM. u, v, nu, fn: node arrays (caller order); sm0: the
    source term SM before the step (zero by default). Host arrays give a host
    SM; float64 torch CUDA tensors keep every operand on the device (the
    element matrices are built from them on the context stream) and give a
    device SM."""
    if not getattr(ctx, "has_geometry", False):
        xe, ye = mesh.xel_yel()
        ctx.set_geometry(xe, ye, mesh.surfac())
    G = (None, None, None)  # the geometry kept on the device
    dev = bief._is_torch(u)
    if not dev:
        fn = np.ascontiguousarray(fn, dtype=np.float64)
    K = bief.Context.SUPG_KSUPG
    # AM2: centred advection plus the SUPG term (cvdftr.F:398-409)
    ctx.matrix_supg(5, 1.0, *G, u, v, op="M=N")
    ctx.matrix_supg(3, 1.0, *G, u, v, flags=K, op="M=M+N")
    t2 = ctx.matvec("X=AY", fn)
    ctx.save_matrix()
    # AM1: mass over dt made non-symmetric, plus the transposed SUPG test
    # term with the element vectors as a P0 velocity (cvdftr.F:448, 513-523)
    ctx.matrix_mt01aa(1.0 / dt, None)
    ctx.matrix_om("M=X(M)")
    ctx.matrix_supg(5, 1.0 / dt, *G, u, v, flags=K, op="M=M+TN")
    rhs = fn if sm0 is None else dt * (sm0 if dev else np.asarray(sm0, dtype=np.float64)) + fn
    sm = ctx.matvec("X=AY", rhs)
    sm = sm + (tetat - 1.0) * t2
    ctx.matrix_om("M=M+CN", tetat)
    # implicit diffusion (MATDIF, P1 isotropic)
    ctx.matrix_mt02aa(1, 1.0, *G, nu, accumulate=True)
    return sm


def estuary_flow(mesh, seed=0, umax=2.0):
    """The synthetic tracer case on an estuary mesh: a divergence-free
    velocity (u, v) = (dpsi/dy, -dpsi/dx) of a streamfunction of 16 seeded
    Gaussians, scaled to a largest speed of umax; a tracer of 8 Gaussian
    patches; and the sea-boundary nodes (x within 1 m of the largest x),
    where the tracer is prescribed. Returns (u, v, fn, sea)."""
    from . import system as S
    rng = np.random.default_rng(seed + 7)
    x, y = mesh.x, mesh.y
    lx, ly = x.max() - x.min(), y.max() - y.min()
    u = np.zeros(mesh.npoin)
    v = np.zeros(mesh.npoin)
    for _ in range(16):
        cx, cy = x.min() + rng.uniform() * lx, y.min() + rng.uniform() * ly
        sg = rng.uniform(0.05, 0.2) * min(lx, ly)
        a = rng.normal()
        e = a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * sg * sg))
        u += -e * (y - cy) / (sg * sg)
        v += e * (x - cx) / (sg * sg)
    sc = umax / np.hypot(u, v).max()
    u, v = u * sc, v * sc
    fn = np.zeros(mesh.npoin)
    for _ in range(8):
        cx, cy = x.min() + rng.uniform() * lx, y.min() + rng.uniform() * ly
        fn += np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * (0.03 * lx) ** 2))
    bnd = S.boundary_nodes(mesh)
    sea = bnd[mesh.x[bnd] > mesh.x.max() - 1.0]
    return u, v, fn, sea
