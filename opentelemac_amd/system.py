"""Host-side (numpy) assembly of the telemac2d wave-equation system that
PROPAG hands to SOLVE (sources/telemac2d/propag.F, TREATMENT OF THE LINEAR
SYSTEM = 2), used to prepare solver inputs for tests and benchmarks.

The element matrices follow BIEF's closed forms:
  * MT01AA mass (sources/utils/bief/mt01aa.f:83-106): off-diagonal S/12,
    diagonal S/6, times XMUL;
  * MT02AA 'UV' diffusion (mt02aa.f:200-240): the stiffness K_ij scaled by
    the exact integral of the P1 product U*V over the element, times XMUL.
The assembled system for the depth increment DH is (propag.F:664-716,
1369-1407)
    AM1 = M/dt + MATDIFUV(XMUL=-1, U=DM1, V=HPROP),
    DM1 = -theta_u theta_h g V2DPAR / AM2D,   AM2D = VOLU2D/dt + friction,
i.e. A = M/dt + theta_u theta_h g int H/(1/dt + f) grad(phi_i).grad(phi_j).
Dirichlet values enter through DIRICH (dirich.f).
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh

GRAV = 9.81


def mt01aa(surfac, xmul=1.0):
    """(T diagonal (3,ne), XM off-diagonal (3,ne)) of the P1 mass matrix."""
    det = surfac * (xmul / 12.0)
    xm = np.stack([det, det, det])
    t = np.stack([det + det, det + det, det + det])
    return t, xm


def mt02aa_uv(xel, yel, surfac, ikle, u, v, xmul=1.0):
    """P1 diffusion with coefficient U*V (mt02aa.f:200-240)."""
    x2, x3 = xel[1], xel[2]
    y2, y3 = yel[1], yel[2]
    i1, i2, i3 = ikle
    g1, g2, g3 = v[i1], v[i2], v[i3]
    coef1 = (xmul / 48.0) / surfac
    g123 = g1 + g2 + g3
    coef2 = u[i1] * (g1 + g123) + u[i2] * (g2 + g123) + u[i3] * (g3 + g123)
    a12 = (y3 * (y2 - y3) + x3 * (x2 - x3)) * coef2 * coef1
    a13 = -(y2 * (y2 - y3) + x2 * (x2 - x3)) * coef2 * coef1
    a23 = -(y2 * y3 + x2 * x3) * coef2 * coef1
    xm = np.stack([a12, a13, a23])
    t = np.stack([-a12 - a13, -a12 - a23, -a13 - a23])
    return t, xm


def assemble(t, ikle, npoin):
    """Assembled diagonal of element diagonal terms (ASSVEC, INIT=.TRUE.)."""
    d = np.zeros(npoin)
    for k in range(t.shape[0]):
        d += np.bincount(ikle[k], weights=t[k], minlength=npoin)
    return d


def volu2d(mesh: Mesh):
    """Integral of each basis function: assembled S/3 (VECTOR 'MASBAS')."""
    s = mesh.surfac()
    return assemble(np.stack([s / 3.0] * 3), mesh.ikle, mesh.npoin)


def strickler_cf(h, k_strickler, hmin=1e-3):
    """Friction coefficient CF = 2 g / (K^2 h^(1/3)) (friction_calc.f:130-134)."""
    hc = np.maximum(h, hmin)
    return 2.0 * GRAV / (k_strickler ** 2 * hc ** (1.0 / 3.0))


def wave_matrix(mesh: Mesh, depth, dt=1.0, theta_u=1.0, theta_h=1.0, friction=None):
    """EBE matrix (DA, XA[3,ne], 'Q', 'S') of the wave-equation system.

    friction: optional per-node term added to AM2%D before inversion, i.e.
    the projected 0.5*CF*|u|/h (fricti.f:129-137) times the basis mass."""
    s = mesh.surfac()
    xel, yel = mesh.xel_yel()
    t1, x1 = mt01aa(s, 1.0 / dt)
    vol = volu2d(mesh)
    am2d = vol / dt
    if friction is not None:
        am2d = am2d + friction
    am2d_inv = np.where(np.abs(am2d) > 1e-6, 1.0 / np.where(am2d == 0, 1, am2d), 0.0)
    dm1 = -theta_u * theta_h * GRAV * vol * am2d_inv
    t2, x2 = mt02aa_uv(xel, yel, s, mesh.ikle, dm1, depth, xmul=-1.0)
    da = assemble(t1 + t2, mesh.ikle, mesh.npoin)
    xa = x1 + x2
    return da, np.ascontiguousarray(xa), "Q", "S"


def ebe_to_segments(xa, eltseg, ikle, gloseg, typext="S"):
    """Assembles element off-diagonal terms into segment storage (the
    operation of ASSEX3, assex3.f): XA(s) = sum over the elements holding
    segment s of their term for that edge. 'S': xa (3, ne) pairs 12, 13, 23
    -> (1, nseg). 'Q': xa (6, ne) with the transposed pairs 21, 31, 32 ->
    (2, nseg), XA1 = A(g1, g2), XA2 = A(g2, g1)."""
    nseg = gloseg.shape[1]
    # element edge k (12, 23, 31) -> pair column and its local nodes
    cols = ((0, 0, 1), (2, 1, 2), (1, 2, 0))  # (column of pair (i<j), node a, node b)
    out = np.zeros((2 if typext == "Q" else 1, nseg))
    for k, (c, na, nb) in enumerate(cols):
        s = eltseg[k]
        if typext == "S":
            out[0] += np.bincount(s, xa[c], minlength=nseg)
        else:
            # xa[c] = A(i, j) for local i < j, xa[c + 3] = A(j, i)
            i_loc, j_loc = min(na, nb), max(na, nb)
            gi = ikle[i_loc]
            forward = gi == gloseg[0, s]  # A(i, j) is A(g1, g2)
            a_ij, a_ji = xa[c], xa[c + 3]
            out[0] += np.bincount(s, np.where(forward, a_ij, a_ji), minlength=nseg)
            out[1] += np.bincount(s, np.where(forward, a_ji, a_ij), minlength=nseg)
    return out


def boundary_nodes(mesh: Mesh):
    """Nodes on edges that belong to a single triangle."""
    ik = mesh.ikle
    e = np.concatenate([ik[[0, 1]], ik[[1, 2]], ik[[2, 0]]], axis=1)
    e = np.sort(e, axis=0)
    key = e[0].astype(np.int64) * mesh.npoin + e[1]
    uniq, cnt = np.unique(key, return_counts=True)
    b = uniq[cnt == 1]
    nodes = np.unique(np.concatenate([b // mesh.npoin, b % mesh.npoin]))
    return nodes


def tidal_dirichlet(mesh: Mesh, nodes, amplitude=0.2, period=600.0, t=150.0, kdir=5):
    """FBOR / LIMDIR of a sinusoidal tide imposed on `nodes`."""
    fbor = np.zeros(mesh.npoin)
    limdir = np.full(mesh.npoin, 2, dtype=np.int32)
    fbor[nodes] = amplitude * np.sin(2 * np.pi * t / period)
    limdir[nodes] = kdir
    return fbor, limdir


def dirich_host(da, xa, ikle, sm, fbor, limdir, kdir=5):
    """DIRICH (dirich.f -> diri01.f -> diraux.f) on host arrays; returns
    (f, da, xa, sm) with xa's rows and columns at prescribed nodes zeroed."""
    npoin = da.size
    # SM -= A FBOR  (A = diag + symmetric off-diagonal terms)
    ax = da * fbor
    i1, i2, i3 = ikle
    a12, a13, a23 = xa
    ax += np.bincount(i1, a12 * fbor[i2] + a13 * fbor[i3], minlength=npoin)
    ax += np.bincount(i2, a12 * fbor[i1] + a23 * fbor[i3], minlength=npoin)
    ax += np.bincount(i3, a13 * fbor[i1] + a23 * fbor[i2], minlength=npoin)
    sm = sm - ax
    w = np.ones(npoin)
    d = limdir == kdir
    sm = sm.copy()
    sm[d] = da[d] * fbor[d]
    w[d] = 0.0
    f = np.zeros(npoin)
    f[d] = fbor[d]
    xa = np.stack([a12 * w[i2] * w[i1], a13 * w[i3] * w[i1], a23 * w[i3] * w[i2]])
    return f, da.copy(), xa, sm


def drop_depth(mesh: Mesh):
    """Initial depth of the reference's gouttedo example, a Gaussian drop of
    2.4 m on 2.4 m of water centred in the 20.1 m basin
    (examples/telemac2d/gouttedo/user_fortran/condin.f:64-68)."""
    eikon = ((mesh.x - 10.05) ** 2 + (mesh.y - 10.05) ** 2) / 4.0
    return 2.4 * (1.0 + np.exp(-eikon))


def drop_system(mesh: Mesh, dt=0.04, theta=0.6):
    """First wave-equation step of the gouttedo example: closed basin (walls
    only, no Dirichlet node), time step 0.04 s and implicitation 0.6 for depth
    and velocity (examples/telemac2d/gouttedo/t2d_gouttedo.cas), linearised
    about the initial depth; the right-hand side is the mass term (M/dt) H0
    of the fluid at rest. Returns dict(da, xa, typdia, typext, b, x0)."""
    h0 = drop_depth(mesh)
    da, xa, td, te = wave_matrix(mesh, h0, dt, theta, theta)
    t1, x1 = mt01aa(mesh.surfac(), 1.0 / dt)
    i1, i2, i3 = mesh.ikle
    a12, a13, a23 = x1
    w = np.stack([t1[0] * h0[i1] + a12 * h0[i2] + a13 * h0[i3],
                  a12 * h0[i1] + t1[1] * h0[i2] + a23 * h0[i3],
                  a13 * h0[i1] + a23 * h0[i2] + t1[2] * h0[i3]])
    b = assemble(w, mesh.ikle, mesh.npoin)
    return {"da": da, "xa": xa, "typdia": td, "typext": te, "b": b, "x0": h0.copy()}


def wave_system(mesh: Mesh, depth, dt=1.0, theta=1.0, inflow=None, amplitude=0.2,
                period=600.0, t=150.0):
    """Matrix and right-hand side of one wave-equation step from rest with a
    tidal Dirichlet condition on `inflow` nodes (default: nodes at min x).
    Returns dict(da, xa, typdia, typext, b, x0, f)."""
    da, xa, td, te = wave_matrix(mesh, depth, dt, theta, theta)
    if inflow is None:
        inflow = np.nonzero(mesh.x <= mesh.x.min() + 1e-9)[0]
    fbor, limdir = tidal_dirichlet(mesh, inflow, amplitude, period, t)
    f, da, xa, b = dirich_host(da, xa, mesh.ikle, np.zeros(mesh.npoin), fbor, limdir)
    return {"da": da, "xa": xa, "typdia": td, "typext": te, "b": b, "x0": f,
            "fbor": fbor, "limdir": limdir}


def sea_boundary(mesh: Mesh):
    """Boundary nodes on the seaward edge (largest x) of an estuary mesh."""
    bnd = boundary_nodes(mesh)
    return bnd[mesh.x[bnd] > mesh.x.max() - 1.0]


def ensemble_sizes(nsys, seed=0, lo=5000, hi=200000):
    """Member sizes drawn log-uniformly from [lo, hi] nodes (seeded)."""
    rng = np.random.default_rng(seed + 3)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), nsys)).astype(int)


def estuary_member(npoin, seed, dt=1.0):
    """One member of an ensemble of estuaries: the unstructured estuary mesh
    generator at npoin points with its own seed (outline, points, depth), a
    wave-equation step (dt, theta = 1) with a 0.2 m tide on its sea boundary.
    Returns (mesh, system dict)."""
    from . import mesh as M
    m = M.estuary_mesh(int(npoin), seed=int(seed))
    depth = M.smooth_random_depth(m, seed=int(seed))
    return m, wave_system(m, depth, dt=dt, theta=1.0, inflow=sea_boundary(m))


def estuary_ensemble(sizes, seed=0, dt=1.0, workers=None):
    """Members estuary_member(sizes[k], seed * 1000 + k) built by a pool of
    threads (the sampler, the triangulation and numpy release the GIL; no
    process is forked, so this is safe once a GPU is in use)."""
    from concurrent.futures import ThreadPoolExecutor
    import os
    n = len(sizes)
    if workers is None:
        workers = max(1, min(n, os.cpu_count() or 1, 16))
    args = [(int(s), seed * 1000 + k) for k, s in enumerate(sizes)]
    # largest members first, so the pool does not end on one big mesh
    order = np.argsort([-a[0] for a in args], kind="stable")
    out = [None] * n
    with ThreadPoolExecutor(workers) as ex:
        futs = {int(k): ex.submit(estuary_member, args[k][0], args[k][1], dt) for k in order}
        for k, f in futs.items():
            out[k] = f.result()
    return [o[0] for o in out], [o[1] for o in out]
