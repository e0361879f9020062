"""Synthetic meshes and fields for the BIEF linear-system path.

A :class:`Mesh` carries exactly what BIEF's ``BIEF_MESH`` provides to the
solver path (``sources/utils/bief/bief_def.f:205-498``): node count NPOIN,
element count NELEM, the connectivity IKLE as structure of arrays
(``ikle[k, e]``, 0-based here), node coordinates X/Y, the element-local
coordinates XEL/YEL with local node 1 at the origin
(``sources/utils/bief/inbief.f:384-397``) and the element areas SURFAC
(``sources/utils/bief/surv11.f``).

Generators:
  * :func:`channel_mesh`  -- the 2000 m x 200 m structured channel
    (400 x 50 cells, 2 triangles per cell: 20,451 nodes, 40,000 elements);
  * :func:`estuary_mesh`  -- seeded Poisson-disk points in an irregular
    50 km x 30 km estuary polygon with a refined 2 km channel, Delaunay
    triangulated;
  * :func:`extrude_prisms` -- sigma extrusion of a triangle mesh into
    6-node prisms (telemac3d's layout: plane-major node numbering).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


@dataclass
class Mesh:
    x: np.ndarray            # (npoin,) float64
    y: np.ndarray            # (npoin,) float64
    ikle: np.ndarray         # (ndp, nelem) int32, 0-based
    z: np.ndarray | None = None   # prisms: node elevations
    nplan: int = 1
    npoin2: int = 0
    nelem2: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def ndp(self) -> int:
        return int(self.ikle.shape[0])

    @property
    def npoin(self) -> int:
        return int(self.x.shape[0])

    @property
    def nelem(self) -> int:
        return int(self.ikle.shape[1])

    def xel_yel(self):
        """Element-local coordinates, node 1 at the origin (inbief.f:384-397)."""
        xe = self.x[self.ikle].astype(np.float64)
        ye = self.y[self.ikle].astype(np.float64)
        xe = xe - xe[0:1]
        ye = ye - ye[0:1]
        return np.ascontiguousarray(xe), np.ascontiguousarray(ye)

    def surfac(self):
        """Triangle areas 0.5*(X2*Y3 - X3*Y2) (surv11.f); for prisms, the
        area of the base triangle."""
        xe, ye = self.xel_yel()
        return 0.5 * (xe[1] * ye[2] - xe[2] * ye[1])


def channel_mesh_n(npoin_target: int, seed: int = 0, length=2000.0, width=200.0,
                   jitter: float = 0.2) -> Mesh:
    """A length x width channel (the 20,451-node channel's shape) with about
    npoin_target nodes: a perturbed structured grid, each cell split along a
    random diagonal. Built without Python loops (ensembles of many meshes)."""
    rng = np.random.default_rng(seed)
    ny = max(2, int(round(np.sqrt(npoin_target * width / length))) - 1)
    nx = max(2, int(round(npoin_target / (ny + 1))) - 1)
    dx, dy = length / nx, width / ny
    X, Y = np.meshgrid(np.arange(nx + 1) * dx, np.arange(ny + 1) * dy, indexing="ij")
    inner_x = (X > 0) & (X < length - 1e-9)
    inner_y = (Y > 0) & (Y < width - 1e-9)
    X = X + jitter * dx * rng.uniform(-1, 1, X.shape) * inner_x
    Y = Y + jitter * dy * rng.uniform(-1, 1, Y.shape) * inner_y
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    a = (i * (ny + 1) + j).ravel()
    b = ((i + 1) * (ny + 1) + j).ravel()
    c, d = b + 1, a + 1
    flip = rng.uniform(size=a.size) < 0.5
    t1 = np.where(flip, np.stack([a, b, c]), np.stack([a, b, d]))
    t2 = np.where(flip, np.stack([a, c, d]), np.stack([b, c, d]))
    ikle = np.concatenate([t1, t2], axis=1).astype(np.int32)
    return Mesh(x=np.ascontiguousarray(X.ravel()), y=np.ascontiguousarray(Y.ravel()),
                ikle=np.ascontiguousarray(ikle), npoin2=X.size, nelem2=ikle.shape[1],
                meta={"kind": "channel", "seed": seed, "nx": nx, "ny": ny})


def segments(ikle):
    """Edges of a triangle mesh for segment storage (MATRIX STORAGE = 3):
    gloseg (2, nseg), 0-based, first node < second node, in order of first
    appearance over the elements and their edges 12, 23, 31; eltseg (3, nelem)
    the segment of each element edge (12, 23, 31)."""
    ik = np.asarray(ikle, dtype=np.int64)
    ne = ik.shape[1]
    a = np.concatenate([ik[0], ik[1], ik[2]])
    b = np.concatenate([ik[1], ik[2], ik[0]])
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (int(ik.max()) + 1) + hi
    # element-major order of the edges (e, then local edge)
    order = np.arange(3 * ne).reshape(3, ne).T.ravel()
    uniq, first, inv = np.unique(key[order], return_index=True, return_inverse=True)
    rank = np.empty(uniq.size, dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(uniq.size)
    seg_of = np.empty(3 * ne, dtype=np.int64)
    seg_of[order] = rank[inv]
    nseg = uniq.size
    gloseg = np.empty((2, nseg), dtype=np.int64)
    gloseg[0, seg_of] = lo
    gloseg[1, seg_of] = hi
    return gloseg, seg_of.reshape(3, ne)


def segment_mesh(mesh: "Mesh"):
    """The segments of a triangle mesh as two-node elements (the layout of a
    matrix in segment storage) and the element-to-segment map."""
    gl, es = segments(mesh.ikle)
    seg = Mesh(x=mesh.x, y=mesh.y, ikle=gl, npoin2=mesh.npoin, nelem2=gl.shape[1],
               meta={"storage": "segments"})
    return seg, es


def triangle_areas(x, y, ikle):
    xe = x[ikle]
    ye = y[ikle]
    return 0.5 * ((xe[1] - xe[0]) * (ye[2] - ye[0]) - (xe[2] - xe[0]) * (ye[1] - ye[0]))


# --------------------------------------------------------------------------
# structured channel
# --------------------------------------------------------------------------

def channel_mesh(nx: int = 400, ny: int = 50, lx: float = 2000.0, ly: float = 200.0) -> Mesh:
    """Rectangular channel, each cell split into 2 triangles along the same
    diagonal. Node (i, j) -> i*(ny+1) + j (i along x)."""
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    x = X.ravel().copy()
    y = Y.ravel().copy()
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    a = i * (ny + 1) + j
    b = (i + 1) * (ny + 1) + j
    c = (i + 1) * (ny + 1) + j + 1
    d = i * (ny + 1) + j + 1
    t1 = np.stack([a, b, c])
    t2 = np.stack([a, c, d])
    ikle = np.empty((3, 2 * nx * ny), dtype=np.int32)
    ikle[:, 0::2] = t1
    ikle[:, 1::2] = t2
    return Mesh(x=x, y=y, ikle=ikle, npoin2=x.size, nelem2=ikle.shape[1],
                meta={"kind": "channel", "nx": nx, "ny": ny})


def channel_depth(mesh: Mesh) -> np.ndarray:
    """Depth 1-5 m on a linear slope plus a Gaussian bump of 1.5 m,
    sigma = 50 m, centred at mid-channel (x=1000, y=100)."""
    x, y = mesh.x, mesh.y
    return (1.0 + 4.0 * x / 2000.0
            + 1.5 * np.exp(-((x - 1000.0) ** 2 + (y - 100.0) ** 2) / (2.0 * 50.0 ** 2)))


def random_small_mesh(nx: int, ny: int, seed: int = 0, jitter: float = 0.25) -> Mesh:
    """Small perturbed structured mesh with randomly oriented diagonals and a
    random node and element numbering (for parity tests)."""
    rng = np.random.default_rng(seed)
    xs = np.arange(nx + 1, dtype=np.float64)
    ys = np.arange(ny + 1, dtype=np.float64)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    X = X + jitter * rng.uniform(-1, 1, X.shape) * (0 < X) * (X < nx)
    Y = Y + jitter * rng.uniform(-1, 1, Y.shape) * (0 < Y) * (Y < ny)
    x = X.ravel()
    y = Y.ravel()
    tris = []
    for i in range(nx):
        for j in range(ny):
            a = i * (ny + 1) + j
            b = (i + 1) * (ny + 1) + j
            c = (i + 1) * (ny + 1) + j + 1
            d = i * (ny + 1) + j + 1
            if rng.uniform() < 0.5:
                tris += [(a, b, c), (a, c, d)]
            else:
                tris += [(a, b, d), (b, c, d)]
    ikle = np.array(tris, dtype=np.int64).T
    # random node renumbering and element order, random local rotation
    perm = rng.permutation(x.size)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(x.size)
    x = x[perm]
    y = y[perm]
    ikle = inv[ikle]
    ikle = ikle[:, rng.permutation(ikle.shape[1])]
    rot = rng.integers(0, 3, ikle.shape[1])
    for r in (1, 2):
        sel = rot == r
        ikle[:, sel] = np.roll(ikle[:, sel], r, axis=0)
    return Mesh(x=np.ascontiguousarray(x), y=np.ascontiguousarray(y),
                ikle=np.ascontiguousarray(ikle.astype(np.int32)),
                npoin2=x.size, nelem2=ikle.shape[1],
                meta={"kind": "random", "seed": seed})


# --------------------------------------------------------------------------
# unstructured estuary
# --------------------------------------------------------------------------

def _meshgen_lib():
    path = os.path.join(_HERE, "libbiefmesh.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    lib.meshgen_poisson_disk.restype = ctypes.c_long
    lib.meshgen_poisson_disk.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return lib


def estuary_polygon(lx: float = 50000.0, ly: float = 30000.0, seed: int = 0, n: int = 96):
    """Irregular estuary outline: a funnel widening towards the sea (x=lx)
    with a seeded wavy coastline."""
    rng = np.random.default_rng(seed + 17)
    t = np.linspace(0.0, 1.0, n // 2)
    amp = rng.uniform(0.02, 0.05, 4)
    ph = rng.uniform(0, 2 * math.pi, 4)
    wave = sum(amp[k] * np.sin((k + 2) * math.pi * t + ph[k]) for k in range(4))
    half = 0.5 * ly * (0.35 + 0.65 * t ** 0.7) * (1.0 + wave)
    half = np.minimum(half, 0.5 * ly)
    xs = t * lx
    top_x, top_y = xs, 0.5 * ly + half
    bot_x, bot_y = xs[::-1], (0.5 * ly - half)[::-1]
    px = np.concatenate([top_x, bot_x])
    py = np.concatenate([top_y, bot_y])
    return px, py


def points_in_polygon(x, y, px, py):
    """Even-odd ray casting, vectorised over the query points."""
    x = np.asarray(x)
    y = np.asarray(y)
    inside = np.zeros(x.shape, dtype=bool)
    n = len(px)
    j = n - 1
    for i in range(n):
        yi, yj, xi, xj = py[i], py[j], px[i], px[j]
        cond = (yi > y) != (yj > y)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (xj - xi) * (y - yi) / (yj - yi) + xi
        inside ^= cond & (x < xint)
        j = i
    return inside


def channel_centerline(x, lx=50000.0, ly=30000.0):
    return 0.5 * ly + 0.12 * ly * np.sin(2 * math.pi * x / (0.6 * lx))


def estuary_mesh(npoin_target: int = 2_000_000, seed: int = 0, lx: float = 50000.0,
                 ly: float = 30000.0, channel_width: float = 2000.0,
                 refine: float = 4.0) -> Mesh:
    """Seeded Poisson-disk (Bridson) points with `refine` times the point
    density inside a meandering channel of width `channel_width`, Delaunay
    triangulated (scipy's Qhull), triangles outside the polygon removed."""
    from scipy.spatial import Delaunay

    lib = _meshgen_lib()
    px, py = estuary_polygon(lx, ly, seed)
    gnx, gny = 501, 301
    gx = np.linspace(px.min(), px.max(), gnx)
    gy = np.linspace(py.min(), py.max(), gny)
    GX, GY = np.meshgrid(gx, gy)
    dist = np.abs(GY - channel_centerline(GX, lx, ly))
    s = np.clip((dist - 0.5 * channel_width) / (0.5 * channel_width), 0.0, 1.0)
    fine = 1.0 / math.sqrt(refine)
    fgrid = np.ascontiguousarray(fine + (1.0 - fine) * (3 * s ** 2 - 2 * s ** 3))
    # area-weighted estimate of the spacing giving npoin_target points
    ins = points_in_polygon(GX.ravel(), GY.ravel(), px, py)
    cell_area = (gx[1] - gx[0]) * (gy[1] - gy[0])
    eff_area = float(np.sum(cell_area / fgrid.ravel()[ins] ** 2))
    # this sampler fills about 0.641 points per r^2 of effective area
    r0 = math.sqrt(0.6412 * eff_area / npoin_target)
    pxc = np.ascontiguousarray(px)
    pyc = np.ascontiguousarray(py)

    def spacing(xq, yq):
        ix = np.clip(((xq - gx[0]) / (gx[-1] - gx[0]) * (gnx - 1)).astype(int), 0, gnx - 1)
        iy = np.clip(((yq - gy[0]) / (gy[-1] - gy[0]) * (gny - 1)).astype(int), 0, gny - 1)
        return fgrid[iy, ix]

    def outline(r0):
        # outline samples at the local spacing, vertices included once
        bx, by = [], []
        for i in range(px.size):
            x0, y0, x1, y1 = px[i], py[i], px[(i + 1) % px.size], py[(i + 1) % px.size]
            length = math.hypot(x1 - x0, y1 - y0)
            if length == 0:
                continue
            h = r0 * float(spacing(np.array([0.5 * (x0 + x1)]), np.array([0.5 * (y0 + y1)]))[0])
            k = max(1, int(math.ceil(length / (1.05 * h))))
            t = np.arange(k) / k
            bx.append(x0 + t * (x1 - x0))
            by.append(y0 + t * (y1 - y0))
        return np.ascontiguousarray(np.concatenate(bx)), np.ascontiguousarray(np.concatenate(by))

    for _ in range(3):
        cap = int(npoin_target * 1.3) + 1000
        ox = np.empty(cap)
        oy = np.empty(cap)
        ix, iy = outline(r0)
        n = lib.meshgen_poisson_disk(pxc.ctypes.data, pyc.ctypes.data, px.size, r0,
                                     fgrid.ctypes.data, gnx, gny, seed, 30,
                                     ix.ctypes.data, iy.ctypes.data, ix.size,
                                     ox.ctypes.data, oy.ctypes.data, cap)
        if n < 0:
            r0 *= 1.1
            continue
        if abs(n - npoin_target) <= 0.005 * npoin_target:
            break
        r0 *= math.sqrt(n / npoin_target)
    x = ox[:n].copy()
    y = oy[:n].copy()
    tri = Delaunay(np.stack([x, y], 1))
    ikle = tri.simplices.T.astype(np.int64)
    cx = x[ikle].mean(0)
    cy = y[ikle].mean(0)
    keep = points_in_polygon(cx, cy, px, py)
    ikle = ikle[:, keep]
    # drop degenerate slivers created along the concave outline
    area = triangle_areas(x, y, ikle)
    keep = area > 1e-6 * np.median(area)
    ikle = ikle[:, keep]
    used = np.zeros(n, dtype=bool)
    used[ikle.ravel()] = True
    remap = -np.ones(n, dtype=np.int64)
    remap[used] = np.arange(used.sum())
    x = x[used]
    y = y[used]
    ikle = remap[ikle]
    # counter-clockwise orientation
    area = triangle_areas(x, y, ikle)
    neg = area < 0
    ikle[1, neg], ikle[2, neg] = ikle[2, neg].copy(), ikle[1, neg].copy()
    # element order: by smallest node id (Poisson-disk ids follow the front)
    order = np.argsort(ikle.min(0), kind="stable")
    ikle = np.ascontiguousarray(ikle[:, order].astype(np.int32))
    return Mesh(x=x, y=y, ikle=ikle, npoin2=x.size, nelem2=ikle.shape[1],
                meta={"kind": "estuary", "seed": seed, "r0": r0})


def smooth_random_depth(mesh: Mesh, seed: int = 0, hmin: float = 0.5, hmax: float = 40.0,
                        ngauss: int = 32) -> np.ndarray:
    """Depth from a seeded sum of Gaussians, scaled into [hmin, hmax]."""
    rng = np.random.default_rng(seed + 101)
    x, y = mesh.x, mesh.y
    lx = x.max() - x.min()
    ly = y.max() - y.min()
    f = np.zeros_like(x)
    for _ in range(ngauss):
        cx = x.min() + rng.uniform() * lx
        cy = y.min() + rng.uniform() * ly
        s = rng.uniform(0.05, 0.25) * max(lx, ly)
        a = rng.uniform(0.2, 1.0)
        f += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    f = (f - f.min()) / max(f.max() - f.min(), 1e-30)
    return hmin + (hmax - hmin) * f


# --------------------------------------------------------------------------
# prisms
# --------------------------------------------------------------------------

def extrude_prisms(base: Mesh, nplan: int, bottom: np.ndarray, surface: np.ndarray) -> Mesh:
    """Sigma extrusion: plane p (0 = bottom) of 2D node i is 3D node
    p*npoin2 + i; prism k of triangle e spans planes k, k+1 with local nodes
    (1,2,3) below and (4,5,6) above, element id k*nelem2 + e (telemac3d's
    numbering)."""
    n2 = base.npoin
    e2 = base.nelem
    sig = np.linspace(0.0, 1.0, nplan)
    z = (bottom[None, :] + sig[:, None] * (surface - bottom)[None, :]).ravel()
    x = np.tile(base.x, nplan)
    y = np.tile(base.y, nplan)
    ik = np.empty((6, e2 * (nplan - 1)), dtype=np.int64)
    for k in range(nplan - 1):
        sl = slice(k * e2, (k + 1) * e2)
        ik[0:3, sl] = base.ikle + k * n2
        ik[3:6, sl] = base.ikle + (k + 1) * n2
    return Mesh(x=x, y=y, ikle=np.ascontiguousarray(ik.astype(np.int32)), z=z,
                nplan=nplan, npoin2=n2, nelem2=e2,
                meta={"kind": "prism", "base": base.meta})
