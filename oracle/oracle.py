"""ctypes access to the CPU oracle (liboracle.so) and, where it was built,
to the reference's own BIEF (oracle/_ref/libbief_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbief_ref.so")

OPS = {"X=AY": 0, "X=CAY": 1, "X=-AY": 2, "X=X+AY": 3, "X=X-AY": 4, "X=X+CAY": 5,
       "X=TAY": 6, "X=-TAY": 7, "X=X+TAY": 8, "X=X-TAY": 9, "X=X+CTAY": 10}

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_C = ctypes.c_char


class OrcInfo(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int), ("relres", ctypes.c_double),
                ("relat", ctypes.c_int), ("converged", ctypes.c_int),
                ("zero_rhs", ctypes.c_int)]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_dot.restype = _D
        L.orc_dot.argtypes = [_I, _P, _P]
        L.orc_matvec.argtypes = [_I, _P, _P, _C, _P, _C, _P, _D, _P, _I, _I, _I, _I, _P, _P]
        L.orc_mt01aa.argtypes = [_P, _P, _D, _P, _I, _I]
        L.orc_mt02aa.argtypes = [_I, _P, _P, _D, _P, _P, _I, _P, _P, _P, _P, _I, _I, _I]
        L.orc_mt01pp.argtypes = [_P, _P, _D, _P, _P, _P, _I, _I]
        L.orc_mt02pp.argtypes = [_P, _P, _D, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P,
                                 _I, _I, _I, _I]
        L.orc_assemble_diag.argtypes = [_P, _P, _P, _I, _I, _I, _I]
        L.orc_kspg11.argtypes = [_P, _P, _P, _P, _P, _P, _P, _I, _I, _D]
        L.orc_mt03aa.argtypes = [_P, _P, _D, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I]
        L.orc_mt04aa.argtypes = [_P, _P, _D, _P, _P, _P, _P, _P, _P, _I, _I]
        L.orc_mt05aa.argtypes = [_P, _P, _D, _P, _P, _I, _I, _P, _P, _P, _I, _I]
        L.orc_om_dmd.argtypes = [_P, _P, _P, _C, _P, _P, _I, _I, _I, _I]
        L.orc_precd1.argtypes = [_P, _P, _P, _P, _C, _P, _P, _P, _I, _I, _I, _I, _I]
        L.orc_solve.restype = _I
        L.orc_solve.argtypes = [_I, _I, _I, _I, _D, _P, _P, _P, _P, _C, _P, _P,
                                _I, _I, _I, _I, ctypes.POINTER(OrcInfo)]
        L.orc_dirich.argtypes = [_P, _P, _C, _P, _C, _P, _P, _P, _I, _P, _I, _I, _I, _I]
        L.orc_lump.argtypes = [_P, _P, _C, _P, _C, _D, _P, _I, _I, _I, _I]
        _lib = L
    return _lib


class Matrix:
    """An EBE matrix as BIEF stores it: diagonal DA, off-diagonal XA
    (ncoef, nelem), TYPDIA, TYPEXT."""

    def __init__(self, da, xa, typdia="Q", typext="S"):
        self.da = _f64(da)
        self.xa = _f64(xa)
        self.typdia = typdia
        self.typext = typext

    def copy(self):
        return Matrix(self.da.copy(), self.xa.copy(), self.typdia, self.typext)


# --------------------------------------------------------------------------
# oracle (C restatement)
# --------------------------------------------------------------------------

def dot(x, y):
    x = _f64(x)
    y = _f64(y)
    return lib().orc_dot(x.size, _ptr(x), _ptr(y))


def matvec(op, A: Matrix, ikle, y, x=None, c=0.0):
    ikle = _i32(ikle)
    ndp, nelem = ikle.shape
    y = _f64(y)
    n = y.size
    x = np.zeros(n) if x is None else _f64(x).copy()
    w = np.empty(ndp * nelem)
    t = np.empty(n)
    lib().orc_matvec(OPS[op], _ptr(x), _ptr(A.da), A.typdia.encode(), _ptr(A.xa),
                     A.typext.encode(), _ptr(y), c, _ptr(ikle), n, nelem, nelem, ndp,
                     _ptr(w), _ptr(t))
    return x


def mt01aa(surfac, xmul=1.0):
    s = _f64(surfac)
    ne = s.size
    t = np.empty((3, ne))
    xm = np.empty((3, ne))
    lib().orc_mt01aa(_ptr(t), _ptr(xm), xmul, _ptr(s), ne, ne)
    return t, xm


def mt02aa(kind, xel, yel, surfac, ikle, u, v=None, xmul=1.0, mon=False, ldu=0):
    s = _f64(surfac)
    ne = s.size
    t = np.empty((3, ne))
    xm = np.empty((3, ne))
    u = _f64(u)
    v = _f64(v) if v is not None else u
    lib().orc_mt02aa(kind, _ptr(t), _ptr(xm), xmul, _ptr(u), _ptr(v), ldu,
                     _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(s), _ptr(_i32(ikle)), ne, ne,
                     1 if mon else 0)
    return t, xm


def kspg11(xel, yel, ikle, u, v, xmul=1.0):
    """KSUPG (ksupg.f -> kspg11.f): the SUPG element vectors (kx, ky)."""
    ik = _i32(ikle)
    ne = ik.shape[1]
    kx, ky = np.empty(ne), np.empty(ne)
    lib().orc_kspg11(_ptr(kx), _ptr(ky), _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(_f64(u)),
                     _ptr(_f64(v)), _ptr(ik), ne, ne, xmul)
    return kx, ky


def mt03aa(f, g, u, v, xel, yel, surfac, ikle, xmul=1.0):
    """MASUPG (mt03aa.f): t (3, ne), xm (6, ne) in the 'Q' order."""
    s = _f64(surfac)
    ne = s.size
    t, xm = np.empty((3, ne)), np.empty((6, ne))
    lib().orc_mt03aa(_ptr(t), _ptr(xm), xmul, _ptr(_f64(f)), _ptr(_f64(g)), _ptr(_f64(u)),
                     _ptr(_f64(v)), _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(s), _ptr(_i32(ikle)),
                     ne, ne)
    return t, xm


def mt04aa(u, v, xel, yel, surfac, ikle, xmul=1.0):
    """MAUGUG (mt04aa.f): t (3, ne), xm (3, ne) symmetric."""
    s = _f64(surfac)
    ne = s.size
    t, xm = np.empty((3, ne)), np.empty((3, ne))
    lib().orc_mt04aa(_ptr(t), _ptr(xm), xmul, _ptr(_f64(u)), _ptr(_f64(v)), _ptr(_f64(xel)),
                     _ptr(_f64(yel)), _ptr(s), _ptr(_i32(ikle)), ne, ne)
    return t, xm


def mt05aa(u, v, xel, yel, ikle, xmul=1.0, u_p0=False, upwind=False):
    """MATVGR (mt05aa.f): t (3, ne), xm (6, ne) in the 'Q' order."""
    ik = _i32(ikle)
    ne = ik.shape[1]
    t, xm = np.empty((3, ne)), np.empty((6, ne))
    lib().orc_mt05aa(_ptr(t), _ptr(xm), xmul, _ptr(_f64(u)), _ptr(_f64(v)), int(u_p0),
                     int(upwind), _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(ik), ne, ne)
    return t, xm


def mt01pp(surfac, z, ikle, xmul=1.0):
    s = _f64(surfac)
    ne = s.size
    t = np.empty((6, ne))
    xm = np.empty((15, ne))
    lib().orc_mt01pp(_ptr(t), _ptr(xm), xmul, _ptr(s), _ptr(_f64(z)), _ptr(_i32(ikle)), ne, ne)
    return t, xm


def mt02pp(f, g, h, xel, yel, z, surfac, ikle, npou0, xmul=1.0, h_zero=False, inchyd=False,
           mon=False):
    """Prism diffusion matrix (MT02PP); xel, yel (3, nelem) local coordinates
    of the base triangle; returns t (6, nelem), xm (15, nelem)."""
    s = _f64(surfac)
    ne = s.size
    t = np.empty((6, ne))
    xm = np.empty((15, ne))
    lib().orc_mt02pp(_ptr(t), _ptr(xm), xmul, _ptr(_f64(f)), _ptr(_f64(g)), _ptr(_f64(h)),
                     int(h_zero), int(npou0), _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(_f64(z)),
                     _ptr(s), _ptr(_i32(ikle)), ne, ne, int(inchyd), int(mon))
    return t, xm


def assemble_diag(t, ikle, npoin):
    ikle = _i32(ikle)
    ndp, ne = ikle.shape
    da = np.empty(npoin)
    lib().orc_assemble_diag(_ptr(da), _ptr(_f64(t)), _ptr(ikle), npoin, ne, ne, ndp)
    return da


def solve(A: Matrix, ikle, b, x0=None, slv=1, precon=2, krylov=5, nitmax=1000, eps=1e-8):
    """Oracle SOLVE. A and b are consumed as the reference does (copies are
    made here); returns x, info, and the preconditioned A, b."""
    ikle = _i32(ikle)
    ndp, ne = ikle.shape
    A2 = A.copy()
    b2 = _f64(b).copy()
    n = b2.size
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    td = ctypes.c_char(A2.typdia.encode())
    info = OrcInfo()
    rc = lib().orc_solve(slv, precon, krylov, nitmax, eps, _ptr(x), _ptr(A2.da),
                         ctypes.byref(td), _ptr(A2.xa), A2.typext.encode(), _ptr(b2),
                         _ptr(ikle), n, ne, ne, ndp, ctypes.byref(info))
    if rc != 0:
        raise ValueError(f"orc_solve returned {rc}")
    A2.typdia = td.value.decode()
    return x, info, A2, b2


OMP_SO = os.path.join(HERE, "liboracle_omp.so")
_omp = None


def omp_lib():
    global _omp
    if _omp is None:
        if not os.path.exists(OMP_SO):
            raise RuntimeError(f"{OMP_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(OMP_SO)
        L.orc_omp_plan_create.restype = _P
        L.orc_omp_plan_create.argtypes = [_P, _I, _I, _I]
        L.orc_omp_plan_free.argtypes = [_P]
        L.orc_omp_solve_cg.restype = _I
        L.orc_omp_solve_cg.argtypes = [_P, _P, _P, _C, _P, _C, _P, _P, _D, _I, _I, _I,
                                       ctypes.POINTER(OrcInfo)]
        _omp = L
    return _omp


class OmpPlan:
    """The multi-threaded SOLVE's node-to-element incidence lists (built once
    per mesh, as the reference's domain decomposition is)."""

    def __init__(self, ikle, npoin):
        self.ikle = _i32(ikle)
        self.ndp, self.nelem = self.ikle.shape
        self.npoin = int(npoin)
        self._h = omp_lib().orc_omp_plan_create(_ptr(self.ikle), self.npoin, self.nelem,
                                                self.ndp)
        if not self._h:
            raise ValueError("orc_omp_plan_create failed (node ids out of range?)")

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _omp is not None:
            _omp.orc_omp_plan_free(h)


def omp_solve(plan: OmpPlan, A: Matrix, b, x0=None, precon=2, nitmax=1000, eps=1e-8,
              nthreads=0):
    """SOLVE with SLV 1 (CG) and PRECON 1/2 on nthreads threads (0: OpenMP's
    default). Copies A and b as solve() does; returns x, info."""
    A2 = A.copy()
    b2 = _f64(b).copy()
    x = np.zeros(plan.npoin) if x0 is None else _f64(x0).copy()
    info = OrcInfo()
    rc = omp_lib().orc_omp_solve_cg(plan._h, _ptr(plan.ikle), _ptr(A2.da), A2.typdia.encode(),
                                    _ptr(A2.xa), A2.typext.encode(), _ptr(b2), _ptr(x), eps,
                                    nitmax, precon, int(nthreads), ctypes.byref(info))
    if rc != 0:
        raise ValueError(f"orc_omp_solve_cg returned {rc}")
    return x, info


def dirich(A: Matrix, ikle, sm, fbor, limdir, kdir=5, f=None):
    ikle = _i32(ikle)
    ndp, ne = ikle.shape
    A2 = A.copy()
    sm = _f64(sm).copy()
    n = sm.size
    f = np.zeros(n) if f is None else _f64(f).copy()
    lib().orc_dirich(_ptr(f), _ptr(A2.da), A2.typdia.encode(), _ptr(A2.xa),
                     A2.typext.encode(), _ptr(sm), _ptr(_f64(fbor)), _ptr(_i32(limdir)),
                     kdir, _ptr(ikle), n, ne, ne, ndp)
    return f, A2, sm


def lump(A: Matrix, ikle, xmul=1.0):
    ikle = _i32(ikle)
    ndp, ne = ikle.shape
    n = A.da.size
    d = np.empty(n)
    lib().orc_lump(_ptr(d), _ptr(A.da), A.typdia.encode(), _ptr(A.xa), A.typext.encode(),
                   xmul, _ptr(ikle), n, ne, ne, ndp)
    return d


# --------------------------------------------------------------------------
# the reference's own BIEF (flang build), where present
# --------------------------------------------------------------------------

_ref = None
_dropin = None
DROPIN_SO = os.path.join(HERE, "_ref", "libbief_dropin.so")


def have_ref():
    return os.path.exists(REF_SO)


def have_dropin():
    return os.path.exists(DROPIN_SO)


def dropin():
    """The reference's BIEF with SOLVE / MATVEC / MATRBL replaced by the
    Fortran host layer over libbiefhip (same driver entry points as ref())."""
    global _dropin
    if _dropin is None:
        if not have_dropin():
            raise RuntimeError(f"{DROPIN_SO} missing: run __graft_entry__.build()")
        _dropin = _bind_driver(ctypes.CDLL(DROPIN_SO))
    return _dropin


def _bind_driver(L):
    L.ref_mesh_init.argtypes = [_I, _I, _I, _I, _P]
    L.ref_matvec.argtypes = [_I, _C, _C, _P, _P, _P, _D, _P]
    L.ref_solve.argtypes = [_I, _I, _I, _I, _D, ctypes.POINTER(ctypes.c_char), _C,
                            _P, _P, _P, _P, ctypes.c_char_p, _I]
    L.ref_dirich.argtypes = [_C, _C, _P, _P, _P, _P, _P, _I, _P]
    L.ref_lump.argtypes = [_C, _C, _P, _P, _D, _P]
    L.ref_mt01aa.argtypes = [_I, _D, _P, _P, _P]
    L.ref_mt02aa.argtypes = [_I, _I, _I, _D, _P, _P, _P, _P, _P, _P, _I, _P, _P]
    L.ref_mt01pp.argtypes = [_I, _D, _P, _P, _P, _P, _P]
    L.ref_mt02pp.argtypes = [_I, _I, _I, _D, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _I, _I,
                             _P, _P]
    L.ref_mesh_segments.argtypes = [_I, _P]
    L.ref_kspg11.argtypes = [_I, _D, _P, _P, _P, _P, _P, _P, _P]
    L.ref_mt03aa.argtypes = [_I, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.ref_mt04aa.argtypes = [_I, _D, _P, _P, _P, _P, _P, _P, _P, _P]
    L.ref_mt05aa.argtypes = [_I, _D, _I, _I, _P, _P, _P, _P, _P, _P, _P]
    if hasattr(L, "ref_propag"):
        L.ref_propag.argtypes = [_I, _I] + [_P] * 10 + [_P, _I, _D, _D, _I, _D] + [_P] * 15 + \
            [ctypes.POINTER(_I)]
    return L


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError(f"{REF_SO} missing: run `make -C oracle ref`")
        _ref = _bind_driver(ctypes.CDLL(REF_SO))
    return _ref


class RefMesh:
    """Binds a mesh to the driver's BIEF_MESH (one at a time per library);
    loader = ref (the reference's BIEF) or dropin (its SOLVE/MATVEC
    replaced by the HIP host layer)."""
    _current = {}

    def __init__(self, ikle, npoin2, nelem2, nplan=1, loader=None, gloseg=None):
        """gloseg (2, nseg), 0-based: the matrix of the driver is in segment
        storage (MATRIX STORAGE = 3) over these segments of the triangles."""
        self.ikle = _i32(ikle)
        self.ndp = self.ikle.shape[0]
        self.args = (npoin2, nelem2, nplan)
        self.loader = loader or ref
        self.gloseg = None if gloseg is None else _i32(gloseg)

    @property
    def lib(self):
        return self.loader()

    def bind(self):
        key = self.loader.__name__
        if RefMesh._current.get(key) is not self:
            self.lib.ref_mesh_init(self.ndp, *self.args, _ptr(self.ikle))
            if self.gloseg is not None:
                self.lib.ref_mesh_segments(self.gloseg.shape[1], _ptr(self.gloseg))
            RefMesh._current[key] = self


def ref_matvec(rm: RefMesh, op, A: Matrix, y, x=None, c=0.0):
    rm.bind()
    y = _f64(y)
    x = np.zeros(y.size) if x is None else _f64(x).copy()
    rm.lib.ref_matvec(OPS[op], A.typdia.encode(), A.typext.encode(), _ptr(A.da), _ptr(A.xa),
                      _ptr(y), c, _ptr(x))
    return x


def _parse_solver_log(path):
    """Iteration count and precision from the reference's listing line
    (gracjg.F:426-435, cgstab.f, gmres.f formats)."""
    last = None
    with open(path) as fh:
        for line in fh:
            if "(BIEF)" in line:
                last = line.strip()
    if last is None:
        return None, None, last
    import re
    m = re.search(r":\s+(\d+)\s+ITERATIONS", last) or re.search(r"ITERATIONS:\s+(\d+)", last)
    niter = int(m.group(1)) if m else None
    m2 = re.search(r"PRECISION:\s*([-+0-9.Ee]+)", last)
    prec = float(m2.group(1)) if m2 else None
    return niter, prec, last


def ref_solve(rm: RefMesh, A: Matrix, b, x0=None, slv=1, precon=2, krylov=5, nitmax=1000,
              eps=1e-8):
    rm.bind()
    A2 = A.copy()
    b2 = _f64(b).copy()
    x = np.zeros(b2.size) if x0 is None else _f64(x0).copy()
    td = ctypes.create_string_buffer(A2.typdia.encode(), 1)
    fd, path = tempfile.mkstemp(suffix=".log")
    os.close(fd)
    try:
        p = path.encode()
        rm.lib.ref_solve(slv, precon, krylov, nitmax, eps, td, A2.typext.encode(),
                         _ptr(A2.da), _ptr(A2.xa), _ptr(b2), _ptr(x), p, len(p))
        niter, prec, line = _parse_solver_log(path)
    finally:
        os.unlink(path)
    A2.typdia = td.raw[:1].decode()
    return x, {"niter": niter, "relres": prec, "line": line}, A2, b2


def ref_dirich(rm: RefMesh, A: Matrix, sm, fbor, limdir, kdir=5, f=None):
    rm.bind()
    A2 = A.copy()
    sm = _f64(sm).copy()
    f = np.zeros(sm.size) if f is None else _f64(f).copy()
    rm.lib.ref_dirich(A2.typdia.encode(), A2.typext.encode(), _ptr(A2.da), _ptr(A2.xa), _ptr(sm),
                     _ptr(_f64(fbor)), _ptr(_i32(limdir)), kdir, _ptr(f))
    return f, A2, sm


def ref_lump(rm: RefMesh, A: Matrix, xmul=1.0):
    rm.bind()
    d = np.empty(A.da.size)
    rm.lib.ref_lump(A.typdia.encode(), A.typext.encode(), _ptr(A.da), _ptr(A.xa), xmul, _ptr(d))
    return d


def ref_mt01aa(surfac, xmul=1.0):
    s = _f64(surfac)
    ne = s.size
    t = np.empty((3, ne))
    xm = np.empty((3, ne))
    ref().ref_mt01aa(ne, xmul, _ptr(s), _ptr(t), _ptr(xm))
    return t, xm


def ref_mt02aa(kind, xel, yel, surfac, ikle, u, v=None, xmul=1.0, mon=False, npoin=None):
    s = _f64(surfac)
    ne = s.size
    u = _f64(u)
    v = _f64(v) if v is not None else u
    t = np.empty((3, ne))
    xm = np.empty((3, ne))
    npoin = npoin if npoin is not None else (u.size // (3 if kind == 4 else 1))
    ref().ref_mt02aa(kind, ne, npoin, xmul, _ptr(u), _ptr(v), _ptr(_f64(xel)), _ptr(_f64(yel)),
                     _ptr(s), _ptr(_i32(ikle)), 1 if mon else 0, _ptr(t), _ptr(xm))
    return t, xm


def ref_kspg11(xel, yel, ikle, u, v, xmul=1.0):
    ik = _i32(ikle)
    ne = ik.shape[1]
    kx, ky = np.empty(ne), np.empty(ne)
    ref().ref_kspg11(ne, xmul, _ptr(_f64(u)), _ptr(_f64(v)), _ptr(_f64(xel)), _ptr(_f64(yel)),
                     _ptr(ik), _ptr(kx), _ptr(ky))
    return kx, ky


def ref_mt03aa(f, g, u, v, xel, yel, surfac, ikle, xmul=1.0):
    s = _f64(surfac)
    ne = s.size
    t, xm = np.empty((3, ne)), np.empty((6, ne))
    ref().ref_mt03aa(ne, xmul, _ptr(_f64(f)), _ptr(_f64(g)), _ptr(_f64(u)), _ptr(_f64(v)),
                     _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(s), _ptr(_i32(ikle)), _ptr(t),
                     _ptr(xm))
    return t, xm


def ref_mt04aa(u, v, xel, yel, surfac, ikle, xmul=1.0):
    s = _f64(surfac)
    ne = s.size
    t, xm = np.empty((3, ne)), np.empty((3, ne))
    ref().ref_mt04aa(ne, xmul, _ptr(_f64(u)), _ptr(_f64(v)), _ptr(_f64(xel)), _ptr(_f64(yel)),
                     _ptr(s), _ptr(_i32(ikle)), _ptr(t), _ptr(xm))
    return t, xm


def ref_mt05aa(u, v, xel, yel, ikle, xmul=1.0, u_p0=False, upwind=False):
    ik = _i32(ikle)
    ne = ik.shape[1]
    t, xm = np.empty((3, ne)), np.empty((6, ne))
    ref().ref_mt05aa(ne, xmul, 10 if u_p0 else 11, int(upwind), _ptr(_f64(u)), _ptr(_f64(v)),
                     _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(ik), _ptr(t), _ptr(xm))
    return t, xm


def ref_mt01pp(surfac, z, ikle, xmul=1.0):
    s = _f64(surfac)
    ne = s.size
    t = np.empty((6, ne))
    xm = np.empty((30, ne))
    ref().ref_mt01pp(ne, xmul, _ptr(_f64(z)), _ptr(s), _ptr(_i32(ikle)), _ptr(t), _ptr(xm))
    return t, xm[:15]


def ref_mt02pp(f, g, h, xel, yel, z, surfac, ikle, nplan, xmul=1.0, h_zero=False,
               p0_vertical=False, inchyd=False, mon=False):
    s = _f64(surfac)
    ne = s.size
    f = _f64(f)
    t = np.empty((6, ne))
    xm = np.empty((15, ne))
    ref().ref_mt02pp(ne, f.size, nplan, xmul, _ptr(f), _ptr(_f64(g)), _ptr(_f64(h)), int(h_zero),
                     int(p0_vertical), _ptr(_f64(xel)), _ptr(_f64(yel)), _ptr(_f64(z)), _ptr(s),
                     _ptr(_i32(ikle)), int(inchyd), int(mon), _ptr(t), _ptr(xm))
    return t, xm


# ---------------------------------------------------------------------------
# CVDFTR's SUPG system (cvdftr.F:398-560), assembled from element matrices
# ---------------------------------------------------------------------------
def _om_xm(xa):
    """OM 'M=X(M)': the non-symmetric form of a symmetric triangle matrix."""
    return np.concatenate([xa, xa]) if xa.shape[0] == 3 else xa.copy()


def supg_system(mesh, u, v, nu, dt, fn, tetat=1.0, reference=False):
    """The tracer system of CVDFTR with classical SUPG (OPTADV = 1), an
    implicit P1 isotropic diffusion nu (MATDIF, TETATD = 1) and a zero
    source, built from the oracle's element routines (or, with
    reference=True, from the reference's own through the driver):

      TE = KSUPG(1, U)                                    (cvdftr.F:406)
      AM2 = MATVGR(1, U) + MASUPG(1, TE, U)               (:398-409)
      AM1 = X(MATMAS(1/dt)) + T(MATVGR(1/dt, TE as P0))   (:448, :513-523)
      SM = AM1 FN + (TETAT - 1) AM2 FN                    (:541-546)
      AM1 = AM1 + TETAT AM2 + MATDIF(nu)                  (:550, :694-733)

    Returns (Matrix 'Q', sm). The OM operations are element-wise and
    restated here with numpy; the diagonal assembly and the products are
    the oracle's (ASSVEC order)."""
    xe, ye = mesh.xel_yel()
    s = mesh.surfac()
    ik = mesh.ikle
    n = mesh.npoin
    P = globals()
    k = (lambda name: P["ref_" + name]) if reference else (lambda name: P[name])
    te1, te2 = k("kspg11")(xe, ye, ik, u, v, 1.0)
    t5, x5 = k("mt05aa")(u, v, xe, ye, ik, 1.0)
    t3, x3 = k("mt03aa")(te1, te2, u, v, xe, ye, s, ik, 1.0)
    am2_d = assemble_diag(t5, ik, n) + assemble_diag(t3, ik, n)
    am2_x = x5 + x3
    am2 = Matrix(am2_d, am2_x, "Q", "Q")
    t2 = matvec("X=AY", am2, ik, fn)
    t1, x1 = k("mt01aa")(s, 1.0 / dt)
    am1_d = assemble_diag(t1, ik, n)
    am1_x = _om_xm(x1)
    tp, xp = k("mt05aa")(te1, te2, xe, ye, ik, 1.0 / dt, u_p0=True)
    am1_d = am1_d + assemble_diag(tp, ik, n)
    am1_x = am1_x + np.concatenate([xp[3:], xp[:3]])
    am1 = Matrix(am1_d, am1_x, "Q", "Q")
    sm = matvec("X=AY", am1, ik, fn)
    sm = sm + (tetat - 1.0) * t2
    am1_d = am1_d + tetat * am2_d
    am1_x = am1_x + tetat * am2_x
    td, xd = k("mt02aa")(1, xe, ye, s, ik, nu, None, 1.0)
    am1_d = am1_d + assemble_diag(td, ik, n)
    am1_x = am1_x + _om_xm(xd)
    return Matrix(am1_d, am1_x, "Q", "Q"), sm


# --------------------------------------------------------------------------
# telemac2d PROPAG: one wave-equation step's operands
# --------------------------------------------------------------------------

def _assemble3(w, ikle, npoin, mask=None):
    """ASSVEC of three element columns (assvec.f: column k, elements in
    order), MASKEL multiplying each term when given (assve1.f:66-84)."""
    x = np.zeros(npoin)
    for k in range(3):
        wk = w[k] if mask is None else w[k] * mask
        np.add.at(x, ikle[k], wk)
    return x


def _vc13aa(xel, yel, ikle, f, xmul):
    """GRADF element terms (vc13aa.f, P1 F): (x, y) components, the same
    value for the three nodes."""
    f1, f2, f3 = f[ikle[0]], f[ikle[1]], f[ikle[2]]
    xs6 = xmul / 6.0
    wx = (yel[1] * (f1 - f3) + yel[2] * (f2 - f1)) * xs6
    wy = (xel[1] * (f3 - f1) + xel[2] * (f1 - f2)) * xs6
    return np.stack([wx, wx, wx]), np.stack([wy, wy, wy])


def _slop10(xel, yel, ikle, z):
    """Bottom-slope area factor per element (slop10.f)."""
    x2, x3, y2, y3 = xel[1], xel[2], yel[1], yel[2]
    z2 = z[ikle[1]] - z[ikle[0]]
    z3 = z[ikle[2]] - z[ikle[0]]
    a = (x2 * y3 - x3 * y2) ** 2
    b = (y2 * z3 - z2 * y3) ** 2
    c = (x3 * z2 - z3 * x2) ** 2
    return np.sqrt((a + b + c) / a)


def _mt08aa(xel, yel, ikle, f, xmul, icoord):
    """MATFGR (mt08aa.f): A(i,j) = -XMUL int PSI2(j) F dPSI1(i)/dx (or /dy),
    as (diagonal (3, ne), off-diagonal in the 'Q' order 12,13,23,21,31,32)."""
    s24 = xmul / 24.0
    f1, f2, f3 = f[ikle[0]] * s24, f[ikle[1]] * s24, f[ikle[2]] * s24
    f123 = f1 + f2 + f3
    if icoord == 1:
        y2, y3 = yel[1], yel[2]
        a12 = (y3 - y2) * (f123 + f2)
        a13 = (y3 - y2) * (f123 + f3)
        a23 = y3 * (-f123 - f3)
        a21 = y3 * (-f123 - f1)
        a31 = y2 * (f123 + f1)
        a32 = y2 * (f123 + f2)
    else:
        x2, x3 = xel[1], xel[2]
        a12 = (x2 - x3) * (f123 + f2)
        a13 = (x2 - x3) * (f123 + f3)
        a23 = x3 * (f123 + f3)
        a21 = x3 * (f123 + f1)
        a31 = x2 * (-f123 - f1)
        a32 = x2 * (-f123 - f2)
    t = np.stack([-a21 - a31, -a12 - a32, -a13 - a23])
    return t, np.stack([a12, a13, a23, a21, a31, a32])


def propag_step(mesh, hn, un, vn, zf, chestr, hbor, limdir, kdir=2, dt=1.0, grav=9.81,
                nitmax=1000, eps=1e-12):
    """One wave-equation step of telemac2d PROPAG (TREATMENT OF THE LINEAR
    SYSTEM = 2) for the options the driver's ref_propag runs (implicitation
    1, no advection or diffusion of velocity, no sources, no tidal flats, one
    sub-iteration, Strickler friction with HFROT 1, the depth hbor prescribed
    where limdir == kdir, walls elsewhere), in PROPAG's sequence of
    operations (propag.F:664-716, 1003-1012, 1150-1170, 1369-1465, 1634-1710):

      VOLU2D = MASBAS; HC = max(HN, 1e-4) (friction_unif.f:124-127);
      CF = 2 GRAV / K^2 / HC^(1/3) (friction_calc.f:133-134);
      FU = -0.5 CF |U| / max(HN, 1e-9), |U| raised below 3 cm (fricti.f:123-135);
      T = FU * MASBAS(-1, slope mask) (slop10.f); AM2D = VOLU2D/dt + T;
      CV2 = GRADF(-GRAV, HN + ZF) + (VOLU2D/dt) UN; AM2D^-1 (guard 1e-6);
      DM1 = -GRAV VOLU2D AM2D^-1; HPROP = max(HN, 0);
      AM1 = MATMAS(1/dt) + MATDIFUV(-1, DM1, HPROP);
      T2 = CV2 AM2D^-1; CV1 = -MATFGR_X(HPROP) T2 - MATFGR_Y(HPROP) T3;
      DIRICH(DH, AM1, CV1, HBOR - HN); SOLVE (CG, PRECON 2);
      U = (CV2 + GRADF(-GRAV, DH)) AM2D^-1.

    The element routines are the oracle's (MT01AA, MT02AA, the EBE product)
    or restated here (VC00AA, VC13AA, SLOP10, MT08AA); sums follow ASSVEC's
    order. Returns a dict of node vectors (and AM1's da, xa (3, ne))."""
    xe, ye = mesh.xel_yel()
    s = mesh.surfac()
    ik = mesh.ikle
    n = mesh.npoin
    hn, un, vn, zf, chestr = (_f64(a) for a in (hn, un, vn, zf, chestr))
    third = np.stack([s / 3.0] * 3)
    volu2d = _assemble3(third, ik, n)
    hc = np.maximum(hn, 1e-4)
    cf = 2.0 * grav / chestr ** 2 / hc ** (1.0 / 3.0)
    unorm = np.sqrt(un ** 2 + vn ** 2)
    h = np.maximum(hn, 1e-9)
    ho = 3e-2
    shallow = h < ho
    unorm = np.where(shallow, np.maximum(unorm, np.sqrt(9.81 * np.maximum(ho - h, 0.0) * h / ho)),
                     unorm)
    fu = -0.5 * cf * unorm / h
    te3 = _slop10(xe, ye, ik, zf)
    tm = _assemble3(-third, ik, n, mask=te3) * fu
    am3d = volu2d * (1.0 / dt)
    am2d = am3d + tm
    t8 = hn + zf
    wx, wy = _vc13aa(xe, ye, ik, t8, -grav)
    cv2 = _assemble3(wx, ik, n) + am3d * un
    cv3 = _assemble3(wy, ik, n) + am3d * vn
    inv = np.where(np.abs(am2d) > 1e-6, 1.0 / np.where(am2d == 0.0, 1.0, am2d), 0.0)
    dm1 = -grav * volu2d * inv
    hprop = np.maximum(hn, 0.0)
    t1, x1 = mt01aa(s, 1.0 / dt)
    t2e, x2e = mt02aa(2, xe, ye, s, ik, dm1, hprop, -1.0)
    da = assemble_diag(t1, ik, n) + assemble_diag(t2e, ik, n)
    xa = x1 + x2e
    t2 = cv2 * inv
    t3 = cv3 * inv
    cv1 = np.zeros(n)
    for icoord, tt in ((1, t2), (2, t3)):
        tb, xb = _mt08aa(xe, ye, ik, hprop, 1.0, icoord)
        B = Matrix(assemble_diag(tb, ik, n), xb, "Q", "Q")
        cv1 = matvec("X=X-AY", B, ik, tt, x=cv1)
    fbor = np.where(_i32(limdir) == kdir, _f64(hbor) - hn, 0.0)
    dh0, A1, cv1 = dirich(Matrix(da, xa, "Q", "S"), ik, cv1, fbor, limdir, kdir=kdir,
                          f=np.zeros(n))
    dh, info, _, _ = solve(A1, ik, cv1, dh0, slv=1, precon=2, eps=eps, nitmax=nitmax)
    gx, gy = _vc13aa(xe, ye, ik, dh, -grav)
    u = (cv2 + _assemble3(gx, ik, n)) * inv
    v = (cv3 + _assemble3(gy, ik, n)) * inv
    return {"volu2d": volu2d, "cf": cf, "fu": fu, "am2d": inv, "dm1": dm1, "hprop": hprop,
            "t2": t2, "t3": t3, "da": A1.da, "xa": A1.xa, "cv1": cv1, "dh0": dh0, "dh": dh,
            "u": u, "v": v, "niter": info.niter}


def ref_propag(mesh, hn, un, vn, zf, chestr, hbor, limdir, kdir=2, dt=1.0, grav=9.81,
               nitmax=1000, eps=1e-12):
    """The same step computed by the reference's own routines in the
    driver (ref_driver.f90: ref_propag). Same dict (niter not reported)."""
    rm = RefMesh(mesh.ikle, mesh.npoin, mesh.nelem)
    rm.bind()
    xe, ye = mesh.xel_yel()
    s = mesh.surfac()
    n, ne = mesh.npoin, mesh.nelem
    out = {k: np.zeros(n) for k in ("volu2d", "cf", "fu", "am2d", "dm1", "hprop", "t2", "t3",
                                    "da", "cv1", "dh0", "dh", "u", "v")}
    xa = np.zeros((3, ne))
    niter = ctypes.c_int(0)
    ik = _i32(mesh.ikle)
    args = [_f64(a) for a in (hn, un, vn, zf, chestr, hbor)]
    rm.lib.ref_propag(ne, n, _ptr(ik), _ptr(_f64(xe)), _ptr(_f64(ye)), _ptr(_f64(s)),
                      *[_ptr(a) for a in args], _ptr(_i32(limdir)), kdir, dt, grav, nitmax, eps,
                      *[_ptr(out[k]) for k in ("volu2d", "cf", "fu", "am2d", "dm1", "hprop",
                                               "t2", "t3", "da")],
                      _ptr(xa), _ptr(out["cv1"]), _ptr(out["dh0"]), _ptr(out["dh"]),
                      _ptr(out["u"]), _ptr(out["v"]), ctypes.byref(niter))
    out["xa"] = xa
    return out
