"""Python host layer over libbiefhip (the C ABI in include/bief_hip.h).

It mirrors BIEF's interface for the linear-system path:

  * ``Context(mesh)``                 -- what SOLVE/MATVEC read from BIEF_MESH
  * ``Context.set_matrix(...)``       -- an EBE BIEF_OBJ matrix (D, X, TYPDIA, TYPEXT)
  * ``Context.matvec(op, x, y, c)``   -- MATVEC(OP, X, A, Y, C, MESH)
                                         (sources/utils/bief/matvec.f:98-103)
  * ``Context.solve(x, b, cfg)``      -- SOLVE(X, A, B, TB, CFG, INFOGR, MESH, AUX)
                                         (sources/utils/bief/solve.f:2-5)
  * ``Context.dirich``, ``Context.lump`` -- DIRICH (dirich.f), LUMP (lump.f)

Vectors are either host numpy arrays (synchronous copies, as BIEF's
routines are) or torch CUDA tensors (device resident, work queued on the
context's stream). Errors the reference would PLANTE on raise ``BiefError``;
non-convergence is reported in the returned info, as the reference only
prints it (gracjg.F:411-418).

The library must be built (``__graft_entry__.build()``); there is no CPU
fallback: a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbiefhip.so")

OPS = {"X=AY": 0, "X=CAY": 1, "X=-AY": 2, "X=X+AY": 3, "X=X-AY": 4, "X=X+CAY": 5,
       "X=TAY": 6, "X=-TAY": 7, "X=X+TAY": 8, "X=X-TAY": 9, "X=X+CTAY": 10}

STATUS = {0: "ok", 1: "maxiter", 2: "bad argument", 3: "zero rhs", 4: "device error"}

PREC_FP64 = 0
PREC_MIXED = 1


class BiefError(RuntimeError):
    pass


class SolverCfg(ctypes.Structure):
    """SLVCFG (sources/utils/bief/bief_def.f:506-532) + precision mode."""
    _fields_ = [("slv", ctypes.c_int), ("precon", ctypes.c_int), ("krylov", ctypes.c_int),
                ("nitmax", ctypes.c_int), ("eps", ctypes.c_double), ("precision", ctypes.c_int),
                ("inner_eps", ctypes.c_double), ("inner_nitmax", ctypes.c_int)]


class WaveParams(ctypes.Structure):
    """bhip_wave_params: time step, gravity, speed floor, tide on sea nodes."""
    _fields_ = [("dt", ctypes.c_double), ("grav", ctypes.c_double), ("eps_u", ctypes.c_double),
                ("tide", ctypes.c_double)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int), ("inner_iters", ctypes.c_int),
                ("outer_iters", ctypes.c_int), ("relat", ctypes.c_int),
                ("relres", ctypes.c_double), ("converged", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_C = ctypes.c_char

_lib = None


def lib():
    """Loads libbiefhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BiefError(f"{LIB_PATH} not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.bhip_ctx_create.argtypes = [_I, _I, _I, _P, _P, _P, _I, _I, ctypes.POINTER(_P)]
        L.bhip_ctx_destroy.argtypes = [_P]
        L.bhip_ctx_stream.argtypes = [_P]
        L.bhip_ctx_stream.restype = _P
        L.bhip_ctx_stats.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                     ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64)]
        L.bhip_matrix_set.argtypes = [_P, _C, _C, _P, _P]
        L.bhip_matrix_set_dev.argtypes = [_P, _C, _C, _P, _P]
        L.bhip_matrix_get.argtypes = [_P, ctypes.c_char_p, _P, _P]
        L.bhip_matrix_info.argtypes = [_P, ctypes.c_char_p, ctypes.c_char_p,
                                       ctypes.POINTER(ctypes.c_int)]
        L.bhip_matvec.argtypes = [_P, _I, _D, _P, _P]
        L.bhip_matvec_dev.argtypes = [_P, _I, _D, _P, _P]
        L.bhip_solve.argtypes = [_P, ctypes.POINTER(SolverCfg), _P, _P, ctypes.POINTER(SolveInfo)]
        L.bhip_solve_dev.argtypes = [_P, ctypes.POINTER(SolverCfg), _P, _P,
                                     ctypes.POINTER(SolveInfo)]
        L.bhip_matrix_save.argtypes = [_P]
        L.bhip_matrix_restore.argtypes = [_P]
        L.bhip_time_cg_kernels.argtypes = [_P, _I, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)]
        L.bhip_time_matvec.argtypes = [_P, _I, _I, ctypes.POINTER(_D)]
        L.bhip_time_cg_dir_cold.argtypes = [_P, _I, _I, _P, ctypes.c_long, ctypes.POINTER(_D)]
        L.bhip_layout_check.argtypes = [_I, _I, _I, _P, _P, _P, _I, _I, _P]
        L.bhip_dirich.argtypes = [_P, _P, _P, _P, _P, _I]
        L.bhip_lump.argtypes = [_P, _D, _P]
        L.bhip_matrix_mt01aa.argtypes = [_P, _D, _P, _I]
        L.bhip_matrix_mt02aa.argtypes = [_P, _I, _D, _P, _P, _P, _P, _P, _I]
        L.bhip_batch_create.argtypes = [_I, _P, _P, _I, _P, _P, _P, _I, ctypes.POINTER(_P)]
        L.bhip_batch_info.argtypes = [_P, _I, ctypes.POINTER(SolveInfo)]
        L.bhip_batch_latency.argtypes = [_P, _I, _P]
        L.bhip_matrix_mt01pp.argtypes = [_P, _D, _P, _P, _I]
        L.bhip_matrix_mt02pp.argtypes = [_P, _D, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I]
        L.bhip_solve_dev_internal.argtypes = [_P, ctypes.POINTER(SolverCfg), _P, _P,
                                              ctypes.POINTER(SolveInfo)]
        L.bhip_solve_dev_restored.argtypes = [_P, ctypes.POINTER(SolverCfg), _P, _P,
                                              ctypes.POINTER(SolveInfo)]
        L.bhip_solve_dev_noalloc.argtypes = [_P, ctypes.POINTER(SolverCfg), _P, _P,
                                             ctypes.POINTER(SolveInfo)]
        L.bhip_ctx_perm.argtypes = [_P, _P]
        L.bhip_matrix_supg.argtypes = [_P, _I, _D, _P, _P, _P, _P, _P, _P, _P, _I, _I]
        L.bhip_matrix_om.argtypes = [_P, ctypes.c_char_p, _D]
        L.bhip_matrix_supg_dev.argtypes = [_P, _I, _D, _P, _P, _P, _P, _I, _I]
        L.bhip_matrix_mt01aa_dev.argtypes = [_P, _D, _I]
        L.bhip_matrix_mt02aa_dev.argtypes = [_P, _I, _D, _P, _P, _I]
        L.bhip_matrix_mt01pp_dev.argtypes = [_P, _D, _P, _I]
        L.bhip_matrix_mt02pp_dev.argtypes = [_P, _D, _P, _P, _P, _P, _I, _I, _I, _I, _I]
        L.bhip_dirich_dev.argtypes = [_P, _P, _P, _P, _P, _I]
        L.bhip_lump_dev.argtypes = [_P, _D, _P]
        L.bhip_ctx_set_geometry.argtypes = [_P, _P, _P, _P]
        _W = ctypes.POINTER(WaveParams)
        L.bhip_wave_setup.argtypes = [_P, _P, _P, _P, _P]
        L.bhip_wave_assemble.argtypes = [_P, _W] + [_P] * 6
        L.bhip_wave_momentum.argtypes = [_P, _W] + [_P] * 7
        L.bhip_wave_momentum_adj.argtypes = [_P, _W] + [_P] * 4
        L.bhip_wave_assemble_adj.argtypes = [_P, _W] + [_P] * 16
        _lib = L
    return _lib


def _check(rc, what):
    if rc in (0, 1, 3):
        return rc
    raise BiefError(f"{what}: {STATUS.get(rc, rc)} (status {rc})")


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _host_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _host_or_none(a):
    return None if a is None else _host_f64(a)


def _host_n(a, n, what):
    """float64 contiguous host array of exactly n values: the C entry points
    read n values from the pointer, so a shorter array is refused here."""
    h = _host_f64(a)
    if h.size != n:
        raise BiefError(f"{what}: {h.size} values, expected {n}")
    return h


def _host_n_or_none(a, n, what):
    return None if a is None else _host_n(a, n, what)


def _dev_n(t, n, what):
    """A torch tensor handed to a device entry point: contiguous float64 on
    a GPU with exactly n values (the kernels read n values from its
    address; a host tensor's address is no device address)."""
    import torch
    if not (_is_torch(t) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise BiefError(f"{what}: expected a contiguous float64 CUDA tensor")
    if t.numel() != n:
        raise BiefError(f"{what}: {t.numel()} values, expected {n}")
    return t


def _ptr(a):
    return None if a is None else a.ctypes.data


def layout_check(mesh, block_nodes=0):
    """Host-only build + validation of the block layout (no GPU needed).
    Returns the statistics dict; raises BiefError if the layout is invalid."""
    st = np.zeros(7, dtype=np.int64)
    ik = np.ascontiguousarray(mesh.ikle, dtype=np.int32)
    x = np.ascontiguousarray(mesh.x, dtype=np.float64)
    y = np.ascontiguousarray(mesh.y, dtype=np.float64)
    _check(lib().bhip_layout_check(mesh.npoin, mesh.nelem, mesh.ndp, ik.ctypes.data,
                                   x.ctypes.data, y.ctypes.data, int(getattr(mesh, "nplan", 1)),
                                   int(block_nodes), st.ctypes.data), "bhip_layout_check")
    keys = ["nblocks", "max_colors", "elem_instances", "halo_nodes", "max_local", "max_owned",
            "lds_bytes_f64"]
    return dict(zip(keys, (int(v) for v in st)))


@dataclass
class MatrixInfo:
    typdia: str
    typext: str


class Context:
    """Per-mesh device context (libbiefhip bhip_ctx)."""

    def __init__(self, mesh, block_nodes: int = 0, use_coords: bool = True):
        self.mesh = mesh
        self.npoin = mesh.npoin
        self.nelem = mesh.nelem
        self.ndp = mesh.ndp
        self._ikle = np.ascontiguousarray(mesh.ikle, dtype=np.int32)
        x = np.ascontiguousarray(mesh.x, dtype=np.float64) if use_coords else None
        y = np.ascontiguousarray(mesh.y, dtype=np.float64) if use_coords else None
        h = _P()
        rc = lib().bhip_ctx_create(self.npoin, self.nelem, self.ndp, self._ikle.ctypes.data,
                                   x.ctypes.data if x is not None else None,
                                   y.ctypes.data if y is not None else None,
                                   int(getattr(mesh, "nplan", 1)), int(block_nodes),
                                   ctypes.byref(h))
        _check(rc, "bhip_ctx_create")
        self._h = h
        self.typdia = "0"
        self.typext = "0"

    def close(self):
        if getattr(self, "_h", None):
            lib().bhip_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def stream_ptr(self) -> int:
        return lib().bhip_ctx_stream(self._h)

    def torch_stream(self):
        import torch
        return torch.cuda.ExternalStream(self.stream_ptr())

    def stats(self):
        nb, mc = _I(), _I()
        ei, hn = ctypes.c_int64(), ctypes.c_int64()
        lib().bhip_ctx_stats(self._h, ctypes.byref(nb), ctypes.byref(mc), ctypes.byref(ei),
                             ctypes.byref(hn))
        return {"nblocks": nb.value, "max_colors": mc.value, "elem_instances": ei.value,
                "halo_nodes": hn.value}

    def synchronize(self):
        import torch
        self.torch_stream().synchronize()

    # ---- matrix ---------------------------------------------------------
    def set_matrix(self, typdia: str, typext: str, da, xa):
        """A%D, A%X (ncoef x nelem), A%TYPDIA, A%TYPEXT."""
        if _is_torch(da):
            npair = self.ndp * (self.ndp - 1) // 2
            ncoef = {"Q": 2 * npair, "S": npair}.get(typext, 0)
            _dev_n(da, self.npoin, "set_matrix: da")
            if ncoef:
                _dev_n(xa, ncoef * self.nelem, "set_matrix: xa")
            self._sync_torch()
            rc = lib().bhip_matrix_set_dev(self._h, typdia.encode(), typext.encode(),
                                           da.data_ptr(), xa.data_ptr() if xa is not None else None)
            # the copy runs later on the context stream: torch's stream waits
            # for it, so memory torch hands out again is only written after
            # the copy has read it. (Not record_stream: the context's stream
            # may be destroyed before the tensor is freed, and torch would
            # then record an event on a stream that no longer exists.)
            self._join_torch()
        else:
            da = _host_n_or_none(da, self.npoin, "set_matrix: da")
            npair = self.ndp * (self.ndp - 1) // 2
            ncoef = {"Q": 2 * npair, "S": npair}.get(typext, 0)
            # (TYPEXT '0' reads no off-diagonal terms)
            xa = _host_n(xa, ncoef * self.nelem, "set_matrix: xa") \
                if xa is not None and ncoef else None
            rc = lib().bhip_matrix_set(self._h, typdia.encode(), typext.encode(), _ptr(da),
                                       _ptr(xa))
        _check(rc, "bhip_matrix_set")
        self.typdia, self.typext = typdia, typext

    def save_matrix(self):
        """OM 'M=N' into the context's saved copy (device to device)."""
        _check(lib().bhip_matrix_save(self._h), "bhip_matrix_save")
        self._saved_typext = self.typext

    def restore_matrix(self):
        """OM 'M=N' back from the saved copy (queued on the context stream)."""
        _check(lib().bhip_matrix_restore(self._h), "bhip_matrix_restore")
        self.typext = getattr(self, "_saved_typext", self.typext)

    def matrix_info(self):
        """(TYPDIA, TYPEXT, number of off-diagonal columns) of the context's
        matrix, as the library holds them."""
        td, te, nc = ctypes.create_string_buffer(1), ctypes.create_string_buffer(1), ctypes.c_int()
        _check(lib().bhip_matrix_info(self._h, td, te, ctypes.byref(nc)), "bhip_matrix_info")
        return td.raw[:1].decode(), te.raw[:1].decode(), nc.value

    def get_matrix(self):
        # sized by what the library holds
        _, self.typext, ncoef = self.matrix_info()
        da = np.empty(self.npoin)
        xa = np.empty((max(ncoef, 1), self.nelem))
        td = ctypes.create_string_buffer(1)
        _check(lib().bhip_matrix_get(self._h, td, da.ctypes.data, xa.ctypes.data if ncoef else None),
               "bhip_matrix_get")
        self.typdia = td.raw[:1].decode()
        return self.typdia, da, xa[:ncoef]

    # ---- operand preparation --------------------------------------------
    def dirich(self, f, sm, fbor, limdir, kdir=5):
        """DIRICH with domain-sized FBOR/LIMDIR (dirich.f -> diri01.f ->
        diraux.f). Returns (f, sm); the context's matrix gets the rows and
        columns of prescribed nodes zeroed. Torch CUDA tensors (limdir
        int32): f and sm are updated in place on the context stream."""
        n = self.npoin
        if _is_torch(f):
            import torch
            for t, w in ((f, "f"), (sm, "sm"), (fbor, "fbor")):
                _dev_n(t, n, "dirich: " + w)
            if not (_is_torch(limdir) and limdir.is_cuda and limdir.dtype == torch.int32 and
                    limdir.is_contiguous() and limdir.numel() == n):
                raise BiefError("dirich: limdir: expected a contiguous int32 CUDA tensor of "
                                f"{n} values")
            self._sync_torch()
            _check(lib().bhip_dirich_dev(self._h, f.data_ptr(), sm.data_ptr(), fbor.data_ptr(),
                                         limdir.data_ptr(), int(kdir)), "bhip_dirich_dev")
            self._join_torch()
            return f, sm
        f = _host_n(f, n, "dirich: f").copy()
        sm = _host_n(sm, n, "dirich: sm").copy()
        fb = _host_n(fbor, n, "dirich: fbor")
        ld = np.ascontiguousarray(limdir, dtype=np.int32)
        if ld.size != n:
            raise BiefError(f"dirich: limdir: {ld.size} values, expected {n}")
        _check(lib().bhip_dirich(self._h, f.ctypes.data, sm.ctypes.data, fb.ctypes.data,
                                 ld.ctypes.data, int(kdir)), "bhip_dirich")
        return f, sm

    def lump(self, xmul=1.0, out=None):
        """LUMP: A * (XMUL everywhere) (lump.f). out: a torch CUDA tensor
        receives it on the context stream."""
        if out is not None:
            _dev_n(out, self.npoin, "lump: out")
            self._sync_torch()
            _check(lib().bhip_lump_dev(self._h, float(xmul), out.data_ptr()), "bhip_lump_dev")
            self._join_torch()
            return out
        d = np.empty(self.npoin)
        _check(lib().bhip_lump(self._h, float(xmul), d.ctypes.data), "bhip_lump")
        return d

    def set_geometry(self, xel, yel, surfac):
        """Keeps the mesh's XEL, YEL, SURFAC on the device (what INBIEF
        computes once, inbief.f:384-397); the matrix_* methods then take
        xel = yel = surfac = None."""
        ne = self.nelem
        arrs = [_host_n(xel, 3 * ne, "xel"), _host_n(yel, 3 * ne, "yel"),
                _host_n(surfac, ne, "surfac")]
        _check(lib().bhip_ctx_set_geometry(self._h, *(a.ctypes.data for a in arrs)),
               "bhip_ctx_set_geometry")
        self.has_geometry = True

    def matrix_mt01aa(self, xmul, surfac, accumulate=False):
        """MATRIX(... 'MATMAS' ...) with OM 'M=N' / 'M=M+N' (mt01aa.f).
        surfac None: the geometry of set_geometry."""
        if surfac is None and getattr(self, "has_geometry", False):
            # no host array at all: queued on the context stream
            self._sync_torch()
            _check(lib().bhip_matrix_mt01aa_dev(self._h, float(xmul), int(accumulate)),
                   "bhip_matrix_mt01aa_dev")
            self._join_torch()
        else:
            s = _host_n_or_none(surfac, self.nelem, "surfac")
            _check(lib().bhip_matrix_mt01aa(self._h, float(xmul), _ptr(s), int(accumulate)),
                   "bhip_matrix_mt01aa")
        self.typdia, self.typext = "Q", ("Q" if accumulate and self.typext == "Q" else "S")

    def matrix_mt02aa(self, kind, xmul, xel, yel, surfac, u, v=None, accumulate=False):
        """MATRIX(... 'MATDIF' / 'MATDIFUV' ...): kind 1 P1 isotropic
        viscosity U, kind 2 coefficient U*V (mt02aa.f)."""
        ne, n = self.nelem, self.npoin
        if _is_torch(u):
            # device node fields, the context's geometry
            if xel is not None or yel is not None or surfac is not None:
                raise BiefError("matrix_mt02aa: device fields take the context's geometry")
            _dev_n(u, n, "u")
            if v is not None:
                _dev_n(v, n, "v")
            self._sync_torch()
            _check(lib().bhip_matrix_mt02aa_dev(self._h, int(kind), float(xmul), u.data_ptr(),
                                                None if v is None else v.data_ptr(),
                                                int(accumulate)), "bhip_matrix_mt02aa_dev")
            self._join_torch()
        else:
            xel, yel = _host_n_or_none(xel, 3 * ne, "xel"), _host_n_or_none(yel, 3 * ne, "yel")
            s = _host_n_or_none(surfac, ne, "surfac")
            u, v = _host_n(u, n, "u"), _host_n_or_none(v, n, "v")
            _check(lib().bhip_matrix_mt02aa(self._h, int(kind), float(xmul), _ptr(xel),
                                            _ptr(yel), _ptr(s), u.ctypes.data, _ptr(v),
                                            int(accumulate)), "bhip_matrix_mt02aa")
        self.typdia, self.typext = "Q", ("Q" if accumulate and self.typext == "Q" else "S")

    SUPG_U_P0, SUPG_UPWIND, SUPG_KSUPG = 1, 2, 4
    _OM_OPS = {"M=N": 0, "M=M+N": 1, "M=M+TN": 2}

    def matrix_supg(self, kind, xmul, xel, yel, surfac, u, v, f=None, g=None, flags=0, op="M=N"):
        """MATRIX(... 'MASUPG' (kind 3, MT03AA) / 'MAUGUG' (4, MT04AA) /
        'MATVGR' (5, MT05AA) ...) with OM op 'M=N', 'M=M+N' or 'M=M+TN'.
        flags: SUPG_U_P0 (u, v per element), SUPG_UPWIND (MATVGR's 'N'
        variant), SUPG_KSUPG (the element vectors from KSUPG(1, u, v))."""
        ne, n = self.nelem, self.npoin
        nu = ne if (flags & self.SUPG_U_P0) and not (flags & self.SUPG_KSUPG) else n
        if _is_torch(u):
            # device fields, the context's geometry
            if xel is not None or yel is not None or surfac is not None:
                raise BiefError("matrix_supg: device fields take the context's geometry")
            _dev_n(u, nu, "u")
            _dev_n(v, nu, "v")
            for a in (f, g):
                if a is not None:
                    _dev_n(a, ne, "f, g")
            self._sync_torch()
            _check(lib().bhip_matrix_supg_dev(
                self._h, int(kind), float(xmul), None if f is None else f.data_ptr(),
                None if g is None else g.data_ptr(), u.data_ptr(), v.data_ptr(), int(flags),
                self._OM_OPS[op]), "bhip_matrix_supg_dev")
            self._join_torch()
        else:
            geo = [_host_n_or_none(xel, 3 * ne, "xel"), _host_n_or_none(yel, 3 * ne, "yel"),
                   _host_n_or_none(surfac, ne, "surfac")]
            fg = [_host_n_or_none(a, ne, "f, g") for a in (f, g)]
            u, v = _host_n(u, nu, "u"), _host_n(v, nu, "v")
            _check(lib().bhip_matrix_supg(self._h, int(kind), float(xmul),
                                          *(_ptr(a) for a in geo), *(_ptr(a) for a in fg),
                                          u.ctypes.data, v.ctypes.data, int(flags),
                                          self._OM_OPS[op]), "bhip_matrix_supg")
        sym = kind == 4 and (op == "M=N" or (op == "M=M+N" and self.typext == "S"))
        self.typdia, self.typext = "Q", ("S" if sym else "Q")

    def matrix_om(self, op, c=0.0):
        """OM on the context's matrix with N = its saved copy (save_matrix):
        'M=X(M)', 'M=CN' (N = M), 'M=M+CN', 'M=M+N', 'M=M+TN' (om1111.f)."""
        _check(lib().bhip_matrix_om(self._h, op.ljust(8).encode(), float(c)), "bhip_matrix_om")
        if op.startswith("M=X(M)") or (op.startswith("M=M+") and
                                        getattr(self, "_saved_typext", "S") == "Q"):
            self.typext = "Q"

    def matrix_mt01pp(self, xmul, surfac, z, accumulate=False):
        """MATRIX(... 'MATMAS' ...) on prisms (mt01pp.f): surfac is the base
        triangle area of each prism, z the node elevations."""
        if _is_torch(z):
            if surfac is not None:
                raise BiefError("matrix_mt01pp: device fields take the context's geometry")
            _dev_n(z, self.npoin, "z")
            self._sync_torch()
            _check(lib().bhip_matrix_mt01pp_dev(self._h, float(xmul), z.data_ptr(),
                                                int(accumulate)), "bhip_matrix_mt01pp_dev")
            self._join_torch()
        else:
            s = _host_n_or_none(surfac, self.nelem, "surfac")
            z = _host_n(z, self.npoin, "z")
            _check(lib().bhip_matrix_mt01pp(self._h, float(xmul), _ptr(s), z.ctypes.data,
                                            int(accumulate)), "bhip_matrix_mt01pp")
        # 'M=M+N' keeps a non-symmetric target non-symmetric (both halves)
        self.typdia, self.typext = "Q", ("Q" if accumulate and self.typext == "Q" else "S")

    def matrix_mt02pp(self, xmul, xel, yel, surfac, z, f, g, h, npou0, h_zero=False,
                      inchyd=False, mon=False, accumulate=False):
        """MATRIX(... 'MATDIF' ...) on prisms (mt02pp.f): diagonal diffusion
        tensor (f, g, h) per node; npou0 = NPOIN2 (P1 vertical viscosity) or
        0 (P0 on the vertical)."""
        ne, n = self.nelem, self.npoin
        if _is_torch(z):
            if xel is not None or yel is not None or surfac is not None:
                raise BiefError("matrix_mt02pp: device fields take the context's geometry")
            for a in (z, f, g, h):
                _dev_n(a, n, "z, f, g, h")
            self._sync_torch()
            _check(lib().bhip_matrix_mt02pp_dev(self._h, float(xmul), z.data_ptr(), f.data_ptr(),
                                                g.data_ptr(), h.data_ptr(), int(h_zero),
                                                int(npou0), int(inchyd), int(mon),
                                                int(accumulate)), "bhip_matrix_mt02pp_dev")
            self._join_torch()
        else:
            arrs = [_host_n_or_none(xel, 3 * ne, "xel"), _host_n_or_none(yel, 3 * ne, "yel"),
                    _host_n_or_none(surfac, ne, "surfac")] + \
                [_host_n(a, n, "z, f, g, h") for a in (z, f, g, h)]
            _check(lib().bhip_matrix_mt02pp(self._h, float(xmul), *[_ptr(a) for a in arrs],
                                            int(h_zero), int(npou0), int(inchyd), int(mon),
                                            int(accumulate)), "bhip_matrix_mt02pp")
        self.typdia, self.typext = "Q", ("Q" if accumulate and self.typext == "Q" else "S")

    # ---- MATVEC ---------------------------------------------------------
    def _on_own_stream(self):
        import torch
        return torch.cuda.current_stream().cuda_stream == self.stream_ptr()

    def _sync_torch(self):
        """Orders the context stream after the work queued so far on torch's
        current stream (an event wait on the device, no host round trip);
        nothing to do when torch already runs on the context's stream."""
        import torch
        if self._on_own_stream():
            return
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream())
        self.torch_stream().wait_event(ev)

    def _join_torch(self):
        """Orders torch's current stream after the context stream's work."""
        import torch
        if self._on_own_stream():
            return
        ev = torch.cuda.Event()
        ev.record(self.torch_stream())
        torch.cuda.current_stream().wait_event(ev)

    def matvec(self, op: str, y, x=None, c: float = 0.0):
        """X = op(A, Y, C) as BIEF MATVEC. Host arrays: returns a new array.
        Torch CUDA tensors: x is updated in place on the context stream."""
        code = OPS[op]
        if _is_torch(y):
            import torch
            if x is None:
                x = torch.zeros_like(y)
            _dev_n(y, self.npoin, "matvec: y")
            _dev_n(x, self.npoin, "matvec: x")
            self._sync_torch()
            _check(lib().bhip_matvec_dev(self._h, code, float(c), y.data_ptr(), x.data_ptr()),
                   "bhip_matvec_dev")
            self._join_torch()
            return x
        y = _host_n(y, self.npoin, "matvec: y")
        x = np.zeros(self.npoin) if x is None else _host_n(x, self.npoin, "matvec: x").copy()
        _check(lib().bhip_matvec(self._h, code, float(c), y.ctypes.data, x.ctypes.data),
               "bhip_matvec")
        return x

    def perm(self):
        """The context's node order: internal node i is the caller's perm[i]."""
        p = np.empty(self.npoin, dtype=np.int32)
        _check(lib().bhip_ctx_perm(self._h, p.ctypes.data), "bhip_ctx_perm")
        return p

    def solve_internal(self, x, b, slv=1, precon=2, krylov=5, nitmax=1000, eps=1e-8,
                       precision=PREC_FP64, inner_eps=1e-4, inner_nitmax=0):
        """SOLVE on torch CUDA vectors in the context's internal order
        (bhip_solve_dev_internal); otherwise as solve()."""
        if not _is_torch(x) or not _is_torch(b):
            raise BiefError("solve_internal takes device tensors")
        _dev_n(x, self.npoin, "solve: x")
        _dev_n(b, self.npoin, "solve: b")
        cfg = SolverCfg(slv, precon, krylov, nitmax, eps, precision, inner_eps, inner_nitmax)
        info = SolveInfo()
        self._sync_torch()
        rc = lib().bhip_solve_dev_internal(self._h, ctypes.byref(cfg), x.data_ptr(),
                                           b.data_ptr(), ctypes.byref(info))
        _check(rc, "bhip_solve_dev_internal")
        self._join_torch()
        self.typdia = "I" if precon in (2,) else self.typdia
        d = info.as_dict()
        d["status"] = STATUS[rc]
        return x, d

    def time_cg_kernels(self, precision=PREC_FP64, reps=50):
        """Average ms per launch of the CG direction/mat-vec kernel and the
        X/R update kernel (HIP events on the context stream)."""
        a, b = _D(), _D()
        _check(lib().bhip_time_cg_kernels(self._h, precision, reps, ctypes.byref(a),
                                          ctypes.byref(b)), "bhip_time_cg_kernels")
        return a.value, b.value

    def time_cg_dir_cold(self, flush, precision=PREC_FP64, reps=20):
        """Average ms of the CG direction/mat-vec kernel launched after a
        read pass over `flush` (a float64 CUDA tensor larger than
        the last-level cache), so that every launch starts with none of its
        data cached: bhip_time_cg_dir_cold."""
        import torch
        if not (_is_torch(flush) and flush.is_cuda and flush.dtype == torch.float64
                and flush.is_contiguous()):
            raise BiefError("time_cg_dir_cold: flush must be a contiguous float64 CUDA tensor")
        ms = ctypes.c_double()
        self._sync_torch()
        _check(lib().bhip_time_cg_dir_cold(self._h, precision, reps,
                                           ctypes.c_void_p(flush.data_ptr()),
                                           flush.numel(), ctypes.byref(ms)),
               "bhip_time_cg_dir_cold")
        return ms.value

    def time_matvec(self, precision=PREC_FP64, reps=20):
        """Average ms of the plain EBE product on the current matrix (FP64, or
        its FP32 copy after a mixed solve): bhip_time_matvec."""
        ms = ctypes.c_double()
        _check(lib().bhip_time_matvec(self._h, precision, reps, ctypes.byref(ms)),
               "bhip_time_matvec")
        return ms.value

    # ---- SOLVE ----------------------------------------------------------
    def solve(self, x, b, slv=1, precon=2, krylov=5, nitmax=1000, eps=1e-8,
              precision=PREC_FP64, inner_eps=1e-4, inner_nitmax=0, restore=False):
        """SOLVE: x (initial guess, overwritten), b (overwritten by D*b as in
        the reference); the context's matrix is left preconditioned.
        restore=True solves with the saved matrix (restore_matrix() then
        SOLVE; with device vectors and diagonal preconditioning the copy is
        fused into the preconditioning, bhip_solve_dev_restored).
        Returns (x, info dict, status)."""
        cfg = SolverCfg(slv, precon, krylov, nitmax, eps, precision, inner_eps, inner_nitmax)
        info = SolveInfo()
        if _is_torch(x):
            _dev_n(x, self.npoin, "solve: x")
            _dev_n(b, self.npoin, "solve: b")
            self._sync_torch()
            fn = lib().bhip_solve_dev_restored if restore else lib().bhip_solve_dev
            rc = fn(self._h, ctypes.byref(cfg), x.data_ptr(), b.data_ptr(), ctypes.byref(info))
            _check(rc, "bhip_solve_dev_restored" if restore else "bhip_solve_dev")
            self._join_torch()
            if restore:
                self.typext = getattr(self, "_saved_typext", self.typext)
        else:
            if restore:
                self.restore_matrix()
            xh = _host_n(x, self.npoin, "solve: x")
            bh = _host_n(b, self.npoin, "solve: b")
            # SOLVE overwrites X and B (B with D*B): a writeable float64
            # array is passed as it is; anything else (a read-only view, a
            # list, another dtype) through a copy, then written back where the
            # caller's object can take it
            if not xh.flags.writeable:
                xh = xh.copy()
            if not bh.flags.writeable:
                bh = bh.copy()
            rc = lib().bhip_solve(self._h, ctypes.byref(cfg), xh.ctypes.data, bh.ctypes.data,
                                  ctypes.byref(info))
            _check(rc, "bhip_solve")
            if xh is not x and isinstance(x, np.ndarray) and x.dtype == np.float64 \
                    and x.flags.writeable:
                x[...] = xh
            if bh is not b and isinstance(b, np.ndarray) and b.dtype == np.float64 \
                    and b.flags.writeable:
                b[...] = bh
            x = xh
        self.typdia = "I" if precon in (2,) else self.typdia
        d = info.as_dict()
        d["status"] = STATUS[rc]
        return x, d


class WaveStep:
    """The device wave-equation step and its adjoint kernels on a context
    (bhip_wave_*): node vectors are torch CUDA float64 tensors in the
    context's internal order (Context.perm)."""

    def __init__(self, ctx: Context, xel, yel, surfac, sea_mask):
        self.ctx = ctx
        xel = _host_f64(xel)
        yel = _host_f64(yel)
        surfac = _host_f64(surfac)
        sea = np.ascontiguousarray(sea_mask, dtype=np.int32)
        if xel.size != 3 * ctx.nelem or sea.size != ctx.npoin:
            raise BiefError("wave setup: array sizes do not match the mesh")
        _check(lib().bhip_wave_setup(ctx._h, xel.ctypes.data, yel.ctypes.data, surfac.ctypes.data,
                                     sea.ctypes.data), "bhip_wave_setup")

    @staticmethod
    def params(dt, grav, eps_u, tide):
        return WaveParams(float(dt), float(grav), float(eps_u), float(tide))

    def _run(self, name, p, *tensors):
        for t in tensors:
            if not (_is_torch(t) and t.is_cuda and t.dtype.is_floating_point and
                    t.element_size() == 8 and t.is_contiguous() and t.numel() == self.ctx.npoin):
                raise BiefError(f"{name}: node vectors must be contiguous float64 CUDA tensors "
                                f"of {self.ctx.npoin} entries")
        self.ctx._sync_torch()
        _check(getattr(lib(), name)(self.ctx._h, ctypes.byref(p),
                                    *(t.data_ptr() for t in tensors)), name)
        self.ctx._join_torch()

    def assemble(self, p, eta, u, v, kstr, zb, rhs):
        self._run("bhip_wave_assemble", p, eta, u, v, kstr, zb, rhs)
        self.ctx.typdia, self.ctx.typext = "Q", "S"

    def momentum(self, p, eta, deta, u, v, eta_n, u_n, v_n):
        self._run("bhip_wave_momentum", p, eta, deta, u, v, eta_n, u_n, v_n)

    def momentum_adj(self, p, ubar_n, vbar_n, ebar_n, etot):
        self._run("bhip_wave_momentum_adj", p, ubar_n, vbar_n, ebar_n, etot)

    def assemble_adj(self, p, lam, deta, eta, u, v, u_n, v_n, ubar_n, vbar_n, etot, kstr,
                     ebar, ubar, vbar, kbar, zbbar):
        self._run("bhip_wave_assemble_adj", p, lam, deta, eta, u, v, u_n, v_n, ubar_n, vbar_n,
                  etot, kstr, ebar, ubar, vbar, kbar, zbbar)


class Ensemble(Context):
    """Several independent meshes in one context (bhip_batch_create): the
    vectors and element terms of all meshes one after the other; one launch
    covers all meshes, and CG (FP64 or mixed) solves every system with its
    own scalars, stop tests and refinement sweeps."""

    def __init__(self, meshes, block_nodes: int = 0):
        ndp = meshes[0].ndp
        if any(m.ndp != ndp for m in meshes):
            raise BiefError("an ensemble holds meshes of one element type")
        self.meshes = list(meshes)
        self.nsys = len(meshes)
        npo = np.array([m.npoin for m in meshes], dtype=np.int32)
        nel = np.array([m.nelem for m in meshes], dtype=np.int32)
        self.node_off = np.concatenate([[0], np.cumsum(npo)]).astype(np.int64)
        self.elem_off = np.concatenate([[0], np.cumsum(nel)]).astype(np.int64)
        self.npoin = int(self.node_off[-1])
        self.nelem = int(self.elem_off[-1])
        self.ndp = ndp
        self.mesh = None
        self._ikle = np.ascontiguousarray(
            np.concatenate([np.ascontiguousarray(m.ikle, dtype=np.int32).ravel() for m in meshes]))
        x = np.ascontiguousarray(np.concatenate([m.x for m in meshes]), dtype=np.float64)
        y = np.ascontiguousarray(np.concatenate([m.y for m in meshes]), dtype=np.float64)
        self._npo, self._nel = npo, nel
        h = _P()
        rc = lib().bhip_batch_create(self.nsys, npo.ctypes.data, nel.ctypes.data, ndp,
                                     self._ikle.ctypes.data, x.ctypes.data, y.ctypes.data,
                                     int(block_nodes), ctypes.byref(h))
        _check(rc, "bhip_batch_create")
        self._h = h
        self.typdia = "0"
        self.typext = "0"

    def split(self, v):
        """The per-mesh pieces of an ensemble node vector."""
        return [v[self.node_off[i]:self.node_off[i + 1]] for i in range(self.nsys)]

    def infos(self):
        """Outcome of the last solve for each mesh."""
        arr = (SolveInfo * self.nsys)()
        _check(lib().bhip_batch_info(self._h, self.nsys, arr), "bhip_batch_info")
        return [a.as_dict() for a in arr]

    def latencies_ms(self):
        """Completion time of each mesh in the last solve (ms of device
        clock from the solve's set-up kernel to the kernel that stopped it)."""
        out = np.zeros(self.nsys)
        _check(lib().bhip_batch_latency(self._h, self.nsys, out.ctypes.data),
               "bhip_batch_latency")
        return out
