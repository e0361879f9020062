"""SERAFIN / SELAFIN mesh and result files (the reference's binary format,
documentation/Misc/developer_guide/latex/A2-selafin_format.tex; reader and
writer in sources/utils/hermes/utils_serafin.F).

Fortran sequential unformatted records, each framed by its byte length (a
4-byte integer before and after). Big-endian by convention; the reader also
takes little-endian files, recognised from the first record (80 bytes). The
last 8 characters of the title give the precision: 'SERAFIND' stores reals
in FP64, anything else ('SERAFIN ') in FP32. Records:
  title (80 characters); NBV1, NBV2; NBV1 + NBV2 variable names (32
  characters each); IPARAM (10 integers); the date (6 integers) if
  IPARAM(10) = 1; NELEM, NPOIN, NDP, 1; IKLE (NDP per element, element by
  element, 1-based); IPOBO (NPOIN); X; Y; then per time step the time and
  one record per variable.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

from .mesh import Mesh


@dataclass
class Serafin:
    title: str
    varnames: list
    ikle: np.ndarray          # (ndp, nelem), 0-based
    ipobo: np.ndarray         # (npoin,)
    x: np.ndarray
    y: np.ndarray
    iparam: np.ndarray        # (10,)
    date: tuple | None = None
    times: list = field(default_factory=list)
    values: list = field(default_factory=list)   # per time step: (nvar, npoin)
    double: bool = False
    big_endian: bool = True

    @property
    def npoin(self):
        return self.x.size

    def mesh(self) -> Mesh:
        """The geometry as a Mesh (IPARAM(7) planes for prism files)."""
        nplan = int(self.iparam[6]) if self.iparam[6] > 0 else 1
        return Mesh(x=self.x.astype(np.float64), y=self.y.astype(np.float64),
                    ikle=np.ascontiguousarray(self.ikle.astype(np.int32)), nplan=nplan,
                    npoin2=self.npoin // nplan, nelem2=self.ikle.shape[1] // max(1, nplan - 1)
                    if nplan > 1 else self.ikle.shape[1], meta={"title": self.title.strip()})


class _Records:
    def __init__(self, buf: bytes):
        self.buf = buf
        self.pos = 0
        first = struct.unpack(">i", buf[:4])[0]
        self.e = ">" if first == 80 else "<"
        if struct.unpack(self.e + "i", buf[:4])[0] != 80:
            raise ValueError("not a SERAFIN file: the first record is not an 80-byte title")

    def more(self):
        # fewer than 8 bytes cannot hold a record: trailing padding (some
        # files in the wild end with a stray byte) is ignored
        return self.pos + 8 <= len(self.buf)

    def next(self) -> bytes:
        if self.pos + 8 > len(self.buf):
            raise ValueError(f"file ends inside a record at byte {self.pos}")
        (n,) = struct.unpack(self.e + "i", self.buf[self.pos:self.pos + 4])
        if n < 0 or self.pos + 8 + n > len(self.buf):
            raise ValueError(f"file ends inside a record at byte {self.pos}")
        body = self.buf[self.pos + 4:self.pos + 4 + n]
        (n2,) = struct.unpack(self.e + "i", self.buf[self.pos + 4 + n:self.pos + 8 + n])
        if n2 != n or len(body) != n:
            raise ValueError(f"corrupt record at byte {self.pos}")
        self.pos += 8 + n
        return body

    def ints(self, count=None):
        b = self.next()
        a = np.frombuffer(b, dtype=self.e + "i4")
        if count is not None and a.size != count:
            raise ValueError(f"expected {count} integers, found {a.size}")
        return a

    def reals(self, double):
        return np.frombuffer(self.next(), dtype=self.e + ("f8" if double else "f4"))


def read(path) -> Serafin:
    with open(path, "rb") as fh:
        r = _Records(fh.read())
    title = r.next().decode("latin-1")
    double = title[72:80] == "SERAFIND"
    nbv = r.ints(2)
    names = [r.next().decode("latin-1") for _ in range(int(nbv[0] + nbv[1]))]
    iparam = r.ints(10).copy()
    date = tuple(int(v) for v in r.ints(6)) if iparam[9] == 1 else None
    nelem, npoin, ndp, _ = (int(v) for v in r.ints(4))
    ikle = r.ints(nelem * ndp).reshape(nelem, ndp).T.astype(np.int64) - 1
    ipobo = r.ints(npoin).copy()
    # the coordinates' record length settles the precision where a title
    # lacks the marker (files written in FP64 without it exist)
    xb = r.next()
    if len(xb) == 8 * npoin:
        double = True
    elif len(xb) != 4 * npoin:
        raise ValueError(f"X record of {len(xb)} bytes for {npoin} points")
    x = np.frombuffer(xb, dtype=r.e + ("f8" if double else "f4")).copy()
    y = r.reals(double).copy()
    out = Serafin(title=title, varnames=names, ikle=ikle, ipobo=ipobo, x=x, y=y,
                  iparam=iparam, date=date, double=double, big_endian=r.e == ">")
    nv = len(names)
    while r.more():
        t = r.reals(double)
        out.times.append(float(t[0]))
        out.values.append(np.stack([r.reals(double).copy() for _ in range(nv)]) if nv
                          else np.zeros((0, npoin)))
    return out


def write(path, s: Serafin):
    """Writes the records in the byte order the object names (big-endian
    unless read from a little-endian file)."""
    e = ">" if s.big_endian else "<"
    rtype = "f8" if s.double else "f4"

    def rec(body: bytes):
        return struct.pack(e + "i", len(body)) + body + struct.pack(e + "i", len(body))

    # the precision marker goes in the title's last 8 characters; other text
    # already there (older files carry any) is kept for FP32 files
    title = s.title.ljust(80)[:80]
    if s.double:
        title = title[:72] + "SERAFIND"
    elif title[72:] in ("SERAFIND", " " * 8):
        title = title[:72] + "SERAFIN "
    out = [rec(title.encode("latin-1")),
           rec(np.array([len(s.varnames), 0], dtype=e + "i4").tobytes())]
    out += [rec(n.ljust(32)[:32].encode("latin-1")) for n in s.varnames]
    out.append(rec(np.asarray(s.iparam, dtype=e + "i4").tobytes()))
    if s.iparam[9] == 1:
        out.append(rec(np.asarray(s.date, dtype=e + "i4").tobytes()))
    ndp, nelem = s.ikle.shape
    out.append(rec(np.array([nelem, s.npoin, ndp, 1], dtype=e + "i4").tobytes()))
    out.append(rec((s.ikle.T + 1).astype(e + "i4").tobytes()))
    out.append(rec(np.asarray(s.ipobo, dtype=e + "i4").tobytes()))
    out.append(rec(np.asarray(s.x, dtype=e + rtype).tobytes()))
    out.append(rec(np.asarray(s.y, dtype=e + rtype).tobytes()))
    for t, vals in zip(s.times, s.values):
        out.append(rec(np.array([t], dtype=e + rtype).tobytes()))
        for v in vals:
            out.append(rec(np.asarray(v, dtype=e + rtype).tobytes()))
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


# ---------------------------------------------------------------------------
# Boundary-conditions file (.cli), read as LECLIM does
# ---------------------------------------------------------------------------
# Codes of the boundary-conditions file (sources/utils/bief/declarations_telemac.f:30-45)
KENT, KENTU, KSORT, KADH, KLOG, KINC = 5, 6, 4, 0, 2, 1
# Technical codes the solvers work with (declarations_telemac.f:58-68)
KNEU, KDIR, KDDL = 1, 2, 3


@dataclass
class Boundary:
    """The columns of a boundary-conditions file, one entry per boundary
    point in file order (the boundary numbering K). nbor is 0-based."""
    lihbor: np.ndarray
    liubor: np.ndarray
    livbor: np.ndarray
    hbor: np.ndarray
    ubor: np.ndarray
    vbor: np.ndarray
    chbord: np.ndarray
    litbor: np.ndarray
    tbor: np.ndarray
    atbor: np.ndarray
    btbor: np.ndarray
    nbor: np.ndarray
    colour: np.ndarray

    @property
    def nptfr(self):
        return self.nbor.size


def read_cli(path, npoin=None) -> Boundary:
    """Reads a boundary-conditions file as GET_BND_VALUE_SRF
    (sources/utils/hermes/utils_serafin.F:1608-1630) and GET_BND_NUMBERING
    do: one list-directed record of 13 values per boundary point, LIHBOR,
    LIUBOR, LIVBOR, HBOR, UBOR, VBOR, AUBOR (CHBORD), LITBOR, TBOR, ATBOR,
    BTBOR, N (the global node number, 1-based) and K (the colour); whatever
    follows the 13th value on a line is ignored. Then LECLIM's corrections
    (sources/utils/bief/leclim.f:158-185): a no-slip code on the depth
    becomes a wall, an incident-wave code on the depth a free exit, and
    UBOR/VBOR are zeroed where the velocity is not prescribed.

    Raises ValueError on a short record, a non-numeric field or (npoin
    given) a node number outside 1..npoin, where the reference stops."""
    rows = []
    with open(path, "r", encoding="latin-1") as fh:
        for ln, line in enumerate(fh, 1):
            tok = line.split()
            if not tok:
                continue
            if len(tok) < 13:
                raise ValueError(f"{path}:{ln}: {len(tok)} values, 13 expected")
            try:
                rows.append([float(t) for t in tok[:13]])
            except ValueError as e:
                raise ValueError(f"{path}:{ln}: {e}") from None
    a = np.array(rows, dtype=np.float64).reshape(-1, 13)
    ints = lambda j: a[:, j].astype(np.int64)  # noqa: E731
    b = Boundary(lihbor=ints(0), liubor=ints(1), livbor=ints(2), hbor=a[:, 3].copy(),
                 ubor=a[:, 4].copy(), vbor=a[:, 5].copy(), chbord=a[:, 6].copy(),
                 litbor=ints(7), tbor=a[:, 8].copy(), atbor=a[:, 9].copy(),
                 btbor=a[:, 10].copy(), nbor=ints(11) - 1, colour=ints(12))
    if b.nbor.size and (b.nbor.min() < 0 or (npoin is not None and b.nbor.max() >= npoin)):
        raise ValueError(f"{path}: boundary node number outside 1..{npoin}")
    # LECLIM (leclim.f:158-185)
    b.lihbor[b.lihbor == KADH] = KLOG
    b.lihbor[b.lihbor == KINC] = KSORT
    b.ubor[(b.liubor != KENT) & (b.liubor != KENTU)] = 0.0
    b.vbor[(b.livbor != KENT) & (b.livbor != KENTU)] = 0.0
    return b


def limpro_depth(b: Boundary) -> np.ndarray:
    """Technical condition of the depth at each boundary point, as telemac2d
    sets LIMPRO(K,1) for the propagation step
    (sources/telemac2d/propin_telemac2d.f:246-258): prescribed value -> KDIR,
    free exit -> KDDL, wall -> KNEU; any other code is an error."""
    out = np.empty(b.nptfr, dtype=np.int32)
    out[b.lihbor == KENT] = KDIR
    out[b.lihbor == KSORT] = KDDL
    out[b.lihbor == KLOG] = KNEU
    bad = ~np.isin(b.lihbor, (KENT, KSORT, KLOG))
    if bad.any():
        k = int(np.nonzero(bad)[0][0])
        raise ValueError(f"boundary point {k + 1}: depth code {int(b.lihbor[k])} "
                         "is not a prescribed value, a free exit or a wall")
    return out


def depth_dirichlet(b: Boundary, npoin, values=None):
    """Node arrays for DIRICH on the depth: (fbor, limdir) of length npoin,
    LIMDIR = KDIR at the boundary nodes whose depth is prescribed and KDDL
    elsewhere, FBOR the prescribed value there (values: per boundary point,
    default HBOR) and 0 elsewhere. Pass kdir=KDIR to DIRICH with them."""
    lim = limpro_depth(b)
    fbor = np.zeros(npoin)
    limdir = np.full(npoin, KDDL, dtype=np.int32)
    d = lim == KDIR
    v = b.hbor if values is None else np.broadcast_to(np.asarray(values, dtype=np.float64),
                                                      (b.nptfr,))
    fbor[b.nbor[d]] = v[d]
    limdir[b.nbor[d]] = KDIR
    return fbor, limdir
