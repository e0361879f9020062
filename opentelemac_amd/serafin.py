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
