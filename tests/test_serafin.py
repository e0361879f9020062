"""SERAFIN reader and writer against the reference's own mesh files.

The fixture tests/golden/geo_gouttedo.slf is the geometry of the reference's
telemac2d gouttedo example (examples/telemac2d/gouttedo/geo_gouttedo.slf),
with its boundary-conditions file geo_gouttedo.cli. What pins the reader:
the example's documentation gives the mesh size (8978 triangles, 4624 nodes;
examples/telemac2d/gouttedo/doc/gouttedo.tex:20); the domain is the square
[0, 20.1]^2, so the element areas sum to 20.1^2 and are all positive
(counter-clockwise IKLE, read element by element); the boundary nodes of the
.cli file (columns N and K) are exactly the nodes on edges that one triangle
alone uses, and IPOBO(N) = K there, 0 elsewhere. A writer that reproduces the
original bytes pins the record layout in the other direction.
"""
import os

import numpy as np
import pytest

from opentelemac_amd import serafin as SF

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SLF = os.path.join(GOLD, "geo_gouttedo.slf")
CLI = os.path.join(GOLD, "geo_gouttedo.cli")
# the reference's own reader/writer test data: one result in four encodings
# (examples/python27/selafin_rw); read only where the reference tree exists
RW = "/root/reference/examples/python27/selafin_rw"


def test_gouttedo_geometry():
    s = SF.read(SLF)
    assert s.ikle.shape == (3, 8978) and s.npoin == 4624
    assert not s.double and s.big_endian
    m = s.mesh()
    xe, ye = m.xel_yel()
    area = 0.5 * (xe[1] * ye[2] - xe[2] * ye[1])
    assert area.min() > 0
    side = float(s.x.max())
    assert side == pytest.approx(20.1, abs=1e-5) and float(s.y.max()) == side
    assert area.sum() == pytest.approx(side * side, rel=1e-9)
    # every node is used
    assert np.array_equal(np.unique(s.ikle), np.arange(s.npoin))


def test_gouttedo_boundary_matches_cli():
    s = SF.read(SLF)
    cli = np.loadtxt(CLI)
    n = cli[:, 11].astype(np.int64)
    k = cli[:, 12].astype(np.int64)
    assert np.array_equal(k, np.arange(1, len(k) + 1))
    assert np.array_equal(s.ipobo[n - 1], k)
    assert np.count_nonzero(s.ipobo) == len(n)
    ik = s.ikle
    edges = np.sort(np.concatenate([ik[[0, 1]], ik[[1, 2]], ik[[2, 0]]], axis=1), axis=0)
    uniq, count = np.unique(edges, axis=1, return_counts=True)
    assert np.array_equal(np.unique(uniq[:, count == 1]), np.sort(n - 1))


def test_writer_reproduces_file_bytes(tmp_path):
    s = SF.read(SLF)
    out = tmp_path / "rt.slf"
    SF.write(out, s)
    assert out.read_bytes() == open(SLF, "rb").read()


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("big", [False, True])
def test_round_trip_results(tmp_path, double, big):
    """A result file with time steps, in both precisions and byte orders."""
    s = SF.read(SLF)
    rng = np.random.default_rng(1)
    s.double, s.big_endian = double, big
    s.title = "GOUTTEDO RESULTS"
    s.varnames = ["WATER DEPTH     M", "VELOCITY U      M/S"]
    s.iparam = s.iparam.copy()
    s.iparam[9] = 1
    s.date = (2024, 5, 17, 12, 30, 0)
    s.times = [0.0, 0.5, 1.25]
    s.values = [rng.normal(size=(2, s.npoin)) for _ in s.times]
    SF.write(tmp_path / "r.slf", s)
    r = SF.read(tmp_path / "r.slf")
    assert r.double == double and r.big_endian == big and r.date == s.date
    assert r.title[72:] == ("SERAFIND" if double else "SERAFIN ")
    assert [v.strip() for v in r.varnames] == s.varnames
    assert r.times == s.times
    cast = np.float64 if double else np.float32
    for a, b in zip(r.values, s.values):
        assert np.array_equal(a, b.astype(cast))
    assert np.array_equal(r.ikle, s.ikle) and np.array_equal(r.ipobo, s.ipobo)


def test_truncated_file_is_rejected(tmp_path):
    data = open(SLF, "rb").read()
    (tmp_path / "t.slf").write_bytes(data[: len(data) - 100])
    with pytest.raises(ValueError):
        SF.read(tmp_path / "t.slf")


def test_not_serafin(tmp_path):
    (tmp_path / "n.slf").write_bytes(b"\x00\x00\x00\x05hello\x00\x00\x00\x05")
    with pytest.raises(ValueError):
        SF.read(tmp_path / "n.slf")


@pytest.mark.skipif(not os.path.isdir(RW), reason="reference tree not present")
def test_four_encodings_agree():
    """One result in FP32/FP64 and big/little-endian, from the reference's
    own reader/writer example: the same mesh, names and values per name."""
    enc = {k: SF.read(os.path.join(RW, f"r2d_bowl_vf_gb_{k}.slf"))
           for k in ("DoublePrecision_BigEndian", "DoublePrecision_LittleEndian",
                     "SinglePrecision_BigEndian", "SinglePrecision_LittleEndian")}
    base = enc["SinglePrecision_BigEndian"]
    for k, s in enc.items():
        assert s.double == k.startswith("Double") and s.big_endian == k.endswith("BigEndian")
        assert np.array_equal(s.ikle, base.ikle) and np.array_equal(s.ipobo, base.ipobo)
        assert np.array_equal(s.x.astype(np.float32), base.x)
        names = [v.strip() for v in s.varnames]
        for j, v in enumerate(base.varnames):
            got = s.values[0][names.index(v.strip())]
            assert np.array_equal(got.astype(np.float32), base.values[0][j])
