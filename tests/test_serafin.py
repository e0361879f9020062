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


def _single_edge_nodes(ik):
    edges = np.sort(np.concatenate([ik[[0, 1]], ik[[1, 2]], ik[[2, 0]]], axis=1), axis=0)
    uniq, count = np.unique(edges, axis=1, return_counts=True)
    return np.unique(uniq[:, count == 1])


def test_gouttedo_boundary_matches_cli():
    s = SF.read(SLF)
    b = SF.read_cli(CLI, s.npoin)
    k = b.colour
    assert np.array_equal(k, np.arange(1, len(k) + 1))
    assert np.array_equal(s.ipobo[b.nbor], k)
    assert np.count_nonzero(s.ipobo) == b.nptfr
    assert np.array_equal(_single_edge_nodes(s.ikle), np.sort(b.nbor))
    # a closed basin: walls everywhere, no prescribed depth
    assert np.all(SF.limpro_depth(b) == SF.KNEU)
    fbor, limdir = SF.depth_dirichlet(b, s.npoin)
    assert not np.any(limdir == SF.KDIR) and not np.any(fbor)


def test_estimation_boundary_file():
    """The channel of the reference's friction-estimation example
    (examples/telemac2d/estimation): 85 boundary points, the velocity
    prescribed on the 11 upstream points (code 4 5 5: free depth), the depth
    prescribed on the 6 downstream points (5 4 4), walls elsewhere; the
    boundary nodes are those of the mesh's single-triangle edges and IPOBO
    numbers them in file order."""
    s = SF.read(os.path.join(GOLD, "geo_estimation.slf"))
    b = SF.read_cli(os.path.join(GOLD, "geo_estimation.cli"), s.npoin)
    assert b.nptfr == 85 and s.npoin == 319
    assert np.array_equal(s.ipobo[b.nbor], np.arange(1, 86))
    assert np.array_equal(_single_edge_nodes(s.ikle), np.sort(b.nbor))
    lim = SF.limpro_depth(b)
    assert np.count_nonzero(lim == SF.KDIR) == 6
    assert np.count_nonzero(lim == SF.KDDL) == 11
    assert np.count_nonzero(lim == SF.KNEU) == 85 - 17
    assert set(b.lihbor[lim == SF.KDIR]) == {SF.KENT}
    assert set(b.liubor[lim == SF.KDIR]) == {SF.KSORT}
    # the prescribed-depth nodes lie on one cross-section of the channel
    xd = s.x[b.nbor[lim == SF.KDIR]]
    assert np.ptp(xd) < 1e-3 * np.ptp(s.x)
    fbor, limdir = SF.depth_dirichlet(b, s.npoin, values=1.5)
    assert np.array_equal(np.nonzero(limdir == SF.KDIR)[0], np.sort(b.nbor[lim == SF.KDIR]))
    assert np.all(fbor[limdir == SF.KDIR] == 1.5) and np.count_nonzero(fbor) == 6


def test_cli_corrections_and_errors(tmp_path):
    """LECLIM's corrections (leclim.f:158-185) and the reader's refusals."""
    line = "{} {} {}  0.000  1.000  2.000  0.0   2  0.000  0.000  0.000 {} {}  extra text\n"
    p = tmp_path / "a.cli"
    p.write_text(line.format(0, 0, 5, 1, 1) + line.format(1, 6, 4, 2, 2) + "\n")
    b = SF.read_cli(p, 2)
    assert list(b.lihbor) == [SF.KLOG, SF.KSORT]   # no-slip -> wall, incident -> free exit
    assert list(b.ubor) == [0.0, 1.0] and list(b.vbor) == [2.0, 0.0]
    assert list(b.nbor) == [0, 1]
    with pytest.raises(ValueError):
        SF.read_cli(p, 1)                            # node 2 of a 1-node mesh
    (tmp_path / "s.cli").write_text("2 2 2 0 0 0 0 2 0 0 0 1\n")
    with pytest.raises(ValueError):
        SF.read_cli(tmp_path / "s.cli")
    (tmp_path / "c.cli").write_text(line.format(3, 2, 2, 1, 1))
    with pytest.raises(ValueError):
        SF.limpro_depth(SF.read_cli(tmp_path / "c.cli"))


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
