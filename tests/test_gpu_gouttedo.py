"""The HIP path on a real mesh: the reference's gouttedo geometry read from its
SERAFIN file (tests/golden/geo_gouttedo.slf), first step of the drop, against
the reference's own solves (tests/golden/ref_gouttedo.npz, written by the
flang build of sources/utils/bief) in element and in segment storage."""
import os

import numpy as np
import pytest

from opentelemac_amd import bief, mesh as M, serafin as SF, system as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def case():
    m = SF.read(os.path.join(GOLD, "geo_gouttedo.slf")).mesh()
    gl, es = M.segments(m.ikle)
    sysd = S.drop_system(m)
    z = np.load(os.path.join(GOLD, "ref_gouttedo.npz"))
    return m, gl, es, sysd, z


@pytest.mark.parametrize("sto", ["ebe", "seg"])
@pytest.mark.parametrize("slv", [1, 6, 7])
def test_drop_step_matches_reference(gpu, case, sto, slv):
    m, gl, es, sysd, z = case
    if sto == "ebe":
        ctx, xa = bief.Context(m), sysd["xa"]
    else:
        sm, _ = M.segment_mesh(m)
        ctx, xa = bief.Context(sm), S.ebe_to_segments(sysd["xa"], es, m.ikle, gl, "S")
    ctx.set_matrix("Q", "S", sysd["da"], xa)
    x, info = ctx.solve(np.zeros(m.npoin), sysd["b"].copy(), slv=slv, precon=2, krylov=3,
                        eps=1e-10, nitmax=500)
    xr = z[f"{sto}_{slv}_x"]
    assert info["converged"]
    assert abs(info["niter"] - int(z[f"{sto}_{slv}_niter"])) <= 1, info
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_drop_step_mixed(gpu, case):
    m, gl, es, sysd, z = case
    ctx = bief.Context(m)
    ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
    x, info = ctx.solve(np.zeros(m.npoin), sysd["b"].copy(), slv=1, precon=2, eps=1e-10,
                        nitmax=500, precision=bief.PREC_MIXED)
    xr = z["ebe_1_x"]
    assert info["converged"]
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
