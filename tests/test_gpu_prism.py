"""GPU parity on 6-node prisms (MV0606 path, MT01PP mass matrix)."""
import numpy as np
import pytest

from opentelemac_amd import bief, mesh as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _prisms(nx=7, ny=6, nplan=5, seed=3):
    base = M.random_small_mesh(nx, ny, seed=seed)
    rng = np.random.default_rng(seed + 100)
    bot = -rng.uniform(2, 6, base.npoin)
    return base, M.extrude_prisms(base, nplan, bot, np.zeros(base.npoin))


@pytest.mark.parametrize("bn", [0, 30])
@pytest.mark.parametrize("typext", ["S", "Q"])
@pytest.mark.parametrize("typdia", ["Q", "I"])
def test_prism_matvec_all_ops(gpu, bn, typext, typdia):
    base, p = _prisms()
    rng = np.random.default_rng(1)
    nc = 30 if typext == "Q" else 15
    da = rng.uniform(1, 2, p.npoin)
    xa = rng.normal(size=(nc, p.nelem))
    ctx = bief.Context(p, block_nodes=bn)
    ctx.set_matrix(typdia, typext, da, xa)
    A = O.Matrix(da, xa, typdia, typext)
    for op in bief.OPS:
        y = rng.normal(size=p.npoin)
        x0 = rng.normal(size=p.npoin)
        got = ctx.matvec(op, y, x0, c=0.3)
        ref = O.matvec(op, A, p.ikle, y, x0, c=0.3)
        assert np.abs(got - ref).max() <= 1e-13 * (np.abs(ref).max() + 1), op


@pytest.mark.parametrize("slv", [1, 6, 7])
def test_prism_solve(gpu, slv):
    base, p = _prisms(9, 8, 6, seed=5)
    s = np.tile(base.surfac(), p.nplan - 1)
    t, xm = O.mt01pp(s, p.z, p.ikle, 1.0)
    da = O.assemble_diag(t, p.ikle, p.npoin)
    b = np.random.default_rng(2).normal(size=p.npoin)
    A = O.Matrix(da, xm, "Q", "S")
    x_ref, info_ref, _, _ = O.solve(A, p.ikle, b, slv=slv, precon=2, krylov=5, eps=1e-10)
    ctx = bief.Context(p, block_nodes=60)
    for prec in ((bief.PREC_FP64, bief.PREC_MIXED) if slv == 1 else (bief.PREC_FP64,)):
        ctx.set_matrix("Q", "S", da, xm)
        x, info = ctx.solve(np.zeros(p.npoin), b.copy(), slv=slv, precon=2, krylov=5,
                            eps=1e-10, precision=prec)
        rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        assert info["converged"], info
        if prec == bief.PREC_FP64:
            assert abs(info["niter"] - info_ref.niter) <= 1
            assert rel <= 1e-10
        else:
            assert rel <= 1e-8
