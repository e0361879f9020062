"""SUPG element matrices (MT03AA, MT04AA, MT05AA, KSUPG) and the OM
operations on the device, CVDFTR's SUPG tracer system assembled on the
device, and its GMRES(5) / BiCGSTAB solves against the reference's own
(tests/golden/ref_supg.npz, written by the flang build of
sources/utils/bief)."""
import os

import numpy as np
import pytest

from opentelemac_amd import bief, mesh as M, serafin as SF, tracer
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
K = bief.Context.SUPG_KSUPG


def _close(got, ref, tol=1e-13):
    return np.abs(got - ref).max() <= tol * (np.abs(ref).max() + 1e-300)


@pytest.fixture(scope="module")
def small():
    m = M.random_small_mesh(14, 10, seed=91)
    rng = np.random.default_rng(7)
    u, v = rng.uniform(-2, 2, m.npoin), rng.uniform(-2, 2, m.npoin)
    return m, m.xel_yel(), m.surfac(), u, v


def _oracle(m, t, xm):
    return O.assemble_diag(t, m.ikle, m.npoin), xm


@pytest.mark.parametrize("bn", [0, 40])
def test_element_matrices_match_oracle(gpu, small, bn):
    m, (xe, ye), sf, u, v = small
    ctx = bief.Context(m, block_nodes=bn)
    kx, ky = O.kspg11(xe, ye, m.ikle, u, v, 1.0)
    cases = [
        (dict(kind=3, xmul=0.7, f=kx, g=ky), O.mt03aa(kx, ky, u, v, xe, ye, sf, m.ikle, 0.7)),
        (dict(kind=3, xmul=0.7, flags=K), O.mt03aa(kx, ky, u, v, xe, ye, sf, m.ikle, 0.7)),
        (dict(kind=4, xmul=0.3), O.mt04aa(u, v, xe, ye, sf, m.ikle, 0.3)),
        (dict(kind=5, xmul=1.3), O.mt05aa(u, v, xe, ye, m.ikle, 1.3)),
        (dict(kind=5, xmul=1.0, flags=bief.Context.SUPG_UPWIND),
         O.mt05aa(u, v, xe, ye, m.ikle, 1.0, upwind=True)),
        (dict(kind=5, xmul=2.0, flags=K), O.mt05aa(kx, ky, xe, ye, m.ikle, 2.0, u_p0=True)),
    ]
    for kw, (t, xm) in cases:
        kind, xmul = kw.pop("kind"), kw.pop("xmul")
        if kw.get("flags") == bief.Context.SUPG_U_P0:
            ctx.matrix_supg(kind, xmul, xe, ye, sf, kx, ky, **kw)
        else:
            ctx.matrix_supg(kind, xmul, xe, ye, sf, u, v, **kw)
        td, da, xa = ctx.get_matrix()
        d_ref, x_ref = _oracle(m, t, xm)
        assert td == "Q" and ctx.typext == ("S" if kind == 4 else "Q")
        assert _close(da, d_ref), (kind, kw)
        assert _close(xa, x_ref), (kind, kw)
    # given per-element velocities (P0 MATVGR)
    ctx.matrix_supg(5, 2.0, xe, ye, sf, kx, ky, flags=bief.Context.SUPG_U_P0)
    _, da, xa = ctx.get_matrix()
    t, xm = O.mt05aa(kx, ky, xe, ye, m.ikle, 2.0, u_p0=True)
    assert _close(da, O.assemble_diag(t, m.ikle, m.npoin)) and _close(xa, xm)


def test_accumulate_transpose_and_om(gpu, small):
    """'M=M+N', 'M=M+TN', 'M=X(M)', 'M=CN', 'M=M+CN' against the same
    combinations of the oracle's element matrices."""
    m, (xe, ye), sf, u, v = small
    ctx = bief.Context(m)
    d = lambda t: O.assemble_diag(t, m.ikle, m.npoin)
    t1, x1 = O.mt01aa(sf, 0.5)
    t5, x5 = O.mt05aa(u, v, xe, ye, m.ikle, 1.0)
    t4, x4 = O.mt04aa(u, v, xe, ye, sf, m.ikle, 0.2)
    # symmetric mass, + symmetric MAUGUG stays 'S'
    ctx.matrix_mt01aa(0.5, sf)
    ctx.matrix_supg(4, 0.2, xe, ye, sf, u, v, op="M=M+N")
    _, da, xa = ctx.get_matrix()
    assert ctx.typext == "S" and _close(da, d(t1) + d(t4)) and _close(xa, x1 + x4)
    # + transposed MATVGR turns it 'Q'
    ctx.matrix_supg(5, 1.0, xe, ye, sf, u, v, op="M=M+TN")
    _, da, xa = ctx.get_matrix()
    sym = np.concatenate([x1 + x4, x1 + x4])
    assert ctx.typext == "Q"
    assert _close(da, d(t1) + d(t4) + d(t5))
    assert _close(xa, sym + np.concatenate([x5[3:], x5[:3]]))
    # save N = that, M = MATMAS made 'Q', M = 2 M, M = M + 0.5 N, M = M + N^T
    n_d, n_x = da.copy(), xa.copy()
    ctx.save_matrix()
    ctx.matrix_mt01aa(0.5, sf)
    ctx.matrix_om("M=X(M)")
    _, da, xa = ctx.get_matrix()
    assert ctx.typext == "Q" and _close(xa, np.concatenate([x1, x1]))
    ctx.matrix_om("M=CN", 2.0)
    ctx.matrix_om("M=M+CN", 0.5)
    _, da, xa = ctx.get_matrix()
    assert _close(da, 2.0 * d(t1) + 0.5 * n_d)
    assert _close(xa, 2.0 * np.concatenate([x1, x1]) + 0.5 * n_x)
    ctx.matrix_om("M=M+TN")
    _, da2, xa2 = ctx.get_matrix()
    assert _close(da2, da + n_d) and _close(xa2, xa + np.concatenate([n_x[3:], n_x[:3]]))
    # the diffusion accumulated into a 'Q' matrix goes to both halves
    td, xd = O.mt02aa(1, xe, ye, sf, m.ikle, np.full(m.npoin, 0.1), None, 1.0)
    ctx.matrix_mt02aa(1, 1.0, xe, ye, sf, np.full(m.npoin, 0.1), accumulate=True)
    _, da3, xa3 = ctx.get_matrix()
    assert ctx.typext == "Q" and _close(da3, da2 + d(td))
    assert _close(xa3, xa2 + np.concatenate([xd, xd]))


@pytest.fixture(scope="module")
def drop():
    g = SF.read(os.path.join(GOLD, "geo_gouttedo.slf")).mesh()
    w = 2.0 / (10.05 * np.sqrt(2.0))
    gu, gv = -w * (g.y - 10.05), w * (g.x - 10.05)
    fn = np.exp(-((g.x - 6.0) ** 2 + (g.y - 10.0) ** 2) / 4.0)
    return g, gu, gv, fn, np.load(os.path.join(GOLD, "ref_supg.npz"))


def test_supg_system_assembled_on_device(gpu, drop):
    g, gu, gv, fn, z = drop
    ctx = bief.Context(g)
    sm = tracer.supg_system(ctx, g, gu, gv, np.full(g.npoin, 0.1), 0.5, fn)
    td, da, xa = ctx.get_matrix()
    assert td == "Q" and ctx.typext == "Q"
    assert _close(da, z["sys_da"]) and _close(xa, z["sys_xa"]) and _close(sm, z["sys_sm"])


@pytest.mark.parametrize("slv", [7, 6])
def test_supg_solves_match_reference(gpu, drop, slv):
    g, gu, gv, fn, z = drop
    ctx = bief.Context(g, block_nodes=600)
    sm = tracer.supg_system(ctx, g, gu, gv, np.full(g.npoin, 0.1), 0.5, fn)
    x, info = ctx.solve(fn.copy(), sm, slv=slv, precon=2, krylov=5, eps=1e-10, nitmax=500)
    xr = z[f"solve_{slv}_x"]
    nr = int(z[f"solve_{slv}_niter"])
    assert info["converged"]
    tol_it = 1 if slv == 7 else max(3, nr // 10)
    assert abs(info["niter"] - nr) <= tol_it, (info, nr)
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_supg_mixed_gmres(gpu, drop):
    """GMRES(5) in mixed precision on the reference's CVDFTR system of the
    gouttedo mesh: FP32 Arnoldi products and basis inside FP64 restarts,
    each restart an FP64 refinement step whose true residual carries the
    reference's stop rule. It stops by that rule (checked again here from
    the solution with the oracle's FP64 product), in the reference's number
    of restarts +-1, at the reference's solution within what the tolerance
    leaves (1e-8 relative)."""
    g, gu, gv, fn, z = drop
    eps = 1e-10
    ctx = bief.Context(g, block_nodes=600)
    sm = tracer.supg_system(ctx, g, gu, gv, np.full(g.npoin, 0.1), 0.5, fn)
    td, da, xa = ctx.get_matrix()
    b0 = sm.copy()
    x, info = ctx.solve(fn.copy(), sm, slv=7, precon=2, krylov=5, eps=eps, nitmax=500,
                        precision=bief.PREC_MIXED)
    xr = z["solve_7_x"]
    nr = int(z["solve_7_niter"])
    assert info["converged"] and info["relres"] < eps, info
    assert abs(info["niter"] - nr) <= 1, (info, nr)
    assert info["outer_iters"] == info["niter"] and info["inner_iters"] >= 5 * info["niter"]
    # the stop rule recomputed: ||D (A x - b)|| / max(||D b||, 1) < eps
    A = O.Matrix(da, xa, td, "Q")
    d = 1.0 / np.sqrt(np.abs(da))
    r = d * (O.matvec("X=AY", A, g.ikle, x) - b0)
    assert np.linalg.norm(r) / max(np.linalg.norm(d * b0), 1.0) < 1.01 * eps
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)


def test_device_geometry_matches_caller_arrays(gpu, small):
    """Element matrices from the geometry kept on the device (set_geometry,
    then xel = yel = surfac = None) are the same bits as from the caller's
    arrays, for every element kind (same layout, same kernels)."""
    m, (xe, ye), sf, u, v = small
    a, b = bief.Context(m, block_nodes=40), bief.Context(m, block_nodes=40)
    b.set_geometry(xe, ye, sf)
    kx, ky = O.kspg11(xe, ye, m.ikle, u, v, 1.0)
    C = bief.Context
    cases = [
        lambda c, g: c.matrix_mt01aa(0.5, g[2]),
        lambda c, g: c.matrix_mt02aa(1, 1.0, *g, u),
        lambda c, g: c.matrix_mt02aa(2, 0.7, *g, u, v),
        lambda c, g: c.matrix_supg(3, 0.7, *g, u, v, f=kx, g=ky),
        lambda c, g: c.matrix_supg(3, 0.7, *g, u, v, flags=K),
        lambda c, g: c.matrix_supg(4, 0.3, *g, u, v),
        lambda c, g: c.matrix_supg(5, 1.3, *g, u, v),
        lambda c, g: c.matrix_supg(5, 1.0, *g, u, v, flags=C.SUPG_UPWIND),
        lambda c, g: c.matrix_supg(5, 2.0, *g, kx, ky, flags=C.SUPG_U_P0),
        lambda c, g: c.matrix_supg(5, 1.0, *g, u, v, flags=K, op="M=M+TN"),
    ]
    for i, build in enumerate(cases):
        build(a, (xe, ye, sf))
        build(b, (None, None, None))
        _, da_a, xa_a = a.get_matrix()
        _, da_b, xa_b = b.get_matrix()
        assert np.array_equal(da_a, da_b) and np.array_equal(xa_a, xa_b), i
    with pytest.raises(bief.BiefError):
        a.matrix_mt01aa(0.5, None)  # no geometry kept on this context


def test_supg_rejects_bad_arguments(gpu, small):
    m, (xe, ye), sf, u, v = small
    ctx = bief.Context(m)
    with pytest.raises(bief.BiefError):
        ctx.matrix_supg(3, 1.0, xe, ye, sf, u, v)               # MT03AA without F, G
    with pytest.raises(bief.BiefError):
        ctx.matrix_supg(4, 1.0, xe, ye, sf, u, v, flags=K)      # KSUPG with MT04AA
    with pytest.raises(bief.BiefError):
        ctx.matrix_supg(5, 1.0, xe, ye, sf, u, v, op="M=M+N")   # nothing to add to
    with pytest.raises(bief.BiefError):
        ctx.matrix_om("M=M+CN", 1.0)                            # no matrix, no saved copy
