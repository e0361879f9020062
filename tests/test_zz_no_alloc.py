"""The allocation-free SOLVE entry point (bhip_solve_dev_noalloc), called
through the C ABI: its result equals bhip_solve_dev's to the bit on two calls
in a row, the second on a smaller system (the shared static scratch keeps the
first call's contents), and sizes beyond its scratch are refused."""
import ctypes

import numpy as np
import pytest

from opentelemac_amd import bief, mesh as M, system as S

pytestmark = pytest.mark.gpu


def _cfg():
    return bief.SolverCfg(1, 2, 5, 1000, 1e-10, bief.PREC_FP64, 1e-4, 0)


def _noalloc(ctx, x, b):
    info = bief.SolveInfo()
    rc = bief.lib().bhip_solve_dev_noalloc(ctx._h, ctypes.byref(_cfg()), x.data_ptr(),
                                           b.data_ptr(), ctypes.byref(info))
    return rc, info


def test_noalloc_solve_equals_solve_dev(gpu):
    import torch
    for nx, ny, seed in ((60, 40, 1), (14, 9, 2)):   # the second call is smaller
        m = M.random_small_mesh(nx, ny, seed=seed)
        sysd = S.wave_system(m, 2.0 + m.x / 10.0, dt=5.0)
        ctx = bief.Context(m, block_nodes=64)
        ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
        ctx.save_matrix()
        x_ref = torch.from_numpy(sysd["x0"].copy()).to(gpu)
        b_ref = torch.from_numpy(sysd["b"].copy()).to(gpu)
        _, info = ctx.solve(x_ref, b_ref, slv=1, precon=2, eps=1e-10, nitmax=1000)
        ctx.restore_matrix()
        x = torch.from_numpy(sysd["x0"].copy()).to(gpu)
        b = torch.from_numpy(sysd["b"].copy()).to(gpu)
        ctx._sync_torch()
        rc, ni = _noalloc(ctx, x, b)
        torch.cuda.synchronize()
        assert rc == 0 and ni.converged
        assert ni.niter == info["niter"] and ni.relres == info["relres"]
        assert torch.equal(x, x_ref) and torch.equal(b, b_ref)
        ctx.close()


def test_noalloc_refuses_what_its_scratch_cannot_hold(gpu):
    import torch
    # 65 systems: one more than the scratch holds states for
    meshes = [M.random_small_mesh(3, 3, seed=k) for k in range(65)]
    ens = bief.Ensemble(meshes)
    systems = [S.wave_system(mm, np.full(mm.npoin, 2.0)) for mm in meshes]
    ens.set_matrix("Q", "S", np.concatenate([s["da"] for s in systems]),
                   np.concatenate([s["xa"] for s in systems], axis=1))
    x = torch.zeros(ens.npoin, dtype=torch.float64, device=gpu)
    b = torch.ones(ens.npoin, dtype=torch.float64, device=gpu)
    rc, _ = _noalloc(ens, x, b)
    assert rc == 2
    # mixed precision is not offered
    m = M.random_small_mesh(5, 4, seed=3)
    sysd = S.wave_system(m, np.full(m.npoin, 2.0))
    ctx = bief.Context(m)
    ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
    cfg = bief.SolverCfg(1, 2, 5, 100, 1e-8, bief.PREC_MIXED, 1e-4, 0)
    xx = torch.zeros(m.npoin, dtype=torch.float64, device=gpu)
    bb = torch.ones(m.npoin, dtype=torch.float64, device=gpu)
    rc = bief.lib().bhip_solve_dev_noalloc(ctx._h, ctypes.byref(cfg), xx.data_ptr(),
                                           bb.data_ptr(), ctypes.byref(bief.SolveInfo()))
    assert rc == 2
