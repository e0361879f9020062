"""The device wave-equation step (bhip_wave_assemble, SOLVE, bhip_wave_momentum)
against telemac2d PROPAG's operands computed by the reference's own routines
in PROPAG's sequence (tests/golden/ref_propag.npz, written by the driver's
ref_propag from MT01AA, MT02AA 'MATDIFUV', MT08AA, VC00AA, VC13AA, SLOP10,
ASSVEC, OV/OVD, FRICTION_CALC, FRICTI, MATVEC, DIRICH and SOLVE), on a
channel with a prescribed free surface and shallow nodes, the gouttedo basin
with dry nodes, and the estimation channel with its own boundary file:
  * AM1 (diagonal and off-diagonal terms, Dirichlet rows and columns
    cleared) and CV1 after DIRICH within 1e-12 of the reference's;
  * the increment solved by the device CG within 1e-10 of the reference's
    SOLVE, and the new velocities within 1e-10."""
import os

import numpy as np
import pytest
import torch

from opentelemac_amd import bief, mesh as M, serafin as SF

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _mesh(case):
    if case == "chan":
        return M.channel_mesh(60, 12, 400.0, 80.0)
    name = {"goutte": "geo_gouttedo.slf", "estim": "geo_estimation.slf"}[case]
    return SF.read(os.path.join(GOLD, name)).mesh()


def _close(a, b, tol):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("case", ["chan", "goutte", "estim"])
def test_device_step_matches_propag(gpu, case):
    z = np.load(os.path.join(GOLD, "ref_propag.npz"))
    g = lambda k: z[f"{case}_{k}"]  # noqa: E731
    m = _mesh(case)
    sea = (g("limdir") == SF.KDIR).astype(np.int32)
    ctx = bief.Context(m)
    xe, ye = m.xel_yel()
    ws = bief.WaveStep(ctx, xe, ye, m.surfac(), sea)
    perm = torch.as_tensor(ctx.perm().astype(np.int64), device=gpu)
    inner = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=gpu)[perm]  # noqa

    def caller(t):
        out = torch.empty_like(t)
        out[perm] = t
        return out.cpu().numpy()

    zb = g("zb")
    eta = g("hn") + zb
    dt = float(g("dt"))
    # the free surface prescribed on sea nodes: HBOR + ZF
    tide = float((g("hbor") + zb)[sea == 1][0]) if sea.any() else 0.0
    p = ws.params(dt, 9.81, 0.0, tide)
    e_, u_, v_, k_, zb_ = (inner(a).contiguous() for a in (eta, g("un"), g("vn"), g("k"), zb))
    rhs = torch.empty_like(e_)
    ws.assemble(p, e_, u_, v_, k_, zb_, rhs)
    td, da, xa = ctx.get_matrix()
    assert td == "Q"
    assert _close(da, g("da"), 1e-12), np.abs(da - g("da")).max()
    assert _close(xa, g("xa"), 1e-12), np.abs(xa - g("xa")).max()
    assert _close(caller(rhs), g("cv1"), 1e-12), np.abs(caller(rhs) - g("cv1")).max()
    # SOLVE from DIRICH's initial increment (the prescribed values), internal order
    deta = torch.where(inner(sea.astype(np.float64)) != 0, torch.full_like(e_, tide) - e_,
                       torch.zeros_like(e_)).contiguous()
    _, info = ctx.solve_internal(deta, rhs.clone(), slv=1, precon=2, eps=1e-13, nitmax=2000)
    assert info["converged"]
    assert _close(caller(deta), g("dh"), 1e-10), np.abs(caller(deta) - g("dh")).max()
    en, un, vn = torch.empty_like(e_), torch.empty_like(e_), torch.empty_like(e_)
    ws.momentum(p, e_, deta, u_, v_, en, un, vn)
    assert _close(caller(un), g("u"), 1e-10) and _close(caller(vn), g("v"), 1e-10)
    assert _close(caller(en), eta + g("dh"), 1e-10)
    ctx.close()
