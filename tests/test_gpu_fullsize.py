"""The 2,003,181-node estuary wave-equation system at its full size (the
unstructured 2M-node workload): properties that hold at any size, checked
where the oracle would take minutes.

  * the EBE product is linear and, for the symmetric matrix, its own
    transpose (<A y, z> = <y, A z>; X=TAY equal to X=AY);
  * FP64 CG + PRECON 2 to 1e-8 takes 41 iterations, the count of the
    reference's own SOLVE (flang build of sources/utils/bief) on this
    system, and its true residual ||b - A x|| / ||b|| is below 1e-8;
  * the mixed path (FP32 inner CG in FP64 refinement) reaches the same
    stop rule and agrees with the FP64 solution to 1e-7;
  * CVDFTR's SUPG tracer system assembled on the device equals the
    oracle's (the C restatement, here at full size) to 1e-13, and GMRES(5)
    solves it in the reference's 3 restarts;
  * on the 5M-node prism mesh, the diffusion and the pressure Poisson
    systems in the reference's CG counts.
Building the mesh and the system takes about a minute on the host."""
import json
import os

import numpy as np
import pytest

from opentelemac_amd import bief, mesh as M, system as S, tracer
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# the reference SOLVE's counts on these systems, written by
# tests/golden/make_fullsize_counts.py from the flang build of sources/utils/bief
with open(os.path.join(os.path.dirname(__file__), "golden", "ref_fullsize_counts.json")) as _fh:
    REF = json.load(_fh)


@pytest.fixture(scope="module")
def estuary():
    m = M.estuary_mesh(2_000_000, seed=0)
    depth = M.smooth_random_depth(m, seed=0)
    bnd = S.boundary_nodes(m)
    sea = bnd[m.x[bnd] > m.x.max() - 1.0]
    return m, S.wave_system(m, depth, dt=2.0, theta=1.0, inflow=sea)


def test_fullsize_matvec_linear_and_symmetric(gpu, estuary):
    m, sysd = estuary
    assert m.npoin == 2_003_181 and m.nelem == 3_999_606
    ctx = bief.Context(m)
    ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
    rng = np.random.default_rng(11)
    y, z = rng.normal(size=m.npoin), rng.normal(size=m.npoin)
    ay, az = ctx.matvec("X=AY", y), ctx.matvec("X=AY", z)
    a_yz = ctx.matvec("X=AY", y + 2.0 * z)
    assert np.linalg.norm(a_yz - (ay + 2.0 * az)) <= 1e-12 * np.linalg.norm(a_yz)
    lhs, rhs = float(ay @ z), float(y @ az)
    assert abs(lhs - rhs) <= 1e-12 * np.abs(ay).sum() * np.abs(z).max()
    aty = ctx.matvec("X=TAY", y)
    assert np.linalg.norm(aty - ay) <= 1e-14 * np.linalg.norm(ay)


def test_fullsize_solves(gpu, estuary):
    m, sysd = estuary
    ctx = bief.Context(m)
    b = sysd["b"]
    sols = {}
    for prec in (bief.PREC_FP64, bief.PREC_MIXED):
        ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
        x, info = ctx.solve(sysd["x0"].copy(), b.copy(), slv=1, precon=2, eps=1e-8,
                            nitmax=5000, precision=prec)
        assert info["converged"], (prec, info)
        assert info["relres"] <= 1e-8
        sols[prec] = x
    ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
    for prec, x in sols.items():
        true = np.linalg.norm(b - ctx.matvec("X=AY", x)) / np.linalg.norm(b)
        assert true <= 1e-8, (prec, true)
    ref = sols[bief.PREC_FP64]
    assert np.linalg.norm(sols[bief.PREC_MIXED] - ref) <= 1e-7 * np.linalg.norm(ref)


def test_fullsize_fp64_iterations_match_reference(gpu, estuary):
    m, sysd = estuary
    ctx = bief.Context(m)
    ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
    _, info = ctx.solve(sysd["x0"].copy(), sysd["b"].copy(), slv=1, precon=2, eps=1e-8,
                        nitmax=5000)
    assert info["niter"] == REF["estuary_2m_cg"]["niter"], info
    assert abs(info["relres"] - REF["estuary_2m_cg"]["relres"]) <= 1e-6 * info["relres"], info


@pytest.fixture(scope="module")
def prisms():
    """The 5,014,910-node sigma mesh: the 500k estuary base extruded to 10
    planes over its seeded depth (8,996,184 prisms), with its geometry."""
    base = M.estuary_mesh(500_000, seed=0)
    depth = M.smooth_random_depth(base, seed=0)
    p = M.extrude_prisms(base, 10, -depth, np.zeros(base.npoin))
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    return base, p, np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb]), \
        base.surfac()[eb]


def test_fullsize_prisms_assembly_and_solve(gpu, prisms):
    """MT01PP integrates the mesh volume (all entries of the mass matrix sum
    to sum S (h1 + h2 + h3) / 3); MT02PP with unit coefficients annihilates
    constants; the diffusion system (MT01PP / dt + MT02PP) solves to the
    1e-8 stop rule (on the diagonally scaled system) in FP64, within one
    iteration of the reference's SOLVE on this system (flang build of
    sources/utils/bief, the reference's own MT01PP/MT02PP: 54 iterations,
    golden/ref_fullsize_counts.json), and in mixed precision, with true
    residuals ||b - A x|| / ||b|| below 1e-7."""
    base, p, xel, yel, s = prisms
    assert p.npoin == 5_014_910 and p.nelem == 8_996_184
    ctx = bief.Context(p, block_nodes=2560)
    ctx.set_geometry(xel, yel, s)
    ctx.matrix_mt01pp(1.0, None, p.z)
    _, da, xa = ctx.get_matrix()
    ik = p.ikle
    h = (p.z[ik[3]] - p.z[ik[0]]) + (p.z[ik[4]] - p.z[ik[1]]) + (p.z[ik[5]] - p.z[ik[2]])
    vol = float(np.sum(s * h / 3.0))
    assert abs(da.sum() + 2.0 * xa.sum() - vol) <= 1e-12 * vol
    one = np.ones(p.npoin)
    ctx.matrix_mt02pp(1.0, None, None, None, p.z, one, one, one, base.npoin)
    _, da2, _ = ctx.get_matrix()
    assert np.abs(ctx.matvec("X=AY", one)).max() <= 1e-12 * np.abs(da2).max()
    sig = np.repeat(np.linspace(0.0, 1.0, 10), base.npoin)
    nu = 10.0 ** (-4.0 + 2.0 * sig)
    b = np.random.default_rng(12).normal(size=p.npoin)
    sols = {}
    for prec in (bief.PREC_FP64, bief.PREC_MIXED):
        ctx.matrix_mt01pp(0.5, None, p.z)
        ctx.matrix_mt02pp(1.0, None, None, None, p.z, nu, nu, nu, base.npoin, accumulate=True)
        if prec == bief.PREC_FP64:
            ctx.save_matrix()
        x, info = ctx.solve(np.zeros(p.npoin), b.copy(), slv=1, precon=2, eps=1e-8, nitmax=5000,
                            precision=prec)
        assert info["converged"], (prec, info)
        if prec == bief.PREC_FP64:
            assert abs(info["niter"] - REF["prisms_5m_diffusion"]["niter"]) <= 1, info
        sols[prec] = x
    ctx.restore_matrix()
    for prec, x in sols.items():
        true = np.linalg.norm(b - ctx.matvec("X=AY", x)) / np.linalg.norm(b)
        assert true <= 1e-7, (prec, true)


def test_fullsize_prisms_pressure_poisson(gpu, prisms):
    """The non-hydrostatic pressure system on the 5M-node prism mesh
    (prediv.F:210: MT02PP with unit coefficients; :384: DIRICH with zero on
    the surface plane), assembled and prescribed on the device from device
    fields, solved by CG + PRECON 2 to 1e-8 from zero: FP64 within one
    iteration of the reference's SOLVE on the reference's own assembly (103,
    golden/ref_fullsize_counts.json) and at its stop-test precision to 1e-3;
    mixed to the same stop rule; true residuals below 1e-7; the surface
    nodes stay at their prescribed zero."""
    import torch

    base, p, xel, yel, s = prisms
    ctx = bief.Context(p, block_nodes=2560)
    ctx.set_geometry(xel, yel, s)
    dev = torch.device("cuda:0")
    zd = torch.from_numpy(np.ascontiguousarray(p.z)).to(dev)
    one = torch.ones(p.npoin, dtype=torch.float64, device=dev)
    limdir = torch.full((p.npoin,), 2, dtype=torch.int32, device=dev)
    limdir[(p.nplan - 1) * base.npoin:] = 5
    sm = torch.from_numpy(np.random.default_rng(13).normal(size=p.npoin)).to(dev)
    ref = REF["prisms_5m_poisson"]
    for prec in (bief.PREC_FP64, bief.PREC_MIXED):
        ctx.matrix_mt02pp(1.0, None, None, None, zd, one, one, one, base.npoin)
        f, b = ctx.dirich(torch.zeros_like(zd), sm.clone(), torch.zeros_like(zd), limdir)
        ctx.save_matrix()
        b0 = b.clone()
        x, info = ctx.solve(f, b, slv=1, precon=2, eps=1e-8, nitmax=5000, precision=prec)
        assert info["converged"], (prec, info)
        if prec == bief.PREC_FP64:
            assert abs(info["niter"] - ref["niter"]) <= 1, info
            assert abs(info["relres"] - ref["relres"]) <= 1e-3 * ref["relres"], info
        ctx.restore_matrix()
        r = b0 - ctx.matvec("X=AY", x)
        true = float(torch.linalg.norm(r) / torch.linalg.norm(b0))
        assert true <= 1e-7, (prec, true)
        assert bool((x[(p.nplan - 1) * base.npoin:] == 0).all())


def test_fullsize_supg_system_and_gmres(gpu, estuary):
    """The SUPG tracer system of the 2M-node estuary (the synthetic flow of
    tracer.estuary_flow, nu = 0.1, dt = 10 s): the device assembly against
    the oracle's at full size (which equals the reference's own routines bit
    for bit on this system), DIRICH on the sea boundary, GMRES(5) + PRECON 2
    to 1e-8 in the reference's count (its SOLVE on the reference's own
    assembly: 3 restarts, relative precision 0.2229581E-08,
    golden/ref_fullsize_counts.json)."""
    m, _ = estuary
    u, v, fn, sea = tracer.estuary_flow(m, seed=0)
    nu = np.full(m.npoin, 0.1)
    ctx = bief.Context(m)
    sm = tracer.supg_system(ctx, m, u, v, nu, 10.0, fn)
    _, da, xa = ctx.get_matrix()
    A, smo = O.supg_system(m, u, v, nu, 10.0, fn)
    for got, ref in ((da, A.da), (xa, A.xa), (sm, smo)):
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    limdir = np.full(m.npoin, 2, dtype=np.int32)
    limdir[sea] = 5
    f, sm2 = ctx.dirich(fn, sm, np.zeros(m.npoin), limdir)
    ctx.save_matrix()
    x, info = ctx.solve(f.copy(), sm2.copy(), slv=7, precon=2, krylov=5, eps=1e-8, nitmax=2000)
    ref = REF["supg_2m_gmres"]
    assert info["converged"] and abs(info["niter"] - ref["restarts"]) <= 1, info
    assert abs(info["relres"] - ref["relres"]) <= 1e-3 * ref["relres"], info
    # mixed precision: FP32 Arnoldi inside FP64 restarts, the stop rule on
    # the FP64 residual; restarts within one of the FP64 count
    ctx.restore_matrix()
    xm, im = ctx.solve(f.copy(), sm2.copy(), slv=7, precon=2, krylov=5, eps=1e-8, nitmax=2000,
                       precision=bief.PREC_MIXED)
    assert im["converged"] and im["relres"] < 1e-8 and abs(im["niter"] - info["niter"]) <= 1, im
    assert np.linalg.norm(xm - x) <= 1e-6 * np.linalg.norm(x)


def test_fullsize_adjoint_gradient(gpu):
    """The gradient of the gauge misfit through the device wave model at
    the adjoint workload's size (the 500k-node estuary, 50 steps of 2 s,
    200 gauges, checkpoints every 10 steps, mixed solves to 1e-10):
      * with checkpoints it equals, to the bit, the gradient that keeps
        every state (each recomputed segment repeats the forward sweep's
        solves from the same initial guesses);
      * along a random direction it matches central differences of the
        cost to 1e-5 of sum |g_i d_i| (Strickler and bathymetry; steps of
        1e-5 and 1e-6 of the parameter's largest value)."""
    import torch
    from opentelemac_amd import wave as W

    m = M.estuary_mesh(500_000, seed=0)
    depth = M.smooth_random_depth(m, seed=0)
    bnd = S.boundary_nodes(m)
    sea = bnd[m.x[bnd] > m.x.max() - 1.0]
    model = W.DeviceWaveModel(m, sea, dt=2.0, amp=0.2, period=600.0,
                              precision=bief.PREC_MIXED, eps=1e-10)
    dev = model.dev
    rng = np.random.default_rng(2)
    k = torch.tensor(rng.uniform(20, 60, m.npoin), dtype=torch.float64, device=dev)
    zb = torch.tensor(-depth, dtype=torch.float64, device=dev)
    gauges = torch.tensor(rng.choice(m.npoin, 200, replace=False), device=dev)
    obs = model.run(k * torch.tensor(rng.uniform(0.8, 1.2, m.npoin), device=dev), zb, 50, gauges)
    _, gk, gz = model.cost_and_gradient(k, zb, 50, gauges, obs, 10)
    _, gk_all, gz_all = model.cost_and_gradient(k, zb, 50, gauges, obs, 50)
    for g, g_all in ((gk, gk_all), (gz, gz_all)):
        rel = float(torch.linalg.norm(g - g_all) / torch.linalg.norm(g_all))
        assert torch.equal(g, g_all), rel
    for p, g, scale in ((k, gk, 1e-5), (zb, gz, 1e-6)):
        d = torch.tensor(rng.normal(size=m.npoin), dtype=torch.float64, device=dev)
        h = scale * float(p.abs().max())
        p0 = p.clone()
        p.copy_(p0 + h * d)
        jp = float(model.run(k, zb, 50, gauges, obs))
        p.copy_(p0 - h * d)
        jm = float(model.run(k, zb, 50, gauges, obs))
        p.copy_(p0)
        fd = (jp - jm) / (2 * h)
        assert abs(fd - float(g @ d)) <= 1e-5 * float((g * d).abs().sum()), (fd, float(g @ d))


@pytest.fixture(scope="module")
def ensemble256():
    sizes = S.ensemble_sizes(256, seed=0)
    return S.estuary_ensemble(sizes, seed=0)


def _scaled_residuals(ens, systems, x):
    """Per member: the residual GRACJG's stop rule tests, recomputed in FP64
    from the solution (D (b - A x) over D b, D = diag(A)^-1/2, with the
    ensemble's FP64 product of the original matrix), and the unscaled
    ||b - A x|| / ||b||."""
    b = np.concatenate([s["b"] for s in systems])
    r = b - ens.matvec("X=AY", x)
    d = 1.0 / np.sqrt(np.concatenate([s["da"] for s in systems]))
    out = []
    for rs, bs, ds in zip(ens.split(r), ens.split(b), ens.split(d)):
        out.append((np.linalg.norm(ds * rs) / max(np.linalg.norm(ds * bs), 1.0),
                    np.linalg.norm(rs) / np.linalg.norm(bs)))
    return np.array(out)


def test_fullsize_ensemble_256_estuaries(gpu, ensemble256):
    """The batched-ensemble workload at its size: 256 unstructured estuary
    meshes of 5k-200k nodes (log-uniform, seeded; each its own outline,
    points, depth field and tide), solved as one ensemble by CG + PRECON 2 to
    1e-8, FP64 and mixed. Every member: converged, its stop-rule residual
    recomputed in FP64 from the solution at most 1e-8 (mixed; FP64 CG tests
    its recursive residual, so 1.05e-8 there), and the unscaled relative
    residual at most 2e-8 (the diagonal scaling moves it by up to 1.3x on
    these members). Eight members (the smallest, the largest and six drawn
    at random) against the oracle's solve of that system alone: FP64 within
    1e-10 and +-1 iteration, mixed within 1e-6."""
    meshes, systems = ensemble256
    sizes = np.array([m.npoin for m in meshes])
    assert len(meshes) == 256 and sizes.min() < 5500 and sizes.max() > 190_000
    assert all(m.meta.get("kind") == "estuary" for m in meshes)
    ens = bief.Ensemble(meshes)
    da = np.concatenate([s["da"] for s in systems])
    xa = np.ascontiguousarray(np.concatenate([s["xa"] for s in systems], axis=1))
    x0 = np.concatenate([s["x0"] for s in systems])
    b = np.concatenate([s["b"] for s in systems])
    rng = np.random.default_rng(7)
    pick = [int(np.argmin(sizes)), int(np.argmax(sizes))]
    pick += [int(k) for k in rng.choice(np.setdiff1d(np.arange(256), pick), 6, replace=False)]
    sols = {}
    for prec in (bief.PREC_FP64, bief.PREC_MIXED):
        ens.set_matrix("Q", "S", da, xa)
        x, info = ens.solve(x0.copy(), b.copy(), slv=1, precon=2, eps=1e-8, nitmax=5000,
                            precision=prec)
        infos = ens.infos()
        lat = ens.latencies_ms()
        assert info["converged"] and all(si["converged"] for si in infos), prec
        assert all(si["relres"] <= 1e-8 for si in infos), prec
        assert np.all(lat > 0) and np.all(np.isfinite(lat))
        ens.set_matrix("Q", "S", da, xa)
        res = _scaled_residuals(ens, systems, x)
        lim = 1e-8 if prec == bief.PREC_MIXED else 1.05e-8
        assert res[:, 0].max() <= lim, (prec, res[:, 0].max(), int(np.argmax(res[:, 0])))
        assert res[:, 1].max() <= 2e-8, (prec, res[:, 1].max())
        sols[prec] = (ens.split(x), infos)
    for k in pick:
        m, sd = meshes[k], systems[k]
        A = O.Matrix(sd["da"], sd["xa"], "Q", "S")
        xr, ir, _, _ = O.solve(A, m.ikle, sd["b"], sd["x0"], slv=1, precon=2, eps=1e-8,
                               nitmax=5000)
        xs, infos = sols[bief.PREC_FP64]
        assert abs(infos[k]["niter"] - ir.niter) <= 1, (k, infos[k], ir.niter)
        assert np.linalg.norm(xs[k] - xr) <= 1e-10 * np.linalg.norm(xr), k
        xm, _ = sols[bief.PREC_MIXED]
        assert np.linalg.norm(xm[k] - xr) <= 1e-6 * np.linalg.norm(xr), k
