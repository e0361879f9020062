"""The wave-equation step on its HIP kernels (bhip_wave_*) and their
hand-written adjoint, against the torch model of the same equations
(opentelemac_amd/wave.py WaveModel, whose gradient is itself checked against
central differences and a tangent-linear run): same gauge series, same
gradient, and central differences through the device model alone."""
import numpy as np
import pytest
import torch

from opentelemac_amd import bief, mesh as M, wave as W

pytestmark = pytest.mark.gpu

NSTEPS = 20


@pytest.fixture(scope="module")
def setup(gpu):
    base = M.random_small_mesh(22, 11, seed=71, jitter=0.2)
    m = M.Mesh(x=base.x * 100.0, y=base.y * 100.0, ikle=base.ikle, npoin2=base.npoin,
               nelem2=base.nelem, meta={})
    sea = np.nonzero(m.x <= 1e-9)[0]
    kw = dict(dt=2.0, amp=0.5, period=120.0, eps=1e-13)
    ref = W.WaveModel(m, sea, **kw)
    dev = W.DeviceWaveModel(m, sea, **kw)
    rng = np.random.default_rng(5)
    d = ref.dev
    k = torch.tensor(rng.uniform(20, 60, m.npoin), dtype=torch.float64, device=d)
    zb = torch.tensor(-(1.0 + 2.0 * m.x / m.x.max()) - rng.uniform(0, 0.3, m.npoin),
                      dtype=torch.float64, device=d)
    gauges = torch.tensor(rng.choice(m.npoin, 15, replace=False), device=d)
    with torch.no_grad():
        obs = ref.run(k * 0.7, zb, NSTEPS, gauges)
    return ref, dev, k, zb, gauges, obs, rng


def _rel(a, b):
    return float((a - b).abs().max()) / max(float(b.abs().max()), 1e-300)


def test_forward_matches_torch_model(setup):
    ref, dev, k, zb, gauges, obs, _ = setup
    with torch.no_grad():
        s_ref = ref.run(k, zb, NSTEPS, gauges)
    s_dev = dev.run(k, zb, NSTEPS, gauges)
    assert s_dev.shape == s_ref.shape
    assert _rel(s_dev, s_ref) <= 1e-10
    # the final state, caller order
    with torch.no_grad():
        e_ref = ref.run(k, zb, 7)
    e_dev = dev.run(k, zb, 7)
    assert _rel(e_dev, e_ref) <= 1e-10


@pytest.mark.parametrize("checkpoint", [NSTEPS, 6])
def test_gradient_matches_torch_adjoint(setup, checkpoint):
    ref, dev, k, zb, gauges, obs, _ = setup
    j_ref, gk_ref, gz_ref = W.cost_and_gradient(ref, k, zb, NSTEPS, gauges, obs, checkpoint=5)
    j, gk, gz = dev.cost_and_gradient(k, zb, NSTEPS, gauges, obs, checkpoint=checkpoint)
    assert abs(float(j) - float(j_ref)) <= 1e-9 * float(j_ref)
    assert _rel(gk, gk_ref) <= 1e-7, _rel(gk, gk_ref)
    assert _rel(gz, gz_ref) <= 1e-7, _rel(gz, gz_ref)


def test_gradient_matches_central_differences(setup):
    _, dev, k, zb, gauges, obs, rng = setup
    _, gk, gz = dev.cost_and_gradient(k, zb, NSTEPS, gauges, obs, checkpoint=6)
    for name, p, g in (("K", k, gk), ("zb", zb, gz)):
        scale = float(g.abs().max())
        for i in rng.choice(dev.n, 5, replace=False):
            h = 1e-4 * max(1.0, abs(float(p[i])))
            pp, pm = p.clone(), p.clone()
            pp[i] += h
            pm[i] -= h
            args_p = (pp, zb) if name == "K" else (k, pp)
            args_m = (pm, zb) if name == "K" else (k, pm)
            fd = (float(dev.run(*args_p, NSTEPS, gauges, obs)) -
                  float(dev.run(*args_m, NSTEPS, gauges, obs))) / (2 * h)
            assert abs(float(g[i]) - fd) <= 1e-5 * scale, (name, i, float(g[i]), fd, scale)


def test_autograd_function(setup):
    _, dev, k, zb, gauges, obs, _ = setup
    kk = k.clone().requires_grad_(True)
    zz = zb.clone().requires_grad_(True)
    j = W.DeviceWaveCost.apply(dev, kk, zz, NSTEPS, gauges, obs, 6)
    (3.0 * j).backward()
    j2, gk, gz = dev.cost_and_gradient(k, zb, NSTEPS, gauges, obs, checkpoint=6)
    assert abs(float(j) - float(j2)) <= 1e-12 * float(j2)
    assert _rel(kk.grad, 3.0 * gk) <= 1e-10 and _rel(zz.grad, 3.0 * gz) <= 1e-10


def test_rejects_bad_vectors(setup):
    _, dev, k, zb, *_ = setup
    p = dev._params(2.0)
    bad = torch.zeros(dev.n - 1, dtype=torch.float64, device=dev.dev)
    z = dev._zeros()
    with pytest.raises(bief.BiefError):
        dev.kern.assemble(p, z, z, z, bad, z, z)


def test_shallow_and_dry_nodes(gpu):
    """Nodes at 5-25 mm (the shallow-water friction speed of fricti.f) and
    nodes above the water (negative depth: the depth floors of
    friction_unif.f and fricti.f, HPROP = 0): the run stays finite, the device
    step equals the torch model and so does its gradient (the floors pass no
    derivative below them)."""
    base = M.random_small_mesh(14, 8, seed=73, jitter=0.2)
    m = M.Mesh(x=base.x * 50.0, y=base.y * 50.0, ikle=base.ikle, npoin2=base.npoin,
               nelem2=base.nelem, meta={})
    sea = np.nonzero(m.x <= 1e-9)[0]
    kw = dict(dt=1.0, amp=0.05, period=60.0, eps=1e-13)
    ref = W.WaveModel(m, sea, **kw)
    dev = W.DeviceWaveModel(m, sea, **kw)
    rng = np.random.default_rng(9)
    d = ref.dev
    zb = -(0.5 + m.x / m.x.max())
    inner = np.setdiff1d(np.arange(m.npoin), sea)
    pick = rng.choice(inner, 12, replace=False)
    zb[pick[:8]] = -rng.uniform(0.005, 0.025, 8)     # shallow
    zb[pick[8:]] = rng.uniform(1e-5, 1e-3, 4)         # dry
    zb = torch.tensor(zb, dtype=torch.float64, device=d)
    k = torch.tensor(rng.uniform(20, 60, m.npoin), dtype=torch.float64, device=d)
    gauges = torch.tensor(rng.choice(m.npoin, 10, replace=False), device=d)
    n = 8
    with torch.no_grad():
        obs = ref.run(k * 0.8, zb, n, gauges)
        s_ref = ref.run(k, zb, n, gauges)
    s_dev = dev.run(k, zb, n, gauges)
    assert torch.isfinite(s_dev).all()
    assert _rel(s_dev, s_ref) <= 1e-10
    j_ref, gk_ref, gz_ref = W.cost_and_gradient(ref, k, zb, n, gauges, obs, checkpoint=4)
    j, gk, gz = dev.cost_and_gradient(k, zb, n, gauges, obs, checkpoint=4)
    assert torch.isfinite(gk).all() and torch.isfinite(gz).all()
    assert abs(float(j) - float(j_ref)) <= 1e-9 * float(j_ref)
    assert _rel(gk, gk_ref) <= 1e-7, _rel(gk, gk_ref)
    assert _rel(gz, gz_ref) <= 1e-7, _rel(gz, gz_ref)
