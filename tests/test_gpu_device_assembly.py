"""Operand preparation from device-resident fields: the element matrices
(MT01AA, MT02AA, MT03AA/MT04AA/MT05AA with KSUPG, MT01PP, MT02PP), DIRICH
and LUMP called with float64 torch CUDA tensors (bhip_*_dev: no host array,
queued on the context stream) give the same bits as the same calls from host
arrays, whose parity with the oracle the element-matrix and preparation
tests establish; CVDFTR's SUPG tracer system assembled from device fields
gives the host-array system's right-hand side and matrix."""
import numpy as np
import pytest
import torch

from opentelemac_amd import bief, mesh as M, tracer
from oracle import oracle as O

pytestmark = pytest.mark.gpu
K = bief.Context.SUPG_KSUPG
P0 = bief.Context.SUPG_U_P0
UPW = bief.Context.SUPG_UPWIND


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def _same(a, b):
    ta, da_a, xa_a = a.get_matrix()
    tb, da_b, xa_b = b.get_matrix()
    return ta == tb and np.array_equal(da_a, da_b) and np.array_equal(xa_a, xa_b)


@pytest.fixture(scope="module")
def tri():
    m = M.random_small_mesh(15, 11, seed=93, jitter=0.25)
    rng = np.random.default_rng(17)
    u, v = rng.uniform(-2, 2, m.npoin), rng.uniform(-2, 2, m.npoin)
    xe, ye = m.xel_yel()
    return m, xe, ye, m.surfac(), u, v, rng


@pytest.mark.parametrize("bn", [0, 40])
def test_triangle_matrices_device_equal_host(gpu, tri, bn):
    m, xe, ye, s, u, v, rng = tri
    a, b = bief.Context(m, block_nodes=bn), bief.Context(m, block_nodes=bn)
    for c in (a, b):
        c.set_geometry(xe, ye, s)
    ud, vd = _d(u), _d(v)
    # MATMAS
    a.matrix_mt01aa(0.37, s)
    b.matrix_mt01aa(0.37, None)
    assert _same(a, b)
    # MATDIF P1 isotropic 'M=M+N', then the PROPAG form MATDIFUV
    a.matrix_mt02aa(1, 0.7, xe, ye, s, u * u)
    b.matrix_mt02aa(1, 0.7, None, None, None, ud * ud)
    assert _same(a, b)
    a.matrix_mt01aa(0.5, s)
    a.matrix_mt02aa(2, -1.0, xe, ye, s, u, v, accumulate=True)
    b.matrix_mt01aa(0.5, None)
    b.matrix_mt02aa(2, -1.0, None, None, None, ud, vd, accumulate=True)
    assert _same(a, b)
    # the SUPG family: nodal and P0 velocities, given or computed KSUPG
    kx, ky = O.kspg11(xe, ye, m.ikle, u, v, 1.0)
    cases = [
        dict(kind=3, xmul=0.7, f=kx, g=ky, u=u, v=v),
        dict(kind=3, xmul=0.7, flags=K, u=u, v=v),
        dict(kind=4, xmul=0.3, u=u, v=v),
        dict(kind=5, xmul=1.3, u=u, v=v),
        dict(kind=5, xmul=1.0, flags=UPW, u=u, v=v),
        dict(kind=5, xmul=2.0, flags=K, u=u, v=v),
        dict(kind=5, xmul=2.0, flags=P0, u=kx, v=ky),
    ]
    for kw in cases:
        kind, xmul, flags = kw["kind"], kw["xmul"], kw.get("flags", 0)
        f, g = kw.get("f"), kw.get("g")
        for op in ("M=N", "M=M+N", "M=M+TN"):
            if op != "M=N":
                a.matrix_mt01aa(0.5, s)
                b.matrix_mt01aa(0.5, None)
            a.matrix_supg(kind, xmul, xe, ye, s, kw["u"], kw["v"], f=f, g=g, flags=flags, op=op)
            b.matrix_supg(kind, xmul, None, None, None, _d(kw["u"]), _d(kw["v"]),
                          f=None if f is None else _d(f), g=None if g is None else _d(g),
                          flags=flags, op=op)
            assert a.typext == b.typext
            assert _same(a, b), (kind, flags, op)


def test_dirich_lump_device_equal_host(gpu, tri):
    m, xe, ye, s, u, v, rng = tri
    a, b = bief.Context(m), bief.Context(m)
    for c in (a, b):
        c.set_geometry(xe, ye, s)
        c.matrix_mt01aa(0.5, None)
        c.matrix_supg(5, 1.0, None, None, None, _d(u), _d(v), op="M=M+N")
    la = a.lump(0.3)
    lb = b.lump(0.3, out=torch.empty(m.npoin, dtype=torch.float64, device="cuda:0"))
    assert np.array_equal(la, lb.cpu().numpy())
    f0, sm = rng.normal(size=m.npoin), rng.normal(size=m.npoin)
    limdir = np.where(rng.uniform(size=m.npoin) < 0.2, 5, 2).astype(np.int32)
    fbor = np.where(limdir == 5, rng.normal(size=m.npoin), 0.0)
    fa, sma = a.dirich(f0, sm, fbor, limdir)
    fd, smd = _d(f0), _d(sm)
    fb, smb = b.dirich(fd, smd, _d(fbor), torch.from_numpy(limdir).to("cuda:0"))
    assert fb is fd and smb is smd  # in place
    assert np.array_equal(fa, fb.cpu().numpy()) and np.array_equal(sma, smb.cpu().numpy())
    assert _same(a, b)


def test_prism_matrices_device_equal_host(gpu):
    base = M.random_small_mesh(8, 6, seed=6)
    rng = np.random.default_rng(56)
    bot = -rng.uniform(1, 5, base.npoin)
    top = rng.uniform(-0.2, 0.2, base.npoin)
    top[:3] = bot[:3] + 5e-5
    p = M.extrude_prisms(base, 5, bot, top)
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    xel, yel, s = (np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb]),
                   base.surfac()[eb])
    f, g = rng.uniform(0.5, 2, p.npoin), rng.uniform(0.5, 2, p.npoin)
    h = 10 ** rng.uniform(-4, -2, p.npoin)
    a, b = bief.Context(p, block_nodes=40), bief.Context(p, block_nodes=40)
    for c in (a, b):
        c.set_geometry(xel, yel, s)
    zd, fd, gd, hd = (_d(t) for t in (p.z, f, g, h))
    a.matrix_mt01pp(0.5, None, p.z)
    b.matrix_mt01pp(0.5, None, zd)
    assert _same(a, b)
    for hz in (0, 1):
        for p0 in (0, 1):
            for inch in (0, 1):
                for mon in (0, 1):
                    npou0 = 0 if p0 else base.npoin
                    a.matrix_mt02pp(-0.7, None, None, None, p.z, f, g, h, npou0, bool(hz),
                                    bool(inch), bool(mon), accumulate=True)
                    b.matrix_mt02pp(-0.7, None, None, None, zd, fd, gd, hd, npou0, bool(hz),
                                    bool(inch), bool(mon), accumulate=True)
                    assert _same(a, b), (hz, p0, inch, mon)


def test_supg_system_from_device_fields(gpu):
    """The whole CVDFTR assembly (five element matrices, OM, two mat-vecs)
    from device tensors: the same matrix and right-hand side bits as from
    host arrays, and the solve of the device system converges to the host
    system's solution."""
    m = M.estuary_mesh(6000, seed=3)
    u, v, fn, sea = tracer.estuary_flow(m, seed=3)
    nu = np.full(m.npoin, 0.1)
    a, b = bief.Context(m), bief.Context(m)
    sm_a = tracer.supg_system(a, m, u, v, nu, 10.0, fn, tetat=0.6, sm0=0.01 * fn)
    fnd = _d(fn)
    sm_b = tracer.supg_system(b, m, _d(u), _d(v), _d(nu), 10.0, fnd, tetat=0.6,
                              sm0=0.01 * fnd)
    assert isinstance(sm_b, torch.Tensor) and sm_b.is_cuda
    assert np.array_equal(sm_a, sm_b.cpu().numpy())
    assert _same(a, b)


def test_device_fields_refused_with_caller_geometry_or_wrong_size(gpu, tri):
    m, xe, ye, s, u, v, _ = tri
    c = bief.Context(m)
    c.set_geometry(xe, ye, s)
    with pytest.raises(bief.BiefError):
        c.matrix_mt02aa(1, 1.0, xe, ye, s, _d(u))       # device fields, host geometry
    with pytest.raises(bief.BiefError):
        c.matrix_supg(5, 1.0, None, None, None, _d(u[:-1]), _d(v[:-1]))
    with pytest.raises(bief.BiefError):
        c.matrix_supg(5, 1.0, None, None, None, _d(u), _d(v), flags=P0)  # P0 needs nelem
    with pytest.raises(bief.BiefError):
        c.dirich(_d(u), _d(v), _d(u), torch.zeros(m.npoin, dtype=torch.int64, device="cuda:0"))
    with pytest.raises(bief.BiefError):
        c.lump(1.0, out=torch.empty(m.npoin - 1, dtype=torch.float64, device="cuda:0"))
    # the device entries need the geometry kept on the context
    d = bief.Context(m)
    with pytest.raises(bief.BiefError):
        d.matrix_mt02aa(1, 1.0, None, None, None, _d(u))
