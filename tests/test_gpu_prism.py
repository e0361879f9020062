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


@pytest.mark.parametrize("typext", ["S", "Q"])
def test_prism_matvec_and_cg_without_column_order(gpu, typext):
    """Prisms whose IKLE does not list the top nodes as the bottom ones one
    plane up (planes swapped) take the six-id path of the product (with the
    element stream prefetched a phase ahead, as in the column path): every
    MATVEC op and a CG solve, FP64 and mixed, against the oracle."""
    base, p = _prisms(8, 7, 5, seed=11)
    sw = M.Mesh(x=p.x, y=p.y, ikle=np.ascontiguousarray(p.ikle[[3, 4, 5, 0, 1, 2]]), z=p.z,
                nplan=p.nplan, npoin2=p.npoin2, nelem2=p.nelem2)
    rng = np.random.default_rng(9)
    nc = 30 if typext == "Q" else 15
    da = rng.uniform(1, 2, sw.npoin)
    xa = rng.normal(size=(nc, sw.nelem))
    ctx = bief.Context(sw, block_nodes=40)
    ctx.set_matrix("Q", typext, da, xa)
    A = O.Matrix(da, xa, "Q", typext)
    for op in bief.OPS:
        y = rng.normal(size=sw.npoin)
        x0 = rng.normal(size=sw.npoin)
        got = ctx.matvec(op, y, x0, c=0.3)
        ref = O.matvec(op, A, sw.ikle, y, x0, c=0.3)
        assert np.abs(got - ref).max() <= 1e-13 * (np.abs(ref).max() + 1), op
    # a symmetric definite system: the MT01PP mass matrix of the prisms, its
    # pair columns renumbered for the swapped local order (node k -> k +- 3)
    s2 = np.tile(base.surfac(), p.nplan - 1)
    t, xp = O.mt01pp(s2, p.z, p.ikle, 1.0)
    dm = O.assemble_diag(t, p.ikle, p.npoin)
    pairs = [(i, j) for i in range(6) for j in range(i + 1, 6)]
    col = {pr: c for c, pr in enumerate(pairs)}
    xm = np.empty_like(xp)
    for c, (i, j) in enumerate(pairs):
        xm[col[tuple(sorted(((i + 3) % 6, (j + 3) % 6)))]] = xp[c]
    b = rng.normal(size=sw.npoin)
    Am = O.Matrix(dm, xm, "Q", "S")
    xr, ir, _, _ = O.solve(Am, sw.ikle, b, np.zeros(sw.npoin), slv=1, precon=2, eps=1e-10,
                           nitmax=2000)
    for prec in (bief.PREC_FP64, bief.PREC_MIXED):
        ctx.set_matrix("Q", "S", dm, xm)
        x, info = ctx.solve(np.zeros(sw.npoin), b.copy(), slv=1, precon=2, eps=1e-10,
                            nitmax=2000, precision=prec)
        rel = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        assert info["converged"]
        if prec == bief.PREC_FP64:
            assert rel <= 1e-10 and abs(info["niter"] - ir.niter) <= 1, (rel, info, ir.niter)
        else:
            assert rel <= 1e-7, rel
    ctx.close()


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


def _sloped_prisms(nx=8, ny=6, nplan=5, seed=6):
    base = M.random_small_mesh(nx, ny, seed=seed)
    rng = np.random.default_rng(seed + 50)
    bot = -rng.uniform(1, 5, base.npoin)
    top = rng.uniform(-0.2, 0.2, base.npoin)
    top[:3] = bot[:3] + 5e-5  # nearly dry columns
    p = M.extrude_prisms(base, nplan, bot, top)
    n2 = base.npoin
    p.z[n2 + 5:n2 + 8] = p.z[5:8] - 0.05  # crossing planes (inconsistency test)
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    return (base, p, np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb]),
            base.surfac()[eb])


def _close(got, ref, tol):
    return np.abs(got - ref).max() <= tol * (np.abs(ref).max() + 1e-300)


@pytest.mark.parametrize("bn", [0, 40])
def test_device_prism_element_matrices(gpu, bn):
    """MT01PP and MT02PP on the device, assembled by the context (diagonal
    colour by colour in LDS), against the oracle over MT02PP's options.
    Tolerance 1e-12 of the largest entry: the device contracts products
    into fused multiply-adds."""
    base, p, xel, yel, s = _sloped_prisms()
    rng = np.random.default_rng(7)
    f = rng.uniform(0.5, 2, p.npoin)
    g = rng.uniform(0.5, 2, p.npoin)
    h = 10 ** rng.uniform(-4, -2, p.npoin)
    ctx = bief.Context(p, block_nodes=bn)
    ctx.matrix_mt01pp(0.5, s, p.z)
    td, da, xa = ctx.get_matrix()
    t1, x1 = O.mt01pp(s, p.z, p.ikle, 0.5)
    assert td == "Q" and _close(xa, x1, 1e-14)
    assert _close(da, O.assemble_diag(t1, p.ikle, p.npoin), 1e-13)
    for xmul in (1.0, -0.7):
        for hz in (0, 1):
            for p0 in (0, 1):
                for inch in (0, 1):
                    for mon in (0, 1):
                        npou0 = 0 if p0 else base.npoin
                        ctx.matrix_mt02pp(xmul, xel, yel, s, p.z, f, g, h, npou0, bool(hz),
                                          bool(inch), bool(mon))
                        _, da, xa = ctx.get_matrix()
                        t, xm = O.mt02pp(f, g, h, xel, yel, p.z, s, p.ikle, npou0, xmul,
                                         bool(hz), bool(inch), bool(mon))
                        case = (xmul, hz, p0, inch, mon)
                        assert _close(xa, xm, 1e-12), case
                        assert _close(da, O.assemble_diag(t, p.ikle, p.npoin), 1e-12), case
    # 'M=M+N': mass / dt + diffusion, as the vertical diffusion step builds it
    ctx.matrix_mt01pp(0.5, s, p.z)
    ctx.matrix_mt02pp(1.0, xel, yel, s, p.z, f, g, h, base.npoin, accumulate=True)
    _, da, xa = ctx.get_matrix()
    t2, x2 = O.mt02pp(f, g, h, xel, yel, p.z, s, p.ikle, base.npoin, 1.0)
    assert _close(xa, x1 + x2, 1e-12)
    assert _close(da, O.assemble_diag(t1, p.ikle, p.npoin) + O.assemble_diag(t2, p.ikle, p.npoin),
                  1e-12)


def test_prism_vertical_diffusion_solve(gpu):
    """The vertical diffusion system of a sigma mesh (MT01PP / dt + MT02PP
    with a viscosity growing log-linearly towards the surface) assembled on
    the device and solved by CG + diagonal preconditioning, FP64 and mixed,
    against the oracle's solve of the oracle's matrix."""
    base = M.random_small_mesh(12, 10, seed=8)
    rng = np.random.default_rng(9)
    bot = -rng.uniform(2, 20, base.npoin)
    p = M.extrude_prisms(base, 6, bot, np.zeros(base.npoin))
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    xel, yel, s = np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb]), \
        base.surfac()[eb]
    sig = (p.z - np.tile(bot, p.nplan)) / np.tile(-bot, p.nplan)
    nu = 10 ** (-4 + 2 * sig)
    dt = 10.0
    ctx = bief.Context(p, block_nodes=64)
    ctx.matrix_mt01pp(1.0 / dt, s, p.z)
    ctx.matrix_mt02pp(1.0, xel, yel, s, p.z, nu, nu, nu, base.npoin, accumulate=True)
    _, da, xa = ctx.get_matrix()
    t1, x1 = O.mt01pp(s, p.z, p.ikle, 1.0 / dt)
    t2, x2 = O.mt02pp(nu, nu, nu, xel, yel, p.z, s, p.ikle, base.npoin, 1.0)
    A = O.Matrix(O.assemble_diag(t1, p.ikle, p.npoin) + O.assemble_diag(t2, p.ikle, p.npoin),
                 x1 + x2, "Q", "S")
    b = rng.normal(size=p.npoin)
    x_ref, info_ref, _, _ = O.solve(A, p.ikle, b, slv=1, precon=2, eps=1e-10, nitmax=1000)
    for prec in (bief.PREC_FP64, bief.PREC_MIXED):
        ctx.set_matrix("Q", "S", da, xa)
        x, info = ctx.solve(np.zeros(p.npoin), b.copy(), slv=1, precon=2, eps=1e-10,
                            nitmax=1000, precision=prec)
        rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        assert info["converged"], info
        if prec == bief.PREC_FP64:
            assert abs(info["niter"] - info_ref.niter) <= 1, (info, info_ref.niter)
            assert rel <= 1e-10, rel
        else:
            assert rel <= 1e-8, rel


def test_prism_pressure_poisson_solve(gpu):
    """The non-hydrostatic pressure system of telemac3d (prediv.F:210 builds
    it with MT02PP and unit coefficients; :384 prescribes the surface plane
    by DIRICH): assembled on the device from device-resident fields, the
    surface nodes set to zero by the device DIRICH, and solved by CG +
    PRECON 2, FP64 and mixed, against the oracle's DIRICH and SOLVE of the
    oracle's matrix. The operator alone annihilates constants; with the
    surface prescribed it is definite and the solve converges."""
    import torch

    base = M.random_small_mesh(13, 9, seed=31)
    rng = np.random.default_rng(32)
    bot = -rng.uniform(2, 15, base.npoin)
    p = M.extrude_prisms(base, 7, bot, rng.uniform(-0.2, 0.2, base.npoin))
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    xel, yel, s = np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb]), \
        base.surfac()[eb]
    one = np.ones(p.npoin)
    t, xm = O.mt02pp(one, one, one, xel, yel, p.z, s, p.ikle, base.npoin, 1.0)
    A = O.Matrix(O.assemble_diag(t, p.ikle, p.npoin), xm, "Q", "S")
    assert np.abs(O.matvec("X=AY", A, p.ikle, one)).max() <= 1e-12 * np.abs(A.da).max()
    limdir = np.full(p.npoin, 2, dtype=np.int32)
    limdir[(p.nplan - 1) * base.npoin:] = 5
    sm = rng.normal(size=p.npoin)
    fbor = np.zeros(p.npoin)
    f_o, A_o, sm_o = O.dirich(A, p.ikle, sm, fbor, limdir)
    x_ref, info_ref, _, _ = O.solve(A_o, p.ikle, sm_o, f_o, slv=1, precon=2, eps=1e-10,
                                    nitmax=2000)
    assert info_ref.converged
    ctx = bief.Context(p, block_nodes=64)
    ctx.set_geometry(xel, yel, s)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    zd, oned = d(p.z), d(one)
    for prec in (bief.PREC_FP64, bief.PREC_MIXED):
        ctx.matrix_mt02pp(1.0, None, None, None, zd, oned, oned, oned, base.npoin)
        f, b = ctx.dirich(torch.zeros_like(zd), d(sm), torch.zeros_like(zd), d(limdir))
        x, info = ctx.solve(f, b, slv=1, precon=2, eps=1e-10, nitmax=2000, precision=prec)
        x = x.cpu().numpy() if isinstance(x, torch.Tensor) else x
        rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        assert info["converged"], info
        assert np.all(x[limdir == 5] == 0.0)
        if prec == bief.PREC_FP64:
            assert abs(info["niter"] - info_ref.niter) <= 1, (info, info_ref.niter)
            assert rel <= 1e-10, rel
        else:
            assert rel <= 1e-8, rel


def test_prism_device_geometry_matches_caller_arrays(gpu):
    """MT01PP / MT02PP from the geometry kept on the device (set_geometry,
    then xel = yel = surfac = None) give the same bits as from the caller's
    arrays (same layout, same kernels, slot-ordered reads)."""
    base, p, xel, yel, s = _sloped_prisms()
    rng = np.random.default_rng(8)
    f, g = rng.uniform(0.5, 2, p.npoin), rng.uniform(0.5, 2, p.npoin)
    h = 10 ** rng.uniform(-4, -2, p.npoin)
    a, b = bief.Context(p, block_nodes=40), bief.Context(p, block_nodes=40)
    b.set_geometry(xel, yel, s)
    for ctx, geo in ((a, (xel, yel, s)), (b, (None, None, None))):
        ctx.matrix_mt01pp(0.5, geo[2], p.z)
        ctx.matrix_mt02pp(1.0, *geo, p.z, f, g, h, base.npoin, accumulate=True)
    _, da_a, xa_a = a.get_matrix()
    _, da_b, xa_b = b.get_matrix()
    assert np.array_equal(da_a, da_b) and np.array_equal(xa_a, xa_b)
    with pytest.raises(bief.BiefError):
        a.matrix_mt01pp(0.5, None, p.z)  # no geometry kept on this context


@pytest.mark.parametrize("bn", [500, 1000])
def test_prism_element_matrices_records_staged_and_gathered(gpu, bn):
    """The prism element kernels copy each block's node records (32 B) into LDS
    after the diagonal accumulator (8 B per owned node) when they fit the
    160 KB a gfx950 workgroup holds, and gather them from global memory
    otherwise. A 500-plane mesh cut into one-column blocks needs 132 KB (staged,
    above the 64 KB default), into two-column blocks 184 KB (gathered); either
    way MT01PP and MT02PP agree with the oracle, and the off-diagonal terms,
    computed per element from the same values, give the same bits."""
    base = M.random_small_mesh(5, 4, seed=6)
    rng = np.random.default_rng(3)
    bot = -rng.uniform(1, 5, base.npoin)
    p = M.extrude_prisms(base, 500, bot, rng.uniform(-0.2, 0.2, base.npoin))
    st = bief.layout_check(p, bn)
    need = (((st["max_owned"] + 1) // 2) * 2 + 4 * st["max_local"]) * 8
    assert (64 * 1024 < need <= 160 * 1024) if bn == 500 else need > 160 * 1024, st
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    xel, yel, s = np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb]), \
        base.surfac()[eb]
    f, g = rng.uniform(0.5, 2, p.npoin), rng.uniform(0.5, 2, p.npoin)
    h = 10 ** rng.uniform(-4, -2, p.npoin)
    ctx = bief.Context(p, block_nodes=bn)
    ctx.matrix_mt01pp(0.5, s, p.z)
    _, da, xa = ctx.get_matrix()
    t1, x1 = O.mt01pp(s, p.z, p.ikle, 0.5)
    assert _close(xa, x1, 1e-14) and _close(da, O.assemble_diag(t1, p.ikle, p.npoin), 1e-13)
    out = {}
    for mon in (0, 1):
        ctx.matrix_mt02pp(-0.7, xel, yel, s, p.z, f, g, h, base.npoin, False, True, bool(mon))
        _, da, xa = ctx.get_matrix()
        t, xm = O.mt02pp(f, g, h, xel, yel, p.z, s, p.ikle, base.npoin, -0.7, False, True,
                         bool(mon))
        assert _close(xa, xm, 1e-12) and _close(da, O.assemble_diag(t, p.ikle, p.npoin), 1e-12)
        out[mon] = xa
    other = bief.Context(p, block_nodes=1500 - bn)
    for mon in (0, 1):
        other.matrix_mt02pp(-0.7, xel, yel, s, p.z, f, g, h, base.npoin, False, True, bool(mon))
        assert np.array_equal(other.get_matrix()[2], out[mon])
    # 'M=M+N' into a symmetric and a non-symmetric target (both halves),
    # MT01PP and MT02PP, in the same layout
    da0 = rng.uniform(1, 2, p.npoin)
    for typext, nc in (("S", 15), ("Q", 30)):
        xa0 = rng.normal(size=(nc, p.nelem))
        ctx.set_matrix("Q", typext, da0, xa0)
        ctx.matrix_mt01pp(0.5, s, p.z, accumulate=True)
        ctx.matrix_mt02pp(-0.7, xel, yel, s, p.z, f, g, h, base.npoin, False, True, False,
                          accumulate=True)
        assert ctx.matrix_info() == ("Q", typext, nc)
        td, da, xa = ctx.get_matrix()
        t2, x2 = O.mt02pp(f, g, h, xel, yel, p.z, s, p.ikle, base.npoin, -0.7, False, True,
                          False)
        want = xa0 + np.vstack([x1 + x2] * (nc // 15))
        assert td == "Q" and xa.shape == want.shape
        assert _close(xa, want, 1e-12), typext
        assert _close(da, da0 + O.assemble_diag(t1, p.ikle, p.npoin)
                      + O.assemble_diag(t2, p.ikle, p.npoin), 1e-12), typext
