"""Writes the golden vectors of the reference's own BIEF (oracle/_ref, the
flang build of sources/utils/bief) for the CPU oracle's tests.

Run in the development container, where /root/reference is present:
    make -C oracle ref && python tests/golden/make_golden.py
Inputs are regenerated from their seeds by the tests; the files hold the
reference's outputs (and the inputs, for self-containedness).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from opentelemac_amd import mesh as M  # noqa: E402
from opentelemac_amd import system as S  # noqa: E402
from oracle import oracle as O  # noqa: E402


def small_case():
    m = M.random_small_mesh(9, 7, seed=3)
    ik, n, ne = m.ikle, m.npoin, m.nelem
    rm = O.RefMesh(ik, n, ne)
    rng = np.random.default_rng(20)
    out = {"x": m.x, "y": m.y, "ikle": ik}
    xe, ye = m.xel_yel()
    s = m.surfac()
    u = rng.uniform(1, 2, n)
    v = rng.uniform(1, 3, n)
    ue = rng.uniform(1, 2, ne)
    out["u"], out["v"], out["ue"] = u, v, ue
    out["mt01aa_t"], out["mt01aa_xm"] = O.ref_mt01aa(s, 0.7)
    for kind, uu in ((0, ue), (1, u), (2, u)):
        for mon in (0, 1):
            t, xm = O.ref_mt02aa(kind, xe, ye, s, ik, uu, v, xmul=-1.3, mon=bool(mon), npoin=n)
            out[f"mt02aa_{kind}_{mon}_t"] = t
            out[f"mt02aa_{kind}_{mon}_xm"] = xm
    da = rng.uniform(2, 3, n)
    out["da"] = da
    for typext, nc in (("S", 3), ("Q", 6)):
        xa = rng.uniform(-1, 1, (nc, ne))
        out[f"xa_{typext}"] = xa
        for typdia in "QI0":
            A = O.Matrix(da, xa, typdia, typext)
            for op, code in O.OPS.items():
                y = rng.normal(size=n)
                x0 = rng.normal(size=n)
                key = f"mv_{typdia}{typext}_{code}"
                out[key + "_y"] = y
                out[key + "_x0"] = x0
                out[key] = O.ref_matvec(rm, op, A, y, x0, c=0.37)
    # SPD system: mass + diffusion
    t1, x1 = O.mt01aa(s, 1.0)
    t2, x2 = O.mt02aa(1, xe, ye, s, ik, u, xmul=2.0)
    Ada = O.assemble_diag(t1 + t2, ik, n)
    Axa = x1 + x2
    b = rng.normal(size=n)
    out["spd_da"], out["spd_xa"], out["spd_b"] = Ada, Axa, b
    A = O.Matrix(Ada, Axa, "Q", "S")
    for slv in (1, 6, 7):
        for precon in (0, 2, 5):
            x, info, A2, b2 = O.ref_solve(rm, A, b, slv=slv, precon=precon, krylov=5,
                                          eps=1e-10, nitmax=500)
            key = f"solve_{slv}_{precon}"
            out[key + "_x"] = x
            out[key + "_niter"] = np.array(info["niter"])
            out[key + "_relres"] = np.array(info["relres"])
            out[key + "_da"] = A2.da
            out[key + "_xa"] = A2.xa
            out[key + "_b"] = b2
            out[key + "_typdia"] = np.array(ord(A2.typdia))
    limdir = np.where(rng.uniform(size=n) < 0.2, 5, 2).astype(np.int32)
    fbor = np.where(limdir == 5, rng.normal(size=n), 0.0)
    out["dir_limdir"], out["dir_fbor"] = limdir, fbor
    f, A3, sm = O.ref_dirich(rm, A, b, fbor, limdir)
    out["dir_f"], out["dir_da"], out["dir_xa"], out["dir_sm"] = f, A3.da, A3.xa, sm
    out["lump"] = O.ref_lump(rm, A, 0.3)
    return out


def channel_case():
    """The 20,451-node channel wave-equation step, CG + PRECON 2, eps 1e-8."""
    m = M.channel_mesh()
    sysd = S.wave_system(m, M.channel_depth(m), dt=1.0)
    rm = O.RefMesh(m.ikle, m.npoin, m.nelem)
    A = O.Matrix(sysd["da"], sysd["xa"], "Q", "S")
    x, info, _, _ = O.ref_solve(rm, A, sysd["b"], sysd["x0"], slv=1, precon=2, eps=1e-8,
                                nitmax=5000)
    return {"x": x, "niter": np.array(info["niter"]), "relres": np.array(info["relres"])}


def prism_case():
    """Prisms (MV0606, MT01PP): 4 planes over a small base mesh."""
    base = M.random_small_mesh(5, 4, seed=3)
    rng = np.random.default_rng(30)
    bot = -rng.uniform(2, 5, base.npoin)
    p = M.extrude_prisms(base, 4, bot, np.zeros(base.npoin))
    s = np.tile(base.surfac(), p.nplan - 1)
    rm = O.RefMesh(p.ikle, base.npoin, base.nelem, p.nplan)
    out = {"z": p.z, "ikle": p.ikle}
    out["mt01pp_t"], out["mt01pp_xm"] = O.ref_mt01pp(s, p.z, p.ikle, 0.7)
    da = rng.uniform(1, 2, p.npoin)
    out["da"] = da
    for typext, nc in (("S", 15), ("Q", 30)):
        xa = rng.normal(size=(nc, p.nelem))
        out[f"xa_{typext}"] = xa
        for typdia in "QI":
            A = O.Matrix(da, xa, typdia, typext)
            for op, code in O.OPS.items():
                y = rng.normal(size=p.npoin)
                x0 = rng.normal(size=p.npoin)
                key = f"mv_{typdia}{typext}_{code}"
                out[key + "_y"], out[key + "_x0"] = y, x0
                out[key] = O.ref_matvec(rm, op, A, y, x0, c=0.3)
    tm, xm = O.mt01pp(s, p.z, p.ikle, 1.0)
    A = O.Matrix(O.assemble_diag(tm, p.ikle, p.npoin), xm, "Q", "S")
    b = rng.normal(size=p.npoin)
    out["spd_da"], out["spd_xa"], out["spd_b"] = A.da, A.xa, b
    for slv in (1, 6, 7):
        x, info, _, _ = O.ref_solve(rm, A, b, slv=slv, precon=2, eps=1e-10)
        out[f"solve_{slv}_x"] = x
        out[f"solve_{slv}_niter"] = np.array(info["niter"])
    return out


def mt02pp_case():
    """Prism diffusion matrix (MT02PP) over 4 planes, with nearly dry columns
    and a crossing plane (for the hydrostatic inconsistency test): every
    option of the routine."""
    base = M.random_small_mesh(6, 5, seed=4)
    rng = np.random.default_rng(40)
    bot = -rng.uniform(1, 5, base.npoin)
    top = rng.uniform(-0.2, 0.2, base.npoin)
    top[:3] = bot[:3] + 5e-5
    p = M.extrude_prisms(base, 4, bot, top)
    # a middle plane pushed below its lower neighbour at a few nodes
    n2 = base.npoin
    p.z[n2 + 5:n2 + 8] = p.z[5:8] - 0.05
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    xel, yel = np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb])
    s = base.surfac()[eb]
    f = rng.uniform(0.5, 2, p.npoin)
    g = rng.uniform(0.5, 2, p.npoin)
    h = 10 ** rng.uniform(-4, -2, p.npoin)
    out = {"z": p.z, "ikle": p.ikle, "xel": xel, "yel": yel, "surfac": s, "f": f, "g": g,
           "h": h, "npoin2": np.array(n2), "nplan": np.array(p.nplan)}
    for xmul in (1.0, -0.7):
        for hz in (0, 1):
            for p0 in (0, 1):
                for inch in (0, 1):
                    for mon in (0, 1):
                        t, xm = O.ref_mt02pp(f, g, h, xel, yel, p.z, s, p.ikle, p.nplan, xmul,
                                             bool(hz), bool(p0), bool(inch), bool(mon))
                        key = f"mt02pp_{'p' if xmul > 0 else 'n'}{hz}{p0}{inch}{mon}"
                        out[key + "_t"], out[key + "_xm"] = t, xm
    return out


def seg_case():
    """Segment storage (MATRIX STORAGE = 3, MVSEG / OMSEG): all MATVEC
    operations, SOLVE (CG, BiCGSTAB, GMRES x PRECON 0/2/5), DIRICH, LUMP on
    the segments of a small triangle mesh."""
    m = M.random_small_mesh(9, 7, seed=3)
    gl, es = M.segments(m.ikle)
    rm = O.RefMesh(m.ikle, m.npoin, m.nelem, gloseg=gl)
    rng = np.random.default_rng(50)
    out = {"gloseg": gl, "eltseg": es}
    da = rng.uniform(2, 3, m.npoin)
    out["da"] = da
    for typext, nc in (("S", 1), ("Q", 2)):
        xa = rng.uniform(-1, 1, (nc, gl.shape[1]))
        out[f"xa_{typext}"] = xa
        for typdia in "QI0":
            A = O.Matrix(da, xa, typdia, typext)
            for op, code in O.OPS.items():
                y = rng.normal(size=m.npoin)
                x0 = rng.normal(size=m.npoin)
                key = f"mv_{typdia}{typext}_{code}"
                out[key + "_y"], out[key + "_x0"] = y, x0
                out[key] = O.ref_matvec(rm, op, A, y, x0, c=0.37)
    sysd = S.wave_system(m, 1.0 + m.x / 4.0, dt=2.0)
    xs = S.ebe_to_segments(sysd["xa"], es, m.ikle, gl, "S")
    out["spd_da"], out["spd_xa"], out["spd_b"], out["spd_x0"] = sysd["da"], xs, sysd["b"], \
        sysd["x0"]
    A = O.Matrix(sysd["da"], xs, "Q", "S")
    for slv in (1, 6, 7):
        for precon in (0, 2, 5):
            x, info, A2, b2 = O.ref_solve(rm, A, sysd["b"], sysd["x0"], slv=slv, precon=precon,
                                          krylov=5, eps=1e-10, nitmax=500)
            key = f"solve_{slv}_{precon}"
            out[key + "_x"], out[key + "_niter"] = x, np.array(info["niter"])
            out[key + "_xa"], out[key + "_b"] = A2.xa, b2
    limdir = np.where(rng.uniform(size=m.npoin) < 0.2, 5, 2).astype(np.int32)
    fbor = np.where(limdir == 5, rng.normal(size=m.npoin), 0.0)
    sm = rng.normal(size=m.npoin)
    out["dir_limdir"], out["dir_fbor"], out["dir_sm"] = limdir, fbor, sm
    f, A3, sm2 = O.ref_dirich(rm, A, sm, fbor, limdir)
    out["dir_f"], out["dir_xa"], out["dir_sm_out"] = f, A3.xa, sm2
    out["lump"] = O.ref_lump(rm, A, 0.3)
    return out


def gouttedo_case():
    """The reference's gouttedo mesh (geo_gouttedo.slf, 4624 nodes) and the
    first step of its drop: CG, BiCGSTAB and GMRES(3) with PRECON 2 at
    eps 1e-10 from a zero initial guess, in element and in segment storage."""
    from opentelemac_amd import serafin as SF
    m = SF.read(os.path.join(HERE, "geo_gouttedo.slf")).mesh()
    gl, es = M.segments(m.ikle)
    sysd = S.drop_system(m)
    out = {}
    for sto, ikle, xa in (("ebe", m.ikle, sysd["xa"]),
                          ("seg", gl, S.ebe_to_segments(sysd["xa"], es, m.ikle, gl, "S"))):
        rm = O.RefMesh(m.ikle, m.npoin, m.nelem, gloseg=gl if sto == "seg" else None)
        A = O.Matrix(sysd["da"], xa, "Q", "S")
        for slv in (1, 6, 7):
            x, info, _, _ = O.ref_solve(rm, A, sysd["b"], np.zeros(m.npoin), slv=slv,
                                        precon=2, krylov=3, eps=1e-10, nitmax=500)
            out[f"{sto}_{slv}_x"], out[f"{sto}_{slv}_niter"] = x, np.array(info["niter"])
    return out


def supg_flow(m):
    """A rigid rotation about the gouttedo basin's centre, |u| <= 2 m/s."""
    w = 2.0 / (10.05 * np.sqrt(2.0))
    return -w * (m.y - 10.05), w * (m.x - 10.05)


def supg_case():
    """SUPG element routines (KSPG11, MT03AA, MT04AA, MT05AA in its P1, P0
    and upwind forms) on the small mesh, and CVDFTR's SUPG tracer system on
    the gouttedo mesh (rotating flow, nu = 0.1, dt = 0.5 s, a Gaussian
    patch) solved by GMRES(5) and BiCGSTAB with PRECON 2 at eps 1e-10."""
    from opentelemac_amd import serafin as SF
    out = {}
    m = M.random_small_mesh(9, 7, seed=3)
    xe, ye = m.xel_yel()
    sf = m.surfac()
    rng = np.random.default_rng(60)
    u, v = rng.uniform(-2, 2, m.npoin), rng.uniform(-2, 2, m.npoin)
    out["small_u"], out["small_v"] = u, v
    kx, ky = O.ref_kspg11(xe, ye, m.ikle, u, v, 1.0)
    out["kspg11_kx"], out["kspg11_ky"] = kx, ky
    out["mt03aa_t"], out["mt03aa_xm"] = O.ref_mt03aa(kx, ky, u, v, xe, ye, sf, m.ikle, 0.7)
    out["mt04aa_t"], out["mt04aa_xm"] = O.ref_mt04aa(u, v, xe, ye, sf, m.ikle, 0.3)
    out["mt05aa_t"], out["mt05aa_xm"] = O.ref_mt05aa(u, v, xe, ye, m.ikle, 1.3)
    out["mt05p0_t"], out["mt05p0_xm"] = O.ref_mt05aa(kx, ky, xe, ye, m.ikle, 2.0, u_p0=True)
    out["mt05n_t"], out["mt05n_xm"] = O.ref_mt05aa(u, v, xe, ye, m.ikle, 1.0, upwind=True)
    g = SF.read(os.path.join(HERE, "geo_gouttedo.slf")).mesh()
    gu, gv = supg_flow(g)
    fn = np.exp(-((g.x - 6.0) ** 2 + (g.y - 10.0) ** 2) / 4.0)
    A, sm = O.supg_system(g, gu, gv, np.full(g.npoin, 0.1), 0.5, fn, reference=True)
    out["sys_da"], out["sys_xa"], out["sys_sm"] = A.da, A.xa, sm
    rm = O.RefMesh(g.ikle, g.npoin, g.nelem)
    for slv in (7, 6):
        x, info, _, _ = O.ref_solve(rm, A, sm, fn, slv=slv, precon=2, krylov=5, eps=1e-10,
                                    nitmax=500)
        out[f"solve_{slv}_x"], out[f"solve_{slv}_niter"] = x, np.array(info["niter"])
    return out


def propag_state(m, seed, shallow=False, dry=False):
    """A state of the wave equation on mesh m (seeded): a sloping bottom with
    a cross-channel ripple, a free surface of a cosine wave and noise, random
    velocities and Strickler coefficients; shallow: 5 % of the nodes at 1-30 mm
    depth (the shallow-water friction speed of fricti.f); dry: 2 % at or
    below the bottom (the depth floors)."""
    rng = np.random.default_rng(seed)
    lx = m.x.max() - m.x.min()
    ly = m.y.max() - m.y.min()
    zb = -(1.0 + 4.0 * (m.x - m.x.min()) / lx) + 0.3 * np.sin(m.y / (0.3 * ly + 1.0))
    eta = 0.05 * np.cos(2 * np.pi * (m.x - m.x.min()) / lx) + 0.01 * rng.normal(size=m.npoin)
    if shallow:
        pick = rng.uniform(size=m.npoin) < 0.05
        zb = np.where(pick, eta - rng.uniform(1e-3, 0.03, m.npoin), zb)
    if dry:
        pick = rng.uniform(size=m.npoin) < 0.02
        zb = np.where(pick, eta + rng.uniform(0.0, 1e-3, m.npoin), zb)
    un = 0.3 * rng.normal(size=m.npoin)
    vn = 0.1 * rng.normal(size=m.npoin)
    k = rng.uniform(20, 60, m.npoin)
    return eta - zb, un, vn, zb, k


def propag_case():
    """One wave-equation step of telemac2d PROPAG computed by the reference's
    own routines (driver ref_propag: MT01AA, MT02AA 'MATDIFUV', MT08AA,
    VC00AA, VC13AA, SLOP10, ASSVEC, OV/OVD, FRICTION_CALC, FRICTI, MATVEC,
    DIRICH, SOLVE) on three meshes:
      chan   a 400 m x 80 m channel (60 x 12 cells), the free surface 0.2 m
             prescribed at x = 0, dt = 1 s, shallow nodes;
      goutte the reference's gouttedo basin (closed), dt = 0.04 s, dry nodes;
      estim  the reference's estimation channel with its boundary file: the
             depth prescribed where the file prescribes it, walls elsewhere
             (the file's prescribed-velocity inflow is a wall here), dt = 2 s.
    Inputs and outputs are stored per case, prefixed by the case name."""
    from opentelemac_amd import serafin as SF
    out = {}
    cases = []
    m = M.channel_mesh(60, 12, 400.0, 80.0)
    sea = np.nonzero(m.x <= m.x.min() + 1e-9)[0]
    cases.append(("chan", m, sea, 1.0, dict(shallow=True)))
    g = SF.read(os.path.join(HERE, "geo_gouttedo.slf")).mesh()
    cases.append(("goutte", g, np.zeros(0, dtype=np.int64), 0.04, dict(dry=True)))
    e = SF.read(os.path.join(HERE, "geo_estimation.slf")).mesh()
    b = SF.read_cli(os.path.join(HERE, "geo_estimation.cli"), e.npoin)
    _, lim = SF.depth_dirichlet(b, e.npoin)
    cases.append(("estim", e, np.nonzero(lim == SF.KDIR)[0], 2.0, {}))
    for seed, (name, mesh, sea, dt, kw) in enumerate(cases):
        hn, un, vn, zb, k = propag_state(mesh, 40 + seed, **kw)
        limdir = np.full(mesh.npoin, SF.KNEU, dtype=np.int32)
        limdir[sea] = SF.KDIR
        hbor = np.where(limdir == SF.KDIR, 0.2 - zb, 0.0)
        r = O.ref_propag(mesh, hn, un, vn, zb, k, hbor, limdir, kdir=SF.KDIR, dt=dt,
                         nitmax=2000, eps=1e-13)
        out.update({f"{name}_{key}": val for key, val in
                    dict(hn=hn, un=un, vn=vn, zb=zb, k=k, limdir=limdir, hbor=hbor,
                         dt=np.array(dt)).items()})
        for key in ("cf", "fu", "am2d", "dm1", "t2", "t3", "da", "xa", "cv1", "dh", "u", "v"):
            out[f"{name}_{key}"] = r[key]
    out["chan_x"], out["chan_y"], out["chan_ikle"] = cases[0][1].x, cases[0][1].y, cases[0][1].ikle
    return out


CASES = {"ref_prism": prism_case, "ref_small": small_case, "ref_channel": channel_case,
         "ref_mt02pp": mt02pp_case, "ref_seg": seg_case, "ref_gouttedo": gouttedo_case,
         "ref_supg": supg_case, "ref_propag": propag_case}


def main(names):
    """Writes the named fixtures (all of them by default)."""
    for name in names or CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **CASES[name]())
        print("golden vectors written:", name)


if __name__ == "__main__":
    main(sys.argv[1:])
