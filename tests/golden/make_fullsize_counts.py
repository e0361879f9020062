"""Writes ref_fullsize_counts.json: the iteration counts and stop-test
precisions of the reference's own SOLVE (oracle/_ref, the flang build of
sources/utils/bief, sequential) on the full-size systems that the GPU tests
solve, each system assembled here exactly as the test assembles it:

  * estuary_2m_cg: the 2,003,181-node estuary wave-equation system, CG +
    PRECON 2 to 1e-8;
  * prisms_5m_diffusion: the 5,014,910-node sigma mesh (the 500k estuary
    base, 10 planes), MT01PP / 2 s + MT02PP with nu = 10^(-4 + 2 sigma),
    assembled by the reference's MT01PP/MT02PP, CG + PRECON 2 to 1e-8 from
    zero, right-hand side seeded normal (seed 12);
  * prisms_5m_poisson: MT02PP with unit coefficients (the non-hydrostatic
    pressure Poisson operator of prediv.F:210), zero Dirichlet values on
    the surface plane by the reference's DIRICH (prediv.F:384), CG + PRECON
    2 to 1e-8, right-hand side seeded normal (seed 13);
  * supg_2m_gmres: CVDFTR's SUPG tracer system of the 2M estuary (synthetic
    flow, nu = 0.1, dt = 10 s) from the reference's element routines, DIRICH
    on the sea boundary, GMRES(5) + PRECON 2 to 1e-8.

Run in the development container, where /root/reference is present (about
five minutes and 20 GB of memory):
    make -C oracle ref && python tests/golden/make_fullsize_counts.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from opentelemac_amd import mesh as M  # noqa: E402
from opentelemac_amd import system as S  # noqa: E402
from opentelemac_amd import tracer  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "ref_fullsize_counts.json")


def _entry(info, t0, **kw):
    d = {"niter": int(info["niter"]), "relres": float(info["relres"]),
         "listing": info["line"], "seconds": round(time.time() - t0, 1)}
    d.update(kw)
    return d


def estuary_cases(out):
    m = M.estuary_mesh(2_000_000, seed=0)
    depth = M.smooth_random_depth(m, seed=0)
    bnd = S.boundary_nodes(m)
    sea = bnd[m.x[bnd] > m.x.max() - 1.0]
    sysd = S.wave_system(m, depth, dt=2.0, theta=1.0, inflow=sea)
    rm = O.RefMesh(m.ikle, m.npoin, m.nelem)
    t0 = time.time()
    _, info, _, _ = O.ref_solve(rm, O.Matrix(sysd["da"], sysd["xa"], "Q", "S"), sysd["b"],
                                sysd["x0"], slv=1, precon=2, eps=1e-8, nitmax=5000)
    out["estuary_2m_cg"] = _entry(info, t0, npoin=m.npoin, nelem=m.nelem)
    print("estuary_2m_cg", out["estuary_2m_cg"], flush=True)
    del sysd
    u, v, fn, sea = tracer.estuary_flow(m, seed=0)
    t0 = time.time()
    A, sm = O.supg_system(m, u, v, np.full(m.npoin, 0.1), 10.0, fn, reference=True)
    limdir = np.full(m.npoin, 2, dtype=np.int32)
    limdir[sea] = 5
    f, A2, sm2 = O.ref_dirich(rm, A, sm, np.zeros(m.npoin), limdir, f=fn)
    _, info, _, _ = O.ref_solve(rm, A2, sm2, f, slv=7, precon=2, krylov=5, eps=1e-8,
                                nitmax=2000)
    out["supg_2m_gmres"] = _entry(info, t0, npoin=m.npoin, restarts=int(info["niter"]))
    print("supg_2m_gmres", out["supg_2m_gmres"], flush=True)


def prism_mesh():
    """The test's 5M sigma mesh and per-prism base geometry."""
    base = M.estuary_mesh(500_000, seed=0)
    depth = M.smooth_random_depth(base, seed=0)
    p = M.extrude_prisms(base, 10, -depth, np.zeros(base.npoin))
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    return base, p, np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb]), \
        base.surfac()[eb]


def prism_cases(out):
    base, p, xel, yel, s = prism_mesh()
    rm = O.RefMesh(p.ikle, base.npoin, base.nelem, p.nplan)
    sig = np.repeat(np.linspace(0.0, 1.0, 10), base.npoin)
    nu = 10.0 ** (-4.0 + 2.0 * sig)
    t0 = time.time()
    t1, x1 = O.ref_mt01pp(s, p.z, p.ikle, 0.5)
    t2, x2 = O.ref_mt02pp(nu, nu, nu, xel, yel, p.z, s, p.ikle, p.nplan, 1.0)
    da = O.assemble_diag(t1, p.ikle, p.npoin) + O.assemble_diag(t2, p.ikle, p.npoin)
    A = O.Matrix(da, x1 + x2, "Q", "S")
    del t1, x1, t2, x2
    b = np.random.default_rng(12).normal(size=p.npoin)
    _, info, _, _ = O.ref_solve(rm, A, b, np.zeros(p.npoin), slv=1, precon=2, eps=1e-8,
                                nitmax=5000)
    out["prisms_5m_diffusion"] = _entry(info, t0, npoin=p.npoin, nelem=p.nelem)
    print("prisms_5m_diffusion", out["prisms_5m_diffusion"], flush=True)
    del A
    t0 = time.time()
    one = np.ones(p.npoin)
    t3, x3 = O.ref_mt02pp(one, one, one, xel, yel, p.z, s, p.ikle, p.nplan, 1.0)
    A = O.Matrix(O.assemble_diag(t3, p.ikle, p.npoin), x3, "Q", "S")
    del t3, x3
    limdir = np.full(p.npoin, 2, dtype=np.int32)
    limdir[(p.nplan - 1) * base.npoin:] = 5
    sm = np.random.default_rng(13).normal(size=p.npoin)
    f, A2, sm2 = O.ref_dirich(rm, A, sm, np.zeros(p.npoin), limdir)
    _, info, _, _ = O.ref_solve(rm, A2, sm2, f, slv=1, precon=2, eps=1e-8, nitmax=5000)
    out["prisms_5m_poisson"] = _entry(info, t0, npoin=p.npoin, nelem=p.nelem)
    print("prisms_5m_poisson", out["prisms_5m_poisson"], flush=True)


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref/libbief_ref.so missing: run `make -C oracle ref`")
    out = {}
    estuary_cases(out)
    prism_cases(out)
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print("written", OUT)


if __name__ == "__main__":
    main()
