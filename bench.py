#!/usr/bin/env python3
"""Benchmark of the BIEF linear-system hot path on one MI355X.

Workload (the unstructured 2M-node configuration): seeded Poisson-disk
points in a 50 km x 30 km estuary polygon with a 4x denser 2 km channel,
Delaunay triangulated (~2.0 M nodes, ~4.0 M P1 triangles); depth 0.5-40 m
from a seeded sum of 32 Gaussians; one telemac2d wave-equation step
(dt = 2 s, theta = 1) with a 0.2 m tide imposed on the sea boundary; the
symmetric system solved by diagonally preconditioned CG to the 1e-8
relative residual of the preconditioned system (GRACJG's stop rule).

A step = restore the assembled EBE matrix from its device copy (SOLVE
consumes it, as BIEF's does; telemac2d keeps a copy, propag.F:1643) + one
SOLVE, data resident in HBM.
`value` is mixed-precision solves/s (FP32 EBE CG inside FP64 iterative
refinement); the FP64 path is timed the same way and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "EBE-CG solves/s at 2M nodes, FP64 vs mixed @1e-8 resid; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(npoin, seed, cache=None):
    from opentelemac_amd import mesh as M, system as S
    t = time.time()
    if cache and os.path.exists(cache):
        z = np.load(cache)
        m = M.Mesh(x=z["x"], y=z["y"], ikle=z["ikle"], npoin2=z["x"].size,
                   nelem2=z["ikle"].shape[1], meta={"kind": "estuary", "seed": seed})
        sysd = {k: z[k] for k in ("da", "xa", "b", "x0")}
        sysd.update(typdia="Q", typext="S")
        log(f"workload loaded from {cache}")
        return m, sysd
    m = M.estuary_mesh(npoin, seed=seed)
    depth = M.smooth_random_depth(m, seed=seed)
    bnd = S.boundary_nodes(m)
    sea = bnd[m.x[bnd] > m.x.max() - 1.0]
    sysd = S.wave_system(m, depth, dt=2.0, theta=1.0, inflow=sea)
    log(f"mesh {m.npoin} nodes {m.nelem} elements, sea nodes {sea.size}, "
        f"built in {time.time() - t:.1f}s")
    if cache:
        np.savez(cache, x=m.x, y=m.y, ikle=m.ikle, da=sysd["da"], xa=sysd["xa"],
                 b=sysd["b"], x0=sysd["x0"])
    return m, sysd


def cpu_baseline(m, sysd, budget_s=12.0):
    """Reference SOLVE (flang-built BIEF, sequential, 1 core) on the same
    system, repeated from the same inputs until about budget_s of CPU time;
    falls back to the C restatement if the reference build is absent."""
    from oracle import oracle as O
    A = O.Matrix(sysd["da"], sysd["xa"], "Q", "S")
    if O.have_ref():
        rm = O.RefMesh(m.ikle, m.npoin, m.nelem)

        def one():
            x, info, _, _ = O.ref_solve(rm, A, sysd["b"], sysd["x0"], slv=1, precon=2,
                                        eps=1e-8, nitmax=5000)
            return x, info["niter"]
        kind = "reference"
    else:
        def one():
            x, info, _, _ = O.solve(A, m.ikle, sysd["b"], sysd["x0"], slv=1, precon=2,
                                    eps=1e-8, nitmax=5000)
            return x, info.niter
        kind = "port"
    n, t0 = 0, time.perf_counter()
    while True:
        x, niter = one()
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            break
    return {"value": n / dt, "unit": "solves/s", "cores": 1, "kind": kind,
            "sample": f"{n} full FP64 SOLVEs (CG, PRECON 2, eps 1e-8) of the same "
                      f"{m.npoin}-node system from the same inputs, host-resident, "
                      f"{niter} iterations each, {dt:.2f} s in total",
            "iterations": niter}, x


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--npoin", type=int, default=2_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--block-nodes", type=int, default=0)
    ap.add_argument("--inner-eps", type=float, default=1e-4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cache", default=None, help="npz file to reuse the synthetic workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 256 MiB (profiler counter scale)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from opentelemac_amd import bief

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": src.numel() * 8,
                          "launches": 5}), flush=True)
        return
    m, sysd = build_workload(args.npoin, args.seed, args.cache)
    t = time.time()
    ctx = bief.Context(m, block_nodes=args.block_nodes)
    st = ctx.stats()
    log(f"context in {time.time() - t:.1f}s: {st}")
    da0 = torch.from_numpy(sysd["da"]).to(dev)
    xa0 = torch.from_numpy(np.ascontiguousarray(sysd["xa"])).to(dev)
    b0 = torch.from_numpy(sysd["b"]).to(dev)
    x00 = torch.from_numpy(sysd["x0"]).to(dev)
    x = torch.empty_like(x00)
    b = torch.empty_like(b0)

    ctx.set_matrix("Q", "S", da0, xa0)
    ctx.save_matrix()

    def step(precision):
        # SOLVE consumes A (D A D in place): restore the assembled copy first,
        # as telemac2d copies AM1 before its SOLVE (propag.F:1643)
        ctx.restore_matrix()
        x.copy_(x00)
        b.copy_(b0)
        _, info = ctx.solve(x, b, slv=1, precon=2, eps=1e-8, nitmax=5000, precision=precision,
                            inner_eps=args.inner_eps)
        return info

    results = {}
    sols = {}
    for name, prec in (("fp64", bief.PREC_FP64), ("mixed", bief.PREC_MIXED)):
        for _ in range(args.warmup):
            step(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = [step(prec) for _ in range(args.steps)]
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        info = infos[-1]
        results[name] = {"ms_per_solve": 1e3 * el / args.steps, "solves_per_s": args.steps / el,
                         "iterations": info["niter"], "outer": info["outer_iters"],
                         "relres": info["relres"], "converged": bool(info["converged"])}
        sols[name] = x.detach().cpu().numpy().copy()
        # kernel timings on this path's data (HIP events, context stream)
        ms_dir, ms_upd = ctx.time_cg_kernels(prec, reps=50)
        results[name]["ms_dir_kernel"] = ms_dir
        results[name]["ms_upd_kernel"] = ms_upd
        its = info["niter"]
        results[name]["cg_iters_per_s"] = its * args.steps / el
        log(name, json.dumps(results[name]))

    # true residual ||b - A x|| / ||b|| of both solutions (FP64 mat-vec)
    ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
    for name in ("fp64", "mixed"):
        ax = ctx.matvec("X=AY", sols[name])
        results[name]["true_relres"] = float(np.linalg.norm(sysd["b"] - ax)
                                             / np.linalg.norm(sysd["b"]))
    rel_mixed_vs_fp64 = float(np.linalg.norm(sols["mixed"] - sols["fp64"])
                              / np.linalg.norm(sols["fp64"]))

    ne, nprime = m.nelem, m.npoin
    # algorithmic bytes of the fused CG direction + EBE mat-vec kernel per
    # launch: coefficients (3 per element) + int32 connectivity (3 per
    # element) + R and D_old read, D_new and AD written (4 node vectors)
    bytes_f64 = ne * (3 * 8 + 12) + nprime * 4 * 8
    bytes_f32 = ne * (3 * 4 + 12) + nprime * 4 * 4
    ach32 = bytes_f32 / (results["mixed"]["ms_dir_kernel"] * 1e-3) / 1e9
    ach64 = bytes_f64 / (results["fp64"]["ms_dir_kernel"] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_cg_dir_estuary2m.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        traffic = tj.get("mixed_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach32 / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_ebe<float,...,STAGE_CGDIR,EPI_CGAD,FIN_CG> (mixed inner CG)",
                "algorithmic_bytes_per_launch": bytes_f32,
                "avg_launch_ms": results["mixed"]["ms_dir_kernel"],
                "fp64_kernel": {"achieved": round(ach64, 1), "frac": round(ach64 / HBM_PEAK_GBS, 4),
                                "algorithmic_bytes_per_launch": bytes_f64,
                                "avg_launch_ms": results["fp64"]["ms_dir_kernel"]}}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu, xcpu = cpu_baseline(m, sysd)
            cpu["gpu_fp64_rel_l2_vs_cpu"] = float(np.linalg.norm(sols["fp64"] - xcpu)
                                                  / np.linalg.norm(xcpu))
            cpu["gpu_mixed_rel_l2_vs_cpu"] = float(np.linalg.norm(sols["mixed"] - xcpu)
                                                   / np.linalg.norm(xcpu))
            cpu["host_cpu"] = _cpu_model()
            cpu["nproc"] = os.cpu_count()
        except Exception as e:  # reported, never silently substituted
            cpu = {"value": None, "error": repr(e)}

    mixed = results["mixed"]
    line = {
        "metric": METRIC,
        "value": round(mixed["solves_per_s"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_solve"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "unstructured estuary wave-equation CG, 2M nodes",
                   "npoin": m.npoin, "nelem": m.nelem, "dt_s": 2.0, "eps": 1e-8,
                   "solver": "CG, diagonal preconditioning (PRECON 2)",
                   "parallelism": "single GPU"},
        "fp64": results["fp64"],
        "mixed": mixed,
        "mixed_speedup_vs_fp64": round(results["fp64"]["ms_per_solve"] / mixed["ms_per_solve"], 3),
        "mixed_vs_fp64_rel_l2": rel_mixed_vs_fp64,
        "layout": st,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
