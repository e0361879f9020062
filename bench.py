#!/usr/bin/env python3
"""Benchmark of the BIEF linear-system hot path on one MI355X.

Headline (the unstructured 2M-node configuration): seeded Poisson-disk
points in a 50 km x 30 km estuary polygon with a 4x denser 2 km channel,
Delaunay triangulated (~2.0 M nodes, ~4.0 M P1 triangles); depth 0.5-40 m
from a seeded sum of 32 Gaussians; one telemac2d wave-equation step
(dt = 2 s, theta = 1) with a 0.2 m tide imposed on the sea boundary; the
symmetric system solved by diagonally preconditioned CG to the 1e-8
relative residual of the preconditioned system (GRACJG's stop rule).

A step = restore the assembled EBE matrix from its device copy (SOLVE
consumes it, as BIEF's does; telemac2d keeps a copy, propag.F:1643) + one
SOLVE, data resident in HBM. The restore is fused into SOLVE's diagonal
preconditioning (bhip_solve_dev_restored), which reads the saved copy.
`value` is mixed-precision solves/s (FP32 EBE CG inside FP64 iterative
refinement); the FP64 path is timed the same way and reported beside it.

The default run (--workload all) measures every other configuration in the
same process and adds them to the one JSON line under "configs": the 20k
channel against the reference's CPU SOLVE, the headline system in segment
storage (telemac2d's default MATRIX STORAGE = 3), the 2M SUPG tracer system by
GMRES(5) (FP64 and mixed), the 500k-node adjoint (gradient/forward cost,
finite-difference and tangent-linear checks), the 256-member ensemble
(systems/s, per-system latency) and the 5M-node prism systems; each with
the roofline figure of its dominant kernel.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|...]
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
    out = {"value": n / dt, "unit": "solves/s", "cores": 1, "kind": kind,
           "sample": f"{n} full FP64 SOLVEs (CG, PRECON 2, eps 1e-8) of the same "
                     f"{m.npoin}-node system from the same inputs, host-resident, "
                     f"{niter} iterations each, {dt:.2f} s in total",
           "iterations": niter}
    try:
        out["multicore"] = cpu_multicore(m, A, sysd, x, niter, budget_s=0.75 * budget_s)
    except Exception as e:  # reported, never silently substituted
        out["multicore"] = {"value": None, "error": repr(e)}
    return out, x


def cpu_multicore(m, A, sysd, x_ref, niter_ref, budget_s=9.0):
    """The same SOLVE on every host core this process may use: the oracle's
    OpenMP CG (node-centric gather mat-vec equal to the sequential product,
    per-thread dot products), checked against the single-core solve above
    by iteration count and solution before it is timed."""
    from oracle import oracle as O
    nt = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    plan = O.OmpPlan(m.ikle, m.npoin)

    def one():
        return O.omp_solve(plan, A, sysd["b"], sysd["x0"], precon=2, eps=1e-8, nitmax=5000,
                           nthreads=nt)
    x, info = one()
    rel = float(np.abs(x - x_ref).max() / np.abs(x_ref).max())
    ok = abs(info.niter - niter_ref) <= 1 and rel <= 1e-10
    n, t0 = 0, time.perf_counter()
    while True:
        one()
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            break
    return {"value": n / dt, "unit": "solves/s", "cores": nt, "kind": "port",
            "sample": f"{n} SOLVEs of the same system on {nt} OpenMP threads, "
                      f"{info.niter} iterations each, {dt:.2f} s in total",
            "iterations": info.niter, "max_rel_vs_single_core": rel,
            "agrees_with_single_core": bool(ok)}


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
    ap.add_argument("--workload", default="all",
                    choices=["all", "estuary", "channel", "prisms", "adjoint", "ensemble", "supg"],
                    help="all (default): the headline 2M-node estuary CG line with every other "
                         "configuration measured in the same run as sub-objects of 'configs'; "
                         "the others run one configuration alone")
    ap.add_argument("--adj-fd", type=int, default=20,
                    help="adjoint workload: nodes per parameter checked by central differences")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the path would "
              "shard by mesh domain; nothing here uses a second GPU); run with --gpus 1",
              file=sys.stderr)
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
    single = {"channel": lambda: bench_channel(args),
              "prisms": lambda: bench_prisms(args),
              "ensemble": lambda: bench_ensemble(args),
              "adjoint": lambda: bench_adjoint(args, n_fd=args.adj_fd),
              "supg": lambda: bench_supg(args)}
    if args.workload in single:
        print(json.dumps(single[args.workload]()), flush=True)
        return
    t_all = time.time()
    m, sysd = build_workload(args.npoin, args.seed, args.cache)
    line = bench_estuary(args, dev, m, sysd)
    if args.workload == "all":
        # every other configuration in the same run, each reported (or its
        # error), never silently left out
        configs = {}
        for name, fn in (("channel_20k", lambda: bench_channel(args)),
                         ("estuary_2m_segments", lambda: bench_segments(args, m, sysd)),
                         ("supg_2m", lambda: bench_supg(args, mesh=m)),
                         ("adjoint_500k", lambda: bench_adjoint(args, n_fd=args.adj_fd)),
                         ("ensemble_256", lambda: bench_ensemble(args)),
                         ("prisms_5m", lambda: bench_prisms(args))):
            t = time.time()
            try:
                configs[name] = fn()
            except Exception as e:  # reported in the line
                configs[name] = {"error": repr(e)}
            configs[name]["wall_s"] = round(time.time() - t, 1)
            log(f"{name} done in {time.time() - t:.1f}s")
            torch.cuda.empty_cache()
        line["configs"] = configs
    line["wall_s"] = round(time.time() - t_all, 1)
    print(json.dumps(line), flush=True)


def bench_estuary(args, dev, m, sysd):
    """The headline line: the 2M-node estuary wave-equation CG, FP64 and
    mixed, its roofline and the CPU baseline."""
    import torch
    from opentelemac_amd import bief
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
    # the step's torch copies run on the context's stream, so SOLVE needs no
    # event handshake with another stream
    ts = ctx.torch_stream()

    def step(precision):
        # SOLVE consumes A (D A D in place): it starts from the assembled copy,
        # as telemac2d copies AM1 before its SOLVE (propag.F:1643); the copy is
        # made by the preconditioning itself (restore=True)
        with torch.cuda.stream(ts):
            x.copy_(x00)
            b.copy_(b0)
            _, info = ctx.solve(x, b, slv=1, precon=2, eps=1e-8, nitmax=5000,
                                precision=precision, inner_eps=args.inner_eps, restore=True)
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
        # the same kernel with cold caches: each launch after a 512 MiB read pass
        # that evicts the last-level cache (its DRAM-bound time)
        results[name]["ms_dir_kernel_cold"] = ctx.time_cg_dir_cold(_flush_buffer(dev), prec,
                                                                   reps=20)
        its = info["niter"]
        results[name]["cg_iters_per_s"] = its * args.steps / el
        # per-solve latency: each solve on its own, synchronised (an idle gap
        # of 2 ms before each, so that every solve stands apart on a kernel
        # trace of this run)
        lat = []
        for _ in range(10):
            time.sleep(0.002)
            t1 = time.perf_counter()
            step(prec)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t1)
        results[name]["ms_per_solve_median"] = 1e3 * float(np.median(lat))
        results[name]["ms_per_solve_max"] = 1e3 * float(np.max(lat))
        log(name, json.dumps(results[name]))

    # per-iteration algorithmic bytes of CG (matrix and connectivity once, ten
    # vector touches per node) and the rate they imply
    ne, nprime = m.nelem, m.npoin
    b_it = {"fp64": ne * (3 * 8 + 12) + nprime * 10 * 8,
            "mixed": ne * (3 * 4 + 12) + nprime * 10 * 4}
    for name in ("fp64", "mixed"):
        results[name]["algorithmic_bytes_per_cg_iteration"] = b_it[name]
        results[name]["algorithmic_GBps_cg"] = results[name]["cg_iters_per_s"] * b_it[name] / 1e9

    # the host-pointer path (what the Fortran SOLVE does each call): upload
    # A, X, B from pageable host memory, solve in FP64, copy X and the
    # scaled B back
    da_h, xa_h = sysd["da"], np.ascontiguousarray(sysd["xa"])
    hl = []
    for k in range(4):
        # the caller's X and B (SOLVE overwrites both), prepared untimed
        xk, bk = sysd["x0"].copy(), sysd["b"].copy()
        t1 = time.perf_counter()
        ctx.set_matrix("Q", "S", da_h, xa_h)
        xh, hinfo = ctx.solve(xk, bk, slv=1, precon=2, eps=1e-8, nitmax=5000,
                              precision=bief.PREC_FP64)
        if k:
            hl.append(time.perf_counter() - t1)
    host_path = {"ms_per_solve_median": 1e3 * float(np.median(hl)), "iterations": hinfo["niter"],
                 "bytes_host_to_device": int(da_h.nbytes + xa_h.nbytes + 2 * sysd["b"].nbytes),
                 "bytes_device_to_host": int(2 * sysd["b"].nbytes),
                 "note": "bhip_matrix_set + bhip_solve from pageable numpy arrays, FP64"}
    log("host path", json.dumps(host_path))

    # true residual ||b - A x|| / ||b|| of both solutions (FP64 mat-vec)
    ctx.set_matrix("Q", "S", sysd["da"], sysd["xa"])
    for name in ("fp64", "mixed"):
        ax = ctx.matvec("X=AY", sols[name])
        results[name]["true_relres"] = float(np.linalg.norm(sysd["b"] - ax)
                                             / np.linalg.norm(sysd["b"]))
    rel_mixed_vs_fp64 = float(np.linalg.norm(sols["mixed"] - sols["fp64"])
                              / np.linalg.norm(sols["fp64"]))

    # algorithmic bytes per launch of the dominant kernels (int32
    # connectivity, 3 coefficients per element, node vectors):
    #  * FP64 CG direction + EBE mat-vec + D.AD: R, D_old read; D, AD written
    #  * mixed inner CG: the same kernel in FP32
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
    kname = "k_ebe<float,...,STAGE_CGDIR,EPI_CGAD,FIN_CG> (mixed inner CG direction)"
    # In the CG loop the FP32 kernel's ~110 MB per launch fits the 256 MB
    # last-level cache, so part of its traffic is served from there; the
    # cold-cache launches below bound what DRAM alone sustains.
    cold32 = _roof(bytes_f32, results["mixed"]["ms_dir_kernel_cold"])
    cold64 = _roof(bytes_f64, results["fp64"]["ms_dir_kernel_cold"])
    roofline = {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach32 / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname,
                "algorithmic_bytes_per_launch": bytes_f32,
                "avg_launch_ms": results["mixed"]["ms_dir_kernel"],
                "fp64_kernel": {"achieved": round(ach64, 1), "frac": round(ach64 / HBM_PEAK_GBS, 4),
                                "algorithmic_bytes_per_launch": bytes_f64,
                                "avg_launch_ms": results["fp64"]["ms_dir_kernel"]},
                "cold_cache": {"note": "each launch after a 512 MiB read pass (last-level "
                                       "cache evicted); only the kernel is timed",
                               "fp32": {k: cold32[k] for k in ("avg_launch_ms", "achieved",
                                                                "frac")},
                               "fp64": {k: cold64[k] for k in ("avg_launch_ms", "achieved",
                                                                "frac")}},
                "copy_same_bytes": _copy_rate(dev, bytes_f32),
                "copy_1gib": _copy_rate(dev, 1 << 30, reps=10)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu, xcpu = cpu_baseline(m, sysd)
            cpu["gpu_fp64_rel_l2_vs_cpu"] = float(np.linalg.norm(sols["fp64"] - xcpu)
                                                  / np.linalg.norm(xcpu))
            cpu["gpu_mixed_rel_l2_vs_cpu"] = float(np.linalg.norm(sols["mixed"] - xcpu)
                                                   / np.linalg.norm(xcpu))
            cpu["cg_iters_per_s"] = cpu["value"] * cpu["iterations"]
            cpu["algorithmic_GBps_cg"] = cpu["cg_iters_per_s"] * b_it["fp64"] / 1e9
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
        "host_pointer_path": host_path,
        "layout": st,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    ctx.close()
    return line




def _copy_rate(dev, nbytes, reps=50):
    """A plain device copy moving nbytes in all (nbytes / 2 read, nbytes / 2
    written), timed with events over reps launches: the practical rate of
    that much traffic on this GPU, beside the kernel's."""
    import torch
    src = torch.ones(max(1, nbytes // 16), dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    moved = 2 * src.numel() * 8
    return {"bytes": moved, "avg_launch_ms": ms, "GBps": round(moved / (ms * 1e-3) / 1e9, 1)}


_FLUSH = {}


def _flush_buffer(dev):
    """512 MiB of device memory, twice the last-level cache, for the
    cold-cache kernel timings."""
    import torch
    if "buf" not in _FLUSH:
        _FLUSH["buf"] = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
    return _FLUSH["buf"]


_MESHES = {}


def estuary(npoin, seed):
    """The estuary mesh generator's mesh, built once per run and size."""
    from opentelemac_amd import mesh as M
    key = (int(npoin), int(seed))
    if key not in _MESHES:
        _MESHES[key] = M.estuary_mesh(int(npoin), seed=int(seed))
    return _MESHES[key]


def _roof(nbytes, ms):
    """Algorithmic rate of a kernel moving nbytes per launch in ms, and its
    fraction of the 8 TB/s HBM peak."""
    gbps = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "algorithmic_bytes_per_launch": int(nbytes), "avg_launch_ms": ms,
            "achieved": round(gbps, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbps / HBM_PEAK_GBS, 4)}


def bench_prisms(args):
    """5,014,910-node prism mesh: a 501,491-node estuary base (the 2M
    generator at 500k points) extruded to 10 sigma planes over the seeded
    depth; 8,996,184 prisms. Two systems, assembled on the device:
      * diffusion: MT01PP / dt + MT02PP with nu = 10^(-4 + 2 sigma) m^2/s in
        all three directions (1e-4 at the bed to 1e-2 at the surface), dt = 2 s;
      * Poisson: MT02PP with unit coefficients, zero Dirichlet values on the
        surface plane (DIRICH).
    Each step restores the assembled matrix and solves with CG + diagonal
    preconditioning to eps = 1e-8, FP64 and mixed."""
    import torch
    from opentelemac_amd import bief, mesh as M
    dev = torch.device("cuda:0")
    t = time.time()
    base = estuary(500_000, args.seed)
    depth = M.smooth_random_depth(base, seed=args.seed)
    nplan = 10
    p = M.extrude_prisms(base, nplan, -depth, np.zeros(base.npoin))
    eb = np.arange(p.nelem) % base.nelem
    xe, ye = base.xel_yel()
    xel, yel = np.ascontiguousarray(xe[:, eb]), np.ascontiguousarray(ye[:, eb])
    s = base.surfac()[eb]
    sig = np.repeat(np.linspace(0.0, 1.0, nplan), base.npoin)
    nu = 10.0 ** (-4.0 + 2.0 * sig)
    log(f"prisms {p.npoin} nodes {p.nelem} elements, built in {time.time() - t:.1f}s")
    t = time.time()
    bn, table = args.block_nodes, None
    if bn == 0:
        # node-block size by the autotuner, on the diffusion system
        from opentelemac_amd import tune

        def load(c):
            c.matrix_mt01pp(1.0 / 2.0, s, p.z)
            c.matrix_mt02pp(1.0, xel, yel, s, p.z, nu, nu, nu, base.npoin, accumulate=True)
            c.solve(np.zeros(p.npoin), np.ones(p.npoin), slv=1, precon=2, eps=1e-8,
                    nitmax=20000)

        bn, table = tune.block_nodes(lambda b: bief.Context(p, block_nodes=b), load,
                                     candidates=(1024, 2048, 2560), reps=10)
        log(f"autotuned block size {bn}: {table}")
    ctx = bief.Context(p, block_nodes=bn)
    st = ctx.stats()
    log(f"context in {time.time() - t:.1f}s: {st}")
    rng = np.random.default_rng(args.seed)
    out = {"metric": "EBE-CG solves/s, 5M-node prisms (MT01PP/dt + MT02PP; MT02PP Poisson)",
           "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "data": "synthetic",
           "config": {"workload": "sigma-mesh prisms, 10 planes x 501,491 nodes",
                      "npoin": p.npoin, "nelem": p.nelem, "dt_s": 2.0, "eps": 1e-8},
           "layout": st, "block_nodes": bn, "autotune_ms_per_iteration": table, "systems": {}}
    top = np.zeros(p.npoin, dtype=np.int32)
    top[(nplan - 1) * base.npoin:] = 5
    # the mesh geometry stays on the device; each assembly uploads the node
    # elevations and viscosities only
    ctx.set_geometry(xel, yel, s)
    G = (None, None, None)
    for name in ("diffusion", "poisson"):
        asm = []
        asm_dev = []
        if name == "diffusion":
            for _ in range(3):
                t1 = time.perf_counter()
                ctx.matrix_mt01pp(1.0 / 2.0, None, p.z)
                ctx.matrix_mt02pp(1.0, *G, p.z, nu, nu, nu, base.npoin, accumulate=True)
                asm.append(time.perf_counter() - t1)
            # the same with the elevations and viscosities resident on the device
            zd = torch.from_numpy(np.ascontiguousarray(p.z, dtype=np.float64)).to(dev)
            nud = torch.from_numpy(np.ascontiguousarray(nu, dtype=np.float64)).to(dev)
            for _ in range(6):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                ctx.matrix_mt01pp(1.0 / 2.0, None, zd)
                ctx.matrix_mt02pp(1.0, *G, zd, nud, nud, nud, base.npoin, accumulate=True)
                torch.cuda.synchronize()
                asm_dev.append(time.perf_counter() - t1)
            b0 = rng.normal(size=p.npoin)
        else:
            one = np.ones(p.npoin)
            ctx.matrix_mt02pp(1.0, *G, p.z, one, one, one, base.npoin)
            _, b0 = ctx.dirich(np.zeros(p.npoin), rng.normal(size=p.npoin), np.zeros(p.npoin),
                               top)
        ctx.save_matrix()
        bd = torch.from_numpy(b0).to(dev)
        x0 = torch.zeros(p.npoin, dtype=torch.float64, device=dev)
        x = torch.empty_like(x0)
        b = torch.empty_like(bd)
        res = {}
        for pname, prec in (("fp64", bief.PREC_FP64), ("mixed", bief.PREC_MIXED)):
            def step():
                x.copy_(x0)
                b.copy_(bd)
                return ctx.solve(x, b, slv=1, precon=2, eps=1e-8, nitmax=20000, precision=prec,
                                 inner_eps=args.inner_eps, restore=True)[1]
            for _ in range(args.warmup):
                step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            infos = [step() for _ in range(args.steps)]
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            ms_dir, ms_upd = ctx.time_cg_kernels(prec, reps=20)
            ms_cold = ctx.time_cg_dir_cold(_flush_buffer(dev), prec, reps=10)
            res[pname] = {"ms_per_solve": 1e3 * el / args.steps,
                          "solves_per_s": args.steps / el, "iterations": infos[-1]["niter"],
                          "outer": infos[-1]["outer_iters"],
                          "converged": bool(infos[-1]["converged"]),
                          "ms_dir_kernel": ms_dir, "ms_upd_kernel": ms_upd,
                          "ms_dir_kernel_cold": ms_cold}
            log(name, pname, json.dumps(res[pname]))
        # algorithmic bytes of the CG direction + mat-vec kernel: 15
        # coefficients and 6 int32 node ids per prism, 4 node vectors
        b64 = p.nelem * (15 * 8 + 24) + p.npoin * 4 * 8
        b32 = p.nelem * (15 * 4 + 24) + p.npoin * 4 * 4
        res["fp64_kernel_GBps"] = round(b64 / (res["fp64"]["ms_dir_kernel"] * 1e-3) / 1e9, 1)
        res["roofline"] = _roof(b32, res["mixed"]["ms_dir_kernel"])
        res["roofline"]["kernel"] = "k_ebe<float,...,6,...,STAGE_CGDIR,EPI_CGAD,FIN_CG>"
        res["roofline_fp64"] = _roof(b64, res["fp64"]["ms_dir_kernel"])
        res["roofline"]["cold_cache_frac"] = _roof(b32, res["mixed"]["ms_dir_kernel_cold"])["frac"]
        res["roofline_fp64"]["cold_cache_frac"] = _roof(b64,
                                                       res["fp64"]["ms_dir_kernel_cold"])["frac"]
        res["mixed_speedup_vs_fp64"] = round(res["fp64"]["ms_per_solve"] /
                                             res["mixed"]["ms_per_solve"], 3)
        if asm:
            res["ms_assembly_device_median"] = 1e3 * float(np.median(asm_dev))
            res["ms_assembly_host_arrays_median"] = 1e3 * float(np.median(asm))
            res["assembly"] = ("MT01PP / dt + MT02PP ('M=M+N') on the device from the geometry "
                               "kept there; 'device': node elevations and viscosities resident "
                               "on the GPU, 'host_arrays': uploaded from host arrays")
        out["systems"][name] = res
    ctx.close()
    return out


def bench_adjoint(args, npoin=500_000, nsteps=50, checkpoint=10, n_fd=20):
    """Gradient of a gauge misfit with respect to the per-node Strickler
    coefficient and the bathymetry through the checkpointed adjoint of the
    time-stepped wave equation: a 500k-node estuary mesh (the 2M generator at
    500k points), 50 steps of dt = 2 s, Strickler uniform in [20, 60]
    (seeded), misfit at 200 random gauges against a run with the friction
    perturbed by a uniform factor in [0.8, 1.2], checkpoints every 10 steps,
    mixed-precision solves (FP32 EBE CG inside FP64 refinement) to 1e-10.

    Timed on the device model (HIP step kernels and their hand-written
    adjoint, DeviceWaveModel): forward run, and forward + reverse sweeps for
    the gradient, with and without checkpoint recomputation. Beside it the
    torch model of the same equations (WaveModel: torch element-wise work
    around the HIP mat-vec and solve, autograd for the rest). Checks: central
    differences at 20 random nodes per parameter on the device model, and
    the device gradient against an FP64 tangent-linear (forward-mode) run of
    the torch model."""
    import torch
    import torch.autograd.forward_ad as fwAD
    from opentelemac_amd import bief, mesh as M, system as S, wave as W
    t = time.time()
    m = estuary(npoin, args.seed)
    depth = M.smooth_random_depth(m, seed=args.seed)
    sea = S.sea_boundary(m)
    log(f"adjoint mesh {m.npoin} nodes, built in {time.time() - t:.1f}s")
    kw = dict(dt=2.0, amp=0.2, period=600.0, precision=bief.PREC_MIXED, eps=1e-10)
    # node-block size by the autotuner, on the wave system of this mesh (the
    # mixed inner CG's kernels): 500k nodes make too few 1024-node blocks to
    # fill 256 compute units
    from opentelemac_amd import tune
    wsys = S.wave_system(m, depth, dt=2.0, theta=1.0, inflow=sea)

    def load(c):
        c.set_matrix("Q", "S", wsys["da"], wsys["xa"])
        c.solve(wsys["x0"].copy(), wsys["b"].copy(), slv=1, precon=2, eps=1e-10, nitmax=5000,
                precision=bief.PREC_MIXED)
    bn, table = tune.block_nodes(lambda b: bief.Context(m, block_nodes=b), load,
                                 candidates=(256, 512, 1024), precision=bief.PREC_MIXED,
                                 reps=20)
    log(f"adjoint block size {bn}: {table}")
    model = W.DeviceWaveModel(m, sea, block_nodes=bn, **kw)
    tmodel = W.WaveModel(m, sea, **kw)
    dev = model.dev
    rng = np.random.default_rng(args.seed + 2)
    k = torch.tensor(rng.uniform(20, 60, m.npoin), dtype=torch.float64, device=dev)
    zb = torch.tensor(-depth, dtype=torch.float64, device=dev)
    gauges = torch.tensor(rng.choice(m.npoin, 200, replace=False), device=dev)
    obs = model.run(k * torch.tensor(rng.uniform(0.8, 1.2, m.npoin), device=dev), zb, nsteps,
                    gauges)

    def timed(fn):
        fn()  # warm-up (graphs, allocator)
        torch.cuda.synchronize()
        model.n_solves = model.n_iters = 0
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    def forward():
        return float(model.run(k, zb, nsteps, gauges, obs))

    t_fwd, j0 = timed(forward)
    fwd_solves, fwd_iters = model.n_solves, model.n_iters
    t_grad, (j, gk, gz) = timed(
        lambda: model.cost_and_gradient(k, zb, nsteps, gauges, obs, checkpoint))
    grad_solves = model.n_solves
    t_grad_all, _ = timed(lambda: model.cost_and_gradient(k, zb, nsteps, gauges, obs, nsteps))
    log(f"device model: forward {t_fwd:.3f}s, gradient {t_grad:.3f}s (checkpoints every "
        f"{checkpoint}), {t_grad_all:.3f}s (all states kept)")
    # the torch model of the same equations
    with torch.no_grad():
        tmodel.run(k, zb, nsteps, gauges, obs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        tmodel.run(k, zb, nsteps, gauges, obs)
    torch.cuda.synchronize()
    tt_fwd = time.perf_counter() - t0
    W.cost_and_gradient(tmodel, k, zb, nsteps, gauges, obs, checkpoint)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    W.cost_and_gradient(tmodel, k, zb, nsteps, gauges, obs, checkpoint)
    torch.cuda.synchronize()
    tt_grad = time.perf_counter() - t0
    # central differences at n_fd random nodes for each parameter
    fd_err = {}
    for name, p, g in (("strickler", k, gk), ("bathymetry", zb, gz)):
        scale = float(g.abs().max())
        worst = 0.0
        for i in rng.choice(m.npoin, n_fd, replace=False):
            h = 1e-4 * max(1.0, abs(float(p[i])))
            saved = float(p[i])
            p[i] = saved + h
            jp = forward()
            p[i] = saved - h
            jm = forward()
            p[i] = saved
            fd = (jp - jm) / (2 * h)
            worst = max(worst, abs(float(g[i]) - fd) / scale)
        fd_err[name] = worst
    # the tangent-linear reference: forward mode through the torch model in
    # FP64 (FP64 CG to 1e-12)
    tl_model = W.WaveModel(m, sea, dt=2.0, amp=0.2, period=600.0, precision=bief.PREC_FP64,
                           eps=1e-12)
    dk = torch.tensor(rng.normal(size=m.npoin), dtype=torch.float64, device=dev)
    dz = torch.tensor(rng.normal(size=m.npoin) * 0.01, dtype=torch.float64, device=dev)
    with fwAD.dual_level():
        jd = tl_model.run(fwAD.make_dual(k, dk), fwAD.make_dual(zb, dz), nsteps, gauges, obs)
        tl = float(fwAD.unpack_dual(jd).tangent)
    adj = float((gk * dk).sum() + (gz * dz).sum())
    # the dominant kernel: the inner CG's direction + EBE product (FP32)
    ms_dir, _ = model.ctx.time_cg_kernels(bief.PREC_MIXED, reps=50)
    roof = _roof(m.nelem * (3 * 4 + 12) + m.npoin * 4 * 4, ms_dir)
    roof["kernel"] = "k_ebe<float,...,STAGE_CGDIR,EPI_CGAD,FIN_CG> (mixed inner CG direction)"
    return {"roofline": roof, "block_nodes": bn, "autotune_ms_per_iteration": table,
            "workload": "wave-equation adjoint, 500k nodes, 50 steps, checkpoints every 10",
            "npoin": m.npoin, "nelem": m.nelem, "steps": nsteps, "cost": j0,
            "forward_s": t_fwd, "gradient_s": t_grad,
            "adjoint_over_forward": round(t_grad / t_fwd, 3),
            "gradient_s_all_states_kept": t_grad_all,
            "adjoint_over_forward_all_states_kept": round(t_grad_all / t_fwd, 3),
            "forward_solves": fwd_solves, "forward_cg_iterations": fwd_iters,
            "gradient_solves": grad_solves,
            "torch_model": {"forward_s": tt_fwd, "gradient_s": tt_grad,
                            "adjoint_over_forward": round(tt_grad / tt_fwd, 3)},
            "fd_max_rel_error": fd_err, "fd_nodes_per_parameter": n_fd,
            "tangent_linear_rel_error": abs(adj - tl) / max(abs(tl), 1e-300),
            "solver": "mixed (FP32 EBE CG, FP64 refinement), eps 1e-10"}


def bench_supg(args, dt=10.0, nu=0.1, mesh=None):
    """The non-symmetric system of the 2M-node estuary: CVDFTR's SUPG tracer
    system (classical SUPG, implicit P1 diffusion nu = 0.1 m^2/s, dt = 10 s)
    with a seeded divergence-free velocity field of at most 2 m/s (a
    streamfunction of 16 Gaussians), a zero Dirichlet value on the sea
    boundary and a tracer of Gaussian patches, assembled on the device
    (MT01AA, MT05AA, MT03AA with KSUPG in place, MT02AA, OM) and solved by
    GMRES(5) with PRECON 2 to eps 1e-8, FP64 and mixed (FP32 Arnoldi inside
    FP64 restarts). A step restores the assembled matrix and solves."""
    import torch
    from opentelemac_amd import bief, tracer
    t = time.time()
    m = mesh if mesh is not None else estuary(args.npoin, args.seed)
    u, v, fn, sea = tracer.estuary_flow(m, seed=args.seed)
    log(f"supg mesh {m.npoin} nodes, ready in {time.time() - t:.1f}s")
    ctx = bief.Context(m)
    t = time.time()
    sm = tracer.supg_system(ctx, m, u, v, np.full(m.npoin, nu), dt, fn)
    t_first = time.time() - t
    # the assembly of a later time step: the geometry stays on the device,
    # the nodal velocities, viscosity and tracer come from the host
    asm = []
    for _ in range(3):
        t1 = time.perf_counter()
        sm = tracer.supg_system(ctx, m, u, v, np.full(m.npoin, nu), dt, fn)
        asm.append(time.perf_counter() - t1)
    # the same assembly with every operand resident on the device
    dev = torch.device("cuda:0")
    ud, vd, fnd = (torch.from_numpy(a).to(dev) for a in (u, v, fn))
    nud = torch.full((m.npoin,), nu, dtype=torch.float64, device=dev)
    asm_dev = []
    for _ in range(6):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        smd = tracer.supg_system(ctx, m, ud, vd, nud, dt, fnd)
        torch.cuda.synchronize()
        asm_dev.append(time.perf_counter() - t1)
    dev_host_rel = float((smd.cpu().numpy() - sm).__abs__().max() / np.abs(sm).max())
    limdir = torch.full((m.npoin,), 2, dtype=torch.int32, device=dev)
    limdir[torch.from_numpy(sea).to(dev)] = 5
    x0, b0 = ctx.dirich(fnd.clone(), smd, torch.zeros_like(fnd), limdir)
    ctx.save_matrix()
    xx, bb = torch.empty_like(x0), torch.empty_like(b0)
    out = {"workload": "SUPG tracer system (CVDFTR), 2M-node estuary, GMRES(5) + PRECON 2",
           "npoin": m.npoin, "nelem": m.nelem, "dt_s": dt, "nu": nu, "eps": 1e-8,
           "umax": 2.0,
           "ms_assembly_device_median": 1e3 * float(np.median(asm_dev)),
           "ms_assembly_host_arrays_median": 1e3 * float(np.median(asm)),
           "ms_first_assembly_with_geometry_upload": 1e3 * t_first,
           "assembly_device_vs_host_max_rel": dev_host_rel,
           "assembly": "tracer.supg_system: MT05AA, MT03AA + KSUPG, MT01AA, OM, MT02AA on the "
                       "device, 2 mat-vecs; 'device' keeps u, v, nu, the tracer and SM on the "
                       "GPU, 'host_arrays' uploads them and downloads SM"}
    sols = {}
    for name, prec in (("fp64", bief.PREC_FP64), ("mixed", bief.PREC_MIXED)):
        def step():
            xx.copy_(x0)
            bb.copy_(b0)
            return ctx.solve(xx, bb, slv=7, precon=2, krylov=5, eps=1e-8, nitmax=2000,
                             precision=prec, restore=True)[1]

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        lat = []
        for _ in range(max(args.steps, 5)):
            t0 = time.perf_counter()
            info = step()
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        sols[name] = xx.cpu().numpy()
        ms = float(np.median(lat)) * 1e3
        out[name] = {"ms_per_solve": ms, "solves_per_s": 1e3 / ms,
                     "gmres_restarts": info["niter"], "ebe_products": info["inner_iters"] or
                     info["niter"] * 5, "converged": bool(info["converged"]),
                     "relres": info["relres"],
                     "ms_ebe_product": ctx.time_matvec(prec, reps=20)}
        log("supg", name, json.dumps(out[name]))
    out["mixed_speedup_vs_fp64"] = round(out["fp64"]["ms_per_solve"] /
                                         out["mixed"]["ms_per_solve"], 3)
    out["mixed_vs_fp64_rel_l2"] = float(np.linalg.norm(sols["mixed"] - sols["fp64"]) /
                                        np.linalg.norm(sols["fp64"]))
    # the Arnoldi steps' EBE product on the 'Q' (non-symmetric) matrix after
    # PRECON 2 (unit diagonal): 6 coefficients and 3 int32 ids per element,
    # Y read and X written
    ne, npo = m.nelem, m.npoin
    out["roofline"] = _roof(ne * (6 * 4 + 12) + npo * 2 * 4, out["mixed"]["ms_ebe_product"])
    out["roofline"]["kernel"] = "k_ebe<float,...,SYMQ=false,...,STAGE_PLAIN,EPI_OP> (FP32 'Q')"
    out["roofline_fp64"] = _roof(ne * (6 * 8 + 12) + npo * 2 * 8, out["fp64"]["ms_ebe_product"])
    ctx.close()
    return out


def bench_ensemble(args, nsys=256):
    """The batched ensemble: 256 unstructured estuary meshes (the estuary
    generator, each member its own seed, outline, points and depth field),
    5k to 200k nodes drawn log-uniformly (seeded), each with a wave-equation
    step (dt = 1 s, theta = 1, a 0.2 m tide on its sea boundary). All
    solved as one ensemble -- one context, every kernel launch covering every
    mesh, each mesh with its own scalars, stop tests and refinement sweeps --
    by CG + PRECON 2 to 1e-8, mixed precision (the configuration) and FP64,
    the node-block size chosen by the autotuner on the first 32 members. A
    step restores the assembled matrices and solves all members. Reported:
    systems/s, and per-system completion latency (device clock from the
    solve's set-up kernel to the kernel that stopped the system) p50/p99 over
    all members of the timed steps. Beside it: the same meshes one context
    each, solved one after the other."""
    import torch
    from opentelemac_amd import bief, system as S, tune
    sizes = S.ensemble_sizes(nsys, seed=args.seed)
    t = time.time()
    meshes, systems = S.estuary_ensemble(sizes, seed=args.seed)
    total = sum(m.npoin for m in meshes)
    log(f"ensemble of {nsys} meshes, {total} nodes, built in {time.time() - t:.1f}s")
    cat = lambda key, sel=None: np.ascontiguousarray(np.concatenate(  # noqa: E731
        [s[key] for s in (systems if sel is None else systems[:sel])], axis=-1))

    def make(bn, sel=32):
        return bief.Ensemble(meshes[:sel], block_nodes=bn)

    def load(ctx, sel=32):
        ctx.set_matrix("Q", "S", cat("da", sel), cat("xa", sel))
        ctx.solve(cat("x0", sel), cat("b", sel).copy(), slv=1, precon=2, eps=1e-8, nitmax=5000,
                  precision=bief.PREC_MIXED)

    best, table = tune.block_nodes(make, load, candidates=(512, 1024, 2048),
                                   precision=bief.PREC_MIXED)
    log(f"block size {best}: {table}")
    t = time.time()
    ens = bief.Ensemble(meshes, block_nodes=best)
    log(f"ensemble context in {time.time() - t:.1f}s")
    dev = torch.device("cuda:0")
    ens.set_matrix("Q", "S", torch.from_numpy(cat("da")).to(dev),
                   torch.from_numpy(cat("xa")).to(dev))
    ens.save_matrix()
    x0 = torch.from_numpy(cat("x0")).to(dev)
    b0 = torch.from_numpy(cat("b")).to(dev)
    x, b = torch.empty_like(x0), torch.empty_like(b0)
    ne = sum(m.nelem for m in meshes)
    out = {"workload": f"ensemble of {nsys} unstructured estuary meshes, 5k-200k nodes "
                       "(log-uniform)",
           "systems": nsys, "total_nodes": int(total), "total_elements": int(ne),
           "block_nodes": best, "autotune_ms_per_iteration": table,
           "workgroup_threads": 256, "eps": 1e-8}
    for name, prec in (("mixed", bief.PREC_MIXED), ("fp64", bief.PREC_FP64)):
        def step():
            x.copy_(x0)
            b.copy_(b0)
            return ens.solve(x, b, slv=1, precon=2, eps=1e-8, nitmax=5000, precision=prec,
                             restore=True)[1]

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        steps, lat_sys = [], []
        for _ in range(max(args.steps, 5)):
            t0 = time.perf_counter()
            info = step()
            torch.cuda.synchronize()
            steps.append(time.perf_counter() - t0)
            lat_sys.append(ens.latencies_ms())
        infos = ens.infos()
        lat_sys = np.concatenate(lat_sys)
        st_ms = np.array(steps) * 1e3
        ms_dir, ms_upd = ens.time_cg_kernels(prec, reps=20)
        sv = 4 if prec == bief.PREC_MIXED else 8
        roof = _roof(ne * (3 * sv + 12) + total * 4 * sv, ms_dir)
        roof["kernel"] = ("k_ebe<%s,...,STAGE_CGDIR,EPI_CGAD,FIN_CG> over all members"
                          % ("float" if sv == 4 else "double"))
        out[name] = {"ms_per_ensemble_step_median": float(np.median(st_ms)),
                     "systems_per_s": float(nsys * 1e3 / st_ms.mean()),
                     "per_system_latency_ms_p50": float(np.percentile(lat_sys, 50)),
                     "per_system_latency_ms_p99": float(np.percentile(lat_sys, 99)),
                     "per_system_latency_ms_max": float(lat_sys.max()),
                     "iterations_max": int(info["niter"]),
                     "iterations_median": float(np.median([i["niter"] for i in infos])),
                     "outer_max": int(info["outer_iters"]),
                     "all_converged": bool(all(i["converged"] for i in infos)),
                     "ms_dir_kernel": ms_dir, "ms_upd_kernel": ms_upd, "roofline": roof}
        log("ensemble", name, json.dumps(out[name]))
    out["roofline"] = out["mixed"]["roofline"]
    # the same systems, one context each, one after the other (mixed)
    ctxs = [bief.Context(m, block_nodes=best) for m in meshes]
    dsys = [(torch.from_numpy(sd["da"]).to(dev),
             torch.from_numpy(np.ascontiguousarray(sd["xa"])).to(dev),
             torch.from_numpy(sd["x0"]).to(dev), torch.from_numpy(sd["b"]).to(dev))
            for sd in systems]
    seq = []
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for ctx, (da, xa, xx, bb) in zip(ctxs, dsys):
            ctx.set_matrix("Q", "S", da, xa)
            ctx.solve(xx.clone(), bb.clone(), slv=1, precon=2, eps=1e-8, nitmax=5000,
                      precision=bief.PREC_MIXED)
        torch.cuda.synchronize()
        seq.append(time.perf_counter() - t0)
    for ctx in ctxs:
        ctx.close()
    ens.close()
    out["sequential_ms_per_round_mixed"] = float(min(seq) * 1e3)
    out["speedup_vs_sequential"] = float(min(seq) * 1e3 / out["mixed"]
                                         ["ms_per_ensemble_step_median"])
    return out


def bench_segments(args, m, sysd):
    """The headline system in segment storage (MATRIX STORAGE = 3, the
    default of telemac2d): the element terms assembled per edge as ASSEX3
    does, 6.0 M segments run as two-node elements of the same kernels; FP64
    and mixed CG + PRECON 2 to 1e-8, same step as the headline line."""
    import torch
    from opentelemac_amd import bief, mesh as M, system as S
    dev = torch.device("cuda:0")
    t = time.time()
    sm, es = M.segment_mesh(m)
    xs = np.ascontiguousarray(S.ebe_to_segments(sysd["xa"], es, m.ikle, sm.ikle, "S"))
    ctx = bief.Context(sm, block_nodes=args.block_nodes)
    log(f"segments {sm.nelem} in {time.time() - t:.1f}s: {ctx.stats()}")
    ctx.set_matrix("Q", "S", torch.from_numpy(sysd["da"]).to(dev), torch.from_numpy(xs).to(dev))
    ctx.save_matrix()
    x00 = torch.from_numpy(sysd["x0"]).to(dev)
    b0 = torch.from_numpy(sysd["b"]).to(dev)
    x, b = torch.empty_like(x00), torch.empty_like(b0)
    ts = ctx.torch_stream()
    out = {"workload": "the 2M-node estuary system in segment storage (MATRIX STORAGE = 3)",
           "npoin": m.npoin, "nseg": sm.nelem, "layout": ctx.stats()}
    nseg, npo = sm.nelem, m.npoin
    alg = {"fp64": nseg * (8 + 8) + npo * 4 * 8, "mixed": nseg * (4 + 8) + npo * 4 * 4}
    for name, prec in (("fp64", bief.PREC_FP64), ("mixed", bief.PREC_MIXED)):
        def step():
            with torch.cuda.stream(ts):
                x.copy_(x00)
                b.copy_(b0)
                return ctx.solve(x, b, slv=1, precon=2, eps=1e-8, nitmax=5000, precision=prec,
                                 inner_eps=args.inner_eps, restore=True)[1]
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = [step() for _ in range(args.steps)]
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms_dir, ms_upd = ctx.time_cg_kernels(prec, reps=50)
        out[name] = {"ms_per_solve": 1e3 * el / args.steps, "solves_per_s": args.steps / el,
                     "iterations": infos[-1]["niter"], "outer": infos[-1]["outer_iters"],
                     "relres": infos[-1]["relres"], "converged": bool(infos[-1]["converged"]),
                     "ms_dir_kernel": ms_dir, "ms_upd_kernel": ms_upd,
                     "roofline": _roof(alg[name], ms_dir)}
        log("segments", name, json.dumps(out[name]))
    out["mixed_speedup_vs_fp64"] = round(out["fp64"]["ms_per_solve"] /
                                         out["mixed"]["ms_per_solve"], 3)
    ctx.close()
    return out


def bench_channel(args):
    """The CPU-runnable configuration: the 2000 m x 200 m channel (400 x 50
    cells, 20,451 nodes, 40,000 triangles), depth 1-5 m with a 1.5 m bump,
    one wave-equation step (dt = 1 s) with a 0.2 m tide at the inflow, CG +
    PRECON 2 to 1e-8: GPU FP64 and mixed, and the reference's own SOLVE on
    one host core, on the same system."""
    import torch
    from opentelemac_amd import bief, mesh as M, system as S
    from oracle import oracle as O
    m = M.channel_mesh()
    sysd = S.wave_system(m, M.channel_depth(m), dt=1.0)
    dev = torch.device("cuda:0")
    ctx = bief.Context(m)
    ctx.set_matrix("Q", "S", torch.from_numpy(sysd["da"]).to(dev),
                   torch.from_numpy(np.ascontiguousarray(sysd["xa"])).to(dev))
    ctx.save_matrix()
    x0 = torch.from_numpy(sysd["x0"]).to(dev)
    b0 = torch.from_numpy(sysd["b"]).to(dev)
    x, b = torch.empty_like(x0), torch.empty_like(b0)
    out = {"workload": "structured channel, 20,451 nodes, wave-equation CG",
           "npoin": m.npoin, "nelem": m.nelem}
    sols = {}
    for name, prec in (("fp64", bief.PREC_FP64), ("mixed", bief.PREC_MIXED)):
        def step():
            x.copy_(x0)
            b.copy_(b0)
            return ctx.solve(x, b, slv=1, precon=2, eps=1e-8, nitmax=5000, precision=prec,
                             restore=True)[1]
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        lat = []
        for _ in range(20):
            t0 = time.perf_counter()
            info = step()
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        sols[name] = x.cpu().numpy()
        ms = 1e3 * float(np.median(lat))
        out[name] = {"ms_per_solve": ms, "solves_per_s": 1e3 / ms, "iterations": info["niter"],
                     "relres": info["relres"], "converged": bool(info["converged"])}
    ctx.close()
    A = O.Matrix(sysd["da"], sysd["xa"], "Q", "S")
    if O.have_ref():
        rm = O.RefMesh(m.ikle, m.npoin, m.nelem)
        one = lambda: O.ref_solve(rm, A, sysd["b"], sysd["x0"], slv=1, precon=2,  # noqa: E731
                                  eps=1e-8, nitmax=5000)
        kind = "reference"
    else:
        one = lambda: O.solve(A, m.ikle, sysd["b"], sysd["x0"], slv=1, precon=2,  # noqa: E731
                              eps=1e-8, nitmax=5000)
        kind = "port"
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        xr = one()[0]
        n += 1
    dt = time.perf_counter() - t0
    out["cpu_1core"] = {"kind": kind, "solves_per_s": n / dt, "sample": f"{n} solves"}
    out["fp64_rel_l2_vs_cpu"] = float(np.linalg.norm(sols["fp64"] - xr) / np.linalg.norm(xr))
    out["mixed_rel_l2_vs_cpu"] = float(np.linalg.norm(sols["mixed"] - xr) / np.linalg.norm(xr))
    return out


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
