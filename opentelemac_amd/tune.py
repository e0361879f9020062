"""Launch-shape autotuning for the EBE CG kernels.

The node-block size sets the work of one workgroup, its LDS (owned + halo
nodes), the share of duplicated border elements and, for ensembles of small
meshes, how evenly the blocks spread over the 256 compute units. The best
size depends on the mesh; it is measured, not guessed: for each candidate a
context is built, the system loaded, and the two kernels of one CG iteration
timed with HIP events (bhip_time_cg_kernels) after one solve. The result is
cached per mesh signature for the process.
"""
from __future__ import annotations

from . import bief

_cache: dict = {}


def block_nodes(make_context, load, candidates=(512, 1024, 2048), precision=bief.PREC_FP64,
                reps=20, key=None):
    """Returns (best block size, {size: ms per CG iteration}).

    make_context(block_nodes) -> Context; load(ctx) loads the system and runs
    one solve of the given precision (so the mixed path's FP32 copies exist).
    Candidates the layout rejects (LDS, 16-bit local ids) are skipped."""
    if key is not None and key in _cache:
        return _cache[key]
    timings = {}
    for bn in candidates:
        try:
            ctx = make_context(bn)
        except bief.BiefError:
            continue
        load(ctx)
        ms_dir, ms_upd = ctx.time_cg_kernels(precision, reps=reps)
        timings[bn] = ms_dir + ms_upd
        ctx.close()
    if not timings:
        raise bief.BiefError("no block size fits this mesh")
    best = min(timings, key=timings.get)
    if key is not None:
        _cache[key] = (best, timings)
    return best, timings
