"""Autotuning of the BVH leaf size around the schedule autotuner.

The leaf size (particles per leaf) is a property of the scene: each
candidate rebuilds the bounds and the BVH (Scene.prepare) and runs the
schedule autotuner (aarbf_autotune: step budget, occupancy, graph rounds,
gather walk budget, step-kernel live-lane threshold) on the views, each leaf size from the same starting
schedule; the fastest (leaf size, schedule) is returned with its scene,
and the context keeps that schedule.
"""
from __future__ import annotations

from typing import Callable, Sequence


def tune_leaf_size(ctx, xyzw, colormap, radiusmap, densitymap, rbf_radius: float, make_requests: Callable,
                   radii=None, leaf_sizes: Sequence[int] = (1, 2, 4), branching: int = 8, reps: int = 1):
    """make_requests(scene) -> list of (rg, pc, width, height, spp) for that
    scene.  Returns (best scene, {"particles_per_leaf", schedule fields,
    "best_ms", "per_leaf_size": {ppl: best_ms}})."""
    from . import Scene
    best = None
    per = {}
    settings = {}
    for ppl in leaf_sizes:
        sc = Scene.prepare(ctx, xyzw, colormap, radiusmap, densitymap, rbf_radius, radii=radii, ppl=ppl,
                           branching=branching)
        reqs = make_requests(sc)
        ctx.set_tuning(ctx.TUNE_STEP_CAP, 8192)
        ctx.set_tuning(ctx.TUNE_STEP_WAVES, 0)
        ctx.set_tuning(ctx.TUNE_GRAPH_ROUNDS, 0)
        ctx.set_tuning(ctx.TUNE_WALK_BUDGET, 256)
        ctx.set_tuning(ctx.TUNE_STEP_MIN_LIVE, 16)
        r = ctx.autotune(reqs, reps=reps)
        per[ppl] = r["best_ms"]
        settings[ppl] = r
        if best is None or r["best_ms"] < per[best[0]]:
            best = (ppl, sc)
    ppl, sc = best
    r = settings[ppl]
    ctx.set_tuning(ctx.TUNE_STEP_CAP, r["step_cap"])
    ctx.set_tuning(ctx.TUNE_STEP_WAVES, r["step_waves"])
    ctx.set_tuning(ctx.TUNE_GRAPH_ROUNDS, r["graph_rounds"])
    ctx.set_tuning(ctx.TUNE_WALK_BUDGET, r["walk_budget"])
    ctx.set_tuning(ctx.TUNE_STEP_MIN_LIVE, r["step_min_live"])
    return sc, {"particles_per_leaf": ppl, "step_cap": r["step_cap"], "step_waves": r["step_waves"],
                "graph_rounds": r["graph_rounds"], "walk_budget": r["walk_budget"],
                "step_min_live": r["step_min_live"], "best_ms": r["best_ms"],
                "per_leaf_size": per}
