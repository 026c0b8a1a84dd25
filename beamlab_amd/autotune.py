"""Autotuner of the step's execution form: workgroup size, beam ends per thread, launch form (one launch per
substep under a hipGraph, one workgroup per body, or the register-resident persistent form) and substeps per
launch. Each candidate is timed on the caller's own world (HIP events on the step's stream); the winner is kept in
a JSON cache keyed by the world's shape, so a later world of the same shape starts from it.
"""
import json
import os

from . import abi

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")

CANDIDATES = [
    dict(mode=abi.MODE_HALO, block=256, incidences_per_thread=8, substeps_per_launch=0, use_graph=0),
    dict(mode=abi.MODE_HALO, block=256, incidences_per_thread=4, substeps_per_launch=0, use_graph=0),
    dict(mode=abi.MODE_RESIDENT, block=256, incidences_per_thread=8, substeps_per_launch=0, use_graph=0),
    dict(mode=abi.MODE_RESIDENT, block=256, incidences_per_thread=4, substeps_per_launch=0, use_graph=0),
    dict(mode=abi.MODE_RESIDENT, block=512, incidences_per_thread=4, substeps_per_launch=0, use_graph=0),
    dict(mode=abi.MODE_RESIDENT, block=256, incidences_per_thread=2, substeps_per_launch=0, use_graph=0),
    dict(mode=abi.MODE_STEPPED, block=256, incidences_per_thread=1, substeps_per_launch=32, use_graph=1),
    dict(mode=abi.MODE_STEPPED, block=256, incidences_per_thread=2, substeps_per_launch=32, use_graph=1),
    dict(mode=abi.MODE_STEPPED, block=256, incidences_per_thread=4, substeps_per_launch=32, use_graph=1),
    dict(mode=abi.MODE_STEPPED, block=512, incidences_per_thread=2, substeps_per_launch=64, use_graph=1),
    dict(mode=abi.MODE_STEPPED, block=256, incidences_per_thread=8, substeps_per_launch=32, use_graph=1),
    dict(mode=abi.MODE_PERSISTENT, block=256, incidences_per_thread=4, substeps_per_launch=100, use_graph=0),
    dict(mode=abi.MODE_PERSISTENT, block=1024, incidences_per_thread=4, substeps_per_launch=100, use_graph=0),
]


def shape_key(world):
    nodes = sum(len(b.nodes) for b in world.bodies)
    beams = sum(len(b.beams) for b in world.bodies)
    return f"bodies={len(world.bodies)},nodes={nodes},beams={beams}"


def load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(world, substeps=200, frames=2, candidates=None, save=True, log=None):
    """Times every candidate on `world` (its state advances) and leaves the fastest in force.

    Returns (best candidate, [(candidate, ms per substep or None when it does not run here)])."""
    results = []
    best, best_ms = None, None
    for cand in candidates or CANDIDATES:
        try:
            world.set_tuning(**cand)
            world.time_substeps(max(1, substeps // 4))  # first launch: graph capture, code load
            ms = min(world.time_substeps(substeps) for _ in range(frames)) / substeps
        except abi.BeamlabError as e:
            results.append((cand, None))
            if log:
                log(f"autotune: {cand} skipped: {e}")
            continue
        results.append((cand, ms))
        if log:
            log(f"autotune: {cand}: {ms * 1e3:.2f} us/substep")
        if best_ms is None or ms < best_ms:
            best, best_ms = cand, ms
    if best is not None:
        world.set_tuning(**best)
        if save:
            cache = load_cache()
            cache[shape_key(world)] = {"tuning": best, "us_per_substep": best_ms * 1e3}
            try:
                with open(CACHE, "w") as f:
                    json.dump(cache, f, indent=1, sort_keys=True)
            except OSError:
                pass
    return best, results


def cached(world):
    """The cached tuning for this world's shape, or None."""
    e = load_cache().get(shape_key(world))
    return e["tuning"] if e else None
