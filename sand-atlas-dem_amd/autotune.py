"""Launch-shape autotuner for the collection stage (no reference counterpart).

The collection kernel is a persistent work-queue launch whose shape has several
knobs (`sad_hip_tune`): workgroups per CU, tasks taken per queue atomic, the
per-wavefront query-queue capacity, the leaf batch of the exact queries, the tasks
and traversals a wavefront queries together, and the gap reach of a traversal.
None of them changes any result (tasks are ranked by id, hits by their running
count within a task, and every query returns the same winner), so the tuner is
free to pick whichever is fastest on the workload at hand.  The temporal-skip re-anchor
distance is left at its default: its cost shows over hundreds of steps, not in a
few timed ones.  It times a few real steps per
candidate from one saved state (coordinate descent over the knobs, defaults
first), restores the state afterwards and optionally records the winner in a JSON
cache keyed by device and workload signature.
"""

import json
import time
from pathlib import Path

import numpy as np

SPACE = {
    "collect_blocks_per_cu": (2, 3, 4, 6, 8),
    "collect_pop": (1, 2, 4),
    "collect_queue": (128, 256),
    "collect_leaf_batch": (8, 16, 32, 64),
    "collect_batch_tasks": (8, 16),
    "collect_batch_fill": (48, 96, 160),
    "gap_reach_micro": (0, 300, 1000),
}


def signature(engine):
    """Workload features that decide the best launch shape."""
    shapes = engine.shapes
    ids = np.asarray(engine.ids)
    return {
        "device": engine.m.torch.cuda.get_device_name(engine.m.torch_device),
        "grains": int(engine.n),
        "templates": len(shapes),
        "mean_triangles": round(float(np.mean([len(shapes[i].faces) for i in ids])), 1),
        "mean_nodes": round(float(np.mean([len(shapes[i].samples) for i in ids])), 1),
    }


def apply(backend, setting):
    for name, value in setting.items():
        backend.tune(name, int(value))


def autotune(engine, *, rate=(0, 0, 0), steps=3, space=SPACE, cache=None, log=None):
    """Pick the fastest collection launch shape for `engine`'s workload.

    Returns {"best": {knob: value}, "ms_per_step": t, "trials": [...], "cached": bool}.
    The engine's state and step count are unchanged on return.
    """
    m = engine.m
    key = json.dumps(signature(engine), sort_keys=True)
    path = Path(cache) if cache else None
    table = json.loads(path.read_text()) if path and path.exists() else {}
    if key in table:
        apply(m, table[key]["best"])
        return dict(table[key], cached=True)
    state, step = engine.state(), engine.step
    current = {name: m.tune(name) for name in space}

    def measure(setting):
        apply(m, setting)
        engine.restore(state, step)
        # Every trial starts from the same temporal-skip state: no anchors, so the first
        # (untimed) step anchors every pair at the restored poses.
        invalidate = getattr(m, "invalidate", None)
        if invalidate:
            invalidate()
        engine.advance(1, rate=rate)  # first step after a change may capture a graph
        m.wait()
        t0 = time.perf_counter()
        engine.advance(steps, rate=rate)
        m.wait()
        return (time.perf_counter() - t0) / steps

    best_t = measure(current)
    trials = [(dict(current), best_t)]
    seen = {tuple(sorted(current.items()))}
    if log:
        log(f"[autotune] {current}: {1e3 * best_t:.3f} ms/step")
    for name, values in space.items():
        for value in values:
            trial = dict(current, **{name: value})
            tag = tuple(sorted(trial.items()))
            if tag in seen:
                continue
            seen.add(tag)
            t = measure(trial)
            trials.append((trial, t))
            if log:
                log(f"[autotune] {trial}: {1e3 * t:.3f} ms/step")
            if t < best_t:
                best_t, current = t, trial
    apply(m, current)
    engine.restore(state, step)
    result = {"best": current, "ms_per_step": 1e3 * best_t,
              "trials": [{"setting": s, "ms_per_step": 1e3 * t} for s, t in trials]}
    if path:
        table[key] = result
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(table, indent=1, sort_keys=True))
    return dict(result, cached=False)
