"""Relaxed packings: drive a generated assembly to small overlaps with the step itself.

The seeded generators place grains at the centres of an overlap-free sphere packing
with random orientations, so irregular grains start with deep overlaps (several
percent of a diameter).  `relax` runs the product's own timestep with linear drag
(the reference's `drag` stage argument, `gpu.py:597-647`) from that state: the
contact forces push the grains apart while the drag removes the kinetic energy, an
overdamped descent on the penalty energy.  It stops once the largest overlap of any
contact is below `target` times the smaller equivalent diameter and the grains are
nearly at rest.  Walls stay fixed meanwhile.

Run as a script on a GPU to write a fixture that `assembly.config_workload(...,
relaxed=True)` loads:

    python sand-atlas-dem_amd/relax.py --config dense --out relaxed_dense.npz

A fixture stores the positions (m) and orientations of the relaxed grains, and an
identity hash of everything else the workload is built from (template vertices,
faces and nodes, shape ids, scales, cell, planes), so a fixture is never applied to
a different assembly.
"""

import hashlib
import json
import time

import numpy as np


def identity(w):
    """SHA-256 over what a relaxed state depends on besides the grain poses."""
    h = hashlib.sha256()
    for shape in w.shapes:
        for a in (shape.vertices, shape.faces, shape.samples, shape.areas):
            a = np.ascontiguousarray(a)
            h.update(str((a.dtype.str, a.shape)).encode())
            h.update(a.tobytes())
    for a in (w.ids.astype(np.uint32), np.asarray(w.scales, np.float64), np.asarray(w.cell, np.float64)):
        h.update(a.tobytes())
    h.update(json.dumps(w.spec.planes, sort_keys=True).encode())
    h.update(json.dumps(w.spec.periodic).encode())
    return h.hexdigest()


def relax(w, backend, *, target=0.008, damping=0.05, chunk=250, max_steps=60_000, rest=0.01, start_phi=None,
          compress_steps=20_000, settle_steps=0, log=print):
    """Relax workload `w`; returns (positions m, rotations, cell m, report).

    damping: velocity loss per step (drag = damping / dt).  rest: stop only once the
    largest grain speed is below `rest` m/s as well.  start_phi (periodic cells):
    first expand the assembly affinely to that solids fraction, relax it there, then
    compress it back to the workload's cell over `compress_steps` steps at a constant
    isotropic log rate (the reference's strain stage, `cell[1]`), relaxing throughout.
    settle_steps: before compressing, hold the expanded assembly for at most this many
    steps, or until its largest overlap is below `target`.
    """
    from dataclasses import replace

    from .spec import periodic_axes

    spec = replace(w.spec, max_overlap_fraction=0.5)
    axes = periodic_axes(spec.periodic)
    stretch = 1.0
    if start_phi is not None:
        if not axes.all():
            raise ValueError("start_phi needs a periodic cell")
        phi = float(np.sum([w.shapes[int(i)].volume * s**3 for i, s in zip(w.ids, w.scales)])) / np.prod(w.cell)
        stretch = (phi / start_phi) ** (1 / 3)
    run = replace(w, spec=spec, positions=w.positions * stretch, cell=np.asarray(w.cell) * stretch)
    e = run.engine(backend)
    e.initialize()
    drag = damping / spec.timestep
    walls = np.zeros((e.wall_count, 3)) if e.wall_count else None
    history = []
    t0 = time.perf_counter()
    steps = 0
    compress = 0 if stretch == 1.0 else compress_steps
    rate = np.full(3, -np.log(stretch) / max(compress, 1) / spec.timestep)
    phase = "settle" if compress and settle_steps else ("compress" if compress else "hold")
    compressed = 0
    while steps < max_steps:
        n = min(chunk, compress - compressed) if phase == "compress" else chunk
        e.advance(n, drag=drag, wall_velocities=walls, rate=rate if phase == "compress" else (0, 0, 0))
        steps += n
        compressed += n if phase == "compress" else 0
        d = e.diagnostics()
        speed = float(np.abs(e.read("velocity")[: e.n, :3]).max()) * e.length_unit / e.time_unit
        row = dict(step=steps, phase=phase, max_overlap=d["max_overlap_fraction"],
                   contacts=d["active_surface_contacts"], grain_pairs=d["active_grain_pairs"],
                   coordination=d["coordination"], max_speed=speed, kinetic_energy=d["kinetic_energy"],
                   pressure=d["p"], volume_ratio=float(np.prod(d["cell"]) / np.prod(w.cell)),
                   seconds=time.perf_counter() - t0)
        history.append(row)
        log(json.dumps(row))
        if phase == "settle":
            if d["max_overlap_fraction"] < target or steps >= settle_steps:
                phase = "compress"
            continue
        if phase == "compress":
            if compressed >= compress:
                phase = "hold"
            continue
        if d["max_overlap_fraction"] < target and speed < rest:
            break
    s = e.state()
    positions = (s["position"][:, :3].astype(np.float64) + s["positionRemainder"][:, :3]) * e.length_unit
    cell = s["cell"][0, :3].astype(np.float64) * e.length_unit
    positions = np.where(axes, np.mod(positions, cell), positions)
    rotations = s["rotation"].astype(np.float64) + s["rotationRemainder"]
    rotations /= np.linalg.norm(rotations, axis=1, keepdims=True)
    report = dict(steps=steps, converged=history[-1]["max_overlap"] < target, start_phi=start_phi,
                  history=history)
    return positions, rotations, cell, report


def save(path, w, positions, rotations, cell, report):
    np.savez_compressed(path, positions=positions.astype(np.float32), rotations=rotations.astype(np.float32),
                        cell=np.asarray(cell, np.float64), identity=np.array(identity(w)),
                        report=np.array(json.dumps(report)))


def main():
    import argparse
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    sys.path.insert(0, str(root))
    from __graft_entry__ import load_package

    load_package()
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd import relax as package
    from sand_atlas_dem_amd.hip import HipBackend

    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--config", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--target", type=float, default=0.008)
    ap.add_argument("--damping", type=float, default=0.05)
    ap.add_argument("--chunk", type=int, default=250)
    ap.add_argument("--max-steps", type=int, default=60_000)
    ap.add_argument("--start-phi", type=float, default=None)
    ap.add_argument("--compress-steps", type=int, default=20_000)
    ap.add_argument("--settle-steps", type=int, default=0)
    args = ap.parse_args()
    t0 = time.perf_counter()
    w = assembly.config_workload(args.config, seed=args.seed, n=args.n, relaxed=False)
    print(f"workload {args.config}: {len(w.ids)} grains in {time.perf_counter() - t0:.1f} s", flush=True)
    backend = HipBackend(0, graph=True)
    positions, rotations, cell, report = package.relax(
        w, backend, target=args.target, damping=args.damping, chunk=args.chunk, max_steps=args.max_steps,
        start_phi=args.start_phi, compress_steps=args.compress_steps, settle_steps=args.settle_steps,
        log=lambda s: print(s, flush=True))
    package.save(args.out, w, positions, rotations, cell, report)
    print(json.dumps({k: v for k, v in report.items() if k != "history"}), flush=True)


if __name__ == "__main__":
    main()
