#!/usr/bin/env python3
"""Benchmark: grain-steps/s of the surface-mesh DEM step on one MI355X.

Headline workload (default `--config uneven`): 50,000 grains from a 424-template
library of sand-like spherical-harmonic grains (320-5,120 triangles, log-uniform;
256-4,096 area-weighted surface nodes in proportion to area), solids fraction 0.60
on a periodic sphere packing, Coulomb mu 0.5, FP32 state, synthetic and seeded.
Further workloads (`--extra`, default the 10,000-grain dense library) are measured
the same way and reported under "extra".  A step is one full explicit timestep
(neighbour check/rebuild, node->triangle BVH contact collection, friction-history
resolve, force/torque gather, two half kicks).  The timed region is
`Engine.advance(K)` with inputs already resident in HBM, bracketed by device
synchronisation, as the reference's benchmark times `advance`
(src/sand_atlas_dem/benchmark.py:244-269).

For each workload: the plain HIP path (hipGraph-captured steps, autotuned launch
shape), a per-kernel HIP-event profile with the dominant kernel's roofline, and the
learned cull trained on the GPU, audited and timed from the same state.  For the
headline workload also the C++ oracle on the host cores (multi-thread and single
thread) with a parity check against the HIP path over the same steps.

Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s

METRIC = "grain-steps/s at 10k & 50k grains, plain vs NN-culled HIP; force rel err ≤1e-5"
WORKLOADS = {
    "dense": "10,000 grains, 64 perturbed icospheres (subdiv 2-4, sigma 0.15), 2,048 nodes/grain, phi 0.60, "
             "isotropic compression 0.5/s, mu 0.5",
    "uneven": "50,000 grains, 424 sand-like spherical-harmonic templates (320-5,120 triangles log-uniform, "
              "256-4,096 nodes by area), phi 0.60, mu 0.5",
    "concave": "1,000 grains, 8 strongly non-convex harmonic templates (1,280 triangles, 1,024 nodes), phi 0.55, mu 0.5",
    "small": "1,000 icospheres (subdiv 2, sigma 0.1), 512 nodes/grain, phi 0.30, v ~ N(0, 0.01 m/s), mu 0.5",
    "walls": "10,000 grains from the dense library at phi 0.58 in a nonperiodic box of six planes, top plane moving "
             "down at 0.5 box heights/s, mu 0.5",
    "layer": "10,000 grains from the dense library at phi 0.58 between two z planes, x and y periodic, top plane "
             "moving down at 0.5 layer heights/s, mu 0.5",
    "sdf": "the dense assembly with SDF-grid contact geometry (grids at resolution 64, trilinear field): 10,000 "
           "grains, 64 perturbed icospheres, 2,048 nodes/grain, phi 0.60, isotropic compression 0.5/s, mu 0.5",
}


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def load_package():
    sys.path.insert(0, str(ROOT))
    from __graft_entry__ import load_package as load

    return load()


def kernel_bytes(name, n, pd, q, c, hash_size, sdf_static=0):
    """Algorithmic HBM bytes per launch: per-unit figures (bytes per grain, directed pair, queued
    node and contact) times the launch's units."""
    table = {
        "k_collect": 92 * pd + 20 * q + 48 * c,
        "k_resolve": 128 * c,
        "k_gather": 64 * c + 112 * n,
        "k_scatter": 112 * c,
        "k_build_hash": 20 * c,
        "k_clear_storage": 4 * hash_size,
        "k_begin_step": 208 * n,
        "k_finish_step": 160 * n,
        # SDF geometry: per directed pair its rows (pair, ids, bodies, poses, grid rows)
        # and its 256-B hit bitmask (2,048 nodes), the templates' nodes and grids once
        # (they stay in the caches), per contact its rows.
        "k_sdf_count": 376 * pd + sdf_static,
        "k_sdf_emit": 376 * pd + 48 * c,
    }
    return table.get(name)


def nn_leg(args, w, e, hip, start, step0, plain_state, log_):
    """Train the learned cull (untimed), audit it, then time the culled path from the
    same saved state as the plain timed run."""
    import numpy as np
    import torch
    from sand_atlas_dem_amd import learned

    t0 = time.perf_counter()
    lc = learned.LearnedCull(e, points=args.nn_points, batch=args.nn_batch, validation=16384, seed=args.seed)
    data_s = time.perf_counter() - t0
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    lc.train(iterations=args.nn_iters, lr=3e-3, log_every=max(1, args.nn_iters // 4))
    ev1.record()
    torch.cuda.synchronize()
    train_ms = ev0.elapsed_time(ev1)
    lc.calibrate()
    rep = lc.validation_report
    points = args.nn_iters * lc.batch * lc.templates
    flops = 61440.0 * points  # MFMA flops per point per Adam step: forward 22,528 + weight grads 22,528 + backprop 16,384
    log_(f"[bench] nn: data {data_s:.1f} s, {args.nn_iters} Adam steps in {train_ms / 1e3:.1f} s "
         f"({flops / (train_ms * 1e-3) / 1e12:.2f} TFLOP/s MFMA), enabled {rep['templates_enabled']}/{rep['templates']} "
         f"templates, max error {rep['max_error_mean']:.3f} (label units)")
    out = {
        "templates_enabled": rep["templates_enabled"],
        "templates": rep["templates"],
        "train": {"points_per_template": lc.points, "batch": lc.batch, "iterations": args.nn_iters,
                  "data_s": data_s, "train_s": train_ms / 1e3, "mfma_tflops": flops / (train_ms * 1e-3) / 1e12,
                  "mfma_peak_tflops": 2500.0, "final_mse": lc.history[-1][1] if lc.history else None},
        "validation": rep,
    }
    if not lc.attach(audit=True):
        out.update(value=None, note="no template admits a safe cull (3 x max error >= label clamp); "
                                    "the NN-culled path is the plain path")
        return out
    e.restore(start, step0)
    before = hip.stats()
    e.advance(args.audit_steps, rate=w.rate)
    hip.wait()
    after = hip.stats()
    queued = after["queried_nodes"] - before["queried_nodes"]
    out["audit"] = {"steps": args.audit_steps, "queued_nodes": queued, "culled": after["nn_culled"],
                    "misses": after["nn_misses"], "cull_fraction": after["nn_culled"] / max(queued, 1)}
    lc.attach(audit=False)
    e.restore(start, step0)
    hip.wait()
    before = hip.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.advance(args.steps, rate=w.rate)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    after = hip.stats()
    lc.detach()
    out["timed"] = {"culled_per_step": (after["nn_culled"] - before["nn_culled"]) / args.steps,
                    "queued_per_step": (after["queried_nodes"] - before["queried_nodes"]) / args.steps,
                    "contacts_per_step": (after["contacts"] - before["contacts"]) / args.steps}
    # Any sound cull drops only queued nodes that are not contacts: (Q - C) / Q of the
    # exact queries is the most it can remove.
    t = out["timed"]
    out["sound_cull_bound"] = max(t["queued_per_step"] - t["contacts_per_step"], 0.0) / max(t["queued_per_step"], 1.0)
    st = e.state()
    same = st["contacts"] == plain_state["contacts"] and all(
        np.array_equal(st[k], plain_state[k]) for k in ("meta", "force", "torque", "position", "rotation"))
    out.update(value=e.n * args.steps / elapsed, ms_per_step=1e3 * elapsed / args.steps, identical_to_plain=bool(same))
    log_(f"[bench] nn-culled: {out['ms_per_step']:.3f} ms/step, audit {out['audit']}, identical {same}")
    return out


def servo_leg(args, w, e, state, step, log_):
    """Config 5's loading on the layer: a triaxial stage holding the x-y stress of the
    layer at 20 kPa with the reference's servo (gain 1,000, filter time 10 us, rates up
    to 1,000/s as in its drained-triaxial case) while the top plane closes at 0.5 layer
    heights/s; diagnostics every 50 steps (the reference default).  Timed from the plain
    run's start state over the same step count, diagnostics and servo included."""
    import numpy as np
    import torch
    from sand_atlas_dem_amd.servo import Stages

    stage = dict(kind="triaxial", steps=10 ** 9, pressure=2e4, gain=1000.0, filter_time=1e-5, max_rate=1000.0,
                 rate=0.5, stop_strain=0.1)
    interval = 50
    e.restore(state, step)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run = Stages(e, [stage], interval=interval)
    run.run(max_chunks=max(1, args.steps // interval))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    steps = e.step - step
    last = run.rows[-1]
    out = {"pressure_pa": 2e4, "interval": interval, "steps": steps, "ms_per_step": 1e3 * elapsed / steps,
           "value": e.n * steps / elapsed,
           "last_rate": last["rate"], "layer_stress_diag_pa": np.diag(last["layer_stress"]).tolist(),
           "layer_axial_strain": last["layer_axial_strain"],
           "wall_reactions_N": [x["reaction"] for x in last["walls"]], "wall_work_J": last["wall_work"]}
    log_(f"[bench] layer servo: {out['ms_per_step']:.3f} ms/step with diagnostics every {interval} steps, "
         f"rates {last['rate']}, layer stress {out['layer_stress_diag_pa']}")
    return out


def _profile_record(kernel, config, kind):
    names = {"dense": "dense10k", "uneven": "uneven50k"}
    path = ROOT / "profiles" / f"{kernel}_{names.get(config, config)}_{kind}.json"
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    data["source"] = str(path.relative_to(ROOT))
    return data if data.get("kernel") == kernel and data.get("workload") == config else None


def measured_traffic(kernel, config):
    """HBM bytes per launch from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by its 1 GiB copy calibration), if one exists."""
    data = _profile_record(kernel, config, "traffic")
    return data["traffic_bytes"] if data else None


def measured_flops(kernel, config):
    """FP32 operations per launch from the committed rocprofv3 PMC summary (VALU FP32
    add, multiply, FMA and transcendental wave instructions, times the active lanes
    per VALU instruction), if one exists."""
    data = _profile_record(kernel, config, "valu")
    if not data:
        return None
    return {"flops_per_launch": data["flops_per_launch"], "active_lanes": data["active_lanes"],
            "peak_tflops": 157.3, "source": data["source"]}


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_legs(args, w, state, step0, hip_rate, log_):
    """The oracle on the host cores: `cpu_steps` steps of the same workload from the
    GPU's state on min(16, cores) threads, checked against the HIP path over the same
    steps (the parity leg); then one step on a single thread, of the same workload or,
    for large ones, of a smaller assembly drawn from the same library."""
    import numpy as np
    from oracle import oracle
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    pkg = load_package()
    cores = host_cores()
    threads = args.cpu_threads or min(16, cores)
    cpu = w.engine(oracle.make_oracle(pkg, threads=threads))
    cpu.restore(state, step0)
    ref = w.engine(HipBackend(0, graph=True))
    ref.restore(state, step0)
    t0 = time.perf_counter()
    cpu.advance(args.cpu_steps, rate=hip_rate)
    cpu_s = time.perf_counter() - t0
    ref.advance(args.cpu_steps, rate=hip_rate)
    a, b = ref.state(), cpu.state()
    errs = {}
    for name in ("force", "torque", "position", "rotation"):
        x, y = a[name][:, :3].astype(np.float64), b[name][:, :3].astype(np.float64)
        errs[name] = float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-30))
    parity = {"contacts_equal": bool(a["contacts"] == b["contacts"] and np.array_equal(a["meta"], b["meta"])),
              "states_identical": all(np.array_equal(a[k], b[k]) for k in ("force", "torque", "position", "rotation",
                                                                              "velocity", "slip")),
              "max_rel_err": errs, "steps": args.cpu_steps, "contacts": int(a["contacts"])}
    n = len(w.ids)
    many = {"value": n * args.cpu_steps / cpu_s, "threads": threads, "steps": args.cpu_steps, "grains": n}
    log_(f"[bench] cpu oracle on {threads} threads: {many['value']:.4g} grain-steps/s; parity {parity}")
    single_n = n if n <= args.cpu_single_max else args.cpu_single_max
    if single_n == n:
        one_w, one_state, one_step = w, state, step0
    else:
        one_w = assembly.config_workload(w.name, seed=args.seed, n=single_n)
        one_state, one_step = None, 0
    one = one_w.engine(oracle.make_oracle(pkg, threads=1))
    if one_state is None:
        one.initialize()
    else:
        one.restore(one_state, one_step)
    t0 = time.perf_counter()
    one.advance(args.cpu_single_steps, rate=hip_rate)
    one_s = time.perf_counter() - t0
    single = {"value": single_n * args.cpu_single_steps / one_s, "threads": 1, "steps": args.cpu_single_steps,
              "grains": single_n}
    log_(f"[bench] cpu oracle on 1 thread: {single['value']:.4g} grain-steps/s ({single_n} grains)")
    baseline = {
        "value": many["value"],
        "unit": "grain-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{args.cpu_steps} steps of the same {n}-grain workload from the GPU's state: the C++ oracle "
                  f"(serial kernels, order-preserving pair-parallel collector) on {threads} threads",
        "cpu_model": cpu_model(),
        "host_cores": cores,
        "single_thread": dict(single, sample=(
            f"{args.cpu_single_steps} steps of the same workload on 1 thread" if single_n == n else
            f"{args.cpu_single_steps} steps of a {single_n}-grain assembly from the same template library and "
            f"solids fraction on 1 thread (a full {n}-grain step takes about {n / single['value']:.0f} s there)")),
    }
    return baseline, parity


def curve(args, log_):
    """Step time over a long run: the temporal skip and the unchanged-winner shortcut
    keep state from step to step, so their gain can drift as the packing evolves."""
    import torch
    from sand_atlas_dem_amd import assembly, autotune
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload(args.config, seed=args.seed, n=args.grains)
    hip = HipBackend(0, graph=not args.no_graph)
    e = w.engine(hip)
    e.initialize()
    e.advance(args.warmup, rate=w.rate)
    if not args.no_autotune:
        tuned = autotune.autotune(e, rate=w.rate, steps=10, log=log_)
        log_(f"[curve] autotuned launch shape {tuned['best']}")
    done = 0
    while done < args.curve:
        k = min(args.curve_window, args.curve - done)
        before = hip.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.advance(k, rate=w.rate)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / k
        after = hip.stats()
        done += k
        d = {name: (after[name] - before[name]) / k for name in after}
        print(json.dumps({"config": args.config, "grains": e.n, "steps_done": done, "ms_per_step": round(ms, 4),
                          "contacts_per_step": round(d["contacts"]), "queried_per_step": round(d["queried_nodes"]),
                          "shortcut_fraction": round(d["unchanged_winners"] / max(d["queried_nodes"], 1), 4),
                          "skipped_per_step": round(d["skipped_nodes"]),
                          "anchored_per_step": round(d["anchored_tasks"]),
                          "directed_pairs": round(d["directed_pairs"])}), flush=True)


def measure(args, config, log_, *, primary):
    """Everything measured on one workload; returns a block of the JSON line."""
    import numpy as np
    import torch
    from sand_atlas_dem_amd import assembly, autotune
    from sand_atlas_dem_amd.hip import HipBackend

    t0 = time.perf_counter()
    n_override = args.grains if primary else None
    w = assembly.config_workload(config, seed=args.seed, n=n_override)
    log_(f"[bench] workload {config}: {len(w.ids)} grains, {len(w.shapes)} templates, "
         f"dt {w.spec.timestep:.3g} s, generated in {time.perf_counter() - t0:.1f} s")
    hip = HipBackend(0, graph=not args.no_graph)
    t0 = time.perf_counter()
    e = w.engine(hip)
    e.initialize()
    log_(f"[bench] engine + initial forces {time.perf_counter() - t0:.1f} s")
    e.advance(args.warmup, rate=w.rate)
    hip.wait()
    tuned = None
    sdf = not e.mesh_geometry
    if not args.no_autotune and not sdf:
        tuned = autotune.autotune(e, rate=w.rate, steps=10, log=log_)
        log_(f"[bench] autotuned launch shape {tuned['best']} ({tuned['ms_per_step']:.3f} ms/step)")
    start, step0 = e.state(), e.step
    e.restore(start, step0)
    e.advance(1, rate=w.rate)  # capture the graphs outside the timed region
    e.restore(start, step0)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.advance(args.steps, rate=w.rate)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n = e.n
    value = n * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log_(f"[bench] {config}: timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.4g} grain-steps/s")
    plain_state = e.state()

    # Per-kernel HIP-event profile over a separate window (same stream as the launches).
    before = hip.stats()
    hip.profile(True)
    e.advance(args.profile_steps, rate=w.rate)
    hip.wait()
    prof = hip.profile_read()
    hip.profile(False)
    after = hip.stats()
    steps_p = after["collections"] - before["collections"]
    pd = (after["directed_pairs"] - before["directed_pairs"]) / steps_p
    q = (after["queried_nodes"] - before["queried_nodes"]) / steps_p
    c = (after["contacts"] - before["contacts"]) / steps_p
    skipped = (after["skipped_nodes"] - before["skipped_nodes"]) / steps_p
    unchanged = (after["unchanged_winners"] - before["unchanged_winners"]) / steps_p
    sdf_static = sum(16 * len(s.samples) + 4 * s.sdf.size for s in w.shapes) if sdf else 0
    kernels = {}
    total_ms = sum(v[0] for v in prof.values())
    for name, (ms, launches) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        avg = ms / launches
        b = kernel_bytes(name, n, pd, q, c, e.hash_size, sdf_static)
        kernels[name] = {"avg_ms": round(avg, 5), "share": round(ms / total_ms, 4),
                         **({"gbps": round(b / (avg * 1e-3) / 1e9, 2)} if b else {})}
    dominant = max(prof, key=lambda k: prof[k][0])
    dom_ms = prof[dominant][0] / prof[dominant][1]
    dom_bytes = kernel_bytes(dominant, n, pd, q, c, e.hash_size, sdf_static) or 0
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = 296 * n + (752 * pd + sdf_static if sdf else 92 * pd + 20 * q) + 240 * c
    log_(f"[bench] dominant kernel {dominant}: {dom_ms:.3f} ms/launch; P_d {pd:.0f} Q {q:.0f} C {c:.0f} per step, "
         f"{unchanged:.0f} queries kept their nearest triangle without a traversal")
    flops = measured_flops(dominant, config)
    block = {
        "workload": f"{config}: {WORKLOADS[config]}",
        "grains": n,
        "value": value,
        "ms_per_step": ms_per_step,
        "launch_shape": tuned["best"] if tuned else None,
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": achieved / (PEAK_HBM / 1e9),
            "traffic": measured_traffic(dominant, config),
            "launch_ms": dom_ms,
        },
        "step_roofline": {
            "bytes_per_step": step_bytes,
            "achieved_gbps": step_bytes / (ms_per_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_per_step * 1e-3) / PEAK_HBM,
            "per_step": {"directed_pairs": pd, "queried_nodes": q, "contacts": c, "skipped_nodes": skipped,
                         "unchanged_winners": unchanged, "contacts_per_grain": c / n},
        },
        "fp32": None if flops is None else dict(flops, achieved_tflops=flops["flops_per_launch"] / (dom_ms * 1e-3) / 1e12),
        "kernels": kernels,
    }
    if primary and not args.no_cpu_baseline:
        block["cpu_baseline"], block["parity"] = cpu_legs(args, w, plain_state, step0 + args.steps, w.rate, log_)
    nn = None if args.no_nn or sdf else nn_leg(args, w, e, hip, start, step0, plain_state, log_)
    if nn and nn.get("value"):
        nn["speedup_vs_plain"] = nn["value"] / value
    block["nn_culled"] = nn
    if config == "layer":
        block["servo"] = servo_leg(args, w, e, start, step0, log_)
    if e.wall_count:  # wall reactions and work after the timed run (SI units)
        e.restore(plain_state, step0 + args.steps)
        block["walls"] = [{"normal": x["normal"], "offset_m": x["offset"], "reaction_N": list(x["reaction"]),
                           "work_J": x["work"], "contacts": x["active_surface_contacts"]} for x in e.walls()]
    return block


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="uneven", choices=sorted(WORKLOADS))
    ap.add_argument("--grains", type=int, default=None, help="override the headline workload's grain count")
    ap.add_argument("--extra", default="dense", help="comma-separated further workloads measured after the "
                                                     "headline (plain + NN-culled), or 'none'")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=5, help="oracle steps in the multi-thread CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
    ap.add_argument("--cpu-single-max", type=int, default=5000,
                    help="largest assembly the single-thread CPU leg steps (a smaller one is drawn beyond)")
    ap.add_argument("--cpu-single-steps", type=int, default=5, help="steps of the single-thread CPU leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels one by one instead of a captured step")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-nn", action="store_true", help="skip the learned-cull leg")
    ap.add_argument("--nn-points", type=int, default=1 << 20)
    ap.add_argument("--nn-batch", type=int, default=16384)
    ap.add_argument("--nn-iters", type=int, default=4000)
    ap.add_argument("--audit-steps", type=int, default=5)
    ap.add_argument("--curve", type=int, default=0,
                    help="instead of the bench line: run this many steps of --config after the autotuned "
                         "warmup and print ms/step per --curve-window steps (one JSON line each)")
    ap.add_argument("--curve-window", type=int, default=100)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 4, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    load_package()
    if args.curve:
        curve(args, log)
        return
    head = measure(args, args.config, log, primary=True)
    extras = {}
    for config in [c for c in args.extra.split(",") if c and c != "none" and c != args.config]:
        extras[config] = measure(args, config, log, primary=False)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "grain-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded templates and packing)",
        "config": {"workload": head["workload"], "grains": head["grains"],
                   "path": "plain HIP" + (" (hipGraph-captured steps)" if not args.no_graph else ""),
                   "launch_shape": head["launch_shape"], "seed": args.seed},
        "roofline": head["roofline"],
        "cpu_baseline": head.get("cpu_baseline"),
        "parity": head.get("parity"),
        "nn_culled": head["nn_culled"],
        "step_roofline": head["step_roofline"],
        "fp32": head["fp32"],
        "kernels": head["kernels"],
        "extra": extras,
    }
    if "walls" in head:
        out["walls"] = head["walls"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
