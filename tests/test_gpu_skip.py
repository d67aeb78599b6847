"""Temporal node skip of the collection stage (needs an MI355X).

A node whose signed distance at the anchor poses exceeds a rigorous bound on its
motion since is passed over; the contacts, forces and poses must be identical to
the run that queries every node, on every workload kind, with the skip limit
tight or loose, and across engines that share one workspace."""

from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("meta", "slip", "normalArea", "position", "positionRemainder", "velocity", "rotation", "momentum", "force",
         "torque", "accumulated")


def _run(w, backend, steps, chunks=3):
    e = w.engine(backend)
    e.initialize()
    states = []
    for _ in range(chunks):
        e.advance(steps, rate=w.rate)
        states.append(e.state())
    return states


def _same(a, b):
    assert a["contacts"] == b["contacts"]
    for name in NAMES:
        assert np.array_equal(a[name], b[name], equal_nan=True), name


@pytest.mark.parametrize("config", ["dense", "uneven", "walls", "layer"])
def test_skip_leaves_results_unchanged(config):
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload(config, n=300, templates=6, nodes=512 if config != "uneven" else None)
    plain = HipBackend(0, graph=True, temporal_skip=False)
    fast = HipBackend(0, graph=True, temporal_skip=True)
    for a, b in zip(_run(w, plain, 8), _run(w, fast, 8)):
        _same(a, b)
    s = fast.stats()
    assert s["skipped_nodes"] > 0
    assert s["queried_nodes"] < plain.stats()["queried_nodes"]


@pytest.mark.parametrize("limit,rebase", [(2000, 0), (250000, 0), (2000, 1)])
def test_skip_limit_does_not_change_results(limit, rebase):
    """Tight and loose re-anchor limits; with `skip_rebase` an expired anchor is moved
    by its motion bound instead of being recomputed (re-anchoring happens every few
    steps at the tight limit)."""
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload("dense", n=300, templates=6, nodes=512)
    plain = HipBackend(0, temporal_skip=False)
    fast = HipBackend(0, temporal_skip=True)
    fast.tune("skip_limit_micro", limit)
    fast.tune("skip_rebase", rebase)
    for a, b in zip(_run(w, plain, 5, chunks=6), _run(w, fast, 5, chunks=6)):
        _same(a, b)
    assert fast.stats()["skipped_nodes"] > 0


def test_engines_sharing_a_workspace_do_not_share_anchors():
    """Two workloads of identical sizes but different templates step alternately on
    one backend; each equals its run on a backend of its own."""
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    wa = assembly.config_workload("dense", n=300, templates=6, nodes=512, seed=11)
    wb = assembly.config_workload("dense", n=300, templates=6, nodes=512, seed=12)
    shared = HipBackend(0, graph=True)
    ea, eb = wa.engine(shared), wb.engine(shared)
    ea.initialize()
    eb.initialize()
    for _ in range(3):
        ea.advance(3, rate=wa.rate)
        eb.advance(3, rate=wb.rate)
    for w, e in ((wa, ea), (wb, eb)):
        alone = w.engine(HipBackend(0, graph=True, temporal_skip=False))
        alone.initialize()
        for _ in range(3):
            alone.advance(3, rate=w.rate)
        _same(e.state(), alone.state())


@pytest.mark.parametrize("config", ["dense", "uneven", "walls"])
def test_unchanged_winner_shortcut_leaves_results_unchanged(config):
    """A queued node whose point moved less than half its stored nearest-triangle gap
    since the previous collection takes that triangle without a traversal; the
    states equal those of full traversals, over many one- and multi-step calls
    (poses uploaded in between by restore()), while most queries take the shortcut."""
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload(config, n=300, templates=6, nodes=512 if config != "uneven" else None)
    full = HipBackend(0, graph=True)
    full.tune("gap_shortcut", 0)
    cut = HipBackend(0, graph=True)
    assert cut.tune("gap_shortcut") == 1
    a, b = w.engine(full), w.engine(cut)
    a.initialize()
    b.initialize()
    for k, steps in enumerate((1, 1, 4, 1, 6, 2, 10)):
        a.advance(steps, rate=w.rate)
        b.advance(steps, rate=w.rate)
        _same(a.state(), b.state())
        if k == 3:  # rewind both to a checkpoint: the stored gaps must stay sound
            state, step = a.state(), a.step
            a.restore(state, step)
            b.restore(state, step)
    s = cut.stats()
    assert full.stats()["unchanged_winners"] == 0
    assert s["unchanged_winners"] > 0.3 * s["queried_nodes"], s
