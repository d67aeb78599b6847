"""Contacts written by the reference's own generated CPU collector, as a fixture.

The reference's Slang-generated mesh collector
(`evidence/hostun424-cpu-pair-plane-counts-v1/native-profile-pair_plane_serial/
instrumented-collector.cpp`) is compiled from the reference tree by `oracle/Makefile`
(`make ref`) in the build container only.  This script runs it, through the
oracle adapter, on a seeded 150-grain dense assembly (4 perturbed icosphere
templates, 512 nodes each): initial forces, then two steps under isotropic
compression.  It stores the committed contact rows (meta, pointDepth, normalArea)
so that GPU tests compare the HIP collection with the reference's collector
without loading it.  Run once where /root/reference exists; the fixture is
committed.
"""

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
WORKLOAD = dict(config="dense", n=150, templates=4, nodes=512, seed=9)
STEPS = 2


def main():
    sys.path.insert(0, str(ROOT))
    from __graft_entry__ import load_package

    pkg = load_package()
    from oracle import oracle
    from sand_atlas_dem_amd import assembly

    oracle.build(reference=True)
    w = assembly.config_workload(WORKLOAD["config"], n=WORKLOAD["n"], templates=WORKLOAD["templates"],
                                 nodes=WORKLOAD["nodes"], seed=WORKLOAD["seed"])
    e = w.engine(oracle.make_oracle(pkg, collector="reference"))
    e.initialize()
    e.advance(STEPS, rate=w.rate)
    meta, point_depth, normal_area = e.contacts()
    np.savez_compressed(HERE / "ref_collector_dense150.npz", meta=meta, pointDepth=point_depth,
                        normalArea=normal_area, steps=np.array(STEPS))
    print(f"{len(meta)} contacts")


if __name__ == "__main__":
    main()
