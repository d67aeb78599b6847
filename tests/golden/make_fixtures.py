"""Extract the reference's golden mesh-query vectors into small test fixtures.

Sources (data files the reference holds, read with numpy's non-pickling loader):
  evidence/coherent-mesh-query-probe-v1/shape-{0,1}.npz
      1,024 normalized FP32 query points per shape; `native` = the reference's
      production Metal meshProbe output (normal xyz, signed distance); `expected` =
      independent FP64 signed distance in metres (analytic box / trimesh prism).
      Shape 0 is a 1 x 2 x 3 mm box, shape 1 the extruded-L concave prism; the
      library's length unit is the median equivalent diameter of the two.
  evidence/mesh-face-normal-probe-v2/baseline.npz
      52 points around a 100 um cube with analytic distance and normal, and the
      production shader's output.
  evidence/{brick-mesh-query,adaptive-mesh-query,adaptive-breadth-query}-probe-v1/
      shape-{0,1}-cap-*.npz
      Three further sets of 2,820 normalized FP32 points per shape (cell boundaries,
      ties, external points, shape offsets) around the same box and prism; `expected`
      = independent FP64 signed distance in metres, `baseline` = the production
      meshProbe output.  The breadth set was taken on a twice-subdivided box of the
      same extents (same distances).
Run once with the reference tree mounted; the fixtures are committed.
"""

from pathlib import Path

import numpy as np

REFERENCE = Path("/root/reference/evidence")
HERE = Path(__file__).resolve().parent


def main():
    for sid, name in ((0, "coherent_probe_box"), (1, "coherent_probe_concave")):
        d = np.load(REFERENCE / "coherent-mesh-query-probe-v1" / f"shape-{sid}.npz")
        np.savez_compressed(HERE / f"{name}.npz", points=d["points"], expected=d["expected"], native=d["native"])
    d = np.load(REFERENCE / "mesh-face-normal-probe-v2" / "baseline.npz")
    np.savez_compressed(
        HERE / "face_normal_probe_box.npz",
        points=d["points"],
        expected_distance=d["expected_distance"],
        expected_normal=d["expected_normal"],
        actual=d["actual"],
    )
    for probe, tag, cap in (("brick-mesh-query-probe-v1", "brick", 96), ("adaptive-mesh-query-probe-v1", "adaptive", 24),
                            ("adaptive-breadth-query-probe-v1", "breadth", 24)):
        for sid, shape in ((0, "box"), (1, "concave")):
            d = np.load(REFERENCE / probe / f"shape-{sid}-cap-{cap}.npz")
            np.savez_compressed(HERE / f"{tag}_probe_{shape}.npz", points=d["points"], expected=d["expected"],
                                baseline=d["baseline"])


if __name__ == "__main__":
    main()
