"""Experiment stages and the stress servo: the cell strain rates of each diagnostic interval.

The reference runs an experiment as a list of stages (`simulation.py:163-315`): each
diagnostic interval advances `diagnostic_interval` steps at the rates that the last
diagnostics row gives (`Simulation.rate`, `simulation.py:138-161`):

* `hold` and `wall` stages: no cell strain (a wall stage sets new plane velocities);
* `strain`: the stage's fixed rates;
* `isotropic` and `triaxial`: the diagonal stress passes through a first-order filter
  (time constant `filter_time`, one update per interval), and each rate is
  gain x (filtered - pressure) / pressure, clipped to +-max_rate, or to +-contact_max_rate
  once any surface contact has formed in the stage; a triaxial stage then compresses z at
  its own `rate`.

An isotropic stage ends once `converged_samples` rows in a row meet the pressure, kinetic
energy and unbalanced-force tolerances, and raises if its step budget runs out first; a
triaxial stage ends at `stop_strain` axial strain or at its step budget.

Per-axis periodicity extends this to a cell whose nonperiodic axes planes bound (the
reference allows only hold and wall stages without periodicity, `spec.py:136-137`):
the servo drives the periodic axes alone, and the planes load the others with their own
velocities.  The layer of config 5 is that case: grains between two z planes, periodic in
x and y, the top plane moving down, the x-y stress held at the confining pressure.  Its
stress, pressure test and axial strain refer to the layer between the planes, not to the
whole cell (the grains' sums are the same, the volume is the layer's).
"""

import numpy as np

from .spec import periodic_axes

KINDS = ("hold", "wall", "strain", "isotropic", "triaxial")


class StressServo:
    """The filter and contact-phase state of the reference's servo (`simulation.py:83-85`)
    and its rate rule (`simulation.py:138-161`)."""

    def __init__(self, interval, timestep):
        self.dt = float(interval) * float(timestep)  # time per diagnostic interval (s)
        self.filter_stress = np.zeros(3)
        self.contact_phase = False

    def rate(self, stage, row):
        """Cell strain rates (1/s) for the interval after diagnostics row `row`."""
        kind = stage["kind"]
        if kind in ("hold", "wall"):
            return np.zeros(3)
        if kind == "strain":
            return np.asarray(stage["rate"], float)
        diagonal = np.diag(np.asarray(row["stress"], float))
        alpha = -np.expm1(-self.dt / stage["filter_time"])
        self.filter_stress += (diagonal - self.filter_stress) * alpha
        self.contact_phase = self.contact_phase or row["active_surface_contacts"] > 0
        limit = stage["max_rate"]
        if self.contact_phase:
            limit = min(limit, stage.get("contact_max_rate", limit))
        rate = np.clip(stage["gain"] * (self.filter_stress - stage["pressure"]) / stage["pressure"], -limit, limit)
        if kind == "triaxial":
            rate[2] = -stage["rate"]
        return rate


def validate_stages(stages, periodic, wall_count):
    """The reference's stage checks (`spec.py:113-166`), with per-axis periodicity: a
    cell with at least one periodic axis takes every stage kind, the servo and strain
    rates acting on the periodic axes only."""
    axes = periodic_axes(periodic)
    if not stages:
        raise ValueError("At least one stage required")
    for s in stages:
        kind = s.get("kind")
        if kind not in KINDS:
            raise ValueError("Unsupported stage or unknown stage fields")
        if not axes.any() and kind not in ("hold", "wall"):
            raise ValueError("Nonperiodic domains support hold or wall stages only")
        if type(s.get("steps")) is not int or s["steps"] < 1:
            raise ValueError("Positive stage steps required")
        drag = s.get("drag", 0)
        if not np.isfinite(drag) or drag < 0:
            raise ValueError("Drag must be finite and nonnegative")
        if kind == "wall":
            if not wall_count:
                raise ValueError("Wall stages require planes")
            velocities = np.asarray(s.get("wall_velocities"), float)
            if velocities.shape != (wall_count, 3) or not np.isfinite(velocities).all():
                raise ValueError("wall_velocities must be finite, one row per plane")
        elif "wall_velocities" in s:
            raise ValueError("wall_velocities require a wall stage")
        if kind in ("isotropic", "triaxial"):
            for key in ("pressure", "gain", "filter_time", "max_rate"):
                if not np.isfinite(s.get(key, np.nan)) or s[key] <= 0:
                    raise ValueError(f"Explicit positive servo {key} required")
        for key in ("pressure_tolerance", "kinetic_tolerance", "force_tolerance", "stop_strain", "contact_max_rate"):
            if key in s and (not np.isfinite(s[key]) or s[key] <= 0):
                raise ValueError(f"Positive finite {key} required")
        if "converged_samples" in s and (type(s["converged_samples"]) is not int or s["converged_samples"] < 1):
            raise ValueError("Positive integer converged_samples required")
        if kind == "strain":
            rate = np.asarray(s.get("rate"), float)
            if rate.shape != (3,) or not np.isfinite(rate).all():
                raise ValueError("A strain stage needs three finite rates")
            if np.any(rate[~axes] != 0):
                raise ValueError("Strain rates require periodic axes")
        if kind == "triaxial" and (not np.isfinite(s.get("rate", np.nan)) or s["rate"] <= 0):
            raise ValueError("Triaxial compression rate must be positive")
        if kind == "triaxial" and not axes[2] and not wall_count:
            raise ValueError("A triaxial stage needs a periodic z axis or planes to load it")


def z_gap(row):
    """Distance between the row's two z planes (normals +z and -z), or None."""
    low = [p["offset"] for p in row.get("walls", []) if p["normal"] == [0.0, 0.0, 1.0]]
    high = [p["offset"] for p in row.get("walls", []) if p["normal"] == [0.0, 0.0, -1.0]]
    if len(low) != 1 or len(high) != 1:
        return None
    return -high[0] - low[0]


class Stages:
    """Run `stages` on `engine` one diagnostic interval at a time (`simulation.py:233-315`
    without its files: rows are kept in memory, no checkpoints or snapshots are written).

    rows[0] is the diagnostics of the state before the first interval; every later row
    carries the stage index, the rates it ran at and the work-balance residual
    (`simulation.py:260-281`).  On a cell whose z axis two planes bound, each row also
    holds the layer's gap, its stress (the cell stress times cell height / gap) and its
    axial strain from the gap, and the servo and stage tests use those.
    """

    def __init__(self, engine, stages, *, interval=None):
        spec = engine.spec
        self.engine, self.stages = engine, [dict(s) for s in stages]
        self.interval = int(interval or spec.diagnostic_interval)
        self.axes = periodic_axes(spec.periodic)
        validate_stages(self.stages, spec.periodic, engine.wall_count)
        self.servo = StressServo(self.interval, spec.timestep)
        self.stage_index = self.stage_steps = self.converged = 0
        self.rows = [self._row(engine.diagnostics())]
        self.gap0 = self.rows[0].get("layer_gap")

    def _row(self, row, **extra):
        row = dict(row, **extra)
        gap = z_gap(row) if not self.axes[2] else None
        if gap is not None:
            row["layer_gap"] = gap
            row["layer_stress"] = (np.asarray(row["stress"]) * row["cell"][2] / gap).tolist()
            if getattr(self, "gap0", None):
                row["layer_axial_strain"] = 1 - gap / self.gap0
        return row

    def _servo_row(self, row):
        if "layer_stress" in row:
            return dict(row, stress=row["layer_stress"])
        return row

    def _axial_strain(self, row):
        return row.get("layer_axial_strain", row["axial_strain"])

    def _next_stage(self, row):
        """Stage change after exhaustion or convergence (`simulation.py:236-247, 302-309`):
        the filter carries over, the contact phase and the strain origin restart."""
        self.stage_index += 1
        self.stage_steps = self.converged = 0
        self.servo.contact_phase = False
        self.engine.initial_cell = np.asarray(row["cell"], float)
        self.gap0 = row.get("layer_gap")

    def run(self, max_chunks=None):
        """Advance until every stage is done or `max_chunks` intervals ran; returns
        "complete" or "paused"."""
        chunks = 0
        while self.stage_index < len(self.stages):
            stage = self.stages[self.stage_index]
            if self.stage_steps >= stage["steps"]:
                if stage["kind"] == "isotropic":
                    raise RuntimeError("Isotropic preparation exhausted its step budget before convergence")
                self._next_stage(self.rows[-1])
                continue
            amount = min(self.interval, stage["steps"] - self.stage_steps)
            rate = self.servo.rate(stage, self._servo_row(self.rows[-1]))
            rate = np.where(self.axes, rate, 0.0)
            self.engine.advance(amount, rate=tuple(rate), drag=stage.get("drag", 0),
                                wall_velocities=stage.get("wall_velocities"))
            self.stage_steps += amount
            chunks += 1
            row = self._row(self.engine.diagnostics(), stage=self.stage_index, rate=rate.tolist())
            origin = self.rows[0]
            row["work_balance_proxy_residual"] = (
                row["cell_work"] - origin["cell_work"] + row["wall_work"] - origin["wall_work"]
                - row["normal_contact_dissipation"] + origin["normal_contact_dissipation"]
                - row["friction_dissipation"] + origin["friction_dissipation"]
                - row["drag_dissipation"] + origin["drag_dissipation"]
                - row["kinetic_energy"] + origin["kinetic_energy"]
                - row["elastic_energy_proxy"] + origin["elastic_energy_proxy"])
            self.rows.append(row)
            if stage["kind"] == "isotropic":
                diagonal = np.diag(np.asarray(self._servo_row(row)["stress"]))[self.axes]
                pressure = np.max(np.abs(diagonal / stage["pressure"] - 1))
                ratio = row["kinetic_energy"] / max(row["elastic_energy_proxy"], 1e-30)
                ok = (pressure <= stage.get("pressure_tolerance", 0.01)
                      and ratio <= stage.get("kinetic_tolerance", 1e-5)
                      and row["unbalanced_force"] <= stage.get("force_tolerance", 1e-3))
                self.converged = self.converged + 1 if ok else 0
                if self.converged >= stage.get("converged_samples", 100):
                    self._next_stage(row)
            elif stage["kind"] == "triaxial" and self._axial_strain(row) >= stage.get("stop_strain", 0.1):
                self.stage_index += 1
                self.stage_steps = 0
            if max_chunks is not None and chunks >= max_chunks:
                return "paused"
        return "complete"
