"""Centreline extraction and reference-table errors of the lid-driven cavity,
as the reference's validation harness computes them
(validation/cases/cavity.hpp:24-66, 111-150)."""
from __future__ import annotations

import bisect
import csv
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / "golden" / "cavity"

# The reference's serial CPU backend on these cavities (FV_CAVITY_REFINEMENT /
# FV_CAVITY_RESULT lines of validation/results/performance_scale/raw/
# slow_baseline_63287b9.log and colab_t4/logs/tests_slow_v3_re100_400.log).
REFERENCE_65_RE1000 = {"iterations": 600, "u_error": 0.0207536182404, "v_error": 0.0309673490243,
                       "u_center": -0.0639748126268, "v_center": 0.0242851041257}
REFERENCE_129_RE100 = {"iterations": 1100, "ghia_u": 0.005044, "ghia_v": 0.008972}


def interpolate(profile, x):
    """cavity.hpp:24-32: linear interpolation, clamped at the ends."""
    xs = [p[0] for p in profile]
    i = bisect.bisect_left(xs, x)
    if i == 0:
        return profile[0][1]
    if i == len(profile):
        return profile[-1][1]
    (x0, y0), (x1, y1) = profile[i - 1], profile[i]
    return y0 + (x - x0) / (x1 - x0) * (y1 - y0)


def centerline(centroids: np.ndarray, value: np.ndarray, vertical: bool):
    """cavity.hpp:34-47: per cell row (vertical) or column, the value
    interpolated to the centre line, plus the wall values at both ends."""
    lines: dict[int, list] = {}
    for c in range(len(value)):
        p = centroids[c]
        key = round((p[1] if vertical else p[0]) * 1e10)
        lines.setdefault(key, []).append((p[0] if vertical else p[1], float(value[c])))
    out = [(key / 1e10, interpolate(sorted(lines[key]), 0.5)) for key in sorted(lines)]
    return [(0.0, 0.0)] + out + [(1.0, 1.0 if vertical else 0.0)]


def table(name):
    with open(DATA / name) as f:
        rows = list(csv.reader(f))[1:]
    return [[float(v) for v in r] for r in rows if r]


def reference_error(profile, name, coordinate, column):
    """cavity.hpp:61-69: the largest deviation from a reference table."""
    return max(abs(interpolate(profile, r[coordinate]) - r[column]) for r in table(name))


def run_to_steady(solver, iterate, report, fields, tolerance, maximum=10000):
    """cavity.hpp:111-140: 100 iterations at a time until the scaled momentum
    defects are below 5e-7, the largest velocity update over the block is
    below `tolerance` and the corrected mass imbalance is below 1e-5."""
    old = None
    iteration = 0
    while iteration < maximum:
        iterate(min(100, maximum - iteration))
        iteration += 100
        r = report()
        u, v = fields()
        update = 1.0
        if old is not None:  # float differences, as the harness forms them
            update = float(max(np.abs(u - old[0]).max(), np.abs(v - old[1]).max()))
        steady = max(r["u_scaled"], r["v_scaled"]) < 5e-7 and update < tolerance and \
            r["corrected_mass_imbalance"] < 1e-5
        if steady:
            return r, u, v, True
        old = (u, v)
    return r, u, v, False
