"""Host side of the batched ensemble: the parameter-sweep generator.

Checked against its specification: particle counts log-uniform in
[20k, 500k] (all inside the range up to lattice rounding, and the median of
64 draws near the geometric mean), aspect ratios in [1, 3], every block and
its indenter inside the 128x64x64 sub-grid with the walls' margin, material
and speed draws inside their ranges, one time step and cell size for all
scenes (the ensemble's requirement), and the same scenes for the same seed.
"""
from __future__ import annotations

import math

import numpy as np

from icempm_3dm_amd import scenes
from icempm_3dm_amd.ensemble import sweep_scenes


def test_sweep_ranges_and_fit():
    sw = sweep_scenes(64, seed=2024)
    assert len(sw) == 64
    counts = np.array([s.n_points for s in sw])
    assert counts.min() >= 0.85 * 20_000 and counts.max() <= 1.15 * 500_000
    assert abs(math.log(np.median(counts)) - 0.5 * math.log(20_000 * 500_000)) < 0.6
    for s in sw:
        gx, gy, gz = s.grid
        cx, cy, cz = s.cells
        assert (gx, gy, gz) == (128, 64, 64)
        assert 0.8 <= cx / math.sqrt(cy * cz) <= 3.3
        ox, oy, oz = s.origin()
        assert ox >= 2 * s.h and oz >= 2 * s.h and abs(oy - 2 * s.h) < 1e-15
        assert ox + cx * s.h <= (gx - 2) * s.h and oz + cz * s.h <= (gz - 2) * s.h
        _, center, vel = s.indenter()
        assert center[1] + s.indenter_radius <= (gy - 2) * s.h
        assert vel[1] < 0 and 0.5 <= -vel[1] <= 2.0
        assert 5e9 <= s.youngs <= 1e10 and 40 <= s.phi_deg <= 70 and 0.5e6 <= s.cohesion <= 4e6
        assert s.dt == 2e-6 and s.h == 1 / 64
    assert {s.seed for s in sw} == {2024 * 1000 + i for i in range(64)}


def test_sweep_is_deterministic_and_params_build():
    a, b = sweep_scenes(8, seed=5), sweep_scenes(8, seed=5)
    assert a == b
    assert a != sweep_scenes(8, seed=6)
    p = scenes.make_params(a[0])
    assert p.GridXTotal == 128 and p.InitialTimeStep == 2e-6 and p.YoungsModulus == a[0].youngs
