"""Batched ensemble of independent block-indentation scenes.

The reference runs one scene per process (Model3D owns one parameter set and
one HostSideSOA, simulation/model_3d.h:22-60); a material-parameter sweep is
as many runs. Here every scene of a sweep occupies a slot of whole 4^3-node
bricks of one stored grid, slots one empty brick apart, so the same segment
and brick launches that advance one scene advance all of them (scene lookup
per brick from a per-slot parameter table on the device). Each scene keeps
its own grid size, walls, material, indenter and force accumulators.

``sweep_scenes`` draws the material sweep used for throughput runs: particle
counts log-uniform in [20k, 500k] on the 2-points-per-cell-axis jittered
lattice, block aspect ratio (x : y : z = a : 1 : 1) uniform in [1, 3], a
128x64x64 sub-grid of h = 1/64 m per scene, E uniform in [5e9, 1e10] Pa,
friction angle in [40, 70] deg, cohesion in [0.5e6, 4e6] Pa and indenter
speed in [0.5, 2] m/s.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import replace

import numpy as np

from . import capi, scenes
from .capi import SimParams3D, check, dptr, lib
from .hssoa import HostSideSOA
from .scenes import BlockScene


def sweep_scenes(n_scenes: int = 64, seed: int = 2024, dt: float = 2e-6, grid=(128, 64, 64),
                 n_range=(20_000, 500_000), aspect_range=(1.0, 3.0), youngs=(5e9, 1e10), phi=(40.0, 70.0),
                 cohesion=(0.5e6, 4e6), speed=(0.5, 2.0)) -> list[BlockScene]:
    """Scenes of a parameter sweep, drawn from numpy's PCG64 with `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    gx, gy, gz = grid
    for i in range(n_scenes):
        n_target = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
        a = rng.uniform(*aspect_range)
        side = (n_target / 8.0 / a) ** (1.0 / 3.0)
        cx = min(gx - 8, max(1, int(round(a * side))))
        cy = min(gy - 24, max(1, int(round(side))))
        cz = min(gz - 8, max(1, int(round(side))))
        out.append(BlockScene(f"sweep_{i:03d}", (gx, gy, gz), (cx, cy, cz), ppc_axis=2, dt=dt, steps=500,
                              seed=seed * 1000 + i, youngs=float(rng.uniform(*youngs)),
                              phi_deg=float(rng.uniform(*phi)), cohesion=float(rng.uniform(*cohesion)),
                              indenter_speed=float(rng.uniform(*speed))))
    return out


class BatchedEnsemble:
    """Scenes advanced together on one device through icempm_allocate_ensemble."""

    def __init__(self, block_scenes: list[BlockScene], device: int = 0, capacity: int | None = None,
                 clouds: list[HostSideSOA] | None = None, params: list[SimParams3D] | None = None, indenters=None):
        """clouds / params / indenters (shape, center, velocity per scene)
        default to the scenes' own lattice, parameters and sphere."""
        self.scenes = list(block_scenes)
        self.params = list(params) if params is not None else [scenes.make_params(s) for s in self.scenes]
        self.n_scenes = len(self.scenes)
        if clouds is None:
            clouds = [scenes.make_points(s) for s in self.scenes]
        if indenters is None:
            indenters = [s.indenter() for s in self.scenes]
        self.first = np.zeros(self.n_scenes + 1, dtype=np.int32)
        self.first[1:] = np.cumsum([c.size for c in clouds])
        n = int(self.first[-1])
        self.hssoa = HostSideSOA(max(capacity or n, n))
        L = self.hssoa.lines()
        for i, c in enumerate(clouds):
            L[:, self.first[i]:self.first[i + 1]] = c.lines()[:, :c.size]
        self.hssoa.size = n
        self.ctx = C.c_void_p()
        check(lib().icempm_create(int(device), C.byref(self.ctx)))
        arr = (SimParams3D * self.n_scenes)(*self.params)
        check(lib().icempm_allocate_ensemble(self.ctx, arr, self.n_scenes, self.hssoa.capacity))
        for i, (shape, center, vel) in enumerate(indenters):
            ind = capi.Indenter.make(shape, center, vel)
            check(lib().icempm_ensemble_set_indenter(self.ctx, i, C.byref(ind)))
        self.n_bins = sum(p.IndenterArraySize() for p in self.params)
        self.simulation_time = 0.0
        self.simulation_step = 0
        self.transfer_points_to_device()

    def _read_layout(self) -> None:
        """Each scene's node offset in the stored grid and first accumulator."""
        self.bins_offset = []
        self.origin = []
        for i in range(self.n_scenes):
            o = (C.c_int * 3)()
            off = C.c_longlong(0)
            check(lib().icempm_ensemble_layout(self.ctx, i, o, C.byref(off)))
            self.origin.append(tuple(o))
            self.bins_offset.append(off.value)

    @property
    def n_points(self) -> int:
        return int(self.first[-1])

    def transfer_points_to_device(self) -> None:
        first = (C.c_int * (self.n_scenes + 1))(*[int(x) for x in self.first])
        check(lib().icempm_ensemble_transfer_points(self.ctx, dptr(self.hssoa.host_buffer), self.hssoa.capacity,
                                                    first))
        self._read_layout()  # the upload places each scene in its slot

    def set_fusion(self, on: bool) -> None:
        check(lib().icempm_set_fusion(self.ctx, 1 if on else 0))

    def set_points_per_chunk(self, points_per_chunk: int) -> None:
        check(lib().icempm_set_points_per_chunk(self.ctx, int(points_per_chunk)))

    def set_sort_period(self, period: int) -> None:
        """Regroup the points by grid brick every `period` substeps (0: never)."""
        check(lib().icempm_set_sort_period(self.ctx, int(period)))

    def set_damage(self, damage: "capi.Damage | None") -> None:
        """The damage model of the return mapping (None: off)."""
        d = damage if damage is not None else capi.Damage(0, 1.0, 1.0)
        check(lib().icempm_set_damage(self.ctx, C.byref(d)))

    def set_graph(self, on: bool) -> None:
        """Replay each run of substeps between re-sorts from a captured hipGraph."""
        check(lib().icempm_set_graph(self.ctx, 1 if on else 0))

    def graph_captures(self) -> tuple[int, int]:
        """(graphs captured so far, graphs kept)."""
        n = C.c_longlong(0)
        k = C.c_int(0)
        check(lib().icempm_graph_captures(self.ctx, C.byref(n), C.byref(k)))
        return n.value, k.value

    def autotune(self, trial_substeps: int = 10) -> dict:
        """icempm_autotune over all scenes together (the points are left as they are)."""
        t = capi.Tuning()
        check(lib().icempm_autotune(self.ctx, int(trial_substeps), C.byref(t)))
        return t.as_dict()

    def reset_indenter_force_accumulator(self) -> None:
        check(lib().icempm_reset_indenter_force_accumulator(self.ctx))

    def step(self, n_substeps: int) -> float:
        """n_substeps substeps of every scene; returns the new simulation time."""
        t = C.c_double(self.simulation_time)
        check(lib().icempm_step(self.ctx, int(n_substeps), C.byref(t), int(self.simulation_step)))
        self.simulation_time = t.value
        self.simulation_step += n_substeps
        return t.value

    def synchronize(self) -> None:
        check(lib().icempm_synchronize(self.ctx))

    def transfer_from_device(self) -> np.ndarray:
        """Points back (scene coordinates, upload order); returns the raw
        accumulators of all scenes concatenated in scene order."""
        n = C.c_int(0)
        raw = np.zeros(self.n_bins)
        check(lib().icempm_transfer_from_device(self.ctx, dptr(self.hssoa.host_buffer), self.hssoa.capacity, 0,
                                                C.byref(n), dptr(raw)))
        self.hssoa.size = n.value
        return raw

    def scene_points(self, i: int) -> np.ndarray:
        """(28, n_i) view of scene i's points in the host buffer."""
        return self.hssoa.lines()[:, self.first[i]:self.first[i + 1]]

    def scene_bins(self, raw: np.ndarray, i: int) -> np.ndarray:
        off = self.bins_offset[i]
        return raw[off:off + self.params[i].IndenterArraySize()]

    def indenter_forces(self, raw: np.ndarray) -> np.ndarray:
        """(n_scenes, 3): each scene's accumulated force divided by its
        UpdateEveryNthStep (gpu_implementation6.cpp:499-510)."""
        return np.stack([self.scene_bins(raw, i).reshape(-1, 3).sum(axis=0) / float(self.params[i].UpdateEveryNthStep)
                         for i in range(self.n_scenes)])

    def stored_grid_dims(self) -> tuple[int, int, int]:
        """Nodes per axis of the stored grid that holds every scene's slot:
        slots of ceil((largest scene grid + 4) / 4) bricks per axis, filled z
        first, then y, then x, at most 1024 nodes per axis (the rule of
        icempm_allocate_ensemble)."""
        maxg = [max(getattr(p, a) for p in self.params) for a in ("GridXTotal", "GridY", "GridZ")]
        slot_b = [(g + 4 + 3) // 4 for g in maxg]
        maxslots = [1024 // (b * 4) for b in slot_b]
        n = self.n_scenes
        nz = min(n, maxslots[2])
        ny = min((n + nz - 1) // nz, maxslots[1])
        nx = (n + ny * nz - 1) // (ny * nz)
        return slot_b[0] * nx * 4, slot_b[1] * ny * 4, slot_b[2] * nz * 4

    def download_grid(self) -> np.ndarray:
        """The stored grid as (4, X, Y, Z): mass and momentum (or velocity
        after the grid update) of every node of every slot."""
        dims = self.stored_grid_dims()
        out = np.zeros(4 * dims[0] * dims[1] * dims[2])
        check(lib().icempm_download_grid(self.ctx, dptr(out)))
        return out.reshape(4, *dims)

    def scene_grid(self, grid: np.ndarray, i: int) -> np.ndarray:
        """Scene i's own nodes (4, GridXTotal, GridY, GridZ) of a stored grid."""
        o, p = self.origin[i], self.params[i]
        return grid[:, o[0]:o[0] + p.GridXTotal, o[1]:o[1] + p.GridY, o[2]:o[2] + p.GridZ]

    def read_timings(self) -> dict:
        ms = np.zeros(len(capi.PHASES))
        cnt = np.zeros(len(capi.PHASES), dtype=np.int64)
        check(lib().icempm_read_timings(self.ctx, dptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_longlong))))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(capi.PHASES)}

    def close(self) -> None:
        if self.ctx:
            lib().icempm_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def with_dt(block_scenes: list[BlockScene], dt: float) -> list[BlockScene]:
    return [replace(s, dt=dt) for s in block_scenes]
