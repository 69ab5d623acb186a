"""Synthetic block-indentation scenes.

A scene is a block of ice on a jittered lattice resting on the sticky
bottom band of the grid, with a rigid indenter. Axes follow the reference's
internal frame: y is vertical (gravity on -y, sticky band y <= 2 nodes,
mpm_kernels.cuh:170, :210). Horizontal block placement and the bottom at
y = 2h follow SnapshotManager::LoadRawPoints (snapshotmanager.cpp:56-75).

Configurations (sizes in cells of h = 1/64 m):
  block_100k  grid 64x32x32,    block 25x20x25 cells, 2^3 points/cell -> 100,000
  block_2m    grid 256x128x128, block 100x50x50 cells, 2^3           -> 2,000,000
  block_16m   grid 512x256x256, block 200x100x100, 2^3 (also 3^3, 4^3 at
              134x67x66 and 100x50x50 cells)                          -> ~16,000,000
Material: density 917 kg/m^3, E 9e9 Pa, nu 0.3, Drucker-Prager friction
angle 60 deg, cohesion c = 2e6 Pa mapped to DP_threshold_p = c / tan(phi)
(the reference's DP line is q = (p + DP_threshold_p) tan(phi),
mpm_kernels.cuh:525). Indenter: sphere r = 0.1 m centred over the block,
initially touching the top face, moving down at 1 m/s.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, replace

import numpy as np

from . import capi
from .capi import SimParams3D
from .hssoa import HostSideSOA


@dataclass(frozen=True)
class BlockScene:
    name: str
    grid: tuple  # (GX, GY, GZ) nodes
    cells: tuple  # block size in cells (x, vertical y, z)
    ppc_axis: int = 2
    h: float = 1.0 / 64.0
    dt: float = 1e-5
    steps: int = 200
    jitter: float = 0.5  # fraction of the lattice spacing: +-0.25h at 2 points per cell-axis
    seed: int = 1
    n_grains: int = 1000
    density: float = 917.0
    youngs: float = 9e9
    poisson: float = 0.3
    phi_deg: float = 60.0
    cohesion: float = 2e6
    gravity: float = 9.81
    indenter_radius: float = 0.1
    indenter_speed: float = 1.0
    indenter_shape: int = capi.INDENTER_SPHERE

    @property
    def n_points(self) -> int:
        return self.cells[0] * self.cells[1] * self.cells[2] * self.ppc_axis ** 3

    def origin(self) -> tuple:
        gx, gy, gz = self.grid
        cx, cy, cz = self.cells
        h = self.h
        return ((gx - cx) / 2.0 * h, 2.0 * h, (gz - cz) / 2.0 * h)

    def indenter(self):
        """(shape, center at t=0, velocity): tangent to the top face, moving down."""
        ox, oy, oz = self.origin()
        cx, cy, cz = self.cells
        h = self.h
        top = oy + cy * h
        center = (ox + cx * h / 2.0, top + self.indenter_radius, oz + cz * h / 2.0)
        return self.indenter_shape, center, (0.0, -self.indenter_speed, 0.0)


BLOCK_100K = BlockScene("block_100k", (64, 32, 32), (25, 20, 25), steps=200)
BLOCK_2M = BlockScene("block_2m", (256, 128, 128), (100, 50, 50), dt=5e-6, steps=2000)
# the damage model of the 2M-point block run: full damage at 1% equivalent
# plastic shear strain, softening to 10% of the cohesion
BLOCK_2M_DAMAGE = {"softening_strain": 0.01, "residual_cohesion": 0.1}
BLOCK_16M = BlockScene("block_16m", (512, 256, 256), (200, 100, 100), steps=500)
BLOCK_16M_PPC27 = replace(BLOCK_16M, name="block_16m_ppc27", cells=(134, 67, 66), ppc_axis=3)
BLOCK_16M_PPC64 = replace(BLOCK_16M, name="block_16m_ppc64", cells=(100, 50, 50), ppc_axis=4)

SCENES = {s.name: s for s in (BLOCK_100K, BLOCK_2M, BLOCK_16M, BLOCK_16M_PPC27, BLOCK_16M_PPC64)}


def make_params(scene: BlockScene) -> SimParams3D:
    """SimParams3D for the scene: Reset defaults, then the scene's values and
    the derived constants (ComputeCamClayParams, ComputeHelperVariables)."""
    p = SimParams3D().Reset()
    gx, gy, gz = scene.grid
    p.GridXTotal, p.GridY, p.GridZ = gx, gy, gz
    p.DomainDimensionX = gx * scene.h
    p.InitialTimeStep = scene.dt
    p.YoungsModulus = scene.youngs
    p.PoissonsRatio = scene.poisson
    p.Density = scene.density
    p.Gravity = scene.gravity
    p.SetPhi(scene.phi_deg)
    p.DP_threshold_p = scene.cohesion / p.DP_tan_phi
    p.IndDiameter = 2.0 * scene.indenter_radius
    p.IndVelocity = scene.indenter_speed
    p.nPartitions = 1
    p.nPtsTotal = scene.n_points
    spacing = scene.h / scene.ppc_axis
    p.Volume = scene.n_points * spacing ** 3
    p.ParticleVolume = p.Volume / scene.n_points
    p.ComputeCamClayParams()
    p.ComputeHelperVariables()
    shape, center, vel = scene.indenter()
    p.indenter_x = p.indenter_x_initial = center[0]
    p.indenter_y = p.indenter_y_initial = center[1]
    ox, oy, oz = scene.origin()
    cx, cy, cz = scene.cells
    p.xmin, p.xmax = ox, ox + cx * scene.h
    p.ymin, p.ymax = oy, oy + cy * scene.h
    p.zmin, p.zmax = oz, oz + cz * scene.h
    return p


def make_points(scene: BlockScene, capacity: int | None = None) -> HostSideSOA:
    """Jittered lattice of the scene in a HostSideSOA (Fe = I, Jp_inv = 1)."""
    n = scene.n_points
    hs = HostSideSOA(capacity or n)
    cells = (C.c_int * 3)(*scene.cells)
    origin = (C.c_double * 3)(*scene.origin())
    out = C.c_int(0)
    capi.check(capi.lib().icempm_scene_block(capi.dptr(hs.host_buffer), hs.capacity, cells, scene.ppc_axis,
                                             origin, scene.h, scene.jitter, scene.seed, scene.n_grains,
                                             C.byref(out)))
    hs.size = out.value
    return hs


def scaled(scene: BlockScene, factor: float, name: str | None = None) -> BlockScene:
    """The same scene with the grid and block scaled (for quick tests)."""
    g = tuple(max(8, int(round(x * factor))) for x in scene.grid)
    c = tuple(max(1, int(round(x * factor))) for x in scene.cells)
    return replace(scene, name=name or f"{scene.name}_x{factor}", grid=g, cells=c)


def reference_cylinder_params(scene: BlockScene) -> SimParams3D:
    """Parameters for the reference's own indenter law on this block: a
    z-cylinder of diameter IndDiameter moving along +x, placed as
    SnapshotManager::LoadRawPoints does (snapshotmanager.cpp:82-96) for
    SetupType 0, with IndDepth = r / 2."""
    p = make_params(scene)
    h = p.cellsize
    r = p.IndDiameter / 2
    p.IndDepth = r / 2
    ht = r - p.IndDepth
    x_ind_offset = math.sqrt(r * r - ht * ht)
    block_left = p.xmin
    block_top = p.ymax
    p.indenter_x = p.indenter_x_initial = math.floor((block_left - x_ind_offset) / h) * h
    p.indenter_y = p.indenter_y_initial = block_top + ht
    return p


def perturbed_points(scene: BlockScene, seed: int = 7, amp_F: float = 2e-3, amp_v: float = 0.05,
                     amp_B: float = 5.0, crushed_fraction: float = 0.05,
                     liquid_fraction: float = 0.0) -> HostSideSOA:
    """Scene points with a non-trivial state for one-step parity tests:
    Fe = I + noise, random velocities and affine matrices, some points
    pre-crushed (so the return mapping runs) and, optionally, liquid."""
    hs = make_points(scene)
    rng = np.random.default_rng(seed)
    n = hs.size
    L = hs.lines()
    L[capi.FE00:capi.FE00 + 9, :n] += amp_F * rng.standard_normal((9, n))
    L[capi.VELX:capi.VELX + 3, :n] = amp_v * rng.standard_normal((3, n))
    L[capi.BP00:capi.BP00 + 9, :n] = amp_B * rng.standard_normal((9, n))
    L[capi.IDX_JP_INV, :n] = 1.0 + 0.01 * rng.standard_normal(n)
    util = hs.utility()
    u = rng.random(n)
    util[:n][u < crushed_fraction] |= capi.STATUS_CRUSHED
    if liquid_fraction > 0:
        util[:n][(u >= crushed_fraction) & (u < crushed_fraction + liquid_fraction)] |= capi.STATUS_LIQUID
    return hs
