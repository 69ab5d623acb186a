"""The C-ABI library loads without a GPU, exports every declared entry
point, and its host-only functions (parameters, scene generator) behave as
the reference's SimParams3D / HostSideSOA code they restate."""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.capi import SimParams3D

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 36
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ELF dynamic table agrees (no stale Python-side prototypes)
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported


def test_params_layout_matches_reference_struct():
    # sizeof(SimParams3D) on x86-64 for the member list of parameters_sim_3d.h:40-90
    assert C.sizeof(SimParams3D) == 520
    assert capi.lib().icempm_params_sizeof() == 520
    off = {name: getattr(SimParams3D, name).offset for name, _ in SimParams3D._fields_}
    assert off["SetupType"] == 0 and off["DomainDimensionX"] == 24 and off["Gravity"] == 72
    assert off["cellsize"] == 240 and off["indenter_x"] == 352 and off["nPartitions"] == 392
    assert off["dt_vol_Dpinv"] == 440 and off["gbOffset"] == 472 and off["cellsize_inv"] == 480
    assert off["tpb_P2G"] == 496 and off["animation_threshold_pos"] == 512


def test_tuning_layout_matches_header():
    # icempm_Tuning: four ints (points_per_chunk in the padding before the
    # double), then ms_per_substep at 16
    off = {name: getattr(capi.Tuning, name).offset for name, _ in capi.Tuning._fields_}
    assert C.sizeof(capi.Tuning) == 24
    assert off == {"fuse": 0, "sort_period": 4, "blocks_per_cu": 8, "points_per_chunk": 12, "ms_per_substep": 16}
    header = (ROOT / "include" / "icempm_3dm.h").read_text()
    body = header[header.index("typedef struct icempm_Tuning"):header.index("} icempm_Tuning;")]
    assert [ln.split()[-1].rstrip(";") for ln in body.splitlines()[1:] if ln.strip()] == list(off)


def test_params_reset_and_derived_constants():
    p = SimParams3D().Reset()
    # Reset defaults (parameters_sim_3d.cpp:5-62)
    assert (p.GridXTotal, p.GridY, p.GridZ) == (128, 55, 55)
    assert p.InitialTimeStep == 3e-5 and p.YoungsModulus == 5e8 and p.nPartitions == 4
    assert p.GridHaloSize == 15 and (p.tpb_P2G, p.tpb_Upd, p.tpb_G2P) == (256, 512, 128)
    assert p.DP_tan_phi == pytest.approx(1.0, abs=1e-15) and p.DP_threshold_p == 1e2
    # Lame (closed forms)
    E, nu = p.YoungsModulus, p.PoissonsRatio
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    assert p.lambda_ == lam and p.mu == mu and p.kappa == mu * 2.0 / 3.0 + lam
    assert p.NACC_beta == p.IceTensileStrength / p.IceCompressiveStrength
    # helper variables
    assert p.cellsize == p.DomainDimensionX / p.GridXTotal
    assert p.Dp_inv == 4.0 / (p.cellsize * p.cellsize)
    assert p.vmax == 0.5 * p.cellsize / p.InitialTimeStep
    assert p.UpdateEveryNthStep == int(1.0 / (200 * p.InitialTimeStep))
    assert p.IndenterSubdivisions == int(math.pi * p.IndDiameter / p.cellsize)
    assert p.gbOffset == p.GridZ * p.GridY * p.GridHaloSize
    p.SetPhi(60)
    assert p.DP_tan_phi == pytest.approx(math.sqrt(3), rel=1e-15)


def test_block_100k_scene_parameters():
    sc = scenes.BLOCK_100K
    p = scenes.make_params(sc)
    assert sc.n_points == 100_000
    h = 1 / 64
    assert p.cellsize == h and (p.GridXTotal, p.GridY, p.GridZ) == (64, 32, 32)
    assert p.ParticleMass == pytest.approx(917 * (h / 2) ** 3, rel=1e-15)
    assert p.DP_threshold_p == pytest.approx(2e6 / math.sqrt(3), rel=1e-14)
    assert p.mu == pytest.approx(9e9 / 2.6, rel=1e-15) and p.kappa == pytest.approx(7.5e9, rel=1e-15)
    # IndDiameter 0.2 m at h = 1/64: (int)(pi * 0.2 * 64) = 40 sensor columns
    assert p.IndenterSubdivisions == 40
    shape, center, vel = sc.indenter()
    assert shape == capi.INDENTER_SPHERE and vel == (0.0, -1.0, 0.0)
    assert center[1] - 0.1 == pytest.approx(p.ymax)


def test_scene_generator_matches_splitmix_restatement():
    """The lattice + SplitMix64 jitter of icempm_scene_block, restated in numpy."""
    sc = scenes.scaled(scenes.BLOCK_100K, 0.25)
    hs = scenes.make_points(sc)
    n = sc.n_points
    assert hs.size == n

    def mix(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        key = mix(np.uint64(sc.seed))
        idx = np.arange(n, dtype=np.uint64)
        q = sc.ppc_axis
        nx, ny, nz = (c * q for c in sc.cells)
        lat = [idx // np.uint64(nz * ny), (idx // np.uint64(nz)) % np.uint64(ny), idx % np.uint64(nz)]
        s = sc.h / q
        o = sc.origin()
        for d in range(3):
            u = (mix(key + np.uint64(4) * idx + np.uint64(d)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
            want = o[d] + (lat[d].astype(np.float64) + 0.5) * s + (2.0 * u - 1.0) * sc.jitter * s
            assert np.array_equal(hs.lines()[capi.POSX + d, :n], want)
        ug = (mix(key + np.uint64(4) * idx + np.uint64(3)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        grains = np.minimum((ug * sc.n_grains).astype(np.int64), sc.n_grains - 1)
    assert np.array_equal(hs.utility()[:n], grains)
    L = hs.lines()
    assert np.all(L[capi.FE00, :n] == 1) and np.all(L[capi.FE00 + 1, :n] == 0) and np.all(L[capi.IDX_JP_INV, :n] == 1)


def test_gpu_entry_points_fail_cleanly_without_a_device(gpu_available):
    if gpu_available:
        pytest.skip("a GPU is visible")
    ctx = C.c_void_p()
    rc = capi.lib().icempm_create(0, C.byref(ctx))
    assert rc == capi.ERR_NO_DEVICE and not ctx
    assert b"no HIP device" in capi.lib().icempm_last_error()
    assert capi.lib().icempm_step(None, 1, None, 0) == capi.ERR_INVALID_ARG


def test_cpp_wrapper_header_compiles():
    """include/icempm_gpu_implementation6.hpp (the C++ mirror of GPU_Implementation6) builds
    against the C-ABI with a plain host compiler."""
    src = ROOT / "tests" / "cpp" / "wrapper_smoke.cpp"
    out = Path("/tmp") / "icempm_wrapper_smoke.o"
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}", "-c", str(src),
                          "-o", str(out)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
