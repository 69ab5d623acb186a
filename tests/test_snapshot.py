"""Snapshot save / read (SnapshotManager::SaveSnapshot / ReadSnapshot,
snapshotmanager.cpp:299-405) without an HDF5 library.

Pinned by: bit-exact round trips of the whole host buffer (denormal utility
bit patterns included), the indenter accumulators and every saved parameter,
in the contiguous and the chunked + deflated layout (with more chunks than
one B-tree node holds, so the two-level index is exercised); the derived
parameters against ComputeCamClayParams / ComputeHelperVariables; and the
file's fixed structures against the same structures of a libhdf5-written
file from the reference tree (superblock fields, root group B-tree and heap
placement), where that file is mounted.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np
import pytest

from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.capi import check, dptr, lib
from icempm_3dm_amd.hssoa import HostSideSOA

RAW = Path(__file__).resolve().parent / "golden" / "3d_200k.h5"

INT_FIELDS = ["SetupType", "nPtsTotal", "GridXTotal", "GridY", "GridZ", "IndenterSubdivisions", "UpdateEveryNthStep",
              "SnapshotPeriod", "SimulationStep", "nPartitions", "GridHaloSize", "PointTransferFrequency"]
DOUBLE_FIELDS = ["cellsize", "InitialTimeStep", "SimulationEndTime", "SimulationTime", "Gravity", "Density",
                 "PoissonsRatio", "YoungsModulus", "IceCompressiveStrength", "IceTensileStrength",
                 "IceTensileStrength2", "IceShearStrength", "DP_tan_phi", "DP_threshold_p", "GrainVariability",
                 "indenter_x", "indenter_y", "indenter_x_initial", "indenter_y_initial", "IndDiameter", "IndVelocity",
                 "ParticleVolume", "Volume", "ExtraSpaceForIncomingPoints", "PointsTransferBufferFraction",
                 "RebalanceThresholdFreeSpaceRemaining", "RebalanceThresholdDisabledPercentage",
                 "animation_threshold_pos"]


def _state(capacity=3000, size=2500, seed=1):
    rng = np.random.default_rng(seed)
    hs = HostSideSOA(capacity)
    hs.size = size
    L = hs.lines()
    L[:, :] = rng.standard_normal(L.shape)
    util = rng.integers(0, 1000, capacity).astype(np.int64) | np.where(rng.random(capacity) < 0.3, 0x10000, 0)
    L[capi.IDX_UTILITY, :] = util.view(np.float64)  # denormal doubles
    L[capi.BP00:capi.BP00 + 9, :] = 0.0  # a fresh block: C = 0 (compressible)
    L[:, size:] = 0.0
    p = scenes.make_params(scenes.BLOCK_100K)
    p.SimulationStep = 1234
    p.SimulationTime = 0.01234
    p.DP_threshold_p = 3.21e6
    p.SetupType = 1
    p.nPartitions = 1
    p.animation_threshold_pos = 0.125
    bins = rng.standard_normal(p.IndenterArraySize())
    return hs, p, bins


def _save(path, p, hs, bins, compress):
    check(lib().icempm_save_snapshot(os.fsencode(str(path)), C.byref(p), dptr(hs.host_buffer), hs.capacity, hs.size,
                                     dptr(bins), int(compress)))


def _read(path, capacity=None):
    cap, size = C.c_uint(0), C.c_uint(0)
    p = capi.SimParams3D()
    check(lib().icempm_read_snapshot(os.fsencode(str(path)), C.byref(p), None, 0, C.byref(cap), C.byref(size), None, 0))
    hs = HostSideSOA(capacity or cap.value)
    hs.size = size.value
    bins = np.zeros(p.IndenterArraySize())
    check(lib().icempm_read_snapshot(os.fsencode(str(path)), None, dptr(hs.host_buffer), hs.capacity, C.byref(cap),
                                     C.byref(size), dptr(bins), bins.size))
    return hs, p, bins, cap.value


@pytest.mark.parametrize("compress", [False, True])
def test_round_trip_bit_exact(tmp_path, compress):
    hs, p, bins = _state()
    f = tmp_path / "d00007.h5"
    _save(f, p, hs, bins, compress)
    back, q, bins2, cap = _read(f)
    assert cap == hs.capacity and back.size == hs.size
    assert np.array_equal(back.host_buffer.view(np.int64), hs.host_buffer.view(np.int64))
    assert np.array_equal(bins2.view(np.int64), bins.view(np.int64))
    for name in INT_FIELDS:
        if name in ("UpdateEveryNthStep", "IndenterSubdivisions"):
            continue  # recomputed by ComputeHelperVariables, as in the reference
        assert getattr(q, name) == getattr(p, name), name
    for name in DOUBLE_FIELDS:
        assert getattr(q, name) == getattr(p, name), name
    # derived values: what ReadParametersFromAttributes leaves after
    # DomainDimensionX = GridXTotal * cellsize, ComputeCamClayParams, ComputeHelperVariables
    want = p.copy()
    want.DomainDimensionX = want.GridXTotal * want.cellsize
    want.ComputeCamClayParams()
    want.ComputeHelperVariables()
    for name in ("DomainDimensionX", "lambda_", "mu", "kappa", "NACC_beta", "ParticleMass", "cellsize_inv",
                 "dt_vol_Dpinv", "vmax", "IndRSq", "UpdateEveryNthStep", "IndenterSubdivisions"):
        assert getattr(q, name) == getattr(want, name), name
    if compress:
        assert f.stat().st_size < 0.8 * hs.host_buffer.nbytes
    else:
        assert f.stat().st_size >= hs.host_buffer.nbytes


def test_chunk_index_with_two_levels(tmp_path):
    # 100-element chunks over 300 x 28 values: 84 chunks, more than the 64 a
    # chunk B-tree node holds, so the index has an internal level
    hs, p, bins = _state(capacity=300, size=100, seed=2)
    f = tmp_path / "d00001.h5"
    _save(f, p, hs, bins, True)
    data = f.read_bytes()
    levels = {data[m + 5] for m in range(len(data) - 8) if data[m:m + 4] == b"TREE" and data[m + 4] == 1}
    assert levels == {0, 1}
    back, _, _, _ = _read(f)
    assert np.array_equal(back.host_buffer.view(np.int64), hs.host_buffer.view(np.int64))


def test_read_into_larger_buffer_repitches(tmp_path):
    hs, p, bins = _state(capacity=500, size=400)
    f = tmp_path / "d00002.h5"
    _save(f, p, hs, bins, False)
    back, _, _, cap = _read(f, capacity=800)
    assert cap == 500
    B, L = back.lines(), hs.lines()
    assert np.array_equal(B[:, :500].view(np.int64), L[:, :500].view(np.int64))
    assert np.all(B[:, 500:] == 0)


def test_errors(tmp_path):
    hs, p, bins = _state(capacity=100, size=50)
    f = tmp_path / "d00003.h5"
    _save(f, p, hs, bins, False)
    cap, size = C.c_uint(0), C.c_uint(0)
    small = HostSideSOA(50)
    rc = lib().icempm_read_snapshot(os.fsencode(str(f)), None, dptr(small.host_buffer), small.capacity,
                                    C.byref(cap), C.byref(size), None, 0)
    assert rc == capi.ERR_CAPACITY
    big = HostSideSOA(100)
    short_bins = np.zeros(10)
    rc = lib().icempm_read_snapshot(os.fsencode(str(f)), None, dptr(big.host_buffer), big.capacity,
                                    C.byref(cap), C.byref(size), dptr(short_bins), short_bins.size)
    assert rc == capi.ERR_CAPACITY
    t = tmp_path / "trunc.h5"
    t.write_bytes(f.read_bytes()[:2000])
    rc = lib().icempm_read_snapshot(os.fsencode(str(t)), None, None, 0, C.byref(cap), C.byref(size), None, 0)
    assert rc == capi.ERR_INVALID_ARG
    rc = lib().icempm_read_snapshot(os.fsencode(str(RAW)) if RAW.exists() else os.fsencode(str(t)), None, None, 0,
                                    C.byref(cap), C.byref(size), None, 0)
    assert rc == capi.ERR_INVALID_ARG  # no "Points" dataset / truncated


def test_derived_parameters_saturate_for_absurd_grids(tmp_path):
    """A file's grid sizes go into the reference's int product GridZ * GridY *
    GridHaloSize (gbOffset) and into float-to-int conversions; values no real
    grid has (here 2^30 x 2^10 nodes) saturate at the int range instead of
    overflowing, and the rest of the file reads as written. (A host-only
    AddressSanitizer / UBSan build of the reader also took thousands of
    randomly damaged files without a report.)"""
    hs, p, bins = _state(capacity=100, size=50)
    p.GridY = 1 << 30
    p.GridZ = 1 << 10
    p.GridHaloSize = 15
    f = tmp_path / "d00004.h5"
    _save(f, p, hs, np.zeros(10), False)
    cap, size = C.c_uint(0), C.c_uint(0)
    q = capi.SimParams3D()
    check(lib().icempm_read_snapshot(os.fsencode(str(f)), C.byref(q), None, 0, C.byref(cap), C.byref(size), None, 0))
    assert (q.GridY, q.GridZ, q.GridHaloSize) == (1 << 30, 1 << 10, 15)
    assert q.gbOffset == 2**31 - 1
    assert (cap.value, size.value) == (100, 50)


@pytest.mark.skipif(not RAW.exists(), reason="reference input cloud not mounted here")
def test_fixed_structures_match_a_libhdf5_file(tmp_path):
    hs, p, bins = _state(capacity=100, size=50)
    f = tmp_path / "d00004.h5"
    _save(f, p, hs, bins, False)
    ours, theirs = f.read_bytes(), RAW.read_bytes()
    # superblock: signature, versions, offset/length sizes, group K values,
    # flags, base address, undefined free-space and driver addresses
    assert ours[:24] == theirs[:24]
    assert ours[24:40] == theirs[24:40]
    assert ours[48:56] == theirs[48:56]
    # root symbol table entry: header at 96, cache type 1 with the B-tree at
    # 136 and the local heap right after a full-size group node, at 680
    assert ours[56:96] == theirs[56:96]
    # root object header: one symbol table message pointing there
    assert ours[96:136] == theirs[96:136]
    # group B-tree node prefix: TREE, type 0, level 0, one entry, no siblings
    assert ours[136:160] == theirs[136:160]
    assert ours[680:684] == b"HEAP" and theirs[680:684] == b"HEAP"
