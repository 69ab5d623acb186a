"""Initial-state path: SnapshotManager::LoadRawPoints (snapshotmanager.cpp:20-113).

The HDF5 reader is pinned on the reference's own input cloud
(build/geometries_3d/3d_200k.h5, read where the reference tree is mounted;
skipped elsewhere) by facts of that file that a decoding mistake breaks: the
point count and volume attribute, the 1000 grain ids, the bounding block
whose volume equals the attribute, and the Poisson-disk spacing (no two
points closer than a fraction of the mean spacing, which a mis-ordered or
mis-inflated chunk destroys). The geometry half is checked against the
placement rules written out by hand (block centred in x and z, bottom at
y = 2h, indenter start, ParticleVolume, fresh Fe/Jp) and the SetupType 2
liquid marking against a brute-force box test.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from pathlib import Path

import numpy as np
import pytest

from icempm_3dm_amd import capi, scenes, snapshot
from icempm_3dm_amd.capi import check, dptr, lib
from icempm_3dm_amd.hssoa import HostSideSOA

RAW = Path(__file__).resolve().parent / "golden" / "3d_200k.h5"
need_raw = pytest.mark.skipif(not RAW.exists(), reason="reference input cloud not mounted here")


@need_raw
def test_reader_decodes_reference_cloud():
    hs, vol = snapshot.read_raw_points(str(RAW))
    n = hs.size
    assert n == 202329
    assert vol == 3.75
    g = snapshot.grain_ids(hs)
    assert g.min() == 0 and g.max() == 999 and len(np.unique(g)) == 1000
    P = hs.lines()[capi.POSX:capi.POSX + 3, :n].T.copy()
    assert np.isfinite(P).all()
    lo, hi = P.min(0), P.max(0)
    # the cloud fills a 2.5 x 1.0 x 1.5 m block: its volume is the attribute
    assert np.allclose(lo, 0.0, atol=1e-4) and np.allclose(hi, [2.5, 1.0, 1.5], atol=1e-4)
    assert abs(np.prod(hi - lo) - vol) < 1e-3
    # Poisson-disk sampling: nearest neighbours are never much closer than
    # the mean spacing (a decoding error would scramble or duplicate points)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(P).query(P, k=2)
    mean_spacing = (vol / n) ** (1.0 / 3.0)
    assert d[:, 1].min() > 0.3 * mean_spacing
    # grains are spread over the block, not one id per chunk
    first = g[: n // 4]
    assert len(np.unique(first)) == 1000


@need_raw
def test_reader_size_query_and_capacity_error():
    n = C.c_int(0)
    v = C.c_double(0.0)
    check(lib().icempm_read_raw_points(os.fsencode(str(RAW)), None, 0, C.byref(n), C.byref(v)))
    assert n.value == 202329 and v.value == 3.75
    small = HostSideSOA(1000)
    rc = lib().icempm_read_raw_points(os.fsencode(str(RAW)), dptr(small.host_buffer), small.capacity, C.byref(n),
                                      C.byref(v))
    assert rc == capi.ERR_CAPACITY


def test_reader_rejects_missing_and_non_hdf5(tmp_path):
    n = C.c_int(0)
    rc = lib().icempm_read_raw_points(os.fsencode(str(tmp_path / "nope.h5")), None, 0, C.byref(n), None)
    assert rc == capi.ERR_INVALID_ARG and b"cannot open" in lib().icempm_last_error()
    f = tmp_path / "text.h5"
    f.write_bytes(b"not an hdf5 file" * 16)
    rc = lib().icempm_read_raw_points(os.fsencode(str(f)), None, 0, C.byref(n), None)
    assert rc == capi.ERR_INVALID_ARG and b"signature" in lib().icempm_last_error()
    # a truncated copy of a real file fails with an error, never a crash
    if RAW.exists():
        t = tmp_path / "trunc.h5"
        t.write_bytes(RAW.read_bytes()[:300000])
        rc = lib().icempm_read_raw_points(os.fsencode(str(t)), None, 0, C.byref(n), None)
        hs = HostSideSOA(202329)
        rc2 = lib().icempm_read_raw_points(os.fsencode(str(t)), dptr(hs.host_buffer), hs.capacity, C.byref(n),
                                           None)
        assert rc2 == capi.ERR_INVALID_ARG, (rc, rc2)


def _cloud(n=5000, box=(0.5, 0.25, 0.3), seed=3):
    rng = np.random.default_rng(seed)
    hs = HostSideSOA(n + 100)
    hs.size = n
    L = hs.lines()
    L[:, :] = 7.0  # garbage everywhere: prepare must leave a fresh state
    # like the reference's clouds, the block starts at the origin (the x/z
    # offset of LoadRawPoints centres [0, length], it does not subtract xmin)
    for a in range(3):
        L[capi.POSX + a, :n] = box[a] * rng.random(n)
        L[capi.POSX + a, 0] = 0.0
    L[capi.IDX_UTILITY, :n] = rng.integers(0, 1000, n).astype(np.uint64).view(np.float64)
    return hs


def _params(setup_type=0):
    p = capi.SimParams3D().Reset()
    p.GridXTotal, p.GridY, p.GridZ = 64, 32, 32
    p.DomainDimensionX = 1.0
    p.SetupType = setup_type
    p.IndDiameter = 0.2
    p.IndDepth = 0.05
    p.ComputeCamClayParams()
    p.ComputeHelperVariables()
    return p


@pytest.mark.parametrize("setup_type", [0, 1])
def test_prepare_places_block_and_indenter(setup_type):
    hs = _cloud()
    n = hs.size
    L = hs.lines()
    P0 = L[capi.POSX:capi.POSX + 3, :n].copy()
    util0 = L[capi.IDX_UTILITY, :n].copy()
    p = _params(setup_type)
    p.Volume = 0.5 * 0.25 * 0.3
    check(lib().icempm_prepare_raw_points(C.byref(p), dptr(hs.host_buffer), hs.capacity, n))
    h = p.cellsize
    lo, hi = P0.min(1), P0.max(1)
    # boundaries are stored before the offset, as the reference does
    assert (p.xmin, p.ymin, p.zmin, p.xmax, p.ymax, p.zmax) == (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2])
    P = L[capi.POSX:capi.POSX + 3, :n]
    off = np.array([(p.GridXTotal * h - (hi[0] - lo[0])) / 2, 2 * h - lo[1], (p.GridZ * h - (hi[2] - lo[2])) / 2])
    assert np.array_equal(P, P0 + off[:, None])
    # a block starting at the origin ends up centred in x and z, bottom on y = 2h
    assert lo[0] == 0.0 and lo[2] == 0.0
    assert abs(P[0].min() + P[0].max() - p.GridXTotal * h) < 1e-12
    assert abs(P[2].min() + P[2].max() - p.GridZ * h) < 1e-12
    assert abs(P[1].min() - 2 * h) < 1e-15
    # fresh state: Fe = I, Jp_inv = 1, everything else zero, grain ids kept
    F = L[capi.FE00:capi.FE00 + 9, :n]
    assert np.array_equal(F, np.tile(np.eye(3).reshape(9, 1), (1, n)))
    assert np.all(L[capi.IDX_JP_INV, :n] == 1.0)
    for line in [1, 2] + list(range(capi.VELX, capi.VELX + 3)) + list(range(capi.BP00, capi.BP00 + 9)):
        assert np.all(L[line, :n] == 0.0), line
    assert np.array_equal(L[capi.IDX_UTILITY, :n].view(np.uint64), util0.view(np.uint64))
    # indenter start (snapshotmanager.cpp:86-99)
    r = p.IndDiameter / 2
    ht = r - p.IndDepth
    block_left = off[0]
    block_top = hi[1] + off[1]
    assert p.indenter_x == math.floor((block_left - math.sqrt(r * r - ht * ht)) / h) * h
    want_y = block_top + ht if setup_type == 0 else math.ceil(block_top / h) * h
    assert p.indenter_y == want_y
    assert p.indenter_x_initial == p.indenter_x and p.indenter_y_initial == p.indenter_y
    assert p.ParticleVolume == p.Volume / n
    assert p.nPtsTotal == n
    assert p.ParticleMass == p.ParticleVolume * p.Density  # ComputeHelperVariables ran


def test_prepare_setup2_marks_liquid_outside_cutout():
    hs = _cloud(seed=9)
    n = hs.size
    L = hs.lines()
    P0 = L[capi.POSX:capi.POSX + 3, :n].copy()
    p = _params(2)
    p.BlockCutout[0], p.BlockCutout[1], p.BlockCutout[2] = 0.3, 0.1, 0.2
    p.Volume = 0.1
    check(lib().icempm_prepare_raw_points(C.byref(p), dptr(hs.host_buffer), hs.capacity, n))
    lo, hi = P0.min(1), P0.max(1)
    c = (lo + hi) / 2
    c[1] = hi[1] - 0.1 / 2
    dims = np.array([0.3, 0.1, 0.2])
    inside = np.all((P0 >= (c - dims / 2)[:, None]) & (P0 <= (c + dims / 2)[:, None]), axis=0)
    liquid = (L[capi.IDX_UTILITY, :n].view(np.int64) & capi.STATUS_LIQUID) != 0
    assert np.array_equal(liquid, ~inside)
    assert 0 < inside.sum() < n
