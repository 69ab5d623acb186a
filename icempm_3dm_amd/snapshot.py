"""SnapshotManager mirror (simulation/snapshotmanager.{h,cpp}): the initial
state from a raw HDF5 point cloud, and snapshot save / resume.

``LoadRawPoints`` follows SnapshotManager::LoadRawPoints
(snapshotmanager.cpp:20-113): read the cloud, place the block and the
indenter, fresh point state, then allocate, upload, sort by cell and
Reset/Prepare the model. The file is read and placed by the library
(``icempm_read_raw_points``, ``icempm_prepare_raw_points``); the sort runs on
the device (``icempm_remove_disabled_and_sort``), so the sorted order becomes
the upload order as after the reference's host sort.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi
from .capi import check, dptr, lib
from .hssoa import HostSideSOA
from .model import Model3D


def read_raw_points(path: str, capacity: int | None = None) -> tuple[HostSideSOA, float]:
    """Reads the cloud into a fresh HostSideSOA (capacity defaults to the
    point count); returns (hssoa, volume attribute)."""
    n = C.c_int(0)
    vol = C.c_double(0.0)
    check(lib().icempm_read_raw_points(os.fsencode(path), None, 0, C.byref(n), C.byref(vol)))
    hs = HostSideSOA(max(int(capacity or 0), n.value))
    check(lib().icempm_read_raw_points(os.fsencode(path), dptr(hs.host_buffer), hs.capacity, C.byref(n),
                                       C.byref(vol)))
    hs.size = n.value
    return hs, vol.value


class SnapshotManager:
    def __init__(self, model: Model3D):
        self.model = model
        self.previous_frame_exists = False

    def LoadRawPoints(self, fileName: str) -> None:
        if not os.path.exists(fileName):
            raise capi.IcempmError(capi.ERR_INVALID_ARG, "error reading raw points file - no file 2")
        m = self.model
        p = m.prms
        n = C.c_int(0)
        vol = C.c_double(0.0)
        check(lib().icempm_read_raw_points(os.fsencode(fileName), None, 0, C.byref(n), C.byref(vol)))
        # HostSideSOA::Allocate(nPoints * (1 + ExtraSpaceForIncomingPoints)) (snapshotmanager.cpp:35)
        hs, volume = read_raw_points(fileName, int(n.value * (1 + p.ExtraSpaceForIncomingPoints)))
        p.Volume = volume
        check(lib().icempm_prepare_raw_points(C.byref(p), dptr(hs.host_buffer), hs.capacity, hs.size))
        gpu = m.gpu
        gpu.hssoa = hs
        gpu.initialize()
        gpu.allocate_arrays(p, hs.capacity)
        gpu.update_constants(p)
        gpu.transfer_ponts_to_device()
        gpu.remove_disabled_and_sort()
        m.Reset()
        m.Prepare()
        self.previous_frame_exists = False


    def SaveSnapshot(self, outputDirectory: str, frame: int, compress: bool = False) -> str:
        """SnapshotManager::SaveSnapshot (snapshotmanager.cpp:299-339) from the
        host copy (call transfer_from_device first, as the reference saves
        hssoa after a cycle). Returns the file path."""
        os.makedirs(outputDirectory, exist_ok=True)
        path = os.path.join(outputDirectory, "d%05d.h5" % frame)
        gpu = self.model.gpu
        hs = gpu.hssoa
        bins = np.ascontiguousarray(gpu.indenter_sensor_total, dtype=np.float64)
        if bins.size != self.model.prms.IndenterArraySize():
            bins = np.zeros(self.model.prms.IndenterArraySize())
        check(lib().icempm_save_snapshot(os.fsencode(path), C.byref(self.model.prms), dptr(hs.host_buffer),
                                         hs.capacity, hs.size, dptr(bins), 1 if compress else 0))
        return path

    def ReadSnapshot(self, fileName: str, partitions: int = 0) -> None:
        """SnapshotManager::ReadSnapshot (snapshotmanager.cpp:343-405): params,
        host buffer and indenter accumulators from the file, then allocate,
        upload, sort by cell on the device if SimulationStep != 0, Prepare."""
        if not os.path.exists(fileName):
            raise capi.IcempmError(capi.ERR_INVALID_ARG, "snapshot file does not exist")
        m = self.model
        cap = C.c_uint(0)
        size = C.c_uint(0)
        check(lib().icempm_read_snapshot(os.fsencode(fileName), C.byref(m.prms), None, 0, C.byref(cap),
                                         C.byref(size), None, 0))
        if partitions > 0:
            m.prms.nPartitions = partitions
            m.prms.ComputeHelperVariables()
        hs = HostSideSOA(cap.value)
        bins = np.zeros(m.prms.IndenterArraySize())
        check(lib().icempm_read_snapshot(os.fsencode(fileName), None, dptr(hs.host_buffer), hs.capacity, C.byref(cap),
                                         C.byref(size), dptr(bins), bins.size))
        hs.size = size.value
        gpu = m.gpu
        gpu.hssoa = hs
        gpu.prms = m.prms
        gpu.initialize()
        gpu.allocate_arrays(m.prms, hs.capacity)
        gpu.update_constants(m.prms)
        gpu.transfer_ponts_to_device()
        if m.prms.SimulationStep != 0:
            gpu.remove_disabled_and_sort()
        m.Prepare()
        gpu.indenter_sensor_total = bins
        self.previous_frame_exists = False


def grain_ids(hs: HostSideSOA) -> np.ndarray:
    """The 64-bit grain ids of line 0 (before any utility bits are set)."""
    return hs.lines()[capi.IDX_UTILITY, :hs.size].view(np.uint64).copy()
