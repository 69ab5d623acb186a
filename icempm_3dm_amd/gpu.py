"""Host-side mirror of the reference's device boundary for one partition.

``GPUImplementation6`` keeps the member names and argument meaning of
``GPU_Implementation6`` (simulation/gpu_implementation6.h:24-59) and of the
``GPU_Partition_3D`` calls it sequences, and drives the HIP library through
the C-ABI. Errors raise ``IcempmError`` where the reference throws
``std::runtime_error``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import SimParams3D, check, dptr, lib
from .hssoa import HostSideSOA


class GPUImplementation6:
    def __init__(self, device: int = 0):
        self.device = device
        self.ctx = C.c_void_p()
        self.hssoa = HostSideSOA(0)
        self.prms: SimParams3D | None = None
        self.indenter_force = np.zeros(3)
        self.indenter_sensor_total = np.zeros(0)
        self.error_code = 0

    # -- lifetime -------------------------------------------------------------
    def initialize(self) -> None:
        """GPU_Implementation6::initialize (gpu_implementation6.cpp:319-337), one partition."""
        if self.ctx:
            return
        check(lib().icempm_create(self.device, C.byref(self.ctx)))

    def close(self) -> None:
        if self.ctx:
            lib().icempm_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        self.initialize()
        return self

    def __exit__(self, *exc):
        self.close()

    # -- setup ----------------------------------------------------------------
    def allocate_arrays(self, prms: SimParams3D, capacity: int | None = None) -> None:
        """GPU_Implementation6::allocate_arrays (gpu_implementation6.cpp:385-401)."""
        self.initialize()
        self.prms = prms
        if capacity is None:
            capacity = max(self.hssoa.capacity, self.hssoa.size)
        check(lib().icempm_allocate(self.ctx, C.byref(prms), int(capacity)))
        self.indenter_sensor_total = np.zeros(prms.IndenterArraySize())

    def update_constants(self, prms: SimParams3D | None = None) -> None:
        """GPU_Partition_3D::update_constants (gpu_partition_3d.cu:135-143)."""
        if prms is not None:
            self.prms = prms
        check(lib().icempm_update_constants(self.ctx, C.byref(self.prms)))

    def set_indenter(self, shape: int, center, velocity) -> None:
        ind = capi.Indenter.make(shape, center, velocity)
        check(lib().icempm_set_indenter(self.ctx, C.byref(ind)))

    def set_sort_period(self, period: int) -> None:
        check(lib().icempm_set_sort_period(self.ctx, int(period)))

    def set_fusion(self, on: bool) -> None:
        check(lib().icempm_set_fusion(self.ctx, 1 if on else 0))

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

    def set_blocks_per_cu(self, blocks_per_cu: int) -> None:
        check(lib().icempm_set_blocks_per_cu(self.ctx, int(blocks_per_cu)))

    def set_points_per_chunk(self, points_per_chunk: int) -> None:
        check(lib().icempm_set_points_per_chunk(self.ctx, int(points_per_chunk)))

    def autotune(self, trial_substeps: int = 10) -> dict:
        """icempm_autotune: keeps the fastest fusion / sort period / workgroup
        count / points per chunk for the current points (which it leaves as
        they are)."""
        t = capi.Tuning()
        check(lib().icempm_autotune(self.ctx, int(trial_substeps), C.byref(t)))
        return t.as_dict()

    def transfer_ponts_to_device(self) -> None:
        """GPU_Implementation6::transfer_ponts_to_device (gpu_implementation6.cpp:404-414)."""
        check(lib().icempm_transfer_points_to_device(self.ctx, dptr(self.hssoa.host_buffer), self.hssoa.capacity, 0,
                                                     self.hssoa.size))

    transfer_points_to_device = transfer_ponts_to_device

    # -- the step, phase by phase ---------------------------------------------
    def reset_grid(self) -> None:
        check(lib().icempm_reset_grid(self.ctx))

    def reset_indenter_force_accumulator(self) -> None:
        check(lib().icempm_reset_indenter_force_accumulator(self.ctx))

    def p2g(self) -> None:
        check(lib().icempm_p2g(self.ctx))

    def receive_halos(self) -> None:
        """One partition: nothing to receive (gpu_implementation6.cpp:80-112)."""

    def update_nodes(self, center) -> None:
        check(lib().icempm_update_nodes(self.ctx, float(center[0]), float(center[1]), float(center[2])))

    def g2p(self, recordPQ: bool, enablePointTransfer: bool = False) -> None:
        check(lib().icempm_g2p(self.ctx, 1 if recordPQ else 0))

    def synchronize(self) -> None:
        check(lib().icempm_synchronize(self.ctx))

    def step(self, n_substeps: int, simulation_time: float, simulation_step: int) -> float:
        """Model3D::Step's substep loop (model_3d.cpp:35-51) on the device; returns the new time."""
        t = C.c_double(simulation_time)
        check(lib().icempm_step(self.ctx, int(n_substeps), C.byref(t), int(simulation_step)))
        return t.value

    # -- results --------------------------------------------------------------
    def transfer_from_device(self) -> None:
        """GPU_Implementation6::transfer_from_device (gpu_implementation6.cpp:417-496):
        points back into hssoa (upload order), indenter bins summed, total force
        evaluated, bins scaled by 1/(N h^2), active-count check."""
        n = C.c_int(0)
        raw = np.zeros(self.prms.IndenterArraySize())
        rc = lib().icempm_transfer_from_device(self.ctx, dptr(self.hssoa.host_buffer), self.hssoa.capacity, 0,
                                               C.byref(n), dptr(raw))
        self.hssoa.size = n.value
        if rc == capi.ERR_DEVICE_FLAG:
            code = C.c_int(0)
            lib().icempm_error_code(self.ctx, C.byref(code))
            self.error_code = code.value
        check(rc)
        self.indenter_sensor_total = raw
        self.evaluate_indenter_total_force()
        N = self.prms.UpdateEveryNthStep
        self.indenter_sensor_total = raw / (N * self.prms.cellsize * self.prms.cellsize)
        count = self.hssoa.count_active()
        if count != self.prms.nPtsTotal:
            raise capi.IcempmError(capi.ERR_CAPACITY,
                                   f"transfer_from_device(): active point count mismatch {count} != "
                                   f"{self.prms.nPtsTotal}")

    def evaluate_indenter_total_force(self) -> np.ndarray:
        """gpu_implementation6.cpp:499-510: sum of the bins / UpdateEveryNthStep."""
        b = self.indenter_sensor_total.reshape(-1, 3)
        self.indenter_force = b.sum(axis=0) / float(self.prms.UpdateEveryNthStep)
        return self.indenter_force

    def read_indenter_bins(self) -> np.ndarray:
        """Raw accumulators without moving points (synchronizes)."""
        raw = np.zeros(self.prms.IndenterArraySize())
        check(lib().icempm_read_indenter_bins(self.ctx, dptr(raw)))
        return raw

    def download_grid(self) -> np.ndarray:
        """Grid as (4, GX, GY, GZ): m and momentum (or velocity after update_nodes)."""
        p = self.prms
        out = np.zeros(4 * p.GridXTotal * p.GridY * p.GridZ)
        check(lib().icempm_download_grid(self.ctx, dptr(out)))
        return out.reshape(4, p.GridXTotal, p.GridY, p.GridZ)

    def device_error_code(self) -> int:
        code = C.c_int(0)
        check(lib().icempm_error_code(self.ctx, C.byref(code)))
        return code.value

    def point_count(self) -> int:
        n = C.c_int(0)
        check(lib().icempm_point_count(self.ctx, C.byref(n)))
        return n.value

    # -- sorting ----------------------------------------------------------------
    def remove_disabled_and_sort(self) -> int:
        """HostSideSOA::RemoveDisabledAndSort (host_side_soa.cpp:56-67) on the device."""
        n = C.c_int(0)
        check(lib().icempm_remove_disabled_and_sort(self.ctx, C.byref(n)))
        self.prms.nPtsTotal = n.value
        return n.value

    def sort_by_brick(self) -> None:
        check(lib().icempm_sort_by_brick(self.ctx))

    def cell_keys(self, kind: int) -> np.ndarray:
        n = self.point_count()
        out = np.zeros(n, dtype=np.uint32)
        check(lib().icempm_cell_keys(self.ctx, int(kind), capi.u32ptr(out)))
        return out

    def device_order(self) -> np.ndarray:
        n = self.point_count()
        out = np.zeros(n, dtype=np.int32)
        check(lib().icempm_device_order(self.ctx, capi.i32ptr(out)))
        return out

    def radix_sort_pairs(self, keys: np.ndarray):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        ks = np.zeros_like(keys)
        perm = np.zeros(keys.shape[0], dtype=np.int32)
        check(lib().icempm_radix_sort_pairs(self.ctx, capi.u32ptr(keys), int(keys.shape[0]), capi.u32ptr(ks),
                                            capi.i32ptr(perm)))
        return ks, perm

    def read_stats(self) -> dict:
        """Step counters (synchronizes): spilled points of the last G2P, sorts done, brick segments."""
        out = np.zeros(len(capi.STATS), dtype=np.int64)
        check(lib().icempm_read_stats(self.ctx, out.ctypes.data_as(C.POINTER(C.c_longlong)), len(capi.STATS)))
        return {name: int(out[i]) for i, name in enumerate(capi.STATS)}

    # -- timing -----------------------------------------------------------------
    def enable_timing(self, on: bool = True) -> None:
        check(lib().icempm_enable_timing(self.ctx, 1 if on else 0))

    def reset_timings(self) -> None:
        check(lib().icempm_reset_timings(self.ctx))

    def read_timings(self) -> dict:
        ms = np.zeros(len(capi.PHASES))
        cnt = np.zeros(len(capi.PHASES), dtype=np.int64)
        check(lib().icempm_read_timings(self.ctx, dptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_longlong))))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(capi.PHASES)}
