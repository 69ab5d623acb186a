"""ctypes binding of the C-ABI declared in ``include/icempm_3dm.h``.

The library ``icempm_3dm_amd/lib/libicempm_3dm.so`` is built by
``icempm_3dm_amd.build``. Loading fails loudly if it is missing: there is
no fallback implementation of the step.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_PATH = PKG / "lib" / "libicempm_3dm.so"
HEADERS = [ROOT / "include" / "icempm_params.h", ROOT / "include" / "icempm_3dm.h"]

# ---- constants (icempm_params.h / icempm_3dm.h) -----------------------------
IDX_UTILITY, IDX_P, IDX_Q, IDX_JP_INV = 0, 1, 2, 3
POSX, VELX, FE00, BP00 = 4, 7, 10, 19
N_PTS_ARRAYS = 28
STATUS_CRUSHED = 0x10000
STATUS_DISABLED = 0x20000
STATUS_LIQUID = 0x40000

OK, ERR_INVALID_ARG, ERR_HIP, ERR_NOT_READY, ERR_CAPACITY, ERR_DEVICE_FLAG, ERR_NO_DEVICE = range(7)
DEVERR_NONFINITE = 76
ADJ_UNSUPPORTED = 91
INDENTER_CYLINDER_Z, INDENTER_SPHERE = 0, 1
PHASES = ("sort", "reset", "p2g", "update", "g2p", "g2p2g", "graph")
STATS = ("spilled", "sorts", "segments", "p2g_out_of_tile", "g2p_spilled_total")


class SimParams3D(C.Structure):
    """Field-for-field mirror of the reference's SimParams3D
    (simulation/parameters_sim_3d.h:21-110) / icempm_SimParams3D."""

    _fields_ = [
        ("SetupType", C.c_int),
        ("nPtsTotal", C.c_int),
        ("GridXTotal", C.c_int), ("GridY", C.c_int), ("GridZ", C.c_int),
        ("IndenterSubdivisions", C.c_int),
        ("DomainDimensionX", C.c_double),
        ("InitialTimeStep", C.c_double), ("SimulationEndTime", C.c_double),
        ("SimulationTime", C.c_double),
        ("UpdateEveryNthStep", C.c_int),
        ("SnapshotPeriod", C.c_int),
        ("SimulationStep", C.c_int),
        ("Gravity", C.c_double), ("Density", C.c_double), ("WaterDensity", C.c_double),
        ("PoissonsRatio", C.c_double), ("YoungsModulus", C.c_double),
        ("DragGlobal", C.c_double),
        ("lambda_", C.c_double), ("mu", C.c_double), ("kappa", C.c_double),
        ("IceCompressiveStrength", C.c_double), ("IceTensileStrength", C.c_double),
        ("IceShearStrength", C.c_double), ("IceTensileStrength2", C.c_double),
        ("NACC_beta", C.c_double),
        ("DP_tan_phi", C.c_double), ("DP_threshold_p", C.c_double),
        ("GrainVariability", C.c_double),
        ("IndDiameter", C.c_double), ("IndRSq", C.c_double), ("IndVelocity", C.c_double),
        ("IndDepth", C.c_double),
        ("cellsize", C.c_double),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double),
        ("BlockCutout", C.c_double * 3),
        ("ParticleVolume", C.c_double), ("ParticleMass", C.c_double),
        ("ParticleViewSize", C.c_double), ("LiquidParticleMass", C.c_double),
        ("indenter_x", C.c_double), ("indenter_x_initial", C.c_double),
        ("indenter_y", C.c_double), ("indenter_y_initial", C.c_double),
        ("Volume", C.c_double),
        ("nPartitions", C.c_int),
        ("GridHaloSize", C.c_int),
        ("PointTransferFrequency", C.c_int),
        ("VectorCapacity_transfer", C.c_int),
        ("ExtraSpaceForIncomingPoints", C.c_double),
        ("PointsTransferBufferFraction", C.c_double),
        ("RebalanceThresholdFreeSpaceRemaining", C.c_double),
        ("RebalanceThresholdDisabledPercentage", C.c_double),
        ("dt_vol_Dpinv", C.c_double), ("dt_Gravity", C.c_double), ("vmax", C.c_double),
        ("vmax_squared", C.c_double),
        ("gbOffset", C.c_int),
        ("cellsize_inv", C.c_double), ("Dp_inv", C.c_double),
        ("tpb_P2G", C.c_int), ("tpb_Upd", C.c_int), ("tpb_G2P", C.c_int),
        ("animation_threshold_pos", C.c_double),
    ]

    # SimParams3D member functions (parameters_sim_3d.cpp:5-175), run in the library
    def Reset(self) -> "SimParams3D":
        lib().icempm_params_reset(C.byref(self))
        return self

    def ComputeLame(self) -> None:
        lib().icempm_params_compute_lame(C.byref(self))

    def ComputeCamClayParams(self) -> None:
        lib().icempm_params_compute_camclay(C.byref(self))

    def ComputeHelperVariables(self) -> None:
        lib().icempm_params_compute_helper_variables(C.byref(self))

    def SetPhi(self, phi_degrees: float) -> None:
        lib().icempm_params_set_phi(C.byref(self), C.c_double(phi_degrees))

    # SimParams3D::ParseFile (parameters_sim_3d.cpp:88-160): keys and their
    # meaning as the reference reads them; returns the raw point file named.
    _PARSE_DOUBLE = ("InitialTimeStep", "YoungsModulus", "DomainDimensionX", "ParticleViewSize",
                     "SimulationEndTime", "PoissonsRatio", "Gravity", "Density", "IndDiameter", "IndVelocity",
                     "IndDepth", "IceCompressiveStrength", "IceTensileStrength", "IceTensileStrength2",
                     "IceShearStrength", "DP_threshold_p", "GrainVariability", "ExtraSpaceForIncomingPoints",
                     "RebalanceThresholdFreeSpaceRemaining", "RebalanceThresholdDisabledPercentage",
                     "PointsTransferBufferFraction", "animation_threshold_pos")
    _PARSE_INT = ("SetupType", "tpb_P2G", "tpb_Upd", "tpb_G2P", "GridHaloSize", "nPartitions", "GridY", "GridZ")

    def ParseFile(self, file_name: str) -> str:
        import json
        import os
        self.Reset()
        if not os.path.exists(file_name):
            raise IcempmError(ERR_INVALID_ARG, "configuration file is not found")
        with open(file_name) as f:
            doc = json.load(f)
        if not isinstance(doc, dict):
            raise IcempmError(ERR_INVALID_ARG, "configuration file is not JSON")
        if "InputRawPoints" not in doc:
            raise IcempmError(ERR_INVALID_ARG, "config parameter missing")
        for k in self._PARSE_INT:
            if k in doc:
                setattr(self, k, int(doc[k]))
        if "GridX" in doc:
            self.GridXTotal = int(doc["GridX"])
        for k in self._PARSE_DOUBLE:
            if k in doc:
                setattr(self, k, float(doc[k]))
        if "BlockCutout" in doc:
            for i in range(3):
                self.BlockCutout[i] = float(doc["BlockCutout"][i])
        if "DP_phi" in doc:
            self.SetPhi(float(doc["DP_phi"]))
        self.ComputeCamClayParams()
        self.ComputeHelperVariables()
        return str(doc["InputRawPoints"])

    def IndenterArraySize(self) -> int:
        """Number of doubles of the indenter accumulators (parameters_sim_3d.h:105)."""
        return 3 * self.IndenterSubdivisions * self.GridZ

    def copy(self) -> "SimParams3D":
        other = SimParams3D()
        C.memmove(C.byref(other), C.byref(self), C.sizeof(SimParams3D))
        return other


class Indenter(C.Structure):
    """icempm_Indenter: shape, center at t=0, rigid velocity."""

    _fields_ = [("shape", C.c_int), ("center_initial", C.c_double * 3), ("velocity", C.c_double * 3)]

    @staticmethod
    def make(shape: int, center, velocity) -> "Indenter":
        ind = Indenter()
        ind.shape = shape
        for d in range(3):
            ind.center_initial[d] = float(center[d])
            ind.velocity[d] = float(velocity[d])
        return ind


class Damage(C.Structure):
    """icempm_Damage: softening damage of the return mapping (D in bits 32-63
    of the utility word as an IEEE single)."""

    _fields_ = [("enabled", C.c_int), ("softening_strain", C.c_double), ("residual_cohesion", C.c_double)]

    @staticmethod
    def make(softening_strain: float, residual_cohesion: float = 0.1, enabled: bool = True) -> "Damage":
        return Damage(1 if enabled else 0, float(softening_strain), float(residual_cohesion))


def damage_of(util_bits: np.ndarray) -> np.ndarray:
    """D of int64 utility words: the single in bits 32-63, read into [0, 1] (NaN as 0)."""
    d = (np.asarray(util_bits).astype(np.int64).view(np.uint64) >> np.uint64(32)).astype(np.uint32).view(np.float32)
    d = np.where(np.isnan(d), 0.0, d).astype(np.float64)
    return np.clip(d, 0.0, 1.0)


def with_damage(util_bits: np.ndarray, D) -> np.ndarray:
    """util_bits with D (rounded to single) in bits 32-63."""
    u = np.asarray(util_bits).astype(np.int64).view(np.uint64)
    hi = np.asarray(np.broadcast_to(np.float32(D), u.shape)).astype(np.float32).view(np.uint32).astype(np.uint64)
    return ((u & np.uint64(0xFFFFFFFF)) | (hi << np.uint64(32))).view(np.int64)


class AdjointRun(C.Structure):
    """icempm_AdjointRun: substeps, force sample length, checkpoint period, start time."""

    _fields_ = [("n_steps", C.c_int), ("sample_every", C.c_int), ("checkpoint_every", C.c_int), ("t0", C.c_double)]


class Tuning(C.Structure):
    """icempm_Tuning: the launch settings icempm_autotune keeps."""

    _fields_ = [("fuse", C.c_int), ("sort_period", C.c_int), ("blocks_per_cu", C.c_int),
                ("points_per_chunk", C.c_int), ("ms_per_substep", C.c_double)]

    def as_dict(self) -> dict:
        return {"fuse": bool(self.fuse), "sort_period": self.sort_period, "blocks_per_cu": self.blocks_per_cu,
                "points_per_chunk": self.points_per_chunk, "ms_per_substep": self.ms_per_substep}


class IcempmError(RuntimeError):
    """A C-ABI call returned an error code (the reference throws std::runtime_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"icempm error {code}: {message}")
        self.code = code
        self.message = message


_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_longlong)
_ctxp = C.c_void_p
_pp = C.POINTER(SimParams3D)

_PROTOS = {
    "icempm_last_error": (C.c_char_p, []),
    "icempm_version": (C.c_char_p, []),
    "icempm_get_device_count": (C.c_int, [_ip]),
    "icempm_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "icempm_destroy": (C.c_int, [_ctxp]),
    "icempm_allocate": (C.c_int, [_ctxp, _pp, C.c_int]),
    "icempm_update_constants": (C.c_int, [_ctxp, _pp]),
    "icempm_set_indenter": (C.c_int, [_ctxp, C.POINTER(Indenter)]),
    "icempm_set_damage": (C.c_int, [_ctxp, C.POINTER(Damage)]),
    "icempm_transfer_points_to_device": (C.c_int, [_ctxp, _dp, C.c_uint, C.c_uint, C.c_int]),
    "icempm_transfer_from_device": (C.c_int, [_ctxp, _dp, C.c_uint, C.c_uint, _ip, _dp]),
    "icempm_reset_indenter_force_accumulator": (C.c_int, [_ctxp]),
    "icempm_read_indenter_bins": (C.c_int, [_ctxp, C.POINTER(C.c_double)]),
    "icempm_reset_grid": (C.c_int, [_ctxp]),
    "icempm_p2g": (C.c_int, [_ctxp]),
    "icempm_update_nodes": (C.c_int, [_ctxp, C.c_double, C.c_double, C.c_double]),
    "icempm_g2p": (C.c_int, [_ctxp, C.c_int]),
    "icempm_step": (C.c_int, [_ctxp, C.c_int, _dp, C.c_int]),
    "icempm_synchronize": (C.c_int, [_ctxp]),
    "icempm_error_code": (C.c_int, [_ctxp, _ip]),
    "icempm_point_count": (C.c_int, [_ctxp, _ip]),
    "icempm_remove_disabled_and_sort": (C.c_int, [_ctxp, _ip]),
    "icempm_sort_by_brick": (C.c_int, [_ctxp]),
    "icempm_set_sort_period": (C.c_int, [_ctxp, C.c_int]),
    "icempm_set_fusion": (C.c_int, [_ctxp, C.c_int]),
    "icempm_set_graph": (C.c_int, [_ctxp, C.c_int]),
    "icempm_graph_captures": (C.c_int, [_ctxp, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
    "icempm_read_stats": (C.c_int, [_ctxp, _i64p, C.c_int]),
    "icempm_cell_keys": (C.c_int, [_ctxp, C.c_int, _u32p]),
    "icempm_device_order": (C.c_int, [_ctxp, _i32p]),
    "icempm_radix_sort_pairs": (C.c_int, [_ctxp, _u32p, C.c_int, _u32p, _i32p]),
    "icempm_radix_sort_pairs_noalloc": (C.c_int, [_ctxp, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "icempm_download_grid": (C.c_int, [_ctxp, _dp]),
    "icempm_enable_timing": (C.c_int, [_ctxp, C.c_int]),
    "icempm_reset_timings": (C.c_int, [_ctxp]),
    "icempm_read_timings": (C.c_int, [_ctxp, _dp, _i64p]),
    "icempm_allocate_ensemble": (C.c_int, [_ctxp, _pp, C.c_int, C.c_int]),
    "icempm_ensemble_set_indenter": (C.c_int, [_ctxp, C.c_int, C.POINTER(Indenter)]),
    "icempm_ensemble_transfer_points": (C.c_int, [_ctxp, _dp, C.c_uint, _ip]),
    "icempm_ensemble_layout": (C.c_int, [_ctxp, C.c_int, _ip, C.POINTER(C.c_longlong)]),
    "icempm_autotune": (C.c_int, [_ctxp, C.c_int, C.POINTER(Tuning)]),
    "icempm_set_blocks_per_cu": (C.c_int, [_ctxp, C.c_int]),
    "icempm_set_points_per_chunk": (C.c_int, [_ctxp, C.c_int]),
    "icempm_adjoint_forward": (C.c_int, [_ctxp, C.POINTER(AdjointRun), _dp]),
    "icempm_adjoint_backward": (C.c_int, [_ctxp, _dp, _dp, _dp]),
    "icempm_adjoint_positions": (C.c_int, [_ctxp, C.POINTER(C.c_double)]),
    "icempm_adjoint_backward_ex": (C.c_int, [_ctxp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "icempm_scene_block": (C.c_int, [_dp, C.c_uint, _ip, C.c_int, _dp, C.c_double, C.c_double,
                                     C.c_ulonglong, C.c_int, _ip]),
    "icempm_read_raw_points": (C.c_int, [C.c_char_p, _dp, C.c_uint, _ip, _dp]),
    "icempm_prepare_raw_points": (C.c_int, [_pp, _dp, C.c_uint, C.c_int]),
    "icempm_save_snapshot": (C.c_int, [C.c_char_p, _pp, _dp, C.c_uint, C.c_uint, _dp, C.c_int]),
    "icempm_read_snapshot": (C.c_int, [C.c_char_p, _pp, _dp, C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                       _dp, C.c_int]),
    "icempm_params_reset": (None, [_pp]),
    "icempm_params_compute_lame": (None, [_pp]),
    "icempm_params_compute_camclay": (None, [_pp]),
    "icempm_params_compute_helper_variables": (None, [_pp]),
    "icempm_params_set_phi": (None, [_pp, C.c_double]),
    "icempm_params_sizeof": (C.c_int, []),
}


def lib() -> C.CDLL:
    """Loads the HIP library (once). Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m icempm_3dm_amd.build` "
                "(there is no CPU fallback for the MPM step)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        if handle.icempm_params_sizeof() != C.sizeof(SimParams3D):
            raise RuntimeError("SimParams3D layout mismatch between Python and the library")
        _LIB = handle
    return _LIB


def declared_symbols() -> list[str]:
    """Every icempm_* function declared in include/*.h."""
    names: list[str] = []
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(icempm_[A-Za-z0-9_]+)\s*\(", text):
            if m.group(1) not in names:
                names.append(m.group(1))
    return names


def check(rc: int) -> None:
    if rc != OK:
        raise IcempmError(rc, lib().icempm_last_error().decode(errors="replace"))


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def u32ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_u32p)


def i32ptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i32p)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().icempm_get_device_count(C.byref(n)))
    return n.value
