"""ctypes binding of the CPU oracle (oracle/libmpm_oracle.so).

Test infrastructure: the oracle is what the HIP path is tested against and
is never part of the product. Builds the library with oracle/Makefile on first
use if it is missing.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from icempm_3dm_amd import capi
from icempm_3dm_amd.capi import Indenter, SimParams3D

ROOT = Path(__file__).resolve().parents[1]
ORACLE_LIB = ROOT / "oracle" / "lib" / "libmpm_oracle.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_pp = C.POINTER(SimParams3D)
_indp = C.POINTER(Indenter)

_PROTOS = {
    "oracle_cell_keys_reference": (None, [_pp, _dp, C.c_int, C.c_int, _u32p]),
    "oracle_cell_keys_brick": (None, [_pp, _dp, C.c_int, C.c_int, _u32p]),
    "oracle_stable_sort": (None, [_u32p, C.c_int, _i32p]),
    "oracle_remove_disabled_and_sort": (C.c_int, [_pp, _dp, C.c_int, C.c_int]),
    "oracle_weights": (None, [_pp, _dp, _ip, _dp, _dp]),
    "oracle_kirchhoff": (None, [C.c_double, C.c_double, _dp, _dp]),
    "oracle_svd3": (None, [_dp, _dp, _dp, _dp]),
    "oracle_det3": (C.c_double, [_dp]),
    "oracle_p2g": (C.c_int, [_pp, _dp, C.c_int, C.c_int, _dp, C.c_int]),
    "oracle_update_nodes": (None, [_pp, _indp, _dp, _dp, _dp, C.c_int]),
    "oracle_g2p": (C.c_int, [_pp, _dp, C.c_int, C.c_int, _dp, C.c_int, C.c_int]),
    "oracle_step": (C.c_int, [_pp, _indp, _dp, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, C.c_int]),
    "oracle_adj_forward": (C.c_int, [_pp, _indp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "oracle_adj_gradient": (C.c_int, [_pp, _indp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "oracle_adj_gradient_ex": (C.c_int, [_pp, _indp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
    "oracle_set_damage": (None, [C.POINTER(capi.Damage)]),
    "oracle_set_svd_fp32": (None, [C.c_int]),
    "oracle_svd3_fp32": (None, [_dp, _dp, _dp, _dp]),
}

ORACLE_LIB_NOFMA = ROOT / "oracle" / "lib" / "libmpm_oracle_nofma.so"

_LIB = None
_LIB_NOFMA = None


def _load(path: Path) -> C.CDLL:
    if not path.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    h = C.CDLL(str(path))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    return h


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _load(ORACLE_LIB)
    return _LIB


def lib_nofma() -> C.CDLL:
    """The same oracle built without FMA contraction (rounding-spread probe)."""
    global _LIB_NOFMA
    if _LIB_NOFMA is None:
        _LIB_NOFMA = _load(ORACLE_LIB_NOFMA)
    return _LIB_NOFMA


def set_damage(damage) -> None:
    """The damage model of both oracle builds (None: off)."""
    for h in (lib(), lib_nofma()):
        h.oracle_set_damage(C.byref(damage) if damage is not None else None)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def grid_zeros(p: SimParams3D) -> np.ndarray:
    return np.zeros((4, p.GridXTotal, p.GridY, p.GridZ))


def keys_reference(p: SimParams3D, soa: np.ndarray, pitch: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_cell_keys_reference(C.byref(p), _d(soa), pitch, n, out.ctypes.data_as(_u32p))
    return out


def keys_brick(p: SimParams3D, soa: np.ndarray, pitch: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_cell_keys_brick(C.byref(p), _d(soa), pitch, n, out.ctypes.data_as(_u32p))
    return out


def stable_sort(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    perm = np.zeros(keys.shape[0], dtype=np.int32)
    lib().oracle_stable_sort(keys.ctypes.data_as(_u32p), keys.shape[0], perm.ctypes.data_as(_i32p))
    return perm


def remove_disabled_and_sort(p: SimParams3D, soa: np.ndarray, pitch: int, n: int) -> int:
    return lib().oracle_remove_disabled_and_sort(C.byref(p), _d(soa), pitch, n)


def weights(p: SimParams3D, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    base = (C.c_int * 3)()
    fx = np.zeros(3)
    w = np.zeros(9)
    lib().oracle_weights(C.byref(p), _d(pos), base, _d(fx), _d(w))
    return np.array(list(base)), fx, w.reshape(3, 3)  # w[i, axis]


def kirchhoff(mu: float, kappa: float, F) -> np.ndarray:
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    out = np.zeros(9)
    lib().oracle_kirchhoff(mu, kappa, _d(F), _d(out))
    return out.reshape(3, 3)


def svd3(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    U, S, V = np.zeros(9), np.zeros(3), np.zeros(9)
    lib().oracle_svd3(_d(F), _d(U), _d(S), _d(V))
    return U.reshape(3, 3), S, V.reshape(3, 3)


def det3(F) -> float:
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    return lib().oracle_det3(_d(F))


def p2g(p: SimParams3D, soa: np.ndarray, pitch: int, n: int, grid: np.ndarray, nthreads: int = 1) -> int:
    assert grid.shape == (4, p.GridXTotal, p.GridY, p.GridZ)
    return lib().oracle_p2g(C.byref(p), _d(soa), pitch, n, _d(grid), nthreads)


def update_nodes(p: SimParams3D, ind: Indenter, center, grid: np.ndarray, bins: np.ndarray, nthreads: int = 1):
    c = np.ascontiguousarray(center, dtype=np.float64)
    lib().oracle_update_nodes(C.byref(p), C.byref(ind), _d(c), _d(grid), _d(bins), nthreads)


def g2p(p: SimParams3D, soa: np.ndarray, pitch: int, n: int, grid: np.ndarray, record_pq: bool,
        nthreads: int = 1) -> int:
    return lib().oracle_g2p(C.byref(p), _d(soa), pitch, n, _d(grid), 1 if record_pq else 0, nthreads)


def step(p: SimParams3D, ind: Indenter, soa: np.ndarray, pitch: int, n: int, grid: np.ndarray, bins: np.ndarray,
         n_substeps: int, sim_time: float, sim_step: int, nthreads: int = 1, nofma: bool = False):
    t = C.c_double(sim_time)
    h = lib_nofma() if nofma else lib()
    code = h.oracle_step(C.byref(p), C.byref(ind), _d(soa), pitch, n, _d(grid), _d(bins), n_substeps,
                         C.byref(t), sim_step, nthreads)
    return code, t.value


def reference_indenter(p: SimParams3D) -> Indenter:
    """The reference law: z-cylinder from (indenter_x_initial, indenter_y_initial), +x at IndVelocity."""
    return Indenter.make(capi.INDENTER_CYLINDER_Z, (p.indenter_x_initial, p.indenter_y_initial, 0.0),
                         (p.IndVelocity, 0.0, 0.0))


def adj_forward(p: SimParams3D, ind: Indenter, soa: np.ndarray, pitch: int, n: int, n_steps: int,
                sample_every: int):
    """Force samples of an n_steps run from soa (left untouched): (code, (n_samples, 3))."""
    f = np.zeros((n_steps // sample_every, 3))
    code = lib().oracle_adj_forward(C.byref(p), C.byref(ind), _d(soa), pitch, n, n_steps, sample_every, _d(f))
    return code, f


def adj_gradient(p: SimParams3D, ind: Indenter, soa: np.ndarray, pitch: int, n: int, n_steps: int,
                 sample_every: int, dL_dF: np.ndarray):
    """Reverse mode of adj_forward's run: (code, dL/dmu, dL/dkappa, dL/dDP_threshold_p, dL/dDP_tan_phi)."""
    g = np.zeros(4)
    w = np.ascontiguousarray(dL_dF, dtype=np.float64)
    code = lib().oracle_adj_gradient(C.byref(p), C.byref(ind), _d(soa), pitch, n, n_steps, sample_every, _d(w),
                                     _d(g))
    return code, g[0], g[1], g[2], g[3]


def adj_gradient_ex(p: SimParams3D, ind: Indenter, soa: np.ndarray, pitch: int, n: int, n_steps: int,
                    sample_every: int, dL_dF, dL_dx):
    """adj_gradient with a final-position term: dL_dx (3, n) = dL/dx_N in point
    order; either may be None. Returns (code, 4 parameter adjoints)."""
    g = np.zeros(4)
    w = None if dL_dF is None else np.ascontiguousarray(dL_dF, dtype=np.float64)
    x = None if dL_dx is None else np.ascontiguousarray(dL_dx, dtype=np.float64).reshape(3, n)
    code = lib().oracle_adj_gradient_ex(C.byref(p), C.byref(ind), _d(soa), pitch, n, n_steps, sample_every,
                                        None if w is None else _d(w), None if x is None else _d(x), _d(g))
    return code, g[0], g[1], g[2], g[3]
