"""Comparison helpers for the parity tests.

Tolerances (FP64): per field, ||a - b||_inf / ||b||_inf <= 1e-12. Atomic
summation order on the device and FMA contraction make bit equality
impossible for the floating-point lines; utility bits, keys and orders are
compared exactly.
"""
from __future__ import annotations

import numpy as np

from icempm_3dm_amd import capi

RTOL = 1e-12

FIELDS = {
    "Jp_inv": [capi.IDX_JP_INV],
    "pos": list(range(capi.POSX, capi.POSX + 3)),
    "vel": list(range(capi.VELX, capi.VELX + 3)),
    "Fe": list(range(capi.FE00, capi.FE00 + 9)),
    "Bp": list(range(capi.BP00, capi.BP00 + 9)),
}


def normwise(a: np.ndarray, b: np.ndarray) -> float:
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    diff = float(np.max(np.abs(a - b))) if a.size else 0.0
    if scale == 0.0:
        return diff
    return diff / scale


def compare_points(got: np.ndarray, want: np.ndarray, n: int, rtol: float = RTOL, check_pq: bool = False):
    """got/want: (28, capacity) line views. Returns a dict of errors; asserts."""
    errs = {}
    for name, lines in FIELDS.items():
        errs[name] = normwise(got[lines, :n], want[lines, :n])
    if check_pq:
        errs["P"] = normwise(got[capi.IDX_P, :n], want[capi.IDX_P, :n])
        errs["Q"] = normwise(got[capi.IDX_Q, :n], want[capi.IDX_Q, :n])
    ug = got[capi.IDX_UTILITY, :n].view(np.int64)
    uw = want[capi.IDX_UTILITY, :n].view(np.int64)
    flips = int(np.count_nonzero(ug != uw))
    bad = {k: v for k, v in errs.items() if not (v <= rtol)}
    assert flips == 0, f"utility bits differ on {flips} points; errors {errs}"
    assert not bad, f"fields beyond {rtol}: {bad} (all: {errs})"
    return errs


def compare_grid(got: np.ndarray, want: np.ndarray, rtol: float = RTOL):
    errs = {f"line{k}": normwise(got[k], want[k]) for k in range(4)}
    bad = {k: v for k, v in errs.items() if not (v <= rtol)}
    assert not bad, f"grid lines beyond {rtol}: {bad} (all: {errs})"
    return errs
