"""icempm_radix_sort_pairs_noalloc against icempm_radix_sort_pairs.

The no-allocation variant sorts device arrays with the library's fixed
device array as its only scratch; two calls in a row on caller-owned device
buffers (the second smaller, so stale contents of the fixed array from the
first call are in place) must give exactly the keys and the stable order that
the allocating entry point gives for the same keys, which is numpy's stable
argsort. Sizes beyond ICEMPM_NOALLOC_MAX_PAIRS are refused before any work.
Device buffers come from hipMalloc of the HIP runtime the library itself
loaded (no torch).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from icempm_3dm_amd import capi
from icempm_3dm_amd.capi import lib

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 22  # ICEMPM_NOALLOC_MAX_PAIRS
H2D, D2H = 1, 2      # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    lib()  # the library first, so that its HIP runtime is the one found by soname
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _keys(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 1 << 25, n, dtype=np.uint32)  # cell keys of a 512x256x256 grid
    k[rng.random(n) < 0.05] = 0xFFFFFFFF              # disabled points
    k[: n // 7] = k[n // 3]                             # a long run of one key
    return k


def _reference(ctx, keys: np.ndarray):
    ks = np.zeros_like(keys)
    perm = np.zeros(keys.shape[0], dtype=np.int32)
    capi.check(lib().icempm_radix_sort_pairs(ctx, capi.u32ptr(keys), int(keys.shape[0]), capi.u32ptr(ks),
                                             capi.i32ptr(perm)))
    return ks, perm


def test_noalloc_sort_matches_allocating_sort(require_gpu):
    hip = _hip()
    ctx = C.c_void_p()
    capi.check(lib().icempm_create(0, C.byref(ctx)))
    bufs = []
    try:
        n_max = 3_000_000
        d = {}
        for name in ("keys", "sorted", "perm"):
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), 4 * n_max) == 0
            bufs.append(p)
            d[name] = p
        for call, (n, seed, in_place) in enumerate(((n_max, 11, False), (123_457, 12, False), (70_001, 13, True))):
            keys = _keys(n, seed)
            assert hip.hipMemcpy(d["keys"], keys.ctypes.data, 4 * n, H2D) == 0
            out_keys = d["keys"] if in_place else d["sorted"]
            rc = lib().icempm_radix_sort_pairs_noalloc(ctx, d["keys"], n, out_keys, d["perm"])
            assert rc == 0, lib().icempm_last_error()
            capi.check(lib().icempm_synchronize(ctx))
            got_k = np.zeros(n, dtype=np.uint32)
            got_p = np.zeros(n, dtype=np.int32)
            assert hip.hipMemcpy(got_k.ctypes.data, out_keys, 4 * n, D2H) == 0
            assert hip.hipMemcpy(got_p.ctypes.data, d["perm"], 4 * n, D2H) == 0
            want_k, want_p = _reference(ctx, keys)
            assert np.array_equal(got_k, want_k), call
            assert np.array_equal(got_p, want_p), call
            assert np.array_equal(want_p, np.argsort(keys, kind="stable")), call
            print(f"call {call}: n={n} in_place={in_place}: equal to the allocating sort")
        # beyond the fixed array: refused, nothing launched
        rc = lib().icempm_radix_sort_pairs_noalloc(ctx, d["keys"], MAX_PAIRS + 1, d["sorted"], d["perm"])
        assert rc == capi.ERR_INVALID_ARG and b"ICEMPM_NOALLOC_MAX_PAIRS" in lib().icempm_last_error()
        assert lib().icempm_radix_sort_pairs_noalloc(ctx, d["keys"], 0, d["sorted"], d["perm"]) == 0
    finally:
        for p in bufs:
            hip.hipFree(p)
        lib().icempm_destroy(ctx)
