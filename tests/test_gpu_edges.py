"""Edge cases of the step on the device: an empty cloud, a single point, a
cloud whose every point is disabled, and points spread one per brick.

  empty          no points uploaded: fused and unfused steps advance the time
                 exactly as the oracle does, raise no error, add no force
                 and download nothing;
  single point   one point of the 100k block, stepped against the oracle with
                 the multi-substep bar of the other parity tests;
  all disabled   every point disabled: the step leaves the points as uploaded
                 (bit for bit), and RemoveDisabledAndSort keeps none;
  one per brick  a ragged cloud with one point in each of a set of bricks
                 (every brick segment a single point, every chunk nearly
                 empty), fused and unfused, against the oracle;
  capacity       a capacity above the point-addressing limit is refused.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as orc
from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.hssoa import HostSideSOA
from parity_util import FIELDS, normwise
from test_gpu_parity import _download, _gpu_with, _oracle_spread

pytestmark = pytest.mark.gpu

SC = replace(scenes.BLOCK_100K, dt=2e-6)
LINES = dict(FIELDS, P=[capi.IDX_P], Q=[capi.IDX_Q])


def _subset(hs: HostSideSOA, idx: np.ndarray, capacity: int | None = None) -> HostSideSOA:
    out = HostSideSOA(capacity or max(len(idx), 1))
    out.lines()[:, :len(idx)] = hs.lines()[:, idx]
    out.size = len(idx)
    return out


def _against_oracle(p, hs: HostSideSOA, fuse: bool, nsub: int, tag: str):
    shape, center, vel = SC.indenter()
    ind = capi.Indenter.make(shape, center, vel)
    start, ref = hs.copy(), hs.copy()
    g = _gpu_with(p, hs, indenter=(shape, center, vel))
    g.set_fusion(fuse)
    g.set_sort_period(2)
    g.reset_indenter_force_accumulator()
    t_dev = g.step(nsub, 0.0, 0)
    bins_ref = np.zeros(p.IndenterArraySize())
    code, t_ref = orc.step(p, ind, ref.host_buffer, ref.capacity, ref.size, orc.grid_zeros(p), bins_ref, nsub, 0.0,
                           0, 1)
    assert code == 0 and t_dev == t_ref
    back = _download(g)
    assert g.device_error_code() == 0
    n = ref.size
    spread = _oracle_spread(p, ind, start, nsub, ref)
    report = {}
    for name, ls in LINES.items():
        e = normwise(back.lines()[ls, :n], ref.lines()[ls, :n])
        report[name] = (e, spread[name])
        assert e <= max(1e-12, 10 * spread[name]), (tag, name, report)
    assert np.array_equal(back.utility()[:n], ref.utility()[:n]), tag
    print(tag, n, "points", report)
    g.close()


@pytest.mark.parametrize("fuse", [True, False])
def test_empty_cloud(require_gpu, fuse):
    p = scenes.make_params(SC)
    hs = HostSideSOA(64)
    g = _gpu_with(p, hs)
    g.set_fusion(fuse)
    g.set_sort_period(2)
    g.reset_indenter_force_accumulator()
    t = g.step(5, 0.0, 0)
    t_ref = 0.0
    for _ in range(5):
        t_ref += p.InitialTimeStep
    assert t == t_ref
    assert g.device_error_code() == 0
    assert g.point_count() == 0
    assert np.all(g.read_indenter_bins() == 0)
    back = _download(g)
    assert back.size == 0
    assert g.remove_disabled_and_sort() == 0
    g.close()


def test_single_point(require_gpu):
    p = scenes.make_params(SC)
    hs = scenes.perturbed_points(SC, seed=5, amp_F=5e-4, amp_v=0.02, amp_B=1.0)
    one = _subset(hs, np.array([hs.size // 2]), capacity=8)
    _against_oracle(p, one, True, 6, "single point")


def test_all_points_disabled(require_gpu):
    p = scenes.make_params(SC)
    hs = scenes.perturbed_points(SC, seed=6, amp_F=5e-4, amp_v=0.02, amp_B=1.0)
    n = hs.size
    hs.utility()[:n] |= capi.STATUS_DISABLED
    start = hs.copy()
    g = _gpu_with(p, hs)
    g.reset_indenter_force_accumulator()
    g.step(4, 0.0, 0)
    back = _download(g)
    assert g.device_error_code() == 0
    assert np.array_equal(back.lines()[:, :n].view(np.int64), start.lines()[:, :n].view(np.int64))
    assert np.all(g.read_indenter_bins() == 0)
    assert g.remove_disabled_and_sort() == 0
    g.close()


@pytest.mark.parametrize("fuse", [True, False])
def test_one_point_per_brick(require_gpu, fuse):
    p = scenes.make_params(SC)
    hs = scenes.perturbed_points(SC, seed=7, amp_F=5e-4, amp_v=0.02, amp_B=1.0)
    n = hs.size
    pos = hs.lines()[capi.POSX:capi.POSX + 3, :n]
    cells = np.floor(pos * p.cellsize_inv + 0.5).astype(np.int64)
    bricks = (cells[0] >> 2) * 1_000_000 + (cells[1] >> 2) * 1_000 + (cells[2] >> 2)
    _, first = np.unique(bricks, return_index=True)
    idx = np.sort(first)
    assert 100 < len(idx) < n // 100
    ragged = _subset(hs, idx)
    _against_oracle(p, ragged, fuse, 6, f"one point per brick fuse={fuse}")


def test_capacity_above_the_point_addressing_limit_is_refused(require_gpu):
    """The kernels address a point line with a 32-bit byte offset: a capacity
    above 2^29 - 64 points is refused before any point storage is allocated,
    and the context stays usable."""
    from icempm_3dm_amd.gpu import GPUImplementation6

    p = scenes.make_params(SC)
    g = GPUImplementation6(0)
    g.hssoa = HostSideSOA(64)
    with pytest.raises(capi.IcempmError) as e:
        g.allocate_arrays(p, (1 << 29) - 63)
    assert e.value.code == capi.ERR_CAPACITY
    g.allocate_arrays(p, 64)
    g.close()
