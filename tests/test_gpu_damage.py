"""The damage model on the device (icempm_set_damage) against the oracle's.

Every point of a small block starts crushed and the indenter comes down at
5 m/s with a cohesion of 0.3e6 Pa, so the return mapping softens and damages
points from the first substeps. After 10 substeps (re-sorting every 3, fused
and unfused) the floating-point lines match the oracle to the bar of the
multi-substep parity tests (the larger of 1e-12 and 10x the oracle's own
rounding spread), the low 32 bits of the utility word exactly and the
damage (a single) within 1e-6. Over 200 substeps the indenter force
samples (every 10 substeps) match the oracle's to 1e-6 of the peak force
(the oracle's own rounding spread over this run, from a different summation
order or no FMA contraction, is 1.3e-12 of the peak force and 3e-8 in D).
The same holds for the scenes of an ensemble. Invalid settings are
refused, and the adjoint refuses to run with the model on.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as orc
from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.calibration import ForceCurveModel
from icempm_3dm_amd.ensemble import BatchedEnsemble
from parity_util import FIELDS, normwise
from test_gpu_ensemble import _three_scenes
from test_gpu_parity import _download, _gpu_with, _oracle_spread

pytestmark = pytest.mark.gpu

MID = replace(scenes.BLOCK_100K, name="damage_mid", grid=(48, 32, 32), cells=(16, 10, 12), dt=2e-6,
              indenter_radius=0.1, indenter_speed=5.0)
DAMAGE = capi.Damage.make(0.01, 0.1)
LINES = dict(FIELDS, P=[capi.IDX_P], Q=[capi.IDX_Q])


@pytest.fixture
def oracle_damage():
    orc.set_damage(DAMAGE)
    yield
    orc.set_damage(None)


def _start(scene):
    p = scenes.make_params(scene)
    p.DP_threshold_p = 0.3e6 / p.DP_tan_phi
    p.ComputeHelperVariables()
    hs = scenes.make_points(scene)
    hs.utility()[:hs.size] |= capi.STATUS_CRUSHED
    return p, hs


def _compare_points(got_lines, got_util, ref, spread, n, tag):
    for name, ls in LINES.items():
        e = normwise(got_lines[ls, :n], ref.lines()[ls, :n])
        assert e <= max(1e-12, 10 * spread[name]), (tag, name, e, spread[name])
    assert np.array_equal(got_util & 0xFFFFFFFF, ref.utility()[:n] & 0xFFFFFFFF), tag
    dd = np.abs(capi.damage_of(got_util) - capi.damage_of(ref.utility()[:n]))
    assert dd.max() <= 1e-6, (tag, dd.max())


@pytest.mark.parametrize("fuse,chunk", [(True, 256), (False, 256), (True, 128)])
def test_damage_step_matches_oracle(require_gpu, oracle_damage, fuse, chunk):
    p, hs = _start(MID)
    shape, center, vel = MID.indenter()
    ind = capi.Indenter.make(shape, center, vel)
    start, ref = hs.copy(), hs.copy()
    g = _gpu_with(p, hs, indenter=(shape, center, vel))
    g.set_damage(DAMAGE)
    g.set_sort_period(3)
    g.set_fusion(fuse)
    g.set_points_per_chunk(chunk)
    nsub = 10
    g.reset_indenter_force_accumulator()
    t = g.step(nsub, 0.0, 0)
    bins_ref = np.zeros(p.IndenterArraySize())
    code, t_ref = orc.step(p, ind, ref.host_buffer, ref.capacity, ref.size, orc.grid_zeros(p), bins_ref, nsub, 0.0, 0,
                           1)
    assert code == 0 and t == t_ref
    n = ref.size
    D = capi.damage_of(ref.utility()[:n])
    assert D.max() > 0.05 and (D > 0).sum() > 500, (D.max(), (D > 0).sum())
    spread = _oracle_spread(p, ind, start, nsub, ref)
    back = _download(g)
    _compare_points(back.lines(), back.utility()[:n], ref, spread, n, ("fuse", fuse))
    fd = g.read_indenter_bins().reshape(-1, 3).sum(0)
    fr = bins_ref.reshape(-1, 3).sum(0)
    assert normwise(fd, fr) <= 1e-9, (fd, fr)
    print("damage max", D.max(), "damaged points", int((D > 0).sum()), "of", n)
    g.close()


def test_damage_force_curve_matches_oracle(require_gpu, oracle_damage):
    p, hs = _start(MID)
    shape, center, vel = MID.indenter()
    ind = capi.Indenter.make(shape, center, vel)
    ref = hs.copy()
    g = _gpu_with(p, hs, indenter=(shape, center, vel))
    g.set_damage(DAMAGE)
    g.set_sort_period(20)
    grid_ref = orc.grid_zeros(p)
    fd, fr = [], []
    t = t_ref = 0.0
    for k in range(20):
        g.reset_indenter_force_accumulator()
        t = g.step(10, t, 10 * k)
        fd.append(g.read_indenter_bins().reshape(-1, 3).sum(0) / 10)
        bins = np.zeros(p.IndenterArraySize())
        code, t_ref = orc.step(p, ind, ref.host_buffer, ref.capacity, ref.size, grid_ref, bins, 10, t_ref, 10 * k,
                               1)
        assert code == 0
        fr.append(bins.reshape(-1, 3).sum(0) / 10)
    fd, fr = np.array(fd), np.array(fr)
    peak = np.abs(fr).max()
    assert peak > 0
    dev = np.abs(fd - fr).max() / peak
    assert dev <= 1e-6, dev
    back = _download(g)
    n = ref.size
    D, Dr = capi.damage_of(back.utility()[:n]), capi.damage_of(ref.utility()[:n])
    assert Dr.max() > 0.5 and np.abs(D - Dr).max() <= 1e-5, (Dr.max(), np.abs(D - Dr).max())
    print("force curve deviation / peak", dev, "damage max", Dr.max())
    g.close()


def test_damage_in_ensemble_scenes(require_gpu, oracle_damage):
    sc, params, indenters, clouds = _three_scenes()
    for c in clouds:
        c.utility()[:c.size] |= capi.STATUS_CRUSHED
    starts = [c.copy() for c in clouds]
    ens = BatchedEnsemble(sc, clouds=clouds, params=params, indenters=indenters)
    ens.set_damage(DAMAGE)
    nsub = 8
    t_dev = ens.step(nsub)
    ens.transfer_from_device()
    any_damage = False
    for i in range(3):
        p, start = params[i], starts[i]
        ind = capi.Indenter.make(*indenters[i])
        ref = start.copy()
        code, t_ref = orc.step(p, ind, ref.host_buffer, ref.capacity, ref.size, orc.grid_zeros(p),
                               np.zeros(p.IndenterArraySize()), nsub, 0.0, 0, 1)
        assert code == 0 and t_dev == t_ref
        spread = _oracle_spread(p, ind, start, nsub, ref)
        got = ens.scene_points(i)
        n = ref.size
        _compare_points(got, got[capi.IDX_UTILITY].view(np.int64), ref, spread, n, ("scene", i))
        any_damage |= bool((capi.damage_of(ref.utility()[:n]) > 0).any())
    assert any_damage
    ens.close()


def test_damage_settings_checked_and_adjoint_refuses(require_gpu):
    p, hs = _start(MID)
    g = _gpu_with(p, hs, indenter=MID.indenter())
    lib = capi.lib()
    for bad in (capi.Damage.make(0.0, 0.1), capi.Damage.make(-1.0, 0.1), capi.Damage.make(0.01, 1.5),
                capi.Damage.make(0.01, -0.1)):
        assert lib.icempm_set_damage(g.ctx, C.byref(bad)) == capi.ERR_INVALID_ARG
    g.close()
    m = ForceCurveModel(replace(MID, indenter_speed=1.0), 20, 10, 10)
    m.gpu.set_damage(DAMAGE)
    out = np.zeros((2, 3))
    assert lib.icempm_adjoint_forward(m.gpu.ctx, C.byref(m.run), capi.dptr(out)) == capi.ADJ_UNSUPPORTED
    m.gpu.set_damage(None)
    m.forward()
    m.close()
