"""Batched ensemble through the C-ABI: every scene of an ensemble must end
where the oracle, run on that scene alone with its own parameters and
indenter, ends.

Three scenes of different grid sizes, block sizes (uneven point counts),
materials and indenter laws (two spheres at different speeds, the
reference's z-cylinder) advance 8 substeps together; each scene's points are
compared with its own oracle run to the multi-substep bar of the single-scene
parity tests (the larger of 1e-12 and 10x the oracle's own rounding spread),
utility bits exactly, and each scene's indenter accumulators within 1e-9.
The stored grid puts the scenes at node offsets of a few hundred, so the
device computes weights from positions with a few more integer bits than the
oracle does; the comparison bar absorbs that (a few ulps of 200 in cell units).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as orc
from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.ensemble import BatchedEnsemble
from parity_util import FIELDS, normwise
from test_gpu_parity import _oracle_spread

pytestmark = pytest.mark.gpu

DT = 2e-6


def _three_scenes():
    a = replace(scenes.BLOCK_100K, name="ens_a", dt=DT, seed=11)
    b = replace(scenes.BLOCK_100K, name="ens_b", grid=(48, 40, 36), cells=(20, 15, 18), dt=DT, seed=12,
                youngs=6e9, phi_deg=45.0, cohesion=1e6, indenter_speed=1.7, indenter_radius=0.08)
    c = replace(scenes.BLOCK_100K, name="ens_c", grid=(72, 28, 40), cells=(30, 12, 16), dt=DT, seed=13,
                youngs=1e10, phi_deg=68.0, cohesion=3.5e6)
    params = [scenes.make_params(a), scenes.make_params(b), scenes.reference_cylinder_params(c)]
    pc = params[2]
    indenters = [a.indenter(), b.indenter(),
                 (capi.INDENTER_CYLINDER_Z, (pc.indenter_x_initial, pc.indenter_y_initial, 0.0),
                  (pc.IndVelocity, 0.0, 0.0))]
    clouds = [scenes.perturbed_points(s, seed=30 + i, amp_F=5e-4, amp_v=0.02, amp_B=1.0)
              for i, s in enumerate((a, b, c))]
    return [a, b, c], params, indenters, clouds


@pytest.mark.parametrize("fuse,chunk", [(True, 256), (False, 256), (True, 128)])
def test_ensemble_scenes_match_their_own_oracle_runs(require_gpu, fuse, chunk):
    sc, params, indenters, clouds = _three_scenes()
    starts = [c.copy() for c in clouds]
    ens = BatchedEnsemble(sc, clouds=clouds, params=params, indenters=indenters)
    assert len({o for o in ens.origin}) == 3
    ens.set_fusion(fuse)
    ens.set_points_per_chunk(chunk)
    ens.reset_indenter_force_accumulator()
    nsub = 8
    t_dev = ens.step(nsub)
    raw = ens.transfer_from_device()
    for i in range(3):
        p, start = params[i], starts[i]
        ind = capi.Indenter.make(*indenters[i])
        ref = start.copy()
        bins_ref = np.zeros(p.IndenterArraySize())
        code, t_ref = orc.step(p, ind, ref.host_buffer, ref.capacity, ref.size, orc.grid_zeros(p), bins_ref, nsub,
                               0.0, 0, 1)
        assert code == 0
        assert t_dev == t_ref
        spread = _oracle_spread(p, ind, start, nsub, ref)
        got = ens.scene_points(i)
        n = ref.size
        assert got.shape[1] == n
        report = {}
        for name, ls in dict(FIELDS, P=[capi.IDX_P], Q=[capi.IDX_Q]).items():
            e = normwise(got[ls, :], ref.lines()[ls, :n])
            report[name] = (e, spread[name])
            assert e <= max(1e-12, 10 * spread[name]), (i, name, report)
        assert np.array_equal(got[capi.IDX_UTILITY].view(np.int64), ref.utility()[:n]), i
        fd = ens.scene_bins(raw, i).reshape(-1, 3).sum(0)
        fr = bins_ref.reshape(-1, 3).sum(0)
        assert np.abs(fr).sum() > 0, i
        assert normwise(fd, fr) <= 1e-9, (i, fd, fr)
        print("scene", i, "n", n, report)
    ens.close()


def test_ensemble_scene_boundary_flags_and_upload_check(require_gpu):
    sc, params, indenters, clouds = _three_scenes()
    # a point of scene 1 inside its grid box but with a stencil past the
    # scene's last node: the device must flag it, as a single scene would
    L = clouds[1].lines()
    L[capi.POSX + 2, 5] = (params[1].GridZ - 1.2) * params[1].cellsize
    ens = BatchedEnsemble(sc, clouds=clouds, params=params, indenters=indenters)
    ens.step(1)
    with pytest.raises(capi.IcempmError):
        ens.transfer_from_device()
    code = C.c_int(0)
    capi.check(capi.lib().icempm_error_code(ens.ctx, C.byref(code)))
    assert code.value == 75
    # a point outside its scene's grid box is refused at upload
    ens.hssoa.lines()[capi.POSX, ens.first[2] + 3] = -0.01
    with pytest.raises(capi.IcempmError, match="outside the grid of scene 2"):
        ens.transfer_points_to_device()
    # single-scene entry points are refused on an ensemble context
    assert capi.lib().icempm_update_constants(ens.ctx, C.byref(params[0])) == capi.ERR_INVALID_ARG
    assert capi.lib().icempm_remove_disabled_and_sort(ens.ctx, None) == capi.ERR_INVALID_ARG
    ens.close()


def test_autotune_keeps_the_state_and_picks_a_setting(require_gpu):
    sc, params, indenters, clouds = _three_scenes()
    ens = BatchedEnsemble(sc, clouds=clouds, params=params, indenters=indenters)
    before = ens.transfer_from_device().copy()
    pts_before = ens.hssoa.lines()[:, :ens.n_points].copy()
    t = ens.autotune(4)
    assert t["sort_period"] in (10, 20, 50, 100, 200, 500, 1000) and t["blocks_per_cu"] in (8, 16, 32) and t["ms_per_substep"] > 0
    assert t["points_per_chunk"] in (128, 256) and (t["fuse"] or t["points_per_chunk"] == 256)
    after = ens.transfer_from_device()
    assert np.array_equal(after.view(np.int64), before.view(np.int64))
    assert np.array_equal(ens.hssoa.lines()[:, :ens.n_points].view(np.int64), pts_before.view(np.int64))
    ens.step(3)  # the tuned context still steps
    ens.transfer_from_device()
    ens.close()
