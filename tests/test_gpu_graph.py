"""Graph mode of icempm_step (icempm_set_graph): each run of substeps between
two brick re-sorts is captured once into a hipGraph and replayed, with the
simulation time and the p/q record flag taken from a device-side clock.

Every replay is checked: a graph-mode context and a launched (eager) context
start from the same points and advance call by call in the same process, with
a download and synchronization after each call; after every call both are
compared with the oracle run to the same substep, to the bar of the
multi-substep parity tests (the larger of 1e-12 and 10x the oracle's own
rounding spread), utility bits and the simulation time exactly. The record
period (UpdateEveryNthStep = 4) does not divide the 6-substep calls, so the
substeps that record p/q move within the captured runs from call to call. A
change of the elastic constants between calls must capture anew. Fused runs
between re-sorts every 3 substeps, unfused single substeps and fused runs of
a whole call (no re-sort) are covered.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as orc
from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.ensemble import BatchedEnsemble
from parity_util import FIELDS, normwise
from test_gpu_ensemble import _three_scenes
from test_gpu_parity import _download, _gpu_with, _oracle_spread

pytestmark = pytest.mark.gpu

MID = replace(scenes.BLOCK_100K, name="graph_mid", grid=(48, 32, 32), cells=(16, 10, 12), dt=2e-6,
              indenter_radius=0.1)
LINES = dict(FIELDS, P=[capi.IDX_P], Q=[capi.IDX_Q])


def _spread_from(p, ind, start, nsub, t0, step, ref):
    """How far rounding alone moves the oracle over these nsub substeps (from
    time t0, substep number step): a slab-parallel P2G summation order and
    FMA contraction off, against ref (serial, contraction on)."""
    n = ref.size
    runs = []
    for kw in ({"nthreads": 8}, {"nofma": True}):
        r = start.copy()
        code, _ = orc.step(p, ind, r.host_buffer, r.capacity, n, orc.grid_zeros(p),
                           np.zeros(p.IndenterArraySize()), nsub, t0, step, **kw)
        assert code == 0
        runs.append(r)
    return {name: max(normwise(r.lines()[ls, :n], ref.lines()[ls, :n]) for r in runs) for name, ls in LINES.items()}


def _check(back, ref, spread, n, tag):
    for name, ls in LINES.items():
        e = normwise(back.lines()[ls, :n], ref.lines()[ls, :n])
        assert e <= max(1e-12, 10 * spread[name]), (tag, name, e, spread[name])
    assert np.array_equal(back.utility()[:n], ref.utility()[:n]), tag


@pytest.mark.parametrize("fuse,period", [(True, 3), (False, 3), (True, 0)])
def test_graph_replays_match_launched_path_and_oracle(require_gpu, fuse, period):
    p = scenes.make_params(MID)
    p.UpdateEveryNthStep = 4
    shape, center, vel = MID.indenter()
    ind = capi.Indenter.make(shape, center, vel)
    hs = scenes.perturbed_points(MID, seed=41, amp_F=5e-4, amp_v=0.02, amp_B=1.0)
    ref = hs.copy()
    gr = _gpu_with(p, hs.copy(), indenter=(shape, center, vel))
    ea = _gpu_with(p, hs.copy(), indenter=(shape, center, vel))
    for g in (gr, ea):
        g.set_sort_period(period)
        g.set_fusion(fuse)
        g.reset_indenter_force_accumulator()
    gr.set_graph(True)
    gr.enable_timing(True)
    n = ref.size
    grid_ref = orc.grid_zeros(p)
    bins_ref = np.zeros(p.IndenterArraySize())
    t_gr = t_ea = t_ref = 0.0
    step = 0
    nsub = 6
    for call in range(5):
        if call == 3:
            # new elastic constants: the cached graphs carry the old ones
            p.YoungsModulus *= 1.25
            p.ComputeCamClayParams()
            p.ComputeHelperVariables()
            before = gr.graph_captures()[0]
            for g in (gr, ea):
                g.update_constants(p)
        start, t0 = ref.copy(), t_ref
        t_gr = gr.step(nsub, t_gr, step)
        t_ea = ea.step(nsub, t_ea, step)
        code, t_ref = orc.step(p, ind, ref.host_buffer, ref.capacity, n, grid_ref, bins_ref, nsub, t_ref, step, 1)
        assert code == 0
        assert t_gr == t_ref and t_ea == t_ref, (call, t_gr, t_ea, t_ref)
        # rounding sensitivity of this call's substeps, added to that of the earlier calls
        sp = _spread_from(p, ind, start, nsub, t0, step, ref)
        bar = {k: bar.get(k, 0.0) + sp[k] for k in sp} if call else sp
        b_gr, b_ea = _download(gr), _download(ea)
        _check(b_gr, ref, bar, n, ("graph", call))
        _check(b_ea, ref, bar, n, ("launched", call))
        if call == 3:
            assert gr.graph_captures()[0] > before
        step += nsub
    fd = gr.read_indenter_bins().reshape(-1, 3).sum(0)
    fe = ea.read_indenter_bins().reshape(-1, 3).sum(0)
    fr = bins_ref.reshape(-1, 3).sum(0)
    assert np.abs(fr).sum() > 0
    assert normwise(fd, fr) <= 1e-9 and normwise(fe, fr) <= 1e-9, (fd, fe, fr)
    captures, kept = gr.graph_captures()
    replays = gr.read_timings()["graph"][1]
    # 30 substeps in runs of 3 (fused, re-sort every 3), of 1 (unfused) or of
    # a whole 6-substep call (no re-sort), replayed from a handful of captures
    run = 3 if (fuse and period) else (1 if not fuse else 6)
    assert replays == 30 // run, (replays, run)
    # distinct starting states: at most the runs of one call, once before
    # and once after the parameter change, plus the first call's
    assert 1 <= captures <= 2 * (6 // run) + 2 and kept <= 16, (captures, kept)
    print("graph captures", captures, "kept", kept, "replays", replays)
    gr.set_graph(False)
    assert gr.graph_captures()[1] == 0
    gr.close()
    ea.close()


def test_graph_mode_ensemble_matches_oracle(require_gpu):
    sc, params, indenters, clouds = _three_scenes()
    starts = [c.copy() for c in clouds]
    ens = BatchedEnsemble(sc, clouds=clouds, params=params, indenters=indenters)
    ens.set_sort_period(4)
    ens.set_graph(True)
    ens.reset_indenter_force_accumulator()
    nsub = 8
    t_dev = ens.step(nsub)
    t_dev = ens.step(nsub)  # replays the runs captured by the first call
    assert ens.graph_captures()[0] <= 4
    raw = ens.transfer_from_device()
    for i in range(3):
        p, start = params[i], starts[i]
        ind = capi.Indenter.make(*indenters[i])
        ref = start.copy()
        bins_ref = np.zeros(p.IndenterArraySize())
        code, t_ref = orc.step(p, ind, ref.host_buffer, ref.capacity, ref.size, orc.grid_zeros(p), bins_ref,
                               2 * nsub, 0.0, 0, 1)
        assert code == 0
        assert t_dev == t_ref
        spread = _oracle_spread(p, ind, start, 2 * nsub, ref)
        got = ens.scene_points(i)
        n = ref.size
        for name, ls in LINES.items():
            e = normwise(got[ls, :], ref.lines()[ls, :n])
            assert e <= max(1e-12, 10 * spread[name]), (i, name, e, spread[name])
        assert np.array_equal(got[capi.IDX_UTILITY].view(np.int64), ref.utility()[:n]), i
        fd = ens.scene_bins(raw, i).reshape(-1, 3).sum(0)
        fr = bins_ref.reshape(-1, 3).sum(0)
        assert normwise(fd, fr) <= 1e-9, (i, fd, fr)
    ens.close()
