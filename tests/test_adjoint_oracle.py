"""The oracle's calibration adjoint (oracle_adj_forward / oracle_adj_gradient),
pinned without the device.

Pinned by: central finite differences of the oracle's own forward run
(relative step 1e-6) for dL/dmu and dL/dkappa, and for dL/dE and dL/dnu
through the Lame relations mu = E / (2 (1 + nu)), kappa = 2 mu / 3 +
E nu / ((1 + nu)(1 - 2 nu)) (parameters_sim_3d.cpp:164-169); the forward's
force samples against the indenter accumulators of oracle_step on the same
run (the adjoint's forward is the checked step, not a second model); and the
elastic-regime guard (a run with the reference's crush strengths crushes and
is refused).
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as orc
from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.calibration import apply_theta, elastic_params

TINY = replace(scenes.BLOCK_100K, name="adj_tiny", grid=(24, 20, 20), cells=(6, 5, 6), dt=2e-6, indenter_radius=0.06)
N_STEPS, S = 40, 10


def _setup(theta=None):
    p = elastic_params(TINY, theta)
    ind = capi.Indenter.make(*TINY.indenter())
    hs = scenes.make_points(TINY)
    return p, ind, hs


def _loss(p, ind, hs, w):
    code, F = orc.adj_forward(p, ind, hs.host_buffer, hs.capacity, hs.size, N_STEPS, S)
    assert code == 0
    return float((w * F).sum())


def lame_chain(E, nu):
    """d(mu, kappa)/d(E, nu) as a 2x2 matrix [[dmu/dE, dmu/dnu], [dkappa/dE, dkappa/dnu]]."""
    D = (1 + nu) * (1 - 2 * nu)
    dmu = (1 / (2 * (1 + nu)), -E / (2 * (1 + nu) ** 2))
    dl = (nu / D, E * (1 + 2 * nu * nu) / D ** 2)
    return np.array([dmu, (2 / 3 * dmu[0] + dl[0], 2 / 3 * dmu[1] + dl[1])])


def test_forward_forces_match_step_accumulators():
    p, ind, hs = _setup()
    code, F = orc.adj_forward(p, ind, hs.host_buffer, hs.capacity, hs.size, N_STEPS, S)
    assert code == 0
    soa = hs.copy()
    t = 0.0
    for k in range(N_STEPS // S):
        bins = np.zeros(p.IndenterArraySize())
        code, t = orc.step(p, ind, soa.host_buffer, soa.capacity, soa.size, orc.grid_zeros(p), bins, S, t, 1)
        assert code == 0
        want = bins.reshape(-1, 3).sum(0) / S
        assert np.allclose(F[k], want, rtol=1e-12, atol=1e-12 * np.abs(want).max()), (k, F[k], want)
    assert np.abs(F[:, 1]).min() > 0  # the indenter is in contact from the first sample on


@pytest.mark.parametrize("name,idx", [("mu", 0), ("kappa", 1)])
def test_lame_gradient_matches_central_differences(name, idx):
    p, ind, hs = _setup()
    w = np.random.default_rng(3).standard_normal((N_STEPS // S, 3))
    code, gmu, gk, _, _ = orc.adj_gradient(p, ind, hs.host_buffer, hs.capacity, hs.size, N_STEPS, S, w)
    assert code == 0
    g = (gmu, gk)[idx]
    base = getattr(p, name)
    d = base * 1e-6
    vals = []
    for sgn in (1, -1):
        q = p.copy()
        setattr(q, name, base + sgn * d)
        vals.append(_loss(q, ind, hs, w))
    fd = (vals[0] - vals[1]) / (2 * d)
    assert abs(g - fd) <= 1e-6 * abs(fd), (name, g, fd)


def test_young_and_poisson_gradient_through_lame():
    theta = (8e9, 0.3, 55.0, 1.5e6)
    p, ind, hs = _setup(theta)
    w = np.random.default_rng(4).standard_normal((N_STEPS // S, 3))
    code, gmu, gk, gpt, gtan = orc.adj_gradient(p, ind, hs.host_buffer, hs.capacity, hs.size, N_STEPS, S, w)
    assert code == 0
    assert gpt == 0.0 and gtan == 0.0  # no point crushes
    g = np.array([gmu, gk]) @ lame_chain(theta[0], theta[1])
    for j in range(2):
        d = theta[j] * 1e-6
        vals = []
        for sgn in (1, -1):
            th = list(theta)
            th[j] += sgn * d
            q = p.copy()
            apply_theta(q, th)
            vals.append(_loss(q, ind, hs, w))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g[j] - fd) <= 1e-6 * abs(fd), (j, g[j], fd)
    # friction angle and cohesion do not act on an elastic run
    for th in ((8e9, 0.3, 40.0, 1.5e6), (8e9, 0.3, 55.0, 3e6)):
        q = p.copy()
        apply_theta(q, th)
        assert _loss(q, ind, hs, w) == _loss(p, ind, hs, w)


PLASTIC = replace(TINY, name="adj_tiny_plastic", indenter_speed=5.0)


def plastic_setup():
    """Every point crushed from the start (the crush test cannot switch),
    the reference's crush strengths and cohesion 0.3e6 Pa: under the
    indenter the return mapping projects onto the Drucker-Prager cone."""
    p = scenes.make_params(PLASTIC)
    p.DP_threshold_p = 0.3e6 / p.DP_tan_phi
    ind = capi.Indenter.make(*PLASTIC.indenter())
    hs = scenes.make_points(PLASTIC)
    hs.utility()[:hs.size] |= capi.STATUS_CRUSHED
    return p, ind, hs


@pytest.mark.parametrize("name,idx", [("mu", 0), ("kappa", 1), ("DP_threshold_p", 2), ("DP_tan_phi", 3)])
def test_plastic_gradient_matches_central_differences(name, idx):
    """With the return mapping active the loss is piecewise smooth: the
    projection switches on at q = q_max and min(q_fs, q_dp) switches
    branch. The scene is mild enough (40 substeps, 5 m/s) that no point
    crosses a switch within the 1e-6 steps; the bar is 1e-5 relative."""
    p, ind, hs = plastic_setup()
    w = np.random.default_rng(3).standard_normal((N_STEPS // S, 3))
    code, *g = orc.adj_gradient(p, ind, hs.host_buffer, hs.capacity, hs.size, N_STEPS, S, w)
    assert code == 0
    assert g[2] != 0.0 and g[3] != 0.0
    base = getattr(p, name)
    d = base * 1e-6
    vals = []
    for sgn in (1, -1):
        q = p.copy()
        setattr(q, name, base + sgn * d)
        vals.append(_loss(q, ind, hs, w))
    fd = (vals[0] - vals[1]) / (2 * d)
    assert abs(g[idx] - fd) <= 1e-5 * abs(fd), (name, g[idx], fd)


def _final_positions(p, ind, hs):
    r = hs.copy()
    code, _ = orc.step(p, ind, r.host_buffer, r.capacity, r.size, orc.grid_zeros(p), np.zeros(p.IndenterArraySize()),
                       N_STEPS, 0.0, 0, 1)
    assert code == 0
    return r.lines()[capi.POSX:capi.POSX + 3, :r.size].copy()


@pytest.mark.parametrize("with_forces", [False, True])
def test_displacement_misfit_gradient_matches_central_differences(with_forces):
    """L = sum wx . (x_N - x_0) [+ sum wF . F]: the final-position adjoint
    seeds the reverse pass. Central differences at relative step 3e-4: the
    positions barely depend on kappa over 40 substeps, so smaller steps leave
    the quotient to the positions' rounding (1.4e-6 at 3e-5, 2.8e-6 at 1e-4),
    while the truncation error is still 4e-7 at 1e-3."""
    p, ind, hs = _setup((8e9, 0.3, 55.0, 1.5e6))
    n = hs.size
    rng = np.random.default_rng(11)
    wx = rng.standard_normal((3, n))
    wF = rng.standard_normal((N_STEPS // S, 3)) if with_forces else None
    x0 = hs.lines()[capi.POSX:capi.POSX + 3, :n].copy()

    def loss(pp):
        val = float((wx * (_final_positions(pp, ind, hs) - x0)).sum())
        if with_forces:
            val += _loss(pp, ind, hs, wF)
        return val

    code, gmu, gk, _, _ = orc.adj_gradient_ex(p, ind, hs.host_buffer, hs.capacity, n, N_STEPS, S, wF, wx)
    assert code == 0
    for name, g in (("mu", gmu), ("kappa", gk)):
        v0 = getattr(p, name)
        d = v0 * 3e-4
        vals = []
        for sgn in (1, -1):
            q, _, _ = _setup((8e9, 0.3, 55.0, 1.5e6))
            setattr(q, name, v0 + sgn * d)
            vals.append(loss(q))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g - fd) <= 1e-6 * abs(fd), (name, g, fd)
