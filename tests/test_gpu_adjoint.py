"""The calibration adjoint on the device (icempm_adjoint_forward / _backward).

Checked against: the oracle's forward force samples and its reverse mode
(dL/dmu, dL/dkappa; the oracle adjoint is pinned by finite differences in
test_adjoint_oracle.py); central finite differences of the device's own
forward (relative step 1e-6) for dL/dE and dL/dnu; independence of the
checkpoint period; a plastic run through the return mapping (all four
parameter adjoints against the oracle and against finite differences); and
the torch.autograd wrapper and calibration loop; and a displacement misfit
(a loss of the final point positions, alone and with the forces): the final
positions against the oracle's step, the parameter adjoints against the
oracle's reverse mode and against central differences of the device forward.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as orc
from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.calibration import (ForceCurveModel, IndenterForceCurve, IndenterResponse, TARGET_THETA,
                                        calibrate, elastic_params)
from parity_util import normwise

pytestmark = pytest.mark.gpu

TINY = replace(scenes.BLOCK_100K, name="adj_tiny", grid=(24, 20, 20), cells=(6, 5, 6), dt=2e-6, indenter_radius=0.06)
SMALL = replace(scenes.BLOCK_100K, name="adj_small", grid=(48, 32, 32), cells=(16, 10, 12), dt=2e-6,
                indenter_radius=0.1)


def _oracle(scene, theta, n_steps, S, w):
    p = elastic_params(scene, theta)
    ind = capi.Indenter.make(*scene.indenter())
    hs = scenes.make_points(scene)
    code, F = orc.adj_forward(p, ind, hs.host_buffer, hs.capacity, hs.size, n_steps, S)
    assert code == 0
    code, *g = orc.adj_gradient(p, ind, hs.host_buffer, hs.capacity, hs.size, n_steps, S, w)
    assert code == 0
    return F, np.array(g)


@pytest.mark.parametrize("scene,n_steps,K", [(TINY, 40, 10), (SMALL, 60, 20)])
def test_forward_and_gradient_match_oracle(require_gpu, scene, n_steps, K):
    theta = TARGET_THETA
    S = 10
    w = np.random.default_rng(5).standard_normal((n_steps // S, 3))
    F_ref, g_ref = _oracle(scene, theta, n_steps, S, w)
    m = ForceCurveModel(scene, n_steps, S, K, theta=theta)
    F = m.forward()
    assert normwise(F, F_ref) <= 1e-10, (F, F_ref)
    g, raw = m.backward(w, raw=True)
    assert normwise(raw[:2], g_ref[:2]) <= 1e-8, (raw, g_ref)
    assert g[2] == 0.0 and g[3] == 0.0 and raw[2] == 0.0 and raw[3] == 0.0
    print(scene.name, "forces", F[-1], "dL/dmu,kappa", raw, "oracle", g_ref)
    m.close()


def test_gradient_matches_device_finite_differences(require_gpu):
    theta = np.array(TARGET_THETA)
    n_steps, S, K = 60, 10, 20
    m = ForceCurveModel(SMALL, n_steps, S, K, theta=theta)
    w = np.random.default_rng(6).standard_normal((n_steps // S, 3))
    m.forward(theta)
    g = m.backward(w)
    for j in range(2):
        d = theta[j] * 1e-6
        vals = []
        for sgn in (1, -1):
            th = theta.copy()
            th[j] += sgn * d
            vals.append(float((w * m.forward(th)).sum()))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g[j] - fd) <= 1e-6 * abs(fd), (j, g[j], fd)
    m.close()


def test_checkpoint_period_does_not_change_the_gradient(require_gpu):
    n_steps, S = 40, 10
    w = np.random.default_rng(7).standard_normal((n_steps // S, 3))
    grads = []
    for K in (5, 10, 40):
        m = ForceCurveModel(TINY, n_steps, S, K, theta=TARGET_THETA)
        m.forward()
        grads.append(m.backward(w))
        m.close()
    for g in grads[1:]:
        assert normwise(g, grads[0]) <= 1e-12, grads


PLASTIC = replace(TINY, name="adj_tiny_plastic", indenter_speed=5.0)
PLASTIC_THETA = (8e9, 0.3, 55.0, 0.3e6)


def _plastic_model(n_steps, S, K):
    """Every point crushed from the start, the reference's crush strengths,
    cohesion 0.3e6 Pa: the return mapping projects under the indenter."""
    hs = scenes.make_points(PLASTIC)
    hs.utility()[:hs.size] |= capi.STATUS_CRUSHED
    m = ForceCurveModel(PLASTIC, n_steps, S, K, theta=PLASTIC_THETA, strength_factor=1.0, points=hs)
    return m, hs


def test_plastic_gradient_matches_oracle_and_finite_differences(require_gpu):
    n_steps, S, K = 40, 10, 20
    m, hs = _plastic_model(n_steps, S, K)
    w = np.random.default_rng(8).standard_normal((n_steps // S, 3))
    F = m.forward()
    g, raw = m.backward(w, raw=True)
    # the oracle on the same start state and parameters
    p = m.params.copy()
    ind = capi.Indenter.make(*PLASTIC.indenter())
    code, F_ref = orc.adj_forward(p, ind, hs.host_buffer, hs.capacity, hs.size, n_steps, S)
    assert code == 0
    code, *g_ref = orc.adj_gradient(p, ind, hs.host_buffer, hs.capacity, hs.size, n_steps, S, w)
    assert code == 0
    assert normwise(F, F_ref) <= 1e-9, (F, F_ref)
    for k in range(4):
        assert abs(raw[k] - g_ref[k]) <= 1e-6 * abs(g_ref[k]), (k, raw, g_ref)
    assert g[2] != 0.0 and g[3] != 0.0
    # central differences of the device forward in E, nu, phi, c (relative step
    # 1e-6). Bar 5e-5: the device forward's own rounding (atomic summation
    # order) moves the loss by ~1e-12 relative, which a 1e-6 step turns into
    # ~1e-5 relative noise in the cohesion difference quotient
    theta = np.array(PLASTIC_THETA)
    for j in range(4):
        d = theta[j] * 1e-6
        vals = []
        for sgn in (1, -1):
            th = theta.copy()
            th[j] += sgn * d
            vals.append(float((w * m.forward(th)).sum()))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g[j] - fd) <= 5e-5 * abs(fd), (j, g[j], fd)
    print("plastic dL/d(E, nu, phi, c)", g)
    m.close()


def test_torch_function_and_calibration_loop(require_gpu):
    torch = pytest.importorskip("torch")
    n_steps, S, K = 40, 10, 20
    m = ForceCurveModel(TINY, n_steps, S, K, theta=TARGET_THETA)
    target = m.forward(TARGET_THETA)
    theta0 = np.array(TARGET_THETA) * np.array([1.2, 0.9, 1.0, 1.0])
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    F = IndenterForceCurve.apply(th, m)
    loss = ((F - torch.from_numpy(target)) ** 2).sum()
    loss.backward()
    g_direct = m.backward(2 * (F.detach().numpy() - target))
    assert np.allclose(th.grad.numpy(), g_direct, rtol=1e-12, atol=0)
    hist = calibrate(m, target, theta0, iters=12, lr=0.03)
    best = min(hist, key=lambda h: h[0])
    assert best[0] < 0.2 * hist[0][0], [h[0] for h in hist]
    assert abs(best[1][0] / TARGET_THETA[0] - 1) < abs(theta0[0] / TARGET_THETA[0] - 1)
    m.close()


def test_calibration_scene_gradient_and_fit(require_gpu):
    """The 200,000-point calibration workload (128x64x64 grid, 400 substeps,
    force every 10, checkpoints every 20): the gradient of the force misfit
    against central differences at relative step 1e-6, then Adam from a guess
    with E 30% high and nu 30% low reduces the misfit a thousandfold (the
    curve pins a combination of E and nu), and E alone is recovered to 1%."""
    from icempm_3dm_amd.calibration import CALIBRATION_SCENE
    n_steps, S, K = 400, 10, 20
    m = ForceCurveModel(CALIBRATION_SCENE, n_steps, S, K, theta=TARGET_THETA)
    assert m.gpu.hssoa.size == 200_000
    target = m.forward(TARGET_THETA)
    theta0 = np.array(TARGET_THETA) * np.array([1.3, 0.7, 1.0, 1.0])
    F0 = m.forward(theta0)
    w = 2 * (F0 - target)
    g = m.backward(w)
    for j in range(2):
        d = theta0[j] * 1e-6
        vals = []
        for sgn in (1, -1):
            th = theta0.copy()
            th[j] += sgn * d
            vals.append(float(((m.forward(th) - target) ** 2).sum()))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g[j] - fd) <= 1e-6 * abs(fd), (j, g[j], fd)
    # E and nu together: the force curve pins a combination of the two
    hist = calibrate(m, target, theta0, iters=15, lr=0.04)
    best = min(hist, key=lambda h: h[0])
    print("misfit (E, nu)", [round(h[0], 8) for h in hist], "best theta", best[1])
    assert best[0] < 1e-3 * hist[0][0]
    # E alone (nu at its true value) is recovered
    th0 = np.array(TARGET_THETA) * np.array([1.3, 1.0, 1.0, 1.0])
    hist = calibrate(m, target, th0, iters=15, lr=0.04, fit=(True, False, False, False))
    best = min(hist, key=lambda h: h[0])
    print("misfit (E)", [round(h[0], 8) for h in hist], "best theta", best[1])
    assert abs(best[1][0] / TARGET_THETA[0] - 1) < 0.01
    m.close()


@pytest.mark.parametrize("with_forces", [False, True])
def test_displacement_misfit_gradient(require_gpu, with_forces):
    n_steps, S, K = 40, 10, 20
    theta = np.array(TARGET_THETA)
    m = ForceCurveModel(SMALL, n_steps, S, K, theta=theta)
    hs = m.gpu.hssoa
    n = hs.size
    F = m.forward()
    X = m.positions()
    # the final positions are the step's (upload order)
    p = elastic_params(SMALL, theta)
    ind = capi.Indenter.make(*SMALL.indenter())
    ref = hs.copy()
    code, _ = orc.step(p, ind, ref.host_buffer, ref.capacity, n, orc.grid_zeros(p), np.zeros(p.IndenterArraySize()),
                       n_steps, 0.0, 0, 1)
    assert code == 0
    assert normwise(X, ref.lines()[capi.POSX:capi.POSX + 3, :n]) <= 1e-12
    rng = np.random.default_rng(12)
    wx = rng.standard_normal((3, n))
    wF = rng.standard_normal((n_steps // S, 3)) if with_forces else None
    g, raw = m.backward(wF, wx, raw=True)
    code, *g_ref = orc.adj_gradient_ex(p, ind, hs.host_buffer, hs.capacity, n, n_steps, S, wF, wx)
    assert code == 0
    assert normwise(raw[:2], np.array(g_ref[:2])) <= 1e-8, (raw, g_ref)
    # central differences of the device forward (relative step 3e-4: the
    # positions' own rounding limits smaller steps)
    x0 = hs.lines()[capi.POSX:capi.POSX + 3, :n].copy()

    def loss(th):
        f = m.forward(th)
        val = float((wx * (m.positions() - x0)).sum())
        return val + (float((wF * f).sum()) if with_forces else 0.0)

    for j in range(2):
        d = theta[j] * 3e-4
        vals = []
        for sgn in (1, -1):
            th = theta.copy()
            th[j] += sgn * d
            vals.append(loss(th))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g[j] - fd) <= 1e-6 * abs(fd), (j, g[j], fd)
    m.close()


def test_torch_response_function(require_gpu):
    torch = pytest.importorskip("torch")
    n_steps, S, K = 40, 10, 20
    m = ForceCurveModel(TINY, n_steps, S, K, theta=TARGET_THETA)
    n = m.gpu.hssoa.size
    th = torch.tensor(np.array(TARGET_THETA), dtype=torch.float64, requires_grad=True)
    F, X = IndenterResponse.apply(th, m)
    rng = np.random.default_rng(13)
    wF = torch.from_numpy(rng.standard_normal((n_steps // S, 3)))
    wx = torch.from_numpy(rng.standard_normal((3, n)))
    loss = (wF * F).sum() + (wx * X).sum()
    loss.backward()
    g_direct = m.backward(wF.numpy(), wx.numpy())
    assert np.allclose(th.grad.numpy(), g_direct, rtol=1e-12, atol=0)
    m.close()
