"""Differentiable indenter force curve and a calibration harness.

The reference has no differentiation. ``ForceCurveModel`` runs the
calibration adjoint of the C-ABI (``icempm_adjoint_forward`` /
``icempm_adjoint_backward``): a forward run of ``n_steps`` substeps from a
fixed initial state that returns the indenter force averaged over every
``sample_every`` substeps, and the reverse mode of that run, which returns
dL/d(E, nu, friction angle, cohesion) for a loss L of the force samples.
``IndenterForceCurve`` wraps the pair as a ``torch.autograd.Function``;
``IndenterResponse`` adds the point positions at the end of the run (for a
displacement misfit; ``icempm_adjoint_positions`` /
``icempm_adjoint_backward_ex``); ``calibrate`` fits parameters to a target
curve with Adam.

The differentiated model is the step with its plastic return mapping; the
crush test and the return mapping's branch choices (tension cap or shear
projection, projected or not) are held as the forward took them, so the
gradient is exact where none of them switches under a small change of the
parameters. ``elastic_params`` raises the crush-test strengths so that no
point crushes: then the friction angle and cohesion do not act and their
gradients are exactly zero, and the force curve is smooth in E and nu.

``CALIBRATION_SCENE`` is the 200,000-point calibration workload: a
40x25x25-cell block, 8 points per cell, on a 128x64x64 grid of h = 1/64 m,
sphere indenter r = 0.1 m descending at 1 m/s, dt = 2e-6 s.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import replace

import numpy as np

from . import capi, scenes
from .capi import SimParams3D, check, dptr, lib
from .gpu import GPUImplementation6
from .scenes import BlockScene

PARAMS = ("YoungsModulus", "PoissonsRatio", "phi_deg", "cohesion")

CALIBRATION_SCENE = BlockScene("calibration_200k", (128, 64, 64), (40, 25, 25), ppc_axis=2, dt=2e-6, steps=400)
# the hidden parameters of the target curve (E, nu, friction angle, cohesion)
TARGET_THETA = (8e9, 0.3, 55.0, 1.5e6)


def elastic_params(scene: BlockScene, theta=None, strength_factor: float = 1e4) -> SimParams3D:
    """The scene's parameters with theta = (E, nu, phi_deg, cohesion) applied
    and the crush-test strengths multiplied by strength_factor (no point
    leaves the elastic regime)."""
    p = scenes.make_params(scene)
    for name in ("IceCompressiveStrength", "IceTensileStrength", "IceShearStrength", "IceTensileStrength2"):
        setattr(p, name, getattr(p, name) * strength_factor)
    if theta is not None:
        apply_theta(p, theta)
    p.ComputeCamClayParams()
    p.ComputeHelperVariables()
    return p


def apply_theta(p: SimParams3D, theta) -> None:
    """E, nu, friction angle (degrees) and cohesion into SimParams3D, with the
    derived constants (ComputeCamClayParams, ComputeHelperVariables)."""
    E, nu, phi, c = (float(x) for x in theta)
    p.YoungsModulus = E
    p.PoissonsRatio = nu
    p.SetPhi(phi)
    p.DP_threshold_p = c / p.DP_tan_phi
    p.ComputeCamClayParams()
    p.ComputeHelperVariables()


class ForceCurveModel:
    """Force curve F(theta) of one scene and its gradient, on one device."""

    def __init__(self, scene: BlockScene, n_steps: int, sample_every: int = 10, checkpoint_every: int = 20,
                 theta=None, device: int = 0, strength_factor: float = 1e4, points=None):
        self.scene = scene
        self.strength_factor = strength_factor
        self.params = elastic_params(scene, theta, strength_factor)
        self.theta = np.array(theta if theta is not None else
                              (self.params.YoungsModulus, self.params.PoissonsRatio, scene.phi_deg, scene.cohesion),
                              dtype=np.float64)
        self.run = capi.AdjointRun(int(n_steps), int(sample_every), int(checkpoint_every), 0.0)
        self.n_samples = n_steps // sample_every
        self.gpu = GPUImplementation6(device)
        self.gpu.hssoa = points if points is not None else scenes.make_points(scene)
        self.gpu.allocate_arrays(self.params, self.gpu.hssoa.capacity)
        self.gpu.update_constants(self.params)
        self.gpu.set_indenter(*scene.indenter())
        self.gpu.transfer_ponts_to_device()
        self.evaluations = 0

    def set_theta(self, theta) -> None:
        self.theta = np.array(theta, dtype=np.float64)
        apply_theta(self.params, self.theta)
        self.gpu.update_constants(self.params)

    def forward(self, theta=None) -> np.ndarray:
        """Force samples (n_samples, 3) at theta (default: the current one)."""
        if theta is not None:
            self.set_theta(theta)
        out = np.zeros((self.n_samples, 3))
        check(lib().icempm_adjoint_forward(self.gpu.ctx, C.byref(self.run), dptr(out)))
        self.evaluations += 1
        return out

    def positions(self) -> np.ndarray:
        """Point positions (3, n) at the end of the last forward, in the order of
        the uploaded points."""
        n = self.gpu.hssoa.size
        out = np.zeros((3, n))
        check(lib().icempm_adjoint_positions(self.gpu.ctx, dptr(out)))
        return out

    def backward(self, dL_dF: np.ndarray | None, dL_dx: np.ndarray | None = None, raw: bool = False):
        """dL/d(E, nu, phi_deg, cohesion) of the last forward for dL/dF
        (n_samples, 3) and dL/dx (3, n) of the final positions (either may be
        None); with raw, also dL/d(mu, kappa, DP_threshold_p, DP_tan_phi)."""
        g = np.zeros(4)
        r = np.zeros(4)
        w = None if dL_dF is None else np.ascontiguousarray(dL_dF, dtype=np.float64).reshape(self.n_samples, 3)
        if dL_dx is None:
            if w is None:
                w = np.zeros((self.n_samples, 3))
            check(lib().icempm_adjoint_backward(self.gpu.ctx, dptr(w), dptr(g), dptr(r)))
        else:
            x = np.ascontiguousarray(dL_dx, dtype=np.float64).reshape(3, self.gpu.hssoa.size)
            check(lib().icempm_adjoint_backward_ex(self.gpu.ctx, None if w is None else dptr(w), dptr(x), dptr(g),
                                                   dptr(r)))
        return (g, r) if raw else g

    def close(self) -> None:
        self.gpu.close()


def _torch():
    import torch
    return torch


class IndenterForceCurve:
    """torch.autograd.Function: theta (4,) float64 -> force samples (n, 3)."""

    _fn = None

    @classmethod
    def apply(cls, theta, model: ForceCurveModel):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, th, mdl):
                    ctx.model = mdl
                    f = mdl.forward(th.detach().cpu().numpy())
                    return torch.from_numpy(f).to(th.dtype)

                @staticmethod
                def backward(ctx, gF):
                    g = ctx.model.backward(gF.detach().cpu().double().numpy())
                    return torch.from_numpy(g).to(gF.dtype), None

            cls._fn = _Fn
        return cls._fn.apply(theta, model)


class IndenterResponse:
    """torch.autograd.Function: theta (4,) float64 -> (force samples (n, 3),
    final point positions (3, n_points)), for force and displacement misfits."""

    _fn = None

    @classmethod
    def apply(cls, theta, model: ForceCurveModel):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, th, mdl):
                    ctx.model = mdl
                    f = mdl.forward(th.detach().cpu().numpy())
                    x = mdl.positions()
                    return torch.from_numpy(f).to(th.dtype), torch.from_numpy(x).to(th.dtype)

                @staticmethod
                def backward(ctx, gF, gX):
                    g = ctx.model.backward(None if gF is None else gF.detach().cpu().double().numpy(),
                                           None if gX is None else gX.detach().cpu().double().numpy())
                    return torch.from_numpy(g).to(gF.dtype if gF is not None else gX.dtype), None

            cls._fn = _Fn
        return cls._fn.apply(theta, model)


def calibrate(model: ForceCurveModel, target: np.ndarray, theta0, iters: int = 20, lr: float = 0.05,
              fit=(True, True, False, False)):
    """Fits the selected parameters to the target force curve with Adam on
    theta / theta0 (loss: sum of squared force differences over the samples,
    divided by the sum of squared target forces). A step that raises the
    loss is undone and the step size halved (Adam restarts from the best
    point). Returns the history of (loss, theta) of every evaluation; the
    best entry is the fit."""
    torch = _torch()
    theta0 = torch.tensor(np.asarray(theta0, dtype=np.float64))
    u = torch.ones(4, dtype=torch.float64, requires_grad=True)
    mask = torch.tensor([1.0 if f else 0.0 for f in fit], dtype=torch.float64)
    tgt = torch.from_numpy(np.asarray(target, dtype=np.float64))
    scale = float((tgt ** 2).sum())
    opt = torch.optim.Adam([u], lr=lr)
    history = []
    best_loss, best_u = float("inf"), u.detach().clone()
    for _ in range(iters):
        opt.zero_grad()
        theta = theta0 * (1 + (u - 1) * mask)
        F = IndenterForceCurve.apply(theta, model)
        loss = ((F - tgt) ** 2).sum() / scale
        history.append((float(loss), theta.detach().numpy().copy()))
        if float(loss) > best_loss:
            with torch.no_grad():
                u.copy_(best_u)
            lr = opt.param_groups[0]["lr"] * 0.5
            opt = torch.optim.Adam([u], lr=lr)
            continue
        best_loss, best_u = float(loss), u.detach().clone()
        loss.backward()
        opt.step()
    return history


def scaled_calibration_scene(factor: float, name: str | None = None) -> BlockScene:
    """The calibration scene with grid and block scaled (quick tests)."""
    return replace(scenes.scaled(CALIBRATION_SCENE, factor), name=name or f"calibration_x{factor}")
