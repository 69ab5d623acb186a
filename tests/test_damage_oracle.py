"""The oracle's damage model (no reference counterpart), pinned by closed
forms and invariants computed here from first principles:

- one crushed point, G2P on a grid at rest (so the trial F is the stored F):
  * shear branch: the returned F lies on the softened Drucker-Prager line,
    sqrt(3/2) mu J_tr^(-2/3) |dev(F F^T)| = (p_tr + f p_t) tan(phi) with
    f = 1 - 0.9 D (p_tr, J_tr from the trial F by numpy);
  * tension branch: det F_new = sqrt(2 f p_t / kappa + 1) (to the SVD's
    orthogonality, 1e-11);
  * the damage increment (q_tr - q_new) / (3 mu) / softening strain, with
    q_tr = sqrt(3/2) mu J^(-2/3) |dev(F F^T)| from the trial F, D capped
    at 1, an upper word that is not a number read as D = 0;
- whole substeps: with no softening (residual cohesion 1) every line but
  the damage word is bitwise the run without the model; D stays in [0, 1],
  never decreases, and is zero for points that never crushed.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as orc
from icempm_3dm_amd import capi, scenes
from icempm_3dm_amd.hssoa import HostSideSOA

SC = scenes.BLOCK_100K
EPS_S = 0.01


@pytest.fixture
def damage():
    d = capi.Damage.make(EPS_S, 0.1)
    orc.set_damage(d)
    yield d
    orc.set_damage(None)


def _one_point(p, F, D0=0.0, upper=None):
    hs = HostSideSOA(4)
    hs.size = 1
    L = hs.lines()
    L[capi.POSX:capi.POSX + 3, 0] = [0.5, 0.25, 0.25]
    L[capi.FE00:capi.FE00 + 9, 0] = np.asarray(F, dtype=np.float64).ravel()
    L[capi.IDX_JP_INV, 0] = 1.0
    util = np.array([capi.STATUS_CRUSHED], dtype=np.int64)
    util = capi.with_damage(util, D0)
    if upper is not None:
        util = ((util.view(np.uint64) & np.uint64(0xFFFFFFFF)) | (np.uint64(upper) << np.uint64(32))).view(np.int64)
    hs.utility()[0] = util[0]
    grid = orc.grid_zeros(p)
    assert orc.g2p(p, hs.host_buffer, hs.capacity, 1, grid, False) == 0
    Fn = L[capi.FE00:capi.FE00 + 9, 0].reshape(3, 3).copy()
    return Fn, float(capi.damage_of(hs.utility()[:1])[0]), int(hs.utility()[0])


def _pq(p, F):
    F = np.asarray(F, dtype=np.float64)
    J = np.linalg.det(F)
    b = F @ F.T
    dev = b - np.trace(b) / 3 * np.eye(3)
    q = np.sqrt(1.5) * p.mu * J ** (-2.0 / 3.0) * np.linalg.norm(dev)
    return -0.5 * p.kappa * (J * J - 1), q, J, dev


def _shear(g):
    return np.array([[1.0, g, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


@pytest.mark.parametrize("D0", [0.0, 0.5])
def test_shear_return_lands_on_softened_surface(damage, D0):
    p = scenes.make_params(SC)
    F = _shear(1e-3)
    p_tr, q_tr, J, _ = _pq(p, F)
    f = 1 - 0.9 * D0
    q_dp = (p_tr + f * p.DP_threshold_p) * p.DP_tan_phi
    pmax, pmin = p.IceCompressiveStrength, -p.IceTensileStrength
    q_fs = 2 * np.sqrt((pmax - p_tr) * (p_tr - pmin)) * p.IceShearStrength / (pmax - pmin)
    assert q_dp < q_fs and q_tr > q_dp  # the softened Drucker-Prager line is the active bound
    Fn, D, util = _one_point(p, F, D0)
    # q of the returned F at the trial J (the projection keeps the mean of b's eigenvalues)
    bn = Fn @ Fn.T
    q_ret = np.sqrt(1.5) * p.mu * J ** (-2.0 / 3.0) * np.linalg.norm(bn - np.trace(bn) / 3 * np.eye(3))
    assert abs(q_ret - q_dp) <= 1e-9 * q_dp, (q_ret, q_dp)
    D_expect = np.float32(min(1.0, D0 + (q_tr - q_dp) / (3 * p.mu) / EPS_S))
    assert abs(D - float(D_expect)) <= 2 * np.spacing(np.float32(D_expect)), (D, D_expect)
    assert util & 0xFFFFFFFF == capi.STATUS_CRUSHED


def test_tension_branch_uses_softened_cutoff(damage):
    p = scenes.make_params(SC)
    D0 = 0.8
    s = 1.001 ** (1 / 3)
    F = np.diag([s, s, s]) @ _shear(2e-4)
    p_tr, q_tr, J, _ = _pq(p, F)
    f = 1 - 0.9 * D0
    assert p_tr < -f * p.DP_threshold_p
    Fn, D, _ = _one_point(p, F, D0)
    je_new = np.sqrt(2 * f * p.DP_threshold_p / p.kappa + 1)
    # det of U diag(c, c, c) V^T: c^3 up to the orthogonality of the FP64 SVD factors
    assert abs(np.linalg.det(Fn) - je_new) <= 1e-11, (np.linalg.det(Fn), je_new)
    je_plain = np.sqrt(2 * p.DP_threshold_p / p.kappa + 1)
    assert abs(je_new - je_plain) > 1e-5  # softening is visible at this tolerance
    D_expect = np.float32(min(1.0, D0 + q_tr / (3 * p.mu) / EPS_S))
    assert abs(D - float(D_expect)) <= 2 * np.spacing(np.float32(D_expect)), (D, D_expect)


def test_damage_caps_at_one_and_reads_nan_as_zero(damage):
    p = scenes.make_params(SC)
    _, D, _ = _one_point(p, _shear(1e-2), 0.99)
    assert D == 1.0
    F = _shear(1e-3)
    p_tr, q_tr, _, _ = _pq(p, F)
    _, D_nan, util = _one_point(p, F, upper=0xFFFFFFFF)  # NaN upper word
    _, D_zero, _ = _one_point(p, F, 0.0)
    assert D_nan == D_zero and 0 < D_zero < 1
    assert util & 0xFFFFFFFF == capi.STATUS_CRUSHED


def test_inside_the_surface_damage_is_unchanged(damage):
    p = scenes.make_params(SC)
    _, D, _ = _one_point(p, _shear(1e-5), 0.3)
    assert D == float(np.float32(0.3))


def _run(p, hs, nsub, damage):
    ind = capi.Indenter.make(*PLASTIC.indenter())
    orc.set_damage(damage)
    try:
        r = hs.copy()
        bins = np.zeros(p.IndenterArraySize())
        code, _ = orc.step(p, ind, r.host_buffer, r.capacity, r.size, orc.grid_zeros(p), bins, nsub, 0.0, 0, 1)
        assert code == 0
        return r, bins
    finally:
        orc.set_damage(None)


PLASTIC = replace(SC, name="damage_plastic", grid=(24, 20, 20), cells=(6, 5, 6), dt=2e-6, indenter_radius=0.06,
                  indenter_speed=5.0)


def _plastic_start():
    p = scenes.make_params(PLASTIC)
    p.DP_threshold_p = 0.3e6 / p.DP_tan_phi
    p.ComputeHelperVariables()
    hs = scenes.make_points(PLASTIC)
    hs.utility()[:hs.size] |= capi.STATUS_CRUSHED
    return p, hs


def test_no_softening_leaves_the_state_bitwise_unchanged():
    p, hs = _plastic_start()
    plain, b0 = _run(p, hs, 20, None)
    damp, b1 = _run(p, hs, 20, capi.Damage.make(EPS_S, 1.0))
    n = hs.size
    other = [k for k in range(capi.N_PTS_ARRAYS) if k != capi.IDX_UTILITY]
    assert np.array_equal(plain.lines()[other, :n], damp.lines()[other, :n])
    assert np.array_equal(plain.utility()[:n] & 0xFFFFFFFF, damp.utility()[:n] & 0xFFFFFFFF)
    assert np.array_equal(b0, b1)
    assert (capi.damage_of(damp.utility()[:n]) > 0).any()


def test_damage_bounded_monotone_and_softens_the_force():
    p, hs = _plastic_start()
    d = capi.Damage.make(EPS_S, 0.1)
    n = hs.size
    prev = np.zeros(n)
    cur = hs
    for _ in range(4):
        cur, _ = _run(p, cur, 10, d)
        D = capi.damage_of(cur.utility()[:n])
        assert (D >= 0).all() and (D <= 1).all()
        assert (D >= prev).all()
        prev = D
    assert D.max() > 0.1
    # points never crushed keep D = 0
    p2 = scenes.make_params(PLASTIC)
    hs2 = scenes.make_points(PLASTIC)
    r2, _ = _run(p2, hs2, 10, d)
    crushed = (r2.utility()[:n] & capi.STATUS_CRUSHED) != 0
    assert (capi.damage_of(r2.utility()[:n])[~crushed] == 0).all()
    # softening lowers the indenter force against the undamaged run
    _, bp = _run(p, hs, 40, None)
    _, bd = _run(p, hs, 40, d)
    fp = np.linalg.norm(bp.reshape(-1, 3).sum(0))
    fd = np.linalg.norm(bd.reshape(-1, 3).sum(0))
    assert fd < fp, (fd, fp)
