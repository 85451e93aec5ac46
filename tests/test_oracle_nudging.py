"""Pins of the oracle's nudging and Lyapunov perturbation (CPU): alpha = 0 / 1
limits, exact bilinear interpolation of linear fields from the lattice, the
quadrant masks, conservation of the member's density without an observed one,
and the statistics of the perturbation draws."""
import numpy as np
import pytest

from conftest import small_cfg


def _state(lb, orc, n=32, M=2):
    cfg = small_cfg(lb, "baseline", n=n)
    fs = []
    for k in range(M):
        u, v = orc.initial_velocity(cfg, 10 + k)
        fs.append(orc.feq(cfg, 1.0 + 0.01 * np.cos(u), u, v))
    return cfg, np.stack(fs)


def test_nudge_limits_and_interpolation(lb, orc):
    cfg, f = _state(lb, orc)
    n = cfg.nx
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    # fields linear inside each lattice cell are reproduced exactly between lattice points
    uo = 0.01 + 1e-3 * ii + 2e-4 * jj
    vo = -0.02 + 5e-4 * ii - 1e-4 * jj
    ro = 1.0 + 1e-4 * ii + 3e-5 * jj
    g = orc.nudge(cfg, f, uo, vo, ro, alpha=0.0, xyprune=4)
    np.testing.assert_array_equal(g, f)
    g = orc.nudge(cfg, f, uo, vo, ro, alpha=1.0, xyprune=4)
    # away from the periodic seam (the last lattice cell interpolates towards column 0)
    inner = (ii < n - 4) & (jj < n - 4)
    want = orc.feq(cfg, ro, uo, vo)
    for m in range(f.shape[0]):
        np.testing.assert_allclose(g[m][:, inner], want[:, inner], rtol=1e-12, atol=1e-15)
    # at the seam the east/north corners wrap to column/row 0
    i, j = n - 2, 5
    ir = 0.5
    u_seam = (1 - ir) * uo[4, n - 4] + ir * uo[4, 0]  # j = 5 lies in the cell of rows 4..8, jr = 0.25
    u_seam = 0.75 * u_seam + 0.25 * ((1 - ir) * uo[8, n - 4] + ir * uo[8, 0])
    _, u_g, _ = orc.macro(cfg, g[0])
    assert u_g[j, i] == pytest.approx(u_seam, rel=1e-12)
    # alpha blends linearly
    a = 0.3
    g = orc.nudge(cfg, f, uo, vo, ro, alpha=a, xyprune=1)
    np.testing.assert_allclose(g, a * orc.feq(cfg, ro, uo, vo)[None] + (1 - a) * f, rtol=1e-13, atol=1e-16)


def test_nudge_quadrants_and_own_density(lb, orc):
    cfg, f = _state(lb, orc)
    n = cfg.nx
    uo = np.full((n, n), 0.03)
    vo = np.full((n, n), -0.01)
    for quadra, (ie, jn) in {1: (True, False), 2: (False, True), 3: (True, True)}.items():
        g = orc.nudge(cfg, f, uo, vo, None, alpha=0.5, xyprune=2, quadra=quadra)
        jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        frozen = ((ii >= n // 2) & ie) | ((jj >= n // 2) & jn)
        np.testing.assert_array_equal(g[:, :, frozen], f[:, :, frozen])
        assert np.abs(g[:, :, ~frozen] - f[:, :, ~frozen]).max() > 1e-4
        # without an observed density the member's own density is kept (feq conserves it)
        r0, _, _ = orc.macro(cfg, f[1])
        r1, _, _ = orc.macro(cfg, g[1])
        np.testing.assert_allclose(r1, r0, rtol=1e-14)


def test_perturbation_draws_and_update(lb, orc):
    cfg, f = _state(lb, orc, n=64, M=1)
    ur, vr = orc.perturb_draws(cfg, 3)
    ur2, _ = orc.perturb_draws(cfg, 3)
    np.testing.assert_array_equal(ur, ur2)
    assert not np.array_equal(ur, orc.perturb_draws(cfg, 4)[0])
    both = np.concatenate([ur.ravel(), vr.ravel()])
    assert abs(both.mean()) < 0.02 and abs(both.std() - cfg.u_ref) < 0.02
    # draws are FP32 values
    np.testing.assert_array_equal(both, both.astype(np.float32))
    eps = 1e-3
    g = orc.perturb(cfg, f[0], ur, vr, eps)
    r0, _, _ = orc.macro(cfg, f[0])
    np.testing.assert_allclose(g, eps * orc.feq(cfg, r0, ur, vr) + (1 - eps) * f[0], rtol=1e-13, atol=1e-17)
    r1, _, _ = orc.macro(cfg, g)
    np.testing.assert_allclose(r1, r0, rtol=1e-14)
