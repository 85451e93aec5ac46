"""Pins of the oracle's learned-closure functions (CPU): central finite
differences for the gradient, closed forms for the strain stencil and the box
filter, special values of the activations, and the Smagorinsky fallback."""
import math

import numpy as np
import pytest

from conftest import small_cfg

NP, NS = 5441, 22
W1, B1, W2, B2, W3, B3 = 0, 1152, 1216, 5312, 5376, 5440


def _params(rng, scale=0.3):
    p = np.zeros(NP)
    p[W1:B1] = rng.standard_normal(B1 - W1) * scale / math.sqrt(18)
    p[B1:W2] = rng.standard_normal(64) * 0.1
    p[W2:B2] = rng.standard_normal(B2 - W2) * scale / math.sqrt(64)
    p[B2:W3] = rng.standard_normal(64) * 0.1
    p[W3:B3] = rng.standard_normal(64) * scale / math.sqrt(64)
    p[B3] = 0.2
    return p


def _samples(rng, n):
    s = np.empty((n, NS))
    s[:, :18] = rng.standard_normal((n, 18))
    s[:, 18:20] = rng.standard_normal((n, 2)) * 1e-2
    s[:, 20:22] = rng.standard_normal((n, 2)) * 1e-5
    return s


def test_loss_grad_matches_central_differences(orc):
    rng = np.random.default_rng(0)
    p, s = _params(rng), _samples(rng, 40)
    nu0, wn = 1e-3, 1e8
    loss, g = orc.closure_loss_grad(p, s, nu0, wn)
    assert loss > 0
    idx = np.concatenate([rng.choice(np.arange(W1, B1), 12, replace=False), rng.choice(np.arange(B1, W2), 6),
                          rng.choice(np.arange(W2, B2), 16, replace=False), rng.choice(np.arange(B2, W3), 6),
                          rng.choice(np.arange(W3, B3), 6), [B3]])
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (orc.closure_loss_grad(pp, s, nu0, wn)[0] - orc.closure_loss_grad(pm, s, nu0, wn)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=2e-5, abs=1e-9 * abs(loss)), k


def test_activation_special_values(orc):
    # all weights 0: h2 = gelu(b2), z = w3 . h2 + b3 -> nut = nu0 softplus(gelu(b2_0)) with w3 = e_0
    p = np.zeros(NP)
    p[W3] = 1.0
    feat = np.zeros((3, 18))
    nu0 = 1.0
    # gelu(0) = 0 and gelu(-30) ~ 0 give softplus(0) = ln 2; gelu(30) ~ 30 gives ~30
    for x0, expect in ((0.0, math.log(2.0)), (30.0, 30.0), (-30.0, math.log(2.0))):
        p[B2] = x0
        nut = orc.closure_nut(p, feat, nu0)
        assert nut[0] == pytest.approx(expect, abs=1e-12), x0
    # softplus(b3) with w3 = 0
    p[W3] = 0.0
    for z in (-5.0, 0.3, 40.0):
        p[B3] = z
        assert orc.closure_nut(p, feat, 2.0)[1] == pytest.approx(2.0 * math.log1p(math.exp(z)), rel=1e-12)


def test_features_closed_form_and_stencil_order(orc):
    nx, ny, dx, s0 = 16, 12, 0.5, 0.01
    A, B = 0.03, 0.02
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    u = A * np.sin(2 * math.pi * jj / ny)
    v = B * np.sin(2 * math.pi * ii / nx)
    feat, ab = orc.closure_features(nx, ny, dx, s0, u, v)
    # uy = A cos(2 pi j/ny) sin(2 pi/ny)/dx, vx = B cos(2 pi i/nx) sin(2 pi/nx)/dx, ux = vy = 0
    uy = A * np.cos(2 * math.pi * jj / ny) * math.sin(2 * math.pi / ny) / dx
    vx = B * np.cos(2 * math.pi * ii / nx) * math.sin(2 * math.pi / nx) / dx
    np.testing.assert_allclose(ab[:, 0], 0.0, atol=1e-15)
    np.testing.assert_allclose(ab[:, 1], (0.5 * (uy + vx)).ravel(), rtol=1e-12, atol=1e-15)
    b = (0.5 * (uy + vx))
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            s = (dj + 1) * 3 + (di + 1)
            shifted = np.roll(np.roll(b, -dj, axis=0), -di, axis=1)  # b at (i+di, j+dj)
            np.testing.assert_allclose(feat[:, 2 * s + 1], shifted.ravel() / s0, rtol=1e-12, atol=1e-13)
            np.testing.assert_allclose(feat[:, 2 * s], 0.0, atol=1e-13)


def test_coarse_grain_closed_form(orc):
    F, nxf, nyf, A = 4, 32, 64, 0.04
    jf = np.arange(nyf)
    u = np.repeat((A * np.sin(2 * math.pi * jf / nyf))[:, None], nxf, axis=1)
    v = np.zeros_like(u)
    rho = 1.0 + 0.0 * u
    samples, coarse = orc.coarse_grain(rho, u, v, F, dx_coarse=1.0, strain_scale=0.01)
    d = 2 * math.pi / nyf
    J = np.arange(nyf // F)
    th = 2 * math.pi * (F * J + (F - 1) / 2) / nyf
    ubar = A * np.sin(th) * math.sin(F * d / 2) / (F * math.sin(d / 2))
    u2 = A * A / 2 * (1 - np.cos(2 * th) * math.sin(F * d) / (F * math.sin(d)))
    np.testing.assert_allclose(coarse[1][:, 0], ubar, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(coarse[0], 1.0, rtol=1e-15)
    t1 = (u2 - ubar**2) / 2  # tau22 = 0
    np.testing.assert_allclose(samples.reshape(nyf // F, nxf // F, NS)[:, 0, 20], t1, rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(samples[:, 21], 0.0, atol=1e-18)
    assert (samples[:, 20] >= 0).all()


def test_lbm_step_closure_constant_nut_and_fallback(lb, orc):
    cfg = small_cfg(lb, "baseline", n=32)
    cfg.closure_kind = lb.CLOSURE_MLP
    cfg.closure_nu_scale = 1e-3
    cfg.closure_nu_max = 0.01
    rng = np.random.default_rng(1)
    u, v = orc.initial_velocity(cfg, 5)
    f = orc.feq(cfg, np.ones_like(u), u, v)[None]
    # w = 0, b3 = z0: nu_t = nu0 softplus(z0) everywhere == constant extra viscosity, no Smagorinsky
    p = np.zeros(NP)
    p[B3] = 0.7
    nut = 1e-3 * math.log1p(math.exp(0.7))
    got = orc.lbm_step_closure(cfg, f, 0, p)
    c2 = cfg.copy()
    c2.cs_smag = 0.0
    c2.nu = cfg.nu + nut
    np.testing.assert_allclose(got, orc.lbm_step(c2, f, 0), rtol=1e-13, atol=1e-15)
    # nu_t above closure_nu_max -> Smagorinsky everywhere
    p[B3] = 50.0
    np.testing.assert_array_equal(orc.lbm_step_closure(cfg, f, 0, p), orc.lbm_step(cfg, f, 0))
    # bf16 rounding of the operands changes nu_t by a few bf16 ulps at most
    p = _params(rng)
    feat, _ = orc.closure_features(32, 32, 1.0, cfg.closure_strain_scale, u, v)
    a, b = orc.closure_nut(p, feat, 1e-3), orc.closure_nut(p, feat, 1e-3, bf16=True)
    assert np.max(np.abs(a - b) / a) < 3e-2
