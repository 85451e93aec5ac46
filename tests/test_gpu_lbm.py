"""GPU parity of the forecast kernels against the FP64 oracle (run on an MI355X).

FP32 kernels against an FP64 restatement: the relative L2 difference of the
populations must stay below 1e-5 over the compared steps (a few ulp of FP32 per
step, far below any formula slip, which shows up at >= 1e-3)."""
import numpy as np
import pytest

from conftest import f32_velocity_atol, small_cfg

pytestmark = pytest.mark.gpu
TOL_F = 1e-5


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _ensemble(lb, cfg, M, seed0=1000):
    ctx = lb.Context(cfg, M)
    for k in range(M):
        ctx.init_member(k, seed0 + k)
    return ctx


def _forcing_cfg(lb, n):
    cfg = small_cfg(lb, "ref", n=n)
    # the reference's physical units (c = 20, tau ~ 0.5006) with Watanabe97 forcing
    return cfg


@pytest.mark.parametrize("variant", ["plain", "mixed"])
@pytest.mark.parametrize("kind", ["baseline", "ref_forced"])
def test_forecast_matches_oracle(lb, orc, variant, kind):
    cfg = small_cfg(lb, "baseline", n=64) if kind == "baseline" else _forcing_cfg(lb, 64)
    cfg.variant = lb.VARIANT_PLAIN if variant == "plain" else lb.VARIANT_MIXED
    M = 3
    print(f"case: forecast {kind} {variant} 64x64 M={M}", flush=True)
    with _ensemble(lb, cfg, M) as ctx:
        f = ctx.get_all_states().astype(np.float64)
        for s in range(12):
            ctx.forecast(1)
            f = orc.lbm_step(cfg, f, s)
            if s in (0, 3, 11):
                err = _rel(ctx.get_all_states(), f)
                print(f"  step {s + 1}: rel L2 {err:.2e}", flush=True)
                assert err < TOL_F
        assert ctx.get_step() == 12


def test_fused_and_plain_agree(lb):
    cfg = small_cfg(lb, "ref", n=128)
    M = 4
    out = {}
    for name, var in (("plain", lb.VARIANT_PLAIN), ("mixed", lb.VARIANT_MIXED)):
        cfg.variant = var
        with _ensemble(lb, cfg, M) as ctx:
            ctx.forecast(25)
            out[name] = ctx.get_all_states()
    err = _rel(out["mixed"], out["plain"].astype(np.float64))
    print(f"fused vs plain after 25 steps: {err:.2e}")
    assert err < 1e-6


@pytest.mark.parametrize("shape", [(48, 40), (64, 16), (128, 96)])
def test_odd_grid_shapes(lb, orc, shape):
    # (48,40) is not a multiple of the 64x16 tile: the mixed variant falls back to the plain kernels
    cfg = small_cfg(lb, "baseline", n=8)
    cfg.nx, cfg.ny = shape
    cfg.obs_stride = 8
    cfg.loc_cutoff = 7.0
    cfg.variant = lb.VARIANT_MIXED
    with _ensemble(lb, cfg, 2) as ctx:
        f = ctx.get_all_states().astype(np.float64)
        ctx.forecast(5)
        for s in range(5):
            f = orc.lbm_step(cfg, f, s)
        assert _rel(ctx.get_all_states(), f) < TOL_F


def test_mass_is_conserved(lb):
    cfg = small_cfg(lb, "ref", n=128)
    cfg.variant = lb.VARIANT_MIXED
    with _ensemble(lb, cfg, 2) as ctx:
        m0 = ctx.get_all_states().astype(np.float64).sum()
        ctx.forecast(300)
        m1 = ctx.get_all_states().astype(np.float64).sum()
    assert abs(m1 - m0) / m0 < 1e-6


def test_initial_condition_matches_oracle(lb, orc):
    for preset, seed in (("ref", 334), ("baseline", 1003)):
        cfg = small_cfg(lb, preset, n=64)
        with lb.Context(cfg, 1) as ctx:
            ctx.init_member(0, seed)
            r, u, v = ctx.get_macro(0)
        uo, vo = orc.initial_velocity(cfg, seed)
        # velocities are read back through FP32 populations (u = c sum f e / rho)
        np.testing.assert_allclose(u, uo, rtol=1e-5, atol=f32_velocity_atol(cfg))
        np.testing.assert_allclose(v, vo, rtol=1e-5, atol=f32_velocity_atol(cfg))
        np.testing.assert_allclose(r, cfg.rho_ref, rtol=1e-6)


def test_macro_roundtrip(lb, orc):
    cfg = small_cfg(lb, "baseline", n=64)
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.standard_normal((64, 64))
    u = 0.05 * rng.standard_normal((64, 64))
    v = 0.05 * rng.standard_normal((64, 64))
    with lb.Context(cfg, 2) as ctx:
        ctx.set_macro(1, rho, u, v)
        f = ctx.get_state(1)
        np.testing.assert_allclose(f, orc.feq(cfg, rho, u, v), rtol=2e-6, atol=1e-8)
        r2, u2, v2 = ctx.get_macro(1)
        np.testing.assert_allclose(r2, rho, rtol=1e-6)
        np.testing.assert_allclose(u2, u, rtol=1e-4, atol=1e-7)
        ctx.set_state(0, f)
        np.testing.assert_array_equal(ctx.get_state(0), f)


def test_full_size_batched_step(lb, orc):
    # 256^2 x 64 members, the benchmark shape: compare two members after 3 steps
    cfg = lb.preset("baseline")
    M = 64
    with _ensemble(lb, cfg, M) as ctx:
        f0 = np.stack([ctx.get_state(k) for k in (0, 63)]).astype(np.float64)
        ctx.forecast(3)
        f3 = np.stack([ctx.get_state(k) for k in (0, 63)])
    for s in range(3):
        f0 = orc.lbm_step(cfg, f0, s)
    assert _rel(f3, f0) < TOL_F


@pytest.mark.parametrize("shape", [-116, -132, -216, -232])
def test_multi_step_strips_are_bit_identical(lb, orc, shape):
    """Two or three steps per launch (no body force, no friction): the same
    per-node operations as single steps, so 7 steps (three two-step launches and a
    single step, or two three-step launches and a single step) give the single
    steps' state bit for bit, and the oracle's to TOL_F."""
    cfg = small_cfg(lb, "baseline", n=128)
    cfg.variant = lb.VARIANT_MIXED
    M = 3
    print(f"case: multi-step strips {shape} 128x128 M={M}", flush=True)
    out = []
    for sh in (1, -16, shape):
        with _ensemble(lb, cfg, M) as ctx:
            ctx.set_tuning(sh, 8)
            f0 = ctx.get_all_states().astype(np.float64)
            ctx.forecast(7)
            out.append(ctx.get_all_states())
    np.testing.assert_array_equal(out[1], out[0])  # tiles and one-step strips
    np.testing.assert_array_equal(out[2], out[0])
    f = f0
    for s in range(7):
        f = orc.lbm_step(cfg, f, s)
    assert _rel(out[1], f) < TOL_F


@pytest.mark.parametrize("nx", [512, 1024])
def test_wide_two_step_strips(lb, nx):
    """Rows of 512 and 1024 nodes: one thread per column, 80 and 160 KB of LDS
    per workgroup; the same bits as the tile shape over 5 steps."""
    cfg = small_cfg(lb, "baseline", n=64)
    cfg.nx, cfg.ny = nx, 64
    cfg.variant = lb.VARIANT_MIXED
    print(f"case: two-step strips {nx}x64 M=2", flush=True)
    out = []
    for sh in (1, -116, -132):
        with _ensemble(lb, cfg, 2) as ctx:
            ctx.set_tuning(sh, 8)
            ctx.forecast(5)
            out.append(ctx.get_all_states())
    np.testing.assert_array_equal(out[1], out[0])
    np.testing.assert_array_equal(out[2], out[0])


@pytest.mark.parametrize("steps", [1, 2, 3, 5, 100])
def test_three_step_strips_take_any_step_count(lb, steps):
    """A forecast of any length in the three-step shape (three-step launches, then
    a two-step or one-step launch for what is left) equals the tile shape's."""
    cfg = small_cfg(lb, "baseline", n=64)
    cfg.nx, cfg.ny = 256, 64
    cfg.variant = lb.VARIANT_MIXED
    print(f"case: three-step strips, {steps} steps, 256x64 M=2", flush=True)
    out = []
    for sh in (1, -216):
        with _ensemble(lb, cfg, 2) as ctx:
            ctx.set_tuning(sh, 8)
            ctx.forecast(steps)
            out.append(ctx.get_all_states())
    np.testing.assert_array_equal(out[1], out[0])


def test_wide_three_step_strips(lb):
    """Rows of 512 nodes in the three-step strips (110 KB of LDS per workgroup):
    the same bits as the tile shape over 7 steps."""
    cfg = small_cfg(lb, "baseline", n=64)
    cfg.nx, cfg.ny = 512, 64
    cfg.variant = lb.VARIANT_MIXED
    print("case: three-step strips 512x64 M=2", flush=True)
    out = []
    for sh in (1, -216, -232):
        with _ensemble(lb, cfg, 2) as ctx:
            ctx.set_tuning(sh, 8)
            ctx.forecast(7)
            out.append(ctx.get_all_states())
    np.testing.assert_array_equal(out[1], out[0])
    np.testing.assert_array_equal(out[2], out[0])


def test_two_step_strips_with_forcing_fall_back(lb):
    """A forced flow (ref preset) steps one step per launch in the two-step shape:
    the state equals the one-step strips' bit for bit."""
    cfg = small_cfg(lb, "ref", n=128)
    cfg.variant = lb.VARIANT_MIXED
    out = []
    for sh in (-16, -116, -216):
        with _ensemble(lb, cfg, 2) as ctx:
            ctx.set_tuning(sh, 8)
            ctx.forecast(4)
            out.append(ctx.get_all_states())
    np.testing.assert_array_equal(out[1], out[0])
    np.testing.assert_array_equal(out[2], out[0])


@pytest.mark.parametrize("shape", [1, 2, -8, -16, -32])
def test_launch_shapes_are_bit_identical(lb, orc, shape):
    """Every shape of the fused step (tile bands and full-row strips) computes the
    same per-node operations: bit-identical states, and the oracle's to TOL_F."""
    cfg = small_cfg(lb, "ref", n=128)
    cfg.variant = lb.VARIANT_MIXED
    M = 3
    print(f"case: launch shape {shape} 128x128 M={M}", flush=True)
    out = []
    for sh in (1, shape):
        with _ensemble(lb, cfg, M) as ctx:
            ctx.set_tuning(sh, 8)
            f0 = ctx.get_all_states().astype(np.float64)
            ctx.forecast(6)
            out.append(ctx.get_all_states())
    np.testing.assert_array_equal(out[1], out[0])
    f = f0
    for s in range(6):
        f = orc.lbm_step(cfg, f, s)
    assert _rel(out[1], f) < TOL_F
