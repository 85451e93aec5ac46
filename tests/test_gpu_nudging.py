"""GPU parity of the nudging and Lyapunov-perturbation kernels against the FP64
oracle, and the driver's da_nudging / lyapunov modes end to end."""
import os

import numpy as np
import pytest

from conftest import small_cfg
from test_gpu_driver import BIN, COMMON, N, RUN, STAT, _read, _run

pytestmark = pytest.mark.gpu


def _ens(lb, cfg, M):
    ctx = lb.Context(cfg, M)
    for k in range(M):
        ctx.init_member(k, 500 + k)
    return ctx


@pytest.mark.parametrize("xyprune,with_rho,quadra", [(1, True, 0), (4, True, 0), (4, False, 3), (8, False, 1)])
def test_nudge_matches_oracle(lb, orc, xyprune, with_rho, quadra):
    cfg = small_cfg(lb, "ref", n=64)
    M = 3
    print(f"case: nudge 64x64 M={M} xyprune={xyprune} rho={with_rho} quadra={quadra}", flush=True)
    rng = np.random.default_rng(xyprune)
    uo = (0.01 * rng.standard_normal((64, 64))).astype(np.float32)
    vo = (0.01 * rng.standard_normal((64, 64))).astype(np.float32)
    ro = (1 + 0.001 * rng.standard_normal((64, 64))).astype(np.float32)
    on = np.zeros((64, 64), bool)
    on[::xyprune, ::xyprune] = True
    for a in (uo, vo, ro):
        a[~on] = np.nan  # off-lattice values are never read
    with _ens(lb, cfg, M) as ctx:
        f = ctx.get_all_states().astype(np.float64)
        ctx.nudge(uo, vo, ro if with_rho else None, alpha=0.2, xyprune=xyprune, quadra=quadra)
        got = ctx.get_all_states()
    want = orc.nudge(cfg, f, np.nan_to_num(uo), np.nan_to_num(vo), np.nan_to_num(ro) if with_rho else None,
                     alpha=0.2, xyprune=xyprune, quadra=quadra)
    assert np.isfinite(got).all()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"  rel L2 {err:.2e}", flush=True)
    assert err < 1e-6


def test_perturb_matches_oracle(lb, orc):
    cfg = small_cfg(lb, "ref", n=64)
    print("case: perturb 64x64 M=2", flush=True)
    with _ens(lb, cfg, 2) as ctx:
        f = ctx.get_all_states().astype(np.float64)
        ctx.perturb(1, 7, 1e-2)
        got = ctx.get_all_states()
    ur, vr = orc.perturb_draws(cfg, 7)
    want1 = orc.perturb(cfg, f[1], ur, vr, 1e-2)
    np.testing.assert_array_equal(got[0], f[0].astype(np.float32))
    err = np.linalg.norm(got[1] - want1) / np.linalg.norm(want1 - f[1])
    print(f"  rel L2 of the increment {err:.2e}", flush=True)
    assert err < 1e-4


def test_driver_nudging_and_lyapunov(tmp_path):
    io = str(tmp_path / "io")
    _run([RUN, "observe", "--prefix", io] + COMMON)
    free = str(tmp_path / "free")
    os.makedirs(free)
    os.symlink(os.path.join(io, "observed"), os.path.join(free, "observed"))
    _run([RUN, "da_nudging", "--prefix", io, "--members", "2", "--nud-rate", "0.5"] + COMMON)
    _run([RUN, "free", "--prefix", free, "--members", "2"] + COMMON)

    def rmse(prefix, step, k=1):
        u = _read(prefix, "calc", k, "u", step)
        ut = _read(io, "nature", 0, "u", step)
        return float(np.sqrt(np.mean((u - ut) ** 2)))

    r_nud, r_free = rmse(io, 30), rmse(free, 30)
    print(f"u rmse at step 30: nudging {r_nud:.4e}  free {r_free:.4e}", flush=True)
    assert r_nud < 0.8 * r_free

    # lyapunov: perturbed copies of the nature run (same seed) diverge from it,
    # starting from an error of order epsilon
    ly = {}
    for eps in ("1e-6", "1e-3"):
        pre = str(tmp_path / f"ly{eps}")
        _run([RUN, "lyapunov", "--prefix", pre, "--members", "1", "--ens-offset", "0", "--ly-epsilon", eps]
             + COMMON)
        ly[eps] = [float(np.sqrt(np.mean((_read(pre, "calc", 0, "u", s) - _read(io, "nature", 0, "u", s)) ** 2)))
                   for s in (0, 15, 30)]
    print("lyapunov u rmse (steps 0, 15, 30):", ly, flush=True)
    assert 0 < ly["1e-6"][0] < ly["1e-3"][0]
    assert ly["1e-3"][0] < 1e-2
