"""The committed 100-cycle oracle series (tests/golden/osse_256x256_M*_100cycles.json,
written by tests/golden/make_osse_series.py) is what the current oracle
produces: its first cycle is recomputed here on the CPU (nature run, observations,
the M = 16 ensemble's forecast and analysis statistics).  A change of the oracle
that the fixtures do not follow fails here, before any GPU test compares with
them."""
import os

import numpy as np
import pytest

import osse_series as osse


def _golden(M):
    path = osse.golden_path(M)
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} not written")
    return osse.load_golden(M)


@pytest.fixture(scope="module")
def first_cycle(lb, orc):
    orc.set_threads(min(8, os.cpu_count() or 1))
    truth, yo, _ = osse.nature_run(lb, orc, cycles=1)
    return truth, yo


@pytest.mark.parametrize("M", [16, 64])
def test_series_layout(M):
    g = _golden(M)
    assert g["members"] == M and g["cycles"] == osse.CYCLES and g["steps_per_cycle"] == osse.STEPS
    for key in ("forecast_stats", "analysis_stats", "truth_sum_u2_v2"):
        assert len(g[key]) == osse.CYCLES
    assert sorted(int(s) for s in g["nature_spectra"]) == list(range(0, 10001, osse.SPECTRUM_EVERY))
    an = np.array(g["analysis_stats"])
    fc = np.array(g["forecast_stats"])
    assert np.isfinite(an).all() and np.isfinite(fc).all()
    # the analysis reduces the velocity error of the forecast while the ensemble
    # converges, and on average afterwards (near the observation-noise floor single
    # cycles can go either way: 15% of them at M = 64)
    assert (an[:10, 0] < fc[:10, 0]).all()
    assert an[:, 0].mean() < fc[:, 0].mean()


def test_first_cycle_reproduces(lb, orc, first_cycle):
    truth, yo = first_cycle
    g = _golden(16)
    np.testing.assert_allclose(osse.truth_checksum(truth)[0], g["truth_sum_u2_v2"][0], rtol=1e-12)
    cfg = osse.nature_config(lb)
    n, M = cfg.nx, 16
    f = np.empty((M, 9, n, n))
    for k in range(M):
        u, v = orc.initial_velocity(cfg, 1000 + k)
        f[k] = orc.feq(cfg, np.full((n, n), cfg.rho_ref), u, v)
    for s in range(osse.STEPS):
        f = orc.lbm_step(cfg, f, s)
    tr, tu, tv = truth[0]
    np.testing.assert_allclose(orc.ensemble_stats(cfg, f, tr, tu, tv), g["forecast_stats"][0], rtol=1e-9)
    orc.letkf_analysis(cfg, f, yo[0])
    np.testing.assert_allclose(orc.ensemble_stats(cfg, f, tr, tu, tv), g["analysis_stats"][0], rtol=1e-9)


def test_m64_series_shares_the_nature_run(first_cycle):
    truth, _ = first_cycle
    g = _golden(64)
    np.testing.assert_allclose(osse.truth_checksum(truth)[0], g["truth_sum_u2_v2"][0], rtol=1e-12)
