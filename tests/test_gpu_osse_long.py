"""The benchmark OSSE at its full size against the FP64 oracle.

* The 10,000-step nature run at 256^2 (tau = 0.503, Cs = 0.1, K4 spectrum): the
  FP32 GPU run against the FP64 oracle run, through the energy spectrum E(k) of
  postprocess/fourier.ipynb and the vorticity of src/io.cu:61-86, every 1,000
  steps.
* 100 DA cycles of 100 steps at 256^2 (M = 16 plain and mixed, M = 64 mixed): the
  GPU ensemble assimilates the observations of the FP64 nature run, and its
  analysis RMSE against that nature run must follow the FP64 oracle ensemble's
  (tests/golden/osse_256x256_M*_100cycles.json, written by
  tests/golden/make_osse_series.py) within 1% at every cycle.

The FP64 nature run is re-computed here on the host (module fixture) and checked
against the per-cycle sums of the golden file before anything is compared.
"""
import os

import numpy as np
import pytest

import osse_series as osse

pytestmark = pytest.mark.gpu

RMSE_RTOL = 1e-2        # analysis RMSE vs the FP64 oracle, every cycle
VORTICITY_RTOL = 1e-2   # vorticity rms of the nature run vs the FP64 run, every 1,000 steps
# The FP32 and FP64 runs of this turbulent flow separate as FP32 rounding is
# amplified: the relative L2 difference of the velocity fields grows by ~1.7x per
# 1,000 steps (7e-5 at step 1,000, 5e-4 at 3,000, 2e-2 at 10,000 on MI355X).  The
# pointwise gate therefore holds to step 3,000; the per-shell spectra are held to
# 1% to step 5,000 and to 15% to step 10,000, where shells of a few tens of modes
# reflect the separated flows (measured 5e-3 and 9e-2).  Both spectra are checked:
# the phase-invariant (|u_hat|^2 + |v_hat|^2) / 2 and the form of
# postprocess/fourier.ipynb, |u_hat^2 + v_hat^2| / 2, per shell of the quarter
# plane, for shells holding at least SHELL_FLOOR of the energy.
FIELD_RTOL, FIELD_UNTIL = 1e-3, 3000
SHELL_RTOL_EARLY, SHELL_EARLY_UNTIL = 1e-2, 5000
SHELL_RTOL_LATE = 1.5e-1
SHELL_FLOOR = 1e-6


@pytest.fixture(scope="module")
def nature(lb, orc):
    orc.set_threads(min(16, os.cpu_count() or 1))
    truth, yo, spectra = osse.nature_run(lb, orc)
    return truth, yo, spectra


def _golden(M):
    path = osse.golden_path(M)
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} not written")
    return osse.load_golden(M)


def test_nature_run_reproduces_golden(nature):
    truth, _, spectra = nature
    for M in (16, 64):
        if not os.path.exists(osse.golden_path(M)):
            continue
        g = osse.load_golden(M)
        np.testing.assert_allclose(osse.truth_checksum(truth), g["truth_sum_u2_v2"], rtol=1e-10)
        for s, rec in g["nature_spectra"].items():
            np.testing.assert_allclose(spectra[int(s)][0], rec["E_k"], rtol=1e-9, atol=1e-30)


def test_nature_run_10000_steps_matches_oracle(lb, orc, nature):
    truth, _, spectra = nature
    cfg = osse.nature_config(lb)
    u0, v0 = orc.initial_velocity(cfg, osse.NATURE_SEED)
    worst_vor = 0.0
    with lb.Context(cfg, 1) as nat:
        nat.init_member(0, osse.NATURE_SEED)
        for step in sorted(spectra):
            if step:
                nat.forecast(step - nat.get_step())
            assert nat.get_step() == step
            _, u, v = (x.astype(np.float64) for x in nat.get_macro(0))
            uo, vo = (u0, v0) if step == 0 else truth[step // osse.STEPS - 1][1:]
            diff = np.sqrt(np.sum((u - uo) ** 2 + (v - vo) ** 2) / np.sum(uo ** 2 + vo ** 2))
            vor = float(np.sqrt(np.mean(osse.vorticity(u, v, cfg.dx) ** 2)))
            ek_o, vor_o, _ = spectra[step]
            keep = ek_o >= SHELL_FLOOR * ek_o.sum()
            rel_nb = np.abs(osse.energy_shells(u, v)[keep] - ek_o[keep]) / ek_o[keep]
            pw, pw_o = osse.power_shells(u, v), osse.power_shells(uo, vo)
            keep_p = pw_o >= SHELL_FLOOR * pw_o.sum()
            rel_pw = np.abs(pw[keep_p] - pw_o[keep_p]) / pw_o[keep_p]
            rel_vor = abs(vor - vor_o) / vor_o
            print(f"step {step:5d}: velocity field rel. L2 difference {diff:.2e}; vorticity rms {vor:.6e} "
                  f"(rel {rel_vor:.2e}); shells: power worst {rel_pw.max():.2e} median {np.median(rel_pw):.2e} "
                  f"({keep_p.sum()} shells), notebook worst {rel_nb.max():.2e} median {np.median(rel_nb):.2e}")
            worst_vor = max(worst_vor, rel_vor)
            assert rel_vor <= VORTICITY_RTOL
            if step <= FIELD_UNTIL:
                assert diff <= FIELD_RTOL
            shell_tol = SHELL_RTOL_EARLY if step <= SHELL_EARLY_UNTIL else SHELL_RTOL_LATE
            assert rel_pw.max() <= shell_tol and rel_nb.max() <= shell_tol
    print(f"worst vorticity rms difference over 10,000 steps: {worst_vor:.2e}")


@pytest.mark.parametrize("M,variant", [(16, "plain"), (16, "mixed"), (64, "mixed")])
def test_osse_100_cycles_follows_oracle(lb, nature, M, variant):
    g = _golden(M)
    truth, yo, _ = nature
    cfg = osse.nature_config(lb)
    cfg.variant = lb.VARIANT_PLAIN if variant == "plain" else lb.VARIANT_MIXED
    an_o = np.array(g["analysis_stats"])
    fc_o = np.array(g["forecast_stats"])
    an, fc = [], []
    with lb.Context(cfg, M) as ens:
        for k in range(M):
            ens.init_member(k, 1000 + k)
        for c in range(osse.CYCLES):
            ens.forecast(osse.STEPS)
            fc.append(ens.ensemble_stats(*truth[c]))
            ens.analysis(yo[c])
            an.append(ens.ensemble_stats(*truth[c]))
    an, fc = np.array(an), np.array(fc)
    rel = np.abs(an[:, 0] - an_o[:, 0]) / an_o[:, 0]
    rel_sp = np.abs(an[:, 1] - an_o[:, 1]) / an_o[:, 1]
    rel_fc = np.abs(fc[:, 0] - fc_o[:, 0]) / fc_o[:, 0]
    for c in list(range(0, osse.CYCLES, 10)) + [osse.CYCLES - 1]:
        print(f"cycle {c:3d}: analysis rmse {an[c, 0]:.5e} oracle {an_o[c, 0]:.5e} (rel {rel[c]:.2e}); "
              f"spread rel {rel_sp[c]:.2e}; forecast rmse rel {rel_fc[c]:.2e}")
    print(f"M={M} {variant}: analysis rmse rel. difference max {rel.max():.3e} (cycle {rel.argmax()}), "
          f"mean {rel.mean():.3e}; spread max {rel_sp.max():.3e}; forecast rmse max {rel_fc.max():.3e}")
    assert rel.max() <= RMSE_RTOL
