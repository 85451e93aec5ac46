"""Writes the FP64 oracle's series of the 100-cycle OSSE at 256^2 (see
tests/osse_series.py) to tests/golden/osse_256x256_M{M}_100cycles.json:

* the nature run's energy spectrum E(k), vorticity rms and velocity rms every
  1,000 steps (0 ... 10,000), and per cycle the sums of u^2 and v^2 of the
  nature state (the GPU tests re-run the FP64 nature run and check these);
* per cycle the oracle ensemble's forecast and analysis statistics against the
  nature run: [rmse_uv, spread_uv, rmse_rho, spread_rho]
  (ensemble_stat/main.cpp:108-177, real grid size).

Usage: python tests/golden/make_osse_series.py --members 16 [--threads 8]
The oracle library must be built (python -c "import __graft_entry__ as g; g.build()").
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
sys.path.insert(0, TESTS)


def load_binding():
    spec = importlib.util.spec_from_file_location("lbmletkf", os.path.join(ROOT, "lbm2d-letkf_amd", "lbmletkf.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lbmletkf"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, required=True)
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--cycles", type=int, default=None)
    args = ap.parse_args()
    import oracle_ref as orc
    import osse_series as osse

    lb = load_binding()
    orc.set_threads(args.threads)
    cycles = args.cycles or osse.CYCLES
    t0 = time.perf_counter()
    truth, yo, spectra = osse.nature_run(lb, orc, cycles)
    print(f"nature run: {time.perf_counter() - t0:.1f} s", flush=True)
    cfg = osse.nature_config(lb)
    n, M = cfg.nx, args.members
    f = np.empty((M, 9, n, n))
    for k in range(M):
        u, v = orc.initial_velocity(cfg, 1000 + k)
        f[k] = orc.feq(cfg, np.full((n, n), cfg.rho_ref), u, v)
    fc_stats, an_stats = [], []
    for c in range(cycles):
        for s in range(osse.STEPS):
            f = orc.lbm_step(cfg, f, c * osse.STEPS + s)
        tr, tu, tv = truth[c]
        fc_stats.append([float(x) for x in orc.ensemble_stats(cfg, f, tr, tu, tv)])
        orc.letkf_analysis(cfg, f, yo[c])
        an_stats.append([float(x) for x in orc.ensemble_stats(cfg, f, tr, tu, tv)])
        print(f"cycle {c}: forecast rmse {fc_stats[-1][0]:.5e}  analysis rmse {an_stats[-1][0]:.5e} "
              f"spread {an_stats[-1][1]:.5e}  ({time.perf_counter() - t0:.0f} s)", flush=True)
    out = {
        "what": "FP64 oracle OSSE: baseline preset 256^2, (u,v) every 8th node, sigma 0.01, Gaussian L=12 "
                "cutoff 43.8, inflation 1.1, 100 LBM steps per cycle",
        "members": M, "cycles": cycles, "steps_per_cycle": osse.STEPS,
        "nature_seed": osse.NATURE_SEED, "member_seeds": "1000 + k", "obs_seed": osse.OBS_SEED,
        "stats_columns": ["rmse_uv", "spread_uv", "rmse_rho", "spread_rho"],
        "forecast_stats": fc_stats,
        "analysis_stats": an_stats,
        "truth_sum_u2_v2": osse.truth_checksum(truth),
        "nature_spectra": {str(s): {"E_k": ek.tolist(), "vorticity_rms": vr, "velocity_rms": ur}
                           for s, (ek, vr, ur) in sorted(spectra.items())},
        "generated_in_s": round(time.perf_counter() - t0, 1),
        "threads": args.threads,
    }
    path = osse.golden_path(M) if cycles == osse.CYCLES else osse.golden_path(M).replace(
        f"{osse.CYCLES}cycles", f"{cycles}cycles")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", path, flush=True)


if __name__ == "__main__":
    main()
