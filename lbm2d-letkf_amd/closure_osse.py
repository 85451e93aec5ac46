"""The learned closure inside the LETKF OSSE, compared with Smagorinsky.

The nature run is a fine run of the same LBM (grid refined by ``factor``, same c,
Smagorinsky off), box-filtered to the coarse grid at every observation time:
the filtered high-resolution truth.  Observations of (u, v) are taken from that
filtered truth on the coarse configuration's network with its error sigma.  Two
coarse ensembles, identical apart from their eddy-viscosity closure (Smagorinsky
or the trained network), assimilate the same observations; the analysis RMSE of
each against the filtered truth decides between them.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
import closure_train  # noqa: E402
import lbmletkf as lb  # noqa: E402


def closure_osse(coarse: lb.Config, params: np.ndarray, *, members: int = 64, cycles: int = 10,
                 factor: int = 4, nature_seed: int = 99, spinup: int = 200, obs_seed: int = 5):
    """Run both OSSEs; returns {"smagorinsky": [...], "mlp": [...], "free": [...]} of
    (analysis RMSE of u, v against the filtered truth) per cycle, and the forecast
    RMSE before each analysis under "<name>_forecast"."""
    fine_cfg = closure_train.fine_config(coarse, factor)
    rng = np.random.default_rng(obs_seed)
    out = {}
    with lb.Context(fine_cfg, 1) as nat:
        nat.init_member(0, nature_seed)
        nat.forecast(spinup * factor)
        truths = []
        for c in range(cycles + 1):
            truths.append(nat.filtered_macro(0, factor))
            nat.forecast(coarse.da_interval * factor)
    ctxs = {}
    for name, kind in (("smagorinsky", lb.CLOSURE_SMAGORINSKY), ("mlp", lb.CLOSURE_MLP), ("free", None)):
        cfg = coarse.copy()
        cfg.closure_kind = lb.CLOSURE_SMAGORINSKY if kind is None else kind
        ens = lb.Context(cfg, members)
        for m in range(members):
            ens.init_member(m, 1000 + m)
        if kind == lb.CLOSURE_MLP:
            ens.set_closure(params)
        ctxs[name] = ens
        out[name] = []
        out[name + "_forecast"] = []
    # one observation vector per cycle, shared by the assimilating ensembles
    any_ctx = ctxs["smagorinsky"]
    n_sites, n_obs = any_ctx.obs_count()
    sig_var = [s for bit, s in ((lb.OBS_U, coarse.obs_sigma_u), (lb.OBS_V, coarse.obs_sigma_u),
                                (lb.OBS_RHO, coarse.obs_sigma_rho)) if coarse.obs_vars & bit]
    sig = np.tile(np.array(sig_var), n_sites)
    try:
        for c in range(cycles):
            r, u, v = truths[c + 1]
            for ens in ctxs.values():
                ens.forecast(coarse.da_interval)
            yo = any_ctx.obs_from_fields(r, u, v)
            yo = (yo + sig * rng.standard_normal(n_obs)).astype(np.float32)
            for name, ens in ctxs.items():
                out[name + "_forecast"].append(ens.ensemble_stats(r, u, v)[0])
                if name != "free":
                    ens.analysis(yo)
                out[name].append(ens.ensemble_stats(r, u, v)[0])
    finally:
        for ens in ctxs.values():
            ens.close()
    return out
