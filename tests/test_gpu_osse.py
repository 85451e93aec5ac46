"""End-to-end OSSE on the GPU: analysis RMSE against the nature run follows the
FP64 oracle cycle by cycle, and assimilation beats the free ensemble."""
import numpy as np
import pytest

from conftest import small_cfg

pytestmark = pytest.mark.gpu


def _run_osse(lb, orc, cfg, M, cycles, steps, with_oracle=True):
    ens = lb.Context(cfg, M)
    nat = lb.Context(cfg, 1)
    for k in range(M):
        ens.init_member(k, 1000 + k)
    nat.init_member(0, 7)
    f_orc = ens.get_all_states().astype(np.float64) if with_oracle else None
    rm_gpu, rm_orc = [], []
    with ens, nat:
        for c in range(cycles):
            nat.forecast(steps)
            ens.forecast(steps)
            if with_oracle:
                for s in range(steps):
                    f_orc = orc.lbm_step(cfg, f_orc, c * steps + s)
            yo = nat.observe(0, seed=99, counter=c, noise=True)
            ens.analysis(yo)
            tr, tu, tv = nat.get_macro(0)
            rm_gpu.append(ens.ensemble_stats(tr, tu, tv))
            if with_oracle:
                orc.letkf_analysis(cfg, f_orc, yo.astype(np.float64))
                rm_orc.append(orc.ensemble_stats(cfg, f_orc, tr, tu, tv))
    return np.array(rm_gpu), (np.array(rm_orc) if with_oracle else None)


@pytest.mark.parametrize("variant", ["plain", "mixed"])
def test_osse_rmse_follows_oracle(lb, orc, variant):
    cfg = small_cfg(lb, "baseline", n=64, obs_stride=4, loc_scale=6.0, loc_cutoff=16.0)
    cfg.variant = lb.VARIANT_PLAIN if variant == "plain" else lb.VARIANT_MIXED
    g, o = _run_osse(lb, orc, cfg, M=16, cycles=6, steps=20)
    for c in range(len(g)):
        print(f"cycle {c}: rmse gpu {g[c, 0]:.6e} oracle {o[c, 0]:.6e}  spread gpu {g[c, 1]:.4e} oracle {o[c, 1]:.4e}")
    np.testing.assert_allclose(g[:, 0], o[:, 0], rtol=1e-2)
    np.testing.assert_allclose(g[:, 1], o[:, 1], rtol=1e-2)


def test_assimilation_beats_free_run(lb, orc):
    cfg = small_cfg(lb, "baseline", n=128, obs_stride=8, loc_scale=12.0, loc_cutoff=40.0)
    cfg.variant = lb.VARIANT_MIXED
    M, cycles, steps = 16, 8, 50
    g, _ = _run_osse(lb, orc, cfg, M, cycles, steps, with_oracle=False)
    # free ensemble: same members, no analysis
    ens = lb.Context(cfg, M)
    nat = lb.Context(cfg, 1)
    for k in range(M):
        ens.init_member(k, 1000 + k)
    nat.init_member(0, 7)
    with ens, nat:
        ens.forecast(cycles * steps)
        nat.forecast(cycles * steps)
        free = ens.ensemble_stats(*nat.get_macro(0))
    print(f"rmse DA {g[-1, 0]:.4e} vs free {free[0]:.4e}; first cycle {g[0, 0]:.4e}")
    assert g[-1, 0] < 0.5 * free[0]
    assert g[-1, 0] < g[0, 0]
