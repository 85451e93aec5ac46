"""The autotuner: it times its candidates without changing the ensemble, keeps a
valid choice, writes and then hits its cache, and a tuned analysis still agrees
with the oracle to the mixed variant's tolerance."""
import numpy as np
import pytest

from conftest import small_cfg
from test_gpu_letkf import TOLS, _osse, _rel

pytestmark = pytest.mark.gpu


def test_autotune_keeps_state_and_caches(lb, orc, tmp_path):
    cfg = small_cfg(lb, "baseline", n=128, loc_scale=6.0, loc_cutoff=20.0, obs_stride=8)
    M = 16
    cache = str(tmp_path / "tune.txt")
    print("case: autotune 128x128 M=16", flush=True)
    ens, nat = _osse(lb, cfg, M, spin=5)
    with ens, nat:
        f0 = ens.get_all_states()
        rep = ens.autotune(cache)
        print("  " + rep, flush=True)
        np.testing.assert_array_equal(ens.get_all_states(), f0)
        tpb, chain = ens.get_tuning()
        assert tpb in (1, 2, 4, -8, -16) and chain in (1, 2, 4, 8, 16)
        assert "lbm us/step" in rep and "solve us" in rep
        # the second call is a cache hit with the same choice
        ens.set_tuning(1, 1)
        rep2 = ens.autotune(cache)
        assert rep2.startswith("cached") and ens.get_tuning() == (tpb, chain)
        assert open(cache).read().count("\n") == 1
        # a tuned analysis against the oracle
        yo = nat.observe(0, seed=2, counter=5, noise=True)
        fb = ens.get_all_states().astype(np.float64)
        g0, nd = 128 * 64 + 5, 16
        ens.set_dump(g0, nd)
        ens.analysis(yo)
        W, w, ev = ens.get_dump()
        fa = ens.get_all_states().astype(np.float64)
    fo = fb.copy()
    Wo, wo, evo = orc.letkf_analysis(cfg, fo, yo.astype(np.float64), g0, nd)
    errW = max(_rel(W[i], Wo[i]) for i in range(nd))
    inc = _rel(fa - fb, fo - fb)
    print(f"  tuned: W rel {errW:.2e}  increment rel {inc:.2e}", flush=True)
    assert errW < TOLS["mixed3"][0] and inc < TOLS["mixed3"][1]


def test_set_tuning_rejects(lb):
    cfg = small_cfg(lb, "baseline", n=64, loc_scale=4.0, loc_cutoff=10.0)
    with lb.Context(cfg, 3) as ctx:  # 64x64 x 3 members = 12 tiles
        with pytest.raises(lb.LbmLetkfError):
            ctx.set_tuning(0, 8)
        with pytest.raises(lb.LbmLetkfError):
            ctx.set_tuning(8, 8)  # 12 tiles are not a multiple of 8
        ctx.set_tuning(4, 2)
        assert ctx.get_tuning() == (4, 2)
        with pytest.raises(lb.LbmLetkfError):
            ctx.set_tuning(-12, 2)  # strip heights are 8, 16 or 32
        ctx.set_tuning(-16, 2)
        assert ctx.get_tuning() == (-16, 2)
        ctx.set_tuning(-116, 2)  # two-step strips of 16 rows
        assert ctx.get_tuning() == (-116, 2)
        for bad in (-108, -100, -50):  # two-step strips are 16 or 32 rows
            with pytest.raises(lb.LbmLetkfError):
                ctx.set_tuning(bad, 2)
        ctx.set_tuning(-216, 2)  # three-step strips of 16 rows
        assert ctx.get_tuning() == (-216, 2)
        for bad in (-208, -200, -264):  # three-step strips are 16 or 32 rows
            with pytest.raises(lb.LbmLetkfError):
                ctx.set_tuning(bad, 2)
