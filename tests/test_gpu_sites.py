"""GPU parity of explicit (uneven) observation networks against the FP64 oracle:
local sets bit-exact, and the analysis (plain and mixed) including gridpoints
without any local observation, whose analysis is xbar + sqrt(beta) dX."""
import math

import numpy as np
import pytest

from conftest import small_cfg
from test_gpu_letkf import TOLS, _osse, _rel

pytestmark = pytest.mark.gpu


@pytest.fixture
def sites_reset(orc):
    yield
    orc.set_obs_sites(None)


def test_uneven_sets_bit_exact(lb, orc, sites_reset):
    for n, seed, cut in ((64, 3, 12.5), (256, 8, 43.8)):
        cfg = small_cfg(lb, "baseline", n=n, loc_scale=cut / 3.65, loc_cutoff=cut)
        nodes = lb.uneven_sites(n, n, seed=seed, regions=4, empty_regions=1)
        print(f"case: uneven sets {n}x{n}, {nodes.size} sites", flush=True)
        orc.set_obs_sites(nodes)
        with lb.Context(cfg, 2) as ctx:
            ctx.set_obs_sites(nodes)
            assert ctx.obs_count() == (nodes.size, 2 * nodes.size)
            counts = []
            for g in list(range(0, n * n, 97 if n == 64 else 1031)) + [n * n - 1]:
                s_gpu, w_gpu = ctx.local_obs(g)
                s_orc, w_orc = orc.local_obs(cfg, g)
                np.testing.assert_array_equal(s_gpu, s_orc)
                np.testing.assert_allclose(w_gpu, w_orc.astype(np.float32), rtol=2e-7, atol=1e-30)
                counts.append(len(s_gpu))
            print(f"  local site counts {min(counts)}..{max(counts)}", flush=True)


@pytest.mark.parametrize("form", ["plain", "mixed3"])
def test_uneven_analysis_matches_oracle(lb, orc, form, sites_reset):
    n, M = 64, 8
    variant = lb.VARIANT_PLAIN if form == "plain" else lb.VARIANT_MIXED
    cfg = small_cfg(lb, "baseline", n=n, loc_scale=4.0, loc_cutoff=10.0, variant=variant, gram_passes=3)
    nodes = lb.uneven_sites(n, n, seed=5, regions=2, density=(0.01, 0.06), empty_regions=1)
    orc.set_obs_sites(nodes)
    tol_w, tol_inc, tol_wv = TOLS[form]
    print(f"case: uneven analysis {form} {n}x{n} M={M}, {nodes.size} sites", flush=True)
    ens, nat = _osse(lb, cfg, M)
    with ens, nat:
        ens.set_obs_sites(nodes)
        nat.set_obs_sites(nodes)
        yo = nat.observe(0, seed=4, counter=2, noise=True)
        fb = ens.get_all_states().astype(np.float64)
        # dump a row through the empty region's centre and into observed regions
        g0, nd = n * 16, 64
        ens.set_dump(g0, nd)
        ens.analysis(yo)
        fa = ens.get_all_states().astype(np.float64)
        W, w, ev = ens.get_dump()
    fo = fb.copy()
    Wo, wo, evo = orc.letkf_analysis(cfg, fo, yo.astype(np.float64), g0, nd)
    inc = _rel(fa - fb, fo - fb)
    errW = max(_rel(W[i], Wo[i]) for i in range(nd))
    print(f"  W rel {errW:.2e}  increment rel {inc:.2e}", flush=True)
    assert errW < tol_w and inc < tol_inc
    beta = cfg.inflation
    empty = [i for i in range(nd) if orc.local_obs(cfg, g0 + i)[0].size == 0]
    assert len(empty) > 5
    for i in empty:
        # C = (M-1)/beta I is formed in FP32, so W = sqrt(beta) I to a few FP32 ulps
        # (MIXED: V is the warm-start basis made orthogonal to FP32 rounding)
        np.testing.assert_allclose(W[i], math.sqrt(beta) * np.eye(M), atol=4e-7)
        np.testing.assert_allclose(w[i], 0.0, atol=1e-15)
        g = g0 + i
        X = fb[:, :, g // n, g % n]
        xb = X.mean(0)
        np.testing.assert_allclose(fa[:, :, g // n, g % n], xb + math.sqrt(beta) * (X - xb), rtol=2e-6)


def test_site_list_errors(lb):
    cfg = small_cfg(lb, "baseline", n=64, loc_scale=4.0, loc_cutoff=10.0)
    with lb.Context(cfg, 4) as ctx:
        for bad in ([0, 5, 5], [-1], [64 * 64]):
            with pytest.raises(lb.LbmLetkfError) as e:
                ctx.set_obs_sites(bad)
            assert e.value.status == lb.ERR_ARG
        ctx.set_obs_sites([])  # no sites at all: every analysis is the inflated prior
        assert ctx.obs_count() == (0, 0)
        fb = ctx.get_all_states().astype(np.float64)
        ctx.analysis(np.zeros(1, np.float32))
        fa = ctx.get_all_states().astype(np.float64)
        xb = fb.mean(0, keepdims=True)
        np.testing.assert_allclose(fa, xb + math.sqrt(cfg.inflation) * (fb - xb), rtol=2e-6, atol=1e-9)
    gc = small_cfg(lb, "ref", n=64)  # GC-window localization: lattice only
    with lb.Context(gc, 2) as ctx:
        with pytest.raises(lb.LbmLetkfError):
            ctx.set_obs_sites([1, 2, 3])
