"""The uneven batch at its benchmark sizes: four OSSEs (128^2 M=8, 256^2 M=16,
384^2 M=32, 512^2 M=64) with observation sites at a random density of 0.5-4% per
region of a 4x4 cut, the first region without any site, in shared launches.

* The group's forecast, analysis and graph-replayed cycle equal the same
  contexts run one at a time, bit for bit (states and transform dumps).
* The transform W, the mean weights w and the analysed state along one row of
  every grid equal the FP64 oracle's (oracle/oracle.cpp, letkf_solver.h:235-349)
  to the mixed variant's tolerance.  The row crosses the empty region, where the
  larger grids have gridpoints without any local observation, and regions dense
  enough for ~400 scalar observations in one local problem (the site seeds are
  chosen for that, and the test checks it).
"""
import numpy as np
import pytest

from test_gpu_letkf import TOLS, _rel

pytestmark = pytest.mark.gpu

# (n, M, site seed): the seeds give the dumped row y = n/8 a dense region
GRIDS = [(128, 8, 23), (256, 16, 0), (384, 32, 0), (512, 64, 1)]
SPIN = 20        # forecast steps before the analysis (six three-step launches + two)
DENSE = 350      # scalar observations that some local problem on a row must reach


def _build(lb, n, M, seed):
    cfg = lb.preset("baseline")
    cfg.nx = cfg.ny = n
    sites = lb.uneven_sites(n, n, seed=seed, regions=4, density=(0.005, 0.04), empty_regions=1)
    ens = lb.Context(cfg, M)
    for m in range(M):
        ens.init_member(m, 1000 + m)
    ens.set_obs_sites(sites)
    return cfg, sites, ens


@pytest.fixture
def sites_reset(orc):
    yield
    orc.set_obs_sites(None)


def test_uneven_batch_group_matches_contexts_and_oracle(lb, orc, hipbuf, sites_reset):
    tol_w, tol_inc, tol_wv = TOLS["mixed3"]
    a, b, nats, info = [], [], [], []
    try:
        for n, M, seed in GRIDS:
            print(f"case: uneven batch {n}x{n} M={M}, site seed {seed}", flush=True)
            cfg, sites, ea = _build(lb, n, M, seed)
            _, _, eb = _build(lb, n, M, seed)
            nat = lb.Context(cfg, 1)
            nat.init_member(0, 7 + n)
            nat.set_obs_sites(sites)
            a.append(ea)
            b.append(eb)
            nats.append(nat)
            info.append((cfg, sites))
        with lb.Group(b) as grp:
            for ea, nat in zip(a, nats):
                ea.forecast(SPIN)
                nat.forecast(SPIN)
            grp.forecast(SPIN)
            grp.synchronize()
            for ea, eb in zip(a, b):
                np.testing.assert_array_equal(eb.get_all_states(), ea.get_all_states())
            print(f"  group == contexts after {SPIN} steps", flush=True)

            # one analysis with a dumped row per grid
            rows = []
            for (n, M, _), ea, eb in zip(GRIDS, a, b):
                g0 = (n // 8) * n
                ea.set_dump(g0, n)
                eb.set_dump(g0, n)
                rows.append((g0, n))
            yo = [nat.observe(0, seed=5, counter=1, noise=True) for nat in nats]
            fb = [ea.get_all_states() for ea in a]
            for ea, y in zip(a, yo):
                ea.analysis(y)
            grp.analysis_device([hipbuf.from_host(y) for y in yo])
            grp.synchronize()
            dense = 0
            for (n, M, _), (cfg, sites), ea, eb, y, f0, (g0, nd) in zip(GRIDS, info, a, b, yo, fb, rows):
                fa = ea.get_all_states()
                np.testing.assert_array_equal(eb.get_all_states(), fa)
                W, w, ev = ea.get_dump()
                Wb, wb, evb = eb.get_dump()
                np.testing.assert_array_equal(Wb, W)
                np.testing.assert_array_equal(wb, w)
                # the oracle on the same row, from the same forecast
                orc.set_obs_sites(sites)
                counts = np.array([2 * len(orc.local_obs(cfg, g0 + i)[0]) for i in range(nd)])
                fo = f0.astype(np.float64)
                Wo, wo, evo = orc.letkf_analysis_strip(cfg, fo, y.astype(np.float64), g0, nd)
                errW = max(_rel(W[i], Wo[i]) for i in range(nd))
                errw = max(np.abs(w[i] - wo[i]).max() / max(np.abs(wo[i]).max(), 1e-30) for i in range(nd))
                errev = max(np.abs(ev[i] - evo[i]).max() / np.abs(evo[i]).max() for i in range(nd))
                row = (slice(None), slice(None), g0 // n)
                inc = _rel(fa[row].astype(np.float64) - f0[row], fo[row] - f0[row])
                print(f"  {n}x{n} M={M}: {sites.size} sites; row {g0 // n}: local scalar obs "
                      f"{counts.min()}..{counts.max()}; W rel {errW:.2e}  w rel {errw:.2e}  "
                      f"evals rel {errev:.2e}  increment rel {inc:.2e}", flush=True)
                assert errW < tol_w and errev < tol_w
                assert errw < tol_wv
                assert inc < tol_inc
                if n >= 384:  # the empty region is wider than the cutoff diameter
                    assert counts.min() == 0
                dense = max(dense, counts.max())
            assert dense >= DENSE

            # three cycles as one graph each: analysis, then da_interval steps
            for counter in (2, 3, 4):
                yo = [nat.observe(0, seed=5, counter=counter, noise=True) for nat in nats]
                for ea, y in zip(a, yo):
                    ea.analysis(y)
                    ea.forecast(ea.cfg.da_interval)
                grp.cycle_device([hipbuf.from_host(y) for y in yo])
                grp.synchronize()
            for ea, eb in zip(a, b):
                np.testing.assert_array_equal(eb.get_all_states(), ea.get_all_states())
                assert eb.get_step() == ea.get_step()
            print("  group == contexts after 3 graph cycles of 100 steps", flush=True)
    finally:
        for c in a + b + nats:
            c.close()
