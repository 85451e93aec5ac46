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
                fb = ens.get_all_states().astype(np.float64)
                print(f"cycle {c}: forecast rel L2 gpu vs oracle {np.linalg.norm(fb - f_orc) / np.linalg.norm(f_orc):.2e}")
            yo = nat.observe(0, seed=99, counter=c, noise=True)
            ens.analysis(yo)
            tr, tu, tv = nat.get_macro(0)
            rm_gpu.append(ens.ensemble_stats(tr, tu, tv))
            if with_oracle:
                orc.letkf_analysis(cfg, f_orc, yo.astype(np.float64))
                rm_orc.append(orc.ensemble_stats(cfg, f_orc, tr, tu, tv))
                fa = ens.get_all_states().astype(np.float64)
                print(f"cycle {c}: analysis rel L2 gpu vs oracle {np.linalg.norm(fa - f_orc) / np.linalg.norm(f_orc):.2e}"
                      f"  stats of gpu state by oracle {orc.ensemble_stats(cfg, fa, tr, tu, tv)[:2]}"
                      f"  numpy {_numpy_stats(cfg, fa, tr, tu, tv)[:2]}  gpu again {ens.ensemble_stats(tr, tu, tv)[:2]}"
                      f"  truth dtype {tu.dtype} shape {tu.shape}")
    return np.array(rm_gpu), (np.array(rm_orc) if with_oracle else None)


def _numpy_stats(cfg, f, tr, tu, tv):
    """ensemble_stat/main.cpp:108-177 written with numpy (biased spread)."""
    c = cfg.c
    ex = np.array([q % 3 - 1 for q in range(9)], float)
    ey = np.array([q // 3 - 1 for q in range(9)], float)
    rho = f.sum(1)
    u = c * np.tensordot(f, ex, axes=([1], [0])) / rho
    v = c * np.tensordot(f, ey, axes=([1], [0])) / rho
    n2 = 2 * tu.size
    rmse = np.sqrt((((u.mean(0) - tu) ** 2) + ((v.mean(0) - tv) ** 2)).sum() / n2)
    spread = np.sqrt((u.var(0) + v.var(0)).sum() / n2)
    return np.array([rmse, spread, np.sqrt(((rho.mean(0) - tr) ** 2).mean()), np.sqrt(rho.var(0).mean())])


@pytest.mark.parametrize("truth", ["zero", "random", "member"])
def test_ensemble_stats_matches_numpy(lb, orc, truth):
    cfg = small_cfg(lb, "baseline", n=64)
    M = 16
    with lb.Context(cfg, M) as ens:
        for k in range(M):
            ens.init_member(k, 1000 + k)
        ens.forecast(5)
        f = ens.get_all_states().astype(np.float64)
        rng = np.random.default_rng(3)
        if truth == "zero":
            tr, tu, tv = np.ones((64, 64)), np.zeros((64, 64)), np.zeros((64, 64))
        elif truth == "random":
            tr, tu, tv = (1 + 1e-3 * rng.standard_normal((64, 64)), 0.05 * rng.standard_normal((64, 64)),
                          0.05 * rng.standard_normal((64, 64)))
        else:
            tr, tu, tv = (x.astype(np.float64) for x in ens.get_macro(3))
        got = ens.ensemble_stats(tr, tu, tv)
    want = _numpy_stats(cfg, f, tr, tu, tv)
    orc_out = orc.ensemble_stats(cfg, f, tr, tu, tv)
    print(f"truth={truth}: gpu {got}  numpy {want}  oracle {orc_out}")
    np.testing.assert_allclose(orc_out[:2], want[:2], rtol=1e-9)
    np.testing.assert_allclose(got[:2], want[:2], rtol=1e-5)
    np.testing.assert_allclose(got[2:], want[2:], rtol=1e-3, atol=1e-9)


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
    # the benchmark OSSE scaled to 128^2: (u,v) every 8th node, L = 6, cutoff 21.9
    cfg = small_cfg(lb, "baseline", n=128)
    cfg.variant = lb.VARIANT_MIXED
    M, cycles, steps = 32, 8, 100
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


def test_graph_cycle_matches_eager(lb, hipbuf):
    # MIXED cycles replayed from the captured graph (phase timing off) against the same
    # cycles launched eagerly (phase timing on): identical kernels, identical bits.
    # Replays are separated by host synchronisation and state reads; an odd interval
    # makes the two parities of the double buffer alternate, and set_dump in between
    # forces a fresh capture.  The same runs with the two-step strips (three two-step
    # launches and one single step per interval of 7) give the same bits.
    cfg = small_cfg(lb, "baseline", n=64, obs_stride=4, loc_scale=6.0, loc_cutoff=14.0, da_interval=7)
    M = 16
    nat = lb.Context(cfg, 1)
    nat.init_member(0, 7)
    obs = []
    for c in range(6):
        nat.forecast(cfg.da_interval)
        obs.append(nat.observe(0, seed=5, counter=c, noise=True))
    nat.close()
    yo_dev = [hipbuf.from_host(y) for y in obs]
    runs = []
    for shape, timed in ((1, False), (1, True), (-116, False), (-116, True)):
        print(f"case: shape {shape}, {'eager' if timed else 'graph'}", flush=True)
        ens = lb.Context(cfg, M)
        for k in range(M):
            ens.init_member(k, 1000 + k)
        ens.set_tuning(shape, 8)
        ens.profile_enable(timed)
        states, dumps = [], []
        for c in range(6):
            if c == 3:
                ens.set_dump(64 * 32 + 5, 4)
            ens.cycle_device(yo_dev[c])
            ens.synchronize()
            states.append(ens.get_all_states())
            if c >= 3:
                dumps.append(ens.get_dump())
        assert ens.get_step() == 6 * cfg.da_interval
        ens.close()
        runs.append((states, dumps))
    for r in range(1, len(runs)):
        for c in range(6):
            print(f"run {r} cycle {c}: max |diff| to run 0 {np.abs(runs[0][0][c] - runs[r][0][c]).max():.3e}")
            np.testing.assert_array_equal(runs[0][0][c], runs[r][0][c])
        for a, b in zip(runs[0][1], runs[r][1]):
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)


def test_plain_cycle_matches_analysis_then_forecast(lb, hipbuf):
    # lbmletkf_cycle on the PLAIN variant (eager launches; its per-gridpoint matrices
    # are allocated by the first analysis, whichever entry point runs it) against
    # analysis() followed by forecast(): identical kernels, identical bits
    cfg = small_cfg(lb, "baseline", n=64, obs_stride=8, loc_scale=6.0, loc_cutoff=20.0, da_interval=5,
                    variant=lb.VARIANT_PLAIN)
    M = 8
    nat = lb.Context(cfg, 1)
    nat.init_member(0, 7)
    nat.forecast(cfg.da_interval)
    yo = nat.observe(0, seed=5, counter=0, noise=True)
    nat.close()
    yo_dev = hipbuf.from_host(yo)
    out = []
    for use_cycle in (True, False):
        ens = lb.Context(cfg, M)
        for k in range(M):
            ens.init_member(k, 1000 + k)
        if use_cycle:
            ens.cycle_device(yo_dev)
        else:
            ens.analysis(yo)
            ens.forecast(cfg.da_interval)
        ens.synchronize()
        out.append(ens.get_all_states())
        ens.close()
    np.testing.assert_array_equal(out[0], out[1])


def test_ensemble_fields_match_numpy(lb):
    """lbmletkf_ensemble_fields against the members' own macroscopic fields: the
    mean of u and v, the mean and standard deviation of the vorticity (central
    differences, src/io.cu:61-86) and the spread of u and v, per node."""
    cfg = small_cfg(lb, "baseline", n=64)
    cfg.nx, cfg.ny = 128, 64
    M = 12
    print("case: ensemble fields 128x64 M=12", flush=True)
    with lb.Context(cfg, M) as ens:
        for k in range(M):
            ens.init_member(k, 500 + k)
        ens.forecast(7)
        mu, mv, wm, ws, su, sv = ens.ensemble_fields()
        uvs = [ens.get_macro(k)[1:] for k in range(M)]
    u = np.stack([a for a, _ in uvs]).astype(np.float64)
    v = np.stack([b for _, b in uvs]).astype(np.float64)
    dx = np.float32(cfg.dx)
    vor = []
    for a, b in uvs:  # float32 as the device forms it
        uy = (np.roll(a, -1, axis=0) - np.roll(a, 1, axis=0)) / (np.float32(2) * dx)
        vx = (np.roll(b, -1, axis=1) - np.roll(b, 1, axis=1)) / (np.float32(2) * dx)
        vor.append((vx - uy).astype(np.float64))
    vor = np.stack(vor)
    scale_u = np.abs(u).max()
    np.testing.assert_allclose(mu, u.mean(0), rtol=0, atol=1e-7 * scale_u)
    np.testing.assert_allclose(mv, v.mean(0), rtol=0, atol=1e-7 * scale_u)
    np.testing.assert_allclose(su, u.std(0), rtol=0, atol=1e-6 * scale_u)
    np.testing.assert_allclose(sv, v.std(0), rtol=0, atol=1e-6 * scale_u)
    scale_w = np.abs(vor).max()
    np.testing.assert_allclose(wm, vor.mean(0), rtol=0, atol=1e-6 * scale_w)
    np.testing.assert_allclose(ws, vor.std(0), rtol=0, atol=1e-5 * scale_w)
    print(f"  spread of u {su.mean():.3e}, vorticity sd {ws.mean():.3e}", flush=True)
