"""Groups: several OSSEs of different grids and ensemble sizes in shared launches.

The group kernels run the same per-tile and per-gridpoint code as a single
context's, so a group step, analysis and graph-replayed cycle must reproduce the
contexts' own results bit for bit; explicit uneven networks (some gridpoints with
no observation) take part like lattices, and so do a context with the learned
closure and a forced one (the ref preset's Watanabe-97 forcing and friction)."""
import numpy as np
import pytest

from conftest import small_cfg

pytestmark = pytest.mark.gpu

# (nx, ny, M): analysis widths 16, 16, 32, 64 and unequal grids
SHAPES = [(64, 64, 8), (128, 64, 16), (64, 128, 32), (128, 128, 64)]


def _cfg(lb, nx, ny, uneven, forced=False):
    if forced:
        # the ref preset: Watanabe-97 forcing and friction (one-step launches with a
        # force field per step on the context's side stream), GC-window localization
        cfg = small_cfg(lb, "ref", n=64, obs_stride=4, variant=lb.VARIANT_MIXED)
    else:
        cfg = small_cfg(lb, "baseline", n=64, loc_scale=4.0, loc_cutoff=10.0, obs_stride=8)
    cfg.nx, cfg.ny = nx, ny
    cfg.da_interval = 7
    return cfg


def _make(lb, uneven, mlp=False, forced=False):
    out = []
    for k, (nx, ny, M) in enumerate(SHAPES):
        cfg = _cfg(lb, nx, ny, uneven, forced=forced and k == 2)
        if mlp and k == 1:  # one context with the learned closure
            cfg.closure_kind = lb.CLOSURE_MLP
        ens = lb.Context(cfg, M)
        if mlp and k == 1:
            ens.set_closure(lb.closure_init_params(7))
        nat = lb.Context(cfg, 1)
        for m in range(M):
            ens.init_member(m, 100 * k + m)
        nat.init_member(0, 99 + k)
        if uneven:
            nodes = lb.uneven_sites(nx, ny, seed=k, regions=2, density=(0.01, 0.05), empty_regions=1)
            ens.set_obs_sites(nodes)
            nat.set_obs_sites(nodes)
        out.append((ens, nat))
    return out


def _obs(pairs, counter):
    return [nat.observe(0, seed=3, counter=counter, noise=True) for _, nat in pairs]


@pytest.mark.parametrize("mode", ["lattice", "uneven", "mlp", "forced"])
def test_group_matches_contexts(lb, hipbuf, mode):
    uneven, mlp, forced = mode == "uneven", mode == "mlp", mode == "forced"
    print(f"case: group of {len(SHAPES)} contexts, {mode}", flush=True)
    # a: one context at a time; b: the group
    a, b = _make(lb, uneven, mlp, forced), _make(lb, uneven, mlp, forced)
    try:
        # forecast
        for ens, _ in a:
            ens.forecast(5)
        with lb.Group([ens for ens, _ in b]) as grp:
            grp.forecast(5)
            grp.synchronize()
            for (ea, _), (eb, _) in zip(a, b):
                np.testing.assert_array_equal(eb.get_all_states(), ea.get_all_states())
                assert eb.get_step() == ea.get_step() == 5
            # analysis
            yo = _obs(a, 1)
            yo_dev = [hipbuf.from_host(y) for y in yo]
            for (ea, _), y in zip(a, yo):
                ea.analysis(y)
            grp.analysis_device(yo_dev)
            grp.synchronize()
            for (ea, _), (eb, _) in zip(a, b):
                np.testing.assert_array_equal(eb.get_all_states(), ea.get_all_states())
            # cycles (graph) against analysis + forecast, both parities of the buffers
            for counter in (2, 3, 4):
                yo = _obs(a, counter)
                yo_dev = [hipbuf.from_host(y) for y in yo]
                for (ea, _), y in zip(a, yo):
                    ea.analysis(y)
                    ea.forecast(ea.cfg.da_interval)
                grp.cycle_device(yo_dev)
                grp.synchronize()
                for (ea, _), (eb, _) in zip(a, b):
                    np.testing.assert_array_equal(eb.get_all_states(), ea.get_all_states())
                    assert eb.get_step() == ea.get_step()
            print("  group == contexts after forecast, analysis and 3 graph cycles", flush=True)
    finally:
        for ens, nat in a + b:
            ens.close()
            nat.close()


def test_group_rejects(lb):
    cfg = small_cfg(lb, "baseline", n=64, loc_scale=4.0, loc_cutoff=10.0)
    with lb.Context(cfg, 4) as c1:
        with pytest.raises(lb.LbmLetkfError):
            lb.Group([c1, c1])
        p = cfg.copy()
        p.variant = lb.VARIANT_PLAIN
        with lb.Context(p, 4) as c2:
            with pytest.raises(lb.LbmLetkfError):
                lb.Group([c1, c2])
        q = cfg.copy()
        q.closure_kind = lb.CLOSURE_MLP
        with lb.Context(q, 4) as c3:
            # a context with the learned closure joins, but steps only with its network
            with lb.Group([c1, c3]) as grp:
                s1, s3 = c1.get_all_states(), c3.get_all_states()
                with pytest.raises(lb.LbmLetkfError, match="set_closure"):
                    grp.forecast(1)
                np.testing.assert_array_equal(c1.get_all_states(), s1)  # nothing launched
                np.testing.assert_array_equal(c3.get_all_states(), s3)
        with pytest.raises(lb.LbmLetkfError):
            lb.Group([])
