"""Pins of the oracle's explicit observation networks (CPU): the local sets are
exactly the sites inside the cutoff (brute force) in window-scan order, and an
explicit list of the lattice nodes gives the lattice's analysis (the LETKF does
not depend on the order of the observations)."""
import numpy as np
import pytest

from conftest import small_cfg


@pytest.fixture
def sites_reset(orc):
    yield
    orc.set_obs_sites(None)


def test_local_sets_brute_force(lb, orc, sites_reset):
    cfg = small_cfg(lb, "baseline", n=64, loc_scale=5.0, loc_cutoff=12.5)
    nodes = lb.uneven_sites(64, 64, seed=4, regions=4, density=(0.01, 0.08), empty_regions=1)
    assert nodes.size > 50 and np.all(np.diff(nodes) > 0)
    orc.set_obs_sites(nodes)
    ns, no = orc.obs_count(cfg)
    assert (ns, no) == (nodes.size, 2 * nodes.size)
    si, sj = nodes % 64, nodes // 64
    for g in (0, 63, 64 * 20 + 7, 64 * 40 + 40, 64 * 64 - 1):
        i, j = g % 64, g // 64
        dx = (si - i + 32) % 64 - 32
        dy = (sj - j + 32) % 64 - 32
        inside = dx * dx + dy * dy <= 12.5**2
        s, w = orc.local_obs(cfg, g)
        assert sorted(s) == sorted(np.nonzero(inside)[0].tolist())
        # window-scan order: (dy, dx) ascending
        key = [(dy[k], dx[k]) for k in s]
        assert key == sorted(key)
        np.testing.assert_allclose(w, np.exp(-(dx[s] ** 2 + dy[s] ** 2) / (2 * 25.0)), rtol=1e-15)
    # an empty region (2x2 regions of 32^2, the first empty): its centre sees no site
    orc.set_obs_sites(lb.uneven_sites(64, 64, seed=4, regions=2, density=(0.01, 0.08), empty_regions=1))
    assert orc.local_obs(cfg, 64 * 16 + 16)[0].size == 0
    assert orc.local_obs(cfg, 64 * 48 + 48)[0].size > 0


def test_explicit_lattice_equals_lattice(lb, orc, sites_reset):
    cfg = small_cfg(lb, "baseline", n=32, obs_stride=4, loc_scale=3.0, loc_cutoff=8.0)
    M = 6
    rng = np.random.default_rng(2)
    f = np.stack([orc.feq(cfg, np.ones((32, 32)), *orc.initial_velocity(cfg, 20 + k)) for k in range(M)])
    f += 1e-4 * rng.standard_normal(f.shape)
    hx = orc.obs_moments(cfg, f[:1])
    yo = hx[:, 0] + 0.003 * rng.standard_normal(hx.shape[0])
    fa_lat = f.copy()
    W1, w1, e1 = orc.letkf_analysis(cfg, fa_lat, yo, 100, 5)
    lattice = np.array([(sx * 4) + 32 * (sy * 4) for sy in range(8) for sx in range(8)], np.int32)
    orc.set_obs_sites(lattice)
    fa_exp = f.copy()
    W2, w2, e2 = orc.letkf_analysis(cfg, fa_exp, yo, 100, 5)
    np.testing.assert_allclose(fa_exp, fa_lat, rtol=0, atol=1e-13)
    np.testing.assert_allclose(W2, W1, atol=1e-11)
    np.testing.assert_allclose(e2, e1, rtol=1e-12)
