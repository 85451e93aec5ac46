"""Pins of the FP64 oracle's LETKF half (no GPU): Gaspari-Cohn values, local
observation sets against brute-force enumeration, the Jacobi eigensolver against
LAPACK (numpy.linalg.eigh), the transform against matrix-function identities and
the analysis against the Kalman-gain form of the same update."""
import math

import numpy as np
import pytest

from conftest import small_cfg


def test_gaspari_cohn_published_values(orc):
    # Gaspari & Cohn (1999) eq. 4.10: 1 at 0, 5/24 at c, 0 at 2c, C^1 at c and 2c
    assert orc.gaspari_cohn(0.0) == 1.0
    assert orc.gaspari_cohn(1.0) == pytest.approx(5 / 24, rel=1e-14)
    assert orc.gaspari_cohn(2.0) == pytest.approx(0.0, abs=1e-14)
    assert orc.gaspari_cohn(2.5) == 0.0
    h = 1e-6
    for x in (1.0, 2.0 - 2 * h):
        left = (orc.gaspari_cohn(x - h) - orc.gaspari_cohn(x - 2 * h)) / h
        right = (orc.gaspari_cohn(x + 2 * h) - orc.gaspari_cohn(x + h)) / h
        assert left == pytest.approx(right, abs=1e-4)
    # compactly supported correlation function: its correlation matrix on any
    # point set is positive semi-definite (a sign slip in a term breaks this)
    x = np.linspace(0, 6, 121)
    C = np.array([[orc.gaspari_cohn(abs(a - b)) for b in x] for a in x])
    assert np.linalg.eigvalsh(C).min() > -1e-10


def _brute_window_sites(cfg, i, j):
    """Sites of the reference window (src/letkf.cu:45-55,164-165) found by scanning
    every lattice site and keeping those inside the window's box."""
    p = cfg.obs_stride
    if cfg.loc_rloc_len == 0 and p == 1:
        lo, hi, cut = 0, 0, 1e-6  # pointwise: the gridpoint's own site only
    elif cfg.loc_rloc_len == 0:
        lo, hi, cut = -p, 2 * p, p - 1 + 1e-6
    else:
        L = cfg.loc_rloc_len
        lo, hi, cut = -2 * L * p, 2 * L * p, L * p + 1e-6
    bi, bj = (i // p) * p, (j // p) * p
    out = {}
    for dj in range(lo, hi + 1, p):
        for di in range(lo, hi + 1, p):
            io, jo = bi + di, bj + dj
            s = (io % cfg.nx) // p + ((jo % cfg.ny) // p) * (cfg.nx // p)
            d = math.hypot(i - io, j - jo) / cut
            out[(di, dj)] = (s, d)
    return out


@pytest.mark.parametrize("p,rloc", [(1, 0), (2, 0), (4, 0), (8, 0), (4, 1)])
def test_gc_window_sets_match_brute_force(lb, orc, p, rloc):
    cfg = small_cfg(lb, "ref", n=32, obs_stride=p, loc_rloc_len=rloc)
    for g in range(0, 32 * 32, 7):
        i, j = g % 32, g // 32
        sites, w = orc.local_obs(cfg, g)
        bf = _brute_window_sites(cfg, i, j)
        assert len(sites) == len(bf)
        # gather order of the reference: r = ir + jr*n_obs_x (ir fastest)
        keys = sorted(bf, key=lambda k: (k[1], k[0]))
        assert [bf[k][0] for k in keys] == list(sites)
        for k, wi in zip(keys, w):
            want = 1.0 if p == 1 else orc.gaspari_cohn(bf[k][1])
            assert wi == pytest.approx(want, abs=1e-15)


@pytest.mark.parametrize("n,p,L,R", [(32, 4, 3.0, 7.5), (64, 8, 12.0, 20.0), (48, 3, 2.0, 5.0)])
def test_gauss_sets_match_brute_force(lb, orc, n, p, L, R):
    cfg = small_cfg(lb, "baseline", n=n, obs_stride=p, loc_scale=L, loc_cutoff=R)
    ns = n // p
    sx, sy = np.meshgrid(np.arange(ns) * p, np.arange(ns) * p)
    sx, sy = sx.ravel(), sy.ravel()
    for g in range(0, n * n, 37):
        i, j = g % n, g // n
        dx = (i - sx + n // 2) % n - n // 2  # minimum-image distance
        dy = (j - sy + n // 2) % n - n // 2
        d2 = dx * dx + dy * dy
        keep = np.nonzero(d2 <= R * R)[0]
        sites, w = orc.local_obs(cfg, g)
        assert sorted(sites.tolist()) == sorted(keep.tolist())
        lookup = dict(zip(keep.tolist(), np.exp(-d2[keep] / (2 * L * L))))
        for s, wi in zip(sites, w):
            assert wi == pytest.approx(lookup[s], rel=1e-14)


def test_baseline_local_counts(lb, orc):
    # 1,024 (u,v) sites every 8th node, Gaussian cutoff 43.8 -> 88..97 sites per gridpoint
    cfg = lb.preset("baseline")
    counts = [len(orc.local_obs(cfg, g)[0]) for g in range(0, 256 * 256, 997)]
    assert 85 <= min(counts) and max(counts) <= 100
    assert orc.obs_count(cfg) == (1024, 2048)


def _rand_spd(rng, n, cond=1e3):
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.exp(rng.uniform(0, math.log(cond), n))
    return (Qm * d) @ Qm.T


@pytest.mark.parametrize("n", [2, 3, 5, 16, 33, 64])
def test_jacobi_matches_lapack(orc, n):
    rng = np.random.default_rng(n)
    A = _rand_spd(rng, n)
    d, V, sweeps = orc.jacobi_eig(A, tol=1e-14)
    assert sweeps < 30
    np.testing.assert_allclose(np.sort(d), np.linalg.eigvalsh(A), rtol=1e-11)
    np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)
    np.testing.assert_allclose((V * d) @ V.T, A, rtol=0, atol=1e-11 * np.abs(A).max())


def test_jacobi_repeated_eigenvalues(orc):
    # (M-1)/beta I + rank-deficient term: clustered spectrum as in the LETKF
    rng = np.random.default_rng(5)
    Y = rng.standard_normal((6, 20))
    A = 19 * np.eye(20) + Y.T @ Y
    d, V, _ = orc.jacobi_eig(A, tol=1e-14)
    np.testing.assert_allclose(np.sort(d), np.linalg.eigvalsh(A), rtol=1e-12)
    np.testing.assert_allclose((V * d) @ V.T, A, atol=1e-10)


def _ensemble_problem(rng, M, p, spread=1.0):
    Y = spread * rng.standard_normal((p, M))
    dY = Y - Y.mean(1, keepdims=True)
    rinv = rng.uniform(0.5, 2.0, p)
    dyo = rng.standard_normal(p)
    return dY, rinv, dyo


@pytest.mark.parametrize("M,p,beta", [(4, 3, 1.0), (16, 40, 1.1), (64, 190, 1.1), (8, 0, 1.0)])
def test_transform_against_lapack_matrix_functions(orc, M, p, beta):
    rng = np.random.default_rng(M + p)
    dY, rinv, dyo = _ensemble_problem(rng, M, max(p, 1))
    if p == 0:
        rinv[:] = 0.0
    W, w, ev = orc.letkf_transform(dY, rinv, dyo, beta)
    C = (M - 1) / beta * np.eye(M) + dY.T @ (rinv[:, None] * dY)
    lam, U = np.linalg.eigh(C)
    W_ref = math.sqrt(M - 1) * (U / np.sqrt(lam)) @ U.T
    w_ref = np.linalg.solve(C, dY.T @ (rinv * dyo))
    np.testing.assert_allclose(W, W_ref, rtol=0, atol=1e-11 * np.abs(W_ref).max())
    np.testing.assert_allclose(w, w_ref, rtol=0, atol=1e-11 * max(np.abs(w_ref).max(), 1e-300) + 1e-300)
    np.testing.assert_allclose(ev, lam, rtol=1e-11)
    # identities: W symmetric, W^2 = (M-1) C^-1, W 1 = sqrt(beta) 1, 1^T w = 0
    np.testing.assert_allclose(W, W.T, atol=1e-12)
    np.testing.assert_allclose(W @ W, (M - 1) * np.linalg.inv(C), atol=1e-10)
    np.testing.assert_allclose(W @ np.ones(M), math.sqrt(beta) * np.ones(M), atol=1e-10)
    assert abs(w.sum()) < 1e-10 * max(1.0, np.abs(w).max())
    if p == 0:
        np.testing.assert_allclose(W, math.sqrt(beta) * np.eye(M), atol=1e-13)


def test_transform_tiny_hand_case(orc):
    # M=2, one observation: C = 1/beta + r*(a^2+a^2) on the perturbation direction; closed form
    a, r, beta, dyo = 0.3, 4.0, 1.0, 0.5
    dY = np.array([[a, -a]])
    W, w, ev = orc.letkf_transform(dY, np.array([r]), np.array([dyo]), beta)
    # eigenvector (1,-1)/sqrt2 has eigenvalue 1 + 2 r a^2, (1,1)/sqrt2 has 1
    lam = 1 + 2 * r * a * a
    e = np.array([1, -1]) / math.sqrt(2)
    o = np.array([1, 1]) / math.sqrt(2)
    np.testing.assert_allclose(W, np.outer(e, e) / math.sqrt(lam) + np.outer(o, o), atol=1e-14)
    np.testing.assert_allclose(w, e * (math.sqrt(2) * a * r * dyo) / lam, atol=1e-14)
    np.testing.assert_allclose(ev, [1.0, lam], rtol=1e-14)


def _small_osse(lb, orc, n=32, M=8, **kw):
    opts = dict(obs_stride=4, loc_scale=4.0, loc_cutoff=9.0)
    opts.update(kw)
    cfg = small_cfg(lb, "baseline", n=n, **opts)
    f = np.stack([orc.feq(cfg, np.ones((n, n)), *orc.initial_velocity(cfg, 1000 + k)) for k in range(M)])
    truth = orc.feq(cfg, np.ones((n, n)), *orc.initial_velocity(cfg, 7))[None]
    yo = orc.observe(cfg, orc.obs_moments(cfg, truth)[:, 0], seed=3, counter=0, noise=True)
    return cfg, f, truth, yo


def test_analysis_matches_kalman_gain_form(lb, orc):
    cfg, f, _, yo = _small_osse(lb, orc)
    M, n = f.shape[0], cfg.nx
    fa = f.copy()
    orc.letkf_analysis(cfg, fa, yo)
    hx = orc.obs_moments(cfg, f)
    ybar = hx.mean(1)
    dYall = hx - ybar[:, None]
    nv = 2
    rinv_var = 1 / cfg.obs_sigma_u**2
    beta = cfg.inflation
    for g in (0, 5, 100, 555, n * n - 1):
        sites, wl = orc.local_obs(cfg, g)
        rows = np.array([s * nv + a for s in sites for a in range(nv)])
        rinv = np.repeat(wl, nv) * rinv_var
        dY = dYall[rows]
        X = f[:, :, g // n, g % n].T  # [9][M]
        xbar = X.mean(1, keepdims=True)
        dX = X - xbar
        # Kalman gain form: xbar_a = xbar + dX (dY^T R dY + (M-1)/beta I)^-1 dY^T R (yo - ybar)
        C = dY.T @ (rinv[:, None] * dY) + (M - 1) / beta * np.eye(M)
        xa_mean = xbar[:, 0] + dX @ np.linalg.solve(C, dY.T @ (rinv * (yo[rows] - ybar[rows])))
        Xa = fa[:, :, g // n, g % n].T
        np.testing.assert_allclose(Xa.mean(1), xa_mean, rtol=1e-12, atol=1e-15)
        # perturbations: dXa dXa^T /(M-1) = dX C^-1 dX^T (ETKF covariance identity)
        dXa = Xa - Xa.mean(1, keepdims=True)
        np.testing.assert_allclose(dXa @ dXa.T / (M - 1), dX @ np.linalg.solve(C, dX.T), rtol=1e-9,
                                   atol=1e-18)


def test_analysis_without_local_obs_is_inflation(lb, orc):
    # cutoff below the site spacing: gridpoints off the lattice see no observation,
    # so Xa = xbar + sqrt(beta) dX there
    cfg, f, _, yo = _small_osse(lb, orc, loc_cutoff=0.5)
    fa = f.copy()
    orc.letkf_analysis(cfg, fa, yo)
    n = cfg.nx
    g = 1 + 2 * n  # (1, 2): not a site
    assert len(orc.local_obs(cfg, g)[0]) == 0
    X = f[:, :, 2, 1]
    xbar = X.mean(0)
    np.testing.assert_allclose(fa[:, :, 2, 1], xbar + math.sqrt(cfg.inflation) * (X - xbar), rtol=1e-13)


def test_oracle_wrapper_keeps_converted_inputs_alive(lb, orc):
    # float32 inputs are converted to float64 temporaries inside the wrapper; the
    # temporaries must outlive the C call (a freed copy once gave wrong truths)
    cfg, f, truth, _ = _small_osse(lb, orc, n=16, M=4)
    tr, tu, tv = orc.macro(cfg, truth[0])
    a = orc.ensemble_stats(cfg, f, tr, tu, tv)
    b = orc.ensemble_stats(cfg, f, tr.astype(np.float32), tu.astype(np.float32), tv.astype(np.float32))
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-12)


def test_ensemble_stats_definition(lb, orc):
    cfg, f, truth, _ = _small_osse(lb, orc, n=16, M=5)
    rng = np.random.default_rng(9)
    for k in range(f.shape[0]):  # give the density a spread too
        r0, u0, v0 = orc.macro(cfg, f[k])
        f[k] = orc.feq(cfg, r0 + 0.01 * rng.standard_normal(r0.shape), u0, v0)
    tr, tu, tv = orc.macro(cfg, truth[0])
    out = orc.ensemble_stats(cfg, f, tr, tu, tv)
    mac = [orc.macro(cfg, f[k]) for k in range(f.shape[0])]
    U = np.stack([m[1] for m in mac])
    V = np.stack([m[2] for m in mac])
    R = np.stack([m[0] for m in mac])
    rmse = np.sqrt((((U.mean(0) - tu) ** 2) + ((V.mean(0) - tv) ** 2)).sum() / (2 * U[0].size))
    spread = np.sqrt((U.var(0) + V.var(0)).sum() / (2 * U[0].size))
    np.testing.assert_allclose(out[:2], [rmse, spread], rtol=1e-10)
    np.testing.assert_allclose(out[2:], [np.sqrt(((R.mean(0) - tr) ** 2).mean()), np.sqrt(R.var(0).mean())],
                               rtol=1e-8)


def test_dsyev_route_matches_the_jacobi_route(orc):
    """The CPU baseline routes the local eigensolves through LAPACK dsyev; on the same
    local problem it gives the W, w and eigenvalues of the oracle's own Jacobi."""
    from threadpoolctl import threadpool_limits

    rng = np.random.default_rng(11)
    for M, p in ((8, 3), (16, 40), (64, 188)):
        dY = rng.standard_normal((p, M))
        dY -= dY.mean(1, keepdims=True)
        rinv = rng.uniform(0.5, 2, p) * 1e2
        dyo = rng.standard_normal(p)
        W0, w0, e0 = orc.letkf_transform(dY, rinv, dyo, 1.1)
        try:
            with threadpool_limits(1):
                assert "dsyev" in orc.set_dsyev(True)
                W1, w1, e1 = orc.letkf_transform(dY, rinv, dyo, 1.1)
        finally:
            orc.set_dsyev(False)
        assert np.linalg.norm(W1 - W0) <= 1e-12 * np.linalg.norm(W0)
        assert np.linalg.norm(w1 - w0) <= 1e-11 * np.linalg.norm(w0)
        np.testing.assert_allclose(e1, e0, rtol=1e-12)


def test_strip_analysis_is_the_full_analysis_there(lb, orc):
    # the strip form analyses [g0, g0+n) only: equal to the full analysis on the
    # strip (and its dumped W, w, evals), the rest of f untouched
    cfg, f, _, yo = _small_osse(lb, orc)
    n, g0, nd = cfg.nx, 37, 150
    fa = f.copy()
    Wf, wf, evf = orc.letkf_analysis(cfg, fa, yo, g0, nd)
    fs = f.copy()
    Ws, ws, evs = orc.letkf_analysis_strip(cfg, fs, yo, g0, nd)
    np.testing.assert_array_equal(Ws, Wf)
    np.testing.assert_array_equal(ws, wf)
    np.testing.assert_array_equal(evs, evf)
    flat_a, flat_s, flat_b = (x.reshape(x.shape[0], 9, n * n) for x in (fa, fs, f))
    np.testing.assert_array_equal(flat_s[:, :, g0:g0 + nd], flat_a[:, :, g0:g0 + nd])
    np.testing.assert_array_equal(flat_s[:, :, :g0], flat_b[:, :, :g0])
    np.testing.assert_array_equal(flat_s[:, :, g0 + nd:], flat_b[:, :, g0 + nd:])
    with pytest.raises(ValueError):
        orc.letkf_analysis_strip(cfg, f.copy(), yo, n * n - 10, 11)
