"""GPU parity of the LETKF analysis against the FP64 oracle (run on an MI355X).

* local observation sets: bit-exact site indices, weights to FP32 rounding;
* H(x), observations: to FP32 rounding;
* W, w, eigenvalues of the dumped gridpoints: relative Frobenius error of W below
  TOL_W (FP32 Gram matrix, FP64 Jacobi and transform);
* the analysed state: relative L2 error of the analysis increment below TOL_INC.
"""
import math

import numpy as np
import pytest

from conftest import f32_velocity_atol, small_cfg

pytestmark = pytest.mark.gpu
TOL_W = 1e-5
TOL_INC = 1e-4
# (W and eigenvalue relative error, analysis-increment relative error, w relative error)
# plain: FP32 Gram + FP64 Jacobi.  mixed3: split-bf16 MFMA Gram + FP32/FP64 Jacobi.
# mixed1: single bf16 operands (8-bit mantissa) in the Gram matrix; w, a solve
# with C, carries cond(C) times that rounding (3.4e-2 measured on ref_pointwise_M4).
# mixed3f64: the same with one FP64 sweep after the FP32 phase.
TOLS = {"plain": (TOL_W, TOL_INC, 1e-4), "mixed3": (1e-3, 1e-3, 1e-3), "mixed3f64": (1e-3, 1e-3, 1e-3),
        "mixed1": (3e-2, 3e-2, 6e-2)}


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _osse(lb, cfg, M, spin=20):
    ens = lb.Context(cfg, M)
    for k in range(M):
        ens.init_member(k, 1000 + k)
    nat = lb.Context(cfg, 1)
    nat.init_member(0, 7)
    if spin:
        ens.forecast(spin)
        nat.forecast(spin)
    return ens, nat


@pytest.mark.parametrize("p,rloc", [(1, 0), (2, 0), (4, 0), (8, 0), (4, 1)])
def test_gc_window_sets_bit_exact(lb, orc, p, rloc):
    cfg = small_cfg(lb, "ref", n=64, obs_stride=p, loc_rloc_len=rloc)
    with lb.Context(cfg, 2) as ctx:
        for g in range(0, 64 * 64, 13):
            s_gpu, w_gpu = ctx.local_obs(g)
            s_orc, w_orc = orc.local_obs(cfg, g)
            np.testing.assert_array_equal(s_gpu, s_orc)
            # GC(2c) is a cancellation to ~1e-16 in either evaluation order
            np.testing.assert_allclose(w_gpu, w_orc.astype(np.float32), rtol=2e-7, atol=1e-12)


def test_gauss_sets_bit_exact_full_size(lb, orc):
    cfg = lb.preset("baseline")
    with lb.Context(cfg, 2) as ctx:
        counts = []
        for g in list(range(0, 256 * 256, 331)) + [0, 255, 256 * 255, 256 * 256 - 1]:
            s_gpu, w_gpu = ctx.local_obs(g)
            s_orc, w_orc = orc.local_obs(cfg, g)
            np.testing.assert_array_equal(s_gpu, s_orc)
            np.testing.assert_allclose(w_gpu, w_orc.astype(np.float32), rtol=2e-7, atol=0)
            counts.append(len(s_gpu))
    print(f"local sites per gridpoint: {min(counts)}..{max(counts)}")
    assert 85 <= min(counts) and max(counts) <= 100


def test_observe_matches_oracle(lb, orc):
    cfg = small_cfg(lb, "ref", n=64, obs_stride=4)
    ens, nat = _osse(lb, cfg, 2, spin=5)
    with ens, nat:
        f = nat.get_all_states()
        hx = orc.obs_moments(cfg, f)[:, 0]
        # H(x) is evaluated in FP64 from the FP32 populations, as the oracle does
        np.testing.assert_allclose(nat.observe(0, 5, 1, noise=False), hx, rtol=2e-7, atol=1e-12)
        y = nat.observe(0, 5, 1, noise=True)
        np.testing.assert_allclose(y, orc.observe(cfg, hx, 5, 1, noise=True), rtol=2e-7, atol=1e-9)
        # observations taken from reference-layout fields (FP32 moments) give the same vector
        r, u, v = nat.get_macro(0)
        np.testing.assert_allclose(nat.obs_from_fields(r, u, v), nat.observe(0, 5, 1, noise=False), rtol=1e-6,
                                   atol=f32_velocity_atol(cfg))


CASES = {
    "baseline_gauss_M16": dict(preset="baseline", n=64, M=16, obs_stride=8, loc_scale=6.0, loc_cutoff=20.0),
    "baseline_gauss_M64": dict(preset="baseline", n=64, M=64, obs_stride=4, loc_scale=6.0, loc_cutoff=14.0),
    "ref_gc_p4_M8": dict(preset="ref", n=64, M=8, obs_stride=4),
    "ref_pointwise_M4": dict(preset="ref", n=64, M=4, obs_stride=1),
    "ref_rloc_M32": dict(preset="ref", n=64, M=32, obs_stride=4, loc_rloc_len=1, inflation=1.05),
    # member counts that fill the last tile of 16 only in part (N = 32, 64)
    "baseline_gauss_M20": dict(preset="baseline", n=64, M=20, obs_stride=8, loc_scale=6.0, loc_cutoff=20.0),
    "baseline_gauss_M40": dict(preset="baseline", n=64, M=40, obs_stride=4, loc_scale=6.0, loc_cutoff=14.0),
    # C scaled by 1e10 (R^-1 and (M-1)/beta alike; same eigenvectors and condition):
    # squared column norms ~1e22 take the solvers' power-of-two prescaling (above 1e18)
    "baseline_gauss_M16_scaled": dict(preset="baseline", n=64, M=16, obs_stride=8, loc_scale=6.0, loc_cutoff=20.0,
                                      obs_sigma_u=1e-7, obs_sigma_rho=1e-7, inflation=1.1e-10),
    "baseline_gauss_M64_scaled": dict(preset="baseline", n=64, M=64, obs_stride=4, loc_scale=6.0, loc_cutoff=14.0,
                                      obs_sigma_u=1e-7, obs_sigma_rho=1e-7, inflation=1.1e-10),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("form", ["plain", "mixed3", "mixed3f64", "mixed1"])
def test_analysis_matches_oracle(lb, orc, case, form):
    kw = dict(CASES[case])
    preset, n, M = kw.pop("preset"), kw.pop("n"), kw.pop("M")
    variant = lb.VARIANT_PLAIN if form == "plain" else lb.VARIANT_MIXED
    cfg = small_cfg(lb, preset, n=n, variant=variant, gram_passes=1 if form == "mixed1" else 3,
                    jacobi_fp64_sweeps=1 if form == "mixed3f64" else 0, **kw)
    tol_w, tol_inc, tol_wv = TOLS[form]
    print(f"case: {case} {form}", flush=True)
    ens, nat = _osse(lb, cfg, M)
    with ens, nat:
        yo = nat.observe(0, seed=11, counter=3, noise=True)
        fb = ens.get_all_states().astype(np.float64)
        g0, nd = n * n // 2 + 3, 24
        ens.set_dump(g0, nd)
        ens.analysis(yo)
        fa = ens.get_all_states().astype(np.float64)
        W, w, ev = ens.get_dump()
        Cd, Vd, evd = ens.get_dump_evd()
    fo = fb.copy()
    Wo, wo, evo = orc.letkf_analysis(cfg, fo, yo.astype(np.float64), g0, nd)
    # the eigensolve side of the dump: C against the oracle's C of the same local
    # problem, and V diag(evals) V^T = C with V orthonormal (solver order, unsorted)
    hx = orc.obs_moments(cfg, fb)
    dy_all = hx - hx.mean(axis=1, keepdims=True)
    nv = len(dy_all) // orc.obs_count(cfg)[0]
    rinv_var = {0: 1 / cfg.obs_sigma_u ** 2, 1: 1 / cfg.obs_sigma_u ** 2, 2: 1 / cfg.obs_sigma_rho ** 2}
    var_ids = [v for v, bit in ((0, 1), (1, 2), (2, 4)) if cfg.obs_vars & bit]
    # V's orthogonality: FP64 rotations (plain) keep it to rounding; the FP32 phase
    # of the mixed solver accumulates ~sweeps * M rotations of FP32 rounding each
    # (measured <= 2.3e-6 over these cases, the near-singular rloc problem at M = 32
    # the largest).  The bound also catches rotations that the two lanes holding the
    # halves of a column apply differently (1.8e-4 on rloc when their inner products
    # differed in the last bit).
    tol_orth = 1e-12 if form == "plain" else 2e-5
    errC = errEVD = errOrth = 0.0
    for i in range(nd):
        sites, wl = orc.local_obs(cfg, g0 + i)
        rows = [s_ * nv + a for s_ in sites for a in range(nv)]
        rinv = np.array([wl[e] * rinv_var[var_ids[a]] for e in range(len(sites)) for a in range(nv)])
        dY = dy_all[rows]
        C_o = (M - 1) / cfg.inflation * np.eye(M) + dY.T @ (rinv[:, None] * dY)
        errC = max(errC, _rel(Cd[i], C_o))
        errEVD = max(errEVD, _rel(Vd[i] @ np.diag(evd[i]) @ Vd[i].T, Cd[i]))
        errOrth = max(errOrth, float(np.abs(Vd[i].T @ Vd[i] - np.eye(M)).max()))
        np.testing.assert_allclose(np.sort(evd[i]), ev[i], rtol=1e-12)
    print(f"  C rel {errC:.2e}  V diag(evals) V^T vs C rel {errEVD:.2e}  max|V^T V - I| {errOrth:.2e}",
          flush=True)
    assert errC < tol_w
    assert errOrth < tol_orth
    assert errEVD < 1e-5
    errW = max(_rel(W[i], Wo[i]) for i in range(nd))
    errw = max(np.abs(w[i] - wo[i]).max() / max(np.abs(wo[i]).max(), 1e-30) for i in range(nd))
    errev = max(np.abs(ev[i] - evo[i]).max() / np.abs(evo[i]).max() for i in range(nd))
    inc = _rel(fa - fb, fo - fb)
    # the analysed state is stored in FP32: allow 2 ulp of f on top of TOL_INC of the increment
    ulp = np.spacing(np.abs(fo).astype(np.float32)).astype(np.float64)
    excess = np.abs(fa - fo) - (2 * ulp + tol_inc * np.abs(fo - fb).max())
    inc_el = np.max(np.maximum(np.abs(fa - fo) - 2 * ulp, 0)) / np.abs(fo - fb).max()
    print(f"  W rel {errW:.2e}  w rel {errw:.2e}  evals rel {errev:.2e}  increment rel {inc:.2e}  "
          f"elementwise (beyond 2 ulp) / max|inc| {inc_el:.2e}  max |fa-fo|/ulp {np.max(np.abs(fa - fo) / ulp):.2f}",
          flush=True)
    assert errW < tol_w
    assert errev < tol_w
    assert errw < tol_wv
    assert excess.max() <= 0
    # identities of the transform (hold for any dY with zero row means)
    beta = cfg.inflation
    for i in range(nd):
        np.testing.assert_allclose(W[i], W[i].T, atol=1e-12 * np.abs(W[i]).max())
        np.testing.assert_allclose(W[i].sum(1), math.sqrt(beta), rtol=max(1e-4, tol_w))
        assert ev[i].min() >= (M - 1) / beta * (1 - max(1e-5, tol_w))


@pytest.mark.parametrize("form", ["plain", "mixed3"])
def test_full_size_transform_parity(lb, orc, form):
    # 256^2 x 64, the benchmark shape: compare W of a strip of gridpoints with the
    # oracle's transform of the same local problems
    cfg = lb.preset("baseline")
    cfg.variant = lb.VARIANT_PLAIN if form == "plain" else lb.VARIANT_MIXED
    cfg.gram_passes = 3
    tol_w = TOLS[form][0]
    M = 64
    ens, nat = _osse(lb, cfg, M, spin=10)
    with ens, nat:
        yo = nat.observe(0, seed=1, counter=1, noise=True)
        f = ens.get_all_states().astype(np.float64)
        g0, nd = 256 * 100 + 40, 16
        ens.set_dump(g0, nd)
        ens.analysis(yo)
        W, w, ev = ens.get_dump()
        fa = ens.get_all_states()
    assert np.isfinite(fa).all()
    hx = orc.obs_moments(cfg, f)
    dYall = hx - hx.mean(1, keepdims=True)
    dyoall = yo - hx.mean(1)
    errs = []
    for i in range(nd):
        sites, wl = orc.local_obs(cfg, g0 + i)
        rows = np.array([s * 2 + a for s in sites for a in range(2)])
        rinv = np.repeat(wl, 2) / cfg.obs_sigma_u**2
        Wo, wo, evo = orc.letkf_transform(dYall[rows], rinv, dyoall[rows], cfg.inflation)
        errs.append(_rel(W[i], Wo))
        np.testing.assert_allclose(ev[i], evo, rtol=tol_w)
    print(f"full size {form}: W rel max {max(errs):.2e}")
    assert max(errs) < tol_w


def test_no_local_obs_gives_inflated_prior(lb):
    cfg = small_cfg(lb, "baseline", n=64, obs_stride=8, loc_cutoff=0.5, loc_scale=1.0, inflation=1.2,
                    variant=lb.VARIANT_PLAIN)
    ens, nat = _osse(lb, cfg, 8, spin=3)
    with ens, nat:
        yo = nat.observe(0, 1, 1)
        fb = ens.get_all_states().astype(np.float64)
        ens.analysis(yo)
        fa = ens.get_all_states().astype(np.float64)
    X = fb[:, :, 3, 5]  # node (5,3) is not a site
    xbar = X.mean(0)
    np.testing.assert_allclose(fa[:, :, 3, 5], xbar + math.sqrt(1.2) * (X - xbar), rtol=1e-6)


def test_nan_observation_is_reported(lb):
    cfg = small_cfg(lb, "baseline", n=64, obs_stride=8, loc_cutoff=20.0, loc_scale=6.0)
    ens, nat = _osse(lb, cfg, 4, spin=0)
    with ens, nat:
        yo = nat.observe(0, 1, 1)
        yo[5] = np.nan
        with pytest.raises(lb.LbmLetkfError) as e:
            ens.analysis(yo)
        assert e.value.status == lb.ERR_NUMERIC


def test_single_member_analysis_is_rejected(lb):
    cfg = small_cfg(lb, "baseline", n=64, obs_stride=8, loc_cutoff=20.0, loc_scale=6.0)
    with lb.Context(cfg, 1) as ctx:
        yo = ctx.observe(0, 1, 1)
        with pytest.raises(lb.LbmLetkfError) as e:
            ctx.analysis(yo)
        assert e.value.status == lb.ERR_ARG
