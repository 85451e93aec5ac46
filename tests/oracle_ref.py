"""ctypes access to the FP64 oracle (oracle/liblbmletkf_oracle.so).

Test infrastructure: the checker used by tests/, by ``__graft_entry__.smoke`` and
by the ``cpu_baseline`` leg of bench.py.  Arrays are float64, f is laid out
[member][q][ny][nx].  The configuration is the binding's ``Config`` (same field
order as ``orc_config``; checked by ``sizeof_config``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liblbmletkf_oracle.so")
if not os.path.exists(LIB):
    raise ImportError(f"{LIB} not built (run __graft_entry__.build())")
_o = C.CDLL(LIB)

P, I, D = C.c_void_p, C.c_int, C.c_double
_sig = {
    "orc_sizeof_config": ([], I),
    "orc_set_threads": ([I], None),
    "orc_set_dsyev": ([P], None),
    "orc_philox4x32_10": ([P, P, P], None),
    "orc_philox_normal4": ([C.c_uint64, P, P], None),
    "orc_force_modes": ([P, P, P, P, I], I),
    "orc_force_field": ([P, C.c_longlong, P, P], None),
    "orc_feq": ([P, P, P, P, P], None),
    "orc_macro": ([P, P, P, P, P], None),
    "orc_stream": ([P, P, P], None),
    "orc_les": ([P, P, P, P], None),
    "orc_lbm_step": ([P, I, P, P, C.c_longlong], None),
    "orc_lbm_step_forced": ([P, I, P, P, P, P], None),
    "orc_initial_velocity": ([P, C.c_ulonglong, P, P], None),
    "orc_obs_count": ([P, C.POINTER(I), C.POINTER(I)], None),
    "orc_local_obs": ([P, I, P, P, I], I),
    "orc_gaspari_cohn": ([D], D),
    "orc_obs_moments": ([P, I, P, P], None),
    "orc_observe": ([P, P, C.c_ulonglong, C.c_ulonglong, I, P], None),
    "orc_jacobi_eig": ([I, P, P, P, D, I], I),
    "orc_letkf_transform": ([I, I, P, P, P, D, D, I, P, P, P], I),
    "orc_letkf_analysis": ([P, I, P, P, I, I, P, P, P, I], I),
    "orc_letkf_analysis_strip": ([P, I, P, P, I, I, P, P, P], I),
    "orc_ensemble_stats": ([P, I, P, P, P, P, P], None),
    "orc_closure_features": ([I, I, D, D, P, P, P, P], None),
    "orc_closure_nut": ([P, P, C.c_longlong, D, I, P], None),
    "orc_closure_loss_grad": ([P, P, C.c_longlong, D, D, P], D),
    "orc_coarse_grain": ([I, I, I, P, P, P, D, D, P, P], None),
    "orc_lbm_step_closure": ([P, I, P, P, C.c_longlong, P, I], None),
    "orc_nudge": ([P, I, P, P, P, P, D, I, I], None),
    "orc_set_obs_sites": ([I, P], None),
    "orc_perturb_draws": ([P, C.c_ulonglong, P, P], None),
    "orc_perturb": ([P, P, P, P, D], None),
}
CLOSURE_NPARAM = 5441
CLOSURE_SAMPLE = 22
for _n, (_a, _r) in _sig.items():
    _f = getattr(_o, _n)
    _f.argtypes = _a
    _f.restype = _r


def _p(a):
    # data_as() returns a pointer object that holds a reference to the array, so
    # temporaries such as _p(_f64(x)) stay alive for the duration of the call
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _c(cfg):
    return C.c_void_p(C.addressof(cfg))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def sizeof_config() -> int:
    return _o.orc_sizeof_config()


def set_threads(n: int):
    _o.orc_set_threads(n)


def set_dsyev(on: bool) -> str:
    """Route the oracle's local eigensolves through LAPACK dsyev (SciPy's LAPACK,
    reached through scipy.linalg.cython_lapack) or back to its own cyclic Jacobi.
    Returns the name of the eigensolver in use."""
    if not on:
        _o.orc_set_dsyev(None)
        return "oracle cyclic Jacobi (FP64)"
    from scipy.linalg import cython_lapack

    cap = cython_lapack.__pyx_capi__["dsyev"]
    get = C.pythonapi.PyCapsule_GetPointer
    get.restype = C.c_void_p
    get.argtypes = [C.py_object, C.c_char_p]
    name = C.pythonapi.PyCapsule_GetName
    name.restype = C.c_char_p
    name.argtypes = [C.py_object]
    _o.orc_set_dsyev(C.c_void_p(get(cap, name(cap))))
    return "LAPACK dsyev (SciPy " + __import__("scipy").__version__ + ")"


def philox(ctr, key):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    out = np.empty(4, np.uint32)
    _o.orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def philox_normal4(key: int, ctr):
    c = np.array(ctr, np.uint32)
    out = np.empty(4)
    _o.orc_philox_normal4(key, _p(c), _p(out))
    return out


def force_modes(cfg):
    kx, ky, amp = (np.empty(100) for _ in range(3))
    n = _o.orc_force_modes(_c(cfg), _p(kx), _p(ky), _p(amp), 100)
    return kx[:n], ky[:n], amp[:n]


def force_field(cfg, step: int):
    fx = np.empty((cfg.ny, cfg.nx))
    fy = np.empty((cfg.ny, cfg.nx))
    _o.orc_force_field(_c(cfg), step, _p(fx), _p(fy))
    return fx, fy


def feq(cfg, rho, u, v):
    f = np.empty((9, cfg.ny, cfg.nx))
    _o.orc_feq(_c(cfg), _p(_f64(rho)), _p(_f64(u)), _p(_f64(v)), _p(f))
    return f


def macro(cfg, f):
    r, u, v = (np.empty((cfg.ny, cfg.nx)) for _ in range(3))
    _o.orc_macro(_c(cfg), _p(_f64(f)), _p(r), _p(u), _p(v))
    return r, u, v


def stream(cfg, f):
    fn = np.empty((9, cfg.ny, cfg.nx))
    _o.orc_stream(_c(cfg), _p(_f64(f)), _p(fn))
    return fn


def les(cfg, u, v):
    nus = np.empty((cfg.ny, cfg.nx))
    _o.orc_les(_c(cfg), _p(_f64(u)), _p(_f64(v)), _p(nus))
    return nus


def lbm_step(cfg, f, step: int):
    fin = _f64(f)
    M = fin.shape[0]
    fout = np.empty_like(fin)
    _o.orc_lbm_step(_c(cfg), M, _p(fin), _p(fout), step)
    return fout


def lbm_step_forced(cfg, f, Fx, Fy):
    """One step with the given body-force field [ny][nx] in place of the configured forcing."""
    fin = _f64(f)
    M = fin.shape[0]
    fout = np.empty_like(fin)
    _o.orc_lbm_step_forced(_c(cfg), M, _p(fin), _p(fout), _p(_f64(Fx)), _p(_f64(Fy)))
    return fout


def initial_velocity(cfg, seed: int):
    u = np.empty((cfg.ny, cfg.nx))
    v = np.empty((cfg.ny, cfg.nx))
    _o.orc_initial_velocity(_c(cfg), seed, _p(u), _p(v))
    return u, v


def obs_count(cfg):
    ns, no = C.c_int(), C.c_int()
    _o.orc_obs_count(_c(cfg), C.byref(ns), C.byref(no))
    return ns.value, no.value


def local_obs(cfg, g: int, cap: int = 4096):
    s = np.empty(cap, np.int32)
    w = np.empty(cap)
    n = _o.orc_local_obs(_c(cfg), g, _p(s), _p(w), cap)
    n = min(n, cap)
    return s[:n].copy(), w[:n].copy()


def gaspari_cohn(d: float) -> float:
    return _o.orc_gaspari_cohn(d)


def obs_moments(cfg, f):
    fin = _f64(f)
    M = fin.shape[0]
    _, n_obs = obs_count(cfg)
    hx = np.empty((n_obs, M))
    _o.orc_obs_moments(_c(cfg), M, _p(fin), _p(hx))
    return hx


def observe(cfg, hx_member, seed: int, counter: int, noise: bool = True):
    _, n_obs = obs_count(cfg)
    yo = np.empty(n_obs)
    _o.orc_observe(_c(cfg), _p(_f64(hx_member)), seed, counter, int(noise), _p(yo))
    return yo


def jacobi_eig(A, tol=1e-14, max_sweeps=100):
    A = _f64(A).copy()
    n = A.shape[0]
    V = np.empty((n, n))
    d = np.empty(n)
    sweeps = _o.orc_jacobi_eig(n, _p(A), _p(V), _p(d), tol, max_sweeps)
    return d, V, sweeps


def letkf_transform(dY, rinv, dyo, beta, tol=1e-14, max_sweeps=100):
    dY = _f64(dY)
    p, M = dY.shape
    W = np.empty((M, M))
    w = np.empty(M)
    ev = np.empty(M)
    _o.orc_letkf_transform(M, p, _p(dY), _p(_f64(rinv)), _p(_f64(dyo)), beta, tol, max_sweeps,
                           _p(W), _p(w), _p(ev))
    return W, w, ev


def letkf_analysis(cfg, f, yo, dump_g0=0, dump_n=0, g_stride=1):
    """Analyse f (float64 [M][9][ny][nx]) in place; returns (W, w, evals) of the dump range."""
    assert f.dtype == np.float64 and f.flags.c_contiguous
    M = f.shape[0]
    W = np.empty((max(dump_n, 1), M, M))
    w = np.empty((max(dump_n, 1), M))
    ev = np.empty((max(dump_n, 1), M))
    st = _o.orc_letkf_analysis(_c(cfg), M, _p(f), _p(_f64(yo)), dump_g0, dump_n, _p(W), _p(w),
                               _p(ev), g_stride)
    if st != 0:
        raise FloatingPointError("oracle analysis produced a non-finite state")
    return W[:dump_n], w[:dump_n], ev[:dump_n]


def letkf_analysis_strip(cfg, f, yo, g0, n):
    """Analyse gridpoints [g0, g0+n) of f (float64 [M][9][ny][nx]) in place; returns
    (W, w, evals) of each."""
    assert f.dtype == np.float64 and f.flags.c_contiguous
    M = f.shape[0]
    W = np.empty((max(n, 1), M, M))
    w = np.empty((max(n, 1), M))
    ev = np.empty((max(n, 1), M))
    st = _o.orc_letkf_analysis_strip(_c(cfg), M, _p(f), _p(_f64(yo)), g0, n, _p(W), _p(w), _p(ev))
    if st != 0:
        raise (ValueError("strip outside the grid") if st == -1 else
               FloatingPointError("oracle analysis produced a non-finite state"))
    return W[:n], w[:n], ev[:n]


def ensemble_stats(cfg, f, trho, tu, tv):
    fin = _f64(f)
    out = np.empty(4)
    _o.orc_ensemble_stats(_c(cfg), fin.shape[0], _p(fin), _p(_f64(trho)), _p(_f64(tu)),
                          _p(_f64(tv)), _p(out))
    return out


# ---- learned eddy-viscosity closure ------------------------------------------
def closure_features(nx, ny, dx, strain_scale, u, v):
    """(feat [ny*nx][18], ab [ny*nx][2])"""
    feat = np.empty((nx * ny, 18))
    ab = np.empty((nx * ny, 2))
    _o.orc_closure_features(nx, ny, dx, strain_scale, _p(_f64(u)), _p(_f64(v)), _p(feat), _p(ab))
    return feat, ab


def closure_nut(params, feat, nu_scale, bf16=False):
    feat = _f64(feat)
    nut = np.empty(feat.shape[0])
    _o.orc_closure_nut(_p(_f64(params)), _p(feat), feat.shape[0], nu_scale, int(bf16), _p(nut))
    return nut


def closure_loss_grad(params, samples, nu_scale, w_norm=1.0):
    samples = _f64(samples)
    grad = np.empty(CLOSURE_NPARAM)
    loss = _o.orc_closure_loss_grad(_p(_f64(params)), _p(samples), samples.shape[0], nu_scale, w_norm, _p(grad))
    return loss, grad


def coarse_grain(rho, u, v, F, dx_coarse, strain_scale):
    """fine [nyf][nxf] fields -> (samples [n][22], coarse [3][ny][nx])"""
    nyf, nxf = rho.shape
    ny, nx = nyf // F, nxf // F
    samples = np.empty((nx * ny, CLOSURE_SAMPLE))
    coarse = np.empty((3, ny, nx))
    _o.orc_coarse_grain(nxf, nyf, F, _p(_f64(rho)), _p(_f64(u)), _p(_f64(v)), dx_coarse, strain_scale,
                        _p(samples), _p(coarse))
    return samples, coarse


def lbm_step_closure(cfg, f, step: int, params, bf16=False):
    fin = _f64(f)
    fout = np.empty_like(fin)
    _o.orc_lbm_step_closure(_c(cfg), fin.shape[0], _p(fin), _p(fout), step, _p(_f64(params)), int(bf16))
    return fout


# ---- nudging and the Lyapunov perturbation -----------------------------------------
def nudge(cfg, f, u, v, rho=None, alpha=0.1, xyprune=1, quadra=0):
    out = _f64(f).copy()
    M = out.shape[0]
    _o.orc_nudge(_c(cfg), M, _p(out), None if rho is None else _p(_f64(rho)), _p(_f64(u)), _p(_f64(v)),
                 alpha, xyprune, quadra)
    return out


def perturb_draws(cfg, ensemble_id):
    ur = np.empty((cfg.ny, cfg.nx))
    vr = np.empty((cfg.ny, cfg.nx))
    _o.orc_perturb_draws(_c(cfg), ensemble_id, _p(ur), _p(vr))
    return ur, vr


def perturb(cfg, f, ur, vr, eps):
    out = _f64(f).copy()
    _o.orc_perturb(_c(cfg), _p(out), _p(_f64(ur)), _p(_f64(vr)), eps)
    return out


def set_obs_sites(nodes):
    """Use an explicit site list (None: back to the configuration's lattice)."""
    global _sites_keep
    if nodes is None:
        _sites_keep = None
        _o.orc_set_obs_sites(-1, None)
        return
    _sites_keep = np.ascontiguousarray(nodes, dtype=np.int32)
    _o.orc_set_obs_sites(_sites_keep.size, _sites_keep.ctypes.data_as(C.c_void_p))


_sites_keep = None
