"""ctypes binding of liblbmletkf.so (the C-ABI in include/lbmletkf.h).

The names mirror the reference's host interface so that callers read like the
reference's driver (src/main.cu:22-155):

* ``update(ctx, n)``                 -- n calls of ``update(data&, t)`` (src/lbm.h:16-29)
* ``DataAssimilator(ctx).assimilate(yo)`` -- ``DataAssimilator::assimilate`` (src/assimilate.h:54-68)

Every call raises :class:`LbmLetkfError` on a non-zero status.  There is no
fallback: if the HIP library is missing the import fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbmletkf.so")

OK, ERR_ARG, ERR_HIP, ERR_NUMERIC, ERR_STATE = 0, -1, -2, -3, -4
FORCE_NONE, FORCE_WATANABE97 = 0, 1
IC_REFERENCE, IC_K4_SPECTRUM = 0, 1
OBS_U, OBS_V, OBS_RHO = 1, 2, 4
LOC_GC_WINDOW, LOC_GAUSS = 0, 1
VARIANT_PLAIN, VARIANT_MIXED = 0, 1
CLOSURE_SMAGORINSKY, CLOSURE_MLP = 0, 1
CLOSURE_NPARAM, CLOSURE_SAMPLE = 5441, 22


class Config(C.Structure):
    """Field-for-field mirror of ``lbmletkf_config``."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("c", C.c_double), ("dx", C.c_double), ("dt", C.c_double), ("nu", C.c_double),
        ("rho_ref", C.c_double), ("cs_smag", C.c_double), ("u_ref", C.c_double),
        ("friction_rate", C.c_double),
        ("force_kind", C.c_int),
        ("force_kf", C.c_double), ("force_amp_total", C.c_double), ("force_sigma_k", C.c_double),
        ("force_dk", C.c_double), ("force_rand_stdev", C.c_double),
        ("force_seed", C.c_ulonglong),
        ("ic_kind", C.c_int),
        ("ic_amp", C.c_double), ("ic_k0", C.c_double), ("ic_dk", C.c_double), ("ic_sigma", C.c_double),
        ("obs_stride", C.c_int), ("obs_offset", C.c_int), ("obs_vars", C.c_int),
        ("obs_sigma_u", C.c_double), ("obs_sigma_rho", C.c_double),
        ("loc_kind", C.c_int), ("loc_rloc_len", C.c_int),
        ("loc_scale", C.c_double), ("loc_cutoff", C.c_double), ("inflation", C.c_double),
        ("da_interval", C.c_int), ("variant", C.c_int),
        ("jacobi_tol", C.c_double), ("jacobi_max_sweeps", C.c_int),
        ("gram_passes", C.c_int), ("jacobi_fp64_sweeps", C.c_int),
        ("closure_kind", C.c_int), ("closure_strain_scale", C.c_double), ("closure_nu_scale", C.c_double),
        ("closure_nu_max", C.c_double),
        ("jacobi_tol32", C.c_double),
    ]

    def copy(self) -> "Config":
        c = Config()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(Config))
        return c

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class LbmLetkfError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"lbmletkf status {status}: {message}")
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P, I, D, F = C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_float)
    sig = {
        "lbmletkf_version": ([], C.c_char_p),
        "lbmletkf_config_preset": ([C.c_char_p, C.POINTER(Config)], I),
        "lbmletkf_config_check": ([C.POINTER(Config), I], I),
        "lbmletkf_create": ([C.POINTER(Config), I, C.POINTER(P)], I),
        "lbmletkf_destroy": ([P], I),
        "lbmletkf_last_error": ([P], C.c_char_p),
        "lbmletkf_init_member": ([P, I, C.c_ulonglong], I),
        "lbmletkf_set_state": ([P, I, P], I),
        "lbmletkf_get_state": ([P, I, P], I),
        "lbmletkf_set_macro": ([P, I, P, P, P], I),
        "lbmletkf_get_macro": ([P, I, P, P, P], I),
        "lbmletkf_get_macro_device": ([P, I, P, P, P], I),
        "lbmletkf_forecast": ([P, I], I),
        "lbmletkf_set_step": ([P, C.c_longlong], I),
        "lbmletkf_get_step": ([P, C.POINTER(C.c_longlong)], I),
        "lbmletkf_setup_obs": ([P], I),
        "lbmletkf_obs_count": ([P, C.POINTER(I), C.POINTER(I)], I),
        "lbmletkf_observe": ([P, I, C.c_ulonglong, C.c_ulonglong, I, P, I], I),
        "lbmletkf_obs_from_fields": ([P, P, P, P, P], I),
        "lbmletkf_analysis": ([P, P, I], I),
        "lbmletkf_cycle": ([P, P], I),
        "lbmletkf_ensemble_stats": ([P, P, P, P, I, P], I),
        "lbmletkf_ensemble_stats_noalloc": ([P, P, P, P, P], I),
        "lbmletkf_local_obs": ([P, I, C.POINTER(I), P, P, I], I),
        "lbmletkf_set_dump": ([P, I, I], I),
        "lbmletkf_get_dump": ([P, P, P, P], I),
        "lbmletkf_get_dump_evd": ([P, P, P, P], I),
        "lbmletkf_get_obs_ensemble": ([P, P, P], I),
        "lbmletkf_profile_enable": ([P, I], I),
        "lbmletkf_profile_read": ([P, C.c_char_p, C.POINTER(D), C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_longlong)], I),
        "lbmletkf_nudge": ([P, P, P, P, I, D, I, I], I),
        "lbmletkf_set_obs_sites": ([P, I, P], I),
        "lbmletkf_perturb": ([P, I, C.c_ulonglong, D], I),
        "lbmletkf_closure_init_params": ([C.c_ulonglong, P], I),
        "lbmletkf_set_closure": ([P, P], I),
        "lbmletkf_closure_nut": ([P, I, P], I),
        "lbmletkf_filtered_macro": ([P, I, I, P, P, P], I),
        "lbmletkf_trainer_create": ([C.POINTER(Config), C.c_longlong, C.POINTER(P)], I),
        "lbmletkf_trainer_destroy": ([P], I),
        "lbmletkf_trainer_add_samples": ([P, P, I, I], I),
        "lbmletkf_trainer_set_samples": ([P, P, C.c_longlong], I),
        "lbmletkf_trainer_get_samples": ([P, C.c_longlong, C.c_longlong, P], I),
        "lbmletkf_trainer_count": ([P, C.POINTER(C.c_longlong)], I),
        "lbmletkf_trainer_set_params": ([P, P], I),
        "lbmletkf_trainer_get_params": ([P, P], I),
        "lbmletkf_trainer_loss_grad": ([P, C.c_longlong, C.c_longlong, D, C.POINTER(D), P], I),
        "lbmletkf_trainer_train": ([P, I, I, D, D, P], I),
        "lbmletkf_set_tuning": ([P, I, I], I),
        "lbmletkf_get_tuning": ([P, C.POINTER(I), C.POINTER(I)], I),
        "lbmletkf_autotune": ([P, C.c_char_p, C.c_char_p, I], I),
        "lbmletkf_group_create": ([P, I, C.POINTER(P)], I),
        "lbmletkf_group_destroy": ([P], I),
        "lbmletkf_group_forecast": ([P, I], I),
        "lbmletkf_group_analysis": ([P, P], I),
        "lbmletkf_group_cycle": ([P, P], I),
        "lbmletkf_group_synchronize": ([P], I),
        "lbmletkf_group_get_stream": ([P], P),
        "lbmletkf_ensemble_fields": ([P, P, P, P, P, P, P], I),
        "lbmletkf_get_stream": ([P], P),
        "lbmletkf_set_stream": ([P, P], I),
        "lbmletkf_synchronize": ([P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def _ptr(a: np.ndarray) -> C.c_void_p:
    # the pointer object keeps a reference to the array for the duration of the call
    return a.ctypes.data_as(C.c_void_p)


def _check(status: int, ctx_handle=None):
    if status != OK:
        msg = lib.lbmletkf_last_error(ctx_handle)
        raise LbmLetkfError(status, msg.decode() if msg else "")


def preset(name: str) -> Config:
    cfg = Config()
    _check(lib.lbmletkf_config_preset(name.encode(), C.byref(cfg)))
    return cfg


def config_check(cfg: Config, n_members: int) -> None:
    _check(lib.lbmletkf_config_check(C.byref(cfg), n_members))


def version() -> str:
    return lib.lbmletkf_version().decode()


class Context:
    """An ensemble of M members on the current GPU (``data`` x M + ``DataAssimilator``)."""

    def __init__(self, cfg: Config, n_members: int):
        self.cfg = cfg.copy()
        self.M = n_members
        self.nx, self.ny = cfg.nx, cfg.ny
        self.nodes = cfg.nx * cfg.ny
        h = C.c_void_p()
        _check(lib.lbmletkf_create(C.byref(self.cfg), n_members, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.lbmletkf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ck(self, st):
        _check(st, self.h)

    # state ------------------------------------------------------------------
    def init_member(self, member: int, seed: int):
        self._ck(lib.lbmletkf_init_member(self.h, member, seed))

    def set_state(self, member: int, f: np.ndarray):
        a = np.ascontiguousarray(f, dtype=np.float32).reshape(9, self.ny, self.nx)
        self._ck(lib.lbmletkf_set_state(self.h, member, _ptr(a)))

    def get_state(self, member: int) -> np.ndarray:
        a = np.empty((9, self.ny, self.nx), np.float32)
        self._ck(lib.lbmletkf_get_state(self.h, member, _ptr(a)))
        return a

    def get_all_states(self) -> np.ndarray:
        return np.stack([self.get_state(k) for k in range(self.M)])

    def set_all_states(self, f: np.ndarray):
        for k in range(self.M):
            self.set_state(k, f[k])

    def set_macro(self, member: int, rho, u, v):
        arrs = [np.ascontiguousarray(x, dtype=np.float32).reshape(self.ny, self.nx) for x in (rho, u, v)]
        self._ck(lib.lbmletkf_set_macro(self.h, member, *[_ptr(a) for a in arrs]))

    def get_macro(self, member: int):
        r, u, v = (np.empty((self.ny, self.nx), np.float32) for _ in range(3))
        self._ck(lib.lbmletkf_get_macro(self.h, member, _ptr(r), _ptr(u), _ptr(v)))
        return r, u, v

    # forecast -----------------------------------------------------------------
    def forecast(self, n_steps: int = 1):
        self._ck(lib.lbmletkf_forecast(self.h, n_steps))

    def set_step(self, step: int):
        self._ck(lib.lbmletkf_set_step(self.h, step))

    def get_step(self) -> int:
        s = C.c_longlong()
        self._ck(lib.lbmletkf_get_step(self.h, C.byref(s)))
        return s.value

    # observations and analysis -----------------------------------------------
    def obs_count(self):
        ns, no = C.c_int(), C.c_int()
        self._ck(lib.lbmletkf_obs_count(self.h, C.byref(ns), C.byref(no)))
        return ns.value, no.value

    def observe(self, member: int, seed: int, counter: int, noise: bool = True) -> np.ndarray:
        _, n_obs = self.obs_count()
        yo = np.empty(n_obs, np.float32)
        self._ck(lib.lbmletkf_observe(self.h, member, seed, counter, int(noise), _ptr(yo), 0))
        return yo

    def observe_device(self, member: int, seed: int, counter: int, yo_dev: int, noise: bool = True):
        self._ck(lib.lbmletkf_observe(self.h, member, seed, counter, int(noise), C.c_void_p(yo_dev), 1))

    def obs_from_fields(self, rho=None, u=None, v=None) -> np.ndarray:
        _, n_obs = self.obs_count()
        yo = np.empty(n_obs, np.float32)
        arrs = [None if x is None else np.ascontiguousarray(x, np.float32) for x in (rho, u, v)]
        ptrs = [None if a is None else _ptr(a) for a in arrs]
        self._ck(lib.lbmletkf_obs_from_fields(self.h, *ptrs, _ptr(yo)))
        return yo

    def analysis(self, yo: np.ndarray):
        a = np.ascontiguousarray(yo, dtype=np.float32)
        self._ck(lib.lbmletkf_analysis(self.h, _ptr(a), 0))
        self.synchronize()

    def analysis_device(self, yo_dev: int):
        self._ck(lib.lbmletkf_analysis(self.h, C.c_void_p(yo_dev), 1))

    def cycle_device(self, yo_dev: int):
        self._ck(lib.lbmletkf_cycle(self.h, C.c_void_p(yo_dev)))

    def local_obs(self, gridpoint: int, cap: int = 4096):
        n = C.c_int()
        sites = np.empty(cap, np.int32)
        w = np.empty(cap, np.float32)
        self._ck(lib.lbmletkf_local_obs(self.h, gridpoint, C.byref(n), _ptr(sites), _ptr(w), cap))
        k = min(n.value, cap)
        return sites[:k].copy(), w[:k].copy()

    def set_dump(self, g0: int, count: int):
        self._ck(lib.lbmletkf_set_dump(self.h, g0, count))
        self._dump = (g0, count)

    def get_dump(self):
        g0, count = self._dump
        M = self.M
        W = np.empty((count, M, M))
        w = np.empty((count, M))
        ev = np.empty((count, M))
        self._ck(lib.lbmletkf_get_dump(self.h, _ptr(W), _ptr(w), _ptr(ev)))
        return W, w, ev

    def get_dump_evd(self):
        """(C, V, evals) of the dumped gridpoints: C and V [count][M][M] (row-major, column
        j of V the eigenvector of evals[:, j]), eigenvalues in the solver's order."""
        g0, count = self._dump
        M = self.M
        Cm = np.empty((count, M, M))
        V = np.empty((count, M, M))
        ev = np.empty((count, M))
        self._ck(lib.lbmletkf_get_dump_evd(self.h, _ptr(Cm), _ptr(V), _ptr(ev)))
        return Cm, V, ev

    def ensemble_fields(self):
        """Per-node (mean_u, mean_v, vor_mean, vor_std, sd_u, sd_v) over the members,
        each [ny, nx] (src/io.cu:90-145 ENSEMBLE_STAT and the spread fields)."""
        out = [np.empty((self.ny, self.nx), np.float32) for _ in range(6)]
        self._ck(lib.lbmletkf_ensemble_fields(self.h, *[_ptr(a) for a in out]))
        return tuple(out)

    def get_obs_ensemble(self):
        """(dy [n_obs][M], dyo [n_obs]) of the last analysis."""
        _, n_obs = self.obs_count()
        dy = np.empty((n_obs, self.M), np.float32)
        dyo = np.empty(n_obs, np.float32)
        self._ck(lib.lbmletkf_get_obs_ensemble(self.h, _ptr(dy), _ptr(dyo)))
        return dy, dyo

    def ensemble_stats(self, truth_rho, truth_u, truth_v):
        arrs = [np.ascontiguousarray(x, np.float32) for x in (truth_rho, truth_u, truth_v)]
        out = np.empty(4)
        self._ck(lib.lbmletkf_ensemble_stats(self.h, *[_ptr(a) for a in arrs], 0, _ptr(out)))
        return out

    def ensemble_stats_device(self, d_rho: int, d_u: int, d_v: int):
        out = np.empty(4)
        self._ck(lib.lbmletkf_ensemble_stats(self.h, C.c_void_p(d_rho), C.c_void_p(d_u),
                                             C.c_void_p(d_v), 1, _ptr(out)))
        return out

    def get_macro_device(self, member: int, d_rho: int, d_u: int, d_v: int):
        self._ck(lib.lbmletkf_get_macro_device(self.h, member, C.c_void_p(d_rho), C.c_void_p(d_u),
                                               C.c_void_p(d_v)))

    def profile_enable(self, on: bool = True):
        self._ck(lib.lbmletkf_profile_enable(self.h, int(on)))

    def profile_read(self, phase: str):
        """(seconds, kernel launches, work count) accumulated for one phase."""
        sec, n, w = C.c_double(), C.c_longlong(), C.c_longlong()
        self._ck(lib.lbmletkf_profile_read(self.h, phase.encode(), C.byref(sec), C.byref(n), C.byref(w)))
        return sec.value, n.value, w.value

    def set_tuning(self, lbm_tiles_per_block: int, mix_chain: int):
        """LBM shape (k > 0 tiles per workgroup, -h row strips, -(100 + h) two-step
        and -(200 + h) three-step row strips) and analysis chain length; see
        lbmletkf_set_tuning."""
        self._ck(lib.lbmletkf_set_tuning(self.h, lbm_tiles_per_block, mix_chain))

    def get_tuning(self):
        """(lbm_tiles_per_block, mix_chain)"""
        a, b = C.c_int(), C.c_int()
        self._ck(lib.lbmletkf_get_tuning(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, cache_path: str | None = None) -> str:
        """Time the launch-shape candidates and keep the fastest; returns the report."""
        buf = C.create_string_buffer(1024)
        self._ck(lib.lbmletkf_autotune(self.h, cache_path.encode() if cache_path else None, buf, 1024))
        return buf.value.decode()

    def set_obs_sites(self, nodes):
        """Explicit observation network: sites at the given node indices."""
        a = np.ascontiguousarray(nodes, dtype=np.int32).ravel()
        self._ck(lib.lbmletkf_set_obs_sites(self.h, a.size, _ptr(a)))

    def nudge(self, u, v, rho=None, alpha: float = 0.1, xyprune: int = 1, quadra: int = 0):
        """Newtonian nudging of every member towards observed (rho,) u, v fields."""
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float32).reshape(self.ny, self.nx)
             for x in (rho, u, v)]
        ptrs = [None if x is None else _ptr(x) for x in a]
        self._ck(lib.lbmletkf_nudge(self.h, ptrs[0], ptrs[1], ptrs[2], 0, alpha, xyprune, quadra))

    def perturb(self, member: int, ensemble_id: int, epsilon: float):
        """The Lyapunov runs' random-velocity perturbation of one member."""
        self._ck(lib.lbmletkf_perturb(self.h, member, ensemble_id, epsilon))

    def set_closure(self, params: np.ndarray):
        """Load the learned closure (closure_kind = MLP contexts)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != CLOSURE_NPARAM:
            raise ValueError(f"expected {CLOSURE_NPARAM} parameters")
        self._ck(lib.lbmletkf_set_closure(self.h, _ptr(p)))

    def closure_nut(self, member: int) -> np.ndarray:
        """nu_t [ny, nx] the next step would use for one member (after the fallback)."""
        out = np.empty((self.cfg.ny, self.cfg.nx), np.float32)
        self._ck(lib.lbmletkf_closure_nut(self.h, member, _ptr(out)))
        return out

    def filtered_macro(self, member: int, factor: int):
        """Box-filtered (rho, u, v) of one member on the grid coarsened by factor."""
        shape = (self.cfg.ny // factor, self.cfg.nx // factor)
        r, u, v = (np.empty(shape, np.float32) for _ in range(3))
        self._ck(lib.lbmletkf_filtered_macro(self.h, member, factor, _ptr(r), _ptr(u), _ptr(v)))
        return r, u, v

    @property
    def stream(self) -> int:
        return lib.lbmletkf_get_stream(self.h) or 0

    def set_stream(self, stream: int):
        """Run on the caller's hipStream_t (an integer handle) from now on; 0 returns
        to a stream the context owns."""
        self._ck(lib.lbmletkf_set_stream(self.h, C.c_void_p(stream or None)))

    def synchronize(self):
        self._ck(lib.lbmletkf_synchronize(self.h))


def uneven_sites(nx: int, ny: int, seed: int, regions: int = 4, density=(0.005, 0.04),
                 empty_regions: int = 0) -> np.ndarray:
    """Observation nodes at a seeded random density per region: the grid is cut
    into regions x regions blocks, each block drawing its density uniformly from
    `density` and its sites without replacement; the first `empty_regions` blocks
    (row-major) get none.  Returns the sorted node indices i + nx*j."""
    rng = np.random.default_rng(seed)
    bx, by = nx // regions, ny // regions
    out = []
    for r in range(regions * regions):
        rx, ry = r % regions, r // regions
        x0, y0 = rx * bx, ry * by
        x1 = nx if rx == regions - 1 else x0 + bx
        y1 = ny if ry == regions - 1 else y0 + by
        d = rng.uniform(*density)
        nodes = (np.arange(y0, y1)[:, None] * nx + np.arange(x0, x1)[None, :]).ravel()
        k = int(round(d * nodes.size))
        pick = rng.choice(nodes, size=k, replace=False)
        if r >= empty_regions:
            out.append(pick)
    return np.sort(np.concatenate(out)).astype(np.int32) if out else np.zeros(0, np.int32)


def closure_init_params(seed: int) -> np.ndarray:
    """Initial closure parameters (Glorot-uniform from a Philox stream)."""
    p = np.empty(CLOSURE_NPARAM, np.float32)
    _check(lib.lbmletkf_closure_init_params(seed, _ptr(p)))
    return p


class Trainer:
    """Closure trainer: samples, parameters and Adam state on the GPU."""

    def __init__(self, coarse_cfg: Config, capacity: int):
        self.h = None
        h = C.c_void_p()
        _check(lib.lbmletkf_trainer_create(C.byref(coarse_cfg), capacity, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib.lbmletkf_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_samples(self, fine: Context, member: int, factor: int):
        _check(lib.lbmletkf_trainer_add_samples(self.h, fine.h, member, factor))

    def set_samples(self, samples: np.ndarray):
        s = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, CLOSURE_SAMPLE)
        _check(lib.lbmletkf_trainer_set_samples(self.h, _ptr(s), s.shape[0]))

    def count(self) -> int:
        n = C.c_longlong()
        _check(lib.lbmletkf_trainer_count(self.h, C.byref(n)))
        return n.value

    def get_samples(self, first: int = 0, n: int | None = None) -> np.ndarray:
        n = self.count() - first if n is None else n
        out = np.empty((n, CLOSURE_SAMPLE), np.float32)
        _check(lib.lbmletkf_trainer_get_samples(self.h, first, n, _ptr(out)))
        return out

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != CLOSURE_NPARAM:
            raise ValueError(f"expected {CLOSURE_NPARAM} parameters")
        _check(lib.lbmletkf_trainer_set_params(self.h, _ptr(p)))

    def get_params(self) -> np.ndarray:
        p = np.empty(CLOSURE_NPARAM, np.float32)
        _check(lib.lbmletkf_trainer_get_params(self.h, _ptr(p)))
        return p

    def loss_grad(self, first: int = 0, n: int | None = None, w_norm: float = 0.0):
        """(mean loss, gradient) over samples [first, first+n)."""
        n = self.count() - first if n is None else n
        loss = C.c_double()
        g = np.empty(CLOSURE_NPARAM, np.float32)
        _check(lib.lbmletkf_trainer_loss_grad(self.h, first, n, w_norm, C.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, batch: int, steps: int, lr: float, w_norm: float = 0.0) -> np.ndarray:
        """Adam steps; returns the loss of each minibatch before its update."""
        hist = np.empty(max(steps, 1), np.float32)
        _check(lib.lbmletkf_trainer_train(self.h, batch, steps, lr, w_norm, _ptr(hist)))
        return hist[:steps]


class Group:
    """Several contexts stepped and analysed by shared launches (see the C header)."""

    def __init__(self, contexts):
        self.h = None
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        _check(lib.lbmletkf_group_create(arr, len(self.contexts), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib.lbmletkf_group_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _yo(self, yo_devs):
        if len(yo_devs) != len(self.contexts):
            raise ValueError("one device observation vector per context")
        return (C.c_void_p * len(yo_devs))(*[int(p) for p in yo_devs])

    def forecast(self, n_steps: int = 1):
        _check(lib.lbmletkf_group_forecast(self.h, n_steps))

    def analysis_device(self, yo_devs):
        _check(lib.lbmletkf_group_analysis(self.h, self._yo(yo_devs)))

    def cycle_device(self, yo_devs):
        _check(lib.lbmletkf_group_cycle(self.h, self._yo(yo_devs)))

    def synchronize(self):
        _check(lib.lbmletkf_group_synchronize(self.h))

    @property
    def stream(self) -> int:
        return lib.lbmletkf_group_get_stream(self.h) or 0


def update(ctx: Context, n_steps: int = 1):
    """``update(data&, t)`` applied n times to every member (src/lbm.h:16-29)."""
    ctx.forecast(n_steps)


class DataAssimilator:
    """``DataAssimilator`` (src/assimilate.h:14-82) bound to one context."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        ctx.obs_count()  # init_letkf: builds the local observation sets

    def assimilate(self, yo: np.ndarray):
        self.ctx.analysis(yo)
