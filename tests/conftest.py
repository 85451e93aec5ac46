import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _load_binding():
    if "lbmletkf" in sys.modules:
        return sys.modules["lbmletkf"]
    path = os.path.join(ROOT, "lbm2d-letkf_amd", "lbmletkf.py")
    spec = importlib.util.spec_from_file_location("lbmletkf", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lbmletkf"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def lb():
    return _load_binding()


@pytest.fixture(scope="session")
def orc():
    import oracle_ref

    return oracle_ref


def small_cfg(lb, preset="baseline", n=64, **kw):
    """A preset on an n x n grid.  The baseline's Gaussian localization (L=12,
    cutoff 43.8 on 256^2) is scaled with the grid when it would not fit."""
    cfg = lb.preset(preset)
    cfg.nx = cfg.ny = n
    if cfg.loc_kind == lb.LOC_GAUSS and 2 * int(cfg.loc_cutoff) + 1 > n:
        cfg.loc_scale = 12.0 * n / 256
        cfg.loc_cutoff = 3.65 * cfg.loc_scale
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def f32_velocity_atol(cfg):
    """Velocity resolution of FP32 populations: u = c sum(f e)/rho with f ~ 1/9,
    each f rounded once when feq is evaluated and once when stored, carries an
    absolute error of up to ~12 half-ulps of 1/9, i.e. ~5e-8 c; 1e-7 c bounds it."""
    return 1e-7 * cfg.c


class HipBuffers:
    """Device buffers for tests that pass device pointers through the C-ABI
    (libamdhip64 through ctypes: hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self):
        import ctypes as C

        self.C = C
        self.l = C.CDLL("libamdhip64.so")
        self.l.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.l.hipFree.argtypes = [C.c_void_p]
        self.l.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def _ck(self, st, what):
        if st != 0:
            raise RuntimeError(f"{what} failed with HIP status {st}")

    def from_host(self, a):
        """A device copy of the contiguous array a; freed by free_all()."""
        import numpy as np

        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        self._ck(self.l.hipMalloc(self.C.byref(p), a.nbytes), "hipMalloc")
        self.live.append(p.value)
        self._ck(self.l.hipMemcpy(p, a.ctypes.data_as(self.C.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def free_all(self):
        for p in self.live:
            self.l.hipFree(self.C.c_void_p(p))
        self.live = []


@pytest.fixture
def hipbuf():
    h = HipBuffers()
    yield h
    h.free_all()
