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
    cfg = lb.preset(preset)
    cfg.nx = cfg.ny = n
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg
