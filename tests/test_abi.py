"""The C-ABI library loads on a host without a GPU, exports every entry point the
public header declares, and its host-only entry points behave (presets,
validation, error reporting)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(lbmletkf_\w+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported(lb):
    names = _declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(lb.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_presets_and_validation(lb):
    for name in ("ref", "baseline"):
        cfg = lb.preset(name)
        lb.config_check(cfg, 64)
    with pytest.raises(lb.LbmLetkfError) as e:
        lb.preset("nope")
    assert e.value.status == lb.ERR_ARG
    cfg = lb.preset("baseline")
    with pytest.raises(lb.LbmLetkfError):
        lb.config_check(cfg, 65)  # more members than lanes of a wavefront
    bad = cfg.copy()
    bad.obs_stride = 7  # does not divide 256 (src/letkf.cu:30 static_assert)
    with pytest.raises(lb.LbmLetkfError, match="obs_stride"):
        lb.config_check(bad, 16)
    bad = cfg.copy()
    bad.loc_cutoff = 200.0
    with pytest.raises(lb.LbmLetkfError, match="cutoff"):
        lb.config_check(bad, 16)
    bad = lb.preset("ref")
    bad.obs_offset = 1
    with pytest.raises(lb.LbmLetkfError, match="offset"):
        lb.config_check(bad, 16)


def test_null_context_is_an_error_not_a_crash(lb):
    lib = lb.lib
    assert lib.lbmletkf_forecast(None, 1) == lb.ERR_ARG
    assert b"NULL" in lib.lbmletkf_last_error(None)
    assert lib.lbmletkf_destroy(None) == lb.OK
