"""A context on the caller's stream (lbmletkf_set_stream): the same results bit for
bit as a context on its own stream, through eager steps, an analysis and captured
cycles; and back to an owned stream."""
import ctypes as C

import numpy as np
import pytest

from conftest import small_cfg

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    return h


def _run(lb, cfg, M, nat_yo, hipbuf, ens):
    for k in range(M):
        ens.init_member(k, 1000 + k)
    ens.forecast(10)
    ens.analysis(nat_yo[0])
    for y in nat_yo[1:]:
        ens.cycle_device(hipbuf.from_host(y))  # analysis, then da_interval steps (graph)
    ens.synchronize()
    return ens.get_all_states()


def test_context_on_the_callers_stream(lb, hipbuf):
    hip = _hip()
    cfg = small_cfg(lb, "baseline", n=64, loc_scale=4.0, loc_cutoff=10.0, obs_stride=8)
    cfg.da_interval = 5
    M = 16
    print("case: caller-supplied stream, 64x64 M=16", flush=True)
    with lb.Context(cfg, 1) as nat:
        nat.init_member(0, 7)
        yo = []
        for c in range(4):
            nat.forecast(10 if c == 0 else cfg.da_interval)
            yo.append(nat.observe(0, seed=3, counter=c, noise=True))
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    try:
        with lb.Context(cfg, M) as own:
            want = _run(lb, cfg, M, yo, hipbuf, own)
        with lb.Context(cfg, M) as ens:
            ens.set_stream(s.value)
            assert ens.stream == s.value
            got = _run(lb, cfg, M, yo, hipbuf, ens)
            np.testing.assert_array_equal(got, want)
            # work is queued on the caller's stream: a forecast is complete once that
            # stream is
            ens.forecast(3)
            assert hip.hipStreamSynchronize(s) == 0
            after = ens.get_all_states()
            # back to an owned stream
            ens.set_stream(0)
            assert ens.stream not in (0, s.value)
            ens.forecast(2)
            ens.synchronize()
            final = ens.get_all_states()
        with lb.Context(cfg, M) as ref:
            _run(lb, cfg, M, yo, hipbuf, ref)
            ref.forecast(3)
            ref.synchronize()
            np.testing.assert_array_equal(after, ref.get_all_states())
            ref.forecast(2)
            ref.synchronize()
            np.testing.assert_array_equal(final, ref.get_all_states())
        print("  caller stream == own stream bit for bit; back to an owned stream", flush=True)
    finally:
        assert hip.hipStreamDestroy(s) == 0
