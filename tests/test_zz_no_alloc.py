"""lbmletkf_ensemble_stats_noalloc -- the partial sums in the library's static
device array, no allocation -- against lbmletkf_ensemble_stats with the truth on
the device: the same bits on two calls in a row, the second on a smaller grid and
ensemble (it reuses the slots the first wrote), and a clear error beyond the
array's grid limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import small_cfg

pytestmark = pytest.mark.gpu


def _noalloc(lb, ens, d):
    out = np.empty(4)
    st = lb.lib.lbmletkf_ensemble_stats_noalloc(ens.h, *[C.c_void_p(p) for p in d],
                                               out.ctypes.data_as(C.c_void_p))
    return st, out


def _pair(lb, hipbuf, n, M, seed):
    cfg = small_cfg(lb, "baseline", n=n)
    with lb.Context(cfg, M) as ens:
        for k in range(M):
            ens.init_member(k, 1000 + 37 * seed + k)
        ens.forecast(3)
        rng = np.random.default_rng(seed)
        truth = (1 + 1e-3 * rng.standard_normal((n, n)), 0.05 * rng.standard_normal((n, n)),
                 0.05 * rng.standard_normal((n, n)))
        d = [hipbuf.from_host(np.ascontiguousarray(x, np.float32)) for x in truth]
        ref = ens.ensemble_stats_device(*d)
        st, out = _noalloc(lb, ens, d)
        assert st == 0, lb.lib.lbmletkf_last_error(ens.h)
    return ref, out


def test_noalloc_stats_equal_the_allocating_entry_point(lb, hipbuf):
    r1, o1 = _pair(lb, hipbuf, 256, 16, 1)
    r2, o2 = _pair(lb, hipbuf, 64, 8, 2)
    print(f"256^2 M=16: {o1} (allocating {r1}); 64^2 M=8: {o2} (allocating {r2})")
    np.testing.assert_array_equal(o1, r1)
    np.testing.assert_array_equal(o2, r2)
    assert not np.array_equal(o1, o2)


def test_noalloc_stats_reject_grids_beyond_the_array(lb, hipbuf):
    n = 4160  # 4160^2 = 17,305,600 nodes > 2^24
    cfg = lb.preset("baseline")
    cfg.nx = cfg.ny = n
    with lb.Context(cfg, 2) as ens:
        # full-size truth, although the call must return before any launch
        d = [hipbuf.from_host(np.zeros(n * n, np.float32)) for _ in range(3)]
        st, _ = _noalloc(lb, ens, d)
        msg = lb.lib.lbmletkf_last_error(ens.h)
    assert st != 0
    assert b"too large" in msg
