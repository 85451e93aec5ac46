"""Host logic of bench.py (CPU): the one-GPU contract and the committed HBM-traffic
summary that the bench line's roofline reads."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_host", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "not supported" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_steps_per_launch_of_the_shape_codes():
    b = _bench()
    assert b.lbm_steps_per_launch([-216, 32]) == 3
    assert b.lbm_steps_per_launch([-232, 32]) == 3
    assert b.lbm_steps_per_launch([-116, 32]) == 2
    assert b.lbm_steps_per_launch([-132, 16]) == 2
    assert b.lbm_steps_per_launch([-16, 32]) == 1
    assert b.lbm_steps_per_launch([2, 8]) == 1
    assert b.lbm_steps_per_launch(None) == 1


def test_traffic_summary_holds_the_benchmark_shapes():
    """The rows the roofline looks up at 256^2 x 64 for the shapes the autotuner
    picks -- the fused analysis at chains of 16 and 32 (4,096 and 2,048 workgroups),
    three-step strips of 16 rows (1,024 workgroups) -- are in the committed
    summary, with byte counts no smaller than what the kernels must move."""
    b = _bench()
    nodes, M = 256 * 256, 64
    algorithmic = b.LBM_BYTES_PER_NODE_UPDATE * nodes * M  # the state read and written once
    for kernel, wg in (("k_mixed_analysis<64, 3>", nodes // 16), ("k_mixed_analysis<64, 3>", nodes // 32),
                       ("k_lbm_rows3<16, 256>", (256 // 16) * M)):
        t = b.measured_traffic(kernel, wg)
        assert t is not None, (kernel, wg)
        assert algorithmic <= t <= 2 * algorithmic, (kernel, wg, t)
    assert b.measured_traffic("k_mixed_analysis<64, 3>", 12345) is None
