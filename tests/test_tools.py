"""Host tools of the OSSE front-end (CPU): the offline statistics of
ensemble_stat/main.cpp and bindiff/main.cpp restated with numpy, the raw file
layout of src/io.cu, and the observation-noise stream of the OBSERVE mode."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "lbm2d-letkf_amd", "bin")
STAT = os.path.join(BIN, "lbmletkf_stat")
DRIVER_DIR = os.path.join(ROOT, "lbm2d-letkf_amd", "driver")

pytestmark = pytest.mark.skipif(not os.path.exists(STAT), reason="lbmletkf_stat is not built (run build())")


def _write(path, a):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.ascontiguousarray(a, dtype=np.float32).tofile(path)


def _step(d, var, t):
    return os.path.join(d, f"{var}_step{t:010d}.dat")


def _make_osse_files(tmp, nx, ny, members, steps, seed=0):
    rng = np.random.default_rng(seed)
    nat = os.path.join(tmp, "nature")
    cal = os.path.join(tmp, "calc")
    truth, ens = {}, {}
    for t in steps:
        for var, scale, off in (("u", 0.05, 0.0), ("v", 0.05, 0.0), ("rho", 0.01, 1.0)):
            x = (off + scale * rng.standard_normal((ny, nx))).astype(np.float32)
            truth[(var, t)] = x
            _write(_step(os.path.join(nat, "ens0000"), var, t), x)
            for k in range(members):
                y = (x + scale * 0.3 * rng.standard_normal((ny, nx))).astype(np.float32)
                ens[(var, t, k)] = y
                _write(_step(os.path.join(cal, f"ens{k:04d}"), var, t), y)
    return nat, cal, truth, ens


def _numpy_stats(truth, ens, t, members, count):
    """ensemble_stat/main.cpp:108-177 with numpy: float32 mean / biased sd buffers,
    sums of the errors over `count` leading values."""
    def mean_sd(var):
        x = np.stack([ens[(var, t, k)].ravel()[:count] for k in range(members)]).astype(np.float64)
        m = x.mean(0)
        ms = (x * x).mean(0)
        sd = np.sqrt(np.maximum(ms - m * m, 0.0))
        return m.astype(np.float32).astype(np.float64), sd.astype(np.float32).astype(np.float64)

    mu, su = mean_sd("u")
    mv, sv = mean_sd("v")
    mr, sr = mean_sd("rho")
    tu = truth[("u", t)].ravel()[:count].astype(np.float64)
    tv = truth[("v", t)].ravel()[:count].astype(np.float64)
    tr = truth[("rho", t)].ravel()[:count].astype(np.float64)
    spread = math.sqrt((su**2 + sv**2).sum() / (2 * count))
    rmse = math.sqrt(((mu - tu) ** 2 + (mv - tv) ** 2).sum() / (2 * count))
    return spread, rmse, math.sqrt((sr**2).sum() / count), math.sqrt(((mr - tr) ** 2).sum() / count)


def _parse_csv(text):
    rows = {}
    for line in text.strip().splitlines():
        parts = [p.strip() for p in line.split(",")]
        rows[int(parts[0])] = [float(p) for p in parts[1:]]
    return rows


@pytest.mark.parametrize("quirk", [False, True])
def test_ensemble_stat_matches_numpy(tmp_path, quirk):
    nx, ny, members, steps = (160, 128, 4, [0, 10]) if quirk else (12, 10, 5, [0, 10, 20])
    nat, cal, truth, ens = _make_osse_files(str(tmp_path), nx, ny, members, steps)
    cmd = [STAT, "ensemble", nat, cal, str(max(steps)), "--members", str(members), "--nx", str(nx), "--ny", str(ny)]
    if quirk:
        cmd.append("--grid-quirk")
    p = subprocess.run(cmd, capture_output=True, text=True, check=True)
    rows = _parse_csv(p.stderr)
    assert sorted(rows) == steps
    count = min(nx * ny, 128 * 128) if quirk else nx * ny
    for t in steps:
        np.testing.assert_allclose(rows[t], _numpy_stats(truth, ens, t, members, count), rtol=1e-6)
    assert "t_step, spread, rmse, rho_spread, rho_rmse, max(debug) =" in p.stdout


def test_ensemble_stat_first_member_offset(tmp_path):
    nat, cal, truth, ens = _make_osse_files(str(tmp_path), 8, 8, 3, [0])
    # members ens0001..ens0002 only
    p = subprocess.run([STAT, "ensemble", nat, cal, "0", "--members", "2", "--first-member", "1", "--nx", "8"],
                       capture_output=True, text=True, check=True)
    shifted = {(v, t, k - 1): a for (v, t, k), a in ens.items() if k >= 1}
    np.testing.assert_allclose(_parse_csv(p.stderr)[0], _numpy_stats(truth, shifted, 0, 2, 64), rtol=1e-6)


def test_ensemble_stat_missing_member_fails(tmp_path):
    nat, cal, _, _ = _make_osse_files(str(tmp_path), 8, 8, 2, [0])
    p = subprocess.run([STAT, "ensemble", nat, cal, "0", "--members", "3", "--nx", "8"], capture_output=True, text=True)
    assert p.returncode == 1 and "could not open file" in p.stderr


def test_bindiff_matches_numpy(tmp_path):
    rng = np.random.default_rng(4)
    d0, d1 = str(tmp_path / "a"), str(tmp_path / "b")
    ref = {}
    for t in (0, 5, 10, 15, 20, 25):
        sse, cnt = 0.0, 0
        for var in ("u", "v"):
            a = rng.standard_normal(70).astype(np.float32)
            b = (a + 0.1 * rng.standard_normal(70)).astype(np.float32)
            _write(_step(d0, var, t), a)
            _write(_step(d1, var, t), b)
            e = (a - b).astype(np.float32).astype(np.float64)  # the difference is a float
            sse += (e * e).sum()
            cnt += a.size
        ref[t] = math.sqrt(sse / cnt)
    p = subprocess.run([STAT, "bindiff", d0, d1, "25"], capture_output=True, text=True, check=True)
    out = {int(l.split(",")[0]): float(l.split(",")[1]) for l in p.stdout.strip().splitlines()}
    assert sorted(out) == [0, 10, 20]  # every 10th output step, as the reference's skip
    for t in out:
        assert out[t] == pytest.approx(ref[t], rel=1e-6)
    p = subprocess.run([STAT, "bindiff", d0, d1, "25", "--skip", "5", "--vars", "u"], capture_output=True,
                       text=True, check=True)
    assert [int(l.split(",")[0]) for l in p.stdout.strip().splitlines()] == [0, 5, 10, 15, 20, 25]


def _libstdcxx_normals(seed, n):
    """std::normal_distribution<float>(0,1) over std::mt19937(seed), restated:
    Marsaglia's polar method as libstdc++ implements it (bits/random.tcc) -- two
    canonical floats x, y -> 2c - 1, reject r2 > 1 or r2 == 0, return y*m and keep
    x*m for the next call, m = sqrt(-2 log r2 / r2), all in float.  Each canonical
    float takes one 32-bit draw, float(draw) / 2^32 clamped below 1."""
    rs = np.random.RandomState(seed)  # legacy seeding == std::mt19937(seed)
    raw = rs.randint(0, 2**32, size=8 * n + 64, dtype=np.uint64)
    pos = 0

    def canon():
        nonlocal pos
        r = np.float32(raw[pos]) / np.float32(2.0**32)
        pos += 1
        return r if r < 1 else np.nextafter(np.float32(1), np.float32(0))

    out = []
    while len(out) < n:
        while True:
            x = np.float32(2) * canon() - np.float32(1)
            y = np.float32(2) * canon() - np.float32(1)
            r2 = np.float32(x * x + y * y)
            if not (r2 > 1 or r2 == 0):
                break
        m = np.sqrt(np.float32(-2) * np.log(r2) / r2, dtype=np.float32)
        out += [y * m, x * m]
    return np.array(out[:n], dtype=np.float64)


def test_observation_noise_stream(tmp_path):
    # the OBSERVE-mode noise of the driver (driver/datio.hpp add_obs_noise) against the
    # restated std::mt19937 + std::normal_distribution<float> stream: order rho, u, v per node
    src = tmp_path / "noise.cpp"
    src.write_text(
        '#include <cstdio>\n#include <random>\n#include <vector>\n#include "datio.hpp"\n'
        "int main() {\n"
        "    const size_t n = 500;\n"
        "    std::vector<float> z(n, 0.f), r(n, 1.f), ro(n), uo(n), vo(n);\n"
        "    std::mt19937 eng(334);\n"
        "    std::normal_distribution<float> nd(0.f, 1.f);\n"
        "    lbmletkf_io::add_obs_noise(eng, nd, r.data(), z.data(), z.data(), n, 0.5f, 2.f, ro.data(), uo.data(), vo.data());\n"
        '    for (size_t i = 0; i < n; ++i) std::printf("%.9g %.9g %.9g\\n", ro[i], uo[i], vo[i]);\n'
        "}\n")
    exe = tmp_path / "noise"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{DRIVER_DIR}", str(src), "-o", str(exe)], check=True)
    out = np.loadtxt(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    z = _libstdcxx_normals(334, 3 * 500).reshape(500, 3)
    # a few ulp: logf/sqrtf of the two libraries may round differently; a wrong draw
    # order or pairing would differ by O(1)
    np.testing.assert_allclose(out[:, 0], (np.float32(1) + np.float32(0.5) * z[:, 0].astype(np.float32)), rtol=6e-7)
    np.testing.assert_allclose(out[:, 1], 2 * z[:, 1], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(out[:, 2], 2 * z[:, 2], rtol=2e-6, atol=1e-6)
    assert abs(z.mean()) < 0.1 and abs(z.std() - 1) < 0.05
