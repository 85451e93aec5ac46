"""The OSSE front-end end to end on the GPU: observe -> da_letkf / free with the
driver, then the offline statistics tool, in the reference's io/ layout."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "lbm2d-letkf_amd", "bin")
RUN = os.path.join(BIN, "lbmletkf_run")
STAT = os.path.join(BIN, "lbmletkf_stat")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(RUN), reason="lbmletkf_run is not built (run build())")]

N = 64
COMMON = ["--preset", "baseline", "--nx", str(N), "--obs-stride", "4", "--loc-scale", "6", "--loc-cutoff", "14",
          "--iiter", "10", "--daprune", "5", "--ioprune", "5", "--outputs", "30"]


def _run(args, timeout=120):
    print("+", " ".join(os.path.basename(a) if i == 0 else a for i, a in enumerate(args)), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(p.stdout[-1500:], p.stderr[-1500:], flush=True)
    assert p.returncode == 0, p.stderr
    return p


def _read(prefix, kind, ens, var, step):
    path = os.path.join(prefix, kind, f"ens{ens:04d}", f"{var}_step{step:010d}.dat")
    return np.fromfile(path, dtype=np.float32).reshape(N, N)


def test_osse_front_end(tmp_path):
    io = str(tmp_path / "io")
    _run([RUN, "observe", "--prefix", io] + COMMON)
    cfg_sigma = 0.2 * 0.05  # baseline preset: sigma_u = 0.2 u_rms
    for step in (0, 15, 30):
        nat_u = _read(io, "nature", 0, "u", step)
        obs_u = _read(io, "observed", 0, "u", step)
        on = np.zeros((N, N), bool)
        on[::4, ::4] = True
        assert np.isnan(obs_u[~on]).all() and np.isfinite(obs_u[on]).all()
        err = (obs_u - nat_u)[on]
        assert abs(err.std() / cfg_sigma - 1) < 0.2, err.std()
        assert np.isfinite(_read(io, "nature", 0, "vor", step)).all()

    # the same 8-member ensemble (ids 1..8) with and without analyses
    free = str(tmp_path / "free")
    os.makedirs(free)
    os.symlink(os.path.join(io, "observed"), os.path.join(free, "observed"))
    _run([RUN, "da_letkf", "--prefix", io, "--members", "8", "--stats"] + COMMON)
    _run([RUN, "free", "--prefix", free, "--members", "8"] + COMMON)
    assert os.path.exists(os.path.join(io, "calc", "ens_vor_std_step0000000030.dat"))

    def stats(prefix):
        p = _run([STAT, "ensemble", os.path.join(io, "nature"), os.path.join(prefix, "calc"), "30", "--members", "8",
                  "--first-member", "1", "--nx", str(N), "--skip", "5"])
        rows = {}
        for line in p.stderr.strip().splitlines():
            v = [float(x) for x in line.split(",")]
            rows[int(v[0])] = v[1:]
        return rows

    da, fr = stats(io), stats(free)
    assert sorted(da) == [0, 5, 10, 15, 20, 25, 30]
    # identical before the first analysis, then the analyses pull the mean to the truth
    np.testing.assert_allclose(da[0], fr[0], rtol=1e-6)
    print("rmse da  :", [round(da[t][1], 6) for t in sorted(da)])
    print("rmse free:", [round(fr[t][1], 6) for t in sorted(fr)])
    assert da[30][1] < 0.7 * fr[30][1]
    # every member file exists for every output step
    for k in range(1, 9):
        for step in range(0, 31, 5):
            assert os.path.getsize(os.path.join(io, "calc", f"ens{k:04d}", f"u_step{step:010d}.dat")) == 4 * N * N


def test_driver_rejects_bad_arguments(tmp_path):
    p = subprocess.run([RUN, "bogus"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "unknown mode" in p.stderr
    # analyses need the observation files
    p = subprocess.run([RUN, "da_letkf", "--prefix", str(tmp_path / "none"), "--members", "4"] + COMMON,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 1 and "could not open file" in p.stderr
