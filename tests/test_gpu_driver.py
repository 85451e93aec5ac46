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
    # 8 members on 64^2 with (u,v) every 4th node: the analyses cut the error of the
    # free ensemble by about a fifth after three cycles
    assert da[30][1] < 0.85 * fr[30][1]
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


def _read_matrix(path, rows, cols, batch):
    """A matrix file of the driver: HDF5 dataset "data" (dims rows, cols, batch) or
    the raw .bin twin; returns [batch][cols][rows] (column-major per matrix)."""
    import ctypes as C

    is_f = ".float_" in os.path.basename(path)
    if path.endswith(".bin"):
        a = np.fromfile(path, dtype=np.float32 if is_f else np.float64)
    else:
        hl = C.CDLL("/opt/conda/lib/libhdf5_hl.so")
        h5 = C.CDLL("/opt/conda/lib/libhdf5.so")
        h5.H5Fopen.restype = C.c_int64
        h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
        h5.H5Fclose.argtypes = [C.c_int64]
        fid = h5.H5Fopen(path.encode(), 0, 0)
        assert fid > 0, path
        dims = (C.c_ulonglong * 3)()
        cls, tsize = C.c_int(), C.c_size_t()
        hl.H5LTget_dataset_info.argtypes = [C.c_int64, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p]
        assert hl.H5LTget_dataset_info(fid, b"data", dims, C.byref(cls), C.byref(tsize)) >= 0
        assert list(dims) == [rows, cols, batch], list(dims)
        a = np.empty(rows * cols * batch, np.float32 if is_f else np.float64)
        fn = hl.H5LTread_dataset_float if is_f else hl.H5LTread_dataset_double
        fn.argtypes = [C.c_int64, C.c_char_p, C.c_void_p]
        assert fn(fid, b"data", a.ctypes.data_as(C.c_void_p)) >= 0
        h5.H5Fclose(fid)
    return a.reshape(batch, cols, rows)


def test_savemat_letkf_matrices(tmp_path, lb, orc):
    io = str(tmp_path / "io")
    short = [a for a in COMMON]
    short[short.index("--outputs") + 1] = "5"
    _run([RUN, "observe", "--prefix", io] + short)
    M, g0, n, t = 8, 100, 300, 50
    _run([RUN, "da_letkf", "--prefix", io, "--members", str(M), "--savemat", f"{g0}:{n}"] + short)
    d = os.path.join(io, "letkf", "mat", "ens0")
    names = sorted(os.listdir(d))
    print(names)
    ext = names[0].rsplit(".", 1)[1]
    assert ext in ("h5", "bin") and len(names) == 14

    def rd(label, rows, typ):
        return _read_matrix(os.path.join(d, f"{label}.{typ}_{rows}_{M}_{n}.{t}.{ext}"), rows, M, n)

    Xf, dXf, Xa, AI = (rd(k, 9, "float").astype(np.float64) for k in ("_Xf", "_dXf", "_Xa", "_AI"))
    dWa, Wa, Z, Lam, Wma, Pa = (rd(k, M, "double") for k in ("_dWa", "_Wa", "_Z", "_Lambda", "_Wma", "_Pa"))
    # arrays are [batch][col][row]; transpose to [batch][row][col]
    T = lambda a: np.swapaxes(a, 1, 2)  # noqa: E731
    Xf, dXf, Xa, AI, dWa, Wa, Z, Lam, Wma, Pa = map(T, (Xf, dXf, Xa, AI, dWa, Wa, Z, Lam, Wma, Pa))
    np.testing.assert_allclose(dXf, Xf - Xf.mean(axis=2, keepdims=True), atol=1e-7)
    np.testing.assert_allclose(AI, Xa - Xf, atol=1e-7)
    np.testing.assert_allclose(Z, Wa - np.eye(M), atol=1e-15)
    np.testing.assert_allclose(Wa, dWa + Wma, atol=1e-15)
    np.testing.assert_allclose(dWa, np.swapaxes(dWa, 1, 2), atol=1e-12)
    # the LETKF update the kernels applied: Xa = xbar 1^T + dX Wa
    xa = Xf.mean(axis=2, keepdims=True) + dXf @ Wa
    np.testing.assert_allclose(Xa, xa, rtol=0, atol=2e-6 * np.abs(Xa).max())
    # W 1 = sqrt(beta) 1 (the perturbations keep a zero mean); eig(Pa) = 1 / eig(C)
    beta = 1.1
    # (dY 1 = 0 up to the FP32 rounding of the perturbations and the split-bf16 Gram)
    print("max |W 1 - sqrt(beta)|:", np.abs(dWa.sum(axis=2) - np.sqrt(beta)).max())
    np.testing.assert_allclose(dWa.sum(axis=2), np.sqrt(beta), rtol=1e-4)
    ev = np.diagonal(Lam, axis1=1, axis2=2)  # the solver's order (unsorted), that of the columns of _V
    np.testing.assert_allclose(np.sort(1 / np.linalg.eigvalsh(Pa), axis=1), np.sort(ev, axis=1), rtol=1e-8)

    # solve_evd side: _Q = C, _V its eigenvectors in the order of _Lambda, _dYf and _dyo
    # the local rows of dY and of the innovation in the gridpoint's gather order
    rows = int(next(f for f in names if f.startswith("_dYf.")).split("_")[2])
    Q, V = (T(rd(k, M, "double")) for k in ("_Q", "_V"))
    dYf = T(rd("_dYf", rows, "float")).astype(np.float64)
    dyo = _read_matrix(os.path.join(d, f"_dyo.float_{rows}_1_{n}.{t}.{ext}"), rows, 1, n)[:, 0, :]
    np.testing.assert_allclose(np.swapaxes(V, 1, 2) @ V, np.broadcast_to(np.eye(M), V.shape), atol=1e-5)
    np.testing.assert_allclose(V @ (ev[:, :, None] * np.swapaxes(V, 1, 2)), Q, rtol=0, atol=1e-5 * np.abs(Q).max())
    np.testing.assert_allclose(dYf.sum(axis=2), 0, atol=1e-6)
    # against the FP64 oracle: C and the transform (W, w) of each gridpoint formed from
    # the dumped dY and innovation with the oracle's local weights
    cfg = lb.preset("baseline")
    cfg.nx = cfg.ny = N
    cfg.obs_stride, cfg.loc_scale, cfg.loc_cutoff = 4, 6.0, 14.0
    nv = 2  # (u, v)
    rinv_u = 1.0 / cfg.obs_sigma_u ** 2
    worst = dict(C=0.0, W=0.0, w=0.0)
    for b in range(n):
        sites, wl = orc.local_obs(cfg, g0 + b)
        p = len(sites) * nv
        assert p <= rows and not dYf[b, p:].any()
        dY = dYf[b, :p]
        rinv = np.repeat(wl, nv) * rinv_u
        C_o = (M - 1) / beta * np.eye(M) + dY.T @ (rinv[:, None] * dY)
        W_o, w_o, _ = orc.letkf_transform(dY, rinv, dyo[b, :p].astype(np.float64), beta)
        worst["C"] = max(worst["C"], np.linalg.norm(Q[b] - C_o) / np.linalg.norm(C_o))
        worst["W"] = max(worst["W"], np.linalg.norm(dWa[b] - W_o) / np.linalg.norm(W_o))
        worst["w"] = max(worst["w"], np.linalg.norm(Wma[b][:, 0] - w_o) / max(np.linalg.norm(w_o), 1e-30))
    print("worst rel. difference from the oracle over the dumped gridpoints:", worst)
    assert worst["C"] < 1e-4 and worst["W"] < 1e-4 and worst["w"] < 1e-3
    # the states of the matrices are those of the run's own output at that step
    u_calc = _read(io, "calc", 1, "u", 5)
    assert np.isfinite(u_calc).all()
