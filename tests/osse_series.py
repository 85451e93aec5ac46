"""The long OSSE of the benchmark configuration, shared by the golden-series
script (tests/golden/make_osse_series.py) and the GPU tests that compare with it.

Nature run: the baseline preset (256^2, tau = 0.503, Cs = 0.1, K4 spectrum with
u_rms = 0.05, nature seed 0) stepped by the FP64 oracle.  Every ``steps`` steps
the nature state is observed -- (u, v) every 8th node, Philox noise of sigma
0.01, counter = cycle index -- and the ensemble (members seeded 1000 + k)
assimilates those observations.  Ensemble statistics follow
ensemble_stat/main.cpp:108-177; the energy spectrum and vorticity follow the
reference's post-processing (postprocess/fourier.ipynb ``get_fft``,
src/io.cu:61-86).
"""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NATURE_SEED = 0
OBS_SEED = 2024
CYCLES = 100
STEPS = 100
SPECTRUM_EVERY = 1000  # nature steps between spectra


def golden_path(M: int) -> str:
    return os.path.join(GOLDEN, f"osse_256x256_M{M}_{CYCLES}cycles.json")


def load_golden(M: int) -> dict:
    with open(golden_path(M)) as fh:
        return json.load(fh)


def energy_shells(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """E(k) as postprocess/fourier.ipynb ``get_fft`` sums it: the coefficients
    F = (u_hat^2 + v_hat^2) / 2 of the normalised 2-D FFTs (complex squares, as
    written there), |F| summed over the quarter plane 0 <= i, j <= n/2 into shells
    k = int(sqrt(i^2 + j^2) + 0.5) <= n/2."""
    ny, nx = u.shape
    fu = np.fft.fft2(u) / nx / ny
    fv = np.fft.fft2(v) / nx / ny
    F = (fu ** 2 + fv ** 2) / 2
    kmax = nx // 2
    i, j = np.meshgrid(np.arange(kmax + 1), np.arange(kmax + 1), indexing="ij")
    k = (np.sqrt(i ** 2 + j ** 2) + 0.5).astype(int)
    sel = k <= kmax
    ek = np.zeros(kmax + 1)
    np.add.at(ek, k[sel], np.abs(F[: kmax + 1, : kmax + 1][sel]))
    return ek


def power_shells(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """The phase-invariant counterpart of energy_shells: (|u_hat|^2 + |v_hat|^2) / 2
    summed over the same quarter plane and shells."""
    ny, nx = u.shape
    fu = np.fft.fft2(u) / nx / ny
    fv = np.fft.fft2(v) / nx / ny
    F = (np.abs(fu) ** 2 + np.abs(fv) ** 2) / 2
    kmax = nx // 2
    i, j = np.meshgrid(np.arange(kmax + 1), np.arange(kmax + 1), indexing="ij")
    k = (np.sqrt(i ** 2 + j ** 2) + 0.5).astype(int)
    sel = k <= kmax
    ek = np.zeros(kmax + 1)
    np.add.at(ek, k[sel], F[: kmax + 1, : kmax + 1][sel])
    return ek


def vorticity(u: np.ndarray, v: np.ndarray, dx: float) -> np.ndarray:
    """src/io.cu:61-86: central differences on the periodic grid, v_x - u_y."""
    uy = (np.roll(u, -1, axis=0) - np.roll(u, 1, axis=0)) / (2 * dx)
    vx = (np.roll(v, -1, axis=1) - np.roll(v, 1, axis=1)) / (2 * dx)
    return vx - uy


def nature_config(lb):
    return lb.preset("baseline")


def nature_run(lb, orc, cycles: int = CYCLES, steps: int = STEPS):
    """The FP64 nature run.  Returns (truth, yo, spectra): truth[c] = (rho, u, v) and
    yo[c] the noisy observation vector after cycle c's ``steps`` steps; spectra maps
    the nature step (0, 1000, ...) to (E(k), vorticity rms, u rms)."""
    cfg = nature_config(lb)
    n = cfg.nx
    u0, v0 = orc.initial_velocity(cfg, NATURE_SEED)
    f = orc.feq(cfg, np.full((n, n), cfg.rho_ref), u0, v0)[None]
    truth, yo, spectra = [], [], {}

    def spectrum(step, u, v):
        spectra[step] = (energy_shells(u, v), float(np.sqrt(np.mean(vorticity(u, v, cfg.dx) ** 2))),
                         float(np.sqrt(np.mean(u * u + v * v))))

    spectrum(0, u0, v0)
    for c in range(cycles):
        for s in range(steps):
            f = orc.lbm_step(cfg, f, c * steps + s)
        r, u, v = orc.macro(cfg, f[0])
        truth.append((r, u, v))
        hx = orc.obs_moments(cfg, f)
        yo.append(orc.observe(cfg, hx[:, 0], OBS_SEED, c, True))
        step = (c + 1) * steps
        if step % SPECTRUM_EVERY == 0:
            spectrum(step, u, v)
    return truth, yo, spectra


def truth_checksum(truth) -> list:
    """Per cycle: sums of u^2 and v^2 of the nature state (reproduction check)."""
    return [[float(np.sum(u * u)), float(np.sum(v * v))] for _, u, v in truth]
