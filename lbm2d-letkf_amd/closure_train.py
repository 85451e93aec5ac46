"""Training of the learned eddy-viscosity closure on coarse-grained fine runs.

A fine run of the same LBM -- the grid refined by ``factor`` at the same lattice
speed c (so dx and dt both shrink by ``factor``), the same physical viscosity and
the Smagorinsky model off -- is box-filtered back to the coarse grid every
``every`` coarse steps.  Each snapshot gives one training sample per coarse node:
the 3x3 stencil of the filtered strain and the subgrid stress
tau_ij = <u_i u_j> - <u_i><u_j> of the filter.  The network is trained on the GPU
by the library's own kernels (forward, backward, Adam) through
``lbmletkf.Trainer``; the result is a parameter vector for
``Context.set_closure``.

Command line:  python closure_train.py --preset baseline --out closure.npy
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
import lbmletkf as lb  # noqa: E402


def fine_config(coarse: lb.Config, factor: int) -> lb.Config:
    """The coarse configuration refined by factor: same c, dx/factor,
    dt/factor, same nu, no Smagorinsky term; friction per step scaled with dt."""
    cfg = coarse.copy()
    cfg.nx, cfg.ny = coarse.nx * factor, coarse.ny * factor
    cfg.dx = coarse.dx / factor
    cfg.dt = coarse.dt / factor
    cfg.cs_smag = 0.0
    cfg.friction_rate = coarse.friction_rate / factor
    cfg.closure_kind = lb.CLOSURE_SMAGORINSKY
    cfg.variant = lb.VARIANT_MIXED
    return cfg


def collect(trainer: lb.Trainer, coarse: lb.Config, *, factor: int = 4, seeds=(7,), n_snap: int = 4,
            every: int = 50, spinup: int = 200) -> None:
    """Append snapshots of fine runs to the trainer: one run per seeded spectrum,
    all stepped together as the members of one context; after `spinup` coarse
    steps, n_snap snapshots `every` coarse steps apart, each giving one sample per
    coarse node of every run."""
    fine = fine_config(coarse, factor)
    with lb.Context(fine, len(seeds)) as ctx:
        for k, sd in enumerate(seeds):
            ctx.init_member(k, sd)
        ctx.forecast(spinup * factor)
        for s in range(n_snap):
            for k in range(len(seeds)):
                trainer.add_samples(ctx, k, factor)
            if s + 1 < n_snap:
                ctx.forecast(every * factor)


def train_closure(coarse: lb.Config, *, factor: int = 4, seeds=tuple(range(1, 21)), n_snap: int = 4, every: int = 50,
                  spinup: int = 200, batch: int = 4096, steps: int = 3000, lr: float = 2e-3, init_seed: int = 1,
                  params: np.ndarray | None = None):
    """Collect snapshots of fine runs (20 seeded spectra by default) and train from
    init_params(init_seed) (or from params).  Returns (params, loss history)."""
    n_nodes = coarse.nx * coarse.ny
    with lb.Trainer(coarse, n_nodes * n_snap * len(seeds)) as tr:
        collect(tr, coarse, factor=factor, seeds=seeds, n_snap=n_snap, every=every, spinup=spinup)
        tr.set_params(lb.closure_init_params(init_seed) if params is None else params)
        hist = tr.train(batch, steps, lr)
        return tr.get_params(), hist


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--preset", default="baseline")
    ap.add_argument("--factor", type=int, default=4)
    ap.add_argument("--spectra", type=int, default=20, help="fine runs from seeds 1..N")
    ap.add_argument("--snapshots", type=int, default=4, help="snapshots per fine run")
    ap.add_argument("--every", type=int, default=50, help="coarse steps between snapshots")
    ap.add_argument("--spinup", type=int, default=200, help="coarse steps before the first snapshot")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="closure.npy")
    a = ap.parse_args(argv)
    cfg = lb.preset(a.preset)
    p, hist = train_closure(cfg, factor=a.factor, seeds=tuple(range(1, a.spectra + 1)), n_snap=a.snapshots,
                            every=a.every, spinup=a.spinup, batch=a.batch, steps=a.steps, lr=a.lr)
    np.save(a.out, p)
    k = max(1, len(hist) // 10)
    print(f"loss: first {hist[:k].mean():.4f}  last {hist[-k:].mean():.4f}  -> {a.out}")


if __name__ == "__main__":
    main()
