"""DA-cycle benchmark of the LBM-LETKF assimilation cycle on one MI355X.

Workload: the 256^2 x M=64 OSSE (decaying D2Q9 BGK+Smagorinsky turbulence,
tau=0.503, Cs=0.1; (u,v) observed every 8th node with sigma=0.01; Gaussian
localization L=12 with a 43.8-cell cutoff; inflation 1.1).  One DA cycle = 100
LBM steps of all 64 members + one LETKF analysis at all 65,536 gridpoints.

The nature run and its observations for every cycle are generated first and kept
in HBM; the timed region covers K cycles of the ensemble only (the reference's
`forecast` + `DA` timers, src/main.cu:95-115).  The variants:
plain (per-member LBM launches, FP32 Gram, FP64 two-sided Jacobi, eager launches) at
the preset's eigensolver tolerance (1e-10) and at the reference's (syevj tolerance
1e-6, src/util/cu_matrix_eigensolver_wrapper.hpp:192,202), and mixed (all members'
LBM steps in one launch, one fused analysis kernel -- split-bf16 MFMA Gram, FP32
one-sided Jacobi, FP64 eigenvalue coefficients, FP32 MFMA update -- each cycle
replayed from a hipGraph).  `value` is the mixed variant's DA cycles/s.  The
analysis RMSE of each variant is taken at the observation time after the timed
cycles; a few further cycles with phase timing on give the per-kernel breakdown.

The mixed variant's launch shapes come from the autotuner (lbmletkf_autotune)
before timing.  The JSON line also carries the uneven batch (four OSSEs of
128^2..512^2 and M = 8..64 with random site densities, stepped and analysed by
shared launches; --no-uneven skips it), one LBM step with the learned closure
against the Smagorinsky step, and the mixed cycle with the learned closure in
every forecast step (variant mixed_mlp: Glorot-initialised weights, whose values
do not change the work of a step).

Prints ONE JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3      # published FP32 vector peak (counts packed FP32 issue)
FP32_UNPACKED_TFLOPS = 78.6   # the same without packed FP32, which this build never emits
LBM_BYTES_PER_NODE_UPDATE = 72  # 9 x 4 B read + 9 x 4 B written


def load_binding():
    spec = importlib.util.spec_from_file_location("lbmletkf", os.path.join(ROOT, "lbm2d-letkf_amd", "lbmletkf.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lbmletkf"] = mod
    spec.loader.exec_module(mod)
    return mod


class Hip:
    """The few HIP runtime calls the harness needs (device buffers, events)."""

    def __init__(self):
        self.l = C.CDLL("libamdhip64.so")
        self.l.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.l.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.l.hipEventSynchronize.argtypes = [C.c_void_p]
        self.l.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.l.hipFree.argtypes = [C.c_void_p]

    def ck(self, st, what):
        if st != 0:
            raise RuntimeError(f"{what} failed with HIP status {st}")

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ck(self.l.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        return p.value

    def free(self, p):
        self.l.hipFree(C.c_void_p(p))

    def d2h(self, dst: np.ndarray, src: int):
        self.ck(self.l.hipMemcpy(C.c_void_p(dst.ctypes.data), C.c_void_p(src), C.c_size_t(dst.nbytes), 2),
                "hipMemcpy")

    def sync(self):
        self.ck(self.l.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def event(self):
        e = C.c_void_p()
        self.ck(self.l.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e

    def record(self, e, stream):
        self.ck(self.l.hipEventRecord(e, C.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, a, b):
        self.ck(self.l.hipEventSynchronize(b), "hipEventSynchronize")
        ms = C.c_float()
        self.ck(self.l.hipEventElapsedTime(C.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value

    def device_name(self):
        props = C.create_string_buffer(4096)
        try:
            self.l.hipDeviceGetName(props, 256, 0)
            return props.value.decode(errors="replace")
        except Exception:
            return "unknown"


def bench_variant(lb, hip, cfg, variant, M, K, W, P, obs, truth, n_obs, log, repeats=1, label=None,
                  fixed_tuning=None, closure=None):
    """K timed DA cycles (after W untimed ones) through lbmletkf_cycle -- for the
    mixed variant the captured graph is replayed -- then the analysis RMSE at the
    next observation time; then repeats-1 further timed runs of K cycles (their
    median and spread are reported), then one untimed cycle and P more cycles with
    phase timing on (eager launches) for the per-kernel breakdown.  With `closure` (network parameters) the
    members step with the learned closure instead of Smagorinsky."""
    cfg = cfg.copy()
    cfg.variant = variant
    if closure is not None:
        cfg.closure_kind = lb.CLOSURE_MLP
    ens = lb.Context(cfg, M)
    if closure is not None:
        ens.set_closure(closure)
    for k in range(M):
        ens.init_member(k, 1000 + k)
    steps = cfg.da_interval
    nodes = cfg.nx * cfg.ny

    def yo(c):
        return obs + 4 * n_obs * c

    ens.forecast(steps)  # to the time of the first observation
    tuning = None
    name = label or ("plain" if variant == lb.VARIANT_PLAIN else "mixed")
    if variant == lb.VARIANT_MIXED:
        if fixed_tuning:
            ens.set_tuning(*fixed_tuning)
            log(f"[{name}] tuning fixed at {fixed_tuning}")
        else:
            log(f"[{name}] autotune: " + ens.autotune())
        tuning = list(ens.get_tuning())
    for c in range(W):
        ens.cycle_device(yo(c))
    ens.synchronize()

    def timed(c0):
        e0, e1 = hip.event(), hip.event()
        hip.sync()
        t0 = time.perf_counter()
        hip.record(e0, ens.stream)
        for c in range(c0, c0 + K):
            ens.cycle_device(yo(c))
        hip.record(e1, ens.stream)
        ens.synchronize()
        hip.sync()
        return time.perf_counter() - t0, hip.elapsed_ms(e0, e1)

    walls, evs = [], []
    w, e = timed(W)
    walls.append(w)
    evs.append(e)
    # analysis at the next observation time, against the nature run there
    ca = W + K
    ens.analysis_device(yo(ca))
    tp = truth + 4 * 3 * nodes * ca
    stats = ens.ensemble_stats_device(tp, tp + 4 * nodes, tp + 8 * nodes)
    final_state = ens.get_all_states()
    c = ca + 1
    for _ in range(repeats - 1):
        ens.forecast(steps)
        w, e = timed(c)
        walls.append(w)
        evs.append(e)
        c += K
    # one more cycle before the phase-timed pass: the copies and statistics above
    # leave the GPU idle, and the first kernels after an idle spell run at lower clocks
    ens.forecast(steps)
    ens.analysis_device(yo(c))
    c += 1
    ens.profile_enable(True)
    for cc in range(c, c + P):
        ens.forecast(steps)
        ens.analysis_device(yo(cc))
    ens.synchronize()
    phases = {ph: ens.profile_read(ph) for ph in ("forecast", "obs", "gram", "solve", "finish", "solve_fp64")}
    ens.close()
    ws = sorted(walls)
    wall = ws[len(ws) // 2]
    q1, q3 = (np.percentile(ws, 25), np.percentile(ws, 75)) if len(ws) > 1 else (ws[0], ws[0])
    log(f"[{name}] {K} cycles x {len(ws)} run(s): median {wall:.3f} s wall ({np.median(evs) / 1e3:.3f} s events): "
        f"{K / wall:.2f} cycles/s; analysis rmse {stats[0]:.4e} spread {stats[1]:.4e}")
    for ph, (sec, n, wk) in phases.items():
        log(f"    {ph:10s} {sec * 1e3 / P:9.3f} ms/cycle  launches {n}  work {wk}   (phase-timed pass, {P} cycles)")
    return dict(wall=wall, walls=walls, iqr_cycles_per_s=[K / q3, K / q1], event_s=float(np.median(evs)) / 1e3,
                phases=phases, stats=stats, steps=steps, state=final_state, yo_index=ca, prof_cycles=P,
                variant="plain" if variant == lb.VARIANT_PLAIN else "mixed", tuning=tuning, K=K,
                jacobi_tol=cfg.jacobi_tol, closure="smagorinsky" if closure is None else "mlp")


def rmse_probe(lb, cfg, M, n_cycles, obs, truth, n_obs):
    """Analysis statistics of the mixed variant after n_cycles DA cycles, at the same
    observation time as a shorter run (the plain variants' RMSE is compared there)."""
    cfg = cfg.copy()
    cfg.variant = lb.VARIANT_MIXED
    nodes = cfg.nx * cfg.ny
    with lb.Context(cfg, M) as ens:
        for k in range(M):
            ens.init_member(k, 1000 + k)
        ens.forecast(cfg.da_interval)
        for c in range(n_cycles):
            ens.cycle_device(obs + 4 * n_obs * c)
        ens.analysis_device(obs + 4 * n_obs * n_cycles)
        tp = truth + 4 * 3 * nodes * n_cycles
        return ens.ensemble_stats_device(tp, tp + 4 * nodes, tp + 8 * nodes)


def measured_traffic(kernel, workgroups):
    """HBM MB per launch of one kernel shape, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (a separate profiled run of this benchmark)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_256x256_M64.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0] == kernel and cells[1] == str(workgroups):
            return float(cells[6]) * 1e6
    return None


def lbm_steps_per_launch(tune):
    """Steps one launch of the tuned LBM shape takes: three for the three-step row
    strips (shape code -(200 + h)), two for the two-step strips (-(100 + h)); the
    benchmark flow has no body force."""
    if tune and tune[0] <= -200:
        return 3
    return 2 if tune and tune[0] <= -100 else 1


def kernel_table(r, M, nodes):
    """Average per-launch duration and algorithmic rate of the main kernels, from the
    phase-timed cycles (HIP events on the context's stream around each phase)."""
    out = {}
    P = r["prof_cycles"]
    sec, n, work = r["phases"]["forecast"]
    n_steps = r["steps"] * P
    spl = lbm_steps_per_launch(r.get("tuning"))
    avg = sec / n_steps  # one forecast step (all launches of the step)
    # algorithmic HBM bytes of one launch: every population read once and written
    # once, whether the launch takes one step or two
    out["lbm_step"] = dict(avg_ms=avg * 1e3, steps_per_launch=spl, avg_launch_ms=avg * spl * 1e3,
                           achieved=LBM_BYTES_PER_NODE_UPDATE * M * nodes / (avg * spl) / 1e9, unit="GB/s",
                           node_updates_per_s=M * nodes / avg,
                           bytes_of_single_steps_GBs=LBM_BYTES_PER_NODE_UPDATE * M * nodes / avg / 1e9,
                           peak=HBM_PEAK_GBS, bound="hbm", share=sec)
    sec, n, sweeps = r["phases"]["solve"]
    sweeps_gp = sweeps / (nodes * P) if P else 0.0
    if r["variant"] == "mixed":
        # the fused analysis kernel is bound by its FP32 vector ALU (the one-sided
        # sweeps).  Algorithmic flops of a one-sided sweep: N-1 rounds x N columns x
        # (N FMA for the pair inner product + N FMA for the update) ~ 4 N^3 (N = M = 64).
        # The two-sided count (M(M-1)/2 rotations x 18 M per sweep) is given beside it.
        t = sec / max(n, 1)
        flops_own = 4.0 * M**3 * sweeps_gp * nodes
        flops_2s = M * (M - 1) / 2 * 18.0 * M * sweeps_gp * nodes
        out["letkf_analysis"] = dict(
            avg_ms=t * 1e3, achieved=flops_own / t / 1e12 if t else 0.0, unit="TFLOP/s",
            peak=FP32_PEAK_TFLOPS, bound="valu",
            kernel="k_mixed_analysis<64, 3> (gather + split-bf16 MFMA Gram + FP32 one-sided sweeps + "
                   "FP64 coefficients + FP32 MFMA update)",
            flop_count="one-sided Jacobi: 4 N^3 per sweep per gridpoint x measured sweeps",
            peak_unpacked_fp32=FP32_UNPACKED_TFLOPS,
            frac_of_unpacked_fp32=(flops_own / t / 1e12 / FP32_UNPACKED_TFLOPS) if t else 0.0,
            achieved_two_sided_count=flops_2s / t / 1e12 if t else 0.0,
            frac_two_sided_count=(flops_2s / t / 1e12 / FP32_PEAK_TFLOPS) if t else 0.0,
            sweeps_per_gridpoint=sweeps_gp, fp64_sweeps_per_gridpoint=0.0, share=sec)
    else:
        flops = 9.0 * M**3 * sweeps  # two-sided rotation update (12M) + V update (6M) per rotation
        out["letkf_solve"] = dict(avg_ms=sec / max(n, 1) * 1e3, achieved=flops / sec / 1e12 if sec else 0.0,
                                  unit="TFLOP/s", peak=FP64_PEAK_TFLOPS, bound="valu",
                                  engine="fp64 vector ALU + LDS (two-sided Jacobi)",
                                  sweeps_per_gridpoint=sweeps_gp, share=sec)
    sec, n, pobs = r["phases"]["gram"]
    if n:
        gflops = 2.0 * M * M * pobs
        out["letkf_gram"] = dict(avg_ms=sec / max(n, 1) * 1e3, achieved=gflops / sec / 1e12 if sec else 0.0,
                                 unit="TFLOP/s", peak=FP32_PEAK_TFLOPS, bound="valu",
                                 engine="fp32 vector ALU", local_obs_per_gridpoint=pobs / (nodes * P), share=sec)
    sec, n, _ = r["phases"]["obs"]
    out["obs_prepare"] = dict(avg_ms=sec / max(n, 1) * 1e3, share=sec)
    total = sum(v["share"] for v in out.values())
    for v in out.values():
        v["share"] = v["share"] / total if total else 0.0
    return out


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(cfg, f_state, yo, log):
    """The FP64 oracle (OpenMP on every host core, LAPACK dsyev for the local
    eigensolves) over one whole DA cycle: da_interval LBM steps of all members, then
    the analysis at every gridpoint, starting from the GPU run's final ensemble and
    its last observations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as orc

    nproc, model = cpu_info()
    orc.set_threads(nproc)
    try:
        from threadpoolctl import threadpool_limits
        limits = threadpool_limits(1)  # one BLAS thread per OpenMP thread
    except Exception:
        limits = None
    try:
        solver = orc.set_dsyev(True)
    except Exception as e:  # SciPy's LAPACK not reachable: the oracle's own Jacobi
        solver = orc.set_dsyev(False) + f" (dsyev unavailable: {e})"
    n = cfg.nx
    M = f_state.shape[0]
    f = f_state.astype(np.float64)
    t0 = time.perf_counter()
    for s in range(cfg.da_interval):
        f = orc.lbm_step(cfg, f, s)
    t_lbm = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.letkf_analysis(cfg, f, yo)
    t_ana = time.perf_counter() - t0
    orc.set_dsyev(False)
    if limits is not None:
        limits.unregister()
    cycle = t_lbm + t_ana
    sample = (f"one whole DA cycle: {cfg.da_interval} LBM steps of all {M} members at {n}^2, then the LETKF "
              f"analysis at all {n * n} gridpoints; FP64; OpenMP; eigensolver {solver}")
    log(f"[cpu] {nproc} threads ({model}): LBM {t_lbm:.2f} s, analysis {t_ana:.2f} s, cycle {cycle:.2f} s")
    return dict(value=1.0 / cycle, unit="DA cycles/s", cores=nproc, kind="port", sample=sample,
                nproc=nproc, cpu_model=model, eigensolver=solver, lbm_s=t_lbm, analysis_s=t_ana,
                lbm_node_updates_per_s=cfg.da_interval * M * n * n / t_lbm,
                local_analyses_per_s=n * n / t_ana)


def bench_closure_step(lb, hip, M, log):
    """One LBM step of all members with the learned closure (fused MFMA network in
    the collide kernel) against the Smagorinsky step, same workload."""
    out = {}
    for name, kind in (("smagorinsky", lb.CLOSURE_SMAGORINSKY), ("mlp", lb.CLOSURE_MLP)):
        cfg = lb.preset("baseline")
        cfg.closure_kind = kind
        ens = lb.Context(cfg, M)
        for m in range(M):
            ens.init_member(m, 1000 + m)
        if kind == lb.CLOSURE_MLP:
            ens.set_closure(lb.closure_init_params(1))
        ens.forecast(5)
        ens.synchronize()
        e0, e1 = hip.event(), hip.event()
        n = 50
        hip.record(e0, ens.stream)
        ens.forecast(n)
        hip.record(e1, ens.stream)
        ens.synchronize()
        us = hip.elapsed_ms(e0, e1) * 1e3 / n
        ens.close()
        nodes = cfg.nx * cfg.ny * M
        out[name] = {"us_per_step": us, "node_updates_per_s": nodes / us * 1e6,
                     "hbm_GBs_at_72B": LBM_BYTES_PER_NODE_UPDATE * nodes / us * 1e-3}
        log(f"[closure] {name}: {us:.1f} us per step of {M} members")
    out["mlp_over_smagorinsky_time"] = out["mlp"]["us_per_step"] / out["smagorinsky"]["us_per_step"]
    return out


UNEVEN_GRIDS = [(128, 8), (256, 16), (384, 32), (512, 64)]


def bench_uneven(lb, hip, K, W, log, tune_cache=None):
    """The uneven batch: four OSSEs (128^2..512^2, M = 8..64) with observation sites
    at a seeded random density of 0.5-4% per region (4x4 regions), stepped and
    analysed by shared launches (lbmletkf_group_cycle, one hipGraph per cycle);
    each context's launch shapes come from the autotuner first."""
    ctxs, nats, obs, n_obs, infos = [], [], [], [], []
    for k, (n, M) in enumerate(UNEVEN_GRIDS):
        cfg = lb.preset("baseline")
        cfg.nx = cfg.ny = n
        sites = lb.uneven_sites(n, n, seed=100 + k, regions=4, density=(0.005, 0.04))
        ens = lb.Context(cfg, M)
        for m in range(M):
            ens.init_member(m, 1000 + m)
        nat = lb.Context(cfg, 1)
        nat.init_member(0, k)
        ens.set_obs_sites(sites)
        nat.set_obs_sites(sites)
        no = nat.obs_count()[1]
        buf = hip.malloc(4 * no * (W + K + 1))
        for c in range(W + K + 1):
            nat.forecast(cfg.da_interval)
            nat.observe_device(0, 77 + k, c, buf + 4 * no * c, noise=True)
        ens.forecast(cfg.da_interval)
        rep = ens.autotune(tune_cache)
        counts = [len(ens.local_obs(g)[0]) for g in range(0, n * n, max(1, n * n // 1500))]
        infos.append(dict(grid=f"{n}x{n}", members=M, sites=int(sites.size), local_sites_min=min(counts),
                          local_sites_max=max(counts), tuning=list(ens.get_tuning())))
        log(f"[uneven] {n}^2 M={M}: {sites.size} sites, local sites {min(counts)}..{max(counts)}; {rep}")
        ctxs.append(ens)
        nats.append(nat)
        obs.append(buf)
        n_obs.append(no)
    grp = lb.Group(ctxs)

    def yos(c):
        return [obs[k] + 4 * n_obs[k] * c for k in range(len(ctxs))]

    for c in range(W):
        grp.cycle_device(yos(c))
    grp.synchronize()
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    t0 = time.perf_counter()
    hip.record(e0, grp.stream)
    for c in range(W, W + K):
        grp.cycle_device(yos(c))
    hip.record(e1, grp.stream)
    grp.synchronize()
    hip.sync()
    wall = time.perf_counter() - t0
    grp.analysis_device(yos(W + K))
    grp.synchronize()
    for info, ens, nat in zip(infos, ctxs, nats):
        r, u, v = nat.get_macro(0)
        info["analysis_rmse"] = ens.ensemble_stats(r, u, v)[0]
    grp.close()
    for ens, nat, buf in zip(ctxs, nats, obs):
        ens.close()
        nat.close()
        hip.free(buf)
    node_updates = sum(n * n * M for n, M in UNEVEN_GRIDS) * 100
    log(f"[uneven] {K} group cycles in {wall:.3f} s: {K / wall:.2f} cycles/s ({hip.elapsed_ms(e0, e1) / K:.1f} ms/cycle)")
    return {"workload": "4 OSSEs in shared launches: " + ", ".join(f"{n}^2 M={M}" for n, M in UNEVEN_GRIDS) +
                        "; sites at 0.5-4% random density per region; 100 LBM steps + 1 analysis per cycle",
            "cycles_per_s": K / wall, "ms_per_cycle": wall / K * 1e3, "steps": K, "warmup": W,
            "lbm_node_updates_per_cycle": node_updates, "grids": infos}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed DA cycles per run (mixed variant)")
    ap.add_argument("--warmup", type=int, default=10, help="untimed DA cycles")
    ap.add_argument("--runs", type=int, default=3, help="timed runs of the mixed variant (median reported)")
    ap.add_argument("--plain-steps", type=int, default=10,
                    help="timed DA cycles of each plain variant (~0.5 s per cycle)")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--variants", default="plain,plain_tol1e-6,mixed,mixed_mlp",
                    help="plain (eigensolver tolerance 1e-10), plain_tol1e-6 (the reference's syevj "
                         "tolerance), mixed, mixed_mlp (mixed with the learned closure in the forecast)")
    ap.add_argument("--mlp-steps", type=int, default=20, help="timed DA cycles of mixed_mlp")
    ap.add_argument("--prof-cycles", type=int, default=3, help="extra phase-timed cycles (eager launches)")
    ap.add_argument("--tuning", default=None, metavar="LBM,CHAIN",
                    help="fixed launch shapes of the mixed variant instead of the autotuner (for profiling "
                         "runs: LBM shape code, analysis chain), e.g. --tuning=-116,32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch group measurement")
    ap.add_argument("--uneven-steps", type=int, default=3, help="timed group cycles of the uneven batch")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (HIP's blit kernel): the "
                         "calibration of the FETCH_SIZE / WRITE_SIZE counters, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build targets one MI355X; "
              "each GPU would step its own members and analyse its own strip of grid columns", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        hip = Hip()
        n = args.copy_probe << 20
        a, b = hip.malloc(n), hip.malloc(n)
        hip.l.hipMemset(C.c_void_p(a), 1, C.c_size_t(n))
        for _ in range(3):
            hip.ck(hip.l.hipMemcpy(C.c_void_p(b), C.c_void_p(a), C.c_size_t(n), 3), "hipMemcpy D2D")
        hip.sync()
        hip.free(a)
        hip.free(b)
        print(json.dumps({"copy_probe_bytes": n, "copies": 3}), flush=True)
        return

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    lb = load_binding()
    hip = Hip()
    cfg = lb.preset("baseline")
    M, K, W = args.members, args.steps, args.warmup
    nodes = cfg.nx * cfg.ny
    log(f"device: {hip.device_name()}  {lb.version()}  workload 256^2 x M={M}, K={K}, W={W}")

    # nature run and observations of every cycle, resident in HBM
    nat = lb.Context(cfg, 1)
    nat.init_member(0, 0)
    _, n_obs = nat.obs_count()
    P = args.prof_cycles
    R = max(1, args.runs)
    n_cyc = W + K * R + 2 + P
    obs = hip.malloc(4 * n_obs * n_cyc)
    truth = hip.malloc(4 * 3 * nodes * n_cyc)
    for c in range(n_cyc):
        nat.forecast(cfg.da_interval)
        nat.observe_device(0, 2024, c, obs + 4 * n_obs * c, noise=True)
        t = truth + 4 * 3 * nodes * c
        nat.get_macro_device(0, t, t + 4 * nodes, t + 8 * nodes)
    nat.synchronize()
    nat.close()
    log(f"nature run + {n_cyc} observation vectors ready ({n_obs} obs each)")

    res = {}
    for v in args.variants.split(","):
        vcfg = cfg.copy()
        if v.startswith("plain"):
            var, k = lb.VARIANT_PLAIN, min(K, args.plain_steps)
            if v == "plain_tol1e-6":
                vcfg.jacobi_tol = 1e-6
        elif v == "mixed_mlp":
            var, k = lb.VARIANT_MIXED, min(K, args.mlp_steps)
        else:
            var, k = lb.VARIANT_MIXED, K
        single = var == lb.VARIANT_PLAIN or v == "mixed_mlp"
        res[v] = bench_variant(lb, hip, vcfg, var, M, k, min(W, 2) if single else W, P, obs,
                               truth, n_obs, log, repeats=1 if single else R, label=v,
                               fixed_tuning=[int(x) for x in args.tuning.split(",")] if args.tuning else None,
                               closure=lb.closure_init_params(1) if v == "mixed_mlp" else None)
    head = res.get("mixed") or res[list(res)[0]]
    kern = kernel_table(head, M, nodes)
    dom_name = max(kern, key=lambda k: kern[k]["share"])
    dom = kern[dom_name]
    tune = head.get("tuning") or [1, 8]
    solve_wg = (nodes + tune[1] - 1) // tune[1]
    mixed = head["variant"] == "mixed"
    traffic_ana = measured_traffic("k_mixed_analysis<64, 3>", solve_wg) if mixed else None
    if tune[0] <= -200:
        lbm_kernel = "k_lbm_rows3<%d, %d>" % (-tune[0] - 200, 256 if cfg.nx == 256 else 0)
        lbm_wg = (cfg.ny // (-tune[0] - 200)) * M
    elif tune[0] <= -100:
        lbm_kernel, lbm_wg = "k_lbm_rows2<%d>" % (-tune[0] - 100), (cfg.ny // (-tune[0] - 100)) * M
    elif tune[0] < 0:
        lbm_kernel, lbm_wg = "k_lbm_rows<false, %d>" % -tune[0], (cfg.ny // -tune[0]) * M
    else:
        lbm_kernel, lbm_wg = "k_lbm_fused<false>", cfg.nx // 64 * (cfg.ny // 16) * M // tune[0]
    traffic_lbm = measured_traffic(lbm_kernel, lbm_wg) if mixed else None
    roof = dict(kernel=dom.get("kernel", dom_name), bound=dom.get("bound", "hbm"), achieved=dom.get("achieved"),
                peak=dom.get("peak"), unit=dom.get("unit"),
                frac=(dom["achieved"] / dom["peak"]) if dom.get("peak") else None,
                traffic=traffic_ana if dom_name == "letkf_analysis" else None,
                traffic_unit="bytes per launch, rocprofv3 FETCH_SIZE x2 + WRITE_SIZE "
                             "(profiles/hbm_traffic_256x256_M64.md)",
                avg_launch_ms=dom["avg_ms"], share_of_cycle=dom["share"])
    for key in ("flop_count", "peak_unpacked_fp32", "frac_of_unpacked_fp32", "achieved_two_sided_count",
                "frac_two_sided_count", "sweeps_per_gridpoint"):
        if key in dom:
            roof[key] = dom[key]
    lbm = kern["lbm_step"]
    lbm_bytes = LBM_BYTES_PER_NODE_UPDATE * M * nodes
    out = {
        "metric": "DA cycles/s at 256²×M=64 (plain vs mixed+batched) at equal analysis RMSE",
        "value": K / head["wall"],
        "unit": "DA cycles/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": head["wall"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("FP32 LBM; analysis: split-bf16x3 MFMA Gram (FP32 accumulation), FP32 one-sided Jacobi "
                  "(0 FP64 sweeps), FP64 eigenvalue norms and coefficients, FP32 MFMA W.Xb update") if mixed
                 else "FP32 LBM; FP32 Gram; FP64 two-sided Jacobi and transform",
        "data": "synthetic",
        "config": {"workload": "LETKF OSSE 256x256, M=64 members, 100 LBM steps per DA cycle, "
                               "(u,v) every 8th node, Gaussian L=12 cutoff 43.8, beta=1.1",
                   "nx": cfg.nx, "ny": cfg.ny, "members": M, "steps_per_cycle": cfg.da_interval,
                   "n_obs": n_obs, "variant_timed": "mixed" if "mixed" in res else list(res)[0]},
        "runs": len(head["walls"]),
        "cycles_per_s_iqr": head["iqr_cycles_per_s"],
        "variants": {k: {"cycles_per_s": r["K"] / r["wall"], "ms_per_cycle": r["wall"] / r["K"] * 1e3,
                         "timed_cycles": r["K"], "timed_runs": len(r["walls"]),
                         "jacobi_tol": r["jacobi_tol"], "closure": r["closure"],
                         "analysis_rmse": r["stats"][0], "analysis_spread": r["stats"][1],
                         "tuning_lbm_tiles_per_block_and_chain": r["tuning"],
                         "phase_ms_per_cycle": {p: r["phases"][p][0] / P * 1e3 for p in r["phases"]
                                                if p != "solve_fp64"}}
                     for k, r in res.items()},
        "roofline": roof,
        "roofline_lbm": dict(kernel=lbm_kernel, bound="hbm", achieved=lbm["achieved"], peak=HBM_PEAK_GBS,
                             unit="GB/s", frac=lbm["achieved"] / HBM_PEAK_GBS, traffic=traffic_lbm,
                             algorithmic_bytes=lbm_bytes,
                             bytes_note="algorithmic bytes per launch: every population read and written once "
                                        "(72 B per node), whether the launch takes one step or several",
                             steps_per_launch=lbm["steps_per_launch"], avg_launch_ms=lbm["avg_launch_ms"],
                             ms_per_step=lbm["avg_ms"],
                             per_step_72B_GBs=lbm_bytes / (lbm["avg_ms"] * 1e-3) / 1e9,
                             per_step_72B_frac=lbm_bytes / (lbm["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             per_step_note="72 B per node update per step over the time per step (a launch "
                                           "of several steps moves the bytes of one)"),
        "kernels": kern,
    }
    if "mixed" in res:
        probes = {}
        for pv in ("plain", "plain_tol1e-6"):
            if pv in res:
                out["speedup_mixed_over_" + pv] = (res[pv]["wall"] / res[pv]["K"]) / (res["mixed"]["wall"] / K)
                ca = res[pv]["yo_index"]  # the plain run's analysis time
                if ca not in probes:
                    probes[ca] = rmse_probe(lb, cfg, M, ca, obs, truth, n_obs)
                a, b = probes[ca][0], res[pv]["stats"][0]
                out["rmse_rel_diff_mixed_vs_" + pv] = abs(a - b) / b
                out["rmse_compared_at_cycle"] = ca
                log(f"[rmse] cycle {ca}: mixed {a:.6e} {pv} {b:.6e} (rel {abs(a - b) / b:.2e})")
    if not args.no_uneven:
        out["uneven_batch"] = bench_uneven(lb, hip, args.uneven_steps, 1, log)
    out["closure_step"] = bench_closure_step(lb, hip, M, log)
    if not args.no_cpu:
        yo_last = np.empty(n_obs, np.float32)
        hip.d2h(yo_last, obs + 4 * n_obs * head["yo_index"])
        out["cpu_baseline"] = cpu_baseline(cfg, head["state"], yo_last.astype(np.float64), log)
    hip.free(obs)
    hip.free(truth)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
