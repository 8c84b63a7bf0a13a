#!/usr/bin/env python3
"""Residual + Jacobian throughput of the ternary tabulated EAM workload on one MI355X.

One step = one residual vector plus the full analytic Jacobian at one
parameter vector over the whole dataset (500 configurations, 32-2048 atoms,
240 spline knots), with every input resident in HBM. The metric is
configs*params/s = C * D / (seconds per step).

Prints one JSON line with the step throughput, the roofline figure of the
dominant kernel (k_jac_de: the density/embedding-column pass of the Jacobian),
and a CPU baseline:
the FP64 oracle (a restatement of the reference's oneTBB calculator: central
differences, two residual evaluations per column, configurations in parallel)
timed on a bounded sample of columns on this host.

--workload runs the same step on the other workloads (single-element EAM,
binary ADP with the roofline of its quadrupole-column pass, Si
Stillinger-Weber); the default line is the ternary EAM one.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


# workload name -> (dataset generator, description); the generators live in
# forcesmith_amd.datasets
WORKLOADS = {
    "eam-ternary": ("config2", "ternary tabulated EAM, {C} configs of 32-2048 atoms at 0.06/A^3, "
                               "12 tables x 20 knots (D={D}), residual + full Jacobian"),
    "eam-single": ("config1", "single-element tabulated EAM, {C} FCC configs of 108 atoms, "
                              "15/15/12 knots (D={D}), residual + full Jacobian"),
    "adp-binary": ("config3", "binary tabulated ADP, {C} configs of 250-1024 atoms, 13 tables x 20 "
                              "knots (D={D} free: dipole, quadrupole, embedding), residual + full "
                              "Jacobian"),
    "sw-si": ("config4", "Si Stillinger-Weber, {C} configs of 64-512 atoms (D={D}), residual + "
                         "full Jacobian"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--configs", type=int, default=500,
                    help="configurations of the ternary EAM workload")
    ap.add_argument("--workload", default="eam-ternary", choices=sorted(WORKLOADS),
                    help="dataset the step runs on; the headline figure is eam-ternary, the "
                         "others are the single-element EAM, binary ADP and Si SW workloads")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    # differential-evolution measurement (separate from the default line)
    ap.add_argument("--de", choices=["plain", "surrogate", "both"], default=None,
                    help="measure DE wall time to the plain run's final cost instead")
    ap.add_argument("--de-np", type=int, default=3600)
    ap.add_argument("--de-generations", type=int, default=1000)
    ap.add_argument("--de-box", type=float, default=0.2,
                    help="box half-width as a fraction of |x0| (plus 1e-3)")
    ap.add_argument("--de-reference-box", action="store_true",
                    help="BoostDESolver's box instead: unbounded knots become x0 -+ max(2|x0|, 5e-4) "
                         "(solver.cpp:118-125)")
    ap.add_argument("--de-target", type=float, default=0.0,
                    help="surrogate run: the plain run's final cost")
    ap.add_argument("--de-plain-seconds", type=float, default=0.0,
                    help="surrogate run: the plain run's wall time")
    ap.add_argument("--de-lanes", type=int, default=4)
    ap.add_argument("--lm-iterations", type=int, default=6,
                    help="fs_lm_solve iterations timed in the default line (0: none)")
    ap.add_argument("--de-sample-generations", type=int, default=10,
                    help="generations of the DE sample in the default line (0: none)")
    ap.add_argument("--de-plain-precision", default="fp64", choices=["mixed", "fp64"],
                    help="precision of the plain DE run's batched costs (the baseline: FP64)")
    ap.add_argument("--de-surrogate-precision", default="mixed", choices=["mixed", "fp64"],
                    help="precision of the screened run's exact costs")
    ap.add_argument("--jacobian-tile", type=int, default=0,
                    help="atoms per Jacobian tile (4-64); 0 lets the autotuner time the sizes")
    ap.add_argument("--pmc-driver", action="store_true",
                    help="counter-collection driver: a 1 GiB copy (calibration) then three "
                         "residual + Jacobian evaluations; prints no JSON")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": aff,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def baseline_threads(host):
    """Threads of the CPU baseline: every CPU this process may use, within the
    thread budget the host sets for it (OMP_NUM_THREADS; a GPU box allots 16
    CPUs to each GPU)."""
    n = host["affinity"]
    try:
        n = min(n, int(host["omp_num_threads"]))
    except (TypeError, ValueError):
        pass
    return max(1, n)


def cpu_baseline(w, x, seconds, threads, host):
    """The oracle (an FP64 restatement of the reference's calculator) on this
    host, following the baseline protocol: untimed setup, then per figure one
    warm-up run and the median of three timed runs.
      value: the reference algorithm, central differences (h = 1e-5) with the
        configurations in parallel: 2D + 1 residual evaluations per unit,
        timed on a bounded sample of k columns and scaled to 2D + 1;
      single_thread: the same on one thread (one run, one column);
      analytic_jacobian: labelled second baseline, the oracle's analytic
        per-knot Jacobian plus one residual (not the reference's algorithm)."""
    import statistics

    import oracle
    oracle.set_threads(threads)
    o = oracle.Oracle(w.dataset, w.model)
    t0 = time.time()
    o.prepare()
    t_prep = time.time() - t0
    D = o.param_count
    C = w.dataset.n_configs
    t_res = o.time_residual(x, 1)
    # one run of 2k + 1 residuals takes about a quarter of the budget
    k = int(max(1, min(D, (seconds / 4 / max(t_res, 1e-6) - 1) / 2)))
    o.time_fd_jacobian(x, k)  # warm-up
    runs = [o.time_fd_jacobian(x, k) for _ in range(3)]
    t_sample = statistics.median(runs)
    t_unit = t_sample / (2 * k + 1) * (2 * D + 1)
    out = {"value": C * D / t_unit, "unit": "configs*params/s", "cores": threads, "kind": "port",
           "sample": f"oracle FP64 central-difference Jacobian (reference algorithm, h=1e-5) on "
                     f"{threads} threads: 1 residual + {k} of {D} columns over all {C} configs "
                     f"({2 * k + 1} residual evaluations), 1 warm-up + median of 3 runs "
                     f"({', '.join(f'{r:.2f}' for r in runs)} s); scaled to {2 * D + 1} evaluations "
                     f"= {t_unit:.1f} s per residual+Jacobian; neighbour list + site priming "
                     f"{t_prep:.1f} s untimed",
           "runs_s": runs, "host": host}
    oracle.set_threads(1)
    t1 = o.time_fd_jacobian(x, 1)
    oracle.set_threads(threads)
    t_unit1 = t1 / 3 * (2 * D + 1)
    out["single_thread"] = {"value": C * D / t_unit1, "cores": 1,
                            "sample": f"1 residual + 1 column (3 residual evaluations, {t1:.2f} s) "
                                      f"scaled to {2 * D + 1}"}
    J = np.zeros((o.residual_count, D))

    def analytic():
        t = time.perf_counter()
        o.residual(x)
        o._lib.or_jacobian_analytic(o._h, oracle._p(np.ascontiguousarray(x, dtype=np.float64), oracle._dp),
                                    oracle._p(J, oracle._dp))
        return time.perf_counter() - t

    analytic()  # warm-up
    ar = [analytic() for _ in range(3)]
    ta = statistics.median(ar)
    out["analytic_jacobian"] = {"value": C * D / ta, "cores": threads, "runs_s": ar,
                                "label": "oracle FP64 residual + analytic per-knot Jacobian (a second "
                                         "baseline; the reference computes these columns by central "
                                         "differences)"}
    return out


def de_measure(a):
    """Surrogate-screened DE vs plain batched DE on the ternary EAM workload
    (D = 240, NP = 15 D, F = 0.65, CR = 0.5, seed 42): wall time of the
    screened run until its best cost is <= (1 + 1e-6) x the plain run's final
    cost, over the plain run's wall time."""
    import forcesmith_amd as fa
    from forcesmith_amd import datasets

    log(f"generating the ternary EAM workload ({a.configs} configs, seed 2)")
    w = datasets.config2(n_configs=a.configs)

    pairs = [0]

    def calculator(precision):
        g = fa.GpuCalculator(w.dataset, w.model, precision=precision)
        pairs[0] = g.prepare()
        g.set_launch(a.de_lanes)
        return g

    def pair_member_rate(seconds, gens):
        # one generation evaluates NP members over every neighbour pair
        return a.de_np * pairs[0] * gens / seconds if seconds > 0 else None

    g = calculator(a.de_plain_precision if a.de in ("plain", "both") else a.de_surrogate_precision)
    x0 = g.gather_params()
    if a.de_reference_box:
        lo, hi = g.gather_bounds()  # +-inf: the driver derives the reference's box
        box = "reference: x0 -+ max(2|x0|, 5e-4)"
    else:
        lo = x0 - a.de_box * np.abs(x0) - 1e-3
        hi = x0 + a.de_box * np.abs(x0) + 1e-3
        box = f"x0 +- ({a.de_box}|x0| + 1e-3)"
    common = dict(F=0.65, CR=0.5, seed=42, progress_every=25)
    out = {"metric": "DE wall-time to target cost vs plain GPU", "unit": "x", "higher_is_better": False,
           "n_gpus": 1, "data": "synthetic (seeded ternary analytic EAM truth + noise)",
           "dtype": f"plain {a.de_plain_precision}, surrogate-run exact costs {a.de_surrogate_precision}",
           "config": {"workload": f"ternary tabulated EAM, {a.configs} configs, D={x0.size}",
                      "np": a.de_np, "generations": a.de_generations, "F": 0.65, "CR": 0.5, "seed": 42,
                      "box": box, "lanes_per_atom": a.de_lanes,
                      "plain_precision": a.de_plain_precision,
                      "surrogate_precision": a.de_surrogate_precision,
                      "surrogate": {"hidden": 256, "oversample": 4, "retrain_every": 10}}}
    t_plain, c_plain = a.de_plain_seconds, a.de_target
    if a.de in ("plain", "both"):
        t0 = time.perf_counter()
        _, c_plain, hp = g.de(x0, lo, hi, a.de_np, a.de_generations, **common)
        t_plain = time.perf_counter() - t0
        out["plain"] = {"seconds": t_plain, "final_cost": c_plain, "generations": len(hp),
                        "initial_best": float(hp[0]) if len(hp) else None,
                        "ms_per_generation": 1e3 * t_plain / max(len(hp), 1),
                        "pair_member_evals_per_s": pair_member_rate(t_plain, len(hp))}
        log(f"plain: {len(hp)} generations {t_plain:.1f} s final {c_plain:.9e}")
    # a target the plain run did not move below its initial best is no
    # target: any member at x0 meets it, and a mixed-precision cost of x0
    # can sit below its FP64 cost by rounding alone
    progress = None
    if a.de in ("plain", "both"):
        progress = (out["plain"]["initial_best"] - c_plain) / out["plain"]["initial_best"]
        out["plain"]["relative_progress"] = progress
    if a.de in ("surrogate", "both"):
        if c_plain <= 0:
            raise SystemExit("--de surrogate needs --de-target (the plain run's final cost)")
        if progress is not None and progress <= 1e-6:
            out["value"] = None
            out["no_target"] = ("the plain run ended within 1e-6 of its initial best: no target "
                                "below the starting point to race to")
            log(out["no_target"])
            print(json.dumps(out), flush=True)
            return
        if a.de == "both" and a.de_surrogate_precision != a.de_plain_precision:
            g.close()
            g = calculator(a.de_surrogate_precision)
        target = c_plain * (1 + 1e-6)
        t0 = time.perf_counter()
        xs, c_s, hs = g.de(x0, lo, hi, a.de_np, a.de_generations, surrogate=True, oversample=4,
                           retrain_every=10, hidden=256, target_cost=target, **common)
        t_s = time.perf_counter() - t0
        # the screened run's best judged by the plain run's precision
        if a.de_surrogate_precision != a.de_plain_precision:
            gv = calculator(a.de_plain_precision)
            c_v = gv.cost(xs)
            gv.close()
        else:
            c_v = c_s
        reached = bool(c_v <= target)
        out["surrogate"] = {"seconds": t_s, "final_cost": c_s, "final_cost_plain_precision": c_v,
                            "generations": len(hs), "reached": reached, "stats": g.de_stats(),
                            "ms_per_generation": 1e3 * t_s / max(len(hs), 1),
                            "pair_member_evals_per_s": pair_member_rate(t_s, len(hs))}
        log(f"surrogate: {len(hs)} generations {t_s:.1f} s best {c_s:.9e} ({a.de_plain_precision} "
            f"{c_v:.9e}) reached {reached}")
        if t_plain > 0:
            out["value"] = t_s / t_plain if reached else None
    print(json.dumps(out), flush=True)


def de_sample(w, P, gens, np_, lanes):
    """A few generations of plain FP64 DE and of surrogate-screened mixed DE
    on the step's workload (box x0 +- (0.2|x0| + 1e-3), F = 0.65, CR = 0.5,
    seed 42): time per generation, neighbour-pair x member evaluations per
    second and the screened run's statistics. The 1000-generation comparison
    is bench.py --de both."""
    import forcesmith_amd as fa
    res = {"workload": "the step's workload", "np": np_, "generations": gens,
           "box": "x0 +- (0.2|x0| + 1e-3)", "F": 0.65, "CR": 0.5, "seed": 42,
           "full_comparison": "bench.py --de both (1000 generations)"}
    for key, prec, kw in (("plain_fp64", "fp64", {}),
                          ("screened_mixed", "mixed",
                           dict(surrogate=True, oversample=4, retrain_every=10, hidden=256))):
        g = fa.GpuCalculator(w.dataset, w.model, precision=prec)
        g.prepare()
        g.set_launch(lanes)
        x0 = g.gather_params()
        lo = x0 - 0.2 * np.abs(x0) - 1e-3
        hi = x0 + 0.2 * np.abs(x0) + 1e-3
        t0 = time.perf_counter()
        _, best, h = g.de(x0, lo, hi, np_, gens, F=0.65, CR=0.5, seed=42, **kw)
        dt = time.perf_counter() - t0
        r = {"seconds": dt, "ms_per_generation": 1e3 * dt / max(len(h), 1),
             "pair_member_evals_per_s": np_ * P * len(h) / dt,
             "best_first": float(h[0]), "best_last": float(best)}
        if kw:
            r["stats"] = g.de_stats()
        res[key] = r
        log(f"DE sample {key}: {len(h)} generations {dt:.2f} s, best {h[0]:.6e} -> {best:.6e}")
        g.close()
    return res


def measured_traffic(kernel="k_jac_de"):
    """HBM bytes per launch of the kernel from its committed two-pass PMC
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by a 1 GiB copy)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{kernel}_traffic_eam_ternary_500cfg.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_limiter(kernel="k_jac_de", summary=None):
    """What the counters say bounds the kernel, from its committed summary
    (the line 'limiter: ...'), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        summary or f"{kernel}_traffic_eam_ternary_500cfg.md")
    try:
        for line in open(path):
            if line.startswith("limiter:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def pmc_driver(a):
    """Workload for rocprofv3 --pmc passes: the copy kernel of known size that
    scales FETCH_SIZE / WRITE_SIZE, then the Jacobian kernels."""
    import torch

    n = 1 << 28  # 2^28 FP32 = 1 GiB read and 1 GiB written per dispatch
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.add(x, 1.0, out=y)
    torch.cuda.synchronize()
    del x, y
    torch.cuda.empty_cache()
    log("calibration copies done")
    import forcesmith_amd as fa
    from forcesmith_amd import datasets

    w = datasets.config2(n_configs=a.configs)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    if a.jacobian_tile:
        g.set_jacobian_tile(a.jacobian_tile)
    g.autotune()
    log(f"launch shape {g.launch_shape()}")
    lib, h = g._lib, g._h
    for _ in range(3):
        g._check(lib.fs_eval_residual(h, None, None))
        g._check(lib.fs_eval_jacobian(h, None))
        g.synchronize()
    log("residual + Jacobian evaluations done")


def main():
    a = parse()
    if a.pmc_driver:
        pmc_driver(a)
        return
    if a.de:
        if a.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        de_measure(a)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported; this build runs the "
              "evaluator on a single device", file=sys.stderr)
        sys.exit(2)
    import forcesmith_amd as fa
    from forcesmith_amd import datasets

    gen, desc = WORKLOADS[a.workload]
    log(f"generating the {a.workload} workload")
    w = getattr(datasets, gen)(n_configs=a.configs) if gen == "config2" else getattr(datasets, gen)()
    ds = w.dataset
    kind = w.model.kind
    g = fa.GpuCalculator(ds, w.model, precision=a.precision)
    t0 = time.time()
    P = g.prepare()
    log(f"neighbour list: {P} pairs in {time.time() - t0:.2f} s")
    if a.jacobian_tile:
        g.set_jacobian_tile(a.jacobian_tile)
    lpa = g.autotune()
    shape = g.launch_shape()
    log(f"autotuned lanes per atom: {lpa}, blocks per CU: {shape['blocks_per_cu']}, "
        f"Jacobian tile: {shape['jacobian_tile_atoms']} atoms, Jacobian blocks per CU: "
        f"{shape['jacobian_blocks_per_cu']}")
    x = g.gather_params()
    C, D, N, R = ds.n_configs, g.param_count, ds.n_atoms, g.residual_count
    lib, h = g._lib, g._h

    def step():
        g.scatter_params(x)  # new parameter vector: slopes, residual and Jacobian all recomputed
        g._check(lib.fs_eval_residual(h, None, None))
        g._check(lib.fs_eval_jacobian(h, None))

    for _ in range(a.warmup):
        step()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    g.synchronize()
    ms_eager = 1e3 * (time.perf_counter() - t0) / a.steps
    log(f"eager step {ms_eager:.3f} ms")

    # the same step replayed from its captured hipGraph (fs_eval_step): the
    # parameter vector is copied to a device buffer the graph reads, then
    # scatter, slopes, residual, Jacobian and cost run as one graph launch
    g.step(x)  # eager warm-up and capture
    for _ in range(a.warmup):
        g.step(x, want_cost=False)
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g.step(x, want_cost=False)
    g.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / a.steps
    value = C * D * a.steps / dt
    log(f"captured step {ms:.3f} ms -> {value:.4g} configs*params/s")

    # kernel times with HIP events on the context stream
    t_res = g.time_step(0, 10)
    t_jac = g.time_step(1, 10)
    t_ne = g.time_step(2, 5)
    m = w.model
    free = [int((~t.fixed).sum()) for t in m.tables] if kind != "sw" else []
    roofline = None
    if kind == "eam":
        t_main = g.time_step(4, 10)
        t_pair = g.time_step(5, 10)
        kernels = {"residual": t_res, "jacobian": t_jac, "jac_de": t_main, "jac_pair": t_pair,
                   "normal_equations": t_ne}
        # compulsory bytes of k_jac_de (the dominant kernel) per launch: the
        # pair records (int32 partner + 3 x FP32 displacement, 16 B/pair) and the
        # FP32 force rows of the density/embedding columns it writes
        # (3N x D_de x 4 B). Intermediates it also reads, reported apart: the
        # FP64 sorted distances (8 B/pair) and the per-atom density-column rows
        # (Hs FP32 + H FP64, 12 B per atom and density column)
        D_dens = sum(free[m.npair:m.npair + m.ntypes])
        D_de = sum(free[m.npair:])
        b_main = 16 * P + 12 * N * D_de
        b_inter = 8 * P + 12 * N * D_dens
        achieved = b_main / (t_main * 1e-3) / 1e9
        ternary500 = a.workload == "eam-ternary" and C == 500
        roofline = {"bound": "hbm", "kernel": "k_jac_de", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic() if ternary500 else None,
                    "traffic_source": ("profiles/k_jac_de_traffic_eam_ternary_500cfg.md (rocprofv3 "
                                       "FETCH_SIZE + WRITE_SIZE, copy-calibrated)") if ternary500 else None,
                    "bytes_per_launch": b_main, "intermediate_bytes_per_launch": b_inter,
                    "ms_per_launch": t_main,
                    "limiter": measured_limiter() if ternary500 else None}
    elif kind == "adp":
        t_main = g.time_step(4, 10)
        t_quad = g.time_step(6, 10)
        kernels = {"residual": t_res, "jacobian": t_jac, "jac_de_and_moment_passes": t_main,
                   "adp_quadrupole_pass": t_quad, "normal_equations": t_ne}
        # compulsory bytes of the quadrupole-column force pass (the dominant
        # kernel) per launch: pair records (16 B/pair) and the FP32 force rows
        # it writes (3N x D_w x 4 B); the lambda rows (48 B per atom) and the
        # d lambda rows (24 B per atom and column) it gathers are intermediates
        D_w = sum(free[2 * m.npair + 2 * m.ntypes:])
        b_main = 16 * P + 12 * N * D_w
        b_inter = 8 * P + 48 * N + 24 * N * D_w
        achieved = b_main / (t_quad * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": "k_adp_dq (quadrupole columns)", "achieved": achieved,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": None, "bytes_per_launch": b_main, "intermediate_bytes_per_launch": b_inter,
                    "ms_per_launch": t_quad}
    else:
        t_main = g.time_step(4, 10)
        kernels = {"residual": t_res, "jacobian": t_jac, "jac_centre": t_main, "normal_equations": t_ne}
        # compulsory bytes of the SW triplet-column centre pass (the dominant
        # kernel) per launch: the pair records (16 B/pair) and the force
        # tangents of its columns (3N x D_tri x 4 B). The per-pair records it
        # reads (unit vector, 1/r, radial factors: 64 B/pair) and the triplet
        # slots it hands to the tile pass (24 B per pair and column) are
        # intermediates. The pass is FP64 arithmetic and latency bound
        pf = np.asarray(m.pair_fixed, bool).reshape(-1, 8)
        D_tri = int((~pf[:, 6:8]).sum() + (~np.asarray(m.lam_fixed, bool)).sum())  # gamma, a2, lambda
        b_main = 16 * P + 12 * N * D_tri
        b_inter = 64 * P + 24 * P * D_tri
        achieved = b_main / (t_main * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": "k_sw_jac_centre (triplet columns)", "achieved": achieved,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": None, "bytes_per_launch": b_main, "intermediate_bytes_per_launch": b_inter,
                    "ms_per_launch": t_main}
    b_alg = 16 * P + 28 * N + 56 * C + 8 * R + 4 * R * D
    out = {
        "metric": "residual+Jacobian configs*params/s",
        "value": value,
        "unit": "configs*params/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # SW is FP64 arithmetic in both modes
        "dtype": "fp32" if a.precision == "mixed" and kind != "sw" else "f64",
        "data": f"synthetic (seeded analytic {kind.upper()} truth + noise)",
        "config": {"workload": desc.format(C=C, D=D),
                   "configs": C, "atoms": N, "pairs": P, "params": D, "rows": R,
                   "precision": a.precision, "lanes_per_atom": lpa,
                   "blocks_per_cu": shape["blocks_per_cu"],
                   "jacobian_tile_atoms": shape["jacobian_tile_atoms"],
                   "jacobian_blocks_per_cu": shape["jacobian_blocks_per_cu"]},
        "roofline": roofline,
        "kernels_ms": kernels,
        "eager_ms_per_step": ms_eager,
        "step": "fs_eval_step: parameter copy + one hipGraph launch (scatter, slopes, residual, "
                "Jacobian, cost)",
        "step_bytes_alg": b_alg,
        "step_hbm_frac": b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    if a.lm_iterations > 0:
        # EigenLMSolver's loop (fs_lm_solve) from x: device calls against the
        # host's D x D trust-region algebra
        _, status, counts, hist = g.lm_solve(x, max_iter=a.lm_iterations)
        tm = g.lm_timing()
        out["lm_solve"] = {"iterations": a.lm_iterations, "status": status,
                           "nfev": counts[0], "njev": counts[1], "accepted": counts[2],
                           "cost_first": float(hist[0]) if len(hist) else None,
                           "cost_last": float(hist[len(hist) - 1]) if len(hist) else None,
                           **tm, "host_algebra_share": tm["host_algebra_ms"] / tm["total_ms"]}
        log(f"lm_solve {a.lm_iterations} iterations: {tm['total_ms']:.1f} ms, host algebra "
            f"{tm['host_algebra_ms']:.2f} ms ({100 * tm['host_algebra_ms'] / tm['total_ms']:.1f} %)")
    if kind == "eam" and a.de_sample_generations > 0:
        out["de"] = de_sample(w, P, a.de_sample_generations, a.de_np, a.de_lanes)
        if ternary500:
            # the population passes' bound, from their committed counter summary
            out["de"]["limiter"] = measured_limiter(summary="k_pop_ml_eam_ternary_500cfg.md")
    if not a.no_cpu_baseline:
        host = host_cpu()
        threads = a.cpu_threads or baseline_threads(host)
        log(f"CPU baseline on {threads} threads ({host['cpu_model']}, nproc {host['nproc']}, "
            f"affinity {host['affinity']})")
        cb = cpu_baseline(w, x, a.cpu_seconds, threads, host)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
        out["speedup_vs_cpu_analytic"] = value / cb["analytic_jacobian"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
