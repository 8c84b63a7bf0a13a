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
    ap.add_argument("--de-target", type=float, default=0.0,
                    help="surrogate run: the plain run's final cost")
    ap.add_argument("--de-plain-seconds", type=float, default=0.0,
                    help="surrogate run: the plain run's wall time")
    ap.add_argument("--de-lanes", type=int, default=4)
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


def cpu_baseline(w, x, seconds, threads):
    """FD Jacobian of the oracle on a bounded sample of columns."""
    import oracle
    oracle.set_threads(threads)
    o = oracle.Oracle(w.dataset, w.model)
    t0 = time.time()
    o.prepare()
    t_prep = time.time() - t0
    D = o.param_count
    t_res = o.time_residual(x, 1)
    k = int(max(1, min(D, (seconds / max(t_res, 1e-6) - 1) / 2)))
    t_sample = o.time_fd_jacobian(x, k)  # 1 residual + 2k residuals
    t_per_res = t_sample / (2 * k + 1)
    t_unit = t_per_res * (2 * D + 1)
    C = w.dataset.n_configs
    return {"value": C * D / t_unit, "unit": "configs*params/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 central-difference Jacobian (reference algorithm, h=1e-5): "
                      f"1 residual + {k} of {D} columns over all {C} configs ({2 * k + 1} residual "
                      f"evaluations, {t_sample:.1f} s); full residual+Jacobian extrapolated to "
                      f"{2 * D + 1} evaluations = {t_unit:.1f} s; neighbour list + site priming "
                      f"{t_prep:.1f} s untimed"}


def de_measure(a):
    """Surrogate-screened DE vs plain batched DE on the ternary EAM workload
    (D = 240, NP = 15 D, F = 0.65, CR = 0.5, seed 42): wall time of the
    screened run until its best cost is <= (1 + 1e-6) x the plain run's final
    cost, over the plain run's wall time."""
    import numpy as np

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
    lo = x0 - a.de_box * np.abs(x0) - 1e-3
    hi = x0 + a.de_box * np.abs(x0) + 1e-3
    common = dict(F=0.65, CR=0.5, seed=42, progress_every=25)
    out = {"metric": "DE wall-time to target cost vs plain GPU", "unit": "x", "higher_is_better": False,
           "n_gpus": 1, "data": "synthetic (seeded ternary analytic EAM truth + noise)",
           "dtype": f"plain {a.de_plain_precision}, surrogate-run exact costs {a.de_surrogate_precision}",
           "config": {"workload": f"ternary tabulated EAM, {a.configs} configs, D={x0.size}",
                      "np": a.de_np, "generations": a.de_generations, "F": 0.65, "CR": 0.5, "seed": 42,
                      "box": f"x0 +- ({a.de_box}|x0| + 1e-3)", "lanes_per_atom": a.de_lanes,
                      "plain_precision": a.de_plain_precision,
                      "surrogate_precision": a.de_surrogate_precision,
                      "surrogate": {"hidden": 256, "oversample": 4, "retrain_every": 10}}}
    t_plain, c_plain = a.de_plain_seconds, a.de_target
    if a.de in ("plain", "both"):
        t0 = time.perf_counter()
        _, c_plain, hp = g.de(x0, lo, hi, a.de_np, a.de_generations, **common)
        t_plain = time.perf_counter() - t0
        out["plain"] = {"seconds": t_plain, "final_cost": c_plain, "generations": len(hp),
                        "ms_per_generation": 1e3 * t_plain / max(len(hp), 1),
                        "pair_member_evals_per_s": pair_member_rate(t_plain, len(hp))}
        log(f"plain: {len(hp)} generations {t_plain:.1f} s final {c_plain:.9e}")
    if a.de in ("surrogate", "both"):
        if c_plain <= 0:
            raise SystemExit("--de surrogate needs --de-target (the plain run's final cost)")
        if a.de == "both" and a.de_surrogate_precision != a.de_plain_precision:
            g.close()
            g = calculator(a.de_surrogate_precision)
        target = c_plain * (1 + 1e-6)
        t0 = time.perf_counter()
        _, c_s, hs = g.de(x0, lo, hi, a.de_np, a.de_generations, surrogate=True, oversample=4,
                          retrain_every=10, hidden=256, target_cost=target, **common)
        t_s = time.perf_counter() - t0
        out["surrogate"] = {"seconds": t_s, "final_cost": c_s, "generations": len(hs),
                            "reached": bool(c_s <= target),
                            "ms_per_generation": 1e3 * t_s / max(len(hs), 1),
                            "pair_member_evals_per_s": pair_member_rate(t_s, len(hs))}
        log(f"surrogate: {len(hs)} generations {t_s:.1f} s best {c_s:.9e} reached {c_s <= target}")
        if t_plain > 0:
            out["value"] = t_s / t_plain if c_s <= target else None
    print(json.dumps(out), flush=True)


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
    import numpy as np

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
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / a.steps
    value = C * D * a.steps / dt
    log(f"step {ms:.3f} ms -> {value:.4g} configs*params/s")

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
        # algorithmic bytes of k_jac_de (the dominant kernel) per launch, each
        # datum once: pair records (float4 + FP32 distance, 20 B/pair), the FP32
        # force rows of the density/embedding columns it writes (3N x D_de x 4 B)
        # and the per-atom density-column rows it reads (Hs FP32 + H FP64, 12 B
        # per atom and density column)
        D_dens = sum(free[m.npair:m.npair + m.ntypes])
        D_de = sum(free[m.npair:])
        b_main = 20 * P + 12 * N * D_de + 12 * N * D_dens
        achieved = b_main / (t_main * 1e-3) / 1e9
        ternary500 = a.workload == "eam-ternary" and C == 500
        roofline = {"bound": "hbm", "kernel": "k_jac_de", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic() if ternary500 else None,
                    "traffic_source": ("profiles/k_jac_de_traffic_eam_ternary_500cfg.md (rocprofv3 "
                                       "FETCH_SIZE + WRITE_SIZE, copy-calibrated)") if ternary500 else None,
                    "bytes_per_launch": b_main, "ms_per_launch": t_main}
    elif kind == "adp":
        t_main = g.time_step(4, 10)
        t_quad = g.time_step(6, 10)
        kernels = {"residual": t_res, "jacobian": t_jac, "jac_de_and_moment_passes": t_main,
                   "adp_quadrupole_pass": t_quad, "normal_equations": t_ne}
        # algorithmic bytes of the quadrupole-column force pass (the dominant
        # kernel) per launch: pair records (20 B/pair), the lambda rows of
        # the atoms (48 B), their d lambda rows over the quadrupole columns
        # (24 B per atom and column) and the FP32 force rows it writes
        # (3N x D_w x 4 B)
        D_w = sum(free[2 * m.npair + 2 * m.ntypes:])
        b_main = 20 * P + 48 * N + 24 * N * D_w + 12 * N * D_w
        achieved = b_main / (t_quad * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": "k_adp_dq (quadrupole columns)", "achieved": achieved,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": None, "bytes_per_launch": b_main, "ms_per_launch": t_quad}
    else:
        kernels = {"residual": t_res, "jacobian": t_jac, "normal_equations": t_ne}
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
        "dtype": "fp32" if a.precision == "mixed" else "f64",
        "data": f"synthetic (seeded analytic {kind.upper()} truth + noise)",
        "config": {"workload": desc.format(C=C, D=D),
                   "configs": C, "atoms": N, "pairs": P, "params": D, "rows": R,
                   "precision": a.precision, "lanes_per_atom": lpa,
                   "blocks_per_cu": shape["blocks_per_cu"],
                   "jacobian_tile_atoms": shape["jacobian_tile_atoms"],
                   "jacobian_blocks_per_cu": shape["jacobian_blocks_per_cu"]},
        "roofline": roofline,
        "kernels_ms": kernels,
        "step_bytes_alg": b_alg,
        "step_hbm_frac": b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(w, x, a.cpu_seconds, threads)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
