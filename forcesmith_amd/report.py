"""Evaluation report and its forces-only comparison gate.

``evaluate_report`` writes the per-configuration JSON of the reference's
``--evaluate`` option (src/cli/evaluate_report.cpp:21-93) from one batched
``fs_eval_forces`` on the device: calc/ref energy, stress (Voigt order xx, yy,
zz, xy, xz, yz as include/forcesmith/core/voigt.hpp:9-13), per-atom calc/ref
forces, the limit term and the per-configuration sum of squares

    sumsq = sum |F - F_ref|^2 + (w_E (E - E_ref))^2
            [+ sum_v (w_S (S_v - S_ref_v))^2 if w_S > 0] + limit^2,

which is the squared norm of that configuration's residual rows.

``diff_reports`` is the forces-only gate of the reference's parity harness
(tests/integration/harness/diff_eval.py:95-163): the weight-free sum of squared
force residuals must agree to rtol or atol, and every calc force component must
satisfy |a - b| <= atol + rtol |b| (defaults rtol = atol = 1e-5).
"""
from __future__ import annotations

import json

import numpy as np

from .model import Dataset

_VOIGT = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def _voigt6(S) -> list:
    S = np.asarray(S, dtype=np.float64).reshape(3, 3)
    return [float(S[i, j]) for i, j in _VOIGT]


def build_report(ds: Dataset, F, E, S, L, energy_weight: float, stress_weight: float,
                 names=None) -> dict:
    """The report dictionary from evaluated forces F (N, 3), energies E (C),
    stresses S (C, 3, 3) and limit terms L (C) of dataset ``ds``."""
    ew, sw = float(energy_weight), float(stress_weight)
    F = np.asarray(F, dtype=np.float64).reshape(-1, 3)
    E = np.asarray(E, dtype=np.float64)
    S = np.asarray(S, dtype=np.float64).reshape(-1, 3, 3)
    L = np.asarray(L, dtype=np.float64)
    if F.shape[0] != ds.n_atoms or E.shape[0] != ds.n_configs:
        raise ValueError("evaluated arrays do not match the dataset")
    configs, total = [], 0.0
    for c in range(ds.n_configs):
        a0, a1 = int(ds.atom_off[c]), int(ds.atom_off[c + 1])
        dF = F[a0:a1] - ds.ref_force[a0:a1]
        csq = float(np.sum(dF * dF))
        de = ew * (E[c] - ds.ref_energy[c])
        csq += de * de
        calc6 = _voigt6(S[c])
        ref6 = [float(v) for v in ds.ref_stress[c]]  # stored in the same Voigt order
        if sw > 0.0:
            csq += sum((sw * (a - b)) ** 2 for a, b in zip(calc6, ref6))
        csq += float(L[c]) ** 2
        total += csq
        configs.append({
            "index": c,
            "name": names[c] if names is not None else "",
            "natoms": a1 - a0,
            "calc_energy": float(E[c]),
            "ref_energy": float(ds.ref_energy[c]),
            "calc_stress": calc6,
            "ref_stress": ref6,
            "atoms": [{"calc_force": F[i].tolist(), "ref_force": ds.ref_force[i].tolist()}
                      for i in range(a0, a1)],
            "limit": float(L[c]),
            "sumsq": csq,
        })
    return {"energy_weight": ew, "stress_weight": sw, "nconf": ds.n_configs, "configs": configs,
            "total_sumsq": total}


def evaluate_report(calc, energy_weight: float, stress_weight: float, names=None,
                    path=None) -> dict:
    """Evaluate every configuration of ``calc`` (a GpuCalculator) at its
    current parameters and build the report; written as indented JSON to
    ``path`` when given."""
    F, E, S, L = calc.eval_forces()
    rep = build_report(calc.dataset, F, E, S, L, energy_weight, stress_weight, names)
    if path is not None:
        with open(path, "w") as f:
            f.write(json.dumps(rep, indent=2) + "\n")
    return rep


def force_sumsq(rep: dict) -> float:
    """Weight-free sum of squared force residuals of a report."""
    s = 0.0
    for cfg in rep["configs"]:
        for a in cfg["atoms"]:
            d = np.asarray(a["calc_force"]) - np.asarray(a["ref_force"])
            s += float(d @ d)
    return s


def diff_reports(port: dict, ref: dict, rtol: float = 1e-5, atol: float = 1e-5) -> dict:
    """Forces-only gate between two reports. Returns a summary with ``ok``,
    the scalar check and the per-atom check (n, unmatched, fail, max_abs,
    worst as (config, atom, component))."""
    p_tot, r_tot = force_sumsq(port), force_sumsq(ref)
    rel = abs(p_tot - r_tot) / max(abs(r_tot), atol)
    scalar_ok = rel <= rtol or abs(p_tot - r_tot) <= atol
    ref_atoms = {}
    for cfg in ref["configs"]:
        for ai, a in enumerate(cfg["atoms"]):
            ref_atoms[(cfg["index"], ai)] = a["calc_force"]
    n = unmatched = fail = 0
    max_abs, worst = 0.0, None
    for cfg in port["configs"]:
        for ai, a in enumerate(cfg["atoms"]):
            rc = ref_atoms.get((cfg["index"], ai))
            if rc is None:
                unmatched += 1
                continue
            for k in range(3):
                d = abs(a["calc_force"][k] - rc[k])
                if d > max_abs:
                    max_abs, worst = d, (cfg["index"], ai, k)
                if d > atol + rtol * abs(rc[k]):
                    fail += 1
                n += 1
    forces_ok = fail == 0 and unmatched == 0
    return {"ok": bool(scalar_ok and forces_ok), "scalar_ok": bool(scalar_ok), "port_sumsq": p_tot,
            "ref_sumsq": r_tot, "rel": rel, "forces_ok": bool(forces_ok), "n": n,
            "unmatched": unmatched, "fail": fail, "max_abs": max_abs, "worst": worst}
