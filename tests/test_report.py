"""Evaluation report (src/cli/evaluate_report.cpp:21-93) and the forces-only
gate of the reference's parity harness (tests/integration/harness/
diff_eval.py:95-163), on the CPU: a hand-computed report, the identity
total_sumsq == |F(x)|^2 against the oracle's residual, and the gate's
tolerance rule."""
import json

import numpy as np
import pytest

import oracle
from forcesmith_amd import datasets as dsm
from forcesmith_amd import io as fio
from forcesmith_amd import report as rp
from forcesmith_amd.model import Dataset


def tiny_dataset():
    off = np.array([0, 2, 3], dtype=np.int64)
    types = np.zeros(3, dtype=np.int32)
    pos = np.zeros((3, 3))
    ref_force = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 3.0]])
    cell = np.stack([np.eye(3) * 10.0] * 2)
    ref_energy = np.array([-1.0, -2.0])
    # Voigt order xx, yy, zz, xy, xz, yz
    ref_stress = np.array([[1.0, 2.0, 3.0, 4.0, 5.0, 6.0], [0.0] * 6])
    return Dataset(off, types, pos, ref_force, cell, ref_energy, ref_stress, "tiny")


def test_hand_computed_report():
    ds = tiny_dataset()
    F = ds.ref_force + np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 2.0], [0.0, 0.0, 0.0]])
    E = np.array([-0.5, -2.0])
    S = np.zeros((2, 3, 3))
    S[0] = [[11.0, 14.0, 15.0], [14.0, 12.0, 16.0], [15.0, 16.0, 13.0]]
    L = np.array([0.0, 0.25])
    rep = rp.build_report(ds, F, E, S, L, energy_weight=2.0, stress_weight=0.1, names=["a", "b"])
    c0, c1 = rep["configs"]
    assert c0["calc_stress"] == [11.0, 12.0, 13.0, 14.0, 15.0, 16.0]
    assert c0["ref_stress"] == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    # forces 1 + 4, energy (2 * 0.5)^2, stress 6 * (0.1 * 10)^2
    assert c0["sumsq"] == pytest.approx(5.0 + 1.0 + 6.0, rel=1e-14)
    assert c1["sumsq"] == pytest.approx(0.0625, rel=1e-14)
    assert rep["total_sumsq"] == pytest.approx(12.0625, rel=1e-14)
    assert (c0["name"], c0["natoms"], c1["index"], c1["limit"]) == ("a", 2, 1, 0.25)
    assert c0["atoms"][1] == {"calc_force": [0.0, 2.0, 2.0], "ref_force": [0.0, 2.0, 0.0]}
    assert rep["nconf"] == 2 and rep["energy_weight"] == 2.0 and rep["stress_weight"] == 0.1
    # without a stress weight the stress term drops out
    rep0 = rp.build_report(ds, F, E, S, L, energy_weight=2.0, stress_weight=0.0)
    assert rep0["configs"][0]["sumsq"] == pytest.approx(6.0, rel=1e-14)


@pytest.mark.parametrize("sw", [0.0, 0.5])
def test_total_sumsq_is_the_residual_norm(sw):
    w = dsm.config2(n_configs=4, nmin=32, nmax=80)
    o = oracle.Oracle(w.dataset, w.model, energy_weight=1.0, stress_weight=sw)
    o.prepare()
    x = o.gather()
    cost = o.cost(x)
    F, E, S, L = o.eval_forces()
    rep = rp.build_report(w.dataset, F, E, S, L, 1.0, sw)
    assert rep["total_sumsq"] == pytest.approx(cost, rel=1e-12)
    ro = o.row_offsets()
    r = o.residual(x)
    for c, cfg in enumerate(rep["configs"]):
        seg = r[ro[c]:ro[c + 1]]
        assert cfg["sumsq"] == pytest.approx(float(seg @ seg), rel=1e-12)


def test_report_file_and_names(tmp_path):
    ds = tiny_dataset()
    text = json.dumps([{"name": "first", "X": [1, 0, 0], "Y": [0, 1, 0], "Z": [0, 0, 1], "E": 0,
                        "atoms": []}, {"X": [1, 0, 0], "Y": [0, 1, 0], "Z": [0, 0, 1], "E": 0,
                                       "atoms": []}])
    assert fio.config_names(text) == ["first", ""]

    class Calc:
        dataset = ds

        def eval_forces(self):
            return ds.ref_force.copy(), ds.ref_energy.copy(), np.zeros((2, 3, 3)), np.zeros(2)

    path = tmp_path / "eval.json"
    rep = rp.evaluate_report(Calc(), 1.0, 0.0, names=["p", "q"], path=path)
    back = json.loads(path.read_text())
    assert back == rep
    assert set(back["configs"][0]) == {"index", "name", "natoms", "calc_energy", "ref_energy", "calc_stress",
                                       "ref_stress", "atoms", "limit", "sumsq"}


def test_diff_gate():
    ds = tiny_dataset()
    F = ds.ref_force + 0.5
    args = (np.zeros(2), np.zeros((2, 3, 3)), np.zeros(2), 1.0, 0.0)
    a = rp.build_report(ds, F, *args)
    assert rp.diff_reports(a, a)["ok"]
    # one component off by just under / over atol + rtol |ref|
    for delta, ok in ((0.9e-5 * (1 + 2.5), True), (1.1e-5 * (1 + 2.5), False)):
        G = F.copy()
        G[1, 1] += delta
        d = rp.diff_reports(rp.build_report(ds, G, *args), a)
        assert d["forces_ok"] == ok
        assert d["worst"] == (0, 1, 1)
        assert d["fail"] == (0 if ok else 1)
    # a missing atom is unmatched
    short = json.loads(json.dumps(a))
    short["configs"][1]["atoms"] = []
    d = rp.diff_reports(a, short)
    assert d["unmatched"] == 1 and not d["ok"]
