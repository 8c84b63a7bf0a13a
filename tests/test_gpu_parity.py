"""HIP path against the FP64 oracle, through the C ABI.

Bars (written here, per quantity):
  neighbour list: bit-exact (same pairs, same order, identical FP64 displacement bits)
  mixed precision: |dF| <= 1e-5 eV/A; energy |dE| <= 1e-6 |E|; stress
      max|ds| <= 1e-6 * max|s| per configuration (+1e-12); limit 1e-6 rel;
      cost |dc| <= 1e-6 c
  FP64 mode: every quantity within 1e-11 relative (same scales).
"""
import numpy as np
import pytest

import forcesmith_amd as fa
import oracle
from forcesmith_amd import datasets as dsm
from forcesmith_amd.model import Dataset, SWModel
from tests.helpers import box, dimer, eam_model, make_dataset, random_cluster

pytestmark = pytest.mark.gpu


def _print(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def workloads():
    ws = {
        "eam1": dsm.config1(n_configs=4),
        "eam3": dsm.config2(n_configs=12, nmin=32, nmax=400),
        "adp2": dsm.config3(n_configs=3),
        "sw": dsm.config4(n_configs=8),
    }
    ws["sw"].model.stress_weight = 0.5
    return ws


def pair_oracle(w, **kw):
    o = oracle.Oracle(w.dataset, w.model, **kw)
    o.prepare()
    return o


def compare(tag, o, g, mixed):
    Fo, Eo, So, Lo = o.eval_forces()
    Fg, Eg, Sg, Lg = g.eval_forces()
    x = o.gather()
    co = o.cost(x)
    cg = g.cost()
    fmax = np.abs(Fg - Fo).max()
    erel = np.max(np.abs(Eg - Eo) / np.abs(Eo))
    sscale = np.abs(So).reshape(len(So), -1).max(1)[:, None, None] + 1e-12
    srel = np.max(np.abs(Sg - So) / sscale)
    lrel = np.max(np.abs(Lg - Lo) / (np.abs(Lo) + 1e-12)) if np.any(Lo) else np.max(np.abs(Lg))
    crel = abs(cg - co) / co
    _print(f"[{tag}] |dF|max={fmax:.3e} dE/E={erel:.3e} ds/s={srel:.3e} dL={lrel:.3e} dcost/cost={crel:.3e}")
    if mixed:
        assert fmax <= 1e-5
        assert erel <= 1e-6
        assert srel <= 1e-6
        assert lrel <= 1e-6
        assert crel <= 1e-6
    else:
        assert np.max(np.abs(Fg - Fo) / (1 + np.abs(Fo))) <= 1e-11
        assert erel <= 1e-11 and srel <= 1e-11 and crel <= 1e-11
        assert lrel <= 1e-11


@pytest.mark.parametrize("name", ["eam1", "eam3", "adp2", "sw"])
def test_neighbor_list_bit_exact(workloads, name):
    w = workloads[name]
    _print(f"nl {name}")
    o = oracle.Oracle(w.dataset, w.model)
    g = fa.GpuCalculator(w.dataset, w.model)
    if name == "sw":
        # SW lists carry a 5 % skin on the GPU; compare both at one explicit cutoff
        rc = w.model.max_cutoff()
        o.build_nl(rc * 1.05)
    else:
        o.prepare()
    g.prepare()
    oo, oj, oi, od = o.neighbor_list()
    go, gj, gi, gd = g.neighbor_list()
    assert np.array_equal(oo, go)
    assert np.array_equal(oj, gj)
    assert np.array_equal(oi, gi)
    assert np.array_equal(od.view(np.uint64), gd.view(np.uint64))


def test_neighbor_list_small_and_triclinic_boxes():
    rng = np.random.default_rng(9)
    cfgs = []
    for L, tri in ((6.0, False), (8.1, True), (3.0, False), (12.0, True)):
        cell = np.eye(3) * L
        if tri:
            cell[1, 0] = 0.3 * L
            cell[2, 0] = -0.2 * L
            cell[2, 1] = 0.25 * L
        n = int(rng.integers(1, 30))
        cfgs.append((rng.uniform(-0.5, 1.5, size=(n, 3)) @ cell, cell))
    ds = make_dataset(cfgs)
    m = eam_model(lambda x: np.exp(-x), lambda x: np.exp(-x), lambda x: -np.sqrt(x), span=(0.5, 5.5),
                  emb_span=(0.01, 50.0))
    o = oracle.Oracle(ds, m)
    o.prepare()
    g = fa.GpuCalculator(ds, m)
    g.prepare()
    for a, b in zip(o.neighbor_list(), g.neighbor_list()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", ["eam1", "eam3", "adp2", "sw"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_forces_energy_stress_cost(workloads, name, precision):
    w = workloads[name]
    _print(f"eval {name} {precision}")
    o = pair_oracle(w)
    g = fa.GpuCalculator(w.dataset, w.model, precision=precision)
    g.prepare()
    compare(f"{name}/{precision}", o, g, precision == "mixed")


@pytest.mark.parametrize("name", ["eam1", "adp2"])
def test_residual_rows_match(workloads, name):
    w = workloads[name]
    o = pair_oracle(w)
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    x = o.gather()
    Fo = o.residual(x)
    Fg = g.residual(x)
    assert Fo.shape == Fg.shape
    assert np.max(np.abs(Fg - Fo) / (1 + np.abs(Fo))) <= 1e-11
    assert np.array_equal(o.row_offsets(), g.row_offsets())


def test_scatter_params_then_evaluate(workloads):
    w = workloads["eam3"]
    o = pair_oracle(w)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    rng = np.random.default_rng(1)
    x = o.gather() * (1 + 0.05 * rng.normal(size=o.param_count))
    co = o.cost(x)
    cg = g.cost(x)
    assert abs(cg - co) <= 1e-6 * co
    assert np.array_equal(g.gather_params(), x)


def test_eval_cached_matches_eval_forces(workloads):
    w = workloads["eam3"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    F, E, S, L = g.eval_forces()
    for c in (0, 5, 11):
        a0, a1 = w.dataset.atom_off[c], w.dataset.atom_off[c + 1]
        f, e, s, l = g.eval_cached(c)
        assert np.array_equal(f, F[a0:a1]) and e == E[c] and np.array_equal(s, S[c]) and l == L[c]


def test_deterministic(workloads):
    w = workloads["eam3"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    a = g.residual(w.model.gather())
    b = g.residual(w.model.gather() + 0.0)
    assert np.array_equal(a, b)


def test_launch_shapes_agree(workloads):
    w = workloads["eam3"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    ref = None
    for lpa in (4, 8, 16, 32, 64):
        g.set_launch(lpa, 4)
        F, c = g.residual(w.model.gather(), with_cost=True)
        if ref is None:
            ref = (F, c)
        assert np.max(np.abs(F - ref[0])) < 1e-5
        assert abs(c - ref[1]) <= 1e-7 * ref[1]
    lpa = g.autotune()
    assert lpa in (4, 8, 16, 32, 64)


def test_edge_cases_clamp_empty_and_single_atom():
    # clamped density (limit row), a one-atom config, an empty config, a dimer
    cfgs = [dimer(2.5), ([[1.0, 1.0, 1.0]], box(20.0)), (np.zeros((0, 3)), box(10.0)),
            (random_cluster(np.random.default_rng(3), 6, 5.0), box(30.0))]
    ds = make_dataset(cfgs)
    m = eam_model(lambda x: 3.0 * np.exp(-2.0 * x), lambda x: 2.0 - 0.3 * x, lambda x: 1.5 - 2.0 * x,
                  emb_span=(1.0, 5.0))
    o = oracle.Oracle(ds, m, 1.0, 0.3)
    o.prepare()
    for prec in ("mixed", "fp64"):
        g = fa.GpuCalculator(ds, m, precision=prec, energy_weight=1.0, stress_weight=0.3)
        g.prepare()
        Fo = o.residual(o.gather())
        Fg = g.residual()
        tol = 1e-5 if prec == "mixed" else 1e-11
        assert np.max(np.abs(Fg - Fo) / (1 + np.abs(Fo))) <= tol
        _, _, _, L = g.eval_forces()
        assert L[1] > 0.0  # lone atom: rho = 0 is clamped to the span start
        # the batched population cost on the same configurations (member-lane
        # passes: FP64 in either precision mode)
        x0 = o.gather()
        X = np.stack([x0, 1.01 * x0, 0.98 * x0])
        assert np.allclose(g.population_cost(X), o.population_cost(X), rtol=1e-10, atol=0)


def test_exact_edge_pairs_follow_linear_sites():
    # a pair exactly at the span end uses the linear site (secant slope),
    # mixed precision must route it through the FP64 path
    xn = float(np.float64(5.5))
    cfgs = [([[0.0, 0.0, 0.0], [xn, 0.0, 0.0], [0.0, 2.7, 0.0]], box(40.0))]
    ds = make_dataset(cfgs)
    m = eam_model(lambda x: np.sin(x), lambda x: np.exp(-x), lambda x: -np.sqrt(x), n=12, span=(1.0, 5.5),
                  emb_span=(0.01, 3.0))
    o = oracle.Oracle(ds, m)
    o.prepare()
    g = fa.GpuCalculator(ds, m)
    g.prepare()
    Fo = o.eval_forces()[0]
    Fg = g.eval_forces()[0]
    assert np.abs(Fg - Fo).max() <= 1e-5
    # the population passes take the same linear site for that pair
    x0 = o.gather()
    X = np.stack([x0, 1.03 * x0, 0.97 * x0])
    assert np.allclose(g.population_cost(X), o.population_cost(X), rtol=1e-10, atol=0)


@pytest.mark.parametrize("name", ["eam3", "adp2", "sw"])
def test_evaluate_report_passes_the_force_gate(workloads, name):
    # the report built from the device evaluation against the oracle's, through
    # the reference harness's forces-only gate (rtol = atol = 1e-5); its
    # total_sumsq is the device cost |F|^2
    from forcesmith_amd import report as rp
    w = workloads[name]
    ew, sw = w.model.energy_weight, w.model.stress_weight
    o = pair_oracle(w)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    rg = rp.evaluate_report(g, ew, sw)
    ro = rp.build_report(w.dataset, *o.eval_forces(), ew, sw)
    d = rp.diff_reports(rg, ro)
    _print(f"[report {name}] n={d['n']} max|dF|={d['max_abs']:.3e} rel={d['rel']:.3e}")
    assert d["ok"], d
    assert rg["total_sumsq"] == pytest.approx(g.cost(), rel=1e-6)
    assert rg["total_sumsq"] == pytest.approx(ro["total_sumsq"], rel=1e-6)
