"""fs_eval_step: the residual + Jacobian step replayed from a captured
hipGraph, against the eager entry points (fs_eval_residual, fs_eval_jacobian)
at the same parameters.

Every replay is checked against an eager evaluation made between the
replays (the graph's kernels are the eager ones with the same launch shapes:
bit-identical); the graph is recaptured when the weights, the launch shape or
the Jacobian tile change; Stillinger-Weber runs the step eagerly.
"""
import numpy as np
import pytest

import forcesmith_amd as fa
from forcesmith_amd import datasets as dsm

pytestmark = pytest.mark.gpu


def _xs(x0, n):
    return [x0 * (1.0 + 0.01 * k * np.cos(np.arange(x0.size) + k)) for k in range(n)]


@pytest.mark.parametrize("name", ["eam3", "eam1", "adp2", "sw"])
def test_replays_equal_eager_evaluations(name):
    w = {"eam3": lambda: dsm.config2(n_configs=6, nmin=32, nmax=300), "eam1": lambda: dsm.config1(n_configs=3),
         "adp2": lambda: dsm.config3(n_configs=2), "sw": lambda: dsm.config4(n_configs=4)}[name]()
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x0 = g.gather_params()
    for k, x in enumerate(_xs(x0, 4)):
        cost, F, J = g.step(x, want_residual=True, want_jacobian=True)
        Fe, ce = g.residual(x, with_cost=True)
        Je = g.jacobian(x)
        print(f"[{name}] replay {k}: cost {cost:.9e}", flush=True)
        assert np.array_equal(F, Fe) and np.array_equal(J, Je) and cost == ce
        assert np.array_equal(g.gather_params(), x)


def _inner_ends_moved(model, dr):
    # the same layout on other grids: the pair and density tables start dr
    # further out (the cutoffs, and so the neighbour list, are unchanged)
    from forcesmith_amd.model import Table, TabulatedModel
    tabs = []
    for (role, _), t in zip(model.table_roles(), model.tables):
        lo = t.rmin + dr if role in ("pair", "density") else t.rmin
        tabs.append(Table(t.n, lo, t.rmax, t.y.copy(), t.fixed.copy()))
    return TabulatedModel(model.kind, model.ntypes, tabs, model.energy_weight, model.stress_weight)


def test_recapture_after_layout_changes():
    # weights, launch shape, Jacobian tile, the form of the pair-table
    # columns, and a model upload on other grids (the knot sums and the
    # sorted pair flags are rebuilt by an eager step before the recapture)
    w = dsm.config2(n_configs=6, nmin=32, nmax=300)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x = g.gather_params()
    g.step(x)
    for change in (lambda: g._check(g._lib.fs_set_weights(g._h, 0.7, 0.2)),
                   lambda: g.set_launch(16, 4), lambda: g.set_jacobian_tile(8),
                   lambda: g.set_jacobian_form(1), lambda: g.set_jacobian_form(0),
                   lambda: g.set_model(_inner_ends_moved(w.model, 0.05))):
        change()
        for xx in _xs(x, 2):
            cost, F, J = g.step(xx, want_residual=True, want_jacobian=True)
            Fe, ce = g.residual(xx, with_cost=True)
            assert np.array_equal(F, Fe) and np.array_equal(J, g.jacobian(xx)) and cost == ce


def test_queued_steps():
    w = dsm.config1(n_configs=3)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    xs = _xs(g.gather_params(), 5)
    g.step(xs[0])
    for x in xs:
        assert g.step(x, want_cost=False)[0] is None
    g.synchronize()
    Fe, ce = g.residual(xs[-1], with_cost=True)
    cost, F, _ = g.step(xs[-1], want_residual=True)
    assert np.array_equal(F, Fe) and cost == ce
