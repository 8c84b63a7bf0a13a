"""Jacobian column layouts beyond the throughput workload's, against the
oracle's analytic Jacobian (itself pinned to the reference's central
differences, fd_jacobian.hpp:37-49).

The Jacobian passes give each lane of a wave one column per lane slot; the
slot count of a pass follows the free columns of its table kinds, and the
density-chain windows have a balanced form only when every density table has
the same number of free knots and they fit one slot. The cases here take each
other branch:
  - 150 pair columns (four pair slots);
  - 75 density columns (more than one slot, so the unbalanced windows);
  - density + embedding columns in four slots, and in six;
  - density tables with unequal free-knot counts (one table partly fixed).
Each case runs both forms of the pair-table columns and density rows (window
walks and knot sums, fs_set_jacobian_form). A pair table of 150 knots is
beyond the knot sums' width: there the knot-sum form is refused and the
default falls back to the windows.

Bars: max |J - J_oracle| / (1 + |J_oracle|) <= 5e-5 and
|J - J_oracle|_F <= 1e-6 |J_oracle|_F.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import forcesmith_amd as fa
import oracle
from forcesmith_amd import datasets as dsm
from forcesmith_amd.model import Table, TabulatedModel

pytestmark = pytest.mark.gpu


def resample(t: Table, n: int, free: bool) -> Table:
    # through a smooth cubic: linear interpolation would put runs of knots on
    # one line, where the makima weights |d_k+1 - d_k| sit at their kink and
    # the Jacobian is undefined
    x = np.linspace(t.rmin, t.rmax, n)
    return Table(n, t.rmin, t.rmax, CubicSpline(t.x, t.y)(x), np.full(n, not free))


def relayout(model: TabulatedModel, spec):
    """spec: per role ('pair', 'density', 'embedding') (knots or None to keep,
    free)."""
    tabs = []
    for (role, _), t in zip(model.table_roles(), model.tables):
        n, free = spec[role]
        if n is None:
            tabs.append(Table(t.n, t.rmin, t.rmax, t.y.copy(), np.full(t.n, not free)))
        else:
            tabs.append(resample(t, n, free))
    return TabulatedModel(model.kind, model.ntypes, tabs, model.energy_weight, model.stress_weight)


def check(ds, model, tag, forms=(1, 2)):
    o = oracle.Oracle(ds, model)
    o.prepare()
    x = o.gather()
    Jo = o.jacobian(x)
    g = fa.GpuCalculator(ds, model)
    g.prepare()
    for form in forms:
        g.set_jacobian_form(form)
        Jg = g.jacobian(x).astype(np.float64)
        assert Jg.shape == Jo.shape
        m = float(np.max(np.abs(Jg - Jo) / (1.0 + np.abs(Jo))))
        fro = float(np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo))
        print(f"[{tag}] form {form} D={Jg.shape[1]} J metric {m:.3e} relative Frobenius {fro:.3e}",
              flush=True)
        assert m <= 5e-5
        assert fro <= 1e-6
    Fo = o.residual(x)
    Fg = g.residual(x)
    assert np.max(np.abs(Fg - Fo) / (1 + np.abs(Fo))) <= 1e-5
    return g, x


@pytest.fixture(scope="module")
def ternary():
    return dsm.config2(n_configs=5, nmin=32, nmax=300)


def test_four_pair_slots_and_unbalanced_density_windows(ternary):
    m = relayout(ternary.model, {"pair": (25, True), "density": (25, True), "embedding": (None, False)})
    assert m.param_count == 225
    check(ternary.dataset, m, "pair 150 + density 75")


def test_four_density_embedding_slots(ternary):
    m = relayout(ternary.model, {"pair": (None, False), "density": (25, True), "embedding": (25, True)})
    assert m.param_count == 150
    check(ternary.dataset, m, "density 75 + embedding 75")


def test_unequal_free_density_counts(ternary):
    m = relayout(ternary.model, {"pair": (None, True), "density": (None, True), "embedding": (None, True)})
    dens = m.tables[m.npair + 1]
    dens.fixed[:10] = True
    assert m.param_count == 230
    check(ternary.dataset, m, "density free 20/10/20")


def test_six_density_embedding_slots():
    w = dsm.config1(n_configs=3)
    m = relayout(w.model, {"pair": (None, False), "density": (100, True), "embedding": (150, True)})
    assert m.param_count == 250
    check(w.dataset, m, "single element density 100 + embedding 150")


def test_knot_sums_refused_beyond_their_width():
    w = dsm.config1(n_configs=2)
    m = relayout(w.model, {"pair": (150, True), "density": (None, False), "embedding": (None, False)})
    assert m.param_count == 150
    g, x = check(w.dataset, m, "single element pair 150", forms=(1, 0))
    g.set_jacobian_form(2)
    with pytest.raises(fa.FsError):
        g.jacobian(x)
