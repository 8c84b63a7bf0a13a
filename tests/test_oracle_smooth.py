"""Curvature (smoothness) rows of the oracle, pinned on the CPU.

The rows are w (y[k-1] - 2 y[k] + y[k+1]), k = 1..n-2 (spline.cpp:205-213),
for every table with a free knot and n >= 3 (spline.hpp:104-106), tables in
declaration order, after the configuration rows
(forcesmith_functor.cpp:53-55). Checks: the closed form 2 c h^2 for quadratic
knot values on a uniform grid (a wrong sign, offset or stencil fails), the
count rules, and the analytic Jacobian rows against central differences of the
residual.
"""
import numpy as np
import pytest

import oracle
from forcesmith_amd import datasets as dsm
from forcesmith_amd.model import Table, TabulatedModel, knot_grid
from tests.helpers import box, make_dataset


def quad_table(n, rmin, rmax, a, b, c, fixed=None):
    x = knot_grid(n, rmin, rmax)
    return Table(n, rmin, rmax, a + b * x + c * x * x,
                 np.zeros(n, dtype=bool) if fixed is None else np.asarray(fixed, dtype=bool))


def small_dataset():
    pos = [[0.0, 0.0, 0.0], [2.3, 0.0, 0.0], [0.0, 2.5, 0.3]]
    return make_dataset([(pos, box(30.0))])


def test_quadratic_tables_give_constant_rows():
    ds = small_dataset()
    # phi, rho, F with distinct curvatures c and sizes n
    specs = [(12, 1.0, 6.0, 0.3, -0.2, 0.7), (9, 1.0, 6.0, 1.0, 0.5, -0.4), (7, 0.0, 4.0, -1.0, 0.1, 2.5)]
    m = TabulatedModel("eam", 1, [quad_table(*s) for s in specs])
    o = oracle.Oracle(ds, m)
    o.prepare()
    w = 0.37
    o.set_smooth_weight(w)
    assert o.smooth_count == sum(n - 2 for n, *_ in specs)
    R0 = o.row_offsets()[-1]
    assert o.residual_count == R0 + o.smooth_count
    r = o.residual(o.gather())
    row = R0
    for n, rmin, rmax, a, b, c in specs:
        h = (rmax - rmin) / (n - 1)
        blk = r[row:row + n - 2]
        assert np.allclose(blk, w * 2.0 * c * h * h, rtol=1e-12, atol=1e-13)
        row += n - 2
    assert row == len(r)


def test_count_rules():
    ds = small_dataset()
    fixed_all = np.ones(8, dtype=bool)
    tabs = [quad_table(8, 1.0, 6.0, 0, 0, 1.0, fixed=fixed_all),  # no free knot: no rows
            quad_table(2, 1.0, 6.0, 0, 1.0, 0),                    # n < 3: no rows
            quad_table(5, 0.0, 4.0, 0, 0, 1.0, fixed=[1, 0, 1, 1, 1])]
    m = TabulatedModel("eam", 1, tabs)
    o = oracle.Oracle(ds, m)
    o.prepare()
    assert o.smooth_count == 0  # weight 0
    o.set_smooth_weight(1.0)
    assert o.smooth_count == 3
    # the fixed knots still enter the value: rows of y = x^2 on h = 1
    r = o.residual(o.gather())
    assert np.allclose(r[-3:], 2.0, rtol=1e-13)
    sw = dsm.config4(n_configs=2)
    osw = oracle.Oracle(sw.dataset, sw.model)
    osw.set_smooth_weight(1.0)
    assert osw.smooth_count == 0


@pytest.mark.parametrize("name", ["eam", "adp"])
def test_analytic_rows_match_central_differences(name):
    w = dsm.config1(n_configs=1) if name == "eam" else dsm.config3(n_configs=1)
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    o.set_smooth_weight(0.25)
    x = o.gather()
    R0 = o.row_offsets()[-1]
    J = o.jacobian(x)
    Jfd = o.jacobian_fd(x)
    S, Sfd = J[R0:], Jfd[R0:]
    assert S.shape[0] == o.smooth_count > 0
    assert np.max(np.abs(S - Sfd)) <= 1e-8
    # each row holds w (1, -2, 1) on free knots only
    assert set(np.round(np.unique(S), 12)) <= {0.0, 0.25, -0.5}
    assert np.all(np.count_nonzero(S, axis=1) <= 3)
