"""Oracle makima spline against the reference's spline known answers
(tests/unit/test_spline.cpp:13-196) and scipy's independent makima."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "makima.npz")


def ev(x, y, r, cached=False):
    return oracle.spline_eval(np.array(x, float), np.array(y, float), r, cached)


def test_eval_at_knots_equals_values():  # test_spline.cpp:13-21
    xs, ys = [0.0, 1.5, 3.0, 5.0, 6.0], [2.0, 0.5, 3.0, 1.0, 4.0]
    for x, y in zip(xs, ys):
        assert abs(ev(xs, ys, x)[0] - y) < 1e-10


def test_linear_function_exact():  # :23-30
    a, b = 3.0, -1.5
    xs = [0.0, 1.0, 2.5, 4.0]
    ys = [b, a + b, 2.5 * a + b, 4.0 * a + b]
    for r in (0.5, 1.75, 3.0):
        assert abs(ev(xs, ys, r)[0] - (a * r + b)) < 1e-10


def test_linear_deriv_exact():  # :32-39
    s = 2.5
    xs = [0.0, 1.0, 3.0, 5.0]
    ys = [0.0, s, 3.0 * s, 5.0 * s]
    for r in (0.5, 2.0, 4.0):
        assert abs(ev(xs, ys, r)[1] - s) < 1e-10


def test_deriv_consistent_with_eval():  # :42-50
    xs, ys = [0.0, 1.0, 2.0, 3.0, 4.0], [0.0, 1.0, 4.0, 9.0, 16.0]
    eps = 1e-5
    for r in (0.5, 1.5, 2.5, 3.5):
        fd = (ev(xs, ys, r + eps)[0] - ev(xs, ys, r - eps)[0]) / (2 * eps)
        assert abs(ev(xs, ys, r)[1] - fd) < 1e-5


def test_boundary_extrapolation():  # :54-64
    xs, ys = [1.0, 2.0, 3.0], [5.0, 3.0, 1.0]
    v, d = ev(xs, ys, 0.0)
    assert abs(v - 7.0) < 1e-15 and abs(d + 2.0) < 1e-15
    v, d = ev(xs, ys, 4.0)
    assert abs(v + 1.0) < 1e-15 and abs(d + 2.0) < 1e-15


def test_two_knot_linear_case():  # :73-77
    v, d = ev([0.0, 1.0], [0.0, 1.0], 0.5)
    assert abs(v - 0.5) < 1e-15 and abs(d - 1.0) < 1e-15


def test_monotone_data_monotone_spline():  # :79-87
    xs = [0.0, 1.0, 2.0, 3.0]
    for k in range(1, 10):
        r = 0.3 * k
        assert ev(xs, xs, r + 0.05)[0] > ev(xs, xs, r - 0.05)[0]


def test_construct_from_knots():  # :89-95
    assert abs(ev([1.0, 2.0, 3.0], [2.0, 0.0, 2.0], 2.0)[0]) < 1e-10


@pytest.mark.parametrize("xs,ys,rs", [
    ([0.0, 1.0, 2.0, 3.0, 4.0, 5.0], [2.0, 0.5, 3.0, 1.0, 4.0, 2.5], [-0.3, 0.0, 0.4, 1.0, 2.7, 3.9999, 5.0, 5.6]),
    ([1.0, 2.0, 3.0], [5.0, 3.0, 1.0], [0.0, 1.5, 2.5, 4.0]),
])
def test_cached_site_matches_direct(xs, ys, rs):  # :101-121
    for r in rs:
        vc, dc = ev(xs, ys, r, cached=True)
        v, d = ev(xs, ys, r)
        assert abs(vc - v) < 1e-11 and abs(dc - d) < 1e-11


def test_site_survives_new_values():  # :126-141 (sites depend on x only)
    xs = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0])
    y2 = np.array([1.0, -2.0, 0.5, 3.5, -1.0, 2.0])
    for r in (0.4, 2.7, 3.5):
        vc, dc = ev(xs, y2, r, cached=True)
        v, d = ev(xs, y2, r)
        assert abs(vc - v) < 1e-11 and abs(dc - d) < 1e-11


def test_interior_matches_scipy_makima():
    g = np.load(GOLD)
    for k in range(int(g["ncases"])):
        x, y = g[f"x{k}"], g[f"y{k}"]
        for r, v, d in zip(g[f"r{k}"], g[f"v{k}"], g[f"d{k}"]):
            vo, do = ev(x, y, r)
            assert abs(vo - v) <= 1e-12 * (1 + abs(v)), (k, r)
            if r <= x[0] or r >= x[-1]:
                # the reference's r <= x0 / r >= xn branch returns the end
                # secant, not the Hermite end slope (spline.cpp:236-241)
                i = 0 if r <= x[0] else len(x) - 2
                assert do == (y[i + 1] - y[i]) / (x[i + 1] - x[i])
            else:
                assert abs(do - d) <= 1e-10 * (1 + abs(d)), (k, r)


def test_slope_jacobian_matches_fd():
    rng = np.random.default_rng(7)
    for n in (4, 5, 8, 20):
        x = np.sort(rng.uniform(0, 5, n))
        y = rng.normal(size=n)
        S = oracle.spline_slope_jacobian(x, y)
        h = 1e-7
        for m in range(n):
            yp, ym = y.copy(), y.copy()
            yp[m] += h
            ym[m] -= h
            fd = (oracle.spline_slopes(x, yp) - oracle.spline_slopes(x, ym)) / (2 * h)
            assert np.allclose(S[:, m], fd, atol=1e-5, rtol=1e-5)


def test_param_grad_matches_fd():
    rng = np.random.default_rng(11)
    x = np.linspace(1.0, 4.0, 9)
    y = rng.normal(size=9)
    for r in (0.5, 1.0, 1.3, 2.2, 3.99, 4.0, 4.5):
        dv, dd, d2 = oracle.spline_param_grad(x, y, r)
        h = 1e-7
        for m in range(9):
            yp, ym = y.copy(), y.copy()
            yp[m] += h
            ym[m] -= h
            vp, dp = oracle.spline_eval(x, yp, r)
            vm, dm = oracle.spline_eval(x, ym, r)
            assert abs(dv[m] - (vp - vm) / (2 * h)) < 1e-6
            assert abs(dd[m] - (dp - dm) / (2 * h)) < 1e-5
        if x[0] < r < x[-1]:
            e = 1e-6
            fd2 = (oracle.spline_eval(x, y, r + e)[1] - oracle.spline_eval(x, y, r - e)[1]) / (2 * e)
            assert abs(d2 - fd2) < 1e-4 * (1 + abs(d2))
