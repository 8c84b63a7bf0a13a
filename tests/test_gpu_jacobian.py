"""GPU analytic Jacobian, FP64 normal equations and LM against the oracle.

Bars: the FP32 Jacobian within max |J - J_oracle| / (1 + |J_oracle|) <= 5e-5
(the reference FD metric, test_param_jacobian.cpp:181-193, at FP32 rounding:
every interval is located from the FP64 distance, measured <= 2.4e-5);
J^T J and J^T F within 1e-4 of the oracle's FP64 products relative to their
largest entry; LM cost history non-increasing and its final cost within 1e-3
of the oracle LM after the same number of iterations.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import forcesmith_amd as fa
import oracle
from forcesmith_amd import datasets as dsm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    w4 = dsm.config4(n_configs=6)
    w4.model.stress_weight = 0.4
    # ADP with pair and density knots freed as well (240 columns of all five
    # table kinds) next to the C3 free set (dipole, quadrupole, embedding)
    w5 = dsm.config3(n_configs=2)
    for t in w5.model.tables[:w5.model.npair + w5.model.ntypes]:
        t.fixed[:] = False
    return {
        "eam1": dsm.config1(n_configs=3),
        "eam3s": dsm.config2(n_configs=6, nmin=32, nmax=300),
        "sw": w4,
        "adp2": dsm.config3(n_configs=3),
        "adp2all": w5,
    }


def metric(a, b):
    return float(np.max(np.abs(a - b) / (1.0 + np.abs(b))))


@pytest.mark.parametrize("name", ["eam1", "eam3s", "sw", "adp2", "adp2all"])
def test_jacobian_matches_oracle(wl, name):
    w = wl[name]
    print(f"jac {name}", flush=True)
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x = o.gather()
    Jo = o.jacobian(x)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    Jg = g.jacobian(x).astype(np.float64)
    m = metric(Jg, Jo)
    print(f"[{name}] J metric {m:.3e} shape {Jg.shape}", flush=True)
    assert Jg.shape == Jo.shape
    assert m <= 5e-5


@pytest.mark.parametrize("name", ["eam1", "eam3s", "adp2", "adp2all"])
def test_jacobian_forms_agree(wl, name):
    # the knot-sum and window forms of the pair-table columns and density
    # rows sum the same FP32 terms in another grouping; a repeat of either is
    # bit-identical
    w = wl[name]
    print(f"jac forms {name}", flush=True)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x = g.gather_params()
    g.set_jacobian_form(1)
    Jw = g.jacobian(x).astype(np.float64)
    g.set_jacobian_form(2)
    Jk = g.jacobian(x).astype(np.float64)
    Jk2 = g.jacobian(x).astype(np.float64)
    m = metric(Jk, Jw)
    print(f"[{name}] knot sums vs windows {m:.3e}", flush=True)
    assert m <= 2e-5
    assert np.array_equal(Jk, Jk2)


@pytest.mark.parametrize("name", ["eam3s", "sw", "adp2"])
def test_jacobian_tile_sizes_agree(wl, name):
    # every atom's force rows come from one wave whatever the tile, so they
    # are bit-identical across tile sizes; the energy, stress and limit rows
    # sum per-tile FP64 partials in another grouping (FP64 rounding only)
    w = wl[name]
    print(f"jac tiles {name}", flush=True)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x = g.gather_params()
    nrow = 3 * np.asarray(w.dataset.config_sizes())
    ro = g.row_offsets()
    force = np.zeros(g.residual_count, bool)
    for c in range(len(nrow)):
        force[ro[c]:ro[c] + nrow[c]] = True
    ref = None
    for jt in (4, 8, 16, 32, 64):
        g.set_jacobian_tile(jt)
        J = g.jacobian(x).astype(np.float64)
        assert g.launch_shape()["jacobian_tile_atoms"] == jt
        if ref is None:
            ref = J
            continue
        assert np.array_equal(J[force], ref[force]), jt
        assert np.max(np.abs(J[~force] - ref[~force]) / (1.0 + np.abs(ref[~force]))) <= 1e-6, jt
    g.set_jacobian_tile(0)
    g.autotune()
    shape = g.launch_shape()
    assert shape["jacobian_tile_atoms"] in (4, 8, 16, 32, 64)
    assert shape["lanes_per_atom"] in (4, 8, 16, 32, 64)
    J = g.jacobian(x).astype(np.float64)
    assert np.array_equal(J[force], ref[force])
    # a tile fixed by hand survives the autotuner
    g.set_jacobian_tile(16)
    g.autotune()
    g.jacobian(x)
    assert g.launch_shape()["jacobian_tile_atoms"] == 16
    with pytest.raises(fa.FsError):
        g.set_jacobian_tile(12)


def test_sw_factors_follow_the_parameters(wl):
    # the SW per-pair factors are shared by the residual and the Jacobian of
    # one parameter vector; a residual at x1 followed by a Jacobian at x2 (and
    # back) must give the Jacobian and residual of x2 and x1
    w = wl["sw"]
    print("sw factor cache", flush=True)
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x1 = g.gather_params()
    x2 = x1 * (1.0 + 0.02 * np.cos(np.arange(x1.size)))
    g.residual(x1)
    J2 = g.jacobian(x2).astype(np.float64)
    assert metric(J2, o.jacobian(x2)) <= 5e-5
    F1 = g.residual(x1)
    assert np.max(np.abs(F1 - o.residual(x1))) <= 1e-5
    J1 = g.jacobian().astype(np.float64)  # same parameters: the kept factors
    assert metric(J1, o.jacobian(x1)) <= 5e-5


def test_jacobian_rows_by_kind(wl):
    # force, energy, stress and limit rows are each exercised
    w = wl["eam3s"]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x = o.gather()
    Jo = o.jacobian(x)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    Jg = g.jacobian(x).astype(np.float64)
    ro = o.row_offsets()
    sizes = w.dataset.config_sizes()
    for c in range(len(sizes)):
        e = ro[c] + 3 * sizes[c]
        for r in range(e, ro[c + 1]):
            assert metric(Jg[r], Jo[r]) <= 5e-5, (c, r - e)


def test_normal_equations(wl):
    w = wl["eam1"]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x = o.gather()
    Jo = o.jacobian(x)
    Fo = o.residual(x)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    A, b, c = g.normal_equations()
    Ao, bo = Jo.T @ Jo, Jo.T @ Fo
    assert np.max(np.abs(A - Ao)) <= 1e-4 * np.max(np.abs(Ao))
    assert np.max(np.abs(b - bo)) <= 1e-4 * np.max(np.abs(bo))
    assert np.array_equal(A, A.T)
    assert abs(c - Fo @ Fo) <= 1e-6 * (Fo @ Fo)


def test_lm_tracks_oracle(wl):
    w = wl["eam1"]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    xo, ho, _ = o.lm(x0, 8)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    xg, hg, acc = g.lm(x0, 8)
    print("oracle", ho, "\ngpu", hg, flush=True)
    assert acc > 0
    assert np.all(np.diff(hg) <= 0)
    assert abs(hg[-1] - ho[-1]) <= 1e-3 * ho[-1]
    assert abs(g.cost(xg) - hg[-1]) <= 1e-6 * hg[-1]


def test_lm_50_iterations_single_element():
    """The single-element FCC workload's run: one residual + Jacobian, then 50
    LM iterations; the GPU and oracle drivers share the algorithm, so the
    cost histories agree to the FP32-Jacobian level and both descend."""
    w = dsm.config1()
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    xo, ho, ao = o.lm(x0, 50)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    xg, hg, ag = g.lm(x0, 50)
    print(f"LM 50: oracle {ho[0]:.6e} -> {ho[-1]:.6e} ({ao} accepted); "
          f"gpu {hg[0]:.6e} -> {hg[-1]:.6e} ({ag} accepted)", flush=True)
    assert np.all(np.diff(hg) <= 0)
    assert hg[-1] < 0.1 * hg[0]
    assert abs(hg[-1] - ho[-1]) <= 1e-2 * ho[-1]
    assert abs(g.cost(xg) - hg[-1]) <= 1e-6 * hg[-1]


def test_lm_stillinger_weber(wl):
    """Batched LM on SW with analytic parameter columns (tangents)."""
    w = wl["sw"]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather() * 1.02
    xo, ho, _ = o.lm(x0, 10)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    xg, hg, acc = g.lm(x0, 10)
    print(f"SW LM: oracle {ho[0]:.6e} -> {ho[-1]:.6e}; gpu {hg[0]:.6e} -> {hg[-1]:.6e}", flush=True)
    assert acc > 0 and np.all(np.diff(hg) <= 0)
    assert abs(hg[-1] - ho[-1]) <= 1e-3 * ho[-1]
    assert np.allclose(xg, xo, rtol=1e-3, atol=1e-6)


def test_lm_adp(wl):
    w = wl["adp2"]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    xo, ho, _ = o.lm(x0, 5)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    xg, hg, acc = g.lm(x0, 5)
    print(f"ADP LM: oracle {ho[0]:.6e} -> {ho[-1]:.6e}; gpu {hg[0]:.6e} -> {hg[-1]:.6e}", flush=True)
    assert acc > 0 and np.all(np.diff(hg) <= 0)
    assert abs(hg[-1] - ho[-1]) <= 1e-3 * ho[-1]


@pytest.mark.parametrize("name", ["eam3s", "adp2all"])
def test_smoothness_rows(wl, name):
    """Curvature rows (smooth_weight > 0): residual rows, constant Jacobian
    rows, the batched population cost and LM, each against the oracle."""
    w = wl[name]
    sw = 0.05
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    o.set_smooth_weight(sw)
    g = fa.GpuCalculator(w.dataset, w.model, smooth_weight=sw)
    g.prepare()
    print(f"smooth {name}: rows {o.smooth_count}", flush=True)
    assert g.residual_count == o.residual_count
    R0 = int(g.row_offsets()[-1])
    assert o.smooth_count > 0 and g.residual_count - R0 == o.smooth_count
    x = o.gather()
    rng = np.random.default_rng(5)
    x1 = x * (1.0 + 0.01 * rng.standard_normal(x.shape))
    Fo, (Fg, cg) = o.residual(x1), g.residual(x1, with_cost=True)
    # curvature rows are FP64 on both sides, same operation order
    assert np.allclose(Fg[R0:], Fo[R0:], rtol=1e-13, atol=1e-15)
    # the model-level hooks (model_smoothness_count / model_write_smoothness)
    assert g.smoothness_count() == o.smooth_count
    assert np.allclose(g.write_smoothness(sw), Fo[R0:], rtol=1e-13, atol=1e-15)
    assert abs(cg - Fo @ Fo) <= 1e-6 * (Fo @ Fo)
    Jo, Jg = o.jacobian(x1), g.jacobian(x1).astype(np.float64)
    assert np.array_equal(Jg[R0:], Jo[R0:].astype(np.float32).astype(np.float64))
    assert metric(Jg, Jo) <= 5e-5
    if name.startswith("eam"):
        X = np.stack([x, x1, x * 0.99])
        co, cp = o.population_cost(X), g.population_cost(X)
        print(f"population {cp} oracle {co}", flush=True)
        assert np.allclose(cp, co, rtol=1e-6)
    xo, ho, _ = o.lm(x, 4)
    xg, hg, acc = g.lm(x, 4)
    print(f"smooth LM: oracle {ho} gpu {hg}", flush=True)
    assert acc > 0 and np.all(np.diff(hg) <= 0)
    assert abs(hg[-1] - ho[-1]) <= 1e-3 * ho[-1]


def test_lm_solve_reference_semantics():
    """fs_lm_solve (EigenLMSolver semantics on the GPU normal equations)
    against the oracle's MINPACK restatement (pivoted QR of the FP64 J) on
    the single-element workload, FP64 residuals. With the pair knots free
    (a well-conditioned fit): same status, evaluation counts within a few,
    final cost within 1e-6. With every table free the EAM gauge freedom makes
    J^T J nearly singular, and the FP32-J and FP64-J runs stop at different
    points of the flat valley: both must still end at a comparable cost."""
    w = dsm.config1(n_configs=3)
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    xo, so, co, ho = o.lm_minpack(x0, max_iter=100)
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    xg, sg, cg, hg = g.lm_solve(x0, max_iter=100)
    print(f"all tables: oracle status {so} counts {co} cost {o.cost(xo):.10e}; "
          f"gpu status {sg} counts {cg} cost {g.cost(xg):.10e}", flush=True)
    assert 1 <= sg <= 3 and 1 <= so <= 3
    assert g.cost(xg) <= 1.02 * o.cost(xo)
    for t in w.model.tables[1:]:
        t.fixed[:] = True
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    xo, so, co, ho = o.lm_minpack(x0, max_iter=100)
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    xg, sg, cg, hg = g.lm_solve(x0, max_iter=100)
    print(f"oracle status {so} counts {co} cost {ho[co[1] - 1]:.10e}; "
          f"gpu status {sg} counts {cg} cost {hg[cg[1] - 1]:.10e}", flush=True)
    assert sg == so and 1 <= sg <= 3
    assert abs(cg[0] - co[0]) <= 3 and abs(cg[1] - co[1]) <= 3
    assert abs(g.cost(xg) - o.cost(xo)) <= 1e-6 * o.cost(xo)
    assert abs(g.cost(xg) - hg[cg[1] - 1]) <= 1e-9 * g.cost(xg)


def test_lm_solve_status_rules():
    # pair knots free; the first calls follow the oracle's costs to the
    # FP32-Jacobian level (the host algebra alone matches the oracle's QR
    # form to rounding with FP64 Jacobians, tests/test_lm_host.py)
    w = dsm.config1(n_configs=3)
    for t in w.model.tables[1:]:
        t.fixed[:] = True
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    # still running after max_iter calls: 2 (solver.cpp:84); the first calls
    # follow the oracle's costs
    xo, so, co, ho = o.lm_minpack(x0, max_iter=3)
    xg, sg, cg, hg = g.lm_solve(x0, max_iter=3)
    print(f"3 calls: oracle {so} {co} {ho}; gpu {sg} {cg} {hg}", flush=True)
    assert so == sg == 2 and cg[1] == co[1] == 3
    assert np.allclose(hg, ho, rtol=1e-4)
    # improper input: 0, nothing evaluated, x untouched
    for bad in (dict(xtol=-1.0), dict(ftol=-1.0), dict(gtol=-1.0), dict(factor=0.0)):
        xb, sb, cb, _ = g.lm_solve(x0, **bad)
        assert sb == 0 and cb == (0, 0, 0) and np.array_equal(xb, x0)
    # an evaluation budget: 5
    xg, sg, cg, _ = g.lm_solve(x0, max_iter=50, maxfev=3)
    assert sg == 5 and cg[0] == 3


def test_lm_solve_stillinger_weber(wl):
    w = wl["sw"]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather() * 1.02
    xo, so, co, ho = o.lm_minpack(x0, max_iter=30)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    xg, sg, cg, hg = g.lm_solve(x0, max_iter=30)
    print(f"SW: oracle {so} {co} {o.cost(xo):.8e}; gpu {sg} {cg} {g.cost(xg):.8e}", flush=True)
    assert 1 <= sg <= 3 or sg == so
    assert g.cost(xg) <= o.cost(xo) * (1 + 1e-3)


def test_normal_equations_solve():
    """NormalEquationsSolver step (fs_normal_equations_solve) against the
    oracle's, pair knots free: the new parameters within the FP32-Jacobian
    level and the same cost after the step; with the automatic ridge and with
    a large one (a scaled gradient step)."""
    w = dsm.config1(n_configs=3)
    for t in w.model.tables[1:]:
        t.fixed[:] = True
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    for ridge in (0.0, 1e3):
        xo = o.ne_solve(x0, ridge)
        xg = g.ne_solve(x0, ridge)
        d = np.linalg.norm(xg - xo) / np.linalg.norm(xo - x0)
        co, cg = o.cost(xo), g.cost(xg)
        print(f"ridge {ridge}: |dx| rel {d:.3e}, cost oracle {co:.10e} gpu {cg:.10e}", flush=True)
        assert d <= 1e-3
        assert abs(cg - co) <= 1e-5 * co


def test_lm_step(wl):
    """fs_lm_step: dx = -(J^T J + lambda I)^-1 J^T F against the oracle's FP64
    normal equations, pair knots free (well conditioned)."""
    w = dsm.config1(n_configs=3)
    for t in w.model.tables[1:]:
        t.fixed[:] = True
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    x0 = o.gather()
    J, F = o.jacobian(x0), o.residual(x0)
    A, gr = J.T @ J, J.T @ F
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    for lam in (1e-3 * np.mean(np.diag(A)), 10.0 * np.mean(np.diag(A))):
        ref = -np.linalg.solve(A + lam * np.eye(len(x0)), gr)
        dx = g.lm_step(lam, x0)
        rel = np.linalg.norm(dx - ref) / np.linalg.norm(ref)
        print(f"lambda {lam:.3e}: |dx - ref| / |ref| = {rel:.3e}", flush=True)
        assert rel <= 1e-4
    assert np.array_equal(g.gather_params(), x0)  # parameters unchanged


def test_jacobian_edge_cases_clamp_empty_and_single_atom():
    # clamped densities (limit rows and their density chain), a one-atom
    # configuration, an empty configuration and a small cluster; every knot free
    from tests.helpers import box, dimer, eam_model, make_dataset, random_cluster
    cfgs = [dimer(2.5), ([[1.0, 1.0, 1.0]], box(20.0)), (np.zeros((0, 3)), box(10.0)),
            (random_cluster(np.random.default_rng(3), 6, 5.0), box(30.0))]
    ds = make_dataset(cfgs)
    m = eam_model(lambda x: 3.0 * np.exp(-2.0 * x), lambda x: 2.0 - 0.3 * x, lambda x: 1.5 - 2.0 * x,
                  emb_span=(1.0, 5.0))
    print("jac edge cases", flush=True)
    o = oracle.Oracle(ds, m, 1.0, 0.3)
    o.prepare()
    x = o.gather()
    Jo = o.jacobian(x)
    g = fa.GpuCalculator(ds, m, energy_weight=1.0, stress_weight=0.3)
    g.prepare()
    Jg = g.jacobian(x).astype(np.float64)
    assert Jg.shape == Jo.shape
    assert metric(Jg, Jo) <= 5e-5
    # the limit rows carry the clamp chain (the lone atom's rho = 0 is clamped)
    ro = o.row_offsets()
    lim = [ro[c + 1] - 1 for c in range(len(ro) - 1)]
    assert np.any(Jo[lim] != 0.0)
    assert metric(Jg[lim], Jo[lim]) <= 5e-5


def test_adp_two_lane_slots_per_kind():
    # 30-knot dipole and quadrupole tables, every knot free: 90 columns per
    # kind, two lane slots each in the moment and force passes
    w = dsm.config3(n_configs=2)
    m = w.model
    first = m.npair + 2 * m.ntypes  # first dipole table
    for k in range(first, len(m.tables)):
        t = m.tables[k]
        xn = np.linspace(t.rmin, t.rmax, 30)
        m.tables[k] = type(t)(30, t.rmin, t.rmax, CubicSpline(t.x, t.y)(xn), np.zeros(30, dtype=bool))
    print("jac adp wide maps", flush=True)
    o = oracle.Oracle(w.dataset, m)
    o.prepare()
    x = o.gather()
    assert x.size == 6 * 30 + sum(t.free_count for t in m.tables[m.npair + m.ntypes:first])
    Jo = o.jacobian(x)
    g = fa.GpuCalculator(w.dataset, m)
    g.prepare()
    Jg = g.jacobian(x).astype(np.float64)
    mt = metric(Jg, Jo)
    print(f"[adp wide] J metric {mt:.3e} shape {Jg.shape}", flush=True)
    assert mt <= 5e-5
