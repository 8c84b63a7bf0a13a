"""EAM gauge on the GPU (fs_eam_* entry points, the kernels' per-type
gauge scalars) against the oracle's restatement of rescale_eam and of the
wrapped tables (src/core/rescale.cpp, include/forcesmith/core/
scale_potentials.hpp), through the C ABI.

Bars (written here):
  rescale_eam: scale a, gauge A and reference densities within 1e-12
      relative; slopes and intercepts within 1e-10 (they are spline values at
      the reference densities, which are sums over every atom);
  residual: the parity bars of the ungauged paths (mixed: |dF| <= 1e-5,
      energy, stress, limit and cost 1e-6 relative; FP64 1e-11);
  Jacobian: the reference FD metric max |a - b| / (1 + |b|) <= 5e-5 (FP32
      storage), Frobenius norm of the difference <= 1e-6 of the oracle's;
  population costs: 1e-6 relative (mixed);
  captured step: equal to the eager residual to 1e-12 relative after the
      gauge changes between replays.
"""
import numpy as np
import pytest

import forcesmith_amd as fa
import oracle
from forcesmith_amd import datasets as dsm
from forcesmith_amd import spline as hspline
from forcesmith_amd.model import Table, TabulatedModel
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def _print(msg):
    print(msg, flush=True)


def metric(a, b):
    return float(np.max(np.abs(a - b) / (1.0 + np.abs(b))))


@pytest.fixture(scope="module")
def wl():
    return {"eam1": dsm.config1(n_configs=4), "eam3": dsm.config2(n_configs=12, nmin=32, nmax=400)}


def _oracle(w, model=None):
    o = oracle.Oracle(w.dataset, w.model if model is None else model)
    o.prepare()
    return o


def narrowed(w):
    """The workload's model with each embedding table resampled onto a span
    whose ends fall midway between atoms' densities at the 20th and 75th
    percentiles: a quarter of the atoms clamp low, a quarter high."""
    o = _oracle(w)
    o.eval_forces()
    sigma = o.scratch()[0]
    types = np.asarray(w.dataset.types)
    m = w.model
    tabs = [Table(t.n, t.rmin, t.rmax, t.y.copy(), t.fixed.copy()) for t in m.tables]
    for t in range(m.ntypes):
        v = np.sort(sigma[types == t])
        k0, k1 = int(0.2 * (len(v) - 1)), int(0.75 * (len(v) - 1))
        lo, hi = 0.5 * (v[k0] + v[k0 + 1]), 0.5 * (v[k1] + v[k1 + 1])
        e = m.tables[m.npair + m.ntypes + t]
        s = hspline.makima_slopes(e.x, e.y)
        xs = np.linspace(lo, hi, e.n)
        ys = np.array([hspline.evaluate(e.x, e.y, s, float(r)) for r in xs])
        tabs[m.npair + m.ntypes + t] = Table(e.n, lo, hi, ys, e.fixed.copy())
    return TabulatedModel("eam", m.ntypes, tabs, m.energy_weight, m.stress_weight)


def gauge_of(ntypes):
    S = np.array([-0.4, 0.3, -0.15][:ntypes])
    F0 = np.array([0.6, -0.15, 0.25][:ntypes])
    return 1.25, S, F0


@pytest.mark.parametrize("name", ["eam1", "eam3"])
def test_rescale_eam_matches_oracle(wl, name):
    w = wl[name]
    _print(f"rescale {name}")
    o = _oracle(w)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    ao, ro = o.rescale_eam()
    ag, rg = g.rescale_eam()
    Ao, So, Fo = o.gauge()
    Ag, Sg, Fg = g.gauge()
    _print(f"[{name}] a {ag:.15g} / {ao:.15g}; rho_ref {rg} / {ro}; S {Sg} / {So}; F0 {Fg} / {Fo}")
    assert ao != 1.0
    assert abs(ag - ao) <= 1e-12 * abs(ao) and abs(Ag - Ao) <= 1e-12 * abs(Ao)
    assert np.allclose(rg, ro, rtol=1e-12, atol=0)
    assert np.allclose(Sg, So, rtol=1e-10, atol=1e-13)
    assert np.allclose(Fg, Fo, rtol=1e-10, atol=1e-13)
    # both evaluate with their own gauge from here; the residuals agree
    compare(f"{name} rescaled", o, g, True)
    # the reference's postconditions on the GPU's own densities
    ref = g.compute_rho_ref()
    for t in range(w.model.ntypes):
        F, dF = o.embedding(t, ref[t])
        assert abs(F) < 1e-10 and abs(dF) < 1e-10


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_gauged_residual_with_clamped_atoms(wl, precision):
    w = wl["eam3"]
    m = narrowed(w)
    A, S, F0 = gauge_of(m.ntypes)
    _print(f"gauged residual {precision}")
    o = _oracle(w, m)
    o.set_gauge(A, S, F0)
    g = fa.GpuCalculator(w.dataset, m, precision=precision)
    g.prepare()
    g.set_gauge(A, S, F0)
    L = o.eval_forces()[3]
    assert np.count_nonzero(L) >= len(L) // 2  # clamped atoms in most configurations
    compare(f"eam3 narrowed gauge {precision}", o, g, precision == "mixed")


def test_gauged_jacobian(wl):
    w = wl["eam3"]
    m = narrowed(w)
    A, S, F0 = gauge_of(m.ntypes)
    _print("gauged jacobian")
    o = _oracle(w, m)
    o.set_gauge(A, S, F0)
    x = o.gather()
    Jo = o.jacobian(x)
    g = fa.GpuCalculator(w.dataset, m)
    g.prepare()
    g.set_gauge(A, S, F0)
    Jg = g.jacobian(x).astype(np.float64)
    mt = metric(Jg, Jo)
    fro = np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo)
    _print(f"[gauge] J metric {mt:.3e} frobenius {fro:.3e} shape {Jg.shape}")
    assert mt <= 5e-5 and fro <= 1e-6
    # the gauge reaches the energy and limit rows of clamped atoms
    o.set_gauge()
    J0 = o.jacobian(x)
    off = o.row_offsets()
    e_rows = off[:-1] + 3 * np.diff(np.asarray(w.dataset.atom_off))  # energy row of each configuration
    l_rows = off[1:] - 1                                             # limit row
    assert np.max(np.abs(Jo[e_rows] - J0[e_rows])) > 1e-3
    assert np.max(np.abs(Jo[l_rows] - J0[l_rows])) > 1e-3


def test_gauged_population_and_lm(wl):
    w = wl["eam3"]
    m = narrowed(w)
    A, S, F0 = gauge_of(m.ntypes)
    _print("gauged population")
    o = _oracle(w, m)
    o.set_gauge(A, S, F0)
    g = fa.GpuCalculator(w.dataset, m)
    g.prepare()
    g.set_gauge(A, S, F0)
    x0 = g.gather_params()
    rng = np.random.default_rng(11)
    X = np.array([x0] + [x0 + rng.normal(size=x0.size) * (0.02 * np.abs(x0) + 1e-3) for _ in range(9)])
    cg, co = g.population_cost(X), o.population_cost(X)
    rel = np.abs(cg - co) / co
    _print(f"[gauge] population max rel {rel.max():.3e}")
    assert rel.max() <= 1e-6
    xo, ho, _ = o.lm(x0, 3)
    xg, hg, acc = g.lm(x0, 3)
    _print(f"[gauge] lm oracle {ho} gpu {hg}")
    assert acc > 0 and abs(hg[-1] - ho[-1]) <= 1e-3 * ho[-1]


def test_captured_step_reads_the_gauge(wl):
    w = wl["eam1"]
    _print("gauge step graph")
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    o = _oracle(w)
    x = g.gather_params()
    c1, F1, _ = g.step(x, want_residual=True)
    for gauge in [(1.5, [-0.3], [0.2]), (0.8, [0.1], [-0.4]), (1.0, [0.0], [0.0])]:
        g.set_gauge(*gauge)
        o.set_gauge(*gauge)
        cs, Fs, _ = g.step(x, want_residual=True)
        Fe = g.residual(x)
        Fo = o.residual(x)
        assert np.max(np.abs(Fs - Fe)) <= 1e-12 * (1 + np.max(np.abs(Fe)))
        assert abs(cs - Fo @ Fo) <= 1e-6 * (Fo @ Fo)
    assert np.array_equal(Fs, F1)  # back at the identity


def test_gauge_refusals_and_reset(wl):
    w = wl["eam3"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    g.set_gauge(1.3, [0.1, 0.2, 0.3], [0.0, 0.0, 0.0])
    with pytest.raises(fa.FsError):
        g.set_gauge(0.0)
    with pytest.raises(fa.FsError):
        g.set_gauge(1.0, [np.nan, 0.0, 0.0])
    assert g.gauge()[0] == 1.3  # refused calls leave the gauge
    # a pair table on another span than the density it pairs with
    m = w.model.copy()
    p = m.tables[0]
    m.tables[0] = Table(p.n, p.rmin, p.rmax - 0.5, p.y.copy(), p.fixed.copy())
    g.set_model(m)
    A, S, F0 = g.gauge()
    assert A == 1.0 and not S.any() and not F0.any()  # a new model: identity
    g.set_gauge(1.3)                                   # a scale alone is fine
    with pytest.raises(fa.FsError):
        g.set_gauge(1.0, [0.1, 0.0, 0.0])             # slope 0 compensates pair (0, 0)
    g.set_gauge(1.0, [0.0, 0.2, 0.0])                  # type 1 pairs are on the density span
    adp = dsm.config3(n_configs=1)
    ga = fa.GpuCalculator(adp.dataset, adp.model)
    with pytest.raises(fa.FsError):
        ga.set_gauge(1.1)
