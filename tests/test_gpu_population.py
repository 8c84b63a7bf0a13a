"""Batched population cost and the DE driver on the GPU against the FP64 oracle.

The member-lane passes (one member per lane, FP64) give every member's cost
within 1e-10 of the oracle and of the FP64 single-member path; the
member-block passes (members x tile groups) within the mixed-precision cost
bar (1e-6 relative) of the oracle and the mixed single-member path to
rounding (1e-10 relative). FP64 DE uses the oracle's counter RNG, so with FP64
costs its trajectory equals the oracle's.
"""
import numpy as np
import pytest

import forcesmith_amd as fa
import oracle
from forcesmith_amd import datasets as dsm

pytestmark = pytest.mark.gpu


def _print(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def wl():
    return {"eam1": dsm.config1(n_configs=3), "eam3": dsm.config2(n_configs=10, nmin=32, nmax=300)}


def members(x0, n, seed, clamp_member=None):
    rng = np.random.default_rng(seed)
    X = [x0.copy()]
    for _ in range(n - 1):
        X.append(x0 + rng.normal(size=x0.size) * (0.02 * np.abs(x0) + 1e-3))
    if clamp_member is not None:
        X.append(clamp_member)
    return np.array(X)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("name", ["eam1", "eam3"])
def test_batched_population_matches_oracle(wl, name, form):
    """form 0: the member-lane passes (both workloads have their pair and
    density tables on one grid), FP64 arithmetic whatever the context's
    precision: within 1e-10 of the oracle and of the FP64 single-member path.
    form 1: the member-block passes in the context's mixed precision: within
    the mixed bar of the oracle, and the mixed single-member path to
    rounding."""
    w = wl[name]
    _print(f"population {name} form {form}")
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    g.set_population_form(form)
    x0 = g.gather_params()
    # a member whose density tables are tripled pushes rho outside the
    # embedding span: the limit rows are exercised
    xc = x0.copy()
    off = 0
    for t, tab in enumerate(w.model.tables):
        nfree = tab.free_count
        if w.model.npair <= t < w.model.npair + w.model.ntypes:
            xc[off:off + nfree] *= 3.0
        off += nfree
    X = members(x0, 17, 5, xc)
    cg = g.population_cost(X)
    co = o.population_cost(X)
    rel = np.abs(cg - co) / co
    _print(f"[{name} form {form}] members {len(X)} max rel vs oracle {rel.max():.3e}; "
           f"clamp member cost {co[-1]:.4e}")
    assert rel.max() <= (1e-10 if form == 0 else 1e-6)
    assert co[-1] > 10 * co[0]  # the clamp member really is penalised
    # the single-member residual path of the same arithmetic
    gs = g if form == 1 else fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    if form == 0:
        gs.prepare()
    cs = np.array([gs.cost(x) for x in X])
    assert np.max(np.abs(cg - cs) / cs) <= 1e-10


@pytest.mark.parametrize("name", ["eam1", "eam3"])
def test_member_cost_independent_of_batch(wl, name):
    """A member's cost does not depend on its place in the population: the
    member-lane sums are per lane, in the same order for every lane."""
    w = wl[name]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    X = members(g.gather_params(), 70, 13)
    full = g.population_cost(X)
    assert np.array_equal(g.population_cost(X[5:6]), full[5:6])
    assert np.array_equal(g.population_cost(X[::-1]), full[::-1])


def test_population_deterministic(wl):
    w = wl["eam3"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    X = members(g.gather_params(), 9, 11)
    a, b = g.population_cost(X), g.population_cost(X)
    assert np.array_equal(a, b)


def test_de_fp64_matches_oracle(wl):
    w = wl["eam1"]
    _print("de fp64 vs oracle")
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    x0 = g.gather_params()
    lo, hi = x0 - 0.05 * np.abs(x0) - 1e-3, x0 + 0.05 * np.abs(x0) + 1e-3
    xo, bo, ho = o.de(x0, lo, hi, 16, 5, seed=42)
    xg, bg, hg = g.de(x0, lo, hi, 16, 5, seed=42)
    _print(f"best oracle {bo:.12e} gpu {bg:.12e}")
    assert abs(bg - bo) <= 1e-10 * bo
    assert np.allclose(xg, xo, rtol=0, atol=1e-12 * (1 + np.abs(xo).max()))
    assert np.allclose(hg, ho, rtol=1e-10)


@pytest.mark.parametrize("surrogate", [False, True])
def test_de_mixed_batched(wl, surrogate):
    w = wl["eam3"]
    _print(f"de mixed surrogate={surrogate}")
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x0 = g.gather_params()
    lo, hi = x0 - 0.05 * np.abs(x0) - 1e-3, x0 + 0.05 * np.abs(x0) + 1e-3
    x, best, hist = g.de(x0, lo, hi, 48, 6, seed=7, surrogate=surrogate, oversample=4,
                         retrain_every=2, hidden=64)
    assert np.all(x >= lo) and np.all(x <= hi)
    assert np.all(np.diff(hist) <= 0)
    # the batched exact cost judges every member: the reported best is the
    # population cost of x, at most that of the clamped initial guess, and
    # within the mixed bar of the single-member residual path
    pc = lambda v: g.population_cost(v[None, :])[0]
    assert best <= pc(np.clip(x0, lo, hi))
    assert best == pc(x)
    assert abs(best - g.cost(x)) <= 1e-6 * best
    x2, best2, _ = g.de(x0, lo, hi, 48, 6, seed=7, surrogate=surrogate, oversample=4,
                        retrain_every=2, hidden=64)
    assert best2 == best and np.array_equal(x, x2)


def test_de_device_driver_matches_oracle(wl):
    """The device-resident driver (hipGraph-replayed generations) draws the
    oracle's trials; with costs within the mixed bar the selections agree."""
    w = wl["eam1"]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x0 = g.gather_params()
    lo, hi = x0 - 0.05 * np.abs(x0) - 1e-3, x0 + 0.05 * np.abs(x0) + 1e-3
    xo, bo, ho = o.de(x0, lo, hi, 16, 5, seed=3)
    xg, bg, hg = g.de(x0, lo, hi, 16, 5, seed=3)
    _print(f"device DE best {bg:.10e} oracle {bo:.10e}")
    assert np.array_equal(xg, xo)
    assert abs(bg - bo) <= 1e-6 * bo
    assert np.allclose(hg, ho, rtol=1e-6)


def test_graph_replay_equals_eager_generations(wl):
    """Plain DE replays one captured generation; the screened driver with a
    single candidate runs the same kernels eagerly: identical trajectories."""
    w = wl["eam3"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x0 = g.gather_params()
    lo, hi = x0 - 0.1 * np.abs(x0) - 1e-3, x0 + 0.1 * np.abs(x0) + 1e-3
    xa, ba, ha = g.de(x0, lo, hi, 40, 20, seed=9)
    xb, bb, hb = g.de(x0, lo, hi, 40, 20, seed=9, surrogate=True, oversample=1)
    assert np.array_equal(xa, xb) and ba == bb and np.array_equal(ha, hb)
    assert np.all(np.diff(ha) <= 0)


def test_de_target_stops_early(wl):
    w = wl["eam3"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x0 = g.gather_params()
    lo, hi = x0 - 0.1 * np.abs(x0) - 1e-3, x0 + 0.1 * np.abs(x0) + 1e-3
    _, _, ha = g.de(x0, lo, hi, 40, 40, seed=9)
    target = ha[9]
    _, bt, ht = g.de(x0, lo, hi, 40, 40, seed=9, target_cost=target)
    assert len(ht) == int(np.argmax(ha <= target)) + 1
    assert bt <= target


def test_de_reference_box_and_population_size(wl):
    """Infinite bounds take BoostDESolver's box v -+ max(2|v|, 5e-4)
    (solver.cpp:118-125) and np <= 0 takes 15 D (solver.hpp:78-81): the same
    run as with that box and size given explicitly."""
    w = wl["eam1"]
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x0 = g.gather_params()
    lo, hi = g.gather_bounds()
    assert np.all(np.isinf(lo)) and np.all(np.isinf(hi))
    xa, ba, ha = g.de(x0, lo, hi, 0, 3, seed=5)
    blo, bhi = dsm.de_box(x0)
    xb, bb, hb = g.de(x0, blo, bhi, 15 * x0.size, 3, seed=5)
    _print(f"reference box DE best {ba:.10e}")
    assert np.array_equal(xa, xb) and ba == bb and np.array_equal(ha, hb)
    # a finite bound is kept as given
    lo2 = lo.copy()
    lo2[0] = x0[0] - 1e-3
    xc, bc, hc = g.de(x0, lo2, hi, 0, 3, seed=5)
    blo2 = blo.copy()
    blo2[0] = x0[0] - 1e-3
    xd, bd, hd = g.de(x0, blo2, bhi, 15 * x0.size, 3, seed=5)
    assert np.array_equal(xc, xd) and bc == bd


@pytest.mark.parametrize("name", ["eam1", "eam3"])
def test_batched_population_fp64_matches_oracle(wl, name):
    """FP64 batched population costs (every pair through the FP64 path) within
    the FP64-mode bar of 1e-11, including a clamped member."""
    w = wl[name]
    o = oracle.Oracle(w.dataset, w.model)
    o.prepare()
    g = fa.GpuCalculator(w.dataset, w.model, precision="fp64")
    g.prepare()
    x0 = g.gather_params()
    xc = x0.copy()
    off = 0
    for t, tab in enumerate(w.model.tables):
        nfree = tab.free_count
        if w.model.npair <= t < w.model.npair + w.model.ntypes:
            xc[off:off + nfree] *= 3.0
        off += nfree
    X = members(x0, 9, 5, xc)
    cg, co = g.population_cost(X), o.population_cost(X)
    rel = np.abs(cg - co) / co
    _print(f"[{name} fp64] members {len(X)} max rel vs oracle {rel.max():.3e}")
    assert rel.max() <= 1e-11
