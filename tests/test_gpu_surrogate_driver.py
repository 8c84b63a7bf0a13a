"""Surrogate-screened DE inside the device driver (fs_de_minimize with
surrogate = 1), with the driver's own statistics (fs_de_last_stats).

  - At the screened run's shape (D = 240 ternary EAM, population 3600, hidden
    width 256, 4 candidates per member, retraining every 10 generations) on a
    reduced dataset: the network is retrained on the GPU twice and its training
    loss falls; the best cost never rises.
  - The rank-error fallback: with rank_tau_min above any reachable Kendall
    tau, every screened generation is followed by a fallback (screened ==
    fallbacks == retrains), and the generations before the first retraining
    are the plain driver's generations exactly (candidate 0 is the plain
    trial).
  - Several screened runs on one context with the retraining schedule whose
    minibatches grow (1008, 1848, 2048 rows: the schedule under which the
    former order of the backward pass wrote into a freed workspace): every run
    gives the same result.
"""
import numpy as np
import pytest

import forcesmith_amd as fa
from forcesmith_amd import datasets as dsm

pytestmark = pytest.mark.gpu


def _print(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ternary():
    w = dsm.config2(n_configs=30, nmin=32, nmax=600)
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    return w, g


def _box(x0):
    return x0 - 0.2 * np.abs(x0) - 1e-3, x0 + 0.2 * np.abs(x0) + 1e-3


@pytest.mark.timeout(600)
def test_screened_run_trains_in_driver(ternary):
    _, g = ternary
    x0 = g.gather_params()
    lo, hi = _box(x0)
    _print("screened DE, population 3600, hidden 256, 22 generations")
    x, best, hist = g.de(x0, lo, hi, 3600, 22, seed=42, surrogate=True, oversample=4,
                         retrain_every=10, hidden=256, rank_tau_min=-2.0)
    st = g.de_stats()
    _print(f"stats {st}; best {hist[0]:.6e} -> {best:.6e}")
    assert st["generations"] == 22 and st["retrains"] == 2
    assert st["screened"] == 12 and st["fallbacks"] == 0
    assert 0.0 < st["last_loss"] < 0.5 * st["first_loss"]
    assert np.all(np.diff(hist) <= 0)
    assert np.all(x >= lo) and np.all(x <= hi)
    # the batched exact cost judged the best; the single-member residual path
    # agrees within the mixed bar
    assert best == g.population_cost(x[None, :])[0]
    assert abs(best - g.cost(x)) <= 1e-6 * best


@pytest.mark.timeout(600)
def test_rank_error_fallback(ternary):
    _, g = ternary
    x0 = g.gather_params()
    lo, hi = _box(x0)
    _print("screened DE with an unreachable rank bound")
    _, _, hist = g.de(x0, lo, hi, 720, 22, seed=5, surrogate=True, oversample=4, retrain_every=10,
                      hidden=256, rank_tau_min=2.0)
    st = g.de_stats()
    _print(f"stats {st}")
    assert st["retrains"] == 2 and st["screened"] == 2 and st["fallbacks"] == 2
    _, _, plain = g.de(x0, lo, hi, 720, 10, seed=5)
    assert g.de_stats()["retrains"] == 0
    assert np.array_equal(hist[:10], plain)


@pytest.mark.timeout(600)
def test_repeated_screened_runs_with_growing_minibatches():
    w = dsm.config1()
    g = fa.GpuCalculator(w.dataset, w.model)
    g.prepare()
    x0 = g.gather_params()
    lo, hi = _box(x0)
    res = []
    for call in range(3):
        _print(f"screened DE call {call}: population 168, hidden 64, retraining every 5")
        x, best, hist = g.de(x0, lo, hi, 168, 20, seed=11, surrogate=True, oversample=4,
                             retrain_every=5, hidden=64, rank_tau_min=-2.0)
        st = g.de_stats()
        assert st["retrains"] == 3 and st["screened"] == 15
        res.append((x, best, hist))
    for x, best, hist in res[1:]:
        assert np.array_equal(x, res[0][0]) and best == res[0][1] and np.array_equal(hist, res[0][2])
