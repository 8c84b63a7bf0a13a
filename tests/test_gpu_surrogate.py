"""BF16-MFMA MLP surrogate kernels against the FP64 CPU MLP of the oracle.

Bars: forward outputs within 2e-2 of the FP64 network relative to the output
scale; gradients within 3e-2 relative to each parameter block's largest
entry (BF16 operands, FP32 accumulation); the loss within 1e-2 relative.
"""
import numpy as np
import pytest

import forcesmith_amd as fa
import oracle
from forcesmith_amd import datasets as dsm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def calc():
    w = dsm.config1(n_configs=1)
    return fa.GpuCalculator(w.dataset, w.model)


def net(rng, d, h):
    return (rng.normal(size=(h, d)) / np.sqrt(d), rng.normal(size=h) * 0.1,
            rng.normal(size=(h, h)) / np.sqrt(h), rng.normal(size=h) * 0.1,
            rng.normal(size=h) / np.sqrt(h), 0.2)


@pytest.mark.parametrize("n,d,h", [(37, 40, 64), (300, 240, 256), (1000, 24, 96)])
def test_forward_matches_fp64(calc, n, d, h):
    rng = np.random.default_rng(n)
    W = net(rng, d, h)
    X = rng.uniform(-1, 1, size=(n, d))
    y = calc.mlp_forward(fa.mlp_pack(*W), X)
    yo = oracle.mlp_forward(X, *W)
    err = np.max(np.abs(y - yo)) / (np.max(np.abs(yo)) + 1e-12)
    print(f"forward n={n} d={d} h={h} rel err {err:.3e}", flush=True)
    assert err <= 2e-2


@pytest.mark.parametrize("n,d,h", [(64, 40, 64), (512, 240, 256)])
def test_gradients_match_fp64(calc, n, d, h):
    rng = np.random.default_rng(d)
    W = net(rng, d, h)
    X = rng.uniform(-1, 1, size=(n, d))
    t = rng.normal(size=n)
    loss, g = calc.mlp_loss_grad(fa.mlp_pack(*W), X, t)
    lo, go = oracle.mlp_loss_grad(X, t, *W)
    assert abs(loss - lo) <= 1e-2 * abs(lo)
    parts = fa.mlp_unpack(g, d, h)
    for k, (a, b) in enumerate(zip(parts[:5], go[:5])):
        err = np.max(np.abs(np.asarray(a) - b)) / (np.max(np.abs(b)) + 1e-12)
        print(f"grad block {k} rel err {err:.3e}", flush=True)
        assert err <= 3e-2, k
    assert abs(parts[5] - go[5]) <= 3e-2 * (abs(go[5]) + 1e-3)
