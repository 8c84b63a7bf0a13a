"""The host trust-region algebra of fs_lm_solve (forcesmith_amd/csrc/
lm_tr.hpp, normal-equation form) against the oracle's QR restatement of
MINPACK lmder, on the CPU: both are driven with the same FP64 residuals and
Jacobians (the product's GPU supplies these on the device), so on
well-conditioned problems they take the same steps: equal status and counts,
costs and minimizers within rounding. The wrapper tests/lm_host_harness.cpp is
compiled here with g++.
"""
import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.test_oracle_lm import MGH, fd_jac

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_dp = ct.POINTER(ct.c_double)
COST = ct.CFUNCTYPE(ct.c_double, _dp)
NORMAL = ct.CFUNCTYPE(None, _dp, _dp)


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    out = tmp_path_factory.mktemp("lmh") / "liblmh.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "forcesmith_amd", "csrc"),
                    os.path.join(ROOT, "tests", "lm_host_harness.cpp"), "-o", str(out)], check=True)
    lib = ct.CDLL(str(out))
    lib.lmh_minimize.restype = ct.c_int
    lib.lmh_minimize.argtypes = [ct.c_int, ct.c_int64, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                 ct.c_double, ct.c_double, _dp, COST, NORMAL, ct.POINTER(ct.c_int32), _dp]
    lib.lmh_lmpar.restype = ct.c_int
    lib.lmh_lmpar.argtypes = [ct.c_int, _dp, _dp, _dp, ct.c_double, _dp, _dp]
    return lib


def run_host(lib, fun, jac, x0, m, max_iter=400, maxfev=0, xtol=1e-10, ftol=1e-10, gtol=0.0, factor=100.0):
    x = np.array(x0, dtype=np.float64)
    n = x.size
    last = {}

    def cost(xp):
        v = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        last["x"] = v
        f = fun(v)
        return float(f @ f)

    def normal(Ap, gp):
        v = last["x"]
        J, f = jac(v), fun(v)
        np.ctypeslib.as_array(Ap, shape=(n * n,))[:] = (J.T @ J).reshape(-1)
        np.ctypeslib.as_array(gp, shape=(n,))[:] = J.T @ f

    cb_c, cb_n = COST(cost), NORMAL(normal)
    cnt = (ct.c_int32 * 3)()
    hist = np.zeros(max(max_iter, 1))
    st = lib.lmh_minimize(n, m, max_iter, maxfev, xtol, ftol, gtol, factor, x.ctypes.data_as(_dp), cb_c, cb_n,
                          cnt, hist.ctypes.data_as(_dp))
    return x, st, tuple(cnt), hist


WELL = ["rosenbrock", "freudenstein_roth", "helical_valley", "linear_full_rank", "bard", "kowalik_osborne",
        "jennrich_sampson", "osborne1"]


@pytest.mark.parametrize("name,fun,m,x0,fmin", [p for p in MGH if p[0] in WELL],
                         ids=[p[0] for p in MGH if p[0] in WELL])
def test_same_steps_as_the_qr_form(host, name, fun, m, x0, fmin):
    jac = lambda v: fd_jac(fun, v)
    xo, so, co, ho = oracle.lmder(fun, jac, np.array(x0), m, max_iter=400, xtol=1e-10, ftol=1e-10)
    xh, sh, ch, hh = run_host(host, fun, jac, x0, m)
    print(f"{name}: qr {so} {co} {ho[co[1] - 1]:.12e} | ne {sh} {ch} {hh[ch[1] - 1]:.12e}")
    # the last steps run at the rounding level of actred / prered, where the
    # two factorizations may accept one step more or less
    assert sh == so and all(abs(a - b) <= 1 for a, b in zip(ch, co))
    k = min(co[1], ch[1]) - 1
    scale = max(ho[0], 1e-300)
    # forming J^T J squares the condition number (Osborne 1: cond(J) ~ 1e4),
    # so intermediate costs agree to ~1e-6 there and ~1e-10 elsewhere
    assert np.allclose(hh[:k], ho[:k], rtol=1e-5, atol=1e-12 * scale)
    assert hh[ch[1] - 1] == pytest.approx(ho[co[1] - 1], rel=1e-9, abs=1e-20 * scale)
    assert np.allclose(xh, xo, rtol=1e-5, atol=1e-8 * (1 + np.abs(xo).max()))


# box3d ends at a zero residual, where the status depends on whether the last
# step lands on 0 exactly (cosine test, 4) or at 1e-33 (relative error, 2)
@pytest.mark.parametrize("name", ["powell_singular", "linear_rank1", "linear_rank1_zero", "brown_dennis", "box3d"])
def test_rank_deficient_and_hard_problems_reach_the_minimum(host, name):
    _, fun, m, x0, fmin = next(p for p in MGH if p[0] == name)
    xh, sh, ch, hh = run_host(host, fun, lambda v: fd_jac(fun, v), x0, m)
    cost = float(fun(xh) @ fun(xh))
    print(f"{name}: status {sh} counts {ch} cost {cost:.10e}")
    assert 1 <= sh <= 8
    if fmin == 0.0:
        assert cost <= 1e-10
    else:
        assert cost == pytest.approx(fmin, rel=2e-6)


def test_lmpar_matches_the_qr_form(host):
    rng = np.random.default_rng(3)
    for trial in range(40):
        m, n = 25, 5
        J = rng.standard_normal((m, n)) * rng.uniform(0.2, 5.0, n)
        f = rng.standard_normal(m)
        dg = rng.uniform(0.5, 4.0, n)
        rank, perm, R, qtf = oracle.qr_pivot(J, f)
        gnn = np.linalg.norm(dg * np.linalg.lstsq(J, f, rcond=None)[0])
        delta = gnn * rng.choice([0.02, 0.5, 0.99, 3.0])
        par_o, x_o, _ = oracle.lmpar(R, perm, rank, dg, qtf, delta)
        A, g = np.ascontiguousarray(J.T @ J), J.T @ f
        par = ct.c_double(0.0)
        x = np.zeros(n)
        host.lmh_lmpar(n, A.ctypes.data_as(_dp), g.ctypes.data_as(_dp), dg.ctypes.data_as(_dp), delta,
                       ct.byref(par), x.ctypes.data_as(_dp))
        assert par.value == pytest.approx(par_o, rel=1e-8, abs=1e-300)
        assert np.allclose(x, x_o, rtol=1e-8, atol=1e-12)


def test_status_rules(host):
    f = MGH[0][1]
    jac = lambda v: fd_jac(f, v)
    x0 = np.array([-1.2, 1.0])
    for bad in (dict(xtol=-1.0), dict(ftol=-1.0), dict(gtol=-1.0), dict(factor=0.0)):
        x, st, cnt, _ = run_host(host, f, jac, x0, 2, **bad)
        assert st == 0 and cnt == (0, 0, 0) and np.array_equal(x, x0)
    x, st, cnt, _ = run_host(host, f, jac, x0, 2, max_iter=3)
    assert st == 2 and cnt[1] == 3
    x, st, cnt, _ = run_host(host, f, jac, [1.0, 1.0], 2)
    assert st == 4 and cnt[:2] == (1, 1)
    x, st, cnt, _ = run_host(host, f, jac, x0, 2, max_iter=100, maxfev=4)
    assert st == 5 and cnt[0] == 4
