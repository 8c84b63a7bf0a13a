"""FP64 CPU oracle (test infrastructure only).

A plain-C restatement of the reference calculator's hot path
(src/core/neighbor_list.cpp, src/potentials/spline.cpp, src/force/eam_force.cpp,
src/force/adp_force.cpp, src/force/stiweb_force.cpp,
include/forcesmith/optimization/{residual_layout,fd_jacobian}.hpp), loaded
through ctypes. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline; the product path never touches it.

Parity-pinning status per function is listed in ``oracle/oracle.h``.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

EAM, ADP, SW = 0, 1, 2

_dp = ct.POINTER(ct.c_double)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)
_u8p = ct.POINTER(ct.c_uint8)

# or_lmder callbacks: f(user, x, out) -> int
FCN = ct.CFUNCTYPE(ct.c_int, ct.c_void_p, _dp, _dp)
POPCOST = ct.CFUNCTYPE(None, ct.c_void_p, _dp, ct.c_int, _dp)


class LmOpts(ct.Structure):
    """or_lm_opts; defaults are EigenLMSolver's (solver.hpp:65-68) and Eigen's
    factor 100, gtol 0; maxfev 0 = max(max_iter, 1) * 100 + 100."""
    _fields_ = [("max_iter", ct.c_int32), ("maxfev", ct.c_int32), ("xtol", ct.c_double),
                ("ftol", ct.c_double), ("gtol", ct.c_double), ("factor", ct.c_double)]

    def __init__(self, max_iter=500, maxfev=0, xtol=1e-7, ftol=1e-7, gtol=0.0, factor=100.0):
        super().__init__(int(max_iter), int(maxfev), float(xtol), float(ftol), float(gtol), float(factor))


def build() -> str:
    """Compile the oracle with its committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ct.CDLL(_LIB_PATH)
    sig = {
        "or_create": (ct.c_void_p, [ct.c_int, ct.c_int, ct.c_int, _i64p, _i32p, _dp, _dp, _dp, _dp, _dp]),
        "or_destroy": (None, [ct.c_void_p]),
        "or_set_threads": (None, [ct.c_int]),
        "or_get_threads": (ct.c_int, []),
        "or_set_tables": (ct.c_int, [ct.c_void_p, ct.c_int, _i32p, _dp, _dp, _dp, _u8p]),
        "or_set_sw": (ct.c_int, [ct.c_void_p, _dp, _u8p, _dp, _u8p]),
        "or_param_count": (ct.c_int, [ct.c_void_p]),
        "or_gather_params": (None, [ct.c_void_p, _dp]),
        "or_scatter_params": (None, [ct.c_void_p, _dp]),
        "or_max_cutoff": (ct.c_double, [ct.c_void_p]),
        "or_set_weights": (None, [ct.c_void_p, ct.c_double, ct.c_double]),
        "or_residual_count": (ct.c_int64, [ct.c_void_p]),
        "or_set_smooth_weight": (None, [ct.c_void_p, ct.c_double]),
        "or_smooth_count": (ct.c_int64, [ct.c_void_p]),
        "or_row_offsets": (None, [ct.c_void_p, _i64p]),
        "or_prepare": (ct.c_int64, [ct.c_void_p]),
        "or_build_nl": (ct.c_int64, [ct.c_void_p, ct.c_double]),
        "or_get_nl": (None, [ct.c_void_p, _i64p, _i32p, _i32p, _dp]),
        "or_eval_forces": (None, [ct.c_void_p, _dp, _dp, _dp, _dp]),
        "or_get_scratch": (None, [ct.c_void_p, _dp, _dp, _dp, _dp]),
        "or_residual": (None, [ct.c_void_p, _dp, _dp]),
        "or_jacobian_fd": (None, [ct.c_void_p, _dp, _dp]),
        "or_jacobian_fd_cols": (None, [ct.c_void_p, _dp, ct.c_int, ct.c_int, _dp]),
        "or_jacobian_analytic": (ct.c_int, [ct.c_void_p, _dp, _dp]),
        "or_spline_slopes": (ct.c_int, [ct.c_int, _dp, _dp, _dp]),
        "or_spline_eval": (None, [ct.c_int, _dp, _dp, _dp, ct.c_double, _dp, _dp]),
        "or_spline_eval_cached": (None, [ct.c_int, _dp, _dp, _dp, ct.c_double, _dp, _dp]),
        "or_spline_slope_jacobian": (None, [ct.c_int, _dp, _dp, _dp]),
        "or_spline_param_grad": (None, [ct.c_int, _dp, _dp, ct.c_double, _dp, _dp, _dp]),
        "or_knot_grid": (None, [ct.c_int, ct.c_double, ct.c_double, _dp]),
        "or_lm": (ct.c_int, [ct.c_void_p, _dp, ct.c_int, ct.c_int, _dp]),
        "or_lmder": (ct.c_int, [ct.c_int64, ct.c_int, FCN, FCN, ct.c_void_p, _dp, ct.POINTER(LmOpts), _i32p,
                                _dp]),
        "or_lm_minpack": (ct.c_int, [ct.c_void_p, _dp, ct.POINTER(LmOpts), _i32p, _dp]),
        "or_normal_equations_solve": (ct.c_int, [ct.c_void_p, ct.c_double, _dp]),
        "or_qr_pivot": (ct.c_int, [ct.c_int64, ct.c_int, _dp, _dp, _i32p, _dp, _dp]),
        "or_lmpar": (ct.c_int, [ct.c_int, _dp, _i32p, ct.c_int, _dp, _dp, ct.c_double, _dp, _dp]),
        "or_population_cost": (None, [ct.c_void_p, _dp, ct.c_int, _dp]),
        "or_rng_u64": (ct.c_uint64, [ct.c_uint64, ct.c_uint64, ct.c_uint64]),
        "or_rng_unit": (ct.c_double, [ct.c_uint64, ct.c_uint64, ct.c_uint64]),
        "or_de": (ct.c_double, [ct.c_void_p, _dp, _dp, _dp, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                ct.c_uint64, _dp]),
        "or_de_init": (None, [_dp, _dp, _dp, ct.c_int, ct.c_int, ct.c_uint64, _dp]),
        "or_de_trial": (None, [_dp, ct.c_int, ct.c_int, _dp, _dp, ct.c_double, ct.c_double, ct.c_uint64,
                               ct.c_int, ct.c_int, _dp, _i32p]),
        "or_de_select": (ct.c_int, [ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _u8p]),
        "or_de_fn": (ct.c_double, [POPCOST, ct.c_void_p, ct.c_int, _dp, _dp, _dp, ct.c_int, ct.c_int,
                                   ct.c_double, ct.c_double, ct.c_uint64, _dp]),
        "or_mlp_forward": (None, [ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _dp, _dp, ct.c_double, _dp]),
        "or_mlp_loss_grad": (ct.c_double, [ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                           ct.c_double, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_time_fd_jacobian": (ct.c_double, [ct.c_void_p, _dp, ct.c_int, _dp]),
        "or_time_residual": (ct.c_double, [ct.c_void_p, _dp, ct.c_int, _dp]),
        "or_set_gauge": (ct.c_int, [ct.c_void_p, ct.c_double, _dp, _dp]),
        "or_get_gauge": (None, [ct.c_void_p, _dp, _dp, _dp]),
        "or_gauge_spans_match": (ct.c_int, [ct.c_void_p]),
        "or_emb_gauged": (None, [ct.c_void_p, ct.c_int, ct.c_double, _dp, _dp]),
        "or_rho_stats": (None, [ct.c_void_p, _dp, _dp, _dp, _i64p]),
        "or_rescale_rho_axis": (ct.c_double, [ct.c_void_p]),
        "or_embed_shift": (None, [ct.c_void_p, _dp]),
        "or_rescale_eam": (ct.c_double, [ct.c_void_p, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a, typ):
    return a.ctypes.data_as(typ) if a is not None else ct.cast(None, typ)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    _load().or_set_threads(int(n))


def get_threads() -> int:
    return int(_load().or_get_threads())


# --------------------------------------------------------------------------
# spline primitives
# --------------------------------------------------------------------------
def knot_grid(n, rmin, rmax):
    x = np.empty(n)
    _load().or_knot_grid(n, rmin, rmax, _p(x, _dp))
    return x


def spline_slopes(x, y):
    x, y = _f64(x), _f64(y)
    s = np.zeros(len(x))
    ok = _load().or_spline_slopes(len(x), _p(x, _dp), _p(y, _dp), _p(s, _dp))
    return s if ok else None


def spline_eval(x, y, r, cached=False):
    x, y = _f64(x), _f64(y)
    s = spline_slopes(x, y)
    if s is None:
        s = np.zeros(len(x))
    v, d = ct.c_double(), ct.c_double()
    fn = _load().or_spline_eval_cached if cached else _load().or_spline_eval
    fn(len(x), _p(x, _dp), _p(y, _dp), _p(s, _dp), float(r), ct.byref(v), ct.byref(d))
    return v.value, d.value


def spline_slope_jacobian(x, y):
    x, y = _f64(x), _f64(y)
    S = np.zeros((len(x), len(x)))
    _load().or_spline_slope_jacobian(len(x), _p(x, _dp), _p(y, _dp), _p(S, _dp))
    return S


def spline_param_grad(x, y, r):
    x, y = _f64(x), _f64(y)
    n = len(x)
    dv, dd, d2 = np.zeros(n), np.zeros(n), ct.c_double()
    _load().or_spline_param_grad(n, _p(x, _dp), _p(y, _dp), float(r), _p(dv, _dp), _p(dd, _dp),
                                 ct.byref(d2))
    return dv, dd, d2.value


# --------------------------------------------------------------------------
# rng shared with the GPU DE
# --------------------------------------------------------------------------
def rng_unit(seed, stream, index):
    return _load().or_rng_unit(seed, stream, index)


def rng_u64(seed, stream, index):
    return _load().or_rng_u64(seed, stream, index)


# --------------------------------------------------------------------------
# MLP
# --------------------------------------------------------------------------
def mlp_forward(X, W1, b1, W2, b2, W3, b3):
    X = _f64(X)
    n, d = X.shape
    h = W1.shape[0]
    y = np.zeros(n)
    _load().or_mlp_forward(n, d, h, _p(X, _dp), _p(_f64(W1), _dp), _p(_f64(b1), _dp), _p(_f64(W2), _dp),
                           _p(_f64(b2), _dp), _p(_f64(W3), _dp), float(b3), _p(y, _dp))
    return y


def mlp_loss_grad(X, t, W1, b1, W2, b2, W3, b3):
    X, t = _f64(X), _f64(t)
    n, d = X.shape
    h = W1.shape[0]
    g = [np.zeros_like(_f64(W1)), np.zeros(h), np.zeros((h, h)), np.zeros(h), np.zeros(h)]
    gb3 = ct.c_double()
    loss = _load().or_mlp_loss_grad(n, d, h, _p(X, _dp), _p(t, _dp), _p(_f64(W1), _dp), _p(_f64(b1), _dp),
                                    _p(_f64(W2), _dp), _p(_f64(b2), _dp), _p(_f64(W3), _dp), float(b3),
                                    *[_p(a, _dp) for a in g], ct.byref(gb3))
    return loss, g + [gb3.value]


# --------------------------------------------------------------------------
# calculator context
# --------------------------------------------------------------------------
class Oracle:
    """One dataset + one model, mirroring the reference's functor state."""

    def __init__(self, ds, model, energy_weight=None, stress_weight=None):
        lib = _load()
        self._lib = lib
        self.kind = {"eam": EAM, "adp": ADP, "sw": SW}[model.kind]
        self._keep = []
        off = np.ascontiguousarray(ds.atom_off, dtype=np.int64)
        types = np.ascontiguousarray(ds.types, dtype=np.int32)
        pos = _f64(ds.pos)
        rf = _f64(ds.ref_force)
        cell = _f64(ds.cell)
        re = _f64(ds.ref_energy)
        rs = _f64(ds.ref_stress)
        self._h = lib.or_create(self.kind, model.ntypes, ds.n_configs, _p(off, _i64p), _p(types, _i32p),
                                _p(pos, _dp), _p(rf, _dp), _p(cell, _dp), _p(re, _dp), _p(rs, _dp))
        self.n_atoms = int(off[-1])
        self.n_configs = ds.n_configs
        self.ntypes = int(model.ntypes)
        if model.kind == "sw":
            pp = _f64(model.pair_params)
            pf = np.ascontiguousarray(model.pair_fixed, dtype=np.uint8)
            lam = _f64(model.lam)
            lf = np.ascontiguousarray(model.lam_fixed, dtype=np.uint8)
            rc = lib.or_set_sw(self._h, _p(pp, _dp), _p(pf, _u8p), _p(lam, _dp), _p(lf, _u8p))
        else:
            n = np.array([t.n for t in model.tables], dtype=np.int32)
            rmin = np.array([t.rmin for t in model.tables], dtype=np.float64)
            rmax = np.array([t.rmax for t in model.tables], dtype=np.float64)
            y = np.concatenate([np.asarray(t.y, dtype=np.float64) for t in model.tables])
            fx = np.concatenate([np.asarray(t.fixed, dtype=np.uint8) for t in model.tables])
            rc = lib.or_set_tables(self._h, len(model.tables), _p(n, _i32p), _p(rmin, _dp), _p(rmax, _dp),
                                   _p(y, _dp), _p(fx, _u8p))
        if rc != 0:
            raise ValueError(f"oracle rejected the model (code {rc})")
        ew = model.energy_weight if energy_weight is None else energy_weight
        sw = model.stress_weight if stress_weight is None else stress_weight
        lib.or_set_weights(self._h, float(ew), float(sw))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.or_destroy(h)
            self._h = None

    # --- parameter surface (TableParams) ---
    @property
    def param_count(self) -> int:
        return int(self._lib.or_param_count(self._h))

    def gather(self):
        x = np.zeros(self.param_count)
        self._lib.or_gather_params(self._h, _p(x, _dp))
        return x

    def scatter(self, x):
        x = _f64(x)
        self._lib.or_scatter_params(self._h, _p(x, _dp))

    @property
    def max_cutoff(self) -> float:
        return float(self._lib.or_max_cutoff(self._h))

    @property
    def residual_count(self) -> int:
        return int(self._lib.or_residual_count(self._h))

    def set_smooth_weight(self, w: float):
        """Curvature rows w (y[k-1] - 2 y[k] + y[k+1]) after the config rows."""
        self._lib.or_set_smooth_weight(self._h, float(w))

    @property
    def smooth_count(self) -> int:
        return int(self._lib.or_smooth_count(self._h))

    def row_offsets(self):
        r = np.zeros(self.n_configs + 1, dtype=np.int64)
        self._lib.or_row_offsets(self._h, _p(r, _i64p))
        return r

    # --- neighbour list ---
    def prepare(self) -> int:
        return int(self._lib.or_prepare(self._h))

    def build_nl(self, rcut) -> int:
        return int(self._lib.or_build_nl(self._h, float(rcut)))

    def neighbor_list(self):
        P = self._npairs()
        off = np.zeros(self.n_atoms + 1, dtype=np.int64)
        j = np.zeros(P, dtype=np.int32)
        img = np.zeros((P, 3), dtype=np.int32)
        d = np.zeros((P, 3))
        self._lib.or_get_nl(self._h, _p(off, _i64p), _p(j, _i32p), _p(img, _i32p), _p(d, _dp))
        return off, j, img, d

    def _npairs(self):
        off = np.zeros(self.n_atoms + 1, dtype=np.int64)
        self._lib.or_get_nl(self._h, _p(off, _i64p), ct.cast(None, _i32p), ct.cast(None, _i32p),
                            ct.cast(None, _dp))
        return int(off[-1])

    # --- evaluation ---
    def eval_forces(self):
        F = np.zeros((self.n_atoms, 3))
        E = np.zeros(self.n_configs)
        S = np.zeros((self.n_configs, 3, 3))
        L = np.zeros(self.n_configs)
        self._lib.or_eval_forces(self._h, _p(F, _dp), _p(E, _dp), _p(S, _dp), _p(L, _dp))
        return F, E, S, L

    def scratch(self):
        n = self.n_atoms
        rho, gF, mu, lam = np.zeros(n), np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3, 3))
        self._lib.or_get_scratch(self._h, _p(rho, _dp), _p(gF, _dp), _p(mu, _dp), _p(lam, _dp))
        return rho, gF, mu, lam

    def residual(self, x):
        x = _f64(x)
        F = np.zeros(self.residual_count)
        self._lib.or_residual(self._h, _p(x, _dp), _p(F, _dp))
        return F

    def cost(self, x):
        F = self.residual(x)
        return float(F @ F)

    def jacobian_fd(self, x):
        x = _f64(x)
        J = np.zeros((self.residual_count, self.param_count))
        self._lib.or_jacobian_fd(self._h, _p(x, _dp), _p(J, _dp))
        return J

    def jacobian_fd_cols(self, x, col0, ncols):
        x = _f64(x)
        J = np.zeros((self.residual_count, ncols))
        self._lib.or_jacobian_fd_cols(self._h, _p(x, _dp), int(col0), int(ncols), _p(J, _dp))
        return J

    def jacobian(self, x):
        x = _f64(x)
        J = np.zeros((self.residual_count, self.param_count))
        rc = self._lib.or_jacobian_analytic(self._h, _p(x, _dp), _p(J, _dp))
        if rc != 0:
            raise ValueError(f"oracle analytic Jacobian refused the model (code {rc})")
        return J

    # --- EAM gauge (src/core/rescale.cpp, core/scale_potentials.hpp) ---
    def set_gauge(self, A=1.0, S=None, F0=None):
        nt = self.ntypes
        S = _f64(np.zeros(nt) if S is None else S)
        F0 = _f64(np.zeros(nt) if F0 is None else F0)
        rc = self._lib.or_set_gauge(self._h, float(A), _p(S, _dp), _p(F0, _dp))
        if rc != 0:
            raise ValueError(f"oracle rejected the gauge (code {rc})")

    def gauge(self):
        A = np.zeros(1)
        S, F0 = np.zeros(self.ntypes), np.zeros(self.ntypes)
        self._lib.or_get_gauge(self._h, _p(A, _dp), _p(S, _dp), _p(F0, _dp))
        return float(A[0]), S, F0

    def gauge_spans_match(self) -> bool:
        return bool(self._lib.or_gauge_spans_match(self._h))

    def embedding(self, t, rho):
        """Value and derivative of the gauged embedding of type t at rho."""
        F, dF = np.zeros(1), np.zeros(1)
        self._lib.or_emb_gauged(self._h, int(t), float(rho), _p(F, _dp), _p(dF, _dp))
        return float(F[0]), float(dF[0])

    def rho_stats(self):
        """Per type: min, max, mean of the clamped densities and atom count."""
        nt = self.ntypes
        lo, hi, mean = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        cnt = np.zeros(nt, dtype=np.int64)
        self._lib.or_rho_stats(self._h, _p(lo, _dp), _p(hi, _dp), _p(mean, _dp), _p(cnt, _i64p))
        return lo, hi, mean, cnt

    def compute_rho_ref(self):
        return self.rho_stats()[2]

    def rescale_rho_axis(self) -> float:
        return float(self._lib.or_rescale_rho_axis(self._h))

    def embed_shift(self, rho_ref):
        r = _f64(rho_ref)
        self._lib.or_embed_shift(self._h, _p(r, _dp))

    def rescale_eam(self):
        """Returns (a, rho_ref)."""
        ref = np.zeros(self.ntypes)
        a = self._lib.or_rescale_eam(self._h, _p(ref, _dp))
        return float(a), ref

    def lm(self, x, max_iter, use_fd=False):
        x = _f64(x).copy()
        hist = np.zeros(max_iter)
        acc = self._lib.or_lm(self._h, _p(x, _dp), int(max_iter), int(use_fd), _p(hist, _dp))
        return x, hist, acc

    def ne_solve(self, x, ridge=0.0):
        """NormalEquationsSolver: one ridge step; returns the new x."""
        x = _f64(x).copy()
        rc = self._lib.or_normal_equations_solve(self._h, float(ridge), _p(x, _dp))
        if rc != 0:
            raise ValueError("normal equations not positive definite")
        return x

    def lm_minpack(self, x, **opts):
        """LM with the reference's stopping semantics; returns (x, status,
        (nfev, njev, accepted), cost history per call)."""
        o = LmOpts(**opts)
        x = _f64(x).copy()
        cnt = np.zeros(3, dtype=np.int32)
        hist = np.zeros(max(o.max_iter, 1))
        st = self._lib.or_lm_minpack(self._h, _p(x, _dp), ct.byref(o), _p(cnt, _i32p), _p(hist, _dp))
        return x, int(st), tuple(int(v) for v in cnt), hist

    def population_cost(self, X):
        X = _f64(X)
        c = np.zeros(X.shape[0])
        self._lib.or_population_cost(self._h, _p(X, _dp), X.shape[0], _p(c, _dp))
        return c

    def de(self, x, lo, hi, np_, generations, F=0.65, CR=0.5, seed=42):
        x = _f64(x).copy()
        hist = np.zeros(generations)
        best = self._lib.or_de(self._h, _p(x, _dp), _p(_f64(lo), _dp), _p(_f64(hi), _dp), int(np_),
                               int(generations), float(F), float(CR), int(seed), _p(hist, _dp))
        return x, best, hist

    def time_fd_jacobian(self, x, ncols):
        x = _f64(x)
        J = np.zeros((self.residual_count, ncols))
        return float(self._lib.or_time_fd_jacobian(self._h, _p(x, _dp), int(ncols), _p(J, _dp)))

    def time_residual(self, x, reps=1):
        x = _f64(x)
        F = np.zeros(self.residual_count)
        return float(self._lib.or_time_residual(self._h, _p(x, _dp), int(reps), _p(F, _dp)))


def de_init(x, lo, hi, np_, seed):
    """or_de_init: the initial population (np_ x D)."""
    lib = _load()
    x = _f64(x)
    pop = np.zeros((int(np_), x.size))
    lib.or_de_init(_p(x, _dp), _p(_f64(lo), _dp), _p(_f64(hi), _dp), int(np_), x.size, int(seed), _p(pop, _dp))
    return pop


def de_trial(pop, lo, hi, F, CR, seed, g, m):
    """or_de_trial: (trial of member m in generation g, (r1, r2, r3, jrand))."""
    lib = _load()
    pop = _f64(pop)
    np_, D = pop.shape
    out = np.zeros(D)
    picks = np.zeros(4, dtype=np.int32)
    lib.or_de_trial(_p(pop, _dp), np_, D, _p(_f64(lo), _dp), _p(_f64(hi), _dp), float(F), float(CR),
                    int(seed), int(g), int(m), _p(out, _dp), _p(picks, _i32p))
    return out, tuple(int(v) for v in picks)


def de_select(pop, cost, trial, tcost):
    """or_de_select: (new population, new costs, kept flags, best index)."""
    lib = _load()
    pop = _f64(pop).copy()
    cost = _f64(cost).copy()
    np_, D = pop.shape
    kept = np.zeros(np_, dtype=np.uint8)
    bm = lib.or_de_select(np_, D, _p(pop, _dp), _p(cost, _dp), _p(_f64(trial), _dp), _p(_f64(tcost), _dp),
                          _p(kept, _u8p))
    return pop, cost, kept.astype(bool), int(bm)


def de_fn(fun, x, lo, hi, np_, generations, F=0.65, CR=0.5, seed=42):
    """or_de_fn on a Python cost fun(X[np, D]) -> cost[np]: (x, best, hist)."""
    lib = _load()
    x = _f64(x).copy()
    D = x.size

    def cb(_u, Xp, n, out):
        X = np.ctypeslib.as_array(Xp, shape=(n * D,)).reshape(n, D).copy()
        np.ctypeslib.as_array(out, shape=(n,))[:] = fun(X)

    fcb = POPCOST(cb)
    hist = np.zeros(max(int(generations), 1))
    best = lib.or_de_fn(fcb, None, D, _p(x, _dp), _p(_f64(lo), _dp), _p(_f64(hi), _dp), int(np_),
                        int(generations), float(F), float(CR), int(seed), _p(hist, _dp))
    return x, float(best), hist[:int(generations)]


def lmder(fun, jac, x0, m, **opts):
    """or_lmder on Python callables fun(x) -> f[m], jac(x) -> J[m, n]:
    returns (x, status, (nfev, njev, accepted), cost history)."""
    lib = _load()
    x = _f64(x0).copy()
    n = x.size

    def f_cb(_u, xp, out):
        xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        np.ctypeslib.as_array(out, shape=(m,))[:] = fun(xv)
        return 0

    def j_cb(_u, xp, out):
        xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        np.ctypeslib.as_array(out, shape=(m * n,))[:] = np.asarray(jac(xv), dtype=np.float64).reshape(-1)
        return 0

    fcb, jcb = FCN(f_cb), FCN(j_cb)
    o = LmOpts(**opts)
    cnt = np.zeros(3, dtype=np.int32)
    hist = np.zeros(max(o.max_iter, 1))
    st = lib.or_lmder(int(m), int(n), fcb, jcb, None, _p(x, _dp), ct.byref(o), _p(cnt, _i32p), _p(hist, _dp))
    return x, int(st), tuple(int(v) for v in cnt), hist


def qr_pivot(J, f):
    """Pivoted QR of J (m x n): (rank, perm, R (n x n upper), Q^T f [n])."""
    lib = _load()
    J = _f64(J)
    m, n = J.shape
    perm = np.zeros(n, dtype=np.int32)
    Rc = np.zeros(n * n)
    qtf = np.zeros(n)
    rank = lib.or_qr_pivot(m, n, _p(J, _dp), _p(_f64(f), _dp), _p(perm, _i32p), _p(Rc, _dp), _p(qtf, _dp))
    return int(rank), perm, Rc.reshape(n, n).T.copy(), qtf


def lmpar(R, perm, rank, diag, qtf, delta, par=0.0):
    """lmpar on a pivoted QR: (par, x, damped solves)."""
    lib = _load()
    R = _f64(R)
    n = R.shape[0]
    Rc = np.ascontiguousarray(R.T)
    p = ct.c_double(par)
    x = np.zeros(n)
    k = lib.or_lmpar(n, _p(Rc, _dp), _p(np.ascontiguousarray(perm, dtype=np.int32), _i32p), int(rank),
                     _p(_f64(diag), _dp), _p(_f64(qtf), _dp), float(delta), ct.byref(p), _p(x, _dp))
    return p.value, x, int(k)
