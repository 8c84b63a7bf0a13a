"""MI355X-native force-matching evaluator.

Python host mirror of the reference's force-calculator surface over the
C ABI of ``libforcesmith_amd.so`` (declared in ``include/forcesmith_amd.h``).
``GpuCalculator`` mirrors the erased ``ForceCalculator`` of the reference
(include/forcesmith/force/force_calculator.hpp:84-159): ``prepare``,
``eval_forces``, ``eval_cached``, ``param_count``, ``gather_params``,
``scatter_params``, ``gather_bounds``, ``max_cutoff``; ``residual`` and
``jacobian`` mirror ``ForcesmithFunctor::operator()`` / ``df``
(src/optimization/forcesmith_functor.cpp:221-292).

There is no CPU fallback: importing works anywhere, but constructing a
calculator requires the built HIP library and a HIP device and raises
otherwise.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from .model import Dataset, SWModel, Table, TabulatedModel, knot_grid, pair_ordinal  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libforcesmith_amd.so")

FS_EAM, FS_ADP, FS_SW = 0, 1, 2
FS_MIXED, FS_FP64 = 0, 1
STATUS = {0: "FS_OK", -1: "FS_ERR_ARG", -2: "FS_ERR_STATE", -3: "FS_ERR_HIP", -4: "FS_ERR_NOMEM",
          -5: "FS_ERR_UNSUPPORTED", -6: "FS_ERR_NUMERIC"}

# every symbol include/forcesmith_amd.h declares
EXPORTS = (
    "fs_abi_version", "fs_ctx_create", "fs_ctx_retain", "fs_ctx_release", "fs_last_error",
    "fs_upload_dataset", "fs_set_tabulated_model", "fs_set_sw_model", "fs_set_weights", "fs_set_smooth_weight", "fs_smoothness_count", "fs_write_smoothness",
    "fs_param_count", "fs_gather_params", "fs_scatter_params", "fs_gather_bounds", "fs_max_cutoff",
    "fs_residual_count", "fs_row_offsets", "fs_prepare", "fs_neighbor_count", "fs_get_neighbor_list",
    "fs_eval_forces", "fs_eval_config", "fs_eval_residual", "fs_eval_jacobian", "fs_eval_step", "fs_normal_equations", "fs_normal_equations_noalloc",
    "fs_lm_minimize", "fs_lm_solve", "fs_normal_equations_solve", "fs_lm_step", "fs_population_cost", "fs_de_minimize", "fs_de_last_stats", "fs_device_buffer", "fs_synchronize",
    "fs_eam_set_gauge", "fs_eam_get_gauge", "fs_eam_rho_stats", "fs_eam_rescale_rho_axis",
    "fs_eam_embed_shift", "fs_eam_rescale", "fs_lm_last_timing",
    "fs_set_launch", "fs_set_jacobian_tile", "fs_set_jacobian_form", "fs_set_population_form", "fs_get_launch", "fs_autotune", "fs_time_step", "fs_mlp_forward", "fs_mlp_loss_grad",
)


class FsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class fs_opts(ct.Structure):
    _fields_ = [("device", ct.c_int32), ("precision", ct.c_int32), ("reserved", ct.c_int32 * 6)]


class fs_lm_opts(ct.Structure):
    _fields_ = [("max_iter", ct.c_int32), ("maxfev", ct.c_int32), ("xtol", ct.c_double),
                ("ftol", ct.c_double), ("gtol", ct.c_double), ("factor", ct.c_double)]


class fs_de_opts(ct.Structure):
    _fields_ = [("np", ct.c_int32), ("generations", ct.c_int32), ("F", ct.c_double), ("CR", ct.c_double),
                ("seed", ct.c_uint64), ("surrogate", ct.c_int32), ("oversample", ct.c_int32),
                ("retrain_every", ct.c_int32), ("hidden", ct.c_int32), ("target_cost", ct.c_double),
                ("rank_tau_min", ct.c_double), ("progress_every", ct.c_int32), ("reserved", ct.c_int32 * 3)]


class fs_de_stats(ct.Structure):
    _fields_ = [("generations", ct.c_int32), ("retrains", ct.c_int32), ("screened", ct.c_int32),
                ("fallbacks", ct.c_int32), ("first_loss", ct.c_float), ("last_loss", ct.c_float),
                ("reserved", ct.c_int32 * 2)]


_lib = None
_dp = ct.POINTER(ct.c_double)
_fp = ct.POINTER(ct.c_float)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)
_u8p = ct.POINTER(ct.c_uint8)


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m forcesmith_amd.build` "
                                "(there is no CPU fallback)")
    lib = ct.CDLL(path)
    vp = ct.c_void_p
    sig = {
        "fs_abi_version": (ct.c_int, []),
        "fs_ctx_create": (ct.c_int, [ct.POINTER(fs_opts), ct.POINTER(vp)]),
        "fs_ctx_retain": (ct.c_int, [vp]),
        "fs_ctx_release": (ct.c_int, [vp]),
        "fs_last_error": (ct.c_char_p, [vp]),
        "fs_upload_dataset": (ct.c_int, [vp, ct.c_int32, _i64p, _i32p, _dp, _dp, _dp, _dp, _dp]),
        "fs_set_tabulated_model": (ct.c_int, [vp, ct.c_int32, ct.c_int32, ct.c_int32, _i32p, _dp, _dp, _dp,
                                              _u8p]),
        "fs_set_sw_model": (ct.c_int, [vp, ct.c_int32, _dp, _u8p, _dp, _u8p]),
        "fs_set_weights": (ct.c_int, [vp, ct.c_double, ct.c_double]),
        "fs_set_smooth_weight": (ct.c_int, [vp, ct.c_double]),
        "fs_smoothness_count": (ct.c_int, [vp, _i64p]),
        "fs_write_smoothness": (ct.c_int, [vp, ct.c_double, _dp]),
        "fs_param_count": (ct.c_int, [vp, _i32p]),
        "fs_gather_params": (ct.c_int, [vp, _dp]),
        "fs_scatter_params": (ct.c_int, [vp, _dp]),
        "fs_gather_bounds": (ct.c_int, [vp, _dp, _dp]),
        "fs_max_cutoff": (ct.c_int, [vp, _dp]),
        "fs_residual_count": (ct.c_int, [vp, _i64p]),
        "fs_row_offsets": (ct.c_int, [vp, _i64p]),
        "fs_prepare": (ct.c_int, [vp]),
        "fs_neighbor_count": (ct.c_int, [vp, _i64p]),
        "fs_get_neighbor_list": (ct.c_int, [vp, _i64p, _i32p, _i32p, _dp]),
        "fs_eval_forces": (ct.c_int, [vp, _dp, _dp, _dp, _dp]),
        "fs_eval_config": (ct.c_int, [vp, ct.c_int32, _dp, _dp, _dp, _dp]),
        "fs_eval_residual": (ct.c_int, [vp, _dp, _dp]),
        "fs_eval_jacobian": (ct.c_int, [vp, _fp]),
        "fs_eval_step": (ct.c_int, [vp, _dp, _dp, _fp, _dp]),
        "fs_normal_equations": (ct.c_int, [vp, _dp, _dp, _dp]),
        "fs_normal_equations_noalloc": (ct.c_int, [vp, vp, vp, ct.c_int64, ct.c_int32, vp, vp, vp]),
        "fs_lm_minimize": (ct.c_int, [vp, _dp, ct.c_int32, _dp, _i32p]),
        "fs_lm_solve": (ct.c_int, [vp, ct.POINTER(fs_lm_opts), _dp, _i32p, _i32p, _dp]),
        "fs_normal_equations_solve": (ct.c_int, [vp, ct.c_double, _dp]),
        "fs_lm_step": (ct.c_int, [vp, ct.c_double, _dp]),
        "fs_population_cost": (ct.c_int, [vp, _dp, ct.c_int32, _dp]),
        "fs_de_minimize": (ct.c_int, [vp, ct.POINTER(fs_de_opts), _dp, _dp, _dp, _dp, _dp, _i32p]),
        "fs_de_last_stats": (ct.c_int, [vp, ct.POINTER(fs_de_stats)]),
        "fs_eam_set_gauge": (ct.c_int, [vp, ct.c_double, _dp, _dp]),
        "fs_eam_get_gauge": (ct.c_int, [vp, _dp, _dp, _dp]),
        "fs_eam_rho_stats": (ct.c_int, [vp, _dp, _dp, _dp, _i64p]),
        "fs_eam_rescale_rho_axis": (ct.c_int, [vp, _dp]),
        "fs_eam_embed_shift": (ct.c_int, [vp, _dp]),
        "fs_eam_rescale": (ct.c_int, [vp, _dp, _dp]),
        "fs_lm_last_timing": (ct.c_int, [vp, _dp]),
        "fs_device_buffer": (ct.c_int, [vp, ct.c_int32, ct.POINTER(vp), _i64p]),
        "fs_synchronize": (ct.c_int, [vp]),
        "fs_set_launch": (ct.c_int, [vp, ct.c_int32, ct.c_int32]),
        "fs_set_jacobian_tile": (ct.c_int, [vp, ct.c_int32]),
        "fs_set_jacobian_form": (ct.c_int, [vp, ct.c_int32]),
        "fs_set_population_form": (ct.c_int, [vp, ct.c_int32]),
        "fs_get_launch": (ct.c_int, [vp, _i32p, _i32p, _i32p, _i32p]),
        "fs_autotune": (ct.c_int, [vp, _i32p]),
        "fs_time_step": (ct.c_int, [vp, ct.c_int32, ct.c_int32, _dp]),
        "fs_mlp_forward": (ct.c_int, [vp, ct.c_int32, ct.c_int32, ct.c_int32, _fp, _fp, _fp]),
        "fs_mlp_loss_grad": (ct.c_int, [vp, ct.c_int32, ct.c_int32, ct.c_int32, _fp, _fp, _fp, _fp, _fp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a, typ):
    return a.ctypes.data_as(typ) if a is not None else ct.cast(None, typ)


class GpuCalculator:
    """A dataset plus a model resident on one HIP device (one fs_ctx)."""

    def __init__(self, dataset: Dataset, model, precision: str = "mixed", device: int = 0,
                 energy_weight=None, stress_weight=None, smooth_weight: float = 0.0):
        lib = load_library()
        self._lib = lib
        opts = fs_opts(device, FS_FP64 if precision == "fp64" else FS_MIXED)
        h = ct.c_void_p()
        rc = lib.fs_ctx_create(ct.byref(opts), ct.byref(h))
        if rc != 0:
            raise FsError(rc, lib.fs_last_error(None).decode())
        self._h = h
        self.precision = precision
        self.dataset = dataset
        self.kind = model.kind
        off = np.ascontiguousarray(dataset.atom_off, dtype=np.int64)
        self._check(lib.fs_upload_dataset(h, dataset.n_configs, _p(off, _i64p),
                                          _p(np.ascontiguousarray(dataset.types, np.int32), _i32p),
                                          _p(np.ascontiguousarray(dataset.pos, np.float64), _dp),
                                          _p(np.ascontiguousarray(dataset.ref_force, np.float64), _dp),
                                          _p(np.ascontiguousarray(dataset.cell, np.float64), _dp),
                                          _p(np.ascontiguousarray(dataset.ref_energy, np.float64), _dp),
                                          _p(np.ascontiguousarray(dataset.ref_stress, np.float64), _dp)))
        self.set_model(model)
        ew = model.energy_weight if energy_weight is None else energy_weight
        sw = model.stress_weight if stress_weight is None else stress_weight
        self._check(lib.fs_set_weights(h, float(ew), float(sw)))
        if smooth_weight:
            self.set_smooth_weight(smooth_weight)
        self.n_atoms = dataset.n_atoms
        self.n_configs = dataset.n_configs

    def _check(self, rc):
        if rc != 0:
            raise FsError(rc, self._lib.fs_last_error(self._h).decode())

    def set_model(self, model):
        lib, h = self._lib, self._h
        self.ntypes = int(model.ntypes)
        if model.kind == "sw":
            self._check(lib.fs_set_sw_model(h, model.ntypes, _p(np.ascontiguousarray(model.pair_params), _dp),
                                            _p(np.ascontiguousarray(model.pair_fixed, np.uint8), _u8p),
                                            _p(np.ascontiguousarray(model.lam), _dp),
                                            _p(np.ascontiguousarray(model.lam_fixed, np.uint8), _u8p)))
        else:
            kind = FS_EAM if model.kind == "eam" else FS_ADP
            n = np.array([t.n for t in model.tables], dtype=np.int32)
            rmin = np.array([t.rmin for t in model.tables], dtype=np.float64)
            rmax = np.array([t.rmax for t in model.tables], dtype=np.float64)
            y = np.concatenate([t.y for t in model.tables]).astype(np.float64)
            fx = np.concatenate([t.fixed for t in model.tables]).astype(np.uint8)
            self._check(lib.fs_set_tabulated_model(h, kind, model.ntypes, len(model.tables), _p(n, _i32p),
                                                   _p(rmin, _dp), _p(rmax, _dp), _p(y, _dp), _p(fx, _u8p)))

    def smoothness_count(self) -> int:
        v = ct.c_int64()
        self._check(self._lib.fs_smoothness_count(self._h, ct.byref(v)))
        return int(v.value)

    def write_smoothness(self, weight: float):
        out = np.zeros(max(self.smoothness_count(), 1))
        self._check(self._lib.fs_write_smoothness(self._h, float(weight), _p(out, _dp)))
        return out[:self.smoothness_count()]

    def lm_step(self, lam: float, x=None):
        """dx = -(J^T J + lam I)^-1 J^T F at the current (or given) parameters."""
        if x is not None:
            self.scatter_params(x)
        dx = np.zeros(max(self.param_count, 1))
        self._check(self._lib.fs_lm_step(self._h, float(lam), _p(dx, _dp)))
        return dx[:self.param_count]

    def set_smooth_weight(self, w: float):
        """Curvature rows w (y[k-1] - 2 y[k] + y[k+1]) after the config rows."""
        self._check(self._lib.fs_set_smooth_weight(self._h, float(w)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fs_ctx_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- CFittable surface ----
    @property
    def param_count(self) -> int:
        v = ct.c_int32()
        self._check(self._lib.fs_param_count(self._h, ct.byref(v)))
        return int(v.value)

    def gather_params(self):
        x = np.zeros(self.param_count)
        self._check(self._lib.fs_gather_params(self._h, _p(x, _dp)))
        return x

    def scatter_params(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._check(self._lib.fs_scatter_params(self._h, _p(x, _dp)))

    def gather_bounds(self):
        lo, hi = np.zeros(self.param_count), np.zeros(self.param_count)
        self._check(self._lib.fs_gather_bounds(self._h, _p(lo, _dp), _p(hi, _dp)))
        return lo, hi

    @property
    def max_cutoff(self) -> float:
        v = ct.c_double()
        self._check(self._lib.fs_max_cutoff(self._h, ct.byref(v)))
        return v.value

    @property
    def residual_count(self) -> int:
        v = ct.c_int64()
        self._check(self._lib.fs_residual_count(self._h, ct.byref(v)))
        return int(v.value)

    def row_offsets(self):
        r = np.zeros(self.n_configs + 1, dtype=np.int64)
        self._check(self._lib.fs_row_offsets(self._h, _p(r, _i64p)))
        return r

    # ---- prepare / neighbour list ----
    def prepare(self) -> int:
        self._check(self._lib.fs_prepare(self._h))
        v = ct.c_int64()
        self._check(self._lib.fs_neighbor_count(self._h, ct.byref(v)))
        return int(v.value)

    def neighbor_list(self):
        v = ct.c_int64()
        self._check(self._lib.fs_neighbor_count(self._h, ct.byref(v)))
        P = int(v.value)
        off = np.zeros(self.n_atoms + 1, dtype=np.int64)
        j = np.zeros(P, dtype=np.int32)
        img = np.zeros((P, 3), dtype=np.int32)
        d = np.zeros((P, 3))
        self._check(self._lib.fs_get_neighbor_list(self._h, _p(off, _i64p), _p(j, _i32p), _p(img, _i32p),
                                                   _p(d, _dp)))
        return off, j, img, d

    # ---- evaluation ----
    def eval_forces(self):
        F = np.zeros((self.n_atoms, 3))
        E = np.zeros(self.n_configs)
        S = np.zeros((self.n_configs, 3, 3))
        L = np.zeros(self.n_configs)
        self._check(self._lib.fs_eval_forces(self._h, _p(F, _dp), _p(E, _dp), _p(S, _dp), _p(L, _dp)))
        return F, E, S, L

    def eval_cached(self, config: int):
        na = int(self.dataset.atom_off[config + 1] - self.dataset.atom_off[config])
        F = np.zeros((na, 3))
        E, L = ct.c_double(), ct.c_double()
        S = np.zeros((3, 3))
        self._check(self._lib.fs_eval_config(self._h, int(config), _p(F, _dp), ct.byref(E), _p(S, _dp),
                                             ct.byref(L)))
        return F, E.value, S, L.value

    def residual(self, x=None, with_cost=False):
        if x is not None:
            self.scatter_params(x)
        F = np.zeros(self.residual_count)
        c = ct.c_double()
        self._check(self._lib.fs_eval_residual(self._h, _p(F, _dp), ct.byref(c)))
        return (F, c.value) if with_cost else F

    def cost(self, x=None) -> float:
        if x is not None:
            self.scatter_params(x)
        c = ct.c_double()
        self._check(self._lib.fs_eval_residual(self._h, ct.cast(None, _dp), ct.byref(c)))
        return c.value

    def jacobian(self, x=None):
        if x is not None:
            self.scatter_params(x)
        J = np.zeros((self.residual_count, self.param_count), dtype=np.float32)
        self._check(self._lib.fs_eval_jacobian(self._h, _p(J, _fp)))
        return J

    def step(self, x, want_residual=False, want_jacobian=False, want_cost=True):
        """One residual + Jacobian step at x, replayed from a captured graph
        after the first call: cost (None with want_cost=False: the step is
        then only queued), and F / J when asked for (else None)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        F = np.zeros(self.residual_count) if want_residual else None
        J = np.zeros((self.residual_count, self.param_count), np.float32) if want_jacobian else None
        c = ct.c_double()
        self._check(self._lib.fs_eval_step(self._h, _p(x, _dp), _p(F, _dp), _p(J, _fp),
                                           ct.byref(c) if want_cost else None))
        return (c.value if want_cost else None), F, J

    def normal_equations(self):
        D = self.param_count
        A, g, c = np.zeros((D, D)), np.zeros(D), ct.c_double()
        self._check(self._lib.fs_normal_equations(self._h, _p(A, _dp), _p(g, _dp), ct.byref(c)))
        return A, g, c.value

    def lm(self, x, max_iter):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        hist = np.zeros(max_iter)
        acc = ct.c_int32()
        self._check(self._lib.fs_lm_minimize(self._h, _p(x, _dp), int(max_iter), _p(hist, _dp), ct.byref(acc)))
        return x, hist, int(acc.value)

    def ne_solve(self, x, ridge=0.0):
        """NormalEquationsSolver: one step x - (J^T J + lambda I)^-1 J^T F."""
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        self._check(self._lib.fs_normal_equations_solve(self._h, float(ridge), _p(x, _dp)))
        return x

    def lm_solve(self, x, max_iter=500, maxfev=0, xtol=1e-7, ftol=1e-7, gtol=0.0, factor=100.0):
        """EigenLMSolver semantics: (x, status, (nfev, njev, accepted), cost
        history per call of one step)."""
        o = fs_lm_opts(int(max_iter), int(maxfev), float(xtol), float(ftol), float(gtol), float(factor))
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        st = ct.c_int32()
        cnt = np.zeros(3, dtype=np.int32)
        hist = np.zeros(max(int(max_iter), 1))
        self._check(self._lib.fs_lm_solve(self._h, ct.byref(o), _p(x, _dp), ct.byref(st), _p(cnt, _i32p),
                                          _p(hist, _dp)))
        return x, int(st.value), tuple(int(v) for v in cnt), hist

    def lm_timing(self) -> dict:
        """Wall time of the last lm_solve (ms): whole, device calls, host algebra."""
        ms = np.zeros(3)
        self._check(self._lib.fs_lm_last_timing(self._h, _p(ms, _dp)))
        return {"total_ms": float(ms[0]), "device_ms": float(ms[1]), "host_algebra_ms": float(ms[2])}

    def population_cost(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        c = np.zeros(X.shape[0])
        self._check(self._lib.fs_population_cost(self._h, _p(X, _dp), X.shape[0], _p(c, _dp)))
        return c

    def de(self, x, lo, hi, np_, generations, F=0.65, CR=0.5, seed=42, surrogate=False, oversample=4,
           retrain_every=10, hidden=256, target_cost=0.0, rank_tau_min=0.3, progress_every=0):
        o = fs_de_opts(int(np_), int(generations), float(F), float(CR), int(seed), int(bool(surrogate)),
                       int(oversample), int(retrain_every), int(hidden), float(target_cost),
                       float(rank_tau_min), int(progress_every))
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        best = ct.c_double()
        hist = np.zeros(generations)
        gens = ct.c_int32()
        self._check(self._lib.fs_de_minimize(self._h, ct.byref(o), _p(np.ascontiguousarray(lo, np.float64), _dp),
                                             _p(np.ascontiguousarray(hi, np.float64), _dp), _p(x, _dp),
                                             ct.byref(best), _p(hist, _dp), ct.byref(gens)))
        return x, best.value, hist[:gens.value]

    def de_stats(self) -> dict:
        """Statistics of the last de() call: generations, retrains, screened
        generations, rank-error fallbacks, first and last training loss."""
        st = fs_de_stats()
        self._check(self._lib.fs_de_last_stats(self._h, ct.byref(st)))
        return {"generations": st.generations, "retrains": st.retrains, "screened": st.screened,
                "fallbacks": st.fallbacks, "first_loss": st.first_loss, "last_loss": st.last_loss}

    # ---- EAM gauge (src/core/rescale.cpp, core/scale_potentials.hpp) ----
    def set_gauge(self, scale=1.0, slope=None, intercept=None):
        """Density scale A, per-type embedding slopes S and intercepts F0:
        density A g, embedding F(rho / A) - S rho - F0, pairs compensated
        by S_i A g_j + S_j A g_i."""
        nt = self.ntypes
        S = np.ascontiguousarray(np.zeros(nt) if slope is None else slope, dtype=np.float64)
        F0 = np.ascontiguousarray(np.zeros(nt) if intercept is None else intercept, dtype=np.float64)
        self._check(self._lib.fs_eam_set_gauge(self._h, float(scale), _p(S, _dp), _p(F0, _dp)))

    def gauge(self):
        """(A, S[ntypes], F0[ntypes])."""
        A = np.zeros(1)
        S, F0 = np.zeros(self.ntypes), np.zeros(self.ntypes)
        self._check(self._lib.fs_eam_get_gauge(self._h, _p(A, _dp), _p(S, _dp), _p(F0, _dp)))
        return float(A[0]), S, F0

    def rho_stats(self):
        """Per type: min, max and mean of the clamped densities (the gauge's
        units) over every configuration, and the atom count."""
        nt = self.ntypes
        lo, hi, mean = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        cnt = np.zeros(nt, dtype=np.int64)
        self._check(self._lib.fs_eam_rho_stats(self._h, _p(lo, _dp), _p(hi, _dp), _p(mean, _dp),
                                               _p(cnt, _i64p)))
        return lo, hi, mean, cnt

    def compute_rho_ref(self):
        """compute_rho_ref (rescale.cpp:11-33): mean clamped density per type."""
        return self.rho_stats()[2]

    def rescale_rho_axis(self) -> float:
        """rescale_rho_axis (rescale.cpp:35-101); returns the scale applied."""
        a = np.zeros(1)
        self._check(self._lib.fs_eam_rescale_rho_axis(self._h, _p(a, _dp)))
        return float(a[0])

    def embed_shift(self, rho_ref):
        """embed_shift (rescale.cpp:103-136) at the reference densities."""
        r = np.ascontiguousarray(rho_ref, dtype=np.float64)
        self._check(self._lib.fs_eam_embed_shift(self._h, _p(r, _dp)))

    def rescale_eam(self):
        """rescale_eam (rescale.cpp:138-158); returns (a, rho_ref)."""
        a, ref = np.zeros(1), np.zeros(self.ntypes)
        self._check(self._lib.fs_eam_rescale(self._h, _p(a, _dp), _p(ref, _dp)))
        return float(a[0]), ref

    # ---- launch shape / timing ----
    def set_launch(self, lanes_per_atom: int, blocks_per_cu: int = 8):
        self._check(self._lib.fs_set_launch(self._h, int(lanes_per_atom), int(blocks_per_cu)))

    def set_jacobian_tile(self, atoms: int = 0):
        """Atoms per Jacobian tile (4, 8, 16, 32 or 64; 0 picks it from the
        atom count)."""
        self._check(self._lib.fs_set_jacobian_tile(self._h, int(atoms)))

    def set_jacobian_form(self, form: int = 0):
        """Form of the pair-table columns and density rows: 0 knot sums while
        they fit (else window passes), 1 window passes, 2 knot sums."""
        self._check(self._lib.fs_set_jacobian_form(self._h, int(form)))

    def set_population_form(self, form: int = 0):
        """Form of the batched population cost: 0 member-lane passes where
        they apply (one member per lane, FP64), 1 member-block passes."""
        self._check(self._lib.fs_set_population_form(self._h, int(form)))

    def launch_shape(self) -> dict:
        lpa, bpc, jt, jb = ct.c_int32(), ct.c_int32(), ct.c_int32(), ct.c_int32()
        self._check(self._lib.fs_get_launch(self._h, ct.byref(lpa), ct.byref(bpc), ct.byref(jt), ct.byref(jb)))
        return {"lanes_per_atom": int(lpa.value), "blocks_per_cu": int(bpc.value),
                "jacobian_tile_atoms": int(jt.value), "jacobian_blocks_per_cu": int(jb.value)}

    def autotune(self) -> int:
        """Times the launch shapes on the resident dataset and keeps the
        fastest: lanes per atom and blocks per CU over the residual, then
        atoms per Jacobian tile over the whole Jacobian. Returns the lanes per
        atom (launch_shape() has the whole shape)."""
        v = ct.c_int32()
        self._check(self._lib.fs_autotune(self._h, ct.byref(v)))
        return int(v.value)

    def time_step(self, what: int = 0, reps: int = 10) -> float:
        v = ct.c_double()
        self._check(self._lib.fs_time_step(self._h, int(what), int(reps), ct.byref(v)))
        return v.value

    def synchronize(self):
        self._check(self._lib.fs_synchronize(self._h))

    # ---- MLP surrogate kernels (host buffers) ----
    def mlp_forward(self, params, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, d = X.shape
        h = mlp_hidden(len(params), d)
        y = np.zeros(n, dtype=np.float32)
        self._check(self._lib.fs_mlp_forward(self._h, n, d, h, _p(np.ascontiguousarray(params, np.float32), _fp),
                                             _p(X, _fp), _p(y, _fp)))
        return y

    def mlp_loss_grad(self, params, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, d = X.shape
        h = mlp_hidden(len(params), d)
        g = np.zeros(len(params), dtype=np.float32)
        loss = ct.c_float()
        self._check(self._lib.fs_mlp_loss_grad(self._h, n, d, h, _p(np.ascontiguousarray(params, np.float32), _fp),
                                               _p(X, _fp), _p(np.ascontiguousarray(t, np.float32), _fp),
                                               _p(g, _fp), ct.byref(loss)))
        return loss.value, g


def mlp_hidden(n_params: int, d: int) -> int:
    """Width h of the D -> h -> h -> 1 MLP with n_params parameters."""
    # n = h*d + h + h*h + h + h + 1
    a, b, c = 1.0, d + 3.0, 1.0 - n_params
    h = int(round((-b + (b * b - 4 * a * c) ** 0.5) / 2))
    assert h * d + h * h + 3 * h + 1 == n_params
    return h


def mlp_pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, W3, [b3]]).astype(np.float32)


def mlp_unpack(p, d, h):
    o = 0
    W1 = p[o:o + h * d].reshape(h, d); o += h * d
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    W3 = p[o:o + h]; o += h
    return W1, b1, W2, b2, W3, float(p[o])


def stress_voigt(S):
    """3x3 stresses -> residual order xx, yy, zz, xy, xz, yz."""
    S = np.asarray(S)
    return np.stack([S[..., 0, 0], S[..., 1, 1], S[..., 2, 2], S[..., 0, 1], S[..., 0, 2], S[..., 1, 2]], -1)
