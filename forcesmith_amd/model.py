"""Host-side data model mirroring the reference's dataset and parameter layout.

* ``Dataset`` is the structure-of-arrays form of the reference's
  ``std::vector<Configuration>`` (include/forcesmith/core/atom.hpp:38-100):
  per-atom type/position/reference force, per-configuration box, reference
  energy and reference stress.  ``cell[c]`` holds the three lattice vectors as
  rows (the reference reads them as "X", "Y", "Z" and stores them as the
  columns of its box matrix, src/io/config_reader.cpp:58-66).  ``ref_stress``
  is kept in the residual order xx, yy, zz, xy, xz, yz
  (include/forcesmith/core/voigt.hpp:9-13).
* ``Table`` is one tabulated potential as built by ``make_tabulated``
  (src/io/potential_reader.cpp:162-191): ``n`` knot values on the uniform grid
  ``x_k = rmin + k*(rmax-rmin)/(n-1)`` with per-knot fixed flags.
* ``TabulatedModel`` lists the tables in the reference declaration order
  (EAM: pair, density, embedding, include/forcesmith/force/eam_force.hpp:24-32;
  ADP adds dipole, quadrupole, adp_force.hpp:37-51); pair tables are in
  ``pair_ordinal`` order (include/forcesmith/force/potential_table.hpp:31-37).
* ``SWModel`` holds the Stillinger-Weber pair slots (A, B, p, q, delta, a1,
  gamma, a2; include/forcesmith/force/stiweb_force.hpp:19-29) and the per-triplet
  lambda table indexed ``ti*P + slot(tj, tk)`` (stiweb_force.hpp:52-58).

The free-parameter vector ``x`` follows ``TableParams``
(include/forcesmith/force/force_calculator_concept.hpp:169-208): tables in
declaration order, free knots in index order; for SW the free fields of every
pair slot in slot order, then the free lambdas (src/force/stiweb_force.cpp:
169-211).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SW_FIELDS = ("A", "B", "p", "q", "delta", "a1", "gamma", "a2")


def pair_ordinal(a: int, b: int, ntypes: int) -> int:
    """Slot of the unordered type pair (a, b) (potential_table.hpp:31-37)."""
    if a > b:
        a, b = b, a
    return a * ntypes - a * (a - 1) // 2 + (b - a)


def knot_grid(n: int, rmin: float, rmax: float) -> np.ndarray:
    """x_k = rmin + k*h with h = (rmax-rmin)/(n-1), exactly as
    src/io/potential_reader.cpp:176-182 computes it (so x[-1] may differ from
    rmax in the last bit; the table span and the cutoff use x[-1])."""
    h = (rmax - rmin) / float(n - 1)
    return np.array([rmin + float(k) * h for k in range(n)], dtype=np.float64)


@dataclass
class Dataset:
    atom_off: np.ndarray
    types: np.ndarray
    pos: np.ndarray
    ref_force: np.ndarray
    cell: np.ndarray
    ref_energy: np.ndarray
    ref_stress: np.ndarray
    name: str = ""

    @property
    def n_configs(self) -> int:
        return int(len(self.atom_off) - 1)

    @property
    def n_atoms(self) -> int:
        return int(self.atom_off[-1])

    def config_sizes(self) -> np.ndarray:
        return np.diff(self.atom_off)

    def subset(self, idx) -> "Dataset":
        """Configurations ``idx`` as a new dataset (atom order preserved)."""
        idx = list(idx)
        parts = [np.arange(self.atom_off[c], self.atom_off[c + 1]) for c in idx]
        atoms = np.concatenate(parts) if parts else np.zeros(0, dtype=np.int64)
        sizes = np.array([len(p) for p in parts], dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        return Dataset(off, self.types[atoms].copy(), self.pos[atoms].copy(), self.ref_force[atoms].copy(),
                       self.cell[idx].copy(), self.ref_energy[idx].copy(), self.ref_stress[idx].copy(),
                       self.name)


@dataclass
class Table:
    n: int
    rmin: float
    rmax: float
    y: np.ndarray
    fixed: np.ndarray = None

    def __post_init__(self):
        self.y = np.asarray(self.y, dtype=np.float64).copy()
        if self.fixed is None:
            self.fixed = np.zeros(self.n, dtype=bool)
        self.fixed = np.asarray(self.fixed, dtype=bool).copy()
        assert len(self.y) == self.n and len(self.fixed) == self.n and self.n >= 2

    @property
    def x(self) -> np.ndarray:
        return knot_grid(self.n, self.rmin, self.rmax)

    @property
    def free_count(self) -> int:
        return int((~self.fixed).sum())


@dataclass
class TabulatedModel:
    kind: str  # "eam" or "adp"
    ntypes: int
    tables: list
    energy_weight: float = 1.0
    stress_weight: float = 0.0

    @property
    def npair(self) -> int:
        return self.ntypes * (self.ntypes + 1) // 2

    def expected_tables(self) -> int:
        p, t = self.npair, self.ntypes
        return p + 2 * t if self.kind == "eam" else 3 * p + 2 * t

    def table_roles(self):
        """(role, index) per table in declaration order."""
        p, t = self.npair, self.ntypes
        roles = [("pair", k) for k in range(p)] + [("density", k) for k in range(t)] + \
                [("embedding", k) for k in range(t)]
        if self.kind == "adp":
            roles += [("dipole", k) for k in range(p)] + [("quadrupole", k) for k in range(p)]
        return roles

    @property
    def param_count(self) -> int:
        return sum(t.free_count for t in self.tables)

    def gather(self) -> np.ndarray:
        return np.concatenate([t.y[~t.fixed] for t in self.tables]) if self.tables else np.zeros(0)

    def scatter(self, x) -> None:
        x = np.asarray(x, dtype=np.float64)
        o = 0
        for t in self.tables:
            k = t.free_count
            t.y[~t.fixed] = x[o:o + k]
            o += k

    def copy(self) -> "TabulatedModel":
        return TabulatedModel(self.kind, self.ntypes,
                              [Table(t.n, t.rmin, t.rmax, t.y.copy(), t.fixed.copy()) for t in self.tables],
                              self.energy_weight, self.stress_weight)

    def max_cutoff(self) -> float:
        """Largest span end over the radial tables (embedding excluded),
        TableParams::max_cutoff (force_calculator_concept.hpp:210-221)."""
        m = 0.0
        for (role, _), t in zip(self.table_roles(), self.tables):
            if role != "embedding":
                m = max(m, float(t.x[-1]))
        return m


@dataclass
class SWModel:
    ntypes: int
    pair_params: np.ndarray  # (npair, 8)
    pair_fixed: np.ndarray   # (npair, 8) bool
    lam: np.ndarray          # (ntypes * npair,)
    lam_fixed: np.ndarray    # bool
    energy_weight: float = 1.0
    stress_weight: float = 0.0
    kind: str = field(default="sw")

    def __post_init__(self):
        self.pair_params = np.asarray(self.pair_params, dtype=np.float64).reshape(-1, 8).copy()
        self.pair_fixed = np.asarray(self.pair_fixed, dtype=bool).reshape(-1, 8).copy()
        self.lam = np.asarray(self.lam, dtype=np.float64).ravel().copy()
        self.lam_fixed = np.asarray(self.lam_fixed, dtype=bool).ravel().copy()

    @property
    def npair(self) -> int:
        return self.ntypes * (self.ntypes + 1) // 2

    @property
    def param_count(self) -> int:
        return int((~self.pair_fixed).sum() + (~self.lam_fixed).sum())

    def gather(self) -> np.ndarray:
        return np.concatenate([self.pair_params[~self.pair_fixed], self.lam[~self.lam_fixed]])

    def scatter(self, x) -> None:
        x = np.asarray(x, dtype=np.float64)
        k = int((~self.pair_fixed).sum())
        self.pair_params[~self.pair_fixed] = x[:k]
        self.lam[~self.lam_fixed] = x[k:]

    def copy(self) -> "SWModel":
        return SWModel(self.ntypes, self.pair_params, self.pair_fixed, self.lam, self.lam_fixed,
                       self.energy_weight, self.stress_weight)

    def max_cutoff(self) -> float:
        return float(np.max(np.maximum(self.pair_params[:, 5], self.pair_params[:, 7])))
