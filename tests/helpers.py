"""Small hand-built configurations and models for the tests (the reference
builds the same kind of fixtures inline: dimers, triangles, tetrahedra and
FCC cells in 100 A boxes, tests/unit/test_eam_force.cpp:70-90)."""
import numpy as np

from forcesmith_amd.model import Dataset, SWModel, Table, TabulatedModel, knot_grid


def make_dataset(configs, types=None, ref_force=None, ref_energy=None, ref_stress=None):
    """configs: list of (positions (n,3), cell (3,3) rows = lattice vectors)."""
    sizes = [len(p) for p, _ in configs]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pos = np.concatenate([np.asarray(p, dtype=float) for p, _ in configs]).reshape(-1, 3)
    n = len(pos)
    if types is None:
        types = np.zeros(n, dtype=np.int32)
    C = len(configs)
    return Dataset(off, np.asarray(types, dtype=np.int32), pos,
                   np.zeros((n, 3)) if ref_force is None else np.asarray(ref_force, float),
                   np.array([np.asarray(c, dtype=float) for _, c in configs]),
                   np.zeros(C) if ref_energy is None else np.asarray(ref_energy, float),
                   np.zeros((C, 6)) if ref_stress is None else np.asarray(ref_stress, float))


def box(side):
    return np.eye(3) * side


def dimer(r, side=100.0):
    return ([[0.0, 0.0, 0.0], [r, 0.0, 0.0]], box(side))


def fcc_cell(a, m=2):
    b = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    pts = [np.array([x, y, z]) * a + bb for x in range(m) for y in range(m) for z in range(m) for bb in b]
    return (np.array(pts), box(a * m))


def table_from(fn, n, rmin, rmax, fixed=None):
    return Table(n, rmin, rmax, fn(knot_grid(n, rmin, rmax)), fixed)


def eam_model(phi, rho, emb, n=30, span=(0.5, 6.0), emb_span=(0.0, 5.0), ntypes=1, **kw):
    p = ntypes * (ntypes + 1) // 2
    tabs = [table_from(phi, n, *span) for _ in range(p)] + [table_from(rho, n, *span) for _ in range(ntypes)] + \
           [table_from(emb, n, *emb_span) for _ in range(ntypes)]
    return TabulatedModel("eam", ntypes, tabs, **kw)


def random_cluster(rng, n, side, dmin=1.9, lo=0.0):
    pts = []
    while len(pts) < n:
        p = rng.uniform(lo, side, size=3)
        if all(np.linalg.norm(p - q) >= dmin for q in pts):
            pts.append(p)
    return np.array(pts)
