"""Seeded synthetic workloads for the five benchmark configurations.

Each generator returns a ``Workload``: the dataset (positions, boxes and
reference forces/energies/stresses from an analytic truth model plus Gaussian
noise), the starting tabulated (or SW) model to fit, and the truth parameters.
Reference data follow the recipe of the reference's NIST-style fixture
(tests/integration/test_nist_eam.cpp:115-140): evaluate a truth model, then
add N(0, sigma) noise to every force component and to the energy.

The truth models use the reference's analytic forms
(include/forcesmith/potentials/analytic_potential.hpp): Morse
De*(1-exp(-a(r-re)))^2 - De (:316-326), ExpDecay A*exp(-B r) (:366-373) and
SqrtFunc A*sqrt(r/B) (:526-533), gated to their span as the reference gates
every non-cached table evaluation (src/force/eam_force.cpp:19-43).

Everything here is host-side NumPy/SciPy in float64; it builds inputs, it is
not the evaluator.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import cKDTree

from .model import Dataset, SWModel, Table, TabulatedModel, knot_grid, pair_ordinal

# noise levels: forces from the configuration text, energy as the reference
# fixture (test_nist_eam.cpp:203), stress a small fraction of typical values
SIGMA_F = 0.01
SIGMA_E = 0.001
SIGMA_S = 1e-4


@dataclass
class Workload:
    name: str
    dataset: Dataset
    model: object  # TabulatedModel or SWModel (start point of the fit)
    truth: dict = field(default_factory=dict)
    meta: dict = field(default_factory=dict)


# ----------------------------------------------------------------------------
# geometry helpers
# ----------------------------------------------------------------------------
def _fcc(a: float, m) -> np.ndarray:
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(k) for k in m], indexing="ij"), -1).reshape(-1, 3)
    return ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a


def _b2(a: float, m) -> np.ndarray:
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(k) for k in m], indexing="ij"), -1).reshape(-1, 3)
    return ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a


def _diamond(a: float, m) -> np.ndarray:
    f = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    basis = np.concatenate([f, f + 0.25])
    g = np.stack(np.meshgrid(*[np.arange(k) for k in m], indexing="ij"), -1).reshape(-1, 3)
    return ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a


def random_packing(rng, n: int, L: float, dmin: float) -> np.ndarray:
    """Random sequential addition of n points in a periodic cube of side L with
    a hard-core minimum distance dmin (minimum-image; dmin < L/2)."""
    acc = np.zeros((0, 3))
    while len(acc) < n:
        need = n - len(acc)
        cand = rng.uniform(0.0, L, size=(max(4 * need, 64), 3))
        if len(acc):
            t = cKDTree(acc, boxsize=L)
            dd, _ = t.query(cand, k=1)
            cand = cand[dd >= dmin]
        if len(cand) == 0:
            continue
        ct = cKDTree(cand, boxsize=L)
        pairs = ct.query_pairs(dmin, output_type="ndarray")
        bad = [[] for _ in range(len(cand))]
        for a, b in pairs:
            lo, hi = (a, b) if a < b else (b, a)
            bad[hi].append(lo)
        keep = np.zeros(len(cand), dtype=bool)
        for k in range(len(cand)):
            if not any(keep[q] for q in bad[k]):
                keep[k] = True
        acc = np.concatenate([acc, cand[keep][:need]])
    return acc


def neighbor_pairs(pos: np.ndarray, cell: np.ndarray, rc: float):
    """All directed pairs (i, j, d = x_j + n.cell - x_i) with |d| <= rc over
    periodic images, for orthorhombic or general cells (image replication)."""
    inv = np.linalg.inv(cell.T)
    s = [int(np.ceil(rc * np.linalg.norm(inv[a]))) for a in range(3)]
    shifts = np.stack(np.meshgrid(*[np.arange(-k, k + 1) for k in s], indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    rep = (pos[None, :, :] + (shifts @ cell)[:, None, :]).reshape(-1, 3)
    jrep = np.tile(np.arange(n), len(shifts))
    home = np.repeat(np.all(shifts == 0, axis=1), n)
    t_rep = cKDTree(rep)
    t_pos = cKDTree(pos)
    m = t_pos.sparse_distance_matrix(t_rep, rc, output_type="coo_matrix")
    i, k = m.row, m.col
    keep = ~(home[k] & (jrep[k] == i))
    # sparse_distance_matrix drops exact zeros; coincident atoms are not generated
    i, k = i[keep], k[keep]
    d = rep[k] - pos[i]
    order = np.lexsort((k, i))
    return i[order], jrep[k][order], d[order]


def _finish(rng, ds_parts, name, with_stress=True):
    """Assemble per-config parts and add noise."""
    sizes = np.array([len(p["pos"]) for p in ds_parts], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    types = np.concatenate([p["types"] for p in ds_parts]).astype(np.int32)
    pos = np.concatenate([p["pos"] for p in ds_parts])
    force = np.concatenate([p["force"] for p in ds_parts])
    force = force + rng.normal(0.0, SIGMA_F, size=force.shape)
    energy = np.array([p["energy"] for p in ds_parts]) + rng.normal(0.0, SIGMA_E, size=len(ds_parts))
    stress = np.array([p["stress"] for p in ds_parts])
    if with_stress:
        stress = stress + rng.normal(0.0, SIGMA_S, size=stress.shape)
    cell = np.array([p["cell"] for p in ds_parts])
    return Dataset(off, types, pos, force, cell, energy, stress, name)


def _voigt(S):
    return np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])


# ----------------------------------------------------------------------------
# analytic truth models (float64, NumPy)
# ----------------------------------------------------------------------------
def morse(r, De, a, re):
    e = np.exp(-a * (r - re))
    return De * (1.0 - e) * (1.0 - e) - De, 2.0 * De * (1.0 - e) * a * e


def expdecay(r, A, B):
    v = A * np.exp(-B * r)
    return v, -B * v


def sqrtf(r, A, B):
    v = A * np.sqrt(r / B)
    return v, 0.5 * A / np.sqrt(r * B)


def _gate(r, span, v, dv):
    m = (r >= span[0]) & (r <= span[1])
    return np.where(m, v, 0.0), np.where(m, dv, 0.0)


def eval_truth_eam(pos, types, cell, T, rc, adp=False):
    """EAM (and ADP) forces/energy/stress with analytic tables, same algebra as
    the reference's eval_forces (eam_force.cpp:45-102, adp_force.cpp:73-135)."""
    i, j, d = neighbor_pairs(pos, cell, rc)
    r = np.linalg.norm(d, axis=1)
    n = len(pos)
    ti, tj = types[i], types[j]
    slot = np.array([[pair_ordinal(a, b, T["ntypes"]) for b in range(T["ntypes"])]
                     for a in range(T["ntypes"])])[ti, tj]
    phi = np.zeros(len(r)); dphi = np.zeros(len(r))
    for s, (De, a, re) in enumerate(T["pair"]):
        m = slot == s
        v, dv = _gate(r[m], T["span"], *morse(r[m], De, a, re))
        phi[m], dphi[m] = v, dv
    g_j = np.zeros(len(r)); dg_j = np.zeros(len(r)); dg_i = np.zeros(len(r))
    for t, (A, B) in enumerate(T["density"]):
        m = tj == t
        v, dv = _gate(r[m], T["span"], *expdecay(r[m], A, B))
        g_j[m], dg_j[m] = v, dv
        m = ti == t
        _, dv = _gate(r[m], T["span"], *expdecay(r[m], A, B))
        dg_i[m] = dv
    rho = np.bincount(i, weights=g_j, minlength=n)
    F = np.zeros(n); dF = np.zeros(n)
    for t, (A, B) in enumerate(T["embedding"]):
        m = types == t
        F[m], dF[m] = sqrtf(rho[m], A, B)
    E = F.sum() + 0.5 * phi.sum()
    fsc = (dphi + dF[i] * dg_j + dF[j] * dg_i) / r
    f = fsc[:, None] * d
    if adp:
        u = np.zeros(len(r)); du = np.zeros(len(r)); w = np.zeros(len(r)); dw = np.zeros(len(r))
        for s, (A, B) in enumerate(T["dipole"]):
            m = slot == s
            u[m], du[m] = _gate(r[m], T["span"], *expdecay(r[m], A, B))
        for s, (A, B) in enumerate(T["quadrupole"]):
            m = slot == s
            w[m], dw[m] = _gate(r[m], T["span"], *expdecay(r[m], A, B))
        mu = np.zeros((n, 3)); lam = np.zeros((n, 3, 3))
        np.add.at(mu, i, u[:, None] * d)
        np.add.at(lam, i, w[:, None, None] * (d[:, :, None] * d[:, None, :]))
        tr = np.trace(lam, axis1=1, axis2=2)
        E += 0.5 * (mu * mu).sum() + 0.5 * ((lam * lam).sum(axis=(1, 2)) - tr * tr / 3.0).sum()
        dot = np.einsum("pa,pa->p", mu[i] - mu[j], d)
        f = f + (du / r * dot)[:, None] * d + u[:, None] * (mu[i] - mu[j])
        L = lam[i] + lam[j]
        trL = np.trace(L, axis1=1, axis2=2)
        Ld = np.einsum("pab,pb->pa", L, d)
        nu = np.einsum("pa,pa->p", d, Ld) - (r * r) / 3.0 * trL
        xi = Ld - (trL / 3.0)[:, None] * d
        f = f + (dw / r * nu)[:, None] * d + 2.0 * w[:, None] * xi
    forces = np.zeros((n, 3))
    np.add.at(forces, i, f)
    V = abs(np.linalg.det(cell))
    S = -0.5 * np.einsum("pa,pb->ab", d, f) / V
    return forces, E, S


def sw_v2(r, p):
    A, B, pp, q, delta, a1 = p[:6]
    v = np.zeros_like(r); dv = np.zeros_like(r)
    m = r < a1
    rm = r[m]
    dd = rm - a1
    with np.errstate(under="ignore"):
        e = np.exp(delta / dd)
    poly = A * rm ** (-pp) - B * rm ** (-q)
    dpoly = -A * pp * rm ** (-pp - 1.0) + B * q * rm ** (-q - 1.0)
    v[m] = poly * e
    dv[m] = dpoly * e + poly * e * (-delta / (dd * dd))
    return v, dv


def sw_h(r, p):
    gamma, a2 = p[6], p[7]
    h = np.zeros_like(r); dh = np.zeros_like(r)
    m = r < a2
    dd = r[m] - a2
    with np.errstate(under="ignore"):
        e = np.exp(gamma / dd)
    h[m] = e
    dh[m] = e * (-gamma / (dd * dd))
    return h, dh


def eval_truth_sw(pos, types, cell, params, lam, ntypes):
    """Stillinger-Weber pair + triplet sums (stiweb_force.cpp:241-271,
    kernels.hpp:40-68), vectorised over pairs and triplets."""
    rc = float(np.max(np.maximum(params[:, 5], params[:, 7])))
    i, j, d = neighbor_pairs(pos, cell, rc)
    n = len(pos)
    r = np.linalg.norm(d, axis=1)
    npair = ntypes * (ntypes + 1) // 2
    slot = np.array([[pair_ordinal(a, b, ntypes) for b in range(ntypes)] for a in range(ntypes)])
    sl = slot[types[i], types[j]]
    v2 = np.zeros(len(r)); dv2 = np.zeros(len(r)); h = np.zeros(len(r)); dh = np.zeros(len(r))
    for s in range(npair):
        m = sl == s
        v2[m], dv2[m] = sw_v2(r[m], params[s])
        h[m], dh[m] = sw_h(r[m], params[s])
    f = (dv2 / r)[:, None] * d
    forces = np.zeros((n, 3))
    np.add.at(forces, i, f)
    E = 0.5 * v2.sum()
    S = -0.5 * np.einsum("pa,pb->ab", d, f)
    # triplets (jj < kk within each centre's list)
    off = np.concatenate([[0], np.cumsum(np.bincount(i, minlength=n))])
    deg = np.diff(off)
    P1, P2 = [], []
    for m in np.unique(deg):
        if m < 2:
            continue
        a, b = np.triu_indices(m, 1)
        centres = np.nonzero(deg == m)[0]
        P1.append((off[centres][:, None] + a[None, :]).ravel())
        P2.append((off[centres][:, None] + b[None, :]).ravel())
    if P1:
        p1 = np.concatenate(P1); p2 = np.concatenate(P2)
        keep = ((h[p1] != 0) | (dh[p1] != 0)) & ((h[p2] != 0) | (dh[p2] != 0))
        p1, p2 = p1[keep], p2[keep]
        d1, d2 = d[p1], d[p2]
        ir1, ir2 = 1.0 / r[p1], 1.0 / r[p2]
        cs = np.einsum("pa,pa->p", d1, d2) * ir1 * ir2
        ci = i[p1]
        lidx = types[ci] * npair + slot[types[j[p1]], types[j[p2]]]
        lm = lam[lidx]
        cp = cs + 1.0 / 3.0
        g, dg = lm * cp * cp, 2.0 * lm * cp
        f1, df1, f2, df2 = h[p1], dh[p1], h[p2], dh[p2]
        E += (f1 * f2 * g).sum()
        ir12 = ir1 * ir2
        A_ = cs * ir1 * ir1 - ir12
        B_ = cs * ir2 * ir2 - ir12
        s1 = df1 * ir1 * f2 * g
        s2 = f1 * df2 * ir2 * g
        s3 = f1 * f2 * dg
        fi = s1[:, None] * d1 + s2[:, None] * d2 - s3[:, None] * (A_[:, None] * d1 + B_[:, None] * d2)
        fj = -s1[:, None] * d1 - s3[:, None] * (ir12[:, None] * d2 - (cs * ir1 * ir1)[:, None] * d1)
        fk = -s2[:, None] * d2 - s3[:, None] * (ir12[:, None] * d1 - (cs * ir2 * ir2)[:, None] * d2)
        np.add.at(forces, ci, fi)
        np.add.at(forces, j[p1], fj)
        np.add.at(forces, j[p2], fk)
        S = S + np.einsum("pa,pb->ab", d1, fj) + np.einsum("pa,pb->ab", d2, fk)
    V = abs(np.linalg.det(cell))
    return forces, E, S / V


# ----------------------------------------------------------------------------
# workloads
# ----------------------------------------------------------------------------
def _sample_table(fn, n, rmin, rmax):
    x = knot_grid(n, rmin, rmax)
    return fn(x)[0]


def _emb_span(rhos, margin=0.25):
    lo, hi = float(np.min(rhos)), float(np.max(rhos))
    w = hi - lo
    return max(lo - margin * w, 0.05 * lo), hi + margin * w


def _truth_rho(pos, types, cell, T, rc):
    i, j, d = neighbor_pairs(pos, cell, rc)
    r = np.linalg.norm(d, axis=1)
    g = np.zeros(len(r))
    for t, (A, B) in enumerate(T["density"]):
        m = types[j] == t
        g[m] = _gate(r[m], T["span"], *expdecay(r[m], A, B))[0]
    return np.bincount(i, weights=g, minlength=len(pos))


def config1(seed: int = 1, n_configs: int = 20) -> Workload:
    """Single-element tabulated EAM: 20 FCC 3x3x3 cubic supercells (108 atoms,
    a = 3.615 A), Gaussian displacements 0.1 A, isotropic strain +-2 %; truth =
    Morse pair + exponential density + sqrt embedding; start model phi, rho:
    15 knots on [1.8, 5.5], F: 12 knots over the observed density +- 25 %."""
    rng = np.random.default_rng(seed)
    T = dict(ntypes=1, span=(1.0, 5.5), pair=[(0.15, 1.5, 2.7)], density=[(4.0, 1.2)],
             embedding=[(-2.0, 1.0)])
    start = dict(pair=[(0.18, 1.42, 2.75)], density=[(3.6, 1.26)], embedding=[(-2.2, 1.0)])
    parts, rhos = [], []
    for _ in range(n_configs):
        eps = rng.uniform(-0.02, 0.02)
        a = 3.615 * (1.0 + eps)
        pos = _fcc(a, (3, 3, 3)) + rng.normal(0.0, 0.1, size=(108, 3))
        cell = np.eye(3) * 3 * a
        types = np.zeros(108, dtype=np.int32)
        f, E, S = eval_truth_eam(pos, types, cell, T, 5.5)
        parts.append(dict(pos=pos, types=types, cell=cell, force=f, energy=E, stress=_voigt(S)))
        rhos.append(_truth_rho(pos, types, cell, T, 5.5))
        rhos.append(_truth_rho(pos, types, cell, dict(T, density=start["density"]), 5.5))
    ds = _finish(rng, parts, "config1", with_stress=False)
    lo, hi = _emb_span(np.concatenate(rhos))
    tabs = [Table(15, 1.8, 5.5, _sample_table(lambda x: morse(x, *start["pair"][0]), 15, 1.8, 5.5)),
            Table(15, 1.8, 5.5, _sample_table(lambda x: expdecay(x, *start["density"][0]), 15, 1.8, 5.5)),
            Table(12, lo, hi, _sample_table(lambda x: sqrtf(x, *start["embedding"][0]), 12, lo, hi))]
    model = TabulatedModel("eam", 1, tabs, energy_weight=1.0, stress_weight=0.0)
    return Workload("config1", ds, model, truth=T, meta=dict(seed=seed, emb_span=(lo, hi)))


def config2(seed: int = 2, n_configs: int = 500, nmin: int = 32, nmax: int = 2048,
            stress_weight: float = 1.0) -> Workload:
    """Ternary tabulated EAM on random dense packings: N ~ round(exp(U[ln 32,
    ln 2048])) atoms per config at 0.06 atoms/A^3 in periodic cubes, species
    fractions 0.5/0.4/0.1, hard core 2.0 A; 6 pair + 3 density + 3 embedding
    tables of 20 knots (240 parameters)."""
    rng = np.random.default_rng(seed)
    T = dict(ntypes=3, span=(1.0, 5.5),
             pair=[(0.15, 1.5, 2.7), (0.18, 1.4, 2.85), (0.12, 1.6, 2.75),
                   (0.20, 1.3, 3.0), (0.14, 1.45, 2.9), (0.10, 1.7, 2.8)],
             density=[(4.0, 1.2), (5.0, 1.1), (3.0, 1.3)],
             embedding=[(-2.0, 1.0), (-2.5, 1.0), (-1.6, 1.0)])
    start = dict(pair=[(De * 1.2, a * 0.95, re * 1.02) for De, a, re in T["pair"]],
                 density=[(A * 0.9, B * 1.05) for A, B in T["density"]],
                 embedding=[(A * 1.1, B) for A, B in T["embedding"]])
    parts, rhos = [], [[] for _ in range(3)]
    for _ in range(n_configs):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        L = (n / 0.06) ** (1.0 / 3.0)
        pos = random_packing(rng, n, L, 2.0)
        types = rng.choice(3, size=n, p=[0.5, 0.4, 0.1]).astype(np.int32)
        cell = np.eye(3) * L
        f, E, S = eval_truth_eam(pos, types, cell, T, 5.5)
        parts.append(dict(pos=pos, types=types, cell=cell, force=f, energy=E, stress=_voigt(S)))
        for dens in (T["density"], start["density"]):
            rr = _truth_rho(pos, types, cell, dict(T, density=dens), 5.5)
            for t in range(3):
                rhos[t].append(rr[types == t])
    ds = _finish(rng, parts, "config2", with_stress=True)
    tabs = []
    for s in range(6):
        tabs.append(Table(20, 1.8, 5.5, _sample_table(lambda x: morse(x, *start["pair"][s]), 20, 1.8, 5.5)))
    for t in range(3):
        tabs.append(Table(20, 1.8, 5.5, _sample_table(lambda x: expdecay(x, *start["density"][t]), 20, 1.8, 5.5)))
    spans = []
    for t in range(3):
        lo, hi = _emb_span(np.concatenate(rhos[t]))
        spans.append((lo, hi))
        tabs.append(Table(20, lo, hi, _sample_table(lambda x: sqrtf(x, *start["embedding"][t]), 20, lo, hi)))
    model = TabulatedModel("eam", 3, tabs, energy_weight=1.0, stress_weight=stress_weight)
    return Workload("config2", ds, model, truth=T, meta=dict(seed=seed, emb_spans=spans))


def config3(seed: int = 3, n_configs: int = 200) -> Workload:
    """Binary tabulated ADP: B2 (n^3 x 2 atoms, n = 5..8) and random-alloy FCC
    (n^3 x 4 atoms, n = 4..6) supercells, 250-1024 atoms, displacements 0.15 A,
    random axial strain +-3 %. 13 tables of 20 knots; the free set is the six
    dipole/quadrupole tables (120 knots) plus every other knot of the two
    embedding tables (20), 140 parameters; pair and density tables are held at
    the truth samples. Stress weight 0.1."""
    rng = np.random.default_rng(seed)
    T = dict(ntypes=2, span=(1.0, 5.5),
             pair=[(0.16, 1.5, 2.65), (0.22, 1.45, 2.6), (0.12, 1.55, 2.75)],
             density=[(4.0, 1.2), (3.2, 1.1)],
             embedding=[(-2.0, 1.0), (-1.7, 1.0)],
             dipole=[(0.4, 1.6), (-0.5, 1.5), (0.3, 1.7)],
             quadrupole=[(0.15, 1.4), (0.2, 1.3), (-0.1, 1.5)])
    start = dict(dipole=[(0.6 * A, 1.1 * B) for A, B in T["dipole"]],
                 quadrupole=[(0.6 * A, 1.1 * B) for A, B in T["quadrupole"]],
                 embedding=[(1.08 * A, B) for A, B in T["embedding"]])
    parts, rhos = [], [[] for _ in range(2)]
    for _ in range(n_configs):
        if rng.uniform() < 0.5:
            m = int(rng.integers(5, 9))
            a = 2.9
            pos = _b2(a, (m, m, m))
            types = np.tile(np.array([0, 1], dtype=np.int32), m ** 3)
        else:
            m = int(rng.integers(4, 7))
            a = 3.6
            pos = _fcc(a, (m, m, m))
            types = rng.integers(0, 2, size=len(pos)).astype(np.int32)
        strain = 1.0 + rng.uniform(-0.03, 0.03, size=3)
        pos = pos * strain[None, :] + rng.normal(0.0, 0.15, size=pos.shape)
        cell = np.diag(m * a * strain)
        f, E, S = eval_truth_eam(pos, types, cell, T, 5.5, adp=True)
        parts.append(dict(pos=pos, types=types, cell=cell, force=f, energy=E, stress=_voigt(S)))
        rr = _truth_rho(pos, types, cell, T, 5.5)
        for t in range(2):
            rhos[t].append(rr[types == t])
    ds = _finish(rng, parts, "config3", with_stress=True)
    tabs = []
    for s in range(3):
        tabs.append(Table(20, 1.8, 5.5, _sample_table(lambda x: morse(x, *T["pair"][s]), 20, 1.8, 5.5),
                          np.ones(20, dtype=bool)))
    for t in range(2):
        tabs.append(Table(20, 1.8, 5.5, _sample_table(lambda x: expdecay(x, *T["density"][t]), 20, 1.8, 5.5),
                          np.ones(20, dtype=bool)))
    for t in range(2):
        lo, hi = _emb_span(np.concatenate(rhos[t]))
        fx = np.zeros(20, dtype=bool)
        fx[1::2] = True
        tabs.append(Table(20, lo, hi, _sample_table(lambda x: sqrtf(x, *start["embedding"][t]), 20, lo, hi), fx))
    for s in range(3):
        tabs.append(Table(20, 1.8, 5.5, _sample_table(lambda x: expdecay(x, *start["dipole"][s]), 20, 1.8, 5.5)))
    for s in range(3):
        tabs.append(Table(20, 1.8, 5.5,
                          _sample_table(lambda x: expdecay(x, *start["quadrupole"][s]), 20, 1.8, 5.5)))
    model = TabulatedModel("adp", 2, tabs, energy_weight=1.0, stress_weight=0.1)
    return Workload("config3", ds, model, truth=T, meta=dict(seed=seed))


# Stillinger-Weber Si (1985) in the reference's parameterisation
# (tests/unit/test_stiweb_force.cpp:14-38)
SI_EPS, SI_SIGMA, SI_A, SI_B, SI_P, SI_Q = 2.1683, 2.0951, 7.049556277, 0.6022245584, 4.0, 0.0
SI_ACUT, SI_GAMMA, SI_LAMBDA = 1.80, 1.20, 21.0


def si_sw_params():
    return np.array([SI_A * SI_EPS * SI_B * SI_SIGMA ** SI_P, SI_A * SI_EPS * SI_SIGMA ** SI_Q, SI_P, SI_Q,
                     SI_SIGMA, SI_ACUT * SI_SIGMA, SI_GAMMA * SI_SIGMA, SI_ACUT * SI_SIGMA])


def config4(seed: int = 4, n_configs: int = 256) -> Workload:
    """Stillinger-Weber Si: diamond-cubic supercells of 64-512 atoms with
    displacements 0.2 A, plus 10 % liquid-like random packings at 2.57 g/cm^3;
    free {A, B, p, q, a1, gamma, lambda}, fixed {delta, a2} (7 parameters)."""
    rng = np.random.default_rng(seed)
    p_true = si_sw_params()
    lam_true = np.array([SI_LAMBDA * SI_EPS])
    shapes = [(2, 2, 2), (2, 2, 3), (2, 3, 3), (3, 3, 3), (3, 3, 4), (3, 4, 4), (4, 4, 4)]
    a = 5.431
    n_liq = max(1, int(round(0.1 * n_configs)))
    parts = []
    for c in range(n_configs):
        if c < n_configs - n_liq:
            m = shapes[int(rng.integers(0, len(shapes)))]
            pos = _diamond(a, m) + rng.normal(0.0, 0.2, size=(8 * m[0] * m[1] * m[2], 3))
            cell = np.diag(np.array(m, dtype=float) * a)
        else:
            n = int(rng.integers(64, 513))
            rho_n = 2.57 / 28.0855 * 0.6022140857  # atoms per A^3
            L = (n / rho_n) ** (1.0 / 3.0)
            pos = random_packing(rng, n, L, 2.1)
            cell = np.eye(3) * L
        types = np.zeros(len(pos), dtype=np.int32)
        f, E, S = eval_truth_sw(pos, types, cell, p_true[None, :], lam_true, 1)
        parts.append(dict(pos=pos, types=types, cell=cell, force=f, energy=E, stress=_voigt(S)))
    ds = _finish(rng, parts, "config4", with_stress=True)
    p0 = p_true.copy()
    p0[0] *= 1.05
    p0[1] *= 0.95
    p0[2] = 4.1
    p0[3] = 0.05
    p0[5] *= 1.01
    p0[6] *= 0.96
    fixed = np.zeros((1, 8), dtype=bool)
    fixed[0, 4] = True  # delta
    fixed[0, 7] = True  # a2
    model = SWModel(1, p0[None, :], fixed, lam_true * 1.1, np.zeros(1, dtype=bool),
                    energy_weight=1.0, stress_weight=0.0)
    return Workload("config4", ds, model, truth=dict(params=p_true, lam=lam_true), meta=dict(seed=seed))


def de_box(x: np.ndarray):
    """BoostDESolver box for unbounded parameters: v +- max(2|v|, 5e-4)
    (src/optimization/solver.cpp:118-125)."""
    half = np.maximum(2.0 * np.abs(x), 5e-4)
    return x - half, x + half
