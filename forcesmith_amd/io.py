"""The reference's JSON dataset and model files.

Configurations (src/io/config_reader.cpp:58-206): a top-level array of
objects with lattice vectors "X", "Y", "Z" (the columns of the box,
:58-66), energy "E" (required, :68-72), an optional weight "W" (read, not
applied to residuals), an optional stress "S" = [xx, yy, zz, xy, yz, zx]
(:74-88) and "atoms", each with "element", "position" and an optional
"force" (:90-110). Species are registered in atomic-number order, so the type
slot of an element is its rank by Z among the elements present
(include/forcesmith/core/species.hpp:163-181).

Models (src/io/force_model_reader.cpp, src/io/potential_reader.cpp): an
object with "model" ("eam", "adp" or "stiweb") and "ntypes". Tabulated
families carry one section per table kind in declaration order ("pair",
"density", "embedding", then "dipole", "quadrupole" for ADP;
include/forcesmith/force/eam_force.hpp:24-32, adp_force.hpp:37-51), each
{"format": "tabulated", "potentials": [{"rmin", "rmax", "knots", "fixed"}]}
with ntypes (ntypes + 1) / 2 pair-indexed or ntypes type-indexed entries
(force_model_reader.cpp:41-56); the knot grid is rmin + k (rmax - rmin) /
(n - 1) (potential_reader.cpp:162-191). Stillinger-Weber carries
"potentials" (one parameter object per pair slot: A, B, p, q, delta, a1,
gamma, a2) and "lambda" (ntypes x pair slots); each parameter is a number or
{"value", "min", "max", "fixed"} (include/forcesmith/core/json.hpp:65-74,
force_model_reader.cpp:181-205).

Only tabulated EAM/ADP tables and Stillinger-Weber are evaluated by the GPU
path; an analytic table section is rejected with a message.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .model import Dataset, SWModel, Table, TabulatedModel

ELEMENTS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se "
    "Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy "
    "Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf "
    "Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
ATOMIC_NUMBER = {s: z + 1 for z, s in enumerate(ELEMENTS)}

SW_FIELDS = ("A", "B", "p", "q", "delta", "a1", "gamma", "a2")


def _load(src):
    """A JSON value from a path, a JSON string or an already parsed value."""
    if isinstance(src, (dict, list)):
        return src
    if isinstance(src, (str, os.PathLike)) and os.path.exists(src):
        with open(src) as f:
            return json.load(f)
    return json.loads(src)


def _vec3(v, what):
    a = np.asarray(v, dtype=np.float64)
    if a.shape != (3,):
        raise ValueError(f"{what} must be an array of 3 numbers")
    return a


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def species_order(symbols):
    """Registered species: unique symbols sorted by atomic number."""
    uniq = []
    for s in symbols:
        if s not in ATOMIC_NUMBER:
            raise ValueError(f"unknown element symbol '{s}'")
        if s not in uniq:
            uniq.append(s)
    return sorted(uniq, key=lambda s: ATOMIC_NUMBER[s])


def read_configs(src, name: str = ""):
    """Dataset and the species list (slot order) from a configuration file or
    JSON text."""
    j = _load(src)
    if not isinstance(j, list):
        raise ValueError("top-level JSON must be an array of configurations")
    symbols = []
    for obj in j:
        if isinstance(obj, dict) and isinstance(obj.get("atoms"), list):
            for a in obj["atoms"]:
                if isinstance(a, dict) and isinstance(a.get("element"), str):
                    symbols.append(a["element"])
    species = species_order(symbols)
    slot = {s: i for i, s in enumerate(species)}
    off, types, pos, frc, cells, energy, stress = [0], [], [], [], [], [], []
    for ci, obj in enumerate(j):
        ctx = f"config[{ci}]"
        if not isinstance(obj, dict):
            raise ValueError(f"{ctx}: each configuration must be a JSON object")
        for k in ("X", "Y", "Z", "E", "atoms"):
            if k not in obj:
                raise ValueError(f"{ctx}: missing required key '{k}'")
        cells.append(np.concatenate([_vec3(obj[k], f"{ctx} '{k}'") for k in ("X", "Y", "Z")]))
        energy.append(float(obj["E"]))
        if "S" in obj:
            s = np.asarray(obj["S"], dtype=np.float64)
            if s.shape != (6,):
                raise ValueError(f"{ctx} 'S': expected 6 numbers")
            # file order xx, yy, zz, xy, yz, zx -> residual order xx, yy, zz, xy, xz, yz
            stress.append([s[0], s[1], s[2], s[3], s[5], s[4]])
        else:
            stress.append([0.0] * 6)
        if not isinstance(obj["atoms"], list):
            raise ValueError(f"{ctx}: 'atoms' must be an array")
        for ai, a in enumerate(obj["atoms"]):
            actx = f"{ctx} atom[{ai}]"
            if "element" not in a or "position" not in a:
                raise ValueError(f"{actx}: missing 'element' or 'position'")
            types.append(slot[a["element"]])
            pos.append(_vec3(a["position"], f"{actx} 'position'"))
            frc.append(_vec3(a["force"], f"{actx} 'force'") if "force" in a else np.zeros(3))
        off.append(off[-1] + len(obj["atoms"]))
    ds = Dataset(np.asarray(off, dtype=np.int64), np.asarray(types, dtype=np.int32),
                 np.asarray(pos, dtype=np.float64).reshape(-1, 3),
                 np.asarray(frc, dtype=np.float64).reshape(-1, 3),
                 np.asarray(cells, dtype=np.float64).reshape(-1, 3, 3), np.asarray(energy, dtype=np.float64),
                 np.asarray(stress, dtype=np.float64).reshape(-1, 6), name)
    return ds, species


def config_names(src) -> list:
    """The optional "name" of every configuration ("" when absent), as
    config_reader.cpp:119 reads it."""
    j = _load(src)
    if not isinstance(j, list):
        raise ValueError("top-level JSON must be an array of configurations")
    return [str(o.get("name", "")) if isinstance(o, dict) else "" for o in j]


def write_configs(ds: Dataset, species, path=None) -> str:
    """The inverse of read_configs (weights written as 1)."""
    out = []
    for c in range(ds.n_configs):
        a0, a1 = int(ds.atom_off[c]), int(ds.atom_off[c + 1])
        cell = np.asarray(ds.cell[c], dtype=np.float64).reshape(3, 3)
        s = ds.ref_stress[c]
        out.append({
            "X": cell[0].tolist(), "Y": cell[1].tolist(), "Z": cell[2].tolist(),
            "E": float(ds.ref_energy[c]), "W": 1.0,
            "S": [float(s[0]), float(s[1]), float(s[2]), float(s[3]), float(s[5]), float(s[4])],
            "atoms": [{"element": species[int(ds.types[i])], "position": ds.pos[i].tolist(),
                       "force": ds.ref_force[i].tolist()} for i in range(a0, a1)],
        })
    text = json.dumps(out)
    if path is not None:
        with open(path, "w") as f:
            f.write(text)
    return text


# ---------------------------------------------------------------------------
# models
# ---------------------------------------------------------------------------
def _param(v, what):
    """(value, fixed, lo, hi) of a parameter given as a number or an object."""
    if isinstance(v, dict):
        if "value" not in v:
            raise ValueError(f"{what}: parameter object missing 'value'")
        lo, hi = float(v.get("min", -np.inf)), float(v.get("max", np.inf))
        if lo > hi:
            raise ValueError(f"{what}: min > max")
        return float(v["value"]), bool(v.get("fixed", False)), lo, hi
    return float(v), False, -np.inf, np.inf


def _tables(section, want, label):
    if not isinstance(section, dict) or "format" not in section:
        raise ValueError(f"{label}: missing 'format' key")
    if section["format"] != "tabulated":
        raise ValueError(f"{label}: format '{section['format']}' is not evaluated by this library "
                         "(tabulate the table first)")
    pots = section.get("potentials")
    if not isinstance(pots, list):
        raise ValueError(f"{label}: missing or invalid 'potentials' array")
    if len(pots) != want:
        raise ValueError(f"{label}: expected {want} potentials, got {len(pots)}")
    out = []
    for k, p in enumerate(pots):
        knots = p.get("knots")
        if not isinstance(knots, list) or len(knots) < 2:
            raise ValueError(f"{label}[{k}]: 'knots' must be an array of >= 2 values")
        rmin, rmax = float(p["rmin"]), float(p["rmax"])
        if not rmax > rmin:
            raise ValueError(f"{label}[{k}]: tabulated potential requires rmax > rmin")
        n = len(knots)
        fixed = np.full(n, bool(p.get("fixed", False)))
        out.append(Table(n, rmin, rmax, np.asarray(knots, dtype=np.float64), fixed))
    return out


def read_model(src, energy_weight: float = 1.0, stress_weight: float = 0.0):
    """TabulatedModel (EAM / ADP) or SWModel from a model file or JSON text."""
    j = _load(src)
    if not isinstance(j, dict) or "model" not in j:
        raise ValueError("model file must be an object with a 'model' key")
    kind = j["model"]
    T = int(j.get("ntypes", 1))
    P = T * (T + 1) // 2
    if kind in ("eam", "adp"):
        sections = [("pair", P), ("density", T), ("embedding", T)]
        if kind == "adp":
            sections += [("dipole", P), ("quadrupole", P)]
        tables = []
        for key, want in sections:
            if key not in j:
                raise ValueError(f"{kind}: missing '{key}' section")
            tables += _tables(j[key], want, f"{kind}.{key}")
        return TabulatedModel(kind, T, tables, energy_weight, stress_weight)
    if kind == "stiweb":
        pots, lam = j.get("potentials"), j.get("lambda")
        if not isinstance(pots, list) or len(pots) != P:
            raise ValueError(f"stiweb: expected {P} entries for ntypes={T}")
        if not isinstance(lam, list) or len(lam) != T * P:
            raise ValueError(f"stiweb: expected {T * P} lambda entries for ntypes={T}")
        pp, pf = np.zeros((P, 8)), np.zeros((P, 8), dtype=bool)
        for s, obj in enumerate(pots):
            for f, key in enumerate(SW_FIELDS):
                if key not in obj:
                    raise ValueError(f"stiweb.potentials[{s}]: missing '{key}'")
                pp[s, f], pf[s, f], _, _ = _param(obj[key], f"stiweb.potentials[{s}].{key}")
        lv, lf = np.zeros(T * P), np.zeros(T * P, dtype=bool)
        for k, v in enumerate(lam):
            lv[k], lf[k], _, _ = _param(v, f"stiweb.lambda[{k}]")
        return SWModel(T, pp, pf, lv, lf, energy_weight, stress_weight)
    raise ValueError(f"model '{kind}' is not evaluated by this library (eam, adp, stiweb)")


def write_model(model, path=None, nknots=None) -> str:
    """The model file of a TabulatedModel or SWModel (current parameters).
    nknots: resample every table to that many knots over its span, as the
    reference's writer does (io::write_native, src/io/output_writer.cpp:23-43,
    500 knots by default there); the tables are then written free."""
    if isinstance(model, TabulatedModel) and nknots is not None:
        from .spline import resample
        if int(nknots) < 2:
            raise ValueError("nknots must be >= 2")
        tabs = []
        for t in model.tables:
            lo, hi, vals = resample(t.x, t.y, int(nknots))
            tabs.append(Table(int(nknots), lo, hi, vals))
        model = TabulatedModel(model.kind, model.ntypes, tabs, model.energy_weight, model.stress_weight)
    if isinstance(model, TabulatedModel):
        P, T = model.npair, model.ntypes
        keys = [("pair", P), ("density", T), ("embedding", T)]
        if model.kind == "adp":
            keys += [("dipole", P), ("quadrupole", P)]
        j = {"model": model.kind, "ntypes": T}
        o = 0
        for key, n in keys:
            pots = []
            for t in model.tables[o:o + n]:
                e = {"rmin": float(t.rmin), "rmax": float(t.rmax), "knots": [float(v) for v in t.y]}
                if bool(np.all(t.fixed)):
                    e["fixed"] = True
                pots.append(e)
            j[key] = {"format": "tabulated", "potentials": pots}
            o += n
    else:
        j = {"model": "stiweb", "ntypes": model.ntypes,
             "potentials": [{k: ({"value": float(model.pair_params[s, f]), "fixed": True}
                                 if model.pair_fixed[s, f] else float(model.pair_params[s, f]))
                             for f, k in enumerate(SW_FIELDS)} for s in range(model.npair)],
             "lambda": [({"value": float(v), "fixed": True} if fx else float(v))
                        for v, fx in zip(model.lam, model.lam_fixed)]}
    text = json.dumps(j)
    if path is not None:
        with open(path, "w") as f:
            f.write(text)
    return text
