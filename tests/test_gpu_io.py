"""Reference-format files through the HIP path: a dataset written in the
configuration JSON of src/io/config_reader.cpp:58-142 and a model in the
model JSON of src/io/force_model_reader.cpp / potential_reader.cpp:126-191,
read back by forcesmith_amd.io and evaluated on the device against the
oracle built from the same files.

Bars: residual max |dF| / (1 + |F|) <= 1e-5 (mixed) and 1e-11 (FP64);
Jacobian max |J - J_oracle| / (1 + |J_oracle|) <= 5e-5.
"""
import numpy as np
import pytest

import forcesmith_amd as fa
import oracle
from forcesmith_amd import datasets as dsm
from forcesmith_amd import io as fio

pytestmark = pytest.mark.gpu


def _files(tmp_path, w, species):
    cpath, mpath = tmp_path / "configs.json", tmp_path / "model.json"
    fio.write_configs(w.dataset, species, cpath)
    m = w.model.copy()
    for t in getattr(m, "tables", []):  # the file format fixes a whole table or none of it
        t.fixed[:] = t.fixed.all()
    fio.write_model(m, mpath)
    ds, sp = fio.read_configs(str(cpath))
    assert sp == species
    return ds, fio.read_model(str(mpath), w.model.energy_weight, w.model.stress_weight)


@pytest.mark.parametrize("name", ["eam-ternary", "adp-binary", "sw-si"])
def test_file_loaded_workload_matches_oracle(tmp_path, name):
    if name == "eam-ternary":
        w, species = dsm.config2(n_configs=6, nmin=32, nmax=400), ["Al", "Cu", "Zr"]
    elif name == "adp-binary":
        w, species = dsm.config3(n_configs=2), ["Al", "Ni"]
    else:
        w, species = dsm.config4(n_configs=6), ["Si"]
    ds, model = _files(tmp_path, w, species)
    o = oracle.Oracle(ds, model)
    o.prepare()
    x = o.gather()
    Fo = o.residual(x)
    Jo = o.jacobian(x)
    for prec, tol in (("mixed", 1e-5), ("fp64", 1e-11)):
        g = fa.GpuCalculator(ds, model, precision=prec)
        g.prepare()
        Fg = g.residual(x)
        d = float(np.max(np.abs(Fg - Fo) / (1 + np.abs(Fo))))
        print(f"[{name} from files/{prec}] residual {d:.3e}", flush=True)
        assert d <= tol
    Jg = fa.GpuCalculator(ds, model).jacobian(x).astype(np.float64)
    m = float(np.max(np.abs(Jg - Jo) / (1 + np.abs(Jo))))
    print(f"[{name} from files] J metric {m:.3e}", flush=True)
    assert m <= 5e-5
