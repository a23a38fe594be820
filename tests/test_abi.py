"""CPU-side checks of the C ABI and host logic (no GPU calls)."""
import ctypes as C
import importlib
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
cep = importlib.import_module("2ddrugsimulation_amd")
workload = importlib.import_module("2ddrugsimulation_amd.workload")


def _declared_symbols():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(cep_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    path = cep.LIB_PATH
    assert os.path.exists(path), "libcepgpu.so not built"
    lib = C.CDLL(path)          # loading needs the HIP runtime library, not a GPU
    declared = _declared_symbols()
    assert len(declared) >= 19
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(cep.EXPORTED)


def test_binding_refuses_missing_library(tmp_path):
    try:
        cep.library(str(tmp_path / "nope.so"))
    except cep.CepError as e:
        assert "no CPU fallback" in str(e)
    else:
        raise AssertionError("expected CepError")


def test_population_workload_shapes_and_ranges():
    pop = workload.ord_population(20251204, 50, 4, 50)
    n = 10000
    assert pop["herg"].shape == (n, 19) and pop["gscale"].shape == (n, 7)
    assert set(np.unique(pop["bcl"])) <= {500.0, 1000.0, 2000.0}
    ic50 = pop["herg"][:, 0:14:2]
    assert (ic50 >= 100).all() and (ic50 <= 1e5).all()          # 0.1-100 uM in nM
    hill = pop["herg"][:, 1:14:2]
    assert (hill >= 0.5).all() and (hill <= 2).all()
    # 4 concentrations at 1-4 x Cmax per drug
    c = pop["conc"].reshape(50, 4, 50)
    np.testing.assert_allclose(c[:, 1, 0] / c[:, 0, 0], 2.0)
    # lognormal(0, 0.2) scales
    assert abs(np.log(pop["gscale"]).std() - 0.2) < 0.02
    # same seed, same population
    pop2 = workload.ord_population(20251204, 50, 4, 50)
    np.testing.assert_array_equal(pop["herg"], pop2["herg"])


def test_single_drug_config():
    wl = workload.ord_single_drug(seed=20251203)
    assert wl["conc"][0] == 0.0
    np.testing.assert_allclose(wl["conc"][1:] / wl["cmax"], [1, 2, 3, 4])


def test_qinward_closed_form():
    assert cep.qinward(2.0, 3.0, 4.0, 6.0) == 0.5
