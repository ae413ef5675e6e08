"""Shared fixtures.  The package directory `mcphylo.jl_amd/` is loaded under
the module name `mcphylo_amd`; the CPU oracle under `oracle`."""

import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _load_package():
    if "mcphylo_amd" in sys.modules:
        return sys.modules["mcphylo_amd"]
    pkg = os.path.join(ROOT, "mcphylo.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "mcphylo_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mcphylo_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402
    return oracle


mcp = _load_package()
oracle = _load_oracle()
import mlp_oracle  # noqa: E402  (oracle/ is on sys.path after _load_oracle)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


_GPU = None


def gpu_available() -> bool:
    """HIP device present (asks the HIP runtime, not torch)."""
    global _GPU
    if _GPU is None:
        _GPU = mcp._lib.device_count() > 0
    return _GPU


def pytest_collection_modifyitems(config, items):
    for item in items:
        if "gpu" in item.keywords and not gpu_available():
            item.add_marker(pytest.mark.skip(reason="no HIP device"))


@pytest.fixture(scope="session")
def pkg():
    return mcp


@pytest.fixture(scope="session")
def orc():
    return oracle


# Reference fixtures: test/likelihood/felsenstein.jl:8-14,24-26 (MIT licence,
# MCPhylo.jl).
PRIMATES_NEWICK = """(((Tarsius_syrichta:0.0510942,(Lemur_catta:0.0136013,Homo_sapiens:0.0370755)12:0.0343822)
        13:0.224569,(Pan:0.0712342,Gorilla:0.03754)14:0.0295151)15:0.0768634,((Pongo:
        0.020513,Hylobates:0.159117)16:0.239429,Macaca_fuscata:0.454752)
        17:0.0902988,((M_mulatta:0.0644278,M_fascicularis:0.318016)
        18:0.015879,(M_sylvanus:0.100663,Saimiri_sciureus:0.0112774)19:0.2727)20:0.0448203);"""
SIMU_NEWICK = """(((0:0.110833,1:0.0137979)10:0.146124,(2:0.197891,(3:0.132967,(4:0.0378759,5:0.089252)11:0.101833)12:0.184301)
         13:0.0450774)14:0.335725,6:0.153197,(7:0.0216218,(8:0.0781687,9:0.120419)15:0.0209114)16:0.0209771);"""

# test/likelihood/felsenstein.jl:20,35,37
PRIMATES_LL = -8677.360274116634
SIMU_LL = -738.7363911756175
SIMU_GRAD = np.array([
    -56.25542148325748, -38.05203887880975, 48.05792385792187, -13.52161136915132,
    -6.157297983096069, -19.400861758279206, -13.271503408059996, 23.78407661010017,
    -2.7933830668575474, -10.342020138166133, -6.786353636249929, -32.05726418827399,
    -3.742827673853779, 5.710607498000256, -2.174093469729043, 106.0017224580945,
    57.19513048979947])


def load_fixture(name: str):
    """(tree, x) for 'primates' or 'simudata' exactly as the reference test
    builds them (ParseNexus + ParseNewick + datafortree)."""
    nwk = PRIMATES_NEWICK if name == "primates" else SIMU_NEWICK
    t = mcp.ParseNewick(nwk)
    ntax, nchar, gap, miss, sym, df, langs = mcp.ParseNexus(os.path.join(GOLDEN, f"{name}.nex"))
    x = mcp.datafortree(df, langs, t, sym, gap, miss)
    return t, x


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))
