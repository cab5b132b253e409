"""Registers the package directory feicar3d_3.0_febio_amd as feicar3d_amd and
puts oracle/ on the import path."""
import importlib.util
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "feicar3d_3.0_febio_amd")


def load_package():
    if "feicar3d_amd" in sys.modules:
        return sys.modules["feicar3d_amd"]
    spec = importlib.util.spec_from_file_location(
        "feicar3d_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["feicar3d_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def oracle_dir():
    d = os.path.join(REPO, "oracle")
    if d not in sys.path:
        sys.path.insert(0, d)
    return d
