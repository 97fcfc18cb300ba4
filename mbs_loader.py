"""Imports the package directory ``mbs-fast_amd`` as module ``mbs_fast_amd``."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "mbs-fast_amd")


def load():
    if "mbs_fast_amd" in sys.modules:
        return sys.modules["mbs_fast_amd"]
    spec = importlib.util.spec_from_file_location(
        "mbs_fast_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mbs_fast_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
