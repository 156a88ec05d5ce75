"""Imports the svo-tracer_amd directory as the module ``svo_tracer_amd``.

The directory name is not a Python identifier, so callers (tests, bench.py,
__graft_entry__.py) load it by path through this helper.
"""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent


def load():
    mod = sys.modules.get("svo_tracer_amd")
    if mod is not None:
        return mod
    spec = importlib.util.spec_from_file_location(
        "svo_tracer_amd", PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)]
    )
    mod = importlib.util.module_from_spec(spec)
    sys.modules["svo_tracer_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
