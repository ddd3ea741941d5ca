"""Loads the C++ host extension (_keldy_host), which links libkeldy_amd.so.

There is no pure-Python fallback: a missing build raises ImportError here.
"""
import importlib
import os

from . import _build

if not os.path.exists(_build.host_ext_path()):
    raise ImportError(f"{_build.host_ext_path()} is not built: run __graft_entry__.build()")
ext = importlib.import_module("keldy_amd._keldy_host")
KeldyError = ext.KeldyError
