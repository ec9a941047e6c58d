"""The reference's import names for the host helpers.

Code written against the reference (e.g. its experiment driver, phantom_
experiment.py:12-16) imports ``helpers.cl_helpers``, ``helpers.misc`` and
their siblings; each name here is the module of ``miccai2016_amd.helpers``
itself (the same module object, registered under both names), not a copy.
"""
import importlib as _importlib
import sys as _sys

_MODULES = ("cl_helpers", "helix_phantom", "misc", "parameters", "status")

for _name in _MODULES:
    _mod = _importlib.import_module(f"miccai2016_amd.helpers.{_name}")
    _sys.modules[f"{__name__}.{_name}"] = _mod
    globals()[_name] = _mod
