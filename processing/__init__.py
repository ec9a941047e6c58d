"""The reference's import names for the drop-in processing modules.

Code written against the reference (e.g. its experiment driver, phantom_
experiment.py:17-20) imports ``processing.qfrangi_gpu`` and its siblings;
each name here is the module of ``miccai2016_amd.processing`` itself (the same
module object, registered under both names), not a copy.
"""
import importlib as _importlib
import sys as _sys

_MODULES = ("adjust_directions", "continuous_cut_anisotropic_gpu", "gradient_vector_flow",
            "qfrangi_gpu", "ridge_detection", "slicing")

for _name in _MODULES:
    _mod = _importlib.import_module(f"miccai2016_amd.processing.{_name}")
    _sys.modules[f"{__name__}.{_name}"] = _mod
    globals()[_name] = _mod
