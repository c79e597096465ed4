"""MI355X-native Wilson-Dslash solver (HIP, gfx950).

The compiled pieces live beside this file:
  libwilson.so    HIP kernels + the C-ABI of include/wilson.h and include/wilson_gen.h
  libipcc_api.so  the reference's C++ operator API (lattice_fermion, DslashNew, CGinvert, ...)
  wilson_main     command-line driver with the reference's argv (src/main.cpp:31-45)

`wilson_hip` is the Python (ctypes) binding of the C-ABI.  The directory name is
not a Python identifier, so load it by path, e.g.
    spec = importlib.util.spec_from_file_location("ipcc_amd", ".../ipcc2021-final_amd/__init__.py")
"""
import importlib.util as _ilu
import os as _os

_here = _os.path.dirname(_os.path.abspath(__file__))
_spec = _ilu.spec_from_file_location("ipcc_amd_wilson_hip", _os.path.join(_here, "wilson_hip.py"))
wilson_hip = _ilu.module_from_spec(_spec)
_spec.loader.exec_module(wilson_hip)

__all__ = ["wilson_hip"]
