"""pd_ipc_amd - MI355X-native Projective-Dynamics cloth substep engine.

The product is the HIP/C++ library ``libpdx.so`` (C ABI declared in
``include/pdx.h``); this package holds its sources (``csrc/``), the build step
and a thin ctypes host mirror (``pdx.py``) plus the benchmark scene builders
(``scenes.py``).
"""
from .pdx import Engine, Cloth, SimConfig, PdxError, load_library  # noqa: F401
from . import scenes  # noqa: F401
