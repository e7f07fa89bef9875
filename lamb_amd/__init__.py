"""MI355X-native sector block-sparse multiply C = A*B with norm filtering.

The compute path is the HIP library ``liblamb_hip.so`` (C-ABI in
``include/lamb_hip.h``), reached through :mod:`lamb_amd.device`. There is no
CPU fallback.
"""
from . import systems  # noqa: F401
from .matrix import Distribution, LambMatrix, Sector  # noqa: F401

__all__ = ["systems", "Distribution", "LambMatrix", "Sector"]
