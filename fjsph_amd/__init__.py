"""MI355X-native WCSPH step engine (FJSPH hot path) -- Python host side.

The compute path is the HIP library fjsph_amd/libfjsph_engine.so, reached
through its C-ABI (include/fjsph_engine.h); see fjsph_amd.engine.
"""
from . import abi, cases  # noqa: F401

__all__ = ["abi", "cases"]
