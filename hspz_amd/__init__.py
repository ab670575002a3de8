"""MI355X-native seed-and-filter path (spaced-seed lookup, X-drop score gate,
HSP materialization, scans and compaction) behind the reference's Engine API.

The compute lives in the native library libhspz.so (HIP kernels for gfx950 and
the C ABI in include/hspz.h); this package is its host-side binding.
"""
from . import sequence  # noqa: F401

__all__ = ["sequence", "engine", "synth"]
