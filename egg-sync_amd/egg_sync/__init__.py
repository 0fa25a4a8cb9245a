"""MI355X-native EGG-SynC: exact grid-based clustering by synchronization on gfx950.

Import with ``egg-sync_amd`` on ``sys.path``.  The compute path is the HIP
library libeggsync.so (C ABI in include/egg_sync.h); ``algorithms`` mirrors the
reference's Python API on top of the PyTorch extension.
"""
from . import data  # noqa: F401

__all__ = ["data", "algorithms", "solver"]
