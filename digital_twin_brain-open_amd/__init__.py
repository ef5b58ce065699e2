"""MI355X-native DTB single-block step: a C++/HIP library (lib/libdtb.so,
built from csrc/) behind a C ABI (include/dtb_abi.h), with a Python mirror of
the reference block client.

Import with ``importlib.import_module("digital_twin_brain-open_amd")`` after
putting the repository root on ``sys.path``.
"""
from ._lib import DtbError, load, EXPORTED, LIB_PATH  # noqa: F401
from . import engine  # noqa: F401
from .engine import (Block, Engine, NetworkArrays, gen_uniform, gen_voxel, gen_fanout, gen_gamma, device_count,  # noqa: F401
                     ou_normals, load_npz, save_npz, device_memory_cached, device_memory_reserve, device_memory_trim,
                     COLUMN_TO_PROP,
                     DEFAULT_G_UI, NCOL)
from .blockwrapper import BlockWrapper  # noqa: F401

__all__ = ["Block", "Engine", "NetworkArrays", "gen_uniform", "gen_voxel", "device_count", "ou_normals",
           "load_npz", "save_npz", "BlockWrapper", "gen_fanout", "gen_gamma", "device_memory_cached",
           "device_memory_reserve", "device_memory_trim",
           "DtbError", "load", "EXPORTED", "LIB_PATH", "COLUMN_TO_PROP", "DEFAULT_G_UI", "NCOL"]
