"""ctypes binding of libdtb.so (declarations in include/dtb_abi.h).

The library is built in-tree by ``__graft_entry__.build()`` into
``digital_twin_brain-open_amd/lib/libdtb.so``.  There is no fallback: if the
library is missing or a call fails, an exception is raised.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdtb.so")

# include/dtb_abi.h status codes (0-5 are protos/snn.proto:44-51 SnnStatus).
STATUS_NAMES = {
    0: "SNN_OK",
    1: "SNN_INVALID_PARAMETER",
    2: "SNN_UNSET_SAMPLE_FAILURE",
    3: "SNN_UNSUPPORTED",
    4: "SNN_NOT_ENABLE_METRIC",
    5: "SNN_INVALID_NETWORK",
    100: "HIP_ERROR",
    101: "OUT_OF_MEMORY",
    102: "NO_DEVICE",
    103: "INTERNAL",
}

RUN_USE_OU = 1 << 0
RUN_FREQ = 1 << 1
RUN_SAMPLE_SPIKE = 1 << 2
RUN_SAMPLE_VMEMB = 1 << 3
RUN_SAMPLE_ISYN = 1 << 4
RUN_SAMPLE_IOU = 1 << 5
RUN_VMEAN = 1 << 6
RUN_IMEAN = 1 << 7
RUN_GRAPH = 1 << 8
RUN_RECEPTOR_IMEAN = 1 << 9

OK = 0
INVALID_PARAMETER = 1
UNSET_SAMPLE_FAILURE = 2
UNSUPPORTED = 3
INVALID_NETWORK = 5
HIP_ERROR = 100
OUT_OF_MEMORY = 101
NO_DEVICE = 102


class DtbError(RuntimeError):
    """A libdtb call returned a non-zero status."""

    def __init__(self, status, message):
        self.status = status
        super().__init__("%s (%d): %s" % (STATUS_NAMES.get(status, "ERROR"), status, message))


class Network(ctypes.Structure):
    _fields_ = [
        ("n", c_int64),
        ("props", c_void_p),
        ("nsyn", c_int64),
        ("rowptr", c_void_p),
        ("colinds", c_void_p),
        ("weights", c_void_p),
        ("connkinds", c_void_p),
        ("weight_type", c_int32),
    ]


class RunOutputs(ctypes.Structure):
    _fields_ = [
        ("freq", c_void_p),
        ("sample_spike", c_void_p),
        ("sample_vmemb", c_void_p),
        ("sample_isynaptic", c_void_p),
        ("sample_iou", c_void_p),
        ("vmean", c_void_p),
        ("imean", c_void_p),
        ("ampa_imean", c_void_p),
        ("nmda_imean", c_void_p),
        ("gabaa_imean", c_void_p),
        ("gabab_imean", c_void_p),
    ]


# int (*)(void *user, int32_t first_iter, int32_t rows, const dtb_run_outputs *out)
RUN_CHUNK_FN = ctypes.CFUNCTYPE(c_int, c_void_p, c_int32, c_int32, POINTER(RunOutputs))

_SIGS = {
    "dtb_version": (c_int, []),
    "dtb_last_error": (c_char_p, []),
    "dtb_device_count": (c_int, [POINTER(c_int)]),
    "dtb_device_memory_cached": (c_int, [c_int, POINTER(c_uint64)]),
    "dtb_device_memory_trim": (c_int, []),
    "dtb_device_memory_reserve": (c_int, [c_int, c_uint64]),
    "dtb_engine_create": (c_int, [c_int, POINTER(c_void_p)]),
    "dtb_engine_destroy": (c_int, [c_void_p]),
    "dtb_instance_add": (c_int, [c_void_p, POINTER(Network), c_float, c_uint64, POINTER(c_int)]),
    "dtb_instance_remove": (c_int, [c_void_p, c_int]),
    "dtb_instance_info": (c_int, [c_void_p, c_int, POINTER(c_int64), POINTER(c_int64), POINTER(c_int32),
                                  c_void_p, c_void_p]),
    "dtb_set_ou": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int32]),
    "dtb_set_samples": (c_int, [c_void_p, c_int, c_void_p, c_int32]),
    "dtb_update_props": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int32]),
    "dtb_update_prop_cols": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int32, c_int32]),
    "dtb_run": (c_int, [c_void_p, c_int, c_int32, c_int32, c_uint32, POINTER(RunOutputs)]),
    "dtb_run_noalloc": (c_int, [c_void_p, c_int, c_int32, c_int32, c_uint32, POINTER(RunOutputs)]),
    "dtb_run_stream": (c_int, [c_void_p, c_int, c_int32, c_int32, c_uint32, c_int32, RUN_CHUNK_FN, c_void_p]),
    "dtb_engine_run": (c_int, [c_void_p, c_int64, c_int32]),
    "dtb_configure": (c_int, [c_void_p, c_int, c_int32, c_int32]),
    "dtb_engine_sync": (c_int, [c_void_p]),
    "dtb_gen_gamma": (c_int, [c_uint64, c_uint32, c_uint32, c_uint32, c_float, c_float, c_void_p]),
    "dtb_gamma_prop_cols": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int32]),
    "dtb_read_props": (c_int, [c_void_p, c_int, c_void_p]),
    "dtb_engine_queue_capacity": (c_int, [c_void_p, c_uint32, c_uint32]),
    "dtb_engine_schedule": (c_int, [c_void_p, c_int32, c_int32]),
    "dtb_engine_pull": (c_int, [c_void_p, c_int32, c_uint32]),
    "dtb_engine_split": (c_int, [c_void_p, c_int32, c_uint32]),
    "dtb_engine_get_split": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_uint32)]),
    "dtb_engine_get_pull": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_uint32)]),
    "dtb_instance_pull_info": (c_int, [c_void_p, c_int, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                       POINTER(c_uint64)]),
    "dtb_engine_get_schedule": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    "dtb_engine_autotune": (c_int, [c_void_p, c_int64, POINTER(c_int32), POINTER(c_int32), c_void_p]),
    "dtb_instance_propagation": (c_int, [c_void_p, c_int, POINTER(c_int32), POINTER(c_int32), POINTER(c_float)]),
    "dtb_instance_tile_synapses": (c_int, [c_void_p, c_int, c_void_p]),
    "dtb_engine_spilled": (c_int, [c_void_p, POINTER(c_uint64)]),
    "dtb_engine_propagation_stats": (c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
    "dtb_engine_run_span": (c_int, [c_void_p, c_int64, c_int32, POINTER(c_float)]),
    "dtb_engine_run_span_steps": (c_int, [c_void_p, c_int64, c_int32, POINTER(c_float), c_void_p]),
    "dtb_engine_time_kernels": (c_int, [c_void_p, c_int64, POINTER(c_float)]),
    "dtb_engine_check": (c_int, [c_void_p]),
    "dtb_read_state": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p]),
    "dtb_write_state": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dtb_set_step": (c_int, [c_void_p, c_int, c_int64]),
    "dtb_counters": (c_int, [c_void_p, c_int, POINTER(c_int64), POINTER(c_uint64), POINTER(c_uint64)]),
    "dtb_ou_normals": (c_int, [c_int, c_uint64, c_uint32, c_uint64, c_uint32, c_void_p]),
    "dtb_gen_uniform": (c_int, [c_int64, c_uint32, c_uint64, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p]),
    "dtb_gen_voxel": (c_int, [c_int64, c_uint32, c_uint32, c_float, c_float, c_uint64, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dtb_gen_fanout": (c_int, [c_int64, c_uint32, c_uint64, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p]),
    "dtb_instance_add_fanout": (c_int, [c_void_p, c_int64, c_uint32, c_uint64, c_void_p, c_uint32, c_float, c_uint64,
                                        POINTER(c_int)]),
    "dtb_npz_size": (c_int, [c_char_p, POINTER(c_int64), POINTER(c_int64), POINTER(c_int32)]),
    "dtb_npz_read": (c_int, [c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dtb_npz_weight_type": (c_int, [c_char_p, POINTER(c_int32)]),
    "dtb_npz_read_typed": (c_int, [c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
}

EXPORTED = tuple(sorted(_SIGS))

_lib = None


def load():
    """Loads libdtb.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libdtb.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != 0:
        msg = load().dtb_last_error()
        raise DtbError(status, msg.decode() if msg else "")
    return status
