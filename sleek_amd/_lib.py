"""ctypes binding of libsleek_amd.so, the HIP codec's C ABI (include/sleek_gpu.h).

The library is built in-tree by ``make -C sleek_amd/csrc`` (or
``__graft_entry__.build()``).  Importing this module without the library
raises ImportError: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsleek_amd.so")

OK = 0
EINVAL = -1
EHIP = -2
ECORRUPT = -3
EEB = -4
ECAPACITY = -5
EWORKSPACE = -6
EALIGN = -7
ETOOLARGE = -8

c_ll = ctypes.c_longlong
c_p = ctypes.c_void_p
c_i = ctypes.c_int
c_u = ctypes.c_uint
c_ull = ctypes.c_ulonglong

# (name, restype, argtypes) of every symbol include/sleek_gpu.h declares
SIGNATURES = [
    ("sleek_gpu_upper_bound_size", c_ll, [c_ll]),
    ("sleek_gpu_system_checks", c_i, []),
    ("sleek_gpu_thread_blocks", c_i, []),
    ("sleek_gpu_error_bound_f32", c_i, [ctypes.c_float, ctypes.POINTER(c_i), ctypes.POINTER(c_i), ctypes.POINTER(c_u)]),
    ("sleek_gpu_error_bound_f64", c_i, [ctypes.c_double, ctypes.POINTER(c_i), ctypes.POINTER(c_i), ctypes.POINTER(c_ull)]),
    ("sleek_gpu_strerror", ctypes.c_char_p, [c_i]),
    ("sleek_gpu_last_error", ctypes.c_char_p, []),
    ("sleek_gpu_compress_f32", c_i, [c_i, c_p, c_ll, c_p, c_p]),
    ("sleek_gpu_decompress_f32", c_i, [c_i, c_p, c_p, c_p]),
    ("sleek_gpu_compress_f64", c_i, [c_i, c_p, c_ll, c_p, c_p]),
    ("sleek_gpu_decompress_f64", c_i, [c_i, c_p, c_p, c_p]),
    ("sleek_gpu_compress_lossy_f32", c_i, [c_i, c_i, c_i, c_u, c_p, c_ll, c_p, c_p]),
    ("sleek_gpu_decompress_lossy_f32", c_i, [c_i, c_i, c_i, c_u, c_p, c_p, c_p]),
    ("sleek_gpu_compress_lossy_f64", c_i, [c_i, c_i, c_i, c_ull, c_p, c_ll, c_p, c_p]),
    ("sleek_gpu_decompress_lossy_f64", c_i, [c_i, c_i, c_i, c_ull, c_p, c_p, c_p]),
    ("sleek_gpu_workspace_size", c_ll, [c_ll]),
    ("sleek_gpu_compress_async", c_i, [c_i, c_i, c_i, c_i, c_ull, c_p, c_ll, c_p, c_p, c_p, c_ll, c_i, c_p]),
    ("sleek_gpu_decompress_async", c_i, [c_i, c_i, c_i, c_ull, c_p, c_ll, c_p, c_ll, c_p, c_p, c_ll, c_i, c_p]),
    ("sleek_gpu_set_tuning", c_i, [c_i, c_i, c_i]),
    ("sleek_gpu_set_lookback", c_i, [c_i, c_i]),
    ("sleek_gpu_set_decoder", c_i, [c_i]),
    ("sleek_gpu_set_help_after", c_i, [c_i]),
    ("sleek_gpu_help_events", c_ll, []),
    ("sleek_gpu_set_kernel_events", c_i, [c_p, c_p]),
    ("sleek_gpu_workspace_status", c_i, [c_p, c_p]),
    ("sleek_gpu_compress_noalloc", c_i, [c_i, c_i, c_i, c_i, c_ull, c_p, c_ll, c_p, c_p, c_i, c_p]),
    ("sleek_gpu_batch_plan", c_i, [c_p, c_i, c_i, ctypes.POINTER(c_ll), ctypes.POINTER(c_ll)]),
    ("sleek_gpu_batch_compress_async", c_i, [c_p, c_i, c_ll, c_p, c_ll, c_i, c_p]),
    ("sleek_gpu_batch_decompress_async", c_i, [c_p, c_i, c_ll, c_p, c_ll, c_i, c_p]),
    ("sleek_gpu_batch_status", c_i, [c_p, c_i, c_p, c_ll, c_p, c_p]),
    ("sleek_gpu_batch_graph_create", c_i, [c_i, c_p, c_i, c_ll, c_p, c_ll, c_i, ctypes.POINTER(c_p)]),
    ("sleek_gpu_graph_launch", c_i, [c_p, c_p]),
    ("sleek_gpu_graph_destroy", c_i, [c_p]),
]

_LIB = None


def load():
    """Load the codec library (ImportError if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C sleek_amd/csrc` "
                              "(the codec has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


class SleekError(RuntimeError):
    def __init__(self, status, what=""):
        lib = load()
        msg = lib.sleek_gpu_strerror(status).decode()
        if status == EHIP:
            msg += ": " + lib.sleek_gpu_last_error().decode()
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")
        self.status = status


def check(status, what=""):
    if status != OK:
        raise SleekError(status, what)
    return status
