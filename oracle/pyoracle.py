"""ctypes binding of the CPU oracle (test infrastructure only).

``libsleek_oracle.so`` is the plain-C restatement of the reference CPU codec
(include/sleek_common_cpu.h, include/sleek_lossy_cpu.h).  ``ref()`` returns a
binding of the reference codec itself, compiled from the reference headers by
oracle/Makefile into oracle/_ref/ where the reference tree is present (it is
absent where only the built files are shipped), or None.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

CS = 16384
E_CORRUPT = -3
E_EB = -4
E_CAP = -5

c_ll = ctypes.c_longlong
c_p = ctypes.c_void_p


def _load():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsleek_oracle.so")
        if not os.path.exists(path):
            raise ImportError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        for w, ct in (("u32", ctypes.c_uint32), ("u64", ctypes.c_uint64)):
            getattr(lib, f"sleek_oracle_encode_{w}").argtypes = [c_p, c_ll, c_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ct]
            getattr(lib, f"sleek_oracle_encode_{w}").restype = c_ll
            getattr(lib, f"sleek_oracle_decode_{w}").argtypes = [c_p, c_ll, c_p, c_ll, ctypes.POINTER(c_ll), ctypes.c_int, ctypes.c_int, ct]
            getattr(lib, f"sleek_oracle_decode_{w}").restype = ctypes.c_int
            for fn in ("fwd", "inv"):
                getattr(lib, f"sleek_oracle_{fn}_{w}").argtypes = [ct]
                getattr(lib, f"sleek_oracle_{fn}_{w}").restype = ct
            getattr(lib, f"sleek_oracle_quantize_{w}").argtypes = [ct, ctypes.c_int, ctypes.c_int, ct]
            getattr(lib, f"sleek_oracle_quantize_{w}").restype = ct
            getattr(lib, f"sleek_oracle_dequantize_{w}").argtypes = [ct, ctypes.c_int, ct]
            getattr(lib, f"sleek_oracle_dequantize_{w}").restype = ct
        lib.sleek_oracle_error_bound_f32.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint32)]
        lib.sleek_oracle_error_bound_f64.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]
        lib.sleek_oracle_upper_bound_size.argtypes = [c_ll]
        lib.sleek_oracle_upper_bound_size.restype = c_ll
        lib.sleek_oracle_check_bijection_u32.restype = c_ll
        _LIB = lib
    return _LIB


def lib():
    return _load()


def _word(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "u32"
    if dtype == np.float64:
        return "u64"
    raise TypeError(f"unsupported dtype {dtype}")


def upper_bound_size(insize):
    return _load().sleek_oracle_upper_bound_size(insize)


def error_bound(eb, dtype):
    """(eb_e, thr_e, offs_e) for an absolute bound; ValueError when too large
    (the reference exits there, sleek_lossy_cpu.h:68)."""
    L = _load()
    e, t = ctypes.c_int(), ctypes.c_int()
    if _word(dtype) == "u32":
        o = ctypes.c_uint32()
        rc = L.sleek_oracle_error_bound_f32(float(np.float32(eb)), ctypes.byref(e), ctypes.byref(t), ctypes.byref(o))
    else:
        o = ctypes.c_uint64()
        rc = L.sleek_oracle_error_bound_f64(float(eb), ctypes.byref(e), ctypes.byref(t), ctypes.byref(o))
    if rc != 0:
        raise ValueError("error bound is too large")
    return e.value, t.value, o.value


def compress(arr, eb=None):
    """Encode a float32/float64 numpy array; returns the stream as uint8."""
    arr = np.ascontiguousarray(arr)
    w = _word(arr.dtype)
    lossy, ebi = 0, (0, 0, 0)
    if eb is not None:
        lossy, ebi = 1, error_bound(eb, arr.dtype)
    out = np.zeros(max(upper_bound_size(arr.nbytes), 8), dtype=np.uint8)
    n = getattr(_load(), f"sleek_oracle_encode_{w}")(arr.ctypes.data, arr.nbytes, out.ctypes.data, lossy, *ebi)
    if n < 0:
        raise ValueError(f"oracle encode failed ({n})")
    return out[:n].copy()


def decompress(stream, dtype, eb=None, capacity=None):
    """Decode a stream; returns (array, status).  status 0 = ok."""
    stream = np.ascontiguousarray(stream, dtype=np.uint8)
    w = _word(dtype)
    lossy, thr, offs = 0, 0, 0
    if eb is not None:
        lossy = 1
        _, thr, offs = error_bound(eb, dtype)
    hdr = int(stream[:8].view(np.int64)[0]) if stream.size >= 8 else 0
    cap = hdr if capacity is None else capacity
    cap = max(cap, 0)
    out = np.zeros(max(cap, 8), dtype=np.uint8)
    osz = c_ll()
    rc = getattr(_load(), f"sleek_oracle_decode_{w}")(stream.ctypes.data, stream.size, out.ctypes.data, cap, ctypes.byref(osz), lossy, thr, offs)
    nbytes = min(max(osz.value, 0), cap)
    return out[:nbytes].view(np.dtype(dtype)).copy(), rc


def ref():
    """Binding of the compiled reference CPU codec, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libsleek_ref.so")
        if not os.path.exists(path):
            return None
        r = ctypes.CDLL(path)
        for d, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            getattr(r, f"sleek_ref_compress_{d}").argtypes = [c_p, c_ll, c_p]
            getattr(r, f"sleek_ref_compress_{d}").restype = c_ll
            getattr(r, f"sleek_ref_decompress_{d}").argtypes = [c_p, c_p]
            getattr(r, f"sleek_ref_decompress_{d}").restype = c_ll
            getattr(r, f"sleek_ref_compress_lossy_{d}").argtypes = [ct, c_p, c_ll, c_p]
            getattr(r, f"sleek_ref_compress_lossy_{d}").restype = c_ll
            getattr(r, f"sleek_ref_decompress_lossy_{d}").argtypes = [ct, c_p, c_p]
            getattr(r, f"sleek_ref_decompress_lossy_{d}").restype = c_ll
        _REF = r
    return _REF


def ref_compress(arr, eb=None):
    r = ref()
    arr = np.ascontiguousarray(arr)
    d = "f32" if arr.dtype == np.float32 else "f64"
    out = np.zeros(max(upper_bound_size(arr.nbytes), 8), dtype=np.uint8)
    if eb is None:
        n = getattr(r, f"sleek_ref_compress_{d}")(arr.ctypes.data, arr.size, out.ctypes.data)
    else:
        n = getattr(r, f"sleek_ref_compress_lossy_{d}")(eb, arr.ctypes.data, arr.size, out.ctypes.data)
    return out[:n].copy()


def ref_decompress(stream, dtype, eb=None):
    r = ref()
    stream = np.ascontiguousarray(stream, dtype=np.uint8)
    d = "f32" if np.dtype(dtype) == np.float32 else "f64"
    hdr = int(stream[:8].view(np.int64)[0])
    out = np.zeros(max(hdr, 8), dtype=np.uint8)
    if eb is None:
        n = getattr(r, f"sleek_ref_decompress_{d}")(stream.ctypes.data, out.ctypes.data)
    else:
        n = getattr(r, f"sleek_ref_decompress_lossy_{d}")(eb, stream.ctypes.data, out.ctypes.data)
    return out[:n].view(np.dtype(dtype)).copy()
