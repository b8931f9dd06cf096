"""The C ABI library: it loads, exports every function include/*.h declares,
and its host-side helpers and argument checks behave (no device needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        # C declarations of the ABI: "<type> sleek_gpu_xxx(" at line start
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(sleek_gpu_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from sleek_amd import _lib

    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_error_bound_matches_oracle():
    from sleek_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(3)
    for eb in list(10.0 ** rng.uniform(-40, 30, 200)) + [1e-45, 2.0 ** 103, 0.01]:
        e, t, o = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint()
        rc = lib.sleek_gpu_error_bound_f32(float(np.float32(eb)), ctypes.byref(e), ctypes.byref(t), ctypes.byref(o))
        try:
            ref = po.error_bound(eb, np.float32)
        except ValueError:
            assert rc == _lib.EEB
            continue
        assert rc == 0 and (e.value, t.value, o.value) == ref
    for eb in list(10.0 ** rng.uniform(-300, 300, 200)) + [5e-324, 2.0 ** 900]:
        e, t, o = ctypes.c_int(), ctypes.c_int(), ctypes.c_ulonglong()
        rc = lib.sleek_gpu_error_bound_f64(float(eb), ctypes.byref(e), ctypes.byref(t), ctypes.byref(o))
        try:
            ref = po.error_bound(eb, np.float64)
        except ValueError:
            assert rc == _lib.EEB
            continue
        assert rc == 0 and (e.value, t.value, o.value) == ref
    e, t, o = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint()
    assert lib.sleek_gpu_error_bound_f32(2.0 ** 104, ctypes.byref(e), ctypes.byref(t), ctypes.byref(o)) == _lib.EEB


def test_sizes_and_messages():
    from sleek_amd import _lib

    lib = _lib.load()
    for n in (0, 1, 16384, 16385, 123456789):
        assert lib.sleek_gpu_upper_bound_size(n) == po.upper_bound_size(n)
    assert lib.sleek_gpu_workspace_size(0) > 0
    assert lib.sleek_gpu_workspace_size(16384 * 10) >= (10 + 9) * 8
    assert lib.sleek_gpu_system_checks() == 0
    for st in (0, -1, -2, -3, -4, -5, -6, -7):
        assert lib.sleek_gpu_strerror(st)
    assert lib.sleek_gpu_set_tuning(-1, 0, 0) == _lib.EINVAL
    assert lib.sleek_gpu_set_tuning(3, 0, 0) == _lib.EINVAL
    assert lib.sleek_gpu_set_tuning(0, -1, 0) == _lib.EINVAL
    assert lib.sleek_gpu_set_tuning(0, 0, 0) == 0


def test_argument_checks_before_any_device_work():
    from sleek_amd import _lib

    lib = _lib.load()
    # bad element size, odd byte count, null outputs, misaligned pointers
    assert lib.sleek_gpu_compress_async(2, 0, 0, 0, 0, None, 0, 8, 8, 8, 1024, 0, None) == _lib.EINVAL
    assert lib.sleek_gpu_compress_async(4, 0, 0, 0, 0, 16, 6, 8, 8, 8, 1024, 0, None) == _lib.EINVAL
    assert lib.sleek_gpu_compress_async(4, 0, 0, 0, 0, 16, 16, None, 8, 8, 1024, 0, None) == _lib.EINVAL
    assert lib.sleek_gpu_compress_async(4, 0, 0, 0, 0, 20, 16, 8, 8, 8, 1024, 0, None) == _lib.EALIGN
    assert lib.sleek_gpu_compress_async(4, 0, 0, 0, 0, 16, 16, 9, 8, 8, 1024, 0, None) == _lib.EALIGN
    # workspace too small for the chunk count
    assert lib.sleek_gpu_compress_async(4, 0, 0, 0, 0, 16, 16384 * 100, 8, 8, 8, 64, 0, None) == _lib.EWORKSPACE
    assert lib.sleek_gpu_decompress_async(4, 0, 0, 0, 8, 4, 16, 0, 8, 8, 1024, 0, None) == _lib.ECORRUPT
    assert lib.sleek_gpu_decompress_async(4, 0, 0, 0, 8, -1, 20, 0, 8, 8, 1024, 0, None) == _lib.EALIGN
    assert lib.sleek_gpu_decompress_async(4, 0, 0, 0, 8, -1, 16, 16384 * 100, 8, 8, 64, 0, None) == _lib.EWORKSPACE
    assert lib.sleek_gpu_workspace_status(None, None) == _lib.EINVAL
