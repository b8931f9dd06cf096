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
    assert lib.sleek_gpu_set_tuning(8, 0, 0) == _lib.EINVAL
    assert lib.sleek_gpu_set_tuning(3, 0, 0) == _lib.EINVAL   # retired variant
    assert lib.sleek_gpu_set_help_after(-1) == _lib.EINVAL
    assert lib.sleek_gpu_set_help_after(10_000_001) == _lib.EINVAL
    assert lib.sleek_gpu_set_help_after(100) == 0
    assert lib.sleek_gpu_set_tuning(0, -1, 0) == _lib.EINVAL
    assert lib.sleek_gpu_set_tuning(0, 0, 0) == 0
    for g, w in ((12, 4), (4, 4), (128, 4), (64, 65), (-1, 0), (8, -1)):
        assert lib.sleek_gpu_set_lookback(g, w) == _lib.EINVAL, (g, w)
    for g, w in ((8, 1), (16, 64), (32, 7), (0, 0)):
        assert lib.sleek_gpu_set_lookback(g, w) == 0, (g, w)
    assert lib.sleek_gpu_set_decoder(2) == _lib.EINVAL and lib.sleek_gpu_set_decoder(-1) == _lib.EINVAL
    assert lib.sleek_gpu_set_decoder(1) == 0 and lib.sleek_gpu_set_decoder(0) == 0


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


def _items(specs, decompress):
    """Host batch items with fake aligned device pointers (the plan never dereferences them)."""
    from sleek_amd.batch import ITEM

    items = np.zeros(len(specs), dtype=ITEM)
    for i, (nbytes, eb_, lossy) in enumerate(specs):
        base = 0x10000000 + i * 0x1000000
        items[i]["src"], items[i]["dst"], items[i]["d_size"] = base, base + 0x800000, 0x2000 + 8 * i
        items[i]["elem_bytes"], items[i]["lossy"] = eb_, lossy
        if decompress:
            items[i]["src_bytes"], items[i]["dst_bytes"] = -1, nbytes
        else:
            items[i]["src_bytes"] = nbytes
    return items


def test_batch_item_layout_matches_header():
    from sleek_amd.batch import ITEM

    hdr = open(os.path.join(ROOT, "include", "sleek_gpu.h")).read()
    assert "} sleek_gpu_batch_item;     /* 96 bytes */" in hdr
    assert ITEM.itemsize == 96
    assert [ITEM.fields[f][1] for f in ("src", "offs", "eb_e", "elem_bytes", "lossy", "tile0", "ws_words")] == \
        [0, 40, 48, 56, 60, 64, 88]


@pytest.mark.parametrize("decompress", [0, 1])
def test_batch_plan_layout(decompress):
    from sleek_amd.batch import plan

    specs = [(0, 4, 0), (4, 4, 1), (16384, 8, 0), (16388, 4, 0), (400000, 8, 1), (16384 * 65, 4, 1)]
    items = _items(specs, decompress)
    tiles, wsb = plan(items, decompress)
    even = lambda w: (w + 1) & ~1
    t = 0
    word = 16 + even(len(specs)) + even((sum(-(-b // 16384) for b, _, _ in specs) + 1) // 2)
    for it, (b, _, _) in zip(items, specs):
        ch = -(-b // 16384)
        assert it["tile0"] == t and it["tiles"] == ch
        assert it["ws_words"] == even(16 + ch + 2 * (-(-ch // 8)))   # groups as small as 8 chunks
        assert it["ws_word"] == word and it["ws_word"] % 2 == 0
        t += ch
        word += it["ws_words"]
    assert tiles == t and wsb == (word + 1024) * 8   # then 8 KiB of tile-ticket counters (64, 128 B apart)


def test_batch_plan_rejects_bad_items():
    from sleek_amd import _lib
    from sleek_amd.batch import plan

    def rc_of(items, d):
        try:
            plan(items, d)
            return 0
        except _lib.SleekError as e:
            return e.status

    it = _items([(100, 4, 0)], 0)
    it[0]["src_bytes"] = 102            # not a multiple of the element size
    assert rc_of(it, 0) == _lib.EINVAL
    it = _items([(100, 6, 0)], 0)       # element size
    assert rc_of(it, 0) == _lib.EINVAL
    it = _items([(100, 4, 0)], 0)
    it[0]["src"] += 8                   # arrays need 16-byte alignment
    assert rc_of(it, 0) == _lib.EALIGN
    it = _items([(100, 4, 0)], 1)
    it[0]["dst_bytes"] = -1             # a capacity is required to decompress
    assert rc_of(it, 1) == _lib.EINVAL
    it = _items([(100, 4, 0)], 1)
    it[0]["src_bytes"] = 7              # shorter than the header
    assert rc_of(it, 1) == _lib.EINVAL
    it = _items([(0, 4, 0)], 0)
    it[0]["src"] = 0                    # an empty array may have no data pointer
    assert rc_of(it, 0) == 0
