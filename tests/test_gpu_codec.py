"""Parity of the HIP codec with the oracle and the reference's golden streams.

Every stream the GPU writes and every array it decodes must be byte-identical
to the reference CPU codec's (golden fixtures) and to the oracle's on the same
input; lossy output must also satisfy the reference's point-wise check
(MAXABS_f32.h:40-95).  All GPU calls go through the C ABI of libsleek_amd.so.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from cases import build_input, eb_value
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))["cases"]


def sha(b):
    return hashlib.sha256(np.ascontiguousarray(b).view(np.uint8).tobytes()).hexdigest()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_roundtrip(a, eb, blocks=0, sync=False):
    """Compress + decompress a numpy array on the GPU; returns (stream, decoded) as numpy."""
    import torch
    from sleek_amd import codec

    x = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    if sync:
        s = codec.compress_sync(x, eb, blocks=blocks)
        torch.cuda.synchronize()
        stream = s.cpu().numpy()
        d = codec.decompress_sync(s, x.dtype, eb, blocks=blocks)
    else:
        out, n = codec.compress(x, eb, blocks=blocks)
        ws = codec.Workspace(a.nbytes)
        torch.cuda.synchronize()
        stream = out[: int(n.item())].cpu().numpy()
        d, dn = codec.decompress(out, x.dtype, eb, numel=a.size, blocks=blocks, workspace=ws)
        assert ws.status() == 0
        assert int(dn.item()) == a.nbytes
    torch.cuda.synchronize()
    return stream, d.cpu().numpy()


def maxabs_ok(orig, recon, eb):
    """The reference's point-wise check (MAXABS_f32.h:56-81), in the data's precision:
    non-finite values must match (NaN matches NaN); finite ones need
    |o - r| <= eb and o - eb <= r <= o + eb."""
    o, r = orig, recon
    fin = np.isfinite(o) & np.isfinite(r)
    same = (o == r) | (np.isnan(o) & np.isnan(r))
    if np.any(~fin & ~same):
        return False
    eb = o.dtype.type(eb)
    of, rf = o[fin], r[fin]
    with np.errstate(over="ignore", invalid="ignore"):
        bad = (rf < of - eb) | (rf > of + eb) | (np.abs(of - rf) > eb)
    return not np.any(bad)


@pytest.mark.parametrize("case", GOLDEN, ids=[c["name"] for c in GOLDEN])
def test_golden_streams(gpu, case):
    say(f"golden {case['name']} n={case['n']}")
    a = build_input(case)
    assert sha(a) == case["input_sha256"], "input generator drifted"
    eb = eb_value(case, a)
    ebf = None if eb is None else float(eb)
    stream, dec = gpu_roundtrip(a, ebf)
    assert stream.size == case["stream_len"]
    if case["stream_hex"] is not None:
        assert stream.tobytes().hex() == case["stream_hex"]
    assert sha(stream) == case["stream_sha256"]
    assert sha(dec) == case["decoded_sha256"]
    if eb is None:
        assert dec.tobytes() == a.tobytes()
    elif float(eb) > 0:
        assert maxabs_ok(a, dec, eb)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("blocks", [1, 7, 0])
def test_blocks_invariance(gpu, dt, blocks):
    say(f"blocks={blocks} {np.dtype(dt).name}")
    from sleek_amd.datagen import exact_field

    a = exact_field(300_001 if dt == np.float32 else 150_001, 42, dt)
    for eb in (None, 1e-3):
        ref = po.compress(a, eb)
        for sync in (False, True):
            s, d = gpu_roundtrip(a, eb, blocks=blocks, sync=sync)
            assert s.tobytes() == ref.tobytes()
            od, rc = po.decompress(ref, dt, eb)
            assert rc == 0 and d.tobytes() == od.tobytes()


def help_events():
    from sleek_amd import _lib

    n = _lib.load().sleek_gpu_help_events()
    assert n >= 0, n
    return n


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("variant", [0, 1, 2, 4, 5, 6, 7])
def test_encoder_variant_invariance(gpu, dt, variant):
    """Every encoder kernel (sleek_gpu_set_tuning's encode_variant) writes the oracle's bytes,
    at the automatic grid and at grids of 1, 3, 5, 8 and 37 workgroups that make each one
    walk many chunks (the ticket encoders then run 1, 2, 4, 8 or 32 ticket sets, never more
    than the workgroups; at the automatic grid, 64 sets of one or two tickets each).  The ticket encoders (0 = 6, and 7) hand every chunk to a running
    workgroup, so their look-backs never need to size a missing chunk themselves: no help
    event."""
    say(f"encoder variant={variant} {np.dtype(dt).name}")
    from sleek_amd import _lib
    from sleek_amd.datagen import exact_field, inject_specials

    lib = _lib.load()
    a = exact_field(400_003 if dt == np.float32 else 200_003, 5, dt)
    inject_specials(a, 1e-3, 11)
    try:
        for grid in (0, 1, 3, 5, 8, 37):
            _lib.check(lib.sleek_gpu_set_tuning(variant, grid, 0), "tuning")
            h0 = help_events()
            for eb in (None, 1e-2, 1e-5):
                ref = po.compress(a, eb)
                s, d = gpu_roundtrip(a, eb)
                assert s.tobytes() == ref.tobytes(), (grid, eb)
                od, rc = po.decompress(ref, dt, eb)
                assert rc == 0 and d.tobytes() == od.tobytes()
            if variant in (0, 6, 7):
                assert help_events() == h0, f"ticket encoder needed look-back help at grid {grid}"
    finally:
        lib.sleek_gpu_set_tuning(0, 0, 0)


@pytest.mark.parametrize("variant", [0, 1, 4, 5, 7])
def test_lookback_help_always_on(gpu, variant):
    """With the help threshold at 0 every look-back that does not settle at its first
    poll sizes the missing chunks and groups from the input itself (wave_chunk_size,
    wave_group_size) and publishes them next to their owners: the bytes stay the
    oracle's, and the help path is counted."""
    say(f"look-back help always on, encoder variant={variant}")
    from sleek_amd import _lib
    from sleek_amd.datagen import exact_field, inject_specials

    lib = _lib.load()
    h0 = help_events()
    try:
        _lib.check(lib.sleek_gpu_set_help_after(0), "help")
        for dt, n in ((np.float32, 4096 * 300 + 7), (np.float64, 2048 * 150 + 3)):
            a = exact_field(n, 17, dt)
            inject_specials(a, 1e-3, 3)
            for grid in (0, 3):
                _lib.check(lib.sleek_gpu_set_tuning(variant, grid, 0), "tuning")
                for eb in (None, 1e-3):
                    s, _ = gpu_roundtrip(a, eb)
                    assert s.tobytes() == po.compress(a, eb).tobytes(), (np.dtype(dt).name, grid, eb)
    finally:
        lib.sleek_gpu_set_help_after(100)
        lib.sleek_gpu_set_tuning(0, 0, 0)
    say(f"help events: {help_events() - h0}")


@pytest.mark.parametrize("variant", [0, 1])
def test_decoder_variant_invariance(gpu, variant):
    """Both decoders (sleek_gpu_set_decoder), at the automatic grid and at 1 and 7
    workgroups (the fused decoder's offset windows then span many chunks)."""
    say(f"decoder variant={variant}")
    from sleek_amd import _lib
    from sleek_amd.datagen import exact_field, inject_specials

    lib = _lib.load()
    try:
        _lib.check(lib.sleek_gpu_set_decoder(variant), "decoder")
        for dt, n in ((np.float32, 4096 * 37 + 5), (np.float64, 2048 * 21), (np.float32, 3)):
            a = exact_field(n, 31, dt)
            inject_specials(a, 1e-3, 5)
            for eb in (None, 1e-3):
                ref = po.compress(a, eb)
                od, rc = po.decompress(ref, dt, eb)
                assert rc == 0
                for grid in (0, 1, 7):
                    _lib.check(lib.sleek_gpu_set_tuning(0, 0, grid), "tuning")
                    _, d = gpu_roundtrip(a, eb)
                    assert d.tobytes() == od.tobytes(), (np.dtype(dt).name, n, eb, grid)
    finally:
        lib.sleek_gpu_set_decoder(0)
        lib.sleek_gpu_set_tuning(0, 0, 0)


@pytest.mark.parametrize("variant", [1, 4, 5, 6, 7])
def test_lookback_geometry_invariance(gpu, variant):
    """Look-back group sizes 8..64 and initial windows 1..64 (sleek_gpu_set_lookback)
    change no byte, at the automatic grid and at a 3-workgroup grid."""
    say(f"look-back geometry, encoder variant={variant}")
    from sleek_amd import _lib
    from sleek_amd.datagen import exact_field

    lib = _lib.load()
    a = exact_field(16384 // 4 * 700 + 33, 9, np.float32)   # 701 chunks: 11 groups of 64, 88 of 8
    refs = {eb: po.compress(a, eb) for eb in (None, 1e-3)}
    try:
        for group, window in ((8, 1), (16, 7), (32, 64), (64, 4), (8, 64)):
            _lib.check(lib.sleek_gpu_set_lookback(group, window), "lookback")
            for grid in (0, 3):
                _lib.check(lib.sleek_gpu_set_tuning(variant, grid, 0), "tuning")
                for eb, ref in refs.items():
                    s, _ = gpu_roundtrip(a, eb)
                    assert s.tobytes() == ref.tobytes(), (group, window, grid, eb)
    finally:
        lib.sleek_gpu_set_lookback(0, 0)
        lib.sleek_gpu_set_tuning(0, 0, 0)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_every_exponent_class(gpu, dt):
    """Every (sign, exponent) pair with random mantissas, lossless and at several
    bounds: the transforms' case analysis runs only on the exponent."""
    say(f"every exponent class {np.dtype(dt).name}")
    rng = np.random.default_rng(8)
    if dt == np.float32:
        se = np.arange(512, dtype=np.uint32)
        bits = (se[:, None] << np.uint32(23)) | rng.integers(0, 1 << 23, (512, 8), dtype=np.uint32)
        a = bits.reshape(-1).view(np.float32)
    else:
        se = np.arange(4096, dtype=np.uint64)
        bits = (se[:, None] << np.uint64(52)) | rng.integers(0, 1 << 52, (4096, 4), dtype=np.uint64)
        a = bits.reshape(-1).view(np.float64)
    for eb in (None, 1e-30, 1e-3, 1e3):
        ref = po.compress(a, eb)
        s, d = gpu_roundtrip(a, eb)
        assert s.tobytes() == ref.tobytes(), eb
        od, rc = po.decompress(ref, dt, eb)
        assert rc == 0 and d.tobytes() == od.tobytes(), eb


@pytest.mark.parametrize("config", [1, 2])
def test_config_fields_scaled(gpu, config):
    """The benchmark recipes (sinusoids, ramp, lognormal noise, specials) at 1/4 scale."""
    say(f"config {config} at 1/4 scale")
    from sleek_amd.datagen import config_field, finite_range

    a = config_field(config, seed=3, scale=0.25).reshape(-1)
    dt = a.dtype
    rel = (1e-2, 1e-3, 1e-4) if dt == np.float32 else (1e-6, 1e-9)
    r = finite_range(a)
    for eb in (None,) + tuple(dt.type(r * dt.type(e)) for e in rel):
        ebf = None if eb is None else float(eb)
        s, d = gpu_roundtrip(a, ebf)
        ref = po.compress(a, ebf)
        assert s.tobytes() == ref.tobytes(), f"stream mismatch eb={ebf}"
        od, rc = po.decompress(ref, dt, ebf)
        assert rc == 0 and d.tobytes() == od.tobytes()
        if ebf is None:
            assert d.tobytes() == a.tobytes()
        else:
            assert maxabs_ok(a, d, eb)


@pytest.mark.parametrize("chunks", [2048, 2049, 32768, 32769, 70001])
def test_many_chunks_round_trip(gpu, chunks):
    """Chunk counts around the offset scan's rounds (8,192 table entries each, one
    workgroup per round up to 64), up to 1.1 GiB: lossless round trips are bit-exact,
    and both decoders (offset scan and fused) give the same bytes; the offset scan
    also at a decode grid of 3, whose 3 workgroups then take several rounds each."""
    import torch
    from sleek_amd import _lib, codec

    n = chunks * 4096 - 3
    say(f"{chunks} chunks ({n} float32)")
    g = torch.Generator(device="cuda").manual_seed(chunks)
    x = (torch.randn(n, device="cuda", generator=g) * 0.01).cumsum(0)   # a compressible walk
    lib = _lib.load()
    try:
        for eb in (None, 1e-3):
            out, sz = codec.compress(x, eb)
            outs = []
            for dec, grid in ((0, 0), (1, 0), (0, 3)):
                _lib.check(lib.sleek_gpu_set_decoder(dec), "decoder")
                _lib.check(lib.sleek_gpu_set_tuning(0, 0, grid), "tuning")
                d, dn = codec.decompress(out, x.dtype, eb, numel=n)
                torch.cuda.synchronize()
                assert int(dn.item()) == 4 * n
                outs.append(d)
            for d in outs[1:]:
                assert torch.equal(outs[0].view(torch.int32), d.view(torch.int32)), eb
            if eb is None:
                assert torch.equal(outs[0].view(torch.int32), x.view(torch.int32))
    finally:
        lib.sleek_gpu_set_decoder(0)
        lib.sleek_gpu_set_tuning(0, 0, 0)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_empty_and_tiny(gpu, dt):
    say(f"empty/tiny {np.dtype(dt).name}")
    for n in (0, 1, 2, 3, 5):
        a = (np.arange(n) * 1.5 - 2).astype(dt)
        for eb in (None, 1e-2):
            s, d = gpu_roundtrip(a, eb)
            assert s.tobytes() == po.compress(a, eb).tobytes()
            if eb is None:
                assert d.tobytes() == a.tobytes()


def test_corrupt_stream_detected(gpu):
    say("corrupt stream")
    import torch
    from sleek_amd import codec, _lib

    a = np.linspace(0, 1, 4096 * 4, dtype=np.float32)
    ref = po.compress(a)
    # damage the trailer of chunk 0 (compressed: size < 16384)
    sizes = ref[8:8 + 2 * 4].view(np.uint16)
    assert sizes[0] < 16384
    bad = ref.copy()
    base = 8 + 2 * 4
    bad[base + int(sizes[0]) - 2] ^= 0x5A
    s = torch.from_numpy(bad).cuda()
    ws = codec.Workspace(a.nbytes)
    codec.decompress(s, torch.float32, numel=a.size, workspace=ws)
    assert ws.status() == _lib.ECORRUPT
    # a size-table entry larger than the chunk
    bad = ref.copy()
    bad[8:10] = np.frombuffer(np.uint16(20000).tobytes(), np.uint8)
    s = torch.from_numpy(bad).cuda()
    codec.decompress(s, torch.float32, numel=a.size, workspace=ws, enclen=bad.size)
    assert ws.status() == _lib.ECORRUPT
    # decoded length above the capacity
    s = torch.from_numpy(ref).cuda()
    codec.decompress(s, torch.float32, numel=a.size - 1, workspace=ws)
    assert ws.status() == _lib.ECAPACITY
    # a clean decode after the failures
    d, _ = codec.decompress(s, torch.float32, numel=a.size, workspace=ws)
    assert ws.status() == 0
    assert d.cpu().numpy().tobytes() == a.tobytes()


def edge_code_stream(dt, seed=7):
    """A one-chunk lossy stream, well formed, whose first subchunk is a full
    word wide and holds the edge codes (0, 1, 2, the largest codes and, for
    float32, the code whose enc + 1 wraps) and random ones; the other
    subchunks are 0 bits wide.  No encoder writes the largest codes, but a
    decoder meets them in damaged streams and must map them as the reference
    does (sleek_common_gpu.h:1077-1108).  The float64 decoder deviates on the
    wrapping code 2^64 - 1 (lossy_value in sleek_device.h), which is left out."""
    wt = np.uint32 if dt == "f32" else np.uint64
    tb = 32 if dt == "f32" else 64
    per_sub = 128 if dt == "f32" else 64
    top = wt(np.iinfo(wt).max)
    codes = np.random.default_rng(seed).integers(0, np.iinfo(wt).max, per_sub, dtype=wt, endpoint=False)
    codes[:8] = [0, 1, 2, 3, top if dt == "f32" else top - wt(3), top - wt(1), top - wt(2), top >> wt(1)]
    widths = np.zeros(32, np.uint8)
    widths[0] = tb
    payload = widths.tobytes() + codes.tobytes() + np.int16(16384).tobytes()
    return np.frombuffer(np.int64(16384).tobytes() + np.uint16(len(payload)).tobytes() + payload, np.uint8).copy()


@pytest.mark.parametrize("dt,eb", [("f32", 1e-3), ("f32", 1e-40), ("f32", 9e30), ("f64", 1e-6), ("f64", 1e-310)])
def test_lossy_decode_of_edge_codes(gpu, dt, eb):
    """Every code decodes as the reference's deQuantize does, the ones that
    only damaged streams hold included."""
    import torch
    from sleek_amd import codec

    say(f"edge codes {dt} eb={eb}")
    stream = edge_code_stream(dt)
    dtype = np.float32 if dt == "f32" else np.float64
    want, st = po.decompress(stream, dtype, eb)
    assert st == 0
    tdt = torch.float32 if dt == "f32" else torch.float64
    ws = codec.Workspace(16384)
    d, dn = codec.decompress(torch.from_numpy(stream).cuda(), tdt, eb, numel=16384 // np.dtype(dtype).itemsize,
                             workspace=ws, enclen=stream.size)
    assert ws.status() == 0 and int(dn.item()) == 16384
    assert d.cpu().numpy().tobytes() == want.tobytes()


@pytest.mark.parametrize("dt,n,eb", [("f32", 100_003, 0.0), ("f32", 100_003, 1e-3), ("f64", 50_001, 0.0),
                                     ("f64", 50_001, 1e-6), ("f32", 0, 0.0)])
def test_cpp_dropin_headers(gpu, tmp_path, dt, n, eb):
    """A C++ caller of the reference API (drop-in headers, g++-built) gets the oracle's bytes."""
    import subprocess

    say(f"drop-in C++ {dt} n={n} eb={eb}")
    exe = os.path.join(os.path.dirname(__file__), "cpp", "dropin_check")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/dropin_check not built")
    so, do = tmp_path / "s.bin", tmp_path / "d.bin"
    r = subprocess.run([exe, dt, str(n), str(eb), str(so), str(do)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    npdt = np.float32 if dt == "f32" else np.float64
    i = np.arange(n, dtype=np.int64)
    a = (i.astype(np.float64) * 0.001 + ((i.astype(np.uint64) * np.uint64(2654435761)) % np.uint64(1000))
         .astype(np.float64) * 1e-6).astype(npdt)
    ebv = None if eb == 0 else float(npdt(eb))
    ref = po.compress(a, ebv)
    assert open(so, "rb").read() == ref.tobytes()
    od, rc = po.decompress(ref, npdt, ebv)
    assert rc == 0 and open(do, "rb").read() == od.tobytes()


def test_kernel_events_bracket_the_coding_kernels(gpu):
    """sleek_gpu_set_kernel_events: the next call records the caller's events right
    around its coding kernel (encoder; decoder after its offset scan), once; the bytes
    are unchanged, the kernel's time is positive and within the whole call's, and a
    later call records nothing (the setting is used once)."""
    say("kernel events around the encoder and the decoder")
    import torch
    from sleek_amd import _lib, codec
    from sleek_amd.datagen import exact_field

    lib = _lib.load()
    a = exact_field(4096 * 300 + 5, 23, np.float32)
    x = torch.from_numpy(a).cuda()
    ref = po.compress(a, 1e-3)
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(8)]
    for e in ev:
        e.record(st)   # creates the HIP events
    torch.cuda.synchronize()
    ev[0].record(st)
    _lib.check(lib.sleek_gpu_set_kernel_events(ev[2].cuda_event, ev[3].cuda_event), "events")
    out, n = codec.compress(x, 1e-3)
    ev[1].record(st)
    _lib.check(lib.sleek_gpu_set_kernel_events(ev[6].cuda_event, ev[7].cuda_event), "events")
    d, _ = codec.decompress(out, x.dtype, 1e-3, numel=a.size)
    ev[4].record(st)
    torch.cuda.synchronize()
    assert out[: int(n.item())].cpu().numpy().tobytes() == ref.tobytes()
    od, rc = po.decompress(ref, np.float32, 1e-3)
    assert rc == 0 and d.cpu().numpy().tobytes() == od.tobytes()
    k_enc, call_enc = ev[2].elapsed_time(ev[3]), ev[0].elapsed_time(ev[1])
    k_dec, call_dec = ev[6].elapsed_time(ev[7]), ev[1].elapsed_time(ev[4])
    assert 0 < k_enc <= call_enc and 0 < k_dec <= call_dec, (k_enc, call_enc, k_dec, call_dec)
    # used once: a second call leaves the events where the first one recorded them
    ev[5].record(st)
    codec.compress(x, 1e-3)
    torch.cuda.synchronize()
    assert ev[7].elapsed_time(ev[5]) > 0   # the decoder's events were not recorded again after ev[5]
