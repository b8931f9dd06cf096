"""The oracle, pinned: against the reference codec's golden streams, hand-
derived known answers, exhaustive and randomized transform properties, and
(where it is built here) the compiled reference codec itself."""
import hashlib
import json
import os

import numpy as np
import pytest

from cases import build_input, eb_value
from oracle import pyoracle as po

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))["cases"]


def sha(b):
    return hashlib.sha256(np.ascontiguousarray(b).view(np.uint8).tobytes()).hexdigest()


@pytest.mark.parametrize("case", GOLDEN, ids=[c["name"] for c in GOLDEN])
def test_oracle_matches_reference_golden(case):
    a = build_input(case)
    assert sha(a) == case["input_sha256"], "input generator drifted"
    eb = eb_value(case, a)
    if eb is not None:
        bits = int(np.array(eb).view(np.uint32 if a.dtype == np.float32 else np.uint64))
        assert bits == case["eb_bits"]
    ebf = None if eb is None else float(eb)
    s = po.compress(a, ebf)
    assert s.size == case["stream_len"]
    if case["stream_hex"] is not None:
        assert s.tobytes().hex() == case["stream_hex"]
    assert sha(s) == case["stream_sha256"]
    d, rc = po.decompress(s, a.dtype, ebf)
    assert rc == 0
    assert sha(d) == case["decoded_sha256"]


# ---- known answers derived by hand from sleek_common_cpu.h ----

@pytest.mark.parametrize("x,code", [
    (0x00000000, 0x00000000), (0x80000000, 0x00000002), (0x3F800000, 0x03FFFFFF),
    (0x40000000, 0x01FFFFFF), (0x40800000, 0x02000000), (0x7F800000, 0xFE000000),
    (0x00000001, 0x00000004)])
def test_lossless_code_f32(x, code):
    L = po.lib()
    assert L.sleek_oracle_fwd_u32(x) == code
    assert L.sleek_oracle_inv_u32(code) == x


def test_lossless_code_f64():
    L = po.lib()
    assert L.sleek_oracle_fwd_u64(0x3FF0000000000000) == 0x007FFFFFFFFFFFFF
    assert L.sleek_oracle_inv_u64(0x007FFFFFFFFFFFFF) == 0x3FF0000000000000


def test_error_bound_params():
    assert po.error_bound(0.01, np.float32) == (120, 144, 0x47800000)
    # eb = 1.0 (f64): biased exponent 1023
    e, t, o = po.error_bound(1.0, np.float64)
    assert (e, t) == (1023, 1023 + 53) and o == ((1023 + 53) << 52) - (1 << 52)
    with pytest.raises(ValueError):
        po.error_bound(2.0 ** 104, np.float32)   # thr_e = 255: the reference exits here
    po.error_bound(2.0 ** 103, np.float32)
    with pytest.raises(ValueError):
        po.error_bound(float("inf"), np.float32)


@pytest.mark.parametrize("x,code,back", [
    (1.0, 128, 0x3F800000), (-1.0, 127, 0xBF800000), (0.01, 2, 0x3C800000),
    (0.005, 0, 0x00000000), (0.0, 0, 0), (-0.0, 0, 0)])
def test_quantize_f32_kat(x, code, back):
    L = po.lib()
    e, t, o = po.error_bound(0.01, np.float32)
    bits = int(np.array(x, np.float32).view(np.uint32))
    c = L.sleek_oracle_quantize_u32(bits, e, t, o)
    assert c == code
    assert L.sleek_oracle_dequantize_u32(c, t, o) == back


def test_quantize_special_values_exact():
    L = po.lib()
    e, t, o = po.error_bound(0.01, np.float32)
    for bits in (0x7F800000, 0xFF800000, 0x7FC00000, 0x7FC12345, 0xFFC00001):
        c = L.sleek_oracle_quantize_u32(bits, e, t, o)
        assert L.sleek_oracle_dequantize_u32(c, t, o) == bits
    assert L.sleek_oracle_quantize_u32(0x7F800000, e, t, o) == 0x70000000
    assert L.sleek_oracle_quantize_u32(0x7FC00000, e, t, o) == 0x70800000


def test_stream_kats():
    # one chunk of zeros: 44 bytes (header, size 34, 32 zero bit lengths, trailer 16384)
    s = po.compress(np.zeros(4096, np.float32))
    assert s.tobytes().hex() == "0040000000000000" + "2200" + "00" * 32 + "0040"
    assert po.compress(np.zeros(2048, np.float64)).tobytes() == s.tobytes()
    # a single 1.0f is stored raw
    s = po.compress(np.ones(1, np.float32))
    assert s.tobytes().hex() == "0400000000000000" + "0400" + "0000803f"
    # 4096 x 1.0f: every bit length 26, size 34 + 13312
    s = po.compress(np.ones(4096, np.float32))
    assert s.size == 13356
    assert s[8:10].view(np.uint16)[0] == 13346
    assert np.all(s[10:42] == 26)


def test_empty_input():
    s = po.compress(np.zeros(0, np.float32))
    assert s.tobytes() == b"\0" * 8
    d, rc = po.decompress(s, np.float32)
    assert rc == 0 and d.size == 0


def test_lossless_bijection_f32_exhaustive():
    """inv(fwd(x)) == x and fwd injective over all 2^32 float32 patterns."""
    assert po.lib().sleek_oracle_check_bijection_u32() == 0


def test_lossless_roundtrip_f64_random():
    L = po.lib()
    rng = np.random.default_rng(0)
    xs = rng.integers(0, np.iinfo(np.uint64).max, size=20000, dtype=np.uint64, endpoint=True)
    specials = [0, 1 << 63, 0x7FF0000000000000, 0xFFF0000000000000, 0x7FF8000000000001, 1, 0x000FFFFFFFFFFFFF]
    for x in list(map(int, xs)) + specials:
        assert L.sleek_oracle_inv_u64(L.sleek_oracle_fwd_u64(x)) == x


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_lossy_bound_and_monotone(dt):
    """|deq(q(x)) - x| <= 2^(eb exponent) <= eb for finite x; NaN/Inf exact."""
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 10, 5000), rng.normal(0, 1e-3, 2000), [0.0, -0.0, 1e-30, -5e-3]]).astype(dt)
    for eb in (1e-2, 1e-4, 3e-7):
        s = po.compress(x, eb)
        d, rc = po.decompress(s, dt, eb)
        assert rc == 0
        ebt = dt(eb)
        assert np.all(np.abs(x - d) <= ebt)


def test_decoder_rejects_malformed_streams():
    a = np.linspace(0, 1, 4096 * 3, dtype=np.float32)
    s = po.compress(a)
    # trailer damage
    b = s.copy()
    c0 = int(s[8:10].view(np.uint16)[0])
    b[8 + 6 + c0 - 2] ^= 0x11
    assert po.decompress(b, np.float32)[1] == po.E_CORRUPT
    # size entry above the chunk size
    b = s.copy()
    b[8:10] = np.frombuffer(np.uint16(16385).tobytes(), np.uint8)
    assert po.decompress(b, np.float32)[1] == po.E_CORRUPT
    # truncated stream
    assert po.decompress(s[:-5], np.float32)[1] == po.E_CORRUPT
    # capacity
    assert po.decompress(s, np.float32, capacity=a.nbytes - 4)[1] == po.E_CAP


def test_oracle_vs_compiled_reference_random():
    if po.ref() is None:
        pytest.skip("reference codec not built here (oracle/_ref)")
    rng = np.random.default_rng(7)
    for dt in (np.float32, np.float64):
        per = 16384 // np.dtype(dt).itemsize
        for trial in range(6):
            n = int(rng.integers(1, per * 9))
            mix = rng.random()
            if mix < 0.3:
                a = rng.normal(0, 10 ** rng.uniform(-3, 3), n).astype(dt)
            elif mix < 0.6:
                bits = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32) if dt == np.float32 \
                    else rng.integers(0, np.iinfo(np.uint64).max, size=n, dtype=np.uint64, endpoint=True)
                a = bits.view(dt)
            else:
                a = (np.cumsum(rng.normal(0, 1, n)) * 1e-2).astype(dt)
            for eb in (None, 1e-1, 1e-4, 1e-7):
                s1 = po.compress(a, eb)
                s2 = po.ref_compress(a, eb)
                assert s1.tobytes() == s2.tobytes(), (np.dtype(dt).name, n, eb)
                d1, rc = po.decompress(s1, dt, eb)
                assert rc == 0
                assert d1.tobytes() == po.ref_decompress(s2, dt, eb).tobytes()


def test_upper_bound_size():
    for n in (0, 1, 16384, 16385, 1 << 30):
        ch = (n + 16383) // 16384
        assert po.upper_bound_size(n) == 12 + 2 * ch + 16384 * ch
