"""The stream container: header round trip, the streams inside unchanged, and
malformed headers rejected (host only)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from sleek_amd import container as c


def test_pack_parse_round_trip():
    a = np.linspace(-3, 7, 50_000, dtype=np.float32)
    halves = [a[:32768], a[32768:]]
    streams = [po.compress(h, 1e-3) for h in halves]
    buf = c.pack(streams, [h.nbytes for h in halves], np.float32, 1e-3)
    m = c.parse(buf)
    assert m.dtype == np.float32 and m.eb == pytest.approx(1e-3) and m.total_bytes == a.nbytes
    for (off, ln, raw), s, h in zip(m.segments, streams, halves):
        assert off % 8 == 0 and raw == h.nbytes
        assert buf[off: off + ln].tobytes() == s.tobytes()   # the stream is unchanged
        d, rc = po.decompress(buf[off: off + ln], np.float32, m.eb)
        assert rc == 0 and np.all(np.abs(d - h) <= np.float32(1e-3))


def test_wrap_unwrap_lossless_f64():
    a = np.arange(1000, dtype=np.float64) * 0.25
    s = po.compress(a)
    w = c.wrap(s, np.float64)
    st, dt, eb = c.unwrap(w)
    assert dt == np.float64 and eb is None and st.tobytes() == s.tobytes()
    assert c.parse(w).total_bytes == a.nbytes


@pytest.mark.parametrize("damage", ["magic", "version", "elem", "segment", "total", "eb"])
def test_malformed_headers_are_rejected(damage):
    a = np.ones(4096, dtype=np.float32)
    buf = c.pack([po.compress(a, 0.5)], [a.nbytes], np.float32, 0.5).copy()
    if damage == "magic":
        buf[0] ^= 1
    elif damage == "version":
        buf[4] = 9
    elif damage == "elem":
        buf[8] = 3
    elif damage == "segment":
        buf[c.HEAD + 8: c.HEAD + 16].view(np.uint64)[0] = buf.size   # stream length beyond the buffer
    elif damage == "total":
        buf[24:32].view(np.uint64)[0] += 4
    elif damage == "eb":
        buf[16:24].view(np.float64)[0] = -1.0
    with pytest.raises(ValueError):
        c.parse(buf)


def test_unwrap_needs_one_segment():
    a = np.ones(10, dtype=np.float32)
    buf = c.pack([po.compress(a), po.compress(a)], [a.nbytes, a.nbytes], np.float32, None)
    with pytest.raises(ValueError):
        c.unwrap(buf)
