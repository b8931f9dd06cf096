"""Compressing device<->host copies: offload() leaves a container in pinned host
memory whose per-segment streams are the oracle's streams of those segments,
and reload() brings back the oracle's decode (the input itself when lossless)."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dt,eb,n,seg", [
    (np.float32, None, 300_001, 256 << 10),
    (np.float32, 1e-3, 300_001, 128 << 10),
    (np.float64, None, 90_017, 512 << 10),
    (np.float64, 1e-7, 90_017, 64 << 10),
    (np.float32, 1e-2, 5_000, 64 << 20),     # one segment
    (np.float32, None, 0, 1 << 20),          # empty array
])
def test_offload_reload(gpu, dt, eb, n, seg):
    import torch
    from sleek_amd import container, transfer
    from sleek_amd.datagen import exact_field

    say(f"transfer {np.dtype(dt).name} eb={eb} n={n} segment={seg}")
    a = exact_field(n, 11, dt) if n else np.zeros(0, dtype=dt)
    x = torch.from_numpy(a).cuda()
    host = transfer.offload(x, eb, segment_bytes=seg)
    assert host.is_pinned()
    buf = host.numpy()
    m = container.parse(buf)
    assert m.total_bytes == a.nbytes and m.dtype == dt
    pos = 0
    for off, ln, raw in m.segments:
        part = a.view(np.uint8)[pos: pos + raw].view(dt)
        assert buf[off: off + ln].tobytes() == po.compress(part, eb).tobytes()
        pos += raw
    y = transfer.reload(host)
    torch.cuda.synchronize()
    if eb is None or not n:
        want = a
    else:   # segments decode independently: the oracle's decode, segment by segment
        want = np.concatenate([po.decompress(buf[o: o + ln], dt, eb)[0] for o, ln, _ in m.segments])
    assert y.cpu().numpy().tobytes() == want.tobytes()


def test_sdr_file_to_device(gpu, tmp_path):
    """A raw input file reaches the device intact through the pinned slabs."""
    import torch
    from sleek_amd import sdr

    say("sdr to_device")
    a = np.random.default_rng(3).standard_normal(300_000).astype(np.float64)
    p = tmp_path / "x.d64"
    a.tofile(p)
    x = sdr.to_device(str(p), slab_bytes=1 << 20)
    torch.cuda.synchronize()
    assert x.dtype == torch.float64 and x.cpu().numpy().tobytes() == a.tobytes()


@pytest.mark.parametrize("where", ["trailer", "header"])
def test_reload_reports_a_damaged_early_segment(gpu, where):
    """Damage in segment 0 of a multi-segment container is reported although the
    two workspaces have served later segments since: each segment's status and
    decoded length are checked."""
    import torch
    from sleek_amd import _lib, container, transfer
    from sleek_amd.datagen import exact_field

    say(f"transfer: damaged segment 0 ({where})")
    a = exact_field(300_001, 12, np.float32)
    x = torch.from_numpy(a).cuda()
    host = transfer.offload(x, None, segment_bytes=256 << 10)
    buf = host.numpy()
    m = container.parse(buf)
    assert len(m.segments) >= 4
    off, ln, raw = m.segments[0]
    if where == "trailer":
        nch = (raw + 16383) // 16384
        c0 = int(buf[off + 8: off + 10].view(np.uint16)[0])
        assert c0 < 16384, "chunk 0 of segment 0 must be a packed chunk"
        buf[off + 8 + 2 * nch + c0 - 2] ^= 0x5A   # chunk 0's int16 trailer
    else:
        buf[off: off + 8] = np.frombuffer(np.int64(raw - 4).tobytes(), np.uint8)   # a shorter decoded length
    with pytest.raises(_lib.SleekError):
        transfer.reload(host)
    torch.cuda.synchronize()
