"""A thin container around unchanged SLEEK streams.

A SLEEK stream records neither its element type nor its mode nor the error
bound (sleek_lossy_gpu.h:90-99 takes eb_info from the caller), so a decoder
given the wrong ones produces garbage without noticing.  The container
records them next to one or more streams, each the byte-identical stream of
one contiguous segment of the array, so segments can be coded and moved
independently (``transfer``).  Layout, little-endian:

    0   4   magic b"SLKC"
    4   2   version (1)
    6   2   flags: bit 0 = lossy
    8   4   element bytes (4 or 8)
    12  4   segment count n
    16  8   absolute error bound as float64 (0 when lossless)
    24  8   original bytes (sum over segments)
    32  24n per segment: stream offset, stream length, original bytes (uint64)
    ...     the streams, each starting at a multiple of 8
"""
import numpy as np

MAGIC = b"SLKC"
VERSION = 1
HEAD = 32
SEG = 24


def header_size(nseg):
    return HEAD + SEG * int(nseg)


def align8(n):
    return (int(n) + 7) & ~7


class Meta:
    """What a container says about its streams."""

    def __init__(self, dtype, eb, total_bytes, segments):
        self.dtype = np.dtype(dtype)
        self.eb = eb                    # None = lossless
        self.total_bytes = int(total_bytes)
        self.segments = segments        # [(stream offset, stream length, original bytes)]

    def __repr__(self):
        return (f"Meta(dtype={self.dtype.name}, eb={self.eb}, total_bytes={self.total_bytes}, "
                f"segments={len(self.segments)})")


def write_header(buf, dtype, eb, segments):
    """Writes the header for ``segments`` [(offset, length, original bytes)] into
    the first header_size(len(segments)) bytes of the uint8 array ``buf``."""
    dtype = np.dtype(dtype)
    if dtype not in (np.float32, np.float64):
        raise TypeError(f"SLEEK codes float32 and float64 only, not {dtype}")
    n = len(segments)
    h = np.zeros(header_size(n), dtype=np.uint8)
    h[0:4] = np.frombuffer(MAGIC, dtype=np.uint8)
    h[4:8].view(np.uint16)[:] = (VERSION, 0 if eb is None else 1)
    h[8:16].view(np.uint32)[:] = (dtype.itemsize, n)
    h[16:24].view(np.float64)[0] = 0.0 if eb is None else float(eb)
    h[24:32].view(np.uint64)[0] = sum(int(s[2]) for s in segments)
    if n:
        h[HEAD:].view(np.uint64)[:] = np.asarray(segments, dtype=np.uint64).reshape(-1)
    buf[: h.size] = h
    return h.size


def pack(streams, raw_bytes, dtype, eb):
    """A container (uint8 numpy array) holding the given streams (uint8 arrays)
    of segments of ``raw_bytes`` original bytes each."""
    n = len(streams)
    off = align8(header_size(n))
    segs = []
    for s, r in zip(streams, raw_bytes):
        segs.append((off, int(np.asarray(s).size), int(r)))
        off = align8(off + int(np.asarray(s).size))
    out = np.zeros(off, dtype=np.uint8)
    write_header(out, dtype, eb, segs)
    for s, (o, ln, _) in zip(streams, segs):
        out[o: o + ln] = np.asarray(s, dtype=np.uint8)
    return out


def parse(buf):
    """Meta of a container (uint8 array-like); ValueError if the header is
    malformed or names bytes beyond the buffer."""
    b = np.asarray(buf, dtype=np.uint8)
    if b.size < HEAD or bytes(b[:4]) != MAGIC:
        raise ValueError("not a SLEEK container")
    version, flags = (int(v) for v in b[4:8].view(np.uint16))
    if version != VERSION:
        raise ValueError(f"unsupported container version {version}")
    elem, n = (int(v) for v in b[8:16].view(np.uint32))
    if elem not in (4, 8):
        raise ValueError(f"bad element size {elem}")
    if b.size < header_size(n):
        raise ValueError("truncated container header")
    eb = float(b[16:24].view(np.float64)[0])
    lossy = bool(flags & 1)
    if lossy and not (eb > 0 and np.isfinite(eb)):
        raise ValueError("lossy container without a valid error bound")
    total = int(b[24:32].view(np.uint64)[0])
    segs = [tuple(int(v) for v in row) for row in b[HEAD: header_size(n)].view(np.uint64).reshape(n, 3)] if n else []
    if sum(s[2] for s in segs) != total:
        raise ValueError("segment sizes do not add up")
    for off, ln, raw in segs:
        if off < header_size(n) or off + ln > b.size or ln < 8 or raw % elem:
            raise ValueError("segment outside the container")
    return Meta(np.float32 if elem == 4 else np.float64, eb if lossy else None, total, segs)


def wrap(stream, dtype, eb=None):
    """One stream (of a whole array) in a container."""
    s = np.asarray(stream, dtype=np.uint8)
    raw = int(s[:8].view(np.int64)[0]) if s.size >= 8 else 0
    return pack([s], [raw], dtype, eb)


def unwrap(buf):
    """(stream, dtype, eb) of a single-segment container; the stream is a view."""
    m = parse(buf)
    if len(m.segments) != 1:
        raise ValueError(f"container holds {len(m.segments)} segments, not one")
    off, ln, _ = m.segments[0]
    return np.asarray(buf, dtype=np.uint8)[off: off + ln], m.dtype, m.eb
