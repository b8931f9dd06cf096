"""Compressing copies between device and host memory.

The reference measures what plain copies cost (memcpy.cu:114-139); SLEEK's
use is to move compressed bytes instead.  ``offload`` copies a device array
to pinned host memory as a container (``container``) of per-segment SLEEK
streams: segment i is compressed on a compute stream while segment i-1's
stream crosses PCIe on a copy stream.  ``reload`` brings a container back:
segment i's stream is copied to the device while segment i-1 decompresses
into the output.  Each segment's stream is byte-identical to the
single-array codec's stream of that segment, so a container's streams can
also be decoded one by one with ``codec.decompress``.

Two device stream buffers and workspaces (one per segment in flight) are
the only device scratch; the host side is one pinned buffer.  A workspace is
reused every other segment, so each segment's status word and decoded length
are copied to pinned host memory on the compute stream right after its
decode, and every segment is checked before ``reload`` returns.
"""
import numpy as np
import torch

from . import _lib, codec, container

DEFAULT_SEGMENT = 128 << 20


def _segments(nbytes, elem, segment_bytes):
    seg = max(int(segment_bytes) // codec.CHUNK, 1) * codec.CHUNK   # whole chunks per segment
    seg -= seg % elem
    return [(o, min(seg, nbytes - o)) for o in range(0, nbytes, seg)] or [(0, 0)]


def offload(x, eb=None, segment_bytes=DEFAULT_SEGMENT):
    """Compresses the contiguous float32/float64 CUDA tensor x into a container
    in pinned host memory.  Returns the container as a pinned uint8 CPU tensor
    (its length is the container's).  eb: absolute error bound (None =
    lossless).  Synchronous: returns once the bytes are on the host."""
    if not x.is_cuda or not x.is_contiguous():
        raise ValueError("x must be a contiguous CUDA tensor")
    elem = codec._elem_bytes(x.dtype)
    nbytes = x.numel() * elem
    segs = _segments(nbytes, elem, segment_bytes)
    dev = x.device
    flat = x.view(torch.uint8).view(-1) if nbytes else torch.empty(0, dtype=torch.uint8, device=dev)
    head = container.align8(container.header_size(len(segs)))
    cap = head + sum(container.align8(codec.upper_bound_size(ln)) for _, ln in segs)
    host = torch.empty(cap, dtype=torch.uint8, pin_memory=True)
    seg_max = max(ln for _, ln in segs)
    ub = max(codec.upper_bound_size(seg_max), 8)
    dbuf = [torch.empty(ub, dtype=torch.uint8, device=dev) for _ in range(2)]
    dsize = [torch.empty(1, dtype=torch.int64, device=dev) for _ in range(2)]
    hsize = torch.empty(len(segs), dtype=torch.int64, pin_memory=True)
    ws = [codec.Workspace(seg_max, dev) for _ in range(2)]
    comp, copy = torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)
    comp.wait_stream(torch.cuda.current_stream(dev))
    sized = [torch.cuda.Event() for _ in segs]
    freed = [None, None]   # copy of the stream buffer's previous segment finished

    def enqueue(i):
        b = i & 1
        o, ln = segs[i]
        with torch.cuda.stream(comp):
            if freed[b] is not None:
                comp.wait_event(freed[b])
            seg = flat[o: o + ln].view(x.dtype)
            codec.compress(seg, eb, out=dbuf[b], encsize=dsize[b], workspace=ws[b], stream=comp)
            hsize[i: i + 1].copy_(dsize[b], non_blocking=True)
            sized[i].record(comp)

    meta = []
    pos = head
    enqueue(0)
    for i in range(len(segs)):
        if i + 1 < len(segs):
            enqueue(i + 1)            # the next segment codes while this one is copied
        sized[i].synchronize()
        n = int(hsize[i])
        with torch.cuda.stream(copy):
            copy.wait_event(sized[i])
            host[pos: pos + n].copy_(dbuf[i & 1][:n], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(copy)
            freed[i & 1] = ev
        meta.append((pos, n, segs[i][1]))
        pos = container.align8(pos + n)
    copy.synchronize()
    for w in ws[: len(segs)]:   # an unused workspace holds no status
        _lib.check(w.status(comp), "offload")
    h = host.numpy()
    for o, n, _ in meta:
        h[o + n: container.align8(o + n)] = 0   # padding between streams
    container.write_header(h, np.float32 if elem == 4 else np.float64, eb, meta)
    return host[:pos]


def reload(host, device=None, out=None):
    """Decompresses a container held in (preferably pinned) host memory onto the
    device; returns the array as a CUDA tensor.  The dtype and error bound come
    from the container.  Synchronous."""
    buf = host.numpy() if isinstance(host, torch.Tensor) else np.asarray(host, dtype=np.uint8)
    meta = container.parse(buf)
    if not isinstance(host, torch.Tensor):
        host = torch.from_numpy(buf)
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    tdt = torch.float32 if meta.dtype == np.float32 else torch.float64
    elem = meta.dtype.itemsize
    if out is None:
        out = torch.empty(meta.total_bytes // elem, dtype=tdt, device=dev)
    elif out.dtype != tdt or out.numel() * elem != meta.total_bytes or not out.is_contiguous():
        raise ValueError("out does not match the container")
    flat = out.view(torch.uint8).view(-1)
    if not meta.segments:
        return out
    slen = max(ln for _, ln, _ in meta.segments)
    raw_max = max(r for _, _, r in meta.segments)
    dbuf = [torch.empty(slen, dtype=torch.uint8, device=dev) for _ in range(2)]
    dsz = [torch.empty(1, dtype=torch.int64, device=dev) for _ in range(2)]
    ws = [codec.Workspace(raw_max, dev) for _ in range(2)]
    # per segment: the workspace's status word and the decoded length, taken
    # before the workspace and the length slot serve the segment after next
    seen = torch.zeros((len(meta.segments), 2), dtype=torch.int64, pin_memory=True)
    comp, copy = torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)
    comp.wait_stream(torch.cuda.current_stream(dev))
    copy.wait_stream(torch.cuda.current_stream(dev))
    done = [None, None]   # decode of the buffer's previous segment finished
    pos = 0
    for i, (off, ln, raw) in enumerate(meta.segments):
        b = i & 1
        with torch.cuda.stream(copy):
            if done[b] is not None:
                copy.wait_event(done[b])
            dbuf[b][:ln].copy_(host[off: off + ln], non_blocking=True)
            landed = torch.cuda.Event()
            landed.record(copy)
        with torch.cuda.stream(comp):
            comp.wait_event(landed)
            seg = flat[pos: pos + raw].view(tdt)
            codec.decompress(dbuf[b], meta.dtype, meta.eb, out=seg, decsize=dsz[b], workspace=ws[b], enclen=ln,
                             stream=comp)
            seen[i, 0:1].copy_(ws[b].status_word(), non_blocking=True)
            seen[i, 1:2].copy_(dsz[b], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(comp)
            done[b] = ev
        pos += raw
    comp.synchronize()
    for i, (_, _, raw) in enumerate(meta.segments):
        st, n = int(seen[i, 0]), int(seen[i, 1])
        if st:
            _lib.check(codec.status_of_word(st), f"reload: segment {i}")
        if n != raw:
            raise _lib.SleekError(_lib.ECORRUPT, f"reload: segment {i} decoded to {n} bytes, the container says {raw}")
    torch.cuda.current_stream(dev).wait_stream(comp)
    return out
