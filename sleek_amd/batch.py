"""Batched front end: many arrays in one launch per direction.

Arrays may mix float32/float64 and lossless/lossy; every array's stream and
decoded bytes are exactly those of a single-array call (and of the reference
codec).  A ``Batch`` is planned once (tile numbering and workspace layout by
``sleek_gpu_batch_plan``), its descriptors are copied to the device, and it
can then be launched eagerly (``launch``) or captured into a HIP graph
(``capture``) and replayed.  All coding runs in libsleek_amd.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .codec import _elem_bytes, error_bound, upper_bound_size

# layout of sleek_gpu_batch_item (include/sleek_gpu.h)
ITEM = np.dtype([
    ("src", "<u8"), ("dst", "<u8"), ("src_bytes", "<i8"), ("dst_bytes", "<i8"), ("d_size", "<u8"),
    ("offs", "<u8"), ("eb_e", "<i4"), ("thr_e", "<i4"), ("elem_bytes", "<i4"), ("lossy", "<i4"),
    ("tile0", "<i8"), ("tiles", "<i8"), ("ws_word", "<i8"), ("ws_words", "<i8"),
])
assert ITEM.itemsize == 96


def _stream_handle(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def fill_item(rec, dtype, eb):
    """Type and error-bound fields of one item (eb None = lossless)."""
    rec["elem_bytes"] = _elem_bytes(dtype)
    if eb is None:
        rec["lossy"], rec["eb_e"], rec["thr_e"], rec["offs"] = 0, 0, 0, 0
    else:
        e, t, o = error_bound(eb, dtype)
        rec["lossy"], rec["eb_e"], rec["thr_e"], rec["offs"] = 1, e, t, o


def plan(items, decompress):
    """Runs sleek_gpu_batch_plan on a host ITEM array in place; returns (tiles, workspace bytes)."""
    tiles, wsb = ctypes.c_longlong(), ctypes.c_longlong()
    items = np.ascontiguousarray(items)
    _lib.check(_lib.load().sleek_gpu_batch_plan(items.ctypes.data, len(items), int(decompress),
                                                ctypes.byref(tiles), ctypes.byref(wsb)), "batch plan")
    return tiles.value, wsb.value


class Graph:
    """An executable HIP graph of one batched launch sequence."""

    def __init__(self, handle, batch):
        self.handle = handle
        self.batch = batch   # keeps the items, buffers and workspace alive

    def launch(self, stream=None):
        _lib.check(_lib.load().sleek_gpu_graph_launch(self.handle, _stream_handle(stream)), "graph launch")

    def close(self):
        if self.handle:
            _lib.load().sleek_gpu_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A planned batched call over n arrays in one direction."""

    def __init__(self, items, decompress, device, keep=(), stream=None):
        """stream: when given, the descriptors are copied (pinned, asynchronous)
        and the workspace allocated in that stream's order, for launches on it."""
        self.items = np.ascontiguousarray(items)
        self.n = len(self.items)
        self.decompress = bool(decompress)
        self.tiles, self.ws_bytes = plan(self.items, decompress)
        raw = torch.from_numpy(self.items.view(np.uint8).copy())
        if stream is None:
            self.d_items = raw.to(device)
            self.ws = torch.empty(max(self.ws_bytes, 16), dtype=torch.uint8, device=device)
        else:
            with torch.cuda.stream(stream):
                self.d_items = raw.pin_memory().to(device, non_blocking=True)
                self.ws = torch.empty(max(self.ws_bytes, 16), dtype=torch.uint8, device=device)
        self.keep = list(keep)

    def launch(self, grid=0, stream=None):
        lib = _lib.load()
        fn = lib.sleek_gpu_batch_decompress_async if self.decompress else lib.sleek_gpu_batch_compress_async
        _lib.check(fn(self.d_items.data_ptr(), self.n, self.tiles, self.ws.data_ptr(), self.ws_bytes, int(grid),
                      _stream_handle(stream)), "batch launch")
        return self

    def status(self, stream=None):
        """Per-array status (0, ECORRUPT, ECAPACITY); synchronises the stream."""
        st = np.zeros(self.n, dtype=np.int32)
        _lib.check(_lib.load().sleek_gpu_batch_status(self.d_items.data_ptr(), self.n, self.ws.data_ptr(),
                                                      self.ws_bytes, st.ctypes.data, _stream_handle(stream)),
                   "batch status")
        return st

    def capture(self, grid=0):
        h = ctypes.c_void_p()
        _lib.check(_lib.load().sleek_gpu_batch_graph_create(int(self.decompress), self.d_items.data_ptr(), self.n,
                                                            self.tiles, self.ws.data_ptr(), self.ws_bytes, int(grid),
                                                            ctypes.byref(h)), "graph capture")
        return Graph(h.value, self)


def plan_compress(xs, ebs, outs=None, sizes=None):
    """Batch compressing the contiguous float32/float64 CUDA tensors xs
    (ebs[i]: None = lossless, else the absolute error bound).  outs[i]: uint8
    buffers of at least the upper-bound size (allocated if None); sizes: an
    int64 tensor of n stream lengths.  Returns the Batch with .outs, .sizes."""
    n = len(xs)
    dev = xs[0].device if n else torch.device("cuda")
    if outs is None:
        outs = [torch.empty(upper_bound_size(x.numel() * x.element_size()), dtype=torch.uint8, device=dev) for x in xs]
    if sizes is None:
        sizes = torch.zeros(max(n, 1), dtype=torch.int64, device=dev)
    items = np.zeros(n, dtype=ITEM)
    for i, (x, eb) in enumerate(zip(xs, ebs)):
        if not x.is_contiguous():
            raise ValueError("batched compress needs contiguous arrays")
        nb = x.numel() * x.element_size()
        if outs[i].numel() < upper_bound_size(nb):
            raise ValueError(f"output buffer {i} is smaller than the upper bound")
        items[i]["src"] = x.data_ptr()
        items[i]["dst"] = outs[i].data_ptr()
        items[i]["src_bytes"] = nb
        items[i]["dst_bytes"] = outs[i].numel()
        items[i]["d_size"] = sizes.data_ptr() + 8 * i
        fill_item(items[i], x.dtype, eb)
    b = Batch(items, False, dev, keep=list(xs) + list(outs) + [sizes])
    b.outs, b.sizes = outs, sizes
    return b


def plan_decompress(streams, dtypes, ebs, numels, outs=None, sizes=None, enclens=None):
    """Batch decompressing streams[i] (uint8 CUDA tensors) into numels[i]
    values of dtypes[i] (the capacity; the stream header gives the length).
    Returns the Batch with .outs, .sizes."""
    n = len(streams)
    dev = streams[0].device if n else torch.device("cuda")
    tdt = [torch.float32 if _elem_bytes(d) == 4 else torch.float64 for d in dtypes]
    if outs is None:
        outs = [torch.empty(int(m), dtype=t, device=dev) for m, t in zip(numels, tdt)]
    if sizes is None:
        sizes = torch.zeros(max(n, 1), dtype=torch.int64, device=dev)
    items = np.zeros(n, dtype=ITEM)
    for i in range(n):
        items[i]["src"] = streams[i].data_ptr()
        items[i]["dst"] = outs[i].data_ptr()
        items[i]["src_bytes"] = streams[i].numel() if enclens is None else int(enclens[i])
        items[i]["dst_bytes"] = outs[i].numel() * outs[i].element_size()
        items[i]["d_size"] = sizes.data_ptr() + 8 * i
        fill_item(items[i], tdt[i], ebs[i])
    b = Batch(items, True, dev, keep=list(streams) + list(outs) + [sizes])
    b.outs, b.sizes = outs, sizes
    return b


def compress_batch(xs, ebs, grid=0, stream=None):
    """Eager batched compress; returns (stream buffers, int64 sizes tensor, Batch)."""
    b = plan_compress(xs, ebs).launch(grid, stream)
    return b.outs, b.sizes, b


def decompress_batch(streams, dtypes, ebs, numels, grid=0, stream=None):
    """Eager batched decompress; returns (arrays, int64 decoded-byte sizes, Batch)."""
    b = plan_decompress(streams, dtypes, ebs, numels).launch(grid, stream)
    return b.outs, b.sizes, b
