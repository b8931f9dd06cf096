"""Torch-tensor front end of the HIP codec.

Device memory and streams come from PyTorch (plumbing only); all coding runs
in the HIP kernels of libsleek_amd.so through its C ABI.  Two layers:

* ``compress`` / ``decompress``: stream-ordered calls on the current torch
  stream with a reusable ``Workspace``; results stay on the device.
* ``compress_sync`` / ``decompress_sync``: the reference's synchronous
  calling convention (sleek_lossless_gpu.h:47-100, sleek_lossy_gpu.h:74-127)
  with library-owned scratch.
"""
import ctypes

import numpy as np
import torch

from . import _lib

CHUNK = 16384


def _elem_bytes(dtype):
    if dtype in (torch.float32, np.float32):
        return 4
    if dtype in (torch.float64, np.float64):
        return 8
    raise TypeError(f"SLEEK codes float32 and float64 only, not {dtype}")


def error_bound(eb, dtype):
    """(eb_e, thr_e, offs_e) of an absolute error bound (sleek_lossy_gpu.h:56-71)."""
    lib = _lib.load()
    e, t = ctypes.c_int(), ctypes.c_int()
    if _elem_bytes(dtype) == 4:
        o = ctypes.c_uint()
        rc = lib.sleek_gpu_error_bound_f32(float(np.float32(eb)), ctypes.byref(e), ctypes.byref(t), ctypes.byref(o))
    else:
        o = ctypes.c_ulonglong()
        rc = lib.sleek_gpu_error_bound_f64(float(eb), ctypes.byref(e), ctypes.byref(t), ctypes.byref(o))
    _lib.check(rc, "error_bound")
    return e.value, t.value, o.value


def upper_bound_size(nbytes):
    return _lib.load().sleek_gpu_upper_bound_size(int(nbytes))


def workspace_size(max_bytes):
    return _lib.load().sleek_gpu_workspace_size(int(max_bytes))


def thread_blocks():
    """Suggested persistent grid size for the current device."""
    n = _lib.load().sleek_gpu_thread_blocks()
    if n < 0:
        _lib.check(n, "thread_blocks")
    return n


class Workspace:
    """Device scratch of one in-flight codec call (scan words, ticket, status)."""

    def __init__(self, max_bytes, device=None):
        self.nbytes = int(workspace_size(max_bytes))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device or "cuda")

    def ensure(self, max_bytes):
        need = workspace_size(max_bytes)
        if need > self.nbytes:
            self.nbytes = int(need)
            self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=self.buf.device)
        return self

    def ptr(self):
        return self.buf.data_ptr()

    def status(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream()
        return _lib.load().sleek_gpu_workspace_status(self.ptr(), s.cuda_stream)

    def status_word(self):
        """The device status word (int64 view, one element) that the last call on
        this workspace left: its bits decode with status_of_word."""
        return self.buf[8:16].view(torch.int64)


def status_of_word(word):
    """Status code of a workspace status word (see sleek_gpu_workspace_status)."""
    if word & 2:
        return _lib.ECAPACITY
    if word & 1:
        return _lib.ECORRUPT
    return _lib.OK


def _stream_handle(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def compress(x, eb=None, *, out=None, encsize=None, workspace=None, blocks=0, stream=None):
    """Enqueue a compression of the contiguous float32/float64 CUDA tensor x.

    Returns (out, encsize): ``out`` is a uint8 buffer of the upper-bound size
    holding the stream, ``encsize`` a device int64 tensor with its length.
    ``eb`` is an absolute error bound (None = lossless)."""
    if not x.is_cuda or not x.is_contiguous():
        raise ValueError("x must be a contiguous CUDA tensor")
    eb_bytes = _elem_bytes(x.dtype)
    nbytes = x.numel() * eb_bytes
    ub = max(upper_bound_size(nbytes), 8)
    if out is None:
        out = torch.empty(ub, dtype=torch.uint8, device=x.device)
    elif out.numel() < ub:
        raise ValueError(f"output buffer holds {out.numel()} bytes, needs {ub}")
    if encsize is None:
        encsize = torch.empty(1, dtype=torch.int64, device=x.device)
    ws = (workspace or Workspace(nbytes, x.device)).ensure(nbytes)
    lossy, ebi = (0, (0, 0, 0)) if eb is None else (1, error_bound(eb, x.dtype))
    rc = _lib.load().sleek_gpu_compress_async(eb_bytes, lossy, ebi[0], ebi[1], ebi[2], x.data_ptr(), nbytes,
                                              out.data_ptr(), encsize.data_ptr(), ws.ptr(), ws.nbytes, int(blocks),
                                              _stream_handle(stream))
    _lib.check(rc, "compress")
    return out, encsize


def decompress(stream_buf, dtype, eb=None, *, numel=None, out=None, decsize=None, workspace=None, blocks=0,
               enclen=-1, stream=None):
    """Enqueue a decompression of the uint8 CUDA tensor ``stream_buf``.

    ``numel`` (or ``out``) gives the capacity; without either the header is
    read back to size the output.  Returns (out, decsize)."""
    eb_bytes = _elem_bytes(dtype)
    tdt = torch.float32 if eb_bytes == 4 else torch.float64
    if out is None:
        if numel is None:
            numel = int(stream_buf[:8].cpu().numpy().view(np.int64)[0]) // eb_bytes
        out = torch.empty(max(int(numel), 0), dtype=tdt, device=stream_buf.device)
    cap = out.numel() * eb_bytes
    if decsize is None:
        decsize = torch.empty(1, dtype=torch.int64, device=stream_buf.device)
    ws = (workspace or Workspace(cap, stream_buf.device)).ensure(cap)
    thr, offs = (0, 0) if eb is None else error_bound(eb, dtype)[1:]
    rc = _lib.load().sleek_gpu_decompress_async(eb_bytes, 0 if eb is None else 1, thr, offs, stream_buf.data_ptr(),
                                                int(enclen), out.data_ptr(), cap, decsize.data_ptr(), ws.ptr(),
                                                ws.nbytes, int(blocks), _stream_handle(stream))
    _lib.check(rc, "decompress")
    return out, decsize


def compress_sync(x, eb=None, blocks=0):
    """Synchronous compression (library-owned scratch); returns the trimmed stream tensor."""
    lib = _lib.load()
    eb_bytes = _elem_bytes(x.dtype)
    out = torch.empty(max(upper_bound_size(x.numel() * eb_bytes), 8), dtype=torch.uint8, device=x.device)
    encsize = torch.zeros(1, dtype=torch.int64, device=x.device)
    torch.cuda.synchronize(x.device)
    if eb is None:
        fn = lib.sleek_gpu_compress_f32 if eb_bytes == 4 else lib.sleek_gpu_compress_f64
        rc = fn(int(blocks), x.data_ptr(), x.numel(), out.data_ptr(), encsize.data_ptr())
    else:
        e, t, o = error_bound(eb, x.dtype)
        fn = lib.sleek_gpu_compress_lossy_f32 if eb_bytes == 4 else lib.sleek_gpu_compress_lossy_f64
        rc = fn(int(blocks), e, t, o, x.data_ptr(), x.numel(), out.data_ptr(), encsize.data_ptr())
    _lib.check(rc, "compress_sync")
    return out[: int(encsize.item())]


def decompress_sync(stream_buf, dtype, eb=None, blocks=0):
    """Synchronous decompression (library-owned scratch); returns the array."""
    lib = _lib.load()
    eb_bytes = _elem_bytes(dtype)
    tdt = torch.float32 if eb_bytes == 4 else torch.float64
    hdr = int(stream_buf[:8].cpu().numpy().view(np.int64)[0])
    out = torch.empty(max(hdr // eb_bytes, 1), dtype=tdt, device=stream_buf.device)
    decsize = torch.zeros(1, dtype=torch.int64, device=stream_buf.device)
    torch.cuda.synchronize(stream_buf.device)
    if eb is None:
        fn = lib.sleek_gpu_decompress_f32 if eb_bytes == 4 else lib.sleek_gpu_decompress_f64
        rc = fn(int(blocks), stream_buf.data_ptr(), out.data_ptr(), decsize.data_ptr())
    else:
        e, t, o = error_bound(eb, dtype)
        fn = lib.sleek_gpu_decompress_lossy_f32 if eb_bytes == 4 else lib.sleek_gpu_decompress_lossy_f64
        rc = fn(int(blocks), e, t, o, stream_buf.data_ptr(), out.data_ptr(), decsize.data_ptr())
    _lib.check(rc, "decompress_sync")
    return out[: hdr // eb_bytes]
