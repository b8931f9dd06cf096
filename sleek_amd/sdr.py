"""Reading scientific-data files of the kind the reference's harnesses take
(SDRBench: raw little-endian float32 or float64 values, one array per file,
any dimensions; download_inputs_*.py, compress_decompress_inputs_*.py:144-149
walk a directory tree and code every file).

``list_inputs`` walks a tree the same way; ``load`` memory-maps one file;
``to_device`` moves it to the GPU through pinned memory in slabs, so a large
file is never read into pageable memory whole.
"""
import glob
import os

import numpy as np

SUFFIX_DTYPE = {".f32": np.float32, ".bin": np.float32, ".dat": np.float32, ".d64": np.float64, ".f64": np.float64}


def list_inputs(directory):
    """Every regular file under ``directory`` (recursively), sorted."""
    return sorted(p for p in glob.glob(os.path.join(directory, "**"), recursive=True) if os.path.isfile(p))


def guess_dtype(path, default=np.float32):
    """The element type a file name suggests (.f32 / .d64 and the like)."""
    for suffix, dt in SUFFIX_DTYPE.items():
        if path.endswith(suffix):
            return dt
    return default


def load(path, dtype=None, shape=None):
    """The file's values as a read-only memory map.  ValueError if its size is
    not a multiple of the element size or does not match ``shape``."""
    dt = np.dtype(dtype or guess_dtype(path))
    size = os.path.getsize(path)
    if size % dt.itemsize:
        raise ValueError(f"{path}: {size} bytes is not a multiple of {dt.itemsize}")
    n = size // dt.itemsize
    if shape is not None and int(np.prod(shape)) != n:
        raise ValueError(f"{path}: {n} values do not fill shape {tuple(shape)}")
    if n == 0:
        return np.zeros(0 if shape is None else shape, dtype=dt)
    return np.memmap(path, dtype=dt, mode="r", shape=tuple(shape) if shape is not None else (n,))


def to_device(path, dtype=None, device="cuda", slab_bytes=256 << 20):
    """The file as a contiguous CUDA tensor, copied through a pinned slab."""
    import torch

    a = load(path, dtype).reshape(-1)
    tdt = torch.float32 if a.dtype == np.float32 else torch.float64
    out = torch.empty(a.size, dtype=tdt, device=device)
    per = max(slab_bytes // a.dtype.itemsize, 1)
    slab = torch.empty(min(per, max(a.size, 1)), dtype=tdt, pin_memory=True)
    for o in range(0, a.size, per):
        n = min(per, a.size - o)
        slab[:n].numpy()[:] = a[o: o + n]
        out[o: o + n].copy_(slab[:n], non_blocking=True)
        torch.cuda.current_stream(out.device).synchronize()   # the slab is reused
    return out
