"""The no-allocation compression entry point (sleek_gpu_compress_noalloc):
through the C ABI, two calls in a row, the second on a smaller input, each
byte-identical to sleek_gpu_compress_async with a caller workspace (and to
the oracle); and the size limit is refused before any device work."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_noalloc_matches_async_on_two_calls(gpu):
    import torch
    from sleek_amd import _lib, codec
    from sleek_amd.datagen import exact_field, inject_specials

    lib = _lib.load()
    stream = torch.cuda.current_stream()
    cases = [(np.float32, 1e-3, 4096 * 300 + 11), (np.float64, None, 2048 * 37 + 5), (np.float32, None, 3)]
    for dt, eb, n in cases:   # sizes shrink from call to call
        say(f"no-alloc compress {np.dtype(dt).name} eb={eb} n={n}")
        a = exact_field(n, 23, dt)
        if n > 1000:
            inject_specials(a, 1e-3, 4)
        x = torch.from_numpy(a).cuda()
        nbytes = a.nbytes
        ub = codec.upper_bound_size(nbytes)
        e, t, o = (0, 0, 0) if eb is None else codec.error_bound(eb, x.dtype)
        lossy = int(eb is not None)
        outs = []
        for use_noalloc in (True, False):
            out = torch.full((ub,), 0xA5, dtype=torch.uint8, device=x.device)
            sz = torch.full((1,), -1, dtype=torch.int64, device=x.device)
            if use_noalloc:
                rc = lib.sleek_gpu_compress_noalloc(a.itemsize, lossy, e, t, o, x.data_ptr(), nbytes, out.data_ptr(),
                                                    sz.data_ptr(), 0, stream.cuda_stream)
            else:
                ws = codec.Workspace(nbytes, x.device)
                rc = lib.sleek_gpu_compress_async(a.itemsize, lossy, e, t, o, x.data_ptr(), nbytes, out.data_ptr(),
                                                  sz.data_ptr(), ws.ptr(), ws.nbytes, 0, stream.cuda_stream)
            _lib.check(rc, "compress")
            torch.cuda.synchronize()
            k = int(sz.item())
            outs.append(out[:k].cpu().numpy())
        assert outs[0].size == outs[1].size and np.array_equal(outs[0], outs[1])
        assert np.array_equal(outs[0], po.compress(a, eb))


def test_noalloc_refuses_inputs_above_its_maximum(gpu):
    import torch
    from sleek_amd import _lib

    lib = _lib.load()
    x = torch.zeros(4, dtype=torch.float32, device="cuda")
    out = torch.zeros(64, dtype=torch.uint8, device="cuda")
    sz = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = lib.sleek_gpu_compress_noalloc(4, 0, 0, 0, 0, x.data_ptr(), (4 << 30) + 4, out.data_ptr(), sz.data_ptr(), 0,
                                        None)
    assert rc == _lib.ETOOLARGE
    assert lib.sleek_gpu_compress_noalloc(4, 0, 0, 0, 0, x.data_ptr(), -4, out.data_ptr(), sz.data_ptr(), 0,
                                          None) == _lib.EINVAL
