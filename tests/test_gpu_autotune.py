"""The autotuner on the GPU: every candidate launch shape gives the same bytes
(the tuner raises otherwise), and the chosen one is applied."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def test_tune_small_field(gpu, tmp_path):
    import torch
    from sleek_amd import autotune, codec
    from sleek_amd.datagen import exact_field

    print("autotune 8 MiB float32", flush=True)
    sys.stdout.flush()
    a = exact_field(2 << 20, 17, np.float32)
    x = torch.from_numpy(a).cuda()
    try:
        best, trace = autotune.tune(x, 1e-3, reps=2, cache=True, path=str(tmp_path / "t.json"))
        n_cand = 1 + (len(autotune.ENCODE_SHAPES) - 1) + len(autotune.LOOKBACK) + len(autotune.DECODE_SHAPES)
        assert len(trace) >= n_cand - 2
        assert set(best) == set(autotune.DEFAULT)
        # cached: a second call applies without searching
        best2, trace2 = autotune.tune(x, 1e-3, cache=True, path=str(tmp_path / "t.json"))
        assert best2 == best and trace2 == []
        out, n = codec.compress(x, 1e-3)
        torch.cuda.synchronize()
        assert out[: int(n.item())].cpu().numpy().tobytes() == po.compress(a, 1e-3).tobytes()
    finally:
        autotune.apply(autotune.DEFAULT)


def test_tune_mixed_modes(gpu, tmp_path):
    """A list of bounds is tuned as one mix (the benchmark's four modes) and
    cached under its own key; the bytes of every mode stay the oracle's."""
    import torch
    from sleek_amd import autotune, codec
    from sleek_amd.datagen import exact_field

    print("autotune 4 MiB float32, lossless + 2 lossy modes", flush=True)
    sys.stdout.flush()
    a = exact_field(1 << 20, 23, np.float32)
    x = torch.from_numpy(a).cuda()
    ebs = [None, 1e-2, 1e-4]
    path = str(tmp_path / "t.json")
    try:
        best, trace = autotune.tune(x, ebs, reps=1, cache=True, path=path)
        assert trace and set(best) == set(autotune.DEFAULT)
        assert any("|mixed3|" in k for k in autotune.load_cache(path))
        for eb in ebs:
            out, n = codec.compress(x, eb)
            torch.cuda.synchronize()
            assert out[: int(n.item())].cpu().numpy().tobytes() == po.compress(a, eb).tobytes()
    finally:
        autotune.apply(autotune.DEFAULT)
