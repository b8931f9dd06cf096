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
