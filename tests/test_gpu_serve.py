"""The serving loop on the GPU: requests that arrive while others run are
batched and overlap on the device, and every result is byte-identical to the
same request served alone (a single-array call) and to the oracle."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_results_same_alone_and_in_the_mix(gpu):
    import torch
    from sleek_amd import codec
    from sleek_amd.datagen import exact_field, inject_specials
    from sleek_amd.serve import COMPRESS, DECOMPRESS, GpuEngine, Request, Server, poisson_arrivals, summarize

    rng = np.random.default_rng(4)
    n = 160
    say(f"serving {n} mixed requests")
    arrivals = poisson_arrivals(n, 50000.0, 4)      # far faster than batches finish: they mix
    reqs, host = [], []
    for i in range(n):
        dt = np.float32 if rng.random() < 0.6 else np.float64
        m = int(np.exp(rng.uniform(np.log(16), np.log(600_000))))
        a = exact_field(m, 500 + i, dt)
        if m > 2000:
            inject_specials(a, 1e-3, i)
        eb = [None, 1e-2, 1e-4][int(rng.integers(0, 3))]
        x = torch.from_numpy(a).cuda()
        if rng.random() < 0.5:
            reqs.append(Request(i, COMPRESS, x, dt, eb, arrival=float(arrivals[i])))
        else:
            s = torch.from_numpy(po.compress(a, eb)).cuda()
            reqs.append(Request(i, DECOMPRESS, s, dt, eb, numel=m, arrival=float(arrivals[i])))
        host.append(a)
    torch.cuda.synchronize()
    srv = Server(GpuEngine(streams=3), budget_bytes=64 << 20, max_inflight=3)
    srv.run(reqs)
    torch.cuda.synchronize()
    stats = summarize(reqs, srv)
    say(str(stats))
    assert srv.max_concurrent > 1 and max(srv.batch_sizes) > 1
    assert srv.peak_used <= 64 << 20
    for r, a in zip(reqs, host):
        assert r.status == 0, r.rid
        if r.kind == COMPRESS:
            mix = r.out.cpu().numpy()
            alone, k = codec.compress(r.src, r.eb)
            torch.cuda.synchronize()
            assert mix.tobytes() == alone[: int(k.item())].cpu().numpy().tobytes(), r.rid
            assert mix.tobytes() == po.compress(a, r.eb).tobytes(), r.rid
        else:
            mix = r.out.cpu().numpy()
            alone, _ = codec.decompress(r.src, r.out.dtype, r.eb, numel=r.numel)
            torch.cuda.synchronize()
            assert mix.tobytes() == alone.cpu().numpy().tobytes(), r.rid
            od, rc = po.decompress(r.src.cpu().numpy(), a.dtype, r.eb)
            assert rc == 0 and mix.tobytes() == od.tobytes(), r.rid
