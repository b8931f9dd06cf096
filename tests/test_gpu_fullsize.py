"""Parity at the benchmark workloads' full sizes.

The reference's verify harness compresses and decompresses every input with
the GPU and the CPU codecs and diffs both the compressed and the decompressed
files, lossless and at six bounds (verify_cpu_gpu_single.py:89-104,
verify_cpu_gpu_double.py).  Here the same comparison runs at the sizes the
benchmark times:

* the 256^3 float32 field of the CPU-path workload, lossless;
* the 512^3 float32 field with ramp and specials, lossless and eb = {1e-2,
  1e-3, 1e-4} x range;
* the 384^3 float64 field with lognormal noise and specials, lossless and
  eb = {1e-6, 1e-9} x range;
* the 4,096-array uneven batch (about 29 GB) in one batched launch per
  direction: every array's stream and decoded bytes;
* the serving workload (2,000 Poisson requests, 64 KiB .. 256 MiB): every
  result equals its single call, and a size-stratified sample equals the
  oracle.

Every GPU stream must equal the oracle's stream of the same host array, and
the reference CPU codec's where oracle/_ref is built; every GPU decode must
equal the oracle's decode (lossless: the input itself); lossy decodes must
pass the reference's point-wise check (MAXABS_f32.h:40-95).  All GPU calls go
through the C ABI of libsleek_amd.so.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def maxabs_ok(orig, recon, eb):
    """The reference's point-wise check (MAXABS_f32.h:56-81) in the data's precision."""
    o, r = orig, recon
    fin = np.isfinite(o) & np.isfinite(r)
    same = (o == r) | (np.isnan(o) & np.isnan(r))
    if np.any(~fin & ~same):
        return False
    eb = o.dtype.type(eb)
    of, rf = o[fin], r[fin]
    with np.errstate(over="ignore", invalid="ignore"):
        bad = (rf < of - eb) | (rf > of + eb) | (np.abs(of - rf) > eb)
    return not np.any(bad)


def check_array(x, eb, label):
    """One GPU compress + decompress of the CUDA tensor x against the oracle
    (and the compiled reference codec when present) on the same host bytes."""
    import torch
    from sleek_amd import codec

    a = x.cpu().numpy()
    nbytes = a.nbytes
    ws = codec.Workspace(nbytes, x.device)
    out, n = codec.compress(x, eb, workspace=ws)
    assert ws.status() == 0
    k = int(n.item())
    stream = out[:k].cpu().numpy()
    ref = po.compress(a, eb)
    assert stream.size == ref.size, f"{label}: stream length {stream.size} != oracle {ref.size}"
    assert np.array_equal(stream, ref), f"{label}: stream differs from the oracle at byte " \
        f"{int(np.argmax(stream != ref))}"
    if po.ref() is not None:
        assert np.array_equal(stream, po.ref_compress(a, eb)), f"{label}: stream differs from the reference codec"
    d, dn = codec.decompress(out, x.dtype, eb, numel=x.numel(), workspace=ws, enclen=k)
    assert ws.status() == 0
    assert int(dn.item()) == nbytes
    dec = d.cpu().numpy()
    od, rc = po.decompress(ref, a.dtype, eb)
    assert rc == 0
    assert dec.tobytes() == od.tobytes(), f"{label}: decoded array differs from the oracle's"
    if eb is None:
        assert dec.tobytes() == a.tobytes(), f"{label}: lossless round trip"
    else:
        assert maxabs_ok(a, dec, eb), f"{label}: point-wise error bound"
    del out, d
    torch.cuda.empty_cache()
    return nbytes / k


def test_cpu_path_field_256cubed_lossless(gpu):
    """256^3 float32, 8 sinusoids + N(0, 1e-3): the CPU-path workload, lossless."""
    from sleek_amd.datagen import config_field

    x = config_field(0, seed=1, device=gpu).reshape(-1).contiguous()
    say(f"256^3 float32 lossless ({x.numel() * 4 >> 20} MiB)")
    ratio = check_array(x, None, "256^3 lossless")
    say(f"  ratio {ratio:.3f}")


@pytest.mark.parametrize("rel", [None, 1e-2, 1e-3, 1e-4])
def test_lossy_field_512cubed(gpu, rel):
    """512^3 float32 with ramp and 0.01 % specials: lossless and the three bounds."""
    from sleek_amd.datagen import config_field, finite_range

    x = config_field(1, seed=1, device=gpu).reshape(-1).contiguous()
    eb = None if rel is None else float(np.float32(np.float32(finite_range(x)) * np.float32(rel)))
    say(f"512^3 float32 eb={eb} ({'lossless' if rel is None else f'{rel:g} x range'})")
    ratio = check_array(x, eb, f"512^3 f32 rel={rel}")
    say(f"  ratio {ratio:.3f}")


@pytest.mark.parametrize("rel", [None, 1e-6, 1e-9])
def test_double_field_384cubed(gpu, rel):
    """384^3 float64, 12 sinusoids + lognormal noise + 0.01 % specials: lossless and two bounds."""
    from sleek_amd.datagen import config_field, finite_range

    x = config_field(2, seed=1, device=gpu).reshape(-1).contiguous()
    eb = None if rel is None else float(np.float64(finite_range(x)) * np.float64(rel))
    say(f"384^3 float64 eb={eb} ({'lossless' if rel is None else f'{rel:g} x range'})")
    ratio = check_array(x, eb, f"384^3 f64 rel={rel}")
    say(f"  ratio {ratio:.3f}")


BATCH_SHARDS = 4


@pytest.fixture(scope="module")
def uneven_batch(gpu):
    """The 4,096-array batch of the batched workload, compressed and decompressed
    in one batched launch per direction; results kept on the device."""
    import torch
    from sleek_amd.batch import plan_compress, plan_decompress
    from sleek_amd.datagen import array_field, batch_sizes, finite_range

    specs = batch_sizes(4096, seed=1 + 77)
    rng = np.random.default_rng(1 + 78)
    xs, ebs = [], []
    for i, (nb, dt) in enumerate(specs):
        x = array_field(nb, dt, seed=100003 + i, device=gpu)
        xs.append(x)
        if rng.random() < 0.5:
            ebs.append(None)
        else:
            rel = float(rng.choice([1e-2, 1e-3, 1e-4]))
            r = finite_range(x)
            ebs.append(float(np.dtype(dt).type(np.dtype(dt).type(r) * np.dtype(dt).type(rel))) if r > 0 else rel)
    say(f"uneven batch: {len(xs)} arrays, {sum(x.numel() * x.element_size() for x in xs) / 1e9:.2f} GB")
    bc = plan_compress(xs, ebs).launch()
    torch.cuda.synchronize()
    assert not bc.status().any(), "batched compress status"
    sizes = bc.sizes.cpu().numpy()[: len(xs)]
    streams = [bc.outs[i][: int(sizes[i])] for i in range(len(xs))]
    bd = plan_decompress(streams, [x.dtype for x in xs], ebs, [x.numel() for x in xs]).launch()
    torch.cuda.synchronize()
    assert not bd.status().any(), "batched decompress status"
    dsz = bd.sizes.cpu().numpy()[: len(xs)]
    yield xs, ebs, streams, bd.outs, dsz
    del xs, streams, bc, bd
    torch.cuda.empty_cache()


@pytest.mark.parametrize("shard", range(BATCH_SHARDS))
def test_uneven_batch_every_array(gpu, uneven_batch, shard):
    """Arrays i = shard (mod 4) of the batch: stream and decoded bytes equal the oracle's."""
    xs, ebs, streams, decs, dsz = uneven_batch
    idx = range(shard, len(xs), BATCH_SHARDS)
    say(f"uneven batch shard {shard}: {len(idx)} arrays")
    for i in idx:
        a = xs[i].cpu().numpy()
        eb = ebs[i]
        s = streams[i].cpu().numpy()
        ref = po.compress(a, eb)
        assert np.array_equal(s, ref), (i, a.size, a.dtype, eb)
        od, rc = po.decompress(ref, a.dtype, eb)
        assert rc == 0 and int(dsz[i]) == a.nbytes, (i, int(dsz[i]))
        d = decs[i].cpu().numpy()
        assert d.tobytes() == od.tobytes(), (i, a.size, a.dtype, eb)
        if eb is None:
            assert d.tobytes() == a.tobytes(), i
        elif a.size:
            assert maxabs_ok(a, d, eb), i


def test_serving_workload_every_result(gpu):
    """2,000 Poisson requests (lambda = 2,000/s) of the serving workload under a 96 GB
    budget: every compress result equals the same request compressed alone and every
    decompress result the same stream decompressed alone; one request per size decile
    of each kind also equals the oracle."""
    import torch
    from sleek_amd import codec
    from sleek_amd.batch import plan_compress
    from sleek_amd.datagen import array_field, finite_range
    from sleek_amd.serve import COMPRESS, DECOMPRESS, GpuEngine, Request, Server, poisson_arrivals

    rng = np.random.default_rng(7)
    hi = 256 << 20
    pools = {}
    for dt in (np.float32, np.float64):
        pools[dt] = []
        for k in range(3):
            f = array_field(hi, dt, seed=9000 + k + (100 if dt == np.float64 else 0), device=gpu)
            pools[dt].append((f, finite_range(f)))
    n = 2000
    arrivals = poisson_arrivals(n, 2000.0, 8)
    reqs, pre_x, pre_eb, pre_idx = [], [], [], []
    for i in range(n):
        dt = np.float32 if rng.random() < 0.7 else np.float64
        it = np.dtype(dt).itemsize
        nb = int(np.exp(rng.uniform(np.log(64 << 10), np.log(hi))))
        m = max(1, nb // it)
        f, r = pools[dt][int(rng.integers(0, 3))]
        x = f[:m]
        eb = None
        if rng.random() < 0.5:
            eb = float(np.dtype(dt).type(np.dtype(dt).type(r) * np.dtype(dt).type(float(rng.choice([1e-2, 1e-3, 1e-4])))))
        if rng.random() < 0.5:
            reqs.append(Request(i, COMPRESS, x, dt, eb, arrival=float(arrivals[i])))
        else:
            reqs.append(Request(i, DECOMPRESS, None, dt, eb, numel=m, arrival=float(arrivals[i])))
            pre_x.append(x)
            pre_eb.append(eb)
            pre_idx.append(i)
    pb = plan_compress(pre_x, pre_eb).launch()
    torch.cuda.synchronize()
    psz = pb.sizes.cpu().numpy()
    for j, i in enumerate(pre_idx):
        reqs[i].src = pb.outs[j][: int(psz[j])]
    say(f"serving: {n} requests prepared")
    srv = Server(GpuEngine(gpu, streams=4), budget_bytes=96e9, max_inflight=4)
    srv.run(reqs)
    torch.cuda.synchronize()
    assert all(r.status == 0 for r in reqs), [r.rid for r in reqs if r.status != 0][:8]
    say(f"serving: {srv.batches} batches, mean batch {np.mean(srv.batch_sizes):.2f}, "
        f"max in flight {srv.max_concurrent}")
    ws = codec.Workspace(hi, gpu)
    for r in reqs:
        if r.kind == COMPRESS:
            alone, k = codec.compress(r.src, r.eb, workspace=ws)
            k = int(k.item())
            assert r.size == k and torch.equal(r.out, alone[:k]), f"request {r.rid} differs from its single call"
        else:
            tdt = torch.float32 if np.dtype(r.dtype) == np.float32 else torch.float64
            alone, _ = codec.decompress(r.src, tdt, r.eb, numel=r.numel, workspace=ws)
            torch.cuda.synchronize()
            assert torch.equal(r.out.view(torch.uint8), alone.view(torch.uint8)), f"request {r.rid} differs"
        assert ws.status() == 0
    say("serving: every result equals its single call")
    # size-stratified sample against the oracle: per kind, the request nearest each size decile
    for kind in (COMPRESS, DECOMPRESS):
        group = sorted((r for r in reqs if r.kind == kind), key=lambda r: r.numel if kind == DECOMPRESS else r.in_bytes)
        picks = {group[min(len(group) - 1, int(q * len(group)))].rid for q in np.linspace(0, 1, 10)}
        for rid in sorted(picks):
            r = reqs[rid]
            if kind == COMPRESS:
                a = r.src.cpu().numpy()
                assert np.array_equal(r.out.cpu().numpy(), po.compress(a, r.eb)), rid
            else:
                od, rc = po.decompress(r.src.cpu().numpy(), r.dtype, r.eb)
                assert rc == 0 and r.out.cpu().numpy().tobytes() == od.tobytes(), rid
    say("serving: size-stratified sample equals the oracle")
    del pools, reqs, pb
    torch.cuda.empty_cache()
