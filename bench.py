"""Benchmark of the SLEEK HIP codec on one MI355X.

Workload (default): the 512^3 float32 field of the lossy benchmark
configuration (8 random sinusoids + diagonal ramp U[-1e3, 1e3] + N(0, 1e-3)
noise, 0.01 % NaN/Inf/denormal/-0 values), resident in HBM.  One step
compresses and then decompresses it in each of four modes: lossless and
absolute error bounds {1e-2, 1e-3, 1e-4} x (max - min over finite values).

Printed (one JSON line): whole-step throughput in GB/s (input bytes per
direction / time, the reference's insize/t definition), per-kernel effective
and algorithmic GB/s, the roofline entry of the dominant kernel (measured
with HIP events on the launch stream), a hipMemcpy device-to-device
comparator of the same bytes, and the reference CPU codec on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
MODES = (None, 1e-2, 1e-3, 1e-4)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--scale", type=float, default=1.0, help="shrink every axis (tests)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--encode-variant", type=int, default=0,
                   help="0 default (6), 1 pipelined 256-thread, 2 the same at the compiler's register choice, "
                        "4 one wave per chunk pipelined over its chunks, 5 two waves per chunk, 6 two waves per "
                        "chunk with chunk tickets, 7 one wave per chunk with chunk tickets (3 is retired)")
    p.add_argument("--encode-grid", type=int, default=0, help="encoder persistent grid (0 = automatic)")
    p.add_argument("--decode-grid", type=int, default=0, help="decoder grid (0 = one workgroup per chunk)")
    p.add_argument("--decoder", type=int, default=-1, help="decoder variant (-1 = library default)")
    p.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    p.add_argument("--no-f64", action="store_true", help="skip the float64 workload")
    p.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    p.add_argument("--no-transfer", action="store_true", help="skip the compressing host<->device copies")
    p.add_argument("--batch-arrays", type=int, default=4096)
    p.add_argument("--batch-max-mb", type=float, default=64.0, help="largest array of the batched workload")
    p.add_argument("--batch-steps", type=int, default=3)
    p.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    p.add_argument("--serve-requests", type=int, default=2000)
    p.add_argument("--serve-rates", default="2000,10000,50000,200000",
                   help="Poisson arrival rates (requests/s) of the serving sweep; the first is the headline rate "
                        "(the last offers more than the GPU serves: its saturation point)")
    p.add_argument("--serve-max-mb", type=float, default=256.0)
    p.add_argument("--serve-budget-gb", type=float, default=96.0)
    return p.parse_args()


def cpu_baseline(fields):
    """The reference CPU codec (oracle/_ref, compiled from the reference's own
    headers, OpenMP on every host core) if built, else the oracle
    restatement, on whole benchmark fields.  fields: (label, host array,
    [(mode label, absolute eb or None)]).  Per mode: one warm-up call, then
    the median of 3 timed calls per direction (steady clock around the codec
    call only, as sleek_single_compressor.cpp:88-96).  Rates are insize / t."""
    from oracle import pyoracle as po

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    use_ref = po.ref() is not None
    rows = []
    for label, a, modes in fields:
        a = np.ascontiguousarray(a)
        for mode, eb in modes:
            comp = (lambda x: po.ref_compress(x, eb)) if use_ref else (lambda x: po.compress(x, eb))
            dec = (lambda s: po.ref_decompress(s, a.dtype, eb)) if use_ref else \
                (lambda s: po.decompress(s, a.dtype, eb)[0])
            st = comp(a)
            dec(st)  # warm-up
            tc, td = [], []
            for _ in range(3):
                t0 = time.perf_counter()
                st = comp(a)
                t1 = time.perf_counter()
                dec(st)
                t2 = time.perf_counter()
                tc.append(t1 - t0)
                td.append(t2 - t1)
            tcm, tdm = float(np.median(tc)), float(np.median(td))
            rows.append({"field": label, "mode": mode, "bytes": a.nbytes, "ratio": round(a.nbytes / st.size, 4),
                         "compress_GBps": round(a.nbytes / tcm / 1e9, 3),
                         "decompress_GBps": round(a.nbytes / tdm / 1e9, 3)})
            print(f"cpu baseline: {label} {mode}: {rows[-1]['compress_GBps']} / {rows[-1]['decompress_GBps']} GB/s",
                  file=sys.stderr, flush=True)
    head = rows[0]
    return {
        "value": round(2 / (1 / head["compress_GBps"] + 1 / head["decompress_GBps"]), 3),
        "unit": "GB/s",
        "cores": cores,
        "kind": "reference" if use_ref else "port",
        "sample": f"whole fields: {', '.join(sorted(set(r['field'] for r in rows), key=[f[0] for f in fields].index))}; "
                  "value = the 256^3 float32 field lossless, compress+decompress (2 x insize / total time); one "
                  f"warm-up then the median of 3 calls per direction; OMP threads="
                  f"{os.environ.get('OMP_NUM_THREADS', cores)}",
        "compress_GBps": head["compress_GBps"],
        "decompress_GBps": head["decompress_GBps"],
        "rows": rows,
    }


def bench_batch(args, dev):
    """The uneven batch: log-uniform 4 KiB .. 64 MiB arrays, 70 % float32, half
    lossless and half lossy at {1e-2, 1e-3, 1e-4} x range.  One graph-captured
    batched launch per direction against per-array plain launches (the async
    C ABI in a loop on one stream, each array with its own workspace, from a
    C++ loop and from a Python loop); both forms produce the same bytes,
    which is checked.  Rates are aggregate input bytes / time."""
    import ctypes

    import torch

    from sleek_amd import _lib, codec
    from sleek_amd.batch import ITEM, fill_item, plan_compress, plan_decompress
    from sleek_amd.datagen import array_field, batch_sizes, finite_range

    lib = _lib.load()
    t_gen = time.perf_counter()
    specs = batch_sizes(args.batch_arrays, seed=args.seed + 77, hi=int(args.batch_max_mb * (1 << 20)))
    rng = np.random.default_rng(args.seed + 78)
    xs, ebs = [], []
    for i, (nb, dt) in enumerate(specs):
        x = array_field(nb, dt, seed=args.seed * 100003 + i, device=dev)
        xs.append(x)
        if rng.random() < 0.5:
            ebs.append(None)
        else:
            rel = float(rng.choice([1e-2, 1e-3, 1e-4]))
            r = finite_range(x)
            ebs.append(float(np.dtype(dt).type(np.dtype(dt).type(r) * np.dtype(dt).type(rel))) if r > 0 else rel)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    total = sum(x.numel() * x.element_size() for x in xs)
    print(f"batched workload: {len(xs)} arrays, {total / 1e9:.2f} GB generated in {t_gen:.1f} s",
          file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    h = stream.cuda_stream

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a0.record(stream)
            fn()
            a1.record(stream)
            torch.cuda.synchronize()
            ts.append(a0.elapsed_time(a1) / 1e3)
        return float(np.median(ts))

    # batched, graph-captured
    bc = plan_compress(xs, ebs)
    gc = bc.capture()
    gc.launch()
    torch.cuda.synchronize()
    assert not bc.status().any(), "batched compress status"
    sizes = bc.sizes.cpu().numpy()[: len(xs)]
    streams = [bc.outs[i][: int(sizes[i])] for i in range(len(xs))]
    bd = plan_decompress(streams, [x.dtype for x in xs], ebs, [x.numel() for x in xs])
    gd = bd.capture()
    gd.launch()
    torch.cuda.synchronize()
    assert not bd.status().any(), "batched decompress status"
    print("batched workload: graphs captured and checked", file=sys.stderr, flush=True)
    t_bc = timed(gc.launch, args.batch_steps)
    t_bd = timed(gd.launch, args.batch_steps)
    t_bc_eager = timed(lambda: bc.launch(), args.batch_steps)
    t_bd_eager = timed(lambda: bd.launch(), args.batch_steps)

    # per-array plain launches through the C ABI, each call with its own
    # workspace (256-byte aligned slices of one buffer) so that every call's
    # status is read back: from a Python ctypes loop and from a C++ loop
    # (sleek_amd/cli/bin/libsleek_plain_each.so)
    pouts = [torch.empty_like(o) for o in bc.outs]
    psz = torch.zeros(len(xs), dtype=torch.int64, device=dev)
    pdec = [torch.empty_like(x) for x in xs]
    pdsz = torch.zeros(len(xs), dtype=torch.int64, device=dev)
    wsb = np.array([codec.workspace_size(x.numel() * x.element_size()) for x in xs], dtype=np.int64)
    wsoff = np.concatenate([[0], np.cumsum((wsb + 255) // 256 * 256)])
    wsbuf = torch.empty(int(wsoff[-1]), dtype=torch.uint8, device=dev)
    wsptr = np.array([wsbuf.data_ptr() + int(o) for o in wsoff[:-1]], dtype=np.uint64)
    citems = np.zeros(len(xs), dtype=ITEM)
    ditems = np.zeros(len(xs), dtype=ITEM)
    cargs, dargs = [], []
    for i, (x, eb) in enumerate(zip(xs, ebs)):
        eb_b = x.element_size()
        e, t, o = (0, 0, 0) if eb is None else codec.error_bound(eb, x.dtype)
        nb = x.numel() * eb_b
        cargs.append((eb_b, int(eb is not None), e, t, o, x.data_ptr(), nb, pouts[i].data_ptr(),
                      psz.data_ptr() + 8 * i, int(wsptr[i]), int(wsb[i]), 0, h))
        dargs.append((eb_b, int(eb is not None), t, o, streams[i].data_ptr(), int(sizes[i]), pdec[i].data_ptr(), nb,
                      pdsz.data_ptr() + 8 * i, int(wsptr[i]), int(wsb[i]), 0, h))
        for rec, src, dst, sb, db, dsz_p in ((citems[i], x.data_ptr(), pouts[i].data_ptr(), nb, pouts[i].numel(),
                                              psz.data_ptr() + 8 * i),
                                             (ditems[i], streams[i].data_ptr(), pdec[i].data_ptr(), int(sizes[i]), nb,
                                              pdsz.data_ptr() + 8 * i)):
            rec["src"], rec["dst"], rec["src_bytes"], rec["dst_bytes"], rec["d_size"] = src, dst, sb, db, dsz_p
            fill_item(rec, x.dtype, eb)
    comp, decomp = lib.sleek_gpu_compress_async, lib.sleek_gpu_decompress_async
    each = ctypes.CDLL(os.path.join(ROOT, "sleek_amd", "cli", "bin", "libsleek_plain_each.so")).sleek_plain_each
    each.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    wsb_c = np.ascontiguousarray(wsb)
    status_idx = torch.from_numpy((wsoff[:-1] // 8 + 1).astype(np.int64)).to(dev)

    def statuses():
        torch.cuda.synchronize()
        return wsbuf.view(torch.int64)[status_idx].cpu().numpy()

    def plain_c():
        for a in cargs:
            comp(*a)

    def plain_d():
        for a in dargs:
            decomp(*a)

    def plain_c_native():
        _lib.check(each(0, citems.ctypes.data, len(xs), wsptr.ctypes.data, wsb_c.ctypes.data, h), "plain loop")

    def plain_d_native():
        _lib.check(each(1, ditems.ctypes.data, len(xs), wsptr.ctypes.data, wsb_c.ctypes.data, h), "plain loop")

    for a in cargs:
        _lib.check(comp(*a), "plain compress")
    assert not statuses().any(), "plain compress status"
    for a in dargs:
        _lib.check(decomp(*a), "plain decompress")
    assert not statuses().any(), "plain decompress status"
    print("batched workload: plain launches warmed up", file=sys.stderr, flush=True)
    t_pc = timed(plain_c, args.batch_steps)
    assert not statuses().any(), "plain compress status"
    t_pc_native = timed(plain_c_native, args.batch_steps)
    assert not statuses().any(), "plain compress status (C++ loop)"
    t_pd = timed(plain_d, args.batch_steps)
    assert not statuses().any(), "plain decompress status"
    t_pd_native = timed(plain_d_native, args.batch_steps)
    assert not statuses().any(), "plain decompress status (C++ loop)"

    # the two forms wrote the same bytes
    assert torch.equal(psz[: len(xs)], bc.sizes[: len(xs)]), "batched and plain stream sizes differ"
    for i in range(len(xs)):
        n = int(sizes[i])
        assert torch.equal(pouts[i][:n], bc.outs[i][:n]), f"stream {i} differs"
        assert torch.equal(pdec[i].view(torch.uint8), bd.outs[i].view(torch.uint8)), f"decoded array {i} differs"
        if ebs[i] is None:
            assert torch.equal(bd.outs[i].view(torch.uint8), xs[i].view(torch.uint8)), f"lossless array {i} differs"
    gc.close()
    gd.close()
    del wsbuf
    enc_total = int(sizes.sum())
    out = {
        "arrays": len(xs), "input_bytes": total, "stream_bytes": enc_total,
        "f32_arrays": sum(1 for x in xs if x.dtype == torch.float32),
        "lossy_arrays": sum(1 for e in ebs if e is not None),
        "sizes": f"log-uniform 4 KiB .. {args.batch_max_mb:g} MiB",
        "tiles": bc.tiles,
        "compress": {"batched_graph_ms": round(t_bc * 1e3, 3), "batched_eager_ms": round(t_bc_eager * 1e3, 3),
                     "plain_ms": round(t_pc_native * 1e3, 3), "plain_python_loop_ms": round(t_pc * 1e3, 3),
                     "batched_GBps": round(total / t_bc / 1e9, 1), "plain_GBps": round(total / t_pc_native / 1e9, 1),
                     "speedup": round(t_pc_native / t_bc, 3), "speedup_vs_python_loop": round(t_pc / t_bc, 3),
                     "batched_hbm_frac": round((total + enc_total) / t_bc / HBM_PEAK, 4)},
        "decompress": {"batched_graph_ms": round(t_bd * 1e3, 3), "batched_eager_ms": round(t_bd_eager * 1e3, 3),
                       "plain_ms": round(t_pd_native * 1e3, 3), "plain_python_loop_ms": round(t_pd * 1e3, 3),
                       "batched_GBps": round(total / t_bd / 1e9, 1), "plain_GBps": round(total / t_pd_native / 1e9, 1),
                       "speedup": round(t_pd_native / t_bd, 3), "speedup_vs_python_loop": round(t_pd / t_bd, 3),
                       "batched_hbm_frac": round((total + enc_total) / t_bd / HBM_PEAK, 4)},
        "plain_loop": "one stream-ordered C ABI call per array, each with its own workspace; plain_ms from a C++ "
                      "loop (sleek_plain_each), plain_python_loop_ms from a Python ctypes loop",
        "generate_s": round(t_gen, 2),
        "identical": True,
    }
    del xs, pouts, pdec, bc, bd, streams
    torch.cuda.empty_cache()
    return out


def bench_serve(args, dev):
    """Serving: Poisson arrivals, request sizes log-uniform 64 KiB .. 256 MiB,
    half compress and half decompress, 70 % float32, half lossless and half
    lossy at {1e-2, 1e-3, 1e-4} x range, admission under a device-memory
    budget, continuous batching over 4 streams.  Inputs are resident on the
    device: compress requests read prefixes of pooled fields; decompress
    requests read streams prepared before the run.  The same requests are
    served at each rate of the sweep (arrival times scaled); after each run
    every result is compared with the same request served alone."""
    import torch

    from sleek_amd import codec
    from sleek_amd.batch import plan_compress
    from sleek_amd.datagen import array_field, finite_range
    from sleek_amd.serve import COMPRESS, DECOMPRESS, GpuEngine, Request, Server, poisson_arrivals, summarize

    rates = [float(r) for r in args.serve_rates.split(",")]
    rng = np.random.default_rng(args.seed + 91)
    hi = int(args.serve_max_mb * (1 << 20))
    pools = {}
    for dt in (np.float32, np.float64):
        pools[dt] = []
        for k in range(3):
            f = array_field(hi, dt, seed=args.seed * 7919 + k + (100 if dt == np.float64 else 0), device=dev)
            pools[dt].append((f, finite_range(f)))
    n = args.serve_requests
    unit = poisson_arrivals(n, 1.0, args.seed + 92)   # arrivals of a rate-1 process, scaled per rate
    reqs = []
    pre_x, pre_eb, pre_idx = [], [], []
    for i in range(n):
        dt = np.float32 if rng.random() < 0.7 else np.float64
        it = np.dtype(dt).itemsize
        nb = int(np.exp(rng.uniform(np.log(64 << 10), np.log(hi))))
        m = max(1, nb // it)
        f, r = pools[dt][int(rng.integers(0, 3))]
        x = f[:m]
        eb = None
        if rng.random() < 0.5:
            rel = float(rng.choice([1e-2, 1e-3, 1e-4]))
            eb = float(np.dtype(dt).type(np.dtype(dt).type(r) * np.dtype(dt).type(rel)))
        if rng.random() < 0.5:
            reqs.append(Request(i, COMPRESS, x, dt, eb))
        else:
            reqs.append(Request(i, DECOMPRESS, None, dt, eb, numel=m))
            pre_x.append(x)
            pre_eb.append(eb)
            pre_idx.append(i)
    # the decompress requests' streams, prepared with one batched compress
    pb = plan_compress(pre_x, pre_eb).launch()
    torch.cuda.synchronize()
    psz = pb.sizes.cpu().numpy()
    for j, i in enumerate(pre_idx):
        reqs[i].src = pb.outs[j][: int(psz[j])]
    print(f"serving workload: {n} requests prepared", file=sys.stderr, flush=True)

    # each request served alone (single-array calls), the reference for every run
    ws = codec.Workspace(hi, dev)
    alone = {}
    for r in reqs:
        if r.kind == COMPRESS:
            o, k = codec.compress(r.src, r.eb, workspace=ws)
            alone[r.rid] = o[: int(k.item())].clone()
        else:
            tdt = torch.float32 if np.dtype(r.dtype) == np.float32 else torch.float64
            alone[r.rid] = codec.decompress(r.src, tdt, r.eb, numel=r.numel, workspace=ws)[0]
        assert ws.status() == 0
    torch.cuda.synchronize()

    engine = GpuEngine(dev, streams=4)

    def serve_once(rs):
        srv = Server(engine, budget_bytes=args.serve_budget_gb * 1e9, max_inflight=4)
        srv.run(rs)
        torch.cuda.synchronize()
        return srv

    # warm-up pass over the first requests (allocator, pinned host buffers)
    serve_once([Request(r.rid, r.kind, r.src, r.dtype, r.eb, r.numel, float(unit[r.rid]) / rates[0])
                for r in reqs[: min(64, n)]])
    sweep = []
    for rate in rates:
        run = [Request(r.rid, r.kind, r.src, r.dtype, r.eb, r.numel, float(unit[r.rid]) / rate) for r in reqs]
        srv = serve_once(run)
        stats = summarize(run, srv)
        bad = [r.rid for r in run if r.status != 0]
        assert not bad, f"requests with errors at {rate:g}/s: {bad[:8]}"
        for r in run:
            ref = alone[r.rid]
            same = torch.equal(r.out, ref) if r.kind == COMPRESS else \
                torch.equal(r.out.view(torch.uint8), ref.view(torch.uint8))
            assert same, f"request {r.rid} at {rate:g}/s differs from its single call"
        stats.update({"rate_rps": rate, "checked_alone": len(run), "identical": True})
        sweep.append(stats)
        print(f"serving: {rate:g}/s -> {stats['GBps']} GB/s, p50 {stats['p50_ms']} ms, p99 {stats['p99_ms']} ms, "
              f"mean batch {stats['mean_batch']}", file=sys.stderr, flush=True)
        del run
    out = dict(sweep[0])
    out.update({"sizes": f"log-uniform 64 KiB .. {args.serve_max_mb:g} MiB", "streams": 4, "sweep": sweep,
                "saturation": max(sweep, key=lambda s: s["GBps"] or 0)})
    del pools, reqs, pb, alone
    torch.cuda.empty_cache()
    return out


def profiled(workload, mode):
    """The rocprofv3 figures of one mode of a benchmark field (profiles/codec_kernels.json:
    kernel trace durations and calibrated FETCH_SIZE / WRITE_SIZE per dispatch, for the
    encoder and the decoder), or None."""
    p = os.path.join(ROOT, "profiles", "codec_kernels.json")
    if not os.path.exists(p):
        return None
    return json.load(open(p)).get("workloads", {}).get(workload, {}).get("modes", {}).get(mode)


def add_traffic(per, workload):
    """HBM bytes per dispatch (rocprofv3, calibrated) and the kernels' own durations next to
    each mode's event-timed calls."""
    for row in per:
        prof = profiled(workload, row["mode"])
        if prof is None:
            continue
        for side in ("encode", "decode"):
            k = prof[side]
            row[f"{side}_traffic_bytes"] = k["fetch_bytes"] + k["write_bytes"]
            row[f"{side}_fetch_bytes"] = k["fetch_bytes"]
            row[f"{side}_write_bytes"] = k["write_bytes"]
            row[f"{side}_kernel_ms_rocprof"] = round(k["avg_us"] / 1e3, 4)


def help_events():
    """Look-back help events on the device so far (missing look-back words a
    waiting wave sized itself; the ticket encoders should need none)."""
    from sleek_amd import _lib

    n = _lib.load().sleek_gpu_help_events()
    _lib.check(0 if n >= 0 else int(n), "help events")
    return int(n)


def run_modes(x, ebs, labels, steps, warmup):
    """Compress then decompress x once per error bound per step, with HIP
    events on the launch stream around every call; then the same steps again
    with events that the library records around each call's coding kernel
    alone (sleek_gpu_set_kernel_events), outside the timed region.  Returns
    the per-mode medians, the stream sizes, the timed region's GPU time (s)
    and its wall time (s).  The lossless round trip is checked exact."""
    import torch

    from sleek_amd import _lib, codec

    dev = x.device
    nbytes = x.numel() * x.element_size()
    ub = codec.upper_bound_size(nbytes)
    stream = torch.cuda.current_stream()
    bufs = [torch.empty(ub, dtype=torch.uint8, device=dev) for _ in ebs]
    encs = [torch.zeros(1, dtype=torch.int64, device=dev) for _ in ebs]
    decs = [torch.empty_like(x) for _ in ebs]
    dsz = torch.zeros(1, dtype=torch.int64, device=dev)
    # one workspace per timed call: each decode's status word (the call's own,
    # cleared by its first kernel) is read back after the timed region
    wss = [[codec.Workspace(nbytes, dev) for _ in ebs] for _ in range(max(steps, 1))]

    lib = _lib.load()

    def step(k, events=None, kevents=None):
        for i, eb in enumerate(ebs):
            if events is not None:
                events[i][0].record(stream)
            if kevents is not None:   # around the encoder kernel alone
                lib.sleek_gpu_set_kernel_events(kevents[i][0].cuda_event, kevents[i][1].cuda_event)
            codec.compress(x, eb, out=bufs[i], encsize=encs[i], workspace=wss[k][i])
            if events is not None:
                events[i][1].record(stream)
            if kevents is not None:   # around the decoder kernel alone (after the offset scan)
                lib.sleek_gpu_set_kernel_events(kevents[i][2].cuda_event, kevents[i][3].cuda_event)
            codec.decompress(bufs[i], x.dtype, eb, out=decs[i], decsize=dsz, workspace=wss[k][i])
            if events is not None:
                events[i][2].record(stream)

    for _ in range(warmup):
        step(0)
        torch.cuda.synchronize()
        assert all(w.status() == 0 for w in wss[0]), "warm-up call status"
    iv = torch.int32 if x.element_size() == 4 else torch.int64
    assert torch.equal(decs[0].view(iv), x.view(iv)), "lossless round trip mismatch"
    enc_sizes = [int(e.item()) for e in encs]
    want = [d.clone() for d in decs]

    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in ebs] for _ in range(steps)]
    kev = [[[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in ebs] for _ in range(steps)]
    for e in (e for k in kev for m in k for e in m):
        e.record(stream)   # creates the HIP event, which the library records again around its kernels
    torch.cuda.synchronize()
    # the timed region: plain calls, events only between them
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for k in range(steps):
        step(k, ev[k])
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_total = start.elapsed_time(end) / 1e3

    def check(what):
        bad = [(k, i) for k in range(steps) for i in range(len(ebs)) if wss[k][i].status() != 0]
        assert not bad, f"{what} calls with a bad status: {bad[:8]}"
        assert [int(e.item()) for e in encs] == enc_sizes, f"stream sizes changed between {what} steps"
        for i in range(len(ebs)):
            assert torch.equal(decs[i].view(iv), want[i].view(iv)), f"mode {labels[i]}: decode changed ({what})"

    check("timed")
    # a second pass, untimed as a whole, for the coding kernels' own durations:
    # the events inside a call add a gap of their own, so they stay out of the
    # timed region
    for k in range(steps):
        step(k, None, kev[k])
    torch.cuda.synchronize()
    check("kernel-timing")

    per = []
    for i, eb in enumerate(ebs):
        tc = np.median([ev[k][i][0].elapsed_time(ev[k][i][1]) for k in range(steps)]) / 1e3
        td = np.median([ev[k][i][1].elapsed_time(ev[k][i][2]) for k in range(steps)]) / 1e3
        kc = np.median([kev[k][i][0].elapsed_time(kev[k][i][1]) for k in range(steps)]) / 1e3
        kd = np.median([kev[k][i][2].elapsed_time(kev[k][i][3]) for k in range(steps)]) / 1e3
        alg = nbytes + enc_sizes[i]
        per.append({
            "mode": labels[i], "eb_abs": eb, "ratio": round(nbytes / enc_sizes[i], 4), "encsize": enc_sizes[i],
            "compress_ms": round(tc * 1e3, 4), "decompress_ms": round(td * 1e3, 4),
            "encode_kernel_ms": round(kc * 1e3, 4), "decode_kernel_ms": round(kd * 1e3, 4),
            "encode_kernel_hbm_frac": round(alg / kc / HBM_PEAK, 4), "decode_kernel_hbm_frac": round(alg / kd / HBM_PEAK, 4),
            "compress_GBps": round(nbytes / tc / 1e9, 1), "decompress_GBps": round(nbytes / td / 1e9, 1),
            "compress_alg_GBps": round(alg / tc / 1e9, 1), "decompress_alg_GBps": round(alg / td / 1e9, 1),
            "compress_hbm_frac": round(alg / tc / HBM_PEAK, 4), "decompress_hbm_frac": round(alg / td / HBM_PEAK, 4),
        })
    del bufs, decs, want, wss
    return per, enc_sizes, t_total, wall


def memcpy_d2d(x, reps=10):
    """hipMemcpy device-to-device of x's bytes (torch copy_): warmed up, median of reps."""
    import torch

    stream = torch.cuda.current_stream()
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    mt = []
    for _ in range(reps):
        a0 = torch.cuda.Event(enable_timing=True)
        a1 = torch.cuda.Event(enable_timing=True)
        a0.record(stream)
        y.copy_(x)
        a1.record(stream)
        torch.cuda.synchronize()
        mt.append(a0.elapsed_time(a1) / 1e3)
    nbytes = x.numel() * x.element_size()
    mtm = float(np.median(mt))
    return {"bytes": nbytes, "ms": round(mtm * 1e3, 4), "effective_GBps": round(nbytes / mtm / 1e9, 1),
            "traffic_GBps": round(2 * nbytes / mtm / 1e9, 1), "hbm_frac": round(2 * nbytes / mtm / HBM_PEAK, 4)}


def bench_transfer(x, eb_lossy, reps=3):
    """Compressing copies of the field between device and pinned host memory
    (sleek_amd.transfer: segments compressed on one stream while the previous
    one crosses PCIe on another), next to plain pinned copies of the same
    bytes (the reference's memcpy.cu:114-139 legs).  Effective GB/s =
    insize / t; every round trip is checked."""
    import torch

    from sleek_amd import transfer

    nbytes = x.numel() * x.element_size()
    host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    flat = x.view(torch.uint8).view(-1)

    def timed(fn):
        ts = []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts[1:])), r

    t_d2h, _ = timed(lambda: host.copy_(flat, non_blocking=True))
    dev_back = torch.empty_like(flat)
    t_h2d, _ = timed(lambda: dev_back.copy_(host, non_blocking=True))
    out = {"bytes": nbytes, "segment_bytes": transfer.DEFAULT_SEGMENT,
           "plain_d2h_GBps": round(nbytes / t_d2h / 1e9, 2), "plain_h2d_GBps": round(nbytes / t_h2d / 1e9, 2),
           "modes": []}
    del dev_back
    for label, eb in (("lossless", None), ("eb=1e-3*range", eb_lossy)):
        t_off, cont = timed(lambda: transfer.offload(x, eb))
        t_rel, y = timed(lambda: transfer.reload(cont, x.device))
        if eb is None:
            assert torch.equal(y.view(torch.uint8), x.view(torch.uint8)), "offload/reload round trip mismatch"
        out["modes"].append({"mode": label, "container_bytes": int(cont.numel()),
                             "ratio": round(nbytes / cont.numel(), 4),
                             "offload_ms": round(t_off * 1e3, 3), "reload_ms": round(t_rel * 1e3, 3),
                             "offload_GBps": round(nbytes / t_off / 1e9, 2),
                             "reload_GBps": round(nbytes / t_rel / 1e9, 2),
                             "vs_plain_d2h": round(t_d2h / t_off, 3), "vs_plain_h2d": round(t_h2d / t_rel, 3)})
        del cont, y
    return out


def bench_f64(args, dev):
    """The float64 configuration: a 384^3 field (12 sinusoids + lognormal
    noise, 0.01 % specials), lossless and eb = {1e-6, 1e-9} x range, next to
    a device-to-device copy of the same bytes."""
    from sleek_amd.datagen import config_field, finite_range

    x = config_field(2, seed=args.seed, device=dev, scale=args.scale).reshape(-1).contiguous()
    r = finite_range(x)
    rels = (1e-6, 1e-9)
    ebs = [None] + [float(np.float64(r) * np.float64(e)) for e in rels]
    labels = ["lossless"] + [f"eb={e:g}*range" for e in rels]
    per, _, t_total, _ = run_modes(x, ebs, labels, max(3, args.steps // 2), 1)
    if args.scale == 1.0:
        add_traffic(per, "f64")
    nbytes = x.numel() * 8
    out = {"workload": "384^3 float64, 12 sinusoids + lognormal(-10, 2) noise, 0.01% specials",
           "bytes": nbytes, "per_mode": per, "memcpy_d2d": memcpy_d2d(x),
           "GBps": round(2 * nbytes * len(ebs) / (t_total / max(3, args.steps // 2)) / 1e9, 2)}
    host = None if args.no_cpu else (x.cpu().numpy(), list(zip(labels, ebs)))
    del x
    return out, host


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this codec runs on one GPU; multi-GPU runs are out of scope "
                                   "(a field would shard into contiguous runs of whole chunks, one per GPU, "
                                   "with one exchange of the runs' compressed sizes)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from sleek_amd import codec
    from sleek_amd.datagen import config_field, finite_range

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    x = config_field(1, seed=args.seed, device=dev, scale=args.scale).reshape(-1).contiguous()
    if args.decoder >= 0:
        from sleek_amd import _lib

        _lib.check(_lib.load().sleek_gpu_set_decoder(args.decoder), "decoder")
    tuning = None
    if args.encode_variant or args.encode_grid or args.decode_grid:
        from sleek_amd import _lib

        _lib.check(_lib.load().sleek_gpu_set_tuning(args.encode_variant, args.encode_grid, args.decode_grid), "tuning")
        tuning = {"encode_variant": args.encode_variant, "encode_grid": args.encode_grid,
                  "decode_grid": args.decode_grid, "source": "command line"}
    nbytes = x.numel() * 4
    rng = finite_range(x)
    ebs = [None] + [float(np.float32(np.float32(rng) * np.float32(e))) for e in MODES[1:]]
    labels = ["lossless"] + [f"eb={m:g}*range" for m in MODES[1:]]
    if tuning is None and not args.no_autotune:
        # launch shape for this input and these modes, searched before the timed
        # region (bytes are checked invariant)
        from sleek_amd import autotune

        best, trace = autotune.tune(x, ebs, reps=3, cache=False)
        tuning = dict(best, source="autotune", candidates=len(trace))
        for cfg, te, td in trace:
            print(f"autotune {json.dumps(cfg)} encode {te * 1e3:.4f} ms decode {td * 1e3:.4f} ms (all modes)",
                  file=sys.stderr, flush=True)
    h0 = help_events()
    per, enc_sizes, t_total, wall = run_modes(x, ebs, labels, args.steps, args.warmup)
    helps = {"plain": help_events() - h0}
    ms_step = t_total / args.steps * 1e3
    stream = torch.cuda.current_stream()

    # dominant kernel: the launch with the largest median time
    cands = []
    for i, p in enumerate(per):
        alg = nbytes + enc_sizes[i]
        cands.append((p["encode_kernel_ms"], f"encode ({p['mode']})", alg, p["compress_ms"]))
        cands.append((p["decode_kernel_ms"], f"decode ({p['mode']})", alg, p["decompress_ms"]))
    tms, kname, alg, call_ms = max(cands)
    achieved = alg / (tms / 1e3) / 1e9
    if args.scale == 1.0:
        add_traffic(per, "f32")
    dmode = kname.split("(", 1)[1].rstrip(")")
    dside = kname.split(" ", 1)[0]
    prof = profiled("f32", dmode) if args.scale == 1.0 else None
    kprof = prof[dside] if prof else None

    # hipMemcpy device-to-device comparator of the same bytes
    mc = memcpy_d2d(x)

    total_in = 2 * nbytes * len(ebs)
    out = {
        "metric": "compress+decompress GB/s & % HBM peak; batched vs plain; serve p50/p99",
        "value": round(total_in / (t_total / args.steps) / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "512^3 float32 sinusoid+ramp field with 0.01% specials; per step compress and "
                               "decompress in lossless and eb={1e-2,1e-3,1e-4}*range modes; value = total "
                               "input bytes of the 8 calls / step time (GB/s); % of HBM peak in roofline and "
                               "per_mode; batched vs plain in batched; serve p50/p99 in serve",
                   "shape": [int(round(512 * args.scale))] * 3, "bytes": nbytes},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved * 1e9 / HBM_PEAK, 4),
                     "traffic": (kprof["fetch_bytes"] + kprof["write_bytes"]) if kprof else None,
                     "traffic_detail": ({"fetch_bytes": kprof["fetch_bytes"], "write_bytes": kprof["write_bytes"],
                                         "kernel": kprof["kernel"], "source": "profiles/codec_kernels.json"}
                                        if kprof else None),
                     "alg_bytes_per_launch": alg, "avg_launch_ms": round(tms, 4),
                     "timing": "avg_launch_ms: median over a second pass of the same steps of the kernel's own "
                               "duration, HIP events that the library records on the launch stream right before "
                               "and after the kernel (sleek_gpu_set_kernel_events); call_ms: the whole call (with "
                               "the workspace clear or offset scan) in the timed region, events around it only",
                     "call_ms": round(call_ms, 4), "frac_call": round(alg / (call_ms / 1e3) / HBM_PEAK, 4),
                     "kernel_ms_rocprof": round(kprof["avg_us"] / 1e3, 4) if kprof else None,
                     "frac_kernel_rocprof": (round(alg / (kprof["avg_us"] * 1e-6) / HBM_PEAK, 4) if kprof else None)},
        "per_mode": per,
        "tuning": tuning,
        "memcpy_d2d": mc,
        "wall_s": round(wall, 3),
    }
    if not args.no_transfer:
        h0 = help_events()
        out["transfer"] = bench_transfer(x, ebs[2])
        helps["transfer"] = help_events() - h0
    cpu_fields = []
    if not args.no_cpu:
        f0 = config_field(0, seed=args.seed, device=dev, scale=args.scale).reshape(-1).cpu().numpy()
        cpu_fields.append((f"{f0.size}-value float32 (256^3 field, 8 sinusoids + noise)", f0, [("lossless", None)]))
        cpu_fields.append((f"{x.numel()}-value float32 (the 512^3 field above)", x.cpu().numpy(),
                           list(zip(labels, ebs))))
    del x
    torch.cuda.empty_cache()
    if not args.no_f64:
        h0 = help_events()
        out["float64"], f64_host = bench_f64(args, dev)
        helps["float64"] = help_events() - h0
        if f64_host is not None:
            cpu_fields.append((f"{f64_host[0].size}-value float64 (384^3 field)", f64_host[0], f64_host[1]))
    if not args.no_batch:
        h0 = help_events()
        out["batched"] = bench_batch(args, dev)
        helps["batched"] = help_events() - h0
    if not args.no_serve:
        h0 = help_events()
        out["serve"] = bench_serve(args, dev)
        helps["serve"] = help_events() - h0
    out["lookback_help_events"] = helps
    if cpu_fields:
        out["cpu_baseline"] = cpu_baseline(cpu_fields)
        del cpu_fields
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
