"""Serving loop: continuous batching of compress/decompress requests on one GPU.

Requests arrive over time (each carries its arrival time).  The loop admits
them in arrival order while two limits hold: the device-memory budget (the
bytes a request keeps on the device while in flight: its input, its output
buffer and its share of the workspace) and the number of batches in flight.
Every pass of the loop coalesces what it admitted into at most one batched
launch per direction (sleek_amd.batch, one launch sequence for all arrays),
issued on the next of several HIP streams, so batches admitted while others
run overlap on the device.  A batch's completion (an event on its stream)
completes its requests and returns their memory to the budget.  Latency is
measured from a request's arrival to the loop's observation of its batch's
completion.

The scheduling logic (``Server``) is independent of the device: ``GpuEngine``
launches through the HIP codec; tests drive ``Server`` with a simulated
engine on the CPU.  There is no CPU codec path.
"""
import collections
import dataclasses
import math
import time

import numpy as np

COMPRESS = "compress"
DECOMPRESS = "decompress"


@dataclasses.dataclass
class Request:
    """One codec request.  For compress, ``src`` is the array (a contiguous
    float32/float64 CUDA tensor); for decompress, ``src`` is the stream (a
    uint8 CUDA tensor), ``dtype`` the array type and ``numel`` its value
    count (the capacity).  ``eb`` None = lossless, else the absolute bound."""
    rid: int
    kind: str
    src: object
    dtype: object
    eb: object = None
    numel: int = 0
    arrival: float = 0.0
    # filled by the server
    admitted: float = math.nan
    done: float = math.nan
    out: object = None
    size: int = -1
    status: int = 0
    batch: int = -1

    @property
    def in_bytes(self):
        return int(self.src.numel() * self.src.element_size())

    @property
    def latency(self):
        return self.done - self.arrival


class Server:
    """Admission, batching and completion of requests (device-independent).

    engine: an object with ``footprint(req) -> bytes``,
    ``launch(kind, reqs, slot) -> handle`` and ``poll(handle) -> bool``
    (True once the batch has completed and its requests carry results)."""

    def __init__(self, engine, budget_bytes=96e9, max_inflight=4, max_batch=512, clock=time.perf_counter,
                 on_done=None):
        self.engine = engine
        self.on_done = on_done      # called with each completed request
        self.budget = float(budget_bytes)
        self.max_inflight = int(max_inflight)
        self.max_batch = int(max_batch)
        self.clock = clock
        self.pending = collections.deque()
        self.inflight = []          # [handle, requests, bytes]
        self.used = 0.0
        self.peak_used = 0.0
        self.batches = 0
        self.batch_sizes = []
        self.max_concurrent = 0
        self.slot = 0

    def submit(self, req):
        if self.engine.footprint(req) > self.budget:
            raise ValueError(f"request {req.rid} needs more than the whole device-memory budget")
        self.pending.append(req)

    def _complete(self, now):
        done = []
        still = []
        for h, reqs, nbytes in self.inflight:
            if self.engine.poll(h):
                for r in reqs:
                    r.done = now
                    if self.on_done is not None:
                        self.on_done(r)
                done.extend(reqs)
                self.used -= nbytes
            else:
                still.append([h, reqs, nbytes])
        self.inflight = still
        return done

    def _admit(self, now):
        launched = 0
        while self.pending and len(self.inflight) < self.max_inflight:
            take = {COMPRESS: [], DECOMPRESS: []}
            nb = {COMPRESS: 0.0, DECOMPRESS: 0.0}
            count = 0
            # in arrival order; the head waits for memory rather than being overtaken
            while self.pending and count < self.max_batch:
                r = self.pending[0]
                fp = self.engine.footprint(r)
                if self.used + nb[COMPRESS] + nb[DECOMPRESS] + fp > self.budget:
                    break
                self.pending.popleft()
                take[r.kind].append(r)
                nb[r.kind] += fp
                count += 1
            if count == 0:
                break
            for kind in (COMPRESS, DECOMPRESS):
                reqs = take[kind]
                if not reqs:
                    continue
                if len(self.inflight) >= self.max_inflight:
                    # no slot for the second direction: return it to the queue head, in order
                    for r in reversed(reqs):
                        self.pending.appendleft(r)
                    continue
                for r in reqs:
                    r.admitted = now
                    r.batch = self.batches
                h = self.engine.launch(kind, reqs, self.slot)
                self.slot += 1
                self.inflight.append([h, reqs, nb[kind]])
                self.used += nb[kind]
                self.batches += 1
                self.batch_sizes.append(len(reqs))
                launched += 1
            self.peak_used = max(self.peak_used, self.used)
            self.max_concurrent = max(self.max_concurrent, len(self.inflight))
        return launched

    def run(self, requests, t0=None):
        """Serves `requests` (sorted by arrival, in seconds from t0) to the end;
        returns them with admission/completion times filled in."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = self.clock() if t0 is None else t0
        i = 0
        finished = 0
        while finished < len(reqs):
            now = self.clock() - t0
            while i < len(reqs) and reqs[i].arrival <= now:
                self.submit(reqs[i])
                i += 1
            finished += len(self._complete(now))
            self._admit(now)
        return reqs


def summarize(reqs, server):
    """Throughput and latency figures of a finished run."""
    lat = np.array([r.latency for r in reqs]) * 1e3
    nbytes = sum(r.in_bytes if r.kind == COMPRESS else r.numel * np.dtype(r.dtype).itemsize for r in reqs)
    first = min(r.arrival for r in reqs)
    last = max(r.done for r in reqs)
    span = last - first
    return {
        "requests": len(reqs),
        "compress": sum(1 for r in reqs if r.kind == COMPRESS),
        "decompress": sum(1 for r in reqs if r.kind == DECOMPRESS),
        "bytes": int(nbytes),
        "span_s": round(span, 4),
        "GBps": round(nbytes / span / 1e9, 2) if span > 0 else None,
        "offered_rps": round(len(reqs) / max(max(r.arrival for r in reqs) - first, 1e-9), 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "max_ms": round(float(lat.max()), 3),
        "mean_ms": round(float(lat.mean()), 3),
        "batches": server.batches,
        "mean_batch": round(float(np.mean(server.batch_sizes)), 2) if server.batch_sizes else 0,
        "max_batches_in_flight": server.max_concurrent,
        "peak_budget_bytes": int(server.peak_used),
        "budget_bytes": int(server.budget),
    }


class GpuEngine:
    """Launches batches through the HIP codec on a ring of HIP streams."""

    def __init__(self, device="cuda", streams=4):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(streams)]

    def footprint(self, r):
        from .codec import upper_bound_size, workspace_size

        if r.kind == COMPRESS:
            out = upper_bound_size(r.in_bytes)
            ws = workspace_size(r.in_bytes)
        else:
            out = r.numel * np.dtype(r.dtype).itemsize
            ws = workspace_size(out)
        return r.in_bytes + out + ws + 96 + 16

    def launch(self, kind, reqs, slot):
        torch = self.torch
        from .batch import Batch, ITEM, fill_item
        from .codec import upper_bound_size

        s = self.streams[slot % len(self.streams)]
        n = len(reqs)
        items = np.zeros(n, dtype=ITEM)
        with torch.cuda.stream(s):
            sizes = torch.zeros(n, dtype=torch.int64, device=self.device)
            outs = []
            for i, r in enumerate(reqs):
                if kind == COMPRESS:
                    o = torch.empty(max(upper_bound_size(r.in_bytes), 8), dtype=torch.uint8, device=self.device)
                    items[i]["src_bytes"] = r.in_bytes
                    items[i]["dst_bytes"] = o.numel()
                    fill_item(items[i], r.src.dtype, r.eb)
                else:
                    tdt = torch.float32 if np.dtype(r.dtype) == np.float32 else torch.float64
                    o = torch.empty(int(r.numel), dtype=tdt, device=self.device)
                    items[i]["src_bytes"] = r.in_bytes
                    items[i]["dst_bytes"] = o.numel() * o.element_size()
                    fill_item(items[i], tdt, r.eb)
                items[i]["src"] = r.src.data_ptr()
                items[i]["dst"] = o.data_ptr()
                items[i]["d_size"] = sizes.data_ptr() + 8 * i
                outs.append(o)
            b = Batch(items, kind == DECOMPRESS, self.device, stream=s)
            b.launch(stream=s)
            ev = torch.cuda.Event()
            ev.record(s)
        return [ev, b, reqs, outs, sizes, s]

    def poll(self, h):
        ev, b, reqs, outs, sizes, s = h
        if not ev.query():
            return False
        st = b.status(stream=s)
        sz = sizes.cpu().numpy()
        for i, r in enumerate(reqs):
            r.status = int(st[i])
            r.size = int(sz[i])
            r.out = outs[i][: r.size] if r.kind == COMPRESS else outs[i]
        return True


def poisson_arrivals(n, rate, seed):
    """Arrival times (s) of a Poisson process of `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))
