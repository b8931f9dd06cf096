"""The serving loop's scheduling (sleek_amd.serve.Server) on the CPU, with a
simulated engine: every request completes, admission follows arrival order,
the device-memory budget is never exceeded, batches overlap, and requests
that arrive while others run are batched with each other."""
import numpy as np
import pytest

from sleek_amd.serve import COMPRESS, DECOMPRESS, Request, Server, poisson_arrivals


class FakeSrc:
    def __init__(self, nbytes):
        self.n = nbytes

    def numel(self):
        return self.n

    def element_size(self):
        return 1


class Clock:
    """Deterministic clock: every reading advances it by dt seconds."""

    def __init__(self, dt=1e-5):
        self.t, self.dt = 0.0, dt

    def __call__(self):
        self.t += self.dt
        return self.t


class FakeEngine:
    """A batch takes 20 us plus 1 us per MB; footprint = 2x the input bytes."""

    def __init__(self, clock, budget):
        self.clock, self.budget = clock, budget
        self.inflight = []
        self.launches = []
        self.max_used = 0

    def footprint(self, r):
        return 2 * r.in_bytes

    def launch(self, kind, reqs, slot):
        assert len({r.kind for r in reqs}) == 1 and reqs[0].kind == kind
        fp = sum(self.footprint(r) for r in reqs)
        self.inflight.append(fp)
        self.max_used = max(self.max_used, sum(self.inflight))
        assert sum(self.inflight) <= self.budget
        end = self.clock.t + 20e-6 + sum(r.in_bytes for r in reqs) / 1e12
        self.launches.append((self.clock.t, end, [r.rid for r in reqs]))
        return [end, fp]

    def poll(self, h):
        if self.clock.t >= h[0]:
            if h[1] is not None:
                self.inflight.remove(h[1])
                h[1] = None
            return True
        return False


def make_requests(n, rate, seed, lo=64 << 10, hi=256 << 20):
    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(n, rate, seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)
    kinds = np.where(rng.random(n) < 0.5, COMPRESS, DECOMPRESS)
    return [Request(i, str(k), FakeSrc(int(s)), np.float32, numel=int(s) // 4, arrival=float(a))
            for i, (a, s, k) in enumerate(zip(arr, sizes, kinds))]


@pytest.mark.parametrize("budget", [96e9, 2e9, 600e6])
def test_all_complete_in_order_within_budget(budget):
    clock = Clock()
    eng = FakeEngine(clock, budget)
    srv = Server(eng, budget_bytes=budget, max_inflight=4, clock=clock)
    reqs = srv.run(make_requests(400, 2000, 1), t0=0.0)
    assert all(np.isfinite(r.done) and r.done >= r.admitted >= r.arrival for r in reqs)
    adm = [r.admitted for r in sorted(reqs, key=lambda r: r.arrival)]
    assert adm == sorted(adm)                        # first come, first admitted
    assert eng.max_used <= budget and srv.peak_used <= budget
    assert sorted(sum((ids for _, _, ids in eng.launches), [])) == list(range(400))


def test_batches_overlap_and_coalesce():
    clock = Clock()
    eng = FakeEngine(clock, 96e9)
    srv = Server(eng, budget_bytes=96e9, max_inflight=4, clock=clock)
    srv.run(make_requests(300, 200000, 2), t0=0.0)   # arrivals faster than batches finish
    assert srv.max_concurrent > 1                    # batches in flight together
    assert max(srv.batch_sizes) > 1                  # arrivals during a batch share the next one
    starts = [(s, e) for s, e, _ in eng.launches]
    assert any(s2 < e1 for (s1, e1), (s2, e2) in zip(starts, starts[1:]))


def test_head_of_line_waits_for_memory():
    """A request that does not fit waits, and later small requests do not overtake it."""
    clock = Clock()
    eng = FakeEngine(clock, 1000)
    srv = Server(eng, budget_bytes=1000, max_inflight=4, clock=clock)
    reqs = [Request(0, COMPRESS, FakeSrc(300), np.float32, arrival=0.0),
            Request(1, COMPRESS, FakeSrc(450), np.float32, arrival=1e-6),
            Request(2, COMPRESS, FakeSrc(10), np.float32, arrival=2e-6)]
    srv.run(reqs, t0=0.0)
    assert reqs[1].admitted >= reqs[0].done
    assert reqs[2].admitted >= reqs[1].admitted


def test_oversized_request_is_rejected():
    clock = Clock()
    srv = Server(FakeEngine(clock, 100), budget_bytes=100, clock=clock)
    with pytest.raises(ValueError):
        srv.submit(Request(0, COMPRESS, FakeSrc(51), np.float32))
