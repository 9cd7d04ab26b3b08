"""Host logic of the continuous-batching server (php_fdtd_amd/serve.py) on
the CPU: admission by device footprint and member slots, chunk boundaries at
the next member's finish, per-request errors, and the release of every
context after an error.  The batch and the contexts are stand-ins that record
what the server asks of them; no GPU call is made."""
import numpy as np
import pytest

from php_fdtd_amd import serve


class FakeSim:
    live = []

    def __init__(self, req):
        self.req = req
        self.steps = 0
        self.closed = False
        FakeSim.live.append(self)

    def monitor(self, n):
        assert self.steps == n == self.req.steps
        return np.arange(n, dtype=np.float64)[:, None] * np.ones((1, 12))

    def sync(self):
        return type("D", (), {"diverged": False})()

    def close(self):
        self.closed = True


class FakeBatch:
    fail_after = None  # raise in run() after this many runs
    log = []

    def __init__(self, precision, lib=None):
        self.members = []
        self.runs = 0
        FakeBatch.log = []

    def add(self, sim):
        self.members.append(sim)

    def remove(self, sim):
        self.members.remove(sim)

    def run(self, n):
        if FakeBatch.fail_after is not None and self.runs >= FakeBatch.fail_after:
            raise RuntimeError("batch run failed")
        self.runs += 1
        FakeBatch.log.append((n, [s.req.id for s in self.members], sum(s.req.footprint for s in self.members)))
        for s in self.members:
            s.steps += n
            assert s.steps <= s.req.steps

    def wait(self, keep=0):
        pass

    def sync(self):
        pass

    def close(self):
        self.members = []


@pytest.fixture
def fakes(monkeypatch):
    FakeSim.live = []
    FakeBatch.fail_after = None
    monkeypatch.setattr(serve, "Batch", FakeBatch)
    monkeypatch.setattr(serve, "_build", lambda req, precision: FakeSim(req))
    monkeypatch.setattr(serve, "footprint_bytes", lambda spec, med, prec, lib=None: float(spec["fp"]))
    monkeypatch.setattr(serve._abi, "load_library", lambda: None)
    yield


def make_reqs(steps, fps, gap=0.0):
    out = []
    for i, (n, fp) in enumerate(zip(steps, fps)):
        spec = {"grid": {"tsteps": n}, "fp": fp}
        out.append(serve.Request(i, spec, np.zeros((2, 2, 2), np.int32), arrival=gap * i))
    return out


def test_all_requests_complete_within_limits(fakes):
    rng = np.random.default_rng(0)
    steps = [int(s) for s in rng.integers(1, 300, 24)]
    fps = [float(f) for f in rng.integers(1, 5, 24)]
    reqs = make_reqs(steps, fps, gap=0.0005)
    srv = serve.Server(hbm_budget=8.0, max_members=4, max_chunk=64, workers=2)
    srv.serve(reqs)
    assert all(r.error is None and r.rows is not None and len(r.rows) == r.steps for r in reqs)
    assert all(s.closed for s in FakeSim.live) and len(FakeSim.live) == len(reqs)
    for n, members, fp in FakeBatch.log:
        assert 1 <= n <= 64
        assert len(members) <= 4
        assert fp <= 8.0
    # a chunk ends when the next member finishes: every member's last chunk lands exactly on its step count
    assert sum(n * len(m) for n, m, _ in FakeBatch.log) == sum(steps)
    st = serve.summarize(reqs)
    assert st["completed"] == len(reqs) and st["p50_s"] <= st["p99_s"]


def test_admission_is_fifo_by_footprint(fakes):
    # budget 10: requests 0 (6) and 1 (6) cannot share the device; 2 (3) waits behind 1 (FIFO)
    reqs = make_reqs([100, 100, 100], [6.0, 6.0, 3.0])
    srv = serve.Server(hbm_budget=10.0, max_members=8, max_chunk=1000, workers=1)
    srv.serve(reqs)
    assert all(r.error is None for r in reqs)
    assert reqs[0].t_start <= reqs[1].t_start <= reqs[2].t_start
    assert max(fp for _, _, fp in FakeBatch.log) <= 10.0
    assert not any({0, 1} <= set(m) for _, m, _ in FakeBatch.log)


def test_oversized_request_and_build_failure_are_per_request(fakes, monkeypatch):
    def build(req, precision):
        if req.id == 1:
            raise RuntimeError("context build failed")
        return FakeSim(req)

    monkeypatch.setattr(serve, "_build", build)
    reqs = make_reqs([50, 50, 50, 50], [2.0, 2.0, 99.0, 2.0])
    srv = serve.Server(hbm_budget=10.0, max_members=4, max_chunk=16, workers=2)
    srv.serve(reqs)
    assert reqs[1].error == "context build failed"
    assert "exceeds the HBM budget" in reqs[2].error
    assert reqs[0].error is None and reqs[3].error is None
    assert len(reqs[0].rows) == 50 and len(reqs[3].rows) == 50
    assert srv._in_use == 0.0  # every footprint returned to the budget


def test_contexts_released_after_an_error(fakes):
    FakeBatch.fail_after = 2
    reqs = make_reqs([200] * 6, [1.0] * 6)
    srv = serve.Server(hbm_budget=100.0, max_members=3, max_chunk=32, workers=2)
    with pytest.raises(RuntimeError, match="batch run failed"):
        srv.serve(reqs)
    assert FakeSim.live and all(s.closed for s in FakeSim.live)
