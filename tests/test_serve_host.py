"""Host logic of the serving loop on the CPU: admission by free HBM against a
fake budget, the request stream, and the latency percentile."""
import collections

import numpy as np
import pytest

from ntamd.serve import Request, admit, make_requests, percentile


def req(rid, points):
    """A derivative request whose field holds `points` points (16 B each admitted)."""
    return Request(rid=rid, kind="deriv", arrival=0.0, shape=(points, 1, 1))


def queue(*sizes):
    return collections.deque(req(i, p) for i, p in enumerate(sizes))


def test_admit_fills_the_budget_in_arrival_order():
    w = queue(10, 20, 30, 40)                       # 160, 320, 480, 640 bytes
    got, used = admit(w, 0, 1000, 64)
    assert [q.rid for q in got] == [0, 1, 2] and used == 960
    assert [q.rid for q in w] == [3]                 # 960 + 640 > 1000: waits


def test_admit_head_of_line_blocks():
    # the head does not fit beside what is in flight; a smaller one behind it
    # is not admitted ahead of it
    w = queue(50, 1)
    got, used = admit(w, 400, 1000, 64)
    assert got == [] and used == 400 and len(w) == 2


def test_admit_oversized_request_alone_when_idle():
    w = queue(1000, 1)                              # 16,000 bytes > budget
    got, used = admit(w, 0, 1000, 64)
    assert [q.rid for q in got] == [0] and used == 16000
    got, used = admit(w, used, 1000, 64)             # nothing more while it runs
    assert got == [] and used == 16000


def test_admit_never_exceeds_budget_when_busy():
    rng = np.random.default_rng(0)
    w = collections.deque(req(i, int(p)) for i, p in enumerate(rng.integers(1, 200, 500)))
    used, budget, inflight, order = 0, 2000, [], []
    while w or inflight:
        got, used = admit(w, used, budget, 8)
        assert len(got) <= 8
        assert used <= budget or len(inflight) == 0 and len(got) == 1
        inflight += got
        order += [q.rid for q in got]
        if inflight:                                # the oldest finishes
            used -= inflight.pop(0).nbytes()
    assert order == list(range(500))                 # first come, first served
    assert used == 0


def test_admit_max_batch():
    w = queue(*([1] * 100))
    got, used = admit(w, 0, 1 << 30, 64)
    assert len(got) == 64 and len(w) == 36 and used == 64 * 16


def test_request_bytes_and_points():
    d = Request(rid=0, kind="deriv", arrival=0.0, shape=(4, 5, 6))
    a = Request(rid=1, kind="adi", arrival=0.0, side=32, steps=3)
    assert d.nbytes() == 4 * 5 * 6 * 16 and d.points() == 120
    assert a.nbytes() == 32 * 32 * 16 and a.points() == 1024


def test_make_requests_is_seeded_and_poisson():
    a = make_requests(2000, 1000.0, seed=5)
    b = make_requests(2000, 1000.0, seed=5)
    assert [(q.kind, q.arrival, q.shape, q.side) for q in a] == [(q.kind, q.arrival, q.shape, q.side) for q in b]
    t = np.array([q.arrival for q in a])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(t)
    # exponential gaps at 1000 req/s: mean 1 ms, coefficient of variation 1
    assert gaps.mean() == pytest.approx(1e-3, rel=0.1)
    assert gaps.std() / gaps.mean() == pytest.approx(1.0, rel=0.15)
    up = make_requests(1000, 100.0, seed=5, upload_share=0.5)
    assert 400 < sum(q.upload for q in up) < 600
    # the uploads come from a generator of their own: the rest of the stream is unchanged
    base = make_requests(1000, 100.0, seed=5)
    assert [(q.kind, q.arrival, q.shape, q.side, q.seed) for q in up] == \
        [(q.kind, q.arrival, q.shape, q.side, q.seed) for q in base]


def test_percentile_matches_numpy():
    v = list(np.random.default_rng(1).exponential(size=999))
    for q in (50, 90, 99):
        assert percentile(v, q) == pytest.approx(np.percentile(v, q))
