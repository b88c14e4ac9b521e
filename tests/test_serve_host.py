"""Host-side logic of the continuous-batching server (no GPU): admission by HBM budget,
batch size limit, arrival order, scenes that are busy, retirement freeing the budget."""
import pytest
import torch

from diffvg_amd import serve


class FakeScene:
    pass


class Harness(serve.Server):
    def __init__(self, **kw):
        super().__init__(stream_factory=lambda: object(), **kw)
        self.launched = []

    def _launch(self, group, stream):
        self.launched.append([r.rid for r in group])
        self.flight.append(serve._InFlight(group, None, None, None, stream, None))

    def finish_all(self):
        for f in self.flight:
            for r in f.reqs:
                self.in_use -= r.bytes
                self.done.append(r)
            self.free_streams.append(f.stream)
        self.flight = []


def req(i, scene=None, w=64, h=64):
    return serve.Request(i, scene or FakeScene(), w, h, 2, 2, 0, False, False, None, 0.0)


def test_request_bytes_grow_with_size():
    assert serve.request_bytes(64, 64, 4, False) < serve.request_bytes(128, 128, 4, False)
    assert serve.request_bytes(64, 64, 1, True) < serve.request_bytes(64, 64, 4, False)


def test_admission_respects_budget_batch_size_and_order():
    one = req(0).bytes
    h = Harness(hbm_budget=3 * one, max_batch=2, streams=4)
    for i in range(6):
        h.submit(req(i))
    h._admit(0.0)
    h._admit(0.0)
    assert h.launched == [[0, 1], [2]]          # batch of 2, then the budget allows one more
    assert [r.rid for r in h.waiting] == [3, 4, 5]
    h.finish_all()
    h._admit(0.0)
    h._admit(0.0)
    assert h.launched[2:] == [[3, 4], [5]]


def test_busy_scene_waits_for_next_batch():
    sc = FakeScene()
    h = Harness(hbm_budget=1 << 40, max_batch=8, streams=4)
    h.submit(req(0, sc))
    h.submit(req(1, sc))
    h.submit(req(2))
    h._admit(0.0)
    assert h.launched == [[0, 2]]
    assert [r.rid for r in h.waiting] == [1]
    h._admit(0.0)
    assert h.launched == [[0, 2]]               # still busy while request 0 is in flight
    h.finish_all()
    h._admit(0.0)
    assert h.launched[-1] == [1]


def test_oversized_request_is_rejected():
    h = Harness(hbm_budget=10, max_batch=8, streams=1)
    with pytest.raises(ValueError):
        h.submit(req(0))
