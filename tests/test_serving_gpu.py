"""Serving loop: concurrent streams of different shapes, dtypes and codecs
admitted against an HBM budget. Each request's shards must be identical when
it runs alone and when it runs in the mix, and the budget must bound how many
run at once."""
import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Dim, Server
from tests.helpers import random_typed, thirds_volume, xor_volume

pytestmark = pytest.mark.gpu


def _requests():
    return [
        (ca.Config([Dim(64, 16, 2), Dim(64, 32, 2), Dim(64, 32, 2)], dtype="u16", codec="zstd"),
         xor_volume((64, 64, 64))),
        (ca.Config([Dim(48, 8, 2), Dim(96, 32, 0), Dim(96, 32, 0)], dtype="u8", codec="lz4"),
         random_typed(48 * 96 * 96, "u8", seed=1)),
        (ca.Config([Dim(32, 8, 2, downsample=True), Dim(64, 16, 0, downsample=True),
                    Dim(64, 16, 0, downsample=True)], dtype="f32", codec="zstd", append_reduce="mean"),
         random_typed(32 * 64 * 64, "f32", seed=2)),
        (ca.Config([Dim(32, 16, 1), Dim(96, 32, 2), Dim(192, 64, 0)], dtype="u16", codec="zstd"),
         thirds_volume((32, 96, 192))),
        (ca.Config([Dim(40, 8, 0), Dim(40, 8, 0, downsample=True), Dim(40, 8, 0, downsample=True)],
                   dtype="i16", codec="none", reduce="median"),
         random_typed(40 * 40 * 40, "i16", seed=3)),
        (ca.Config([Dim(16, 4, 4), Dim(128, 32, 0), Dim(128, 32, 0)], dtype="u32", codec="lz4"),
         random_typed(16 * 128 * 128, "u32", seed=4) % 1000),
    ]


def _alone(cfg, data, pipeline=ca.PIPELINE_PLAIN):
    sink = ca.MemSink()
    s = ca.GpuStream(cfg, sink, pipeline=pipeline)
    assert s.append(data) == 0 and s.flush() == 0, ca.last_error()
    s.close()
    return sink.shards()


@pytest.mark.parametrize("device", [False, True])
def test_each_result_same_alone_and_in_the_mix(device):
    # the serving loop runs the batched pipeline: each request's shards are
    # byte-identical to the same pipeline's run of that request alone (the
    # plain pipeline's decode to the same chunks; for the epochs that emit
    # coarse levels the batched pipeline compresses those in half blocks, so
    # its zstd bytes there are its own)
    reqs = _requests()
    want = [_alone(c, d, ca.PIPELINE_BATCHED) for c, d in reqs]
    srv = Server(hbm_budget_bytes=64 << 30, max_concurrent=8, pipeline=ca.PIPELINE_BATCHED)
    sinks, bufs = [], []
    for c, d in reqs:
        sink = ca.MemSink()
        sinks.append(sink)
        payload = ca.DeviceBuffer.from_host(d) if device else d
        bufs.append(payload)
        srv.submit(c, sink, payload, device=device)
    assert srv.wait() == 0, ca.last_error()
    reps = srv.reports()
    assert len(reps) == len(reqs) and all(r.status == 0 for r in reps)
    running, _ = srv.peaks()
    assert running >= 2, "requests never overlapped"
    srv.close()
    for i, sink in enumerate(sinks):
        got = sink.shards()
        assert got.keys() == want[i].keys(), i
        for k in want[i]:
            assert got[k] == want[i][k], (i, k)


def test_budget_bounds_concurrency():
    # a budget below any two requests: they run one at a time, and a request
    # larger than the whole budget still runs (alone)
    reqs = _requests()[:3]
    est = [ca.memory_estimate(c, ca.PIPELINE_BATCHED).device_bytes for c, _ in reqs]
    plain = [ca.memory_estimate(c).device_bytes for c, _ in reqs]
    assert all(e > p for e, p in zip(est, plain))  # the batched pipeline's staging is counted
    srv = Server(hbm_budget_bytes=min(est), max_concurrent=8)
    sinks = []
    for c, d in reqs:
        sinks.append(ca.MemSink())
        srv.submit(c, sinks[-1], d)
    assert srv.wait() == 0, ca.last_error()
    running, peak_bytes = srv.peaks()
    assert running == 1 and peak_bytes == max(est)
    reps = srv.reports()
    assert [r.id for r in reps] == [0, 1, 2]  # admitted in arrival order
    lat = srv.latencies()
    assert len(lat) == sum(r.epochs for r in reps) and min(lat) > 0
    srv.close()


def test_request_past_capacity_finishes_cleanly():
    # a bounded request handed more input than its capacity: the stream
    # delivers everything at capacity and says finished; the request counts
    # as a success and its shards equal the reference's on the same input
    from oracle import ref
    from tests.helpers import assert_same_shards, geometry, run_ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    cfg = ca.Config([Dim(24, 8, 2), Dim(32, 16, 1), Dim(32, 16, 2)], dtype="u16", codec="zstd")
    v = xor_volume((40, 32, 32))  # 16 frames past the 24-frame capacity
    srv = Server(hbm_budget_bytes=8 << 30, max_concurrent=2, pipeline=ca.PIPELINE_BATCHED)
    sink = ca.MemSink()
    srv.submit(cfg, sink, v)
    assert srv.wait() == 0, ca.last_error()
    reps = srv.reports()
    srv.close()
    assert len(reps) == 1 and reps[0].status == 0
    g, _, cb = geometry(cfg)
    rs, _, _ = run_ref(cfg, v[:24])
    assert_same_shards(sink.shards(), rs, lambda lv: g[lv][0], "zstd", cb, "past capacity")
