"""The reference's stream edge cases (tests/test_stream.c:466-816) and sink
failure propagation (tests/test_shard_sink.c), on the GPU stream in both
pipelines. Where the reference's CPU path (oracle/_ref) can run the same
sequence, its return codes, cursor and shards are the expected values."""
import ctypes as C

import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd import Slice
from chucky_amd.api import Dim
from oracle import ref
from tests.helpers import random_typed, run_ref, xor_volume

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]

PIPES = [ca.PIPELINE_PLAIN, ca.PIPELINE_BATCHED]


def _dims3():
    # tests/test_gpu_helpers.c:11-17: zyx 4x4x6, chunks 2x2x3, 2x2x2 chunks a shard
    return [Dim(4, 2, 2, name="z"), Dim(4, 2, 2, name="y"), Dim(6, 3, 2, name="x")]


def _dims3_unbounded(cps0=2):
    # tests/test_gpu_helpers.c:20-27: dim 0 unbounded with 2 chunks a shard
    return [Dim(0, 2, cps0, name="z"), Dim(4, 2, 2, name="y"), Dim(6, 3, 2, name="x")]


def _append(w_append, arr):
    buf = np.ascontiguousarray(arr)
    beg = buf.ctypes.data
    r = w_append(Slice(beg, beg + buf.nbytes))
    consumed = (r.rest.beg - beg) if r.rest.beg else buf.nbytes
    return r.error, consumed


class Both:
    """The same calls on the GPU stream and on the reference's CPU stream."""

    def __init__(self, cfg, pipeline):
        self.gsink, self.rsink = ca.MemSink(), ca.MemSink()
        self.g = ca.GpuStream(cfg, self.gsink, pipeline=pipeline)
        self.r = ref.RefCpuStream(cfg, self.rsink)

    def append(self, arr):
        ge = _append(lambda s: ca.lib.writer_append(self.g.w, s), arr)
        re_ = _append(lambda s: ref.lib().writer_append(self.r.w, s), arr)
        return ge, re_

    def flush(self):
        return self.g.flush(), self.r.flush()

    def check_same(self):
        assert self.g.cursor == self.r.cursor
        gs, rs = self.gsink.shards(), self.rsink.shards()
        assert set(gs) == set(rs)
        for k in rs:
            same = gs[k] == rs[k]
            assert same, f"shard {k} differs"
        return gs

    def close(self):
        self.g.close()
        self.r.close()


@pytest.mark.parametrize("pipeline", PIPES)
def test_zero_length_append(pipeline):
    b = Both(ca.Config(_dims3(), dtype="u16", buffer_capacity_bytes=96 * 2), pipeline)
    (ge, gc), (re_, rc) = b.append(np.zeros(0, dtype=np.uint16))
    assert ge == re_ == 0 and gc == rc == 0
    assert b.g.cursor == 0
    (ge, _), (re_, _) = b.append(np.arange(96, dtype=np.uint16))
    assert ge == re_ == 0
    assert b.flush() == (0, 0)
    gs = b.check_same()
    assert len(gs[(0, 0)]) > 0
    b.close()


def test_create_rejects_null_sink():
    cfg = ca.Config([Dim(4, 2), Dim(6, 3)], dtype="u16", buffer_capacity_bytes=24 * 2)
    c = cfg.to_c()
    assert not ca.lib.tile_stream_gpu_create(C.byref(c), None)
    assert "sink" in ca.last_error()


@pytest.mark.parametrize("pipeline", PIPES)
def test_rank_1(pipeline):
    b = Both(ca.Config([Dim(12, 4)], dtype="u16", buffer_capacity_bytes=12 * 2), pipeline)
    (ge, gc), (re_, rc) = b.append(np.arange(12, dtype=np.uint16))
    assert ge == re_ == 0 and gc == rc == 24
    assert b.flush() == (0, 0)
    gs = b.check_same()
    assert len(gs[(0, 0)]) > 0
    b.close()


@pytest.mark.parametrize("pipeline", PIPES)
def test_flush_empty(pipeline):
    b = Both(ca.Config([Dim(4, 2), Dim(6, 3)], dtype="u16", buffer_capacity_bytes=24 * 2), pipeline)
    assert b.flush() == (0, 0)
    assert b.g.cursor == 0
    b.check_same()
    b.close()


@pytest.mark.parametrize("pipeline", PIPES)
@pytest.mark.parametrize("codec", ["none", "zstd"])
def test_unbounded_dim0(pipeline, codec):
    cfg = ca.Config(_dims3_unbounded(), dtype="u16", codec=codec, buffer_capacity_bytes=96 * 2)
    s = ca.GpuStream(cfg, ca.MemSink(), pipeline=pipeline)
    assert s.layout.chunks_per_epoch == 4 and s.epoch_elements == 48
    s.close()
    data = (np.arange(4 * 48) % 65536).astype(np.uint16)
    if codec == "none":
        b = Both(cfg, pipeline)
        (ge, _), (re_, _) = b.append(data)
        assert ge == re_ == 0
        assert b.flush() == (0, 0)
        gs = b.check_same()
        assert len(gs[(0, 0)]) > 0
        b.close()
    else:
        from tests.helpers import assert_same_shards, geometry, run_gpu
        g, _, cb = geometry(cfg)
        gs, _, gcur = run_gpu(cfg, data, pieces=3, pipeline=pipeline)
        rs, _, rcur = run_ref(cfg, data, pieces=3)
        assert gcur == rcur == data.size
        assert_same_shards(gs, rs, lambda lv: g[lv][0], codec, cb, "unbounded zstd")


def test_unbounded_requires_chunks_per_shard():
    cfg = ca.Config(_dims3_unbounded(cps0=0), dtype="u16", buffer_capacity_bytes=96 * 2)
    c = cfg.to_c()
    sink = ca.MemSink()
    assert not ca.lib.tile_stream_gpu_create(C.byref(c), sink.shard_sink)
    assert not ref.lib().tile_stream_cpu_create(C.byref(c), sink.shard_sink)


@pytest.mark.parametrize("pipeline", PIPES)
@pytest.mark.parametrize("device", [False, True])
def test_bounded_capacity_finishes(pipeline, device):
    # tests/test_stream.c:727-784: 150 elements into a 96-element stream
    cfg = ca.Config(_dims3(), dtype="u16", buffer_capacity_bytes=96 * 2)
    src = np.arange(150, dtype=np.uint16)
    b = Both(cfg, pipeline)
    if device:
        dev = ca.DeviceBuffer.from_host(src)
        r = ca.lib.tile_stream_gpu_append_device(b.g.h, dev.ptr, src.nbytes)
        ge, gc = r.error, (r.rest.beg - dev.ptr) if r.rest.beg else src.nbytes
        re_, rc = _append(lambda s: ref.lib().writer_append(b.r.w, s), src)
    else:
        (ge, gc), (re_, rc) = b.append(src)
    assert ge == re_ == 2  # writer_error_finished
    assert gc == rc == 96 * 2
    # the stream delivered everything at capacity; the caller's flush then
    # succeeds (src/cpu/stream.c:545-548, 764-771) and later appends say finished
    assert b.flush() == (0, 0)
    (ge, gc), (re_, rc) = b.append(src[:4])
    assert ge == re_ == 2 and gc == rc == 0
    gs = b.check_same()
    assert len(gs[(0, 0)]) > 0
    b.close()


@pytest.mark.parametrize("pipeline", PIPES)
def test_append_after_flush_is_finished(pipeline):
    b = Both(ca.Config(_dims3_unbounded(), dtype="u16", buffer_capacity_bytes=96 * 2), pipeline)
    (ge, _), (re_, _) = b.append(np.arange(60, dtype=np.uint16))
    assert ge == re_ == 0
    assert b.flush() == (0, 0)
    (ge, gc), (re_, rc) = b.append(np.arange(10, dtype=np.uint16))
    assert ge == re_ == 2 and gc == rc == 0
    b.check_same()
    b.close()


@pytest.mark.parametrize("pipeline", PIPES)
@pytest.mark.parametrize("fail", ["open", "write"])
def test_sink_failure_propagates(pipeline, fail):
    # a sink that cannot open a shard or write to it fails the append that
    # delivers (or the flush), and the stream stays usable for destroy
    cfg = ca.Config([Dim(16, 4, 2), Dim(32, 16, 1), Dim(32, 16, 2)], dtype="u16", codec="zstd",
                    epochs_per_batch=1)
    sink = ca.MemSink()
    sink.inject_failure(fail_open=(fail == "open"), fail_write=(fail == "write"))
    s = ca.GpuStream(cfg, sink, pipeline=pipeline)
    v = xor_volume((16, 32, 32))
    errs = [s.append(v[i:i + 4]) for i in range(0, 16, 4)]
    errs.append(s.flush())
    assert any(e == 1 for e in errs), errs
    s.close()


@pytest.mark.parametrize("pipeline", PIPES)
@pytest.mark.parametrize("alignment", [512, 4096])
def test_shard_alignment_bytes(pipeline, alignment):
    # CODEC_NONE with page-aligned shards (src/zarr/shard_delivery.c): the
    # same file lengths, chunk offsets, chunk bytes, zero pad before the index,
    # index and CRC as the reference. Each run is written up to its page
    # boundary; the bytes past a run's data are whatever the reference's
    # malloc'd aggregate buffer held there (src/cpu/stream.c:105), so only ours
    # are checked there: zeros.
    cfg = ca.Config([Dim(20, 4, 2), Dim(24, 8, 2), Dim(40, 16, 2)], dtype="u16", codec="none",
                    shard_alignment=alignment, epochs_per_batch=2)
    data = random_typed(20 * 24 * 40, "u16", seed=71)
    b = Both(cfg, pipeline)
    for part in np.array_split(data, 3):
        (ge, _), (re_, _) = b.append(part)
        assert ge == re_ == 0
    assert b.flush() == (0, 0)
    assert b.g.cursor == b.r.cursor
    gs, rs = b.gsink.shards(), b.rsink.shards()
    assert set(gs) == set(rs)
    from tests.helpers import geometry
    g, _, _ = geometry(cfg)
    cps = g[0][0]
    tail = 16 * cps + 4
    for k in rs:
        got, want = gs[k], rs[k]
        assert len(got) == len(want) and len(got) % alignment == 0, k
        assert got[-tail:] == want[-tail:], f"index of shard {k} differs"
        pad0 = len(got) - align_up(tail, alignment)
        covered = np.zeros(len(got), dtype=bool)
        covered[pad0:] = True
        assert got[pad0:-tail] == want[pad0:-tail] == bytes(len(got) - pad0 - tail)
        for slot in range(cps):
            off, n = np.frombuffer(got[-tail:-4], dtype="<u8").reshape(-1, 2)[slot]
            if off == 2**64 - 1:
                continue
            assert got[off:off + n] == want[off:off + n], f"shard {k} slot {slot}"
            covered[off:off + n] = True
        assert not np.frombuffer(got, dtype=np.uint8)[~covered].any(), f"shard {k}: nonzero padding"
    b.close()


def align_up(v, a):
    return -(-v // a) * a


@pytest.mark.parametrize("pipeline", PIPES)
@pytest.mark.parametrize("device", [False, True])
def test_partial_last_epoch_reuses_slot(pipeline, device):
    # 7 epochs and a partial one with 2 epochs a batch: the last epoch's pool
    # slot held an earlier epoch; the cells past the cursor keep that epoch's
    # values, as in the reference's CPU path (it clears the pool only when
    # chunk strides are padded, src/cpu/stream.c:620-627)
    cfg = ca.Config([Dim(30, 4, 2), Dim(8, 4, 1), Dim(12, 4, 0)], dtype="u8", epochs_per_batch=2)
    data = random_typed(30 * 8 * 12, "u8", seed=73)
    from tests.helpers import run_gpu
    gs, _, gcur = run_gpu(cfg, data, pieces=4, pipeline=pipeline, device_input=device)
    rs, _, rcur = run_ref(cfg, data, pieces=4)
    assert gcur == rcur
    assert set(gs) == set(rs)
    for k in rs:
        same = gs[k] == rs[k]
        assert same, f"shard {k} differs"


def test_drop_in_create_and_advise_chunk_sizes():
    # the exact reference entry points: tile_stream_gpu_create (no options)
    # driven through the writer vtable, and advise_chunk_sizes halving the
    # chunk target until the memory estimate fits the budget
    # (src/gpu/stream.init.c:490-511)
    dims = [Dim(64, 16, 2), Dim(64, 32, 2), Dim(64, 32, 2)]
    cfg = ca.Config(dims, dtype="u16", codec="lz4")
    c = cfg.to_c()
    sink = ca.MemSink()
    h = ca.lib.tile_stream_gpu_create(C.byref(c), sink.shard_sink)
    assert h, ca.last_error()
    w = ca.lib.tile_stream_gpu_writer(h)
    v = xor_volume((64, 64, 64))
    beg = v.ctypes.data
    assert ca.lib.writer_append(w, Slice(beg, beg + v.nbytes)).error == 0
    assert ca.lib.writer_flush(w).error == 0
    assert ca.lib.tile_stream_gpu_cursor(h) == v.size
    ca.lib.tile_stream_gpu_destroy(h)
    rs, _, _ = run_ref(cfg, v)
    from tests.helpers import assert_same_shards, geometry
    g, _, cb = geometry(cfg)
    assert_same_shards(sink.shards(), rs, lambda lv: g[lv][0], "lz4", cb, "drop-in create")

    adv = ca.Config([Dim(0, 1, 16), Dim(2048, 2048, 0), Dim(2048, 2048, 0)], dtype="u16", codec="zstd")
    ac = adv.to_c()
    ratios = (C.c_uint8 * 3)(0, 1, 1)
    budget = 2 << 30
    assert ca.lib.tile_stream_gpu_advise_chunk_sizes(C.byref(ac), 64 << 20, ratios, budget) == 0
    chunks = [ac.dimensions[i].chunk_size for i in range(3)]
    info = ca.MemoryInfo()
    assert ca.lib.tile_stream_gpu_memory_estimate(C.byref(ac), C.byref(info)) == 0
    assert info.device_bytes <= budget
    assert int(np.prod(chunks)) * 2 <= 64 << 20
    assert ca.lib.tile_stream_gpu_advise_chunk_sizes(C.byref(ac), 64 << 20, ratios, 1) != 0


@pytest.mark.parametrize("pipeline", PIPES)
@pytest.mark.parametrize("codec", ["none", "zstd"])
def test_zero_copy_writes_outlive_the_next_copy(pipeline, codec):
    # a sink that reads zero-copy buffers only when the level's fence is
    # waited for (as the file sink's I/O queue does): every level's bytes
    # must still be intact then, across batches and levels
    cfg = ca.Config([Dim(48, 4, 2, downsample=True), Dim(64, 16, 0, downsample=True),
                     Dim(48, 16, 0, downsample=True)], dtype="u16", codec=codec, epochs_per_batch=2)
    data = random_typed(48 * 64 * 48, "u16", seed=75)
    sink = ca.MemSink(deferred=True)
    s = ca.GpuStream(cfg, sink, pipeline=pipeline)
    for part in np.array_split(data, 5):
        assert s.append(part) == 0, ca.last_error()
    assert s.flush() == 0, ca.last_error()
    nlod = s.status().nlod
    s.close()
    assert nlod == 3
    from tests.helpers import assert_same_shards, geometry
    g, _, cb = geometry(cfg)
    rs, _, _ = run_ref(cfg, data, pieces=2)
    assert_same_shards(sink.shards(), rs, lambda lv: g[lv][0], codec, cb, "deferred sink")


@pytest.mark.parametrize("opts,msg", [({"lz4_hash_log": 8}, "lz4_hash_log"),
                                      ({"lz4_hash_log": 14}, "lz4_hash_log"),
                                      ({"lz4_segment_bytes": 512}, "lz4_segment_bytes"),
                                      ({"lz4_segment_bytes": 3000}, "lz4_segment_bytes")])
def test_create_rejects_bad_lz4_options(opts, msg):
    cfg = ca.Config([Dim(16, 8, 1), Dim(64, 64, 1), Dim(64, 64, 1)], dtype="u16", codec="lz4")
    sink = ca.MemSink()
    with pytest.raises(RuntimeError, match=msg):
        ca.GpuStream(cfg, sink, **opts)
    sink.close()
