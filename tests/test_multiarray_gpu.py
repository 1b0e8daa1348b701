"""Multi-array writer over GPU tile streams (reference behaviour:
src/multiarray/stream.c:552-726 update, :866-887 flush; the reference's own
test tests/test_multiarray.c switches arrays at epoch boundaries and expects
multiarray_writer_not_flushable for a switch mid-epoch)."""
import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Dim
from tests.helpers import random_typed, run_gpu, xor_volume

ARRAYS = [
    (ca.Config([Dim(32, 8, 2), Dim(32, 16, 1), Dim(48, 16, 3)], dtype="u16", codec="zstd"),
     lambda: xor_volume((32, 32, 48))),
    (ca.Config([Dim(24, 4, 2), Dim(16, 8, 2), Dim(20, 8, 0)], dtype="u8", codec="none"),
     lambda: random_typed(24 * 16 * 20, "u8", seed=3)),
    (ca.Config([Dim(16, 4, 2, downsample=True), Dim(32, 8, 0, downsample=True),
                Dim(32, 8, 0, downsample=True)], dtype="f32", codec="lz4"),
     lambda: random_typed(16 * 32 * 32, "f32", seed=4)),
]


def test_multiarray_symbols_exported():
    for name in ("multiarray_tile_stream_gpu_create", "multiarray_tile_stream_gpu_destroy",
                 "multiarray_tile_stream_gpu_writer", "multiarray_tile_stream_gpu_get_metrics"):
        assert hasattr(ca.lib, name)


@pytest.mark.gpu
def test_multiarray_interleaved_epochs_match_single_streams():
    cfgs = [c for c, _ in ARRAYS]
    datas = [np.ascontiguousarray(make()).reshape(-1) for _, make in ARRAYS]
    sinks = [ca.MemSink() for _ in cfgs]
    m = ca.GpuMultiArray(cfgs, sinks)
    epoch = []
    for c in cfgs:
        s = ca.GpuStream(c, ca.MemSink())
        epoch.append(s.epoch_elements)
        s.close()
    # round robin, one epoch of each array per turn
    pos = [0] * len(cfgs)
    while any(p < d.size for p, d in zip(pos, datas)):
        for i, d in enumerate(datas):
            if pos[i] >= d.size:
                continue
            part = d[pos[i]:pos[i] + epoch[i]]
            err, used = m.update(i, part)
            assert err == 0 and used == part.nbytes, (i, err, used, ca.last_error())
            pos[i] += part.size
    assert m.flush() == 0, ca.last_error()
    met = m.metrics()
    assert met.scatter.count > 0
    m.close()
    for i, (c, d) in enumerate(zip(cfgs, datas)):
        want, _, _ = run_gpu(c, d, pieces=1)
        got = sinks[i].shards()
        assert set(got) == set(want), i
        for k in want:
            same = got[k] == want[k]
            assert same, (i, k)


@pytest.mark.gpu
def test_multiarray_switch_mid_epoch_not_flushable():
    cfgs = [c for c, _ in ARRAYS[:2]]
    datas = [np.ascontiguousarray(make()).reshape(-1) for _, make in ARRAYS[:2]]
    sinks = [ca.MemSink() for _ in cfgs]
    m = ca.GpuMultiArray(cfgs, sinks)
    s = ca.GpuStream(cfgs[0], ca.MemSink())
    e0 = s.epoch_elements
    s.close()
    half = e0 // 2
    assert m.update(0, datas[0][:half]) == (0, half * 2)
    # array 0 is mid-epoch: nothing of array 1 is consumed
    assert m.update(1, datas[1][:16]) == (3, 0)
    # an unknown array consumes nothing
    assert m.update(5, datas[1][:16]) == (1, 0)
    assert m.update(0, datas[0][half:e0]) == (0, (e0 - half) * 2)
    assert m.update(1, datas[1]) == (0, datas[1].nbytes)
    assert m.update(0, datas[0][e0:]) == (0, (datas[0].size - e0) * 2)
    assert m.flush() == 0, ca.last_error()
    m.close()
    for i, (c, d) in enumerate(zip(cfgs, datas)):
        want, _, _ = run_gpu(c, d, pieces=3)
        assert sinks[i].shards() == want, i


def _script(datas, epochs, seed):
    """A seeded sequence of (array, start, stop): whole epochs, partial
    epochs, switches away from a partial epoch, and appends past a bounded
    array's capacity."""
    rng = np.random.default_rng(seed)
    pos = [0] * len(datas)
    out = []
    for _ in range(40):
        i = int(rng.integers(len(datas)))
        e = epochs[i]
        kind = rng.integers(4)
        n = e if kind < 2 else (e // 2 if kind == 2 else 3 * e)
        out.append((i, pos[i], pos[i] + n))
        pos[i] += n
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_multiarray_matches_reference_multiarray(seed):
    # the GPU multi-array writer and the reference's multiarray_tile_stream_cpu
    # (oracle/_ref) take the same scripted updates: the same error codes
    # (including not_flushable for a switch away from a partial epoch and
    # finished past a bounded array's capacity), the same bytes consumed,
    # and after the final flush (active array mid-epoch) the same shards
    from oracle import ref
    from tests.helpers import assert_same_shards, geometry
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    cfgs = [c for c, _ in ARRAYS]
    datas = [np.ascontiguousarray(make()).reshape(-1) for _, make in ARRAYS]
    datas = [np.concatenate([d, d, d]) for d in datas]  # more than capacity
    epochs = []
    for c in cfgs:
        s = ca.GpuStream(c, ca.MemSink())
        epochs.append(s.epoch_elements)
        s.close()
    gsinks = [ca.MemSink() for _ in cfgs]
    rsinks = [ca.MemSink() for _ in cfgs]
    g = ca.GpuMultiArray(cfgs, gsinks)
    r = ref.RefMultiArray(cfgs, rsinks)
    codes = set()
    for i, a, b in _script(datas, epochs, seed):
        part = datas[i][a:b]
        ge, gu = g.update(i, part)
        re_, ru = r.update(i, part)
        assert (ge, gu) == (re_, ru), (i, a, b, ge, gu, re_, ru, ca.last_error())
        codes.add(ge)
    assert g.flush() == r.flush()
    g.close()
    r.close()
    assert {0, 3} <= codes, codes
    for i, c in enumerate(cfgs):
        geo, _, cb = geometry(c)
        assert_same_shards(gsinks[i].shards(), rsinks[i].shards(), lambda lv: geo[lv][0], c.codec, cb,
                           f"array {i}")
