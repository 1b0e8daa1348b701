"""Dtype / reduction sweep (the 512x1024x1024 sweep's cases at reduced size):
u8, i16, u32, f16, f32, f64 x mean, median, max_sup, min_sup x three chunk
shapes (cubes of 8 and 16, 4x32x32 slabs), every one with padded edges (dims
not divisible by the chunk) and five LOD levels, codecs none / lz4 / zstd in
turn. The GPU stream is compared with the reference's CPU path built from its
sources (oracle/_ref): CODEC_NONE shards byte for byte, compressed shards by
index layout and decoded chunks. f16 has no CPU path in the reference; it is
compared with the numpy restatement within 1 ULP."""
import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Dim
from oracle import chucky_oracle as O
from oracle import ref
from tests.helpers import decoded_chunks, parse_shard, random_typed, run_gpu, run_ref

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]

SHAPES = {
    # 15 chunks along y and x: levels 120, 60, 30, 15, 8
    "c8": [Dim(24, 8, 0), Dim(120, 8, 0, downsample=True), Dim(120, 8, 0, downsample=True)],
    # 16 chunks: 250, 125, 63, 32, 16
    "c16": [Dim(32, 16, 0), Dim(250, 16, 0, downsample=True), Dim(250, 16, 0, downsample=True)],
    # 16 chunks: 500, 250, 125, 63, 32
    "c4x32": [Dim(8, 4, 0), Dim(500, 32, 0, downsample=True), Dim(500, 32, 0, downsample=True)],
}
DTYPES = ["u8", "i16", "u32", "f32", "f64"]
METHODS = ["mean", "median", "max_sup", "min_sup"]
CODECS = ["none", "lz4", "zstd"]


def _data(dims, dtype, seed):
    n = int(np.prod([d.size for d in dims]))
    if dtype in ("f16", "f32", "f64"):
        rng = np.random.default_rng(seed)
        return (rng.standard_normal(n) * 100).astype(ca.NP_DTYPE[dtype])
    return random_typed(n, dtype, seed=seed)  # uniform over the full range


def _levels(cfg):
    s = ca.GpuStream(cfg, ca.MemSink())
    n = s.status().nlod
    s.close()
    return n


@pytest.mark.parametrize("shape", sorted(SHAPES))
@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("dtype", DTYPES)
def test_sweep_matches_reference(shape, method, dtype):
    case = (sorted(SHAPES).index(shape) * 31 + METHODS.index(method) * 7 + DTYPES.index(dtype))
    codec = CODECS[case % 3]
    dims = SHAPES[shape]
    cfg = ca.Config(dims, dtype=dtype, reduce=method, codec=codec)
    assert _levels(cfg) == 5
    data = _data(dims, dtype, seed=1000 + case)
    gs, gmeta, gcur = run_gpu(cfg, data, pieces=2)
    rs, rmeta, rcur = run_ref(cfg, data, pieces=2)
    assert gcur == rcur
    assert set(gs) == set(rs)
    for lv in gmeta:
        assert gmeta[lv] == rmeta[lv]
    if codec == "none":
        for k in rs:
            same = gs[k] == rs[k]
            assert same, f"shard {k} differs"
        return
    cps = O.Plan([O.ODim(d.size, d.chunk, d.cps, d.downsample) for d in dims], 1).cps_total
    chunk_bytes = int(np.prod([d.chunk for d in dims])) * cfg.np_dtype.itemsize
    for k in rs:
        gi = [e is None for e in parse_shard(gs[k], cps[k[0]])]
        ri = [e is None for e in parse_shard(rs[k], cps[k[0]])]
        assert gi == ri, f"index layout of shard {k} differs"
    gd = decoded_chunks(gs, lambda lv: cps[lv], codec, chunk_bytes)
    rd = decoded_chunks(rs, lambda lv: cps[lv], codec, chunk_bytes)
    assert gd.keys() == rd.keys()
    for k in rd:
        same = gd[k] == rd[k]
        assert same, f"decoded chunk {k} differs"


def _ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in units in the last place between two float16 arrays."""
    ia = a.view(np.int16).astype(np.int32)
    ib = b.view(np.int16).astype(np.int32)
    # map sign-magnitude bit patterns onto a monotonic integer line
    ia = np.where(ia < 0, -32768 - ia, ia)
    ib = np.where(ib < 0, -32768 - ib, ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("codec", ["none", "zstd"])
def test_sweep_f16_within_one_ulp(method, codec):
    dims = SHAPES["c8"]
    cfg = ca.Config(dims, dtype="f16", reduce=method, codec=codec)
    assert _levels(cfg) == 5
    data = _data(dims, "f16", seed=77 + METHODS.index(method))
    gs, _, gcur = run_gpu(cfg, data, pieces=2)
    want, wcur = O.write_stream([O.ODim(d.size, d.chunk, d.cps, d.downsample) for d in dims],
                                np.float16, data, reduce_method=ca.REDUCE_ID[method])
    assert gcur == wcur
    assert set(gs) == set(want)
    cps = O.Plan([O.ODim(d.size, d.chunk, d.cps, d.downsample) for d in dims], 2).cps_total
    chunk_bytes = int(np.prod([d.chunk for d in dims])) * 2
    gd = decoded_chunks(gs, lambda lv: cps[lv], codec, chunk_bytes)
    wd = decoded_chunks(want, lambda lv: cps[lv], "none", chunk_bytes)
    assert gd.keys() == wd.keys()
    worst = 0
    for k in wd:
        a = np.frombuffer(gd[k], dtype=np.float16)
        b = np.frombuffer(wd[k], dtype=np.float16)
        worst = max(worst, int(_ulp_diff(a, b).max()))
    # the stated tolerance is 1 ULP; the GPU performs the restatement's
    # operations in the same order, so it must be exact
    assert worst == 0, worst


F16_DIM0 = [Dim(24, 2, 2, downsample=True), Dim(40, 8, 0, downsample=True),
            Dim(36, 8, 0, downsample=True)]


@pytest.mark.parametrize("append_reduce", ["mean", "min", "max"])
@pytest.mark.parametrize("pipeline", [0, 1])
def test_f16_dim0_fold_emit_exact(append_reduce, pipeline):
    # f16 with the append dim downsampled: the fold (half adds) and the emit
    # (half sum / count) of every level, drained at the end, against the
    # restatement bit for bit
    dims = F16_DIM0
    cfg = ca.Config(dims, dtype="f16", reduce="mean", append_reduce=append_reduce, codec="none")
    n = int(np.prod([d.size for d in dims]))
    data = _data(dims, "f16", seed=91 + ["mean", "min", "max"].index(append_reduce))
    gs, _, gcur = run_gpu(cfg, data, pieces=3, pipeline=pipeline)
    want, wcur = O.write_stream([O.ODim(d.size, d.chunk, d.cps, d.downsample) for d in dims],
                                np.float16, data, reduce_method=O.MEAN,
                                append_method={"mean": O.MEAN, "min": O.MIN, "max": O.MAX}[append_reduce])
    assert gcur == wcur == n
    assert set(gs) == set(want)
    assert len({k[0] for k in want}) >= 3
    for k in want:
        same = gs[k] == want[k]
        assert same, f"shard {k} differs"
