"""GPU parity: the HIP write path against the reference's own CPU path
(oracle/_ref, built from the reference sources) on the same inputs.

CODEC_NONE shards must be byte-identical (the reference's definition of
parity, tests/test_cross_validate.c:168-622). With lz4/zstd the compressed
bytes may differ; every shard's index layout (which slots are filled), CRC and
every decoded chunk must match.
"""
import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Dim
from oracle import ref
from tests.helpers import (decoded_chunks, parse_shard, random_typed, run_gpu, run_ref,
                           thirds_volume, xor_volume)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]


def _geom(cfg):
    s = ca.GpuStream(cfg, ca.MemSink())
    n = s.status().nlod
    g = {lv: s.level_geometry(lv) for lv in range(n)}
    E = s.epoch_elements
    chunk_bytes = s.layout.chunk_elements * cfg.np_dtype.itemsize
    s.close()
    return g, E, chunk_bytes


def _where(got: bytes, want: bytes, cps_total: int, chunk_bytes: int, dt) -> str:
    """Names the slots and the first elements where two CODEC_NONE shards differ."""
    try:
        gi, ri = parse_shard(got, cps_total), parse_shard(want, cps_total)
    except AssertionError as e:
        return f"unparseable ({e})"
    if gi != ri:
        return f"index differs: got {gi} want {ri}"
    msgs = []
    for slot, ent in enumerate(ri):
        if ent is None:
            continue
        off, n = ent
        a = np.frombuffer(got[off:off + n], dtype=dt)
        b = np.frombuffer(want[off:off + n], dtype=dt)
        bad = np.flatnonzero(a != b)
        if bad.size:
            msgs.append(f"slot {slot}: {bad.size}/{a.size} differ, first at {bad[:4].tolist()} "
                        f"got {a[bad[:4]].tolist()} want {b[bad[:4]].tolist()}")
    return "; ".join(msgs) or "data equal, bytes outside chunks differ"


def _compare(cfg, data, pieces=1, device_input=False, pipeline=0, **options):
    g, _, chunk_bytes = _geom(cfg)
    gs, gmeta, gcur = run_gpu(cfg, data, pieces=pieces, device_input=device_input,
                              pipeline=pipeline, **options)
    rs, rmeta, rcur = run_ref(cfg, data, pieces=pieces)
    assert gcur == rcur
    assert set(gs) == set(rs), (sorted(gs), sorted(rs))
    for lv in gmeta:
        assert gmeta[lv] == rmeta[lv], (lv, gmeta[lv], rmeta[lv])
    if cfg.codec == "none":
        for k in rs:
            same = gs[k] == rs[k]  # a bool, so a failure does not diff whole shards
            assert same, f"shard {k} differs: " + _where(gs[k], rs[k], g[k[0]][0], chunk_bytes,
                                                         cfg.np_dtype)
        return gs
    cps = lambda lv: g[lv][0]  # noqa: E731
    for k in rs:
        gi = [e is None for e in parse_shard(gs[k], cps(k[0]))]
        ri = [e is None for e in parse_shard(rs[k], cps(k[0]))]
        assert gi == ri, f"index layout of shard {k} differs"
    gd = decoded_chunks(gs, cps, cfg.codec, chunk_bytes)
    rd = decoded_chunks(rs, cps, cfg.codec, chunk_bytes)
    assert gd.keys() == rd.keys()
    for k in rd:
        same = gd[k] == rd[k]
        assert same, f"decoded chunk {k} differs"
    return gs


def test_cross_validate_basic():
    # tests/test_cross_validate.c:168-252: 4x4x6, chunk 2x2x3, 2 epochs
    cfg = ca.Config([Dim(4, 2, 2), Dim(4, 2, 2), Dim(6, 3, 2)], dtype="u16",
                    buffer_capacity_bytes=4096)
    _, E, _ = _geom(cfg)
    data = (np.arange(2 * E) % 65536).astype(np.uint16)
    _compare(cfg, data)


def test_cross_validate_multishard():
    # tests/test_cross_validate.c:254-330: 6x8x9, chunk 2x4x3, cps 1
    cfg = ca.Config([Dim(6, 2, 1), Dim(8, 4, 1), Dim(9, 3, 1)], dtype="u16",
                    buffer_capacity_bytes=8192)
    _, E, _ = _geom(cfg)
    data = random_typed(3 * E, "u16", seed=3)
    _compare(cfg, data)


@pytest.mark.parametrize("codec", ["none", "lz4", "zstd"])
def test_volume_codecs(codec):
    v = xor_volume((64, 64, 64))
    cfg = ca.Config([Dim(64, 16, 2), Dim(64, 16, 2), Dim(64, 16, 4)], dtype="u16", codec=codec,
                    buffer_capacity_bytes=96 * 1024)
    _compare(cfg, v, pieces=5)


@pytest.mark.parametrize("codec", ["lz4", "zstd"])
def test_thirds_codecs(codec):
    v = thirds_volume((32, 96, 192))
    cfg = ca.Config([Dim(32, 16, 0), Dim(96, 32, 2), Dim(192, 64, 0)], dtype="u16", codec=codec)
    _compare(cfg, v, pieces=2)


def test_partial_last_epoch_and_small_buffer():
    cfg = ca.Config([Dim(20, 8, 2), Dim(16, 8, 1), Dim(24, 8, 3)], dtype="u8",
                    buffer_capacity_bytes=4096)
    data = random_typed(20 * 16 * 24, "u8", seed=11)
    _compare(cfg, data, pieces=7)


@pytest.mark.parametrize("pipeline", [0, 1])
def test_h2d_copy_kernel(pipeline):
    # staged pieces moved to the device by the copy kernel (h2d_copy = 1):
    # u8 pieces through a 4 KiB staging buffer start at every alignment
    cfg = ca.Config([Dim(20, 8, 2), Dim(16, 8, 1), Dim(24, 8, 3)], dtype="u8",
                    buffer_capacity_bytes=4096)
    data = random_typed(20 * 16 * 24, "u8", seed=11)
    _compare(cfg, data, pieces=7, pipeline=pipeline, h2d_copy=1)


@pytest.mark.parametrize("pipeline", [0, 1])
def test_h2d_copy_kernel_epochs(pipeline):
    # whole 8 MiB epochs of a multiscale stream by the copy kernel, the first
    # one in 2 MiB sub-ranges as it is staged
    v = xor_volume((128, 256, 256))
    cfg = ca.Config([Dim(128, 64, 2, name="z", downsample=True), Dim(256, 64, 4, name="y", downsample=True),
                     Dim(256, 64, 4, name="x", downsample=True)], dtype="u16", codec="zstd", level=1,
                    reduce="mean", append_reduce="mean", buffer_capacity_bytes=16 << 20)
    _compare(cfg, v, pieces=1, pipeline=pipeline, h2d_copy=1)


def test_device_input_matches_host_input():
    v = xor_volume((32, 64, 64))
    cfg = ca.Config([Dim(32, 8, 2), Dim(64, 32, 1), Dim(64, 32, 2)], dtype="u16", codec="zstd")
    _compare(cfg, v, pieces=3, device_input=True)


@pytest.mark.parametrize("dtype", ["u8", "u32", "i16", "f32", "f64", "u64"])
def test_dtypes_none(dtype):
    cfg = ca.Config([Dim(12, 4, 0), Dim(10, 5, 1), Dim(16, 8, 0)], dtype=dtype)
    data = random_typed(12 * 10 * 16, dtype, seed=5)
    _compare(cfg, data, pieces=2)


def test_storage_order_permuted():
    # storage order z, x, y (the append dim stays first)
    cfg = ca.Config([Dim(8, 4, 1), Dim(12, 4, 0, storage_position=2), Dim(16, 8, 0, storage_position=1)],
                    dtype="u16")
    data = random_typed(8 * 12 * 16, "u16", seed=9)
    _compare(cfg, data)


def test_shard_alignment_padding():
    cfg = ca.Config([Dim(16, 4, 2), Dim(32, 16, 1), Dim(32, 16, 2)], dtype="u16", codec="lz4",
                    shard_alignment=4096)
    v = xor_volume((16, 32, 32))
    _compare(cfg, v)


@pytest.mark.parametrize("reduce", ["mean", "min", "max", "median", "max_sup", "min_sup"])
def test_multiscale_inner(reduce):
    # LOD over y, x without dim0 downsampling: all levels every epoch (CPU path)
    cfg = ca.Config([Dim(16, 8, 1), Dim(48, 8, 0, downsample=True), Dim(40, 8, 0, downsample=True)],
                    dtype="u16", reduce=reduce)
    data = random_typed(16 * 48 * 40, "u16", seed=21)
    _compare(cfg, data, pieces=3)


@pytest.mark.parametrize("dtype", ["u8", "i8", "u16", "i32", "u32", "f32", "f64", "i64"])
@pytest.mark.parametrize("append_reduce", ["mean", "min", "max"])
def test_multiscale_dim0(dtype, append_reduce):
    # append dim downsampled: cross-epoch fold with the level-shift mean
    cfg = ca.Config([Dim(40, 4, 2, downsample=True), Dim(24, 4, 0, downsample=True),
                     Dim(20, 4, 0, downsample=True)],
                    dtype=dtype, append_reduce=append_reduce, reduce="mean")
    data = random_typed(40 * 24 * 20, dtype, seed=33)
    _compare(cfg, data, pieces=4)


BATCHED_CASES = {
    "volume_zstd": (lambda: xor_volume((64, 64, 64)),
                    [Dim(64, 16, 2), Dim(64, 16, 2), Dim(64, 16, 4)], "u16", "zstd", {}),
    "volume_lz4_k2": (lambda: xor_volume((64, 64, 64)),
                      [Dim(64, 16, 2), Dim(64, 16, 2), Dim(64, 16, 4)], "u16", "lz4",
                      {"epochs_per_batch": 2}),
    # 256 KiB chunks, 4 per epoch: 8 KiB segments, two epochs compress side by side
    "volume_lz4_seg": (lambda: xor_volume((128, 128, 128)),
                       [Dim(128, 32, 2), Dim(128, 64, 1), Dim(128, 64, 1)], "u16", "lz4",
                       {"epochs_per_batch": 2}),
    "partial_none": (lambda: random_typed(20 * 16 * 24, "u8", seed=11),
                     [Dim(20, 8, 2), Dim(16, 8, 1), Dim(24, 8, 3)], "u8", "none", {}),
    "lod_inner": (lambda: random_typed(16 * 48 * 40, "u16", seed=21),
                  [Dim(16, 8, 1), Dim(48, 8, 0, downsample=True), Dim(40, 8, 0, downsample=True)],
                  "u16", "zstd", {}),
    # 32 KiB chunks of several blocks, 3 levels emitted every epoch: with direct
    # launches each epoch is split (coarse chunks compress beside level 0)
    "lod_zstd_split": (lambda: np.clip(np.random.default_rng(6).poisson(100, 64 * 128 * 128), 0, 4095)
                       .astype(np.uint16),
                       [Dim(64, 16, 2), Dim(128, 32, 0, downsample=True), Dim(128, 32, 0, downsample=True)],
                       "u16", "zstd", {}),
    "lod_dim0": (lambda: random_typed(40 * 24 * 20, "f32", seed=33),
                 [Dim(40, 4, 2, downsample=True), Dim(24, 4, 0, downsample=True),
                  Dim(20, 4, 0, downsample=True)], "f32", "none", {"epochs_per_batch": 4}),
}


@pytest.mark.parametrize("name", sorted(BATCHED_CASES))
@pytest.mark.parametrize("device_input", [False, True])
@pytest.mark.parametrize("launch", [1, 2])
def test_batched_pipeline_matches_reference(name, device_input, launch):
    # launch 1: per-epoch compress and aggregate as captured graphs; 2: direct launches
    make, dims, dtype, codec, extra = BATCHED_CASES[name]
    cfg = ca.Config(dims, dtype=dtype, codec=codec, **extra)
    _compare(cfg, make(), pieces=3, device_input=device_input, pipeline=ca.PIPELINE_BATCHED,
             batched_launch=launch)


@pytest.mark.parametrize("rows,K,codec", [(40, 4, "none"), (40, 8, "lz4"), (12, 4, "none"),
                                          (24, 8, "zstd"), (40, 2, "none")])
@pytest.mark.parametrize("pipeline", [0, 1])
def test_drain_batch_lut_reuse(rows, K, codec, pipeline):
    # levels whose fold period equals K drain pool slot K-1 as last compressed
    cfg = ca.Config([Dim(rows, 4, 2, downsample=True), Dim(24, 4, 0, downsample=True),
                     Dim(20, 4, 0, downsample=True)], dtype="f32", codec=codec, epochs_per_batch=K)
    _compare(cfg, random_typed(rows * 24 * 20, "f32", seed=33), pieces=3, pipeline=pipeline)


LZ4_SEGMENT_DATA = {
    "xor": lambda: xor_volume((64, 64, 64)),
    "thirds": lambda: thirds_volume((64, 64, 64)),
    "random_u16": lambda: random_typed(64 * 64 * 64, "u16", seed=5),
    "sensor_like": lambda: np.clip(np.random.default_rng(3).poisson(100, 64 * 64 * 64), 0, 4095)
    .astype(np.uint16),
}


@pytest.mark.parametrize("name", sorted(LZ4_SEGMENT_DATA))
@pytest.mark.parametrize("pipeline", [0, 1])
def test_lz4_multi_wavefront_chunks(name, pipeline):
    # 128 KiB chunks, few per launch: each chunk is parsed in 8 KiB segments
    # by separate wavefronts and stitched into one LZ4 block
    cfg = ca.Config([Dim(64, 32, 2), Dim(64, 32, 2), Dim(64, 64, 1)], dtype="u16", codec="lz4")
    _compare(cfg, LZ4_SEGMENT_DATA[name](), pieces=2, pipeline=pipeline)


LZ4_SEGMENT_SHAPES = {
    # 128 KiB chunks: up to 128 segments a chunk (two wavefront windows of the layout pass)
    "pow2": ([Dim(64, 32, 2), Dim(64, 32, 2), Dim(64, 64, 1)], "u16", (64, 64, 64)),
    # 46080-byte chunks: a partial last segment
    "ragged": ([Dim(48, 24, 2), Dim(48, 24, 2), Dim(40, 40, 1)], "u16", (48, 48, 40)),
    # 1029-byte chunks: with 1 KiB segments the last one is too short to parse
    "tiny_tail": ([Dim(6, 3, 2), Dim(14, 7, 2), Dim(98, 49, 2)], "u8", (6, 14, 98)),
}


@pytest.mark.parametrize("shape", sorted(LZ4_SEGMENT_SHAPES))
@pytest.mark.parametrize("seg,hash_log", [(1024, 9), (2048, 10), (4096, 11), (8192, 12), (16384, 12),
                                          (32768, 13), (1 << 20, 9), (1 << 20, 12)])
@pytest.mark.parametrize("data", ["xor", "thirds"])
@pytest.mark.parametrize("pipeline", [0, 1])
def test_lz4_segment_sizes_and_hash_logs(shape, seg, hash_log, data, pipeline):
    # every LZ4 segment size (LDS-staged up to 16 KiB, parsed from HBM above,
    # one wavefront a chunk when the segment covers the chunk) and table size
    dims, dtype, vol = LZ4_SEGMENT_SHAPES[shape]
    v = (xor_volume if data == "xor" else thirds_volume)(vol)
    if dtype == "u8":
        v = (v & 0xFF).astype(np.uint8)
    cfg = ca.Config(dims, dtype=dtype, codec="lz4")
    _compare(cfg, v.reshape(-1), pieces=2, pipeline=pipeline, lz4_segment_bytes=seg,
             lz4_hash_log=hash_log)


ZSTD_MODE_DATA = {
    "xor": lambda: xor_volume((32, 128, 128)),
    "sensor_like": lambda: np.clip(np.random.default_rng(4).poisson(100, 32 * 128 * 128), 0, 4095)
    .astype(np.uint16),
    "thirds": lambda: thirds_volume((32, 128, 192))[:, :, :128].copy(),
    "random_u8": lambda: random_typed(32 * 128 * 128 * 2, "u8", seed=9),
}


@pytest.mark.parametrize("name", sorted(ZSTD_MODE_DATA))
@pytest.mark.parametrize("block,tables", [(4096, 2), (8192, 1), (8192, 2), (32768, 1), (131072, 1),
                                          (131072, 2)])
def test_zstd_block_sizes_and_table_modes(name, block, tables):
    # per-block Huffman tables (2) or one table per chunk carried by the first
    # block and reused as treeless literals (1), at several block sizes
    dt = "u8" if name == "random_u8" else "u16"
    shape_x = 256 if dt == "u8" else 128
    cfg = ca.Config([Dim(32, 16, 1), Dim(128, 64, 2), Dim(shape_x, 64 if dt == "u16" else 128, 0)],
                    dtype=dt, codec="zstd")
    data = ZSTD_MODE_DATA[name]()
    _compare(cfg, data, pieces=2, zstd_block_bytes=block, zstd_chunk_tables=tables)


def test_multiscale_dim0_zstd():
    v = xor_volume((64, 64, 64))
    cfg = ca.Config([Dim(64, 16, 2, downsample=True), Dim(64, 16, 0, downsample=True),
                     Dim(64, 16, 0, downsample=True)], dtype="u16", codec="zstd")
    _compare(cfg, v, pieces=2)


# LOD shapes that form full tiles (about 4-8 KiB: 64x64 u16, 32x32 u64,
# 16^3 u16, 8^3 f32, 4096 u16 in 1-D) for the tiled gather/scatter, next to
# ragged edges, tiles that straddle chunk boundaries, a permuted storage order
# and rows that are not 16-byte aligned.
TILE_CASES = {
    "2d_ragged": ([Dim(8, 4, 1), Dim(100, 64, 0, downsample=True), Dim(130, 64, 0, downsample=True)],
                  "u16", {"reduce": "mean"}),
    "2d_straddle_u8": ([Dim(4, 2, 1), Dim(192, 96, 0, downsample=True), Dim(192, 96, 0, downsample=True)],
                       "u8", {"reduce": "max"}),
    "2d_median_i16": ([Dim(4, 4, 1), Dim(70, 64, 0, downsample=True), Dim(66, 64, 0, downsample=True)],
                      "i16", {"reduce": "median"}),
    "2d_minsup_u64": ([Dim(4, 4, 1), Dim(64, 32, 0, downsample=True), Dim(96, 32, 0, downsample=True)],
                      "u64", {"reduce": "min_sup"}),
    "2d_maxsup_f32": ([Dim(4, 4, 1), Dim(64, 32, 0, downsample=True), Dim(64, 32, 0, downsample=True)],
                      "f32", {"reduce": "max_sup"}),
    "2d_permuted": ([Dim(4, 2, 1), Dim(128, 64, 0, downsample=True, storage_position=2),
                     Dim(160, 64, 0, downsample=True, storage_position=1)], "u16", {"reduce": "mean"}),
    "2d_odd_rows": ([Dim(4, 2, 1), Dim(64, 64, 0, downsample=True), Dim(65, 65, 0, downsample=True)],
                    "u16", {"reduce": "min"}),
    "3d_dim0_f32": ([Dim(32, 16, 2, downsample=True), Dim(40, 16, 0, downsample=True),
                     Dim(36, 16, 0, downsample=True)], "f32", {"reduce": "mean", "append_reduce": "mean"}),
    "3d_dim0_u16": ([Dim(32, 16, 2, downsample=True), Dim(40, 16, 0, downsample=True),
                     Dim(34, 16, 0, downsample=True)], "u16", {"reduce": "mean", "append_reduce": "mean"}),
    "3d_u8_median": ([Dim(32, 16, 2, downsample=True), Dim(32, 16, 0, downsample=True),
                      Dim(48, 16, 0, downsample=True)], "u8", {"reduce": "median", "append_reduce": "min"}),
    "3d_f64_max": ([Dim(24, 16, 2, downsample=True), Dim(16, 16, 0, downsample=True),
                    Dim(32, 16, 0, downsample=True)], "f64", {"reduce": "max", "append_reduce": "max"}),
    "1d": ([Dim(6, 2, 1), Dim(3, 3, 0), Dim(9000, 4096, 0, downsample=True)], "u16", {"reduce": "mean"}),
}


@pytest.mark.parametrize("name", sorted(TILE_CASES))
@pytest.mark.parametrize("pipeline", [0, 1])
def test_multiscale_tiles(name, pipeline):
    dims, dtype, extra = TILE_CASES[name]
    cfg = ca.Config(dims, dtype=dtype, **extra)
    n = int(np.prod([d.size for d in dims]))
    _compare(cfg, random_typed(n, dtype, seed=41), pieces=3, pipeline=pipeline)


@pytest.mark.parametrize("name", ["2d_ragged", "3d_dim0_f32", "2d_odd_rows", "1d"])
@pytest.mark.parametrize("pieces", [1, 3])
@pytest.mark.parametrize("pipeline", [0, 1])
def test_multiscale_device_epochs(name, pieces, pipeline):
    # device input: a whole epoch in one piece is gathered in place, pieces
    # that split epochs go through the stream's epoch buffer
    dims, dtype, extra = TILE_CASES[name]
    cfg = ca.Config(dims, dtype=dtype, **extra)
    n = int(np.prod([d.size for d in dims]))
    _compare(cfg, random_typed(n, dtype, seed=43), pieces=pieces, device_input=True, pipeline=pipeline)


@pytest.mark.parametrize("name", ["2d_ragged", "3d_dim0_u16"])
def test_multiscale_device_then_host(name):
    # whole epochs read in place from device memory, then the rest through
    # host staging: the cells of the last (partial) epoch that no append
    # reaches hold the previous epoch's values, as in the reference
    dims, dtype, extra = TILE_CASES[name]
    cfg = ca.Config(dims, dtype=dtype, **extra)
    n = int(np.prod([d.size for d in dims]))
    data = random_typed(n, dtype, seed=47)
    flat = np.ascontiguousarray(data).reshape(-1)
    sink = ca.MemSink()
    s = ca.GpuStream(cfg, sink)
    E = s.epoch_elements
    k = min(n, E * max(1, (n // E) - 1)) if n > E else 0
    if k:
        dev = ca.DeviceBuffer.from_host(flat[:k])
        assert s.append_device(dev.ptr, k * flat.itemsize) == 0, ca.last_error()
    assert s.append(flat[k:]) == 0, ca.last_error()
    assert s.flush() == 0, ca.last_error()
    gs = sink.shards()
    gcur = s.cursor
    s.close()
    rs, _, rcur = run_ref(cfg, data, pieces=1)
    assert gcur == rcur
    assert set(gs) == set(rs)
    for key in rs:
        same = gs[key] == rs[key]
        assert same, f"shard {key} differs"
