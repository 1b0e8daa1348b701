"""Host-side checks that need no GPU: the C-ABI library loads and exports every
function the public headers declare; dimension builders, memory estimates and
plan geometry agree with the reference's CPU build; sinks behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd import Dimension, TileStreamConfiguration
from chucky_amd.api import Dim
from oracle import ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def _declared_functions():
    names = set()
    for f in os.listdir(INCLUDE):
        if not f.endswith(".h"):
            continue
        txt = open(os.path.join(INCLUDE, f)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"static inline[^{]*\{.*?\n\}", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", txt, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_drop_in_api():
    names = _declared_functions()
    for fn in ("tile_stream_gpu_create", "tile_stream_gpu_destroy", "tile_stream_gpu_memory_estimate",
               "tile_stream_gpu_advise_chunk_sizes", "tile_stream_gpu_get_metrics",
               "tile_stream_gpu_layout", "tile_stream_gpu_writer", "tile_stream_gpu_cursor",
               "tile_stream_gpu_status", "writer_append", "writer_flush", "writer_append_wait",
               "dims_create", "dims_n_append", "dims_validate", "zarr_fs_sink_create",
               "multiarray_tile_stream_gpu_create", "multiarray_tile_stream_gpu_destroy",
               "multiarray_tile_stream_gpu_writer", "multiarray_tile_stream_gpu_get_metrics"):
        assert fn in names, fn


@pytest.mark.parametrize("fn", _declared_functions())
def test_library_exports_declared_symbol(fn):
    assert hasattr(ca.lib, fn), f"{fn} declared in include/ but not exported"


def _dims_array(spec):
    arr = (Dimension * len(spec))()
    for i, (size, chunk, cps, ds) in enumerate(spec):
        arr[i].size, arr[i].chunk_size, arr[i].chunks_per_shard = size, chunk, cps
        arr[i].downsample = ds
        arr[i].storage_position = i
        arr[i].name = b"tzyxcw"[i:i + 1]
    return arr


@needs_ref
def test_dims_builders_match_reference():
    R = ref.lib()
    for name in ("dims_n_append", "dims_validate"):
        getattr(R, name).argtypes = [C.POINTER(Dimension), C.c_uint8]
        getattr(ca.lib, name).argtypes = [C.POINTER(Dimension), C.c_uint8]
    R.dims_budget_chunk_size.argtypes = [C.POINTER(Dimension), C.c_uint8, C.c_uint64, C.POINTER(C.c_uint8)]
    ca.lib.dims_budget_chunk_size.argtypes = R.dims_budget_chunk_size.argtypes
    R.dims_set_shard_counts.argtypes = [C.POINTER(Dimension), C.c_uint8, C.POINTER(C.c_uint64)]
    ca.lib.dims_set_shard_counts.argtypes = R.dims_set_shard_counts.argtypes
    rng = np.random.default_rng(0)
    for _ in range(200):
        rank = int(rng.integers(1, 6))
        spec = [(int(rng.integers(1, 300)), int(rng.choice([1, 1, 2, 8, 16, 64])), int(rng.integers(0, 4)),
                 int(rng.integers(0, 2))) for _ in range(rank)]
        a, b = _dims_array(spec), _dims_array(spec)
        assert ca.lib.dims_n_append(a, rank) == R.dims_n_append(b, rank)
        assert ca.lib.dims_validate(a, rank) == R.dims_validate(b, rank)
        ratios = (C.c_uint8 * rank)(*[int(v) for v in rng.integers(0, 4, rank)])
        nelem = int(rng.integers(1, 1 << 24))
        ca.lib.dims_budget_chunk_size(a, rank, nelem, ratios)
        R.dims_budget_chunk_size(b, rank, nelem, ratios)
        assert [a[i].chunk_size for i in range(rank)] == [b[i].chunk_size for i in range(rank)]
        counts = (C.c_uint64 * rank)(*[int(v) for v in rng.integers(0, 4, rank)])
        ca.lib.dims_set_shard_counts(a, rank, counts)
        R.dims_set_shard_counts(b, rank, counts)
        assert [a[i].chunks_per_shard for i in range(rank)] == [b[i].chunks_per_shard for i in range(rank)]


class CpuMemInfo(C.Structure):
    _fields_ = [(n, C.c_size_t) for n in ("heap_bytes", "chunk_pool_bytes", "compressed_pool_bytes",
                                          "comp_sizes_bytes", "aggregate_bytes", "lod_bytes",
                                          "shard_bytes")] + \
               [("chunks_per_epoch", C.c_uint64), ("total_chunks", C.c_uint64),
                ("max_output_size", C.c_size_t), ("nlod", C.c_int), ("epochs_per_batch", C.c_uint32)]


GEOMETRY = [
    [Dim(512, 64, 1), Dim(4096, 256, 8), Dim(2304, 256, 9)],                       # sensor frames
    [Dim(256, 64, 4), Dim(256, 64, 4), Dim(256, 64, 4)],                           # 256^3 volume
    [Dim(256, 64, 4, downsample=True), Dim(256, 64, 4, downsample=True),
     Dim(256, 64, 4, downsample=True)],                                            # nlod = 3 rule
    [Dim(0, 1, 64), Dim(256, 8, 8, downsample=True), Dim(256, 8, 8, downsample=True),
     Dim(256, 8, 8, downsample=True)],                                             # 6 levels
    [Dim(512, 16, 0), Dim(1024, 128, 0, downsample=True), Dim(1024, 128, 0, downsample=True)],
    [Dim(512, 32, 0), Dim(1000, 32, 0, downsample=True), Dim(1000, 32, 0, downsample=True)],
]


@needs_ref
@pytest.mark.parametrize("g", range(len(GEOMETRY)))
def test_memory_estimate_geometry_matches_reference(g):
    cfg = ca.Config(GEOMETRY[g], dtype="u16", codec="zstd", max_nlod=0)
    c = cfg.to_c()
    mine = ca.memory_estimate(cfg)
    R = ref.lib()
    R.tile_stream_cpu_memory_estimate.argtypes = [C.POINTER(TileStreamConfiguration), C.POINTER(CpuMemInfo)]
    info = CpuMemInfo()
    assert R.tile_stream_cpu_memory_estimate(C.byref(c), C.byref(info)) == 0
    assert (mine.chunks_per_epoch, mine.total_chunks, mine.nlod, mine.epochs_per_batch) == \
        (info.chunks_per_epoch, info.total_chunks, info.nlod, info.epochs_per_batch)


def test_known_geometry_answers():
    # 256^3 with 64^3 chunks stops at 3 levels (level dims fit one chunk), not 6
    m = ca.memory_estimate(ca.Config(GEOMETRY[2], dtype="u16"))
    assert m.nlod == 3 and m.total_chunks == 16 + 4 + 1
    # 512 frames of 4096x2304: 144 chunks per epoch, K = nextpow2(ceil(1024/144)) = 8
    m = ca.memory_estimate(ca.Config(GEOMETRY[0], dtype="u16", codec="zstd"))
    assert m.chunks_per_epoch == 144 and m.epochs_per_batch == 8 and m.nlod == 1


def test_invalid_configs_rejected():
    bad = [
        ca.Config([Dim(10, 0, 0)]),                                   # chunk 0
        ca.Config([Dim(0, 4, 0), Dim(8, 4, 0)]),                      # unbounded without cps
        ca.Config([Dim(8, 4, 0), Dim(0, 4, 0)]),                      # inner unbounded
        ca.Config([Dim(8, 4, 0), Dim(8, 4, 0)], codec="lz4", level=0),  # lz4 level 0
        ca.Config([Dim(8, 1, 0, downsample=True), Dim(8, 4, 0)], append_reduce="median"),
    ]
    for cfg in bad:
        with pytest.raises(ValueError):
            ca.memory_estimate(cfg)


def test_mem_sink_records_and_fails_on_request():
    s = ca.MemSink()
    sink = C.cast(s.shard_sink, C.c_void_p)
    assert sink.value
    s.inject_failure(fail_open=True)
    assert s.shards() == {}


def test_crc32c_matches_reference_table():
    data = bytes(range(256)) * 5
    from oracle.chucky_oracle import crc32c
    assert ca.crc32c(data) == crc32c(data)


def test_autotuner_choice_varies_with_dtype_and_chunk_shape():
    # rows recorded by the sweep in chucky_amd/tune.py (MI355X): the choice
    # depends on the dtype size, the chunk bytes and the chunks per launch
    z = {k: ca.tuned_codec("zstd", *k) for k in [(1, 262144, 256), (2, 524288, 256), (4, 1048576, 256),
                                                 (2, 524288, 16), (2, 524288, 64)]}
    assert z[(1, 262144, 256)] == (32768, 0)
    assert z[(2, 524288, 256)] == (65536, 0)
    assert z[(4, 1048576, 256)] == (8192, 1)
    assert z[(2, 524288, 16)] == (8192, 0)
    assert z[(2, 524288, 64)] == (16384, 0)
    lz = {k: ca.tuned_codec("lz4", *k) for k in [(1, 262144, 256), (2, 524288, 64), (4, 1048576, 256)]}
    assert lz[(1, 262144, 256)] == (4096, 9)
    assert lz[(2, 524288, 64)] == (2048, 9)
    assert lz[(4, 1048576, 256)] == (1024, 9)
    # multiscale streams: every level in one launch, the coarse levels' dense
    # sequences set the tail, so smaller blocks and segments win
    assert ca.tuned_codec("zstd", 2, 524288, 84, multiscale=True) == (8192, 0)
    assert ca.tuned_codec("lz4", 2, 524288, 84, multiscale=True) == (1024, 9)
    # between rows: the nearest on a log scale; a segment never covers the chunk
    assert ca.tuned_codec("zstd", 2, 600000, 200) == (65536, 0)
    assert ca.tuned_codec("lz4", 2, 1024, 64)[0] == 0
    for codec in ("zstd", "lz4"):
        with pytest.raises(ValueError):
            ca.tuned_codec(codec, 0, 1024, 1)
    with pytest.raises(ValueError):
        ca.tuned_codec("none", 2, 1024, 1)


def test_autotuner_tables_match_the_recorded_sweep():
    # the C++ tables are the choices tune.py derives from the committed sweep
    from chucky_amd import tune
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    rows = tune.load(os.path.join(root, "chucky_amd", "tune_sweep_mi355x.csv"))
    src = open(os.path.join(root, "chucky_amd", "csrc", "autotune.cpp")).read()
    for codec, name in (("lz4", "kLz4Table"), ("zstd", "kZstdTable")):
        body = src[src.index(name + "[] = {"):]
        body = body[:body.index("};")]
        got = {tuple(int(x) for x in m[:4]): tuple(int(x) for x in m[4:])
               for m in re.findall(r"\{ (\d+), (\d+), (\d+), (\d+), (\d+), (\d+) \}", body)}
        want = {}
        for (lod, bpe, cb, pl), (o, _, _) in tune.choose([r for r in rows if r["codec"] == codec], codec).items():
            if codec == "lz4":
                seg = o["lz4_segment_bytes"]
                want[(lod, bpe, cb, pl)] = (0 if seg >= cb else seg, o["lz4_hash_log"])
            else:
                want[(lod, bpe, cb, pl)] = (o["zstd_block_bytes"], int(o["zstd_chunk_tables"] == 1))
        assert got == want, codec
