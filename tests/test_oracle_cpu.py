"""Pins the numpy restatement (oracle/chucky_oracle.py) on the CPU:
against the reference's own CPU path built from its sources (oracle/_ref),
against the committed golden shard files, and against closed forms."""
import json
import os
import struct

import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Dim
from oracle import chucky_oracle as O
from oracle import ref
from tests.helpers import random_typed, run_ref

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def test_crc32c_check_value():
    # standard CRC-32C check value (Castagnoli): "123456789" -> 0xE3069283
    assert O.crc32c(b"123456789") == 0xE3069283
    assert ca.crc32c(b"123456789") == 0xE3069283
    rng = np.random.default_rng(1)
    for n in (0, 1, 7, 8, 9, 63, 1000):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert O.crc32c(b) == ca.crc32c(b)


def test_morton_table_3x5():
    # compacted Morton order of a 3x5 box (docs/design.md): rows y, cols x
    want = [[0, 1, 4, 5, 12], [2, 3, 6, 7, 13], [8, 9, 10, 11, 14]]
    for y in range(3):
        for x in range(5):
            assert O.morton_rank([5, 3], [x, y]) == want[y][x]


@pytest.mark.parametrize("shape", [[3, 5], [4, 4], [7, 2], [5, 3, 2], [1, 9], [6, 6, 3]])
def test_morton_rank_brute_force(shape):
    # rank = number of in-bounds cells with a smaller interleaved Z key
    def zkey(c):
        k = 0
        for bit in range(8):
            for d in range(len(c)):
                k |= ((c[d] >> bit) & 1) << (bit * len(c) + d)
        return k
    cells = list(np.ndindex(*shape))
    keys = sorted(zkey(c) for c in cells)
    for c in cells:
        assert O.morton_rank(list(shape), list(c)) == keys.index(zkey(c))


def test_window_ends_cover_children():
    # windows of consecutive parents tile the child Morton range
    for child in ([6, 5], [4, 4], [7, 3], [3, 3, 5]):
        parent = [(c + 1) // 2 for c in child]
        e = O.window_ends(child, parent)
        assert e == sorted(e) and e[-1] == int(np.prod(child))
        lens = np.diff([0] + e)
        assert lens.min() >= 1 and lens.max() <= 2 ** len(child)


def test_fold_level2_chain_pinned():
    # tests/test_lod.c:919-928: level-2 mean fold of a constant v over its
    # 4-epoch period is v -> v/2 -> 3v/8 -> 11v/32 (shift-by-level fold)
    dt = np.dtype(np.uint16)
    v = np.array([3200], dtype=dt)
    acc = v.copy()
    for _ in range(3):
        acc = O.add_shift(acc, v, 2, dt)
    assert int(acc[0]) == 3200 * 11 // 32


def test_reduce_window_textbook():
    x = np.array([5, 1, 9, 3], dtype=np.uint16)
    assert O.reduce_window(x, O.MEAN) == 4
    assert O.reduce_window(x, O.MIN) == 1 and O.reduce_window(x, O.MAX) == 9
    assert O.reduce_window(x, O.MEDIAN) == 5       # sorted [1,3,5,9], n/2 = 2
    assert O.reduce_window(x, O.MAX_SUP) == 5 and O.reduce_window(x, O.MIN_SUP) == 3
    xi = np.array([-3, -4], dtype=np.int16)
    assert O.reduce_window(xi, O.MEAN) == -3      # C truncation toward zero


CASES = [
    ([Dim(4, 2, 2), Dim(4, 2, 2), Dim(6, 3, 2)], "u16", "mean", "mean", 2 * 48),
    ([Dim(6, 2, 1), Dim(8, 4, 1), Dim(9, 3, 1)], "u16", "mean", "mean", 432),
    ([Dim(20, 8, 2), Dim(16, 8, 1), Dim(24, 8, 3)], "u8", "mean", "mean", 20 * 16 * 24),
    ([Dim(8, 4, 0), Dim(12, 4, 0, storage_position=2), Dim(16, 8, 0, storage_position=1)],
     "i32", "mean", "mean", 8 * 12 * 16),
    ([Dim(8, 4, 1), Dim(12, 4, 0, downsample=True), Dim(10, 4, 0, downsample=True)],
     "u16", "median", "mean", 8 * 12 * 10),
    ([Dim(16, 2, 4, downsample=True), Dim(8, 4, 0, downsample=True), Dim(12, 4, 0, downsample=True)],
     "u8", "mean", "mean", 16 * 8 * 12),
    ([Dim(12, 2, 2, downsample=True), Dim(8, 4, 0, downsample=True), Dim(8, 4, 0, downsample=True)],
     "f32", "max_sup", "max", 12 * 8 * 8),
    ([Dim(10, 2, 2, downsample=True), Dim(6, 4, 0, downsample=True), Dim(10, 4, 0, downsample=True)],
     "i16", "min", "min", 10 * 6 * 10),
]


def _odims(dims):
    return [O.ODim(d.size, d.chunk, d.cps, d.downsample,
                   -1 if d.storage_position is None else d.storage_position) for d in dims]


@needs_ref
@pytest.mark.parametrize("case", range(len(CASES)))
def test_restatement_matches_reference(case):
    dims, dtype, red, ared, n = CASES[case]
    cfg = ca.Config(dims, dtype=dtype, reduce=red, append_reduce=ared, buffer_capacity_bytes=4096)
    data = random_typed(n, dtype, seed=100 + case)
    want, _, wcur = run_ref(cfg, data, pieces=3)
    got, gcur = O.write_stream(_odims(dims), ca.NP_DTYPE[dtype], data,
                               reduce_method=O.__dict__[{"mean": "MEAN", "min": "MIN", "max": "MAX",
                                                         "median": "MEDIAN", "max_sup": "MAX_SUP",
                                                         "min_sup": "MIN_SUP"}[red]],
                               append_method={"mean": O.MEAN, "min": O.MIN, "max": O.MAX}[ared])
    assert gcur == wcur
    assert set(got) == set(want)
    for k in want:
        assert got[k] == want[k], f"shard {k}"


def _golden_files():
    if not os.path.isdir(GOLDEN):
        return []
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("name", _golden_files())
def test_restatement_matches_golden(name):
    g = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    spec = json.loads(g["spec"].tobytes().decode())
    dims = [O.ODim(int(a), int(b), int(c), bool(d), int(e)) for a, b, c, d, e in spec["dims"]]
    data = g["input"]
    got, _ = O.write_stream(dims, data.dtype, data, reduce_method=int(spec["reduce"]),
                            append_method=int(spec["append"]))
    keys = [tuple(int(v) for v in k) for k in g["keys"]]
    assert sorted(got) == sorted(keys)
    for i, k in enumerate(keys):
        assert got[k] == g[f"shard{i}"].tobytes(), f"shard {k}"


def test_shard_index_layout_closed_form():
    # one epoch of a 2x2x2-chunk volume in one shard: index entries are
    # consecutive chunk-sized runs and the trailer is the CRC of the index
    dims = [O.ODim(2, 2, 1), O.ODim(4, 2, 0), O.ODim(4, 2, 0)]
    data = np.arange(32, dtype=np.uint16)
    out, _ = O.write_stream(dims, np.uint16, data)
    (k, blob), = out.items()
    idx = blob[-(16 * 4 + 4):-4]
    assert struct.unpack("<I", blob[-4:])[0] == O.crc32c(idx)
    ents = [struct.unpack_from("<QQ", idx, 16 * i) for i in range(4)]
    assert ents == [(0, 16), (16, 16), (32, 16), (48, 16)]


# The drain of a dim-0 downsampled stream, for every level whose fold period
# equals the batch length K, delivers pool slot K-1 as last compressed (the
# reference reuses its init-time batch LUT, src/cpu/pipeline.c:159-180): the
# previous emission, or empty chunks when no full batch ever ran.
DRAIN_CASES = [(40, 4), (40, 8), (40, 2), (8, 4), (12, 4), (16, 4), (24, 8), (20, 4)]


@needs_ref
@pytest.mark.parametrize("rows,K", DRAIN_CASES)
def test_drain_batch_lut_reuse_matches_reference(rows, K):
    dims = [Dim(rows, 4, 2, downsample=True), Dim(24, 4, 0, downsample=True),
            Dim(20, 4, 0, downsample=True)]
    cfg = ca.Config(dims, dtype="f32", epochs_per_batch=K, buffer_capacity_bytes=4096)
    data = random_typed(rows * 24 * 20, "f32", seed=33)
    want, _, _ = run_ref(cfg, data, pieces=3)
    got, _ = O.write_stream([O.ODim(rows, 4, 2, True), O.ODim(24, 4, 0, True),
                             O.ODim(20, 4, 0, True)], np.float32, data, K=K)
    assert set(got) == set(want)
    for k in want:
        assert got[k] == want[k], f"shard {k}"


def test_drain_without_full_batch_is_empty():
    # two epochs, K=4: level 2 (period 4) drains into a slot no batch compressed
    dims = [O.ODim(8, 4, 2, True), O.ODim(24, 4, 0, True), O.ODim(20, 4, 0, True)]
    data = np.arange(8 * 24 * 20, dtype=np.float32)
    out, _ = O.write_stream(dims, np.float32, data, K=4)
    blob = out[(2, 0)]
    assert len(blob) == 16 * 4 + 4  # index only
    idx = blob[:-4]
    assert all(struct.unpack_from("<QQ", idx, 16 * i) == (2**64 - 1, 2**64 - 1) for i in range(4))


# ---- f16: pinned against float16 of the f32 reference --------------------

def _chunk_arrays(shards, cps_of_level, np_dtype):
    """{(level, shard, slot): chunk values} of CODEC_NONE shards."""
    from tests.helpers import parse_shard
    out = {}
    for (lv, si), blob in shards.items():
        for k, ent in enumerate(parse_shard(blob, cps_of_level(lv))):
            if ent is not None:
                off, n = ent
                out[(lv, si, k)] = np.frombuffer(blob[off:off + n], dtype=np_dtype)
    return out


def _cps(dims):
    cps = O.Plan(_odims(dims), 2).cps_total
    return lambda lv: cps[lv]


F16_DIMS = [Dim(8, 4, 1), Dim(40, 8, 0, downsample=True), Dim(36, 8, 0, downsample=True)]


@needs_ref
@pytest.mark.parametrize("red", ["mean", "min", "max", "median", "max_sup", "min_sup"])
def test_f16_restatement_matches_f32_reference(red):
    n = 8 * 40 * 36
    x16 = (np.random.default_rng(5).standard_normal(n) * 100).astype(np.float16)
    cfg = ca.Config(F16_DIMS, dtype="f32", reduce=red)
    want, _, _ = run_ref(cfg, x16.astype(np.float32), pieces=2)
    got, _ = O.write_stream(_odims(F16_DIMS), np.float16, x16,
                            reduce_method=ca.REDUCE_ID[red])
    w = _chunk_arrays(want, _cps(F16_DIMS), np.float32)
    g = _chunk_arrays(got, _cps(F16_DIMS), np.float16)
    assert g.keys() == w.keys()
    levels = {k[0] for k in w}
    assert len(levels) >= 3
    for k in w:
        if red == "mean" and k[0] >= 2:
            continue  # level 2 averages rounded halfs: not a function of the f32 path
        np.testing.assert_array_equal(g[k], w[k].astype(np.float16), err_msg=str(k))


@needs_ref
@pytest.mark.parametrize("ared", ["mean", "min", "max"])
def test_f16_fold_exact_on_short_fractions(ared):
    # multiples of 256 below 2048: every sum, mean, fold and emit is a short
    # binary fraction, exact in half precision
    dims = [Dim(16, 2, 2, downsample=True), Dim(16, 4, 0, downsample=True),
            Dim(12, 4, 0, downsample=True)]
    n = 16 * 16 * 12
    x16 = (np.random.default_rng(9).integers(0, 8, n) * 256).astype(np.float16)
    cfg = ca.Config(dims, dtype="f32", reduce="mean", append_reduce=ared)
    want, _, _ = run_ref(cfg, x16.astype(np.float32), pieces=1)
    got, _ = O.write_stream(_odims(dims), np.float16, x16, reduce_method=O.MEAN,
                            append_method={"mean": O.MEAN, "min": O.MIN, "max": O.MAX}[ared])
    w = _chunk_arrays(want, _cps(dims), np.float32)
    g = _chunk_arrays(got, _cps(dims), np.float16)
    assert g.keys() == w.keys()
    for k in w:
        assert np.all(w[k] == w[k].astype(np.float16).astype(np.float32)), "f32 value not a half"
        np.testing.assert_array_equal(g[k], w[k].astype(np.float16), err_msg=str(k))


@pytest.mark.parametrize("dtype", ["u8", "i8", "u16", "i16", "u32", "i32", "u64", "i64",
                                   "f16", "f32", "f64"])
@pytest.mark.parametrize("method", range(6))
def test_vectorized_windows_match_scalar(dtype, method):
    # ragged windows of 1..8 cells, two batch rows, values over the full range
    rng = np.random.default_rng(31 + method)
    lens = rng.integers(1, 9, 60)
    ends = np.cumsum(lens)
    src = random_typed(2 * int(ends[-1]), dtype, seed=3 + method).reshape(2, -1)
    got = O.reduce_windows(src, ends, method)
    for b in range(2):
        for i, e in enumerate(ends):
            s = ends[i - 1] if i else 0
            want = O.reduce_window(src[b, s:e], method)
            assert got[b, i].tobytes() == np.array(want, dtype=src.dtype).tobytes(), (b, i)
