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


@pytest.mark.parametrize("shape", [[3, 5], [4, 4], [7, 2], [5, 3, 2], [1, 9], [6, 6, 3], [13, 11],
                                   [1000, 17], [9, 5, 7]])
def test_morton_rank_array_brute_force(shape):
    # the array form against the brute-force rank, every cell at once
    cells = np.stack(np.unravel_index(np.arange(int(np.prod(shape))), shape))
    nd = len(shape)
    key = np.zeros(cells.shape[1], dtype=np.int64)
    for bit in range(11):
        for d in range(nd):
            key |= ((cells[d] >> bit) & 1) << (bit * nd + d)
    want = np.argsort(np.argsort(key, kind="stable"), kind="stable")
    np.testing.assert_array_equal(O.morton_rank_array(shape, cells), want)


@pytest.mark.parametrize("child", [[6, 5], [4, 4], [7, 3], [3, 3, 5], [25, 19], [9, 8, 5], [1, 7]])
def test_window_ends_array_matches_scalar(child):
    parent = [(c + 1) // 2 for c in child]
    np.testing.assert_array_equal(O.window_ends_array(child, parent), O.window_ends(child, parent))
    # depth-1 ranks of coordinates past the box (the z_increment carry)
    c = np.array([[0, 1, 3, 2], [1, 1, 0, 3]])
    for j in range(c.shape[1]):
        inc = O.z_increment(list(c[:, j]), 2)
        assert list(O.z_increment_array(c[:, j:j + 1], 2)[:, 0]) == inc
        assert O.morton_rank_array([4, 4], np.array(inc)[:, None], 1)[0] == O.morton_rank([4, 4], inc, 1)


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


# ---- f16: exact-rational pin of the restatement's half arithmetic ---------
#
# The reference computes f16 LOD on the GPU only (src/gpu/lod.cu): the window
# mean sums the halfs in a float accumulator, divides in float and converts
# to half (:71-75, :413-528); the append-dim fold adds halfs (:847-860) and
# the emit divides the half sum by the count (:827-845). Each step is an
# IEEE-754 operation rounded to nearest, ties to even. The model below
# evaluates those steps on exact rationals with its own rounding function,
# independent of numpy's casts, and must equal the restatement bit for bit
# on general data (all levels of the pyramid, every fold and emit).

from fractions import Fraction  # noqa: E402

F16 = (11, -14, 15)   # significand bits, emin, emax
F32 = (24, -126, 127)
INF = None


def rne(x, fmt):
    """x rounded to the nearest value of binary format fmt, ties to even
    (subnormals included; None for overflow to infinity)."""
    if x is INF:
        return INF
    p, emin, emax = fmt
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    q = Fraction(2) ** (max(e, emin) - p + 1)  # spacing of representable values near a
    m = a / q
    n = m.numerator // m.denominator
    r = m - n
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
        n += 1
    v = n * q
    if v > (2 - Fraction(2) ** (1 - p)) * Fraction(2) ** emax:
        return INF
    return sign * v


def _frac(h):
    return Fraction(float(h))


def model_window_mean_f16(vals):
    s = Fraction(0)
    for v in vals:
        s = rne(s + _frac(v), F32)
    return rne(rne(s / len(vals), F32), F16)


def _as_half(x):
    return np.float16(float(x))


@pytest.mark.parametrize("shape", [[40, 36], [17, 23], [8, 8, 6]])
def test_f16_pyramid_every_level_exact_rational(shape):
    # a general-data f16 epoch through every level of the restatement's
    # pyramid; each level is the model's window means of the level below
    ch = 4 if len(shape) == 2 else 2
    dims = [O.ODim(4, 4, 1)] + [O.ODim(s, ch, 0, True) for s in shape]
    plan = O.Plan(dims, 2)
    assert plan.nlod >= 3
    lod = O.LodState(plan, np.float16)
    E = plan.levels[0].epoch_elements
    x = (np.random.default_rng(sum(shape)).standard_normal(E) * 100).astype(np.float16)
    x[::97] = np.float16(60000.0)   # large and tiny magnitudes: float sums round
    x[5::89] = np.float16(3e-5)
    vals = lod.reduce_levels(x, O.MEAN)
    for lv in range(1, plan.nlod):
        src = vals[lv - 1].reshape(plan.fixed_count, -1)
        ends = np.asarray(lod.ends[lv - 1])
        got = vals[lv].reshape(plan.fixed_count, -1)
        for b in range(plan.fixed_count):
            s = 0
            for i, e in enumerate(ends[:got.shape[1]]):
                want = model_window_mean_f16(src[b, s:e])
                assert want is not INF
                assert got[b, i] == _as_half(want) and np.signbit(got[b, i]) == (want < 0), (lv, b, i)
                s = e


def test_f16_fold_emit_chain_exact_rational():
    # eight epochs of general-data level values through the restatement's
    # fold and emit (mean, min, max) against the model's chain
    dims = [O.ODim(16, 2, 2, True), O.ODim(20, 4, 0, True), O.ODim(12, 4, 0, True)]
    plan = O.Plan(dims, 2)
    assert plan.nlod >= 3
    rng = np.random.default_rng(17)
    for method in (O.MEAN, O.MIN, O.MAX):
        lod = O.LodState(plan, np.float16)
        acc = {lv: None for lv in range(1, plan.nlod)}
        cnt = {lv: 0 for lv in range(1, plan.nlod)}
        for epoch in range(8):
            vals = [(rng.standard_normal(plan.fixed_count * L.lod_nelem) * 100).astype(np.float16)
                    for L in plan.levels]
            vals[1][::7] = np.float16(30000.0)
            new = [v.copy() for v in vals]
            mask = lod.fold_emit(new, method)
            want_mask = 1
            for lv in range(1, plan.nlod):
                v = [_frac(h) for h in vals[lv]]
                if cnt[lv] == 0:
                    acc[lv] = v
                elif method == O.MEAN:
                    acc[lv] = [rne(a + b, F16) for a, b in zip(acc[lv], v)]
                elif method == O.MIN:
                    acc[lv] = [b if b < a else a for a, b in zip(acc[lv], v)]
                else:
                    acc[lv] = [b if b > a else a for a, b in zip(acc[lv], v)]
                cnt[lv] += 1
                if cnt[lv] >= (1 << lv):
                    out = [rne(a / cnt[lv], F16) if method == O.MEAN else a for a in acc[lv]]
                    got = new[lv]
                    for i, w in enumerate(out):
                        assert w is not INF
                        assert got[i] == _as_half(w), (method, epoch, lv, i)
                    cnt[lv] = 0
                    want_mask |= 1 << lv
            assert mask == want_mask
            for lv in range(1, plan.nlod):
                if acc[lv] is not None and all(a is not INF for a in acc[lv]):
                    np.testing.assert_array_equal(lod.accum[lv], np.array([float(a) for a in acc[lv]],
                                                                          dtype=np.float16))


def test_f16_model_catches_a_half_accumulator():
    # the model distinguishes the reference's float accumulator from a
    # half-precision one on these inputs (a planted mistake on the CPU)
    vals = np.array([2048, 1, 1, 1], dtype=np.float16)
    good = model_window_mean_f16(vals)
    s = Fraction(0)
    for v in vals:
        s = rne(s + _frac(v), F16)
    bad = rne(s / len(vals), F16)
    assert good != bad
    got = O.reduce_windows(vals.reshape(1, -1), np.array([4]), O.MEAN)[0, 0]
    assert got == _as_half(good)


def test_rne_matches_ieee_on_known_values():
    # closed forms of the rounding itself: half max, the smallest subnormal,
    # ties to even at 2049 -> 2048 and 2051 -> 2052, overflow
    assert rne(Fraction(65504), F16) == 65504
    assert rne(Fraction(65520), F16) is INF
    assert rne(Fraction(2049), F16) == 2048
    assert rne(Fraction(2051), F16) == 2052
    assert rne(Fraction(1, 2 ** 24), F16) == Fraction(1, 2 ** 24)
    assert rne(Fraction(1, 2 ** 25), F16) == 0  # tie between 0 and 2^-24: even (0)
    assert rne(Fraction(3, 2 ** 26), F16) == Fraction(1, 2 ** 24)
    assert rne(Fraction(1, 3), F32) == Fraction(11184811, 2 ** 25)
