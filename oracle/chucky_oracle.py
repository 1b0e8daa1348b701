"""CPU restatement of the reference's streaming write path (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker, never by the product.

Restates, function by function, the reference's CPU backend for CODEC_NONE
output (compressed chunks are pinned by decode round trips instead):

  layout      src/stream/config.c:84-138, 232-376; src/util/index.ops.c:97-113
  dims        src/dimension.c:178-196; src/stream/dim_info.c:25-60
  LOD plan    src/lod/lod_plan.c:21-72 (morton_rank), 105-125 (ends),
              213-275 (level shapes / nlod), 328-345 (epoch dims)
  transpose   src/cpu/transpose.cpp:43-105 (closed form of the odometer)
  LOD         src/cpu/lod.cpp:54-127 (reduce windows), 182-268 (gather LUT),
              298-359 (chunk LUT), 363-438 (dim0 fold / emit)
  pipeline    src/cpu/pipeline.c:112-460, src/cpu/stream.c:527-762
  aggregate   src/cpu/aggregate.c:10-24, 129-188; src/stream/types.aggregate.c:28-143
  shards      src/zarr/shard_delivery.c:10-204; src/zarr/crc32c.c:8-28

Pinning: tests/test_oracle_cpu.py checks every part against the reference
build (oracle/_ref) on randomized configurations, against the committed golden
shard files (tests/golden/), and against closed forms: the standard CRC-32C
check value, the 3x5 compacted Morton table of docs/design.md, brute-force
Z-order ranks on tiny boxes, and the level-2 fold chain 11v/32 of
tests/test_lod.c:919-928.

f16 has no CPU path in the reference; the restatement follows its GPU path
(float accumulator for the mean, half-precision fold and emit;
src/gpu/lod.cu:71-75, 413-528, 827-941). Pinned: every level of the
order-only reductions (min, max, median, 2nd extremes) and levels 0-1 of the
mean equal float16 of the f32 reference; on general data every pyramid level
and every fold / emit equal an exact-rational model of that arithmetic (sums,
divides and half adds on fractions with an independent round-to-nearest-even)
bit for bit.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

MEAN, MIN, MAX, MEDIAN, MAX_SUP, MIN_SUP = range(6)


def ceildiv(a, b):
    return (a + b - 1) // b


def ceil_log2(v):
    p = 0
    while (1 << p) < v:
        p += 1
    return p


# ---- CRC-32C (src/zarr/crc32c.c:8-28) --------------------------------------

def _crc_table():
    t = []
    for i in range(256):
        c = i
        for _ in range(8):
            c = (c >> 1) ^ (0x82F63B78 & (-(c & 1) & 0xFFFFFFFF))
        t.append(c)
    return t


_CRC = _crc_table()


def crc32c(data: bytes) -> int:
    c = 0xFFFFFFFF
    for b in data:
        c = _CRC[(c ^ b) & 0xFF] ^ (c >> 8)
    return c ^ 0xFFFFFFFF


# ---- Morton rank and window ends (src/lod/lod_plan.c:21-125) ---------------

def morton_rank(shape, coords, depth=0):
    nd = len(shape)
    p = ceil_log2(max(shape)) if shape else 0
    for c in coords:
        if c > 0:
            p = max(p, ceil_log2(c + 1))
    levels = p + depth
    rank = 0
    prefix = [0] * nd
    for lv in range(levels):
        scale = 1 << (levels - 1 - lv)
        digit = 0
        if lv < p:
            for d in range(nd):
                digit |= ((coords[d] >> (p - 1 - lv)) & 1) << d

        def inside(ext, lo):
            return 0 if lo >= ext else min(ext - lo, scale)
        lo_h = [inside(shape[d], prefix[d] * 2 * scale) for d in range(nd)]
        hi_h = [inside(shape[d], (prefix[d] * 2 + 1) * scale) for d in range(nd)]
        below = [1] * (nd + 1)
        for d in range(nd):
            below[d + 1] = below[d] * (lo_h[d] + hi_h[d])
        fixed = 1
        for d in range(nd - 1, -1, -1):
            hi = (digit >> d) & 1
            if hi:
                rank += fixed * lo_h[d] * below[d]
            fixed *= hi_h[d] if hi else lo_h[d]
        for d in range(nd):
            prefix[d] = prefix[d] * 2 + ((digit >> d) & 1)
    return rank


def z_increment(coords, p):
    c = list(coords)
    for bit in range(p):
        for d in range(len(c)):
            m = 1 << bit
            c[d] ^= m
            if c[d] & m:
                return c
    c = [0] * len(c)
    c[0] = 1 << p
    return c


def _ceil_log2_array(v):
    """ceil_log2 elementwise: the least p with 2^p >= v (integer steps only)."""
    v = np.asarray(v, dtype=np.int64)
    out = np.zeros(v.shape, dtype=np.int64)
    while True:
        short = (np.int64(1) << out) < v
        if not short.any():
            return out
        out += short


def _morton_rank_fixed_p(shape, c, p, depth):
    nd = len(shape)
    levels = p + depth
    n = c.shape[1]
    rank = np.zeros(n, dtype=np.int64)
    prefix = np.zeros((nd, n), dtype=np.int64)
    for lv in range(levels):
        scale = 1 << (levels - 1 - lv)
        bits = [((c[d] >> (p - 1 - lv)) & 1) if lv < p else np.zeros(n, dtype=np.int64)
                for d in range(nd)]

        def inside(ext, lo):
            return np.where(lo >= ext, 0, np.minimum(ext - lo, scale))
        lo_h = [inside(shape[d], prefix[d] * 2 * scale) for d in range(nd)]
        hi_h = [inside(shape[d], (prefix[d] * 2 + 1) * scale) for d in range(nd)]
        below = [np.ones(n, dtype=np.int64)]
        for d in range(nd):
            below.append(below[d] * (lo_h[d] + hi_h[d]))
        fixed = np.ones(n, dtype=np.int64)
        for d in range(nd - 1, -1, -1):
            hi = bits[d] == 1
            rank += np.where(hi, fixed * lo_h[d] * below[d], 0)
            fixed = fixed * np.where(hi, hi_h[d], lo_h[d])
        for d in range(nd):
            prefix[d] = prefix[d] * 2 + bits[d]
    return rank


def morton_rank_array(shape, coords, depth=0):
    """morton_rank over arrays: coords is (len(shape), N); the same per-element
    arithmetic, elements grouped by their bit depth p."""
    c = np.asarray(coords, dtype=np.int64).reshape(len(shape), -1)
    base = ceil_log2(max(shape)) if len(shape) else 0
    p_el = np.maximum(base, _ceil_log2_array(c.max(axis=0) + 1)) if c.size else np.zeros(0, np.int64)
    out = np.zeros(c.shape[1], dtype=np.int64)
    for p in np.unique(p_el):
        sel = p_el == p
        out[sel] = _morton_rank_fixed_p(shape, c[:, sel], int(p), depth)
    return out


def z_increment_array(coords, p):
    """z_increment over arrays (coords: (nd, N))."""
    c = np.array(coords, dtype=np.int64).reshape(len(coords), -1)
    done = np.zeros(c.shape[1], dtype=bool)
    for bit in range(p):
        m = 1 << bit
        for d in range(c.shape[0]):
            act = ~done
            c[d, act] ^= m
            done |= act & ((c[d] & m) != 0)
    rest = ~done
    c[:, rest] = 0
    c[0, rest] = 1 << p
    return c


def window_ends_array(child_shape, parent_shape):
    """window_ends computed over all parents at once."""
    n_par = int(np.prod(parent_shape))
    j = np.arange(n_par, dtype=np.int64)
    c, r = [], j
    for s in parent_shape:
        c.append(r % s)
        r = r // s
    c = np.stack(c)
    zp = ceil_log2(max(parent_shape))
    ends = np.zeros(n_par, dtype=np.int64)
    ends[morton_rank_array(parent_shape, c, 0)] = morton_rank_array(child_shape, z_increment_array(c, zp), 1)
    return ends


def window_ends(child_shape, parent_shape):
    """ends[morton_rank(parent)] = child rank of the next parent's first child."""
    n_par = int(np.prod(parent_shape))
    ends = [0] * n_par
    zp = ceil_log2(max(parent_shape))
    for j in range(n_par):
        c, r = [], j
        for s in parent_shape:
            c.append(r % s)
            r //= s
        pos = morton_rank(parent_shape, c, 0)
        ends[pos] = morton_rank(child_shape, z_increment(c, zp), 1)
    return ends


# ---- stream plan ------------------------------------------------------------

@dataclass
class ODim:
    size: int
    chunk: int
    cps: int = 0
    downsample: bool = False
    storage_position: int = -1


@dataclass
class Level:
    shape: List[int]
    lod_nelem: int = 1
    lifted_shape: List[int] = field(default_factory=list)
    lifted_strides: List[int] = field(default_factory=list)
    chunk_elements: int = 1
    chunks_per_epoch: int = 1
    epoch_elements: int = 1
    chunk_offset: int = 0
    cps_append: int = 1
    cps_inner: int = 1
    shard_inner_count: int = 1
    agg_shape: List[int] = field(default_factory=list)
    agg_strides: List[int] = field(default_factory=list)
    covering: int = 1


class Plan:
    def __init__(self, dims: List[ODim], bpe: int, K: int = 0, target_batch_chunks: int = 0,
                 max_nlod: int = 0, append_downsample_method: int = MEAN):
        self.dims = dims
        self.rank = r = len(dims)
        self.bpe = bpe
        for i, d in enumerate(dims):
            if d.storage_position < 0:
                d.storage_position = i
        # n_append (src/dimension.c:178-196)
        max_n = 0
        for d in range(r):
            if dims[d].chunk != 1:
                break
            max_n = d + 1
        na = 1
        if max_n > 1:
            na = max_n
            for d in range(max_n):
                if dims[d].downsample:
                    na = d + 1
                    break
        self.na = na
        self.lod_mask = [d for d in range(na, r) if dims[d].downsample]
        self.append_downsample = bool(dims[na - 1].downsample)
        self.bounded_append_chunks = 1
        for d in range(1, na):
            self.bounded_append_chunks *= ceildiv(dims[d].size, dims[d].chunk)
        self.order = [0] * r
        for d in range(r):
            self.order[dims[d].storage_position] = d
        # LOD plan over epoch dims
        shape0 = [(d.size if d.size else d.chunk) for d in dims]
        for d in range(na):
            shape0[d] = dims[d].chunk
        self.lod_dims = list(self.lod_mask)
        self.fixed_dims = [d for d in range(r) if d not in self.lod_mask]
        self.fixed_shape = [shape0[d] for d in self.fixed_dims]
        self.fixed_count = int(np.prod(self.fixed_shape)) if self.fixed_shape else 1
        max_levels = 32 if max_nlod == 0 else max_nlod
        shapes = [shape0]
        while len(shapes) < max_levels:
            cur = shapes[-1]
            if all(ceildiv(cur[d], dims[d].chunk) <= 1 for d in self.lod_dims):
                break
            nxt = list(cur)
            for d in self.lod_dims:
                nxt[d] = (cur[d] + 1) // 2
            shapes.append(nxt)
        self.multiscale = len(self.lod_dims) > 0
        self.nlod = len(shapes) if self.multiscale else 1
        self.levels = [Level(shape=list(shapes[i])) for i in range(self.nlod)]
        for L in self.levels:
            L.lod_nelem = int(np.prod([L.shape[d] for d in self.lod_dims])) if self.lod_dims else 1
        # chunk layouts (src/stream/config.c:84-138)
        total = 0
        for L in self.levels:
            cc = [1 if L.shape[d] == 0 else ceildiv(L.shape[d], dims[d].chunk) for d in range(r)]
            L.lifted_shape = [v for d in range(r) for v in (cc[d], dims[d].chunk)]
            L.chunk_elements = int(np.prod([d.chunk for d in dims]))
            st = [0] * (2 * r)
            es, cs = 1, L.chunk_elements
            for j in range(r - 1, -1, -1):
                d = self.order[j]
                st[2 * d + 1] = es
                es *= dims[d].chunk
                st[2 * d] = cs
                cs *= cc[d]
            for d in range(na):
                st[2 * d] = 0
            L.lifted_strides = st
            L.chunks_per_epoch = int(np.prod(cc[na:])) if na < r else 1
            L.epoch_elements = L.chunks_per_epoch * L.chunk_elements
            L.chunk_offset = total
            total += L.chunks_per_epoch
        self.total_chunks = total
        if K == 0:
            t = target_batch_chunks or 1024
            K = 1
            while K < ceildiv(t, total):
                K <<= 1
        self.K = min(K, 128)
        # shard geometry and aggregate layout (config.c:298-372, types.aggregate.c:65-143)
        for lv, L in enumerate(self.levels):
            ext = []
            for d in range(r):
                s = dims[d].size if lv == 0 else L.shape[d]
                cc = 1 if s == 0 else ceildiv(s, dims[d].chunk)
                cps = dims[d].cps or cc
                if s > 0 and cc < cps:
                    cps = cc
                ext.append((cc, cps, ceildiv(cc, cps)))
            L.shard_inner_count = int(np.prod([ext[d][2] for d in range(na, r)])) if na < r else 1
            L.cps_append = int(np.prod([ext[d][1] for d in range(na)]))
            if self.append_downsample and lv > 0:
                L.cps_append = L.cps_append // (1 << lv) if L.cps_append > (1 << lv) else 1
            L.cps_inner = int(np.prod([ext[d][1] for d in range(na, r)])) if na < r else 1
            so = [ext[self.order[j]] for j in range(r)]
            L.agg_shape, sc = [], []
            for d in range(na, r):
                L.agg_shape += [so[d][2], so[d][1]]
                sc.append(so[d][2])
            L.covering = int(np.prod(L.agg_shape)) if L.agg_shape else 1
            strides = [0] * len(L.agg_shape)
            shard_acc, within_acc = 1, 1
            for d in range(r - 1, na - 1, -1):
                k = d - na
                strides[2 * k] = shard_acc * L.cps_inner
                shard_acc *= so[d][2]
                strides[2 * k + 1] = within_acc
                within_acc *= so[d][1]
            L.agg_strides = strides
        self.max_cursor = 0
        if dims[0].size > 0:
            self.max_cursor = self.levels[0].epoch_elements
            for d in range(na):
                self.max_cursor *= ceildiv(dims[d].size, dims[d].chunk)

    @property
    def cps_total(self):
        return [L.cps_append * L.cps_inner for L in self.levels]

    def perm(self, lv, j):
        L = self.levels[lv]
        o, rest = 0, j
        for d in range(len(L.agg_shape) - 1, -1, -1):
            o += (rest % L.agg_shape[d]) * L.agg_strides[d]
            rest //= L.agg_shape[d]
        return o


def ravel(shape, strides, idx):
    """src/util/index.ops.c:42-53, vectorised over idx."""
    idx = np.asarray(idx, dtype=np.uint64)
    o = np.zeros_like(idx)
    rest = idx.copy()
    for d in range(len(shape) - 1, -1, -1):
        o += (rest % np.uint64(shape[d])) * np.uint64(strides[d])
        rest //= np.uint64(shape[d])
    return o


# ---- LOD -------------------------------------------------------------------

# Widened accumulators of src/cpu/lod.cpp:54-127. f16 has no CPU path; it
# follows the reference's GPU path (src/gpu/lod.cu:71-75: float accumulator
# for the mean; fold a+b and emit acc/count in half precision,
# src/gpu/lod.cu:827-845), which numpy's float16 reproduces to within 1 ULP
# (its half arithmetic rounds a float32 result).
_ACC = {np.dtype(np.uint8): np.uint32, np.dtype(np.uint16): np.uint32,
        np.dtype(np.uint32): np.uint64, np.dtype(np.int8): np.int32,
        np.dtype(np.int16): np.int32, np.dtype(np.int32): np.int64,
        np.dtype(np.float16): np.float32}


def reduce_window(vals: np.ndarray, method: int):
    """src/cpu/lod.cpp:54-127 for one window (numpy scalar semantics)."""
    dt = vals.dtype
    n = len(vals)
    if method == MEAN:
        acc_t = _ACC.get(dt, dt.type)
        s = acc_t(0)
        with np.errstate(over="ignore"):  # 64-bit sums wrap as in C
            for v in vals:
                s = acc_t(s + acc_t(v))
        if dt.kind in "iu":
            q = int(s) // n if int(s) >= 0 else -((-int(s)) // n)  # C truncation
            return dt.type(np.array(q).astype(dt))
        return dt.type(s / acc_t(n))
    if method == MIN:
        b = vals[0]
        for v in vals[1:]:
            if v < b:
                b = v
        return b
    if method == MAX:
        b = vals[0]
        for v in vals[1:]:
            if v > b:
                b = v
        return b
    if method == MEDIAN:
        buf = sorted(list(vals[:16]))
        return dt.type(buf[min(n, 16) // 2])
    if method == MAX_SUP:
        t1 = t2 = vals[0]
        for v in vals[1:]:
            if v >= t1:
                t2, t1 = t1, v
            elif v > t2:
                t2 = v
        return t2
    b1 = b2 = vals[0]
    for v in vals[1:]:
        if v <= b1:
            b2, b1 = b1, v
        elif v < b2:
            b2 = v
    return b2


def reduce_windows(src: np.ndarray, ends: np.ndarray, method: int) -> np.ndarray:
    """reduce_window over every window [ends[i-1], ends[i]) of each row of
    src (batches x n), vectorized across windows of equal length; each
    window is still reduced element by element in its own order, with the
    same operations as reduce_window (src/cpu/lod.cpp:54-127)."""
    dt = src.dtype
    ends = np.asarray(ends, dtype=np.int64)
    starts = np.concatenate([[0], ends[:-1]]).astype(np.int64)
    lens = ends - starts
    out = np.empty((src.shape[0], len(ends)), dtype=dt)
    for n in np.unique(lens):
        n = int(n)
        idx = np.flatnonzero(lens == n)
        w = src[:, starts[idx][:, None] + np.arange(n)[None, :]]  # (B, k, n)
        if method == MEAN:
            acc_t = _ACC.get(dt, dt.type)
            s = np.zeros(w.shape[:2], dtype=acc_t)
            with np.errstate(over="ignore"):
                for j in range(n):
                    s = (s + w[..., j].astype(acc_t)).astype(acc_t)
            if dt.kind in "iu":
                si = s.astype(object) if dt.itemsize == 8 else s.astype(np.int64)
                q = np.where(si >= 0, si // n, -((-si) // n))  # C truncation
                r = np.array(q, dtype=object).astype(dt) if dt.itemsize == 8 else q.astype(dt)
            else:
                r = (s / acc_t(n)).astype(dt)
        elif method in (MIN, MAX):
            r = w[..., 0].copy()
            for j in range(1, n):
                x = w[..., j]
                r = np.where(x < r if method == MIN else x > r, x, r)
        elif method == MEDIAN:
            m = min(n, 16)
            r = np.sort(w[..., :m], axis=-1, kind="stable")[..., m // 2]
        else:
            t1 = w[..., 0].copy()
            t2 = w[..., 0].copy()
            for j in range(1, n):
                x = w[..., j]
                if method == MAX_SUP:
                    first, second = x >= t1, x > t2
                else:
                    first, second = x <= t1, x < t2
                t2 = np.where(first, t1, np.where(second, x, t2))
                t1 = np.where(first, x, t1)
            r = t2
        out[:, idx] = r
    return out


def add_shift(a, b, s, dt):
    """overflow_safe_add_shift of src/cpu/lod.cpp:363-374 (C promotion rules)."""
    if dt.kind == "f" or dt.itemsize >= 8:
        with np.errstate(over="ignore"):
            return (a + b).astype(dt)
    ai, bi = a.astype(np.int64), b.astype(np.int64)
    m = (1 << s) - 1
    if dt == np.uint32:  # unsigned int arithmetic wraps at 2^32
        r = ((ai >> s) + (bi >> s) + (((ai & m) + (bi & m)) >> s)) & 0xFFFFFFFF
    else:
        r = (ai >> s) + (bi >> s) + (((ai & m) + (bi & m)) >> s)
    return r.astype(dt)


def emit_accum(acc, count, append_method):
    """Accumulator emit (src/cpu/lod.cpp:545-602; f16: src/gpu/lod.cu:827-845):
    a float mean divides the running sum by the epoch count, everything else
    is already final."""
    dt = acc.dtype
    if dt.kind == "f" and append_method == MEAN:
        return (acc / dt.type(count)).astype(dt)
    return acc.copy()


class LodState:
    def __init__(self, plan: Plan, dtype):
        self.p = plan
        self.dt = np.dtype(dtype)
        p = plan
        full = p.levels[0].shape
        fstride = [1] * p.rank
        for d in range(p.rank - 2, -1, -1):
            fstride[d] = fstride[d + 1] * full[d + 1]
        self.fstride = fstride
        # the LUT loops of src/cpu/lod.cpp:182-268 (gather) and 298-359
        # (chunk scatter), evaluated over every cell at once: row-major
        # coordinates (last dim fastest), Morton rank -> offset
        lshape = [p.levels[0].shape[d] for d in p.lod_dims]
        n0 = p.levels[0].lod_nelem
        c = np.unravel_index(np.arange(n0, dtype=np.int64), lshape)
        off = sum(c[k].astype(np.int64) * fstride[p.lod_dims[k]] for k in range(len(lshape)))
        self.gather_lut = np.zeros(n0, dtype=np.int64)
        self.gather_lut[morton_rank_array(lshape, np.stack(c))] = off
        fc = np.unravel_index(np.arange(p.fixed_count, dtype=np.int64), p.fixed_shape) if p.fixed_shape else ()
        self.gather_fixed = np.zeros(p.fixed_count, dtype=np.int64)
        for k in range(len(p.fixed_shape)):
            self.gather_fixed += fc[k].astype(np.int64) * fstride[p.fixed_dims[k]]
        self.ends = []
        for l in range(p.nlod - 1):
            ch = [p.levels[l].shape[d] for d in p.lod_dims]
            pa = [p.levels[l + 1].shape[d] for d in p.lod_dims]
            self.ends.append(window_ends_array(ch, pa))
        self.chunk_lut, self.chunk_fixed = [], []
        for lv, L in enumerate(p.levels):
            ls = [L.shape[d] for d in p.lod_dims]
            c = np.unravel_index(np.arange(L.lod_nelem, dtype=np.int64), ls)
            off = np.zeros(L.lod_nelem, dtype=np.int64)
            for k in range(len(ls)):
                d = p.lod_dims[k]
                cs = L.lifted_shape[2 * d + 1]
                ck = c[k].astype(np.int64)
                off += (ck // cs) * L.lifted_strides[2 * d] + (ck % cs) * L.lifted_strides[2 * d + 1]
            lut = np.zeros(L.lod_nelem, dtype=np.int64)
            lut[morton_rank_array(ls, np.stack(c))] = off
            fx = np.full(p.fixed_count, L.chunk_offset * L.chunk_elements, dtype=np.int64)
            for k in range(len(p.fixed_shape)):
                d = p.fixed_dims[k]
                cs = L.lifted_shape[2 * d + 1]
                ck = fc[k].astype(np.int64)
                fx += (ck // cs) * L.lifted_strides[2 * d] + (ck % cs) * L.lifted_strides[2 * d + 1]
            self.chunk_lut.append(lut)
            self.chunk_fixed.append(fx)
        self.accum = [None] * p.nlod
        self.counts = [0] * p.nlod

    def reduce_levels(self, linear: np.ndarray, method: int):
        p = self.p
        n0 = p.levels[0].lod_nelem
        vals = [linear[(self.gather_fixed[:, None] + self.gather_lut[None, :]).reshape(-1)]]
        for l in range(p.nlod - 1):
            sn, dn = p.levels[l].lod_nelem, p.levels[l + 1].lod_nelem
            src = vals[l].reshape(p.fixed_count, sn)
            out = reduce_windows(src, np.asarray(self.ends[l][:dn]), method)
            vals.append(out.reshape(-1))
        del n0
        return vals

    def fold_emit(self, vals, append_method: int) -> int:
        """Append-dim fold of one epoch's levels >= 1 into the accumulators
        (src/cpu/lod.cpp:363-438) and the emit of every level whose window of
        2^level epochs completes (src/cpu/pipeline.c:258-288); emitted values
        replace vals[level]. Returns the active-level mask."""
        p, dt = self.p, self.dt
        for lv in range(1, p.nlod):
            v = vals[lv]
            if self.counts[lv] == 0:
                self.accum[lv] = v.copy()
            elif append_method == MEAN:
                self.accum[lv] = add_shift(self.accum[lv], v, lv, dt)
            elif append_method == MIN:
                self.accum[lv] = np.where(v < self.accum[lv], v, self.accum[lv])
            else:
                self.accum[lv] = np.where(v > self.accum[lv], v, self.accum[lv])
        mask = 1
        for lv in range(1, p.nlod):
            self.counts[lv] += 1
            if self.counts[lv] >= (1 << lv):
                vals[lv] = emit_accum(self.accum[lv], self.counts[lv], append_method)
                self.counts[lv] = 0
                mask |= 1 << lv
        return mask

    def scatter(self, pool_epoch: np.ndarray, lv: int, vals: np.ndarray):
        idx = (self.chunk_fixed[lv][:, None] + self.chunk_lut[lv][None, :]).reshape(-1)
        pool_epoch[idx] = vals


# ---- aggregate + shards -----------------------------------------------------

class ShardState:
    def __init__(self, L: Level):
        self.L = L
        self.eis = 0
        self.shard_epoch = 0
        self.cursor = [0] * L.shard_inner_count
        self.index = [[(2**64 - 1, 2**64 - 1)] * (L.cps_append * L.cps_inner)
                      for _ in range(L.shard_inner_count)]
        self.open = [False] * L.shard_inner_count


def finalize(ss: ShardState, out: Dict, level: int):
    """src/zarr/shard_delivery.c:35-102 (no alignment)."""
    for si in range(ss.L.shard_inner_count):
        if not ss.open[si]:
            continue
        idx = b"".join(struct.pack("<QQ", o, n) for o, n in ss.index[si])
        key = (level, ss.shard_epoch * ss.L.shard_inner_count + si)
        out[key] = out.get(key, b"") + idx + struct.pack("<I", crc32c(idx))
        ss.open[si] = False
        ss.cursor[si] = 0
        ss.index[si] = [(2**64 - 1, 2**64 - 1)] * len(ss.index[si])
    ss.eis = 0
    ss.shard_epoch += 1


def deliver(plan: Plan, lv: int, ss: ShardState, chunks: List[bytes], n_active: int, out: Dict):
    """Aggregate (src/cpu/aggregate.c:129-188, no padding) + deliver
    (src/zarr/shard_delivery.c:104-204). chunks[a*M + j] = bytes of epoch a, chunk j."""
    L = plan.levels[lv]
    M, C, cpsi = L.chunks_per_epoch, L.covering, L.cps_inner
    sizes = [0] * (n_active * C)
    data = [b""] * (n_active * C)
    for a in range(n_active):
        for j in range(M):
            q = plan.perm(lv, j)
            t = (q // cpsi) * n_active * cpsi + (q % cpsi) + a * cpsi
            sizes[t] = len(chunks[a * M + j])
            data[t] = chunks[a * M + j]
    offs = [0]
    for s in sizes:
        offs.append(offs[-1] + s)
    blob = b"".join(data)
    a = 0
    while a < n_active:
        run = min(L.cps_append - ss.eis, n_active - a)
        for si in range(L.shard_inner_count):
            key = (lv, ss.shard_epoch * L.shard_inner_count + si)
            ss.open[si] = True
            j0 = si * n_active * cpsi + a * cpsi
            j1 = j0 + run * cpsi
            rb = offs[j1] - offs[j0]
            cur = out.get(key, b"")
            if len(cur) < ss.cursor[si]:
                cur += b"\0" * (ss.cursor[si] - len(cur))
            out[key] = cur[:ss.cursor[si]] + blob[offs[j0]:offs[j1]]
            for r in range(run):
                eis = ss.eis + r
                js = j0 + r * cpsi
                for j in range(js, js + cpsi):
                    if sizes[j] > 0:
                        slot = eis * cpsi + (j - js)
                        ss.index[si][slot] = (ss.cursor[si] + offs[j] - offs[j0], sizes[j])
            ss.cursor[si] += rb
        ss.eis += run
        a += run
        if ss.eis >= L.cps_append:
            finalize(ss, out, lv)


# ---- the stream ---------------------------------------------------------------

def write_stream(dims: List[ODim], dtype, data: np.ndarray, reduce_method=MEAN,
                 append_method=MEAN, K=0, max_nlod=0) -> Tuple[Dict, int]:
    """CODEC_NONE shards of the whole stream: {(level, shard_index): bytes}."""
    dt = np.dtype(dtype)
    plan = Plan(dims, dt.itemsize, K=K, max_nlod=max_nlod)
    flat = np.ascontiguousarray(data).view(dt).reshape(-1)
    if plan.max_cursor:
        flat = flat[:plan.max_cursor]
    E = plan.levels[0].epoch_elements
    CE = plan.levels[0].chunk_elements
    pool = np.zeros((plan.K, plan.total_chunks * CE), dtype=dt)
    linear = np.zeros(E, dtype=dt)
    lod = LodState(plan, dt) if plan.multiscale else None
    shards = [ShardState(L) for L in plan.levels]
    out: Dict = {}
    masks: List[int] = []

    def scatter_epoch(e):
        if not plan.multiscale:
            return 1
        vals = lod.reduce_levels(linear, reduce_method)
        mask = (1 << plan.nlod) - 1
        if plan.append_downsample:
            mask = lod.fold_emit(vals, append_method)
        for lv in range(plan.nlod):
            if mask & (1 << lv):
                lod.scatter(pool[e], lv, vals[lv])
        return mask

    def emit(acc, count):
        return emit_accum(acc, count, append_method)

    flushed = [False] * plan.K  # pool slot compressed by some flush so far

    def flush_batch(n_epochs, mks, drain=False):
        for e in range(n_epochs):
            flushed[e] = True
        for lv in range(plan.nlod):
            L = plan.levels[lv]
            act = [e for e in range(n_epochs) if mks[e] & (1 << lv)]
            if not act:
                continue
            chunks = []
            if drain and (1 << lv) == plan.K:
                # src/cpu/pipeline.c:159-180: the per-level batch LUT built at
                # init (pool epoch (a+1)*period-1, src/cpu/pipeline.c:14-31)
                # is reused whenever the active count equals the level's
                # batch_active_count (K / period). For the one-epoch drain
                # batch that holds exactly when period == K, so the drain
                # delivers pool slot K-1 as last compressed -- the level's
                # previous emission -- or zero-size (empty) chunks if no
                # batch ever compressed that slot (comp_sizes is calloc'd,
                # src/cpu/stream.c:70-71).
                src = plan.K - 1
                base = L.chunk_offset * CE
                for j in range(L.chunks_per_epoch):
                    chunks.append(pool[src, base + j * CE: base + (j + 1) * CE].tobytes()
                                  if flushed[src] else b"")
                deliver(plan, lv, shards[lv], chunks, 1, out)
                continue
            for e in act:
                base = L.chunk_offset * CE
                for j in range(L.chunks_per_epoch):
                    chunks.append(pool[e, base + j * CE: base + (j + 1) * CE].tobytes())
            deliver(plan, lv, shards[lv], chunks, len(act), out)

    cursor = 0
    while cursor < flat.size:
        n = min(E - cursor % E, flat.size - cursor)
        seg = flat[cursor:cursor + n]
        off = cursor % E
        if plan.multiscale:
            linear[off:off + n] = seg
        else:
            L0 = plan.levels[0]
            pos = ravel(L0.lifted_shape, L0.lifted_strides, np.arange(off, off + n, dtype=np.uint64))
            pool[len(masks)][pos.astype(np.int64)] = seg
        cursor += n
        if cursor % E == 0:
            masks.append(scatter_epoch(len(masks)))
            if len(masks) == plan.K:
                flush_batch(len(masks), masks)
                masks = []
    if cursor % E != 0:
        masks.append(scatter_epoch(len(masks)))
    if masks:
        flush_batch(len(masks), masks)
    if plan.multiscale and plan.append_downsample:
        drain = 0
        for lv in range(1, plan.nlod):
            if lod.counts[lv] > 0:
                v = emit(lod.accum[lv], lod.counts[lv])
                lod.counts[lv] = 0
                lod.scatter(pool[0], lv, v)
                drain |= 1 << lv
        if drain:
            flush_batch(1, [drain], drain=True)
    for lv in range(plan.nlod):
        if shards[lv].eis > 0:
            finalize(shards[lv], out, lv)
    return out, cursor
