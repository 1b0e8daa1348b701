"""Shared test helpers: synthetic data, codec decoders, shard parsing and the
runners that push the same input through the GPU stream and the reference's
CPU stream (oracle/_ref)."""
from __future__ import annotations

import ctypes as C
import os
import struct

import numpy as np

import chucky_amd as ca
from chucky_amd.api import Dim


# ---- synthetic inputs ------------------------------------------------------

def splitmix64(state: int, n: int) -> np.ndarray:
    """n outputs of splitmix64 (the PRNG family of the reference's
    tests/test_data.c:11-18), vectorised."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(state) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def xor_volume(shape, seed=1234) -> np.ndarray:
    """((x ^ y ^ z) * 257 + n) mod 65536, n uniform in [0, 16) (u16)."""
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.uint32) for s in shape], indexing="ij")
    n = (splitmix64(seed, int(np.prod(shape))) % np.uint64(16)).astype(np.uint32).reshape(shape)
    return (((x ^ y ^ z) * 257 + n) % 65536).astype(np.uint16)


def thirds_volume(shape, seed=99) -> np.ndarray:
    """x thirds: zeros, constant 4095, uniform u16 noise."""
    v = np.zeros(shape, dtype=np.uint16)
    X = shape[-1]
    a, b = X // 3, 2 * X // 3
    v[..., a:b] = 4095
    rng = np.random.default_rng(seed)
    v[..., b:] = rng.integers(0, 65536, size=v[..., b:].shape, dtype=np.uint16)
    return v


def random_typed(n: int, dtype: str, seed=7) -> np.ndarray:
    rng = np.random.default_rng(seed)
    dt = np.dtype(ca.NP_DTYPE[dtype])
    if dt.kind == "f":
        return (rng.standard_normal(n) * 100).astype(dt)
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)


# ---- codecs (stock decoders) -----------------------------------------------

def _load(names):
    for n in names:
        try:
            return C.CDLL(n)
        except OSError:
            continue
    raise OSError(f"none of {names} loadable")


_zstd = None
_lz4 = None


def zstd():
    global _zstd
    if _zstd is None:
        _zstd = _load(["/opt/conda/lib/libzstd.so.1", "libzstd.so.1"])
        _zstd.ZSTD_decompress.restype = C.c_size_t
        _zstd.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        _zstd.ZSTD_isError.restype = C.c_uint
        _zstd.ZSTD_isError.argtypes = [C.c_size_t]
    return _zstd


def lz4():
    global _lz4
    if _lz4 is None:
        _lz4 = _load(["/opt/conda/lib/liblz4.so.1", "liblz4.so.1"])
        _lz4.LZ4_decompress_safe.restype = C.c_int
        _lz4.LZ4_decompress_safe.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    return _lz4


def decode(codec: str, blob: bytes, nbytes: int) -> bytes:
    if codec == "none":
        return blob
    out = C.create_string_buffer(nbytes + 64)
    src = C.create_string_buffer(bytes(blob), len(blob))
    if codec == "zstd":
        r = zstd().ZSTD_decompress(out, nbytes + 64, src, len(blob))
        if zstd().ZSTD_isError(r):
            raise ValueError("zstd decode error")
        return out.raw[:r]
    r = lz4().LZ4_decompress_safe(src, out, len(blob), nbytes)
    if r < 0:
        raise ValueError(f"lz4 decode error {r}")
    return out.raw[:r]


# ---- shards -----------------------------------------------------------------

def parse_shard(blob: bytes, cps_total: int):
    """-> list of (offset, nbytes) per slot (None = empty); checks the CRC."""
    idx_bytes = 16 * cps_total
    assert len(blob) >= idx_bytes + 4, "shard shorter than its index"
    idx = blob[-(idx_bytes + 4):-4]
    crc = struct.unpack("<I", blob[-4:])[0]
    assert ca.crc32c(idx) == crc, "index CRC mismatch"
    out = []
    for i in range(cps_total):
        off, n = struct.unpack_from("<QQ", idx, 16 * i)
        out.append(None if (off == 2**64 - 1 and n == 2**64 - 1) else (off, n))
    return out


def decoded_chunks(shards, cps_total_of_level, codec, chunk_bytes):
    """{(level, shard, slot): decoded bytes} for every non-empty slot."""
    out = {}
    for (lv, si), blob in shards.items():
        for k, ent in enumerate(parse_shard(blob, cps_total_of_level(lv))):
            if ent is None:
                continue
            off, n = ent
            out[(lv, si, k)] = decode(codec, blob[off:off + n], chunk_bytes)
    return out


def geometry(cfg):
    """({level: (cps_total, shard_inner_count, chunks_per_epoch)}, epoch
    elements, chunk bytes) of a configuration, from the product's plan."""
    s = ca.GpuStream(cfg, ca.MemSink())
    n = s.status().nlod
    g = {lv: s.level_geometry(lv) for lv in range(n)}
    E = s.epoch_elements
    chunk_bytes = s.layout.chunk_elements * cfg.np_dtype.itemsize
    s.close()
    return g, E, chunk_bytes


def assert_same_shards(got, want, cps_of_level, codec, chunk_bytes, what=""):
    """The reference's definition of equal output: the same shard keys; with
    no codec every shard byte-identical; with lz4 / zstd every shard's index
    fills the same slots, its CRC holds and every chunk decodes (stock
    liblz4 / libzstd) to the same bytes. Chunks are decoded one at a time,
    so large outputs are compared without holding decoded copies."""
    assert set(got) == set(want), f"{what}: shard keys differ: {sorted(set(got) ^ set(want))[:8]}"
    for k in sorted(want):
        if codec == "none":
            same = got[k] == want[k]  # a bool, so a failure does not diff whole shards
            assert same, f"{what}: shard {k} differs"
            continue
        cps = cps_of_level(k[0])
        gi, ri = parse_shard(got[k], cps), parse_shard(want[k], cps)
        assert [e is None for e in gi] == [e is None for e in ri], f"{what}: index layout of shard {k} differs"
        for slot, (ge, re_) in enumerate(zip(gi, ri)):
            if re_ is None:
                continue
            a = decode(codec, got[k][ge[0]:ge[0] + ge[1]], chunk_bytes)
            b = decode(codec, want[k][re_[0]:re_[0] + re_[1]], chunk_bytes)
            same = a == b
            assert same, f"{what}: decoded chunk {slot} of shard {k} differs"


# ---- runners ------------------------------------------------------------------

def run_gpu(cfg, data, pieces=1, pipeline=0, device_input=False, **options):
    sink = ca.MemSink()
    s = ca.GpuStream(cfg, sink, pipeline=pipeline, **options)
    flat = np.ascontiguousarray(data).reshape(-1)
    if device_input:
        dev = ca.DeviceBuffer.from_host(flat)
        for part in np.array_split(np.arange(flat.size), pieces):
            if part.size == 0:
                continue
            b = int(part[0]) * flat.itemsize
            assert s.append_device(dev.ptr + b, part.size * flat.itemsize) == 0, ca.last_error()
    else:
        for part in np.array_split(flat, pieces):
            assert s.append(part) == 0, ca.last_error()
    assert s.flush() == 0, ca.last_error()
    shards = sink.shards()
    meta = {lv: sink.append_sizes(lv) for lv in range(s.status().nlod)}
    cursor = s.cursor
    s.close()
    return shards, meta, cursor


def run_ref(cfg, data, pieces=1):
    from oracle.ref import RefCpuStream
    sink = ca.MemSink()
    r = RefCpuStream(cfg, sink)
    flat = np.ascontiguousarray(data).reshape(-1)
    for part in np.array_split(flat, pieces):
        assert r.append(part) == 0
    assert r.flush() == 0
    shards = sink.shards()
    meta = {lv: sink.append_sizes(lv) for lv in range(8)}
    cursor = r.cursor
    r.close()
    return shards, meta, cursor
