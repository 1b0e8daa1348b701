"""Kernel-level GPU checks through the C ABI: the chunk scatter (transpose)
against the closed form dst[ravel(lifted_shape, lifted_strides, i)] = src[i]
(src/util/index.ops.c:42-53, tests/test_transpose.c:99-136), including
stream offsets beyond 2^32 elements, and the batch compressors by round trip
through stock libzstd / liblz4."""
import ctypes as C

import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Dim
from oracle import chucky_oracle as O
from tests.helpers import decode, xor_volume

pytestmark = pytest.mark.gpu


def _layout(cfg):
    s = ca.GpuStream(cfg, ca.MemSink())
    L = s.layout
    r = L.lifted_rank
    out = (list(L.lifted_shape[:r]), list(L.lifted_strides[:r]), L.epoch_elements,
           L.chunk_stride * L.chunks_per_epoch)
    s.close()
    return out


CASES = [
    ("u16_vec", [Dim(64, 8, 1), Dim(48, 16, 0), Dim(96, 32, 0)], "u16"),
    ("u8_elem", [Dim(30, 5, 1), Dim(21, 7, 0), Dim(13, 6, 0)], "u8"),
    ("u32_vec", [Dim(16, 4, 1), Dim(32, 8, 0), Dim(64, 16, 0)], "u32"),
    ("f64_storage", [Dim(8, 2, 1), Dim(12, 4, 0, storage_position=2), Dim(16, 8, 0, storage_position=1)], "f64"),
]


@pytest.mark.parametrize("name", [c[0] for c in CASES])
@pytest.mark.parametrize("epoch_index", [0, 3, 1 << 20])
def test_transpose_closed_form(name, epoch_index):
    _, dims, dt = next(c for c in CASES if c[0] == name)
    cfg = ca.Config(dims, dtype=dt)
    shape, strides, E, pool_elems = _layout(cfg)
    npdt = cfg.np_dtype
    rng = np.random.default_rng(epoch_index)
    src = rng.integers(0, 2**31, E, dtype=np.int64).astype(npdt)
    # a whole epoch in two pieces (second one starts mid-epoch, unaligned)
    i0 = epoch_index * E  # beyond 2^32 elements for the largest index
    cut = E // 3 + 1
    d_src = ca.DeviceBuffer.from_host(src)
    d_dst = ca.DeviceBuffer(pool_elems * npdt.itemsize)
    sh = (C.c_uint64 * len(shape))(*shape)
    st = (C.c_int64 * len(strides))(*strides)
    for a, b in ((0, cut), (cut, E)):
        assert ca.lib.chucky_gpu_transpose(d_dst.ptr + 0, d_src.ptr + a * npdt.itemsize, b - a,
                                           npdt.itemsize, i0 + a, len(shape), sh, st) == 0
    got = d_dst.to_host(pool_elems * npdt.itemsize).view(npdt)
    pos = O.ravel(shape, strides, np.arange(i0, i0 + E, dtype=np.uint64)).astype(np.int64)
    want = np.zeros(pool_elems, dtype=npdt)
    want[pos] = src
    assert np.array_equal(got[pos], src)
    assert np.array_equal(got, want)


DATA = {
    "xor_u16": lambda: xor_volume((16, 64, 64)).tobytes(),
    "zeros": lambda: bytes(300_000),
    "random": lambda: np.random.default_rng(1).integers(0, 256, 200_000, dtype=np.uint8).tobytes(),
    "sensor_like": lambda: np.clip(np.random.default_rng(2).poisson(100, 150_000), 0, 4095)
    .astype(np.uint16).tobytes(),
    # Huffman depths beyond 11 bits: length limiting on the device
    "geometric": lambda: np.random.default_rng(6).geometric(0.25, 400_000).clip(0, 255)
    .astype(np.uint8).tobytes(),
    "dominant_rare_tail": lambda: np.where(np.random.default_rng(7).random(400_000) < 0.99, 0,
                                           np.random.default_rng(8).integers(1, 256, 400_000))
    .astype(np.uint8).tobytes(),
    "two_symbols": lambda: np.random.default_rng(9).integers(0, 2, 300_000, dtype=np.uint8).tobytes(),
    "all_256": lambda: np.tile(np.arange(256, dtype=np.uint8), 1200).tobytes(),
}


@pytest.mark.parametrize("codec", ["zstd", "lz4"])
@pytest.mark.parametrize("name", sorted(DATA))
def test_compress_round_trip(codec, name):
    raw = DATA[name]()
    n_chunks = 3
    chunk = len(raw) // n_chunks
    cid = ca.CODEC_ID[codec]
    bound = ca.lib.chucky_gpu_max_output_size(cid, chunk)
    d_src = ca.DeviceBuffer.from_host(np.frombuffer(raw, dtype=np.uint8))
    d_dst = ca.DeviceBuffer(bound * n_chunks)
    d_sizes = ca.DeviceBuffer(8 * n_chunks)
    assert ca.lib.chucky_gpu_compress(cid, 1, d_src.ptr, chunk, chunk, n_chunks, d_dst.ptr, bound,
                                      d_sizes.ptr) == 0, ca.last_error()
    sizes = d_sizes.to_host(8 * n_chunks).view(np.uint64)
    out = d_dst.to_host(bound * n_chunks).tobytes()
    for i in range(n_chunks):
        blob = out[i * bound: i * bound + int(sizes[i])]
        assert decode(codec, blob, chunk) == raw[i * chunk:(i + 1) * chunk], f"chunk {i}"
