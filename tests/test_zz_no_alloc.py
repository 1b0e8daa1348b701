"""chucky_gpu_compress_noalloc (no counterpart in the reference; for callers
that cannot allocate) against chucky_gpu_compress, through the C ABI: the
same sizes and the same compressed bytes for every chunk, frames that stock
liblz4 / libzstd decode to the input; two calls in a row, the second with a
smaller input, so the shared scratch array is reused; and a call past the
scratch limits is refused with an error, launching nothing."""
import ctypes as C

import numpy as np
import pytest

import chucky_amd as ca
from tests.helpers import decode, random_typed, xor_volume

pytestmark = pytest.mark.gpu


def _inputs():
    # structured and incompressible data, 1.25 MiB
    xor = xor_volume((64, 64, 64)).reshape(-1).view(np.uint8)
    noise = random_typed(64 * 64 * 64, "u16", seed=5).view(np.uint8)
    return np.concatenate([xor, noise, xor[: len(xor) // 2]])


def _run(fn, codec, host, chunk_bytes, n, noalloc):
    bound = ca.lib.chucky_gpu_max_output_size(codec, chunk_bytes)
    stride = (bound + 255) // 256 * 256
    src = ca.DeviceBuffer.from_host(host[: n * chunk_bytes])
    dst = ca.DeviceBuffer(n * stride)
    sizes = ca.DeviceBuffer(n * 8)
    if noalloc:
        rc = fn(codec, 1, src.ptr, chunk_bytes, chunk_bytes, n, dst.ptr, stride, sizes.ptr, None)
        assert rc == 0, ca.last_error()
        assert ca.lib.chucky_device_synchronize() == 0
    else:
        assert fn(codec, 1, src.ptr, chunk_bytes, chunk_bytes, n, dst.ptr, stride, sizes.ptr) == 0, \
            ca.last_error()
    sz = sizes.to_host(n * 8).view(np.uint64)
    out = dst.to_host(n * stride)
    frames = [out[i * stride:i * stride + int(sz[i])].tobytes() for i in range(n)]
    return sz, frames


@pytest.mark.parametrize("codec", [ca.CODEC_ZSTD, ca.CODEC_LZ4])
def test_noalloc_matches_the_allocating_entry_point(codec):
    host = _inputs()
    chunk = 128 * 1024 if codec == ca.CODEC_LZ4 else 512 * 1024
    name = {ca.CODEC_ZSTD: "zstd", ca.CODEC_LZ4: "lz4"}[codec]
    for n in (len(host) // chunk, 1):  # then a smaller call on the same scratch
        print(f"[noalloc] codec={name} chunks={n} chunk_bytes={chunk}", flush=True)
        want_sz, want = _run(ca.lib.chucky_gpu_compress, codec, host, chunk, n, False)
        got_sz, got = _run(ca.lib.chucky_gpu_compress_noalloc, codec, host, chunk, n, True)
        for i in range(n):
            src = host[i * chunk:(i + 1) * chunk].tobytes()
            assert decode(name, want[i], chunk) == src, f"allocating entry point, chunk {i}"
            assert decode(name, got[i], chunk) == src, f"no-alloc entry point, chunk {i}"
        np.testing.assert_array_equal(got_sz, want_sz)
        for i in range(n):
            assert got[i] == want[i], f"chunk {i} differs"


def test_noalloc_refuses_calls_past_its_limits():
    # 65537 one-block zstd chunks: refused before any launch
    rc = ca.lib.chucky_gpu_compress_noalloc(ca.CODEC_ZSTD, 1, C.c_void_p(16), 4096, 4096, 65537,
                                            C.c_void_p(16), 1 << 20, C.c_void_p(16), None)
    assert rc != 0
    assert "exceed the limit" in ca.last_error()
