"""The zstd encoding core (chucky_amd/csrc/zstd_core.hpp: FSE tables, Huffman
codes and weights, sequence and literal sections, frame/block headers),
compiled for the host with a serial parser, must emit frames that stock
libzstd decodes to the original bytes. Runs on the CPU only."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import decode, thirds_volume, xor_volume

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "zcore_host.cpp")
LIB = os.path.join(HERE, "native", "libzcore_host.so")


def _lib():
    hdr = os.path.join(ROOT, "chucky_amd", "csrc", "zstd_core.hpp")
    if (not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC)
            or os.path.getmtime(LIB) < os.path.getmtime(hdr)):
        subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I",
                        os.path.join(ROOT, "chucky_amd", "csrc"), SRC, "-o", LIB], check=True)
    lib = C.CDLL(LIB)
    lib.zcore_host_compress.restype = C.c_size_t
    lib.zcore_host_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_int]
    return lib


def _roundtrip(data: bytes, force_raw=0):
    lib = _lib()
    src = C.create_string_buffer(data, len(data))
    dst = C.create_string_buffer(len(data) + len(data) // 64 + 1024)
    n = lib.zcore_host_compress(src, len(data), dst, force_raw)
    blob = dst.raw[:n]
    assert decode("zstd", blob, len(data)) == data
    return len(data) / n


def _sensor(n_px, seed=3):
    rng = np.random.default_rng(seed)
    v = rng.poisson(100, n_px).astype(np.float64)
    v[: n_px // 10] += 1500 * np.exp(-np.linspace(-3, 3, n_px // 10) ** 2)
    return np.clip(np.rint(v), 0, 4095).astype(np.uint16).tobytes()


CASES = {
    "zeros": lambda: bytes(300_000),
    "random": lambda: np.random.default_rng(1).integers(0, 256, 200_000, dtype=np.uint8).tobytes(),
    "xor": lambda: xor_volume((16, 64, 64)).tobytes(),
    "thirds": lambda: thirds_volume((8, 64, 192)).tobytes(),
    "sensor": lambda: _sensor(150_000),
    "ramp": lambda: (np.arange(100_000) % 251).astype(np.uint8).tobytes(),
    "two_symbols": lambda: np.random.default_rng(2).integers(0, 2, 70_000, dtype=np.uint8).tobytes(),
    "tiny": lambda: b"abc",
    "just_over_block": lambda: np.random.default_rng(4).integers(0, 40, 131_073, dtype=np.uint8).tobytes(),
    "geometric_lengths": lambda: np.random.default_rng(6).geometric(0.25, 200_000).clip(0, 255).astype(np.uint8).tobytes(),
    "dominant_zero_rare_tail": lambda: np.where(np.random.default_rng(7).random(200_000) < 0.99, 0,
                                                np.random.default_rng(8).integers(1, 256, 200_000)).astype(np.uint8).tobytes(),
    "skewed_257_symbols": lambda: np.random.default_rng(5).zipf(1.3, 90_000).clip(0, 255).astype(np.uint8).tobytes(),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_core_frames_decode_with_libzstd(name):
    r = _roundtrip(CASES[name]())
    assert r > 0


def test_huffman_pays_on_12bit_data():
    # 12-bit little-endian samples: Huffman literals must beat raw literals
    data = _sensor(200_000)
    r_huf = _roundtrip(data)
    r_raw = _roundtrip(data, force_raw=1)
    assert r_huf > 1.5 and r_huf > r_raw * 1.4
