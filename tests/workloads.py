"""The benchmark workloads, defined once for the parity tests and bench.py.

Each workload is named by what it is and its sizes:

  volume_lz4          256^3 u16 xor+noise, chunks 64^3, shards 4^3 chunks, one
                      scale, lz4 (the reference's CPU-runnable case)
  sensor_zstd         512 frames of 4096x2304 u16 (12-bit Poisson(100)
                      background + 64 Gaussian blobs a frame), chunks
                      64x256x256, shards 1x8x9 chunks, one scale, zstd level 1
  multiscale_zstd     256^3 u16 (xor+noise or thirds), mean LOD over z, y, x with
                      dim-0 downsampling, chunks 64^3, shards 4^3 chunks, zstd 1:
                      three levels (the reference stops the pyramid once every
                      LOD dim fits one chunk, src/lod/lod_plan.c:213-275)
  multiscale6_zstd    the same volumes as timepoints of a leading append dim t
                      (chunk 1, downsampled), 2x2x2 mean windows over z, y, x
                      in 8^3 chunks: six levels (256 ... 8), shards 4^3 chunks
  sweep               512x1024x1024 per dtype (u8, i16, u32, f16, f32, f64),
                      mean / median / max_sup / min_sup over y, x, chunks 32^3,
                      64^3 or 16x128x128 with a padded edge, lz4 or zstd
  serving             24 seeded Poisson arrivals (4/s) over u8 2048x512x512,
                      u16 256x2048x2048 and f32 256^3

Reference shapes: bench/bench_stream_256cube_single.c,
bench/bench_stream_256cube_multiscale_dim0.c:6-14 (t, z, y, x downsampled).
"""
from __future__ import annotations

import numpy as np

import chucky_amd as ca
from chucky_amd.api import Dim
from tests.helpers import random_typed, thirds_volume, xor_volume

SENSOR_SEED = 20260403


# ---- configurations -----------------------------------------------------------

def volume_lz4(n=256):
    return ca.Config([Dim(n, 64, 4, name="z"), Dim(n, 64, 4, name="y"), Dim(n, 64, 4, name="x")],
                     dtype="u16", codec="lz4", level=1, buffer_capacity_bytes=16 << 20)


def sensor_zstd(frames=512, height=4096, width=2304):
    return ca.Config([Dim(frames, 64, 1, name="t"), Dim(height, 256, 8, name="y"),
                      Dim(width, 256, 9, name="x")], dtype="u16", codec="zstd", level=1,
                     buffer_capacity_bytes=64 << 20)


def multiscale_zstd(n=256):
    return ca.Config([Dim(n, 64, 4, name="z", downsample=True), Dim(n, 64, 4, name="y", downsample=True),
                      Dim(n, 64, 4, name="x", downsample=True)], dtype="u16", codec="zstd", level=1,
                     reduce="mean", append_reduce="mean", buffer_capacity_bytes=16 << 20)


def multiscale6_zstd(timepoints=2, n=256):
    return ca.Config([Dim(timepoints, 1, 1, name="t", downsample=True),
                      Dim(n, 8, 4, name="z", downsample=True), Dim(n, 8, 4, name="y", downsample=True),
                      Dim(n, 8, 4, name="x", downsample=True)], dtype="u16", codec="zstd", level=1,
                     reduce="mean", append_reduce="mean", buffer_capacity_bytes=16 << 20)


SWEEP_DTYPES = ["u8", "i16", "u32", "f16", "f32", "f64"]
SWEEP_METHODS = ["mean", "median", "max_sup", "min_sup"]
SWEEP_CHUNKS = {"c32": (32, 32, 32), "c64": (64, 64, 64), "c16x128": (16, 128, 128)}


def sweep(dtype, chunk, method, codec, frames=512, edge=1000):
    """The dtype / reduction sweep: y and x of `edge` (not divisible by the
    chunk: a padded edge), LOD over y, x, at most 5 levels."""
    ct, cy, cx = SWEEP_CHUNKS[chunk]
    return ca.Config([Dim(frames, ct, 1, name="t"), Dim(edge, cy, 4, name="y", downsample=True),
                      Dim(edge, cx, 4, name="x", downsample=True)], dtype=dtype, codec=codec, level=1,
                     reduce=method, max_nlod=5, buffer_capacity_bytes=16 << 20)


def sweep_cases():
    """(dtype, chunk, method, codec) for every dtype x chunk shape, methods
    and codecs rotated."""
    out = []
    for i, dt in enumerate(SWEEP_DTYPES):
        for j, ch in enumerate(sorted(SWEEP_CHUNKS)):
            out.append((dt, ch, SWEEP_METHODS[(i + j) % 4], "lz4" if (i * 3 + j) % 2 == 0 else "zstd"))
    return out


# ---- serving mix ---------------------------------------------------------------

SERVING_SHAPES = [("u8", (2048, 512, 512)), ("u16", (256, 2048, 2048)), ("f32", (256, 256, 256))]


def serving_requests(n=24, rate=4.0, seed=20260404):
    """Seeded request mix: (arrival_s, dtype, shape, chunk, codec, nlod)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    out = []
    for i in range(n):
        dt, shape = SERVING_SHAPES[rng.integers(3)]
        chunk = (64, 64, 64) if rng.integers(2) == 0 else (64, 256, 256)
        codec = "zstd" if rng.integers(2) == 0 else "lz4"
        nlod = int(rng.integers(1, 6))
        out.append((float(t[i]), dt, shape, chunk, codec, nlod))
    return out


def serving_config(dt, shape, chunk, codec, nlod):
    lod = nlod > 1
    dims = [Dim(shape[0], chunk[0], 1, name="t"),
            Dim(shape[1], min(chunk[1], shape[1]), 4, name="y", downsample=lod),
            Dim(shape[2], min(chunk[2], shape[2]), 4, name="x", downsample=lod)]
    cfg = ca.Config(dims, dtype=dt, codec=codec, level=1, max_nlod=nlod if lod else 0,
                    buffer_capacity_bytes=64 << 20)
    # a bounded stream never needs a batch longer than itself
    epochs = -(-shape[0] // chunk[0])
    k = ca.api.memory_estimate(cfg).epochs_per_batch
    if k > epochs:
        cfg.epochs_per_batch = 1 << (epochs - 1).bit_length()
    return cfg


# ---- data ------------------------------------------------------------------------

def sweep_data(dtype, n, seed):
    """Integers uniform over the full range, floats N(0,1)*100."""
    if dtype in ("f16", "f32", "f64"):
        rng = np.random.default_rng(seed)
        return (rng.standard_normal(n) * 100).astype(ca.NP_DTYPE[dtype])
    return random_typed(n, dtype, seed=seed)


def multiscale_volumes(n=256):
    """The xor+noise volume and the thirds volume (timepoints 0 and 1)."""
    return np.stack([xor_volume((n, n, n)), thirds_volume((n, n, n))])


def sensor_blobs(frames, height, width, seed=SENSOR_SEED, nblob=64):
    rng = np.random.default_rng(seed)
    blobs = np.empty((frames, nblob, 4), dtype=np.float32)
    blobs[..., 0] = rng.uniform(0, height, size=(frames, nblob))
    blobs[..., 1] = rng.uniform(0, width, size=(frames, nblob))
    blobs[..., 2] = rng.uniform(3, 20, size=(frames, nblob))
    blobs[..., 3] = rng.uniform(200, 3500, size=(frames, nblob))
    return blobs


def sensor_frames_device(frames, height=4096, width=2304, seed=SENSOR_SEED):
    """12-bit sensor frames generated in device memory (a DeviceBuffer)."""
    import ctypes as C
    blobs = sensor_blobs(frames, height, width, seed)
    buf = ca.DeviceBuffer(frames * height * width * 2)
    if ca.lib.chucky_synth_sensor_frames(buf.ptr, frames, height, width, seed, 100.0,
                                         blobs.ctypes.data_as(C.POINTER(C.c_float)), blobs.shape[1]):
        raise RuntimeError("synthetic frame generation failed: " + ca.last_error())
    return buf
