#!/usr/bin/env python3
"""End-to-end ingest benchmark of the streaming Zarr v3 write path on one GPU.

Workload (default): 512 sensor frames of 4096 x 2304 u16 (12-bit: seeded
Poisson(100) background + 64 Gaussian blobs per frame, sigma 3-20 px,
amplitude 200-3500), chunks 64 x 256 x 256 (t, y, x), shards of 1 x 8 x 9
chunks, one scale, zstd level 1 -- 9.66 GB of input.

One step = one acquisition: the whole 512-frame stream appended from device
memory (frames already resident in HBM before the timed region) and flushed
into a discard sink: scatter -> compress -> aggregate -> D2H -> shard
assembly. value = input GB / wall seconds over the timed steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def workload_cfg1(frames):
    import chucky_amd as ca
    from chucky_amd.api import Dim
    return ca.Config([Dim(frames, 64, 1, name="t"), Dim(4096, 256, 8, name="y"),
                      Dim(2304, 256, 9, name="x")], dtype="u16", codec="zstd", level=1,
                     buffer_capacity_bytes=64 << 20)


def make_frames(frames, H, W, seed=20260403):
    """Synthetic frames generated on the device (no host copy)."""
    import chucky_amd as ca
    import ctypes as C
    rng = np.random.default_rng(seed)
    nb = 64
    blobs = np.empty((frames, nb, 4), dtype=np.float32)
    blobs[..., 0] = rng.uniform(0, H, size=(frames, nb))
    blobs[..., 1] = rng.uniform(0, W, size=(frames, nb))
    blobs[..., 2] = rng.uniform(3, 20, size=(frames, nb))
    blobs[..., 3] = rng.uniform(200, 3500, size=(frames, nb))
    buf = ca.DeviceBuffer(frames * H * W * 2)
    rc = ca.lib.chucky_synth_sensor_frames(
        buf.ptr, frames, H, W, seed, 100.0,
        blobs.ctypes.data_as(C.POINTER(C.c_float)), nb)
    if rc:
        raise RuntimeError("synthetic frame generation failed: " + ca.last_error())
    return buf


def metric_snapshot(s):
    m = s.metrics()
    return {n: (getattr(m, n).ms, getattr(m, n).count, getattr(m, n).input_bytes)
            for n in ("scatter", "compress", "aggregate", "d2h", "sink")}


def run_gpu(frames, steps, warmup, pipeline):
    import chucky_amd as ca
    H, W = 4096, 2304
    cfg = workload_cfg1(frames)
    data = make_frames(frames, H, W)
    nbytes = frames * H * W * 2
    sink = ca.DiscardSink()
    s = ca.GpuStream(cfg, sink, pipeline=pipeline)

    def one_step():
        assert s.reset(sink) == 0
        r = s.append_device(data.ptr, nbytes)
        assert r == 0, ca.last_error()
        assert s.flush() == 0, ca.last_error()

    for _ in range(warmup):
        one_step()
    b0 = sink.bytes
    m0 = metric_snapshot(s)
    lat = []
    ca.lib.chucky_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
        lat.extend(s.latencies())
    ca.lib.chucky_device_synchronize()
    dt = time.perf_counter() - t0
    m1 = metric_snapshot(s)
    out_bytes = (sink.bytes - b0) / max(steps, 1)
    d = {k: (m1[k][0] - m0[k][0], m1[k][1] - m0[k][1], m1[k][2] - m0[k][2]) for k in m1}
    s.close()
    res = dict(seconds=dt, bytes=nbytes, out_bytes=out_bytes, metrics=d, latencies=lat,
               ms_per_step=dt * 1e3 / max(steps, 1))
    return res, data


def cpu_baseline(data, frames_total, sample_frames):
    """The reference's CPU path (oracle/_ref) on the first sample_frames of the
    same synthetic frames, all host threads, discard sink."""
    from oracle import ref
    if not ref.available():
        return None
    import chucky_amd as ca
    H, W = 4096, 2304
    n = sample_frames * H * W * 2
    host = data.to_host(n).view(np.uint16)
    cfg = workload_cfg1(sample_frames)
    sink = ca.DiscardSink()
    r = ref.RefCpuStream(cfg, sink)
    t0 = time.perf_counter()
    assert r.append(host) == 0
    assert r.flush() == 0
    dt = time.perf_counter() - t0
    r.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    return {"value": round(n / dt / 1e9, 4), "unit": "GB/s", "cores": cores,
            "kind": "reference",
            "sample": f"first {sample_frames} of {frames_total} frames "
                      f"({sample_frames // 64} epochs, {n / 1e9:.2f} GB), same data and "
                      f"config, zstd level 1 via libzstd, discard sink",
            "ratio": round(n / max(sink.bytes, 1), 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=512)
    ap.add_argument("--pipeline", choices=["plain", "batched"], default="plain")
    ap.add_argument("--cpu-sample-frames", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    pipeline = 1 if args.pipeline == "batched" else 0
    res, data = run_gpu(args.frames, args.steps, args.warmup, pipeline)
    value = res["bytes"] * args.steps / res["seconds"] / 1e9
    m = res["metrics"]
    comp_ms, comp_n, comp_in = m["compress"]
    avg_ms = comp_ms / max(comp_n, 1)
    comp_bytes_per_launch = comp_in / max(comp_n, 1)
    achieved = comp_bytes_per_launch / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    sc_ms, sc_n, sc_in = m["scatter"]
    sc_avg = sc_ms / max(sc_n, 1)
    sc_gbps = 2 * (sc_in / max(sc_n, 1)) / (sc_avg / 1e3) / 1e9 if sc_avg > 0 else 0.0
    lat = sorted(res["latencies"])
    p = (lambda q: lat[min(len(lat) - 1, int(q * len(lat)))] * 1e3) if lat else (lambda q: None)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(data, args.frames, min(args.cpu_sample_frames, args.frames))

    line = {
        "metric": "end-to-end ingest GB/s (batched vs plain, bit-exact) + p99 shard latency",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic (seeded Poisson(100) + 64 Gaussian blobs per frame, 12-bit), "
                "generated in HBM",
        "config": {"workload": f"{args.frames} frames of 4096x2304 u16, chunks 64x256x256, "
                               "shards 1x8x9 chunks, 1 scale, zstd level 1, discard sink",
                   "pipeline": args.pipeline, "input_GB": round(res["bytes"] / 1e9, 3)},
        "GiB_per_s": round(value * 1e9 / 2**30, 3),
        "compression_ratio": round(res["bytes"] / max(res["out_bytes"], 1), 4),
        "p50_epoch_latency_ms": p(0.50),
        "p99_epoch_latency_ms": p(0.99),
        "roofline": {"bound": "hbm", "kernel": "zstd_k (compress, input bytes)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "avg_launch_ms": round(avg_ms, 4),
                     "bytes_per_launch": comp_bytes_per_launch},
        "transpose": {"kernel": "transpose_vec_k", "achieved_GBps": round(sc_gbps, 1),
                      "avg_launch_ms": round(sc_avg, 4), "frac_of_hbm_peak": round(sc_gbps / HBM_PEAK_GBPS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
