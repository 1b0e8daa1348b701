#!/usr/bin/env python3
"""End-to-end ingest benchmark of the streaming Zarr v3 write path on one GPU.

Default workload (the configuration the metric is quoted on): 512 sensor
frames of 4096 x 2304 u16 (12-bit: seeded Poisson(100) background + 64
Gaussian blobs per frame, sigma 3-20 px, amplitude 200-3500), chunks
64 x 256 x 256 (t, y, x), shards of 1 x 8 x 9 chunks, one scale, zstd level 1
-- 9.66 GB of input.

One step = one acquisition: the whole stream appended from device memory
(frames already resident in HBM before the timed region) and flushed into a
discard sink: scatter -> compress -> aggregate -> D2H -> shard assembly.
value = input GB / wall seconds over the timed steps of the batched pipeline;
the plain (reference-shaped, one launch per stage, host syncs) pipeline is
timed the same way and reported beside it.

Other workloads (--workload): `volume` = 256^3 u16 xor+noise volume, 64^3
chunks, 4^3 chunks per shard, lz4 (the reference's CPU-runnable case);
`multiscale` = the same volume with a mean 2x2 LOD over y, x and dim-0
downsampling (3 levels), zstd level 1. `--host-input` measures the
PCIe-inclusive rate through the pinned staging ring instead.

`serving`: 24 requests arriving as a seeded Poisson process (mean 4/s), each
drawn uniformly from u8 2048 frames of 512x512, u16 256 frames of 2048x2048
and f32 256^3, with 64^3 or 64x256x256 chunks, zstd or lz4 and 1-5 LOD levels
over y, x; admitted by the serving loop against a 200 GB HBM budget. value =
total input GB / (last request done - first arrival); reported beside the
same requests run one after another through the plain pipeline.
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


def workload(name, frames):
    import chucky_amd as ca
    from chucky_amd.api import Dim
    if name == "sensor":
        cfg = ca.Config([Dim(frames, 64, 1, name="t"), Dim(4096, 256, 8, name="y"),
                         Dim(2304, 256, 9, name="x")], dtype="u16", codec="zstd", level=1,
                        buffer_capacity_bytes=64 << 20)
        desc = (f"{frames} frames of 4096x2304 u16, chunks 64x256x256, shards 1x8x9 chunks, "
                "1 scale, zstd level 1, discard sink")
        return cfg, (frames, 4096, 2304), desc
    if name == "lod":
        cfg = ca.Config([Dim(frames, 16, 1, name="t"), Dim(2048, 256, 8, name="y", downsample=True),
                         Dim(2048, 256, 8, name="x", downsample=True)], dtype="u16", codec="none",
                        reduce="mean", buffer_capacity_bytes=64 << 20)
        return cfg, (frames, 2048, 2048), (f"{frames} frames of 2048x2048 u16, mean 2x2 LOD over y,x "
                                          "(4 levels), chunks 16x256x256, no codec, discard sink")
    n = 256
    if name == "volume":
        cfg = ca.Config([Dim(n, 64, 4, name="z"), Dim(n, 64, 4, name="y"), Dim(n, 64, 4, name="x")],
                        dtype="u16", codec="lz4", level=1, buffer_capacity_bytes=16 << 20)
        return cfg, (n, n, n), "256^3 u16 xor+noise, chunks 64^3, shards 4^3 chunks, lz4, discard sink"
    cfg = ca.Config([Dim(n, 64, 4, name="z", downsample=True), Dim(n, 64, 4, name="y", downsample=True),
                     Dim(n, 64, 4, name="x", downsample=True)], dtype="u16", codec="zstd", level=1,
                    buffer_capacity_bytes=16 << 20)
    return cfg, (n, n, n), ("256^3 u16 xor+noise, mean LOD y,x + dim0 fold (3 levels), chunks 64^3, "
                            "shards 4^3 chunks, zstd level 1, discard sink")


def make_input(name, shape, seed=20260403):
    """Synthetic input generated on (sensor) or copied to the device."""
    import ctypes as C
    import chucky_amd as ca
    if name == "sensor":
        frames, H, W = shape
        rng = np.random.default_rng(seed)
        nb = 64
        blobs = np.empty((frames, nb, 4), dtype=np.float32)
        blobs[..., 0] = rng.uniform(0, H, size=(frames, nb))
        blobs[..., 1] = rng.uniform(0, W, size=(frames, nb))
        blobs[..., 2] = rng.uniform(3, 20, size=(frames, nb))
        blobs[..., 3] = rng.uniform(200, 3500, size=(frames, nb))
        buf = ca.DeviceBuffer(frames * H * W * 2)
        if ca.lib.chucky_synth_sensor_frames(buf.ptr, frames, H, W, seed, 100.0,
                                             blobs.ctypes.data_as(C.POINTER(C.c_float)), nb):
            raise RuntimeError("synthetic frame generation failed: " + ca.last_error())
        return buf, None
    if name == "lod":
        n = int(np.prod(shape))
        buf = ca.DeviceBuffer(n * 2)
        if ca.lib.chucky_synth_fill(buf.ptr, n, ca.DTYPE_ID["u16"], seed, 4096):
            raise RuntimeError("fill failed: " + ca.last_error())
        return buf, None
    from tests.helpers import xor_volume
    vol = xor_volume(shape)
    return ca.DeviceBuffer.from_host(vol), vol


def run(name, frames, steps, warmup, pipelines, host_input, options=None):
    """Times `steps` steps of each pipeline on the same input. The pipelines'
    steps alternate (in alternating order), each step bracketed by device
    synchronizations, so that both see the same state of the machine: the
    device-to-host copy rate of these hosts falls by about half for a minute
    after a few seconds of full load, which would otherwise favour whichever
    pipeline runs first."""
    import chucky_amd as ca
    cfg, shape, desc = workload(name, frames)
    data, host = make_input(name, shape)
    nbytes = int(np.prod(shape)) * 2
    if host_input and host is None:
        host = data.to_host(nbytes).view(np.uint16)
    streams = {}
    for p in pipelines:
        sink = ca.DiscardSink()
        streams[p] = (ca.GpuStream(cfg, sink, pipeline=1 if p == "batched" else 0, **(options or {})), sink)

    def one_step(p):
        s, sink = streams[p]
        assert s.reset(sink) == 0
        if host_input:
            assert s.append(host) == 0, ca.last_error()
        else:
            assert s.append_device(data.ptr, nbytes) == 0, ca.last_error()
        assert s.flush() == 0, ca.last_error()

    for _ in range(warmup):
        for p in pipelines:
            one_step(p)
    b0 = {p: streams[p][1].bytes for p in pipelines}
    m0 = {p: snapshot(streams[p][0]) for p in pipelines}
    lat = {p: [] for p in pipelines}
    step_ms = {p: [] for p in pipelines}
    for i in range(steps):
        for p in (pipelines if i % 2 == 0 else pipelines[::-1]):
            ca.lib.chucky_device_synchronize()
            ts = time.perf_counter()
            one_step(p)
            ca.lib.chucky_device_synchronize()
            step_ms[p].append((time.perf_counter() - ts) * 1e3)
            lat[p].extend(streams[p][0].latencies())
    res = {}
    for p in pipelines:
        s, sink = streams[p]
        m1 = snapshot(s)
        dt = sum(step_ms[p]) / 1e3
        res[p] = dict(seconds=dt, bytes=nbytes, out_bytes=(sink.bytes - b0[p]) / max(steps, 1), desc=desc,
                      metrics={k: tuple(m1[k][i] - m0[p][k][i] for i in range(4)) for k in m1},
                      latencies=lat[p], ms_per_step=dt * 1e3 / max(steps, 1),
                      step_ms=[round(v, 2) for v in step_ms[p]])
        s.close()
    return res, data, cfg


LOD_PASSES = {"lod_gather": "tiles_k (linear epoch -> level-0 Morton)",
              "lod_reduce": "pyramid_k (every level from level 0, one launch)",
              "lod_append_fold": "fold_emit_k (append-dim fold and emit, all levels, one launch)",
              "lod_morton_chunk": "tiles_levels_k (Morton levels -> chunk pool, levels of one "
                                  "chunk geometry in one launch)"}


def snapshot(s):
    m = s.metrics()
    return {n: (getattr(m, n).ms, getattr(m, n).count, getattr(m, n).input_bytes,
                getattr(m, n).output_bytes)
            for n in ("scatter", "compress", "aggregate", "d2h", "sink", *LOD_PASSES)}


def pass_rate(metric):
    """(avg ms per launch, algorithmic bytes read+written per launch, GB/s)."""
    ms, n, rd, wr = metric
    if not n or ms <= 0:
        return 0.0, 0.0, 0.0
    avg = ms / n
    per = (rd + wr) / n
    return avg, per, per / (avg / 1e3) / 1e9


def cpu_baseline(name, data, frames_total, sample_frames):
    """The reference's CPU path (oracle/_ref) on a bounded sample of the same
    input, all host threads, discard sink."""
    from oracle import ref
    if not ref.available():
        return None
    import chucky_amd as ca
    cfg, shape, _ = workload(name, sample_frames if name == "sensor" else frames_total)
    n = int(np.prod(shape)) * 2
    host = data.to_host(n).view(np.uint16)
    sink = ca.DiscardSink()
    r = ref.RefCpuStream(cfg, sink)
    t0 = time.perf_counter()
    assert r.append(host) == 0
    assert r.flush() == 0
    dt = time.perf_counter() - t0
    r.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    sample = (f"first {sample_frames} of {frames_total} frames ({n / 1e9:.2f} GB), same data and "
              f"config" if name == "sensor" else f"whole input ({n / 1e9:.3f} GB)")
    return {"value": round(n / dt / 1e9, 4), "unit": "GB/s", "cores": cores, "kind": "reference",
            "sample": sample + ", reference CPU path built from its sources (libzstd/liblz4), "
                               "discard sink",
            "ratio": round(n / max(sink.bytes, 1), 4)}


def summarize(res, steps):
    value = res["bytes"] * steps / res["seconds"] / 1e9
    m = res["metrics"]
    comp_ms, comp_n, comp_in, _ = m["compress"]
    lat = sorted(res["latencies"])
    pct = (lambda q: round(lat[min(len(lat) - 1, int(q * len(lat)))] * 1e3, 3)) if lat else (lambda q: None)
    stages = {k: {"ms": round(v[0] / steps, 3), "launches": v[1] // max(steps, 1),
                  "GBps": round(v[2] / v[0] / 1e6, 2) if v[0] > 0 else None}
              for k, v in m.items() if v[1] and k not in LOD_PASSES}
    return value, dict(GBps=round(value, 3), ms_per_step=round(res["ms_per_step"], 3),
                       ratio=round(res["bytes"] / max(res["out_bytes"], 1), 4),
                       p50_epoch_latency_ms=pct(0.50), p99_epoch_latency_ms=pct(0.99),
                       step_ms=res.get("step_ms"), stages_per_step=stages)


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
    import chucky_amd as ca
    from chucky_amd.api import Dim
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


def serving(args):
    import chucky_amd as ca
    from chucky_amd.api import Server
    reqs = serving_requests()
    inputs = {}
    for k, (dt, shape) in enumerate(SERVING_SHAPES):
        n = int(np.prod(shape))
        buf = ca.DeviceBuffer(n * np.dtype(ca.NP_DTYPE[dt]).itemsize)
        if ca.lib.chucky_synth_fill(buf.ptr, n, ca.DTYPE_ID[dt], 1000 + k, 0):
            raise RuntimeError("fill failed: " + ca.last_error())
        inputs[dt] = buf
    total = sum(inputs[r[1]].nbytes for r in reqs)

    # serial, plain pipeline: one stream at a time
    t0 = time.perf_counter()
    serial_lat = []
    for _, dt, shape, chunk, codec, nlod in reqs:
        sink = ca.DiscardSink()
        s = ca.GpuStream(serving_config(dt, shape, chunk, codec, nlod), sink, pipeline=ca.PIPELINE_PLAIN)
        eb = s.epoch_elements * np.dtype(ca.NP_DTYPE[dt]).itemsize
        buf = inputs[dt]
        for off in range(0, buf.nbytes, eb):
            assert s.append_device(buf.ptr + off, min(eb, buf.nbytes - off)) == 0, ca.last_error()
        assert s.flush() == 0, ca.last_error()
        serial_lat.extend(s.latencies())
        s.close()
    serial_s = time.perf_counter() - t0

    srv = Server(hbm_budget_bytes=200 * 10**9, max_concurrent=24, pipeline=ca.PIPELINE_BATCHED)
    sinks = []
    t0 = time.perf_counter()
    for arrival, dt, shape, chunk, codec, nlod in reqs:
        wait = arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        sinks.append(ca.DiscardSink())
        srv.submit(serving_config(dt, shape, chunk, codec, nlod), sinks[-1], inputs[dt], device=True)
    assert srv.wait() == 0, ca.last_error()
    reps = srv.reports()
    lat = sorted(srv.latencies())
    running, peak_bytes = srv.peaks()
    makespan = max(r.end_s for r in reps) - min(r.submit_s for r in reps)
    srv.close()
    pct = lambda v, q: round(v[min(len(v) - 1, int(q * len(v)))] * 1e3, 3) if v else None
    serial_lat.sort()
    value = total / makespan / 1e9
    return {
        "metric": "end-to-end ingest GB/s (batched vs plain, bit-exact) + p99 shard latency",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(makespan * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/u16/f32", "data": "synthetic, generated in HBM",
        "config": {"workload": "serving: 24 Poisson arrivals (4/s) over u8 2048x512x512, "
                               "u16 256x2048x2048, f32 256^3; 64^3 or 64x256x256 chunks; zstd/lz4; "
                               "1-5 LOD levels; 200 GB HBM budget; discard sinks",
                   "input_GB": round(total / 1e9, 3)},
        "p50_epoch_latency_ms": pct(lat, 0.50), "p99_epoch_latency_ms": pct(lat, 0.99),
        "peak_concurrent": running, "peak_admitted_GB": round(peak_bytes / 1e9, 3),
        "serial_plain": {"GBps": round(total / serial_s / 1e9, 3), "seconds": round(serial_s, 3),
                         "p50_epoch_latency_ms": pct(serial_lat, 0.50),
                         "p99_epoch_latency_ms": pct(serial_lat, 0.99)},
        "served_over_serial": round(serial_s / makespan, 3),
        "arrival_span_s": round(reqs[-1][0], 3),
    }


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE,
    scaled by a copy of known size), or None when none covers the pair."""
    files = {"sensor": "zstd_k_sensor512x4096x2304_pmc.md", "lod": "tiles_k_lod64x2048x2048_pmc.md"}
    if workload not in files:
        return None
    path = os.path.join(ROOT, "profiles", files[workload])
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def calibrate_copy(nbytes):
    """One identity-layout scatter of nbytes (a plain 16-byte-vector copy):
    the known-size kernel that scales FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    import chucky_amd as ca
    n = nbytes // 4
    src, dst = ca.DeviceBuffer(n * 4), ca.DeviceBuffer(n * 4)
    shape, strides = (C.c_uint64 * 1)(n), (C.c_int64 * 1)(1)
    t0 = time.perf_counter()
    if ca.lib.chucky_gpu_transpose(dst.ptr, src.ptr, n, 4, 0, 1, shape, strides):
        raise RuntimeError("copy failed: " + ca.last_error())
    dt = time.perf_counter() - t0
    return {"copy_bytes": n * 4, "kernel": "transpose_vec_k<4>", "seconds": round(dt, 6)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=512)
    ap.add_argument("--workload", choices=["sensor", "volume", "multiscale", "lod", "serving"],
                    default="sensor")
    ap.add_argument("--pipeline", choices=["plain", "batched", "both"], default="both")
    ap.add_argument("--host-input", action="store_true")
    ap.add_argument("--cpu-sample-frames", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--zstd-block", type=int, default=0, help="zstd block bytes (0 = autotuned)")
    ap.add_argument("--gather-wg", type=int, default=0,
                    help="workgroups of the gather into pinned host memory (0 = autotuned)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one plain device copy of BYTES (the vector scatter kernel over an "
                         "identity layout), the reference for counter-based HBM traffic")
    ap.add_argument("--batched-d2h", type=int, default=0,
                    help="batched pipeline D2H: 0 = device gather + DMA, 1 = gather kernel into host")
    ap.add_argument("--batched-launch", type=int, default=0,
                    help="batched pipeline per-epoch work: 0 = autotuned, 1 = captured graphs, "
                         "2 = direct launches")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        print(json.dumps(calibrate_copy(args.calibrate_copy)))
        return
    if args.workload == "serving":
        print(json.dumps(serving(args)))
        return

    order = ["plain", "batched"] if args.pipeline == "both" else [args.pipeline]
    runs, data, _ = run(args.workload, args.frames, args.steps, args.warmup, order, args.host_input,
                        {"zstd_block_bytes": args.zstd_block,
                         "host_gather_workgroups": args.gather_wg,
                         "batched_d2h": args.batched_d2h,
                         "batched_launch": args.batched_launch})
    head = runs["batched" if "batched" in runs else order[0]]
    value, head_sum = summarize(head, args.steps)
    m = head["metrics"]
    comp_ms, comp_n, comp_in, _ = m["compress"]
    avg_ms = comp_ms / max(comp_n, 1)
    bytes_per_launch = comp_in / max(comp_n, 1)
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    sc_ms, sc_n, sc_in, _ = m["scatter"]
    sc_avg = sc_ms / max(sc_n, 1)
    sc_gbps = 2 * (sc_in / max(sc_n, 1)) / (sc_avg / 1e3) / 1e9 if sc_avg > 0 else 0.0
    lod = {}
    for k, label in LOD_PASSES.items():
        avg, per, gbps = pass_rate(m[k])
        if per:
            lod[k] = {"kernels": label, "avg_launch_ms": round(avg, 4), "bytes_per_launch": per,
                      "achieved_GBps": round(gbps, 1), "frac_of_hbm_peak": round(gbps / HBM_PEAK_GBPS, 4)}
    codec_none = "no codec" in head["desc"]
    if codec_none and "lod_gather" in lod:
        # no compressor: the dominant kernel is the LOD gather (one launch per epoch)
        g = lod["lod_gather"]
        roofline = {"bound": "hbm", "kernel": "tiles_k (LOD gather)", "achieved": g["achieved_GBps"],
                    "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": g["frac_of_hbm_peak"],
                    "traffic": measured_traffic(args.workload, "tiles_k"),
                    "avg_launch_ms": g["avg_launch_ms"], "bytes_per_launch": g["bytes_per_launch"],
                    "note": "epoch bytes read + Morton bytes written per launch / event-timed launch duration"}
    else:
        kname = "zstd_k" if "zstd" in head["desc"] else "lz4_chunk_k"
        roofline = {"bound": "hbm", "kernel": kname,
                    "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "traffic": measured_traffic(args.workload, kname),
                    "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": bytes_per_launch,
                    "note": "compressor input bytes per launch / event-timed launch duration"}

    cpu = None
    if not args.no_cpu_baseline and data is not None:
        cpu = cpu_baseline(args.workload, data, args.frames,
                           min(args.cpu_sample_frames, args.frames))

    line = {
        "metric": "end-to-end ingest GB/s (batched vs plain, bit-exact) + p99 shard latency",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head_sum["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": ("synthetic, generated in HBM" if args.workload == "sensor" else "synthetic, copied to HBM")
                + (" (appended from host memory: PCIe-inclusive)" if args.host_input else ""),
        "config": {"workload": head["desc"], "pipeline": "batched" if "batched" in runs else order[0],
                   "input_GB": round(head["bytes"] / 1e9, 3)},
        "GiB_per_s": round(value * 1e9 / 2**30, 3),
        "compression_ratio": head_sum["ratio"],
        "p50_epoch_latency_ms": head_sum["p50_epoch_latency_ms"],
        "p99_epoch_latency_ms": head_sum["p99_epoch_latency_ms"],
        "pipelines": {k: summarize(v, args.steps)[1] for k, v in runs.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if sc_n:
        line["transpose"] = {"kernel": "transpose_vec_k", "achieved_GBps": round(sc_gbps, 1),
                             "avg_launch_ms": round(sc_avg, 4),
                             "frac_of_hbm_peak": round(sc_gbps / HBM_PEAK_GBPS, 4)}
    if lod:
        line["lod"] = lod
    if "plain" in runs and "batched" in runs:
        line["batched_over_plain"] = round(runs["plain"]["seconds"] / runs["batched"]["seconds"], 3)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
