#!/usr/bin/env python3
"""End-to-end ingest benchmark of the streaming Zarr v3 write path on one GPU.

Headline workload (the multiscale batching case the batched-vs-plain target
is quoted on): a 256^3 u16 volume, ((x^y^z)*257 + n) mod 65536 with seeded
noise n in [0, 16), mean LOD over z, y, x with dim-0 downsampling (3 levels:
the pyramid stops once every LOD dim fits one 64^3 chunk), chunks 64^3,
shards of 4^3 chunks, zstd level 1, discard sink -- 33.5 MB of input.

One step = one acquisition appended from HOST memory through the writer
(pinned staging -> H2D -> LOD -> Morton -> chunks -> compress -> aggregate ->
D2H -> shard assembly) and flushed: the reference's end-to-end definition,
input bytes / wall time of append + flush (bench/bench_util.c:671-692).
value = the batched pipeline's GB/s; the plain pipeline (one launch per stage,
host waits; the reference's shape) is timed the same way beside it, the two
pipelines' steps alternating.

Outside the timed region both pipelines' shards are compared with the
reference CPU path's (oracle/_ref) chunk by chunk after decode, and the
reference CPU path is timed on the same input (cpu_baseline).

Beside the headline (extra keys, never `value`): the same workload with the
input already in HBM; the six-level 2x2x2 variant (two timepoints of 8^3
chunks); 512 sensor frames of 4096x2304 (64x256x256 chunks, 1x8x9-chunk
shards, zstd 1) from host memory; the 256^3 lz4 volume from host memory; the
headline workload written to a zarr filesystem store (page cache and
O_DIRECT); the serving mix.

Other workloads (--workload): `sensor`, `volume`, `multiscale6`, `lod`
(2048^2 frames, 4-level mean LOD, no codec), `serving` (24 Poisson
arrivals), `fs` (the filesystem-store leg alone); `--device-input` times
HBM-resident input instead.
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
METRIC = "end-to-end ingest GB/s (batched vs plain, bit-exact) + p99 shard latency"


def workload(name, frames=512):
    import chucky_amd as ca
    from chucky_amd.api import Dim
    from tests import workloads as W
    if name == "sensor":
        return (W.sensor_zstd(frames), (frames, 4096, 2304),
                f"{frames} frames of 4096x2304 u16 (12-bit sensor), chunks 64x256x256, shards 1x8x9 "
                "chunks, 1 scale, zstd level 1, discard sink")
    if name == "lod":
        cfg = ca.Config([Dim(frames, 16, 1, name="t"), Dim(2048, 256, 8, name="y", downsample=True),
                         Dim(2048, 256, 8, name="x", downsample=True)], dtype="u16", codec="none",
                        reduce="mean", buffer_capacity_bytes=64 << 20)
        return cfg, (frames, 2048, 2048), (f"{frames} frames of 2048x2048 u16, mean 2x2 LOD over y,x "
                                          "(4 levels), chunks 16x256x256, no codec, discard sink")
    if name == "volume":
        return (W.volume_lz4(), (256, 256, 256),
                "256^3 u16 xor+noise, chunks 64^3, shards 4^3 chunks, 1 scale, lz4, discard sink")
    if name == "multiscale6":
        return (W.multiscale6_zstd(2), (2, 256, 256, 256),
                "two 256^3 u16 timepoints (xor+noise, thirds), mean 2x2x2 LOD over z,y,x + t fold "
                "(6 levels), chunks 8^3, shards 4^3 chunks, zstd level 1, discard sink")
    return (W.multiscale_zstd(), (256, 256, 256),
            "256^3 u16 xor+noise, mean LOD over z,y,x with dim-0 downsampling (3 levels), chunks 64^3, "
            "shards 4^3 chunks, zstd level 1, discard sink")


def make_input(name, shape, need_host):
    """(device buffer, host array or None) of the workload's synthetic input."""
    import chucky_amd as ca
    from tests import workloads as W
    if name == "sensor":
        dev = W.sensor_frames_device(shape[0], shape[1], shape[2])
        return dev, (dev.to_host().view(np.uint16) if need_host else None)
    if name == "lod":
        n = int(np.prod(shape))
        dev = ca.DeviceBuffer(n * 2)
        if ca.lib.chucky_synth_fill(dev.ptr, n, ca.DTYPE_ID["u16"], 20260403, 4096):
            raise RuntimeError("fill failed: " + ca.last_error())
        return dev, (dev.to_host().view(np.uint16) if need_host else None)
    if name == "multiscale6":
        host = W.multiscale_volumes()
    else:
        from tests.helpers import xor_volume
        host = xor_volume(shape)
    return ca.DeviceBuffer.from_host(host), host


def run(cfg, desc, data, host, steps, warmup, pipelines, host_input, options=None):
    """Times `steps` steps of each pipeline on the same input. The pipelines'
    steps alternate (in alternating order), each step bracketed by device
    synchronizations, so that both see the same state of the machine: the
    host link's rate on these machines drops for a while after seconds of
    full load, which would otherwise favour whichever pipeline runs first."""
    import chucky_amd as ca
    nbytes = data.nbytes
    streams = {}
    # the batched pipeline's streams are created first: HIP hands its hardware
    # queues to streams in creation order, and the batched pipeline's
    # kernel-launching streams are laid out to take distinct ones
    for p in sorted(pipelines, key=lambda q: q != "batched"):
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
                      step_ms=[round(v, 3) for v in step_ms[p]])
        s.close()
    return res


LOD_PASSES = {"lod_gather": "tiles_k (linear epoch -> level-0 Morton)",
              "lod_reduce": "pyramid_k (every level from level 0, one launch)",
              "lod_append_fold": "fold_emit_k (append-dim fold and emit, all levels, one launch)",
              "lod_morton_chunk": "tiles_levels_k (Morton levels -> chunk pool, levels of one "
                                  "chunk geometry in one launch)"}
STAGES = ("memcpy", "h2d", "scatter", "compress", "aggregate", "d2h", "sink")


def snapshot(s):
    m = s.metrics()
    return {n: (getattr(m, n).ms, getattr(m, n).count, getattr(m, n).input_bytes,
                getattr(m, n).output_bytes)
            for n in (*STAGES, *LOD_PASSES)}


def pass_rate(metric):
    """(avg ms per launch, algorithmic bytes read+written per launch, GB/s)."""
    ms, n, rd, wr = metric
    if not n or ms <= 0:
        return 0.0, 0.0, 0.0
    avg = ms / n
    per = (rd + wr) / n
    return avg, per, per / (avg / 1e3) / 1e9


def summarize(res, steps):
    value = res["bytes"] * steps / res["seconds"] / 1e9
    m = res["metrics"]
    lat = sorted(res["latencies"])
    pct = (lambda q: round(lat[min(len(lat) - 1, int(q * len(lat)))] * 1e3, 3)) if lat else (lambda q: None)
    stages = {k: {"ms": round(v[0] / steps, 3), "launches": v[1] // max(steps, 1),
                  "GBps": round(v[2] / v[0] / 1e6, 2) if v[0] > 0 else None}
              for k, v in m.items() if v[1] and k not in LOD_PASSES}
    return value, dict(GBps=round(value, 3), GiBps=round(value * 1e9 / 2**30, 3),
                       ms_per_step=round(res["ms_per_step"], 3),
                       ratio=round(res["bytes"] / max(res["out_bytes"], 1), 4),
                       p50_epoch_latency_ms=pct(0.50), p99_epoch_latency_ms=pct(0.99),
                       step_ms=res.get("step_ms"), stages_per_step=stages)


def compress_roofline(res, name):
    m = res["metrics"]
    comp_ms, comp_n, comp_in, _ = m["compress"]
    if not comp_n or comp_ms <= 0:
        return None
    avg_ms = comp_ms / comp_n
    per = comp_in / comp_n
    achieved = per / (avg_ms / 1e3) / 1e9
    kname = "zstd_k" if "zstd" in res["desc"] else "lz4_chunk_k"
    return {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
            "traffic": measured_traffic(name, kname), "avg_launch_ms": round(avg_ms, 4),
            "bytes_per_launch": per,
            "note": "compressor input bytes per launch (one epoch's chunks, every level) / "
                    "event-timed launch duration on the compression queue"}


def lod_rates(res):
    out = {}
    for k, label in LOD_PASSES.items():
        avg, per, gbps = pass_rate(res["metrics"][k])
        if per:
            out[k] = {"kernels": label, "avg_launch_ms": round(avg, 4), "bytes_per_launch": per,
                      "achieved_GBps": round(gbps, 1), "frac_of_hbm_peak": round(gbps / HBM_PEAK_GBPS, 4)}
    return out


def transpose_rate(res):
    sc_ms, sc_n, sc_in, _ = res["metrics"]["scatter"]
    if not sc_n or sc_ms <= 0:
        return None
    avg = sc_ms / sc_n
    gbps = 2 * (sc_in / sc_n) / (avg / 1e3) / 1e9
    return {"kernel": "transpose_vec_k", "achieved_GBps": round(gbps, 1), "avg_launch_ms": round(avg, 4),
            "frac_of_hbm_peak": round(gbps / HBM_PEAK_GBPS, 4)}


def bit_exact(cfg, data, host):
    """Both pipelines' shards against the reference CPU path's on the bench
    input (outside the timed region): same keys and filled index slots, every
    chunk decoding to the same bytes."""
    import chucky_amd as ca
    from oracle import ref
    from tests.helpers import assert_same_shards, geometry, run_ref
    if not ref.available():
        return None
    g, _, cb = geometry(cfg)
    want = run_ref(cfg, host)[0]
    out = {"vs": "reference CPU path (oracle/_ref) on the same input"}
    for name, p in (("plain", ca.PIPELINE_PLAIN), ("batched", ca.PIPELINE_BATCHED)):
        sink = ca.MemSink()
        s = ca.GpuStream(cfg, sink, pipeline=p)
        assert s.append(host) == 0 and s.flush() == 0, ca.last_error()
        s.close()
        got = sink.shards()
        try:
            assert_same_shards(got, want, lambda lv: g[lv][0], cfg.codec, cb, name)
            out[name] = True
        except AssertionError as e:
            out[name] = False
            out[name + "_error"] = str(e)[:200]
        out["shards"] = len(want)
    return out


def cpu_baseline(cfg, host, sample_desc, repeat=1):
    """The reference's CPU path (oracle/_ref) on a bounded sample of the same
    input, all host threads, discard sink."""
    from oracle import ref
    if not ref.available():
        return None
    import chucky_amd as ca
    n = host.nbytes
    dt = 0.0
    out = 0
    for _ in range(repeat):
        sink = ca.DiscardSink()
        r = ref.RefCpuStream(cfg, sink)
        t0 = time.perf_counter()
        assert r.append(host) == 0
        assert r.flush() == 0
        dt += time.perf_counter() - t0
        r.close()
        out = sink.bytes
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    return {"value": round(n * repeat / dt / 1e9, 4), "unit": "GB/s", "cores": cores, "kind": "reference",
            "sample": f"{sample_desc} ({n / 1e9:.3f} GB) x{repeat}, same config; the reference CPU path "
                      "built from its sources (libzstd 1.4.9 / liblz4 1.9.3), discard sink",
            "ratio": round(n / max(out, 1), 4)}


def leg(name, steps, warmup, host_input, frames=512, options=None):
    """One workload through both pipelines: {plain, batched, ratio, ...}."""
    cfg, shape, desc = workload(name, frames)
    data, host = make_input(name, shape, host_input)
    runs = run(cfg, desc, data, host, steps, warmup, ["plain", "batched"], host_input, options)
    out = {"workload": desc, "input_GB": round(data.nbytes / 1e9, 4),
           "input_from": "host memory" if host_input else "HBM",
           "steps": steps, "warmup": warmup}
    for p in ("plain", "batched"):
        out[p] = summarize(runs[p], steps)[1]
    out["batched_over_plain"] = round(runs["plain"]["seconds"] / runs["batched"]["seconds"], 3)
    rl = compress_roofline(runs["batched"], name)
    if rl:
        out["compressor"] = {k: rl[k] for k in ("kernel", "achieved", "frac", "avg_launch_ms")}
    lod = lod_rates(runs["batched"])
    if lod:
        out["lod"] = lod
    t = transpose_rate(runs["batched"])
    if t:
        out["transpose"] = t
    data.free()
    return out


def mount_of(path):
    """(mount point, filesystem type) holding `path`, from /proc/mounts."""
    best = ("", "?")
    try:
        real = os.path.realpath(path)
        for line in open("/proc/mounts"):
            f = line.split()
            if len(f) > 2 and (real == f[1] or real.startswith(f[1].rstrip("/") + "/")) and len(f[1]) >= len(best[0]):
                best = (f[1], f[2])
    except OSError:
        pass
    return best


def fs_sink_leg(steps, warmup):
    """The headline workload (host input, batched pipeline) written by the
    zarr filesystem sink, each step into a fresh store as an acquisition
    would be (the sink is created before the timer; the step is reset +
    append + flush, and the flush waits for every shard write and the
    metadata): through the page cache, and as the reference's benchmark
    writes to disk (bench/bench_util.c:551-587: O_DIRECT shards, shard
    alignment of a page, which keeps the stream on the plain pipeline). The
    stores live in a temporary directory, removed afterwards."""
    import shutil
    import tempfile
    import chucky_amd as ca
    from chucky_amd.api import FsSink
    _, shape, desc = workload("multiscale")
    data, host = make_input("multiscale", shape, True)
    out = {"workload": desc.replace("discard sink", "zarr filesystem store (OME-NGFF group)"),
           "input_GB": round(data.nbytes / 1e9, 4), "input_from": "host memory",
           "steps": steps, "warmup": warmup}
    root = tempfile.mkdtemp(prefix="chucky_fs_")
    mnt, fstype = mount_of(root)
    out["store_filesystem"] = {"mount": mnt, "type": fstype}
    page = os.sysconf("SC_PAGE_SIZE")
    try:
        for mode, unbuffered in (("page_cache", False), ("o_direct", True)):
            c = workload("multiscale")[0]
            if unbuffered:
                c.shard_alignment = page
            prev = FsSink(os.path.join(root, mode, "0"), c, name="vol", unbuffered=unbuffered)
            s = ca.GpuStream(c, prev, pipeline=1, pass_timing=2)
            ms, written, err = [], 0, None
            for i in range(warmup + steps):
                store = os.path.join(root, mode, str(i))
                sink = prev if i == 0 else FsSink(store, c, name="vol", unbuffered=unbuffered)
                ca.lib.chucky_device_synchronize()
                t0 = time.perf_counter()
                if s.reset(sink) or s.append(host) or s.flush():
                    err = ca.last_error()
                ca.lib.chucky_device_synchronize()
                dt = (time.perf_counter() - t0) * 1e3
                if prev is not sink:
                    prev.close()  # the reset has let go of it
                prev = sink
                if err:
                    break
                if i >= warmup:
                    ms.append(dt)
                    written = sum(os.path.getsize(os.path.join(dp, f)) for dp, _, fs in os.walk(store) for f in fs)
            s.close()
            prev.close()
            if err or not ms:
                out[mode] = {"error": (err or "no steps")[:200]}
                continue
            mean = sorted(ms)[len(ms) // 2] if len(ms) > 4 else sum(ms) / len(ms)
            out[mode] = {"GBps": round(data.nbytes / (mean * 1e-3) / 1e9, 3), "ms_per_step": round(mean, 3),
                         "store_bytes": written, "shard_alignment": c.shard_alignment,
                         "pipeline": "plain (page-aligned shards)" if c.shard_alignment else "batched",
                         "step_ms": [round(v, 3) for v in ms]}
    finally:
        shutil.rmtree(root, ignore_errors=True)
        data.free()
    return out


SERVING_DESC = ("serving: 24 Poisson arrivals (4/s) over u8 2048x512x512, u16 256x2048x2048, f32 256^3; "
                "64^3 or 64x256x256 chunks; zstd/lz4; 1-5 LOD levels; 200 GB HBM budget; discard sinks")


def serving(args):
    """The serving mix: 24 requests arriving as a seeded Poisson process. The
    serving loop (batched pipeline, concurrent requests admitted against the
    HBM budget) and the plain pipeline run one request at a time see the same
    arrival times; both report aggregate GB/s over the makespan (first
    arrival to last shard written), per-epoch append->shard-written latency
    and per-request latency (arrival -> last shard written, queueing
    included)."""
    import chucky_amd as ca
    from chucky_amd.api import Server
    from tests import workloads as W
    reqs = W.serving_requests()
    inputs = {}
    for k, (dt, shape) in enumerate(W.SERVING_SHAPES):
        n = int(np.prod(shape))
        buf = ca.DeviceBuffer(n * np.dtype(ca.NP_DTYPE[dt]).itemsize)
        if ca.lib.chucky_synth_fill(buf.ptr, n, ca.DTYPE_ID[dt], 1000 + k, 0):
            raise RuntimeError("fill failed: " + ca.last_error())
        inputs[dt] = buf
    total = sum(inputs[r[1]].nbytes for r in reqs)
    pct = lambda v, q: round(sorted(v)[min(len(v) - 1, int(q * len(v)))] * 1e3, 3) if v else None  # noqa: E731

    # plain pipeline, one request at a time, each started at its arrival or
    # when the previous one ends
    serial_lat, serial_req = [], []
    t0 = time.perf_counter()
    for arrival, dt, shape, chunk, codec, nlod in reqs:
        wait = arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        sink = ca.DiscardSink()
        s = ca.GpuStream(W.serving_config(dt, shape, chunk, codec, nlod), sink, pipeline=ca.PIPELINE_PLAIN)
        eb = s.epoch_elements * np.dtype(ca.NP_DTYPE[dt]).itemsize
        buf = inputs[dt]
        for off in range(0, buf.nbytes, eb):
            assert s.append_device(buf.ptr + off, min(eb, buf.nbytes - off)) == 0, ca.last_error()
        assert s.flush() == 0, ca.last_error()
        serial_lat.extend(s.latencies())
        s.close()
        serial_req.append(time.perf_counter() - t0 - arrival)
    serial_s = time.perf_counter() - t0

    srv = Server(hbm_budget_bytes=200 * 10**9, max_concurrent=24, pipeline=ca.PIPELINE_BATCHED)
    sinks = []
    t0 = time.perf_counter()
    for arrival, dt, shape, chunk, codec, nlod in reqs:
        wait = arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        sinks.append(ca.DiscardSink())
        srv.submit(W.serving_config(dt, shape, chunk, codec, nlod), sinks[-1], inputs[dt], device=True)
    assert srv.wait() == 0, ca.last_error()
    reps = srv.reports()
    lat = srv.latencies()
    running, peak_bytes = srv.peaks()
    first = min(r.submit_s for r in reps)
    makespan = max(r.end_s for r in reps) - first
    req_lat = [r.end_s - r.submit_s for r in reps]
    srv.close()
    for b in inputs.values():
        b.free()
    return {
        "GBps": round(total / makespan / 1e9, 3), "makespan_s": round(makespan, 3),
        "input_GB": round(total / 1e9, 3),
        "p50_epoch_latency_ms": pct(lat, 0.50), "p99_epoch_latency_ms": pct(lat, 0.99),
        "p50_request_latency_ms": pct(req_lat, 0.50), "p99_request_latency_ms": pct(req_lat, 0.99),
        "peak_concurrent": running, "peak_admitted_GB": round(peak_bytes / 1e9, 3),
        "serial_plain": {"GBps": round(total / serial_s / 1e9, 3), "makespan_s": round(serial_s, 3),
                         "p50_epoch_latency_ms": pct(serial_lat, 0.50),
                         "p99_epoch_latency_ms": pct(serial_lat, 0.99),
                         "p50_request_latency_ms": pct(serial_req, 0.50),
                         "p99_request_latency_ms": pct(serial_req, 0.99)},
        "served_over_serial": round(serial_s / makespan, 3),
        "arrival_span_s": round(reqs[-1][0], 3),
        "workload": SERVING_DESC,
    }


def serving_line(args):
    r = serving(args)
    return {
        "metric": METRIC, "value": r["GBps"], "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(r["makespan_s"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/u16/f32", "data": "synthetic, generated in HBM",
        "config": {"workload": SERVING_DESC, "input_GB": r["input_GB"]},
        "p50_epoch_latency_ms": r["p50_epoch_latency_ms"], "p99_epoch_latency_ms": r["p99_epoch_latency_ms"],
        "serving": r,
    }


PMC_FILES = {"sensor": "zstd_k_sensor512x4096x2304_pmc.md", "lod": "tiles_k_lod64x2048x2048_pmc.md",
             "multiscale": "zstd_k_multiscale256_pmc.md"}


def measured_traffic(workload_name, kernel):
    """HBM bytes per launch of `kernel` on the workload from the committed
    counter summary (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE,
    scaled by a copy of known size), or None when none covers the pair."""
    f = PMC_FILES.get(workload_name)
    if not f:
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", f)):
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=512, help="frames of the sensor / lod workloads")
    ap.add_argument("--workload", choices=["multiscale", "multiscale6", "sensor", "volume", "lod", "serving", "fs"],
                    default="multiscale")
    ap.add_argument("--pipeline", choices=["plain", "batched", "both"], default="both",
                    help="time one pipeline alone (profiling runs) or both")
    ap.add_argument("--device-input", action="store_true",
                    help="time input already resident in HBM instead of host-memory appends")
    ap.add_argument("--no-extras", action="store_true", help="skip the extra workloads beside the headline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--zstd-block", type=int, default=0, help="zstd block bytes (0 = autotuned)")
    ap.add_argument("--zstd-chunk-tables", type=int, default=0,
                    help="zstd Huffman tables: 0 = autotuned, 1 = one per chunk, 2 = one per block")
    ap.add_argument("--batched-launch", type=int, default=0,
                    help="batched pipeline per-epoch work: 0 = autotuned, 1 = captured graphs, 2 = direct")
    ap.add_argument("--pass-timing", type=int, default=2,
                    help="event timing of each scatter / LOD launch in the timed steps: 1 = on, 2 = off "
                         "(the default: the per-pass rates then come from a separate HBM-input leg)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one plain device copy of BYTES (the vector scatter kernel over an "
                         "identity layout), the reference for counter-based HBM traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; runs on several GPUs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        print(json.dumps(calibrate_copy(args.calibrate_copy)))
        return
    if args.workload == "serving":
        print(json.dumps(serving_line(args)))
        return
    if args.workload == "fs":
        print(json.dumps(fs_sink_leg(args.steps, args.warmup)))
        return

    import chucky_amd as ca
    host_input = not args.device_input
    options = {"zstd_block_bytes": args.zstd_block, "batched_launch": args.batched_launch,
               "zstd_chunk_tables": args.zstd_chunk_tables, "pass_timing": args.pass_timing}
    cfg, shape, desc = workload(args.workload, args.frames)
    data, host = make_input(args.workload, shape, True)
    pipes = ["plain", "batched"] if args.pipeline == "both" else [args.pipeline]
    runs = run(cfg, desc, data, host, args.steps, args.warmup, pipes, host_input, options)
    # the host link's peak, measured after the timed steps, so that no other
    # pinned allocation precedes the streams' own
    h2d_peak = ca.lib.chucky_measure_link_gbps(256 << 20, 8, 1)
    d2h_peak = ca.lib.chucky_measure_link_gbps(256 << 20, 8, 0)
    lead = "batched" if "batched" in runs else "plain"
    value, head = summarize(runs[lead], args.steps)
    roofline = compress_roofline(runs[lead], args.workload)
    if roofline is None and lod_rates(runs[lead]).get("lod_gather"):
        g = lod_rates(runs[lead])["lod_gather"]
        roofline = {"bound": "hbm", "kernel": "tiles_k (LOD gather)", "achieved": g["achieved_GBps"],
                    "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": g["frac_of_hbm_peak"],
                    "traffic": measured_traffic(args.workload, "tiles_k"),
                    "avg_launch_ms": g["avg_launch_ms"], "bytes_per_launch": g["bytes_per_launch"],
                    "note": "epoch bytes read + Morton bytes written per launch / event-timed launch duration"}

    exact = (bit_exact(cfg, data, host) if args.workload in ("multiscale", "multiscale6", "volume")
             and args.pipeline == "both" else None)
    cpu = None
    if not args.no_cpu_baseline:
        if args.workload == "sensor":
            ncpu = 128
            scfg = workload("sensor", ncpu)[0]
            cpu = cpu_baseline(scfg, host[:ncpu * 4096 * 2304], f"first {ncpu} of {args.frames} frames")
        elif args.workload == "lod":
            cpu = cpu_baseline(cfg, host, "whole input")
        else:
            cpu = cpu_baseline(cfg, host, "whole input", repeat=10)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic, " + ("appended from host memory (pinned staging + H2D inside the timed region)"
                                 if host_input else "resident in HBM before the timed region"),
        "config": {"workload": desc, "pipeline": lead, "input_GB": round(data.nbytes / 1e9, 4),
                   "input_from": "host memory" if host_input else "HBM",
                   "per_pass_event_timing": args.pass_timing != 2},
        "GiB_per_s": head["GiBps"],
        "compression_ratio": head["ratio"],
        "p50_epoch_latency_ms": head["p50_epoch_latency_ms"],
        "p99_epoch_latency_ms": head["p99_epoch_latency_ms"],
        "batched_over_plain": (round(runs["plain"]["seconds"] / runs["batched"]["seconds"], 3)
                               if len(runs) == 2 else None),
        "pipelines": {p: summarize(runs[p], args.steps)[1] for p in runs},
        "bit_exact": exact,
        "link": {"h2d_peak_GBps": round(h2d_peak, 2), "d2h_peak_GBps": round(d2h_peak, 2),
                 "value_frac_of_h2d_peak": round(value / h2d_peak, 4) if h2d_peak > 0 else None},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    t = transpose_rate(runs[lead])
    if t:
        line["transpose"] = t
    lod = lod_rates(runs[lead])
    if lod:
        line["lod"] = lod
    data.free()
    if not args.no_extras and args.workload == "multiscale" and args.pipeline == "both":
        extras = {}
        # input resident in HBM, each scatter / LOD launch timed by events:
        # the per-pass HBM rates of the headline workload
        extras["multiscale_hbm_input"] = leg("multiscale", args.steps, args.warmup, False,
                                             options={"pass_timing": 1})
        if "lod" in extras["multiscale_hbm_input"]:
            line["lod"] = dict(extras["multiscale_hbm_input"]["lod"],
                               note="from the HBM-input leg (extras.multiscale_hbm_input), every pass timed")
        extras["multiscale6_host_input"] = leg("multiscale6", max(3, args.steps // 4), 1, True)
        extras["volume_lz4_host_input"] = leg("volume", max(3, args.steps // 4), 1, True)
        extras["sensor_host_input"] = leg("sensor", 2, 1, True, frames=512)
        try:
            extras["fs_sink_host_input"] = fs_sink_leg(max(3, args.steps // 2), 2)
            extras["fs_sink_host_input"]["discard_sink_GBps"] = line["value"]
        except Exception as e:  # the headline stands on its own
            extras["fs_sink_host_input"] = {"error": str(e)[:300]}
        try:
            extras["serving"] = serving(args)
        except Exception as e:  # the headline stands on its own
            extras["serving"] = {"error": str(e)[:300]}
        line["extras"] = extras
    print(json.dumps(line))


if __name__ == "__main__":
    main()
