"""Codec launch-shape sweep behind the tables in csrc/autotune.cpp.

    python -m chucky_amd.tune --codec lz4 --out lz4.json
    python -m chucky_amd.tune --codec zstd --out zstd.json
    python -m chucky_amd.tune --table lz4.json zstd.json --csv chucky_amd/tune_sweep_mi355x.csv
    python -m chucky_amd.tune --table chucky_amd/tune_sweep_mi355x.csv

The sweep streams synthetic volumes through the plain pipeline (device-
resident input, discard sink) for every dtype size x chunk shape x chunks per
compress launch below, multiscale volumes for the LOD chunk shapes and
sensor frames at their own geometry, once per candidate setting, and records the compress
stage's event-timed milliseconds per launch and its compression ratio. --table
turns recorded sweeps into the C++ rows of the tables: for each measured case
(rates and ratios averaged over the data kinds) the candidate with the least
time per input byte through the compressor plus the time its compressed bytes
take over the device-to-host link (LINK_GBPS, the measured MI355X PCIe rate),
so a faster setting wins unless it costs more in output bytes than it saves.
"""
import argparse
import json
import sys
import time

import numpy as np

DTYPES = ("u8", "u16", "f32")
CHUNKS = ((32, 32, 32), (64, 64, 64), (16, 128, 128))
# (chunks per epoch as a x b grid, epochs per batch): chunks per compress launch = a*b*K
LAUNCHES = (((4, 4), 1), ((4, 4), 4), ((8, 8), 4))
# multiscale streams (mean 2x2x2 pyramid over every dim of a 256^3 u16 volume,
# all levels compressed in one launch): the coarse levels' smooth chunks parse
# into dense sequences, so their blocks set the launch's tail
LOD_CHUNKS = ((64, 64, 64), (32, 32, 32), (8, 8, 8))
LOD_EPOCHS = (1, 4)
# sensor frames: 4096x2304 u16, 64x256x256 chunks (8 MiB), one 1.2 GB epoch
# of 144 chunks a launch (or two), Poisson counts synthesized on the device
SENSOR_EPOCHS = (1, 2)
LZ4_CANDIDATES = [(seg, h) for seg in (1024, 2048, 4096, 8192, 16384, 1 << 30) for h in (9, 10, 11, 12)]
ZSTD_CANDIDATES = [(b, t) for b in (8192, 16384, 32768, 65536, 131072) for t in (2, 1)]
LINK_GBPS = 55.0


def volume(kind, shape, dtype, seed=7):
    """xor: ((x ^ y ^ z) * 257 + noise) like the volume benchmarks; sensor:
    Poisson(100) counts clipped to 12 bits. Cast to dtype (u8 keeps the low byte)."""
    rng = np.random.default_rng(seed)
    if kind == "xor":
        z, y, x = np.meshgrid(*[np.arange(s, dtype=np.uint32) for s in shape], indexing="ij")
        v = ((x ^ y ^ z) * 257 + rng.integers(0, 16, shape, dtype=np.uint32)) % 65536
    else:
        v = np.clip(rng.poisson(100, shape), 0, 4095).astype(np.uint32)
    if dtype == "u8":
        return (v & 0xFF).astype(np.uint8)
    if dtype == "u16":
        return v.astype(np.uint16)
    return v.astype(np.float32)


def run_candidates(codec, cfg, host, option_sets, case, steps=3, warmup=1, dev=None):
    """One row per option set: compress ms per launch, GB/s and ratio.
    `host` is the input, or (with `dev`, the input already on the device) an
    array of its dtype and size only."""
    import chucky_amd as ca
    info = ca.memory_estimate(cfg)
    per_launch = info.total_chunks * info.epochs_per_batch
    if dev is None:
        dev = ca.DeviceBuffer.from_host(host)
    rows = []
    for opts in option_sets:
        sink = ca.DiscardSink()
        s = ca.GpuStream(cfg, sink, pipeline=0, **opts)
        m0 = None
        for i in range(warmup + steps):
            if i == warmup:
                m = s.metrics().compress
                m0 = (m.ms, m.count, m.input_bytes, sink.bytes)
            assert s.reset(sink) == 0
            assert s.append_device(dev.ptr, host.nbytes) == 0, ca.last_error()
            assert s.flush() == 0, ca.last_error()
        m = s.metrics().compress
        # ratio: input bytes over the shard bytes delivered (chunks and indices)
        ms, n, rd, wr = (m.ms - m0[0], m.count - m0[1], m.input_bytes - m0[2], sink.bytes - m0[3])
        s.close()
        sink.close()
        rows.append(dict(case, codec=codec, bpe=host.itemsize, per_launch=int(per_launch), options=opts,
                         ms_per_launch=ms / max(n, 1), GBps=rd / (ms / 1e3) / 1e9 if ms > 0 else 0.0,
                         ratio=rd / max(wr, 1)))
        print(json.dumps(rows[-1]), flush=True)
    dev.free()
    return rows


def measure(codec, dtype, chunk, grid, K, option_sets, kind):
    from chucky_amd.api import Config, Dim
    a, b = grid
    cz, cy, cx = chunk
    epochs = 2 * K  # two compress launches per step
    shape = (cz * epochs, cy * a, cx * b)
    cfg = Config([Dim(shape[0], cz, 1), Dim(shape[1], cy, a), Dim(shape[2], cx, b)], dtype=dtype,
                 codec=codec, epochs_per_batch=K)
    case = dict(dtype=dtype, chunk=list(chunk), chunk_bytes=int(np.prod(chunk)) * np.dtype(
        {"u8": np.uint8, "u16": np.uint16, "f32": np.float32}[dtype]).itemsize, data=kind, lod=0)
    return run_candidates(codec, cfg, volume(kind, shape, dtype), option_sets, case)


def measure_lod(codec, chunk, K, option_sets, kind, n=256):
    from chucky_amd.api import Config, Dim
    cfg = Config([Dim(n, c, 4, downsample=True) for c in chunk], dtype="u16", codec=codec,
                 reduce="mean", append_reduce="mean", epochs_per_batch=K)
    case = dict(dtype="u16", chunk=list(chunk), chunk_bytes=int(np.prod(chunk)) * 2, data=kind, lod=1)
    return run_candidates(codec, cfg, volume(kind, (n, n, n), "u16"), option_sets, case)


def measure_sensor(codec, K, option_sets, height=4096, width=2304):
    import ctypes as C
    import chucky_amd as ca
    from chucky_amd.api import Config, Dim
    frames = 64 * 2 * K
    cfg = Config([Dim(frames, 64, 1), Dim(height, 256, 8), Dim(width, 256, 9)], dtype="u16", codec=codec,
                 epochs_per_batch=K, buffer_capacity_bytes=64 << 20)
    dev = ca.DeviceBuffer(frames * height * width * 2)
    if ca.lib.chucky_synth_sensor_frames(dev.ptr, frames, height, width, 7, 100.0, C.POINTER(C.c_float)(), 0):
        raise RuntimeError("frame synthesis failed: " + ca.last_error())
    shape = np.empty((frames, height, width), dtype=np.uint16)  # size and dtype only
    case = dict(dtype="u16", chunk=[64, 256, 256], chunk_bytes=64 * 256 * 256 * 2, data="sensor", lod=0)
    return run_candidates(codec, cfg, shape, option_sets, case, dev=dev)


def sweep(codec):
    kinds = ("xor", "sensor")
    cands = ([dict(lz4_segment_bytes=s, lz4_hash_log=h) for s, h in LZ4_CANDIDATES] if codec == "lz4" else
             [dict(zstd_block_bytes=b, zstd_chunk_tables=t) for b, t in ZSTD_CANDIDATES])
    out = []
    for dtype in DTYPES:
        for chunk in CHUNKS:
            for grid, K in LAUNCHES:
                for kind in kinds:
                    out += measure(codec, dtype, chunk, grid, K, cands, kind)
    for chunk in LOD_CHUNKS:
        for K in LOD_EPOCHS:
            for kind in kinds:
                out += measure_lod(codec, chunk, K, cands, kind)
    for K in SENSOR_EPOCHS:
        out += measure_sensor(codec, K, cands)
    return out


def choose(rows, codec):
    """{(lod, bpe, chunk_bytes, per_launch): (options, GB/s, ratio)} from recorded
    rows: mean rate and ratio over the data kinds, least cost per input byte."""
    cases = {}
    for r in rows:
        key = (r.get("lod", 0), r["bpe"], r["chunk_bytes"], r["per_launch"])
        opt = tuple(sorted(r["options"].items()))
        cases.setdefault(key, {}).setdefault(opt, []).append(r)
    best = {}
    for key, by_opt in sorted(cases.items()):
        agg = {o: (np.mean([r["GBps"] for r in rs]), np.mean([r["ratio"] for r in rs])) for o, rs in by_opt.items()}
        o = min(agg, key=lambda o: 1 / agg[o][0] + 1 / (agg[o][1] * LINK_GBPS))
        best[key] = (dict(o), agg[o][0], agg[o][1])
    return best


CSV_FIELDS = ("codec", "lod", "bpe", "chunk_bytes", "per_launch", "data", "opt_a", "opt_b", "GBps", "ratio")


def load(path):
    """Sweep rows from a recorded JSON-lines file or its CSV summary."""
    if path.endswith(".csv"):
        rows = []
        for line in open(path):
            if line.startswith("#") or line.startswith("codec,"):
                continue
            v = line.strip().split(",")
            r = dict(zip(CSV_FIELDS, v))
            for k in ("lod", "bpe", "chunk_bytes", "per_launch", "opt_a", "opt_b"):
                r[k] = int(r[k])
            r["GBps"], r["ratio"] = float(r["GBps"]), float(r["ratio"])
            keys = ("lz4_segment_bytes", "lz4_hash_log") if r["codec"] == "lz4" else \
                   ("zstd_block_bytes", "zstd_chunk_tables")
            r["options"] = {keys[0]: r.pop("opt_a"), keys[1]: r.pop("opt_b")}
            rows.append(r)
        return rows
    return [json.loads(line) for line in open(path) if line.startswith("{")]


def write_csv(rows, path):
    with open(path, "w") as f:
        f.write("# chucky_amd.tune sweep on MI355X: compress stage GB/s and ratio per run\n")
        f.write(",".join(CSV_FIELDS) + "\n")
        for r in rows:
            a, b = r["options"].values()
            f.write(",".join(str(x) for x in (r["codec"], r.get("lod", 0), r["bpe"], r["chunk_bytes"],
                                              r["per_launch"], r["data"], a, b)) +
                    f",{r['GBps']:.2f},{r['ratio']:.4f}\n")


def table(paths, csv_out=None):
    rows = []
    for p in paths:
        rows += load(p)
    if csv_out:
        write_csv(rows, csv_out)
    for codec in ("lz4", "zstd"):
        rs = [r for r in rows if r["codec"] == codec]
        if not rs:
            continue
        print(f"// {codec}: lod, bpe, chunk bytes, chunks per launch -> choice (GB/s, ratio)")
        for (lod, bpe, cb, pl), (o, gbps, ratio) in choose(rs, codec).items():
            if codec == "lz4":
                seg = o["lz4_segment_bytes"]
                seg = 0 if seg >= cb else seg
                print(f"  {{ {lod}, {bpe}, {cb}, {pl}, {seg}, {o['lz4_hash_log']} }}, "
                      f"// {gbps:.1f} GB/s, {ratio:.3f}")
            else:
                print(f"  {{ {lod}, {bpe}, {cb}, {pl}, {o['zstd_block_bytes']}, "
                      f"{int(o['zstd_chunk_tables'] == 1)} }}, // {gbps:.1f} GB/s, {ratio:.3f}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--codec", choices=["lz4", "zstd"])
    ap.add_argument("--out")
    ap.add_argument("--table", nargs="+", metavar="SWEEP", help="recorded sweeps (JSON lines or CSV)")
    ap.add_argument("--csv", help="with --table: also write the rows as one CSV summary")
    a = ap.parse_args(argv)
    if a.table:
        table(a.table, a.csv)
        return 0
    if not a.codec or not a.out:
        ap.error("--codec and --out, or --table")
    t0 = time.time()
    rows = sweep(a.codec)
    with open(a.out, "w") as f:
        for r in rows:
            f.write(json.dumps(r) + "\n")
    print(f"{len(rows)} runs in {time.time() - t0:.0f} s -> {a.out}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
