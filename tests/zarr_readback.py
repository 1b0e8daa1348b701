"""Offline readback of the Zarr v3 stores the file sink writes: the project's
replacement for the reference's tests/validate_zarr.py (which needs
zarr-python, absent here). Test infrastructure only.

It reads what a Zarr v3 reader would: the array's zarr.json (shape, data
type, regular chunk grid = shard shape, default "/" key encoding, one
sharding_indexed codec with bytes + optional lz4 / zstd inner codecs and a
bytes + crc32c index at the end), every shard file c/<i>/<j>/..., checks each
shard index's CRC-32C, decodes every chunk with the stock decoders and
assembles the array (fill value where a chunk or shard is absent). The
multiscale form follows the OME-NGFF group's dataset paths.

    python tests/zarr_readback.py STORE [ARRAY]   # prints shape, dtype, counts
"""
from __future__ import annotations

import json
import os
import struct
import sys

import numpy as np

if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle.chucky_oracle import crc32c  # noqa: E402
from tests.helpers import decode  # noqa: E402

DTYPES = {"uint8": np.uint8, "uint16": np.uint16, "uint32": np.uint32, "uint64": np.uint64,
          "int8": np.int8, "int16": np.int16, "int32": np.int32, "int64": np.int64,
          "float16": np.float16, "float32": np.float32, "float64": np.float64}
BYTES_LE = {"name": "bytes", "configuration": {"endian": "little"}}
EMPTY = 2 ** 64 - 1


def _ceildiv(a, b):
    return -(-a // b)


def read_array(path):
    """(array, metadata, stats) of the array at `path`."""
    with open(os.path.join(path, "zarr.json")) as f:
        meta = json.load(f)
    assert meta["zarr_format"] == 3 and meta["node_type"] == "array", "not a zarr v3 array"
    shape = [int(s) for s in meta["shape"]]
    dt = np.dtype(DTYPES[meta["data_type"]]).newbyteorder("<")
    assert meta["chunk_grid"]["name"] == "regular"
    shard_shape = [int(s) for s in meta["chunk_grid"]["configuration"]["chunk_shape"]]
    assert meta["chunk_key_encoding"] == {"name": "default", "configuration": {"separator": "/"}}
    assert len(meta["codecs"]) == 1 and meta["codecs"][0]["name"] == "sharding_indexed"
    cfg = meta["codecs"][0]["configuration"]
    chunk_shape = [int(s) for s in cfg["chunk_shape"]]
    inner = cfg["codecs"]
    assert inner[0] == BYTES_LE and len(inner) <= 2
    codec = inner[1]["name"] if len(inner) == 2 else "none"
    assert codec in ("none", "lz4", "zstd"), codec
    assert cfg["index_codecs"] == [BYTES_LE, {"name": "crc32c", "configuration": {}}]
    assert cfg["index_location"] == "end"
    assert all(s % c == 0 for s, c in zip(shard_shape, chunk_shape)), "shard not a multiple of the chunk"
    per = [s // c for s, c in zip(shard_shape, chunk_shape)]
    n_inner = int(np.prod(per))
    index_bytes = 16 * n_inner
    chunk_bytes = int(np.prod(chunk_shape)) * dt.itemsize
    out = np.full(shape, meta["fill_value"], dtype=dt)
    stats = {"shards": 0, "chunks": 0, "empty_slots": 0, "codec": codec}
    grid = [_ceildiv(s, ss) if s else 0 for s, ss in zip(shape, shard_shape)]
    for sidx in np.ndindex(*grid):
        p = os.path.join(path, "c", *map(str, sidx))
        if not os.path.exists(p):
            continue
        with open(p, "rb") as f:
            blob = f.read()
        stats["shards"] += 1
        assert len(blob) >= index_bytes + 4, f"{p}: shorter than its index"
        idx = blob[-(index_bytes + 4):-4]
        assert crc32c(idx) == struct.unpack("<I", blob[-4:])[0], f"{p}: index CRC mismatch"
        data_end = len(blob) - index_bytes - 4
        for j, cidx in enumerate(np.ndindex(*per)):
            off, nb = struct.unpack_from("<QQ", idx, 16 * j)
            if off == EMPTY and nb == EMPTY:
                stats["empty_slots"] += 1
                continue
            assert off + nb <= data_end, f"{p}: chunk {j} runs into the index"
            raw = decode(codec, blob[off:off + nb], chunk_bytes)
            assert len(raw) == chunk_bytes, f"{p}: chunk {j} decodes to {len(raw)} bytes"
            origin = [s * ss + c * cs for s, ss, c, cs in zip(sidx, shard_shape, cidx, chunk_shape)]
            if any(o >= S for o, S in zip(origin, shape)):
                continue  # a chunk wholly past the array edge carries no array cells
            region = tuple(slice(o, min(o + cs, S)) for o, cs, S in zip(origin, chunk_shape, shape))
            arr = np.frombuffer(raw, dtype=dt).reshape(chunk_shape)
            out[region] = arr[tuple(slice(0, r.stop - r.start) for r in region)]
            stats["chunks"] += 1
    return out, meta, stats


def read_multiscale(path):
    """[level arrays] of an OME-NGFF v0.5 multiscale group at `path`."""
    with open(os.path.join(path, "zarr.json")) as f:
        meta = json.load(f)
    assert meta["zarr_format"] == 3 and meta["node_type"] == "group"
    ms = meta["attributes"]["ome"]["multiscales"]
    assert meta["attributes"]["ome"]["version"] == "0.5" and len(ms) == 1
    return [read_array(os.path.join(path, ds["path"]))[0] for ds in ms[0]["datasets"]]


def main():
    store = sys.argv[1]
    path = os.path.join(store, sys.argv[2]) if len(sys.argv) > 2 else store
    with open(os.path.join(path, "zarr.json")) as f:
        kind = json.load(f)["node_type"]
    if kind == "group":
        for lv, a in enumerate(read_multiscale(path)):
            print(f"level {lv}: shape {a.shape} dtype {a.dtype}")
        return
    a, _, st = read_array(path)
    print(f"shape {a.shape} dtype {a.dtype} shards {st['shards']} chunks {st['chunks']} "
          f"empty slots {st['empty_slots']} codec {st['codec']}")


if __name__ == "__main__":
    main()
