"""Writes the golden shard fixtures of tests/golden/*.npz.

Each fixture holds an input array, the stream specification and the
CODEC_NONE shard files that the reference's own CPU path (oracle/_ref, built
from the reference sources by oracle/Makefile) writes for it. Run from the
repository root in the container that has /root/reference:

    make -C oracle && python tests/golden/make_golden.py
"""
import json
import os
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import chucky_amd as ca  # noqa: E402
from chucky_amd.api import Dim  # noqa: E402
from tests.helpers import random_typed, run_ref, xor_volume  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

SPECS = {
    # tests/test_cross_validate.c:168-252 geometry
    "basic_4x4x6_u16": ([Dim(4, 2, 2), Dim(4, 2, 2), Dim(6, 3, 2)], "u16", "mean", "mean", "seq96"),
    # uneven shards along y, partial last epoch along z
    "uneven_20x16x24_u8": ([Dim(20, 8, 2), Dim(16, 8, 1), Dim(24, 8, 3)], "u8", "mean", "mean", "rand"),
    # LOD over y, x (3 levels), median
    "lod_yx_median_u16": ([Dim(8, 4, 1), Dim(24, 4, 0, downsample=True),
                           Dim(20, 4, 0, downsample=True)], "u16", "median", "mean", "rand"),
    # dim0 downsampling with the shift-by-level integer mean fold
    "lod_dim0_mean_u16": ([Dim(16, 2, 4, downsample=True), Dim(8, 4, 0, downsample=True),
                           Dim(12, 4, 0, downsample=True)], "u16", "mean", "mean", "xor"),
    # float dim0 max fold, max_sup in-epoch
    "lod_dim0_maxsup_f32": ([Dim(12, 2, 2, downsample=True), Dim(8, 4, 0, downsample=True),
                             Dim(8, 4, 0, downsample=True)], "f32", "max_sup", "max", "rand"),
}
RED = {"mean": 0, "min": 1, "max": 2, "median": 3, "max_sup": 4, "min_sup": 5}


def main():
    for name, (dims, dtype, red, ared, fill) in SPECS.items():
        n = int(np.prod([d.size for d in dims]))
        if fill == "seq96":
            n = 96
            data = (np.arange(n) % 65536).astype(np.uint16)
        elif fill == "xor":
            data = xor_volume([d.size for d in dims]).reshape(-1)
        else:
            data = random_typed(n, dtype, seed=zlib.crc32(name.encode()) & 0xFFFF)
        cfg = ca.Config(dims, dtype=dtype, reduce=red, append_reduce=ared,
                        buffer_capacity_bytes=4096)
        shards, _, _ = run_ref(cfg, data, pieces=2)
        keys = sorted(shards)
        spec = {"dims": [(d.size, d.chunk, d.cps, int(d.downsample),
                          -1 if d.storage_position is None else d.storage_position) for d in dims],
                "reduce": RED[red], "append": RED[ared], "dtype": dtype}
        arrays = {"input": data, "keys": np.array(keys, dtype=np.uint64),
                  "spec": np.frombuffer(json.dumps(spec).encode(), dtype=np.uint8)}
        for i, k in enumerate(keys):
            arrays[f"shard{i}"] = np.frombuffer(shards[k], dtype=np.uint8)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, len(keys), "shards", sum(len(v) for v in shards.values()), "bytes")


if __name__ == "__main__":
    main()
