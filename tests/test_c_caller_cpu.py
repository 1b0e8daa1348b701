"""The reference's caller pattern as a C11 program (tests/native/caller.c):
the headers in include/ compile as C with warnings as errors, the program
links against libchucky_amd.so, and the host-only entry points
(tile_stream_gpu_memory_estimate, tile_stream_gpu_advise_chunk_sizes) answer
from C the same as through the Python mirror of the structures."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

import chucky_amd as ca
from chucky_amd.api import Dim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "chucky_amd")


def build_caller(out_dir: str) -> str:
    cc = shutil.which("cc") or shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = os.path.join(out_dir, "caller")
    subprocess.run([cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "caller.c"), "-o", exe, "-L", LIBDIR, "-lchucky_amd",
                    f"-Wl,-rpath,{LIBDIR}"], check=True)
    return exe


def test_c_caller_compiles_links_and_estimates(tmp_path):
    exe = build_caller(str(tmp_path))
    out = subprocess.run([exe, "estimate"], check=True, capture_output=True, text=True).stdout.split("\n")
    est = dict(zip(out[0].split()[0::2], map(int, out[0].split()[1::2])))
    # the same configuration through the Python mirror of the structures
    cfg = ca.Config([Dim(10, 4, 1, name="t"), Dim(64, 32, 2, name="y"), Dim(96, 32, 3, name="x")],
                    dtype="u16", codec="zstd", level=1, buffer_capacity_bytes=64 << 10)
    info = ca.memory_estimate(cfg)
    assert est["chunks_per_epoch"] == info.chunks_per_epoch == 6
    assert est["total_chunks"] == info.total_chunks
    assert est["nlod"] == info.nlod == 1
    assert est["epochs_per_batch"] == info.epochs_per_batch
    assert est["device_bytes"] == info.device_bytes
    adv = out[1].split()
    chunks = [int(v) for v in adv[2:5]]
    assert chunks[0] == 1  # ratio 0: the append dim gets no bits
    assert all(c & (c - 1) == 0 for c in chunks)
    assert int(adv[6]) <= int(adv[8])  # the advised configuration fits the budget
