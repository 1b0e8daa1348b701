"""tests/native/caller.c on the GPU: the reference's caller pattern in C --
dims_create, zarr_fs_sink_create, tile_stream_gpu_create, writer_append from
host memory in three uneven pieces, writer_flush, destroy -- writes a store
that the offline reader decodes back to the appended frames."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from tests.test_c_caller_cpu import build_caller
from tests.zarr_readback import read_array

pytestmark = pytest.mark.gpu


def expected():
    t, y, x = np.meshgrid(np.arange(10), np.arange(64), np.arange(96), indexing="ij")
    return ((t * 7919 + y * 131 + x * 17) & 0xFFFF).astype(np.uint16)


@pytest.mark.parametrize("codec", ["zstd", "lz4", "none"])
def test_c_caller_writes_readable_store(tmp_path, codec):
    exe = build_caller(str(tmp_path))
    store = tmp_path / "acq.zarr"
    r = subprocess.run([exe, "run", str(store), codec], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    arr, meta, stats = read_array(str(store / "0"))
    assert arr.shape == (10, 64, 96) and stats["codec"] == codec
    assert np.array_equal(arr, expected())
