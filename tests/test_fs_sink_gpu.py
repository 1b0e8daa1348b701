"""The HIP stream writing into the product's Zarr v3 file sink, against the
reference's CPU stream writing into the reference's own file sink (both built
from sources): the same file tree, zarr.json documents byte-identical, shards
byte-identical with no codec and otherwise equal after decode; level 0 reads
back (tests/zarr_readback.py) as the input. Buffered and O_DIRECT stores."""
import errno
import os

import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Dim
from oracle import ref
from tests.helpers import assert_same_shards, geometry, random_typed, xor_volume
from tests.test_fs_sink_cpu import _Sink, _tree
from tests.zarr_readback import read_array

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]

CASES = {
    # several inner shards, ragged dim 0, zstd
    "single_zstd": (False, [Dim(40, 8, 2), Dim(64, 16, 2), Dim(96, 32, 1)], "u16", "zstd", 0),
    # unbounded dim 0, lz4
    "unbounded_lz4": (False, [Dim(0, 4, 2), Dim(32, 16, 1), Dim(48, 16, 3)], "u8", "lz4", 0),
    # no codec: every file byte-identical
    "single_none": (False, [Dim(24, 4, 3), Dim(40, 8, 0), Dim(36, 12, 0)], "i16", "none", 0),
    # LOD over y, x with the append dim downsampled: four levels, each level's
    # zero-copy writes pending in the sink's I/O queue while the next level's
    # bytes arrive
    "multiscale_zstd": (True, [Dim(32, 4, 2, downsample=True), Dim(64, 16, 0, downsample=True),
                                Dim(64, 16, 0, downsample=True)], "u16", "zstd", 0),
    "multiscale_none": (True, [Dim(24, 2, 1), Dim(48, 8, 0, downsample=True),
                               Dim(40, 8, 0, downsample=True)], "f32", "none", 0),
    # page-aligned shards for O_DIRECT
    "aligned_zstd": (False, [Dim(32, 8, 2), Dim(64, 32, 1), Dim(64, 32, 2)], "u16", "zstd", 4096),
}


def _odirect_dir(tmp_path):
    """A directory whose file system takes O_DIRECT (tmpfs does not), or None."""
    for base in (str(tmp_path), os.path.join(os.path.dirname(os.path.dirname(__file__)), "gpurun_out")):
        try:
            os.makedirs(base, exist_ok=True)
            probe = os.path.join(base, ".odirect_probe")
            fd = os.open(probe, os.O_WRONLY | os.O_CREAT | os.O_DIRECT, 0o644)
            os.close(fd)
            os.unlink(probe)
            return base
        except OSError as e:
            if e.errno not in (errno.EINVAL, errno.EOPNOTSUPP, errno.ENOENT, errno.EACCES):
                raise
    return None


def _data(dims, dtype):
    n = int(np.prod([d.size or 40 for d in dims]))
    if dtype == "u16" and len(dims) == 3:
        shape = [d.size or 40 for d in dims]
        return xor_volume(shape).reshape(-1)
    return random_typed(n, dtype, seed=11)


def _write(lib_name, store, cfg, data, multiscale, pipeline, unbuffered):
    lib = ca.lib if lib_name == "mine" else ref.lib()
    sink = _Sink(lib, store, cfg, multiscale, unbuffered=int(unbuffered))
    if lib_name == "mine":
        s = ca.GpuStream(cfg, sink, pipeline=pipeline)
    else:
        s = ref.RefCpuStream(cfg, sink)
    for part in np.array_split(data, 3):
        assert s.append(part) == 0, ca.last_error()
    assert s.flush() == 0, ca.last_error()
    s.close()
    sink.close()
    return _tree(store)


def _compare(cfg, got, want, multiscale):
    assert sorted(got) == sorted(want)
    g, _, chunk_bytes = geometry(cfg)
    shards_g, shards_w = {}, {}
    for path in want:
        if path.endswith("zarr.json"):
            assert got[path] == want[path], path
            continue
        parts = path.split("/")
        lv = int(parts[1]) if multiscale else 0
        shards_g[(lv, path)] = got[path]
        shards_w[(lv, path)] = want[path]
    assert shards_w
    assert_same_shards(shards_g, shards_w, lambda lv: g[lv][0], cfg.codec, chunk_bytes, "fs sink")


@pytest.mark.parametrize("pipeline", [ca.PIPELINE_PLAIN, ca.PIPELINE_BATCHED])
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_stream_into_fs_sink(tmp_path, name, pipeline):
    multiscale, dims, dtype, codec, align = CASES[name]
    cfg = ca.Config(dims, dtype=dtype, codec=codec, shard_alignment=align, buffer_capacity_bytes=1 << 16)
    data = _data(dims, dtype)
    base = str(tmp_path)
    unbuffered = False
    if align:
        base = _odirect_dir(tmp_path)
        if base is None:
            pytest.skip("no file system here takes O_DIRECT")
        unbuffered = True
        base = os.path.join(base, f"fs_{name}_{pipeline}_{os.getpid()}")
    print(f"[fs sink] {name} pipeline={pipeline} unbuffered={unbuffered}", flush=True)
    got = _write("mine", os.path.join(base, "mine"), cfg, data, multiscale, pipeline, unbuffered)
    want = _write("ref", os.path.join(base, "ref"), cfg, data, multiscale, pipeline, False)
    _compare(cfg, got, want, multiscale)
    if not multiscale and dims[0].size:
        arr, _, _ = read_array(os.path.join(base, "mine", "arr"))
        np.testing.assert_array_equal(arr.reshape(-1), data.astype(arr.dtype))
    if unbuffered:
        import shutil
        shutil.rmtree(base, ignore_errors=True)


@pytest.mark.parametrize("pipeline", [ca.PIPELINE_PLAIN, ca.PIPELINE_BATCHED])
def test_reset_into_another_fs_sink(tmp_path, pipeline):
    # one stream, two acquisitions: the second reset hands it a new store.
    # The new sink's write fences count from zero, so none of the first
    # sink's fences may be waited for on it; the first sink is destroyed only
    # after the reset has let go of it. Both stores come out identical.
    multiscale, dims, dtype, codec, _ = CASES["multiscale_zstd"]
    cfg = ca.Config(dims, dtype=dtype, codec=codec, buffer_capacity_bytes=1 << 16)
    data = _data(dims, dtype)
    print(f"[fs sink reset] pipeline={pipeline}", flush=True)
    stores = [str(tmp_path / "a"), str(tmp_path / "b")]
    sinks = [_Sink(ca.lib, p, cfg, multiscale) for p in stores]
    s = ca.GpuStream(cfg, sinks[0], pipeline=pipeline)
    for k, sink in enumerate(sinks):
        if k:
            assert s.reset(sink) == 0, ca.last_error()
            sinks[k - 1].close()
        for part in np.array_split(data, 3):
            assert s.append(part) == 0, ca.last_error()
        assert s.flush() == 0, ca.last_error()
    s.close()
    sinks[-1].close()
    a, b = _tree(stores[0]), _tree(stores[1])
    assert sorted(a) == sorted(b) and a
    for path in a:
        assert a[path] == b[path], path
