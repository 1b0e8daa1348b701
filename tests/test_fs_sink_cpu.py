"""The product's zarr filesystem sink against the reference's (both driven by
the reference's CPU stream, so only the sink differs): identical file trees,
zarr.json documents and shard bytes. CPU only."""
import ctypes as C
import os

import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd import CodecConfig, Dimension
from chucky_amd.api import Dim
from oracle import ref
from tests.helpers import random_typed

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


class ZarrConfig(C.Structure):
    _fields_ = [("store_path", C.c_char_p), ("array_name", C.c_char_p), ("data_type", C.c_int),
                ("fill_value", C.c_double), ("rank", C.c_uint8),
                ("dimensions", C.POINTER(Dimension)), ("unbuffered", C.c_int),
                ("codec", CodecConfig)]


class ZarrMultiscaleConfig(C.Structure):
    _fields_ = [("store_path", C.c_char_p), ("array_name", C.c_char_p), ("data_type", C.c_int),
                ("fill_value", C.c_double), ("rank", C.c_uint8),
                ("dimensions", C.POINTER(Dimension)), ("nlod", C.c_int), ("unbuffered", C.c_int),
                ("codec", CodecConfig)]


def _bind(lib, multiscale):
    pre = "zarr_fs_multiscale_sink" if multiscale else "zarr_fs_sink"
    cfg_t = ZarrMultiscaleConfig if multiscale else ZarrConfig
    getattr(lib, pre + "_create").restype = C.c_void_p
    getattr(lib, pre + "_create").argtypes = [C.POINTER(cfg_t)]
    getattr(lib, pre + "_as_shard_sink").restype = C.c_void_p
    getattr(lib, pre + "_as_shard_sink").argtypes = [C.c_void_p]
    getattr(lib, pre + "_destroy").argtypes = [C.c_void_p]
    return pre, cfg_t


class _Sink:
    def __init__(self, lib, store, cfg, multiscale, name=b"arr", unbuffered=0):
        self.lib = lib
        self.pre, cfg_t = _bind(lib, multiscale)
        c = cfg.to_c()
        z = cfg_t()
        z.store_path = store.encode()
        z.array_name = name
        z.data_type = c.dtype
        z.fill_value = 0.0
        z.rank = c.rank
        z.dimensions = c.dimensions
        z.unbuffered = unbuffered
        z.codec = c.codec
        if multiscale:
            z.nlod = 0
        self._keep = (c, z, cfg)
        self.h = getattr(lib, self.pre + "_create")(C.byref(z))
        assert self.h, "sink create failed"

    @property
    def shard_sink(self):
        return getattr(self.lib, self.pre + "_as_shard_sink")(self.h)

    def close(self):
        getattr(self.lib, self.pre + "_destroy")(self.h)


def _tree(root):
    out = {}
    for dp, _, fs in os.walk(root):
        for f in fs:
            p = os.path.join(dp, f)
            out[os.path.relpath(p, root)] = open(p, "rb").read()
    return out


CASES = [
    (False, [Dim(20, 8, 2), Dim(16, 8, 1), Dim(24, 8, 3)], "u8", "none"),
    (False, [Dim(0, 4, 2), Dim(12, 4, 0), Dim(16, 8, 0)], "u16", "none"),      # unbounded dim 0
    (True, [Dim(16, 2, 4, downsample=True), Dim(8, 4, 0, downsample=True),
            Dim(12, 4, 0, downsample=True)], "u16", "none"),
    (True, [Dim(8, 4, 1), Dim(24, 4, 0, downsample=True), Dim(20, 4, 0, downsample=True)],
     "f32", "none"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fs_sink_tree_matches_reference(tmp_path, case):
    multiscale, dims, dtype, codec = CASES[case]
    cfg = ca.Config(dims, dtype=dtype, codec=codec, buffer_capacity_bytes=4096)
    n = int(np.prod([d.size or 40 for d in dims]))
    data = random_typed(n, dtype, seed=case)
    trees = []
    for which, lib in (("ref", ref.lib()), ("mine", ca.lib)):
        store = str(tmp_path / which)
        sink = _Sink(lib, store, cfg, multiscale)
        r = ref.RefCpuStream(cfg, sink)
        assert r.append(data) == 0
        assert r.flush() == 0
        r.close()
        sink.close()
        trees.append(_tree(store))
    want, got = trees
    assert sorted(got) == sorted(want)
    for k in want:
        assert got[k] == want[k], k


def test_rewrite_into_existing_store_truncates(tmp_path):
    # a second acquisition into the same store, more compressible than the
    # first: every shard file must equal a fresh write (no stale tail, and so
    # no stale index, past the new index)
    dims = [Dim(16, 4, 2), Dim(32, 16, 1), Dim(32, 16, 2)]
    cfg = ca.Config(dims, dtype="u16", codec="zstd", buffer_capacity_bytes=4096)
    noisy = random_typed(16 * 32 * 32, "u16", seed=3)
    flat = np.zeros(16 * 32 * 32, dtype=np.uint16)

    def write(store, data):
        sink = _Sink(ca.lib, store, cfg, False)
        r = ref.RefCpuStream(cfg, sink)
        assert r.append(data) == 0
        assert r.flush() == 0
        r.close()
        sink.close()
        return _tree(store)

    first = write(str(tmp_path / "a"), noisy)
    again = write(str(tmp_path / "a"), flat)
    fresh = write(str(tmp_path / "b"), flat)
    assert sorted(again) == sorted(fresh)
    assert any(len(first[k]) > len(fresh[k]) for k in fresh if k in first and "/c/" in "/" + k)
    for k in fresh:
        assert again[k] == fresh[k], k


READBACK = [
    ([Dim(0, 1, 4), Dim(8, 4, 0), Dim(12, 6, 0)], "u16", "none"),   # tests/test_zarr_readback.c
    ([Dim(0, 1, 4), Dim(8, 4, 0), Dim(12, 6, 0)], "u16", "zstd"),
    ([Dim(0, 1, 4), Dim(8, 4, 0), Dim(12, 6, 0)], "u16", "lz4"),
    ([Dim(20, 8, 2), Dim(16, 8, 1), Dim(24, 8, 3)], "f32", "zstd"),  # ragged dim 0, several shards
]


@pytest.mark.parametrize("case", range(len(READBACK)))
def test_readback_returns_the_input(tmp_path, case):
    from tests.zarr_readback import read_array
    dims, dtype, codec = READBACK[case]
    cfg = ca.Config(dims, dtype=dtype, codec=codec, buffer_capacity_bytes=4096)
    nt = 4 if dims[0].size == 0 else dims[0].size
    shape = (nt,) + tuple(d.size for d in dims[1:])
    full = data = (np.arange(int(np.prod(shape))) & 0xFFFF).astype(ca.NP_DTYPE[dtype]).reshape(shape)
    for which, lib in (("ref", ref.lib()), ("mine", ca.lib)):
        store = str(tmp_path / which)
        sink = _Sink(lib, store, cfg, False, name=b"0")
        r = ref.RefCpuStream(cfg, sink)
        assert r.append(full) == 0 and r.flush() == 0
        r.close()
        sink.close()
        got, meta, st = read_array(os.path.join(store, "0"))
        assert list(got.shape) == list(shape), (which, got.shape)
        np.testing.assert_array_equal(got, data, err_msg=which)
        assert st["chunks"] > 0 and st["codec"] == codec


def test_readback_multiscale_and_corruption(tmp_path):
    from tests.zarr_readback import read_array, read_multiscale
    # LOD over y, x with one chunk a shard along dim 0: every level's
    # metadata describes its shards
    dims = [Dim(6, 2, 1), Dim(16, 4, 0, downsample=True), Dim(24, 4, 0, downsample=True)]
    cfg = ca.Config(dims, dtype="u16", codec="zstd", buffer_capacity_bytes=4096)
    data = random_typed(6 * 16 * 24, "u16", seed=8).reshape(6, 16, 24)
    levels = {}
    for which, lib in (("ref", ref.lib()), ("mine", ca.lib)):
        store = str(tmp_path / which)
        sink = _Sink(lib, store, cfg, True)
        r = ref.RefCpuStream(cfg, sink)
        assert r.append(data) == 0 and r.flush() == 0
        r.close()
        sink.close()
        levels[which] = read_multiscale(os.path.join(store, "arr"))
    assert len(levels["mine"]) == 4
    np.testing.assert_array_equal(levels["mine"][0], data)
    assert [a.shape for a in levels["mine"]] == [(6, 16, 24), (6, 8, 12), (6, 4, 6), (6, 2, 3)]
    # level 1 is the 2x2 mean of level 0 (u16: truncating integer mean)
    l0 = data.astype(np.uint32)
    want1 = ((l0[:, 0::2, 0::2] + l0[:, 0::2, 1::2] + l0[:, 1::2, 0::2] + l0[:, 1::2, 1::2]) // 4)
    np.testing.assert_array_equal(levels["mine"][1], want1.astype(np.uint16))
    for a, b in zip(levels["mine"], levels["ref"]):
        np.testing.assert_array_equal(a, b)
    # a flipped bit in a shard index is caught by its CRC
    shard = os.path.join(str(tmp_path / "mine"), "arr", "0", "c", "0", "0", "0")
    blob = bytearray(open(shard, "rb").read())
    blob[-10] ^= 1
    open(shard, "wb").write(bytes(blob))
    with pytest.raises(AssertionError, match="CRC"):
        read_array(os.path.join(str(tmp_path / "mine"), "arr", "0"))


@pytest.mark.parametrize("dim0_downsample", [False, True])
def test_coarse_levels_metadata_quirk(tmp_path, dim0_downsample):
    # With several chunks a shard along dim 0, the reference delivers levels
    # >= 1 one epoch a shard (their epoch extent is one chunk, so their
    # chunks-per-shard along dim 0 clamps to 1, src/stream/config.c:298-372)
    # but describes them in zarr.json with level 0's shard shape
    # (src/zarr/zarr_metadata.c:33-207): a reader expects a longer index than
    # the file holds. Both sinks write the same bytes; level 0 reads back.
    from tests.zarr_readback import read_array
    dims = [Dim(16, 2, 4, downsample=dim0_downsample), Dim(8, 4, 0, downsample=True),
            Dim(12, 4, 0, downsample=True)]
    cfg = ca.Config(dims, dtype="u16", codec="zstd", buffer_capacity_bytes=4096)
    data = random_typed(16 * 8 * 12, "u16", seed=8).reshape(16, 8, 12)
    trees = {}
    for which, lib in (("ref", ref.lib()), ("mine", ca.lib)):
        store = str(tmp_path / which)
        sink = _Sink(lib, store, cfg, True)
        r = ref.RefCpuStream(cfg, sink)
        assert r.append(data) == 0 and r.flush() == 0
        r.close()
        sink.close()
        trees[which] = _tree(store)
        np.testing.assert_array_equal(read_array(os.path.join(store, "arr", "0"))[0], data)
        with pytest.raises(AssertionError, match="CRC"):
            read_array(os.path.join(store, "arr", "1"))
    assert trees["mine"] == trees["ref"]
