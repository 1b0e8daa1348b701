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
    def __init__(self, lib, store, cfg, multiscale, name=b"arr"):
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
        z.unbuffered = 0
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
