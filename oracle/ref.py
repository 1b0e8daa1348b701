"""ctypes access to the reference's own CPU write path (oracle/_ref).

TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles the reference's CPU
backend (src/cpu/*, src/stream/*, src/lod/*, src/zarr/{shard_delivery,crc32c}
...) from its sources into oracle/_ref/libchucky_ref.so. Only tests/, the
smoke check and bench.py's cpu_baseline leg may load it, and only as the
checker / the timed CPU baseline; the product never does.

The reference's tile_stream_cpu_* functions take the same C structures as the
product's API (struct tile_stream_configuration, struct shard_sink), so the
same configuration and the product's in-memory sink can be handed to both.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(_HERE, "_ref", "libchucky_ref.so")

_lib = None


def available() -> bool:
    return os.path.exists(REF_LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"{REF_LIB} not built (make -C oracle)")
        _lib = C.CDLL(REF_LIB)
        from chucky_amd import (TileStreamConfiguration, Slice, WriterResult, Writer,
                                TileStreamLayout)
        P = C.c_void_p
        _lib.tile_stream_cpu_create.restype = P
        _lib.tile_stream_cpu_create.argtypes = [C.POINTER(TileStreamConfiguration), P]
        _lib.tile_stream_cpu_destroy.argtypes = [P]
        _lib.tile_stream_cpu_writer.restype = C.POINTER(Writer)
        _lib.tile_stream_cpu_writer.argtypes = [P]
        _lib.tile_stream_cpu_cursor.restype = C.c_uint64
        _lib.tile_stream_cpu_cursor.argtypes = [P]
        _lib.tile_stream_cpu_layout.restype = C.POINTER(TileStreamLayout)
        _lib.tile_stream_cpu_layout.argtypes = [P]
        _lib.writer_append.restype = WriterResult
        _lib.writer_append.argtypes = [C.POINTER(Writer), Slice]
        _lib.writer_flush.restype = WriterResult
        _lib.writer_flush.argtypes = [C.POINTER(Writer)]
        from chucky_amd import MultiarrayWriter
        _lib.multiarray_tile_stream_cpu_create.restype = P
        _lib.multiarray_tile_stream_cpu_create.argtypes = [C.c_int, C.POINTER(TileStreamConfiguration),
                                                           C.POINTER(C.c_void_p), C.c_int]
        _lib.multiarray_tile_stream_cpu_destroy.argtypes = [P]
        _lib.multiarray_tile_stream_cpu_writer.restype = C.POINTER(MultiarrayWriter)
        _lib.multiarray_tile_stream_cpu_writer.argtypes = [P]
    return _lib


class RefCpuStream:
    """The reference's tile_stream_cpu_* writer."""

    def __init__(self, cfg, sink):
        L = lib()
        self._c = cfg.to_c()
        self.h = L.tile_stream_cpu_create(C.byref(self._c), sink.shard_sink)
        if not self.h:
            raise RuntimeError("tile_stream_cpu_create failed")
        self.w = L.tile_stream_cpu_writer(self.h)

    @property
    def epoch_elements(self) -> int:
        return lib().tile_stream_cpu_layout(self.h).contents.epoch_elements

    @property
    def cursor(self) -> int:
        return lib().tile_stream_cpu_cursor(self.h)

    def append(self, data) -> int:
        import numpy as np
        from chucky_amd import Slice
        buf = np.ascontiguousarray(data)
        beg = buf.ctypes.data
        return lib().writer_append(self.w, Slice(beg, beg + buf.nbytes)).error

    def flush(self) -> int:
        return lib().writer_flush(self.w).error

    def close(self):
        if self.h:
            lib().tile_stream_cpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class RefMultiArray:
    """The reference's multiarray_tile_stream_cpu_* writer
    (src/multiarray.cpu.h:31-46), same calls as chucky_amd.GpuMultiArray."""

    def __init__(self, cfgs, sinks):
        from chucky_amd import TileStreamConfiguration
        L = lib()
        self._c = [c.to_c() for c in cfgs]
        arr = (TileStreamConfiguration * len(cfgs))()
        for i, c in enumerate(self._c):
            arr[i] = c
        self._arr = arr
        ss = (C.c_void_p * len(sinks))(*[s.shard_sink for s in sinks])
        self._ss = ss
        self._sinks = list(sinks)  # the sinks must outlive the writer
        self.h = L.multiarray_tile_stream_cpu_create(len(cfgs), arr, ss, 0)
        if not self.h:
            raise RuntimeError("multiarray_tile_stream_cpu_create failed")
        self.w = L.multiarray_tile_stream_cpu_writer(self.h)

    def update(self, index: int, data):
        """Returns (error code, bytes consumed)."""
        import numpy as np
        from chucky_amd import Slice
        buf = np.ascontiguousarray(data)
        beg = buf.ctypes.data
        r = self.w.contents.update(self.w, index, Slice(beg, beg + buf.nbytes))
        rest = (r.rest.end or 0) - (r.rest.beg or 0) if r.rest.beg else 0
        return r.error, buf.nbytes - rest

    def flush(self) -> int:
        return self.w.contents.flush(self.w).error

    def close(self):
        if self.h:
            lib().multiarray_tile_stream_cpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
