"""Pythonic wrappers over the C ABI (configuration, streams, sinks)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import (CODEC_ID, DTYPE_ID, NP_DTYPE, REDUCE_ID, CodecConfig, Dimension, MemoryInfo,
               Slice, StreamOptions, TileStreamConfiguration, lib)


def last_error() -> str:
    v = lib.chucky_last_error()
    return v.decode() if v else ""


def crc32c(data: bytes) -> int:
    buf = C.create_string_buffer(bytes(data), len(data))
    return lib.chucky_crc32c(C.cast(buf, C.c_void_p), len(data))


@dataclass
class Dim:
    size: int
    chunk: int
    cps: int = 0
    name: str = ""
    downsample: bool = False
    storage_position: Optional[int] = None


@dataclass
class Config:
    """Mirror of struct tile_stream_configuration (include/types.stream.h)."""
    dims: Sequence[Dim]
    dtype: str = "u16"
    codec: str = "none"
    level: int = 1
    reduce: str = "mean"
    append_reduce: str = "mean"
    max_nlod: int = 0
    epochs_per_batch: int = 0
    target_batch_chunks: int = 0
    buffer_capacity_bytes: int = 1 << 20
    shard_alignment: int = 0
    metadata_update_interval_s: float = 1e9
    _keep: list = field(default_factory=list, repr=False)

    def to_c(self) -> TileStreamConfiguration:
        rank = len(self.dims)
        arr = (Dimension * rank)()
        names = []
        for i, d in enumerate(self.dims):
            nm = (d.name or "abcdefghijklmnopqrstuvwxyz"[i]).encode()
            names.append(nm)
            arr[i].size = d.size
            arr[i].chunk_size = d.chunk
            arr[i].chunks_per_shard = d.cps
            arr[i].name = nm
            arr[i].downsample = 1 if d.downsample else 0
            arr[i].storage_position = i if d.storage_position is None else d.storage_position
            arr[i].ngff.scale = 1.0
        cfg = TileStreamConfiguration()
        cfg.buffer_capacity_bytes = self.buffer_capacity_bytes
        cfg.dtype = DTYPE_ID[self.dtype]
        cfg.rank = rank
        cfg.dimensions = C.cast(arr, C.POINTER(Dimension))
        cfg.codec = CodecConfig(CODEC_ID[self.codec], self.level, 0)
        cfg.reduce_method = REDUCE_ID[self.reduce]
        cfg.append_reduce_method = REDUCE_ID[self.append_reduce]
        cfg.max_nlod = self.max_nlod
        cfg.epochs_per_batch = self.epochs_per_batch
        cfg.target_batch_chunks = self.target_batch_chunks
        cfg.metadata_update_interval_s = self.metadata_update_interval_s
        cfg.shard_alignment = self.shard_alignment
        cfg.max_threads = 0
        self._keep[:] = [arr, names, cfg]
        return cfg

    @property
    def np_dtype(self):
        return np.dtype(NP_DTYPE[self.dtype])


class MemSink:
    """In-memory shard sink (chucky_mem_sink_*)."""

    def __init__(self, deferred: bool = False):
        self.h = lib.chucky_mem_sink_create()
        if deferred:
            lib.chucky_mem_sink_set_deferred(self.h, 1)

    @property
    def shard_sink(self):
        return lib.chucky_mem_sink_as_shard_sink(self.h)

    def shards(self) -> Dict[Tuple[int, int], bytes]:
        lib.chucky_mem_sink_drain(self.h)
        out = {}
        n = lib.chucky_mem_sink_count(self.h)
        for i in range(n):
            lv = C.c_uint8()
            si = C.c_uint64()
            data = C.POINTER(C.c_uint8)()
            size = C.c_size_t()
            fin = C.c_int()
            lib.chucky_mem_sink_get(self.h, i, C.byref(lv), C.byref(si), C.byref(data),
                                    C.byref(size), C.byref(fin))
            out[(lv.value, si.value)] = C.string_at(data, size.value) if size.value else b""
        return out

    def finalized(self) -> Dict[Tuple[int, int], bool]:
        out = {}
        for i in range(lib.chucky_mem_sink_count(self.h)):
            lv = C.c_uint8()
            si = C.c_uint64()
            fin = C.c_int()
            lib.chucky_mem_sink_get(self.h, i, C.byref(lv), C.byref(si), None, None, C.byref(fin))
            out[(lv.value, si.value)] = bool(fin.value)
        return out

    def append_sizes(self, level: int) -> Optional[List[int]]:
        buf = (C.c_uint64 * 64)()
        n = C.c_uint8()
        if lib.chucky_mem_sink_append_sizes(self.h, level, buf, C.byref(n)):
            return None
        return list(buf[: n.value])

    def inject_failure(self, fail_open=False, fail_write=False):
        lib.chucky_mem_sink_inject_failure(self.h, int(fail_open), int(fail_write))

    def close(self):
        if self.h:
            lib.chucky_mem_sink_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class _ZarrFsConfig(C.Structure):
    _fields_ = [("store_path", C.c_char_p), ("array_name", C.c_char_p), ("data_type", C.c_int),
                ("fill_value", C.c_double), ("rank", C.c_uint8),
                ("dimensions", C.POINTER(Dimension)), ("unbuffered", C.c_int),
                ("codec", CodecConfig)]


class _ZarrFsMultiscaleConfig(C.Structure):
    _fields_ = [("store_path", C.c_char_p), ("array_name", C.c_char_p), ("data_type", C.c_int),
                ("fill_value", C.c_double), ("rank", C.c_uint8),
                ("dimensions", C.POINTER(Dimension)), ("nlod", C.c_int), ("unbuffered", C.c_int),
                ("codec", CodecConfig)]


class FsSink:
    """Zarr v3 filesystem store (zarr_fs_sink_* of zarr_fs_sink.h): one array
    `name` under `store`, or an OME-NGFF multiscale group of arrays 0, 1, ...
    when the stream downsamples (zarr_fs_multiscale_sink_*, levels from the
    stream's configuration). `unbuffered` writes the shards with O_DIRECT
    (the store must be on a filesystem that supports it)."""

    def __init__(self, store: str, cfg: "Config", name: str = "arr", unbuffered: bool = False):
        multiscale = any(d.downsample for d in cfg.dims)
        self._pre = "zarr_fs_multiscale_sink" if multiscale else "zarr_fs_sink"
        cfg_t = _ZarrFsMultiscaleConfig if multiscale else _ZarrFsConfig
        fn = getattr(lib, self._pre + "_create")
        fn.restype = C.c_void_p
        fn.argtypes = [C.POINTER(cfg_t)]
        getattr(lib, self._pre + "_as_shard_sink").restype = C.c_void_p
        getattr(lib, self._pre + "_as_shard_sink").argtypes = [C.c_void_p]
        getattr(lib, self._pre + "_destroy").argtypes = [C.c_void_p]
        c = cfg.to_c()
        z = cfg_t()
        z.store_path = store.encode()
        z.array_name = name.encode()
        z.data_type = c.dtype
        z.fill_value = 0.0
        z.rank = c.rank
        z.dimensions = c.dimensions
        z.unbuffered = int(unbuffered)
        z.codec = c.codec
        if multiscale:
            z.nlod = 0
        self._keep = (c, z, cfg)
        self.h = fn(C.byref(z))
        if not self.h:
            raise RuntimeError("fs sink: " + last_error())

    @property
    def shard_sink(self):
        return getattr(lib, self._pre + "_as_shard_sink")(self.h)

    def close(self):
        if self.h:
            getattr(lib, self._pre + "_destroy")(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DiscardSink:
    def __init__(self):
        self.h = lib.chucky_discard_sink_create()

    @property
    def shard_sink(self):
        return lib.chucky_discard_sink_as_shard_sink(self.h)

    @property
    def bytes(self) -> int:
        return lib.chucky_discard_sink_bytes(self.h)

    def close(self):
        if self.h:
            lib.chucky_discard_sink_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def tuned_codec(codec: str, bpe: int, chunk_bytes: int, chunks_per_launch: int,
                multiscale: bool = False) -> Tuple[int, int]:
    """The autotuner's launch shape: zstd (block bytes, chunk tables 0/1), lz4
    (segment bytes or 0 for whole chunks, hash log)."""
    out = (C.c_uint32 * 2)()
    if lib.chucky_tune_codec(CODEC_ID[codec], int(multiscale), bpe, chunk_bytes, chunks_per_launch, out):
        raise ValueError("chucky_tune_codec: invalid arguments")
    return out[0], out[1]


def memory_estimate(cfg: Config, pipeline: Optional[int] = None) -> MemoryInfo:
    """tile_stream_gpu_memory_estimate; with a pipeline, the estimate of what
    that pipeline allocates (tile_stream_gpu_memory_estimate_ex)."""
    info = MemoryInfo()
    c = cfg.to_c()
    if pipeline is None:
        rc = lib.tile_stream_gpu_memory_estimate(C.byref(c), C.byref(info))
    else:
        opts = StreamOptions()
        lib.tile_stream_gpu_options_default(C.byref(opts))
        opts.pipeline = pipeline
        rc = lib.tile_stream_gpu_memory_estimate_ex(C.byref(c), C.byref(opts), C.byref(info))
    if rc:
        raise ValueError("invalid configuration")
    return info


class Server:
    """chucky_server_*: concurrent streams admitted against an HBM budget."""

    def __init__(self, hbm_budget_bytes: int, max_concurrent: int, pipeline: int = 1):
        from . import RequestReport  # noqa: F401
        self.h = lib.chucky_server_create(hbm_budget_bytes, max_concurrent, pipeline)
        self._keep = []

    def submit(self, cfg: Config, sink, data, device: bool = False) -> int:
        c = cfg.to_c()
        rid = C.c_uint64()
        if device:
            ptr, nbytes = data.ptr, data.nbytes
        else:
            arr = np.ascontiguousarray(data)
            ptr, nbytes = arr.ctypes.data, arr.nbytes
            data = arr
        self._keep.append((cfg, c, sink, data))
        if lib.chucky_server_submit(self.h, C.byref(c), sink.shard_sink, ptr, nbytes, int(device),
                                    C.byref(rid)):
            raise RuntimeError("submit failed: " + last_error())
        return rid.value

    def wait(self) -> int:
        return lib.chucky_server_wait_all(self.h)

    def reports(self):
        from . import RequestReport
        buf = (RequestReport * 4096)()
        n = lib.chucky_server_reports(self.h, buf, 4096)
        return [buf[i] for i in range(n)]

    def latencies(self) -> List[float]:
        buf = (C.c_double * (1 << 20))()
        n = lib.chucky_server_latencies(self.h, buf, 1 << 20)
        return list(buf[:n])

    def peaks(self):
        r, b = C.c_int(), C.c_size_t()
        lib.chucky_server_peaks(self.h, C.byref(r), C.byref(b))
        return r.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib.chucky_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class GpuStream:
    """tile_stream_gpu_* through its writer vtable."""

    def __init__(self, cfg: Config, sink, pipeline: int = 0, workers: int = 0, **options):
        """options: fields of struct tile_stream_gpu_options (zstd_block_bytes,
        lz4_hash_log, host_gather_workgroups, device); 0 keeps the autotuned value."""
        self.cfg = cfg
        self._c = cfg.to_c()
        self.sink = sink
        opts = StreamOptions()
        lib.tile_stream_gpu_options_default(C.byref(opts))
        opts.pipeline = pipeline
        opts.compress_workers = workers
        for k, v in options.items():
            setattr(opts, k, v)
        self.h = lib.tile_stream_gpu_create_ex(C.byref(self._c), sink.shard_sink, C.byref(opts))
        if not self.h:
            raise RuntimeError("tile_stream_gpu_create failed: " + last_error())
        self.w = lib.tile_stream_gpu_writer(self.h)

    @property
    def layout(self):
        return lib.tile_stream_gpu_layout(self.h).contents

    @property
    def epoch_elements(self) -> int:
        return self.layout.epoch_elements

    @property
    def cursor(self) -> int:
        return lib.tile_stream_gpu_cursor(self.h)

    def status(self):
        return lib.tile_stream_gpu_status(self.h)

    def level_geometry(self, level: int):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        if lib.tile_stream_gpu_level_geometry(self.h, level, C.byref(a), C.byref(b), C.byref(c)):
            raise ValueError("bad level")
        return a.value, b.value, c.value

    def metrics(self):
        return lib.tile_stream_gpu_get_metrics(self.h)

    def append(self, data) -> int:
        """Append host bytes (numpy array or bytes). Returns writer error code."""
        buf = np.ascontiguousarray(data)
        beg = buf.ctypes.data
        r = lib.writer_append(self.w, Slice(beg, beg + buf.nbytes))
        self._last_rest = (r.rest.beg or 0) - beg if r.rest.beg else 0
        return r.error

    def append_device(self, ptr: int, nbytes: int) -> int:
        r = lib.tile_stream_gpu_append_device(self.h, ptr, nbytes)
        return r.error

    def flush(self) -> int:
        return lib.writer_flush(self.w).error

    def reset(self, sink=None) -> int:
        if sink is not None:
            self.sink = sink
        return lib.tile_stream_gpu_reset(self.h, sink.shard_sink if sink is not None else None)

    def latencies(self) -> List[float]:
        buf = (C.c_double * 65536)()
        n = lib.tile_stream_gpu_epoch_latencies(self.h, buf, 65536)
        return list(buf[:n])

    def close(self):
        if getattr(self, "h", None):
            lib.tile_stream_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class GpuMultiArray:
    """multiarray_tile_stream_gpu_* through its multiarray_writer vtable: one
    writer, several arrays, each with its own GPU tile stream."""

    def __init__(self, cfgs, sinks):
        self.cfgs = list(cfgs)
        self.sinks = list(sinks)
        arr = (TileStreamConfiguration * len(self.cfgs))()
        self._c = [c.to_c() for c in self.cfgs]
        for i, c in enumerate(self._c):
            arr[i] = c
        self._arr = arr
        ss = (C.c_void_p * len(self.sinks))(*[s.shard_sink for s in self.sinks])
        self.h = lib.multiarray_tile_stream_gpu_create(len(self.cfgs), arr, ss, 1)
        if not self.h:
            raise RuntimeError("multiarray_tile_stream_gpu_create failed: " + last_error())
        self.w = lib.multiarray_tile_stream_gpu_writer(self.h)

    def update(self, index: int, data):
        """Returns (error code, bytes consumed)."""
        buf = np.ascontiguousarray(data)
        beg = buf.ctypes.data
        r = self.w.contents.update(self.w, index, Slice(beg, beg + buf.nbytes))
        rest = (r.rest.end or 0) - (r.rest.beg or 0) if r.rest.beg else 0
        return r.error, buf.nbytes - rest

    def flush(self) -> int:
        return self.w.contents.flush(self.w).error

    def metrics(self):
        return lib.multiarray_tile_stream_gpu_get_metrics(self.h)

    def close(self):
        if self.h:
            lib.multiarray_tile_stream_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
