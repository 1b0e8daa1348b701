"""chucky-amd: MI355X-native streaming Zarr v3 writer.

Python mirror of the C API in include/stream.gpu.h (the reference's
src/stream.gpu.h) over ctypes. The native library libchucky_amd.so (gfx950
HIP kernels + C++ host runtime) is built in-tree by `make -C chucky_amd`; the
package refuses to import without it -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchucky_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"chucky_amd: native library missing at {LIB_PATH}; build it with "
        "`make -C chucky_amd` (hipcc, gfx950). No CPU fallback exists.")

lib = C.CDLL(LIB_PATH)

# ---- enums (include/dtype.h, types.codec.h, types.lod.h) -------------------

DTYPES = ["u8", "u16", "u32", "u64", "i8", "i16", "i32", "i64", "f16", "f32", "f64"]
DTYPE_ID = {n: i for i, n in enumerate(DTYPES)}
NP_DTYPE = {"u8": "uint8", "u16": "uint16", "u32": "uint32", "u64": "uint64",
            "i8": "int8", "i16": "int16", "i32": "int32", "i64": "int64",
            "f16": "float16", "f32": "float32", "f64": "float64"}
CODEC_NONE, CODEC_LZ4, CODEC_ZSTD, CODEC_BLOSC_LZ4, CODEC_BLOSC_ZSTD = range(5)
CODEC_ID = {"none": CODEC_NONE, "lz4": CODEC_LZ4, "zstd": CODEC_ZSTD}
REDUCE = ["mean", "min", "max", "median", "max_sup", "min_sup"]
REDUCE_ID = {n: i for i, n in enumerate(REDUCE)}
PIPELINE_PLAIN, PIPELINE_BATCHED = 0, 1

MAX_RANK = 64


class OmengffAxis(C.Structure):
    _fields_ = [("unit", C.c_char_p), ("scale", C.c_double), ("type", C.c_int)]


class Dimension(C.Structure):
    _fields_ = [("size", C.c_uint64), ("chunk_size", C.c_uint64),
                ("chunks_per_shard", C.c_uint64), ("name", C.c_char_p),
                ("downsample", C.c_int), ("storage_position", C.c_uint8),
                ("ngff", OmengffAxis)]


class CodecConfig(C.Structure):
    _fields_ = [("id", C.c_int), ("level", C.c_uint8), ("shuffle", C.c_int)]


class TileStreamConfiguration(C.Structure):
    _fields_ = [("buffer_capacity_bytes", C.c_size_t), ("dtype", C.c_int),
                ("rank", C.c_uint8), ("dimensions", C.POINTER(Dimension)),
                ("codec", CodecConfig), ("reduce_method", C.c_int),
                ("append_reduce_method", C.c_int), ("max_nlod", C.c_int),
                ("epochs_per_batch", C.c_uint8), ("target_batch_chunks", C.c_uint32),
                ("metadata_update_interval_s", C.c_float),
                ("shard_alignment", C.c_size_t), ("max_threads", C.c_int)]


class Slice(C.Structure):
    _fields_ = [("beg", C.c_void_p), ("end", C.c_void_p)]


class WriterResult(C.Structure):
    _fields_ = [("error", C.c_int), ("rest", Slice)]


class Writer(C.Structure):
    pass


Writer._fields_ = [("append", C.CFUNCTYPE(WriterResult, C.POINTER(Writer), Slice)),
                   ("flush", C.CFUNCTYPE(WriterResult, C.POINTER(Writer)))]


class MultiarrayWriter(C.Structure):
    pass


MultiarrayWriter._fields_ = [
    ("update", C.CFUNCTYPE(WriterResult, C.POINTER(MultiarrayWriter), C.c_int, Slice)),
    ("flush", C.CFUNCTYPE(WriterResult, C.POINTER(MultiarrayWriter)))]


class StreamMetric(C.Structure):
    _fields_ = [("name", C.c_char_p), ("ms", C.c_float), ("best_ms", C.c_float),
                ("input_bytes", C.c_double), ("output_bytes", C.c_double),
                ("count", C.c_int)]


METRIC_NAMES = ["memcpy", "h2d", "lod_gather", "lod_reduce", "lod_append_fold",
                "lod_morton_chunk", "scatter", "compress", "aggregate", "d2h", "sink"]


class StreamMetrics(C.Structure):
    _fields_ = [(n, StreamMetric) for n in METRIC_NAMES]


class TileStreamLayout(C.Structure):
    _fields_ = [("lifted_rank", C.c_uint8), ("lifted_shape", C.c_uint64 * MAX_RANK),
                ("lifted_strides", C.c_int64 * MAX_RANK), ("chunk_elements", C.c_uint64),
                ("chunk_stride", C.c_uint64), ("chunks_per_epoch", C.c_uint64),
                ("epoch_elements", C.c_uint64), ("chunk_pool_bytes", C.c_size_t)]


class TileStreamStatus(C.Structure):
    _fields_ = [("nlod", C.c_int), ("append_downsample", C.c_int),
                ("epochs_per_batch", C.c_uint32), ("max_compressed_size", C.c_size_t),
                ("dtype", C.c_int), ("codec", CodecConfig), ("codec_batch_size", C.c_size_t),
                ("batch_accumulated", C.c_uint32), ("pool_current", C.c_int),
                ("flush_pending", C.c_int)]


class MemoryInfo(C.Structure):
    _fields_ = [(n, C.c_size_t) for n in (
        "device_bytes", "host_pinned_bytes", "staging_bytes", "chunk_pool_bytes",
        "compressed_pool_bytes", "aggregate_bytes", "lod_bytes", "codec_bytes",
        "shard_bytes")] + [("chunks_per_epoch", C.c_uint64), ("total_chunks", C.c_uint64),
                           ("max_output_size", C.c_size_t), ("nlod", C.c_int),
                           ("epochs_per_batch", C.c_uint32)]


class StreamOptions(C.Structure):
    _fields_ = [("pipeline", C.c_int), ("device", C.c_int), ("compress_workers", C.c_uint32),
                ("lz4_hash_log", C.c_uint32), ("zstd_block_bytes", C.c_uint32),
                ("host_gather_workgroups", C.c_uint32), ("zstd_chunk_tables", C.c_uint32),
                ("batched_d2h", C.c_uint32), ("batched_launch", C.c_uint32),
                ("lz4_segment_bytes", C.c_uint32), ("pass_timing", C.c_uint32),
                ("h2d_copy", C.c_uint32)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_P = C.c_void_p
_sig("tile_stream_gpu_create", _P, C.POINTER(TileStreamConfiguration), _P)
_sig("tile_stream_gpu_create_ex", _P, C.POINTER(TileStreamConfiguration), _P,
     C.POINTER(StreamOptions))
_sig("tile_stream_gpu_destroy", None, _P)
_sig("tile_stream_gpu_writer", C.POINTER(Writer), _P)
_sig("tile_stream_gpu_cursor", C.c_uint64, _P)
_sig("tile_stream_gpu_layout", C.POINTER(TileStreamLayout), _P)
_sig("tile_stream_gpu_status", TileStreamStatus, _P)
_sig("tile_stream_gpu_get_metrics", StreamMetrics, _P)
_sig("tile_stream_gpu_memory_estimate", C.c_int, C.POINTER(TileStreamConfiguration),
     C.POINTER(MemoryInfo))
_sig("tile_stream_gpu_memory_estimate_ex", C.c_int, C.POINTER(TileStreamConfiguration),
     C.POINTER(StreamOptions), C.POINTER(MemoryInfo))
_sig("tile_stream_gpu_advise_chunk_sizes", C.c_int, C.POINTER(TileStreamConfiguration),
     C.c_size_t, C.POINTER(C.c_uint8), C.c_size_t)
_sig("tile_stream_gpu_append_device", WriterResult, _P, _P, C.c_size_t)
_sig("tile_stream_gpu_reset", C.c_int, _P, _P)
_sig("tile_stream_gpu_epoch_latencies", C.c_size_t, _P, C.POINTER(C.c_double), C.c_size_t)
_sig("tile_stream_gpu_level_geometry", C.c_int, _P, C.c_int, C.POINTER(C.c_uint64),
     C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
_sig("tile_stream_gpu_options_default", None, C.POINTER(StreamOptions))
_sig("multiarray_tile_stream_gpu_create", _P, C.c_int, C.POINTER(TileStreamConfiguration),
     C.POINTER(C.c_void_p), C.c_int)
_sig("multiarray_tile_stream_gpu_destroy", None, _P)
_sig("multiarray_tile_stream_gpu_writer", C.POINTER(MultiarrayWriter), _P)
_sig("multiarray_tile_stream_gpu_get_metrics", StreamMetrics, _P)
_sig("writer_append", WriterResult, C.POINTER(Writer), Slice)
_sig("writer_flush", WriterResult, C.POINTER(Writer))
_sig("chucky_last_error", C.c_char_p)
_sig("chucky_mem_sink_create", _P)
_sig("chucky_mem_sink_destroy", None, _P)
_sig("chucky_mem_sink_as_shard_sink", _P, _P)
_sig("chucky_mem_sink_count", C.c_size_t, _P)
_sig("chucky_mem_sink_get", C.c_int, _P, C.c_size_t, C.POINTER(C.c_uint8),
     C.POINTER(C.c_uint64), C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t),
     C.POINTER(C.c_int))
_sig("chucky_mem_sink_append_sizes", C.c_int, _P, C.c_uint8, C.POINTER(C.c_uint64),
     C.POINTER(C.c_uint8))
_sig("chucky_mem_sink_inject_failure", None, _P, C.c_int, C.c_int)
_sig("chucky_mem_sink_set_deferred", None, _P, C.c_int)
_sig("chucky_mem_sink_drain", None, _P)
_sig("chucky_discard_sink_create", _P)
_sig("chucky_discard_sink_destroy", None, _P)
_sig("chucky_discard_sink_as_shard_sink", _P, _P)
_sig("chucky_discard_sink_bytes", C.c_uint64, _P)
_sig("chucky_discard_sink_finalized", C.c_uint64, _P)
_sig("chucky_gpu_max_output_size", C.c_size_t, C.c_int, C.c_size_t)
_sig("chucky_gpu_compress", C.c_int, C.c_int, C.c_int, _P, C.c_size_t, C.c_size_t,
     C.c_uint64, _P, C.c_size_t, _P)
_sig("chucky_gpu_compress_noalloc", C.c_int, C.c_int, C.c_int, _P, C.c_size_t, C.c_size_t,
     C.c_uint64, _P, C.c_size_t, _P, _P)
_sig("chucky_gpu_transpose", C.c_int, _P, _P, C.c_uint64, C.c_uint8, C.c_uint64, C.c_uint8,
     C.POINTER(C.c_uint64), C.POINTER(C.c_int64))
_sig("chucky_crc32c", C.c_uint32, _P, C.c_size_t)
_sig("dims_n_append", C.c_uint8, C.POINTER(Dimension), C.c_uint8)
_sig("chucky_tune_codec", C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint64, C.c_uint64,
     C.POINTER(C.c_uint32))
_sig("chucky_device_alloc", _P, C.c_size_t)
_sig("chucky_device_free", None, _P)
_sig("chucky_memcpy_to_device", C.c_int, _P, _P, C.c_size_t)
_sig("chucky_memcpy_to_host", C.c_int, _P, _P, C.c_size_t)
_sig("chucky_device_synchronize", C.c_int)
_sig("chucky_measure_link_gbps", C.c_double, C.c_size_t, C.c_int, C.c_int)
_sig("chucky_synth_sensor_frames", C.c_int, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64,
     C.c_double, C.POINTER(C.c_float), C.c_uint32)
_sig("chucky_synth_fill", C.c_int, _P, C.c_uint64, C.c_int, C.c_uint64, C.c_uint64)


class RequestReport(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("end_s", C.c_double), ("input_bytes", C.c_uint64),
                ("epochs", C.c_uint64), ("p50_latency_s", C.c_double),
                ("p99_latency_s", C.c_double), ("max_latency_s", C.c_double)]


_sig("chucky_server_create", _P, C.c_size_t, C.c_int, C.c_int)
_sig("chucky_server_submit", C.c_int, _P, C.POINTER(TileStreamConfiguration), _P, _P, C.c_size_t,
     C.c_int, C.POINTER(C.c_uint64))
_sig("chucky_server_wait_all", C.c_int, _P)
_sig("chucky_server_reports", C.c_size_t, _P, C.POINTER(RequestReport), C.c_size_t)
_sig("chucky_server_latencies", C.c_size_t, _P, C.POINTER(C.c_double), C.c_size_t)
_sig("chucky_server_peaks", None, _P, C.POINTER(C.c_int), C.POINTER(C.c_size_t))
_sig("chucky_server_destroy", None, _P)


class DeviceBuffer:
    """Device allocation owned by Python (hipMalloc through the library)."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        self.ptr = lib.chucky_device_alloc(nbytes)
        if not self.ptr:
            raise MemoryError(f"device allocation of {nbytes} bytes failed")

    @classmethod
    def from_host(cls, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        if lib.chucky_memcpy_to_device(b.ptr, a.ctypes.data, a.nbytes):
            raise RuntimeError("host->device copy failed")
        return b

    def to_host(self, nbytes=None):
        import numpy as np
        n = self.nbytes if nbytes is None else nbytes
        out = np.empty(n, dtype=np.uint8)
        if lib.chucky_memcpy_to_host(out.ctypes.data, self.ptr, n):
            raise RuntimeError("device->host copy failed")
        return out

    def free(self):
        if self.ptr:
            lib.chucky_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()

from .api import (Config, GpuStream, GpuMultiArray, MemSink, DiscardSink, last_error,  # noqa: E402
                  crc32c, memory_estimate, tuned_codec)

__all__ = ["lib", "PIPELINE_PLAIN", "PIPELINE_BATCHED", "DeviceBuffer", "Config", "GpuStream", "GpuMultiArray", "MemSink", "DiscardSink", "last_error", "crc32c",
           "memory_estimate", "tuned_codec", "DTYPES", "CODEC_ID", "REDUCE_ID"]
