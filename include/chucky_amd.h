/* Entry points of chucky-amd that have no counterpart in the reference.
 *
 * They exist for callers the reference API cannot serve on this hardware:
 * producers whose frames are already in device memory, benchmarks that reuse
 * one stream across acquisitions, the batched (graph-captured) pipeline, the
 * multi-stream serving loop, in-memory / discard sinks for tests and benches,
 * and kernel-level entry points used by parity tests and the autotuner. */
#pragma once

#include "stream.gpu.h"
#include "writer.h"

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stream options ------------------------------------------------------ */

enum chucky_pipeline
{
  CHUCKY_PIPELINE_PLAIN = 0,   /* one launch per stage per epoch, host syncs like the reference */
  CHUCKY_PIPELINE_BATCHED = 1, /* per-epoch hipGraph, persistent compressor, no host sync on the device path */
};

struct tile_stream_gpu_options
{
  int pipeline;              /* enum chucky_pipeline */
  int device;                /* HIP device ordinal; -1 = current device */
  uint32_t compress_workers; /* persistent compressor workgroups; 0 = autotuned */
  uint32_t lz4_hash_log;     /* LDS hash table log2 entries; 0 = autotuned */
  uint32_t zstd_block_bytes; /* zstd block size (<= 131072); 0 = autotuned */
  uint32_t host_gather_workgroups; /* CU budget of the gather into pinned host memory
                                      (batched pipeline); 0 = autotuned */
  uint32_t zstd_chunk_tables;      /* one Huffman table per chunk: 0 = autotuned, 1 = on, 2 = off */
  uint32_t batched_d2h;            /* batched pipeline: 0/2 = gather on the device + DMA copy,
                                      1 = gather kernel writing pinned host memory */
  uint32_t batched_launch;         /* batched pipeline per-epoch work: 0 = autotuned,
                                      1 = captured graphs, 2 = direct kernel launches */
  uint32_t lz4_segment_bytes;      /* LZ4 input parsed by one wavefront: a power of two
                                      >= 1024 below the chunk size, or >= the chunk size for
                                      one wavefront per chunk; 0 = autotuned */
  uint32_t pass_timing;            /* event timing of each scatter and LOD launch for the
                                      stream metrics: 0/1 = on, 2 = off (those launches then
                                      run back to back with no event markers between them;
                                      the compress and copy metrics are kept) */
  uint32_t h2d_copy;               /* host appends, staging -> device: 0/2 = the DMA engines
                                      (hipMemcpyAsync), 1 = a copy kernel reading the pinned
                                      staging buffer */
};

void tile_stream_gpu_options_default(struct tile_stream_gpu_options* o);

/* tile_stream_gpu_memory_estimate for the pipeline opts selects: the batched
 * pipeline's pinned host slots, their device staging and its per-slot
 * aggregate tables are included (device_bytes, aggregate_bytes,
 * host_pinned_bytes). NULL opts = the plain pipeline. 0 on success. */
int tile_stream_gpu_memory_estimate_ex(const struct tile_stream_configuration* config,
                                       const struct tile_stream_gpu_options* opts,
                                       struct tile_stream_memory_info* info);

struct tile_stream_gpu* tile_stream_gpu_create_ex(const struct tile_stream_configuration* config,
                                                  struct shard_sink* sink,
                                                  const struct tile_stream_gpu_options* opts);

/* Append nbytes that already live in device memory (no staging copy).
 * Same epoch / capacity / finished semantics as writer->append. */
struct writer_result tile_stream_gpu_append_device(struct tile_stream_gpu* s, const void* d_src,
                                                   size_t nbytes);

/* Re-arm a flushed stream for a new acquisition into `sink` (NULL = same sink),
 * keeping every allocation. 0 on success. */
int tile_stream_gpu_reset(struct tile_stream_gpu* s, struct shard_sink* sink);

/* Latency record of the most recent epochs: for each finished epoch, seconds
 * from the append call that completed it to the return of the last shard write
 * covering it. Returns the number of samples copied (<= cap). */
size_t tile_stream_gpu_epoch_latencies(const struct tile_stream_gpu* s, double* out, size_t cap);

/* Shard geometry of one level: chunks per shard (all dims), inner shards per
 * append step, chunks per epoch. 0 on success. */
int tile_stream_gpu_level_geometry(const struct tile_stream_gpu* s, int level,
                                   uint64_t* chunks_per_shard_total, uint64_t* shard_inner_count,
                                   uint64_t* chunks_per_epoch);

/* Last error message of this thread ("" if none). */
const char* chucky_last_error(void);

/* ---- serving loop: many concurrent streams on one GPU ---------------------
 * Requests (one acquisition each: a configuration, a sink and its input) are
 * admitted in arrival order while the number of running requests is below
 * max_concurrent and the sum of their device-memory estimates stays within
 * hbm_budget_bytes. Each admitted request runs on its own host thread and HIP
 * queues with the batched pipeline, so work of different requests overlaps on
 * the device; its partial-epoch and LOD state stay resident until it
 * finishes. Input is appended epoch by epoch (from host or device memory) to
 * model a live acquisition. */

struct chucky_server;

struct chucky_request_report
{
  uint64_t id;
  int status;               /* 0 ok, non-zero failure */
  double submit_s;          /* submit time, seconds since server creation */
  double start_s;           /* admission time */
  double end_s;             /* last shard written */
  uint64_t input_bytes;
  uint64_t epochs;
  double p50_latency_s;     /* append -> shard written, per epoch */
  double p99_latency_s;
  double max_latency_s;
};

struct chucky_server* chucky_server_create(size_t hbm_budget_bytes, int max_concurrent, int pipeline);
/* config, sink and data are borrowed until the request's report is out.
 * data_is_device: the input lives in device memory. 0 on success. */
int chucky_server_submit(struct chucky_server* srv, const struct tile_stream_configuration* config,
                         struct shard_sink* sink, const void* data, size_t nbytes,
                         int data_is_device, uint64_t* request_id);
/* Blocks until every submitted request finished. 0 if all succeeded. */
int chucky_server_wait_all(struct chucky_server* srv);
/* Copies reports of finished requests (in completion order); returns count. */
size_t chucky_server_reports(struct chucky_server* srv, struct chucky_request_report* out, size_t cap);
/* All per-epoch latency samples of finished requests (seconds). */
size_t chucky_server_latencies(struct chucky_server* srv, double* out, size_t cap);
/* Peak concurrently running requests and peak admitted device bytes. */
void chucky_server_peaks(struct chucky_server* srv, int* running, size_t* device_bytes);
void chucky_server_destroy(struct chucky_server* srv);

/* ---- in-memory and discard sinks ----------------------------------------- */

struct chucky_mem_sink;
struct chucky_mem_sink* chucky_mem_sink_create(void);
void chucky_mem_sink_destroy(struct chucky_mem_sink* s);
struct shard_sink* chucky_mem_sink_as_shard_sink(struct chucky_mem_sink* s);
/* Number of shards opened so far (all levels). */
size_t chucky_mem_sink_count(const struct chucky_mem_sink* s);
/* i-th shard in open order. data stays valid until the sink is destroyed. */
int chucky_mem_sink_get(const struct chucky_mem_sink* s, size_t i, uint8_t* level,
                        uint64_t* shard_index, const uint8_t** data, size_t* size,
                        int* finalized);
/* Last append extents published through update_append for a level (n values). */
int chucky_mem_sink_append_sizes(const struct chucky_mem_sink* s, uint8_t level,
                                 uint64_t* out, uint8_t* n);
/* Make open() fail (fail_open) or write() fail (fail_write) from now on. */
void chucky_mem_sink_inject_failure(struct chucky_mem_sink* s, int fail_open, int fail_write);
/* Deferred mode: zero-copy writes (write_direct) are held and their bytes
 * read from the caller's buffer only when the level's fence is waited for or
 * the sink is drained, as an asynchronous I/O queue reads them; fences are
 * counted per level. Exposes buffers reused before their fence retires. */
void chucky_mem_sink_set_deferred(struct chucky_mem_sink* s, int deferred);
/* Copies every held-back write (call before reading shards in deferred mode). */
void chucky_mem_sink_drain(struct chucky_mem_sink* s);

struct chucky_discard_sink;
struct chucky_discard_sink* chucky_discard_sink_create(void);
void chucky_discard_sink_destroy(struct chucky_discard_sink* s);
struct shard_sink* chucky_discard_sink_as_shard_sink(struct chucky_discard_sink* s);
uint64_t chucky_discard_sink_bytes(const struct chucky_discard_sink* s);
uint64_t chucky_discard_sink_finalized(const struct chucky_discard_sink* s);

/* The autotuner's codec launch shape for a stream (multiscale != 0: several
 * levels of detail) whose compress launches hold chunks_per_launch chunks of
 * chunk_bytes of a bpe-byte dtype (one epoch in the batched pipeline, a batch
 * of epochs in the plain one); the choice is the nearest row of tables
 * recorded by a sweep on MI355X. zstd: out[0] = block bytes, out[1] = 1 for
 * one Huffman table per chunk. lz4: out[0] = segment bytes parsed by one
 * wavefront (0 = whole chunks), out[1] = hash table log2 entries. 0 on
 * success. No device is touched. */
int chucky_tune_codec(enum compression_codec codec, int multiscale, uint32_t bpe, uint64_t chunk_bytes,
                      uint64_t chunks_per_launch, uint32_t out[2]);

/* ---- kernel-level entry points (device pointers, synchronous) ------------- */

/* Worst-case compressed bytes of one chunk for a GPU codec. */
size_t chucky_gpu_max_output_size(enum compression_codec codec, size_t chunk_bytes);

/* Compress n chunks: chunk i is chunk_bytes at d_src + i*src_stride, output at
 * d_dst + i*dst_stride (dst_stride >= chucky_gpu_max_output_size), sizes into
 * d_sizes[i] (uint64). The compressors read whole dwords and 16-byte pieces,
 * so d_src must stay readable for 64 bytes past the last chunk's end (the
 * bytes there do not change the output). 0 on success. */
int chucky_gpu_compress(enum compression_codec codec, int level, const void* d_src,
                        size_t src_stride, size_t chunk_bytes, uint64_t n, void* d_dst,
                        size_t dst_stride, uint64_t* d_sizes);

/* No counterpart in the reference: chucky_gpu_compress for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops). The
 * same inputs (with the same 64 bytes of read slack), outputs and compressed
 * bytes, with every other buffer in one
 * file-scope device array of the library (about 25 MiB a device, held until
 * the process ends): no hipMalloc, hipFree or stream creation. Asynchronous
 * on `stream` (a hipStream_t; NULL = the null stream): synchronize it before
 * reading d_dst / d_sizes. At most 65536 chunks x zstd blocks of 128 KiB per
 * call (lz4: 65536 chunks); larger calls fail with an error and launch
 * nothing. The first call on a device prepares that device's tables and must
 * not be under graph capture. Every call on a device shares the array, so
 * calls that overlap on one device race. */
int chucky_gpu_compress_noalloc(enum compression_codec codec, int level, const void* d_src,
                                size_t src_stride, size_t chunk_bytes, uint64_t n, void* d_dst,
                                size_t dst_stride, uint64_t* d_sizes, void* stream);

/* Scatter n elements (element index i_offset + i of the lifted row-major space)
 * into a chunk pool: dst[ravel(shape, strides, i_offset+i)] = src[i]. */
int chucky_gpu_transpose(void* d_dst, const void* d_src, uint64_t n, uint8_t bpe,
                         uint64_t i_offset, uint8_t lifted_rank, const uint64_t* lifted_shape,
                         const int64_t* lifted_strides);

/* CRC-32C (Castagnoli, reflected, init/xorout 0xFFFFFFFF) of host bytes. */
uint32_t chucky_crc32c(const void* data, size_t len);

/* ---- device memory helpers (current device, synchronous) ----------------- */

void* chucky_device_alloc(size_t bytes);
void chucky_device_free(void* p);
int chucky_memcpy_to_device(void* d_dst, const void* h_src, size_t n);
int chucky_memcpy_to_host(void* h_dst, const void* d_src, size_t n);
int chucky_device_synchronize(void);
/* DMA rate of `bytes` between pinned host memory (on the GPU's NUMA node)
 * and device memory of the current device: GB/s over `reps` back-to-back
 * copies after one untimed copy; to_device 1 = host to device. -1 on error. */
double chucky_measure_link_gbps(size_t bytes, int reps, int to_device);

/* Synthetic 12-bit sensor frames (u16, frames x height x width, row-major) in
 * device memory: per pixel a Poisson(lambda) background drawn by inversion of
 * the CDF from a counter-based generator keyed by (seed, pixel index), plus
 * Gaussian blobs. blobs (host) holds frames*nblob records of 4 floats
 * (y, x, sigma, amplitude); the sum is rounded and clipped to [0, 4095]. */
int chucky_synth_sensor_frames(void* d_dst, uint64_t frames, uint32_t height, uint32_t width,
                               uint64_t seed, double lambda, const float* blobs, uint32_t nblob);

/* Seeded synthetic fill of n elements of dt in device memory: integers
 * uniform over [0, int_range) (0 = every bit pattern), floats N(0,1) * 100. */
int chucky_synth_fill(void* d_dst, uint64_t n, enum dtype dt, uint64_t seed, uint64_t int_range);

#ifdef __cplusplus
}
#endif
