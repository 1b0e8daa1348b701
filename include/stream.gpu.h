/* GPU streaming Zarr v3 writer: the drop-in C API.
 *
 * Every declaration replaces the one at the cited line of the reference's
 * src/stream.gpu.h with the same name, argument meaning and error behaviour.
 * No HIP or CUDA type appears here; device pointers are plain pointers. */
#pragma once

#include "types.stream.h"
#include "writer.h"

#ifdef __cplusplus
extern "C" {
#endif

struct tile_stream_gpu;

/* src/stream.gpu.h:9-13 — device copies of the L0 lifted shape/strides. */
struct tile_stream_layout_gpu
{
  uint64_t* d_lifted_shape;
  int64_t* d_lifted_strides;
};

/* src/stream.gpu.h:15-37 */
struct tile_stream_memory_info
{
  size_t device_bytes;
  size_t host_pinned_bytes;
  size_t staging_bytes;
  size_t chunk_pool_bytes;
  size_t compressed_pool_bytes;
  size_t aggregate_bytes;
  size_t lod_bytes;
  size_t codec_bytes;
  size_t shard_bytes;
  uint64_t chunks_per_epoch;
  uint64_t total_chunks;
  size_t max_output_size;
  int nlod;
  uint32_t epochs_per_batch;
};

/* src/stream.gpu.h:41-44. 0 on success, non-zero on an invalid config. */
int tile_stream_gpu_memory_estimate(const struct tile_stream_configuration* config,
                                    struct tile_stream_memory_info* info);

/* src/stream.gpu.h:49-53. Halves target_chunk_bytes until the estimate fits
 * budget_bytes; rewrites config->dimensions[].chunk_size. 0 on success. */
int tile_stream_gpu_advise_chunk_sizes(struct tile_stream_configuration* config,
                                       size_t target_chunk_bytes, const uint8_t* ratios,
                                       size_t budget_bytes);

/* src/stream.gpu.h:59-61. NULL on any failure. config is copied;
 * config->dimensions and sink are borrowed for the stream's lifetime. */
struct tile_stream_gpu* tile_stream_gpu_create(const struct tile_stream_configuration* config,
                                               struct shard_sink* sink);

/* src/stream.gpu.h:63-64. Waits for all queued device work, then frees. */
void tile_stream_gpu_destroy(struct tile_stream_gpu* stream);

/* src/stream.gpu.h:67-68 */
struct stream_metrics tile_stream_gpu_get_metrics(const struct tile_stream_gpu* s);

/* src/stream.gpu.h:72-73 */
const struct tile_stream_layout* tile_stream_gpu_layout(const struct tile_stream_gpu* s);

/* src/stream.gpu.h:75-76. append() takes host memory; flush() is terminal. */
struct writer* tile_stream_gpu_writer(struct tile_stream_gpu* s);

/* src/stream.gpu.h:78-79. Elements consumed so far. */
uint64_t tile_stream_gpu_cursor(const struct tile_stream_gpu* s);

/* src/stream.gpu.h:81-82 */
struct tile_stream_status tile_stream_gpu_status(const struct tile_stream_gpu* s);

#ifdef __cplusplus
}
#endif
