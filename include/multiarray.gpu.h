/* Multi-array writer over GPU tile streams (reference: src/multiarray.cpu.h:6-46,
 * behaviour of src/multiarray/stream.c:552-726,866-887). One writer feeds several
 * arrays; the caller names the array on every update. The active array may
 * change only at an epoch boundary of the departing array. Every array keeps
 * its own GPU stream state resident in HBM, so a switch moves no data. */
#pragma once

#include "types.stream.h"
#include "writer.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/multiarray.cpu.h:6-12 */
enum multiarray_writer_error
{
  multiarray_writer_ok = 0,
  multiarray_writer_fail = 1,
  multiarray_writer_finished = 2,
  multiarray_writer_not_flushable = 3, /* switch requested mid-epoch */
};

/* src/multiarray.cpu.h:14-18 */
struct multiarray_writer_result
{
  int error;
  struct slice rest;
};

/* src/multiarray.cpu.h:20-26 */
struct multiarray_writer
{
  struct multiarray_writer_result (*update)(struct multiarray_writer* self, int array_index,
                                            struct slice data);
  struct multiarray_writer_result (*flush)(struct multiarray_writer* self);
};

struct multiarray_tile_stream_gpu;

/* Counterpart of multiarray_tile_stream_cpu_create (src/multiarray.cpu.h:31-36).
 * configs[i] and sinks[i] describe array i; each array gets its own GPU tile
 * stream (plain pipeline). NULL on any failure. Metrics are always collected;
 * enable_metrics is accepted for interface parity. */
struct multiarray_tile_stream_gpu* multiarray_tile_stream_gpu_create(
  int n_arrays, const struct tile_stream_configuration configs[], struct shard_sink* sinks[],
  int enable_metrics);

/* src/multiarray.cpu.h:38-39 */
void multiarray_tile_stream_gpu_destroy(struct multiarray_tile_stream_gpu* ms);

/* src/multiarray.cpu.h:41-42 */
struct multiarray_writer* multiarray_tile_stream_gpu_writer(struct multiarray_tile_stream_gpu* ms);

/* src/multiarray.cpu.h:44-46: the per-stage metrics summed over the arrays. */
struct stream_metrics multiarray_tile_stream_gpu_get_metrics(
  const struct multiarray_tile_stream_gpu* ms);

#ifdef __cplusplus
}
#endif
