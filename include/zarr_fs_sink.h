/* Zarr v3 filesystem sink (reference: src/zarr_fs_sink.h:10-80).
 * Same structures, functions and on-disk layout: {store}/{array}/zarr.json,
 * shard files at c/<i>/<j>/..., sharding_indexed codec with a crc32c index at
 * the end; the multiscale form writes one array per level ("0", "1", ...) plus
 * OME-NGFF v0.5 group metadata. Shard bytes are written by an I/O worker
 * thread; write_direct keeps the caller's buffer until the next fence. */
#pragma once

#include "dimension.h"
#include "dtype.h"
#include "types.codec.h"
#include "writer.h"

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct zarr_config
{
  const char* store_path;
  const char* array_name; /* e.g. "0"; NULL = write at store_path */
  enum dtype data_type;
  double fill_value;
  uint8_t rank;
  const struct dimension* dimensions;
  int unbuffered; /* O_DIRECT */
  struct codec_config codec;
};

struct zarr_fs_sink;

/* src/zarr_fs_sink.h:25-26 */
struct zarr_fs_sink* zarr_fs_sink_create(const struct zarr_config* cfg);
/* src/zarr_fs_sink.h:28-29 */
void zarr_fs_sink_destroy(struct zarr_fs_sink* s);
/* src/zarr_fs_sink.h:32-33: blocks until queued I/O is done */
void zarr_fs_sink_flush(struct zarr_fs_sink* s);
/* src/zarr_fs_sink.h:36-37 */
size_t zarr_fs_sink_pending_bytes(struct zarr_fs_sink* s);
/* src/zarr_fs_sink.h:40-41 */
struct shard_sink* zarr_fs_sink_as_shard_sink(struct zarr_fs_sink* s);

struct zarr_multiscale_config
{
  const char* store_path;
  const char* array_name; /* group name; NULL = store_path */
  enum dtype data_type;
  double fill_value;
  uint8_t rank;
  const struct dimension* dimensions; /* L0 */
  int nlod;                           /* 0 = auto */
  int unbuffered;
  struct codec_config codec;
};

struct zarr_fs_multiscale_sink;

/* src/zarr_fs_sink.h:63-64 */
struct zarr_fs_multiscale_sink* zarr_fs_multiscale_sink_create(const struct zarr_multiscale_config* cfg);
/* src/zarr_fs_sink.h:66-67 */
void zarr_fs_multiscale_sink_destroy(struct zarr_fs_multiscale_sink* s);
/* src/zarr_fs_sink.h:69-70 */
void zarr_fs_multiscale_sink_flush(struct zarr_fs_multiscale_sink* s);
/* src/zarr_fs_sink.h:73-74 */
size_t zarr_fs_multiscale_sink_pending_bytes(struct zarr_fs_multiscale_sink* s);
/* src/zarr_fs_sink.h:77-78 */
struct shard_sink* zarr_fs_multiscale_sink_as_shard_sink(struct zarr_fs_multiscale_sink* s);

#ifdef __cplusplus
}
#endif
