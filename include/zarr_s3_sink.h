/* Zarr v3 S3 sink interface (reference: src/zarr_s3_sink.h:10-109).
 * Configuration structs, defaults and validation behave as in the reference
 * (part-count limits, key rules). This build has no S3 transport (no network
 * client library in the image), so the create functions log an error and
 * return NULL; any other shard_sink implementation plugs into the stream. */
#pragma once

#include "dimension.h"
#include "dtype.h"
#include "types.codec.h"
#include "writer.h"

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct zarr_s3_config
{
  const char* bucket;
  const char* prefix;
  const char* array_name;
  const char* region;
  const char* endpoint;
  enum dtype data_type;
  double fill_value;
  uint8_t rank;
  const struct dimension* dimensions;
  struct codec_config codec;
  size_t part_size;          /* 0 = 8 MiB */
  double throughput_gbps;    /* 0 = 10 */
  size_t max_retries;
  uint32_t backoff_scale_ms;
  uint32_t max_backoff_secs;
  uint64_t timeout_ns;
};

struct zarr_s3_sink;

/* src/zarr_s3_sink.h:32-33 */
void zarr_s3_config_set_defaults(struct zarr_s3_config* cfg);
/* src/zarr_s3_sink.h:37-38: 0 when valid */
int zarr_s3_config_validate(const struct zarr_s3_config* cfg);
/* src/zarr_s3_sink.h:43-44: NULL in this build (no S3 transport) */
struct zarr_s3_sink* zarr_s3_sink_create(struct zarr_s3_config* cfg);
/* src/zarr_s3_sink.h:47-48 */
int zarr_s3_sink_has_error(const struct zarr_s3_sink* s);
/* src/zarr_s3_sink.h:51-52 */
int zarr_s3_sink_destroy(struct zarr_s3_sink* s);
/* src/zarr_s3_sink.h:56-57 */
int zarr_s3_sink_flush(struct zarr_s3_sink* s);
/* src/zarr_s3_sink.h:61-62 */
struct shard_sink* zarr_s3_sink_as_shard_sink(struct zarr_s3_sink* s);

struct zarr_s3_multiscale_config
{
  const char* bucket;
  const char* prefix;
  const char* array_name;
  const char* region;
  const char* endpoint;
  enum dtype data_type;
  double fill_value;
  uint8_t rank;
  const struct dimension* dimensions;
  int nlod;
  struct codec_config codec;
  size_t part_size;
  double throughput_gbps;
  size_t max_retries;
  uint32_t backoff_scale_ms;
  uint32_t max_backoff_secs;
  uint64_t timeout_ns;
};

struct zarr_s3_multiscale_sink;

void zarr_s3_multiscale_config_set_defaults(struct zarr_s3_multiscale_config* cfg);
int zarr_s3_multiscale_config_validate(const struct zarr_s3_multiscale_config* cfg);
struct zarr_s3_multiscale_sink* zarr_s3_multiscale_sink_create(struct zarr_s3_multiscale_config* cfg);
int zarr_s3_multiscale_sink_destroy(struct zarr_s3_multiscale_sink* s);
int zarr_s3_multiscale_sink_flush(struct zarr_s3_multiscale_sink* s);
struct shard_sink* zarr_s3_multiscale_sink_as_shard_sink(struct zarr_s3_multiscale_sink* s);

#ifdef __cplusplus
}
#endif
