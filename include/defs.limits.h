/* Compile-time limits. Values equal the reference's (src/defs.limits.h:5-18)
 * because they size fixed arrays inside ABI structs (tile_stream_layout). */
#pragma once

#define MAX_RANK 64
#define HALF_MAX_RANK (MAX_RANK / 2)
#define LOD_MAX_NDIM HALF_MAX_RANK
#define LOD_MAX_LEVELS 32
#define MAX_BATCH_EPOCHS 128
#define MAX_ZARR_RANK (HALF_MAX_RANK)
#define ZARR_GROUP_JSON_MAX_LENGTH 8192
#define S3_MAX_PARTS 10000
#define S3_DEFAULT_PART_SIZE (8 * 1024 * 1024)
#define S3_DEFAULT_THROUGHPUT_GBPS 10.0
