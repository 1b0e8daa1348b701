/* Array dimension description and builders.
 * struct dimension has the reference's layout (src/dimension.h:17-38); the
 * builder functions replace src/dimension.c:39-250 with the same semantics. */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dimension_axis_type
{
  dimension_axis_space = 0,
  dimension_axis_time,
  dimension_axis_channel,
  dimension_axis_other,
};

struct omengff_axis
{
  const char* unit; /* NULL -> no unit in metadata */
  double scale;     /* physical pixel scale, 0 treated as 1.0 */
  enum dimension_axis_type type;
};

struct dimension
{
  uint64_t size;             /* 0 = unbounded (dim 0 only) */
  uint64_t chunk_size;
  uint64_t chunks_per_shard; /* 0 = all chunks along this dim */
  const char* name;
  int downsample;            /* part of the LOD pyramid */
  uint8_t storage_position;  /* 0 = outermost; dims[0] must be 0 */
  struct omengff_axis ngff;
};

/* One dim per character of names; chunk = size, identity storage order. */
uint8_t dims_create(struct dimension* dims, const char* names, const uint64_t* sizes);
int dims_set_storage_order(struct dimension* dims, uint8_t rank, const char* order);
void dims_set_downsample_by_name(struct dimension* dims, uint8_t rank, const char* names);
void dims_set_chunk_sizes(struct dimension* dims, uint8_t rank, const uint64_t* chunk_sizes);
/* Greedy power-of-two split of nelem over dims by ratio (ties to higher index). */
void dims_budget_chunk_size(struct dimension* dims, uint8_t rank, uint64_t nelem,
                            const uint8_t* ratios);
void dims_set_shard_counts(struct dimension* dims, uint8_t rank, const uint64_t* shard_counts);
void dims_budget_chunk_bytes(struct dimension* dims, uint8_t rank, size_t target_chunk_bytes,
                             size_t bytes_per_element, const uint8_t* ratios);
/* Leftmost run of chunk_size==1 dims, cut after the first downsampled one; >= 1. */
uint8_t dims_n_append(const struct dimension* dims, uint8_t rank);
int dims_validate(const struct dimension* dims, uint8_t rank);
void dims_print(const struct dimension* dims, uint8_t rank);

#ifdef __cplusplus
}
#endif
