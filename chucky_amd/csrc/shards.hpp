// Shard assembly on the host: per-level open shards, the (offset, nbytes)
// index, CRC-32C trailer and the writes into a shard_sink.
// Reproduces src/zarr/shard_delivery.c:10-204 byte for byte.
#pragma once

#include "plan.hpp"
#include "writer.h"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace chucky {

struct OpenShard
{
  uint64_t data_cursor = 0;
  std::vector<uint64_t> index; // 2 * cps_total, 0xFF.. = empty slot
  shard_writer* writer = nullptr;
};

struct ShardState
{
  uint64_t epoch_in_shard = 0;
  uint64_t shard_epoch = 0;
  uint64_t shard_inner_count = 1;
  uint64_t cps_inner = 1, cps_total = 1, cps_append = 1;
  std::vector<OpenShard> shards;

  void init(const LevelInfo& L);
  void reset();
};

// View of one batch's aggregated bytes for one level:
//   data + offsets[i] .. offsets[i+1] is slot i in [shard][epoch][chunk] order,
//   chunk_sizes[i] is the unpadded compressed size (0 = no chunk).
struct AggView
{
  const uint8_t* data;
  const uint64_t* offsets; // n_active * C + 1
  const uint64_t* chunk_sizes;
};

// Write index + CRC and finalize every open shard, then advance shard_epoch.
int finalize_shards(ShardState& ss, size_t shard_alignment);

// Deliver n_active epochs of one level. *out_bytes = bytes handed to writers.
int deliver_batch(uint8_t level, ShardState& ss, const AggView& v, uint32_t n_active,
                  shard_sink* sink, size_t shard_alignment, size_t* out_bytes);

} // namespace chucky
