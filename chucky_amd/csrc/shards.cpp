#include "shards.hpp"

#include "chucky_amd.h"

#include <cstdlib>
#include <cstring>
#include <string>

namespace chucky {

void set_error(const std::string& msg);

void ShardState::init(const LevelInfo& L)
{
  epoch_in_shard = 0;
  shard_epoch = 0;
  shard_inner_count = L.shard_inner_count;
  cps_inner = L.cps_inner;
  cps_total = L.cps_total;
  cps_append = L.cps_append;
  shards.assign(shard_inner_count, OpenShard{});
  for (auto& s : shards)
    s.index.assign(2 * cps_total, ~0ull);
}

void ShardState::reset()
{
  epoch_in_shard = 0;
  shard_epoch = 0;
  for (auto& s : shards) {
    s.data_cursor = 0;
    s.writer = nullptr;
    std::fill(s.index.begin(), s.index.end(), ~0ull);
  }
}

// src/zarr/shard_delivery.c:35-102. Layout written at data_cursor:
//   [zero pad to alignment] [index: cps_total x (u64 offset, u64 nbytes) LE] [crc32c u32 LE]
int finalize_shards(ShardState& ss, size_t sa)
{
  int err = 0;
  for (auto& sh : ss.shards) {
    if (!sh.writer)
      continue;
    const size_t index_bytes = ss.cps_total * 16;
    const size_t total = index_bytes + 4;
    const size_t write_bytes = sa ? align_up(total, sa) : total;
    const size_t pad = write_bytes - total;
    void* raw = nullptr;
    if (sa) {
      if (posix_memalign(&raw, sa < sizeof(void*) ? sizeof(void*) : sa, write_bytes))
        raw = nullptr;
    } else {
      raw = std::malloc(write_bytes);
    }
    uint8_t* buf = (uint8_t*)raw;
    if (!buf) {
      set_error("finalize_shards: index allocation failed");
      err = 1;
    } else {
      std::memset(buf, 0, pad);
      std::memcpy(buf + pad, sh.index.data(), index_bytes);
      const uint32_t crc = chucky_crc32c(buf + pad, index_bytes);
      std::memcpy(buf + pad + index_bytes, &crc, 4);
      if (sh.writer->write(sh.writer, sh.data_cursor, buf, buf + write_bytes)) {
        set_error("finalize_shards: index write failed");
        err = 1;
      }
      std::free(buf);
    }
    if (sh.writer->finalize(sh.writer)) {
      set_error("finalize_shards: finalize failed");
      err = 1;
    }
    sh.writer = nullptr;
    sh.data_cursor = 0;
    std::fill(sh.index.begin(), sh.index.end(), ~0ull);
  }
  ss.epoch_in_shard = 0;
  ss.shard_epoch++;
  return err;
}

// src/zarr/shard_delivery.c:104-204. Epochs are delivered in runs that stay
// inside one shard; each run is one write per inner shard.
int deliver_batch(uint8_t level, ShardState& ss, const AggView& v, uint32_t n_active,
                  shard_sink* sink, size_t sa, size_t* out_bytes)
{
  const uint64_t cpsi = ss.cps_inner;
  size_t total = 0;
  uint32_t a = 0;
  while (a < n_active) {
    const uint64_t left_in_shard = ss.cps_append - ss.epoch_in_shard;
    const uint32_t run = (uint32_t)std::min<uint64_t>(left_in_shard, n_active - a);
    for (uint64_t si = 0; si < ss.shard_inner_count; ++si) {
      OpenShard& sh = ss.shards[si];
      if (!sh.writer) {
        sh.writer = sink->open(sink, level, ss.shard_epoch * ss.shard_inner_count + si);
        if (!sh.writer) {
          set_error("shard sink open failed");
          return 1;
        }
      }
      const uint64_t j0 = si * n_active * cpsi + (uint64_t)a * cpsi;
      const uint64_t j1 = j0 + (uint64_t)run * cpsi;
      const uint64_t run_bytes = v.offsets[j1] - v.offsets[j0];
      if (run_bytes > 0) {
        const uint8_t* src = v.data + v.offsets[j0];
        const size_t wb = sa ? align_up(run_bytes, sa) : run_bytes;
        total += wb;
        const bool aligned = sa == 0 || ((uintptr_t)src % sa) == 0;
        int rc;
        if (aligned && sh.writer->write_direct)
          rc = sh.writer->write_direct(sh.writer, sh.data_cursor, src, src + wb);
        else
          rc = sh.writer->write(sh.writer, sh.data_cursor, src, src + wb);
        if (rc) {
          set_error("shard write failed");
          return 1;
        }
      }
      for (uint32_t r = 0; r < run; ++r) {
        const uint64_t eis = ss.epoch_in_shard + r;
        const uint64_t js = j0 + (uint64_t)r * cpsi;
        for (uint64_t j = js; j < js + cpsi; ++j) {
          const uint64_t cs = v.chunk_sizes[j];
          if (cs > 0) {
            const uint64_t slot = eis * cpsi + (j - js);
            sh.index[2 * slot] = sh.data_cursor + (v.offsets[j] - v.offsets[j0]);
            sh.index[2 * slot + 1] = cs;
          }
        }
      }
      sh.data_cursor += sa ? align_up(run_bytes, sa) : run_bytes;
    }
    ss.epoch_in_shard += run;
    a += run;
    if (ss.epoch_in_shard >= ss.cps_append)
      if (finalize_shards(ss, sa))
        return 1;
  }
  if (out_bytes)
    *out_bytes = total;
  return 0;
}

} // namespace chucky
