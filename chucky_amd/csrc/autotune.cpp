// Launch-shape choices for the codec kernels.
//
// The compressors give one wavefront to each block of a chunk, and a
// persistent grid of `compressor_slots()` wavefronts pulls blocks from a
// device queue. A launch finishes after ceil(blocks / slots) rounds of the
// slowest block, so when a launch holds few chunks (a per-epoch launch of a
// small epoch, a short bounded stream) smaller blocks put more wavefronts to
// work; when it holds many, larger blocks cost less per byte (one Huffman
// table and one header per block) and compress better.
#include "stream_gpu.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace chucky {

uint32_t compressor_slots()
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // the zstd kernel's LDS footprint (~10.5 KiB) fits 15 wavefronts per CU
  return (uint32_t)cus * 15;
}

uint32_t tune_zstd_block(uint64_t chunk_bytes, uint64_t chunks_per_launch)
{
  const uint64_t slots = compressor_slots();
  // Largest block up to 32 KiB that still fills every slot once, else the
  // smallest block. Measured on 12-bit sensor tiles: 32 KiB blocks compress
  // faster than 128 KiB ones (more wavefronts in flight, shorter serial
  // Huffman steps per byte) and slightly better (more local tables).
  // 8 KiB is the floor: 4 KiB blocks (allowed when asked for) measured slower
  // on latency-bound launches too (twice the per-block Huffman work) and
  // compress 1% worse.
  constexpr uint32_t kPreferred = 32 * 1024, kFloor = 8 * 1024;
  for (uint32_t b = kPreferred; b > kFloor; b >>= 1) {
    const uint64_t blocks = chunks_per_launch * ((chunk_bytes + b - 1) / b);
    if (blocks >= slots)
      return b;
  }
  return kFloor;
}

uint64_t lz4_wave_slots(int hash_log)
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // one LZ4 wavefront holds a (4 << hash_log)-byte LDS table
  const uint64_t per_cu = std::max<uint64_t>(1, (160u * 1024u) / (4ull << hash_log));
  return (uint64_t)cus * std::min<uint64_t>(per_cu, 16);
}

uint32_t tune_lz4_segment(uint64_t chunk_bytes, uint64_t chunks_per_launch, int hash_log)
{
  const uint64_t slots = lz4_wave_slots(hash_log);
  if (chunks_per_launch >= slots)
    return 0;
  // 8 KiB measured fastest end to end on one-epoch launches of 512 KiB
  // chunks; 4 KiB segments parse faster but lose more to the stitch pass
  constexpr uint32_t kMinSeg = 8 * 1024;
  uint32_t seg = 0;
  for (uint64_t b = 1ull << 30; b >= kMinSeg; b >>= 1) {
    if (b >= chunk_bytes)
      continue;
    seg = (uint32_t)b;
    if (chunks_per_launch * ((chunk_bytes + b - 1) / b) >= slots)
      break;
  }
  return seg;
}

// Per-block Huffman tables adapt to local statistics; since the code
// construction runs wave-parallel (register queues, pointer-jumping depths)
// they beat one table per chunk in both speed and ratio on every measured
// workload, so the per-chunk tables are opt-in.
bool tune_zstd_chunk_tables(uint32_t block)
{
  (void)block;
  return false;
}

uint32_t tune_host_gather_workgroups()
{
  // PCIe-bound stores: a quarter of the CUs keeps the link busy
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return (uint32_t)(cus / 4 > 0 ? cus / 4 : 1);
}

} // namespace chucky
