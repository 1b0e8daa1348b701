// Launch-shape choices for the codec kernels.
//
// The compressors give one wavefront to each block (zstd) or segment (LZ4)
// of a chunk. A launch finishes after its slowest rounds of blocks, so a
// launch that holds few chunks (a per-epoch launch of a small epoch, a short
// bounded stream) wants small blocks to put every CU to work, and one that
// holds many can afford larger blocks, which cost less per byte and compress
// better. LZ4 speed is also set by how many wavefronts fit a CU: each holds
// its hash table and its staged segment in LDS.
//
// The choices come from recorded sweeps (chucky_amd/tune.py, run on MI355X:
// `python -m chucky_amd.tune --codec lz4|zstd --out F`, then `--table F`):
// every dtype size x chunk size x chunks per launch below was streamed once
// per candidate over two synthetic data kinds (xor+noise volumes, Poisson
// sensor counts; 4096x2304 sensor frames in 8 MiB chunks), and the row keeps the candidate with the least time per
// input byte through the compressor plus its output's time over the 55 GB/s
// device-to-host link. Multiscale streams have rows of their own (every
// level in one launch; the coarse levels' smooth chunks parse into dense
// sequences and set the launch's tail). A stream takes the row nearest its
// own case (multiscale or not, then dtype size, then chunk bytes and chunks
// per launch on a log scale).
#include "chucky_amd.h"
#include "kernels.hpp"
#include "stream_gpu.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

namespace chucky {

namespace {

struct Lz4Row
{
  uint32_t lod, bpe, chunk_bytes, per_launch, seg_bytes, hash_log;
};

// multiscale (1) or single-level (0) stream, bpe, chunk bytes, chunks per
// launch -> segment bytes (0 = whole chunk), hash log; the comment is the
// sweep's mean rate and ratio for the choice
const Lz4Row kLz4Table[] = {
  { 0, 1, 32768, 16, 1024, 9 }, // 28.8 GB/s, 0.996
  { 0, 1, 32768, 64, 1024, 9 }, // 101.3 GB/s, 0.996
  { 0, 1, 32768, 256, 1024, 9 }, // 259.6 GB/s, 0.996
  { 0, 1, 262144, 16, 2048, 9 }, // 136.9 GB/s, 0.996
  { 0, 1, 262144, 64, 2048, 9 }, // 382.0 GB/s, 0.996
  { 0, 1, 262144, 256, 4096, 9 }, // 659.7 GB/s, 0.996
  { 0, 2, 65536, 16, 1024, 9 }, // 35.3 GB/s, 1.012
  { 0, 2, 65536, 64, 1024, 9 }, // 107.5 GB/s, 1.012
  { 0, 2, 65536, 256, 1024, 9 }, // 199.5 GB/s, 1.012
  { 0, 2, 524288, 16, 2048, 9 }, // 136.0 GB/s, 1.013
  { 0, 2, 524288, 64, 2048, 9 }, // 266.4 GB/s, 1.013
  { 0, 2, 524288, 256, 2048, 9 }, // 326.3 GB/s, 1.013
  { 0, 2, 8388608, 144, 2048, 9 }, // 161.6 GB/s, 1.029
  { 0, 2, 8388608, 288, 1024, 9 }, // 169.2 GB/s, 1.025
  { 0, 4, 131072, 16, 1024, 9 }, // 41.9 GB/s, 1.072
  { 0, 4, 131072, 64, 1024, 9 }, // 92.3 GB/s, 1.072
  { 0, 4, 131072, 256, 1024, 9 }, // 125.8 GB/s, 1.073
  { 0, 4, 1048576, 16, 1024, 9 }, // 118.9 GB/s, 1.061
  { 0, 4, 1048576, 64, 1024, 9 }, // 166.2 GB/s, 1.061
  { 0, 4, 1048576, 256, 1024, 9 }, // 172.1 GB/s, 1.061
  { 1, 2, 1024, 1365, 0, 9 }, // 21.9 GB/s, 1.168
  { 1, 2, 1024, 5460, 0, 9 }, // 71.9 GB/s, 1.168
  { 1, 2, 65536, 85, 1024, 9 }, // 76.4 GB/s, 1.180
  { 1, 2, 65536, 340, 1024, 9 }, // 162.1 GB/s, 1.180
  { 1, 2, 524288, 21, 1024, 9 }, // 111.1 GB/s, 1.167
  { 1, 2, 524288, 84, 1024, 9 }, // 207.7 GB/s, 1.167
};

struct ZstdRow
{
  uint32_t lod, bpe, chunk_bytes, per_launch, block, chunk_tables;
};

// multiscale or not, bpe, chunk bytes, chunks per launch -> block bytes, one
// Huffman table per chunk
const ZstdRow kZstdTable[] = {
  { 0, 1, 32768, 16, 8192, 0 }, // 6.3 GB/s, 1.470
  { 0, 1, 32768, 64, 8192, 0 }, // 23.4 GB/s, 1.470
  { 0, 1, 32768, 256, 8192, 0 }, // 73.2 GB/s, 1.470
  { 0, 1, 262144, 16, 8192, 0 }, // 41.7 GB/s, 1.333
  { 0, 1, 262144, 64, 16384, 0 }, // 112.3 GB/s, 1.339
  { 0, 1, 262144, 256, 32768, 0 }, // 247.7 GB/s, 1.342
  { 0, 2, 65536, 16, 8192, 0 }, // 10.9 GB/s, 1.804
  { 0, 2, 65536, 64, 8192, 0 }, // 38.5 GB/s, 1.803
  { 0, 2, 65536, 256, 8192, 0 }, // 99.6 GB/s, 1.803
  { 0, 2, 524288, 16, 8192, 0 }, // 64.5 GB/s, 1.665
  { 0, 2, 524288, 64, 16384, 0 }, // 158.9 GB/s, 1.675
  { 0, 2, 524288, 256, 65536, 0 }, // 274.3 GB/s, 1.682
  { 0, 2, 8388608, 144, 65536, 0 }, // 411.6 GB/s, 2.154
  { 0, 2, 8388608, 288, 131072, 0 }, // 450.7 GB/s, 2.155
  { 0, 4, 131072, 16, 8192, 0 }, // 8.2 GB/s, 2.329
  { 0, 4, 131072, 64, 8192, 0 }, // 29.3 GB/s, 2.309
  { 0, 4, 131072, 256, 8192, 1 }, // 55.6 GB/s, 2.336
  { 0, 4, 1048576, 16, 8192, 0 }, // 55.7 GB/s, 2.185
  { 0, 4, 1048576, 64, 8192, 0 }, // 67.8 GB/s, 2.184
  { 0, 4, 1048576, 256, 8192, 1 }, // 99.4 GB/s, 2.211
  { 1, 2, 1024, 1365, 8192, 0 }, // 11.8 GB/s, 2.116
  { 1, 2, 1024, 5460, 16384, 0 }, // 20.8 GB/s, 2.116
  { 1, 2, 65536, 85, 8192, 0 }, // 11.8 GB/s, 2.131
  { 1, 2, 65536, 340, 8192, 0 }, // 38.3 GB/s, 2.131
  { 1, 2, 524288, 21, 8192, 0 }, // 25.0 GB/s, 1.993
  { 1, 2, 524288, 84, 8192, 0 }, // 64.8 GB/s, 1.993
};

template<typename Row, size_t N>
const Row& nearest(const Row (&table)[N], bool lod, uint32_t bpe, uint64_t chunk_bytes, uint64_t per_launch)
{
  const auto dist = [&](const Row& r) {
    const double dm = (r.lod != 0) != lod ? 16.0 : 0.0;
    const double db = std::fabs(std::log2((double)std::max<uint32_t>(bpe, 1) / r.bpe));
    const double dc = std::fabs(std::log2((double)std::max<uint64_t>(chunk_bytes, 1) / r.chunk_bytes));
    const double dl = std::fabs(std::log2((double)std::max<uint64_t>(per_launch, 1) / r.per_launch));
    return dm + 4 * db + dc + dl;
  };
  size_t best = 0;
  for (size_t i = 1; i < N; ++i)
    if (dist(table[i]) < dist(table[best]))
      best = i;
  return table[best];
}

} // namespace

uint32_t compressor_slots()
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // the zstd kernel's LDS footprint (~10.5 KiB) fits 15 wavefronts per CU
  return (uint32_t)cus * 15;
}

uint32_t tune_zstd_block(uint64_t chunk_bytes, uint64_t chunks_per_launch, uint32_t bpe, bool lod)
{
  return nearest(kZstdTable, lod, bpe, chunk_bytes, chunks_per_launch).block;
}

bool tune_zstd_chunk_tables(uint64_t chunk_bytes, uint64_t chunks_per_launch, uint32_t bpe, bool lod)
{
  return nearest(kZstdTable, lod, bpe, chunk_bytes, chunks_per_launch).chunk_tables != 0;
}

Lz4Choice tune_lz4(uint64_t chunk_bytes, uint64_t chunks_per_launch, uint32_t bpe, bool lod)
{
  const Lz4Row& r = nearest(kLz4Table, lod, bpe, chunk_bytes, chunks_per_launch);
  return Lz4Choice{ r.seg_bytes && r.seg_bytes < chunk_bytes ? r.seg_bytes : 0u, (int)r.hash_log };
}

uint64_t lz4_wave_slots(int hash_log, uint32_t seg_bytes)
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint64_t per_cu = std::max<uint64_t>(1, (160u * 1024u) / lz4_lds_bytes(hash_log, seg_bytes));
  return (uint64_t)cus * std::min<uint64_t>(per_cu, 16);
}

uint32_t tune_host_gather_workgroups()
{
  // PCIe-bound stores: a quarter of the CUs keeps the link busy
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return (uint32_t)(cus / 4 > 0 ? cus / 4 : 1);
}

uint32_t tune_h2d_copy_workgroups()
{
  // PCIe-bound loads from pinned memory, measured on MI355X for 2 and 8 MiB
  // pieces: 32 workgroups of 256 lanes reach 50 and 55 GB/s; 16 fall short
  // on 8 MiB (48 GB/s) and 512 or more lose to contention (39-52 GB/s)
  return 32;
}

} // namespace chucky

extern "C" int chucky_tune_codec(enum compression_codec codec, int multiscale, uint32_t bpe,
                                 uint64_t chunk_bytes, uint64_t chunks_per_launch, uint32_t out[2])
{
  if (!out || bpe == 0 || chunk_bytes == 0 || chunks_per_launch == 0)
    return 1;
  const bool lod = multiscale != 0;
  if (codec == CODEC_ZSTD) {
    out[0] = chucky::tune_zstd_block(chunk_bytes, chunks_per_launch, bpe, lod);
    out[1] = chucky::tune_zstd_chunk_tables(chunk_bytes, chunks_per_launch, bpe, lod) ? 1u : 0u;
    return 0;
  }
  if (codec == CODEC_LZ4_NON_STANDARD) {
    const chucky::Lz4Choice c = chucky::tune_lz4(chunk_bytes, chunks_per_launch, bpe, lod);
    out[0] = c.seg_bytes;
    out[1] = (uint32_t)c.hash_log;
    return 0;
  }
  return 1;
}
