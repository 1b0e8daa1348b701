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
  // largest block that still fills every slot once; else the smallest block
  for (uint32_t b = kZstdBlockMax; b > kZstdBlockMin; b >>= 1) {
    const uint64_t blocks = chunks_per_launch * ((chunk_bytes + b - 1) / b);
    if (blocks >= slots)
      return b;
  }
  return kZstdBlockMin;
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
