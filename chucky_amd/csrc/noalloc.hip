// chucky_gpu_compress_noalloc: the batch chunk compressor of
// chucky_gpu_compress with every buffer beyond the caller's inputs and
// outputs in one file-scope device array, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).
//
// Layout of `scratch` (per device, shared by every call on that device):
//   [0, kPreBytes)          zstd predefined FSE tables (written once)
//   [kQueueOff, +256)       the zstd work counter (zero between launches)
//   [kSegOff, +4*kMaxSeg)   compressed size of every segment of the call
//   [kWorkOff, +...)        zstd worker scratch for kWorkers wavefronts
// Calls that overlap on one device race on it.
#include "chucky_amd.h"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

namespace chucky {
void set_error(const std::string& msg);
}

using namespace chucky;

namespace {

constexpr uint32_t kWorkers = 64;            // persistent zstd wavefronts per call
constexpr uint64_t kMaxSegments = 1u << 16;  // chunks x blocks per call
constexpr size_t kPreBytes = 64 * 1024;      // >= zstd_predefined_bytes(), checked at setup
constexpr size_t kQueueOff = kPreBytes;
constexpr size_t kSegOff = kQueueOff + 256;
constexpr size_t kWorkOff = kSegOff + kMaxSegments * 4;
constexpr size_t kWorkBytes = (size_t)kWorkers * 393728; // zstd_scratch_bytes(kWorkers), checked at setup
constexpr size_t kScratchBytes = kWorkOff + kWorkBytes;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};

__global__ void zero_k(uint32_t* p, uint32_t n)
{
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    p[i] = 0;
}

// The device address of `scratch` on the current device, found (and the
// zstd tables and work counter prepared) at the first call on each device.
uint8_t* scratch_on_device(hipStream_t st)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("chucky_gpu_compress_noalloc: no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[dev])
    return g_ptr[dev];
  if (zstd_predefined_bytes() > kPreBytes || zstd_scratch_bytes(kWorkers) > kWorkBytes) {
    set_error("chucky_gpu_compress_noalloc: scratch layout too small for this build");
    return nullptr;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    set_error("chucky_gpu_compress_noalloc: the first call on a device prepares its tables and "
              "must run outside graph capture");
    return nullptr;
  }
  uint8_t* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("chucky_gpu_compress_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    return nullptr;
  }
  if (zstd_build_predefined(ptr, st) != hipSuccess) {
    set_error("chucky_gpu_compress_noalloc: table setup launch failed");
    return nullptr;
  }
  hipLaunchKernelGGL(zero_k, dim3(1), dim3(64), 0, st, (uint32_t*)(ptr + kQueueOff), 64u);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    set_error("chucky_gpu_compress_noalloc: table setup failed");
    return nullptr;
  }
  g_ptr[dev] = ptr;
  return ptr;
}

} // namespace


namespace chucky {

namespace {
// Moves segments 1.. of every chunk's zstd frame down behind segment 0 and
// writes the frame's size. Destinations lie below their sources, and each
// round reads all of its bytes before writing any, so the moves never
// overwrite bytes still to be read.
__global__ __launch_bounds__(256) void compact_k(uint8_t* dst, uint64_t dst_stride, const uint32_t* seg,
                                                 uint32_t nseg, uint64_t seg_stride, uint64_t* sizes)
{
  const uint64_t c = blockIdx.x;
  uint8_t* base = dst + c * dst_stride;
  const uint32_t* s = seg + c * nseg;
  uint64_t off = s[0];
  for (uint32_t k = 1; k < nseg; ++k) {
    const uint32_t len = s[k];
    const uint8_t* from = base + (uint64_t)k * seg_stride;
    uint8_t* to = base + off;
    for (uint32_t r = 0; r < len; r += 256 * 16) {
      uint8_t v[16];
      const uint32_t b = r + threadIdx.x * 16;
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (b + j < len)
          v[j] = from[b + j];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (b + j < len)
          to[b + j] = v[j];
      __syncthreads();
    }
    off += len;
  }
  if (threadIdx.x == 0)
    sizes[c] = off;
}

__global__ __launch_bounds__(256) void widen_k(const uint32_t* seg, uint64_t* sizes, uint64_t n)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    sizes[i] = seg[i];
}

} // namespace

hipError_t launch_frame_sizes(uint8_t* dst, uint64_t dst_stride, const uint32_t* seg, uint32_t nseg,
                              uint64_t seg_stride, uint64_t n, uint64_t* sizes, hipStream_t stream)
{
  if (n == 0)
    return hipSuccess;
  if (nseg == 1)
    hipLaunchKernelGGL(widen_k, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, seg, sizes, n);
  else
    hipLaunchKernelGGL(compact_k, dim3((uint32_t)n), dim3(256), 0, stream, dst, dst_stride, seg, nseg,
                       seg_stride, sizes);
  return hipGetLastError();
}

} // namespace chucky

extern "C" {

int chucky_gpu_compress_noalloc(enum compression_codec codec, int level, const void* d_src,
                                size_t src_stride, size_t chunk_bytes, uint64_t n, void* d_dst,
                                size_t dst_stride, uint64_t* d_sizes, void* stream)
{
  hipStream_t st = (hipStream_t)stream;
  if (n == 0)
    return 0;
  if (codec != CODEC_LZ4_NON_STANDARD && codec != CODEC_ZSTD) {
    set_error("chucky_gpu_compress_noalloc: codec must be lz4 or zstd");
    return 1;
  }
  if (dst_stride < chucky_gpu_max_output_size(codec, chunk_bytes)) {
    set_error("chucky_gpu_compress_noalloc: dst_stride below the codec bound");
    return 1;
  }
  const uint32_t nseg = codec == CODEC_ZSTD ? zstd_segments(chunk_bytes, kZstdBlockMax) : 1;
  if (n > kMaxSegments / nseg) {
    set_error("chucky_gpu_compress_noalloc: " + std::to_string(n) + " chunks of " +
              std::to_string(nseg) + " segments exceed the limit of " + std::to_string(kMaxSegments) +
              " segments per call");
    return 1;
  }
  uint8_t* ptr = scratch_on_device(st);
  if (!ptr)
    return 1;
  uint32_t* seg = (uint32_t*)(ptr + kSegOff);
  hipError_t e;
  if (codec == CODEC_LZ4_NON_STANDARD) {
    e = launch_lz4((const uint8_t*)d_src, src_stride, chunk_bytes, n, (uint8_t*)d_dst, dst_stride, seg, 12, 0,
                   nullptr, st);
    if (e == hipSuccess)
      e = launch_frame_sizes((uint8_t*)d_dst, dst_stride, seg, 1, 0, n, d_sizes, st);
  } else {
    const uint32_t workers = (uint32_t)std::min<uint64_t>(n * nseg, kWorkers);
    e = launch_zstd_pre(ptr, (const uint8_t*)d_src, src_stride, chunk_bytes, n, (uint8_t*)d_dst, dst_stride,
                        seg, level, kZstdBlockMax, ptr + kWorkOff, workers, (uint32_t*)(ptr + kQueueOff),
                        nullptr, st);
    if (e == hipSuccess)
      e = launch_frame_sizes((uint8_t*)d_dst, dst_stride, seg, nseg, zstd_segment_stride(kZstdBlockMax), n,
                             d_sizes, st);
  }
  if (e != hipSuccess) {
    set_error(std::string("chucky_gpu_compress_noalloc: launch failed: ") + hipGetErrorString(e));
    return 1;
  }
  return 0;
}

} // extern "C"
