// Host-to-device copy of a staged piece by a kernel that reads the pinned
// staging buffer over the host link.
//
// On MI355X a few dozen workgroups reading 16 bytes per lane from pinned host
// memory reach the link's rate (55 GB/s for 8 MiB with 32 workgroups, the
// DMA engines' is 52-57 GB/s), and the copy is then an ordinary kernel: the
// launches that wait for it (the LOD gather, the chunk scatter) are queued on
// the device behind an event of a kernel, whereas a launch queued behind an
// in-flight DMA copy holds the launching host thread until the copy ends.
// The reads bypass the caches (non-temporal): every byte is read once.
#include "kernels.hpp"

namespace chucky {
namespace {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

// dst and src share their address mod 16 (checked by the launcher): bytes
// before the first 16-byte boundary and after the last are copied by the
// first workgroup, the 16-byte vectors between them by the whole grid.
__global__ __launch_bounds__(kThreads) void h2d_copy_k(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src,
                                                        uint64_t n, uint64_t head, uint64_t n16)
{
  const v4u* s = (const v4u*)(src + head);
  v4u* d = (v4u*)(dst + head);
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n16; i += stride)
    d[i] = __builtin_nontemporal_load(s + i);
  if (blockIdx.x == 0) {
    const uint64_t tail = head + n16 * 16;
    if (threadIdx.x < head)
      dst[threadIdx.x] = src[threadIdx.x];
    if (threadIdx.x < n - tail)
      dst[tail + threadIdx.x] = src[tail + threadIdx.x];
  }
}

} // namespace

bool h2d_copy_ok(const void* dst, const void* src) { return (((uintptr_t)dst ^ (uintptr_t)src) & 15) == 0; }

hipError_t launch_h2d_copy(void* dst, const void* src, uint64_t n, uint32_t groups, hipStream_t stream)
{
  if (n == 0)
    return hipSuccess;
  if (!h2d_copy_ok(dst, src) || groups == 0)
    return hipErrorInvalidValue;
  const uint64_t head = std::min<uint64_t>(n, (16 - ((uintptr_t)dst & 15)) & 15);
  const uint64_t n16 = (n - head) / 16;
  // fewer than 16 bytes remain after the vectors, copied by lanes of group 0
  const uint64_t need = (n16 + kThreads - 1) / kThreads;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(groups, need));
  hipLaunchKernelGGL(h2d_copy_k, dim3(grid), dim3(kThreads), 0, stream, (uint8_t*)dst, (const uint8_t*)src, n, head,
                     n16);
  return hipGetLastError();
}

} // namespace chucky
