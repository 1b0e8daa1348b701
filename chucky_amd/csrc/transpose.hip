// Scatter of row-major input into chunk-major order.
//
// Semantics of the reference's transpose (src/cpu/transpose.cpp:43-105,
// src/gpu/transpose.cu:84-137): element i of the stream goes to
//   dst[ sum_d coord_d(i) * lifted_strides[d] ],  coords = unravel(i, lifted_shape)
// where the lifted shape interleaves (chunk index, within-chunk index) per
// dimension and append-dimension chunk strides are zero.
//
// gfx950 design: when the innermost lifted coordinate has unit stride and the
// innermost chunk extent is a multiple of 16 bytes, every aligned 16-byte
// vector of input lands contiguously in the output, so each lane moves one
// 128-bit vector with one 128-bit load and one 128-bit store (both coalesced
// across the wavefront for chunk rows >= 64 bytes). Index math uses 32-bit
// multiply-high division by precomputed magic numbers; unit dimensions are
// dropped on the host. Everything else takes the per-element path.
#include "kernels.hpp"

namespace chucky {
namespace {

struct FastDiv32
{
  uint32_t d, m, s; // q = (umulhi(n, m) + n) >> s, exact for n < 2^31 ... see make()
};

// Round-up magic number (Granlund-Montgomery): for d >= 1 and n < 2^32,
// t = umulhi(n, m); q = (t + ((n - t) >> 1)) >> (l - 1) with l = ceil(log2 d).
FastDiv32 make_fastdiv(uint32_t d)
{
  FastDiv32 f{ d, 0, 0 };
  if (d == 1)
    return f;
  uint32_t l = 0;
  while ((1ull << l) < d)
    ++l;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  f.s = l;
  return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv32& f)
{
  if (f.m == 0)
    return n;
  const uint32_t t = __umulhi(n, f.m);
  return (t + ((n - t) >> 1)) >> (f.s - 1);
}

constexpr int kMaxDims = 64;

struct Space32
{
  int rank;
  FastDiv32 div[kMaxDims];
  uint64_t stride[kMaxDims];
};

struct Space64
{
  int rank;
  uint64_t shape[kMaxDims];
  int64_t stride[kMaxDims];
};

__device__ __forceinline__ uint64_t ravel32(const Space32& sp, uint32_t idx)
{
  uint64_t o = 0;
  uint32_t rest = idx;
  for (int d = sp.rank - 1; d >= 0; --d) {
    const uint32_t q = fdiv(rest, sp.div[d]);
    o += (uint64_t)(rest - q * sp.div[d].d) * sp.stride[d];
    rest = q;
  }
  return o;
}

__device__ __forceinline__ uint64_t ravel64(const Space64& sp, uint64_t idx)
{
  uint64_t o = 0, rest = idx;
  for (int d = sp.rank - 1; d >= 0; --d) {
    o += (rest % sp.shape[d]) * (uint64_t)sp.stride[d];
    rest /= sp.shape[d];
  }
  return o;
}

// One 16-byte vector per lane. n_vec vectors starting at element i_offset.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template<int V>
__global__ __launch_bounds__(256) void transpose_vec_k(u32x4* __restrict__ dst_vec,
                                                       const u32x4* __restrict__ src,
                                                       uint64_t n_vec, uint32_t i_offset,
                                                       Space32 sp)
{
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_vec)
    return;
  const u32x4 x = __builtin_nontemporal_load(src + v);
  const uint64_t o = ravel32(sp, i_offset + (uint32_t)(v * V)); // element offset, multiple of V
  dst_vec[o / V] = x;
}

template<typename T>
__global__ __launch_bounds__(256) void transpose_elem32_k(T* __restrict__ dst,
                                                          const T* __restrict__ src, uint64_t n,
                                                          uint32_t i_offset, Space32 sp)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  dst[ravel32(sp, i_offset + (uint32_t)i)] = src[i];
}

template<typename T>
__global__ __launch_bounds__(256) void transpose_elem64_k(T* __restrict__ dst,
                                                          const T* __restrict__ src, uint64_t n,
                                                          uint64_t i_offset, Space64 sp)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  dst[ravel64(sp, i_offset + i)] = src[i];
}

template<typename T>
hipError_t elem_launch(void* dst, const void* src, uint64_t n, uint64_t off, bool fits32,
                       const Space32& s32, const Space64& s64, hipStream_t st)
{
  // A dispatch counts work-items in 32 bits: launch in pieces of <= 2^31.
  const uint64_t piece = 1ull << 31;
  for (uint64_t done = 0; done < n; done += piece) {
    const uint64_t m = n - done < piece ? n - done : piece;
    const uint64_t blocks = (m + 255) / 256;
    if (fits32)
      hipLaunchKernelGGL(transpose_elem32_k<T>, dim3((uint32_t)blocks), dim3(256), 0, st, (T*)dst,
                         (const T*)src + done, m, (uint32_t)(off + done), s32);
    else
      hipLaunchKernelGGL(transpose_elem64_k<T>, dim3((uint32_t)blocks), dim3(256), 0, st, (T*)dst,
                         (const T*)src + done, m, off + done, s64);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

} // namespace

hipError_t launch_transpose(void* dst, const void* src, uint64_t n, int bpe, uint64_t i_offset,
                            const LiftedSpace& space, hipStream_t stream)
{
  if (n == 0)
    return hipSuccess;
  // Drop unit dims, and leading dims of stride 0 (the append dimension's
  // chunk index: its coordinate never moves the destination, so only the
  // index modulo the product of the remaining shapes -- one epoch -- matters,
  // which keeps long streams on the 32-bit index path). The period of the
  // lifted space is the product of the kept shapes.
  Space32 s32{};
  Space64 s64{};
  uint64_t period = 1;
  int r = 0;
  bool leading = true;
  for (int d = 0; d < space.rank; ++d) {
    if (space.shape[d] == 1)
      continue;
    if (leading && space.strides[d] == 0)
      continue;
    leading = false;
    s64.shape[r] = space.shape[d];
    s64.stride[r] = space.strides[d];
    ++r;
    period *= space.shape[d];
  }
  if (r == 0) { // a single element position
    s64.shape[0] = 1;
    s64.stride[0] = 0;
    r = 1;
  }
  s64.rank = r;
  const uint64_t off = i_offset % period;
  const bool fits32 = period <= 0xFFFFFFFFull && off + n <= 0xFFFFFFFFull;
  if (fits32) {
    s32.rank = r;
    for (int d = 0; d < r; ++d) {
      s32.div[d] = make_fastdiv((uint32_t)s64.shape[d]);
      s32.stride[d] = (uint64_t)s64.stride[d];
    }
  }
  const int V = 16 / bpe;
  const bool unit_inner = s64.stride[r - 1] == 1 && s64.shape[r - 1] % V == 0;
  const bool aligned = ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 16 == 0);
  if (fits32 && unit_inner && aligned && off % V == 0) {
    const uint64_t n_vec = n / V; // < 2^32 / V work-items: one dispatch
    if (n_vec) {
      const uint64_t blocks = (n_vec + 255) / 256;
      switch (bpe) {
        case 1:
          hipLaunchKernelGGL(transpose_vec_k<16>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                             (u32x4*)dst, (const u32x4*)src, n_vec, (uint32_t)off, s32);
          break;
        case 2:
          hipLaunchKernelGGL(transpose_vec_k<8>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                             (u32x4*)dst, (const u32x4*)src, n_vec, (uint32_t)off, s32);
          break;
        case 4:
          hipLaunchKernelGGL(transpose_vec_k<4>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                             (u32x4*)dst, (const u32x4*)src, n_vec, (uint32_t)off, s32);
          break;
        default:
          hipLaunchKernelGGL(transpose_vec_k<2>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                             (u32x4*)dst, (const u32x4*)src, n_vec, (uint32_t)off, s32);
          break;
      }
      hipError_t e = hipGetLastError();
      if (e != hipSuccess)
        return e;
    }
    const uint64_t done = n_vec * V;
    if (done == n)
      return hipSuccess;
    src = (const uint8_t*)src + done * bpe;
    n -= done;
    i_offset = off + done;
  } else {
    i_offset = off;
  }
  switch (bpe) {
    case 1: return elem_launch<uint8_t>(dst, src, n, i_offset, fits32, s32, s64, stream);
    case 2: return elem_launch<uint16_t>(dst, src, n, i_offset, fits32, s32, s64, stream);
    case 4: return elem_launch<uint32_t>(dst, src, n, i_offset, fits32, s32, s64, stream);
    case 8: return elem_launch<uint64_t>(dst, src, n, i_offset, fits32, s32, s64, stream);
  }
  return hipErrorInvalidValue;
}

} // namespace chucky
