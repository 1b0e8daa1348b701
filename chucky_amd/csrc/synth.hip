// Device memory helpers and the synthetic sensor-frame generator used by
// benchmarks and tests (so that no second GPU runtime is needed in-process).
#include "chucky_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

namespace chucky {
void set_error(const std::string& msg);
hipError_t host_alloc_unpooled(void** p, size_t bytes, int device);
}

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr int kCdfLen = 512;

// cdf[k] = P(X <= k) in double; binary search of a uniform in (0, 1) gives an
// exact Poisson draw up to the table end.
__device__ __forceinline__ void sensor_pixel(uint16_t* __restrict__ out, uint64_t i,
                                             uint64_t per_frame, uint32_t W, uint64_t seed,
                                             const double* __restrict__ cdf, int cdf_len,
                                             const float4* __restrict__ blobs, uint32_t nblob)
{
  const uint64_t f = i / per_frame;
  const uint32_t r = (uint32_t)(i - f * per_frame);
  const float y = (float)(r / W), x = (float)(r % W);
  const double u = ((double)(mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  int lo = 0, hi = cdf_len - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] < u)
      lo = mid + 1;
    else
      hi = mid;
  }
  float v = (float)lo;
  const float4* b = blobs + f * nblob;
  for (uint32_t k = 0; k < nblob; ++k) {
    const float4 p = b[k];
    const float dy = y - p.x, dx = x - p.y, s = p.z;
    const float r2 = dy * dy + dx * dx;
    if (r2 < 36.f * s * s)
      v += p.w * expf(-r2 / (2.f * s * s));
  }
  v = rintf(v);
  v = v < 0.f ? 0.f : (v > 4095.f ? 4095.f : v);
  out[i] = (uint16_t)v;
}

// Grid-stride: a dispatch counts work-items in 32 bits, so the grid is capped
// on the host and each lane walks the range.
__global__ __launch_bounds__(256) void sensor_k(uint16_t* __restrict__ out, uint64_t frames,
                                                uint32_t H, uint32_t W, uint64_t seed,
                                                const double* __restrict__ cdf, int cdf_len,
                                                const float4* __restrict__ blobs, uint32_t nblob)
{
  const uint64_t per_frame = (uint64_t)H * W;
  const uint64_t total = frames * per_frame;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * 256)
    sensor_pixel(out, i, per_frame, W, seed, cdf, cdf_len, blobs, nblob);
}

// Seeded fill: integers uniform over [0, range) (range 0 = all bit patterns of
// the type), floats N(0, 1) * 100 by the Box-Muller transform of two uniforms.
template<typename T>
__global__ __launch_bounds__(256) void fill_k(T* __restrict__ out, uint64_t n, uint64_t seed,
                                              uint64_t range)
{
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    const uint64_t r = mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
    if constexpr (std::is_floating_point<T>::value || std::is_same<T, _Float16>::value) {
      const uint64_t r2 = mix64(r ^ 0xD1B54A32D192ED03ull);
      const float u1 = ((float)(r >> 40) + 0.5f) * (1.f / 16777216.f);
      const float u2 = (float)(r2 >> 40) * (1.f / 16777216.f);
      const float z = sqrtf(-2.f * logf(u1)) * cospif(2.f * u2);
      out[i] = (T)(z * 100.f);
    } else {
      out[i] = (T)(range ? r % range : r);
    }
  }
}

template<typename T>
int fill_t(void* d, uint64_t n, uint64_t seed, uint64_t range)
{
  hipLaunchKernelGGL(fill_k<T>, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 1u << 20)),
                     dim3(256), 0, nullptr, (T*)d, n, seed, range);
  return hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

} // namespace

extern "C" {

void* chucky_device_alloc(size_t bytes)
{
  void* p = nullptr;
  if (hipMalloc(&p, bytes + 256) != hipSuccess) {
    chucky::set_error("chucky_device_alloc failed");
    return nullptr;
  }
  return p;
}

void chucky_device_free(void* p) { (void)hipFree(p); }

// Copies run on a non-blocking queue of the calling thread: legacy-stream
// copies are refused while another thread captures a graph.
static int copy_sync(void* dst, const void* src, size_t n, hipMemcpyKind kind)
{
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return 1;
  const bool ok = hipMemcpyAsync(dst, src, n, kind, st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
  (void)hipStreamDestroy(st);
  return ok ? 0 : 1;
}

int chucky_memcpy_to_device(void* d, const void* h, size_t n)
{
  return copy_sync(d, h, n, hipMemcpyHostToDevice);
}

int chucky_memcpy_to_host(void* h, const void* d, size_t n)
{
  return copy_sync(h, d, n, hipMemcpyDeviceToHost);
}

int chucky_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 1; }

double chucky_measure_link_gbps(size_t bytes, int reps, int to_device)
{
  void* h = nullptr;
  void* d = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  double gbps = -1;
  float ms = 0;
  int dev = 0;
  if (reps < 1 || hipGetDevice(&dev) != hipSuccess || chucky::host_alloc_unpooled(&h, bytes, dev) != hipSuccess || hipMalloc(&d, bytes) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&a) != hipSuccess ||
      hipEventCreate(&b) != hipSuccess)
    goto done;
  std::memset(h, 1, bytes);
  // one untimed copy, then `reps` timed back to back
  for (int r = 0; r <= reps; ++r) {
    if (r == 1 && hipEventRecord(a, st) != hipSuccess)
      goto done;
    if (hipMemcpyAsync(to_device ? d : h, to_device ? h : d, bytes,
                       to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, st) != hipSuccess)
      goto done;
  }
  if (hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
      hipEventElapsedTime(&ms, a, b) != hipSuccess || ms <= 0)
    goto done;
  gbps = (double)bytes * reps / (ms * 1e-3) / 1e9;
done:
  if (a)
    (void)hipEventDestroy(a);
  if (b)
    (void)hipEventDestroy(b);
  if (st)
    (void)hipStreamDestroy(st);
  if (d)
    (void)hipFree(d);
  if (h)
    (void)hipHostFree(h);
  if (gbps < 0)
    chucky::set_error("chucky_measure_link_gbps failed");
  return gbps;
}

int chucky_synth_sensor_frames(void* d_dst, uint64_t frames, uint32_t height, uint32_t width,
                               uint64_t seed, double lambda, const float* blobs, uint32_t nblob)
{
  // Poisson CDF table (log-space terms to avoid underflow of exp(-lambda)).
  std::vector<double> cdf(kCdfLen);
  double acc = 0;
  for (int k = 0; k < kCdfLen; ++k) {
    acc += std::exp(-lambda + k * std::log(lambda) - std::lgamma((double)k + 1));
    cdf[k] = acc;
  }
  cdf[kCdfLen - 1] = 2.0; // catch-all
  double* d_cdf = nullptr;
  float4* d_blobs = nullptr;
  int rc = 1;
  const uint64_t n = frames * (uint64_t)height * width;
  if (hipMalloc((void**)&d_cdf, kCdfLen * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&d_blobs, std::max<uint64_t>(1, frames * nblob) * sizeof(float4)) != hipSuccess)
    goto done;
  if (hipMemcpy(d_cdf, cdf.data(), kCdfLen * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    goto done;
  if (nblob && hipMemcpy(d_blobs, blobs, frames * nblob * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess)
    goto done;
  hipLaunchKernelGGL(sensor_k, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 1u << 20)),
                     dim3(256), 0, nullptr,
                     (uint16_t*)d_dst, frames, height, width, seed, d_cdf, kCdfLen, d_blobs, nblob);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    goto done;
  rc = 0;
done:
  (void)hipFree(d_cdf);
  (void)hipFree(d_blobs);
  return rc;
}

int chucky_synth_fill(void* d_dst, uint64_t n, enum dtype dt, uint64_t seed, uint64_t int_range)
{
  switch (dt) {
    case dtype_u8: return fill_t<uint8_t>(d_dst, n, seed, int_range);
    case dtype_u16: return fill_t<uint16_t>(d_dst, n, seed, int_range);
    case dtype_u32: return fill_t<uint32_t>(d_dst, n, seed, int_range);
    case dtype_u64: return fill_t<uint64_t>(d_dst, n, seed, int_range);
    case dtype_i8: return fill_t<int8_t>(d_dst, n, seed, int_range);
    case dtype_i16: return fill_t<int16_t>(d_dst, n, seed, int_range);
    case dtype_i32: return fill_t<int32_t>(d_dst, n, seed, int_range);
    case dtype_i64: return fill_t<int64_t>(d_dst, n, seed, int_range);
    case dtype_f16: return fill_t<_Float16>(d_dst, n, seed, int_range);
    case dtype_f32: return fill_t<float>(d_dst, n, seed, int_range);
    case dtype_f64: return fill_t<double>(d_dst, n, seed, int_range);
  }
  chucky::set_error("chucky_synth_fill: unknown dtype");
  return 1;
}

} // extern "C"
