// Device memory helpers and the synthetic sensor-frame generator used by
// benchmarks and tests (so that no second GPU runtime is needed in-process).
#include "chucky_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

namespace chucky {
void set_error(const std::string& msg);
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

int chucky_memcpy_to_device(void* d, const void* h, size_t n)
{
  return hipMemcpy(d, h, n, hipMemcpyHostToDevice) == hipSuccess ? 0 : 1;
}

int chucky_memcpy_to_host(void* h, const void* d, size_t n)
{
  return hipMemcpy(h, d, n, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 1;
}

int chucky_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 1; }

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

} // extern "C"
