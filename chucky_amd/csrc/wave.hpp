// Wavefront (64-lane) helpers shared by the codec kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace chucky {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// 4 bytes at any address, from two aligned dword loads funnel-shifted (no
// branch on the alignment, so lanes never diverge). Reads up to 7 bytes past
// p: every codec input buffer carries >= 64 bytes of slack.
__device__ __forceinline__ uint32_t ld32u(const uint8_t* p)
{
  // stepping back from p (rather than masking its integer value) keeps the
  // pointer's address space known, so LDS input compiles to ds_read and
  // global input to global_load instead of flat loads
  const uint32_t mis = (uint32_t)(uintptr_t)p & 3u;
  const uint32_t* w = (const uint32_t*)__builtin_assume_aligned(p - mis, 4);
  return __builtin_amdgcn_alignbit(w[1], w[0], mis * 8);
}

// 8 bytes at any address, from three aligned dword loads (same slack rule).
__device__ __forceinline__ uint64_t ld64u(const uint8_t* p)
{
  const uint32_t mis = (uint32_t)(uintptr_t)p & 3u;
  const uint32_t* w = (const uint32_t*)__builtin_assume_aligned(p - mis, 4);
  const uint32_t sh = mis * 8;
  const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
  const uint32_t lo = __builtin_amdgcn_alignbit(w1, w0, sh);
  const uint32_t hi = __builtin_amdgcn_alignbit(w2, w1, sh);
  return ((uint64_t)hi << 32) | lo;
}

// Wave-cooperative copy of n bytes (any alignment): byte head to a dword
// boundary of dst, dword body (funnel-shifted source), byte tail.
__device__ __forceinline__ void wave_copy(uint8_t* dst, const uint8_t* src, uint32_t n)
{
  const int lane = lane_id();
  const uint32_t head = (uint32_t)((4 - ((uintptr_t)dst & 3)) & 3);
  const uint32_t h = head < n ? head : n;
  if ((uint32_t)lane < h)
    dst[lane] = src[lane];
  const uint32_t body = (n - h) >> 2;
  uint32_t* dw = (uint32_t*)(dst + h);
  const uint8_t* s = src + h;
  for (uint32_t w = lane; w < body; w += kWave)
    dw[w] = ld32u(s + 4 * w);
  const uint32_t t0 = h + body * 4;
  if ((uint32_t)lane < n - t0)
    dst[t0 + lane] = src[t0 + lane];
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Orders the LDS accesses of one wavefront across lanes (a wavefront issues
// its LDS operations in order), for single-wavefront algorithms that may run
// inside a larger workgroup.
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of lane `src_lane` (a wave-uniform index) for every lane, as a scalar:
// code that follows it stays uniform (scalar registers, no exec-mask flow).
template<typename T>
__device__ __forceinline__ T wave_bcast(T v, int src_lane)
{
  static_assert(sizeof(T) == 4, "32-bit values");
  return (T)__builtin_amdgcn_readlane((int)v, src_lane);
}

// A value every lane holds alike, made visibly uniform to the compiler.
__device__ __forceinline__ uint32_t uniform(uint32_t v)
{
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Cross-lane steps by DPP (data-parallel primitives: a VALU operand read
// from another lane of the same row of 16 or broadcast across rows), a few
// cycles each, where a shuffle is an LDS-unit permute with its full latency.
// Source lanes outside the row read as 0 (bound_ctrl), lanes outside
// row_mask keep their value. The reductions and scans below need every lane
// of the wave active (they are called from converged code only).
template<int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_or0(uint32_t v)
{
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false);
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

// Inclusive prefix sum over the wave: Hillis-Steele within each row of 16
// lanes, then row 0's total into row 1 and row 2's into row 3, then the
// first half's total into the second half.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
  v += dpp_or0<kDppRowShr1>(v);
  v += dpp_or0<kDppRowShr2>(v);
  v += dpp_or0<kDppRowShr4>(v);
  v += dpp_or0<kDppRowShr8>(v);
  v += dpp_or0<kDppRowBcast15, 0xA>(v);
  v += dpp_or0<kDppRowBcast31, 0xC>(v);
  return v;
}

// Sum over the wave, as a wave-uniform (scalar) value.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), kWave - 1);
}

// Maximum over the wave, as a wave-uniform value (the same running scheme
// with max in place of +; 0 is its identity for unsigned values).
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
  uint32_t u;
  u = dpp_or0<kDppRowShr1>(v);
  v = u > v ? u : v;
  u = dpp_or0<kDppRowShr2>(v);
  v = u > v ? u : v;
  u = dpp_or0<kDppRowShr4>(v);
  v = u > v ? u : v;
  u = dpp_or0<kDppRowShr8>(v);
  v = u > v ? u : v;
  u = dpp_or0<kDppRowBcast15, 0xA>(v);
  v = u > v ? u : v;
  u = dpp_or0<kDppRowBcast31, 0xC>(v);
  v = u > v ? u : v;
  return (uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1);
}

// A 512-entry array spread over the wave's registers: entry i lives in lane
// i % 64, register i / 64. Reads and writes take a wave-uniform index, so a
// sequential algorithm over the array (one that lane 0 would run out of LDS)
// runs out of registers at a few cycles per access.
struct Dist512
{
  uint32_t r[8];
};

__device__ __forceinline__ uint32_t dist_get(const Dist512& a, uint32_t i)
{
  uint32_t v;
  switch (i >> 6) {
    case 0: v = a.r[0]; break;
    case 1: v = a.r[1]; break;
    case 2: v = a.r[2]; break;
    case 3: v = a.r[3]; break;
    case 4: v = a.r[4]; break;
    case 5: v = a.r[5]; break;
    case 6: v = a.r[6]; break;
    default: v = a.r[7]; break;
  }
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)(i & 63));
}

__device__ __forceinline__ void dist_set(Dist512& a, uint32_t i, uint32_t v)
{
  if ((uint32_t)lane_id() != (i & 63))
    return;
  switch (i >> 6) {
    case 0: a.r[0] = v; break;
    case 1: a.r[1] = v; break;
    case 2: a.r[2] = v; break;
    case 3: a.r[3] = v; break;
    case 4: a.r[4] = v; break;
    case 5: a.r[5] = v; break;
    case 6: a.r[6] = v; break;
    default: a.r[7] = v; break;
  }
}

// Number of leading equal bytes (0..4) of two little-endian dwords.
__device__ __forceinline__ uint32_t eq_bytes(uint32_t a, uint32_t b)
{
  const uint32_t x = a ^ b;
  return x ? (uint32_t)(__builtin_ctz(x) >> 3) : 4u;
}

// Extend a match forward: bytes equal at p+len.. and c+len.. up to limit
// (exclusive end of p's range). Returns the final length. Uniform result.
__device__ __forceinline__ uint32_t wave_match_extend(const uint8_t* base, uint32_t p, uint32_t c,
                                                      uint32_t len, uint32_t limit)
{
  const int lane = lane_id();
  for (;;) {
    const uint32_t q = p + len + 4u * lane;
    uint32_t nb = 0;
    if (q < limit) {
      nb = eq_bytes(ld32u(base + q), ld32u(base + c + len + 4u * lane));
      if (nb > limit - q)
        nb = limit - q;
    }
    const uint64_t short_mask = wave_ballot(nb < 4);
    if (short_mask == 0) {
      len += 4u * kWave;
      continue;
    }
    const int first = __ffsll((unsigned long long)short_mask) - 1;
    const uint32_t nb_first = wave_bcast(nb, first);
    return len + 4u * (uint32_t)first + nb_first;
  }
}

} // namespace chucky
