// Shard aggregation on the device.
//
// Reference semantics (src/cpu/aggregate.c:129-188, src/gpu/aggregate.cu:357-438):
//   permuted[perm[i]] = size(gather[i]); chunk_sizes = permuted (unpadded);
//   the last slot of every group of n_active*cps_inner slots is padded so the
//   group ends on a page boundary; offsets = exclusive scan of the padded sizes;
//   bytes of item i are copied to data + offsets[perm[i]].
//
// gfx950 design: the size permutation, padding and scan are one single-
// workgroup kernel (slot counts are thousands, so one CU finishes it in a few
// microseconds and no host round trip or device-wide scan is needed). The byte
// gather runs one workgroup per (item, segment) and writes aligned dwords:
// each lane funnel-shifts two source dwords into one destination dword, so
// unaligned destination offsets still produce full-width stores.
#include "kernels.hpp"
#include "wave.hpp"

namespace chucky {
namespace {

constexpr int kPlanThreads = 1024;
constexpr uint64_t kGatherPiece = 64 * 1024; // bytes one workgroup copies at most

uint32_t pieces_of(const SegmentSource& s)
{
  const uint64_t seg = s.seg_sizes ? s.seg_stride : s.const_size;
  const uint64_t p = (seg + kGatherPiece - 1) / kGatherPiece;
  return (uint32_t)(p ? p : 1);
}

__device__ __forceinline__ uint64_t item_bytes(const SegmentSource& s, uint32_t chunk)
{
  if (!s.seg_sizes)
    return s.const_total;
  uint64_t t = 0;
  for (uint32_t k = 0; k < s.nseg; ++k)
    t += s.seg_sizes[(uint64_t)chunk * s.nseg + k];
  return t;
}

__device__ __forceinline__ void plan_body(
  const SegmentSource& src, const uint32_t* __restrict__ gather, const uint32_t* __restrict__ perm,
  uint64_t n_items, uint64_t n_slots, uint64_t group, uint64_t page, uint64_t* __restrict__ chunk_sizes,
  uint64_t* __restrict__ offsets, uint64_t* __restrict__ host_sizes, uint64_t* __restrict__ host_offsets)
{
  // Every pass issues the loads of kBatch slots before using them: a pass
  // that loaded, used and stored one slot at a time would wait a global
  // memory latency per slot (tens of slots per thread for small chunks).
  constexpr int kBatch = 8;
  __shared__ uint64_t part[kPlanThreads];
  const int t = threadIdx.x;
  for (uint64_t i = t; i < n_slots; i += kPlanThreads)
    chunk_sizes[i] = 0;
  __syncthreads();
  if (src.seg_sizes && src.nseg > 1) {
    // Items of several segments: G lanes (a power of two, up to the whole
    // wavefront) sum one item's segment sizes side by side, so an item costs
    // one round of loads instead of one load latency per segment; each
    // thread holds kBatch items' loads in flight.
    const uint32_t nseg = src.nseg;
    uint32_t G = 1;
    while (G < nseg && G < 64)
      G <<= 1;
    const uint32_t lane = (uint32_t)t & 63u, gl = lane & (G - 1);
    const uint64_t per_wave = 64 / G;
    const uint64_t round = (uint64_t)(kPlanThreads / 64) * per_wave;
    const uint64_t mine = (uint64_t)(t >> 6) * per_wave + lane / G;
    for (uint64_t r0 = 0; r0 < n_items; r0 += round * kBatch) {
      uint32_t g[kBatch], pm[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const uint64_t i = r0 + (uint64_t)q * round + mine;
        g[q] = i < n_items ? gather[i] : 0u;
        pm[q] = i < n_items ? perm[i] : 0u;
      }
      uint64_t v[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const uint64_t i = r0 + (uint64_t)q * round + mine;
        const uint32_t* __restrict__ ss = src.seg_sizes + (uint64_t)g[q] * nseg;
        uint64_t acc = 0;
        if (i < n_items)
          for (uint32_t k = gl; k < nseg; k += G)
            acc += ss[k];
        v[q] = acc;
      }
      // sum within each group of G lanes (every lane takes part)
      for (uint32_t off = G >> 1; off; off >>= 1)
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
          const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v[q], (int)off, 64);
          const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v[q] >> 32), (int)off, 64);
          v[q] += ((uint64_t)hi << 32) | lo;
        }
#pragma unroll
      for (int q = 0; q < kBatch; ++q)
        if (gl == 0 && r0 + (uint64_t)q * round + mine < n_items)
          chunk_sizes[pm[q]] = v[q];
    }
  } else {
    for (uint64_t i0 = t; i0 < n_items; i0 += (uint64_t)kPlanThreads * kBatch) {
      uint32_t g[kBatch], pm[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const uint64_t i = i0 + (uint64_t)q * kPlanThreads;
        g[q] = i < n_items ? gather[i] : 0u;
        pm[q] = i < n_items ? perm[i] : 0u;
      }
      uint64_t v[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q)
        v[q] = item_bytes(src, g[q]);
#pragma unroll
      for (int q = 0; q < kBatch; ++q)
        if (i0 + (uint64_t)q * kPlanThreads < n_items)
          chunk_sizes[pm[q]] = v[q];
    }
  }
  __syncthreads();

  // Contiguous range per thread; padding is added on the fly to the value of
  // the last slot in each group (the unpadded size stays in chunk_sizes).
  const uint64_t per = (n_slots + kPlanThreads - 1) / kPlanThreads;
  const uint64_t b = (uint64_t)t * per;
  const uint64_t e = b + per < n_slots ? b + per : n_slots;
  auto pad = [&](uint64_t i, uint64_t v) -> uint64_t {
    if (page && group && (i % group) == group - 1) {
      uint64_t tot = 0;
      for (uint64_t j = i + 1 - group; j <= i; ++j)
        tot += chunk_sizes[j];
      v += (tot + page - 1) / page * page - tot;
    }
    return v;
  };
  uint64_t local = 0;
  for (uint64_t i0 = b; i0 < e; i0 += kBatch) {
    uint64_t v[kBatch];
#pragma unroll
    for (int q = 0; q < kBatch; ++q)
      v[q] = i0 + q < e ? chunk_sizes[i0 + q] : 0ull;
#pragma unroll
    for (int q = 0; q < kBatch; ++q)
      local += i0 + q < e ? pad(i0 + q, v[q]) : 0ull;
  }
  part[t] = local;
  __syncthreads();
  // Hillis-Steele inclusive scan over the 1024 partials.
  for (int off = 1; off < kPlanThreads; off <<= 1) {
    const uint64_t add = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  uint64_t run = part[t] - local;
  for (uint64_t i0 = b; i0 < e; i0 += kBatch) {
    uint64_t v[kBatch];
#pragma unroll
    for (int q = 0; q < kBatch; ++q)
      v[q] = i0 + q < e ? chunk_sizes[i0 + q] : 0ull;
#pragma unroll
    for (int q = 0; q < kBatch; ++q)
      if (i0 + q < e) {
        offsets[i0 + q] = run;
        run += pad(i0 + q, v[q]);
      }
  }
  if (t == kPlanThreads - 1)
    offsets[n_slots] = part[t];
  if (host_offsets || host_sizes) {
    // the copies for the host go out with consecutive threads on consecutive
    // slots: a thread's own contiguous range would write pinned memory at a
    // 8 * per byte stride across the wavefront, one PCIe transaction per
    // 8-byte value
    // (kBatch slots' loads are issued before their stores, so a thread waits
    // one load latency per kBatch slots, not per slot)
    __syncthreads();
    for (uint64_t i0 = t; i0 <= n_slots; i0 += (uint64_t)kPlanThreads * kBatch) {
      uint64_t o[kBatch], z[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const uint64_t i = i0 + (uint64_t)q * kPlanThreads;
        o[q] = i <= n_slots ? offsets[i] : 0ull;
        z[q] = i < n_slots ? chunk_sizes[i] : 0ull;
      }
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const uint64_t i = i0 + (uint64_t)q * kPlanThreads;
        if (host_offsets && i <= n_slots)
          host_offsets[i] = o[q];
        if (host_sizes && i < n_slots)
          host_sizes[i] = z[q];
      }
    }
  }
}

__global__ __launch_bounds__(kPlanThreads) void aggregate_plan_k(
  SegmentSource src, const uint32_t* __restrict__ gather, const uint32_t* __restrict__ perm,
  uint64_t n_items, uint64_t n_slots, uint64_t group, uint64_t page, uint64_t* chunk_sizes,
  uint64_t* offsets, uint64_t* host_sizes, uint64_t* host_offsets)
{
  plan_body(src, gather, perm, n_items, n_slots, group, page, chunk_sizes, offsets, host_sizes,
            host_offsets);
}

// One workgroup per level.
__global__ __launch_bounds__(kPlanThreads) void aggregate_plan_levels_k(SegmentSource src, AggLevels a)
{
  const int k = blockIdx.x;
  plan_body(src, a.gather[k], a.perm[k], a.n_items[k], a.n_slots[k], a.group[k], 0, a.sizes[k],
            a.offsets[k], a.host_sizes[k], a.host_offsets[k]);
}

__device__ __forceinline__ uint32_t load_u32_unaligned(const uint8_t* p)
{
  const uintptr_t a = (uintptr_t)p;
  const uint32_t* w = (const uint32_t*)(a & ~(uintptr_t)3);
  const uint32_t sh = (uint32_t)(a & 3) * 8;
  if (sh == 0)
    return w[0];
  return (uint32_t)(((uint64_t)w[1] << 32 | w[0]) >> sh);
}

// Workgroup-cooperative copy of n bytes to any destination alignment: byte
// head to a dword boundary of dst, dword body (aligned or funnel-shifted
// source loads), byte tail.
__device__ __forceinline__ void block_copy(uint8_t* out, const uint8_t* in, uint64_t nbytes)
{
  const uint32_t head = (uint32_t)((4 - ((uintptr_t)out & 3)) & 3);
  const uint64_t h = head < nbytes ? head : nbytes;
  if (threadIdx.x < h)
    out[threadIdx.x] = in[threadIdx.x];
  const uint64_t body = (nbytes - h) / 4;
  uint32_t* ow = (uint32_t*)(out + h);
  const uint8_t* ib = in + h;
  if (((uintptr_t)ib & 3) == 0) {
    const uint32_t* iw = (const uint32_t*)ib;
    for (uint64_t w = threadIdx.x; w < body; w += blockDim.x)
      ow[w] = iw[w];
  } else {
    for (uint64_t w = threadIdx.x; w < body; w += blockDim.x)
      ow[w] = load_u32_unaligned(ib + 4 * w);
  }
  const uint64_t tail0 = h + body * 4;
  if (threadIdx.x < nbytes - tail0)
    out[tail0 + threadIdx.x] = in[tail0 + threadIdx.x];
}

// Workgroups take (item, segment, piece) triples with a grid stride: a
// compressed segment lands at its chunk's offset plus the sizes of the
// chunk's earlier segments (summed by the workgroup), and a segment larger
// than kGatherPiece is copied by several workgroups, so a batch of few large
// chunks (one LZ4 segment per chunk) still spreads over the whole GPU.
// Constant-size sources (CODEC_NONE chunks split into fixed segments) are
// walked the same way.
// A gather into pinned host memory runs at PCIe speed, so it is given only a
// few CUs and leaves the rest to the compressor and transpose beside it.
// One (item, segment) pair of a level (uniform across the workgroup).
__device__ __forceinline__ void gather_pair(const SegmentSource& src, const uint32_t* __restrict__ gather,
                                            const uint32_t* __restrict__ perm,
                                            const uint64_t* __restrict__ offsets, uint8_t* __restrict__ dst,
                                            uint64_t tr, uint32_t* part)
{
  const uint64_t pr = tr / src.pieces;
  const uint64_t piece_off = (uint64_t)(uint32_t)(tr - pr * src.pieces) * kGatherPiece;
  const uint64_t item = pr / src.nseg;
  const uint32_t seg = (uint32_t)(pr - item * src.nseg);
  const uint32_t chunk = gather[item];
  if (src.seg_sizes) {
    const uint32_t* sz = src.seg_sizes + (uint64_t)chunk * src.nseg;
    const uint32_t nb = sz[seg];
    if (nb <= piece_off)
      return;  // uniform across the workgroup: no barrier is skipped by part of it
    // a chunk's compressed bytes (below its slot stride) fit 32 bits
    uint32_t before = 0;
    for (uint32_t j = threadIdx.x; j < seg; j += 256)
      before += sz[j];
    before = wave_sum(before);
    if ((threadIdx.x & 63) == 0)
      part[threadIdx.x >> 6] = before;
    __syncthreads();
    before = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    const uint64_t n = nb - piece_off < kGatherPiece ? nb - piece_off : kGatherPiece;
    block_copy(dst + offsets[perm[item]] + before + piece_off,
               src.base + (uint64_t)chunk * src.chunk_stride + (uint64_t)seg * src.seg_stride + piece_off, n);
    return;
  }
  const uint64_t seg_off = (uint64_t)seg * src.const_size;
  const uint64_t nbytes = seg + 1 == src.nseg ? src.const_total - seg_off : src.const_size;
  if (nbytes <= piece_off || seg_off >= src.const_total)
    return;
  const uint64_t n = nbytes - piece_off < kGatherPiece ? nbytes - piece_off : kGatherPiece;
  block_copy(dst + offsets[perm[item]] + seg_off + piece_off,
             src.base + (uint64_t)chunk * src.chunk_stride + (uint64_t)seg * src.seg_stride + piece_off, n);
}

__global__ __launch_bounds__(256) void aggregate_gather_k(SegmentSource src,
                                                          const uint32_t* __restrict__ gather,
                                                          const uint32_t* __restrict__ perm,
                                                          uint64_t n_items,
                                                          const uint64_t* __restrict__ offsets,
                                                          uint8_t* __restrict__ dst)
{
  __shared__ uint32_t part[256 / 64];
  const uint64_t pairs = n_items * src.nseg * src.pieces;
  for (uint64_t pr = blockIdx.x; pr < pairs; pr += gridDim.x)
    gather_pair(src, gather, perm, offsets, dst, pr, part);
}

// Every level's pairs, level after level, in one grid.
__global__ __launch_bounds__(256) void aggregate_gather_levels_k(SegmentSource src, AggLevels a)
{
  __shared__ uint32_t part[256 / 64];
  uint64_t first = 0, base = 0;
  for (int k = 0; k < a.nlev; ++k) {
    const uint64_t pairs = a.n_items[k] * src.nseg * src.pieces;
    // this workgroup's first pair of level k under the grid stride
    uint64_t pr = first <= blockIdx.x ? blockIdx.x - first
                                      : (gridDim.x - (first - blockIdx.x) % gridDim.x) % gridDim.x;
    uint8_t* dst = a.packed ? a.dst[0] + base : a.dst[k];
    for (; pr < pairs; pr += gridDim.x)
      gather_pair(src, a.gather[k], a.perm[k], a.offsets[k], dst, pr, part);
    first += pairs;
    if (a.packed)
      base += a.offsets[k][a.n_slots[k]]; // level k's total, from the plan launch
  }
}

} // namespace

hipError_t launch_aggregate_plan(const SegmentSource& src, const uint32_t* gather,
                                 const uint32_t* perm, uint64_t n_items, uint64_t n_slots,
                                 uint64_t group, uint64_t page_size, uint64_t* chunk_sizes,
                                 uint64_t* offsets, hipStream_t stream, uint64_t* host_sizes,
                                 uint64_t* host_offsets)
{
  hipLaunchKernelGGL(aggregate_plan_k, dim3(1), dim3(kPlanThreads), 0, stream, src, gather, perm,
                     n_items, n_slots, group, page_size, chunk_sizes, offsets, host_sizes, host_offsets);
  return hipGetLastError();
}

hipError_t launch_aggregate_levels(const SegmentSource& src, const AggLevels& a, hipStream_t stream,
                                   uint32_t max_workgroups)
{
  if (a.nlev <= 0 || a.nlev > kAggMaxLevels)
    return hipErrorInvalidValue;
  SegmentSource s = src;
  s.pieces = pieces_of(src);
  hipLaunchKernelGGL(aggregate_plan_levels_k, dim3((uint32_t)a.nlev), dim3(kPlanThreads), 0, stream, s, a);
  uint64_t pairs = 0;
  for (int k = 0; k < a.nlev; ++k)
    pairs += a.n_items[k] * s.nseg * s.pieces;
  if (pairs == 0)
    return hipGetLastError();
  const uint64_t cap = max_workgroups ? max_workgroups : (1u << 20);
  const uint64_t wg = pairs < cap ? pairs : cap;
  hipLaunchKernelGGL(aggregate_gather_levels_k, dim3((uint32_t)wg), dim3(256), 0, stream, s, a);
  return hipGetLastError();
}

hipError_t launch_aggregate_gather(const SegmentSource& src, const uint32_t* gather,
                                   const uint32_t* perm, uint64_t n_items, const uint64_t* offsets,
                                   uint8_t* dst, hipStream_t stream, uint32_t max_workgroups)
{
  if (n_items == 0)
    return hipSuccess;
  SegmentSource s = src;
  s.pieces = pieces_of(src);
  uint64_t wg = n_items * s.nseg * s.pieces;
  const uint64_t cap = max_workgroups ? max_workgroups : (1u << 20);
  if (wg > cap)
    wg = cap;
  hipLaunchKernelGGL(aggregate_gather_k, dim3((uint32_t)wg), dim3(256), 0, stream, s, gather, perm,
                     n_items, offsets, dst);
  return hipGetLastError();
}

} // namespace chucky
