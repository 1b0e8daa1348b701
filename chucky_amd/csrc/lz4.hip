// LZ4 raw-block compressor: one wavefront per chunk.
//
// Output format: one LZ4 block (no frame) per chunk, the same container the
// reference writes with LZ4_compress_HC (src/cpu/compress.c:55-76); any
// conforming decoder (LZ4_decompress_safe) restores the chunk. The bytes
// themselves differ from LZ4-HC, which is allowed: chunks are checked by
// round trip.
//
// Parse: greedy, 64 candidate positions per step. Every lane hashes the four
// bytes at its position into an LDS table (4-byte keys, multiplicative hash),
// verifies its candidate, and the wave takes the first lane with a match
// (ballot + ffs). The match is extended 256 bytes per step with a ballot over
// per-lane dword compares, extended backwards over pending literals, and
// emitted: lane 0 writes token / length bytes / offset, the wave copies the
// literal run with dword stores. Block-end rules of the LZ4 format are kept:
// matches start at most 12 bytes before the end and end at least 5 bytes
// before it; offsets are at most 65535.
#include "kernels.hpp"

#include <algorithm>
#include "wave.hpp"

namespace chucky {
namespace {

constexpr uint32_t kMaxOffset = 65535;
constexpr uint32_t kMfLimit = 12;
constexpr uint32_t kLastLiterals = 5;
// LDS staging capacities of the segment parser (bytes of input per
// wavefront): a segment is staged in the smallest that holds it, so small
// segments keep the wavefront's LDS small and more wavefronts fit a CU.
constexpr int kLz4StageMax = 16 * 1024;

constexpr int lz4_stage_for(uint32_t seg_bytes)
{
  return seg_bytes <= 2048 ? 2048 : seg_bytes <= 4096 ? 4096 : seg_bytes <= 8192 ? 8192
                                                     : seg_bytes <= (uint32_t)kLz4StageMax ? kLz4StageMax : 0;
}

__device__ __forceinline__ uint32_t lz4_hash(uint32_t v, int hash_log)
{
  return (v * 2654435761u) >> (32 - hash_log);
}

__device__ __forceinline__ uint32_t put_len_ext(uint8_t* out, uint32_t op, uint32_t v)
{
  // v >= 15 was flagged in the token; write (v - 15) as 255-runs.
  v -= 15;
  while (v >= 255) {
    out[op++] = 255;
    v -= 255;
  }
  out[op++] = (uint8_t)v;
  return op;
}

// Emits one sequence (literals [anchor, anchor+lit) then optionally a match).
// Uniform arguments; returns the new output cursor.
__device__ __forceinline__ uint32_t emit(uint8_t* out, uint32_t op, const uint8_t* in,
                                         uint32_t anchor, uint32_t lit, bool has_match,
                                         uint32_t mlen, uint32_t offset)
{
  const int lane = lane_id();
  const uint32_t ml4 = has_match ? mlen - 4 : 0;
  // header length is uniform; lane 0 writes it
  uint32_t hdr = 1 + (lit >= 15 ? 1 + (lit - 15) / 255 : 0);
  if (lane == 0) {
    const uint8_t tok = (uint8_t)(((lit >= 15 ? 15 : lit) << 4) | (ml4 >= 15 ? 15 : ml4));
    out[op] = tok;
    if (lit >= 15)
      put_len_ext(out, op + 1, lit);
  }
  op += hdr;
  wave_copy(out + op, in + anchor, lit);
  op += lit;
  if (has_match) {
    if (lane == 0) {
      out[op] = (uint8_t)offset;
      out[op + 1] = (uint8_t)(offset >> 8);
      if (ml4 >= 15)
        put_len_ext(out, op + 2, ml4);
    }
    op += 2 + (ml4 >= 15 ? 1 + (ml4 - 15) / 255 : 0);
  }
  return op;
}

template<int HASH_LOG>
__global__ __launch_bounds__(64) void lz4_chunk_k(const uint8_t* __restrict__ src,
                                                  uint64_t src_stride, uint32_t n, uint64_t nchunks,
                                                  uint8_t* __restrict__ dst, uint64_t dst_stride,
                                                  uint32_t* __restrict__ sizes)
{
  __shared__ uint32_t table[1 << HASH_LOG]; // position + 1, 0 = empty
  const uint64_t chunk = blockIdx.x;
  if (chunk >= nchunks)
    return;
  const int lane = lane_id();
  const uint8_t* in = src + chunk * src_stride;
  uint8_t* out = dst + chunk * dst_stride;
  for (int i = lane; i < (1 << HASH_LOG); i += kWave)
    table[i] = 0;
  __syncthreads();

  uint32_t op = 0, anchor = 0;
  if (n > kMfLimit) {
    const uint32_t last_start = n - kMfLimit;   // matches start at <= this
    const uint32_t match_end = n - kLastLiterals; // and end at <= this
    uint32_t pos = 0;
    while (pos <= last_start) {
      // search acceleration: lanes spread out over long literal runs
      const uint32_t stride = 1 + ((pos - anchor) >> 10);
      const uint32_t p = pos + (uint32_t)lane * stride;
      const bool active = p <= last_start;
      uint32_t v = 0, h = 0, cand = 0;
      if (active) {
        v = ld32u(in + p);
        h = lz4_hash(v, HASH_LOG);
        cand = table[h];
      }
      // Lanes past the previous match may have left entries at or after the
      // current positions: only strictly earlier candidates are matches.
      bool hit = false;
      if (active && cand != 0) {
        const uint32_t c = cand - 1;
        hit = c < p && (p - c) <= kMaxOffset && ld32u(in + c) == v;
      }
      // all lanes read before any lane writes (wave executes in lockstep)
      __builtin_amdgcn_wave_barrier();
      if (active)
        table[h] = p + 1;
      __builtin_amdgcn_wave_barrier();
      const uint64_t mask = wave_ballot(hit);
      if (mask == 0) {
        pos += kWave * stride;
        continue;
      }
      const int first = __ffsll((unsigned long long)mask) - 1;
      uint32_t mp = pos + (uint32_t)first * stride;
      uint32_t mc = wave_bcast(cand, first) - 1;
      uint32_t len = wave_match_extend(in, mp, mc, 4, match_end);
      // backward extension over pending literals
      while (mp > anchor && mc > 0 && in[mp - 1] == in[mc - 1]) {
        --mp;
        --mc;
        ++len;
      }
      op = emit(out, op, in, anchor, mp - anchor, true, len, mp - mc);
      pos = mp + len;
      anchor = pos;
      if (lane == 0 && pos - 2 <= last_start) {
        const uint32_t q = pos - 2;
        table[lz4_hash(ld32u(in + q), HASH_LOG)] = q + 1;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  op = emit(out, op, in, anchor, n - anchor, false, 0, 0);
  if (lane == 0)
    sizes[chunk] = op;
}

struct SegRec
{
  uint32_t nseq; // sequences with a match
  uint32_t ll0;  // literals of the first sequence inside the segment
  uint32_t tok0; // match-length nibble of the first sequence
  uint32_t tail; // trailing literals (carried into the next segment)
  uint32_t body; // bytes written to scratch
};

// Writes a sequence's match part: offset, match-length extension.
__device__ __forceinline__ uint32_t emit_match_part(uint8_t* out, uint32_t op, uint32_t mlen,
                                                    uint32_t offset)
{
  const uint32_t ml4 = mlen - 4;
  if (lane_id() == 0) {
    out[op] = (uint8_t)offset;
    out[op + 1] = (uint8_t)(offset >> 8);
    if (ml4 >= 15)
      put_len_ext(out, op + 2, ml4);
  }
  return op + 2 + (ml4 >= 15 ? 1 + (ml4 - 15) / 255 : 0);
}

// One wavefront per segment of a chunk. With STAGE > 0 (segments of at most
// STAGE bytes) the segment and the 8 bytes after it are first copied into
// LDS with wide loads issued together, so the parse loop's position and
// candidate loads and the match extension read LDS instead of waiting on
// HBM latency step after step; only the backward extension, which may reach
// before the segment, still reads the chunk in global memory.
template<int HASH_LOG, int STAGE>
__global__ __launch_bounds__(64) void lz4_segment_k(const uint8_t* __restrict__ src,
                                                    uint64_t src_stride, uint32_t n,
                                                    uint32_t seg_len, uint32_t nseg,
                                                    uint64_t total, uint8_t* __restrict__ tmp,
                                                    uint64_t seg_bound, SegRec* __restrict__ recs)
{
  __shared__ uint32_t table[1 << HASH_LOG]; // position + 1, 0 = empty
  __shared__ uint4 stage_v[STAGE ? (STAGE + 16) / 16 : 1];
  const uint64_t item = blockIdx.x;
  if (item >= total)
    return;
  const int lane = lane_id();
  const uint64_t chunk = item / nseg;
  const uint32_t seg = (uint32_t)(item - chunk * nseg);
  const uint8_t* in = src + chunk * src_stride;
  uint8_t* out = tmp + item * seg_bound;
  const uint32_t s0 = seg * seg_len;
  const uint32_t s1 = n - s0 < seg_len ? n : s0 + seg_len;
  for (int i = lane; i < (1 << HASH_LOG); i += kWave)
    table[i] = 0;
  const uint8_t* base = in + s0; // input byte s0 + i is base[i] for i < s1 - s0 + 8
  if constexpr (STAGE > 0) {
    const uint32_t nb = s1 - s0 + 8; // <= STAGE + 8: fits stage_v rounded up to 16
    const uint8_t* g = in + s0;
    if (((uintptr_t)g & 15) == 0) {
      const uint4* gv = (const uint4*)g;
      for (uint32_t i = lane; i < (nb + 15) / 16; i += kWave)
        stage_v[i] = gv[i];
    } else {
      uint32_t* sw = (uint32_t*)stage_v;
      for (uint32_t i = lane; i < (nb + 3) / 4; i += kWave)
        sw[i] = ld32u(g + 4 * i);
    }
    base = (const uint8_t*)stage_v;
  }
  __syncthreads();

  uint32_t op = 0, anchor = s0, nseq = 0, ll0 = 0, tok0 = 0;
  // block-end rules against the chunk end; keys and matches stay in [s0, s1)
  const uint32_t last_start = n < kMfLimit ? 0 : (s1 - 4 < n - kMfLimit ? s1 - 4 : n - kMfLimit);
  const uint32_t match_end = s1 < n - kLastLiterals ? s1 : n - kLastLiterals;
  if (n > kMfLimit && s1 - s0 >= 8 && last_start >= s0) {
    uint32_t pos = s0;
    while (pos <= last_start) {
      const uint32_t stride = 1 + ((pos - anchor) >> 10);
      const uint32_t p = pos + (uint32_t)lane * stride;
      const bool active = p <= last_start;
      uint32_t v = 0, h = 0, cand = 0;
      if (active) {
        v = ld32u(base + (p - s0));
        h = lz4_hash(v, HASH_LOG);
        cand = table[h];
      }
      // table entries are positions of this segment: c >= s0
      bool hit = false;
      if (active && cand != 0) {
        const uint32_t c = cand - 1;
        hit = c < p && (p - c) <= kMaxOffset && ld32u(base + (c - s0)) == v;
      }
      __builtin_amdgcn_wave_barrier();
      if (active)
        table[h] = p + 1;
      __builtin_amdgcn_wave_barrier();
      const uint64_t mask = wave_ballot(hit);
      if (mask == 0) {
        pos += kWave * stride;
        continue;
      }
      const int first = __ffsll((unsigned long long)mask) - 1;
      uint32_t mp = pos + (uint32_t)first * stride;
      uint32_t mc = wave_bcast(cand, first) - 1;
      uint32_t len = wave_match_extend(base, mp - s0, mc - s0, 4, match_end - s0);
      // mp - 1 >= anchor >= s0 is staged; mc - 1 may lie before the segment
      while (mp > anchor && mc > 0 && (mp - mc) <= kMaxOffset &&
             base[mp - 1 - s0] == (mc - 1 >= s0 ? base[mc - 1 - s0] : in[mc - 1])) {
        --mp;
        --mc;
        ++len;
      }
      if (nseq == 0) {
        ll0 = mp - s0;
        tok0 = len - 4 >= 15 ? 15 : len - 4;
        op = emit_match_part(out, op, len, mp - mc);
      } else {
        op = emit(out, op, base + (anchor - s0), 0, mp - anchor, true, len, mp - mc);
      }
      ++nseq;
      pos = mp + len;
      anchor = pos;
      if (lane == 0 && pos - 2 <= last_start) {
        const uint32_t q = pos - 2;
        table[lz4_hash(ld32u(base + (q - s0)), HASH_LOG)] = q + 1;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (lane == 0)
    recs[item] = SegRec{ nseq, ll0, tok0, s1 - anchor, op };
}

// Workgroup-cooperative copy (any alignment): byte head to a dword boundary of
// dst, dword body from funnel-shifted source, byte tail.
__device__ __forceinline__ void block_copy(uint8_t* dst, const uint8_t* src, uint32_t n)
{
  const uint32_t t = threadIdx.x, nt = blockDim.x;
  const uint32_t head = (uint32_t)((4 - ((uintptr_t)dst & 3)) & 3);
  const uint32_t h = head < n ? head : n;
  if (t < h)
    dst[t] = src[t];
  const uint32_t body = (n - h) >> 2;
  uint32_t* dw = (uint32_t*)(dst + h);
  const uint8_t* s = src + h;
  for (uint32_t w = t; w < body; w += nt)
    dw[w] = ld32u(s + 4 * w);
  const uint32_t t0 = h + body * 4;
  if (t < n - t0)
    dst[t0 + t] = src[t0 + t];
}

// Where the stitched block puts the bytes of one input segment. Part a: the
// literals of the segment's first sequence that lie inside it, input
// [a_src, a_src + a_len). Part b: the input bytes after its last match (all
// of it when it holds no match), [b_src, b_src + b_len), which open the
// literal run of the next segment with a match (or the block's closing run).
// Body: the segment's sequences from scratch. Every input byte that is a
// literal lies in exactly one part a or b, so the copies are disjoint.
struct SegPlace
{
  uint32_t a_src, a_len, a_dst, b_src, b_len, b_dst, body_dst, body_len;
};

// One wavefront per chunk, 64 segments per step. A segment with matches
// ("live") opens a literal run that starts where the previous live segment's
// matches ended (its carry-in, found by ballot), so its token and length
// bytes take a known number of bytes and an exclusive scan over the wave
// gives every segment's output position. A second, backward pass gives each
// segment's part b the output position of the run it belongs to: delta =
// output position - input position is constant along one literal run.
__global__ __launch_bounds__(64) void lz4_layout_k(uint32_t n, uint32_t seg_len, uint32_t nseg,
                                                   const SegRec* __restrict__ recs_all,
                                                   SegPlace* __restrict__ places_all,
                                                   uint8_t* __restrict__ dst, uint64_t dst_stride,
                                                   uint32_t* __restrict__ sizes)
{
  const uint64_t chunk = blockIdx.x;
  const int lane = lane_id();
  uint8_t* out = dst + chunk * dst_stride;
  const SegRec* recs = recs_all + chunk * nseg;
  SegPlace* places = places_all + chunk * nseg;
  const uint64_t below = (1ull << lane) - 1;
  uint32_t op = 0, carry = 0; // uniform: output cursor, start of the pending literals
  for (uint32_t w0 = 0; w0 < nseg; w0 += kWave) {
    const uint32_t seg = w0 + (uint32_t)lane;
    SegRec r{ 0, 0, 0, 0, 0 };
    if (seg < nseg)
      r = recs[seg];
    const bool live = seg < nseg && r.nseq != 0;
    const uint32_t s0 = seg * seg_len;
    const uint32_t s1 = seg < nseg ? (n - s0 < seg_len ? n : s0 + seg_len) : s0;
    const uint32_t end = s1 - r.tail; // where this segment's pending literals start, if live
    const uint64_t live_mask = wave_ballot(live);
    const uint64_t lower = live_mask & below;
    const int j = lower ? 63 - __clzll((long long)lower) : 0;
    const uint32_t end_j = __shfl(end, j, kWave);
    const uint32_t cin = lower ? end_j : carry;
    const uint32_t lit = live ? s0 + r.ll0 - cin : 0;
    const uint32_t hdr = live ? 1 + (lit >= 15 ? 1 + (lit - 15) / 255 : 0) : 0;
    const uint32_t bytes = hdr + lit + (live ? r.body : 0);
    const uint32_t incl = wave_incl_scan(bytes);
    const uint32_t at = op + incl - bytes;
    if (live) {
      out[at] = (uint8_t)(((lit >= 15 ? 15 : lit) << 4) | r.tok0);
      if (lit >= 15)
        put_len_ext(out, at + 1, lit);
    }
    if (seg < nseg) {
      const uint32_t delta = at + hdr - cin; // output - input along this run (mod 2^32)
      places[seg] = live ? SegPlace{ s0, r.ll0, s0 + delta, end, r.tail, 0, at + hdr + lit, r.body }
                         : SegPlace{ s0, 0, 0, s0, s1 - s0, 0, 0, 0 };
    }
    op += wave_bcast(incl, kWave - 1);
    if (live_mask)
      carry = wave_bcast(end, 63 - __clzll((long long)live_mask));
  }
  const uint32_t lit = n - carry;
  const uint32_t hdr = 1 + (lit >= 15 ? 1 + (lit - 15) / 255 : 0);
  if (lane == 0) {
    out[op] = (uint8_t)((lit >= 15 ? 15 : lit) << 4);
    if (lit >= 15)
      put_len_ext(out, op + 1, lit);
    sizes[chunk] = op + hdr + lit;
  }
  uint32_t next_delta = op + hdr - carry; // the closing run's
  // the places read back here were written by the same lanes above
  for (uint32_t w0 = (nseg - 1) / kWave * kWave;; w0 -= kWave) {
    const uint32_t seg = w0 + (uint32_t)lane;
    SegPlace pl{};
    if (seg < nseg)
      pl = places[seg];
    const bool is_live = seg < nseg && recs[seg].nseq != 0;
    const uint32_t delta = pl.a_dst - pl.a_src;
    const uint64_t live_mask = wave_ballot(is_live);
    const uint64_t upper = lane == kWave - 1 ? 0 : live_mask & (~0ull << (lane + 1));
    const int k = upper ? __ffsll((unsigned long long)upper) - 1 : 0;
    const uint32_t delta_k = __shfl(delta, k, kWave);
    if (seg < nseg)
      places[seg].b_dst = pl.b_src + (upper ? delta_k : next_delta);
    if (live_mask)
      next_delta = wave_bcast(delta, __ffsll((unsigned long long)live_mask) - 1);
    if (w0 == 0)
      break;
  }
}

// One workgroup per (chunk, segment): the segment's literal parts a and b
// from the input and its sequences from scratch, each to its place.
__global__ __launch_bounds__(256) void lz4_copy_k(const uint8_t* __restrict__ src, uint64_t src_stride,
                                                  uint32_t nseg, uint64_t total, const uint8_t* __restrict__ tmp,
                                                  uint64_t seg_bound, const SegPlace* __restrict__ places,
                                                  uint8_t* __restrict__ dst, uint64_t dst_stride)
{
  for (uint64_t item = blockIdx.x; item < total; item += gridDim.x) {
    const uint64_t chunk = item / nseg;
    const SegPlace pl = places[item];
    const uint8_t* in = src + chunk * src_stride;
    uint8_t* out = dst + chunk * dst_stride;
    if (pl.a_len)
      block_copy(out + pl.a_dst, in + pl.a_src, pl.a_len);
    if (pl.b_len)
      block_copy(out + pl.b_dst, in + pl.b_src, pl.b_len);
    if (pl.body_len)
      block_copy(out + pl.body_dst, tmp + item * seg_bound, pl.body_len);
  }
}

} // namespace

size_t lz4_max_output_size(size_t chunk_bytes)
{
  // LZ4_COMPRESSBOUND: n + n/255 + 16
  return chunk_bytes + chunk_bytes / 255 + 16;
}

uint32_t lz4_lds_bytes(int hash_log, uint32_t seg_bytes)
{
  const int stage = seg_bytes == 0 ? 0 : lz4_stage_for(seg_bytes);
  return (4u << hash_log) + (stage ? (uint32_t)stage + 16 : 0u);
}

uint64_t lz4_segment_bound(uint32_t seg_bytes) { return (lz4_max_output_size(seg_bytes) + 15) & ~15ull; }

size_t lz4_scratch_bytes(uint64_t chunk_bytes, uint64_t n, uint32_t seg_bytes)
{
  if (!seg_bytes || seg_bytes >= chunk_bytes)
    return 0;
  const uint64_t nseg = (chunk_bytes + seg_bytes - 1) / seg_bytes;
  return n * nseg * (lz4_segment_bound(seg_bytes) + sizeof(SegRec)) + n * nseg * sizeof(SegPlace) + 256;
}

hipError_t launch_lz4(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                      uint8_t* dst, uint64_t dst_stride, uint32_t* sizes, int hash_log,
                      uint32_t seg_bytes, void* scratch, hipStream_t stream)
{
  if (n == 0)
    return hipSuccess;
  if (chunk_bytes > 0x7E000000ull)
    return hipErrorInvalidValue; // LZ4 block size limit
  const bool segmented = seg_bytes && seg_bytes < chunk_bytes;
  if (!segmented) {
#define LZ4_CHUNK(H)                                                                          \
  hipLaunchKernelGGL(lz4_chunk_k<H>, dim3((uint32_t)n), dim3(64), 0, stream, src, src_stride,   \
                     (uint32_t)chunk_bytes, n, dst, dst_stride, sizes)
    switch (hash_log) {
      case 9: LZ4_CHUNK(9); break;
      case 10: LZ4_CHUNK(10); break;
      case 11: LZ4_CHUNK(11); break;
      case 13: LZ4_CHUNK(13); break;
      default: LZ4_CHUNK(12); break;
    }
#undef LZ4_CHUNK
    return hipGetLastError();
  }
  if (!scratch || seg_bytes < 1024)
    return hipErrorInvalidValue;
  const uint32_t nseg = (uint32_t)((chunk_bytes + seg_bytes - 1) / seg_bytes);
  const uint64_t total = n * nseg;
  const uint64_t bound = lz4_segment_bound(seg_bytes);
  uint8_t* tmp = (uint8_t*)scratch;
  SegRec* recs = (SegRec*)(tmp + total * bound);
  // segments that fit the staging buffer are parsed out of LDS
#define LZ4_SEG(H, S)                                                                         \
  hipLaunchKernelGGL((lz4_segment_k<H, S>), dim3((uint32_t)total), dim3(64), 0, stream, src,    \
                     src_stride, (uint32_t)chunk_bytes, seg_bytes, nseg, total, tmp, bound, recs)
#define LZ4_SEG_H(S)                                                                          \
  switch (hash_log) {                                                                         \
    case 9: LZ4_SEG(9, S); break;                                                             \
    case 10: LZ4_SEG(10, S); break;                                                           \
    case 11: LZ4_SEG(11, S); break;                                                           \
    case 13: LZ4_SEG(13, S); break;                                                           \
    default: LZ4_SEG(12, S); break;                                                           \
  }
  switch (lz4_stage_for(seg_bytes)) {
    case 2048: LZ4_SEG_H(2048) break;
    case 4096: LZ4_SEG_H(4096) break;
    case 8192: LZ4_SEG_H(8192) break;
    case kLz4StageMax: LZ4_SEG_H(kLz4StageMax) break;
    default: LZ4_SEG_H(0) break;
  }
#undef LZ4_SEG_H
#undef LZ4_SEG
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  SegPlace* places = (SegPlace*)(((uintptr_t)(recs + total) + 15) & ~(uintptr_t)15);
  hipLaunchKernelGGL(lz4_layout_k, dim3((uint32_t)n), dim3(64), 0, stream, (uint32_t)chunk_bytes, seg_bytes, nseg,
                     recs, places, dst, dst_stride, sizes);
  hipLaunchKernelGGL(lz4_copy_k, dim3((uint32_t)std::min<uint64_t>(total, 1u << 20)), dim3(256), 0, stream, src,
                     src_stride, nseg, total, tmp, bound, places, dst, dst_stride);
  return hipGetLastError();
}

} // namespace chucky
