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

template<int HASH_LOG>
__global__ __launch_bounds__(64) void lz4_segment_k(const uint8_t* __restrict__ src,
                                                    uint64_t src_stride, uint32_t n,
                                                    uint32_t seg_len, uint32_t nseg,
                                                    uint64_t total, uint8_t* __restrict__ tmp,
                                                    uint64_t seg_bound, SegRec* __restrict__ recs)
{
  __shared__ uint32_t table[1 << HASH_LOG]; // position + 1, 0 = empty
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
        v = ld32u(in + p);
        h = lz4_hash(v, HASH_LOG);
        cand = table[h];
      }
      bool hit = false;
      if (active && cand != 0) {
        const uint32_t c = cand - 1;
        hit = c < p && (p - c) <= kMaxOffset && ld32u(in + c) == v;
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
      uint32_t len = wave_match_extend(in, mp, mc, 4, match_end);
      while (mp > anchor && mc > 0 && (mp - mc) <= kMaxOffset && in[mp - 1] == in[mc - 1]) {
        --mp;
        --mc;
        ++len;
      }
      if (nseq == 0) {
        ll0 = mp - s0;
        tok0 = len - 4 >= 15 ? 15 : len - 4;
        op = emit_match_part(out, op, len, mp - mc);
      } else {
        op = emit(out, op, in, anchor, mp - anchor, true, len, mp - mc);
      }
      ++nseq;
      pos = mp + len;
      anchor = pos;
      if (lane == 0 && pos - 2 <= last_start) {
        const uint32_t q = pos - 2;
        table[lz4_hash(ld32u(in + q), HASH_LOG)] = q + 1;
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

// Where the stitched block puts one segment's pieces: its literal run (input
// bytes [lit_src, lit_src + lit_len), carried from earlier segments without
// a match) at lit_dst, its sequences (scratch) at body_dst. Entry nseg of a
// chunk is the block's closing literal run (body_len 0).
struct SegPlace
{
  uint32_t lit_src, lit_len, lit_dst, body_dst, body_len;
};

// One thread per chunk: walks the chunk's segment records in order, writes
// each literal run's token (with the first match's length nibble) and length
// extension into the block, and records where the literal and sequence bytes
// go; the closing literal-only sequence ends the block (LZ4 end rules).
__global__ __launch_bounds__(64) void lz4_layout_k(uint32_t n, uint32_t seg_len, uint32_t nseg,
                                                   uint64_t nchunks, const SegRec* __restrict__ recs,
                                                   SegPlace* __restrict__ places, uint8_t* __restrict__ dst,
                                                   uint64_t dst_stride, uint32_t* __restrict__ sizes)
{
  const uint64_t chunk = (uint64_t)blockIdx.x * 64 + threadIdx.x;
  if (chunk >= nchunks)
    return;
  uint8_t* out = dst + chunk * dst_stride;
  SegPlace* pl = places + chunk * (nseg + 1);
  uint32_t op = 0, carry = 0; // carry: input position where pending literals start
  for (uint32_t seg = 0; seg < nseg; ++seg) {
    const SegRec r = recs[chunk * nseg + seg];
    if (r.nseq == 0) {
      pl[seg] = SegPlace{ 0, 0, 0, 0, 0 };
      continue;
    }
    const uint32_t s0 = seg * seg_len;
    const uint32_t lit = s0 + r.ll0 - carry;
    out[op] = (uint8_t)(((lit >= 15 ? 15 : lit) << 4) | r.tok0);
    op = lit >= 15 ? put_len_ext(out, op + 1, lit) : op + 1;
    pl[seg] = SegPlace{ carry, lit, op, op + lit, r.body };
    op += lit + r.body;
    const uint32_t s1 = n - s0 < seg_len ? n : s0 + seg_len;
    carry = s1 - r.tail;
  }
  const uint32_t lit = n - carry;
  out[op] = (uint8_t)((lit >= 15 ? 15 : lit) << 4);
  op = lit >= 15 ? put_len_ext(out, op + 1, lit) : op + 1;
  pl[nseg] = SegPlace{ carry, lit, op, op + lit, 0 };
  sizes[chunk] = op + lit;
}

constexpr uint32_t kStitchPiece = 16 * 1024;

// Grid-stride over (chunk, segment, part, piece): part 0 copies a piece of
// the segment's literal run from the input, part 1 a piece of its sequences
// from scratch, each to its place in the block.
__global__ __launch_bounds__(256) void lz4_copy_k(const uint8_t* __restrict__ src, uint64_t src_stride,
                                                  uint32_t nseg, uint64_t nchunks, uint32_t lit_pieces,
                                                  uint32_t body_pieces, const uint8_t* __restrict__ tmp,
                                                  uint64_t seg_bound, const SegPlace* __restrict__ places,
                                                  uint8_t* __restrict__ dst, uint64_t dst_stride)
{
  const uint64_t per_seg = (uint64_t)lit_pieces + body_pieces;
  const uint64_t total = nchunks * (nseg + 1) * per_seg;
  for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
    const uint64_t cs = w / per_seg; // chunk * (nseg + 1) + seg
    const uint32_t k = (uint32_t)(w - cs * per_seg);
    const uint64_t chunk = cs / (nseg + 1);
    const uint32_t seg = (uint32_t)(cs - chunk * (nseg + 1));
    const SegPlace pl = places[cs];
    uint8_t* out = dst + chunk * dst_stride;
    if (k < lit_pieces) {
      const uint32_t off = k * kStitchPiece;
      if (off < pl.lit_len)
        block_copy(out + pl.lit_dst + off, src + chunk * src_stride + pl.lit_src + off,
                   pl.lit_len - off < kStitchPiece ? pl.lit_len - off : kStitchPiece);
    } else {
      const uint32_t off = (k - lit_pieces) * kStitchPiece;
      if (off < pl.body_len)
        block_copy(out + pl.body_dst + off, tmp + (chunk * nseg + seg) * seg_bound + off,
                   pl.body_len - off < kStitchPiece ? pl.body_len - off : kStitchPiece);
    }
  }
}

} // namespace

size_t lz4_max_output_size(size_t chunk_bytes)
{
  // LZ4_COMPRESSBOUND: n + n/255 + 16
  return chunk_bytes + chunk_bytes / 255 + 16;
}

uint64_t lz4_segment_bound(uint32_t seg_bytes) { return (lz4_max_output_size(seg_bytes) + 15) & ~15ull; }

size_t lz4_scratch_bytes(uint64_t chunk_bytes, uint64_t n, uint32_t seg_bytes)
{
  if (!seg_bytes || seg_bytes >= chunk_bytes)
    return 0;
  const uint64_t nseg = (chunk_bytes + seg_bytes - 1) / seg_bytes;
  return n * nseg * (lz4_segment_bound(seg_bytes) + sizeof(SegRec)) + n * (nseg + 1) * sizeof(SegPlace) + 256;
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
#define LZ4_SEG(H)                                                                            \
  hipLaunchKernelGGL(lz4_segment_k<H>, dim3((uint32_t)total), dim3(64), 0, stream, src,         \
                     src_stride, (uint32_t)chunk_bytes, seg_bytes, nseg, total, tmp, bound, recs)
  switch (hash_log) {
    case 10: LZ4_SEG(10); break;
    case 11: LZ4_SEG(11); break;
    case 13: LZ4_SEG(13); break;
    default: LZ4_SEG(12); break;
  }
#undef LZ4_SEG
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  SegPlace* places = (SegPlace*)(((uintptr_t)(recs + total) + 15) & ~(uintptr_t)15);
  hipLaunchKernelGGL(lz4_layout_k, dim3((uint32_t)((n + 63) / 64)), dim3(64), 0, stream, (uint32_t)chunk_bytes,
                     seg_bytes, nseg, n, recs, places, dst, dst_stride, sizes);
  const uint32_t lit_pieces = (uint32_t)((chunk_bytes + kStitchPiece - 1) / kStitchPiece);
  const uint32_t body_pieces = (uint32_t)((bound + kStitchPiece - 1) / kStitchPiece);
  const uint64_t items = n * (nseg + 1) * ((uint64_t)lit_pieces + body_pieces);
  hipLaunchKernelGGL(lz4_copy_k, dim3((uint32_t)std::min<uint64_t>(items, 1u << 16)), dim3(256), 0, stream,
                     src, src_stride, nseg, n, lit_pieces, body_pieces, tmp, bound, places, dst, dst_stride);
  return hipGetLastError();
}

} // namespace chucky
