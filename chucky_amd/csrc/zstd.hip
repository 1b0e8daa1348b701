// zstd compressor: one frame per chunk, built from independent 128 KiB blocks
// that are compressed by one wavefront each.
//
// Output container is the one the reference produces with ZSTD_compress
// (src/cpu/compress.c:78-96): a single zstd frame per chunk, decodable by
// ZSTD_decompress. Block layout inside a chunk's output slot:
//   segment 0 = frame header (10 bytes) + block 0, segment b = block b,
// each segment at a fixed stride; the aggregate gather concatenates them.
//
// Per block (RFC 8878 3.1.1):
//  * greedy wave-parallel LZ77 parse: 64 candidate positions per step, a
//    6-byte key hashed into an LDS table, matches of >= 6 bytes inside the
//    block; after long literal runs the 64 lanes spread out (zstd-style
//    search acceleration: stride 1 + run/256), so incompressible stretches
//    cost little; literal runs are compacted to scratch while a single
//    256-bin LDS histogram counts them;
//  * RLE block when the parse finds one offset-1 match covering the block;
//  * literals: Huffman, 4 streams, table built per block (FSE-compressed or
//    direct weights); each stream is encoded by the whole wave (per-lane bit
//    lengths -> wave prefix sum -> LDS atomic-OR staging -> bytes copied out)
//    straight to its final position; the jump table is filled in after the
//    streams. Raw literals when Huffman does not pay;
//  * sequences: predefined FSE distributions;
//  * raw block whenever the compressed form is not smaller.
// The work list is a persistent grid pulling block indices from a device
// counter (one queue per launch).
#include "kernels.hpp"
#include "wave.hpp"
#include "zstd_core.hpp"

namespace chucky {
namespace {

using zcore::Seq;

constexpr uint32_t kBlock = zcore::kBlockMax;
constexpr uint64_t kSegStride = kBlock + 64; // >= frame header + block header + raw block
constexpr int kHashLog = 12;
constexpr uint32_t kMinMatch = 6;
constexpr uint32_t kMaxSeq = 16384;
constexpr uint64_t kLitsBytes = kBlock + 256;
constexpr uint64_t kSeqBytes = (uint64_t)kMaxSeq * sizeof(Seq);
constexpr uint64_t kSeqOutBytes = (uint64_t)kMaxSeq * 12 + 256;
constexpr uint64_t kWorkerScratch = (kLitsBytes + kSeqBytes + kSeqOutBytes + 255) & ~255ull;
constexpr int kStageWords = 104;
constexpr int kSymPerLane = 4;

struct Lds
{
  union
  {
    uint32_t table[1 << kHashLog];
    zcore::HufWork huf;
  } u;
  uint32_t hist[256];
  zcore::HufCode code;
  zcore::Predefined pre;
  uint8_t tree[192];
  uint32_t stage[kStageWords + 4];
  // decisions published by lane 0
  uint32_t mode, tree_size, lit_size, lit_hsize, seq_size;
};

__device__ __forceinline__ uint32_t key_hash(uint64_t key48)
{
  return (uint32_t)((key48 * 0xCF1BBCDCB7A56463ull) >> (64 - kHashLog));
}

// Ascending bitonic sort of 256 keys in LDS by one wavefront (two
// compare-exchange pairs per lane per stage).
__device__ void wave_sort256(uint32_t* keys)
{
  const uint32_t lane = (uint32_t)lane_id();
  for (uint32_t k = 2; k <= 256; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = lane; t < 128; t += kWave) {
        const uint32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        const uint32_t l = i | j;
        const bool asc = (i & k) == 0;
        const uint32_t a = keys[i], b = keys[l];
        if ((a > b) == asc) {
          keys[i] = b;
          keys[l] = a;
        }
      }
      __syncthreads();
    }
}

// Compacts in[0, n) to dst and adds every byte to the LDS histogram.
__device__ __forceinline__ void copy_count(uint8_t* dst, const uint8_t* src, uint32_t n,
                                           uint32_t* hist)
{
  wave_copy(dst, src, n);
  const int lane = lane_id();
  for (uint32_t i = 4u * lane; i < n; i += 4u * kWave) {
    const uint32_t w = ld32u(src + i);
    const uint32_t m = n - i < 4 ? n - i : 4;
    for (uint32_t k = 0; k < m; ++k)
      atomicAdd(&hist[(w >> (8 * k)) & 0xFF], 1u);
  }
}

// Encodes lits[a, b) as one backward Huffman stream at out; returns its size.
__device__ uint32_t huf_stream_wave(uint8_t* out, const uint8_t* lits, uint32_t a, uint32_t b,
                                    const zcore::HufCode& h, uint32_t* stage)
{
  const int lane = lane_id();
  for (int i = lane; i < kStageWords + 4; i += kWave)
    stage[i] = 0;
  __syncthreads();
  uint32_t flushed = 0; // bytes already copied out
  uint32_t rel = 0;     // bits held in stage (stage bit 0 = stream bit 8*flushed)
  long top = (long)b - 1;
  for (;;) {
    // lane L takes symbols top - 4L - 3 .. top - 4L (one dword), highest first
    const long lo = top - (long)lane * kSymPerLane - (kSymPerLane - 1);
    uint64_t bits = 0;
    uint32_t nb = 0;
    if (lo + kSymPerLane - 1 >= (long)a) {
      const uint32_t w = ld32u(lits + (lo >= (long)a ? lo : (long)a));
      const int skip = lo >= (long)a ? 0 : (int)((long)a - lo);
      for (int j = kSymPerLane - 1; j >= skip; --j) {
        const uint8_t s = (uint8_t)(w >> (8 * (j - skip)));
        bits |= (uint64_t)h.val[s] << nb;
        nb += h.len[s];
      }
    }
    const uint32_t incl = wave_incl_scan(nb);
    const uint32_t total = wave_bcast(incl, kWave - 1);
    if (nb) {
      const uint32_t p = rel + incl - nb;
      const uint32_t w = p >> 5, sh = p & 31;
      const uint32_t w0 = (uint32_t)(bits << sh);
      const uint32_t w1 = sh ? (uint32_t)(bits >> (32 - sh)) : (uint32_t)(bits >> 32);
      const uint32_t w2 = sh ? (uint32_t)(bits >> (64 - sh)) : 0u;
      atomicOr(&stage[w], w0);
      if (w1)
        atomicOr(&stage[w + 1], w1);
      if (w2)
        atomicOr(&stage[w + 2], w2);
    }
    __syncthreads();
    rel += total;
    top -= (long)kWave * kSymPerLane;
    if (top < (long)a) {
      if (lane == 0)
        stage[rel >> 5] |= 1u << (rel & 31); // end mark
      __syncthreads();
      const uint32_t nbytes = (rel + 1 + 7) >> 3;
      wave_copy(out + flushed, (const uint8_t*)stage, nbytes);
      __syncthreads();
      return flushed + nbytes;
    }
    const uint32_t full = rel >> 3;
    wave_copy(out + flushed, (const uint8_t*)stage, full);
    __syncthreads();
    const uint32_t carry = ((const uint8_t*)stage)[full];
    __syncthreads();
    for (int i = lane; i < kStageWords + 4; i += kWave)
      stage[i] = (i == 0) ? carry : 0u;
    __syncthreads();
    flushed += full;
    rel &= 7;
  }
}

// Compresses one block; returns bytes written at out (header included).
__device__ uint32_t compress_block(Lds& L, const uint8_t* in, uint32_t len, uint8_t* out, bool last,
                                   uint8_t* lits, Seq* seqs, uint8_t* seqout)
{
  const int lane = lane_id();

  // ---- parse + literal compaction + histogram ----
  for (int i = lane; i < (1 << kHashLog); i += kWave)
    L.u.table[i] = 0;
  for (int i = lane; i < 256; i += kWave)
    L.hist[i] = 0;
  __syncthreads();
  uint32_t nseq = 0, nlits = 0, anchor = 0;
  if (len >= 16) {
    const uint32_t last_start = len - kMinMatch; // a key needs 6 readable bytes
    uint32_t pos = 0;
    while (pos <= last_start && nseq < kMaxSeq) {
      const uint32_t stride = 1 + ((pos - anchor) >> 8);
      const uint32_t p = pos + (uint32_t)lane * stride;
      const bool active = p <= last_start;
      uint32_t v0 = 0, v1 = 0, h = 0, cand = 0;
      if (active) {
        v0 = ld32u(in + p);
        v1 = ld32u(in + p + 4) & 0xFFFFu;
        h = key_hash(((uint64_t)v1 << 32) | v0);
        cand = L.u.table[h];
      }
      // entries at or after p can be left by lanes past the previous match
      const bool hit = active && cand != 0 && cand - 1 < p && ld32u(in + cand - 1) == v0 &&
                       (ld32u(in + cand + 3) & 0xFFFFu) == v1;
      __builtin_amdgcn_wave_barrier();
      if (active)
        L.u.table[h] = p + 1;
      __builtin_amdgcn_wave_barrier();
      const uint64_t mask = wave_ballot(hit);
      if (mask == 0) {
        pos += kWave * stride;
        continue;
      }
      const int first = __ffsll((unsigned long long)mask) - 1;
      uint32_t mp = pos + (uint32_t)first * stride;
      uint32_t mc = wave_bcast(cand, first) - 1;
      uint32_t ml = wave_match_extend(in, mp, mc, kMinMatch, len);
      while (mp > anchor && mc > 0 && in[mp - 1] == in[mc - 1]) {
        --mp;
        --mc;
        ++ml;
      }
      const uint32_t ll = mp - anchor;
      copy_count(lits + nlits, in + anchor, ll, L.hist);
      nlits += ll;
      if (lane == 0)
        seqs[nseq] = Seq{ ll, ml, mp - mc };
      ++nseq;
      pos = mp + ml;
      anchor = pos;
      if (lane == 0 && pos >= 2 && pos - 2 <= last_start) {
        const uint32_t q = pos - 2;
        L.u.table[key_hash(((uint64_t)(ld32u(in + q + 4) & 0xFFFFu) << 32) | ld32u(in + q))] = q + 1;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  // one literal plus one offset-1 match to the end: every byte equals in[0]
  if (nseq == 1 && anchor == len) {
    uint32_t rle = 0;
    if (lane == 0)
      rle = seqs[0].ll == 1 && seqs[0].off == 1;
    if (wave_bcast(rle, 0)) {
      if (lane == 0) {
        zcore::write_block_header(out, last, 1, len);
        out[3] = in[0];
      }
      return 4;
    }
  }
  copy_count(lits + nlits, in + anchor, len - anchor, L.hist);
  nlits += len - anchor;
  __syncthreads();

  // ---- literal mode and Huffman code ----
  const uint32_t raw_h = nlits < 32 ? 1 : (nlits < 4096 ? 2 : 3);
  uint32_t used_local = 0;
  for (int s = lane; s < 256; s += kWave)
    used_local += L.hist[s] != 0;
  const uint32_t used = wave_sum(used_local);
  if (lane == 0) {
    L.mode = 0;
    L.lit_size = raw_h + nlits;
    L.tree_size = 0;
    if (nlits >= 256 && used == 1) {
      L.mode = 1;
      L.lit_size = raw_h + 1;
    }
  }
  if (nlits >= 256 && used > 1) {
    // used symbols in ascending count order: bitonic sort of (count, symbol)
    uint32_t* keys = L.u.huf.w;
    for (int s = lane; s < 256; s += kWave)
      keys[s] = (L.hist[s] << 8) | (uint32_t)s;
    __syncthreads();
    wave_sort256(keys);
    for (uint32_t i = lane; i < used; i += kWave)
      L.u.huf.sym[i] = (uint16_t)(keys[256 - used + i] & 0xFF);
    __syncthreads();
    if (lane == 0) {
      zcore::huf_build_sorted(L.code, L.hist, (int)used, L.u.huf);
      L.tree_size = (uint32_t)zcore::huf_write_tree(L.tree, L.code, L.u.huf);
    }
    __syncthreads();
    uint32_t bits_local = 0;
    for (int s = lane; s < 256; s += kWave)
      bits_local += L.hist[s] * L.code.len[s];
    const uint32_t bits = wave_sum(bits_local);
    if (lane == 0 && L.tree_size) {
      // streams: <= bits + 4 end marks, rounded up per stream; a stream holds
      // <= nlits/4+1 symbols of <= 11 bits, so 16-bit jump entries suffice
      const uint64_t comp_bound = L.tree_size + 6 + (bits + 4 * 8 + 7) / 8;
      const uint32_t hs = (uint32_t)zcore::huf_lits_header_size(nlits, nlits);
      if (comp_bound < nlits && hs + comp_bound < L.lit_size) {
        L.mode = 2;
        L.lit_hsize = hs;
        L.lit_size = hs + (uint32_t)comp_bound; // bound; exact size known after encoding
      }
    }
  }
  // ---- sequences section into scratch (lane 0) ----
  if (lane == 0)
    L.seq_size = (uint32_t)zcore::write_sequences(seqout, seqs, nseq, L.pre);
  __syncthreads();
  const uint32_t mode = L.mode, seq_size = L.seq_size;
  if (L.lit_size + seq_size >= len) {
    if (lane == 0)
      zcore::write_block_header(out, last, 0, len);
    wave_copy(out + 3, in, len);
    return 3 + len;
  }
  uint8_t* o = out + 3;
  uint32_t lit_size = L.lit_size;
  if (mode == 0) {
    uint32_t h = 0;
    if (lane == 0)
      h = (uint32_t)zcore::write_raw_lits_header(o, nlits, 0);
    h = wave_bcast(h, 0);
    wave_copy(o + h, lits, nlits);
  } else if (mode == 1) {
    if (lane == 0) {
      const size_t h = zcore::write_raw_lits_header(o, nlits, 1);
      o[h] = lits[0];
    }
  } else {
    const uint32_t hs = L.lit_hsize, ts = L.tree_size;
    if (lane == 0)
      for (uint32_t i = 0; i < ts; ++i)
        o[hs + i] = L.tree[i];
    const uint32_t seg = (nlits + 3) / 4;
    uint32_t sp = hs + ts + 6, sizes[4];
    for (int k = 0; k < 4; ++k) {
      const uint32_t a = (uint32_t)k * seg;
      const uint32_t b = k == 3 ? nlits : a + seg;
      sizes[k] = huf_stream_wave(o + sp, lits, a, b, L.code, L.stage);
      sp += sizes[k];
    }
    const uint32_t comp = sp - hs;
    if (lane == 0) {
      zcore::write_huf_lits_header(o, hs, nlits, comp);
      uint8_t* jt = o + hs + ts;
      for (int k = 0; k < 3; ++k) {
        jt[2 * k] = (uint8_t)sizes[k];
        jt[2 * k + 1] = (uint8_t)(sizes[k] >> 8);
      }
    }
    lit_size = hs + comp;
  }
  wave_copy(o + lit_size, seqout, seq_size);
  const uint32_t body = lit_size + seq_size;
  if (lane == 0)
    zcore::write_block_header(out, last, 2, body);
  return 3 + body;
}

__global__ __launch_bounds__(64) void zstd_k(const uint8_t* __restrict__ src, uint64_t src_stride,
                                             uint32_t chunk_bytes, uint32_t nseg,
                                             uint64_t total_blocks, uint8_t* __restrict__ dst,
                                             uint64_t chunk_dst_stride,
                                             uint32_t* __restrict__ seg_sizes, uint8_t* scratch,
                                             uint32_t* queue)
{
  __shared__ Lds L;
  const int lane = lane_id();
  if (lane == 0)
    zcore::build_predefined(L.pre, L.u.huf.spread, L.u.huf.cumul);
  __syncthreads();
  uint8_t* ws = scratch + (uint64_t)blockIdx.x * kWorkerScratch;
  uint8_t* lits = ws;
  Seq* seqs = (Seq*)(ws + kLitsBytes);
  uint8_t* seqout = ws + kLitsBytes + kSeqBytes;
  uint64_t item = blockIdx.x;
  for (;;) {
    if (queue) {
      uint32_t q = 0;
      if (lane == 0)
        q = atomicAdd(queue, 1u);
      item = wave_bcast(q, 0);
    }
    if (item >= total_blocks)
      break;
    const uint64_t chunk = item / nseg;
    const uint32_t b = (uint32_t)(item % nseg);
    const uint8_t* in = src + chunk * src_stride + (uint64_t)b * kBlock;
    const uint32_t len = (chunk_bytes - b * kBlock) < kBlock ? chunk_bytes - b * kBlock : kBlock;
    uint8_t* out = dst + chunk * chunk_dst_stride + (uint64_t)b * kSegStride;
    uint32_t hdr = 0;
    if (b == 0) {
      if (lane == 0)
        zcore::write_frame_header(out, chunk_bytes);
      hdr = (uint32_t)zcore::kFrameHeaderBytes;
    }
    const uint32_t sz = compress_block(L, in, len, out + hdr, b + 1 == nseg, lits, seqs, seqout);
    if (lane == 0)
      seg_sizes[chunk * nseg + b] = hdr + sz;
    __syncthreads();
    if (!queue)
      break;
  }
}

__global__ void zero_u32_k(uint32_t* p) { *p = 0; }

} // namespace

uint32_t zstd_segments(size_t chunk_bytes) { return (uint32_t)((chunk_bytes + kBlock - 1) / kBlock); }
uint64_t zstd_segment_stride() { return kSegStride; }
size_t zstd_max_output_size(size_t chunk_bytes) { return (size_t)zstd_segments(chunk_bytes) * kSegStride; }
size_t zstd_scratch_bytes(uint32_t workers) { return (size_t)workers * kWorkerScratch; }

hipError_t launch_zstd(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                       uint8_t* dst, uint64_t chunk_dst_stride, uint32_t* seg_sizes, int level,
                       void* scratch, uint32_t workers, uint32_t* queue_counter,
                       hipStream_t stream)
{
  (void)level;
  if (n == 0)
    return hipSuccess;
  if (chunk_bytes == 0 || chunk_bytes > 0xFFFFFFFFull)
    return hipErrorInvalidValue;
  const uint32_t nseg = zstd_segments(chunk_bytes);
  const uint64_t total = n * nseg;
  if (queue_counter) {
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(1), 0, stream, queue_counter);
    hipLaunchKernelGGL(zstd_k, dim3(workers), dim3(64), 0, stream, src, src_stride,
                       (uint32_t)chunk_bytes, nseg, total, dst, chunk_dst_stride, seg_sizes,
                       (uint8_t*)scratch, queue_counter);
  } else {
    if (total > workers)
      return hipErrorInvalidValue; // one scratch slice per block in this mode
    hipLaunchKernelGGL(zstd_k, dim3((uint32_t)total), dim3(64), 0, stream, src, src_stride,
                       (uint32_t)chunk_bytes, nseg, total, dst, chunk_dst_stride, seg_sizes,
                       (uint8_t*)scratch, nullptr);
  }
  return hipGetLastError();
}

} // namespace chucky
