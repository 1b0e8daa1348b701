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
//  * greedy wave-parallel LZ77 parse: 64 candidate positions per step, an
//    8-byte key hashed into an LDS table, matches of >= 8 bytes inside the
//    block; after long literal runs the 64 lanes spread out (zstd-style
//    search acceleration: stride 1 + run/256), so incompressible stretches
//    cost little. The 8-byte minimum is deliberate: on noisy integer data
//    (sensor frames, small-range integers) 6- and 7-byte matches cost more
//    in the sequence section than their Huffman-coded literals, and every
//    sequence is serial work; literal runs are compacted to scratch while a
//    single 256-bin LDS histogram counts them (a block without matches is
//    its own literal buffer);
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

#include <map>
#include <mutex>

namespace chucky {
namespace {

using zcore::Seq;

constexpr uint32_t kBlock = zcore::kBlockMax; // largest block (scratch is sized for it)
// segment = block + slack >= frame header + (chunk-table carrier block) +
// block header + raw block
constexpr uint64_t kSegSlack = 320;
// bytes of a chunk's first block that carry its Huffman table
constexpr uint32_t kCarrier = 64;
constexpr int kHashLog = 12;
constexpr uint32_t kMinMatch = 8;
constexpr uint32_t kLaneExt = 32; // bytes a lane extends its own match before the wave takes over
constexpr uint32_t kMaxSeq = 8192;
constexpr uint64_t kLitsBytes = kBlock + 256;
constexpr uint64_t kSeqBytes = (uint64_t)kMaxSeq * sizeof(Seq);
constexpr uint64_t kSeqOutBytes = (uint64_t)kMaxSeq * 12 + 256;
constexpr uint64_t kAuxBytes = (uint64_t)kMaxSeq * 4;
constexpr uint64_t kWorkerScratch = (kLitsBytes + kSeqBytes + kSeqOutBytes + kAuxBytes + 255) & ~255ull;
constexpr int kSymPerLane = 8; // Huffman symbols per lane per round
// one round: <= 64 lanes * 8 symbols * 11 bits, plus < 8 carried bits
constexpr int kStageWords = (kWave * kSymPerLane * zcore::kHufMaxBits + 8 + 31) / 32 + 4;
// The literal streams stage up to 4 rounds before copying complete bytes out.
constexpr int kRoundBits = kWave * kSymPerLane * zcore::kHufMaxBits;
constexpr int kStreamWords = (4 * kRoundBits + 8 + 31) / 32 + 4;
constexpr uint32_t kStreamFlushBits = (uint32_t)(kStreamWords - 4) * 32 - (uint32_t)kRoundBits;

// ~10.2 KiB per wavefront, so that 15 wavefronts fit a CU's LDS. The union
// is the parse's hash table, then the Huffman builder's scratch, then the
// literal encoder's packed codes and bit staging.
// Scratch of the package-merge code-length builder; overlays HufWork and
// keeps its symbol list (sym) in place.
struct PmWork
{
  uint16_t sym[256];
  // leaf weights ascending at [0, 512), packages at [512, 1024), each area
  // padded with ~0u so that a binary search needs no bounds
  uint32_t lp[1024];
  uint32_t list[512];
  uint32_t bits[zcore::kHufMaxBits][16]; // leaf positions in each level's list
};

struct Lds
{
  union
  {
    uint16_t table[1 << kHashLog]; // low 16 bits of the last position per hash
    zcore::HufWork huf;
    PmWork pm;
    struct
    {
      uint32_t cw[256]; // code value | length << 16
      uint32_t stage[kStreamWords]; // >= kStageWords
    } enc;
  } u;
  uint32_t hist[256];
  zcore::HufCode code;
  uint8_t tree[192];
  // decisions published by lane 0
  uint32_t mode, tree_size, lit_size, lit_hsize, seq_size;
};
static_assert(offsetof(PmWork, sym) == offsetof(zcore::HufWork, sym), "PmWork keeps HufWork's symbol list");
static_assert(sizeof(PmWork) <= sizeof(uint16_t) << kHashLog, "the builder fits the hash table's space");

// Huffman table shared by every block of a chunk (chunk-table mode).
struct ChunkTable
{
  uint32_t cw[256]; // code value | length << 16
  uint8_t tree[192];
  uint32_t tree_size; // 0 = no table: blocks use raw literals
  uint32_t pad[3];
};

__device__ __forceinline__ uint32_t key_hash(uint64_t key)
{
  return (uint32_t)((key * 0xCF1BBCDCB7A56463ull) >> (64 - kHashLog));
}

// Adds every byte of src[0, n) to the LDS histogram. Each lane keeps one
// (symbol, run) pair per byte position within a dword and only issues an LDS
// atomic when the symbol changes: the high bytes of small 16/32-bit integers
// (most often 0) would otherwise make every atomic of the wave hit one bin.
__device__ __forceinline__ void count_bytes_any(const uint8_t* src, uint32_t n, uint32_t* hist)
{
  const int lane = lane_id();
  uint32_t sym[4] = { 0, 0, 0, 0 }, run[4] = { 0, 0, 0, 0 };
  const uint32_t body = n & ~63u; // whole 16-byte pieces per lane
  for (uint32_t i = 16u * lane; i < body; i += 16u * kWave) {
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      w[j] = ld32u(src + i + 4 * j);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t b = (w[j] >> (8 * k)) & 0xFF;
        if (b == sym[k]) {
          ++run[k];
        } else {
          if (run[k])
            atomicAdd(&hist[sym[k]], run[k]);
          sym[k] = b;
          run[k] = 1;
        }
      }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (run[k])
      atomicAdd(&hist[sym[k]], run[k]);
  for (uint32_t i = body + lane; i < n; i += kWave)
    atomicAdd(&hist[src[i]], 1u);
}

template<int AS> using BytePtr = const __attribute__((address_space(AS))) uint8_t*;
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
template<int AS> using Vec4Ptr = const __attribute__((address_space(AS))) v4u32*;

// The same for a 16-byte aligned src in global memory (AS 1) or LDS (3): one
// 16-byte load per lane and round, the next round's in flight while this
// one's bytes are counted; the part below a whole round by count_bytes_any.
template<int AS>
__device__ __forceinline__ void count_bytes(BytePtr<AS> src, uint32_t n, uint32_t* hist)
{
  const int lane = lane_id();
  uint32_t sym[4] = { 0, 0, 0, 0 }, run[4] = { 0, 0, 0, 0 };
  const uint32_t body = n & ~1023u; // whole rounds of 16 bytes per lane
  v4u32 cur = body ? ((Vec4Ptr<AS>)src)[lane] : v4u32{ 0, 0, 0, 0 };
  for (uint32_t i = 16u * lane; i < body; i += 16u * kWave) {
    const uint32_t nx = i + 16u * kWave;
    const v4u32 nxt = ((Vec4Ptr<AS>)(src + (nx < body ? nx : i)))[0];
    const uint32_t w[4] = { cur.x, cur.y, cur.z, cur.w };
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t b = (w[j] >> (8 * k)) & 0xFF;
        if (b == sym[k]) {
          ++run[k];
        } else {
          if (run[k])
            atomicAdd(&hist[sym[k]], run[k]);
          sym[k] = b;
          run[k] = 1;
        }
      }
    cur = nxt;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (run[k])
      atomicAdd(&hist[sym[k]], run[k]);
  if (body < n)
    count_bytes_any((const uint8_t*)(src + body), n - body, hist);
}

// count_bytes for a whole workgroup (blockDim.x threads).
__device__ __forceinline__ void count_bytes_block(const uint8_t* src, uint64_t n, uint32_t* hist)
{
  uint32_t sym[4] = { 0, 0, 0, 0 }, run[4] = { 0, 0, 0, 0 };
  const uint64_t body = n & ~(uint64_t)(16 * 256 - 1);
  for (uint64_t i = 16ull * threadIdx.x; i < body; i += 16ull * blockDim.x) {
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      w[j] = ld32u(src + i + 4 * j);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t b = (w[j] >> (8 * k)) & 0xFF;
        if (b == sym[k]) {
          ++run[k];
        } else {
          if (run[k])
            atomicAdd(&hist[sym[k]], run[k]);
          sym[k] = b;
          run[k] = 1;
        }
      }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (run[k])
      atomicAdd(&hist[sym[k]], run[k]);
  for (uint64_t i = body + threadIdx.x; i < n; i += blockDim.x)
    atomicAdd(&hist[src[i]], 1u);
}

// The raw dwords holding the up-to-8 symbols lane `lane` encodes in the round
// whose highest symbol is `top` (symbols top-8*lane-7 .. top-8*lane; those
// below `a` are skipped). The load is unconditional (a lane with nothing to
// encode reads at `a`) and its bytes are only aligned when the round uses
// them, so it stays in flight over the previous round.
struct RoundLoad
{
  uint32_t w0, w1, w2, sh;
  int skip;
};

// The literals' memory, named for the compiler: global (address space 1) or
// LDS (3), so that the load is a global or an LDS load and not a flat one (a
// flat load counts against the LDS counter too, so waiting for the round's
// table lookups would also wait for it).
template<int AS> using LitPtr = const __attribute__((address_space(AS))) uint8_t*;
template<int AS> using LitWord = const __attribute__((address_space(AS))) uint32_t*;

template<int AS>
__device__ __forceinline__ RoundLoad round_load(LitPtr<AS> lits, long a, long top, int lane)
{
  const long lo = top - (long)lane * kSymPerLane - (kSymPerLane - 1);
  const int sk = lo + kSymPerLane - 1 < a ? kSymPerLane : (lo >= a ? 0 : (int)(a - lo));
  const long at = sk >= kSymPerLane ? a : lo + sk;
  const uint32_t mis = (uint32_t)(uintptr_t)(lits + at) & 3u;
  const LitWord<AS> w = (LitWord<AS>)(lits + (at - (long)mis));
  return RoundLoad{ w[0], w[1], w[2], mis * 8, sk };
}

// byte j of the result = symbol lo + j; skipped symbols are zero
__device__ __forceinline__ uint64_t round_syms(const RoundLoad& r)
{
  const uint32_t lo = __builtin_amdgcn_alignbit(r.w1, r.w0, r.sh);
  const uint32_t hi = __builtin_amdgcn_alignbit(r.w2, r.w1, r.sh);
  const uint64_t v = ((uint64_t)hi << 32) | lo;
  return r.skip >= kSymPerLane ? 0ull : v << (8 * r.skip);
}

// Encodes lits[a, b) as one backward Huffman stream at out; returns its size.
// Each round the wave takes 512 symbols, highest first: every lane packs its
// 8 codes (at most 88 bits) into two 64-bit accumulators, a wave prefix sum
// of the bit counts places the lanes' words, funnel shifts align them and
// LDS atomic-ORs assemble them in `stage` (a lane ORs all four words, zero
// ones included, so no lane branches); complete bytes are copied out once
// the stage could not take another round. The next round's symbols are
// loaded while this round's codes are looked up and placed.
template<int AS>
__device__ uint32_t huf_stream_wave(uint8_t* out, LitPtr<AS> lits, uint32_t a, uint32_t b,
                                    const uint32_t* cw, uint32_t* stage)
{
  static_assert(5 * zcore::kHufMaxBits <= 64, "five codes fit one accumulator");
  const int lane = lane_id();
  for (int i = lane; i < kStreamWords; i += kWave)
    stage[i] = 0;
  __syncthreads();
  uint32_t flushed = 0; // bytes already copied out
  uint32_t rel = 0;     // bits held in stage (stage bit 0 = stream bit 8*flushed)
  long top = (long)b - 1;
  RoundLoad cur = round_load<AS>(lits, a, top, lane);
  for (;;) {
    const long next = top - (long)kWave * kSymPerLane;
    const RoundLoad nxt = round_load<AS>(lits, a, next >= (long)a ? next : (long)a - 1, lane);
    const uint64_t q = round_syms(cur);
    const int skip = cur.skip;
    // every lookup is made (a skipped symbol reads the entry of 0) and its
    // code dropped afterwards: no exec-mask changes around the LDS reads
    uint32_t c[kSymPerLane];
#pragma unroll
    for (int j = 0; j < kSymPerLane; ++j)
      c[j] = cw[(uint32_t)(q >> (8 * j)) & 0xFF];
#pragma unroll
    for (int j = 0; j < kSymPerLane; ++j)
      c[j] = j >= skip ? c[j] : 0u;
    // symbols 7..3 into A, 2..0 into B (a skipped symbol adds no bits)
    uint64_t A = 0, B = 0;
    uint32_t na = 0, nbb = 0;
#pragma unroll
    for (int j = kSymPerLane - 1; j >= 3; --j) {
      A |= (uint64_t)(c[j] & 0xFFFF) << na;
      na += c[j] >> 16;
    }
#pragma unroll
    for (int j = 2; j >= 0; --j) {
      B |= (uint64_t)(c[j] & 0xFFFF) << nbb;
      nbb += c[j] >> 16;
    }
    const uint32_t nb = na + nbb;
    // the lane's bits: lo | hi << 64, hi below 2^24
    const uint64_t lo = A | (B << na);
    const uint32_t hi = na ? (uint32_t)(B >> (64 - na)) : 0u;
    const uint32_t incl = wave_incl_scan(nb);
    const uint32_t total = wave_bcast(incl, kWave - 1);
    {
      // rel + total stays below the flush bound plus one round: the four
      // words lie inside the stage for every lane, nb = 0 included
      const uint32_t p = rel + incl - nb;
      const uint32_t sh = p & 31;
      const uint32_t v0 = (uint32_t)lo, v1 = (uint32_t)(lo >> 32);
      const uint32_t w0 = v0 << sh;
      const uint32_t w1 = sh ? __builtin_amdgcn_alignbit(v1, v0, 32 - sh) : v1;
      const uint32_t w2 = sh ? __builtin_amdgcn_alignbit(hi, v1, 32 - sh) : hi;
      const uint32_t w3 = sh ? hi >> (32 - sh) : 0u;
      uint32_t* w = stage + (p >> 5);
      atomicOr(&w[0], w0);
      atomicOr(&w[1], w1);
      atomicOr(&w[2], w2);
      atomicOr(&w[3], w3);
    }
    rel += total;
    if (next < (long)a) {
      __syncthreads();
      if (lane == 0)
        stage[rel >> 5] |= 1u << (rel & 31); // end mark
      __syncthreads();
      const uint32_t nbytes = (rel + 1 + 7) >> 3;
      wave_copy(out + flushed, (const uint8_t*)stage, nbytes);
      __syncthreads();
      return flushed + nbytes;
    }
    if (rel > kStreamFlushBits) {
      __syncthreads();
      const uint32_t full = rel >> 3;
      wave_copy(out + flushed, (const uint8_t*)stage, full);
      const uint32_t carry = ((const uint8_t*)stage)[full];
      __syncthreads();
      const uint32_t used_words = (rel + 31) >> 5;
      for (uint32_t i = lane; i < used_words; i += kWave)
        stage[i] = (i == 0) ? (carry & ((1u << (rel & 7)) - 1u)) : 0u;
      __syncthreads();
      flushed += full;
      rel &= 7;
    }
    top = next;
    cur = nxt;
  }
}

// The n >= 2 used symbols in ascending count order, ties by symbol, into
// L.u.pm.sym[0..n) and their counts into L.u.pm.lp[0..n): a bitonic sort of
// the 256 keys (count << 8 | symbol) held in registers, element 64r + lane in
// register r; partners in another register are compared within the lane,
// partners in another lane are exchanged by cross-lane shuffles. Unused
// symbols (count 0) sort first and are dropped. Counts stay below 2^24.
__device__ void huf_sort_wave(Lds& L, uint32_t n)
{
  const uint32_t lane = (uint32_t)lane_id();
  uint32_t v[4];
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t sy = 64u * r + lane;
    v[r] = (L.hist[sy] << 8) | sy;
  }
#pragma unroll
  for (uint32_t k = 2; k <= 256; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= (uint32_t)kWave) {
        // partner in register r ^ (j / 64); k > j >= 64, so the direction
        // depends on the register only
        const uint32_t rj = j / kWave;
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
          if (r & rj)
            continue;
          const bool asc = ((64u * r) & k) == 0;
          const uint32_t a = v[r], b = v[r | rj];
          const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
          v[r] = asc ? lo : hi;
          v[r | rj] = asc ? hi : lo;
        }
      } else {
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
          const uint32_t y = (uint32_t)__shfl_xor((int)v[r], (int)j, kWave);
          const bool asc = ((64u * r + lane) & k) == 0;
          const bool keep_min = ((lane & j) == 0) == asc;
          const uint32_t lo = v[r] < y ? v[r] : y, hi = v[r] < y ? y : v[r];
          v[r] = keep_min ? lo : hi;
        }
      }
    }
  }
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t e = 64u * r + lane;
    if (e >= 256u - n) {
      L.u.pm.sym[e - (256u - n)] = (uint16_t)(v[r] & 0xFF);
      L.u.pm.lp[e - (256u - n)] = v[r] >> 8;
    }
  }
  wave_sync();
}

// Huffman code lengths (at most kHufMaxBits, a complete prefix code) and
// canonical values for the n >= 2 used symbols in L.u.pm.sym[0..n), counts in
// L.u.pm.lp[0..n) ascending, by the package-merge algorithm (Larmore and
// Hirschberg, "A fast algorithm for optimal length-restricted Huffman codes",
// 1990) run by the whole wave. Level kHufMaxBits is the leaves alone; level
// d < kHufMaxBits merges the leaves with the packages (sums of adjacent
// pairs) of level d + 1, every item ranked in the other sorted array by a
// binary search (ties: leaves first). The first 2n - 2 items of level 1 are
// selected; the packages among the first m items of a level select the first
// 2 (m - leaves) items of the next, so each level only records where its
// leaves sit, and a leaf's code length is the number of levels that select
// it. The code is optimal among codes of at most kHufMaxBits bits.
__device__ void huf_build_wave(Lds& L, uint32_t n_in)
{
  const uint32_t lane = (uint32_t)lane_id();
  constexpr uint32_t Lmax = zcore::kHufMaxBits;
  PmWork& pm = L.u.pm;
  const uint32_t n = uniform(n_in);
  for (uint32_t k = n + lane; k < 512; k += kWave)
    pm.lp[k] = ~0u;
  for (uint32_t k = lane; k < 512; k += kWave)
    pm.lp[512 + k] = ~0u;
  wave_sync();
  uint32_t items = n; // size of the previous level's list
  bool from_leaves = true;
  for (uint32_t d = Lmax - 1; d >= 1; --d) {
    // the package count never shrinks from level to level (it stays below
    // n), so the padding above it is never overwritten
    const uint32_t np = items >> 1;
    for (uint32_t k = lane; k < np; k += kWave)
      pm.lp[512 + k] = from_leaves ? pm.lp[2 * k] + pm.lp[2 * k + 1] : pm.list[2 * k] + pm.list[2 * k + 1];
    if (lane < 16)
      pm.bits[d][lane] = 0;
    wave_sync();
    const uint32_t total = n + np; // <= 511: at most 8 items per lane
    uint32_t thr[8], off[8], c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t t = lane + (uint32_t)(kWave * q);
      const bool leaf = t < n;
      const uint32_t w = pm.lp[leaf ? t : 512 + (t - n)];
      // a leaf counts the packages lighter than it, a package the leaves
      // at most as heavy (padding items are never counted)
      thr[q] = leaf ? w : w + 1;
      off[q] = leaf ? 512u : 0u;
      c[q] = 0;
    }
#pragma unroll
    for (uint32_t step = 256; step; step >>= 1) {
      uint32_t v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        v[q] = pm.lp[off[q] + c[q] + step - 1];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        c[q] += v[q] < thr[q] ? step : 0u;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t t = lane + (uint32_t)(kWave * q);
      if (t < total) {
        const bool leaf = t < n;
        const uint32_t pos = (leaf ? t : t - n) + c[q];
        pm.list[pos] = leaf ? thr[q] : thr[q] - 1;
        if (leaf)
          atomicOr(&pm.bits[d][pos >> 5], 1u << (pos & 31));
      }
    }
    wave_sync();
    items = total;
    from_leaves = false;
  }
  // selection: m items of level d hold a_d leaves and select 2 (m - a_d)
  // items of level d + 1; level kHufMaxBits holds leaves only
  uint32_t a[Lmax + 1];
  uint32_t m = 2 * n - 2;
#pragma unroll
  for (uint32_t d = 1; d < Lmax; ++d) {
    uint32_t cnt = 0;
    if (lane < 16) {
      const uint32_t lo = 32u * lane;
      const uint32_t mk = m <= lo ? 0u : (m - lo >= 32 ? ~0u : (1u << (m - lo)) - 1u);
      cnt = (uint32_t)__popc(pm.bits[d][lane] & mk);
    }
    a[d] = uniform(wave_sum(cnt));
    m = 2 * (m - a[d]);
  }
  a[Lmax] = m;
  uint32_t len[4];
  uint32_t maxd = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t idx = (uint32_t)k * kWave + lane;
    uint32_t l = 0;
#pragma unroll
    for (uint32_t d = 1; d <= Lmax; ++d)
      l += idx < a[d] ? 1u : 0u;
    len[k] = idx < n ? l : 0u;
    maxd = len[k] > maxd ? len[k] : maxd;
  }
  maxd = uniform(wave_max(maxd));
  wave_sync();
  // scatter lengths to symbols
  for (uint32_t s = lane; s < 256; s += kWave)
    L.code.len[s] = 0;
  wave_sync();
  uint32_t maxsym = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t idx = (uint32_t)k * kWave + lane;
    if (idx < n) {
      const uint32_t s = L.u.huf.sym[idx];
      L.code.len[s] = (uint8_t)len[k];
      maxsym = s > maxsym ? s : maxsym;
    }
  }
  maxsym = wave_max(maxsym);
  wave_sync();
  // canonical values: per length, consecutive in symbol order, starting
  // values from the longest length down
  uint32_t sl[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    sl[k] = L.code.len[(uint32_t)k * kWave + lane];
  uint32_t val[4] = { 0, 0, 0, 0 };
  uint32_t minv = 0;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (uint32_t l = maxd; l >= 1; --l) {
    uint32_t rank = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t m = wave_ballot(sl[k] == l);
      if (sl[k] == l)
        val[k] = minv + rank + (uint32_t)__popcll(m & below);
      rank += (uint32_t)__popcll(m);
    }
    minv = (minv + rank) >> 1;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    L.code.val[(uint32_t)k * kWave + lane] = (uint16_t)val[k];
  if (lane == 0) {
    L.code.max_bits = (int)maxd;
    L.code.max_sym = (int)maxsym;
  }
  wave_sync();
}

// Huffman code lengths (at most kHufMaxBits, a complete prefix code) and
// canonical values for a block's literals, counts in L.hist (at least two
// used symbols, every count below 2^17), without sorting: each lane holds
// symbols 64r + lane, r < 4.
//  1. Shannon lengths l = ceil(log2(M / c)) capped at kHufMaxBits, with
//     M = N (the literal count) when no symbol falls below the cap, else
//     M = N + ceil(N / 7): then the uncapped terms of the Kraft sum add up
//     to at most 7/8 and the (at most 256) capped ones to at most 1/8, so
//     the code is never over-full.
//  2. The Kraft sum is filled to exactly 1 greedily, shortening the codes
//     with the largest count per unit of Kraft space they take (count * 2^l;
//     shortening a length-l code takes 2^(11-l) units of 2^-11). A round
//     takes the best code's length class at once: every code of that class
//     whose ratio is still at least the best of the other classes, as many
//     as fit, in symbol order. A round always finds a code that fits (the
//     remaining space is a multiple of the longest code's unit), so the loop
//     ends with a complete code.
// Measured against package-merge (optimal length-limited codes) on 8 KiB
// blocks of imaging data, skewed and uniform synthetic bytes: 0.00-0.4% more
// literal bits, in 2-45 rounds (about 14 on imaging data). Returns false
// (no code) only if the fill found no code to shorten, which the argument
// above rules out.
__device__ bool huf_build_fast_wave(Lds& L)
{
  const uint32_t lane = (uint32_t)lane_id();
  constexpr uint32_t Lmax = zcore::kHufMaxBits;
  constexpr uint32_t kFull = 1u << Lmax;
  uint32_t c[4], l[4];
  uint32_t nloc = 0, rare = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    c[r] = L.hist[64u * r + lane];
    nloc += c[r];
  }
  const uint32_t N = wave_sum(nloc);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    rare |= (c[r] && (c[r] << Lmax) < N) ? 1u : 0u;
  const uint32_t M = wave_ballot(rare != 0) ? N + (N + 6) / 7 : N;
  const uint32_t clz_m = (uint32_t)__builtin_clz(M);
  uint32_t kloc = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    uint32_t v = 0;
    if (c[r]) {
      const uint32_t l0 = (uint32_t)__builtin_clz(c[r]) - clz_m; // c <= M
      v = (c[r] << l0) >= M ? l0 : l0 + 1;
      v = v < 1 ? 1 : (v > Lmax ? Lmax : v);
      kloc += kFull >> v;
    }
    l[r] = v;
  }
  uint32_t K = wave_sum(kloc);
  const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  while (K < kFull) {
    const uint32_t R = kFull - K;
    uint32_t key[4], kmax = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      key[r] = (l[r] > 1 && (kFull >> l[r]) <= R) ? c[r] << l[r] : 0u; // c < 2^17: below 2^28
      kmax = key[r] > kmax ? key[r] : kmax;
    }
    kmax = wave_max(kmax);
    if (kmax == 0)
      return false;
    // the class of the best ratio (the shortest, when classes tie)
    uint32_t inv = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      inv = key[r] == kmax && 16u - l[r] > inv ? 16u - l[r] : inv;
    const uint32_t ls = 16u - wave_max(inv);
    uint32_t k2 = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      k2 = l[r] != ls && key[r] > k2 ? key[r] : k2;
    k2 = wave_max(k2);
    const uint32_t m = R >> (Lmax - ls); // codes of the class that fit
    uint32_t base = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const bool e = key[r] != 0 && l[r] == ls && key[r] >= k2;
      const uint64_t b = wave_ballot(e);
      if (e && base + (uint32_t)__popcll(b & below) < m)
        l[r] -= 1;
      base += (uint32_t)__popcll(b);
    }
    K += (base < m ? base : m) << (Lmax - ls);
  }
  uint32_t maxd = 0, maxsym = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    L.code.len[64u * r + lane] = (uint8_t)l[r];
    maxd = l[r] > maxd ? l[r] : maxd;
    maxsym = c[r] ? 64u * r + lane : maxsym;
  }
  maxd = wave_max(maxd);
  maxsym = wave_max(maxsym);
  // canonical values: per length, consecutive in symbol order, starting
  // values from the longest length down
  uint32_t val[4] = { 0, 0, 0, 0 };
  uint32_t minv = 0;
  for (uint32_t d = maxd; d >= 1; --d) {
    uint32_t rank = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint64_t b = wave_ballot(l[r] == d);
      if (l[r] == d)
        val[r] = minv + rank + (uint32_t)__popcll(b & below);
      rank += (uint32_t)__popcll(b);
    }
    minv = (minv + rank) >> 1;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
    L.code.val[64u * r + lane] = (uint16_t)val[r];
  if (lane == 0) {
    L.code.max_bits = (int)maxd;
    L.code.max_sym = (int)maxsym;
  }
  wave_sync();
  return true;
}

// Huffman tree description (RFC 8878 4.2.1.1) of L.code into L.tree by the
// whole wave; the same bytes as zcore::huf_write_tree. Weights and their
// histogram come from ballots; the FSE table of the weights (table log 6) is
// built in parallel: cell j of the symbol spread is step*j mod 64 (step odd,
// so every cell once; cells above `high` hold the low-probability symbols),
// state-table ranks come from ballots; the two interleaved state chains then
// run in scalar registers over the weights with the tables one entry per
// lane. Normalization and the table header stay on lane 0 (16 symbols).
// Returns the description's size, 0 when no description fits (raw literals).
__device__ uint32_t huf_write_tree_wave(Lds& L)
{
  const int lane = lane_id();
  zcore::HufWork& ws = L.u.huf;
  const int nw = (int)uniform((uint32_t)L.code.max_sym); // the last symbol's weight is implied
  const int mb = (int)uniform((uint32_t)L.code.max_bits);
  constexpr uint32_t kNone = 16;
  uint32_t wt[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int s = k * kWave + lane;
    const uint32_t len = s < nw ? L.code.len[s] : 0u;
    wt[k] = s < nw ? (len ? (uint32_t)(mb + 1) - len : 0u) : kNone;
    if (s < nw)
      ws.wt[s] = (uint8_t)wt[k];
  }
  // weight histogram: lane w holds the count of weight w
  uint32_t cnt_lane = 0;
#pragma unroll
  for (uint32_t w = 0; w < 16; ++w) {
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      c += (uint32_t)__popcll(wave_ballot(wt[k] == w));
    if ((uint32_t)lane == w)
      cnt_lane = c;
  }
  const uint64_t present = wave_ballot(lane < 16 && cnt_lane > 0);
  const int distinct = __popcll(present);
  const int maxw = present ? 63 - __clzll(present) : 0;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t fse_size = 0;
  if (nw > 2 && distinct >= 2) {
    if (lane < 16)
      ws.wcount[lane] = cnt_lane;
    wave_sync();
    if (lane == 0)
      zcore::fse_normalize(ws.wnorm, ws.wcount, maxw, (uint32_t)nw, 6);
    wave_sync();
    const bool sym_lane = lane <= maxw;
    const int nrm = sym_lane ? ws.wnorm[lane] : 0;
    // cumul[s]: cells before symbol s (a low-probability symbol takes one)
    const uint32_t cells = nrm == -1 ? 1u : (nrm > 0 ? (uint32_t)nrm : 0u);
    const uint32_t cum = wave_incl_scan(cells) - cells;
    const uint32_t nc = nrm > 0 ? (uint32_t)nrm : 0u;
    const uint32_t n_incl = wave_incl_scan(nc), n_excl = n_incl - nc;
    const uint64_t lowm = wave_ballot(sym_lane && nrm == -1);
    const uint32_t high = 63u - (uint32_t)__popcll(lowm);
    // spread: lane j takes the j-th position of the stepping sequence
    const uint32_t pos = ((uint32_t)lane * 43u) & 63u; // step = 32 + 8 + 3
    const uint64_t validm = wave_ballot(pos <= high);
    if (pos <= high) {
      const uint32_t t = (uint32_t)__popcll(validm & below);
      uint32_t s = 0;
      for (int q = 0; q <= maxw; ++q) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)n_excl, q);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)n_incl, q);
        if (t >= lo && t < hi)
          s = (uint32_t)q;
      }
      ws.spread[pos] = (uint8_t)s;
    }
    if ((uint32_t)lane > high) {
      // cells from the top down take the low-probability symbols in order
      uint64_t m = lowm;
      for (uint32_t r = 63u - (uint32_t)lane; r > 0; --r)
        m &= m - 1;
      ws.spread[lane] = (uint8_t)(__ffsll((unsigned long long)m) - 1);
    }
    wave_sync();
    // state table: cell u is entry cumul[s] + (cells of s before u)
    const uint32_t su = ws.spread[lane];
    uint32_t rank = 0;
    for (int q = 0; q <= maxw; ++q) {
      const uint64_t m = wave_ballot(su == (uint32_t)q);
      if (su == (uint32_t)q)
        rank = (uint32_t)__popcll(m & below);
    }
    const uint32_t base = (uint32_t)__shfl((int)cum, (int)su, kWave);
    wave_sync();
    ws.ct.state[base + rank] = (uint16_t)(64u + (uint32_t)lane);
    // symbol transforms, lane s
    uint32_t dnb = 0;
    int dfs = 0;
    if (nrm == 0) {
      dnb = (7u << 16) - 64u;
    } else if (nrm == -1 || nrm == 1) {
      dnb = (6u << 16) - 64u;
      dfs = (int)cum - 1;
    } else if (nrm > 1) {
      const uint32_t mbo = 6u - (uint32_t)zcore::highbit32((uint32_t)nrm - 1);
      dnb = (mbo << 16) - ((uint32_t)nrm << mbo);
      dfs = (int)cum - nrm;
    }
    const uint32_t symv = dnb | ((uint32_t)(dfs + 128) << 20);
    wave_sync();
    const uint32_t stv = ws.ct.state[lane];
    uint32_t hsz = 0;
    if (lane == 0)
      hsz = (uint32_t)zcore::fse_write_ncount(ws.fse_buf, ws.wnorm, maxw, 6);
    hsz = wave_bcast(hsz, 0);
    // the weights' two-state stream, backward, even positions on state `se`.
    // Every weight's symbol transform is gathered first (one lane-indexed
    // shuffle per register) and split into its two fields, so a step of the
    // state chains -- which run in scalar registers -- reads two lanes for
    // the transform and one for its next state, and records only the state it
    // leaves (a lane write into the lane of the weight's own index, no LDS
    // store, no exec mask); the (bits, count) records follow from the states
    // afterwards, one lane per weight.
    auto rl = [](uint32_t v, uint32_t i) -> uint32_t {
      return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)i);
    };
    uint32_t ed[4], ef[4], rec[4] = { 0u, 0u, 0u, 0u };
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t e = (uint32_t)__shfl((int)symv, (int)(wt[k] & 63u), kWave);
      ed[k] = e & 0xFFFFFu;
      ef[k] = (e >> 20) - 128u; // signed, modulo 2^32
    }
    auto init = [&](int i) -> uint32_t {
      const int k = i >> 6;
      const uint32_t d = rl(k == 0 ? ed[0] : k == 1 ? ed[1] : k == 2 ? ed[2] : ed[3], (uint32_t)(i & 63));
      const uint32_t f = rl(k == 0 ? ef[0] : k == 1 ? ef[1] : k == 2 ? ef[2] : ef[3], (uint32_t)(i & 63));
      const uint32_t nb = (d + (1u << 15)) >> 16;
      const uint32_t v = (nb << 16) - d;
      return rl(stv, (v >> nb) + f);
    };
    int i = nw - 1;
    uint32_t se, so;
    if (i & 1) {
      so = init(i);
      se = init(i - 1);
    } else {
      se = init(i);
      so = init(i - 1);
    }
    i -= 2;
    const int top = i; // records at weight indices top .. 0, in stream order
    // one step of a state chain over weight index `wi`: the state it leaves
    // into lane wi % 64 of rk, the next state returned
    auto step = [&](uint32_t dk, uint32_t fk, uint32_t& rk, int wi, uint32_t st) -> uint32_t {
      const uint32_t li = (uint32_t)(wi & 63);
      const uint32_t d = rl(dk, li), f = rl(fk, li);
      const uint32_t nb = (st + d) >> 16;
      // (one scalar source per instruction on this target: the lane goes in m0)
      asm volatile("v_writelane_b32 %0, %1, m0" : "+v"(rk) : "s"(st), "{m0}"(li));
      return rl(stv, (st >> nb) + f);
    };
    // Odd indices run on `so`, even ones on `se`. After a first even step
    // (when top is even) the steps go in (odd, even) pairs, whose two chains
    // are independent and overlap; a pair never straddles a register (the
    // 64-index boundaries are even).
#pragma unroll
    for (int k = 3; k >= 0; --k) {
      const uint32_t dk = ed[k], fk = ef[k];
      uint32_t rk = rec[k];
      if (i >= 64 * k && !(i & 1)) {
        se = step(dk, fk, rk, i, se);
        --i;
      }
      for (; i >= 64 * k + 1; i -= 2) {
        const uint32_t no = step(dk, fk, rk, i, so);
        se = step(dk, fk, rk, i - 1, se);
        so = no;
      }
      // the record of each weight: (bits << 3) | count from the state it left
      const uint32_t nb = (rk + ed[k]) >> 16;
      rec[k] = ((rk & ((1u << nb) - 1u)) << 3) | nb;
    }
    const uint32_t nsteps = (uint32_t)(top + 1);
    // pack: lane l takes steps 4l .. 4l+3 (nsteps <= 253), step j being the
    // record of weight index top - j
    uint32_t bits = 0, nb = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t sj = 4u * (uint32_t)lane + (uint32_t)q;
      const int wi = top - (int)sj;
      const int src = wi & 63;
      uint32_t r = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t v = (uint32_t)__shfl((int)rec[k], src, kWave);
        r = (wi >> 6) == k ? v : r;
      }
      if (sj < nsteps) {
        bits |= (r >> 3) << nb;
        nb += r & 7;
      }
    }
    const uint32_t incl = wave_incl_scan(nb);
    const uint32_t total = wave_bcast(incl, kWave - 1);
    uint32_t* stage = (uint32_t*)ws.spread; // 128 words
    for (int q = lane; q < 128; q += kWave)
      stage[q] = 0;
    wave_sync();
    if (nb) {
      const uint32_t p = incl - nb;
      const uint64_t sh = (uint64_t)bits << (p & 31);
      atomicOr(&stage[p >> 5], (uint32_t)sh);
      if ((uint32_t)(sh >> 32))
        atomicOr(&stage[(p >> 5) + 1], (uint32_t)(sh >> 32));
    }
    wave_sync();
    if (lane == 0) {
      // final states and the end mark
      uint64_t tail = (uint64_t)(so & 63u) | ((uint64_t)(se & 63u) << 6) | (1ull << 12);
      const uint32_t p = total;
      tail <<= (p & 31);
      stage[p >> 5] |= (uint32_t)tail;
      if ((uint32_t)(tail >> 32))
        stage[(p >> 5) + 1] |= (uint32_t)(tail >> 32);
    }
    wave_sync();
    const uint32_t nout = (total + 13 + 7) / 8;
    const uint32_t cap = 127u - hsz; // a description must stay below 128 bytes
    const bool over = nout > cap;
    if (!over) {
      uint8_t* dst = ws.fse_buf + hsz;
      for (uint32_t q = lane; q < nout; q += kWave)
        dst[q] = ((const uint8_t*)stage)[q];
    }
    if (!over)
      fse_size = hsz + nout;
  }
  wave_sync();
  const bool direct_ok = nw <= 128;
  const uint32_t direct_size = 1u + (uint32_t)(nw + 1) / 2;
  if (fse_size > 1 && fse_size < 128 && (!direct_ok || fse_size + 1 < direct_size)) {
    if (lane == 0)
      L.tree[0] = (uint8_t)fse_size;
    for (uint32_t k = lane; k < fse_size; k += kWave)
      L.tree[1 + k] = ws.fse_buf[k];
    wave_sync();
    return fse_size + 1;
  }
  if (!direct_ok)
    return 0;
  if (lane == 0)
    L.tree[0] = (uint8_t)(127 + nw);
  for (int k = lane; 2 * k < nw; k += kWave) {
    const uint32_t a = ws.wt[2 * k], b = 2 * k + 1 < nw ? ws.wt[2 * k + 1] : 0u;
    L.tree[1 + k] = (uint8_t)((a << 4) | b);
  }
  wave_sync();
  return direct_size;
}

// The block's literals (every byte outside the matches, in order) into lits,
// with their byte histogram. Per 64 KiB of the block: the matches mark their
// bytes in an LDS bitmap (the parse's hash table space, free by now), then
// each round every lane takes 16 input bytes and their 16 bitmap bits, a
// wave prefix sum of the unmarked counts places them, and the next round's
// bytes are loaded before this round's are stored.
__device__ void gather_literals(const uint8_t* in, uint32_t len, const Seq* seqs, uint32_t nseq,
                                uint32_t* bitmap, uint8_t* lits, uint32_t* hist)
{
  constexpr uint32_t kPart = 65536; // bytes one bitmap covers
  static_assert(kPart / 8 <= sizeof(uint16_t) << kHashLog, "bitmap fits the hash table space");
  const uint32_t lane = (uint32_t)lane_id();
  __threadfence_block();
  uint32_t out = 0;
  for (uint32_t base = 0; base < len; base += kPart) {
    const uint32_t n = min(len - base, kPart);
    const uint32_t words = (n + 31) / 32;
    for (uint32_t w = lane; w < words; w += kWave)
      bitmap[w] = 0;
    wave_sync();
    for (uint32_t i = lane; i < nseq; i += kWave) {
      const Seq sq = seqs[i];
      uint32_t lo = sq.pos, e = sq.pos + sq.ml;
      if (e <= base || lo >= base + n)
        continue;
      lo = lo > base ? lo - base : 0u;
      e = min(e - base, n);
      while (lo < e) {
        const uint32_t w = lo >> 5, bit = lo & 31;
        const uint32_t c = min(32u - bit, e - lo);
        const uint32_t m = (c == 32 ? ~0u : ((1u << c) - 1u)) << bit;
        atomicOr(&bitmap[w], m);
        lo += c;
      }
    }
    wave_sync();
    const uint8_t* src = in + base;
    // round r: lane bytes [16 (64 r + lane), +16); inputs carry 64 bytes of
    // slack, so a partial last piece reads past n harmlessly
    uint32_t p = 16u * lane;
    uint32_t d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    if (p < n) {
      d0 = ld32u(src + p);
      d1 = ld32u(src + p + 4);
      d2 = ld32u(src + p + 8);
      d3 = ld32u(src + p + 12);
    }
    for (uint32_t r0 = 0; r0 < n; r0 += 16u * kWave) {
      const uint32_t q = p + 16u * kWave;
      uint32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;
      if (q < n) {
        e0 = ld32u(src + q);
        e1 = ld32u(src + q + 4);
        e2 = ld32u(src + q + 8);
        e3 = ld32u(src + q + 12);
      }
      uint32_t lit = 0;
      if (p < n) {
        const uint32_t valid = n - p >= 16 ? 0xFFFFu : ((1u << (n - p)) - 1u);
        lit = ~(bitmap[p >> 5] >> (p & 16)) & valid;
      }
      const uint32_t cnt = (uint32_t)__popc(lit);
      const uint32_t incl = wave_incl_scan(cnt);
      uint32_t dst = out + incl - cnt;
      const uint32_t dw[4] = { d0, d1, d2, d3 };
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (lit & (1u << j)) {
          const uint32_t byte = (dw[j >> 2] >> (8 * (j & 3))) & 0xFF;
          lits[dst++] = (uint8_t)byte;
          atomicAdd(&hist[byte], 1u);
        }
      out += wave_bcast(incl, kWave - 1);
      p = q;
      d0 = e0;
      d1 = e1;
      d2 = e2;
      d3 = e3;
    }
    wave_sync();
  }
  __threadfence_block();
  wave_sync();
}

// Sequences section (predefined LL/ML/OF tables) by the whole wave; the same
// bytes as zcore::write_sequences. (A) lanes compute the three codes of every
// sequence into aux; (B) the three FSE state chains run backward over the
// sequences in scalar registers, the tables held one entry per lane and read
// with readlane, and aux[i] receives the state bits sequence i emits (value
// and count per table); (C) every lane assembles one sequence's fields
// (state bits, then the LL/ML/OF extra bits) into a 128-bit word in stream
// order, a wave prefix sum places the words, LDS atomic-ORs assemble them,
// complete bytes are copied out (as in huf_stream_wave); the final states
// and the end mark close the stream.
__device__ uint32_t write_sequences_wave(Lds& L, uint8_t* out, const Seq* seqs, uint32_t* aux,
                                         uint32_t nseq_in, const zcore::Predefined& t)
{
  const int lane = lane_id();
  const uint32_t nseq = uniform(nseq_in);
  uint32_t h;
  if (nseq < 128)
    h = 1;
  else if (nseq < 0x7F00)
    h = 2;
  else
    h = 3;
  if (lane == 0) {
    if (h == 1) {
      out[0] = (uint8_t)nseq;
    } else if (h == 2) {
      out[0] = (uint8_t)((nseq >> 8) + 128);
      out[1] = (uint8_t)nseq;
    } else {
      out[0] = 0xFF;
      out[1] = (uint8_t)(nseq - 0x7F00);
      out[2] = (uint8_t)((nseq - 0x7F00) >> 8);
    }
    if (nseq)
      out[h] = 0; // LL, OF, ML: predefined
  }
  if (nseq == 0)
    return h;
  uint8_t* bs = out + h + 1;

  // (A) codes
  for (uint32_t i = lane; i < nseq; i += kWave) {
    const Seq s = seqs[i];
    const uint32_t llc = (uint32_t)zcore::ll_code(s.ll), mlc = (uint32_t)zcore::ml_code(s.ml - 3);
    const uint32_t ofc = (uint32_t)zcore::highbit32(s.off + 3);
    aux[i] = llc | (mlc << 8) | (ofc << 16);
  }
  // tables, one entry per lane: symbol (delta_nb_bits | (delta_find_state + 128) << 20), state
  auto sym_of = [&](const zcore::FseCTable& ct, int n) -> uint32_t {
    return lane < n ? ct.sym[lane].delta_nb_bits | ((uint32_t)(ct.sym[lane].delta_find_state + 128) << 20) : 0u;
  };
  const uint32_t ll_sym = sym_of(t.ll, 36), ml_sym = sym_of(t.ml, 53), of_sym = sym_of(t.of, 29);
  const uint32_t ll_st = t.ll.state[lane], ml_st = t.ml.state[lane];
  const uint32_t of_st = t.of.state[lane & 31];
  __threadfence_block();
  __syncthreads();

  // (B) state chains
  auto rl = [](uint32_t v, uint32_t i) -> uint32_t {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)i);
  };
  auto init = [&](uint32_t symv, uint32_t stv, uint32_t code) -> uint32_t {
    const uint32_t e = rl(symv, code);
    const uint32_t dnb = e & 0xFFFFFu;
    const int dfs = (int)(e >> 20) - 128;
    const uint32_t nb = (dnb + (1u << 15)) >> 16;
    const uint32_t v = (nb << 16) - dnb;
    return rl(stv, (uint32_t)((int)(v >> nb) + dfs));
  };
  // one encode step with the symbol's transform fields (d, f): advances the
  // state; the state it leaves is recorded by the caller
  auto adv = [&](uint32_t stv, uint32_t d, uint32_t f, uint32_t st) -> uint32_t {
    const uint32_t nb = (st + d) >> 16;
    return rl(stv, (st >> nb) + f);
  };
  // (bits << 3) | count that state st emits for a symbol with field d
  auto emitted = [](uint32_t st, uint32_t d) -> uint32_t {
    const uint32_t nb = (st + d) >> 16;
    return ((st & ((1u << nb) - 1u)) << 3) | nb;
  };
  uint32_t sll, sml, sof;
  {
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)aux[nseq - 1]);
    sml = init(ml_sym, ml_st, (c >> 8) & 0xFF);
    sof = init(of_sym, of_st, c >> 16);
    sll = init(ll_sym, ll_st, c & 0xFF);
  }
  // Per round of 64 sequences every lane first looks up its own sequence's
  // three symbol transforms and splits them into fields (dnb and dfs), so a
  // step reads two lanes and the state table; the three states a sequence
  // leaves go into its lane by lane writes, and the (bits, count) fields
  // follow from them afterwards, one lane per sequence.
  for (long base = (long)nseq - 2; base >= 0; base -= kWave) {
    const long my = base - lane;
    const uint32_t v = my >= 0 ? aux[my] : 0u;
    const uint32_t eo = (uint32_t)__shfl((int)of_sym, (int)(v >> 16), kWave);
    const uint32_t em = (uint32_t)__shfl((int)ml_sym, (int)((v >> 8) & 0xFF), kWave);
    const uint32_t el = (uint32_t)__shfl((int)ll_sym, (int)(v & 0xFF), kWave);
    const uint32_t d_o = eo & 0xFFFFFu, d_m = em & 0xFFFFFu, d_l = el & 0xFFFFFu;
    const uint32_t f_o = (eo >> 20) - 128u, f_m = (em >> 20) - 128u, f_l = (el >> 20) - 128u;
    uint32_t r_o = 0, r_m = 0, r_l = 0;
    const int cnt = base + 1 < kWave ? (int)base + 1 : kWave;
    for (int j = 0; j < cnt; ++j) {
      // (one scalar source per instruction on this target: the lane goes in m0)
      asm volatile("v_writelane_b32 %0, %3, m0\n\tv_writelane_b32 %1, %4, m0\n\tv_writelane_b32 %2, %5, m0"
                   : "+v"(r_o), "+v"(r_m), "+v"(r_l)
                   : "s"(sof), "s"(sml), "s"(sll), "{m0}"(j));
      sof = adv(of_st, rl(d_o, (uint32_t)j), rl(f_o, (uint32_t)j), sof);
      sml = adv(ml_st, rl(d_m, (uint32_t)j), rl(f_m, (uint32_t)j), sml);
      sll = adv(ll_st, rl(d_l, (uint32_t)j), rl(f_l, (uint32_t)j), sll);
    }
    if (my >= 0)
      aux[my] = emitted(r_o, d_o) | (emitted(r_m, d_m) << 9) | (emitted(r_l, d_l) << 18);
  }
  __threadfence_block();
  __syncthreads();

  // (C) bitstream, sequences in encoding order k = 0.. (sequence nseq-1-k)
  uint32_t* stage = L.u.enc.stage;
  for (int i = lane; i < kStageWords; i += kWave)
    stage[i] = 0;
  __syncthreads();
  uint32_t flushed = 0, rel = 0;
  for (uint32_t k0 = 0; k0 < nseq; k0 += kWave) {
    const uint32_t k = k0 + (uint32_t)lane;
    unsigned __int128 bits = 0;
    uint32_t nb = 0;
    if (k < nseq) {
      const uint32_t i = nseq - 1 - k;
      const Seq s = seqs[i];
      if (k > 0) {
        const uint32_t sb = aux[i];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const uint32_t field = (sb >> (9 * f)) & 0x1FF;
          bits |= (unsigned __int128)(field >> 3) << nb;
          nb += field & 7;
        }
      }
      const int llc = zcore::ll_code(s.ll), mlc = zcore::ml_code(s.ml - 3);
      const uint32_t ofv = s.off + 3;
      const uint32_t ofc = (uint32_t)zcore::highbit32(ofv);
      const uint32_t llb = zcore::ll_bits(llc), mlb = zcore::ml_bits(mlc);
      bits |= (unsigned __int128)(s.ll & ((1u << llb) - 1u)) << nb;
      nb += llb;
      bits |= (unsigned __int128)((s.ml - 3) & ((1u << mlb) - 1u)) << nb;
      nb += mlb;
      bits |= (unsigned __int128)(ofv & (uint32_t)((1ull << ofc) - 1)) << nb;
      nb += ofc;
    }
    const uint32_t incl = wave_incl_scan(nb);
    const uint32_t total = wave_bcast(incl, kWave - 1);
    if (nb) {
      const uint32_t p = rel + incl - nb;
      const unsigned __int128 sh = bits << (p & 31);
      uint32_t* w = stage + (p >> 5);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t v = (uint32_t)(sh >> (32 * q));
        if (v)
          atomicOr(&w[q], v);
      }
    }
    __syncthreads();
    rel += total;
    const uint32_t full = rel >> 3;
    wave_copy(bs + flushed, (const uint8_t*)stage, full);
    const uint32_t carry = ((const uint8_t*)stage)[full];
    __syncthreads();
    const uint32_t used_words = (rel + 31) >> 5;
    for (uint32_t i = lane; i < used_words; i += kWave)
      stage[i] = (i == 0) ? (carry & ((1u << (rel & 7)) - 1u)) : 0u;
    __syncthreads();
    flushed += full;
    rel &= 7;
  }
  // final states (ML, OF, LL) and the end mark: <= 7 + 6 + 5 + 6 + 1 bits
  if (lane == 0) {
    uint64_t acc = stage[0] & ((1u << rel) - 1u);
    uint32_t n = rel;
    acc |= (uint64_t)(sml & 63u) << n;
    n += 6;
    acc |= (uint64_t)(sof & 31u) << n;
    n += 5;
    acc |= (uint64_t)(sll & 63u) << n;
    n += 6;
    acc |= 1ull << n;
    n += 1;
    for (uint32_t b = 0; b < (n + 7) / 8; ++b)
      bs[flushed + b] = (uint8_t)(acc >> (8 * b));
    flushed += (n + 7) / 8;
  }
  flushed = wave_bcast(flushed, 0);
  __syncthreads();
  return h + 1 + flushed;
}

// Compresses one block; returns bytes written at out (header included).
// in_lds: `in` is the LDS-staged block (else global memory); the literal
// scratch `lits` is global memory.
__device__ __forceinline__ uint32_t compress_block(Lds& L, const uint8_t* in, uint32_t len, uint8_t* out, bool last,
                                   uint8_t* lits, Seq* seqs, uint8_t* seqout, uint32_t* aux,
                                   const zcore::Predefined& pre, const ChunkTable* tab, bool in_lds)
{
  const int lane = lane_id();

  // ---- parse + literal compaction + histogram ----
  static_assert(sizeof(L.u.table) == 512 * sizeof(uint4), "hash table clear");
  for (int i = lane; i < 512; i += kWave)
    ((uint4*)L.u.table)[i] = make_uint4(0, 0, 0, 0);
  for (int i = lane; i < 256; i += kWave)
    L.hist[i] = 0;
  __syncthreads();
  uint32_t nseq = 0, nlits = 0, anchor = 0;
  if (len >= 16) {
    const uint32_t last_start = len - kMinMatch; // a key needs 8 readable bytes
    uint32_t pos = 0;
    while (pos <= last_start && nseq < kMaxSeq) {
      const uint32_t stride = 1 + ((pos - anchor) >> 8);
      const uint32_t p = pos + (uint32_t)lane * stride;
      const bool active = p <= last_start;
      uint64_t key = 0;
      uint32_t h = 0, dist = 0;
      if (active) {
        key = ld64u(in + p);
        h = key_hash(key);
        // entries hold positions mod 2^16: the candidate is the last position
        // with this hash if it lies within 64 KiB before p. Stale, empty or
        // foreign entries only ever name in-block positions before p, and the
        // 8-byte compare rejects every candidate that is not a match.
        dist = (p - (uint32_t)L.u.table[h]) & 0xFFFFu;
      }
      const bool hit = active && dist != 0 && dist <= p && ld64u(in + p - dist) == key;
      // every hit lane extends its own match up to kLaneExt bytes forward and
      // up to 8 bytes backward, so one window can yield several sequences
      // without new lookups
      uint32_t ml_lane = 0, bx_lane = 0;
      if (hit && p - dist >= 8) {
        const uint64_t x = ld64u(in + p - 8) ^ ld64u(in + p - dist - 8);
        bx_lane = x ? (uint32_t)(__builtin_clzll(x) >> 3) : 8u;
      }
      if (hit) {
        const uint32_t cap = min(len - p, kLaneExt);
        ml_lane = kMinMatch;
        while (ml_lane < cap) {
          const uint64_t x = ld64u(in + p + ml_lane) ^ ld64u(in + p - dist + ml_lane);
          const uint32_t nb = x ? (uint32_t)(__builtin_ctzll(x) >> 3) : 8u;
          ml_lane = min(ml_lane + nb, cap);
          if (nb < 8)
            break;
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (active)
        L.u.table[h] = (uint16_t)p;
      __builtin_amdgcn_wave_barrier();
      uint64_t mask = wave_ballot(hit);
      if (mask == 0) {
        pos += kWave * stride;
        continue;
      }
      // Greedy over the window, wave-parallel: the first hit, then the first
      // hit at or past the end of the previous match, and so on. Every hit
      // lane knows its own match end, hence its successor in that chain; the
      // scalar unit walks the chain (one lane read per match) and the
      // selected lanes then finish their sequences side by side. A lane whose
      // forward match reached kLaneExt has an end only the wave extension
      // knows: the chain stops there, and resumes past its extended end.
      const uint32_t end_l = p + ml_lane;
      uint32_t succ = kWave;
      if (hit && ml_lane < kLaneExt) {
        const uint32_t sp = stride == 1 ? end_l - pos : (end_l - pos + stride - 1) / stride;
        const uint64_t rest = sp >= (uint32_t)kWave ? 0ull : mask & (~0ull << sp);
        if (rest)
          succ = (uint32_t)(__ffsll((unsigned long long)rest) - 1);
      }
      const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
      do {
        uint64_t sel = 0;
        uint32_t cur = (uint32_t)(__ffsll((unsigned long long)mask) - 1), last = cur;
        for (uint32_t room = kMaxSeq - nseq; cur < (uint32_t)kWave && room; --room) {
          sel |= 1ull << cur;
          last = cur;
          cur = (uint32_t)__builtin_amdgcn_readlane((int)succ, (int)cur);
        }
        const uint32_t pl = pos + last * stride;
        uint32_t ml_f = ml_lane;
        if ((uint32_t)__builtin_amdgcn_readlane((int)ml_lane, (int)last) == kLaneExt) {
          const uint32_t dl = (uint32_t)__builtin_amdgcn_readlane((int)dist, (int)last);
          const uint32_t ext = wave_match_extend(in, pl, pl - dl, kLaneExt, len);
          if ((uint32_t)lane == last)
            ml_f = ext;
        }
        // the previous selected match's end (the anchor for the first)
        const uint64_t before = sel & below;
        const int prev = before ? 63 - __clzll(before) : 0;
        const uint32_t prev_lane_end = (uint32_t)__shfl((int)end_l, prev, kWave);
        const uint32_t prev_end = before ? prev_lane_end : anchor;
        uint32_t ll = 0, end_sel = 0;
        if ((sel >> lane) & 1) {
          uint32_t mp = p, mc = p - dist, ml = ml_f;
          const uint32_t bx = min(bx_lane, p - prev_end);
          mp -= bx;
          mc -= bx;
          ml += bx;
          if (bx == 8)
            while (mp > prev_end && mc > 0 && in[mp - 1] == in[mc - 1]) {
              --mp;
              --mc;
              ++ml;
            }
          ll = mp - prev_end;
          seqs[nseq + (uint32_t)__popcll(before)] = Seq{ ll, ml, mp - mc, mp };
          end_sel = mp + ml;
          // the position two bytes before the match end goes into the table
          if (end_sel >= 2 && end_sel - 2 <= last_start)
            L.u.table[key_hash(ld64u(in + end_sel - 2))] = (uint16_t)(end_sel - 2);
        }
        __builtin_amdgcn_wave_barrier();
        nlits += wave_sum(ll);
        nseq += (uint32_t)__popcll(sel);
        anchor = (uint32_t)__builtin_amdgcn_readlane((int)end_sel, (int)last);
        // lanes whose position lies inside a selected match are spent
        const uint32_t spent = stride == 1 ? anchor - pos : (anchor - pos + stride - 1) / stride;
        mask = spent >= (uint32_t)kWave ? 0ull : mask & (~0ull << spent);
      } while (mask && nseq < kMaxSeq);
      pos = anchor;
    }
  }
  // sequences were stored by the lanes that found them; later passes read
  // them from every lane
  wave_sync();
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
  // the literals are read from the block itself when it has no matches
  const bool lits_lds = nseq == 0 && in_lds;
  if (nseq == 0) {
    if (((uintptr_t)in & 15) != 0)
      count_bytes_any(in, len, L.hist);
    else if (in_lds)
      count_bytes<3>((BytePtr<3>)in, len, L.hist);
    else
      count_bytes<1>((BytePtr<1>)in, len, L.hist);
    lits = const_cast<uint8_t*>(in);
  } else {
    gather_literals(in, len, seqs, nseq, (uint32_t*)L.u.table, lits, L.hist);
  }
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
  if (tab) {
    // chunk-table mode: treeless literals with the chunk's code when that is
    // smaller than raw (the carrier block at the start of the frame holds the
    // tree description)
    if (tab->tree_size && nlits >= 256 && used > 1) {
      uint32_t bits_local = 0;
      for (int s = lane; s < 256; s += kWave) {
        const uint32_t c = tab->cw[s];
        L.code.val[s] = (uint16_t)c;
        L.code.len[s] = (uint8_t)(c >> 16);
        bits_local += L.hist[s] * (c >> 16);
      }
      const uint32_t bits = wave_sum(bits_local);
      if (lane == 0) {
        const uint64_t comp_bound = 6 + (bits + 4 * 8 + 7) / 8;
        const uint32_t hs = (uint32_t)zcore::huf_lits_header_size(nlits, nlits);
        if (comp_bound < nlits && hs + comp_bound < L.lit_size) {
          L.mode = 3;
          L.lit_hsize = hs;
          L.lit_size = hs + (uint32_t)comp_bound;
        }
      }
    }
  } else if (nlits >= 256 && used > 1 && huf_build_fast_wave(L)) {
    {
      const uint32_t ts = huf_write_tree_wave(L);
      if (lane == 0)
        L.tree_size = ts;
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
  // ---- sequences section into scratch ----
  {
    const uint32_t ss = write_sequences_wave(L, seqout, seqs, aux, nseq, pre);
    if (lane == 0)
      L.seq_size = ss;
  }
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
    // mode 2: tree description + streams; mode 3: streams only (treeless)
    const uint32_t hs = L.lit_hsize, ts = mode == 2 ? L.tree_size : 0u;
    if (lane == 0)
      for (uint32_t i = 0; i < ts; ++i)
        o[hs + i] = L.tree[i];
    for (int sy = lane; sy < 256; sy += kWave)
      L.u.enc.cw[sy] = (uint32_t)L.code.val[sy] | ((uint32_t)L.code.len[sy] << 16);
    __syncthreads();
    const uint32_t seg = (nlits + 3) / 4;
    uint32_t sp = hs + ts + 6, sizes[4];
    for (int k = 0; k < 4; ++k) {
      const uint32_t a = (uint32_t)k * seg;
      const uint32_t b = k == 3 ? nlits : a + seg;
      sizes[k] = lits_lds ? huf_stream_wave<3>(o + sp, (LitPtr<3>)lits, a, b, L.u.enc.cw, L.u.enc.stage)
                          : huf_stream_wave<1>(o + sp, (LitPtr<1>)lits, a, b, L.u.enc.cw, L.u.enc.stage);
      sp += sizes[k];
    }
    const uint32_t comp = sp - hs;
    if (lane == 0) {
      zcore::write_huf_lits_header(o, hs, nlits, comp, mode);
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

// First block of a chunk-table frame: the chunk's first `k` bytes as Huffman
// literals in one stream, preceded by the tree description, and no sequences.
// Every later treeless block of the frame decodes with this table. Lane 0.
__device__ uint32_t write_carrier(uint8_t* out, const uint8_t* in, uint32_t k, const ChunkTable& t)
{
  uint8_t* lit = out + 3;
  uint8_t* p = lit + 3;
  for (uint32_t i = 0; i < t.tree_size; ++i)
    p[i] = t.tree[i];
  p += t.tree_size;
  zcore::BitW bw;
  zcore::bw_init(bw, p);
  for (long i = (long)k - 1; i >= 0; --i) {
    const uint32_t c = t.cw[in[i]];
    zcore::bw_add(bw, c & 0xFFFF, c >> 16);
  }
  const uint32_t comp = t.tree_size + (uint32_t)zcore::bw_close(bw);
  // compressed literals, size format 0: one stream, 10-bit sizes
  const uint32_t v = 2u | (k << 4) | (comp << 14);
  lit[0] = (uint8_t)v;
  lit[1] = (uint8_t)(v >> 8);
  lit[2] = (uint8_t)(v >> 16);
  lit[3 + comp] = 0; // no sequences
  const uint32_t body = 3 + comp + 1;
  zcore::write_block_header(out, false, 2, body);
  return 3 + body;
}

// One workgroup per chunk: byte histogram of the whole chunk, then the first
// wavefront builds its Huffman code and tree description. The table is used
// when its estimated stream size beats raw literals over the chunk.
__global__ __launch_bounds__(256) void zstd_table_k(const uint8_t* __restrict__ src,
                                                    uint64_t src_stride, uint32_t chunk_bytes,
                                                    uint32_t nseg, uint64_t nchunks,
                                                    ChunkTable* __restrict__ tables)
{
  __shared__ Lds L;
  const uint64_t chunk = blockIdx.x;
  if (chunk >= nchunks)
    return;
  for (uint32_t s = threadIdx.x; s < 256; s += blockDim.x)
    L.hist[s] = 0;
  __syncthreads();
  count_bytes_block(src + chunk * src_stride, chunk_bytes, L.hist);
  __syncthreads();
  if (threadIdx.x >= kWave)
    return;
  const int lane = lane_id();
  ChunkTable& t = tables[chunk];
  uint32_t used_local = 0;
  for (int s = lane; s < 256; s += kWave)
    used_local += L.hist[s] != 0;
  const uint32_t used = wave_sum(used_local);
  uint32_t tree_size = 0;
  if (used > 1 && chunk_bytes >= 4 * kCarrier) {
    huf_sort_wave(L, used);
    huf_build_wave(L, used);
    {
      const uint32_t ts = huf_write_tree_wave(L);
      if (lane == 0)
        L.tree_size = ts;
    }
    wave_sync();
    uint64_t bits_local = 0;
    for (int s = lane; s < 256; s += kWave)
      bits_local += (uint64_t)L.hist[s] * L.code.len[s];
    uint32_t bits_hi = wave_sum((uint32_t)(bits_local >> 16)), bits_lo = wave_sum((uint32_t)(bits_local & 0xFFFF));
    const uint64_t bytes = (((uint64_t)bits_hi << 16) + bits_lo + 7) / 8;
    tree_size = L.tree_size;
    // worth it when the coded literals beat raw by more than the per-block
    // stream overhead
    if (!tree_size || bytes + tree_size + 16ull * nseg >= (uint64_t)chunk_bytes)
      tree_size = 0;
    for (int s = lane; s < 256; s += kWave)
      t.cw[s] = (uint32_t)L.code.val[s] | ((uint32_t)L.code.len[s] << 16);
    for (uint32_t i = lane; i < tree_size; i += kWave)
      t.tree[i] = L.tree[i];
  }
  if (lane == 0)
    t.tree_size = tree_size;
}

// LDS of one worker. Staged workers (blocks of at most kZstdStageBytes)
// also hold their block's input, so the parse and the literal passes read LDS
// instead of global memory: 18.3 KiB a wavefront, 8 per CU, for the short
// launches the smallest block size is chosen for.
template<bool Staged> struct WorkerLds
{
  Lds L;
};
template<> struct WorkerLds<true>
{
  Lds L;
  alignas(16) uint8_t in[kZstdStageBytes + 64];
};

// Work item = one block of `block` bytes. Output segments (one per `sub`
// consecutive blocks) sit seg_pitch apart in a chunk's slot; block h of a
// segment writes at h * sub_off in it (sub > 1: the blocks of a segment by
// as many wavefronts, joined afterwards by zstd_join_k). sizes[chunk * nseg * sub
// + block index] receives each block's bytes (frame header included).
template<bool Staged>
__global__ __launch_bounds__(64, 4) void zstd_k(const uint8_t* __restrict__ src, uint64_t src_stride,
                                             uint32_t chunk_bytes, uint32_t block, uint32_t nseg,
                                             uint32_t sub, uint32_t sub_off, uint64_t seg_pitch,
                                             uint64_t total_blocks, uint8_t* __restrict__ dst,
                                             uint64_t chunk_dst_stride,
                                             uint32_t* __restrict__ seg_sizes, uint8_t* scratch,
                                             uint32_t* queue, const zcore::Predefined* __restrict__ pre,
                                             const ChunkTable* __restrict__ tables)
{
  __shared__ WorkerLds<Staged> W;
  Lds& L = W.L;
  const int lane = lane_id();
  uint8_t* ws = scratch + (uint64_t)blockIdx.x * kWorkerScratch;
  uint8_t* lits = ws;
  Seq* seqs = (Seq*)(ws + kLitsBytes);
  uint8_t* seqout = ws + kLitsBytes + kSeqBytes;
  uint32_t* aux = (uint32_t*)(ws + kLitsBytes + kSeqBytes + kSeqOutBytes);
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
    const uint32_t per_chunk = nseg * sub;
    const uint64_t chunk = item / per_chunk;
    const uint32_t b = (uint32_t)(item % per_chunk); // block index in the chunk
    const uint8_t* in = src + chunk * src_stride + (uint64_t)b * block;
    const uint32_t len = (chunk_bytes - b * block) < block ? chunk_bytes - b * block : block;
    uint8_t* out = dst + chunk * chunk_dst_stride + (uint64_t)(b / sub) * seg_pitch + (b % sub) * sub_off;
    const ChunkTable* tab = tables ? tables + chunk : nullptr;
    uint32_t hdr = 0, skip = 0;
    if (b == 0) {
      if (lane == 0)
        zcore::write_frame_header(out, chunk_bytes);
      hdr = (uint32_t)zcore::kFrameHeaderBytes;
      if (tab && tab->tree_size) {
        uint32_t c = 0;
        if (lane == 0)
          c = write_carrier(out + hdr, in, kCarrier, *tab);
        hdr += wave_bcast(c, 0);
        skip = kCarrier;
      }
    }
    const uint8_t* blk = in;
    if constexpr (Staged) {
      // block starts are 16-byte aligned (chunk strides and block sizes are)
      for (uint32_t o = 16u * (uint32_t)lane; o < len; o += 16u * kWave)
        *(uint4*)&W.in[o] = *(const uint4*)(in + o);
      __syncthreads();
      blk = W.in;
    }
    const uint32_t sz = compress_block(L, blk + skip, len - skip, out + hdr, b + 1 == per_chunk, lits, seqs,
                                       seqout, aux, *pre, tab, Staged);
    if (lane == 0)
      seg_sizes[chunk * per_chunk + b] = hdr + sz;
    __syncthreads();
    if (!queue)
      break;
  }
  // queue[0] hands out blocks, queue[1] counts the workers that found none
  // left; the last of them clears both for the next launch on this queue
  // (every worker has taken its last item by then)
  if (queue && lane == 0 && atomicAdd(queue + 1, 1u) == gridDim.x - 1) {
    atomicExch(queue + 1, 0u);
    atomicExch(queue, 0u);
  }
}

std::mutex g_pre_mu;
std::map<int, zcore::Predefined*> g_pre; // per device, kept until the process ends

const zcore::Predefined* predefined_on_current_device()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return nullptr;
  std::lock_guard<std::mutex> lk(g_pre_mu);
  auto it = g_pre.find(dev);
  return it == g_pre.end() ? nullptr : it->second;
}

} // namespace

hipError_t zstd_prepare()
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return e;
  std::lock_guard<std::mutex> lk(g_pre_mu);
  if (g_pre.count(dev))
    return hipSuccess;
  zcore::Predefined h;
  uint8_t spread[512];
  uint32_t cumul[66];
  zcore::build_predefined(h, spread, cumul);
  zcore::Predefined* d = nullptr;
  hipStream_t st = nullptr;
  if ((e = hipMalloc((void**)&d, sizeof h)) != hipSuccess)
    return e;
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) == hipSuccess) {
    e = hipMemcpyAsync(d, &h, sizeof h, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  if (e != hipSuccess) {
    (void)hipFree(d);
    return e;
  }
  g_pre[dev] = d;
  return hipSuccess;
}

bool zstd_block_ok(uint32_t block)
{
  return block >= kZstdBlockMin && block <= kBlock && (block & (block - 1)) == 0;
}
uint32_t zstd_segments(size_t chunk_bytes, uint32_t block)
{
  return (uint32_t)((chunk_bytes + block - 1) / block);
}
uint64_t zstd_segment_stride(uint32_t block) { return block + kSegSlack; }
size_t zstd_max_output_size(size_t chunk_bytes, uint32_t block)
{
  return (size_t)zstd_segments(chunk_bytes, block) * zstd_segment_stride(block);
}
size_t zstd_scratch_bytes(uint32_t workers) { return (size_t)workers * kWorkerScratch; }

size_t zstd_table_bytes(uint64_t n) { return (size_t)n * sizeof(ChunkTable); }

size_t zstd_predefined_bytes() { return sizeof(zcore::Predefined); }

namespace {
__global__ void predefined_k(zcore::Predefined* out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint8_t spread[512];
    uint32_t cumul[66];
    zcore::build_predefined(*out, spread, cumul);
  }
}
} // namespace

hipError_t zstd_build_predefined(void* dst, hipStream_t stream)
{
  hipLaunchKernelGGL(predefined_k, dim3(1), dim3(1), 0, stream, (zcore::Predefined*)dst);
  return hipGetLastError();
}

hipError_t launch_zstd(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                       uint8_t* dst, uint64_t chunk_dst_stride, uint32_t* seg_sizes, int level,
                       uint32_t block, void* scratch, uint32_t workers, uint32_t* queue_counter,
                       void* tables, hipStream_t stream)
{
  const zcore::Predefined* pre = predefined_on_current_device();
  if (!pre && n)
    return hipErrorNotInitialized; // zstd_prepare() was not called on this device
  return launch_zstd_pre(pre, src, src_stride, chunk_bytes, n, dst, chunk_dst_stride, seg_sizes, level,
                         block, scratch, workers, queue_counter, tables, stream);
}

hipError_t launch_zstd_pre(const void* predefined, const uint8_t* src, uint64_t src_stride,
                           uint64_t chunk_bytes, uint64_t n, uint8_t* dst, uint64_t chunk_dst_stride,
                           uint32_t* seg_sizes, int level, uint32_t block, void* scratch, uint32_t workers,
                           uint32_t* queue_counter, void* tables, hipStream_t stream)
{
  (void)level;
  if (n == 0)
    return hipSuccess;
  if (chunk_bytes == 0 || chunk_bytes > 0xFFFFFFFFull || !zstd_block_ok(block))
    return hipErrorInvalidValue;
  const zcore::Predefined* pre = (const zcore::Predefined*)predefined;
  if (!pre)
    return hipErrorInvalidValue;
  const uint32_t nseg = zstd_segments(chunk_bytes, block);
  const uint64_t total = n * nseg;
  ChunkTable* tabs = (ChunkTable*)tables;
  if (tabs) {
    hipLaunchKernelGGL(zstd_table_k, dim3((uint32_t)n), dim3(256), 0, stream, src, src_stride,
                       (uint32_t)chunk_bytes, nseg, n, tabs);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
  }
  if (!queue_counter && total > workers)
    return hipErrorInvalidValue; // one scratch slice per block in this mode
  const dim3 grid(queue_counter ? workers : (uint32_t)total);
  // staged workers need 16-byte aligned block starts
  const bool staged = block <= kZstdStageBytes && (src_stride % 16) == 0 && ((uintptr_t)src % 16) == 0;
  const uint64_t pitch = block + kSegSlack;
  if (staged)
    hipLaunchKernelGGL(zstd_k<true>, grid, dim3(64), 0, stream, src, src_stride, (uint32_t)chunk_bytes,
                       block, nseg, 1u, 0u, pitch, total, dst, chunk_dst_stride, seg_sizes, (uint8_t*)scratch,
                       queue_counter, pre, tabs);
  else
    hipLaunchKernelGGL(zstd_k<false>, grid, dim3(64), 0, stream, src, src_stride, (uint32_t)chunk_bytes,
                       block, nseg, 1u, 0u, pitch, total, dst, chunk_dst_stride, seg_sizes, (uint8_t*)scratch,
                       queue_counter, pre, tabs);
  return hipGetLastError();
}

namespace {
// One wavefront per segment of a several-blocks-per-segment launch: blocks
// 1 .. parts-1 move down, in order, behind the blocks before them (a block's
// source lies at or above its destination, each round reads its 1 KiB before
// writing it, and a block's destination ends below the next block's source,
// so no byte is overwritten before it is read) and the segment's size becomes
// their sum.
__global__ __launch_bounds__(64) void zstd_join_k(uint8_t* __restrict__ dst, uint64_t chunk_dst_stride,
                                                  uint32_t nseg, uint32_t parts, uint64_t seg_pitch,
                                                  uint32_t sub_off, const uint32_t* __restrict__ sub_sizes,
                                                  uint32_t* __restrict__ seg_sizes, uint64_t n_segments)
{
  const uint64_t g = blockIdx.x;
  if (g >= n_segments)
    return;
  const uint64_t chunk = g / nseg;
  const uint32_t b = (uint32_t)(g % nseg);
  uint8_t* seg = dst + chunk * chunk_dst_stride + (uint64_t)b * seg_pitch;
  const int lane = lane_id();
  uint32_t at = sub_sizes[(uint64_t)parts * g];
  for (uint32_t j = 1; j < parts; ++j) {
    const uint32_t n = sub_sizes[(uint64_t)parts * g + j];
    const uint8_t* from = seg + (uint64_t)j * sub_off;
    uint8_t* to = seg + at;
    for (uint32_t o = 0; o < n; o += 16u * kWave) {
      uint8_t v[16];
      const uint32_t q = o + 16u * (uint32_t)lane;
      const uint32_t k = q < n ? min(16u, n - q) : 0u;
      for (uint32_t i = 0; i < k; ++i)
        v[i] = from[q + i];
      __builtin_amdgcn_wave_barrier();
      for (uint32_t i = 0; i < k; ++i)
        to[q + i] = v[i];
      __builtin_amdgcn_wave_barrier();
    }
    at += n;
  }
  if (lane == 0)
    seg_sizes[g] = at;
}
} // namespace

uint32_t zstd_parts(uint64_t chunk_bytes, uint32_t seg_block)
{
  if (seg_block > 64 * 1024 || chunk_bytes % seg_block)
    return 0;
  for (uint32_t parts : { kZstdMaxParts, 2u })
    if (seg_block / parts >= kZstdPartMin && seg_block / parts <= kZstdStageBytes)
      return parts;
  return 0;
}

hipError_t launch_zstd_parts(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                             uint8_t* dst, uint64_t chunk_dst_stride, uint32_t* seg_sizes,
                             uint32_t* sub_sizes, uint32_t seg_block, void* scratch, uint32_t workers,
                             uint32_t* queue_counter, hipStream_t stream, hipEvent_t between)
{
  if (n == 0)
    return between ? hipEventRecord(between, stream) : hipSuccess;
  const zcore::Predefined* pre = predefined_on_current_device();
  if (!pre)
    return hipErrorNotInitialized;
  const uint32_t parts = zstd_parts(chunk_bytes, seg_block);
  if (!parts || !queue_counter || (src_stride % 16) || ((uintptr_t)src % 16))
    return hipErrorInvalidValue;
  const uint32_t block = seg_block / parts;
  const uint32_t nseg = zstd_segments(chunk_bytes, seg_block);
  const uint64_t pitch = zstd_segment_stride(seg_block);
  // a block writes at most its frame header, block header and raw bytes:
  // block j of a segment starts at j * sub_off, which leaves that much room
  const uint32_t sub_off = (uint32_t)(pitch / parts) & ~15u;
  const uint64_t total = n * nseg * parts;
  hipLaunchKernelGGL(zstd_k<true>, dim3(workers), dim3(64), 0, stream, src, src_stride, (uint32_t)chunk_bytes,
                     block, nseg, parts, sub_off, pitch, total, dst, chunk_dst_stride, sub_sizes,
                     (uint8_t*)scratch, queue_counter, pre, (const ChunkTable*)nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  if (between && (e = hipEventRecord(between, stream)) != hipSuccess)
    return e;
  const uint64_t segs = n * nseg;
  hipLaunchKernelGGL(zstd_join_k, dim3((uint32_t)segs), dim3(64), 0, stream, dst, chunk_dst_stride, nseg, parts,
                     pitch, sub_off, sub_sizes, seg_sizes, segs);
  return hipGetLastError();
}

} // namespace chucky
