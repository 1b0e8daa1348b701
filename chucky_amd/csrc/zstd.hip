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
constexpr uint32_t kMaxSeq = 8192;
constexpr uint64_t kLitsBytes = kBlock + 256;
constexpr uint64_t kSeqBytes = (uint64_t)kMaxSeq * sizeof(Seq);
constexpr uint64_t kSeqOutBytes = (uint64_t)kMaxSeq * 12 + 256;
constexpr uint64_t kWorkerScratch = (kLitsBytes + kSeqBytes + kSeqOutBytes + 255) & ~255ull;
constexpr int kSymPerLane = 8; // Huffman symbols per lane per round
// one round: <= 64 lanes * 8 symbols * 11 bits, plus < 8 carried bits
constexpr int kStageWords = (kWave * kSymPerLane * zcore::kHufMaxBits + 8 + 31) / 32 + 4;

// ~10.2 KiB per wavefront, so that 15 wavefronts fit a CU's LDS. The union
// is the parse's hash table, then the Huffman builder's scratch, then the
// literal encoder's packed codes and bit staging.
struct Lds
{
  union
  {
    uint16_t table[1 << kHashLog]; // low 16 bits of the last position per hash
    zcore::HufWork huf;
    struct
    {
      uint32_t cw[256]; // code value | length << 16
      uint32_t stage[kStageWords];
    } enc;
  } u;
  uint32_t hist[256];
  zcore::HufCode code;
  uint8_t tree[192];
  // decisions published by lane 0
  uint32_t mode, tree_size, lit_size, lit_hsize, seq_size;
};

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
      wave_sync();
    }
}

// Adds every byte of src[0, n) to the LDS histogram. Each lane keeps one
// (symbol, run) pair per byte position within a dword and only issues an LDS
// atomic when the symbol changes: the high bytes of small 16/32-bit integers
// (most often 0) would otherwise make every atomic of the wave hit one bin.
__device__ __forceinline__ void count_bytes(const uint8_t* src, uint32_t n, uint32_t* hist)
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

// Compacts in[0, n) to dst and adds every byte to the LDS histogram.
__device__ __forceinline__ void copy_count(uint8_t* dst, const uint8_t* src, uint32_t n,
                                           uint32_t* hist)
{
  wave_copy(dst, src, n);
  count_bytes(src, n, hist);
}

// The up-to-8 symbols lane `lane` encodes in the round whose highest symbol
// is `top`: symbols top-8*lane-7 .. top-8*lane, as bytes of a 64-bit word
// (byte j = symbol lo+j; bytes below `a` are zero and skipped by the caller).
__device__ __forceinline__ uint64_t round_syms(const uint8_t* lits, long a, long top, int lane,
                                               int* skip)
{
  const long lo = top - (long)lane * kSymPerLane - (kSymPerLane - 1);
  if (lo + kSymPerLane - 1 < a) {
    *skip = kSymPerLane;
    return 0;
  }
  const int sk = lo >= a ? 0 : (int)(a - lo);
  *skip = sk;
  return ld64u(lits + lo + sk) << (8 * sk);
}

// Encodes lits[a, b) as one backward Huffman stream at out; returns its size.
// Each round the wave takes 512 symbols, highest first: every lane packs its
// 8 codes into a 128-bit word, a wave prefix sum of the bit counts places the
// words, LDS atomic-ORs assemble them in `stage`, and complete bytes are
// copied out. The next round's symbols are loaded before this round's codes
// are looked up, so the global load latency overlaps the LDS work.
__device__ uint32_t huf_stream_wave(uint8_t* out, const uint8_t* lits, uint32_t a, uint32_t b,
                                    const uint32_t* cw, uint32_t* stage)
{
  const int lane = lane_id();
  for (int i = lane; i < kStageWords; i += kWave)
    stage[i] = 0;
  __syncthreads();
  uint32_t flushed = 0; // bytes already copied out
  uint32_t rel = 0;     // bits held in stage (stage bit 0 = stream bit 8*flushed)
  long top = (long)b - 1;
  int skip = 0;
  uint64_t q = round_syms(lits, a, top, lane, &skip);
  for (;;) {
    const long next = top - (long)kWave * kSymPerLane;
    int nskip = kSymPerLane;
    const uint64_t nq = next >= (long)a ? round_syms(lits, a, next, lane, &nskip) : 0;
    unsigned __int128 bits = 0;
    uint32_t nb = 0;
#pragma unroll
    for (int j = kSymPerLane - 1; j >= 0; --j) {
      if (j >= skip) {
        const uint32_t c = cw[(uint32_t)(q >> (8 * j)) & 0xFF];
        bits |= (unsigned __int128)(c & 0xFFFF) << nb;
        nb += c >> 16;
      }
    }
    const uint32_t incl = wave_incl_scan(nb);
    const uint32_t total = wave_bcast(incl, kWave - 1);
    if (nb) {
      const uint32_t p = rel + incl - nb;
      const unsigned __int128 sh = bits << (p & 31);
      uint32_t* w = stage + (p >> 5);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t v = (uint32_t)(sh >> (32 * k));
        if (v)
          atomicOr(&w[k], v);
      }
    }
    __syncthreads();
    rel += total;
    if (next < (long)a) {
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
    const uint32_t carry = ((const uint8_t*)stage)[full];
    __syncthreads();
    const uint32_t used_words = (rel + 31) >> 5;
    for (uint32_t i = lane; i < used_words; i += kWave)
      stage[i] = (i == 0) ? (carry & ((1u << (rel & 7)) - 1u)) : 0u;
    __syncthreads();
    flushed += full;
    rel &= 7;
    top = next;
    q = nq;
    skip = nskip;
  }
}

// Huffman code lengths and canonical values for the n >= 2 used symbols in
// L.u.huf.sym[0..n) (ascending count, ties by symbol), built by the whole
// wave: two-queue Huffman merge and the depth pass over register-resident
// arrays, length limiting to kHufMaxBits on per-length counts (Kraft sum
// restored exactly), lengths handed out in count order, canonical values by
// wave ballots. Same code as zcore::huf_build_sorted produces on the host.
__device__ void huf_build_wave(Lds& L, uint32_t n)
{
  const uint32_t lane = (uint32_t)lane_id();
  constexpr uint32_t Lmax = zcore::kHufMaxBits;
  // Two-queue Huffman merge. Leaf weights (ascending) and internal-node
  // weights (created in ascending order) live in registers, 64 per register,
  // lane j of register k holding entry 64k+j; the merge keeps both queue
  // heads in scalars and refreshes only the popped head with one lane read.
  // Parent links go to LDS (fire-and-forget stores by lane 0).
  uint32_t lw[4], nw[4] = { 0, 0, 0, 0 };
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t idx = (uint32_t)k * kWave + lane;
    lw[k] = idx < n ? L.hist[L.u.huf.sym[idx]] : 0u;
  }
  auto pick4 = [](const uint32_t (&r)[4], uint32_t i) -> uint32_t {
    uint32_t v;
    switch (i >> 6) {
      case 0: v = r[0]; break;
      case 1: v = r[1]; break;
      case 2: v = r[2]; break;
      default: v = r[3]; break;
    }
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)(i & 63));
  };
  int16_t* parent = L.u.huf.parent;
  uint32_t leaf = 0, node = 0, made = 0;
  uint32_t hl = pick4(lw, 0), hn = 0xFFFFFFFFu;
  for (uint32_t it = 0; it + 1 < n; ++it) {
    uint32_t pk0 = 0, pk1 = 0, ws = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      uint32_t pk;
      if (leaf < n && hl <= hn) { // ties go to the leaf
        pk = leaf;
        ws += hl;
        ++leaf;
        hl = leaf < n ? pick4(lw, leaf) : 0xFFFFFFFFu;
      } else {
        pk = n + node;
        ws += hn;
        ++node;
        hn = node < made ? pick4(nw, node) : 0xFFFFFFFFu;
      }
      if (c == 0)
        pk0 = pk;
      else
        pk1 = pk;
    }
    const bool mine = lane == (made & 63);
    switch (made >> 6) {
      case 0: nw[0] = mine ? ws : nw[0]; break;
      case 1: nw[1] = mine ? ws : nw[1]; break;
      case 2: nw[2] = mine ? ws : nw[2]; break;
      default: nw[3] = mine ? ws : nw[3]; break;
    }
    if (lane == 0) {
      parent[pk0] = (int16_t)(n + made);
      parent[pk1] = (int16_t)(n + made);
    }
    if (node == made)
      hn = ws; // the node queue was empty: the new node is its head
    ++made;
  }
  const uint32_t nodes = n + made; // root = nodes - 1
  if (lane == 0)
    parent[nodes - 1] = (int16_t)(nodes - 1);
  wave_sync();
  // depths by pointer jumping: every node adds its parent's distance to the
  // root and jumps to its grandparent; 9 rounds cover depths < 512
  uint8_t* dep = L.u.huf.depth;
  uint32_t pj[8], dj[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t i = (uint32_t)k * kWave + lane;
    pj[k] = i < nodes ? (uint32_t)parent[i] : i;
    dj[k] = i < nodes && i != nodes - 1 ? 1u : 0u;
  }
  for (int round = 0; round < 9; ++round) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t i = (uint32_t)k * kWave + lane;
      if (i < nodes) {
        dep[i] = (uint8_t)dj[k];
        parent[i] = (int16_t)pj[k];
      }
    }
    wave_sync();
    uint32_t dp[8], pp[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t i = (uint32_t)k * kWave + lane;
      dp[k] = i < nodes ? dep[pj[k]] : 0u;
      pp[k] = i < nodes ? (uint32_t)parent[pj[k]] : pj[k];
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      dj[k] += dp[k];
      pj[k] = pp[k];
    }
  }
  struct
  {
    uint32_t r[4];
  } D;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    D.r[k] = dj[k];

  // leaf depths of this lane (leaves k*64 + lane, k < 4)
  uint32_t len[4];
  uint32_t maxd = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t idx = (uint32_t)k * kWave + lane;
    len[k] = idx < n ? D.r[k] : 0u;
    maxd = len[k] > maxd ? len[k] : maxd;
  }
  maxd = wave_max(maxd);
  if (maxd > Lmax) {
    // per-length code counts live in lanes: lane b holds the count of length b
    uint32_t blv = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t d = len[k] > Lmax ? Lmax : len[k];
#pragma unroll
      for (uint32_t b = 1; b <= Lmax; ++b) {
        const uint32_t c = (uint32_t)__popcll(wave_ballot(len[k] != 0 && d == b));
        if (lane == b)
          blv += c;
      }
    }
    const int64_t full = (int64_t)1 << Lmax;
    int64_t kraft = (int64_t)wave_sum(lane >= 1 && lane <= Lmax ? blv << (Lmax - lane) : 0u);
    while (kraft > full) {
      // lengthen codes at the longest length below Lmax (smallest decrement)
      const uint64_t nz = wave_ballot(lane >= 1 && lane < Lmax && blv != 0);
      const uint32_t b = 63u - (uint32_t)__clzll(nz);
      const uint32_t cnt = (uint32_t)__builtin_amdgcn_readlane((int)blv, (int)b);
      const int64_t step = (int64_t)1 << (Lmax - b - 1);
      int64_t k = (kraft - full + step - 1) / step;
      if (k > (int64_t)cnt)
        k = cnt;
      if (lane == b)
        blv -= (uint32_t)k;
      if (lane == b + 1)
        blv += (uint32_t)k;
      kraft -= k * step;
    }
    while (kraft < full) {
      // shorten codes at the longest length whose gain still fits
      const int64_t gap = full - kraft;
      const uint64_t ok = wave_ballot(lane > 1 && lane <= Lmax && blv != 0 &&
                                      ((int64_t)1 << (Lmax - lane)) <= gap);
      if (!ok)
        break;
      const uint32_t b = 63u - (uint32_t)__clzll(ok);
      const uint32_t cnt = (uint32_t)__builtin_amdgcn_readlane((int)blv, (int)b);
      const int64_t step = (int64_t)1 << (Lmax - b);
      int64_t k = gap / step;
      if (k > (int64_t)cnt)
        k = cnt;
      if (lane == b)
        blv -= (uint32_t)k;
      if (lane == b - 1)
        blv += (uint32_t)k;
      kraft += k * step;
    }
    // longest codes to the least frequent leaves: leaf ranks [above, above +
    // count) of length b, with `above` the number of codes longer than b
    uint32_t above = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      len[k] = 0;
    for (uint32_t b = Lmax; b >= 1; --b) {
      const uint32_t cnt = (uint32_t)__builtin_amdgcn_readlane((int)blv, (int)b);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t idx = (uint32_t)k * kWave + lane;
        if (idx < n && idx >= above && idx < above + cnt)
          len[k] = b;
      }
      above += cnt;
    }
    const uint64_t nz = wave_ballot(lane >= 1 && lane <= Lmax && blv != 0);
    maxd = 63u - (uint32_t)__clzll(nz);
  }
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

// Compresses one block; returns bytes written at out (header included).
__device__ uint32_t compress_block(Lds& L, const uint8_t* in, uint32_t len, uint8_t* out, bool last,
                                   uint8_t* lits, Seq* seqs, uint8_t* seqout,
                                   const zcore::Predefined& pre, const ChunkTable* tab)
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
      __builtin_amdgcn_wave_barrier();
      if (active)
        L.u.table[h] = (uint16_t)p;
      __builtin_amdgcn_wave_barrier();
      const uint64_t mask = wave_ballot(hit);
      if (mask == 0) {
        pos += kWave * stride;
        continue;
      }
      const int first = __ffsll((unsigned long long)mask) - 1;
      uint32_t mp = pos + (uint32_t)first * stride;
      uint32_t mc = mp - wave_bcast(dist, first);
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
        L.u.table[key_hash(ld64u(in + q))] = (uint16_t)q;
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
  if (nseq == 0) {
    count_bytes(in, len, L.hist);
    lits = const_cast<uint8_t*>(in);
  } else {
    copy_count(lits + nlits, in + anchor, len - anchor, L.hist);
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
  } else if (nlits >= 256 && used > 1) {
    // used symbols in ascending count order: bitonic sort of (count, symbol)
    uint32_t* keys = L.u.huf.w;
    for (int s = lane; s < 256; s += kWave)
      keys[s] = (L.hist[s] << 8) | (uint32_t)s;
    __syncthreads();
    wave_sort256(keys);
    for (uint32_t i = lane; i < used; i += kWave)
      L.u.huf.sym[i] = (uint16_t)(keys[256 - used + i] & 0xFF);
    __syncthreads();
    huf_build_wave(L, used);
    if (lane == 0)
      L.tree_size = (uint32_t)zcore::huf_write_tree(L.tree, L.code, L.u.huf);
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
    L.seq_size = (uint32_t)zcore::write_sequences(seqout, seqs, nseq, pre);
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
      sizes[k] = huf_stream_wave(o + sp, lits, a, b, L.u.enc.cw, L.u.enc.stage);
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
    uint32_t* keys = L.u.huf.w;
    for (int s = lane; s < 256; s += kWave)
      keys[s] = (L.hist[s] << 8) | (uint32_t)s;
    wave_sync();
    wave_sort256(keys);
    for (uint32_t i = lane; i < used; i += kWave)
      L.u.huf.sym[i] = (uint16_t)(keys[256 - used + i] & 0xFF);
    wave_sync();
    huf_build_wave(L, used);
    if (lane == 0)
      L.tree_size = (uint32_t)zcore::huf_write_tree(L.tree, L.code, L.u.huf);
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

__global__ __launch_bounds__(64, 4) void zstd_k(const uint8_t* __restrict__ src, uint64_t src_stride,
                                             uint32_t chunk_bytes, uint32_t block, uint32_t nseg,
                                             uint64_t total_blocks, uint8_t* __restrict__ dst,
                                             uint64_t chunk_dst_stride,
                                             uint32_t* __restrict__ seg_sizes, uint8_t* scratch,
                                             uint32_t* queue, const zcore::Predefined* __restrict__ pre,
                                             const ChunkTable* __restrict__ tables)
{
  __shared__ Lds L;
  const int lane = lane_id();
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
    const uint8_t* in = src + chunk * src_stride + (uint64_t)b * block;
    const uint32_t len = (chunk_bytes - b * block) < block ? chunk_bytes - b * block : block;
    uint8_t* out = dst + chunk * chunk_dst_stride + (uint64_t)b * (block + kSegSlack);
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
    const uint32_t sz = compress_block(L, in + skip, len - skip, out + hdr, b + 1 == nseg, lits, seqs,
                                       seqout, *pre, tab);
    if (lane == 0)
      seg_sizes[chunk * nseg + b] = hdr + sz;
    __syncthreads();
    if (!queue)
      break;
  }
}

__global__ void zero_u32_k(uint32_t* p) { *p = 0; }

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

hipError_t launch_zstd(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                       uint8_t* dst, uint64_t chunk_dst_stride, uint32_t* seg_sizes, int level,
                       uint32_t block, void* scratch, uint32_t workers, uint32_t* queue_counter,
                       void* tables, hipStream_t stream)
{
  (void)level;
  if (n == 0)
    return hipSuccess;
  if (chunk_bytes == 0 || chunk_bytes > 0xFFFFFFFFull || !zstd_block_ok(block))
    return hipErrorInvalidValue;
  const zcore::Predefined* pre = predefined_on_current_device();
  if (!pre)
    return hipErrorNotInitialized; // zstd_prepare() was not called on this device
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
  if (queue_counter) {
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(1), 0, stream, queue_counter);
    hipLaunchKernelGGL(zstd_k, dim3(workers), dim3(64), 0, stream, src, src_stride,
                       (uint32_t)chunk_bytes, block, nseg, total, dst, chunk_dst_stride, seg_sizes,
                       (uint8_t*)scratch, queue_counter, pre, tabs);
  } else {
    if (total > workers)
      return hipErrorInvalidValue; // one scratch slice per block in this mode
    hipLaunchKernelGGL(zstd_k, dim3((uint32_t)total), dim3(64), 0, stream, src, src_stride,
                       (uint32_t)chunk_bytes, block, nseg, total, dst, chunk_dst_stride, seg_sizes,
                       (uint8_t*)scratch, nullptr, pre, tabs);
  }
  return hipGetLastError();
}

} // namespace chucky
