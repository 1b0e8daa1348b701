// zstd (RFC 8878) block encoding primitives, written for both the host and
// gfx950 device code.
//
// The GPU compressor uses these pieces for the serial parts of a block
// (Huffman code construction, FSE table builds, sequence-section encoding);
// a host build of the same header lets the unit tests check the bitstream
// format against stock libzstd without a GPU.
//
// Methods (all public): FSE tables and the tANS state machine follow Yann
// Collet's Finite State Entropy as specified by RFC 8878 section 4.1;
// Huffman codes are length-limited with a Kraft-sum repair and assigned
// canonically per RFC 8878 section 4.2.1; sequence codes and the predefined
// distributions are the tables of RFC 8878 section 3.1.1.3.2.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define ZHD __host__ __device__
#else
#define ZHD
#endif

namespace zcore {

constexpr uint32_t kMagic = 0xFD2FB528u;
constexpr uint32_t kBlockMax = 128 * 1024;
constexpr int kHufMaxBits = 11;
constexpr int kMinMatch = 4; // the match finders hash 4 bytes

ZHD inline int highbit32(uint32_t v) { return 31 - __builtin_clz(v); }

// ---------------------------------------------------------------------------
// Forward bit writer (LSB first). Streams that RFC 8878 reads backwards are
// produced by writing the last-decoded element first and closing with a
// single 1 bit.
struct BitW
{
  uint8_t* p;
  uint8_t* start;
  uint64_t acc;
  uint32_t n;
};

ZHD inline void bw_init(BitW& b, uint8_t* dst)
{
  b.p = b.start = dst;
  b.acc = 0;
  b.n = 0;
}
ZHD inline void bw_flush(BitW& b)
{
  while (b.n >= 8) {
    *b.p++ = (uint8_t)b.acc;
    b.acc >>= 8;
    b.n -= 8;
  }
}
// nb <= 32
ZHD inline void bw_add(BitW& b, uint64_t v, uint32_t nb)
{
  if (b.n + nb > 64)
    bw_flush(b);
  b.acc |= (v & ((nb == 64) ? ~0ull : ((1ull << nb) - 1))) << b.n;
  b.n += nb;
}
ZHD inline size_t bw_close(BitW& b)
{
  bw_add(b, 1, 1);
  bw_flush(b);
  if (b.n)
    *b.p++ = (uint8_t)b.acc;
  b.n = 0;
  b.acc = 0;
  return (size_t)(b.p - b.start);
}

// ---------------------------------------------------------------------------
// FSE compression tables (tableLog <= 9, at most 64 symbols).
struct FseSym
{
  int32_t delta_find_state;
  uint32_t delta_nb_bits;
};
struct FseCTable
{
  uint32_t table_log;
  uint16_t state[512];
  FseSym sym[64];
};

// norm[s] in {-1, 0, 1, ...}, sum of |norm| = 1 << table_log.
// spread: 512 bytes, cumul: 66 words of caller workspace (LDS on the device).
ZHD inline void fse_build_ctable(FseCTable& ct, const int16_t* norm, int max_sym, int table_log,
                                 uint8_t* spread, uint32_t* cumul)
{
  const uint32_t size = 1u << table_log, mask = size - 1;
  const uint32_t step = (size >> 1) + (size >> 3) + 3;
  uint32_t high = size - 1;
  ct.table_log = (uint32_t)table_log;
  cumul[0] = 0;
  for (int u = 1; u <= max_sym + 1; ++u) {
    if (norm[u - 1] == -1) {
      cumul[u] = cumul[u - 1] + 1;
      spread[high--] = (uint8_t)(u - 1);
    } else {
      cumul[u] = cumul[u - 1] + (uint32_t)norm[u - 1];
    }
  }
  uint32_t pos = 0;
  for (int s = 0; s <= max_sym; ++s)
    for (int k = 0; k < norm[s]; ++k) {
      spread[pos] = (uint8_t)s;
      do
        pos = (pos + step) & mask;
      while (pos > high);
    }
  for (uint32_t u = 0; u < size; ++u) {
    const uint8_t s = spread[u];
    ct.state[cumul[s]++] = (uint16_t)(size + u);
  }
  int total = 0;
  for (int s = 0; s <= max_sym; ++s) {
    const int c = norm[s];
    if (c == 0) {
      ct.sym[s].delta_nb_bits = (uint32_t)(((table_log + 1) << 16) - (1 << table_log));
      ct.sym[s].delta_find_state = 0;
    } else if (c == -1 || c == 1) {
      ct.sym[s].delta_nb_bits = (uint32_t)((table_log << 16) - (1 << table_log));
      ct.sym[s].delta_find_state = total - 1;
      total += 1;
    } else {
      const uint32_t max_bits_out = (uint32_t)table_log - (uint32_t)highbit32((uint32_t)c - 1);
      const uint32_t min_state_plus = (uint32_t)c << max_bits_out;
      ct.sym[s].delta_nb_bits = (max_bits_out << 16) - min_state_plus;
      ct.sym[s].delta_find_state = total - c;
      total += c;
    }
  }
}

struct FseState
{
  uint32_t value;
};

// State whose decoder cell emits `sym`, with no bits written.
ZHD inline void fse_init2(FseState& st, const FseCTable& ct, int sym)
{
  const FseSym& tt = ct.sym[sym];
  const uint32_t nb = (tt.delta_nb_bits + (1u << 15)) >> 16;
  const uint32_t v = (nb << 16) - tt.delta_nb_bits;
  st.value = ct.state[(v >> nb) + tt.delta_find_state];
}
ZHD inline void fse_encode(BitW& b, FseState& st, const FseCTable& ct, int sym)
{
  const FseSym& tt = ct.sym[sym];
  const uint32_t nb = (st.value + tt.delta_nb_bits) >> 16;
  bw_add(b, st.value, nb);
  st.value = ct.state[(st.value >> nb) + tt.delta_find_state];
}
ZHD inline void fse_flush_state(BitW& b, const FseState& st, const FseCTable& ct)
{
  bw_add(b, st.value, ct.table_log);
}

// ---------------------------------------------------------------------------
// Normalized-count header (RFC 8878 4.1.1). Returns bytes written.
ZHD inline size_t fse_write_ncount(uint8_t* out, const int16_t* norm, int max_sym, int table_log)
{
  uint8_t* const start = out;
  const int size = 1 << table_log;
  uint32_t bits = 0;
  int nbits = 0;
  bits |= (uint32_t)(table_log - 5) << nbits;
  nbits += 4;
  int remaining = size + 1;
  int threshold = size;
  int nb = table_log + 1;
  int sym = 0;
  bool prev0 = false;
  const int alphabet = max_sym + 1;
  while (sym < alphabet && remaining > 1) {
    if (prev0) {
      int start_sym = sym;
      while (sym < alphabet && norm[sym] == 0)
        ++sym;
      while (sym >= start_sym + 24) {
        start_sym += 24;
        bits |= 0xFFFFu << nbits;
        *out++ = (uint8_t)bits;
        *out++ = (uint8_t)(bits >> 8);
        bits >>= 16;
      }
      while (sym >= start_sym + 3) {
        start_sym += 3;
        bits |= 3u << nbits;
        nbits += 2;
      }
      bits |= (uint32_t)(sym - start_sym) << nbits;
      nbits += 2;
      if (nbits > 16) {
        *out++ = (uint8_t)bits;
        *out++ = (uint8_t)(bits >> 8);
        bits >>= 16;
        nbits -= 16;
      }
    }
    int count = norm[sym++];
    const int max = (2 * threshold - 1) - remaining;
    remaining -= count < 0 ? -count : count;
    count++;
    if (count >= threshold)
      count += max;
    bits |= (uint32_t)count << nbits;
    nbits += nb;
    nbits -= (count < max);
    prev0 = (count == 1);
    while (remaining < threshold) {
      nb--;
      threshold >>= 1;
    }
    if (nbits > 16) {
      *out++ = (uint8_t)bits;
      *out++ = (uint8_t)(bits >> 8);
      bits >>= 16;
      nbits -= 16;
    }
  }
  *out++ = (uint8_t)bits;
  *out++ = (uint8_t)(bits >> 8);
  out -= 2;
  out += (nbits + 7) / 8;
  return (size_t)(out - start);
}

// Two-state FSE stream of `n` symbols (even indices on state 1); 0 if n <= 2.
ZHD inline size_t fse_encode_two_state(uint8_t* dst, const uint8_t* src, size_t n,
                                       const FseCTable& ct)
{
  if (n <= 2)
    return 0;
  BitW b;
  bw_init(b, dst);
  // state of the even and of the odd positions (named, not an indexed pair:
  // on the device an indexed local array would live in scratch memory)
  FseState se, so;
  long i = (long)n - 1;
  if (i & 1) {
    fse_init2(so, ct, src[i]);
    fse_init2(se, ct, src[i - 1]);
  } else {
    fse_init2(se, ct, src[i]);
    fse_init2(so, ct, src[i - 1]);
  }
  for (i -= 2; i >= 0; --i) {
    if (i & 1)
      fse_encode(b, so, ct, src[i]);
    else
      fse_encode(b, se, ct, src[i]);
  }
  fse_flush_state(b, so, ct);
  fse_flush_state(b, se, ct);
  return bw_close(b);
}

// Scale `count` (sum total, max_sym+1 symbols) to 1 << table_log, every
// present symbol >= 1.
ZHD inline void fse_normalize(int16_t* norm, const uint32_t* count, int max_sym, uint32_t total,
                              int table_log)
{
  const int size = 1 << table_log;
  int sum = 0, largest = -1;
  for (int s = 0; s <= max_sym; ++s) {
    if (!count[s]) {
      norm[s] = 0;
      continue;
    }
    int v = (int)(((uint64_t)count[s] * size + total / 2) / total);
    if (v < 1)
      v = 1;
    norm[s] = (int16_t)v;
    sum += v;
    if (largest < 0 || count[s] > count[largest])
      largest = s;
  }
  int diff = size - sum;
  if (norm[largest] + diff >= 1) {
    norm[largest] = (int16_t)(norm[largest] + diff);
    return;
  }
  // take the excess from the biggest entries one by one
  while (diff < 0) {
    int big = -1;
    for (int s = 0; s <= max_sym; ++s)
      if (norm[s] > 1 && (big < 0 || norm[s] > norm[big]))
        big = s;
    norm[big]--;
    diff++;
  }
}

// ---------------------------------------------------------------------------
// Sequence codes (RFC 8878 3.1.1.3.2.1).
ZHD inline int ll_code(uint32_t ll)
{
  if (ll < 16)
    return (int)ll;
  if (ll < 64) {
    if (ll < 24)
      return 16 + (int)((ll - 16) >> 1);
    if (ll < 32)
      return 20 + (int)((ll - 24) >> 2);
    if (ll < 48)
      return 22 + (int)((ll - 32) >> 3);
    return 24;
  }
  return highbit32(ll) + 19;
}
ZHD inline uint32_t ll_bits(int code)
{
  if (code < 16)
    return 0;
  if (code < 20)
    return 1;
  if (code < 22)
    return 2;
  if (code < 24)
    return 3;
  if (code == 24)
    return 4;
  return (uint32_t)(code - 19);
}
// argument: match length - 3
ZHD inline int ml_code(uint32_t mlb)
{
  if (mlb < 32)
    return (int)mlb;
  if (mlb < 128) {
    if (mlb < 40)
      return 32 + (int)((mlb - 32) >> 1);
    if (mlb < 48)
      return 36 + (int)((mlb - 40) >> 2);
    if (mlb < 64)
      return 38 + (int)((mlb - 48) >> 3);
    if (mlb < 96)
      return 40 + (int)((mlb - 64) >> 4);
    return 42;
  }
  return highbit32(mlb) + 36;
}
ZHD inline uint32_t ml_bits(int code)
{
  if (code < 32)
    return 0;
  if (code < 36)
    return 1;
  if (code < 38)
    return 2;
  if (code < 40)
    return 3;
  if (code < 42)
    return 4;
  if (code == 42)
    return 5;
  return (uint32_t)(code - 36);
}

// Predefined distributions (RFC 8878 3.1.1.3.2.2).
struct Predefined
{
  FseCTable ll, ml, of;
};
ZHD inline void build_predefined(Predefined& p, uint8_t* spread, uint32_t* cumul)
{
  const int16_t ll[36] = { 4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2,
                           2, 2, 2, 2, 2, 2, 2, 3, 2, 1, 1, 1, 1, 1, -1, -1, -1, -1 };
  const int16_t ml[53] = { 1, 4, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                           1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                           1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1 };
  const int16_t of[29] = { 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1,
                           1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1 };
  fse_build_ctable(p.ll, ll, 35, 6, spread, cumul);
  fse_build_ctable(p.ml, ml, 52, 6, spread, cumul);
  fse_build_ctable(p.of, of, 28, 5, spread, cumul);
}

struct Seq
{
  uint32_t ll;  // literals before the match
  uint32_t ml;  // match length (>= 3)
  uint32_t off; // match distance (>= 1)
  uint32_t pos; // match start in the block (device parser; unused by the encoder)
};

// Sequences section with predefined tables. Returns bytes written.
ZHD inline size_t write_sequences(uint8_t* out, const Seq* seq, uint32_t nseq, const Predefined& t)
{
  uint8_t* const start = out;
  if (nseq < 128) {
    *out++ = (uint8_t)nseq;
  } else if (nseq < 0x7F00) {
    *out++ = (uint8_t)((nseq >> 8) + 128);
    *out++ = (uint8_t)nseq;
  } else {
    *out++ = 0xFF;
    *out++ = (uint8_t)(nseq - 0x7F00);
    *out++ = (uint8_t)((nseq - 0x7F00) >> 8);
  }
  if (nseq == 0)
    return (size_t)(out - start);
  *out++ = 0; // LL, OF, ML: predefined
  BitW b;
  bw_init(b, out);
  FseState sll, sml, sof;
  {
    const Seq& s = seq[nseq - 1];
    const int llc = ll_code(s.ll), mlc = ml_code(s.ml - 3);
    const uint32_t ofv = s.off + 3;
    const int ofc = highbit32(ofv);
    fse_init2(sml, t.ml, mlc);
    fse_init2(sof, t.of, ofc);
    fse_init2(sll, t.ll, llc);
    bw_add(b, s.ll, ll_bits(llc));
    bw_add(b, s.ml - 3, ml_bits(mlc));
    bw_add(b, ofv, (uint32_t)ofc);
    bw_flush(b);
  }
  for (long n = (long)nseq - 2; n >= 0; --n) {
    const Seq& s = seq[n];
    const int llc = ll_code(s.ll), mlc = ml_code(s.ml - 3);
    const uint32_t ofv = s.off + 3;
    const int ofc = highbit32(ofv);
    fse_encode(b, sof, t.of, ofc);
    fse_encode(b, sml, t.ml, mlc);
    fse_encode(b, sll, t.ll, llc);
    bw_add(b, s.ll, ll_bits(llc));
    bw_add(b, s.ml - 3, ml_bits(mlc));
    bw_add(b, ofv, (uint32_t)ofc);
    bw_flush(b);
  }
  fse_flush_state(b, sml, t.ml);
  fse_flush_state(b, sof, t.of);
  fse_flush_state(b, sll, t.ll);
  out += bw_close(b);
  return (size_t)(out - start);
}

// ---------------------------------------------------------------------------
// Huffman literals.
struct HufCode
{
  uint16_t val[256];
  uint8_t len[256];
  int max_bits;
  int max_sym;
};

// Scratch of the Huffman builder (LDS on the device).
struct HufWork
{
  uint16_t sym[256];
  uint32_t w[512];
  int16_t parent[512];
  uint8_t depth[512];
  uint8_t wt[256];
  uint8_t fse_buf[192];
  uint8_t spread[512];
  uint32_t cumul[66];
  uint16_t blc[kHufMaxBits + 2];
  uint32_t wcount[16]; // weight histogram of the tree description
  int16_t wnorm[16];   // its normalized counts
  FseCTable ct;
};

// Code lengths <= kHufMaxBits forming a complete prefix code. count has 256
// entries; needs >= 2 used symbols.
// Same as huf_build, with ws.sym[0..n) already holding the used symbols in
// ascending count order (the GPU kernel sorts them with the whole wave).
ZHD inline void huf_build_sorted(HufCode& h, const uint32_t* count, int n, HufWork& ws);

ZHD inline void huf_build(HufCode& h, const uint32_t* count, HufWork& ws)
{
  uint16_t* sym = ws.sym;
  int n = 0;
  for (int s = 0; s < 256; ++s)
    if (count[s])
      sym[n++] = (uint16_t)s;
  // ascending by count (stable insertion sort; n <= 256)
  for (int i = 1; i < n; ++i) {
    const uint16_t k = sym[i];
    int j = i;
    while (j > 0 && count[sym[j - 1]] > count[k]) {
      sym[j] = sym[j - 1];
      --j;
    }
    sym[j] = k;
  }
  huf_build_sorted(h, count, n, ws);
}

ZHD inline void huf_build_sorted(HufCode& h, const uint32_t* count, int n, HufWork& ws)
{
  uint16_t* sym = ws.sym;
  uint32_t* w = ws.w;
  int16_t* parent = ws.parent;
  h.max_sym = 0;
  for (int s = 0; s < 256; ++s)
    h.len[s] = 0;
  for (int i = 0; i < n; ++i)
    h.max_sym = sym[i] > h.max_sym ? sym[i] : h.max_sym;
  for (int i = 0; i < n; ++i)
    w[i] = count[sym[i]];
  // two-queue Huffman: leaves 0..n-1, internal nodes n..2n-2
  int leaf = 0, node = n, next = n;
  for (int k = 0; k < n - 1; ++k) {
    int pick[2];
    for (int c = 0; c < 2; ++c) {
      if (leaf < n && (node >= next || w[leaf] <= w[node]))
        pick[c] = leaf++;
      else
        pick[c] = node++;
    }
    w[next] = w[pick[0]] + w[pick[1]];
    parent[pick[0]] = (int16_t)next;
    parent[pick[1]] = (int16_t)next;
    ++next;
  }
  uint8_t* depth = ws.depth;
  depth[next - 1] = 0;
  for (int i = next - 2; i >= 0; --i)
    depth[i] = (uint8_t)(depth[parent[i]] + 1);
  int maxlen = 0;
  for (int i = 0; i < n; ++i)
    maxlen = depth[i] > maxlen ? depth[i] : maxlen;
  if (maxlen > kHufMaxBits) {
    // Length limiting on the per-length counts (O(n + L^2), no per-symbol
    // search): clamp to L, then move codes between lengths until the Kraft
    // sum is exactly 2^L (complete code), and hand the lengths back out in
    // count order, longest to the least frequent symbols.
    const int L = kHufMaxBits;
    uint16_t* bl = ws.blc;
    for (int b = 0; b <= L; ++b)
      bl[b] = 0;
    int64_t kraft = 0;
    for (int i = 0; i < n; ++i) {
      const int d = depth[i] > L ? L : depth[i];
      bl[d]++;
      kraft += (int64_t)1 << (L - d);
    }
    const int64_t full = (int64_t)1 << L;
    while (kraft > full) {
      // lengthen codes at the longest length below L (smallest decrement)
      int b = L - 1;
      while (bl[b] == 0)
        --b;
      const int64_t step = (int64_t)1 << (L - b - 1);
      int64_t k = (kraft - full + step - 1) / step;
      if (k > bl[b])
        k = bl[b];
      bl[b] = (uint16_t)(bl[b] - k);
      bl[b + 1] = (uint16_t)(bl[b + 1] + k);
      kraft -= k * step;
    }
    while (kraft < full) {
      // shorten codes at the longest length whose gain still fits
      int b = L;
      while (b > 1 && (bl[b] == 0 || ((int64_t)1 << (L - b)) > full - kraft))
        --b;
      if (b <= 1)
        break; // cannot happen for n >= 2
      const int64_t step = (int64_t)1 << (L - b);
      int64_t k = (full - kraft) / step;
      if (k > bl[b])
        k = bl[b];
      bl[b] = (uint16_t)(bl[b] - k);
      bl[b - 1] = (uint16_t)(bl[b - 1] + k);
      kraft += k * step;
    }
    int i = 0;
    for (int b = L; b >= 1; --b)
      for (int k = 0; k < bl[b]; ++k)
        depth[i++] = (uint8_t)b;
    maxlen = 0;
    for (int b = L; b >= 1 && !maxlen; --b)
      if (bl[b])
        maxlen = b;
  }
  for (int i = 0; i < n; ++i)
    h.len[sym[i]] = depth[i];
  h.max_bits = maxlen;
  // canonical values: per length, starting value from the longest length down
  uint16_t nb_per[kHufMaxBits + 2] = {}, val_per[kHufMaxBits + 2] = {};
  for (int s = 0; s < 256; ++s)
    if (h.len[s])
      nb_per[h.len[s]]++;
  uint16_t minv = 0;
  for (int l = maxlen; l > 0; --l) {
    val_per[l] = minv;
    minv = (uint16_t)((minv + nb_per[l]) >> 1);
  }
  for (int s = 0; s < 256; ++s)
    h.val[s] = h.len[s] ? val_per[h.len[s]]++ : 0;
}

// Huffman tree description (RFC 8878 4.2.1). Returns bytes written, or 0 if
// the weights cannot be described.
ZHD inline size_t huf_write_tree(uint8_t* out, const HufCode& h, HufWork& ws)
{
  uint8_t* wt = ws.wt;
  const int nw = h.max_sym; // the last symbol's weight is implied
  for (int s = 0; s < nw; ++s)
    wt[s] = h.len[s] ? (uint8_t)(h.max_bits + 1 - h.len[s]) : 0;
  // FSE-compressed weights
  size_t fse_size = 0;
  uint8_t* fse_buf = ws.fse_buf;
  if (nw > 2) {
    uint32_t* cnt = ws.wcount;
    for (int w = 0; w < 16; ++w)
      cnt[w] = 0;
    int maxw = 0, distinct = 0;
    for (int s = 0; s < nw; ++s) {
      if (!cnt[wt[s]]++)
        ++distinct;
      maxw = wt[s] > maxw ? wt[s] : maxw;
    }
    if (distinct >= 2) {
      int16_t* norm = ws.wnorm;
      fse_normalize(norm, cnt, maxw, (uint32_t)nw, 6);
      FseCTable& ct = ws.ct;
      fse_build_ctable(ct, norm, maxw, 6, ws.spread, ws.cumul);
      size_t hsz = fse_write_ncount(fse_buf, norm, maxw, 6);
      const size_t ssz = fse_encode_two_state(fse_buf + hsz, wt, (size_t)nw, ct);
      if (ssz)
        fse_size = hsz + ssz;
    }
  }
  const bool direct_ok = nw <= 128;
  const size_t direct_size = 1 + (size_t)(nw + 1) / 2;
  if (fse_size > 1 && fse_size < 128 && (!direct_ok || fse_size + 1 < direct_size)) {
    out[0] = (uint8_t)fse_size;
    for (size_t i = 0; i < fse_size; ++i)
      out[1 + i] = fse_buf[i];
    return fse_size + 1;
  }
  if (!direct_ok)
    return 0;
  out[0] = (uint8_t)(127 + nw);
  for (int i = 0; i < nw; i += 2)
    out[1 + i / 2] = (uint8_t)((wt[i] << 4) | (i + 1 < nw ? wt[i + 1] : 0));
  return direct_size;
}

// One backward Huffman stream of symbols src[0..n) (serial reference).
ZHD inline size_t huf_encode_stream(uint8_t* out, const uint8_t* src, size_t n, const HufCode& h)
{
  BitW b;
  bw_init(b, out);
  for (long i = (long)n - 1; i >= 0; --i)
    bw_add(b, h.val[src[i]], h.len[src[i]]);
  return bw_close(b);
}

// Literals section header (RFC 8878 3.1.1.3.1.1).
ZHD inline size_t write_raw_lits_header(uint8_t* out, uint32_t n, int type /*0 raw, 1 rle*/)
{
  if (n < 32) {
    out[0] = (uint8_t)(type | (n << 3));
    return 1;
  }
  if (n < 4096) {
    const uint32_t v = (uint32_t)type | (1u << 2) | (n << 4);
    out[0] = (uint8_t)v;
    out[1] = (uint8_t)(v >> 8);
    return 2;
  }
  const uint32_t v = (uint32_t)type | (3u << 2) | (n << 4);
  out[0] = (uint8_t)v;
  out[1] = (uint8_t)(v >> 8);
  out[2] = (uint8_t)(v >> 16);
  return 3;
}
ZHD inline size_t huf_lits_header_size(uint32_t regen, uint32_t comp)
{
  const uint32_t m = regen > comp ? regen : comp;
  return m < 1024 ? 3 : (m < 16384 ? 4 : 5);
}
// type 2: compressed literals (tree description follows); type 3: treeless
// (the previous Huffman table of the frame is reused). Four streams.
ZHD inline void write_huf_lits_header(uint8_t* out, size_t hsize, uint32_t regen, uint32_t comp,
                                      uint32_t type = 2)
{
  uint64_t v;
  if (hsize == 3)
    v = type | (1u << 2) | ((uint64_t)regen << 4) | ((uint64_t)comp << 14);
  else if (hsize == 4)
    v = type | (2u << 2) | ((uint64_t)regen << 4) | ((uint64_t)comp << 18);
  else
    v = type | (3u << 2) | ((uint64_t)regen << 4) | ((uint64_t)comp << 22);
  for (size_t i = 0; i < hsize; ++i)
    out[i] = (uint8_t)(v >> (8 * i));
}

// Frame header: magic, descriptor (4-byte content size, not single segment),
// window descriptor (128 KiB), content size.
constexpr size_t kFrameHeaderBytes = 10;
ZHD inline void write_frame_header(uint8_t* out, uint32_t content_size)
{
  out[0] = 0x28;
  out[1] = 0xB5;
  out[2] = 0x2F;
  out[3] = 0xFD;
  out[4] = 0x80; // FCS field 4 bytes, no checksum, no dictionary
  out[5] = 0x38; // window = 2^(10+7) = 128 KiB
  out[6] = (uint8_t)content_size;
  out[7] = (uint8_t)(content_size >> 8);
  out[8] = (uint8_t)(content_size >> 16);
  out[9] = (uint8_t)(content_size >> 24);
}
ZHD inline void write_block_header(uint8_t* out, bool last, int type, uint32_t size)
{
  const uint32_t v = (last ? 1u : 0u) | ((uint32_t)type << 1) | (size << 3);
  out[0] = (uint8_t)v;
  out[1] = (uint8_t)(v >> 8);
  out[2] = (uint8_t)(v >> 16);
}

} // namespace zcore
