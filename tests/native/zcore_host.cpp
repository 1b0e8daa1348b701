// Host build of the zstd encoding core (chucky_amd/csrc/zstd_core.hpp) with a
// serial greedy parser, so the CPU test suite can check the emitted frames
// against stock libzstd without a GPU. Test-only code.
#include "zstd_core.hpp"

#include <cstdint>
#include <cstring>
#include <vector>

using namespace zcore;

static uint32_t rd32(const uint8_t* p)
{
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

static size_t block(const uint8_t* in, uint32_t len, uint8_t* out, bool last,
                    const Predefined& pre, int force_raw_lits)
{
  bool same = true;
  for (uint32_t i = 1; i < len && same; ++i)
    same = in[i] == in[0];
  if (same) {
    write_block_header(out, last, 1, len);
    out[3] = in[0];
    return 4;
  }
  std::vector<Seq> seqs;
  std::vector<uint8_t> lits;
  std::vector<uint32_t> table(4096, 0);
  uint32_t anchor = 0;
  if (len >= 16) {
    for (uint32_t p = 0; p + 4 <= len;) {
      const uint32_t v = rd32(in + p);
      const uint32_t h = (v * 2654435761u) >> 20;
      const uint32_t c = table[h];
      table[h] = p + 1;
      uint32_t ml = 0;
      if (c && rd32(in + c - 1) == v) {
        ml = 4;
        while (p + ml < len && in[p + ml] == in[c - 1 + ml])
          ++ml;
      }
      if (ml >= 6) { // minimum match of the GPU parser
        uint32_t mc = c - 1, mp = p;
        while (mp > anchor && mc > 0 && in[mp - 1] == in[mc - 1]) {
          --mp;
          --mc;
          ++ml;
        }
        lits.insert(lits.end(), in + anchor, in + mp);
        seqs.push_back(Seq{ mp - anchor, ml, mp - mc });
        p = anchor = mp + ml;
      } else {
        ++p;
      }
    }
  }
  lits.insert(lits.end(), in + anchor, in + len);
  const uint32_t nl = (uint32_t)lits.size();
  std::vector<uint8_t> lsec(nl + 1024), ssec(seqs.size() * 16 + 64);
  size_t lsz = write_raw_lits_header(lsec.data(), nl, 0);
  std::memcpy(lsec.data() + lsz, lits.data(), nl);
  lsz += nl;
  if (nl >= 256 && !force_raw_lits) {
    uint32_t count[256] = {};
    for (uint8_t b : lits)
      count[b]++;
    int used = 0;
    for (int s = 0; s < 256; ++s)
      used += count[s] != 0;
    if (used == 1) {
      lsz = write_raw_lits_header(lsec.data(), nl, 1);
      lsec[lsz++] = lits[0];
    } else {
      HufCode h;
      HufWork ws;
      huf_build(h, count, ws);
      uint8_t tree[192];
      const size_t tsz = huf_write_tree(tree, h, ws);
      if (tsz) {
        const uint32_t seg = (nl + 3) / 4;
        std::vector<uint8_t> st[4];
        uint32_t comp = (uint32_t)tsz + 6;
        bool fits = true;
        for (int k = 0; k < 4; ++k) {
          const uint32_t a = k * seg, b = k == 3 ? nl : a + seg;
          st[k].resize((b - a) * 2 + 16);
          st[k].resize(huf_encode_stream(st[k].data(), lits.data() + a, b - a, h));
          fits &= k == 3 || st[k].size() <= 0xFFFF;
          comp += (uint32_t)st[k].size();
        }
        const size_t hs = huf_lits_header_size(nl, comp);
        if (fits && hs + comp < lsz) {
          write_huf_lits_header(lsec.data(), hs, nl, comp);
          size_t o = hs;
          std::memcpy(lsec.data() + o, tree, tsz);
          o += tsz;
          for (int k = 0; k < 3; ++k) {
            lsec[o++] = (uint8_t)st[k].size();
            lsec[o++] = (uint8_t)(st[k].size() >> 8);
          }
          for (int k = 0; k < 4; ++k) {
            std::memcpy(lsec.data() + o, st[k].data(), st[k].size());
            o += st[k].size();
          }
          lsz = o;
        }
      }
    }
  }
  const size_t ssz = write_sequences(ssec.data(), seqs.data(), (uint32_t)seqs.size(), pre);
  const size_t body = lsz + ssz;
  if (body >= len) {
    write_block_header(out, last, 0, len);
    std::memcpy(out + 3, in, len);
    return 3 + len;
  }
  write_block_header(out, last, 2, (uint32_t)body);
  std::memcpy(out + 3, lsec.data(), lsz);
  std::memcpy(out + 3 + lsz, ssec.data(), ssz);
  return 3 + body;
}

extern "C" size_t zcore_host_compress(const uint8_t* src, size_t n, uint8_t* dst, int force_raw_lits)
{
  static Predefined pre;
  static bool init = false;
  if (!init) {
    uint8_t spread[512];
    uint32_t cumul[66];
    build_predefined(pre, spread, cumul);
    init = true;
  }
  write_frame_header(dst, (uint32_t)n);
  size_t o = kFrameHeaderBytes;
  for (size_t b = 0; b < n; b += kBlockMax) {
    const uint32_t len = (uint32_t)(n - b < kBlockMax ? n - b : kBlockMax);
    o += block(src + b, len, dst + o, b + len == n, pre, force_raw_lits);
  }
  return o;
}
