// Host-only C entry points: writer helpers, dimension builders, codec
// predicates, CRC-32C and the thread-local error string.
#include "chucky_amd.h"
#include "dimension.h"
#include "types.codec.h"
#include "writer.h"

#include "plan.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

namespace chucky {
thread_local std::string g_last_error;
void set_error(const std::string& msg)
{
  g_last_error = msg;
  std::fprintf(stderr, "chucky-amd: %s\n", msg.c_str());
}
} // namespace chucky

extern "C" {

const char* chucky_last_error(void) { return chucky::g_last_error.c_str(); }

// ---- writer (src/writer.c:6-80) ----

struct writer_result writer_ok(void) { return writer_result{ 0, { nullptr, nullptr } }; }
struct writer_result writer_error(void) { return writer_result{ writer_error_fail, { nullptr, nullptr } }; }
struct writer_result writer_error_at(const void* beg, const void* end)
{
  return writer_result{ writer_error_fail, { beg, end } };
}
struct writer_result writer_finished_at(const void* beg, const void* end)
{
  return writer_result{ writer_error_finished, { beg, end } };
}
struct writer_result writer_append(struct writer* w, struct slice data) { return w->append(w, data); }
struct writer_result writer_flush(struct writer* w) { return w->flush(w); }

struct writer_result writer_append_wait(struct writer* w, struct slice data)
{
  int stalls = 0;
  while (data.beg < data.end) {
    struct writer_result r = writer_append(w, data);
    if (r.error)
      return r;
    if (r.rest.beg == data.beg) {
      if (++stalls >= 10) {
        chucky::set_error("writer_append_wait: no progress after 10 retries");
        return writer_error_at(data.beg, data.end);
      }
      std::this_thread::sleep_for(std::chrono::milliseconds(1ll << (stalls < 6 ? stalls : 6)));
    } else {
      stalls = 0;
    }
    data = r.rest;
  }
  return writer_ok();
}

// ---- codec predicates (src/types.codec.c:3-13) ----

int codec_is_blosc(enum compression_codec c) { return c == CODEC_BLOSC_LZ4 || c == CODEC_BLOSC_ZSTD; }
int codec_is_gpu_supported(enum compression_codec c)
{
  return c == CODEC_NONE || c == CODEC_LZ4_NON_STANDARD || c == CODEC_ZSTD;
}

// ---- dimension builders (src/dimension.c:39-250) ----

static const char* dim_name(unsigned char ch)
{
  // Stable one-character strings: dims[i].name may outlive the caller's buffer.
  static char table[128][2];
  static bool init = false;
  if (!init) {
    for (int i = 0; i < 128; ++i) {
      table[i][0] = (char)i;
      table[i][1] = 0;
    }
    init = true;
  }
  const bool ok = (ch >= 'a' && ch <= 'z') || (ch >= 'A' && ch <= 'Z') || (ch >= '0' && ch <= '9');
  return ok ? table[ch] : nullptr;
}

uint8_t dims_create(struct dimension* dims, const char* names, const uint64_t* sizes)
{
  if (!dims || !names || !sizes)
    return 0;
  const size_t len = std::strlen(names);
  if (len == 0 || len > HALF_MAX_RANK)
    return 0;
  for (size_t i = 0; i < len; ++i) {
    const char* nm = dim_name((unsigned char)names[i]);
    if (!nm)
      return 0;
    dims[i] = dimension{};
    dims[i].size = sizes[i];
    dims[i].chunk_size = sizes[i];
    dims[i].name = nm;
    dims[i].storage_position = (uint8_t)i;
    dims[i].ngff.scale = 1.0;
    dims[i].ngff.type = dimension_axis_space;
  }
  return (uint8_t)len;
}

int dims_set_storage_order(struct dimension* dims, uint8_t rank, const char* order)
{
  if (!order) {
    for (uint8_t i = 0; i < rank; ++i)
      dims[i].storage_position = i;
    return 0;
  }
  if (std::strlen(order) != rank)
    return 1;
  const uint8_t na = dims_n_append(dims, rank);
  for (uint8_t d = 0; d < na; ++d)
    if (order[d] != dims[d].name[0])
      return 1;
  for (uint8_t i = 0; i < rank; ++i) {
    bool found = false;
    for (uint8_t j = 0; j < rank && !found; ++j)
      if (dims[j].name[0] == order[i]) {
        dims[j].storage_position = i;
        found = true;
      }
    if (!found)
      return 1;
  }
  return 0;
}

void dims_set_downsample_by_name(struct dimension* dims, uint8_t rank, const char* names)
{
  for (uint8_t i = 0; i < rank; ++i)
    dims[i].downsample = (names && std::strchr(names, dims[i].name[0])) ? 1 : 0;
}

void dims_set_chunk_sizes(struct dimension* dims, uint8_t rank, const uint64_t* chunk_sizes)
{
  for (uint8_t i = 0; i < rank; ++i)
    dims[i].chunk_size = chunk_sizes[i];
}

void dims_budget_chunk_size(struct dimension* dims, uint8_t rank, uint64_t nelem,
                            const uint8_t* ratios)
{
  if (nelem == 0)
    return;
  const int total_bits = 63 - __builtin_clzll(nelem);
  uint32_t sum = 0;
  for (uint8_t i = 0; i < rank; ++i)
    sum += ratios[i];
  if (sum == 0)
    return;
  int bits[HALF_MAX_RANK] = {};
  for (int b = 0; b < total_bits; ++b) {
    int best = -1;
    for (int i = 0; i < rank; ++i) {
      if (!ratios[i])
        continue;
      // most underserved: lowest bits/ratio, ties to the higher index
      if (best < 0 || bits[i] * ratios[best] <= bits[best] * ratios[i])
        best = i;
    }
    bits[best]++;
  }
  for (uint8_t i = 0; i < rank; ++i)
    dims[i].chunk_size = 1ull << bits[i];
}

void dims_budget_chunk_bytes(struct dimension* dims, uint8_t rank, size_t target_chunk_bytes,
                             size_t bytes_per_element, const uint8_t* ratios)
{
  if (bytes_per_element == 0 || target_chunk_bytes < bytes_per_element)
    return;
  dims_budget_chunk_size(dims, rank, target_chunk_bytes / bytes_per_element, ratios);
}

void dims_set_shard_counts(struct dimension* dims, uint8_t rank, const uint64_t* shard_counts)
{
  for (uint8_t i = 0; i < rank; ++i) {
    if (!shard_counts[i])
      continue;
    const uint64_t n = chucky::ceildiv(dims[i].size, dims[i].chunk_size);
    dims[i].chunks_per_shard = chucky::ceildiv(n, shard_counts[i]);
  }
}

uint8_t dims_n_append(const struct dimension* dims, uint8_t rank)
{
  uint8_t max_n = 0;
  for (uint8_t d = 0; d < rank; ++d) {
    if (dims[d].chunk_size != 1)
      break;
    max_n = d + 1;
  }
  if (max_n <= 1)
    return 1;
  for (uint8_t d = 0; d < max_n; ++d)
    if (dims[d].downsample)
      return d + 1;
  return max_n;
}

int dims_validate(const struct dimension* dims, uint8_t rank)
{
  if (!dims || rank == 0 || rank > HALF_MAX_RANK)
    return 1;
  for (int d = 0; d < rank; ++d)
    if (dims[d].chunk_size == 0)
      return 1;
  for (int d = 1; d < rank; ++d)
    if (dims[d].size == 0)
      return 1;
  if (dims[0].size == 0 && dims[0].chunks_per_shard == 0)
    return 1;
  const uint8_t na = dims_n_append(dims, rank);
  for (int d = 0; d < na; ++d)
    if (dims[d].storage_position != d)
      return 1;
  uint64_t seen = 0;
  for (int d = 0; d < rank; ++d) {
    const unsigned j = dims[d].storage_position;
    if (j >= rank || (seen & (1ull << j)))
      return 1;
    seen |= 1ull << j;
  }
  return 0;
}

void dims_print(const struct dimension* dims, uint8_t rank)
{
  std::fprintf(stderr, "dim  name        size       chunk    chunks     cps    shards  storage  ds\n");
  for (uint8_t i = 0; i < rank; ++i) {
    const uint64_t tc = chucky::ceildiv(dims[i].size, dims[i].chunk_size);
    const uint64_t cps = dims[i].chunks_per_shard ? dims[i].chunks_per_shard : tc;
    std::fprintf(stderr, "%3d  %-4s  %10llu  %10llu  %8llu  %6llu  %8llu  %7d  %s\n", i,
                 dims[i].name ? dims[i].name : "?", (unsigned long long)dims[i].size,
                 (unsigned long long)dims[i].chunk_size, (unsigned long long)tc,
                 (unsigned long long)cps, (unsigned long long)(cps ? chucky::ceildiv(tc, cps) : 0),
                 (int)dims[i].storage_position, dims[i].downsample ? "Y" : ".");
  }
}

// ---- CRC-32C, slicing-by-8 over a table built at first use ----

static uint32_t g_crc_table[8][256];
static void crc_init()
{
  for (uint32_t i = 0; i < 256; ++i) {
    uint32_t c = i;
    for (int k = 0; k < 8; ++k)
      c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
    g_crc_table[0][i] = c;
  }
  for (uint32_t i = 0; i < 256; ++i)
    for (int t = 1; t < 8; ++t)
      g_crc_table[t][i] = (g_crc_table[t - 1][i] >> 8) ^ g_crc_table[0][g_crc_table[t - 1][i] & 0xFF];
}

uint32_t chucky_crc32c(const void* data, size_t len)
{
  static const bool once = (crc_init(), true);
  (void)once;
  const uint8_t* p = (const uint8_t*)data;
  uint32_t c = 0xFFFFFFFFu;
  while (len >= 8) {
    uint32_t lo, hi;
    std::memcpy(&lo, p, 4);
    std::memcpy(&hi, p + 4, 4);
    lo ^= c;
    c = g_crc_table[7][lo & 0xFF] ^ g_crc_table[6][(lo >> 8) & 0xFF] ^
        g_crc_table[5][(lo >> 16) & 0xFF] ^ g_crc_table[4][lo >> 24] ^
        g_crc_table[3][hi & 0xFF] ^ g_crc_table[2][(hi >> 8) & 0xFF] ^
        g_crc_table[1][(hi >> 16) & 0xFF] ^ g_crc_table[0][hi >> 24];
    p += 8;
    len -= 8;
  }
  while (len--)
    c = g_crc_table[0][(c ^ *p++) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

} // extern "C"
