// Stream geometry: every layout number the write path needs, computed once on
// the host from a tile_stream_configuration.
//
// The arithmetic reproduces the reference's planning layer so that chunk
// placement, LOD windows and shard files come out identical:
//   dim partition      src/stream/dim_info.c:25-60, src/dimension.c:178-196
//   LOD level shapes   src/lod/lod_plan.c:213-275   (nlod stopping rule)
//   Morton rank, ends  src/lod/lod_plan.c:21-125
//   level layouts      src/stream/config.c:84-138, src/util/index.ops.c:97-113
//   batch size K       src/stream/config.c:31-45
//   aggregate layout   src/stream/types.aggregate.c:65-143
//   shard geometry     src/stream/config.c:298-372, src/lod/lod_plan.c:395-418
#pragma once

#include "defs.limits.h"
#include "types.stream.h"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace chucky {

inline uint64_t ceildiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }
inline int ceil_log2(uint64_t v)
{
  int p = 0;
  while ((1ull << p) < v)
    ++p;
  return p;
}

// Compacted Morton rank of `coords` inside a box of `shape` (not necessarily
// powers of two), counting only in-bounds cells that precede it in Z order.
// `depth` extra trailing levels rank a coordinate one level finer.
uint64_t morton_rank(int ndim, const uint64_t* shape, const uint64_t* coords, int depth);

struct DimExtent
{
  uint64_t size = 0;
  uint64_t chunk_size = 1;
  uint64_t chunks_per_shard = 1;
  uint64_t chunk_count = 1;
  uint64_t shard_count = 1;
};

// Resolve chunks_per_shard (0 = all, clamped to chunk_count for bounded dims)
// and shard counts; returns the product of shard counts over inner dims.
uint64_t compute_shards(std::vector<DimExtent>& de, int n_append, const uint64_t* config_cps);

struct AggLayout
{
  int lifted_rank = 0; // 2 * (rank - n_append)
  uint64_t lifted_shape[MAX_RANK] = {};
  int64_t lifted_strides[MAX_RANK] = {};
  uint64_t chunks_per_epoch = 0; // M
  uint64_t covering_count = 0;   // C >= M
  uint64_t cps_inner = 1;
  size_t max_comp_chunk_bytes = 0;
  size_t page_size = 0;
  uint64_t num_shards() const { return covering_count / cps_inner; }
  // Position of epoch-chunk j inside the per-epoch covering (shard-major) order.
  uint64_t perm(uint64_t j) const;
};

struct LevelInfo
{
  tile_stream_layout layout{};       // chunk layout of this level's epoch
  uint64_t shape[HALF_MAX_RANK] = {}; // level shape (append dims = chunk size)
  uint64_t lod_nelem = 1;            // product of LOD-dim sizes
  uint64_t chunk_count = 0;          // chunks per epoch
  uint64_t chunk_offset = 0;         // first chunk of this level in an epoch's pool
  AggLayout agg;
  uint32_t batch_active_count = 0;
  uint64_t cps_append = 1, cps_inner = 1, cps_total = 1;
  uint64_t shard_inner_count = 1;
  std::vector<DimExtent> extents; // shard geometry in acquisition order
};

struct Plan
{
  // configuration copy
  tile_stream_configuration config{};
  const dimension* dims = nullptr;
  int rank = 0;
  size_t bpe = 0;

  // dim partition
  int n_append = 1;
  int append_downsample = 0;
  uint32_t lod_mask = 0;
  uint64_t bounded_append_chunks = 1;
  uint8_t storage_order[HALF_MAX_RANK] = {}; // storage position -> dim

  // LOD plan
  int enable_multiscale = 0;
  int nlod = 1;
  int lod_ndim = 0;
  int lod_to_dim[HALF_MAX_RANK] = {};
  int fixed_ndim = 0;
  int fixed_to_dim[HALF_MAX_RANK] = {};
  uint64_t fixed_shape[HALF_MAX_RANK] = {};
  uint64_t fixed_count = 1;
  std::vector<uint64_t> lod_level_ends;   // cumulative lod_nelem
  std::vector<uint64_t> level_span_ends;  // fixed_count * lod_level_ends
  std::vector<uint64_t> ends;             // window ends, all segments packed
  uint64_t segment_begin(int l) const;    // offset of segment l (l -> l+1) in ends

  std::vector<LevelInfo> levels;
  uint64_t total_chunks = 0;
  uint32_t epochs_per_batch = 1;
  size_t max_output_size = 0;
  uint64_t max_cursor_elements = 0; // 0 = unbounded

  uint64_t lod_shape(int lv, int k) const { return levels[lv].shape[lod_to_dim[k]]; }
  uint64_t level_span_begin(int lv) const { return lv ? level_span_ends[lv - 1] : 0; }
  uint64_t append_accum_elements() const;
  uint64_t append_accum_offset(int lv) const;
};

// Fills `out`. codec_alignment pads chunk strides; max_out(codec, chunk_bytes)
// bounds one compressed chunk. Returns 0 on success, otherwise an error string
// is left in *err.
int build_plan(const tile_stream_configuration* config, size_t codec_alignment,
               size_t (*max_out)(enum compression_codec, size_t), Plan* out, std::string* err);

// Host LUTs of the LOD path (built once per stream):
//  gather_lut[m]        source offset (within a fixed-dims batch) of L0 Morton cell m
//  gather_fixed[b]      base offset of fixed-dims batch b in the linear epoch
//  chunk_lut[lv][m]     pool offset of level-lv Morton cell m (within its batch)
//  chunk_fixed[lv][b]   pool offset of fixed-dims batch b at level lv
// Side (log2) of the LOD tile for `ndim` LOD dims and `bpe`-byte elements:
// about 4-8 KiB a tile, rows of at least 16 bytes.
constexpr int lod_tile_side_log(int ndim, size_t bpe)
{
  return ndim == 1 ? (bpe == 1 ? 13 : bpe == 2 ? 12 : bpe == 4 ? 11 : 10)
       : ndim == 2 ? (bpe <= 2 ? 6 : 5)
       : ndim == 3 ? (bpe <= 2 ? 4 : 3)
                   : 0;
}

// Tiled form of a Morton <-> array mapping. A tile is an aligned cube of
// side 2^side_log over the LOD dims lying inside the box (and, for the chunk
// pool, inside one chunk): its cells hold consecutive Morton ranks in local
// Z-order (coordinate k takes bits k, k+D, k+2D, ... of the local index) and
// their offsets are origin + sum(c_k * stride[k]). Cells outside full tiles
// are listed one by one.
struct LodTiles
{
  int ndim = 0;                 // LOD dims (1..3); 0: no tiles (every cell in `edge`)
  int side_log = 0;
  int64_t stride[3] = {};       // offset step per LOD dim inside a tile
  bool vec = false;             // array side moves in 16-byte pieces (rows contiguous
                                // and every row start 16-byte aligned relative to the base)
  std::vector<uint32_t> tiles;  // (first Morton rank, origin offset) pairs
  std::vector<uint32_t> edge;   // (Morton rank, offset) pairs
};

struct LodLuts
{
  std::vector<uint32_t> gather_lut;
  std::vector<uint64_t> gather_fixed;
  std::vector<std::vector<uint32_t>> chunk_lut;
  std::vector<std::vector<uint64_t>> chunk_fixed;
  LodTiles gather_tiles;              // linear epoch -> level-0 Morton order
  std::vector<LodTiles> chunk_tiles;  // level Morton order -> chunk pool, per level
};
void build_lod_luts(const Plan& p, LodLuts* out);
// The tiled forms and batch offsets of build_lod_luts (rank LUTs left
// empty), from a process-wide cache of the geometries built so far.
void cached_lod_luts(const Plan& p, LodLuts* out);

// Aggregate LUTs for one level and one batch: entry a*M+j maps the j-th chunk
// of active epoch a to its pool chunk (gather) and to its slot in the
// shard-major batch order (perm).
void batch_luts(const Plan& p, int lv, uint32_t active_count, const uint32_t* pool_epochs,
                uint32_t* gather, uint32_t* perm);

} // namespace chucky
