// Host-side stream geometry. See plan.hpp for the reference functions each
// part reproduces.
#include "plan.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace chucky {

namespace {

// fn(lo, hi) over [0, n) split into contiguous ranges on up to 16 threads
// (the calling thread takes the first); small ranges stay on the caller.
// The per-element LUTs of a large LOD level are millions of Morton ranks.
template<typename F>
void parallel_ranges(uint64_t n, F fn)
{
  constexpr uint64_t kMinPerThread = 1u << 15;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const uint64_t nt = std::min<uint64_t>({ (uint64_t)std::min(hw, 16u), n / kMinPerThread, 64 });
  if (nt <= 1) {
    fn((uint64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  const uint64_t step = (n + nt - 1) / nt;
  for (uint64_t t = 1; t < nt; ++t) {
    const uint64_t lo = std::min(n, t * step), hi = std::min(n, lo + step);
    th.emplace_back([=] { fn(lo, hi); });
  }
  fn((uint64_t)0, std::min(n, step));
  for (auto& x : th)
    x.join();
}

uint64_t max_of(int n, const uint64_t* v)
{
  uint64_t m = 0;
  for (int i = 0; i < n; ++i)
    m = std::max(m, v[i]);
  return m;
}

// Cells of [lo, lo+scale) that fall inside [0, extent).
uint64_t cells_inside(uint64_t extent, uint64_t lo, uint64_t scale)
{
  if (lo >= extent)
    return 0;
  return std::min(extent - lo, scale);
}

// Next cell in (uncompacted) Z order over a 2^p box; dim 0 is the lowest bit
// of each digit. Past the last cell the result is (2^p, 0, 0, ...).
void z_increment(int ndim, int p, uint64_t* c)
{
  for (int bit = 0; bit < p; ++bit)
    for (int d = 0; d < ndim; ++d) {
      const uint64_t m = 1ull << bit;
      c[d] ^= m;
      if (c[d] & m)
        return;
    }
  for (int d = 0; d < ndim; ++d)
    c[d] = 0;
  c[0] = 1ull << p;
}

uint32_t next_pow2(uint32_t v)
{
  uint32_t r = 1;
  while (r < v)
    r <<= 1;
  return r;
}

} // namespace

// src/lod/lod_plan.c:21-72. Walk the Z-order digits from the top; at each
// level add the in-bounds cells of the sibling sub-boxes that precede the
// coordinate's own sub-box.
uint64_t morton_rank(int ndim, const uint64_t* shape, const uint64_t* coords, int depth)
{
  int p = ceil_log2(max_of(ndim, shape));
  for (int d = 0; d < ndim; ++d)
    if (coords[d] > 0)
      p = std::max(p, ceil_log2(coords[d] + 1));
  const int levels = p + depth;

  uint64_t rank = 0;
  uint64_t prefix[LOD_MAX_NDIM] = {};
  for (int lv = 0; lv < levels; ++lv) {
    const uint64_t scale = 1ull << (levels - 1 - lv);
    unsigned digit = 0;
    if (lv < p)
      for (int d = 0; d < ndim; ++d)
        digit |= (unsigned)((coords[d] >> (p - 1 - lv)) & 1) << d;

    uint64_t lo_half[LOD_MAX_NDIM], hi_half[LOD_MAX_NDIM];
    for (int d = 0; d < ndim; ++d) {
      lo_half[d] = cells_inside(shape[d], (prefix[d] * 2) * scale, scale);
      hi_half[d] = cells_inside(shape[d], (prefix[d] * 2 + 1) * scale, scale);
    }
    // below[d] = cells of a sub-box whose dims < d range over both halves
    uint64_t below[LOD_MAX_NDIM + 1];
    below[0] = 1;
    for (int d = 0; d < ndim; ++d)
      below[d + 1] = below[d] * (lo_half[d] + hi_half[d]);
    uint64_t fixed_above = 1;
    for (int d = ndim - 1; d >= 0; --d) {
      const bool hi = (digit >> d) & 1;
      if (hi)
        rank += fixed_above * lo_half[d] * below[d];
      fixed_above *= hi ? hi_half[d] : lo_half[d];
    }
    for (int d = 0; d < ndim; ++d)
      prefix[d] = prefix[d] * 2 + ((digit >> d) & 1);
  }
  return rank;
}

// src/lod/lod_plan.c:395-418
uint64_t compute_shards(std::vector<DimExtent>& de, int n_append, const uint64_t* config_cps)
{
  uint64_t inner = 1;
  for (size_t d = 0; d < de.size(); ++d) {
    DimExtent& e = de[d];
    e.chunk_count = e.size == 0 ? 1 : ceildiv(e.size, e.chunk_size);
    uint64_t cps = config_cps[d] ? config_cps[d] : e.chunk_count;
    if (e.size > 0 && e.chunk_count < cps)
      cps = e.chunk_count;
    e.chunks_per_shard = cps;
    e.shard_count = ceildiv(e.chunk_count, cps);
    if ((int)d >= n_append)
      inner *= e.shard_count;
  }
  return inner;
}

// src/util/index.ops.c:42-53 (ravel) over the aggregate's lifted space.
uint64_t AggLayout::perm(uint64_t j) const
{
  uint64_t o = 0, rest = j;
  for (int d = lifted_rank - 1; d >= 0; --d) {
    o += (rest % lifted_shape[d]) * (uint64_t)lifted_strides[d];
    rest /= lifted_shape[d];
  }
  return o;
}

uint64_t Plan::segment_begin(int l) const
{
  // segment l holds the windows of level l+1 cells, packed after level 0.
  return lod_level_ends[l] - lod_level_ends[0];
}

uint64_t Plan::append_accum_elements() const
{
  uint64_t n = 0;
  for (int lv = 1; lv < nlod; ++lv)
    n += fixed_count * levels[lv].lod_nelem;
  return n;
}

uint64_t Plan::append_accum_offset(int lv) const
{
  uint64_t n = 0;
  for (int k = 1; k < lv; ++k)
    n += fixed_count * levels[k].lod_nelem;
  return n;
}

namespace {

int fail(std::string* err, const char* msg)
{
  if (err)
    *err = msg;
  return 1;
}

// src/dimension.c:178-196
int n_append_of(const dimension* dims, int rank)
{
  int max_n = 0;
  for (int d = 0; d < rank; ++d) {
    if (dims[d].chunk_size != 1)
      break;
    max_n = d + 1;
  }
  if (max_n <= 1)
    return 1;
  for (int d = 0; d < max_n; ++d)
    if (dims[d].downsample)
      return d + 1;
  return max_n;
}

// src/stream/config.c:84-138 with src/util/index.ops.c:97-113.
int level_layout(tile_stream_layout* L, int rank, int na, size_t bpe, const dimension* dims,
                 const uint64_t* shape, size_t alignment, const uint8_t* order)
{
  memset(L, 0, sizeof(*L));
  L->lifted_rank = (uint8_t)(2 * rank);
  L->chunk_elements = 1;
  uint64_t cc[HALF_MAX_RANK];
  for (int d = 0; d < rank; ++d) {
    cc[d] = shape[d] == 0 ? 1 : ceildiv(shape[d], dims[d].chunk_size);
    L->lifted_shape[2 * d] = cc[d];
    L->lifted_shape[2 * d + 1] = dims[d].chunk_size;
    L->chunk_elements *= dims[d].chunk_size;
  }
  L->chunk_stride = align_up(L->chunk_elements * bpe, alignment) / bpe;
  // strides assigned innermost storage position first
  int64_t elem_stride = 1, chunk_stride = (int64_t)L->chunk_stride;
  for (int j = rank - 1; j >= 0; --j) {
    const int d = order[j];
    L->lifted_strides[2 * d + 1] = elem_stride;
    elem_stride *= (int64_t)dims[d].chunk_size;
    L->lifted_strides[2 * d] = chunk_stride;
    chunk_stride *= (int64_t)cc[d];
  }
  L->chunks_per_epoch = 1;
  for (int d = na; d < rank; ++d)
    L->chunks_per_epoch *= cc[d];
  if (L->chunk_elements && L->chunks_per_epoch > UINT64_MAX / L->chunk_elements)
    return 1;
  L->epoch_elements = L->chunks_per_epoch * L->chunk_elements;
  for (int d = 0; d < na; ++d)
    L->lifted_strides[2 * d] = 0;
  L->chunk_pool_bytes = L->chunks_per_epoch * L->chunk_stride * bpe;
  return 0;
}

// src/stream/types.aggregate.c:65-143 (inputs in storage order).
int agg_layout(AggLayout* a, int rank, int na, const uint64_t* cc, const uint64_t* cps,
               uint64_t M, size_t max_comp, size_t page)
{
  *a = AggLayout{};
  a->chunks_per_epoch = M;
  a->max_comp_chunk_bytes = max_comp;
  a->page_size = page;
  a->lifted_rank = 2 * (rank - na);
  uint64_t sc[HALF_MAX_RANK];
  a->covering_count = 1;
  a->cps_inner = 1;
  for (int d = na; d < rank; ++d) {
    if (cps[d] < 1)
      return 1;
    sc[d] = ceildiv(cc[d], cps[d]);
    const int k = d - na;
    a->lifted_shape[2 * k] = sc[d];
    a->lifted_shape[2 * k + 1] = cps[d];
    a->covering_count *= sc[d] * cps[d];
    a->cps_inner *= cps[d];
  }
  if (a->covering_count > UINT32_MAX)
    return 1;
  uint64_t shard_acc = 1, within_acc = 1;
  for (int d = rank - 1; d >= na; --d) {
    const int k = d - na;
    a->lifted_strides[2 * k] = (int64_t)(shard_acc * a->cps_inner);
    shard_acc *= sc[d];
    a->lifted_strides[2 * k + 1] = (int64_t)within_acc;
    within_acc *= cps[d];
  }
  return 0;
}

int validate_dims(const dimension* dims, int rank, std::string* err)
{
  if (!dims || rank <= 0 || rank > HALF_MAX_RANK)
    return fail(err, "invalid rank or dimensions");
  for (int d = 0; d < rank; ++d)
    if (dims[d].chunk_size == 0)
      return fail(err, "chunk_size must be > 0");
  for (int d = 1; d < rank; ++d)
    if (dims[d].size == 0)
      return fail(err, "only dim 0 may be unbounded");
  if (dims[0].size == 0 && dims[0].chunks_per_shard == 0)
    return fail(err, "unbounded dim 0 requires chunks_per_shard > 0");
  const int na = n_append_of(dims, rank);
  for (int d = 0; d < na; ++d)
    if (dims[d].storage_position != d)
      return fail(err, "append dims must keep their storage position");
  uint64_t seen = 0;
  for (int d = 0; d < rank; ++d) {
    const unsigned j = dims[d].storage_position;
    if ((int)j >= rank || (seen & (1ull << j)))
      return fail(err, "storage_position is not a permutation");
    seen |= 1ull << j;
  }
  return 0;
}

} // namespace

int build_plan(const tile_stream_configuration* config, size_t codec_alignment,
               size_t (*max_out)(enum compression_codec, size_t), Plan* out, std::string* err)
{
  *out = Plan{};
  Plan& p = *out;
  if (!config)
    return fail(err, "null configuration");
  p.config = *config;
  p.dims = config->dimensions;
  p.rank = config->rank;
  p.bpe = dtype_bpe(config->dtype);
  const dimension* dims = p.dims;
  const int rank = p.rank;

  // ---- validation (src/stream/config.c:144-221) ----
  if (p.bpe == 0)
    return fail(err, "unknown dtype");
  if (config->buffer_capacity_bytes == 0)
    return fail(err, "buffer_capacity_bytes must be > 0");
  if (config->max_threads < 0)
    return fail(err, "max_threads must be >= 0");
  if (validate_dims(dims, rank, err))
    return 1;
  if (config->codec.id == CODEC_LZ4_NON_STANDARD && config->codec.level == 0)
    return fail(err, "LZ4 requires level >= 1");
  if ((config->codec.id == CODEC_BLOSC_LZ4 || config->codec.id == CODEC_BLOSC_ZSTD) &&
      config->codec.level > 9)
    return fail(err, "blosc level must be 0..9");
  if (config->max_nlod < 0 || config->max_nlod > LOD_MAX_LEVELS)
    return fail(err, "max_nlod out of range");

  // ---- dim partition (src/stream/dim_info.c:25-60) ----
  const int na = n_append_of(dims, rank);
  p.n_append = na;
  for (int d = na; d < rank; ++d)
    if (dims[d].downsample)
      p.lod_mask |= 1u << d;
  p.append_downsample = dims[na - 1].downsample ? 1 : 0;
  for (int d = 1; d < na; ++d)
    p.bounded_append_chunks *= ceildiv(dims[d].size, dims[d].chunk_size);
  if (p.append_downsample) {
    const auto m = config->append_reduce_method;
    if (m != lod_reduce_mean && m != lod_reduce_min && m != lod_reduce_max)
      return fail(err, "append reduce method must be mean, min or max");
    if (p.lod_mask == 0)
      return fail(err, "append downsample requires a downsampled inner dim");
  }
  for (int d = 0; d < rank; ++d)
    p.storage_order[dims[d].storage_position] = (uint8_t)d;

  // ---- LOD plan on epoch dims (src/lod/lod_plan.c:151-275,328-345) ----
  uint64_t shape0[HALF_MAX_RANK], chunk[HALF_MAX_RANK];
  for (int d = 0; d < rank; ++d) {
    shape0[d] = dims[d].size == 0 ? dims[d].chunk_size : dims[d].size;
    chunk[d] = dims[d].chunk_size;
  }
  for (int d = 0; d < na; ++d)
    shape0[d] = dims[d].chunk_size;
  for (int d = 0; d < rank; ++d) {
    if (p.lod_mask & (1u << d))
      p.lod_to_dim[p.lod_ndim++] = d;
    else {
      p.fixed_to_dim[p.fixed_ndim] = d;
      p.fixed_shape[p.fixed_ndim++] = shape0[d];
    }
  }
  for (int k = 0; k < p.fixed_ndim; ++k)
    p.fixed_count *= p.fixed_shape[k];

  const int max_levels = config->max_nlod == 0 ? LOD_MAX_LEVELS : config->max_nlod;
  std::vector<std::vector<uint64_t>> shapes{ std::vector<uint64_t>(shape0, shape0 + rank) };
  for (;;) {
    if ((int)shapes.size() >= max_levels)
      break;
    const auto& cur = shapes.back();
    bool all_one = true;
    for (int k = 0; k < p.lod_ndim; ++k)
      if (ceildiv(cur[p.lod_to_dim[k]], chunk[p.lod_to_dim[k]]) > 1)
        all_one = false;
    if (all_one)
      break;
    auto next = cur;
    for (int k = 0; k < p.lod_ndim; ++k)
      next[p.lod_to_dim[k]] = (cur[p.lod_to_dim[k]] + 1) / 2;
    shapes.push_back(next);
  }
  p.enable_multiscale = p.lod_mask != 0;
  p.nlod = p.enable_multiscale ? (int)shapes.size() : 1;
  if (!p.enable_multiscale)
    shapes.resize(1);
  p.levels.resize(p.nlod);
  for (int lv = 0; lv < p.nlod; ++lv) {
    std::copy(shapes[lv].begin(), shapes[lv].end(), p.levels[lv].shape);
    p.levels[lv].lod_nelem = 1;
    for (int k = 0; k < p.lod_ndim; ++k)
      p.levels[lv].lod_nelem *= p.lod_shape(lv, k);
  }
  if (p.enable_multiscale) {
    p.lod_level_ends.resize(p.nlod);
    p.level_span_ends.resize(p.nlod);
    uint64_t acc = 0;
    for (int lv = 0; lv < p.nlod; ++lv) {
      acc += p.levels[lv].lod_nelem;
      p.lod_level_ends[lv] = acc;
      p.level_span_ends[lv] = p.fixed_count * acc;
    }
    p.ends.assign(p.lod_level_ends[p.nlod - 1] - p.lod_level_ends[0], 0);
    // src/lod/lod_plan.c:105-125: window end of each parent cell
    for (int l = 0; l + 1 < p.nlod; ++l) {
      uint64_t child[LOD_MAX_NDIM], parent[LOD_MAX_NDIM];
      for (int k = 0; k < p.lod_ndim; ++k) {
        child[k] = p.lod_shape(l, k);
        parent[k] = p.lod_shape(l + 1, k);
      }
      const int zp = ceil_log2(max_of(p.lod_ndim, parent));
      const uint64_t n_par = p.levels[l + 1].lod_nelem;
      uint64_t* seg = p.ends.data() + p.segment_begin(l);
      const int nd = p.lod_ndim;
      parallel_ranges(n_par, [&](uint64_t j0, uint64_t j1) {
        for (uint64_t j = j0; j < j1; ++j) {
          uint64_t c[LOD_MAX_NDIM], nx[LOD_MAX_NDIM];
          uint64_t r = j;
          for (int k = 0; k < nd; ++k) {
            c[k] = r % parent[k];
            r /= parent[k];
          }
          const uint64_t pos = morton_rank(nd, parent, c, 0);
          std::copy(c, c + nd, nx);
          z_increment(nd, zp, nx);
          seg[pos] = morton_rank(nd, child, nx, 1);
        }
      });
    }
  }

  // ---- per-level chunk layouts ----
  for (int lv = 0; lv < p.nlod; ++lv)
    if (level_layout(&p.levels[lv].layout, rank, na, p.bpe, dims, p.levels[lv].shape,
                     codec_alignment, p.storage_order))
      return fail(err, "epoch size overflows");
  for (int lv = 0; lv < p.nlod; ++lv) {
    p.levels[lv].chunk_count = p.levels[lv].layout.chunks_per_epoch;
    p.levels[lv].chunk_offset = p.total_chunks;
    p.total_chunks += p.levels[lv].chunk_count;
  }

  // ---- K (src/stream/config.c:31-45) ----
  {
    uint32_t K = config->epochs_per_batch;
    if (K == 0) {
      const uint32_t target = config->target_batch_chunks ? config->target_batch_chunks : 1024;
      K = next_pow2((uint32_t)ceildiv(target, p.total_chunks));
    }
    p.epochs_per_batch = std::min<uint32_t>(K, MAX_BATCH_EPOCHS);
  }

  // ---- codec bound ----
  p.max_output_size = max_out(config->codec.id, p.levels[0].layout.chunk_stride * p.bpe);
  if (config->codec.id != CODEC_NONE && p.max_output_size == 0)
    return fail(err, "codec not supported on the GPU");

  // ---- aggregate layouts and shard geometry (src/stream/config.c:298-372) ----
  uint64_t cps_cfg[HALF_MAX_RANK];
  for (int d = 0; d < rank; ++d)
    cps_cfg[d] = dims[d].chunks_per_shard;
  for (int lv = 0; lv < p.nlod; ++lv) {
    LevelInfo& L = p.levels[lv];
    L.extents.assign(rank, DimExtent{});
    for (int d = 0; d < rank; ++d) {
      L.extents[d].size = lv == 0 ? dims[d].size : L.shape[d];
      L.extents[d].chunk_size = dims[d].chunk_size;
    }
    L.shard_inner_count = compute_shards(L.extents, na, cps_cfg);
    uint32_t bc = p.epochs_per_batch;
    if (p.append_downsample && lv > 0) {
      const uint32_t period = 1u << lv;
      bc = p.epochs_per_batch >= period ? p.epochs_per_batch / period : 0;
    }
    L.batch_active_count = bc;
    uint64_t so_cc[HALF_MAX_RANK], so_cps[HALF_MAX_RANK];
    for (int j = 0; j < rank; ++j) {
      so_cc[j] = L.extents[p.storage_order[j]].chunk_count;
      so_cps[j] = L.extents[p.storage_order[j]].chunks_per_shard;
    }
    if (agg_layout(&L.agg, rank, na, so_cc, so_cps, L.chunk_count, p.max_output_size,
                   config->shard_alignment))
      return fail(err, "aggregate layout overflow");
    L.cps_append = 1;
    for (int d = 0; d < na; ++d)
      L.cps_append *= L.extents[d].chunks_per_shard;
    if (p.append_downsample && lv > 0) {
      const uint64_t div = 1ull << lv;
      L.cps_append = L.cps_append > div ? L.cps_append / div : 1;
    }
    L.cps_inner = 1;
    for (int d = na; d < rank; ++d)
      L.cps_inner *= L.extents[d].chunks_per_shard;
    L.cps_total = L.cps_append * L.cps_inner;
  }

  // ---- capacity of bounded append dims (src/cpu/stream.c:234-246) ----
  if (dims[0].size > 0) {
    p.max_cursor_elements = p.levels[0].layout.epoch_elements;
    for (int d = 0; d < na; ++d)
      p.max_cursor_elements *= ceildiv(dims[d].size, dims[d].chunk_size);
  }
  return 0;
}

// src/stream/types.aggregate.c:28-62
void batch_luts(const Plan& p, int lv, uint32_t active_count, const uint32_t* pool_epochs,
                uint32_t* gather, uint32_t* perm)
{
  const LevelInfo& L = p.levels[lv];
  const uint64_t M = L.agg.chunks_per_epoch;
  const uint64_t cpsi = L.agg.cps_inner;
  for (uint32_t a = 0; a < active_count; ++a)
    for (uint64_t j = 0; j < M; ++j) {
      const uint64_t idx = (uint64_t)a * M + j;
      gather[idx] = (uint32_t)(pool_epochs[a] * p.total_chunks + L.chunk_offset + j);
      const uint64_t q = L.agg.perm(j);
      // [shard][active epoch][chunk in shard]
      perm[idx] = (uint32_t)((q / cpsi) * active_count * cpsi + (q % cpsi) + a * cpsi);
    }
}

namespace {

// Tiles of a Morton mapping `lut` (rank -> offset) over the box `shape`
// (ndim <= 3 LOD dims). `in_one_piece(origin)` says whether the tile at that
// origin keeps one linear stride pattern (same chunk); every accepted tile is
// checked cell by cell against the LUT, so a tile the formula gets wrong falls
// back to the per-cell list instead of producing wrong bytes.
template<typename OnePiece>
void build_tiles(int ndim, size_t bpe, const uint64_t* shape, const std::vector<uint32_t>& lut,
                 const int64_t* stride, const std::vector<uint64_t>& fixed, OnePiece in_one_piece,
                 LodTiles* t)
{
  t->tiles.clear();
  t->edge.clear();
  t->ndim = ndim >= 1 && ndim <= 3 ? ndim : 0;
  t->side_log = lod_tile_side_log(t->ndim, bpe);
  t->vec = false;
  for (int k = 0; k < 3; ++k)
    t->stride[k] = k < ndim ? stride[k] : 0;
  const uint64_t n = lut.size();
  std::vector<uint8_t> covered(n, 0);
  if (t->side_log) {
    const uint64_t side = 1ull << t->side_log;
    const uint64_t cells = 1ull << (t->side_log * ndim);
    uint64_t grid[3] = { 1, 1, 1 }, ntiles = 1;
    for (int k = 0; k < ndim; ++k) {
      grid[k] = shape[k] / side;
      ntiles *= grid[k];
    }
    // each tile is checked on its own (in parallel); accepted tiles are
    // then listed in tile order
    const int side_log = t->side_log;
    std::vector<uint64_t> tile_base(ntiles, ~0ull);
    parallel_ranges(ntiles * cells, [&](uint64_t c0, uint64_t c1) {
      for (uint64_t ti = c0 / cells; ti < ntiles && ti * cells < c1; ++ti) {
        if (ti * cells < c0)
          continue; // a tile belongs to the range holding its first cell
        uint64_t origin[3] = { 0, 0, 0 }, r = ti;
        for (int k = ndim - 1; k >= 0; --k) {
          origin[k] = (r % grid[k]) * side;
          r /= grid[k];
        }
        if (!in_one_piece(origin))
          continue;
        const uint64_t base = morton_rank(ndim, shape, origin, 0);
        if (base + cells > n || base > 0xFFFFFFFFull)
          continue;
        const uint32_t off0 = lut[base];
        bool ok = true;
        for (uint64_t z = 0; z < cells && ok; ++z) {
          int64_t off = off0;
          for (int k = 0; k < ndim; ++k) {
            uint64_t c = 0;
            for (int j = 0; j < side_log; ++j)
              c |= ((z >> (j * ndim + k)) & 1ull) << j;
            off += (int64_t)c * stride[k];
          }
          ok = (int64_t)lut[base + z] == off;
        }
        if (ok)
          tile_base[ti] = base;
      }
    });
    for (uint64_t ti = 0; ti < ntiles; ++ti) {
      const uint64_t base = tile_base[ti];
      if (base == ~0ull)
        continue;
      t->tiles.push_back((uint32_t)base);
      t->tiles.push_back(lut[base]);
      std::fill(covered.begin() + (int64_t)base, covered.begin() + (int64_t)(base + cells), (uint8_t)1);
    }
  }
  for (uint64_t m = 0; m < n; ++m)
    if (!covered[m]) {
      t->edge.push_back((uint32_t)m);
      t->edge.push_back(lut[m]);
    }
  // 16-byte pieces on the array side: the fastest coordinate has stride 1,
  // a row holds whole pieces, and every other step (other strides, tile
  // origins, batch offsets) is a multiple of a piece
  if (t->ndim && !t->tiles.empty()) {
    const uint64_t V = 16 / bpe;
    int fast = 0;
    for (int k = 1; k < ndim; ++k)
      if (std::llabs(stride[k]) < std::llabs(stride[fast]))
        fast = k;
    bool ok = stride[fast] == 1 && ((1ull << t->side_log) % V) == 0;
    for (int k = 0; k < ndim && ok; ++k)
      ok = k == fast || (uint64_t)std::llabs(stride[k]) % V == 0;
    for (size_t i = 1; i < t->tiles.size() && ok; i += 2)
      ok = t->tiles[i] % V == 0;
    for (size_t b = 0; b < fixed.size() && ok; ++b)
      ok = fixed[b] % V == 0;
    t->vec = ok;
  }
}

} // namespace

// src/cpu/lod.cpp:182-268 (gather LUT and batch offsets), :298-359 and
// src/cpu/pipeline.c:369-397 (chunk LUT and batch offsets).
void build_lod_luts(const Plan& p, LodLuts* out)
{
  const int rank = p.rank;
  const uint64_t* full = p.levels[0].shape;
  uint64_t fstride[HALF_MAX_RANK];
  fstride[rank - 1] = 1;
  for (int d = rank - 2; d >= 0; --d)
    fstride[d] = fstride[d + 1] * full[d + 1];

  uint64_t lshape[LOD_MAX_NDIM];
  for (int k = 0; k < p.lod_ndim; ++k)
    lshape[k] = p.lod_shape(0, k);
  const uint64_t n0 = p.levels[0].lod_nelem;
  out->gather_lut.assign(n0, 0);
  uint32_t* glut = out->gather_lut.data();
  parallel_ranges(n0, [&](uint64_t g0, uint64_t g1) {
    for (uint64_t g = g0; g < g1; ++g) {
      uint64_t c[LOD_MAX_NDIM], r = g, off = 0;
      for (int k = p.lod_ndim - 1; k >= 0; --k) {
        c[k] = r % lshape[k];
        r /= lshape[k];
        off += c[k] * fstride[p.lod_to_dim[k]];
      }
      glut[morton_rank(p.lod_ndim, lshape, c, 0)] = (uint32_t)off;
    }
  });
  out->gather_fixed.assign(p.fixed_count, 0);
  for (uint64_t b = 0; b < p.fixed_count; ++b) {
    uint64_t r = b, off = 0;
    for (int k = p.fixed_ndim - 1; k >= 0; --k) {
      off += (r % p.fixed_shape[k]) * fstride[p.fixed_to_dim[k]];
      r /= p.fixed_shape[k];
    }
    out->gather_fixed[b] = off;
  }

  out->chunk_lut.assign(p.nlod, {});
  out->chunk_fixed.assign(p.nlod, {});
  for (int lv = 0; lv < p.nlod; ++lv) {
    const tile_stream_layout& L = p.levels[lv].layout;
    uint64_t ls[LOD_MAX_NDIM];
    for (int k = 0; k < p.lod_ndim; ++k)
      ls[k] = p.lod_shape(lv, k);
    const uint64_t n = p.levels[lv].lod_nelem;
    auto& lut = out->chunk_lut[lv];
    lut.assign(n, 0);
    uint32_t* clut = lut.data();
    parallel_ranges(n, [&](uint64_t g0, uint64_t g1) {
      for (uint64_t g = g0; g < g1; ++g) {
        uint64_t c[LOD_MAX_NDIM], r = g;
        int64_t off = 0;
        for (int k = p.lod_ndim - 1; k >= 0; --k) {
          c[k] = r % ls[k];
          r /= ls[k];
          const int d = p.lod_to_dim[k];
          const uint64_t cs = L.lifted_shape[2 * d + 1];
          off += (int64_t)(c[k] / cs) * L.lifted_strides[2 * d] +
                 (int64_t)(c[k] % cs) * L.lifted_strides[2 * d + 1];
        }
        clut[morton_rank(p.lod_ndim, ls, c, 0)] = (uint32_t)off;
      }
    });
    auto& fx = out->chunk_fixed[lv];
    fx.assign(p.fixed_count, 0);
    for (uint64_t b = 0; b < p.fixed_count; ++b) {
      uint64_t r = b;
      int64_t off = 0;
      for (int k = p.fixed_ndim - 1; k >= 0; --k) {
        const uint64_t coord = r % p.fixed_shape[k];
        r /= p.fixed_shape[k];
        const int d = p.fixed_to_dim[k];
        const uint64_t cs = L.lifted_shape[2 * d + 1];
        off += (int64_t)(coord / cs) * L.lifted_strides[2 * d] +
               (int64_t)(coord % cs) * L.lifted_strides[2 * d + 1];
      }
      fx[b] = (uint64_t)off + p.levels[lv].chunk_offset * L.chunk_stride;
    }
  }

  // tiled forms (LOD over at most 3 dims)
  out->chunk_tiles.assign(p.nlod, LodTiles{});
  if (p.lod_ndim >= 1 && p.lod_ndim <= 3) {
    int64_t gs[3] = {};
    for (int k = 0; k < p.lod_ndim; ++k)
      gs[k] = (int64_t)fstride[p.lod_to_dim[k]];
    build_tiles(p.lod_ndim, p.bpe, lshape, out->gather_lut, gs, out->gather_fixed,
                [](const uint64_t*) { return true; }, &out->gather_tiles);
    for (int lv = 0; lv < p.nlod; ++lv) {
      const tile_stream_layout& L = p.levels[lv].layout;
      uint64_t ls[3], cs[3];
      int64_t st[3];
      for (int k = 0; k < p.lod_ndim; ++k) {
        const int d = p.lod_to_dim[k];
        ls[k] = p.lod_shape(lv, k);
        cs[k] = L.lifted_shape[2 * d + 1];
        st[k] = L.lifted_strides[2 * d + 1];
      }
      const int nd = p.lod_ndim;
      const uint64_t side = 1ull << lod_tile_side_log(nd, p.bpe);
      build_tiles(nd, p.bpe, ls, out->chunk_lut[lv], st, out->chunk_fixed[lv],
                  [&](const uint64_t* o) {
                    for (int k = 0; k < nd; ++k)
                      if (o[k] / cs[k] != (o[k] + side - 1) / cs[k])
                        return false;
                    return true;
                  },
                  &out->chunk_tiles[lv]);
    }
  } else {
    // no tiled path: every cell in the per-cell lists
    build_tiles(0, p.bpe, lshape, out->gather_lut, nullptr, out->gather_fixed,
                [](const uint64_t*) { return true; }, &out->gather_tiles);
    for (int lv = 0; lv < p.nlod; ++lv)
      build_tiles(0, p.bpe, lshape, out->chunk_lut[lv], nullptr, out->chunk_fixed[lv],
                  [](const uint64_t*) { return true; }, &out->chunk_tiles[lv]);
  }
}

// Tile lists, edge lists and batch offsets of every geometry seen, keyed by
// the plan values they follow from (a serving process creates many streams
// of a few geometries; a 2048x2048 level costs ~10^7 Morton ranks to build).
// The per-cell rank LUTs are not kept.
void cached_lod_luts(const Plan& p, LodLuts* out)
{
  std::string key;
  auto put = [&](uint64_t v) { key.append((const char*)&v, sizeof v); };
  put((uint64_t)p.rank);
  put(p.bpe);
  put((uint64_t)p.nlod);
  put((uint64_t)p.lod_ndim);
  put((uint64_t)p.fixed_ndim);
  put(p.fixed_count);
  for (int k = 0; k < HALF_MAX_RANK; ++k) {
    put((uint64_t)p.lod_to_dim[k]);
    put((uint64_t)p.fixed_to_dim[k]);
    put(p.fixed_shape[k]);
  }
  for (int lv = 0; lv < p.nlod; ++lv) {
    const LevelInfo& L = p.levels[lv];
    put(L.lod_nelem);
    put(L.chunk_offset);
    put(L.layout.chunk_stride);
    for (int d = 0; d < p.rank; ++d)
      put(L.shape[d]);
    for (int d = 0; d < 2 * p.rank; ++d) {
      put(L.layout.lifted_shape[d]);
      put((uint64_t)L.layout.lifted_strides[d]);
    }
  }
  static std::mutex mu;
  static std::map<std::string, std::shared_ptr<const LodLuts>> cache;
  static std::deque<std::string> order;
  std::shared_ptr<const LodLuts> hit;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end())
      hit = it->second;
  }
  if (!hit) {
    auto built = std::make_shared<LodLuts>();
    build_lod_luts(p, built.get());
    built->gather_lut = {};
    built->chunk_lut = {};
    std::lock_guard<std::mutex> lk(mu);
    if (!cache.count(key)) {
      cache[key] = built;
      order.push_back(key);
      while (order.size() > 32) {
        cache.erase(order.front());
        order.pop_front();
      }
    }
    hit = built;
  }
  *out = *hit;
}

} // namespace chucky
