// tile_stream_gpu_*: the drop-in streaming writer.
//
// Control flow mirrors the reference's stream (src/gpu/stream.c:65-245 for
// the GPU, src/cpu/stream.c:527-762 for the CPU path whose output is the
// parity target): appends are clipped at epoch boundaries and at the bounded
// capacity; each full epoch is scattered (or LOD-reduced) into its slot of a
// K-epoch chunk pool; every K epochs the batch is compressed, aggregated per
// level and delivered to the shard sink; flush handles the partial epoch, the
// partial batch, the append-dim drain and the final shard indices.
//
// Device work of the plain pipeline runs on one compute queue with a host
// wait after the size scan (as the reference does at
// src/gpu/flush.d2h_deliver.c:103-147); host-memory appends are staged through
// a pinned double buffer on a separate copy queue.
#include "stream_gpu.hpp"

#include <algorithm>
#include <cstring>
#include <new>

using namespace chucky;

namespace {

#define HIPCHECK(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                         \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

size_t gpu_max_output(enum compression_codec c, size_t chunk_bytes)
{
  switch (c) {
    case CODEC_NONE: return chunk_bytes;
    case CODEC_LZ4_NON_STANDARD: return lz4_max_output_size(chunk_bytes);
    // bound over every block size the stream may pick
    case CODEC_ZSTD: return zstd_max_output_size(chunk_bytes, kZstdBlockMin);
    default: return 0;
  }
}

constexpr size_t kSlack = 256; // every device buffer carries read slack for the codecs

int dalloc(tile_stream_gpu* s, void** p, size_t bytes)
{
  *p = nullptr;
  if (bytes == 0)
    return 0;
  HIPCHECK(hipMalloc(p, bytes + kSlack));
  s->allocations.push_back(*p);
  return 0;
}

int halloc(tile_stream_gpu* s, void** p, size_t bytes)
{
  *p = nullptr;
  if (bytes == 0)
    return 0;
  HIPCHECK(host_alloc_near_gpu(p, bytes + kSlack, s->device));
  s->host_allocations.push_back(*p);
  return 0;
}

stream_metric mk_metric(const char* name)
{
  stream_metric m{};
  m.name = name;
  m.best_ms = 1e30f;
  return m;
}

void add_metric(stream_metric& m, float ms, double in_bytes, double out_bytes)
{
  m.ms += ms;
  m.input_bytes += in_bytes;
  m.output_bytes += out_bytes;
  m.count++;
  if (ms < m.best_ms)
    m.best_ms = ms;
}

float event_ms(hipEvent_t a, hipEvent_t b)
{
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess)
    return 0;
  return ms;
}

// Timing events are recycled: creating one costs the host tens of
// microseconds, several times per epoch.
hipEvent_t take_timing(tile_stream_gpu* s)
{
  if (!s->timing_free.empty()) {
    hipEvent_t e = s->timing_free.back();
    s->timing_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess)
    return nullptr;
  return e;
}

void give_timing(tile_stream_gpu* s, hipEvent_t e)
{
  if (e)
    s->timing_free.push_back(e);
}

LiftedSpace space_of(const tile_stream_layout& L)
{
  LiftedSpace sp{};
  sp.rank = L.lifted_rank;
  for (int d = 0; d < L.lifted_rank; ++d) {
    sp.shape[d] = L.lifted_shape[d];
    sp.strides[d] = L.lifted_strides[d];
  }
  return sp;
}

uint8_t* pool_epoch(tile_stream_gpu* s, uint32_t e) { return s->d_pool + (uint64_t)e * s->epoch_pool_bytes; }

uint64_t epoch_elements(const tile_stream_gpu* s) { return s->plan.levels[0].layout.epoch_elements; }

SegmentSource segment_source(const tile_stream_gpu* s)
{
  SegmentSource src{};
  if (s->plan.config.codec.id == CODEC_NONE) {
    src.base = s->d_pool;
    src.chunk_stride = s->chunk_stride_bytes;
    const uint64_t piece = 256 * 1024;
    src.nseg = (uint32_t)std::max<uint64_t>(1, (s->chunk_bytes + piece - 1) / piece);
    src.seg_stride = piece;
    src.seg_sizes = nullptr;
    src.const_size = piece;
    src.const_total = s->chunk_bytes;
  } else {
    src.base = s->d_comp;
    src.chunk_stride = s->chunk_slot;
    src.seg_stride = s->seg_stride;
    src.nseg = s->nseg;
    src.seg_sizes = s->d_seg_sizes;
  }
  return src;
}

// ---- compression of a batch ----

int compress_batch(tile_stream_gpu* s, uint32_t n_epochs)
{
  const uint64_t n = (uint64_t)n_epochs * s->plan.total_chunks;
  const auto codec = s->plan.config.codec;
  if (codec.id == CODEC_NONE || n == 0)
    return 0;
  hipEvent_t e0 = take_timing(s), e1 = take_timing(s);
  if (!e0 || !e1) {
    set_error("event creation failed");
    return 1;
  }
  HIPCHECK(hipEventRecord(e0, s->st_comp));
  if (codec.id == CODEC_LZ4_NON_STANDARD) {
    HIPCHECK(launch_lz4(s->d_pool, s->chunk_stride_bytes, s->chunk_bytes, n, s->d_comp,
                        s->chunk_slot, s->d_seg_sizes, s->lz4_hash_log, s->lz4_seg, s->d_lz4_scratch,
                        s->st_comp));
  } else {
    HIPCHECK(launch_zstd(s->d_pool, s->chunk_stride_bytes, s->chunk_bytes, n, s->d_comp,
                         s->chunk_slot, s->d_seg_sizes, codec.level, s->zblock, s->d_zscratch, s->zworkers,
                         s->d_queue, s->d_ztables, s->st_comp));
  }
  HIPCHECK(hipEventRecord(e1, s->st_comp));
  HIPCHECK(hipEventSynchronize(e1));
  add_metric(s->metrics.compress, event_ms(e0, e1), (double)n * s->chunk_bytes, 0);
  give_timing(s, e0);
  give_timing(s, e1);
  return 0;
}

// ---- aggregate + D2H + delivery of one level ----

int aggregate_level(tile_stream_gpu* s, int lv, uint32_t active, const uint32_t* pool_epochs,
                    bool empty = false)
{
  const LevelInfo& L = s->plan.levels[lv];
  const uint64_t M = L.agg.chunks_per_epoch, C = L.agg.covering_count;
  const uint64_t n_items = (uint64_t)active * M, n_slots = (uint64_t)active * C;
  if (n_items > s->lut_cap || n_slots > s->slot_cap) {
    set_error("aggregate: batch exceeds the allocated LUT capacity");
    return 1;
  }
  uint32_t* hg = s->h_lut;
  uint32_t* hp = s->h_lut + s->lut_cap;
  batch_luts(s->plan, lv, active, pool_epochs, hg, hp);
  HIPCHECK(hipMemcpyAsync(s->d_gather, hg, n_items * 4, hipMemcpyHostToDevice, s->st_comp));
  HIPCHECK(hipMemcpyAsync(s->d_perm, hp, n_items * 4, hipMemcpyHostToDevice, s->st_comp));

  hipEvent_t e0 = take_timing(s), e1 = take_timing(s), e2 = take_timing(s);
  if (!e0 || !e1 || !e2) {
    set_error("event creation failed");
    return 1;
  }
  HIPCHECK(hipEventRecord(e0, s->st_comp));
  SegmentSource src = segment_source(s);
  if (empty) {
    // every chunk of the batch has size 0 (an index of empty entries)
    src = SegmentSource{};
    src.base = s->d_pool;
    src.nseg = 1;
  }
  const uint64_t group = (uint64_t)active * L.agg.cps_inner;
  HIPCHECK(launch_aggregate_plan(src, s->d_gather, s->d_perm, n_items, n_slots, group,
                                 s->plan.config.shard_alignment, s->d_chunk_sizes, s->d_offsets,
                                 s->st_comp));
  uint64_t* h_off = s->h_sizes;
  uint64_t* h_cs = s->h_sizes + (s->slot_cap + 1);
  HIPCHECK(hipMemcpyAsync(h_off, s->d_offsets, (n_slots + 1) * 8, hipMemcpyDeviceToHost, s->st_comp));
  HIPCHECK(hipMemcpyAsync(h_cs, s->d_chunk_sizes, n_slots * 8, hipMemcpyDeviceToHost, s->st_comp));
  HIPCHECK(hipStreamSynchronize(s->st_comp));
  const uint64_t total = h_off[n_slots];
  const uint64_t copy = total + (s->plan.config.shard_alignment ? s->plan.config.shard_alignment : 0);
  if (copy > s->agg_cap) {
    set_error("aggregate: output exceeds the aggregate buffer");
    return 1;
  }
  // Page-aligned streams write whole pages: the bytes between a run's end and
  // its page boundary are written too. The reference leaves them as whatever
  // its malloc'd aggregate buffer held (src/cpu/stream.c:105); here they are
  // zeros.
  if (s->plan.config.shard_alignment)
    HIPCHECK(hipMemsetAsync(s->d_agg, 0, copy, s->st_comp));
  HIPCHECK(launch_aggregate_gather(src, s->d_gather, s->d_perm, n_items, s->d_offsets, s->d_agg,
                                   s->st_comp));
  HIPCHECK(hipEventRecord(e1, s->st_comp));
  // Every level's zero-copy writes read h_agg until their fence retires
  // (write_direct contract), so all levels' fences must pass before the
  // copy overwrites it.
  if (s->sink->wait_fence)
    for (int j = 0; j < s->plan.nlod; ++j)
      s->sink->wait_fence(s->sink, (uint8_t)j, s->io_done[j]);
  HIPCHECK(hipMemcpyAsync(s->h_agg, s->d_agg, copy, hipMemcpyDeviceToHost, s->st_comp));
  HIPCHECK(hipEventRecord(e2, s->st_comp));
  HIPCHECK(hipEventSynchronize(e2));
  add_metric(s->metrics.aggregate, event_ms(e0, e1), (double)total, (double)total);
  add_metric(s->metrics.d2h, event_ms(e1, e2), (double)copy, (double)copy);
  give_timing(s, e0);
  give_timing(s, e1);
  give_timing(s, e2);

  const auto t0 = std::chrono::steady_clock::now();
  AggView v{ s->h_agg, h_off, h_cs };
  size_t out_bytes = 0;
  if (deliver_batch((uint8_t)lv, s->shards[lv], v, active, s->sink, s->plan.config.shard_alignment,
                    &out_bytes))
    return 1;
  const float ms =
    std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  add_metric(s->metrics.sink, ms, (double)out_bytes, 0);
  if (s->sink->record_fence)
    s->io_done[lv] = s->sink->record_fence(s->sink, (uint8_t)lv);
  return 0;
}

// Launches on st_comp are bracketed by events; their durations are collected
// once the queue has passed them.
int time_begin(tile_stream_gpu* s, hipEvent_t* a)
{
  *a = nullptr;
  if (!s->pass_timing)
    return 0;
  *a = take_timing(s);
  if (!*a) {
    set_error("event creation failed");
    return 1;
  }
  HIPCHECK(hipEventRecord(*a, s->st_comp));
  return 0;
}

int time_end(tile_stream_gpu* s, hipEvent_t a, stream_metric stream_metrics::*which, double in_bytes,
             double out_bytes)
{
  if (!a)
    return 0;
  hipEvent_t b = take_timing(s);
  if (!b) {
    set_error("event creation failed");
    return 1;
  }
  HIPCHECK(hipEventRecord(b, s->st_comp));
  s->pending_timing.push_back({ a, b, in_bytes, out_bytes, which });
  return 0;
}

int scatter_begin(tile_stream_gpu* s, hipEvent_t* a) { return time_begin(s, a); }

int scatter_end(tile_stream_gpu* s, hipEvent_t a, double bytes)
{
  return time_end(s, a, &stream_metrics::scatter, bytes, bytes);
}

// Records the timed launches that have completed, keeps the others.
void collect_scatter_ready(tile_stream_gpu* s)
{
  size_t keep = 0;
  for (size_t i = 0; i < s->pending_timing.size(); ++i) {
    auto& t = s->pending_timing[i];
    if (hipEventQuery(t.b) == hipSuccess) {
      add_metric(s->metrics.*t.which, event_ms(t.a, t.b), t.in_bytes, t.out_bytes);
      give_timing(s, t.a);
      give_timing(s, t.b);
    } else {
      s->pending_timing[keep++] = t;
    }
  }
  s->pending_timing.resize(keep);
}

void collect_scatter(tile_stream_gpu* s)
{
  for (auto& t : s->pending_timing) {
    if (hipEventSynchronize(t.b) == hipSuccess)
      add_metric(s->metrics.*t.which, event_ms(t.a, t.b), t.in_bytes, t.out_bytes);
    give_timing(s, t.a);
    give_timing(s, t.b);
  }
  s->pending_timing.clear();
}

// src/cpu/pipeline.c:112-190: compress the batch, then aggregate + deliver
// every level that has active epochs.
int flush_batch(tile_stream_gpu* s, uint32_t n_epochs, const uint32_t* masks)
{
  if (compress_batch(s, n_epochs))
    return 1;
  for (uint32_t e = 0; e < n_epochs; ++e)
    s->slot_compressed[e] = 1;
  collect_scatter(s);
  for (int lv = 0; lv < s->plan.nlod; ++lv) {
    uint32_t pool_epochs[MAX_BATCH_EPOCHS];
    uint32_t active = 0;
    for (uint32_t e = 0; e < n_epochs; ++e)
      if (masks[e] & (1u << lv))
        pool_epochs[active++] = e;
    if (!active)
      continue;
    if (aggregate_level(s, lv, active, pool_epochs))
      return 1;
  }
  // every epoch appended so far has reached the sink once L0 is delivered
  const auto now = std::chrono::steady_clock::now();
  while (s->epochs_written < s->epochs_done) {
    const auto& t0 = s->epoch_t0[s->epochs_written % s->epoch_t0.size()];
    s->latencies.push_back(std::chrono::duration<double>(now - t0).count());
    ++s->epochs_written;
  }
  return 0;
}

// The one-epoch batch that delivers the drained dim-0 accumulators (pool
// slot 0). src/cpu/pipeline.c:159-180 reuses a level's init-time batch LUT
// (pool epoch (a+1)*period-1, src/cpu/pipeline.c:14-31) whenever the active
// count equals the level's batch_active_count = K / period; for this batch
// that is the case exactly when period == K, and the reference then delivers
// pool slot K-1 as last compressed (the level's previous emission), or
// zero-size chunks when no batch has compressed that slot yet.
int flush_drain(tile_stream_gpu* s, uint32_t mask)
{
  if (compress_batch(s, 1))
    return 1;
  s->slot_compressed[0] = 1;
  collect_scatter(s);
  const uint32_t K = s->plan.epochs_per_batch;
  for (int lv = 0; lv < s->plan.nlod; ++lv) {
    if (!(mask & (1u << lv)))
      continue;
    uint32_t pe[1] = { 0 };
    bool empty = false;
    if ((1u << lv) == K) {
      pe[0] = K - 1;
      empty = !s->slot_compressed[K - 1];
    }
    if (aggregate_level(s, lv, 1, pe, empty))
      return 1;
  }
  return 0;
}

// ---- ingest ----

// d_linear is about to take (part of) an epoch. Cells of an epoch that the
// appends do not reach keep the previous epoch's values (the reference reuses
// its epoch buffer); when that epoch was read in place from the caller's
// buffer, put its values back from the Morton buffer first.
int linear_prepare(tile_stream_gpu* s)
{
  if (!s->linear_in_lod)
    return 0;
  if (lod_restore_linear(s)) {
    set_error("lod: restoring the epoch buffer failed");
    return 1;
  }
  s->linear_in_lod = false;
  return 0;
}

int end_epoch(tile_stream_gpu* s);

// Queues the launches of an epoch whose end was deferred (epoch_pending).
int settle_epoch(tile_stream_gpu* s)
{
  if (!s->epoch_pending)
    return 0;
  s->epoch_pending = false;
  return end_epoch(s);
}

// The staged piece, in two steps. dispatch_h2d: the H2D copy into the
// device staging slot. dispatch_rest: the launches of a finished epoch that
// waited for that copy to be queued, then the piece's move into the epoch:
// the chunk scatter, or for multiscale streams a device copy into the epoch
// buffer -- unless the piece is a whole epoch, which the LOD gather then
// reads straight from the staging slot.
hipError_t h2d_copy(tile_stream_gpu* s, int i, size_t at, size_t n)
{
  if (s->h2d_groups)
    return launch_h2d_copy(s->d_stage[i] + at, s->h_stage_dev[i] + at, n, s->h2d_groups, s->st_h2d);
  return hipMemcpyAsync(s->d_stage[i] + at, s->h_stage[i] + at, n, hipMemcpyHostToDevice, s->st_h2d);
}

int dispatch_h2d(tile_stream_gpu* s)
{
  const size_t bytes = s->stage_fill;
  if (bytes == 0)
    return 0;
  const int i = s->stage_cur;
  const size_t sent = s->stage_sent;
  if (sent < bytes)
    HIPCHECK(h2d_copy(s, i, sent, bytes - sent));
  HIPCHECK(hipEventRecord(s->ev_h2d[i], s->st_h2d));
  return 0;
}

int dispatch_rest(tile_stream_gpu* s)
{
  const size_t bytes = s->stage_fill;
  if (settle_epoch(s))
    return 1;
  if (bytes == 0)
    return 0;
  const int i = s->stage_cur;
  const uint64_t n = bytes / s->plan.bpe;
  const uint64_t E = epoch_elements(s);
  HIPCHECK(hipStreamWaitEvent(s->st_comp, s->ev_h2d[i], 0));
  bool held = false;
  if (s->plan.enable_multiscale) {
    const uint64_t off = s->cursor % E;
    if (off == 0 && n == E) {
      // the gather records the slot's release once it has read it
      s->lod_src = s->d_stage[i];
      s->lod_src_free = s->ev_stage_free[i];
      held = true;
    } else {
      if (linear_prepare(s))
        return 1;
      HIPCHECK(hipMemcpyAsync(s->d_linear + off * s->plan.bpe, s->d_stage[i], bytes, hipMemcpyDeviceToDevice,
                              s->st_comp));
    }
  } else {
    hipEvent_t ta = nullptr;
    if (scatter_begin(s, &ta))
      return 1;
    HIPCHECK(launch_transpose(pool_epoch(s, s->batch_acc), s->d_stage[i], n, (int)s->plan.bpe,
                              s->cursor, space_of(s->plan.levels[0].layout), s->st_comp));
    if (scatter_end(s, ta, (double)bytes))
      return 1;
  }
  if (!held)
    HIPCHECK(hipEventRecord(s->ev_stage_free[i], s->st_comp));
  s->cursor += n;
  s->stage_fill = 0;
  s->stage_sent = 0;
  s->stage_cur = (s->stage_cur + 1) % tile_stream_gpu::kStageSlots;
  s->metrics.h2d.input_bytes += (double)bytes;
  s->metrics.h2d.count++;
  return 0;
}

int update_metadata(tile_stream_gpu* s, bool final)
{
  if (!s->sink->update_append)
    return 0;
  const Plan& p = s->plan;
  const int na = p.n_append;
  for (int lv = 0; lv < p.nlod; ++lv) {
    uint64_t sizes[HALF_MAX_RANK] = {};
    for (int d = 1; d < na; ++d)
      sizes[d] = p.dims[d].size;
    if (final) {
      // src/stream/dim_info.h:77-99
      uint64_t inner = 1;
      for (int d = 1; d < p.rank; ++d)
        inner *= p.dims[d].size;
      uint64_t dim0 = inner ? s->cursor / inner : 0;
      if (lv > 0 && p.append_downsample)
        dim0 = (dim0 + ((1ull << lv) - 1)) >> lv;
      sizes[0] = dim0;
    } else {
      // src/stream/dim_info.c:7-22
      const ShardState& ss = s->shards[lv];
      const uint64_t total_ac = ss.shard_epoch * ss.cps_append + ss.epoch_in_shard;
      sizes[0] = ceildiv(total_ac, p.bounded_append_chunks) * p.dims[0].chunk_size;
    }
    if (s->sink->update_append(s->sink, (uint8_t)lv, (uint8_t)na, sizes)) {
      set_error("update_append failed");
      return 1;
    }
  }
  return 0;
}

// Epoch boundary: LOD (multiscale), record the active mask, flush at K.
int end_epoch(tile_stream_gpu* s)
{
  uint32_t mask = 1;
  if (s->plan.enable_multiscale) {
    if (lod_run_epoch(s, s->batch_acc, &mask))
      return 1;
  }
  if (s->batched) {
    s->epochs_done++;
    if (batched_submit(s, s->batch_acc, mask))
      return 1;
    s->batch_acc = (s->batch_acc + 1) % s->plan.epochs_per_batch;
    if (batched_acquire_slot(s, s->batch_acc))
      return 1;
  } else if (s->batch_acc >= MAX_BATCH_EPOCHS) {
    set_error("batch overflow");
    return 1;
  } else {
    s->masks[s->batch_acc++] = mask;
    s->epoch_t0[s->epochs_done % s->epoch_t0.size()] = std::chrono::steady_clock::now();
    s->epochs_done++;
  }
  if (!s->batched && s->batch_acc == s->plan.epochs_per_batch) {
    if (flush_batch(s, s->batch_acc, s->masks))
      return 1;
    s->batch_acc = 0;
  }
  if (s->sink->update_append) {
    const float elapsed = std::chrono::duration<float>(std::chrono::steady_clock::now() - s->meta_clock).count();
    if (elapsed >= s->plan.config.metadata_update_interval_s) {
      s->meta_clock = std::chrono::steady_clock::now();
      if (update_metadata(s, false))
        return 1;
    }
  }
  return 0;
}

int dispatch_stage(tile_stream_gpu* s) { return dispatch_h2d(s) || dispatch_rest(s); }

int flush_all(tile_stream_gpu* s)
{
  if (dispatch_stage(s))
    return 1;
  const uint64_t E = epoch_elements(s);
  if (s->cursor % E != 0) {
    uint32_t mask = 1;
    if (s->plan.enable_multiscale && lod_run_epoch(s, s->batch_acc, &mask))
      return 1;
    if (s->batched) {
      s->epochs_done++;
      if (batched_submit(s, s->batch_acc, mask))
        return 1;
      s->batch_acc = (s->batch_acc + 1) % s->plan.epochs_per_batch;
    } else {
      if (s->batch_acc >= MAX_BATCH_EPOCHS)
        return 1;
      s->masks[s->batch_acc++] = mask;
      s->epoch_t0[s->epochs_done % s->epoch_t0.size()] = std::chrono::steady_clock::now();
      s->epochs_done++;
    }
  }
  if (s->batched) {
    if (batched_drain(s))
      return 1;
    s->batch_acc = 0;
  } else if (s->batch_acc > 0) {
    if (flush_batch(s, s->batch_acc, s->masks))
      return 1;
    s->batch_acc = 0;
  }
  if (s->plan.enable_multiscale && s->plan.append_downsample) {
    uint32_t drain = 0;
    if (lod_drain(s, &drain))
      return 1;
    if (drain && flush_drain(s, drain))
      return 1;
  }
  for (int lv = 0; lv < s->plan.nlod; ++lv) {
    if (s->sink->wait_fence)
      s->sink->wait_fence(s->sink, (uint8_t)lv, s->io_done[lv]);
    if (s->shards[lv].epoch_in_shard > 0 && finalize_shards(s->shards[lv], s->plan.config.shard_alignment))
      return 1;
  }
  return update_metadata(s, true);
}

// ---- writer vtable ----

// Joins and forgets a staging copy begun for input this call did not reach:
// the caller's buffer is only valid during its append.
void prefetch_drop(tile_stream_gpu* s)
{
  if (s->pf_n) {
    s->copy_pool->wait();
    s->pf_n = 0;
  }
}

// Staging copies that nothing ran ahead of go in sub-ranges of this size,
// each H2D-copied as soon as it is staged.
constexpr size_t kStageSub = 2u << 20;

// Starts the host copy of the piece the append loop takes next: the rest of
// the current epoch segment (`seg_left` bytes at `next`), else the start of
// the next segment. Called between a piece's H2D copy and its launches, when
// the cursor has not yet moved past the piece (stage_fill elements pending).
// The next slot was released by an earlier piece's launches, all queued.
void prefetch_next(tile_stream_gpu* s, const uint8_t* next, uint64_t seg_left, const uint8_t* end)
{
  if (!s->copy_pool || s->pf_n)
    return;
  const size_t bpe = s->plan.bpe;
  uint64_t nb = seg_left;
  if (nb == 0) {
    const uint64_t E = epoch_elements(s);
    const uint64_t maxc = s->plan.max_cursor_elements;
    const uint64_t cur = s->cursor + s->stage_fill / bpe;
    uint64_t ne = std::min<uint64_t>(E - cur % E, (uint64_t)(end - next) / bpe);
    if (maxc)
      ne = std::min<uint64_t>(ne, maxc > cur ? maxc - cur : 0);
    nb = ne * bpe;
  }
  nb = std::min<uint64_t>(nb, s->stage_cap);
  const int j = (s->stage_cur + 1) % tile_stream_gpu::kStageSlots;
  if (!nb || hipEventSynchronize(s->ev_stage_free[j]) != hipSuccess)
    return;
  s->copy_pool->start(s->h_stage[j], next, (size_t)nb);
  s->pf_src = next;
  s->pf_n = (size_t)nb;
  s->pf_stage = j;
}

template<bool DEVICE>
writer_result append_impl(tile_stream_gpu* s, const uint8_t* src, const uint8_t* end)
{
  if (s->flushed)
    return writer_finished_at(src, end);
  if (hipSetDevice(s->device) != hipSuccess)
    return writer_error_at(src, end);
  const size_t bpe = s->plan.bpe;
  const uint64_t E = epoch_elements(s);
  const uint64_t maxc = s->plan.max_cursor_elements;
  while (src < end) {
    if (maxc && s->cursor >= maxc) {
      prefetch_drop(s);
      // capacity reached: everything is delivered once and the append says
      // finished; the caller's own flush still succeeds afterwards, as in
      // the reference (src/cpu/stream.c:545-548, 764-771)
      if (!s->at_capacity) {
        const int rc = flush_all(s);
        s->at_capacity = 1;
        if (rc)
          return writer_error_at(src, end);
      }
      return writer_finished_at(src, end);
    }
    const uint64_t in_left = (uint64_t)(end - src) / bpe;
    if (in_left == 0)
      break; // trailing partial element stays unconsumed
    uint64_t elems = std::min<uint64_t>(E - s->cursor % E, in_left);
    if (maxc)
      elems = std::min<uint64_t>(elems, maxc - s->cursor);
    const uint64_t bytes = elems * bpe;
    if (DEVICE) {
      const uint64_t off = s->cursor % E;
      if (s->plan.enable_multiscale) {
        // a whole epoch in one device piece feeds the LOD gather directly
        if (off == 0 && elems == E && ((uintptr_t)src % bpe) == 0)
          s->lod_src = src;
        else if (linear_prepare(s) ||
                 hipMemcpyAsync(s->d_linear + off * bpe, src, bytes, hipMemcpyDeviceToDevice, s->st_comp) != hipSuccess)
          return writer_error_at(src, end);
      } else {
        hipEvent_t ta = nullptr;
        if (scatter_begin(s, &ta))
          return writer_error_at(src, end);
        if (launch_transpose(pool_epoch(s, s->batch_acc), src, elems, (int)bpe, s->cursor,
                             space_of(s->plan.levels[0].layout), s->st_comp) != hipSuccess) {
          set_error("transpose launch failed");
          return writer_error_at(src, end);
        }
        if (scatter_end(s, ta, (double)bytes))
          return writer_error_at(src, end);
      }
      s->cursor += elems;
    } else {
      uint64_t written = 0;
      while (written < bytes) {
        const int i = s->stage_cur;
        const size_t payload = (size_t)std::min<uint64_t>(s->stage_cap - s->stage_fill, bytes - written);
        if (!s->copy_pool)
          s->copy_pool.reset(new (std::nothrow) CopyPool(default_copy_threads()));
        if (!s->copy_pool) {
          set_error("append: staging copy threads could not start");
          return writer_error_at(src, end);
        }
        const auto t0 = std::chrono::steady_clock::now();
        if (s->pf_n && s->pf_src == src + written && s->pf_stage == i && s->stage_fill == 0 &&
            s->pf_n == payload) {
          // begun while the previous piece's launches were queued
          s->copy_pool->wait();
          s->pf_n = 0;
        } else {
          prefetch_drop(s);
          if (s->stage_fill == 0 && hipEventSynchronize(s->ev_stage_free[i]) != hipSuccess)
            return writer_error_at(src, end);
          if (payload >= 2 * kStageSub && s->stage_sent == s->stage_fill) {
            // nothing copied ahead: the H2D copy starts after the first
            // sub-range instead of after the whole piece
            for (size_t o = 0; o < payload; o += kStageSub) {
              const size_t m = std::min(kStageSub, payload - o);
              const size_t at = s->stage_fill + o;
              s->copy_pool->copy(s->h_stage[i] + at, src + written + o, m);
              if (h2d_copy(s, i, at, m) != hipSuccess) {
                set_error("append: H2D copy failed");
                return writer_error_at(src, end);
              }
            }
            s->stage_sent = s->stage_fill + payload;
          } else {
            s->copy_pool->copy(s->h_stage[i] + s->stage_fill, src + written, payload);
          }
        }
        add_metric(s->metrics.memcpy,
                   std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                   (double)payload, (double)payload);
        s->stage_fill += payload;
        written += payload;
        if (s->stage_fill == s->stage_cap || written == bytes) {
          if (dispatch_h2d(s))
            return writer_error_at(src, end);
          // the next piece's host copy runs on the copy threads while this
          // thread queues the launches below
          prefetch_next(s, src + written, bytes - written, end);
          if (dispatch_rest(s))
            return writer_error_at(src, end);
        }
      }
    }
    src += bytes;
    if (s->cursor % E == 0 && s->cursor > 0) {
      // host appends with more input: the epoch's launches follow the next
      // piece's H2D copy (dispatch_stage); otherwise they go now
      const bool more = !DEVICE && (uint64_t)(end - src) >= bpe && !(maxc && s->cursor >= maxc);
      if (more) {
        s->epoch_pending = true;
      } else if (end_epoch(s)) {
        prefetch_drop(s);
        return writer_error_at(src, end);
      }
    }
  }
  prefetch_drop(s);
  if (settle_epoch(s))
    return writer_error_at(src, end);
  return writer_result{ 0, { src, end } };
}

writer_result gpu_append(writer* self, slice data)
{
  tile_stream_gpu* s = reinterpret_cast<tile_stream_gpu*>(self);
  return append_impl<false>(s, (const uint8_t*)data.beg, (const uint8_t*)data.end);
}

writer_result gpu_flush(writer* self)
{
  tile_stream_gpu* s = reinterpret_cast<tile_stream_gpu*>(self);
  if (s->flushed)
    return writer_finished_at(nullptr, nullptr);
  if (hipSetDevice(s->device) != hipSuccess)
    return writer_error();
  const int rc = s->at_capacity ? 0 : flush_all(s);
  s->flushed = 1;
  return rc ? writer_error() : writer_ok();
}

void free_stream(tile_stream_gpu* s)
{
  if (!s)
    return;
  (void)hipSetDevice(s->device);
  batched_free(s);
  if (s->st_comp)
    (void)hipStreamSynchronize(s->st_comp);
  if (s->st_h2d)
    (void)hipStreamSynchronize(s->st_h2d);
  if (s->st_d2h)
    (void)hipStreamSynchronize(s->st_d2h);
  for (auto& t : s->pending_timing) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t e : s->timing_free)
    (void)hipEventDestroy(e);
  for (int i = 0; i < tile_stream_gpu::kStageSlots; ++i) {
    if (s->ev_h2d[i])
      (void)hipEventDestroy(s->ev_h2d[i]);
    if (s->ev_stage_free[i])
      (void)hipEventDestroy(s->ev_stage_free[i]);
  }
  for (void* p : s->allocations)
    (void)hipFree(p);
  for (void* p : s->host_allocations)
    host_release(p);
  if (s->st_comp)
    (void)hipStreamDestroy(s->st_comp);
  if (s->st_h2d)
    (void)hipStreamDestroy(s->st_h2d);
  if (s->st_d2h)
    (void)hipStreamDestroy(s->st_d2h);
  delete s;
}

uint32_t default_workers() { return compressor_slots(); }

// Chunks handed to one compressor launch: one epoch in the batched pipeline,
// a batch of K epochs (at most the stream's length when dim 0 is bounded)
// in the plain one.
uint64_t chunks_per_launch(const tile_stream_gpu* s)
{
  const Plan& p = s->plan;
  if (s->pipeline == CHUCKY_PIPELINE_BATCHED && p.config.shard_alignment == 0)
    return p.total_chunks;
  uint64_t epochs = p.epochs_per_batch;
  const uint64_t E = p.levels[0].layout.epoch_elements;
  if (p.max_cursor_elements && E)
    epochs = std::min<uint64_t>(epochs, (p.max_cursor_elements + E - 1) / E);
  return std::max<uint64_t>(epochs, 1) * p.total_chunks;
}

int create_impl(tile_stream_gpu* s, const tile_stream_configuration* config, shard_sink* sink)
{
  std::string err;
  if (!config || !sink) {
    set_error("create: null configuration or sink");
    return 1;
  }
  if (!codec_is_gpu_supported(config->codec.id)) {
    set_error("create: codec not supported on the GPU");
    return 1;
  }
  if (build_plan(config, 1, gpu_max_output, &s->plan, &err)) {
    set_error("create: " + err);
    return 1;
  }
  const Plan& p = s->plan;
  s->sink = sink;
  s->w.append = gpu_append;
  s->w.flush = gpu_flush;
  if (s->opts.device >= 0)
    HIPCHECK(hipSetDevice(s->opts.device));
  s->pass_timing = s->opts.pass_timing != 2;
  HIPCHECK(hipGetDevice(&s->device));
  HIPCHECK(hipStreamCreateWithFlags(&s->st_h2d, hipStreamNonBlocking));
  HIPCHECK(hipStreamCreateWithFlags(&s->st_comp, hipStreamNonBlocking));

  const uint32_t K = p.epochs_per_batch;
  s->host_gather_wg = s->opts.host_gather_workgroups ? s->opts.host_gather_workgroups
                                                     : tune_host_gather_workgroups();
  s->slot_compressed.assign(K, 0);
  const tile_stream_layout& L0 = p.levels[0].layout;
  s->chunk_stride_bytes = L0.chunk_stride * p.bpe;
  s->chunk_bytes = L0.chunk_elements * p.bpe;
  s->epoch_pool_bytes = p.total_chunks * s->chunk_stride_bytes;


  // staging
  s->stage_cap = std::max<size_t>(align_up(config->buffer_capacity_bytes, 4096), 4096);
  s->stage_cap -= s->stage_cap % p.bpe;
  for (int i = 0; i < tile_stream_gpu::kStageSlots; ++i) {
    if (halloc(s, (void**)&s->h_stage[i], s->stage_cap))
      return 1;
    if (dalloc(s, (void**)&s->d_stage[i], s->stage_cap))
      return 1;
    HIPCHECK(hipEventCreateWithFlags(&s->ev_h2d[i], hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&s->ev_stage_free[i], hipEventDisableTiming));
    HIPCHECK(hipEventRecord(s->ev_stage_free[i], s->st_comp));
  }
  // The H2D copy of staged pieces: the DMA engines, or on request a kernel
  // reading the pinned buffers through their device addresses (launches that
  // wait for it are then queued on the device instead of holding the host
  // thread until a DMA copy ends; in the pipelines it runs slower, sharing
  // the CUs and the fabric with the compressors)
  if (s->opts.h2d_copy > 2) {
    set_error("create: h2d_copy must be 0, 1 or 2");
    return 1;
  }
  if (s->opts.h2d_copy == 1) {
    for (int i = 0; i < tile_stream_gpu::kStageSlots; ++i) {
      void* dp = nullptr;
      if (hipHostGetDevicePointer(&dp, s->h_stage[i], 0) != hipSuccess || !dp ||
          !h2d_copy_ok(s->d_stage[i], dp)) {
        (void)hipGetLastError();
        set_error("create: h2d_copy = 1 but the pinned staging buffers have no device address");
        return 1;
      }
      s->h_stage_dev[i] = (uint8_t*)dp;
    }
    s->h2d_groups = tune_h2d_copy_workgroups();
  }

  // pool (zeroed once; padding positions stay zero for the stream's life)
  if (dalloc(s, (void**)&s->d_pool, (uint64_t)K * s->epoch_pool_bytes))
    return 1;
  HIPCHECK(hipMemsetAsync(s->d_pool, 0, (uint64_t)K * s->epoch_pool_bytes + kSlack, s->st_comp));

  // codec
  const uint64_t nchunks = (uint64_t)K * p.total_chunks;
  if (config->codec.id == CODEC_LZ4_NON_STANDARD) {
    s->nseg = 1;
    s->chunk_slot = align_up(lz4_max_output_size(s->chunk_bytes), 256);
    s->seg_stride = s->chunk_slot;
    if (s->opts.lz4_hash_log && (s->opts.lz4_hash_log < 9 || s->opts.lz4_hash_log > 13)) {
      set_error("create: lz4_hash_log must be in [9, 13]");
      return 1;
    }
    const uint64_t per_launch = chunks_per_launch(s);
    const Lz4Choice tuned = tune_lz4(s->chunk_bytes, per_launch, (uint32_t)s->plan.bpe, s->plan.nlod > 1);
    s->lz4_hash_log = s->opts.lz4_hash_log ? (int)s->opts.lz4_hash_log : tuned.hash_log;
    if (s->opts.lz4_segment_bytes) {
      const uint32_t sb = s->opts.lz4_segment_bytes;
      if (sb < s->chunk_bytes && (sb < 1024 || (sb & (sb - 1)))) {
        set_error("create: lz4_segment_bytes must be a power of two >= 1024, or >= the chunk bytes");
        return 1;
      }
      s->lz4_seg = sb < s->chunk_bytes ? sb : 0;
    } else {
      s->lz4_seg = tuned.seg_bytes;
    }
    if (s->lz4_seg && dalloc(s, (void**)&s->d_lz4_scratch,
                             lz4_scratch_bytes(s->chunk_bytes, per_launch, s->lz4_seg)))
      return 1;
  } else if (config->codec.id == CODEC_ZSTD) {
    HIPCHECK(zstd_prepare());
    if (s->opts.zstd_block_bytes && !zstd_block_ok(s->opts.zstd_block_bytes)) {
      set_error("create: zstd_block_bytes must be a power of two in [8 KiB, 128 KiB]");
      return 1;
    }
    s->zblock = s->opts.zstd_block_bytes ? s->opts.zstd_block_bytes
                                         : tune_zstd_block(s->chunk_bytes, chunks_per_launch(s),
                                                           (uint32_t)s->plan.bpe, s->plan.nlod > 1);
    const uint32_t ct = s->opts.zstd_chunk_tables;
    const bool chunk_tables =
      ct == 1 || (ct == 0 && tune_zstd_chunk_tables(s->chunk_bytes, chunks_per_launch(s), (uint32_t)s->plan.bpe,
                                                    s->plan.nlod > 1));
    if (chunk_tables && dalloc(s, (void**)&s->d_ztables, zstd_table_bytes(chunks_per_launch(s))))
      return 1;
    s->nseg = zstd_segments(s->chunk_bytes, s->zblock);
    s->seg_stride = zstd_segment_stride(s->zblock);
    s->chunk_slot = zstd_max_output_size(s->chunk_bytes, s->zblock);
    s->zworkers = s->opts.compress_workers ? s->opts.compress_workers : default_workers();
    const uint64_t blocks = nchunks * s->nseg;
    if (s->zworkers > blocks)
      s->zworkers = (uint32_t)blocks;
    if (dalloc(s, &s->d_zscratch, zstd_scratch_bytes(s->zworkers)))
      return 1;
    if (dalloc(s, (void**)&s->d_queue, 64))
      return 1;
    // the compressor's work counter starts at zero and each launch leaves it so
    HIPCHECK(hipMemsetAsync(s->d_queue, 0, 64, s->st_comp));
  }
  if (config->codec.id != CODEC_NONE) {
    if (dalloc(s, (void**)&s->d_comp, nchunks * s->chunk_slot))
      return 1;
    if (dalloc(s, (void**)&s->d_seg_sizes, nchunks * s->nseg * 4))
      return 1;
  }

  // aggregate buffers sized for the largest level batch
  for (int lv = 0; lv < p.nlod; ++lv) {
    const LevelInfo& L = p.levels[lv];
    const uint64_t slots = std::max<uint32_t>(L.batch_active_count, 1);
    const uint64_t kk = std::max<uint64_t>(slots, K); // partial batches recompute with <= K
    s->lut_cap = std::max<uint64_t>(s->lut_cap, kk * L.agg.chunks_per_epoch);
    s->slot_cap = std::max<uint64_t>(s->slot_cap, kk * L.agg.covering_count);
    const uint64_t pad_total = config->shard_alignment
                                 ? (kk * L.agg.covering_count / std::max<uint64_t>(L.agg.cps_inner, 1) + 2) *
                                     config->shard_alignment
                                 : 0;
    s->agg_cap = std::max<uint64_t>(s->agg_cap, kk * L.agg.chunks_per_epoch * p.max_output_size + pad_total);
  }
  if (dalloc(s, (void**)&s->d_gather, s->lut_cap * 4) || dalloc(s, (void**)&s->d_perm, s->lut_cap * 4) ||
      halloc(s, (void**)&s->h_lut, s->lut_cap * 8) ||
      dalloc(s, (void**)&s->d_chunk_sizes, s->slot_cap * 8) ||
      dalloc(s, (void**)&s->d_offsets, (s->slot_cap + 1) * 8) ||
      halloc(s, (void**)&s->h_sizes, (2 * s->slot_cap + 1) * 8) ||
      dalloc(s, (void**)&s->d_agg, s->agg_cap) || halloc(s, (void**)&s->h_agg, s->agg_cap))
    return 1;

  // shards
  s->shards.resize(p.nlod);
  for (int lv = 0; lv < p.nlod; ++lv)
    s->shards[lv].init(p.levels[lv]);
  s->io_done.assign(p.nlod, io_event{ 0 });

  // LOD
  if (p.enable_multiscale && lod_init(s))
    return 1;

  // device copy of the L0 lifted layout
  if (dalloc(s, (void**)&s->layout_gpu.d_lifted_shape, MAX_RANK * 8) ||
      dalloc(s, (void**)&s->layout_gpu.d_lifted_strides, MAX_RANK * 8))
    return 1;
  HIPCHECK(hipMemcpyAsync(s->layout_gpu.d_lifted_shape, L0.lifted_shape, MAX_RANK * 8,
                          hipMemcpyHostToDevice, s->st_comp));
  HIPCHECK(hipMemcpyAsync(s->layout_gpu.d_lifted_strides, L0.lifted_strides, MAX_RANK * 8,
                          hipMemcpyHostToDevice, s->st_comp));
  HIPCHECK(hipStreamSynchronize(s->st_comp));

  s->metrics.memcpy = mk_metric("memcpy");
  s->metrics.h2d = mk_metric("h2d");
  s->metrics.scatter = mk_metric("scatter");
  s->metrics.compress = mk_metric("compress");
  s->metrics.aggregate = mk_metric("aggregate");
  s->metrics.d2h = mk_metric("d2h");
  s->metrics.sink = mk_metric("sink");
  s->metrics.lod_gather = mk_metric("lod_gather");
  s->metrics.lod_reduce = mk_metric("lod_reduce");
  s->metrics.lod_append_fold = mk_metric("lod_append_fold");
  s->metrics.lod_morton_chunk = mk_metric("lod_morton");
  s->meta_clock = std::chrono::steady_clock::now();
  s->epoch_t0.resize(MAX_BATCH_EPOCHS * 2);
  // The batched pipeline delivers per epoch, which equals per-batch delivery
  // only without shard page alignment; aligned streams stay on the plain path.
  if (s->pipeline == CHUCKY_PIPELINE_BATCHED && config->shard_alignment == 0 && batched_init(s))
    return 1;
  // created last: a copy-only stream, while the streams that launch kernels
  // (compute, then the batched pipeline's) take HIP's hardware queues in
  // creation order
  HIPCHECK(hipStreamCreateWithFlags(&s->st_d2h, hipStreamNonBlocking));
  HIPCHECK(hipStreamSynchronize(s->st_comp));
  return 0;
}

} // namespace

namespace chucky {
std::recursive_mutex& graph_mutex()
{
  static std::recursive_mutex m;
  return m;
}
void metric_add(stream_metric& m, float ms, double in_bytes, double out_bytes)
{
  add_metric(m, ms, in_bytes, out_bytes);
}
void scatter_collect_ready(tile_stream_gpu* s) { collect_scatter_ready(s); }
int timing_begin(tile_stream_gpu* s, hipEvent_t* a) { return time_begin(s, a); }
int timing_pair(tile_stream_gpu* s, hipEvent_t* a, hipEvent_t* b)
{
  *a = *b = nullptr;
  if (!s->pass_timing)
    return 0;
  *a = take_timing(s);
  *b = take_timing(s);
  if (!*a || !*b) {
    give_timing(s, *a);
    give_timing(s, *b);
    *a = *b = nullptr;
    set_error("event creation failed");
    return 1;
  }
  return 0;
}
void timing_add(tile_stream_gpu* s, hipEvent_t a, hipEvent_t b, stream_metric stream_metrics::*which,
                double in_bytes, double out_bytes)
{
  if (!a || !b)
    return;
  s->pending_timing.push_back({ a, b, in_bytes, out_bytes, which });
}
int timing_end(tile_stream_gpu* s, hipEvent_t a, stream_metric stream_metrics::*which, double in_bytes,
               double out_bytes)
{
  return time_end(s, a, which, in_bytes, out_bytes);
}
} // namespace chucky

extern "C" {

void tile_stream_gpu_options_default(struct tile_stream_gpu_options* o)
{
  *o = tile_stream_gpu_options{};
  o->pipeline = CHUCKY_PIPELINE_PLAIN;
  o->device = -1;
}

struct tile_stream_gpu* tile_stream_gpu_create_ex(const struct tile_stream_configuration* config,
                                                  struct shard_sink* sink,
                                                  const struct tile_stream_gpu_options* opts)
{
  std::lock_guard<std::recursive_mutex> lk(graph_mutex());
  tile_stream_gpu* s = new (std::nothrow) tile_stream_gpu();
  if (!s)
    return nullptr;
  if (opts)
    s->opts = *opts;
  else
    tile_stream_gpu_options_default(&s->opts);
  s->pipeline = s->opts.pipeline;
  if (create_impl(s, config, sink)) {
    free_stream(s);
    return nullptr;
  }
  return s;
}

struct tile_stream_gpu* tile_stream_gpu_create(const struct tile_stream_configuration* config,
                                               struct shard_sink* sink)
{
  return tile_stream_gpu_create_ex(config, sink, nullptr);
}

void tile_stream_gpu_destroy(struct tile_stream_gpu* stream)
{
  std::lock_guard<std::recursive_mutex> lk(graph_mutex());
  free_stream(stream);
}

struct stream_metrics tile_stream_gpu_get_metrics(const struct tile_stream_gpu* s) { return s->metrics; }

const struct tile_stream_layout* tile_stream_gpu_layout(const struct tile_stream_gpu* s)
{
  return &s->plan.levels[0].layout;
}

struct writer* tile_stream_gpu_writer(struct tile_stream_gpu* s) { return &s->w; }

uint64_t tile_stream_gpu_cursor(const struct tile_stream_gpu* s) { return s->cursor; }

struct tile_stream_status tile_stream_gpu_status(const struct tile_stream_gpu* s)
{
  tile_stream_status st{};
  st.nlod = s->plan.nlod;
  st.append_downsample = s->plan.append_downsample;
  st.epochs_per_batch = s->plan.epochs_per_batch;
  st.max_compressed_size = s->plan.max_output_size;
  st.dtype = s->plan.config.dtype;
  st.codec = s->plan.config.codec;
  st.codec_batch_size = (size_t)s->plan.epochs_per_batch * s->plan.total_chunks;
  st.batch_accumulated = s->batch_acc;
  st.pool_current = 0;
  st.flush_pending = 0;
  return st;
}

struct writer_result tile_stream_gpu_append_device(struct tile_stream_gpu* s, const void* d_src,
                                                   size_t nbytes)
{
  const uint8_t* b = (const uint8_t*)d_src;
  return append_impl<true>(s, b, b + nbytes);
}

int tile_stream_gpu_reset(struct tile_stream_gpu* s, struct shard_sink* sink)
{
  if (hipSetDevice(s->device) != hipSuccess || hipStreamSynchronize(s->st_comp) != hipSuccess)
    return 1;
  // writes of the previous acquisition out of the stream's pinned buffers
  // end before those buffers are reused or the sink is replaced (its fences
  // count in its own sequence)
  if (s->sink && s->sink->wait_fence)
    for (int lv = 0; lv < s->plan.nlod; ++lv)
      s->sink->wait_fence(s->sink, (uint8_t)lv, s->io_done[lv]);
  batched_reset(s);
  if (sink)
    s->sink = sink;
  s->cursor = 0;
  s->batch_acc = 0;
  s->flushed = 0;
  s->at_capacity = 0;
  std::fill(s->slot_compressed.begin(), s->slot_compressed.end(), (uint8_t)0);
  s->stage_fill = 0;
  s->stage_sent = 0;
  s->epoch_pending = false;
  s->lod_src = nullptr;
  s->lod_src_free = nullptr;
  std::fill(std::begin(s->append_counts), std::end(s->append_counts), 0u);
  for (auto& ss : s->shards)
    ss.reset();
  std::fill(s->io_done.begin(), s->io_done.end(), io_event{ 0 });
  s->latencies.clear();
  s->epochs_done = s->epochs_written = 0;
  return 0;
}

int tile_stream_gpu_level_geometry(const struct tile_stream_gpu* s, int level,
                                   uint64_t* cps_total, uint64_t* shard_inner_count,
                                   uint64_t* chunks_per_epoch)
{
  if (!s || level < 0 || level >= s->plan.nlod)
    return 1;
  const LevelInfo& L = s->plan.levels[level];
  if (cps_total)
    *cps_total = L.cps_total;
  if (shard_inner_count)
    *shard_inner_count = L.shard_inner_count;
  if (chunks_per_epoch)
    *chunks_per_epoch = L.chunk_count;
  return 0;
}

size_t tile_stream_gpu_epoch_latencies(const struct tile_stream_gpu* s, double* out, size_t cap)
{
  const size_t n = std::min(cap, s->latencies.size());
  for (size_t i = 0; i < n; ++i)
    out[i] = s->latencies[i];
  return n;
}

namespace {
int estimate_core(const struct tile_stream_configuration* config, struct tile_stream_memory_info* info,
                  uint32_t lz4_segment_override)
{
  if (!info || !config)
    return 1;
  *info = tile_stream_memory_info{};
  Plan p;
  std::string err;
  if (build_plan(config, 1, gpu_max_output, &p, &err))
    return 1;
  const uint32_t K = p.epochs_per_batch;
  const uint64_t csb = p.levels[0].layout.chunk_stride * p.bpe;
  const uint64_t cb = p.levels[0].layout.chunk_elements * p.bpe;
  const uint64_t nchunks = (uint64_t)K * p.total_chunks;
  info->staging_bytes = tile_stream_gpu::kStageSlots * align_up(config->buffer_capacity_bytes, 4096);
  info->chunk_pool_bytes = nchunks * csb;
  info->compressed_pool_bytes = config->codec.id == CODEC_NONE ? 0 : nchunks * p.max_output_size;
  uint64_t agg = 0;
  for (int lv = 0; lv < p.nlod; ++lv) {
    const LevelInfo& L = p.levels[lv];
    const uint64_t kk = std::max<uint64_t>(std::max<uint32_t>(L.batch_active_count, 1), K);
    agg = std::max<uint64_t>(agg, kk * L.agg.chunks_per_epoch * p.max_output_size);
  }
  info->aggregate_bytes = agg;
  uint64_t lod = 0;
  if (p.enable_multiscale) {
    lod += p.levels[0].layout.epoch_elements * p.bpe;
    lod += p.level_span_ends[p.nlod - 1] * p.bpe;
    lod += p.levels[0].lod_nelem * 8; // level-0 gather lists
    lod += p.append_accum_elements() * p.bpe;
    lod += p.ends.size() * 8;
    for (int lv = 0; lv < p.nlod; ++lv)
      lod += p.levels[lv].lod_nelem * 8 + p.fixed_count * 8; // tile/edge lists, at most 8 B a cell
  }
  info->lod_bytes = lod;
  if (config->codec.id == CODEC_ZSTD) {
    const uint64_t blocks = nchunks * zstd_segments(cb, kZstdBlockMin);
    info->codec_bytes = zstd_scratch_bytes((uint32_t)std::min<uint64_t>(blocks, compressor_slots())) +
                        zstd_table_bytes(nchunks);
  }
  info->codec_bytes += nchunks * (config->codec.id == CODEC_ZSTD ? zstd_segments(cb, kZstdBlockMin) : 1) * 4;
  if (config->codec.id == CODEC_LZ4_NON_STANDARD) {
    // segment scratch of the multi-wavefront LZ4 path (plain pipeline bound)
    uint64_t epochs = K;
    const uint64_t E = p.levels[0].layout.epoch_elements;
    if (p.max_cursor_elements && E)
      epochs = std::min<uint64_t>(epochs, (p.max_cursor_elements + E - 1) / E);
    const uint64_t per_launch = std::max<uint64_t>(epochs, 1) * p.total_chunks;
    const uint32_t seg = lz4_segment_override
                           ? (lz4_segment_override < cb ? lz4_segment_override : 0)
                           : tune_lz4(cb, per_launch, (uint32_t)p.bpe, p.nlod > 1).seg_bytes;
    if (seg)
      info->codec_bytes += lz4_scratch_bytes(cb, per_launch, seg);
  }
  uint64_t shard = 0;
  for (int lv = 0; lv < p.nlod; ++lv)
    shard += p.levels[lv].shard_inner_count * (sizeof(OpenShard) + p.levels[lv].cps_total * 16);
  info->shard_bytes = shard;
  info->device_bytes = info->staging_bytes + info->chunk_pool_bytes + info->compressed_pool_bytes +
                       info->aggregate_bytes + info->lod_bytes + info->codec_bytes;
  info->host_pinned_bytes = info->staging_bytes + info->aggregate_bytes;
  info->chunks_per_epoch = p.levels[0].layout.chunks_per_epoch;
  info->total_chunks = p.total_chunks;
  info->max_output_size = p.max_output_size;
  info->nlod = p.nlod;
  info->epochs_per_batch = K;
  return 0;
}

} // namespace

int tile_stream_gpu_memory_estimate(const struct tile_stream_configuration* config,
                                    struct tile_stream_memory_info* info)
{
  return estimate_core(config, info, 0);
}

int tile_stream_gpu_memory_estimate_ex(const struct tile_stream_configuration* config,
                                       const struct tile_stream_gpu_options* opts,
                                       struct tile_stream_memory_info* info)
{
  if (estimate_core(config, info, opts ? opts->lz4_segment_bytes : 0))
    return 1;
  // the batched pipeline (which aligned streams do not use) adds, per host
  // slot, one epoch's worst-case bytes of every level in pinned memory and
  // in device staging, and per pool slot its aggregate LUTs and sizes
  if (!opts || opts->pipeline != CHUCKY_PIPELINE_BATCHED || config->shard_alignment)
    return 0;
  Plan p;
  std::string err;
  if (build_plan(config, 1, gpu_max_output, &p, &err))
    return 1;
  const uint32_t K = p.epochs_per_batch;
  const uint64_t R = std::min<uint32_t>(K, 4);
  const uint64_t cb = p.levels[0].layout.chunk_elements * p.bpe;
  const uint64_t per = config->codec.id == CODEC_NONE ? cb : p.max_output_size;
  uint64_t epoch_bytes = 0, luts = 0, sizes = 0;
  for (int lv = 0; lv < p.nlod; ++lv) {
    const LevelInfo& L = p.levels[lv];
    epoch_bytes += align_up(L.agg.chunks_per_epoch * per, 256);
    luts += L.agg.chunks_per_epoch * 8 + (2 * L.agg.covering_count + 1) * 8;
    sizes += (2 * L.agg.covering_count + 1) * 8;
  }
  const bool dma = opts->batched_d2h != 1;
  const uint64_t device_extra = (dma ? R * epoch_bytes : 0) + (uint64_t)K * luts;
  const uint64_t host_extra = R * (epoch_bytes + sizes);
  info->aggregate_bytes += device_extra;
  info->device_bytes += device_extra;
  info->host_pinned_bytes += host_extra;
  return 0;
}

int tile_stream_gpu_advise_chunk_sizes(struct tile_stream_configuration* config,
                                       size_t target_chunk_bytes, const uint8_t* ratios,
                                       size_t budget_bytes)
{
  const size_t bpe = dtype_bpe(config->dtype);
  if (bpe == 0 || budget_bytes == 0)
    return 1;
  for (size_t target = target_chunk_bytes; target >= bpe; target >>= 1) {
    dims_budget_chunk_bytes(config->dimensions, config->rank, target, bpe, ratios);
    tile_stream_memory_info mem;
    if (tile_stream_gpu_memory_estimate(config, &mem))
      return 1;
    if (mem.device_bytes <= budget_bytes)
      return 0;
  }
  return 1;
}

} // extern "C"
