// Batched pipeline: every epoch is compressed, aggregated and handed to the
// sink on its own, as soon as it is complete, with no host wait on the device
// path.
//
//   compute queue:  scatter/LOD(e) -> [graph: compress all chunks of epoch e]
//   aggregate queue: wait(compress e) -> [graph: per active level, size plan +
//                    exclusive scan, offsets to pinned memory, gather of the
//                    compressed bytes into a device staging slot]
//   copy queue:      one DMA copy per active level of exactly the gathered
//                    bytes into a pinned host slot, issued by the host as soon
//                    as the epoch's offsets are known (the copy engines move
//                    the data at link speed without taking CUs from the
//                    compressor; a gather kernel writing host memory directly
//                    is kept as an option)
//   host:           delivers epochs in order once their copy has completed
//                   (sink writes, index entries, shard finalisation)
//
// The compress graph of epoch e+1 runs while the gather of epoch e streams
// over PCIe. Epochs rotate through the K pool slots of the plan and through R
// pinned host slots; a slot is reused only after the work that read it has
// completed (events on the device side, delivery + sink fences on the host).
//
// Output equals the per-K-batch delivery of the plain pipeline byte for byte
// when shard_alignment == 0 (src/zarr/shard_delivery.c:104-204 appends each
// shard's runs in epoch order either way); with page alignment the padding
// depends on the batch grouping, so such streams use the plain pipeline.
#include "stream_gpu.hpp"

#include <algorithm>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <thread>

namespace chucky {

namespace {

#define BCHECK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      set_error(std::string("batched: ") + #expr + ": " + hipGetErrorString(e_));           \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

} // namespace

// compression queues at most: with the compute and copy streams they still
// fit HIP's 4 hardware queues
constexpr uint32_t kMaxZip = 2;

struct BatchedState
{
  hipStream_t st_agg = nullptr;
  // DMA copies of host slot h go on st_copy[h], and nothing else does: a
  // copy-only stream carries no event markers, which on a hardware queue
  // shared with a compression stream would wait behind its kernels; a
  // copy's completion is its stream's (hipStreamQuery / Synchronize)
  std::vector<hipStream_t> st_copy;
  // Compression queues: when one epoch's blocks fill only part of the
  // compressor's slots, successive epochs compress side by side on NZ queues,
  // each with its own work counter and share of the worker scratch.
  uint32_t NZ = 1;
  uint32_t zip_workers = 0;
  // per-epoch compress and aggregate as captured graphs (else direct launches)
  bool graphs = false;
  std::vector<hipStream_t> st_zip;
  uint32_t* d_queues = nullptr; // NZ work counters, 64 bytes apart
  // zstd split epochs: per compression queue, the sizes of the part-segment
  // blocks the coarse levels are compressed in (kZstdMaxParts u32 per segment)
  uint32_t* d_sub_sizes = nullptr;
  uint64_t sub_stride = 0;
  // LZ4 segment scratch, one slice of lz4_stride bytes per compression queue
  uint8_t* d_lz4 = nullptr;
  uint64_t lz4_stride = 0;
  bool dma = true; // gathered bytes reach host memory by DMA copies
  uint32_t R = 0; // host slots
  struct HostSlot
  {
    uint8_t* h_agg = nullptr;
    uint8_t* d_stage = nullptr; // device staging of the gathered bytes (DMA mode)
    std::vector<uint64_t> level_base; // byte offset of each level's bytes
    std::vector<uint64_t*> h_offsets, h_sizes;
    std::vector<io_event> fence;
  };
  struct PoolSlot
  {
    std::vector<uint32_t*> d_gather, d_perm;
    std::vector<uint64_t*> d_sizes, d_offsets;
    hipGraphExec_t compress_exec = nullptr;
    hipEvent_t ev_filled = nullptr;     // scatter / LOD of the slot's epoch done
    hipEvent_t ev_compressed = nullptr; // compress of the slot's epoch done
    hipEvent_t ev_gathered = nullptr;   // gather of the slot's epoch done
  };
  std::vector<HostSlot> host;
  std::vector<PoolSlot> pool;
  // aggregate graphs by (pool slot, host slot, mask)
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, hipGraphExec_t> agg_exec;
  struct Ticket
  {
    uint64_t epoch;
    uint32_t pool_slot, host_slot, mask;
    hipEvent_t done;   // aggregate graph finished
    bool copied;       // DMA copies queued (DMA mode) on the host slot's copy stream
    hipEvent_t c0, c1; // compress launch start / end (timing events)
    std::chrono::steady_clock::time_point t0;
    double bytes01 = 0; // compressor input bytes between c0 and c1
    // split epochs: the second compress launch (level 0) and its input bytes
    hipEvent_t c2 = nullptr, c3 = nullptr;
    double bytes23 = 0;
    // split epochs: level 0 (mask_a) is compressed, gathered and copied on
    // its own, beside the coarse levels (done / copied)
    uint32_t mask_a = 0;
    hipEvent_t done_a = nullptr;
    bool copied_a = false;
  };
  std::deque<Ticket> inflight;
  uint64_t next_epoch = 0;
  std::vector<hipEvent_t> event_pool, timing_pool;
  // DMA mode: a copier thread waits for each gather in epoch order and
  // queues its copy at once, whatever the submitting thread is doing. The
  // mutex guards inflight, the event pools and the tickets' copy events.
  std::recursive_mutex mu;
  std::condition_variable_any cv; // a ticket was added, or stop
  std::thread copier;
  bool stop = false;
  std::string copy_error; // first failure of the copier thread
};

namespace {

SegmentSource source_of(const tile_stream_gpu* s)
{
  SegmentSource src{};
  if (s->plan.config.codec.id == CODEC_NONE) {
    const uint64_t piece = 256 * 1024;
    src.base = s->d_pool;
    src.chunk_stride = s->chunk_stride_bytes;
    src.nseg = (uint32_t)std::max<uint64_t>(1, (s->chunk_bytes + piece - 1) / piece);
    src.seg_stride = piece;
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

template<typename T>
int dev_alloc(tile_stream_gpu* s, T** p, size_t count)
{
  void* q = nullptr;
  BCHECK(hipMalloc(&q, count * sizeof(T) + 256));
  s->allocations.push_back(q);
  *p = (T*)q;
  return 0;
}

template<typename T>
int host_alloc(tile_stream_gpu* s, T** p, size_t count)
{
  void* q = nullptr;
  BCHECK(host_alloc_near_gpu(&q, count * sizeof(T) + 256, s->device));
  s->host_allocations.push_back(q);
  *p = (T*)q;
  return 0;
}

hipEvent_t take_event(BatchedState* b)
{
  std::lock_guard<std::recursive_mutex> lk(b->mu);
  if (!b->event_pool.empty()) {
    hipEvent_t e = b->event_pool.back();
    b->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
    return nullptr;
  return e;
}

hipEvent_t take_timing_event(BatchedState* b)
{
  std::lock_guard<std::recursive_mutex> lk(b->mu);
  if (!b->timing_pool.empty()) {
    hipEvent_t e = b->timing_pool.back();
    b->timing_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess)
    return nullptr;
  return e;
}

// Chunks [0, n) of the epoch in pool slot `slot` (the captured graphs take
// every chunk of the slot, n = total_chunks).
int launch_compress(tile_stream_gpu* s, uint32_t slot, hipStream_t st, uint64_t n)
{
  const BatchedState* b = s->batched;
  const uint32_t z = slot % b->NZ;
  const auto codec = s->plan.config.codec;
  const uint64_t T = s->plan.total_chunks;
  if (codec.id == CODEC_LZ4_NON_STANDARD)
    BCHECK(launch_lz4(s->d_pool + (uint64_t)slot * s->epoch_pool_bytes, s->chunk_stride_bytes,
                      s->chunk_bytes, n, s->d_comp + (uint64_t)slot * T * s->chunk_slot, s->chunk_slot,
                      s->d_seg_sizes + (uint64_t)slot * T * s->nseg, s->lz4_hash_log, s->lz4_seg,
                      b->d_lz4 ? b->d_lz4 + z * b->lz4_stride : s->d_lz4_scratch, st));
  else if (codec.id == CODEC_ZSTD)
    BCHECK(launch_zstd(s->d_pool + (uint64_t)slot * s->epoch_pool_bytes, s->chunk_stride_bytes,
                       s->chunk_bytes, n, s->d_comp + (uint64_t)slot * T * s->chunk_slot, s->chunk_slot,
                       s->d_seg_sizes + (uint64_t)slot * T * s->nseg, codec.level, s->zblock,
                       (uint8_t*)s->d_zscratch + zstd_scratch_bytes(z * b->zip_workers), b->zip_workers,
                       b->d_queues + 16 * z, s->d_ztables, st));
  return 0;
}

// End of the chunks of the levels in `mask` in an epoch's pool slot: levels
// are stored one after another, level 0 first.
uint64_t active_chunk_end(const Plan& p, uint32_t mask)
{
  int top = 0;
  for (int lv = 0; lv < p.nlod; ++lv)
    if (mask & (1u << lv))
      top = lv;
  return top + 1 < p.nlod ? p.levels[top + 1].chunk_offset : p.total_chunks;
}

// zstd on chunks [c0, c1) of pool slot `slot`, with compression queue z's
// work counter and scratch share, on stream st.
int launch_zstd_range(tile_stream_gpu* s, uint32_t slot, uint32_t z, uint64_t c0, uint64_t c1, hipStream_t st)
{
  const BatchedState* b = s->batched;
  const uint64_t T = s->plan.total_chunks;
  if (c1 <= c0)
    return 0;
  BCHECK(launch_zstd(s->d_pool + (uint64_t)slot * s->epoch_pool_bytes + c0 * s->chunk_stride_bytes,
                     s->chunk_stride_bytes, s->chunk_bytes, c1 - c0,
                     s->d_comp + ((uint64_t)slot * T + c0) * s->chunk_slot, s->chunk_slot,
                     s->d_seg_sizes + ((uint64_t)slot * T + c0) * s->nseg, s->plan.config.codec.level, s->zblock,
                     (uint8_t*)s->d_zscratch + zstd_scratch_bytes(z * b->zip_workers), b->zip_workers,
                     b->d_queues + 16 * z, nullptr, st));
  return 0;
}

// zstd on the coarse chunks [c0, c1) of a split epoch: up to four wavefronts
// per block-sized segment when the layout allows (these launches hold few
// chunks, and coarse blocks are the slowest), else as launch_zstd_range.
// `done` is recorded once the compression kernel has run (before the join).
int launch_zstd_coarse(tile_stream_gpu* s, uint32_t slot, uint32_t z, uint64_t c0, uint64_t c1, hipStream_t st,
                       hipEvent_t done)
{
  const BatchedState* b = s->batched;
  if (!b->d_sub_sizes || c1 <= c0) {
    if (launch_zstd_range(s, slot, z, c0, c1, st))
      return 1;
    BCHECK(hipEventRecord(done, st));
    return 0;
  }
  const uint64_t T = s->plan.total_chunks;
  BCHECK(launch_zstd_parts(s->d_pool + (uint64_t)slot * s->epoch_pool_bytes + c0 * s->chunk_stride_bytes,
                            s->chunk_stride_bytes, s->chunk_bytes, c1 - c0,
                            s->d_comp + ((uint64_t)slot * T + c0) * s->chunk_slot, s->chunk_slot,
                            s->d_seg_sizes + ((uint64_t)slot * T + c0) * s->nseg, b->d_sub_sizes + z * b->sub_stride,
                            s->zblock, (uint8_t*)s->d_zscratch + zstd_scratch_bytes(z * b->zip_workers),
                            b->zip_workers, b->d_queues + 16 * z, st, done));
  return 0;
}

// Aggregate of the levels in `mask`; in DMA mode their bytes are packed from
// d_stage + base on.
int launch_aggregate(tile_stream_gpu* s, BatchedState* b, uint32_t ps, uint32_t hs, uint32_t mask,
                     hipStream_t st, uint64_t base = 0)
{
  const SegmentSource src = source_of(s);
  auto& P = b->pool[ps];
  auto& H = b->host[hs];
  if (s->plan.nlod <= kAggMaxLevels) {
    // every active level in two kernels (each graph node costs host time at
    // every launch); the plan writes sizes and offsets to pinned memory too
    AggLevels a = {};
    for (int lv = 0; lv < s->plan.nlod; ++lv) {
      if (!(mask & (1u << lv)))
        continue;
      const LevelInfo& L = s->plan.levels[lv];
      const int k = a.nlev++;
      a.gather[k] = P.d_gather[lv];
      a.perm[k] = P.d_perm[lv];
      a.n_items[k] = L.agg.chunks_per_epoch;
      a.n_slots[k] = L.agg.covering_count;
      a.group[k] = L.agg.cps_inner;
      a.sizes[k] = P.d_sizes[lv];
      a.offsets[k] = P.d_offsets[lv];
      a.host_sizes[k] = H.h_sizes[lv];
      a.host_offsets[k] = H.h_offsets[lv];
      a.dst[k] = b->dma ? H.d_stage + base : H.h_agg + H.level_base[lv];
    }
    // DMA mode packs the levels back to back, so one copy moves the epoch
    a.packed = b->dma ? 1 : 0;
    if (a.nlev)
      BCHECK(launch_aggregate_levels(src, a, st, b->dma ? 0 : s->host_gather_wg));
    return 0;
  }
  for (int lv = 0; lv < s->plan.nlod; ++lv) {
    if (!(mask & (1u << lv)))
      continue;
    const LevelInfo& L = s->plan.levels[lv];
    const uint64_t M = L.agg.chunks_per_epoch, C = L.agg.covering_count;
    // the plan kernel writes sizes and offsets to pinned memory as well: no
    // copy nodes (each node costs host time at every graph launch)
    BCHECK(launch_aggregate_plan(src, P.d_gather[lv], P.d_perm[lv], M, C, L.agg.cps_inner, 0,
                                 P.d_sizes[lv], P.d_offsets[lv], st, H.h_sizes[lv], H.h_offsets[lv]));
    if (b->dma)
      BCHECK(launch_aggregate_gather(src, P.d_gather[lv], P.d_perm[lv], M, P.d_offsets[lv],
                                     H.d_stage + H.level_base[lv], st, 0));
    else
      BCHECK(launch_aggregate_gather(src, P.d_gather[lv], P.d_perm[lv], M, P.d_offsets[lv],
                                     H.h_agg + H.level_base[lv], st, s->host_gather_wg));
  }
  return 0;
}

// Capture `fn(stream)` into an executable graph.
template<typename F>
int capture(hipStream_t st, hipGraphExec_t* exec, F fn)
{
  std::lock_guard<std::recursive_mutex> lk(graph_mutex());
  hipGraph_t g = nullptr;
  BCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = fn(st);
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc || e != hipSuccess) {
    if (g)
      (void)hipGraphDestroy(g);
    set_error("batched: graph capture failed");
    return 1;
  }
  const hipError_t ie = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    set_error(std::string("batched: graph instantiate: ") + hipGetErrorString(ie));
    return 1;
  }
  return 0;
}

// DMA mode: queue the copies of a gathered epoch (its offsets are in pinned
// memory once the aggregate graph has finished).
// Byte offset of level lv's bytes in a host slot for an epoch with `mask`:
// packed (DMA mode, at most kAggMaxLevels levels) right after the previous
// active levels' bytes, else at the level's worst-case base.
uint64_t level_offset(const tile_stream_gpu* s, const BatchedState* b, const BatchedState::HostSlot& H,
                      const BatchedState::Ticket& t, int lv)
{
  if (!b->dma || s->plan.nlod > kAggMaxLevels)
    return H.level_base[lv];
  // a split epoch keeps level 0 at 0 and packs the coarse levels from
  // level 1's base on
  const int from = t.mask_a && lv > 0 ? 1 : 0;
  uint64_t off = from ? H.level_base[1] : 0;
  for (int j = from; j < lv; ++j)
    if (t.mask & (1u << j))
      off += H.h_offsets[j][s->plan.levels[j].agg.covering_count];
  return off;
}

// Queues the DMA copy of one part of a ticket: level 0 of a split epoch
// (part_a), else the rest (every level of an unsplit epoch).
int issue_copies(tile_stream_gpu* s, BatchedState* b, BatchedState::Ticket& t, bool part_a)
{
  auto& H = b->host[t.host_slot];
  const uint32_t mask = part_a ? t.mask_a : (t.mask & ~t.mask_a);
  if (s->plan.nlod <= kAggMaxLevels) {
    // the part's levels sit back to back: one copy
    const uint64_t base = !part_a && t.mask_a ? H.level_base[1] : 0;
    uint64_t bytes = 0;
    for (int lv = 0; lv < s->plan.nlod; ++lv)
      if (mask & (1u << lv))
        bytes += H.h_offsets[lv][s->plan.levels[lv].agg.covering_count];
    if (bytes)
      BCHECK(hipMemcpyAsync(H.h_agg + base, H.d_stage + base, bytes, hipMemcpyDeviceToHost,
                            b->st_copy[t.host_slot]));
  } else {
    for (int lv = 0; lv < s->plan.nlod; ++lv) {
      if (!(mask & (1u << lv)))
        continue;
      const uint64_t C = s->plan.levels[lv].agg.covering_count;
      const uint64_t bytes = H.h_offsets[lv][C];
      if (bytes)
        BCHECK(hipMemcpyAsync(H.h_agg + H.level_base[lv], H.d_stage + H.level_base[lv], bytes,
                              hipMemcpyDeviceToHost, b->st_copy[t.host_slot]));
    }
  }
  (part_a ? t.copied_a : t.copied) = true;
  return 0;
}

// The copier thread: the next part without a copy (in epoch order) is
// waited for and its copy queued; sleeps while there is none.
void copier_loop(tile_stream_gpu* s, BatchedState* b)
{
  (void)hipSetDevice(s->device);
  std::unique_lock<std::recursive_mutex> lk(b->mu);
  for (;;) {
    BatchedState::Ticket* t = nullptr;
    bool part_a = false;
    for (auto& x : b->inflight) {
      if (x.mask_a && !x.copied_a) {
        t = &x;
        part_a = true;
        break;
      }
      if (!x.copied) {
        t = &x;
        break;
      }
    }
    if (!t) {
      if (b->stop)
        return;
      b->cv.wait(lk);
      continue;
    }
    const hipEvent_t ev = part_a ? t->done_a : t->done;
    lk.unlock();
    const hipError_t e = hipEventSynchronize(ev);
    lk.lock();
    // the ticket stays in flight until its copies are done, so t is valid
    if (e != hipSuccess || issue_copies(s, b, *t, part_a)) {
      // the submitting thread reports it at its next delivery; no further copies
      b->copy_error = std::string("batched: copy thread: ") +
                      (e != hipSuccess ? hipGetErrorString(e) : chucky_last_error());
      b->cv.notify_all();
      while (!b->stop)
        b->cv.wait(lk);
      return;
    }
    // a delivery may be waiting for this copy to be queued
    b->cv.notify_all();
  }
}

// Hands the oldest in-flight epoch to the sink once its bytes are in host
// memory. Returns 2 when `block` is false and it is not ready yet.
int deliver_one(tile_stream_gpu* s, BatchedState* b, bool block)
{
  if (b->inflight.empty())
    return 0;
  auto& t = b->inflight.front();
  if (b->dma) {
    // the copier thread queues the copies; wait for them to finish
    std::unique_lock<std::recursive_mutex> lk(b->mu);
    for (;;) {
      if (!b->copy_error.empty()) {
        set_error(b->copy_error);
        return 1;
      }
      if (t.copied && (!t.mask_a || t.copied_a)) {
        // every copy of the ticket is queued on its host slot's stream, and
        // only those (the slot is not reused before this delivery)
        const hipStream_t cs = b->st_copy[t.host_slot];
        const hipError_t q = hipStreamQuery(cs);
        if (q == hipSuccess)
          break;
        if (q != hipErrorNotReady)
          BCHECK(q);
        // a stream query's "not ready" stays behind as the thread's last
        // error, which the next launch check would take for its own
        (void)hipGetLastError();
        if (!block)
          return 2;
        lk.unlock();
        BCHECK(hipStreamSynchronize(cs));
        lk.lock();
        continue;
      }
      if (!block)
        return 2;
      // the copier thread signals once it has queued a copy (or failed)
      b->cv.wait(lk);
    }
    // (the ticket keeps its copy events until it leaves the queue: the
    // copier thread takes a ticket without them for one still to copy)
  } else if (block) {
    BCHECK(hipEventSynchronize(t.done));
  } else {
    const hipError_t q = hipEventQuery(t.done);
    if (q == hipErrorNotReady)
      return 2;
    BCHECK(q);
  }
  auto& H = b->host[t.host_slot];
  // one sample per compress launch: its input bytes and event-timed duration
  float ms = 0;
  if (t.c0 && t.c1 && hipEventElapsedTime(&ms, t.c0, t.c1) == hipSuccess)
    metric_add(s->metrics.compress, ms, t.bytes01, 0);
  if (t.c2 && t.c3 && hipEventElapsedTime(&ms, t.c2, t.c3) == hipSuccess)
    metric_add(s->metrics.compress, ms, t.bytes23, 0);
  {
    std::lock_guard<std::recursive_mutex> lk(b->mu);
    for (hipEvent_t e : { t.c0, t.c1, t.c2, t.c3 })
      if (e)
        b->timing_pool.push_back(e);
  }
  scatter_collect_ready(s);
  const auto t1 = std::chrono::steady_clock::now();
  for (int lv = 0; lv < s->plan.nlod; ++lv) {
    if (!(t.mask & (1u << lv)))
      continue;
    AggView v{ H.h_agg + level_offset(s, b, H, t, lv), H.h_offsets[lv], H.h_sizes[lv] };
    size_t out = 0;
    if (deliver_batch((uint8_t)lv, s->shards[lv], v, 1, s->sink, 0, &out))
      return 1;
    s->metrics.sink.input_bytes += (double)out;
    s->metrics.d2h.input_bytes += (double)out;
    if (s->sink->record_fence)
      H.fence[lv] = s->sink->record_fence(s->sink, (uint8_t)lv);
  }
  const auto t2 = std::chrono::steady_clock::now();
  s->metrics.sink.ms += std::chrono::duration<float, std::milli>(t2 - t1).count();
  s->metrics.sink.count++;
  s->latencies.push_back(std::chrono::duration<double>(t2 - t.t0).count());
  s->epochs_written++;
  std::lock_guard<std::recursive_mutex> lk(b->mu);
  b->event_pool.push_back(t.done);
  if (t.done_a)
    b->event_pool.push_back(t.done_a);
  b->inflight.pop_front();
  return 0;
}

} // namespace

int batched_init(tile_stream_gpu* s)
{
  auto* b = new BatchedState();
  s->batched = b;
  const Plan& p = s->plan;
  const uint32_t K = p.epochs_per_batch;
  b->dma = s->opts.batched_d2h != 1;
  b->graphs = s->opts.batched_launch == 1;
  b->R = std::min<uint32_t>(K, 4);
  b->zip_workers = s->zworkers;
  if (p.config.codec.id == CODEC_ZSTD && !s->d_ztables) {
    const uint64_t blocks = std::max<uint64_t>((uint64_t)p.total_chunks * s->nseg, 1);
    b->NZ = (uint32_t)std::min<uint64_t>(std::min<uint32_t>(K, kMaxZip), std::max<uint64_t>(1, s->zworkers / blocks));
    b->zip_workers = (uint32_t)std::min<uint64_t>(blocks, s->zworkers / b->NZ);
  } else if (p.config.codec.id == CODEC_LZ4_NON_STANDARD) {
    // LZ4 launches one wavefront per chunk or segment; epochs whose
    // wavefronts fill at most half of the device compress side by side
    const uint64_t per_chunk = s->lz4_seg ? (s->chunk_bytes + s->lz4_seg - 1) / s->lz4_seg : 1;
    const uint64_t waves = std::max<uint64_t>((uint64_t)p.total_chunks * per_chunk, 1);
    b->NZ = (uint32_t)std::min<uint64_t>(std::min<uint32_t>(K, kMaxZip),
                                         std::max<uint64_t>(1, lz4_wave_slots(s->lz4_hash_log, s->lz4_seg) / waves));
    b->lz4_stride = align_up(lz4_scratch_bytes(s->chunk_bytes, p.total_chunks, s->lz4_seg), 256);
  }
  // HIP maps streams onto its hardware queues (GPU_MAX_HW_QUEUES, 4 by
  // default) in creation order, and kernels of streams sharing a queue run
  // in submission order. The streams that launch kernels are created right
  // after the compute stream, so that they land on distinct queues: the
  // compression queues (each also runs its epochs' aggregates) or, with one
  // compression queue (the compute stream itself), the aggregate stream; the
  // copy-only stream follows.
  if (b->NZ > 1) {
    b->st_zip.assign(b->NZ, nullptr);
    for (auto& z : b->st_zip)
      BCHECK(hipStreamCreateWithFlags(&z, hipStreamNonBlocking));
  } else {
    BCHECK(hipStreamCreateWithFlags(&b->st_agg, hipStreamNonBlocking));
  }
  b->st_copy.assign(b->R, nullptr);
  for (auto& c : b->st_copy)
    BCHECK(hipStreamCreateWithFlags(&c, hipStreamNonBlocking));
  if (dev_alloc(s, &b->d_queues, 16 * b->NZ))
    return 1;
  // split epochs (see batched_submit) compress their coarse levels in
  // blocks of a part of a segment
  if (b->NZ > 1 && b->dma && !b->graphs && p.nlod > 1 && p.nlod <= kAggMaxLevels &&
      p.config.codec.id == CODEC_ZSTD && zstd_parts(s->chunk_bytes, s->zblock)) {
    b->sub_stride = align_up((uint64_t)p.total_chunks * s->nseg * kZstdMaxParts, 64);
    if (dev_alloc(s, &b->d_sub_sizes, b->NZ * b->sub_stride))
      return 1;
  }
  if (b->NZ > 1 && b->lz4_stride && dev_alloc(s, &b->d_lz4, b->NZ * b->lz4_stride))
    return 1;
  // the compressors' work counters start at zero and each launch leaves them so
  BCHECK(hipMemsetAsync(b->d_queues, 0, 16 * b->NZ * sizeof(uint32_t), s->st_comp));
  BCHECK(hipStreamSynchronize(s->st_comp));
  // worst-case bytes of one epoch per level
  std::vector<uint64_t> level_bytes(p.nlod);
  uint64_t total = 0;
  for (int lv = 0; lv < p.nlod; ++lv) {
    const uint64_t per = p.config.codec.id == CODEC_NONE ? s->chunk_bytes : p.max_output_size;
    level_bytes[lv] = align_up(p.levels[lv].agg.chunks_per_epoch * per, 256);
    total += level_bytes[lv];
  }
  b->host.resize(b->R);
  for (auto& H : b->host) {
    if (host_alloc(s, &H.h_agg, total))
      return 1;
    if (b->dma && dev_alloc(s, &H.d_stage, total))
      return 1;
    H.level_base.assign(p.nlod, 0);
    H.h_offsets.assign(p.nlod, nullptr);
    H.h_sizes.assign(p.nlod, nullptr);
    H.fence.assign(p.nlod, io_event{ 0 });
    uint64_t base = 0;
    for (int lv = 0; lv < p.nlod; ++lv) {
      H.level_base[lv] = base;
      base += level_bytes[lv];
      const uint64_t C = p.levels[lv].agg.covering_count;
      if (host_alloc(s, &H.h_offsets[lv], C + 1) || host_alloc(s, &H.h_sizes[lv], C))
        return 1;
    }
  }
  b->pool.resize(K);
  std::vector<uint32_t> g, pm;
  for (uint32_t ps = 0; ps < K; ++ps) {
    auto& P = b->pool[ps];
    P.d_gather.assign(p.nlod, nullptr);
    P.d_perm.assign(p.nlod, nullptr);
    P.d_sizes.assign(p.nlod, nullptr);
    P.d_offsets.assign(p.nlod, nullptr);
    for (int lv = 0; lv < p.nlod; ++lv) {
      const uint64_t M = p.levels[lv].agg.chunks_per_epoch, C = p.levels[lv].agg.covering_count;
      if (dev_alloc(s, &P.d_gather[lv], M) || dev_alloc(s, &P.d_perm[lv], M) ||
          dev_alloc(s, &P.d_sizes[lv], C) || dev_alloc(s, &P.d_offsets[lv], C + 1))
        return 1;
      g.assign(M, 0);
      pm.assign(M, 0);
      const uint32_t pe[1] = { ps };
      batch_luts(p, lv, 1, pe, g.data(), pm.data());
      BCHECK(hipMemcpyAsync(P.d_gather[lv], g.data(), M * 4, hipMemcpyHostToDevice, s->st_comp));
      BCHECK(hipMemcpyAsync(P.d_perm[lv], pm.data(), M * 4, hipMemcpyHostToDevice, s->st_comp));
      BCHECK(hipStreamSynchronize(s->st_comp));
    }
    BCHECK(hipEventCreateWithFlags(&P.ev_filled, hipEventDisableTiming));
    BCHECK(hipEventCreateWithFlags(&P.ev_compressed, hipEventDisableTiming));
    BCHECK(hipEventCreateWithFlags(&P.ev_gathered, hipEventDisableTiming));
    BCHECK(hipEventRecord(P.ev_gathered, s->st_comp));
    if (b->graphs && p.config.codec.id != CODEC_NONE &&
        capture(s->st_comp, &P.compress_exec,
                [&](hipStream_t st) { return launch_compress(s, ps, st, p.total_chunks); }))
      return 1;
  }
  if (b->dma)
    b->copier = std::thread(copier_loop, s, b);
  return 0;
}

void batched_free(tile_stream_gpu* s)
{
  BatchedState* b = s->batched;
  if (!b)
    return;
  if (b->copier.joinable()) {
    {
      std::lock_guard<std::recursive_mutex> lk(b->mu);
      b->stop = true;
    }
    b->cv.notify_all();
    b->copier.join();
  }
  if (b->st_agg)
    (void)hipStreamSynchronize(b->st_agg);
  for (auto z : b->st_zip)
    if (z)
      (void)hipStreamSynchronize(z);
  for (auto& P : b->pool) {
    if (P.compress_exec)
      (void)hipGraphExecDestroy(P.compress_exec);
    if (P.ev_filled)
      (void)hipEventDestroy(P.ev_filled);
    if (P.ev_compressed)
      (void)hipEventDestroy(P.ev_compressed);
    if (P.ev_gathered)
      (void)hipEventDestroy(P.ev_gathered);
  }
  for (auto& kv : b->agg_exec)
    (void)hipGraphExecDestroy(kv.second);
  for (auto c : b->st_copy)
    if (c)
      (void)hipStreamSynchronize(c);
  for (auto& t : b->inflight) {
    (void)hipEventDestroy(t.done);
    if (t.done_a)
      (void)hipEventDestroy(t.done_a);
    if (t.c0)
      (void)hipEventDestroy(t.c0);
    if (t.c1)
      (void)hipEventDestroy(t.c1);
  }
  for (auto e : b->event_pool)
    (void)hipEventDestroy(e);
  for (auto e : b->timing_pool)
    (void)hipEventDestroy(e);
  if (b->st_agg)
    (void)hipStreamDestroy(b->st_agg);
  for (auto c : b->st_copy)
    if (c)
      (void)hipStreamDestroy(c);
  for (auto z : b->st_zip)
    if (z)
      (void)hipStreamDestroy(z);
  delete b;
  s->batched = nullptr;
}

// Epoch `pool_slot` of the pool is complete (scatter / LOD already queued on
// the compute queue): compress, aggregate and queue it for delivery.
int batched_submit(tile_stream_gpu* s, uint32_t pool_slot, uint32_t mask)
{
  BatchedState* b = s->batched;
  const uint64_t ticket = b->next_epoch++;
  const uint32_t hs = (uint32_t)(ticket % b->R);
  // the host slot's previous epoch must be delivered and its writes retired
  while (!b->inflight.empty() && b->inflight.front().epoch + b->R <= ticket)
    if (deliver_one(s, b, true))
      return 1;
  auto& H = b->host[hs];
  if (s->sink->wait_fence)
    for (int lv = 0; lv < s->plan.nlod; ++lv)
      s->sink->wait_fence(s->sink, (uint8_t)lv, H.fence[lv]);
  auto& P = b->pool[pool_slot];
  s->slot_compressed[pool_slot] = 1;
  const Plan& p = s->plan;
  // An epoch that also emits coarse levels is split over both compression
  // queues: the coarse levels' chunks (their blocks are the slow ones) start
  // first on the slot's queue, level 0 beside them on the other queue, and
  // level 0's gather and copy leave while the coarse levels still compress.
  if (b->NZ > 1 && b->dma && !b->graphs && p.nlod > 1 && p.nlod <= kAggMaxLevels &&
      p.config.codec.id == CODEC_ZSTD && (mask & 1u) && (mask & ~1u)) {
    hipStream_t zb = b->st_zip[pool_slot % b->NZ], za = b->st_zip[(pool_slot + 1) % b->NZ];
    const uint64_t n0 = p.levels[1].chunk_offset; // level 0 chunks come first
    BCHECK(hipEventRecord(P.ev_filled, s->st_comp));
    for (hipStream_t z : { zb, za }) {
      BCHECK(hipStreamWaitEvent(z, P.ev_filled, 0));
      BCHECK(hipStreamWaitEvent(z, P.ev_gathered, 0));
    }
    const uint64_t n1 = active_chunk_end(p, mask);
    hipEvent_t c0 = take_timing_event(b), c1 = take_timing_event(b);
    hipEvent_t c2 = take_timing_event(b), c3 = take_timing_event(b);
    hipEvent_t done_a = take_event(b), done = take_event(b);
    if (!c0 || !c1 || !c2 || !c3 || !done_a || !done) {
      set_error("batched: event creation failed");
      return 1;
    }
    BCHECK(hipEventRecord(c0, zb));
    if (launch_zstd_coarse(s, pool_slot, pool_slot % b->NZ, n0, n1, zb, c1))
      return 1;
    BCHECK(hipEventRecord(c2, za));
    if (launch_zstd_range(s, pool_slot, (pool_slot + 1) % b->NZ, 0, n0, za))
      return 1;
    BCHECK(hipEventRecord(c3, za));
    BCHECK(hipEventRecord(P.ev_compressed, za));
    if (launch_aggregate(s, b, pool_slot, hs, 1u, za, 0))
      return 1;
    BCHECK(hipEventRecord(done_a, za));
    BCHECK(hipStreamWaitEvent(zb, P.ev_compressed, 0));
    if (launch_aggregate(s, b, pool_slot, hs, mask & ~1u, zb, H.level_base[1]))
      return 1;
    BCHECK(hipStreamWaitEvent(zb, done_a, 0));
    BCHECK(hipEventRecord(P.ev_gathered, zb));
    BCHECK(hipEventRecord(done, zb));
    BatchedState::Ticket t{ ticket, pool_slot, hs, mask, done, false, c0, c1,
                            std::chrono::steady_clock::now() };
    t.bytes01 = (double)(n1 - n0) * s->chunk_bytes;
    t.c2 = c2;
    t.c3 = c3;
    t.bytes23 = (double)n0 * s->chunk_bytes;
    t.mask_a = 1u;
    t.done_a = done_a;
    {
      std::lock_guard<std::recursive_mutex> lk(b->mu);
      b->inflight.push_back(t);
    }
    b->cv.notify_all();
    for (;;) {
      const int r = deliver_one(s, b, false);
      if (r == 1)
        return 1;
      if (r == 2 || b->inflight.empty())
        break;
    }
    return 0;
  }
  // the slot's compression queue takes over once the epoch is in the pool;
  // compressed slots of this pool slot are free once its last gather is done
  hipStream_t zs = s->st_comp;
  if (b->NZ > 1) {
    zs = b->st_zip[pool_slot % b->NZ];
    BCHECK(hipEventRecord(P.ev_filled, s->st_comp));
    BCHECK(hipStreamWaitEvent(zs, P.ev_filled, 0));
  }
  BCHECK(hipStreamWaitEvent(zs, P.ev_gathered, 0));
  hipEvent_t c0 = nullptr, c1 = nullptr;
  // direct launches compress the active levels' chunks only (levels are
  // stored in order, so they are the chunks below the last active level's
  // end); the captured graph takes every chunk of the slot
  const uint64_t n_comp = b->graphs ? p.total_chunks : active_chunk_end(p, mask);
  if (s->plan.config.codec.id != CODEC_NONE) {
    c0 = take_timing_event(b);
    c1 = take_timing_event(b);
    if (!c0 || !c1) {
      set_error("batched: event creation failed");
      return 1;
    }
    BCHECK(hipEventRecord(c0, zs));
    if (b->graphs)
      BCHECK(hipGraphLaunch(P.compress_exec, zs));
    else if (launch_compress(s, pool_slot, zs, n_comp))
      return 1;
    BCHECK(hipEventRecord(c1, zs));
  }
  BCHECK(hipEventRecord(P.ev_compressed, zs));
  // the aggregate runs on the compression queue after the compress, or on its
  // own stream when the compute stream compresses
  hipStream_t as = b->NZ > 1 ? zs : b->st_agg;
  if (as != zs)
    BCHECK(hipStreamWaitEvent(as, P.ev_compressed, 0));
  if (b->graphs) {
    const auto key = std::make_tuple(pool_slot, hs, mask);
    auto it = b->agg_exec.find(key);
    if (it == b->agg_exec.end()) {
      hipGraphExec_t ex = nullptr;
      if (capture(as, &ex, [&](hipStream_t st) { return launch_aggregate(s, b, pool_slot, hs, mask, st); }))
        return 1;
      it = b->agg_exec.emplace(key, ex).first;
    }
    BCHECK(hipGraphLaunch(it->second, as));
  } else if (launch_aggregate(s, b, pool_slot, hs, mask, as)) {
    return 1;
  }
  BCHECK(hipEventRecord(P.ev_gathered, as));
  hipEvent_t done = take_event(b);
  if (!done) {
    set_error("batched: event creation failed");
    return 1;
  }
  BCHECK(hipEventRecord(done, as));
  {
    std::lock_guard<std::recursive_mutex> lk(b->mu);
    b->inflight.push_back({ ticket, pool_slot, hs, mask, done, false, c0, c1,
                            std::chrono::steady_clock::now(), (double)n_comp * s->chunk_bytes });
  }
  b->cv.notify_all();
  // opportunistic in-order delivery of whatever has finished
  for (;;) {
    const int r = deliver_one(s, b, false);
    if (r == 1)
      return 1;
    if (r == 2 || b->inflight.empty())
      break;
  }
  return 0;
}

// The next epoch will be written into pool slot `pool_slot` (transpose, LOD
// scatter): queue a wait on the compute queue until the slot's previous
// epoch has been gathered. Uncompressed streams gather straight from the pool,
// so without it a slot could be overwritten while the aggregate queue reads it.
int batched_acquire_slot(tile_stream_gpu* s, uint32_t pool_slot)
{
  BCHECK(hipStreamWaitEvent(s->st_comp, s->batched->pool[pool_slot].ev_gathered, 0));
  return 0;
}

int batched_drain(tile_stream_gpu* s)
{
  BatchedState* b = s->batched;
  while (!b->inflight.empty())
    if (deliver_one(s, b, true))
      return 1;
  if (s->sink->wait_fence)
    for (auto& H : b->host)
      for (int lv = 0; lv < s->plan.nlod; ++lv)
        s->sink->wait_fence(s->sink, (uint8_t)lv, H.fence[lv]);
  BCHECK(hipStreamSynchronize(s->st_comp));
  for (auto z : b->st_zip)
    BCHECK(hipStreamSynchronize(z));
  if (b->st_agg)
    BCHECK(hipStreamSynchronize(b->st_agg));
  for (auto c : b->st_copy)
    BCHECK(hipStreamSynchronize(c));
  scatter_collect_ready(s);
  return 0;
}

void batched_reset(tile_stream_gpu* s)
{
  BatchedState* b = s->batched;
  if (!b)
    return;
  b->next_epoch = 0;
  // The host slots' fences belong to the sink of the previous acquisition:
  // its writes out of the slots end before the slots are reused, and a reset
  // may hand the stream another sink, whose fences count from zero.
  if (s->sink && s->sink->wait_fence)
    for (auto& H : b->host)
      for (int lv = 0; lv < s->plan.nlod; ++lv)
        s->sink->wait_fence(s->sink, (uint8_t)lv, H.fence[lv]);
  for (auto& H : b->host)
    std::fill(H.fence.begin(), H.fence.end(), io_event{ 0 });
}

} // namespace chucky
