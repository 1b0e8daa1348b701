// Internal state of a GPU stream (both pipelines).
#pragma once

#include "chucky_amd.h"
#include "host_copy.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "shards.hpp"

#include <hip/hip_runtime.h>

#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <mutex>
#include <vector>

namespace chucky {
struct BatchedState;
}

struct tile_stream_gpu
{
  writer w; // must stay the first member (writer* <-> stream*)
  chucky::Plan plan;
  chucky::LodLuts luts;
  shard_sink* sink = nullptr;
  tile_stream_gpu_options opts{};
  int device = 0;
  int pipeline = CHUCKY_PIPELINE_PLAIN;

  hipStream_t st_h2d = nullptr, st_comp = nullptr, st_d2h = nullptr;

  // host staging ring (host-memory appends): kStageSlots pinned buffers and
  // their device twins, so that a piece's host copy overlaps the H2D copy of
  // the one before and the launches of the epoch before that
  static constexpr int kStageSlots = 3;
  size_t stage_cap = 0;
  uint8_t* h_stage[kStageSlots] = {};
  uint8_t* d_stage[kStageSlots] = {};
  // device addresses of h_stage, read by the H2D copy kernel (option
  // h2d_copy = 1); h2d_groups is its grid (0: the DMA engines copy instead)
  uint8_t* h_stage_dev[kStageSlots] = {};
  uint32_t h2d_groups = 0;
  hipEvent_t ev_h2d[kStageSlots] = {}, ev_stage_free[kStageSlots] = {};
  size_t stage_fill = 0;
  size_t stage_sent = 0; // bytes of the current slot whose H2D copy is already queued
  int stage_cur = 0;
  // host appends: a finished epoch's launches are queued right after the
  // next piece's H2D copy (so the copy engine moves the next epoch while the
  // device works on this one); set while that copy is still to be queued
  bool epoch_pending = false;
  std::unique_ptr<chucky::CopyPool> copy_pool; // staging copy threads, started at the first host append
  // a staging copy begun before an epoch's launches: source, bytes, stage
  const uint8_t* pf_src = nullptr;
  size_t pf_n = 0;
  int pf_stage = -1;

  // chunk pool: K epochs x total_chunks x chunk_stride elements
  uint8_t* d_pool = nullptr;
  uint64_t epoch_pool_bytes = 0;
  uint64_t chunk_stride_bytes = 0;
  uint64_t chunk_bytes = 0;

  // compressed slots
  uint8_t* d_comp = nullptr;
  uint32_t* d_seg_sizes = nullptr;
  uint32_t nseg = 1;
  uint64_t seg_stride = 0, chunk_slot = 0;
  void* d_zscratch = nullptr;
  uint32_t zworkers = 0;
  uint32_t zblock = chucky::kZstdBlockMax;
  uint8_t* d_ztables = nullptr; // per-chunk Huffman tables (chunk-table mode), else null
  uint32_t host_gather_wg = 64;
  uint32_t* d_queue = nullptr;
  int lz4_hash_log = 12;
  uint32_t lz4_seg = 0; // segment bytes of the multi-wavefront LZ4 path (0 = one per chunk)
  uint8_t* d_lz4_scratch = nullptr;

  // aggregate
  uint32_t* d_gather = nullptr;
  uint32_t* d_perm = nullptr;
  uint32_t* h_lut = nullptr; // pinned: gather then perm
  uint64_t lut_cap = 0;
  uint64_t* d_chunk_sizes = nullptr;
  uint64_t* d_offsets = nullptr;
  uint64_t* h_sizes = nullptr; // pinned: offsets (slots+1) then chunk sizes (slots)
  uint64_t slot_cap = 0;
  uint8_t* d_agg = nullptr;
  uint8_t* h_agg = nullptr;
  uint64_t agg_cap = 0;

  // LOD (multiscale)
  uint8_t* d_linear = nullptr;
  const uint8_t* lod_src = nullptr; // device epoch read in place of d_linear (caller's buffer or staging)
  hipEvent_t lod_src_free = nullptr; // recorded by the LOD gather once it has read lod_src, or null
  bool linear_in_lod = false;       // the last epoch was read in place: d_linear lacks its values
  uint8_t* d_lod = nullptr;     // Morton values, all levels packed
  uint8_t* d_accum = nullptr;   // append-dim accumulators, levels >= 1
  uint32_t* d_gather_tiles = nullptr;  // (Morton base, offset) per full tile
  uint32_t* d_gather_edge = nullptr;   // (Morton rank, offset) per remaining cell
  uint64_t* d_gather_fixed = nullptr;
  uint64_t* d_ends = nullptr;
  std::vector<uint32_t*> d_chunk_tiles;
  std::vector<uint32_t*> d_chunk_edge;
  std::vector<uint64_t*> d_chunk_fixed;
  uint32_t append_counts[LOD_MAX_LEVELS] = {};
  std::vector<uint8_t> slot_compressed; // per pool slot: compressed by a batch since the start

  // stream state
  uint64_t cursor = 0;
  uint32_t batch_acc = 0;
  uint32_t masks[MAX_BATCH_EPOCHS] = {};
  int flushed = 0;
  int at_capacity = 0; // a bounded stream reached its capacity and delivered everything
  std::vector<chucky::ShardState> shards;
  std::vector<io_event> io_done;
  stream_metrics metrics{};
  tile_stream_layout_gpu layout_gpu{};
  std::chrono::steady_clock::time_point meta_clock;

  // latency bookkeeping: completion time of each epoch's append, and the
  // latency samples once its shards are written
  std::vector<std::chrono::steady_clock::time_point> epoch_t0;
  std::vector<double> latencies;
  uint64_t epochs_done = 0;     // epochs whose append completed
  uint64_t epochs_written = 0;  // epochs whose shards have been written

  // event pairs around st_comp launches (scatter, LOD passes) whose
  // durations are not yet accounted, with the metric they go to
  struct PendingTiming
  {
    hipEvent_t a, b;
    double in_bytes, out_bytes;
    stream_metric stream_metrics::*which;
  };
  std::vector<PendingTiming> pending_timing;
  // per-pass event timing of the scatter and LOD launches (stream metrics);
  // off, those launches run back to back without event markers between them
  bool pass_timing = true;
  std::vector<hipEvent_t> timing_free; // recycled timing events

  chucky::BatchedState* batched = nullptr; // non-null: batched pipeline active

  std::vector<void*> allocations;      // device
  std::vector<void*> host_allocations; // pinned host
};

namespace chucky {
void set_error(const std::string& msg);
int lod_run_epoch(tile_stream_gpu* s, uint32_t epoch_in_batch, uint32_t* out_mask);
// Writes level 0 of the Morton buffer back into d_linear (the cells the
// gather reads), on st_comp.
int lod_restore_linear(tile_stream_gpu* s);
int lod_drain(tile_stream_gpu* s, uint32_t* out_mask);
int lod_init(tile_stream_gpu* s);
void metric_add(stream_metric& m, float ms, double in_bytes, double out_bytes);
// Held while a graph is captured and while streams are created or destroyed,
// so that one thread's capture never sees another thread's allocations,
// frees or initial copies.
std::recursive_mutex& graph_mutex();
void scatter_collect_ready(tile_stream_gpu* s);
// Brackets launches on st_comp with events; the duration is added to
// s->metrics.*which once the queue has passed the end event.
int timing_begin(tile_stream_gpu* s, hipEvent_t* a);
int timing_end(tile_stream_gpu* s, hipEvent_t a, stream_metric stream_metrics::*which, double in_bytes,
               double out_bytes);
// Two timing events for launches that record them themselves (the start of
// the first and the end of the last kernel of a pass), and their hand-over
// to the same collection once those launches are queued.
int timing_pair(tile_stream_gpu* s, hipEvent_t* a, hipEvent_t* b);
void timing_add(tile_stream_gpu* s, hipEvent_t a, hipEvent_t b, stream_metric stream_metrics::*which,
                double in_bytes, double out_bytes);
uint32_t compressor_slots();
// hipHostMalloc on the NUMA node of `device` (default policy if unknown).
hipError_t host_alloc_near_gpu(void** p, size_t bytes, int device);
// Returns a buffer of host_alloc_near_gpu to the process's pinned pool (or
// frees it when the pool is full).
void host_release(void* p);
// A new pinned buffer on the GPU's node that bypasses the pool (freed with
// hipHostFree): for measurements of the host link itself.
hipError_t host_alloc_unpooled(void** p, size_t bytes, int device);
// Codec launch shapes from the recorded sweep tables (autotune.cpp), by
// multiscale or not, dtype size, chunk bytes and chunks per compress launch.
uint32_t tune_zstd_block(uint64_t chunk_bytes, uint64_t chunks_per_launch, uint32_t bpe, bool lod);
bool tune_zstd_chunk_tables(uint64_t chunk_bytes, uint64_t chunks_per_launch, uint32_t bpe, bool lod);
struct Lz4Choice
{
  uint32_t seg_bytes; // 0 = one wavefront per chunk
  int hash_log;
};
Lz4Choice tune_lz4(uint64_t chunk_bytes, uint64_t chunks_per_launch, uint32_t bpe, bool lod);
uint32_t tune_host_gather_workgroups();
// grid of the H2D copy kernel of staged pieces
uint32_t tune_h2d_copy_workgroups();
// LZ4 wavefronts resident at once on the device (LDS-bound: hash table and
// staged segment)
uint64_t lz4_wave_slots(int hash_log, uint32_t seg_bytes);
int batched_init(tile_stream_gpu* s);
void batched_free(tile_stream_gpu* s);
int batched_submit(tile_stream_gpu* s, uint32_t pool_slot, uint32_t mask);
int batched_acquire_slot(tile_stream_gpu* s, uint32_t pool_slot);
int batched_drain(tile_stream_gpu* s);
void batched_reset(tile_stream_gpu* s);
} // namespace chucky
