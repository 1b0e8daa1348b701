// Launch wrappers of the gfx950 kernels. All launches are asynchronous on the
// given stream and return hipSuccess or the launch error.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace chucky {

// ---- transpose (scatter into chunk-major pool) ----

struct LiftedSpace
{
  int rank;                // lifted rank (2 * array rank)
  uint64_t shape[64];
  int64_t strides[64];
};

// dst[ravel(space, i_offset + i)] = src[i] for i < n. i_offset may be any
// element position of the stream; only its value modulo the epoch matters.
hipError_t launch_transpose(void* dst, const void* src, uint64_t n, int bpe, uint64_t i_offset,
                            const LiftedSpace& space, hipStream_t stream);

// ---- host-to-device copy by a kernel ----

// True when launch_h2d_copy can copy between these addresses (they agree
// modulo 16).
bool h2d_copy_ok(const void* dst, const void* src);
// dst[0, n) = src[0, n), src a device address of pinned host memory, read by
// at most `groups` workgroups of 256 lanes.
hipError_t launch_h2d_copy(void* dst, const void* src, uint64_t n, uint32_t groups, hipStream_t stream);

// ---- aggregate ----

// Sizes of the compressed segments. A chunk is nseg segments; segment s of
// chunk c lives at base + c*chunk_stride + s*seg_stride and holds
// seg_sizes[c*nseg+s] bytes (or const_size bytes when seg_sizes is null).
struct SegmentSource
{
  const uint8_t* base;
  uint64_t chunk_stride;
  uint64_t seg_stride;
  uint32_t nseg;
  const uint32_t* seg_sizes;
  uint64_t const_size;      // used when seg_sizes == nullptr (last segment gets the remainder)
  uint64_t const_total;     // chunk bytes when seg_sizes == nullptr
  uint32_t pieces;          // set by the gather launchers: pieces a segment is copied in
};

// One workgroup: chunk_sizes[perm[i]] = size(gather[i]) for i < n_items,
// zero elsewhere; optional page padding on the last slot of each group of
// `group` slots; offsets = exclusive scan (n_slots + 1 entries). When
// host_sizes / host_offsets (pinned host memory) are given, the kernel also
// writes both arrays there, so no copy has to follow it.
hipError_t launch_aggregate_plan(const SegmentSource& src, const uint32_t* gather,
                                 const uint32_t* perm, uint64_t n_items, uint64_t n_slots,
                                 uint64_t group, uint64_t page_size, uint64_t* chunk_sizes,
                                 uint64_t* offsets, hipStream_t stream, uint64_t* host_sizes = nullptr,
                                 uint64_t* host_offsets = nullptr);

// Several levels' aggregates of one epoch: a plan workgroup per level, then
// one gather grid over every level's (item, segment) pairs. Host mirrors of
// the sizes and offsets are written by the plan kernel when given.
constexpr int kAggMaxLevels = 16;

struct AggLevels
{
  const uint32_t* gather[kAggMaxLevels];
  const uint32_t* perm[kAggMaxLevels];
  uint64_t n_items[kAggMaxLevels];
  uint64_t n_slots[kAggMaxLevels];
  uint64_t group[kAggMaxLevels];
  uint64_t* sizes[kAggMaxLevels];
  uint64_t* offsets[kAggMaxLevels];
  uint64_t* host_sizes[kAggMaxLevels];
  uint64_t* host_offsets[kAggMaxLevels];
  uint8_t* dst[kAggMaxLevels];
  int nlev;
  // packed: every level's bytes follow the previous level's in dst[0] (one
  // copy moves the whole epoch); dst[k] is then ignored for k > 0
  int packed;
};

hipError_t launch_aggregate_levels(const SegmentSource& src, const AggLevels& a, hipStream_t stream,
                                   uint32_t max_workgroups);

// Copy every item's segments to dst + offsets[perm[i]] (dst may be device or
// pinned host memory).
// max_workgroups caps the grid (0 = one workgroup per (item, segment));
// gathers into host memory use a small cap so they do not hold every CU.
hipError_t launch_aggregate_gather(const SegmentSource& src, const uint32_t* gather,
                                   const uint32_t* perm, uint64_t n_items, const uint64_t* offsets,
                                   uint8_t* dst, hipStream_t stream, uint32_t max_workgroups = 0);

// ---- codecs ----

size_t lz4_max_output_size(size_t chunk_bytes);
// zstd block sizes: powers of two in [kZstdBlockMin, kZstdBlockMax]; each
// block is compressed by one wavefront, so the block size sets the number of
// wavefronts per chunk.
constexpr uint32_t kZstdBlockMin = 4 * 1024;
// blocks up to this size are staged whole in LDS by their worker
constexpr uint32_t kZstdStageBytes = 8 * 1024;
constexpr uint32_t kZstdBlockMax = 128 * 1024;
bool zstd_block_ok(uint32_t block);
size_t zstd_max_output_size(size_t chunk_bytes, uint32_t block);
uint32_t zstd_segments(size_t chunk_bytes, uint32_t block);
uint64_t zstd_segment_stride(uint32_t block);

// LZ4 raw block per chunk: one wavefront per chunk, or (seg_bytes < chunk
// bytes) one wavefront per segment of seg_bytes plus a per-chunk stitch,
// using `scratch` of lz4_scratch_bytes(chunk_bytes, n, seg_bytes).
size_t lz4_scratch_bytes(uint64_t chunk_bytes, uint64_t n, uint32_t seg_bytes);
// LDS bytes of one LZ4 wavefront: the hash table, plus the staged input of a
// segment when segments of seg_bytes (0 = whole chunks) are parsed from LDS.
uint32_t lz4_lds_bytes(int hash_log, uint32_t seg_bytes);
hipError_t launch_lz4(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                      uint8_t* dst, uint64_t dst_stride, uint32_t* sizes, int hash_log,
                      uint32_t seg_bytes, void* scratch, hipStream_t stream);

// zstd frame per chunk, split into independent blocks (segments) of `block` bytes.
// Uploads the predefined FSE tables to the current device (once per device;
// call before the first launch_zstd on a device, outside graph capture).
hipError_t zstd_prepare();
// scratch: see zstd_scratch_bytes(workers). queue_counter (two u32, zero
// before the first launch on them, e.g. zeroed with their allocation): the
// persistent workers take blocks from it and the last worker out sets both
// words back to zero, so launches on one counter need no clearing between
// them. Without it, one workgroup per block (blocks <= workers).
size_t zstd_scratch_bytes(uint32_t workers);
hipError_t launch_zstd(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                       uint8_t* dst, uint64_t chunk_dst_stride, uint32_t* seg_sizes, int level,
                       uint32_t block, void* scratch, uint32_t workers, uint32_t* queue_counter,
                       void* tables, hipStream_t stream);
// launch_zstd with the predefined FSE tables at `predefined` (device memory
// holding zstd_predefined_bytes() bytes written by zstd_build_predefined)
// instead of the per-device copy zstd_prepare() allocates.
hipError_t launch_zstd_pre(const void* predefined, const uint8_t* src, uint64_t src_stride,
                           uint64_t chunk_bytes, uint64_t n, uint8_t* dst, uint64_t chunk_dst_stride,
                           uint32_t* seg_sizes, int level, uint32_t block, void* scratch, uint32_t workers,
                           uint32_t* queue_counter, void* tables, hipStream_t stream);
size_t zstd_predefined_bytes();
// launch_zstd over segments of seg_block bytes, each compressed as `parts`
// blocks of seg_block / parts by as many wavefronts (more parallelism for
// launches of few chunks, e.g. coarse LOD levels: a launch lasts as long as
// its slowest block), the later blocks then moved behind the first: the same
// segment layout and seg_sizes as launch_zstd with block seg_block. parts =
// zstd_parts(chunk_bytes, seg_block): the most blocks per segment (up to
// kZstdMaxParts) of at least kZstdPartMin bytes that the worker stages in LDS,
// 0 when the segments cannot be split. sub_sizes: kZstdMaxParts u32 per
// segment of scratch. Needs a queue counter and 16-byte aligned chunks.
// `between` (optional) is recorded on the stream after the compression and
// before the join.
constexpr uint32_t kZstdPartMin = 2 * 1024;
constexpr uint32_t kZstdMaxParts = 4;
uint32_t zstd_parts(uint64_t chunk_bytes, uint32_t seg_block);
hipError_t launch_zstd_parts(const uint8_t* src, uint64_t src_stride, uint64_t chunk_bytes, uint64_t n,
                             uint8_t* dst, uint64_t chunk_dst_stride, uint32_t* seg_sizes,
                             uint32_t* sub_sizes, uint32_t seg_block, void* scratch, uint32_t workers,
                             uint32_t* queue_counter, hipStream_t stream, hipEvent_t between = nullptr);
// Per-chunk frame sizes into sizes[0, n) from the segment sizes of a batch
// compress; with nseg > 1 (zstd blocks at seg_stride) segments 1.. of each
// chunk are first moved down behind segment 0, so every frame is contiguous
// at dst + i*dst_stride.
hipError_t launch_frame_sizes(uint8_t* dst, uint64_t dst_stride, const uint32_t* seg, uint32_t nseg,
                              uint64_t seg_stride, uint64_t n, uint64_t* sizes, hipStream_t stream);
// Writes the predefined FSE tables into device memory at dst (one thread).
hipError_t zstd_build_predefined(void* dst, hipStream_t stream);
// Chunk-table mode (tables != nullptr, zstd_table_bytes(n) of device memory):
// one Huffman table per chunk, built from the chunk's byte histogram by a
// workgroup per chunk, carried by a small first block of the frame and reused
// by every block as treeless literals; otherwise every block builds its own.
size_t zstd_table_bytes(uint64_t n);

} // namespace chucky
