// Kernel-level C entry points (device pointers, synchronous on the null
// stream of the current device). Used by parity tests and the autotuner.
#include "chucky_amd.h"
#include "kernels.hpp"

#include <string>

namespace chucky {
void set_error(const std::string& msg);
}
using namespace chucky;

extern "C" {

size_t chucky_gpu_max_output_size(enum compression_codec codec, size_t chunk_bytes)
{
  switch (codec) {
    case CODEC_NONE: return chunk_bytes;
    case CODEC_LZ4_NON_STANDARD: return lz4_max_output_size(chunk_bytes);
    case CODEC_ZSTD: return zstd_max_output_size(chunk_bytes, kZstdBlockMax);
    default: return 0;
  }
}

int chucky_gpu_compress(enum compression_codec codec, int level, const void* d_src,
                        size_t src_stride, size_t chunk_bytes, uint64_t n, void* d_dst,
                        size_t dst_stride, uint64_t* d_sizes)
{
  if (n == 0)
    return 0;
  if (dst_stride < chucky_gpu_max_output_size(codec, chunk_bytes)) {
    set_error("chucky_gpu_compress: dst_stride below the codec bound");
    return 1;
  }
  hipStream_t st = nullptr;
  uint32_t nseg = codec == CODEC_ZSTD ? zstd_segments(chunk_bytes, kZstdBlockMax) : 1;
  uint32_t* seg = nullptr;
  void* scratch = nullptr;
  uint32_t* queue = nullptr;
  int rc = 1;
  if (hipStreamCreate(&st) != hipSuccess)
    return 1;
  if (hipMalloc((void**)&seg, n * nseg * 4 + 256) != hipSuccess)
    goto done;
  if (codec == CODEC_NONE) {
    set_error("chucky_gpu_compress: CODEC_NONE has no compressor");
    goto done;
  } else if (codec == CODEC_LZ4_NON_STANDARD) {
    if (launch_lz4((const uint8_t*)d_src, src_stride, chunk_bytes, n, (uint8_t*)d_dst, dst_stride,
                   seg, 12, 0, nullptr, st) != hipSuccess)
      goto done;
  } else if (codec == CODEC_ZSTD) {
    const uint32_t workers = (uint32_t)(n * nseg < 2048 ? n * nseg : 2048);
    if (zstd_prepare() != hipSuccess)
      goto done;
    if (hipMalloc(&scratch, zstd_scratch_bytes(workers)) != hipSuccess ||
        hipMalloc((void**)&queue, 256) != hipSuccess || hipMemsetAsync(queue, 0, 256, st) != hipSuccess)
      goto done;
    if (launch_zstd((const uint8_t*)d_src, src_stride, chunk_bytes, n, (uint8_t*)d_dst, dst_stride,
                    seg, level, kZstdBlockMax, scratch, workers, queue, nullptr, st) != hipSuccess)
      goto done;
  } else {
    set_error("chucky_gpu_compress: unsupported codec");
    goto done;
  }
  // per-chunk frame sizes; zstd segments moved behind segment 0 on the
  // device (the moves overlap, which a copy call does not allow)
  if (launch_frame_sizes((uint8_t*)d_dst, dst_stride, seg, nseg,
                         codec == CODEC_ZSTD ? zstd_segment_stride(kZstdBlockMax) : 0, n, d_sizes, st) !=
        hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    goto done;
  rc = 0;
done:
  if (rc && chucky_last_error()[0] == 0)
    set_error("chucky_gpu_compress failed");
  (void)hipFree(seg);
  (void)hipFree(scratch);
  (void)hipFree(queue);
  (void)hipStreamDestroy(st);
  return rc;
}

int chucky_gpu_transpose(void* d_dst, const void* d_src, uint64_t n, uint8_t bpe, uint64_t i_offset,
                         uint8_t lifted_rank, const uint64_t* lifted_shape,
                         const int64_t* lifted_strides)
{
  if (lifted_rank == 0 || lifted_rank > 64 || (bpe != 1 && bpe != 2 && bpe != 4 && bpe != 8))
    return 1;
  LiftedSpace sp{};
  sp.rank = lifted_rank;
  for (int d = 0; d < lifted_rank; ++d) {
    sp.shape[d] = lifted_shape[d];
    sp.strides[d] = lifted_strides[d];
  }
  if (launch_transpose(d_dst, d_src, n, bpe, i_offset, sp, nullptr) != hipSuccess)
    return 1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

} // extern "C"
