/* Chunk codec selection (reference: src/types.codec.h:10-38). */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum compression_codec
{
  CODEC_NONE,
  CODEC_LZ4_NON_STANDARD, /* raw LZ4 block per chunk (no frame) */
  CODEC_ZSTD,             /* one zstd frame per chunk */
  CODEC_BLOSC_LZ4,        /* CPU-only in the reference; rejected here */
  CODEC_BLOSC_ZSTD,       /* CPU-only in the reference; rejected here */
};

enum codec_shuffle
{
  CODEC_SHUFFLE_NONE = 0,
  CODEC_SHUFFLE_BYTE = 1,
  CODEC_SHUFFLE_BIT = 2,
};

struct codec_config
{
  enum compression_codec id;
  uint8_t level;              /* LZ4: 1..12 (0 rejected). ZSTD: any. */
  enum codec_shuffle shuffle; /* blosc only */
};

/* Replaces src/types.codec.c:3-7. */
int codec_is_blosc(enum compression_codec c);

/* Replaces src/types.codec.c:9-13: NONE, LZ4 and ZSTD run on the GPU. */
int codec_is_gpu_supported(enum compression_codec c);

#ifdef __cplusplus
}
#endif
