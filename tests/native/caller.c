/* The unchanged-caller pattern of the reference (bench/bench_util.c:644-692,
 * tests/test_stream.c) written against include/ and linked to
 * libchucky_amd.so, as a C11 program:
 *
 *   caller estimate          memory estimate + chunk-size advice (host only)
 *   caller run STORE [zstd|lz4|none]
 *                            writes a t,y,x u16 acquisition of 10 frames of
 *                            64x96 (chunks 4x32x32, shards 1x2x3 chunks, a
 *                            partial last epoch) through zarr_fs_sink, from
 *                            host memory in three appends
 *
 * The frame's inner dims are chunk multiples: the reference reads appended
 * elements in the chunk-padded (lifted) shape of the inner dims
 * (src/stream/config.c:97-103), so unpadded frames of other sizes would land
 * shifted, in the reference and here alike.
 * Pixel (t, y, x) = (t * 7919 + y * 131 + x * 17) mod 65536. Exit status 0 on
 * success; failures print the failing call. */
#include "stream.gpu.h"
#include "zarr_fs_sink.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { T = 10, Y = 64, X = 96 };

static void set_dims(struct dimension* dims)
{
  const uint64_t sizes[3] = { T, Y, X };
  const uint64_t chunks[3] = { 4, 32, 32 };
  dims_create(dims, "tyx", sizes);
  dims_set_chunk_sizes(dims, 3, chunks);
  dims[0].chunks_per_shard = 1;
  dims[1].chunks_per_shard = 2;
  dims[2].chunks_per_shard = 3;
}

static struct tile_stream_configuration make_config(struct dimension* dims, enum compression_codec codec)
{
  struct tile_stream_configuration cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.buffer_capacity_bytes = 64 << 10;
  cfg.dtype = dtype_u16;
  cfg.rank = 3;
  cfg.dimensions = dims;
  cfg.codec.id = codec;
  cfg.codec.level = 1;
  cfg.metadata_update_interval_s = 1e9f;
  return cfg;
}

static int estimate(void)
{
  struct dimension dims[3];
  set_dims(dims);
  struct tile_stream_configuration cfg = make_config(dims, CODEC_ZSTD);
  struct tile_stream_memory_info info;
  if (tile_stream_gpu_memory_estimate(&cfg, &info)) {
    fprintf(stderr, "tile_stream_gpu_memory_estimate failed\n");
    return 1;
  }
  printf("chunks_per_epoch %llu total_chunks %llu nlod %d epochs_per_batch %u device_bytes %llu\n",
         (unsigned long long)info.chunks_per_epoch, (unsigned long long)info.total_chunks, info.nlod,
         (unsigned)info.epochs_per_batch, (unsigned long long)info.device_bytes);
  /* advice: halve a 1 MiB target until the estimate fits the budget */
  const uint8_t ratios[3] = { 0, 1, 1 };
  const size_t budget = (size_t)512 << 20;
  if (tile_stream_gpu_advise_chunk_sizes(&cfg, 1 << 20, ratios, budget)) {
    fprintf(stderr, "tile_stream_gpu_advise_chunk_sizes failed\n");
    return 1;
  }
  if (tile_stream_gpu_memory_estimate(&cfg, &info)) {
    fprintf(stderr, "estimate of the advised configuration failed\n");
    return 1;
  }
  printf("advised chunks %llu %llu %llu device_bytes %llu budget %llu\n", (unsigned long long)dims[0].chunk_size,
         (unsigned long long)dims[1].chunk_size, (unsigned long long)dims[2].chunk_size,
         (unsigned long long)info.device_bytes, (unsigned long long)budget);
  return 0;
}

static int run(const char* store, enum compression_codec codec)
{
  struct dimension dims[3];
  set_dims(dims);
  uint16_t* frames = malloc(sizeof(uint16_t) * T * Y * X);
  if (!frames)
    return 1;
  for (int t = 0; t < T; ++t)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x)
        frames[((size_t)t * Y + y) * X + x] = (uint16_t)((t * 7919u + y * 131u + x * 17u) & 0xFFFFu);

  struct zarr_config zc;
  memset(&zc, 0, sizeof zc);
  zc.store_path = store;
  zc.array_name = "0";
  zc.data_type = dtype_u16;
  zc.rank = 3;
  zc.dimensions = dims;
  zc.codec.id = codec;
  zc.codec.level = 1;
  struct zarr_fs_sink* zs = zarr_fs_sink_create(&zc);
  if (!zs) {
    fprintf(stderr, "zarr_fs_sink_create failed\n");
    free(frames);
    return 1;
  }
  struct tile_stream_configuration cfg = make_config(dims, codec);
  struct tile_stream_gpu* s = tile_stream_gpu_create(&cfg, zarr_fs_sink_as_shard_sink(zs));
  if (!s) {
    fprintf(stderr, "tile_stream_gpu_create failed\n");
    zarr_fs_sink_destroy(zs);
    free(frames);
    return 1;
  }
  struct writer* w = tile_stream_gpu_writer(s);
  const char* p = (const char*)frames;
  const size_t total = sizeof(uint16_t) * T * Y * X, cut[4] = { 0, total / 3 + 6, 2 * total / 3, total };
  int rc = 0;
  for (int k = 0; k < 3 && !rc; ++k) {
    struct slice in = { p + cut[k], p + cut[k + 1] };
    struct writer_result r = writer_append(w, in);
    if (r.error != writer_error_ok) {
      fprintf(stderr, "writer_append failed (%d)\n", r.error);
      rc = 1;
    }
  }
  if (!rc && writer_flush(w).error != writer_error_ok) {
    fprintf(stderr, "writer_flush failed\n");
    rc = 1;
  }
  if (!rc && tile_stream_gpu_cursor(s) != (uint64_t)T * Y * X) {
    fprintf(stderr, "cursor %llu\n", (unsigned long long)tile_stream_gpu_cursor(s));
    rc = 1;
  }
  tile_stream_gpu_destroy(s);
  zarr_fs_sink_flush(zs);
  zarr_fs_sink_destroy(zs);
  free(frames);
  if (!rc)
    printf("wrote %s\n", store);
  return rc;
}

int main(int argc, char** argv)
{
  if (argc >= 2 && strcmp(argv[1], "estimate") == 0)
    return estimate();
  if (argc >= 3 && strcmp(argv[1], "run") == 0) {
    enum compression_codec c = CODEC_ZSTD;
    if (argc >= 4 && strcmp(argv[3], "lz4") == 0)
      c = CODEC_LZ4_NON_STANDARD;
    else if (argc >= 4 && strcmp(argv[3], "none") == 0)
      c = CODEC_NONE;
    return run(argv[2], c);
  }
  fprintf(stderr, "usage: caller estimate | caller run STORE [zstd|lz4|none]\n");
  return 2;
}
