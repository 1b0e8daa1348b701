/* Streaming writer and shard sink vtables (reference: src/writer.h:6-98).
 * These vtables are the plugin boundary: layouts and argument meaning are
 * identical so that a sink written for the reference plugs in unchanged. */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct io_event
{
  uint64_t seq;
};

struct slice
{
  const void* beg;
  const void* end;
};

enum writer_error_code
{
  writer_error_ok = 0,
  writer_error_fail = 1,
  writer_error_finished = 2, /* capacity reached or flush already called */
};

struct writer_result
{
  int error;
  struct slice rest; /* unconsumed input */
};

struct writer
{
  struct writer_result (*append)(struct writer* self, struct slice data);
  struct writer_result (*flush)(struct writer* self); /* terminal */
};

struct shard_writer
{
  int (*write)(struct shard_writer* self, uint64_t offset, const void* beg, const void* end);
  /* Zero-copy: buffer stays valid until the next fence retires. NULL = use write. */
  int (*write_direct)(struct shard_writer* self, uint64_t offset, const void* beg,
                      const void* end);
  int (*finalize)(struct shard_writer* self);
};

struct shard_sink
{
  struct shard_writer* (*open)(struct shard_sink* self, uint8_t level, uint64_t shard_index);
  /* Optional: publish append-dim extents (n_append values). */
  int (*update_append)(struct shard_sink* self, uint8_t level, uint8_t n_append,
                       const uint64_t* append_sizes);
  /* Optional IO fences. */
  struct io_event (*record_fence)(struct shard_sink* self, uint8_t level);
  void (*wait_fence)(struct shard_sink* self, uint8_t level, struct io_event ev);
};

struct writer_result writer_ok(void);
struct writer_result writer_error(void);
struct writer_result writer_error_at(const void* beg, const void* end);
struct writer_result writer_finished_at(const void* beg, const void* end);
struct writer_result writer_append(struct writer* w, struct slice data);
struct writer_result writer_flush(struct writer* w);
/* Retries a stalled append with 1..64 ms back-off, at most 10 stalls. */
struct writer_result writer_append_wait(struct writer* w, struct slice data);

#ifdef __cplusplus
}
#endif
