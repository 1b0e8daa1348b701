// Multi-array writer: one GPU tile stream per array, resident for the
// writer's lifetime, behind the reference's multiarray_writer vtable
// (src/multiarray/stream.c:552-726 update, :866-887 flush).
//
// The reference shares one chunk pool and LOD buffer between the arrays, so
// a switch flushes the departing array's partial batch and clears the shared
// buffers. With 288 GB of HBM each array here keeps its own pool, LOD state
// and compressed slots, so a switch only checks the epoch boundary; each
// array's batches are written when they fill or at flush, and its shards hold
// the same bytes either way.
#include "multiarray.gpu.h"
#include "stream.gpu.h"

#include <new>
#include <vector>

struct multiarray_tile_stream_gpu
{
  multiarray_writer writer; // first member: the vtable's self points here
  std::vector<tile_stream_gpu*> arrays;
  int active = -1;
};

namespace {

multiarray_writer_result result(int error, slice rest) { return multiarray_writer_result{ error, rest }; }

int map_error(int e)
{
  switch (e) {
    case writer_error_ok: return multiarray_writer_ok;
    case writer_error_finished: return multiarray_writer_finished;
    default: return multiarray_writer_fail;
  }
}

multiarray_writer_result update_impl(multiarray_writer* self, int array_index, slice data)
{
  auto* ms = reinterpret_cast<multiarray_tile_stream_gpu*>(self);
  // src/multiarray/stream.c:626-627: an unknown array consumes nothing
  if (array_index < 0 || array_index >= (int)ms->arrays.size())
    return result(multiarray_writer_fail, data);
  if (array_index != ms->active) {
    // src/multiarray/stream.c:555-558: the departing array must sit on an
    // epoch boundary
    if (ms->active >= 0) {
      tile_stream_gpu* dep = ms->arrays[ms->active];
      const tile_stream_layout* lay = tile_stream_gpu_layout(dep);
      if (!lay || (lay->epoch_elements && tile_stream_gpu_cursor(dep) % lay->epoch_elements))
        return result(multiarray_writer_not_flushable, data);
    }
    ms->active = array_index;
  }
  const writer_result r = writer_append(tile_stream_gpu_writer(ms->arrays[array_index]), data);
  return result(map_error(r.error), r.rest);
}

multiarray_writer_result flush_impl(multiarray_writer* self)
{
  auto* ms = reinterpret_cast<multiarray_tile_stream_gpu*>(self);
  // src/multiarray/stream.c:866-887: every array's partial epoch, batches,
  // dim-0 LOD drain and shards; fail if any of them fails
  int error = multiarray_writer_ok;
  for (tile_stream_gpu* s : ms->arrays) {
    const writer_result r = writer_flush(tile_stream_gpu_writer(s));
    if (r.error != writer_error_ok && r.error != writer_error_finished)
      error = multiarray_writer_fail;
  }
  return result(error, slice{ nullptr, nullptr });
}

void add_metric(stream_metric& a, const stream_metric& b)
{
  if (!a.name)
    a.name = b.name;
  a.ms += b.ms;
  if (b.best_ms < a.best_ms)
    a.best_ms = b.best_ms;
  a.input_bytes += b.input_bytes;
  a.output_bytes += b.output_bytes;
  a.count += b.count;
}

} // namespace

extern "C" {

multiarray_tile_stream_gpu* multiarray_tile_stream_gpu_create(int n_arrays,
                                                              const tile_stream_configuration configs[],
                                                              shard_sink* sinks[], int enable_metrics)
{
  (void)enable_metrics;
  if (n_arrays <= 0 || !configs || !sinks)
    return nullptr;
  auto* ms = new (std::nothrow) multiarray_tile_stream_gpu();
  if (!ms)
    return nullptr;
  ms->writer.update = update_impl;
  ms->writer.flush = flush_impl;
  for (int i = 0; i < n_arrays; ++i) {
    tile_stream_gpu* s = tile_stream_gpu_create(&configs[i], sinks[i]);
    if (!s) {
      multiarray_tile_stream_gpu_destroy(ms);
      return nullptr;
    }
    ms->arrays.push_back(s);
  }
  return ms;
}

void multiarray_tile_stream_gpu_destroy(multiarray_tile_stream_gpu* ms)
{
  if (!ms)
    return;
  for (tile_stream_gpu* s : ms->arrays)
    tile_stream_gpu_destroy(s);
  delete ms;
}

multiarray_writer* multiarray_tile_stream_gpu_writer(multiarray_tile_stream_gpu* ms)
{
  return ms ? &ms->writer : nullptr;
}

stream_metrics multiarray_tile_stream_gpu_get_metrics(const multiarray_tile_stream_gpu* ms)
{
  stream_metrics m{};
  stream_metric* out = reinterpret_cast<stream_metric*>(&m);
  constexpr int kN = sizeof(stream_metrics) / sizeof(stream_metric);
  for (int k = 0; k < kN; ++k)
    out[k].best_ms = 1e30f;
  if (!ms)
    return m;
  for (tile_stream_gpu* s : ms->arrays) {
    const stream_metrics a = tile_stream_gpu_get_metrics(s);
    const stream_metric* in = reinterpret_cast<const stream_metric*>(&a);
    for (int k = 0; k < kN; ++k)
      add_metric(out[k], in[k]);
  }
  return m;
}

} // extern "C"
