// In-memory and discard shard sinks. The memory sink keeps every shard's bytes
// (the test fake of tests/test_shard_sink.c:6-109 in the reference); the
// discard sink only counts bytes (bench/bench_util.c:27-70).
#include "chucky_amd.h"

#include <atomic>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <vector>

namespace {

struct MemWriter
{
  shard_writer base;
  struct chucky_mem_sink_impl* owner;
  uint8_t level;
  uint64_t shard_index;
  std::vector<uint8_t> bytes;
  int finalized = 0;
};

} // namespace

// A zero-copy write held back until its level's fence is waited for (or the
// sink is drained): the caller's buffer is read only then, as an
// asynchronous I/O queue would.
struct PendingWrite
{
  MemWriter* w;
  uint64_t offset;
  const uint8_t* src;
  size_t n;
  uint64_t seq;
};

struct chucky_mem_sink
{
  shard_sink base;
  std::mutex mu;
  std::deque<MemWriter> writers; // stable addresses
  std::map<std::pair<uint8_t, uint64_t>, MemWriter*> by_key;
  std::map<uint8_t, std::vector<uint64_t>> append_sizes;
  int fail_open = 0, fail_write = 0;
  int deferred = 0;
  std::map<uint8_t, uint64_t> level_seq;             // zero-copy writes issued per level
  std::map<uint8_t, std::deque<PendingWrite>> pending; // not yet copied, in issue order
};

namespace {

chucky_mem_sink* mem_of(shard_sink* s) { return reinterpret_cast<chucky_mem_sink*>(s); }

void store(MemWriter* w, uint64_t offset, const void* beg, size_t n)
{
  if (w->bytes.size() < offset + n)
    w->bytes.resize(offset + n);
  if (n)
    std::memcpy(w->bytes.data() + offset, beg, n);
}

int mem_write(shard_writer* self, uint64_t offset, const void* beg, const void* end)
{
  MemWriter* w = reinterpret_cast<MemWriter*>(self);
  chucky_mem_sink* s = reinterpret_cast<chucky_mem_sink*>(w->owner);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fail_write)
    return 1;
  store(w, offset, beg, (const uint8_t*)end - (const uint8_t*)beg);
  return 0;
}

int mem_write_direct(shard_writer* self, uint64_t offset, const void* beg, const void* end)
{
  MemWriter* w = reinterpret_cast<MemWriter*>(self);
  chucky_mem_sink* s = reinterpret_cast<chucky_mem_sink*>(w->owner);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fail_write)
    return 1;
  const size_t n = (const uint8_t*)end - (const uint8_t*)beg;
  s->pending[w->level].push_back({ w, offset, (const uint8_t*)beg, n, ++s->level_seq[w->level] });
  return 0;
}

// Copies the level's held-back writes with sequence <= upto (mu held).
void apply_pending(chucky_mem_sink* s, uint8_t level, uint64_t upto)
{
  auto& q = s->pending[level];
  while (!q.empty() && q.front().seq <= upto) {
    const PendingWrite& p = q.front();
    store(p.w, p.offset, p.src, p.n);
    q.pop_front();
  }
}

io_event mem_record_fence(shard_sink* self, uint8_t level)
{
  chucky_mem_sink* s = reinterpret_cast<chucky_mem_sink*>(self);
  std::lock_guard<std::mutex> lk(s->mu);
  return io_event{ s->level_seq[level] };
}

void mem_wait_fence(shard_sink* self, uint8_t level, io_event ev)
{
  chucky_mem_sink* s = reinterpret_cast<chucky_mem_sink*>(self);
  std::lock_guard<std::mutex> lk(s->mu);
  apply_pending(s, level, ev.seq);
}

int mem_finalize(shard_writer* self)
{
  MemWriter* w = reinterpret_cast<MemWriter*>(self);
  chucky_mem_sink* s = reinterpret_cast<chucky_mem_sink*>(w->owner);
  std::lock_guard<std::mutex> lk(s->mu);
  w->finalized = 1;
  return 0;
}

shard_writer* mem_open(shard_sink* self, uint8_t level, uint64_t shard_index)
{
  chucky_mem_sink* s = mem_of(self);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fail_open)
    return nullptr;
  auto key = std::make_pair(level, shard_index);
  auto it = s->by_key.find(key);
  if (it != s->by_key.end()) {
    it->second->finalized = 0;
    return &it->second->base;
  }
  s->writers.emplace_back();
  MemWriter& w = s->writers.back();
  w.base.write = mem_write;
  w.base.write_direct = s->deferred ? mem_write_direct : nullptr;
  w.base.finalize = mem_finalize;
  w.owner = reinterpret_cast<struct chucky_mem_sink_impl*>(s);
  w.level = level;
  w.shard_index = shard_index;
  s->by_key[key] = &w;
  return &w.base;
}

int mem_update_append(shard_sink* self, uint8_t level, uint8_t n, const uint64_t* sizes)
{
  chucky_mem_sink* s = mem_of(self);
  std::lock_guard<std::mutex> lk(s->mu);
  s->append_sizes[level].assign(sizes, sizes + n);
  return 0;
}

} // namespace

extern "C" {

struct chucky_mem_sink* chucky_mem_sink_create(void)
{
  chucky_mem_sink* s = new chucky_mem_sink();
  s->base.open = mem_open;
  s->base.update_append = mem_update_append;
  s->base.record_fence = nullptr;
  s->base.wait_fence = nullptr;
  return s;
}

void chucky_mem_sink_destroy(struct chucky_mem_sink* s) { delete s; }

struct shard_sink* chucky_mem_sink_as_shard_sink(struct chucky_mem_sink* s) { return &s->base; }

size_t chucky_mem_sink_count(const struct chucky_mem_sink* s) { return s->writers.size(); }

int chucky_mem_sink_get(const struct chucky_mem_sink* s, size_t i, uint8_t* level,
                        uint64_t* shard_index, const uint8_t** data, size_t* size, int* finalized)
{
  if (i >= s->writers.size())
    return 1;
  const MemWriter& w = s->writers[i];
  if (level)
    *level = w.level;
  if (shard_index)
    *shard_index = w.shard_index;
  if (data)
    *data = w.bytes.data();
  if (size)
    *size = w.bytes.size();
  if (finalized)
    *finalized = w.finalized;
  return 0;
}

int chucky_mem_sink_append_sizes(const struct chucky_mem_sink* s, uint8_t level, uint64_t* out,
                                 uint8_t* n)
{
  auto it = s->append_sizes.find(level);
  if (it == s->append_sizes.end())
    return 1;
  for (size_t i = 0; i < it->second.size(); ++i)
    out[i] = it->second[i];
  *n = (uint8_t)it->second.size();
  return 0;
}

void chucky_mem_sink_inject_failure(struct chucky_mem_sink* s, int fail_open, int fail_write)
{
  std::lock_guard<std::mutex> lk(s->mu);
  s->fail_open = fail_open;
  s->fail_write = fail_write;
}

void chucky_mem_sink_set_deferred(struct chucky_mem_sink* s, int deferred)
{
  std::lock_guard<std::mutex> lk(s->mu);
  s->deferred = deferred;
  s->base.record_fence = deferred ? mem_record_fence : nullptr;
  s->base.wait_fence = deferred ? mem_wait_fence : nullptr;
  for (auto& w : s->writers)
    w.base.write_direct = deferred ? mem_write_direct : nullptr;
}

void chucky_mem_sink_drain(struct chucky_mem_sink* s)
{
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto& kv : s->pending)
    apply_pending(s, kv.first, ~0ull);
}

} // extern "C"

// ---- discard sink ----

struct chucky_discard_sink
{
  shard_sink base;
  shard_writer writer;
  std::atomic<uint64_t> bytes{ 0 };
  std::atomic<uint64_t> finalized{ 0 };
};

namespace {
chucky_discard_sink* discard_of_writer(shard_writer* w)
{
  return reinterpret_cast<chucky_discard_sink*>(reinterpret_cast<char*>(w) -
                                                offsetof(chucky_discard_sink, writer));
}
int discard_write(shard_writer* self, uint64_t, const void* beg, const void* end)
{
  discard_of_writer(self)->bytes += (uint64_t)((const uint8_t*)end - (const uint8_t*)beg);
  return 0;
}
int discard_finalize(shard_writer* self)
{
  discard_of_writer(self)->finalized += 1;
  return 0;
}
shard_writer* discard_open(shard_sink* self, uint8_t, uint64_t)
{
  return &reinterpret_cast<chucky_discard_sink*>(self)->writer;
}
} // namespace

extern "C" {

struct chucky_discard_sink* chucky_discard_sink_create(void)
{
  chucky_discard_sink* s = new chucky_discard_sink();
  s->base.open = discard_open;
  s->base.update_append = nullptr;
  s->base.record_fence = nullptr;
  s->base.wait_fence = nullptr;
  s->writer.write = discard_write;
  s->writer.write_direct = discard_write;
  s->writer.finalize = discard_finalize;
  return s;
}
void chucky_discard_sink_destroy(struct chucky_discard_sink* s) { delete s; }
struct shard_sink* chucky_discard_sink_as_shard_sink(struct chucky_discard_sink* s) { return &s->base; }
uint64_t chucky_discard_sink_bytes(const struct chucky_discard_sink* s) { return s->bytes.load(); }
uint64_t chucky_discard_sink_finalized(const struct chucky_discard_sink* s) { return s->finalized.load(); }

} // extern "C"
