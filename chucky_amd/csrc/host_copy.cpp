#include "host_copy.hpp"

#include <emmintrin.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace chucky {

namespace {

// Copy into pinned staging with non-temporal stores: the staging bytes are
// read next by a DMA engine, not by this CPU, so they need not be cached, and
// streaming stores skip the read-for-ownership that ordinary stores make,
// taking host-memory traffic from three times the bytes to two. The fence
// drains the write-combining buffers before the piece counts as copied.
void stream_copy(uint8_t* dst, const uint8_t* src, size_t n)
{
  constexpr size_t kMinStream = 64u << 10;
  if (n < kMinStream) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t head = (16 - ((uintptr_t)dst & 15)) & 15;
  std::memcpy(dst, src, head);
  dst += head;
  src += head;
  n -= head;
  const size_t body = n & ~(size_t)63;
  for (size_t i = 0; i < body; i += 64) {
    const __m128i a = _mm_loadu_si128((const __m128i*)(src + i));
    const __m128i b = _mm_loadu_si128((const __m128i*)(src + i + 16));
    const __m128i c = _mm_loadu_si128((const __m128i*)(src + i + 32));
    const __m128i d = _mm_loadu_si128((const __m128i*)(src + i + 48));
    _mm_stream_si128((__m128i*)(dst + i), a);
    _mm_stream_si128((__m128i*)(dst + i + 16), b);
    _mm_stream_si128((__m128i*)(dst + i + 32), c);
    _mm_stream_si128((__m128i*)(dst + i + 48), d);
  }
  std::memcpy(dst + body, src + body, n - body);
  _mm_sfence();
}

} // namespace

CopyPool::CopyPool(unsigned threads)
{
  for (unsigned i = 0; i < threads; ++i)
    workers_.emplace_back([this] { run(); });
}

CopyPool::~CopyPool()
{
  wait();
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  start_.notify_all();
  for (auto& t : workers_)
    t.join();
}

void CopyPool::work()
{
  for (;;) {
    const size_t off = next_.fetch_add(piece_);
    if (off >= n_)
      return;
    stream_copy(dst_ + off, src_ + off, std::min(piece_, n_ - off));
  }
}

void CopyPool::run()
{
  uint64_t seen = 0;
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    start_.wait(lk, [&] { return stop_ || generation_ != seen; });
    if (stop_)
      return;
    seen = generation_;
    lk.unlock();
    work();
    lk.lock();
    if (--busy_ == 0)
      done_.notify_all();
  }
}

void CopyPool::start(void* dst, const void* src, size_t n)
{
  wait();
  if (workers_.empty() || n < 2 * kMinPiece) {
    stream_copy((uint8_t*)dst, (const uint8_t*)src, n);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    const size_t per = n / (2 * (workers_.size() + 1));
    piece_ = std::min(kMaxPiece, std::max(kMinPiece, per & ~(size_t)4095));
    dst_ = (uint8_t*)dst;
    src_ = (const uint8_t*)src;
    n_ = n;
    next_.store(0);
    busy_ = (unsigned)workers_.size();
    ++generation_;
  }
  pending_ = true;
  start_.notify_all();
}

void CopyPool::wait()
{
  if (!pending_)
    return;
  work();
  std::unique_lock<std::mutex> lk(mu_);
  done_.wait(lk, [&] { return busy_ == 0; });
  pending_ = false;
}

void CopyPool::copy(void* dst, const void* src, size_t n)
{
  start(dst, src, n);
  wait();
}

unsigned default_copy_threads()
{
  if (const char* e = std::getenv("CHUCKY_COPY_THREADS"))
    return (unsigned)std::max(0, std::atoi(e));
  const unsigned hw = std::thread::hardware_concurrency();
  return std::min(7u, std::max(1u, hw / 2) - 1);
}

} // namespace chucky
