#include "host_copy.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace chucky {

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
    const size_t off = next_.fetch_add(kPiece);
    if (off >= n_)
      return;
    std::memcpy(dst_ + off, src_ + off, std::min(kPiece, n_ - off));
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
  if (workers_.empty() || n < 2 * kPiece) {
    std::memcpy(dst, src, n);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
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
