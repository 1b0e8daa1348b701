// Host-memory appends pass through pinned staging buffers before the DMA
// engines move them to the device. One thread copies at about a third of the
// host link's rate, so the staging copy is split over a few threads.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>

namespace chucky {

class CopyPool
{
public:
  // threads: helper threads besides the caller (0 = copy on the caller only)
  explicit CopyPool(unsigned threads);
  ~CopyPool();
  CopyPool(const CopyPool&) = delete;
  CopyPool& operator=(const CopyPool&) = delete;

  // dst[0, n) = src[0, n); returns once every piece is copied.
  void copy(void* dst, const void* src, size_t n);
  // The same copy, begun on the helper threads; the caller may do other
  // work and must call wait() before touching dst or src again.
  void start(void* dst, const void* src, size_t n);
  // Joins the copy begun by start() and returns once it is complete.
  void wait();
  unsigned threads() const { return (unsigned)workers_.size(); }

private:
  void run();
  void work();

  // bytes a thread takes at a time: about two pieces per thread, so that
  // every thread has work, within [kMinPiece, kMaxPiece]
  static constexpr size_t kMinPiece = 256u << 10, kMaxPiece = 2u << 20;
  size_t piece_ = kMaxPiece;

  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable start_, done_;
  uint64_t generation_ = 0;
  unsigned busy_ = 0; // helpers still on the current job
  bool stop_ = false;
  uint8_t* dst_ = nullptr;
  const uint8_t* src_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{ 0 };
  bool pending_ = false; // a copy begun by start() not yet joined
};

// Helper threads for staging copies: the CHUCKY_COPY_THREADS environment
// variable if set, else half the hardware threads, at most 7.
unsigned default_copy_threads();

} // namespace chucky
