// Pinned host memory on the NUMA node of the GPU.
//
// Compressed shards leave the device by DMA (or by kernel stores) into pinned
// host buffers; when such a buffer sits on the other CPU socket the traffic
// crosses the inter-socket link and the device-to-host rate drops to about
// half of the PCIe link's. The buffers are therefore allocated under a
// preferred-node memory policy naming the GPU's node (read from the PCI
// device's sysfs entry); where that node is unknown the default policy applies.
#include "stream_gpu.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <sys/syscall.h>
#include <unistd.h>

namespace chucky {

namespace {

constexpr int kMpolDefault = 0;
constexpr int kMpolPreferred = 1;

int gpu_numa_node(int device)
{
  static std::mutex mu;
  static int cache[64];
  static bool init = false;
  std::lock_guard<std::mutex> lk(mu);
  if (!init) {
    for (int& c : cache)
      c = -2;
    init = true;
  }
  if (device < 0 || device >= 64)
    return -1;
  if (cache[device] != -2)
    return cache[device];
  int node = -1;
  char bus[64] = { 0 };
  if (hipDeviceGetPCIBusId(bus, sizeof bus, device) == hipSuccess) {
    for (char* c = bus; *c; ++c)
      if (*c >= 'A' && *c <= 'F')
        *c = (char)(*c - 'A' + 'a');
    char path[160];
    std::snprintf(path, sizeof path, "/sys/bus/pci/devices/%s/numa_node", bus);
    if (FILE* f = std::fopen(path, "r")) {
      if (std::fscanf(f, "%d", &node) != 1)
        node = -1;
      std::fclose(f);
    }
  }
  cache[device] = node;
  return node;
}


// Pinned buffers of finished streams are kept for the next streams of the
// process (a serving loop creates a stream per request, and pinning GBs of
// host memory takes a large part of a stream's set-up): a request takes the
// smallest free buffer of its device between its size and twice that.
// Contents are not cleared (every user writes a buffer before reading it).
struct PinnedPool
{
  std::mutex mu;
  std::map<void*, std::pair<size_t, int>> live; // every buffer made here: size, device
  std::multimap<size_t, std::pair<void*, int>> free_by_size;
  size_t free_bytes = 0;
  static constexpr size_t kMaxFree = (size_t)16 << 30;
};

PinnedPool& pool()
{
  static PinnedPool* p = new PinnedPool(); // kept past static destruction
  return *p;
}

hipError_t pin_new(void** p, size_t bytes, int device)
{
  const int node = gpu_numa_node(device);
  if (node < 0 || node >= 64)
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
  const unsigned long mask = 1ul << node;
  if (syscall(SYS_set_mempolicy, kMpolPreferred, &mask, 64ul) != 0)
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
  const hipError_t e = hipHostMalloc(p, bytes, hipHostMallocNumaUser);
  (void)syscall(SYS_set_mempolicy, kMpolDefault, nullptr, 0ul);
  return e;
}

} // namespace

hipError_t host_alloc_near_gpu(void** p, size_t bytes, int device)
{
  PinnedPool& P = pool();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    for (auto it = P.free_by_size.lower_bound(bytes); it != P.free_by_size.end() && it->first <= 2 * bytes;
         ++it)
      if (it->second.second == device) {
        *p = it->second.first;
        P.free_bytes -= it->first;
        P.free_by_size.erase(it);
        return hipSuccess;
      }
  }
  const hipError_t e = pin_new(p, bytes, device);
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lk(P.mu);
    P.live[*p] = { bytes, device };
  }
  return e;
}

hipError_t host_alloc_unpooled(void** p, size_t bytes, int device) { return pin_new(p, bytes, device); }

void host_release(void* p)
{
  if (!p)
    return;
  PinnedPool& P = pool();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it != P.live.end() && P.free_bytes + it->second.first <= PinnedPool::kMaxFree) {
      P.free_by_size.emplace(it->second.first, std::make_pair(p, it->second.second));
      P.free_bytes += it->second.first;
      return;
    }
    if (it != P.live.end())
      P.live.erase(it);
  }
  (void)hipHostFree(p);
}

} // namespace chucky
