// Serving loop: admission of concurrent write streams against an HBM budget.
//
// A scheduler thread admits queued requests in arrival order when a running
// slot and enough of the device-memory budget (tile_stream_gpu_memory_estimate)
// are free. Every admitted request gets its own host thread and its own GPU
// stream object (its own HIP queues, pools, LOD accumulators), appends its
// input epoch by epoch and flushes; the kernels of different requests
// interleave on the device. When a request finishes its memory returns to the
// budget and the next queued request is admitted. Reports carry the request's
// append -> shard-written latency percentiles.
#include "chucky_amd.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace chucky {
void set_error(const std::string& msg);
}

namespace {

using Clock = std::chrono::steady_clock;

struct Request
{
  uint64_t id;
  tile_stream_configuration config;
  shard_sink* sink;
  const uint8_t* data;
  size_t nbytes;
  int device;
  size_t estimate;
  Clock::time_point submitted;
};

double pct(std::vector<double> v, double q)
{
  if (v.empty())
    return 0;
  std::sort(v.begin(), v.end());
  return v[std::min(v.size() - 1, (size_t)(q * v.size()))];
}

} // namespace

struct chucky_server
{
  size_t budget;
  int max_concurrent;
  int pipeline;
  int hip_device = 0;
  Clock::time_point t0 = Clock::now();
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> queue;
  std::vector<std::thread> workers;
  std::vector<chucky_request_report> reports;
  std::vector<double> latencies;
  uint64_t next_id = 0;
  int running = 0, peak_running = 0;
  size_t in_use = 0, peak_bytes = 0;
  uint64_t submitted = 0, finished = 0;
  int failures = 0;
  std::string error; // first failure, reported by wait_all
  bool stop = false;
  std::thread scheduler;

  double since(Clock::time_point t) const { return std::chrono::duration<double>(t - t0).count(); }

  void run_request(Request r)
  {
    (void)hipSetDevice(hip_device);
    chucky_request_report rep{};
    rep.id = r.id;
    rep.submit_s = since(r.submitted);
    rep.start_s = since(Clock::now());
    rep.input_bytes = r.nbytes;
    tile_stream_gpu_options o;
    tile_stream_gpu_options_default(&o);
    o.pipeline = pipeline;
    o.device = hip_device;
    int status = 1;
    std::vector<double> lat;
    tile_stream_gpu* s = tile_stream_gpu_create_ex(&r.config, r.sink, &o);
    if (!s)
      chucky::set_error(std::string("create failed: ") + chucky_last_error());
    if (s) {
      status = 0;
      const size_t epoch_bytes =
        tile_stream_gpu_layout(s)->epoch_elements * dtype_bpe(r.config.dtype);
      writer* w = tile_stream_gpu_writer(s);
      for (size_t off = 0; off < r.nbytes && status == 0; off += epoch_bytes) {
        const size_t n = std::min(epoch_bytes, r.nbytes - off);
        const writer_result res =
          r.device ? tile_stream_gpu_append_device(s, r.data + off, n)
                   : writer_append(w, slice{ r.data + off, r.data + off + n });
        if (res.error == writer_error_finished)
          break;
        status = res.error;
        if (status)
          chucky::set_error(std::string("append failed at byte ") + std::to_string(off) + ": " +
                            chucky_last_error());
      }
      if (status == 0 && writer_flush(w).error) {
        status = 1;
        chucky::set_error(std::string("flush failed: ") + chucky_last_error());
      }
      lat.resize(1 << 16);
      lat.resize(tile_stream_gpu_epoch_latencies(s, lat.data(), lat.size()));
      tile_stream_gpu_destroy(s);
    }
    rep.status = status;
    const std::string why = status ? std::string(chucky_last_error()) : std::string();
    rep.end_s = since(Clock::now());
    rep.epochs = lat.size();
    rep.p50_latency_s = pct(lat, 0.50);
    rep.p99_latency_s = pct(lat, 0.99);
    rep.max_latency_s = lat.empty() ? 0 : *std::max_element(lat.begin(), lat.end());
    std::lock_guard<std::mutex> lk(mu);
    reports.push_back(rep);
    latencies.insert(latencies.end(), lat.begin(), lat.end());
    if (status && error.empty())
      error = "request " + std::to_string(r.id) + ": " + (why.empty() ? "failed" : why);
    failures += status != 0;
    --running;
    in_use -= r.estimate;
    ++finished;
    cv.notify_all();
  }

  void schedule()
  {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv.wait(lk, [&] {
        if (stop && queue.empty())
          return true;
        if (queue.empty() || running >= max_concurrent)
          return false;
        // admit in arrival order; a request larger than the whole budget
        // runs alone
        const size_t need = queue.front().estimate;
        return in_use + need <= budget || running == 0;
      });
      if (queue.empty())
        return;
      Request r = queue.front();
      queue.pop_front();
      ++running;
      in_use += r.estimate;
      peak_running = std::max(peak_running, running);
      peak_bytes = std::max(peak_bytes, in_use);
      workers.emplace_back([this, r] { run_request(r); });
    }
  }
};

extern "C" {

struct chucky_server* chucky_server_create(size_t hbm_budget_bytes, int max_concurrent, int pipeline)
{
  auto* s = new chucky_server();
  s->budget = hbm_budget_bytes;
  s->max_concurrent = max_concurrent > 0 ? max_concurrent : 1;
  s->pipeline = pipeline;
  (void)hipGetDevice(&s->hip_device);
  s->scheduler = std::thread([s] { s->schedule(); });
  return s;
}

int chucky_server_submit(struct chucky_server* srv, const struct tile_stream_configuration* config,
                         struct shard_sink* sink, const void* data, size_t nbytes,
                         int data_is_device, uint64_t* request_id)
{
  // admission counts what the request's pipeline allocates
  tile_stream_memory_info info;
  tile_stream_gpu_options o;
  tile_stream_gpu_options_default(&o);
  if (srv)
    o.pipeline = srv->pipeline;
  if (!srv || !config || !sink || tile_stream_gpu_memory_estimate_ex(config, &o, &info)) {
    chucky::set_error("chucky_server_submit: invalid request");
    return 1;
  }
  Request r{};
  r.config = *config;
  r.sink = sink;
  r.data = (const uint8_t*)data;
  r.nbytes = nbytes;
  r.device = data_is_device;
  r.estimate = info.device_bytes;
  r.submitted = Clock::now();
  std::lock_guard<std::mutex> lk(srv->mu);
  r.id = srv->next_id++;
  if (request_id)
    *request_id = r.id;
  srv->queue.push_back(r);
  ++srv->submitted;
  srv->cv.notify_all();
  return 0;
}

int chucky_server_wait_all(struct chucky_server* srv)
{
  std::unique_lock<std::mutex> lk(srv->mu);
  srv->cv.wait(lk, [&] { return srv->finished == srv->submitted; });
  if (srv->failures)
    chucky::set_error(srv->error);
  return srv->failures ? 1 : 0;
}

size_t chucky_server_reports(struct chucky_server* srv, struct chucky_request_report* out, size_t cap)
{
  std::lock_guard<std::mutex> lk(srv->mu);
  const size_t n = std::min(cap, srv->reports.size());
  std::copy(srv->reports.begin(), srv->reports.begin() + n, out);
  return n;
}

size_t chucky_server_latencies(struct chucky_server* srv, double* out, size_t cap)
{
  std::lock_guard<std::mutex> lk(srv->mu);
  const size_t n = std::min(cap, srv->latencies.size());
  std::copy(srv->latencies.begin(), srv->latencies.begin() + n, out);
  return n;
}

void chucky_server_peaks(struct chucky_server* srv, int* running, size_t* device_bytes)
{
  std::lock_guard<std::mutex> lk(srv->mu);
  if (running)
    *running = srv->peak_running;
  if (device_bytes)
    *device_bytes = srv->peak_bytes;
}

void chucky_server_destroy(struct chucky_server* srv)
{
  if (!srv)
    return;
  chucky_server_wait_all(srv);
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv.notify_all();
  srv->scheduler.join();
  for (auto& t : srv->workers)
    t.join();
  delete srv;
}

} // extern "C"
