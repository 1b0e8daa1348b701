// Zarr v3 filesystem sink with an asynchronous I/O worker.
//
// Reproduces the reference's on-disk format and sink behaviour
// (src/zarr/zarr_fs_sink.c:51-830, src/zarr/zarr_metadata.c:33-411,
// src/zarr/io_queue.posix.c:35-194): array and group zarr.json documents
// with the same keys, order and number formatting; shard files at
// {array}/c/<coords> with coordinates unravelled from the flat shard index;
// one writer per inner shard reused across shard epochs; pwrite jobs run in
// posting order on one worker thread, fences are sequence numbers.
#include "zarr_fs_sink.h"

#include "plan.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cerrno>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace chucky {
void set_error(const std::string& msg);

// ---- I/O queue --------------------------------------------------------------

class IoQueue
{
public:
  IoQueue() : worker_([this] { run(); }) {}
  ~IoQueue()
  {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    worker_.join();
  }
  void post(std::function<void()> fn)
  {
    {
      std::lock_guard<std::mutex> lk(mu_);
      jobs_.push_back(std::move(fn));
      ++posted_;
    }
    cv_.notify_all();
  }
  io_event record()
  {
    std::lock_guard<std::mutex> lk(mu_);
    return io_event{ posted_ };
  }
  void wait(io_event ev)
  {
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return done_ >= ev.seq; });
  }

private:
  void run()
  {
    for (;;) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
        if (jobs_.empty())
          return;
        job = std::move(jobs_.front());
        jobs_.pop_front();
      }
      job();
      {
        std::lock_guard<std::mutex> lk(mu_);
        ++done_;
      }
      done_cv_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<std::function<void()>> jobs_;
  uint64_t posted_ = 0, done_ = 0;
  bool stop_ = false;
  std::thread worker_;
};

// ---- JSON (compact, src/zarr/json_writer.c semantics) --------------------

class Json
{
public:
  void obj_begin() { comma(); s_ += '{'; need_ = false; }
  void obj_end() { s_ += '}'; need_ = true; }
  void arr_begin() { comma(); s_ += '['; need_ = false; }
  void arr_end() { s_ += ']'; need_ = true; }
  void key(const char* k) { comma(); s_ += '"'; s_ += k; s_ += "\":"; need_ = false; }
  void str(const char* v)
  {
    comma();
    s_ += '"';
    for (const char* p = v; *p; ++p) {
      const unsigned char c = (unsigned char)*p;
      switch (c) {
        case '"': s_ += "\\\""; break;
        case '\\': s_ += "\\\\"; break;
        case '\n': s_ += "\\n"; break;
        case '\r': s_ += "\\r"; break;
        case '\t': s_ += "\\t"; break;
        default:
          if (c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", c);
            s_ += b;
          } else {
            s_ += (char)c;
          }
      }
    }
    s_ += '"';
    need_ = true;
  }
  void i64(long long v) { comma(); s_ += std::to_string(v); need_ = true; }
  void u64(unsigned long long v) { comma(); s_ += std::to_string(v); need_ = true; }
  void f64(double v)
  {
    comma();
    char b[64];
    if (v == (double)(long long)v && v >= -1e15 && v <= 1e15)
      std::snprintf(b, sizeof b, "%.1f", v);
    else
      std::snprintf(b, sizeof b, "%g", v);
    s_ += b;
    need_ = true;
  }
  void null() { comma(); s_ += "null"; need_ = true; }
  void boolean(bool v) { comma(); s_ += v ? "true" : "false"; need_ = true; }
  const std::string& str() const { return s_; }

private:
  void comma()
  {
    if (need_)
      s_ += ',';
  }
  std::string s_;
  bool need_ = false;
};

std::string root_json()
{
  Json j;
  j.obj_begin();
  j.key("zarr_format");
  j.i64(3);
  j.key("node_type");
  j.str("group");
  j.key("consolidated_metadata");
  j.null();
  j.key("attributes");
  j.obj_begin();
  j.obj_end();
  j.obj_end();
  return j.str();
}

// src/zarr/zarr_metadata.c:33-207
std::string array_json(int rank, const dimension* dims, enum dtype dt, double fill,
                       const uint64_t* cps, codec_config codec)
{
  Json j;
  j.obj_begin();
  j.key("zarr_format");
  j.i64(3);
  j.key("node_type");
  j.str("array");
  j.key("shape");
  j.arr_begin();
  for (int d = 0; d < rank; ++d)
    j.u64(dims[d].size);
  j.arr_end();
  j.key("data_type");
  j.str(dtype_zarr_string(dt));
  j.key("chunk_grid");
  j.obj_begin();
  j.key("name");
  j.str("regular");
  j.key("configuration");
  j.obj_begin();
  j.key("chunk_shape");
  j.arr_begin();
  for (int d = 0; d < rank; ++d)
    j.u64(dims[d].chunk_size * cps[d]);
  j.arr_end();
  j.obj_end();
  j.obj_end();
  j.key("chunk_key_encoding");
  j.obj_begin();
  j.key("name");
  j.str("default");
  j.key("configuration");
  j.obj_begin();
  j.key("separator");
  j.str("/");
  j.obj_end();
  j.obj_end();
  j.key("codecs");
  j.arr_begin();
  j.obj_begin();
  j.key("name");
  j.str("sharding_indexed");
  j.key("configuration");
  j.obj_begin();
  j.key("chunk_shape");
  j.arr_begin();
  for (int d = 0; d < rank; ++d)
    j.u64(dims[d].chunk_size);
  j.arr_end();
  j.key("codecs");
  j.arr_begin();
  j.obj_begin();
  j.key("name");
  j.str("bytes");
  j.key("configuration");
  j.obj_begin();
  j.key("endian");
  j.str("little");
  j.obj_end();
  j.obj_end();
  if (codec.id != CODEC_NONE) {
    const bool blosc = codec.id == CODEC_BLOSC_LZ4 || codec.id == CODEC_BLOSC_ZSTD;
    j.obj_begin();
    j.key("name");
    j.str(blosc ? "blosc" : (codec.id == CODEC_LZ4_NON_STANDARD ? "lz4" : "zstd"));
    j.key("configuration");
    j.obj_begin();
    if (codec.id == CODEC_ZSTD) {
      j.key("level");
      j.i64(codec.level);
      j.key("checksum");
      j.boolean(false);
    }
    if (blosc) {
      j.key("cname");
      j.str(codec.id == CODEC_BLOSC_LZ4 ? "lz4" : "zstd");
      j.key("clevel");
      j.i64(codec.level);
      j.key("shuffle");
      j.str(codec.shuffle == CODEC_SHUFFLE_BIT ? "bitshuffle"
                                                : (codec.shuffle == CODEC_SHUFFLE_BYTE ? "shuffle"
                                                                                       : "noshuffle"));
      j.key("typesize");
      j.i64((long long)dtype_bpe(dt));
      j.key("blocksize");
      j.i64(0);
    }
    j.obj_end();
    j.obj_end();
  }
  j.arr_end();
  j.key("index_codecs");
  j.arr_begin();
  j.obj_begin();
  j.key("name");
  j.str("bytes");
  j.key("configuration");
  j.obj_begin();
  j.key("endian");
  j.str("little");
  j.obj_end();
  j.obj_end();
  j.obj_begin();
  j.key("name");
  j.str("crc32c");
  j.key("configuration");
  j.obj_begin();
  j.obj_end();
  j.obj_end();
  j.arr_end();
  j.key("index_location");
  j.str("end");
  j.obj_end();
  j.obj_end();
  j.arr_end();
  j.key("fill_value");
  j.f64(fill);
  bool named = false;
  for (int d = 0; d < rank; ++d)
    named |= dims[d].name != nullptr;
  if (named) {
    j.key("dimension_names");
    j.arr_begin();
    for (int d = 0; d < rank; ++d) {
      if (dims[d].name)
        j.str(dims[d].name);
      else
        j.null();
    }
    j.arr_end();
  }
  j.obj_end();
  return j.str();
}

// src/zarr/zarr_metadata.c:209-346
std::string multiscale_group_json(int rank, int nlod, const std::vector<std::vector<dimension>>& lv)
{
  Json j;
  const auto& l0 = lv[0];
  j.obj_begin();
  j.key("zarr_format");
  j.i64(3);
  j.key("node_type");
  j.str("group");
  j.key("consolidated_metadata");
  j.null();
  j.key("attributes");
  j.obj_begin();
  j.key("ome");
  j.obj_begin();
  j.key("version");
  j.str("0.5");
  j.key("multiscales");
  j.arr_begin();
  j.obj_begin();
  j.key("axes");
  j.arr_begin();
  for (int d = 0; d < rank; ++d) {
    j.obj_begin();
    j.key("name");
    if (l0[d].name)
      j.str(l0[d].name);
    else
      j.str(("d" + std::to_string(d)).c_str());
    j.key("type");
    const char* t = "space";
    switch (l0[d].ngff.type) {
      case dimension_axis_time: t = "time"; break;
      case dimension_axis_channel: t = "channel"; break;
      case dimension_axis_other: t = "custom"; break;
      default: break;
    }
    j.str(t);
    if (l0[d].ngff.unit) {
      j.key("unit");
      j.str(l0[d].ngff.unit);
    }
    j.obj_end();
  }
  j.arr_end();
  j.key("datasets");
  j.arr_begin();
  for (int l = 0; l < nlod; ++l) {
    j.obj_begin();
    j.key("path");
    j.str(std::to_string(l).c_str());
    std::vector<double> scale(rank), trans(rank);
    for (int d = 0; d < rank; ++d) {
      const double phys = l0[d].ngff.scale > 0 ? l0[d].ngff.scale : 1.0;
      double factor = 1.0;
      if (l0[d].size > 0 && lv[l][d].size > 0)
        factor = (double)l0[d].size / (double)lv[l][d].size;
      scale[d] = phys * factor;
      trans[d] = 0.5 * phys * (factor - 1.0);
    }
    j.key("coordinateTransformations");
    j.arr_begin();
    j.obj_begin();
    j.key("type");
    j.str("scale");
    j.key("scale");
    j.arr_begin();
    for (int d = 0; d < rank; ++d)
      j.f64(scale[d]);
    j.arr_end();
    j.obj_end();
    j.obj_begin();
    j.key("type");
    j.str("translation");
    j.key("translation");
    j.arr_begin();
    for (int d = 0; d < rank; ++d)
      j.f64(trans[d]);
    j.arr_end();
    j.obj_end();
    j.arr_end();
    j.obj_end();
  }
  j.arr_end();
  j.obj_end();
  j.arr_end();
  j.obj_end();
  j.obj_end();
  j.obj_end();
  return j.str();
}

// src/zarr/zarr_metadata.c:381-411: "c/i/j/k" with dim 0 taking the rest.
std::string shard_key(int rank, const uint64_t* shard_count, uint64_t flat)
{
  std::vector<uint64_t> c(rank);
  uint64_t rem = flat;
  for (int d = rank - 1; d >= 0; --d) {
    if (d == 0)
      c[d] = rem;
    else {
      c[d] = rem % shard_count[d];
      rem /= shard_count[d];
    }
  }
  std::string k = "c";
  for (int d = 0; d < rank; ++d)
    k += "/" + std::to_string(c[d]);
  return k;
}

int mkdirp(const std::string& path)
{
  std::string cur;
  size_t i = 0;
  while (i <= path.size()) {
    const size_t j = path.find('/', i);
    cur = path.substr(0, j == std::string::npos ? path.size() : j);
    if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
      return 1;
    if (j == std::string::npos)
      break;
    i = j + 1;
  }
  return 0;
}

int write_file(const std::string& path, const std::string& data)
{
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f)
    return 1;
  const size_t n = std::fwrite(data.data(), 1, data.size(), f);
  std::fclose(f);
  return n == data.size() ? 0 : 1;
}

int pwrite_all(int fd, const void* buf, size_t n, uint64_t off)
{
  const char* p = (const char*)buf;
  while (n > 0) {
    const ssize_t w = ::pwrite(fd, p, n, (off_t)off);
    if (w < 0) {
      if (errno == EINTR)
        continue;
      return errno ? errno : 1;
    }
    p += w;
    n -= (size_t)w;
    off += (uint64_t)w;
  }
  return 0;
}

} // namespace chucky

using namespace chucky;

struct FsWriter
{
  shard_writer base;
  struct zarr_fs_sink* owner;
  int fd = -1;
};

struct zarr_fs_sink
{
  shard_sink base;
  IoQueue* queue = nullptr;
  int unbuffered = 0;
  size_t page = 4096;
  std::atomic<uint64_t> queued{ 0 }, retired{ 0 };
  std::atomic<int> io_error{ 0 };
  std::mutex err_mu;
  std::string err_msg; // the first failed write, for the caller's error message
  int rank = 0;
  std::vector<DimExtent> ext;
  uint64_t shard_inner_count = 1;
  std::string array_dir;
  std::vector<FsWriter> writers;
  std::vector<dimension> dims;
  enum dtype dt = dtype_u8;
  double fill = 0;
  codec_config codec{};
};

namespace {

zarr_fs_sink* fs_of(shard_sink* s) { return reinterpret_cast<zarr_fs_sink*>(s); }

// Runs on the I/O worker: records the first failed write.
void io_failed(zarr_fs_sink* zs, int err, const void* buf, size_t n, uint64_t off)
{
  std::lock_guard<std::mutex> lk(zs->err_mu);
  if (zs->err_msg.empty())
    zs->err_msg = "zarr_fs_sink: pwrite of " + std::to_string(n) + " bytes at offset " +
                  std::to_string(off) + " (buffer address mod 4096 = " +
                  std::to_string((uintptr_t)buf % 4096) + ") failed: " + std::strerror(err);
  zs->io_error = 1;
}

int io_status(zarr_fs_sink* zs)
{
  if (!zs->io_error.load())
    return 0;
  std::lock_guard<std::mutex> lk(zs->err_mu);
  set_error(zs->err_msg.empty() ? std::string("zarr_fs_sink: write failed") : zs->err_msg);
  return 1;
}

int fs_write(shard_writer* self, uint64_t offset, const void* beg, const void* end)
{
  FsWriter* w = reinterpret_cast<FsWriter*>(self);
  zarr_fs_sink* zs = w->owner;
  const size_t n = (size_t)((const char*)end - (const char*)beg);
  if (n == 0)
    return 0;
  void* copy = nullptr;
  const size_t align = zs->unbuffered ? zs->page : 16;
  if (posix_memalign(&copy, align, n ? n : 1))
    return 1;
  std::memcpy(copy, beg, n);
  const int fd = w->fd;
  zs->queued += n;
  zs->queue->post([zs, fd, copy, n, offset] {
    if (const int e = pwrite_all(fd, copy, n, offset))
      io_failed(zs, e, copy, n, offset);
    std::free(copy);
    zs->retired += n;
  });
  return io_status(zs);
}

int fs_write_direct(shard_writer* self, uint64_t offset, const void* beg, const void* end)
{
  FsWriter* w = reinterpret_cast<FsWriter*>(self);
  zarr_fs_sink* zs = w->owner;
  const size_t n = (size_t)((const char*)end - (const char*)beg);
  if (n == 0)
    return 0;
  const int fd = w->fd;
  zs->queued += n;
  // O_DIRECT has the kernel DMA from the caller's pages. The stream hands
  // over pinned buffers from the HIP runtime, which are not ordinary user
  // pages: O_DIRECT writes from them came out as zeros or failed, so
  // unbuffered stores write from an aligned bounce copy -- made by the I/O
  // worker, since the caller's buffer stays valid until the next fence.
  if (zs->unbuffered) {
    const size_t page = zs->page;
    zs->queue->post([zs, fd, beg, n, offset, page] {
      void* copy = nullptr;
      if (posix_memalign(&copy, page, n)) {
        io_failed(zs, ENOMEM, beg, n, offset);
      } else {
        std::memcpy(copy, beg, n);
        if (const int e = pwrite_all(fd, copy, n, offset))
          io_failed(zs, e, copy, n, offset);
        std::free(copy);
      }
      zs->retired += n;
    });
    return io_status(zs);
  }
  zs->queue->post([zs, fd, beg, n, offset] {
    if (const int e = pwrite_all(fd, beg, n, offset))
      io_failed(zs, e, beg, n, offset);
    zs->retired += n;
  });
  return io_status(zs);
}

int fs_finalize(shard_writer* self)
{
  FsWriter* w = reinterpret_cast<FsWriter*>(self);
  if (w->fd < 0)
    return 0;
  const int fd = w->fd;
  w->owner->queue->post([fd] { ::close(fd); });
  w->fd = -1;
  return 0;
}

shard_writer* fs_open(shard_sink* self, uint8_t, uint64_t shard_index)
{
  zarr_fs_sink* zs = fs_of(self);
  FsWriter& w = zs->writers[shard_index % zs->shard_inner_count];
  std::vector<uint64_t> sc(zs->rank);
  for (int d = 0; d < zs->rank; ++d)
    sc[d] = zs->ext[d].shard_count;
  const std::string path = zs->array_dir + "/" + shard_key(zs->rank, sc.data(), shard_index);
  // a shard file starts empty: an older, longer file at the same key would
  // keep its tail (and its stale index) past the new index
  const int flags = O_WRONLY | O_CREAT | O_TRUNC | (zs->unbuffered ? O_DIRECT : 0);
  int fd = ::open(path.c_str(), flags, 0644);
  if (fd < 0) {
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos && mkdirp(path.substr(0, slash)) == 0)
      fd = ::open(path.c_str(), flags, 0644);
  }
  if (fd < 0) {
    set_error("zarr_fs_sink: open(" + path + ") failed");
    return nullptr;
  }
  w.fd = fd;
  return &w.base;
}

io_event fs_record(shard_sink* self, uint8_t) { return fs_of(self)->queue->record(); }
void fs_wait(shard_sink* self, uint8_t, io_event ev) { fs_of(self)->queue->wait(ev); }

int fs_write_metadata(zarr_fs_sink* zs)
{
  std::vector<uint64_t> cps(zs->rank);
  for (int d = 0; d < zs->rank; ++d)
    cps[d] = zs->ext[d].chunks_per_shard;
  return write_file(zs->array_dir + "/zarr.json",
                    array_json(zs->rank, zs->dims.data(), zs->dt, zs->fill, cps.data(), zs->codec));
}

int fs_update_append(shard_sink* self, uint8_t, uint8_t n_append, const uint64_t* sizes)
{
  zarr_fs_sink* zs = fs_of(self);
  if (n_append == 0 || n_append > zs->rank)
    return 1;
  bool changed = false;
  for (int d = 0; d < n_append; ++d)
    changed |= zs->dims[d].size != sizes[d];
  if (!changed)
    return 0;
  for (int d = 0; d < n_append; ++d)
    zs->dims[d].size = sizes[d];
  if (fs_write_metadata(zs)) {
    set_error("zarr_fs_sink: rewriting zarr.json failed");
    return 1;
  }
  return 0;
}

int write_intermediate_groups(const std::string& store, const char* array_name)
{
  if (!array_name)
    return 0;
  const std::string name = array_name;
  if (name.empty() || name[0] == '/' || name.back() == '/' || name.find("//") != std::string::npos)
    return 1;
  for (size_t i = 0; i < name.size(); ++i)
    if (name[i] == '/') {
      const std::string dir = store + "/" + name.substr(0, i);
      if (mkdirp(dir) || write_file(dir + "/zarr.json", root_json()))
        return 1;
    }
  return 0;
}

} // namespace

extern "C" {

struct zarr_fs_sink* zarr_fs_sink_create(const struct zarr_config* cfg)
{
  if (!cfg || !cfg->store_path || cfg->rank == 0 || cfg->rank > MAX_ZARR_RANK || !cfg->dimensions)
    return nullptr;
  zarr_fs_sink* zs = new zarr_fs_sink();
  zs->base.open = fs_open;
  zs->base.update_append = fs_update_append;
  zs->base.record_fence = fs_record;
  zs->base.wait_fence = fs_wait;
  zs->unbuffered = cfg->unbuffered;
  zs->page = (size_t)sysconf(_SC_PAGESIZE);
  zs->rank = cfg->rank;
  zs->dt = cfg->data_type;
  zs->fill = cfg->fill_value;
  zs->codec = cfg->codec;
  zs->dims.assign(cfg->dimensions, cfg->dimensions + cfg->rank);
  zs->ext.assign(cfg->rank, DimExtent{});
  std::vector<uint64_t> cps(cfg->rank);
  for (int d = 0; d < cfg->rank; ++d) {
    zs->ext[d].size = cfg->dimensions[d].size;
    zs->ext[d].chunk_size = cfg->dimensions[d].chunk_size;
    cps[d] = cfg->dimensions[d].chunks_per_shard;
  }
  const int na = dims_n_append(cfg->dimensions, cfg->rank);
  zs->shard_inner_count = compute_shards(zs->ext, na, cps.data());
  zs->array_dir = cfg->array_name ? std::string(cfg->store_path) + "/" + cfg->array_name
                                  : std::string(cfg->store_path);
  // shard directories for every shard when dim 0 is bounded, else the first
  // shard epoch only (later ones are created on open)
  {
    uint64_t total = zs->shard_inner_count;
    for (int d = 1; d < na; ++d)
      total *= zs->ext[d].shard_count;
    if (cfg->dimensions[0].size > 0)
      total *= zs->ext[0].shard_count;
    std::vector<uint64_t> sc(zs->rank);
    for (int d = 0; d < zs->rank; ++d)
      sc[d] = zs->ext[d].shard_count;
    std::string last;
    for (uint64_t f = 0; f < total; ++f) {
      std::string p = zs->array_dir + "/" + shard_key(zs->rank, sc.data(), f);
      p = p.substr(0, p.rfind('/'));
      if (p != last && mkdirp(p)) {
        set_error("zarr_fs_sink: mkdir " + p + " failed");
        delete zs;
        return nullptr;
      }
      last = p;
    }
  }
  if (cfg->array_name) {
    if (mkdirp(cfg->store_path) || write_file(std::string(cfg->store_path) + "/zarr.json", root_json()) ||
        write_intermediate_groups(cfg->store_path, cfg->array_name)) {
      delete zs;
      return nullptr;
    }
  }
  if (mkdirp(zs->array_dir) || fs_write_metadata(zs)) {
    delete zs;
    return nullptr;
  }
  zs->writers.resize(zs->shard_inner_count);
  for (auto& w : zs->writers) {
    w.base.write = fs_write;
    w.base.write_direct = fs_write_direct;
    w.base.finalize = fs_finalize;
    w.owner = zs;
  }
  zs->queue = new IoQueue();
  return zs;
}

void zarr_fs_sink_flush(struct zarr_fs_sink* s)
{
  if (s && s->queue)
    s->queue->wait(s->queue->record());
}

void zarr_fs_sink_destroy(struct zarr_fs_sink* s)
{
  if (!s)
    return;
  zarr_fs_sink_flush(s);
  delete s->queue;
  for (auto& w : s->writers)
    if (w.fd >= 0)
      ::close(w.fd);
  delete s;
}

size_t zarr_fs_sink_pending_bytes(struct zarr_fs_sink* s)
{
  return s ? (size_t)(s->queued.load() - s->retired.load()) : 0;
}

struct shard_sink* zarr_fs_sink_as_shard_sink(struct zarr_fs_sink* s) { return &s->base; }

} // extern "C"

// ---- multiscale -----------------------------------------------------------

struct zarr_fs_multiscale_sink
{
  shard_sink base;
  std::vector<zarr_fs_sink*> levels;
  std::string group;
  int rank = 0;
};

namespace {

zarr_fs_multiscale_sink* ms_of(shard_sink* s) { return reinterpret_cast<zarr_fs_multiscale_sink*>(s); }

shard_writer* ms_open(shard_sink* self, uint8_t level, uint64_t shard_index)
{
  auto* ms = ms_of(self);
  if (level >= ms->levels.size())
    return nullptr;
  return fs_open(&ms->levels[level]->base, level, shard_index);
}

io_event ms_record(shard_sink* self, uint8_t level)
{
  auto* ms = ms_of(self);
  return level < ms->levels.size() ? ms->levels[level]->queue->record() : io_event{ 0 };
}

void ms_wait(shard_sink* self, uint8_t level, io_event ev)
{
  auto* ms = ms_of(self);
  if (level < ms->levels.size())
    ms->levels[level]->queue->wait(ev);
}

int ms_write_group(zarr_fs_multiscale_sink* ms)
{
  std::vector<std::vector<dimension>> lv;
  for (auto* z : ms->levels)
    lv.push_back(z->dims);
  return write_file(ms->group + "/zarr.json", multiscale_group_json(ms->rank, (int)lv.size(), lv));
}

int ms_update_append(shard_sink* self, uint8_t level, uint8_t n_append, const uint64_t* sizes)
{
  auto* ms = ms_of(self);
  if (level >= ms->levels.size())
    return 1;
  const uint64_t old = ms->levels[level]->dims[0].size;
  if (fs_update_append(&ms->levels[level]->base, level, n_append, sizes))
    return 1;
  if (old == sizes[0])
    return 0;
  return ms_write_group(ms);
}

} // namespace

extern "C" {

struct zarr_fs_multiscale_sink* zarr_fs_multiscale_sink_create(const struct zarr_multiscale_config* cfg)
{
  if (!cfg || !cfg->store_path || cfg->rank == 0 || cfg->rank > MAX_ZARR_RANK || !cfg->dimensions)
    return nullptr;
  const std::string group = cfg->array_name ? std::string(cfg->store_path) + "/" + cfg->array_name
                                            : std::string(cfg->store_path);
  // Level shapes of the full array (src/lod/lod_plan.c:347-366): LOD over
  // downsampled inner dims, stopping when every LOD dim fits one chunk.
  const int rank = cfg->rank;
  const int na = dims_n_append(cfg->dimensions, cfg->rank);
  std::vector<int> lod;
  for (int d = na; d < rank; ++d)
    if (cfg->dimensions[d].downsample)
      lod.push_back(d);
  std::vector<std::vector<uint64_t>> shapes(1);
  for (int d = 0; d < rank; ++d)
    shapes[0].push_back(cfg->dimensions[d].size ? cfg->dimensions[d].size : cfg->dimensions[d].chunk_size);
  const int max_lev = cfg->nlod > 0 ? cfg->nlod : LOD_MAX_LEVELS;
  while ((int)shapes.size() < max_lev && !lod.empty()) {
    bool all_one = true;
    for (int d : lod)
      all_one &= ceildiv(shapes.back()[d], cfg->dimensions[d].chunk_size) <= 1;
    if (all_one)
      break;
    auto nx = shapes.back();
    for (int d : lod)
      nx[d] = (nx[d] + 1) / 2;
    shapes.push_back(nx);
  }
  auto* ms = new zarr_fs_multiscale_sink();
  ms->base.open = ms_open;
  ms->base.update_append = ms_update_append;
  ms->base.record_fence = ms_record;
  ms->base.wait_fence = ms_wait;
  ms->group = group;
  ms->rank = rank;
  bool ok = true;
  if (cfg->array_name) {
    ok = !mkdirp(cfg->store_path) && !mkdirp(group) &&
         !write_file(std::string(cfg->store_path) + "/zarr.json", root_json()) &&
         !write_intermediate_groups(cfg->store_path, cfg->array_name);
  } else {
    ok = !mkdirp(group);
  }
  for (size_t lv = 0; ok && lv < shapes.size(); ++lv) {
    std::vector<dimension> ld(cfg->dimensions, cfg->dimensions + rank);
    std::vector<DimExtent> ext(rank);
    std::vector<uint64_t> cps(rank);
    for (int d = 0; d < rank; ++d) {
      ext[d].size = shapes[lv][d];
      ext[d].chunk_size = cfg->dimensions[d].chunk_size;
      cps[d] = cfg->dimensions[d].chunks_per_shard;
    }
    compute_shards(ext, na, cps.data());
    for (int d = 0; d < rank; ++d) {
      ld[d].size = (d == 0 && cfg->dimensions[0].size == 0) ? 0 : shapes[lv][d];
      ld[d].chunks_per_shard = ext[d].chunks_per_shard;
    }
    const std::string name = std::to_string(lv);
    zarr_config z{};
    z.store_path = group.c_str();
    z.array_name = name.c_str();
    z.data_type = cfg->data_type;
    z.fill_value = cfg->fill_value;
    z.rank = cfg->rank;
    z.dimensions = ld.data();
    z.unbuffered = cfg->unbuffered;
    z.codec = cfg->codec;
    zarr_fs_sink* s = zarr_fs_sink_create(&z);
    if (!s)
      ok = false;
    else
      ms->levels.push_back(s);
  }
  if (!ok || ms_write_group(ms)) {
    for (auto* s : ms->levels)
      zarr_fs_sink_destroy(s);
    delete ms;
    return nullptr;
  }
  return ms;
}

void zarr_fs_multiscale_sink_destroy(struct zarr_fs_multiscale_sink* s)
{
  if (!s)
    return;
  for (auto* l : s->levels)
    zarr_fs_sink_destroy(l);
  delete s;
}

void zarr_fs_multiscale_sink_flush(struct zarr_fs_multiscale_sink* s)
{
  if (s)
    for (auto* l : s->levels)
      zarr_fs_sink_flush(l);
}

size_t zarr_fs_multiscale_sink_pending_bytes(struct zarr_fs_multiscale_sink* s)
{
  size_t t = 0;
  if (s)
    for (auto* l : s->levels)
      t += zarr_fs_sink_pending_bytes(l);
  return t;
}

struct shard_sink* zarr_fs_multiscale_sink_as_shard_sink(struct zarr_fs_multiscale_sink* s)
{
  return &s->base;
}

} // extern "C"
