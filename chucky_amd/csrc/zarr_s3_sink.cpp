// S3 sink: configuration defaults and validation as in the reference
// (src/zarr/zarr_s3_sink.c:27-140). The transport itself needs an S3 client
// library that this image does not provide, so creation reports an error.
#include "zarr_s3_sink.h"

#include "defs.limits.h"
#include "plan.hpp"

#include <cstring>
#include <string>

namespace chucky {
void set_error(const std::string& msg);
}
using namespace chucky;

namespace {

void transport_defaults(size_t* part, double* gbps)
{
  if (*part == 0)
    *part = S3_DEFAULT_PART_SIZE;
  if (*gbps == 0.0)
    *gbps = S3_DEFAULT_THROUGHPUT_GBPS;
}

// Uncompressed shard size bounds the multipart part count.
int part_count_ok(uint8_t rank, const dimension* dims, enum dtype dt, size_t part)
{
  const size_t bpe = dtype_bpe(dt);
  if (!bpe)
    return 0;
  uint64_t elems = 1, cps_total = 1;
  for (int d = 0; d < rank; ++d) {
    uint64_t cps = dims[d].chunks_per_shard;
    if (cps == 0)
      cps = dims[d].size == 0 ? 1 : ceildiv(dims[d].size, dims[d].chunk_size);
    cps_total *= cps;
    elems *= dims[d].chunk_size * cps;
  }
  const uint64_t max_bytes = elems * bpe + cps_total * 16 + 4;
  if (ceildiv(max_bytes, part) > S3_MAX_PARTS) {
    set_error("shard too large for S3 multipart upload");
    return 0;
  }
  return 1;
}

bool key_part_ok(const char* s)
{
  if (!s || !s[0] || s[0] == '/')
    return false;
  return s[std::strlen(s) - 1] != '/';
}

template<typename Cfg>
int validate(const Cfg* c, bool name_required)
{
  if (!c || !key_part_ok(c->bucket) || !key_part_ok(c->prefix))
    return 1;
  if (name_required ? !key_part_ok(c->array_name) : (c->array_name && !key_part_ok(c->array_name)))
    return 1;
  if (!c->region || !c->region[0] || !c->endpoint || !c->endpoint[0])
    return 1;
  if (c->rank == 0 || c->rank > MAX_ZARR_RANK || !c->dimensions || dims_validate(c->dimensions, c->rank))
    return 1;
  if (!dtype_bpe(c->data_type) || c->part_size == 0)
    return 1;
  if (c->array_name && std::strlen(c->prefix) + 1 + std::strlen(c->array_name) + sizeof("/zarr.json") >= 4096)
    return 1;
  return part_count_ok(c->rank, c->dimensions, c->data_type, c->part_size) ? 0 : 1;
}

} // namespace

extern "C" {

void zarr_s3_config_set_defaults(struct zarr_s3_config* cfg)
{
  if (cfg)
    transport_defaults(&cfg->part_size, &cfg->throughput_gbps);
}

int zarr_s3_config_validate(const struct zarr_s3_config* cfg) { return validate(cfg, true); }

struct zarr_s3_sink* zarr_s3_sink_create(struct zarr_s3_config* cfg)
{
  zarr_s3_config_set_defaults(cfg);
  if (zarr_s3_config_validate(cfg))
    return nullptr;
  set_error("zarr_s3_sink_create: no S3 transport in this build");
  return nullptr;
}

int zarr_s3_sink_has_error(const struct zarr_s3_sink*) { return 1; }
int zarr_s3_sink_destroy(struct zarr_s3_sink*) { return 0; }
int zarr_s3_sink_flush(struct zarr_s3_sink*) { return 1; }
struct shard_sink* zarr_s3_sink_as_shard_sink(struct zarr_s3_sink*) { return nullptr; }

void zarr_s3_multiscale_config_set_defaults(struct zarr_s3_multiscale_config* cfg)
{
  if (cfg)
    transport_defaults(&cfg->part_size, &cfg->throughput_gbps);
}

int zarr_s3_multiscale_config_validate(const struct zarr_s3_multiscale_config* cfg)
{
  return validate(cfg, false);
}

struct zarr_s3_multiscale_sink* zarr_s3_multiscale_sink_create(struct zarr_s3_multiscale_config* cfg)
{
  zarr_s3_multiscale_config_set_defaults(cfg);
  if (zarr_s3_multiscale_config_validate(cfg))
    return nullptr;
  set_error("zarr_s3_multiscale_sink_create: no S3 transport in this build");
  return nullptr;
}

int zarr_s3_multiscale_sink_destroy(struct zarr_s3_multiscale_sink*) { return 0; }
int zarr_s3_multiscale_sink_flush(struct zarr_s3_multiscale_sink*) { return 1; }
struct shard_sink* zarr_s3_multiscale_sink_as_shard_sink(struct zarr_s3_multiscale_sink*) { return nullptr; }

} // extern "C"
