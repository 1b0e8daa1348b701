/* Element types of a stream.
 * Same enumerators, order and byte widths as the reference (src/dtype.h:10-75):
 * the enum is part of the C ABI of tile_stream_configuration. */
#pragma once

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dtype
{
  dtype_u8,
  dtype_u16,
  dtype_u32,
  dtype_u64,
  dtype_i8,
  dtype_i16,
  dtype_i32,
  dtype_i64,
  dtype_f16,
  dtype_f32,
  dtype_f64,
};

/* Bytes per element; 0 for an unknown value. */
static inline size_t
dtype_bpe(enum dtype dt)
{
  static const unsigned char w[] = { 1, 2, 4, 8, 1, 2, 4, 8, 2, 4, 8 };
  return ((unsigned)dt < sizeof(w)) ? (size_t)w[dt] : 0;
}

/* Zarr v3 data_type name. */
static inline const char*
dtype_zarr_string(enum dtype dt)
{
  static const char* const n[] = { "uint8",   "uint16",  "uint32", "uint64",
                                   "int8",    "int16",   "int32",  "int64",
                                   "float16", "float32", "float64" };
  return ((unsigned)dt < 11u) ? n[dt] : "unknown";
}

#ifdef __cplusplus
}
#endif
