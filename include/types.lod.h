/* Level-of-detail reduction methods (reference: src/types.lod.h:10-27). */
#pragma once

#include "dtype.h"

#ifdef __cplusplus
extern "C" {
#endif

enum lod_reduce_method
{
  lod_reduce_mean,
  lod_reduce_min,
  lod_reduce_max,
  lod_reduce_median,
  lod_reduce_max_suppressed, /* second-highest value of the window */
  lod_reduce_min_suppressed, /* second-lowest value of the window */
};

/* Accumulators for the append-dimension fold keep the element's own width. */
static inline size_t
dtype_accum_bpe(enum dtype dt, enum lod_reduce_method method)
{
  (void)method;
  return dtype_bpe(dt);
}

#ifdef __cplusplus
}
#endif
