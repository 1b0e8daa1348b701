// Multiscale (LOD) pyramid on the device.
//
// Per epoch (reference: src/cpu/pipeline.c:194-336, src/cpu/lod.cpp:54-633):
//   gather   linear epoch -> level-0 Morton order, per fixed-dims batch
//   reduce   level l -> l+1 over variable windows [ends[i-1], ends[i])
//            (mean: sequential widened sum then one divide; min/max keep-first;
//             median: insertion sort <= 16; max/min_suppressed: 2nd extreme)
//   fold     across epochs for levels >= 1 when the append dim is downsampled
//            (integer mean uses the shift-by-level running fold of
//             src/cpu/lod.cpp:363-414, which is pinned by the reference tests)
//   emit     accumulator -> level values every 2^l epochs
//   scatter  Morton level values -> chunk pool of the active levels
// Each output element is owned by one lane and its window is reduced
// sequentially in Morton order, so float sums are bit-identical to the CPU
// path. The pass is HBM-bound integer/byte work: no LDS staging is needed for
// the windowed reads, which are contiguous runs of 2^lod_ndim elements.
#include "kernels.hpp"
#include "stream_gpu.hpp"

#include <type_traits>

namespace chucky {
namespace {

template<typename T> struct Acc { using type = T; };
template<> struct Acc<uint8_t> { using type = uint32_t; };
template<> struct Acc<uint16_t> { using type = uint32_t; };
template<> struct Acc<uint32_t> { using type = uint64_t; };
template<> struct Acc<int8_t> { using type = int32_t; };
template<> struct Acc<int16_t> { using type = int32_t; };
template<> struct Acc<int32_t> { using type = int64_t; };
template<> struct Acc<_Float16> { using type = float; };

template<typename T> constexpr bool is_fp = std::is_floating_point<T>::value || std::is_same<T, _Float16>::value;

template<typename T>
__global__ __launch_bounds__(256) void gather_k(const T* __restrict__ linear, T* __restrict__ morton,
                                                const uint32_t* __restrict__ lut,
                                                const uint64_t* __restrict__ fixed, uint64_t n0,
                                                uint64_t total)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total)
    return;
  const uint64_t b = g / n0, m = g - b * n0;
  morton[g] = linear[fixed[b] + lut[m]];
}

template<typename T, int M>
__device__ __forceinline__ T reduce_window(const T* v, uint64_t s, uint64_t e)
{
  const uint64_t len = e - s;
  if constexpr (M == lod_reduce_mean) {
    using A = typename Acc<T>::type;
    A sum = 0;
    for (uint64_t j = s; j < e; ++j)
      sum += (A)v[j];
    return (T)(sum / (A)len);
  } else if constexpr (M == lod_reduce_min) {
    T best = v[s];
    for (uint64_t j = s + 1; j < e; ++j)
      if (v[j] < best)
        best = v[j];
    return best;
  } else if constexpr (M == lod_reduce_max) {
    T best = v[s];
    for (uint64_t j = s + 1; j < e; ++j)
      if (v[j] > best)
        best = v[j];
    return best;
  } else if constexpr (M == lod_reduce_median) {
    T buf[16];
    const uint64_t n = len <= 16 ? len : 16;
    for (uint64_t j = 0; j < n; ++j)
      buf[j] = v[s + j];
    for (uint64_t i = 1; i < n; ++i) {
      const T key = buf[i];
      uint64_t k = i;
      while (k > 0 && buf[k - 1] > key) {
        buf[k] = buf[k - 1];
        --k;
      }
      buf[k] = key;
    }
    return buf[n / 2];
  } else if constexpr (M == lod_reduce_max_suppressed) {
    T t1 = v[s], t2 = v[s];
    for (uint64_t j = s + 1; j < e; ++j) {
      const T x = v[j];
      if (x >= t1) {
        t2 = t1;
        t1 = x;
      } else if (x > t2)
        t2 = x;
    }
    return t2;
  } else {
    T b1 = v[s], b2 = v[s];
    for (uint64_t j = s + 1; j < e; ++j) {
      const T x = v[j];
      if (x <= b1) {
        b2 = b1;
        b1 = x;
      } else if (x < b2)
        b2 = x;
    }
    return b2;
  }
}

template<typename T, int M>
__global__ __launch_bounds__(256) void reduce_k(T* __restrict__ values, uint64_t src_beg,
                                                uint64_t dst_beg, const uint64_t* __restrict__ ends,
                                                uint64_t src_n, uint64_t dst_n, uint64_t total)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total)
    return;
  const uint64_t b = g / dst_n, i = g - b * dst_n;
  const uint64_t s = i ? ends[i - 1] : 0, e = ends[i];
  values[dst_beg + b * dst_n + i] = reduce_window<T, M>(values + src_beg + b * src_n, s, e);
}

template<typename T>
__device__ __forceinline__ T add_shift(T a, T b, int s)
{
  if constexpr (is_fp<T> || sizeof(T) >= 8) {
    return (T)(a + b);
  } else {
    const T mask = (T)((1u << s) - 1);
    return (T)((a >> s) + (b >> s) + (((a & mask) + (b & mask)) >> s));
  }
}

template<typename T>
__global__ __launch_bounds__(256) void fold_k(T* __restrict__ accum, const T* __restrict__ x,
                                              uint64_t n, uint32_t count, int level, int method)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n)
    return;
  const T v = x[g];
  if (count == 0) {
    accum[g] = v;
    return;
  }
  if (method == lod_reduce_mean)
    accum[g] = add_shift<T>(accum[g], v, level);
  else if (method == lod_reduce_min) {
    if (v < accum[g])
      accum[g] = v;
  } else if (method == lod_reduce_max) {
    if (v > accum[g])
      accum[g] = v;
  }
}

template<typename T>
__global__ __launch_bounds__(256) void emit_k(T* __restrict__ dst, const T* __restrict__ accum,
                                              uint64_t n, uint32_t count, int method)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n)
    return;
  if constexpr (is_fp<T>) {
    if (method == lod_reduce_mean) {
      dst[g] = (T)(accum[g] / (T)count);
      return;
    }
  }
  dst[g] = accum[g];
}

template<typename T>
__global__ __launch_bounds__(256) void scatter_k(T* __restrict__ pool, const T* __restrict__ morton,
                                                 const uint32_t* __restrict__ lut,
                                                 const uint64_t* __restrict__ fixed, uint64_t n_l,
                                                 uint64_t total)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total)
    return;
  const uint64_t b = g / n_l, m = g - b * n_l;
  pool[fixed[b] + lut[m]] = morton[g];
}

// One work-item per element; a dispatch counts work-items in 32 bits, which
// lod_init checks against the largest per-epoch launch.
inline dim3 grid_of(uint64_t n) { return dim3((uint32_t)((n + 255) / 256)); }

template<typename T>
hipError_t reduce_launch(int method, T* values, uint64_t sb, uint64_t db, const uint64_t* ends,
                         uint64_t sn, uint64_t dn, uint64_t total, hipStream_t st)
{
  switch (method) {
#define R(M)                                                                                   \
  case M:                                                                                      \
    hipLaunchKernelGGL((reduce_k<T, M>), grid_of(total), dim3(256), 0, st, values, sb, db, ends, \
                       sn, dn, total);                                                         \
    break;
    R(lod_reduce_mean)
    R(lod_reduce_min)
    R(lod_reduce_max)
    R(lod_reduce_median)
    R(lod_reduce_max_suppressed)
    R(lod_reduce_min_suppressed)
#undef R
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// One epoch of the LOD path for element type T.
template<typename T>
int run_epoch_t(tile_stream_gpu* s, uint32_t e, uint32_t* out_mask)
{
  const Plan& p = s->plan;
  hipStream_t st = s->st_comp;
  T* lod = (T*)s->d_lod;
  const uint64_t n0 = p.levels[0].lod_nelem;
  const uint64_t tot0 = p.fixed_count * n0;
  hipLaunchKernelGGL(gather_k<T>, grid_of(tot0), dim3(256), 0, st, (const T*)s->d_linear, lod,
                     s->d_gather_lut, s->d_gather_fixed, n0, tot0);
  if (hipGetLastError() != hipSuccess)
    return 1;
  for (int l = 0; l + 1 < p.nlod; ++l) {
    const uint64_t sn = p.levels[l].lod_nelem, dn = p.levels[l + 1].lod_nelem;
    if (reduce_launch<T>(p.config.reduce_method, lod, p.level_span_begin(l), p.level_span_begin(l + 1),
                         s->d_ends + p.segment_begin(l), sn, dn, p.fixed_count * dn, st) != hipSuccess)
      return 1;
  }
  uint32_t mask = p.append_downsample ? 1u : (uint32_t)((1ull << p.nlod) - 1);
  if (p.append_downsample) {
    T* acc = (T*)s->d_accum;
    for (int lv = 1; lv < p.nlod; ++lv) {
      const uint64_t n = p.fixed_count * p.levels[lv].lod_nelem;
      hipLaunchKernelGGL(fold_k<T>, grid_of(n), dim3(256), 0, st, acc + p.append_accum_offset(lv),
                         lod + p.level_span_begin(lv), n, s->append_counts[lv], lv,
                         (int)p.config.append_reduce_method);
    }
    for (int lv = 1; lv < p.nlod; ++lv) {
      s->append_counts[lv]++;
      if (s->append_counts[lv] >= (1u << lv)) {
        const uint64_t n = p.fixed_count * p.levels[lv].lod_nelem;
        hipLaunchKernelGGL(emit_k<T>, grid_of(n), dim3(256), 0, st, lod + p.level_span_begin(lv),
                           (const T*)(acc + p.append_accum_offset(lv)), n, s->append_counts[lv],
                           (int)p.config.append_reduce_method);
        s->append_counts[lv] = 0;
        mask |= 1u << lv;
      }
    }
  }
  T* pool = (T*)(s->d_pool + (uint64_t)e * s->epoch_pool_bytes);
  for (int lv = 0; lv < p.nlod; ++lv) {
    if (!(mask & (1u << lv)))
      continue;
    const uint64_t n = p.levels[lv].lod_nelem, tot = p.fixed_count * n;
    hipLaunchKernelGGL(scatter_k<T>, grid_of(tot), dim3(256), 0, st, pool,
                       (const T*)(lod + p.level_span_begin(lv)), s->d_chunk_lut[lv],
                       s->d_chunk_fixed[lv], n, tot);
  }
  if (hipGetLastError() != hipSuccess)
    return 1;
  *out_mask = mask;
  return 0;
}

template<typename T>
int drain_t(tile_stream_gpu* s, uint32_t* out_mask)
{
  const Plan& p = s->plan;
  hipStream_t st = s->st_comp;
  T* lod = (T*)s->d_lod;
  T* acc = (T*)s->d_accum;
  uint32_t mask = 0;
  for (int lv = 1; lv < p.nlod; ++lv) {
    if (s->append_counts[lv] == 0)
      continue;
    const uint64_t n = p.levels[lv].lod_nelem, tot = p.fixed_count * n;
    hipLaunchKernelGGL(emit_k<T>, grid_of(tot), dim3(256), 0, st, lod + p.level_span_begin(lv),
                       (const T*)(acc + p.append_accum_offset(lv)), tot, s->append_counts[lv],
                       (int)p.config.append_reduce_method);
    s->append_counts[lv] = 0;
    hipLaunchKernelGGL(scatter_k<T>, grid_of(tot), dim3(256), 0, st, (T*)s->d_pool,
                       (const T*)(lod + p.level_span_begin(lv)), s->d_chunk_lut[lv],
                       s->d_chunk_fixed[lv], n, tot);
    mask |= 1u << lv;
  }
  if (hipGetLastError() != hipSuccess)
    return 1;
  *out_mask = mask;
  return 0;
}

#define DTYPE_DISPATCH(dt, CALL)                                                               \
  switch (dt) {                                                                                \
    case dtype_u8: return CALL(uint8_t);                                                       \
    case dtype_u16: return CALL(uint16_t);                                                     \
    case dtype_u32: return CALL(uint32_t);                                                     \
    case dtype_u64: return CALL(uint64_t);                                                     \
    case dtype_i8: return CALL(int8_t);                                                        \
    case dtype_i16: return CALL(int16_t);                                                      \
    case dtype_i32: return CALL(int32_t);                                                      \
    case dtype_i64: return CALL(int64_t);                                                      \
    case dtype_f16: return CALL(_Float16);                                                     \
    case dtype_f32: return CALL(float);                                                        \
    case dtype_f64: return CALL(double);                                                       \
  }                                                                                            \
  return 1

template<typename T>
int upload(tile_stream_gpu* s, T** d, const std::vector<T>& h)
{
  void* p = nullptr;
  const size_t bytes = h.size() * sizeof(T);
  if (hipMalloc(&p, bytes + 256) != hipSuccess) {
    set_error("lod: hipMalloc failed");
    return 1;
  }
  s->allocations.push_back(p);
  // copies go through the stream's own (non-blocking) queue: legacy-stream
  // operations are refused while another thread captures a graph
  if (bytes && (hipMemcpyAsync(p, h.data(), bytes, hipMemcpyHostToDevice, s->st_comp) != hipSuccess ||
                hipStreamSynchronize(s->st_comp) != hipSuccess)) {
    set_error("lod: upload failed");
    return 1;
  }
  *d = (T*)p;
  return 0;
}

} // namespace

int lod_init(tile_stream_gpu* s)
{
  const Plan& p = s->plan;
  if (p.level_span_ends[p.nlod - 1] >= (1ull << 32) - 256) {
    set_error("lod: an epoch of more than 2^32 LOD elements is not supported");
    return 1;
  }
  build_lod_luts(p, &s->luts);
  const size_t bpe = p.bpe;
  void* lin = nullptr;
  void* lod = nullptr;
  if (hipMalloc(&lin, p.levels[0].layout.epoch_elements * bpe + 256) != hipSuccess ||
      (s->allocations.push_back(lin), false) ||
      hipMalloc(&lod, p.level_span_ends[p.nlod - 1] * bpe + 256) != hipSuccess) {
    set_error("lod: allocation failed");
    return 1;
  }
  s->allocations.push_back(lod);
  s->d_linear = (uint8_t*)lin;
  s->d_lod = (uint8_t*)lod;
  if (hipMemsetAsync(lin, 0, p.levels[0].layout.epoch_elements * bpe, s->st_comp) != hipSuccess ||
      hipMemsetAsync(lod, 0, p.level_span_ends[p.nlod - 1] * bpe, s->st_comp) != hipSuccess)
    return 1;
  if (p.append_downsample && p.append_accum_elements()) {
    void* acc = nullptr;
    if (hipMalloc(&acc, p.append_accum_elements() * bpe + 256) != hipSuccess)
      return 1;
    s->allocations.push_back(acc);
    s->d_accum = (uint8_t*)acc;
    if (hipMemsetAsync(acc, 0, p.append_accum_elements() * bpe, s->st_comp) != hipSuccess)
      return 1;
  }
  if (upload(s, &s->d_gather_lut, s->luts.gather_lut) || upload(s, &s->d_gather_fixed, s->luts.gather_fixed))
    return 1;
  std::vector<uint64_t> ends(p.ends.begin(), p.ends.end());
  if (ends.empty())
    ends.push_back(0);
  if (upload(s, &s->d_ends, ends))
    return 1;
  s->d_chunk_lut.assign(p.nlod, nullptr);
  s->d_chunk_fixed.assign(p.nlod, nullptr);
  for (int lv = 0; lv < p.nlod; ++lv)
    if (upload(s, &s->d_chunk_lut[lv], s->luts.chunk_lut[lv]) ||
        upload(s, &s->d_chunk_fixed[lv], s->luts.chunk_fixed[lv]))
      return 1;
  return 0;
}

int lod_run_epoch(tile_stream_gpu* s, uint32_t e, uint32_t* out_mask)
{
#define CALL(T) run_epoch_t<T>(s, e, out_mask)
  DTYPE_DISPATCH(s->plan.config.dtype, CALL);
#undef CALL
}

int lod_drain(tile_stream_gpu* s, uint32_t* out_mask)
{
#define CALL(T) drain_t<T>(s, out_mask)
  DTYPE_DISPATCH(s->plan.config.dtype, CALL);
#undef CALL
}

} // namespace chucky
