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
// path. The pass is HBM-bound integer/byte work.
//
// gather and scatter move whole tiles: an aligned 2^s cube of the LOD dims
// (1024 cells in 1-D and 2-D, 512 in 3-D) that lies inside the box (and, in the
// chunk pool, inside one chunk) is a contiguous run of Morton ranks, and its
// cells sit at origin + sum(c_k * stride_k) on the array side. A workgroup
// stages one tile in LDS so that both the strided side (row by row) and the
// Morton side (one run) are read and written in whole rows. Cells outside full
// tiles go through a per-cell (rank, offset) list. The host builds both lists
// and checks every tile against the rank LUT.
//
// reduce: a block of 256 outputs whose windows all have 2^lod_ndim cells (the
// interior of the box) skips the window-end lookups and reduces fixed-length
// runs; ragged blocks read the ends table.
#include "kernels.hpp"
#include "stream_gpu.hpp"


#include <algorithm>
#include <cstdlib>
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

// Float mean emit: the accumulated sum over `count` epochs, rounded once to
// T. Halfs divide in float (correctly rounded) and convert to half, which is
// the correctly rounded half quotient (float carries more than 2x11+2 bits).
template<typename T>
__device__ __forceinline__ T div_count(T a, uint32_t count)
{
  if constexpr (std::is_same<T, _Float16>::value)
    return (T)((float)a / (float)count);
  else
    return (T)(a / (T)count);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct TileGeom
{
  uint32_t shift[3];  // position of coordinate k in the row-major tile index
  int64_t stride[3];  // offset step of row-major position q (0 = slowest)
};

// Local Z index -> row-major index: coordinate k takes bits k, k+D, k+2D, ...
template<int D, int S>
__device__ __forceinline__ uint32_t z_to_row(uint32_t z, const TileGeom& g)
{
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < S; ++j)
      c |= ((z >> (j * D + k)) & 1u) << j;
    r |= c << g.shift[k];
  }
  return r;
}

template<int D, int S>
__device__ __forceinline__ int64_t row_offset(uint32_t r, const TileGeom& g)
{
  int64_t off = 0;
#pragma unroll
  for (int q = 0; q < D; ++q)
    off += (int64_t)((r >> (S * (D - 1 - q))) & ((1u << S) - 1)) * g.stride[q];
  return off;
}

// One tile per workgroup and batch: ToMorton gathers array -> Morton run,
// otherwise scatters Morton run -> array. U is an unsigned integer of the
// element size (the pass only moves bytes). Both sides move in 16-byte pieces
// when aligned: the array side when VecArray (the host found every row start
// aligned, and the base pointer is), the Morton side when its run start is.
// A piece of V consecutive Z indices maps to rows z_to_row(z) + z_to_row(j),
// j < V, since the bits of z (a multiple of V) and of j are disjoint. LDS
// rows are padded by one piece, which spreads the Z-order accesses of a wave
// over the banks; the padded index is linear over disjoint bits as well.
template<typename U, int D, bool ToMorton, bool VecArray>
__device__ __forceinline__ void tiles_body(U* __restrict__ morton, U* __restrict__ array,
                                           const uint32_t* __restrict__ tiles,
                                           const uint64_t* __restrict__ fixed, uint64_t n, uint32_t nb,
                                           const TileGeom& g, uint32_t tile)
{
  constexpr int S = lod_tile_side_log(D, sizeof(U));
  constexpr uint32_t cells = 1u << (S * D);
  constexpr uint32_t V = 16 / sizeof(U);
  constexpr uint32_t pieces = cells / V;
  constexpr uint32_t kPieceIters = (pieces + 255) / 256;
  constexpr uint32_t pad = D > 1 ? V : 0;
  static_assert(cells % 256 == 0, "a tile spans whole workgroups");
  __shared__ __align__(16) U buf[cells + (cells >> S) * pad];
  auto at = [](uint32_t r) { return r + (r >> S) * pad; };
  uint32_t jr[V];
#pragma unroll
  for (uint32_t j = 0; j < V; ++j)
    jr[j] = at(z_to_row<D, S>(j, g));
  const uint32_t base = tiles[2 * tile], origin = tiles[2 * tile + 1];
  for (uint32_t b = blockIdx.y; b < nb; b += gridDim.y) {
    U* m = morton + (uint64_t)b * n + base;
    U* a = array + fixed[b] + origin;
    const bool mvec = ((uintptr_t)m & 15) == 0;
    if constexpr (ToMorton) {
      if constexpr (VecArray) {
#pragma unroll
        for (uint32_t i = 0; i < kPieceIters; ++i) {
          const uint32_t u = i * 256 + threadIdx.x;
          if (pieces % 256 == 0 || u < pieces)
            *(u32x4*)&buf[at(u * V)] = *(const u32x4*)(a + row_offset<D, S>(u * V, g));
        }
      } else {
#pragma unroll
        for (uint32_t i = 0; i < cells / 256; ++i) {
          const uint32_t r = i * 256 + threadIdx.x;
          buf[at(r)] = a[row_offset<D, S>(r, g)];
        }
      }
      __syncthreads();
      if (mvec) {
#pragma unroll
        for (uint32_t i = 0; i < kPieceIters; ++i) {
          const uint32_t u = i * 256 + threadIdx.x;
          if (pieces % 256 != 0 && u >= pieces)
            break;
          const uint32_t zr = at(z_to_row<D, S>(u * V, g));
          U e[V];
#pragma unroll
          for (uint32_t j = 0; j < V; ++j)
            e[j] = buf[zr + jr[j]];
          u32x4 w;
          __builtin_memcpy(&w, e, 16);
          // plain stores: the reduce reads level 0 next, from the caches
          *(u32x4*)(m + u * V) = w;
        }
      } else {
#pragma unroll
        for (uint32_t i = 0; i < cells / 256; ++i) {
          const uint32_t z = i * 256 + threadIdx.x;
          m[z] = buf[at(z_to_row<D, S>(z, g))];
        }
      }
    } else {
      if (mvec) {
#pragma unroll
        for (uint32_t i = 0; i < kPieceIters; ++i) {
          const uint32_t u = i * 256 + threadIdx.x;
          if (pieces % 256 != 0 && u >= pieces)
            break;
          const uint32_t zr = at(z_to_row<D, S>(u * V, g));
          const u32x4 w = *(const u32x4*)(m + u * V);
          U e[V];
          __builtin_memcpy(e, &w, 16);
#pragma unroll
          for (uint32_t j = 0; j < V; ++j)
            buf[zr + jr[j]] = e[j];
        }
      } else {
#pragma unroll
        for (uint32_t i = 0; i < cells / 256; ++i) {
          const uint32_t z = i * 256 + threadIdx.x;
          buf[at(z_to_row<D, S>(z, g))] = m[z];
        }
      }
      __syncthreads();
      if constexpr (VecArray) {
#pragma unroll
        for (uint32_t i = 0; i < kPieceIters; ++i) {
          const uint32_t u = i * 256 + threadIdx.x;
          // the chunk pool is read back only by the compressor / aggregate,
          // later: streamed out with nontemporal stores (measured faster)
          if (pieces % 256 == 0 || u < pieces)
            __builtin_nontemporal_store(*(const u32x4*)&buf[at(u * V)], (u32x4*)(a + row_offset<D, S>(u * V, g)));
        }
      } else {
#pragma unroll
        for (uint32_t i = 0; i < cells / 256; ++i) {
          const uint32_t r = i * 256 + threadIdx.x;
          a[row_offset<D, S>(r, g)] = buf[at(r)];
        }
      }
    }
    __syncthreads();
  }
}

template<typename U, int D, bool ToMorton, bool VecArray>
__global__ __launch_bounds__(256) void tiles_k(U* __restrict__ morton, U* __restrict__ array,
                                               const uint32_t* __restrict__ tiles,
                                               const uint64_t* __restrict__ fixed, uint64_t n,
                                               uint32_t nb, TileGeom g)
{
  tiles_body<U, D, ToMorton, VecArray>(morton, array, tiles, fixed, n, nb, g, blockIdx.x);
}

// The Morton -> chunk-pool scatter of several levels in one launch (levels
// whose tiles share one geometry): workgroup x takes tile x - first[k] of
// level k, the level whose range [first[k], first[k + 1]) holds x.
constexpr int kMoveMax = LOD_MAX_LEVELS;

template<typename U>
struct MoveLevels
{
  U* morton[kMoveMax];
  const uint32_t* tiles[kMoveMax];
  const uint64_t* fixed[kMoveMax];
  uint64_t n[kMoveMax];
  uint32_t first[kMoveMax + 1];
  int nlev;
};

template<typename U, int D, bool VecArray>
__global__ __launch_bounds__(256) void tiles_levels_k(MoveLevels<U> m, U* __restrict__ array, uint32_t nb,
                                                      TileGeom g)
{
  int k = 0;
  while (k + 1 < m.nlev && blockIdx.x >= m.first[k + 1])
    ++k;
  tiles_body<U, D, false, VecArray>(m.morton[k], array, m.tiles[k], m.fixed[k], m.n[k], nb, g,
                                    blockIdx.x - m.first[k]);
}

// Cells outside full tiles: one (rank, offset) pair per work-item.
template<typename T, bool ToMorton>
__global__ __launch_bounds__(256) void edge_k(T* __restrict__ morton, T* __restrict__ array,
                                              const uint32_t* __restrict__ edge, uint32_t ne,
                                              const uint64_t* __restrict__ fixed, uint64_t n,
                                              uint32_t nb)
{
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ne)
    return;
  const uint32_t mi = edge[2 * i], off = edge[2 * i + 1];
  for (uint32_t b = blockIdx.y; b < nb; b += gridDim.y) {
    if constexpr (ToMorton)
      morton[(uint64_t)b * n + mi] = array[fixed[b] + off];
    else
      array[fixed[b] + off] = morton[(uint64_t)b * n + mi];
  }
}

// Reduce the run v[0, len). With len a compile-time constant (the regular
// path) the loops unroll; the operations and their order are the same.
template<typename T, int M>
__device__ __forceinline__ T reduce_run(const T* v, uint32_t len)
{
  const uint32_t s = 0, e = len;
  if constexpr (M == lod_reduce_mean) {
    using A = typename Acc<T>::type;
    A sum = 0;
    for (uint32_t j = s; j < e; ++j)
      sum += (A)v[j];
    return (T)(sum / (A)len);
  } else if constexpr (M == lod_reduce_min) {
    T best = v[s];
    for (uint32_t j = s + 1; j < e; ++j)
      if (v[j] < best)
        best = v[j];
    return best;
  } else if constexpr (M == lod_reduce_max) {
    T best = v[s];
    for (uint32_t j = s + 1; j < e; ++j)
      if (v[j] > best)
        best = v[j];
    return best;
  } else if constexpr (M == lod_reduce_median) {
    T buf[16];
    const uint32_t n = len <= 16 ? len : 16;
    for (uint32_t j = 0; j < n; ++j)
      buf[j] = v[s + j];
    for (uint32_t i = 1; i < n; ++i) {
      const T key = buf[i];
      uint32_t k = i;
      while (k > 0 && buf[k - 1] > key) {
        buf[k] = buf[k - 1];
        --k;
      }
      buf[k] = key;
    }
    return buf[n / 2];
  } else if constexpr (M == lod_reduce_max_suppressed) {
    T t1 = v[s], t2 = v[s];
    for (uint32_t j = s + 1; j < e; ++j) {
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
    for (uint32_t j = s + 1; j < e; ++j) {
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

// reduce_run over a register window x[0, W) with compile-time indices, so it
// stays in registers. The median sorts by compare-and-swap passes that make
// exactly the moves of the insertion sort (each pass carries the new element
// down while its left neighbour is greater; the prefix has no adjacent
// inversion, so the pass makes no further swap).
template<typename T, int M, int W>
__device__ __forceinline__ T reduce_fixed(const T* x)
{
  if constexpr (M == lod_reduce_median) {
    T buf[W];
#pragma unroll
    for (int j = 0; j < W; ++j)
      buf[j] = x[j];
#pragma unroll
    for (int i = 1; i < W; ++i)
#pragma unroll
      for (int k = i; k > 0; --k)
        if (buf[k - 1] > buf[k]) {
          const T t = buf[k - 1];
          buf[k - 1] = buf[k];
          buf[k] = t;
        }
    return buf[W / 2];
  } else {
    return reduce_run<T, M>(x, W);
  }
}

// N bytes between pointers aligned to min(N, 16) (the callers check), as
// memcpy so that register arrays of any element type may take part (no
// type punning); it compiles to 16-byte or narrower vector moves.
template<int N>
__device__ __forceinline__ void copy_bytes(void* __restrict__ dst, const void* __restrict__ src)
{
  constexpr int A = N < 16 ? N : 16;
  __builtin_memcpy(__builtin_assume_aligned(dst, A), __builtin_assume_aligned(src, A), N);
}

constexpr uint32_t outputs_per_lane(int W, size_t sz)
{
  return W == 0 || W * sz >= 32 ? 1u : (uint32_t)(32 / (W * sz));
}

// Level l -> l+1 for dst_n outputs per batch; W = 2^lod_ndim. A workgroup owns
// 256*P consecutive outputs. When all their windows have W cells (the
// interior) each lane reduces P consecutive windows read as one run of
// W*P elements (32 bytes or more, in 16-byte pieces when aligned); otherwise
// lanes take outputs i0 + p*256 + lane and read the window ends.
template<typename T, int M, int W>
__global__ __launch_bounds__(256) void reduce_k(T* __restrict__ values, uint64_t src_beg,
                                                uint64_t dst_beg, const uint64_t* __restrict__ ends,
                                                uint32_t src_n, uint32_t dst_n, uint32_t nb)
{
  constexpr uint32_t P = outputs_per_lane(W, sizeof(T));
  const uint32_t i0 = blockIdx.x * (256 * P);
  const uint32_t i1 = min(i0 + 256 * P, dst_n);
  const uint64_t s0 = i0 ? ends[i0 - 1] : 0;
  // windows hold 1..W cells, so W cells per window on average means all W
  const bool regular = W && ends[i1 - 1] - s0 == (uint64_t)W * (i1 - i0);
  if (regular) {
    if constexpr (W > 0) {
      const uint32_t ib = i0 + threadIdx.x * P;
      if (ib >= i1)
        return;
      const uint64_t s = s0 + (uint64_t)W * (ib - i0);
      const uint32_t cnt = min(P, i1 - ib);
      constexpr int NIN = W * P * sizeof(T), NOUT = P * sizeof(T);
      for (uint32_t b = blockIdx.y; b < nb; b += gridDim.y) {
        const T* v = values + src_beg + (uint64_t)b * src_n + s;
        T* o = values + dst_beg + (uint64_t)b * dst_n + ib;
        if (cnt == P && ((uintptr_t)v & 15) == 0 && ((uintptr_t)o & (min(NOUT, 16) - 1)) == 0) {
          alignas(16) T x[W * P];
          alignas(16) T y[P];
          copy_bytes<NIN>(x, v);
#pragma unroll
          for (uint32_t p = 0; p < P; ++p)
            y[p] = reduce_fixed<T, M, W>(x + p * W);
          copy_bytes<NOUT>(o, y);
        } else {
          for (uint32_t p = 0; p < cnt; ++p) {
            T x[W];
#pragma unroll
            for (int j = 0; j < W; ++j)
              x[j] = v[p * W + j];
            o[p] = reduce_fixed<T, M, W>(x);
          }
        }
      }
    }
    return;
  }
#pragma unroll
  for (uint32_t p = 0; p < P; ++p) {
    const uint32_t i = i0 + p * 256 + threadIdx.x;
    if (i >= i1)
      break;
    const uint64_t s = i ? ends[i - 1] : 0;
    const uint32_t len = (uint32_t)(ends[i] - s);
    for (uint32_t b = blockIdx.y; b < nb; b += gridDim.y)
      values[dst_beg + (uint64_t)b * dst_n + i] =
          reduce_run<T, M>(values + src_beg + (uint64_t)b * src_n + s, len);
  }
}

// Several level pairs l -> l+1 in one launch. Windows partition every
// level, so the top-level outputs [k0, k1) of a workgroup own one span per
// level below them, disjoint from every other workgroup's. A full workgroup
// whose windows all have W cells (the interior; at most W each, so W^F cells
// for each top output means all) is regular: every lane loads 64 contiguous
// bytes of level 0 into registers and reduces them through as many levels as
// they cover, writing each level's values straight out; the values left per
// lane go to LDS for the levels above. Intermediate levels are never read
// back from memory. A ragged workgroup (edges, the last one of a batch)
// reduces level by level through the window ends, in memory. Each window
// reduces as in reduce_k.
constexpr int kPyrMax = 6; // level pairs one launch fuses at most

template<typename T>
constexpr uint32_t pyr_cells() { return 64 / sizeof(T); } // level-0 cells per lane

template<typename T>
constexpr uint32_t pyr_span() { return 256 * pyr_cells<T>(); } // level-0 cells per workgroup

// values a lane holds after reducing its cells through every level they cover
template<typename T, int W>
constexpr uint32_t pyr_rest()
{
  uint32_t c = pyr_cells<T>();
  while (c >= (uint32_t)W)
    c /= (uint32_t)W;
  return c;
}

struct PyrArgs
{
  uint64_t beg[kPyrMax + 1]; // element offset of each level's values
  uint64_t ends[kPyrMax];    // offset of the l -> l+1 window ends
  uint32_t n[kPyrMax + 1];   // elements per batch of each level
  uint32_t levels;           // level pairs fused (F)
  uint32_t group;            // top-level outputs per workgroup: span / W^F
};

// n values of x to dst (16-byte pieces, or narrower, when aligned).
template<typename T, uint32_t N>
__device__ __forceinline__ void store_run(T* dst, const T* x)
{
  constexpr uint32_t B = N * sizeof(T);
  constexpr uint32_t A = B < 16 ? B : 16;
  if (((uintptr_t)dst & (A - 1)) == 0) {
    copy_bytes<B>(dst, x);
  } else {
#pragma unroll
    for (uint32_t i = 0; i < N; ++i)
      dst[i] = x[i];
  }
}

template<typename T, int M, int W>
__global__ __launch_bounds__(256) void pyramid_k(T* __restrict__ values, const uint64_t* __restrict__ ends,
                                                 PyrArgs a, uint32_t nb)
{
  constexpr uint32_t C = pyr_cells<T>(), SPAN = pyr_span<T>();
  __shared__ alignas(16) T buf0[256 * W];
  __shared__ alignas(16) T buf1[256 * W / 2];
  __shared__ uint32_t lo[kPyrMax + 1], hi[kPyrMax + 1];
  const uint32_t F = a.levels;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    uint32_t l0 = blockIdx.x * a.group, h0 = min(l0 + a.group, a.n[F]);
    lo[F] = l0;
    hi[F] = h0;
    for (int l = (int)F - 1; l >= 0; --l) {
      const uint64_t* e = ends + a.ends[l];
      l0 = l0 ? (uint32_t)e[l0 - 1] : 0u;
      h0 = (uint32_t)e[h0 - 1];
      lo[l] = l0;
      hi[l] = h0;
    }
  }
  __syncthreads();
  const bool regular = hi[0] - lo[0] == SPAN && hi[F] - lo[F] == a.group;
  for (uint32_t b = blockIdx.y; b < nb; b += gridDim.y) {
    if (regular) {
      const T* src = values + a.beg[0] + (uint64_t)b * a.n[0] + lo[0] + tid * C;
      alignas(16) T x[C];
      if (((uintptr_t)src & 15) == 0) {
        copy_bytes<C * sizeof(T)>(x, src);
      } else {
#pragma unroll
        for (uint32_t i = 0; i < C; ++i)
          x[i] = src[i];
      }
      // levels within the lane's cells, in place: x[0, cnt) holds the
      // current level
      uint32_t done = 0;
#pragma unroll
      for (uint32_t l = 1, cnt = C; l <= kPyrMax; ++l) {
        if (cnt < (uint32_t)W || l > F)
          break;
        cnt /= (uint32_t)W;
#pragma unroll
        for (uint32_t i = 0; i < cnt; ++i)
          x[i] = reduce_fixed<T, M, W>(x + i * W);
        T* dst = values + a.beg[l] + (uint64_t)b * a.n[l] + lo[l] + tid * cnt;
        switch (cnt) {
          case 1: store_run<T, 1>(dst, x); break;
          case 2: store_run<T, 2>(dst, x); break;
          case 4: store_run<T, 4>(dst, x); break;
          case 8: store_run<T, 8>(dst, x); break;
          case 16: store_run<T, 16>(dst, x); break;
          default: store_run<T, 32>(dst, x); break;
        }
        done = l;
      }
      if (done == F)
        continue;
      // the levels above (the lane's cells are used up: R values remain):
      // the lanes' values in LDS, W of them per output
      constexpr uint32_t R = pyr_rest<T, W>();
#pragma unroll
      for (uint32_t i = 0; i < R; ++i)
        buf0[tid * R + i] = x[i];
      uint32_t n_in = 256 * R;
      __syncthreads();
      const T* in = buf0;
      T* out = buf1;
      for (uint32_t l = done + 1; l <= F; ++l) {
        const uint32_t n_out = n_in / (uint32_t)W;
        T* dst = values + a.beg[l] + (uint64_t)b * a.n[l] + lo[l];
        for (uint32_t i = tid; i < n_out; i += 256) {
          const T y = reduce_fixed<T, M, W>(in + i * W);
          out[i] = y;
          dst[i] = y;
        }
        __syncthreads();
        const T* t = in;
        in = out;
        out = const_cast<T*>(t);
        n_in = n_out;
      }
    } else {
      for (uint32_t l = 1; l <= F; ++l) {
        const T* src = values + a.beg[l - 1] + (uint64_t)b * a.n[l - 1];
        T* dst = values + a.beg[l] + (uint64_t)b * a.n[l];
        const uint64_t* e = ends + a.ends[l - 1];
        for (uint32_t i = lo[l] + tid; i < hi[l]; i += 256) {
          const uint64_t s0 = i ? e[i - 1] : 0;
          dst[i] = reduce_run<T, M>(src + s0, (uint32_t)(e[i] - s0));
        }
        // the next level reads what this workgroup just wrote
        __threadfence_block();
        __syncthreads();
      }
    }
  }
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
__global__ __launch_bounds__(256) void emit_k(T* __restrict__ dst, const T* __restrict__ accum,
                                              uint64_t n, uint32_t count, int method)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n)
    return;
  if constexpr (is_fp<T>) {
    if (method == lod_reduce_mean) {
      dst[g] = div_count<T>(accum[g], count);
      return;
    }
  }
  dst[g] = accum[g];
}

// The append-dim fold of every level >= 1 in one launch: level lv's
// elements follow the previous levels' (end[] holds the running totals).
// Each element folds the epoch's value into its accumulator (the first
// epoch of a window starts it) and, when the level's window of 2^lv epochs
// completes, also writes the emitted value over the level's Morton values:
// the same operations as fold_k followed by emit_k.
constexpr int kFoldMax = LOD_MAX_LEVELS;

template<typename T>
struct FoldLevels
{
  T* x[kFoldMax];         // the level's Morton values (the emitted value goes here)
  T* acc[kFoldMax];       // its accumulator
  uint64_t end[kFoldMax]; // running element totals
  uint32_t count[kFoldMax];
  uint32_t emit[kFoldMax]; // epochs in the window when it completes now, else 0
  int level[kFoldMax];
  int nlev;
};

template<typename T>
__global__ __launch_bounds__(256) void fold_emit_k(FoldLevels<T> f, int method)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  int k = 0;
  while (k < f.nlev && g >= f.end[k])
    ++k;
  if (k >= f.nlev)
    return;
  const uint64_t i = g - (k ? f.end[k - 1] : 0);
  const T v = f.x[k][i];
  T a = v;
  if (f.count[k]) {
    a = f.acc[k][i];
    if (method == lod_reduce_mean)
      a = add_shift<T>(a, v, f.level[k]);
    else if (method == lod_reduce_min)
      a = v < a ? v : a;
    else if (method == lod_reduce_max)
      a = v > a ? v : a;
  }
  f.acc[k][i] = a;
  if (f.emit[k]) {
    T out = a;
    if constexpr (is_fp<T>) {
      if (method == lod_reduce_mean)
        out = div_count<T>(a, f.emit[k]);
    }
    f.x[k][i] = out;
  }
}

// Element-wise kernels count work-items in 32 bits, which lod_init checks
// against the largest per-epoch launch.
inline dim3 grid_of(uint64_t n) { return dim3((uint32_t)((n + 255) / 256)); }

inline uint32_t batch_grid(uint64_t nb) { return (uint32_t)std::min<uint64_t>(nb, 65535); }

// A launch that records the timing events it is given (either may be null):
// `start` on the stream right before the kernel, `stop` right after it, so
// the pair brackets the kernel's execution. (hipExtLaunchKernelGGL's
// in-kernel events measure the same span for about twice the host time per
// launch; a host-paced pipeline of small epochs feels that.) A failed record
// leaves its error for the caller's hipGetLastError.
template<typename K, typename... A>
void launch(K kernel, dim3 grid, dim3 block, hipStream_t st, hipEvent_t start, hipEvent_t stop, A... args)
{
  if (start)
    (void)hipEventRecord(start, st);
  hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
  if (stop)
    (void)hipEventRecord(stop, st);
}

TileGeom geom_of(const LodTiles& t)
{
  // row-major order puts the coordinate with the smallest stride fastest
  int order[3] = { 0, 1, 2 };
  std::stable_sort(order, order + t.ndim, [&](int a, int b) {
    return std::llabs(t.stride[a]) > std::llabs(t.stride[b]);
  });
  TileGeom g{};
  for (int q = 0; q < t.ndim; ++q) {
    g.shift[order[q]] = (uint32_t)(t.side_log * (t.ndim - 1 - q));
    g.stride[q] = t.stride[order[q]];
  }
  return g;
}

template<size_t N> struct Bits;
template<> struct Bits<1> { using type = uint8_t; };
template<> struct Bits<2> { using type = uint16_t; };
template<> struct Bits<4> { using type = uint32_t; };
template<> struct Bits<8> { using type = uint64_t; };

// Move one level between its Morton run (n cells per batch) and the array.
// `start`/`stop` are recorded by the first/last kernel launched.
template<typename T, bool ToMorton>
hipError_t move_launch(const LodTiles& t, const uint32_t* d_tiles, const uint32_t* d_edge, T* morton_t,
                       T* array_t, const uint64_t* fixed, uint64_t n, uint64_t nb, hipStream_t st,
                       hipEvent_t start = nullptr, hipEvent_t stop = nullptr)
{
  using U = typename Bits<sizeof(T)>::type;
  U* morton = (U*)morton_t;
  U* array = (U*)array_t;
  if (nb == 0)
    return hipSuccess;
  const uint32_t ntiles = (uint32_t)(t.tiles.size() / 2), nedge = (uint32_t)(t.edge.size() / 2);
  const uint32_t gy = batch_grid(nb);
  if (ntiles) {
    const TileGeom g = geom_of(t);
    const bool vec = t.vec && ((uintptr_t)array & 15) == 0;
    hipEvent_t e1 = nedge ? nullptr : stop;
    switch (t.ndim) {
#define K(D)                                                                                     \
  case D:                                                                                        \
    if (vec)                                                                                     \
      launch(tiles_k<U, D, ToMorton, true>, dim3(ntiles, gy), dim3(256), st, start, e1, morton,  \
             array, d_tiles, fixed, n, (uint32_t)nb, g);                                         \
    else                                                                                         \
      launch(tiles_k<U, D, ToMorton, false>, dim3(ntiles, gy), dim3(256), st, start, e1, morton, \
             array, d_tiles, fixed, n, (uint32_t)nb, g);                                         \
    break;
      K(1)
      K(2)
      K(3)
#undef K
      default: return hipErrorInvalidValue;
    }
  }
  if (nedge)
    launch(edge_k<U, ToMorton>, dim3((nedge + 255) / 256, gy), dim3(256), st, ntiles ? nullptr : start,
           stop, morton, array, d_edge, nedge, fixed, n, (uint32_t)nb);
  return hipGetLastError();
}

// Morton -> chunk-pool scatter of the levels in `lvs`: one tiles_levels_k
// launch when their tiles share one geometry (the usual case: every level's
// chunks have the same shape), then the edge cells level by level; otherwise
// one move_launch per level.
template<typename T>
hipError_t scatter_levels(tile_stream_gpu* s, const std::vector<int>& lvs, T* lod, T* pool, hipStream_t st,
                          hipEvent_t start, hipEvent_t stop)
{
  using U = typename Bits<sizeof(T)>::type;
  const Plan& p = s->plan;
  const uint64_t nb = p.fixed_count;
  if (lvs.empty() || nb == 0)
    return hipSuccess;
  const LodTiles& t0 = s->luts.chunk_tiles[lvs[0]];
  bool same = lvs.size() <= (size_t)kMoveMax && t0.ndim >= 1 && t0.ndim <= 3;
  for (int lv : lvs) {
    const LodTiles& t = s->luts.chunk_tiles[lv];
    same = same && t.ndim == t0.ndim && t.side_log == t0.side_log && t.vec == t0.vec;
    for (int q = 0; q < t0.ndim && same; ++q)
      same = t.stride[q] == t0.stride[q];
  }
  if (!same) {
    for (size_t i = 0; i < lvs.size(); ++i) {
      const int lv = lvs[i];
      if (move_launch<T, false>(s->luts.chunk_tiles[lv], s->d_chunk_tiles[lv], s->d_chunk_edge[lv],
                                lod + p.level_span_begin(lv), pool, s->d_chunk_fixed[lv], p.levels[lv].lod_nelem,
                                nb, st, i == 0 ? start : nullptr, i + 1 == lvs.size() ? stop : nullptr) !=
          hipSuccess)
        return hipErrorLaunchFailure;
    }
    return hipSuccess;
  }
  MoveLevels<U> m = {};
  uint32_t tiles = 0;
  int last_edge = -1;
  for (size_t i = 0; i < lvs.size(); ++i) {
    const int lv = lvs[i];
    const LodTiles& t = s->luts.chunk_tiles[lv];
    m.morton[i] = (U*)(lod + p.level_span_begin(lv));
    m.tiles[i] = s->d_chunk_tiles[lv];
    m.fixed[i] = s->d_chunk_fixed[lv];
    m.n[i] = p.levels[lv].lod_nelem;
    m.first[i] = tiles;
    tiles += (uint32_t)(t.tiles.size() / 2);
    if (!t.edge.empty())
      last_edge = (int)i;
  }
  m.first[lvs.size()] = tiles;
  m.nlev = (int)lvs.size();
  const uint32_t gy = batch_grid(nb);
  if (tiles) {
    const TileGeom g = geom_of(t0);
    const bool vec = t0.vec && ((uintptr_t)pool & 15) == 0;
    hipEvent_t e1 = last_edge < 0 ? stop : nullptr;
    switch (t0.ndim) {
#define K(D)                                                                                     \
  case D:                                                                                        \
    if (vec)                                                                                     \
      launch(tiles_levels_k<U, D, true>, dim3(tiles, gy), dim3(256), st, start, e1, m, (U*)pool, \
             (uint32_t)nb, g);                                                                   \
    else                                                                                         \
      launch(tiles_levels_k<U, D, false>, dim3(tiles, gy), dim3(256), st, start, e1, m,          \
             (U*)pool, (uint32_t)nb, g);                                                         \
    break;
      K(1)
      K(2)
      K(3)
#undef K
      default: return hipErrorInvalidValue;
    }
  }
  for (size_t i = 0; i < lvs.size(); ++i) {
    const int lv = lvs[i];
    const LodTiles& t = s->luts.chunk_tiles[lv];
    const uint32_t nedge = (uint32_t)(t.edge.size() / 2);
    if (!nedge)
      continue;
    launch(edge_k<U, false>, dim3((nedge + 255) / 256, gy), dim3(256), st, tiles || (int)i > 0 ? nullptr : start,
           (int)i == last_edge ? stop : nullptr, m.morton[i], (U*)pool, s->d_chunk_edge[lv], nedge,
           s->d_chunk_fixed[lv], m.n[i], (uint32_t)nb);
  }
  return hipGetLastError();
}

template<typename T, int W>
hipError_t reduce_launch_w(int method, T* values, uint64_t sb, uint64_t db, const uint64_t* ends,
                           uint64_t sn, uint64_t dn, uint64_t nb, hipStream_t st, hipEvent_t start,
                           hipEvent_t stop)
{
  constexpr uint32_t span = 256 * outputs_per_lane(W, sizeof(T));
  const dim3 grid((uint32_t)((dn + span - 1) / span), batch_grid(nb));
  switch (method) {
#define R(M)                                                                                   \
  case M:                                                                                      \
    launch(reduce_k<T, M, W>, grid, dim3(256), st, start, stop, values, sb, db, ends,           \
           (uint32_t)sn, (uint32_t)dn, (uint32_t)nb);                                          \
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

template<typename T>
hipError_t reduce_launch(int ndim, int method, T* values, uint64_t sb, uint64_t db, const uint64_t* ends,
                         uint64_t sn, uint64_t dn, uint64_t nb, hipStream_t st, hipEvent_t start,
                         hipEvent_t stop)
{
  if (dn == 0 || nb == 0)
    return hipSuccess;
  switch (ndim) {
    case 1: return reduce_launch_w<T, 2>(method, values, sb, db, ends, sn, dn, nb, st, start, stop);
    case 2: return reduce_launch_w<T, 4>(method, values, sb, db, ends, sn, dn, nb, st, start, stop);
    case 3: return reduce_launch_w<T, 8>(method, values, sb, db, ends, sn, dn, nb, st, start, stop);
    // more LOD dims: windows of up to 2^ndim cells; the block check never
    // passes for W = 0, so every block reads the ends table
    default: return reduce_launch_w<T, 0>(method, values, sb, db, ends, sn, dn, nb, st, start, stop);
  }
}

// Every level pair l -> l+1 with l in [l0, l1) by pyramid_k launches of up to
// kPyrMax pairs (W^F <= span / 4, so a workgroup owns at least 4 top
// outputs); for more than 3 LOD dims (no fixed window) the per-level reduce.
// Sets *rin / *rout to the bytes the launches read and write.
template<typename T, int W>
hipError_t pyramid_launch_w(const Plan& p, int method, T* values, const uint64_t* d_ends, int l0, int l1,
                            hipStream_t st, hipEvent_t start, hipEvent_t stop, double* rin, double* rout)
{
  const uint64_t nb = p.fixed_count;
  int l = l0;
  while (l < l1) {
    PyrArgs a = {};
    uint32_t F = 0, wf = 1;
    while (l + (int)F < l1 && (int)F < kPyrMax && wf * W <= pyr_span<T>() / 4) {
      ++F;
      wf *= W;
    }
    a.levels = F;
    a.group = pyr_span<T>() / wf;
    for (uint32_t k = 0; k <= F; ++k) {
      a.beg[k] = p.level_span_begin(l + (int)k);
      a.n[k] = (uint32_t)p.levels[l + k].lod_nelem;
      if (k < F)
        a.ends[k] = p.segment_begin(l + (int)k);
    }
    *rin += (double)nb * a.n[0] * sizeof(T);
    for (uint32_t k = 1; k <= F; ++k)
      *rout += (double)nb * a.n[k] * sizeof(T);
    const dim3 grid((a.n[F] + a.group - 1) / a.group, batch_grid(nb));
    const bool first = l == l0, last = l + (int)F == l1;
    switch (method) {
#define R(M)                                                                                   \
  case M:                                                                                      \
    launch(pyramid_k<T, M, W>, grid, dim3(256), st, first ? start : nullptr, last ? stop : nullptr, \
           values, d_ends, a, (uint32_t)nb);                                                   \
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
    if (hipGetLastError() != hipSuccess)
      return hipErrorLaunchFailure;
    l += (int)F;
  }
  return hipSuccess;
}

template<typename T>
hipError_t pyramid_launch(const Plan& p, int method, T* values, const uint64_t* d_ends, int l0, int l1,
                          hipStream_t st, hipEvent_t start, hipEvent_t stop, double* rin, double* rout)
{
  switch (p.lod_ndim) {
    case 1: return pyramid_launch_w<T, 2>(p, method, values, d_ends, l0, l1, st, start, stop, rin, rout);
    case 2: return pyramid_launch_w<T, 4>(p, method, values, d_ends, l0, l1, st, start, stop, rin, rout);
    case 3: return pyramid_launch_w<T, 8>(p, method, values, d_ends, l0, l1, st, start, stop, rin, rout);
    default: break;
  }
  for (int l = l0; l < l1; ++l) {
    const uint64_t sn = p.levels[l].lod_nelem, dn = p.levels[l + 1].lod_nelem;
    if (reduce_launch<T>(p.lod_ndim, method, values, p.level_span_begin(l), p.level_span_begin(l + 1),
                         d_ends + p.segment_begin(l), sn, dn, p.fixed_count, st, l == l0 ? start : nullptr,
                         l + 1 == l1 ? stop : nullptr) != hipSuccess)
      return hipErrorLaunchFailure;
    *rin += (double)p.fixed_count * sn * sizeof(T);
    *rout += (double)p.fixed_count * dn * sizeof(T);
  }
  return hipSuccess;
}

// One epoch of the LOD path for element type T.
template<typename T>
int run_epoch_t(tile_stream_gpu* s, uint32_t e, uint32_t* out_mask)
{
  const Plan& p = s->plan;
  hipStream_t st = s->st_comp;
  T* lod = (T*)s->d_lod;
  const uint64_t n0 = p.levels[0].lod_nelem;
  T* linear = (T*)(s->lod_src ? s->lod_src : s->d_linear);
  s->linear_in_lod = s->lod_src != nullptr;
  s->lod_src = nullptr;
  const double bpe = (double)sizeof(T), nb = (double)p.fixed_count;
  // each pass is timed by events that its first and last kernel record
  // (start and end of execution) and credited its algorithmic bytes
  hipEvent_t ta = nullptr, tb = nullptr;
  if (timing_pair(s, &ta, &tb) ||
      move_launch<T, true>(s->luts.gather_tiles, s->d_gather_tiles, s->d_gather_edge, lod, linear,
                           s->d_gather_fixed, n0, p.fixed_count, st, ta, tb) != hipSuccess)
    return 1;
  timing_add(s, ta, tb, &stream_metrics::lod_gather, nb * n0 * bpe, nb * n0 * bpe);
  if (s->lod_src_free) {
    // a staging buffer read in place: free for the next host copy once the
    // gather has read it (the rest of the pass works on the Morton levels)
    if (hipEventRecord(s->lod_src_free, st) != hipSuccess)
      return 1;
    s->lod_src_free = nullptr;
  }
  int last_reduce = -1;
  for (int l = 0; l + 1 < p.nlod; ++l)
    if (p.levels[l + 1].lod_nelem)
      last_reduce = l;
  if (last_reduce >= 0) {
    double rin = 0, rout = 0;
    if (timing_pair(s, &ta, &tb) ||
        pyramid_launch<T>(p, p.config.reduce_method, lod, s->d_ends, 0, last_reduce + 1, st, ta, tb, &rin,
                          &rout) != hipSuccess)
      return 1;
    timing_add(s, ta, tb, &stream_metrics::lod_reduce, rin, rout);
  }
  uint32_t mask = p.append_downsample ? 1u : (uint32_t)((1ull << p.nlod) - 1);
  if (p.append_downsample && p.nlod > 1) {
    T* acc = (T*)s->d_accum;
    double fin = 0, fout = 0;
    if (timing_begin(s, &ta))
      return 1;
    FoldLevels<T> f = {};
    uint64_t tot = 0;
    for (int lv = 1; lv < p.nlod; ++lv) {
      const uint64_t n = p.fixed_count * p.levels[lv].lod_nelem;
      if (!n)
        continue;
      const int k = f.nlev++;
      f.x[k] = lod + p.level_span_begin(lv);
      f.acc[k] = acc + p.append_accum_offset(lv);
      tot += n;
      f.end[k] = tot;
      f.count[k] = s->append_counts[lv];
      f.level[k] = lv;
      fin += (s->append_counts[lv] ? 2.0 : 1.0) * n * bpe;
      fout += (double)n * bpe;
      s->append_counts[lv]++;
      if (s->append_counts[lv] >= (1u << lv)) {
        f.emit[k] = s->append_counts[lv];
        fout += (double)n * bpe;
        s->append_counts[lv] = 0;
        mask |= 1u << lv;
      }
    }
    if (tot)
      hipLaunchKernelGGL(fold_emit_k<T>, grid_of(tot), dim3(256), 0, st, f, (int)p.config.append_reduce_method);
    if (timing_end(s, ta, &stream_metrics::lod_append_fold, fin, fout))
      return 1;
  }
  T* pool = (T*)(s->d_pool + (uint64_t)e * s->epoch_pool_bytes);
  double sbytes = 0;
  int first_lv = -1, last_lv = -1;
  for (int lv = 0; lv < p.nlod; ++lv)
    if ((mask & (1u << lv)) && p.levels[lv].lod_nelem) {
      if (first_lv < 0)
        first_lv = lv;
      last_lv = lv;
    }
  if (first_lv >= 0) {
    std::vector<int> lvs;
    for (int lv = first_lv; lv <= last_lv; ++lv)
      if ((mask & (1u << lv)) && p.levels[lv].lod_nelem) {
        lvs.push_back(lv);
        sbytes += nb * p.levels[lv].lod_nelem * bpe;
      }
    if (timing_pair(s, &ta, &tb) || scatter_levels<T>(s, lvs, lod, pool, st, ta, tb) != hipSuccess)
      return 1;
    timing_add(s, ta, tb, &stream_metrics::lod_morton_chunk, sbytes, sbytes);
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
    if (move_launch<T, false>(s->luts.chunk_tiles[lv], s->d_chunk_tiles[lv], s->d_chunk_edge[lv],
                              lod + p.level_span_begin(lv), (T*)s->d_pool, s->d_chunk_fixed[lv], n,
                              p.fixed_count, st) != hipSuccess)
      return 1;
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
  cached_lod_luts(p, &s->luts);
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
  if (upload(s, &s->d_gather_tiles, s->luts.gather_tiles.tiles) ||
      upload(s, &s->d_gather_edge, s->luts.gather_tiles.edge) ||
      upload(s, &s->d_gather_fixed, s->luts.gather_fixed))
    return 1;
  std::vector<uint64_t> ends(p.ends.begin(), p.ends.end());
  if (ends.empty())
    ends.push_back(0);
  if (upload(s, &s->d_ends, ends))
    return 1;
  s->d_chunk_tiles.assign(p.nlod, nullptr);
  s->d_chunk_edge.assign(p.nlod, nullptr);
  s->d_chunk_fixed.assign(p.nlod, nullptr);
  for (int lv = 0; lv < p.nlod; ++lv)
    if (upload(s, &s->d_chunk_tiles[lv], s->luts.chunk_tiles[lv].tiles) ||
        upload(s, &s->d_chunk_edge[lv], s->luts.chunk_tiles[lv].edge) ||
        upload(s, &s->d_chunk_fixed[lv], s->luts.chunk_fixed[lv]))
      return 1;
  // the rank LUTs only served to build and check the lists
  s->luts.gather_lut = {};
  s->luts.chunk_lut = {};
  return 0;
}

int lod_run_epoch(tile_stream_gpu* s, uint32_t e, uint32_t* out_mask)
{
#define CALL(T) run_epoch_t<T>(s, e, out_mask)
  DTYPE_DISPATCH(s->plan.config.dtype, CALL);
#undef CALL
}

template<typename T>
int restore_t(tile_stream_gpu* s)
{
  const Plan& p = s->plan;
  return move_launch<T, false>(s->luts.gather_tiles, s->d_gather_tiles, s->d_gather_edge, (T*)s->d_lod,
                               (T*)s->d_linear, s->d_gather_fixed, p.levels[0].lod_nelem, p.fixed_count,
                               s->st_comp) != hipSuccess;
}

int lod_restore_linear(tile_stream_gpu* s)
{
#define CALL(T) restore_t<T>(s)
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
