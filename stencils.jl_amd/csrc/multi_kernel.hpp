// Multi-array gather: dest[I] = t_0 (+|*) t_1 (+|*) ... with one term per
// argument of `f`, the reference's
//   dest[I] = f(stencil(source, I), _getarg.(args, (I,))...)
// (src/gatherstencil.jl:105-109, 111-113) for the fixed family of `f` that
// combines one fold per argument:
//   t_k = g_k(hood_k)          or   coef_k * g_k(hood_k)
//   g_k = sum / mean / minimum / maximum / kernelproduct / Laplacian / Life
//         over the argument's stencil, or its centre value (center(hood),
//         or a plain array read at I),
// folded left in argument order in the element type T: `a + b + c` in Julia
// is (a + b) + c, every operation rounded (no contraction).
//
// One workgroup computes a TX x TY x TZ output tile.  For each term it stages
// that argument's source box (its own stencil reach and boundary) in LDS,
// folds each cell with the same expressions as the single-array kernels, and
// combines into registers; centre-value terms read their array directly.
// The term table (boundary, padval, offsets, weights per argument) is in
// device memory: wave-uniform scalar loads.
#pragma once

#include "registry.hpp"
#include "tile_kernel.hpp"

namespace sth {

constexpr int kMaxMultiTerms = 8;

struct MultiTerm {
  DevArr src;
  int32_t bc;
  int32_t read_halo;
  int32_t rx, ry, rz;
  int32_t L;
  int32_t opk;     // OpKind of the fold
  int32_t direct;  // 1: the centre value (no fold, no LDS)
  int32_t scaled;  // 1: coef * fold
  int32_t pad_;
  uint64_t padval;
  OpParams op;
  double coef;
  int32_t offs[kMaxRuntimeL];
  union {
    float f[kMaxRuntimeL];
    double d[kMaxRuntimeL];
    uint64_t u64[kMaxRuntimeL];
  } w;
};

struct MultiParams {
  DevArr dst;
  int32_t nterms;
  int32_t combine;  // 0: +, 1: *
  int32_t tiles_x, tiles_y, tiles_z;
  int32_t xcd_remap;
};

// Fold of one term for the cell at LDS index c, converted to the element
// type as the single-array kernels store it.
template <typename S, class T>
__device__ __forceinline__ T multi_fold(const MultiTerm& q, const typename Staged<S>::type* __restrict__ tile,
                                        int c, int W, int HW) {
  switch (q.opk) {
    case OP_SUM: return narrow<T>(reduce_cell<S, OP_SUM, RuntimeShape, 0, 0>(tile, c, q, W, HW));
    case OP_MIN: return narrow<T>(reduce_cell<S, OP_MIN, RuntimeShape, 0, 0>(tile, c, q, W, HW));
    case OP_MAX: return narrow<T>(reduce_cell<S, OP_MAX, RuntimeShape, 0, 0>(tile, c, q, W, HW));
    default: break;
  }
  if constexpr (Elem<S>::is_float) {
    switch (q.opk) {
      case OP_MEAN: return narrow<T>(reduce_cell<S, OP_MEAN, RuntimeShape, 0, 0>(tile, c, q, W, HW));
      case OP_KP: return narrow<T>(reduce_cell<S, OP_KP, RuntimeShape, 0, 0>(tile, c, q, W, HW));
      case OP_LAP: return narrow<T>(reduce_cell<S, OP_LAP, RuntimeShape, 0, 0>(tile, c, q, W, HW));
      default: break;
    }
  } else {
    switch (q.opk) {
      case OP_KP: return narrow<T>(reduce_cell<S, OP_KP, RuntimeShape, 0, 0>(tile, c, q, W, HW));
      case OP_LIFE: return narrow<T>(reduce_cell<S, OP_LIFE, RuntimeShape, 0, 0>(tile, c, q, W, HW));
      default: break;
    }
  }
  return T(0);
}

// The element type's `+` and `*` (integers wrap).
template <class T>
__device__ __forceinline__ T multi_combine(T a, T b, int mul) {
  if constexpr (Elem<T>::is_float) {
    return mul ? a * b : a + b;
  } else {
    using U = typename Elem<T>::uacc;
    return static_cast<T>(mul ? static_cast<U>(static_cast<U>(a) * static_cast<U>(b))
                              : static_cast<U>(static_cast<U>(a) + static_cast<U>(b)));
  }
}

template <typename S, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void multi_kernel(MultiParams mp, const MultiTerm* __restrict__ terms) {
  using T = typename Elem<S>::raw;
  using Val = typename Staged<S>::type;
  extern __shared__ __align__(16) unsigned char smem_dyn[];
  auto* tile = reinterpret_cast<Val*>(smem_dyn);
  constexpr int kThreads = 256;
  constexpr int G = kThreads / TX;
  constexpr int KR = TY * TZ / G;
  static_assert(kThreads % TX == 0 && (TY * TZ) % G == 0, "tile shape");

  const int nblocks = mp.tiles_x * mp.tiles_y * mp.tiles_z;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, mp.xcd_remap);
  const int bx = bid % mp.tiles_x;
  const int byz = bid / mp.tiles_x;
  const int by = byz % mp.tiles_y;
  const int bz = byz / mp.tiles_y;
  const int lx = threadIdx.x % TX;
  const int grp = threadIdx.x / TX;
  const int64_t n0 = mp.dst.n0, n1 = mp.dst.n1, n2 = mp.dst.n2;
  const int64_t gx = static_cast<int64_t>(bx) * TX + lx;

  T acc[KR];
  for (int t = 0; t < mp.nterms; ++t) {
    const MultiTerm& q = terms[t];
    T val[KR];
    if (q.direct) {
      const T* sp = reinterpret_cast<const T*>(q.src.base);
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        const int row = grp * KR + j;
        const int64_t gy = static_cast<int64_t>(by) * TY + row % TY;
        const int64_t gz = static_cast<int64_t>(bz) * TZ + row / TY;
        val[j] = gx < n0 && gy < n1 && gz < n2 ? sp[gz * q.src.s2 + gy * q.src.s1 + gx] : T(0);
      }
    } else {
      const int W = TX + 2 * q.rx, H = TY + 2 * q.ry, D = TZ + 2 * q.rz;
      const int HW = H * W;
      __syncthreads();  // the previous term's folds have read the tile
      stage_box<S>(q, static_cast<int64_t>(bx) * TX - q.rx, static_cast<int64_t>(by) * TY - q.ry,
                   static_cast<int64_t>(bz) * TZ - q.rz, W, H, D, q.rx, q.ry, q.rz, tile);
      __syncthreads();
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        const int row = grp * KR + j;
        const int c = (row / TY + q.rz) * HW + (row % TY + q.ry) * W + (lx + q.rx);
        val[j] = multi_fold<S, T>(q, tile, c, W, HW);
      }
    }
    if (q.scaled) {
      if constexpr (Elem<T>::is_float) {
#pragma unroll
        for (int j = 0; j < KR; ++j) val[j] = static_cast<T>(q.coef) * val[j];
      }
    }
#pragma unroll
    for (int j = 0; j < KR; ++j) acc[j] = t == 0 ? val[j] : multi_combine<T>(acc[j], val[j], mp.combine);
  }

  T* dp = reinterpret_cast<T*>(mp.dst.base);
#pragma unroll
  for (int j = 0; j < KR; ++j) {
    const int row = grp * KR + j;
    const int64_t gy = static_cast<int64_t>(by) * TY + row % TY;
    const int64_t gz = static_cast<int64_t>(bz) * TZ + row / TY;
    if (gx < n0 && gy < n1 && gz < n2) dp[gz * mp.dst.s2 + gy * mp.dst.s1 + gx] = acc[j];
  }
}

using MultiLaunchFn = void (*)(const MultiParams&, const MultiTerm*, size_t lds, hipStream_t);

struct MultiKernelEntry {
  int dtype;
  int cfg;  // CFG_RT1D / CFG_RT2D / CFG_RT3D tile shape
  MultiLaunchFn fn;
};
std::vector<MultiKernelEntry>& multi_registry();

template <typename S, int TX, int TY, int TZ>
void launch_multi(const MultiParams& mp, const MultiTerm* terms, size_t lds, hipStream_t s) {
  const unsigned nblocks = static_cast<unsigned>(mp.tiles_x) * mp.tiles_y * mp.tiles_z;
  hipLaunchKernelGGL((multi_kernel<S, TX, TY, TZ>), dim3(nblocks), dim3(256), lds, s, mp, terms);
}

template <typename S>
void register_multi_type() {
  multi_registry().push_back(MultiKernelEntry{Elem<S>::code, CFG_RT1D, &launch_multi<S, 256, 1, 1>});
  multi_registry().push_back(MultiKernelEntry{Elem<S>::code, CFG_RT2D, &launch_multi<S, 64, 16, 1>});
  multi_registry().push_back(MultiKernelEntry{Elem<S>::code, CFG_RT3D, &launch_multi<S, 32, 8, 4>});
}

}  // namespace sth
