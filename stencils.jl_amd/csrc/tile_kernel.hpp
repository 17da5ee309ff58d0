// Tiled gather kernel: dst[I] = op(neighbours of src around I), one template
// for 1-D, 2-D and 3-D arrays.
//
// Per workgroup (256 threads = 4 wave64s):
//   1. stage the (TX+2rx) x (TY+2ry) x (TZ+2rz) source box around the output
//      tile in LDS.  Boundary conditions are resolved here, per staged
//      element, and only in blocks that touch an edge (a uniform branch), so
//      the compute loop below never branches on boundaries.
//   2. each thread owns one x column of the tile and KR consecutive rows; it
//      folds the stencil over LDS in offset order (the reference's left fold,
//      src/gatherstencil.jl:105-109 + StaticArrays reductions).  For shapes
//      fixed at build time (`Sh::kStatic`) the offsets are constants, the
//      loops unroll and identical LDS reads across the thread's rows are
//      shared in registers.
//   3. results are written with coalesced stores along dim 1.
#pragma once

#include "common.hpp"

namespace sth {

// Arithmetic type of the fold for (storage type, op).
template <typename S, int OPK>
struct AccOf {
  using type = S;
};
template <int OPK> struct AccOf<bf16_t, OPK> { using type = float; };
template <int OPK> struct AccOf<f16_t, OPK> { using type = float; };
#define STH_INT_ACC(T)                                                        \
  template <> struct AccOf<T, OP_SUM> { using type = Elem<T>::uacc; };        \
  template <> struct AccOf<T, OP_KP> { using type = Elem<T>::uacc; };         \
  template <> struct AccOf<T, OP_MEAN> { using type = Elem<T>::uacc; };       \
  template <> struct AccOf<T, OP_LIFE> { using type = uint32_t; };
STH_INT_ACC(bool)
STH_INT_ACC(uint8_t)
STH_INT_ACC(int8_t)
STH_INT_ACC(uint16_t)
STH_INT_ACC(int16_t)
STH_INT_ACC(uint32_t)
STH_INT_ACC(int32_t)
STH_INT_ACC(uint64_t)
STH_INT_ACC(int64_t)
#undef STH_INT_ACC

// Type fold_finish returns: the accumulator's, except for an integer mean,
// which folds in the (wrapping) integer accumulator and divides in Float64.
template <typename S, int OPK>
struct FinOf {
  using type = typename AccOf<S, OPK>::type;
};
#define STH_INT_FIN(T) \
  template <> struct FinOf<T, OP_MEAN> { using type = double; };
STH_INT_FIN(bool)
STH_INT_FIN(uint8_t)
STH_INT_FIN(int8_t)
STH_INT_FIN(uint16_t)
STH_INT_FIN(int16_t)
STH_INT_FIN(uint32_t)
STH_INT_FIN(int32_t)
STH_INT_FIN(uint64_t)
STH_INT_FIN(int64_t)
#undef STH_INT_FIN

// Storage type of the result (the reference's _return_type).
template <typename S, int OPK>
struct OutOf {
  using type = S;
};
template <> struct OutOf<bool, OP_SUM> { using type = int64_t; };
#define STH_INT_MEAN(T) \
  template <> struct OutOf<T, OP_MEAN> { using type = double; };
STH_INT_MEAN(bool)
STH_INT_MEAN(uint8_t)
STH_INT_MEAN(int8_t)
STH_INT_MEAN(uint16_t)
STH_INT_MEAN(int16_t)
STH_INT_MEAN(uint32_t)
STH_INT_MEAN(int32_t)
STH_INT_MEAN(uint64_t)
STH_INT_MEAN(int64_t)
#undef STH_INT_MEAN

// Element type staged in LDS: BF16 / F16 storage is widened to f32 once,
// when the tile is loaded, so the fold reads plain floats.
template <typename S> struct Staged { using type = typename Elem<S>::raw; };
template <> struct Staged<bf16_t> { using type = float; };
template <> struct Staged<f16_t> { using type = float; };

template <typename S>
__device__ __forceinline__ typename Staged<S>::type stage(typename Elem<S>::raw v) {
  if constexpr (__is_same(S, bf16_t)) return bf16_to_float(v);
  else if constexpr (__is_same(S, f16_t)) return f16_to_float(v);
  else return v;
}

template <typename S>
constexpr bool kStorageMode = __is_same(S, bf16_t) || __is_same(S, f16_t);

// A stencil whose offsets arrive at run time (Positional, NamedStencil and
// shapes outside the build-time catalogue).
struct RuntimeShape {
  static constexpr bool kStatic = false;
  static constexpr int L = 0, RX = 0, RY = 0, RZ = 0;
};

template <typename A, class P>
__device__ __forceinline__ A weight(const P& p, int k) {
  if constexpr (__is_same(A, double)) {
    return p.w.d[k];
  } else if constexpr (__is_same(A, float)) {
    return p.w.f[k];
  } else {
    return static_cast<A>(p.w.u64[k]);
  }
}

template <typename A>
__device__ __forceinline__ A fold_min(A a, A b) {
  if constexpr (__is_same(A, float) || __is_same(A, double)) return jl_min(a, b);
  else return b < a ? b : a;
}
template <typename A>
__device__ __forceinline__ A fold_max(A a, A b) {
  if constexpr (__is_same(A, float) || __is_same(A, double)) return jl_max(a, b);
  else return b > a ? b : a;
}

// One step of the left fold over the stencil values, in offset order: `k` is
// the offset's position, `rv` its staged value.  Shared by every kernel so
// that all paths compute each cell with the same expressions.
template <typename S, int OPK, class P>
__device__ __forceinline__ void fold_step(typename AccOf<S, OPK>::type& acc, uint32_t& cnt, int k,
                                          typename Staged<S>::type rv, const P& p) {
  using A = typename AccOf<S, OPK>::type;
  if constexpr (OPK == OP_LIFE) {
    cnt += nonzero(rv) ? 1u : 0u;
  } else {
    A v;
    if constexpr (kStorageMode<S>) v = rv;
    else v = widen<A, S>(rv);
    if constexpr (OPK == OP_SUM || OPK == OP_MEAN || OPK == OP_LAP) {
      acc = k == 0 ? v : acc + v;
    } else if constexpr (OPK == OP_MIN) {
      acc = k == 0 ? v : fold_min(acc, v);
    } else if constexpr (OPK == OP_MAX) {
      acc = k == 0 ? v : fold_max(acc, v);
    } else if constexpr (OPK == OP_KP) {
      if constexpr (kStorageMode<S>) {
        // storage precision mode: f32 fused multiply-add, offset order
        acc = __builtin_fmaf(v, weight<A>(p, k), acc);
      } else {
        // native types: separate multiply and add, as the reference
        const A pr = v * weight<A>(p, k);
        acc = acc + pr;
      }
    }
  }
}

// Finish the fold: mean divides by L, the Laplacian and Life rules combine
// with the centre value `cv`.  An integer mean is the sum folded in the
// element type (wrapping; Bool counts in Int64) converted to Float64 and
// divided by L, which is `sum(a) / length(a)` over a StaticVector.
template <typename S, int OPK, class P>
__device__ __forceinline__ typename FinOf<S, OPK>::type fold_finish(typename AccOf<S, OPK>::type acc,
                                                                    uint32_t cnt,
                                                                    typename Staged<S>::type cv,
                                                                    int L, const P& p) {
  using A = typename AccOf<S, OPK>::type;
  if constexpr (OPK == OP_MEAN && !Elem<S>::is_float) {
    double s;
    if constexpr (__is_same(S, bool)) s = static_cast<double>(static_cast<int64_t>(acc));
    else s = static_cast<double>(static_cast<typename Elem<S>::raw>(acc));
    return s / static_cast<double>(L);
  } else if constexpr (OPK == OP_MEAN) {
    acc = acc / static_cast<A>(L);
  } else if constexpr (OPK == OP_LAP) {
    A c;
    if constexpr (kStorageMode<S>) c = cv;
    else c = widen<A, S>(cv);
    const A t = static_cast<A>(L) * c;
    const A d = acc - t;
    const A a = static_cast<A>(p.op.alpha) * d;
    acc = c + a;
  } else if constexpr (OPK == OP_LIFE) {
    const bool alive = nonzero(cv);
    const uint64_t mask = alive ? p.op.survive : p.op.birth;
    acc = cnt < 64u ? static_cast<A>((mask >> cnt) & 1u) : A(0);
  }
  return acc;
}

// One output value from the staged box; `c` is the LDS index of the centre.
template <typename S, int OPK, class Sh, int W, int HW, class P>
__device__ __forceinline__ typename FinOf<S, OPK>::type reduce_cell(
    const typename Staged<S>::type* __restrict__ tile, int c, const P& p,
    int wdyn, int hwdyn) {
  using A = typename AccOf<S, OPK>::type;
  const int L = Sh::kStatic ? Sh::L : p.L;
  A acc = A(0);
  uint32_t cnt = 0;
  if constexpr (Sh::kStatic) {
#pragma unroll
    for (int k = 0; k < Sh::L; ++k)
      fold_step<S, OPK>(acc, cnt, k, tile[c + Sh::off[k][2] * HW + Sh::off[k][1] * W + Sh::off[k][0]], p);
  } else {
    for (int k = 0; k < L; ++k) {
      const int o = p.offs[k];
      const int dx = static_cast<int8_t>(o & 0xff);
      const int dy = static_cast<int8_t>((o >> 8) & 0xff);
      const int dz = static_cast<int8_t>((o >> 16) & 0xff);
      fold_step<S, OPK>(acc, cnt, k, tile[c + dz * hwdyn + dy * wdyn + dx], p);
    }
  }
  return fold_finish<S, OPK>(acc, cnt, tile[c], L, p);
}

template <typename Raw>
__device__ __forceinline__ Raw load_raw(const DevArr& a, int64_t x, int64_t y, int64_t z) {
  const Raw* p = reinterpret_cast<const Raw*>(a.base);
  return p[z * a.s2 + y * a.s1 + x];
}

// Stage the source box [X0, X0 + W) x [Y0, Y0 + H) x [Z0, Z0 + D) of `p`'s
// array into `tile` (x fastest) with the workgroup's 256 threads, applying
// the boundary outside the memory that may be read directly (the logical
// array, plus the halo when the padding keeps a valid one).  (rx, ry, rz) is
// the stencil's reach: cells beyond the reach box of the array feed no
// stored result and are not loaded.  The caller synchronises afterwards.
//
// Each thread takes the elements tid, tid + 256, ... in batches of kB, all
// loads of a batch issued before the first LDS store (kB loads in flight
// per thread rather than one), with coordinates advanced incrementally.
// (Staging by rows, a wave-uniform row pointer plus the lane, measured
// slower: the 64-bit row address is scalar work per chunk, and a row of
// 64 + 2r elements takes two mostly empty chunk loads.)
template <typename S, class P>
__device__ __forceinline__ void stage_box(const P& p, int64_t X0, int64_t Y0, int64_t Z0, int W, int H, int D,
                                          int rx, int ry, int rz, typename Staged<S>::type* __restrict__ tile) {
  using Raw = typename Elem<S>::raw;
  constexpr int kThreads = 256;
  constexpr int kB = 8;
  const int HW = H * W;
  const int64_t n0 = p.src.n0, n1 = p.src.n1, n2 = p.src.n2;
  const int total = HW * D;
  const int64_t h0 = p.read_halo ? p.src.h0 : 0;
  const int64_t h1 = p.read_halo ? p.src.h1 : 0;
  const int64_t h2 = p.read_halo ? p.src.h2 : 0;
  const bool direct = X0 >= -h0 && X0 + W <= n0 + h0 && Y0 >= -h1 && Y0 + H <= n1 + h1 &&
                      Z0 >= -h2 && Z0 + D <= n2 + h2;
  const int tid = static_cast<int>(threadIdx.x);
  // (lx, ly, lz) of element tid, and the step of 256 elements as (sx, sy, sz)
  int lz = tid / HW, ly = (tid - lz * HW) / W, lx = tid - lz * HW - ly * W;
  const int sz = kThreads / HW, sy = (kThreads - sz * HW) / W, sx = kThreads - sz * HW - sy * W;
  auto advance = [&]() {
    lx += sx;
    ly += sy;
    lz += sz;
    if (lx >= W) {
      lx -= W;
      ++ly;
    }
    if (ly >= H) {
      ly -= H;
      ++lz;
    }
  };
  const Raw pv = raw_from_bits<Raw>(p.padval);
  for (int base = tid; base < total; base += kB * kThreads) {
    Raw v[kB];
#pragma unroll
    for (int e = 0; e < kB; ++e) {
      v[e] = Raw(0);
      if (base + e * kThreads < total) {
        const int64_t gx = X0 + lx, gy = Y0 + ly, gz = Z0 + lz;
        if (direct) {
          v[e] = load_raw<Raw>(p.src, gx, gy, gz);
        } else if (gx >= -rx && gx < n0 + rx && gy >= -ry && gy < n1 + ry && gz >= -rz && gz < n2 + rz) {
          // only the box [-r, n + r) feeds valid outputs; the rest of a
          // partial tile is never read by a stored result
          const bool inb = gx >= 0 && gx < n0 && gy >= 0 && gy < n1 && gz >= 0 && gz < n2;
          if (inb || p.read_halo) {
            v[e] = load_raw<Raw>(p.src, gx, gy, gz);
          } else if (p.bc == STH_REMOVE) {
            v[e] = pv;
          } else {
            v[e] = load_raw<Raw>(p.src, map_index(gx, n0, p.bc), map_index(gy, n1, p.bc), map_index(gz, n2, p.bc));
          }
        }
      }
      advance();
    }
#pragma unroll
    for (int e = 0; e < kB; ++e)
      if (base + e * kThreads < total) tile[base + e * kThreads] = stage<S>(v[e]);
  }
}

// One output tile: stage the source box in LDS, fold, store.  `P` carries
// the array descriptors, boundary and op parameters (MapParams for a single
// array, BatchJob for one job of a batch); (bx, by, bz) is the tile.
template <typename S, int OPK, class Sh, int TX, int TY, int TZ, class P>
__device__ __forceinline__ void tile_body(const P& p, int bx, int by, int bz,
                                          typename Staged<S>::type* __restrict__ tile) {
  using Raw = typename Elem<S>::raw;
  using O = typename OutOf<S, OPK>::type;
  using ORaw = typename Elem<O>::raw;
  constexpr int kThreads = 256;
  static_assert(kThreads % TX == 0, "TX must divide the block");
  constexpr int G = kThreads / TX;
  static_assert((TY * TZ) % G == 0, "rows must split evenly over thread groups");
  constexpr int KR = TY * TZ / G;

  const int rx = Sh::kStatic ? Sh::RX : p.rx;
  const int ry = Sh::kStatic ? Sh::RY : p.ry;
  const int rz = Sh::kStatic ? Sh::RZ : p.rz;
  constexpr int WS = TX + 2 * Sh::RX, HS = TY + 2 * Sh::RY, DS = TZ + 2 * Sh::RZ;
  const int W = Sh::kStatic ? WS : TX + 2 * rx;
  const int H = Sh::kStatic ? HS : TY + 2 * ry;
  const int D = Sh::kStatic ? DS : TZ + 2 * rz;
  const int HW = H * W;

  const int64_t n0 = p.src.n0, n1 = p.src.n1, n2 = p.src.n2;
  const int64_t X0 = static_cast<int64_t>(bx) * TX - rx;
  const int64_t Y0 = static_cast<int64_t>(by) * TY - ry;
  const int64_t Z0 = static_cast<int64_t>(bz) * TZ - rz;

  // ---- 1. stage the source box -------------------------------------
  stage_box<S>(p, X0, Y0, Z0, W, H, D, rx, ry, rz, tile);
  __syncthreads();

  // ---- 2. fold -------------------------------------------------------
  const int lx = threadIdx.x % TX;
  const int grp = threadIdx.x / TX;
  const int64_t gx = static_cast<int64_t>(bx) * TX + lx;
  ORaw* dstp = reinterpret_cast<ORaw*>(p.dst.base);
  if constexpr (!Sh::kStatic) {
    // runtime offsets: each offset decoded once and applied to the thread's
    // KR rows together (KR independent LDS reads in flight per offset)
    using A = typename AccOf<S, OPK>::type;
    A acc[KR];
    uint32_t cnt[KR];
    int c[KR];
#pragma unroll
    for (int j = 0; j < KR; ++j) {
      const int row = grp * KR + j;
      c[j] = (row / TY + rz) * HW + (row % TY + ry) * W + (lx + rx);
      acc[j] = A(0);
      cnt[j] = 0;
    }
    auto delta = [&](int k) {
      const int o = p.offs[k];
      return static_cast<int8_t>((o >> 16) & 0xff) * HW + static_cast<int8_t>((o >> 8) & 0xff) * W +
             static_cast<int8_t>(o & 0xff);
    };
    if (p.L > 0) {
      const int d = delta(0);
#pragma unroll
      for (int j = 0; j < KR; ++j) fold_step<S, OPK>(acc[j], cnt[j], 0, tile[c[j] + d], p);
    }
    for (int k = 1; k < p.L; ++k) {
      __builtin_assume(k > 0);  // the fold's first step is peeled above
      const int d = delta(k);
#pragma unroll
      for (int j = 0; j < KR; ++j) fold_step<S, OPK>(acc[j], cnt[j], k, tile[c[j] + d], p);
    }
#pragma unroll
    for (int j = 0; j < KR; ++j) {
      const int row = grp * KR + j;
      const int64_t gy = static_cast<int64_t>(by) * TY + row % TY;
      const int64_t gz = static_cast<int64_t>(bz) * TZ + row / TY;
      const auto r = fold_finish<S, OPK>(acc[j], cnt[j], tile[c[j]], p.L, p);
      if (gx < n0 && gy < n1 && gz < n2) dstp[gz * p.dst.s2 + gy * p.dst.s1 + gx] = narrow<O>(r);
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < KR; ++j) {
    const int row = grp * KR + j;
    const int ly = row % TY;
    const int lz = row / TY;
    const int c = (lz + rz) * HW + (ly + ry) * W + (lx + rx);
    const auto acc = reduce_cell<S, OPK, Sh, WS, WS * HS>(tile, c, p, W, HW);
    const int64_t gy = static_cast<int64_t>(by) * TY + ly;
    const int64_t gz = static_cast<int64_t>(bz) * TZ + lz;
    if (gx < n0 && gy < n1 && gz < n2)
      dstp[gz * p.dst.s2 + gy * p.dst.s1 + gx] = narrow<O>(acc);
  }
}

template <typename S, int OPK, class Sh, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void tile_kernel(MapParams p) {
  extern __shared__ __align__(16) unsigned char smem_dyn[];
  auto* tile = reinterpret_cast<typename Staged<S>::type*>(smem_dyn);
  const int nblocks = p.tiles_x * p.tiles_y * p.tiles_z;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bx = bid % p.tiles_x;
  const int byz = bid / p.tiles_x;
  const int by = byz % p.tiles_y;
  const int bz = byz / p.tiles_y;
  tile_body<S, OPK, Sh, TX, TY, TZ>(p, bx, by, bz, tile);
}

// Host-side launcher with a uniform signature for the kernel registry.
template <typename S, int OPK, class Sh, int TX, int TY, int TZ>
void launch_tile(const MapParams& p, hipStream_t stream) {
  using Raw = typename Staged<S>::type;
  const int rx = Sh::kStatic ? Sh::RX : p.rx;
  const int ry = Sh::kStatic ? Sh::RY : p.ry;
  const int rz = Sh::kStatic ? Sh::RZ : p.rz;
  const size_t lds = static_cast<size_t>(TX + 2 * rx) * (TY + 2 * ry) * (TZ + 2 * rz) * sizeof(Raw);
  const unsigned nblocks = static_cast<unsigned>(p.tiles_x) * p.tiles_y * p.tiles_z;
  hipLaunchKernelGGL((tile_kernel<S, OPK, Sh, TX, TY, TZ>), dim3(nblocks), dim3(256), lds,
                     stream, p);
}

}  // namespace sth
