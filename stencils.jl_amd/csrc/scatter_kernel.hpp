// Scatter stencil (scatterstencil!, src/scatterstencil.jl): every source
// cell s sends one value per stencil offset k to dest at s + offset_k, folded
// with `op`.  The reference runs it on the CPU in passes over source columns
// (dim 2) p, p + S, p + 2S, ... (S = 2R + 1) for p = 1..S, rows ascending
// within a column and offsets in order; within one pass no two source
// columns reach the same dest cell, so that loop order is the order in which
// each dest cell receives its values (src/scatterstencil.jl:51-75).
//
// Here one thread owns one dest cell and pulls its values in exactly that
// order (a gather of the scatter), so `+` rounds as the reference's loop
// does and no atomics or column passes are needed:
//   - interior cells (more than R from every edge) receive only from the
//     (2R+1)^2 cells around them: for each pass p the one source column of
//     residue p, rows ascending, and a table maps the offset d - s to its k;
//   - edge cells also receive values whose targets were wrapped or reflected
//     onto them (Wrap, Reflect): the source columns and rows that can reach
//     them are the windows around each preimage of the cell, walked in
//     ascending order per pass, every offset's mapped target checked.
#pragma once

#include "tile_kernel.hpp"

namespace sth {

constexpr int kScatterMaxR = 7;
constexpr int kScatterMaxS = 2 * kScatterMaxR + 1;
constexpr int kScatterMaxL = 128;

struct ScatterParams {
  DevArr src;  // A (interior base; halo widths when read_halo)
  DevArr aux;  // B, when has_aux
  DevArr dst;
  int32_t bc;         // the source's boundary: maps scatter targets, and A's neighbours are not read
  int32_t aux_bc;
  int32_t aux_read_halo;
  int32_t has_aux;
  uint64_t aux_padval;
  int32_t kind;       // STH_SCATTER_*
  int32_t combine;    // STH_COMBINE_ADD / MAX / MIN
  int32_t zero_dest;  // start from zero instead of dest's values
  int32_t R, S, L;
  int32_t general;    // 1: every cell takes the general path (repeated offsets)
  double coef;
  int32_t tiles_x, tiles_y;
  int16_t offs[kScatterMaxL][2];                   // (dim 1, dim 2) offsets in fold order
  int8_t rev[kScatterMaxS * kScatterMaxS];         // k of offset (a, b) at (b + R) * S + (a + R), or -1
};

// Position p of one axis (0-based, |p - x| <= R for an in-range x) mapped
// like the reference's scatter targets: Wrap once, Reflect once (2 - i /
// 2n - i in 1-based terms), Remove / Use: -1 outside.
__device__ __forceinline__ int64_t scatter_map(int64_t p, int64_t n, int bc) {
  if (p >= 0 && p < n) return p;
  if (bc == STH_WRAP) return p < 0 ? p + n : p - n;
  if (bc == STH_REFLECT) return p < 0 ? -p : 2 * n - 2 - p;
  return -1;
}

// Preimages of coordinate x under scatter_map (positions s + o, |o| <= R,
// that land on x), ascending; returns how many (1..3).
__device__ __forceinline__ int scatter_preimages(int64_t x, int64_t n, int R, int bc, int64_t q[3]) {
  int m = 0;
  if (bc == STH_WRAP) {
    if (x - n >= -R) q[m++] = x - n;
    q[m++] = x;
    if (x + n <= n - 1 + R) q[m++] = x + n;
  } else if (bc == STH_REFLECT) {
    if (x > 0 && -x >= -R) q[m++] = -x;
    q[m++] = x;
    if (x < n - 1 && 2 * n - 2 - x <= n - 1 + R) q[m++] = 2 * n - 2 - x;
  } else {
    q[m++] = x;
  }
  return m;
}

template <typename T>
__device__ __forceinline__ T scatter_combine(T acc, T v, int combine) {
  if (combine == STH_COMBINE_MAX) return jl_max(acc, v);
  if (combine == STH_COMBINE_MIN) return jl_min(acc, v);
  return acc + v;
}

// B's neighbour of source cell (si, sj) at offset (oi, oj), under B's own
// boundary and padding (the value neighbors(stencil(B, s))[k] holds).
template <typename T>
__device__ __forceinline__ T scatter_aux_neighbor(const ScatterParams& p, int64_t si, int64_t sj, int oi, int oj) {
  const T* b = reinterpret_cast<const T*>(p.aux.base);
  const int64_t n0 = p.aux.n0, n1 = p.aux.n1;
  int64_t x = si + oi, y = sj + oj;
  const bool in = x >= 0 && x < n0 && y >= 0 && y < n1;
  if (!in && !p.aux_read_halo) {
    if (p.aux_bc == STH_REMOVE) return raw_from_bits<T>(p.aux_padval);
    x = map_index(x, n0, p.aux_bc);
    y = map_index(y, n1, p.aux_bc);
  }
  return b[y * p.aux.s1 + x];
}

// The value source cell (si, sj) sends along offset k.
template <typename T>
__device__ __forceinline__ T scatter_value(const ScatterParams& p, int64_t si, int64_t sj, int k) {
  const T c = static_cast<T>(p.coef);
  if (p.kind == STH_SCATTER_CONST) return c;
  const T a = reinterpret_cast<const T*>(p.src.base)[sj * p.src.s1 + si];
  if (p.kind == STH_SCATTER_CENTER) return a;
  if (p.kind == STH_SCATTER_SCALED) return a * c;
  const T bc = reinterpret_cast<const T*>(p.aux.base)[sj * p.aux.s1 + si];
  const T d = bc - scatter_aux_neighbor<T>(p, si, sj, p.offs[k][0], p.offs[k][1]);
  if (!(d > T(0))) return T(0);
  if (p.kind == STH_SCATTER_DOWNHILL) return a * c;
  const T ad = a * d;
  return ad * c;
}

template <typename T>
__global__ __launch_bounds__(256) void scatter_kernel(ScatterParams p) {
  const int64_t n0 = p.dst.n0, n1 = p.dst.n1;
  const int64_t di = static_cast<int64_t>(blockIdx.x % p.tiles_x) * 64 + (threadIdx.x & 63);
  const int64_t dj = static_cast<int64_t>(blockIdx.x / p.tiles_x) * 4 + (threadIdx.x >> 6);
  if (di >= n0 || dj >= n1) return;
  T* dp = reinterpret_cast<T*>(p.dst.base) + dj * p.dst.s1 + di;
  T acc = p.zero_dest ? T(0) : *dp;
  const int R = p.R, S = p.S;
  if (!p.general && di > R && di < n0 - 1 - R && dj > R && dj < n1 - 1 - R) {
    // interior: only the cells within R send here, each along offset d - s
    for (int ps = 0; ps < S; ++ps) {
      const int64_t lo = dj - R;
      const int64_t sj = lo + (((ps - lo) % S) + S) % S;
      for (int64_t si = di - R; si <= di + R; ++si) {
        const int k = p.rev[(dj - sj + R) * S + (di - si + R)];
        if (k >= 0) acc = scatter_combine<T>(acc, scatter_value<T>(p, si, sj, k), p.combine);
      }
    }
    *dp = acc;
    return;
  }
  // edge: every source cell whose mapped target can be this cell
  int64_t qi[3], qj[3];
  const int mi = scatter_preimages(di, n0, R, p.bc, qi);
  const int mj = scatter_preimages(dj, n1, R, p.bc, qj);
  for (int ps = 0; ps < S; ++ps) {
    // the source columns of residue ps in the windows around dj's
    // preimages, ascending (at most one per window)
    int64_t cols[3];
    int nc = 0;
    for (int w = 0; w < mj; ++w) {
      const int64_t lo = qj[w] - R;
      const int64_t c = lo + (((ps - lo) % S) + S) % S;
      if (c >= 0 && c < n1 && (nc == 0 || c > cols[nc - 1])) cols[nc++] = c;
    }
    for (int ci = 0; ci < nc; ++ci) {
      const int64_t sj = cols[ci];
      // rows: the windows around di's preimages, merged ascending
      int64_t last = -1;
      for (int w = 0; w < mi; ++w) {
        const int64_t r0 = qi[w] - R > last + 1 ? qi[w] - R : last + 1;
        const int64_t r1 = qi[w] + R < n0 - 1 ? qi[w] + R : n0 - 1;
        for (int64_t si = r0 < 0 ? 0 : r0; si <= r1; ++si) {
          for (int k = 0; k < p.L; ++k) {
            if (scatter_map(si + p.offs[k][0], n0, p.bc) == di && scatter_map(sj + p.offs[k][1], n1, p.bc) == dj)
              acc = scatter_combine<T>(acc, scatter_value<T>(p, si, sj, k), p.combine);
          }
        }
        if (r1 > last) last = r1;
      }
    }
  }
  *dp = acc;
}

template <typename T>
void launch_scatter(const ScatterParams& p, hipStream_t s) {
  hipLaunchKernelGGL((scatter_kernel<T>), dim3(static_cast<unsigned>(p.tiles_x) * p.tiles_y), dim3(256), 0, s, p);
}

}  // namespace sth
