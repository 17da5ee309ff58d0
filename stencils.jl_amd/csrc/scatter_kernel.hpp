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
// does and no atomics or column passes are needed.  A dest cell d receives
// from the senders q - offset_k for each preimage q of d under the target
// mapping (d itself; under Wrap or Reflect, cells within R of an edge have
// one more per axis):
//   - cells whose only preimage is d receive from the cells d - offset_k,
//     one per k (those inside the array, for cells within R of an edge).
//     Their order depends only on the dest column's residue modulo S, so a
//     host-built table lists the offsets in arrival order per residue.  A
//     workgroup stages its 64 x 16 cells' senders (A, and B for the
//     downhill rules) and the table in LDS once; a wave shares its dest
//     column, so table reads are LDS broadcasts.  The second array's
//     neighbour of each sender is the dest cell itself;
//   - cells with a wrapped or reflected preimage walk the source columns of
//     each pass (one per preimage window) and their rows ascending, taking
//     every offset whose target maps onto the cell, in fold order.
// The value rule and `op` are template parameters: the loop over offsets
// holds no branch on them.
#pragma once

#include "tile_kernel.hpp"

namespace sth {

constexpr int kScatterMaxR = 7;
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
  double coef;
  int32_t tiles_x, tiles_y;
  // cells with a wrapped or reflected preimage: the coordinates of each axis
  // that have one (ascending), and the leading workgroups that own them
  int32_t nbi, nbj;
  int32_t band_i[4 * kScatterMaxR], band_j[4 * kScatterMaxR];
  int32_t gen_cells, gen_blocks;
};

// Offsets packed as (dim 1 & 0xffff) | (dim 2 << 16).
__device__ __forceinline__ int scatter_o1(int32_t o) { return static_cast<int16_t>(o & 0xffff); }
__device__ __forceinline__ int scatter_o2(int32_t o) { return o >> 16; }

// The positions p in [-R, n - 1 + R] (a cell plus an offset) that the
// reference's scatter target mapping sends to coordinate x of one axis: x
// itself, and under Wrap (one wrap: p +- n) or Reflect (one reflection: -p,
// 2n - 2 - p in 0-based terms) at most one more on each side.  Three
// ascending slots with valid flags (fixed slots stay in registers).
struct ScatterPre {
  int q0, q1, q2;
  bool v0, v2;  // q1 = x is always valid
};
__device__ __forceinline__ ScatterPre scatter_preimages(int x, int n, int R, int bc) {
  ScatterPre r{x, x, x, false, false};
  if (bc == STH_WRAP) {
    r.q0 = x - n;
    r.v0 = x - n >= -R;
    r.q2 = x + n;
    r.v2 = x + n <= n - 1 + R;
  } else if (bc == STH_REFLECT) {
    r.q0 = -x;
    r.v0 = x > 0 && -x >= -R;
    r.q2 = 2 * n - 2 - x;
    r.v2 = x < n - 1 && 2 * n - 2 - x <= n - 1 + R;
  }
  return r;
}

// The offsets o in [-R, R] with s + o a preimage of x, as bits o + R.
__device__ __forceinline__ uint32_t scatter_offset_mask(const ScatterPre& q, int s, int R) {
  uint32_t m = 0;
  const int d0 = q.q0 - s, d1 = q.q1 - s, d2 = q.q2 - s;
  if (q.v0 && d0 >= -R && d0 <= R) m |= 1u << (d0 + R);
  if (d1 >= -R && d1 <= R) m |= 1u << (d1 + R);
  if (q.v2 && d2 >= -R && d2 <= R) m |= 1u << (d2 + R);
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

// The value a sender with value a (and second-array value b) sends along
// an offset whose target's second-array value is bk.
template <typename T, int KIND>
__device__ __forceinline__ T scatter_rule(T a, T b, T bk, T c) {
  if constexpr (KIND == STH_SCATTER_CONST) return c;
  if constexpr (KIND == STH_SCATTER_CENTER) return a;
  if constexpr (KIND == STH_SCATTER_SCALED) return a * c;
  const T d = b - bk;
  if constexpr (KIND == STH_SCATTER_DOWNHILL) return d > T(0) ? a * c : T(0);
  const T ad = a * d;
  return d > T(0) ? ad * c : T(0);
}

// A workgroup owns kScatterTX x kScatterTY dest cells: lane -> dim 1, each
// of the four waves every fourth dest column (dim 2).  Its senders (the
// tile and R around it) are staged in LDS rows of kScatterTX + 2R.
constexpr int kScatterTX = 64;
constexpr int kScatterTY = 16;

// Device table (int32): [0, L) the offsets packed in fold order; then for
// each dest column residue r = 0..S-1 the L packed offsets in the order a
// cell receives them from the cells d - offset; then the same S x L list
// as LDS deltas -(o2 * (kScatterTX + 2R) + o1) of those senders.
__host__ __device__ constexpr int scatter_table_len(int L, int S) { return L * (1 + 2 * S); }
constexpr int kScatterTabPerThread = 4;  // table entries each thread loads alongside the box

// A dest cell whose only preimage in both axes is itself receives from the
// cells d - offset that lie in the array, in the residue's order; `checked`
// drops the senders outside the array (cells within R of an edge).  The
// senders sit in the LDS box (c: the dest cell's slot).
template <typename T, int KIND, int COMB, bool CHECKED>
__device__ __forceinline__ T scatter_direct(T acc, T cf, const int32_t* __restrict__ pk,
                                            const int32_t* __restrict__ dl, int L, const T* __restrict__ la,
                                            const T* __restrict__ lb, int c, int di, int dj, int n0, int n1) {
  T bd = T(0);
  if constexpr (KIND >= STH_SCATTER_DOWNHILL) bd = lb[c];
  for (int t = 0; t < L; ++t) {
    if constexpr (CHECKED) {
      const int32_t o = pk[t];
      const int si = di - scatter_o1(o), sj = dj - scatter_o2(o);
      if (si < 0 || si >= n0 || sj < 0 || sj >= n1) continue;
    }
    T a = T(0), b = T(0);
    if constexpr (KIND != STH_SCATTER_CONST) {
      const int sc = c + dl[t];
      a = la[sc];
      if constexpr (KIND >= STH_SCATTER_DOWNHILL) b = lb[sc];
    }
    acc = scatter_combine<T>(acc, scatter_rule<T, KIND>(a, b, bd, cf), COMB);
  }
  return acc;
}

// A dest cell with a wrapped or reflected preimage: the source columns of
// each pass in ascending order (one per preimage window), their rows
// ascending over the merged row windows, every offset whose target maps
// onto the cell (both axes' offset masks) in fold order.
template <typename T, int KIND, int COMB>
__device__ __forceinline__ T scatter_general(const ScatterParams& p, const int32_t* __restrict__ fold, T acc, T cf,
                                             int di, int dj) {
  const int n0 = static_cast<int>(p.dst.n0), n1 = static_cast<int>(p.dst.n1);
  const int R = p.R, S = p.S, L = p.L;
  const ScatterPre pi = scatter_preimages(di, n0, R, p.bc);
  const ScatterPre pj = scatter_preimages(dj, n1, R, p.bc);
  const T* __restrict__ a = reinterpret_cast<const T*>(p.src.base);
  const T* __restrict__ b = reinterpret_cast<const T*>(p.aux.base);
  for (int ps = 0; ps < S; ++ps) {
    int last_col = -1;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      const bool valid = w == 0 ? pj.v0 : (w == 1 ? true : pj.v2);
      const int q = w == 0 ? pj.q0 : (w == 1 ? pj.q1 : pj.q2);
      if (!valid) continue;
      const int lo = q - R;
      const int sj = lo + (((ps - lo) % S) + S) % S;
      if (sj < 0 || sj >= n1 || sj <= last_col) continue;
      last_col = sj;
      const uint32_t mj = scatter_offset_mask(pj, sj, R);
      int last = -1;
#pragma unroll
      for (int wi = 0; wi < 3; ++wi) {
        const bool vi = wi == 0 ? pi.v0 : (wi == 1 ? true : pi.v2);
        const int qi = wi == 0 ? pi.q0 : (wi == 1 ? pi.q1 : pi.q2);
        if (!vi) continue;
        const int r0 = qi - R > last + 1 ? qi - R : last + 1;
        const int r1 = qi + R < n0 - 1 ? qi + R : n0 - 1;
        for (int si = r0 < 0 ? 0 : r0; si <= r1; ++si) {
          const uint32_t mi = scatter_offset_mask(pi, si, R);
          T av = T(0), bv = T(0);
          if constexpr (KIND != STH_SCATTER_CONST) av = a[static_cast<int64_t>(sj) * p.src.s1 + si];
          if constexpr (KIND >= STH_SCATTER_DOWNHILL) bv = b[static_cast<int64_t>(sj) * p.aux.s1 + si];
          for (int k = 0; k < L; ++k) {
            const int32_t o = fold[k];
            const int oi = scatter_o1(o), oj = scatter_o2(o);
            if (!((mi >> (oi + R)) & (mj >> (oj + R)) & 1u)) continue;
            T bk = T(0);
            if constexpr (KIND >= STH_SCATTER_DOWNHILL) bk = scatter_aux_neighbor<T>(p, si, sj, oi, oj);
            acc = scatter_combine<T>(acc, scatter_rule<T, KIND>(av, bv, bk, cf), COMB);
          }
        }
        if (r1 > last) last = r1;
      }
    }
  }
  return acc;
}

// One step of an interior cell's fold with its sender at LDS delta D, known
// at compile time (runtime-compiled kernels, whose interior order is code).
template <typename T, int KIND, int COMB, int D>
__device__ __forceinline__ T scatter_step(T acc, T cf, const T* __restrict__ la, const T* __restrict__ lb, int c, T bd) {
  T a = T(0), b = T(0);
  if constexpr (KIND != STH_SCATTER_CONST) {
    a = la[c + D];
    if constexpr (KIND >= STH_SCATTER_DOWNHILL) b = lb[c + D];
  }
  return scatter_combine<T>(acc, scatter_rule<T, KIND>(a, b, bd, cf), COMB);
}

// The interior order read from the LDS table (the build-time kernels).
struct TableOrder {
  static constexpr bool kStatic = false;
};

// Copy the box [x0, x0 + W) x [y0, y0 + H) of A (and B) into LDS, zero
// outside the array: batches of kB elements per thread, every load of a
// batch issued before the first LDS store, so a workgroup keeps kB loads in
// flight per thread instead of one.
template <typename T, bool AUX>
__device__ __forceinline__ void scatter_stage(const ScatterParams& p, T* __restrict__ la, T* __restrict__ lb,
                                              int64_t x0, int64_t y0, int W, int H) {
  constexpr int kB = 8;
  const T* __restrict__ a = reinterpret_cast<const T*>(p.src.base);
  const T* __restrict__ b = reinterpret_cast<const T*>(p.aux.base);
  const int64_t n0 = p.dst.n0, n1 = p.dst.n1;
  const int n = W * H;
  const int tid = static_cast<int>(threadIdx.x);
  int ly = tid / W, lx = tid - (tid / W) * W;
  const int dly = 256 / W, dlx = 256 - dly * W;
  for (int base = tid; base < n; base += kB * 256) {
    T va[kB], vb[kB];
#pragma unroll
    for (int e = 0; e < kB; ++e) {
      const int64_t x = x0 + lx, y = y0 + ly;
      const bool in = base + e * 256 < n && x >= 0 && x < n0 && y >= 0 && y < n1;
      va[e] = in ? a[y * p.src.s1 + x] : T(0);
      if constexpr (AUX) vb[e] = in ? b[y * p.aux.s1 + x] : T(0);
      lx += dlx;
      ly += dly;
      if (lx >= W) {
        lx -= W;
        ++ly;
      }
    }
#pragma unroll
    for (int e = 0; e < kB; ++e) {
      const int idx = base + e * 256;
      if (idx < n) {
        la[idx] = va[e];
        if constexpr (AUX) lb[idx] = vb[e];
      }
    }
  }
}

// One thread per cell with a wrapped or reflected preimage: cells g of the
// column band (dj in band_j, every di), then of the row band (di in
// band_i, dj outside band_j).
template <typename T, int KIND, int COMB>
__device__ __forceinline__ void scatter_band_cell(const ScatterParams& p, const int32_t* __restrict__ fold, int g) {
  const int n0 = static_cast<int>(p.dst.n0);
  int di, dj;
  const int colcells = p.nbj * n0;
  if (g < colcells) {
    dj = p.band_j[g / n0];
    di = g - (g / n0) * n0;
  } else {
    const int h = g - colcells;
    di = p.band_i[h % p.nbi];
    dj = h / p.nbi;  // the (h / nbi)-th column outside band_j
    for (int b = 0; b < p.nbj; ++b)
      if (p.band_j[b] <= dj) ++dj;
  }
  T* dp = reinterpret_cast<T*>(p.dst.base) + static_cast<int64_t>(dj) * p.dst.s1 + di;
  const T acc = p.zero_dest ? T(0) : *dp;
  *dp = scatter_general<T, KIND, COMB>(p, fold, acc, static_cast<T>(p.coef), di, dj);
}

template <typename T, int KIND, int COMB, class ORD = TableOrder>
__global__ __launch_bounds__(256) void scatter_kernel(ScatterParams p, const int32_t* __restrict__ tbl) {
  extern __shared__ __align__(16) unsigned char smem_dyn[];
  if (static_cast<int>(blockIdx.x) < p.gen_blocks) {
    // band cells: the leading workgroups, so their longer walks overlap the tiles
    const int g = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (g < p.gen_cells) scatter_band_cell<T, KIND, COMB>(p, tbl, g);
    return;
  }
  // Workgroups go to the 8 XCDs round robin, each XCD with its own L2:
  // XCD k takes the k-th contiguous run of tiles, so neighbouring tiles that
  // share halo rows and cache lines share an L2 (gen_blocks is a multiple
  // of 8, so the tile index mod 8 is the XCD).
  const int tile = xcd_block_id(static_cast<int>(blockIdx.x) - p.gen_blocks, p.tiles_x * p.tiles_y, 1);
  const int n0 = static_cast<int>(p.dst.n0), n1 = static_cast<int>(p.dst.n1);
  const int di0 = (tile % p.tiles_x) * kScatterTX;
  const int dj0 = (tile / p.tiles_x) * kScatterTY;
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const int di = di0 + lane;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  const int R = p.R, S = p.S, L = p.L;
  const int W = kScatterTX + 2 * R, H = kScatterTY + 2 * R;
  constexpr bool kNeedA = KIND != STH_SCATTER_CONST;
  constexpr bool kAux = KIND >= STH_SCATTER_DOWNHILL;
  // LDS: A box, B box (as the rule reads them), then the device table
  T* __restrict__ la = reinterpret_cast<T*>(smem_dyn);
  T* __restrict__ lb = la + (kNeedA ? W * H : 0);
  int32_t* __restrict__ ltab = reinterpret_cast<int32_t*>(lb + (kAux ? W * H : 0));
  // the dest values of this wave's columns, loaded up front
  constexpr int kCols = kScatterTY / 4;
  T* __restrict__ dst = reinterpret_cast<T*>(p.dst.base);
  T dv[kCols];
#pragma unroll
  for (int c = 0; c < kCols; ++c) {
    const int dj = dj0 + wave + 4 * c;
    dv[c] = !p.zero_dest && dj < n1 && di < n0 ? dst[static_cast<int64_t>(dj) * p.dst.s1 + di] : T(0);
  }
  // the table (loaded alongside the box) and the box
  const int ntab = scatter_table_len(L, S);
  int32_t tv[kScatterTabPerThread];
#pragma unroll
  for (int e = 0; e < kScatterTabPerThread; ++e) {
    const int i = static_cast<int>(threadIdx.x) + e * 256;
    tv[e] = i < ntab ? tbl[i] : 0;
  }
  if constexpr (kNeedA) scatter_stage<T, kAux>(p, la, lb, di0 - R, dj0 - R, W, H);
#pragma unroll
  for (int e = 0; e < kScatterTabPerThread; ++e) {
    const int i = static_cast<int>(threadIdx.x) + e * 256;
    if (i < ntab) ltab[i] = tv[e];
  }
  for (int i = static_cast<int>(threadIdx.x) + kScatterTabPerThread * 256; i < ntab; i += 256) ltab[i] = tbl[i];
  __syncthreads();
  const T cf = static_cast<T>(p.coef);
  const bool row_inner = di > R && di < n0 - 1 - R;
  const ScatterPre pi = scatter_preimages(di, n0, R, p.bc);
  const bool row_general = pi.v0 || pi.v2;
  // dest column residues modulo S, stepped by 4 (no division per column)
  int r = (dj0 + wave) % S;
  const int r4 = 4 % S;
  for (int jj = wave; jj < kScatterTY; jj += 4) {
    const int dj = dj0 + jj;
    if (dj >= n1) break;
    T acc = dv[0];
#pragma unroll
    for (int c = 0; c + 1 < kCols; ++c) dv[c] = dv[c + 1];
    if (di < n0) {
      const int32_t* __restrict__ pk = ltab + L + r * L;
      const int32_t* __restrict__ dl = ltab + L + S * L + r * L;
      const int c = (jj + R) * W + lane + R;
      const ScatterPre pj = scatter_preimages(dj, n1, R, p.bc);
      if (row_inner && dj > R && dj < n1 - 1 - R) {
        if constexpr (ORD::kStatic)
          acc = ORD::template run<T, KIND, COMB>(r, acc, cf, la, lb, c);
        else
          acc = scatter_direct<T, KIND, COMB, false>(acc, cf, pk, dl, L, la, lb, c, di, dj, n0, n1);
        dst[static_cast<int64_t>(dj) * p.dst.s1 + di] = acc;
      } else if (!row_general && !pj.v0 && !pj.v2) {
        acc = scatter_direct<T, KIND, COMB, true>(acc, cf, pk, dl, L, la, lb, c, di, dj, n0, n1);
        dst[static_cast<int64_t>(dj) * p.dst.s1 + di] = acc;
      }  // else a band cell: its leading workgroup writes it
    }
    r += r4;
    if (r >= S) r -= S;
  }
}

// LDS bytes one workgroup of scatter_kernel takes.
inline size_t scatter_lds_bytes(int kind, int R, int L, size_t elem) {
  const size_t box = static_cast<size_t>(kScatterTX + 2 * R) * (kScatterTY + 2 * R) * elem;
  const size_t boxes = kind == STH_SCATTER_CONST ? 0 : (kind >= STH_SCATTER_DOWNHILL ? 2 : 1);
  return box * boxes + static_cast<size_t>(scatter_table_len(L, 2 * R + 1)) * sizeof(int32_t);
}

template <typename T, int KIND, int COMB>
void launch_scatter_k(const ScatterParams& p, const int32_t* tbl, hipStream_t s) {
  hipLaunchKernelGGL((scatter_kernel<T, KIND, COMB>),
                     dim3(static_cast<unsigned>(p.gen_blocks) + static_cast<unsigned>(p.tiles_x) * p.tiles_y), dim3(256),
                     scatter_lds_bytes(KIND, p.R, p.L, sizeof(T)), s, p, tbl);
}

template <typename T, int KIND>
void launch_scatter_c(const ScatterParams& p, const int32_t* tbl, hipStream_t s) {
  if (p.combine == STH_COMBINE_MAX) launch_scatter_k<T, KIND, STH_COMBINE_MAX>(p, tbl, s);
  else if (p.combine == STH_COMBINE_MIN) launch_scatter_k<T, KIND, STH_COMBINE_MIN>(p, tbl, s);
  else launch_scatter_k<T, KIND, STH_COMBINE_ADD>(p, tbl, s);
}

template <typename T>
void launch_scatter(const ScatterParams& p, const int32_t* tbl, hipStream_t s) {
  switch (p.kind) {
    case STH_SCATTER_CONST: launch_scatter_c<T, STH_SCATTER_CONST>(p, tbl, s); break;
    case STH_SCATTER_CENTER: launch_scatter_c<T, STH_SCATTER_CENTER>(p, tbl, s); break;
    case STH_SCATTER_SCALED: launch_scatter_c<T, STH_SCATTER_SCALED>(p, tbl, s); break;
    case STH_SCATTER_DOWNHILL: launch_scatter_c<T, STH_SCATTER_DOWNHILL>(p, tbl, s); break;
    default: launch_scatter_c<T, STH_SCATTER_DOWNHILL_FLOW>(p, tbl, s); break;
  }
}

}  // namespace sth
