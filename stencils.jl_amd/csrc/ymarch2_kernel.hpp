// Two steps of a 2-D radius-1 stencil update per launch (temporal blocking)
// for the steady steps of an iteration: generation t+2 from generation t,
// generation t+1 held in registers.
//
// Every wave works alone (no LDS, no barriers).  Each lane holds one 16-byte
// vector of V cells of a row; lane l of the wave covers columns
// X0 - V + V l .. X0 + V l - 1, so lanes 1..56 own the 56 V output columns
// X0 .. X0 + 56 V - 1 (whole aligned vectors), lane 0 and lane 57 are the
// rims and lanes 58..63 idle along.
// The wave marches down a chunk of rows: generation-t rows sit in an
// eight-deep register queue (loads run six rows ahead of the row computed: a
// row is little work, so one row ahead would leave the load latency exposed);
// generation t+1 is computed on
// every lane into a second queue, generation t+2 of the row behind on lanes
// 1..56, and stored.  Neighbours along x come from the lane's own vector and,
// at its ends, from the adjacent lanes (DPP wave shifts); along y from the
// queue.  Chunks overlap by two rows on each side.
//
// Exactness as in the 3-D z-marching kernel: every value uses the shared
// fold expressions in offset order, generation t+1 is rounded to the storage
// type as a store would round it, Wrap outside cells are translations of
// inside ones (src/array.jl:167-179) and Remove outside cells read padval in
// every generation (src/array.jl:133-135).  Under Reflect the outside
// generation-t+1 cells next to the array are replaced by the cells they
// mirror (src/array.jl:154-166) before generation t+2 reads them.
#pragma once

#include <type_traits>

#include "stream_kernel.hpp"
#include "tile_kernel.hpp"
#include "zmarch2_kernel.hpp"

namespace sth {

template <typename Raw>
constexpr int ymarch2_v() {
  return 16 / static_cast<int>(sizeof(Raw));
}

// Output lanes per wave (lanes 1 .. kYm2Out): 56 lanes of 16 bytes make each
// row segment a whole number of 128-byte lines (62 lanes measured 7% slower
// on 16384^2 f32 from partial-line writes).
constexpr int kYm2Out = 56;

// One fold over the stencil at element i of a lane's vector: row slots q0
// (dy = -1), q1, q2; lm / rp are the element left of element 0 and right of
// element V-1 (from the neighbouring lanes) for each row slot.
template <typename S, int OPK, class Sh, int V, int I, class P>
__device__ __forceinline__ void ym2_fold_terms(typename AccOf<S, OPK>::type& acc, uint32_t& cnt,
                                               const typename Staged<S>::type (&q0)[V],
                                               const typename Staged<S>::type (&q1)[V],
                                               const typename Staged<S>::type (&q2)[V],
                                               const typename Staged<S>::type (&lm)[3],
                                               const typename Staged<S>::type (&rp)[3], int i, const P& p) {
  if constexpr (I < Sh::L) {
    constexpr int dx = Sh::off[I][0], dy = Sh::off[I][1];
    const auto& q = dy < 0 ? q0 : (dy == 0 ? q1 : q2);
    const int k = i + dx;
    const auto v = k < 0 ? lm[dy + 1] : (k >= V ? rp[dy + 1] : q[k < 0 ? 0 : (k >= V ? V - 1 : k)]);
    fold_step<S, OPK>(acc, cnt, I, v, p);
    ym2_fold_terms<S, OPK, Sh, V, I + 1>(acc, cnt, q0, q1, q2, lm, rp, i, p);
  }
}

// c ? x : y on values that stay values: a select between two elements of one
// register array would otherwise become a load through a selected address,
// for which the compiler moves the array to LDS.
template <typename Val>
__device__ __forceinline__ Val select_val(bool c, Val x, Val y) {
  STH_OPAQUE(x);
  STH_OPAQUE(y);
  return c ? x : y;
}

// element `k` of q for the one bit k set in `bits` (bits is a run-time mask)
template <int V, typename Val>
__device__ __forceinline__ Val pick(const Val (&q)[V], uint32_t bits) {
  Val r = q[0];
#pragma unroll
  for (int k = 1; k < V; ++k) r = select_val(((bits >> k) & 1u) != 0, q[k], r);
  return r;
}

template <typename S, int OPK, class Sh, int BC>
__global__ __launch_bounds__(256) void ymarch2_kernel(MapParams p) {
  constexpr bool REMOVE = BC == STH_REMOVE, REFLECT = BC == STH_REFLECT;
  using Raw = typename Elem<S>::raw;
  using Val = typename Staged<S>::type;
  using A = typename AccOf<S, OPK>::type;
  using O = typename OutOf<S, OPK>::type;
  using ORaw = typename Elem<O>::raw;
  constexpr int V = ymarch2_v<Raw>();
  static_assert(sizeof(ORaw) == sizeof(Raw), "iterations keep the element type");
  static_assert(Sh::RX <= 1 && Sh::RY <= 1 && Sh::RZ == 0, "2-D radius-1 stencils only");

  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nblocks = p.tiles_x * p.tiles_y;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bx = bid % p.tiles_x;
  const int by = bid / p.tiles_x;

  const int32_t n0 = static_cast<int32_t>(p.src.n0), n1 = static_cast<int32_t>(p.src.n1);
  // four waves of a block take four consecutive row chunks
  const int32_t rc = p.tiles_z;  // rows per chunk
  const int32_t yb = (by * 4 + wave) * rc;
  const int32_t ye = yb + rc < n1 ? yb + rc : n1;
  if (yb >= ye) return;  // uniform: the wave has no rows
  const int32_t X0 = bx * kYm2Out * V;
  const Val pv = stage<S>(raw_from_bits<Raw>(p.padval));
  constexpr int64_t es = sizeof(Raw);

  // 32-bit byte offsets (the host checks the extent), loads always inside
  // the array, stores of lanes and cells that own no result dropped by the
  // buffer range check (offset 2^31 is past the destination's extent)
  const __amdgpu_buffer_rsrc_t src = buf_rsrc(p.src.base, 0xffffffffu);
  const int64_t dext = ((p.dst.n1 - 1) * p.dst.s1 + p.dst.n0) * es;
  const __amdgpu_buffer_rsrc_t dst = buf_rsrc(p.dst.base, static_cast<uint32_t>(dext));
  constexpr uint32_t kDrop = 0x80000000u;

  // mapped coordinate of c; cells beyond the two-cell rim feed no stored
  // result and are clamped into it first (n >= 3, so every mapped rim cell
  // lies inside the array)
  auto map_c = [&](int32_t c, int32_t n) -> int32_t {
    c = c < -2 ? -2 : (c > n + 1 ? n + 1 : c);
    if (c < 0) return REMOVE ? 0 : (REFLECT ? -c : c + n);
    if (c >= n) return REMOVE ? 0 : (REFLECT ? 2 * n - 2 - c : c - n);
    return c;
  };
  // this lane's columns: x0 .. x0 + V - 1
  const int32_t x0 = X0 - V + V * lane;
  const bool x_direct = x0 >= 0 && x0 + V <= n0;  // one vector load
  uint32_t xb[V];
  bool xout[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const int32_t x = x0 + i;
    xout[i] = x < 0 || x >= n0;
    xb[i] = static_cast<uint32_t>(map_c(x, n0)) * static_cast<uint32_t>(es);
  }
  const bool store_lane = lane >= 1 && lane <= kYm2Out;
  uint32_t xs[V];  // store offsets; kDrop where the lane owns no result
#pragma unroll
  for (int i = 0; i < V; ++i)
    xs[i] = store_lane && x0 + i < n0 ? static_cast<uint32_t>(x0 + i) * static_cast<uint32_t>(es) : kDrop;
  const bool store_vec = store_lane && x0 + V <= n0;

  constexpr int QD = 8;  // generation-t queue depth
  Val g0[QD][V];  // generation t: row i in slot i % QD (i = 0 is yb - 2)
  Val g1[4][V];   // generation t+1: row j in slot j % 4 (j = 0 is yb - 1)
  Val e0[QD][2];  // generation t: the cell left of element 0 / right of element V-1
  Val e1[4][2];   // the same for generation t+1

  auto load_row = [&](int i, Val (&q)[V]) {
    const int32_t y = yb - 2 + i;
    const bool yout = y < 0 || y >= n1;
    const uint32_t yoff = static_cast<uint32_t>(static_cast<int64_t>(map_c(y, n1)) * p.src.s1 * es);
    if (x_direct) {
      const RawVec<Raw, V> m = buf_load<RawVec<Raw, V>>(src, xb[0], yoff);
#pragma unroll
      for (int k = 0; k < V; ++k) q[k] = stage<S>(m.e[k]);
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) q[k] = stage<S>(buf_load<Raw>(src, xb[k], yoff));
    }
    if constexpr (REMOVE) {
#pragma unroll
      for (int k = 0; k < V; ++k)
        if (xout[k] || yout) q[k] = pv;
    }
  };
  // neighbours of a row's end elements from the adjacent lanes
  auto edges = [&](const Val (&q)[V], Val (&e)[2]) {
    e[0] = lane_shift<-1>(q[V - 1]);
    e[1] = lane_shift<1>(q[0]);
  };
  // Reflect: the generation-(t+1) cells just outside the array, x = -1 and
  // x = n0, take the values of the cells they mirror, x = 1 and x = n0 - 2 --
  // the reference reads those cells there (src/array.jl:154-166), and a fold
  // over the mirrored neighbourhood would sum in the reverse order.
  // The lanes and elements concerned are fixed per wave, so the fix-up is
  // selects on per-lane masks, with no branch in the row loop.
  const int32_t pr = n0 - (X0 - V);  // x = n0 in this wave's cells (uniform)
  const bool rim_in = pr >= 0 && pr < 64 * V;
  const int32_t pr_c = rim_in ? pr : 2;  // any in-range position when unused
  const int rim_lane = pr_c / V, rim_elem = pr_c % V;
  const int mir_lane = (pr_c - 2) / V, mir_elem = (pr_c - 2) % V;  // x = n0 - 2
  const bool left_fix = X0 == 0 && lane == 0;
  const bool right_fix = rim_in && lane == rim_lane;
  const bool from_prev = mir_lane != rim_lane;
  // Element choices as bit masks the compiler cannot see through: a runtime
  // element index would make it index the register arrays, which it then
  // moves to LDS.
  uint32_t mir_bits = 1u << mir_elem;
  uint32_t rim_bits = right_fix ? 1u << rim_elem : 0u;
  STH_OPAQUE(mir_bits);
  STH_OPAQUE(rim_bits);
  auto reflect_columns = [&](Val (&h)[V]) {
    const Val from_right = lane_shift<1>(h[1]);  // lane l: lane l+1's element 1 (x = 1 for lane 0)
    const Val same = pick<V>(h, mir_bits);
    const Val prev = lane_shift<-1>(same);  // lane l: lane l-1's element mir_elem
    const Val mirrored = from_prev ? prev : same;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      if (k == V - 1) h[k] = left_fix ? from_right : h[k];
      h[k] = ((rim_bits >> k) & 1u) ? mirrored : h[k];
    }
  };

  // Reflect fix-ups run only in waves that touch the array's boundary; the
  // others take the same branch-free body as Wrap (branches inside the row
  // loop make the compiler wait for the prefetched rows early).
  const bool edge_wave = REFLECT && (X0 == 0 || rim_in || yb == 0 || ye == n1);
  auto march = [&](auto edge_tag) {
    constexpr bool EDGE = decltype(edge_tag)::value;
    const int nit = (ye - yb) + 2;  // generation t+1 rows yb-1 .. ye; generation-t rows 0 .. nit+1
  #pragma unroll
    for (int i = 0; i < QD - 1; ++i)
      if (i <= nit + 1) load_row(i, g0[i]);
    edges(g0[0], e0[0]);
    edges(g0[1], e0[1]);
    for (int j0 = 0; j0 < nit; j0 += QD) {
  #pragma unroll
      for (int u = 0; u < QD; ++u) {
        const int j = j0 + u;
        if (j >= nit) break;
        // row j + QD - 1 into the slot row j - 1 held
        if (j + QD - 1 <= nit + 1) load_row(j + QD - 1, g0[(u + QD - 1) % QD]);
        edges(g0[(u + 2) % QD], e0[(u + 2) % QD]);
        // generation t+1 on row yb - 1 + j
        {
          const auto& q0 = g0[u % QD];
          const auto& q1 = g0[(u + 1) % QD];
          const auto& q2 = g0[(u + 2) % QD];
          const Val lm[3] = {e0[u % QD][0], e0[(u + 1) % QD][0], e0[(u + 2) % QD][0]};
          const Val rp[3] = {e0[u % QD][1], e0[(u + 1) % QD][1], e0[(u + 2) % QD][1]};
          auto& h = g1[u & 3];
  #pragma unroll
          for (int i = 0; i < V; ++i) {
            A acc = A(0);
            uint32_t cnt = 0;
            ym2_fold_terms<S, OPK, Sh, V, 0>(acc, cnt, q0, q1, q2, lm, rp, i, p);
            h[i] = stage<S>(narrow<O>(fold_finish<S, OPK>(acc, cnt, q1[i], Sh::L, p)));
          }
          if constexpr (REMOVE) {
            const int32_t y = yb - 1 + j;
            const bool yout = y < 0 || y >= n1;
  #pragma unroll
            for (int i = 0; i < V; ++i)
              if (xout[i] || yout) h[i] = pv;
          }
          if constexpr (EDGE) reflect_columns(h);
          edges(h, e1[u & 3]);
        }
        // generation t+2 on row yb - 2 + j
        if (j >= 2) {
          const int64_t y = yb - 2 + j;
          const auto& q1 = g1[(u + 3) & 3];
          RawVec<ORaw, V> out;
          auto fold_row = [&](const Val (&q0)[V], const Val (&q2)[V], const Val (&lm)[3], const Val (&rp)[3]) {
  #pragma unroll
            for (int i = 0; i < V; ++i) {
              A acc = A(0);
              uint32_t cnt = 0;
              ym2_fold_terms<S, OPK, Sh, V, 0>(acc, cnt, q0, q1, q2, lm, rp, i, p);
              out.e[i] = narrow<O>(fold_finish<S, OPK>(acc, cnt, q1[i], Sh::L, p));
            }
          };
          const Val lm[3] = {e1[(u + 2) & 3][0], e1[(u + 3) & 3][0], e1[u & 3][0]};
          const Val rp[3] = {e1[(u + 2) & 3][1], e1[(u + 3) & 3][1], e1[u & 3][1]};
          // Reflect: rows y = -1 and y = n1 read rows 1 and n1 - 2 (uniform
          // selects, effective on the array's first and last rows only)
          if constexpr (EDGE) {
            const bool top = y == 0, bottom = y == n1 - 1;
            Val q0[V], q2[V];
  #pragma unroll
            for (int i = 0; i < V; ++i) {
              q0[i] = select_val(top, g1[u & 3][i], g1[(u + 2) & 3][i]);
              q2[i] = select_val(bottom, g1[(u + 2) & 3][i], g1[u & 3][i]);
            }
            const Val lr[3] = {select_val(top, lm[2], lm[0]), lm[1], select_val(bottom, lm[0], lm[2])};
            const Val rr[3] = {select_val(top, rp[2], rp[0]), rp[1], select_val(bottom, rp[0], rp[2])};
            fold_row(q0, q2, lr, rr);
          } else {
            fold_row(g1[(u + 2) & 3], g1[u & 3], lm, rp);
          }
          const uint32_t yoff = static_cast<uint32_t>(y * p.dst.s1 * es);
          if (store_vec) {
            buf_store<RawVec<ORaw, V>, kStoreNT>(out, dst, xs[0], yoff);
          } else {
  #pragma unroll
            for (int i = 0; i < V; ++i) buf_store<ORaw, kStoreNT>(out.e[i], dst, xs[i], yoff);
          }
        }
      }
    }
  };
  if (edge_wave)
    march(std::true_type{});
  else
    march(std::false_type{});
}

// Grid of ymarch2_kernel for an n0 x n1 array: 56 V-column tiles x groups of
// four row chunks, chunks sized so the launch has about `waves` waves (at
// least 16 rows each).  The chunk length travels in tiles_z.
inline void ymarch2_grid(int64_t n0, int64_t n1, int v, int64_t waves, int32_t* tx, int32_t* ty, int32_t* rows) {
  const int64_t gx = (n0 + kYm2Out * v - 1) / (kYm2Out * v);
  int64_t chunks = (waves + gx - 1) / gx;
  const int64_t max_chunks = (n1 + 15) / 16;
  chunks = chunks < 1 ? 1 : (chunks > max_chunks ? max_chunks : chunks);
  const int64_t rc = (n1 + chunks - 1) / chunks;
  const int64_t nch = (n1 + rc - 1) / rc;
  *tx = static_cast<int32_t>(gx);
  *ty = static_cast<int32_t>((nch + 3) / 4);
  *rows = static_cast<int32_t>(rc);
}

template <typename S, int OPK, class Sh>
void launch_ymarch2(const MapParams& p, hipStream_t s) {
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y;
  if (p.bc == STH_REMOVE)
    hipLaunchKernelGGL((ymarch2_kernel<S, OPK, Sh, STH_REMOVE>), dim3(nb), dim3(256), 0, s, p);
  else if (p.bc == STH_REFLECT)
    hipLaunchKernelGGL((ymarch2_kernel<S, OPK, Sh, STH_REFLECT>), dim3(nb), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((ymarch2_kernel<S, OPK, Sh, STH_WRAP>), dim3(nb), dim3(256), 0, s, p);
}

}  // namespace sth
