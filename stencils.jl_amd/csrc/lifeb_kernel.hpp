// Conway's Life (B3/S23) on 1-byte cells, G generations per launch, with the
// cells of a row held one bit each: the steady steps of a Life iteration.
//
// A wave owns a strip of 1024 columns (16 per lane) and a chunk of rows, and
// marches down the chunk.  Per row it loads each lane's 16 bytes once, packs
// them to 16 bits (v_dot4_u32_u8 against bit weights: cells are 0 or 1 in the
// steady state) and forms a 32-bit window: 8 cells of the lane to the left,
// its own 16, 8 cells of the lane to the right (DPP wave shifts; lanes 0 and
// 63 load the neighbouring strip's cells themselves).  Generation k of row y
// is computed from generation k - 1 of rows y - 1, y, y + 1 as soon as row
// y + 1 of generation k - 1 exists, so a wave keeps three rows per
// generation in registers and every row is loaded and stored exactly once
// (plus 2 G rows of overlap per chunk).  Horizontal neighbours are the
// window shifted by one bit; every generation the outermost bit on each
// side loses its neighbours, so after G <= 8 generations the lane's own 16
// bits are still exact.  The neighbour count is a bit-parallel adder tree
// over the 3 x 3 block (3-input full adders as v_bitop3), and the rule is
// `S == 3 or (S == 4 and alive)` on the 3 x 3 sum S including the centre.
// Generation G is unpacked back to bytes (a 24-bit multiply spreads four
// bits into four bytes) and stored with one 16-byte store per lane.
//
// Cells outside the array follow the reference's getneighbor
// (src/array.jl:129-179) in every generation:
//   - Wrap: rows and the window's outside cells are read at their wrapped
//     coordinates; Wrap is a translation of a torus, so every generation of
//     an outside cell equals the wrapped cell's;
//   - Reflect: outside rows are read at their mirrored rows, and the window
//     cells beyond x = 0 or x = n0 - 1 are the lane's own bits reversed; the
//     mirrored grid evolves as the mirror image (the Life rule does not
//     depend on neighbour order), so outside cells stay the mirrored cells;
//   - Remove: outside rows and cells are padval in every generation.
// Only interior cells of the destination are written; the host finishes the
// iteration with single-generation steps (halo refresh included).
//
// Preconditions checked on the host: 2-D, n0 % 16 == 0, n0 >= 16, n1 >= 3,
// interior rows of source and destination on 16-byte boundaries, B3/S23,
// Wrap / Remove / Reflect, source cells already 0 or 1.
#pragma once

#include "life_kernel.hpp"
#include "zmarch2_kernel.hpp"

namespace sth {

// v_bitop3: bitwise function of three operands given by its truth table on
// (0xf0, 0xcc, 0xaa).
template <uint32_t LUT>
__device__ __forceinline__ uint32_t bop3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, static_cast<unsigned char>(LUT & 0xffu));
}
constexpr uint32_t kA3 = 0xf0u, kB3 = 0xccu, kC3 = 0xaau;

// 16 bytes of 0/1 cells -> 16 bits (bit i = cell i).
__device__ __forceinline__ uint32_t pack16(const uint4 v) {
  const uint32_t lo = __builtin_amdgcn_udot4(v.y, 0x80402010u, __builtin_amdgcn_udot4(v.x, 0x08040201u, 0u, false),
                                             false);
  const uint32_t hi = __builtin_amdgcn_udot4(v.w, 0x80402010u, __builtin_amdgcn_udot4(v.z, 0x08040201u, 0u, false),
                                             false);
  return lo | (hi << 8);
}

// Bits 8..23 of a window -> 16 bytes of 0/1 cells.  n * 0x204081 places bit
// k of a nibble n at bit 8 k (the other partial products land elsewhere).
__device__ __forceinline__ uint4 unpack16(uint32_t w) {
  uint4 r;
  r.x = __umul24(__builtin_amdgcn_ubfe(w, 8, 4), 0x204081u) & 0x01010101u;
  r.y = __umul24(__builtin_amdgcn_ubfe(w, 12, 4), 0x204081u) & 0x01010101u;
  r.z = __umul24(__builtin_amdgcn_ubfe(w, 16, 4), 0x204081u) & 0x01010101u;
  r.w = __umul24(__builtin_amdgcn_ubfe(w, 20, 4), 0x204081u) & 0x01010101u;
  return r;
}

// Horizontal 3-sum of a window: per bit, s + 2 c = w[i-1] + w[i] + w[i+1].
__device__ __forceinline__ void hsum3(uint32_t w, uint32_t& s, uint32_t& c) {
  const uint32_t l = w << 1, r = w >> 1;
  s = bop3<kA3 ^ kB3 ^ kC3>(l, w, r);
  c = bop3<(kA3 & kB3) | (kA3 & kC3) | (kB3 & kC3)>(l, w, r);
}

// One generation of B3/S23 for a row from the horizontal 3-sums of the rows
// above (a), at (b) and below (c) and the row's own bits.
__device__ __forceinline__ uint32_t conway_bits(uint32_t sa, uint32_t ca, uint32_t sb, uint32_t cb, uint32_t sc,
                                                uint32_t cc, uint32_t alive) {
  // S = s1 + 2 T with T = u0 + 2 u1 + c1
  const uint32_t s1 = bop3<kA3 ^ kB3 ^ kC3>(sa, sb, sc);
  const uint32_t c1 = bop3<(kA3 & kB3) | (kA3 & kC3) | (kB3 & kC3)>(sa, sb, sc);
  const uint32_t u0 = bop3<kA3 ^ kB3 ^ kC3>(ca, cb, cc);
  const uint32_t u1 = bop3<(kA3 & kB3) | (kA3 & kC3) | (kB3 & kC3)>(ca, cb, cc);
  // T == 1 <=> !u1 & (u0 ^ c1);  T == 2 <=> (u1 & !u0 & !c1) | (!u1 & u0 & c1)
  const uint32_t t1 = bop3<~kA3 & (kB3 ^ kC3)>(u1, u0, c1);
  const uint32_t t2 = bop3<(kA3 & ~kB3 & ~kC3) | (~kA3 & kB3 & kC3)>(u1, u0, c1);
  // S == 3 <=> s1 & T == 1;  S == 4 <=> !s1 & T == 2 (then alive)
  const uint32_t q = bop3<kA3 & kB3 & ~kC3>(t2, alive, s1);
  return bop3<(kA3 & kB3) | kC3>(s1, t1, q);
}

// Row coordinate y (any integer) mapped into [0, n) for Wrap (torus) and
// Reflect (mirror at 0 and n - 1, period 2 n - 2).
__device__ __forceinline__ int32_t lifeb_map_row(int32_t y, int32_t n, bool reflect) {
  if (y >= 0 && y < n) return y;
  if (!reflect) {
    const int32_t m = y % n;
    return m < 0 ? m + n : m;
  }
  const int32_t per = 2 * n - 2;
  int32_t m = y % per;
  if (m < 0) m += per;
  return m < n ? m : per - m;
}

template <int G, int P, int NT, int BC>
__global__ __launch_bounds__(256) void lifeb_kernel(MapParams p) {
  static_assert(G >= 1 && G <= 8, "a 32-bit window keeps 16 exact cells for at most 8 generations");
  static_assert(P == 3 || P == 6, "prefetch ring must divide the unroll of 6");
  constexpr bool REMOVE = BC == STH_REMOVE, REFLECT = BC == STH_REFLECT;
  constexpr int U = 6;  // unroll: rings of 3 (generations) and P (loads) are static

  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nblocks = p.tiles_x * p.tiles_y;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int strip = bid % p.tiles_x;
  const int group = bid / p.tiles_x;

  const int32_t n0 = static_cast<int32_t>(p.src.n0), n1 = static_cast<int32_t>(p.src.n1);
  const int32_t nch = p.tiles_y * 4;
  const int32_t CH = (n1 + nch - 1) / nch;
  const int32_t ya = (group * 4 + wave) * CH;
  if (ya >= n1) return;  // uniform: the wave has no rows
  const int32_t yb = ya + CH < n1 ? ya + CH : n1;
  const int32_t X0 = strip * 1024;
  const int32_t x0 = X0 + 16 * lane;
  const bool active = x0 < n0;
  // Raw buffer accesses with 32-bit offsets (the host runs this kernel only
  // on grids below 2 GiB).  Every load and store is unconditional, so the
  // compiler can count the accesses in flight and wait only for the row a
  // step consumes: loads are always in range, and stores of lanes or rows
  // that own no result use offset 2^31, which the destination's range check
  // drops.
  const __amdgpu_buffer_rsrc_t srs = buf_rsrc(p.src.base, 0xffffffffu);
  const int64_t dext = (p.dst.n1 - 1) * p.dst.s1 + p.dst.n0;
  const __amdgpu_buffer_rsrc_t drs = buf_rsrc(p.dst.base, static_cast<uint32_t>(dext));
  constexpr uint32_t kDrop = 0x80000000u;
  const int64_t s1 = p.src.s1, d1 = p.dst.s1;
  const uint32_t pb = (p.padval & 0xffu) ? 1u : 0u;
  const uint32_t padwin = pb ? 0xffffffffu : 0u;

  // Column of the lane's own 16-byte load: inactive lanes read chunk 0; the
  // first lane past the end stands for x = n0 .. n0 + 15, which under Wrap
  // are cells 0 .. 15.
  const int32_t xl = active ? x0 : 0;
  // Lanes 0 and 63 also read the neighbouring strip's chunk (Wrap maps it).
  const int32_t xe = lane == 0 ? (X0 == 0 ? n0 - 16 : X0 - 16) : (X0 + 1024 < n0 ? X0 + 1024 : 0);
  const bool edge_lane = lane == 0 || lane == 63;
  // the second load of lanes 1..62 re-reads their own chunk (a cache hit)
  const uint32_t xeo = static_cast<uint32_t>(edge_lane ? xe : xl);
  const uint32_t xlo = static_cast<uint32_t>(xl);
  const uint32_t xso = active ? static_cast<uint32_t>(x0) : kDrop;
  const bool at_left = x0 == 0, at_right = x0 + 16 == n0;  // lanes at the array's x edges
  // Remove: window bits at x < 0 (left lane) or x >= n0 (right lane and beyond)
  uint32_t keep = 0xffffffffu;
  if (at_left) keep &= 0xffffff00u;
  if (at_right) keep &= 0x00ffffffu;
  if (!active) keep = 0u;
  const bool xedge_wave = X0 == 0 || X0 + 1024 >= n0;  // uniform
  const int32_t NI = (yb - ya) + 2 * G;  // gen-0 rows ya - G .. yb - 1 + G


  // gen-0 row loads, P rows ahead (rows past the chunk reload its last row)
  using U4 = unsigned int __attribute__((ext_vector_type(4)));
  U4 own[P], edg[P];
  auto issue = [&](int i, int slot) {
    const int32_t y = ya - G + (i < NI ? i : NI - 1);
    const int32_t ym = REMOVE ? (y < 0 ? 0 : (y >= n1 ? n1 - 1 : y)) : lifeb_map_row(y, n1, REFLECT);
    const uint32_t soff = static_cast<uint32_t>(static_cast<int64_t>(ym) * s1);
    own[slot] = buf_load<U4>(srs, xlo, soff);
    edg[slot] = buf_load<U4>(srs, xeo, soff);
  };
  auto as_uint4 = [](U4 v) {
    uint4 r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    r.w = v.w;
    return r;
  };
  // the 32-bit window of gen-0 row ya - G + i from its loaded chunks
  auto window = [&](int i, int slot) -> uint32_t {
    const uint32_t m = pack16(as_uint4(own[slot]));
    const uint32_t me = pack16(as_uint4(edg[slot]));
    uint32_t ml = from_left(m, me), mr = from_right(m, me);
    if (xedge_wave) {
      if constexpr (REFLECT) {
        const uint32_t rv = __builtin_bitreverse32(m);
        if (at_left) ml = rv >> 15;   // bits 8..15 <- cells 8 .. 1
        if (at_right) mr = rv >> 17;  // bits 0..7 <- cells 14 .. 7
      }
      if constexpr (REMOVE) {
        if (at_left) ml = pb ? 0xff00u : 0u;
        if (at_right) mr = pb ? 0xffu : 0u;
      }
    }
    uint32_t w = __builtin_amdgcn_ubfe(ml, 8, 8) | (m << 8) | (mr << 24);
    if constexpr (REMOVE) {
      const int32_t y = ya - G + i;
      if (y < 0 || y >= n1) w = padwin;
    }
    return w;
  };

  // generation k (0 .. G-1) rows: window and horizontal 3-sums, ring of 3
  uint32_t wr[G][3], hs[G][3], hc[G][3];
#pragma unroll
  for (int k = 0; k < G; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) wr[k][j] = hs[k][j] = hc[k][j] = 0u;

#pragma unroll
  for (int i = 0; i < P; ++i) issue(i, i);

  for (int i0 = 0; i0 < NI; i0 += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u;
      if (i >= NI) break;
      const int sl = u % P;
      const uint32_t w0 = window(i, sl);
      issue(i + P, sl);
      // level 0: gen-0 row of iteration i
      wr[0][u % 3] = w0;
      hsum3(w0, hs[0][u % 3], hc[0][u % 3]);
      // gen k of row ya - G + i - k from gen k - 1 of iterations i-2, i-1, i
      uint32_t out = 0;
#pragma unroll
      for (int k = 1; k <= G; ++k) {
        const int a = (u + 1) % 3, b = (u + 2) % 3, c = u % 3;  // iterations i-2, i-1, i
        uint32_t nw = conway_bits(hs[k - 1][a], hc[k - 1][a], hs[k - 1][b], hc[k - 1][b], hs[k - 1][c],
                                  hc[k - 1][c], wr[k - 1][b]);
        if constexpr (REMOVE) {
          if (xedge_wave) nw = bop3<(kA3 & kB3) | (~kB3 & kC3)>(nw, keep, padwin);
          const int32_t y = ya - G + i - k;
          if (y < 0 || y >= n1) nw = padwin;
        }
        if (k < G) {
          wr[k][c] = nw;
          hsum3(nw, hs[k][c], hc[k][c]);
        } else {
          out = nw;
        }
      }
      // gen G of row ya + i - 2G (the first 2G steps only fill the pipeline)
      const int32_t yo = ya + i - 2 * G;
      const bool st = i >= 2 * G;
      const uint32_t soff = st ? static_cast<uint32_t>(static_cast<int64_t>(yo) * d1) : 0u;
      const uint4 o = unpack16(out);
      U4 ov;
      ov.x = o.x;
      ov.y = o.y;
      ov.z = o.z;
      ov.w = o.w;
      buf_store<U4, NT ? kStoreNT : 0>(ov, drs, st ? xso : kDrop, soff);
    }
  }
}

// Grid of lifeb_kernel: 1024-column strips x groups of four row chunks,
// chunks sized so the launch has about `waves` waves and at least 4 G rows.
inline void lifeb_grid(int64_t n0, int64_t n1, int g, int64_t waves, int32_t* tx, int32_t* ty) {
  const int64_t strips = (n0 + 1023) / 1024;
  int64_t chunks = (waves + strips - 1) / strips;
  const int64_t max_chunks = n1 / (4 * g) > 0 ? n1 / (4 * g) : 1;
  chunks = chunks < 1 ? 1 : (chunks > max_chunks ? max_chunks : chunks);
  *tx = static_cast<int32_t>(strips);
  *ty = static_cast<int32_t>((chunks + 3) / 4);
}

template <int G, int P, int NT>
void launch_lifeb(const MapParams& p, hipStream_t s) {
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y;
  if (p.bc == STH_REMOVE)
    hipLaunchKernelGGL((lifeb_kernel<G, P, NT, STH_REMOVE>), dim3(nb), dim3(256), 0, s, p);
  else if (p.bc == STH_REFLECT)
    hipLaunchKernelGGL((lifeb_kernel<G, P, NT, STH_REFLECT>), dim3(nb), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((lifeb_kernel<G, P, NT, STH_WRAP>), dim3(nb), dim3(256), 0, s, p);
}

}  // namespace sth
