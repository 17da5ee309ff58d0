// Two Life generations per launch (temporal blocking) for the steady steps
// of an iteration: gen t+2 is computed from gen t with gen t+1 held in
// registers, so HBM sees one read and one write per two generations.
//
// Each lane owns 16 cells of a row (as in life_kernel) and KY output rows.
// It loads gen-t rows y0-2 .. y0+KY+1 and takes two extra cells on each side
// (x0-2, x0-1, x0+16, x0+17) from its neighbouring lanes, computes gen t+1 on rows y0-1 .. y0+KY for
// its 16 cells plus the two edge cells x0-1 and x0+16, and gen t+2 on its
// KY rows.  Cells outside the array are boundary-mapped exactly as the
// reference's getneighbor does (src/array.jl:129-179):
//   - Wrap: an outside gen t+1 cell is computed from its (mapped) gen-t
//     neighbourhood, which is the neighbourhood of the wrapped cell, so it
//     equals the wrapped cell's gen t+1 value;
//   - Reflect: the mirrored neighbourhood holds the same cells (in mirrored
//     order) and the Life count does not depend on order, so again equal;
//   - Remove: outside cells read padval in every generation.
// The Use boundary (halo read as-is) is not fused.  Only interior cells of
// the destination are written; the host finishes every iteration with at
// least two single-generation steps, which leave both buffers and their
// halos exactly as the step-by-step reference does.
//
// Preconditions checked on the host: the life_kernel ones (2-D, n0 % 16 == 0,
// 16-byte aligned rows), n1 >= 3, bytes already 0/1 (steady state).
#pragma once

#include "life_kernel.hpp"

namespace sth {

template <bool CONWAY>
__device__ __forceinline__ uint32_t life_rule(uint32_t s, uint32_t c, uint32_t tbl) {
  return CONWAY ? rule_conway(s, c) : rule_table(s, c, tbl);
}

// Horizontal 3-sums of 16 cells (4 words) with left / right neighbour bytes.
__device__ __forceinline__ void life_hsum(const uint32_t (&v)[4], uint32_t l, uint32_t r, uint32_t (&h)[4]) {
  h[0] = v[0] + __builtin_amdgcn_alignbyte(v[0], l << 24, 3) + __builtin_amdgcn_alignbyte(v[1], v[0], 1);
  h[1] = v[1] + __builtin_amdgcn_alignbyte(v[1], v[0], 3) + __builtin_amdgcn_alignbyte(v[2], v[1], 1);
  h[2] = v[2] + __builtin_amdgcn_alignbyte(v[2], v[1], 3) + __builtin_amdgcn_alignbyte(v[3], v[2], 1);
  h[3] = v[3] + __builtin_amdgcn_alignbyte(v[3], v[2], 3) + __builtin_amdgcn_alignbyte(r, v[3], 1);
}

template <int KY, bool CONWAY, bool EDGE>
__device__ __forceinline__ void life2_rows(const MapParams& p, const uint8_t* __restrict__ src,
                                           uint8_t* __restrict__ dst, int32_t x0, int32_t y0, bool active,
                                           int32_t xs, const int32_t (&xm)[4], const bool (&xpad)[4],
                                           bool lvirt, bool rvirt, bool fill_l, bool fill_r) {
  const int32_t n1 = static_cast<int32_t>(p.src.n1);
  const int64_t s1 = p.src.s1;
  const int bc = p.bc;
  const bool remove = bc == STH_REMOVE;
  const uint32_t pb = (p.padval & 0xffu) ? 1u : 0u;
  const uint32_t padn = pb * 0x01010101u;
  const uint32_t tbl = static_cast<uint32_t>(p.op.birth & 0x1ffu) |
                       (static_cast<uint32_t>(p.op.survive & 0x1ffu) << 9);

  constexpr int NR = KY + 4;  // gen-t rows y0-2 .. y0+KY+1
  uint32_t w[NR][4];
  uint32_t e[NR][4];  // gen-t cells x0-2, x0-1, x0+16, x0+17
  bool virt[NR];      // row outside the array
  bool padrow[NR];
  const uint8_t* rowp[NR];
  // 1. every row's 16-byte chunk (all loads in flight together)
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int32_t y = y0 - 2 + r;
    int32_t ym = y;
    padrow[r] = false;
    virt[r] = false;
    if constexpr (EDGE) {
      // rows beyond [-2, n1 + 1] exist only in the last, partial block;
      // they are clamped and feed no stored result
      const int32_t yc = y < -2 ? -2 : (y > n1 + 1 ? n1 + 1 : y);
      const bool out = yc < 0 || yc >= n1;
      virt[r] = out;
      if (out) {
        if (remove) {
          padrow[r] = true;
          ym = 0;
        } else {
          ym = static_cast<int32_t>(map_index(yc, n1, bc));
        }
      } else {
        ym = yc;
      }
    }
    rowp[r] = src + static_cast<int64_t>(ym) * s1;
    const uint4 v = *reinterpret_cast<const uint4*>(rowp[r] + xs);
    w[r][0] = v.x;
    w[r][1] = v.y;
    w[r][2] = v.z;
    w[r][3] = v.w;
  }
  // 2. lanes at a wave or array edge load their outer cells (boundary-mapped)
  uint32_t fl[NR], fr[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) fl[r] = fr[r] = 0;
  if (fill_l || fill_r) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      auto cell = [&](int k) { return xpad[k] ? pb : static_cast<uint32_t>(rowp[r][xm[k]]); };
      fl[r] = (cell(0) << 16) | (cell(1) << 24);
      fr[r] = cell(2) | (cell(3) << 8);
    }
  }
  // 3. the others take them from the neighbouring lanes' chunks
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (EDGE && padrow[r]) {
#pragma unroll
      for (int k = 0; k < 4; ++k) w[r][k] = padn;
      fl[r] = (pb << 16) | (pb << 24);
      fr[r] = pb | (pb << 8);
    }
    const uint32_t lw = from_left(w[r][3], fl[r]);
    const uint32_t rw = from_right(w[r][0], fr[r]);
    const uint32_t lv = fill_l ? fl[r] : lw, rv = fill_r ? fr[r] : rw;
    e[r][0] = (lv >> 16) & 0xffu;
    e[r][1] = lv >> 24;
    e[r][2] = rv & 0xffu;
    e[r][3] = (rv >> 8) & 0xffu;
  }

  // ---- gen t+1 on rows 1 .. NR-2 (16 cells + the two edge cells)
  uint32_t h0[NR][4];
#pragma unroll
  for (int r = 0; r < NR; ++r) life_hsum(w[r], e[r][1], e[r][2], h0[r]);
  uint32_t g[NR][4], gl[NR], gr[NR];
#pragma unroll
  for (int r = 1; r < NR - 1; ++r) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t s = h0[r - 1][i] + h0[r][i] + h0[r + 1][i] - w[r][i];
      g[r][i] = life_rule<CONWAY>(s, w[r][i], tbl);
    }
    uint32_t sl = 0, sr = 0;
#pragma unroll
    for (int q = r - 1; q <= r + 1; ++q) {
      sl += e[q][0] + e[q][1] + (w[q][0] & 0xffu);
      sr += (w[q][3] >> 24) + e[q][2] + e[q][3];
    }
    sl -= e[r][1];
    sr -= e[r][2];
    gl[r] = life_rule<CONWAY>(sl, e[r][1], tbl) & 0xffu;
    gr[r] = life_rule<CONWAY>(sr, e[r][2], tbl) & 0xffu;
    // Remove: cells outside the array stay padval in every generation
    if (remove) {
      if (lvirt) gl[r] = pb;
      if (rvirt) gr[r] = pb;
      if (EDGE && virt[r]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) g[r][i] = padn;
        gl[r] = gr[r] = pb;
      }
    }
  }

  // ---- gen t+2 on rows 2 .. NR-3
  uint32_t h1[NR][4];
#pragma unroll
  for (int r = 1; r < NR - 1; ++r) life_hsum(g[r], gl[r], gr[r], h1[r]);
  if (active) {
    uint8_t* d = dst + static_cast<int64_t>(y0) * p.dst.s1 + x0;
#pragma unroll
    for (int r = 2; r < NR - 2; ++r) {
      if (EDGE && y0 + (r - 2) >= n1) break;
      uint32_t o[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t s = h1[r - 1][i] + h1[r][i] + h1[r + 1][i] - g[r][i];
        o[i] = life_rule<CONWAY>(s, g[r][i], tbl);
      }
      *reinterpret_cast<uint4*>(d) = make_uint4(o[0], o[1], o[2], o[3]);
      d += p.dst.s1;
    }
  }
}

template <int KY, bool CONWAY>
__global__ __launch_bounds__(256) void life2_kernel(MapParams p) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nblocks = p.tiles_x * p.tiles_y;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bcol = bid % p.tiles_x;
  const int brow = bid / p.tiles_x;

  const int32_t n0 = static_cast<int32_t>(p.src.n0), n1 = static_cast<int32_t>(p.src.n1);
  const int32_t x0 = (bcol * 64 + lane) * 16;
  const int32_t y0 = (brow * 4 + wave) * KY;
  if (y0 >= n1) return;  // uniform: the wave has no rows
  const bool active = x0 < n0;
  const int bc = p.bc;
  const uint8_t* src = reinterpret_cast<const uint8_t*>(p.src.base);
  uint8_t* dst = reinterpret_cast<uint8_t*>(p.dst.base);

  // boundary-mapped columns of the four edge cells; inactive lanes read a
  // valid chunk and store nothing
  const int32_t xs = active ? x0 : 0;
  const int32_t xo[4] = {xs - 2, xs - 1, xs + 16, xs + 17};
  int32_t xm[4];
  bool xpad[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool out = xo[k] < 0 || xo[k] >= n0;
    xpad[k] = out && bc == STH_REMOVE;
    xm[k] = xpad[k] ? 0 : static_cast<int32_t>(map_index(xo[k], n0, bc));
  }
  const bool lvirt = xs == 0, rvirt = xs + 16 == n0;
  const bool fill_l = lane == 0 || xs == 0;
  const bool fill_r = lane == 63 || xs + 16 >= n0 || x0 + 16 >= n0;
  if (y0 >= 2 && y0 + KY + 1 < n1)
    life2_rows<KY, CONWAY, false>(p, src, dst, x0, y0, active, xs, xm, xpad, lvirt, rvirt, fill_l, fill_r);
  else
    life2_rows<KY, CONWAY, true>(p, src, dst, x0, y0, active, xs, xm, xpad, lvirt, rvirt, fill_l, fill_r);
}

template <int KY, bool CONWAY>
void launch_life2(const MapParams& p, hipStream_t s) {
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y;
  hipLaunchKernelGGL((life2_kernel<KY, CONWAY>), dim3(nb), dim3(256), 0, s, p);
}

}  // namespace sth
