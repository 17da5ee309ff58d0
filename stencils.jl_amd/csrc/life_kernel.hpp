// Life-like automata on 1-byte cells (UInt8 / Bool) with Moore(1) in 2-D:
// the Game of Life hot loop, bandwidth-bound at 2 bytes per cell-step.
//
// Each lane owns a 16-byte column chunk (16 cells) and marches down KY rows,
// so every source row is loaded once per lane (one 16-byte vector load) and
// reused for three output rows.  Cells are processed four per 32-bit word
// (SWAR): bytes are normalised to 0/1, horizontal neighbours come from
// byte-aligned funnel shifts (v_alignbyte) with the chunk's edge bytes taken
// from the neighbouring lanes, vertical sums are plain word adds (a count
// never exceeds 8, so no byte carries), and the rule is a bit lookup
// (Conway's B3/S23 reduces to `(s | c) == 3`).
//
// Boundaries: rows -1 and n1 and columns -1 and n0 are mapped per Remove /
// Wrap / Reflect exactly as the reference's getneighbor does
// (src/array.jl:129-179), or read from the halo as it is (`Use`).  For Halo
// padding the same launch also refreshes the source halo, which is what
// update_boundary! does before the reference's kernel (src/array.jl:195-242):
// no thread reads the halo during the step, so the writes cannot race.
//
// Preconditions checked on the host: 2-D, n0 % 16 == 0, interior rows of
// source and destination 16-byte aligned.
#pragma once

#include "common.hpp"

namespace sth {

// Cross-lane moves (DPP wave shifts, full rate): lane i receives lane i - 1
// (from_left) or lane i + 1 (from_right); the lane with no such neighbour
// (0, resp. 63) receives `fill`.
__device__ __forceinline__ uint32_t from_left(uint32_t v, uint32_t fill) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(fill), static_cast<int>(v),
                                                          0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}
__device__ __forceinline__ uint32_t from_right(uint32_t v, uint32_t fill) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(fill), static_cast<int>(v),
                                                          0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}

__device__ __forceinline__ uint32_t norm01(uint32_t x) {
  // 0x01 in every byte that is nonzero, 0x00 elsewhere.
  const uint32_t t = ((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x;
  return (t >> 7) & 0x01010101u;
}

__device__ __forceinline__ uint32_t rule_conway(uint32_t s, uint32_t c) {
  // byte-wise (s | c) == 3  <=>  s == 3, or s == 2 and alive
  const uint32_t x = (s | c) ^ 0x03030303u;          // bytes < 0x10
  const uint32_t nz = (x + 0x7f7f7f7fu) | x;         // bit 7 set iff byte != 0
  return (~nz >> 7) & 0x01010101u;
}

__device__ __forceinline__ uint32_t rule_table(uint32_t s, uint32_t c, uint32_t tbl) {
  // tbl bit (s + 9 c): birth mask in bits 0..8, survive mask in bits 9..17
  const uint32_t idx = s + (c << 3) + c;  // bytes <= 17
  uint32_t out = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const uint32_t i = (idx >> (8 * b)) & 0xffu;
    out |= ((tbl >> i) & 1u) << (8 * b);
  }
  return out;
}

// One wave's rows.  EDGE selects the general path (rows mapped per boundary,
// padding rows for Remove); interior waves read consecutive rows.  NORM
// normalises bytes to 0/1; it can be skipped when the source was written by
// this kernel (every interior byte is then 0 or 1).
template <int KY, bool CONWAY, bool NORM, bool EDGE>
__device__ __forceinline__ void life_rows(const MapParams& p, const uint8_t* __restrict__ src,
                                          uint8_t* __restrict__ dst, int32_t x0, int32_t y0,
                                          bool active, int32_t xs, int32_t xlm, int32_t xrm,
                                          bool lpad, bool rpad, bool fill_l, bool fill_r) {
  const int32_t n1 = static_cast<int32_t>(p.src.n1);
  const int64_t s1 = p.src.s1;
  const int bc = p.bc;
  const bool halo = p.read_halo != 0;
  const uint32_t padn = (p.padval & 0xffu) ? 0x01010101u : 0u;
  const uint32_t tbl = static_cast<uint32_t>(p.op.birth & 0x1ffu) |
                       (static_cast<uint32_t>(p.op.survive & 0x1ffu) << 9);

  constexpr int NR = KY + 2;
  uint4 raw[NR];
  uint32_t lb[NR], rb[NR];
  bool padrow[NR];
  const uint8_t* rowp[NR];
  if constexpr (!EDGE) {
    const uint8_t* row = src + static_cast<int64_t>(y0 - 1) * s1;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      padrow[r] = false;
      rowp[r] = row;
      raw[r] = *reinterpret_cast<const uint4*>(row + xs);
      row += s1;
    }
  } else {
    // Rows beyond [-1, n1] exist only in the last, partial block; they are
    // clamped to a valid row and their results never stored.
    const bool wrap = bc == STH_WRAP, refl = bc == STH_REFLECT;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int32_t y = y0 - 1 + r;
      const int32_t yc = y < -1 ? -1 : (y > n1 ? n1 : y);
      const bool out = yc < 0 || yc >= n1;
      const int32_t yw = yc < 0 ? yc + n1 : (yc >= n1 ? yc - n1 : yc);
      const int32_t yr = yc < 0 ? -yc : (yc >= n1 ? 2 * n1 - 2 - yc : yc);
      const int32_t ym = halo ? yc : (wrap ? yw : (refl ? yr : (out ? 0 : yc)));
      padrow[r] = !halo && !wrap && !refl && out;
      rowp[r] = src + static_cast<int64_t>(ym) * s1;
      raw[r] = *reinterpret_cast<const uint4*>(rowp[r] + xs);
    }
  }
  // The edge bytes (x0 - 1, x0 + 16) come from the neighbouring lanes'
  // chunks; lanes at a wave or array edge (fill_l / fill_r) load them,
  // boundary-mapped, after every chunk load is in flight.
  uint32_t fl[NR], fr[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) fl[r] = fr[r] = 0;
  if (fill_l || fill_r) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      fl[r] = static_cast<uint32_t>(rowp[r][xlm]) << 24;
      fr[r] = rowp[r][xrm];
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const uint32_t lw = from_left(raw[r].w, fl[r]);
    const uint32_t rw = from_right(raw[r].x, fr[r]);
    lb[r] = fill_l ? (fl[r] >> 24) : (lw >> 24);
    rb[r] = fill_r ? fr[r] : (rw & 0xffu);
  }

  auto hsum = [&](int r, uint32_t* w, uint32_t* h) {
    uint32_t v0 = raw[r].x, v1 = raw[r].y, v2 = raw[r].z, v3 = raw[r].w, l = lb[r], rr = rb[r];
    if constexpr (NORM) {
      v0 = norm01(v0);
      v1 = norm01(v1);
      v2 = norm01(v2);
      v3 = norm01(v3);
      l = l ? 1u : 0u;
      rr = rr ? 1u : 0u;
    }
    l = lpad ? (padn & 1u) : l;
    rr = rpad ? (padn & 1u) : rr;
    if (EDGE && padrow[r]) {
      v0 = v1 = v2 = v3 = padn;
      l = rr = padn & 1u;
    }
    w[0] = v0;
    w[1] = v1;
    w[2] = v2;
    w[3] = v3;
    h[0] = v0 + __builtin_amdgcn_alignbyte(v0, l << 24, 3) + __builtin_amdgcn_alignbyte(v1, v0, 1);
    h[1] = v1 + __builtin_amdgcn_alignbyte(v1, v0, 3) + __builtin_amdgcn_alignbyte(v2, v1, 1);
    h[2] = v2 + __builtin_amdgcn_alignbyte(v2, v1, 3) + __builtin_amdgcn_alignbyte(v3, v2, 1);
    h[3] = v3 + __builtin_amdgcn_alignbyte(v3, v2, 3) + __builtin_amdgcn_alignbyte(rr, v3, 1);
  };

  uint32_t wa[4], ha[4], wb[4], hb[4], wc[4], hc[4];
  uint4 outs[KY];
  hsum(0, wa, ha);
  hsum(1, wb, hb);
#pragma unroll
  for (int j = 0; j < KY; ++j) {
    hsum(j + 2, wc, hc);
    uint32_t o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t s = ha[i] + hb[i] + hc[i] - wb[i];  // Moore count, bytes <= 8
      o[i] = CONWAY ? rule_conway(s, wb[i]) : rule_table(s, wb[i], tbl);
    }
    outs[j] = make_uint4(o[0], o[1], o[2], o[3]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ha[i] = hb[i];
      hb[i] = hc[i];
      wb[i] = wc[i];
    }
  }
  if (active) {
    uint8_t* d = dst + static_cast<int64_t>(y0) * p.dst.s1 + x0;
#pragma unroll
    for (int j = 0; j < KY; ++j) {
      if (EDGE && y0 + j >= n1) break;
      // (plain stores: nontemporal ones measured 45% slower on MI355X here)
      *reinterpret_cast<uint4*>(d) = outs[j];
      d += p.dst.s1;
    }
  }
}

template <int KY, bool CONWAY, bool NORM>
__global__ __launch_bounds__(256) void life_kernel(MapParams p) {
  const int lane = threadIdx.x & 63;
  // wave index is uniform: keep the row arithmetic on the scalar unit
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nblocks = p.tiles_x * p.tiles_y;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bcol = bid % p.tiles_x;
  const int brow = bid / p.tiles_x;

  const int32_t n0 = static_cast<int32_t>(p.src.n0), n1 = static_cast<int32_t>(p.src.n1);
  const int32_t x0 = (bcol * 64 + lane) * 16;
  const int32_t y0 = (brow * 4 + wave) * KY;
  const bool active = x0 < n0;
  const int bc = p.bc;
  const bool halo = p.read_halo != 0;
  const uint8_t* src = reinterpret_cast<const uint8_t*>(p.src.base);
  uint8_t* dst = reinterpret_cast<uint8_t*>(p.dst.base);
  const int64_t s1 = p.src.s1;

  // Column positions of the two edge bytes each lane needs (left of x0,
  // right of x0 + 15), boundary-mapped; every address stays inside memory so
  // the loads are unconditional and can all be in flight at once.
  const int32_t xs = active ? x0 : 0;
  const int32_t xl = xs - 1, xr = xs + 16;
  const bool lpad = !halo && xl < 0 && bc == STH_REMOVE;
  const bool rpad = !halo && xr >= n0 && bc == STH_REMOVE;
  const int32_t xlm = halo ? xl : (lpad ? 0 : static_cast<int32_t>(map_index(xl, n0, bc)));
  const int32_t xrm = halo ? xr : (rpad ? 0 : static_cast<int32_t>(map_index(xr, n0, bc)));

  // lanes whose neighbour chunk is outside this wave or the array load the
  // edge byte themselves (every lane takes part in the cross-lane moves)
  const bool fill_l = lane == 0 || xs == 0;
  const bool fill_r = lane == 63 || xs + 16 >= n0 || x0 + 16 >= n0;
  if (y0 >= 1 && y0 + KY < n1)
    life_rows<KY, CONWAY, NORM, false>(p, src, dst, x0, y0, active, xs, xlm, xrm, lpad, rpad, fill_l, fill_r);
  else
    life_rows<KY, CONWAY, NORM, true>(p, src, dst, x0, y0, active, xs, xlm, xrm, lpad, rpad, fill_l, fill_r);

  if (p.fuse_halo && active) {
    // update_boundary! of the source: halo cell (x, y) takes the value of the
    // boundary-mapped interior cell, or padval.
    uint8_t* sw = reinterpret_cast<uint8_t*>(p.src.base);
    const uint8_t pv = static_cast<uint8_t>(p.padval & 0xffu);
    auto raw_at = [&](int64_t x, int64_t y) -> uint8_t {
      const bool out = x < 0 || x >= n0 || y < 0 || y >= n1;
      if (out && bc == STH_REMOVE) return pv;
      return src[map_index(y, n1, bc) * s1 + map_index(x, n0, bc)];
    };
    for (int j = 0; j < KY; ++j) {
      const int64_t y = y0 + j;
      if (y >= n1) break;
      if (x0 == 0) sw[y * s1 - 1] = raw_at(-1, y);
      if (x0 + 16 == n0) sw[y * s1 + n0] = raw_at(n0, y);
    }
    const bool top = y0 == 0;
    const bool bottom = y0 <= n1 - 1 && n1 - 1 < y0 + KY;
    for (int side = 0; side < 2; ++side) {
      if (!(side == 0 ? top : bottom)) continue;
      const int64_t y = side == 0 ? -1 : n1;
      for (int b = 0; b < 16; ++b) sw[y * s1 + x0 + b] = raw_at(x0 + b, y);
      if (x0 == 0) sw[y * s1 - 1] = raw_at(-1, y);
      if (x0 + 16 == n0) sw[y * s1 + n0] = raw_at(n0, y);
    }
  }
}

template <int KY, bool CONWAY, bool NORM>
void launch_life(const MapParams& p, hipStream_t s) {
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y;
  hipLaunchKernelGGL((life_kernel<KY, CONWAY, NORM>), dim3(nb), dim3(256), 0, s, p);
}

}  // namespace sth
