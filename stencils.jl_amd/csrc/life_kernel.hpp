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

struct Row16 {
  uint32_t w[4];  // normalised cells
  uint32_t h[4];  // horizontal sums (left + centre + right), bytes <= 3
};

template <int KY>
__global__ __launch_bounds__(256) void life_kernel(MapParams p) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nblocks = p.tiles_x * p.tiles_y;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bcol = bid % p.tiles_x;
  const int brow = bid / p.tiles_x;

  const int64_t n0 = p.src.n0, n1 = p.src.n1;
  const int64_t x0 = (static_cast<int64_t>(bcol) * 64 + lane) * 16;
  const int64_t y0 = (static_cast<int64_t>(brow) * 4 + wave) * KY;
  const bool active = x0 < n0;
  const int bc = p.bc;
  const bool halo = p.read_halo != 0;
  const uint32_t padn = (p.padval & 0xffu) ? 0x01010101u : 0u;
  const bool conway = p.op.birth == (1ull << 3) && p.op.survive == ((1ull << 2) | (1ull << 3));
  const uint32_t tbl = static_cast<uint32_t>(p.op.birth & 0x1ffu) |
                       (static_cast<uint32_t>(p.op.survive & 0x1ffu) << 9);
  const uint8_t* src = reinterpret_cast<const uint8_t*>(p.src.base);
  uint8_t* dst = reinterpret_cast<uint8_t*>(p.dst.base);

  // Byte of row y (already mapped / halo row) at column x (-1 .. n0).
  auto edge_byte = [&](const uint8_t* row, bool padrow, int64_t x) -> uint32_t {
    if (padrow) return padn & 1u;
    if (halo || (x >= 0 && x < n0)) return row[x] ? 1u : 0u;
    if (bc == STH_REMOVE) return padn & 1u;
    return row[map_index(x, n0, bc)] ? 1u : 0u;
  };

  auto load_row = [&](int64_t y, Row16& r) {
    const bool need = y >= -1 && y <= n1;
    bool padrow = false;
    int64_t ym = y;
    if (!halo && (y < 0 || y >= n1)) {
      if (bc == STH_REMOVE) padrow = true;
      else ym = map_index(y, n1, bc);
    }
    const uint8_t* row = src + ym * p.src.s1;
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    if (need && active) {
      if (padrow) {
        w0 = w1 = w2 = w3 = padn;
      } else {
        const uint4 v = *reinterpret_cast<const uint4*>(row + x0);
        w0 = norm01(v.x);
        w1 = norm01(v.y);
        w2 = norm01(v.z);
        w3 = norm01(v.w);
      }
    }
    // Neighbouring lanes hold the adjacent chunks.
    uint32_t left = __shfl_up(w3, 1) >> 24;
    uint32_t right = __shfl_down(w0, 1) & 0xffu;
    if (need && active) {
      if (lane == 0 || x0 == 0) left = edge_byte(row, padrow, x0 - 1);
      if (lane == 63 || x0 + 16 == n0) right = edge_byte(row, padrow, x0 + 16);
    }
    r.w[0] = w0;
    r.w[1] = w1;
    r.w[2] = w2;
    r.w[3] = w3;
    r.h[0] = w0 + __builtin_amdgcn_alignbyte(w0, left << 24, 3) + __builtin_amdgcn_alignbyte(w1, w0, 1);
    r.h[1] = w1 + __builtin_amdgcn_alignbyte(w1, w0, 3) + __builtin_amdgcn_alignbyte(w2, w1, 1);
    r.h[2] = w2 + __builtin_amdgcn_alignbyte(w2, w1, 3) + __builtin_amdgcn_alignbyte(w3, w2, 1);
    r.h[3] = w3 + __builtin_amdgcn_alignbyte(w3, w2, 3) + __builtin_amdgcn_alignbyte(right, w3, 1);
  };

  Row16 a, b, c;
  load_row(y0 - 1, a);
  load_row(y0, b);
#pragma unroll
  for (int j = 0; j < KY; ++j) {
    const int64_t y = y0 + j;
    load_row(y + 1, c);
    uint4 out;
    uint32_t o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t s = a.h[i] + b.h[i] + c.h[i] - b.w[i];  // Moore count, bytes <= 8
      o[i] = conway ? rule_conway(s, b.w[i]) : rule_table(s, b.w[i], tbl);
    }
    out.x = o[0];
    out.y = o[1];
    out.z = o[2];
    out.w = o[3];
    if (active && y < n1) *reinterpret_cast<uint4*>(dst + y * p.dst.s1 + x0) = out;
    a = b;
    b = c;
  }

  if (p.fuse_halo && active) {
    // update_boundary! of the source: halo cell (x, y) takes the value of the
    // boundary-mapped interior cell, or padval.
    uint8_t* sw = reinterpret_cast<uint8_t*>(p.src.base);
    const uint8_t pv = static_cast<uint8_t>(p.padval & 0xffu);
    const int64_t s1 = p.src.s1;
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

template <int KY>
void launch_life(const MapParams& p, hipStream_t s) {
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y;
  hipLaunchKernelGGL(life_kernel<KY>, dim3(nb), dim3(256), 0, s, p);
}

}  // namespace sth
