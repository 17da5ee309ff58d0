// Two steps of a 3-D radius-1 stencil update per launch (temporal blocking)
// for the steady steps of an iteration: generation t+2 from generation t,
// with generation t+1 held in registers, so HBM sees about one read and one
// write per two steps.
//
// Every wave works alone (no LDS, no barriers).  It owns 64 columns
// x = X0-2 .. X0+61 (one per lane) and R rows y = Y0-2 .. Y0+R-3 and marches
// through a chunk of z-planes.  Per plane it keeps generation-t planes in a
// four-deep register queue (the fourth slot receives the plane two ahead,
// so loads are in flight while the current plane is computed), computes
// generation t+1 on lanes 1..62 and rows 1..R-2 into a second queue, and
// generation t+2 on lanes 2..61 and rows 2..R-3 of the plane behind, which
// it stores: 60 x (R-4) results per plane.  Neighbours along x come from the
// adjacent lanes (DPP wave shifts), along y from the lane's other rows and
// along z from the queue.  Chunks overlap by two planes on each side.
//
// Every value is computed with the shared fold expressions (fold_step /
// fold_finish) in offset order, and generation t+1 is rounded to the storage
// type exactly as a store and reload would, so results are bit-identical to
// two single steps.  Cells outside the array follow the reference's
// getneighbor (src/array.jl:129-179) in every generation:
//   - Wrap: an outside cell of generation t+1 is computed from its mapped
//     neighbourhood -- Wrap is a translation, so that is the wrapped cell's
//     value, with the same fold order;
//   - Remove: outside cells read padval in every generation;
//   - Reflect: a fold over the mirrored neighbourhood would sum in the
//     reverse order, so the outside generation-t+1 cells next to the array
//     are replaced by the cells they mirror, which the reference reads.
// Lanes, rows and planes beyond
// the two-cell rim of a partial tile are clamped into the rim; they feed no
// stored result.
#pragma once

#include "tile_kernel.hpp"

namespace sth {

constexpr int kZm2TX = 60;  // output columns per wave

template <typename Val>
constexpr int zmarch2_rows() {
  return sizeof(Val) == 8 ? 10 : 16;
}

// Lane l receives lane l - 1's value (DX = -1) or lane l + 1's (DX = +1);
// the lane with no such neighbour receives 0 (it is a rim lane whose result
// is never used).
template <int DX, typename Val>
__device__ __forceinline__ Val lane_shift(Val v) {
  // old value undefined and bound_ctrl set: the move can fold into the
  // instruction that consumes it as a DPP operand
  constexpr int ctrl = DX < 0 ? 0x138 /* wave_shr:1 */ : 0x130 /* wave_shl:1 */;
  if constexpr (sizeof(Val) == 8) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t l2 = static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(static_cast<uint32_t>(u)), ctrl, 0xf, 0xf, true));
    const uint32_t h2 = static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(static_cast<uint32_t>(u >> 32)), ctrl, 0xf, 0xf, true));
    return __builtin_bit_cast(Val, (static_cast<uint64_t>(h2) << 32) | l2);
  } else if constexpr (sizeof(Val) == 4) {
    return __builtin_bit_cast(Val, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xf, 0xf, true));
  } else if constexpr (sizeof(Val) == 2) {
    const int w = __builtin_bit_cast(uint16_t, v);
    return __builtin_bit_cast(Val, static_cast<uint16_t>(__builtin_amdgcn_mov_dpp(w, ctrl, 0xf, 0xf, true)));
  } else {
    static_assert(sizeof(Val) == 1, "1-, 2-, 4- or 8-byte values");
    const int w = __builtin_bit_cast(uint8_t, v);
    return __builtin_bit_cast(Val, static_cast<uint8_t>(__builtin_amdgcn_mov_dpp(w, ctrl, 0xf, 0xf, true)));
  }
}

template <int DX, typename Val>
__device__ __forceinline__ Val shifted(Val v) {
  if constexpr (DX == 0) return v;
  else return lane_shift<DX>(v);
}

// One fold over the stencil at row r: plane slots q0 (dz = -1), q1, q2.
template <typename S, int OPK, class Sh, int R, int I, class P>
__device__ __forceinline__ void zm2_fold_terms(typename AccOf<S, OPK>::type& acc, uint32_t& cnt,
                                               const typename Staged<S>::type (&q0)[R],
                                               const typename Staged<S>::type (&q1)[R],
                                               const typename Staged<S>::type (&q2)[R], int r, const P& p) {
  if constexpr (I < Sh::L) {
    constexpr int dx = Sh::off[I][0], dy = Sh::off[I][1], dz = Sh::off[I][2];
    const auto& q = dz < 0 ? q0 : (dz == 0 ? q1 : q2);
    fold_step<S, OPK>(acc, cnt, I, shifted<dx>(q[r + dy]), p);
    zm2_fold_terms<S, OPK, Sh, R, I + 1>(acc, cnt, q0, q1, q2, r, p);
  }
}

template <typename S, int OPK, class Sh, int R, int BC>
__global__ __launch_bounds__(256) void zmarch2_kernel(MapParams p) {
  constexpr bool REMOVE = BC == STH_REMOVE, REFLECT = BC == STH_REFLECT;
  using Raw = typename Elem<S>::raw;
  using Val = typename Staged<S>::type;
  using A = typename AccOf<S, OPK>::type;
  using O = typename OutOf<S, OPK>::type;
  using ORaw = typename Elem<O>::raw;
  static_assert(Sh::RX <= 1 && Sh::RY <= 1 && Sh::RZ <= 1, "radius-1 stencils only");

  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nblocks = p.tiles_x * p.tiles_y * p.tiles_z;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bx = bid % p.tiles_x;
  const int bt = bid / p.tiles_x;
  const int by = bt % p.tiles_y;
  const int bz = bt / p.tiles_y;

  const int32_t n0 = static_cast<int32_t>(p.src.n0), n1 = static_cast<int32_t>(p.src.n1),
                n2 = static_cast<int32_t>(p.src.n2);
  const int32_t Y0 = (by * 4 + wave) * (R - 4);
  const int32_t zc = (n2 + p.tiles_z - 1) / p.tiles_z;
  const int32_t zb = bz * zc;
  const int32_t ze = zb + zc < n2 ? zb + zc : n2;
  if (Y0 >= n1 || zb >= ze) return;  // uniform: the wave has no output
  const int32_t X0 = bx * kZm2TX;
  const Val pv = stage<S>(raw_from_bits<Raw>(p.padval));
  constexpr int64_t es = sizeof(Raw), eo = sizeof(ORaw);

  // Byte offsets are 32-bit: the host runs this kernel only on arrays whose
  // cells all lie below 2 GiB from the base.  Loads never leave the array
  // (every coordinate is mapped into it).  Stores from lanes and rows that
  // own no result use offset 2^31, past the destination's extent, which the
  // buffer unit's range check drops.
  const __amdgpu_buffer_rsrc_t src = buf_rsrc(p.src.base, 0xffffffffu);
  const int64_t dext = ((p.dst.n2 - 1) * p.dst.s2 + (p.dst.n1 - 1) * p.dst.s1 + p.dst.n0) * eo;
  const __amdgpu_buffer_rsrc_t dst = buf_rsrc(p.dst.base, static_cast<uint32_t>(dext));
  constexpr uint32_t kDrop = 0x80000000u;

  // mapped coordinate of c in [-2, n + 1] (beyond: clamped first)
  auto map_c = [&](int32_t c, int32_t n) -> int32_t {
    c = c < -2 ? -2 : (c > n + 1 ? n + 1 : c);
    if (c < 0) return REMOVE ? 0 : (REFLECT ? -c : c + n);
    if (c >= n) return REMOVE ? 0 : (REFLECT ? 2 * n - 2 - c : c - n);
    return c;
  };
  const int32_t x = X0 - 2 + lane;
  const bool xout = x < 0 || x >= n0;
  const uint32_t xb = static_cast<uint32_t>(map_c(x, n0)) * static_cast<uint32_t>(es);
  const bool store_lane = lane >= 2 && lane < 62 && x < n0;
  const uint32_t xbo = store_lane ? static_cast<uint32_t>(x) * static_cast<uint32_t>(eo) : kDrop;
  // per-row load offsets (lane column + mapped row), per-row pad flags
  uint32_t vrow[R];
  bool yout[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int32_t y = Y0 - 2 + r;
    yout[r] = y < 0 || y >= n1;
    vrow[r] = xb + static_cast<uint32_t>(static_cast<int64_t>(map_c(y, n1)) * p.src.s1 * es);
  }

  Val g0[4][R];  // generation t: plane i in slot i % 4 (i = 0 is zb - 2)
  Val g1[4][R];  // generation t+1: plane j in slot j % 4 (j = 0 is zb - 1)

  auto load_plane = [&](int i, Val (&q)[R]) {
    const int32_t z = zb - 2 + i;
    const uint32_t zoff = static_cast<uint32_t>(static_cast<int64_t>(map_c(z, n2)) * p.src.s2 * es);
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = stage<S>(buf_load<Raw>(src, vrow[r], zoff));
    if constexpr (REMOVE) {
      const bool zout = z < 0 || z >= n2;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (xout || yout[r] || zout) q[r] = pv;
    }
  };

  const int nit = (ze - zb) + 2;  // generation t+1 planes zb-1 .. ze
  load_plane(0, g0[0]);
  load_plane(1, g0[1]);
  load_plane(2, g0[2]);
  for (int j0 = 0; j0 < nit; j0 += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = j0 + u;
      if (j >= nit) break;
      // prefetch generation-t plane j + 3 into the slot plane j - 1 held
      if (j + 3 <= nit + 1) load_plane(j + 3, g0[(u + 3) & 3]);
      // generation t+1 on plane zb - 1 + j, rows 1 .. R-2
      {
        const auto& q0 = g0[u & 3];
        const auto& q1 = g0[(u + 1) & 3];
        const auto& q2 = g0[(u + 2) & 3];
        auto& h = g1[u & 3];
#pragma unroll
        for (int r = 1; r < R - 1; ++r) {
          A acc = A(0);
          uint32_t cnt = 0;
          zm2_fold_terms<S, OPK, Sh, R, 0>(acc, cnt, q0, q1, q2, r, p);
          h[r] = stage<S>(narrow<O>(fold_finish<S, OPK>(acc, cnt, q1[r], Sh::L, p)));
        }
        if constexpr (REMOVE) {
          const int32_t z = zb - 1 + j;
          const bool zout = z < 0 || z >= n2;
#pragma unroll
          for (int r = 1; r < R - 1; ++r)
            if (xout || yout[r] || zout) h[r] = pv;
        }
        if constexpr (REFLECT) {
          // The generation-(t+1) cells just outside the array take the values
          // of the cells they mirror (x = -1 <- 1, x = n0 <- n0 - 2, and the
          // same in y), which is where the reference reads them
          // (src/array.jl:154-166); planes are handled where generation t+2
          // reads them.
          if (X0 == 0) {  // x = -1 is lane 1, x = 1 lane 3
#pragma unroll
            for (int r = 1; r < R - 1; ++r) {
              const Val v = lane_shift<1>(lane_shift<1>(h[r]));
              if (lane == 1) h[r] = v;
            }
          }
          const int ln = n0 - X0 + 2;  // lane of x = n0
          if (ln >= 1 && ln <= 62) {
#pragma unroll
            for (int r = 1; r < R - 1; ++r) {
              const Val v = lane_shift<-1>(lane_shift<-1>(h[r]));
              if (lane == ln) h[r] = v;
            }
          }
          if (Y0 == 0) h[1] = h[3];  // y = -1 <- y = 1
          const int rn = n1 - Y0 + 2;  // row of y = n1
#pragma unroll
          for (int r = 3; r < R - 1; ++r)
            if (r == rn) h[r] = h[r - 2];
        }
      }
      // generation t+2 on plane zb - 2 + j, rows 2 .. R-3
      if (j >= 2) {
        const int64_t z = zb - 2 + j;
        // Reflect: planes z = -1 and z = n2 read planes 1 and n2 - 2
        const bool zlo = REFLECT && z == 0, zhi = REFLECT && z == n2 - 1;
        const auto& q0 = zlo ? g1[u & 3] : g1[(u + 2) & 3];
        const auto& q1 = g1[(u + 3) & 3];
        const auto& q2 = zhi ? g1[(u + 2) & 3] : g1[u & 3];
        const uint32_t zoff = static_cast<uint32_t>(z * p.dst.s2 * eo);
#pragma unroll
        for (int r = 2; r < R - 2; ++r) {
          A acc = A(0);
          uint32_t cnt = 0;
          zm2_fold_terms<S, OPK, Sh, R, 0>(acc, cnt, q0, q1, q2, r, p);
          const auto res = narrow<O>(fold_finish<S, OPK>(acc, cnt, q1[r], Sh::L, p));
          const int32_t y = Y0 - 2 + r;
          const uint32_t vo = y < n1 ? xbo + static_cast<uint32_t>(static_cast<int64_t>(y) * p.dst.s1 * eo) : kDrop;
          buf_store<ORaw>(res, dst, xbo == kDrop ? kDrop : vo, zoff);
        }
      }
    }
  }
}

// Grid of zmarch2_kernel for an n0 x n1 x n2 array: 60-column tiles, groups of
// four R-4-row strips, and z-chunks sized so the launch has about `waves`
// waves (at least 8 planes per chunk).
inline void zmarch2_grid(int64_t n0, int64_t n1, int64_t n2, int rows, int64_t waves, int32_t* tx,
                         int32_t* ty, int32_t* tz) {
  const int64_t gx = (n0 + kZm2TX - 1) / kZm2TX;
  const int64_t gy = (n1 + 4 * (rows - 4) - 1) / (4 * (rows - 4));
  const int64_t xy_waves = gx * gy * 4;
  int64_t chunks = (waves + xy_waves - 1) / xy_waves;
  const int64_t max_chunks = (n2 + 7) / 8;
  chunks = chunks < 1 ? 1 : (chunks > max_chunks ? max_chunks : chunks);
  const int64_t zc = (n2 + chunks - 1) / chunks;
  *tx = static_cast<int32_t>(gx);
  *ty = static_cast<int32_t>(gy);
  *tz = static_cast<int32_t>((n2 + zc - 1) / zc);
}

// Largest byte extent (from the base) the kernel's 32-bit offsets address.
constexpr int64_t kZm2MaxBytes = int64_t(1) << 31;

template <typename S, int OPK, class Sh>
void launch_zmarch2(const MapParams& p, hipStream_t s) {
  constexpr int R = zmarch2_rows<typename Staged<S>::type>();
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y * p.tiles_z;
  if (p.bc == STH_REMOVE)
    hipLaunchKernelGGL((zmarch2_kernel<S, OPK, Sh, R, STH_REMOVE>), dim3(nb), dim3(256), 0, s, p);
  else if (p.bc == STH_REFLECT)
    hipLaunchKernelGGL((zmarch2_kernel<S, OPK, Sh, R, STH_REFLECT>), dim3(nb), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((zmarch2_kernel<S, OPK, Sh, R, STH_WRAP>), dim3(nb), dim3(256), 0, s, p);
}

}  // namespace sth
