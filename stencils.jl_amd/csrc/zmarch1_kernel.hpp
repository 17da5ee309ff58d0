// One step of a 3-D stencil of reach <= 1 per dim (VonNeumann(1,3),
// Moore(1,3)), marching along z: every source plane is loaded once per wave
// and feeds the three output planes that read it.
//
// Every wave works alone (no LDS, no barriers).  A lane owns V consecutive
// cells along x (one V * sizeof(T)-byte load per row), so a wave covers 64 V
// columns x = X0 .. X0 + 64 V - 1, and R output rows y = Y0 .. Y0 + R - 1.
// Per plane it loads rows Y0 - 1 .. Y0 + R (R + 2 row loads) into a D-deep
// register queue -- three planes feed the plane being computed and the
// other D - 3 receive the planes ahead, so their loads are in flight while
// the current plane is folded -- plus, per row a stencil reads with an x
// offset, one more load whose lane 0 gets x = X0 - 1 and whose lane 63 gets
// x = X0 + 64 V (the cells beyond the wave).  x neighbours inside the wave
// come from the lane's own vector or the adjacent lane's (DPP wave shifts),
// y neighbours from the lane's other rows and z neighbours from the queue.
// A wave marches through a chunk of planes; chunks read one plane beyond
// each end.
//
// Every value is computed with the shared fold expressions (fold_step /
// fold_finish) in offset order, so results are bit-identical to the
// streaming and LDS kernels.  Coordinates outside the array are resolved by
// the reference's getneighbor (src/array.jl:129-179): Wrap and Reflect map
// the index, Remove substitutes padval.  With Halo padding the source halo
// is refreshed first (update_boundary!) by the caller, and this kernel maps
// indices itself, which reads the same values.
#pragma once

#include "zmarch2_kernel.hpp"

namespace sth {

// Raw vector of V elements of a storage type, loaded or stored as one
// V * sizeof(Raw)-byte buffer access.
template <typename Raw, int V>
struct ZVec {
  Raw e[V];
};

template <typename Raw, int V>
__device__ __forceinline__ ZVec<Raw, V> zvec_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  constexpr int B = V * static_cast<int>(sizeof(Raw));
  ZVec<Raw, V> v;
  if constexpr (B == 2) {
    v = __builtin_bit_cast(ZVec<Raw, V>, buf_load<uint16_t>(r, voff, soff));
  } else if constexpr (B == 4) {
    v = __builtin_bit_cast(ZVec<Raw, V>, buf_load<uint32_t>(r, voff, soff));
  } else if constexpr (B == 8) {
    v = __builtin_bit_cast(ZVec<Raw, V>, buf_load<uint64_t>(r, voff, soff));
  } else {
    static_assert(B == 16, "2-, 4-, 8- or 16-byte vectors");
    using U4 = unsigned int __attribute__((ext_vector_type(4)));
    v = __builtin_bit_cast(ZVec<Raw, V>, buf_load<U4>(r, voff, soff));
  }
  return v;
}

template <typename Raw, int V, int AUX>
__device__ __forceinline__ void zvec_store(const ZVec<Raw, V>& v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                           uint32_t soff) {
  constexpr int B = V * static_cast<int>(sizeof(Raw));
  if constexpr (B == 2) {
    buf_store<uint16_t, AUX>(__builtin_bit_cast(uint16_t, v), r, voff, soff);
  } else if constexpr (B == 4) {
    buf_store<uint32_t, AUX>(__builtin_bit_cast(uint32_t, v), r, voff, soff);
  } else if constexpr (B == 8) {
    buf_store<uint64_t, AUX>(__builtin_bit_cast(uint64_t, v), r, voff, soff);
  } else {
    static_assert(B == 16, "2-, 4-, 8- or 16-byte vectors");
    using U4 = unsigned int __attribute__((ext_vector_type(4)));
    buf_store<U4, AUX>(__builtin_bit_cast(U4, v), r, voff, soff);
  }
}

// Which rows (r = 0 .. R+1, i.e. y = Y0 - 1 + r) and planes a stencil reads
// with an x offset: those need the wave's edge cells.
template <class Sh>
__host__ __device__ constexpr bool zm1_reads_x() {
  for (int k = 0; k < Sh::L; ++k)
    if (Sh::off[k][0] != 0) return true;
  return false;
}

// The value at x offset DX (-1, 0, +1) from cell I of a lane's vector, for
// one row of one plane: inside the vector, from the adjacent lane, or, at the
// wave's ends, from the edge load (`e`: lane 0 holds x = X0 - 1, lane 63
// holds x = X0 + 64 V).
template <int DX, int I, int V, typename Val>
__device__ __forceinline__ Val zm1_at(const Val (&v)[V], Val e, int lane) {
  if constexpr (I + DX >= 0 && I + DX < V) {
    return v[I + DX];
  } else if constexpr (I + DX < 0) {
    const Val t = lane_shift<-1>(v[V - 1]);
    return lane == 0 ? e : t;
  } else {
    const Val t = lane_shift<1>(v[0]);
    return lane == 63 ? e : t;
  }
}

template <typename S, int OPK, class Sh, int V, int R, int I, int K>
__device__ __forceinline__ void zm1_fold(typename AccOf<S, OPK>::type& acc, uint32_t& cnt,
                                         const typename Staged<S>::type (&q0)[R + 2][V],
                                         const typename Staged<S>::type (&q1)[R + 2][V],
                                         const typename Staged<S>::type (&q2)[R + 2][V],
                                         const typename Staged<S>::type (&e0)[R + 2],
                                         const typename Staged<S>::type (&e1)[R + 2],
                                         const typename Staged<S>::type (&e2)[R + 2], int r, int lane,
                                         const MapParams& p) {
  if constexpr (K < Sh::L) {
    constexpr int dx = Sh::off[K][0], dy = Sh::off[K][1], dz = Sh::off[K][2];
    const auto& q = dz < 0 ? q0 : (dz == 0 ? q1 : q2);
    const auto& e = dz < 0 ? e0 : (dz == 0 ? e1 : e2);
    fold_step<S, OPK>(acc, cnt, K, zm1_at<dx, I, V>(q[r + dy], e[r + dy], lane), p);
    zm1_fold<S, OPK, Sh, V, R, I, K + 1>(acc, cnt, q0, q1, q2, e0, e1, e2, r, lane, p);
  }
}

template <typename S, int OPK, class Sh, int V, int R, int I>
__device__ __forceinline__ void zm1_cells(typename Elem<typename OutOf<S, OPK>::type>::raw (&out)[V],
                                          const typename Staged<S>::type (&q0)[R + 2][V],
                                          const typename Staged<S>::type (&q1)[R + 2][V],
                                          const typename Staged<S>::type (&q2)[R + 2][V],
                                          const typename Staged<S>::type (&e0)[R + 2],
                                          const typename Staged<S>::type (&e1)[R + 2],
                                          const typename Staged<S>::type (&e2)[R + 2], int r, int lane,
                                          const MapParams& p) {
  if constexpr (I < V) {
    using A = typename AccOf<S, OPK>::type;
    using O = typename OutOf<S, OPK>::type;
    A acc = A(0);
    uint32_t cnt = 0;
    zm1_fold<S, OPK, Sh, V, R, I, 0>(acc, cnt, q0, q1, q2, e0, e1, e2, r, lane, p);
    out[I] = narrow<O>(fold_finish<S, OPK>(acc, cnt, q1[r][I], Sh::L, p));
    zm1_cells<S, OPK, Sh, V, R, I + 1>(out, q0, q1, q2, e0, e1, e2, r, lane, p);
  }
}

template <typename S, int OPK, class Sh, int V, int R, int D, int NT, int BC>
__global__ __launch_bounds__(256) void zmarch1_kernel(MapParams p) {
  constexpr bool REMOVE = BC == STH_REMOVE, REFLECT = BC == STH_REFLECT;
  using Raw = typename Elem<S>::raw;
  using Val = typename Staged<S>::type;
  using O = typename OutOf<S, OPK>::type;
  using ORaw = typename Elem<O>::raw;
  static_assert(Sh::RX <= 1 && Sh::RY <= 1 && Sh::RZ <= 1, "reach-1 stencils only");
  static_assert(D >= 4, "three planes in use and at least one ahead");
  constexpr bool XN = zm1_reads_x<Sh>();

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
  const int32_t Y0 = (by * 4 + wave) * R;
  const int32_t zc = (n2 + p.tiles_z - 1) / p.tiles_z;
  const int32_t zb = bz * zc;
  const int32_t ze = zb + zc < n2 ? zb + zc : n2;
  if (Y0 >= n1 || zb >= ze) return;  // uniform: the wave has no output
  const int32_t X0 = bx * 64 * V;
  const Val pv = stage<S>(raw_from_bits<Raw>(p.padval));
  constexpr int64_t es = sizeof(Raw), eo = sizeof(ORaw);

  // Byte offsets are 32-bit: the host runs this kernel only on arrays whose
  // cells all lie below 2 GiB from the base.  Loads never leave the array
  // (every coordinate is mapped into it).  Stores of cells outside the array
  // use offset 2^31, past the destination's extent, which the buffer unit's
  // range check drops.
  const __amdgpu_buffer_rsrc_t src = buf_rsrc(p.src.base, 0xffffffffu);
  const int64_t dext = ((p.dst.n2 - 1) * p.dst.s2 + (p.dst.n1 - 1) * p.dst.s1 + p.dst.n0) * eo;
  const __amdgpu_buffer_rsrc_t dst = buf_rsrc(p.dst.base, static_cast<uint32_t>(dext));
  constexpr uint32_t kDrop = 0x80000000u;

  // mapped coordinate of c in [-1, n] (beyond: clamped first; those cells
  // feed no stored result)
  auto map_c = [&](int32_t c, int32_t n) -> int32_t {
    c = c < -1 ? -1 : (c > n ? n : c);
    if (c < 0) return REMOVE ? 0 : (REFLECT ? 1 : n - 1);
    if (c >= n) return REMOVE ? 0 : (REFLECT ? n - 2 : 0);
    return c;
  };
  // The lane's vector starts at x0 = X0 + V lane.  A vector either lies
  // wholly inside the row or starts at or past its end (n0 is a multiple of
  // V, checked on the host).  A lane past the end feeds only its cell 0, as
  // the right neighbour x = n0 of the lane before it, so it loads the vector
  // that starts at the mapped x0: Wrap gives x = 0 and Reflect x = n0 - 2,
  // both multiples of V, and cell 0 is then the mapped cell.
  const int32_t x0 = X0 + V * lane;
  const bool vin = x0 < n0;
  const uint32_t xb = static_cast<uint32_t>(map_c(x0, n0)) * static_cast<uint32_t>(es);
  const int32_t xe = lane == 0 ? X0 - 1 : X0 + 64 * V;  // edge cell of lanes 0 and 63
  const bool eout = xe < 0 || xe >= n0;
  const uint32_t eb = static_cast<uint32_t>(map_c(xe, n0)) * static_cast<uint32_t>(es);
  const bool xout_vec = !vin;  // Remove: the lane's cells are all padval
  const bool store_lane = vin;
  const uint32_t xbo = store_lane ? static_cast<uint32_t>(x0) * static_cast<uint32_t>(eo) : kDrop;

  uint32_t roff[R + 2];
  bool yout[R + 2];
#pragma unroll
  for (int r = 0; r < R + 2; ++r) {
    const int32_t y = Y0 - 1 + r;
    yout[r] = y < 0 || y >= n1;
    roff[r] = static_cast<uint32_t>(static_cast<int64_t>(map_c(y, n1)) * p.src.s1 * es);
  }

  Val q[D][R + 2][V];  // plane i (i = 0 is zb - 1) in slot i % D
  Val e[D][R + 2];     // its edge cells (lanes 0 and 63)

  auto load_plane = [&](int i, Val (&qq)[R + 2][V], Val (&ee)[R + 2]) {
    const int32_t z = zb - 1 + i;
    const uint32_t zoff = static_cast<uint32_t>(static_cast<int64_t>(map_c(z, n2)) * p.src.s2 * es);
    const bool zout = z < 0 || z >= n2;
#pragma unroll
    for (int r = 0; r < R + 2; ++r) {
      const ZVec<Raw, V> w = zvec_load<Raw, V>(src, roff[r] + xb, zoff);
#pragma unroll
      for (int c = 0; c < V; ++c) qq[r][c] = stage<S>(w.e[c]);
      if constexpr (XN) ee[r] = stage<S>(buf_load<Raw>(src, roff[r] + eb, zoff));
      if constexpr (REMOVE) {
        const bool rowout = yout[r] || zout;
#pragma unroll
        for (int c = 0; c < V; ++c)
          if (rowout || xout_vec) qq[r][c] = pv;
        if constexpr (XN)
          if (rowout || eout) ee[r] = pv;
      }
    }
  };

  const int nit = ze - zb;  // output planes zb .. ze - 1
  // Every load is unconditional (planes past the chunk's end reload its last
  // plane, which nothing reads), so the compiler can count the loads in
  // flight and wait for exactly the ones a plane needs.
  const int last = nit + 1;
#pragma unroll
  for (int i = 0; i < D - 1; ++i) load_plane(i < last ? i : last, q[i], e[i]);
  for (int j0 = 0; j0 < nit; j0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 + u;
      if (j >= nit) break;
      // output plane zb + j reads planes j, j + 1, j + 2 (slots u, u+1, u+2);
      // prefetch plane j + D - 1 into the slot plane j - 1 held
      load_plane(j + D - 1 < last ? j + D - 1 : last, q[(u + D - 1) % D], e[(u + D - 1) % D]);
      const int32_t z = zb + j;
      const uint32_t zoff = static_cast<uint32_t>(static_cast<int64_t>(z) * p.dst.s2 * eo);
      const auto& q0 = q[u % D];
      const auto& q1 = q[(u + 1) % D];
      const auto& q2 = q[(u + 2) % D];
      const auto& e0 = e[u % D];
      const auto& e1 = e[(u + 1) % D];
      const auto& e2 = e[(u + 2) % D];
#pragma unroll
      for (int r = 1; r <= R; ++r) {
        ZVec<ORaw, V> out;
        zm1_cells<S, OPK, Sh, V, R, 0>(out.e, q0, q1, q2, e0, e1, e2, r, lane, p);
        const int32_t y = Y0 - 1 + r;
        const uint32_t vo = (y < n1 && xbo != kDrop)
                                ? xbo + static_cast<uint32_t>(static_cast<int64_t>(y) * p.dst.s1 * eo)
                                : kDrop;
        zvec_store<ORaw, V, NT ? 2 : 0>(out, dst, vo, zoff);
      }
    }
  }
}

// Grid of zmarch1_kernel for an n0 x n1 x n2 array: 64 V-column tiles,
// groups of four R-row strips, and z-chunks sized so the launch has about
// `waves` waves (at least 8 planes per chunk).
inline void zmarch1_grid(int64_t n0, int64_t n1, int64_t n2, int v, int rows, int64_t waves, int32_t* tx,
                         int32_t* ty, int32_t* tz) {
  const int64_t gx = (n0 + 64 * v - 1) / (64 * v);
  const int64_t gy = (n1 + 4 * rows - 1) / (4 * rows);
  const int64_t xy_waves = gx * gy * 4;
  int64_t chunks = (waves + xy_waves - 1) / xy_waves;
  const int64_t max_chunks = (n2 + 7) / 8;
  chunks = chunks < 1 ? 1 : (chunks > max_chunks ? max_chunks : chunks);
  const int64_t zc = (n2 + chunks - 1) / chunks;
  *tx = static_cast<int32_t>(gx);
  *ty = static_cast<int32_t>(gy);
  *tz = static_cast<int32_t>((n2 + zc - 1) / zc);
}

template <typename S, int OPK, class Sh, int V, int R, int D, int NT>
void launch_zmarch1(const MapParams& p, hipStream_t s) {
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y * p.tiles_z;
  if (p.bc == STH_REMOVE)
    hipLaunchKernelGGL((zmarch1_kernel<S, OPK, Sh, V, R, D, NT, STH_REMOVE>), dim3(nb), dim3(256), 0, s, p);
  else if (p.bc == STH_REFLECT)
    hipLaunchKernelGGL((zmarch1_kernel<S, OPK, Sh, V, R, D, NT, STH_REFLECT>), dim3(nb), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((zmarch1_kernel<S, OPK, Sh, V, R, D, NT, STH_WRAP>), dim3(nb), dim3(256), 0, s, p);
}

}  // namespace sth
