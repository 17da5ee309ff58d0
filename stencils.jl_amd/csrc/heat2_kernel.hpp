// Two steps of a 3-D stencil update per launch (temporal blocking) for the
// steady steps of an iteration: generation t+2 from generation t, with
// generation t+1 held in LDS, so HBM sees about one read and one write per
// two steps.
//
// A workgroup owns a TX x TY tile of (x, y) and marches through a chunk of ZC
// z-planes.  Per input plane it loads the generation-t plane of the tile plus
// a two-cell rim into an LDS ring of three planes, computes generation t+1 on
// the tile plus a one-cell rim into a second ring of three planes, and
// generation t+2 on the tile, which it stores.  The next plane is prefetched
// into registers while the current one is computed.  Chunks overlap by two
// planes on each side (recomputed), so chunks are independent.
//
// Every value is computed with the shared fold expressions (fold_step /
// fold_finish) and generation t+1 is rounded to the storage type exactly as a
// store and reload would, so results are bit-identical to two single steps.
// Cells outside the array follow the reference's getneighbor
// (src/array.jl:129-179) in every generation:
//   - Wrap: an outside cell of generation t+1 is computed from its mapped
//     neighbourhood -- Wrap is a translation, so that is the wrapped cell's
//     value, with the same fold order;
//   - Remove: outside cells read padval in every generation.
// (Reflect mirrors the neighbourhood, which reverses the fold order of
// floating-point sums, so it is not fused.)
#pragma once

#include "tile_kernel.hpp"

namespace sth {

constexpr int kHeat2TX = 64, kHeat2TY = 16, kHeat2ZC = 16;

template <typename S, int OPK, class Sh, class P>
__device__ __forceinline__ void heat2_body(const P& p, int bx, int by, int bzc,
                                           typename Staged<S>::type* __restrict__ lds) {
  using Raw = typename Elem<S>::raw;
  using Val = typename Staged<S>::type;
  using A = typename AccOf<S, OPK>::type;
  using O = typename OutOf<S, OPK>::type;
  static_assert(Sh::RX <= 1 && Sh::RY <= 1 && Sh::RZ <= 1, "radius-1 stencils only");
  constexpr int TX = kHeat2TX, TY = kHeat2TY, ZC = kHeat2ZC;
  constexpr int W0 = TX + 4, H0 = TY + 4, N0 = W0 * H0;  // generation t: tile + 2-cell rim
  constexpr int W1 = TX + 2, H1 = TY + 2, N1 = W1 * H1;  // generation t+1: tile + 1-cell rim
  constexpr int K0 = (N0 + 255) / 256, K1 = (N1 + 255) / 256, K2 = TX * TY / 256;
  Val* g0 = lds;            // 3 planes of generation t
  Val* g1 = lds + 3 * N0;   // 3 planes of generation t+1

  const int t = static_cast<int>(threadIdx.x);
  const int64_t n0 = p.src.n0, n1 = p.src.n1, n2 = p.src.n2;
  const int64_t X0 = static_cast<int64_t>(bx) * TX, Y0 = static_cast<int64_t>(by) * TY;
  const int64_t zb = static_cast<int64_t>(bzc) * ZC;
  const int64_t ze = zb + ZC < n2 ? zb + ZC : n2;
  const int bc = p.bc;
  const bool remove = bc == STH_REMOVE;
  const Val pv = stage<S>(raw_from_bits<Raw>(p.padval));
  const Raw* __restrict__ src = reinterpret_cast<const Raw*>(p.src.base);
  const int64_t s1 = p.src.s1, s2 = p.src.s2;
  auto wrapc = [](int64_t i, int64_t n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
  auto inside = [&](int64_t x, int64_t y, int64_t z) {
    return x >= 0 && x < n0 && y >= 0 && y < n1 && z >= 0 && z < n2;
  };

  // generation-t loads of one plane (K0 cells per thread), boundary-mapped
  Val pre[K0];
  auto load_plane = [&](int64_t z) {
#pragma unroll
    for (int m = 0; m < K0; ++m) {
      const int idx = t + 256 * m;
      Val v = pv;
      if (idx < N0) {
        const int ly = idx / W0, lx = idx - ly * W0;
        const int64_t x = X0 - 2 + lx, y = Y0 - 2 + ly;
        if (inside(x, y, z)) {
          v = stage<S>(src[z * s2 + y * s1 + x]);
        } else if (!remove) {
          // Wrap.  A partial tile reaches past the array's 2-cell rim; those
          // cells feed no stored result and are clamped into the rim first.
          const int64_t xc = x < -2 ? -2 : (x > n0 + 1 ? n0 + 1 : x);
          const int64_t yc = y < -2 ? -2 : (y > n1 + 1 ? n1 + 1 : y);
          const int64_t zc = z < -2 ? -2 : (z > n2 + 1 ? n2 + 1 : z);
          v = stage<S>(src[wrapc(zc, n2) * s2 + wrapc(yc, n1) * s1 + wrapc(xc, n0)]);
        }
      }
      pre[m] = v;
    }
  };

  const int nit = static_cast<int>(ze - zb) + 4;
  load_plane(zb - 2);
  for (int k = 0; k < nit; ++k) {
    const int64_t zi = zb - 2 + k;
    // A. the prefetched plane zi into ring slot k % 3; prefetch zi + 1
    Val* g0k = g0 + (k % 3) * N0;
#pragma unroll
    for (int m = 0; m < K0; ++m) {
      const int idx = t + 256 * m;
      if (idx < N0) g0k[idx] = pre[m];
    }
    if (k + 1 < nit) load_plane(zi + 1);
    __syncthreads();
    // B. generation t+1 on plane zi - 1 (tile + 1-cell rim)
    if (k >= 2) {
      const int64_t zp = zi - 1;
      const Val* pl[3] = {g0 + ((k - 2) % 3) * N0, g0 + ((k - 1) % 3) * N0, g0k};
      Val* g1k = g1 + ((k - 1) % 3) * N1;
#pragma unroll
      for (int m = 0; m < K1; ++m) {
        const int idx = t + 256 * m;
        if (idx < N1) {
          const int ly = idx / W1, lx = idx - ly * W1;  // generation t+1 cell (lx - 1, ly - 1) of the tile
          const int c0 = (ly + 1) * W0 + (lx + 1);      // its generation-t index in a plane
          Val v;
          if (remove && !inside(X0 - 1 + lx, Y0 - 1 + ly, zp)) {
            v = pv;
          } else {
            A acc = A(0);
            uint32_t cnt = 0;
#pragma unroll
            for (int q = 0; q < Sh::L; ++q)
              fold_step<S, OPK>(acc, cnt, q, pl[1 + Sh::off[q][2]][c0 + Sh::off[q][1] * W0 + Sh::off[q][0]], p);
            // rounded to the storage type as a store would
            v = stage<S>(narrow<O>(fold_finish<S, OPK>(acc, cnt, pl[1][c0], Sh::L, p)));
          }
          g1k[idx] = v;
        }
      }
    }
    __syncthreads();
    // C. generation t+2 on plane zi - 2 (the tile)
    if (k >= 4) {
      const int64_t zq = zi - 2;
      const Val* pl[3] = {g1 + ((k - 3) % 3) * N1, g1 + ((k - 2) % 3) * N1, g1 + ((k - 1) % 3) * N1};
      using ORaw = typename Elem<O>::raw;
      ORaw* dst = reinterpret_cast<ORaw*>(p.dst.base);
#pragma unroll
      for (int m = 0; m < K2; ++m) {
        const int idx = t + 256 * m;
        const int ly = idx / TX, lx = idx - ly * TX;
        const int64_t x = X0 + lx, y = Y0 + ly;
        const int c1 = (ly + 1) * W1 + (lx + 1);
        A acc = A(0);
        uint32_t cnt = 0;
#pragma unroll
        for (int q = 0; q < Sh::L; ++q)
          fold_step<S, OPK>(acc, cnt, q, pl[1 + Sh::off[q][2]][c1 + Sh::off[q][1] * W1 + Sh::off[q][0]], p);
        const auto r = fold_finish<S, OPK>(acc, cnt, pl[1][c1], Sh::L, p);
        if (x < n0 && y < n1) dst[zq * p.dst.s2 + y * p.dst.s1 + x] = narrow<O>(r);
      }
    }
    // Ring reuse: A of the next iteration overwrites the generation-t slot B
    // read here, after every thread passed the barrier that follows B; B of
    // the next iteration overwrites the generation-t+1 slot C read here, after
    // the next barrier A, which every thread reaches only once C is done.
  }
}

template <typename S, int OPK, class Sh>
__global__ __launch_bounds__(256) void heat2_kernel(MapParams p) {
  extern __shared__ __align__(16) unsigned char smem_dyn[];
  auto* lds = reinterpret_cast<typename Staged<S>::type*>(smem_dyn);
  const int nblocks = p.tiles_x * p.tiles_y * p.tiles_z;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bx = bid % p.tiles_x;
  const int byz = bid / p.tiles_x;
  const int by = byz % p.tiles_y;
  const int bz = byz / p.tiles_y;
  heat2_body<S, OPK, Sh>(p, bx, by, bz, lds);
}

template <typename S>
constexpr size_t heat2_lds() {
  return (3 * (kHeat2TX + 4) * (kHeat2TY + 4) + 3 * (kHeat2TX + 2) * (kHeat2TY + 2)) *
         sizeof(typename Staged<S>::type);
}

template <typename S, int OPK, class Sh>
void launch_heat2(const MapParams& p, hipStream_t s) {
  const unsigned nb = static_cast<unsigned>(p.tiles_x) * p.tiles_y * p.tiles_z;
  hipLaunchKernelGGL((heat2_kernel<S, OPK, Sh>), dim3(nb), dim3(256), heat2_lds<S>(), s, p);
}

}  // namespace sth
