// Streaming gather kernel for stencils whose offsets are fixed at build time,
// on 2-D and 3-D arrays: no LDS, no barriers.
//
// Work split (one workgroup = 4 wave64s):
//   lane  -> V consecutive cells along dim 1 (one 16-byte vector for the
//            natural V of the element type) and KY consecutive rows along
//            dim 2,
//   wave  -> LPR = 2^lpr_shift lanes along a row x 64 / LPR lane rows (the
//            host picks LPR from the array width so narrow arrays keep every
//            lane busy),
//   block -> LPR V cells x 4 (64 / LPR) KY rows of one dim-3 plane.
// Each lane loads every source row its KY output rows need -- the vector
// part plus RX cells on either side -- into registers, folds the stencil for
// its KY x V outputs in offset order (the reference's left fold,
// src/gatherstencil.jl:105-109) and stores whole vectors.  Rows no offset
// reads are never loaded: the value box is indexed with compile-time
// constants, so unused entries and their loads are dead code.  Rows shared by
// neighbouring waves come from L2, so HBM sees each cell about once.
//
// Boundaries (src/array.jl:129-179) are resolved per row for the whole wave
// (a uniform decision) and per cell only in lanes whose x range leaves the
// array; waves entirely inside the array read straight rows.
#pragma once

#include "tile_kernel.hpp"

namespace sth {

template <typename Raw, int V>
struct RawVec {
  Raw e[V];
};

template <typename Raw, int V>
__device__ __forceinline__ RawVec<Raw, V> load_vec(const Raw* p) {
  RawVec<Raw, V> r;
  __builtin_memcpy(&r, p, sizeof(r));
  return r;
}

template <typename Raw, int V>
__device__ __forceinline__ void store_vec(Raw* p, const RawVec<Raw, V>& v) {
  __builtin_memcpy(p, &v, sizeof(v));
}

// The same vectors at their natural alignment (V * sizeof(Raw) bytes): one
// load or store instruction even for 16-bit elements.
template <typename Raw, int V>
struct alignas(V * sizeof(Raw)) AlignedVec {
  Raw e[V];
};

template <typename Raw, int V>
__device__ __forceinline__ RawVec<Raw, V> load_avec(const Raw* p) {
  const AlignedVec<Raw, V> a = *reinterpret_cast<const AlignedVec<Raw, V>*>(p);
  RawVec<Raw, V> r;
#pragma unroll
  for (int i = 0; i < V; ++i) r.e[i] = a.e[i];
  return r;
}

template <typename Raw, int V>
__device__ __forceinline__ void store_avec(Raw* p, const RawVec<Raw, V>& v) {
  AlignedVec<Raw, V> a;
#pragma unroll
  for (int i = 0; i < V; ++i) a.e[i] = v.e[i];
  *reinterpret_cast<AlignedVec<Raw, V>*>(p) = a;
}

// True when `base` and every row and plane start are multiples of `bytes`.
__device__ __forceinline__ bool rows_aligned(const void* base, int64_t s1, int64_t s2, int64_t esize,
                                             int64_t bytes) {
  return ((reinterpret_cast<uint64_t>(base) | static_cast<uint64_t>(s1 * esize) |
           static_cast<uint64_t>(s2 * esize)) % static_cast<uint64_t>(bytes)) == 0;
}

__device__ __forceinline__ int64_t clamp64(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Resolve one source row coordinate.  Returns the memory coordinate, or sets
// `pad` when the whole row reads padval (Remove outside the array).
__device__ __forceinline__ int64_t resolve_row(int64_t y, int64_t n, int64_t h, int r, bool halo, int bc,
                                               bool& pad) {
  if (halo) return clamp64(y, -h, n + h - 1);
  // rows beyond the reach box only feed results that are never stored
  const int64_t yc = clamp64(y, -r, n - 1 + r);
  if (yc >= 0 && yc < n) return yc;
  if (bc == STH_REMOVE) {
    pad = true;
    return 0;
  }
  return map_index(yc, n, bc);
}

// What happens to a lane's V results of row j (row y of the array) before
// they are stored: nothing (a plain gatherstencil!), or the combination with
// further terms of a multi-array gather (runtime-compiled kernels define
// their own).  An epilogue with kStore false keeps the results itself and
// nothing is stored.
struct NoEpilogue {
  static constexpr bool kStore = true;
  template <typename R, int V>
  __device__ __forceinline__ void apply(RawVec<R, V>&, int64_t, int64_t, int64_t, int64_t, int) const {}
};

template <typename S, int OPK, class Sh, int V, int KY, class P, class Epi = NoEpilogue>
__device__ __forceinline__ void stream_body(const P& p, int bx, int by, int bz, int lane, int wave,
                                            const Epi& epi = Epi{}) {
  using Raw = typename Elem<S>::raw;
  using Val = typename Staged<S>::type;
  using A = typename AccOf<S, OPK>::type;
  using O = typename OutOf<S, OPK>::type;
  using ORaw = typename Elem<O>::raw;
  constexpr int RX = Sh::RX, RY = Sh::RY, RZ = Sh::RZ;
  constexpr int NX = V + 2 * RX, NY = KY + 2 * RY, NZ = 2 * RZ + 1;

  const int64_t n0 = p.src.n0, n1 = p.src.n1, n2 = p.src.n2;
  // A wave covers 2^lsh lanes of a row and 64 >> lsh rows side by side
  // (narrow arrays keep every lane busy); each lane owns KY consecutive rows.
  const int lsh = p.lpr_shift;
  const int LPR = 1 << lsh, RPW = 64 >> lsh;
  const int64_t wy0 = (static_cast<int64_t>(by) * 4 + wave) * KY * RPW;  // first row of the wave
  const int64_t z0 = bz;
  if (wy0 >= n1) return;  // uniform: the wave has no rows (no barrier follows)
  const int64_t y0 = wy0 + static_cast<int64_t>(lane >> lsh) * KY;        // first row of the lane
  const int64_t wx0 = static_cast<int64_t>(bx) * LPR * V;                  // first cell of the wave
  const int64_t x0 = wx0 + static_cast<int64_t>(lane & (LPR - 1)) * V;
  const bool halo = p.read_halo != 0;
  const int64_t h0 = halo ? p.src.h0 : 0, h1 = halo ? p.src.h1 : 0, h2 = halo ? p.src.h2 : 0;
  const int bc = p.bc;
  const Raw* __restrict__ sbase = reinterpret_cast<const Raw*>(p.src.base);
  const int64_t s1 = p.src.s1, s2 = p.src.s2;

  Val v[NZ][NY][NX];

  const bool rows_direct = wy0 - RY >= -h1 && wy0 + RPW * KY - 1 + RY < n1 + h1 && z0 - RZ >= -h2 &&
                           z0 + RZ < n2 + h2;
  const bool wave_x_direct = wx0 - RX >= -h0 && wx0 + LPR * V + RX <= n0 + h0;
  // 16-bit elements on aligned rows whose neighbouring vectors lie inside the
  // row: the RX cells on either side come from the neighbouring aligned
  // vectors, three vector loads per row instead of one vector and 2 RX
  // single 16-bit loads.  (For 4- and 8-byte elements the single loads are
  // cheaper than the wider vectors: measured 12% slower on a 7x7 f32 kernel.)
  const bool vec3 = sizeof(Raw) == 2 && RX <= V && rows_direct && wx0 - V >= -h0 &&
                    wx0 + (LPR + 1) * V <= n0 + h0 && rows_aligned(sbase, s1, s2, sizeof(Raw), V * sizeof(Raw));
  if (vec3) {
#pragma unroll
    for (int iz = 0; iz < NZ; ++iz) {
#pragma unroll
      for (int iy = 0; iy < NY; ++iy) {
        const Raw* row = sbase + (z0 - RZ + iz) * s2 + (y0 - RY + iy) * s1 + x0;
        const RawVec<Raw, V> m = load_avec<Raw, V>(row);
#pragma unroll
        for (int i = 0; i < V; ++i) v[iz][iy][RX + i] = stage<S>(m.e[i]);
        if constexpr (RX > 0) {
          const RawVec<Raw, V> l = load_avec<Raw, V>(row - V);
          const RawVec<Raw, V> r = load_avec<Raw, V>(row + V);
#pragma unroll
          for (int i = 0; i < RX; ++i) {
            v[iz][iy][i] = stage<S>(l.e[V - RX + i]);
            v[iz][iy][RX + V + i] = stage<S>(r.e[i]);
          }
        }
      }
    }
  } else if (rows_direct && wave_x_direct) {
    // ---- interior wave: straight rows, one vector + 2 RX cells per row
#pragma unroll
    for (int iz = 0; iz < NZ; ++iz) {
#pragma unroll
      for (int iy = 0; iy < NY; ++iy) {
        const Raw* row = sbase + (z0 - RZ + iz) * s2 + (y0 - RY + iy) * s1 + x0;
        const RawVec<Raw, V> m = load_vec<Raw, V>(row);
#pragma unroll
        for (int i = 0; i < V; ++i) v[iz][iy][RX + i] = stage<S>(m.e[i]);
#pragma unroll
        for (int i = 0; i < RX; ++i) {
          v[iz][iy][i] = stage<S>(row[i - RX]);
          v[iz][iy][RX + V + i] = stage<S>(row[V + i]);
        }
      }
    }
  } else {
    // ---- edge wave: rows resolved per boundary for the whole wave; every
    // lane maps its NX column positions once (lanes inside the array keep
    // them), so loads are unconditional and Remove is a select.
    const Val spv = stage<S>(raw_from_bits<Raw>(p.padval));
    int32_t xm[NX];
    bool xpad[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      const int64_t x = x0 - RX + i;
      if (halo) {
        xm[i] = static_cast<int32_t>(clamp64(x, -h0, n0 + h0 - 1));
        xpad[i] = false;
      } else {
        // lanes past the end clamp into the reach box; their results are not stored
        const int64_t xc = clamp64(x, -RX, n0 - 1 + RX);
        const bool inside = xc >= 0 && xc < n0;
        xpad[i] = !inside && bc == STH_REMOVE;
        xm[i] = static_cast<int32_t>(inside ? xc : (bc == STH_REMOVE ? 0 : map_index(xc, n0, bc)));
      }
    }
    // the vector part can be one load if no lane straddles the end of a row
    // (lanes wholly past the end load from the row start instead)
    const bool vec_lane = x0 >= n0 || x0 + V <= n0 + h0;
    const bool vec_ok = __all(vec_lane);
    const int64_t xv = x0 < n0 ? x0 : 0;
#pragma unroll
    for (int iz = 0; iz < NZ; ++iz) {
#pragma unroll
      for (int iy = 0; iy < NY; ++iy) {
        bool pad = false;
        const int64_t zm = resolve_row(z0 - RZ + iz, n2, h2, RZ, halo, bc, pad);
        const int64_t ym = resolve_row(y0 - RY + iy, n1, h1, RY, halo, bc, pad);
        const Raw* row = sbase + zm * s2 + ym * s1;
        Raw r[NX];
        if (vec_ok) {
          const RawVec<Raw, V> m = load_vec<Raw, V>(row + xv);
#pragma unroll
          for (int i = 0; i < V; ++i) r[RX + i] = m.e[i];
#pragma unroll
          for (int i = 0; i < RX; ++i) {
            r[i] = row[xm[i]];
            r[RX + V + i] = row[xm[RX + V + i]];
          }
        } else {
#pragma unroll
          for (int i = 0; i < NX; ++i) r[i] = row[xm[i]];
        }
#pragma unroll
        for (int i = 0; i < NX; ++i) v[iz][iy][i] = (pad || xpad[i]) ? spv : stage<S>(r[i]);
      }
    }
  }

  // ---- fold and store
  ORaw* __restrict__ dbase = reinterpret_cast<ORaw*>(p.dst.base) + z0 * p.dst.s2;
  const bool dst_aligned = rows_aligned(p.dst.base, p.dst.s1, p.dst.s2, sizeof(ORaw), V * sizeof(ORaw));
#pragma unroll
  for (int j = 0; j < KY; ++j) {
    RawVec<ORaw, V> out;
#pragma unroll
    for (int i = 0; i < V; ++i) {
      A acc = A(0);
      uint32_t cnt = 0;
#pragma unroll
      for (int k = 0; k < Sh::L; ++k)
        fold_step<S, OPK>(acc, cnt, k, v[RZ + Sh::off[k][2]][RY + j + Sh::off[k][1]][RX + i + Sh::off[k][0]], p);
      out.e[i] = narrow<O>(fold_finish<S, OPK>(acc, cnt, v[RZ][RY + j][RX + i], Sh::L, p));
    }
    const int64_t y = y0 + j;
    if (y < n1) {
      epi.apply(out, x0, y, z0, n0, j);
      if constexpr (Epi::kStore) {
        ORaw* drow = dbase + y * p.dst.s1;
        if (x0 + V <= n0) {
          if (dst_aligned)
            store_avec<ORaw, V>(drow + x0, out);
          else
            store_vec<ORaw, V>(drow + x0, out);
        } else {
#pragma unroll
          for (int i = 0; i < V; ++i)
            if (x0 + i < n0) drow[x0 + i] = out.e[i];
        }
      }
    }
  }
}

template <typename S, int OPK, class Sh, int V, int KY>
__global__ __launch_bounds__(256) void stream_kernel(MapParams p) {
  const int nblocks = p.tiles_x * p.tiles_y * p.tiles_z;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bx = bid % p.tiles_x;
  const int byz = bid / p.tiles_x;
  const int by = byz % p.tiles_y;
  const int bz = byz / p.tiles_y;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  stream_body<S, OPK, Sh, V, KY>(p, bx, by, bz, static_cast<int>(threadIdx.x) & 63, wave);
}

// The streaming kernel with an epilogue (Epi::Params is its second
// argument); instantiated only by runtime-compiled specialisations.
template <typename S, int OPK, class Sh, int V, int KY, class Epi>
__global__ __launch_bounds__(256) void stream_kernel_epi(MapParams p, typename Epi::Params e) {
  const int nblocks = p.tiles_x * p.tiles_y * p.tiles_z;
  const int bid = xcd_block_id(static_cast<int>(blockIdx.x), nblocks, p.xcd_remap);
  const int bx = bid % p.tiles_x;
  const int byz = bid / p.tiles_x;
  const int by = byz % p.tiles_y;
  const int bz = byz / p.tiles_y;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  stream_body<S, OPK, Sh, V, KY>(p, bx, by, bz, static_cast<int>(threadIdx.x) & 63, wave, Epi{e});
}

// One stencil argument of a multi-array call whose arguments are fused into
// one streaming kernel: what stream_body reads of MapParams (the offsets are
// compiled in; weighted folds are not fused, so no weight table).
struct StreamTerm {
  DevArr src;
  DevArr dst;
  int32_t bc;
  int32_t read_halo;
  int32_t lpr_shift;
  int32_t pad_;
  uint64_t padval;
  OpParams op;
};

// Parameters of such a kernel (runtime-compiled: the fold and centre terms
// are code): the stencil arguments in call order, the centre arguments in
// call order (interior element (0,0,0) and strides), each argument's
// coefficient by its position in the call, and the tile grid.
struct StreamMultiParams {
  StreamTerm term[STH_MAX_TERMS];
  const void* cbase[STH_MAX_TERMS];
  int64_t cs1[STH_MAX_TERMS], cs2[STH_MAX_TERMS];
  double coef[STH_MAX_TERMS];
  int32_t tiles_x, tiles_y, tiles_z, xcd_remap;
};

// Natural vector width: 16 bytes per lane.
template <typename S>
constexpr int stream_v() {
  return 16 / static_cast<int>(sizeof(typename Elem<S>::raw));
}

template <typename S, int OPK, class Sh, int V, int KY>
void launch_stream(const MapParams& p, hipStream_t stream) {
  const unsigned nblocks = static_cast<unsigned>(p.tiles_x) * p.tiles_y * p.tiles_z;
  hipLaunchKernelGGL((stream_kernel<S, OPK, Sh, V, KY>), dim3(nblocks), dim3(256), 0, stream, p);
}

}  // namespace sth
