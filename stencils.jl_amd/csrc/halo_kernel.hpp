// update_boundary! for Halo padding (src/array.jl:195-242): every halo cell
// of a padded parent receives `padval` (Remove) or the interior value at the
// boundary-mapped index (Wrap, Reflect).  Values only ever come from the
// interior, so all faces are written by one launch with no ordering; cells
// shared by two faces (corners) receive the same value from both.
#pragma once

#include "common.hpp"

namespace sth {

struct HaloParams {
  DevArr a;        // logical array; base = logical (0, 0, 0)
  int32_t ndim;
  int32_t R;       // halo width
  int32_t bc;
  int32_t pad_;
  uint64_t padval;
  int64_t cells[3];  // number of cells in the two slabs of dim d
  int64_t blocks[3]; // blocks used by dim d (prefix over gridDim.x)
};

template <typename Raw>
__global__ __launch_bounds__(256) void halo_kernel(HaloParams p) {
  // Which face slab does this block serve?
  int64_t b = blockIdx.x;
  int d = 0;
  while (d < 2 && b >= p.blocks[d]) {
    b -= p.blocks[d];
    ++d;
  }
  const int64_t idx = b * 256 + threadIdx.x;
  if (d >= p.ndim || idx >= p.cells[d]) return;
  const int64_t n[3] = {p.a.n0, p.a.n1, p.a.n2};
  const int R = p.R;
  int64_t t = idx;
  int64_t c[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int re = e < p.ndim ? R : 0;
    const int64_t ext = e == d ? 2 * R : n[e] + 2 * re;
    const int64_t q = t % ext;
    t /= ext;
    if (e == d) c[e] = q < R ? q - R : n[e] + (q - R);
    else c[e] = q - re;
  }
  Raw* base = reinterpret_cast<Raw*>(p.a.base);
  Raw v;
  if (p.bc == STH_REMOVE) {
    v = raw_from_bits<Raw>(p.padval);
  } else {
    const int64_t m0 = map_index(c[0], n[0], p.bc);
    const int64_t m1 = map_index(c[1], n[1], p.bc);
    const int64_t m2 = map_index(c[2], n[2], p.bc);
    v = base[m2 * p.a.s2 + m1 * p.a.s1 + m0];
  }
  base[c[2] * p.a.s2 + c[1] * p.a.s1 + c[0]] = v;
}

template <typename Raw>
void launch_halo(const HaloParams& p, hipStream_t s) {
  const int64_t nb = p.blocks[0] + p.blocks[1] + p.blocks[2];
  if (nb == 0) return;
  hipLaunchKernelGGL(halo_kernel<Raw>, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, p);
}

}  // namespace sth
