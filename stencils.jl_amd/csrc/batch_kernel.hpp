// Batched gather: many independent mapstencil calls of mixed sizes, shapes
// and reductions in one persistent launch.
//
// The host flattens the tiles of every job into one index space (job j owns
// tiles [tile_begin_j, tile_begin_{j+1})).  Each workgroup takes a contiguous
// chunk of that space -- consecutive tiles of one job usually, so neighbouring
// tiles share L2 -- and runs the same tile body as the single-array kernel,
// selected per job by a switch over the build-time (shape, op) variants.
// Results are bit-identical to one sth_map call per job.
#pragma once

#include "tile_kernel.hpp"

namespace sth {

struct BatchJob {
  DevArr src;
  DevArr dst;
  int32_t bc;
  int32_t read_halo;
  uint64_t padval;
  OpParams op;
  int32_t variant;  // case of the generated dispatch switch
  int32_t tiles_x, tiles_y, tiles_z;
  int64_t tile_begin;  // first tile of this job in the flattened index space
  // Only read by runtime-offset shapes, which batches do not contain; kept so
  // the shared tile body compiles against this type.
  int32_t rx, ry, rz, L;
};

struct BatchParams {
  const BatchJob* jobs;
  int32_t njobs;
  int32_t pad_;
  int64_t total_tiles;
};

// `Dispatch::run(job, bx, by, bz, tile)` runs the tile body of job.variant;
// it is generated per storage type and force-inlined into the kernel.  (An
// out-of-line dispatch function of this size gets long-branch sequences that
// reuse the return-address registers, so the kernel must not call one.)
template <typename S, class Dispatch>
__global__ __launch_bounds__(256) void batch_kernel(BatchParams bp) {
  extern __shared__ __align__(16) unsigned char smem_dyn[];
  auto* tile = reinterpret_cast<typename Staged<S>::type*>(smem_dyn);
  const int G = static_cast<int>(gridDim.x);
  const int b = xcd_block_id(static_cast<int>(blockIdx.x), G, 1);
  const int64_t per = (bp.total_tiles + G - 1) / G;
  const int64_t t0 = static_cast<int64_t>(b) * per;
  const int64_t t1 = t0 + per < bp.total_tiles ? t0 + per : bp.total_tiles;
  if (t0 >= t1) return;
  // last job whose first tile is <= t0
  int lo = 0, hi = bp.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bp.jobs[mid].tile_begin <= t0) lo = mid;
    else hi = mid - 1;
  }
  int j = lo;
  for (int64_t t = t0; t < t1; ++t) {
    while (j + 1 < bp.njobs && bp.jobs[j + 1].tile_begin <= t) ++j;
    const BatchJob& job = bp.jobs[j];
    const int64_t local = t - job.tile_begin;
    const int bx = static_cast<int>(local % job.tiles_x);
    const int64_t byz = local / job.tiles_x;
    const int by = static_cast<int>(byz % job.tiles_y);
    const int bz = static_cast<int>(byz / job.tiles_y);
    Dispatch::run(job, bx, by, bz, tile);
    __syncthreads();  // the next tile restages LDS
  }
}

using BatchLaunchFn = void (*)(const BatchParams&, unsigned blocks, size_t lds, hipStream_t);

struct BatchVariant {
  int dtype;
  int op;
  int L;
  int ndim;
  const int8_t* offs;
  int variant;
  int tx, ty;  // tile of this variant's body
};

struct BatchKernelEntry {
  int dtype;
  BatchLaunchFn fn;
  size_t lds;                     // dynamic LDS of the largest variant
  int (*occupancy)(size_t lds);   // resident workgroups per CU at that LDS size
};

std::vector<BatchVariant>& batch_variants();
std::vector<BatchKernelEntry>& batch_kernels();

template <typename S, class Dispatch>
void launch_batch(const BatchParams& bp, unsigned blocks, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((batch_kernel<S, Dispatch>), dim3(blocks), dim3(256), lds, s, bp);
}

template <typename S, class Dispatch>
int batch_occupancy(size_t lds) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel<S, Dispatch>, 256, lds) != hipSuccess)
    n = 1;
  return n > 0 ? n : 1;
}

}  // namespace sth
