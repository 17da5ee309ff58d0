// Batched gather: many independent mapstencil calls of mixed sizes, shapes
// and reductions in one persistent launch.
//
// The host flattens the tiles of every job into one index space (job j owns
// tiles [tile_begin_j, tile_begin_{j+1})) and launches as many workgroups as
// fit on the device at once.  Each workgroup pulls tiles from a device-side
// counter -- one atomic per tile, issued for the next tile before the
// current one is processed -- until the counter passes the last tile, so
// cheap and expensive jobs balance across the device without a host loop.
// A tile runs the streaming body of its job's (shape, op) variant, selected
// by a switch generated per element type.  Results are bit-identical to one
// sth_map call per job (same fold expressions per cell).
#pragma once

#include "stream_kernel.hpp"

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
};

struct BatchParams {
  const BatchJob* jobs;
  unsigned int* queue;  // tile counter; zero when the launch starts
  int32_t njobs;
  int32_t pad_;
  int64_t total_tiles;
};

// `Dispatch::run(job, bx, by, bz, lane, wave)` runs the streaming body of
// job.variant; it is generated per storage type and force-inlined into the
// kernel.  (An out-of-line dispatch function of this size gets long-branch
// sequences that reuse the return-address registers, so the kernel must not
// call one.)
template <typename S, class Dispatch>
__global__ __launch_bounds__(256) void batch_kernel(BatchParams bp) {
  extern __shared__ __align__(16) unsigned char smem_dyn[];
  unsigned int& slot = *reinterpret_cast<unsigned int*>(smem_dyn);  // the block's next tile
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  if (threadIdx.x == 0) slot = atomicAdd(bp.queue, 1u);
  __syncthreads();
  int64_t t = slot;
  __syncthreads();
  int j = 0;
  while (t < bp.total_tiles) {
    unsigned int next = 0;
    if (threadIdx.x == 0) next = atomicAdd(bp.queue, 1u);  // in flight during this tile
    // tiles are handed out in increasing order, so the job index only advances
    while (j + 1 < bp.njobs && bp.jobs[j + 1].tile_begin <= t) ++j;
    const BatchJob& job = bp.jobs[j];
    const int64_t local = t - job.tile_begin;
    const int bx = static_cast<int>(local % job.tiles_x);
    const int64_t byz = local / job.tiles_x;
    const int by = static_cast<int>(byz % job.tiles_y);
    const int bz = static_cast<int>(byz / job.tiles_y);
    Dispatch::run(job, bx, by, bz, lane, wave);
    if (threadIdx.x == 0) slot = next;
    __syncthreads();
    t = slot;
    __syncthreads();  // every thread has read the slot before it is rewritten
  }
}

using BatchLaunchFn = void (*)(const BatchParams&, unsigned blocks, hipStream_t);

struct BatchVariant {
  int dtype;
  int op;
  int L;
  int ndim;
  const int8_t* offs;
  int variant;
  int sv, sky;  // streaming tile: 64*sv cells x 4*sky rows
};

struct BatchKernelEntry {
  int dtype;
  BatchLaunchFn fn;
  int (*occupancy)();  // resident workgroups per CU
};

std::vector<BatchVariant>& batch_variants();
std::vector<BatchKernelEntry>& batch_kernels();

template <typename S, class Dispatch>
void launch_batch(const BatchParams& bp, unsigned blocks, hipStream_t s) {
  hipLaunchKernelGGL((batch_kernel<S, Dispatch>), dim3(blocks), dim3(256), 16, s, bp);
}

template <typename S, class Dispatch>
int batch_occupancy() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel<S, Dispatch>, 256, 16) != hipSuccess)
    n = 1;
  return n > 0 ? n : 1;
}

}  // namespace sth
