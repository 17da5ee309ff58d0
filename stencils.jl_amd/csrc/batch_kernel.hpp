// Batched gather: many independent mapstencil calls of mixed sizes, shapes
// and reductions in one launch per (element type, stencil shape, op).
//
// The host flattens the tiles of every job into one index space (job j owns
// tiles [tile_begin_j, tile_begin_{j+1})) and launches one workgroup per
// strip of `tiles_per_block` consecutive tiles; the hardware dispatcher hands
// strips to CUs as they free up, so cheap and expensive jobs balance without
// atomics or barriers (a single device-wide atomic tile counter measured much
// slower: same-address atomics serialise in L2).  A workgroup finds its first
// job by binary search over the job table (uniform, scalar loads) and walks
// forward.  Each tile runs the streaming body of the kernel's (shape, op)
// variant: one small kernel per variant keeps the code in the instruction
// cache (a kernel switching over a shape's ops measured up to 40 % slower).
// Results are bit-identical to one sth_map call per job (same fold
// expressions per cell).
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
  int32_t lpr_shift;   // streaming body: log2 of the lanes per row
  int32_t pad_;
  int64_t tile_begin;  // first tile of this job in the flattened index space
};

struct BatchParams {
  const BatchJob* jobs;
  int32_t njobs;
  int32_t tiles_per_block;
  int64_t total_tiles;
};

// `Dispatch::run(job, bx, by, bz, lane, wave)` runs the streaming body of the
// kernel's variant; it is generated per (storage type, shape, op) and
// force-inlined into the kernel, which makes no calls.
template <typename S, class Dispatch>
__global__ __launch_bounds__(256) void batch_kernel(BatchParams bp, const BatchJob* __restrict__ jobs) {
  // `jobs` is read-only and aliases nothing the kernel writes, so its fields
  // stay in scalar registers instead of being reloaded after every store.
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  const int G = static_cast<int>(gridDim.x);
  const int64_t b = xcd_block_id(static_cast<int>(blockIdx.x), G, 1);
  int64_t t = b * bp.tiles_per_block;
  const int64_t tend = t + bp.tiles_per_block < bp.total_tiles ? t + bp.tiles_per_block : bp.total_tiles;
  if (t >= tend) return;
  // last job whose first tile is <= t
  int lo = 0, hi = bp.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].tile_begin <= t) lo = mid;
    else hi = mid - 1;
  }
  int j = lo;
  for (; t < tend; ++t) {
    while (j + 1 < bp.njobs && jobs[j + 1].tile_begin <= t) ++j;
    const BatchJob job = jobs[j];  // a register copy for the tile
    // a job's own tile count fits in 32 bits (checked on the host)
    const uint32_t local = static_cast<uint32_t>(t - job.tile_begin);
    const uint32_t tx = static_cast<uint32_t>(job.tiles_x), ty = static_cast<uint32_t>(job.tiles_y);
    const int bx = static_cast<int>(local % tx);
    const uint32_t byz = local / tx;
    const int by = static_cast<int>(byz % ty);
    const int bz = static_cast<int>(byz / ty);
    Dispatch::run(job, bx, by, bz, lane, wave);
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
  int klass;    // the batched kernel that runs this variant (one per variant)
};

struct BatchKernelEntry {
  int dtype;
  int klass;
  BatchLaunchFn fn;
  int (*occupancy)();  // resident workgroups per CU
};

std::vector<BatchVariant>& batch_variants();
std::vector<BatchKernelEntry>& batch_kernels();

template <typename S, class Dispatch>
void launch_batch(const BatchParams& bp, unsigned blocks, hipStream_t s) {
  hipLaunchKernelGGL((batch_kernel<S, Dispatch>), dim3(blocks), dim3(256), 0, s, bp, bp.jobs);
}

template <typename S, class Dispatch>
int batch_occupancy() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel<S, Dispatch>, 256, 0) != hipSuccess)
    n = 1;
  return n > 0 ? n : 1;
}

}  // namespace sth
