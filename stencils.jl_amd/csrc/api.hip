// C ABI of libstencils_hip.so (include/stencils_hip.h): validation that
// mirrors the reference's ArgumentErrors, kernel selection over the
// catalogue, Halo refresh, and the ping-pong iteration engine with hipGraph
// replay for SwitchingStencilArray (src/gatherstencil.jl:77-103).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "batch_kernel.hpp"
#include "halo_kernel.hpp"
#include "jit.hpp"
#include "zmarch2_kernel.hpp"
#include "zmarch1_kernel.hpp"
#include "ymarch2_kernel.hpp"
#include "life2_kernel.hpp"
#include "lifeb_kernel.hpp"
#include "life_kernel.hpp"
#include "multi_kernel.hpp"
#include "registry.hpp"
#include "scatter_kernel.hpp"
#include "stencils_hip.h"

namespace sth {
constexpr int kStaticBatchJobs = STH_STATIC_BATCH_MAX_JOBS;
}  // namespace sth

// Device memory of sth_map_batch_static: the job tables of its batched
// launches (819,200 bytes; every device the library runs on holds one copy
// for the life of the process).  Host code never uses this name as an
// address: sth_ctx_create looks up each device's copy with
// hipGetSymbolAddress and keeps that pointer in the context.
static __device__ __align__(256) sth::BatchJob scratch[sth::kStaticBatchJobs];

namespace sth {

// Jobs per fill_jobs_kernel launch: they travel in the kernel arguments
// (3,200 bytes of the 4 KiB argument space).
constexpr int kFillJobs = 16;
struct JobChunk {
  BatchJob j[kFillJobs];
};

// Writes `n` jobs of a chunk into a device job table with vector stores.
__global__ __launch_bounds__(256) void fill_jobs_kernel(BatchJob* __restrict__ dst, JobChunk c, int n) {
  static_assert(sizeof(BatchJob) % 4 == 0, "jobs copy as 32-bit words");
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&c);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  const int words = n * static_cast<int>(sizeof(BatchJob) / 4);
  for (int i = static_cast<int>(threadIdx.x); i < words; i += 256) d[i] = src[i];
}

void launch_fill_jobs(BatchJob* dst, const JobChunk& c, int n, hipStream_t s) {
  hipLaunchKernelGGL(fill_jobs_kernel, dim3(1), dim3(256), 0, s, dst, c, n);
}

std::vector<KernelEntry>& registry() {
  static std::vector<KernelEntry> r;
  return r;
}
std::vector<MultiKernelEntry>& multi_registry() {
  static std::vector<MultiKernelEntry> r;
  return r;
}
std::vector<Fused2Entry>& fused2_registry() {
  static std::vector<Fused2Entry> r;
  return r;
}
std::vector<BatchVariant>& batch_variants() {
  static std::vector<BatchVariant> r;
  return r;
}
std::vector<BatchKernelEntry>& batch_kernels() {
  static std::vector<BatchKernelEntry> r;
  return r;
}

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define STH_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) return fail(STH_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

int dtype_size(int dt) {
  switch (dt) {
    case STH_BOOL: case STH_U8: case STH_I8: return 1;
    case STH_U16: case STH_I16: case STH_F16: case STH_BF16: return 2;
    case STH_U32: case STH_I32: case STH_F32: return 4;
    case STH_U64: case STH_I64: case STH_F64: return 8;
    default: return 0;
  }
}
bool is_float(int dt) { return dt == STH_F16 || dt == STH_BF16 || dt == STH_F32 || dt == STH_F64; }
bool is_signed(int dt) { return dt == STH_I8 || dt == STH_I16 || dt == STH_I32 || dt == STH_I64; }

// Result type of `op` over `dt`, as the reference's _return_type gives it
// (src/gatherstencil.jl:41-56); -1 if the GPU path does not cover it.
int out_dtype(int dt, const sth_op& op, int wdt) {
  const bool storage = dt == STH_BF16 || dt == STH_F16;
  if (storage && !(op.flags & STH_OPF_ACC_F32)) return -1;
  switch (op.kind) {
    case STH_OP_SUM: return dt == STH_BOOL ? STH_I64 : dt;
    case STH_OP_MEAN: return is_float(dt) ? dt : STH_F64;
    case STH_OP_MIN: case STH_OP_MAX: return dt;
    case STH_OP_KERNELPRODUCT:
      if (dt == STH_BOOL) return -1;
      if (storage) return wdt == STH_F32 ? dt : -1;
      return wdt == dt ? dt : -1;
    case STH_OP_LIFE: return is_float(dt) ? -1 : dt;
    case STH_OP_LAPLACIAN: return is_float(dt) ? dt : -1;
    default: return -1;
  }
}

struct Plan {
  MapParams p;
  LaunchFn fn;
  LaunchFn fn_steady;  // valid when the source was written by the previous step
  hipFunction_t jit;   // a runtime-compiled specialisation (jit.hpp) instead of fn
  const char* kname;
  int32_t halo;
  int32_t bvariant;  // case of the batched kernel for this call, -1 if none
  int32_t bsv, bsky; // its streaming tile (64*bsv cells x 4*bsky rows)
  int32_t bklass;    // which batched kernel (one per (shape, op) variant)
  int32_t dtype;     // sth_dtype of the source
  LaunchFn fn_fused;      // several steps per launch (steady steps), or null
  hipFunction_t jit_fused; // a runtime-compiled two-step kernel instead of fn_fused
  int32_t jit_fused_3d;    // its grid: tiles_x * tiles_y (2-D) or * tiles_z (3-D)
  int32_t fused_gens;      // steps fn_fused advances per launch
  int32_t fused_tiles_x, fused_tiles_y, fused_tiles_z;  // its grid
  HaloParams hp;
  void (*halo_fn)(const HaloParams&, hipStream_t);
};

// One launchable kernel choice for a plan; make_plan lists every variant that
// can run the call (the first is the default) and the autotuner picks one.
struct Variant {
  LaunchFn fn = nullptr;
  LaunchFn fn_steady = nullptr;
  hipFunction_t jit = nullptr;
  const char* name = "";
  int32_t tiles_x = 0, tiles_y = 0, tiles_z = 0;
  int32_t lpr_shift = 6;  // streaming kernels: log2 of the lanes per row
};

// Lanes per row of the streaming kernels: the fewest (a power of two, 8..64)
// that cover a row of n0 cells at V cells per lane.
int lpr_shift_for(int64_t n0, int v) {
  const int64_t lanes = (n0 + v - 1) / v;
  int sh = 3;
  while (sh < 6 && (int64_t(1) << sh) < lanes) ++sh;
  return sh;
}

void apply_variant(Plan& pl, const Variant& v) {
  pl.fn = v.fn;
  pl.fn_steady = v.fn_steady;
  pl.jit = v.jit;
  pl.kname = v.name;
  pl.p.tiles_x = v.tiles_x;
  pl.p.tiles_y = v.tiles_y;
  pl.p.tiles_z = v.tiles_z;
  pl.p.lpr_shift = v.lpr_shift;
  if (pl.p.src.n0 * pl.p.src.n1 * pl.p.src.n2 == 0) {  // nothing to launch
    pl.fn = pl.fn_steady = nullptr;
    pl.jit = nullptr;
  }
}

DevArr make_devarr(const sth_array& a, bool with_halo) {
  DevArr d{};
  const int es = dtype_size(a.dtype);
  const int h = with_halo ? a.halo : 0;
  const int64_t h0 = h, h1 = a.ndim > 1 ? h : 0, h2 = a.ndim > 2 ? h : 0;
  const int64_t s1 = a.ndim > 1 ? a.stride[1] : a.size[0];
  const int64_t s2 = a.ndim > 2 ? a.stride[2] : s1 * (a.ndim > 1 ? a.size[1] + 2 * h1 : 1);
  d.base = static_cast<char*>(a.data) + (h0 + h1 * s1 + h2 * s2) * es;
  d.n0 = a.size[0];
  d.n1 = a.ndim > 1 ? a.size[1] : 1;
  d.n2 = a.ndim > 2 ? a.size[2] : 1;
  d.s1 = s1;
  d.s2 = s2;
  d.h0 = static_cast<int32_t>(h0);
  d.h1 = static_cast<int32_t>(h1);
  d.h2 = static_cast<int32_t>(h2);
  return d;
}

int check_array(const sth_array* a, const char* what) {
  if (!a) return fail(STH_EINVAL, "%s descriptor is NULL", what);
  if (a->ndim < 1 || a->ndim > 3) return fail(STH_EINVAL, "%s: ndim %d not in 1..3", what, a->ndim);
  if (dtype_size(a->dtype) == 0) return fail(STH_EINVAL, "%s: unknown dtype %d", what, a->dtype);
  if (!a->data) return fail(STH_EINVAL, "%s: data is NULL", what);
  if (a->halo < 0) return fail(STH_EINVAL, "%s: negative halo", what);
  for (int d = 0; d < a->ndim; ++d)
    if (a->size[d] < 0) return fail(STH_EINVAL, "%s: negative size", what);
  if (a->stride[0] != 1)
    return fail(STH_ENOTSUP, "%s: dim 1 must be contiguous (stride 1), got %lld", what,
                static_cast<long long>(a->stride[0]));
  return STH_OK;
}

// Byte range [lo, hi) a descriptor covers (parent, halo included).
void byte_range(const sth_array& a, uintptr_t* lo, uintptr_t* hi) {
  const int es = dtype_size(a.dtype);
  int64_t last = 0;
  for (int d = 0; d < a.ndim; ++d) last += (a.size[d] + 2 * a.halo - 1) * a.stride[d];
  *lo = reinterpret_cast<uintptr_t>(a.data);
  *hi = *lo + static_cast<uintptr_t>((last + 1) * es);
}

bool offsets_equal(const int8_t* tbl, int L, const std::vector<int>& offs3) {
  for (int k = 0; k < L; ++k)
    for (int d = 0; d < 3; ++d)
      if (tbl[3 * k + d] != offs3[3 * k + d]) return false;
  return true;
}

// Bytes from an array descriptor's base to the end of its last cell.
int64_t zm2_extent(const DevArr& a, int64_t esize) {
  return ((a.n2 - 1) * a.s2 + (a.n1 - 1) * a.s1 + a.n0) * esize;
}

int make_plan(const sth_array* src, const sth_array* dst, const sth_stencil* st, int32_t bc,
              const void* padval, int32_t padding, const sth_op* op, uint32_t flags, Plan* out,
              std::vector<Variant>* cands_out = nullptr, std::string* group_out = nullptr) {
  int rc;
  if ((rc = check_array(src, "source"))) return rc;
  if ((rc = check_array(dst, "dest"))) return rc;
  if (!st || !op) return fail(STH_EINVAL, "stencil or op descriptor is NULL");
  if (src->ndim != dst->ndim) return fail(STH_EINVAL, "Source array sizes must match (ndims differ)");
  for (int d = 0; d < src->ndim; ++d)
    if (src->size[d] != dst->size[d])
      return fail(STH_EINVAL, "Source array sizes must match. Found: dim %d %lld and %lld", d + 1,
                  static_cast<long long>(src->size[d]), static_cast<long long>(dst->size[d]));
  if (padding < STH_CONDITIONAL || padding > STH_HALO_IN) return fail(STH_EINVAL, "unknown padding %d", padding);
  if (bc < STH_REMOVE || bc > STH_USE) return fail(STH_EINVAL, "unknown boundary %d", bc);
  const bool halo_pad = padding != STH_CONDITIONAL;
  if (!halo_pad && bc == STH_USE)
    return fail(STH_EINVAL, "Use() boundary needs Halo padding (no getneighbor method for Conditional)");
  if (!halo_pad && src->halo != 0) return fail(STH_EINVAL, "Conditional padding with a halo in the source descriptor");
  if (bc == STH_REMOVE && !padval) return fail(STH_EINVAL, "Remove boundary without a padval");
  if (st->length < 1 || !st->offsets) return fail(STH_EINVAL, "stencil has no offsets");
  if (st->ndim < 1 || st->ndim > src->ndim)
    return fail(STH_EINVAL, "stencil has %d dims, array %d", st->ndim, src->ndim);

  // The reference checks R < every parent axis at construction (src/array.jl:451-453).
  for (int d = 0; d < src->ndim; ++d) {
    const int64_t ps = src->size[d] + 2 * static_cast<int64_t>(src->halo);
    if (!(std::abs(st->radius) < ps))
      return fail(STH_EINVAL, "stencil radius is larger than array axis %lld", static_cast<long long>(ps));
  }
  if (halo_pad && src->halo != st->radius)
    return fail(STH_EINVAL, "Halo padding: halo width %d differs from the stencil radius %d", src->halo,
                st->radius);
  if (halo_pad && (bc == STH_WRAP || bc == STH_REFLECT))
    for (int d = 0; d < src->ndim; ++d)
      if (src->halo >= src->size[d])
        return fail(STH_EINVAL, "BoundsError: halo %d >= axis %lld under Wrap/Reflect", src->halo,
                    static_cast<long long>(src->size[d]));

  const int sdt = src->dtype;
  const int wdt = st->weights ? st->weights_dtype : -1;
  if (op->kind == STH_OP_KERNELPRODUCT && !st->weights)
    return fail(STH_EINVAL, "kernelproduct needs a Kernel stencil (weights)");
  const int odt = out_dtype(sdt, *op, wdt);
  if (odt < 0) return fail(STH_ENOTSUP, "op %d over dtype %d (weights %d) is not accelerated", op->kind, sdt, wdt);
  // Integer sum / mean whose result depends on how the caller's `sum` folds
  // a StaticVector (wrapping in T or widening): only with STH_OPF_INT_WRAP.
  if (st->length >= 2 && !(op->flags & STH_OPF_INT_WRAP) && !is_float(sdt) && sdt != STH_BOOL) {
    const bool narrow_int = dtype_size(sdt) < 8;
    if ((op->kind == STH_OP_SUM && narrow_int) || op->kind == STH_OP_MEAN)
      return fail(STH_ENOTSUP,
                  "%s over integer dtype %d with %d offsets: overflow semantics unpinned (pass STH_OPF_INT_WRAP "
                  "if the caller's sum wraps in the element type)",
                  op->kind == STH_OP_SUM ? "sum" : "mean", sdt, st->length);
  }
  if (odt != dst->dtype)
    return fail(STH_EINVAL, "dest dtype %d does not match the result type %d", dst->dtype, odt);

  // offsets, zero-extended to 3 dims; reach per dim
  const int L = st->length;
  std::vector<int> offs3(3 * L, 0);
  int reach[3] = {0, 0, 0};
  for (int k = 0; k < L; ++k)
    for (int d = 0; d < st->ndim; ++d) {
      const int v = st->offsets[k * st->ndim + d];
      offs3[3 * k + d] = v;
      reach[d] = std::max(reach[d], std::abs(v));
    }
  const int64_t n[3] = {src->size[0], src->ndim > 1 ? src->size[1] : 1, src->ndim > 2 ? src->size[2] : 1};
  for (int d = 0; d < 3; ++d) {
    if (halo_pad && d < src->ndim && reach[d] > src->halo)
      return fail(STH_ENOTSUP, "stencil reach %d exceeds the halo %d (reference reads out of bounds)", reach[d],
                  src->halo);
    if (!halo_pad && reach[d] >= n[d] && reach[d] > 0)
      return fail(STH_ENOTSUP, "stencil reach %d >= axis %lld (reference reads out of bounds)", reach[d],
                  static_cast<long long>(n[d]));
    if (reach[d] > 127) return fail(STH_ENOTSUP, "stencil reach %d too large", reach[d]);
  }

  // aliasing
  {
    uintptr_t a0, a1, b0, b1;
    byte_range(*src, &a0, &a1);
    byte_range(*dst, &b0, &b1);
    if (a0 < b1 && b0 < a1) return fail(STH_EINVAL, "dest overlaps source");
  }

  // Zero every byte (padding included): plans are compared bytewise as graph keys.
  Plan& pl = *out;
  std::memset(static_cast<void*>(&pl), 0, sizeof(Plan));
  MapParams& p = pl.p;
  p.src = make_devarr(*src, halo_pad);
  p.dst = make_devarr(*dst, true);
  p.bc = bc;
  p.read_halo = halo_pad ? 1 : 0;
  p.rx = reach[0];
  p.ry = reach[1];
  p.rz = reach[2];
  p.L = L;
  p.padval = 0;
  if (bc == STH_REMOVE) std::memcpy(&p.padval, padval, dtype_size(sdt));
  p.op.alpha = op->alpha;
  p.op.birth = op->birth_mask;
  p.op.survive = op->survive_mask;
  p.xcd_remap = 1;

  if (op->kind == STH_OP_KERNELPRODUCT) {
    if (L > kMaxRuntimeL) return fail(STH_ENOTSUP, "kernel with %d weights (max %d)", L, kMaxRuntimeL);
    const char* w = static_cast<const char*>(st->weights);
    for (int k = 0; k < L; ++k) {
      if (sdt == STH_F64) {
        std::memcpy(&p.w.d[k], w + 8 * k, 8);
      } else if (is_float(sdt)) {
        std::memcpy(&p.w.f[k], w + 4 * k, 4);
      } else {
        const int es = dtype_size(wdt);
        uint64_t u = 0;
        std::memcpy(&u, w + es * k, es);
        if (is_signed(wdt) && es < 8) {
          const int sh = 64 - 8 * es;
          u = static_cast<uint64_t>(static_cast<int64_t>(u << sh) >> sh);
        }
        p.w.u64[k] = u;
      }
    }
  }

  // ---- kernel choice ------------------------------------------------
  const bool moore1 = st->ndim == 2 && L == 8 && [&] {
    static const int m[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}};
    for (int k = 0; k < 8; ++k)
      if (offs3[3 * k] != m[k][0] || offs3[3 * k + 1] != m[k][1]) return false;
    return true;
  }();
  const auto aligned16 = [](const DevArr& d) {
    return (reinterpret_cast<uintptr_t>(d.base) % 16 == 0) && (d.s1 % 16 == 0);
  };
  std::vector<Variant> cands;
  std::string group;
  const int64_t cells = n[0] * n[1] * n[2];
  auto ilog2 = [](int64_t v) { int r = 0; while (v > 1) { v >>= 1; ++r; } return r; };
  const std::string bucket = "|c" + std::to_string(ilog2(cells)) + "|x" + std::to_string(ilog2(n[0]));
  if (!(flags & STH_GENERIC) && op->kind == STH_OP_LIFE && (sdt == STH_U8 || sdt == STH_BOOL) &&
      moore1 && src->ndim == 2 && n[0] % 16 == 0 && aligned16(p.src) && aligned16(p.dst)) {
    const bool conway = op->birth_mask == (1ull << 3) && op->survive_mask == ((1ull << 2) | (1ull << 3));
    // Every byte this kernel writes is 0 or 1, so later steps of an iteration
    // can skip normalising the source -- unless they read a halo (Use).
    const bool steady_norm = halo_pad && bc == STH_USE;
    auto add = [&](int ky, LaunchFn norm, LaunchFn raw, const char* name) {
      Variant v;
      v.fn = norm;
      v.fn_steady = steady_norm ? norm : raw;
      v.name = name;
      v.tiles_x = static_cast<int32_t>((n[0] + 1023) / 1024);
      v.tiles_y = static_cast<int32_t>((n[1] + 4 * ky - 1) / (4 * ky));
      v.tiles_z = 1;
      cands.push_back(v);
    };
    // rows per wave: 4 is the default; the autotuner may pick another
    if (conway) {
      add(4, &launch_life<4, true, true>, &launch_life<4, true, false>, "life_conway/ky4");
      add(2, &launch_life<2, true, true>, &launch_life<2, true, false>, "life_conway/ky2");
      add(8, &launch_life<8, true, true>, &launch_life<8, true, false>, "life_conway/ky8");
      add(16, &launch_life<16, true, true>, &launch_life<16, true, false>, "life_conway/ky16");
    } else {
      add(4, &launch_life<4, false, true>, &launch_life<4, false, false>, "life_table/ky4");
      add(2, &launch_life<2, false, true>, &launch_life<2, false, false>, "life_table/ky2");
      add(8, &launch_life<8, false, true>, &launch_life<8, false, false>, "life_table/ky8");
      add(16, &launch_life<16, false, true>, &launch_life<16, false, false>, "life_table/ky16");
    }
    group = std::string(conway ? "life_conway" : "life_table") + (sdt == STH_BOOL ? "_bool" : "_u8") + bucket;
    if (halo_pad && bc != STH_USE) {
      // map boundaries in registers and refresh the source halo in the same launch
      p.read_halo = 0;
      p.fuse_halo = 1;
    }
    // Conway's rule under Wrap / Remove / Reflect on grids below 2 GiB:
    // several generations per launch with one bit per cell (lifeb_kernel)
    const bool bitlife = conway && (bc == STH_WRAP || bc == STH_REMOVE || bc == STH_REFLECT) && n[0] >= 16 &&
                         zm2_extent(p.src, 1) < kZm2MaxBytes && zm2_extent(p.dst, 1) < kZm2MaxBytes;
    if (bc != STH_USE && n[1] >= 3 && bitlife) {
      // generations per launch (STH_LIFEB_G: 2, 4 or 8) and waves per launch
      // (STH_LIFEB_WAVES), read per call
      // (8 measured fastest on 8192^2: 3.6 us per generation, against 5.8
      // with 4 and 10.7 with 2)
      const char* eg = std::getenv("STH_LIFEB_G");
      const int lg0 = eg ? std::atoi(eg) : 8;
      const int lg = lg0 == 2 || lg0 == 4 ? lg0 : 8;
      const char* ew = std::getenv("STH_LIFEB_WAVES");
      const int64_t lwaves = ew ? std::max<int64_t>(1, std::atoll(ew)) : int64_t(2048);
      // plain stores: the next launch reads them back, from the Infinity
      // Cache when the grid fits (STH_LIFEB_NT=1: non-temporal, to compare)
      const char* en = std::getenv("STH_LIFEB_NT");
      const bool lnt = en && std::strcmp(en, "1") == 0;
      pl.fn_fused = lg == 2 ? &launch_lifeb<2, 6, 0>
                  : lg == 4 ? &launch_lifeb<4, 6, 0>
                            : (lnt ? &launch_lifeb<8, 6, 1> : &launch_lifeb<8, 6, 0>);
      pl.fused_gens = lg;
      lifeb_grid(n[0], n[1], lg, lwaves, &pl.fused_tiles_x, &pl.fused_tiles_y);
      pl.fused_tiles_z = 1;
    } else if (bc != STH_USE && n[1] >= 3) {
      // temporal blocking for steady iteration steps (two generations per launch)
      int ky2 = 8;  // rows per wave of the fused kernel (measured best; STH_LIFE2_KY: 2, 4 or 8)
      if (const char* e = std::getenv("STH_LIFE2_KY")) ky2 = std::atoi(e);
      if (ky2 != 2 && ky2 != 4) ky2 = 8;
      pl.fn_fused = ky2 == 2 ? (conway ? &launch_life2<2, true> : &launch_life2<2, false>)
                   : ky2 == 8 ? (conway ? &launch_life2<8, true> : &launch_life2<8, false>)
                              : (conway ? &launch_life2<4, true> : &launch_life2<4, false>);
      pl.fused_gens = 2;
      pl.fused_tiles_x = static_cast<int32_t>((n[0] + 1023) / 1024);
      pl.fused_tiles_y = static_cast<int32_t>((n[1] + 4 * ky2 - 1) / (4 * ky2));
      pl.fused_tiles_z = 1;
    }
  } else {
    if (!(flags & STH_GENERIC)) {
      std::string family;  // shapes with equal offsets (e.g. Window(1), Circle(1)) share kernels
      for (const auto& e : registry()) {
        if (e.L == 0 || e.dtype != sdt || e.op != op->kind || e.L != L || e.ndim != st->ndim) continue;
        if (!offsets_equal(e.offs, L, offs3)) continue;
        const char* slash = std::strchr(e.name, '/');
        const std::string fam(e.name, slash ? static_cast<size_t>(slash - e.name) : std::strlen(e.name));
        if (family.empty()) family = fam;
        else if (fam != family) continue;
        Variant v;
        v.fn = v.fn_steady = e.fn;
        v.name = e.name;
        if (e.cfg == CFG_STREAM) {
          v.lpr_shift = lpr_shift_for(n[0], e.sv);
          const int64_t lpr = int64_t(1) << v.lpr_shift, rpw = 64 / lpr;
          v.tiles_x = static_cast<int32_t>((n[0] + lpr * e.sv - 1) / (lpr * e.sv));
          v.tiles_y = static_cast<int32_t>((n[1] + 4 * e.sky * rpw - 1) / (4 * e.sky * rpw));
          if (n[2] > (int64_t(1) << 31) / 2) continue;
          v.tiles_z = static_cast<int32_t>(n[2]);
        } else if (e.cfg == CFG_ZMARCH1) {
          // z-marching one-step kernel: 3-D arrays under Wrap / Remove /
          // Reflect, rows a whole number of V-cell vectors on V-cell
          // boundaries, byte extents within its 32-bit buffer offsets
          const int es = dtype_size(sdt), eo = dtype_size(odt);
          const auto vec_aligned = [&](const DevArr& a, int esz) {
            const int64_t vb = static_cast<int64_t>(e.sv) * esz;
            return reinterpret_cast<uintptr_t>(a.base) % vb == 0 && (a.s1 * esz) % vb == 0 &&
                   (a.s2 * esz) % vb == 0;
          };
          if (src->ndim != 3 || !(bc == STH_WRAP || bc == STH_REMOVE || bc == STH_REFLECT) ||
              n[0] % e.sv != 0 || !vec_aligned(p.src, es) || !vec_aligned(p.dst, eo) ||
              zm2_extent(p.src, es) >= kZm2MaxBytes || zm2_extent(p.dst, eo) >= kZm2MaxBytes)
            continue;
          static const int64_t zm1_waves = [] {
            const char* w = std::getenv("STH_ZM1_WAVES");
            return w ? std::max<int64_t>(1, std::atoll(w)) : int64_t(8192);
          }();
          zmarch1_grid(n[0], n[1], n[2], e.sv, e.sky, zm1_waves, &v.tiles_x, &v.tiles_y, &v.tiles_z);
        } else {
          const TileDims td = tile_dims(e.cfg);
          const int staged = (sdt == STH_BF16 || sdt == STH_F16) ? 4 : dtype_size(sdt);
          const int64_t lds = static_cast<int64_t>(td.tx + 2 * reach[0]) * (td.ty + 2 * reach[1]) *
                              (td.tz + 2 * reach[2]) * staged;
          if (lds > 65536) continue;
          v.tiles_x = static_cast<int32_t>((n[0] + td.tx - 1) / td.tx);
          v.tiles_y = static_cast<int32_t>((n[1] + td.ty - 1) / td.ty);
          v.tiles_z = static_cast<int32_t>((n[2] + td.tz - 1) / td.tz);
        }
        if (static_cast<int64_t>(v.tiles_x) * v.tiles_y * v.tiles_z > (int64_t(1) << 31) - 1) continue;
        cands.push_back(v);
      }
      if (!cands.empty()) {
        const char* slash = std::strchr(cands[0].name, '/');
        group.assign(cands[0].name, slash ? static_cast<size_t>(slash - cands[0].name) : std::strlen(cands[0].name));
        group += bucket;
      }
    }
    if (cands.empty()) {
      if (L > kMaxRuntimeL) return fail(STH_ENOTSUP, "stencil with %d offsets (runtime max %d)", L, kMaxRuntimeL);
      const int cfg = reach[2] > 0 ? CFG_RT3D : (n[1] == 1 && reach[1] == 0 ? CFG_RT1D : CFG_RT2D);
      const KernelEntry* pick = nullptr;
      for (const auto& e : registry())
        if (e.L == 0 && e.dtype == sdt && e.op == op->kind && e.cfg == cfg) {
          pick = &e;
          break;
        }
      if (!pick) return fail(STH_ENOTSUP, "no kernel for dtype %d op %d", sdt, op->kind);
      // a runtime-compiled streaming kernel for this shape, ahead of the
      // runtime-offset kernel (which stays the tuner's alternative)
      if (!(flags & STH_GENERIC) && jit_wanted(n[0] * n[1] * n[2]) && src->ndim >= 2 && op->kind != STH_OP_LIFE &&
          std::max(reach[0], std::max(reach[1], reach[2])) <= kJitMaxReach && n[2] <= (int64_t(1) << 30)) {
        const int es = dtype_size(sdt);
        const int rmax = std::max(reach[0], std::max(reach[1], reach[2]));
        int v = 16 / es;
        if (es == 2 && (rmax >= 2 || reach[2] > 0)) v = 4;
        if (es == 1) v = 8;  // 16 one-byte cells per lane take 140+ VGPRs
        // four rows per lane (two at reach 3 for 1- and 8-byte cells, whose
        // value box would take too many registers)
        const int ky = (reach[2] > 0 || rmax <= 2 || es == 2 || es == 4) ? 4 : 2;
        std::vector<int8_t> o8(3 * L);
        for (int i = 0; i < 3 * L; ++i) o8[i] = static_cast<int8_t>(offs3[i]);
        std::string why;
        const JitKernel* jk = jit_stream_kernel(sdt, op->kind, L, o8.data(), reach[0], reach[1], reach[2], v, ky, &why);
        if (!jk) jit_note("stream kernel", why);
        if (jk) {
          Variant jv;
          jv.jit = jk->fn;
          jv.name = jk->name.c_str();
          jv.lpr_shift = lpr_shift_for(n[0], v);
          const int64_t lpr = int64_t(1) << jv.lpr_shift, rpw = 64 / lpr;
          const int64_t jtx = (n[0] + lpr * v - 1) / (lpr * v), jty = (n[1] + 4 * ky * rpw - 1) / (4 * ky * rpw);
          if (jtx * jty * n[2] <= (int64_t(1) << 31) - 1) {
            jv.tiles_x = static_cast<int32_t>(jtx);
            jv.tiles_y = static_cast<int32_t>(jty);
            jv.tiles_z = static_cast<int32_t>(n[2]);
            cands.push_back(jv);
          }
        }
      }
      for (int k = 0; k < L; ++k)
        p.offs[k] = (offs3[3 * k] & 0xff) | ((offs3[3 * k + 1] & 0xff) << 8) | ((offs3[3 * k + 2] & 0xff) << 16);
      const TileDims td = tile_dims(pick->cfg);
      const int staged = (sdt == STH_BF16 || sdt == STH_F16) ? 4 : dtype_size(sdt);
      const int64_t lds = static_cast<int64_t>(td.tx + 2 * reach[0]) * (td.ty + 2 * reach[1]) *
                          (td.tz + 2 * reach[2]) * staged;
      if (lds > 65536)
        return fail(STH_ENOTSUP, "stencil reach needs %lld bytes of LDS per tile", static_cast<long long>(lds));
      const int64_t tx = (n[0] + td.tx - 1) / td.tx, ty = (n[1] + td.ty - 1) / td.ty,
                    tz = (n[2] + td.tz - 1) / td.tz;
      if (tx * ty * tz > (int64_t(1) << 31) - 1) return fail(STH_ENOTSUP, "array too large for one launch");
      Variant v;
      v.fn = v.fn_steady = pick->fn;
      v.name = pick->name;
      v.tiles_x = static_cast<int32_t>(tx);
      v.tiles_y = static_cast<int32_t>(ty);
      v.tiles_z = static_cast<int32_t>(tz);
      cands.push_back(v);
      if (cands.size() > 1) {  // a specialisation and the runtime-offset kernel: the tuner picks
        const char* slash = std::strchr(cands[0].name, '/');
        group.assign(cands[0].name, slash ? static_cast<size_t>(slash - cands[0].name) : std::strlen(cands[0].name));
        group += bucket;
      }
    }
  }
  apply_variant(pl, cands[0]);
  if (cands_out) *cands_out = std::move(cands);

  // two steps per launch for 3-D radius-1 catalogue stencils (Wrap / Remove /
  // Reflect)
  // on arrays large enough for the z-marching kernel to beat two streaming
  // steps: from 128^3 cells it measured 1.0-1.5x faster per step, from 256^3
  // 1.5-2.8x (STH_FUSE3D_MIN_CELLS overrides the threshold)
  static const int64_t fuse3d_min = [] {
    const char* e = std::getenv("STH_FUSE3D_MIN_CELLS");
    return e ? std::atoll(e) : (int64_t(1) << 21);
  }();
  if (!pl.fn_fused && !(flags & (STH_NO_FUSE | STH_GENERIC)) && src->ndim == 3 && st->ndim == 3 &&
      (bc == STH_WRAP || bc == STH_REMOVE || bc == STH_REFLECT) && n[0] >= 3 && n[1] >= 3 && n[2] >= 3 &&
      n[0] * n[1] * n[2] >= fuse3d_min &&
      zm2_extent(pl.p.src, dtype_size(sdt)) < kZm2MaxBytes && zm2_extent(pl.p.dst, dtype_size(odt)) < kZm2MaxBytes) {
    for (const auto& e : fused2_registry()) {
      if (e.dtype != sdt || e.op != op->kind || e.L != L || e.ndim != st->ndim) continue;
      if (!offsets_equal(e.offs, L, offs3)) continue;
      static const int64_t zm_waves = [] {
        const char* v = std::getenv("STH_ZM2_WAVES");
        return v ? std::max<int64_t>(1, std::atoll(v)) : int64_t(8192);
      }();
      pl.fn_fused = e.fn;
      pl.fused_gens = 2;
      zmarch2_grid(n[0], n[1], n[2], e.rows, zm_waves, &pl.fused_tiles_x, &pl.fused_tiles_y, &pl.fused_tiles_z);
      break;
    }
    // other radius-1 shapes and element types: the same kernel, runtime-compiled
    if (!pl.fn_fused && sdt == odt && dtype_size(sdt) >= 2 && reach[0] <= 1 && reach[1] <= 1 && reach[2] <= 1 &&
        op->kind != STH_OP_LIFE && jit_wanted(n[0] * n[1] * n[2])) {
      std::vector<int8_t> o8(3 * static_cast<size_t>(L));
      for (size_t i = 0; i < o8.size(); ++i) o8[i] = static_cast<int8_t>(offs3[i]);
      std::string why;
      if (const JitKernel* jk = jit_fused_kernel(sdt, op->kind, L, o8.data(), 1, 1, 1, 3, bc, &why)) {
        const char* v = std::getenv("STH_ZM2_WAVES");
        const int64_t zw = v ? std::max<int64_t>(1, std::atoll(v)) : int64_t(8192);
        pl.jit_fused = jk->fn;
        pl.jit_fused_3d = 1;
        pl.fused_gens = 2;
        zmarch2_grid(n[0], n[1], n[2], jit_zmarch2_rows(sdt), zw, &pl.fused_tiles_x, &pl.fused_tiles_y,
                     &pl.fused_tiles_z);
      } else {
        jit_note("two-step 3-D kernel", why);
      }
    }
  }

  // two steps per launch for 2-D radius-1 catalogue stencils (Wrap / Remove /
  // Reflect)
  // on arrays whose rows start on 16-byte boundaries (the y-marching kernel
  // loads whole vectors) and of 2^21 cells or more: 1.3-1.7x faster per step
  // from 4096^2, mixed (0.6-1.3x) at 1024^2 (STH_FUSE2D_MIN_CELLS overrides)
  static const int64_t fuse2d_min = [] {
    const char* e = std::getenv("STH_FUSE2D_MIN_CELLS");
    return e ? std::atoll(e) : (int64_t(1) << 21);
  }();
  auto rows16 = [](const DevArr& a, int64_t es) {
    return reinterpret_cast<uintptr_t>(a.base) % 16 == 0 && (a.s1 * es) % 16 == 0;
  };
  if (!pl.fn_fused && !(flags & (STH_NO_FUSE | STH_GENERIC)) && src->ndim == 2 && st->ndim == 2 &&
      (bc == STH_WRAP || bc == STH_REMOVE || bc == STH_REFLECT) && n[0] >= 3 && n[1] >= 3 &&
      n[0] * n[1] >= fuse2d_min &&
      sdt == odt && rows16(pl.p.src, dtype_size(sdt)) && rows16(pl.p.dst, dtype_size(odt)) &&
      zm2_extent(pl.p.src, dtype_size(sdt)) < kZm2MaxBytes && zm2_extent(pl.p.dst, dtype_size(odt)) < kZm2MaxBytes) {
    for (const auto& e : fused2_registry()) {
      if (e.dtype != sdt || e.op != op->kind || e.L != L || e.ndim != 2) continue;
      if (!offsets_equal(e.offs, L, offs3)) continue;
      static const int64_t ym_waves = [] {
        const char* v = std::getenv("STH_YM2_WAVES");
        return v ? std::max<int64_t>(1, std::atoll(v)) : int64_t(8192);
      }();
      pl.fn_fused = e.fn;
      pl.fused_gens = 2;
      ymarch2_grid(n[0], n[1], e.rows, ym_waves, &pl.fused_tiles_x, &pl.fused_tiles_y, &pl.fused_tiles_z);
      break;
    }
    if (!pl.fn_fused && dtype_size(sdt) >= 2 && reach[0] <= 1 && reach[1] <= 1 && op->kind != STH_OP_LIFE &&
        jit_wanted(n[0] * n[1])) {
      std::vector<int8_t> o8(3 * static_cast<size_t>(L));
      for (size_t i = 0; i < o8.size(); ++i) o8[i] = static_cast<int8_t>(offs3[i]);
      std::string why;
      if (const JitKernel* jk = jit_fused_kernel(sdt, op->kind, L, o8.data(), 1, 1, 0, 2, bc, &why)) {
        const char* v = std::getenv("STH_YM2_WAVES");
        const int64_t yw = v ? std::max<int64_t>(1, std::atoll(v)) : int64_t(8192);
        pl.jit_fused = jk->fn;
        pl.jit_fused_3d = 0;
        pl.fused_gens = 2;
        ymarch2_grid(n[0], n[1], 16 / dtype_size(sdt), yw, &pl.fused_tiles_x, &pl.fused_tiles_y,
                     &pl.fused_tiles_z);
      } else {
        jit_note("two-step 2-D kernel", why);
      }
    }
  }

  // the batched kernel's variant for this (dtype, op, offsets), if any
  pl.dtype = sdt;
  pl.bvariant = -1;
  for (const auto& bv : batch_variants()) {
    if (bv.dtype != sdt || bv.op != op->kind || bv.L != L || bv.ndim != st->ndim) continue;
    if (!offsets_equal(bv.offs, L, offs3)) continue;
    pl.bvariant = bv.variant;
    pl.bsv = bv.sv;
    pl.bsky = bv.sky;
    pl.bklass = bv.klass;
    break;
  }
  if (group_out) *group_out = group;

  // ---- halo refresh (update_boundary!) --------------------------------
  if (halo_pad && bc != STH_USE && src->halo > 0 && !p.fuse_halo) {
    pl.halo = 1;
    HaloParams& h = pl.hp;
    h.a = make_devarr(*src, true);
    h.ndim = src->ndim;
    h.R = src->halo;
    h.bc = bc;
    h.padval = p.padval;
    for (int d = 0; d < 3; ++d) {
      int64_t cells = 0;
      if (d < src->ndim) {
        cells = 2 * static_cast<int64_t>(h.R);
        for (int e = 0; e < 3; ++e)
          if (e != d) cells *= n[e] + (e < src->ndim ? 2 * h.R : 0);
      }
      h.cells[d] = cells;
      h.blocks[d] = (cells + 255) / 256;
    }
    switch (dtype_size(sdt)) {
      case 1: pl.halo_fn = &launch_halo<uint8_t>; break;
      case 2: pl.halo_fn = &launch_halo<uint16_t>; break;
      case 4: pl.halo_fn = &launch_halo<uint32_t>; break;
      default: pl.halo_fn = &launch_halo<uint64_t>; break;
    }
  }
  return STH_OK;
}

int run_plan(const Plan& pl, hipStream_t s, bool steady = false) {
  if (pl.halo) pl.halo_fn(pl.hp, s);
  if (pl.jit) {
    STH_HIP(jit_launch(pl.jit, pl.p, s));
  } else {
    const LaunchFn f = steady ? pl.fn_steady : pl.fn;
    if (f) f(pl.p, s);
  }
  STH_HIP(hipGetLastError());
  return STH_OK;
}

// Several steps in one launch (pl.fn_fused, steady state).  The fused kernels
// map every boundary themselves and write only the destination's interior.
int run_fused(const Plan& pl, hipStream_t s) {
  MapParams q = pl.p;
  q.tiles_x = pl.fused_tiles_x;
  q.tiles_y = pl.fused_tiles_y;
  q.tiles_z = pl.fused_tiles_z;
  q.read_halo = 0;
  q.fuse_halo = 0;
  if (pl.jit_fused) {
    const unsigned nb = static_cast<unsigned>(q.tiles_x) * q.tiles_y * (pl.jit_fused_3d ? q.tiles_z : 1);
    void* args[] = {&q};
    STH_HIP(hipModuleLaunchKernel(pl.jit_fused, nb, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  } else {
    pl.fn_fused(q, s);
  }
  STH_HIP(hipGetLastError());
  return STH_OK;
}

// A job that fills the device by itself (a full wave of resident workgroups)
// runs as its own launch of the tuned single-array kernel; batching pays off
// for the jobs whose separate launches would be overhead-bound.
bool runs_alone(const Plan& pl, int cus) {
  const int64_t own_tiles = static_cast<int64_t>(pl.p.tiles_x) * pl.p.tiles_y * pl.p.tiles_z;
  return own_tiles >= 8 * static_cast<int64_t>(cus);
}

}  // namespace
}  // namespace sth

// Job-table slot of sth_map_batch: pinned host staging + device copy; the
// event marks the launch that last read it.
struct BatchSlot {
  void* host = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;     // capacity
  size_t content = 0;   // bytes of the table the device copy holds
  uint64_t last_use = 0;
  hipEvent_t done = nullptr;
  bool used = false;
};

// Job tables kept on the device.  A call whose table equals one already
// there launches on it without a copy (the serving loop alternates between
// two tables per group, one per step parity); otherwise the least recently
// used slot is overwritten, once its last launch has finished with it.
constexpr int kBatchSlots = 32;

constexpr size_t kMaxGraphs = 512;  // cached graph executables per context
constexpr int kBatchStreams = 4;    // most streams a batched call spreads its launches over

struct sth_ctx {
  int device = 0;
  int cus = 0;
  sth::BatchJob* scratch = nullptr;  // this device's copy of `scratch` (sth_map_batch_static)
  hipStream_t capture = nullptr;  // private stream used only for graph capture
  std::mutex mu;
  std::map<std::string, hipGraphExec_t> graphs;
  // device job tables a captured batched iteration's graph launches read
  // (owned by that graph, freed with it)
  std::map<std::string, std::vector<void*>> graph_tables;
  std::vector<void*>* capture_tables = nullptr;  // set while such a graph is captured
  hipStream_t upload = nullptr;                  // their uploads (non-blocking, outside the graph)
  // streams a batched call spreads its launches over (default priority),
  // created on first use, with the fork and join events that order them
  // after and before the caller's stream
  std::mutex batch_mu;
  hipStream_t batch_streams[kBatchStreams] = {};
  hipEvent_t batch_fork = nullptr;
  hipEvent_t batch_join[kBatchStreams] = {};
  BatchSlot slots[kBatchSlots];
  uint64_t slot_clock = 0;
  // autotuner: shape family + size bucket -> chosen variant name
  std::map<std::string, std::string> tuned;
  bool autotune = true;    // STH_AUTOTUNE=0 disables timing (defaults are used)
  std::string forced;      // sth_set_variant: pick the first variant whose name contains it
  std::string tune_file;   // STH_TUNE_FILE: decisions are loaded from and appended to it
};

using namespace sth;

#define STH_GUARD_BEGIN try {
#define STH_GUARD_END                                                   \
  }                                                                     \
  catch (const std::exception& e) {                                     \
    return fail(STH_EDEVICE, "internal error: %s", e.what());           \
  }                                                                     \
  catch (...) {                                                         \
    return fail(STH_EDEVICE, "internal error");                         \
  }

namespace {
int set_device(sth_ctx* ctx) {
  if (!ctx) return fail(STH_EINVAL, "context is NULL");
  STH_HIP(hipSetDevice(ctx->device));
  return STH_OK;
}

// Streams a batched call uses (STH_BATCH_STREAMS, 1..kBatchStreams; 1 keeps
// every launch on the caller's stream).
int batch_streams_wanted() {
  static const int v = [] {
    const char* e = std::getenv("STH_BATCH_STREAMS");
    const int k = e ? std::atoi(e) : kBatchStreams;
    return k < 1 ? 1 : (k > kBatchStreams ? kBatchStreams : k);
  }();
  return v;
}

// Creates the first `n` batch streams and the fork / join events (under
// ctx->batch_mu).
int ensure_batch_streams(sth_ctx* ctx, int n) {
  if (!ctx->batch_fork) STH_HIP(hipEventCreateWithFlags(&ctx->batch_fork, hipEventDisableTiming));
  for (int k = 0; k < n; ++k) {
    if (!ctx->batch_streams[k]) STH_HIP(hipStreamCreateWithFlags(&ctx->batch_streams[k], hipStreamNonBlocking));
    if (!ctx->batch_join[k]) STH_HIP(hipEventCreateWithFlags(&ctx->batch_join[k], hipEventDisableTiming));
  }
  return STH_OK;
}

// Autotuner: the variants of a call compute identical results (same fold
// expressions per cell), so the fastest may be chosen freely.  Each variant
// runs once to warm up and three times between events; the fastest average
// wins and is remembered per (shape family, element type, op, size bucket).
// Timing is skipped while the stream is being captured, with STH_NO_TUNE,
// or when `lookup_only` is set; the default (first) variant is used then.
int tune_pick(sth_ctx* ctx, const Plan& pl, const std::vector<Variant>& c, const std::string& group,
              uint32_t flags, hipStream_t s, bool force, bool lookup_only, int* idx) {
  *idx = 0;
  if (c.size() <= 1 || group.empty()) return STH_OK;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!ctx->forced.empty()) {
      for (size_t i = 0; i < c.size(); ++i)
        if (std::strstr(c[i].name, ctx->forced.c_str())) {
          *idx = static_cast<int>(i);
          return STH_OK;
        }
    }
  }
  if (!force) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tuned.find(group);
    if (it != ctx->tuned.end()) {
      for (size_t i = 0; i < c.size(); ++i)
        if (it->second == c[i].name) {
          *idx = static_cast<int>(i);
          return STH_OK;
        }
    }
  }
  if (lookup_only || (flags & STH_NO_TUNE) || (!ctx->autotune && !force)) return STH_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  STH_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) return STH_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  STH_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(STH_EDEVICE, "hipEventCreate failed");
  }
  int rc = STH_OK;
  // two interleaved passes (clocks settle during the first), the best time
  // per launch of each variant; the second pass times enough launches to
  // span about a quarter of a millisecond, so that kernels of a few
  // microseconds are not compared on noise
  std::vector<float> best(c.size(), 0.0f);
  int reps = 3;
  for (int pass = 0; pass < 2 && rc == STH_OK; ++pass) {
    for (size_t i = 0; i < c.size() && rc == STH_OK; ++i) {
      Plan t = pl;
      apply_variant(t, c[i]);
      if (pass == 0) rc = run_plan(t, s);  // warm-up (first launch loads the code object)
      if (rc == STH_OK && hipEventRecord(e0, s) != hipSuccess) rc = fail(STH_EDEVICE, "hipEventRecord failed");
      for (int r = 0; r < reps && rc == STH_OK; ++r) rc = run_plan(t, s);
      if (rc == STH_OK && hipEventRecord(e1, s) != hipSuccess) rc = fail(STH_EDEVICE, "hipEventRecord failed");
      if (rc == STH_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(STH_EDEVICE, "hipEventSynchronize failed");
      float ms = 0.0f;
      if (rc == STH_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        rc = fail(STH_EDEVICE, "hipEventElapsedTime failed");
      const float per = ms / static_cast<float>(reps);
      if (rc == STH_OK && (pass == 0 || per < best[i])) best[i] = per;
    }
    if (pass == 0) {
      const float fastest = *std::min_element(best.begin(), best.end());
      const float want = fastest > 0.0f ? 0.25f / fastest : 64.0f;
      reps = static_cast<int>(std::min(64.0f, std::max(3.0f, std::ceil(want))));
    }
  }
  int bi = 0;
  for (size_t i = 1; i < c.size(); ++i)
    if (best[i] < best[bi]) bi = static_cast<int>(i);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *idx = bi;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->tuned[group] = c[bi].name;
  if (!ctx->tune_file.empty()) {
    if (FILE* f = std::fopen(ctx->tune_file.c_str(), "a")) {
      std::fprintf(f, "%s\t%s\n", group.c_str(), c[bi].name);
      std::fclose(f);
    }
  }
  return STH_OK;
}

// One step of many independent calls, first half: calls without a batched
// variant, and calls that fill the device alone, run here as single launches
// on `s`; the rest are grouped per batched kernel ((element type, variant)
// key), each after its halo refresh.  `serial_only` runs every call alone.
using BatchGroups = std::map<int64_t, std::vector<BatchJob>>;

// Relative cost of one call, for spreading launches over streams: its cells,
// weighted by the stencil length (the fold's work).
double plan_cost(const Plan& pl) {
  return static_cast<double>(pl.p.src.n0 * pl.p.src.n1 * pl.p.src.n2) * (1.0 + pl.p.L / 8.0);
}

// Whether calls that fill the device alone run as their own launches of the
// tuned kernel (STH_BATCH_ALONE=0: every catalogue call joins its group).
bool batch_alone_enabled() {
  static const bool v = [] {
    const char* e = std::getenv("STH_BATCH_ALONE");
    return !(e && std::strcmp(e, "0") == 0);
  }();
  return v;
}

// `alone` (if not null) receives the calls that run as single launches
// instead of running them on `s`; `cost` (if not null) the cost per group.
int group_batched(sth_ctx* ctx, const Plan* const* plans, int32_t n, bool serial_only, hipStream_t s,
                  BatchGroups* groups, std::vector<const Plan*>* alone = nullptr,
                  std::map<int64_t, double>* cost = nullptr) {
  int rc = STH_OK;
  for (int32_t i = 0; i < n; ++i) {
    const Plan& pl = *plans[i];
    if (serial_only || pl.bvariant < 0 || !(pl.fn || pl.jit) || (batch_alone_enabled() && runs_alone(pl, ctx->cus))) {
      if (!(pl.fn || pl.jit || pl.halo)) continue;
      if (alone) alone->push_back(&pl);
      else if ((rc = run_plan(pl, s))) return rc;
      continue;
    }
    if (cost) (*cost)[static_cast<int64_t>(pl.dtype) * 65536 + pl.bklass] += plan_cost(pl);
    if (pl.halo) {  // update_boundary! before the batched gather
      pl.halo_fn(pl.hp, s);
      STH_HIP(hipGetLastError());
    }
    BatchJob bj;
    std::memset(static_cast<void*>(&bj), 0, sizeof bj);
    bj.src = pl.p.src;
    bj.dst = pl.p.dst;
    bj.bc = pl.p.bc;
    bj.read_halo = pl.p.read_halo;
    bj.padval = pl.p.padval;
    bj.op = pl.p.op;
    bj.variant = pl.bvariant;
    bj.lpr_shift = lpr_shift_for(pl.p.src.n0, pl.bsv);
    const int64_t lpr = int64_t(1) << bj.lpr_shift, rpw = 64 / lpr;
    bj.tiles_x = static_cast<int32_t>((pl.p.src.n0 + lpr * pl.bsv - 1) / (lpr * pl.bsv));
    bj.tiles_y = static_cast<int32_t>((pl.p.src.n1 + 4 * pl.bsky * rpw - 1) / (4 * pl.bsky * rpw));
    bj.tiles_z = static_cast<int32_t>(pl.p.src.n2);
    (*groups)[static_cast<int64_t>(pl.dtype) * 65536 + pl.bklass].push_back(bj);
  }
  return STH_OK;
}

// Second half, per group: its kernel, the jobs' first tiles and the launch
// shape (one tile per workgroup: a workgroup's tiles run one after another
// with no overlap inside a wave; STH_BATCH_TPB sets strips).  Sets *total to
// 0 for a group with no tiles.
int prepare_group(int64_t key, std::vector<BatchJob>& bj, const BatchKernelEntry** ke, BatchParams* bp,
                  int64_t* blocks) {
  *ke = nullptr;
  for (const auto& e : batch_kernels())
    if (static_cast<int64_t>(e.dtype) * 65536 + e.klass == key) *ke = &e;
  if (!*ke) return fail(STH_EDEVICE, "no batched kernel for dtype %d variant %d", static_cast<int>(key / 65536),
                        static_cast<int>(key % 65536));
  int64_t total = 0;
  for (auto& j : bj) {
    j.tile_begin = total;
    total += static_cast<int64_t>(j.tiles_x) * j.tiles_y * j.tiles_z;
  }
  int64_t tpb = 1;
  if (const char* e = std::getenv("STH_BATCH_TPB")) tpb = std::max(1, std::atoi(e));
  *blocks = (total + tpb - 1) / tpb;
  if (*blocks > (int64_t(1) << 31) - 1)
    return fail(STH_ENOTSUP, "batch of %lld tiles is too large", static_cast<long long>(total));
  bp->jobs = nullptr;
  bp->njobs = static_cast<int32_t>(bj.size());
  bp->tiles_per_block = static_cast<int32_t>(tpb);
  bp->total_tiles = total;
  return STH_OK;
}

// Runs `launch(table)` on stream `st` with `table` the device copy of `bytes`
// bytes at `data`, held in one of the context's slots (pinned staging plus a
// device buffer): a slot already holding the same bytes is reused (the
// stream waits for its upload and earlier launches), else the least recently
// used slot is refilled once its last launch has finished.
template <class F>
int with_device_table(sth_ctx* ctx, const void* data, size_t bytes, hipStream_t st, F&& launch) {
  if (ctx->capture_tables) {
    // a graph is being captured: the table gets device memory of its own,
    // filled now (before any replay), which the graph keeps until evicted
    void* dev = nullptr;
    STH_HIP(hipMalloc(&dev, std::max<size_t>(bytes, 256)));
    ctx->capture_tables->push_back(dev);
    if (!ctx->upload) STH_HIP(hipStreamCreateWithFlags(&ctx->upload, hipStreamNonBlocking));
    STH_HIP(hipMemcpyAsync(dev, data, bytes, hipMemcpyHostToDevice, ctx->upload));
    STH_HIP(hipStreamSynchronize(ctx->upload));
    launch(static_cast<const void*>(dev));
    STH_HIP(hipGetLastError());
    return STH_OK;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  BatchSlot* hit = nullptr;
  for (BatchSlot& c : ctx->slots)
    if (c.used && c.content == bytes && std::memcmp(c.host, data, bytes) == 0) {
      hit = &c;
      break;
    }
  BatchSlot* lru = &ctx->slots[0];
  for (BatchSlot& c : ctx->slots)
    if (c.last_use < lru->last_use) lru = &c;
  BatchSlot& sl = hit ? *hit : *lru;
  sl.last_use = ++ctx->slot_clock;
  if (!hit && sl.used) STH_HIP(hipEventSynchronize(sl.done));  // its last launch has finished with it
  if (!hit && sl.bytes < bytes) {
    if (sl.host) STH_HIP(hipHostFree(sl.host));
    if (sl.dev) STH_HIP(hipFree(sl.dev));
    sl.host = sl.dev = nullptr;
    sl.bytes = 0;
    const size_t cap = std::max<size_t>(bytes, 16384);
    STH_HIP(hipHostMalloc(&sl.host, cap, hipHostMallocDefault));
    STH_HIP(hipMalloc(&sl.dev, cap));
    sl.bytes = cap;
  }
  if (!sl.done) STH_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  if (!hit) {
    std::memcpy(sl.host, data, bytes);
    STH_HIP(hipMemcpyAsync(sl.dev, sl.host, bytes, hipMemcpyHostToDevice, st));
    sl.content = bytes;
  } else {
    // after the slot's upload and every launch on it so far, whatever stream
    // they ran on (so an overwrite that waits for `done` waits for all)
    STH_HIP(hipStreamWaitEvent(st, sl.done, 0));
  }
  launch(static_cast<const void*>(sl.dev));
  STH_HIP(hipGetLastError());
  STH_HIP(hipEventRecord(sl.done, st));
  sl.used = true;
  return STH_OK;
}

// One step of many independent calls: calls with a batched variant go into
// one launch per (element type, shape, reduction) after their halo refresh;
// the rest, and calls that fill the device alone, run as single launches.
// The job tables are uploaded into device slots the context owns.  With
// more than one launch, the launches are spread over the context's batch
// streams (largest first, each to the least loaded stream) between a fork
// and a join on `s`, so one launch's tail overlaps the next launch's start.
int launch_batched(sth_ctx* ctx, const Plan* const* plans, int32_t n, bool serial_only, hipStream_t s) {
  BatchGroups groups;
  std::vector<const Plan*> alone;
  std::map<int64_t, double> gcost;
  // the unbatched path keeps one launch per call, in order, on `s`
  int rc = group_batched(ctx, plans, n, serial_only, s, &groups, serial_only ? nullptr : &alone, &gcost);
  if (rc) return rc;
  auto launch_group = [&](BatchGroups::value_type& kv, hipStream_t st) -> int {
    std::vector<BatchJob>& bj = kv.second;
    const BatchKernelEntry* ke = nullptr;
    BatchParams bp;
    int64_t blocks = 0;
    int r2 = prepare_group(kv.first, bj, &ke, &bp, &blocks);
    if (r2) return r2;
    if (bp.total_tiles == 0) return STH_OK;
    return with_device_table(ctx, bj.data(), bj.size() * sizeof(BatchJob), st, [&](const void* dev) {
      bp.jobs = static_cast<const BatchJob*>(dev);
      ke->fn(bp, static_cast<unsigned>(blocks), st);
    });
  };
  // the launches, most costly first
  struct Item {
    const Plan* alone;
    BatchGroups::value_type* group;
    double cost;
  };
  std::vector<Item> items;
  for (const Plan* pl : alone) items.push_back({pl, nullptr, plan_cost(*pl)});
  for (auto& kv : groups) items.push_back({nullptr, &kv, gcost[kv.first]});
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });
  auto run_item = [&](const Item& it, hipStream_t st) -> int {
    return it.alone ? run_plan(*it.alone, st) : launch_group(*it.group, st);
  };
  const int nst = batch_streams_wanted();
  if (items.size() <= 1 || nst <= 1) {
    for (const Item& it : items)
      if ((rc = run_item(it, s))) return rc;
    return STH_OK;
  }
  std::lock_guard<std::mutex> fork_lock(ctx->batch_mu);
  if ((rc = ensure_batch_streams(ctx, nst))) return rc;
  STH_HIP(hipEventRecord(ctx->batch_fork, s));
  double load[kBatchStreams] = {};
  bool used[kBatchStreams] = {};
  for (const Item& it : items) {
    int k = 0;
    for (int i = 1; i < nst; ++i)
      if (load[i] < load[k]) k = i;
    if (!used[k]) {
      STH_HIP(hipStreamWaitEvent(ctx->batch_streams[k], ctx->batch_fork, 0));
      used[k] = true;
    }
    load[k] += it.cost;
    if ((rc = run_item(it, ctx->batch_streams[k]))) return rc;
  }
  for (int k = 0; k < nst; ++k) {
    if (!used[k]) continue;
    STH_HIP(hipEventRecord(ctx->batch_join[k], ctx->batch_streams[k]));
    STH_HIP(hipStreamWaitEvent(s, ctx->batch_join[k], 0));
  }
  return STH_OK;
}

// The same step with every job table in `table` (the device's fixed
// `scratch` array, at most kStaticBatchJobs entries): each group's table is
// written by fill_jobs_kernel launches whose arguments carry the jobs, so the
// call copies nothing from host memory and allocates nothing.
int launch_batched_static(sth_ctx* ctx, const Plan* const* plans, int32_t n, bool serial_only, hipStream_t s,
                          BatchJob* table) {
  BatchGroups groups;
  int rc = group_batched(ctx, plans, n, serial_only, s, &groups);
  if (rc) return rc;
  int64_t used = 0;
  for (auto& kv : groups) used += static_cast<int64_t>(kv.second.size());
  if (used > kStaticBatchJobs)
    return fail(STH_EINVAL, "%lld batched jobs exceed the fixed table of %d", static_cast<long long>(used),
                kStaticBatchJobs);
  int64_t off = 0;
  for (auto& kv : groups) {
    std::vector<BatchJob>& bj = kv.second;
    const BatchKernelEntry* ke = nullptr;
    BatchParams bp;
    int64_t blocks = 0;
    if ((rc = prepare_group(kv.first, bj, &ke, &bp, &blocks))) return rc;
    if (bp.total_tiles == 0) continue;
    for (size_t i = 0; i < bj.size(); i += kFillJobs) {
      JobChunk c;
      std::memset(static_cast<void*>(&c), 0, sizeof c);
      const int m = static_cast<int>(std::min<size_t>(kFillJobs, bj.size() - i));
      std::memcpy(static_cast<void*>(c.j), bj.data() + i, m * sizeof(BatchJob));
      launch_fill_jobs(table + off + static_cast<int64_t>(i), c, m, s);
      STH_HIP(hipGetLastError());
    }
    bp.jobs = table + off;
    ke->fn(bp, static_cast<unsigned>(blocks), s);
    STH_HIP(hipGetLastError());
    off += static_cast<int64_t>(bj.size());
  }
  return STH_OK;
}

void load_tune_file(sth_ctx* ctx) {
  if (ctx->tune_file.empty()) return;
  FILE* f = std::fopen(ctx->tune_file.c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    char* tab = std::strchr(line, '\t');
    if (!tab) continue;
    *tab = 0;
    char* name = tab + 1;
    name[std::strcspn(name, "\r\n")] = 0;
    ctx->tuned[line] = name;
  }
  std::fclose(f);
}
}  // namespace

// Drop every cached graph (and the job tables batched ones own) once the
// device is idle: the cache's bound, and context teardown.  ctx->mu held.
static int clear_graphs(sth_ctx* ctx) {
  STH_HIP(hipDeviceSynchronize());
  for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
  ctx->graphs.clear();
  for (auto& kv : ctx->graph_tables)
    for (void* d : kv.second) (void)hipFree(d);
  ctx->graph_tables.clear();
  return STH_OK;
}

extern "C" {

int sth_version(void) { return STH_ABI_VERSION * 10000 + 0; }

const char* sth_last_error(void) { return g_err.c_str(); }

int sth_ctx_create(int device, sth_ctx** out) {
  STH_GUARD_BEGIN
  if (!out) return fail(STH_EINVAL, "out is NULL");
  int n = 0;
  STH_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(STH_EINVAL, "device %d not in 0..%d", device, n - 1);
  STH_HIP(hipSetDevice(device));
  auto* c = new sth_ctx();
  c->device = device;
  if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->cus < 1)
    c->cus = 256;
  hipError_t e = hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(STH_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  // the device address of this device's `scratch` array (the name itself is
  // a host-side stand-in that kernels must never receive)
  e = hipGetSymbolAddress(reinterpret_cast<void**>(&c->scratch), HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->capture);
    delete c;
    return fail(STH_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  }
  if (const char* v = std::getenv("STH_AUTOTUNE")) c->autotune = std::strcmp(v, "0") != 0;
  if (const char* v = std::getenv("STH_TUNE_FILE")) c->tune_file = v;
  load_tune_file(c);
  *out = c;
  g_err.clear();
  return STH_OK;
  STH_GUARD_END
}

int sth_ctx_destroy(sth_ctx* ctx) {
  STH_GUARD_BEGIN
  if (!ctx) return STH_OK;
  (void)hipSetDevice(ctx->device);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)clear_graphs(ctx);
  }
  if (ctx->upload) (void)hipStreamDestroy(ctx->upload);
  for (auto& sl : ctx->slots) {
    if (sl.done) (void)hipEventSynchronize(sl.done);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.host) (void)hipHostFree(sl.host);
    if (sl.dev) (void)hipFree(sl.dev);
  }
  if (ctx->capture) (void)hipStreamDestroy(ctx->capture);
  for (int k = 0; k < kBatchStreams; ++k) {
    if (ctx->batch_streams[k]) (void)hipStreamDestroy(ctx->batch_streams[k]);
    if (ctx->batch_join[k]) (void)hipEventDestroy(ctx->batch_join[k]);
  }
  if (ctx->batch_fork) (void)hipEventDestroy(ctx->batch_fork);
  delete ctx;
  return STH_OK;
  STH_GUARD_END
}

int sth_alloc(sth_ctx* ctx, size_t bytes, void** dptr) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (!dptr) return fail(STH_EINVAL, "dptr is NULL");
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
  if (e == hipErrorOutOfMemory) return fail(STH_ENOMEM, "hipMalloc(%zu): out of device memory", bytes);
  if (e != hipSuccess) return fail(STH_EDEVICE, "hipMalloc: %s", hipGetErrorString(e));
  return STH_OK;
  STH_GUARD_END
}

int sth_free(sth_ctx* ctx, void* dptr) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (dptr) STH_HIP(hipFree(dptr));
  return STH_OK;
  STH_GUARD_END
}

int sth_copy(sth_ctx* ctx, void* dst, const void* src, size_t bytes, int kind, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (kind < 0 || kind > 2) return fail(STH_EINVAL, "copy kind %d not in 0..2", kind);
  auto s = static_cast<hipStream_t>(stream);
  STH_HIP(hipMemcpyAsync(dst, src, bytes, k, s));
  STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

int sth_synchronize(sth_ctx* ctx, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  STH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return STH_OK;
  STH_GUARD_END
}

int sth_update_boundary(sth_ctx* ctx, const sth_array* a, int32_t radius, int32_t boundary,
                        const void* padval, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if ((rc = check_array(a, "array"))) return rc;
  if (radius != a->halo) return fail(STH_EINVAL, "radius %d differs from the halo %d", radius, a->halo);
  if (boundary < STH_REMOVE || boundary > STH_USE) return fail(STH_EINVAL, "unknown boundary %d", boundary);
  if (boundary == STH_USE || radius == 0) return STH_OK;
  if (boundary == STH_REMOVE && !padval) return fail(STH_EINVAL, "Remove boundary without a padval");
  if (boundary == STH_WRAP || boundary == STH_REFLECT)
    for (int d = 0; d < a->ndim; ++d)
      if (radius >= a->size[d])
        return fail(STH_EINVAL, "BoundsError: halo %d >= axis %lld under Wrap/Reflect", radius,
                    static_cast<long long>(a->size[d]));
  HaloParams h{};
  h.a = make_devarr(*a, true);
  h.ndim = a->ndim;
  h.R = radius;
  h.bc = boundary;
  if (boundary == STH_REMOVE) std::memcpy(&h.padval, padval, dtype_size(a->dtype));
  const int64_t n[3] = {h.a.n0, h.a.n1, h.a.n2};
  for (int d = 0; d < 3; ++d) {
    int64_t cells = 0;
    if (d < a->ndim) {
      cells = 2 * static_cast<int64_t>(radius);
      for (int e = 0; e < 3; ++e)
        if (e != d) cells *= n[e] + (e < a->ndim ? 2 * radius : 0);
    }
    h.cells[d] = cells;
    h.blocks[d] = (cells + 255) / 256;
  }
  auto s = static_cast<hipStream_t>(stream);
  switch (dtype_size(a->dtype)) {
    case 1: launch_halo<uint8_t>(h, s); break;
    case 2: launch_halo<uint16_t>(h, s); break;
    case 4: launch_halo<uint32_t>(h, s); break;
    default: launch_halo<uint64_t>(h, s); break;
  }
  STH_HIP(hipGetLastError());
  STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

int sth_map(sth_ctx* ctx, const sth_array* src, const sth_array* dst, const sth_stencil* st,
            int32_t boundary, const void* padval, int32_t padding, const sth_op* op, uint32_t flags,
            void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  Plan pl;
  std::vector<Variant> cands;
  std::string group;
  if ((rc = make_plan(src, dst, st, boundary, padval, padding, op, flags, &pl, &cands, &group))) return rc;
  auto s = static_cast<hipStream_t>(stream);
  int idx = 0;
  if ((rc = tune_pick(ctx, pl, cands, group, flags, s, false, false, &idx))) return rc;
  apply_variant(pl, cands[idx]);
  if ((rc = run_plan(pl, s))) return rc;
  if (!(flags & STH_ASYNC)) STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

static_assert(kMaxMultiTerms == STH_MAX_TERMS, "term table bound");

int sth_map_multi(sth_ctx* ctx, int32_t nterms, const sth_term* terms, const sth_array* dst, int32_t combine,
                  uint32_t flags, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (nterms < 1 || nterms > kMaxMultiTerms || !terms)
    return fail(STH_EINVAL, "multi-array gather with %d arguments (1..%d)", nterms, kMaxMultiTerms);
  if (combine != STH_COMBINE_ADD && combine != STH_COMBINE_MUL) return fail(STH_EINVAL, "unknown combine %d", combine);
  if ((rc = check_array(dst, "dest"))) return rc;
  const int dt = dst->dtype;
  if (dt != STH_F32 && dt != STH_F64 && dt != STH_U8 && dt != STH_I32 && dt != STH_I64)
    return fail(STH_ENOTSUP, "multi-array gather over dtype %d is not accelerated", dt);
  std::vector<MultiTerm> tbl(static_cast<size_t>(nterms));
  std::vector<Plan> plans(static_cast<size_t>(nterms));
  for (int32_t t = 0; t < nterms; ++t) {
    const sth_term& a = terms[t];
    if ((rc = check_array(&a.src, "argument"))) return rc;
    if (a.src.ndim != dst->ndim) return fail(STH_EINVAL, "Source array sizes must match (argument %d ndims)", t);
    for (int d = 0; d < dst->ndim; ++d)
      if (a.src.size[d] != dst->size[d])
        return fail(STH_EINVAL, "Source array sizes must match. Found: dim %d %lld and %lld (argument %d)", d + 1,
                    static_cast<long long>(dst->size[d]), static_cast<long long>(a.src.size[d]), t);
    if (a.src.dtype != dt)
      return fail(STH_ENOTSUP, "argument %d has dtype %d, dest %d (promotion is not accelerated)", t, a.src.dtype, dt);
    const bool centre = a.stencil.length == 0;
    sth_stencil st = a.stencil;
    sth_op op = a.op;
    int32_t bc = a.boundary;
    const int32_t zero3[3] = {0, 0, 0};
    if (centre) {
      // the value at I: a one-offset stencil at the centre (radius = halo, so
      // a Halo argument keeps its refresh)
      st = sth_stencil{};
      st.radius = a.src.halo;
      st.ndim = a.src.ndim;
      st.length = 1;
      st.offsets = zero3;
      op = sth_op{};
      op.kind = STH_OP_SUM;
      if (a.padding == STH_CONDITIONAL) bc = STH_WRAP;  // nothing outside is read
    } else if (out_dtype(dt, op, st.weights ? st.weights_dtype : -1) != dt) {
      return fail(STH_ENOTSUP, "argument %d: op %d over dtype %d does not return the element type", t, op.kind, dt);
    }
    if (a.scaled) {
      if (!is_float(dt)) return fail(STH_ENOTSUP, "a scaled term over an integer dtype promotes to Float64");
      if (dt == STH_F32 && static_cast<double>(static_cast<float>(a.coef)) != a.coef)
        return fail(STH_ENOTSUP, "coefficient %.17g is not a Float32 (the term would promote)", a.coef);
    }
    Plan& pl = plans[t];
    if ((rc = make_plan(&a.src, dst, &st, bc, a.padval, a.padding, &op, flags | STH_GENERIC | STH_NO_FUSE, &pl)))
      return rc;
    MultiTerm& q = tbl[t];
    std::memset(static_cast<void*>(&q), 0, sizeof(MultiTerm));
    q.src = pl.p.src;
    q.bc = pl.p.bc;
    q.read_halo = pl.p.read_halo;
    q.rx = pl.p.rx;
    q.ry = pl.p.ry;
    q.rz = pl.p.rz;
    q.L = pl.p.L;
    q.opk = op.kind;
    q.direct = centre ? 1 : 0;
    if (centre) q.src = make_devarr(a.src, true);
    q.scaled = a.scaled ? 1 : 0;
    q.padval = pl.p.padval;
    q.op = pl.p.op;
    q.coef = a.coef;
    std::memcpy(q.offs, pl.p.offs, sizeof(q.offs));
    std::memcpy(&q.w, &pl.p.w, sizeof(q.w));
  }
  // tile shape by array rank (the runtime-offset kernel's)
  const int cfg = dst->ndim == 1 ? CFG_RT1D : (dst->ndim == 2 ? CFG_RT2D : CFG_RT3D);
  const TileDims td = tile_dims(cfg);
  int64_t max_box = 0;  // LDS elements of the largest staged box
  for (const MultiTerm& q : tbl)
    if (!q.direct)
      max_box = std::max<int64_t>(max_box, int64_t(td.tx + 2 * q.rx) * (td.ty + 2 * q.ry) * (td.tz + 2 * q.rz));
  const size_t lds = static_cast<size_t>(max_box) * dtype_size(dt);
  if (lds > 65536) return fail(STH_ENOTSUP, "argument reach needs %zu bytes of LDS per tile", lds);
  const MultiKernelEntry* ke = nullptr;
  for (const auto& e : multi_registry())
    if (e.dtype == dt && e.cfg == cfg) ke = &e;
  if (!ke) return fail(STH_ENOTSUP, "no multi-array kernel for dtype %d", dt);
  MultiParams mp{};
  mp.dst = make_devarr(*dst, true);
  mp.nterms = nterms;
  mp.combine = combine;
  const int64_t n0 = mp.dst.n0, n1 = mp.dst.n1, n2 = mp.dst.n2;
  const int64_t tx = (n0 + td.tx - 1) / td.tx, ty = (n1 + td.ty - 1) / td.ty, tz = (n2 + td.tz - 1) / td.tz;
  if (tx * ty * tz > (int64_t(1) << 31) - 1) return fail(STH_ENOTSUP, "array too large for one launch");
  mp.tiles_x = static_cast<int32_t>(tx);
  mp.tiles_y = static_cast<int32_t>(ty);
  mp.tiles_z = static_cast<int32_t>(tz);
  mp.xcd_remap = 1;
  auto s = static_cast<hipStream_t>(stream);
  // update_boundary! on every Halo argument first (src/gatherstencil.jl:93-95)
  for (const Plan& pl : plans)
    if (pl.halo) {
      pl.halo_fn(pl.hp, s);
      STH_HIP(hipGetLastError());
    }
  // Large calls: one runtime-compiled streaming pass per stencil argument,
  // each folding its term into the running value (read back from dest after
  // the first pass) together with the centre terms up to the next stencil
  // argument.  Any pass that cannot be built leaves the call on the tiled
  // kernel below.
  bool done = false;
  if (!(flags & STH_GENERIC) && dst->ndim >= 2 && jit_wanted(n0 * n1 * n2) && n0 * n1 * n2 > 0) {
    std::vector<int32_t> folds;
    for (int32_t t = 0; t < nterms; ++t)
      if (!tbl[t].direct) folds.push_back(t);
    bool ok = !folds.empty();
    for (int32_t t : folds)
      ok = ok && std::max(tbl[t].rx, std::max(tbl[t].ry, tbl[t].rz)) <= kJitMaxReach && tbl[t].opk != STH_OP_LIFE;
    if (!ok) jit_note("multi-array call", "no stencil argument, a reach above 3 or the Life rule");
    const int es = dtype_size(dt);
    const int v = es == 1 ? 8 : 16 / es;
    // Two stencil arguments or more, none weighted: one kernel for the
    // call, every argument's fold and the centre terms combined in registers
    // (STH_NO_FUSE keeps one pass per stencil argument).
    bool fused_ok = ok && folds.size() >= 2 && !(flags & STH_NO_FUSE);
    for (int32_t t : folds) fused_ok = fused_ok && tbl[t].opk != STH_OP_KERNELPRODUCT;
    if (fused_ok) {
      int ky = 4;
      std::vector<std::vector<int8_t>> o8s;
      std::vector<JitFold> jf;
      std::string items;
      for (int32_t u = 0; u < nterms; ++u) items.push_back(tbl[u].direct ? (tbl[u].scaled ? 'C' : 'c')
                                                                          : (tbl[u].scaled ? 'F' : 'f'));
      for (int32_t t : folds) {
        const MultiTerm& q = tbl[t];
        if (q.rz == 0 && std::max(q.rx, q.ry) > 2) ky = 2;
        std::vector<int8_t> o8(3 * static_cast<size_t>(q.L));
        for (int k = 0; k < q.L; ++k) {
          o8[3 * k] = static_cast<int8_t>(q.offs[k] & 0xff);
          o8[3 * k + 1] = static_cast<int8_t>((q.offs[k] >> 8) & 0xff);
          o8[3 * k + 2] = static_cast<int8_t>((q.offs[k] >> 16) & 0xff);
        }
        o8s.push_back(std::move(o8));
      }
      for (size_t f = 0; f < folds.size(); ++f) {
        const MultiTerm& q = tbl[folds[f]];
        jf.push_back(JitFold{q.opk, q.L, q.rx, q.ry, q.rz, o8s[f].data()});
      }
      const int lsh = lpr_shift_for(n0, v);
      const int64_t lpr = int64_t(1) << lsh, rpw = 64 / lpr;
      const int64_t jtx = (n0 + lpr * v - 1) / (lpr * v), jty = (n1 + 4 * ky * rpw - 1) / (4 * ky * rpw);
      std::string why;
      const JitKernel* k = nullptr;
      if (jtx * jty * n2 <= (int64_t(1) << 31) - 1 && n2 <= (int64_t(1) << 30))
        k = jit_multi_fused(dt, v, ky, jf, items, combine == STH_COMBINE_MUL, &why);
      else
        why = "array too large for one launch";
      if (k) {
        StreamMultiParams sp;
        std::memset(static_cast<void*>(&sp), 0, sizeof sp);
        int nc = 0;
        for (int32_t u = 0; u < nterms; ++u) {
          sp.coef[u] = tbl[u].coef;
          if (tbl[u].direct) {
            sp.cbase[nc] = tbl[u].src.base;
            sp.cs1[nc] = tbl[u].src.s1;
            sp.cs2[nc] = tbl[u].src.s2;
            ++nc;
          }
        }
        for (size_t f = 0; f < folds.size(); ++f) {
          const MapParams& pp = plans[folds[f]].p;
          StreamTerm& st = sp.term[f];
          st.src = pp.src;
          st.dst = pp.dst;
          st.bc = pp.bc;
          st.read_halo = pp.read_halo;
          st.lpr_shift = lsh;
          st.padval = pp.padval;
          st.op = pp.op;
        }
        sp.tiles_x = static_cast<int32_t>(jtx);
        sp.tiles_y = static_cast<int32_t>(jty);
        sp.tiles_z = static_cast<int32_t>(n2);
        sp.xcd_remap = 1;
        void* args[] = {&sp};
        STH_HIP(hipModuleLaunchKernel(k->fn, static_cast<unsigned>(jtx * jty * n2), 1, 1, 256, 1, 1, 0, s, args,
                                      nullptr));
        done = true;
      } else {
        jit_note("fused multi-array kernel", why);
      }
    }
    struct Pass {
      const JitKernel* k;
      MapParams p;
      MultiEpiParams e;
    };
    std::vector<Pass> passes;
    for (size_t f = 0; ok && !done && f < folds.size(); ++f) {
      const int32_t t = folds[f];
      const int32_t lo = f == 0 ? 0 : t, hi = f + 1 < folds.size() ? folds[f + 1] : nterms;
      const MultiTerm& q = tbl[t];
      const int rmax = std::max(q.rx, std::max(q.ry, q.rz));
      const int ky = (q.rz > 0 || rmax <= 2) ? 4 : 2;
      Pass ps;
      std::memset(static_cast<void*>(&ps.e), 0, sizeof ps.e);
      std::string items;
      int nc = 0;
      for (int32_t u = lo; u < hi; ++u) {
        if (u == t) {
          items.push_back(tbl[u].scaled ? 'F' : 'f');
          ps.e.coef[0] = tbl[u].coef;
        } else {
          items.push_back(tbl[u].scaled ? 'C' : 'c');
          ps.e.base[nc] = tbl[u].src.base;
          ps.e.s1[nc] = tbl[u].src.s1;
          ps.e.s2[nc] = tbl[u].src.s2;
          ps.e.coef[1 + nc] = tbl[u].coef;
          ++nc;
        }
      }
      ps.e.acc = mp.dst.base;
      ps.e.acc_s1 = mp.dst.s1;
      ps.e.acc_s2 = mp.dst.s2;
      std::vector<int8_t> o8(3 * static_cast<size_t>(q.L));
      for (int k = 0; k < q.L; ++k) {
        o8[3 * k] = static_cast<int8_t>(q.offs[k] & 0xff);
        o8[3 * k + 1] = static_cast<int8_t>((q.offs[k] >> 8) & 0xff);
        o8[3 * k + 2] = static_cast<int8_t>((q.offs[k] >> 16) & 0xff);
      }
      std::string why;
      ps.k = jit_multi_pass(dt, q.opk, q.L, o8.data(), q.rx, q.ry, q.rz, v, ky, f > 0, items, combine == STH_COMBINE_MUL,
                            &why);
      if (!ps.k) {
        jit_note("multi-array pass", why);
        ok = false;
        break;
      }
      ps.p = plans[t].p;
      ps.p.lpr_shift = lpr_shift_for(n0, v);
      const int64_t lpr = int64_t(1) << ps.p.lpr_shift, rpw = 64 / lpr;
      const int64_t jtx = (n0 + lpr * v - 1) / (lpr * v), jty = (n1 + 4 * ky * rpw - 1) / (4 * ky * rpw);
      if (jtx * jty * n2 > (int64_t(1) << 31) - 1 || n2 > (int64_t(1) << 30)) {
        ok = false;
        break;
      }
      ps.p.tiles_x = static_cast<int32_t>(jtx);
      ps.p.tiles_y = static_cast<int32_t>(jty);
      ps.p.tiles_z = static_cast<int32_t>(n2);
      passes.push_back(ps);
    }
    if (ok && !done) {
      for (const Pass& ps : passes) STH_HIP(jit_launch_multi(ps.k->fn, ps.p, ps.e, s));
      done = true;
    }
  }
  if (!done && n0 * n1 * n2 > 0) {
    rc = with_device_table(ctx, tbl.data(), tbl.size() * sizeof(MultiTerm), s, [&](const void* dev) {
      ke->fn(mp, static_cast<const MultiTerm*>(dev), lds, s);
    });
    if (rc) return rc;
  }
  if (!(flags & STH_ASYNC)) STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

int sth_scatter(sth_ctx* ctx, const sth_term* source, const sth_term* aux, const sth_array* dst,
                const sth_scatter_rule* rule, uint32_t flags, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (!source || !rule) return fail(STH_EINVAL, "source or rule is NULL");
  if ((rc = check_array(dst, "dest"))) return rc;
  const int dt = dst->dtype;
  if (dt != STH_F32 && dt != STH_F64) return fail(STH_ENOTSUP, "scatter over dtype %d is not accelerated", dt);
  if (dst->ndim != 2 || source->src.ndim != 2)
    return fail(STH_ENOTSUP, "scatterstencil! is defined for 2-D arrays (src/scatterstencil.jl:51)");
  if (rule->kind < STH_SCATTER_CONST || rule->kind > STH_SCATTER_DOWNHILL_FLOW)
    return fail(STH_EINVAL, "unknown scatter rule %d", rule->kind);
  if (rule->combine != STH_COMBINE_ADD && rule->combine != STH_COMBINE_MAX && rule->combine != STH_COMBINE_MIN)
    return fail(STH_ENOTSUP, "scatter op %d is not accelerated", rule->combine);
  const bool needs_aux = rule->kind >= STH_SCATTER_DOWNHILL;
  if (needs_aux && !aux) return fail(STH_EINVAL, "the downhill rules need the second array");
  if (dt == STH_F32 && static_cast<double>(static_cast<float>(rule->coef)) != rule->coef)
    return fail(STH_ENOTSUP, "coefficient %.17g is not a Float32 (the values would promote)", rule->coef);
  const sth_stencil& st = source->stencil;
  if (st.length < 1 || st.length > kScatterMaxL || !st.offsets || st.ndim != 2)
    return fail(STH_ENOTSUP, "scatter stencil with %d offsets in %d dims", st.length, st.ndim);
  const int R = st.radius;
  if (R < 0 || R > kScatterMaxR) return fail(STH_ENOTSUP, "scatter radius %d (max %d)", R, kScatterMaxR);
  if (dst->size[0] >= (int64_t(1) << 30) || dst->size[1] >= (int64_t(1) << 30))
    return fail(STH_ENOTSUP, "scatter axis longer than 2^30 cells");
  for (int k = 0; k < st.length; ++k)
    for (int d = 0; d < 2; ++d)
      if (std::abs(st.offsets[2 * k + d]) > R)
        return fail(STH_ENOTSUP, "stencil reach beyond its radius %d: the reference's column passes would overlap", R);
  // validate the arguments as gathers (sizes, boundary, padding, radius)
  sth_op sum_op{};
  sum_op.kind = STH_OP_SUM;
  Plan ps, pa;
  if ((rc = make_plan(&source->src, dst, &st, source->boundary, source->padval, source->padding, &sum_op,
                      flags | STH_GENERIC | STH_NO_FUSE, &ps)))
    return rc;
  if (source->src.dtype != dt) return fail(STH_ENOTSUP, "source and dest element types differ");
  if (needs_aux) {
    if (aux->src.dtype != dt) return fail(STH_ENOTSUP, "the second array's element type differs");
    if (aux->stencil.length != st.length || !aux->stencil.offsets ||
        std::memcmp(aux->stencil.offsets, st.offsets, sizeof(int32_t) * 2 * st.length) != 0)
      return fail(STH_ENOTSUP, "the second array's stencil differs from the source's");
    if ((rc = make_plan(&aux->src, dst, &aux->stencil, aux->boundary, aux->padval, aux->padding, &sum_op,
                        flags | STH_GENERIC | STH_NO_FUSE, &pa)))
      return rc;
  }
  ScatterParams p;
  std::memset(static_cast<void*>(&p), 0, sizeof p);
  p.src = ps.p.src;
  p.dst = make_devarr(*dst, true);
  p.bc = source->boundary;
  if (needs_aux) {
    p.aux = pa.p.src;
    p.aux_bc = pa.p.bc;
    p.aux_read_halo = pa.p.read_halo;
    p.aux_padval = pa.p.padval;
    p.has_aux = 1;
  }
  p.kind = rule->kind;
  p.combine = rule->combine;
  p.zero_dest = (flags & STH_SCATTER_ZERO_DEST) ? 1 : 0;
  p.R = R;
  p.S = 2 * R + 1;
  p.L = st.length;
  p.coef = rule->coef;
  // device table: the offsets in fold order; then per dest-column residue r
  // the order a cell receives them from the cells d - offset: the sender
  // sits in source column dj - b (b = the offset's dim-2 part), whose pass
  // is its residue modulo S; within a pass rows ascend (dim-1 part
  // descending), then k.  The order depends on d only through dj mod S.
  // Listed once as packed offsets and once as LDS deltas of the senders.
  const int L = st.length;
  auto pack = [&](int k) {
    return static_cast<int32_t>((static_cast<uint32_t>(st.offsets[2 * k]) & 0xffffu) |
                                (static_cast<uint32_t>(st.offsets[2 * k + 1]) << 16));
  };
  std::vector<int32_t> tbl(static_cast<size_t>(scatter_table_len(L, p.S)));
  for (int k = 0; k < L; ++k) tbl[k] = pack(k);
  const int W = kScatterTX + 2 * R;
  for (int r = 0; r < p.S; ++r) {
    std::vector<int> ks(static_cast<size_t>(L));
    for (int k = 0; k < L; ++k) ks[k] = k;
    auto pass = [&](int k) { return ((r - st.offsets[2 * k + 1]) % p.S + p.S) % p.S; };
    std::stable_sort(ks.begin(), ks.end(), [&](int x, int y) {
      if (pass(x) != pass(y)) return pass(x) < pass(y);
      return st.offsets[2 * x] > st.offsets[2 * y];
    });
    for (int t = 0; t < L; ++t) {
      tbl[static_cast<size_t>(L) * (1 + r) + t] = pack(ks[t]);
      tbl[static_cast<size_t>(L) * (1 + p.S + r) + t] = -(st.offsets[2 * ks[t] + 1] * W + st.offsets[2 * ks[t]]);
    }
  }
  const int64_t tx = (p.dst.n0 + kScatterTX - 1) / kScatterTX, ty = (p.dst.n1 + kScatterTY - 1) / kScatterTY;
  // the coordinates of each axis with a wrapped or reflected preimage
  // (Wrap: within R - 1 of the low edge or R of the high one; Reflect:
  // 1..R from either edge), as the kernel's scatter_preimages finds them
  auto band = [&](int64_t n, int32_t* out) {
    int cnt = 0;
    for (int64_t x = 0; x < n; ++x) {
      if (x >= 2 * R + 2 && x < n - 2 * R - 2) x = n - 2 * R - 2;
      bool nontrivial = false;
      if (p.bc == STH_WRAP) nontrivial = x - n >= -R || x + n <= n - 1 + R;
      else if (p.bc == STH_REFLECT) nontrivial = (x > 0 && x <= R) || (x < n - 1 && 2 * n - 2 - x <= n - 1 + R);
      if (nontrivial) out[cnt++] = static_cast<int32_t>(x);
    }
    return cnt;
  };
  p.nbi = band(p.dst.n0, p.band_i);
  p.nbj = band(p.dst.n1, p.band_j);
  const int64_t gen = p.nbj * p.dst.n0 + p.nbi * (p.dst.n1 - p.nbj);
  const int64_t gen_blocks = (((gen + 255) / 256) + 7) & ~int64_t(7);  // a multiple of 8: tiles start on XCD 0
  if (gen > (int64_t(1) << 30) || tx * ty + gen_blocks > (int64_t(1) << 31) - 1)
    return fail(STH_ENOTSUP, "array too large for one launch");
  p.gen_cells = static_cast<int32_t>(gen);
  p.gen_blocks = static_cast<int32_t>(gen_blocks);
  p.tiles_x = static_cast<int32_t>(tx);
  p.tiles_y = static_cast<int32_t>(ty);
  auto s = static_cast<hipStream_t>(stream);
  // update_boundary! on the source and on a StencilArray argument first
  // (src/scatterstencil.jl:41-43)
  if (ps.halo) ps.halo_fn(ps.hp, s);
  if (needs_aux && pa.halo) pa.halo_fn(pa.hp, s);
  STH_HIP(hipGetLastError());
  if (tx * ty > 0) {
    // large calls: the interior order compiled in (immediate LDS offsets)
    const JitKernel* jk = nullptr;
    if (!(flags & STH_GENERIC) && jit_wanted(p.dst.n0 * p.dst.n1)) {
      std::string why;
      jk = jit_scatter_kernel(dt, p.kind, p.combine, p.S, L, tbl.data() + static_cast<size_t>(L) * (1 + p.S), &why);
      if (!jk) jit_note("scatter kernel", why);
    }
    hipError_t lerr = hipSuccess;
    rc = with_device_table(ctx, tbl.data(), tbl.size() * sizeof(int32_t), s, [&](const void* dev) {
      if (jk) {
        const int32_t* dtbl = static_cast<const int32_t*>(dev);
        void* args[] = {&p, &dtbl};
        const unsigned nb = static_cast<unsigned>(p.gen_blocks) + static_cast<unsigned>(p.tiles_x) * p.tiles_y;
        const size_t lds = scatter_lds_bytes(p.kind, p.R, p.L, dtype_size(dt));
        lerr = hipModuleLaunchKernel(jk->fn, nb, 1, 1, 256, 1, 1, static_cast<unsigned>(lds), s, args, nullptr);
      } else if (dt == STH_F32) {
        launch_scatter<float>(p, static_cast<const int32_t*>(dev), s);
      } else {
        launch_scatter<double>(p, static_cast<const int32_t*>(dev), s);
      }
    });
    if (rc) return rc;
    STH_HIP(lerr);
  }
  if (!(flags & STH_ASYNC)) STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

int sth_autotune(sth_ctx* ctx, const sth_array* src, const sth_array* dst, const sth_stencil* st,
                 int32_t boundary, const void* padval, int32_t padding, const sth_op* op, void* stream,
                 char* chosen, size_t chosen_len) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  Plan pl;
  std::vector<Variant> cands;
  std::string group;
  if ((rc = make_plan(src, dst, st, boundary, padval, padding, op, 0, &pl, &cands, &group))) return rc;
  auto s = static_cast<hipStream_t>(stream);
  int idx = 0;
  if ((rc = tune_pick(ctx, pl, cands, group, 0, s, true, false, &idx))) return rc;
  apply_variant(pl, cands[idx]);
  if ((rc = run_plan(pl, s))) return rc;
  STH_HIP(hipStreamSynchronize(s));
  if (chosen && chosen_len) {
    std::snprintf(chosen, chosen_len, "%s", cands[idx].name);
  }
  return STH_OK;
  STH_GUARD_END
}

int sth_set_variant(sth_ctx* ctx, const char* pattern) {
  STH_GUARD_BEGIN
  if (!ctx) return fail(STH_EINVAL, "context is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->forced = pattern ? pattern : "";
  return STH_OK;
  STH_GUARD_END
}

int sth_tuning_table(sth_ctx* ctx, char* buf, size_t len, size_t* needed) {
  STH_GUARD_BEGIN
  if (!ctx) return fail(STH_EINVAL, "context is NULL");
  std::string out;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (const auto& kv : ctx->tuned) out += kv.first + "\t" + kv.second + "\n";
  }
  if (needed) *needed = out.size() + 1;
  if (buf && len) std::snprintf(buf, len, "%s", out.c_str());
  return STH_OK;
  STH_GUARD_END
}

int sth_iterate(sth_ctx* ctx, const sth_array* a, const sth_array* b, const sth_stencil* st,
                int32_t boundary, const void* padval, int32_t padding, const sth_op* op, int64_t steps,
                uint32_t flags, int32_t* current, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (steps < 0) return fail(STH_EINVAL, "negative step count");
  if (!current) return fail(STH_EINVAL, "current is NULL");
  if (!a || !b) return fail(STH_EINVAL, "buffer descriptor is NULL");
  if (a->dtype != b->dtype || a->halo != b->halo)
    return fail(STH_EINVAL, "source and dest buffers differ in type or halo");
  if (op && op->kind == STH_OP_MEAN && !is_float(a->dtype))
    return fail(STH_EINVAL, "mapstencil! cannot store a Float64 mean into an integer buffer");
  Plan ab, ba;
  std::vector<Variant> cab, cba;
  std::string group;
  if ((rc = make_plan(a, b, st, boundary, padval, padding, op, flags, &ab, &cab, &group))) return rc;
  if ((rc = make_plan(b, a, st, boundary, padval, padding, op, flags, &ba, &cba))) return rc;
  auto s = static_cast<hipStream_t>(stream);
  if (steps > 0) {
    // choose the variant on the a -> b step (idempotent: it only writes b)
    int idx = 0;
    if ((rc = tune_pick(ctx, ab, cab, group, flags, s, false, false, &idx))) return rc;
    apply_variant(ab, cab[idx]);
    apply_variant(ba, cba[idx]);
  }
  *current = static_cast<int32_t>(steps % 2);
  if (steps == 0) return STH_OK;
  // The first step runs the normalising kernel (it may read data this
  // library did not write); steps 1.. use the steady-state kernels.  With
  // graphs it is recorded at the head of the first graph launched.
  const int64_t rest = steps - 1;
  // Temporal blocking: G = fused_gens steps per launch.  A fused launch
  // writes only the newest generation, so the buffer it reads is left G - 1
  // generations behind the step-by-step order; fused launches come in pairs
  // (a unit: b -> a, a -> b, 2 G steps, the write buffer alternating as in
  // the step-by-step order), and every iteration ends with at least two
  // single steps, which leave both buffers and their halos as the
  // step-by-step reference does.
  const int64_t G = ab.fused_gens;
  const bool fuse = !(flags & STH_NO_FUSE) && (ab.fn_fused || ab.jit_fused) && (ba.fn_fused || ba.jit_fused) &&
                    G >= 2 && rest >= 2 * G + 2;
  const int64_t units = fuse ? (rest - 2) / (2 * G) : 0;  // fused pairs
  const int64_t tail = rest - 2 * G * units;              // single steps (2 .. 2G + 1 when fused)
  auto enqueue = [&](hipStream_t st, bool first, int64_t nunits, int64_t nplain) -> int {
    int r2 = first ? run_plan(ab, st) : STH_OK;
    for (int64_t u = 0; u < nunits && r2 == STH_OK; ++u) {
      r2 = run_fused(ba, st);
      if (r2 == STH_OK) r2 = run_fused(ab, st);
    }
    for (int64_t i = 0; i < nplain && r2 == STH_OK; ++i) r2 = run_plan(i % 2 ? ab : ba, st, true);
    return r2;
  };
  // Calls of a few launches go straight to the stream even with STH_GRAPH:
  // a graph launch costs the host about 27 us against about 16 us for six
  // direct launches, and the device is the bound either way (Life 8192^2,
  // 20 steps: 152 us direct, 159 us replayed).
  constexpr int64_t kGraphMinLaunches = 17;
  const int64_t launches = 1 + 2 * units + tail;
  if (!(flags & STH_GRAPH) || launches < kGraphMinLaunches) {
    if ((rc = enqueue(s, true, units, tail))) return rc;
  } else {
    // Graphs of steady steps (starting from buffer b) are captured once per
    // parameter set and replayed.  The lookup and the launches happen under
    // ctx->mu, so a cache eviction on another thread cannot destroy an
    // executable between the two.
    auto launch_graph = [&](const char* tag, bool first, int64_t nunits, int64_t nplain, int64_t times) -> int {
      std::string key(reinterpret_cast<const char*>(&ab), sizeof(Plan));
      key.append(reinterpret_cast<const char*>(&ba), sizeof(Plan));
      key.append(tag);
      key.push_back(first ? '1' : '0');
      key.append(reinterpret_cast<const char*>(&nunits), sizeof nunits);
      key.append(reinterpret_cast<const char*>(&nplain), sizeof nplain);
      std::lock_guard<std::mutex> lk(ctx->mu);
      auto it = ctx->graphs.find(key);
      if (it != ctx->graphs.end()) {
        for (int64_t i = 0; i < times; ++i) STH_HIP(hipGraphLaunch(it->second, s));
        return STH_OK;
      }
      STH_HIP(hipStreamBeginCapture(ctx->capture, hipStreamCaptureModeThreadLocal));
      const int r2 = enqueue(ctx->capture, first, nunits, nplain);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(ctx->capture, &g);
      if (r2) return r2;
      if (e != hipSuccess) return fail(STH_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
      hipGraphExec_t x = nullptr;
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(STH_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
      if (ctx->graphs.size() >= kMaxGraphs) {
        // bound the cache (a long-running caller iterating ever new buffers
        // would grow it without end): wait for the device, then drop it all
        if (int r3 = clear_graphs(ctx)) return r3;
      }
      ctx->graphs.emplace(key, x);
      for (int64_t i = 0; i < times; ++i) STH_HIP(hipGraphLaunch(x, s));
      return STH_OK;
    };
    // the first graph launched carries the first step; a call's graphs
    // (head, repeated body, tail) are one graph when the call is short
    if (fuse) {
      constexpr int64_t FU = 16;  // units per replayed graph
      const int64_t reps = units / FU, remu = units % FU;
      if (reps > 0) {
        if ((rc = launch_graph("f", true, FU, 0, 1))) return rc;
        if (reps > 1 && (rc = launch_graph("f", false, FU, 0, reps - 1))) return rc;
      }
      if ((rc = launch_graph("f", reps == 0, remu, tail, 1))) return rc;
    } else {
      const int64_t U = rest >= 64 ? 16 : std::max<int64_t>(1, rest / 2);
      const int64_t G = 2 * U;
      const int64_t reps = rest / G, rem = rest % G;
      if (reps > 0) {
        if ((rc = launch_graph("p", true, 0, G, 1))) return rc;
        if (reps > 1 && (rc = launch_graph("p", false, 0, G, reps - 1))) return rc;
      }
      if ((reps == 0 || rem > 0) && (rc = launch_graph("p", reps == 0, 0, rem, 1))) return rc;
    }
  }
  STH_HIP(hipGetLastError());
  if (!(flags & STH_ASYNC)) STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

int sth_map_batch(sth_ctx* ctx, int32_t njobs, const sth_job* jobs, uint32_t flags, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (njobs < 0 || (njobs > 0 && !jobs)) return fail(STH_EINVAL, "bad job list");
  auto s = static_cast<hipStream_t>(stream);
  const bool serial_only = (flags & STH_BATCH_SERIAL) != 0;
  std::vector<Plan> plans(static_cast<size_t>(njobs));
  std::vector<const Plan*> ptrs(static_cast<size_t>(njobs));
  for (int32_t i = 0; i < njobs; ++i) {
    const sth_job& jb = jobs[i];
    std::vector<Variant> cands;
    std::string group;
    if ((rc = make_plan(&jb.src, &jb.dst, &jb.stencil, jb.boundary, jb.padval, jb.padding, &jb.op,
                        flags & ~STH_GENERIC, &plans[i], &cands, &group))) {
      std::string m = g_err;
      return fail(rc, "job %d: %s", i, m.c_str());
    }
    // one launch per job: the tuned single-array kernel (tuning may time it
    // here); inside a batched call only decisions already made are used
    int idx = 0;
    if ((rc = tune_pick(ctx, plans[i], cands, group, flags, s, false, !serial_only, &idx))) return rc;
    apply_variant(plans[i], cands[idx]);
    ptrs[i] = &plans[i];
  }
  if ((rc = launch_batched(ctx, ptrs.data(), njobs, serial_only, s))) return rc;
  if (!(flags & STH_ASYNC)) STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

int sth_iterate_batch(sth_ctx* ctx, int32_t njobs, const sth_job* jobs, const int64_t* steps, uint32_t flags,
                      int32_t* current, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (njobs < 0 || (njobs > 0 && (!jobs || !steps || !current))) return fail(STH_EINVAL, "bad job list");
  auto s = static_cast<hipStream_t>(stream);
  const bool serial_only = (flags & STH_BATCH_SERIAL) != 0;
  // each array's two steps: source -> dest (even rounds), dest -> source
  std::vector<Plan> fwd(static_cast<size_t>(njobs)), bwd(static_cast<size_t>(njobs));
  std::vector<int32_t> order;
  int64_t maxsteps = 0;
  for (int32_t i = 0; i < njobs; ++i) {
    const sth_job& jb = jobs[i];
    if (steps[i] < 0) return fail(STH_EINVAL, "job %d: negative step count", i);
    if (jb.src.dtype != jb.dst.dtype || jb.src.halo != jb.dst.halo)
      return fail(STH_EINVAL, "job %d: source and dest buffers differ in type or halo", i);
    if (jb.op.kind == STH_OP_MEAN && !is_float(jb.src.dtype))
      return fail(STH_EINVAL, "job %d: mapstencil! cannot store a Float64 mean into an integer buffer", i);
    std::vector<Variant> cf, cb;
    std::string group;
    if ((rc = make_plan(&jb.src, &jb.dst, &jb.stencil, jb.boundary, jb.padval, jb.padding, &jb.op,
                        flags & ~STH_GENERIC, &fwd[i], &cf, &group)) ||
        (rc = make_plan(&jb.dst, &jb.src, &jb.stencil, jb.boundary, jb.padval, jb.padding, &jb.op,
                        flags & ~STH_GENERIC, &bwd[i], &cb))) {
      std::string m = g_err;
      return fail(rc, "job %d: %s", i, m.c_str());
    }
    int idx = 0;
    if ((rc = tune_pick(ctx, fwd[i], cf, group, flags, s, false, true, &idx))) return rc;
    apply_variant(fwd[i], cf[idx]);
    apply_variant(bwd[i], cb[idx]);
    current[i] = static_cast<int32_t>(steps[i] % 2);
    if (steps[i] > 0) order.push_back(i);
    maxsteps = std::max(maxsteps, steps[i]);
  }
  // round r advances every array with more than r steps: with the arrays
  // ordered by step count (descending) that is a prefix of `order`, the
  // same prefix until its shortest member is done
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return steps[x] > steps[y]; });
  auto enqueue = [&](hipStream_t st) -> int {
    std::vector<const Plan*> even, odd;
    size_t m = order.size();
    for (int64_t r = 0; r < maxsteps;) {
      while (m > 0 && steps[order[m - 1]] <= r) --m;
      even.clear();
      odd.clear();
      for (size_t j = 0; j < m; ++j) {
        even.push_back(&fwd[order[j]]);
        odd.push_back(&bwd[order[j]]);
      }
      for (const int64_t r_end = steps[order[m - 1]]; r < r_end; ++r)
        if (int r2 = launch_batched(ctx, (r % 2 ? odd : even).data(), static_cast<int32_t>(m), serial_only, st))
          return r2;
    }
    return STH_OK;
  };
  if (!(flags & STH_GRAPH) || maxsteps == 0) {
    if ((rc = enqueue(s))) return rc;
  } else {
    // the whole schedule as one graph, captured once per (arrays, functions,
    // step counts) and replayed; its job tables are device memory it owns
    std::string key("B", 1);
    key.push_back(serial_only ? 's' : 'b');
    for (int32_t i : order) {
      key.append(reinterpret_cast<const char*>(&fwd[i]), sizeof(Plan));
      key.append(reinterpret_cast<const char*>(&bwd[i]), sizeof(Plan));
      key.append(reinterpret_cast<const char*>(&steps[i]), sizeof(int64_t));
    }
    bool done = false;
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      auto it = ctx->graphs.find(key);
      if (it != ctx->graphs.end()) {
        STH_HIP(hipGraphLaunch(it->second, s));
        done = true;
      }
    }
    if (!done) {
      std::vector<void*> tables;
      ctx->capture_tables = &tables;
      hipError_t e = hipStreamBeginCapture(ctx->capture, hipStreamCaptureModeRelaxed);
      const int r2 = e == hipSuccess ? enqueue(ctx->capture) : STH_OK;
      hipGraph_t g = nullptr;
      if (e == hipSuccess) e = hipStreamEndCapture(ctx->capture, &g);
      ctx->capture_tables = nullptr;
      hipGraphExec_t x = nullptr;
      if (r2 == STH_OK && e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (r2 || e != hipSuccess) {
        for (void* d : tables) (void)hipFree(d);
        if (r2) return r2;
        return fail(STH_EDEVICE, "batched iteration graph: %s", hipGetErrorString(e));
      }
      std::lock_guard<std::mutex> lk(ctx->mu);
      if (ctx->graphs.size() >= kMaxGraphs && (rc = clear_graphs(ctx))) return rc;
      ctx->graphs.emplace(key, x);
      ctx->graph_tables.emplace(key, std::move(tables));
      STH_HIP(hipGraphLaunch(x, s));
    }
  }
  STH_HIP(hipGetLastError());
  if (!(flags & STH_ASYNC)) STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

int sth_jit_selftest(char* log, size_t cap) {
  STH_GUARD_BEGIN
  // representative specialisations of every generator: streaming kernels
  // over each storage type, multi-array passes over each of their types and
  // combine forms; compiled only (no device is needed or touched)
  static const int8_t shape[6][3] = {{-2, 0, 0}, {0, -1, 0}, {0, 0, 0}, {1, 1, 0}, {2, -2, 0}, {-1, 2, 0}};
  std::string out;
  int failures = 0;
  auto one = [&](const std::string& src, const std::string& expr, const std::string& what) {
    std::string code, lowered, lg;
    if (!jit_compile(src, expr, &code, &lowered, &lg)) {
      ++failures;
      out += what + ": " + lg.substr(0, 600) + "\n";
    }
  };
  const int dts[] = {STH_BOOL, STH_U8, STH_I16, STH_I32, STH_I64, STH_F32, STH_F64, STH_BF16};
  for (int dt : dts) {
    const std::string targs = std::string(jit_type_name(dt)) + ", " + std::to_string(STH_OP_MAX) +
                              ", sth::jit::Shape, 4, 4";
    one("#include \"stream_kernel.hpp\"\n" + jit_shape_source(6, &shape[0][0], 2, 2, 0) +
            "namespace sth {\ntemplate __global__ void stream_kernel<" + targs + ">(MapParams);\n}\n",
        "&sth::stream_kernel<" + targs + ">", "stream " + std::string(jit_type_name(dt)));
  }
  struct M {
    int dt, op;
    const char* items;
    bool acc, mul;
  };
  const M ms[] = {{STH_F32, STH_OP_LAPLACIAN, "fC", false, false}, {STH_F64, STH_OP_SUM, "Ccfc", false, true},
                  {STH_U8, STH_OP_MAX, "cf", true, true},           {STH_I32, STH_OP_SUM, "fc", false, false},
                  {STH_I64, STH_OP_MIN, "f", true, false}};
  for (const M& m : ms) {
    std::string src, expr;
    const int v = m.dt == STH_U8 ? 8 : 16 / dtype_size(m.dt);
    jit_multi_source(m.dt, m.op, 6, &shape[0][0], 2, 2, 0, v, 4, m.acc, m.items, m.mul, &src, &expr);
    one(src, expr, std::string("multi pass ") + jit_type_name(m.dt) + " " + m.items);
  }
  static const int8_t vn2[4][3] = {{0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}};
  static const int8_t vn3[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  {
    // multi-array kernels with every argument in one pass
    const std::vector<JitFold> two = {{STH_OP_LAPLACIAN, 4, 1, 1, 0, &vn2[0][0]}, {STH_OP_MAX, 6, 2, 2, 0, &shape[0][0]}};
    const std::vector<JitFold> three = {{STH_OP_SUM, 6, 1, 1, 1, &vn3[0][0]}, {STH_OP_MIN, 4, 1, 1, 0, &vn2[0][0]},
                                        {STH_OP_MEAN, 6, 1, 1, 1, &vn3[0][0]}};
    struct F {
      int dt;
      const std::vector<JitFold>* folds;
      const char* items;
      bool mul;
    };
    const F fs[] = {{STH_F32, &two, "CfFc", false}, {STH_F64, &three, "fcfF", true}, {STH_U8, &two, "fcf", true},
                    {STH_I64, &three, "cfff", false}};
    for (const F& m : fs) {
      std::string src, expr;
      const int v = m.dt == STH_U8 ? 8 : 16 / dtype_size(m.dt);
      jit_multi_fused_source(m.dt, v, 4, *m.folds, m.items, m.mul, &src, &expr);
      one(src, expr, std::string("fused multi ") + jit_type_name(m.dt) + " " + m.items);
    }
  }
  const int fdt[][3] = {{STH_I32, STH_OP_MAX, 2}, {STH_I64, STH_OP_SUM, 2}, {STH_F32, STH_OP_SUM, 2},
                        {STH_I16, STH_OP_MIN, 3}, {STH_I64, STH_OP_SUM, 3}, {STH_F64, STH_OP_MAX, 3}};
  for (const auto& c : fdt) {
    std::string src, expr;
    jit_fused_source(c[0], c[1], c[2] == 3 ? 6 : 4, c[2] == 3 ? &vn3[0][0] : &vn2[0][0], 1, 1, c[2] == 3 ? 1 : 0, c[2],
                     STH_WRAP, &src, &expr);
    one(src, expr, std::string("two-step ") + std::to_string(c[2]) + "-D " + jit_type_name(c[0]));
  }
  {
    // Moore(1) in a 66-wide box, residues in some order
    int32_t deltas[3 * 8];
    for (int i = 0; i < 24; ++i) deltas[i] = ((i % 8) - 4) * 66 + (i % 3) - 1;
    for (int kind : {STH_SCATTER_CONST, STH_SCATTER_DOWNHILL_FLOW}) {
      std::string src, expr;
      jit_scatter_source(STH_F64, kind, STH_COMBINE_ADD, 3, 8, deltas, &src, &expr);
      one(src, expr, "scatter f64 rule " + std::to_string(kind));
    }
  }
  if (log && cap) {
    const size_t n = std::min(cap - 1, out.size());
    std::memcpy(log, out.data(), n);
    log[n] = '\0';
  }
  return failures;
  STH_GUARD_END
}

int sth_map_batch_static(sth_ctx* ctx, int32_t njobs, const sth_job* jobs, uint32_t flags, void* stream) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (njobs < 0 || (njobs > 0 && !jobs)) return fail(STH_EINVAL, "bad job list");
  if (njobs > kStaticBatchJobs)
    return fail(STH_EINVAL, "%d jobs exceed STH_STATIC_BATCH_MAX_JOBS (%d)", njobs, kStaticBatchJobs);
  if (!ctx->scratch) return fail(STH_EDEVICE, "the fixed job table of this device is not mapped");
  auto s = static_cast<hipStream_t>(stream);
  const bool serial_only = (flags & STH_BATCH_SERIAL) != 0;
  std::vector<Plan> plans(static_cast<size_t>(njobs));
  std::vector<const Plan*> ptrs(static_cast<size_t>(njobs));
  for (int32_t i = 0; i < njobs; ++i) {
    const sth_job& jb = jobs[i];
    std::vector<Variant> cands;
    std::string group;
    if ((rc = make_plan(&jb.src, &jb.dst, &jb.stencil, jb.boundary, jb.padval, jb.padding, &jb.op,
                        flags & ~STH_GENERIC, &plans[i], &cands, &group))) {
      std::string m = g_err;
      return fail(rc, "job %d: %s", i, m.c_str());
    }
    // decisions the tuner already made only: timing would create events
    int idx = 0;
    if ((rc = tune_pick(ctx, plans[i], cands, group, flags, s, false, true, &idx))) return rc;
    apply_variant(plans[i], cands[idx]);
    ptrs[i] = &plans[i];
  }
  if ((rc = launch_batched_static(ctx, ptrs.data(), njobs, serial_only, s, ctx->scratch))) return rc;
  if (!(flags & STH_ASYNC)) STH_HIP(hipStreamSynchronize(s));
  return STH_OK;
  STH_GUARD_END
}

}  // extern "C"

#include "server.hpp"
