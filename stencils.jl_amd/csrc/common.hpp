// Device-side building blocks shared by every stencil kernel: element traits,
// Julia-exact reductions, boundary index mapping and the kernel parameter
// block.  Everything here compiles with -ffp-contract=off so that `a + b * c`
// rounds twice, as Julia does without @fastmath (src/stencils/kernel.jl:37-43).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "stencils_hip.h"

namespace sth {

// ------------------------------------------------------------ element types
// Storage types.  BF16 and F16 are kept as raw 16-bit patterns in memory and
// in LDS; arithmetic happens in the accumulator type.
struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };

__device__ __forceinline__ float bf16_to_float(uint16_t b) {
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}
// Round-to-nearest-even f32 -> bf16; NaN stays NaN (quiet bit forced).
__device__ __forceinline__ uint16_t float_to_bf16(float f) {
  uint32_t x = __float_as_uint(f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return static_cast<uint16_t>(x >> 16);
}
__device__ __forceinline__ float f16_to_float(uint16_t h) {
  _Float16 v = __builtin_bit_cast(_Float16, h);
  return static_cast<float>(v);
}
// Keeps a value materialised in a VGPR (an empty asm the optimiser cannot
// see through).  A host build of these headers for testing defines it away.
#ifndef STH_OPAQUE
#define STH_OPAQUE(x) __asm__("" : "+v"(x))
#endif

__device__ __forceinline__ uint16_t float_to_f16(float f) {
  // Without the barrier the compiler folds the conversion into the producing
  // FMA (v_fma_mix*_f16), which rounds the exact FMA result to f16 once
  // instead of rounding the f32 value, and differs at f16 ties.
  STH_OPAQUE(f);
  _Float16 v = static_cast<_Float16>(f);  // v_cvt_f16_f32: round to nearest even
  return __builtin_bit_cast(uint16_t, v);
}

// Julia `max` / `min` on IEEE floats (base/math.jl): NaN if either argument
// is NaN, and -0.0 orders below +0.0 -- the IEEE 754-2019 maximum/minimum
// operations, which gfx950 executes as v_maximum3_f32 / v_minimum3_f32 (f64:
// v_max_f64 plus a NaN select).  Which NaN comes out is not specified; the
// reference returns the NaN argument, so results agree up to "is a NaN".
template <typename F>
__device__ __forceinline__ F jl_max(F x, F y) {
  return __builtin_elementwise_maximum(x, y);
}
template <typename F>
__device__ __forceinline__ F jl_min(F x, F y) {
  return __builtin_elementwise_minimum(x, y);
}

// ------------------------------------------------------------- op kinds
enum OpKind : int {
  OP_SUM = STH_OP_SUM,
  OP_MEAN = STH_OP_MEAN,
  OP_MIN = STH_OP_MIN,
  OP_MAX = STH_OP_MAX,
  OP_KP = STH_OP_KERNELPRODUCT,
  OP_LIFE = STH_OP_LIFE,
  OP_LAP = STH_OP_LAPLACIAN
};

// Traits of a (storage in, storage out, op) triple.
//   S   - storage type read from memory and staged in LDS
//   O   - storage type written
//   A   - accumulator type of the fold
// load(): S -> A;  store(): A -> O.
template <typename S> struct Elem;
#define STH_INT_ELEM(T, CODE, SIGNED, UT)                       \
  template <> struct Elem<T> {                                  \
    using raw = T;                                              \
    using uacc = UT;                                            \
    static constexpr int code = CODE;                           \
    static constexpr bool is_float = false;                     \
    static constexpr bool is_signed = SIGNED;                   \
  };
STH_INT_ELEM(bool, STH_BOOL, false, uint64_t)
STH_INT_ELEM(uint8_t, STH_U8, false, uint32_t)
STH_INT_ELEM(int8_t, STH_I8, true, uint32_t)
STH_INT_ELEM(uint16_t, STH_U16, false, uint32_t)
STH_INT_ELEM(int16_t, STH_I16, true, uint32_t)
STH_INT_ELEM(uint32_t, STH_U32, false, uint32_t)
STH_INT_ELEM(int32_t, STH_I32, true, uint32_t)
STH_INT_ELEM(uint64_t, STH_U64, false, uint64_t)
STH_INT_ELEM(int64_t, STH_I64, true, uint64_t)
#undef STH_INT_ELEM
template <> struct Elem<float> {
  using raw = float;
  static constexpr int code = STH_F32;
  static constexpr bool is_float = true;
  static constexpr bool is_signed = true;
};
template <> struct Elem<double> {
  using raw = double;
  static constexpr int code = STH_F64;
  static constexpr bool is_float = true;
  static constexpr bool is_signed = true;
};
template <> struct Elem<bf16_t> {
  using raw = uint16_t;
  static constexpr int code = STH_BF16;
  static constexpr bool is_float = true;
  static constexpr bool is_signed = true;
};
template <> struct Elem<f16_t> {
  using raw = uint16_t;
  static constexpr int code = STH_F16;
  static constexpr bool is_float = true;
  static constexpr bool is_signed = true;
};

// Widen a staged value to the arithmetic type of the fold.
template <typename A, typename S>
__device__ __forceinline__ A widen(typename Elem<S>::raw v) {
  if constexpr (__is_same(S, bf16_t)) {
    return static_cast<A>(bf16_to_float(v));
  } else if constexpr (__is_same(S, f16_t)) {
    return static_cast<A>(f16_to_float(v));
  } else {
    return static_cast<A>(v);
  }
}
template <typename O, typename A>
__device__ __forceinline__ typename Elem<O>::raw narrow(A v) {
  if constexpr (__is_same(O, bf16_t)) {
    return float_to_bf16(static_cast<float>(v));
  } else if constexpr (__is_same(O, f16_t)) {
    return float_to_f16(static_cast<float>(v));
  } else {
    return static_cast<typename Elem<O>::raw>(v);
  }
}

// Is a staged raw value nonzero (Life counts nonzero neighbours)?
template <typename R>
__device__ __forceinline__ bool nonzero(R v) {
  return v != R(0);
}

// ----------------------------------------------------- boundary mapping
// bounded_index for one component (src/array.jl:154-179), 0-based.
__device__ __forceinline__ int64_t map_index(int64_t i, int64_t n, int bc) {
  if (i >= 0 && i < n) return i;
  if (bc == STH_WRAP) return i < 0 ? i + n : i - n;
  // STH_REFLECT
  return i < 0 ? -i : 2 * n - 2 - i;
}

// ------------------------------------------------------- parameter block
// A logical array on the device: element (i0, i1, i2) is at
// base + i0 + i1 * s1 + i2 * s2 elements (dim 1 is contiguous).  `h` is the
// halo width actually present in memory around the logical array.
struct DevArr {
  char* base;  // address of logical element (0, 0, 0)
  int64_t n0, n1, n2;
  int64_t s1, s2;
  int32_t h0, h1, h2;  // halo cells present in memory on each side, per dim
  int32_t pad_;
};

template <typename Raw>
__device__ __forceinline__ Raw raw_from_bits(uint64_t b) {
  if constexpr (sizeof(Raw) == 1) return __builtin_bit_cast(Raw, static_cast<uint8_t>(b));
  else if constexpr (sizeof(Raw) == 2) return __builtin_bit_cast(Raw, static_cast<uint16_t>(b));
  else if constexpr (sizeof(Raw) == 4) return __builtin_bit_cast(Raw, static_cast<uint32_t>(b));
  else return __builtin_bit_cast(Raw, b);
}

constexpr int kMaxRuntimeL = 128;  // runtime-offset tables live in kernargs

struct OpParams {
  double alpha;
  uint64_t birth;
  uint64_t survive;
};

struct MapParams {
  DevArr src;
  DevArr dst;
  int32_t bc;         // sth_boundary (Conditional mapping)
  int32_t read_halo;  // 1: the halo in memory is valid; read it, no mapping
  int32_t rx, ry, rz; // reach of the stencil per dim (tile halo)
  int32_t L;          // runtime stencils: number of offsets
  uint64_t padval;    // raw bits of the Remove value (storage type)
  OpParams op;
  int32_t tiles_x, tiles_y, tiles_z;  // tile grid
  int32_t xcd_remap;                  // 1: remap block ids so each XCD owns a contiguous range
  int32_t fuse_halo;                  // Life kernel: read the interior only and refresh the
                                      // source halo in the same launch (update_boundary!)
  int32_t lpr_shift;                  // streaming kernels: log2 of the lanes per row (3..6)
  // runtime stencils: packed (dx, dy, dz) as int8 each in the low 24 bits
  int32_t offs[kMaxRuntimeL];
  // Kernel weights in the accumulator type (f64 uses both halves).
  union {
    float f[kMaxRuntimeL];
    double d[kMaxRuntimeL];
    uint64_t u64[kMaxRuntimeL];
  } w;
};

// Block id after spreading consecutive tiles over one XCD: the hardware
// dispatches workgroup i to XCD (i mod 8), so tile t of XCD k is
// k * ceil(n/8) + t, keeping spatial neighbours in one L2.
__device__ __forceinline__ int xcd_block_id(int bid, int nblocks, int remap) {
  if (!remap) return bid;
  const int per = (nblocks + 7) / 8;
  const int xcd = bid & 7;
  const int t = bid >> 3;
  int v = xcd * per + t;
  // Ragged last XCDs: fall back to identity for ids the formula overflows.
  const int full = nblocks - (per - 1) * 8;  // XCDs that own `per` tiles
  if (xcd >= full) v = full * per + (xcd - full) * (per - 1) + t;
  return v;
}

// Raw buffer accesses: wave-uniform byte offset (soffset) + 32-bit per-lane
// byte offset, one instruction per access and no 64-bit address math.
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 2) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else if constexpr (sizeof(T) == 8) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  } else {
    static_assert(sizeof(T) == 16, "2-, 4-, 8- or 16-byte accesses");
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  }
}
// Cache policy of stores whose data this launch never reads back: the
// non-temporal hint (`nt`) streams them past the caches.  It pays for grids
// larger than the 256 MB Infinity Cache (the z-marching one-step kernel at
// 512^3: 1.05-1.3x; the 2-D two-step kernel at 16384^2: 1.03x) and costs
// where the next launch would find the data there (Game of Life 8192^2, two
// 67 MB buffers: 1.3x slower; the 3-D two-step kernel on f32 512^3: 1.13x
// slower), so each kernel chooses.
constexpr int kStoreNT = 2;

// AUX is the cache policy of the store (kStoreNT: non-temporal).
template <typename T, int AUX = 0>
__device__ __forceinline__ void buf_store(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 2) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, voff, soff, AUX);
  } else if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, AUX);
  } else if constexpr (sizeof(T) == 8) {
    using V2 = unsigned int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, voff, soff, AUX);
  } else {
    static_assert(sizeof(T) == 16, "2-, 4-, 8- or 16-byte accesses");
    using V4 = unsigned int __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, v), r, voff, soff, AUX);
  }
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
  // gfx9 raw buffer (stride 0): accesses at byte offsets >= `bytes` are
  // dropped (stores) or return 0 (loads); word 3 is 0x00020000
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}

}  // namespace sth
