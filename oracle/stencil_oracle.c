/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it as the checker and as the reported CPU baseline.
 *
 * A CPU restatement, in plain C, of the per-cell algorithm of Stencils.jl
 * v0.3.6's gather path (`mapstencil` = `gatherstencil`, src/gatherstencil.jl:
 * 127-128).  It follows the reference step by step:
 *
 *   gatherstencil!(f, dest, source)       src/gatherstencil.jl:89-103
 *     update_boundary!(source)            src/array.jl:195-242   -> orc_update_boundary
 *     per cell: dest[I] = f(stencil(source, I))  src/gatherstencil.jl:105-109
 *       neighbors / getneighbor           src/array.jl:68-138    -> fetch()
 *       bounded_index (Wrap, Reflect)     src/array.jl:146-179   -> bound_wrap/bound_reflect
 *       Halo shifted indexing (add_halo)  src/array.jl:331-377
 *       _center = A[I]                    src/array.jl:32
 *   mapstencil!(f, ::SwitchingStencilArray) src/gatherstencil.jl:77-83 -> orc_iterate
 *
 * Reductions (the fixed set) fold left over the offsets in offset order:
 *   sum      StaticArrays `sum` (external, StaticArrays 1.x): left fold of `+`
 *            in the element type (integers wrap); Bool sums to Int64.
 *   mean     `mean` over the Stencil, a StaticVector: floats fold in T and
 *            divide by L (Statistics' `mean` does the same for floats);
 *            integers take the element-type `sum` above (wrapping; Bool
 *            counts in Int64) converted to Float64, then `/ L` -- what
 *            StaticArrays' `mean(a) = sum(a) / length(a)` gives.  Which of
 *            StaticArrays' or Statistics' generic method (every element
 *            promoted to Float64 first) a Stencil reaches is not pinned by
 *            anything the reference holds; the library accelerates integer
 *            sum/mean where the two can differ only when its caller says
 *            which (STH_OPF_INT_WRAP).  PARITY UNPINNED for those cases.
 *   min/max  StaticArrays `minimum`/`maximum`: left fold of Julia `min`/`max`
 *            (NaN-propagating, -0.0 < +0.0).
 *   kernelproduct  src/stencils/kernel.jl:37-43: acc = zero(T);
 *            acc += hood[i] * kernel[i]  (separate multiply and add, no FMA).
 *   life     c' = (c != 0 ? survive : birth) bit s, s = # nonzero neighbours.
 *   laplacian c + alpha * (sum(neighbours) - L * c), each operation rounded.
 *
 * Floating-point contraction is off (compile with -ffp-contract=off), as Julia
 * never fuses `a + b * c` without @fastmath / muladd.
 *
 * Conditional padding: the reference takes an unchecked fast path when the
 * corners I +- R are in bounds (src/array.jl:101-115).  Whenever the stencil's
 * reach is <= R that path reads exactly what the checked path reads, so this
 * restatement always takes the checked path; for Positional stencils whose
 * reach exceeds the reference's signed-max radius the reference reads out of
 * bounds (undefined), and this oracle applies the boundary instead.
 *
 * BF16 / F16 with STH_OPF_ACC_F32 are this project's storage precision mode
 * (the reference has no BF16 type): values are widened to f32, folded in f32
 * in offset order (kernelproduct with fused multiply-adds, fmaf) and rounded
 * once (round-to-nearest-even) at the store.
 * ORC_MODE_F64 computes any float input in double: the FP64 oracle that
 * reduced-precision results are measured against.
 */
#include "../include/stencils_hip.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Contraction is disabled by -ffp-contract=off in oracle/Makefile. */

#define ORC_MODE_NATIVE 0
#define ORC_MODE_F64 1

/* ---------------------------------------------------------------- helpers */

static float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1fu;
  uint32_t man = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    if (man == 0) {
      bits = sign;
    } else { /* subnormal: value = man * 2^-24 */
      float v = (float)man * 5.9604644775390625e-08f;
      memcpy(&bits, &v, 4);
      bits |= sign;
    }
  } else if (exp == 31) {
    bits = sign | 0x7f800000u | (man << 13);
  } else {
    bits = sign | ((exp + 112u) << 23) | (man << 13);
  }
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

static uint16_t f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) /* inf or nan */
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) /* rounds to >= 65520 -> inf */
    return (uint16_t)(sign | 0x7c00u);
  if (ax < 0x38800000u) { /* result subnormal or zero: round(|f| * 2^24) */
    float af;
    memcpy(&af, &ax, 4);
    float scaled = af * 16777216.0f; /* exact */
    float r = nearbyintf(scaled);   /* ties to even in the default mode */
    return (uint16_t)(sign | (uint32_t)r);
  }
  uint32_t e = ((ax >> 23) - 112u) << 10;
  uint32_t m = (ax >> 13) & 0x3ffu;
  uint32_t rest = ax & 0x1fffu;
  uint32_t h = e | m;
  if (rest > 0x1000u || (rest == 0x1000u && (m & 1u))) h += 1u;
  return (uint16_t)(sign | h);
}

static float bf16_to_f32(uint16_t b) {
  uint32_t bits = (uint32_t)b << 16;
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

static uint16_t f32_to_bf16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

static int dtype_size(int dt) {
  switch (dt) {
    case STH_BOOL: case STH_U8: case STH_I8: return 1;
    case STH_U16: case STH_I16: case STH_F16: case STH_BF16: return 2;
    case STH_U32: case STH_I32: case STH_F32: return 4;
    case STH_U64: case STH_I64: case STH_F64: return 8;
    default: return 0;
  }
}

static int is_float(int dt) {
  return dt == STH_F16 || dt == STH_BF16 || dt == STH_F32 || dt == STH_F64;
}
static int is_signed_int(int dt) {
  return dt == STH_I8 || dt == STH_I16 || dt == STH_I32 || dt == STH_I64;
}

/* Value categories for the fold. */
enum { CAT_F64 = 0, CAT_F32 = 1, CAT_SINT = 2, CAT_UINT = 3 };

static double load_f64(const void* p, int dt) {
  switch (dt) {
    case STH_F64: return *(const double*)p;
    case STH_F32: return (double)*(const float*)p;
    case STH_F16: return (double)f16_to_f32(*(const uint16_t*)p);
    case STH_BF16: return (double)bf16_to_f32(*(const uint16_t*)p);
    case STH_BOOL: case STH_U8: return (double)*(const uint8_t*)p;
    case STH_I8: return (double)*(const int8_t*)p;
    case STH_U16: return (double)*(const uint16_t*)p;
    case STH_I16: return (double)*(const int16_t*)p;
    case STH_U32: return (double)*(const uint32_t*)p;
    case STH_I32: return (double)*(const int32_t*)p;
    case STH_U64: return (double)*(const uint64_t*)p;
    case STH_I64: return (double)*(const int64_t*)p;
  }
  return 0.0;
}
static float load_f32(const void* p, int dt) {
  switch (dt) {
    case STH_F32: return *(const float*)p;
    case STH_F16: return f16_to_f32(*(const uint16_t*)p);
    case STH_BF16: return bf16_to_f32(*(const uint16_t*)p);
  }
  return 0.0f;
}
static int64_t load_sint(const void* p, int dt) {
  switch (dt) {
    case STH_I8: return *(const int8_t*)p;
    case STH_I16: return *(const int16_t*)p;
    case STH_I32: return *(const int32_t*)p;
    case STH_I64: return *(const int64_t*)p;
  }
  return 0;
}
static uint64_t load_uint(const void* p, int dt) {
  switch (dt) {
    case STH_BOOL: case STH_U8: return *(const uint8_t*)p;
    case STH_U16: return *(const uint16_t*)p;
    case STH_U32: return *(const uint32_t*)p;
    case STH_U64: return *(const uint64_t*)p;
  }
  return 0;
}
/* A 64-bit two's-complement accumulator reduced to the width of `dt`, as a
 * fold with that type's own `+` leaves it. */
static int64_t wrap_sint(uint64_t u, int dt) {
  switch (dt) {
    case STH_I8: return (int8_t)u;
    case STH_I16: return (int16_t)u;
    case STH_I32: return (int32_t)u;
  }
  return (int64_t)u;
}
static uint64_t wrap_uint(uint64_t u, int dt) {
  switch (dt) {
    case STH_U8: return (uint8_t)u;
    case STH_U16: return (uint16_t)u;
    case STH_U32: return (uint32_t)u;
  }
  return u;
}

/* Julia max/min on IEEE floats: NaN if either is NaN, -0.0 < +0.0. */
static double jmax_d(double x, double y) {
  if (isnan(x) || isnan(y)) return x - y;
  if (x > y) return x;
  if (y > x) return y;
  return signbit(x) ? y : x;
}
static double jmin_d(double x, double y) {
  if (isnan(x) || isnan(y)) return x - y;
  if (x < y) return x;
  if (y < x) return y;
  return signbit(x) ? x : y;
}
static float jmax_f(float x, float y) {
  if (isnan(x) || isnan(y)) return x - y;
  if (x > y) return x;
  if (y > x) return y;
  return signbit(x) ? y : x;
}
static float jmin_f(float x, float y) {
  if (isnan(x) || isnan(y)) return x - y;
  if (x < y) return x;
  if (y < x) return y;
  return signbit(x) ? x : y;
}

/* ------------------------------------------------------------- the job */

typedef struct {
  const sth_array* src;
  const sth_array* dst;
  int L;
  int32_t offs[4096 * 3]; /* zero-extended to 3 dims */
  int bc;
  const void* padval;
  int halo_mode; /* 1: read the (refreshed) halo directly */
  sth_op op;
  const void* weights;
  int wdtype;
  int cat;
  int mode;
  int64_t row0, row1; /* rows (dims 2..3 flattened) owned by a thread */
} job_t;

static inline const char* elem_ptr(const sth_array* a, int64_t i0, int64_t i1,
                                   int64_t i2) {
  int64_t h0 = a->halo, h1 = a->ndim > 1 ? a->halo : 0, h2 = a->ndim > 2 ? a->halo : 0;
  int64_t off = (i0 + h0) * a->stride[0] + (i1 + h1) * a->stride[1] +
                (i2 + h2) * a->stride[2];
  return (const char*)a->data + off * dtype_size(a->dtype);
}

/* bounded_index for one component, src/array.jl:154-179 (0-based). */
static inline int64_t bound_wrap(int64_t i, int64_t s) {
  if (i < 0) return i + s;
  if (i >= s) return i - s;
  return i;
}
static inline int64_t bound_reflect(int64_t i, int64_t s) {
  if (i < 0) return -i;
  if (i >= s) return 2 * s - 2 - i;
  return i;
}

/* Neighbour at logical J: pointer to the element, or NULL for `padval`. */
static inline const char* fetch(const job_t* j, int64_t J0, int64_t J1, int64_t J2) {
  const sth_array* a = j->src;
  if (j->halo_mode) return elem_ptr(a, J0, J1, J2);
  int64_t J[3] = {J0, J1, J2};
  int inb = 1;
  for (int d = 0; d < 3; ++d)
    if (J[d] < 0 || J[d] >= a->size[d]) inb = 0;
  if (inb) return elem_ptr(a, J0, J1, J2);
  if (j->bc == STH_REMOVE) return NULL;
  for (int d = 0; d < 3; ++d)
    J[d] = j->bc == STH_WRAP ? bound_wrap(J[d], a->size[d]) : bound_reflect(J[d], a->size[d]);
  return elem_ptr(a, J[0], J[1], J[2]);
}

static void store_f64(void* p, int dt, double v) {
  if (dt == STH_F64) *(double*)p = v;
  else if (dt == STH_F32) *(float*)p = (float)v;
}

static void cell(const job_t* j, int64_t i0, int64_t i1, int64_t i2) {
  const sth_array* s = j->src;
  const int dt = s->dtype;
  const int L = j->L;
  const int kind = j->op.kind;
  char* out = (char*)elem_ptr(j->dst, i0, i1, i2);
  const char* cptr = elem_ptr(s, i0, i1, i2);

  if (kind == STH_OP_LIFE) {
    uint64_t cnt = 0;
    for (int k = 0; k < L; ++k) {
      const char* p = fetch(j, i0 + j->offs[3 * k], i1 + j->offs[3 * k + 1], i2 + j->offs[3 * k + 2]);
      const char* q = p ? p : (const char*)j->padval;
      int nz = 0;
      for (int b = 0; b < dtype_size(dt); ++b) nz |= q[b] != 0;
      cnt += (uint64_t)nz;
    }
    int cz = 0;
    for (int b = 0; b < dtype_size(dt); ++b) cz |= cptr[b] != 0;
    uint64_t mask = cz ? j->op.survive_mask : j->op.birth_mask;
    uint64_t r = cnt < 64 ? (mask >> cnt) & 1u : 0u;
    memset(out, 0, (size_t)dtype_size(j->dst->dtype));
    *(uint8_t*)out = (uint8_t)r; /* little-endian: lowest byte carries 0/1 */
    return;
  }

  if (j->cat == CAT_F64) {
    double acc = 0.0, c = load_f64(cptr, dt);
    for (int k = 0; k < L; ++k) {
      const char* p = fetch(j, i0 + j->offs[3 * k], i1 + j->offs[3 * k + 1], i2 + j->offs[3 * k + 2]);
      double v = load_f64(p ? p : (const char*)j->padval, dt);
      switch (kind) {
        case STH_OP_SUM: case STH_OP_MEAN: case STH_OP_LAPLACIAN:
          acc = k == 0 ? v : acc + v; break;
        case STH_OP_MIN: acc = k == 0 ? v : jmin_d(acc, v); break;
        case STH_OP_MAX: acc = k == 0 ? v : jmax_d(acc, v); break;
        case STH_OP_KERNELPRODUCT: {
          double w = load_f64((const char*)j->weights + (size_t)k * dtype_size(j->wdtype), j->wdtype);
          double pr = v * w;
          acc = acc + pr;
          break;
        }
      }
    }
    if (kind == STH_OP_MEAN) acc = acc / (double)L;
    if (kind == STH_OP_LAPLACIAN) {
      double t = (double)L * c;
      double d = acc - t;
      double a = j->op.alpha * d;
      acc = c + a;
    }
    store_f64(out, j->dst->dtype, acc);
    return;
  }

  if (j->cat == CAT_F32) {
    float acc = 0.0f, c = load_f32(cptr, dt);
    for (int k = 0; k < L; ++k) {
      const char* p = fetch(j, i0 + j->offs[3 * k], i1 + j->offs[3 * k + 1], i2 + j->offs[3 * k + 2]);
      float v = load_f32(p ? p : (const char*)j->padval, dt);
      switch (kind) {
        case STH_OP_SUM: case STH_OP_MEAN: case STH_OP_LAPLACIAN:
          acc = k == 0 ? v : acc + v; break;
        case STH_OP_MIN: acc = k == 0 ? v : jmin_f(acc, v); break;
        case STH_OP_MAX: acc = k == 0 ? v : jmax_f(acc, v); break;
        case STH_OP_KERNELPRODUCT: {
          float w = load_f32((const char*)j->weights + (size_t)k * 4, STH_F32);
          if (dt == STH_BF16 || dt == STH_F16) {
            acc = fmaf(v, w, acc); /* storage mode: fused multiply-add */
          } else {
            float pr = v * w;
            acc = acc + pr;
          }
          break;
        }
      }
    }
    if (kind == STH_OP_MEAN) acc = acc / (float)L;
    if (kind == STH_OP_LAPLACIAN) {
      float t = (float)L * c;
      float d = acc - t;
      float a = (float)j->op.alpha * d;
      acc = c + a;
    }
    int odt = j->dst->dtype;
    if (odt == STH_F32) *(float*)out = acc;
    else if (odt == STH_F16) *(uint16_t*)out = f32_to_f16(acc);
    else if (odt == STH_BF16) *(uint16_t*)out = f32_to_bf16(acc);
    return;
  }

  /* integers (and Bool) */
  int sgn = j->cat == CAT_SINT;
  if (kind == STH_OP_MEAN) {
    /* sum(a) / length(a) over the StaticVector: the sum folds with the
     * element type's wrapping `+` (Bool + Bool is an Int), then Float64 / L */
    uint64_t uacc = 0;
    for (int k = 0; k < L; ++k) {
      const char* p = fetch(j, i0 + j->offs[3 * k], i1 + j->offs[3 * k + 1], i2 + j->offs[3 * k + 2]);
      const char* q = p ? p : (const char*)j->padval;
      uint64_t uv = sgn ? (uint64_t)load_sint(q, dt) : load_uint(q, dt);
      uacc = k == 0 ? uv : uacc + uv;
    }
    double s;
    if (dt == STH_BOOL) s = (double)(int64_t)uacc;
    else if (sgn) s = (double)wrap_sint(uacc, dt);
    else s = (double)wrap_uint(uacc, dt);
    *(double*)out = s / (double)L;
    return;
  }
  uint64_t uacc = 0;
  int64_t sacc = 0;
  for (int k = 0; k < L; ++k) {
    const char* p = fetch(j, i0 + j->offs[3 * k], i1 + j->offs[3 * k + 1], i2 + j->offs[3 * k + 2]);
    const char* q = p ? p : (const char*)j->padval;
    uint64_t uv = sgn ? (uint64_t)load_sint(q, dt) : load_uint(q, dt);
    int64_t sv = (int64_t)uv;
    switch (kind) {
      case STH_OP_SUM: uacc = k == 0 ? uv : uacc + uv; break;
      case STH_OP_KERNELPRODUCT: {
        const char* wp = (const char*)j->weights + (size_t)k * dtype_size(j->wdtype);
        uint64_t w = sgn ? (uint64_t)load_sint(wp, j->wdtype) : load_uint(wp, j->wdtype);
        uacc = uacc + uv * w;
        break;
      }
      case STH_OP_MIN:
        if (sgn) sacc = k == 0 ? sv : (sv < sacc ? sv : sacc);
        else uacc = k == 0 ? uv : (uv < uacc ? uv : uacc);
        break;
      case STH_OP_MAX:
        if (sgn) sacc = k == 0 ? sv : (sv > sacc ? sv : sacc);
        else uacc = k == 0 ? uv : (uv > uacc ? uv : uacc);
        break;
    }
  }
  if (sgn && (kind == STH_OP_MIN || kind == STH_OP_MAX)) uacc = (uint64_t)sacc;
  /* store, truncating to the destination width (wrapping arithmetic) */
  switch (dtype_size(j->dst->dtype)) {
    case 1: *(uint8_t*)out = (uint8_t)uacc; break;
    case 2: *(uint16_t*)out = (uint16_t)uacc; break;
    case 4: *(uint32_t*)out = (uint32_t)uacc; break;
    case 8: *(uint64_t*)out = uacc; break;
  }
}

static void* run_rows(void* arg) {
  const job_t* j = (const job_t*)arg;
  const int64_t n0 = j->dst->size[0], n1 = j->dst->size[1];
  for (int64_t r = j->row0; r < j->row1; ++r) {
    int64_t i1 = r % n1, i2 = r / n1;
    for (int64_t i0 = 0; i0 < n0; ++i0) cell(j, i0, i1, i2);
  }
  return NULL;
}

/* Output dtype of `op` over `src_dt`, as the reference's _return_type gives
 * it (src/gatherstencil.jl:41-56), or -1 if unsupported. */
int orc_out_dtype(int src_dt, int opkind, unsigned flags, int wdtype, int mode) {
  if (mode == ORC_MODE_F64) return STH_F64;
  int storage = src_dt == STH_BF16 || src_dt == STH_F16;
  if (storage && !(flags & STH_OPF_ACC_F32)) return -1;
  switch (opkind) {
    case STH_OP_SUM: return src_dt == STH_BOOL ? STH_I64 : src_dt;
    case STH_OP_MEAN: return is_float(src_dt) ? src_dt : STH_F64;
    case STH_OP_MIN: case STH_OP_MAX: return src_dt;
    case STH_OP_KERNELPRODUCT:
      if (src_dt == STH_BOOL) return -1;
      if (storage) return wdtype == STH_F32 ? src_dt : -1;
      return wdtype == src_dt ? src_dt : -1;
    case STH_OP_LIFE: return is_float(src_dt) ? -1 : src_dt;
    case STH_OP_LAPLACIAN: return is_float(src_dt) ? src_dt : -1;
  }
  return -1;
}

/* update_boundary! for Halo padding, src/array.jl:195-242. */
int orc_update_boundary(const sth_array* a, int bc, const void* padval) {
  if (bc == STH_USE || a->halo == 0) return STH_OK;
  const int R = a->halo, nd = a->ndim, es = dtype_size(a->dtype);
  /* halo_val maps -R..-1 into the interior: a BoundsError when R >= axis */
  if (bc != STH_REMOVE)
    for (int d = 0; d < nd; ++d)
      if (R >= a->size[d]) return STH_EINVAL;
  int64_t ps[3] = {1, 1, 1};
  for (int d = 0; d < nd; ++d) ps[d] = a->size[d] + 2 * R;
  char* base = (char*)a->data;
  for (int n = 0; n < nd; ++n) {
    for (int64_t p2 = 0; p2 < ps[2]; ++p2)
      for (int64_t p1 = 0; p1 < ps[1]; ++p1)
        for (int64_t p0 = 0; p0 < ps[0]; ++p0) {
          int64_t p[3] = {p0, p1, p2};
          int in_slab = p[n] < R || p[n] >= a->size[n] + R;
          if (!in_slab) continue;
          char* dstp = base + (p0 * a->stride[0] + p1 * a->stride[1] + p2 * a->stride[2]) * es;
          if (bc == STH_REMOVE) {
            memcpy(dstp, padval, (size_t)es);
            continue;
          }
          int64_t I[3] = {0, 0, 0};
          for (int d = 0; d < nd; ++d) {
            int64_t i = p[d] - R;
            I[d] = bc == STH_WRAP ? bound_wrap(i, a->size[d]) : bound_reflect(i, a->size[d]);
          }
          memcpy(dstp, elem_ptr(a, I[0], I[1], I[2]), (size_t)es);
        }
  }
  return STH_OK;
}

/*
 * gatherstencil!(f, dest, source): `offsets` holds L * st_ndim ints in fold
 * order; `nthreads` host threads split the rows.  Returns STH_* status.
 */
int orc_map(const sth_array* src, const sth_array* dst, int L, int st_ndim,
            const int32_t* offsets, int bc, const void* padval, int padding,
            const sth_op* op, const void* weights, int wdtype, int mode,
            int nthreads) {
  if (L < 1 || L > 4096 || st_ndim < 1 || st_ndim > src->ndim) return STH_EINVAL;
  for (int d = 0; d < 3; ++d)
    if (src->size[d] != dst->size[d]) return STH_EINVAL;
  int odt = orc_out_dtype(src->dtype, op->kind, op->flags, wdtype, mode);
  if (odt < 0 || odt != dst->dtype) return STH_ENOTSUP;
  if (padding == STH_CONDITIONAL && bc == STH_USE) return STH_EINVAL;
  if (padding != STH_CONDITIONAL) {
    int rc = orc_update_boundary(src, bc, padval);
    if (rc) return rc;
  }
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  if (!j) return STH_ENOMEM;
  j->src = src;
  j->dst = dst;
  j->L = L;
  for (int k = 0; k < L; ++k)
    for (int d = 0; d < 3; ++d) j->offs[3 * k + d] = d < st_ndim ? offsets[k * st_ndim + d] : 0;
  j->bc = bc;
  j->padval = padval;
  j->halo_mode = padding != STH_CONDITIONAL;
  j->op = *op;
  j->weights = weights;
  j->wdtype = wdtype;
  j->mode = mode;
  int dt = src->dtype;
  if (mode == ORC_MODE_F64 || dt == STH_F64) j->cat = CAT_F64;
  else if (is_float(dt)) j->cat = CAT_F32;
  else if (is_signed_int(dt)) j->cat = CAT_SINT;
  else j->cat = CAT_UINT;

  int64_t rows = dst->size[1] * dst->size[2];
  if (nthreads < 1) nthreads = 1;
  if (nthreads > rows) nthreads = (int)rows;
  if (nthreads <= 1) {
    j->row0 = 0;
    j->row1 = rows;
    run_rows(j);
  } else {
    job_t* jobs = (job_t*)malloc(sizeof(job_t) * (size_t)nthreads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
      jobs[t] = *j;
      jobs[t].row0 = rows * t / nthreads;
      jobs[t].row1 = rows * (t + 1) / nthreads;
      pthread_create(&th[t], NULL, run_rows, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
  }
  free(j);
  return STH_OK;
}

/*
 * `steps` calls of mapstencil!(f, A::SwitchingStencilArray): each step
 * refreshes the source halo (Halo padding, re-wrapped as Halo{:in}) and
 * writes the destination interior, then the two buffers switch.  Returns the
 * index (0 = a, 1 = b) of the buffer holding the newest state in *current.
 */
int orc_iterate(const sth_array* a, const sth_array* b, int L, int st_ndim,
                const int32_t* offsets, int bc, const void* padval, int padding,
                const sth_op* op, const void* weights, int wdtype, int64_t steps,
                int nthreads, int32_t* current) {
  const sth_array* bufs[2] = {a, b};
  int cur = 0;
  for (int64_t s = 0; s < steps; ++s) {
    int rc = orc_map(bufs[cur], bufs[1 - cur], L, st_ndim, offsets, bc, padval,
                     padding, op, weights, wdtype, ORC_MODE_NATIVE, nthreads);
    if (rc) return rc;
    cur = 1 - cur;
  }
  *current = cur;
  return STH_OK;
}

/*
 * scatterstencil!(f, op, dest, source, args...) for the fixed value rules
 * of sth_scatter, 2-D, in the reference's loop order
 * (src/scatterstencil.jl:37-47, 51-75): update_boundary! on the source and
 * the second array; passes p = 1..S (S = 2R + 1) over source columns
 * j = p, p + S, ... (1-based), rows i ascending, offsets k in order; target
 * (i, j) + offset_k mapped by the source's boundary (:78-114: Wrap mod1,
 * Reflect once, Remove / Use skip), dest[t] = op(dest[t], v_k).  `aux` is
 * the second StencilArray (NULL if the rule does not read one).  Float32 /
 * Float64.  zero_dest clears dest first (the SwitchingStencilArray forms).
 */
static int64_t scat_map(int64_t p, int64_t n, int bc) {
  if (p >= 0 && p < n) return p;
  if (bc == STH_WRAP) return ((p % n) + n) % n;
  if (bc == STH_REFLECT) return p < 0 ? -p : 2 * n - 2 - p;
  return -1;
}

int orc_scatter(const sth_array* src, int L, const int32_t* offsets, int R, int bc, const void* padval,
                int padding, const sth_array* aux, int aux_bc, const void* aux_padval, int aux_padding,
                int kind, double coef, int combine, const sth_array* dst, int zero_dest) {
  if (src->ndim != 2 || dst->ndim != 2 || L < 1 || R < 0) return STH_EINVAL;
  for (int d = 0; d < 2; ++d)
    if (src->size[d] != dst->size[d] || (aux && aux->size[d] != dst->size[d])) return STH_EINVAL;
  const int dt = dst->dtype;
  if ((dt != STH_F32 && dt != STH_F64) || src->dtype != dt || (aux && aux->dtype != dt)) return STH_ENOTSUP;
  if (padding != STH_CONDITIONAL) {
    int rc = orc_update_boundary(src, bc, padval);
    if (rc) return rc;
  }
  job_t* bj = NULL;
  if (aux) {
    if (aux_padding != STH_CONDITIONAL) {
      int rc = orc_update_boundary(aux, aux_bc, aux_padval);
      if (rc) return rc;
    }
    bj = (job_t*)calloc(1, sizeof(job_t));
    if (!bj) return STH_ENOMEM;
    bj->src = aux;
    bj->bc = aux_bc;
    bj->padval = aux_padval;
    bj->halo_mode = aux_padding != STH_CONDITIONAL;
  }
  const int64_t ny = dst->size[0], nx = dst->size[1];
  const int64_t S = 2 * R + 1;
  if (zero_dest)
    for (int64_t j = 0; j < nx; ++j)
      for (int64_t i = 0; i < ny; ++i) memset((char*)elem_ptr(dst, i, j, 0), 0, (size_t)dtype_size(dt));
  for (int64_t pass = 0; pass < S; ++pass)
    for (int64_t j = pass; j < nx; j += S)
      for (int64_t i = 0; i < ny; ++i) {
        const double a = load_f64(elem_ptr(src, i, j, 0), dt);
        const double b = aux ? load_f64(elem_ptr(aux, i, j, 0), dt) : 0.0;
        for (int k = 0; k < L; ++k) {
          const int oi = offsets[2 * k], oj = offsets[2 * k + 1];
          const int64_t ti = scat_map(i + oi, ny, bc), tj = scat_map(j + oj, nx, bc);
          if (ti < 0 || tj < 0) continue;
          char* t = (char*)elem_ptr(dst, ti, tj, 0);
          if (dt == STH_F64) {
            double v = coef;
            if (kind == STH_SCATTER_CENTER) v = a;
            else if (kind == STH_SCATTER_SCALED) v = a * coef;
            else if (kind >= STH_SCATTER_DOWNHILL) {
              const char* q = fetch(bj, i + oi, j + oj, 0);
              const double dz = b - load_f64(q ? q : (const char*)aux_padval, dt);
              v = dz > 0 ? (kind == STH_SCATTER_DOWNHILL ? a * coef : (a * dz) * coef) : 0.0;
            }
            double* o = (double*)t;
            *o = combine == STH_COMBINE_MAX ? jmax_d(*o, v) : combine == STH_COMBINE_MIN ? jmin_d(*o, v) : *o + v;
          } else {
            const float af = (float)a, bf = (float)b, cf = (float)coef;
            float v = cf;
            if (kind == STH_SCATTER_CENTER) v = af;
            else if (kind == STH_SCATTER_SCALED) v = af * cf;
            else if (kind >= STH_SCATTER_DOWNHILL) {
              const char* q = fetch(bj, i + oi, j + oj, 0);
              const float dz = bf - load_f32(q ? q : (const char*)aux_padval, dt);
              v = dz > 0 ? (kind == STH_SCATTER_DOWNHILL ? af * cf : (af * dz) * cf) : 0.0f;
            }
            float* o = (float*)t;
            *o = combine == STH_COMBINE_MAX ? jmax_f(*o, v) : combine == STH_COMBINE_MIN ? jmin_f(*o, v) : *o + v;
          }
        }
      }
  free(bj);
  return STH_OK;
}

/* Rounding helpers exported for the tests. */
uint16_t orc_f32_to_bf16(float f) { return f32_to_bf16(f); }
uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
float orc_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }
