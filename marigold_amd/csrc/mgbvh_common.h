// Device-side arithmetic shared by every mgbvh kernel.
//
// Bit-exactness with marigold's builder (bvh/src/lib.rs) rests on:
//   * fp32 ops in the reference's order, no FMA contraction (-ffp-contract=off),
//   * correctly rounded fp32 division (HIP default),
//   * min/max as `a < b ? a : b` (x86 minps/maxps semantics of glam Vec3A),
//   * order-independent min/max folds through an order-preserving u32 key, so
//     atomics give the same bits as the sequential fold,
//   * f32::total_cmp ordering and first-minimum tie-breaks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MG_DEPTH_COST 1.0f   // BvhNode::DEPTH_COST, bvh/src/lib.rs:217
#define MG_OBJECT_COST 8.0f  // BvhNode::OBJECT_COST, bvh/src/lib.rs:218
#define MG_NBINS 32          // len.clamp(16, 32) with len > 32, bvh/src/lib.rs:343
#define MG_SMALL 32          // adaptive sweep for len <= 32, bvh/src/lib.rs:518

namespace mg {

__host__ __device__ inline float fmin_x86(float a, float b) { return a < b ? a : b; }
__host__ __device__ inline float fmax_x86(float a, float b) { return a > b ? a : b; }

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ inline float bitsf(uint32_t u) { return __uint_as_float(u); }

// Order-preserving map float -> u32 (for non-NaN values): a < b <=> key(a) < key(b).
__device__ inline uint32_t okey(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float okey_inv(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}
// Accumulators are zero-initialised and only ever atomically max-ed:
// a min is kept as ~okey(x), a max as okey(x).
__device__ inline uint32_t acc_min_enc(float f) { return ~okey(f); }
__device__ inline float acc_min_dec(uint32_t v) { return okey_inv(~v); }
__device__ inline uint32_t acc_max_enc(float f) { return okey(f); }
__device__ inline float acc_max_dec(uint32_t v) { return okey_inv(v); }

// f32::total_cmp as an unsigned key (ascending).
__device__ inline uint32_t tc_ukey(float f) {
  int32_t i = (int32_t)__float_as_uint(f);
  i ^= (int32_t)(((uint32_t)(i >> 31)) >> 1);
  return (uint32_t)i ^ 0x80000000u;
}
// The float (bits) a tc_ukey came from.
__device__ inline float tc_ukey_inv(uint32_t k) {
  const uint32_t i = k ^ 0x80000000u;
  return __uint_as_float((i & 0x80000000u) ? i ^ 0x7fffffffu : i);
}

// x86-64 SSE produces the default NaN 0xFFC00000 for invalid operations; all
// NaN costs are that value (inputs are finite).
__device__ inline float nan_canon(float f) { return f != f ? __uint_as_float(0xFFC00000u) : f; }

__device__ inline float canon_zero(float x) { return x == 0.0f ? 0.0f : x; }

struct Box {
  float mn[3];
  float mx[3];
};

__device__ inline Box box_empty() {
  Box b;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b.mn[a] = __uint_as_float(0x7f800000u);
    b.mx[a] = __uint_as_float(0xff800000u);
  }
  return b;
}

__device__ inline void box_grow(Box& acc, const Box& o) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    acc.mn[a] = fmin_x86(acc.mn[a], o.mn[a]);
    acc.mx[a] = fmax_x86(acc.mx[a], o.mx[a]);
  }
}

// BoundingVolume::surface_area, bvh/src/lib.rs:79-91.
__device__ inline float box_sa(const Box& b) {
  float w = __fsub_rn(b.mx[0], b.mn[0]);
  float h = __fsub_rn(b.mx[1], b.mn[1]);
  float d = __fsub_rn(b.mx[2], b.mn[2]);
  bool empty = (b.mx[0] < b.mn[0]) | (b.mx[1] < b.mn[1]) | (b.mx[2] < b.mn[2]);
  if (empty) return 0.0f;
  float s = __fadd_rn(__fmul_rn(w, h), __fmul_rn(w, d));
  s = __fadd_rn(s, __fmul_rn(h, d));
  return __fmul_rn(2.0f, s);
}

// box_sa of a box known to hold at least one object (never empty): the same
// value without the emptiness test.
__device__ inline float box_sa_ne(const Box& b) {
  const float w = __fsub_rn(b.mx[0], b.mn[0]);
  const float h = __fsub_rn(b.mx[1], b.mn[1]);
  const float d = __fsub_rn(b.mx[2], b.mn[2]);
  float s = __fadd_rn(__fmul_rn(w, h), __fmul_rn(w, d));
  s = __fadd_rn(s, __fmul_rn(h, d));
  return __fmul_rn(2.0f, s);
}

// (1 + ((SA_lt / SA_p) * 8) * n_lt) + ((SA_gt / SA_p) * 8) * n_gt, lib.rs:278-288.
__device__ inline float split_cost(float sa_p, float sa_lt, float sa_gt, uint32_t n_lt,
                                   uint32_t n_gt) {
  float l = __fdiv_rn(sa_lt, sa_p);
  l = __fmul_rn(l, MG_OBJECT_COST);
  l = __fmul_rn(l, (float)n_lt);
  float g = __fdiv_rn(sa_gt, sa_p);
  g = __fmul_rn(g, MG_OBJECT_COST);
  g = __fmul_rn(g, (float)n_gt);
  float c = __fadd_rn(MG_DEPTH_COST, l);
  c = __fadd_rn(c, g);
  return nan_canon(c);
}

// BoundingVolume::center component, lib.rs:71-73.
__device__ inline float center(float mn, float mx) { return __fmul_rn(__fadd_rn(mn, mx), 0.5f); }

// Bin index, lib.rs:358-362: min(floor(((c - cmin) / cext) * B) as usize, B - 1).
__device__ inline uint32_t bin_index(float c, float cmin, float cext) {
  float pct = __fdiv_rn(__fsub_rn(c, cmin), cext);
  float f = floorf(__fmul_rn(pct, (float)MG_NBINS));
  uint32_t b;
  if (!(f > 0.0f)) b = 0;           // NaN and negatives saturate to 0
  else if (f >= 31.0f) b = MG_NBINS - 1;
  else b = (uint32_t)f;
  return b;
}

// Adaptive threshold i of S steps, lib.rs:430-437.
__device__ inline float adaptive_threshold(float cmin, float cmax, uint32_t steps, uint32_t i) {
  float step = __fdiv_rn(__fsub_rn(cmax, cmin), (float)steps);
  float fi = __fadd_rn((float)i, 0.5f);
  return __fadd_rn(cmin, __fmul_rn(step, fi));
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// ---- DPP cross-lane primitives (CDNA data-parallel primitives: no LDS round trip)
// Segmented inclusive scans inside 32-lane halves (a segment [ls, last] never
// crosses lane 31/32) with DPP row shifts: row_shr 1, 2, 4, 8 inside each
// 16-lane row, then row_bcast:15 carries lane 15 (47) into row 1 (3) for
// segments that started in the row before.  Lanes outside the segment see the
// identity `id`.
template <int CTRL, int ROWS>
__device__ inline uint32_t dpp_mov(uint32_t id, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, CTRL, ROWS, 0xF, false);
}

template <class Op>
__device__ inline uint32_t seg_scan32(uint32_t v, uint32_t id, uint32_t lane, uint32_t ls, Op op) {
  const uint32_t rs = lane & ~15u, s = ls > rs ? ls : rs;
  uint32_t t;
  t = dpp_mov<0x111, 0xF>(id, v); if (lane >= s + 1) v = op(v, t);
  t = dpp_mov<0x112, 0xF>(id, v); if (lane >= s + 2) v = op(v, t);
  t = dpp_mov<0x114, 0xF>(id, v); if (lane >= s + 4) v = op(v, t);
  t = dpp_mov<0x118, 0xF>(id, v); if (lane >= s + 8) v = op(v, t);
  t = dpp_mov<0x142, 0xA>(id, v); if (ls < rs) v = op(v, t);
  return v;
}

// Inclusive segmented minimum scan of u64 keys inside 32-lane halves.
__device__ inline uint64_t seg_scan_min_u64(uint64_t v, uint32_t lane, uint32_t ls) {
  const uint32_t rs = lane & ~15u, s = ls > rs ? ls : rs;
  uint64_t t;
#define MG_SEG_MIN_STEP(C, R, COND)                                                   \
  t = ((uint64_t)dpp_mov<C, R>(0xFFFFFFFFu, (uint32_t)(v >> 32)) << 32) |             \
      dpp_mov<C, R>(0xFFFFFFFFu, (uint32_t)v);                                        \
  if (COND) v = t < v ? t : v;
  MG_SEG_MIN_STEP(0x111, 0xF, lane >= s + 1)
  MG_SEG_MIN_STEP(0x112, 0xF, lane >= s + 2)
  MG_SEG_MIN_STEP(0x114, 0xF, lane >= s + 4)
  MG_SEG_MIN_STEP(0x118, 0xF, lane >= s + 8)
  MG_SEG_MIN_STEP(0x142, 0xA, ls < rs)
#undef MG_SEG_MIN_STEP
  return v;
}

// Minimum of u64 keys over the segment, returned on every lane of it.
__device__ inline uint64_t seg_min_u64(uint64_t v, uint32_t lane, uint32_t ls, uint32_t last) {
  return __shfl(seg_scan_min_u64(v, lane, ls), (int)last, 64);
}

// Minimum of u64 keys over the wavefront (the two half minima sit in lanes 31, 63).
__device__ inline uint64_t wave_min_u64_dpp(uint64_t v, uint32_t lane) {
  v = seg_scan_min_u64(v, lane, lane & 32u);
  const uint64_t a = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), 31) << 32) |
                     (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 31);
  const uint64_t b = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), 63) << 32) |
                     (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  return a < b ? a : b;
}

// Maximum over the whole wavefront (every lane active): row prefix maxima
// with row_shr, then row_bcast:15 / row_bcast:31 carry them to lane 63.
__device__ inline uint32_t wave_max_u32_dpp(uint32_t v) {
  const auto umax = [](uint32_t a, uint32_t b) { return a > b ? a : b; };
  v = umax(v, dpp_mov<0x111, 0xF>(0u, v));
  v = umax(v, dpp_mov<0x112, 0xF>(0u, v));
  v = umax(v, dpp_mov<0x114, 0xF>(0u, v));
  v = umax(v, dpp_mov<0x118, 0xF>(0u, v));
  v = umax(v, dpp_mov<0x142, 0xA>(0u, v));
  v = umax(v, dpp_mov<0x143, 0xC>(0u, v));
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Inclusive prefix sum over the whole wavefront (every lane active): row
// prefixes with row_shr, then row_bcast:15 adds rows 0 / 2 into rows 1 / 3 and
// row_bcast:31 adds rows 0-1 into rows 2-3.
__device__ inline uint32_t wave_scan_u32_dpp(uint32_t v) {
  v += dpp_mov<0x111, 0xF>(0u, v);
  v += dpp_mov<0x112, 0xF>(0u, v);
  v += dpp_mov<0x114, 0xF>(0u, v);
  v += dpp_mov<0x118, 0xF>(0u, v);
  v += dpp_mov<0x142, 0xA>(0u, v);
  v += dpp_mov<0x143, 0xC>(0u, v);
  return v;
}

// Sum over the whole wavefront (every lane active).
__device__ inline uint32_t wave_sum_u32_dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_u32_dpp(v), 63);
}

__device__ inline uint32_t lanemask_lt_popc(uint64_t mask) {
  uint32_t lo = (uint32_t)mask, hi = (uint32_t)(mask >> 32);
  return __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
}

}  // namespace mg
