// Closest-hit traversal over the 32-byte nodes (one BLAS, or a TLAS over
// instanced BLASes) and the ray generators of the path tracer's first bounce.
//
// Hit semantics (tree independent, so trees built by different methods give
// identical hits): Moller-Trumbore exactly as scene.slang:19-89 (cross/dot
// expanded left to right, 1/det correctly rounded, |det| < 1e-6 and t < 1e-6
// rejected); the closest t wins and equal t goes to the lowest original
// (instance, triangle) index.  The reference's Hit.merge
// (intersect.slang:57-82) keeps whichever equal hit was tested later, which
// depends on the tree.  The slab test keeps a box that holds the closest hit
// even when Moller-Trumbore's rounding puts a hit outside the triangle's box.
// That error grows with |o - v1| (the hit is computed from s = o - v1), with
// 1/cos of the angle between the ray and the triangle's plane, and, through
// the edges, with the vertices' magnitude; so each box is padded per axis by
// max(2^-20 x max(|min|, |max|), 2^-16 x D), D the largest |min - o| or
// |max - o| over the three axes.  2^-16 holds on a stress set of far origins
// (distance 30 to 1e7 from a unit scene) with rays nearly parallel to a face,
// where 2^-17 already culls a few closest hits.  Far distances are widened by
// 1 + 2^-20, near distances compared against the closest t widened by
// 1 + 2^-10, and an axis whose plane distance is 0 * inf (a ray inside a face
// plane) left unbounded.
#include <cmath>
#include <cstring>

#include "mgbvh_common.h"
#include "mgbvh_internal.h"

namespace mg {

constexpr float TF_WIDEN = 1.00000095367431640625f;  // 1 + 2^-20
constexpr float BOX_PAD = 9.5367431640625e-07f;      // 2^-20 of the box's coordinate size
constexpr float DIST_PAD = 1.52587890625e-05f;       // 2^-16 of the origin-to-box distance
constexpr float TN_SLACK = 1.0009765625f;            // 1 + 2^-10
constexpr int TRACE_STACK = 64;                      // > tree height + 1
constexpr uint32_t NONE = 0xffffffffu;

__device__ inline float f_inf() { return __uint_as_float(0x7f800000u); }

struct RayD {
  float o[3], d[3], inv[3];
};

struct Best {
  float t;
  uint32_t inst, prim, slot;
};

__device__ inline void set_inv(RayD& r) {
#pragma unroll
  for (int a = 0; a < 3; ++a) r.inv[a] = __fdiv_rn(1.0f, r.d[a]);  // intersect.slang:9
}

__device__ inline bool slab(const RayD& r, const float* bmn, const float* bmx, float best,
                            float& tn_out) {
  float tn = -f_inf(), tf = f_inf();
  float lo[3], hi[3], dist = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = __fsub_rn(bmn[a], r.o[a]);
    hi[a] = __fsub_rn(bmx[a], r.o[a]);
    dist = fmaxf(dist, fmaxf(fabsf(lo[a]), fabsf(hi[a])));
  }
  const float dp = __fmul_rn(dist, DIST_PAD);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float pad = fmaxf(__fmul_rn(fmaxf(fabsf(bmn[a]), fabsf(bmx[a])), BOX_PAD), dp);
    const float t0 = __fmul_rn(__fsub_rn(lo[a], pad), r.inv[a]);
    const float t1 = __fmul_rn(__fadd_rn(hi[a], pad), r.inv[a]);
    // 0 * inf: the ray runs inside a face plane, so this axis does not bound t
    const bool free_axis = (t0 != t0) | (t1 != t1);
    tn = fmaxf(tn, free_axis ? -f_inf() : fminf(t0, t1));
    tf = fminf(tf, free_axis ? f_inf() : fmaxf(t0, t1));
  }
  tf = __fmul_rn(tf, TF_WIDEN);
  tn_out = tn;
  return tn <= tf && tf >= 0.0f && tn <= __fmul_rn(best, TN_SLACK);
}

__device__ inline float dot3_rn(const float* a, const float* b) {
  return __fadd_rn(__fadd_rn(__fmul_rn(a[0], b[0]), __fmul_rn(a[1], b[1])), __fmul_rn(a[2], b[2]));
}

// scene.slang:19-89 (distance only): true and t on a hit.
__device__ inline bool ray_tri_hit(const RayD& r, const float* v1, const float* v2,
                                   const float* v3, float& t_out) {
  float e1[3], e2[3], h[3], s[3], q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) { e1[a] = __fsub_rn(v2[a], v1[a]); e2[a] = __fsub_rn(v3[a], v1[a]); }
  h[0] = __fsub_rn(__fmul_rn(r.d[1], e2[2]), __fmul_rn(r.d[2], e2[1]));
  h[1] = __fsub_rn(__fmul_rn(r.d[2], e2[0]), __fmul_rn(r.d[0], e2[2]));
  h[2] = __fsub_rn(__fmul_rn(r.d[0], e2[1]), __fmul_rn(r.d[1], e2[0]));
  const float det = dot3_rn(h, e1);
  if (fabsf(det) < 1e-6f) return false;
  const float f = __fdiv_rn(1.0f, det);
#pragma unroll
  for (int a = 0; a < 3; ++a) s[a] = __fsub_rn(r.o[a], v1[a]);
  const float u = __fmul_rn(f, dot3_rn(s, h));
  if (u < 0.0f || u > 1.0f) return false;
  q[0] = __fsub_rn(__fmul_rn(s[1], e1[2]), __fmul_rn(s[2], e1[1]));
  q[1] = __fsub_rn(__fmul_rn(s[2], e1[0]), __fmul_rn(s[0], e1[2]));
  q[2] = __fsub_rn(__fmul_rn(s[0], e1[1]), __fmul_rn(s[1], e1[0]));
  const float v = __fmul_rn(f, dot3_rn(r.d, q));
  if (v < 0.0f || __fadd_rn(u, v) > 1.0f) return false;
  const float t = __fmul_rn(f, dot3_rn(e2, q));
  if (t < 1e-6f) return false;
  t_out = t;
  return true;
}

// t of a hit, +inf on a miss.
__device__ inline float ray_tri(const RayD& r, const float* v1, const float* v2, const float* v3) {
  float t;
  return ray_tri_hit(r, v1, v2, v3, t) ? t : f_inf();
}

__device__ inline void load_node(const mgbvh_node32* nodes, uint32_t v, float* bmn, float* bmx,
                                 uint32_t& start, uint32_t& child) {
  const uint4* p = reinterpret_cast<const uint4*>(&nodes[v]);
  const uint4 a = p[0], b = p[1];
  bmn[0] = __uint_as_float(a.x); bmn[1] = __uint_as_float(a.y); bmn[2] = __uint_as_float(a.z);
  start = a.w;
  bmx[0] = __uint_as_float(b.x); bmx[1] = __uint_as_float(b.y); bmx[2] = __uint_as_float(b.z);
  child = b.w;
}

// (t, inst, prim) lexicographic; a miss (+inf) never replaces anything.
__device__ inline bool better(float t, uint32_t inst, uint32_t prim, const Best& b) {
  if (t < b.t) return true;
  if (t != b.t || t == f_inf()) return false;
  return inst < b.inst || (inst == b.inst && prim < b.prim);
}

// Pushes the children of an internal node that pass the slab test, nearer on
// top.  Only node indices are kept: a pop reads the node again and repeats its
// slab test against the closest hit found since (same ray, same box: the same
// entry distance), which costs less than a stack of distances in scratch.
__device__ inline void push_children(const RayD& r, const mgbvh_node32* nodes, uint32_t child,
                                     float best, uint32_t* stk, int& sp, uint32_t& overflow) {
  float amn[3], amx[3], cmn[3], cmx[3], ta, tb;
  uint32_t s0, c0;
  load_node(nodes, child, amn, amx, s0, c0);
  load_node(nodes, child + 1, cmn, cmx, s0, c0);
  const bool ha = slab(r, amn, amx, best, ta);
  const bool hb = slab(r, cmn, cmx, best, tb);
  if (sp + 2 > TRACE_STACK) {
    if (ha || hb) overflow = 1;
    return;
  }
  if (ha && hb) {
    const bool a_near = ta <= tb;
    stk[sp++] = a_near ? child + 1 : child;
    stk[sp++] = a_near ? child : child + 1;
  } else if (ha) {
    stk[sp++] = child;
  } else if (hb) {
    stk[sp++] = child + 1;
  }
}

// One BLAS: nodes/lens/tris/perm already offset to the BLAS's first entries;
// slot0 = the BLAS's triangle offset in the concatenated list.
__device__ inline void trace_blas(const RayD& r, const mgbvh_node32* __restrict__ nodes,
                                  const uint32_t* __restrict__ lens,
                                  const float* __restrict__ pos, uint32_t vs, uint32_t voff,
                                  const uint32_t* __restrict__ tris,
                                  const uint32_t* __restrict__ perm, uint32_t slot0,
                                  uint32_t inst, Best& best, uint32_t& overflow) {
  uint32_t stk[TRACE_STACK];
  int sp = 1;
  float bmn[3], bmx[3], tn;
  uint32_t start, child;
  stk[0] = 0;
  while (sp > 0) {
    const uint32_t v = stk[--sp];
    load_node(nodes, v, bmn, bmx, start, child);
    if (!slab(r, bmn, bmx, best.t, tn)) continue;
    if (child == 0) {
      const uint32_t len = lens ? lens[v] : 1u;
      for (uint32_t k = start; k < start + len; ++k) {
        const uint32_t* tr = tris + 3 * (size_t)k;
        const float t = ray_tri(r, pos + (size_t)(tr[0] + voff) * vs,
                                pos + (size_t)(tr[1] + voff) * vs,
                                pos + (size_t)(tr[2] + voff) * vs);
        const uint32_t id = perm[k];
        if (better(t, inst, id, best)) {
          best.t = t;
          best.inst = inst;
          best.prim = id;
          best.slot = slot0 + k;
        }
      }
    } else {
      push_children(r, nodes, child, best.t, stk, sp, overflow);
    }
  }
}

__device__ inline RayD load_ray(const mgbvh_ray* rays, uint32_t i, float& tmax) {
  const float4* p = reinterpret_cast<const float4*>(&rays[i]);
  const float4 a = p[0], b = p[1];
  RayD r;
  r.o[0] = a.x; r.o[1] = a.y; r.o[2] = a.z;
  r.d[0] = b.x; r.d[1] = b.y; r.d[2] = b.z;
  tmax = b.w;
  set_inv(r);
  return r;
}

__device__ inline void store_hit(mgbvh_hit* hits, uint32_t i, const Best& b) {
  reinterpret_cast<uint4*>(hits)[i] = make_uint4(__float_as_uint(b.t), b.prim, b.inst, b.slot);
}

__global__ __launch_bounds__(256) void k_trace_blas(const mgbvh_node32* __restrict__ nodes,
                                                    const uint32_t* __restrict__ lens,
                                                    const float* __restrict__ pos, uint32_t vs,
                                                    const uint32_t* __restrict__ tris,
                                                    const uint32_t* __restrict__ perm,
                                                    const mgbvh_ray* __restrict__ rays,
                                                    uint32_t nrays, mgbvh_hit* __restrict__ hits,
                                                    uint32_t* overflow_flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrays) return;
  Best b;
  const RayD r = load_ray(rays, i, b.t);
  b.inst = NONE; b.prim = NONE; b.slot = NONE;
  uint32_t overflow = 0;
  trace_blas(r, nodes, lens, pos, vs, 0, tris, perm, 0, 0, b, overflow);
  store_hit(hits, i, b);
  if (overflow) atomicOr(overflow_flag, 1u);
}

// World ray -> object space: rows of the 3x4 world_to_object, summed left to right.
__device__ inline RayD to_object(const RayD& w, const float* m) {
  RayD r;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float* row = m + 4 * i;
    r.o[i] = __fadd_rn(dot3_rn(row, w.o), row[3]);
    r.d[i] = dot3_rn(row, w.d);
  }
  set_inv(r);
  return r;
}

__global__ __launch_bounds__(256) void k_trace_scene(
    const mgbvh_node32* __restrict__ tnodes, const uint32_t* __restrict__ tlens,
    const uint32_t* __restrict__ tperm, const mgbvh_instance* __restrict__ insts,
    const mgbvh_blas_desc* __restrict__ blas, const mgbvh_node32* __restrict__ bnodes,
    const uint32_t* __restrict__ blens, const float* __restrict__ pos, uint32_t vs,
    const uint32_t* __restrict__ tris, const uint32_t* __restrict__ bperm,
    const mgbvh_ray* __restrict__ rays, uint32_t nrays, uint32_t ninst, uint32_t nblas,
    mgbvh_hit* __restrict__ hits, uint32_t* overflow_flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrays) return;
  Best b;
  const RayD r = load_ray(rays, i, b.t);
  b.inst = NONE; b.prim = NONE; b.slot = NONE;
  uint32_t overflow = 0;
  uint32_t stk[TRACE_STACK];
  int sp = 1;
  float bmn[3], bmx[3], tn;
  uint32_t start, child;
  stk[0] = 0;
  while (sp > 0) {
    const uint32_t v = stk[--sp];
    load_node(tnodes, v, bmn, bmx, start, child);
    if (!slab(r, bmn, bmx, b.t, tn)) continue;
    if (child == 0) {
      const uint32_t len = tlens ? tlens[v] : 1u;
      for (uint32_t k = start; k < start + len; ++k) {
        const uint32_t id = tperm[k];
        if (id >= ninst) { overflow |= 2; continue; }
        const mgbvh_instance& in = insts[id];
        if (in.blas >= nblas) { overflow |= 2; continue; }
        const mgbvh_blas_desc bd = blas[in.blas];
        const RayD lr = to_object(r, in.world_to_object);
        trace_blas(lr, bnodes + bd.node_offset, blens ? blens + bd.node_offset : nullptr, pos,
                   vs, bd.vertex_offset, tris + 3 * (size_t)bd.tri_offset, bperm + bd.tri_offset,
                   bd.tri_offset, id, b, overflow);
      }
    } else {
      push_children(r, tnodes, child, b.t, stk, sp, overflow);
    }
  }
  store_hit(hits, i, b);
  if (overflow) atomicOr(overflow_flag, overflow);
}

// ------------------------------------------------ the path tracer's traversal
// scene.slang:91-221 step for step: one 64-entry stack of node payloads
// (objectIndex, childNode) shared by both levels (BVH_STACK_SIZE,
// scene.slang:5-7); children pushed when their box distance is strictly below
// the closest distance, nearer on top; a popped node is not culled again; the
// BLAS root is pushed without a box test; the BLAS pass starts from the world
// distance times length(object-space direction) and its result is merged back
// without conversion; Hit.merge keeps the earlier hit only when strictly
// closer.  min/max of a NaN take the other operand (fminf/fmaxf).
constexpr int SHADER_STACK = 64;
constexpr uint32_t SHADER_NO_TRI = 0xfffffffeu;

struct SHit {  // Hit.success, Hit.distance and the triangle behind them
  float t;
  uint32_t succ, inst, prim, slot;
};

__device__ inline float f_max_finite() { return __uint_as_float(0x7f7fffffu); }

// rayIntersectBoundingVolume, scene.slang:296-310.
__device__ inline bool shader_box(const RayD& r, const float* bmn, const float* bmx, float& dist) {
  float t1[3], t2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lo = __fmul_rn(__fsub_rn(bmn[a], r.o[a]), r.inv[a]);
    const float hi = __fmul_rn(__fsub_rn(bmx[a], r.o[a]), r.inv[a]);
    t1[a] = fminf(lo, hi);
    t2[a] = fmaxf(lo, hi);
  }
  const float tn = fmaxf(fmaxf(t1[0], t1[1]), t1[2]);
  const float tf = fminf(fminf(t2[0], t2[1]), t2[2]);
  const bool front = tn == 0.0f;
  dist = front ? tf : tn;
  return front || (tf >= tn && tf > 0.0f);
}

// Hit.merge, intersect.slang:73-82.
__device__ inline void shader_merge(SHit& a, const SHit& b) {
  const float da = a.succ ? a.t : f_max_finite();
  const float db = b.succ ? b.t : f_max_finite();
  const uint32_t s = a.succ | b.succ;
  if (!(da < db)) a = b;
  a.succ = s;
}

__device__ inline void shader_push(uint32_t* ps, uint32_t* pc, int& sp, uint32_t start,
                                   uint32_t child, uint32_t& err) {
  if (sp >= SHADER_STACK) { err |= 1u; return; }
  ps[sp] = start;
  pc[sp] = child;
  ++sp;
}

// The child step of both loops (scene.slang:112-134, 180-202).
__device__ inline void shader_children(const RayD& r, const mgbvh_node32* nodes, uint32_t child,
                                       const SHit& closest, uint32_t* ps, uint32_t* pc, int& sp,
                                       uint32_t& err) {
  float amn[3], amx[3], bmn[3], bmx[3], da, db;
  uint32_t as, ac, bs, bc;
  load_node(nodes, child, amn, amx, as, ac);
  load_node(nodes, child + 1, bmn, bmx, bs, bc);
  const float cd = closest.succ ? closest.t : f_max_finite();
  const bool pa = shader_box(r, amn, amx, da) && da < cd;
  const bool pb = shader_box(r, bmn, bmx, db) && db < cd;
  if (pa && pb) {
    const bool a_first = da < db;  // the nearer child is pushed last
    shader_push(ps, pc, sp, a_first ? bs : as, a_first ? bc : ac, err);
    shader_push(ps, pc, sp, a_first ? as : bs, a_first ? ac : bc, err);
  } else if (pa || pb) {
    shader_push(ps, pc, sp, pa ? as : bs, pa ? ac : bc, err);
  }
}

__global__ __launch_bounds__(256) void k_raytrace_tlas(
    const mgbvh_node32* __restrict__ tnodes, const uint32_t* __restrict__ tperm,
    const mgbvh_instance* __restrict__ insts, const mgbvh_blas_desc* __restrict__ blas,
    const mgbvh_node32* __restrict__ bnodes, const float* __restrict__ pos, uint32_t vs,
    const uint32_t* __restrict__ tris, const uint32_t* __restrict__ bperm,
    const mgbvh_ray* __restrict__ rays, uint32_t nrays, uint32_t ninst, uint32_t nblas,
    mgbvh_hit* __restrict__ hits, uint32_t* err_flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrays) return;
  float unused_tmax;
  const RayD r = load_ray(rays, i, unused_tmax);
  SHit best = {0.0f, 0u, NONE, NONE, NONE};  // Hit.NO_HIT
  uint32_t err = 0;
  uint32_t ps[SHADER_STACK], pc[SHADER_STACK];
  int sp = 0;
  float bmn[3], bmx[3], dist;
  uint32_t start, child;
  load_node(tnodes, 0, bmn, bmx, start, child);
  if (shader_box(r, bmn, bmx, dist)) shader_push(ps, pc, sp, start, child, err);
  while (sp > 0) {
    --sp;
    const uint32_t s0 = ps[sp], c0 = pc[sp];
    if (c0 != 0) {
      shader_children(r, tnodes, c0, best, ps, pc, sp, err);
      continue;
    }
    // raytraceBlas (scene.slang:143-220) for meshes[s0]
    const uint32_t id = tperm[s0];
    if (id >= ninst) { err |= 2u; continue; }
    const mgbvh_instance& in = insts[id];
    if (in.blas >= nblas) { err |= 2u; continue; }
    const mgbvh_blas_desc bd = blas[in.blas];
    RayD lr;
#pragma unroll
    for (int k = 0; k < 3; ++k) {  // mul(inverseTransform, float4(v, w)), w = 1 and 0
      const float* row = in.world_to_object + 4 * k;
      lr.o[k] = __fadd_rn(dot3_rn(row, r.o), __fmul_rn(row[3], 1.0f));
      lr.d[k] = __fadd_rn(dot3_rn(row, r.d), __fmul_rn(row[3], 0.0f));
    }
    set_inv(lr);
    SHit loc = best;
    loc.t = __fmul_rn(best.t, sqrtf(dot3_rn(lr.d, lr.d)));  // correctly rounded
    const mgbvh_node32* bn = bnodes + bd.node_offset;
    const int base = sp;
    load_node(bn, 0, bmn, bmx, start, child);
    shader_push(ps, pc, sp, start, child, err);
    while (sp > base) {
      --sp;
      const uint32_t s1 = ps[sp], c1 = pc[sp];
      if (c1 != 0) {
        shader_children(lr, bn, c1, loc, ps, pc, sp, err);
      } else {
        const uint32_t slot = bd.tri_offset + s1;
        const uint32_t* tr = tris + 3 * (size_t)slot;
        SHit h = {0.0f, 0u, NONE, NONE, NONE};
        float t;
        if (ray_tri_hit(lr, pos + (size_t)(tr[0] + bd.vertex_offset) * vs,
                        pos + (size_t)(tr[1] + bd.vertex_offset) * vs,
                        pos + (size_t)(tr[2] + bd.vertex_offset) * vs, t)) {
          h.t = t; h.succ = 1u; h.inst = id; h.prim = bperm[slot]; h.slot = slot;
        }
        shader_merge(loc, h);
      }
    }
    shader_merge(best, loc);
  }
  const bool no_tri = best.succ && best.prim == NONE;
  reinterpret_cast<uint4*>(hits)[i] =
      make_uint4(__float_as_uint(best.t), best.succ ? best.prim : NONE,
                 best.succ ? best.inst : NONE, !best.succ ? NONE : no_tri ? SHADER_NO_TRI : best.slot);
  if (err) atomicOr(err_flag, err);
}

// ---------------------------------------------------------------- rays
// PCG hash of random.slang:5-9.
__device__ inline uint32_t pcg(uint32_t& seed) {
  const uint32_t state = seed * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  seed = (word >> 22u) ^ word;
  return seed;
}
__device__ inline float rnd(uint32_t& s) { return (float)pcg(s) / (float)0xffffffffu; }

__device__ inline void normalize3(float* v) {
  const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= l; v[1] /= l; v[2] /= l;
}

__device__ inline void store_ray(mgbvh_ray* rays, size_t i, const float* o, const float* d) {
  float4* p = reinterpret_cast<float4*>(&rays[i]);
  p[0] = make_float4(o[0], o[1], o[2], 0.0f);
  p[1] = make_float4(d[0], d[1], d[2], f_inf());
}

// The path tracer's primary ray (pathtrace.slang:46-66, screen.slang:48-72):
// texcoord = (pixel + jitter) / (width, height), clip = (2 tx - 1, 1 - 2 ty, 1)
// (screen depth 1), world = (M clip).xyz / (M clip).w with M the inverse
// view-projection matrix (column-major, as glam stores it), direction =
// normalize(world - position).  Every operation is rounded once, in the
// order written (sums left to right, no FMA, correctly rounded division and
// square root), so the rays are reproducible bit for bit.
struct ViewParams {
  float m[16];
  float pos[3];
  float jx, jy;
};

__global__ void k_gen_view(ViewParams vp, uint32_t w, uint32_t h, mgbvh_ray* rays) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y;
  if (x >= w || y >= h) return;
  const float tx = __fdiv_rn(__fadd_rn((float)x, vp.jx), (float)w);
  const float ty = __fdiv_rn(__fadd_rn((float)y, vp.jy), (float)h);
  const float cl[3] = {__fsub_rn(__fmul_rn(tx, 2.0f), 1.0f), __fsub_rn(1.0f, __fmul_rn(ty, 2.0f)),
                       1.0f};
  float t[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    t[i] = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(vp.m[i], cl[0]), __fmul_rn(vp.m[4 + i], cl[1])),
                               __fmul_rn(vp.m[8 + i], cl[2])),
                     vp.m[12 + i]);
  float d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) d[a] = __fsub_rn(__fdiv_rn(t[a], t[3]), vp.pos[a]);
  // sqrtf, not __fsqrt_rn: without OCML_BASIC_ROUNDED_OPERATIONS the latter is
  // the native (approximate) square root; sqrtf is correctly rounded
  const float l = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(d[0], d[0]), __fmul_rn(d[1], d[1])),
                                  __fmul_rn(d[2], d[2])));
#pragma unroll
  for (int a = 0; a < 3; ++a) d[a] = __fdiv_rn(d[a], l);
  store_ray(rays, (size_t)y * w + x, vp.pos, d);
}

// random.slang:29-41 getUnitVector / getCosineVector; seeded as pathtrace.slang:49.
__global__ void k_gen_bounce(const mgbvh_ray* __restrict__ rin, const mgbvh_hit* __restrict__ hits,
                             uint32_t n, const mgbvh_instance* __restrict__ insts,
                             const mgbvh_blas_desc* __restrict__ blas,
                             const float* __restrict__ pos, uint32_t vs,
                             const uint32_t* __restrict__ tris, uint32_t seed,
                             mgbvh_ray* __restrict__ rout) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const mgbvh_ray q = rin[i];
  const mgbvh_hit hh = hits[i];
  uint32_t s = i + seed;
  const float xy0 = rnd(s) * 6.283185307179586f;
  const float xy1 = 2.0f * rnd(s) - 1.0f;
  const float rr = sqrtf(fmaxf(0.0f, 1.0f - xy1 * xy1));
  const float uvec[3] = {sinf(xy0) * rr, cosf(xy0) * rr, xy1};
  float o[3], d[3];
  if (hh.prim != NONE) {
    uint32_t voff = 0;
    const float* m = nullptr;
    if (insts) {
      m = insts[hh.inst].world_to_object;
      voff = blas[insts[hh.inst].blas].vertex_offset;
    }
    const uint32_t* tr = tris + 3 * (size_t)hh.tri_slot;
    const float* v1 = pos + (size_t)(tr[0] + voff) * vs;
    const float* v2 = pos + (size_t)(tr[1] + voff) * vs;
    const float* v3 = pos + (size_t)(tr[2] + voff) * vs;
    float e1[3], e2[3], nl[3], nrm[3];
    for (int a = 0; a < 3; ++a) { e1[a] = v2[a] - v1[a]; e2[a] = v3[a] - v1[a]; }
    nl[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nl[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nl[2] = e1[0] * e2[1] - e1[1] * e2[0];
    if (m) {  // object normal to world: transpose of world_to_object's linear part
      for (int a = 0; a < 3; ++a) nrm[a] = m[a] * nl[0] + m[4 + a] * nl[1] + m[8 + a] * nl[2];
    } else {
      for (int a = 0; a < 3; ++a) nrm[a] = nl[a];
    }
    normalize3(nrm);
    if (nrm[0] * q.dir[0] + nrm[1] * q.dir[1] + nrm[2] * q.dir[2] > 0.0f)
      for (int a = 0; a < 3; ++a) nrm[a] = -nrm[a];
    for (int a = 0; a < 3; ++a) {
      o[a] = q.org[a] + hh.t * q.dir[a] + 0.001f * nrm[a];
      d[a] = nrm[a] + uvec[a];
    }
    normalize3(d);
  } else {
    for (int a = 0; a < 3; ++a) { o[a] = q.org[a]; d[a] = uvec[a]; }
  }
  store_ray(rout, i, o, d);
}

// Host copy of the PCG stream (random.slang:5-27).
uint32_t pcg_host(uint32_t& seed) {
  const uint32_t state = seed * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  seed = (word >> 22u) ^ word;
  return seed;
}
float pcg_float(uint32_t& s) { return (float)pcg_host(s) / (float)0xffffffffu; }

// inverse(perspective_rh(fov_y, aspect, 0.1, 100) * look_at_rh) of a pinhole
// camera given by its position and orthonormal basis, column-major floats.
void pinhole_inverse_view_proj(const mgbvh_camera& cam, float* out16) {
  const double n = 0.1, f = 100.0;
  const double th = std::tan(0.5 * (double)cam.fov_y);
  // inverse projection (view-space point of a clip point, then / w)
  double ip[4][4] = {{th * cam.aspect, 0, 0, 0}, {0, th, 0, 0}, {0, 0, 0, -1},
                     {0, 0, (n - f) / (f * n), 1.0 / n}};
  // inverse view: view space (x right, y up, -z forward) to world
  double iv[4][4] = {};
  for (int r = 0; r < 3; ++r) {
    iv[r][0] = cam.right[r];
    iv[r][1] = cam.up[r];
    iv[r][2] = -(double)cam.fwd[r];
    iv[r][3] = cam.pos[r];
  }
  iv[3][3] = 1.0;
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) {
      double acc = 0.0;
      for (int j = 0; j < 4; ++j) acc += iv[r][j] * ip[j][k];
      out16[4 * k + r] = (float)acc;  // column k, row r
    }
}

int gen_view(mgbvh_ctx* c, const ViewParams& vp, uint32_t width, uint32_t height,
             mgbvh_ray* d_rays) {
  if (!width || !height) return MGBVH_OK;
  if ((uint64_t)width * height > 0xffffffffull || height > 65535) {
    set_error("image larger than 2^32-1 pixels or 65535 rows");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  k_gen_view<<<dim3((width + 255) / 256, height), 256, 0, c->stream>>>(vp, width, height, d_rays);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "k_gen_view");
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_gen_view");
}

int launch_done(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, what);
}

// Reads the overflow flag after the stream drains.
int finish_trace(mgbvh_ctx* c, uint32_t* flag,
                 const char* overflow = "traversal stack overflow (tree height above 62)") {
  hipError_t e = hipMemcpyAsync(&c->h_ctr->err, flag, 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  int rc = prof_collect(c);
  if (rc) return rc;
  if (c->h_ctr->err & 2u) {
    set_error("a TLAS leaf names an instance >= ninstances or an instance names a BLAS >= nblas");
    return MGBVH_ERR_INDEX;
  }
  if (c->h_ctr->err) {
    set_error(overflow);
    return MGBVH_ERR_INTERNAL;
  }
  return MGBVH_OK;
}

int trace_begin(mgbvh_ctx* c, const void* rays, const void* hits, uint64_t nrays,
                uint32_t vstride_bytes, uint32_t** flag) {
  if (!rays || !hits) { set_error("null ray or hit buffer"); return MGBVH_ERR_INVALID_ARG; }
  if (vstride_bytes < 12 || vstride_bytes % 4 || nrays > 0xffffffffull) {
    set_error("bad vstride_bytes or more than 2^32-1 rays");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  int rc = ensure(c->sc.ctr, sizeof(Counters));
  if (rc) return rc;
  *flag = reinterpret_cast<uint32_t*>(c->sc.ctr.p);
  e = hipMemsetAsync(*flag, 0, 4, c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
  return MGBVH_OK;
}

#define MG_TRACE_PROF(end)                                                   \
  do {                                                                       \
    if (c->prof_mask & (1u << KK_TRACE)) {                                   \
      int r_ = (end) ? prof_end(c, KK_TRACE) : prof_begin(c, KK_TRACE);     \
      if (r_) return r_;                                                     \
    }                                                                        \
  } while (0)

}  // namespace mg

using namespace mg;

extern "C" {

int mgbvh_trace_closest_device(mgbvh_ctx* c, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                               const float* d_pos, uint32_t vstride_bytes,
                               const uint32_t* d_tris_perm, const uint32_t* d_perm,
                               const mgbvh_ray* d_rays, uint64_t nrays, mgbvh_hit* d_hits) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (nrays == 0) return MGBVH_OK;
  if (!d_nodes || !d_pos || !d_tris_perm || !d_perm) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  uint32_t* flag = nullptr;
  int rc = trace_begin(c, d_rays, d_hits, nrays, vstride_bytes, &flag);
  if (rc) return rc;
  MG_TRACE_PROF(0);
  k_trace_blas<<<(uint32_t)((nrays + 255) / 256), 256, 0, c->stream>>>(
      d_nodes, d_lens, d_pos, vstride_bytes / 4, d_tris_perm, d_perm, d_rays, (uint32_t)nrays,
      d_hits, flag);
  MG_TRACE_PROF(1);
  if ((rc = launch_done("k_trace_blas"))) return rc;
  return finish_trace(c, flag);
}

int mgbvh_trace_scene_device(mgbvh_ctx* c, const mgbvh_node32* d_tlas_nodes,
                             const uint32_t* d_tlas_lens, const uint32_t* d_tlas_perm,
                             const mgbvh_instance* d_instances, uint32_t ninstances,
                             const mgbvh_blas_desc* d_blas, uint32_t nblas,
                             const mgbvh_node32* d_blas_nodes, const uint32_t* d_blas_lens,
                             const float* d_pos, uint32_t vstride_bytes,
                             const uint32_t* d_tris_perm, const uint32_t* d_blas_perm,
                             const mgbvh_ray* d_rays, uint64_t nrays, mgbvh_hit* d_hits) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (nrays == 0) return MGBVH_OK;
  if (ninstances == 0 || nblas == 0) {
    set_error("a scene needs at least one instance and one BLAS");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (!d_tlas_nodes || !d_tlas_perm || !d_instances || !d_blas || !d_blas_nodes || !d_pos ||
      !d_tris_perm || !d_blas_perm) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  uint32_t* flag = nullptr;
  int rc = trace_begin(c, d_rays, d_hits, nrays, vstride_bytes, &flag);
  if (rc) return rc;
  MG_TRACE_PROF(0);
  k_trace_scene<<<(uint32_t)((nrays + 255) / 256), 256, 0, c->stream>>>(
      d_tlas_nodes, d_tlas_lens, d_tlas_perm, d_instances, d_blas, d_blas_nodes, d_blas_lens,
      d_pos, vstride_bytes / 4, d_tris_perm, d_blas_perm, d_rays, (uint32_t)nrays, ninstances,
      nblas, d_hits, flag);
  MG_TRACE_PROF(1);
  if ((rc = launch_done("k_trace_scene"))) return rc;
  return finish_trace(c, flag);
}

int mgbvh_raytrace_tlas_device(mgbvh_ctx* c, const mgbvh_node32* d_tlas_nodes,
                               const uint32_t* d_tlas_perm, const mgbvh_instance* d_instances,
                               uint32_t ninstances, const mgbvh_blas_desc* d_blas, uint32_t nblas,
                               const mgbvh_node32* d_blas_nodes, const float* d_pos,
                               uint32_t vstride_bytes, const uint32_t* d_tris_perm,
                               const uint32_t* d_blas_perm, const mgbvh_ray* d_rays,
                               uint64_t nrays, mgbvh_hit* d_hits) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (nrays == 0) return MGBVH_OK;
  if (ninstances == 0 || nblas == 0) {
    set_error("a scene needs at least one instance and one BLAS");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (!d_tlas_nodes || !d_tlas_perm || !d_instances || !d_blas || !d_blas_nodes || !d_pos ||
      !d_tris_perm || !d_blas_perm) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  uint32_t* flag = nullptr;
  int rc = trace_begin(c, d_rays, d_hits, nrays, vstride_bytes, &flag);
  if (rc) return rc;
  MG_TRACE_PROF(0);
  k_raytrace_tlas<<<(uint32_t)((nrays + 255) / 256), 256, 0, c->stream>>>(
      d_tlas_nodes, d_tlas_perm, d_instances, d_blas, d_blas_nodes, d_pos, vstride_bytes / 4,
      d_tris_perm, d_blas_perm, d_rays, (uint32_t)nrays, ninstances, nblas, d_hits, flag);
  MG_TRACE_PROF(1);
  if ((rc = launch_done("k_raytrace_tlas"))) return rc;
  return finish_trace(c, flag, "the 64-entry traversal stack of the path tracer overflowed");
}

int mgbvh_gen_view_rays_device(mgbvh_ctx* c, const float* inv_view_proj, const float* position,
                               uint32_t width, uint32_t height, uint32_t frame_count,
                               int accumulate, mgbvh_ray* d_rays) {
  if (!c || !inv_view_proj || !position || !d_rays) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  ViewParams vp;
  std::memcpy(vp.m, inv_view_proj, sizeof vp.m);
  std::memcpy(vp.pos, position, sizeof vp.pos);
  // pathtrace.slang:50-61: the frame's jitter, two draws of Random(frameCount)
  // minus 0.5, applied only while the image accumulates
  uint32_t s = frame_count;
  const float j0 = pcg_float(s) - 0.5f;
  const float j1 = pcg_float(s) - 0.5f;
  vp.jx = accumulate ? j0 : 0.0f;
  vp.jy = accumulate ? j1 : 0.0f;
  return gen_view(c, vp, width, height, d_rays);
}

int mgbvh_gen_primary_rays_device(mgbvh_ctx* c, const mgbvh_camera* cam, uint32_t width,
                                  uint32_t height, mgbvh_ray* d_rays) {
  if (!c || !cam || !d_rays) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  // the pinhole camera as glam's look_at_rh and perspective_rh (depth 0..1,
  // near 0.1, far 100) would give it, inverted in double; a jitter of +0.5
  // puts every ray through its pixel's centre
  ViewParams vp;
  pinhole_inverse_view_proj(*cam, vp.m);
  std::memcpy(vp.pos, cam->pos, sizeof vp.pos);
  vp.jx = 0.5f;
  vp.jy = 0.5f;
  return gen_view(c, vp, width, height, d_rays);
}

int mgbvh_gen_bounce_rays_device(mgbvh_ctx* c, const mgbvh_ray* d_rays_in,
                                 const mgbvh_hit* d_hits, uint64_t nrays,
                                 const mgbvh_instance* d_instances, uint32_t ninstances,
                                 const mgbvh_blas_desc* d_blas, uint32_t nblas,
                                 const float* d_pos, uint32_t vstride_bytes,
                                 const uint32_t* d_tris_perm, uint32_t seed,
                                 mgbvh_ray* d_rays_out) {
  if (!c || !d_rays_in || !d_hits || !d_pos || !d_tris_perm || !d_rays_out) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  if ((d_instances == nullptr) != (d_blas == nullptr) ||
      (d_instances && (ninstances == 0 || nblas == 0))) {
    set_error("pass both instances and BLAS table (non-empty) or neither");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (nrays == 0) return MGBVH_OK;
  if (vstride_bytes < 12 || vstride_bytes % 4 || nrays > 0xffffffffull) {
    set_error("bad vstride_bytes or more than 2^32-1 rays");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  k_gen_bounce<<<(uint32_t)((nrays + 255) / 256), 256, 0, c->stream>>>(
      d_rays_in, d_hits, (uint32_t)nrays, d_instances, d_blas, d_pos, vstride_bytes / 4,
      d_tris_perm, seed, d_rays_out);
  int rc = launch_done("k_gen_bounce");
  if (rc) return rc;
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_gen_bounce");
}

}  // extern "C"
