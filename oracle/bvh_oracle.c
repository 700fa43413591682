/*
 * CPU oracle for the marigold `bvh` crate -- TEST INFRASTRUCTURE ONLY.
 *
 * Op-for-op C restatement of `bvh/src/lib.rs` (the reference is Rust and cannot
 * be compiled in this image: no cargo/rustc, crates not vendored).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
 * library, and only as the checker / CPU baseline, never as the product.
 *
 * Semantics pinned to the reference (file:line in bvh/src/lib.rs unless noted):
 *   AABB algebra .................. :61-111
 *   binned sweep (32 bins) ........ :331-389, cost :255-290
 *   adaptive threshold sweep ...... :391-442, cost :292-329
 *   median fallback ............... :444-491
 *   select_split .................. :493-534
 *   split / node order ............ :536-665
 *   new() ......................... :698-878
 *
 * Deliberate, documented choices where Rust's behaviour is not pinned:
 *   - `select_nth_unstable_by` and `partition_in_place` leave an unspecified
 *     order inside each side.  The oracle uses the canonical STABLE variant
 *     (ties by current list position).  Choices that depend on that order are
 *     counted in or_stats.n_tie_events so tests can assert they never happen.
 *   - glam min/max on +-0 is operand-order dependent; triangle AABBs map -0.0
 *     to +0.0 so that every min/max fold is order independent.
 *   - NaN costs follow x86-64 SSE: an invalid operation yields the default
 *     NaN 0xFFC00000, which f32::total_cmp orders below -inf.
 */
#include "bvh_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* OR_PLANT selects a deliberately wrong variant of one formula.  Only the
 * tests build such variants (into separate libraries) to show that their
 * checks fail on a plausible mistake; the oracle proper is built with 0.
 *   1: adaptive sweep with S = 32 thresholds for every n (lib.rs:430)
 *   2: adaptive thresholds without the + 0.5 (lib.rs:434)
 *   3: tree SAH leaf term without len
 *   4: tree SAH internal term weighted by OBJECT_COST instead of DEPTH_COST
 *   5: traversal box pad without the origin-distance term
 *   6, 7: ray generation variants (rays_oracle.c) */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

#define DEPTH_COST 1.0f  /* lib.rs:217 */
#define OBJECT_COST 8.0f /* lib.rs:218 */

static inline float f_min(float a, float b) { return a < b ? a : b; } /* minps */
static inline float f_max(float a, float b) { return a > b ? a : b; } /* maxps */

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* f32::total_cmp key: i ^ (((i >> 31) as u32) >> 1) as i32, compared signed. */
int32_t or_total_cmp_key(float f) {
  int32_t i = (int32_t)fbits(f);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}

static inline float x86_nan_canon(float f) { return f != f ? bitsf(0xFFC00000u) : f; }

static const or_bv BV_EMPTY = {{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};

/* lib.rs:71-73 */
static inline float bv_center(const or_bv* b, int a) { return (b->mn[a] + b->mx[a]) * 0.5f; }
/* lib.rs:75-77 */
static inline float bv_extent(const or_bv* b, int a) { return b->mx[a] - b->mn[a]; }
/* lib.rs:103-106 */
static inline int bv_is_empty(const or_bv* b) {
  return (b->mx[0] < b->mn[0]) | (b->mx[1] < b->mn[1]) | (b->mx[2] < b->mn[2]);
}
/* lib.rs:79-91: 2 * ((w*h + w*d) + h*d), no FMA. */
static float bv_sa(const or_bv* b) {
  float w = b->mx[0] - b->mn[0];
  float h = b->mx[1] - b->mn[1];
  float d = b->mx[2] - b->mn[2];
  if (!bv_is_empty(b)) {
    float wh = w * h;
    float wd = w * d;
    float hd = h * d;
    float s = wh + wd;
    s = s + hd;
    return 2.0f * s;
  }
  return 0.0f;
}
/* lib.rs:98-101: self.min = self.min.min(other.min) ... */
static inline void bv_grow(or_bv* acc, const or_bv* o) {
  for (int a = 0; a < 3; ++a) {
    acc->mn[a] = f_min(acc->mn[a], o->mn[a]);
    acc->mx[a] = f_max(acc->mx[a], o->mx[a]);
  }
}
/* lib.rs:108-110 */
static int bv_contains(const or_bv* b, const float* p) {
  if (bv_is_empty(b)) return 0;
  for (int a = 0; a < 3; ++a)
    if (!(p[a] >= b->mn[a]) || !(p[a] <= b->mx[a])) return 0;
  return 1;
}

float or_surface_area(const float* bv6) {
  or_bv b;
  memcpy(&b, bv6, sizeof b);
  return bv_sa(&b);
}

/* ((SA_lt / SA_p) * 8) * n_lt ; cost = (1 + lt) + gt   (lib.rs:278-288) */
static float split_cost(const or_bv* parent, const or_bv* lt, const or_bv* gt, uint32_t n_lt,
                        uint32_t n_gt) {
  float sp = bv_sa(parent);
  float lt_cost = bv_sa(lt) / sp;
  lt_cost = lt_cost * OBJECT_COST;
  lt_cost = lt_cost * (float)n_lt;
  float gt_cost = bv_sa(gt) / sp;
  gt_cost = gt_cost * OBJECT_COST;
  gt_cost = gt_cost * (float)n_gt;
  float c = DEPTH_COST + lt_cost;
  c = c + gt_cost;
  return x86_nan_canon(c);
}

float or_binned_cost(const float* p6, const float* l6, const float* g6, uint32_t n_lt,
                     uint32_t n_gt) {
  or_bv p, l, g;
  memcpy(&p, p6, sizeof p);
  memcpy(&l, l6, sizeof l);
  memcpy(&g, g6, sizeof g);
  return split_cost(&p, &l, &g, n_lt, n_gt);
}

/* `(x).floor() as usize` saturating cast, then .min(B-1)   (lib.rs:358-362) */
uint32_t or_bin_index(float c, float cmin, float cext, uint32_t nbins) {
  float pct = (c - cmin) / cext;
  float f = floorf(pct * (float)nbins);
  uint64_t bi;
  if (!(f > 0.0f)) bi = 0; /* NaN and negatives saturate to 0 */
  else if (f >= 18446744073709551616.0f) bi = UINT64_MAX;
  else bi = (uint64_t)f;
  if (bi > nbins - 1) bi = nbins - 1;
  return (uint32_t)bi;
}

/* ---------------------------------------------------------------- AABBs */

static inline float canon_zero(float x) { return x == 0.0f ? 0.0f : x; }

void or_tri_aabbs(const float* pos, uint32_t vs, const uint32_t* tris, uint64_t n, float* out) {
  for (uint64_t t = 0; t < n; ++t) {
    const float* v1 = pos + (uint64_t)tris[3 * t + 0] * vs;
    const float* v2 = pos + (uint64_t)tris[3 * t + 1] * vs;
    const float* v3 = pos + (uint64_t)tris[3 * t + 2] * vs;
    for (int a = 0; a < 3; ++a) {
      /* v1.min(v2).min(v3), mesh_interface/src/lib.rs:25-26 */
      float mn = f_min(f_min(v1[a], v2[a]), v3[a]);
      float mx = f_max(f_max(v1[a], v2[a]), v3[a]);
      out[6 * t + a] = canon_zero(mn);
      out[6 * t + 3 + a] = canon_zero(mx);
    }
  }
}

/* ---------------------------------------------------------------- builder */

typedef struct {
  const or_bv* aabb; /* per object id */
  uint32_t* list;    /* object ids */
  or_node* nodes;
  uint32_t nnodes;
  or_rec* recs;
  or_stats* st;
  uint32_t min_leaf, max_leaf, max_depth;
  uint32_t* tmp;     /* scratch, n entries */
  uint64_t* sortkey; /* scratch, n entries */
} ctx_t;

typedef struct {
  or_bv blt, bgt;
  uint32_t n_lt, n_gt;
  float cost;
  int kind; /* 0 = Index (binned), 1 = Threshold (adaptive), 2 = Exact (median) */
  int axis;
  uint32_t value; /* bin cut i / threshold bits / median index */
  float threshold;
  uint32_t index;
} split_t;

static inline void obj_center(const ctx_t* c, uint32_t id, float* ctr) {
  const or_bv* b = &c->aabb[id];
  for (int a = 0; a < 3; ++a) ctr[a] = bv_center(b, a);
}

static inline int better(const split_t* cand, const split_t* best, int have_best) {
  /* Iterator::min_by keeps the first of equal minima (strictly less replaces). */
  return !have_best || or_total_cmp_key(cand->cost) < or_total_cmp_key(best->cost);
}

/* evaluate_binned_split, lib.rs:255-290 */
static void eval_binned(const or_bv* parent, const uint32_t* cnt, const or_bv* bb, uint32_t B,
                        uint32_t i, split_t* s) {
  s->blt = BV_EMPTY;
  s->bgt = BV_EMPTY;
  uint32_t nl = 0, ng = 0;
  for (uint32_t k = 0; k < i; ++k) { nl += cnt[k]; bv_grow(&s->blt, &bb[k]); }
  for (uint32_t k = i; k < B; ++k) { ng += cnt[k]; bv_grow(&s->bgt, &bb[k]); }
  s->n_lt = nl;
  s->n_gt = ng;
  s->cost = split_cost(parent, &s->blt, &s->bgt, nl, ng);
}

/* binned_sweep, lib.rs:331-389 */
static int binned_sweep(ctx_t* c, const or_bv* parent, const or_bv* cb, const uint32_t* list,
                        uint32_t n, int axis, split_t* out) {
  if (bv_extent(cb, axis) == 0.0f) return 0;
  uint32_t B = n < 16 ? 16 : (n > 32 ? 32 : n);
  uint32_t cnt[32];
  or_bv bb[32];
  for (uint32_t k = 0; k < B; ++k) { cnt[k] = 0; bb[k] = BV_EMPTY; }
  for (uint32_t k = 0; k < n; ++k) {
    const or_bv* ob = &c->aabb[list[k]];
    float ctr = bv_center(ob, axis);
    uint32_t bi = or_bin_index(ctr, cb->mn[axis], bv_extent(cb, axis), B);
    cnt[bi] += 1;
    bv_grow(&bb[bi], ob);
  }
  int have = 0;
  for (uint32_t i = 1; i < B; ++i) {
    split_t s;
    eval_binned(parent, cnt, bb, B, i, &s);
    if (s.n_lt < c->min_leaf || s.n_gt < c->min_leaf) continue;
    s.kind = 0;
    s.axis = axis;
    s.value = i;
    s.index = s.n_lt;
    if (better(&s, out, have)) { *out = s; have = 1; }
  }
  return have;
}

/* evaluate_threshold_split, lib.rs:292-329 */
static void eval_threshold(const ctx_t* c, const or_bv* parent, const uint32_t* list, uint32_t n,
                           int axis, float t, split_t* s) {
  s->blt = BV_EMPTY;
  s->bgt = BV_EMPTY;
  uint32_t nl = 0, ng = 0;
  for (uint32_t k = 0; k < n; ++k) {
    const or_bv* ob = &c->aabb[list[k]];
    if (bv_center(ob, axis) < t) { bv_grow(&s->blt, ob); nl++; }
    else { bv_grow(&s->bgt, ob); ng++; }
  }
  s->n_lt = nl;
  s->n_gt = ng;
  s->cost = split_cost(parent, &s->blt, &s->bgt, nl, ng);
}

/* adaptive_sweep, lib.rs:391-442.  *tie is set when the first-minimum choice
 * depended on list order (an equal-cost candidate with a different partition). */
static int adaptive_sweep(ctx_t* c, const or_bv* parent, const or_bv* cb, const uint32_t* list,
                          uint32_t n, int axis, split_t* out, int* tie) {
  const uint32_t accurate = 8;
  int have = 0;
  split_t cands[32];
  uint32_t nc = 0;
  if (n < accurate) {
    for (uint32_t k = 0; k < n; ++k) {
      float t = bv_center(&c->aabb[list[k]], axis);
      split_t s;
      eval_threshold(c, parent, list, n, axis, t, &s);
      if (s.n_lt < c->min_leaf || s.n_gt < c->min_leaf) continue;
      s.kind = 1; s.axis = axis; s.threshold = t; s.value = fbits(t);
      cands[nc++] = s;
      if (better(&s, out, have)) { *out = s; have = 1; }
    }
    if (have) {
      for (uint32_t k = 0; k < nc; ++k)
        if (fbits(cands[k].cost) == fbits(out->cost) && cands[k].n_lt != out->n_lt) *tie = 1;
    }
  } else {
    uint32_t steps = n < accurate ? accurate : (n > 32 ? 32 : n);
    if (OR_PLANT == 1) steps = 32;
    float step = bv_extent(cb, axis) / (float)steps;
    for (uint32_t i = 0; i < steps; ++i) {
      float fi = (float)i + (OR_PLANT == 2 ? 0.0f : 0.5f);
      float off = step * fi;
      float t = cb->mn[axis] + off;
      split_t s;
      eval_threshold(c, parent, list, n, axis, t, &s);
      if (s.n_lt < c->min_leaf || s.n_gt < c->min_leaf) continue;
      s.kind = 1; s.axis = axis; s.threshold = t; s.value = fbits(t);
      if (better(&s, out, have)) { *out = s; have = 1; }
    }
  }
  return have;
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* Canonical stable select_nth_unstable_by(m, total_cmp(center[axis])):
 * the m smallest (ties by current position) go first, both sides keep their
 * relative order.  Returns 1 when keys tie across the m boundary. */
static int stable_select(ctx_t* c, uint32_t* list, uint32_t n, uint32_t m, int axis) {
  if (n == 0) return 0;
  for (uint32_t k = 0; k < n; ++k) {
    float ctr = bv_center(&c->aabb[list[k]], axis);
    uint32_t key = (uint32_t)or_total_cmp_key(ctr) ^ 0x80000000u;
    c->sortkey[k] = ((uint64_t)key << 32) | k;
  }
  qsort(c->sortkey, n, sizeof(uint64_t), cmp_u64);
  int tie = 0;
  if (m > 0 && m < n && (c->sortkey[m - 1] >> 32) == (c->sortkey[m] >> 32)) {
    /* A tie across the boundary matters only if the tied objects' boxes differ;
     * identical boxes give the same children whichever goes left. */
    uint32_t kb = (uint32_t)(c->sortkey[m] >> 32);
    const or_bv* ref = NULL;
    tie = 2; /* benign unless proven otherwise */
    for (uint32_t k = 0; k < n; ++k) {
      if ((uint32_t)(c->sortkey[k] >> 32) != kb) continue;
      const or_bv* b = &c->aabb[list[(uint32_t)(c->sortkey[k] & 0xffffffffu)]];
      if (!ref) ref = b;
      else if (memcmp(ref, b, sizeof *b) != 0) { tie = 1; break; }
    }
  }
  /* mark: tmp[pos] = 1 if selected */
  for (uint32_t k = 0; k < n; ++k) c->tmp[k] = 0;
  for (uint32_t k = 0; k < m; ++k) c->tmp[(uint32_t)(c->sortkey[k] & 0xffffffffu)] = 1;
  /* stable partition into sortkey (reuse as id buffer) */
  uint32_t w = 0;
  uint32_t* ids = (uint32_t*)c->sortkey; /* n u64 entries hold >= n u32 ids */
  uint32_t* flags = c->tmp;
  uint32_t* out2 = ids + n;              /* second half of the u64 scratch */
  for (uint32_t k = 0; k < n; ++k) out2[k] = flags[k];
  for (uint32_t k = 0; k < n; ++k) if (out2[k]) ids[w++] = list[k];
  for (uint32_t k = 0; k < n; ++k) if (!out2[k]) ids[w++] = list[k];
  memcpy(list, ids, n * sizeof(uint32_t));
  return tie;
}

/* median_split, lib.rs:444-491 (mutates the list) */
static void median_split(ctx_t* c, const or_bv* parent, uint32_t* list, uint32_t n, int axis,
                         split_t* s, int* tie) {
  uint32_t m = n / 2;
  int t = stable_select(c, list, n, m, axis);
  if (t == 1) *tie = 1;
  else if (t == 2 && *tie == 0) *tie = 2;
  s->blt = BV_EMPTY;
  s->bgt = BV_EMPTY;
  for (uint32_t k = 0; k < m; ++k) bv_grow(&s->blt, &c->aabb[list[k]]);
  for (uint32_t k = m; k < n; ++k) bv_grow(&s->bgt, &c->aabb[list[k]]);
  s->n_lt = m;
  s->n_gt = n - m;
  /* lib.rs:470-482: the lengths come from the slices */
  s->cost = split_cost(parent, &s->blt, &s->bgt, m, n - m);
  s->kind = 2;
  s->axis = axis;
  s->value = m;
  s->index = m;
}

/* select_split, lib.rs:493-534 */
static int select_split(ctx_t* c, const or_bv* parent, uint32_t* list, uint32_t n, split_t* best,
                        int* median_quirk) {
  or_bv cb;
  for (int a = 0; a < 3; ++a) { cb.mn[a] = INFINITY; cb.mx[a] = -INFINITY; }
  for (uint32_t k = 0; k < n; ++k) {
    float ctr[3];
    obj_center(c, list[k], ctr);
    for (int a = 0; a < 3; ++a) {
      cb.mn[a] = f_min(cb.mn[a], ctr[a]);
      cb.mx[a] = f_max(cb.mx[a], ctr[a]);
    }
  }
  int have = 0;
  int tie_axis[3] = {0, 0, 0};
  for (int axis = 0; axis < 3; ++axis) {
    split_t s;
    int ok;
    if (n <= 32) ok = adaptive_sweep(c, parent, &cb, list, n, axis, &s, &tie_axis[axis]);
    else ok = binned_sweep(c, parent, &cb, list, n, axis, &s);
    if (ok && better(&s, best, have)) { *best = s; have = 1; }
  }
  if (have && best->kind == 1 && tie_axis[best->axis]) c->st->n_tie_events++;
  *median_quirk = 0;
  if (!have && n > c->max_leaf) {
    /* each call reorders the list; the final list is the axis-2 one (R11) */
    uint32_t* memb[3];
    uint32_t* snap = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)n);
    int tie = 0;
    for (int axis = 0; axis < 3; ++axis) {
      split_t s;
      median_split(c, parent, list, n, axis, &s, &tie);
      memb[axis] = snap + (size_t)axis * n;
      memcpy(memb[axis], list, sizeof(uint32_t) * n);
      if (better(&s, best, have)) { *best = s; have = 1; }
    }
    if (tie == 1) c->st->n_tie_events++;
    else if (tie == 2) c->st->n_tie_benign++;
    /* Does the winning axis' lt set differ from the final (axis 2) one? */
    if (best->axis != 2) {
      uint32_t m = n / 2;
      /* sets differ iff some element of memb[best][0..m) is not in memb[2][0..m) */
      for (uint32_t k = 0; k < n; ++k) c->tmp[k] = 0;
      /* map ids -> mark via sort of copies */
      uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)m + 1);
      uint32_t* b2 = a + m;
      memcpy(a, memb[best->axis], m * sizeof(uint32_t));
      memcpy(b2, memb[2], m * sizeof(uint32_t));
      for (uint32_t k = 0; k < m; ++k) { c->sortkey[k] = a[k]; }
      qsort(c->sortkey, m, sizeof(uint64_t), cmp_u64);
      for (uint32_t k = 0; k < m; ++k) a[k] = (uint32_t)c->sortkey[k];
      for (uint32_t k = 0; k < m; ++k) { c->sortkey[k] = b2[k]; }
      qsort(c->sortkey, m, sizeof(uint64_t), cmp_u64);
      for (uint32_t k = 0; k < m; ++k)
        if (a[k] != (uint32_t)c->sortkey[k]) { *median_quirk = 1; break; }
      free(a);
    }
    free(snap);
  }
  return have;
}

static inline int u32_lt_threshold(const ctx_t* c, uint32_t id, int axis, float t) {
  return bv_center(&c->aabb[id], axis) < t;
}

/* split, lib.rs:536-665.  `self` lives at nodes[self_idx]. */
static void split_node(ctx_t* c, uint32_t self_idx, uint32_t depth) {
  or_node self = c->nodes[self_idx];
  if (self.len <= 1) return;
  if (depth == c->max_depth || (self.len <= c->max_leaf && self.len < c->min_leaf * 2)) {
    if (depth == c->max_depth && self.len > 1) c->st->n_cutoff_leaves++;
    return;
  }
  or_node child_gt = {BV_EMPTY, self.start, 0, 0};
  or_node child_lt = {BV_EMPTY, self.start, 0, 0};
  uint32_t* list = c->list + self.start;
  uint32_t n = self.len;

  split_t sp;
  int quirk = 0;
  if (!select_split(c, &self.b, list, n, &sp, &quirk)) { c->st->n_refused++; return; }
  if (self.len <= c->max_leaf && sp.cost >= OBJECT_COST * (float)self.len) {
    c->st->n_cost_leaf++;
    return;
  }
  if (quirk) c->st->n_median_quirk++;

  child_lt.b = sp.blt;
  child_gt.b = sp.bgt;
  child_lt.len = sp.n_lt;
  child_gt.len = sp.n_gt;

  uint32_t split_index;
  if (sp.kind == 1) {
    /* partition_in_place(center[axis] < t): canonical stable partition */
    uint32_t w = 0;
    for (uint32_t k = 0; k < n; ++k)
      if (u32_lt_threshold(c, list[k], sp.axis, sp.threshold)) c->tmp[w++] = list[k];
    uint32_t nl = w;
    for (uint32_t k = 0; k < n; ++k)
      if (!u32_lt_threshold(c, list[k], sp.axis, sp.threshold)) c->tmp[w++] = list[k];
    memcpy(list, c->tmp, n * sizeof(uint32_t));
    split_index = nl;
  } else if (sp.kind == 0) {
    int t = stable_select(c, list, n, sp.index, sp.axis);
    if (t == 1) c->st->n_tie_events++;
    else if (t == 2) c->st->n_tie_benign++;
    split_index = sp.index;
  } else {
    split_index = sp.index;
  }

  /* asserts lib.rs:618-646 */
  if (split_index == 0 || split_index == n) c->st->n_assert_fail++;
  if (child_lt.len != split_index || child_gt.len != n - split_index) c->st->n_assert_fail++;
  for (uint32_t k = 0; k < n; ++k) {
    float ctr[3];
    obj_center(c, list[k], ctr);
    const or_bv* cb = k < split_index ? &child_lt.b : &child_gt.b;
    if (!bv_contains(cb, ctr)) { c->st->n_assert_fail++; break; }
  }

  child_gt.start = self.start + child_lt.len;

  if (child_gt.len > 0 && child_lt.len > 0) {
    self.child = c->nnodes;
    uint32_t gi = c->nnodes, li = c->nnodes + 1;
    c->nodes[gi] = child_gt;
    c->nodes[li] = child_lt;
    c->nnodes += 2;
    c->nodes[self_idx] = self;
    if (depth > c->st->height) c->st->height = depth;
    if (c->recs) {
      or_rec* r = &c->recs[self_idx];
      r->axis_method = (uint32_t)sp.axis | ((uint32_t)sp.kind << 8);
      r->value = sp.value;
      r->count_lt = sp.n_lt;
      r->cost_bits = fbits(sp.cost);
    }
    if (sp.kind == 0) c->st->n_binned++;
    else if (sp.kind == 1) c->st->n_adaptive++;
    else c->st->n_median++;
    split_node(c, gi, depth + 1);
    split_node(c, li, depth + 1);
  }
}

int or_build(const float* aabbs, uint32_t n, const or_settings* s, uint32_t* list,
             or_node* nodes, or_rec* recs, or_stats* st) {
  memset(st, 0, sizeof *st);
  if (n == 0) return 0; /* lib.rs:703-708: empty BVH */
  ctx_t c;
  c.aabb = (const or_bv*)aabbs;
  c.list = list;
  c.nodes = nodes;
  c.nnodes = 0;
  c.recs = recs;
  c.st = st;
  c.min_leaf = s->min_leaf;
  c.max_leaf = s->max_leaf;
  c.max_depth = s->max_depth;
  c.tmp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  c.sortkey = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n + 8);
  if (!c.tmp || !c.sortkey) { free(c.tmp); free(c.sortkey); return -1; }
  if (recs) memset(recs, 0xff, sizeof(or_rec) * (2 * (size_t)n - 1));

  or_node root;
  if (s->has_bounds) {
    root.b = s->bounds; /* root_with_bounds, lib.rs:230-243 */
  } else {
    root.b = BV_EMPTY; /* root, lib.rs:220-228 */
    for (uint32_t k = 0; k < n; ++k) bv_grow(&root.b, &c.aabb[list[k]]);
  }
  root.start = 0;
  root.len = n;
  root.child = 0;
  nodes[0] = root;
  c.nnodes = 1;
  split_node(&c, 0, 0);
  st->nnodes = c.nnodes;
  free(c.tmp);
  free(c.sortkey);
  return 0;
}

void or_encode_nodes(const or_node* nodes, uint32_t nn, uint32_t* out8) {
  for (uint32_t i = 0; i < nn; ++i) {
    uint32_t* o = out8 + 8 * (size_t)i;
    for (int a = 0; a < 3; ++a) o[a] = fbits(nodes[i].b.mn[a]);
    o[3] = nodes[i].start;
    for (int a = 0; a < 3; ++a) o[4 + a] = fbits(nodes[i].b.mx[a]);
    o[7] = nodes[i].child;
  }
}

double or_tree_sah(const or_node* nodes, uint32_t nn) {
  if (nn == 0) return 0.0;
  double sr = (double)bv_sa(&nodes[0].b);
  if (sr <= 0.0) return 0.0;
  double acc = 0.0;
  for (uint32_t i = 0; i < nn; ++i) {
    double sa = (double)bv_sa(&nodes[i].b) / sr;
    if (nodes[i].child != 0) acc += sa * (OR_PLANT == 4 ? 8.0 : 1.0);
    else acc += sa * 8.0 * (OR_PLANT == 3 ? 1.0 : (double)nodes[i].len);
  }
  return acc;
}

/* ---------------------------------------------------------------- closest hit */

/* Moller-Trumbore as in scene.slang:19-89 (float3 cross/dot expanded left to
 * right, no FMA).  Returns 1 and t on a hit, 0 on a miss. */
static int or_ray_tri_hit(const float* r, const float* v1, const float* v2, const float* v3,
                          float* t_out) {
  float e1[3], e2[3], h[3], s[3], q[3];
  for (int a = 0; a < 3; ++a) { e1[a] = v2[a] - v1[a]; e2[a] = v3[a] - v1[a]; }
  const float* o = r;
  const float* d = r + 3;
  h[0] = d[1] * e2[2] - d[2] * e2[1];
  h[1] = d[2] * e2[0] - d[0] * e2[2];
  h[2] = d[0] * e2[1] - d[1] * e2[0];
  float det = h[0] * e1[0] + h[1] * e1[1];
  det = det + h[2] * e1[2];
  if (fabsf(det) < 1e-6f) return 0;
  float f = 1.0f / det;
  for (int a = 0; a < 3; ++a) s[a] = o[a] - v1[a];
  float du = s[0] * h[0] + s[1] * h[1];
  du = du + s[2] * h[2];
  float u = f * du;
  if (u < 0.0f || u > 1.0f) return 0;
  q[0] = s[1] * e1[2] - s[2] * e1[1];
  q[1] = s[2] * e1[0] - s[0] * e1[2];
  q[2] = s[0] * e1[1] - s[1] * e1[0];
  float dv = d[0] * q[0] + d[1] * q[1];
  dv = dv + d[2] * q[2];
  float v = f * dv;
  if (v < 0.0f || u + v > 1.0f) return 0;
  float dt = e2[0] * q[0] + e2[1] * q[1];
  dt = dt + e2[2] * q[2];
  float t = f * dt;
  if (t < 1e-6f) return 0;
  *t_out = t;
  return 1;
}

/* t of a hit, +inf on a miss. */
float or_ray_tri(const float* r, const float* v1, const float* v2, const float* v3) {
  float t;
  return or_ray_tri_hit(r, v1, v2, v3, &t) ? t : INFINITY;
}

void or_closest_hit_brute(const float* pos, uint32_t vs, const uint32_t* tris, uint32_t ntris,
                          const float* rays, uint32_t nrays, float* t_out, uint32_t* prim_out) {
  for (uint32_t r = 0; r < nrays; ++r) {
    float best = INFINITY;
    uint32_t bp = 0xffffffffu;
    for (uint32_t k = 0; k < ntris; ++k) {
      float t = or_ray_tri(rays + 6 * (size_t)r, pos + (size_t)tris[3 * k] * vs,
                           pos + (size_t)tris[3 * k + 1] * vs, pos + (size_t)tris[3 * k + 2] * vs);
      if (t < best) { best = t; bp = k; } /* ascending k: strict < keeps lowest id */
    }
    t_out[r] = best;
    prim_out[r] = bp;
  }
}

/* Slab test of the traversal (not the reference's box test, which may cull a
 * box holding the closest hit by rounding).  Moller-Trumbore works on
 * s = o - v1, so the hit it reports can lie off the float ray by a rounding
 * error that grows with |o - v1| (and with 1/cos of the angle between the ray
 * and the triangle's plane) and, through the edges v2 - v1 and v3 - v1, with
 * the vertices' own magnitude.  Each box is therefore padded per axis by
 *   max(2^-20 x max(|min|, |max|), 2^-16 x D),
 * D the largest |min - o| or |max - o| over the three axes (the farthest a
 * vertex inside the box can be from the origin, per coordinate), the far
 * distance is widened by (1 + 2^-20), an axis whose plane distance is
 * 0 * inf (the ray runs inside a face plane) does not bound t, and the near
 * distance must not exceed the current closest t widened by (1 + 2^-10). */
#define OR_TF_WIDEN 1.00000095367431640625f /* 1 + 2^-20 */
#define OR_BOX_PAD 9.5367431640625e-07f    /* 2^-20 */
#ifndef OR_DIST_PAD
#define OR_DIST_PAD 1.52587890625e-05f     /* 2^-16 */
#endif
#define OR_TN_SLACK 1.0009765625f           /* 1 + 2^-10 */


static int slab(const float* o, const float* inv, const float* bmn, const float* bmx, float best,
                float* tn_out) {
  float tn = -INFINITY, tf = INFINITY;
  float lo_d[3], hi_d[3], dist = 0.0f;
  for (int a = 0; a < 3; ++a) {
    lo_d[a] = bmn[a] - o[a];
    hi_d[a] = bmx[a] - o[a];
    dist = fmaxf(dist, fmaxf(fabsf(lo_d[a]), fabsf(hi_d[a])));
  }
  const float dp = OR_PLANT == 5 ? 0.0f : dist * OR_DIST_PAD;
  for (int a = 0; a < 3; ++a) {
    float pad = fmaxf(fmaxf(fabsf(bmn[a]), fabsf(bmx[a])) * OR_BOX_PAD, dp);
    float t0 = (lo_d[a] - pad) * inv[a];
    float t1 = (hi_d[a] + pad) * inv[a];
    if (t0 != t0 || t1 != t1) continue; /* 0 * inf: the ray runs inside a face plane */
    float lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
    tn = tn > lo ? tn : lo;
    tf = tf < hi ? tf : hi;
  }
  tf = tf * OR_TF_WIDEN;
  *tn_out = tn;
  return tn <= tf && tf >= 0.0f && tn <= best * OR_TN_SLACK;
}

void or_closest_hit_bvh(const uint32_t* nodes8, const uint32_t* lens, const float* pos,
                        uint32_t vs, const uint32_t* tris, const uint32_t* perm, const float* rays,
                        uint32_t nrays, float* t_out, uint32_t* prim_out) {
  for (uint32_t r = 0; r < nrays; ++r) {
    const float* ray = rays + 6 * (size_t)r;
    float inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / ray[3 + a];
    float best = INFINITY;
    uint32_t bp = 0xffffffffu;
    uint32_t stack[256];
    int sp = 0;
    float tn;
    const float* b0 = (const float*)nodes8;
    if (slab(ray, inv, b0, b0 + 4, best, &tn)) stack[sp++] = 0;
    while (sp > 0) {
      uint32_t v = stack[--sp];
      const uint32_t* nd = nodes8 + 8 * (size_t)v;
      const float* bf = (const float*)nd;
      if (!slab(ray, inv, bf, bf + 4, best, &tn)) continue;
      if (nd[7] == 0) {
        uint32_t len = lens ? lens[v] : 1;
        for (uint32_t k = nd[3]; k < nd[3] + len; ++k) {
          const uint32_t* tr = tris + 3 * (size_t)k;
          float t = or_ray_tri(ray, pos + (size_t)tr[0] * vs, pos + (size_t)tr[1] * vs,
                               pos + (size_t)tr[2] * vs);
          uint32_t id = perm[k];
          if (t < best || (t == best && t != INFINITY && id < bp)) { best = t; bp = id; }
        }
      } else {
        if (sp + 2 > 256) continue;
        stack[sp++] = nd[7] + 1;
        stack[sp++] = nd[7];
      }
    }
    t_out[r] = best;
    prim_out[r] = bp;
  }
}

/* World ray to object space, as the traversal does (row dot products summed
 * left to right, no FMA). */
static void to_object(const float* m, const float* w, float* r) {
  for (int i = 0; i < 3; ++i) {
    const float* row = m + 4 * i;
    float o = row[0] * w[0] + row[1] * w[1];
    o = o + row[2] * w[2];
    r[i] = o + row[3];
    float d = row[0] * w[3] + row[1] * w[4];
    r[3 + i] = d + row[2] * w[5];
  }
}

void or_closest_hit_scene_brute(const float* pos, uint32_t vs, const uint32_t* tris,
                                const uint32_t* mesh3, const uint32_t* inst16, uint32_t ninst,
                                const float* rays, uint32_t nrays, float* t_out,
                                uint32_t* prim_out, uint32_t* inst_out) {
  for (uint32_t r = 0; r < nrays; ++r) {
    float best = INFINITY;
    uint32_t bp = 0xffffffffu, bi = 0xffffffffu;
    for (uint32_t i = 0; i < ninst; ++i) {
      const uint32_t* rec = inst16 + 16 * (size_t)i;
      float m[12];
      memcpy(m, rec, sizeof m);
      const uint32_t* mesh = mesh3 + 3 * (size_t)rec[12];
      float lr[6];
      to_object(m, rays + 6 * (size_t)r, lr);
      for (uint32_t k = 0; k < mesh[1]; ++k) {
        const uint32_t* tr = tris + 3 * ((size_t)mesh[0] + k);
        float t = or_ray_tri(lr, pos + (size_t)(tr[0] + mesh[2]) * vs,
                             pos + (size_t)(tr[1] + mesh[2]) * vs,
                             pos + (size_t)(tr[2] + mesh[2]) * vs);
        if (t < best) { best = t; bp = k; bi = i; } /* ascending (i, k): strict < keeps lowest */
      }
    }
    t_out[r] = best;
    prim_out[r] = bp;
    inst_out[r] = bi;
  }
}

/* ------------------------------------------------- the path tracer's traversal */

/* Hit of scene.slang's traversal: the fields the closest-hit decision reads
 * (Hit.success, Hit.distance, intersect.slang:27-41) plus the triangle that
 * made it.  NO_HIT has distance 0 and no triangle. */
typedef struct {
  float t;
  uint32_t succ, inst, prim, slot;
} sh_hit;

#define SH_STACK 64 /* BVH_STACK_SIZE = TLAS_MAX_DEPTH + BLAS_MAX_DEPTH, scene.slang:5-7 */
#define SH_NONE 0xffffffffu

/* rayIntersectBoundingVolume, scene.slang:296-310.  min/max take the other
 * operand when one is NaN (IEEE minNum, fminf). */
static int sh_box(const float* o, const float* inv, const float* bmn, const float* bmx,
                  float* dist) {
  float t1[3], t2[3];
  for (int a = 0; a < 3; ++a) {
    float lo = (bmn[a] - o[a]) * inv[a];
    float hi = (bmx[a] - o[a]) * inv[a];
    t1[a] = fminf(lo, hi);
    t2[a] = fmaxf(lo, hi);
  }
  float tn = fmaxf(fmaxf(t1[0], t1[1]), t1[2]);
  float tf = fminf(fminf(t2[0], t2[1]), t2[2]);
  int front = tn == 0.0f;
  *dist = front ? tf : tn;
  return front || (tf >= tn && tf > 0.0f);
}

/* Hit.merge, intersect.slang:73-82: b replaces a unless a is strictly closer
 * (a missing hit counts as float.maxValue); success is the OR of both. */
static void sh_merge(sh_hit* a, const sh_hit* b) {
  float da = a->succ ? a->t : FLT_MAX;
  float db = b->succ ? b->t : FLT_MAX;
  uint32_t s = a->succ | b->succ;
  if (!(da < db)) *a = *b;
  a->succ = s;
}

/* Stack.push, stack.slang: the shader writes past a full stack; here the push
 * is dropped and reported. */
static void sh_push(uint32_t* ps, uint32_t* pc, int* sp, uint32_t start, uint32_t child,
                    int* err) {
  if (*sp >= SH_STACK) { *err |= 1; return; }
  ps[*sp] = start;
  pc[*sp] = child;
  ++*sp;
}

/* The child step of both loops, scene.slang:112-134 and 180-202. */
static void sh_children(const uint32_t* nodes8, uint32_t child, const float* o, const float* inv,
                        const sh_hit* closest, uint32_t* ps, uint32_t* pc, int* sp, int* err) {
  const uint32_t* na = nodes8 + 8 * (size_t)child;
  const uint32_t* nb = na + 8;
  float cd = closest->succ ? closest->t : FLT_MAX, da, db;
  int ha = sh_box(o, inv, (const float*)na, (const float*)na + 4, &da);
  int hb = sh_box(o, inv, (const float*)nb, (const float*)nb + 4, &db);
  int pa = ha && da < cd, pb = hb && db < cd;
  if (pa && pb) {
    const uint32_t* first = da < db ? nb : na;
    const uint32_t* second = da < db ? na : nb;
    sh_push(ps, pc, sp, first[3], first[7], err);
    sh_push(ps, pc, sp, second[3], second[7], err);
  } else if (pa || pb) {
    const uint32_t* one = pa ? na : nb;
    sh_push(ps, pc, sp, one[3], one[7], err);
  }
}

int or_raytrace_tlas(const uint32_t* tlas8, const uint32_t* tlas_perm, const uint32_t* inst16,
                     uint32_t ninst, const uint32_t* blas4, uint32_t nblas, const uint32_t* blas8,
                     const float* pos, uint32_t vs, const uint32_t* tris_perm,
                     const uint32_t* blas_perm, const float* rays, uint32_t nrays,
                     uint32_t* hits4) {
  int err = 0;
  for (uint32_t r = 0; r < nrays; ++r) {
    const float* o = rays + 6 * (size_t)r;
    const float* d = o + 3;
    float inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / d[a]; /* Ray.__init, intersect.slang:9 */
    sh_hit best = {0.0f, 0, SH_NONE, SH_NONE, SH_NONE};
    uint32_t ps[SH_STACK], pc[SH_STACK];
    int sp = 0;
    float dist;
    if (sh_box(o, inv, (const float*)tlas8, (const float*)tlas8 + 4, &dist))
      sh_push(ps, pc, &sp, tlas8[3], tlas8[7], &err);
    while (sp > 0) {
      --sp;
      const uint32_t start = ps[sp], child = pc[sp];
      if (child != 0) {
        sh_children(tlas8, child, o, inv, &best, ps, pc, &sp, &err);
        continue;
      }
      /* raytraceBlas, scene.slang:143-220, for meshes[start] */
      const uint32_t id = tlas_perm[start];
      if (id >= ninst) { err |= 2; continue; }
      const uint32_t* rec = inst16 + 16 * (size_t)id;
      if (rec[12] >= nblas) { err |= 2; continue; }
      const uint32_t* bd = blas4 + 4 * (size_t)rec[12];
      float m[12], lo[3], ld[3], linv[3];
      memcpy(m, rec, sizeof m);
      for (int i = 0; i < 3; ++i) { /* mul(inverseTransform, float4(p, w)), :158-161 */
        const float* row = m + 4 * i;
        float x = row[0] * o[0] + row[1] * o[1];
        x = x + row[2] * o[2];
        lo[i] = x + row[3] * 1.0f;
        float y = row[0] * d[0] + row[1] * d[1];
        y = y + row[2] * d[2];
        ld[i] = y + row[3] * 0.0f;
      }
      for (int a = 0; a < 3; ++a) linv[a] = 1.0f / ld[a];
      float dd = ld[0] * ld[0] + ld[1] * ld[1];
      dd = dd + ld[2] * ld[2];
      sh_hit loc = best; /* :165-166: the world distance scaled by length(localRay.dir) */
      loc.t = best.t * sqrtf(dd);
      const uint32_t* bn = blas8 + 8 * (size_t)bd[0];
      const int base = sp;
      sh_push(ps, pc, &sp, bn[3], bn[7], &err); /* the root, without a box test (:172) */
      float lr[6] = {lo[0], lo[1], lo[2], ld[0], ld[1], ld[2]};
      while (sp > base) {
        --sp;
        const uint32_t s2 = ps[sp], c2 = pc[sp];
        if (c2 != 0) {
          sh_children(bn, c2, lo, linv, &loc, ps, pc, &sp, &err);
        } else {
          const uint32_t slot = bd[1] + s2;
          const uint32_t* tr = tris_perm + 3 * (size_t)slot;
          float t;
          int ok = or_ray_tri_hit(lr, pos + (size_t)(tr[0] + bd[2]) * vs,
                                  pos + (size_t)(tr[1] + bd[2]) * vs,
                                  pos + (size_t)(tr[2] + bd[2]) * vs, &t);
          sh_hit h = {0.0f, 0, SH_NONE, SH_NONE, SH_NONE};
          if (ok) { h.t = t; h.succ = 1; h.inst = id; h.prim = blas_perm[slot]; h.slot = slot; }
          sh_merge(&loc, &h);
        }
      }
      sh_merge(&best, &loc); /* :216, no conversion back to world distance (:211-213) */
    }
    uint32_t* out = hits4 + 4 * (size_t)r;
    out[0] = fbits(best.t);
    out[1] = best.succ ? best.prim : SH_NONE;
    out[2] = best.succ ? best.inst : SH_NONE;
    out[3] = !best.succ ? SH_NONE : best.prim == SH_NONE ? 0xfffffffeu : best.slot;
  }
  return err;
}

/* ------------------------------------------- compressed 8-wide BVH (collapse) */

/* Surface area of a node box for the collapse cost: 2 * ((dx*dy + dy*dz) + dz*dx). */
static float w8_area(const float* mn, const float* mx) {
  float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
  float a = dx * dy;
  a = a + dy * dz;
  a = a + dz * dx;
  return 2.0f * a;
}

typedef struct {
  float c[7];       /* C(v, i), i = 1..7 */
  uint8_t dec[8];   /* [0]: 1 = leaf at one slot; [i-1] (i = 2..7): 0 reduce, k split; [7]: k of dist(8) */
  uint32_t p, w, t; /* objects, wide internal nodes and triangles of v's wide subtree */
} w8_rec;

typedef struct {
  const uint32_t* n8;
  const uint32_t* lens;
  w8_rec* r;
  float c_node, c_prim;
  uint32_t p_max;
} w8_ctx;

static inline uint32_t w8_child(const w8_ctx* c, uint32_t v) { return c->n8[8 * (size_t)v + 7]; }

/* The slots of a budget j at binary node m, in order (left child first). */
static int w8_expand(const w8_ctx* c, uint32_t m, uint32_t j, uint32_t* out, int n) {
  while (j > 1 && w8_child(c, m) != 0 && c->r[m].dec[j - 1] == 0) --j;
  if (j == 1 || w8_child(c, m) == 0) { out[n++] = m; return n; }
  uint32_t k = c->r[m].dec[j - 1], ch = w8_child(c, m);
  n = w8_expand(c, ch, k, out, n);
  return w8_expand(c, ch + 1, j - k, out, n);
}

/* Slots of a wide internal node built from binary node v. */
static int w8_slots(const w8_ctx* c, uint32_t v, uint32_t* out) {
  uint32_t ch = w8_child(c, v);
  if (ch == 0) { out[0] = v; return 1; }
  uint32_t k = c->r[v].dec[7];
  int n = w8_expand(c, ch, k, out, 0);
  return w8_expand(c, ch + 1, 8 - k, out, n);
}

static inline int w8_is_leaf_slot(const w8_ctx* c, uint32_t m) {
  return w8_child(c, m) == 0 || c->r[m].dec[0] == 1;
}

static void w8_dp(w8_ctx* c, uint32_t v) {
  const uint32_t* nd = c->n8 + 8 * (size_t)v;
  const float* b = (const float*)nd;
  float area = w8_area(b, b + 4);
  w8_rec* r = &c->r[v];
  memset(r, 0, sizeof *r);
  if (nd[7] == 0) {
    r->p = c->lens ? c->lens[v] : 1;
    float cl = (area * (float)r->p) * c->c_prim;
    for (int i = 0; i < 7; ++i) r->c[i] = cl;
    r->dec[0] = 1;
    r->t = r->p;
    return;
  }
  uint32_t L = nd[7], R = L + 1;
  w8_dp(c, L);
  w8_dp(c, R);
  const w8_rec *a = &c->r[L], *bb = &c->r[R];
  r->p = a->p + bb->p;
  float dist[9];
  uint32_t arg[9];
  for (uint32_t i = 2; i <= 8; ++i) {
    float best = INFINITY;
    uint32_t bk = 1;
    for (uint32_t k = 1; k < i; ++k) {
      if (k > 7 || i - k > 7) continue;
      float s = a->c[k - 1] + bb->c[i - k - 1];
      if (s < best) { best = s; bk = k; }
    }
    dist[i] = best;
    arg[i] = bk;
  }
  float cint = area * c->c_node + dist[8];
  float cleaf = r->p <= c->p_max ? (area * (float)r->p) * c->c_prim : INFINITY;
  /* p <= p_max is part of the rule: when the sums overflow, cint is +inf too */
  const int leaf = r->p <= c->p_max && cleaf <= cint;
  r->dec[0] = leaf ? 1 : 0;
  r->c[0] = leaf ? cleaf : cint;
  for (uint32_t i = 2; i <= 7; ++i) {
    if (dist[i] < r->c[i - 2]) { r->c[i - 1] = dist[i]; r->dec[i - 1] = (uint8_t)arg[i]; }
    else { r->c[i - 1] = r->c[i - 2]; r->dec[i - 1] = 0; }
  }
  r->dec[7] = (uint8_t)arg[8];
  /* wide subtree sizes if v becomes a wide internal node */
  uint32_t s[8];
  int ns = w8_slots(c, v, s);
  r->w = 1;
  r->t = 0;
  for (int k = 0; k < ns; ++k) {
    if (w8_is_leaf_slot(c, s[k])) r->t += c->r[s[k]].p;
    else { r->w += c->r[s[k]].w; r->t += c->r[s[k]].t; }
  }
}

/* Smallest biased exponent E in [1, 254] with 255 * 2^(E-127) >= ext. */
static uint32_t w8_exponent(float lo, float hi) {
  double ext = (double)hi - (double)lo;
  int ex;
  frexp(ext, &ex);
  int e = ex - 8;
  if (e < -126) e = -126;
  while (e < 127 && ldexp(255.0, e) < ext) ++e;
  while (e > -126 && ldexp(255.0, e - 1) >= ext) --e;
  return (uint32_t)(e + 127);
}

/* Conservative 8-bit plane of value x relative to origin p and scale s. */
static uint32_t w8_qlo(float p, float s, float x) {
  double q = floor(((double)x - (double)p) / (double)s);
  uint32_t qi = q < 0 ? 0 : q > 255 ? 255 : (uint32_t)q;
  while (qi > 0 && p + (float)qi * s > x) --qi;
  return qi;
}
static uint32_t w8_qhi(float p, float s, float x) {
  double q = ceil(((double)x - (double)p) / (double)s);
  uint32_t qi = q < 0 ? 0 : q > 255 ? 255 : (uint32_t)q;
  while (qi < 255 && p + (float)qi * s < x) ++qi;
  return qi;
}

static void w8_emit(const w8_ctx* c, uint32_t v, uint32_t idx, uint32_t region, uint32_t tri,
                    uint8_t* out, uint32_t* tri_index) {
  uint8_t* o = out + 80 * (size_t)idx;
  memset(o, 0, 80);
  const float* b = (const float*)(c->n8 + 8 * (size_t)v);
  uint32_t s[8];
  int ns = w8_slots(c, v, s);
  float p[3], sc[3];
  for (int a = 0; a < 3; ++a) {
    uint32_t E = w8_exponent(b[a], b[4 + a]);
    p[a] = b[a];
    sc[a] = bitsf(E << 23);
    o[12 + a] = (uint8_t)E;
  }
  memcpy(o, p, 12);
  uint32_t nint = 0, own = 0;
  for (int k = 0; k < ns; ++k) {
    if (w8_is_leaf_slot(c, s[k])) own += c->r[s[k]].p;
    else ++nint;
  }
  uint32_t base_child = region, base_tri = tri;
  memcpy(o + 16, &base_child, 4);
  memcpy(o + 20, &base_tri, 4);
  uint32_t ci = 0, toff = 0, sub = region + nint, tsub = tri + own;
  uint8_t imask = 0;
  for (int k = 0; k < ns; ++k) {
    uint32_t m = s[k];
    const float* cb = (const float*)(c->n8 + 8 * (size_t)m);
    for (int a = 0; a < 3; ++a) {
      o[32 + 8 * a + k] = (uint8_t)w8_qlo(p[a], sc[a], cb[a]);
      o[56 + 8 * a + k] = (uint8_t)w8_qhi(p[a], sc[a], cb[4 + a]);
    }
    if (w8_is_leaf_slot(c, m)) {
      uint32_t np = c->r[m].p;
      o[24 + k] = (uint8_t)((((1u << np) - 1u) << 5) | toff);
      /* the objects of m's binary subtree in order, child before child + 1 */
      uint32_t st[64];
      int sp = 0;
      st[sp++] = m;
      while (sp > 0) {
        uint32_t u = st[--sp];
        const uint32_t* un = c->n8 + 8 * (size_t)u;
        if (un[7] == 0) {
          uint32_t len = c->lens ? c->lens[u] : 1;
          for (uint32_t j = 0; j < len; ++j) tri_index[tri + toff++] = un[3] + j;
        } else {
          st[sp++] = un[7] + 1;
          st[sp++] = un[7];
        }
      }
    } else {
      imask |= (uint8_t)(1u << k);
      o[24 + k] = (uint8_t)((1u << 5) | (24 + ci));
      w8_emit(c, m, region + ci, sub, tsub, out, tri_index);
      sub += c->r[m].w - 1;
      tsub += c->r[m].t;
      ++ci;
    }
  }
  o[15] = imask;
}

int or_wide8_collapse(const uint32_t* nodes8, const uint32_t* lens, uint32_t nnodes,
                      float c_node, float c_prim, uint32_t p_max, uint8_t* out80,
                      uint32_t* tri_index, uint32_t* out_nwide, float* out_cost) {
  if (nnodes == 0) { *out_nwide = 0; return 0; }
  if (p_max < 1 || p_max > 3) return -3; /* a leaf's count is 3 unary bits of meta */
  if (!(c_node >= 0.0f && c_node < INFINITY && c_prim >= 0.0f && c_prim < INFINITY)) return -4;
  w8_ctx c = {nodes8, lens, (w8_rec*)calloc(nnodes, sizeof(w8_rec)), c_node, c_prim, p_max};
  if (!c.r) return -1;
  for (uint32_t v = 0; v < nnodes; ++v)
    if (nodes8[8 * (size_t)v + 7] == 0 && (lens ? lens[v] : 1) > p_max) { free(c.r); return -2; }
  w8_dp(&c, 0);
  uint32_t nw = nodes8[7] == 0 ? 1 : c.r[0].w;
  w8_emit(&c, 0, 0, 1, 0, out80, tri_index);
  *out_nwide = nw;
  if (out_cost) {
    float a = w8_area((const float*)nodes8, (const float*)nodes8 + 4);
    uint32_t k = c.r[0].dec[7];
    *out_cost = nodes8[7] == 0 ? c.r[0].c[0]
                               : a * c_node + (c.r[nodes8[7]].c[k - 1] + c.r[nodes8[7] + 1].c[8 - k - 1]);
  }
  free(c.r);
  return 0;
}

/* Closest hit through the wide tree: child boxes decoded as p + q * 2^e in
 * fp32, the conservative slab test of the binary traversal, lowest original
 * triangle on equal t. */
void or_trace_wide8(const uint8_t* w80, const uint32_t* tri_index, const float* pos, uint32_t vs,
                    const uint32_t* tris, const uint32_t* perm, const float* rays, uint32_t nrays,
                    float* t_out, uint32_t* prim_out) {
  for (uint32_t r = 0; r < nrays; ++r) {
    const float* ray = rays + 6 * (size_t)r;
    float inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / ray[3 + a];
    float best = INFINITY;
    uint32_t bp = 0xffffffffu;
    uint32_t stack[512];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const uint8_t* o = w80 + 80 * (size_t)stack[--sp];
      float p[3], sc[3];
      memcpy(p, o, 12);
      for (int a = 0; a < 3; ++a) sc[a] = bitsf((uint32_t)o[12 + a] << 23);
      uint32_t bc, bt;
      memcpy(&bc, o + 16, 4);
      memcpy(&bt, o + 20, 4);
      for (int k = 0; k < 8; ++k) {
        uint8_t meta = o[24 + k];
        if (meta == 0) continue;
        float mn[3], mx[3], tn;
        for (int a = 0; a < 3; ++a) {
          mn[a] = p[a] + (float)o[32 + 8 * a + k] * sc[a];
          mx[a] = p[a] + (float)o[56 + 8 * a + k] * sc[a];
        }
        if (!slab(ray, inv, mn, mx, best, &tn)) continue;
        if ((o[15] >> k) & 1) {
          if (sp < 512) stack[sp++] = bc + (meta & 31u) - 24u;
        } else {
          uint32_t cnt = (uint32_t)__builtin_popcount(meta >> 5), off = meta & 31u;
          for (uint32_t j = 0; j < cnt; ++j) {
            uint32_t slot = tri_index[bt + off + j];
            const uint32_t* tr = tris + 3 * (size_t)slot;
            float t = or_ray_tri(ray, pos + (size_t)tr[0] * vs, pos + (size_t)tr[1] * vs,
                                 pos + (size_t)tr[2] * vs);
            uint32_t id = perm[slot];
            if (t < best || (t == best && t != INFINITY && id < bp)) { best = t; bp = id; }
          }
        }
      }
    }
    t_out[r] = best;
    prim_out[r] = bp;
  }
}
