// Compressed 8-wide BVH from the binary tree, and closest-hit traversal of it.
//
// The wide format and its collapse are those of Ylitie, Karras and Laine,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs"
// (HPG 2017), which marigold's README names as the step after the binary
// builder (the binary tree keeps one object per leaf for it).  The collapse
// is the paper's dynamic program over the binary nodes: C(v, i) is the cost of
// v's subtree as at most i subtrees (i = 1..7); a wide leaf of p <= p_max
// objects costs area * p * c_prim, a wide internal node area * c_node plus the
// best split of 8 slots between the children.  Minima keep the first k on
// ties, a leaf wins a leaf/internal tie, and every cost is fp32 in one fixed
// order, so the tree is a pure function of the binary tree.
//
// GPU shape: the node list of every binary depth, built top-down from the
// roots (each level's nodes append their children), then one launch per depth,
// deepest first, each node combining its children's records (no fences or
// atomics between nodes: a per-node device-scope fence, which writes back and
// invalidates the XCD's L2, made a bottom-up atomic walk 10x slower).  A record
// carries, per slot budget, the wide nodes and objects below its slots, so no
// node looks further down than its children.  The layout is then written one
// wide level per launch from the subtree sizes (W wide nodes, T objects): a node's internal
// children sit contiguously at base_child, followed by the children's own
// subtrees in child order, so every index follows from the sizes and the
// result does not depend on thread order.
#include "mgbvh_common.h"
#include "mgbvh_internal.h"

namespace mg {

constexpr uint32_t W8_NONE = 0xffffffffu;
constexpr int W8_GSTACK = 64;  // traversal: one node group per wide level at most

// Per binary node.  The objects below any set of slots that covers v's
// subtree are all of its p objects (each object sits in exactly one wide
// leaf), so only the wide-node counts need a value per budget.
struct W8Rec {
  float c[7];    // C(v, i), i = 1..7
  float c8;      // area * c_node + the best split of 8 slots (v as a wide internal node)
  uint8_t dec[8];  // [0]: 1 = leaf at one slot; [i-1], i = 2..7: 0 reduce, k split; [7]: k of 8
  uint32_t p, w;   // objects; wide nodes of the wide subtree rooted here
  uint32_t ws[7];  // wide nodes below the slots of budget i, i = 1..7
};
static_assert(sizeof(W8Rec) == 76, "W8Rec is 76 bytes");

// What the layout reads of a binary node: decisions, objects, wide nodes.
struct W8E {
  uint8_t dec[8];
  uint32_t p, w;
};
static_assert(sizeof(W8E) == 16, "W8E is 16 bytes");

struct W8Item {  // a wide internal node to write
  uint32_t bin;           // binary node (absolute index)
  uint32_t idx, region;   // wide node and its descendants' first index, relative to wbase
  uint32_t tri;           // first object slot of its leaves, relative to tbase
  uint32_t wbase, tbase;  // the tree's first wide node and object slot
  uint32_t bbase, pad;    // the tree's first binary node (child indices are relative to it)
};

// Absolute index of v's first child (0: a leaf).  In a forest of BLASes the
// child indices are relative to the BLAS's first node, bbase.
__device__ inline uint32_t w8_child(const mgbvh_node32* nodes, uint32_t bbase, uint32_t v) {
  const uint32_t ch = nodes[v].child_node;
  return ch == 0 ? 0u : bbase + ch;
}

__device__ inline float w8_area(const float* mn, const float* mx) {
  const float dx = __fsub_rn(mx[0], mn[0]), dy = __fsub_rn(mx[1], mn[1]);
  const float dz = __fsub_rn(mx[2], mn[2]);
  float a = __fmul_rn(dx, dy);
  a = __fadd_rn(a, __fmul_rn(dy, dz));
  a = __fadd_rn(a, __fmul_rn(dz, dx));
  return __fmul_rn(2.0f, a);
}

__device__ inline void w8_box(const mgbvh_node32* nodes, uint32_t v, float* mn, float* mx,
                              uint32_t& start, uint32_t& child) {
  const uint4* q = reinterpret_cast<const uint4*>(&nodes[v]);
  const uint4 a = q[0], b = q[1];
  mn[0] = __uint_as_float(a.x); mn[1] = __uint_as_float(a.y); mn[2] = __uint_as_float(a.z);
  start = a.w;
  mx[0] = __uint_as_float(b.x); mx[1] = __uint_as_float(b.y); mx[2] = __uint_as_float(b.z);
  child = b.w;
}

__device__ inline float w8_node_area(const mgbvh_node32* nodes, uint32_t v) {
  float mn[3], mx[3];
  uint32_t st, ch;
  w8_box(nodes, v, mn, mx, st, ch);
  return w8_area(mn, mx);
}

// Slots of budget j at binary node m, left child first (an explicit stack).
__device__ inline int w8_expand(const mgbvh_node32* nodes, uint32_t bbase, uint32_t nn,
                                const W8E* E, uint32_t m, uint32_t j, uint32_t* out, int n) {
  uint32_t sm[16], sj[16];
  int sp = 0;
  sm[sp] = m; sj[sp] = j; ++sp;
  while (sp > 0) {
    --sp;
    uint32_t u = sm[sp], b = sj[sp];
    const uint32_t ch = w8_child(nodes, bbase, u);
    if (ch != 0 && ch + 1 >= nn) return 99;
    while (b > 1 && ch != 0 && E[u].dec[b - 1] == 0) --b;
    if (b == 1 || ch == 0) {
      if (n < 8) out[n] = u;
      ++n;
      continue;
    }
    const uint32_t k = E[u].dec[b - 1];
    if (sp + 2 > 16) return 99;  // cannot happen: a budget of 8 splits at most 7 times
    sm[sp] = ch + 1; sj[sp] = b - k; ++sp;  // right pushed first: left expands first
    sm[sp] = ch; sj[sp] = k; ++sp;
  }
  return n;
}

// Slots of a wide internal node built from v.
__device__ inline int w8_slots(const mgbvh_node32* nodes, uint32_t bbase, uint32_t nn,
                               const W8E* E, uint32_t v, uint32_t* out) {
  const uint32_t ch = w8_child(nodes, bbase, v);
  if (ch == 0) { out[0] = v; return 1; }
  if (ch + 1 >= nn) return 99;
  const uint32_t k = E[v].dec[7];
  if (k < 1 || k > 7) return 99;
  const int n = w8_expand(nodes, bbase, nn, E, ch, k, out, 0);
  return n > 8 ? n : w8_expand(nodes, bbase, nn, E, ch + 1, 8 - k, out, n);
}

__device__ inline bool w8_leaf_slot(const mgbvh_node32* nodes, const W8E* E, uint32_t m) {
  return nodes[m].child_node == 0 || E[m].dec[0] == 1;
}

// The record of a binary leaf of p objects.
__device__ inline W8Rec w8_leaf_rec(float area, uint32_t p, float c_prim) {
  W8Rec r;
  r.p = p;
  const float cl = __fmul_rn(__fmul_rn(area, (float)p), c_prim);
#pragma unroll
  for (int i = 0; i < 7; ++i) { r.c[i] = cl; r.ws[i] = 0; }
#pragma unroll
  for (int i = 0; i < 8; ++i) r.dec[i] = 0;
  r.dec[0] = 1;
  r.w = 0;
  r.c8 = cl;
  return r;
}

// The record of an internal node of surface area `area` from its children's
// records A (first child) and B (second).  Costs and counts are read into
// registers at fixed indices; the per-budget wide-node counts are read where
// the decisions pick them, through the pointers (A and B may be global or
// LDS memory; a local copy indexed at run time would live in scratch).
template <class Rec>
__device__ inline W8Rec w8_combine(const Rec* A, const Rec* B, float area, float c_node,
                                   float c_prim, uint32_t p_max) {
  float ac[7], bc[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) { ac[k] = A->c[k]; bc[k] = B->c[k]; }
  W8Rec r;
  r.p = A->p + B->p;
  float dist[9];
  uint32_t arg[9];
#pragma unroll
  for (uint32_t i = 2; i <= 8; ++i) {
    float best = __uint_as_float(0x7f800000u);
    uint32_t bk = 1;
#pragma unroll
    for (uint32_t k = 1; k < i; ++k) {
      if (k > 7 || i - k > 7) continue;
      const float s = __fadd_rn(ac[k - 1], bc[i - k - 1]);
      if (s < best) { best = s; bk = k; }
    }
    dist[i] = best;
    arg[i] = bk;
  }
  const float cint = __fadd_rn(__fmul_rn(area, c_node), dist[8]);
  const float cleaf = r.p <= p_max ? __fmul_rn(__fmul_rn(area, (float)r.p), c_prim)
                                   : __uint_as_float(0x7f800000u);
  // p <= p_max is part of the rule: when the sums overflow, cint is +inf too
  const bool leaf = r.p <= p_max && cleaf <= cint;
  r.c8 = cint;
  r.dec[0] = leaf ? 1 : 0;
  r.c[0] = leaf ? cleaf : cint;
#pragma unroll
  for (uint32_t i = 2; i <= 7; ++i) {
    const bool split = dist[i] < r.c[i - 2];
    r.c[i - 1] = split ? dist[i] : r.c[i - 2];
    r.dec[i - 1] = split ? (uint8_t)arg[i] : 0;
  }
  r.dec[7] = (uint8_t)arg[8];
  // sizes from the children's per-budget sums: no walk below the children
  const uint32_t k8 = arg[8];  // 1..7
  r.w = 1 + A->ws[k8 - 1] + B->ws[8 - k8 - 1];
  r.ws[0] = leaf ? 0u : r.w;
#pragma unroll
  for (uint32_t i = 2; i <= 7; ++i) {
    const uint32_t k = r.dec[i - 1];  // 0 or 1..i-1
    const uint32_t kk = k == 0 ? 1u : k;
    const uint32_t split = A->ws[kk - 1] + B->ws[i - kk - 1];
    r.ws[i - 1] = k == 0 ? r.ws[i - 2] : split;
  }
  return r;
}

__device__ inline W8E w8_e(const W8Rec& r) {
  W8E e;
#pragma unroll
  for (int i = 0; i < 8; ++i) e.dec[i] = r.dec[i];
  e.p = r.p;
  e.w = r.w;
  return e;
}

// Level lists top-down: every node of level l appends its two children to the
// list of level l + 1 (one atomic per block).  A child index must follow
// its parent and fit the array, so a malformed tree cannot loop or overrun.
__global__ __launch_bounds__(1024) void k_w8_next_level(const mgbvh_node32* __restrict__ nodes,
                                                        const uint32_t* __restrict__ base,
                                                        uint32_t nn, const uint32_t* __restrict__ in,
                                                        uint32_t n_in, uint32_t* __restrict__ out,
                                                        uint32_t out_cap, uint32_t* out_count,
                                                        uint32_t* err_flag) {
  __shared__ uint32_t wsum[16], wbase[16], bbase;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = __lane_id();
  const uint32_t i = blockIdx.x * blockDim.x + tid;
  uint32_t v = 0, ch = 0;
  if (i < n_in) {
    v = in[i];
    ch = v < nn ? w8_child(nodes, base ? base[v] : 0u, v) : 0u;
    if (ch != 0 && (ch <= v || ch + 1 >= nn)) { atomicOr(err_flag, 16u); ch = 0; }
  }
  const uint64_t mask = __ballot(ch != 0);
  if (lane == 0) wsum[wave] = 2u * (uint32_t)__popcll(mask);
  __syncthreads();
  if (tid == 0) {  // one atomic per block: same-address atomics serialise in L2
    uint32_t acc = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) { wbase[w] = acc; acc += wsum[w]; }
    bbase = acc ? atomicAdd(out_count, acc) : 0u;
  }
  __syncthreads();
  if (ch != 0) {
    const uint32_t pos = bbase + wbase[wave] + 2u * (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (pos + 1 < out_cap) { out[pos] = ch; out[pos + 1] = ch + 1; }
    else atomicOr(err_flag, 16u);
  }
}

// One binary depth: leaves get their leaf record, internal nodes combine their
// children's (one depth deeper, written by the previous launch).  Every node's
// layout entry (decisions, objects, wide nodes) goes to E as well.
__global__ __launch_bounds__(256) void k_w8_dp_level(const mgbvh_node32* __restrict__ nodes,
                                                     const uint32_t* __restrict__ base,
                                                     const uint32_t* __restrict__ lens,
                                                     uint32_t nn, const uint32_t* __restrict__ order,
                                                     uint32_t count, W8Rec* rec, W8E* E,
                                                     float c_node, float c_prim, uint32_t p_max,
                                                     uint32_t* err_flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t v = order[i];
  if (v >= nn) return;
  const float area = w8_node_area(nodes, v);
  const uint32_t L = w8_child(nodes, base ? base[v] : 0u, v);
  W8Rec r;
  if (L != 0) {
    r = w8_combine(&rec[L], &rec[L + 1], area, c_node, c_prim, p_max);
  } else {
    r = w8_leaf_rec(area, lens ? lens[v] : 1u, c_prim);
    if (r.p > p_max) atomicOr(err_flag, 8u);
  }
  rec[v] = r;
  E[v] = w8_e(r);
}

// ------------------------------------------------------------ subtree-local DP
// The dynamic program without global level lists, for trees built here: a
// node's descendants occupy one contiguous index range starting at its first
// child (children follow their parents, pairs in depth-first order).
//   phase 0 (k_w8_top, one block per tree): the nodes with more than
//     W8_SUB objects form the top of the tree; a breadth-first pass over
//     them lists, level by level, the top nodes and, as subtree roots, their
//     children with at most W8_SUB objects;
//   phase 1 (k_w8_sub, one wavefront per W8_SUBG subtree roots): the
//     subtrees' levels and their DP, bottom up, with every record in LDS;
//     each node's layout entry and the roots' records go to memory;
//   phase 2 (k_w8_topdp, one block per tree): the top nodes' DP, bottom up.
// Object counts come from lens or, without lens (one object per leaf), from
// the first children's start indices (a node's lt child starts where it
// does, its gt child n_lt objects later).  A tree that breaks these
// assumptions (a subtree past its range, a count that does not add up, a
// capacity exceeded) raises a flag and the level path runs instead.
constexpr uint32_t W8_SUB = 64;                // objects of a subtree built in LDS
constexpr uint32_t W8_SUBN = 2 * W8_SUB - 1;   // its nodes at most
constexpr uint32_t W8_MAXLV = 64;              // levels of a top part or a subtree
constexpr uint32_t W8_TAG = 0x80000000u;       // a top entry's child: a subtree slot
constexpr uint32_t W8_FAST_FAIL = 32u;         // err flag: the fast path does not apply

struct W8Top {   // a top node: n > W8_SUB objects
  uint32_t v, n;
  uint32_t ch[2];  // first and second child: top entry index, or W8_TAG | subtree slot
};
struct W8Sub {   // a subtree root: n <= W8_SUB objects, or a leaf
  uint32_t root, bbase, n, pad;
};
struct W8Tree {  // one tree of the forest
  uint32_t first, count;  // its nodes [first, first + count); child indices relative to bbase
  uint32_t bbase, toff, tcap;
  uint32_t ntop, nlev;    // top entries and their levels (phase 0)
  uint32_t root_src;      // the root's record: top entry 0, or W8_TAG | subtree slot
};

// Object count of u (first child c of a node of n objects at k = 0, second at k = 1).
__device__ inline uint32_t w8_count(const mgbvh_node32* nodes, const uint32_t* lens, uint32_t c,
                                    uint32_t n, int k) {
  if (lens) return lens[c + k];
  const uint32_t n_lt = nodes[c].start_index - nodes[c + 1].start_index;
  return k == 1 ? n_lt : n - n_lt;
}

__global__ __launch_bounds__(256) void k_w8_top(const mgbvh_node32* __restrict__ nodes,
                                                const uint32_t* __restrict__ lens, uint32_t nn,
                                                W8Tree* __restrict__ trees, W8Top* __restrict__ top,
                                                uint32_t* __restrict__ lvstart,
                                                W8Sub* __restrict__ sub, uint32_t sub_cap,
                                                uint32_t* ctr) {
  __shared__ uint32_t s_top, s_sub, s_base;
  const uint32_t tid = threadIdx.x;
  W8Tree& T = trees[blockIdx.x];
  const uint32_t first = T.first, count = T.count, bb = T.bbase;
  if (count == 0) return;
  uint32_t* lv = lvstart + (size_t)blockIdx.x * (W8_MAXLV + 1);
  const uint32_t root = first;
  const uint32_t nroot = lens ? lens[root] : (count + 1) / 2;
  if (nroot <= W8_SUB || nodes[root].child_node == 0) {
    if (tid == 0) {
      const uint32_t slot = atomicAdd(&ctr[4], 1u);
      if (slot < sub_cap) sub[slot] = W8Sub{root, bb, nroot, 0u};
      else atomicOr(&ctr[0], W8_FAST_FAIL);
      T.ntop = 0;
      T.nlev = 0;
      T.root_src = W8_TAG | slot;
    }
    return;
  }
  W8Top* E0 = top + T.toff;
  const uint32_t cap = T.tcap;
  if (tid == 0) {
    E0[0] = W8Top{root, nroot, {0u, 0u}};
    lv[0] = 0;
    lv[1] = 1;
    T.root_src = 0;
  }
  __syncthreads();
  uint32_t lo = 0, hi = 1, level = 0;
  bool fail = false;
  while (lo < hi && !fail) {
    if (level + 2 > W8_MAXLV) { fail = true; break; }
    if (tid == 0) { s_top = hi; s_sub = 0; }
    __syncthreads();
    // children: top nodes appended to the next level, subtree roots counted
    for (uint32_t i = lo + tid; i < hi; i += blockDim.x) {
      W8Top e = E0[i];
      const uint32_t ch = nodes[e.v].child_node;
      const uint32_t c = bb + ch;
      bool bad = ch == 0 || c <= e.v || c + 1 >= first + count || c + 1 >= nn;
      uint32_t nc[2] = {0u, 0u};
      if (!bad) {
        nc[0] = w8_count(nodes, lens, c, e.n, 0);
        nc[1] = w8_count(nodes, lens, c, e.n, 1);
        bad = nc[0] == 0 || nc[1] == 0 || nc[0] >= e.n || nc[1] >= e.n || nc[0] + nc[1] != e.n;
      }
      if (bad) { atomicOr(&ctr[0], W8_FAST_FAIL); continue; }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const uint32_t u = c + k;
        if (nc[k] > W8_SUB && nodes[u].child_node != 0) {
          const uint32_t t = atomicAdd(&s_top, 1u);
          if (t < cap) { E0[t] = W8Top{u, nc[k], {0u, 0u}}; e.ch[k] = t; }
          else atomicOr(&ctr[0], W8_FAST_FAIL);
        } else {
          e.ch[k] = W8_TAG | atomicAdd(&s_sub, 1u);  // level-local, rebased below
        }
      }
      E0[i].ch[0] = e.ch[0];
      E0[i].ch[1] = e.ch[1];
    }
    __syncthreads();
    if (tid == 0) s_base = s_sub ? atomicAdd(&ctr[4], s_sub) : 0u;
    __syncthreads();
    // subtree roots of this level into their slots
    for (uint32_t i = lo + tid; i < hi; i += blockDim.x) {
      const W8Top e = E0[i];
      const uint32_t c = bb + nodes[e.v].child_node;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!(e.ch[k] & W8_TAG)) continue;
        const uint32_t slot = s_base + (e.ch[k] & ~W8_TAG);
        if (slot < sub_cap)
          sub[slot] = W8Sub{c + k, bb, w8_count(nodes, lens, c, e.n, k), 0u};
        else atomicOr(&ctr[0], W8_FAST_FAIL);
        E0[i].ch[k] = W8_TAG | slot;
      }
    }
    const uint32_t nhi = min(s_top, cap);
    __syncthreads();
    lo = hi;
    hi = nhi;
    ++level;
    if (tid == 0) lv[level + 1] = hi;
  }
  if (tid == 0) {
    T.ntop = hi;
    T.nlev = level;
    if (fail) atomicOr(&ctr[0], W8_FAST_FAIL);
  }
}

// Phase 1: W8_SUBG subtree roots per wavefront (one wavefront per block,
// persistent over the list).  The wavefront runs the levels of its subtrees
// together, so a level holds the nodes of all of them at that depth.  Local
// index of a node: g * W8_SUBN for the root of the batch's subtree g, plus
// 1 + (u - first child) for a descendant.  The surface area, the object count
// and the local index of the first child of every node of the subtrees'
// ranges are loaded into LDS first, in one coalesced pass, so the level pass
// and the DP wait on no memory round trip per level.  Only internal nodes
// keep a record in LDS, and only what a parent reads of it: a leaf's record
// follows from its area and object count, and the DP forms it on the fly.
constexpr uint32_t W8_SUBG = 4;
constexpr uint32_t W8_SUBL = W8_SUBG * W8_SUBN;       // local nodes of a batch at most
constexpr uint32_t W8_SUBI = W8_SUBG * (W8_SUB - 1);  // internal nodes of a batch at most

// Record of an internal binary node inside a subtree of at most W8_SUB
// objects, as its parent reads it: costs per budget, wide-node counts per
// budget (at most the subtree's 63 internal nodes) and objects (at most 64
// leaves of p_max <= 3 objects; a larger leaf raises the p_max error).
struct W8L {
  float c[7];
  uint8_t ws[7];
  uint8_t p;
};
static_assert(sizeof(W8L) == 36, "W8L is 36 bytes");

__device__ inline W8L w8_small(const W8Rec& r) {
  W8L s;
#pragma unroll
  for (int i = 0; i < 7; ++i) { s.c[i] = r.c[i]; s.ws[i] = (uint8_t)r.ws[i]; }
  s.p = (uint8_t)r.p;
  return s;
}

// w8_combine for children that may be leaves: a leaf child (la / lb) reads as
// the record w8_leaf_rec gives it (every cost its leaf cost cl, no wide
// nodes, its p objects), an internal child as its LDS record (A / B; for a
// leaf child any valid slot, whose values are not used).  Selects, not
// branches, so lanes with different kinds of children stay together; the
// result equals w8_combine on the full records bit for bit.
__device__ inline W8Rec w8_combine_lr(const W8L* A, bool la, float cla, uint32_t pa, const W8L* B,
                                      bool lb, float clb, uint32_t pb, float area, float c_node,
                                      float c_prim, uint32_t p_max) {
  float ac[7], bc[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    ac[k] = la ? cla : A->c[k];
    bc[k] = lb ? clb : B->c[k];
  }
  W8Rec r;
  r.p = (la ? pa : (uint32_t)A->p) + (lb ? pb : (uint32_t)B->p);
  float dist[9];
  uint32_t arg[9];
#pragma unroll
  for (uint32_t i = 2; i <= 8; ++i) {
    float best = __uint_as_float(0x7f800000u);
    uint32_t bk = 1;
#pragma unroll
    for (uint32_t k = 1; k < i; ++k) {
      if (k > 7 || i - k > 7) continue;
      const float s = __fadd_rn(ac[k - 1], bc[i - k - 1]);
      if (s < best) { best = s; bk = k; }
    }
    dist[i] = best;
    arg[i] = bk;
  }
  const float cint = __fadd_rn(__fmul_rn(area, c_node), dist[8]);
  const float cleaf = r.p <= p_max ? __fmul_rn(__fmul_rn(area, (float)r.p), c_prim)
                                   : __uint_as_float(0x7f800000u);
  const bool leaf = r.p <= p_max && cleaf <= cint;
  r.c8 = cint;
  r.dec[0] = leaf ? 1 : 0;
  r.c[0] = leaf ? cleaf : cint;
#pragma unroll
  for (uint32_t i = 2; i <= 7; ++i) {
    const bool split = dist[i] < r.c[i - 2];
    r.c[i - 1] = split ? dist[i] : r.c[i - 2];
    r.dec[i - 1] = split ? (uint8_t)arg[i] : 0;
  }
  r.dec[7] = (uint8_t)arg[8];
  const uint32_t k8 = arg[8];  // 1..7
  const uint32_t wa8 = la ? 0u : (uint32_t)A->ws[k8 - 1];
  const uint32_t wb8 = lb ? 0u : (uint32_t)B->ws[8 - k8 - 1];
  r.w = 1 + wa8 + wb8;
  r.ws[0] = leaf ? 0u : r.w;
#pragma unroll
  for (uint32_t i = 2; i <= 7; ++i) {
    const uint32_t k = r.dec[i - 1];  // 0 or 1..i-1
    const uint32_t kk = k == 0 ? 1u : k;
    const uint32_t wa = la ? 0u : (uint32_t)A->ws[kk - 1];
    const uint32_t wb = lb ? 0u : (uint32_t)B->ws[i - kk - 1];
    r.ws[i - 1] = k == 0 ? r.ws[i - 2] : wa + wb;
  }
  return r;
}

struct W8SubLds {
  W8L R[W8_SUBI];                 // records of the internal nodes, by internal ordinal
  float lar[W8_SUBL];             // surface area of each local node
  uint16_t iord[W8_SUBL];         // internal ordinal of an internal node the level pass reached
  uint16_t order[W8_SUBL + 1];    // local indices level by level
  uint8_t lcl[W8_SUBL];           // local index of its first child in its subtree (0: a leaf)
  uint8_t lp[W8_SUBL];            // objects of a leaf, saturated at 255
  uint32_t lvs[W8_MAXLV + 2];
  uint32_t root[W8_SUBG], c0[W8_SUBG], nloc[W8_SUBG], slot[W8_SUBG];
};
static_assert(W8_SUBN < 256 && W8_SUBL < 65536, "local indices fit");

// LDS writes of the wavefront's lanes visible to all of them.
__device__ inline void w8_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64) void k_w8_sub(const mgbvh_node32* __restrict__ nodes,
                                               const uint32_t* __restrict__ lens, uint32_t nn,
                                               const W8Sub* __restrict__ sub, const uint32_t* ctr_in,
                                               W8Rec* __restrict__ rsub, W8E* __restrict__ E,
                                               float c_node, float c_prim, uint32_t p_max,
                                               uint32_t* ctr) {
  __shared__ W8SubLds L;
  const uint32_t lane = lane_id();
  const uint32_t nsub = ctr_in[4];
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t s0 = blockIdx.x * W8_SUBG; s0 < nsub; s0 += gridDim.x * W8_SUBG) {
    // the batch's subtrees: lane g sets up subtree s0 + g
    if (lane < W8_SUBG) {
      const uint32_t slot = s0 + lane;
      uint32_t r = 0, c0 = 0, nloc = 0;
      if (slot < nsub) {
        const W8Sub S = sub[slot];
        r = S.root;
        const uint32_t rch = r < nn ? nodes[r].child_node : 0u;
        if (r >= nn) {
          atomicOr(&ctr[0], W8_FAST_FAIL);
        } else if (rch == 0) {  // a leaf
          const W8Rec x = w8_leaf_rec(w8_node_area(nodes, r), lens ? lens[r] : 1u, c_prim);
          if (x.p > p_max) atomicOr(&ctr[0], 8u);
          rsub[slot] = x;
          E[r] = w8_e(x);
        } else {
          c0 = S.bbase + rch;
          if (S.n > W8_SUB || S.n < 2 || c0 <= r || c0 >= nn) atomicOr(&ctr[0], W8_FAST_FAIL);
          else nloc = min(2u * S.n - 1u, W8_SUBN);  // descendants: at most 2 n - 2 after the root
        }
      }
      L.root[lane] = r;
      L.c0[lane] = c0;
      L.nloc[lane] = nloc;
      L.slot[lane] = slot;
    }
    w8_wave_sync();
    // the ranges' nodes, flattened over the subtrees (those past the array
    // are never reached: their parents' child test fails first)
    uint32_t tot = 0, po[W8_SUBG];
#pragma unroll
    for (uint32_t g = 0; g < W8_SUBG; ++g) { po[g] = tot; tot += L.nloc[g]; }
    // all of the batch's reads are issued before any is used (the kernel's
    // resident wavefronts are bounded by LDS, not registers)
    constexpr uint32_t W8_LOADS = (W8_SUBL + 63) / 64;
    uint4 xa[W8_LOADS], xb[W8_LOADS];
    uint32_t xl[W8_LOADS];
#pragma unroll
    for (uint32_t j = 0; j < W8_LOADS; ++j) {
      const uint32_t i = lane + 64 * j;
      xa[j] = xb[j] = make_uint4(0u, 0u, 0u, 0u);
      xl[j] = 0;
      if (i < tot) {
        uint32_t g = 0;
#pragma unroll
        for (uint32_t h = 1; h < W8_SUBG; ++h) g = i >= po[h] ? h : g;
        const uint32_t loc = i - po[g];
        const uint32_t u = loc == 0 ? L.root[g] : L.c0[g] + loc - 1;
        if (u < nn) {
          xa[j] = reinterpret_cast<const uint4*>(&nodes[u])[0];
          xb[j] = reinterpret_cast<const uint4*>(&nodes[u])[1];
          xl[j] = lens ? lens[u] : 1u;
        }
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < W8_LOADS; ++j) {
      const uint32_t i = lane + 64 * j;
      if (i >= tot) continue;
      uint32_t g = 0;
#pragma unroll
      for (uint32_t h = 1; h < W8_SUBG; ++h) g = i >= po[h] ? h : g;
      const uint32_t loc = i - po[g], bb = sub[s0 + g].bbase;
      const uint32_t c0 = L.c0[g], nloc = L.nloc[g];
      const uint32_t u = loc == 0 ? L.root[g] : c0 + loc - 1;
      const uint32_t ch = u < nn ? xb[j].w : 0u;
      uint32_t cl = 0;
      float ar = 0.0f;
      if (u < nn) {
        const float mn[3] = {__uint_as_float(xa[j].x), __uint_as_float(xa[j].y),
                             __uint_as_float(xa[j].z)};
        const float mx[3] = {__uint_as_float(xb[j].x), __uint_as_float(xb[j].y),
                             __uint_as_float(xb[j].z)};
        ar = w8_area(mn, mx);
      }
      if (ch != 0) {
        // the first child is usable: after u, inside the subtree's range and
        // the array
        cl = bb + ch - c0 + 1;
        if (!(bb + ch > u && cl >= 1 && cl + 1 < nloc && bb + ch + 1 < nn)) {
          atomicOr(&ctr[0], W8_FAST_FAIL);
          cl = 0;
        }
      }
      const uint32_t loc2 = g * W8_SUBN + loc;
      L.lcl[loc2] = (uint8_t)cl;
      L.lar[loc2] = ar;
      const uint32_t np = ch == 0 && u < nn ? xl[j] : 0u;
      L.lp[loc2] = (uint8_t)min(np, 255u);
    }
    // level 0: the roots of the subtrees that are not leaves (internal
    // ordinals 0, 1, ...)
    uint32_t n = 0, ni = 0;  // nodes listed and internal ordinals given so far (wave-uniform)
    if (lane == 0) {
      for (uint32_t g = 0; g < W8_SUBG; ++g)
        if (L.nloc[g] != 0) {
          L.iord[g * W8_SUBN] = (uint16_t)n;
          L.order[n++] = (uint16_t)(g * W8_SUBN);
        }
      L.lvs[0] = 0;
      L.lvs[1] = n;
    }
    n = __shfl(n, 0, 64);
    ni = n;
    w8_wave_sync();
    // levels, top down: a level's internal nodes append their children in
    // lane order and give the internal ones their ordinals (ballot and
    // popcount, no atomics)
    uint32_t nlev = 1;
    bool fail = false;
    for (;;) {
      const uint32_t a = L.lvs[nlev - 1], b = L.lvs[nlev];
      if (a == b) break;
      for (uint32_t i0 = a; i0 < b; i0 += 64) {
        const uint32_t i = i0 + lane;
        uint32_t cl = 0, loc = 0;
        if (i < b) {
          loc = L.order[i];
          cl = L.lcl[loc];
        }
        const uint32_t c = loc - loc % W8_SUBN + cl;  // combined local index of the first child
        const bool i0c = cl != 0 && L.lcl[c] != 0, i1c = cl != 0 && L.lcl[c + 1] != 0;
        const uint64_t m = __ballot(cl != 0), m0 = __ballot(i0c), m1 = __ballot(i1c);
        const uint32_t k = n + 2u * (uint32_t)__popcll(m & below);
        const uint32_t q = ni + (uint32_t)__popcll(m0 & below) + (uint32_t)__popcll(m1 & below);
        n += 2u * (uint32_t)__popcll(m);
        ni += (uint32_t)__popcll(m0) + (uint32_t)__popcll(m1);
        if (cl != 0) {
          if (k + 1 < W8_SUBL) { L.order[k] = (uint16_t)c; L.order[k + 1] = (uint16_t)(c + 1); }
          else atomicOr(&ctr[0], W8_FAST_FAIL);
          if (i0c) L.iord[c] = (uint16_t)q;
          if (i1c) L.iord[c + 1] = (uint16_t)(q + (i0c ? 1u : 0u));
        }
      }
      n = min(n, W8_SUBL);
      if (lane == 0) L.lvs[nlev + 1] = n;
      w8_wave_sync();
      ++nlev;
      if (nlev + 1 > W8_MAXLV + 1 || ni > W8_SUBI) {
        fail = true;
        if (lane == 0) atomicOr(&ctr[0], W8_FAST_FAIL);
        break;
      }
    }
    // the DP, bottom up
    for (int l = (int)nlev - 1; l >= 0 && !fail; --l) {
      for (uint32_t i = L.lvs[l] + lane; i < L.lvs[l + 1]; i += 64) {
        const uint32_t loc = L.order[i];
        const uint32_t g = loc / W8_SUBN, lo = loc - g * W8_SUBN;
        const uint32_t u = lo == 0 ? L.root[g] : L.c0[g] + lo - 1;
        const uint32_t cl = L.lcl[loc];
        if (cl == 0) {  // a leaf (never a root here): its record follows from area and count
          W8E e;
#pragma unroll
          for (int k = 0; k < 8; ++k) e.dec[k] = k == 0 ? 1 : 0;
          e.p = L.lp[loc];  // exact up to 254, so a leaf over p_max <= 3 is still caught
          e.w = 0;
          if (e.p > p_max) atomicOr(&ctr[0], 8u);
          E[u] = e;
          continue;
        }
        const uint32_t c = g * W8_SUBN + cl;
        const bool la = L.lcl[c] == 0, lb = L.lcl[c + 1] == 0;
        const uint32_t pa = L.lp[c], pb = L.lp[c + 1];
        const float cla = __fmul_rn(__fmul_rn(L.lar[c], (float)pa), c_prim);
        const float clb = __fmul_rn(__fmul_rn(L.lar[c + 1], (float)pb), c_prim);
        const uint32_t ia = la ? 0u : L.iord[c], ib = lb ? 0u : L.iord[c + 1];
        const W8Rec x = w8_combine_lr(&L.R[ia], la, cla, pa, &L.R[ib], lb, clb, pb, L.lar[loc],
                                      c_node, c_prim, p_max);
        L.R[L.iord[loc]] = w8_small(x);
        E[u] = w8_e(x);
        if (lo == 0) rsub[L.slot[g]] = x;
      }
      w8_wave_sync();
    }
    w8_wave_sync();  // the wavefront's LDS is rewritten by its next batch
  }
}

// Phase 2: the top nodes' DP, level by level bottom up, one block per tree.
__global__ __launch_bounds__(256) void k_w8_topdp(const mgbvh_node32* __restrict__ nodes,
                                                  const W8Tree* __restrict__ trees,
                                                  const W8Top* __restrict__ top,
                                                  const uint32_t* __restrict__ lvstart,
                                                  const W8Rec* __restrict__ rsub,
                                                  W8Rec* __restrict__ rtop, W8E* __restrict__ E,
                                                  float c_node, float c_prim, uint32_t p_max) {
  const W8Tree T = trees[blockIdx.x];
  if (T.count == 0 || T.ntop == 0) return;
  const uint32_t* lv = lvstart + (size_t)blockIdx.x * (W8_MAXLV + 1);
  const W8Top* E0 = top + T.toff;
  W8Rec* R0 = rtop + T.toff;
  for (int l = (int)T.nlev - 1; l >= 0; --l) {
    for (uint32_t i = lv[l] + threadIdx.x; i < lv[l + 1]; i += blockDim.x) {
      const W8Top e = E0[i];
      const W8Rec* A = (e.ch[0] & W8_TAG) ? &rsub[e.ch[0] & ~W8_TAG] : &R0[e.ch[0]];
      const W8Rec* B = (e.ch[1] & W8_TAG) ? &rsub[e.ch[1] & ~W8_TAG] : &R0[e.ch[1]];
      const W8Rec x = w8_combine(A, B, w8_node_area(nodes, e.v), c_node, c_prim, p_max);
      R0[i] = x;
      E[e.v] = w8_e(x);
    }
    __syncthreads();  // the next level up reads these records (one block: workgroup scope)
  }
}

// Each tree's root record, wherever its DP left it.
__global__ void k_w8_rootrec(const W8Tree* __restrict__ trees, uint32_t nt,
                             const W8Rec* __restrict__ rsub, const W8Rec* __restrict__ rtop,
                             W8Rec* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt || trees[t].count == 0) return;
  const uint32_t s = trees[t].root_src;
  out[t] = (s & W8_TAG) ? rsub[s & ~W8_TAG] : rtop[trees[t].toff];
}

// The same from the level path's node records.
__global__ void k_w8_rootrec_nodes(const W8Tree* __restrict__ trees, uint32_t nt,
                                   const W8Rec* __restrict__ rec, W8Rec* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt || trees[t].count == 0) return;
  out[t] = rec[trees[t].first];
}

// Largest q with p + q * s <= x (from the double floor), smallest with >= x.
// s is a power of two, 2^(E - 127) with E in 1..254, and inv = 1 / s exactly,
// so the product below is the quotient (x - p) / s bit for bit.
__device__ inline double w8_inv_scale(uint32_t E) {
  return __longlong_as_double((long long)(1150u - E) << 52);  // 2^(127 - E)
}
__device__ inline uint32_t w8_qlo(float p, float s, double inv, float x) {
  const double q = floor(((double)x - (double)p) * inv);
  uint32_t qi = q < 0.0 ? 0u : q > 255.0 ? 255u : (uint32_t)q;
  while (qi > 0 && __fadd_rn(p, __fmul_rn((float)qi, s)) > x) --qi;
  return qi;
}
__device__ inline uint32_t w8_qhi(float p, float s, double inv, float x) {
  const double q = ceil(((double)x - (double)p) * inv);
  uint32_t qi = q < 0.0 ? 0u : q > 255.0 ? 255u : (uint32_t)q;
  while (qi < 255 && __fadd_rn(p, __fmul_rn((float)qi, s)) < x) ++qi;
  return qi;
}
__device__ inline uint32_t w8_exponent(float lo, float hi) {
  const double ext = (double)hi - (double)lo;
  int ex;
  frexp(ext, &ex);
  int e = ex - 8;
  if (e < -126) e = -126;
  while (e < 127 && ldexp(255.0, e) < ext) ++e;
  while (e > -126 && ldexp(255.0, e - 1) >= ext) --e;
  return (uint32_t)(e + 127);
}

// One wide node per group of 8 lanes.  The lanes expand the node's slots
// together: an expansion of budget b is held by the lane at its first
// position and owns positions [pos, pos + b); a split of budget b into k and
// b - k keeps the left part and hands the right part to lane pos + k, so the
// final slots sit in left-first order across the group after as many rounds
// as the expansion is deep.  Each lane then quantises its slot's box and lays
// out its leaf's objects or its internal child's frontier item; prefix sums
// over the group give each slot what the sequential layout gave it (objects
// of the leaf slots before it, internal slots before it, and the wide nodes
// and objects of the internal children before it).  The node's words are
// assembled in LDS and stored as five 16-byte writes.  The frontier items of
// the next level are staged in LDS and reserved in the output list one flush
// at a time (at least W8E_STAGE - 8 * W8E_NODES + 1 items each, and the
// block's remainder at its end): same-address atomics serialise device-wide,
// and the list's order is free (an item carries its own layout offsets).
constexpr uint32_t W8E_THREADS = 256;           // threads of a layout block
constexpr uint32_t W8E_NODES = W8E_THREADS / 8;  // wide nodes per block and step
constexpr uint32_t W8E_STAGE = 512;             // staged frontier items per block
static_assert(W8E_STAGE >= 8 * W8E_NODES, "a step's items fit after a flush");

// Copies the block's staged items to out[base, base + n) (all threads).
__device__ inline void w8_flush(const W8Item* stage, uint32_t n, uint32_t base, W8Item* out,
                                uint32_t front_cap, uint32_t* err_flag) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    if ((uint64_t)base + i < front_cap) out[base + i] = stage[i];
    else atomicOr(err_flag, 4u);
  }
}
__global__ __launch_bounds__(W8E_THREADS) void k_w8_emit(
    const mgbvh_node32* __restrict__ nodes, const uint32_t* __restrict__ lens,
    const W8E* __restrict__ E, const W8Item* __restrict__ in, const uint32_t* n_in,
    W8Item* __restrict__ out, uint32_t* n_out, uint32_t front_cap, uint8_t* __restrict__ wide,
    uint32_t* __restrict__ tri_index, uint32_t nn, uint32_t wide_cap, uint32_t tri_cap,
    uint32_t* err_flag) {
  __shared__ uint32_t wtot[W8E_THREADS / 64], bbase, bslot, bflush, bstaged;
  __shared__ __align__(16) W8Item stage[W8E_STAGE];
  __shared__ __align__(16) uint32_t nw[W8E_NODES][20];  // the node words of each group
  __shared__ uint32_t box_u[W8E_THREADS], box_b[W8E_THREADS];  // hand-over of right parts, per lane
  const uint32_t count = *n_in;
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, t = threadIdx.x;
  const uint32_t k = lane & 7u, g = t >> 3, g0 = lane & ~7u;
  uint32_t* W = nw[g];
  uint8_t* Wb = reinterpret_cast<uint8_t*>(W);
  box_b[t] = 0;
  if (t == 0) bstaged = 0;
  w8_wave_sync();
  // block-uniform trip count: every thread takes part in the staging and the
  // flushes below
  for (uint32_t b0 = blockIdx.x * W8E_NODES; b0 < count; b0 += gridDim.x * W8E_NODES) {
    const uint32_t it = b0 + g;
    bool ok = it < count;  // group-uniform
    W8Item item{};
    uint32_t u = 0, b = 0;  // this lane's expansion (b == 0: none)
    bool fin = false, bad = false;
    // the 32-byte node and the layout entry of u, as last read (a finished
    // slot keeps those of its own node, so nothing is read again after the
    // expansion), and the wide node's own box, read before the expansion
    uint4 na = make_uint4(0u, 0u, 0u, 0u), nb = na, ee = na, pa = na, pb = na;
    if (ok) {
      item = in[it];
      if (item.bin >= nn) {
        bad = true;
      } else {
        const uint32_t v = item.bin;
        pa = reinterpret_cast<const uint4*>(&nodes[v])[0];
        pb = reinterpret_cast<const uint4*>(&nodes[v])[1];
        const uint4 ev = *reinterpret_cast<const uint4*>(&E[v]);  // read beside the node
        const uint32_t ch = pb.w == 0 ? 0u : item.bbase + pb.w;
        if (ch == 0) {
          if (k == 0) {
            u = v; b = 1; fin = true;
            na = pa; nb = pb;
            ee = ev;
          }
        } else if (ch + 1 >= nn) {
          bad = true;
        } else {
          const uint32_t kk = ev.y >> 24;  // dec[7]
          if (kk < 1 || kk > 7) bad = true;
          else if (k == 0) { u = ch; b = kk; }
          else if (k == kk) { u = ch + 1; b = 8 - kk; }
        }
      }
    }
    for (;;) {  // one round per level of the expansion (wave-uniform exit)
      const bool act = b != 0 && !fin && !bad;
      if (__ballot(act) == 0) break;
      bool send = false;
      uint32_t su = 0, sb = 0, kk = 0;
      if (act) {
        na = reinterpret_cast<const uint4*>(&nodes[u])[0];
        nb = reinterpret_cast<const uint4*>(&nodes[u])[1];
        ee = *reinterpret_cast<const uint4*>(&E[u]);
        const uint32_t c = nb.w == 0 ? 0u : item.bbase + nb.w;
        if (c == 0) {
          fin = true;
        } else if (c + 1 >= nn) {
          bad = true;
        } else {
          const uint64_t d = ((uint64_t)ee.y << 32) | ee.x;
          while (b > 1 && ((d >> (8 * (b - 1))) & 0xffu) == 0) --b;
          if (b == 1) {
            fin = true;
          } else {
            kk = (uint32_t)(d >> (8 * (b - 1))) & 0xffu;
            if (kk < 1 || kk >= b) {
              bad = true;
            } else {
              send = true;
              su = c + 1;
              sb = b - kk;
              u = c;
              b = kk;
            }
          }
        }
      }
      // k + b <= 8 holds for every expansion, so lane k + kk is in the group
      // and holds none
      if (send) { box_u[t + kk] = su; box_b[t + kk] = sb; }
      w8_wave_sync();
      const uint32_t rb = box_b[t];
      if (rb != 0) { u = box_u[t]; b = rb; box_b[t] = 0; }
      w8_wave_sync();
    }
    if ((uint32_t)(__ballot(bad) >> g0) & 0xffu) {
      if (k == 0) atomicOr(err_flag, 4u);
      ok = false;
    }
    uint64_t widx = 0;
    if (ok) {
      widx = (uint64_t)item.wbase + item.idx;
      if (widx >= wide_cap) { if (k == 0) atomicOr(err_flag, 4u); ok = false; }
    }
    const bool has = ok && fin;
    const uint32_t fmask = (uint32_t)(__ballot(has) >> g0) & 0xffu;
    const uint32_t sidx = __popc(fmask & ((1u << k) - 1u));  // this lane's slot
    // w8_leaf_slot from the registers: no child, or a leaf at one slot
    const bool leaf = has && (nb.w == 0 || (ee.x & 0xffu) == 1u);
    const bool inner = has && !leaf;
    const uint32_t p = leaf ? ee.z : 0u;        // objects of a leaf slot
    const uint32_t wi = inner ? ee.w - 1 : 0u;  // wide nodes below an internal child
    const uint32_t ti = inner ? ee.z : 0u;      // its objects
    const uint32_t ii = inner ? 1u : 0u;
    // inclusive scans over the group's 8 lanes (lane order is slot order)
    uint32_t sp = p, sw = wi, st_ = ti, si = ii;
#pragma unroll
    for (uint32_t d = 1; d < 8; d <<= 1) {
      const uint32_t a = __shfl_up(sp, d, 8), bb = __shfl_up(sw, d, 8);
      const uint32_t c = __shfl_up(st_, d, 8), e = __shfl_up(si, d, 8);
      if (k >= d) { sp += a; sw += bb; st_ += c; si += e; }
    }
    const uint32_t own = __shfl(sp, (int)(g0 + 7), 64);   // objects of all leaf slots
    const uint32_t nint = __shfl(si, (int)(g0 + 7), 64);  // internal slots
    const uint32_t toff = sp - p, ci = si - ii;
    // the group's first frontier slot: block exclusive scan of nint (group leaders)
    const uint32_t lead = k == 0 ? nint : 0u;
    const uint32_t incl = wave_scan_u32_dpp(lead);
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    if (t == 0) {
      uint32_t tot = 0;
      for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
        const uint32_t x = wtot[w];
        wtot[w] = tot;
        tot += x;
      }
      uint32_t s0 = bstaged, fl = 0;
      if (s0 + tot > W8E_STAGE) {  // flush the staged items first
        fl = s0;
        bbase = atomicAdd(n_out, s0);
        s0 = 0;
      }
      bflush = fl;
      bslot = s0;
      bstaged = s0 + tot;
    }
    __syncthreads();
    if (bflush) {  // block-uniform
      w8_flush(stage, bflush, bbase, out, front_cap, err_flag);
      __syncthreads();  // the stage is refilled below
    }
    const uint32_t gslot = bslot + wtot[wave] + __shfl(incl - lead, (int)g0, 64);
    __syncthreads();  // wtot, bslot and bflush are rewritten by the next step
    if (!ok) continue;
    // the node's frame (every lane of the group: the same values)
    const float pmn[3] = {__uint_as_float(pa.x), __uint_as_float(pa.y), __uint_as_float(pa.z)};
    const float pmx[3] = {__uint_as_float(pb.x), __uint_as_float(pb.y), __uint_as_float(pb.z)};
    // lane k of the group finds the exponent of axis k mod 3, the group shares them
    const uint32_t ka = k % 3u;
    const float lo_k = ka == 0 ? pmn[0] : (ka == 1 ? pmn[1] : pmn[2]);
    const float hi_k = ka == 0 ? pmx[0] : (ka == 1 ? pmx[1] : pmx[2]);
    const uint32_t Ek = w8_exponent(lo_k, hi_k);
    float sc[3];
    double inv[3];
    uint32_t ebytes = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const uint32_t Ex = __shfl(Ek, (int)(g0 + a), 64);
      sc[a] = __uint_as_float(Ex << 23);
      inv[a] = w8_inv_scale(Ex);
      ebytes |= Ex << (8 * a);
    }
    // words 6-19 (meta and planes) start at zero: empty slots stay zero
    W[6 + k] = 0u;
    if (k < 6) W[14 + k] = 0u;
    w8_wave_sync();
    uint32_t ibit = 0;
    if (has) {
      const float cmn[3] = {__uint_as_float(na.x), __uint_as_float(na.y), __uint_as_float(na.z)};
      const float cmx[3] = {__uint_as_float(nb.x), __uint_as_float(nb.y), __uint_as_float(nb.z)};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        Wb[32 + 8 * a + sidx] = (uint8_t)w8_qlo(pmn[a], sc[a], inv[a], cmn[a]);
        Wb[56 + 8 * a + sidx] = (uint8_t)w8_qhi(pmn[a], sc[a], inv[a], cmx[a]);
      }
    }
    if (leaf) {
      Wb[24 + sidx] = (uint8_t)((((1u << p) - 1u) << 5) | toff);
      // the leaf's objects, child before child + 1: p <= 3 objects lie in at
      // most 3 binary leaves, at most two levels below the slot.  A node's
      // start index and child are read together with its object count; a slot
      // that is a binary leaf has its node in registers and its count in p
      // (a leaf slot holds p <= p_max objects, so p is exact)
      uint32_t to = toff;
      const uint32_t c = nb.w == 0 ? 0u : item.bbase + nb.w;
      for (uint32_t i = 0; i < (c == 0 ? 1u : 2u); ++i) {
        const uint32_t x = c == 0 ? u : c + i;
        uint32_t xs = na.w, xc = 0u, xl = p;
        if (c != 0) {
          xs = nodes[x].start_index;
          xc = nodes[x].child_node;
          xl = lens ? lens[x] : 1u;
        }
        const uint32_t cx = xc == 0 ? 0u : item.bbase + xc;
        if (cx != 0 && cx + 1 >= nn) { atomicOr(err_flag, 4u); continue; }
        for (uint32_t j = 0; j < (cx == 0 ? 1u : 2u); ++j) {
          uint32_t ys = xs, len = xl;
          if (cx != 0) {
            const uint32_t y = cx + j;
            ys = nodes[y].start_index;
            len = lens ? lens[y] : 1u;
            if (nodes[y].child_node != 0) { atomicOr(err_flag, 4u); continue; }
          }
          for (uint32_t q = 0; q < len; ++q) {
            const uint64_t at = (uint64_t)item.tbase + item.tri + to++;
            if (at < tri_cap) tri_index[at] = ys + q;
            else atomicOr(err_flag, 4u);
          }
        }
      }
    } else if (inner) {
      Wb[24 + sidx] = (uint8_t)((1u << 5) | (24 + ci));
      ibit = 1u << sidx;
      stage[gslot + ci] = W8Item{u, item.region + ci, item.region + nint + (sw - wi),
                                 item.tri + own + (st_ - ti), item.wbase, item.tbase,
                                 item.bbase, 0u};
    }
#pragma unroll
    for (uint32_t d = 1; d < 8; d <<= 1) ibit |= __shfl_xor(ibit, (int)d, 8);
    if (k == 0) {
      W[0] = __float_as_uint(pmn[0]);
      W[1] = __float_as_uint(pmn[1]);
      W[2] = __float_as_uint(pmn[2]);
      W[3] = ebytes | (ibit << 24);
      W[4] = item.region;
      W[5] = item.tri;
    }
    w8_wave_sync();
    if (k < 5) {
      const uint4 x = reinterpret_cast<const uint4*>(W)[k];
      reinterpret_cast<uint4*>(wide + 80 * (size_t)widx)[k] = x;
    }
    w8_wave_sync();  // W is rewritten by the group's next node
  }
  __syncthreads();  // the last step's staged items
  if (t == 0) {
    const uint32_t s0 = bstaged;
    bflush = s0;
    bbase = s0 ? atomicAdd(n_out, s0) : 0u;
  }
  __syncthreads();
  w8_flush(stage, bflush, bbase, out, front_cap, err_flag);
}

// ------------------------------------------------------------------ traversal
// The binary traversal's slab test and margins (the box padded per axis by
// max(2^-20 x the box's coordinate size, 2^-16 x its farthest coordinate from
// the origin), far distances widened by 1 + 2^-20, near ones compared with
// the closest hit widened by 1 + 2^-10), with the pad given: a wide node's
// pad from its quantisation frame [p, p + 255 s], which holds every child's
// box, is at least each child's own pad (a larger box and a farther corner),
// so the test only keeps more.
__device__ inline bool w8_slab_pad(const float* o, const float* inv, const float* bmn,
                                   const float* bmx, const float* pad, float best) {
  float tn = -__uint_as_float(0x7f800000u), tf = __uint_as_float(0x7f800000u);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float t0 = __fmul_rn(__fsub_rn(__fsub_rn(bmn[a], o[a]), pad[a]), inv[a]);
    const float t1 = __fmul_rn(__fadd_rn(__fsub_rn(bmx[a], o[a]), pad[a]), inv[a]);
    const bool free_axis = (t0 != t0) | (t1 != t1);
    tn = fmaxf(tn, free_axis ? -__uint_as_float(0x7f800000u) : fminf(t0, t1));
    tf = fminf(tf, free_axis ? __uint_as_float(0x7f800000u) : fmaxf(t0, t1));
  }
  tf = __fmul_rn(tf, 1.00000095367431640625f);
  return tn <= tf && tf >= 0.0f && tn <= __fmul_rn(best, 1.0009765625f);
}

__device__ inline float w8_dot(const float* a, const float* b) {
  return __fadd_rn(__fadd_rn(__fmul_rn(a[0], b[0]), __fmul_rn(a[1], b[1])), __fmul_rn(a[2], b[2]));
}

__device__ inline float w8_tri(const float* o, const float* d, const float* v1, const float* v2,
                               const float* v3) {
  float e1[3], e2[3], h[3], s[3], q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) { e1[a] = __fsub_rn(v2[a], v1[a]); e2[a] = __fsub_rn(v3[a], v1[a]); }
  h[0] = __fsub_rn(__fmul_rn(d[1], e2[2]), __fmul_rn(d[2], e2[1]));
  h[1] = __fsub_rn(__fmul_rn(d[2], e2[0]), __fmul_rn(d[0], e2[2]));
  h[2] = __fsub_rn(__fmul_rn(d[0], e2[1]), __fmul_rn(d[1], e2[0]));
  const float det = w8_dot(h, e1);
  const float inf = __uint_as_float(0x7f800000u);
  if (fabsf(det) < 1e-6f) return inf;
  const float f = __fdiv_rn(1.0f, det);
#pragma unroll
  for (int a = 0; a < 3; ++a) s[a] = __fsub_rn(o[a], v1[a]);
  const float u = __fmul_rn(f, w8_dot(s, h));
  if (u < 0.0f || u > 1.0f) return inf;
  q[0] = __fsub_rn(__fmul_rn(s[1], e1[2]), __fmul_rn(s[2], e1[1]));
  q[1] = __fsub_rn(__fmul_rn(s[2], e1[0]), __fmul_rn(s[0], e1[2]));
  q[2] = __fsub_rn(__fmul_rn(s[0], e1[1]), __fmul_rn(s[1], e1[0]));
  const float vv = __fmul_rn(f, w8_dot(d, q));
  if (vv < 0.0f || __fadd_rn(u, vv) > 1.0f) return inf;
  const float t = __fmul_rn(f, w8_dot(e2, q));
  return t < 1e-6f ? inf : t;
}

__global__ __launch_bounds__(256) void k_trace_w8(
    const uint8_t* __restrict__ wide, const uint32_t* __restrict__ tri_index,
    const float* __restrict__ pos, uint32_t vs, const uint32_t* __restrict__ tris,
    const uint32_t* __restrict__ perm, const mgbvh_ray* __restrict__ rays, uint32_t nrays,
    uint32_t nwide, mgbvh_hit* __restrict__ hits, uint32_t* err_flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrays) return;
  const float4* rp = reinterpret_cast<const float4*>(&rays[i]);
  const float4 ra = rp[0], rb = rp[1];
  const float o[3] = {ra.x, ra.y, ra.z}, d[3] = {rb.x, rb.y, rb.z};
  float inv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) inv[a] = __fdiv_rn(1.0f, d[a]);
  float best = rb.w;
  uint32_t bp = W8_NONE, bslot = W8_NONE, err = 0;
  // node groups (Ylitie et al.): the current group is a first internal child
  // and a mask of the hit internal children after it (bit ci: child base +
  // ci), kept in registers; a group goes on the stack only when the
  // traversal descends while siblings remain, so the stack is no deeper than
  // the wide tree.  Children are tested against the closest hit when their
  // parent is opened; the hits do not depend on the visiting order.
  uint2 stk[W8_GSTACK];
  int sp = 0;
  uint32_t gbase = 0, gmask = 1;  // the root as a group of one
  for (;;) {
    if (gmask == 0) {
      if (sp == 0) break;
      --sp;
      gbase = stk[sp].x;
      gmask = stk[sp].y;
    }
    const uint32_t w = gbase + (uint32_t)(__ffs(gmask) - 1);
    gmask &= gmask - 1;
    if (w >= nwide) { err |= 2u; continue; }
    const uint4* np = reinterpret_cast<const uint4*>(wide + 80 * (size_t)w);
    uint32_t words[20];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const uint4 x = np[k];
      words[4 * k] = x.x; words[4 * k + 1] = x.y; words[4 * k + 2] = x.z; words[4 * k + 3] = x.w;
    }
    const float p[3] = {__uint_as_float(words[0]), __uint_as_float(words[1]),
                        __uint_as_float(words[2])};
    const float sc[3] = {__uint_as_float((words[3] & 0xffu) << 23),
                         __uint_as_float(((words[3] >> 8) & 0xffu) << 23),
                         __uint_as_float(((words[3] >> 16) & 0xffu) << 23)};
    const uint32_t imask = words[3] >> 24, bc = words[4], bt = words[5];
    // the node's pad, from its frame [p, p + 255 s]
    float pad[3];
    {
      float dist = 0.0f, fr[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        fr[a] = __fadd_rn(p[a], __fmul_rn(255.0f, sc[a]));
        dist = fmaxf(dist, fmaxf(fabsf(__fsub_rn(p[a], o[a])), fabsf(__fsub_rn(fr[a], o[a]))));
      }
      const float dp = __fmul_rn(dist, 1.52587890625e-05f);
#pragma unroll
      for (int a = 0; a < 3; ++a)
        pad[a] = fmaxf(__fmul_rn(fmaxf(fabsf(p[a]), fabsf(fr[a])), 9.5367431640625e-07f), dp);
    }
    uint32_t cmask = 0;  // hit internal children, by internal ordinal
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t meta = (words[6 + (k >> 2)] >> (8 * (k & 3))) & 0xffu;
      if (meta == 0) continue;
      float mn[3], mx[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const uint32_t ql = (words[8 + 2 * a + (k >> 2)] >> (8 * (k & 3))) & 0xffu;
        const uint32_t qh = (words[14 + 2 * a + (k >> 2)] >> (8 * (k & 3))) & 0xffu;
        mn[a] = __fadd_rn(p[a], __fmul_rn((float)ql, sc[a]));
        mx[a] = __fadd_rn(p[a], __fmul_rn((float)qh, sc[a]));
      }
      if (!w8_slab_pad(o, inv, mn, mx, pad, best)) continue;
      if ((imask >> k) & 1u) {
        cmask |= 1u << (((meta & 31u) - 24u) & 7u);
      } else {
        // a leaf's (at most 3) objects: every read of a step issued together
        const uint32_t cnt = __popc(meta >> 5), off = meta & 31u;
        uint32_t sl[3], tv[3][3];
#pragma unroll
        for (uint32_t j = 0; j < 3; ++j) sl[j] = j < cnt ? tri_index[bt + off + j] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < 3; ++j)
#pragma unroll
          for (uint32_t q = 0; q < 3; ++q) tv[j][q] = j < cnt ? tris[3 * (size_t)sl[j] + q] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < 3; ++j) {
          if (j >= cnt) break;
          const float t = w8_tri(o, d, pos + (size_t)tv[j][0] * vs, pos + (size_t)tv[j][1] * vs,
                                 pos + (size_t)tv[j][2] * vs);
          const uint32_t id = perm[sl[j]];
          if (t < best || (t == best && t != __uint_as_float(0x7f800000u) && id < bp)) {
            best = t; bp = id; bslot = sl[j];
          }
        }
      }
    }
    if (cmask != 0) {  // descend; the siblings left wait on the stack
      if (gmask != 0) {
        if (sp < W8_GSTACK) stk[sp++] = make_uint2(gbase, gmask);
        else err |= 1u;
      }
      gbase = bc;
      gmask = cmask;
    }
  }
  reinterpret_cast<uint4*>(hits)[i] = make_uint4(__float_as_uint(best), bp,
                                                 bp == W8_NONE ? W8_NONE : 0u, bslot);
  if (err) atomicOr(err_flag, err);
}

// Every tree's first node: base[v] for the nodes of tree blockIdx.x (level path).
__global__ void k_w8_base(const W8Tree* __restrict__ trees, uint32_t nt, uint32_t* __restrict__ base) {
  if (blockIdx.x >= nt) return;
  const uint32_t f = trees[blockIdx.x].first, n = trees[blockIdx.x].count;
  const uint32_t b = trees[blockIdx.x].bbase;
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) base[f + k] = b;
}

// Per tree: wide nodes and objects of its collapsed tree.
__global__ void k_w8_roots(const mgbvh_node32* __restrict__ nodes, const W8Tree* __restrict__ trees,
                           const W8Rec* __restrict__ rootrec, uint32_t nt, uint32_t* __restrict__ out) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nt) return;
  uint32_t w = 0, t = 0;
  if (trees[m].count) {
    t = rootrec[m].p;
    w = nodes[trees[m].first].child_node == 0 ? 1u : rootrec[m].w;
  }
  out[2 * m] = w;
  out[2 * m + 1] = t;
}

struct W8Forest {
  W8E* E = nullptr;          // layout entries, per binary node
  W8Rec* rootrec = nullptr;  // per tree: its root's record
  uint32_t* ctr = nullptr;   // [0] err, [1..2] frontier counts, [3] level count, [4] subtrees
};

static inline size_t w8_align(size_t b) { return (b + 255) & ~(size_t)255; }

// The layout entries and root records of every tree: the subtree-local path
// (w8_dp_fast) for trees built here, else the level path (w8_dp_levels).
// d_trees: the device copy of `trees`.
static int w8_dp_levels(mgbvh_ctx* c, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                        uint32_t nn, const std::vector<W8Tree>& trees, W8Tree* d_trees,
                        const mgbvh_wide8_params& P, W8Forest& f) {
  Scratch& sc = c->sc;
  int rc;
  const uint32_t nt = (uint32_t)trees.size();
  if ((rc = ensure(sc.w8rec, sizeof(W8Rec) * (size_t)nn))) return rc;
  if ((rc = ensure(sc.w8par, 4 * (size_t)nn + 8))) return rc;  // node ids, level by level
  if ((rc = ensure(sc.w8base, 4 * (size_t)nn))) return rc;     // each node's tree base
  W8Rec* rec = reinterpret_cast<W8Rec*>(sc.w8rec.p);
  uint32_t* order = reinterpret_cast<uint32_t*>(sc.w8par.p);
  uint32_t* base = reinterpret_cast<uint32_t*>(sc.w8base.p);
  hipStream_t st = c->stream;
  hipError_t e;
  std::vector<uint32_t> roots;
  for (const W8Tree& t : trees)
    if (t.count) roots.push_back(t.first);
  if (roots.size() > nn) { set_error("more roots than nodes"); return MGBVH_ERR_INVALID_ARG; }
  if ((e = hipMemsetAsync(f.ctr, 0, 32, st)) != hipSuccess ||
      (e = hipMemsetAsync(base, 0, 4 * (size_t)nn, st)) != hipSuccess ||
      (e = hipMemcpyAsync(order, roots.data(), 4 * roots.size(), hipMemcpyHostToDevice, st)) !=
          hipSuccess)
    return hip_fail(e, "wide8 roots");
  // a node outside every tree keeps base 0 and an absolute child index
  k_w8_base<<<nt, 256, 0, st>>>(d_trees, nt, base);
  std::vector<uint32_t> start = {0}, count = {(uint32_t)roots.size()};
  for (;;) {
    const uint32_t s0 = start.back(), n0 = count.back();
    const uint32_t s1 = s0 + n0;
    if (s1 > nn) { set_error("more nodes reached than the array holds"); return MGBVH_ERR_INDEX; }
    if (start.size() > 127) { set_error("binary tree deeper than 126"); return MGBVH_ERR_INVALID_ARG; }
    if ((e = hipMemsetAsync(f.ctr + 3, 0, 4, st)) != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    k_w8_next_level<<<(n0 + 1023) / 1024, 1024, 0, st>>>(d_nodes, base, nn, order + s0, n0,
                                                       order + s1, nn - s1, f.ctr + 3, f.ctr);
    uint32_t hd[4];
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(hd, f.ctr, 16, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
      return hip_fail(e, "wide8 levels");
    if (hd[0] & 16u) {
      set_error("a child index is out of range, not after its parent, or shared");
      return MGBVH_ERR_INDEX;
    }
    if (hd[3] == 0) break;
    start.push_back(s1);
    count.push_back(hd[3]);
  }
  for (int level = (int)start.size() - 1; level >= 0; --level)
    k_w8_dp_level<<<(count[level] + 255) / 256, 256, 0, st>>>(
        d_nodes, base, d_lens, nn, order + start[level], count[level], rec, f.E, P.c_node,
        P.c_prim, P.p_max, f.ctr);
  k_w8_rootrec_nodes<<<(nt + 255) / 256, 256, 0, st>>>(d_trees, nt, rec, f.rootrec);
  uint32_t err = 0;
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(&err, f.ctr, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return hip_fail(e, "k_w8_dp_level");
  if (err & 8u) {
    set_error("a binary leaf holds more than p_max objects");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (err) { set_error("wide8 collapse: slot expansion out of range"); return MGBVH_ERR_INTERNAL; }
  return MGBVH_OK;
}

// Returns MGBVH_OK, an error, or 1: the trees do not fit the subtree-local
// path (the level path then computes the same records).
static int w8_dp_fast(mgbvh_ctx* c, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                      uint32_t nn, std::vector<W8Tree>& trees, W8Tree* d_trees,
                      const mgbvh_wide8_params& P, W8Forest& f) {
  Scratch& sc = c->sc;
  const uint32_t nt = (uint32_t)trees.size();
  // top nodes of a tree of n <= (count + 1) / 2 objects: each has more than
  // W8_SUB objects, so at most 2 n / W8_SUB of them; subtree roots at most
  // 2 per top node + 1
  uint64_t ttot = 0;
  for (W8Tree& t : trees) {
    t.toff = (uint32_t)ttot;
    t.tcap = t.count ? t.count / W8_SUB + 2u : 0u;
    ttot += t.tcap;
  }
  const uint64_t sub_cap = 2 * ttot + nt;
  if (ttot + sub_cap > 0x7fffffffull) return 1;
  const size_t o_top = 0, o_rtop = o_top + w8_align(sizeof(W8Top) * ttot);
  const size_t o_sub = o_rtop + w8_align(sizeof(W8Rec) * ttot);
  const size_t o_rsub = o_sub + w8_align(sizeof(W8Sub) * sub_cap);
  const size_t o_lv = o_rsub + w8_align(sizeof(W8Rec) * sub_cap);
  const size_t bytes = o_lv + w8_align(4 * (size_t)nt * (W8_MAXLV + 1));
  int rc = ensure(sc.w8fast, bytes);
  if (rc) return rc;
  char* fb = reinterpret_cast<char*>(sc.w8fast.p);
  W8Top* top = reinterpret_cast<W8Top*>(fb + o_top);
  W8Rec* rtop = reinterpret_cast<W8Rec*>(fb + o_rtop);
  W8Sub* sub = reinterpret_cast<W8Sub*>(fb + o_sub);
  W8Rec* rsub = reinterpret_cast<W8Rec*>(fb + o_rsub);
  uint32_t* lv = reinterpret_cast<uint32_t*>(fb + o_lv);
  hipStream_t st = c->stream;
  hipError_t e;
  if ((e = hipMemcpyAsync(d_trees, trees.data(), sizeof(W8Tree) * nt, hipMemcpyHostToDevice,
                          st)) != hipSuccess ||
      (e = hipMemsetAsync(f.ctr, 0, 32, st)) != hipSuccess)
    return hip_fail(e, "wide8 trees");
  k_w8_top<<<nt, 256, 0, st>>>(d_nodes, d_lens, nn, d_trees, top, lv, sub, (uint32_t)sub_cap,
                               f.ctr);
  uint32_t hc[8];
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(hc, f.ctr, 32, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return hip_fail(e, "k_w8_top");
  if ((hc[0] & W8_FAST_FAIL) || hc[4] > sub_cap) return 1;
  const uint32_t nsub = hc[4];
  const uint32_t g1 = std::max<uint32_t>(
      1, std::min<uint32_t>((nsub + W8_SUBG - 1) / W8_SUBG, (uint32_t)c->cus * 16));
  k_w8_sub<<<g1, 64, 0, st>>>(d_nodes, d_lens, nn, sub, f.ctr, rsub, f.E, P.c_node, P.c_prim,
                              P.p_max, f.ctr);
  k_w8_topdp<<<nt, 256, 0, st>>>(d_nodes, d_trees, top, lv, rsub, rtop, f.E, P.c_node,
                                 P.c_prim, P.p_max);
  k_w8_rootrec<<<(nt + 255) / 256, 256, 0, st>>>(d_trees, nt, rsub, rtop, f.rootrec);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(hc, f.ctr, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return hip_fail(e, "k_w8_sub");
  if (hc[0] & W8_FAST_FAIL) return 1;
  if (hc[0] & 8u) {
    set_error("a binary leaf holds more than p_max objects");
    return MGBVH_ERR_INVALID_ARG;
  }
  return MGBVH_OK;
}

static int w8_dp(mgbvh_ctx* c, const mgbvh_node32* d_nodes, const uint32_t* d_lens, uint32_t nn,
                 std::vector<W8Tree>& trees, const mgbvh_wide8_params& P, W8Forest& f) {
  Scratch& sc = c->sc;
  const uint32_t nt = (uint32_t)trees.size();
  int rc;
  if ((rc = ensure(sc.w8e, sizeof(W8E) * (size_t)nn))) return rc;
  if ((rc = ensure(sc.w8ctr, 32 + w8_align(sizeof(W8Tree) * nt) + sizeof(W8Rec) * (size_t)nt)))
    return rc;
  f.E = reinterpret_cast<W8E*>(sc.w8e.p);
  f.ctr = reinterpret_cast<uint32_t*>(sc.w8ctr.p);
  W8Tree* d_trees = reinterpret_cast<W8Tree*>(reinterpret_cast<char*>(sc.w8ctr.p) + 32);
  f.rootrec = reinterpret_cast<W8Rec*>(reinterpret_cast<char*>(d_trees) +
                                       w8_align(sizeof(W8Tree) * nt));
  rc = w8_dp_fast(c, d_nodes, d_lens, nn, trees, d_trees, P, f);
  if (rc != 1) return rc;
  return w8_dp_levels(c, d_nodes, d_lens, nn, trees, d_trees, P, f);
}

// Writes the wide trees whose roots are `roots`, one wide level per launch.
static int w8_emit_forest(mgbvh_ctx* c, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                          uint32_t nn, const W8Forest& f, const std::vector<W8Item>& roots,
                          mgbvh_wide_node* d_wide, uint64_t wide_cap, uint32_t* d_tri_index,
                          uint64_t tri_cap) {
  if (roots.empty()) return MGBVH_OK;
  Scratch& sc = c->sc;
  const uint32_t front_cap = nn / 2 + (uint32_t)roots.size() + 1;
  const size_t front = sizeof(W8Item) * (size_t)front_cap;
  int rc;
  if ((rc = ensure(sc.w8front[0], front)) || (rc = ensure(sc.w8front[1], front))) return rc;
  W8Item* fr[2] = {reinterpret_cast<W8Item*>(sc.w8front[0].p),
                   reinterpret_cast<W8Item*>(sc.w8front[1].p)};
  hipStream_t st = c->stream;
  // Levels go in batches without reading counts back in between: a level's
  // grid covers up to 8 internal children per node of the last count read,
  // and a launch with nothing to do returns at once.
  uint32_t n_in = (uint32_t)roots.size();
  hipError_t e;
  if ((e = hipMemsetAsync(f.ctr, 0, 4, st)) != hipSuccess ||
      (e = hipMemcpyAsync(fr[0], roots.data(), sizeof(W8Item) * roots.size(),
                          hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(f.ctr + 1, &n_in, 4, hipMemcpyHostToDevice, st)) != hipSuccess)
    return hip_fail(e, "wide8 frontier");
  const uint32_t egrid = std::max<uint32_t>(
      1, std::min<uint32_t>((nn / 2 + W8E_NODES - 1) / W8E_NODES, 8192 * 32 / W8E_NODES));
  constexpr int W8_EMIT_BATCH = 4;
  int level = 0;
  while (n_in > 0 && level < 256) {
    uint64_t bound = n_in;  // items of this level at most
    for (int b = 0; b < W8_EMIT_BATCH && level < 256; ++b, ++level) {
      uint32_t* cin = f.ctr + 1 + (level & 1);
      uint32_t* cout = f.ctr + 1 + ((level + 1) & 1);
      const uint32_t grid = (uint32_t)std::min<uint64_t>(egrid, (bound + W8E_NODES - 1) / W8E_NODES);
      if ((e = hipMemsetAsync(cout, 0, 4, st)) != hipSuccess) return hip_fail(e, "hipMemsetAsync");
      k_w8_emit<<<std::max<uint32_t>(grid, 1), W8E_THREADS, 0, st>>>(
          d_nodes, d_lens, f.E, fr[level & 1], cin, fr[(level + 1) & 1], cout, front_cap,
          reinterpret_cast<uint8_t*>(d_wide), d_tri_index, nn,
          (uint32_t)std::min<uint64_t>(wide_cap, 0xffffffffu),
          (uint32_t)std::min<uint64_t>(tri_cap, 0xffffffffu), f.ctr);
      if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_w8_emit");
      bound = std::min<uint64_t>(bound * 8, front_cap);
    }
    if ((e = hipMemcpyAsync(&n_in, f.ctr + 1 + (level & 1), 4, hipMemcpyDeviceToHost, st)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
      return hip_fail(e, "wide8 level");
  }
  uint32_t err = 0;
  if ((e = hipMemcpy(&err, f.ctr, 4, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "err");
  if (err || n_in != 0) { set_error("wide8 collapse: layout out of range"); return MGBVH_ERR_INTERNAL; }
  return MGBVH_OK;
}

static int w8_params(const mgbvh_wide8_params* params, mgbvh_wide8_params& P) {
  P = params ? *params : mgbvh_wide8_params{1.0f, 0.3f, 3u, 0u};
  if (P.p_max < 1 || P.p_max > 3 || !(P.c_node >= 0.0f && P.c_node < INFINITY) ||
      !(P.c_prim >= 0.0f && P.c_prim < INFINITY)) {
    set_error("wide8 parameters: p_max in 1..3 (a leaf's count is 3 unary bits), costs finite "
              "and >= 0");
    return MGBVH_ERR_INVALID_ARG;
  }
  return MGBVH_OK;
}

}  // namespace mg

using namespace mg;

extern "C" {

int mgbvh_wide8_collapse_device(mgbvh_ctx* c, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                                uint64_t nnodes, const mgbvh_wide8_params* params,
                                mgbvh_wide_node* d_wide, uint64_t wide_cap,
                                uint32_t* d_tri_index, uint64_t tri_cap, uint64_t* out_nwide,
                                float* out_cost) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (out_nwide) *out_nwide = 0;
  if (nnodes == 0) return MGBVH_OK;
  if (!d_nodes || !d_wide || !d_tri_index) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (nnodes >= 0x7fffffffull) { set_error("more than 2^31-1 nodes"); return MGBVH_ERR_INVALID_ARG; }
  mgbvh_wide8_params P;
  int rc = w8_params(params, P);
  if (rc) return rc;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const uint32_t nn = (uint32_t)nnodes;
  W8Forest f;
  std::vector<W8Tree> trees(1);
  trees[0] = W8Tree{0u, nn, 0u, 0u, 0u, 0u, 0u, 0u};
  if ((rc = w8_dp(c, d_nodes, d_lens, nn, trees, P, f))) return rc;
  mgbvh_node32 root;
  W8Rec r0{};
  if ((e = hipMemcpy(&root, d_nodes, sizeof root, hipMemcpyDeviceToHost)) != hipSuccess ||
      (e = hipMemcpy(&r0, f.rootrec, sizeof r0, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "wide8 root readback");
  const uint32_t L = root.child_node;
  if (L != 0 && L + 1 >= nn) { set_error("root child index out of range"); return MGBVH_ERR_INDEX; }
  const uint64_t nwide = L == 0 ? 1 : r0.w;
  const uint64_t nobj = r0.p;
  if (nwide > wide_cap || nobj > tri_cap) {
    set_error("wide8 output capacity too small (nodes " + std::to_string(nwide) + ", objects " +
              std::to_string(nobj) + ")");
    return MGBVH_ERR_INVALID_ARG;
  }
  const std::vector<W8Item> roots = {W8Item{0u, 0u, 1u, 0u, 0u, 0u, 0u, 0u}};
  if ((rc = w8_emit_forest(c, d_nodes, d_lens, nn, f, roots, d_wide, wide_cap, d_tri_index,
                           tri_cap)))
    return rc;
  if (out_nwide) *out_nwide = nwide;
  // a leaf root's cost is its leaf cost; an internal root is a wide internal
  // node: area * c_node + the best split of its 8 slots
  if (out_cost) *out_cost = L == 0 ? r0.c[0] : r0.c8;
  return MGBVH_OK;
}

int mgbvh_wide8_collapse_batch_device(mgbvh_ctx* c, const mgbvh_node32* d_nodes,
                                      const uint32_t* d_lens, uint64_t nnodes,
                                      const mgbvh_blas_desc* blas, const uint32_t* node_counts,
                                      uint32_t nblas, const mgbvh_wide8_params* params,
                                      mgbvh_wide_node* d_wide, uint64_t wide_cap,
                                      uint32_t* d_tri_index, uint64_t tri_cap,
                                      uint32_t* out_wide_first, uint32_t* out_wide_count,
                                      uint64_t* out_nwide) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (out_nwide) *out_nwide = 0;
  if (nblas == 0) return MGBVH_OK;
  if (!blas || !node_counts || !out_wide_first || !out_wide_count) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  bool any = false;
  for (uint32_t m = 0; m < nblas; ++m) {
    out_wide_first[m] = 0;
    out_wide_count[m] = 0;
    any |= node_counts[m] != 0;
    if ((uint64_t)blas[m].node_offset + node_counts[m] > nnodes) {
      set_error("BLAS " + std::to_string(m) + " runs past nnodes");
      return MGBVH_ERR_INVALID_ARG;
    }
  }
  mgbvh_wide8_params P;
  int rc = w8_params(params, P);
  if (rc) return rc;
  if (!any) return MGBVH_OK;  // every BLAS empty: nothing to collapse
  if (!d_nodes || !d_wide || !d_tri_index) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (nnodes >= 0x7fffffffull) { set_error("more than 2^31-1 nodes"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const uint32_t nn = (uint32_t)nnodes;
  Scratch& sc = c->sc;
  std::vector<W8Tree> trees(nblas);
  for (uint32_t m = 0; m < nblas; ++m)
    trees[m] = W8Tree{blas[m].node_offset, node_counts[m], blas[m].node_offset, 0u, 0u, 0u, 0u, 0u};
  W8Forest f;
  if ((rc = w8_dp(c, d_nodes, d_lens, nn, trees, P, f))) return rc;
  // root sizes of every tree (after the tree table in w8ctr)
  if ((rc = ensure(sc.w8base, 8 * (size_t)nblas))) return rc;
  uint32_t* d_sizes = reinterpret_cast<uint32_t*>(sc.w8base.p);
  hipStream_t st = c->stream;
  const W8Tree* d_trees =
      reinterpret_cast<const W8Tree*>(reinterpret_cast<const char*>(sc.w8ctr.p) + 32);
  k_w8_roots<<<(nblas + 255) / 256, 256, 0, st>>>(d_nodes, d_trees, f.rootrec, nblas, d_sizes);
  std::vector<uint32_t> sizes(2 * (size_t)nblas);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(sizes.data(), d_sizes, 8 * (size_t)nblas, hipMemcpyDeviceToHost, st)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return hip_fail(e, "wide8 root sizes");
  std::vector<W8Item> roots;
  roots.reserve(nblas);
  uint64_t wsum = 0;
  for (uint32_t m = 0; m < nblas; ++m) {
    const uint32_t w = sizes[2 * m], t = sizes[2 * m + 1];
    out_wide_first[m] = (uint32_t)wsum;
    out_wide_count[m] = w;
    if (w == 0) continue;
    if ((uint64_t)blas[m].tri_offset + t > tri_cap) {
      set_error("wide8 tri_index capacity too small for BLAS " + std::to_string(m));
      return MGBVH_ERR_INVALID_ARG;
    }
    roots.push_back(W8Item{blas[m].node_offset, 0u, 1u, 0u, (uint32_t)wsum, blas[m].tri_offset,
                           blas[m].node_offset, 0u});
    wsum += w;
  }
  if (wsum > wide_cap || wsum > 0xffffffffull) {
    set_error("wide8 output capacity too small (nodes " + std::to_string(wsum) + ")");
    return MGBVH_ERR_INVALID_ARG;
  }
  if ((rc = w8_emit_forest(c, d_nodes, d_lens, nn, f, roots, d_wide, wide_cap, d_tri_index,
                           tri_cap)))
    return rc;
  if (out_nwide) *out_nwide = wsum;
  return MGBVH_OK;
}

int mgbvh_trace_wide8_device(mgbvh_ctx* c, const mgbvh_wide_node* d_wide, uint64_t nwide,
                             const uint32_t* d_tri_index, const float* d_pos,
                             uint32_t vstride_bytes, const uint32_t* d_tris_perm,
                             const uint32_t* d_perm, const mgbvh_ray* d_rays, uint64_t nrays,
                             mgbvh_hit* d_hits) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (nrays == 0) return MGBVH_OK;
  if (!d_wide || !d_tri_index || !d_pos || !d_tris_perm || !d_perm || !d_rays || !d_hits ||
      nwide == 0 || nwide > 0xffffffffull) {
    set_error("null device buffer or no wide nodes");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (vstride_bytes < 12 || vstride_bytes % 4 || nrays > 0xffffffffull) {
    set_error("bad vstride_bytes or more than 2^32-1 rays");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  int rc = ensure(c->sc.w8ctr, 16);
  if (rc) return rc;
  uint32_t* flag = reinterpret_cast<uint32_t*>(c->sc.w8ctr.p);
  if ((e = hipMemsetAsync(flag, 0, 4, c->stream)) != hipSuccess) return hip_fail(e, "hipMemsetAsync");
  if (c->prof_mask & (1u << KK_TRACE)) {
    if ((rc = prof_begin(c, KK_TRACE))) return rc;
  }
  k_trace_w8<<<(uint32_t)((nrays + 255) / 256), 256, 0, c->stream>>>(
      reinterpret_cast<const uint8_t*>(d_wide), d_tri_index, d_pos, vstride_bytes / 4,
      d_tris_perm, d_perm, d_rays, (uint32_t)nrays, (uint32_t)nwide, d_hits, flag);
  if (c->prof_mask & (1u << KK_TRACE)) {
    if ((rc = prof_end(c, KK_TRACE))) return rc;
  }
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_trace_w8");
  uint32_t err = 0;
  if ((e = hipMemcpyAsync(&err, flag, 4, hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return hip_fail(e, "hipStreamSynchronize");
  if ((rc = prof_collect(c))) return rc;
  if (err & 2u) { set_error("a wide child index is past nwide"); return MGBVH_ERR_INDEX; }
  if (err) { set_error("wide traversal stack overflow"); return MGBVH_ERR_INTERNAL; }
  return MGBVH_OK;
}

}  // extern "C"
