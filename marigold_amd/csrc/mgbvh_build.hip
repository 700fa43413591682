// mgbvh plain builder: breadth-first SAH BVH construction on gfx950 that
// reproduces marigold's recursive builder (bvh/src/lib.rs:536-665) bit for bit.
//
// Pipeline for one object list of n boxes:
//   k_setup_*     object AABBs (mesh_interface/src/lib.rs:19-30) into 28-byte
//                 records, root bounds and root centroid bounds (lib.rs:220-228, 499-509)
//   level loop    for every node with > 32 objects (binned sweep, lib.rs:331-389):
//     k_bin       per chunk of 2048 objects: LDS-staged 3-axis x 32-bin histograms
//                 (count + bin box), merged into a per-node histogram
//     k_sweep     3 x 31 prefix-SAH candidates per node, first minimum under
//                 f32::total_cmp (lib.rs:373-388, 515-524), median fallback
//                 (lib.rs:527-531), node record, child allocation (one wavefront
//                 per node, or one lane per node on crowded levels)
//     k_partition stable coalesced partition by bin (select_nth_unstable at
//                 lib.rs:601-607 restated as a stable split), children's centroid
//                 bounds; a node that fits one block also bins its children and
//                 takes their splits
//   k_small       one wavefront per two subtrees of <= 32 objects: adaptive sweep
//                 (lib.rs:391-442) down to the leaves, entirely in LDS
//   k_compact     only when max_depth cut a subtree short: squeeze the slots
//                 reserved for the missing nodes
//
// Node numbering: the reference pushes each split's (gt, lt) pair at
// nodes.len() in depth-first, gt-first order (lib.rs:648-663).  With one object
// per leaf a subtree of n objects owns exactly 2(n-1) descendant slots, so a
// node's pair base P gives its children P(gt) = P + 2 and
// P(lt) = P + 2 + 2(n_gt - 1) without any global order.
#include "mgbvh_common.h"
#include "mgbvh_internal.h"

namespace mg {

// --------------------------------------------------------------- records
typedef float f4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float f3_a4 __attribute__((ext_vector_type(3), aligned(4)));
typedef uint32_t u3_a4 __attribute__((ext_vector_type(3), aligned(4)));
__device__ inline void rec_load(const Rec* r, Box& b, uint32_t& id) {
  const float* p = reinterpret_cast<const float*>(r);
  const f4_a4 a = *reinterpret_cast<const f4_a4*>(p);
  const f3_a4 c = *reinterpret_cast<const f3_a4*>(p + 4);
  b.mn[0] = a.x; b.mn[1] = a.y; b.mn[2] = a.z;
  b.mx[0] = a.w; b.mx[1] = c.x; b.mx[2] = c.y;
  id = __float_as_uint(c.z);
}
__device__ inline void rec_store(Rec* r, const Box& b, uint32_t id) {
  float* p = reinterpret_cast<float*>(r);
  f4_a4 a;
  a.x = b.mn[0]; a.y = b.mn[1]; a.z = b.mn[2]; a.w = b.mx[0];
  f3_a4 c;
  c.x = b.mx[1]; c.y = b.mx[2]; c.z = __uint_as_float(id);
  *reinterpret_cast<f4_a4*>(p) = a;
  *reinterpret_cast<f3_a4*>(p + 4) = c;
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of lane k (wave-uniform k) broadcast to the whole wavefront.
__device__ inline float rl(float v, uint32_t k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)k));
}

// min/max of coordinates that are finite and never -0 (object boxes are
// canonicalised at setup; centroids and unions of such boxes, and the empty
// box's infinities, keep that): the IEEE minimum/maximum then give the bits of
// the x86 min/max, in one instruction without operand canonicalisation.
__device__ inline float min_nz(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ inline float max_nz(float a, float b) { return __builtin_elementwise_maximum(a, b); }
__device__ inline void box_grow_nz(Box& acc, const Box& o) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    acc.mn[a] = min_nz(acc.mn[a], o.mn[a]);
    acc.mx[a] = max_nz(acc.mx[a], o.mx[a]);
  }
}

// Union of the boxes of lanes [ls, last] that have `take`, on every lane of
// the segment.  Box coordinates are finite and carry no -0 (object boxes are
// canonicalised at setup, centroids of such boxes cannot be -0), so the fold
// order does not change the result and the IEEE minimum/maximum (one
// instruction each, no operand canonicalisation) equal the x86 min/max.
__device__ inline Box seg_box(bool take, const float* mn, const float* mx, uint32_t lane,
                              uint32_t ls, uint32_t last) {
  const auto fmn = [](uint32_t a, uint32_t b) { return fbits(min_nz(bitsf(a), bitsf(b))); };
  const auto fmx = [](uint32_t a, uint32_t b) { return fbits(max_nz(bitsf(a), bitsf(b))); };
  uint32_t lo[3], hi[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    lo[x] = seg_scan32(take ? fbits(mn[x]) : 0x7f800000u, 0x7f800000u, lane, ls, fmn);
    hi[x] = seg_scan32(take ? fbits(mx[x]) : 0xff800000u, 0xff800000u, lane, ls, fmx);
  }
  Box r;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    r.mn[x] = bitsf((uint32_t)__shfl((int)lo[x], (int)last, 64));
    r.mx[x] = bitsf((uint32_t)__shfl((int)hi[x], (int)last, 64));
  }
  return r;
}

// Wave-wide maximum of a value in [0, 63] from six ballots.
__device__ inline uint32_t wave_max6(uint32_t v) {
  uint32_t m = 0;
#pragma unroll
  for (int b = 5; b >= 0; --b) {
    const uint32_t t = m | (1u << b);
    if (__ballot(v >= t)) m = t;
  }
  return m;
}

// Adds each valid lane's object to bin `bi` of axis `a` of an LDS histogram
// row (count and/or the box words e: ~okey(min) x3, okey(max) x3).  LDS
// serialises atomics on one address, and spatially coherent chunks put most
// of a wavefront into one bin (a heightfield row: all 64 lanes), so a bin
// shared by at least half the wavefront (at most two such bins) is merged
// first: the wave maximum of its box words and its popcount go in with one set
// of atomics.  Smaller groups go per lane: with four histogram copies their
// address conflicts cost less than a merge's DPP reductions (measured: a
// threshold of 8 lanes was 0.3 ms per scene build slower).  Counts and max
// folds are order independent: the row is the same as with per-lane atomics.
// Call with every lane of the wavefront active.
__device__ inline void hist_add(uint32_t* h, uint32_t a, uint32_t bi, bool valid,
                                const uint32_t (&e)[6], bool do_count, bool do_boxes) {
  const uint32_t lane = lane_id();
  uint64_t rem = __ballot(valid);
#pragma unroll 1
  for (int round = 0; round < 2 && rem; ++round) {
    const uint32_t lead = (uint32_t)__ffsll((unsigned long long)rem) - 1;
    const uint32_t B = (uint32_t)__builtin_amdgcn_readlane((int)bi, (int)lead);
    const bool in = ((rem >> lane) & 1ull) && bi == B;
    const uint64_t m = __ballot(in);
    const uint32_t k = (uint32_t)__popcll(m);
    if (k < 32) break;
    uint32_t* hb = &h[(a * MG_NBINS + B) * BIN_WORDS];
    if (do_boxes) {
      uint32_t r[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) r[j] = wave_max_u32_dpp(in ? e[j] : 0u);
      if (lane == lead) {
#pragma unroll
        for (int j = 0; j < 6; ++j) atomicMax(&hb[1 + j], r[j]);
      }
    }
    if (do_count && lane == lead) atomicAdd(&hb[0], k);
    rem &= ~m;
  }
  if ((rem >> lane) & 1ull) {
    uint32_t* hb = &h[(a * MG_NBINS + bi) * BIN_WORDS];
    if (do_count) atomicAdd(&hb[0], 1u);
    if (do_boxes) {
#pragma unroll
      for (int j = 0; j < 6; ++j) atomicMax(&hb[1 + j], e[j]);
    }
  }
}

__device__ inline bool coord_ok(float x) { return fabsf(x) < 1.152921504606846976e18f; }  // 2^60

// Block-wide max of 12 accumulator words, then one atomicMax each.
__device__ inline void block_acc12_flush(uint32_t (&acc)[12], uint32_t* dst) {
  __shared__ uint32_t red[4][12];
  const uint32_t w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = wave_max_u32_dpp(acc[k]);
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < 12; ++k) red[w][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    uint32_t v = red[0][threadIdx.x];
    for (uint32_t k = 1; k < (blockDim.x >> 6); ++k) v = max(v, red[k][threadIdx.x]);
    if (v) atomicMax(&dst[threadIdx.x], v);
  }
}

__device__ inline void acc12_add(uint32_t (&acc)[12], const Box& b) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    acc[a] = max(acc[a], acc_min_enc(b.mn[a]));
    acc[3 + a] = max(acc[3 + a], acc_max_enc(b.mx[a]));
    float c = center(b.mn[a], b.mx[a]);
    acc[6 + a] = max(acc[6 + a], acc_min_enc(c));
    acc[9 + a] = max(acc[9 + a], acc_max_enc(c));
  }
}

// ------------------------------------------------------------------ setup
// Mesh of setup chunk c: the last mesh whose first chunk is <= c (empty meshes
// own no chunks and share their first chunk with the next mesh).
__device__ inline uint32_t chunk_mesh(const MeshRow* rows, uint32_t nm, uint32_t c) {
  uint32_t lo = 0, hi = nm;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rows[mid].setup_chunk0 <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// MeshTriangle::bounding_volume (mesh_interface/src/lib.rs:19-30):
// min = v1.min(v2).min(v3), max likewise; -0.0 is stored as +0.0.  One block
// per chunk of CH triangles of one mesh; the mesh's root bounds and root
// centroid bounds (lib.rs:220-228, 499-509) are accumulated per mesh.
__global__ __launch_bounds__(256) void k_setup_tris(const float* __restrict__ pos, uint32_t vs,
                                                    const uint32_t* __restrict__ tris,
                                                    const MeshRow* __restrict__ rows,
                                                    uint32_t nm, Rec* __restrict__ recs,
                                                    uint32_t* __restrict__ macc, Counters* ctr) {
  __shared__ uint32_t s_m;
  if (threadIdx.x == 0) s_m = chunk_mesh(rows, nm, blockIdx.x);
  __syncthreads();
  const MeshRow& mr = rows[s_m];
  const uint32_t vcount = mr.vertex_count, voff = mr.vertex_offset;
  const uint32_t lo = mr.tri_offset + (blockIdx.x - mr.setup_chunk0) * CH;
  const uint32_t hi = min(lo + CH, mr.tri_offset + mr.count);
  uint32_t acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0;
  uint32_t err = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    uint32_t i0 = tris[3 * (size_t)i + 0];
    uint32_t i1 = tris[3 * (size_t)i + 1];
    uint32_t i2 = tris[3 * (size_t)i + 2];
    if (i0 >= vcount || i1 >= vcount || i2 >= vcount) {
      err |= ERR_INDEX;
      i0 = i1 = i2 = 0;
    }
    const float* v1 = pos + (size_t)(voff + i0) * vs;
    const float* v2 = pos + (size_t)(voff + i1) * vs;
    const float* v3 = pos + (size_t)(voff + i2) * vs;
    Box b;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float x1 = v1[a], x2 = v2[a], x3 = v3[a];
      if (!coord_ok(x1) || !coord_ok(x2) || !coord_ok(x3)) err |= ERR_NONFINITE;
      b.mn[a] = canon_zero(fmin_x86(fmin_x86(x1, x2), x3));
      b.mx[a] = canon_zero(fmax_x86(fmax_x86(x1, x2), x3));
    }
    rec_store(&recs[i], b, i);
    acc12_add(acc, b);
  }
  if (err) atomicOr(&ctr->err, err);
  block_acc12_flush(acc, macc + 12 * (size_t)s_m);
}

// Objects given as boxes (AsBoundingVolume, e.g. UploadedMesh at
// mesh_interface/src/lib.rs:87-91).
__global__ __launch_bounds__(256) void k_setup_aabbs(const float* __restrict__ boxes,
                                                     const MeshRow* __restrict__ rows,
                                                     uint32_t nm, Rec* __restrict__ recs,
                                                     uint32_t* __restrict__ macc, Counters* ctr) {
  __shared__ uint32_t s_m;
  if (threadIdx.x == 0) s_m = chunk_mesh(rows, nm, blockIdx.x);
  __syncthreads();
  const MeshRow& mr = rows[s_m];
  const uint32_t lo = mr.tri_offset + (blockIdx.x - mr.setup_chunk0) * CH;
  const uint32_t hi = min(lo + CH, mr.tri_offset + mr.count);
  uint32_t acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0;
  uint32_t err = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    Box b;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float lo_a = boxes[6 * (size_t)i + a], hi_a = boxes[6 * (size_t)i + 3 + a];
      if (!coord_ok(lo_a) || !coord_ok(hi_a)) err |= ERR_NONFINITE;
      b.mn[a] = canon_zero(lo_a);
      b.mx[a] = canon_zero(hi_a);
    }
    rec_store(&recs[i], b, i);
    acc12_add(acc, b);
  }
  if (err) atomicOr(&ctr->err, err);
  block_acc12_flush(acc, macc + 12 * (size_t)s_m);
}

// Root node of every mesh, lib.rs:220-243 and 731-735: a node with > 32
// objects starts the breadth-first levels, a smaller one (or any root when
// max_depth == 0) goes straight to the small-subtree kernel.  One wave per mesh.
__global__ void k_root(const MeshRow* __restrict__ rows, const uint32_t* __restrict__ macc,
                       Counters* ctr, LNode* L, uint32_t* cnode, uint32_t* ck, SNode* small,
                       uint32_t n_small_roots, uint32_t learn_min, uint32_t* lrows,
                       uint32_t* hist, uint32_t* lv0, uint32_t n_large, uint32_t n_chunks,
                       uint32_t n_counted) {
  const uint32_t m = blockIdx.x;
  if (m == 0 && threadIdx.x == 0) {
    ctr->n_small = n_small_roots;
    lv0[LV_NL] = n_large;
    lv0[LV_NC] = n_chunks;
    lv0[LV_NCOUNTED] = n_counted;
    lv0[LV_NBIN] = n_chunks;  // the roots are binned by k_bin
  }
  const MeshRow mr = rows[m];
  if (mr.root_kind == 0) return;
  const uint32_t* acc = macc + 12 * (size_t)m;
  if (threadIdx.x == 0) {
    float bmn[3], bmx[3];
    for (int a = 0; a < 3; ++a) {
      bmn[a] = mr.has_bounds ? mr.bmn[a] : acc_min_dec(acc[a]);
      bmx[a] = mr.has_bounds ? mr.bmx[a] : acc_max_dec(acc[3 + a]);
    }
    if (mr.root_kind == 1) {
      LNode r = {};
      r.start = mr.tri_offset; r.count = mr.count; r.idx = mr.nbase; r.P = mr.nbase + 1;
      r.depth = 0; r.chunk0 = mr.lchunk0; r.nchunk = (mr.count + CH - 1) / CH;
      for (int a = 0; a < 3; ++a) {
        r.bmn[a] = bmn[a]; r.bmx[a] = bmx[a];
        r.cacc[a] = acc[6 + a];
        r.cacc[3 + a] = acc[9 + a];
      }
      r.child_slot[0] = r.child_slot[1] = -1;
      r.nbase = mr.nbase; r.obase = mr.tri_offset;
      r.mode = (learn_min && mr.count >= learn_min) ? LM_COUNTED : LM_PLAIN;
      if (r.mode == LM_COUNTED) lrows[mr.lrow] = mr.root_slot;
      L[mr.root_slot] = r;
    } else {
      SNode sn = {};
      sn.start = mr.tri_offset; sn.count = mr.count; sn.idx = mr.nbase; sn.P = mr.nbase + 1;
      sn.depth = 0; sn.parity = 0; sn.nbase = mr.nbase; sn.obase = mr.tri_offset;
      for (int a = 0; a < 3; ++a) { sn.bmn[a] = bmn[a]; sn.bmx[a] = bmx[a]; }
      small[mr.root_slot] = sn;
    }
  }
  if (mr.root_kind == 1) {
    const uint32_t nch = (mr.count + CH - 1) / CH;
    for (uint32_t k = threadIdx.x; k < nch; k += blockDim.x) {
      cnode[mr.lchunk0 + k] = mr.root_slot;
      ck[mr.lchunk0 + k] = k;
    }
    // several chunks merge into this row with atomics: it starts at zero
    if (nch > 1)
      for (uint32_t t = threadIdx.x; t < HIST_WORDS; t += blockDim.x)
        hist[(size_t)mr.root_slot * HIST_WORDS + t] = 0;
  }
}

// ------------------------------------------------------------------ binning
// One block per chunk of <= CH objects of one node: 3 x 32 bins of
// {count, bin box} in LDS, flushed into the node histogram; per-chunk counts
// are kept for the stable partition.
// Every node is binned completely (counts and bin boxes on all three axes),
// also a learned-method node: its features read the counts and its window
// restricts the sweep to cuts of the predicted axis, and binning the boxes of
// the other axes in the same pass costs less than a second read of the
// records would.
__device__ inline void bin_chunk(uint32_t c, const LNode* __restrict__ L,
                                 const uint32_t* __restrict__ cnode,
                                 const uint32_t* __restrict__ ck, const Rec* __restrict__ recs,
                                 uint32_t* __restrict__ hist, uint32_t* __restrict__ blk) {
  // BIN_COPIES copies of the row, lane l adding to copy l % BIN_COPIES: lanes of
  // one wavefront that share a bin mostly hit different addresses (LDS
  // serialises same-address atomics).  The odd stride puts the copies of a
  // word in different banks.  The flush merges the copies (sums and max folds
  // are order independent).
  constexpr uint32_t BIN_COPIES = 4, COPY_STRIDE = HIST_WORDS + 1;
  __shared__ uint32_t h[BIN_COPIES * COPY_STRIDE];
  const uint32_t slot = cnode[c];
  const uint32_t k = ck[c];
  const LNode& nd = L[slot];
  if (nd.flags & LF_PREBINNED) return;  // block-uniform: binned by the parent's partition
  const bool single = nd.nchunk == 1;
  for (uint32_t t = threadIdx.x; t < BIN_COPIES * COPY_STRIDE; t += blockDim.x) h[t] = 0;
  uint32_t* hl = h + (lane_id() % BIN_COPIES) * COPY_STRIDE;
  float cmin[3], cext[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    cmin[a] = acc_min_dec(nd.cacc[a]);
    float cmax = acc_max_dec(nd.cacc[3 + a]);
    cext[a] = __fsub_rn(cmax, cmin[a]);
  }
  __syncthreads();
  const uint32_t lo = nd.start + k * CH;
  const uint32_t hi = min(lo + CH, nd.start + nd.count);
  // whole-block steps (every lane active for the wavefront merges in hist_add)
  for (uint32_t base = lo; base < hi; base += blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < hi;
    Box b;
    uint32_t id = 0;
#pragma unroll
    for (int x = 0; x < 3; ++x) { b.mn[x] = 0.0f; b.mx[x] = 0.0f; }
    if (valid) rec_load(&recs[i], b, id);
    uint32_t e[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) { e[j] = acc_min_enc(b.mn[j]); e[3 + j] = acc_max_enc(b.mx[j]); }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const uint32_t bi = valid ? bin_index(center(b.mn[a], b.mx[a]), cmin[a], cext[a]) : 0u;
      hist_add(hl, (uint32_t)a, bi, valid, e, true, true);
    }
  }
  __syncthreads();
  uint32_t* gh = hist + (size_t)slot * HIST_WORDS;
  for (uint32_t t = threadIdx.x; t < HIST_WORDS; t += blockDim.x) {
    uint32_t f = t % BIN_WORDS;
    uint32_t v = h[t];
#pragma unroll
    for (uint32_t q = 1; q < BIN_COPIES; ++q) {
      const uint32_t u = h[q * COPY_STRIDE + t];
      v = f == 0 ? v + u : max(v, u);
    }
    if (f == 0) {
      blk[(size_t)c * 96 + t / BIN_WORDS] = v;
      if (single) gh[t] = v;
      else if (v) atomicAdd(&gh[t], v);
    } else if (single) {
      gh[t] = v;
    } else if (v) {
      atomicMax(&gh[t], v);
    }
  }
}

// Persistent over the level's chunk queue (count in device memory): each block
// takes chunks blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ __launch_bounds__(256) void k_bin(const LNode* __restrict__ L,
                                             const uint32_t* __restrict__ cnode,
                                             const uint32_t* __restrict__ ck,
                                             const Rec* __restrict__ recs,
                                             uint32_t* __restrict__ hist,
                                             uint32_t* __restrict__ blk,
                                             const uint32_t* __restrict__ lvc) {
  // Every level of a recorded graph launches it.  No chunk to bin (the lower
  // levels' nodes were all binned by their parents' partition blocks) means
  // nothing to do.
  if (lvc[LV_NBIN] == 0) return;
  const uint32_t nC = lvc[LV_NC];
  for (uint32_t c = blockIdx.x; c < nC; c += gridDim.x) {
    bin_chunk(c, L, cnode, ck, recs, hist, blk);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ sweep

// Child record of a split: a node with > 32 objects for the next level (slot s,
// chunks [c0, c0 + nch)), or a subtree root of <= 32 objects / a depth-cut leaf.
__device__ inline void put_child(uint32_t kind, uint32_t s, uint32_t c0, uint32_t start,
                                 uint32_t count, const Box& b, uint32_t idx, uint32_t P,
                                 uint32_t depth, uint32_t parity_next, uint32_t nbase,
                                 uint32_t obase, LNode* Ln, SNode* small, SNode* med,
                                 uint32_t learn_min, uint32_t* lvn, uint32_t* lrows,
                                 bool prebinned) {
  if (kind == 1) {
    LNode r = {};
    r.start = start; r.count = count; r.idx = idx; r.P = P; r.depth = depth;
    r.chunk0 = c0; r.nchunk = (count + CH - 1) / CH;
    for (int a = 0; a < 3; ++a) { r.bmn[a] = b.mn[a]; r.bmx[a] = b.mx[a]; }
    r.child_slot[0] = r.child_slot[1] = -1;
    r.nbase = nbase; r.obase = obase;
    r.flags = prebinned ? LF_PREBINNED : 0u;
    r.mode = (!prebinned && learn_min && count >= learn_min) ? LM_COUNTED : LM_PLAIN;
    if (r.mode == LM_COUNTED) lrows[atomicAdd(&lvn[LV_NCOUNTED], 1u)] = s;
    Ln[s] = r;
  } else {
    SNode sn = {};
    sn.start = start; sn.count = count; sn.idx = idx; sn.P = P; sn.depth = depth;
    sn.parity = parity_next; sn.nbase = nbase; sn.obase = obase;
    for (int a = 0; a < 3; ++a) { sn.bmn[a] = b.mn[a]; sn.bmx[a] = b.mx[a]; }
    (kind == 2 ? med : small)[s] = sn;
  }
}

// Where a child of a split goes (lib.rs:648-663): 0 the small-subtree list
// (<= 32 objects, or a depth-cut leaf), 2 the medium list (33..med_max
// objects, not on the learned path), 1 the next level.
__device__ inline uint32_t child_route(uint32_t count, uint32_t cd, uint32_t max_depth,
                                       uint32_t med_max, uint32_t learn_min) {
  if (count <= MG_SMALL || cd >= max_depth) return 0;
  if (count <= med_max && (learn_min == 0 || count < learn_min)) return 2;
  return 1;
}

// The split of a node with > 32 objects (select_split, lib.rs:493-534).
struct SplitDec {
  Box lt, gt;
  uint32_t nl, ng, axis, value, kind;  // kind 0 binned cut, 2 median
  float cost;
  uint32_t pad[2];
};
static_assert(sizeof(SplitDec) == 80, "SplitDec is 80 bytes");
static_assert(sizeof(SplitDec) <= HIST_WORDS * 4, "a SplitDec fits a histogram row");

// Inclusive scans inside each 16-lane DPP row, toward higher lanes (prefix)
// or lower lanes (suffix); lanes past the row's edge contribute the identity,
// so no segment test is needed.
template <class Op>
__device__ inline uint32_t row_prefix(uint32_t v, uint32_t id, Op op) {
  v = op(v, dpp_mov<0x111, 0xF>(id, v));
  v = op(v, dpp_mov<0x112, 0xF>(id, v));
  v = op(v, dpp_mov<0x114, 0xF>(id, v));
  v = op(v, dpp_mov<0x118, 0xF>(id, v));
  return v;
}
template <class Op>
__device__ inline uint32_t row_suffix(uint32_t v, uint32_t id, Op op) {
  v = op(v, dpp_mov<0x101, 0xF>(id, v));
  v = op(v, dpp_mov<0x102, 0xF>(id, v));
  v = op(v, dpp_mov<0x104, 0xF>(id, v));
  v = op(v, dpp_mov<0x108, 0xF>(id, v));
  return v;
}

// One wavefront: the split of a node from its binned histogram h (global
// memory or LDS), node box nb, centroid bounds (cmin, cext), its objects
// [start, start + count) in recs (median fallback only) and its learned-method
// state lm (a window node only considers its window's cuts).  Every lane
// returns the same decision.
//
// Layout: DPP row a (lanes 16a .. 16a + 15) holds axis a, lane j of the row
// bins 2j and 2j + 1, and evaluates the cuts b = 2j and 2j + 1 (lt = bins
// [0, b), gt = bins [b, 32)).  Bin words are a count and ~okey(min) x3,
// okey(max) x3 (0 = empty), so every fold is a sum or a u32 maximum with
// identity 0 and each scan step is one DPP-fused instruction.
__device__ inline SplitDec split_from_hist(const uint32_t* h, const Box& nb,
                                           const float (&cmin)[3], const float (&cext)[3],
                                           uint32_t start, uint32_t count,
                                           const Rec* __restrict__ recs, uint32_t lm) {
  const uint32_t lane = lane_id();
  SplitDec d;
  const float sa_p = box_sa(nb);
  const bool win = lmode_kind(lm) == LM_WINDOW;
  const uint32_t wa = lmode_axis(lm);
  const uint32_t wlo = lmode_lo(lm), whi = lmode_hi(lm);
  const uint32_t row = lane >> 4, j = lane & 15u;
  const bool aact = row < 3;
  const uint32_t a = aact ? row : 0u;
  uint32_t c0 = 0, c1 = 0;
  uint32_t e0[6] = {0u, 0u, 0u, 0u, 0u, 0u}, e1[6] = {0u, 0u, 0u, 0u, 0u, 0u};
  if (aact) {
    const uint32_t* hb = &h[(a * MG_NBINS + 2 * j) * BIN_WORDS];
    c0 = hb[0];
    c1 = hb[BIN_WORDS];
    if (c0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) e0[k] = hb[1 + k];
    }
    if (c1) {
#pragma unroll
      for (int k = 0; k < 6; ++k) e1[k] = hb[BIN_WORDS + 1 + k];
    }
  }
  const auto add = [](uint32_t x, uint32_t y) { return x + y; };
  const auto umax = [](uint32_t x, uint32_t y) { return x > y ? x : y; };
  // counts: E = bins below the lane's pair, S = bins above it
  const uint32_t pc = c0 + c1;
  const uint32_t Ec = dpp_mov<0x111, 0xF>(0u, row_prefix(pc, 0u, add));
  const uint32_t Sc = dpp_mov<0x101, 0xF>(0u, row_suffix(pc, 0u, add));
  const uint32_t nl0 = Ec, ng0 = pc + Sc;         // cut 2j
  const uint32_t nl1 = Ec + c0, ng1 = c1 + Sc;    // cut 2j + 1
  Box lt0, gt0, lt1, gt1;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const uint32_t pe = umax(e0[k], e1[k]);
    const uint32_t E = dpp_mov<0x111, 0xF>(0u, row_prefix(pe, 0u, umax));
    const uint32_t S = dpp_mov<0x101, 0xF>(0u, row_suffix(pe, 0u, umax));
    const uint32_t l0 = E, g0 = umax(pe, S), l1 = umax(E, e0[k]), g1 = umax(e1[k], S);
    if (k < 3) {
      lt0.mn[k] = acc_min_dec(l0); gt0.mn[k] = acc_min_dec(g0);
      lt1.mn[k] = acc_min_dec(l1); gt1.mn[k] = acc_min_dec(g1);
    } else {
      lt0.mx[k - 3] = acc_max_dec(l0); gt0.mx[k - 3] = acc_max_dec(g0);
      lt1.mx[k - 3] = acc_max_dec(l1); gt1.mx[k - 3] = acc_max_dec(g1);
    }
  }
  // candidate key (cost, axis * 32 + cut): the first minimum under total_cmp,
  // axes 0, 1, 2, cuts ascending (lib.rs:373-388, 515-524)
  uint64_t best = ~0ull;
  // the lane's axis extent by selects (a run-time index would put cext in scratch)
  const float ce0 = cext[0], ce1 = cext[1], ce2 = cext[2];
  const float ce = a == 0 ? ce0 : (a == 1 ? ce1 : ce2);
  const auto consider = [&](uint32_t b, uint32_t nl, uint32_t ng, const Box& l, const Box& g) {
    // lib.rs:339-341 (zero extent refused), 379-381 (MIN objects per side)
    const bool inwin = !win || (a == wa && b >= wlo && b <= whi);
    if (aact && inwin && b >= 1 && ce != 0.0f && nl >= 1 && ng >= 1) {
      const float c = split_cost(sa_p, box_sa_ne(l), box_sa_ne(g), nl, ng);  // nl, ng >= 1
      const uint64_t key = ((uint64_t)tc_ukey(c) << 32) | (a * MG_NBINS + b);
      best = key < best ? key : best;
    }
  };
  consider(2 * j, nl0, ng0, lt0, gt0);
  consider(2 * j + 1, nl1, ng1, lt1, gt1);
  best = wave_min_u64_dpp(best, lane);

  if (best != ~0ull) {
    const uint32_t q = (uint32_t)best;
    d.axis = q / MG_NBINS;
    d.value = q % MG_NBINS;
    d.kind = 0;
    // the winning candidate's boxes and counts, from the lane that evaluated it
    // (prefix / suffix unions of exact min/max: the same values as summing bins)
    const uint32_t src = __builtin_amdgcn_readfirstlane(16u * d.axis + (d.value >> 1));
    const bool odd = d.value & 1u;  // wave-uniform
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      d.lt.mn[x] = rl(odd ? lt1.mn[x] : lt0.mn[x], src);
      d.lt.mx[x] = rl(odd ? lt1.mx[x] : lt0.mx[x], src);
      d.gt.mn[x] = rl(odd ? gt1.mn[x] : gt0.mn[x], src);
      d.gt.mx[x] = rl(odd ? gt1.mx[x] : gt0.mx[x], src);
    }
    d.nl = (uint32_t)__builtin_amdgcn_readlane((int)(odd ? nl1 : nl0), (int)src);
    d.ng = (uint32_t)__builtin_amdgcn_readlane((int)(odd ? ng1 : ng0), (int)src);
    d.cost = tc_ukey_inv((uint32_t)(best >> 32));  // the winner's cost, as its lane computed it
  } else {
    // Median fallback (lib.rs:444-491, 527-531).  With > 32 objects it is only
    // reached when every centroid coincides, so each axis' select keeps the
    // list order and axis 0 wins: lt = first n/2 objects.
    const uint32_t m = count / 2;
    uint32_t acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0;
    for (uint32_t j = lane; j < count; j += 64) {
      Box b;
      uint32_t id;
      rec_load(&recs[start + j], b, id);
      const int o = j < m ? 0 : 6;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        acc[o + a] = max(acc[o + a], acc_min_enc(b.mn[a]));
        acc[o + 3 + a] = max(acc[o + 3 + a], acc_max_enc(b.mx[a]));
      }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = wave_max_u32_dpp(acc[k]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      d.lt.mn[a] = acc_min_dec(acc[a]);
      d.lt.mx[a] = acc_max_dec(acc[3 + a]);
      d.gt.mn[a] = acc_min_dec(acc[6 + a]);
      d.gt.mx[a] = acc_max_dec(acc[9 + a]);
    }
    d.nl = m;
    d.ng = count - m;
    d.axis = 0;
    d.value = m;
    d.kind = 2;
    d.cost = split_cost(sa_p, box_sa(d.lt), box_sa(d.gt), d.nl, d.ng);
  }
  d.pad[0] = d.pad[1] = 0;
  return d;
}

constexpr uint32_t SW_WAVES = 16;  // wavefronts per k_sweep block

// Training sample of the split proposer: a node's bin counts, its extents and
// the split it got (whole wavefront; nodes are not decided early while
// samples are collected).
__device__ inline void put_sample(const NodeOut& out, Counters* ctr, const uint32_t* h,
                                  const LNode& nd, const Box& nb, const float (&cext)[3],
                                  const SplitDec& d) {
  if (!out.samples || d.kind != 0 || nd.count < out.sample_min) return;  // wave-uniform
  const uint32_t lane = lane_id();
  uint32_t si = 0;
  if (lane == 0) si = atomicAdd(&ctr->n_samples, 1u);
  si = __shfl(si, 0, 64);
  if (si < out.max_samples) {
    float* f = out.samples + (size_t)si * MLP_IN;
    const float inv = 1.0f / log1pf((float)nd.count);
    for (uint32_t q = lane; q < 3 * MG_NBINS; q += 64) f[q] = log1pf((float)h[q * BIN_WORDS]) * inv;
    if (lane == 0) {
      node_extent_features(nb.mn, nb.mx, cext, f + 3 * MG_NBINS);
      out.labels[si] = d.axis * MG_NBINS + d.value;
    }
  }
}

// select_split (lib.rs:493-534) of a node from its histogram row h, by the
// whole wavefront (every lane returns the decision).
__device__ inline SplitDec sweep_split(const LNode& nd, const uint32_t* h,
                                       const Rec* __restrict__ recs, const NodeOut& out,
                                       Counters* ctr) {
  Box nb;
  float cmin[3], cext[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    nb.mn[a] = nd.bmn[a];
    nb.mx[a] = nd.bmx[a];
    cmin[a] = acc_min_dec(nd.cacc[a]);
    cext[a] = __fsub_rn(acc_max_dec(nd.cacc[3 + a]), cmin[a]);
  }
  const SplitDec d = split_from_hist(h, nb, cmin, cext, nd.start, nd.count, recs, nd.mode);
  put_sample(out, ctr, h, nd, nb, cext, d);
  return d;
}

// What a split node asks of the next level: large slots, chunks, small slots,
// medium slots (the four lists), objects (work counter), window nodes,
// fallback nodes, chunks left to k_bin.  Children go where child_route says.
constexpr int SW_NEEDS = 8, SW_LISTS = 4;
struct SweepCfg {
  uint32_t max_depth, med_max, learn_min;
};
__device__ inline void sweep_need(const LNode& nd, const SplitDec& d, const SweepCfg& cf,
                                  uint32_t (&need)[SW_NEEDS]) {
  const uint32_t cd = nd.depth + 1;
  const uint32_t rg = child_route(d.ng, cd, cf.max_depth, cf.med_max, cf.learn_min);
  const uint32_t rl = child_route(d.nl, cd, cf.max_depth, cf.med_max, cf.learn_min);
  need[0] = (uint32_t)(rg == 1) + (uint32_t)(rl == 1);
  need[1] = (rg == 1 ? (d.ng + CH - 1) / CH : 0) + (rl == 1 ? (d.nl + CH - 1) / CH : 0);
  need[2] = (uint32_t)(rg == 0) + (uint32_t)(rl == 0);
  need[3] = (uint32_t)(rg == 2) + (uint32_t)(rl == 2);
  need[4] = nd.count;
  need[5] = lmode_kind(nd.mode) == LM_WINDOW;
  need[6] = lmode_kind(nd.mode) == LM_FALLBACK;
  need[7] = nd.nchunk > 1 ? need[1] : 0u;  // children of a one-block node are binned there
}

// Adds a block's total of need k to the next level's list or counter; for
// the four lists the base of the block's range comes back in s_gbase[k].
__device__ inline void sweep_alloc(uint32_t k, uint32_t tot, uint32_t* s_gbase, Counters* ctr,
                                   uint32_t* lvn) {
  if (k < SW_LISTS) {
    uint32_t* counter = k == 0 ? &lvn[LV_NL]
                        : k == 1 ? &lvn[LV_NC]
                        : k == 2 ? &ctr->n_small
                                 : &ctr->n_med;
    s_gbase[k] = tot ? atomicAdd(counter, tot) : 0u;
  } else if (tot) {
    if (k == 4) atomicAdd(&ctr->sum_large, (unsigned long long)tot);
    else atomicAdd(k == 5 ? &ctr->n_window : (k == 6 ? &ctr->n_fallback : &lvn[LV_NBIN]), tot);
  }
}

// Where a split node's children go, from its bases in the next level's
// large (bl), chunk (bc), small (bs) and medium (bm) lists.
struct SweepSlots {
  uint32_t rg, rl;  // child_route of the gt and lt child
  uint32_t sgs, sls, c0g, c0l, chg, chl;
};
__device__ inline SweepSlots sweep_slots(const LNode& nd, const SplitDec& d, uint32_t bl,
                                         uint32_t bc, uint32_t bs, uint32_t bm,
                                         const SweepCfg& cf) {
  SweepSlots r;
  const uint32_t cd = nd.depth + 1;
  r.rg = child_route(d.ng, cd, cf.max_depth, cf.med_max, cf.learn_min);
  r.rl = child_route(d.nl, cd, cf.max_depth, cf.med_max, cf.learn_min);
  r.chg = r.rg == 1 ? (d.ng + CH - 1) / CH : 0;
  r.chl = r.rl == 1 ? (d.nl + CH - 1) / CH : 0;
  // gt child slot first in its list, then the lt child's (after the gt child
  // when both go to the same list)
  r.sgs = r.rg == 1 ? bl : (r.rg == 2 ? bm : bs);
  r.sls = (r.rl == 1 ? bl : (r.rl == 2 ? bm : bs)) + (uint32_t)(r.rl == r.rg);
  r.c0g = bc;
  r.c0l = bc + r.chg;
  return r;
}

// Node record, split record, children and the node's split fields (read by
// the partition).  One lane; the chunk lists are written by the caller.
__device__ inline void sweep_emit(LNode* __restrict__ L, uint32_t slot, const LNode& nd,
                                  const SplitDec& d, const SweepSlots& r, LNode* Ln,
                                  SNode* small, SNode* med, Counters* ctr, const NodeOut& out,
                                  uint32_t parity_next, uint32_t learn_min, uint32_t* lrows_next,
                                  uint32_t* lvn) {
  Box nb;
#pragma unroll
  for (int a = 0; a < 3; ++a) { nb.mn[a] = nd.bmn[a]; nb.mx[a] = nd.bmx[a]; }
  const uint32_t cd = nd.depth + 1;
  out.write_node(nd.idx, nb, nd.start - nd.obase, nd.P - nd.nbase, nd.count);
  out.write_split(nd.idx, d.axis | (d.kind << 8), d.value, d.nl, fbits(d.cost));
  if (d.kind == 2) atomicAdd(&ctr->n_median, 1u);
  // gt child at P, lt child at P + 1
  put_child(r.rg, r.sgs, r.c0g, nd.start + d.nl, d.ng, d.gt, nd.P, nd.P + 2, cd, parity_next,
            nd.nbase, nd.obase, Ln, small, med, learn_min, lvn, lrows_next, nd.nchunk == 1);
  put_child(r.rl, r.sls, r.c0l, nd.start, d.nl, d.lt, nd.P + 1, nd.P + 2 + 2 * (d.ng - 1), cd,
            parity_next, nd.nbase, nd.obase, Ln, small, med, learn_min, lvn, lrows_next,
            nd.nchunk == 1);
  L[slot].axis = d.axis;
  L[slot].cut = d.value;
  L[slot].n_lt = d.nl;
  L[slot].kind = d.kind;
  // only next-level children are prebinned by the partition (medium ones bin themselves)
  L[slot].child_slot[0] = r.rg == 1 ? (int32_t)r.sgs : -1;
  L[slot].child_slot[1] = r.rl == 1 ? (int32_t)r.sls : -1;
}

// Levels with few, mostly large nodes: one wavefront per node.  The split
// comes from the node's histogram (whole wavefront) or is the decision its
// parent's partition block already took.  A block appends the children of
// its SW_WAVES nodes with one atomic per list: a same-address atomic costs
// ~11 ns, serialised across the whole device.
__device__ inline void sweep_group(uint32_t group, LNode* __restrict__ L, uint32_t nL,
                                   const uint32_t* __restrict__ hist,
                                   const Rec* __restrict__ recs, LNode* Ln, uint32_t* cn_next,
                                   uint32_t* ck_next, SNode* small, SNode* med, Counters* ctr,
                                   NodeOut out, const SweepCfg& cf, uint32_t parity_next,
                                   uint32_t* lrows_next, uint32_t* lvn) {
  __shared__ uint32_t s_need[SW_WAVES][SW_NEEDS];
  __shared__ uint32_t s_base[SW_WAVES][SW_LISTS];
  __shared__ uint32_t s_gbase[SW_LISTS];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint32_t slot = group * SW_WAVES + w;
  const bool active = slot < nL;  // wave-uniform
  uint32_t need[SW_NEEDS] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  LNode nd = {};
  SplitDec d = {};
  if (active) {
    nd = L[slot];
    const uint32_t* h = hist + (size_t)slot * HIST_WORDS;
    if (nd.flags & LF_DECIDED)  // taken by the parent's partition block (same function, same histogram)
      d = *reinterpret_cast<const SplitDec*>(h);
    else
      d = sweep_split(nd, h, recs, out, ctr);
    sweep_need(nd, d, cf, need);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < SW_NEEDS; ++k) s_need[w][k] = need[k];
  }
  __syncthreads();
  if (threadIdx.x < SW_NEEDS) {
    const uint32_t k = threadIdx.x;
    uint32_t tot = 0;
    for (uint32_t q = 0; q < SW_WAVES; ++q) tot += s_need[q][k];
    sweep_alloc(k, tot, s_gbase, ctr, lvn);
    if (k < SW_LISTS) {
      uint32_t run = s_gbase[k];
      for (uint32_t q = 0; q < SW_WAVES; ++q) {
        s_base[q][k] = run;
        run += s_need[q][k];
      }
    }
  }
  __syncthreads();
  if (!active) return;
  const SweepSlots r =
      sweep_slots(nd, d, s_base[w][0], s_base[w][1], s_base[w][2], s_base[w][3], cf);
  if (lane == 0)
    sweep_emit(L, slot, nd, d, r, Ln, small, med, ctr, out, parity_next, cf.learn_min,
               lrows_next, lvn);
  for (uint32_t k = lane; k < r.chg; k += 64) { cn_next[r.c0g + k] = r.sgs; ck_next[r.c0g + k] = k; }
  for (uint32_t k = lane; k < r.chl; k += 64) { cn_next[r.c0l + k] = r.sls; ck_next[r.c0l + k] = k; }
}

// Crowded levels (the lower levels: hundreds of thousands of small nodes,
// nearly all decided by their parents' partition blocks): one node per lane,
// SW_WAVES * 64 nodes per block.  A decided node needs no cross-lane work, so
// each lane loads its own node and decision and writes its own records, and
// the loads of 64 nodes are in flight together instead of one node's after
// another's.  The block's undecided nodes are listed in LDS and split by its
// wavefronts in turn (whole wavefront per node); each decision goes to the
// start of the node's histogram row (not read again) for its lane.  Child
// slots: exclusive DPP prefix sums inside each wavefront, wavefront totals
// through LDS, one atomic per list per block.
__device__ inline void sweep_lanes(uint32_t group, LNode* __restrict__ L, uint32_t nL,
                                   uint32_t* __restrict__ hist, const Rec* __restrict__ recs,
                                   LNode* Ln, uint32_t* cn_next, uint32_t* ck_next, SNode* small,
                                   SNode* med, Counters* ctr, NodeOut out, const SweepCfg& cf,
                                   uint32_t parity_next, uint32_t* lrows_next, uint32_t* lvn) {
  __shared__ uint32_t s_wtot[SW_WAVES][SW_NEEDS];
  __shared__ uint32_t s_gbase[SW_LISTS];
  __shared__ uint32_t s_und[SW_WAVES * 64];
  __shared__ uint32_t s_nund;
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint32_t slot = group * SW_WAVES * 64 + threadIdx.x;
  const bool active = slot < nL;
  if (threadIdx.x == 0) s_nund = 0;
  LNode nd = {};
  bool decided = false;
  if (active) {
    nd = L[slot];
    decided = (nd.flags & LF_DECIDED) != 0;
  }
  __syncthreads();
  const uint64_t um = __ballot(active && !decided);
  if (um) {  // wave-uniform
    uint32_t at = 0;
    if (lane == 0) at = atomicAdd(&s_nund, (uint32_t)__popcll(um));
    at = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
    if (active && !decided) s_und[at + lanemask_lt_popc(um)] = slot;
  }
  __syncthreads();
  const uint32_t nund = s_nund;
  for (uint32_t i = w; i < nund; i += SW_WAVES) {
    const uint32_t s2 = s_und[i];
    uint32_t* h2 = hist + (size_t)s2 * HIST_WORDS;
    const SplitDec d2 = sweep_split(L[s2], h2, recs, out, ctr);
    if (lane == 0) *reinterpret_cast<SplitDec*>(h2) = d2;
  }
  if (nund) __syncthreads();  // block-uniform: the decisions are visible to their lanes
  SplitDec d = {};
  if (active) d = *reinterpret_cast<const SplitDec*>(hist + (size_t)slot * HIST_WORDS);
  uint32_t need[SW_NEEDS] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  if (active) sweep_need(nd, d, cf, need);
  uint32_t pre[SW_LISTS];
#pragma unroll
  for (int k = 0; k < SW_NEEDS; ++k) {
    const uint32_t inc = wave_scan_u32_dpp(need[k]);
    if (k < SW_LISTS) pre[k] = inc - need[k];
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    if (lane == 0) s_wtot[w][k] = tot;
  }
  __syncthreads();
  if (threadIdx.x < SW_NEEDS) {
    const uint32_t k = threadIdx.x;
    uint32_t tot = 0;
    for (uint32_t q = 0; q < SW_WAVES; ++q) tot += s_wtot[q][k];
    sweep_alloc(k, tot, s_gbase, ctr, lvn);
  }
  __syncthreads();
  if (!active) return;
  uint32_t base[SW_LISTS];
#pragma unroll
  for (int k = 0; k < SW_LISTS; ++k) {
    uint32_t b = s_gbase[k] + pre[k];
    for (uint32_t q = 0; q < w; ++q) b += s_wtot[q][k];
    base[k] = b;
  }
  const SweepSlots r = sweep_slots(nd, d, base[0], base[1], base[2], base[3], cf);
  sweep_emit(L, slot, nd, d, r, Ln, small, med, ctr, out, parity_next, cf.learn_min, lrows_next,
             lvn);
  for (uint32_t k = 0; k < r.chg; ++k) { cn_next[r.c0g + k] = r.sgs; ck_next[r.c0g + k] = k; }
  for (uint32_t k = 0; k < r.chl; ++k) { cn_next[r.c0l + k] = r.sls; ck_next[r.c0l + k] = k; }
}

// Persistent over the level's node queue.  lvc: this level's counts; lvn: the
// next level's, filled by the allocation.  Levels with at least lanes_min
// nodes take one node per lane, the others one node per wavefront (their
// nodes are large and mostly swept from histograms: parallelism first).
__global__ __launch_bounds__(1024) void k_sweep(LNode* __restrict__ L,
                                                const uint32_t* __restrict__ lvc,
                                                uint32_t* __restrict__ hist,
                                                const Rec* __restrict__ recs, LNode* Ln,
                                                uint32_t* cn_next, uint32_t* ck_next,
                                                SNode* small, SNode* med, Counters* ctr,
                                                NodeOut out, uint32_t max_depth,
                                                uint32_t parity_next, uint32_t learn_min,
                                                uint32_t med_max, uint32_t* lrows_next,
                                                uint32_t* lvn, uint32_t lanes_min) {
  const uint32_t nL = lvc[LV_NL];
  const SweepCfg cf{max_depth, med_max, learn_min};
  if (nL >= lanes_min) {
    const uint32_t ngroups = (nL + SW_WAVES * 64 - 1) / (SW_WAVES * 64);
    for (uint32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
      sweep_lanes(g, L, nL, hist, recs, Ln, cn_next, ck_next, small, med, ctr, out, cf,
                  parity_next, lrows_next, lvn);
      __syncthreads();
    }
  } else {
    const uint32_t ngroups = (nL + SW_WAVES - 1) / SW_WAVES;
    for (uint32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
      sweep_group(g, L, nL, hist, recs, Ln, cn_next, ck_next, small, med, ctr, out, cf,
                  parity_next, lrows_next, lvn);
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ partition
// Stable partition of one chunk: lt objects keep their order at the front of
// the node range, gt objects after n_lt.  Offsets of earlier chunks come from
// their per-bin counts.  Children's centroid bounds are accumulated here.
template <int NW>
__device__ inline void partition_chunk(uint32_t c, const LNode* __restrict__ L,
                                       const uint32_t* __restrict__ cnode,
                                       const uint32_t* __restrict__ ck,
                                       const uint32_t* __restrict__ blk,
                                       const Rec* __restrict__ src, Rec* __restrict__ dst,
                                       LNode* Ln, uint32_t* __restrict__ hist, uint32_t fuse) {
  __shared__ uint32_t s_red[NW];
  __shared__ uint32_t s_lt[NW], s_gt[NW];
  const uint32_t slot = cnode[c];
  const uint32_t k = ck[c];
  const LNode& nd = L[slot];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint32_t kind = nd.kind, axis = nd.axis, cut = nd.cut, n_lt = nd.n_lt;
  const uint32_t lo = nd.start + k * CH;
  const uint32_t hi = min(lo + CH, nd.start + nd.count);

  uint32_t lt_before;
  if (kind == 0 && k == 0) {  // block-uniform: the node's first chunk
    lt_before = 0;
  } else if (kind == 0) {
    uint32_t part = 0;
    const uint32_t tot = k * cut;
    for (uint32_t t = threadIdx.x; t < tot; t += blockDim.x) {
      uint32_t kk = t / cut, b = t % cut;
      part += blk[(size_t)(nd.chunk0 + kk) * 96 + axis * MG_NBINS + b];
    }
    part = wave_sum_u32_dpp(part);
    if (lane == 0) s_red[w] = part;
    __syncthreads();
    lt_before = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) lt_before += s_red[q];
  } else {
    lt_before = min(k * CH, n_lt);
  }
  const uint32_t gt_before = k * CH - lt_before;
  const float cmin = acc_min_dec(nd.cacc[axis]);
  const float cext = __fsub_rn(acc_max_dec(nd.cacc[3 + axis]), cmin);

  uint32_t acc_lt[6], acc_gt[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) { acc_lt[j] = 0; acc_gt[j] = 0; }
  uint32_t run_lt = 0, run_gt = 0;
  for (uint32_t base = lo; base < hi; base += blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < hi;
    Box b;
    uint32_t id = 0;
    bool pred = false;
    if (valid) {
      rec_load(&src[i], b, id);
      if (kind == 0) pred = bin_index(center(b.mn[axis], b.mx[axis]), cmin, cext) < cut;
      else pred = (i - nd.start) < n_lt;
    }
    const uint64_t ml = __ballot(valid && pred);
    const uint64_t mg = __ballot(valid && !pred);
    if (lane == 0) { s_lt[w] = (uint32_t)__popcll(ml); s_gt[w] = (uint32_t)__popcll(mg); }
    __syncthreads();
    uint32_t pl = 0, pg = 0, tl = 0, tg = 0;
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)NW; ++q) {
      if (q < w) { pl += s_lt[q]; pg += s_gt[q]; }
      tl += s_lt[q];
      tg += s_gt[q];
    }
    if (valid) {
      uint32_t p;
      if (pred) p = nd.start + lt_before + run_lt + pl + lanemask_lt_popc(ml);
      else p = nd.start + n_lt + gt_before + run_gt + pg + lanemask_lt_popc(mg);
      rec_store(&dst[p], b, id);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float cc = center(b.mn[a], b.mx[a]);
        const uint32_t emn = acc_min_enc(cc), emx = acc_max_enc(cc);
        if (pred) {
          acc_lt[a] = max(acc_lt[a], emn);
          acc_lt[3 + a] = max(acc_lt[3 + a], emx);
        } else {
          acc_gt[a] = max(acc_gt[a], emn);
          acc_gt[3 + a] = max(acc_gt[3 + a], emx);
        }
      }
    }
    run_lt += tl;
    run_gt += tg;
    __syncthreads();
  }
  // children centroid bounds
  __shared__ uint32_t red[NW][12];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    acc_lt[j] = wave_max_u32_dpp(acc_lt[j]);
    acc_gt[j] = wave_max_u32_dpp(acc_gt[j]);
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) { red[w][j] = acc_gt[j]; red[w][6 + j] = acc_lt[j]; }
  }
  __syncthreads();
  const bool single = nd.nchunk == 1;
  if (threadIdx.x < 12) {
    const int which = threadIdx.x < 6 ? 0 : 1;  // 0 = gt child, 1 = lt child
    const int32_t cs = nd.child_slot[which];
    uint32_t v = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) v = max(v, red[q][threadIdx.x]);
    if (cs >= 0) {
      if (single) Ln[cs].cacc[threadIdx.x % 6] = v;  // this block saw all of the child
      else if (v) atomicMax(&Ln[cs].cacc[threadIdx.x % 6], v);
    }
  }
  if (!single) {
    // a child spanning several chunks merges its histogram atomically: zero its row
    if (k == 0) {
      for (int which = 0; which < 2; ++which) {
        const int32_t cs = nd.child_slot[which];
        if (cs >= 0 && Ln[cs].nchunk > 1)
          for (uint32_t t = threadIdx.x; t < HIST_WORDS; t += blockDim.x)
            hist[(size_t)cs * HIST_WORDS + t] = 0;
      }
    }
    return;
  }
  // The whole node was in this chunk: bin each large child now, from the
  // records just written and the child's exact centroid bounds.
  __shared__ uint32_t h[HIST_WORDS];
  for (int which = 0; which < 2; ++which) {
    const int32_t cs = nd.child_slot[which];
    if (cs < 0) continue;  // block-uniform
    const uint32_t cstart = which == 0 ? nd.start + n_lt : nd.start;
    const uint32_t ccount = which == 0 ? nd.count - n_lt : n_lt;
    float cmin3[3], cext3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      uint32_t mnw = 0, mxw = 0;
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        mnw = max(mnw, red[q][6 * which + a]);
        mxw = max(mxw, red[q][6 * which + 3 + a]);
      }
      cmin3[a] = acc_min_dec(mnw);
      cext3[a] = __fsub_rn(acc_max_dec(mxw), cmin3[a]);
    }
    for (uint32_t t = threadIdx.x; t < HIST_WORDS; t += blockDim.x) h[t] = 0;
    __syncthreads();  // also orders this block's record stores before the reads below
    for (uint32_t base = cstart; base < cstart + ccount; base += blockDim.x) {
      const uint32_t i = base + threadIdx.x;
      const bool valid = i < cstart + ccount;
      Box b;
      uint32_t id = 0;
#pragma unroll
      for (int x = 0; x < 3; ++x) { b.mn[x] = 0.0f; b.mx[x] = 0.0f; }
      if (valid) rec_load(&dst[i], b, id);
      uint32_t e[6];
#pragma unroll
      for (int j = 0; j < 3; ++j) { e[j] = acc_min_enc(b.mn[j]); e[3 + j] = acc_max_enc(b.mx[j]); }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        // children of deep nodes spread over the bins: per-lane atomics
        if (valid) {
          const uint32_t bi = bin_index(center(b.mn[a], b.mx[a]), cmin3[a], cext3[a]);
          uint32_t* hb = &h[(a * MG_NBINS + bi) * BIN_WORDS];
          atomicAdd(&hb[0], 1u);
#pragma unroll
          for (int j = 0; j < 6; ++j) atomicMax(&hb[1 + j], e[j]);
        }
      }
    }
    __syncthreads();
    uint32_t* gh = hist + (size_t)cs * HIST_WORDS;
    if (fuse) {
      // the child's split straight from the histogram in LDS: its decision, not
      // the 2688-byte row, goes to memory, and the next level's sweep skips it
      if (w == 0) {
        Box cb;
#pragma unroll
        for (int a = 0; a < 3; ++a) { cb.mn[a] = Ln[cs].bmn[a]; cb.mx[a] = Ln[cs].bmx[a]; }
        const SplitDec d = split_from_hist(h, cb, cmin3, cext3, cstart, ccount, dst, LM_PLAIN);
        if (lane == 0) {
          *reinterpret_cast<SplitDec*>(gh) = d;
          Ln[cs].flags = LF_PREBINNED | LF_DECIDED;
        }
      }
    } else {
      for (uint32_t t = threadIdx.x; t < HIST_WORDS; t += blockDim.x) gh[t] = h[t];
    }
    __syncthreads();
  }
}

// NW wavefronts per block: 1 suits the many small nodes of the lower levels
// (no block barriers between waves), 4 the large chunks (the tuning picks).
template <int NW>
__global__ __launch_bounds__(NW * 64) void k_partition(const LNode* __restrict__ L,
                                                       const uint32_t* __restrict__ cnode,
                                                       const uint32_t* __restrict__ ck,
                                                       const uint32_t* __restrict__ blk,
                                                       const Rec* __restrict__ src,
                                                       Rec* __restrict__ dst, LNode* Ln,
                                                       uint32_t* __restrict__ hist,
                                                       const uint32_t* __restrict__ lvc,
                                                       uint32_t fuse) {
  const uint32_t nC = lvc[LV_NC];
  for (uint32_t c = blockIdx.x; c < nC; c += gridDim.x) {
    partition_chunk<NW>(c, L, cnode, ck, blk, src, dst, Ln, hist, fuse);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ medium nodes
// One wavefront per node of 33..med_max objects builds the part of its
// subtree above the <= 32-object subtrees, depth first, with the level path's
// own steps: bin the node's records into 3 x 32 bins in LDS from its exact
// centroid bounds, take the split from that histogram (split_from_hist, the
// same function the sweep and the partition use), write the node's records,
// and partition its records stably into the other record buffer at the same
// positions while the children's centroid bounds accumulate.  A child with
// > 32 objects below max_depth goes on an LDS stack, any other to the
// small-subtree list (appended in batches, one atomic per batch).  The node's
// records are read and written while it is worked on, so they stay in cache,
// where the level path streams every record of every level through HBM.
constexpr uint32_t MED_STACK = 68;  // depth first: one net push per level, max_depth <= 64
constexpr uint32_t MED_SBUF = 32;   // small children per list append
constexpr uint32_t MED_REG = 4;     // records per lane loaded together (and kept, <= 256 objects)

struct MedEntry {
  uint32_t start, count, idx, P;
  uint32_t depth, parity;
  float bmn[3], bmx[3];
  uint32_t cacc[6];  // centroid bounds: ~okey(min) x3, okey(max) x3
};

__device__ inline void med_flush(SNode* sbuf, uint32_t nsb, SNode* __restrict__ small,
                                 uint32_t small_cap, Counters* ctr, uint32_t& err) {
  const uint32_t lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(&ctr->n_small, nsb);
  base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
  if ((uint64_t)base + nsb > small_cap) { err |= ERR_INTERNAL; return; }
  // 64-byte entries, 16 words each: the wavefront copies them word by word
  const uint32_t* src = reinterpret_cast<const uint32_t*>(sbuf);
  uint32_t* dst = reinterpret_cast<uint32_t*>(small + base);
  for (uint32_t t = lane; t < nsb * 16; t += 64) dst[t] = src[t];
}

// A wavefront's next medium node: the first gridDim.x are the blocks' own, so
// a grid that covers the list takes no atomic at all.
__device__ inline uint32_t med_ticket(Counters* ctr, uint32_t lane, uint32_t nmed) {
  if (gridDim.x >= nmed) return nmed;
  uint32_t t = 0;
  if (lane == 0) t = gridDim.x + atomicAdd(&ctr->n_med_next, 1u);
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
}

__global__ __launch_bounds__(64) void k_medium(const SNode* __restrict__ med,
                                               const uint32_t* __restrict__ n_med_dev,
                                               Rec* __restrict__ rec0, Rec* __restrict__ rec1,
                                               SNode* __restrict__ small, uint32_t small_cap,
                                               Counters* ctr, NodeOut out, uint32_t max_depth,
                                               uint32_t nslots) {
  __shared__ uint32_t h[HIST_WORDS];
  __shared__ MedEntry stk[MED_STACK];
  __shared__ SNode sbuf[MED_SBUF];
  const uint32_t lane = lane_id();
  const uint32_t nmed = *n_med_dev;
  uint32_t nsb = 0, err = 0;  // nsb wave-uniform; err per lane
  unsigned long long visits = 0;
  uint32_t splits = 0, medians = 0;
  // medium nodes range from 33 to med_max objects: block b takes node b, and
  // every later one comes from a queue (one atomic per node, none at the
  // start, where every wavefront would wait on the same address), so that the
  // wavefronts finish together instead of each working through a fixed share
  for (uint32_t m = blockIdx.x; m < nmed; m = med_ticket(ctr, lane, nmed)) {
    const SNode root = med[m];
    const uint32_t nbase = root.nbase, obase = root.obase;
    {  // the root's centroid bounds
      const Rec* X = root.parity ? rec1 : rec0;
      uint32_t acc[6] = {0u, 0u, 0u, 0u, 0u, 0u};
      for (uint32_t j = lane; j < root.count; j += 64) {
        Box b;
        uint32_t id;
        rec_load(&X[root.start + j], b, id);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float cc = center(b.mn[a], b.mx[a]);
          acc[a] = max(acc[a], acc_min_enc(cc));
          acc[3 + a] = max(acc[3 + a], acc_max_enc(cc));
        }
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k] = wave_max_u32_dpp(acc[k]);
      if (lane == 0) {
        MedEntry& e = stk[0];
        e.start = root.start; e.count = root.count; e.idx = root.idx; e.P = root.P;
        e.depth = root.depth; e.parity = root.parity;
#pragma unroll
        for (int a = 0; a < 3; ++a) { e.bmn[a] = root.bmn[a]; e.bmx[a] = root.bmx[a]; }
#pragma unroll
        for (int k = 0; k < 6; ++k) e.cacc[k] = acc[k];
      }
    }
    uint32_t sp = 1;
    __syncthreads();
    while (sp > 0) {
      --sp;
      const MedEntry e = stk[sp];
      const Rec* X = e.parity ? rec1 : rec0;
      Rec* Y = e.parity ? rec0 : rec1;
      Box nb;
      float cmin[3], cext[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        nb.mn[a] = e.bmn[a];
        nb.mx[a] = e.bmx[a];
        cmin[a] = acc_min_dec(e.cacc[a]);
        cext[a] = __fsub_rn(acc_max_dec(e.cacc[3 + a]), cmin[a]);
      }
      // bin (the partition's prebinning).  Each lane issues the loads of up
      // to MED_REG records at once (object j = g0 + 64 k + lane); a sub-node
      // of <= 64 * MED_REG objects keeps them in registers for the partition,
      // so it reads its records from memory once.
      for (uint32_t t = lane; t < HIST_WORDS; t += 64) h[t] = 0;
      __syncthreads();
      const bool cached = e.count <= 64 * MED_REG;
      Box rb[MED_REG];
      uint32_t rid[MED_REG];
      for (uint32_t g0 = 0; g0 < e.count; g0 += 64 * MED_REG) {
#pragma unroll
        for (uint32_t k = 0; k < MED_REG; ++k) {
          const uint32_t j = g0 + 64 * k + lane;
          if (j < e.count) rec_load(&X[e.start + j], rb[k], rid[k]);
        }
#pragma unroll
        for (uint32_t k = 0; k < MED_REG; ++k) {
          const uint32_t j = g0 + 64 * k + lane;
          if (j >= e.count) continue;
          uint32_t en[6];
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            en[q] = acc_min_enc(rb[k].mn[q]);
            en[3 + q] = acc_max_enc(rb[k].mx[q]);
          }
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const uint32_t bi = bin_index(center(rb[k].mn[a], rb[k].mx[a]), cmin[a], cext[a]);
            uint32_t* hb = &h[(a * MG_NBINS + bi) * BIN_WORDS];
            atomicAdd(&hb[0], 1u);
#pragma unroll
            for (int q = 0; q < 6; ++q) atomicMax(&hb[1 + q], en[q]);
          }
        }
      }
      __syncthreads();
      const SplitDec d = split_from_hist(h, nb, cmin, cext, e.start, e.count, X, LM_PLAIN);
      // never happens: a split leaves both sides non-empty, and the node and
      // its children's slots lie in the mesh's node range
      if (d.nl == 0 || d.nl >= e.count || e.idx >= nslots ||
          (uint64_t)e.P + 2 * (uint64_t)e.count - 2 > nslots) {
        err |= ERR_INTERNAL;
        __syncthreads();
        continue;
      }
      if (lane == 0) {
        out.write_node(e.idx, nb, e.start - obase, e.P - nbase, e.count);
        out.write_split(e.idx, d.axis | (d.kind << 8), d.value, d.nl, fbits(d.cost));
      }
      visits += e.count;
      ++splits;
      medians += d.kind == 2 ? 1u : 0u;
      // stable partition X -> Y: lt objects first, then gt (partition_chunk's rule)
      // the split axis' values by selects (a run-time index would put arrays in scratch)
      const uint32_t ax = d.axis;
      const float pcmin = ax == 0 ? cmin[0] : (ax == 1 ? cmin[1] : cmin[2]);
      const float pcext = ax == 0 ? cext[0] : (ax == 1 ? cext[1] : cext[2]);
      uint32_t acc_lt[6] = {0u, 0u, 0u, 0u, 0u, 0u}, acc_gt[6] = {0u, 0u, 0u, 0u, 0u, 0u};
      uint32_t run_lt = 0, run_gt = 0;
      // objects in list order: group g0, then k, then lane (as they were binned)
      for (uint32_t g0 = 0; g0 < e.count; g0 += 64 * MED_REG) {
        if (!cached) {
#pragma unroll
          for (uint32_t k = 0; k < MED_REG; ++k) {
            const uint32_t j = g0 + 64 * k + lane;
            if (j < e.count) rec_load(&X[e.start + j], rb[k], rid[k]);
          }
        }
#pragma unroll
        for (uint32_t k = 0; k < MED_REG; ++k) {
          const uint32_t base = g0 + 64 * k;
          if (base >= e.count) break;
          const uint32_t i = base + lane;
          const bool valid = i < e.count;
          Box b;
          uint32_t id = 0;
          bool pred = false;
          if (valid) {
            b = rb[k];
            id = rid[k];
            const float lo = ax == 0 ? b.mn[0] : (ax == 1 ? b.mn[1] : b.mn[2]);
            const float hi = ax == 0 ? b.mx[0] : (ax == 1 ? b.mx[1] : b.mx[2]);
            pred = d.kind == 0 ? bin_index(center(lo, hi), pcmin, pcext) < d.value : i < d.nl;
          }
          const uint64_t ml = __ballot(valid && pred);
          const uint64_t mg = __ballot(valid && !pred);
          if (valid) {
            const uint32_t pos = pred ? e.start + run_lt + lanemask_lt_popc(ml)
                                      : e.start + d.nl + run_gt + lanemask_lt_popc(mg);
            // the predicate and the histogram count the same objects, so every
            // position lies in the node's range; a store outside it is refused
            const bool in_range = pred ? pos < e.start + d.nl : pos < e.start + e.count;
            if (in_range) rec_store(&Y[pos], b, id);
            else err |= ERR_INTERNAL;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const float cc = center(b.mn[a], b.mx[a]);
              const uint32_t emn = acc_min_enc(cc), emx = acc_max_enc(cc);
              if (pred) {
                acc_lt[a] = max(acc_lt[a], emn);
                acc_lt[3 + a] = max(acc_lt[3 + a], emx);
              } else {
                acc_gt[a] = max(acc_gt[a], emn);
                acc_gt[3 + a] = max(acc_gt[3 + a], emx);
              }
            }
          }
          run_lt += (uint32_t)__popcll(ml);
          run_gt += (uint32_t)__popcll(mg);
        }
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        acc_lt[k] = wave_max_u32_dpp(acc_lt[k]);
        acc_gt[k] = wave_max_u32_dpp(acc_gt[k]);
      }
      // the stores above come before any lane reads Y as a child's X
      __syncthreads();
      // children (lib.rs:648-663): gt at P, lt at P + 1
      const uint32_t cd = e.depth + 1;
#pragma unroll
      for (int which = 0; which < 2; ++which) {
        const bool gt = which == 0;
        const uint32_t cstart = gt ? e.start + d.nl : e.start;
        const uint32_t ccount = gt ? d.ng : d.nl;
        const uint32_t cidx = gt ? e.P : e.P + 1;
        const uint32_t cP = gt ? e.P + 2 : e.P + 2 + 2 * (d.ng - 1);
        Box cb;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          cb.mn[a] = gt ? d.gt.mn[a] : d.lt.mn[a];
          cb.mx[a] = gt ? d.gt.mx[a] : d.lt.mx[a];
        }
        if (ccount > MG_SMALL && cd < max_depth) {
          if (sp >= MED_STACK) { err |= ERR_INTERNAL; continue; }
          if (lane == 0) {
            MedEntry& c = stk[sp];
            c.start = cstart; c.count = ccount; c.idx = cidx; c.P = cP;
            c.depth = cd; c.parity = e.parity ^ 1u;
#pragma unroll
            for (int a = 0; a < 3; ++a) { c.bmn[a] = cb.mn[a]; c.bmx[a] = cb.mx[a]; }
#pragma unroll
            for (int k = 0; k < 6; ++k) c.cacc[k] = gt ? acc_gt[k] : acc_lt[k];
          }
          ++sp;
        } else {
          if (lane == 0) {
            SNode& sn = sbuf[nsb];
            sn = SNode{};
            sn.start = cstart; sn.count = ccount; sn.idx = cidx; sn.P = cP;
            sn.depth = cd; sn.parity = e.parity ^ 1u; sn.nbase = nbase; sn.obase = obase;
#pragma unroll
            for (int a = 0; a < 3; ++a) { sn.bmn[a] = cb.mn[a]; sn.bmx[a] = cb.mx[a]; }
          }
          ++nsb;
          if (nsb == MED_SBUF) {
            __syncthreads();
            med_flush(sbuf, nsb, small, small_cap, ctr, err);
            nsb = 0;
          }
        }
      }
      __syncthreads();
    }
  }
  if (nsb) {
    __syncthreads();
    med_flush(sbuf, nsb, small, small_cap, ctr, err);
  }
  const bool failed = __ballot(err != 0) != 0;  // a lane's refused store counts too
  if (lane == 0) {
    if (visits) atomicAdd(&ctr->sum_large, visits);
    if (splits) atomicAdd(&ctr->n_med_split, splits);
    if (medians) atomicAdd(&ctr->n_median, medians);
    if (failed) atomicOr(&ctr->err, ERR_INTERNAL);
  }
}

// ------------------------------------------------------------------ small subtrees
// One wavefront builds two subtrees of <= 32 objects (lanes 0-31 and 32-63),
// level by level.  Every node of a level is a segment of consecutive lanes
// (lane = position in the node's list), so all nodes of a level are split at
// once: segmented scans give centroid bounds and the first-minimum split, each
// lane evaluates the 3 candidates (axis a, threshold r) of its list position
// (adaptive sweep, lib.rs:391-442), and a stable in-segment partition moves the
// objects for the next level.
//
// Per level the fixed work is kept small, since a subtree pair takes about
// five levels and the candidate loop is the only part that grows with the
// node size:
//   * a node's record is written when its parent splits (everything in it is
//     known then: box, start, count, pair base), so the lane that found the
//     split writes both children's records as 64 contiguous bytes; only the
//     subtree root is written by the root's first lane;
//   * the other lanes of a node get from the winner only the child count, the
//     threshold and the two children's surface areas (the next level's
//     parent SA, the same bits box_sa gives), not the boxes;
//   * object data lives in LDS only (each lane keeps its centroid in
//     registers); the final order is written once, after the last level, by
//     all lanes at consecutive positions.
// One wavefront's objects in LDS, one array per field group, so lanes that
// touch different positions hit different banks, followed by the staged node
// records of its two subtrees.  The candidate loop reads positions up to
// 63 + 32 without a bound test: reads past an array's 64 entries land in the
// next array of the same struct and are never folded in (they belong to no
// node), so the loop addresses the arrays through byte offsets from the
// struct's base.
struct SmallObjs {
  float4 a[64];      // c0 c1 c2 mn0
  float4 b[64];      // mn1 mn2 mx0 mx1
  float m2[64];      // mx2
  uint32_t id[64];   // lane the object was loaded by (its index and triangle stay there)
  // per half: records of the subtree's descendants (slot - pair base of the
  // root, at most 62) and of the root itself (entry 63); written to memory
  // once, after the subtree's last level
  uint4 rec[2][64][2];
  uint32_t len[2][64];
  // per half: constants of its subtree, read by the lanes that write records,
  // and counts kept in LDS so that the level loop issues no memory operation
  uint32_t hdr[2][16];
  uint32_t gap[2][16][2];  // depth-cut leaves: pair base, unused slots (at most 16)
  // the subtree pair's output by final position: index inside the mesh, triangle
  uint32_t operm[64];
  uint32_t otri[64][3];
};
// SmallObjs::hdr fields
constexpr uint32_t H_BASE = 0, H_NBASE = 1, H_OBASE = 2, H_P0 = 3, H_ROOT = 4, H_NREC = 5,
                   H_OWN = 6, H_NMED = 7, H_ERR = 8, H_NGAP = 9;
constexpr uint32_t SO_A = 0, SO_B = 64 * 16, SO_M2 = 128 * 16;
static_assert(offsetof(SmallObjs, b) == SO_B && offsetof(SmallObjs, m2) == SO_M2,
              "SmallObjs field offsets");
static_assert(SO_M2 + 96 * 4 <= sizeof(SmallObjs), "loop reads stay inside SmallObjs");
constexpr uint32_t ROOT_REC = 63;

__device__ inline void obj_put(SmallObjs* S, uint32_t pos, const Box& bx, const float (&c)[3],
                               uint32_t id) {
  S->a[pos] = make_float4(c[0], c[1], c[2], bx.mn[0]);
  S->b[pos] = make_float4(bx.mn[1], bx.mn[2], bx.mx[0], bx.mx[1]);
  S->m2[pos] = bx.mx[2];
  S->id[pos] = id;
}
// The box of the object at `pos`.
__device__ inline Box obj_box(const SmallObjs* S, uint32_t pos) {
  const float4 a = S->a[pos], b = S->b[pos];
  Box r;
  r.mn[0] = a.w; r.mn[1] = b.x; r.mn[2] = b.y;
  r.mx[0] = b.z; r.mx[1] = b.w; r.mx[2] = S->m2[pos];
  return r;
}
// The centroid and box of the object at `pos` <= 95 (what the candidate loop
// reads; positions past 63 give values that are never used).
struct SmallGeo {
  float4 a, b;
  float mx2;
};
__device__ inline SmallGeo geo_get(const SmallObjs* S, uint32_t pos) {
  const char* base = reinterpret_cast<const char*>(S);
  SmallGeo g;
  g.a = *reinterpret_cast<const float4*>(base + SO_A + 16 * pos);
  g.b = *reinterpret_cast<const float4*>(base + SO_B + 16 * pos);
  g.mx2 = *reinterpret_cast<const float*>(base + SO_M2 + 4 * pos);
  return g;
}
// Moves each lane's object to `newpos` (a permutation of the lanes) and
// returns the centroid of the object that arrives at the lane's position.
__device__ inline void lane_move(SmallObjs* S, uint32_t lane, uint32_t newpos, float (&c)[3]) {
  const float4 a = S->a[lane], b = S->b[lane];
  const float m2 = S->m2[lane];
  const uint32_t id = S->id[lane];
  wave_sync();
  S->a[newpos] = a;
  S->b[newpos] = b;
  S->m2[newpos] = m2;
  S->id[newpos] = id;
  wave_sync();
  const float4 n = S->a[lane];
  c[0] = n.x; c[1] = n.y; c[2] = n.z;
}
// A node record into the half's LDS staging area.
__device__ inline void stage_node(SmallObjs* S, uint32_t h, uint32_t e, const Box& b,
                                  uint32_t start, uint32_t child, uint32_t len) {
  S->rec[h][e][0] = make_uint4(fbits(b.mn[0]), fbits(b.mn[1]), fbits(b.mn[2]), start);
  S->rec[h][e][1] = make_uint4(fbits(b.mx[0]), fbits(b.mx[1]), fbits(b.mx[2]), child);
  S->len[h][e] = len;
}

// A depth-cut leaf of n > 1 objects at pair base P leaves 2(n - 1) slots
// unused; k_compact squeezes them out afterwards.
__device__ inline void put_gap(Counters* ctr, uint32_t* gaps, uint32_t P, uint32_t n) {
  const uint32_t g = atomicAdd(&ctr->n_gaps, 1u);
  gaps[2 * g] = P;
  gaps[2 * g + 1] = 2 * (n - 1);
  atomicAdd(&ctr->n_cutoff, 1u);
}

// Writes the final object order: perm (index inside the mesh) and, for
// triangles, the permuted triangle list.
__device__ inline void put_perm(uint32_t pos, uint32_t id, uint32_t obase, uint32_t* perm,
                                const uint32_t* tris_in, uint32_t* tris_out) {
  perm[pos] = id - obase;
  if (tris_out) {
    const uint32_t a = tris_in[3 * (size_t)id], b = tris_in[3 * (size_t)id + 1],
                   c = tris_in[3 * (size_t)id + 2];
    tris_out[3 * (size_t)pos] = a;
    tris_out[3 * (size_t)pos + 1] = b;
    tris_out[3 * (size_t)pos + 2] = c;
  }
}

// A depth-cut leaf of the half's subtree, listed in LDS (a subtree of <= 32
// objects has at most 16 leaves of two or more objects).
__device__ inline void stage_gap(SmallObjs* S, uint32_t h, uint32_t P, uint32_t n) {
  const uint32_t k = atomicAdd(&S->hdr[h][H_NGAP], 1u);
  S->gap[h][k][0] = P;
  S->gap[h][k][1] = 2 * (n - 1);
}

// The record of a node created by a split at depth d - 1, staged at entry e:
// a leaf when it holds one object or d == max_depth (lib.rs:548-558), else an
// internal node whose children will sit at pair base P.
__device__ inline void put_child(SmallObjs* S, uint32_t h, uint32_t e, const Box& b,
                                 uint32_t start, uint32_t n, uint32_t P, uint32_t nbase,
                                 uint32_t d, uint32_t max_depth) {
  const bool inner = n > 1 && d < max_depth;
  stage_node(S, h, e, b, start, inner ? P - nbase : 0u, n);
  if (n > 1 && d >= max_depth) stage_gap(S, h, P, n);
}

// Writes out the previous subtree pair staged in LDS (records, lengths, split
// records, depth-cut gaps, counts, permutation and permuted triangles) and
// marks the staging area empty.  Called with every lane of the wavefront.
__device__ inline void small_flush(SmallObjs* S, uint4* __restrict__ sp, const NodeOut& out,
                                   uint32_t* __restrict__ perm, uint32_t* __restrict__ tris_out,
                                   Counters* ctr, uint32_t* gaps) {
  const uint32_t lane = lane_id();
  const uint32_t h = lane >> 5, hl = lane & 31;
  const uint32_t* hd = S->hdr[h];
  const uint32_t own = hd[H_OWN];
  if (own) {
    const uint32_t P0 = hd[H_P0], root = hd[H_ROOT], nrec = hd[H_NREC];
    // descendants fill slots [P0, P0 + nrec): 16-byte stores at consecutive
    // addresses; the root its own slot
    uint4* dst = reinterpret_cast<uint4*>(out.nodes + P0);
    const uint4* src = &S->rec[h][0][0];
    for (uint32_t j = hl; j < 2 * nrec; j += 32) dst[j] = src[j];
    if (hl < 2) reinterpret_cast<uint4*>(out.nodes + root)[hl] = S->rec[h][ROOT_REC][hl];
    if (out.lens) {
      for (uint32_t j = hl; j < nrec; j += 32) out.lens[P0 + j] = S->len[h][j];
      if (hl == 0) out.lens[root] = S->len[h][ROOT_REC];
    }
    if (sp) {  // split records of the internal nodes (entries without one hold ~0)
      const uint4* ss = sp + 64 * h;
      for (uint32_t j = hl; j < nrec; j += 32) {
        const uint4 v = ss[j];
        if (v.x != ~0u) reinterpret_cast<uint4*>(out.splits)[P0 + j] = v;
      }
      if (hl == 0 && ss[ROOT_REC].x != ~0u)
        reinterpret_cast<uint4*>(out.splits)[root] = ss[ROOT_REC];
    }
    const uint32_t ng = hd[H_NGAP];
    if (ng) {
      uint32_t g = 0;
      if (hl == 0) {
        g = atomicAdd(&ctr->n_gaps, ng);
        atomicAdd(&ctr->n_cutoff, ng);
      }
      g = (uint32_t)__shfl((int)g, (int)(lane & 32u), 64);
      if (hl < ng) {
        gaps[2 * (g + hl)] = S->gap[h][hl][0];
        gaps[2 * (g + hl) + 1] = S->gap[h][hl][1];
      }
    }
    if (hl == 0 && hd[H_NMED]) atomicAdd(&ctr->n_median, hd[H_NMED]);
    if (hl == 0 && hd[H_ERR]) atomicOr(&ctr->err, ERR_INTERNAL);
    if (hl < own) {
      const uint32_t pos = hd[H_BASE] + hl;
      perm[pos] = S->operm[lane];
      if (tris_out) {
        tris_out[3 * (size_t)pos] = S->otri[lane][0];
        tris_out[3 * (size_t)pos + 1] = S->otri[lane][1];
        tris_out[3 * (size_t)pos + 2] = S->otri[lane][2];
      }
    }
  }
  wave_sync();
  if (hl == 0) S->hdr[h][H_OWN] = 0;
  wave_sync();
}

// Subtrees ea (lanes 0-31) and eb (lanes 32-63); an index >= nS is none.
// Order of memory operations: this pair's loads first, then the previous
// pair's stores (small_flush), then the level loop, which issues none.  Vector
// memory operations of a wavefront retire in order, so a wait for a load also
// waits for every store issued before it; stores issued right before the level
// loop have long retired when the next pair waits for its loads.  sp: split
// record staging of the wavefront (null when no split records are written).
__device__ inline void small_pair(uint32_t ea, uint32_t eb, SmallObjs* S, uint4* __restrict__ sp,
                                  const SNode* __restrict__ small,
                                  uint32_t nS, const Rec* __restrict__ rec0,
                                  const Rec* __restrict__ rec1, NodeOut out,
                                  uint32_t* __restrict__ perm, const uint32_t* __restrict__ tris_in,
                                  uint32_t* __restrict__ tris_out, Counters* ctr, uint32_t* gaps,
                                  uint32_t max_depth) {
  const uint32_t lane = lane_id();
  const uint32_t h = lane >> 5, hl = lane & 31, off = 32 * h;
  const uint32_t e = h ? eb : ea;

  // has: the lane's object sits in a node that is still to be split
  bool has = false;
  float c[3] = {0.0f, 0.0f, 0.0f};
  uint32_t ls = lane, lc = 1, idx = 0, P = 0, depth = 0;
  float sa_p = 0.0f;
  // the lane's object: original index and triangle, loaded with its record so
  // the gather's latency overlaps the record load; written out from this lane
  uint32_t oid = 0, tri0 = 0, tri1 = 0, tri2 = 0;
  // the half's subtree constants (written to LDS after the previous flush)
  uint32_t hv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool stage_root = false, forced = false;
  Box nb;
  {
    // Straight-line loads with clamped addresses (every lane loads something
    // valid), so that no branch makes the compiler wait for the triangle
    // gather before the level loop: the gather is first used at the end.
    const SNode sn = small[e < nS ? e : nS - 1];
    const bool mine = e < nS && sn.count <= MG_SMALL && hl < sn.count;
    const Rec* src = sn.parity ? rec1 : rec0;
    const uint32_t j = min(hl, sn.count - 1);
    Box b;
    rec_load(&src[sn.start + j], b, oid);
    // boxes (TLAS) have no triangles: read the record again instead
    const uint32_t* tsrc = tris_out ? tris_in + 3 * (size_t)oid
                                    : reinterpret_cast<const uint32_t*>(&src[sn.start + j]);
    const u3_a4 t = *reinterpret_cast<const u3_a4*>(tsrc);  // one 12-byte load
    tri0 = t.x;
    tri1 = t.y;
    tri2 = t.z;
    // everything but that load has arrived (s_waitcnt vmcnt(1), other counters
    // untouched): the record and node fields used below need no further wait,
    // and the gather stays in flight across the flush and the level loop
    __builtin_amdgcn_s_waitcnt(0x0F71);
    forced = e < nS && sn.count > MG_SMALL;  // written after the level loop
#pragma unroll
    for (int x = 0; x < 3; ++x) { nb.mn[x] = sn.bmn[x]; nb.mx[x] = sn.bmx[x]; }
    if (e < nS && sn.count <= MG_SMALL) {
      stage_root = true;
      hv[H_BASE] = sn.start;
      hv[H_NBASE] = sn.nbase;
      hv[H_OBASE] = sn.obase;
      hv[H_P0] = sn.P;
      hv[H_ROOT] = sn.idx;
      // descendant records (none below a depth-cut root)
      hv[H_NREC] = sn.count > 1 && sn.depth < max_depth ? 2 * (sn.count - 1) : 0u;
      hv[H_OWN] = sn.count;
      hv[7] = sn.depth;
    }
    if (mine) {
      ls = off; lc = sn.count; idx = sn.idx; P = sn.P; depth = sn.depth;
      has = lc > 1 && depth < max_depth;
      sa_p = box_sa(nb);
    } else {  // lanes without an object store zeros (never read as objects)
#pragma unroll
      for (int x = 0; x < 3; ++x) { b.mn[x] = 0.0f; b.mx[x] = 0.0f; }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) c[x] = center(b.mn[x], b.mx[x]);
    obj_put(S, lane, b, c, lane);
  }
  wave_sync();
  small_flush(S, sp, out, perm, tris_out, ctr, gaps);
  if (hl == 0 && stage_root) {  // the subtree root's own record and constants
    const uint32_t count = hv[H_OWN], P0 = hv[H_P0], rdepth = hv[7];
    const bool inner = count > 1 && rdepth < max_depth;
    uint32_t* hd = S->hdr[h];
#pragma unroll
    for (int k = 0; k < 7; ++k) hd[k] = hv[k];
    hd[H_NMED] = 0;
    hd[H_ERR] = 0;
    hd[H_NGAP] = 0;
    stage_node(S, h, ROOT_REC, nb, hv[H_BASE] - hv[H_OBASE], inner ? P0 - hv[H_NBASE] : 0u,
               count);
    if (count > 1 && rdepth >= max_depth) stage_gap(S, h, P0, count);
  }
  if (sp) {
    for (uint32_t j = hl; j < 64; j += 32) sp[64 * h + j] = make_uint4(~0u, ~0u, ~0u, ~0u);
  }
  wave_sync();

  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (;;) {
    const bool act = has;
    if (!__ballot(act)) break;
    const uint32_t last = ls + lc - 1;
    const uint32_t r = lane - ls;
    // thresholds: the centroids in list order below 8 objects, else S = len
    // uniform steps over the centroid bounds (lib.rs:417-441)
    float th[3] = {c[0], c[1], c[2]};
    if (__ballot(act && lc >= 8)) {
      const Box cb = seg_box(act, c, c, lane, ls, last);  // lib.rs:499-509
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (lc >= 8) th[a] = adaptive_threshold(cb.mn[a], cb.mx[a], lc, r);
    }
    // the three candidates (a, r) of this lane over the node's objects; the
    // first object initialises both sides
    Box lt[3], gt[3];
    uint32_t nl[3];
    const uint32_t p0 = act ? ls : lane;
    {
      const SmallGeo g = geo_get(S, p0);
      const float qc[3] = {g.a.x, g.a.y, g.a.z};
      const float qmn[3] = {g.a.w, g.b.x, g.b.y}, qmx[3] = {g.b.z, g.b.w, g.mx2};
      const float pinf = __uint_as_float(0x7f800000u), ninf = __uint_as_float(0xff800000u);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const bool in = qc[a] < th[a];
        nl[a] = in ? 1u : 0u;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          lt[a].mn[x] = in ? qmn[x] : pinf;
          lt[a].mx[x] = in ? qmx[x] : ninf;
          gt[a].mn[x] = in ? pinf : qmn[x];
          gt[a].mx[x] = in ? ninf : qmx[x];
        }
      }
    }
    const uint32_t maxlc = wave_max6(act ? lc : 0u);
    const auto fold = [&](const SmallGeo& g) {
      const float qc[3] = {g.a.x, g.a.y, g.a.z};
      Box qb;
      qb.mn[0] = g.a.w; qb.mn[1] = g.b.x; qb.mn[2] = g.b.y;
      qb.mx[0] = g.b.z; qb.mx[1] = g.b.w; qb.mx[2] = g.mx2;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (qc[a] < th[a]) { box_grow_nz(lt[a], qb); ++nl[a]; }
        else box_grow_nz(gt[a], qb);
      }
    };
    // two objects in flight: the next one's LDS reads are issued before the
    // current one is folded in; p0 + k <= 63 + 32 stays inside the arrays
    SmallGeo g1 = geo_get(S, p0 + 1);
    for (uint32_t k = 1; k < maxlc; k += 2) {
      const SmallGeo g0 = geo_get(S, p0 + k + 1);
      if (act && k < lc) fold(g1);
      if (k + 1 >= maxlc) break;
      g1 = geo_get(S, p0 + k + 2);
      if (act && k + 1 < lc) fold(g0);
    }
    // costs (lib.rs:278-288) and the lane's best candidate
    uint64_t key = ~0ull;
    int ba = 0;
    float bsl = 0.0f, bsg = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const uint32_t ng = lc - nl[a];
      if (act && nl[a] >= 1 && ng >= 1) {
        // both sides hold objects: their boxes are not empty
        const float sl = box_sa_ne(lt[a]), sg = box_sa_ne(gt[a]);
        const float cst = split_cost(sa_p, sl, sg, nl[a], ng);
        const uint64_t k2 = ((uint64_t)tc_ukey(cst) << 32) | (uint32_t)(a * 32 + r);
        if (k2 < key) { key = k2; ba = a; bsl = sl; bsg = sg; }
      }
    }
    // first minimum over (cost, axis, candidate) of the node (lib.rs:515-524)
    const uint64_t skey = seg_min_u64(act ? key : ~0ull, lane, ls, last);
    const bool med = act && skey == ~0ull;
    const uint32_t q32 = (uint32_t)skey;
    const uint32_t wl = (act && !med) ? ls + (q32 & 31u) : lane;
    uint32_t axis = (act && !med) ? (q32 >> 5) : 0;
    const uint32_t mynl = ba == 0 ? nl[0] : (ba == 1 ? nl[1] : nl[2]);
    const float myth = ba == 0 ? th[0] : (ba == 1 ? th[1] : th[2]);
    uint32_t n_lt = __shfl(mynl, (int)wl, 64);
    const float tw = __shfl(myth, (int)wl, 64);
    float sa_lt = __shfl(bsl, (int)wl, 64), sa_gt = __shfl(bsg, (int)wl, 64);
    uint32_t kind = 1, value = fbits(tw);
    float cost = tc_ukey_inv((uint32_t)(skey >> 32));
    bool pred = act && (axis == 0 ? c[0] : (axis == 1 ? c[1] : c[2])) < tw;
    // the lane that writes the children's records, and their boxes
    bool writer = act && !med && lane == wl;
    Box wlt, wgt;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      wlt.mn[x] = ba == 0 ? lt[0].mn[x] : (ba == 1 ? lt[1].mn[x] : lt[2].mn[x]);
      wlt.mx[x] = ba == 0 ? lt[0].mx[x] : (ba == 1 ? lt[1].mx[x] : lt[2].mx[x]);
      wgt.mn[x] = ba == 0 ? gt[0].mn[x] : (ba == 1 ? gt[1].mn[x] : gt[2].mn[x]);
      wgt.mx[x] = ba == 0 ? gt[0].mx[x] : (ba == 1 ? gt[1].mx[x] : gt[2].mx[x]);
    }

    if (__ballot(med)) {
      // median fallback on axes 0,1,2, each reordering the list (lib.rs:444-491,
      // 527-531); the children take the winning axis' boxes and the axis-2 order.
      const uint32_t m = lc / 2;
      const uint64_t segm = (lc >= 64 ? ~0ull : ((1ull << lc) - 1)) << ls;
      bool have = false;
      uint32_t bk = 0, maxis = 0;
      Box mlt = box_empty(), mgt = box_empty();
      float mcost = 0.0f;
      for (uint32_t a = 0; a < 3; ++a) {
        const float ca = a == 0 ? c[0] : (a == 1 ? c[1] : c[2]);
        const uint32_t kme = tc_ukey(ca);
        uint32_t rank = 0;
        for (uint32_t k = 0; k < maxlc; ++k) {
          const uint32_t kk = __shfl(kme, (int)(med && k < lc ? ls + k : lane), 64);
          if (med && k < lc) rank += (kk < kme) || (kk == kme && k < r);
        }
        const bool sel = med && rank < m;
        const uint64_t sm = __ballot(sel);
        const uint32_t np = !med ? lane
                                 : (sel ? ls + (uint32_t)__popcll(sm & segm & below)
                                        : ls + m + (uint32_t)__popcll(~sm & segm & below));
        lane_move(S, lane, np, c);
        const Box ob = obj_box(S, lane);
        const bool inl = med && r < m;
        const Box l2 = seg_box(inl, ob.mn, ob.mx, lane, ls, last);
        const Box g2 = seg_box(med && !inl, ob.mn, ob.mx, lane, ls, last);
        const float c2 = split_cost(sa_p, box_sa(l2), box_sa(g2), m, lc - m);
        const uint32_t k2 = tc_ukey(c2);
        if (!have || k2 < bk) { have = true; bk = k2; maxis = a; mlt = l2; mgt = g2; mcost = c2; }
      }
      if (med) {
        axis = maxis; kind = 2; value = m; n_lt = m; cost = mcost;
        wlt = mlt; wgt = mgt;
        sa_lt = box_sa(mlt);
        sa_gt = box_sa(mgt);
        pred = r < m;
        writer = lane == ls;
      }
    }
    // a split always leaves both children non-empty; anything else is a bug and
    // stops this subtree rather than writing outside its slots
    const bool bad = act && (n_lt == 0 || n_lt >= lc);
    if (__ballot(bad)) {
      if (bad) atomicOr(&S->hdr[h][H_ERR], 1u);
      break;
    }

    // stable partition inside each node (partition_in_place, lib.rs:598-600)
    {
      const uint64_t segm = (lc >= 64 ? ~0ull : ((1ull << lc) - 1)) << ls;
      const uint64_t pm = __ballot(pred);
      uint32_t np = lane;
      if (act && !med)
        np = pred ? ls + (uint32_t)__popcll(pm & segm & below)
                  : ls + n_lt + (uint32_t)__popcll(~pm & segm & below);
      if (__ballot(act && !med)) lane_move(S, lane, np, c);
    }
    const uint32_t ngt = lc - n_lt;
    if (writer) {
      // children (lib.rs:648-663): gt at P, lt at P + 1, 64 contiguous bytes
      const uint32_t* hd = S->hdr[h];
      const uint32_t P0 = hd[H_P0], nbase = hd[H_NBASE];
      const uint32_t start = hd[H_BASE] + (ls - off) - hd[H_OBASE];
      put_child(S, h, P - P0, wgt, start + n_lt, ngt, P + 2, nbase, depth + 1, max_depth);
      put_child(S, h, P + 1 - P0, wlt, start, n_lt, P + 2 + 2 * (ngt - 1), nbase, depth + 1,
                max_depth);
      if (sp)
        sp[64 * h + (idx == hd[H_ROOT] ? ROOT_REC : idx - P0)] =
            make_uint4(axis | (kind << 8), value, n_lt, fbits(cost));
      if (kind == 2) atomicAdd(&S->hdr[h][H_NMED], 1u);
    }
    if (act) {
      if (lane < ls + n_lt) {
        lc = n_lt;
        idx = P + 1;
        P = P + 2 + 2 * (ngt - 1);
        sa_p = sa_lt;
      } else {
        ls = ls + n_lt;
        lc = ngt;
        idx = P;
        P = P + 2;
        sa_p = sa_gt;
      }
      ++depth;
      has = lc > 1 && depth < max_depth;
    }
  }
  // every object's final position is the lane it ended in; the lane that
  // loaded it learns that position through LDS and stages its output there
  uint32_t* fin = reinterpret_cast<uint32_t*>(S->m2);  // object data is no longer needed
  const uint32_t from = S->id[lane];
  wave_sync();
  fin[from] = lane;
  wave_sync();
  if (hv[H_OWN] && hl < hv[H_OWN]) {
    const uint32_t f = fin[lane];
    S->operm[f] = oid - hv[H_OBASE];
    S->otri[f][0] = tri0;
    S->otri[f][1] = tri1;
    S->otri[f][2] = tri2;
  }
  wave_sync();  // the next pair's flush reads the staged output
  if (forced) {
    // leaf forced by max_depth (lib.rs:556) holding more than 32 objects
    const SNode sn = small[e];
    const Rec* src = sn.parity ? rec1 : rec0;
    if (hl == 0) {
      Box fb;
#pragma unroll
      for (int x = 0; x < 3; ++x) { fb.mn[x] = sn.bmn[x]; fb.mx[x] = sn.bmx[x]; }
      out.write_node(sn.idx, fb, sn.start - sn.obase, 0, sn.count);
      put_gap(ctr, gaps, sn.P, sn.count);
    }
    for (uint32_t k = hl; k < sn.count; k += 32)
      put_perm(sn.start + k, src[sn.start + k].id, sn.obase, perm, tris_in, tris_out);
  }
}

// Persistent over the small-subtree queue in windows of `win` (2..SMALL_WIN)
// subtrees: a wavefront ranks its window's subtrees by size and builds them in
// pairs of neighbours in that order, so the two halves of a wavefront mostly
// hold subtrees of similar size and depth (a level costs as much as its
// largest node, and a pair as many levels as its deeper subtree).  Small builds
// take windows of 2 (plain pairs): there a window's pairs, built one after
// another, would lengthen the launch.  Only wave-local synchronisation inside.
// Four wavefronts per SIMD: the pair state and the candidate loop fit in 128
// VGPRs without scratch (at five, spills put vector memory operations, which
// retire in order behind the stores, into the loop and the flush).
// SPLITS: split records are staged too (more LDS, fewer resident blocks).
constexpr uint32_t SMALL_WIN = 8;  // windows of 16 and 32 measured slower: coarser work items
template <bool SPLITS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_small(const SNode* __restrict__ small,
                                               const uint32_t* __restrict__ nS_dev,
                                               const Rec* __restrict__ rec0,
                                               const Rec* __restrict__ rec1, NodeOut out,
                                               uint32_t* __restrict__ perm,
                                               const uint32_t* __restrict__ tris_in,
                                               uint32_t* __restrict__ tris_out, Counters* ctr,
                                               uint32_t* gaps, uint32_t max_depth,
                                               uint32_t win) {
  __shared__ SmallObjs lds_obj[4];
  __shared__ uint32_t lds_order[4][SMALL_WIN];
  __shared__ uint4 lds_split[SPLITS ? 4 * 128 : 1];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint32_t nS = *nS_dev;
  SmallObjs* S = &lds_obj[w];
  uint4* sp = SPLITS ? &lds_split[128 * w] : nullptr;
  if ((lane & 31) == 0) S->hdr[lane >> 5][H_OWN] = 0;  // nothing staged yet
  wave_sync();
  for (uint32_t wi = blockIdx.x * 4 + w; wi * win < nS; wi += gridDim.x * 4) {
    const uint32_t base = wi * win;
    const uint32_t cnt = min(win, nS - base);
    // key: size (forced leaves above 32 as 33), then position: all distinct
    uint32_t key = ~0u;
    if (lane < cnt) key = (min(small[base + lane].count, 33u) << 3) | lane;
    uint32_t rank = 0;
    for (uint32_t j = 0; j < cnt; ++j)
      rank += (uint32_t)__builtin_amdgcn_readlane((int)key, (int)j) < key;
    if (lane < cnt) lds_order[w][rank] = base + lane;
    wave_sync();
    for (uint32_t q = 0; q < cnt; q += 2) {
      const uint32_t ea = lds_order[w][q];
      const uint32_t eb = q + 1 < cnt ? lds_order[w][q + 1] : nS;
      small_pair(ea, eb, S, sp, small, nS, rec0, rec1, out, perm, tris_in, tris_out, ctr, gaps,
                 max_depth);
    }
    wave_sync();  // the next window's order overwrites this one
  }
  small_flush(S, sp, out, perm, tris_out, ctr, gaps);
}

// ------------------------------------------------------------------ compaction
// Only after max_depth cut subtrees short.  A cut leaf of n objects at pair
// base P leaves slots [P, P + 2(n - 1)) unused; inside its mesh, slot i moves
// to i - (unused slots of that mesh before i) and child indices likewise.
// upto(x) = total length of the gaps starting at or before slot x.
__device__ inline uint32_t gap_upto(const uint32_t* gstart, const uint32_t* gcum, uint32_t ng,
                                    uint32_t x, uint32_t& nle) {
  uint32_t lo = 0, hi = ng;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (gstart[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  nle = lo;
  return lo == 0 ? 0 : gcum[lo - 1];
}

__global__ void k_compact(const mgbvh_node32* __restrict__ nin, const uint32_t* __restrict__ lin,
                          const mgbvh_split_rec* __restrict__ sin, mgbvh_node32* nout,
                          uint32_t* lout, mgbvh_split_rec* sout, uint32_t nslots,
                          const uint32_t* gstart, const uint32_t* glen, const uint32_t* gcum,
                          uint32_t ng, const uint32_t* __restrict__ mnb, uint32_t nm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  // mesh of slot i: the last mesh whose first slot is <= i
  uint32_t lo = 0, hi = nm;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (mnb[mid] <= i) lo = mid;
    else hi = mid;
  }
  const uint32_t base = mnb[lo];
  uint32_t nle;
  const uint32_t before = base == 0 ? 0 : gap_upto(gstart, gcum, ng, base - 1, nle);
  const uint32_t cum = gap_upto(gstart, gcum, ng, i, nle);
  if (nle > 0 && i < gstart[nle - 1] + glen[nle - 1]) return;  // an unused slot
  const uint32_t ni = i - (cum - before);
  mgbvh_node32 nd = nin[i];
  if (nd.child_node != 0)
    nd.child_node -= gap_upto(gstart, gcum, ng, base + nd.child_node, nle) - before;
  nout[ni] = nd;
  if (lout) lout[ni] = lin[i];
  if (sout) sout[ni] = sin[i];
}

}  // namespace mg

// ====================================================================== host
#include <algorithm>
#include <vector>

namespace mg {

#define MG_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

// kernel timing events on the enclosing function's stream `st`
#define MG_PROF(kind, end)                                                       \
  do {                                                                           \
    if (c->prof_mask & (1u << (kind))) {                                         \
      int r_ = (end) ? prof_end(c, (kind), st) : prof_begin(c, (kind), st);    \
      if (r_) return r_;                                                         \
    }                                                                            \
  } while (0)

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__global__ void k_fill(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Fill with a kernel, not hipMemsetAsync: the same sequence is captured in graphs.
static int fill_u32(hipStream_t st, void* p, uint32_t v, uint64_t n) {
  if (!n) return MGBVH_OK;
  k_fill<<<std::min<uint64_t>(cdiv(n, 256), 2048), 256, 0, st>>>((uint32_t*)p, v, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_fill");
}

int build_device(mgbvh_ctx* c, const BuildInput& in, uint32_t n, const mgbvh_settings* s,
                 const BuildOutput& o, BuildResult* res) {
  MeshIn m{};
  m.tri_offset = 0;
  m.count = n;
  m.vertex_offset = 0;
  m.vertex_count = in.nverts > 0xffffffffull ? 0xffffffffu : (uint32_t)in.nverts;
  m.has_bounds = s->has_bounds;
  for (int a = 0; a < 3; ++a) { m.bmn[a] = s->bounds_min[a]; m.bmx[a] = s->bounds_max[a]; }
  return build_multi(c, in, std::vector<MeshIn>{m}, s, o, res);
}

int build_multi(mgbvh_ctx* c, const BuildInput& in, const std::vector<MeshIn>& meshes,
                const mgbvh_settings* s, const BuildOutput& o, BuildResult* res) {
  *res = BuildResult{};
  BuildPlan p;
  int rc = plan_build(c, c->sc, in, meshes, s, o, &p);
  if (rc) return rc;
  res->mesh_nodes.assign(p.nm, 0);
  res->mesh_nbase.assign(p.nm, 0);
  for (uint32_t i = 0; i < p.nm; ++i) res->mesh_nbase[i] = p.rows[i].nbase;
  if (p.n == 0) return MGBVH_OK;
  hipStream_t st = c->stream;
  MG_HIP(hipEventRecord(c->ev[62], st));
  uint32_t levels = 0;
  if ((rc = enqueue_build(c, c->sc, p, 0, true, &levels))) return rc;
  if ((rc = finish_build(c, c->sc, p, res))) return rc;
  MG_HIP(hipEventRecord(c->ev[63], st));
  MG_HIP(hipEventSynchronize(c->ev[63]));
  float ms = 0.0f;
  MG_HIP(hipEventElapsedTime(&ms, c->ev[62], c->ev[63]));
  if ((rc = prof_collect(c))) return rc;
  res->ms = ms;
  return MGBVH_OK;
}

int plan_build(mgbvh_ctx* c, Scratch& sc, const BuildInput& in, const std::vector<MeshIn>& meshes,
               const mgbvh_settings* s, const BuildOutput& o, BuildPlan* pp) {
  BuildPlan& p = *pp;
  p = BuildPlan{};
  p.in = in;
  p.o = o;
  p.s = *s;
  p.nm = (uint32_t)meshes.size();
  // chunks, node slots and root slots of every mesh (meshes are contiguous)
  p.rows.resize(p.nm);
  uint64_t n64 = 0, slots64 = 0;
  for (uint32_t i = 0; i < p.nm; ++i) {
    const MeshIn& m = meshes[i];
    if (m.tri_offset != n64) {
      set_error("meshes must be contiguous: tri_offset of each mesh = sum of earlier counts");
      return MGBVH_ERR_INVALID_ARG;
    }
    MeshRow& r = p.rows[i];
    r = MeshRow{};
    r.tri_offset = m.tri_offset;
    r.count = m.count;
    r.vertex_offset = m.vertex_offset;
    r.vertex_count = m.vertex_count;
    r.setup_chunk0 = (uint32_t)p.setup_chunks;
    r.nbase = (uint32_t)slots64;
    r.has_bounds = m.has_bounds ? 1u : 0u;
    for (int a = 0; a < 3; ++a) { r.bmn[a] = m.bmn[a]; r.bmx[a] = m.bmx[a]; }
    if (m.count == 0) {
      r.root_kind = 0;  // lib.rs:703-708: empty BVH
    } else if (m.count > MG_SMALL && s->max_depth > 0) {
      r.root_kind = 1;
      r.root_slot = p.n_large++;
      r.lchunk0 = (uint32_t)p.lchunks;
      p.lchunks += cdiv(m.count, CH);
    } else {
      r.root_kind = 2;
      r.root_slot = p.n_small++;
    }
    n64 += m.count;
    slots64 += m.count ? 2ull * m.count - 1 : 0;
    p.setup_chunks += cdiv(m.count, CH);
    if (n64 > 0x7fffffffull || slots64 > 0xffffffffull) {
      set_error("more than 2^31-1 objects in one build");
      return MGBVH_ERR_INVALID_ARG;
    }
  }
  p.n = (uint32_t)n64;
  p.nslots = slots64;
  if (p.n == 0) return MGBVH_OK;
  p.capL = p.n / (MG_SMALL + 1) + 2;
  p.capC = cdiv(p.n, CH) + p.capL + 2;
  // nodes with > 32 objects have depth < max_depth, one level per depth
  p.max_levels = std::min<uint32_t>(s->max_depth, p.n);
  int rc;
  if (s->method == MGBVH_METHOD_LEARNED) {
    if (!c->model) { set_error("the learned method needs a model"); return MGBVH_ERR_UNSUPPORTED; }
    p.lp.min_objects = s->learned_min_objects
                           ? std::max<uint32_t>(s->learned_min_objects, MG_SMALL + 1)
                           : 512u;
    p.lp.window = s->learned_window ? s->learned_window : 3u;
    p.lp.guard = s->learned_guard;
    p.lp.no_network = s->learned_no_network;
    p.lp.model = c->model;
    for (uint32_t i = 0; i < p.nm; ++i)
      if (p.rows[i].root_kind == 1 && p.rows[i].count >= p.lp.min_objects)
        p.rows[i].lrow = p.n_counted++;
    for (int q = 0; q < 2; ++q)
      if ((rc = ensure(sc.lrows[q], (size_t)p.capL * 4))) return rc;
    if ((rc = ensure(sc.lfeat, (size_t)p.capL * MLP_IN * 4)) ||
        (rc = ensure(sc.llogit, (size_t)p.capL * MLP_OUT * 4)))
      return rc;
  }
  const uint32_t n = p.n;
  if ((rc = ensure(sc.recs[0], (size_t)n * sizeof(Rec)))) return rc;
  if ((rc = ensure(sc.recs[1], (size_t)n * sizeof(Rec)))) return rc;
  for (int q = 0; q < 2; ++q) {
    if ((rc = ensure(sc.L[q], (size_t)p.capL * sizeof(LNode)))) return rc;
    if ((rc = ensure(sc.cn[q], (size_t)p.capC * 4))) return rc;
    if ((rc = ensure(sc.ck[q], (size_t)p.capC * 4))) return rc;
  }
  if ((rc = ensure(sc.hist, (size_t)p.capL * HIST_WORDS * 4))) return rc;
  if ((rc = ensure(sc.blk, (size_t)p.capC * 96 * 4))) return rc;
  if ((rc = ensure(sc.small, (size_t)small_list_cap(n) * sizeof(SNode)))) return rc;
  if ((rc = ensure(sc.ctr, sizeof(Counters)))) return rc;
  if ((rc = ensure(sc.gaps, ((size_t)n / 2 + 1) * 8))) return rc;
  if ((rc = ensure(sc.mrows, (size_t)p.nm * sizeof(MeshRow)))) return rc;
  if ((rc = ensure(sc.macc, (size_t)p.nm * 12 * 4))) return rc;
  if ((rc = ensure(sc.lv, ((size_t)p.max_levels + 2) * LV_WORDS * 4))) return rc;
  p.lens = o.lens;
  if (!p.lens) {
    if ((rc = ensure(sc.lens, p.nslots * 4))) return rc;
    p.lens = (uint32_t*)sc.lens.p;
  }
  return MGBVH_OK;
}

int enqueue_build(mgbvh_ctx* c, Scratch& sc, BuildPlan& p, uint32_t fixed_levels,
                  bool upload_rows, uint32_t* levels_run, hipStream_t stream,
                  const std::function<int()>* after_setup) {
  hipStream_t st = stream ? stream : c->stream;
  const BuildInput& in = p.in;
  const BuildOutput& o = p.o;
  const mgbvh_settings* s = &p.s;
  const uint32_t nm = p.nm, n = p.n;
  const LearnParams& lp = p.lp;
  int rc;
  Rec* recs[2] = {(Rec*)sc.recs[0].p, (Rec*)sc.recs[1].p};
  LNode* L[2] = {(LNode*)sc.L[0].p, (LNode*)sc.L[1].p};
  uint32_t* cn[2] = {(uint32_t*)sc.cn[0].p, (uint32_t*)sc.cn[1].p};
  uint32_t* ck[2] = {(uint32_t*)sc.ck[0].p, (uint32_t*)sc.ck[1].p};
  uint32_t* lrows[2] = {(uint32_t*)sc.lrows[0].p, (uint32_t*)sc.lrows[1].p};
  Counters* ctr = (Counters*)sc.ctr.p;
  SNode* small = (SNode*)sc.small.p;
  SNode* med = small + n;  // the medium list follows the small-subtree list's n entries
  // medium nodes only outside sample collection (samples come from the level path's sweeps)
  const uint32_t med_max = o.samples ? 0u : c->tune.med_max;
  uint32_t* hist = (uint32_t*)sc.hist.p;
  uint32_t* blk = (uint32_t*)sc.blk.p;
  uint32_t* gaps = (uint32_t*)sc.gaps.p;
  uint32_t* lv = (uint32_t*)sc.lv.p;
  MeshRow* drows = (MeshRow*)sc.mrows.p;
  uint32_t* macc = (uint32_t*)sc.macc.p;
  NodeOut out{o.nodes, p.lens, o.splits};
  out.samples = o.samples;
  out.labels = o.labels;
  out.max_samples = o.max_samples;
  out.sample_min = o.sample_min;
  const bool graph = fixed_levels > 0;
  const uint32_t cus = (uint32_t)c->cus;
  const Tuning& tu = c->tune;

  if (upload_rows)
    MG_HIP(hipMemcpyAsync(drows, p.rows.data(), (size_t)nm * sizeof(MeshRow),
                          hipMemcpyHostToDevice, st));
  if ((rc = fill_u32(st, ctr, 0, sizeof(Counters) / 4))) return rc;
  if ((rc = fill_u32(st, macc, 0, (uint64_t)nm * 12))) return rc;
  if ((rc = fill_u32(st, lv, 0, ((uint64_t)p.max_levels + 2) * LV_WORDS))) return rc;
  if (o.splits && (rc = fill_u32(st, o.splits, 0xffffffffu, p.nslots * 4))) return rc;

  MG_PROF(KK_SETUP, 0);
  if (in.kind == 0)
    k_setup_tris<<<(uint32_t)p.setup_chunks, 256, 0, st>>>(in.pos, in.vstride_floats, in.tris,
                                                           drows, nm, recs[0], macc, ctr);
  else
    k_setup_aabbs<<<(uint32_t)p.setup_chunks, 256, 0, st>>>(in.boxes, drows, nm, recs[0], macc,
                                                            ctr);
  MG_HIP(hipGetLastError());
  MG_PROF(KK_SETUP, 1);
  MG_PROF(KK_ROOT, 0);
  k_root<<<nm, 64, 0, st>>>(drows, macc, ctr, L[0], cn[0], ck[0], small, p.n_small,
                            lp.min_objects, lrows[0], hist, lv, p.n_large,
                            (uint32_t)p.lchunks, p.n_counted);
  MG_PROF(KK_ROOT, 1);
  MG_HIP(hipGetLastError());
  // the object boxes and the per-mesh root bounds are in place
  if (after_setup && (rc = (*after_setup)())) return rc;

  // level-by-level: exact grids from the counts read back after each level;
  // fixed levels (graph capture): grids sized for the capacities
  uint32_t nL = p.n_large, nC = (uint32_t)p.lchunks, n_counted = p.n_counted, nBin = nC;
  int cur = 0;
  uint32_t level = 0;
  for (;; ++level) {
    if (graph) {
      if (level >= fixed_levels) break;
      nL = p.capL;
      nC = p.capC;
      n_counted = lp.min_objects ? p.capL : 0;
      nBin = nC;
    } else {
      if (nL == 0) break;
      if (level > p.max_levels + 1) { set_error("level loop did not terminate"); return MGBVH_ERR_INTERNAL; }
    }
    const uint32_t* lvc = lv + (size_t)level * LV_WORDS;
    uint32_t* lvn = lv + (size_t)(level + 1) * LV_WORDS;
    const uint32_t g_bin = std::max<uint32_t>(1, std::min<uint32_t>(nC, cus * tu.bin_bpc));
    const uint32_t g_part = std::max<uint32_t>(1, std::min<uint32_t>(nC, cus * tu.part_bpc));
    const uint32_t g_sweep =
        std::max<uint32_t>(1, std::min<uint32_t>(cdiv(nL, SW_WAVES), cus * tu.sweep_bpc));
    if (nBin) {  // level by level: skipped once every node was binned by its parent
      MG_PROF(KK_BIN, 0);
      k_bin<<<g_bin, 64 * tu.bin_waves, 0, st>>>(L[cur], cn[cur], ck[cur], recs[cur], hist, blk,
                                                  lvc);
      MG_PROF(KK_BIN, 1);
    }
    if (n_counted) {
      MG_PROF(KK_LEARN, 0);
      if ((rc = learned_decide(c, sc, L[cur], lrows[cur], lvc + LV_NCOUNTED, n_counted, hist,
                               ctr, lp)))
        return rc;
      MG_PROF(KK_LEARN, 1);
    }
    MG_PROF(KK_SWEEP, 0);
    k_sweep<<<g_sweep, 1024, 0, st>>>(L[cur], lvc, hist, recs[cur], L[cur ^ 1], cn[cur ^ 1],
                                      ck[cur ^ 1], small, med, ctr, out, s->max_depth,
                                      (uint32_t)(cur ^ 1), lp.min_objects, med_max,
                                      lrows[cur ^ 1], lvn, tu.sweep_pack);
    MG_PROF(KK_SWEEP, 1);
    MG_PROF(KK_PARTITION, 0);
    // children's splits are taken inside the partition unless samples are
    // collected (those need the histogram rows)
    const uint32_t fuse = out.samples ? 0u : 1u;
    // The first levels hold the largest nodes (whole meshes, their halves and
    // quarters: full 2048-object chunks), which four-wavefront blocks move
    // faster; with part_waves 1 the levels below take one-wavefront blocks.
    constexpr uint32_t PART_WIDE_LEVELS = 4;
    if (tu.part_waves == 1 && level >= PART_WIDE_LEVELS)
      k_partition<1><<<g_part, 64, 0, st>>>(L[cur], cn[cur], ck[cur], blk, recs[cur],
                                            recs[cur ^ 1], L[cur ^ 1], hist, lvc, fuse);
    else
      k_partition<4><<<g_part, 256, 0, st>>>(L[cur], cn[cur], ck[cur], blk, recs[cur],
                                             recs[cur ^ 1], L[cur ^ 1], hist, lvc, fuse);
    MG_PROF(KK_PARTITION, 1);
    MG_HIP(hipGetLastError());
    if (!graph) {
      MG_HIP(hipMemcpyAsync(&sc.h_ctr->pad[0], lvn, LV_WORDS * 4, hipMemcpyDeviceToHost, st));
      MG_HIP(hipMemcpyAsync(&sc.h_ctr->err, &ctr->err, 4, hipMemcpyDeviceToHost, st));
      MG_HIP(hipStreamSynchronize(st));
      if (sc.h_ctr->err) { ++level; break; }
      nL = sc.h_ctr->pad[LV_NL];
      nC = sc.h_ctr->pad[LV_NC];
      n_counted = sc.h_ctr->pad[LV_NCOUNTED];
      nBin = sc.h_ctr->pad[LV_NBIN];
    }
    cur ^= 1;
  }
  if (levels_run) *levels_run = level;
  // medium nodes: their subtrees above the <= 32-object subtrees, one wavefront each
  if (med_max) {
    const uint32_t g_med =
        std::max<uint32_t>(1, std::min<uint32_t>(n / (MG_SMALL + 1) + 1, cus * tu.med_bpc));
    MG_PROF(KK_MEDIUM, 0);
    k_medium<<<g_med, 64, 0, st>>>(med, &ctr->n_med, recs[0], recs[1], small, n, ctr, out,
                                   s->max_depth, (uint32_t)p.nslots);
    MG_PROF(KK_MEDIUM, 1);
    MG_HIP(hipGetLastError());
  }
  // subtrees of <= 32 objects (and depth-cut leaves), persistent over their queue
  const uint32_t g_small = std::max<uint32_t>(1, std::min<uint32_t>(cdiv(n, 8), cus * tu.small_bpc));
  MG_PROF(KK_SMALL, 0);
  // size-ranked windows where even the fewest possible subtrees (n / 32) fill
  // every wavefront of the grid with a window; plain pairs otherwise
  const uint32_t win =
      (uint64_t)n / (32u * SMALL_WIN) >= 4ull * g_small ? SMALL_WIN : 2u;
  (out.splits ? k_small<true> : k_small<false>)<<<g_small, 256, 0, st>>>(
      small, &ctr->n_small, recs[0], recs[1], out, o.perm,
                                   in.kind == 0 ? in.tris : nullptr,
                                   in.kind == 0 ? o.tris_out : nullptr, ctr, gaps, s->max_depth,
                                   win);
  MG_PROF(KK_SMALL, 1);
  MG_HIP(hipGetLastError());
  return MGBVH_OK;
}

int finish_build(mgbvh_ctx* c, Scratch& sc, BuildPlan& p, BuildResult* res) {
  hipStream_t st = c->stream;
  const BuildOutput& o = p.o;
  const uint32_t nm = p.nm;
  const uint64_t nslots = p.nslots;
  Counters* ctr = (Counters*)sc.ctr.p;
  uint32_t* gaps = (uint32_t*)sc.gaps.p;
  uint32_t* lens = p.lens;
  int rc;
  MG_HIP(hipMemcpyAsync(sc.h_ctr, ctr, sizeof(Counters), hipMemcpyDeviceToHost, st));
  MG_HIP(hipStreamSynchronize(st));
  const Counters hc = *sc.h_ctr;
  if (hc.err & ERR_INDEX) { set_error("triangle index out of range"); return MGBVH_ERR_INDEX; }
  if (hc.err & ERR_NONFINITE) {
    set_error("non-finite or out-of-range (|x| >= 2^60) input coordinate");
    return MGBVH_ERR_NONFINITE;
  }
  if (hc.err & ERR_INTERNAL) {
    set_error("builder invariant violated (a split count, a list capacity or a stack depth)");
    return MGBVH_ERR_INTERNAL;
  }
  if (p.fixed_levels) {  // a replayed graph: did its levels reach every node?
    uint32_t waiting = 0;
    MG_HIP(hipMemcpy(&waiting, (uint32_t*)sc.lv.p + (size_t)p.fixed_levels * LV_WORDS + LV_NL, 4,
                     hipMemcpyDeviceToHost));
    if (waiting) return MG_LEVELS_SHORT;
  }
  res->mesh_nodes.assign(nm, 0);
  res->mesh_nbase.assign(nm, 0);
  for (uint32_t i = 0; i < nm; ++i) {
    res->mesh_nodes[i] = p.rows[i].count ? 2 * p.rows[i].count - 1 : 0;
    res->mesh_nbase[i] = p.rows[i].nbase;
  }
  uint64_t removed = 0;
  if (hc.n_gaps) {
    std::vector<uint32_t> g(2 * (size_t)hc.n_gaps);
    MG_HIP(hipMemcpy(g.data(), gaps, g.size() * 4, hipMemcpyDeviceToHost));
    std::vector<std::pair<uint32_t, uint32_t>> gp(hc.n_gaps);
    for (uint32_t i = 0; i < hc.n_gaps; ++i) gp[i] = {g[2 * i], g[2 * i + 1]};
    std::sort(gp.begin(), gp.end());
    std::vector<uint32_t> tab(3 * (size_t)hc.n_gaps);
    uint64_t cum = 0;
    uint32_t mi = 0;
    for (uint32_t i = 0; i < hc.n_gaps; ++i) {
      cum += gp[i].second;
      tab[i] = gp[i].first;
      tab[hc.n_gaps + i] = gp[i].second;
      tab[2 * hc.n_gaps + i] = (uint32_t)cum;
      while (mi + 1 < nm && p.rows[mi + 1].nbase <= gp[i].first) ++mi;
      res->mesh_nodes[mi] -= gp[i].second;
    }
    removed = cum;
    std::vector<uint32_t> mnb(nm);
    for (uint32_t i = 0; i < nm; ++i) mnb[i] = p.rows[i].nbase;
    if ((rc = ensure(sc.gtab, tab.size() * 4))) return rc;
    if ((rc = ensure(sc.mnb, (size_t)nm * 4))) return rc;
    if ((rc = ensure(sc.tmpn, nslots * sizeof(mgbvh_node32)))) return rc;
    if ((rc = ensure(sc.tmpl, nslots * 4))) return rc;
    if (o.splits && (rc = ensure(sc.tmps, nslots * sizeof(mgbvh_split_rec)))) return rc;
    MG_HIP(hipMemcpyAsync(sc.gtab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st));
    MG_HIP(hipMemcpyAsync(sc.mnb.p, mnb.data(), mnb.size() * 4, hipMemcpyHostToDevice, st));
    const uint32_t* gt = (const uint32_t*)sc.gtab.p;
    MG_PROF(KK_COMPACT, 0);
    k_compact<<<cdiv(nslots, 256), 256, 0, st>>>(
        o.nodes, lens, o.splits, (mgbvh_node32*)sc.tmpn.p, (uint32_t*)sc.tmpl.p,
        o.splits ? (mgbvh_split_rec*)sc.tmps.p : nullptr, (uint32_t)nslots, gt, gt + hc.n_gaps,
        gt + 2 * hc.n_gaps, hc.n_gaps, (const uint32_t*)sc.mnb.p, nm);
    MG_PROF(KK_COMPACT, 1);
    MG_HIP(hipGetLastError());
    // every mesh keeps its first slot; its compacted nodes fill the front of its range
    MG_HIP(hipMemcpyAsync(o.nodes, sc.tmpn.p, nslots * sizeof(mgbvh_node32),
                          hipMemcpyDeviceToDevice, st));
    MG_HIP(hipMemcpyAsync(lens, sc.tmpl.p, nslots * 4, hipMemcpyDeviceToDevice, st));
    if (o.splits)
      MG_HIP(hipMemcpyAsync(o.splits, sc.tmps.p, nslots * sizeof(mgbvh_split_rec),
                            hipMemcpyDeviceToDevice, st));
    MG_HIP(hipStreamSynchronize(st));
  }
  // level counts: nodes with > 32 objects, and whether a fixed level count sufficed
  std::vector<uint32_t> lvh(((size_t)p.max_levels + 2) * LV_WORDS);
  MG_HIP(hipMemcpy(lvh.data(), sc.lv.p, lvh.size() * 4, hipMemcpyDeviceToHost));
  uint64_t large_nodes = hc.n_med_split;  // split inside medium subtrees
  uint32_t levels = 0;
  for (uint32_t l = 0; l <= p.max_levels; ++l) {
    large_nodes += lvh[(size_t)l * LV_WORDS + LV_NL];
    if (lvh[(size_t)l * LV_WORDS + LV_NL]) levels = l + 1;
  }
  if (lvh[((size_t)p.max_levels + 1) * LV_WORDS + LV_NL]) {
    set_error("builder produced nodes beyond max_depth levels");
    return MGBVH_ERR_INTERNAL;
  }
  res->n_levels = levels;
  res->large_nodes = large_nodes;
  res->nnodes = nslots - removed;
  res->n_median = hc.n_median;
  res->n_cutoff = hc.n_cutoff;
  res->large_visits = hc.sum_large;
  res->n_samples = hc.n_samples;
  res->n_window = hc.n_window;
  res->n_guard = hc.n_fallback;
  return MGBVH_OK;
}

}  // namespace mg
