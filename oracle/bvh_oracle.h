/*
 * CPU oracle for the marigold `bvh` crate (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C, op-for-op restatement of `bvh/src/lib.rs` used by tests/, the
 * smoke check and bench.py's cpu_baseline leg as the checker.  The product
 * (marigold_amd / libmgbvh.so) never links or calls this code.
 *
 * Everything is fp32 with strict left-to-right operation order and no FMA
 * contraction (build with -ffp-contract=off, no -ffast-math).
 */
#ifndef MGBVH_ORACLE_H
#define MGBVH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BoundingVolume, bvh/src/lib.rs:41-46 (x,y,z lanes of two Vec3A). */
typedef struct { float mn[3]; float mx[3]; } or_bv;

/* BvhNode<MIN,MAX>, bvh/src/lib.rs:149-155. */
typedef struct { or_bv b; uint32_t start, len, child; } or_node;

/* Split decision record: one per internal node, indexed by node index.
 * method: 0 = binned (lib.rs:331), 1 = adaptive threshold (lib.rs:391),
 * 2 = median fallback (lib.rs:444).  value: bin cut i / threshold bits / m. */
typedef struct { uint32_t axis_method, value, count_lt, cost_bits; } or_rec;

typedef struct {
  uint32_t min_leaf, max_leaf;  /* const generics MIN_LEAF_OBJECTS / MAX_LEAF_OBJECTS */
  uint32_t max_depth;           /* BvhSettings.max_depth */
  int32_t has_bounds;           /* BvhSettings.bounds is Some */
  or_bv bounds;
} or_settings;

typedef struct {
  uint32_t nnodes, height;
  uint32_t n_binned, n_adaptive, n_median;
  uint32_t n_cutoff_leaves;     /* leaves forced by depth == max_depth holding > 1 object */
  uint32_t n_tie_events;        /* list-order-dependent choices that can change the node array */
  uint32_t n_median_quirk;      /* Exact split whose membership != winning axis' membership */
  uint32_t n_assert_fail;       /* a reference assert! would have panicked */
  uint32_t n_cost_leaf;         /* leaf because split.cost >= leaf_cost (MAX > 1 only) */
  uint32_t n_refused;           /* leaf because select_split returned None */
  uint32_t n_tie_benign;        /* order-dependent choices between objects with identical boxes */
} or_stats;

/* MeshTriangle::bounding_volume (mesh_interface/src/lib.rs:19-30); -0.0 -> +0.0. */
void or_tri_aabbs(const float* pos, uint32_t vstride_floats, const uint32_t* tris,
                  uint64_t n, float* aabbs_out);

/* BoundingVolumeHierarchy::new (lib.rs:698-878) over objects given by AABBs.
 * list: in/out object ids (permuted in place like `&mut [T]`).
 * nodes: capacity 2n-1.  recs: capacity 2n-1 (may be NULL).
 * Returns 0 on success. */
int or_build(const float* aabbs, uint32_t n, const or_settings* s, uint32_t* list,
             or_node* nodes, or_rec* recs, or_stats* st);

/* Pack nodes into the 32-byte GPU record (lib.rs:157-212, MAX == 1 branch). */
void or_encode_nodes(const or_node* nodes, uint32_t nnodes, uint32_t* out8);

/* Tree SAH cost with the reference's constants (lib.rs:217-218), in double. */
double or_tree_sah(const or_node* nodes, uint32_t nnodes);

/* Single helpers exported for known-answer tests. */
float or_surface_area(const float* bv6);
int32_t or_total_cmp_key(float f);
uint32_t or_bin_index(float c, float cmin, float cext, uint32_t nbins);
float or_binned_cost(const float* parent6, const float* lt6, const float* gt6,
                     uint32_t n_lt, uint32_t n_gt);

/* Closest hit (t, prim) by brute force: Moller-Trumbore per scene.slang:19-89,
 * lowest primitive id wins ties.  rays: n x {ox,oy,oz,dx,dy,dz}. */
void or_closest_hit_brute(const float* pos, uint32_t vstride_floats, const uint32_t* tris,
                          uint32_t ntris, const float* rays, uint32_t nrays,
                          float* t_out, uint32_t* prim_out);
float or_ray_tri(const float* ray6, const float* v1, const float* v2, const float* v3);

/* Closest hit by traversing a built tree (32-byte nodes as (k, 8) u32, lens
 * optional for depth-cut leaves): conservative slab test, so the result equals
 * the brute force.  tris_perm: the permuted triangle list, perm: original
 * index of each position.  rays: n x {ox,oy,oz,dx,dy,dz}. */
void or_closest_hit_bvh(const uint32_t* nodes8, const uint32_t* lens, const float* pos,
                        uint32_t vstride_floats, const uint32_t* tris_perm, const uint32_t* perm,
                        const float* rays, uint32_t nrays, float* t_out, uint32_t* prim_out);

/* Closest hit in a two-level scene by brute force: every instance (ascending
 * index) takes the ray to object space with its 3x4 world_to_object rows
 * (((m0*x + m1*y) + m2*z) + m3, direction without m3) and tests every triangle
 * of its mesh (ascending index); (t, inst, prim) lexicographic minimum.
 * inst16: ninst records of 16 words (12 floats, BLAS index at word 12).
 * mesh3: per mesh {first triangle, triangle count, vertex offset} into the
 * concatenated ORIGINAL-order triangle list. */
void or_closest_hit_scene_brute(const float* pos, uint32_t vstride_floats, const uint32_t* tris,
                                const uint32_t* mesh3, const uint32_t* inst16, uint32_t ninst,
                                const float* rays, uint32_t nrays, float* t_out,
                                uint32_t* prim_out, uint32_t* inst_out);

/* The path tracer's own two-level traversal, step for step
 * (scene.slang:91-221; box test :296-310; Hit.merge intersect.slang:73-82;
 * one 64-entry stack shared by both levels, scene.slang:5-7): the BLAS root is
 * pushed without a box test, a popped node is not culled again, children are
 * pushed when their box distance is strictly below the closest distance, the
 * BLAS pass starts from the world distance times length(object-space
 * direction) and its result is merged back unconverted, and an equal distance
 * goes to the hit tested later.  Tree dependent by design.
 * tlas8/blas8: 32-byte nodes as (k, 8) u32 (BLAS child indices relative to
 * the BLAS's first node); tlas_perm: TLAS start_index -> instance;
 * inst16: as or_closest_hit_scene_brute; blas4: per BLAS {node offset,
 * triangle offset, vertex offset, 0}; tris_perm/blas_perm: permuted
 * triangles and their original indices.  hits4 per ray: {distance bits,
 * prim, inst, slot}: a miss is {distance, ~0, ~0, ~0}; a success that holds no
 * triangle (a merge that took NO_HIT) has slot 0xfffffffe.  Returns 0, or
 * bit 0 = a push on the full stack was dropped, bit 1 = an index out of range. */
int or_raytrace_tlas(const uint32_t* tlas8, const uint32_t* tlas_perm, const uint32_t* inst16,
                     uint32_t ninst, const uint32_t* blas4, uint32_t nblas, const uint32_t* blas8,
                     const float* pos, uint32_t vstride_floats, const uint32_t* tris_perm,
                     const uint32_t* blas_perm, const float* rays, uint32_t nrays,
                     uint32_t* hits4);

/* Compressed 8-wide BVH from a binary tree (the wide format of Ylitie,
 * Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs Through
 * Compressed Wide BVHs", HPG 2017, which marigold's README names as its next
 * step).  Collapse by the paper's dynamic program: C(v, i), the cost of v's
 * subtree as at most i subtrees; a wide leaf holds at most p_max objects
 * (cost area * count * c_prim), a wide internal node costs area * c_node plus
 * the best split of 8 slots between the children; first minimum over k, a
 * leaf on a tie.  Layout: the root at 0; a node's internal children
 * contiguous at base_child, then each child's own subtree in child order; its
 * leaves' objects contiguous at base_tri, then each child's objects.  Node of
 * 80 bytes: p[3] f32, e[3] biased exponents, imask, base_child, base_tri,
 * meta[8], qlo x/y/z [8], qhi x/y/z [8]; a child plane decodes as
 * p + q * 2^(e-127) in fp32 and always contains the exact box.  tri_index:
 * the binary leaves' start_index values in wide-leaf order.  p_max is 1..3 (-3
 * otherwise).  Returns 0, -2
 * when a binary leaf holds more than p_max objects. */
int or_wide8_collapse(const uint32_t* nodes8, const uint32_t* lens, uint32_t nnodes,
                      float c_node, float c_prim, uint32_t p_max, uint8_t* out80,
                      uint32_t* tri_index, uint32_t* out_nwide, float* out_cost);

/* Closest hit through such a wide tree (conservative slab test, lowest
 * original triangle on equal t): equals or_closest_hit_bvh on the binary tree. */
void or_trace_wide8(const uint8_t* w80, const uint32_t* tri_index, const float* pos,
                    uint32_t vstride_floats, const uint32_t* tris_perm, const uint32_t* perm,
                    const float* rays, uint32_t nrays, float* t_out, uint32_t* prim_out);

#ifdef __cplusplus
}
#endif
#endif
