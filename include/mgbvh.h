/*
 * mgbvh -- MI355X-native SAH BVH builder with the node format and builder API
 * of marigold's `bvh` crate.  Plain C ABI: plain pointers and sizes, integer
 * status codes, no exceptions or aborts across the boundary.
 *
 * Reference interfaces replaced (marigold, file:line):
 *   mgbvh_build_tris    BoundingVolumeHierarchy::<1,1>::new(&mut [MeshTriangle],
 *                       &[MeshVertex], BvhSettings)          bvh/src/lib.rs:698-878
 *                       (T = MeshTriangle, S = MeshVertex,   mesh_interface/src/lib.rs:19-30;
 *                       caller marigold/src/app/data/scene/model.rs:93)
 *   mgbvh_build_aabbs   BoundingVolumeHierarchy::<1,1>::new(&mut [T], &[()], ..)
 *                       for T: AsBoundingVolume, e.g. the TLAS over UploadedMesh
 *                       (mesh_interface/src/lib.rs:87-91; caller
 *                       marigold/src/app/data/scene/geometry/mod.rs:196)
 *   mgbvh_node32        AsGpuBytes for BvhNode, MAX_LEAF_OBJECTS == 1 branch
 *                       bvh/src/lib.rs:157-169 == scene.slang:261-266 (32 bytes)
 *   mgbvh_settings      BvhSettings                          bvh/src/lib.rs:668-674
 *   mgbvh_stats         BvhProfilingInfo                     bvh/src/lib.rs:676-689
 *   *_device variants   the same builds over device-resident buffers (no host
 *                       copies inside the timed region)
 */
#ifndef MGBVH_H
#define MGBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGBVH_ABI_VERSION 1

/* status codes (0 = ok) */
#define MGBVH_OK 0
#define MGBVH_ERR_INVALID_ARG -1   /* null pointer, bad stride, capacity too small */
#define MGBVH_ERR_UNSUPPORTED -2   /* e.g. MIN/MAX leaf objects other than 1/1 */
#define MGBVH_ERR_NONFINITE -3     /* an input coordinate is NaN/inf or |x| >= 2^60 */
#define MGBVH_ERR_INDEX -4         /* triangle index >= vertex count */
#define MGBVH_ERR_DEVICE -5        /* HIP runtime error (message in mgbvh_last_error) */
#define MGBVH_ERR_NO_DEVICE -6     /* no usable gfx950 device */
#define MGBVH_ERR_INTERNAL -7      /* builder invariant violated */

/* 32-byte GPU node, bvh/src/lib.rs:157-169 (MAX_LEAF_OBJECTS == 1):
 * a node is a leaf iff child_node == 0; a leaf's start_index indexes the
 * permuted object list; internal node children are child_node (gt) and
 * child_node + 1 (lt). */
typedef struct mgbvh_node32 {
  float min[3];
  uint32_t start_index;
  float max[3];
  uint32_t child_node;
} mgbvh_node32;

/* An object's axis-aligned box (BoundingVolume, bvh/src/lib.rs:41-46). */
typedef struct mgbvh_aabb {
  float min[3];
  float max[3];
} mgbvh_aabb;

/* Split decision of one internal node (indexed by node index):
 * axis_method = axis | method << 8, method 0 = binned (lib.rs:331-389),
 * 1 = adaptive threshold (lib.rs:391-442), 2 = median (lib.rs:444-491);
 * value = bin cut i (binned) | threshold bits (adaptive) | m (median);
 * cost_bits = the fp32 SAH cost bits of the chosen split. */
typedef struct mgbvh_split_rec {
  uint32_t axis_method;
  uint32_t value;
  uint32_t count_lt;
  uint32_t cost_bits;
} mgbvh_split_rec;

#define MGBVH_METHOD_PLAIN 0   /* the reference algorithm, bit-exact */
#define MGBVH_METHOD_LEARNED 1 /* learned split-window proposer + guard (needs a model,
                                  mgbvh_ctx_set_model) */

/* BvhSettings (bvh/src/lib.rs:668-674) plus the build method. */
typedef struct mgbvh_settings {
  const char* name;                     /* BvhSettings.name (may be NULL) */
  int32_t has_bounds;                   /* BvhSettings.bounds is Some(..) */
  float bounds_min[3];
  float bounds_max[3];
  uint32_t max_depth;                   /* BvhSettings.max_depth */
  int32_t profiling_info;               /* BvhSettings.profiling_info */
  const char* profiling_info_directory; /* BvhSettings.profiling_info_directory (NULL = None) */
  uint32_t min_leaf_objects;            /* const generic MIN_LEAF_OBJECTS (must be 1) */
  uint32_t max_leaf_objects;            /* const generic MAX_LEAF_OBJECTS (must be 1) */
  uint32_t method;                      /* MGBVH_METHOD_* */
  /* MGBVH_METHOD_LEARNED only (0 = default): nodes with at least
   * learned_min_objects objects take the learned path; the window spans
   * learned_window cuts either side of the predicted cut; a window whose best
   * proxy SAH exceeds (1 + learned_guard) x the best proxy SAH over all cuts
   * falls back to the plain binned sweep. */
  uint32_t learned_min_objects;
  uint32_t learned_window;
  float learned_guard;
  uint32_t learned_no_network;          /* ablation: centre each window on the best proxy
                                           cut instead of the network's choice */
  uint32_t reserved[3];
} mgbvh_settings;

/* BvhProfilingInfo (bvh/src/lib.rs:676-689) plus builder counters. */
typedef struct mgbvh_stats {
  double construction_time;   /* seconds, device time of the build */
  uint32_t node_count;
  uint32_t leaf_node_count;
  uint32_t min_leaf_object_count;
  uint32_t max_leaf_object_count;
  double avg_leaf_object_count;
  double stddev_leaf_object_count;
  uint32_t total_object_count;
  uint32_t height;
  uint32_t max_depth;
  uint32_t n_large_levels;    /* breadth-first levels run over nodes with > 32 objects */
  uint32_t n_median;          /* median fallback splits */
  uint32_t n_cutoff_leaves;   /* leaves forced by max_depth holding > 1 object */
  uint32_t n_guard_fallbacks; /* learned mode: windows rejected by the exact guard */
  uint64_t large_object_visits; /* sum of object counts of the nodes with > 32 objects
                                   (the binned levels read and move each such object once) */
  uint32_t large_nodes;       /* nodes with > 32 objects (binned sweeps) */
  uint32_t n_window_splits;   /* learned mode: nodes split inside their predicted window */
  uint32_t reserved[1];
} mgbvh_stats;

typedef struct mgbvh_ctx mgbvh_ctx;

/* Fills *s with BvhSettings defaults of the marigold app (max_depth 32, MIN=MAX=1). */
void mgbvh_settings_default(mgbvh_settings* s);

/* Creates a builder context on HIP device `device` (owns a stream and scratch). */
int mgbvh_ctx_create(int device, mgbvh_ctx** out);
int mgbvh_ctx_destroy(mgbvh_ctx* ctx);
/* Thread-local message of the last failing call on this thread. */
const char* mgbvh_last_error(void);
/* Stream the context enqueues on (hipStream_t as void*). */
void* mgbvh_ctx_stream(mgbvh_ctx* ctx);
int mgbvh_ctx_synchronize(mgbvh_ctx* ctx);

/* ---- host-buffer entry points (drop-in for BoundingVolumeHierarchy::new) ----
 * positions: vertex positions, xyz floats at `positions + k * vstride_bytes`.
 * tris: n x 3 u32 vertex indices, PERMUTED IN PLACE like `list: &mut [T]`.
 * out_nodes: capacity 2n-1; *out_nnodes receives the node count.
 * out_lens (optional): BvhNode.len per node.  out_perm (optional): original
 * object index at each position of the permuted list.  out_splits (optional,
 * capacity 2n-1): split record per internal node.  out_stats optional. */
int mgbvh_build_tris(mgbvh_ctx* ctx, const float* positions, uint32_t vstride_bytes,
                     uint64_t nverts, uint32_t* tris, uint64_t ntris,
                     const mgbvh_settings* settings, mgbvh_node32* out_nodes,
                     uint32_t* out_lens, uint64_t* out_nnodes, uint32_t* out_perm,
                     mgbvh_split_rec* out_splits, mgbvh_stats* out_stats);

/* Objects given by their boxes (T: AsBoundingVolume).  The caller applies
 * out_perm (required) to its own object list, as the crate permutes `list`. */
int mgbvh_build_aabbs(mgbvh_ctx* ctx, const mgbvh_aabb* aabbs, uint64_t n,
                      const mgbvh_settings* settings, mgbvh_node32* out_nodes,
                      uint32_t* out_lens, uint64_t* out_nnodes, uint32_t* out_perm,
                      mgbvh_split_rec* out_splits, mgbvh_stats* out_stats);

/* ---- device-buffer entry points (all pointers are device pointers) ----
 * Outputs: d_nodes (2n-1), d_perm (n); d_tris_out (optional, n x 3) receives the
 * permuted triangle list; d_lens and d_splits optional.  Enqueued on the
 * context stream; synchronizes only to read the node count. */
int mgbvh_build_tris_device(mgbvh_ctx* ctx, const float* d_positions, uint32_t vstride_bytes,
                            uint64_t nverts, const uint32_t* d_tris, uint64_t ntris,
                            const mgbvh_settings* settings, mgbvh_node32* d_nodes,
                            uint32_t* d_lens, uint32_t* d_perm, uint32_t* d_tris_out,
                            mgbvh_split_rec* d_splits, uint64_t* out_nnodes,
                            mgbvh_stats* out_stats);

/* Allocation-free variant of mgbvh_build_tris_device.  It has no counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * scratch buffer of the build lives in one static device array per GPU
 * (about 24 MB, held until the process ends), so the call allocates no device
 * or pinned memory and creates no HIP object.  At most 65536 triangles
 * (MGBVH_ERR_INVALID_ARG beyond), plain method only, no profiling JSON and no
 * statistics.  All calls on one GPU share that array: calls that overlap on
 * one GPU race.  Same arguments and outputs as mgbvh_build_tris_device. */
int mgbvh_build_tris_device_noalloc(mgbvh_ctx* ctx, const float* d_positions,
                                    uint32_t vstride_bytes, uint64_t nverts,
                                    const uint32_t* d_tris, uint64_t ntris,
                                    const mgbvh_settings* settings, mgbvh_node32* d_nodes,
                                    uint32_t* d_lens, uint32_t* d_perm, uint32_t* d_tris_out,
                                    mgbvh_split_rec* d_splits, uint64_t* out_nnodes);

int mgbvh_build_aabbs_device(mgbvh_ctx* ctx, const mgbvh_aabb* d_aabbs, uint64_t n,
                             const mgbvh_settings* settings, mgbvh_node32* d_nodes,
                             uint32_t* d_lens, uint32_t* d_perm, mgbvh_split_rec* d_splits,
                             uint64_t* out_nnodes, mgbvh_stats* out_stats);

/* ---- device memory helpers (so callers need no other GPU runtime) ---- */
int mgbvh_device_alloc(mgbvh_ctx* ctx, uint64_t bytes, void** out);
int mgbvh_device_free(mgbvh_ctx* ctx, void* p);
int mgbvh_copy_h2d(mgbvh_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int mgbvh_copy_d2h(mgbvh_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy on the context stream (also the known-size transfer
 * the traffic counters are calibrated with). */
int mgbvh_copy_d2d(mgbvh_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- per-kernel timing: HIP event pairs around each launch of the kernel
 * classes selected by `kernel_mask` (bit k = class k), accumulated per class.
 * Classes: 0 setup, 1 root, 2 bin, 3 sweep, 4 partition, 5 small-subtree,
 * 6 compact, 7 permute, 8 trace. ---- */
#define MGBVH_KERNEL_KINDS 16
int mgbvh_ctx_set_profiling(mgbvh_ctx* ctx, uint32_t kernel_mask);
int mgbvh_ctx_kernel_times(mgbvh_ctx* ctx, double* ms /* [16] */, uint64_t* launches /* [16] */,
                           int reset);

/* ---- closest-hit traversal --------------------------------------------------
 * Replaces the GPU traversal of the path tracer, scene.slang:90-221
 * (raytraceTlas / raytraceBlas) with the triangle test of scene.slang:19-89.
 * Hit rules (tree independent, so any valid tree of the same objects gives the
 * same hits): Moller-Trumbore in the reference's operation order; closest t;
 * equal t goes to the lowest (instance, triangle) original index; the box test
 * never culls a box that holds the closest hit.  A leaf holds lens[v] objects
 * (d_lens NULL: one, as the 32-byte node format stores no length). */
typedef struct mgbvh_ray {
  float org[3];
  uint32_t reserved;   /* must be 0 */
  float dir[3];
  float tmax;          /* hits with t > tmax are ignored (+inf: none) */
} mgbvh_ray;

typedef struct mgbvh_hit {
  float t;             /* tmax when there is no hit */
  uint32_t prim;       /* original triangle index in its mesh, 0xffffffff = miss */
  uint32_t inst;       /* original instance index (0 for a single BLAS), 0xffffffff = miss */
  uint32_t tri_slot;   /* position of the hit triangle in the permuted (concatenated)
                          triangle list, the index the reference's shader uses */
} mgbvh_hit;

/* One BLAS inside concatenated arrays (the fields of Mesh, scene.slang:245-254). */
typedef struct mgbvh_blas_desc {
  uint32_t node_offset;    /* first node (Mesh.blasRoot); child indices are relative */
  uint32_t tri_offset;     /* first triangle of the permuted list (Mesh.triangleOffset) */
  uint32_t vertex_offset;  /* added to triangle vertex indices (Mesh.vertexOffset) */
  uint32_t reserved;
} mgbvh_blas_desc;

/* One TLAS object: a BLAS placed by an affine transform (Mesh.transform).
 * world_to_object = Mesh.inverseTransform as 3 rows x 4 floats; the ray is
 * taken to object space unnormalised, so t is the same parameter in both. */
typedef struct mgbvh_instance {
  float world_to_object[12];
  uint32_t blas;
  uint32_t reserved[3];
} mgbvh_instance;

/* Pinhole camera of mgbvh_gen_primary_rays_device (position, orthonormal basis). */
typedef struct mgbvh_camera {
  float pos[3];
  float fwd[3];
  float right[3];
  float up[3];
  float fov_y;   /* radians */
  float aspect;  /* width / height */
} mgbvh_camera;

/* Closest hit of each ray in one BLAS.  d_tris_perm / d_perm: the permuted
 * triangle list and the original index at each position (build outputs). */
int mgbvh_trace_closest_device(mgbvh_ctx* ctx, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                               const float* d_positions, uint32_t vstride_bytes,
                               const uint32_t* d_tris_perm, const uint32_t* d_perm,
                               const mgbvh_ray* d_rays, uint64_t nrays, mgbvh_hit* d_hits);

/* Closest hit of each ray in a two-level scene.  d_instances is in the
 * caller's original order; d_tlas_perm (the TLAS build's perm output) maps a
 * TLAS leaf's start_index to an instance, so the instance list is never
 * reordered.  d_tlas_lens optional.  BLASes are concatenated as d_blas says:
 * node, lens, permuted-triangle and perm arrays share the offsets. */
int mgbvh_trace_scene_device(mgbvh_ctx* ctx, const mgbvh_node32* d_tlas_nodes,
                             const uint32_t* d_tlas_lens, const uint32_t* d_tlas_perm,
                             const mgbvh_instance* d_instances, uint32_t ninstances,
                             const mgbvh_blas_desc* d_blas, uint32_t nblas,
                             const mgbvh_node32* d_blas_nodes, const uint32_t* d_blas_lens,
                             const float* d_positions, uint32_t vstride_bytes,
                             const uint32_t* d_tris_perm, const uint32_t* d_blas_perm,
                             const mgbvh_ray* d_rays, uint64_t nrays, mgbvh_hit* d_hits);

/* The path tracer's own closest-hit traversal, step for step: replaces
 * Scene.raytraceTlas / raytraceBlas (scene.slang:91-221) with the box test of
 * scene.slang:296-310, the triangle test of :19-89 and Hit.merge of
 * intersect.slang:73-82.  Unlike mgbvh_trace_scene_device the result depends
 * on the tree, as the shader's does: one 64-entry stack shared by both levels
 * (BVH_STACK_SIZE, scene.slang:5-7; an overflow returns MGBVH_ERR_INTERNAL
 * where the shader would write past it); children pushed when their box
 * distance is strictly below the closest distance, nearer on top; no cull on
 * pop; the BLAS root pushed without a box test; the BLAS pass starts from the
 * world distance times length(object-space direction) (:165-166) and returns
 * its distance unconverted (:211-213); an equal distance goes to the hit
 * tested later.  Every op is IEEE fp32 round-to-nearest without contraction;
 * min/max of a NaN take the other operand.  One object per leaf (the shader's
 * node has no length); every instance's BLAS must be non-empty (the shader
 * reads its root unconditionally).  ray.tmax is ignored (the shader has none).
 * Per ray: t = Hit.distance (0 for NO_HIT); prim, inst, tri_slot as in
 * mgbvh_hit, 0xffffffff on a miss; tri_slot 0xfffffffe for a success that
 * holds no triangle (a merge that took NO_HIT, only when a distance reaches
 * float.maxValue). */
int mgbvh_raytrace_tlas_device(mgbvh_ctx* ctx, const mgbvh_node32* d_tlas_nodes,
                               const uint32_t* d_tlas_perm, const mgbvh_instance* d_instances,
                               uint32_t ninstances, const mgbvh_blas_desc* d_blas, uint32_t nblas,
                               const mgbvh_node32* d_blas_nodes, const float* d_positions,
                               uint32_t vstride_bytes, const uint32_t* d_tris_perm,
                               const uint32_t* d_blas_perm, const mgbvh_ray* d_rays,
                               uint64_t nrays, mgbvh_hit* d_hits);

/* Primary rays of a width x height image, row-major from the top-left pixel,
 * as the path tracer makes them (pathtrace.slang:46-66, screen.slang:48-72):
 * texcoord = (pixel + jitter) / (width, height); clip = (2 tx - 1, 1 - 2 ty, 1);
 * world = (M clip).xyz / (M clip).w; dir = normalize(world - position);
 * org = position.  inv_view_proj: M, the inverse view-projection matrix, 16
 * floats column-major (glam's Mat4 layout, camera.rs:355-362).  The jitter is
 * the shader's: two draws of the PCG stream seeded with frame_count, minus 0.5,
 * when accumulate is non-zero (shouldAccumulate), else 0.  Each operation is
 * rounded once in the order written (no FMA, correctly rounded division and
 * square root), so the rays are reproducible bit for bit. */
int mgbvh_gen_view_rays_device(mgbvh_ctx* ctx, const float* inv_view_proj, const float* position,
                               uint32_t width, uint32_t height, uint32_t frame_count,
                               int accumulate, mgbvh_ray* d_rays);

/* The same rays for a pinhole camera: M from glam's look_at_rh and
 * perspective_rh (depth 0..1, near 0.1, far 100) of the camera, computed in
 * double, and a jitter of +0.5, so every ray passes through its pixel's centre. */
int mgbvh_gen_primary_rays_device(mgbvh_ctx* ctx, const mgbvh_camera* cam, uint32_t width,
                                  uint32_t height, mgbvh_ray* d_rays);

/* One cosine-weighted bounce per ray (pathtrace.slang:49, random.slang:5-41,
 * intersect.slang:49-51): from the hit point offset 0.001 along the world
 * normal facing the ray, direction normalize(normal + random unit vector); a
 * missed ray is re-cast from its origin in a random direction.  The triangle
 * is d_tris_perm[hit.tri_slot]; for a scene pass the instances and BLAS table
 * (vertex offset, normal taken to world space), else NULL and 0.  seed: the
 * frame offset added to the ray index to seed its PCG stream. */
int mgbvh_gen_bounce_rays_device(mgbvh_ctx* ctx, const mgbvh_ray* d_rays_in,
                                 const mgbvh_hit* d_hits, uint64_t nrays,
                                 const mgbvh_instance* d_instances, uint32_t ninstances,
                                 const mgbvh_blas_desc* d_blas, uint32_t nblas,
                                 const float* d_positions, uint32_t vstride_bytes,
                                 const uint32_t* d_tris_perm, uint32_t seed,
                                 mgbvh_ray* d_rays_out);

/* ---- compressed 8-wide BVH ----------------------------------------------------
 * No counterpart in the reference's code: its README (README.md, "A BVH8
 * implementation is planned, following Efficient Incoherent Ray Traversal on
 * GPUs Through Compressed Wide BVHs (Ylitie et al.)") names it as the step
 * after the binary builder, which keeps one object per leaf for it.  The
 * collapse is that paper's dynamic program: C(v, i) = cost of v's subtree as
 * at most i subtrees; a wide leaf of p <= p_max objects costs
 * area * p * c_prim, a wide internal node area * c_node plus the best split
 * of its 8 slots; first minimum on ties, a leaf on a leaf/internal tie; area
 * = 2 * ((dx*dy + dy*dz) + dz*dx) of the binary node's box, all fp32. */
typedef struct mgbvh_wide8_params {
  float c_node;      /* default 1.0; finite and >= 0 (else MGBVH_ERR_INVALID_ARG) */
  float c_prim;      /* default 0.3; finite and >= 0 */
  uint32_t p_max;    /* objects per wide leaf, 1..3 (default 3) */
  uint32_t reserved;
} mgbvh_wide8_params;

/* 80-byte node.  Child k's box is origin + q * 2^(exponent - 127) per axis in
 * fp32 (q = qlo or qhi), conservative for the binary node it came from.
 * meta[k] = 0: empty slot; internal child (imask bit k): 0x20 | (24 + r), the
 * child is wide node base_child + r; leaf child: (unary object count << 5) |
 * offset, its objects are tri_index[base_tri + offset ...].  Layout: root at
 * 0; a node's internal children contiguous, followed by their own subtrees in
 * child order; a node's leaf objects contiguous, followed by its children's. */
typedef struct mgbvh_wide_node {
  float origin[3];
  uint8_t exponent[3];
  uint8_t imask;
  uint32_t base_child;
  uint32_t base_tri;
  uint8_t meta[8];
  uint8_t qlo[3][8];
  uint8_t qhi[3][8];
} mgbvh_wide_node;

/* Collapses a binary tree (device nodes from any build here; d_lens optional,
 * leaves of more than p_max objects are an error).  d_wide: capacity wide_cap
 * (objects - 1 is always enough, at least 1); d_tri_index: capacity tri_cap
 * (the object count), receives the binary leaves' start_index values in wide
 * leaf order.  out_cost (optional): the root's collapse cost.  params NULL =
 * defaults.  Synchronises the context stream. */
int mgbvh_wide8_collapse_device(mgbvh_ctx* ctx, const mgbvh_node32* d_nodes,
                                const uint32_t* d_lens, uint64_t nnodes,
                                const mgbvh_wide8_params* params, mgbvh_wide_node* d_wide,
                                uint64_t wide_cap, uint32_t* d_tri_index, uint64_t tri_cap,
                                uint64_t* out_nwide, float* out_cost);

/* Every BLAS of a batched build (mgbvh_build_blas_batch_device outputs: BLAS m's
 * nodes at blas[m].node_offset, node_counts[m] of them, child indices relative
 * to that offset) collapsed in one call.  Each BLAS gets the tree
 * mgbvh_wide8_collapse_device gives its nodes alone, at d_wide +
 * out_wide_first[m] (out_wide_count[m] nodes, 0 for an empty BLAS; base_child
 * relative to its first wide node), its tri_index at blas[m].tri_offset
 * (entries relative to the BLAS's permuted triangle list).  Host arrays:
 * blas, node_counts, out_wide_first, out_wide_count (nblas each).  wide_cap:
 * total objects is always enough. */
int mgbvh_wide8_collapse_batch_device(mgbvh_ctx* ctx, const mgbvh_node32* d_nodes,
                                      const uint32_t* d_lens, uint64_t nnodes,
                                      const mgbvh_blas_desc* blas, const uint32_t* node_counts,
                                      uint32_t nblas, const mgbvh_wide8_params* params,
                                      mgbvh_wide_node* d_wide, uint64_t wide_cap,
                                      uint32_t* d_tri_index, uint64_t tri_cap,
                                      uint32_t* out_wide_first, uint32_t* out_wide_count,
                                      uint64_t* out_nwide);

/* Closest hit of each ray through a wide tree of one BLAS, with the hit rules
 * of mgbvh_trace_closest_device (the same hits as the binary tree). */
int mgbvh_trace_wide8_device(mgbvh_ctx* ctx, const mgbvh_wide_node* d_wide, uint64_t nwide,
                             const uint32_t* d_tri_index, const float* d_positions,
                             uint32_t vstride_bytes, const uint32_t* d_tris_perm,
                             const uint32_t* d_perm, const mgbvh_ray* d_rays, uint64_t nrays,
                             mgbvh_hit* d_hits);

/* ---- batched BLAS build ------------------------------------------------------
 * Replaces the per-mesh BLAS loop of the scene upload (model.rs:76-95, one
 * BoundingVolumeHierarchy::new per mesh on rayon workers): every mesh of the
 * concatenated triangle list gets the tree mgbvh_build_tris_device would give
 * it alone, but all meshes share each breadth-first level's launches. */
typedef struct mgbvh_mesh_desc {
  uint32_t tri_offset;     /* first triangle = sum of tri_count of the earlier meshes */
  uint32_t tri_count;      /* 0 = empty mesh (no nodes, lib.rs:703-708) */
  uint32_t vertex_offset;  /* the mesh's vertex indices address vertex_offset + index */
  uint32_t vertex_count;   /* every vertex index of the mesh is < vertex_count */
} mgbvh_mesh_desc;

/* meshes (HOST array, nmeshes entries) describe d_tris (concatenated, indices
 * local to each mesh).  mesh_bounds (HOST, optional) = BvhSettings.bounds per
 * mesh (NULL: the union of the mesh's triangles); settings.has_bounds and
 * settings.profiling_info must be 0.  Mesh m's nodes start at slot
 * out_blas[m].node_offset = sum over earlier meshes of 2 n - 1 (d_nodes, d_lens
 * and d_splits hold sum(2 n - 1) slots); its child indices are relative to
 * that slot (Mesh.blasRoot, scene.slang:182), its leaf start indices and
 * d_perm entries relative to tri_offset.  out_node_counts (HOST, optional)
 * receives each mesh's node count (2 n - 1 unless max_depth cut it short).
 * out_stats (optional): construction_time, node_count, total_object_count,
 * n_large_levels, n_median, n_cutoff_leaves summed over meshes; others 0. */
int mgbvh_build_blas_batch_device(mgbvh_ctx* ctx, const float* d_positions,
                                  uint32_t vstride_bytes, uint64_t nverts, const uint32_t* d_tris,
                                  const mgbvh_mesh_desc* meshes, uint32_t nmeshes,
                                  const mgbvh_aabb* mesh_bounds, const mgbvh_settings* settings,
                                  mgbvh_node32* d_nodes, uint32_t* d_lens, uint32_t* d_perm,
                                  uint32_t* d_tris_out, mgbvh_split_rec* d_splits,
                                  mgbvh_blas_desc* out_blas, uint32_t* out_node_counts,
                                  mgbvh_stats* out_stats);

/* World boxes of instances, the TLAS objects (model.rs:189-221): the 8
 * corners of the instance's BLAS root box through its object_to_world
 * transform (3 rows x 4 floats per instance, each row summed left to right as
 * glam's Mat4 * Vec4), then the component-wise min and max. */
int mgbvh_instance_boxes_device(mgbvh_ctx* ctx, const mgbvh_node32* d_blas_nodes,
                                const mgbvh_blas_desc* d_blas, uint32_t nblas,
                                const float* d_object_to_world, const uint32_t* d_instance_blas,
                                uint32_t ninstances, mgbvh_aabb* d_boxes);

/* ---- learned split proposer ---------------------------------------------------
 * No counterpart in the reference: the network of MGBVH_METHOD_LEARNED, a
 * 102 -> 128 -> 128 -> 96 MLP (ReLU) from a node's centroid-bin counts and
 * extents to split logits (axis * 32 + cut).  Features of a node with n > 32
 * objects: log1p(count) / log1p(n) of the 3 x 32 centroid bins (axis-major),
 * node extent / largest node extent (3), centroid extent / node extent (3).
 * Parameters (floats): W1[128][128] (inputs 102..127 zero), b1[128],
 * W2[128][128], b2[128], W3[96][128], b3[96].  Training runs in FP32 on the
 * device (softmax cross-entropy, Adam); inference in BF16 on MFMA with FP32
 * accumulation. */
#define MGBVH_MODEL_FEATURES 102
#define MGBVH_MODEL_CLASSES 96
#define MGBVH_MODEL_PARAMS 45408

typedef struct mgbvh_model mgbvh_model;

/* Kaiming-uniform weights from `seed`, zero biases. */
int mgbvh_model_create(mgbvh_ctx* ctx, uint64_t seed, mgbvh_model** out);
int mgbvh_model_destroy(mgbvh_model* model);
int mgbvh_model_get_params(mgbvh_ctx* ctx, mgbvh_model* model, float* host_params);
int mgbvh_model_set_params(mgbvh_ctx* ctx, mgbvh_model* model, const float* host_params);

/* Split samples of a plain build: the features and the chosen binned split
 * (axis * 32 + cut) of every node with >= min_objects objects split by the
 * binned sweep.  d_features: max_samples x 102 floats. */
int mgbvh_collect_samples_device(mgbvh_ctx* ctx, const float* d_positions, uint32_t vstride_bytes,
                                 uint64_t nverts, const uint32_t* d_tris, uint64_t ntris,
                                 const mgbvh_settings* settings, uint32_t min_objects,
                                 float* d_features, uint32_t* d_labels, uint32_t max_samples,
                                 uint32_t* out_count);

/* Mean cross-entropy over n samples and its gradient (MGBVH_MODEL_PARAMS
 * floats, parameter layout above) at the current parameters; no update. */
int mgbvh_model_gradients_device(mgbvh_ctx* ctx, mgbvh_model* model, const float* d_features,
                                 const uint32_t* d_labels, uint32_t n, float* d_grads,
                                 float* out_loss);

/* Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled minibatches; returns the
 * last epoch's mean loss and accuracy. */
int mgbvh_model_train_device(mgbvh_ctx* ctx, mgbvh_model* model, const float* d_features,
                             const uint32_t* d_labels, uint32_t n, uint32_t epochs, uint32_t batch,
                             float lr, uint64_t seed, float* out_loss, float* out_accuracy);

/* The network MGBVH_METHOD_LEARNED builds on this context use (NULL: none). */
int mgbvh_ctx_set_model(mgbvh_ctx* ctx, mgbvh_model* model);

/* Logits (n x 96) by the BF16 MFMA path the learned builder uses. */
int mgbvh_model_infer_device(mgbvh_ctx* ctx, mgbvh_model* model, const float* d_features,
                             uint32_t n, float* d_logits);

/* Surface-area cost of each BLAS with the reference's constants
 * (bvh/src/lib.rs:217-218): sum over internal nodes of SA(n) / SA(root) * 1
 * plus sum over leaves of SA(l) / SA(root) * 8 * len(l), in double.  blas and
 * node_counts are HOST arrays (node_offset of each BLAS and its node count). */
int mgbvh_tree_sah_device(mgbvh_ctx* ctx, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                          const mgbvh_blas_desc* blas, const uint32_t* node_counts, uint32_t nblas,
                          double* out_sah);

/* ---- scene plans: a whole scene build as one HIP graph ----------------------
 * No counterpart in the reference (it rebuilds with rayon, model.rs:76-95,
 * geometry/mod.rs:175-199).  mgbvh_scene_plan_create records every BLAS
 * level of a batched build (persistent level kernels over device-side work
 * queues, max_depth levels), the instance boxes and the TLAS build as one
 * graph over buffers the plan owns; mgbvh_scene_plan_launch replays it (new
 * vertex positions in the same buffers, same mesh sizes), waits, checks
 * errors and squeezes depth-cut gaps, and returns the graph's device time.
 * Outputs as mgbvh_build_blas_batch_device, mgbvh_instance_boxes_device and
 * mgbvh_build_aabbs_device.  max_depth <= 64; the TLAS uses the plain method;
 * a learned BLAS method uses the context's model as it was at creation. */
typedef struct mgbvh_scene_plan mgbvh_scene_plan;
int mgbvh_scene_plan_create(mgbvh_ctx* ctx, const float* d_positions, uint32_t vstride_bytes,
                            uint64_t nverts, const uint32_t* d_tris,
                            const mgbvh_mesh_desc* meshes, uint32_t nmeshes,
                            const mgbvh_settings* blas_settings, mgbvh_node32* d_blas_nodes,
                            uint32_t* d_blas_lens, uint32_t* d_blas_perm, uint32_t* d_tris_out,
                            const float* d_object_to_world, const uint32_t* d_instance_blas,
                            uint32_t ninstances, const mgbvh_settings* tlas_settings,
                            mgbvh_aabb* d_boxes, mgbvh_node32* d_tlas_nodes,
                            uint32_t* d_tlas_lens, uint32_t* d_tlas_perm,
                            mgbvh_scene_plan** out);
int mgbvh_scene_plan_launch(mgbvh_scene_plan* plan, float* out_ms,
                            uint32_t* out_blas_node_counts, uint64_t* out_tlas_nodes);
int mgbvh_scene_plan_destroy(mgbvh_scene_plan* plan);
/* Levels the plan's graph runs for the BLASes and the TLAS (sized from a
 * level-by-level build at creation, plus a margin of 2), and how many launches
 * so far found a tree deeper than that and were rebuilt level by level (same
 * results, slower launch). */
int mgbvh_scene_plan_info(const mgbvh_scene_plan* plan, uint32_t* blas_levels,
                          uint32_t* tlas_levels, uint32_t* launches_rebuilt);

/* Grid sizes of the level kernels (blocks per compute unit, 1..256; the
 * kernels loop over their work lists, so any size is correct), the knobs the
 * autotuner (marigold_amd/autotune.py) sets; results never depend on them. */
typedef struct mgbvh_tuning {
  uint32_t bin_blocks_per_cu;
  uint32_t partition_blocks_per_cu;
  uint32_t sweep_blocks_per_cu;
  uint32_t small_blocks_per_cu;
  uint32_t compute_units;  /* read-only (get) */
  uint32_t partition_waves; /* wavefronts per partition block below the first 4 levels
                              (those always take 4): 1 or 4, default 4 (0 = keep) */
  uint32_t bin_waves;       /* wavefronts per binning block: 1..4 (0 = keep) */
  uint32_t sweep_pack_min_nodes; /* levels with at least this many nodes are swept one node
                                    per lane, the others one node per wavefront
                                    (0 = keep; 1 = every level) */
  uint32_t medium_max_objects;   /* nodes of 33..this many objects build their whole subtree
                                    in one wavefront instead of level by level: 33..2048;
                                    1 = off; 0 = keep */
} mgbvh_tuning;
int mgbvh_ctx_set_tuning(mgbvh_ctx* ctx, const mgbvh_tuning* tuning);
int mgbvh_ctx_get_tuning(mgbvh_ctx* ctx, mgbvh_tuning* tuning);

/* ---- timing helpers: HIP events recorded on the context stream ---- */
int mgbvh_event_record(mgbvh_ctx* ctx, int slot);           /* slot 0..59 */
int mgbvh_event_elapsed_ms(mgbvh_ctx* ctx, int a, int b, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* MGBVH_H */
