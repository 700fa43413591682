// Scene-level pieces around the builder: instance world boxes for the TLAS.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "mgbvh_common.h"
#include "mgbvh_internal.h"

namespace mg {

// model.rs:189-221: the 8 corners of the mesh box through the instance
// transform (glam Mat4 * Vec4: x_axis*x + y_axis*y + z_axis*z + w_axis*w,
// summed left to right, w = 1), then the component-wise min and max.  The mesh
// box is the BLAS root's bounds (= mesh.bounds when the BLAS was built with
// them).  object_to_world: 3 rows x 4 floats per instance.
__device__ inline mgbvh_aabb instance_box(const float (&bmin)[3], const float (&bmax)[3],
                                          const float* __restrict__ m) {
  float lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = __uint_as_float(0x7f800000u);
    hi[a] = __uint_as_float(0xff800000u);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float x = (k & 4) ? bmax[0] : bmin[0];
    const float y = (k & 2) ? bmax[1] : bmin[1];
    const float z = (k & 1) ? bmax[2] : bmin[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float* row = m + 4 * a;
      float w = __fmul_rn(row[0], x);
      w = __fadd_rn(w, __fmul_rn(row[1], y));
      w = __fadd_rn(w, __fmul_rn(row[2], z));
      w = __fadd_rn(w, row[3]);
      lo[a] = fmin_x86(lo[a], w);
      hi[a] = fmax_x86(hi[a], w);
    }
  }
  mgbvh_aabb r;
#pragma unroll
  for (int a = 0; a < 3; ++a) { r.min[a] = lo[a]; r.max[a] = hi[a]; }
  return r;
}

// The mesh box from each BLAS's root node.
__global__ __launch_bounds__(256) void k_instance_boxes(const mgbvh_node32* __restrict__ nodes,
                                                        const mgbvh_blas_desc* __restrict__ blas,
                                                        uint32_t nblas,
                                                        const float* __restrict__ o2w,
                                                        const uint32_t* __restrict__ inst_blas,
                                                        uint32_t ninst, mgbvh_aabb* __restrict__ out,
                                                        uint32_t* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  const uint32_t b = inst_blas[i];
  if (b >= nblas) {
    atomicOr(err, 1u);
    return;
  }
  const mgbvh_node32 root = nodes[blas[b].node_offset];
  float bmin[3], bmax[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) { bmin[a] = root.min[a]; bmax[a] = root.max[a]; }
  out[i] = instance_box(bmin, bmax, o2w + 12 * (size_t)i);
}

// The mesh box from the batched build's per-mesh root bounds (the union of
// the mesh's object boxes, the root node's box of a non-empty mesh built
// without given bounds), available as soon as the objects are set up.
__global__ __launch_bounds__(256) void k_instance_boxes_acc(const uint32_t* __restrict__ macc,
                                                            uint32_t nblas,
                                                            const float* __restrict__ o2w,
                                                            const uint32_t* __restrict__ inst_blas,
                                                            uint32_t ninst,
                                                            mgbvh_aabb* __restrict__ out,
                                                            uint32_t* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  const uint32_t b = inst_blas[i];
  if (b >= nblas) {
    atomicOr(err, 1u);
    return;
  }
  const uint32_t* acc = macc + 12 * (size_t)b;
  float bmin[3], bmax[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    bmin[a] = acc_min_dec(acc[a]);
    bmax[a] = acc_max_dec(acc[3 + a]);
  }
  out[i] = instance_box(bmin, bmax, o2w + 12 * (size_t)i);
}

// SAH cost terms of every node slot: mesh m owns slots [off[m], off[m] + cnt[m]).
__global__ __launch_bounds__(256) void k_tree_sah(const mgbvh_node32* __restrict__ nodes,
                                                  const uint32_t* __restrict__ lens,
                                                  const uint32_t* __restrict__ off,
                                                  const uint32_t* __restrict__ cnt, uint32_t nm,
                                                  uint32_t nslots, double* __restrict__ sums) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  uint32_t lo = 0, hi = nm;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  if (i >= off[lo] + cnt[lo]) return;  // unused tail slot of a depth-cut mesh
  auto sa = [](const mgbvh_node32& n) {  // BoundingVolume::surface_area in f32, lib.rs:79-91
    Box b;
    for (int a = 0; a < 3; ++a) { b.mn[a] = n.min[a]; b.mx[a] = n.max[a]; }
    return (double)box_sa(b);
  };
  const mgbvh_node32 nd = nodes[i];
  const double root = sa(nodes[off[lo]]);
  const double w = nd.child_node ? 1.0 : 8.0 * (double)(lens ? lens[i] : 1u);
  double term = root > 0.0 ? sa(nd) / root * w : w;
  // one atomic per wave when the whole wave lies in one mesh (the common case)
  const uint32_t m0 = __shfl(lo, 0, 64);
  const uint64_t active = __ballot(true);
  if (__ballot(lo == m0) == active) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double v = __shfl_xor(term, o, 64);
      term += ((active >> (lane_id() ^ o)) & 1ull) ? v : 0.0;
    }
    if (lane_id() == (uint32_t)__builtin_ctzll(active)) atomicAdd(&sums[lo], term);
  } else {
    atomicAdd(&sums[lo], term);
  }
}

}  // namespace mg

using namespace mg;

extern "C" int mgbvh_tree_sah_device(mgbvh_ctx* c, const mgbvh_node32* d_nodes,
                                     const uint32_t* d_lens, const mgbvh_blas_desc* blas,
                                     const uint32_t* node_counts, uint32_t nblas,
                                     double* out_sah) {
  if (!c || !d_nodes || !blas || !node_counts || !out_sah) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (nblas == 0) return MGBVH_OK;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  std::vector<uint32_t> tab(2 * (size_t)nblas);
  uint64_t nslots = 0;
  for (uint32_t m = 0; m < nblas; ++m) {
    tab[m] = blas[m].node_offset;
    tab[nblas + m] = node_counts[m];
    if (m && blas[m].node_offset < blas[m - 1].node_offset) {
      set_error("BLAS node offsets must be ascending");
      return MGBVH_ERR_INVALID_ARG;
    }
    nslots = std::max<uint64_t>(nslots, (uint64_t)blas[m].node_offset + node_counts[m]);
  }
  if (nslots > 0xffffffffull) { set_error("too many node slots"); return MGBVH_ERR_INVALID_ARG; }
  int rc;
  if ((rc = ensure(c->sc.gtab, tab.size() * 4)) || (rc = ensure(c->sc.tmps, (size_t)nblas * 8))) return rc;
  hipStream_t st = c->stream;
  e = hipMemcpyAsync(c->sc.gtab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(c->sc.tmps.p, 0, (size_t)nblas * 8, st);
  if (e != hipSuccess) return hip_fail(e, "tree sah setup");
  const uint32_t* t = (const uint32_t*)c->sc.gtab.p;
  if (nslots)
    k_tree_sah<<<(uint32_t)((nslots + 255) / 256), 256, 0, st>>>(
        d_nodes, d_lens, t, t + nblas, nblas, (uint32_t)nslots, (double*)c->sc.tmps.p);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_sah, c->sc.tmps.p, (size_t)nblas * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_tree_sah");
}

extern "C" int mgbvh_instance_boxes_device(mgbvh_ctx* c, const mgbvh_node32* d_blas_nodes,
                                           const mgbvh_blas_desc* d_blas, uint32_t nblas,
                                           const float* d_object_to_world,
                                           const uint32_t* d_instance_blas, uint32_t ninst,
                                           mgbvh_aabb* d_boxes) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (ninst == 0) return MGBVH_OK;
  if (!d_blas_nodes || !d_blas || !d_object_to_world || !d_instance_blas || !d_boxes || !nblas) {
    set_error("null device buffer or no BLAS");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  int rc = ensure(c->sc.ctr, sizeof(Counters));
  if (rc) return rc;
  uint32_t* flag = reinterpret_cast<uint32_t*>(c->sc.ctr.p);
  e = hipMemsetAsync(flag, 0, 4, c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
  k_instance_boxes<<<(ninst + 255) / 256, 256, 0, c->stream>>>(
      d_blas_nodes, d_blas, nblas, d_object_to_world, d_instance_blas, ninst, d_boxes, flag);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&c->h_ctr->err, flag, 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(e, "k_instance_boxes");
  if (c->h_ctr->err) {
    set_error("an instance names a BLAS >= nblas");
    return MGBVH_ERR_INDEX;
  }
  return MGBVH_OK;
}

// ---------------------------------------------------------------- scene plan
// The whole build of a scene -- every BLAS level, the instance boxes and the
// TLAS -- recorded once as a HIP graph over buffers the plan owns, replayed for
// each rebuild (new vertex positions, same mesh sizes).  The level kernels are
// persistent over device-side work queues and run a fixed number of levels
// (max_depth: a node with > 32 objects has depth < max_depth), so nothing in
// the graph depends on counts read back by the host.
struct mgbvh_scene_plan {
  mgbvh_ctx* c = nullptr;
  mg::Scratch bsc, tsc;
  mg::BuildPlan bp, tp;
  DevBuf blas_desc, flag;
  const float* o2w = nullptr;
  const uint32_t* inst_blas = nullptr;
  uint32_t nblas = 0, ninst = 0;
  mgbvh_aabb* boxes = nullptr;
  mgbvh_node32* blas_nodes = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<uint32_t> blas_counts;
  uint64_t tlas_nodes = 0;
  uint32_t n_redo = 0;  // launches rebuilt level by level (trees deeper than recorded)
  mgbvh_model* model = nullptr;  // learned method: the model as it was at creation
};

namespace mg {
// The plan's scene build level by level (read-backs after every level): sizes
// the recorded graph, and redoes a replay whose trees needed more levels.
static int plan_levels(mgbvh_scene_plan* pl, BuildResult* rb, BuildResult* rt) {
  mgbvh_ctx* c = pl->c;
  hipStream_t st = c->stream;
  const uint32_t fb = pl->bp.fixed_levels, ft = pl->tp.fixed_levels;
  pl->bp.fixed_levels = pl->tp.fixed_levels = 0;
  uint32_t lv_run = 0;
  int rc = enqueue_build(c, pl->bsc, pl->bp, 0, false, &lv_run);
  if (!rc) rc = finish_build(c, pl->bsc, pl->bp, rb);
  if (!rc && pl->ninst) {
    k_fill<<<1, 64, 0, st>>>((uint32_t*)pl->flag.p, 0u, 1);
    k_instance_boxes<<<(pl->ninst + 255) / 256, 256, 0, st>>>(
        pl->blas_nodes, (const mgbvh_blas_desc*)pl->blas_desc.p, pl->nblas, pl->o2w,
        pl->inst_blas, pl->ninst, pl->boxes, (uint32_t*)pl->flag.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "k_instance_boxes");
    if (!rc) rc = enqueue_build(c, pl->tsc, pl->tp, 0, false, &lv_run);
    if (!rc) rc = finish_build(c, pl->tsc, pl->tp, rt);
  }
  pl->bp.fixed_levels = fb;
  pl->tp.fixed_levels = ft;
  return rc;
}

static void plan_release(mgbvh_scene_plan* pl) {
  if (!pl) return;
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  if (pl->e0) (void)hipEventDestroy(pl->e0);
  if (pl->e1) (void)hipEventDestroy(pl->e1);
  scratch_free(pl->bsc);
  scratch_free(pl->tsc);
  for (DevBuf* b : {&pl->blas_desc, &pl->flag})
    if (b->p) (void)hipFree(b->p);
  if (pl->bsc.h_ctr) (void)hipHostFree(pl->bsc.h_ctr);
  if (pl->tsc.h_ctr) (void)hipHostFree(pl->tsc.h_ctr);
  if (pl->model) (void)mgbvh_model_destroy(pl->model);
  delete pl;
}

}  // namespace mg

extern "C" {

int mgbvh_scene_plan_create(mgbvh_ctx* c, const float* d_pos, uint32_t vstride_bytes,
                            uint64_t nverts, const uint32_t* d_tris,
                            const mgbvh_mesh_desc* meshes, uint32_t nmeshes,
                            const mgbvh_settings* blas_settings, mgbvh_node32* d_blas_nodes,
                            uint32_t* d_blas_lens, uint32_t* d_blas_perm, uint32_t* d_tris_out,
                            const float* d_object_to_world, const uint32_t* d_instance_blas,
                            uint32_t ninstances, const mgbvh_settings* tlas_settings,
                            mgbvh_aabb* d_boxes, mgbvh_node32* d_tlas_nodes,
                            uint32_t* d_tlas_lens, uint32_t* d_tlas_perm,
                            mgbvh_scene_plan** out) {
  if (!c || !out || !blas_settings || (nmeshes && !meshes)) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (blas_settings->has_bounds || blas_settings->profiling_info || blas_settings->max_depth > 64 ||
      (ninstances && (!tlas_settings || tlas_settings->has_bounds || tlas_settings->profiling_info ||
                      tlas_settings->max_depth > 64 ||
                      tlas_settings->method != MGBVH_METHOD_PLAIN))) {
    set_error("scene plans take union bounds, no profiling JSON, max_depth <= 64 and a plain TLAS");
    return MGBVH_ERR_UNSUPPORTED;
  }
  if (blas_settings->min_leaf_objects != 1 || blas_settings->max_leaf_objects != 1 ||
      (blas_settings->method == MGBVH_METHOD_LEARNED && !c->model) ||
      blas_settings->method > MGBVH_METHOD_LEARNED) {
    set_error("unsupported BLAS settings (leaf sizes, method or missing model)");
    return MGBVH_ERR_UNSUPPORTED;
  }
  if (vstride_bytes < 12 || vstride_bytes % 4) {
    set_error("vstride_bytes must be a multiple of 4 and >= 12");
    return MGBVH_ERR_INVALID_ARG;
  }
  std::vector<MeshIn> ms(nmeshes);
  uint64_t total = 0;
  for (uint32_t i = 0; i < nmeshes; ++i) {
    const mgbvh_mesh_desc& d = meshes[i];
    if (d.tri_offset != total || (d.tri_count && (uint64_t)d.vertex_offset + d.vertex_count > nverts)) {
      set_error("mesh " + std::to_string(i) + ": not contiguous or vertices beyond nverts");
      return MGBVH_ERR_INVALID_ARG;
    }
    ms[i] = MeshIn{d.tri_offset, d.tri_count, d.vertex_offset, d.vertex_count, 0, {0, 0, 0},
                   {0, 0, 0}};
    total += d.tri_count;
  }
  if (total == 0 || !d_pos || !d_tris || !d_blas_nodes || !d_blas_perm ||
      (ninstances && (!d_object_to_world || !d_instance_blas || !d_boxes || !d_tlas_nodes ||
                      !d_tlas_perm))) {
    set_error("empty scene or null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  mgbvh_scene_plan* pl = new mgbvh_scene_plan();
  pl->c = c;
  pl->nblas = nmeshes;
  pl->ninst = ninstances;
  pl->o2w = d_object_to_world;
  pl->inst_blas = d_instance_blas;
  pl->boxes = d_boxes;
  pl->blas_nodes = d_blas_nodes;
  int rc = MGBVH_OK;
  auto fail = [&](int r) { plan_release(pl); return r; };
  if ((e = hipHostMalloc((void**)&pl->bsc.h_ctr, sizeof(Counters), hipHostMallocDefault)) ||
      (e = hipHostMalloc((void**)&pl->tsc.h_ctr, sizeof(Counters), hipHostMallocDefault)) ||
      (e = hipEventCreate(&pl->e0)) || (e = hipEventCreate(&pl->e1)))
    return fail(hip_fail(e, "plan setup"));
  BuildInput bin{0, d_pos, vstride_bytes / 4, nverts, d_tris, nullptr};
  BuildOutput bout{d_blas_nodes, d_blas_lens, d_blas_perm, d_tris_out, nullptr};
  if ((rc = plan_build(c, pl->bsc, bin, ms, blas_settings, bout, &pl->bp))) return fail(rc);
  std::vector<mgbvh_blas_desc> desc(nmeshes);
  for (uint32_t i = 0; i < nmeshes; ++i)
    desc[i] = mgbvh_blas_desc{pl->bp.rows[i].nbase, meshes[i].tri_offset, meshes[i].vertex_offset, 0};
  if ((rc = ensure(pl->blas_desc, (size_t)std::max(nmeshes, 1u) * sizeof(mgbvh_blas_desc))) ||
      (rc = ensure(pl->flag, 256)))
    return fail(rc);
  if (ninstances) {
    BuildInput tin{1, nullptr, 0, 0, nullptr, (const float*)d_boxes};
    BuildOutput tout{d_tlas_nodes, d_tlas_lens, d_tlas_perm, nullptr, nullptr};
    std::vector<MeshIn> one{MeshIn{0, ninstances, 0, 0, 0, {0, 0, 0}, {0, 0, 0}}};
    if ((rc = plan_build(c, pl->tsc, tin, one, tlas_settings, tout, &pl->tp))) return fail(rc);
  }
  // constant inputs of the graph, outside it
  hipStream_t st = c->stream;
  if ((e = hipMemcpyAsync(pl->blas_desc.p, desc.data(), desc.size() * sizeof(mgbvh_blas_desc),
                          hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(pl->bsc.mrows.p, pl->bp.rows.data(), pl->bp.rows.size() * sizeof(MeshRow),
                          hipMemcpyHostToDevice, st)) ||
      (ninstances &&
       (e = hipMemcpyAsync(pl->tsc.mrows.p, pl->tp.rows.data(), pl->tp.rows.size() * sizeof(MeshRow),
                           hipMemcpyHostToDevice, st))))
    return fail(hip_fail(e, "plan upload"));
  // the captured kernels and a level-by-level redo both use the plan's own
  // copy of the model, so later training, set_params, set_model or destroy of
  // the context's model change nothing in this plan
  if (blas_settings->method == MGBVH_METHOD_LEARNED && !blas_settings->learned_no_network) {
    if ((rc = model_snapshot(c, c->model, &pl->model))) return fail(rc);
    pl->bp.lp.model = pl->model;
  }
  if ((e = hipStreamSynchronize(st))) return fail(hip_fail(e, "plan upload"));
  // one build level by level gives each tree's level count; the graph runs
  // that many levels plus a margin (launch redoes a replay that needed more)
  {
    BuildResult rb, rt;
    if ((rc = plan_levels(pl, &rb, &rt))) return fail(rc);
    constexpr uint32_t kMargin = 2;
    pl->bp.fixed_levels = std::max<uint32_t>(1, std::min(pl->bp.max_levels, rb.n_levels + kMargin));
    if (ninstances)
      pl->tp.fixed_levels = std::max<uint32_t>(1, std::min(pl->tp.max_levels, rt.n_levels + kMargin));
  }
  // record
  const uint32_t prof = c->prof_mask;
  c->prof_mask = 0;
  e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) { c->prof_mask = prof; return fail(hip_fail(e, "hipStreamBeginCapture")); }
  uint32_t lv_run = 0, lv_run_t = 0;
  // The TLAS needs only each mesh's box, known once the BLAS objects are set
  // up: with every mesh non-empty its branch (instance boxes from the per-mesh
  // root bounds, then the TLAS levels) is recorded on the second stream beside
  // the BLAS levels, whose stream waits for it at the end.  An empty mesh's
  // instance box is read from the node array, so such a scene is recorded in
  // sequence.
  bool all_nonempty = true;
  for (uint32_t i = 0; i < nmeshes; ++i) all_nonempty = all_nonempty && meshes[i].tri_count > 0;
  const bool branch = ninstances && all_nonempty;
  hipStream_t st2 = c->stream2;
  const std::function<int()> tlas_branch = [&]() -> int {
    hipError_t e2;
    if ((e2 = hipEventRecord(c->fork_ev, st)) != hipSuccess ||
        (e2 = hipStreamWaitEvent(st2, c->fork_ev, 0)) != hipSuccess)
      return hip_fail(e2, "TLAS branch fork");
    k_fill<<<1, 64, 0, st2>>>((uint32_t*)pl->flag.p, 0u, 1);
    k_instance_boxes_acc<<<(ninstances + 255) / 256, 256, 0, st2>>>(
        (const uint32_t*)pl->bsc.macc.p, nmeshes, d_object_to_world, d_instance_blas, ninstances,
        d_boxes, (uint32_t*)pl->flag.p);
    if ((e2 = hipGetLastError()) != hipSuccess) return hip_fail(e2, "k_instance_boxes_acc");
    const int r2 = enqueue_build(c, pl->tsc, pl->tp, pl->tp.fixed_levels, false, &lv_run_t, st2);
    if (r2) return r2;
    if ((e2 = hipEventRecord(c->join_ev, st2)) != hipSuccess) return hip_fail(e2, "TLAS branch join");
    return MGBVH_OK;
  };
  rc = enqueue_build(c, pl->bsc, pl->bp, pl->bp.fixed_levels, false, &lv_run, st,
                     branch ? &tlas_branch : nullptr);
  if (!rc && branch) {
    if ((e = hipStreamWaitEvent(st, c->join_ev, 0)) != hipSuccess) rc = hip_fail(e, "TLAS branch join");
  } else if (!rc && ninstances) {
    k_fill<<<1, 64, 0, st>>>((uint32_t*)pl->flag.p, 0u, 1);
    k_instance_boxes<<<(ninstances + 255) / 256, 256, 0, st>>>(
        d_blas_nodes, (const mgbvh_blas_desc*)pl->blas_desc.p, nmeshes, d_object_to_world,
        d_instance_blas, ninstances, d_boxes, (uint32_t*)pl->flag.p);
    e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "k_instance_boxes");
    if (!rc)
      rc = enqueue_build(c, pl->tsc, pl->tp, pl->tp.fixed_levels, false, &lv_run_t);
  }
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(st, &g);
  c->prof_mask = prof;
  pl->graph = g;
  if (rc) return fail(rc);
  if (e != hipSuccess) return fail(hip_fail(e, "hipStreamEndCapture"));
  e = hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) return fail(hip_fail(e, "hipGraphInstantiate"));
  *out = pl;
  return MGBVH_OK;
}

int mgbvh_scene_plan_launch(mgbvh_scene_plan* pl, float* out_ms, uint32_t* out_blas_node_counts,
                            uint64_t* out_tlas_nodes) {
  if (!pl) { set_error("plan is NULL"); return MGBVH_ERR_INVALID_ARG; }
  mgbvh_ctx* c = pl->c;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = c->stream;
  if ((e = hipEventRecord(pl->e0, st)) || (e = hipGraphLaunch(pl->exec, st)) ||
      (e = hipEventRecord(pl->e1, st)) || (e = hipEventSynchronize(pl->e1)))
    return hip_fail(e, "hipGraphLaunch");
  float ms = 0.0f;
  if ((e = hipEventElapsedTime(&ms, pl->e0, pl->e1))) return hip_fail(e, "hipEventElapsedTime");
  BuildResult r, rt;
  int rc = finish_build(c, pl->bsc, pl->bp, &r);
  bool redo = rc == MG_LEVELS_SHORT;
  if (rc && !redo) return rc;
  if (!redo && pl->ninst) {
    rc = finish_build(c, pl->tsc, pl->tp, &rt);
    redo = rc == MG_LEVELS_SHORT;
    if (rc && !redo) return rc;
  }
  if (redo) {
    // the new geometry needs deeper trees than the graph records: build this
    // scene level by level instead (same results; slower for this launch)
    if ((rc = plan_levels(pl, &r, &rt))) return rc;
    pl->n_redo++;
  }
  if (out_blas_node_counts)
    for (uint32_t i = 0; i < pl->nblas; ++i) out_blas_node_counts[i] = r.mesh_nodes[i];
  if (pl->ninst) {
    uint32_t flag = 0;
    if ((e = hipMemcpy(&flag, pl->flag.p, 4, hipMemcpyDeviceToHost)))
      return hip_fail(e, "hipMemcpy(flag)");
    if (flag) { set_error("an instance names a BLAS >= nmeshes"); return MGBVH_ERR_INDEX; }
    if (out_tlas_nodes) *out_tlas_nodes = rt.nnodes;
  } else if (out_tlas_nodes) {
    *out_tlas_nodes = 0;
  }
  if (out_ms) *out_ms = ms;
  return MGBVH_OK;
}

int mgbvh_scene_plan_info(const mgbvh_scene_plan* pl, uint32_t* blas_levels,
                          uint32_t* tlas_levels, uint32_t* launches_rebuilt) {
  if (!pl) { set_error("plan is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (blas_levels) *blas_levels = pl->bp.fixed_levels;
  if (tlas_levels) *tlas_levels = pl->tp.fixed_levels;
  if (launches_rebuilt) *launches_rebuilt = pl->n_redo;
  return MGBVH_OK;
}

int mgbvh_scene_plan_destroy(mgbvh_scene_plan* pl) {
  if (!pl) return MGBVH_OK;
  (void)hipSetDevice(pl->c->device);
  (void)hipStreamSynchronize(pl->c->stream);
  plan_release(pl);
  return MGBVH_OK;
}

int mgbvh_ctx_set_tuning(mgbvh_ctx* c, const mgbvh_tuning* t) {
  if (!c || !t) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  auto ok = [](uint32_t v) { return v >= 1 && v <= 256; };
  auto waves = [](uint32_t v) { return v == 0 || v == 1 || v == 4; };
  if (!ok(t->bin_blocks_per_cu) || !ok(t->partition_blocks_per_cu) ||
      !ok(t->sweep_blocks_per_cu) || !ok(t->small_blocks_per_cu) || !waves(t->partition_waves) ||
      !(t->bin_waves <= 4) ||
      !(t->medium_max_objects <= 1 || (t->medium_max_objects > MG_SMALL && t->medium_max_objects <= CH))) {
    set_error("blocks per compute unit must be in [1, 256]; partition_waves 1 or 4, bin_waves <= 4; "
              "medium_max_objects 0, 1 or 33..2048");
    return MGBVH_ERR_INVALID_ARG;
  }
  c->tune.bin_bpc = t->bin_blocks_per_cu;
  c->tune.part_bpc = t->partition_blocks_per_cu;
  c->tune.sweep_bpc = t->sweep_blocks_per_cu;
  c->tune.small_bpc = t->small_blocks_per_cu;
  if (t->partition_waves) c->tune.part_waves = t->partition_waves;
  if (t->bin_waves) c->tune.bin_waves = t->bin_waves;
  if (t->sweep_pack_min_nodes) c->tune.sweep_pack = t->sweep_pack_min_nodes;
  if (t->medium_max_objects) c->tune.med_max = t->medium_max_objects == 1 ? 0u : t->medium_max_objects;
  return MGBVH_OK;
}

int mgbvh_ctx_get_tuning(mgbvh_ctx* c, mgbvh_tuning* t) {
  if (!c || !t) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  memset(t, 0, sizeof *t);
  t->bin_blocks_per_cu = c->tune.bin_bpc;
  t->partition_blocks_per_cu = c->tune.part_bpc;
  t->sweep_blocks_per_cu = c->tune.sweep_bpc;
  t->small_blocks_per_cu = c->tune.small_bpc;
  t->compute_units = (uint32_t)c->cus;
  t->partition_waves = c->tune.part_waves;
  t->bin_waves = c->tune.bin_waves;
  t->sweep_pack_min_nodes = c->tune.sweep_pack;
  t->medium_max_objects = c->tune.med_max ? c->tune.med_max : 1u;
  return MGBVH_OK;
}

}  // extern "C"
