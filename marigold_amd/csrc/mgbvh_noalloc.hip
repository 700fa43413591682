// Allocation-free variant of the single-mesh device build
// (mgbvh_build_tris_device_noalloc): every scratch buffer of the build lives
// in one file-scope device array sized for NOALLOC_MAX_TRIS triangles, so a
// call allocates no device or pinned memory and creates no HIP object.  For
// callers that cannot allocate: code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc.
//
// The array's name is a host-side stand-in, not a device address: its
// address on each device comes from hipGetSymbolAddress, taken once per
// device (with the device current) when the first context on it is created.
// Every call on a device shares that device's array, which keeps the last
// call's contents; the build clears what it needs cleared (counters, level
// counts) itself, as it does with the context's reused scratch.  Calls that
// overlap on one device race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mgbvh_internal.h"

namespace mg {

constexpr uint64_t NOALLOC_MAX_TRIS = 65536;

constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Capacities for n objects of one mesh (the sizes plan_build and finish_build
// ask for, each monotone in n), in the order noalloc_views lays them out.
struct NoallocSizes {
  size_t recs, L, cn, hist, blk, small, ctr, gaps, mrows, macc, lv, lens, gtab, mnb, tmpn, tmpl,
      tmps;
};
constexpr NoallocSizes noalloc_sizes(uint64_t n) {
  const uint64_t capL = n / (MG_SMALL + 1) + 2;
  const uint64_t capC = (n + CH - 1) / CH + capL + 2;
  const uint64_t slots = 2 * n - 1;
  NoallocSizes z{};
  z.recs = n * sizeof(Rec);
  z.L = capL * sizeof(LNode);
  z.cn = capC * 4;
  z.hist = capL * HIST_WORDS * 4;
  z.blk = capC * 96 * 4;
  z.small = small_list_cap(n) * sizeof(SNode);
  z.ctr = sizeof(Counters);
  z.gaps = (n / 2 + 1) * 8;
  z.mrows = sizeof(MeshRow);
  z.macc = 12 * 4;
  z.lv = (n + 2) * LV_WORDS * 4;  // levels <= min(max_depth, n)
  z.lens = slots * 4;
  z.gtab = 3 * (n / 2 + 1) * 4;
  z.mnb = 4;
  z.tmpn = slots * sizeof(mgbvh_node32);
  z.tmpl = slots * 4;
  z.tmps = slots * sizeof(mgbvh_split_rec);
  return z;
}
constexpr size_t noalloc_total(uint64_t n) {
  const NoallocSizes z = noalloc_sizes(n);
  return 2 * al256(z.recs) + 2 * al256(z.L) + 4 * al256(z.cn) + al256(z.hist) + al256(z.blk) +
         al256(z.small) + al256(z.ctr) + al256(z.gaps) + al256(z.mrows) + al256(z.macc) +
         al256(z.lv) + al256(z.lens) + al256(z.gtab) + al256(z.mnb) + al256(z.tmpn) +
         al256(z.tmpl) + al256(z.tmps);
}
constexpr size_t NOALLOC_BYTES = noalloc_total(NOALLOC_MAX_TRIS);  // about 24 MB

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static char* g_scratch[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

int noalloc_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return MGBVH_OK;  // variant unavailable there
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return MGBVH_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_scratch[device] = static_cast<char*>(ptr);
  return MGBVH_OK;
}

// Fixed views into this device's array (never grown or freed by ensure).
static void noalloc_views(Scratch& sc, char* base) {
  const NoallocSizes z = noalloc_sizes(NOALLOC_MAX_TRIS);
  size_t off = 0;
  const auto view = [&](DevBuf& b, size_t bytes) {
    b.p = base + off;
    b.cap = bytes;
    b.fixed = true;
    off += al256(bytes);
  };
  for (int q = 0; q < 2; ++q) view(sc.recs[q], z.recs);
  for (int q = 0; q < 2; ++q) view(sc.L[q], z.L);
  for (int q = 0; q < 2; ++q) { view(sc.cn[q], z.cn); view(sc.ck[q], z.cn); }
  view(sc.hist, z.hist);
  view(sc.blk, z.blk);
  view(sc.small, z.small);
  view(sc.ctr, z.ctr);
  view(sc.gaps, z.gaps);
  view(sc.mrows, z.mrows);
  view(sc.macc, z.macc);
  view(sc.lv, z.lv);
  view(sc.lens, z.lens);
  view(sc.gtab, z.gtab);
  view(sc.mnb, z.mnb);
  view(sc.tmpn, z.tmpn);
  view(sc.tmpl, z.tmpl);
  view(sc.tmps, z.tmps);
}

}  // namespace mg

using namespace mg;

extern "C" {

int mgbvh_build_tris_device_noalloc(mgbvh_ctx* c, const float* d_pos, uint32_t vstride_bytes,
                                    uint64_t nverts, const uint32_t* d_tris, uint64_t ntris,
                                    const mgbvh_settings* s, mgbvh_node32* d_nodes,
                                    uint32_t* d_lens, uint32_t* d_perm, uint32_t* d_tris_out,
                                    mgbvh_split_rec* d_splits, uint64_t* out_nnodes) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (!s) { set_error("settings is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (s->min_leaf_objects != 1 || s->max_leaf_objects != 1) {
    set_error("only MIN_LEAF_OBJECTS = MAX_LEAF_OBJECTS = 1 is built on the GPU");
    return MGBVH_ERR_UNSUPPORTED;
  }
  if (s->method != MGBVH_METHOD_PLAIN || s->profiling_info) {
    set_error("the allocation-free build takes the plain method and writes no profiling JSON");
    return MGBVH_ERR_UNSUPPORTED;
  }
  if (ntris > NOALLOC_MAX_TRIS) {
    set_error("the allocation-free build takes at most 65536 triangles");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (vstride_bytes < 12 || vstride_bytes % 4) {
    set_error("vstride_bytes must be a multiple of 4 and >= 12");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (ntris && (!d_pos || !d_tris || !d_nodes || !d_perm)) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES || !g_scratch[c->device]) {
    set_error("no allocation-free scratch on this device");
    return MGBVH_ERR_UNSUPPORTED;
  }
  if (out_nnodes) *out_nnodes = 0;
  if (ntris == 0) return MGBVH_OK;  // lib.rs:703-708: empty BVH
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  Scratch sc;
  noalloc_views(sc, g_scratch[c->device]);
  sc.h_ctr = c->h_ctr;
  BuildInput in{0, d_pos, vstride_bytes / 4, nverts, d_tris, nullptr};
  BuildOutput o{d_nodes, d_lens, d_perm, d_tris_out, d_splits};
  MeshIn m{};
  m.count = (uint32_t)ntris;
  m.vertex_count = nverts > 0xffffffffull ? 0xffffffffu : (uint32_t)nverts;
  m.has_bounds = s->has_bounds;
  for (int a = 0; a < 3; ++a) { m.bmn[a] = s->bounds_min[a]; m.bmx[a] = s->bounds_max[a]; }
  // kernel timing would create events: off for this call
  const uint32_t prof = c->prof_mask;
  c->prof_mask = 0;
  BuildPlan p;
  BuildResult r;
  int rc = plan_build(c, sc, in, std::vector<MeshIn>{m}, s, o, &p);
  if (!rc) rc = enqueue_build(c, sc, p, 0, true, nullptr);
  if (!rc) rc = finish_build(c, sc, p, &r);
  c->prof_mask = prof;
  if (rc) return rc;
  if (out_nnodes) *out_nnodes = r.nnodes;
  return MGBVH_OK;
}

}  // extern "C"
