// C ABI of mgbvh (include/mgbvh.h): context, argument checks, host-buffer
// wrappers around the device builder, BvhProfilingInfo statistics and JSON.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <random>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "mgbvh_internal.h"

namespace mg {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
  return MGBVH_ERR_DEVICE;
}

int ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.cap >= bytes) return MGBVH_OK;
  if (b.fixed) {
    set_error("fixed scratch region too small (internal sizing error)");
    return MGBVH_ERR_INTERNAL;
  }
  if (b.p) {
    hipError_t e = hipFree(b.p);
    if (e != hipSuccess) return hip_fail(e, "hipFree");
    b.p = nullptr;
    b.cap = 0;
  }
  size_t want = bytes + bytes / 8;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  b.cap = want;
  return MGBVH_OK;
}

int prof_begin(mgbvh_ctx* c, int kind, hipStream_t st) {
  const size_t k = c->prof_kind.size();
  while (c->prof_pool.size() < 2 * (k + 1)) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return hip_fail(r, "hipEventCreate");
    c->prof_pool.push_back(e);
  }
  c->prof_kind.push_back(kind);
  hipError_t r = hipEventRecord(c->prof_pool[2 * k], st ? st : c->stream);
  return r == hipSuccess ? MGBVH_OK : hip_fail(r, "hipEventRecord");
}

int prof_end(mgbvh_ctx* c, int kind, hipStream_t st) {
  (void)kind;
  const size_t k = c->prof_kind.size() - 1;
  hipError_t r = hipEventRecord(c->prof_pool[2 * k + 1], st ? st : c->stream);
  return r == hipSuccess ? MGBVH_OK : hip_fail(r, "hipEventRecord");
}

int prof_collect(mgbvh_ctx* c) {
  for (size_t k = 0; k < c->prof_kind.size(); ++k) {
    float ms = 0.0f;
    hipError_t r = hipEventSynchronize(c->prof_pool[2 * k + 1]);
    if (r == hipSuccess) r = hipEventElapsedTime(&ms, c->prof_pool[2 * k], c->prof_pool[2 * k + 1]);
    if (r != hipSuccess) return hip_fail(r, "hipEventElapsedTime");
    c->prof_ms[c->prof_kind[k]] += ms;
    c->prof_n[c->prof_kind[k]] += 1;
  }
  c->prof_kind.clear();
  return MGBVH_OK;
}

static void free_buf(DevBuf& b) {
  if (b.p && !b.fixed) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

void scratch_free(Scratch& sc) {
  DevBuf* bufs[] = {&sc.recs[0], &sc.recs[1], &sc.L[0],     &sc.L[1],     &sc.hist,  &sc.blk,
                    &sc.cn[0],   &sc.cn[1],   &sc.ck[0],    &sc.ck[1],    &sc.small, &sc.ctr,
                    &sc.gaps,    &sc.lens,    &sc.mrows,    &sc.macc,     &sc.mnb,   &sc.lrows[0],
                    &sc.lrows[1], &sc.lv,     &sc.lfeat,    &sc.llogit,   &sc.gtab,  &sc.tmpn,
                    &sc.tmpl,    &sc.tmps,    &sc.w8rec,    &sc.w8par,    &sc.w8e,
                    &sc.w8fast,  &sc.w8front[0], &sc.w8front[1], &sc.w8ctr, &sc.w8base};
  for (DevBuf* b : bufs) free_buf(*b);
}

static int check_settings(const mgbvh_settings* s, const mgbvh_ctx* c) {
  if (!s) { set_error("settings is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (s->min_leaf_objects != 1 || s->max_leaf_objects != 1) {
    set_error("only MIN_LEAF_OBJECTS = MAX_LEAF_OBJECTS = 1 is built on the GPU");
    return MGBVH_ERR_UNSUPPORTED;
  }
  if (s->method == MGBVH_METHOD_LEARNED) {
    if (!c || !c->model) {
      set_error("the learned method needs a model (mgbvh_ctx_set_model)");
      return MGBVH_ERR_UNSUPPORTED;
    }
    if (!(s->learned_guard >= 0.0f) || s->learned_window > 31) {
      set_error("learned_guard must be >= 0 and learned_window <= 31");
      return MGBVH_ERR_INVALID_ARG;
    }
  } else if (s->method != MGBVH_METHOD_PLAIN) {
    set_error("unknown build method");
    return MGBVH_ERR_UNSUPPORTED;
  }
  return MGBVH_OK;
}

// Shortest decimal that round-trips, like serde_json's f64 output.
static std::string fmt_f64(double v) {
  if (std::isnan(v) || std::isinf(v)) return "null";
  char buf[64];
  for (int p = 1; p <= 17; ++p) {
    snprintf(buf, sizeof buf, "%.*g", p, v);
    if (strtod(buf, nullptr) == v) break;
  }
  std::string s(buf);
  if (s.find_first_of(".eE") == std::string::npos) s += ".0";
  return s;
}

// BvhProfilingInfo fields, bvh/src/lib.rs:767-848 (leaves counted over nodes[1..]).
static void fill_stats(const std::vector<mgbvh_node32>& nodes, const std::vector<uint32_t>& lens,
                       uint64_t nn, uint64_t nobj, const mgbvh_settings* s,
                       const BuildResult& r, mgbvh_stats* st) {
  memset(st, 0, sizeof *st);
  st->construction_time = (double)r.ms * 1e-3;
  st->node_count = (uint32_t)nn;
  uint32_t cnt = 0, mn = 0, mx = 0;
  uint64_t sum = 0;
  bool first = true;
  for (uint64_t i = 1; i < nn; ++i) {
    if (nodes[i].child_node != 0) continue;
    uint32_t l = lens[i];
    ++cnt;
    sum += l;
    if (first || l < mn) mn = l;
    if (first || l > mx) mx = l;
    first = false;
  }
  const uint32_t root_len = nn ? lens[0] : 0;
  st->leaf_node_count = cnt;
  st->min_leaf_object_count = cnt ? mn : root_len;
  st->max_leaf_object_count = cnt ? mx : root_len;
  double avg = (double)sum / (double)cnt;
  double var = 0.0;
  for (uint64_t i = 1; i < nn; ++i)
    if (nodes[i].child_node == 0) var += ((double)lens[i] - avg) * ((double)lens[i] - avg);
  st->avg_leaf_object_count = avg;
  st->stddev_leaf_object_count = std::sqrt(var / (double)cnt);
  st->total_object_count = (uint32_t)nobj;
  // height: deepest node that was split (the fetch_max at lib.rs:656)
  uint32_t height = 0;
  if (nn > 1) {
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    stack.push_back({0u, 0u});
    while (!stack.empty()) {
      const std::pair<uint32_t, uint32_t> top = stack.back();
      stack.pop_back();
      const uint32_t ch = nodes[top.first].child_node;
      if (ch == 0 || (uint64_t)ch + 1 >= nn) continue;
      height = std::max(height, top.second);
      stack.push_back({ch, top.second + 1});
      stack.push_back({ch + 1, top.second + 1});
    }
  }
  st->height = height;
  st->max_depth = s->max_depth;
  st->n_large_levels = r.n_levels;
  st->n_median = r.n_median;
  st->n_cutoff_leaves = r.n_cutoff;
  st->n_guard_fallbacks = r.n_guard;
  st->large_object_visits = r.large_visits;
  st->large_nodes = (uint32_t)r.large_nodes;
  st->n_window_splits = r.n_window;
}

// bvh/src/lib.rs:851-866: bvh_<name>_<9-digit id>.json
static void write_json(const mgbvh_settings* s, const mgbvh_stats* st) {
  if (!s->profiling_info || !s->profiling_info_directory) return;
  std::string dir = s->profiling_info_directory;
  std::string acc;
  for (size_t i = 0; i <= dir.size(); ++i) {
    if (i == dir.size() || dir[i] == '/') {
      if (!acc.empty()) mkdir(acc.c_str(), 0755);
    }
    if (i < dir.size()) acc.push_back(dir[i]);
  }
  std::random_device rd;
  std::uniform_int_distribution<uint32_t> dist(100000000u, 999999999u);
  const char* name = s->name ? s->name : "";
  std::string path = dir + "/bvh_" + name + "_" + std::to_string(dist(rd)) + ".json";
  struct stat sb;
  if (stat(path.c_str(), &sb) == 0) return;
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return;
  std::string nm;
  for (const char* p = name; *p; ++p) {
    if (*p == '"' || *p == '\\') nm.push_back('\\');
    nm.push_back(*p);
  }
  fprintf(f,
          "{\"name\":\"%s\",\"construction_time\":%s,\"node_count\":%u,\"leaf_node_count\":%u,"
          "\"min_leaf_object_count\":%u,\"max_leaf_object_count\":%u,"
          "\"avg_leaf_object_count\":%s,\"stddev_leaf_object_count\":%s,"
          "\"total_object_count\":%u,\"height\":%u,\"max_depth\":%u}\n",
          nm.c_str(), fmt_f64(st->construction_time).c_str(), st->node_count, st->leaf_node_count,
          st->min_leaf_object_count, st->max_leaf_object_count,
          fmt_f64(st->avg_leaf_object_count).c_str(), fmt_f64(st->stddev_leaf_object_count).c_str(),
          st->total_object_count, st->height, st->max_depth);
  fclose(f);
}

static int stats_from_device(mgbvh_ctx* c, const mgbvh_node32* d_nodes, const uint32_t* d_lens,
                             uint64_t nn, uint64_t nobj, const mgbvh_settings* s,
                             const BuildResult& r, mgbvh_stats* st) {
  std::vector<mgbvh_node32> nodes(nn);
  std::vector<uint32_t> lens(nn);
  hipError_t e = hipMemcpy(nodes.data(), d_nodes, nn * sizeof(mgbvh_node32), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(nodes)");
  e = hipMemcpy(lens.data(), d_lens, nn * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(lens)");
  fill_stats(nodes, lens, nn, nobj, s, r, st);
  return MGBVH_OK;
}

}  // namespace mg

using namespace mg;

extern "C" {

void mgbvh_settings_default(mgbvh_settings* s) {
  if (!s) return;
  memset(s, 0, sizeof *s);
  s->max_depth = 32;  // BLAS_MAX_DEPTH / TLAS_MAX_DEPTH, marigold/src/app/data/scene/mod.rs:28-29
  s->min_leaf_objects = 1;
  s->max_leaf_objects = 1;
  s->method = MGBVH_METHOD_PLAIN;
}

const char* mgbvh_last_error(void) { return g_err.c_str(); }

int mgbvh_ctx_set_model(mgbvh_ctx* c, mgbvh_model* model) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  c->model = model;
  return MGBVH_OK;
}

int mgbvh_ctx_create(int device, mgbvh_ctx** out) {
  if (!out) { set_error("out is NULL"); return MGBVH_ERR_INVALID_ARG; }
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= device || device < 0) {
    set_error("no HIP device available");
    return MGBVH_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
    return MGBVH_ERR_NO_DEVICE;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  mgbvh_ctx* c = new mgbvh_ctx();
  c->device = device;
  c->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate"); }
  // the second stream has the default priority too (mixed priorities are avoided)
  if ((e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming)) != hipSuccess) {
    (void)mgbvh_ctx_destroy(c);
    return hip_fail(e, "hipStreamCreate/hipEventCreate");
  }
  for (int i = 0; i < 64; ++i) {
    e = hipEventCreate(&c->ev[i]);
    if (e != hipSuccess) { (void)mgbvh_ctx_destroy(c); return hip_fail(e, "hipEventCreate"); }
  }
  e = hipHostMalloc((void**)&c->h_ctr, sizeof(Counters), hipHostMallocDefault);
  if (e != hipSuccess) { (void)mgbvh_ctx_destroy(c); return hip_fail(e, "hipHostMalloc"); }
  c->sc.h_ctr = c->h_ctr;
  int rc = noalloc_init(device);  // once per device, with the device current
  if (rc) { (void)mgbvh_ctx_destroy(c); return rc; }
  *out = c;
  return MGBVH_OK;
}

int mgbvh_ctx_destroy(mgbvh_ctx* c) {
  if (!c) return MGBVH_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  scratch_free(c->sc);
  DevBuf* bufs[] = {&c->io_pos, &c->io_tris, &c->io_nodes, &c->io_lens, &c->io_perm,
                    &c->io_tris_out, &c->io_splits, &c->io_boxes};
  for (DevBuf* b : bufs) free_buf(*b);
  for (int i = 0; i < 64; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
  if (c->h_ctr) (void)hipHostFree(c->h_ctr);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->join_ev) (void)hipEventDestroy(c->join_ev);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MGBVH_OK;
}

void* mgbvh_ctx_stream(mgbvh_ctx* c) { return c ? (void*)c->stream : nullptr; }

int mgbvh_ctx_synchronize(mgbvh_ctx* c) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipStreamSynchronize");
}

int mgbvh_build_tris_device(mgbvh_ctx* c, const float* d_pos, uint32_t vstride_bytes,
                            uint64_t nverts, const uint32_t* d_tris, uint64_t ntris,
                            const mgbvh_settings* s, mgbvh_node32* d_nodes, uint32_t* d_lens,
                            uint32_t* d_perm, uint32_t* d_tris_out, mgbvh_split_rec* d_splits,
                            uint64_t* out_nnodes, mgbvh_stats* out_stats) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  int rc = check_settings(s, c);
  if (rc) return rc;
  if (ntris > 0x7fffffffull) { set_error("more than 2^31-1 objects"); return MGBVH_ERR_INVALID_ARG; }
  if (vstride_bytes < 12 || vstride_bytes % 4) {
    set_error("vstride_bytes must be a multiple of 4 and >= 12");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (ntris && (!d_pos || !d_tris || !d_nodes || !d_perm)) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  BuildInput in{0, d_pos, vstride_bytes / 4, nverts, d_tris, nullptr};
  BuildOutput o{d_nodes, d_lens, d_perm, d_tris_out, d_splits};
  if (out_stats && !d_lens) {
    if ((rc = ensure(c->io_lens, (2 * ntris + 1) * 4))) return rc;
    o.lens = (uint32_t*)c->io_lens.p;
  }
  BuildResult r;
  rc = build_device(c, in, (uint32_t)ntris, s, o, &r);
  if (rc) return rc;
  if (out_nnodes) *out_nnodes = r.nnodes;
  if (out_stats) {
    if (r.nnodes) {
      if ((rc = stats_from_device(c, d_nodes, o.lens, r.nnodes, ntris, s, r, out_stats))) return rc;
    } else {
      memset(out_stats, 0, sizeof *out_stats);
    }
    write_json(s, out_stats);
  }
  return MGBVH_OK;
}

int mgbvh_build_blas_batch_device(mgbvh_ctx* c, const float* d_pos, uint32_t vstride_bytes,
                                  uint64_t nverts, const uint32_t* d_tris,
                                  const mgbvh_mesh_desc* meshes, uint32_t nmeshes,
                                  const mgbvh_aabb* mesh_bounds, const mgbvh_settings* s,
                                  mgbvh_node32* d_nodes, uint32_t* d_lens, uint32_t* d_perm,
                                  uint32_t* d_tris_out, mgbvh_split_rec* d_splits,
                                  mgbvh_blas_desc* out_blas, uint32_t* out_node_counts,
                                  mgbvh_stats* out_stats) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  int rc = check_settings(s, c);
  if (rc) return rc;
  if (s->has_bounds || s->profiling_info) {
    set_error("batched builds take per-mesh bounds (mesh_bounds) and write no profiling JSON");
    return MGBVH_ERR_UNSUPPORTED;
  }
  if (nmeshes && !meshes) { set_error("meshes is NULL"); return MGBVH_ERR_INVALID_ARG; }
  if (vstride_bytes < 12 || vstride_bytes % 4) {
    set_error("vstride_bytes must be a multiple of 4 and >= 12");
    return MGBVH_ERR_INVALID_ARG;
  }
  std::vector<MeshIn> ms(nmeshes);
  uint64_t total = 0;
  for (uint32_t i = 0; i < nmeshes; ++i) {
    const mgbvh_mesh_desc& d = meshes[i];
    if (d.tri_offset != total) {
      set_error("mesh " + std::to_string(i) + ": tri_offset must equal the sum of earlier tri_count");
      return MGBVH_ERR_INVALID_ARG;
    }
    if (d.tri_count && (uint64_t)d.vertex_offset + d.vertex_count > nverts) {
      set_error("mesh " + std::to_string(i) + ": vertex_offset + vertex_count > nverts");
      return MGBVH_ERR_INDEX;
    }
    MeshIn& m = ms[i];
    m.tri_offset = d.tri_offset;
    m.count = d.tri_count;
    m.vertex_offset = d.vertex_offset;
    m.vertex_count = d.vertex_count;
    m.has_bounds = mesh_bounds != nullptr;
    for (int a = 0; a < 3; ++a) {
      m.bmn[a] = mesh_bounds ? mesh_bounds[i].min[a] : 0.0f;
      m.bmx[a] = mesh_bounds ? mesh_bounds[i].max[a] : 0.0f;
    }
    total += d.tri_count;
  }
  if (total > 0x7fffffffull) { set_error("more than 2^31-1 triangles"); return MGBVH_ERR_INVALID_ARG; }
  if (total && (!d_pos || !d_tris || !d_nodes || !d_perm)) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  BuildInput in{0, d_pos, vstride_bytes / 4, nverts, d_tris, nullptr};
  BuildOutput o{d_nodes, d_lens, d_perm, d_tris_out, d_splits};
  BuildResult r;
  rc = build_multi(c, in, ms, s, o, &r);
  if (rc) return rc;
  for (uint32_t i = 0; i < nmeshes; ++i) {
    if (out_blas) {
      out_blas[i].node_offset = r.mesh_nbase[i];
      out_blas[i].tri_offset = meshes[i].tri_offset;
      out_blas[i].vertex_offset = meshes[i].vertex_offset;
      out_blas[i].reserved = 0;
    }
    if (out_node_counts) out_node_counts[i] = r.mesh_nodes[i];
  }
  if (out_stats) {
    memset(out_stats, 0, sizeof *out_stats);
    out_stats->construction_time = r.ms * 1e-3;
    out_stats->node_count = (uint32_t)r.nnodes;
    out_stats->total_object_count = (uint32_t)total;
    out_stats->max_depth = s->max_depth;
    out_stats->n_large_levels = r.n_levels;
    out_stats->n_median = r.n_median;
    out_stats->n_cutoff_leaves = r.n_cutoff;
    out_stats->large_object_visits = r.large_visits;
    out_stats->large_nodes = (uint32_t)r.large_nodes;
    out_stats->n_guard_fallbacks = r.n_guard;
    out_stats->n_window_splits = r.n_window;
  }
  return MGBVH_OK;
}

int mgbvh_build_aabbs_device(mgbvh_ctx* c, const mgbvh_aabb* d_aabbs, uint64_t n,
                             const mgbvh_settings* s, mgbvh_node32* d_nodes, uint32_t* d_lens,
                             uint32_t* d_perm, mgbvh_split_rec* d_splits, uint64_t* out_nnodes,
                             mgbvh_stats* out_stats) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  int rc = check_settings(s, c);
  if (rc) return rc;
  if (n > 0x7fffffffull) { set_error("more than 2^31-1 objects"); return MGBVH_ERR_INVALID_ARG; }
  if (n && (!d_aabbs || !d_nodes || !d_perm)) {
    set_error("null device buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  BuildInput in{1, nullptr, 0, 0, nullptr, (const float*)d_aabbs};
  BuildOutput o{d_nodes, d_lens, d_perm, nullptr, d_splits};
  if (out_stats && !d_lens) {
    if ((rc = ensure(c->io_lens, (2 * n + 1) * 4))) return rc;
    o.lens = (uint32_t*)c->io_lens.p;
  }
  BuildResult r;
  rc = build_device(c, in, (uint32_t)n, s, o, &r);
  if (rc) return rc;
  if (out_nnodes) *out_nnodes = r.nnodes;
  if (out_stats) {
    if (r.nnodes) {
      if ((rc = stats_from_device(c, d_nodes, o.lens, r.nnodes, n, s, r, out_stats))) return rc;
    } else {
      memset(out_stats, 0, sizeof *out_stats);
    }
    write_json(s, out_stats);
  }
  return MGBVH_OK;
}

// Host-buffer variants: upload, build, download; `tris` is permuted in place.
int mgbvh_build_tris(mgbvh_ctx* c, const float* positions, uint32_t vstride_bytes,
                     uint64_t nverts, uint32_t* tris, uint64_t ntris, const mgbvh_settings* s,
                     mgbvh_node32* out_nodes, uint32_t* out_lens, uint64_t* out_nnodes,
                     uint32_t* out_perm, mgbvh_split_rec* out_splits, mgbvh_stats* out_stats) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  int rc = check_settings(s, c);
  if (rc) return rc;
  if (ntris == 0) {
    if (out_nnodes) *out_nnodes = 0;
    if (out_stats) memset(out_stats, 0, sizeof *out_stats);
    return MGBVH_OK;
  }
  if (!positions || !tris || !out_nodes) {
    set_error("null host buffer");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (vstride_bytes < 12 || vstride_bytes % 4) {
    set_error("vstride_bytes must be a multiple of 4 and >= 12");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (ntris > 0x7fffffffull) { set_error("more than 2^31-1 objects"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const uint64_t nslots = 2 * ntris - 1;
  const size_t vbytes = (size_t)(nverts ? nverts - 1 : 0) * vstride_bytes + 12;
  if ((rc = ensure(c->io_pos, vbytes))) return rc;
  if ((rc = ensure(c->io_tris, ntris * 12))) return rc;
  if ((rc = ensure(c->io_tris_out, ntris * 12))) return rc;
  if ((rc = ensure(c->io_nodes, nslots * sizeof(mgbvh_node32)))) return rc;
  if ((rc = ensure(c->io_lens, nslots * 4))) return rc;
  if ((rc = ensure(c->io_perm, ntris * 4))) return rc;
  if (out_splits && (rc = ensure(c->io_splits, nslots * sizeof(mgbvh_split_rec)))) return rc;
  if (nverts) {
    e = hipMemcpy(c->io_pos.p, positions, vbytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(positions)");
  }
  e = hipMemcpy(c->io_tris.p, tris, ntris * 12, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(tris)");
  BuildInput in{0, (const float*)c->io_pos.p, vstride_bytes / 4, nverts,
                (const uint32_t*)c->io_tris.p, nullptr};
  BuildOutput o{(mgbvh_node32*)c->io_nodes.p, (uint32_t*)c->io_lens.p, (uint32_t*)c->io_perm.p,
                (uint32_t*)c->io_tris_out.p,
                out_splits ? (mgbvh_split_rec*)c->io_splits.p : nullptr};
  BuildResult r;
  rc = build_device(c, in, (uint32_t)ntris, s, o, &r);
  if (rc) return rc;
  std::vector<mgbvh_node32> hn;
  std::vector<uint32_t> hl;
  e = hipMemcpy(out_nodes, o.nodes, r.nnodes * sizeof(mgbvh_node32), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(nodes)");
  if (out_lens) {
    e = hipMemcpy(out_lens, o.lens, r.nnodes * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(lens)");
  }
  if (out_perm) {
    e = hipMemcpy(out_perm, o.perm, ntris * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(perm)");
  }
  if (out_splits) {
    e = hipMemcpy(out_splits, o.splits, r.nnodes * sizeof(mgbvh_split_rec), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(splits)");
  }
  e = hipMemcpy(tris, o.tris_out, ntris * 12, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(tris)");
  if (out_nnodes) *out_nnodes = r.nnodes;
  if (out_stats) {
    hn.assign(out_nodes, out_nodes + r.nnodes);
    hl.resize(r.nnodes);
    e = hipMemcpy(hl.data(), o.lens, r.nnodes * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(lens)");
    fill_stats(hn, hl, r.nnodes, ntris, s, r, out_stats);
    write_json(s, out_stats);
  }
  return MGBVH_OK;
}

int mgbvh_build_aabbs(mgbvh_ctx* c, const mgbvh_aabb* aabbs, uint64_t n,
                      const mgbvh_settings* s, mgbvh_node32* out_nodes, uint32_t* out_lens,
                      uint64_t* out_nnodes, uint32_t* out_perm, mgbvh_split_rec* out_splits,
                      mgbvh_stats* out_stats) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  int rc = check_settings(s, c);
  if (rc) return rc;
  if (n == 0) {
    if (out_nnodes) *out_nnodes = 0;
    if (out_stats) memset(out_stats, 0, sizeof *out_stats);
    return MGBVH_OK;
  }
  if (!aabbs || !out_nodes || !out_perm) {
    set_error("null host buffer (aabbs, out_nodes and out_perm are required)");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (n > 0x7fffffffull) { set_error("more than 2^31-1 objects"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const uint64_t nslots = 2 * n - 1;
  if ((rc = ensure(c->io_boxes, n * sizeof(mgbvh_aabb)))) return rc;
  if ((rc = ensure(c->io_nodes, nslots * sizeof(mgbvh_node32)))) return rc;
  if ((rc = ensure(c->io_lens, nslots * 4))) return rc;
  if ((rc = ensure(c->io_perm, n * 4))) return rc;
  if (out_splits && (rc = ensure(c->io_splits, nslots * sizeof(mgbvh_split_rec)))) return rc;
  e = hipMemcpy(c->io_boxes.p, aabbs, n * sizeof(mgbvh_aabb), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(aabbs)");
  BuildInput in{1, nullptr, 0, 0, nullptr, (const float*)c->io_boxes.p};
  BuildOutput o{(mgbvh_node32*)c->io_nodes.p, (uint32_t*)c->io_lens.p, (uint32_t*)c->io_perm.p,
                nullptr, out_splits ? (mgbvh_split_rec*)c->io_splits.p : nullptr};
  BuildResult r;
  rc = build_device(c, in, (uint32_t)n, s, o, &r);
  if (rc) return rc;
  e = hipMemcpy(out_nodes, o.nodes, r.nnodes * sizeof(mgbvh_node32), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(nodes)");
  if (out_lens) {
    e = hipMemcpy(out_lens, o.lens, r.nnodes * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(lens)");
  }
  e = hipMemcpy(out_perm, o.perm, n * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(perm)");
  if (out_splits) {
    e = hipMemcpy(out_splits, o.splits, r.nnodes * sizeof(mgbvh_split_rec), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(splits)");
  }
  if (out_nnodes) *out_nnodes = r.nnodes;
  if (out_stats) {
    std::vector<mgbvh_node32> hn(out_nodes, out_nodes + r.nnodes);
    std::vector<uint32_t> hl(r.nnodes);
    e = hipMemcpy(hl.data(), o.lens, r.nnodes * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(lens)");
    fill_stats(hn, hl, r.nnodes, n, s, r, out_stats);
    write_json(s, out_stats);
  }
  return MGBVH_OK;
}

int mgbvh_device_alloc(mgbvh_ctx* c, uint64_t bytes, void** out) {
  if (!c || !out) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  e = hipMalloc(out, bytes ? bytes : 16);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipMalloc");
}

int mgbvh_device_free(mgbvh_ctx* c, void* p) {
  if (!c) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  if (!p) return MGBVH_OK;
  hipError_t e = hipFree(p);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipFree");
}

int mgbvh_copy_h2d(mgbvh_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  if (!bytes) return MGBVH_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipMemcpy H2D");
}

int mgbvh_copy_d2d(mgbvh_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  if (!bytes) return MGBVH_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipMemcpyAsync(d2d)");
}

int mgbvh_copy_d2h(mgbvh_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  if (!bytes) return MGBVH_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipMemcpy D2H");
}

int mgbvh_ctx_set_profiling(mgbvh_ctx* c, uint32_t kernel_mask) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  c->prof_mask = kernel_mask;
  return MGBVH_OK;
}

int mgbvh_ctx_kernel_times(mgbvh_ctx* c, double* ms, uint64_t* launches, int reset) {
  if (!c) { set_error("ctx is NULL"); return MGBVH_ERR_INVALID_ARG; }
  for (int k = 0; k < MGBVH_KERNEL_KINDS; ++k) {
    if (ms) ms[k] = c->prof_ms[k];
    if (launches) launches[k] = c->prof_n[k];
    if (reset) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
  }
  return MGBVH_OK;
}

int mgbvh_event_record(mgbvh_ctx* c, int slot) {
  if (!c || slot < 0 || slot >= 60) { set_error("bad event slot (0..59)"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipEventRecord(c->ev[slot], c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipEventRecord");
}

int mgbvh_event_elapsed_ms(mgbvh_ctx* c, int a, int b, float* ms) {
  if (!c || !ms || a < 0 || a >= 60 || b < 0 || b >= 60) {
    set_error("bad event slot (0..59)");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipEventSynchronize(c->ev[b]);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, c->ev[a], c->ev[b]);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipEventElapsedTime");
}

}  // extern "C"
