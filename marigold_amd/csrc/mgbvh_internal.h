// Internal device data layout of the mgbvh builder (not part of the C ABI).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mgbvh.h"
#include "mgbvh_common.h"

namespace mg {

constexpr uint32_t CH = 2048;                      // objects per chunk (k_bin / k_partition block)
constexpr uint32_t BIN_WORDS = 7;                  // count, ~okey(min) x3, okey(max) x3
constexpr uint32_t HIST_WORDS = 3 * 32 * BIN_WORDS;  // one node: 3 axes x 32 bins

constexpr uint32_t ERR_INDEX = 1;
constexpr uint32_t ERR_NONFINITE = 2;

// One object during the build: its box and original index (32 bytes, two dwordx4).
struct __align__(16) Rec {
  float mn[3];
  float mx[3];
  uint32_t id;
  uint32_t pad;
};

// A node with > 32 objects at the current breadth-first level.  Object
// positions, idx and P are absolute (over all meshes of a batched build);
// nbase / obase are the node slot and object position of the mesh's root, so
// a mesh's node array holds child = P - nbase and start_index = start - obase.
struct LNode {
  uint32_t start, count, idx, P;      // object range, final node index, pair base
  uint32_t depth, chunk0, nchunk, kind;
  float bmn[3], bmx[3];               // node bounds (SA of the parent in the cost)
  uint32_t cacc[6];                   // centroid bounds: ~okey(min) x3, okey(max) x3
  uint32_t axis, cut, n_lt, pad0;     // chosen split
  int32_t child_slot[2];              // next-level slot of (gt, lt) child, -1 if none
  uint32_t nbase, obase;
};

// A subtree root with <= 32 objects, or a leaf forced by max_depth.
struct SNode {
  uint32_t start, count, idx, P;
  uint32_t depth, parity, nbase, obase;
  float bmn[3], bmx[3];
  uint32_t pad2[2];
};

// One mesh of a (batched) build, device copy of the host-side plan.
struct MeshRow {
  uint32_t tri_offset, count, vertex_offset, vertex_count;
  uint32_t setup_chunk0;  // first setup chunk (chunks of CH objects per mesh)
  uint32_t nbase;         // first node slot: sum of 2 n - 1 over earlier meshes
  uint32_t root_slot;     // LNode slot (large root) or SNode slot (small root)
  uint32_t root_kind;     // 0 empty, 1 large root, 2 small root / leaf
  uint32_t lchunk0;       // first level-0 chunk of a large root
  uint32_t has_bounds;
  float bmn[3], bmx[3];
};

struct Counters {
  uint32_t n_large_next, n_chunks_next, n_small, height;
  uint32_t err, n_gaps, n_median, n_cutoff;
  uint32_t root_acc[12];  // bounds (~min, max) and centroid bounds (~min, max)
  unsigned long long sum_large;  // objects of all nodes split on the level path
  uint32_t pad[10];
};

struct NodeOut {
  mgbvh_node32* nodes;
  uint32_t* lens;
  mgbvh_split_rec* splits;
  __device__ inline void write_node(uint32_t idx, const Box& b, uint32_t start,
                                    uint32_t child, uint32_t len) const;
  __device__ inline void write_split(uint32_t idx, uint32_t am, uint32_t value, uint32_t nl,
                                     uint32_t cost_bits) const {
    if (splits) {
      mgbvh_split_rec r;
      r.axis_method = am;
      r.value = value;
      r.count_lt = nl;
      r.cost_bits = cost_bits;
      splits[idx] = r;
    }
  }
};

__device__ inline void NodeOut::write_node(uint32_t idx, const Box& b, uint32_t start,
                                           uint32_t child, uint32_t len) const {
  uint4* p = reinterpret_cast<uint4*>(&nodes[idx]);
  p[0] = make_uint4(__float_as_uint(b.mn[0]), __float_as_uint(b.mn[1]),
                    __float_as_uint(b.mn[2]), start);
  p[1] = make_uint4(__float_as_uint(b.mx[0]), __float_as_uint(b.mx[1]),
                    __float_as_uint(b.mx[2]), child);
  if (lens) lens[idx] = len;
}

}  // namespace mg

// ---------------------------------------------------------------- host side
#include <hip/hip_runtime.h>

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

namespace mg {

struct BuildInput {
  int kind;  // 0 = triangles, 1 = boxes
  const float* pos;
  uint32_t vstride_floats;
  uint64_t nverts;
  const uint32_t* tris;
  const float* boxes;
};

struct BuildOutput {
  mgbvh_node32* nodes;
  uint32_t* lens;
  uint32_t* perm;
  uint32_t* tris_out;
  mgbvh_split_rec* splits;
};

// One mesh of a batched build (host side).  Triangle vertex indices are local
// to the mesh (< vertex_count) and address vertex_offset + index.
struct MeshIn {
  uint32_t tri_offset, count, vertex_offset, vertex_count;
  int has_bounds;
  float bmn[3], bmx[3];
};

struct BuildResult {
  uint64_t nnodes;            // total over meshes
  uint32_t height, n_levels, n_median, n_cutoff, n_guard;
  uint64_t large_nodes, large_visits;  // nodes with > 32 objects, sum of their object counts
  float ms;
  std::vector<uint32_t> mesh_nodes;   // node count per mesh
  std::vector<uint32_t> mesh_nbase;   // first node slot per mesh
};

}  // namespace mg

// Kernel classes timed with HIP events when profiling is enabled.
enum KernelKind { KK_SETUP = 0, KK_ROOT, KK_BIN, KK_SWEEP, KK_PARTITION, KK_SMALL, KK_COMPACT,
                  KK_PERMUTE, KK_TRACE, KK_COUNT = 16 };

struct mgbvh_ctx {
  int device = 0;
  uint32_t prof_mask = 0;                 // bit k: time kernels of kind k
  std::vector<hipEvent_t> prof_pool;      // event pairs
  std::vector<int> prof_kind;             // kind of each recorded pair
  double prof_ms[KK_COUNT] = {};
  uint64_t prof_n[KK_COUNT] = {};
  hipStream_t stream = nullptr;
  hipEvent_t ev[64] = {};
  DevBuf recs[2], L[2], hist, blk, cn[2], ck[2], small, ctr, gaps, gtab, tmpn, tmpl, tmps, lens;
  DevBuf mrows, macc, mnb;
  DevBuf io_pos, io_tris, io_nodes, io_lens, io_perm, io_tris_out, io_splits, io_boxes;
  mg::Counters* h_ctr = nullptr;  // pinned
};

namespace mg {
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int ensure(DevBuf& b, size_t bytes);
int build_device(mgbvh_ctx* c, const BuildInput& in, uint32_t n, const mgbvh_settings* s,
                 const BuildOutput& o, BuildResult* res);
// Many object lists (meshes) built together: each gets the tree build_device
// would give it alone, in node slots [nbase, nbase + 2n - 1).
int build_multi(mgbvh_ctx* c, const BuildInput& in, const std::vector<MeshIn>& meshes,
                const mgbvh_settings* s, const BuildOutput& o, BuildResult* res);
// Event pair around one kernel launch (no-op unless kind is in ctx->prof_mask).
int prof_begin(mgbvh_ctx* c, int kind);
int prof_end(mgbvh_ctx* c, int kind);
int prof_collect(mgbvh_ctx* c);  // after a synchronize: accumulate and recycle
}  // namespace mg
