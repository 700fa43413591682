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
constexpr uint32_t ERR_INTERNAL = 4;   // a device-side invariant failed (split count out of range)

// One object during the build: its box and original index, 28 bytes, packed
// (4-byte aligned; read and written as a 16-byte and a 12-byte access).
struct Rec {
  float mn[3];
  float mx[3];
  uint32_t id;
};
static_assert(sizeof(Rec) == 28, "Rec is 28 bytes");

// A node with > 32 objects at the current breadth-first level.  Object
// positions, idx and P are absolute (over all meshes of a batched build);
// nbase / obase are the node slot and object position of the mesh's root, so
// a mesh's node array holds child = P - nbase and start_index = start - obase.
struct LNode {
  uint32_t start, count, idx, P;      // object range, final node index, pair base
  uint32_t depth, chunk0, nchunk, kind;
  float bmn[3], bmx[3];               // node bounds (SA of the parent in the cost)
  uint32_t cacc[6];                   // centroid bounds: ~okey(min) x3, okey(max) x3
  uint32_t axis, cut, n_lt, mode;     // chosen split; learned-method state (lmode_*)
  int32_t child_slot[2];              // next-level slot of (gt, lt) child, -1 if none
  uint32_t nbase, obase;
  uint32_t flags, pad3[3];            // LF_* (128 bytes in all)
};

// LNode::flags.  A node whose parent fits one chunk is binned by the parent's
// partition block (which knows the node's exact centroid bounds first), so
// its own binning pass is skipped.
constexpr uint32_t LF_PREBINNED = 1;
// ... and, outside sample collection, that block also takes the node's split
// decision from the histogram it holds in LDS and stores the decision (a
// SplitDec) at the start of the node's histogram row instead of the row.
constexpr uint32_t LF_DECIDED = 2;

// Learned-method state of a node with > 32 objects (LNode::mode):
// bits 0-1 kind (0 plain, 1 counted, 2 window, 3 fallback), bits 2-3 predicted
// axis, bits 4-8 first and bits 9-13 last cut of the window.
constexpr uint32_t LM_PLAIN = 0, LM_COUNTED = 1, LM_WINDOW = 2, LM_FALLBACK = 3;
__host__ __device__ inline uint32_t lmode_pack(uint32_t kind, uint32_t axis, uint32_t lo,
                                               uint32_t hi) {
  return kind | (axis << 2) | (lo << 4) | (hi << 9);
}
__host__ __device__ inline uint32_t lmode_kind(uint32_t m) { return m & 3u; }
__host__ __device__ inline uint32_t lmode_axis(uint32_t m) { return (m >> 2) & 3u; }
__host__ __device__ inline uint32_t lmode_lo(uint32_t m) { return (m >> 4) & 31u; }
__host__ __device__ inline uint32_t lmode_hi(uint32_t m) { return (m >> 9) & 31u; }

// Parameters of the learned build method inside one build.
struct LearnParams {
  uint32_t min_objects;  // 0: method off
  uint32_t window;
  float guard;
  uint32_t no_network;   // windows centred on the best proxy cut (ablation)
  // the network: the context's model of the call, or a scene plan's own copy
  // of the model as it was when the plan was created
  mgbvh_model* model = nullptr;
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
  uint32_t lrow;          // learned method: network row of a counted root
};

struct Counters {
  uint32_t n_small, err, n_gaps, n_median;
  uint32_t n_cutoff, n_samples, n_window, n_fallback;
  unsigned long long sum_large;  // objects of all nodes split with the binned sweep
  uint32_t n_med;                // medium-node list (one wavefront builds each subtree)
  uint32_t n_med_split;          // nodes with > 32 objects split inside those subtrees
  uint32_t n_med_next;           // next medium node to take (k_medium's work queue)
  uint32_t pad[11];
};

// The small-subtree list (n entries) is followed in the same buffer by the
// medium-node list: nodes of 33..Tuning::med_max objects, at most n / 33 of them.
__host__ __device__ constexpr uint64_t small_list_cap(uint64_t n) { return n + n / 33 + 2; }

// Per-level work counts in device memory (LV_WORDS words per level): the
// level kernels read their own level's counts and the sweep fills the next
// level's, so a whole build can be enqueued without reading counts back.
// LV_NBIN: chunks of nodes that k_bin bins (a node whose parent's partition
// fitted one block was binned there); 0 lets k_bin return at once.
constexpr uint32_t LV_NL = 0, LV_NC = 1, LV_NCOUNTED = 2, LV_NBIN = 3, LV_WORDS = 4;

// Split proposer network: 102 features -> 128 -> 128 (ReLU) -> 96 logits
// (axis * 32 + cut).  Features of a node with > 32 objects: log1p(count) /
// log1p(n) of its 3 x 32 centroid bins, node extents over the largest node
// extent, centroid extents over node extents.
constexpr int MLP_IN = 102, MLP_INP = 128, MLP_H = 128, MLP_OUT = 96;
constexpr int MLP_NPARAM = MLP_H * MLP_INP + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT;

// Feature 96 + a and 99 + a of a node (the 96 count features come from its bins).
__device__ inline void node_extent_features(const float* bmn, const float* bmx, const float* cext,
                                            float* f6) {
  float ext[3], emax = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ext[a] = bmx[a] - bmn[a];
    emax = fmaxf(emax, ext[a]);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    f6[a] = emax > 0.0f ? ext[a] / emax : 0.0f;
    f6[3 + a] = ext[a] > 0.0f ? fminf(cext[a] / ext[a], 1.0f) : 0.0f;
  }
}

struct NodeOut {
  mgbvh_node32* nodes;
  uint32_t* lens;
  mgbvh_split_rec* splits;
  float* samples = nullptr;      // split samples: MLP_IN features per binned split
  uint32_t* labels = nullptr;    // axis * 32 + cut of each sample
  uint32_t max_samples = 0, sample_min = 0;
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
#include <functional>

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool fixed = false;  // a view into caller-provided memory: never grown or freed
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
  float* samples = nullptr;
  uint32_t* labels = nullptr;
  uint32_t max_samples = 0, sample_min = 0;
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
  uint32_t n_samples;
  uint32_t n_window;                   // learned nodes split inside their window
  float ms;
  std::vector<uint32_t> mesh_nodes;   // node count per mesh
  std::vector<uint32_t> mesh_nbase;   // first node slot per mesh
};

}  // namespace mg

// Kernel classes timed with HIP events when profiling is enabled.
enum KernelKind { KK_SETUP = 0, KK_ROOT, KK_BIN, KK_SWEEP, KK_PARTITION, KK_SMALL, KK_COMPACT,
                  KK_PERMUTE, KK_TRACE, KK_LEARN, KK_MEDIUM, KK_COUNT = 16 };

namespace mg {
// Device scratch of one build: the context keeps one for its calls, a graph
// plan owns its own (a captured graph must keep its addresses).
struct Scratch {
  DevBuf recs[2], L[2], hist, blk, cn[2], ck[2], small, ctr, gaps, lens, mrows, macc, mnb;
  DevBuf lrows[2], lv, lfeat, llogit, gtab, tmpn, tmpl, tmps;
  // 8-wide collapse (its own: plans keep pointers): level-path records and
  // node lists, layout entries, the subtree-local path's tables, frontiers
  DevBuf w8rec, w8par, w8e, w8fast, w8front[2], w8ctr, w8base;
  Counters* h_ctr = nullptr;  // pinned
};
void scratch_free(Scratch& sc);

// Grid sizes of the level kernels, in blocks per compute unit.  The kernels
// loop over their work lists, but grids well past the resident block count
// measured faster for k_bin, k_partition and k_small: blocks that finish
// early free their slots for others, which balances uneven chunks and
// subtrees better than a fixed share per resident block.
struct Tuning {
  uint32_t bin_bpc = 128, part_bpc = 192, sweep_bpc = 4, small_bpc = 64;
  // wavefronts per block; part_waves 1 gives the partition one-wavefront
  // blocks below the first four levels (better when the level path went down
  // to small nodes; with the medium-node kernel taking those, four measured
  // faster on scenes and single meshes alike)
  uint32_t part_waves = 4, bin_waves = 4;
  uint32_t sweep_pack = 65536;             // sweep: one node per lane from this many nodes
  // nodes of 33..med_max objects build their whole subtree in one wavefront
  // (k_medium) instead of level by level; 0 = off.  512 measured best for
  // single meshes of 65 k-1 M objects (~7 % faster builds); a scene of
  // thousands of meshes has enough medium nodes to fill the GPU at 2048
  uint32_t med_max = 512, med_bpc = 64;
};
}  // namespace mg

struct mgbvh_ctx {
  int device = 0;
  int cus = 256;                          // compute units of the device
  uint32_t prof_mask = 0;                 // bit k: time kernels of kind k
  std::vector<hipEvent_t> prof_pool;      // event pairs
  std::vector<int> prof_kind;             // kind of each recorded pair
  double prof_ms[KK_COUNT] = {};
  uint64_t prof_n[KK_COUNT] = {};
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // a scene plan's TLAS branch, recorded beside the BLAS levels
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipEvent_t ev[64] = {};
  mg::Scratch sc;                 // scratch of this context's builds
  DevBuf io_pos, io_tris, io_nodes, io_lens, io_perm, io_tris_out, io_splits, io_boxes;
  mg::Counters* h_ctr = nullptr;  // pinned, the same as sc.h_ctr
  mgbvh_model* model = nullptr;   // network of the learned build method (not owned)
  mg::Tuning tune;
};

namespace mg {
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int ensure(DevBuf& b, size_t bytes);
// Address of this device's allocation-free build scratch (set up by
// mgbvh_ctx_create with the device current); 0 when unavailable.
int noalloc_init(int device);
int build_device(mgbvh_ctx* c, const BuildInput& in, uint32_t n, const mgbvh_settings* s,
                 const BuildOutput& o, BuildResult* res);
// Many object lists (meshes) built together: each gets the tree build_device
// would give it alone, in node slots [nbase, nbase + 2n - 1).
int build_multi(mgbvh_ctx* c, const BuildInput& in, const std::vector<MeshIn>& meshes,
                const mgbvh_settings* s, const BuildOutput& o, BuildResult* res);

// A build split into planning (sizes, slots, buffers), enqueueing (every
// kernel, either level by level with count read-backs or with a fixed level
// count for graph capture) and finishing (errors, depth-cut compaction).
struct BuildPlan {
  BuildInput in{};
  BuildOutput o{};
  mgbvh_settings s{};
  std::vector<MeshRow> rows;
  uint32_t nm = 0, n = 0, n_large = 0, n_small = 0, n_counted = 0;
  uint64_t lchunks = 0, setup_chunks = 0, nslots = 0;
  uint32_t capL = 0, capC = 0, max_levels = 0;
  LearnParams lp{0, 0, 0.0f, 0};
  uint32_t* lens = nullptr;  // o.lens or scratch
  // levels a recorded graph runs (0: not a graph); finish_build reports
  // MG_LEVELS_SHORT when nodes were left waiting at that level
  uint32_t fixed_levels = 0;
};
constexpr int MG_LEVELS_SHORT = 1000;  // internal: a graph ran too few levels
int plan_build(mgbvh_ctx* c, Scratch& sc, const BuildInput& in, const std::vector<MeshIn>& meshes,
               const mgbvh_settings* s, const BuildOutput& o, BuildPlan* p);
// fixed_levels = 0: level by level with read-backs; else exactly that many levels.
int enqueue_build(mgbvh_ctx* c, Scratch& sc, BuildPlan& p, uint32_t fixed_levels,
                  bool upload_rows, uint32_t* levels_run,
                  hipStream_t stream = nullptr,  // nullptr: c->stream
                  const std::function<int()>* after_setup = nullptr);
int finish_build(mgbvh_ctx* c, Scratch& sc, BuildPlan& p, BuildResult* res);
__global__ void k_fill(uint32_t* __restrict__ p, uint32_t v, uint64_t n);
__global__ void k_instance_boxes(const mgbvh_node32* __restrict__ nodes,
                                 const mgbvh_blas_desc* __restrict__ blas, uint32_t nblas,
                                 const float* __restrict__ o2w,
                                 const uint32_t* __restrict__ inst_blas, uint32_t ninst,
                                 mgbvh_aabb* __restrict__ out, uint32_t* err);
// Event pair around one kernel launch (no-op unless kind is in ctx->prof_mask).
int prof_begin(mgbvh_ctx* c, int kind, hipStream_t st = nullptr);  // nullptr: c->stream
int prof_end(mgbvh_ctx* c, int kind, hipStream_t st = nullptr);
int prof_collect(mgbvh_ctx* c);  // after a synchronize: accumulate and recycle
// Learned method, one level: features of the counted nodes, network
// inference, window and guard decision (LNode::mode -> window or fallback).
// n_dev: the row count in device memory; cap: its bound (grid sizes).
// A copy of the model's inference state for a scene plan (mgbvh_model_destroy frees it).
int model_snapshot(mgbvh_ctx* c, mgbvh_model* src, mgbvh_model** out);
int learned_decide(mgbvh_ctx* c, Scratch& sc, LNode* L, const uint32_t* rows,
                   const uint32_t* n_dev, uint32_t cap, const uint32_t* hist, Counters* ctr,
                   const LearnParams& lp);
}  // namespace mg
