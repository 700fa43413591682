// Scene-level pieces around the builder: instance world boxes for the TLAS.
#include <algorithm>
#include <vector>

#include "mgbvh_common.h"
#include "mgbvh_internal.h"

namespace mg {

// model.rs:189-221: the 8 corners of the mesh box through the instance
// transform (glam Mat4 * Vec4: x_axis*x + y_axis*y + z_axis*z + w_axis*w,
// summed left to right, w = 1), then the component-wise min and max.  The mesh
// box is the BLAS root's bounds (= mesh.bounds when the BLAS was built with
// them).  object_to_world: 3 rows x 4 floats per instance.
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
  const float* m = o2w + 12 * (size_t)i;
  float lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = __uint_as_float(0x7f800000u);
    hi[a] = __uint_as_float(0xff800000u);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float x = (k & 4) ? root.max[0] : root.min[0];
    const float y = (k & 2) ? root.max[1] : root.min[1];
    const float z = (k & 1) ? root.max[2] : root.min[2];
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
  out[i] = r;
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
  if ((rc = ensure(c->gtab, tab.size() * 4)) || (rc = ensure(c->tmps, (size_t)nblas * 8))) return rc;
  hipStream_t st = c->stream;
  e = hipMemcpyAsync(c->gtab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(c->tmps.p, 0, (size_t)nblas * 8, st);
  if (e != hipSuccess) return hip_fail(e, "tree sah setup");
  const uint32_t* t = (const uint32_t*)c->gtab.p;
  if (nslots)
    k_tree_sah<<<(uint32_t)((nslots + 255) / 256), 256, 0, st>>>(
        d_nodes, d_lens, t, t + nblas, nblas, (uint32_t)nslots, (double*)c->tmps.p);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_sah, c->tmps.p, (size_t)nblas * 8, hipMemcpyDeviceToHost, st);
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
  int rc = ensure(c->ctr, sizeof(Counters));
  if (rc) return rc;
  uint32_t* flag = reinterpret_cast<uint32_t*>(c->ctr.p);
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
