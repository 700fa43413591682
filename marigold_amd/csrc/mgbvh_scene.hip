// Scene-level pieces around the builder: instance world boxes for the TLAS.
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

}  // namespace mg

using namespace mg;

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
