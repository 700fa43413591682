// Split proposer network of the learned build method: a 102 -> 128 -> 128 ->
// 96 MLP mapping a node's centroid-bin counts and extents to split logits
// (axis * 32 + cut).  Trained here on the split decisions the plain builder
// emits (FP32 master weights, hand-written tiled GEMMs, softmax cross-entropy,
// Adam); inference runs on MFMA with BF16 weights and activations and FP32
// accumulation (v_mfma_f32_16x16x16_bf16).
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "mgbvh_common.h"
#include "mgbvh_internal.h"

// Parameter layout (floats): W1 [128][128] (inputs 102..127 are zero
// padding), b1 [128], W2 [128][128], b2 [128], W3 [96][128], b3 [96].
namespace mg {
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_INP;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_OUT * MLP_H;
static_assert(OFF_B3 + MLP_OUT == MLP_NPARAM, "parameter layout");
constexpr int NW_BF16 = MLP_H * MLP_INP + MLP_H * MLP_H + MLP_OUT * MLP_H;  // bf16 weights
}  // namespace mg

struct mgbvh_model {
  int device = 0;
  DevBuf params, grads, adam_m, adam_v, wbf, act, bidx, blab, stats;
  uint32_t step = 0;
  bool bf_valid = false;
};

namespace mg {

// ------------------------------------------------------------ FP32 training
struct GemmArgs {
  int M, N, K;
  const float* A;
  int lda, ta;  // A(m, k) = ta ? A[k * lda + m] : A[m * lda + k]
  const float* B;
  int ldb, tb;  // B(k, n) = tb ? B[n * ldb + k] : B[k * ldb + n]
  float* C;
  int ldc;
  const float* bias;  // + bias[n]
  int relu;           // max(0, .)
  const float* mask;  // * (mask[m * ldm + n] > 0)
  int ldm;
};

// C = epilogue(A * B), 16 x 16 output tile per block, K staged through LDS.
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ float As[16][17], Bs[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m = blockIdx.y * 16 + ty, n = blockIdx.x * 16 + tx;
  float acc = 0.0f;
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    const int am = blockIdx.y * 16 + ty, ak = k0 + tx;
    As[ty][tx] = (am < g.M && ak < g.K) ? (g.ta ? g.A[(size_t)ak * g.lda + am]
                                                : g.A[(size_t)am * g.lda + ak])
                                        : 0.0f;
    const int bk = k0 + ty, bn = blockIdx.x * 16 + tx;
    Bs[ty][tx] = (bk < g.K && bn < g.N) ? (g.tb ? g.B[(size_t)bn * g.ldb + bk]
                                                : g.B[(size_t)bk * g.ldb + bn])
                                        : 0.0f;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) acc = fmaf(As[ty][kk], Bs[kk][tx], acc);
    __syncthreads();
  }
  if (m < g.M && n < g.N) {
    if (g.bias) acc += g.bias[n];
    if (g.relu) acc = fmaxf(acc, 0.0f);
    if (g.mask && !(g.mask[(size_t)m * g.ldm + n] > 0.0f)) acc = 0.0f;
    g.C[(size_t)m * g.ldc + n] = acc;
  }
}

// out[n] = sum over rows of Y[rows][N] (bias gradients).
// Column sums of a rows x N row-major matrix with row stride ld: a 1024-thread
// block takes 64 columns, its 16 wavefronts every 16th row each (coalesced
// 256-byte row segments), then one wavefront adds the 16 partial sums in a
// fixed order, so the result is the same on every run.
__global__ __launch_bounds__(1024) void k_colsum(const float* __restrict__ Y, int rows, int N,
                                                 int ld, float* __restrict__ out) {
  __shared__ float part[16][64];
  const int col = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + col;
  float s = 0.0f;
  if (n < N)
    for (int r = g; r < rows; r += 16) s += Y[(size_t)r * ld + n];
  part[g][col] = s;
  __syncthreads();
  if (g == 0 && n < N) {
    float t = 0.0f;
    for (int k = 0; k < 16; ++k) t += part[k][col];
    out[n] = t;
  }
}

// Minibatch rows: X[i][0..127] = features[idx[i]] zero-padded, lab[i].
__global__ void k_gather(const float* __restrict__ feat, const uint32_t* __restrict__ labels,
                         const uint32_t* __restrict__ idx, int B, float* __restrict__ X,
                         uint32_t* __restrict__ lab) {
  const int i = blockIdx.x, k = threadIdx.x;  // 128 threads
  if (i >= B) return;
  const uint32_t r = idx ? idx[i] : (uint32_t)i;
  X[(size_t)i * MLP_INP + k] = k < MLP_IN ? feat[(size_t)r * MLP_IN + k] : 0.0f;
  if (k == 0) lab[i] = labels[r];
}

// Softmax cross-entropy over 96 logits: dZ = (softmax - onehot) / B; stats =
// {loss sum, correct count}.
__global__ void k_softmax_ce(const float* __restrict__ Z, const uint32_t* __restrict__ lab, int B,
                             float* __restrict__ dZ, float* stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const float* z = Z + (size_t)i * MLP_OUT;
  float mx = z[0];
  int am = 0;
  for (int k = 1; k < MLP_OUT; ++k)
    if (z[k] > mx) { mx = z[k]; am = k; }
  float se = 0.0f;
  for (int k = 0; k < MLP_OUT; ++k) se += expf(z[k] - mx);
  const uint32_t y = lab[i];
  const float inv = 1.0f / (float)B;
  for (int k = 0; k < MLP_OUT; ++k) {
    const float p = expf(z[k] - mx) / se;
    dZ[(size_t)i * MLP_OUT + k] = (p - (k == (int)y ? 1.0f : 0.0f)) * inv;
  }
  atomicAdd(&stats[0], logf(se) + mx - z[y]);
  if (am == (int)y) atomicAdd(&stats[1], 1.0f);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// ------------------------------------------------------------ BF16 inference
__device__ inline uint16_t f2bf(float f) {  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ void k_to_bf16(const float* __restrict__ params, uint16_t* __restrict__ wbf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NW_BF16) return;
  int src;
  if (i < MLP_H * MLP_INP) src = OFF_W1 + i;
  else if (i < MLP_H * MLP_INP + MLP_H * MLP_H) src = OFF_W2 + (i - MLP_H * MLP_INP);
  else src = OFF_W3 + (i - MLP_H * MLP_INP - MLP_H * MLP_H);
  wbf[i] = f2bf(params[src]);
}

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

// LDS row stride of weights and activations: 128 BF16 values + 8 of padding
// (68 words), so the 16 rows one MFMA operand read touches fall in different
// banks.
constexpr int LDS_ROW = MLP_INP + 8;
constexpr int W_ROWS = 2 * MLP_H + MLP_OUT;  // W1, W2, W3 rows of 128 inputs each

// One dense layer for the 16 rows of a wave: out[16][N] = act(in[16][128] * W^T + b),
// weights and activations in LDS (row stride LDS_ROW).
// MFMA 16x16x16 bf16 lane layout: A[i = l % 16][k = 4 (l / 16) + j],
// B[k = 4 (l / 16) + j][n = l % 16], D[i = 4 (l / 16) + j][n = l % 16].
template <int N, bool LAST>
__device__ inline void mlp_layer(const uint16_t* in_lds, const uint16_t* W,
                                 const float* __restrict__ bias, uint16_t* out_lds,
                                 float* __restrict__ out_logits, int row0, int nrows) {
  const int l = lane_id();
  const int r = l & 15, q = l >> 4;
#pragma unroll
  for (int t = 0; t < N / 16; ++t) {
    v4f32 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int n = 16 * t + r;
#pragma unroll
    for (int kc = 0; kc < MLP_INP / 16; ++kc) {
      const int k = 16 * kc + 4 * q;
      v4i16 a, b;
      const uint2 av = *reinterpret_cast<const uint2*>(&in_lds[r * LDS_ROW + k]);
      const uint2 bv = *reinterpret_cast<const uint2*>(&W[n * LDS_ROW + k]);
      a[0] = (short)(av.x & 0xffff); a[1] = (short)(av.x >> 16);
      a[2] = (short)(av.y & 0xffff); a[3] = (short)(av.y >> 16);
      b[0] = (short)(bv.x & 0xffff); b[1] = (short)(bv.x >> 16);
      b[2] = (short)(bv.y & 0xffff); b[3] = (short)(bv.y >> 16);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const float bn = bias[n];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = 4 * q + j;
      const float z = acc[j] + bn;
      if (LAST) {
        if (i < nrows) out_logits[(size_t)(row0 + i) * MLP_OUT + n] = z;
      } else {
        out_lds[i * LDS_ROW + n] = f2bf(fmaxf(z, 0.0f));
      }
    }
  }
}

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Logits of the (up to) 16 feature rows from row0 (MLP_IN floats each, row
// stride `stride` floats).  One wavefront; activations staged in LDS as BF16.
__device__ inline void mlp_tile(int row0, int n, uint16_t* A, uint16_t* Bf, const uint16_t* Wl,
                                const float* __restrict__ feat, int stride,
                                const float* __restrict__ params, float* __restrict__ logits) {
  const int l = lane_id();
  const int nrows = min(16, n - row0);
  for (int e = l; e < 16 * MLP_INP; e += 64) {
    const int i = e / MLP_INP, k = e % MLP_INP;
    const float x = (i < nrows && k < MLP_IN) ? feat[(size_t)(row0 + i) * stride + k] : 0.0f;
    A[i * LDS_ROW + k] = f2bf(x);
  }
  wave_lds_sync();
  const uint16_t* W1 = Wl;
  const uint16_t* W2 = Wl + MLP_H * LDS_ROW;
  const uint16_t* W3 = W2 + MLP_H * LDS_ROW;
  mlp_layer<MLP_H, false>(A, W1, params + OFF_B1, Bf, nullptr, row0, nrows);
  wave_lds_sync();
  mlp_layer<MLP_H, false>(Bf, W2, params + OFF_B2, A, nullptr, row0, nrows);
  wave_lds_sync();
  mlp_layer<MLP_OUT, true>(A, W3, params + OFF_B3, nullptr, logits, row0, nrows);
  wave_lds_sync();  // the next tile overwrites A
}

constexpr size_t MLP_LDS_BYTES = (size_t)(W_ROWS + 4 * 2 * 16) * LDS_ROW * 2;

// Logits of n feature rows (MLP_IN floats each, row stride `stride` floats);
// n from n_dev when given.  Each block first copies the BF16 weights into LDS
// (one read of 88 KB, then every MFMA operand comes from LDS); then one
// wavefront per 16 rows, persistent over tiles.  The 130,560 bytes of LDS are
// a static allocation (gfx950 gives a workgroup up to 160 KB), so launches
// need no function attribute and no dynamic LDS size, in graphs too.
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ feat, int stride,
                                                   int n_host, const uint32_t* __restrict__ n_dev,
                                                   const uint16_t* __restrict__ wbf,
                                                   const float* __restrict__ params,
                                                   float* __restrict__ logits) {
  __shared__ __align__(16) uint16_t lds[MLP_LDS_BYTES / 2];
  uint16_t* Wl = lds;                                   // W_ROWS x LDS_ROW
  uint16_t* act = lds + W_ROWS * LDS_ROW;               // 4 waves x 2 x 16 x LDS_ROW
  const int w = threadIdx.x >> 6;
  const int n = n_dev ? (int)*n_dev : n_host;
  if (blockIdx.x * 64 >= (unsigned)n) return;          // block-uniform: no tile for this block
  // 16-byte pieces: 16 per 128-value row
  for (int e = threadIdx.x; e < W_ROWS * 16; e += blockDim.x) {
    const int row = e >> 4, c = e & 15;
    *reinterpret_cast<uint4*>(&Wl[row * LDS_ROW + 8 * c]) =
        *reinterpret_cast<const uint4*>(&wbf[(size_t)row * MLP_INP + 8 * c]);
  }
  __syncthreads();
  uint16_t* A = act + (size_t)w * 2 * 16 * LDS_ROW;
  uint16_t* Bf = A + 16 * LDS_ROW;
  for (int row0 = (blockIdx.x * 4 + w) * 16; row0 < n; row0 += gridDim.x * 64)
    mlp_tile(row0, n, A, Bf, Wl, feat, stride, params, logits);
}

// ------------------------------------------------------------ learned build step
// One wavefront per network row (a counted node, listed when it was created):
// its features.
__device__ inline void feat_row(uint32_t row, const LNode* __restrict__ L,
                                const uint32_t* __restrict__ rows,
                                const uint32_t* __restrict__ hist, float* __restrict__ feat) {
  const uint32_t lane = lane_id();
  const uint32_t slot = rows[row];
  const LNode& nd = L[slot];
  const uint32_t* h = hist + (size_t)slot * HIST_WORDS;
  float* f = feat + (size_t)row * MLP_IN;
  const float inv = 1.0f / log1pf((float)nd.count);
  for (uint32_t q = lane; q < 3 * MG_NBINS; q += 64) f[q] = log1pf((float)h[q * BIN_WORDS]) * inv;
  if (lane == 0) {
    float cext[3];
    for (int a = 0; a < 3; ++a) cext[a] = acc_max_dec(nd.cacc[3 + a]) - acc_min_dec(nd.cacc[a]);
    node_extent_features(nd.bmn, nd.bmx, cext, f + 3 * MG_NBINS);
  }
}

// Persistent over the level's network rows (count in device memory).
__global__ __launch_bounds__(256) void k_feat(const LNode* __restrict__ L,
                                              const uint32_t* __restrict__ rows,
                                              const uint32_t* __restrict__ n_dev,
                                              const uint32_t* __restrict__ hist,
                                              float* __restrict__ feat) {
  const uint32_t n = *n_dev;
  for (uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6); row < n; row += gridDim.x * 4)
    feat_row(row, L, rows, hist, feat);
}

// One wavefront per network row: the best valid (axis, cut) by the logits
// (valid = both sides non-empty and a non-zero centroid extent, exactly the
// binned sweep's rule, so a window always holds a valid cut), the window of
// `win` cuts either side, and the guard: the proxy SAH of a cut puts every
// object of a side inside the node box clipped at the cut plane; a window
// whose best proxy exceeds (1 + guard) x the best proxy over all valid cuts
// sends the node to the plain sweep.
__device__ inline void decide_row(uint32_t row, LNode* __restrict__ L,
                                  const uint32_t* __restrict__ list,
                                  const uint32_t* __restrict__ hist,
                                  const float* __restrict__ logits, uint32_t win, float guard,
                                  Counters* ctr) {
  const uint32_t lane = lane_id();
  const uint32_t slot = list[row];
  const LNode& nd = L[slot];
  const uint32_t* h = hist + (size_t)slot * HIST_WORDS;
  float cmin[3], cext[3];
  Box nb;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    cmin[a] = acc_min_dec(nd.cacc[a]);
    cext[a] = __fsub_rn(acc_max_dec(nd.cacc[3 + a]), cmin[a]);
    nb.mn[a] = nd.bmn[a];
    nb.mx[a] = nd.bmx[a];
  }
  const float sa_p = box_sa(nb);
  uint64_t bestk = ~0ull;
  bool valid[2];
  uint32_t nlt[2];
  float proxy[2];
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const uint32_t q = lane + 64 * pass;
    const bool qact = q < 3 * MG_NBINS;
    const uint32_t b = q % MG_NBINS;
    const int a = qact ? (int)(q / MG_NBINS) : 0;
    const uint32_t cnt = qact ? h[q * BIN_WORDS] : 0u;
    uint32_t pc = cnt;  // inclusive prefix over the axis' bins
#pragma unroll
    for (uint32_t d = 1; d < MG_NBINS; d <<= 1) {
      const uint32_t o = __shfl_up(pc, d, MG_NBINS);
      if (b >= d) pc += o;
    }
    const uint32_t lt = pc - cnt;  // bins [0, b)
    const uint32_t gtc = nd.count - lt;
    const bool v = qact && b >= 1 && cext[a] != 0.0f && lt >= 1 && gtc >= 1;
    valid[pass] = v;
    nlt[pass] = lt;
    float pr = 0.0f;
    if (v) {
      const float plane = cmin[a] + cext[a] * ((float)b * (1.0f / MG_NBINS));
      Box bl = nb, bg = nb;
      bl.mx[a] = fmin_x86(fmax_x86(plane, nb.mn[a]), nb.mx[a]);
      bg.mn[a] = bl.mx[a];
      pr = split_cost(sa_p, box_sa(bl), box_sa(bg), lt, gtc);
      // largest logit (or, without a network, smallest proxy), lowest class
      const uint32_t k32 = logits ? ~okey(logits[(size_t)row * MLP_OUT + q]) : okey(pr);
      const uint64_t key = ((uint64_t)k32 << 32) | q;
      bestk = key < bestk ? key : bestk;
    }
    proxy[pass] = pr;
  }
  bestk = wave_min_u64_dpp(bestk, lane);
  uint32_t mode = LM_FALLBACK;
  if (bestk != ~0ull) {
    const uint32_t qb = (uint32_t)bestk;
    const uint32_t pa = qb / MG_NBINS, pcut = qb % MG_NBINS;
    const uint32_t lo = pcut > win ? pcut - win : 1u;
    const uint32_t hi = min(pcut + win, (uint32_t)MG_NBINS - 1);
    uint64_t all = ~0ull, inw = ~0ull;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const uint32_t q = lane + 64 * pass;
      if (!valid[pass]) continue;
      const uint64_t k = okey(proxy[pass]);
      all = k < all ? k : all;
      const uint32_t a = q / MG_NBINS, b = q % MG_NBINS;
      if (a == pa && b >= lo && b <= hi) inw = k < inw ? k : inw;
    }
    all = wave_min_u64_dpp(all, lane);
    inw = wave_min_u64_dpp(inw, lane);
    const float pall = okey_inv((uint32_t)all), pwin = okey_inv((uint32_t)inw);
    if (inw != ~0ull && pwin <= pall * (1.0f + guard)) mode = lmode_pack(LM_WINDOW, pa, lo, hi);
  }
  // (window / fallback nodes are counted by the sweep, one atomic per block)
  if (lane == 0) L[slot].mode = mode;
}

__global__ __launch_bounds__(256) void k_decide(LNode* __restrict__ L,
                                                const uint32_t* __restrict__ list,
                                                const uint32_t* __restrict__ n_dev,
                                                const uint32_t* __restrict__ hist,
                                                const float* __restrict__ logits, uint32_t win,
                                                float guard, Counters* ctr) {
  const uint32_t n = *n_dev;
  for (uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6); row < n; row += gridDim.x * 4)
    decide_row(row, L, list, hist, logits, win, guard, ctr);
}

// ------------------------------------------------------------ host helpers
static inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

static int gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, int ta, const float* B,
                int ldb, int tb, float* C, int ldc, const float* bias, int relu, const float* mask,
                int ldm) {
  GemmArgs g{M, N, K, A, lda, ta, B, ldb, tb, C, ldc, bias, relu, mask, ldm};
  k_gemm<<<dim3(cdivu(N, 16), cdivu(M, 16)), 256, 0, st>>>(g);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_gemm");
}

// Activation scratch for a batch of B rows: X, H1, H2 (B x 128), Z3, dZ3
// (B x 96), dH (B x 128) twice.
struct Acts {
  float *X, *H1, *H2, *Z3, *dZ3, *dA, *dB;
};

static int ensure_acts(mgbvh_model* m, int B, Acts* a) {
  const size_t per = (size_t)B * (5 * MLP_INP + 2 * MLP_OUT);
  int rc = ensure(m->act, per * 4);
  if (rc) return rc;
  float* p = (float*)m->act.p;
  a->X = p; p += (size_t)B * MLP_INP;
  a->H1 = p; p += (size_t)B * MLP_INP;
  a->H2 = p; p += (size_t)B * MLP_INP;
  a->Z3 = p; p += (size_t)B * MLP_OUT;
  a->dZ3 = p; p += (size_t)B * MLP_OUT;
  a->dA = p; p += (size_t)B * MLP_INP;
  a->dB = p;
  return MGBVH_OK;
}

// Forward + backward of one batch already gathered into a.X / lab; gradients
// into m->grads, {loss sum, correct} added to m->stats.
static int fwd_bwd(hipStream_t st, mgbvh_model* m, int B, const Acts& a, const uint32_t* lab) {
  const float* P = (const float*)m->params.p;
  float* G = (float*)m->grads.p;
  int rc;
  // forward
  if ((rc = gemm(st, B, MLP_H, MLP_INP, a.X, MLP_INP, 0, P + OFF_W1, MLP_INP, 1, a.H1, MLP_INP,
                 P + OFF_B1, 1, nullptr, 0))) return rc;
  if ((rc = gemm(st, B, MLP_H, MLP_H, a.H1, MLP_INP, 0, P + OFF_W2, MLP_H, 1, a.H2, MLP_INP,
                 P + OFF_B2, 1, nullptr, 0))) return rc;
  if ((rc = gemm(st, B, MLP_OUT, MLP_H, a.H2, MLP_INP, 0, P + OFF_W3, MLP_H, 1, a.Z3, MLP_OUT,
                 P + OFF_B3, 0, nullptr, 0))) return rc;
  k_softmax_ce<<<cdivu(B, 256), 256, 0, st>>>(a.Z3, lab, B, a.dZ3, (float*)m->stats.p);
  // backward: dW3 = dZ3^T H2, db3, dZ2 = (dZ3 W3) * (H2 > 0)
  if ((rc = gemm(st, MLP_OUT, MLP_H, B, a.dZ3, MLP_OUT, 1, a.H2, MLP_INP, 0, G + OFF_W3, MLP_H,
                 nullptr, 0, nullptr, 0))) return rc;
  k_colsum<<<cdivu(MLP_OUT, 64), 1024, 0, st>>>(a.dZ3, B, MLP_OUT, MLP_OUT, G + OFF_B3);
  if ((rc = gemm(st, B, MLP_H, MLP_OUT, a.dZ3, MLP_OUT, 0, P + OFF_W3, MLP_H, 0, a.dA, MLP_INP,
                 nullptr, 0, a.H2, MLP_INP))) return rc;
  // dW2 = dZ2^T H1, db2, dZ1 = (dZ2 W2) * (H1 > 0)
  if ((rc = gemm(st, MLP_H, MLP_H, B, a.dA, MLP_INP, 1, a.H1, MLP_INP, 0, G + OFF_W2, MLP_H,
                 nullptr, 0, nullptr, 0))) return rc;
  k_colsum<<<cdivu(MLP_H, 64), 1024, 0, st>>>(a.dA, B, MLP_H, MLP_INP, G + OFF_B2);
  if ((rc = gemm(st, B, MLP_H, MLP_H, a.dA, MLP_INP, 0, P + OFF_W2, MLP_H, 0, a.dB, MLP_INP,
                 nullptr, 0, a.H1, MLP_INP))) return rc;
  // dW1 = dZ1^T X, db1
  if ((rc = gemm(st, MLP_H, MLP_INP, B, a.dB, MLP_INP, 1, a.X, MLP_INP, 0, G + OFF_W1, MLP_INP,
                 nullptr, 0, nullptr, 0))) return rc;
  k_colsum<<<cdivu(MLP_H, 64), 1024, 0, st>>>(a.dB, B, MLP_H, MLP_INP, G + OFF_B1);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "fwd_bwd");
}

int model_bf16(mgbvh_ctx* c, mgbvh_model* m) {
  if (m->bf_valid) return MGBVH_OK;
  int rc = ensure(m->wbf, (size_t)NW_BF16 * 2);
  if (rc) return rc;
  k_to_bf16<<<cdivu(NW_BF16, 256), 256, 0, c->stream>>>((const float*)m->params.p,
                                                         (uint16_t*)m->wbf.p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "k_to_bf16");
  m->bf_valid = true;
  return MGBVH_OK;
}

// n rows (or, with n_dev, at most n rows, the count read on the device).
int model_infer(mgbvh_ctx* c, mgbvh_model* m, const float* feat, int stride, int n,
                const uint32_t* n_dev, float* logits) {
  if (n <= 0) return MGBVH_OK;
  int rc = model_bf16(c, m);
  if (rc) return rc;
  // one block per compute unit at most (the weights fill most of its LDS)
  const unsigned grid = std::max(1u, std::min<unsigned>(cdivu(n, 64), (unsigned)c->cus));
  k_mlp_infer<<<grid, 256, 0, c->stream>>>(feat, stride, n, n_dev,
                                                       (const uint16_t*)m->wbf.p,
                                                       (const float*)m->params.p, logits);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_mlp_infer");
}

// A copy of a model's inference state (FP32 parameters, whose biases the
// inference reads, and the BF16 weights), owned by whoever keeps it: a scene
// plan uses the model as it was when the plan was recorded, whatever happens
// to the context's model afterwards.
int model_snapshot(mgbvh_ctx* c, mgbvh_model* src, mgbvh_model** out) {
  *out = nullptr;
  int rc = model_bf16(c, src);
  if (rc) return rc;
  mgbvh_model* m = new mgbvh_model();
  m->device = src->device;
  if ((rc = ensure(m->params, (size_t)MLP_NPARAM * 4)) ||
      (rc = ensure(m->wbf, (size_t)NW_BF16 * 2))) {
    (void)mgbvh_model_destroy(m);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(m->params.p, src->params.p, (size_t)MLP_NPARAM * 4,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->wbf.p, src->wbf.p, (size_t)NW_BF16 * 2, hipMemcpyDeviceToDevice,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    (void)mgbvh_model_destroy(m);
    return hip_fail(e, "model snapshot");
  }
  m->bf_valid = true;
  *out = m;
  return MGBVH_OK;
}

int learned_decide(mgbvh_ctx* c, Scratch& sc, LNode* L, const uint32_t* list,
                   const uint32_t* n_dev, uint32_t cap, const uint32_t* hist, Counters* ctr,
                   const LearnParams& lp) {
  int rc;
  if ((rc = ensure(sc.lfeat, (size_t)cap * MLP_IN * 4)) ||
      (rc = ensure(sc.llogit, (size_t)cap * MLP_OUT * 4)))
    return rc;
  hipStream_t st = c->stream;
  float* feat = (float*)sc.lfeat.p;
  float* logits = (float*)sc.llogit.p;
  const unsigned grid = std::max(1u, std::min<unsigned>(cdivu(cap, 4), (unsigned)c->cus * 8));
  k_feat<<<grid, 256, 0, st>>>(L, list, n_dev, hist, feat);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "k_feat");
  if (!lp.no_network) {
    if (!lp.model) { set_error("the learned method needs a model"); return MGBVH_ERR_UNSUPPORTED; }
    if ((rc = model_infer(c, lp.model, feat, MLP_IN, (int)cap, n_dev, logits))) return rc;
  }
  k_decide<<<grid, 256, 0, st>>>(L, list, n_dev, hist, lp.no_network ? nullptr : logits,
                                 lp.window, lp.guard, ctr);
  e = hipGetLastError();
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_decide");
}

}  // namespace mg

using namespace mg;

// ------------------------------------------------------------ C ABI
extern "C" {

int mgbvh_model_create(mgbvh_ctx* c, uint64_t seed, mgbvh_model** out) {
  if (!c || !out) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  *out = nullptr;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  mgbvh_model* m = new mgbvh_model();
  m->device = c->device;
  int rc = 0;
  const size_t pb = (size_t)MLP_NPARAM * 4;
  if ((rc = ensure(m->params, pb)) || (rc = ensure(m->grads, pb)) || (rc = ensure(m->adam_m, pb)) ||
      (rc = ensure(m->adam_v, pb)) || (rc = ensure(m->stats, 16))) {
    mgbvh_model_destroy(m);
    return rc;
  }
  // Kaiming-uniform weights (fan-in 102, 128, 128), zero biases and zero padding
  std::vector<float> h(MLP_NPARAM, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int rows, int cols, int fan_in) {
    const double lim = std::sqrt(6.0 / fan_in);
    for (int r = 0; r < rows; ++r)
      for (int k = 0; k < fan_in; ++k)
        h[off + r * cols + k] = (float)((double)(rng() >> 11) * (1.0 / 9007199254740992.0) * 2.0 * lim - lim);
  };
  fill(OFF_W1, MLP_H, MLP_INP, MLP_IN);
  fill(OFF_W2, MLP_H, MLP_H, MLP_H);
  fill(OFF_W3, MLP_OUT, MLP_H, MLP_H);
  e = hipMemcpy(m->params.p, h.data(), pb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->adam_m.p, 0, pb);
  if (e == hipSuccess) e = hipMemset(m->adam_v.p, 0, pb);
  if (e != hipSuccess) {
    mgbvh_model_destroy(m);
    return hip_fail(e, "model init");
  }
  *out = m;
  return MGBVH_OK;
}

int mgbvh_model_destroy(mgbvh_model* m) {
  if (!m) return MGBVH_OK;
  hipSetDevice(m->device);
  for (DevBuf* b : {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->wbf, &m->act, &m->bidx,
                    &m->blab, &m->stats})
    if (b->p) hipFree(b->p);
  delete m;
  return MGBVH_OK;
}

int mgbvh_model_get_params(mgbvh_ctx* c, mgbvh_model* m, float* host) {
  if (!c || !m || !host) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipMemcpy(host, m->params.p, (size_t)MLP_NPARAM * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipMemcpy(params)");
}

int mgbvh_model_set_params(mgbvh_ctx* c, mgbvh_model* m, const float* host) {
  if (!c || !m || !host) { set_error("null argument"); return MGBVH_ERR_INVALID_ARG; }
  hipError_t e = hipMemcpy(m->params.p, host, (size_t)MLP_NPARAM * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(params)");
  m->bf_valid = false;
  m->step = 0;
  e = hipMemset(m->adam_m.p, 0, (size_t)MLP_NPARAM * 4);
  if (e == hipSuccess) e = hipMemset(m->adam_v.p, 0, (size_t)MLP_NPARAM * 4);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "hipMemset(adam)");
}

int mgbvh_model_gradients_device(mgbvh_ctx* c, mgbvh_model* m, const float* d_features,
                                 const uint32_t* d_labels, uint32_t n, float* d_grads,
                                 float* out_loss) {
  if (!c || !m || !d_features || !d_labels || !n) {
    set_error("null argument or no samples");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  Acts a;
  int rc = ensure_acts(m, (int)n, &a);
  if (rc || (rc = ensure(m->blab, (size_t)n * 4))) return rc;
  hipStream_t st = c->stream;
  e = hipMemsetAsync(m->stats.p, 0, 16, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
  k_gather<<<n, 128, 0, st>>>(d_features, d_labels, nullptr, (int)n, a.X, (uint32_t*)m->blab.p);
  if ((rc = fwd_bwd(st, m, (int)n, a, (const uint32_t*)m->blab.p))) return rc;
  if (d_grads) {
    e = hipMemcpyAsync(d_grads, m->grads.p, (size_t)MLP_NPARAM * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(grads)");
  }
  float hs[4];
  e = hipMemcpyAsync(hs, m->stats.p, 16, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  if (out_loss) *out_loss = hs[0] / (float)n;
  return MGBVH_OK;
}

int mgbvh_model_train_device(mgbvh_ctx* c, mgbvh_model* m, const float* d_features,
                             const uint32_t* d_labels, uint32_t n, uint32_t epochs, uint32_t batch,
                             float lr, uint64_t seed, float* out_loss, float* out_accuracy) {
  if (!c || !m || !d_features || !d_labels || !n || !batch || !epochs) {
    set_error("null argument, no samples, zero batch or zero epochs");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const int B = (int)std::min<uint32_t>(batch, n);
  Acts a;
  int rc = ensure_acts(m, B, &a);
  if (rc || (rc = ensure(m->bidx, (size_t)n * 4)) || (rc = ensure(m->blab, (size_t)B * 4)))
    return rc;
  hipStream_t st = c->stream;
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; ++i) perm[i] = i;
  std::mt19937_64 rng(seed);
  float last_loss = 0.0f, last_acc = 0.0f;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    for (uint32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng() % (i + 1)]);
    e = hipMemcpyAsync(m->bidx.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(m->stats.p, 0, 16, st);
    if (e != hipSuccess) return hip_fail(e, "epoch setup");
    uint32_t seen = 0;
    for (uint32_t b0 = 0; b0 + (uint32_t)B <= n; b0 += (uint32_t)B) {
      k_gather<<<B, 128, 0, st>>>(d_features, d_labels, (const uint32_t*)m->bidx.p + b0, B, a.X,
                                  (uint32_t*)m->blab.p);
      if ((rc = fwd_bwd(st, m, B, a, (const uint32_t*)m->blab.p))) return rc;
      ++m->step;
      const float c1 = 1.0f - std::pow(0.9f, (float)m->step);
      const float c2 = 1.0f - std::pow(0.999f, (float)m->step);
      k_adam<<<cdivu(MLP_NPARAM, 256), 256, 0, st>>>((float*)m->params.p, (const float*)m->grads.p,
                                                     (float*)m->adam_m.p, (float*)m->adam_v.p,
                                                     MLP_NPARAM, lr, c1, c2);
      seen += (uint32_t)B;
    }
    float hs[4];
    e = hipMemcpyAsync(hs, m->stats.p, 16, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    last_loss = hs[0] / (float)seen;
    last_acc = hs[1] / (float)seen;
  }
  m->bf_valid = false;
  if (out_loss) *out_loss = last_loss;
  if (out_accuracy) *out_accuracy = last_acc;
  return MGBVH_OK;
}

int mgbvh_model_infer_device(mgbvh_ctx* c, mgbvh_model* m, const float* d_features, uint32_t n,
                             float* d_logits) {
  if (!c || !m || (n && (!d_features || !d_logits))) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  int rc = model_infer(c, m, d_features, MLP_IN, (int)n, nullptr, d_logits);
  if (rc) return rc;
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MGBVH_OK : hip_fail(e, "k_mlp_infer");
}

int mgbvh_collect_samples_device(mgbvh_ctx* c, const float* d_pos, uint32_t vstride_bytes,
                                 uint64_t nverts, const uint32_t* d_tris, uint64_t ntris,
                                 const mgbvh_settings* s, uint32_t min_objects, float* d_features,
                                 uint32_t* d_labels, uint32_t max_samples, uint32_t* out_count) {
  if (!c || !s || !d_features || !d_labels || !out_count) {
    set_error("null argument");
    return MGBVH_ERR_INVALID_ARG;
  }
  if (ntris == 0) { *out_count = 0; return MGBVH_OK; }
  if (!d_pos || !d_tris || vstride_bytes < 12 || vstride_bytes % 4 || ntris > 0x7fffffffull) {
    set_error("bad geometry arguments");
    return MGBVH_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const uint64_t nslots = 2 * ntris - 1;
  int rc;
  if ((rc = ensure(c->io_nodes, nslots * sizeof(mgbvh_node32))) ||
      (rc = ensure(c->io_perm, ntris * 4)))
    return rc;
  mgbvh_settings plain = *s;
  plain.method = MGBVH_METHOD_PLAIN;
  BuildInput in{0, d_pos, vstride_bytes / 4, nverts, d_tris, nullptr};
  BuildOutput o{(mgbvh_node32*)c->io_nodes.p, nullptr, (uint32_t*)c->io_perm.p, nullptr, nullptr};
  o.samples = d_features;
  o.labels = d_labels;
  o.max_samples = max_samples;
  o.sample_min = min_objects;
  BuildResult r;
  if ((rc = build_device(c, in, (uint32_t)ntris, &plain, o, &r))) return rc;
  *out_count = std::min(r.n_samples, max_samples);
  return MGBVH_OK;
}

}  // extern "C"
