// MLP surrogate for delayed-acceptance MH: inputs -> H -> H -> 1 with GELU
// (tanh form) after both hidden layers.
//
// Parameters live as FP32 master copies (updated by Adam); the forward and
// backward passes read BF16 copies of the weight matrices (round to nearest
// even, refreshed after every update) and accumulate in FP32.  Biases stay
// FP32.  The loss is the mean squared error of the standardised output.
//
// The layers are small (404 x 256, 256 x 256 at a minibatch of 256): the
// products are 32 x 32-tiled FP32 loops with both operands staged in LDS,
// far below the sizes where matrix cores would pay for themselves against
// launch latency.  A training minibatch is a dozen launches, enqueued
// without a host synchronisation until the end of the epoch.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/b3mcmc.h"

namespace b3h {
b3h_status host_fail(b3h_status s, const std::string &msg);
}

namespace {


__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

__device__ inline float gelu(float z) {
  const float t = tanhf(kGeluC * (z + kGeluA * z * z * z));
  return 0.5f * z * (1.0f + t);
}
__device__ inline float gelu_grad(float z) {
  const float t = tanhf(kGeluC * (z + kGeluA * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * z * z);
}

// X[b][k] = (Xall[idx[b]][k] - mean[k]) / std[k]  (idx null: identity)
__global__ void k_gather(const float *xall, const uint32_t *idx, const float *mean,
                         const float *stdv, float *x, uint32_t B, uint32_t D) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)B * D) return;
  const uint32_t b = (uint32_t)(i / D), k = (uint32_t)(i % D);
  const size_t src = (size_t)(idx ? idx[b] : b) * D + k;
  x[i] = (xall[src] - mean[k]) / stdv[k];
}

__global__ void k_gather_y(const float *yall, const uint32_t *idx, float ymean, float ystd,
                           float *t, uint32_t B) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  t[b] = (yall[idx ? idx[b] : b] - ymean) / ystd;
}

// y[b] = sum_k H[b][k] w[k] + bias: one wavefront per row
__global__ __launch_bounds__(64) void k_out_fwd(const float *Hm, const uint16_t *w,
                                                const float *bias, float *y, uint32_t K) {
  const uint32_t b = blockIdx.x, l = threadIdx.x;
  float s = 0.0f;
  for (uint32_t k = l; k < K; k += 64u) s = fmaf(Hm[(size_t)b * K + k], from_bf16(w[k]), s);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (l == 0) y[b] = s + bias[0];
}

// dY = 2 (y - t) / B; loss_acc += mean (y - t)^2; g_b3 = sum_b dY[b]
// (one workgroup)
__global__ __launch_bounds__(256) void k_loss(const float *y, const float *t, float *dy,
                                              float *loss_acc, float *g_b3, uint32_t B) {
  __shared__ float part[256], gpart[256];
  float s = 0.0f, gs = 0.0f;
  for (uint32_t b = threadIdx.x; b < B; b += 256u) {
    const float e = y[b] - t[b];
    const float d = 2.0f * e / (float)B;
    dy[b] = d;
    s += e * e;
    gs += d;
  }
  part[threadIdx.x] = s;
  gpart[threadIdx.x] = gs;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      part[threadIdx.x] += part[threadIdx.x + o];
      gpart[threadIdx.x] += gpart[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    loss_acc[0] += part[0] / (float)B;
    g_b3[0] = gpart[0];
  }
}

// Products of the layers: C[M][N] = op(A)[M][K] op(B)[K][N] with FP32
// accumulation, A FP32 and B FP32 or BF16 (weights), either transposed in
// memory:
//   op(A)[i][k] = TA ? A[k lda + i] : A[i lda + k]
//   op(B)[k][j] = TB ? B[j ldb + k] : B[k ldb + j]
// Workgroup: a 32 x 32 tile of C and four wavefronts, each summing a quarter
// of every 64-deep chunk of the reduction (16 steps) into 4 x 4 outputs per
// lane; the four partial tiles are added in wavefront order at the end.  The
// chunks pass through LDS, each operand loaded with consecutive threads on
// consecutive addresses whichever way it is stored, and the next chunk's
// loads are in flight while the current one is summed.  At 256 x 404 x 256
// that is 64 workgroups of four busy SIMDs each; a split by output alone
// leaves one wavefront per SIMD waiting on every LDS read.  Epilogue:
//   kEpiBias      C = acc + bias[j], C2 = gelu(C)    (forward layer)
//   kEpiStore     C = acc                            (weight gradient)
//   kEpiGeluGrad  C = acc gelu'(Zp[i][j])            (input gradient)
constexpr int kEpiBias = 0, kEpiStore = 1, kEpiGeluGrad = 2;
constexpr uint32_t kTile = 32, kKT = 64, kLd = kTile + 4;
constexpr uint32_t kPer = kTile * kKT / 256;  // operand elements per thread and chunk
// B entries are held as loaded (FP32 bits, or a zero-extended BF16) and
// widened when stored to LDS: a conversion next to its conditional load
// would make every load wait for the one before it
__device__ __forceinline__ uint32_t ld_raw(const float *p, size_t i) { return __float_as_uint(p[i]); }
__device__ __forceinline__ uint32_t ld_raw(const uint16_t *p, size_t i) { return p[i]; }
__device__ __forceinline__ float widen(const float *, uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ float widen(const uint16_t *, uint32_t u) {
  return __uint_as_float(u << 16);
}
template <bool TA, bool TB, typename TBt, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const float *A, uint32_t lda, const TBt *Bm,
                                              uint32_t ldb, uint32_t M, uint32_t N, uint32_t K,
                                              float *C, float *C2, const float *bias,
                                              const float *Zp) {
  // [k][i] and [k][j] chunks (rows padded to 36 floats: 16-byte aligned
  // reads of four consecutive entries); afterwards the four partial tiles
  __shared__ __align__(16) float sm[2 * kKT * kLd];
  float *As = sm, *Bs = sm + kKT * kLd;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t tx = lane % 8u, ty = lane / 8u;
  const uint32_t i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
  // chunk element e of each operand: (i, k) of A and (k, j) of B, the
  // contiguous index of the stored layout varying fastest
  auto a_ik = [](uint32_t e, uint32_t &i, uint32_t &k) {
    i = TA ? e % kTile : e / kKT;
    k = TA ? e / kTile : e % kKT;
  };
  auto b_kj = [](uint32_t e, uint32_t &k, uint32_t &j) {
    k = TB ? e % kKT : e / kTile;
    j = TB ? e / kKT : e % kTile;
  };
  float ra[kPer];
  uint32_t rb[kPer];
  auto fetch = [&](uint32_t k0) {
#pragma unroll
    for (uint32_t r = 0; r < kPer; r++) {
      const uint32_t e = tid + 256u * r;
      uint32_t i, k, j, kb;
      a_ik(e, i, k);
      b_kj(e, kb, j);
      ra[r] = (i0 + i < M && k0 + k < K)
                  ? (TA ? A[(size_t)(k0 + k) * lda + i0 + i] : A[(size_t)(i0 + i) * lda + k0 + k])
                  : 0.0f;
      rb[r] = (k0 + kb < K && j0 + j < N)
                  ? (TB ? ld_raw(Bm, (size_t)(j0 + j) * ldb + k0 + kb)
                        : ld_raw(Bm, (size_t)(k0 + kb) * ldb + j0 + j))
                  : 0u;
    }
  };
  float acc[4][4] = {};
  fetch(0);
  for (uint32_t k0 = 0; k0 < K; k0 += kKT) {
#pragma unroll
    for (uint32_t r = 0; r < kPer; r++) {
      const uint32_t e = tid + 256u * r;
      uint32_t i, k, j, kb;
      a_ik(e, i, k);
      b_kj(e, kb, j);
      As[k * kLd + i] = ra[r];
      Bs[kb * kLd + j] = widen(Bm, rb[r]);
    }
    __syncthreads();
    if (k0 + kKT < K) fetch(k0 + kKT);
#pragma unroll 4
    for (uint32_t kk = 16u * w; kk < 16u * w + 16u; kk++) {
      const float4 a = *reinterpret_cast<const float4 *>(As + kk * kLd + 4u * ty);
      const float4 b = *reinterpret_cast<const float4 *>(Bs + kk * kLd + 4u * tx);
      const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (uint32_t r = 0; r < 4; r++)
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) acc[r][q] = fmaf(av[r], bv[q], acc[r][q]);
    }
    __syncthreads();
  }
  // partial tiles [w][32][32], then each thread finishes four outputs
  float *red = sm;
#pragma unroll
  for (uint32_t r = 0; r < 4; r++)
#pragma unroll
    for (uint32_t q = 0; q < 4; q++)
      red[(w * kTile + 4u * ty + r) * kTile + 4u * tx + q] = acc[r][q];
  __syncthreads();
#pragma unroll
  for (uint32_t u = 0; u < 4; u++) {
    const uint32_t o_t = tid + 256u * u, ti = o_t / kTile, tj = o_t % kTile;
    const uint32_t i = i0 + ti, j = j0 + tj;
    if (i >= M || j >= N) continue;
    float v = red[ti * kTile + tj];
#pragma unroll
    for (uint32_t q = 1; q < 4; q++) v += red[(q * kTile + ti) * kTile + tj];
    const size_t o = (size_t)i * N + j;
    if (EPI == kEpiBias) {
      const float z = v + bias[j];
      C[o] = z;
      C2[o] = gelu(z);
    } else if (EPI == kEpiStore) {
      C[o] = v;
    } else {
      C[o] = v * gelu_grad(Zp[o]);
    }
  }
}

// g[n] = sum_b s[b] D[b][n] (s null: 1), in a fixed order: 16 wavefronts
// take every 16th row, then one lane per column adds the 16 sums in order.
__global__ __launch_bounds__(1024) void k_colsum(const float *Dm, const float *scale, float *g,
                                                 uint32_t B, uint32_t N) {
  __shared__ float part[16][64];
  const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6, n = blockIdx.x * 64u + l;
  float s = 0.0f;
  if (n < N)
    for (uint32_t b = w; b < B; b += 16u) s = fmaf(scale ? scale[b] : 1.0f, Dm[(size_t)b * N + n], s);
  part[w][l] = s;
  __syncthreads();
  if (w == 0 && n < N) {
    float t = 0.0f;
    for (uint32_t q = 0; q < 16u; q++) t += part[q][l];
    g[n] = t;
  }
}

// dZ2[b][k] = dy[b] w3[k] gelu'(Z2[b][k])
__global__ __launch_bounds__(256) void k_out_bwd(const float *dy, const float *Z2,
                                                 const uint16_t *w3, float *dZ2, uint32_t B,
                                                 uint32_t K) {
  const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
  if (i < (size_t)B * K) {
    const uint32_t b = (uint32_t)(i / K), k = (uint32_t)(i % K);
    dZ2[i] = dy[b] * from_bf16(w3[k]) * gelu_grad(Z2[i]);
  }
}

// Adam (bias-corrected) on the FP32 masters, then the BF16 copy
__global__ void k_adam(float *p, const float *g, float *m, float *v, uint16_t *pb, size_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float q = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = q;
  pb[i] = to_bf16(q);
}

// One sample: the whole forward pass in one launch, with the input passed by
// value (no copy to the device) and the result written to pinned host memory
// (no copy back): one launch per surrogate call in the MCMC loop.  The
// weights are read from transposed copies (one zero-padded row of BF16 per
// hidden unit, k_transpose), eight weights per 16-byte load.  The pass is
// spread over ceil(H / 16) workgroups: workgroup g computes first-layer
// units [16 g, 16 g + 16) (16 lanes per unit, each summing every 16th group
// of eight inputs, then a butterfly over the 16 lanes) and those units'
// contribution to every second-layer unit; the workgroup that takes the
// last ticket adds the contributions in workgroup order and finishes the
// output layer.  (In one workgroup the pass was bound by one CU's load
// bandwidth for the 0.3 MB of weights: 12 us per call.)
constexpr uint32_t kMaxFusedD = 512, kMaxFusedH = 1024;
constexpr uint32_t kUnits = 16;  // first-layer units per workgroup
struct FusedInput {
  float x[kMaxFusedD];
};
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }
// eight BF16 weights dotted with v in index order, starting from z
__device__ __forceinline__ float dot8(const uint4 w, const float *v, float z) {
  z = fmaf(v[0], bf16_lo(w.x), z);
  z = fmaf(v[1], bf16_hi(w.x), z);
  z = fmaf(v[2], bf16_lo(w.y), z);
  z = fmaf(v[3], bf16_hi(w.y), z);
  z = fmaf(v[4], bf16_lo(w.z), z);
  z = fmaf(v[5], bf16_hi(w.z), z);
  z = fmaf(v[6], bf16_lo(w.w), z);
  return fmaf(v[7], bf16_hi(w.w), z);
}
__device__ __forceinline__ void wait_stores() {
  __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
}

// grid ceil(H / kUnits) x 256 threads; `in` holds the standardised input
// (the host applies the input scaling, so no load of it is on the critical
// path).  part [grid][H] holds the second-layer contributions; ticket is zero
// between launches (the last workgroup resets it).  Every row access stays
// inside its zero-padded row: Dp and Hp are multiples of 8 and each
// workgroup's unit range starts at a multiple of 16 below H.  The
// second-layer weights a thread needs first are loaded at the start, so that
// their latency overlaps the first layer's.
__global__ __launch_bounds__(256) void k_predict1(FusedInput in, uint32_t D, uint32_t Dp,
                                                  uint32_t H, uint32_t Hp, const uint16_t *W1t,
                                                  const float *b1, const uint16_t *W2t,
                                                  const float *b2, const uint16_t *w3,
                                                  const float *b3, float ymean, float ystd,
                                                  float *part, uint32_t *ticket, float *out,
                                                  uint32_t *flag, uint32_t seq) {
  __shared__ __align__(16) float xs[kMaxFusedD];
  __shared__ __align__(16) float h1[kUnits];
  __shared__ float red[4];
  __shared__ int last;
  const uint32_t tid = threadIdx.x, g = blockIdx.x, G = gridDim.x;
  const uint32_t m0 = g * kUnits;
  const bool two = m0 + 8u < Hp;
  uint4 wa = make_uint4(0u, 0u, 0u, 0u), wb = wa;
  if (tid < H) {
    const uint4 *row = reinterpret_cast<const uint4 *>(W2t + (size_t)tid * Hp + m0);
    wa = row[0];
    if (two) wb = row[1];
  }
  for (uint32_t k = tid; k < Dp; k += 256u) xs[k] = k < D ? in.x[k] : 0.0f;
  __syncthreads();
  // first layer: unit u of this workgroup, lane j of its 16
  const uint32_t u = tid >> 4, j = tid & 15u, n1 = m0 + u;
  float z = 0.0f;
  if (n1 < H) {
    const uint4 *row = reinterpret_cast<const uint4 *>(W1t + (size_t)n1 * Dp);
    for (uint32_t q = j; q < Dp / 8u; q += 16u) z = dot8(row[q], xs + 8u * q, z);
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) z += __shfl_xor(z, o, 16);
  if (j == 0) h1[u] = n1 < H ? gelu(z + b1[n1]) : 0.0f;
  __syncthreads();
  // these units' contribution to every second-layer unit
  for (uint32_t n = tid; n < H; n += 256u) {
    if (n != tid) {
      const uint4 *row = reinterpret_cast<const uint4 *>(W2t + (size_t)n * Hp + m0);
      wa = row[0];
      if (two) wb = row[1];
    }
    float s = dot8(wa, h1, 0.0f);
    if (two) s = dot8(wb, h1 + 8, s);
    __hip_atomic_store(&part[(size_t)g * H + n], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // the ticket only after every contribution is stored (order, no flush:
  // the agent-scope stores write to the level every XCD reads)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  wait_stores();
  __syncthreads();
  if (tid == 0)
    last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
           G - 1u;
  __syncthreads();
  if (!last) return;
  wait_stores();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  float s = 0.0f;
  for (uint32_t n = tid; n < H; n += 256u) {
    float z2 = b2[n];
    for (uint32_t q = 0; q < G; q++)
      z2 += __hip_atomic_load(&part[(size_t)q * H + n], __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
    s = fmaf(gelu(z2), from_bf16(w3[n]), s);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((tid & 63u) == 0) red[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    out[0] = (red[0] + red[1] + red[2] + red[3] + b3[0]) * ystd + ymean;
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // fine-grained pinned memory: a host that sees seq also sees the value
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ void k_cast(const float *p, uint16_t *pb, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pb[i] = to_bf16(p[i]);
}

// W [K][N] (BF16) -> Wt [N][Kp], zero-padded from K to Kp
__global__ void k_transpose(const uint16_t *W, uint16_t *Wt, uint32_t K, uint32_t N, uint32_t Kp) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * Kp) return;
  const uint32_t n = (uint32_t)(i / Kp), k = (uint32_t)(i % Kp);
  Wt[i] = k < K ? W[(size_t)k * N + n] : (uint16_t)0;
}

}  // namespace

struct b3m_mlp {
  int device = 0;
  uint32_t D = 0, H = 0;
  size_t np = 0, oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, ow3 = 0, ob3 = 0;
  hipStream_t stream = nullptr;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *pb = nullptr;  // BF16 copy (all parameters; the biases' entries unused)
  // transposed BF16 W1 [H][Dp] and W2 [H][Hp] for the single-sample pass,
  // rebuilt from pb before the first prediction after the parameters change
  uint16_t *w1t = nullptr, *w2t = nullptr;
  uint32_t Dp = 0, Hp = 0;
  bool t_stale = true;
  float *xmean = nullptr, *xstd = nullptr;
  std::vector<float> hmean, hstd;  // host copies: the one-sample pass scales on the host
  float ymean = 0.0f, ystd = 1.0f;
  uint32_t bcap = 0;
  float *X = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr,
        *T = nullptr, *dY = nullptr, *dZ1 = nullptr, *dZ2 = nullptr, *loss = nullptr;
  float *xin = nullptr;
  float *part = nullptr;       // [ceil(H / 16)][H] second-layer contributions
  uint32_t *ticket = nullptr;  // last-workgroup ticket of the one-sample pass
  float *h_out = nullptr;  // pinned result of the fused single-sample pass
  uint32_t *h_flag = nullptr;  // sequence number stored after the result
  uint32_t seq = 0;
  uint64_t adam_t = 0, seed = 0;
  std::vector<void *> owned;
};

namespace {

#define MCHK(expr)                                                                \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return b3h::host_fail(B3H_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

b3h_status malloc_dev(b3m_mlp *m, void **ptr, size_t bytes) {
  if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return b3h::host_fail(B3H_ERR_OOM, "hipMalloc failed for the surrogate");
  }
  m->owned.push_back(*ptr);
  return B3H_OK;
}

// activations for batches up to `cap` rows
b3h_status ensure_batch(b3m_mlp *m, uint32_t cap) {
  if (cap <= m->bcap) return B3H_OK;
  float **bufs[] = {&m->X, &m->Z1, &m->H1, &m->Z2, &m->H2, &m->dZ1, &m->dZ2};
  const size_t sizes[] = {(size_t)cap * m->D, (size_t)cap * m->H, (size_t)cap * m->H,
                          (size_t)cap * m->H, (size_t)cap * m->H, (size_t)cap * m->H,
                          (size_t)cap * m->H};
  for (int i = 0; i < 7; i++) {
    b3h_status st = malloc_dev(m, (void **)bufs[i], sizes[i] * 4);
    if (st) return st;
  }
  float **small[] = {&m->Y, &m->T, &m->dY};
  for (int i = 0; i < 3; i++) {
    b3h_status st = malloc_dev(m, (void **)small[i], (size_t)cap * 4);
    if (st) return st;
  }
  b3h_status st = malloc_dev(m, (void **)&m->xin, (size_t)cap * m->D * 4);
  if (st) return st;
  m->bcap = cap;
  return B3H_OK;
}

// forward over X (already standardised, B rows); y in m->Y
hipError_t forward(b3m_mlp *m, uint32_t B) {
  const uint32_t D = m->D, H = m->H;
  const uint16_t *W1 = m->pb + m->oW1, *W2 = m->pb + m->oW2, *w3 = m->pb + m->ow3;
  const dim3 g((H + kTile - 1) / kTile, (B + kTile - 1) / kTile);
  hipLaunchKernelGGL((k_gemm<false, false, uint16_t, kEpiBias>), g, dim3(256), 0, m->stream,
                     m->X, D, W1, H, B, H, D, m->Z1, m->H1, m->p + m->ob1, nullptr);
  hipLaunchKernelGGL((k_gemm<false, false, uint16_t, kEpiBias>), g, dim3(256), 0, m->stream,
                     m->H1, H, W2, H, B, H, H, m->Z2, m->H2, m->p + m->ob2, nullptr);
  hipLaunchKernelGGL(k_out_fwd, dim3(B), dim3(64), 0, m->stream, m->H2, w3, m->p + m->ob3, m->Y,
                     H);
  return hipGetLastError();
}

// loss and gradients (into m->g) for targets m->T
hipError_t backward(b3m_mlp *m, uint32_t B) {
  const uint32_t D = m->D, H = m->H;
  const uint16_t *W2 = m->pb + m->oW2;
  const dim3 cs((H + 63) / 64);
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, m->stream, m->Y, m->T, m->dY, m->loss,
                     m->g + m->ob3, B);
  // output layer: dZ2, then g_w3[k] = sum_b dy[b] H2[b][k]
  hipLaunchKernelGGL(k_out_bwd, dim3((unsigned)(((size_t)B * H + 255) / 256)), dim3(256), 0,
                     m->stream, m->dY, m->Z2, m->pb + m->ow3, m->dZ2, B, H);
  hipLaunchKernelGGL(k_colsum, cs, dim3(1024), 0, m->stream, m->H2, m->dY, m->g + m->ow3, B, H);
  // second layer: g_W2 = H1^T dZ2, g_b2, dZ1 = (dZ2 W2^T) gelu'(Z1)
  hipLaunchKernelGGL((k_gemm<true, false, float, kEpiStore>),
                     dim3((H + kTile - 1) / kTile, (H + kTile - 1) / kTile), dim3(256), 0,
                     m->stream, m->H1, H, m->dZ2, H, H, H, B, m->g + m->oW2, nullptr, nullptr,
                     nullptr);
  hipLaunchKernelGGL(k_colsum, cs, dim3(1024), 0, m->stream, m->dZ2, nullptr, m->g + m->ob2, B,
                     H);
  hipLaunchKernelGGL((k_gemm<false, true, uint16_t, kEpiGeluGrad>),
                     dim3((H + kTile - 1) / kTile, (B + kTile - 1) / kTile), dim3(256), 0,
                     m->stream, m->dZ2, H, W2, H, B, H, H, m->dZ1, nullptr, nullptr, m->Z1);
  // first layer: g_W1 = X^T dZ1, g_b1
  hipLaunchKernelGGL((k_gemm<true, false, float, kEpiStore>),
                     dim3((H + kTile - 1) / kTile, (D + kTile - 1) / kTile), dim3(256), 0,
                     m->stream, m->X, D, m->dZ1, H, D, H, B, m->g + m->oW1, nullptr, nullptr,
                     nullptr);
  hipLaunchKernelGGL(k_colsum, cs, dim3(1024), 0, m->stream, m->dZ1, nullptr, m->g + m->ob1, B,
                     H);
  return hipGetLastError();
}

}  // namespace

// forward pass for the chain's use: x (host, B rows) -> y (host)
namespace b3h {
b3h_status mlp_predict(b3m_mlp *m, const float *x, uint32_t B, float *y) {
  if (B == 1 && m->D <= kMaxFusedD && m->H <= kMaxFusedH && m->h_out) {
    if (m->t_stale) {
      const size_t n1 = (size_t)m->H * m->Dp, n2 = (size_t)m->H * m->Hp;
      hipLaunchKernelGGL(k_transpose, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, m->stream,
                         m->pb + m->oW1, m->w1t, m->D, m->H, m->Dp);
      hipLaunchKernelGGL(k_transpose, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, m->stream,
                         m->pb + m->oW2, m->w2t, m->H, m->H, m->Hp);
      MCHK(hipGetLastError());
      m->t_stale = false;
    }
    // (x - mean) / std in FP32, as k_gather computes it on the device
    FusedInput in;
    for (uint32_t k = 0; k < m->D; k++) in.x[k] = (x[k] - m->hmean[k]) / m->hstd[k];
    hipLaunchKernelGGL(k_predict1, dim3((m->H + kUnits - 1) / kUnits), dim3(256), 0, m->stream,
                       in, m->D, m->Dp, m->H, m->Hp, m->w1t, m->p + m->ob1,
                       m->w2t, m->p + m->ob2, m->pb + m->ow3, m->p + m->ob3, m->ymean, m->ystd,
                       m->part, m->ticket, m->h_out, m->h_flag, ++m->seq);
    MCHK(hipGetLastError());
    // poll for the launch's sequence number (up to 20 ms), else wait for
    // the stream, which also reports any error of the launch
    const uint32_t want = m->seq;
    bool seen = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 1; !seen; i++) {
      seen = __atomic_load_n(m->h_flag, __ATOMIC_ACQUIRE) == want;
      if (!seen) __builtin_ia32_pause();
      if (!seen && (i & 1023u) == 0 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20))
        break;
    }
    if (!seen) {
      MCHK(hipStreamSynchronize(m->stream));
      if (__atomic_load_n(m->h_flag, __ATOMIC_ACQUIRE) != want)
        return b3h::host_fail(B3H_ERR_HIP, "surrogate forward did not complete");
    }
    y[0] = *(volatile float *)m->h_out;
    return B3H_OK;
  }
  b3h_status st = ensure_batch(m, B);
  if (st) return st;
  MCHK(hipMemcpyAsync(m->xin, x, (size_t)B * m->D * 4, hipMemcpyHostToDevice, m->stream));
  const size_t tot = (size_t)B * m->D;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, m->stream,
                     m->xin, nullptr, m->xmean, m->xstd, m->X, B, m->D);
  MCHK(forward(m, B));
  std::vector<float> out(B);
  MCHK(hipMemcpyAsync(out.data(), m->Y, (size_t)B * 4, hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  for (uint32_t b = 0; b < B; b++) y[b] = out[b] * m->ystd + m->ymean;
  return B3H_OK;
}
}  // namespace b3h

extern "C" {

b3h_status b3m_mlp_create(int device, uint32_t inputs, uint32_t hidden, uint64_t seed,
                          b3m_mlp **out) {
  if (!out || inputs == 0 || hidden == 0 || inputs > 4096 || hidden > 1024)
    return b3h::host_fail(B3H_ERR_ARG, "bad surrogate dimensions");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return b3h::host_fail(B3H_ERR_NO_DEVICE, "no HIP device");
  }
  if (device < 0 || device >= count) return b3h::host_fail(B3H_ERR_ARG, "bad device");
  if (hipSetDevice(device) != hipSuccess) return b3h::host_fail(B3H_ERR_HIP, "hipSetDevice failed");
  b3m_mlp *m = new b3m_mlp();
  m->device = device;
  m->D = inputs;
  m->H = hidden;
  m->Dp = (inputs + 7u) & ~7u;
  m->Hp = (hidden + 7u) & ~7u;
  m->seed = seed;
  m->oW1 = 0;
  m->ob1 = (size_t)inputs * hidden;
  m->oW2 = m->ob1 + hidden;
  m->ob2 = m->oW2 + (size_t)hidden * hidden;
  m->ow3 = m->ob2 + hidden;
  m->ob3 = m->ow3 + hidden;
  m->np = m->ob3 + 1;
  auto fail = [&](b3h_status st) {
    b3m_mlp_destroy(m);
    return st;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(b3h::host_fail(B3H_ERR_HIP, "hipStreamCreate failed"));
  b3h_status st;
  if ((st = malloc_dev(m, (void **)&m->p, m->np * 4)) ||
      (st = malloc_dev(m, (void **)&m->g, m->np * 4)) ||
      (st = malloc_dev(m, (void **)&m->m, m->np * 4)) ||
      (st = malloc_dev(m, (void **)&m->v, m->np * 4)) ||
      (st = malloc_dev(m, (void **)&m->pb, m->np * 2)) ||
      (st = malloc_dev(m, (void **)&m->w1t, (size_t)hidden * ((inputs + 7u) & ~7u) * 2)) ||
      (st = malloc_dev(m, (void **)&m->w2t, (size_t)hidden * ((hidden + 7u) & ~7u) * 2)) ||
      (st = malloc_dev(m, (void **)&m->xmean, (size_t)inputs * 4)) ||
      (st = malloc_dev(m, (void **)&m->xstd, (size_t)inputs * 4)) ||
      (st = malloc_dev(m, (void **)&m->loss, 16)) ||
      (st = malloc_dev(m, (void **)&m->part,
                       (size_t)((hidden + kUnits - 1) / kUnits) * hidden * 4)) ||
      (st = malloc_dev(m, (void **)&m->ticket, 64)))
    return fail(st);
  // He-uniform initial weights (GELU), zero biases, identity scaling
  std::vector<float> p(m->np, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](size_t off, size_t count, uint32_t fan_in) {
    std::uniform_real_distribution<float> u(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
    for (size_t i = 0; i < count; i++) p[off + i] = u(rng);
  };
  fill(m->oW1, (size_t)inputs * hidden, inputs);
  fill(m->oW2, (size_t)hidden * hidden, hidden);
  fill(m->ow3, hidden, hidden);
  std::vector<float> zeros(inputs, 0.0f), ones(inputs, 1.0f);
  m->hmean = zeros;
  m->hstd = ones;
  bool ok = hipMemcpy(m->p, p.data(), m->np * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(m->m, 0, m->np * 4) == hipSuccess &&
            hipMemset(m->v, 0, m->np * 4) == hipSuccess &&
            hipMemset(m->ticket, 0, 64) == hipSuccess &&
            hipMemcpy(m->xmean, zeros.data(), inputs * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(m->xstd, ones.data(), inputs * 4, hipMemcpyHostToDevice) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_cast, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->stream,
                       m->p, m->pb, m->np);
    ok = hipStreamSynchronize(m->stream) == hipSuccess;
  }
  if (ok) ok = hipHostMalloc((void **)&m->h_out, 64, hipHostMallocCoherent) == hipSuccess &&
               hipHostMalloc((void **)&m->h_flag, 64, hipHostMallocCoherent) == hipSuccess;
  if (ok) {
    std::memset(m->h_out, 0, 64);
    std::memset(m->h_flag, 0, 64);
  }
  if (!ok) return fail(b3h::host_fail(B3H_ERR_HIP, "surrogate initialisation failed"));
  *out = m;
  return B3H_OK;
}

b3h_status b3m_mlp_destroy(b3m_mlp *m) {
  if (!m) return B3H_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void *p : m->owned) (void)hipFree(p);
  if (m->h_out) (void)hipHostFree(m->h_out);
  if (m->h_flag) (void)hipHostFree(m->h_flag);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return B3H_OK;
}

b3h_status b3m_mlp_set_scaling(b3m_mlp *m, const float *x_mean, const float *x_std, float y_mean,
                               float y_std) {
  if (!m || !x_mean || !x_std || !(y_std > 0.0f))
    return b3h::host_fail(B3H_ERR_ARG, "bad surrogate scaling");
  for (uint32_t k = 0; k < m->D; k++)
    if (!(x_std[k] > 0.0f)) return b3h::host_fail(B3H_ERR_ARG, "input scales must be positive");
  (void)hipSetDevice(m->device);
  MCHK(hipMemcpy(m->xmean, x_mean, (size_t)m->D * 4, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(m->xstd, x_std, (size_t)m->D * 4, hipMemcpyHostToDevice));
  m->hmean.assign(x_mean, x_mean + m->D);
  m->hstd.assign(x_std, x_std + m->D);
  m->ymean = y_mean;
  m->ystd = y_std;
  return B3H_OK;
}

b3h_status b3m_mlp_forward(b3m_mlp *m, const float *x, uint32_t batch, float *y) {
  if (!m || !x || !y || batch == 0) return b3h::host_fail(B3H_ERR_ARG, "bad surrogate forward");
  (void)hipSetDevice(m->device);
  return b3h::mlp_predict(m, x, batch, y);
}

size_t b3m_mlp_num_params(const b3m_mlp *m) { return m ? m->np : 0; }

b3h_status b3m_mlp_get_params(const b3m_mlp *m, float *params) {
  if (!m || !params) return b3h::host_fail(B3H_ERR_ARG, "null argument");
  (void)hipSetDevice(m->device);
  MCHK(hipMemcpy(params, m->p, m->np * 4, hipMemcpyDeviceToHost));
  return B3H_OK;
}

b3h_status b3m_mlp_set_params(b3m_mlp *m, const float *params) {
  if (!m || !params) return b3h::host_fail(B3H_ERR_ARG, "null argument");
  (void)hipSetDevice(m->device);
  MCHK(hipMemcpy(m->p, params, m->np * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_cast, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->stream,
                     m->p, m->pb, m->np);
  m->t_stale = true;
  MCHK(hipStreamSynchronize(m->stream));
  return B3H_OK;
}

b3h_status b3m_mlp_loss_grad(b3m_mlp *m, const float *x, const float *y, uint32_t batch,
                             float *loss, float *grads) {
  if (!m || !x || !y || !loss || !grads || batch == 0)
    return b3h::host_fail(B3H_ERR_ARG, "bad surrogate gradient call");
  (void)hipSetDevice(m->device);
  b3h_status st = ensure_batch(m, batch);
  if (st) return st;
  MCHK(hipMemcpyAsync(m->xin, x, (size_t)batch * m->D * 4, hipMemcpyHostToDevice, m->stream));
  std::vector<float> yin(y, y + batch);
  MCHK(hipMemcpyAsync(m->dY, yin.data(), (size_t)batch * 4, hipMemcpyHostToDevice, m->stream));
  const size_t tot = (size_t)batch * m->D;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, m->stream,
                     m->xin, nullptr, m->xmean, m->xstd, m->X, batch, m->D);
  hipLaunchKernelGGL(k_gather_y, dim3((batch + 255) / 256), dim3(256), 0, m->stream, m->dY,
                     nullptr, m->ymean, m->ystd, m->T, batch);
  MCHK(hipMemsetAsync(m->loss, 0, 4, m->stream));
  MCHK(forward(m, batch));
  MCHK(backward(m, batch));
  MCHK(hipMemcpyAsync(grads, m->g, m->np * 4, hipMemcpyDeviceToHost, m->stream));
  MCHK(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return B3H_OK;
}

b3h_status b3m_mlp_train(b3m_mlp *m, const float *x, const float *y, uint32_t n, uint32_t epochs,
                         uint32_t batch, float lr, float *final_loss) {
  if (!m || !x || !y || n == 0 || batch == 0 || epochs == 0)
    return b3h::host_fail(B3H_ERR_ARG, "bad surrogate training call");
  (void)hipSetDevice(m->device);
  batch = std::min(batch, n);
  b3h_status st = ensure_batch(m, batch);
  if (st) return st;
  float *xall = nullptr, *yall = nullptr;
  uint32_t *dorder = nullptr;
  if (hipMalloc((void **)&xall, (size_t)n * m->D * 4) != hipSuccess ||
      hipMalloc((void **)&yall, (size_t)n * 4) != hipSuccess ||
      hipMalloc((void **)&dorder, (size_t)n * 4) != hipSuccess) {
    (void)hipGetLastError();
    if (xall) (void)hipFree(xall);
    if (yall) (void)hipFree(yall);
    return b3h::host_fail(B3H_ERR_OOM, "hipMalloc failed for the training set");
  }
  auto done = [&](b3h_status s) {
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(xall);
    (void)hipFree(yall);
    (void)hipFree(dorder);
    return s;
  };
  if (hipMemcpy(xall, x, (size_t)n * m->D * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(yall, y, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
    return done(b3h::host_fail(B3H_ERR_HIP, "training set upload failed"));
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::mt19937_64 rng(m->seed ^ 0x9E3779B97F4A7C15ull);
  const uint32_t nb = n / batch;  // full minibatches only
  float last = 0.0f;
  for (uint32_t ep = 0; ep < epochs; ep++) {
    // the stream is idle here (the previous epoch ended with a
    // synchronisation), so the blocking upload of this epoch's order is
    // complete before its first minibatch reads it
    std::shuffle(order.begin(), order.end(), rng);
    if (hipMemcpy(dorder, order.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
      return done(b3h::host_fail(B3H_ERR_HIP, "index upload failed"));
    if (hipMemsetAsync(m->loss, 0, 4, m->stream) != hipSuccess)
      return done(b3h::host_fail(B3H_ERR_HIP, "hipMemsetAsync failed"));
    for (uint32_t bi = 0; bi < nb; bi++) {
      const uint32_t *idx = dorder + (size_t)bi * batch;
      const size_t tot = (size_t)batch * m->D;
      hipLaunchKernelGGL(k_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, m->stream,
                         xall, idx, m->xmean, m->xstd, m->X, batch, m->D);
      hipLaunchKernelGGL(k_gather_y, dim3((batch + 255) / 256), dim3(256), 0, m->stream, yall,
                         idx, m->ymean, m->ystd, m->T, batch);
      if (forward(m, batch) != hipSuccess || backward(m, batch) != hipSuccess)
        return done(b3h::host_fail(B3H_ERR_HIP, "surrogate training launch failed"));
      m->adam_t++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.0f - std::pow(b1, (float)m->adam_t);
      const float c2 = 1.0f - std::pow(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->stream,
                         m->p, m->g, m->m, m->v, m->pb, m->np, lr, b1, b2, 1e-8f, c1, c2);
      m->t_stale = true;
      if (hipGetLastError() != hipSuccess)
        return done(b3h::host_fail(B3H_ERR_HIP, "surrogate training launch failed"));
    }
    float acc = 0.0f;
    if (hipMemcpyAsync(&acc, m->loss, 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      return done(b3h::host_fail(B3H_ERR_HIP, "surrogate training failed"));
    last = acc / (float)std::max(nb, 1u);
  }
  if (final_loss) *final_loss = last;
  return done(B3H_OK);
}

}  // extern "C"
