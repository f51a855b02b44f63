// MLP surrogate for delayed-acceptance MH: inputs -> H -> H -> 1 with GELU
// (tanh form) after both hidden layers.
//
// Parameters live as FP32 master copies (updated by Adam); the forward and
// backward passes read BF16 copies of the weight matrices (round to nearest
// even, refreshed after every update) and accumulate in FP32.  Biases stay
// FP32.  The loss is the mean squared error of the standardised output.
//
// The layers are small (404 x 256, 256 x 256), so every kernel is a plain
// FP32-accumulating loop with the reused operand staged in LDS; a training
// minibatch is a dozen launches.
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

constexpr uint32_t kRows = 16;  // batch rows per forward workgroup

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

// Z[b][n] = sum_k A[b][k] W[k][n] + bias[n], Hout = gelu(Z).  Workgroup:
// kRows batch rows x 256 columns; the rows' inputs are staged in LDS.
__global__ __launch_bounds__(256) void k_dense_fwd(const float *A, const uint16_t *W,
                                                   const float *bias, float *Z, float *Hout,
                                                   uint32_t B, uint32_t K, uint32_t N) {
  extern __shared__ float a_s[];  // [kRows][K]
  const uint32_t b0 = blockIdx.y * kRows, n = blockIdx.x * 256u + threadIdx.x;
  for (uint32_t i = threadIdx.x; i < kRows * K; i += 256u) {
    const uint32_t r = i / K, k = i % K;
    a_s[i] = b0 + r < B ? A[(size_t)(b0 + r) * K + k] : 0.0f;
  }
  __syncthreads();
  if (n >= N) return;
  float acc[kRows];
#pragma unroll
  for (uint32_t r = 0; r < kRows; r++) acc[r] = bias[n];
  for (uint32_t k = 0; k < K; k++) {
    const float w = from_bf16(W[(size_t)k * N + n]);
#pragma unroll
    for (uint32_t r = 0; r < kRows; r++) acc[r] = fmaf(a_s[r * K + k], w, acc[r]);
  }
#pragma unroll
  for (uint32_t r = 0; r < kRows; r++) {
    if (b0 + r >= B) break;
    Z[(size_t)(b0 + r) * N + n] = acc[r];
    Hout[(size_t)(b0 + r) * N + n] = gelu(acc[r]);
  }
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

// dY = 2 (y - t) / B; loss_acc += mean (y - t)^2 (one workgroup)
__global__ __launch_bounds__(256) void k_loss(const float *y, const float *t, float *dy,
                                              float *loss_acc, uint32_t B) {
  __shared__ float part[256];
  float s = 0.0f;
  for (uint32_t b = threadIdx.x; b < B; b += 256u) {
    const float e = y[b] - t[b];
    dy[b] = 2.0f * e / (float)B;
    s += e * e;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_acc[0] += part[0] / (float)B;
}

// output layer backward: g_w3[k] = sum_b dy[b] H2[b][k] (and g_b3), and
// dZ2[b][k] = dy[b] w3[k] gelu'(Z2[b][k])
__global__ __launch_bounds__(256) void k_out_bwd(const float *dy, const float *H2,
                                                 const float *Z2, const uint16_t *w3,
                                                 float *g_w3, float *g_b3, float *dZ2, uint32_t B,
                                                 uint32_t K) {
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k < K) {
    float g = 0.0f;
    const float w = from_bf16(w3[k]);
    for (uint32_t b = 0; b < B; b++) {
      g = fmaf(dy[b], H2[(size_t)b * K + k], g);
      dZ2[(size_t)b * K + k] = dy[b] * w * gelu_grad(Z2[(size_t)b * K + k]);
    }
    g_w3[k] = g;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.0f;
    for (uint32_t b = 0; b < B; b++) s += dy[b];
    g_b3[0] = s;
  }
}

// G[k][n] = sum_b A[b][k] D[b][n]; workgroup = one k, 256 columns
__global__ __launch_bounds__(256) void k_wgrad(const float *A, const float *Dm, float *G,
                                               uint32_t B, uint32_t K, uint32_t N) {
  const uint32_t k = blockIdx.y, n = blockIdx.x * 256u + threadIdx.x;
  if (n >= N) return;
  float g = 0.0f;
  for (uint32_t b = 0; b < B; b++) g = fmaf(A[(size_t)b * K + k], Dm[(size_t)b * N + n], g);
  G[(size_t)k * N + n] = g;
}

__global__ __launch_bounds__(256) void k_bgrad(const float *Dm, float *g, uint32_t B, uint32_t N) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (uint32_t b = 0; b < B; b++) s += Dm[(size_t)b * N + n];
  g[n] = s;
}

// dZp[b][k] = (sum_n D[b][n] W[k][n]) gelu'(Zp[b][k])
__global__ __launch_bounds__(256) void k_igrad(const float *Dm, const uint16_t *W,
                                               const float *Zp, float *dZp, uint32_t B, uint32_t K,
                                               uint32_t N) {
  const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= (size_t)B * K) return;
  const uint32_t b = (uint32_t)(i / K), k = (uint32_t)(i % K);
  float s = 0.0f;
  for (uint32_t n = 0; n < N; n++) s = fmaf(Dm[(size_t)b * N + n], from_bf16(W[(size_t)k * N + n]), s);
  dZp[i] = s * gelu_grad(Zp[i]);
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

// One sample, one workgroup: the whole forward pass with the input passed by
// value (no copy to the device) and the result written to pinned host
// memory (no copy back): one launch per surrogate call in the MCMC loop.
// The hidden units use the same FP32 summation order as k_dense_fwd.  The
// weights are read from transposed copies (one padded row of BF16 per
// hidden unit, k_transpose), so a lane loads eight weights per 16-byte load
// and all of a row's loads are independent: in the [k][n] layout of the
// training kernels a lane read two bytes per load, and the pass was bound
// by one CU's memory latency (about 40 us per call).
constexpr uint32_t kMaxFusedD = 512, kMaxFusedH = 1024;
struct FusedInput {
  float x[kMaxFusedD];
};
// row of weights (BF16, zero-padded from `used` to a multiple of 8) dotted
// with v in index order, starting from z: whole groups of eight weights,
// sixteen 16-byte loads in flight at a time, then the last partial group
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }
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
__device__ __forceinline__ float dot_row(const uint16_t *row, const float *v, uint32_t used,
                                         float z) {
  const uint4 *r = reinterpret_cast<const uint4 *>(row);
  const uint32_t full = used / 8u;
  uint32_t g = 0;
  constexpr uint32_t kBatch = 16;  // 16-byte loads in flight per lane
  for (; g + kBatch <= full; g += kBatch) {
    uint4 w[kBatch];
#pragma unroll
    for (uint32_t q = 0; q < kBatch; q++) w[q] = r[g + q];
#pragma unroll
    for (uint32_t q = 0; q < kBatch; q++) z = dot8(w[q], v + 8u * (g + q), z);
  }
  for (; g < full; g++) z = dot8(r[g], v + 8u * g, z);
  if (used % 8u) {
    const uint4 w = r[full];
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    for (uint32_t j = 0; j < used % 8u; j++)
      z = fmaf(v[8u * full + j], (j & 1u) ? bf16_hi(ws[j / 2]) : bf16_lo(ws[j / 2]), z);
  }
  return z;
}

__global__ __launch_bounds__(256) void k_predict1(FusedInput in, uint32_t D, uint32_t Dp,
                                                  uint32_t H, uint32_t Hp, const float *xmean,
                                                  const float *xstd, const uint16_t *W1t,
                                                  const float *b1, const uint16_t *W2t,
                                                  const float *b2, const uint16_t *w3,
                                                  const float *b3, float ymean, float ystd,
                                                  float *out, uint32_t *flag, uint32_t seq) {
  __shared__ float xs[kMaxFusedD];
  __shared__ float h1[kMaxFusedH];
  __shared__ float part[4];
  const uint32_t tid = threadIdx.x;
  for (uint32_t k = tid; k < D; k += 256u) xs[k] = (in.x[k] - xmean[k]) / xstd[k];
  __syncthreads();
  for (uint32_t n = tid; n < H; n += 256u)
    h1[n] = gelu(dot_row(W1t + (size_t)n * Dp, xs, D, b1[n]));
  __syncthreads();
  float s = 0.0f;
  for (uint32_t n = tid; n < H; n += 256u) {
    const float z = dot_row(W2t + (size_t)n * Hp, h1, H, b2[n]);
    s = fmaf(gelu(z), from_bf16(w3[n]), s);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((tid & 63u) == 0) part[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    out[0] = (part[0] + part[1] + part[2] + part[3] + b3[0]) * ystd + ymean;
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
  float ymean = 0.0f, ystd = 1.0f;
  uint32_t bcap = 0;
  float *X = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr,
        *T = nullptr, *dY = nullptr, *dZ1 = nullptr, *dZ2 = nullptr, *loss = nullptr;
  float *xin = nullptr;
  uint32_t *idx = nullptr;
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
  st = malloc_dev(m, (void **)&m->idx, (size_t)cap * 4);
  if (st) return st;
  m->bcap = cap;
  return B3H_OK;
}

// forward over X (already standardised, B rows); y in m->Y
hipError_t forward(b3m_mlp *m, uint32_t B) {
  const uint32_t D = m->D, H = m->H;
  const uint16_t *W1 = m->pb + m->oW1, *W2 = m->pb + m->oW2, *w3 = m->pb + m->ow3;
  dim3 g1((H + 255) / 256, (B + kRows - 1) / kRows);
  hipLaunchKernelGGL(k_dense_fwd, g1, dim3(256), kRows * D * 4, m->stream, m->X, W1,
                     m->p + m->ob1, m->Z1, m->H1, B, D, H);
  hipLaunchKernelGGL(k_dense_fwd, g1, dim3(256), kRows * H * 4, m->stream, m->H1, W2,
                     m->p + m->ob2, m->Z2, m->H2, B, H, H);
  hipLaunchKernelGGL(k_out_fwd, dim3(B), dim3(64), 0, m->stream, m->H2, w3, m->p + m->ob3, m->Y,
                     H);
  return hipGetLastError();
}

// loss and gradients (into m->g) for targets m->T
hipError_t backward(b3m_mlp *m, uint32_t B) {
  const uint32_t D = m->D, H = m->H;
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, m->stream, m->Y, m->T, m->dY, m->loss, B);
  hipLaunchKernelGGL(k_out_bwd, dim3((H + 255) / 256), dim3(256), 0, m->stream, m->dY, m->H2,
                     m->Z2, m->pb + m->ow3, m->g + m->ow3, m->g + m->ob3, m->dZ2, B, H);
  hipLaunchKernelGGL(k_wgrad, dim3((H + 255) / 256, H), dim3(256), 0, m->stream, m->H1, m->dZ2,
                     m->g + m->oW2, B, H, H);
  hipLaunchKernelGGL(k_bgrad, dim3((H + 255) / 256), dim3(256), 0, m->stream, m->dZ2,
                     m->g + m->ob2, B, H);
  hipLaunchKernelGGL(k_igrad, dim3((B * H + 255) / 256), dim3(256), 0, m->stream, m->dZ2,
                     m->pb + m->oW2, m->Z1, m->dZ1, B, H, H);
  hipLaunchKernelGGL(k_wgrad, dim3((H + 255) / 256, D), dim3(256), 0, m->stream, m->X, m->dZ1,
                     m->g + m->oW1, B, D, H);
  hipLaunchKernelGGL(k_bgrad, dim3((H + 255) / 256), dim3(256), 0, m->stream, m->dZ1,
                     m->g + m->ob1, B, H);
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
    FusedInput in;
    std::memcpy(in.x, x, (size_t)m->D * 4);
    hipLaunchKernelGGL(k_predict1, dim3(1), dim3(256), 0, m->stream, in, m->D, m->Dp, m->H,
                       m->Hp, m->xmean, m->xstd, m->w1t, m->p + m->ob1, m->w2t, m->p + m->ob2,
                       m->pb + m->ow3, m->p + m->ob3, m->ymean, m->ystd, m->h_out, m->h_flag,
                       ++m->seq);
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
      (st = malloc_dev(m, (void **)&m->loss, 16)))
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
  bool ok = hipMemcpy(m->p, p.data(), m->np * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(m->m, 0, m->np * 4) == hipSuccess &&
            hipMemset(m->v, 0, m->np * 4) == hipSuccess &&
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
  if (hipMalloc((void **)&xall, (size_t)n * m->D * 4) != hipSuccess ||
      hipMalloc((void **)&yall, (size_t)n * 4) != hipSuccess) {
    (void)hipGetLastError();
    if (xall) (void)hipFree(xall);
    return b3h::host_fail(B3H_ERR_OOM, "hipMalloc failed for the training set");
  }
  auto done = [&](b3h_status s) {
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(xall);
    (void)hipFree(yall);
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
    std::shuffle(order.begin(), order.end(), rng);
    if (hipMemsetAsync(m->loss, 0, 4, m->stream) != hipSuccess)
      return done(b3h::host_fail(B3H_ERR_HIP, "hipMemsetAsync failed"));
    for (uint32_t bi = 0; bi < nb; bi++) {
      if (hipMemcpyAsync(m->idx, order.data() + (size_t)bi * batch, (size_t)batch * 4,
                         hipMemcpyHostToDevice, m->stream) != hipSuccess)
        return done(b3h::host_fail(B3H_ERR_HIP, "index upload failed"));
      const size_t tot = (size_t)batch * m->D;
      hipLaunchKernelGGL(k_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, m->stream,
                         xall, m->idx, m->xmean, m->xstd, m->X, batch, m->D);
      hipLaunchKernelGGL(k_gather_y, dim3((batch + 255) / 256), dim3(256), 0, m->stream, yall,
                         m->idx, m->ymean, m->ystd, m->T, batch);
      if (forward(m, batch) != hipSuccess || backward(m, batch) != hipSuccess)
        return done(b3h::host_fail(B3H_ERR_HIP, "surrogate training launch failed"));
      m->adam_t++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.0f - std::pow(b1, (float)m->adam_t);
      const float c2 = 1.0f - std::pow(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->stream,
                         m->p, m->g, m->m, m->v, m->pb, m->np, lr, b1, b2, 1e-8f, c1, c2);
      m->t_stale = true;
      // synchronise before the next index upload reuses the host order buffer
      if (hipStreamSynchronize(m->stream) != hipSuccess)
        return done(b3h::host_fail(B3H_ERR_HIP, "surrogate training failed"));
    }
    float acc = 0.0f;
    if (hipMemcpy(&acc, m->loss, 4, hipMemcpyDeviceToHost) != hipSuccess)
      return done(b3h::host_fail(B3H_ERR_HIP, "loss readback failed"));
    last = acc / (float)std::max(nb, 1u);
  }
  if (final_loss) *final_loss = last;
  return done(B3H_OK);
}

}  // extern "C"
