// FP32 4-state pruning kernel with exact power-of-two rescaling.
//
// Computes what b3's calculators compute (crates/b3/src/calculator/cpu4.rs:
// 170-433, cuda_partials/kernels.cu:36-140) for K rate categories at once:
//   x_j = (sum_k PL[j][k] yL_k) * (sum_k PR[j][k] yR_k)   (tip side: PL[j][code],
//         or 1 for an unknown code)
// followed by the root value ln(sum_j pi_j x_j) net of scaling, the category
// mix LSE_c + ln(1/K) (likelihood/gamma.rs:154-166) and the pattern-weighted
// sum (likelihood/mod.rs:134-139).
//
// Layout and schedule (MI355X):
//  - a workgroup is K x T wavefronts; wavefront (c, t) owns category c of the
//    64 patterns of sub-tile t, so one lane = one (pattern, category) and a
//    partial is one float4 (16 B): a wavefront moves 1 KiB per node, coalesced.
//  - partials [slot][K][P] float4 with an int16 cumulative binary exponent per
//    (slot, category, pattern).  After every node the four values are scaled
//    by 2^-e (e = exponent of their maximum), which is exact, so FP32 never
//    underflows and no rounding comes from scaling.
//  - the host orders the update list depth-first with the larger register
//    need first (Strahler order); a child computed in the same pass is taken
//    from a per-wavefront LDS stack instead of HBM.  Every node is still
//    written to HBM because later partial updates read it.
//  - the transition matrices of an entry are wave-uniform (scalar loads).
#include "common.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

namespace b3h {

namespace {

struct Src {
  float4 v;
  int e;
};

__device__ __forceinline__ float col_or_one(float m0, float m1, float m2,
                                            float m3, uint32_t code) {
  float r = 1.0f;
  r = code == 0 ? m0 : r;
  r = code == 1 ? m1 : r;
  r = code == 2 ? m2 : r;
  r = code == 3 ? m3 : r;
  return r;
}

__device__ __forceinline__ Src project(uint32_t src, const float *__restrict__ M,
                                       const F32Args &a, uint32_t c,
                                       uint32_t pp, bool active,
                                       const float4 *stk, const int *stke,
                                       uint32_t lane, size_t KP) {
  float m[16];
#pragma unroll
  for (int k = 0; k < 16; k++) m[k] = M[k];
  const uint32_t tag = src & SRC_TAG, idx = src & SRC_IDX;
  Src out;
  if (tag == SRC_TIP) {
    const uint32_t code = active ? a.codes[(size_t)idx * a.P + pp] : 4u;
    out.v.x = col_or_one(m[0], m[1], m[2], m[3], code);
    out.v.y = col_or_one(m[4], m[5], m[6], m[7], code);
    out.v.z = col_or_one(m[8], m[9], m[10], m[11], code);
    out.v.w = col_or_one(m[12], m[13], m[14], m[15], code);
    out.e = 0;
    return out;
  }
  float4 y;
  if (tag == SRC_LDS) {
    y = stk[idx * 64 + lane];
    out.e = stke[idx * 64 + lane];
  } else {
    const size_t o = (size_t)idx * KP + (size_t)c * a.P + pp;
    y = a.partials[o];
    out.e = a.exps[o];
  }
  out.v.x = m[0] * y.x + m[1] * y.y + m[2] * y.z + m[3] * y.w;
  out.v.y = m[4] * y.x + m[5] * y.y + m[6] * y.z + m[7] * y.w;
  out.v.z = m[8] * y.x + m[9] * y.y + m[10] * y.z + m[11] * y.w;
  out.v.w = m[12] * y.x + m[13] * y.y + m[14] * y.z + m[15] * y.w;
  return out;
}

}  // namespace

__global__ __launch_bounds__(1024) void prune_f32_kernel(F32Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t W = a.K * a.T;
  const uint32_t c = w % a.K, t = w / a.K;
  const uint32_t p = (blockIdx.x * a.T + t) * 64u + lane;
  const bool active = p < a.P;
  const uint32_t pp = active ? p : 0u;
  const uint32_t ns = a.nslots;

  float4 *stk_all = reinterpret_cast<float4 *>(smem);
  int *stke_all = reinterpret_cast<int *>(stk_all + (size_t)W * ns * 64);
  double *mix = reinterpret_cast<double *>(stke_all + (size_t)W * ns * 64);
  double *red = mix + (size_t)W * 64;
  int *flag = reinterpret_cast<int *>(red + a.T);
  float4 *stk = stk_all + (size_t)w * ns * 64;
  int *stke = stke_all + (size_t)w * ns * 64;

  const size_t KP = (size_t)a.K * a.P;
  float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
  int E = 0;
  for (uint32_t i = 0; i < a.n; ++i) {
    const Entry e = a.entries[i];
    const float *ML = a.pmats + ((size_t)(2 * i) * a.K + c) * 16;
    const float *MR = a.pmats + ((size_t)(2 * i + 1) * a.K + c) * 16;
    const Src l = project(e.left, ML, a, c, pp, active, stk, stke, lane, KP);
    const Src r = project(e.right, MR, a, c, pp, active, stk, stke, lane, KP);
    x.x = l.v.x * r.v.x;
    x.y = l.v.y * r.v.y;
    x.z = l.v.z * r.v.z;
    x.w = l.v.w * r.v.w;
    const float mx = fmaxf(fmaxf(x.x, x.y), fmaxf(x.z, x.w));
    const int ex = (mx > 0.0f) ? __builtin_amdgcn_frexp_expf(mx) : 0;
    x.x = __builtin_amdgcn_ldexpf(x.x, -ex);
    x.y = __builtin_amdgcn_ldexpf(x.y, -ex);
    x.z = __builtin_amdgcn_ldexpf(x.z, -ex);
    x.w = __builtin_amdgcn_ldexpf(x.w, -ex);
    E = l.e + r.e + ex;
    if (active) {
      const size_t o = (size_t)e.out * KP + (size_t)c * a.P + p;
      a.partials[o] = x;
      a.exps[o] = (int16_t)E;
    }
    const uint32_t push = e.meta & 0xFFu;
    if (push != NO_PUSH) {
      stk[push * 64 + lane] = x;
      stke[push * 64 + lane] = E;
    }
  }

  // root: ln(pi . x) + E ln 2 in FP64 (cpu4.rs:277-284, cpu4.rs:69-73)
  double site = -INFINITY;
  if (active) {
    double v = (double)x.x * a.pi0 + (double)x.y * a.pi1;
    v = v + (double)x.z * a.pi2;
    v = v + (double)x.w * a.pi3;
    site = log(v) + (double)E * LN2;
    a.site[(size_t)c * a.P + p] = site;
  }
  mix[w * 64 + lane] = site;
  __syncthreads();
  double wv = 0.0;
  if (c == 0) {
    double mxv = -INFINITY;
    for (uint32_t k = 0; k < a.K; k++) mxv = fmax(mxv, mix[(t * a.K + k) * 64 + lane]);
    double mixed = -INFINITY;
    if (mxv != -INFINITY) {
      double se = 0.0;
      for (uint32_t k = 0; k < a.K; k++) se += exp(mix[(t * a.K + k) * 64 + lane] - mxv);
      mixed = mxv + log(se) + a.ln_inv_k;
    }
    if (active) {
      a.mixed[p] = mixed;
      wv = mixed * (double)a.weights[p];
    }
    wv = wave_sum(wv);
    if (lane == 0) red[t] = wv;
  }
  __syncthreads();
  double bsum = 0.0;
  if (threadIdx.x == 0)
    for (uint32_t tt = 0; tt < a.T; tt++) bsum += red[tt];
  last_block_total(bsum, a.blocksum, a.counter, a.total, flag);
}

size_t prune_f32_lds_bytes(uint32_t K, uint32_t T, uint32_t nslots) {
  const size_t W = (size_t)K * T;
  return W * nslots * 64 * (16 + 4) + W * 64 * 8 + T * 8 + 16;
}

hipError_t prepare_prune_f32(int max_lds_bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&prune_f32_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             max_lds_bytes);
}

hipError_t launch_prune_f32(const F32Args &a, uint32_t blocks, hipStream_t s) {
  const size_t lds = prune_f32_lds_bytes(a.K, a.T, a.nslots);
  const uint32_t threads = 64u * a.K * a.T;
  hipLaunchKernelGGL(prune_f32_kernel, dim3(blocks), dim3(threads),
                     (uint32_t)lds, s, a);
  return hipGetLastError();
}

}  // namespace b3h
