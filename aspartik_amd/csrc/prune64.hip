// FP64 plain pruning (the baseline the FP32 kernel is measured against) and
// the category-mix / weighted-sum reduction it uses.
//
// Semantics follow b3's CPU calculator exactly (crates/b3/src/calculator/
// cpu4.rs:170-433): FP64 partials in two slots per internal node, a per-node
// flag set when all four values fall below e^-scale_ln (the values are then
// multiplied by e^scale_ln), a per-pattern count of scaled nodes updated
// incrementally against the other slot's flag, cherries never scaled
// (cpu4.rs:288-318), root ln(pi . x) - count * scale_ln.  Like b3's GPU
// kernel (cuda_partials/kernels.cu:36-140) one lane handles one pattern and
// walks the update list, reading children from HBM; here blockIdx.y selects
// the rate category.  Floating-point contraction is off so each product and
// sum rounds as in the reference.
#include "common.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

namespace b3h {

namespace {

__device__ __forceinline__ double tipv(const double *M, int row,
                                       uint32_t code) {
  return code < 4 ? M[row * 4 + code] : 1.0;  // cpu4.rs:53-61 fifth column
}

}  // namespace

__global__ __launch_bounds__(256) void prune_f64_kernel(F64Args a) {
#pragma clang fp contract(off)
  const uint32_t c = blockIdx.y;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.P) return;
  const size_t KP = (size_t)a.K * a.P;
  const size_t cp = (size_t)c * a.P + p;
  uint32_t ss = a.scale_sums[cp];
  double x0 = 0, x1 = 0, x2 = 0, x3 = 0;
  for (uint32_t i = 0; i < a.n; ++i) {
    const Entry e = a.entries[i];
    const double *tl = a.pmats + ((size_t)(2 * i) * a.K + c) * 16;
    const double *tr = a.pmats + ((size_t)(2 * i + 1) * a.K + c) * 16;
    const size_t onew = (size_t)e.out * KP + cp;
    const size_t oold = (size_t)(e.out ^ 1u) * KP + cp;
    const uint8_t old = a.scales[oold];
    const bool ltip = (e.left & SRC_TAG) == SRC_TIP;
    const bool rtip = (e.right & SRC_TAG) == SRC_TIP;
    double p0, p1, p2, p3;
    if (rtip) {  // both children are tips: calc_sample_sample
      const uint32_t sl = a.codes[(size_t)(e.left & SRC_IDX) * a.cstride + p];
      const uint32_t sr = a.codes[(size_t)(e.right & SRC_IDX) * a.cstride + p];
      p0 = tipv(tl, 0, sl) * tipv(tr, 0, sr);
      p1 = tipv(tl, 1, sl) * tipv(tr, 1, sr);
      p2 = tipv(tl, 2, sl) * tipv(tr, 2, sr);
      p3 = tipv(tl, 3, sl) * tipv(tr, 3, sr);
      if (old) ss -= 1;
      a.scales[onew] = 0;
    } else {
      const double4 y = a.partials[(size_t)(e.right & SRC_IDX) * KP + cp];
      double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
      r0 += tr[0] * y.x; r0 += tr[1] * y.y; r0 += tr[2] * y.z; r0 += tr[3] * y.w;
      r1 += tr[4] * y.x; r1 += tr[5] * y.y; r1 += tr[6] * y.z; r1 += tr[7] * y.w;
      r2 += tr[8] * y.x; r2 += tr[9] * y.y; r2 += tr[10] * y.z; r2 += tr[11] * y.w;
      r3 += tr[12] * y.x; r3 += tr[13] * y.y; r3 += tr[14] * y.z; r3 += tr[15] * y.w;
      double l0, l1, l2, l3;
      if (ltip) {  // calc_sample_partial
        const uint32_t sl = a.codes[(size_t)(e.left & SRC_IDX) * a.cstride + p];
        l0 = tipv(tl, 0, sl); l1 = tipv(tl, 1, sl);
        l2 = tipv(tl, 2, sl); l3 = tipv(tl, 3, sl);
      } else {  // calc_partial_partial
        const double4 z = a.partials[(size_t)(e.left & SRC_IDX) * KP + cp];
        l0 = 0.0; l1 = 0.0; l2 = 0.0; l3 = 0.0;
        l0 += tl[0] * z.x; l0 += tl[1] * z.y; l0 += tl[2] * z.z; l0 += tl[3] * z.w;
        l1 += tl[4] * z.x; l1 += tl[5] * z.y; l1 += tl[6] * z.z; l1 += tl[7] * z.w;
        l2 += tl[8] * z.x; l2 += tl[9] * z.y; l2 += tl[10] * z.z; l2 += tl[11] * z.w;
        l3 += tl[12] * z.x; l3 += tl[13] * z.y; l3 += tl[14] * z.z; l3 += tl[15] * z.w;
      }
      p0 = l0 * r0; p1 = l1 * r1; p2 = l2 * r2; p3 = l3 * r3;
      // scale (cpu4.rs:403-433)
      if (p0 < a.threshold && p1 < a.threshold && p2 < a.threshold &&
          p3 < a.threshold) {
        p0 *= a.mult; p1 *= a.mult; p2 *= a.mult; p3 *= a.mult;
        if (!old) ss += 1;
        a.scales[onew] = 1;
      } else {
        if (old) ss -= 1;
        a.scales[onew] = 0;
      }
    }
    a.partials[onew] = make_double4(p0, p1, p2, p3);
    x0 = p0; x1 = p1; x2 = p2; x3 = p3;
  }
  const double *pi = a.pik ? a.pik + 4 * c : nullptr;
  const double h0 = x0 * (pi ? pi[0] : a.pi0), h1 = x1 * (pi ? pi[1] : a.pi1),
               h2 = x2 * (pi ? pi[2] : a.pi2), h3 = x3 * (pi ? pi[3] : a.pi3);
  double lik = log(((h0 + h1) + h2) + h3);
  lik -= (double)(uint32_t)(ss * a.scale_ln);
  a.site[cp] = lik;
  a.scale_sums[cp] = ss;
}

hipError_t launch_prune_f64(const F64Args &a, hipStream_t s) {
  dim3 grid((a.P + 255) / 256, a.K);
  hipLaunchKernelGGL(prune_f64_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

// One lane per pattern: LSE over categories + ln(1/K), weighted, reduced.
__global__ __launch_bounds__(256) void mix_kernel(MixArgs a) {
  __shared__ double red[4];
  __shared__ int flag;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  double wv = 0.0;
  if (p < a.P) {
    double mixed = -INFINITY;
    if (a.classes) {  // hetero.rs:150-160
      mixed = a.site[(size_t)a.classes[p] * a.P + p];
    } else {
      double mx = -INFINITY;
      for (uint32_t k = 0; k < a.K; k++) mx = fmax(mx, a.site[(size_t)k * a.P + p]);
      if (mx != -INFINITY) {
        double se = 0.0;
        for (uint32_t k = 0; k < a.K; k++) se += exp(a.site[(size_t)k * a.P + p] - mx);
        mixed = mx + log(se) + a.ln_inv_k;
      }
    }
    a.mixed[p] = mixed;
    wv = mixed * (double)a.weights[p];
  }
  wv = wave_sum(wv);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wv;
  __syncthreads();
  double b = 0.0;
  if (threadIdx.x == 0) b = ((red[0] + red[1]) + red[2]) + red[3];
  last_block_total(b, a.blocksum, a.counter, a.total, a.host_total, a.host_flag, a.seq, &flag,
                   blockIdx.x, gridDim.x);
}

uint32_t mix_blocks(uint32_t P) { return (P + 255) / 256; }

hipError_t launch_mix(const MixArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(mix_kernel, dim3(mix_blocks(a.P)), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace b3h
