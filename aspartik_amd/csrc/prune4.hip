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
//  - the host orders the update list depth-first (Sethi-Ullman order); a
//    child computed in the same pass is taken from the lane's registers (the
//    node computed just before) or from a per-wavefront LDS stack instead of
//    HBM.  Every node is still written to HBM because later partial updates
//    read it.
//  - nothing on a node's critical path waits on memory latency: the whole
//    update list is staged in LDS at kernel start; transition matrices and
//    the tile's tip codes stream through a double-buffered LDS chunk whose
//    next chunk is fetched (into registers) while the current one is used;
//    children stored in HBM (partial updates) are fetched one node ahead.
#include "common.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

namespace b3h {

namespace {

// Partials are written with nontemporal stores: they are read again only in
// a later evaluation, and streaming them past the caches measured ~25%
// faster than plain stores for this write pattern.
#ifndef B3H_NT_STORE
#define B3H_NT_STORE 1
#endif

typedef float f4v __attribute__((ext_vector_type(4)));

constexpr uint32_t CH = 8;          // update-list entries per LDS chunk
constexpr uint32_t MAT_REGS = 4;   // chunk matrix dwords per thread (T >= 1)
constexpr uint32_t CODE_REGS = 4;  // chunk code dwords per thread (K >= 1)

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

__device__ __forceinline__ float4 apply(const float4 *M4, float4 y) {
  const float4 r0 = M4[0], r1 = M4[1], r2 = M4[2], r3 = M4[3];
  float4 o;
  o.x = r0.x * y.x + r0.y * y.y + r0.z * y.z + r0.w * y.w;
  o.y = r1.x * y.x + r1.y * y.y + r1.z * y.z + r1.w * y.w;
  o.z = r2.x * y.x + r2.y * y.y + r2.z * y.z + r2.w * y.w;
  o.w = r3.x * y.x + r3.y * y.y + r3.z * y.z + r3.w * y.w;
  return o;
}

__device__ __forceinline__ float4 tip(const float4 *M4, uint32_t code) {
  const float4 r0 = M4[0], r1 = M4[1], r2 = M4[2], r3 = M4[3];
  float4 o;
  o.x = col_or_one(r0.x, r0.y, r0.z, r0.w, code);
  o.y = col_or_one(r1.x, r1.y, r1.z, r1.w, code);
  o.z = col_or_one(r2.x, r2.y, r2.z, r2.w, code);
  o.w = col_or_one(r3.x, r3.y, r3.z, r3.w, code);
  return o;
}

__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

}  // namespace

size_t prune_f32_lds_bytes(uint32_t K, uint32_t T, uint32_t nslots, uint32_t n) {
  const size_t W = (size_t)K * T;
  const size_t TP = 64 * (size_t)T;
  size_t b = (size_t)n * 16;                      // entries
  b += 2 * ((size_t)CH * 2 * K * 64);             // matrix chunks
  b += 2 * ((size_t)CH * 2 * TP);                 // code chunks
  b = (b + 15) & ~(size_t)15;
  b += W * nslots * 64 * (16 + 4);                // stacks
  b += W * 64 * 8 + T * 8 + 16;                   // mix, red, flag
  return b;
}

__global__ __launch_bounds__(1024) void prune_f32_kernel(F32Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t nthr = blockDim.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t w = uni(tid >> 6);
  const uint32_t K = a.K, T = a.T, n = a.n;
  const uint32_t W = K * T;
  const uint32_t c = w % K, t = w / K;
  const uint32_t TP = 64u * T;
  const uint32_t tile0 = blockIdx.x * TP;
  const uint32_t p = tile0 + t * 64u + lane;
  const bool active = p < a.P;
  const uint32_t pp = active ? p : 0u;
  const uint32_t ns = a.nslots;
  const size_t KP = (size_t)K * a.P;
  const uint32_t MATC = CH * 2 * K * 16;  // floats per matrix chunk
  const uint32_t CODEC = CH * 2 * TP;     // bytes per code chunk

  // ---- LDS carve-up (mirrors prune_f32_lds_bytes) ----
  unsigned char *sp = smem;
  uint4 *ent_s = reinterpret_cast<uint4 *>(sp);
  sp += (size_t)n * 16;
  float *mat_s = reinterpret_cast<float *>(sp);  // [2][MATC]
  sp += 2 * (size_t)MATC * 4;
  unsigned char *code_s = sp;  // [2][CODEC]
  sp += 2 * (size_t)CODEC;
  sp = smem + ((sp - smem + 15) & ~(size_t)15);
  float4 *stk_all = reinterpret_cast<float4 *>(sp);
  int *stke_all = reinterpret_cast<int *>(stk_all + (size_t)W * ns * 64);
  double *mix = reinterpret_cast<double *>(stke_all + (size_t)W * ns * 64);
  double *red = mix + (size_t)W * 64;
  int *flag = reinterpret_cast<int *>(red + T);
  float4 *stk = stk_all + (size_t)w * ns * 64;
  int *stke = stke_all + (size_t)w * ns * 64;

  // ---- stage the update list ----
  const uint4 *g_ent = reinterpret_cast<const uint4 *>(a.entries);
  for (uint32_t i = tid; i < n; i += nthr) ent_s[i] = g_ent[i];
  __syncthreads();

  // Chunk fetch into registers / store into an LDS buffer.  Code rows are
  // padded to a multiple of 64 patterns on the device (cstride), so a tile's
  // codes of one leaf are TP/4 aligned dwords.
  float mreg[MAT_REGS];
  uint32_t creg[CODE_REGS];
  const uint32_t TPW = TP / 4;  // code dwords per tip row
#define B3H_FETCH(ck)                                                          \
  do {                                                                         \
    const uint32_t base_ = (ck) * CH;                                          \
    const uint32_t cnt_ = min(CH, n - base_);                                  \
    const uint32_t nm_ = cnt_ * 2 * K * 16;                                    \
    const float *gm_ = a.pmats + (size_t)base_ * 2 * K * 16;                   \
    _Pragma("unroll") for (uint32_t r_ = 0; r_ < MAT_REGS; r_++) {             \
      const uint32_t q_ = tid + r_ * nthr;                                     \
      mreg[r_] = q_ < nm_ ? gm_[q_] : 0.0f;                                    \
    }                                                                          \
    const uint32_t nc_ = cnt_ * 2 * TPW;                                       \
    _Pragma("unroll") for (uint32_t r_ = 0; r_ < CODE_REGS; r_++) {            \
      const uint32_t q_ = tid + r_ * nthr;                                     \
      uint32_t v_ = 0x04040404u;                                               \
      if (q_ < nc_) {                                                          \
        const uint32_t jj_ = q_ / (2 * TPW);                                   \
        const uint32_t side_ = (q_ / TPW) & 1u;                                \
        const uint32_t k_ = q_ % TPW;                                          \
        const uint4 e_ = ent_s[base_ + jj_];                                   \
        const uint32_t src_ = side_ ? e_.z : e_.y;                             \
        if ((src_ & SRC_TAG) == SRC_TIP)                                       \
          v_ = a.codes32[(size_t)(src_ & SRC_IDX) * (a.cstride / 4) +          \
                         tile0 / 4 + k_];                                      \
      }                                                                        \
      creg[r_] = v_;                                                           \
    }                                                                          \
  } while (0)
#define B3H_STORE(buf, ck)                                                     \
  do {                                                                         \
    const uint32_t base_ = (ck) * CH;                                          \
    const uint32_t cnt_ = min(CH, n - base_);                                  \
    const uint32_t nm_ = cnt_ * 2 * K * 16;                                    \
    const uint32_t nc_ = cnt_ * 2 * TPW;                                       \
    float *ms_ = mat_s + (size_t)(buf) * MATC;                                 \
    uint32_t *cs_ = reinterpret_cast<uint32_t *>(code_s + (size_t)(buf) * CODEC); \
    _Pragma("unroll") for (uint32_t r_ = 0; r_ < MAT_REGS; r_++) {             \
      const uint32_t q_ = tid + r_ * nthr;                                     \
      if (q_ < nm_) ms_[q_] = mreg[r_];                                        \
    }                                                                          \
    _Pragma("unroll") for (uint32_t r_ = 0; r_ < CODE_REGS; r_++) {            \
      const uint32_t q_ = tid + r_ * nthr;                                     \
      if (q_ < nc_) cs_[q_] = creg[r_];                                        \
    }                                                                          \
  } while (0)

  const uint32_t nchunks = (n + CH - 1) / CH;
  B3H_FETCH(0u);
  B3H_STORE(0u, 0u);
  __syncthreads();

  // HBM children of the next node, fetched one node ahead
  float4 hl = make_float4(0.f, 0.f, 0.f, 0.f), hr = hl;
  int hel = 0, her = 0;
  auto fetch_hbm = [&](uint32_t i, float4 &vl, int &el, float4 &vr, int &er) {
    const uint4 e = ent_s[i];
    const uint32_t lsrc = uni(e.y), rsrc = uni(e.z);
    if ((lsrc & SRC_TAG) == SRC_HBM) {
      const size_t o = (size_t)(lsrc & SRC_IDX) * KP + (size_t)c * a.P + pp;
      vl = a.partials[o];
      el = a.exps[o];
    }
    if ((rsrc & SRC_TAG) == SRC_HBM) {
      const size_t o = (size_t)(rsrc & SRC_IDX) * KP + (size_t)c * a.P + pp;
      vr = a.partials[o];
      er = a.exps[o];
    }
  };
  fetch_hbm(0, hl, hel, hr, her);

  float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
  int E = 0;
  for (uint32_t ck = 0; ck < nchunks; ck++) {
    const uint32_t buf = ck & 1u;
    const bool more = ck + 1 < nchunks;
    if (more) B3H_FETCH(ck + 1);
    const uint32_t base = ck * CH;
    const uint32_t cnt = min(CH, n - base);
    const float4 *ms = reinterpret_cast<const float4 *>(mat_s + (size_t)buf * MATC);
    const unsigned char *cs = code_s + (size_t)buf * CODEC;
    for (uint32_t j = 0; j < cnt; j++) {
      const uint32_t i = base + j;
      const uint4 e4 = ent_s[i];
      const uint32_t out = uni(e4.x), lsrc = uni(e4.y), rsrc = uni(e4.z),
                     meta = uni(e4.w);
      const float4 cl = hl, cr = hr;
      const int cel = hel, cer = her;
      if (i + 1 < n) fetch_hbm(i + 1, hl, hel, hr, her);
      const float4 *ML = ms + ((size_t)(j * 2 + 0) * K + c) * 4;
      const float4 *MR = ms + ((size_t)(j * 2 + 1) * K + c) * 4;
      Src l, r;
      {
        const uint32_t tag = lsrc & SRC_TAG, idx = lsrc & SRC_IDX;
        if (tag == SRC_TIP) {
          l.v = tip(ML, cs[(j * 2 + 0) * TP + t * 64 + lane]);
          l.e = 0;
        } else if (tag == SRC_PREV) {
          l.v = apply(ML, x);
          l.e = E;
        } else if (tag == SRC_LDS) {
          l.v = apply(ML, stk[idx * 64 + lane]);
          l.e = stke[idx * 64 + lane];
        } else {
          l.v = apply(ML, cl);
          l.e = cel;
        }
      }
      {
        const uint32_t tag = rsrc & SRC_TAG, idx = rsrc & SRC_IDX;
        if (tag == SRC_TIP) {
          r.v = tip(MR, cs[(j * 2 + 1) * TP + t * 64 + lane]);
          r.e = 0;
        } else if (tag == SRC_PREV) {
          r.v = apply(MR, x);
          r.e = E;
        } else if (tag == SRC_LDS) {
          r.v = apply(MR, stk[idx * 64 + lane]);
          r.e = stke[idx * 64 + lane];
        } else {
          r.v = apply(MR, cr);
          r.e = cer;
        }
      }
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
        const size_t o = (size_t)out * KP + (size_t)c * a.P + p;
#if B3H_NT_STORE
        const f4v xv = {x.x, x.y, x.z, x.w};
        __builtin_nontemporal_store(xv, reinterpret_cast<f4v *>(&a.partials[o]));
        __builtin_nontemporal_store((int16_t)E, &a.exps[o]);
#else
        a.partials[o] = x;
        a.exps[o] = (int16_t)E;
#endif
      }
      const uint32_t push = meta & 0xFFu;
      if (push != NO_PUSH) {
        stk[push * 64 + lane] = x;
        stke[push * 64 + lane] = E;
      }
    }
    if (more) B3H_STORE(buf ^ 1u, ck + 1);
    __syncthreads();
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
    for (uint32_t k = 0; k < K; k++) mxv = fmax(mxv, mix[(t * K + k) * 64 + lane]);
    double mixed = -INFINITY;
    if (mxv != -INFINITY) {
      double se = 0.0;
      for (uint32_t k = 0; k < K; k++) se += exp(mix[(t * K + k) * 64 + lane] - mxv);
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
  if (tid == 0)
    for (uint32_t tt = 0; tt < T; tt++) bsum += red[tt];
  last_block_total(bsum, a.blocksum, a.counter, a.total, flag);
#undef B3H_FETCH
#undef B3H_STORE
}

hipError_t prepare_prune_f32(int max_lds_bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&prune_f32_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             max_lds_bytes);
}

hipError_t launch_prune_f32(const F32Args &a, uint32_t blocks, hipStream_t s) {
  const size_t lds = prune_f32_lds_bytes(a.K, a.T, a.nslots, a.n);
  const uint32_t threads = 64u * a.K * a.T;
  hipLaunchKernelGGL(prune_f32_kernel, dim3(blocks), dim3(threads),
                     (uint32_t)lds, s, a);
  return hipGetLastError();
}

}  // namespace b3h
