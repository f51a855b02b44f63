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
//  - a workgroup is K x T wavefronts; wavefront (c, t) owns category c of
//    64 x PPL patterns of sub-tile t.  A lane handles PPL patterns 64 apart,
//    so one partial is one float4 (16 B) and a wavefront moves PPL KiB per
//    node in coalesced 1 KiB rows; PPL independent chains per lane hide
//    latency and halve the wavefront count so a whole evaluation fits on the
//    chip at once.  (prune_tile_s1 is the one-state-per-lane form: a quad of
//    lanes per pattern, 16 patterns per wavefront, bit-identical values.)
//  - partials [slot][K][P] float4 with an int16 cumulative binary exponent per
//    (slot, category, pattern).  After every node the four values are scaled
//    by 2^-e (e = exponent of their maximum), which is exact, so FP32 never
//    underflows and no rounding comes from scaling.
//  - the host orders the update list depth-first (Sethi-Ullman order); a
//    child computed in the same pass is taken from the lane's registers (the
//    node computed just before) or from a per-wavefront LDS stack instead of
//    HBM.  Every node is still written to HBM because later partial updates
//    read it, except cherries (both children leaves): a later reader
//    recomputes a cherry from its two leaves' codes and the two matrix
//    columns the P(t) build recorded for its slot, bit for bit.
//  - transition matrices are [5][4] floats: four columns and an all-ones
//    column for the unknown code, so a tip child is one 16-byte LDS read of
//    column `code` and an internal child four broadcast column reads.
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

constexpr uint32_t CH = 4;  // update-list entries per LDS chunk

// A 4-state transition matrix as its four columns: c_k = P[.][k].  In memory
// every matrix of the FP32 path is [5][4] floats: the four columns and an
// all-ones fifth column, the value of an unknown tip code (4), so a tip's
// projection P[.][code] is one 16-byte read at column `code`.
struct Mat {
  float4 c0, c1, c2, c3;
};

// P y with explicit fused multiply-adds in a fixed order (per component
// j: P[j][0] y_0, then + P[j][1] y_1, + P[j][2] y_2, + P[j][3] y_3), so every
// kernel variant (tile shape, patterns per lane) rounds identically.
__device__ __forceinline__ float4 apply(const Mat &M, float4 y) {
  float4 o;
  o.x = fmaf(M.c3.x, y.w, fmaf(M.c2.x, y.z, fmaf(M.c1.x, y.y, M.c0.x * y.x)));
  o.y = fmaf(M.c3.y, y.w, fmaf(M.c2.y, y.z, fmaf(M.c1.y, y.y, M.c0.y * y.x)));
  o.z = fmaf(M.c3.z, y.w, fmaf(M.c2.z, y.z, fmaf(M.c1.z, y.y, M.c0.z * y.x)));
  o.w = fmaf(M.c3.w, y.w, fmaf(M.c2.w, y.z, fmaf(M.c1.w, y.y, M.c0.w * y.x)));
  return o;
}

// a tip's projection from matrix columns held in registers: column `code`,
// or all ones for the unknown code 4 (selected per component: a select of
// whole vectors makes the compiler index a private copy)
__device__ __forceinline__ float pick(float a0, float a1, float a2, float a3, uint32_t code) {
  float r = 1.0f;
  r = code == 0 ? a0 : r;
  r = code == 1 ? a1 : r;
  r = code == 2 ? a2 : r;
  r = code == 3 ? a3 : r;
  return r;
}
__device__ __forceinline__ float4 tip(const Mat &M, uint32_t code) {
  float4 o;
  o.x = pick(M.c0.x, M.c1.x, M.c2.x, M.c3.x, code);
  o.y = pick(M.c0.y, M.c1.y, M.c2.y, M.c3.y, code);
  o.z = pick(M.c0.z, M.c1.z, M.c2.z, M.c3.z, code);
  o.w = pick(M.c0.w, M.c1.w, M.c2.w, M.c3.w, code);
  return o;
}

__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ Mat load_mat(const float4 *p) {
  Mat m;
  m.c0 = p[0];
  m.c1 = p[1];
  m.c2 = p[2];
  m.c3 = p[3];
  return m;
}

// x <- x / 2^e with e the binary exponent of max_j x_j (exact); returns e.
// When all four values are 0 the site's likelihood is 0 whatever e is (its
// ln L is -inf), so zero needs no separate case.
__device__ __forceinline__ int normalise(float4 &v) {
  const float mx = fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w));
  const int ex = __builtin_amdgcn_frexp_expf(mx);
  v.x = __builtin_amdgcn_ldexpf(v.x, -ex);
  v.y = __builtin_amdgcn_ldexpf(v.y, -ex);
  v.z = __builtin_amdgcn_ldexpf(v.z, -ex);
  v.w = __builtin_amdgcn_ldexpf(v.w, -ex);
  return ex;
}

#if defined(__HIP_DEVICE_COMPILE__)
#define B3H_CONSTANT __attribute__((address_space(4)))
#else
#define B3H_CONSTANT
#endif
typedef const B3H_CONSTANT float4 ConstF4;
typedef const B3H_CONSTANT uint4 ConstU4;
typedef const B3H_CONSTANT uint2 ConstU2;
typedef const B3H_CONSTANT uint32_t ConstU32;

__device__ __forceinline__ Mat load_mat_uniform(ConstF4 *p) {
  Mat m;
  m.c0 = p[0];
  m.c1 = p[1];
  m.c2 = p[2];
  m.c3 = p[3];
  return m;
}

// A cherry slot (both children leaves; its partial is not stored) is
// recomputed from its two leaves' codes and the two child matrices the P(t)
// build recorded for the slot as columns: cmats [slot][side][K][5][4], column
// 4 all ones (the unknown code), so a lane loads the one column its code
// selects -- the values tip() selects, so the bits are those the pass that
// computed the cherry had.
__device__ __forceinline__ float4 cherry_col(const float *cmats, uint32_t slot, uint32_t side,
                                             uint32_t K, uint32_t c, uint32_t code) {
  const float4 *m = reinterpret_cast<const float4 *>(cmats);
  return m[((size_t)(slot * 2 + side) * K + c) * 5 + code];
}
// codes packed a | b << 8
__device__ __forceinline__ float4 cherry_value(const float *cmats, uint32_t slot, uint32_t K,
                                               uint32_t c, uint32_t codes, int &e) {
  const float4 l = cherry_col(cmats, slot, 0, K, c, codes & 0xFFu);
  const float4 r = cherry_col(cmats, slot, 1, K, c, (codes >> 8) & 0xFFu);
  float4 v;
  v.x = l.x * r.x;
  v.y = l.y * r.y;
  v.z = l.z * r.z;
  v.w = l.w * r.w;
  e = exp_clamp(normalise(v));
  return v;
}

}  // namespace

size_t prune_f32_lds_bytes(uint32_t K, uint32_t T, uint32_t nslots, uint32_t n,
                           uint32_t ppl) {
  const size_t W = (size_t)K * T;
  const size_t TP = 64 * (size_t)T * ppl;
  size_t b = (size_t)n * 16;               // entries
  b += 2 * ((size_t)CH * 2 * K * 80);      // matrix chunks ([5][4] floats each)
  b += 2 * ((size_t)CH * 2 * TP);          // code chunks
  b = (b + 15) & ~(size_t)15;
  // the stacks (float4 + int16 per lane and slot) are dead once the node loop
  // ends, and the category-mix buffers reuse their space
  const size_t stacks = W * nslots * ppl * 64 * (16 + 2);
  const size_t mixbuf = W * ppl * 64 * 8 + T * 8 + 16;
  b += stacks > mixbuf ? stacks : mixbuf;
  return b;
}

size_t prune_f32_s1_lds_bytes(uint32_t K, uint32_t T, uint32_t nslots, uint32_t n) {
  const size_t W = (size_t)K * T;
  const size_t TP = 16 * (size_t)T;
  size_t b = (size_t)n * 16;
  b += 2 * ((size_t)CH * 2 * K * 80);
  b += 2 * ((size_t)CH * 2 * TP);
  b = (b + 15) & ~(size_t)15;
  const size_t stacks = W * nslots * 64 * (4 + 2);
  const size_t mixbuf = W * 16 * 8 + T * 8 + 16;
  b += stacks > mixbuf ? stacks : mixbuf;
  return b;
}

// HBM: the update list reads children from HBM (partial updates); those are
// then fetched one node ahead.  Without it (full evaluations) the rare HBM
// child (LDS stack overflow) is loaded where it is used, and the prefetch
// registers are not allocated.
// `A` is the argument record's type: the kernel argument itself, or a record
// in the constant address space (batched form), whose fields the compiler
// may reload instead of keeping them in scalar registers.
template <int PPL, bool HBM, typename A>
__device__ __forceinline__ void prune_tile(A &a, const uint32_t tile, unsigned char *smem) {
  constexpr uint32_t MAT_REGS = 4;        // matrix-chunk dwords per thread
  constexpr uint32_t CODE_REGS = 4 * PPL;  // code-chunk dwords per thread
  const uint32_t nthr = blockDim.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t w = uni(tid >> 6);
  const uint32_t K = a.K, T = a.T, n = a.n;
  const uint32_t W = K * T;
  const uint32_t c = w % K, t = w / K;
  const uint32_t TP = 64u * T * PPL;
  const uint32_t tile0 = tile * TP;
  uint32_t p[PPL], pp[PPL];
  bool active[PPL];
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    p[q] = tile0 + t * 64u * PPL + q * 64u + lane;
    active[q] = p[q] < a.P;
    pp[q] = active[q] ? p[q] : 0u;
  }
  // byte offsets of this lane's patterns in a partial row and an exponent row
  uint32_t poff[PPL], eoff[PPL];
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    poff[q] = p[q] * 16u;
    eoff[q] = p[q] * 2u;
  }
  const uint32_t ns = a.nslots;
  const size_t KP = (size_t)K * a.P;
  const uint8_t *codes8 = reinterpret_cast<const uint8_t *>(a.codes32);
  const uint32_t MATC = CH * 2 * K * 20;  // floats per matrix chunk
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
  float4 *stk_all = reinterpret_cast<float4 *>(sp);  // [W][ns][PPL][64]
  int16_t *stke_all = reinterpret_cast<int16_t *>(stk_all + (size_t)W * ns * PPL * 64);
  double *mix = reinterpret_cast<double *>(sp);  // after the node loop only
  double *red = mix + (size_t)W * PPL * 64;
  int *flag = reinterpret_cast<int *>(red + T);
  float4 *stk = stk_all + (size_t)w * ns * PPL * 64;
  int16_t *stke = stke_all + (size_t)w * ns * PPL * 64;

  // ---- stage the update list ----
  const uint4 *g_ent = reinterpret_cast<const uint4 *>(a.entries);
  for (uint32_t i = tid; i < n; i += nthr) ent_s[i] = g_ent[i];
  __syncthreads();

  // Chunk fetch into registers / store into an LDS buffer.  Code rows are
  // padded to whole tiles on the device (cstride), so a tile's codes of one
  // leaf are TP/4 aligned dwords.
  float mreg[MAT_REGS];
  uint32_t creg[CODE_REGS];
  const uint32_t TPW = TP / 4;
#define B3H_FETCH(ck)                                                          \
  do {                                                                         \
    const uint32_t base_ = (ck) * CH;                                          \
    const uint32_t cnt_ = min(CH, n - base_);                                  \
    const uint32_t nm_ = cnt_ * 2 * K * 20;                                    \
    const float *gm_ = a.pmats + (size_t)base_ * 2 * K * 20;                   \
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
    const uint32_t nm_ = cnt_ * 2 * K * 20;                                    \
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
  float4 hl[PPL], hr[PPL];
  int hel[PPL], her[PPL];
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    hl[q] = hr[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    hel[q] = her[q] = 0;
  }
  // a stored child: its partial and exponent, or for a cherry the two codes
  // (packed in the exponent register) it is recomputed from
  auto fetch_hbm = [&](uint32_t sv, float4 *hv, int *he) {
    const uint32_t slot = sv & SRC_SLOT;
    if (sv & SRC_CHERRY) {
      const uint2 lv = ((ConstU2 *)a.cleaves)[slot];
#pragma unroll
      for (int q = 0; q < PPL; q++) {
        hv[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        he[q] = (int)codes8[(size_t)lv.x * a.cstride + pp[q]] |
                ((int)codes8[(size_t)lv.y * a.cstride + pp[q]] << 8);
      }
    } else {
      const size_t o = (size_t)slot * KP + (size_t)c * a.P;
#pragma unroll
      for (int q = 0; q < PPL; q++) {
        hv[q] = a.partials[o + pp[q]];
        he[q] = exp_unpack(a.exps[o + pp[q]]);
      }
    }
  };
#define B3H_FETCH_HBM(i)                                                       \
  do {                                                                         \
    const uint4 e_ = ent_s[(i)];                                               \
    const uint32_t ls_ = uni(e_.y), rs_ = uni(e_.z);                           \
    if ((ls_ & SRC_TAG) == SRC_HBM) fetch_hbm(ls_, hl, hel);                   \
    if ((rs_ & SRC_TAG) == SRC_HBM) fetch_hbm(rs_, hr, her);                   \
  } while (0)
  if (HBM) B3H_FETCH_HBM(0u);

  float4 x[PPL];
  int E[PPL];
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    x[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    E[q] = 0;
  }
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
      // through the scalar cache: no vector registers, no readfirstlane
      const uint4 e4 = ((ConstU4 *)a.entries)[i];
      const uint32_t out = uni(e4.x), lsrc = uni(e4.y), rsrc = uni(e4.z),
                     meta = uni(e4.w);
      float4 cl[PPL], cr[PPL];
      int cel[PPL], cer[PPL];
      if (HBM) {
#pragma unroll
        for (int q = 0; q < PPL; q++) {
          cl[q] = hl[q];
          cr[q] = hr[q];
          cel[q] = hel[q];
          cer[q] = her[q];
        }
        if (i + 1 < n) B3H_FETCH_HBM(i + 1);
      }
      // one child: a tip is its matrix column `code` read from LDS (the
      // fifth column, all ones, for an unknown code); any other child is
      // projected with the four columns
      auto child = [&](uint32_t src, uint32_t side, const float4 *cl_, const int *cel_,
                       float4 *v, int *ee) {
        const float4 *mc = ms + ((size_t)(j * 2 + side) * K + c) * 5;
        const uint32_t tag = src & SRC_TAG, idx = src & SRC_IDX;
        if (tag == SRC_TIP) {
#pragma unroll
          for (int q = 0; q < PPL; q++) {
            v[q] = mc[cs[(j * 2 + side) * TP + t * 64 * PPL + q * 64 + lane]];
            ee[q] = 0;
          }
          return;
        }
        const Mat M = load_mat(mc);
        if (tag == SRC_PREV) {
#pragma unroll
          for (int q = 0; q < PPL; q++) {
            v[q] = apply(M, x[q]);
            ee[q] = E[q];
          }
        } else if (tag == SRC_LDS) {
#pragma unroll
          for (int q = 0; q < PPL; q++) {
            v[q] = apply(M, stk[(idx * PPL + q) * 64 + lane]);
            ee[q] = exp_unpack(stke[(idx * PPL + q) * 64 + lane]);
          }
        } else if (HBM) {
          // stored cherries come only with the HBM-prefetch variant (host rule)
          if (src & SRC_CHERRY) {
#pragma unroll
            for (int q = 0; q < PPL; q++)
              v[q] = apply(M, cherry_value(a.cmats, src & SRC_SLOT, K, c, (uint32_t)cel_[q], ee[q]));
          } else {
#pragma unroll
            for (int q = 0; q < PPL; q++) {
              v[q] = apply(M, cl_[q]);
              ee[q] = cel_[q];
            }
          }
        } else {
#pragma unroll
          for (int q = 0; q < PPL; q++) {
            const size_t o = (size_t)(idx & SRC_SLOT) * KP + (size_t)c * a.P + pp[q];
            v[q] = apply(M, a.partials[o]);
            ee[q] = exp_unpack(a.exps[o]);
          }
        }
      };
      float4 lv[PPL], rv[PPL];
      int le[PPL], re[PPL];
      child(lsrc, 0, cl, cel, lv, le);
      child(rsrc, 1, cr, cer, rv, re);
      const uint32_t push = meta & 0xFFu;
      const size_t ob = (size_t)out * KP + (size_t)c * a.P;
      float4 *const pout = a.partials + ob;
      int16_t *const eout = a.exps + ob;
      // a cherry's partial is not stored when its matrices are recorded,
      // unless this pass reads it back from memory (META_KEEP)
      const bool store = !(a.cmats && (lsrc & SRC_TAG) == SRC_TIP &&
                           (rsrc & SRC_TAG) == SRC_TIP) || (meta & META_KEEP);
#pragma unroll
      for (int q = 0; q < PPL; q++) {
        float4 v;
        v.x = lv[q].x * rv[q].x;
        v.y = lv[q].y * rv[q].y;
        v.z = lv[q].z * rv[q].z;
        v.w = lv[q].w * rv[q].w;
        const int ex = normalise(v);
        x[q] = v;
        E[q] = exp_clamp(le[q] + re[q] + ex);
        const int16_t ep = exp_pack(E[q]);
        if (active[q] && store) {
          // a uniform base plus the lane's loop-invariant 32-bit byte offset
          f4v *dp = reinterpret_cast<f4v *>(reinterpret_cast<char *>(pout) + poff[q]);
          int16_t *de = reinterpret_cast<int16_t *>(reinterpret_cast<char *>(eout) + eoff[q]);
#if B3H_NT_STORE
          const f4v xv = {v.x, v.y, v.z, v.w};
          __builtin_nontemporal_store(xv, dp);
          __builtin_nontemporal_store(ep, de);
#else
          *dp = f4v{v.x, v.y, v.z, v.w};
          *de = ep;
#endif
        }
        if (push != NO_PUSH) {
          stk[(push * PPL + q) * 64 + lane] = v;
          stke[(push * PPL + q) * 64 + lane] = ep;
        }
      }
    }
    if (more) B3H_STORE(buf ^ 1u, ck + 1);
    __syncthreads();
  }

  // root: ln(pi . x) + E ln 2 in FP64 (cpu4.rs:277-284, cpu4.rs:69-73); the
  // last node-loop barrier separates the stacks from the mix buffers.  With
  // per-category models each category has its own frequencies (hetero.rs:
  // 122-145 proposes each calculator with its own transitions).
  double pi0 = a.pi0, pi1 = a.pi1, pi2 = a.pi2, pi3 = a.pi3;
  if (a.pik) {
    pi0 = a.pik[4 * c];
    pi1 = a.pik[4 * c + 1];
    pi2 = a.pik[4 * c + 2];
    pi3 = a.pik[4 * c + 3];
  }
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    double site = -INFINITY;
    if (active[q]) {
      double v = (double)x[q].x * pi0;
      v = fma((double)x[q].y, pi1, v);
      v = fma((double)x[q].z, pi2, v);
      v = fma((double)x[q].w, pi3, v);
      site = exp_site(v, E[q]);
      a.site[(size_t)c * a.P + p[q]] = site;
    }
    mix[(w * PPL + q) * 64 + lane] = site;
  }
  __syncthreads();
  double wv = 0.0;
  if (c == 0 && a.classes) {
    // HeteroLikelihood: the pattern's class picks one category's value
    // (hetero.rs:150-160)
#pragma unroll
    for (int q = 0; q < PPL; q++) {
      if (active[q]) {
        const uint32_t k = a.classes[p[q]];
        const double mixed = mix[((t * K + k) * PPL + q) * 64 + lane];
        a.mixed[p[q]] = mixed;
        wv = fma(mixed, (double)a.weights[p[q]], wv);
      }
    }
    wv = wave_sum(wv);
    if (lane == 0) red[t] = wv;
  } else if (c == 0) {
#pragma unroll
    for (int q = 0; q < PPL; q++) {
      double mxv = -INFINITY;
      for (uint32_t k = 0; k < K; k++)
        mxv = fmax(mxv, mix[((t * K + k) * PPL + q) * 64 + lane]);
      double mixed = -INFINITY;
      if (mxv != -INFINITY) {
        double se = 0.0;
        for (uint32_t k = 0; k < K; k++)
          se = se + exp(mix[((t * K + k) * PPL + q) * 64 + lane] - mxv);
        mixed = (mxv + log(se)) + a.ln_inv_k;
      }
      if (active[q]) {
        a.mixed[p[q]] = mixed;
        wv = fma(mixed, (double)a.weights[p[q]], wv);
      }
    }
    wv = wave_sum(wv);
    if (lane == 0) red[t] = wv;
  }
  __syncthreads();
  double bsum = 0.0;
  if (tid == 0)
    for (uint32_t tt = 0; tt < T; tt++) bsum += red[tt];
  last_block_total(bsum, a.blocksum, a.counter, a.total, a.host_total, a.host_flag, a.seq, flag,
                   tile, a.ntiles);
#undef B3H_FETCH
#undef B3H_STORE
#undef B3H_FETCH_HBM
}

// One state per lane.  A quad of lanes holds one pattern's four states, so a
// wavefront covers 16 patterns and a tile four times as many wavefronts as
// the four-states-per-lane form: the shape for pattern counts too small to
// fill the chip, where a tile's walk through the update list is bound by
// latency rather than by issue or HBM.  Lane s of a quad keeps state s; a
// projection reads row s of the matrix and the child's four states from the
// quad (DPP broadcasts), in the order apply() uses, and the normalisation
// takes the quad's maximum, so every partial, exponent and site value is
// bit for bit what the other shapes compute.  Lists with stored children in
// HBM beyond the rare in-place load (partial updates, stored cherries) keep
// the four-states-per-lane forms (host rule).
template <int CTRL>
__device__ __forceinline__ float quad_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
constexpr int kQuadB0 = 0x00, kQuadB1 = 0x55, kQuadB2 = 0xAA, kQuadB3 = 0xFF;  // broadcast lane k
constexpr int kQuadSwap1 = 0xB1, kQuadSwap2 = 0x4E;  // [1,0,3,2], [2,3,0,1]

template <typename A>
__device__ __forceinline__ void prune_tile_s1(A &a, const uint32_t tile, unsigned char *smem) {
  constexpr uint32_t MAT_REGS = 4;   // matrix-chunk dwords per thread
  constexpr uint32_t CODE_REGS = 1;  // code-chunk dwords per thread
  const uint32_t nthr = blockDim.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t w = uni(tid >> 6);
  const uint32_t K = a.K, T = a.T, n = a.n;
  const uint32_t W = K * T;
  const uint32_t c = w % K, t = w / K;
  const uint32_t TP = 16u * T;  // patterns per tile
  const uint32_t tile0 = tile * TP;
  const uint32_t s = lane & 3u, pl = lane >> 2;  // state, pattern of the wavefront
  const uint32_t pt = t * 16u + pl;              // pattern of the tile
  const uint32_t p = tile0 + pt;
  const bool active = p < a.P;
  const uint32_t pp = active ? p : 0u;
  // byte offsets of this lane's value in a partial row and of its pattern's
  // exponent in an exponent row
  const uint32_t poff = p * 16u + s * 4u, eoff = p * 2u;
  const uint32_t ns = a.nslots;
  const size_t KP = (size_t)K * a.P;
  const uint32_t MATC = CH * 2 * K * 20;  // floats per matrix chunk
  const uint32_t CODEC = CH * 2 * TP;     // bytes per code chunk

  // ---- LDS carve-up (mirrors prune_f32_s1_lds_bytes) ----
  unsigned char *sp = smem;
  uint4 *ent_s = reinterpret_cast<uint4 *>(sp);
  sp += (size_t)n * 16;
  float *mat_s = reinterpret_cast<float *>(sp);  // [2][MATC]
  sp += 2 * (size_t)MATC * 4;
  unsigned char *code_s = sp;  // [2][CODEC]
  sp += 2 * (size_t)CODEC;
  sp = smem + ((sp - smem + 15) & ~(size_t)15);
  float *stk_all = reinterpret_cast<float *>(sp);  // [W][ns][64]
  int16_t *stke_all = reinterpret_cast<int16_t *>(stk_all + (size_t)W * ns * 64);
  double *mix = reinterpret_cast<double *>(sp);  // [W][16], after the node loop only
  double *red = mix + (size_t)W * 16;
  int *flag = reinterpret_cast<int *>(red + T);
  float *stk = stk_all + (size_t)w * ns * 64;
  int16_t *stke = stke_all + (size_t)w * ns * 64;

  const uint4 *g_ent = reinterpret_cast<const uint4 *>(a.entries);
  for (uint32_t i = tid; i < n; i += nthr) ent_s[i] = g_ent[i];
  __syncthreads();

  // chunk fetch into registers / store into an LDS buffer, as prune_tile
  float mreg[MAT_REGS];
  uint32_t creg[CODE_REGS];
  const uint32_t TPW = TP / 4;
  auto fetch = [&](uint32_t ck) {
    const uint32_t base = ck * CH;
    const uint32_t cnt = min(CH, n - base);
    const uint32_t nm = cnt * 2 * K * 20;
    const float *gm = a.pmats + (size_t)base * 2 * K * 20;
#pragma unroll
    for (uint32_t r = 0; r < MAT_REGS; r++) {
      const uint32_t q = tid + r * nthr;
      mreg[r] = q < nm ? gm[q] : 0.0f;
    }
    const uint32_t nc = cnt * 2 * TPW;
#pragma unroll
    for (uint32_t r = 0; r < CODE_REGS; r++) {
      const uint32_t q = tid + r * nthr;
      uint32_t v = 0x04040404u;
      if (q < nc) {
        const uint32_t jj = q / (2 * TPW);
        const uint32_t side = (q / TPW) & 1u;
        const uint32_t k = q % TPW;
        const uint4 e = ent_s[base + jj];
        const uint32_t src = side ? e.z : e.y;
        if ((src & SRC_TAG) == SRC_TIP)
          v = a.codes32[(size_t)(src & SRC_IDX) * (a.cstride / 4) + tile0 / 4 + k];
      }
      creg[r] = v;
    }
  };
  auto stash = [&](uint32_t buf, uint32_t ck) {
    const uint32_t base = ck * CH;
    const uint32_t cnt = min(CH, n - base);
    const uint32_t nm = cnt * 2 * K * 20;
    const uint32_t nc = cnt * 2 * TPW;
    float *ms = mat_s + (size_t)buf * MATC;
    uint32_t *cs = reinterpret_cast<uint32_t *>(code_s + (size_t)buf * CODEC);
#pragma unroll
    for (uint32_t r = 0; r < MAT_REGS; r++) {
      const uint32_t q = tid + r * nthr;
      if (q < nm) ms[q] = mreg[r];
    }
#pragma unroll
    for (uint32_t r = 0; r < CODE_REGS; r++) {
      const uint32_t q = tid + r * nthr;
      if (q < nc) cs[q] = creg[r];
    }
  };

  const uint32_t nchunks = (n + CH - 1) / CH;
  fetch(0u);
  stash(0u, 0u);
  __syncthreads();

  float x = 0.0f;
  int E = 0;
  for (uint32_t ck = 0; ck < nchunks; ck++) {
    const uint32_t buf = ck & 1u;
    const bool more = ck + 1 < nchunks;
    if (more) fetch(ck + 1);
    const uint32_t base = ck * CH;
    const uint32_t cnt = min(CH, n - base);
    const float *ms = mat_s + (size_t)buf * MATC;
    const unsigned char *cs = code_s + (size_t)buf * CODEC;
    for (uint32_t j = 0; j < cnt; j++) {
      const uint32_t i = base + j;
      const uint4 e4 = ((ConstU4 *)a.entries)[i];
      const uint32_t out = uni(e4.x), lsrc = uni(e4.y), rsrc = uni(e4.z), meta = uni(e4.w);
      // one child: a tip is entry `code` of row s (column `code`, component
      // s; the fifth column, all ones, for an unknown code); any other
      // child is row s dotted with the child's four states
      auto child = [&](uint32_t src, uint32_t side, float &v, int &ee) {
        const float *mc = ms + ((size_t)(j * 2 + side) * K + c) * 20;
        const uint32_t tag = src & SRC_TAG, idx = src & SRC_IDX;
        if (tag == SRC_TIP) {
          v = mc[(uint32_t)cs[(j * 2 + side) * TP + pt] * 4u + s];
          ee = 0;
          return;
        }
        const float m0 = mc[s], m1 = mc[4 + s], m2 = mc[8 + s], m3 = mc[12 + s];
        float y;
        if (tag == SRC_PREV) {
          y = x;
          ee = E;
        } else if (tag == SRC_LDS) {
          y = stk[idx * 64 + lane];
          ee = exp_unpack(stke[idx * 64 + lane]);
        } else {
          const size_t o = (size_t)(idx & SRC_SLOT) * KP + (size_t)c * a.P + pp;
          y = reinterpret_cast<const float *>(a.partials)[o * 4 + s];
          ee = exp_unpack(a.exps[o]);
        }
        const float y0 = quad_dpp<kQuadB0>(y), y1 = quad_dpp<kQuadB1>(y);
        const float y2 = quad_dpp<kQuadB2>(y), y3 = quad_dpp<kQuadB3>(y);
        v = fmaf(m3, y3, fmaf(m2, y2, fmaf(m1, y1, m0 * y0)));
      };
      float lv, rv;
      int le, re;
      child(lsrc, 0, lv, le);
      child(rsrc, 1, rv, re);
      float v = lv * rv;
      float mx = fmaxf(v, quad_dpp<kQuadSwap1>(v));
      mx = fmaxf(mx, quad_dpp<kQuadSwap2>(mx));
      const int ex = __builtin_amdgcn_frexp_expf(mx);
      v = __builtin_amdgcn_ldexpf(v, -ex);
      x = v;
      E = exp_clamp(le + re + ex);
      const int16_t ep = exp_pack(E);
      const uint32_t push = meta & 0xFFu;
      const bool store = !(a.cmats && (lsrc & SRC_TAG) == SRC_TIP &&
                           (rsrc & SRC_TAG) == SRC_TIP) || (meta & META_KEEP);
      if (active && store) {
        const size_t ob = (size_t)out * KP + (size_t)c * a.P;
        float *dp = reinterpret_cast<float *>(reinterpret_cast<char *>(a.partials + ob) + poff);
#if B3H_NT_STORE
        __builtin_nontemporal_store(v, dp);
        if (s == 0)
          __builtin_nontemporal_store(ep, reinterpret_cast<int16_t *>(
                                              reinterpret_cast<char *>(a.exps + ob) + eoff));
#else
        *dp = v;
        if (s == 0) *reinterpret_cast<int16_t *>(reinterpret_cast<char *>(a.exps + ob) + eoff) = ep;
#endif
      }
      if (push != NO_PUSH) {
        stk[push * 64 + lane] = v;
        stke[push * 64 + lane] = ep;
      }
    }
    if (more) stash(buf ^ 1u, ck + 1);
    __syncthreads();
  }

  // root, category mix and weighted sum as prune_tile, one pattern per quad
  double pi0 = a.pi0, pi1 = a.pi1, pi2 = a.pi2, pi3 = a.pi3;
  if (a.pik) {
    pi0 = a.pik[4 * c];
    pi1 = a.pik[4 * c + 1];
    pi2 = a.pik[4 * c + 2];
    pi3 = a.pik[4 * c + 3];
  }
  const float x0 = quad_dpp<kQuadB0>(x), x1 = quad_dpp<kQuadB1>(x);
  const float x2 = quad_dpp<kQuadB2>(x), x3 = quad_dpp<kQuadB3>(x);
  double site = -INFINITY;
  if (active) {
    double v = (double)x0 * pi0;
    v = fma((double)x1, pi1, v);
    v = fma((double)x2, pi2, v);
    v = fma((double)x3, pi3, v);
    site = exp_site(v, E);
    if (s == 0) a.site[(size_t)c * a.P + p] = site;
  }
  if (s == 0) mix[w * 16u + pl] = site;
  __syncthreads();
  double wv = 0.0;
  if (c == 0) {
    if (s == 0 && active) {
      double mixed;
      if (a.classes) {
        mixed = mix[(t * K + (uint32_t)a.classes[p]) * 16u + pl];
      } else {
        double mxv = -INFINITY;
        for (uint32_t k = 0; k < K; k++) mxv = fmax(mxv, mix[(t * K + k) * 16u + pl]);
        mixed = -INFINITY;
        if (mxv != -INFINITY) {
          double se = 0.0;
          for (uint32_t k = 0; k < K; k++) se = se + exp(mix[(t * K + k) * 16u + pl] - mxv);
          mixed = (mxv + log(se)) + a.ln_inv_k;
        }
      }
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
  last_block_total(bsum, a.blocksum, a.counter, a.total, a.host_total, a.host_flag, a.seq, flag,
                   tile, a.ntiles);
}

__global__ __launch_bounds__(1024) void prune_f32_s1_kernel(F32Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
#if defined(__HIP_DEVICE_COMPILE__)
  (void)a;
  typedef const __attribute__((address_space(4))) F32Args ConstArgs;
  ConstArgs &ka = *(ConstArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  prune_tile_s1(ka, blockIdx.x, smem);
#else
  prune_tile_s1(a, blockIdx.x, smem);
#endif
}

template <int PPL, bool HBM>
__global__ __launch_bounds__(1024) void prune_f32_kernel(F32Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
#if defined(__HIP_DEVICE_COMPILE__)
  // The record is read where the dispatch put it (offset 0 of the kernel
  // argument segment) through the constant address space.  Through the
  // by-value parameter the compiler keeps a private copy and merges the two
  // ways of reading the root frequencies into one flat load from it.
  (void)a;
  typedef const __attribute__((address_space(4))) F32Args ConstArgs;
  ConstArgs &ka = *(ConstArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  prune_tile<PPL, HBM>(ka, blockIdx.x, smem);
#else
  prune_tile<PPL, HBM>(a, blockIdx.x, smem);
#endif
}

// Persistent batched form: workgroups (at most 512 threads) take
// (instance, tile) items from a queue (ordered longest first by the host)
// until it is empty.  The queue
// counter is reset by the last workgroup to leave, so the kernel can be
// replayed (also from a captured graph) without a host-side reset.
template <int PPL, bool HBM>
__global__ __launch_bounds__(512) void prune_f32_batch_kernel(const F32Args *insts,
                                                               const uint2 *items,
                                                               uint32_t nitems,
                                                               uint32_t *queue) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t item_s;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(&queue[0], 1u);
    __syncthreads();
    const uint32_t item = item_s;
    __syncthreads();
    if (item >= nitems) break;
    const uint2 it = items[item];
    const uint32_t inst = __builtin_amdgcn_readfirstlane(it.x);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(it.y);
#if defined(__HIP_DEVICE_COMPILE__)
    // the record is read through the constant address space: uniform scalar
    // loads the compiler can repeat rather than hold in registers
    typedef const __attribute__((address_space(4))) F32Args ConstArgs;
    ConstArgs &a = ((ConstArgs *)insts)[inst];
    prune_tile<PPL, HBM>(a, tile, smem);
#else
    prune_tile<PPL, HBM>(insts[inst], tile, smem);
#endif
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
      __hip_atomic_store(&queue[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&queue[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// The persistent batched form with one state per lane (prune_tile_s1).
__global__ __launch_bounds__(512) void prune_f32_batch_s1_kernel(const F32Args *insts,
                                                                 const uint2 *items,
                                                                 uint32_t nitems,
                                                                 uint32_t *queue) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t item_s;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(&queue[0], 1u);
    __syncthreads();
    const uint32_t item = item_s;
    __syncthreads();
    if (item >= nitems) break;
    const uint2 it = items[item];
    const uint32_t inst = __builtin_amdgcn_readfirstlane(it.x);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(it.y);
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) F32Args ConstArgs;
    ConstArgs &a = ((ConstArgs *)insts)[inst];
    prune_tile_s1(a, tile, smem);
#else
    prune_tile_s1(insts[inst], tile, smem);
#endif
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
      __hip_atomic_store(&queue[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&queue[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Partial updates: update lists that form one chain -- entry 0 reads two
// stored children (leaves or partials of an earlier evaluation), every later
// entry reads the entry just before it and one stored child (the path from
// a changed edge to the root).  The host lists the stored children in the
// order they are read, so entry 0 reads sources 0 and 1 and entry i > 0
// reads source i + 1: which register holds a source is known when the
// kernel is compiled.  A wavefront issues the loads of kPathSeg sources at
// once, runs the entries that read them from registers (one wait for the
// whole segment instead of one memory latency per node), and moves on to
// the next segment; eight waves per SIMD hide the rest.  At 128 taxa and
// 100,000 patterns two sources per segment measured 36 us against 42 us for
// four and eight (at 200 taxa and 200,000 patterns 2-4 sources measured
// 58-59 us, 8 sources 70 us), and issuing the next segment's loads before
// computing the current one was slower at every length (42-47 us): more
// loads in flight per wave do not help here.  Entries and matrices are the same for every
// lane of a wavefront and come through the scalar cache; LDS holds only the
// category mix.  One lane per (pattern, category): wavefront (c, t) of a
// workgroup of K x T wavefronts is category c of the t-th 64 patterns of
// the workgroup's tile.  Arithmetic per node, root, mix and sum are those
// of prune_tile -- products and exponent sums commute, so which side
// carries the chain changes no bit -- and values match the other variants.
constexpr uint32_t kPathSeg = 2;

typedef const B3H_CONSTANT F32Args ConstF32Args;

__global__ __launch_bounds__(1024) void prune_path_kernel(F32Args arg, const uint32_t *srcs) {
  extern __shared__ __align__(16) unsigned char smem[];
#if defined(__HIP_DEVICE_COMPILE__)
  // the record as the dispatch placed it (offset 0 of the kernel arguments),
  // read through the constant address space (see prune_f32_kernel)
  (void)arg;
  ConstF32Args &a = *(ConstF32Args *)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const F32Args &a = arg;
#endif
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t w = uni(tid >> 6);
  const uint32_t K = a.K, n = a.n, T = blockDim.x / (64u * K);
  const uint32_t c = w % K, t = w / K;
  const uint32_t p = (blockIdx.x * T + t) * 64u + lane;
  const bool active = p < a.P;
  const uint32_t pp = active ? p : 0u;
  const size_t KP = (size_t)K * a.P;
  const uint8_t *codes8 = reinterpret_cast<const uint8_t *>(a.codes32);
  // read-only for the whole kernel: uniform reads go through the scalar cache
  ConstU4 *ents = (ConstU4 *)a.entries;
  ConstF4 *mats = (ConstF4 *)a.pmats;
  ConstU32 *tab = (ConstU32 *)srcs;

  // a stored child: a partial (value, exponent), a leaf's code (in `e`), or
  // a cherry recomputed from its two leaves' codes (value, exponent)
  auto load_src = [&](uint32_t sv, float4 &v, int &e) {
    const uint32_t idx = sv & SRC_IDX;
    if ((sv & SRC_TAG) == SRC_HBM && (sv & SRC_CHERRY)) {
      const uint2 lv = ((ConstU2 *)a.cleaves)[sv & SRC_SLOT];
      const uint32_t cc = (uint32_t)codes8[(size_t)lv.x * a.cstride + pp] |
                          ((uint32_t)codes8[(size_t)lv.y * a.cstride + pp] << 8);
      v = cherry_value(a.cmats, sv & SRC_SLOT, K, c, cc, e);
    } else if ((sv & SRC_TAG) == SRC_HBM) {
      const size_t o = (size_t)(idx & SRC_SLOT) * KP + (size_t)c * a.P + pp;
      v = a.partials[o];
      e = exp_unpack(a.exps[o]);
    } else {
      v = make_float4(0.f, 0.f, 0.f, 0.f);
      e = codes8[(size_t)idx * a.cstride + pp];
    }
  };
  auto project = [&](const Mat &M, uint32_t sv, const float4 &v, int e, float4 &out, int &oe) {
    if ((sv & SRC_TAG) == SRC_TIP) {
      out = tip(M, (uint32_t)e);
      oe = 0;
    } else {
      out = apply(M, v);
      oe = e;
    }
  };

  const uint32_t s0 = tab[0];
  float4 v0;
  int e0;
  load_src(s0, v0, e0);
  float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
  int E = 0;
  for (uint32_t i0 = 0; i0 < n; i0 += kPathSeg) {
    uint32_t sv[kPathSeg];
    float4 pv[kPathSeg];
    int ev[kPathSeg];
#pragma unroll
    for (uint32_t j = 0; j < kPathSeg; j++) {
      sv[j] = 0;
      pv[j] = make_float4(0.f, 0.f, 0.f, 0.f);
      ev[j] = 0;
      if (i0 + j < n) {
        sv[j] = tab[i0 + j + 1];
        load_src(sv[j], pv[j], ev[j]);
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kPathSeg; j++) {
      const uint32_t i = i0 + j;
      if (i < n) {
        const uint4 e4 = ents[i];
        // side of the chain (entry 0: its left child, source 0)
        const uint32_t cs = (i == 0 || (e4.y & SRC_TAG) == SRC_PREV) ? 0u : 1u;
        const Mat MC = load_mat_uniform(mats + ((size_t)(i * 2 + cs) * K + c) * 5);
        const Mat MS = load_mat_uniform(mats + ((size_t)(i * 2 + (cs ^ 1u)) * K + c) * 5);
        float4 cv, sv4;
        int ce, se;
        if (i == 0) project(MC, s0, v0, e0, cv, ce);
        else {
          cv = apply(MC, x);
          ce = E;
        }
        project(MS, sv[j], pv[j], ev[j], sv4, se);
        float4 v;
        v.x = cv.x * sv4.x;
        v.y = cv.y * sv4.y;
        v.z = cv.z * sv4.z;
        v.w = cv.w * sv4.w;
        const int ex = normalise(v);
        x = v;
        E = exp_clamp(ce + se + ex);
        // entry 0 may be a cherry: its matrices are recorded, its partial not
        const bool store = !(i == 0 && a.cmats && (s0 & SRC_TAG) == SRC_TIP &&
                             (sv[j] & SRC_TAG) == SRC_TIP) || (e4.w & META_KEEP);
        if (active && store) {
          const size_t o = (size_t)e4.x * KP + (size_t)c * a.P + p;
#if B3H_NT_STORE
          const f4v xv = {v.x, v.y, v.z, v.w};
          __builtin_nontemporal_store(xv, reinterpret_cast<f4v *>(&a.partials[o]));
          __builtin_nontemporal_store(exp_pack(E), &a.exps[o]);
#else
          a.partials[o] = v;
          a.exps[o] = exp_pack(E);
#endif
        }
      }
    }
  }

  // root (cpu4.rs:277-284), then the category mix and the weighted sum
  double pi0 = a.pi0, pi1 = a.pi1, pi2 = a.pi2, pi3 = a.pi3;
  if (a.pik) {
    pi0 = a.pik[4 * c];
    pi1 = a.pik[4 * c + 1];
    pi2 = a.pik[4 * c + 2];
    pi3 = a.pik[4 * c + 3];
  }
  double site = -INFINITY;
  if (active) {
    double v = (double)x.x * pi0;
    v = fma((double)x.y, pi1, v);
    v = fma((double)x.z, pi2, v);
    v = fma((double)x.w, pi3, v);
    site = exp_site(v, E);
    a.site[(size_t)c * a.P + p] = site;
  }
  // LDS: mix [T][K][64], then the sub-tile sums [T] and the ticket flag
  double *mix = reinterpret_cast<double *>(smem);
  double *red = mix + (size_t)T * K * 64u;
  int *flag = reinterpret_cast<int *>(red + T);
  double *mt = mix + (size_t)t * K * 64u;
  mt[c * 64u + lane] = site;
  __syncthreads();
  if (c == 0) {
    double wv = 0.0;
    if (active) {
      double mixed;
      if (a.classes) {
        mixed = mt[(uint32_t)a.classes[p] * 64u + lane];
      } else {
        double mxv = -INFINITY;
        for (uint32_t kk = 0; kk < K; kk++) mxv = fmax(mxv, mt[kk * 64u + lane]);
        mixed = -INFINITY;
        if (mxv != -INFINITY) {
          double se = 0.0;
          for (uint32_t kk = 0; kk < K; kk++) se = se + exp(mt[kk * 64u + lane] - mxv);
          mixed = (mxv + log(se)) + a.ln_inv_k;
        }
      }
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
  last_block_total(bsum, a.blocksum, a.counter, a.total, a.host_total, a.host_flag, a.seq, flag,
                   blockIdx.x, gridDim.x);
}

uint32_t path_tiles(uint32_t P, uint32_t K) {
  const uint32_t T = K >= 16 ? 1u : 16u / K;
  return (P + 64u * T - 1) / (64u * T);
}

hipError_t launch_prune_path(const F32Args &a, const uint32_t *srcs, hipStream_t s) {
  if (a.K == 0 || a.K > 16 || a.n == 0) return hipErrorInvalidValue;
  // workgroups of up to 1024 threads: T sub-tiles of 64 patterns per
  // category, so fewer tiles take a ticket
  const uint32_t T = a.K >= 16 ? 1u : 16u / a.K;
  if (a.ntiles != path_tiles(a.P, a.K)) return hipErrorInvalidValue;
  const dim3 g(a.ntiles), b(64u * a.K * T);
  const size_t lds = (size_t)T * a.K * 64u * 8u + (size_t)T * 8u + 16u;
  hipLaunchKernelGGL(prune_path_kernel, g, b, (uint32_t)lds, s, a, srcs);
  return hipGetLastError();
}

hipError_t prepare_prune_f32(int max_lds_bytes) {
  const void *fns[] = {reinterpret_cast<const void *>(&prune_f32_kernel<1, false>),
                       reinterpret_cast<const void *>(&prune_f32_kernel<1, true>),
                       reinterpret_cast<const void *>(&prune_f32_kernel<2, false>),
                       reinterpret_cast<const void *>(&prune_f32_s1_kernel),
                       reinterpret_cast<const void *>(&prune_path_kernel)};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       max_lds_bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t prepare_prune_f32_batch(int max_lds_bytes) {
  const void *fns[] = {reinterpret_cast<const void *>(&prune_f32_batch_kernel<1, false>),
                       reinterpret_cast<const void *>(&prune_f32_batch_kernel<1, true>),
                       reinterpret_cast<const void *>(&prune_f32_batch_kernel<2, false>),
                       reinterpret_cast<const void *>(&prune_f32_batch_s1_kernel)};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       max_lds_bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t batch_grid(uint32_t ppl, bool hbm, bool s1, uint32_t threads, size_t lds,
                      uint32_t *grid) {
  const void *f = s1 ? reinterpret_cast<const void *>(&prune_f32_batch_s1_kernel)
                  : (ppl == 2 && !hbm)
                      ? reinterpret_cast<const void *>(&prune_f32_batch_kernel<2, false>)
                      : hbm ? reinterpret_cast<const void *>(&prune_f32_batch_kernel<1, true>)
                            : reinterpret_cast<const void *>(&prune_f32_batch_kernel<1, false>);
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, (int)threads, lds);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  *grid = (uint32_t)((per_cu > 0 ? per_cu : 1) * cus);
  return hipSuccess;
}

hipError_t launch_prune_f32_batch(const F32Args *insts, const uint2 *items, uint32_t nitems,
                                  uint32_t *queue, uint32_t ppl, bool hbm, bool s1,
                                  uint32_t threads, size_t lds, uint32_t grid, hipStream_t s) {
  const dim3 g(grid), b(threads);
  const uint32_t l = (uint32_t)lds;
  if (s1)
    hipLaunchKernelGGL(prune_f32_batch_s1_kernel, g, b, l, s, insts, items, nitems, queue);
  else if (ppl == 2 && !hbm)
    hipLaunchKernelGGL((prune_f32_batch_kernel<2, false>), g, b, l, s, insts, items, nitems, queue);
  else if (hbm)
    hipLaunchKernelGGL((prune_f32_batch_kernel<1, true>), g, b, l, s, insts, items, nitems, queue);
  else
    hipLaunchKernelGGL((prune_f32_batch_kernel<1, false>), g, b, l, s, insts, items, nitems, queue);
  return hipGetLastError();
}

hipError_t launch_prune_f32_s1(const F32Args &a, uint32_t blocks, hipStream_t s) {
  const size_t lds = prune_f32_s1_lds_bytes(a.K, a.T, a.nslots, a.n);
  hipLaunchKernelGGL(prune_f32_s1_kernel, dim3(blocks), dim3(64u * a.K * a.T), (uint32_t)lds, s,
                     a);
  return hipGetLastError();
}

hipError_t launch_prune_f32(const F32Args &a, uint32_t ppl, bool hbm, uint32_t blocks,
                            hipStream_t s) {
  const size_t lds = prune_f32_lds_bytes(a.K, a.T, a.nslots, a.n, ppl);
  const dim3 g(blocks), b(64u * a.K * a.T);
  const uint32_t l = (uint32_t)lds;
  if (ppl == 2 && !hbm) {  // the HBM-prefetch variant is built with one pattern per lane
    hipLaunchKernelGGL((prune_f32_kernel<2, false>), g, b, l, s, a);
  } else {
    if (hbm) hipLaunchKernelGGL((prune_f32_kernel<1, true>), g, b, l, s, a);
    else hipLaunchKernelGGL((prune_f32_kernel<1, false>), g, b, l, s, a);
  }
  return hipGetLastError();
}

}  // namespace b3h
