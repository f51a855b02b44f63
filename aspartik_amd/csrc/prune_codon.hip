// S-state pruning (S <= 64, e.g. 61 codons) on MFMA: FP16 partials and
// transition matrices, FP32 accumulation, exact power-of-two rescaling.
//
// Per node and category the update is a dense contraction (b3's recursion,
// crates/b3/src/calculator/cpu4.rs:361-401, for S states):
//   X_parent[j][p] = (sum_k PL[j][k] XL[k][p]) * (sum_k PR[j][k] XR[k][p])
// computed per 32-pattern tile with v_mfma_f32_32x32x16_f16 (two 32-row
// blocks x four 16-deep steps per child and matrix part).
//
// Precision.  P is held as two FP16 matrices, P = hi + 2^-20 lo: hi is P
// rounded to FP16 and lo the rounded residual scaled into range, so P keeps
// about 22 significant bits and entries down to ~1e-12 (multi-substitution
// probabilities on short branches, which FP16 alone flushes to zero).  The
// lo product is accumulated first, scaled, and the hi product accumulated on
// top.  Stored partials are FP16 with the per-pattern maximum scaled into
// [2^14, 2^15) so that entries far below the maximum keep normal precision;
// the scale is a per-(node, category, pattern) int16 power of two.  The
// root sum, logarithm and exponent use FP64.
//
// Layout: a partial tile is stored as the MFMA accumulators hold it.  For a
// 32-pattern tile, lane l (pattern l % 32, half h = l / 32) owns 32 states,
// register i holding state
//   row(h, i) = 32 (i / 16) + 8 ((i % 16) / 4) + 4 h + (i % 4).
// The contraction index of the B operand is ordered to match: lane (p, h) at
// step s supplies states row(h, 8s .. 8s+7), i.e. its own registers 8s..8s+7,
// and the matrices' columns are permuted the same way,
//   A[j][16 s + 8 h + m] = P[j][row(h, 8 s + m)].
// So a child's stored (or still in-register) values are directly the next
// product's B operand.  In HBM a tile is [step s][lane][8 halves] (4 KiB), so
// each of its four 16-byte accesses per lane is one contiguous 1 KiB.
// Tips use a one-hot B operand (the product is column P[j][code] exactly); an
// unknown code gives 1 for every real state, as b3's fifth column does
// (cpu4.rs:53-61).  Padded states (>= S) have zero rows and columns in P and
// stay zero.
//
// Schedule: one wavefront per (32-pattern tile, category) walks the whole
// update list in the depth-first order of the host scheduler.  The child
// evaluated just before its parent stays in registers; every other internal
// child is read back from HBM (each node's partial is stored anyway), and
// that read, like the tip codes, is issued ahead of its node.  The four
// wavefronts of a workgroup share the current node's matrices in LDS; each
// side's buffer is refilled with the next node's matrices (loaded into
// registers half a node earlier) once all four have used it.
#include "common.hpp"
#include "kernels.hpp"

namespace b3h {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kRowH = 72;           // halves per LDS matrix row (64 + 8 pad)
constexpr uint32_t kMatH = 64 * kRowH;   // halves per LDS matrix
constexpr float kLoScale = 0x1p-20f;     // P = hi + kLoScale * lo
constexpr int kXTop = 15;                // stored partials: pattern max in [2^14, 2^15)
constexpr uint32_t kWaves = 4;           // tiles (wavefronts) per workgroup

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// state held in register i by a lane of half h
__device__ __forceinline__ uint32_t row_of(uint32_t h, uint32_t i) {
  return 32u * (i / 16u) + 8u * ((i % 16u) / 4u) + 4u * h + (i % 4u);
}

struct Tile {
  h8 v[4];  // 32 halves in register order
};

// row_of(h, i) = row_of(0, i) + 4 h: lane-dependent tests are written against
// the compile-time row_of(0, i) so that no per-lane table stays live.
//
// B operand of step s for a tip: one-hot over the lane's states of that
// step; d = code - 4 h (as an int, so codes below 4 h match nothing).
__device__ __forceinline__ h8 one_hot_step(int d, uint32_t s) {
  h8 b;
#pragma unroll
  for (uint32_t m = 0; m < 8; m++)
    b[m] = (int)row_of(0, 8 * s + m) == d ? (_Float16)1.0f : (_Float16)0.0f;
  return b;
}

// One 32-row block of a child's product, (P x)[32 rb + ..] for the lane's
// pattern, from the child's two matrix parts in LDS (hi at A, lo at
// A + kMatH); bstep(s) gives the B operand of step s.
template <typename BStep>
__device__ __forceinline__ f16v row_block(const _Float16 *A, uint32_t rb, uint32_t col, uint32_t h,
                                          BStep bstep) {
  const _Float16 *hi = A + (32 * rb + col) * kRowH + 8 * h;
  const _Float16 *lo = hi + kMatH;
  f16v acc = {};
#pragma unroll
  for (uint32_t s = 0; s < 4; s++)
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(*reinterpret_cast<const h8 *>(lo + 16 * s),
                                                 bstep(s), acc, 0, 0, 0);
  acc *= kLoScale;
#pragma unroll
  for (uint32_t s = 0; s < 4; s++)
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(*reinterpret_cast<const h8 *>(hi + 16 * s),
                                                 bstep(s), acc, 0, 0, 0);
  return acc;
}

// Sources of one entry as the node loop consumes them: tip codes, and the
// first child read from HBM (tile + exponent), loaded ahead of their node.
struct Pending {
  uint32_t code[2];
  Tile x;
  int e;
};

}  // namespace

// LDS: entries | stationary frequencies (64 doubles) | matrices [side][part]
size_t prune_codon_lds_bytes(uint32_t n) {
  return ((size_t)n * 16 + 15) / 16 * 16 + 64 * 8 + 4 * (size_t)kMatH * 2;
}

// grid (Ppad / 128, K): blockIdx.y is the rate category; 4 wavefronts per
// workgroup, each owning one 32-pattern tile (the pattern count is padded to
// a multiple of 128, so every wavefront's tile exists and every access of
// the padded codes, partials and exponents is in bounds).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void prune_codon_kernel(CodonArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u, h = lane / 32u, col = lane % 32u;
  const uint32_t t = uni(tid >> 6);
  const uint32_t c = blockIdx.y;
  const uint32_t K = a.K, n = a.n, S = a.S;
  const uint32_t tile = blockIdx.x * kWaves + t;
  const uint32_t p = tile * 32u + col;
  const size_t tiles_all = a.Ppad / 32u;
  // [s][lane][8] tile of (slot, category), and the lane's exponent
  auto tile_ptr = [&](uint32_t slot) -> _Float16 * {
    return a.partials + (((size_t)slot * K + c) * tiles_all + tile) * 2048u + (size_t)lane * 8u;
  };
  auto exp_idx = [&](uint32_t slot) -> size_t { return ((size_t)slot * K + c) * a.Ppad + p; };

  uint4 *ent_s = reinterpret_cast<uint4 *>(smem);
  double *pi_s = reinterpret_cast<double *>(smem + ((size_t)n * 16 + 15) / 16 * 16);
  _Float16 *amat = reinterpret_cast<_Float16 *>(pi_s + 64);

  const uint4 *g_ent = reinterpret_cast<const uint4 *>(a.entries);
  for (uint32_t i = tid; i < n; i += blockDim.x) ent_s[i] = g_ent[i];
  if (tid < 64) pi_s[tid] = a.pi[tid];

  // matrices of entry e, side, category c: 2 parts x 64 x 64 halves = 1024
  // units of 16 B, four per thread of the 256-thread workgroup.  The two
  // sides are staged separately: one side's LDS buffer is refilled with the
  // next node's matrices as soon as every wavefront has used it.
  u4 m0, m1, m2, m3;
  auto fetch_side = [&](uint32_t e, uint32_t side) {
    const u4 *src =
        reinterpret_cast<const u4 *>(a.amats + ((size_t)(2 * e + side) * K + c) * 8192u) + tid;
    m0 = src[0];
    m1 = src[256];
    m2 = src[512];
    m3 = src[768];
  };
  // unit w = tid + 256 q: part w / 512, row (w / 8) % 64, 16-byte chunk w % 8
  const uint32_t srow = (tid >> 3) & 31u, schunk = tid & 7u;
  auto store_side = [&](uint32_t side) {
    _Float16 *d = amat + side * 2u * kMatH + srow * kRowH + schunk * 8u;
    *reinterpret_cast<u4 *>(d) = m0;
    *reinterpret_cast<u4 *>(d + 32u * kRowH) = m1;
    *reinterpret_cast<u4 *>(d + kMatH) = m2;
    *reinterpret_cast<u4 *>(d + kMatH + 32u * kRowH) = m3;
  };
  auto load_tile = [&](uint32_t slot, Tile &x) {
    const h8 *g = reinterpret_cast<const h8 *>(tile_ptr(slot));
#pragma unroll
    for (uint32_t s = 0; s < 4; s++) x.v[s] = g[s * 64u];
  };
  // issue the loads of entry i's tip codes and first HBM child
  Pending pend;
  auto prefetch = [&](uint32_t i) {
    const uint4 e4 = ent_s[i];
    const uint32_t srcs[2] = {uni(e4.y), uni(e4.z)};
    bool got = false;
#pragma unroll
    for (uint32_t side = 0; side < 2; side++) {
      const uint32_t tag = srcs[side] & SRC_TAG, idx = srcs[side] & SRC_IDX;
      pend.code[side] = tag == SRC_TIP ? a.codes[(size_t)idx * a.cstride + p] : 0u;
      if (tag == SRC_HBM && !got) {
        load_tile(idx, pend.x);
        pend.e = a.exps[exp_idx(idx)];
        got = true;
      }
    }
  };
  auto has_hbm = [&](uint32_t i) {
    const uint4 e4 = ent_s[i];
    return ((uni(e4.y) & SRC_TAG) == SRC_HBM) || ((uni(e4.z) & SRC_TAG) == SRC_HBM);
  };

  __syncthreads();  // entries
  fetch_side(0, 0);
  store_side(0);
  fetch_side(0, 1);
  store_side(1);
  if (n > 1) fetch_side(1, 0);
  prefetch(0);
  __syncthreads();  // matrices of entry 0

  Tile x;
#pragma unroll
  for (uint32_t s = 0; s < 4; s++) x.v[s] = h8{};
  int E = 0;
  for (uint32_t i = 0; i < n; i++) {
    const uint4 e4 = ent_s[i];
    const uint32_t out = uni(e4.x), srcs[2] = {uni(e4.y), uni(e4.z)};
    const bool more = i + 1 < n;
    const uint32_t code[2] = {pend.code[0], pend.code[1]};
    // without an HBM child here the next entry's loads can go out at once;
    // otherwise right after the HBM child's product has consumed them
    bool issued = !has_hbm(i);
    if (issued && more) prefetch(i + 1);

    // side 0's product into acc, side 1's multiplied into it block by block
    f16v acc[2];
    int ce[2];
#pragma unroll
    for (uint32_t side = 0; side < 2; side++) {
      const uint32_t tag = srcs[side] & SRC_TAG, idx = srcs[side] & SRC_IDX;
      const _Float16 *A = amat + side * 2u * kMatH;
      // one row block at a time (the barrier keeps the scheduler from
      // interleaving the two chains, which would double the live registers)
      auto combine = [&](uint32_t rb, const f16v &r) {
        if (side == 0) acc[rb] = r;
        else acc[rb] *= r;
        __builtin_amdgcn_sched_barrier(0);
      };
      if (tag == SRC_TIP) {
        const uint32_t cd = code[side];
        const int d = (int)cd - 4 * (int)h, lim = (int)S - 4 * (int)h;
#pragma unroll
        for (uint32_t rb = 0; rb < 2; rb++) {
          f16v r = row_block(A, rb, col, h, [&](uint32_t st) { return one_hot_step(d, st); });
          if (cd >= S) {  // unknown: the all-ones column
#pragma unroll
            for (uint32_t q = 0; q < 16; q++) r[q] = (int)row_of(0, 16 * rb + q) < lim ? 1.0f : 0.0f;
          }
          combine(rb, r);
        }
        ce[side] = 0;
      } else if (tag == SRC_PREV) {
#pragma unroll
        for (uint32_t rb = 0; rb < 2; rb++)
          combine(rb, row_block(A, rb, col, h, [&](uint32_t st) { return x.v[st]; }));
        ce[side] = E;
      } else if (!issued) {
#pragma unroll
        for (uint32_t rb = 0; rb < 2; rb++)
          combine(rb, row_block(A, rb, col, h, [&](uint32_t st) { return pend.x.v[st]; }));
        ce[side] = pend.e;
        issued = true;
        if (more) prefetch(i + 1);
      } else {  // a second child from HBM (partial updates only)
        Tile y;
        load_tile(idx, y);
#pragma unroll
        for (uint32_t rb = 0; rb < 2; rb++)
          combine(rb, row_block(A, rb, col, h, [&](uint32_t st) { return y.v[st]; }));
        ce[side] = a.exps[exp_idx(idx)];
      }
      if (more) {
        __syncthreads();  // every wavefront is done with this side's matrices
        store_side(side);  // the next entry's
        if (side == 0) fetch_side(i + 1, 1);
        else if (i + 2 < n) fetch_side(i + 2, 0);
      }
    }
    // product and per-pattern normalisation across the lane pair
    float v[32];
    float mx = 0.0f;
#pragma unroll
    for (uint32_t r = 0; r < 16; r++) {
      v[r] = acc[0][r];
      v[16 + r] = acc[1][r];
      mx = fmaxf(mx, fmaxf(v[r], v[16 + r]));
    }
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const int ex = mx > 0.0f ? __builtin_amdgcn_frexp_expf(mx) : 0;
#pragma unroll
    for (uint32_t r = 0; r < 32; r++) v[r] = __builtin_amdgcn_ldexpf(v[r], kXTop - ex);
    E = ce[0] + ce[1] + ex - kXTop;

    if (more) {
#pragma unroll
      for (uint32_t s = 0; s < 4; s++)
#pragma unroll
        for (uint32_t m = 0; m < 8; m++) x.v[s][m] = (_Float16)v[8 * s + m];
      h8 *g = reinterpret_cast<h8 *>(tile_ptr(out));
#pragma unroll
      for (uint32_t s = 0; s < 4; s++) g[s * 64u] = x.v[s];
      if (h == 0) a.exps[exp_idx(out)] = (int16_t)E;
    } else {
      // root of the update: ln(sum_j pi_j x_j) + E ln 2 in FP64 from the
      // unrounded FP32 values (the root partial is never read as a child)
      double sum = 0.0;
      const double *pl = pi_s + 4 * h;
#pragma unroll
      for (uint32_t r = 0; r < 32; r++) sum = fma((double)v[r], pl[row_of(0, r)], sum);
      sum += __shfl_xor(sum, 32);
      if (p < a.P && h == 0) a.site[(size_t)c * a.P + p] = fma((double)E, LN2, log(sum));
    }
  }
}

hipError_t prepare_prune_codon(int max_lds_bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&prune_codon_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes);
}

hipError_t launch_prune_codon(const CodonArgs &a, hipStream_t s) {
  if (a.n == 0 || a.Ppad % (32u * kWaves) || a.S > 64) return hipErrorInvalidValue;
  const size_t lds = prune_codon_lds_bytes(a.n);
  dim3 grid(a.Ppad / (32u * kWaves), a.K);
  hipLaunchKernelGGL(prune_codon_kernel, grid, dim3(256), (uint32_t)lds, s, a);
  return hipGetLastError();
}

// ---- transition matrices in A layout ------------------------------------
// One workgroup per (entry, side, category): P = V diag(exp(lambda t)) V^-1 in
// FP64 from the device eigensystem, split into FP16 hi and lo parts
// (P = hi + 2^-20 lo) and stored with the contraction (column) index in the
// kernel's order:
//   A[part][j][16 s + 8 h + m] = part(P)[j][row(h, 8 s + m)],
// zero outside the S states.
__global__ __launch_bounds__(256) void pbuild_codon_kernel(const Entry *entries, uint32_t n,
                                                           uint32_t K, const double *times,
                                                           const double *rates,
                                                           const DevModel *model,
                                                           _Float16 *out) {
  __shared__ double ev[64];
  const uint32_t mat = blockIdx.x;  // (entry * 2 + side) * K + c
  const uint32_t c = mat % K, es = mat / K, side = es & 1u, e = es >> 1;
  const uint32_t S = (uint32_t)model->n;
  const uint32_t orig = entries[e].meta >> 8;
  const double t = times[2 * orig + side] * rates[c];
  if (threadIdx.x < 64) ev[threadIdx.x] = threadIdx.x < S ? exp(model->eval[threadIdx.x] * t) : 0.0;
  __syncthreads();
  for (uint32_t el = threadIdx.x; el < 4096u; el += blockDim.x) {
    const uint32_t j = el / 64u, kslot = el % 64u;
    const uint32_t s = kslot / 16u, hh = (kslot / 8u) & 1u, m = kslot % 8u;
    const uint32_t k = row_of(hh, 8 * s + m);
    double v = 0.0;
    if (j < S && k < S) {
      for (uint32_t q = 0; q < S; q++)
        v = fma(model->evec[j * S + q] * ev[q], model->ievc[q * S + k], v);
    }
    const _Float16 hi = (_Float16)(float)v;
    const _Float16 lo = (_Float16)(float)((v - (double)hi) * 0x1p20);
    out[(size_t)mat * 8192u + el] = hi;
    out[(size_t)mat * 8192u + 4096u + el] = lo;
  }
}

hipError_t launch_pbuild_codon(const Entry *entries, uint32_t n, uint32_t K, const double *times,
                               const double *rates, const DevModel *model, _Float16 *out,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(pbuild_codon_kernel, dim3(n * 2 * K), dim3(256), 0, s, entries, n, K, times,
                     rates, model, out);
  return hipGetLastError();
}

}  // namespace b3h
