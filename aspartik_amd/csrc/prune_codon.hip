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
// that read, like the tip codes, is issued a node ahead, straight into LDS.
// The four wavefronts of a workgroup share the current node's matrices in
// LDS; the next node's are loaded into registers a node ahead and written
// between the two barriers of a node.  Each wavefront runs its two
// children's product chains interleaved, one 32-row block at a time.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

// Buffer load of `size` bytes per lane straight into LDS (buffer_load ...
// lds): lane l's data lands at lds + l * size.
typedef int32_t i4 __attribute__((ext_vector_type(4)));
__device__ void buffer_load_lds(i4 rsrc, __attribute__((address_space(3))) uint32_t *lds,
                                int size, int voffset, int soffset, int offset,
                                int aux) __asm("llvm.amdgcn.raw.buffer.load.lds");

namespace b3h {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kMatH = 64 * 64;      // halves per LDS matrix
constexpr float kLoScale = 0x1p-20f;     // P = hi + kLoScale * lo
constexpr int kXTop = 15;                // stored partials: pattern max in [2^14, 2^15)
constexpr uint32_t kWaves = 4;           // tiles (wavefronts) per workgroup

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Workgroup barrier for the LDS matrix buffers: waits for this wavefront's
// LDS operations only.  (__syncthreads also waits for every outstanding
// global load, including the LDS-direct prefetches of later nodes.)
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// state held in register i by a lane of half h
__device__ __forceinline__ uint32_t row_of(uint32_t h, uint32_t i) {
  return 32u * (i / 16u) + 8u * ((i % 16u) / 4u) + 4u * h + (i % 4u);
}

struct Tile {
  h8 v[4];  // 32 halves in register order
};

// LDS matrices are row-major with the 16-byte chunks of row r stored at
// chunk ^ ((r >> 1) & 7).  gfx950's LDS serves a b128 read 16 lanes at a
// time across 64 banks (256 B), and two 128-byte rows share a bank row: the
// 16 lanes of a phase read rows 32 rb + 0..15, so rows 2q and 2q + 1 take
// the two halves of the bank row and q selects distinct chunks.  (The
// swizzle by r & 7 made rows r and r + 8 collide: half the LDS cycles were
// bank conflicts.)

// max / sum with the lane holding the same pattern in the other half
__device__ __forceinline__ float pair_max(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                  false);
  return fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
}
__device__ __forceinline__ double pair_sum(double v) {
  const uint64_t b = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)b, (uint32_t)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(b >> 32), (uint32_t)(b >> 32),
                                                   false, false);
  return __longlong_as_double((int64_t)(((uint64_t)hi[0] << 32) | lo[0])) +
         __longlong_as_double((int64_t)(((uint64_t)hi[1] << 32) | lo[1]));
}

}  // namespace

// All LDS is static (32.5 KiB: the matrices of both sides and the root
// frequencies); with at most 168 VGPRs three 4-wavefront workgroups fit a
// CU, so 20,000 patterns x 4 categories (628 workgroups) run in one round on
// the 256 CUs.
size_t prune_codon_lds_bytes(uint32_t) { return 0; }

namespace {

// Buffer resource over `bytes` bytes at a wave-uniform base: a resource of
// 0 bytes makes every lane's access out of range, which returns zeros and
// moves no data, so a prefetch of an absent source costs an instruction but
// no memory traffic, and every node issues the same loads.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, bytes, 0x00020000);
}

// The same resource as four dwords (base, stride 0 | base high, bytes,
// flags), the form buffer_load_lds takes.
__device__ __forceinline__ i4 wave_rsrc(const void *p, uint32_t bytes) {
  const uint64_t b = reinterpret_cast<uint64_t>(p);
  i4 r;
  r[0] = __builtin_amdgcn_readfirstlane((int32_t)(uint32_t)b);
  r[1] = __builtin_amdgcn_readfirstlane((int32_t)((uint32_t)(b >> 32) & 0xFFFFu));
  r[2] = (int32_t)bytes;
  r[3] = 0x00020000;
  return r;
}
__device__ __forceinline__ __attribute__((address_space(3))) uint32_t *lds_ptr(void *p) {
  return reinterpret_cast<__attribute__((address_space(3))) uint32_t *>(
      reinterpret_cast<uintptr_t>(p));
}

// B operand of a tip: the lane's 32 states (register order) one-hot at
// `code`.  State row(h, 8 s + m) = 16 s + 8 (m / 4) + 4 h + m % 4, so code k
// sits in the half of lane-half (k >> 2) & 1, at register position
// 8 (k >> 4) + 4 ((k >> 3) & 1) + (k & 3); codes >= 64 give zeros.
__device__ __forceinline__ Tile one_hot(uint32_t code, uint32_t h) {
  const uint32_t hit = (code < 64u && ((code >> 2) & 1u) == h)
                           ? 8u * (code >> 4) + 4u * ((code >> 3) & 1u) + (code & 3u)
                           : 64u;
  const uint32_t word = hit >> 1, val = 0x3C00u << (16u * (hit & 1u));  // FP16 1.0
  Tile b;
#pragma unroll
  for (uint32_t s = 0; s < 4; s++) {
    u4 w;
#pragma unroll
    for (uint32_t d = 0; d < 4; d++) w[d] = (4u * s + d == word) ? val : 0u;
    b.v[s] = __builtin_bit_cast(h8, w);
  }
  return b;
}

}  // namespace

// grid (Ppad / 128, K): blockIdx.y is the rate category; 4 wavefronts per
// workgroup, each owning one 32-pattern tile (the pattern count is padded to
// a multiple of 128, so every wavefront's tile exists and every access of
// the padded codes, partials and exponents is in bounds).
//
// Per node i the global memory operations are issued in a fixed order:
//   node start   7 buffer loads, the sources of entry i+1 into registers:
//                one stored child's tile (4) and exponent (1), and the tip
//                codes of both sides (2); a load whose source is absent has
//                an empty range
//   node end     (after the products and the normalisation) 8 loads of
//                entry i+1's matrices, the node's partial (4 stores) and
//                exponent (1 store), barrier 1, the matrices written to LDS,
//                barrier 2
// Entry i+1's sources thus load during node i's products, its matrices
// during node i's stores and first barrier.  A second stored child of one
// entry (two children read from HBM: the first entry of some partial
// updates) is loaded when the entry starts.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void
prune_codon_kernel(CodonArgs a, const uint4 *__restrict__ ents) {
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u, h = lane / 32u, col = lane % 32u;
  const uint32_t t = uni(tid >> 6);
  const uint32_t c = blockIdx.y;
  const uint32_t K = a.K, n = a.n, S = a.S;
  const uint32_t tile = blockIdx.x * kWaves + t;
  const uint32_t p = tile * 32u + col;
  const size_t tiles_all = a.Ppad / 32u;
  // [s][lane][8] tile of (slot, category) and the tile's 32 exponents, at
  // wave-uniform bases
  const size_t tile_stride = (size_t)K * tiles_all * 2048u, exp_stride = (size_t)K * a.Ppad;
  _Float16 *const tile0 = a.partials + ((size_t)c * tiles_all + tile) * 2048u;
  int16_t *const exp0 = a.exps + (size_t)c * a.Ppad + tile * 32u;
  const uint8_t *const code0 = a.codes + (size_t)tile * 32u;

  __shared__ double pi_s[64];
  __shared__ __align__(16) _Float16 amat0[2 * kMatH];  // side 0: hi, lo
  __shared__ __align__(16) _Float16 amat1[2 * kMatH];  // side 1
  if (tid < 64) pi_s[tid] = a.pi[tid];

  // Matrices of entry e, category c: 2 sides x 2 parts x 64 x 64 halves,
  // loaded straight into LDS (global_load_lds, no registers): load q of
  // wavefront t fills the 1 KiB of rows 8 t .. 8 t + 7 of part q / 2, row
  // block q % 2, one 16-byte unit per lane, so lane l writes row
  // r = 32 (q % 2) + 8 t + l / 8, chunk slot l % 8, which holds logical chunk
  // (l % 8) ^ ((r >> 1) & 7); the lane's source address is chosen to match.
  // (the lane's part of the source address is one 32-bit byte offset; the
  // rest, resource base and soffset, is wave-uniform)
  const uint32_t mrow = 8u * t + (lane >> 3);
  const uint32_t moff = mrow * 128u + 16u * ((lane & 7u) ^ ((mrow >> 1) & 7u));
  auto dma_mats = [&](uint32_t e) {
    const i4 rs = wave_rsrc(a.amats + ((size_t)(2 * e) * K + c) * 8192u, (K + 1u) * 16384u);
#pragma unroll
    for (uint32_t side = 0; side < 2; side++) {
      _Float16 *dst = side == 0 ? amat0 : amat1;
#pragma unroll
      for (uint32_t q = 0; q < 4; q++)
        buffer_load_lds(rs, lds_ptr(dst + (q >> 1) * kMatH + (32u * (q & 1u) + 8u * t) * 64u), 16,
                        (int)moff,
                        (int)(side * K * 16384u + (q >> 1) * 8192u + (q & 1u) * 4096u), 0, 0);
    }
  };

  // Sources of entry i in registers: the first stored child's tile and
  // exponent (yside: its side, 2 when there is none) and both sides' codes.
  Tile y;
  int16_t ey = 0;  // packed
  uint32_t yside = 2u;
  uint32_t code[2] = {0u, 0u};
  auto prefetch = [&](uint32_t i, bool live) {
    const uint4 e4 = live ? ents[i] : make_uint4(0u, SRC_PREV, SRC_PREV, 0u);
    const uint32_t src[2] = {e4.y, e4.z};
    const bool hbm0 = (src[0] & SRC_TAG) == SRC_HBM, hbm1 = (src[1] & SRC_TAG) == SRC_HBM;
    yside = hbm0 ? 0u : hbm1 ? 1u : 2u;
    const uint32_t slot = (hbm0 ? src[0] : src[1]) & SRC_IDX;
    const bool any = yside < 2u;
    const auto rt = rsrc(tile0 + (any ? slot : 0u) * tile_stride, any ? 4096u : 0u);
    const auto re = rsrc(exp0 + (any ? slot : 0u) * exp_stride, any ? 64u : 0u);
#pragma unroll
    for (uint32_t s = 0; s < 4; s++)
      y.v[s] = __builtin_bit_cast(
          h8, __builtin_amdgcn_raw_buffer_load_b128(rt, (int)((s * 64u + lane) * 16u), 0, 0));
    // (unpacked where it is used: unpacking here would make the compiler
    // wait for this load before the node's own work)
    ey = (int16_t)__builtin_amdgcn_raw_buffer_load_b16(re, (int)(col * 2u), 0, 0);
#pragma unroll
    for (uint32_t side = 0; side < 2; side++) {
      const bool tip = (src[side] & SRC_TAG) == SRC_TIP;
      const auto rc = rsrc(code0 + (size_t)(tip ? (src[side] & SRC_IDX) : 0u) * a.cstride,
                           tip ? 32u : 0u);
      code[side] = __builtin_amdgcn_raw_buffer_load_b8(rc, (int)col, 0, 0);
    }
  };

  dma_mats(0);
  prefetch(0, true);
  __syncthreads();  // frequencies, matrices of entry 0 (waits for every load)

  Tile x;  // the previous entry's result (FP16, stored layout)
  int ex_prev = 0;
  // entry i; the root (the last entry) is a separate instance: it has no
  // next entry, and its path does not join the others'
  auto node = [&](uint32_t i, auto root_tag) {
    constexpr bool more = !decltype(root_tag)::value;
    const uint4 e4 = ents[i];
    const uint32_t out = e4.x, srcs[2] = {e4.y, e4.z};

    // B operands and exponents of both children: a tip's one-hot column,
    // the previous result, or a stored child (prefetched, or loaded here
    // when it is the entry's second stored child)
    Tile b[2];
    int ce[2];
    // this entry's codes: the prefetch below overwrites code[]
    const uint32_t cd[2] = {code[0], code[1]};
#pragma unroll
    for (uint32_t side = 0; side < 2; side++) {
      const uint32_t tag = srcs[side] & SRC_TAG;
      if (tag == SRC_TIP) {
        b[side] = one_hot(cd[side], h);
        ce[side] = 0;
      } else if (tag == SRC_PREV) {
        b[side] = x;
        ce[side] = ex_prev;
      } else if (side == yside) {
        b[side] = y;
        ce[side] = exp_unpack(ey);
      } else {
        const uint32_t slot = srcs[side] & SRC_IDX;
        const h8 *g = reinterpret_cast<const h8 *>(tile0 + slot * tile_stride) + lane;
#pragma unroll
        for (uint32_t s = 0; s < 4; s++) b[side].v[s] = g[64u * s];
        ce[side] = exp_unpack(exp0[slot * exp_stride + col]);
        // complete here (vmcnt(0)), so that the common path after the join
        // does not wait for the previous entry's stores
        __builtin_amdgcn_s_waitcnt(0x0F70);
      }
    }
    if constexpr (more) prefetch(i + 1, true);

    // per row block, the two children's chains interleaved step by step:
    // lo parts, the 2^-20 scale, then hi parts accumulated on top; unknown
    // tips take the all-ones column (row_of(h, i) = row_of(0, i) + 4 h)
    f16v acc[2];
    const uint32_t sw = (col >> 1) & 7u;
    const int lim = (int)S - 4 * (int)h;
#pragma unroll
    for (uint32_t rb = 0; rb < 2; rb++) {
      f16v r[2] = {};
#pragma unroll
      for (uint32_t part = 0; part < 2; part++) {
        const uint32_t po = (part == 0 ? kMatH : 0u) + (32 * rb + col) * 64;  // lo first
#pragma unroll
        for (uint32_t s = 0; s < 4; s++)
#pragma unroll
          for (uint32_t side = 0; side < 2; side++) {
            const _Float16 *A = (side == 0 ? amat0 : amat1) + po;
            r[side] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                *reinterpret_cast<const h8 *>(A + 8 * ((2 * s + h) ^ sw)), b[side].v[s], r[side],
                0, 0, 0);
          }
        if (part == 0) {
          r[0] *= kLoScale;
          r[1] *= kLoScale;
        }
      }
      // issue order of the row block: four A-fragment reads ahead, then one
      // read per MFMA, so that no MFMA waits out a whole LDS round trip
      __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
      for (uint32_t k = 0; k < 12; k++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
#pragma unroll
      for (uint32_t side = 0; side < 2; side++) {
        if ((srcs[side] & SRC_TAG) == SRC_TIP && __builtin_amdgcn_ballot_w64(cd[side] >= S)) {
          const bool unk = cd[side] >= S;
          // opaque here, so that the 16 selects are made in this (rare)
          // branch instead of being hoisted out of the node loop as 32
          // registers of constants held for the whole kernel
          int l = lim;
          asm volatile("" : "+v"(l));
#pragma unroll
          for (uint32_t q = 0; q < 16; q++)
            if (unk) r[side][q] = (int)row_of(0, 16 * rb + q) < l ? 1.0f : 0.0f;
        }
      }
      acc[rb] = r[0] * r[1];
      // the product now, so that the first row block's two accumulators are
      // not held through the second's
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (more) {
      lds_barrier();    // every wavefront is done with this node's matrices
      dma_mats(i + 1);  // the next entry's, in flight during the normalisation
      __builtin_amdgcn_sched_barrier(0);
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
    mx = pair_max(mx);
    const int ex = mx > 0.0f ? __builtin_amdgcn_frexp_expf(mx) : 0;
#pragma unroll
    for (uint32_t r = 0; r < 32; r++) v[r] = __builtin_amdgcn_ldexpf(v[r], kXTop - ex);
    const int E = exp_clamp(ce[0] + ce[1] + ex - kXTop);

    if constexpr (more) {
#pragma unroll
      for (uint32_t s = 0; s < 4; s++)
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) x.v[s][q] = (_Float16)v[8 * s + q];
      ex_prev = E;
      // streaming stores, which keep the L2 for the matrices every
      // workgroup reads (both halves of a pattern store the same exponent)
      // (buffer stores: a wave-uniform base and a 32-bit lane offset; aux 2
      // is the nontemporal hint)
      const auto rt = rsrc(tile0 + out * tile_stride, 4096u);
#pragma unroll
      for (uint32_t s = 0; s < 4; s++)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, x.v[s]), rt,
                                               (int)((s * 64u + lane) * 16u), 0, 2);
      __builtin_amdgcn_raw_buffer_store_b16((uint16_t)exp_pack(E),
                                            rsrc(exp0 + out * exp_stride, 64u), (int)(col * 2u),
                                            0, 2);
      // this wavefront's matrix loads are complete once at most the five
      // stores after them are outstanding (vector memory operations retire
      // in issue order); then the next entry's matrices are visible to
      // every wavefront
      // (s_waitcnt vmcnt(5) expcnt(7) lgkmcnt(0) as a builtin, which the
      // compiler's own wait insertion takes into account)
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_waitcnt(0x0075);
      __builtin_amdgcn_s_barrier();
    } else {
      // root of the update: ln(sum_j pi_j x_j) + E ln 2 in FP64 from the
      // unrounded FP32 values (the root partial is never read as a child)
      double sum = 0.0;
      const double *pl = pi_s + 4 * h;
#pragma unroll
      for (uint32_t r = 0; r < 32; r++) sum = fma((double)v[r], pl[row_of(0, r)], sum);
      sum = pair_sum(sum);
      if (p < a.P && h == 0) a.site[(size_t)c * a.P + p] = exp_site(sum, E);
    }
  };
  for (uint32_t i = 0; i + 1 < n; i++) node(i, std::false_type{});
  node(n - 1, std::true_type{});
}

// the kernel uses static LDS only: nothing to raise
hipError_t prepare_prune_codon(int) { return hipSuccess; }

hipError_t launch_prune_codon(const CodonArgs &a, hipStream_t s) {
  if (a.n == 0 || a.Ppad % (32u * kWaves) || a.S > 64) return hipErrorInvalidValue;
  dim3 grid(a.Ppad / (32u * kWaves), a.K);
  hipLaunchKernelGGL(prune_codon_kernel, grid, dim3(256), 0, s, a,
                     reinterpret_cast<const uint4 *>(a.entries));
  return hipGetLastError();
}

// ---- transition matrices in A layout ------------------------------------
// Every P(t) of a propose shares the eigenvectors: with
//   U[q][el] = V[j][q] V^-1[q][k]   (el = 64 j + kslot, k = row(h, 8 s + m)
//                                    for kslot = 16 s + 8 h + m; zero outside
//                                    the S states)
// built once per model, the matrices of a propose are one FP64 product
//   P[mat][el] = sum_q E[mat][q] U[q][el],  E[mat][q] = exp(lambda_q t_mat),
// split into FP16 hi and lo parts (P = hi + 2^-20 lo) in the kernel's A
// layout: A[part][j][16 s + 8 h + m] = part(P)[j][row(h, 8 s + m)].

// one workgroup per q (64), 256 threads x 16 elements
__global__ __launch_bounds__(256) void ubuild_codon_kernel(const DevModel *model, double *U) {
  const uint32_t q = blockIdx.x, S = (uint32_t)model->n;
  for (uint32_t el = threadIdx.x; el < 4096u; el += blockDim.x) {
    const uint32_t j = el / 64u, kslot = el % 64u;
    const uint32_t k = row_of((kslot / 8u) & 1u, 8u * (kslot / 16u) + kslot % 8u);
    U[(size_t)q * 4096u + el] =
        (j < S && k < S && q < S) ? model->evec[j * S + q] * model->ievc[q * S + k] : 0.0;
  }
}

hipError_t launch_ubuild_codon(const DevModel *model, double *U, hipStream_t s) {
  hipLaunchKernelGGL(ubuild_codon_kernel, dim3(64), dim3(256), 0, s, model, U);
  return hipGetLastError();
}

// grid (4096 / 256 element tiles, ceil(matrices / 16)): each thread owns one
// element of 16 matrices; U is read once per workgroup, E from LDS.
__global__ __launch_bounds__(256) void pbuild_codon_kernel(const Entry *entries, uint32_t n,
                                                           uint32_t K, const double *times,
                                                           const double *rates,
                                                           const DevModel *model, const double *U,
                                                           _Float16 *out) {
  constexpr uint32_t kM = 16;
  __shared__ double E[kM][64];
  const uint32_t nmat = n * 2u * K, m0 = blockIdx.y * kM;
  const uint32_t S = (uint32_t)model->n;
  for (uint32_t w = threadIdx.x; w < kM * 64u; w += blockDim.x) {
    const uint32_t mi = w / 64u, q = w % 64u, mat = m0 + mi;
    double v = 0.0;
    if (mat < nmat && q < S) {
      const uint32_t c = mat % K, es = mat / K, side = es & 1u, e = es >> 1;
      const uint32_t orig = entry_index(entries[e].meta);
      v = exp(model->eval[q] * (times[2 * orig + side] * rates[c]));
    }
    E[mi][q] = v;
  }
  __syncthreads();
  const uint32_t el = blockIdx.x * 256u + threadIdx.x;
  double acc[kM];
#pragma unroll
  for (uint32_t mi = 0; mi < kM; mi++) acc[mi] = 0.0;
  for (uint32_t q = 0; q < S; q++) {
    const double u = U[(size_t)q * 4096u + el];
#pragma unroll
    for (uint32_t mi = 0; mi < kM; mi++) acc[mi] = fma(E[mi][q], u, acc[mi]);
  }
#pragma unroll
  for (uint32_t mi = 0; mi < kM; mi++) {
    const uint32_t mat = m0 + mi;
    if (mat >= nmat) break;
    const double v = acc[mi];
    const _Float16 hi = (_Float16)(float)v;
    out[(size_t)mat * 8192u + el] = hi;
    out[(size_t)mat * 8192u + 4096u + el] = (_Float16)(float)((v - (double)hi) * 0x1p20);
  }
}

hipError_t launch_pbuild_codon(const Entry *entries, uint32_t n, uint32_t K, const double *times,
                               const double *rates, const DevModel *model, const double *U,
                               _Float16 *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t nmat = n * 2u * K;
  hipLaunchKernelGGL(pbuild_codon_kernel, dim3(16, (nmat + 15) / 16), dim3(256), 0, s, entries, n,
                     K, times, rates, model, U, out);
  return hipGetLastError();
}

}  // namespace b3h
