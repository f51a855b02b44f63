// Host-callable launchers of the b3hip device kernels.
#pragma once

#include "common.hpp"

namespace b3h {

// ---- model.hip ----------------------------------------------------------
hipError_t launch_eigen(DevModel *m, hipStream_t s);
// Transition matrices of one propose, [entry][side][category][S*S]:
//   mode 0: cast of host matrices mats[K][2n][S*S]
//   mode 1: P(times[m] * rates[c]) from the device eigensystem
//   mode 2: P(times[c * 2n + m]), branch lengths given per category
// model_stride 0: one model for every category; 1: model[c] per category.
// copy_units > 0: the launch also copies copy_units 16-byte units from
// copy_src (pinned host memory) to copy_dst, the upload of a staging buffer
// whose inputs the build then reads in place.
// cmats / cleaves (4 states, may be null): for every entry whose children
// are both leaves, its matrices are also written to cmats[out slot]
// ([2][K][5][4]: columns, the fifth all ones) and its two leaves to
// cleaves[out slot].
template <typename T>
hipError_t launch_pbuild(const Entry *entries, uint32_t n, uint32_t K,
                         uint32_t S, int mode, const double *mats,
                         const double *times, const double *rates,
                         const DevModel *model, uint32_t model_stride, T *out,
                         hipStream_t s, const void *copy_src = nullptr,
                         void *copy_dst = nullptr, uint32_t copy_units = 0,
                         T *cmats = nullptr, uint2 *cleaves = nullptr);

struct PbuildDesc {
  const Entry *entries;
  const double *times;
  const double *rates;
  const DevModel *model;
  float *out;
  // or null: an entry whose children are both leaves also writes its
  // matrices to cmats[out slot] ([2][K][5][4]) and its leaves to
  // cleaves[out slot], as launch_pbuild does
  float *cmats;
  uint2 *cleaves;
  uint32_t n;
  uint32_t pad;
};
hipError_t launch_pbuild_batch(const PbuildDesc *desc, uint32_t count, uint32_t max_n,
                               uint32_t K, uint32_t S, hipStream_t s);

// ---- prune4.hip ---------------------------------------------------------
// FP32 partials, 4 states, K categories per workgroup (one wavefront per
// category and 64-pattern sub-tile), LDS stack for in-pass children,
// power-of-two normalisation, fused root / category mix / weighted sum.
struct F32Args {
  const Entry *entries;
  const float *pmats;     // [n][2][K][16]
  const uint32_t *codes32;  // [L][cstride] state codes 0..4, 4 per dword
  float4 *partials;       // [slots][K][P]
  int16_t *exps;          // [slots][K][P] cumulative binary exponents
  double *site;           // [K][P]
  double *mixed;          // [P]
  const uint32_t *weights;
  double *blocksum;       // [ntiles] per-tile partial sums
  uint32_t *counter;      // last-tile ticket (self-resetting)
  double *total;
  double *host_total;     // pinned host mirror of *total, or null
  uint32_t *host_flag;    // with host_total: seq stored after it (system-scope release)
  uint32_t seq;
  double pi0, pi1, pi2, pi3;
  const double *pik;        // [K][4] root frequencies per category, or null (pi0..pi3)
  const uint8_t *classes;   // [P] category selected per pattern, or null (mean over categories)
  double ln_inv_k;
  uint32_t n, P, K, T, nslots;
  uint32_t cstride;       // code row stride (a multiple of 64 T)
  uint32_t ntiles;        // tiles of this evaluation (blocksum entries)
  // Cherries (both children leaves), or null: their partials are then not
  // stored.  cmats [slots][2][K][5][4]: the two child matrices recorded for a
  // cherry slot by the P(t) build, as columns plus an all-ones fifth column;
  // cleaves [slots]: its two leaves.
  const float *cmats;
  const uint2 *cleaves;
};
hipError_t launch_prune_f32(const F32Args &a, uint32_t ppl, bool hbm, uint32_t blocks,
                            hipStream_t s);
size_t prune_f32_lds_bytes(uint32_t K, uint32_t T, uint32_t nslots, uint32_t n,
                           uint32_t ppl);
// One state per lane (a quad per pattern, 16 patterns per wavefront, 16 T
// per tile): lists whose stored children are read in place only (full
// evaluations); values bit-identical to launch_prune_f32's.
hipError_t launch_prune_f32_s1(const F32Args &a, uint32_t blocks, hipStream_t s);
size_t prune_f32_s1_lds_bytes(uint32_t K, uint32_t T, uint32_t nslots, uint32_t n);
// persistent batched form over many instances (one F32Args each)
hipError_t prepare_prune_f32_batch(int max_lds_bytes);
// s1: one state per lane (prune_f32_s1_kernel's form; items of 16 T patterns)
hipError_t batch_grid(uint32_t ppl, bool hbm, bool s1, uint32_t threads, size_t lds,
                      uint32_t *grid);
hipError_t launch_prune_f32_batch(const F32Args *insts, const uint2 *items, uint32_t nitems,
                                  uint32_t *queue, uint32_t ppl, bool hbm, bool s1,
                                  uint32_t threads, size_t lds, uint32_t grid, hipStream_t s);
hipError_t prepare_prune_f32(int max_lds_bytes);
// Partial updates that form one chain (entry 0 reads two stored children,
// entry i > 0 the entry before it and one stored child): srcs[n + 1] lists
// the stored children (SRC_HBM slot or SRC_TIP leaf) in the order the entries
// read them; one 64-pattern workgroup of K wavefronts.
// a.ntiles must be path_tiles(P, K)
hipError_t launch_prune_path(const F32Args &a, const uint32_t *srcs, hipStream_t s);
uint32_t path_tiles(uint32_t P, uint32_t K);

// ---- prune_codon.hip ----------------------------------------------------
// S <= 64 states on MFMA: FP16 partials [slot][K][Ppad/32][64 lanes][32],
// FP16 matrices in the kernel's A layout [n][2][K][64][64].
struct CodonArgs {
  const Entry *entries;
  const _Float16 *amats;   // [entry][side][K][hi, lo][64][64], permuted columns
  const uint8_t *codes;    // [L][cstride] state indices, >= S unknown
  _Float16 *partials;      // [slot][K][Ppad / 32][4][64][8]
  int16_t *exps;           // [slot][K][Ppad]
  double *site;            // [K][P]
  const double *pi;        // [64], zero beyond S
  uint32_t cstride, n, P, Ppad, K, S;
};
size_t prune_codon_lds_bytes(uint32_t n);
hipError_t prepare_prune_codon(int max_lds_bytes);
hipError_t launch_prune_codon(const CodonArgs &a, hipStream_t s);
// U[q][el] = V[j][q] V^-1[q][k] of the model (64 x 4096 doubles), once per
// model; the matrices of a propose are then E U (prune_codon.hip)
hipError_t launch_ubuild_codon(const DevModel *model, double *U, hipStream_t s);
hipError_t launch_pbuild_codon(const Entry *entries, uint32_t n, uint32_t K, const double *times,
                               const double *rates, const DevModel *model, const double *U,
                               _Float16 *out, hipStream_t s);

// ---- prune64.hip --------------------------------------------------------
// FP64 plain baseline: one lane per (pattern, category), b3's e^scale_ln
// flag scaling, children read from HBM.
struct F64Args {
  const Entry *entries;
  const double *pmats;    // [n][2][K][16]
  const uint8_t *codes;   // [L][cstride]
  double4 *partials;      // [slots][K][P]
  uint8_t *scales;        // [slots][K][P]
  uint32_t *scale_sums;   // [K][P]
  double *site;           // [K][P]
  double pi0, pi1, pi2, pi3;
  const double *pik;      // [K][4] per category, or null (pi0..pi3)
  double threshold, mult;
  uint32_t scale_ln;
  uint32_t n, P, K, L;
  uint32_t cstride;
};
hipError_t launch_prune_f64(const F64Args &a, hipStream_t s);

// ---- prune_prop.hip ------------------------------------------------------
// Propagation storage: per-node projections [2N slots][K][P] of double4
// (b3's flag scaling) or float4 (binary exponents).  Entries: updated leaves
// (out slot, left = leaf id), internal nodes (out, left and right child
// slots), the root last (out = its slot, children slots).
struct PropArgs {
  const Entry *entries;
  const void *mats;         // [n - 1][K][16] of double or float, entry order
  const uint8_t *codes;     // [L][cstride] state codes 0..4
  void *proj;               // [2N][K][P] double4 or float4
  uint8_t *scales;          // double: [2N][K][P] flags
  int16_t *exps;            // float: [2N][K][P] cumulative exponents
  uint32_t *scale_sums;     // double: [K][P]
  double *site;             // [K][P]
  double pi0, pi1, pi2, pi3;
  const double *pik;        // [K][4] or null
  double threshold, mult;
  uint32_t scale_ln;
  uint32_t n, nleaves, P, K, cstride;
};
hipError_t launch_prune_prop(const PropArgs &a, bool f64, hipStream_t s);
// One transition per listed node but the root, [m][K][16]: mode 0 casts host
// matrices [K][nm][16]; mode 1 builds P(times[m] * rates[c]).
template <typename T>
hipError_t launch_pbuild_prop(uint32_t nm, uint32_t K, int mode, const double *mats,
                              const double *times, const double *rates, const DevModel *model,
                              T *out, hipStream_t s);

// Category mix + weighted sum with a deterministic last-block reduction.
// With classes, pattern i takes category classes[i]'s value instead of the
// mean over categories.
struct MixArgs {
  const double *site;  // [K][P]
  const uint8_t *classes;  // [P] or null
  double *mixed;       // [P]
  const uint32_t *weights;
  double *blocksum;    // [mix_blocks(P)]
  uint32_t *counter;   // last-block ticket (self-resetting)
  double *total;
  double *host_total;     // pinned host mirror of *total, or null
  uint32_t *host_flag;    // with host_total: seq stored after it (system-scope release)
  uint32_t seq;
  double ln_inv_k;
  uint32_t P, K;
};
hipError_t launch_mix(const MixArgs &a, hipStream_t s);
uint32_t mix_blocks(uint32_t P);

}  // namespace b3h
