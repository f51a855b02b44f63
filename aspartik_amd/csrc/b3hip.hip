// Host side of the b3hip C ABI (include/b3hip.h): handle lifecycle, the
// EditBuf slot bookkeeping, the update-list schedule and kernel launches.
#include "../../include/b3hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "reduce.hpp"
#include "schedule.hpp"

using namespace b3h;

namespace {

thread_local std::string g_err;

b3h_status fail(b3h_status s, const std::string &msg) {
  g_err = msg;
  return s;
}

}  // namespace

namespace b3h {
// shared with the other host translation units (batch.hip)
b3h_status host_fail(b3h_status s, const std::string &msg) { return fail(s, msg); }
uint8_t host_remap_code(uint8_t v, uint32_t S);
}  // namespace b3h

namespace {

#define HIPCHK(expr)                                                   \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess)                                              \
      return fail(B3H_ERR_HIP, std::string(#expr) + ": " +             \
                                   hipGetErrorString(e_));             \
  } while (0)

constexpr int kStages = 2;
constexpr uint32_t kMaxSlots = 32;
constexpr size_t kMaxLds = 160 * 1024;

}  // namespace

struct b3h_calc {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t precision = B3H_F32_SCALED;
  uint32_t L = 0, I = 0, P = 0, K = 0, S = 4;
  uint32_t scale_ln = 30;
  uint32_t T = 1;    // 64-lane pattern rows (wavefronts per category) per workgroup
  uint32_t PPL = 1;  // patterns per lane
  bool S1 = false;   // full evaluations with one state per lane (16 patterns per wavefront)
  uint32_t cstride = 0;  // device code row stride (P rounded up to 64)
  uint32_t blocks = 0;

  // device buffers
  uint8_t *d_codes = nullptr;
  void *d_partials = nullptr;   // float4 or double4 [2I][K][P]
  int16_t *d_exps = nullptr;    // F32 path
  uint8_t *d_scales = nullptr;  // F64 path flags [2I][K][P]
  uint32_t *d_scale_sums = nullptr, *d_scale_sums_bak = nullptr;
  double *d_site = nullptr, *d_mixed = nullptr;  // [2][K][P], [2][P]: result slots
  uint32_t *d_weights = nullptr;
  double *d_blocksum = nullptr, *d_mix_blocksum = nullptr;
  uint32_t *d_counter = nullptr;
  double *d_total = nullptr;  // [2]
  void *d_pmats = nullptr;  // [I][2][K][S*S] float or double
  // FP32 path: cherries' recorded matrices [2I slots][2][K][5][4] and leaves
  // [2I]; slot_cherry[s] is 1 while slot s holds a cherry (host mirror)
  float *d_cmats = nullptr;
  uint2 *d_cleaves = nullptr;
  std::vector<uint8_t> slot_cherry;
  DevModel *d_model = nullptr;  // [K]; only [0] is used while one model serves all
  uint8_t *d_classes = nullptr;  // [P] per-pattern category (HeteroLikelihood)
  double *d_ucodon = nullptr;  // MFMA path: [64][4096] eigenvector products
  bool model_set = false;

  // staging (pinned host -> device), double buffered
  size_t stage_bytes = 0;
  unsigned char *h_stage[kStages] = {nullptr, nullptr};
  double *h_total = nullptr;  // pinned host mirrors of the result slots' totals (kernel-written)
  uint32_t *h_flag = nullptr;  // per slot: sequence number stored after the mirror
  uint32_t seq = 0, expect[2] = {0u, 0u};  // last number issued; expected per slot
  unsigned char *d_stage[kStages] = {nullptr, nullptr};
  hipEvent_t stage_ev[kStages] = {nullptr, nullptr};
  // a stage whose last reader is a kernel that stores a sequence number in
  // h_flag is free once either slot's flag has reached that number (the
  // stream runs in order); the others are released by stage_ev
  uint32_t stage_seq[kStages] = {0u, 0u};
  bool stage_by_flag[kStages] = {false, false};
  int stage_next = 0;

  // host state
  std::vector<uint8_t> editbuf;  // crates/sk/src/editbuf.rs
  std::vector<uint8_t> editbuf_next;
  Scheduler sched;
  std::vector<uint32_t> path_srcs;  // sources of a path-kernel list, in use order
  double freqs[64] = {0};
  // per-category models (hetero.rs:122-135): model[c] and its frequencies
  bool per_category = false;
  double freqs_k[16][4] = {{0}};
  bool classes_on = false;
  // Results (site values, mixed values, total) live in two slots, like the
  // likelihood layer's SkSliceBuf (hetero.rs:30,149-156): a propose writes
  // slot rslot ^ 1, accept makes it the accepted slot, reject keeps rslot.
  uint32_t rslot = 0;
  bool pending = false;   // a propose since the last accept / reject
  bool accepted_valid = false;
  bool has_result = false;
  uint64_t bytes = 0;

  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> ev_start, ev_stop;
  size_t ev_used = 0;
  double prof_ms = 0.0;
  uint64_t prof_launches = 0;
  std::vector<double> prof_each;  // per-launch ms since enable
};

namespace {

b3h_status dev_alloc(b3h_calc *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(B3H_ERR_OOM, "hipMalloc(" + std::to_string(bytes) +
                                 " B) failed: " + hipGetErrorString(e));
  }
  c->bytes += bytes;
  return B3H_OK;
}

void release(b3h_calc *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->d_codes, c->d_partials, c->d_exps, c->d_scales, c->d_classes,
                  c->d_scale_sums, c->d_scale_sums_bak, c->d_site, c->d_mixed,
                  c->d_weights, c->d_blocksum, c->d_mix_blocksum,
                  c->d_counter, c->d_total, c->d_pmats, c->d_model, c->d_ucodon,
                  c->d_cmats, c->d_cleaves};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_total) (void)hipHostFree(c->h_total);
  if (c->h_flag) (void)hipHostFree(c->h_flag);
  for (int s = 0; s < kStages; s++) {
    if (c->h_stage[s]) (void)hipHostFree(c->h_stage[s]);
    if (c->d_stage[s]) (void)hipFree(c->d_stage[s]);
    if (c->stage_ev[s]) (void)hipEventDestroy(c->stage_ev[s]);
  }
  for (auto e : c->ev_start) (void)hipEventDestroy(e);
  for (auto e : c->ev_stop) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// Tip codes as the kernels read them: 4 states use cpu4.rs:117-125's remap
// {1,2,4,8} -> {0,1,2,3}, everything else 4 (unknown); S > 4 keeps indices,
// clamping unknowns to S.
uint8_t remap_code(uint8_t v, uint32_t S) { return b3h::host_remap_code(v, S); }

}  // namespace

uint8_t b3h::host_remap_code(uint8_t v, uint32_t S) {
  if (S == 4) {
    switch (v) {
      case 1: return 0;
      case 2: return 1;
      case 4: return 2;
      case 8: return 3;
      default: return 4;
    }
  }
  return v < S ? v : (uint8_t)S;
}

namespace {

b3h_status check_handle(const b3h_calc *c) {
  if (!c) return fail(B3H_ERR_ARG, "null calculator handle");
  return B3H_OK;
}

b3h_status set_device(const b3h_calc *c) {
  HIPCHK(hipSetDevice(c->device));
  return B3H_OK;
}

bool is_prop(const b3h_calc *c) {
  return c->precision == B3H_F64_PROPAGATION || c->precision == B3H_F32_PROPAGATION;
}

// Slot the result reads use: the pending propose's, else the accepted one.
b3h_status result_slot(const b3h_calc *c, uint32_t *slot) {
  if (c->pending) *slot = c->rslot ^ 1u;
  else if (c->accepted_valid) *slot = c->rslot;
  else return fail(B3H_ERR_STATE, c->has_result ? "no accepted result" : "no propose yet");
  return B3H_OK;
}

// Validate an update list: node ids, child ids, root last, children before
// parents within the list.
b3h_status validate_list(const b3h_calc *c, uint32_t n,
                         const uint32_t *internals, const uint32_t *children) {
  if (n == 0) return fail(B3H_ERR_ARG, "empty update list");
  if (n > c->I) return fail(B3H_ERR_ARG, "update list longer than internal node count");
  if (!internals || !children) return fail(B3H_ERR_ARG, "null update list");
  const uint32_t N = 2 * c->L - 1;
  for (uint32_t i = 0; i < n; i++) {
    if (internals[i] < c->L || internals[i] >= N)
      return fail(B3H_ERR_ARG, "internal node id out of range at " + std::to_string(i));
    for (int s = 0; s < 2; s++)
      if (children[2 * i + s] >= N || children[2 * i + s] == internals[i])
        return fail(B3H_ERR_ARG, "child node id out of range at " + std::to_string(i));
  }
  return B3H_OK;
}

bool seq_reached(const b3h_calc *c, uint32_t q) {
  for (int slot = 0; slot < 2; slot++)
    if ((int32_t)(__atomic_load_n(c->h_flag + slot, __ATOMIC_ACQUIRE) - q) >= 0) return true;
  return false;
}

unsigned char *next_stage(b3h_calc *c, int *which) {
  int s = c->stage_next;
  c->stage_next = (s + 1) % kStages;
  if (c->stage_by_flag[s]) {
    // poll (up to 20 ms) for the last reader's sequence number, else wait
    // for the stream, which a failed launch also ends
    bool done = seq_reached(c, c->stage_seq[s]);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 1; !done; i++) {
      __builtin_ia32_pause();
      done = seq_reached(c, c->stage_seq[s]);
      if (!done && (i & 1023u) == 0 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20))
        break;
    }
    if (!done) (void)hipStreamSynchronize(c->stream);
    c->stage_by_flag[s] = false;
  } else {
    (void)hipEventSynchronize(c->stage_ev[s]);  // previous copy from it is done
  }
  *which = s;
  return c->h_stage[s];
}

// Restores the EditBuf selectors flipped by a propose unless disarmed (the
// pre-propose copy is kept in editbuf_next after the swap).
struct EditbufUndo {
  b3h_calc *c;
  bool armed = true;
  ~EditbufUndo() {
    if (armed) c->editbuf.swap(c->editbuf_next);
  }
};

b3h_status run_propose(b3h_calc *c, uint32_t n, const uint32_t *internals,
                       const uint32_t *children, const double *mats,
                       const double *times, const double *rates,
                       const double *freqs) {
  b3h_status st = set_device(c);
  if (st) return st;
  st = validate_list(c, n, internals, children);
  if (st) return st;
  const uint32_t K = c->K, S = c->S, SS = S * S;

  // EditBuf::set_edited for every node of the list (cpu4.rs:101-106), on a
  // copy that is committed only once the list has been validated
  c->editbuf_next = c->editbuf;
  for (uint32_t i = 0; i < n; i++) {
    uint8_t &m = c->editbuf_next[internals[i] - c->L];
    m = (uint8_t)(((m & 1u) ^ (uint8_t)~(m >> 1)) & 3u);
  }

  const bool f32 = c->precision == B3H_F32_SCALED;
  const bool codon = c->precision == B3H_F16_MFMA;
  const bool reorder = f32 || codon;
  const uint32_t cap = reorder ? c->sched.capacity : 0;
  std::string err;
  // FP32: the new slot of every listed node holds a cherry iff both its
  // children are leaves; stored children in cherry slots are recomputed
  // (set again before every schedule build, which may clear the flags of
  // cherries it has to store)
  uint8_t *cherry = f32 ? c->slot_cherry.data() : nullptr;
  auto mark_cherries = [&]() {
    if (!cherry) return;
    for (uint32_t i = 0; i < n; i++)
      cherry[Scheduler::slot_of(internals[i], c->L, c->editbuf_next.data())] =
          (children[2 * i] < c->L && children[2 * i + 1] < c->L) ? 1 : 0;
  };
  // Partial updates on the FP32 path: a list that forms one chain (entry 0
  // reads two stored children, every later entry the entry before it and
  // one stored child that no earlier entry of this pass wrote) runs in the
  // path kernel, which reads its stored children a segment at a time.
  bool path = false;
  std::vector<uint32_t> &psrc = c->path_srcs;
  psrc.clear();
  if (f32 && n <= 64) {
    mark_cherries();
    if (!c->sched.build(n, internals, children, c->L, c->editbuf_next.data(), true, 0, &err,
                        cherry))
      return fail(B3H_ERR_ARG, err);
    path = true;
    for (uint32_t i = 0; i < n && path; i++) {
      const Entry &e = c->sched.entries[i];
      const uint32_t l = e.left & SRC_TAG, r = e.right & SRC_TAG;
      if (l == SRC_LDS || r == SRC_LDS) {
        path = false;
      } else if (i == 0) {
        if (l == SRC_PREV || r == SRC_PREV) path = false;
        psrc.push_back(e.left);
        psrc.push_back(e.right);
      } else if ((l == SRC_PREV) == (r == SRC_PREV)) {
        path = false;  // a second chain starts here
      } else {
        psrc.push_back(l == SRC_PREV ? e.right : e.left);
      }
    }
    for (uint32_t k = 0; k < psrc.size() && path; k++)
      if ((psrc[k] & SRC_TAG) == SRC_HBM)
        for (uint32_t j = 0; j + 1 < k && j < n; j++)  // written earlier in this pass
          if (c->sched.entries[j].out == (psrc[k] & SRC_SLOT)) path = false;
  }
  if (!path) mark_cherries();
  if (!path &&
      !c->sched.build(n, internals, children, c->L, c->editbuf_next.data(), reorder, cap, &err,
                      cherry))
    return fail(B3H_ERR_ARG, err);
  c->editbuf.swap(c->editbuf_next);
  // every error return below restores the slot selectors, so a failed
  // propose leaves the accepted state readable (no proposal is pending)
  EditbufUndo undo{c};

  // stage: entries | matrices (mode 0) or times + rates (mode 1) or branch
  // lengths [K][2n] (mode 2) | pi (the MFMA path stages host matrices
  // already in its FP16 A layout)
  const int mode = mats ? 0 : rates ? 1 : 2;
  const size_t ebytes = (size_t)n * sizeof(Entry);
  const size_t mbytes = !mats ? 0 : codon ? (size_t)K * 2 * n * 8192 * 2 : (size_t)K * 2 * n * SS * 8;
  const size_t tbytes = !times ? 0 : (mode == 2 ? (size_t)K : 1u) * 2 * n * 8;
  const size_t rbytes = rates ? (size_t)K * 8 : 0;
  const size_t off_m = (ebytes + 255) & ~(size_t)255;
  const size_t off_t = off_m + ((mbytes + 255) & ~(size_t)255);
  const size_t off_r = off_t + ((tbytes + 255) & ~(size_t)255);
  const size_t off_pi = off_r + ((rbytes + 255) & ~(size_t)255);
  const size_t off_src = off_pi + 64 * 8;
  const size_t total = off_src + (path ? psrc.size() * 4 : 0);
  if (total > c->stage_bytes) return fail(B3H_ERR_ARG, "staging overflow");
  int s = 0;
  unsigned char *h = next_stage(c, &s);
  std::memcpy(h, c->sched.entries.data(), ebytes);
  if (mats && !codon) std::memcpy(h + off_m, mats, mbytes);
  if (mats && codon) {
    // FP16 A layout of the kernel: entry order, [side][category][hi, lo][row]
    // [slot], slot 16 s + 8 hh + m holding column row(hh, 8 s + m) and
    // P = hi + 2^-20 lo (prune_codon.hip)
    _Float16 *dst = reinterpret_cast<_Float16 *>(h + off_m);
    for (uint32_t e = 0; e < n; e++) {
      const uint32_t orig = entry_index(c->sched.entries[e].meta);
      for (uint32_t side = 0; side < 2; side++)
        for (uint32_t cat = 0; cat < K; cat++) {
          const double *src = mats + ((size_t)cat * 2 * n + 2 * orig + side) * SS;
          _Float16 *out = dst + ((size_t)(2 * e + side) * K + cat) * 8192;
          for (uint32_t j = 0; j < 64; j++)
            for (uint32_t kslot = 0; kslot < 64; kslot++) {
              const uint32_t q = 8 * (kslot / 16) + kslot % 8, hh = (kslot / 8) & 1u;
              const uint32_t k = 32u * (q / 16u) + 8u * ((q % 16u) / 4u) + 4u * hh + (q % 4u);
              const double v = (j < S && k < S) ? src[j * S + k] : 0.0;
              const _Float16 hi = (_Float16)(float)v;
              out[j * 64 + kslot] = hi;
              out[4096 + j * 64 + kslot] = (_Float16)(float)((v - (double)hi) * 0x1p20);
            }
        }
    }
  }
  if (times) std::memcpy(h + off_t, times, tbytes);
  if (rates) std::memcpy(h + off_r, rates, rbytes);
  // root frequencies: [S], or [K][4] when the device models differ per
  // category (K <= 16, so the 64-entry block holds them)
  const bool pik = c->per_category && mode != 0;
  double *hpi = reinterpret_cast<double *>(h + off_pi);
  for (uint32_t j = 0; j < 64; j++) hpi[j] = j < S ? freqs[j] : 0.0;
  if (pik)
    for (uint32_t k = 0; k < K; k++)
      for (uint32_t j = 0; j < 4; j++) hpi[4 * k + j] = c->freqs_k[k][j];
  if (path && !psrc.empty()) std::memcpy(h + off_src, psrc.data(), psrc.size() * 4);
  unsigned char *d = c->d_stage[s];
  // 4-state FP32: the P(t) build reads its inputs from the pinned staging
  // buffer and writes the device copy that the pruning kernel reads, which
  // saves a copy launch per evaluation; the other paths copy first
  const bool fused_upload = f32 && !codon;
  if (!fused_upload) {
    HIPCHK(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->stage_ev[s], c->stream));
  }
  const Entry *d_entries = reinterpret_cast<const Entry *>(d);
  const double *d_mats = reinterpret_cast<const double *>(d + off_m);
  const double *d_times = reinterpret_cast<const double *>(d + off_t);
  const double *d_rates = reinterpret_cast<const double *>(d + off_r);

  if (mode == 0)
    for (uint32_t j = 0; j < S; j++) c->freqs[j] = freqs[j];
  const double *d_pik = pik ? reinterpret_cast<const double *>(d + off_pi) : nullptr;
  const uint32_t mstride = c->per_category ? 1u : 0u;
  const uint8_t *d_classes = c->classes_on ? c->d_classes : nullptr;
  // results go to the slot that is not the accepted one
  const uint32_t ws = c->rslot ^ 1u;
  double *r_site = c->d_site + (size_t)ws * K * c->P;
  double *r_mixed = c->d_mixed + (size_t)ws * c->P;
  double *r_total = c->d_total + ws;

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->profiling) {
    if (c->ev_used == c->ev_start.size()) {
      hipEvent_t a, b;
      HIPCHK(hipEventCreate(&a));
      HIPCHK(hipEventCreate(&b));
      c->ev_start.push_back(a);
      c->ev_stop.push_back(b);
    }
    e0 = c->ev_start[c->ev_used];
    e1 = c->ev_stop[c->ev_used];
    c->ev_used++;
  }

  if (codon) {
    const _Float16 *amats = reinterpret_cast<const _Float16 *>(d + off_m);
    if (mode == 1) {
      HIPCHK(launch_pbuild_codon(d_entries, n, K, d_times, d_rates, c->d_model, c->d_ucodon,
                                 static_cast<_Float16 *>(c->d_pmats), c->stream));
      amats = static_cast<const _Float16 *>(c->d_pmats);
    }
    CodonArgs a{};
    a.entries = d_entries;
    a.amats = amats;
    a.codes = c->d_codes;
    a.partials = static_cast<_Float16 *>(c->d_partials);
    a.exps = c->d_exps;
    a.site = r_site;
    a.pi = reinterpret_cast<const double *>(d + off_pi);
    a.cstride = c->cstride;
    a.n = n;
    a.P = c->P;
    a.Ppad = c->cstride;
    a.K = K;
    a.S = S;
    if (e0) HIPCHK(hipEventRecord(e0, c->stream));
    HIPCHK(launch_prune_codon(a, c->stream));
    if (e1) HIPCHK(hipEventRecord(e1, c->stream));
    MixArgs m{};
    m.site = r_site;
    m.classes = d_classes;
    m.mixed = r_mixed;
    m.weights = c->d_weights;
    m.blocksum = c->d_mix_blocksum;
    m.counter = c->d_counter;
    m.total = r_total;
    m.host_total = c->h_total + ws;
    m.host_flag = c->h_flag + ws;
    m.seq = c->expect[ws] = ++c->seq;
    m.ln_inv_k = std::log(1.0 / (double)K);
    m.P = c->P;
    m.K = K;
    HIPCHK(launch_mix(m, c->stream));
  } else if (f32) {
    HIPCHK(launch_pbuild<float>(reinterpret_cast<const Entry *>(h), n, K, S, mode,
                                reinterpret_cast<const double *>(h + off_m),
                                reinterpret_cast<const double *>(h + off_t),
                                reinterpret_cast<const double *>(h + off_r), c->d_model, mstride,
                                static_cast<float *>(c->d_pmats), c->stream, h, d,
                                (uint32_t)((total + 15) / 16), c->d_cmats, c->d_cleaves));
    F32Args a{};
    a.entries = d_entries;
    a.pmats = static_cast<const float *>(c->d_pmats);
    a.codes32 = reinterpret_cast<const uint32_t *>(c->d_codes);
    a.cstride = c->cstride;
    a.partials = static_cast<float4 *>(c->d_partials);
    a.exps = c->d_exps;
    a.site = r_site;
    a.mixed = r_mixed;
    a.weights = c->d_weights;
    a.blocksum = c->d_blocksum;
    a.counter = c->d_counter;
    a.total = r_total;
    a.host_total = c->h_total + ws;
    a.host_flag = c->h_flag + ws;
    a.seq = c->expect[ws] = ++c->seq;
    // the P(t) build, the last reader of h, runs before this kernel stores
    // its sequence number
    c->stage_seq[s] = a.seq;
    c->stage_by_flag[s] = true;
    a.pi0 = freqs[0];
    a.pi1 = freqs[1];
    a.pi2 = freqs[2];
    a.pi3 = freqs[3];
    a.pik = d_pik;
    a.classes = d_classes;
    a.ln_inv_k = std::log(1.0 / (double)K);
    a.n = n;
    a.P = c->P;
    a.K = K;
    a.T = c->T;
    a.nslots = std::max<uint32_t>(c->sched.slots_used, 1u);
    a.cmats = c->d_cmats;
    a.cleaves = c->d_cleaves;
    if (e0) HIPCHK(hipEventRecord(e0, c->stream));
    // Lists that read children from HBM (partial updates) use the variant
    // that prefetches them, with one pattern per lane (its registers do not
    // fit two); lists computed from registers and LDS (full evaluations) use
    // the configured patterns per lane.
    if (path) {
      a.ntiles = path_tiles(c->P, K);
      HIPCHK(launch_prune_path(a, reinterpret_cast<const uint32_t *>(d + off_src), c->stream));
    } else {
      // stored cherries are recomputed only by the HBM-prefetch variant
      const bool hbm = c->sched.hbm_sources * 8 > n || c->sched.cherry_sources > 0;
      if (c->S1 && !hbm) {
        const uint32_t tp = 16u * c->T;
        a.ntiles = (c->P + tp - 1) / tp;
        HIPCHK(launch_prune_f32_s1(a, a.ntiles, c->stream));
      } else {
        const uint32_t ppl = hbm ? 1u : c->PPL;
        const uint32_t tp = 64u * c->T * ppl;
        a.ntiles = (c->P + tp - 1) / tp;
        HIPCHK(launch_prune_f32(a, ppl, hbm, a.ntiles, c->stream));  // hbm => ppl 1
      }
    }
    if (e1) HIPCHK(hipEventRecord(e1, c->stream));
  } else {
    HIPCHK(launch_pbuild<double>(d_entries, n, K, S, mode, d_mats, d_times,
                                 d_rates, c->d_model, mstride,
                                 static_cast<double *>(c->d_pmats), c->stream));
    F64Args a{};
    a.entries = d_entries;
    a.pmats = static_cast<const double *>(c->d_pmats);
    a.codes = c->d_codes;
    a.cstride = c->cstride;
    a.partials = static_cast<double4 *>(c->d_partials);
    a.scales = c->d_scales;
    a.scale_sums = c->d_scale_sums;
    a.site = r_site;
    a.pi0 = freqs[0];
    a.pi1 = freqs[1];
    a.pi2 = freqs[2];
    a.pi3 = freqs[3];
    a.pik = d_pik;
    a.scale_ln = c->scale_ln;
    a.mult = std::exp((double)c->scale_ln);
    a.threshold = std::exp(-(double)c->scale_ln);
    a.n = n;
    a.P = c->P;
    a.K = K;
    a.L = c->L;
    if (e0) HIPCHK(hipEventRecord(e0, c->stream));
    HIPCHK(launch_prune_f64(a, c->stream));
    MixArgs m{};
    m.site = r_site;
    m.classes = d_classes;
    m.mixed = r_mixed;
    m.weights = c->d_weights;
    m.blocksum = c->d_mix_blocksum;
    m.counter = c->d_counter;
    m.total = r_total;
    m.host_total = c->h_total + ws;
    m.host_flag = c->h_flag + ws;
    m.seq = c->expect[ws] = ++c->seq;
    m.ln_inv_k = std::log(1.0 / (double)K);
    m.P = c->P;
    m.K = K;
    HIPCHK(launch_mix(m, c->stream));
    if (e1) HIPCHK(hipEventRecord(e1, c->stream));
  }
  undo.armed = false;
  c->has_result = true;
  c->pending = true;
  return B3H_OK;
}

// CudaPropagationsCalculator::propose (cuda_propagations/mod.rs:77-141):
// selectors of every listed node are edited, one transition per node but the
// root, leaves' projections, internal nodes children first, the root.
b3h_status run_propose_prop(b3h_calc *c, uint32_t nn, const uint32_t *nodes, uint32_t nl,
                            const uint32_t *children, const double *mats, const double *times,
                            const double *rates, const double *freqs) {
  b3h_status st = set_device(c);
  if (st) return st;
  const uint32_t L = c->L, N = 2 * L - 1, K = c->K;
  if (nn == 0) return fail(B3H_ERR_ARG, "empty update list");
  if (nn > N) return fail(B3H_ERR_ARG, "update list longer than the node count");
  if (!nodes || !children) return fail(B3H_ERR_ARG, "null update list");
  if (nl >= nn) return fail(B3H_ERR_ARG, "the list must end with the root, an internal node");
  for (uint32_t i = 0; i < nn; i++) {
    const bool leaf = i < nl;
    if (leaf ? nodes[i] >= L : (nodes[i] < L || nodes[i] >= N))
      return fail(B3H_ERR_ARG, "node id out of range at " + std::to_string(i));
  }
  for (uint32_t i = 0; i < 2 * (nn - nl); i++)
    if (children[i] >= N || children[i] == nodes[nl + i / 2])
      return fail(B3H_ERR_ARG, "child node id out of range at " + std::to_string(i / 2));
  // EditBuf::set_edited of every listed node (mod.rs:91-93)
  c->editbuf_next = c->editbuf;
  for (uint32_t i = 0; i < nn; i++) {
    uint8_t &m = c->editbuf_next[nodes[i]];
    m = (uint8_t)(((m & 1u) ^ (uint8_t)~(m >> 1)) & 3u);
  }
  c->editbuf.swap(c->editbuf_next);
  EditbufUndo undo{c};
  auto slot = [&](uint32_t node) { return node * 2u + (c->editbuf[node] & 1u); };

  const uint32_t nm = nn - 1;
  const int mode = mats ? 0 : 1;
  const size_t ebytes = (size_t)nn * sizeof(Entry);
  const size_t mbytes = mats ? (size_t)K * nm * 16 * 8 : 0;
  const size_t tbytes = times ? (size_t)nm * 8 : 0;
  const size_t rbytes = rates ? (size_t)K * 8 : 0;
  const size_t off_m = (ebytes + 255) & ~(size_t)255;
  const size_t off_t = off_m + ((mbytes + 255) & ~(size_t)255);
  const size_t off_r = off_t + ((tbytes + 255) & ~(size_t)255);
  const size_t off_pi = off_r + ((rbytes + 255) & ~(size_t)255);
  const size_t total = off_pi + 64 * 8;
  if (total > c->stage_bytes) return fail(B3H_ERR_ARG, "staging overflow");
  int sidx = 0;
  unsigned char *h = next_stage(c, &sidx);
  Entry *ent = reinterpret_cast<Entry *>(h);
  for (uint32_t i = 0; i < nn; i++) {
    Entry e{};
    e.out = slot(nodes[i]);
    if (i < nl) {
      e.left = nodes[i] | SRC_TIP;
    } else {
      e.left = slot(children[2 * (i - nl)]);
      e.right = slot(children[2 * (i - nl) + 1]);
    }
    e.meta = i << 8;
    ent[i] = e;
  }
  if (mats) std::memcpy(h + off_m, mats, mbytes);
  if (times) std::memcpy(h + off_t, times, tbytes);
  if (rates) std::memcpy(h + off_r, rates, rbytes);
  double *hpi = reinterpret_cast<double *>(h + off_pi);
  for (uint32_t j = 0; j < 64; j++) hpi[j] = j < 4 ? freqs[j] : 0.0;
  unsigned char *d = c->d_stage[sidx];
  HIPCHK(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(c->stage_ev[sidx], c->stream));
  if (mode == 0)
    for (uint32_t j = 0; j < 4; j++) c->freqs[j] = freqs[j];

  const bool f64 = c->precision == B3H_F64_PROPAGATION;
  const double *d_mats = reinterpret_cast<const double *>(d + off_m);
  const double *d_times = reinterpret_cast<const double *>(d + off_t);
  const double *d_rates = reinterpret_cast<const double *>(d + off_r);
  if (f64)
    HIPCHK(launch_pbuild_prop<double>(nm, K, mode, d_mats, d_times, d_rates, c->d_model,
                                      static_cast<double *>(c->d_pmats), c->stream));
  else
    HIPCHK(launch_pbuild_prop<float>(nm, K, mode, d_mats, d_times, d_rates, c->d_model,
                                     static_cast<float *>(c->d_pmats), c->stream));

  const uint32_t ws = c->rslot ^ 1u;
  double *r_site = c->d_site + (size_t)ws * K * c->P;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->profiling) {
    if (c->ev_used == c->ev_start.size()) {
      hipEvent_t ea, eb;
      HIPCHK(hipEventCreate(&ea));
      HIPCHK(hipEventCreate(&eb));
      c->ev_start.push_back(ea);
      c->ev_stop.push_back(eb);
    }
    e0 = c->ev_start[c->ev_used];
    e1 = c->ev_stop[c->ev_used];
    c->ev_used++;
  }
  PropArgs a{};
  a.entries = reinterpret_cast<const Entry *>(d);
  a.mats = c->d_pmats;
  a.codes = c->d_codes;
  a.proj = c->d_partials;
  a.scales = c->d_scales;
  a.exps = c->d_exps;
  a.scale_sums = c->d_scale_sums;
  a.site = r_site;
  a.pi0 = freqs[0];
  a.pi1 = freqs[1];
  a.pi2 = freqs[2];
  a.pi3 = freqs[3];
  a.pik = nullptr;
  a.scale_ln = c->scale_ln;
  a.threshold = std::exp(-(double)c->scale_ln);  // mod.rs:267-268
  a.mult = std::exp((double)c->scale_ln);
  a.n = nn;
  a.nleaves = nl;
  a.P = c->P;
  a.K = K;
  a.cstride = c->cstride;
  if (e0) HIPCHK(hipEventRecord(e0, c->stream));
  HIPCHK(launch_prune_prop(a, f64, c->stream));
  if (e1) HIPCHK(hipEventRecord(e1, c->stream));
  MixArgs m{};
  m.site = r_site;
  m.classes = c->classes_on ? c->d_classes : nullptr;
  m.mixed = c->d_mixed + (size_t)ws * c->P;
  m.weights = c->d_weights;
  m.blocksum = c->d_mix_blocksum;
  m.counter = c->d_counter;
  m.total = c->d_total + ws;
  m.host_total = c->h_total + ws;
  m.host_flag = c->h_flag + ws;
  m.seq = c->expect[ws] = ++c->seq;
  m.ln_inv_k = std::log(1.0 / (double)K);
  m.P = c->P;
  m.K = K;
  HIPCHK(launch_mix(m, c->stream));
  undo.armed = false;
  c->has_result = true;
  c->pending = true;
  return B3H_OK;
}

}  // namespace

extern "C" {

void b3h_config_default(b3h_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof *cfg);
  cfg->device = 0;
  cfg->precision = B3H_F32_SCALED;
  cfg->states = 4;
  cfg->scale_ln = 30;
  cfg->tile_patterns = 0;
}

const char *b3h_last_error(void) { return g_err.c_str(); }

int b3h_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

b3h_status b3h_create(const b3h_config *cfg, const uint8_t *samples,
                      uint32_t num_leaves, uint32_t num_patterns,
                      uint32_t num_categories, b3h_calc **out) {
  if (!out) return fail(B3H_ERR_ARG, "null output handle pointer");
  *out = nullptr;
  b3h_config def;
  b3h_config_default(&def);
  if (!cfg) cfg = &def;
  if (!samples) return fail(B3H_ERR_ARG, "null samples");
  if (num_leaves < 2) return fail(B3H_ERR_ARG, "need at least 2 leaves");
  if (num_patterns == 0) return fail(B3H_ERR_ARG, "need at least 1 pattern");
  if (num_categories == 0 || num_categories > 16)
    return fail(B3H_ERR_ARG, "num_categories must be 1..16");
  if (cfg->scale_ln == 0) return fail(B3H_ERR_ARG, "scale_ln must be nonzero");
  if (cfg->precision != B3H_F32_SCALED && cfg->precision != B3H_F64_PLAIN &&
      cfg->precision != B3H_F16_MFMA && cfg->precision != B3H_F64_PROPAGATION &&
      cfg->precision != B3H_F32_PROPAGATION)
    return fail(B3H_ERR_UNSUPPORTED, "unknown precision");
  if (cfg->precision == B3H_F16_MFMA) {
    if (cfg->states < 2 || cfg->states > 64)
      return fail(B3H_ERR_UNSUPPORTED, "the MFMA path supports 2..64 states");
  } else if (cfg->states != 4) {
    return fail(B3H_ERR_UNSUPPORTED, "the FP32, FP64 and propagation paths support 4 states");
  }
  if ((uint64_t)num_leaves * 2 >= (1u << 24))
    return fail(B3H_ERR_ARG, "too many leaves");

  int ndev = b3h_device_count();
  if (ndev == 0) return fail(B3H_ERR_NO_DEVICE, "no HIP device visible");
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(B3H_ERR_NO_DEVICE, "device ordinal out of range");

  b3h_calc *c = new b3h_calc();
  c->device = cfg->device;
  c->precision = cfg->precision;
  c->L = num_leaves;
  c->I = num_leaves - 1;
  c->P = num_patterns;
  c->K = num_categories;
  c->S = cfg->states;
  c->scale_ln = cfg->scale_ln;
  // partial storage: one EditBuf entry per internal node; propagation
  // storage: one selector per node (cuda_propagations/mod.rs:327)
  const bool prop = c->precision == B3H_F64_PROPAGATION || c->precision == B3H_F32_PROPAGATION;
  c->editbuf.assign(prop ? 2 * c->L - 1 : c->I, 0);
  c->has_result = false;

  auto bail = [&](b3h_status s) {
    std::string keep = g_err;
    release(c);
    g_err = keep;
    return s;
  };
  if (hipSetDevice(c->device) != hipSuccess) {
    fail(B3H_ERR_HIP, "hipSetDevice failed");
    return bail(B3H_ERR_HIP);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fail(B3H_ERR_HIP, "hipStreamCreate failed");
    return bail(B3H_ERR_HIP);
  }

  const uint32_t K = c->K, P = c->P, S = c->S;
  const size_t slots = prop ? (size_t)2 * (2 * c->L - 1) : (size_t)2 * c->I;
  const bool f32 = c->precision == B3H_F32_SCALED;
  const bool codon = c->precision == B3H_F16_MFMA;
  const bool f32_prop = c->precision == B3H_F32_PROPAGATION;
  b3h_status st;

  if (prop) {
    // propagation storage: one lane per (pattern, category), no tiles
    c->T = 1;
    c->PPL = 1;
    c->cstride = (P + 63) / 64 * 64;
    c->blocks = 1;
    c->sched.capacity = 0;
  } else if (codon) {
    // MFMA path: 32-pattern tiles, four per 256-thread workgroup, patterns
    // padded to a multiple of 128; FP16 partials of 64 states; no LDS stack
    // (children not in registers are read back from HBM)
    c->T = 4;
    c->PPL = 1;
    c->cstride = (P + 127) / 128 * 128;
    if (prune_codon_lds_bytes(c->I) > kMaxLds) {
      fail(B3H_ERR_UNSUPPORTED, "tree too large for the LDS-staged update list");
      return bail(B3H_ERR_UNSUPPORTED);
    }
    c->sched.capacity = 0;
  } else {
    // Tile choice: T wavefront rows per category (64*K*T <= 1024 threads) and
    // PPL patterns per lane.  Two patterns per lane once the wavefront count
    // (K*P/64) exceeds what the chip holds at once (~16 per CU x 256 CUs).
    const uint32_t SPL = cfg->lane_states ? cfg->lane_states : 4u;
    if (SPL != 4 && SPL != 1) {
      fail(B3H_ERR_ARG, "lane_states must be 0, 1 or 4");
      return bail(B3H_ERR_ARG);
    }
    uint32_t PPL = cfg->lane_patterns;
    if (SPL == 1) PPL = 1;  // the four-states forms it falls back to take one pattern per lane
    if (PPL == 0) PPL = ((uint64_t)K * P / 64 > 4096) ? 2 : 1;
    // patterns one wavefront row of a tile covers in full evaluations
    const uint32_t row = SPL == 1 ? 16u : 64u * PPL;
    uint32_t T = 1;
    if (cfg->tile_patterns) {
      T = std::max<uint32_t>(1u, cfg->tile_patterns / 64u);
    } else {
      // 1024-thread workgroups (measured best at 100,000 patterns, 196
      // workgroups), narrowed while fewer than kMinTiles workgroups would
      // cover the patterns: a tile walks the whole update list, so a
      // half-empty chip does not finish sooner (128 taxa, 12,500 patterns:
      // T = 4 takes 106 us, T = 1 79 us)
      constexpr uint32_t kMinTiles = 192;
      T = std::max<uint32_t>(1u, 16u / K);
      while (T > 1 && (P + row * T - 1) / (row * T) < kMinTiles) T--;
    }
    while (64u * K * T > 1024u && T > 1) T--;
    if (PPL > 2) {
      fail(B3H_ERR_ARG, "lane_patterns must be 0, 1 or 2");
      return bail(B3H_ERR_ARG);
    }
    c->T = T;
    c->PPL = PPL;  // upper bound; partial updates launch with one pattern per lane
    c->S1 = SPL == 1;
    const uint32_t TP = 64 * T * PPL;  // a multiple of the one-state tile, 16 T
    // the most blocks any launch uses
    c->blocks = c->S1 ? (P + 16 * T - 1) / (16 * T) : (P + 64 * T - 1) / (64 * T);
    // LDS stack capacity for this tile shape (and, with one state per lane,
    // the four-states form partial updates fall back to) and the longest
    // update list
    auto lds = [&](uint32_t cap) {
      const size_t b = prune_f32_lds_bytes(K, T, cap, c->I, PPL);
      return c->S1 ? std::max(b, prune_f32_s1_lds_bytes(K, T, cap, c->I)) : b;
    };
    uint32_t cap = kMaxSlots;
    while (cap > 1 && lds(cap) > kMaxLds) cap--;
    if (lds(cap) > kMaxLds) {
      fail(B3H_ERR_UNSUPPORTED, "tree too large for the LDS-staged update list");
      return bail(B3H_ERR_UNSUPPORTED);
    }
    c->sched.capacity = cap;
    c->cstride = (P + TP - 1) / TP * TP;  // whole tiles, no over-read
  }

  std::vector<uint8_t> codes((size_t)c->L * c->cstride, (uint8_t)S);
  for (size_t l = 0; l < c->L; l++)
    for (size_t i = 0; i < P; i++)
      codes[l * c->cstride + i] = remap_code(samples[l * P + i], S);
  if ((st = dev_alloc(c, (void **)&c->d_codes, codes.size()))) return bail(st);
  // partial store, one "pattern unit" per (slot, category, pattern)
  const size_t punits = codon ? slots * K * c->cstride : slots * K * P;
  const size_t elem = codon ? 128 : (f32 || f32_prop) ? 16 : 32;
  if ((st = dev_alloc(c, &c->d_partials, punits * elem))) return bail(st);
  if (f32 || codon || f32_prop) {
    if ((st = dev_alloc(c, (void **)&c->d_exps, punits * 2))) return bail(st);
  } else {
    if ((st = dev_alloc(c, (void **)&c->d_scales, slots * K * P))) return bail(st);
    if ((st = dev_alloc(c, (void **)&c->d_scale_sums, (size_t)K * P * 4))) return bail(st);
    if ((st = dev_alloc(c, (void **)&c->d_scale_sums_bak, (size_t)K * P * 4))) return bail(st);
  }
  if ((st = dev_alloc(c, (void **)&c->d_site, (size_t)2 * K * P * 8))) return bail(st);
  if ((st = dev_alloc(c, (void **)&c->d_mixed, (size_t)2 * P * 8))) return bail(st);
  if ((st = dev_alloc(c, (void **)&c->d_classes, (size_t)P))) return bail(st);
  if ((st = dev_alloc(c, (void **)&c->d_weights, (size_t)P * 4))) return bail(st);
  // the path kernel (partial updates) uses one block per 64 patterns
  // per-tile partial sums (of the pruning kernel or the path kernel)
  if ((st = dev_alloc(c, (void **)&c->d_blocksum,
                      (size_t)std::max<uint32_t>(std::max<uint32_t>(c->blocks, 1u),
                                                 path_tiles(P, K)) *
                          8)))
    return bail(st);
  if ((st = dev_alloc(c, (void **)&c->d_mix_blocksum, (size_t)mix_blocks(P) * 8))) return bail(st);
  if ((st = dev_alloc(c, (void **)&c->d_counter, 64))) return bail(st);
  if ((st = dev_alloc(c, (void **)&c->d_total, 64))) return bail(st);
  // FP32 4-state matrices as columns plus an all-ones column ([5][4] floats)
  const size_t mat_bytes = codon ? 8192 * 2 : f32 ? 80 : (size_t)S * S * (f32_prop ? 4 : 8);
  if ((st = dev_alloc(c, &c->d_pmats, (size_t)c->I * 2 * K * mat_bytes))) return bail(st);
  if (f32) {
    if ((st = dev_alloc(c, (void **)&c->d_cmats, slots * 2 * K * 20 * 4))) return bail(st);
    if ((st = dev_alloc(c, (void **)&c->d_cleaves, slots * 8))) return bail(st);
    c->slot_cherry.assign(slots, 0);
  }
  if ((st = dev_alloc(c, (void **)&c->d_model, (size_t)K * sizeof(DevModel)))) return bail(st);
  if (codon && (st = dev_alloc(c, (void **)&c->d_ucodon, (size_t)64 * 4096 * 8))) return bail(st);

  // staging: entries + host matrices (the larger of the f64 input and the
  // FP16 A layout) + times + rates + stationary frequencies
  const size_t host_mat = std::max<size_t>((size_t)S * S * 8, codon ? 8192 * 2 : 0);
  c->stage_bytes = ((size_t)2 * c->L * sizeof(Entry) + 255) / 256 * 256 +
                   ((size_t)K * 2 * c->I * host_mat + 255) / 256 * 256 +
                   ((size_t)K * 2 * c->I * 8 + 255) / 256 * 256 + ((size_t)K * 8 + 255) / 256 * 256 +
                   64 * 8 + 512;
  // fine-grained: the kernels publish the total before they retire
  if (hipHostMalloc((void **)&c->h_total, 64, hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&c->h_flag, 64, hipHostMallocCoherent) != hipSuccess) {
    fail(B3H_ERR_OOM, "hipHostMalloc failed");
    return bail(B3H_ERR_OOM);
  }
  std::memset(c->h_total, 0, 64);
  std::memset(c->h_flag, 0, 64);  // sequence number 0: nothing computed yet
  for (int s = 0; s < kStages; s++) {
    if (hipHostMalloc((void **)&c->h_stage[s], c->stage_bytes, hipHostMallocDefault) != hipSuccess) {
      fail(B3H_ERR_OOM, "hipHostMalloc staging failed");
      return bail(B3H_ERR_OOM);
    }
    if ((st = dev_alloc(c, (void **)&c->d_stage[s], c->stage_bytes))) return bail(st);
    if (hipEventCreateWithFlags(&c->stage_ev[s], hipEventDisableTiming) != hipSuccess) {
      fail(B3H_ERR_HIP, "hipEventCreate failed");
      return bail(B3H_ERR_HIP);
    }
  }

  std::vector<uint32_t> ones(P, 1u);
  bool ok = hipMemcpy(c->d_codes, codes.data(), codes.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(c->d_weights, ones.data(), (size_t)P * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(c->d_counter, 0, 64) == hipSuccess &&
            hipMemset(c->d_model, 0, (size_t)K * sizeof(DevModel)) == hipSuccess &&
            hipMemset(c->d_classes, 0, P) == hipSuccess &&
            hipMemset(c->d_partials, 0, punits * elem) == hipSuccess;
  if (ok && (f32 || codon || f32_prop)) ok = hipMemset(c->d_exps, 0, punits * 2) == hipSuccess;
  if (ok && f32)
    ok = hipMemset(c->d_cmats, 0, slots * 2 * K * 20 * 4) == hipSuccess &&
         hipMemset(c->d_cleaves, 0, slots * 8) == hipSuccess;
  if (ok && !f32 && !codon && !f32_prop)
    ok = hipMemset(c->d_scales, 0, slots * K * P) == hipSuccess &&
         hipMemset(c->d_scale_sums, 0, (size_t)K * P * 4) == hipSuccess &&
         hipMemset(c->d_scale_sums_bak, 0, (size_t)K * P * 4) == hipSuccess;
  if (ok) ok = hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    fail(B3H_ERR_HIP, "initial device upload failed");
    return bail(B3H_ERR_HIP);
  }
  if ((f32 && prepare_prune_f32((int)kMaxLds) != hipSuccess) ||
      (codon && prepare_prune_codon((int)kMaxLds) != hipSuccess)) {
    fail(B3H_ERR_HIP, "cannot raise the kernel's LDS limit");
    return bail(B3H_ERR_HIP);
  }
  *out = c;
  return B3H_OK;
}

b3h_status b3h_destroy(b3h_calc *c) {
  if (!c) return B3H_OK;
  (void)hipSetDevice(c->device);
  release(c);
  return B3H_OK;
}

b3h_status b3h_set_weights(b3h_calc *c, const uint32_t *weights) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!weights) return fail(B3H_ERR_ARG, "null weights");
  if ((st = set_device(c))) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->d_weights, weights, (size_t)c->P * 4, hipMemcpyHostToDevice));
  return B3H_OK;
}

b3h_status b3h_propose(b3h_calc *c, uint32_t n, const uint32_t *internals,
                       const uint32_t *children, const double *matrices,
                       const double *frequencies) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!matrices || !frequencies) return fail(B3H_ERR_ARG, "null matrices or frequencies");
  if (is_prop(c)) return fail(B3H_ERR_UNSUPPORTED, "propagation storage takes propagation lists");
  return run_propose(c, n, internals, children, matrices, nullptr, nullptr, frequencies);
}

b3h_status b3h_propose_propagations(b3h_calc *c, uint32_t num_nodes, const uint32_t *nodes,
                                    uint32_t num_leaves, const uint32_t *children,
                                    const double *matrices, const double *frequencies) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!is_prop(c)) return fail(B3H_ERR_UNSUPPORTED, "the handle does not use propagation storage");
  if (!matrices || !frequencies) return fail(B3H_ERR_ARG, "null matrices or frequencies");
  return run_propose_prop(c, num_nodes, nodes, num_leaves, children, matrices, nullptr, nullptr,
                          frequencies);
}

b3h_status b3h_propose_propagations_times(b3h_calc *c, uint32_t num_nodes,
                                          const uint32_t *nodes, uint32_t num_leaves,
                                          const uint32_t *children, const double *edge_times,
                                          const double *category_rates) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!is_prop(c)) return fail(B3H_ERR_UNSUPPORTED, "the handle does not use propagation storage");
  if (!c->model_set) return fail(B3H_ERR_STATE, "b3h_set_model must be called first");
  if (!edge_times || !category_rates) return fail(B3H_ERR_ARG, "null edge times or rates");
  if (c->per_category)
    return fail(B3H_ERR_UNSUPPORTED, "propagation storage uses one model for all categories");
  double fr[4];
  for (uint32_t j = 0; j < 4; j++) fr[j] = c->freqs[j];
  return run_propose_prop(c, num_nodes, nodes, num_leaves, children, nullptr, edge_times,
                          category_rates, fr);
}

b3h_status b3h_set_model(b3h_calc *c, const double *q, const double *freqs,
                         uint32_t flags) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!q || !freqs) return fail(B3H_ERR_ARG, "null model arrays");
  const uint32_t S = c->S;
  for (uint32_t i = 0; i < S; i++)
    if (!(freqs[i] > 0.0)) return fail(B3H_ERR_ARG, "frequencies must be positive");
  if ((st = set_device(c))) return st;
  DevModel hm;
  std::memset(&hm, 0, sizeof hm);
  hm.n = (int32_t)S;
  hm.special = (flags & B3H_MODEL_SPECIAL_CASES) ? 1 : 0;
  for (uint32_t i = 0; i < S; i++) hm.freqs[i] = freqs[i];
  for (uint32_t i = 0; i < S * S; i++) hm.q[i] = q[i];
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpyAsync(c->d_model, &hm, sizeof hm, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_eigen(c->d_model, c->stream));
  if (c->precision == B3H_F16_MFMA) HIPCHK(launch_ubuild_codon(c->d_model, c->d_ucodon, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < S; i++) c->freqs[i] = freqs[i];
  c->model_set = true;
  c->per_category = false;
  return B3H_OK;
}

b3h_status b3h_set_category_model(b3h_calc *c, uint32_t category, const double *q,
                                  const double *freqs, uint32_t flags) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!q || !freqs) return fail(B3H_ERR_ARG, "null model arrays");
  if (category >= c->K) return fail(B3H_ERR_ARG, "category out of range");
  if (c->precision == B3H_F16_MFMA)
    return fail(B3H_ERR_UNSUPPORTED, "per-category models need 4 states");
  if (is_prop(c))
    return fail(B3H_ERR_UNSUPPORTED, "propagation storage uses one model for all categories");
  for (uint32_t i = 0; i < 4; i++)
    if (!(freqs[i] > 0.0)) return fail(B3H_ERR_ARG, "frequencies must be positive");
  if ((st = set_device(c))) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!c->per_category) {
    // leaving the shared model: every category starts from it (or from
    // the new model where none was set yet)
    for (uint32_t k = 0; k < c->K; k++) {
      if (c->model_set && k != 0)
        HIPCHK(hipMemcpyAsync(c->d_model + k, c->d_model, sizeof(DevModel),
                              hipMemcpyDeviceToDevice, c->stream));
      for (uint32_t j = 0; j < 4; j++) c->freqs_k[k][j] = c->model_set ? c->freqs[j] : freqs[j];
    }
  }
  DevModel hm;
  std::memset(&hm, 0, sizeof hm);
  hm.n = 4;
  hm.special = (flags & B3H_MODEL_SPECIAL_CASES) ? 1 : 0;
  for (uint32_t i = 0; i < 4; i++) hm.freqs[i] = freqs[i];
  for (uint32_t i = 0; i < 16; i++) hm.q[i] = q[i];
  HIPCHK(hipMemcpyAsync(c->d_model + category, &hm, sizeof hm, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_eigen(c->d_model + category, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint32_t j = 0; j < 4; j++) c->freqs_k[category][j] = freqs[j];
  if (!c->model_set) {  // categories never given a model copy this one
    for (uint32_t k = 0; k < c->K; k++)
      if (k != category) {
        HIPCHK(hipMemcpyAsync(c->d_model + k, c->d_model + category, sizeof(DevModel),
                              hipMemcpyDeviceToDevice, c->stream));
      }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t j = 0; j < 4; j++) c->freqs[j] = freqs[j];
  }
  c->model_set = true;
  c->per_category = true;
  return B3H_OK;
}

b3h_status b3h_propose_lengths(b3h_calc *c, uint32_t n, const uint32_t *internals,
                               const uint32_t *children, const double *lengths) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!c->model_set) return fail(B3H_ERR_STATE, "b3h_set_model must be called first");
  if (!lengths) return fail(B3H_ERR_ARG, "null branch lengths");
  if (c->precision == B3H_F16_MFMA || is_prop(c))
    return fail(B3H_ERR_UNSUPPORTED,
                "per-category branch lengths need the FP32 or FP64 partial paths");
  double fr[64];
  for (uint32_t j = 0; j < c->S; j++) fr[j] = c->freqs[j];
  return run_propose(c, n, internals, children, nullptr, lengths, nullptr, fr);
}

b3h_status b3h_set_classes(b3h_calc *c, const uint32_t *classes) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if ((st = set_device(c))) return st;
  if (!classes) {
    c->classes_on = false;
  } else {
    std::vector<uint8_t> cls(c->P);
    for (uint32_t i = 0; i < c->P; i++) {
      if (classes[i] >= c->K)
        return fail(B3H_ERR_ARG, "class out of range at pattern " + std::to_string(i));
      cls[i] = (uint8_t)classes[i];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->d_classes, cls.data(), c->P, hipMemcpyHostToDevice));
    c->classes_on = true;
  }
  // recompute the total of the current result from its site values (the
  // path of hetero.rs:150-160 when only the classes changed)
  uint32_t slot;
  if (c->pending) slot = c->rslot ^ 1u;
  else if (c->accepted_valid) slot = c->rslot;
  else return B3H_OK;
  MixArgs m{};
  m.site = c->d_site + (size_t)slot * c->K * c->P;
  m.classes = c->classes_on ? c->d_classes : nullptr;
  m.mixed = c->d_mixed + (size_t)slot * c->P;
  m.weights = c->d_weights;
  m.blocksum = c->d_mix_blocksum;
  m.counter = c->d_counter;
  m.total = c->d_total + slot;
  m.host_total = c->h_total + slot;
  m.host_flag = c->h_flag + slot;
  m.seq = c->expect[slot] = ++c->seq;
  m.ln_inv_k = std::log(1.0 / (double)c->K);
  m.P = c->P;
  m.K = c->K;
  HIPCHK(launch_mix(m, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return B3H_OK;
}

b3h_status b3h_propose_times(b3h_calc *c, uint32_t n, const uint32_t *internals,
                             const uint32_t *children, const double *edge_times,
                             const double *category_rates) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!c->model_set) return fail(B3H_ERR_STATE, "b3h_set_model must be called first");
  if (!edge_times || !category_rates) return fail(B3H_ERR_ARG, "null edge times or rates");
  if (is_prop(c)) return fail(B3H_ERR_UNSUPPORTED, "propagation storage takes propagation lists");
  double fr[64];
  for (uint32_t j = 0; j < c->S; j++) fr[j] = c->freqs[j];
  return run_propose(c, n, internals, children, nullptr, edge_times, category_rates, fr);
}

b3h_status b3h_accept(b3h_calc *c) {
  b3h_status st = check_handle(c);
  if (st) return st;
  for (auto &m : c->editbuf) m &= 1u;  // editbuf.rs:48-53
  if (c->pending) {
    c->rslot ^= 1u;
    c->accepted_valid = true;
    c->pending = false;
  }
  if (c->precision == B3H_F64_PLAIN || c->precision == B3H_F64_PROPAGATION) {  // cpu4.rs:78-83
    if ((st = set_device(c))) return st;
    HIPCHK(hipMemcpyAsync(c->d_scale_sums_bak, c->d_scale_sums, (size_t)c->K * c->P * 4,
                          hipMemcpyDeviceToDevice, c->stream));
  }
  return B3H_OK;
}

b3h_status b3h_reject(b3h_calc *c) {
  b3h_status st = check_handle(c);
  if (st) return st;
  for (auto &m : c->editbuf) m = (uint8_t)((m ^ (m >> 1)) & 1u);  // editbuf.rs:55-66
  c->pending = false;
  if (c->precision == B3H_F64_PLAIN || c->precision == B3H_F64_PROPAGATION) {  // cpu4.rs:85-89
    if ((st = set_device(c))) return st;
    HIPCHK(hipMemcpyAsync(c->d_scale_sums, c->d_scale_sums_bak, (size_t)c->K * c->P * 4,
                          hipMemcpyDeviceToDevice, c->stream));
  }
  return B3H_OK;
}

b3h_status b3h_likelihoods(b3h_calc *c, uint32_t category, double *out) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!out) return fail(B3H_ERR_ARG, "null output");
  if (category >= c->K) return fail(B3H_ERR_ARG, "category out of range");
  uint32_t slot;
  if ((st = result_slot(c, &slot))) return st;
  if ((st = set_device(c))) return st;
  HIPCHK(hipMemcpyAsync(out, c->d_site + ((size_t)slot * c->K + category) * c->P, (size_t)c->P * 8,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return B3H_OK;
}

b3h_status b3h_site_mixed(b3h_calc *c, double *out) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!out) return fail(B3H_ERR_ARG, "null output");
  uint32_t slot;
  if ((st = result_slot(c, &slot))) return st;
  if ((st = set_device(c))) return st;
  HIPCHK(hipMemcpyAsync(out, c->d_mixed + (size_t)slot * c->P, (size_t)c->P * 8,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return B3H_OK;
}

b3h_status b3h_total(b3h_calc *c, double *out) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (!out) return fail(B3H_ERR_ARG, "null output");
  uint32_t slot;
  if ((st = result_slot(c, &slot))) return st;
  if ((st = set_device(c))) return st;
  // every kernel that writes a slot's total also stores it in the slot's
  // pinned host mirror and then the launch's sequence number: poll for the
  // number (up to 20 ms), else wait for the stream, which also reports any
  // error of the launch
  const uint32_t want = c->expect[slot];
  bool seen = __atomic_load_n(c->h_flag + slot, __ATOMIC_ACQUIRE) == want;
  if (!seen) {
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 1; !seen; i++) {
      __builtin_ia32_pause();
      seen = __atomic_load_n(c->h_flag + slot, __ATOMIC_ACQUIRE) == want;
      if (!seen && (i & 1023u) == 0 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20))
        break;
    }
  }
  if (!seen) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (__atomic_load_n(c->h_flag + slot, __ATOMIC_ACQUIRE) != want)
      return fail(B3H_ERR_STATE, "no total was computed for the current state");
  }
  *out = *(volatile double *)(c->h_total + slot);  // ordered by the flag's acquire
  return B3H_OK;
}

uint32_t b3h_num_patterns(const b3h_calc *c) { return c ? c->P : 0; }

uint64_t b3h_device_bytes(const b3h_calc *c) { return c ? c->bytes : 0; }

b3h_status b3h_profile_enable(b3h_calc *c, int enable) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if ((st = set_device(c))) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->profiling = enable != 0;
  c->ev_used = 0;
  c->prof_ms = 0.0;
  c->prof_launches = 0;
  c->prof_each.clear();
  return B3H_OK;
}

namespace {
// Fold the recorded event pairs into the running sums and the per-launch list.
b3h_status profile_collect(b3h_calc *c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < c->ev_used; i++) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev_start[i], c->ev_stop[i]));
    c->prof_ms += ms;
    c->prof_launches++;
    c->prof_each.push_back(ms);
  }
  c->ev_used = 0;
  return B3H_OK;
}
}  // namespace

b3h_status b3h_profile_read(b3h_calc *c, double *kernel_ms, uint64_t *launches) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if ((st = set_device(c))) return st;
  if ((st = profile_collect(c))) return st;
  if (kernel_ms) *kernel_ms = c->prof_ms;
  if (launches) *launches = c->prof_launches;
  return B3H_OK;
}

b3h_status b3h_profile_launch_times(b3h_calc *c, double *ms, uint64_t capacity,
                                    uint64_t *launches) {
  b3h_status st = check_handle(c);
  if (st) return st;
  if (capacity && !ms) return fail(B3H_ERR_ARG, "null output with nonzero capacity");
  if ((st = set_device(c))) return st;
  if ((st = profile_collect(c))) return st;
  const uint64_t n = c->prof_each.size();
  for (uint64_t i = 0; i < n && i < capacity; i++) ms[i] = c->prof_each[i];
  if (launches) *launches = n;
  return B3H_OK;
}

}  // extern "C"
