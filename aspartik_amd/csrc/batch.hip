// Batched evaluation of many independent instances (include/b3hip.h,
// b3h_batch_*): one persistent pruning launch over all (instance, tile) work
// items, longest trees first, optionally replayed from a captured HIP graph.
//
// Per instance the data layout and semantics are those of a single
// b3h_calc (partials [slot][K][P] float4 + int16 exponents, EditBuf slots,
// device eigensystem and P(t)); all instances live in shared device arenas,
// and one staging buffer carries every instance's update list, edge times,
// rates and kernel arguments to the device in a single copy.
#include "../../include/b3hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "schedule.hpp"

namespace b3h {
b3h_status host_fail(b3h_status s, const std::string &msg);
uint8_t host_remap_code(uint8_t v, uint32_t S);
}  // namespace b3h

using namespace b3h;

namespace {

#define BHIP(expr)                                                             \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return host_fail(B3H_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr size_t kAlign = 256;
constexpr uint32_t kMaxSlots = 32;
constexpr size_t kMaxLds = 160 * 1024 - 256;  // the kernel also has a few static LDS bytes

size_t up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct BInst {
  uint32_t L = 0, I = 0, P = 0, cstride = 0;
  std::vector<uint8_t> editbuf, editbuf_next;
  // per partial slot: 1 while the slot holds a cherry (both children
  // leaves), whose partial is not stored but recomputed by its readers from
  // the two leaves' codes and the matrices recorded for the slot
  std::vector<uint8_t> slot_cherry, cherry_next;
  Scheduler sched;
  // offsets into the device arenas (element units of each arena)
  size_t part_off = 0, code_off = 0, site_off = 0, pat_off = 0, bsum_off = 0, pm_off = 0;
  size_t cm_off = 0, cl_off = 0;  // cherry matrices (floats) and leaves (uint2) of its slots
  // offsets into the staging buffer (bytes)
  size_t st_entries = 0, st_times = 0, st_rates = 0;
  bool model_set = false;
  uint32_t n = 0;  // list length of the last propose (0: not updated)
};

}  // namespace

struct b3h_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t count = 0, K = 0, S = 4, T = 1, PPL = 2;
  bool S1 = false;  // full evaluations with one state per lane (items of 16 T patterns)
  bool single = false;
  std::vector<BInst> inst;
  uint32_t cap = 8;
  uint64_t bytes = 0;

  // device arenas
  float4 *partials = nullptr;
  int16_t *exps = nullptr;
  uint8_t *codes = nullptr;
  double *site = nullptr, *mixed = nullptr;
  uint32_t *weights = nullptr;
  double *blocksum = nullptr;
  uint32_t *counters = nullptr;  // one per instance, 16 words apart
  double *totals = nullptr;
  float *pmats = nullptr;
  float *cmats = nullptr;    // [slots][2][K][5][4] per instance
  uint2 *cleaves = nullptr;  // [slots] per instance
  DevModel *models = nullptr;
  uint32_t *queue = nullptr;

  // staging: [F32Args x count][PbuildDesc x count][items][per-instance lists]
  size_t st_args = 0, st_pdesc = 0, st_items = 0, st_bytes = 0;
  unsigned char *h_stage = nullptr, *d_stage = nullptr;
  hipEvent_t stage_ev = nullptr;
  uint32_t max_items = 0;

  // graph replay
  hipGraphExec_t graph = nullptr;
  std::vector<uint64_t> graph_key;

  bool has_result = false;
  bool profiling = false;
  std::vector<hipEvent_t> ev_start, ev_stop;
  size_t ev_used = 0;
  double prof_ms = 0.0;
  uint64_t prof_launches = 0;
};

namespace {

b3h_status balloc(b3h_batch *b, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return host_fail(B3H_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " +
                                      hipGetErrorString(e));
  }
  b->bytes += bytes;
  return B3H_OK;
}

void batch_release(b3h_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  void *ptrs[] = {b->partials, b->exps,  b->codes,    b->site,   b->mixed,
                  b->weights,  b->blocksum, b->counters, b->totals, b->pmats,
                  b->cmats,    b->cleaves, b->models,  b->queue,  b->d_stage};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->h_stage) (void)hipHostFree(b->h_stage);
  if (b->stage_ev) (void)hipEventDestroy(b->stage_ev);
  for (auto e : b->ev_start) (void)hipEventDestroy(e);
  for (auto e : b->ev_stop) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

uint32_t tiles_for(const b3h_batch *b, const BInst &in, uint32_t ppl, bool s1) {
  const uint32_t tp = s1 ? 16u * b->T : 64u * b->T * ppl;
  return (in.P + tp - 1) / tp;
}

}  // namespace

extern "C" {

b3h_status b3h_batch_create(const b3h_config *cfg, uint32_t count, const b3h_instance *insts,
                            uint32_t K, uint32_t flags, b3h_batch **out) {
  if (!out) return host_fail(B3H_ERR_ARG, "null output handle pointer");
  *out = nullptr;
  b3h_config def;
  b3h_config_default(&def);
  if (!cfg) cfg = &def;
  if (count == 0 || !insts) return host_fail(B3H_ERR_ARG, "empty batch");
  if (K == 0 || K > 16) return host_fail(B3H_ERR_ARG, "num_categories must be 1..16");
  if (cfg->precision != B3H_F32_SCALED || cfg->states != 4)
    return host_fail(B3H_ERR_UNSUPPORTED, "batches use the FP32 4-state path");
  for (uint32_t i = 0; i < count; i++) {
    if (!insts[i].samples || insts[i].num_leaves < 2 || insts[i].num_patterns == 0)
      return host_fail(B3H_ERR_ARG, "bad instance " + std::to_string(i));
    if ((uint64_t)insts[i].num_leaves * 2 >= (1u << 24))
      return host_fail(B3H_ERR_ARG, "too many leaves in instance " + std::to_string(i));
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return host_fail(B3H_ERR_NO_DEVICE, "no HIP device visible");
  }
  if (cfg->device < 0 || cfg->device >= ndev)
    return host_fail(B3H_ERR_NO_DEVICE, "device ordinal out of range");

  b3h_batch *b = new b3h_batch();
  auto bail = [&](b3h_status s) {
    batch_release(b);
    return s;
  };
  b->device = cfg->device;
  b->count = count;
  b->K = K;
  b->single = (flags & B3H_BATCH_SINGLE_SLOT) != 0;
  b->T = std::max<uint32_t>(1u, 8u / K);  // 512-thread workgroups
  while (64u * K * b->T > 512u && b->T > 1) b->T--;
  if (cfg->tile_patterns) b->T = std::max<uint32_t>(1u, std::min(cfg->tile_patterns / 64u, 512u / (64u * K)));
  // two patterns per lane unless asked for one: 256 HKY+Gamma4 instances
  // measured 29.5 ms per batch against 33.9 ms with one (the measured choice
  // per workload is tuning.autotune_batch)
  b->PPL = cfg->lane_patterns == 1 ? 1u : 2u;
  if (cfg->lane_states != 0 && cfg->lane_states != 1 && cfg->lane_states != 4) {
    host_fail(B3H_ERR_ARG, "lane_states must be 0, 1 or 4");
    return bail(B3H_ERR_ARG);
  }
  b->S1 = cfg->lane_states == 1;
  if (hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    host_fail(B3H_ERR_HIP, "stream creation failed");
    return bail(B3H_ERR_HIP);
  }

  // layout
  b->inst.resize(count);
  size_t part = 0, code = 0, site = 0, pat = 0, bsum = 0, pm = 0, st = 0, cm = 0, cl = 0;
  uint32_t max_I = 0;
  const uint32_t tp_max = 64u * b->T * 2u;
  for (uint32_t i = 0; i < count; i++) {
    BInst &in = b->inst[i];
    in.L = insts[i].num_leaves;
    in.I = in.L - 1;
    in.P = insts[i].num_patterns;
    in.cstride = (in.P + tp_max - 1) / tp_max * tp_max;
    in.editbuf.assign(in.I, 0);
    in.part_off = part;
    part += (size_t)(b->single ? 1 : 2) * in.I * K * in.P;
    in.code_off = code;
    code += (size_t)in.L * in.cstride;
    in.site_off = site;
    site += (size_t)K * in.P;
    in.pat_off = pat;
    pat += in.P;
    in.bsum_off = bsum;
    bsum += b->S1 ? (in.P + 16u * b->T - 1) / (16u * b->T) : (in.P + 64u * b->T - 1) / (64u * b->T);
    in.pm_off = pm;
    pm += (size_t)in.I * 2 * K * 20;  // [entry][side][K][5][4]: columns, the fifth all ones
    const size_t slots = (size_t)(b->single ? 1 : 2) * in.I;
    in.slot_cherry.assign(slots, 0);
    in.cm_off = cm;
    cm += slots * 2 * K * 20;
    in.cl_off = cl;
    cl += slots;
    max_I = std::max(max_I, in.I);
  }
  b->max_items = (uint32_t)bsum;
  b->st_args = 0;
  b->st_pdesc = up(sizeof(F32Args) * count);
  b->st_items = b->st_pdesc + up(sizeof(PbuildDesc) * count);
  st = b->st_items + up(sizeof(uint2) * b->max_items);
  for (uint32_t i = 0; i < count; i++) {
    BInst &in = b->inst[i];
    in.st_entries = st;
    st += up((size_t)in.I * sizeof(Entry));
    in.st_times = st;
    st += up((size_t)2 * in.I * 8);
    in.st_rates = st;
    st += up((size_t)K * 8);
  }
  b->st_bytes = st;

  // LDS stack capacity shared by all instances
  auto lds_need = [&](uint32_t cap) {
    const size_t four = prune_f32_lds_bytes(K, b->T, cap, max_I, 2);
    return b->S1 ? std::max(four, prune_f32_s1_lds_bytes(K, b->T, cap, max_I)) : four;
  };
  uint32_t cap = kMaxSlots;
  while (cap > 1 && lds_need(cap) > kMaxLds) cap--;
  if (lds_need(cap) > kMaxLds) {
    host_fail(B3H_ERR_UNSUPPORTED, "tree too large for the LDS-staged update list");
    return bail(B3H_ERR_UNSUPPORTED);
  }
  b->cap = cap;

  b3h_status s;
  if ((s = balloc(b, (void **)&b->partials, part * 16))) return bail(s);
  if ((s = balloc(b, (void **)&b->exps, part * 2))) return bail(s);
  if ((s = balloc(b, (void **)&b->codes, code))) return bail(s);
  if ((s = balloc(b, (void **)&b->site, site * 8))) return bail(s);
  if ((s = balloc(b, (void **)&b->mixed, pat * 8))) return bail(s);
  if ((s = balloc(b, (void **)&b->weights, pat * 4))) return bail(s);
  if ((s = balloc(b, (void **)&b->blocksum, bsum * 8))) return bail(s);
  if ((s = balloc(b, (void **)&b->counters, (size_t)count * 64))) return bail(s);
  if ((s = balloc(b, (void **)&b->totals, (size_t)count * 8))) return bail(s);
  if ((s = balloc(b, (void **)&b->pmats, pm * 4))) return bail(s);
  if ((s = balloc(b, (void **)&b->cmats, cm * 4))) return bail(s);
  if ((s = balloc(b, (void **)&b->cleaves, cl * 8))) return bail(s);
  if ((s = balloc(b, (void **)&b->models, (size_t)count * sizeof(DevModel)))) return bail(s);
  if ((s = balloc(b, (void **)&b->queue, 64))) return bail(s);
  if ((s = balloc(b, (void **)&b->d_stage, b->st_bytes))) return bail(s);
  if (hipHostMalloc((void **)&b->h_stage, b->st_bytes, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&b->stage_ev, hipEventDisableTiming) != hipSuccess) {
    host_fail(B3H_ERR_OOM, "pinned staging allocation failed");
    return bail(B3H_ERR_OOM);
  }
  std::memset(b->h_stage, 0, b->st_bytes);

  // upload codes and weights, instance by instance
  bool ok = true;
  std::vector<uint8_t> codes;
  std::vector<uint32_t> w;
  for (uint32_t i = 0; i < count && ok; i++) {
    const BInst &in = b->inst[i];
    codes.assign((size_t)in.L * in.cstride, 4);
    for (size_t l = 0; l < in.L; l++)
      for (size_t p = 0; p < in.P; p++)
        codes[l * in.cstride + p] = host_remap_code(insts[i].samples[l * in.P + p], 4);
    ok = hipMemcpy(b->codes + in.code_off, codes.data(), codes.size(), hipMemcpyHostToDevice) ==
         hipSuccess;
    if (insts[i].weights) w.assign(insts[i].weights, insts[i].weights + in.P);
    else w.assign(in.P, 1u);
    ok = ok && hipMemcpy(b->weights + in.pat_off, w.data(), (size_t)in.P * 4,
                         hipMemcpyHostToDevice) == hipSuccess;
  }
  ok = ok && hipMemset(b->counters, 0, (size_t)count * 64) == hipSuccess &&
       hipMemset(b->queue, 0, 64) == hipSuccess &&
       hipMemset(b->exps, 0, part * 2) == hipSuccess &&
       hipMemset(b->partials, 0, part * 16) == hipSuccess &&
       hipMemset(b->cmats, 0, cm * 4) == hipSuccess &&
       hipMemset(b->cleaves, 0, cl * 8) == hipSuccess &&
       hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    host_fail(B3H_ERR_HIP, "initial device upload failed");
    return bail(B3H_ERR_HIP);
  }
  if (prepare_prune_f32_batch((int)kMaxLds) != hipSuccess) {
    host_fail(B3H_ERR_HIP, "cannot raise the kernel's LDS limit");
    return bail(B3H_ERR_HIP);
  }
  *out = b;
  return B3H_OK;
}

b3h_status b3h_batch_destroy(b3h_batch *b) {
  if (!b) return B3H_OK;
  (void)hipSetDevice(b->device);
  batch_release(b);
  return B3H_OK;
}

uint64_t b3h_batch_device_bytes(const b3h_batch *b) { return b ? b->bytes : 0; }

b3h_status b3h_batch_set_model(b3h_batch *b, uint32_t i, const double *q, const double *freqs,
                               uint32_t flags) {
  if (!b) return host_fail(B3H_ERR_ARG, "null batch");
  if (i >= b->count) return host_fail(B3H_ERR_ARG, "instance out of range");
  if (!q || !freqs) return host_fail(B3H_ERR_ARG, "null model arrays");
  for (uint32_t k = 0; k < 4; k++)
    if (!(freqs[k] > 0.0)) return host_fail(B3H_ERR_ARG, "frequencies must be positive");
  BHIP(hipSetDevice(b->device));
  DevModel hm;
  std::memset(&hm, 0, sizeof hm);
  hm.n = 4;
  hm.special = (flags & B3H_MODEL_SPECIAL_CASES) ? 1 : 0;
  for (int k = 0; k < 4; k++) hm.freqs[k] = freqs[k];
  for (int k = 0; k < 16; k++) hm.q[k] = q[k];
  BHIP(hipStreamSynchronize(b->stream));
  BHIP(hipMemcpyAsync(b->models + i, &hm, sizeof hm, hipMemcpyHostToDevice, b->stream));
  BHIP(launch_eigen(b->models + i, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  b->inst[i].model_set = true;
  // frequencies for the root go into the instance's kernel arguments
  F32Args *args = reinterpret_cast<F32Args *>(b->h_stage + b->st_args);
  args[i].pi0 = freqs[0];
  args[i].pi1 = freqs[1];
  args[i].pi2 = freqs[2];
  args[i].pi3 = freqs[3];
  return B3H_OK;
}

b3h_status b3h_batch_propose_times(b3h_batch *b, const b3h_update *ups, uint32_t flags) {
  if (!b) return host_fail(B3H_ERR_ARG, "null batch");
  if (!ups) return host_fail(B3H_ERR_ARG, "null updates");
  BHIP(hipSetDevice(b->device));
  const uint32_t K = b->K;
  // the previous copy out of the pinned staging buffer must be complete
  BHIP(hipEventSynchronize(b->stage_ev));
  F32Args *args = reinterpret_cast<F32Args *>(b->h_stage + b->st_args);
  PbuildDesc *pd = reinterpret_cast<PbuildDesc *>(b->h_stage + b->st_pdesc);
  uint32_t max_n = 0, max_slots = 1, hbm_lists = 0, active = 0;
  // validate and schedule every updated instance first; commit only if all pass
  for (uint32_t i = 0; i < b->count; i++) {
    BInst &in = b->inst[i];
    const b3h_update &u = ups[i];
    if (u.n == 0) continue;
    if (!in.model_set)
      return host_fail(B3H_ERR_STATE, "instance " + std::to_string(i) + " has no model");
    if (u.n > in.I || !u.internals || !u.children || !u.edge_times || !u.category_rates)
      return host_fail(B3H_ERR_ARG, "bad update for instance " + std::to_string(i));
    const uint32_t N = 2 * in.L - 1;
    for (uint32_t j = 0; j < u.n; j++) {
      if (u.internals[j] < in.L || u.internals[j] >= N)
        return host_fail(B3H_ERR_ARG, "internal node id out of range in instance " + std::to_string(i));
      for (int s = 0; s < 2; s++)
        if (u.children[2 * j + s] >= N || u.children[2 * j + s] == u.internals[j])
          return host_fail(B3H_ERR_ARG, "child node id out of range in instance " + std::to_string(i));
    }
    const uint8_t *eb = nullptr;
    if (!b->single) {
      in.editbuf_next = in.editbuf;
      for (uint32_t j = 0; j < u.n; j++) {
        uint8_t &m = in.editbuf_next[u.internals[j] - in.L];
        m = (uint8_t)(((m & 1u) ^ (uint8_t)~(m >> 1)) & 3u);
      }
      eb = in.editbuf_next.data();
    }
    // the new slot of every listed node holds a cherry iff both its children
    // are leaves (the schedule build clears the flag of a cherry it has to
    // store); the flags are committed with the EditBuf below
    in.cherry_next = in.slot_cherry;
    for (uint32_t j = 0; j < u.n; j++)
      in.cherry_next[Scheduler::slot_of(u.internals[j], in.L, eb)] =
          (u.children[2 * j] < in.L && u.children[2 * j + 1] < in.L) ? 1 : 0;
    std::string err;
    if (!in.sched.build(u.n, u.internals, u.children, in.L, eb, true, b->cap, &err,
                        in.cherry_next.data()))
      return host_fail(B3H_ERR_ARG, "instance " + std::to_string(i) + ": " + err);
    max_n = std::max(max_n, u.n);
    max_slots = std::max(max_slots, in.sched.slots_used);
    // stored cherries are recomputed only by the HBM-prefetch variant
    hbm_lists += in.sched.hbm_sources * 8 > u.n || in.sched.cherry_sources > 0;
    active++;
  }
  if (active == 0) return host_fail(B3H_ERR_ARG, "no instance has an update");
  const bool hbm = hbm_lists > 0;
  const uint32_t ppl = hbm ? 1u : b->PPL;
  const bool s1 = b->S1 && !hbm;
  // commit: EditBuf flips, staging contents, arguments and the work list
  std::vector<uint32_t> order;
  for (uint32_t i = 0; i < b->count; i++) {
    BInst &in = b->inst[i];
    const b3h_update &u = ups[i];
    in.n = u.n;
    pd[i].n = u.n;
    if (u.n == 0) continue;
    order.push_back(i);
    if (!b->single) in.editbuf.swap(in.editbuf_next);
    in.slot_cherry.swap(in.cherry_next);
    std::memcpy(b->h_stage + in.st_entries, in.sched.entries.data(), (size_t)u.n * sizeof(Entry));
    std::memcpy(b->h_stage + in.st_times, u.edge_times, (size_t)2 * u.n * 8);
    std::memcpy(b->h_stage + in.st_rates, u.category_rates, (size_t)K * 8);
    const Entry *d_entries = reinterpret_cast<const Entry *>(b->d_stage + in.st_entries);
    pd[i].entries = d_entries;
    pd[i].times = reinterpret_cast<const double *>(b->d_stage + in.st_times);
    pd[i].rates = reinterpret_cast<const double *>(b->d_stage + in.st_rates);
    pd[i].model = b->models + i;
    pd[i].out = b->pmats + in.pm_off;
    pd[i].cmats = b->cmats + in.cm_off;
    pd[i].cleaves = b->cleaves + in.cl_off;
    F32Args &a = args[i];
    a.entries = d_entries;
    a.pmats = b->pmats + in.pm_off;
    a.codes32 = reinterpret_cast<const uint32_t *>(b->codes + in.code_off);
    a.partials = b->partials + in.part_off;
    a.exps = b->exps + in.part_off;
    a.site = b->site + in.site_off;
    a.mixed = b->mixed + in.pat_off;
    a.weights = b->weights + in.pat_off;
    a.blocksum = b->blocksum + in.bsum_off;
    a.counter = b->counters + (size_t)i * 16;
    a.total = b->totals + i;
    a.host_total = nullptr;
    a.host_flag = nullptr;
    a.seq = 0;
    a.pik = nullptr;
    a.classes = nullptr;
    a.cmats = b->cmats + in.cm_off;
    a.cleaves = b->cleaves + in.cl_off;
    a.ln_inv_k = std::log(1.0 / (double)K);
    a.n = u.n;
    a.P = in.P;
    a.K = K;
    a.T = b->T;
    a.nslots = std::max<uint32_t>(max_slots, 1u);
    a.cstride = in.cstride;
    a.ntiles = tiles_for(b, in, ppl, s1);
  }
  // longest update lists first (their tiles take longest)
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    return b->inst[x].n > b->inst[y].n;
  });
  uint2 *items = reinterpret_cast<uint2 *>(b->h_stage + b->st_items);
  uint32_t nitems = 0;
  for (uint32_t i : order) {
    const uint32_t nt = args[i].ntiles;
    for (uint32_t t = 0; t < nt; t++) items[nitems++] = make_uint2(i, t);
  }
  const size_t lds =
      s1 ? prune_f32_s1_lds_bytes(K, b->T, std::max<uint32_t>(max_slots, 1u), max_n)
         : prune_f32_lds_bytes(K, b->T, std::max<uint32_t>(max_slots, 1u), max_n, ppl);
  const uint32_t threads = 64u * K * b->T;
  uint32_t grid = 0;
  BHIP(batch_grid(ppl, hbm, s1, threads, lds, &grid));
  grid = std::min(grid, nitems);

  const F32Args *d_args = reinterpret_cast<const F32Args *>(b->d_stage + b->st_args);
  const PbuildDesc *d_pd = reinterpret_cast<const PbuildDesc *>(b->d_stage + b->st_pdesc);
  const uint2 *d_items = reinterpret_cast<const uint2 *>(b->d_stage + b->st_items);
  auto enqueue = [&](hipEvent_t e0, hipEvent_t e1) -> b3h_status {
    BHIP(hipMemcpyAsync(b->d_stage, b->h_stage, b->st_bytes, hipMemcpyHostToDevice, b->stream));
    BHIP(launch_pbuild_batch(d_pd, b->count, max_n, K, 4, b->stream));
    if (e0) BHIP(hipEventRecord(e0, b->stream));
    BHIP(launch_prune_f32_batch(d_args, d_items, nitems, b->queue, ppl, hbm, s1, threads, lds,
                                grid, b->stream));
    if (e1) BHIP(hipEventRecord(e1, b->stream));
    return B3H_OK;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (b->profiling) {
    if (b->ev_used == b->ev_start.size()) {
      hipEvent_t x, y;
      BHIP(hipEventCreate(&x));
      BHIP(hipEventCreate(&y));
      b->ev_start.push_back(x);
      b->ev_stop.push_back(y);
    }
    e0 = b->ev_start[b->ev_used];
    e1 = b->ev_stop[b->ev_used];
    b->ev_used++;
  }
  b3h_status st;
  if ((flags & B3H_BATCH_GRAPH) && !b->profiling) {
    std::vector<uint64_t> key = {ppl, hbm, s1, lds, grid, nitems, max_n};
    if (!b->graph || key != b->graph_key) {
      if (b->graph) {
        BHIP(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
      }
      hipGraph_t g = nullptr;
      BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      st = enqueue(nullptr, nullptr);
      hipError_t ce = hipStreamEndCapture(b->stream, &g);
      if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (ce != hipSuccess) return host_fail(B3H_ERR_HIP, "graph capture failed");
      hipError_t ie = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        b->graph = nullptr;
        return host_fail(B3H_ERR_HIP, "graph instantiation failed");
      }
      b->graph_key = key;
    }
    BHIP(hipGraphLaunch(b->graph, b->stream));
  } else {
    if ((st = enqueue(e0, e1))) return st;
  }
  BHIP(hipEventRecord(b->stage_ev, b->stream));
  b->has_result = true;
  return B3H_OK;
}

b3h_status b3h_batch_totals(b3h_batch *b, double *out) {
  if (!b || !out) return host_fail(B3H_ERR_ARG, "null argument");
  if (!b->has_result) return host_fail(B3H_ERR_STATE, "no propose yet");
  BHIP(hipSetDevice(b->device));
  BHIP(hipMemcpyAsync(out, b->totals, (size_t)b->count * 8, hipMemcpyDeviceToHost, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  return B3H_OK;
}

b3h_status b3h_batch_likelihoods(b3h_batch *b, uint32_t i, uint32_t c, double *out) {
  if (!b || !out) return host_fail(B3H_ERR_ARG, "null argument");
  if (i >= b->count || c >= b->K) return host_fail(B3H_ERR_ARG, "index out of range");
  if (!b->has_result) return host_fail(B3H_ERR_STATE, "no propose yet");
  BHIP(hipSetDevice(b->device));
  const BInst &in = b->inst[i];
  BHIP(hipMemcpyAsync(out, b->site + in.site_off + (size_t)c * in.P, (size_t)in.P * 8,
                      hipMemcpyDeviceToHost, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  return B3H_OK;
}

b3h_status b3h_batch_accept(b3h_batch *b) {
  if (!b) return host_fail(B3H_ERR_ARG, "null batch");
  for (auto &in : b->inst)
    for (auto &m : in.editbuf) m &= 1u;
  return B3H_OK;
}

b3h_status b3h_batch_reject(b3h_batch *b) {
  if (!b) return host_fail(B3H_ERR_ARG, "null batch");
  if (b->single) return host_fail(B3H_ERR_UNSUPPORTED, "single-slot batches cannot reject");
  for (auto &in : b->inst)
    for (auto &m : in.editbuf) m = (uint8_t)((m ^ (m >> 1)) & 1u);
  return B3H_OK;
}

b3h_status b3h_batch_profile_enable(b3h_batch *b, int enable) {
  if (!b) return host_fail(B3H_ERR_ARG, "null batch");
  BHIP(hipSetDevice(b->device));
  BHIP(hipStreamSynchronize(b->stream));
  b->profiling = enable != 0;
  b->ev_used = 0;
  b->prof_ms = 0.0;
  b->prof_launches = 0;
  return B3H_OK;
}

b3h_status b3h_batch_profile_read(b3h_batch *b, double *kernel_ms, uint64_t *launches) {
  if (!b) return host_fail(B3H_ERR_ARG, "null batch");
  BHIP(hipSetDevice(b->device));
  BHIP(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->ev_used; i++) {
    float ms = 0.f;
    BHIP(hipEventElapsedTime(&ms, b->ev_start[i], b->ev_stop[i]));
    b->prof_ms += ms;
    b->prof_launches++;
  }
  b->ev_used = 0;
  if (kernel_ms) *kernel_ms = b->prof_ms;
  if (launches) *launches = b->prof_launches;
  return B3H_OK;
}

}  // extern "C"
