// Device-side substitution model: FP64 eigensystem of a reversible Q and
// per-edge transition matrices P(t) = V diag(exp(lambda t)) V^-1.
//
// Replaces the host-side eigendecomposition (crates/b3/src/beast_eigen.rs:
// 66-101, called from substitution.rs:544-554) and the per-edge host build
// (transitions.rs:26-48, substitution.rs:403-420,571-588).  Q is reversible
// for every model b3 exports, so S = D^1/2 Q D^-1/2 (D = diag(pi)) is
// symmetric and is diagonalised by parallel cyclic Jacobi rotations.
#include "common.hpp"
#include "kernels.hpp"

namespace b3h {

// Parallel cyclic Jacobi, n <= 64, one 1024-thread workgroup; A and U in
// LDS (2 x 32 KiB).  Indices are paired by the round-robin ("chess
// tournament") ordering: m = n rounded up to even, m - 1 rounds per sweep,
// m / 2 disjoint (p, q) pairs per round.  Disjoint rotations commute, so a
// round applies all its column rotations, then all its row rotations
// (A <- J^T A J, U <- U J), with one barrier between phases.  A pair with a
// padding index or a negligible a_pq is the identity.
__global__ __launch_bounds__(1024) void eigen_kernel(DevModel *m) {
  // rows padded to 65 doubles: a column walk over k spreads over the banks
  constexpr int LD = 65;
  __shared__ double A[64 * LD];
  __shared__ double U[64 * LD];
  __shared__ double rc[32], rs[32];
  __shared__ int rp[32], rq[32];
  __shared__ double red_off[16], red_tot[16];
  __shared__ int rot_any;
  const int n = m->n;
  const int tid = threadIdx.x;
  const int mm = (n + 1) & ~1;
  const int npairs = mm / 2;
  for (int idx = tid; idx < 64 * 64; idx += blockDim.x) {
    const int k = idx >> 6, j = idx & 63;
    double a = 0.0;
    if (k < n && j < n) {
      const double sk = sqrt(m->freqs[k]), sj = sqrt(m->freqs[j]);
      const double skj = m->q[k * n + j] * sk / sj;
      const double sjk = m->q[j * n + k] * sj / sk;
      a = 0.5 * (skj + sjk);
    }
    A[k * LD + j] = a;
    U[k * LD + j] = (k == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 64; sweep++) {
    // convergence test: off-diagonal mass relative to the total
    double off = 0.0, tot = 0.0;
    for (int idx = tid; idx < n * 64; idx += blockDim.x) {
      const int k = idx >> 6, j = idx & 63;
      if (j < n) {
        const double v = A[k * LD + j] * A[k * LD + j];
        tot += v;
        if (j != k) off += v;
      }
    }
    for (int sh = 32; sh > 0; sh >>= 1) {
      off += __shfl_xor(off, sh);
      tot += __shfl_xor(tot, sh);
    }
    if ((tid & 63) == 0) {
      red_off[tid >> 6] = off;
      red_tot[tid >> 6] = tot;
    }
    if (tid == 0) rot_any = 0;
    __syncthreads();
    off = 0.0;
    tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      off += red_off[w];
      tot += red_tot[w];
    }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int round = 0; round < mm - 1; round++) {
      if (tid < npairs) {
        // positions tid and mm - 1 - tid; position 0 is fixed, the others rotate
        auto player = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + round) % (mm - 1); };
        int p = player(tid), q = player(mm - 1 - tid);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, sn = 0.0;
        if (q < n) {
          const double apq = A[p * LD + q];
          const double app = A[p * LD + p], aqq = A[q * LD + q];
          if (!(fabs(apq) <= 1e-18 * (fabs(app) + fabs(aqq)) || apq == 0.0)) {
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            sn = t * c;
            rot_any = 1;
          }
        }
        rp[tid] = p;
        rq[tid] = q < n ? q : p;  // identity pairs touch only p
        rc[tid] = c;
        rs[tid] = sn;
      }
      __syncthreads();
      // columns p, q of A and U: lane k of a 64-lane row walks row k
      for (int idx = tid; idx < npairs * 64; idx += blockDim.x) {
        const int i = idx >> 6, k = idx & 63;
        const double sn = rs[i];
        if (sn == 0.0) continue;
        const double c = rc[i];
        const int p = rp[i], q = rq[i];
        const double akp = A[k * LD + p], akq = A[k * LD + q];
        A[k * LD + p] = c * akp - sn * akq;
        A[k * LD + q] = sn * akp + c * akq;
        const double ukp = U[k * LD + p], ukq = U[k * LD + q];
        U[k * LD + p] = c * ukp - sn * ukq;
        U[k * LD + q] = sn * ukp + c * ukq;
      }
      __syncthreads();
      // rows p, q of A
      for (int idx = tid; idx < npairs * 64; idx += blockDim.x) {
        const int i = idx >> 6, k = idx & 63;
        const double sn = rs[i];
        if (sn == 0.0) continue;
        const double c = rc[i];
        const int p = rp[i], q = rq[i];
        const double apk = A[p * LD + k], aqk = A[q * LD + k];
        A[p * LD + k] = c * apk - sn * aqk;
        A[q * LD + k] = sn * apk + c * aqk;
      }
      __syncthreads();
    }
    // every thread reads the flag before thread 0 may clear it for the next sweep
    const int any = rot_any;
    __syncthreads();
    if (!any) break;
  }
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += blockDim.x) {
    const int k = idx / n, j = idx % n;
    const double sk = sqrt(m->freqs[k]);
    m->evec[k * n + j] = U[k * LD + j] / sk;  // V = D^-1/2 U
    m->ievc[j * n + k] = U[k * LD + j] * sk;  // V^-1 = U^T D^1/2
  }
  if (tid < n) m->eval[tid] = A[tid * LD + tid];
}

// P(t) element (r, col) for an S-state model: the special cases of
// substitution.rs:245-268,408-414 (when enabled), else the eigen product.
__device__ inline double p_element(const DevModel *m, int S, double t, int r,
                                   int col) {
  if (m->special) {
    if (t > 20.0) return m->freqs[col];
    if (t < 1e-10) return (r == col) ? 1.0 : m->q[r * S + col] * t;
  }
  double acc = 0.0;
  for (int k = 0; k < S; k++)
    acc += m->evec[r * S + k] * exp(m->eval[k] * t) * m->ievc[k * S + col];
  return acc;
}

// Build the transition matrices of one propose.  Output layout
// [entry][side][category][S*S] (FP64, row-major) or [entry][side][category]
// [5][4] (FP32, 4 states: columns and an all-ones fifth column) in the
// kernels' entry order; entry_index(meta) maps
// an entry to the caller's list position.  Mode 0 casts host matrices given
// as [K][2n][S*S] f64; mode 1 builds them from edge_times[2n] x rates[K]
// (gamma.rs:126-134: time length x (clock rate x category rate)); mode 2
// from branch lengths [K][2n] formed by the caller, as transitions.rs:33-35
// forms them for a per-edge (relaxed) clock.  With model_stride 1 category c
// uses model[c] (hetero.rs:122-135: one substitution model per category).
template <typename T>
__global__ void pbuild_kernel(const Entry *entries, uint32_t n, uint32_t K,
                              uint32_t S, int mode, const double *mats,
                              const double *times, const double *rates,
                              const DevModel *model, uint32_t model_stride, T *out,
                              const uint4 *copy_src, uint4 *copy_dst, uint32_t copy_units,
                              T *cmats, uint2 *cleaves) {
  // the upload of the staging buffer, when this launch carries it: the
  // inputs above then point into the pinned host staging buffer itself, and
  // the kernels after this one read the device copy
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < copy_units;
       u += gridDim.x * blockDim.x)
    copy_dst[u] = copy_src[u];
  const uint32_t SS = S * S;
  const uint64_t total = (uint64_t)n * 2 * K * SS;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t el = (uint32_t)(g % SS);
    const uint64_t mat = g / SS;  // (entry * 2 + side) * K + c
    const uint32_t c = (uint32_t)(mat % K);
    const uint64_t es = mat / K;
    const uint32_t side = (uint32_t)(es & 1);
    const uint32_t e = (uint32_t)(es >> 1);
    const Entry en = entries[e];
    const uint32_t orig = entry_index(en.meta);
    double v;
    if (mode == 0) {
      v = mats[((uint64_t)c * 2 * n + 2ull * orig + side) * SS + el];
    } else {
      const double t = mode == 1 ? times[2 * orig + side] * rates[c]
                                 : times[(uint64_t)c * 2 * n + 2 * orig + side];
      v = p_element(model + c * model_stride, (int)S, t, (int)(el / S), (int)(el % S));
    }
    if constexpr (sizeof(T) == 4) {
      // FP32 (4 states): columns, the fifth all ones (the unknown code),
      // [entry][side][K][5][4]; a cherry's copy and its leaves are also kept
      // with its slot for later readers
      const uint32_t r = el / 4, col = el % 4;
      T *o = out + mat * 20;
      o[col * 4 + r] = (T)v;
      if (col == 0) o[16 + r] = (T)1;
      if (cmats && (en.left & SRC_TAG) == SRC_TIP && (en.right & SRC_TAG) == SRC_TIP) {
        T *cm = cmats + (((uint64_t)en.out * 2 + side) * K + c) * 20;
        cm[col * 4 + r] = (T)v;
        if (col == 0) cm[16 + r] = (T)1;
        if (side == 0 && c == 0 && el == 0)
          cleaves[en.out] = make_uint2(en.left & SRC_IDX, en.right & SRC_IDX);
      }
    } else {
      out[g] = (T)v;
    }
  }
}

// Batched P(t) build: blockIdx.y selects the instance, whose entries, edge
// times, rates, model and output come from its descriptor.
__global__ void pbuild_batch_kernel(const PbuildDesc *desc, uint32_t K, uint32_t S) {
  const PbuildDesc d = desc[blockIdx.y];
  const uint32_t SS = S * S;
  const uint64_t total = (uint64_t)d.n * 2 * K * SS;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t el = (uint32_t)(g % SS);
    const uint64_t mat = g / SS;
    const uint32_t c = (uint32_t)(mat % K);
    const uint64_t es = mat / K;
    const uint32_t side = (uint32_t)(es & 1);
    const uint32_t e = (uint32_t)(es >> 1);
    const Entry en = d.entries[e];
    const uint32_t orig = entry_index(en.meta);
    const double t = d.times[2 * orig + side] * d.rates[c];
    // columns, the fifth all ones: [entry][side][K][5][4] (S = 4)
    const uint32_t r = el / 4, col = el % 4;
    float *o = d.out + mat * 20;
    const float v = (float)p_element(d.model, (int)S, t, (int)r, (int)col);
    o[col * 4 + r] = v;
    if (col == 0) o[16 + r] = 1.0f;
    if (d.cmats && (en.left & SRC_TAG) == SRC_TIP && (en.right & SRC_TAG) == SRC_TIP) {
      float *cm = d.cmats + (((uint64_t)en.out * 2 + side) * K + c) * 20;
      cm[col * 4 + r] = v;
      if (col == 0) cm[16 + r] = 1.0f;
      if (side == 0 && c == 0 && el == 0)
        d.cleaves[en.out] = make_uint2(en.left & SRC_IDX, en.right & SRC_IDX);
    }
  }
}

hipError_t launch_pbuild_batch(const PbuildDesc *desc, uint32_t count, uint32_t max_n,
                               uint32_t K, uint32_t S, hipStream_t s) {
  const uint64_t per = (uint64_t)max_n * 2 * K * S * S;
  uint64_t bx = (per + 255) / 256;
  if (bx > 1024) bx = 1024;
  if (bx == 0 || count == 0) return hipSuccess;
  hipLaunchKernelGGL(pbuild_batch_kernel, dim3((uint32_t)bx, count), dim3(256), 0, s, desc,
                     K, S);
  return hipGetLastError();
}

// One transition per propagation-list node (its own edge), [m][K][16].
template <typename T>
__global__ void pbuild_prop_kernel(uint32_t nm, uint32_t K, int mode, const double *mats,
                                   const double *times, const double *rates,
                                   const DevModel *model, T *out) {
  const uint64_t total = (uint64_t)nm * K * 16;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t el = (uint32_t)(g % 16);
    const uint64_t mc = g / 16;
    const uint32_t c = (uint32_t)(mc % K);
    const uint32_t m = (uint32_t)(mc / K);
    double v;
    if (mode == 0)
      v = mats[((uint64_t)c * nm + m) * 16 + el];
    else
      v = p_element(model, 4, times[m] * rates[c], (int)(el / 4), (int)(el % 4));
    out[g] = (T)v;
  }
}

template <typename T>
hipError_t launch_pbuild_prop(uint32_t nm, uint32_t K, int mode, const double *mats,
                              const double *times, const double *rates, const DevModel *model,
                              T *out, hipStream_t s) {
  const uint64_t total = (uint64_t)nm * K * 16;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(pbuild_prop_kernel<T>, dim3((uint32_t)blocks), dim3(256), 0, s, nm, K,
                     mode, mats, times, rates, model, out);
  return hipGetLastError();
}

template hipError_t launch_pbuild_prop<float>(uint32_t, uint32_t, int, const double *,
                                              const double *, const double *, const DevModel *,
                                              float *, hipStream_t);
template hipError_t launch_pbuild_prop<double>(uint32_t, uint32_t, int, const double *,
                                               const double *, const double *, const DevModel *,
                                               double *, hipStream_t);

hipError_t launch_eigen(DevModel *m, hipStream_t s) {
  hipLaunchKernelGGL(eigen_kernel, dim3(1), dim3(1024), 0, s, m);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_pbuild(const Entry *entries, uint32_t n, uint32_t K,
                         uint32_t S, int mode, const double *mats,
                         const double *times, const double *rates,
                         const DevModel *model, uint32_t model_stride, T *out,
                         hipStream_t s, const void *copy_src, void *copy_dst,
                         uint32_t copy_units, T *cmats, uint2 *cleaves) {
  const uint64_t total = (uint64_t)n * 2 * K * S * S;
  uint64_t blocks = std::max<uint64_t>((total + 255) / 256, (copy_units + 255u) / 256u);
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(pbuild_kernel<T>, dim3((uint32_t)blocks), dim3(256), 0, s,
                     entries, n, K, S, mode, mats, times, rates, model, model_stride, out,
                     static_cast<const uint4 *>(copy_src), static_cast<uint4 *>(copy_dst),
                     copy_units, cmats, cleaves);
  return hipGetLastError();
}

template hipError_t launch_pbuild<float>(const Entry *, uint32_t, uint32_t,
                                         uint32_t, int, const double *,
                                         const double *, const double *,
                                         const DevModel *, uint32_t, float *,
                                         hipStream_t, const void *, void *, uint32_t, float *,
                                         uint2 *);
template hipError_t launch_pbuild<double>(const Entry *, uint32_t, uint32_t,
                                          uint32_t, int, const double *,
                                          const double *, const double *,
                                          const DevModel *, uint32_t, double *,
                                          hipStream_t, const void *, void *, uint32_t, double *,
                                          uint2 *);

}  // namespace b3h
