// Propagation storage: every node stores the projection P(edge) x of its
// likelihood x instead of x itself (b3's default GPU calculator,
// crates/b3/src/calculator/cuda_propagations/kernels.cu:73-187 and
// mod.rs:77-164).  An update list from DatedTree::propagation_lists
// (tree.rs:436-454) names the updated leaves (projection = the column of the
// leaf's state), the internal nodes whose projections change (children
// first) and the root last.  Each internal node costs one 4x4 product (its
// own edge) where partial storage costs two (its children's edges); in
// exchange leaves hold 16-byte projections instead of 1-byte codes and a
// changed edge recomputes its child node too.
//
// One lane per (pattern, category), blockIdx.y = category.  Two value types:
//  * double: b3's arithmetic -- a node whose four values all fall below
//    e^-scale_ln is multiplied by e^scale_ln and flagged in its slot; a
//    per-pattern sum of scale_ln over flagged slots is kept incrementally
//    (kernels.cu:110-136); the root's flag is cleared (kernels.cu:170-176).
//    Products and dot products round in b3's order (no contraction).
//  * float: the node's likelihood is normalised by the power of two of its
//    largest value and a cumulative int16 exponent travels with the slot, as
//    in the FP32 partial kernel; the root is evaluated in FP64.
// The projection of the entry just computed stays in registers and is used
// directly when the next entry names it as a child.
#include "common.hpp"
#include "kernels.hpp"

namespace b3h {

namespace {

template <typename T>
struct Vec4 {
  T x, y, z, w;
};

template <typename T>
__device__ __forceinline__ Vec4<T> load4(const void *base, size_t idx) {
  if constexpr (sizeof(T) == 8) {
    const double4 v = reinterpret_cast<const double4 *>(base)[idx];
    return {v.x, v.y, v.z, v.w};
  } else {
    const float4 v = reinterpret_cast<const float4 *>(base)[idx];
    return {v.x, v.y, v.z, v.w};
  }
}

template <typename T>
__device__ __forceinline__ void store4(void *base, size_t idx, const Vec4<T> &v) {
  if constexpr (sizeof(T) == 8)
    reinterpret_cast<double4 *>(base)[idx] = make_double4(v.x, v.y, v.z, v.w);
  else
    reinterpret_cast<float4 *>(base)[idx] = make_float4(v.x, v.y, v.z, v.w);
}

// kernels.cu:5-7: a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w
template <typename T>
__device__ __forceinline__ T row_dot(const T *m, const Vec4<T> &v) {
#pragma clang fp contract(off)
  return m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3] * v.w;
}

}  // namespace

template <typename T>
__global__ __launch_bounds__(256) void prune_prop_kernel(PropArgs a) {
#pragma clang fp contract(off)
  const uint32_t c = blockIdx.y;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.P) return;
  const size_t KP = (size_t)a.K * a.P;
  const size_t cp = (size_t)c * a.P + p;
  const T *mats = static_cast<const T *>(a.mats);
  constexpr bool F64 = sizeof(T) == 8;

  // updated leaves: projection = column `code` of P, unknown -> ones
  for (uint32_t i = 0; i < a.nleaves; i++) {
    const Entry e = a.entries[i];
    const uint32_t code = a.codes[(size_t)(e.left & SRC_IDX) * a.cstride + p];
    const T *M = mats + ((size_t)i * a.K + c) * 16;
    Vec4<T> v;
    if (code < 4) {
      v = {M[code], M[4 + code], M[8 + code], M[12 + code]};
    } else {
      v = {(T)1, (T)1, (T)1, (T)1};
    }
    store4<T>(a.proj, (size_t)e.out * KP + cp, v);
    if (!F64) a.exps[(size_t)e.out * KP + cp] = 0;
  }

  uint32_t ss = F64 ? a.scale_sums[cp] : 0u;
  uint32_t prev = 0xFFFFFFFFu;  // slot of the projection held in registers
  Vec4<T> pv{};
  int pe = 0;
  const uint32_t last = a.n - 1;
  for (uint32_t i = a.nleaves; i < last; i++) {
    const Entry e = a.entries[i];
    const Vec4<T> l = (e.left == prev) ? pv : load4<T>(a.proj, (size_t)e.left * KP + cp);
    const Vec4<T> r = (e.right == prev) ? pv : load4<T>(a.proj, (size_t)e.right * KP + cp);
    Vec4<T> x = {l.x * r.x, l.y * r.y, l.z * r.z, l.w * r.w};
    int E = 0;
    if constexpr (F64) {
      const bool scale = x.x < a.threshold && x.y < a.threshold && x.z < a.threshold &&
                         x.w < a.threshold;
      if (scale) {
        x.x *= a.mult;
        x.y *= a.mult;
        x.z *= a.mult;
        x.w *= a.mult;
      }
      const uint8_t old = a.scales[(size_t)(e.out ^ 1u) * KP + cp];
      a.scales[(size_t)e.out * KP + cp] = scale ? 1 : 0;
      if (old == 0 && scale) ss += a.scale_ln;
      else if (old == 1 && !scale) ss -= a.scale_ln;
    } else {
      const int el = (e.left == prev) ? pe : exp_unpack(a.exps[(size_t)e.left * KP + cp]);
      const int er = (e.right == prev) ? pe : exp_unpack(a.exps[(size_t)e.right * KP + cp]);
      const float mx = fmaxf(fmaxf((float)x.x, (float)x.y), fmaxf((float)x.z, (float)x.w));
      const int ex = (mx > 0.0f) ? __builtin_amdgcn_frexp_expf(mx) : 0;
      x.x = (T)__builtin_amdgcn_ldexpf((float)x.x, -ex);
      x.y = (T)__builtin_amdgcn_ldexpf((float)x.y, -ex);
      x.z = (T)__builtin_amdgcn_ldexpf((float)x.z, -ex);
      x.w = (T)__builtin_amdgcn_ldexpf((float)x.w, -ex);
      E = exp_clamp(el + er + ex);
    }
    const T *M = mats + ((size_t)i * a.K + c) * 16;
    const Vec4<T> pr = {row_dot<T>(M, x), row_dot<T>(M + 4, x), row_dot<T>(M + 8, x),
                        row_dot<T>(M + 12, x)};
    store4<T>(a.proj, (size_t)e.out * KP + cp, pr);
    if (!F64) a.exps[(size_t)e.out * KP + cp] = exp_pack(E);
    prev = e.out;
    pv = pr;
    pe = E;
  }

  // root (kernels.cu:155-187)
  const Entry e = a.entries[last];
  const Vec4<T> l = (e.left == prev) ? pv : load4<T>(a.proj, (size_t)e.left * KP + cp);
  const Vec4<T> r = (e.right == prev) ? pv : load4<T>(a.proj, (size_t)e.right * KP + cp);
  double pi0 = a.pi0, pi1 = a.pi1, pi2 = a.pi2, pi3 = a.pi3;
  if (a.pik) {
    pi0 = a.pik[4 * c];
    pi1 = a.pik[4 * c + 1];
    pi2 = a.pik[4 * c + 2];
    pi3 = a.pik[4 * c + 3];
  }
  double site;
  if constexpr (F64) {
    const double x0 = l.x * r.x, x1 = l.y * r.y, x2 = l.z * r.z, x3 = l.w * r.w;
    const double lk = x0 * pi0 + x1 * pi1 + x2 * pi2 + x3 * pi3;
    a.scales[(size_t)e.out * KP + cp] = 0;
    if (a.scales[(size_t)(e.out ^ 1u) * KP + cp]) ss -= a.scale_ln;
    site = log(lk) - (double)ss;
    a.scale_sums[cp] = ss;
  } else {
    const int el = (e.left == prev) ? pe : exp_unpack(a.exps[(size_t)e.left * KP + cp]);
    const int er = (e.right == prev) ? pe : exp_unpack(a.exps[(size_t)e.right * KP + cp]);
    double lk = ((double)l.x * (double)r.x) * pi0;
    lk = fma((double)l.y * (double)r.y, pi1, lk);
    lk = fma((double)l.z * (double)r.z, pi2, lk);
    lk = fma((double)l.w * (double)r.w, pi3, lk);
    site = exp_site(lk, exp_clamp(el + er));
  }
  a.site[cp] = site;
}

hipError_t launch_prune_prop(const PropArgs &a, bool f64, hipStream_t s) {
  dim3 grid((a.P + 255) / 256, a.K);
  if (f64)
    hipLaunchKernelGGL(prune_prop_kernel<double>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(prune_prop_kernel<float>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace b3h
