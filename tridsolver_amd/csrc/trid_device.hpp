// trid_device.hpp -- device building blocks shared by the tridsolver_amd
// kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tridhip {

constexpr int kMaxDim = 8;
constexpr int kWave = 64;

// Description of a strided batch: every system has length N; system q's
// element i of array t (0=a, 1=b, 2=c, 3=d, 4=u) sits at
// off_t(q) + i * sstr[t], where off_t(q) decomposes q over the non-solve
// dimensions (fastest first) with per-array strides.
struct SysDesc {
  long long nsys;
  long long sstr[5];
  long long ostr[5][kMaxDim];
  int osize[kMaxDim];
  int nother;
  int N;
  int small_index;  // 1 when nsys and every osize fit 32-bit division
};

__device__ __forceinline__ void sys_offsets(const SysDesc &D, long long q,
                                            long long off[5]) {
#pragma unroll
  for (int t = 0; t < 5; t++) off[t] = 0;
  if (D.small_index) {
    uint32_t qq = (uint32_t)q;
    for (int k = 0; k < D.nother; k++) {
      uint32_t sz = (uint32_t)D.osize[k];
      uint32_t ik = qq % sz;
      qq /= sz;
#pragma unroll
      for (int t = 0; t < 5; t++) off[t] += (long long)ik * D.ostr[t][k];
    }
  } else {
    for (int k = 0; k < D.nother; k++) {
      long long sz = D.osize[k];
      long long ik = q % sz;
      q /= sz;
#pragma unroll
      for (int t = 0; t < 5; t++) off[t] += ik * D.ostr[t][k];
    }
  }
}

// IEEE reciprocal (correctly rounded division).
template <typename T>
__device__ __forceinline__ T rcp(T x) {
  return T(1) / x;
}

// Lane permute through the LDS crossbar (ds_bpermute): returns v of lane
// `src` (0..63) of the current wave.
__device__ __forceinline__ float lane_read(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_ds_bpermute(src << 2, lo);
  hi = __builtin_amdgcn_ds_bpermute(src << 2, hi);
  return __hiloint2double(hi, lo);
}

// One normalised tridiagonal row (diagonal 1): a x_{-} + x + c x_{+} = d.
template <typename T>
struct Row {
  T a, c, d;
};

// Parallel cyclic reduction update of row r with its neighbours m (above)
// and p (below) at the current distance; a missing neighbour is the zero row.
template <typename T>
__device__ __forceinline__ Row<T> pcr_update(const Row<T> &r, const Row<T> &m,
                                             const Row<T> &p) {
  T bn = T(1) - r.a * m.c - r.c * p.a;
  T inv = rcp(bn);
  Row<T> o;
  o.a = -r.a * m.a * inv;
  o.c = -r.c * p.c * inv;
  o.d = (r.d - r.a * m.d - r.c * p.d) * inv;
  return o;
}

}  // namespace tridhip
