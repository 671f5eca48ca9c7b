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
// off_t(q) + i * sstr[t] (in elements), where off_t(q) decomposes q over the
// non-solve dimensions (fastest first) with per-array strides.
struct SysDesc {
  long long nsys;
  long long sstr[5];
  long long ostr[5][kMaxDim];
  int osize[kMaxDim];
  int nother;
  int N;
  int small_index;  // 1 when nsys fits 32-bit division
};

// Element offsets of system q in the five arrays, in index type I (uint32_t
// when the host has checked that every array's extent fits, else int64).
// `interior` is set when every non-solve coordinate of the system lies in
// [1, size-2] (the system is a line through the interior of the grid).
template <typename I>
__device__ __forceinline__ void sys_offsets(const SysDesc &D, long long q,
                                            I off[5], bool *interior = nullptr) {
#pragma unroll
  for (int t = 0; t < 5; t++) off[t] = 0;
  bool in = true;
  if (D.small_index) {
    uint32_t qq = (uint32_t)q;
    for (int k = 0; k < D.nother; k++) {
      uint32_t sz = (uint32_t)D.osize[k];
      uint32_t ik = qq % sz;
      qq /= sz;
      in &= ik >= 1 && ik + 2 <= sz;
#pragma unroll
      for (int t = 0; t < 5; t++) off[t] += (I)ik * (I)D.ostr[t][k];
    }
  } else {
    for (int k = 0; k < D.nother; k++) {
      long long sz = D.osize[k];
      long long ik = q % sz;
      q /= sz;
      in &= ik >= 1 && ik + 2 <= sz;
#pragma unroll
      for (int t = 0; t < 5; t++) off[t] += (I)ik * (I)D.ostr[t][k];
    }
  }
  if (interior) *interior = in;
}

// Reciprocal.  Default: the hardware estimate (v_rcp_f64 / v_rcp_f32)
// refined by Newton-Raphson steps y <- y + y(1 - x y) written with fused
// multiply-adds (two steps in FP64, one in FP32), which leaves the result
// within about one ulp of 1/x for finite non-zero x of normal range.  A zero
// or infinite x yields NaN (a singular pivot).  Define TRIDHIP_IEEE_DIV for
// correctly rounded 1/x (the compiler's full division sequence) instead.
__device__ __forceinline__ double rcp(double x) {
#ifdef TRIDHIP_IEEE_DIV
  return 1.0 / x;
#else
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
#endif
}
__device__ __forceinline__ float rcp(float x) {
#ifdef TRIDHIP_IEEE_DIV
  return 1.0f / x;
#else
  float y = __builtin_amdgcn_rcpf(x);
  float e = __builtin_fmaf(-x, y, 1.0f);
  return __builtin_fmaf(y, e, y);
#endif
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

// Value of lane `src` (wave-uniform) broadcast to the whole wave.
__device__ __forceinline__ float lane_bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
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
