// trid_adjoint.hip -- adjoint of the strided-batch solves
// (tridXmtsvStridedBatchAdjoint of include/trid_adjoint.h).
//
// The transposed solve reuses the partitioned Thomas kernels with the
// CoefTransposed coefficient loader of trid_part.hpp; the coefficient
// gradients are one elementwise kernel (one thread per element, no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trid_adjoint.h"
#include "../../include/trid_hip_ext.h"
#include "trid_launch.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

// grad_a_i = -lam_i x_{i-1}, grad_b_i = -lam_i x_i, grad_c_i = -lam_i x_{i+1}
// over a strided batch (one thread per element; rows along solvedim).
template <typename T>
__global__ void __launch_bounds__(256)
    k_adj_grads(const T *__restrict__ lam, const T *__restrict__ x,
                T *__restrict__ ga, T *__restrict__ gb, T *__restrict__ gc,
                SysDesc D, long long total) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  // e enumerates systems (q) and rows (i): e = q * N + i
  const long long q = e / D.N;
  const int i = (int)(e - q * D.N);
  long long off[5];
  sys_offsets<long long>(D, q, off);
  const long long s = D.sstr[0];
  const long long p = off[0] + (long long)i * s;
  const T l = lam[p];
  gb[p] = -l * x[p];
  ga[p] = i > 0 ? -l * x[p - s] : T(0);
  gc[p] = i + 1 < D.N ? -l * x[p + s] : T(0);
}

}  // namespace tridhip

using namespace tridhip;

namespace {

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

template <typename T>
tridStatus_t strided_adjoint(const T *a, const T *b, const T *c, const T *x, const T *g,
                             T *ga, T *gb, T *gc, T *gd, int ndim, int solvedim,
                             const int *dims, const int *pads, int sync) {
  clear_error();
  if (!dims) return fail(TRID_STATUS_INVALID_VALUE, "dims is NULL");
  if (ndim < 1 || ndim > MAXDIM) return fail(TRID_STATUS_INVALID_VALUE, "ndim must be in [1, 8]");
  if (solvedim < 0 || solvedim >= ndim)
    return fail(TRID_STATUS_INVALID_VALUE, "solvedim must be in [0, ndim)");
  if (!pads) pads = dims;
  SysDesc D;
  memset(&D, 0, sizeof D);
  D.N = dims[solvedim];
  D.nsys = 1;
  long long cum = 1, ext = 1;
  for (int k = 0; k < ndim; k++) {
    if (dims[k] < 0 || pads[k] < dims[k])
      return fail(TRID_STATUS_INVALID_VALUE, "bad dims/pads");
    if (k == solvedim) {
      for (int t = 0; t < 5; t++) D.sstr[t] = cum;
    } else {
      D.osize[D.nother] = dims[k];
      for (int t = 0; t < 5; t++) D.ostr[t][D.nother] = cum;
      D.nother++;
      D.nsys *= dims[k];
    }
    ext += (long long)(dims[k] - 1) * cum;
    cum *= pads[k];
  }
  if (D.N == 0 || D.nsys == 0) return TRID_STATUS_SUCCESS;
  if (!a || !b || !c || !x || !g || !ga || !gb || !gc || !gd)
    return fail(TRID_STATUS_INVALID_VALUE, "NULL array");
  D.small_index = D.nsys < (1LL << 31) && !wide_index();
  hipStream_t st = current_stream();
  const int order = solvedim == 0 ? 0 : 1;
  PartCfg cfg = heuristic(order, D.N, 0, sizeof(T));
  if (cfg.kind != 1) return fail(TRID_STATUS_INVALID_VALUE, "system too long for the adjoint kernel");
  if (order == 1) {
    // the strided batches' layout (most systems per work-group)
    const PartCfg s = strided_default(D.N, sizeof(T));
    if (s.kind == 1) cfg = s;
  }
  cfg.vec = 0;
  cfg.idx32 = (unsigned long long)ext * sizeof(T) < (1ULL << 32) && !wide_index();
  // lambda = A^T^{-1} g, written to grad_d (g itself is not modified)
  const CoefTransposed<T> cf{a, b, c};
  tridStatus_t s = check_hip(launch_part<T, 0>(cfg, cf, const_cast<T *>(g), nullptr, D, st, gd),
                             "adjoint solve launch");
  if (s) return s;
  const long long total = (long long)D.N * D.nsys;
  hipLaunchKernelGGL(k_adj_grads<T>, dim3(blocks_for(total)), dim3(256), 0, st, gd, x, ga, gb,
                     gc, D, total);
  s = check_hip(hipGetLastError(), "adjoint gradient launch");
  if (s) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

}  // namespace

EXTERN_C tridStatus_t tridDmtsvStridedBatchAdjoint(
    const double *a, const double *b, const double *c, const double *x,
    const double *g, double *grad_a, double *grad_b, double *grad_c,
    double *grad_d, int ndim, int solvedim, int *dims, int *pads, int sync) {
  return strided_adjoint<double>(a, b, c, x, g, grad_a, grad_b, grad_c, grad_d, ndim,
                                 solvedim, dims, pads, sync);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchAdjoint(
    const float *a, const float *b, const float *c, const float *x,
    const float *g, float *grad_a, float *grad_b, float *grad_c, float *grad_d,
    int ndim, int solvedim, int *dims, int *pads, int sync) {
  return strided_adjoint<float>(a, b, c, x, g, grad_a, grad_b, grad_c, grad_d, ndim,
                                solvedim, dims, pads, sync);
}
