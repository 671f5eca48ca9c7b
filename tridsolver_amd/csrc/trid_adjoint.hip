// trid_adjoint.hip -- adjoints of the batched solves and the differentiable
// variable-diffusivity ADI (include/trid_adjoint.h).
//
// The transposed solves reuse the partitioned Thomas kernels with the
// CoefTransposed / CoefKappa coefficient loaders of trid_part.hpp; the
// gradient formulas are elementwise stencil kernels written as gathers (each
// thread owns one output element), so no atomics are needed.
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

// ---- variable-diffusivity ADI -------------------------------------------

struct Grid {
  int nx, ny, nz;
  long long nxy, n;
};

__device__ __forceinline__ bool grid_interior(const Grid &g, int i, int j, int k) {
  return i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1 && k > 0 && k < g.nz - 1;
}

// r = lam * sum_axes [kf+(u+ - u) - kf-(u - u-)] at interior points, 0 else
template <typename T>
__global__ void __launch_bounds__(256)
    k_vadi_rhs(const T *__restrict__ K, const T *__restrict__ u, T *__restrict__ r,
               Grid g, T lam) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n) return;
  const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / g.nxy);
  if (!grid_interior(g, i, j, k)) {
    r[p] = T(0);
    return;
  }
  const long long st[3] = {1, g.nx, g.nxy};
  T s = T(0);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const long long q = st[a];
    const T kp = T(0.5) * (K[p] + K[p + q]);
    const T km = T(0.5) * (K[p] + K[p - q]);
    s += kp * (u[p + q] - u[p]) - km * (u[p] - u[p - q]);
  }
  r[p] = lam * s;
}

// kappa-gradient of one sweep along `axis` with solution x and adjoint
// solution l, gathered per point: every interior point j contributes
// gm_j = -lam l_j (x_j - x_{j-1}) and gp_j = -lam l_j (x_j - x_{j+1}),
// split half to itself and half to the neighbour the face is shared with.
template <typename T>
__global__ void __launch_bounds__(256)
    k_vadi_coef_grad(T *__restrict__ gk, const T *__restrict__ l,
                     const T *__restrict__ x, Grid g, int axis, T lam) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n) return;
  int c[3] = {(int)(p % g.nx), (int)((p / g.nx) % g.ny), (int)(p / g.nxy)};
  const int size[3] = {g.nx, g.ny, g.nz};
  const long long q = axis == 0 ? 1 : (axis == 1 ? g.nx : g.nxy);
  const int t = c[axis];
  T acc = T(0);
  if (grid_interior(g, c[0], c[1], c[2])) {
    const T gm = lam * (-l[p] * (x[p] - x[p - q]));
    const T gp = lam * (-l[p] * (x[p] - x[p + q]));
    acc += T(0.5) * gm + T(0.5) * gp;
  }
  if (t + 1 < size[axis]) {  // neighbour above: its gm face is shared with p
    c[axis] = t + 1;
    if (grid_interior(g, c[0], c[1], c[2])) {
      const long long pn = p + q;
      acc += T(0.5) * (lam * (-l[pn] * (x[pn] - x[p])));
    }
  }
  if (t >= 1) {  // neighbour below: its gp face is shared with p
    c[axis] = t - 1;
    if (grid_interior(g, c[0], c[1], c[2])) {
      const long long pn = p - q;
      acc += T(0.5) * (lam * (-l[pn] * (x[pn] - x[p])));
    }
  }
  gk[p] += acc;
}

// ubar += D(kappa)^T rbar and gk += (d(D(kappa) u)/d kappa)^T rbar, per point
// as a gather over the point's six faces.  Face f = (a, b = a + q):
// G_f = lam (rbar_a [a interior] - rbar_b [b interior]);
// ubar_b += G kf, ubar_a -= G kf, gk_a += G (u_b - u_a)/2, gk_b += same.
template <typename T>
__global__ void __launch_bounds__(256)
    k_vadi_rhs_adj(const T *__restrict__ K, const T *__restrict__ u,
                   const T *__restrict__ rbar, T *__restrict__ ubar,
                   T *__restrict__ gk, Grid g, T lam) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n) return;
  int c[3] = {(int)(p % g.nx), (int)((p / g.nx) % g.ny), (int)(p / g.nxy)};
  const int size[3] = {g.nx, g.ny, g.nz};
  const long long st[3] = {1, g.nx, g.nxy};
  const bool in_p = grid_interior(g, c[0], c[1], c[2]);
  const T rp = in_p ? rbar[p] : T(0);
  T du = T(0), dk = T(0);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const long long q = st[a];
    const int t = c[a];
    if (t >= 1) {  // face (p - q, p): p is its upper end b
      c[a] = t - 1;
      const T ra = grid_interior(g, c[0], c[1], c[2]) ? rbar[p - q] : T(0);
      c[a] = t;
      const T G = lam * (ra - rp);
      const T kf = T(0.5) * (K[p] + K[p - q]);
      du += G * kf;
      dk += T(0.5) * G * (u[p] - u[p - q]);
    }
    if (t + 1 < size[a]) {  // face (p, p + q): p is its lower end a
      c[a] = t + 1;
      const T rb = grid_interior(g, c[0], c[1], c[2]) ? rbar[p + q] : T(0);
      c[a] = t;
      const T G = lam * (rp - rb);
      const T kf = T(0.5) * (K[p] + K[p + q]);
      du -= G * kf;
      dk += T(0.5) * G * (u[p + q] - u[p]);
    }
  }
  ubar[p] += du;
  gk[p] += dk;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_axpy(T *__restrict__ y, const T *__restrict__ x, long long n) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) y[p] += x[p];
}

// ubar = u - target (ubar may alias u); partial sums of ubar^2 per block
template <typename T>
__global__ void __launch_bounds__(256)
    k_residual(const T *u, const T *__restrict__ t, T *ub, T *__restrict__ partial,
               long long n) {
  __shared__ T sh[256];
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  T v = T(0);
  if (p < n) {
    const T e = u[p] - t[p];
    ub[p] = e;
    v = e * e;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

template <typename T>
__global__ void __launch_bounds__(256) k_sum(const T *__restrict__ in, T *__restrict__ out,
                                            long long n) {
  __shared__ T sh[256];
  T v = T(0);
  for (long long p = threadIdx.x; p < n; p += blockDim.x) v += in[p];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace tridhip

using namespace tridhip;

namespace {

SysDesc line_desc(int nx, int ny, int nz, int axis) {
  SysDesc D;
  memset(&D, 0, sizeof D);
  const long long str[3] = {1, nx, (long long)nx * ny};
  const int dims[3] = {nx, ny, nz};
  D.N = dims[axis];
  D.nother = 2;
  int o = 0;
  D.nsys = 1;
  for (int k = 0; k < 3; k++) {
    if (k == axis) continue;
    D.osize[o] = dims[k];
    for (int t = 0; t < 5; t++) D.ostr[t][o] = str[k];
    D.nsys *= dims[k];
    o++;
  }
  for (int t = 0; t < 5; t++) D.sstr[t] = str[axis];
  D.small_index = D.nsys < (1LL << 31);
  return D;
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

// one kappa sweep: d <- A^{-1} d (or A^{-T}); inc: u += A^{-1} d; out: the
// result goes to out instead of d (non-inc only)
tridStatus_t vadi_sweep(const double *K, double *d, double *u, double *out, int nx,
                        int ny, int nz, int axis, double lam, int transpose, int inc,
                        hipStream_t st) {
  const SysDesc D = line_desc(nx, ny, nz, axis);
  const int order = axis == 0 ? 0 : 1;
  PartCfg cfg = heuristic(order, D.N, 0, sizeof(double));
  if (cfg.kind != 1) return fail(TRID_STATUS_INVALID_VALUE, "vadi: grid too long along an axis");
  cfg.vec = 0;
  cfg.idx32 = (unsigned long long)nx * ny * nz * 8 < (1ULL << 32);
  const CoefKappa<double> cf{K, lam, transpose};
  hipError_t e = inc ? launch_part<double, 1>(cfg, cf, d, u, D, st)
                     : launch_part<double, 0>(cfg, cf, d, u, D, st, out);
  return check_hip(e, "vadi sweep launch");
}

tridStatus_t vadi_rhs(const double *K, const double *u, double *r, const Grid &g,
                      double lam, hipStream_t st) {
  hipLaunchKernelGGL(k_vadi_rhs<double>, dim3(blocks_for(g.n)), dim3(256), 0, st, K, u, r, g,
                     lam);
  return check_hip(hipGetLastError(), "vadi rhs launch");
}

// one forward step: u <- u + Az^-1 Ay^-1 Ax^-1 D u  (r: scratch)
tridStatus_t vadi_step(const double *K, double *u, double *r, const Grid &g, double lam,
                       hipStream_t st) {
  tridStatus_t s = vadi_rhs(K, u, r, g, lam, st);
  if (!s) s = vadi_sweep(K, r, nullptr, nullptr, g.nx, g.ny, g.nz, 0, lam, 0, 0, st);
  if (!s) s = vadi_sweep(K, r, nullptr, nullptr, g.nx, g.ny, g.nz, 1, lam, 0, 0, st);
  if (!s) s = vadi_sweep(K, r, u, nullptr, g.nx, g.ny, g.nz, 2, lam, 0, 1, st);
  return s;
}

Grid make_grid(int nx, int ny, int nz) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nxy = (long long)nx * ny;
  g.n = g.nxy * nz;
  return g;
}

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
  D.small_index = D.nsys < (1LL << 31);
  hipStream_t st = current_stream();
  const int order = solvedim == 0 ? 0 : 1;
  PartCfg cfg = heuristic(order, D.N, 0, sizeof(T));
  if (cfg.kind != 1) return fail(TRID_STATUS_INVALID_VALUE, "system too long for the adjoint kernel");
  cfg.vec = 0;
  cfg.idx32 = (unsigned long long)ext * sizeof(T) < (1ULL << 32);
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

EXTERN_C tridStatus_t tridVadiRhs(const double *kappa, const double *u, double *r,
                                  int nx, int ny, int nz, double lam) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || !kappa || !u || !r)
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  return vadi_rhs(kappa, u, r, make_grid(nx, ny, nz), lam, current_stream());
}

EXTERN_C tridStatus_t tridVadiSweep(const double *kappa, double *d, double *u, int nx,
                                    int ny, int nz, int axis, double lam, int transpose,
                                    int inc) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || !kappa || !d || (inc && !u) || axis < 0 || axis > 2 ||
      (inc && transpose))
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  return vadi_sweep(kappa, d, u, nullptr, nx, ny, nz, axis, lam, transpose, inc,
                    current_stream());
}

EXTERN_C tridStatus_t tridVadiForward(const double *kappa, double *u, int nx, int ny,
                                      int nz, int steps, double lam, int sync) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || steps < 0 || !kappa || !u)
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  const Grid g = make_grid(nx, ny, nz);
  double *r = static_cast<double *>(workspace((size_t)g.n * sizeof(double)));
  if (!r) return fail(TRID_STATUS_ALLOC_FAILED, "workspace allocation failed");
  hipStream_t st = current_stream();
  for (int s = 0; s < steps; s++) {
    tridStatus_t e = vadi_step(kappa, u, r, g, lam, st);
    if (e) return e;
  }
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridVadiLossGrad(const double *kappa, const double *u0,
                                       const double *target, int nx, int ny, int nz,
                                       int steps, double lam, int ckpt, double *loss,
                                       double *grad_kappa, double *grad_u0) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || steps < 1 || ckpt < 1 || !kappa || !u0 || !target ||
      !loss || !grad_kappa || !grad_u0)
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  const Grid g = make_grid(nx, ny, nz);
  const long long n = g.n;
  const int nseg = (steps + ckpt - 1) / ckpt;
  // workspace: nseg checkpoints, per step of a segment u_s, p, q, w (4*ckpt),
  // adjoint buffer L, scratch r, partial sums
  const long long nblk = blocks_for(n);
  const size_t arrays = (size_t)nseg + 4 * (size_t)ckpt + 2;
  double *ws = static_cast<double *>(workspace((arrays * n + nblk + 1) * sizeof(double)));
  if (!ws) return fail(TRID_STATUS_ALLOC_FAILED, "workspace allocation failed");
  double *ck = ws;                         // nseg checkpoints
  double *seg = ck + (size_t)nseg * n;     // 4 * ckpt arrays
  double *L = seg + 4 * (size_t)ckpt * n;  // adjoint solution
  double *r = L + n;                       // scratch
  double *partial = r + n;
  hipStream_t st = current_stream();
  const size_t bytes = (size_t)n * sizeof(double);
  tridStatus_t s = TRID_STATUS_SUCCESS;
  // forward pass keeping checkpoints; u lives in grad_u0 until the end
  double *u = grad_u0;
  s = check_hip(hipMemcpyAsync(u, u0, bytes, hipMemcpyDeviceToDevice, st), "copy");
  for (int k = 0; k < steps && !s; k++) {
    if (k % ckpt == 0)
      s = check_hip(hipMemcpyAsync(ck + (size_t)(k / ckpt) * n, u, bytes,
                                   hipMemcpyDeviceToDevice, st),
                    "checkpoint");
    if (!s) s = vadi_step(kappa, u, r, g, lam, st);
  }
  if (s) return s;
  // loss and ubar = u_T - target (ubar kept in grad_u0)
  hipLaunchKernelGGL(k_residual<double>, dim3((unsigned)nblk), dim3(256), 0, st, u, target, u,
                     partial, n);
  hipLaunchKernelGGL(k_sum<double>, dim3(1), dim3(256), 0, st, partial, partial + nblk, nblk);
  s = check_hip(hipGetLastError(), "loss launch");
  if (s) return s;
  double sq = 0.0;
  s = check_hip(hipMemcpyAsync(&sq, partial + nblk, sizeof(double), hipMemcpyDeviceToHost, st),
                "loss copy");
  if (s) return s;
  s = check_hip(hipMemsetAsync(grad_kappa, 0, bytes, st), "memset");
  double *ubar = u;
  for (int sg = nseg - 1; sg >= 0 && !s; sg--) {
    const int k0 = sg * ckpt, k1 = std::min(steps, k0 + ckpt);
    // recompute the segment, keeping u_k, p, q, w of every step
    double *uk = seg;
    s = check_hip(hipMemcpyAsync(uk, ck + (size_t)sg * n, bytes, hipMemcpyDeviceToDevice, st),
                  "restore");
    for (int k = k0; k < k1 && !s; k++) {
      double *us = seg + (size_t)4 * (k - k0) * n;
      double *p = us + n, *q = p + n, *w = q + n;
      if (k > k0) {
        double *prev = seg + (size_t)4 * (k - 1 - k0) * n;
        s = check_hip(hipMemcpyAsync(us, prev, bytes, hipMemcpyDeviceToDevice, st), "copy");
        if (!s) {
          hipLaunchKernelGGL(k_axpy<double>, dim3((unsigned)nblk), dim3(256), 0, st, us,
                             prev + 3 * n, n);
          s = check_hip(hipGetLastError(), "axpy");
        }
      }
      if (!s) s = vadi_rhs(kappa, us, r, g, lam, st);
      if (!s) s = vadi_sweep(kappa, r, nullptr, p, nx, ny, nz, 0, lam, 0, 0, st);
      if (!s) s = vadi_sweep(kappa, p, nullptr, q, nx, ny, nz, 1, lam, 0, 0, st);
      if (!s) s = vadi_sweep(kappa, q, nullptr, w, nx, ny, nz, 2, lam, 0, 0, st);
    }
    for (int k = k1 - 1; k >= k0 && !s; k--) {
      double *us = seg + (size_t)4 * (k - k0) * n;
      double *p = us + n, *q = p + n, *w = q + n;
      // z: L = Az^{-T} ubar
      s = vadi_sweep(kappa, ubar, nullptr, L, nx, ny, nz, 2, lam, 1, 0, st);
      if (s) break;
      hipLaunchKernelGGL(k_vadi_coef_grad<double>, dim3((unsigned)nblk), dim3(256), 0, st,
                         grad_kappa, L, w, g, 2, lam);
      s = vadi_sweep(kappa, L, nullptr, nullptr, nx, ny, nz, 1, lam, 1, 0, st);
      if (s) break;
      hipLaunchKernelGGL(k_vadi_coef_grad<double>, dim3((unsigned)nblk), dim3(256), 0, st,
                         grad_kappa, L, q, g, 1, lam);
      s = vadi_sweep(kappa, L, nullptr, nullptr, nx, ny, nz, 0, lam, 1, 0, st);
      if (s) break;
      hipLaunchKernelGGL(k_vadi_coef_grad<double>, dim3((unsigned)nblk), dim3(256), 0, st,
                         grad_kappa, L, p, g, 0, lam);
      hipLaunchKernelGGL(k_vadi_rhs_adj<double>, dim3((unsigned)nblk), dim3(256), 0, st, kappa,
                         us, L, ubar, grad_kappa, g, lam);
      s = check_hip(hipGetLastError(), "vadi adjoint launch");
    }
  }
  if (s) return s;
  s = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  if (s) return s;
  *loss = 0.5 * sq;
  return TRID_STATUS_SUCCESS;
}
