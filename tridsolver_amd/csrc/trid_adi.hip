// trid_adi.hip -- ADI heat-diffusion time stepping (include/trid_adi.h).
//
// One time step is four kernels on the library stream:
//   k_adi_rhs           du = lambda * (7-point Laplacian of u), 0 on the
//                       boundary (apps/adi/src/preproc.cu:38-90 without the
//                       nine coefficient arrays)
//   k_part x, ORDER 0   du <- A_x^{-1} du   (wave-staged, contiguous lines)
//   k_part y, ORDER 1   du <- A_y^{-1} du
//   k_part z, ORDER 1   u  <- u + A_z^{-1} du   (the Inc form)
// with constant coefficients (CoefConst) instead of the reference's
// materialised a/b/c arrays: 9 words of HBM traffic per grid point and step
// instead of the reference's 27.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/trid_adi.h"
#include "../../include/trid_hip_ext.h"
#include "trid_launch.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

constexpr int kRhsPlanes = 8;  // z planes per thread of the RHS kernel

// 7-point right-hand side.  Block (64, 4): one thread per (i, j) column
// segment of kRhsPlanes planes, walking z with the k-1/k/k+1 values in
// registers.  The sum is formed in the reference's order
// (apps/adi/src/adi_orig.c:183-186) without contraction, so it matches the
// reference's rounding.
template <typename T>
__global__ void __launch_bounds__(256)
    k_adi_rhs(const T *__restrict__ u, T *__restrict__ du, int nx, int ny,
              int nz, T lam) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k0 = blockIdx.z * kRhsPlanes;
  if (i >= nx || j >= ny || k0 >= nz) return;
  const long long nxy = (long long)nx * ny;
  long long idx = i + (long long)nx * j + nxy * k0;
  const bool ij_in = i > 0 && i < nx - 1 && j > 0 && j < ny - 1;
  T um = k0 > 0 ? u[idx - nxy] : T(0);
  T uc = u[idx];
  const int k1 = min(k0 + kRhsPlanes, nz);
  for (int k = k0; k < k1; k++) {
    const T up = k + 1 < nz ? u[idx + nxy] : T(0);
    T v = T(0);
    if (ij_in && k > 0 && k < nz - 1) {
      T s = u[idx - 1] + u[idx + 1];
      s = s + u[idx - nx];
      s = s + u[idx + nx];
      s = s + um;
      s = s + up;
      const T six = T(6) * uc;
      s = s - six;
      v = lam * s;
    }
    st_nt(du + idx, v);
    um = uc;
    uc = up;
    idx += nxy;
  }
}

namespace {

SysDesc grid_desc(int nx, int ny, int nz, int dir) {
  SysDesc D;
  memset(&D, 0, sizeof D);
  const long long str[3] = {1, nx, (long long)nx * ny};
  const int dims[3] = {nx, ny, nz};
  D.N = dims[dir];
  D.nother = 2;
  int o = 0;
  D.nsys = 1;
  for (int k = 0; k < 3; k++) {
    if (k == dir) continue;
    D.osize[o] = dims[k];
    for (int t = 0; t < 5; t++) D.ostr[t][o] = str[k];
    D.nsys *= dims[k];
    o++;
  }
  for (int t = 0; t < 5; t++) D.sstr[t] = str[dir];
  D.small_index = D.nsys < (1LL << 31);
  return D;
}

}  // namespace
}  // namespace tridhip

using namespace tridhip;

struct tridAdiPlan_st {
  int nx = 0, ny = 0, nz = 0, dbl = 1, mode = 1;
  double lambda = 1.0;
  SysDesc desc[3];
  PartCfg cfg[3];      // x (16-byte aligned arrays), y, z
  PartCfg cfg_x_unaligned;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  std::string text;
};

namespace {

template <typename T>
CoefConst<T> coefs(const tridAdiPlan_st *p) {
  const T lam = (T)p->lambda;
  if (p->mode == 0) return CoefConst<T>{(T)-0.5f * lam, (T)1.0f + lam, (T)-0.5f * lam};
  return CoefConst<T>{-lam, (T)1.0f + (T)2.0f * lam, -lam};
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <typename T>
tridStatus_t stage_t(tridAdiPlan_st *p, T *u, T *du, int stage, hipStream_t st) {
  hipError_t e = hipSuccess;
  const CoefConst<T> cf = coefs<T>(p);
  if (stage == 0) {
    dim3 grid((p->nx + 63) / 64, (p->ny + 3) / 4, (p->nz + kRhsPlanes - 1) / kRhsPlanes);
    hipLaunchKernelGGL(k_adi_rhs<T>, grid, dim3(64, 4), 0, st, u, du, p->nx, p->ny,
                       p->nz, (T)p->lambda);
    e = hipGetLastError();
  } else if (stage == 1) {
    const PartCfg &c = aligned16(du) ? p->cfg[0] : p->cfg_x_unaligned;
    e = launch_part<T, 0>(c, cf, du, du, p->desc[0], st);
  } else if (stage == 2) {
    e = launch_part<T, 0>(p->cfg[1], cf, du, du, p->desc[1], st);
  } else {
    e = launch_part<T, 1>(p->cfg[2], cf, du, u, p->desc[2], st);
  }
  return check_hip(e, "ADI kernel launch");
}

tridStatus_t stage_any(tridAdiPlan_st *p, void *u, void *du, int stage, hipStream_t st) {
  if (p->dbl) return stage_t<double>(p, (double *)u, (double *)du, stage, st);
  return stage_t<float>(p, (float *)u, (float *)du, stage, st);
}

tridStatus_t steps_any(tridAdiPlan_st *p, void *u, void *du, int nsteps, hipStream_t st) {
  for (int s = 0; s < nsteps; s++)
    for (int g = 0; g < 4; g++) {
      tridStatus_t r = stage_any(p, u, du, g, st);
      if (r) return r;
    }
  return TRID_STATUS_SUCCESS;
}

tridStatus_t check_args(tridAdiPlan_t p, void *u, void *du) {
  if (!p) return fail(TRID_STATUS_INVALID_VALUE, "ADI plan is NULL");
  if (!u || !du) return fail(TRID_STATUS_INVALID_VALUE, "u and du must be non-NULL");
  if (u == du) return fail(TRID_STATUS_INVALID_VALUE, "u and du must be distinct arrays");
  return TRID_STATUS_SUCCESS;
}

}  // namespace

EXTERN_C tridStatus_t tridAdiPlanCreate(tridAdiPlan_t *plan, int nx, int ny,
                                        int nz, int is_double, double lambda,
                                        int mode) {
  clear_error();
  if (!plan) return fail(TRID_STATUS_INVALID_VALUE, "plan pointer is NULL");
  *plan = nullptr;
  if (nx < 1 || ny < 1 || nz < 1)
    return fail(TRID_STATUS_INVALID_VALUE, "nx, ny, nz must be >= 1");
  if (mode != 0 && mode != 1) return fail(TRID_STATUS_INVALID_VALUE, "mode must be 0 or 1");
  auto *p = new (std::nothrow) tridAdiPlan_st;
  if (!p) return fail(TRID_STATUS_ALLOC_FAILED, "plan allocation failed");
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->dbl = is_double ? 1 : 0;
  p->lambda = lambda;
  p->mode = mode;
  const size_t ts = p->dbl ? 8 : 4;
  const int idx32 = (unsigned long long)nx * ny * nz * ts < (1ULL << 32);
  const int n[3] = {nx, ny, nz};
  for (int dir = 0; dir < 3; dir++) {
    p->desc[dir] = grid_desc(nx, ny, nz, dir);
    const int order = dir == 0 ? 0 : 1;
    const int vec = dir == 0 && (nx * ts) % 16 == 0;
    PartCfg c = heuristic(order, n[dir], vec, ts);
    if (dir == 0) {
      PartCfg cu = heuristic(order, n[dir], 0, ts);
      cu.idx32 = idx32;
      p->cfg_x_unaligned = cu;
      if (stage_ok(c, true, p->desc[0].nsys, nx)) c.vec = 2;
    }
    c.idx32 = idx32;
    p->cfg[dir] = c;
  }
  p->text = "rhs<" + std::string(p->dbl ? "double" : "float") + "> x:" +
            cfg_name(p->cfg[0], ts) + " y:" + cfg_name(p->cfg[1], ts) +
            " z:" + cfg_name(p->cfg[2], ts);
  *plan = p;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridAdiPlanDestroy(tridAdiPlan_t p) {
  if (!p) return TRID_STATUS_SUCCESS;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  delete p;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridAdiStep(tridAdiPlan_t p, void *u, void *du, int nsteps,
                                  int sync) {
  clear_error();
  tridStatus_t s = check_args(p, u, du);
  if (s) return s;
  if (nsteps < 0) return fail(TRID_STATUS_INVALID_VALUE, "nsteps < 0");
  hipStream_t st = current_stream();
  s = steps_any(p, u, du, nsteps, st);
  if (s) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridAdiStage(tridAdiPlan_t p, void *u, void *du, int stage) {
  clear_error();
  tridStatus_t s = check_args(p, u, du);
  if (s) return s;
  if (stage < 0 || stage > 3) return fail(TRID_STATUS_INVALID_VALUE, "stage must be 0..3");
  return stage_any(p, u, du, stage, current_stream());
}

EXTERN_C tridStatus_t tridAdiGraphCapture(tridAdiPlan_t p, void *u, void *du,
                                          int nsteps) {
  clear_error();
  tridStatus_t s = check_args(p, u, du);
  if (s) return s;
  if (nsteps < 1) return fail(TRID_STATUS_INVALID_VALUE, "nsteps must be >= 1");
  if (!p->cap) {
    s = check_hip(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking),
                  "hipStreamCreate");
    if (s) return s;
  }
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  s = check_hip(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
  if (s) return s;
  tridStatus_t s2 = steps_any(p, u, du, nsteps, p->cap);
  hipGraph_t g = nullptr;
  s = check_hip(hipStreamEndCapture(p->cap, &g), "hipStreamEndCapture");
  if (s2) {
    if (g) (void)hipGraphDestroy(g);
    return s2;
  }
  if (s) return s;
  p->graph = g;
  return check_hip(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0),
                   "hipGraphInstantiate");
}

EXTERN_C tridStatus_t tridAdiGraphLaunch(tridAdiPlan_t p, int sync) {
  clear_error();
  if (!p || !p->exec)
    return fail(TRID_STATUS_NOT_INITIALIZED, "no captured ADI graph in this plan");
  hipStream_t st = current_stream();
  tridStatus_t s = check_hip(hipGraphLaunch(p->exec, st), "hipGraphLaunch");
  if (s) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

EXTERN_C const char *tridAdiPlanDescribe(tridAdiPlan_t p) {
  return p ? p->text.c_str() : "";
}
