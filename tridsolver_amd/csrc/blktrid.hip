// blktrid.hip -- batched block-tridiagonal solves (include/blktrid_hip.h).
//
// Block Thomas per problem (block/src/blktrid_gpu.cuh:265-431 computes the
// same recurrence):
//   X_n = B_n - A_n C'_{n-1},  [C'_n | d'_n] = X_n^{-1} [C_n | d_n - A_n d'_{n-1}]
//   u_{N-1} = d'_{N-1},        u_n = d'_n - C'_n u_{n+1}
// with Gauss-Jordan elimination without pivoting on [X | C | d] (as the
// reference's gaussjordan_shfl) -- the systems are block diagonally dominant.
//
// Lane mapping (wave64): L lanes share one problem; lane j of a group owns
// the block columns j, j+L, ... (M/L of them).  Each lane keeps its columns
// of C'_{n-1} and the whole vector d'_{n-1} in registers from one block row
// to the next, so the forward sweep reads A, B, C, d once; C'/d' go to the
// caller's Cstar/dstar for the backward sweep (the reference's interface).
// The A_n columns and d_n are read by all L lanes of a group (one cache
// line fetch per group); Gauss-Jordan pivot columns are broadcast with
// ds_bpermute; in the back-substitution every lane of a group forms the
// whole C'_n u_{n+1} (the group's lanes read the same addresses).
// With TRANS the same kernel solves the transposed system of the adjoint:
// block row n of A^T is (C_{n-1}^T, B_n^T, A_{n+1}^T).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/blktrid_hip.h"
#include "../../include/trid_hip_ext.h"
#include "trid_device.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

template <typename T>
struct BlkArgs {
  const T *A, *B, *C, *d;
  T *Cs, *ds, *u;
  int N, P;
};

template <typename T, int M, int L, int TRANS>
__device__ __forceinline__ void blk_problem(const BlkArgs<T> &g, long long p,
                                            bool active, int j, int gbase) {
  constexpr int CM = M / L;  // columns per lane
  const long long P = g.P;
  const int N = g.N;
  auto blk = [&](int n) { return ((long long)n * P + p) * (M * M); };
  auto vec = [&](int n) { return ((long long)n * P + p) * M; };
  // (virtual) sub-, main- and super-diagonal block entries of row n
  auto sub = [&](int n, int i, int k) -> T {
    return TRANS ? g.C[blk(n - 1) + k * M + i] : g.A[blk(n) + i * M + k];
  };
  auto diag = [&](int n, int i, int c) -> T {
    return TRANS ? g.B[blk(n) + c * M + i] : g.B[blk(n) + i * M + c];
  };
  auto sup = [&](int n, int i, int c) -> T {
    return TRANS ? g.A[blk(n + 1) + c * M + i] : g.C[blk(n) + i * M + c];
  };

  T Cp[M][CM], dp[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    dp[i] = T(0);
#pragma unroll
    for (int t = 0; t < CM; t++) Cp[i][t] = T(0);
  }

  for (int n = 0; n < N; n++) {
    T X[M][CM], Cn[M][CM], v[M];
    if (active) {
#pragma unroll
      for (int i = 0; i < M; i++) {
#pragma unroll
        for (int t = 0; t < CM; t++) {
          const int c = j + L * t;
          X[i][t] = diag(n, i, c);
          Cn[i][t] = n < N - 1 ? sup(n, i, c) : T(0);
        }
        v[i] = g.d[vec(n) + i];
      }
      if (n > 0) {
#pragma unroll
        for (int k = 0; k < M; k++) {
          T acol[M];
#pragma unroll
          for (int i = 0; i < M; i++) acol[i] = sub(n, i, k);
#pragma unroll
          for (int i = 0; i < M; i++) {
#pragma unroll
            for (int t = 0; t < CM; t++) X[i][t] -= acol[i] * Cp[k][t];
            v[i] -= acol[i] * dp[k];
          }
          // keep at most a few columns of A_n in flight: for large blocks the
          // scheduler would otherwise hoist all M*M loads and spill
          if constexpr (M >= 6) {
            if ((k & 1) == 1) __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < M; i++) {
        v[i] = T(0);
#pragma unroll
        for (int t = 0; t < CM; t++) {
          X[i][t] = (j + L * t == i) ? T(1) : T(0);
          Cn[i][t] = T(0);
        }
      }
    }
    // Gauss-Jordan on [X | Cn | v] without pivoting
#pragma unroll
    for (int q = 0; q < M; q++) {
      T col[M];
      const int tq = q / L;
#pragma unroll
      for (int i = 0; i < M; i++)
        col[i] = (L == 1) ? X[i][tq] : lane_read(X[i][tq], min(gbase + (q % L), kWave - 1));
      const T inv = rcp(col[q]);
#pragma unroll
      for (int t = 0; t < CM; t++) {
        X[q][t] *= inv;
        Cn[q][t] *= inv;
      }
      v[q] *= inv;
#pragma unroll
      for (int i = 0; i < M; i++) {
        if (i == q) continue;
        const T f = col[i];
#pragma unroll
        for (int t = 0; t < CM; t++) {
          X[i][t] -= f * X[q][t];
          Cn[i][t] -= f * Cn[q][t];
        }
        v[i] -= f * v[q];
      }
    }
    if (active) {
#pragma unroll
      for (int t = 0; t < CM; t++) {
        const int c = j + L * t;
        if (n < N - 1) {
#pragma unroll
          for (int i = 0; i < M; i++) g.Cs[blk(n) + i * M + c] = Cn[i][t];
        }
      }
      // entry i of d'_n is stored by the lane owning column i
#pragma unroll
      for (int i = 0; i < M; i++)
        if (i % L == j) g.ds[vec(n) + i] = v[i];
    }
#pragma unroll
    for (int i = 0; i < M; i++) {
      dp[i] = v[i];
#pragma unroll
      for (int t = 0; t < CM; t++) Cp[i][t] = Cn[i][t];
    }
  }

  // backward: u_{N-1} = d'_{N-1} (still in dp), u_n = d'_n - C'_n u_{n+1}.
  // Every lane of a group forms the whole product from C'_n (the group's
  // lanes read the same addresses), so no cross-lane reduction is needed.
  T un[M];
#pragma unroll
  for (int i = 0; i < M; i++) un[i] = dp[i];
  if (active) {
#pragma unroll
    for (int i = 0; i < M; i++)
      if (i % L == j) g.u[vec(N - 1) + i] = un[i];
  }
  for (int n = N - 2; n >= 0; n--) {
    if (!active) continue;
    T nu[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      T acc = g.ds[vec(n) + i];
#pragma unroll
      for (int k = 0; k < M; k++) acc -= g.Cs[blk(n) + i * M + k] * un[k];
      nu[i] = acc;
      if constexpr (M >= 6) {
        if ((i & 1) == 1) __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int i = 0; i < M; i++) {
      un[i] = nu[i];
      if (i % L == j) g.u[vec(n) + i] = un[i];
    }
  }
}

// one work-group of 256 threads = 4 waves; each wave holds 64/L problems
template <typename T, int M, int L, int TRANS>
__device__ __forceinline__ void blk_block(const BlkArgs<T> &g, long long block) {
  constexpr int PW = kWave / L;  // problems per wave
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int grp = lane / L;
  const int j = lane - grp * L;
  const int gbase = grp * L;
  const long long p = (block * 4 + wave) * PW + grp;
  const bool active = grp < PW && p < g.P;
  blk_problem<T, M, L, TRANS>(g, p, active, j, gbase);
}

// Minimum waves per SIMD requested from the register allocator (it would
// otherwise hoist every block load of a row and spill for M >= 6).
constexpr int blk_min_waves(int M) { return M <= 2 ? 8 : (M <= 4 ? 4 : 2); }

template <typename T, int M, int L, int TRANS>
__global__ void __launch_bounds__(256, blk_min_waves(M)) k_blk(BlkArgs<T> g) {
  blk_block<T, M, L, TRANS>(g, blockIdx.x);
}

template <int M, int L>
constexpr long long blk_blocks(long long P) {
  return (P + 4 * (kWave / L) - 1) / (4 * (kWave / L));
}

// ---- mixed block sizes in one launch ---------------------------------------
constexpr int kMaxGroups = 16;
struct MixedArgs {
  BlkArgs<double> g[kMaxGroups];
  int M[kMaxGroups];
  long long first_block[kMaxGroups + 1];
  int ngroups;
};

// The lane count per block size of the mixed kernel (and the default).
__host__ __device__ constexpr int default_lanes(int M) {
  return M == 1 ? 1 : (M == 2 ? 2 : (M == 3 ? 3 : (M == 4 ? 4 : (M == 8 ? 8 : M))));
}

__global__ void __launch_bounds__(256, 2) k_blk_mixed(MixedArgs a) {
  const long long b = blockIdx.x;
  int gi = 0;
  while (gi + 1 < a.ngroups && b >= a.first_block[gi + 1]) gi++;
  const long long lb = b - a.first_block[gi];
  switch (a.M[gi]) {
#define MIXED_CASE(MM)                                                            \
  case MM:                                                                        \
    blk_block<double, MM, default_lanes(MM), 0>(a.g[gi], lb);                     \
    break;
    MIXED_CASE(1)
    MIXED_CASE(2)
    MIXED_CASE(3)
    MIXED_CASE(4)
    MIXED_CASE(5)
    MIXED_CASE(6)
    MIXED_CASE(7)
    MIXED_CASE(8)
#undef MIXED_CASE
    default:
      break;
  }
}

// ---- adjoint gradients: outer products ------------------------------------
// grad_A_n = -lam_n u_{n-1}^T, grad_B_n = -lam_n u_n^T, grad_C_n = -lam_n u_{n+1}^T
// one thread per block entry (n, p, i, c)
template <typename T>
__global__ void __launch_bounds__(256)
    k_blk_grads(const T *__restrict__ lam, const T *__restrict__ u,
                T *__restrict__ gA, T *__restrict__ gB, T *__restrict__ gC,
                int N, long long P, int M) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long MM = (long long)M * M;
  const long long total = (long long)N * P * MM;
  if (e >= total) return;
  const long long bp = e / MM;  // n * P + p
  const int ic = (int)(e - bp * MM);
  const int i = ic / M, c = ic - (ic / M) * M;
  const long long n = bp / P;
  const T l = lam[bp * M + i];
  gB[e] = -l * u[bp * M + c];
  gA[e] = n > 0 ? -l * u[(bp - P) * M + c] : T(0);
  gC[e] = n < N - 1 ? -l * u[(bp + P) * M + c] : T(0);
}

}  // namespace tridhip

using namespace tridhip;

namespace {

std::atomic<int> g_force_lanes{0};

template <typename T, int M, int TRANS>
hipError_t launch_m(const BlkArgs<T> &g, int lanes, hipStream_t st) {
  constexpr int Ld = default_lanes(M);
  const int L = lanes ? lanes : Ld;
#define LCASE(LL)                                                                \
  if constexpr (M % LL == 0)                                                     \
    if (L == LL) {                                                               \
      hipLaunchKernelGGL((k_blk<T, M, LL, TRANS>),                               \
                         dim3((unsigned)blk_blocks<M, LL>(g.P)), dim3(256), 0, st, g); \
      return hipGetLastError();                                                  \
    }
  LCASE(1)
  LCASE(2)
  LCASE(3)
  LCASE(4)
  LCASE(5)
  LCASE(6)
  LCASE(7)
  LCASE(8)
  LCASE(9)
  LCASE(10)
#undef LCASE
  return hipErrorInvalidValue;
}

template <typename T, int TRANS>
hipError_t launch_blk(const BlkArgs<T> &g, int M, int lanes, hipStream_t st) {
  switch (M) {
    case 1: return launch_m<T, 1, TRANS>(g, lanes, st);
    case 2: return launch_m<T, 2, TRANS>(g, lanes, st);
    case 3: return launch_m<T, 3, TRANS>(g, lanes, st);
    case 4: return launch_m<T, 4, TRANS>(g, lanes, st);
    case 5: return launch_m<T, 5, TRANS>(g, lanes, st);
    case 6: return launch_m<T, 6, TRANS>(g, lanes, st);
    case 7: return launch_m<T, 7, TRANS>(g, lanes, st);
    case 8: return launch_m<T, 8, TRANS>(g, lanes, st);
    case 9: return launch_m<T, 9, TRANS>(g, lanes, st);
    case 10: return launch_m<T, 10, TRANS>(g, lanes, st);
  }
  return hipErrorInvalidValue;
}

tridStatus_t check_blk(const void *A, const void *B, const void *C, const void *d,
                       const void *u, int N, int P, int M) {
  if (M < 1 || M > 10) return fail(TRID_STATUS_INVALID_VALUE, "blkdim must be in [1, 10]");
  if (N < 0 || P < 0) return fail(TRID_STATUS_INVALID_VALUE, "N and P must be >= 0");
  if (N > 0 && P > 0 && (!B || !d || !u || (N > 1 && (!A || !C))))
    return fail(TRID_STATUS_INVALID_VALUE, "NULL array");
  const int L = g_force_lanes.load();
  if (L && M % L != 0) return fail(TRID_STATUS_INVALID_VALUE, "forced lanes do not divide blkdim");
  return TRID_STATUS_SUCCESS;
}

template <typename T>
tridStatus_t btsv(const T *A, const T *B, const T *C, const T *d, T *Cs, T *ds,
                  T *u, int N, int P, int M, int sync) {
  clear_error();
  tridStatus_t s = check_blk(A, B, C, d, u, N, P, M);
  if (s) return s;
  if (N == 0 || P == 0) return TRID_STATUS_SUCCESS;
  if (!Cs || !ds) {
    size_t mat = (size_t)N * P * M * M, v = (size_t)N * P * M;
    T *w = static_cast<T *>(workspace((mat + v) * sizeof(T)));
    if (!w) return fail(TRID_STATUS_ALLOC_FAILED, "workspace allocation failed");
    Cs = w;
    ds = w + mat;
  }
  hipStream_t st = current_stream();
  BlkArgs<T> g{A, B, C, d, Cs, ds, u, N, P};
  s = check_hip(launch_blk<T, 0>(g, M, g_force_lanes.load(), st), "block kernel launch");
  if (s) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

}  // namespace

EXTERN_C tridStatus_t tridSbtsvStridedBatch(const float *A, const float *B,
                                            const float *C, const float *d,
                                            float *Cstar, float *dstar,
                                            float *u, int N, int P, int blkdim,
                                            int sync) {
  return btsv<float>(A, B, C, d, Cstar, dstar, u, N, P, blkdim, sync);
}

EXTERN_C tridStatus_t tridDbtsvStridedBatch(const double *A, const double *B,
                                            const double *C, const double *d,
                                            double *Cstar, double *dstar,
                                            double *u, int N, int P,
                                            int blkdim, int sync) {
  return btsv<double>(A, B, C, d, Cstar, dstar, u, N, P, blkdim, sync);
}

EXTERN_C void sbtsv_gpu(float *A, float *B, float *C, float *d, float *Cstar,
                        float *dstar, float *u, const int N, const int P,
                        const int blkdim) {
  (void)btsv<float>(A, B, C, d, Cstar, dstar, u, N, P, blkdim, 1);
}

EXTERN_C void dbtsv_gpu(double *A, double *B, double *C, double *d,
                        double *Cstar, double *dstar, double *u, const int N,
                        const int P, const int blkdim) {
  (void)btsv<double>(A, B, C, d, Cstar, dstar, u, N, P, blkdim, 0);
}

EXTERN_C tridStatus_t tridDbtsvMixedBatch(const tridDbtsvGroup_t *groups,
                                          int ngroups, int sync) {
  clear_error();
  if (ngroups < 0 || ngroups > kMaxGroups || (ngroups > 0 && !groups))
    return fail(TRID_STATUS_INVALID_VALUE, "ngroups must be in [0, 16]");
  MixedArgs a;
  memset(&a, 0, sizeof a);
  long long nb = 0;
  int k = 0;
  for (int i = 0; i < ngroups; i++) {
    const tridDbtsvGroup_t &G = groups[i];
    tridStatus_t s = check_blk(G.A, G.B, G.C, G.d, G.u, G.N, G.P, G.M);
    if (s) return s;
    if (G.M > 8) return fail(TRID_STATUS_INVALID_VALUE, "mixed batch: blkdim must be <= 8");
    if (G.N == 0 || G.P == 0) continue;
    if (!G.Cstar || !G.dstar)
      return fail(TRID_STATUS_INVALID_VALUE, "mixed batch: Cstar/dstar required");
    a.g[k] = BlkArgs<double>{G.A, G.B, G.C, G.d, G.Cstar, G.dstar, G.u, G.N, G.P};
    a.M[k] = G.M;
    a.first_block[k] = nb;
    const int L = default_lanes(G.M);
    const long long per = 4LL * (kWave / L);
    nb += (G.P + per - 1) / per;
    k++;
  }
  a.ngroups = k;
  a.first_block[k] = nb;
  if (nb == 0) return TRID_STATUS_SUCCESS;
  hipStream_t st = current_stream();
  hipLaunchKernelGGL(k_blk_mixed, dim3((unsigned)nb), dim3(256), 0, st, a);
  tridStatus_t s = check_hip(hipGetLastError(), "mixed block kernel launch");
  if (s) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridDbtsvAdjoint(const double *A, const double *B,
                                       const double *C, const double *u,
                                       const double *g, double *grad_A,
                                       double *grad_B, double *grad_C,
                                       double *grad_d, int N, int P,
                                       int blkdim, int sync) {
  clear_error();
  const int M = blkdim;
  tridStatus_t s = check_blk(A, B, C, g, u, N, P, M);
  if (s) return s;
  if (!grad_A || !grad_B || !grad_C || !grad_d)
    return fail(TRID_STATUS_INVALID_VALUE, "NULL gradient array");
  if (N == 0 || P == 0) return TRID_STATUS_SUCCESS;
  size_t mat = (size_t)N * P * M * M, v = (size_t)N * P * M;
  double *w = static_cast<double *>(workspace((mat + v) * sizeof(double)));
  if (!w) return fail(TRID_STATUS_ALLOC_FAILED, "workspace allocation failed");
  hipStream_t st = current_stream();
  // lambda = A^{-T} g  (written straight into grad_d)
  BlkArgs<double> ga{A, B, C, g, w, w + mat, grad_d, N, P};
  s = check_hip(launch_blk<double, 1>(ga, M, g_force_lanes.load(), st), "adjoint kernel launch");
  if (s) return s;
  const long long total = (long long)N * P * M * M;
  hipLaunchKernelGGL(k_blk_grads<double>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     st, grad_d, u, grad_A, grad_B, grad_C, N, (long long)P, M);
  s = check_hip(hipGetLastError(), "gradient kernel launch");
  if (s) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridBlkForceLanes(int lanes) {
  if (lanes < 0 || lanes > 10) return fail(TRID_STATUS_INVALID_VALUE, "lanes must be in [0, 10]");
  g_force_lanes.store(lanes);
  return TRID_STATUS_SUCCESS;
}
