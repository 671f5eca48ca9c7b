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

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/blktrid_hip.h"
#include "../../include/trid_hip_ext.h"
#include "trid_device.hpp"
#include "trid_launch.hpp"
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

// Two-ended ("twisted") form of blk_problem: each problem has two halves of
// L lanes.  Half 0 eliminates rows 0 .. m-1 downward (C'_n, d'_n as in
// blk_problem), half 1 rows N-1 .. m+1 upward -- the same recurrence on the
// rows in reverse order with the sub- and super-diagonal blocks exchanged:
// A''_n = X_n^{-1} A_n, d''_n = X_n^{-1}(d_n - C_n d''_{n+1}), X_n = B_n -
// C_n A''_{n+1} -- so that every problem runs two dependency chains of N/2
// block rows instead of one of N.  Both go to Cs/ds at their own rows (rows
// past m hold A'', d'').  The middle row m (m = N/2) takes both neighbours:
// (B_m - A_m C'_{m-1} - C_m A''_{m+1}) u_m = d_m - A_m d'_{m-1} - C_m
// d''_{m+1}; every lane of both halves forms it with the same operations in
// the same order (the partner half's C' or A'' columns and d' or d'' come
// through lane reads), so both hold the same u_m.  The halves then
// substitute outward: u_n = d'_n - C'_n u_{n+1} above m, u_n = d''_n - A''_n
// u_{n-1} below.  h: the lane's half; gbase: its half's first lane; partner:
// the lane of the other half that holds the same columns.
template <typename T, int M, int L, int TRANS>
__device__ __forceinline__ void blk_problem_tw(const BlkArgs<T> &g, long long p, bool active,
                                               int h, int j, int gbase, int partner) {
  constexpr int CM = M / L;
  const long long P = g.P;
  const int N = g.N;
  const int m = N / 2;
  const int cnt = h == 0 ? m : N - 1 - m;  // rows of this half
  auto blk = [&](int n) { return ((long long)n * P + p) * (M * M); };
  auto vec = [&](int n) { return ((long long)n * P + p) * M; };
  auto sub = [&](int n, int i, int k) -> T {
    return TRANS ? g.C[blk(n - 1) + k * M + i] : g.A[blk(n) + i * M + k];
  };
  auto diag = [&](int n, int i, int c) -> T {
    return TRANS ? g.B[blk(n) + c * M + i] : g.B[blk(n) + i * M + c];
  };
  auto sup = [&](int n, int i, int c) -> T {
    return TRANS ? g.A[blk(n + 1) + c * M + i] : g.C[blk(n) + i * M + c];
  };
  // row of step s of this half, and its couplings to the previous step's
  // row (back) and toward the middle row (fwd)
  auto rowof = [&](int s) { return h == 0 ? s : N - 1 - s; };
  auto back = [&](int r, int i, int k) -> T { return h == 0 ? sub(r, i, k) : sup(r, i, k); };
  auto fwd = [&](int r, int i, int c) -> T { return h == 0 ? sup(r, i, c) : sub(r, i, c); };

  T Cp[M][CM], dp[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    dp[i] = T(0);
#pragma unroll
    for (int t = 0; t < CM; t++) Cp[i][t] = T(0);
  }
  // [X | Cn | v] <- X^{-1} [X | Cn | v] (Gauss-Jordan without pivoting, as
  // blk_problem; the lanes of one half share a block's columns)
  auto gauss_jordan = [&](T (&X)[M][CM], T (&Cn)[M][CM], T (&v)[M]) {
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
  };
  // X -= K Q, v -= K q for the row's coupling block K (entry (i, k) from kf)
  // and the neighbour's eliminated Q (this lane's columns) and q
  auto couple = [&](T (&X)[M][CM], T (&v)[M], auto kf, const T (&Q)[M][CM], const T (&q)[M]) {
#pragma unroll
    for (int k = 0; k < M; k++) {
      T kcol[M];
#pragma unroll
      for (int i = 0; i < M; i++) kcol[i] = kf(i, k);
#pragma unroll
      for (int i = 0; i < M; i++) {
#pragma unroll
        for (int t = 0; t < CM; t++) X[i][t] -= kcol[i] * Q[k][t];
        v[i] -= kcol[i] * q[k];
      }
      if constexpr (M >= 6) {
        if ((k & 1) == 1) __builtin_amdgcn_sched_barrier(0);
      }
    }
  };

  for (int s = 0; s < cnt; s++) {
    const int r = rowof(s);
    T X[M][CM], Cn[M][CM], v[M];
    if (active) {
#pragma unroll
      for (int i = 0; i < M; i++) {
#pragma unroll
        for (int t = 0; t < CM; t++) {
          const int c = j + L * t;
          X[i][t] = diag(r, i, c);
          Cn[i][t] = fwd(r, i, c);
        }
        v[i] = g.d[vec(r) + i];
      }
      if (s > 0) couple(X, v, [&](int i, int k) { return back(r, i, k); }, Cp, dp);
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
    gauss_jordan(X, Cn, v);
    if (active) {
#pragma unroll
      for (int t = 0; t < CM; t++) {
        const int c = j + L * t;
#pragma unroll
        for (int i = 0; i < M; i++) g.Cs[blk(r) + i * M + c] = Cn[i][t];
      }
#pragma unroll
      for (int i = 0; i < M; i++)
        if (i % L == j) g.ds[vec(r) + i] = v[i];
    }
#pragma unroll
    for (int i = 0; i < M; i++) {
      dp[i] = v[i];
#pragma unroll
      for (int t = 0; t < CM; t++) Cp[i][t] = Cn[i][t];
    }
  }

  // the middle row: the upper neighbour's C', d' and the lower one's A'', d''
  // in the order (upper, lower) on both halves
  T Qo[M][CM], qo[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    qo[i] = lane_read(dp[i], partner);
#pragma unroll
    for (int t = 0; t < CM; t++) Qo[i][t] = lane_read(Cp[i][t], partner);
  }
  T un[M];
  {
    T X[M][CM], Cn[M][CM], v[M];
    if (active) {
#pragma unroll
      for (int i = 0; i < M; i++) {
#pragma unroll
        for (int t = 0; t < CM; t++) {
          X[i][t] = diag(m, i, j + L * t);
          Cn[i][t] = T(0);
        }
        v[i] = g.d[vec(m) + i];
      }
      if (m > 0) {
        if (h == 0) couple(X, v, [&](int i, int k) { return sub(m, i, k); }, Cp, dp);
        else couple(X, v, [&](int i, int k) { return sub(m, i, k); }, Qo, qo);
      }
      if (m < N - 1) {
        if (h == 0) couple(X, v, [&](int i, int k) { return sup(m, i, k); }, Qo, qo);
        else couple(X, v, [&](int i, int k) { return sup(m, i, k); }, Cp, dp);
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
    gauss_jordan(X, Cn, v);
#pragma unroll
    for (int i = 0; i < M; i++) un[i] = v[i];
  }
  if (active && h == 0) {
#pragma unroll
    for (int i = 0; i < M; i++)
      if (i % L == j) g.u[vec(m) + i] = un[i];
  }
  // substitution outward from the middle, each half over its own rows
  for (int s = cnt - 1; s >= 0; s--) {
    if (!active) continue;
    const int r = rowof(s);
    T nu[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      T acc = g.ds[vec(r) + i];
#pragma unroll
      for (int k = 0; k < M; k++) acc -= g.Cs[blk(r) + i * M + k] * un[k];
      nu[i] = acc;
      if constexpr (M >= 6) {
        if ((i & 1) == 1) __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int i = 0; i < M; i++) {
      un[i] = nu[i];
      if (i % L == j) g.u[vec(r) + i] = un[i];
    }
  }
}

// Prefetching form of blk_problem (used where blk_prefetch says): row n+1's
// blocks are loaded
// while row n is eliminated, and row n-1's C'/d' while row n is
// substituted.  With one lane per problem a batch of 65,536 problems fills
// one wave per SIMD, so without this every block row waits a full memory
// latency.  Same operations in the same order as blk_problem.
template <typename T, int M, int L, int TRANS>
__device__ __forceinline__ void blk_problem_pf(const BlkArgs<T> &g, long long p,
                                               bool active, int j, int gbase) {
  constexpr int CM = M / L;
  const long long P = g.P;
  const int N = g.N;
  auto blk = [&](int n) { return ((long long)n * P + p) * (M * M); };
  auto vec = [&](int n) { return ((long long)n * P + p) * M; };
  auto sub = [&](int n, int i, int k) -> T {
    return TRANS ? g.C[blk(n - 1) + k * M + i] : g.A[blk(n) + i * M + k];
  };
  auto diag = [&](int n, int i, int c) -> T {
    return TRANS ? g.B[blk(n) + c * M + i] : g.B[blk(n) + i * M + c];
  };
  auto sup = [&](int n, int i, int c) -> T {
    return TRANS ? g.A[blk(n + 1) + c * M + i] : g.C[blk(n) + i * M + c];
  };
  // the inputs of a block row: diagonal and super-diagonal columns of this
  // lane, the right-hand side and the whole sub-diagonal block
  struct RowIn {
    T X[M][CM], C[M][CM], v[M], A[M][M];
  };
  auto load_row = [&](int n, RowIn &r) {
#pragma unroll
    for (int i = 0; i < M; i++) {
#pragma unroll
      for (int t = 0; t < CM; t++) {
        const int c = j + L * t;
        r.X[i][t] = diag(n, i, c);
        r.C[i][t] = n < N - 1 ? sup(n, i, c) : T(0);
      }
      r.v[i] = g.d[vec(n) + i];
#pragma unroll
      for (int k = 0; k < M; k++) r.A[i][k] = n > 0 ? sub(n, i, k) : T(0);
    }
  };

  T Cp[M][CM], dp[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    dp[i] = T(0);
#pragma unroll
    for (int t = 0; t < CM; t++) Cp[i][t] = T(0);
  }
  RowIn cur;
  if (active) load_row(0, cur);
  for (int n = 0; n < N; n++) {
    RowIn nxt;
    if (active && n + 1 < N) load_row(n + 1, nxt);
    T X[M][CM], Cn[M][CM], v[M];
    if (active) {
#pragma unroll
      for (int i = 0; i < M; i++) {
#pragma unroll
        for (int t = 0; t < CM; t++) {
          X[i][t] = cur.X[i][t];
          Cn[i][t] = cur.C[i][t];
        }
        v[i] = cur.v[i];
      }
      if (n > 0) {
#pragma unroll
        for (int k = 0; k < M; k++) {
#pragma unroll
          for (int i = 0; i < M; i++) {
#pragma unroll
            for (int t = 0; t < CM; t++) X[i][t] -= cur.A[i][k] * Cp[k][t];
            v[i] -= cur.A[i][k] * dp[k];
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
    cur = nxt;
  }

  // backward, with row n-1's C'_n-1 and d'_n-1 loaded during row n
  T un[M];
#pragma unroll
  for (int i = 0; i < M; i++) un[i] = dp[i];
  if (active) {
#pragma unroll
    for (int i = 0; i < M; i++)
      if (i % L == j) g.u[vec(N - 1) + i] = un[i];
  }
  if (!active) return;
  T bC[M][M], bd[M];
  auto load_back = [&](int n) {
#pragma unroll
    for (int i = 0; i < M; i++) {
      bd[i] = g.ds[vec(n) + i];
#pragma unroll
      for (int k = 0; k < M; k++) bC[i][k] = g.Cs[blk(n) + i * M + k];
    }
  };
  if (N >= 2) load_back(N - 2);
  for (int n = N - 2; n >= 0; n--) {
    T cC[M][M], cd[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      cd[i] = bd[i];
#pragma unroll
      for (int k = 0; k < M; k++) cC[i][k] = bC[i][k];
    }
    if (n >= 1) load_back(n - 1);
    T nu[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      T acc = cd[i];
#pragma unroll
      for (int k = 0; k < M; k++) acc -= cC[i][k] * un[k];
      nu[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < M; i++) {
      un[i] = nu[i];
      if (i % L == j) g.u[vec(n) + i] = un[i];
    }
  }
}

// Where blk_problem_pf is used: measured on MI355X (N = 128, P = 65,536) it
// helps M = 3 with one lane per problem (0.913 -> 0.870 ms) and costs the
// layouts with more waves per problem set or smaller rows (M = 2, one lane:
// 0.350 -> 0.368; M = 3, three lanes: 1.058 -> 1.091; M = 4, four lanes:
// 1.46 -> 1.49).
__host__ __device__ constexpr bool blk_prefetch(int M, int L) { return M == 3 && L == 1; }

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
  if constexpr (blk_prefetch(M, L))
    blk_problem_pf<T, M, L, TRANS>(g, p, active, j, gbase);
  else
    blk_problem<T, M, L, TRANS>(g, p, active, j, gbase);
}

// Minimum waves per SIMD requested from the register allocator (it would
// otherwise hoist every block load of a row and spill for M >= 6).
constexpr int blk_min_waves(int M) { return M <= 2 ? 8 : (M <= 4 ? 4 : 2); }

// The prefetching form holds two rows' inputs: fewer waves per SIMD (it runs
// with one lane per problem, where a batch fills few waves anyway).
constexpr int blk_waves(int M, int L) {
  return blk_prefetch(M, L) ? 2 : blk_min_waves(M);
}

template <typename T, int M, int L, int TRANS>
__global__ void __launch_bounds__(256, blk_waves(M, L)) k_blk(BlkArgs<T> g) {
  blk_block<T, M, L, TRANS>(g, blockIdx.x);
}

template <int M, int L>
constexpr long long blk_blocks(long long P) {
  return (P + 4 * (kWave / L) - 1) / (4 * (kWave / L));
}

// the two-ended form: 2L lanes per problem (half 0 = lanes 0..L-1 of the
// group, half 1 = lanes L..2L-1), 64/(2L) problems per wave
template <typename T, int M, int L, int TRANS>
__global__ void __launch_bounds__(256, blk_min_waves(M)) k_blk_tw(BlkArgs<T> g) {
  constexpr int PW = kWave / (2 * L);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int grp = lane / (2 * L);
  const int in = lane - grp * 2 * L;
  const int h = in / L, j = in - h * L;
  const int gbase = grp * 2 * L + h * L;
  const int partner = grp * 2 * L + (1 - h) * L + j;
  const long long p = ((long long)blockIdx.x * 4 + wave) * PW + grp;
  const bool active = grp < PW && p < g.P;
  // idle lanes past the last whole group read from themselves
  blk_problem_tw<T, M, L, TRANS>(g, p, active, h, j, min(gbase, kWave - 1),
                                 grp < PW ? partner : lane);
}

template <int M, int L>
constexpr long long blk_tw_blocks(long long P) {
  return (P + 4 * (kWave / (2 * L)) - 1) / (4 * (kWave / (2 * L)));
}

// ---- one lane per block column, tiles through LDS ---------------------------
// The default layout (L = M <= 8).  Every M x M block is moved between
// global memory and the group's LDS tile cooperatively -- lane j of a group
// takes elements j, j+M, j+2M, ... of the contiguous block, so each load or
// store instruction of a wave covers one contiguous stretch of 64 elements --
// and every lane then reads what it needs from the tile: column j of B_n and
// C_n, every entry of A_n (group-wide broadcast reads), row i of C'_n in the
// backward sweep.  The forward recurrence, Gauss-Jordan elimination and
// storage format are those of blk_problem; the backward sweep is
// row-distributed (lane i forms u_n[i] = d'_n[i] - C'_n[i,:] u_{n+1}, and u_n
// is shared through the tile).  The groups of a wave exchange data only
// within their own wave, so wave barriers order the LDS accesses.
template <typename T, int M>
constexpr int blk2_tile() {
  return 3 * M * M + 2 * M;  // A, B, C tiles, d and u vectors
}

// p: the problem (problem P-1 for the lanes of groups past the end, which
// solve it again and store nothing); store: whether the lane writes results;
// lane_ok: the lane belongs to a whole group (idle lanes do not write the
// tile).
template <typename T, int M, int TRANS>
__device__ __forceinline__ void blk2_problem(const BlkArgs<T> &g, long long p, bool store,
                                             bool lane_ok, int j, int gbase, T *tile) {
  const long long P = g.P;
  const int N = g.N;
  auto blk = [&](int n) { return ((long long)n * P + p) * (M * M); };
  auto vec = [&](int n) { return ((long long)n * P + p) * M; };
  T *sA = tile, *sB = tile + M * M, *sC = tile + 2 * M * M, *sd = tile + 3 * M * M;
  T *su = sd + M;
  auto wb = []() { __builtin_amdgcn_wave_barrier(); };

  T Cp[M], dp[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    Cp[i] = T(0);
    dp[i] = T(0);
  }
  // the row's tiles: sub-diagonal (A_n; C_{n-1} transposed), diagonal
  // (B_n; transposed), super-diagonal (C_n; A_{n+1} transposed), this
  // lane's elements j, j+M, ...  Row n+1 is loaded while row n is
  // eliminated: with at most a few problems per SIMD the recurrence would
  // otherwise wait a full memory latency every block row.
  T ra[M], rb[M], rc[M], rd;
  auto load_row = [&](int n) {
    const bool sub_n = n > 0, sup_n = n < N - 1;
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int e = j + M * r;
      ra[r] = sub_n ? (TRANS ? g.C[blk(n - 1) + e] : g.A[blk(n) + e]) : T(0);
      rb[r] = g.B[blk(n) + e];
      rc[r] = sup_n ? (TRANS ? g.A[blk(n + 1) + e] : g.C[blk(n) + e]) : T(0);
    }
    rd = g.d[vec(n) + j];
  };
  load_row(0);
  for (int n = 0; n < N; n++) {
    const bool has_sub = n > 0, has_sup = n < N - 1;
    if (lane_ok) {
#pragma unroll
      for (int r = 0; r < M; r++) {
        const int e = j + M * r;
        sA[e] = ra[r];
        sB[e] = rb[r];
        sC[e] = rc[r];
      }
      sd[j] = rd;
    }
    wb();
    if (n + 1 < N) load_row(n + 1);
    // element (i, c) of the row's sub/diag/sup block
    auto at_t = [&](const T *t, int i, int c) { return TRANS ? t[c * M + i] : t[i * M + c]; };
    T X[M], Cn[M], v[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      X[i] = at_t(sB, i, j);
      Cn[i] = at_t(sC, i, j);
      v[i] = sd[i];
    }
    if (has_sub) {
#pragma unroll
      for (int k = 0; k < M; k++) {
#pragma unroll
        for (int i = 0; i < M; i++) {
          const T aik = at_t(sA, i, k);
          X[i] -= aik * Cp[k];
          v[i] -= aik * dp[k];
        }
        // keep a few columns of A_n in flight: the scheduler would otherwise
        // hoist all M*M tile reads and spill for large blocks
        if constexpr (M >= 5) {
          if ((k & 1) == 1) __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    wb();
    // phase boundary: no tile load or read moves across the elimination
    __builtin_amdgcn_sched_barrier(0);
    // Gauss-Jordan on [X | Cn | v] without pivoting (pivot columns from the
    // owning lane)
#pragma unroll
    for (int q = 0; q < M; q++) {
      T col[M];
#pragma unroll
      for (int i = 0; i < M; i++) col[i] = lane_read(X[i], min(gbase + q, kWave - 1));
      const T inv = rcp(col[q]);
      X[q] *= inv;
      Cn[q] *= inv;
      v[q] *= inv;
#pragma unroll
      for (int i = 0; i < M; i++) {
        if (i == q) continue;
        const T f = col[i];
        X[i] -= f * X[q];
        Cn[i] -= f * Cn[q];
        v[i] -= f * v[q];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // C'_n (column j in this lane) and d'_n to the caller's scratch
    if (has_sup) {
      if (lane_ok) {
#pragma unroll
        for (int i = 0; i < M; i++) sC[i * M + j] = Cn[i];
      }
      wb();
      if (store) {
#pragma unroll
        for (int r = 0; r < M; r++) g.Cs[blk(n) + j + M * r] = sC[j + M * r];
      }
      wb();
    }
    if (store) g.ds[vec(n) + j] = v[j];
#pragma unroll
    for (int i = 0; i < M; i++) {
      Cp[i] = Cn[i];
      dp[i] = v[i];
    }
  }
  // backward: u_{N-1} = d'_{N-1} (dp, in every lane), u_n = d'_n - C'_n u_{n+1}
  T un[M];
#pragma unroll
  for (int i = 0; i < M; i++) un[i] = dp[i];
  if (store) g.u[vec(N - 1) + j] = un[j];
  // C'_n and d'_n as this lane's group stored them (stores by the other
  // lanes of a wave are visible to its later loads); row n-1 is loaded
  // while row n is substituted
  T bc[M], bd = T(0);
  auto load_back = [&](int n) {
#pragma unroll
    for (int r = 0; r < M; r++) bc[r] = g.Cs[blk(n) + j + M * r];
    bd = g.ds[vec(n) + j];
  };
  if (N >= 2) load_back(N - 2);
  for (int n = N - 2; n >= 0; n--) {
    const T dj = bd;
    if (lane_ok) {
#pragma unroll
      for (int r = 0; r < M; r++) sC[j + M * r] = bc[r];
    }
    wb();
    if (n >= 1) load_back(n - 1);
    T acc = dj;
#pragma unroll
    for (int k = 0; k < M; k++) acc -= sC[j * M + k] * un[k];
    if (lane_ok) su[j] = acc;
    wb();
#pragma unroll
    for (int k = 0; k < M; k++) un[k] = su[k];
    wb();
    if (store) g.u[vec(n) + j] = acc;
  }
}

// 256 threads = 4 waves of 64/M groups of M lanes (lanes past the last whole
// group are idle)
template <typename T, int M, int TRANS>
__device__ __forceinline__ void blk2_block(const BlkArgs<T> &g, long long block, T *smem) {
  constexpr int G = kWave / M;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int grp = lane / M;
  const int j = lane - grp * M;
  const long long p = (block * 4 + wave) * G + grp;
  const bool lane_ok = grp < G;
  const bool store = lane_ok && p < g.P;
  T *tile = smem + (wave * G + (lane_ok ? grp : G - 1)) * blk2_tile<T, M>();
  blk2_problem<T, M, TRANS>(g, std::min<long long>(p, g.P - 1), store, lane_ok, j, grp * M, tile);
}

// Minimum waves per SIMD for the register allocator (the live set is about
// 5M values per lane plus the next row's 3M + 1; without a bound the
// scheduler hoists tile loads and doubles it).  65,536 problems fill about
// 3 (M = 3) to 6 (M = 6) waves per SIMD of MI355X, so the bound costs no
// occupancy that a batch of that size would use.
constexpr int blk2_min_waves(int M) {
  return M <= 2 ? 8 : (M == 3 ? 4 : (M == 4 ? 3 : 2));
}

template <typename T, int M, int TRANS>
__global__ void __launch_bounds__(256, blk2_min_waves(M)) k_blk2(BlkArgs<T> g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  blk2_block<T, M, TRANS>(g, blockIdx.x, reinterpret_cast<T *>(smem_raw));
}

// Block sizes for which the tiled kernel is the default (measured faster on
// MI355X; for M = 1, 2 the direct kernel's occupancy wins, for M = 7, 8 the
// tiled kernel's register demand does not fit).
__host__ __device__ constexpr bool blk2_sizes(int M) { return M >= 3 && M <= 6; }
// Block sizes of the two-ended tiled kernel (half the live set of k_blk2 per
// lane: no spills up to M = 8, 200 VGPRs at M = 8 FP64).
__host__ __device__ constexpr bool blk2tw_sizes(int M) { return M >= 3 && M <= 8; }
// layout code of the two-ended tiled kernel (tridBlkForceLanes, MixedArgs::L)
constexpr int kTwistTiled = 200;

// Two-ended form of blk2_problem (the elimination of blk_problem_tw with the
// tiled kernel's cooperative block moves): two groups of M lanes per
// problem, each with its own tile.  Half 0 eliminates rows 0 .. m-1
// downward (C'_n, d'_n), half 1 rows N-1 .. m+1 upward (A''_n, d''_n, the
// sub- and super-diagonal blocks exchanging roles), m = N/2; both halves
// form the middle row's system with the same operations in the same order
// (the other half's eliminated columns through lane reads, partner = the
// lane of the other half with the same column j) and substitute outward
// through their tiles.  Tile roles: K = the coupling to the previous row of
// the half (sub-diagonal for half 0, super-diagonal for half 1; sub(m) for
// the middle row), F = the coupling toward the middle (super / sub; sup(m)).
template <typename T, int M, int TRANS>
__device__ __forceinline__ void blk2_problem_tw(const BlkArgs<T> &g, long long p, bool store,
                                                bool lane_ok, int h, int j, int gbase,
                                                int partner, T *tile) {
  const long long P = g.P;
  const int N = g.N;
  const int m = N / 2;
  const int cnt = h == 0 ? m : N - 1 - m;  // rows of this half
  auto blk = [&](int n) { return ((long long)n * P + p) * (M * M); };
  auto vec = [&](int n) { return ((long long)n * P + p) * M; };
  T *sK = tile, *sB = tile + M * M, *sF = tile + 2 * M * M, *sd = tile + 3 * M * M;
  T *su = sd + M;
  auto wb = []() { __builtin_amdgcn_wave_barrier(); };
  auto rowof = [&](int s) { return h == 0 ? s : N - 1 - s; };
  // element e of the raw sub- and super-diagonal blocks of row n (C_{n-1}
  // and A_{n+1} read transposed through at_t in TRANS mode)
  auto ld_sub = [&](int n, int e) -> T { return TRANS ? g.C[blk(n - 1) + e] : g.A[blk(n) + e]; };
  auto ld_sup = [&](int n, int e) -> T { return TRANS ? g.A[blk(n + 1) + e] : g.C[blk(n) + e]; };
  auto at_t = [&](const T *t, int i, int c) { return TRANS ? t[c * M + i] : t[i * M + c]; };

  T Cp[M], dp[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    Cp[i] = T(0);
    dp[i] = T(0);
  }
  // this lane's elements j, j+M, ... of the next row's K, B, F blocks and
  // entry j of its d (the next row is loaded while this one is eliminated)
  T rk[M], rb[M], rf[M], rd;
  auto load_step = [&](int s) {
    const int n = rowof(s);
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int e = j + M * r;
      rk[r] = s > 0 ? (h == 0 ? ld_sub(n, e) : ld_sup(n, e)) : T(0);
      rb[r] = g.B[blk(n) + e];
      rf[r] = h == 0 ? ld_sup(n, e) : ld_sub(n, e);
    }
    rd = g.d[vec(n) + j];
  };
  auto load_mid = [&]() {
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int e = j + M * r;
      rk[r] = m > 0 ? ld_sub(m, e) : T(0);
      rb[r] = g.B[blk(m) + e];
      rf[r] = m < N - 1 ? ld_sup(m, e) : T(0);
    }
    rd = g.d[vec(m) + j];
  };
  auto put_tiles = [&]() {
    if (lane_ok) {
#pragma unroll
      for (int r = 0; r < M; r++) {
        const int e = j + M * r;
        sK[e] = rk[r];
        sB[e] = rb[r];
        sF[e] = rf[r];
      }
      sd[j] = rd;
    }
    wb();
  };
  // X -= K Q, v -= K q for the tile t's block K and the eliminated columns
  // Q (this lane's column) and vector q
  auto couple = [&](T (&X)[M], T (&v)[M], const T *t, const T (&Q)[M], const T (&q)[M]) {
#pragma unroll
    for (int k = 0; k < M; k++) {
#pragma unroll
      for (int i = 0; i < M; i++) {
        const T kik = at_t(t, i, k);
        X[i] -= kik * Q[k];
        v[i] -= kik * q[k];
      }
      if constexpr (M >= 5) {
        if ((k & 1) == 1) __builtin_amdgcn_sched_barrier(0);
      }
    }
  };
  // [X | Cn | v] <- X^{-1} [X | Cn | v], pivot columns from the half's lanes
  auto gauss_jordan = [&](T (&X)[M], T (&Cn)[M], T (&v)[M]) {
#pragma unroll
    for (int q = 0; q < M; q++) {
      T col[M];
#pragma unroll
      for (int i = 0; i < M; i++) col[i] = lane_read(X[i], min(gbase + q, kWave - 1));
      const T inv = rcp(col[q]);
      X[q] *= inv;
      Cn[q] *= inv;
      v[q] *= inv;
#pragma unroll
      for (int i = 0; i < M; i++) {
        if (i == q) continue;
        const T f = col[i];
        X[i] -= f * X[q];
        Cn[i] -= f * Cn[q];
        v[i] -= f * v[q];
      }
    }
  };

  if (cnt > 0) load_step(0);
  else load_mid();
  for (int s = 0; s < cnt; s++) {
    const int n = rowof(s);
    put_tiles();
    if (s + 1 < cnt) load_step(s + 1);
    else load_mid();
    T X[M], Cn[M], v[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      X[i] = at_t(sB, i, j);
      Cn[i] = at_t(sF, i, j);
      v[i] = sd[i];
    }
    if (s > 0) couple(X, v, sK, Cp, dp);
    wb();
    __builtin_amdgcn_sched_barrier(0);
    gauss_jordan(X, Cn, v);
    __builtin_amdgcn_sched_barrier(0);
    // the eliminated F block (column j in this lane) and d to the scratch
    if (lane_ok) {
#pragma unroll
      for (int i = 0; i < M; i++) sF[i * M + j] = Cn[i];
    }
    wb();
    if (store) {
#pragma unroll
      for (int r = 0; r < M; r++) g.Cs[blk(n) + j + M * r] = sF[j + M * r];
      g.ds[vec(n) + j] = v[j];
    }
    wb();
#pragma unroll
    for (int i = 0; i < M; i++) {
      Cp[i] = Cn[i];
      dp[i] = v[i];
    }
  }

  // the middle row: the upper neighbour's C', d' and the lower one's A'',
  // d'' in the order (upper, lower) on both halves
  T Qo[M], qo[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    Qo[i] = lane_read(Cp[i], partner);
    qo[i] = lane_read(dp[i], partner);
  }
  put_tiles();
  T un[M];
  {
    T X[M], Cn[M], v[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      X[i] = at_t(sB, i, j);
      Cn[i] = T(0);
      v[i] = sd[i];
    }
    if (m > 0) {
      if (h == 0) couple(X, v, sK, Cp, dp);
      else couple(X, v, sK, Qo, qo);
    }
    if (m < N - 1) {
      if (h == 0) couple(X, v, sF, Qo, qo);
      else couple(X, v, sF, Cp, dp);
    }
    wb();
    __builtin_amdgcn_sched_barrier(0);
    gauss_jordan(X, Cn, v);
#pragma unroll
    for (int i = 0; i < M; i++) un[i] = v[i];
  }
  if (store && h == 0) g.u[vec(m) + j] = un[j];

  // substitution outward from the middle over the half's rows: lane i forms
  // u_n[i] = d*_n[i] - F*_n[i, :] u_prev from the row's tile
  T bc[M], bd = T(0);
  auto load_back = [&](int n) {
#pragma unroll
    for (int r = 0; r < M; r++) bc[r] = g.Cs[blk(n) + j + M * r];
    bd = g.ds[vec(n) + j];
  };
  if (cnt > 0) load_back(rowof(cnt - 1));
  for (int s = cnt - 1; s >= 0; s--) {
    const int n = rowof(s);
    const T dj = bd;
    if (lane_ok) {
#pragma unroll
      for (int r = 0; r < M; r++) sF[j + M * r] = bc[r];
    }
    wb();
    if (s >= 1) load_back(rowof(s - 1));
    T acc = dj;
#pragma unroll
    for (int k = 0; k < M; k++) acc -= sF[j * M + k] * un[k];
    if (lane_ok) su[j] = acc;
    wb();
#pragma unroll
    for (int k = 0; k < M; k++) un[k] = su[k];
    wb();
    if (store) g.u[vec(n) + j] = acc;
  }
}

template <typename T, int M>
constexpr size_t blk2_lds() {
  return (size_t)4 * (kWave / M) * blk2_tile<T, M>() * sizeof(T);
}

// the two-ended tiled form: 2M lanes per problem (half 0 = lanes 0..M-1 of
// the group, half 1 = lanes M..2M-1), 64/(2M) problems per wave, one tile
// per half; work-group `block` of 256 threads
template <typename T, int M, int TRANS>
__device__ __forceinline__ void blk2_tw_block(const BlkArgs<T> &g, long long block, T *smem) {
  constexpr int G2 = kWave / (2 * M);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int grp = lane / (2 * M);
  const int in = lane - grp * 2 * M;
  const int h = in / M, j = in - h * M;
  const bool lane_ok = grp < G2;
  const long long p = (block * 4 + wave) * G2 + grp;
  const bool store = lane_ok && p < g.P;
  // idle lanes past the last whole group read the last group's first tile,
  // solve problem P-1 again and write nothing
  T *tile = smem + ((wave * G2 + (lane_ok ? grp : G2 - 1)) * 2 + (lane_ok ? h : 0)) *
                       blk2_tile<T, M>();
  blk2_problem_tw<T, M, TRANS>(g, std::min<long long>(p, g.P - 1), store, lane_ok, h, j,
                               min(grp * 2 * M + h * M, kWave - 1),
                               lane_ok ? grp * 2 * M + (1 - h) * M + j : lane, tile);
}

template <typename T, int M, int TRANS>
__global__ void __launch_bounds__(256, blk2_min_waves(M)) k_blk2_tw(BlkArgs<T> g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  blk2_tw_block<T, M, TRANS>(g, blockIdx.x, reinterpret_cast<T *>(smem_raw));
}

template <typename T, int M>
constexpr size_t blk2_tw_lds() {
  return (size_t)4 * (kWave / (2 * M)) * 2 * blk2_tile<T, M>() * sizeof(T);
}

// ---- mixed block sizes in one launch ---------------------------------------
constexpr int kMaxGroups = 16;
struct MixedArgs {
  BlkArgs<double> g[kMaxGroups];
  int M[kMaxGroups];
  int L[kMaxGroups];  // lanes per problem (k_blk), M for the tiled kernel, or kTwistTiled
  long long first_block[kMaxGroups + 1];
  int ngroups;
};

// The lane count per block size of the mixed kernel (and the default).
__host__ __device__ constexpr int default_lanes(int M) {
  return M == 1 ? 1 : (M == 2 ? 2 : (M == 3 ? 3 : (M == 4 ? 4 : (M == 8 ? 8 : M))));
}

__global__ void __launch_bounds__(256) k_blk_mixed(MixedArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *smem = reinterpret_cast<double *>(smem_raw);
  const long long b = blockIdx.x;
  int gi = 0;
  while (gi + 1 < a.ngroups && b >= a.first_block[gi + 1]) gi++;
  const long long lb = b - a.first_block[gi];
  switch (a.M[gi]) {
#define MIXED_CASE(MM)                                                            \
  case MM:                                                                        \
    if constexpr (MM <= 3) {                                                      \
      if (a.L[gi] == 1) {                                                         \
        blk_block<double, MM, 1, 0>(a.g[gi], lb);                                 \
        break;                                                                    \
      }                                                                           \
    }                                                                             \
    if constexpr (MM == 8) {                                                      \
      if (a.L[gi] == 4) {                                                         \
        blk_block<double, MM, 4, 0>(a.g[gi], lb);                                 \
        break;                                                                    \
      }                                                                           \
    }                                                                             \
    if constexpr (blk2tw_sizes(MM)) {                                             \
      if (a.L[gi] == kTwistTiled) {                                               \
        blk2_tw_block<double, MM, 0>(a.g[gi], lb, smem);                          \
        break;                                                                    \
      }                                                                           \
    }                                                                             \
    if constexpr (blk2_sizes(MM))                                                 \
      blk2_block<double, MM, 0>(a.g[gi], lb, smem);                               \
    else                                                                          \
      blk_block<double, MM, default_lanes(MM), 0>(a.g[gi], lb);                   \
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

// layout: kTiled = the tiled kernel (k_blk2, M = 3..6), L > 0 = k_blk with
// L lanes per problem, kTwist + L = k_blk_tw with L lanes per half,
// kTwistTiled = k_blk2_tw (M = 3..8), 0 = the mixed launch's layout for M
constexpr int kTiled = -1;
constexpr int kTwist = 100;

template <typename T, int M, int TRANS>
hipError_t launch_m(const BlkArgs<T> &g, int lanes, hipStream_t st) {
  constexpr int Ld = default_lanes(M);
  const int L = lanes > 0 ? lanes : Ld;
  if constexpr (blk2tw_sizes(M)) {
    if (lanes == kTwistTiled) {
      const long long G = 4 * (kWave / (2 * M));
      const size_t lds = blk2_tw_lds<T, M>();
      hipLaunchKernelGGL((k_blk2_tw<T, M, TRANS>), dim3((unsigned)((g.P + G - 1) / G)), dim3(256),
                         lds, st, g);
      return hipGetLastError();
    }
  }
  if constexpr (blk2_sizes(M)) {
    // one lane per block column, tiles through LDS
    if (lanes == 0 || lanes == kTiled) {
      const long long G = 4 * (kWave / M);
      const size_t lds = blk2_lds<T, M>();
      hipLaunchKernelGGL((k_blk2<T, M, TRANS>), dim3((unsigned)((g.P + G - 1) / G)), dim3(256),
                         lds, st, g);
      return hipGetLastError();
    }
  }
#define TCASE(LL)                                                                \
  if constexpr (M % LL == 0)                                                     \
    if (lanes == kTwist + LL) {                                                  \
      hipLaunchKernelGGL((k_blk_tw<T, M, LL, TRANS>),                            \
                         dim3((unsigned)blk_tw_blocks<M, LL>(g.P)), dim3(256), 0, st, g); \
      return hipGetLastError();                                                  \
    }
  TCASE(1)
  TCASE(2)
  TCASE(4)
#undef TCASE
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
  if (L == kTwistTiled) {
    if (!blk2tw_sizes(M))
      return fail(TRID_STATUS_INVALID_VALUE, "the two-ended tiled layout needs blkdim 3..8");
    return TRID_STATUS_SUCCESS;
  }
  const int Lh = L > kTwist ? L - kTwist : L;  // lanes per half of the two-ended form
  if (L && M % Lh != 0) return fail(TRID_STATUS_INVALID_VALUE, "forced lanes do not divide blkdim");
  return TRID_STATUS_SUCCESS;
}

// Layout of a batch without tuning, from the tuner's sweeps on MI355X at
// N = 128 over P = 4096 / 16384 / 65536 (one column each below: P < 8192,
// P < 65536, larger), M = 2..10, FP64 and FP32.  A two-ended form wins
// almost all of them (FP64, ms of the tuning runs: M = 2, P = 4096: 0.201
// against 0.370 with one lane; M = 4, P = 65536: 3.82 two-ended tiled
// against 4.00 tiled; M = 5, P = 4096: 0.571 against 1.089 tiled; M = 3,
// P = 65536: 2.32 with one lane per half against 2.59 with one lane; FP32
// M = 5, P = 65536: 3.00 against 4.05 tiled; M = 8, P = 65536: 13.6
// two-ended tiled against 15.1 with four lanes per half; M = 7, P = 65536:
// 11.9 against 16.6 with seven lanes); one lane per problem keeps FP64
// M = 2 at P = 65536 (1.038 against 1.063), M lanes FP64 M = 9, 10 (27.2 /
// 32.0 ms; the two-ended tiled kernel stops at M = 8).  0: M lanes (M = 9,
// 10) or default_lanes.
int default_layout(int M, long long P, int ts) {
  constexpr int W1 = kTwist + 1, W2 = kTwist + 2, W4 = kTwist + 4, TT = kTwistTiled;
  static constexpr int k64[11][3] = {{0, 0, 0},    {0, 0, 0},    {W1, W1, 1},  {TT, TT, W1},
                                     {TT, TT, TT}, {TT, TT, W1}, {TT, W2, TT}, {TT, TT, TT},
                                     {TT, TT, TT}, {0, 0, 0},    {0, 0, 0}};
  static constexpr int k32[11][3] = {{0, 0, 0},    {0, 0, 0},    {W2, W2, W1}, {W1, W1, W1},
                                     {TT, TT, W1}, {TT, W1, W1}, {TT, W2, W1}, {TT, TT, W1},
                                     {TT, TT, TT}, {0, 0, W1},   {0, W2, W2}};
  const int band = P < 8192 ? 0 : (P < 65536 ? 1 : 2);
  const int t = M >= 1 && M <= 10 ? (ts == 8 ? k64 : k32)[M][band] : 0;
  if (t) return t;
  return blk2_sizes(M) ? kTiled : default_lanes(M);
}

std::mutex g_blk_mu;
// tuned layout per (precision, transposed, M, N, P)
std::map<std::tuple<int, int, int, int, int>, int> g_blk_tuned;

// The layout of a call: forced lanes, else the tuned layout (tuned at the
// first call of a shape with autotuning on and the stream not capturing: the
// call's own arrays, which the solve only reads, or only overwrites, are
// used for the timing; a call whose u is its d is not tuned), else the
// default.
template <typename T, int TRANS>
int choose_layout(const BlkArgs<T> &g, int M, hipStream_t st) {
  const int forced = g_force_lanes.load();
  if (forced) return forced;
  const int def = default_layout(M, g.P, (int)sizeof(T));
  if (autotune_mode() != 1) return def;
  const auto key = std::make_tuple((int)sizeof(T), TRANS, M, g.N, g.P);
  {
    std::lock_guard<std::mutex> lk(g_blk_mu);
    auto it = g_blk_tuned.find(key);
    if (it != g_blk_tuned.end()) return it->second;
  }
  // (timing runs the solve several times: not when its output overwrites
  // its right-hand side, and not inside a caller's capture)
  if (capturing(st) || static_cast<const void *>(g.u) == static_cast<const void *>(g.d))
    return def;
  std::vector<int> cands{def};
  if (blk2_sizes(M) && def != kTiled) cands.push_back(kTiled);
  for (int L = 1; L <= M && L <= 8; L++)
    if (M % L == 0 && L != def) cands.push_back(L);
  // the two-ended forms: 1, 2 or 4 lanes per half, and tiled
  for (int L : {1, 2, 4})
    if (M % L == 0 && kTwist + L != def) cands.push_back(kTwist + L);
  if (blk2tw_sizes(M) && def != kTwistTiled) cands.push_back(kTwistTiled);
  std::vector<float> tm;
  auto run = [&](int i) { return launch_blk<T, TRANS>(g, M, cands[i], st); };
  const int i = pick_fastest((int)cands.size(), [](int) { return true; }, run, st, &tm, 3, 2);
  const int pick = i >= 0 ? cands[i] : def;
  const char *log = std::getenv("TRIDHIP_TUNE_LOG");
  if (log && log[0] == '1')
    for (size_t c = 0; c < cands.size(); c++)
      fprintf(stderr, "[tridhip tune] block M=%d N=%d P=%d %s layout %d: %.4f ms\n", M, g.N, g.P,
              TRANS ? "transposed" : "", cands[c], tm[c]);
  std::lock_guard<std::mutex> lk(g_blk_mu);
  g_blk_tuned[key] = pick;
  return pick;
}

template <typename T>
tridStatus_t btsv(const T *A, const T *B, const T *C, const T *d, T *Cs, T *ds,
                  T *u, int N, int P, int M, int sync) {
  clear_error();
  tridStatus_t s = check_blk(A, B, C, d, u, N, P, M);
  if (s) return s;
  if (N == 0 || P == 0) return TRID_STATUS_SUCCESS;
  hipStream_t st = current_stream();
  if (!Cs || !ds) {
    size_t mat = (size_t)N * P * M * M, v = (size_t)N * P * M;
    T *w = static_cast<T *>(workspace((mat + v) * sizeof(T), st));
    if (!w) return TRID_STATUS_ALLOC_FAILED;  // reason recorded by workspace()
    Cs = w;
    ds = w + mat;
  }
  BlkArgs<T> g{A, B, C, d, Cs, ds, u, N, P};
  s = check_hip(launch_blk<T, 0>(g, M, choose_layout<T, 0>(g, M, st), st),
                "block kernel launch");
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
    // the two-ended tiled layout where it is the untuned default; else one
    // lane per problem for M <= 3 and four for M = 8 in batches that fill
    // the chip with them (the one-ended layouts that beat M lanes at
    // P = 65536), else default_lanes
    const bool big = G.P >= 65536;
    int L = big && G.M <= 3 ? 1 : (big && G.M == 8 ? 4 : default_lanes(G.M));
    if (default_layout(G.M, G.P, 8) == kTwistTiled) L = kTwistTiled;
    a.L[k] = L;
    const long long per = L == kTwistTiled ? 4LL * (kWave / (2 * G.M)) : 4LL * (kWave / L);
    nb += (G.P + per - 1) / per;
    k++;
  }
  a.ngroups = k;
  a.first_block[k] = nb;
  if (nb == 0) return TRID_STATUS_SUCCESS;
  // LDS: the largest tile set of the batch's block sizes
  size_t lds = 0;
  for (int i = 0; i < k; i++) {
    const int Mi = a.M[i], tile = 3 * Mi * Mi + 2 * Mi;
    if (a.L[i] == kTwistTiled)
      lds = std::max(lds, (size_t)4 * (kWave / (2 * Mi)) * 2 * tile * 8);
    else if (blk2_sizes(Mi))
      lds = std::max(lds, (size_t)4 * (kWave / Mi) * tile * 8);
  }
  hipStream_t st = current_stream();
  hipLaunchKernelGGL(k_blk_mixed, dim3((unsigned)nb), dim3(256), lds, st, a);
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
  hipStream_t st = current_stream();
  double *w = static_cast<double *>(workspace((mat + v) * sizeof(double), st));
  if (!w) return TRID_STATUS_ALLOC_FAILED;  // reason recorded by workspace()
  // lambda = A^{-T} g  (written straight into grad_d)
  BlkArgs<double> ga{A, B, C, g, w, w + mat, grad_d, N, P};
  s = check_hip(launch_blk<double, 1>(ga, M, choose_layout<double, 1>(ga, M, st), st),
                "adjoint kernel launch");
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
  if ((lanes < 0 || lanes > 10) && lanes != kTwist + 1 && lanes != kTwist + 2 &&
      lanes != kTwist + 4 && lanes != kTwistTiled)
    return fail(TRID_STATUS_INVALID_VALUE, "lanes must be in [0, 10] or 101, 102, 104, 200");
  g_force_lanes.store(lanes);
  return TRID_STATUS_SUCCESS;
}
