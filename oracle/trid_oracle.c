/*
 * trid_oracle.c -- CPU oracle for the tridsolver_amd test-suite.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only the tests, the
 * smoke() entry point and bench.py's cpu_baseline leg load the shared
 * library built from it (oracle/liboracle.so).  The GPU library never links,
 * loads or calls it, and nothing here is on the product path.
 *
 * Every routine restates, in plain C, the reference algorithm it cites
 * (paths relative to the TobyFlynn/tridsolver tree).  The arithmetic order
 * of the scalar Thomas recurrence is kept exactly as in the reference
 * (reciprocal of the pivot, then multiplications), so that the oracle's ADI
 * reproduces the reference's own apps/adi/src/adi_orig.c bit for bit when
 * both are built with gcc -O2 and no FMA contraction.
 *
 * Parity pins (which check fixes which function) are listed in the header
 * of tests/test_oracle.py.
 *
 * Build: oracle/Makefile  (gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXDIM 8

/* status codes, numerically equal to tridStatus_t of
 * scalar/include/trid_common.h:65-76 */
#define OR_SUCCESS 0
#define OR_INVALID_VALUE 3
#define OR_ALLOC_FAILED 2

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------
 * Scalar Thomas for one system.
 * Restates scalar/src/cpu/trid_cpu.cpp:360-401 (trid_scalar<INC>): forward
 * sweep with bb = 1/(b_i - a_i c'_{i-1}), c'_i = bb c_i, d'_i = bb (d_i - a_i
 * d'_{i-1}); backward sweep x_i = d'_i - c'_i x_{i+1}.  a_0 and c_{N-1} are
 * never read.  Non-Inc writes x into d, Inc adds x into u (d untouched).
 * Generalised to one stride per array (for the per-array padded API of
 * scalar/include/trid_cuda.h:60-67).  c2/d2 are caller scratch of length N
 * (the reference caps N at N_MAX; the oracle has no cap).
 * ---------------------------------------------------------------------- */
#define DEFINE_THOMAS(T, SUF)                                                  \
  static void thomas_##SUF(const T *a, long sa, const T *b, long sb,          \
                           const T *c, long sc, T *d, long sd, T *u, long su, \
                           int N, int inc, T *c2, T *d2) {                    \
    T aa, bb, cc, dd;                                                          \
    int i;                                                                     \
    if (N <= 0) return;                                                        \
    bb = (T)1.0 / b[0];                                                        \
    cc = bb * c[0];                                                            \
    dd = bb * d[0];                                                            \
    c2[0] = cc;                                                                \
    d2[0] = dd;                                                                \
    for (i = 1; i < N; i++) {                                                  \
      aa = a[i * sa];                                                          \
      bb = b[i * sb] - aa * cc;                                                \
      dd = d[i * sd] - aa * dd;                                                \
      bb = (T)1.0 / bb;                                                        \
      cc = bb * c[i * sc];                                                     \
      dd = bb * dd;                                                            \
      c2[i] = cc;                                                              \
      d2[i] = dd;                                                              \
    }                                                                          \
    if (inc)                                                                   \
      u[(long)(N - 1) * su] += dd;                                             \
    else                                                                       \
      d[(long)(N - 1) * sd] = dd;                                              \
    for (i = N - 2; i >= 0; i--) {                                             \
      dd = d2[i] - c2[i] * dd;                                                 \
      if (inc)                                                                 \
        u[i * su] += dd;                                                       \
      else                                                                     \
        d[i * sd] = dd;                                                        \
    }                                                                          \
  }
DEFINE_THOMAS(double, d)
DEFINE_THOMAS(float, s)

/* ------------------------------------------------------------------------
 * Multi-dimensional strided batch.
 * Semantics of tridXmtsvStridedBatch[Inc][Padded] (scalar/include/trid_cuda.h
 * :56-84, scalar/include/trid_cpu.h:52-82; dispatch trid_cpu.cpp:406-531 and
 * trid_cuda.cu:191-445): ndim-dimensional arrays with dimension 0 fastest,
 * element (i_0..i_{n-1}) at sum_k i_k * prod_{j<k} pads[j]; one independent
 * system along `solvedim` for every index of the other dimensions.  The
 * reference CPU path always assumes 3 dims (trid_cpu.cpp:414-415); the
 * oracle handles any ndim <= 8 and one pads vector per array.
 * ---------------------------------------------------------------------- */
typedef struct {
  int nother;                 /* number of non-solve dims */
  long osize[OR_MAXDIM];      /* their sizes, fastest first */
  long ostride[5][OR_MAXDIM]; /* per array: element stride of each */
  long sstride[5];            /* per array: stride along solvedim */
  long nsys;
  int N;
} or_layout;

static int or_make_layout(or_layout *L, int ndim, int solvedim, const int *dims,
                          const int *const pads[5]) {
  int k, arr;
  if (ndim < 1 || ndim > OR_MAXDIM || solvedim < 0 || solvedim >= ndim)
    return OR_INVALID_VALUE;
  for (k = 0; k < ndim; k++)
    if (dims[k] < 0) return OR_INVALID_VALUE;
  L->N = dims[solvedim];
  L->nsys = 1;
  L->nother = 0;
  for (k = 0; k < ndim; k++) {
    if (k == solvedim) continue;
    L->osize[L->nother] = dims[k];
    L->nsys *= dims[k];
    L->nother++;
  }
  for (arr = 0; arr < 5; arr++) {
    long cum = 1;
    int o = 0;
    const int *p = pads[arr];
    for (k = 0; k < ndim; k++) {
      if (p && p[k] < dims[k]) return OR_INVALID_VALUE;
      if (k == solvedim)
        L->sstride[arr] = cum;
      else
        L->ostride[arr][o++] = cum;
      cum *= p ? p[k] : dims[k];
    }
  }
  return OR_SUCCESS;
}

static void or_sys_offsets(const or_layout *L, long q, long off[5]) {
  int k, arr;
  for (arr = 0; arr < 5; arr++) off[arr] = 0;
  for (k = 0; k < L->nother; k++) {
    long ik = q % L->osize[k];
    q /= L->osize[k];
    for (arr = 0; arr < 5; arr++) off[arr] += ik * L->ostride[arr][k];
  }
}

#define DEFINE_MTSV(T, SUF)                                                    \
  int oracle_mtsv_##SUF(const T *a, const int *a_pads, const T *b,            \
                        const int *b_pads, const T *c, const int *c_pads,     \
                        T *d, const int *d_pads, T *u, const int *u_pads,     \
                        int ndim, int solvedim, const int *dims, int inc) {   \
    or_layout L;                                                               \
    const int *pads[5] = {a_pads, b_pads, c_pads, d_pads,                     \
                          u_pads ? u_pads : d_pads};                          \
    int st = or_make_layout(&L, ndim, solvedim, dims, pads);                   \
    int err = 0;                                                               \
    if (st) return st;                                                         \
    if (inc && !u) return OR_INVALID_VALUE;                                    \
    if (L.N == 0 || L.nsys == 0) return OR_SUCCESS;                            \
    _Pragma("omp parallel") {                                                  \
      T *c2 = (T *)malloc(sizeof(T) * (size_t)L.N);                            \
      T *d2 = (T *)malloc(sizeof(T) * (size_t)L.N);                            \
      long q;                                                                  \
      if (!c2 || !d2) {                                                        \
        _Pragma("omp atomic write") err = 1;                                   \
      } else {                                                                 \
        _Pragma("omp for schedule(static)") for (q = 0; q < L.nsys; q++) {     \
          long o[5];                                                           \
          or_sys_offsets(&L, q, o);                                            \
          thomas_##SUF(a + o[0], L.sstride[0], b + o[1], L.sstride[1],        \
                       c + o[2], L.sstride[2], d + o[3], L.sstride[3],        \
                       inc ? u + o[4] : NULL, L.sstride[4], L.N, inc, c2,     \
                       d2);                                                    \
        }                                                                      \
      }                                                                        \
      free(c2);                                                                \
      free(d2);                                                                \
    }                                                                          \
    return err ? OR_ALLOC_FAILED : OR_SUCCESS;                                 \
  }
DEFINE_MTSV(double, d)
DEFINE_MTSV(float, s)

/* One system with a single stride: restates the exported trid_scalarD/S
 * (scalar/include/trid_cpu.h:84-98, trid_cpu.cpp:597-602). */
void oracle_trid_scalar_d(const double *a, const double *b, const double *c,
                          double *d, double *u, int N, int stride, int inc) {
  double *c2 = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  double *d2 = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  thomas_d(a, stride, b, stride, c, stride, d, stride, u, stride, N, inc, c2,
           d2);
  free(c2);
  free(d2);
}

void oracle_trid_scalar_s(const float *a, const float *b, const float *c,
                          float *d, float *u, int N, int stride, int inc) {
  float *c2 = (float *)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1));
  float *d2 = (float *)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1));
  thomas_s(a, stride, b, stride, c, stride, d, stride, u, stride, N, inc, c2,
           d2);
  free(c2);
  free(d2);
}

/* ------------------------------------------------------------------------
 * Ragged batch (no reference counterpart): system k occupies elements
 * [offsets[k], offsets[k+1]) of a, b, c, d (and u), stride 1.  Each system
 * is solved exactly as the reference's 1-D entry trid_scalarD
 * (scalar/src/cpu/trid_cpu.cpp:597-602) would solve it.
 * ---------------------------------------------------------------------- */
#define DEFINE_RAGGED(T, SUF)                                                  \
  int oracle_ragged_##SUF(const T *a, const T *b, const T *c, T *d, T *u,     \
                          const int64_t *offsets, long nsys, int inc) {       \
    long k, maxlen = 0;                                                        \
    int err = 0;                                                               \
    for (k = 0; k < nsys; k++) {                                               \
      long len = (long)(offsets[k + 1] - offsets[k]);                          \
      if (len < 0) return OR_INVALID_VALUE;                                    \
      if (len > maxlen) maxlen = len;                                          \
    }                                                                          \
    if (inc && !u) return OR_INVALID_VALUE;                                    \
    _Pragma("omp parallel") {                                                  \
      T *c2 = (T *)malloc(sizeof(T) * (size_t)(maxlen + 1));                   \
      T *d2 = (T *)malloc(sizeof(T) * (size_t)(maxlen + 1));                   \
      if (!c2 || !d2) {                                                        \
        _Pragma("omp atomic write") err = 1;                                   \
      } else {                                                                 \
        _Pragma("omp for schedule(dynamic, 64)") for (k = 0; k < nsys; k++) { \
          long o = (long)offsets[k];                                           \
          int len = (int)(offsets[k + 1] - offsets[k]);                        \
          thomas_##SUF(a + o, 1, b + o, 1, c + o, 1, d + o, 1,                 \
                       inc ? u + o : NULL, 1, len, inc, c2, d2);               \
        }                                                                      \
      }                                                                        \
      free(c2);                                                                \
      free(d2);                                                                \
    }                                                                          \
    return err ? OR_ALLOC_FAILED : OR_SUCCESS;                                 \
  }
DEFINE_RAGGED(double, d)
DEFINE_RAGGED(float, s)

/* ------------------------------------------------------------------------
 * Adjoint of the batched solve (no reference counterpart).
 * Forward: A x = d with A = tridiag(a, b, c) per system.  Given the upstream
 * gradient g = dL/dx, lambda = A^{-T} g is computed by the same Thomas
 * recurrence applied to the transposed system, whose sub-diagonal is
 * c_{i-1}, diagonal b_i and super-diagonal a_{i+1}; then
 *   grad_d = lambda, grad_a_i = -lambda_i x_{i-1}, grad_b_i = -lambda_i x_i,
 *   grad_c_i = -lambda_i x_{i+1}   (terms outside the system are zero).
 * All arrays share one pads vector (the strided-batch layout above).
 * ---------------------------------------------------------------------- */
int oracle_adjoint_d(const double *a, const double *b, const double *c,
                     const double *x, const double *g, double *grad_a,
                     double *grad_b, double *grad_c, double *grad_d, int ndim,
                     int solvedim, const int *dims, const int *pads) {
  or_layout L;
  const int *pp[5] = {pads, pads, pads, pads, pads};
  int st = or_make_layout(&L, ndim, solvedim, dims, pp);
  int err = 0;
  if (st) return st;
  if (L.N == 0 || L.nsys == 0) return OR_SUCCESS;
#pragma omp parallel
  {
    int N = L.N;
    double *at = (double *)malloc(sizeof(double) * (size_t)N);
    double *ct = (double *)malloc(sizeof(double) * (size_t)N);
    double *bt = (double *)malloc(sizeof(double) * (size_t)N);
    double *lam = (double *)malloc(sizeof(double) * (size_t)N);
    double *c2 = (double *)malloc(sizeof(double) * (size_t)N);
    double *d2 = (double *)malloc(sizeof(double) * (size_t)N);
    long q;
    if (!at || !ct || !bt || !lam || !c2 || !d2) {
#pragma omp atomic write
      err = 1;
    } else {
#pragma omp for schedule(static)
      for (q = 0; q < L.nsys; q++) {
        long o[5], s = L.sstride[0];
        int i;
        or_sys_offsets(&L, q, o);
        for (i = 0; i < N; i++) {
          at[i] = i > 0 ? c[o[0] + (long)(i - 1) * s] : 0.0;
          bt[i] = b[o[0] + (long)i * s];
          ct[i] = i < N - 1 ? a[o[0] + (long)(i + 1) * s] : 0.0;
          lam[i] = g[o[0] + (long)i * s];
        }
        thomas_d(at, 1, bt, 1, ct, 1, lam, 1, NULL, 1, N, 0, c2, d2);
        for (i = 0; i < N; i++) {
          long p = o[0] + (long)i * s;
          double xm = i > 0 ? x[p - s] : 0.0;
          double xp = i < N - 1 ? x[p + s] : 0.0;
          grad_d[p] = lam[i];
          grad_a[p] = -lam[i] * xm;
          grad_b[p] = -lam[i] * x[p];
          grad_c[p] = -lam[i] * xp;
        }
      }
    }
    free(at);
    free(ct);
    free(bt);
    free(lam);
    free(c2);
    free(d2);
  }
  return err ? OR_ALLOC_FAILED : OR_SUCCESS;
}

/* ------------------------------------------------------------------------
 * Block tridiagonal solve (block Thomas).
 * Restates block/src/blktrid_cpu.cpp:388-504 (gold_blkThomas_solve):
 *   C'_0 = B_0^{-1} C_0, d'_0 = B_0^{-1} d_0,
 *   X_n = B_n - A_n C'_{n-1}, C'_n = X_n^{-1} C_n,
 *   d'_n = X_n^{-1} (d_n - A_n d'_{n-1}),
 *   u_{N-1} = d'_{N-1}, u_n = d'_n - C'_n u_{n+1}.
 * with the matrix product written as the true product sum_k M1[i,k] M2[k,j];
 * the reference's gold_MMmultiply reads M2[i,j] instead of M2[k,j]
 * (blktrid_cpu.cpp:310-317,323), which is only correct when M1 is diagonal.
 * Inverses follow gold_invert (blktrid_cpu.cpp:48-206): closed forms for 2x2
 * and 3x3, Gauss-Jordan with partial (row) pivoting otherwise.
 * Blocks are M x M row-major.  layout 0 = GPU row-interleaved layout of
 * block/src/blktrid_gpu.cuh:265-431 (block row n of problem p at
 * (n*P + p)*M*M, vector at (n*P + p)*M); layout 1 = CPU system-major layout
 * of blktrid_cpu.cpp:396-400 ((p*N + n)*M*M).
 * ---------------------------------------------------------------------- */
#define BLK_MAXM 16

static void blk_inv(const double *Min, double *I, int M) {
  double a[BLK_MAXM * BLK_MAXM];
  int i, j, p, q;
  if (M == 1) {
    I[0] = 1.0 / Min[0];
    return;
  }
  if (M == 2) {
    double det = Min[0] * Min[3] - Min[1] * Min[2];
    I[0] = Min[3] / det;
    I[1] = -Min[1] / det;
    I[2] = -Min[2] / det;
    I[3] = Min[0] / det;
    return;
  }
  if (M == 3) {
    const double *m = Min;
    double A = m[4] * m[8] - m[5] * m[7], B = -(m[3] * m[8] - m[5] * m[6]),
           C = m[3] * m[7] - m[4] * m[6];
    double D = -(m[1] * m[8] - m[2] * m[7]), E = m[0] * m[8] - m[2] * m[6],
           F = -(m[0] * m[7] - m[1] * m[6]);
    double G = m[1] * m[5] - m[2] * m[4], H = -(m[0] * m[5] - m[2] * m[3]),
           K = m[0] * m[4] - m[1] * m[3];
    double det = m[0] * A + m[1] * B + m[2] * C;
    I[0] = A / det; I[1] = D / det; I[2] = G / det;
    I[3] = B / det; I[4] = E / det; I[5] = H / det;
    I[6] = C / det; I[7] = F / det; I[8] = K / det;
    return;
  }
  memcpy(a, Min, sizeof(double) * (size_t)(M * M));
  for (i = 0; i < M; i++)
    for (j = 0; j < M; j++) I[i * M + j] = (i == j) ? 1.0 : 0.0;
  for (i = 0; i < M; i++) {
    double best = fabs(a[i * M + i]), inv;
    p = i;
    for (j = i + 1; j < M; j++)
      if (fabs(a[j * M + i]) > best) {
        best = fabs(a[j * M + i]);
        p = j;
      }
    if (p != i)
      for (j = 0; j < M; j++) {
        double t = a[i * M + j];
        a[i * M + j] = a[p * M + j];
        a[p * M + j] = t;
        t = I[i * M + j];
        I[i * M + j] = I[p * M + j];
        I[p * M + j] = t;
      }
    inv = 1.0 / a[i * M + i];
    for (j = 0; j < M; j++) {
      a[i * M + j] *= inv;
      I[i * M + j] *= inv;
    }
    for (q = 0; q < M; q++) {
      double f;
      if (q == i) continue;
      f = a[q * M + i];
      for (j = 0; j < M; j++) {
        a[q * M + j] -= f * a[i * M + j];
        I[q * M + j] -= f * I[i * M + j];
      }
    }
  }
}

static void mm(const double *X, const double *Y, double *Z, int M) {
  int i, j, k;
  for (i = 0; i < M; i++)
    for (j = 0; j < M; j++) {
      double s = 0.0;
      for (k = 0; k < M; k++) s += X[i * M + k] * Y[k * M + j];
      Z[i * M + j] = s;
    }
}

static void mv(const double *X, const double *v, double *w, int M) {
  int i, k;
  for (i = 0; i < M; i++) {
    double s = 0.0;
    for (k = 0; k < M; k++) s += X[i * M + k] * v[k];
    w[i] = s;
  }
}

int oracle_btsv_d(const double *A, const double *B, const double *C,
                  const double *d, double *u, int N, int P, int M,
                  int layout) {
  int err = 0;
  if (M < 1 || M > BLK_MAXM || N < 0 || P < 0) return OR_INVALID_VALUE;
  if (N == 0 || P == 0) return OR_SUCCESS;
#pragma omp parallel
  {
    int MM = M * M;
    double *Cs = (double *)malloc(sizeof(double) * (size_t)N * (size_t)MM);
    double *ds = (double *)malloc(sizeof(double) * (size_t)N * (size_t)M);
    double X[BLK_MAXM * BLK_MAXM], Xi[BLK_MAXM * BLK_MAXM],
        T[BLK_MAXM * BLK_MAXM], v[BLK_MAXM], w[BLK_MAXM];
    long p;
    if (!Cs || !ds) {
#pragma omp atomic write
      err = 1;
    } else {
#pragma omp for schedule(static)
      for (p = 0; p < P; p++) {
        int n, i, k;
#define BIDX(n) (layout == 0 ? ((long)(n) * P + p) : ((long)p * N + (n)))
        const double *B0 = B + BIDX(0) * MM;
        blk_inv(B0, Xi, M);
        mm(Xi, C + BIDX(0) * MM, Cs, M);
        mv(Xi, d + BIDX(0) * M, ds, M);
        for (n = 1; n < N; n++) {
          const double *An = A + BIDX(n) * MM;
          const double *Bn = B + BIDX(n) * MM;
          mm(An, Cs + (long)(n - 1) * MM, T, M);
          for (k = 0; k < MM; k++) X[k] = Bn[k] - T[k];
          blk_inv(X, Xi, M);
          mm(Xi, C + BIDX(n) * MM, Cs + (long)n * MM, M);
          mv(An, ds + (long)(n - 1) * M, v, M);
          for (i = 0; i < M; i++) v[i] = d[BIDX(n) * M + i] - v[i];
          mv(Xi, v, ds + (long)n * M, M);
        }
        for (i = 0; i < M; i++)
          u[BIDX(N - 1) * M + i] = ds[(long)(N - 1) * M + i];
        for (n = N - 2; n >= 0; n--) {
          mv(Cs + (long)n * MM, u + BIDX(n + 1) * M, w, M);
          for (i = 0; i < M; i++)
            u[BIDX(n) * M + i] = ds[(long)n * M + i] - w[i];
        }
#undef BIDX
      }
    }
    free(Cs);
    free(ds);
  }
  return err ? OR_ALLOC_FAILED : OR_SUCCESS;
}

/* ------------------------------------------------------------------------
 * ADI heat diffusion time steps on an nx*ny*nz grid (dim 0 = x fastest).
 * Restates apps/adi/src/adi_orig.c:147-243 (adi_cpu): per step the 7-point
 * right-hand side d = lambda*(u[-1]+u[+1]+u[-nx]+u[+nx]+u[-nxny]+u[+nxny]
 * - 6u) at interior points, Dirichlet rows (a=c=0, b=1, d=0) on the
 * boundary, Thomas sweeps along x and y into du, then along z and u += du.
 * mode 0 is the reference's coefficients a=c=-lambda/2, b=1+lambda
 * (adi_orig.c:189-191); mode 1 is a=c=-lambda, b=1+2 lambda.  The initial
 * condition is the caller's u (the reference sets 1 on the boundary and 0
 * inside, adi_orig.c:297-307).  oracle_adi_ws_* takes the four work arrays
 * (du, a, b, c) as one caller buffer of 4*nx*ny*nz values, so that timed
 * calls do not allocate; oracle_adi_* allocates it per call.
 * ---------------------------------------------------------------------- */
#define DEFINE_ADI(T, SUF)                                                     \
  int oracle_adi_ws_##SUF(int mode, T lambda, T *u, int nx, int ny, int nz,    \
                          int iters, T *ws) {                                  \
    long n = (long)nx * ny * nz, ind;                                          \
    int it, i, j, k;                                                           \
    int dims[3] = {nx, ny, nz};                                                \
    T *du, *a, *b, *c;                                                         \
    T ai, bi, ci;                                                              \
    if (nx < 1 || ny < 1 || nz < 1 || !ws) return OR_INVALID_VALUE;            \
    if (mode == 0) {                                                           \
      ai = (T)-0.5f * lambda;                                                  \
      bi = (T)1.0f + lambda;                                                   \
      ci = (T)-0.5f * lambda;                                                  \
    } else {                                                                   \
      ai = -lambda;                                                            \
      bi = (T)1.0f + (T)2.0f * lambda;                                         \
      ci = -lambda;                                                            \
    }                                                                          \
    du = ws;                                                                   \
    a = ws + n;                                                                \
    b = ws + 2 * n;                                                            \
    c = ws + 3 * n;                                                            \
    for (it = 0; it < iters; it++) {                                           \
      _Pragma("omp parallel for collapse(2) private(i, ind)")                  \
      for (k = 0; k < nz; k++)                                                 \
        for (j = 0; j < ny; j++)                                               \
          for (i = 0; i < nx; i++) {                                           \
            ind = (long)k * nx * ny + (long)j * nx + i;                        \
            if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 ||    \
                k == nz - 1) {                                                 \
              du[ind] = (T)0.0f;                                               \
              a[ind] = (T)0.0f;                                                \
              b[ind] = (T)1.0f;                                                \
              c[ind] = (T)0.0f;                                                \
            } else {                                                           \
              du[ind] = lambda * (u[ind - 1] + u[ind + 1] + u[ind - nx] +      \
                                  u[ind + nx] + u[ind - (long)nx * ny] +       \
                                  u[ind + (long)nx * ny] - (T)6.0f * u[ind]);  \
              a[ind] = ai;                                                     \
              b[ind] = bi;                                                     \
              c[ind] = ci;                                                     \
            }                                                                  \
          }                                                                    \
      oracle_mtsv_##SUF(a, dims, b, dims, c, dims, du, dims, u, dims, 3, 0,    \
                        dims, 0);                                              \
      oracle_mtsv_##SUF(a, dims, b, dims, c, dims, du, dims, u, dims, 3, 1,    \
                        dims, 0);                                              \
      oracle_mtsv_##SUF(a, dims, b, dims, c, dims, du, dims, u, dims, 3, 2,    \
                        dims, 1);                                              \
    }                                                                          \
    return OR_SUCCESS;                                                         \
  }                                                                            \
  int oracle_adi_##SUF(int mode, T lambda, T *u, int nx, int ny, int nz,       \
                       int iters) {                                            \
    long n = (long)nx * ny * nz;                                               \
    T *ws;                                                                     \
    int st;                                                                    \
    if (nx < 1 || ny < 1 || nz < 1) return OR_INVALID_VALUE;                   \
    ws = (T *)malloc(sizeof(T) * 4 * (size_t)n);                               \
    if (!ws) return OR_ALLOC_FAILED;                                           \
    st = oracle_adi_ws_##SUF(mode, lambda, u, nx, ny, nz, iters, ws);          \
    free(ws);                                                                  \
    return st;                                                                 \
  }
DEFINE_ADI(double, d)
DEFINE_ADI(float, s)
