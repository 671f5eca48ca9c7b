/*
 * vadi_oracle.c -- CPU FP64 oracle for the differentiable, variable-
 * coefficient ADI heat-diffusion model (forward and reverse mode).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU adjoint path.  Only the
 * tests and bench.py's cpu_baseline leg load it; the GPU library never does.
 *
 * The reference has no differentiable ADI.  The model is the reference's
 * ADI step (apps/adi/src/adi_orig.c:147-243: 7-point right-hand side, x and
 * y sweeps into du, z sweep then u += du, Dirichlet rows on the boundary)
 * with per-point coefficients from a diffusivity field kappa:
 *   face value   kf_{i+1/2} = (kappa_i + kappa_{i+1}) / 2  (per direction)
 *   rhs          r_i = lam * sum_dir [kf_{i+1/2}(u_{i+1}-u_i)
 *                                     - kf_{i-1/2}(u_i-u_{i-1})]
 *   sweep rows   a_i = -lam kf_{i-1/2}, c_i = -lam kf_{i+1/2},
 *                b_i = 1 - a_i - c_i
 * at interior points; boundary points have r = 0 and identity rows.
 * Loss = 0.5 * ||u_T - target||^2.
 * The reverse pass stores every step's u (no checkpointing: this is the
 * plain full-storage reference the checkpointed GPU adjoint is compared
 * against).  Grid is n^3, index i + n*j + n*n*k.
 */
#include <stdlib.h>
#include <string.h>

static inline int interior(int i, int j, int k, int n) {
  return i > 0 && i < n - 1 && j > 0 && j < n - 1 && k > 0 && k < n - 1;
}

static void vadi_rhs(const double *kap, const double *u, double *r, int n,
                     double lam) {
  long nn = (long)n * n;
  int k;
#pragma omp parallel for schedule(static)
  for (k = 0; k < n; k++) {
    int i, j;
    for (j = 0; j < n; j++)
      for (i = 0; i < n; i++) {
        long p = i + (long)n * j + nn * k;
        if (!interior(i, j, k, n)) {
          r[p] = 0.0;
          continue;
        }
        {
          long st[3] = {1, n, nn};
          double s = 0.0;
          int dir;
          for (dir = 0; dir < 3; dir++) {
            long q = st[dir];
            double kp = 0.5 * (kap[p] + kap[p + q]);
            double km = 0.5 * (kap[p] + kap[p - q]);
            s += kp * (u[p + q] - u[p]) - km * (u[p] - u[p - q]);
          }
          r[p] = lam * s;
        }
      }
  }
}

/* line (along dir) coefficients at position t of a line whose element t
 * sits at p */
static inline void vadi_coef(const double *kap, long p, long q, int ipt,
                             double lam, double *a, double *b, double *c) {
  if (!ipt) {
    *a = 0.0;
    *b = 1.0;
    *c = 0.0;
    return;
  }
  *a = -lam * 0.5 * (kap[p] + kap[p - q]);
  *c = -lam * 0.5 * (kap[p] + kap[p + q]);
  *b = 1.0 - *a - *c;
}

static void thomas_line(const double *a, const double *b, const double *c,
                        double *d, int N, double *c2, double *d2) {
  double bb, cc, dd;
  int i;
  bb = 1.0 / b[0];
  cc = bb * c[0];
  dd = bb * d[0];
  c2[0] = cc;
  d2[0] = dd;
  for (i = 1; i < N; i++) {
    bb = b[i] - a[i] * cc;
    dd = d[i] - a[i] * dd;
    bb = 1.0 / bb;
    cc = bb * c[i];
    dd = bb * dd;
    c2[i] = cc;
    d2[i] = dd;
  }
  d[N - 1] = dd;
  for (i = N - 2; i >= 0; i--) {
    dd = d2[i] - c2[i] * dd;
    d[i] = dd;
  }
}

/* x <- A_dir^{-1} x (transpose=0) or A_dir^{-T} x (transpose=1) */
static void vadi_solve(const double *kap, double *x, int n, int dir,
                       double lam, int transpose) {
  long nn = (long)n * n;
  long st[3] = {1, n, nn};
  long q = st[dir];
  int line;
#pragma omp parallel
  {
    double *a = (double *)malloc(sizeof(double) * n * 7);
    double *b = a + n, *c = a + 2 * n, *d = a + 3 * n, *c2 = a + 4 * n,
           *d2 = a + 5 * n, *tmp = a + 6 * n;
#pragma omp for schedule(static)
    for (line = 0; line < n * n; line++) {
      int o1 = line % n, o2 = line / n, t;
      int idx[3];
      long base;
      /* the two non-solve coordinates, fastest first */
      if (dir == 0) {
        idx[1] = o1;
        idx[2] = o2;
      } else if (dir == 1) {
        idx[0] = o1;
        idx[2] = o2;
      } else {
        idx[0] = o1;
        idx[1] = o2;
      }
      idx[dir] = 0;
      base = idx[0] + (long)n * idx[1] + nn * idx[2];
      for (t = 0; t < n; t++) {
        int ii[3] = {idx[0], idx[1], idx[2]};
        ii[dir] = t;
        vadi_coef(kap, base + t * q, q, interior(ii[0], ii[1], ii[2], n), lam,
                  &a[t], &b[t], &c[t]);
        d[t] = x[base + t * q];
      }
      if (transpose) {
        /* transposed tridiagonal: sub_t = c_{t-1}, super_t = a_{t+1} */
        for (t = 0; t < n; t++) tmp[t] = t > 0 ? c[t - 1] : 0.0;
        for (t = 0; t < n; t++) c[t] = t < n - 1 ? a[t + 1] : 0.0;
        memcpy(a, tmp, sizeof(double) * n);
      }
      thomas_line(a, b, c, d, n, c2, d2);
      for (t = 0; t < n; t++) x[base + t * q] = d[t];
    }
    free(a);
  }
}

static void vadi_step(const double *kap, double *u, double *r, int n,
                      double lam) {
  long N3 = (long)n * n * n, p;
  vadi_rhs(kap, u, r, n, lam);
  vadi_solve(kap, r, n, 0, lam, 0);
  vadi_solve(kap, r, n, 1, lam, 0);
  vadi_solve(kap, r, n, 2, lam, 0);
  for (p = 0; p < N3; p++) u[p] += r[p];
}

int oracle_vadi_forward(const double *kap, const double *u0, double *uT, int n,
                        int steps, double lam) {
  long N3 = (long)n * n * n;
  double *r = (double *)malloc(sizeof(double) * N3);
  int s;
  if (!r) return 2;
  memcpy(uT, u0, sizeof(double) * N3);
  for (s = 0; s < steps; s++) vadi_step(kap, uT, r, n, lam);
  free(r);
  return 0;
}

/* kappa-gradient of one sweep along dir: x is the sweep's solution, l the
 * adjoint solution lambda = A^{-T} xbar; accumulates into gk. */
static void vadi_coef_grad(const double *x, const double *l, double *gk, int n,
                           int dir, double lam) {
  long nn = (long)n * n;
  long st[3] = {1, n, nn};
  long q = st[dir];
  int k;
  /* serial over k-planes to keep the scatter race-free */
  for (k = 0; k < n; k++) {
    int i, j;
    for (j = 0; j < n; j++)
      for (i = 0; i < n; i++) {
        long p = i + (long)n * j + nn * k;
        double gm, gp;
        if (!interior(i, j, k, n)) continue;
        /* abar = -l x_{-}, bbar = -l x, cbar = -l x_{+} */
        gm = lam * (-l[p] * (x[p] - x[p - q]));
        gp = lam * (-l[p] * (x[p] - x[p + q]));
        gk[p] += 0.5 * gm + 0.5 * gp;
        gk[p - q] += 0.5 * gm;
        gk[p + q] += 0.5 * gp;
      }
  }
}

int oracle_vadi_grad(const double *kap, const double *u0, const double *target,
                     int n, int steps, double lam, double *loss, double *gk,
                     double *gu0) {
  long N3 = (long)n * n * n, p, nn = (long)n * n;
  double *us = (double *)malloc(sizeof(double) * N3 * (steps + 1));
  double *r = (double *)malloc(sizeof(double) * N3);
  double *px = (double *)malloc(sizeof(double) * N3);
  double *py = (double *)malloc(sizeof(double) * N3);
  double *pz = (double *)malloc(sizeof(double) * N3);
  double *ub = (double *)malloc(sizeof(double) * N3);
  double *lb = (double *)malloc(sizeof(double) * N3);
  int s;
  double L = 0.0;
  if (!us || !r || !px || !py || !pz || !ub || !lb) return 2;
  memcpy(us, u0, sizeof(double) * N3);
  for (s = 0; s < steps; s++) {
    memcpy(us + (s + 1) * N3, us + s * N3, sizeof(double) * N3);
    vadi_step(kap, us + (s + 1) * N3, r, n, lam);
  }
  for (p = 0; p < N3; p++) {
    double e = us[steps * N3 + p] - target[p];
    L += 0.5 * e * e;
    ub[p] = e;
  }
  *loss = L;
  memset(gk, 0, sizeof(double) * N3);
  for (s = steps - 1; s >= 0; s--) {
    const double *un = us + s * N3;
    /* recompute the step's intermediates */
    vadi_rhs(kap, un, px, n, lam);
    vadi_solve(kap, px, n, 0, lam, 0); /* px = Ax^{-1} r */
    memcpy(py, px, sizeof(double) * N3);
    vadi_solve(kap, py, n, 1, lam, 0); /* py = Ay^{-1} px */
    memcpy(pz, py, sizeof(double) * N3);
    vadi_solve(kap, pz, n, 2, lam, 0); /* pz = Az^{-1} py */
    /* z sweep */
    memcpy(lb, ub, sizeof(double) * N3);
    vadi_solve(kap, lb, n, 2, lam, 1);
    vadi_coef_grad(pz, lb, gk, n, 2, lam);
    /* y sweep */
    vadi_solve(kap, lb, n, 1, lam, 1);
    vadi_coef_grad(py, lb, gk, n, 1, lam);
    /* x sweep */
    vadi_solve(kap, lb, n, 0, lam, 1);
    vadi_coef_grad(px, lb, gk, n, 0, lam);
    /* rhs: lb is rbar; ubar += D^T rbar, kappa-bar += dD/dkappa^T rbar */
    {
      long st[3] = {1, n, nn};
      int dir, k;
      for (dir = 0; dir < 3; dir++) {
        long q = st[dir];
        for (k = 0; k < n; k++) {
          int i, j;
          for (j = 0; j < n; j++)
            for (i = 0; i < n; i++) {
              int ii[3] = {i, j, k};
              long pa = i + (long)n * j + nn * k, pb;
              double G, kf;
              if (ii[dir] == n - 1) continue;
              pb = pa + q;
              {
                int jj[3] = {i, j, k};
                jj[dir] += 1;
                G = lam * ((interior(i, j, k, n) ? lb[pa] : 0.0) -
                           (interior(jj[0], jj[1], jj[2], n) ? lb[pb] : 0.0));
              }
              if (G == 0.0) continue;
              kf = 0.5 * (kap[pa] + kap[pb]);
              ub[pb] += G * kf;
              ub[pa] -= G * kf;
              gk[pa] += 0.5 * G * (un[pb] - un[pa]);
              gk[pb] += 0.5 * G * (un[pb] - un[pa]);
            }
        }
      }
    }
  }
  memcpy(gu0, ub, sizeof(double) * N3);
  free(us);
  free(r);
  free(px);
  free(py);
  free(pz);
  free(ub);
  free(lb);
  return 0;
}
