// trid_vadi.hip -- the variable-diffusivity ADI ("vadi") of
// include/trid_adjoint.h: forward steps and the reverse-mode gradient of
// 0.5 ||u_T - target||^2 with respect to kappa and u_0.
//
// One step maps u_k to u_{k+1}:
//   r = D(kappa) u_k                  (7-point stencil, zero on the boundary)
//   p = A_x^{-1} r,  q = A_y^{-1} p,  w = A_z^{-1} q,   u_{k+1} = u_k + w
// with A_axis = tridiag(-lam kf_{i-1/2}, 1 + lam (kf_{i-1/2} + kf_{i+1/2}),
// -lam kf_{i+1/2}) at grid-interior points and identity rows elsewhere;
// kf_{i+1/2} = (kappa_i + kappa_{i+1}) / 2.  The reference ADI step
// (apps/adi/src/adi_orig.c:147-243) has the same structure with constant
// coefficients.
//
// Kernels:
//   k_vadi_x<R, MODE>: whole x-lines per wave, u and kappa staged through
//     the wave's LDS buffer in 16-byte pieces; MODE 0 builds r = D u in
//     registers and solves A_x (the right-hand side never goes to memory),
//     MODE 1 solves the transposed system A_x^T.
//   y and z sweeps: the partitioned Thomas kernel with the CoefKappa loader
//     (coefficients formed in registers from kappa), forward or transposed;
//     the z sweep adds into u, optionally out of place (u_{k+1} into a new
//     buffer, u_k kept).
//   k_vadi_grad_col: one gather per point for everything the adjoint of one
//     step accumulates besides the three transposed solves -- the kappa
//     gradient of the three sweeps and of the stencil, and ubar += D^T rbar.
//
// Reverse mode with checkpoints: the forward pass keeps u at the start of
// every segment of `ckpt` steps (written out of place, no copies) and
// records the last segment's intermediates as it goes; the backward pass
// replays every other segment from its checkpoint, recording u_k, p, q of
// each step, then runs the adjoint steps of the segment in reverse.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/trid_adjoint.h"
#include "../../include/trid_hip_ext.h"
#include "trid_launch.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

struct Grid {
  int nx, ny, nz;
  long long nxy, n;
};

__device__ __forceinline__ bool grid_interior(const Grid &g, int i, int j, int k) {
  return i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1 && k > 0 && k < g.nz - 1;
}

__device__ __forceinline__ void grid_coords(const Grid &g, long long p, int &i, int &j,
                                            int &k) {
  if (g.n < (1LL << 31)) {
    const unsigned up = (unsigned)p, line = up / (unsigned)g.nx;
    i = (int)(up - line * (unsigned)g.nx);
    k = (int)(line / (unsigned)g.ny);
    j = (int)(line - (unsigned)k * (unsigned)g.ny);
  } else {
    const long long line = p / g.nx;
    i = (int)(p - line * g.nx);
    k = (int)(line / g.ny);
    j = (int)(line - (long long)k * g.ny);
  }
}

// r = lam * sum_axes [kf+ (u+ - u) - kf- (u - u-)] at interior points, 0 else
__global__ void __launch_bounds__(256)
    k_vadi_rhs(const double *__restrict__ K, const double *__restrict__ u,
               double *__restrict__ r, Grid g, double lam) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n) return;
  int i, j, k;
  grid_coords(g, p, i, j, k);
  if (!grid_interior(g, i, j, k)) {
    r[p] = 0.0;
    return;
  }
  const long long st[3] = {1, g.nx, g.nxy};
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const long long q = st[a];
    const double kp = 0.5 * (K[p] + K[p + q]);
    const double km = 0.5 * (K[p] + K[p - q]);
    s += kp * (u[p + q] - u[p]) - km * (u[p] - u[p - q]);
  }
  r[p] = lam * s;
}

// x sweep over whole lines, one wave holding 64/cps lines of one z-plane
// (the host requires nx == cps * R, cps <= 64 a power of two, ny a multiple
// of the lines per wave and 16-byte aligned arrays).
//   MODE 0: out = A_x^{-1} D(kappa) in   (in = u_k, out = p)
//   MODE 1: out = A_x^{-T} in
// CP: cache policy of the output stores (kStream / kCached).
template <int R, int MODE, int CP>
__global__ void __launch_bounds__(256)
    k_vadi_x(const double *__restrict__ K, const double *__restrict__ in,
             double *__restrict__ out, int nx, int ny, int nz, double lam, int cps_log2) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int cps = 1 << cps_log2;
  const int kc_idx = tid & (cps - 1);
  const int lines_per_block = 256 >> cps_log2;
  const long long nlines = (long long)ny * nz;
  const long long line = (long long)blockIdx.x * lines_per_block + (tid >> cps_log2);
  const long long wline0 =
      (long long)blockIdx.x * lines_per_block + ((tid & ~(kWave - 1)) >> cps_log2);
  const bool wave_active = wline0 < nlines;
  const long long nxy = (long long)nx * ny;
  const int kw = (int)(wline0 / ny);
  const bool plane_in = kw >= 1 && kw <= nz - 2;
  const int jl = (int)(line % ny);
  const bool line_in = plane_in && jl >= 1 && jl <= ny - 2;
  double *stage = reinterpret_cast<double *>(smem_raw) + (tid >> 6) * (kWave * R);
  const long long wo = wline0 * nx;
  const int row0 = kc_idx * R;
  auto inner = [&](int row) { return line_in && row >= 1 && row + 2 <= nx; };

  double kc[R], kl = 0.0, kr = 0.0;
  double d[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    kc[r] = 1.0;
    d[r] = 0.0;
  }
  if (wave_active) {
    if (MODE == 1 || !plane_in) {
      Span<double, R> sk, sd;
      span_issue<double, R>(K + wo, lane, sk);
      if (MODE == 1) span_issue<double, R>(in + wo, lane, sd);
      span_stage<double, R>(sk, stage, lane, kc, &kl, &kr);
      if (MODE == 1) span_stage<double, R>(sd, stage, lane, d);
    } else {
      // d = lam * (x part + y part + z part) of the stencil, evaluated in the
      // loaded layout (lane holds elements (v*64 + lane)*2 .. +1 of the
      // wave's span): neighbour spans line up element for element, the
      // x-neighbours at piece ends come from the adjacent lanes; only kappa
      // (for the coefficients) and the right-hand side are transposed into
      // chunks.
      constexpr int NP = R / 2;
      const long long nb[4] = {-(long long)nx, (long long)nx, -nxy, nxy};
      Span<double, R> sk, su, pu, pk, sd;
      span_issue<double, R>(K + wo, lane, sk);
      span_issue<double, R>(in + wo, lane, su);
      span_issue<double, R>(in + wo + nb[0], lane, pu);
      span_issue<double, R>(K + wo + nb[0], lane, pk);
      span_stage<double, R>(sk, stage, lane, kc, &kl, &kr);
      const int nx_log2 = __builtin_ctz(nx);  // nx = cps * R, a power of two
      const int jw = (int)(wline0 % ny);      // the wave's first line in its plane
#pragma unroll
      for (int v = 0; v < NP; v++) {
        const auto uv = su.v[v], kv = sk.v[v];
        double ul = lane_read(uv[1], lane > 0 ? lane - 1 : 0);
        double ur = lane_read(uv[0], lane < kWave - 1 ? lane + 1 : kWave - 1);
        double kl_ = lane_read(kv[1], lane > 0 ? lane - 1 : 0);
        double kr_ = lane_read(kv[0], lane < kWave - 1 ? lane + 1 : kWave - 1);
        const double up = v > 0 ? lane_bcast(su.v[v > 0 ? v - 1 : 0][1], kWave - 1) : 0.0;
        const double un = v + 1 < NP ? lane_bcast(su.v[v + 1 < NP ? v + 1 : 0][0], 0) : 0.0;
        const double kp_ = v > 0 ? lane_bcast(sk.v[v > 0 ? v - 1 : 0][1], kWave - 1) : 0.0;
        const double kn_ = v + 1 < NP ? lane_bcast(sk.v[v + 1 < NP ? v + 1 : 0][0], 0) : 0.0;
        if (lane == 0) {
          ul = up;
          kl_ = kp_;
        }
        if (lane == kWave - 1) {
          ur = un;
          kr_ = kn_;
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
          const double upv = j == 0 ? ul : uv[0], unx = j == 1 ? ur : uv[1];
          const double kpv = j == 0 ? kl_ : kv[0], knx = j == 1 ? kr_ : kv[1];
          const double kp = 0.5 * (kv[j] + knx), km = 0.5 * (kv[j] + kpv);
          sd.v[v][j] = kp * (unx - uv[j]) - km * (uv[j] - upv);
        }
      }
#pragma unroll
      for (int t = 0; t < 4; t++) {
        if (t > 0) {
          span_issue<double, R>(in + wo + nb[t], lane, pu);
          span_issue<double, R>(K + wo + nb[t], lane, pk);
        }
#pragma unroll
        for (int v = 0; v < NP; v++)
#pragma unroll
          for (int j = 0; j < 2; j++) {
            const double uc_ = su.v[v][j], kc_ = sk.v[v][j];
            const double un_ = pu.v[v][j], kn_ = pk.v[v][j];
            const double kf = 0.5 * (kc_ + kn_);
            // minus side: -kf (u - u-), plus side: +kf (u+ - u)
            sd.v[v][j] += (t & 1) ? kf * (un_ - uc_) : -(kf * (uc_ - un_));
          }
      }
#pragma unroll
      for (int v = 0; v < NP; v++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          const int e = (v * kWave + lane) * 2 + j;
          const int row = e & (nx - 1);
          const int jl_e = jw + (e >> nx_log2);
          const bool in_e = plane_in && jl_e >= 1 && jl_e <= ny - 2 && row >= 1 && row + 2 <= nx;
          sd.v[v][j] = in_e ? lam * sd.v[v][j] : 0.0;
        }
      span_stage<double, R>(sd, stage, lane, d);
    }
  }
  double a[R], b[R], c[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int row = row0 + r;
    const double kpv = r == 0 ? kl : kc[r - 1], knx = r == R - 1 ? kr : kc[r + 1];
    const double am = -lam * (0.5 * (kc[r] + kpv));
    const double cp = -lam * (0.5 * (kc[r] + knx));
    const bool in_r = inner(row);
    b[r] = in_r ? 1.0 - am - cp : 1.0;
    if (MODE == 0) {
      a[r] = in_r ? am : 0.0;
      c[r] = in_r ? cp : 0.0;
    } else {
      a[r] = inner(row - 1) ? am : 0.0;
      c[r] = inner(row + 1) ? cp : 0.0;
    }
  }
  Row<double> f, l;
  chunk_reduce<double, R>(a, b, c, d, f, l);
  double xf, xl;
  reduced_wave<double>(f, l, kc_idx, cps, lane, xf, xl);
  double y[R];
  chunk_expand<double, R>(a, c, d, xf, xl, y);
  if (wave_active) store_chunk_staged<double, R, CP>(out + wo, stage, lane, y, false);
}

// Adjoint of one step at every point p (gather form, no atomics).  For each
// face f = (lo, hi = lo + stride) along an axis, with that axis' sweep
// solution X (p, q or w = u_{k+1} - u_k) and transposed-sweep solution Lam
// (Lx, Ly, Lz), tilde = value times [point is grid-interior]:
//   sweep term    G1_f = -lam (X_hi - X_lo) (Lam~_hi - Lam~_lo)
//   stencil term  G_f  =  lam (Lx~_lo - Lx~_hi),  G2_f = G_f (u_hi - u_lo)
// and gk_lo, gk_hi += (G1_f + G2_f) / 2,  ubar_hi += G_f kf,  ubar_lo -= G_f kf.
// Launched on a (64, 4) thread tile per (x-tile, 4 rows, ZC planes): rows
// y-1 and y+1 of the tile's interior rows come from the same work-group's
// lines.  Each thread marches along z over ZC planes of one (x, y) column: U, K,
// LX, U1 and LZ of planes z-1, z, z+1 roll through registers, so each point
// loads the z-neighbour values once (28 loads per point instead of 38 for a
// thread per point) and the neighbour planes never come back from L2 or HBM.
// Every load of a point is issued before its first use; neighbours off the
// grid read the point itself and are masked.
// Tiles are taken in XCD-contiguous order: work-group b of the 1-D grid (the
// hardware deals work-groups to the 8 XCDs round-robin) takes tile
// (b mod 8) * (B / 8) + b / 8 of the x-fastest, then y, then z tile order (B
// the grid rounded down to a multiple of 8), so that the tiles of one XCD
// are y-neighbours and find each other's halo rows in that XCD's L2.
template <int ZC>
__global__ void __launch_bounds__(256)
    k_vadi_grad_col(const double *__restrict__ K, const double *__restrict__ U,
                    const double *__restrict__ U1, const double *__restrict__ P,
                    const double *__restrict__ Q, const double *__restrict__ LZ,
                    const double *__restrict__ LY, const double *__restrict__ LX,
                    double *__restrict__ UB, double *__restrict__ GK, Grid g, double lam,
                    int ntx, int nty) {
  const unsigned nb8 = gridDim.x & ~7u, b = blockIdx.x;
  const unsigned t = b < nb8 ? (b & 7u) * (nb8 >> 3) + (b >> 3) : b;
  const int tx = (int)(t % (unsigned)ntx), tyz = (int)(t / (unsigned)ntx);
  const int ty = tyz % nty, tz = tyz / nty;
  const int x = tx * 64 + threadIdx.x;
  const int y = ty * 4 + threadIdx.y;
  if (x >= g.nx || y >= g.ny) return;
  const int z0 = tz * ZC;
  const int z1 = z0 + ZC < g.nz ? z0 + ZC : g.nz;
  const long long col = (long long)y * g.nx + x;
  // plane values: U, K, LX, U1, LZ
  struct Pl {
    double u, k, lx, u1, lz;
  };
  auto load_plane = [&](int z) {
    Pl v{0.0, 0.0, 0.0, 0.0, 0.0};
    if (z >= 0 && z < g.nz) {
      const long long q = (long long)z * g.nxy + col;
      v.u = U[q];
      v.k = K[q];
      v.lx = LX[q];
      v.u1 = U1[q];
      v.lz = LZ[q];
    }
    return v;
  };
  Pl m = load_plane(z0 - 1), c = load_plane(z0);
  const bool xm_in = x >= 1, xp_in = x + 1 < g.nx, ym_in = y >= 1, yp_in = y + 1 < g.ny;
#pragma unroll 1
  for (int z = z0; z < z1; z++) {
    const Pl pl = load_plane(z + 1);
    const long long p = (long long)z * g.nxy + col;
    const long long qxm = xm_in ? p - 1 : p, qxp = xp_in ? p + 1 : p;
    const long long qym = ym_in ? p - g.nx : p, qyp = yp_in ? p + g.nx : p;
    // in-plane neighbours (off-grid ones read the point itself, masked below)
    const double uxm = U[qxm], uxp = U[qxp], uym = U[qym], uyp = U[qyp];
    const double kxm = K[qxm], kxp = K[qxp], kym = K[qym], kyp = K[qyp];
    const double rxm = LX[qxm], rxp = LX[qxp], rym = LX[qym], ryp = LX[qyp];
    const double pc = P[p], pxm = P[qxm], pxp = P[qxp];
    const double qc = Q[p], qym_ = Q[qym], qyp_ = Q[qyp];
    const double lyc = LY[p], lym = LY[qym], lyp = LY[qyp];
    const double ub0 = UB[p], gk0 = GK[p];
    const bool inp = grid_interior(g, x, y, z);
    const double up = c.u, kp = c.k;
    const double rp = inp ? c.lx : 0.0;
    // per axis a and side sd (0: minus, 1: plus): neighbour u, kappa, raw LX,
    // the axis' sweep solution X and adjoint L, presence and interiority
    double un[3][2] = {{uxm, uxp}, {uym, uyp}, {m.u, pl.u}};
    double kn[3][2] = {{kxm, kxp}, {kym, kyp}, {m.k, pl.k}};
    double rn[3][2] = {{rxm, rxp}, {rym, ryp}, {m.lx, pl.lx}};
    double xn[3][2] = {{pxm, pxp}, {qym_, qyp_}, {m.u1, pl.u1}};
    double ln[3][2] = {{rxm, rxp}, {lym, lyp}, {m.lz, pl.lz}};
    const bool has[3][2] = {{xm_in, xp_in}, {ym_in, yp_in}, {z >= 1, z + 1 < g.nz}};
    bool inn[3][2];
    inn[0][0] = xm_in && grid_interior(g, x - 1, y, z);
    inn[0][1] = xp_in && grid_interior(g, x + 1, y, z);
    inn[1][0] = ym_in && grid_interior(g, x, y - 1, z);
    inn[1][1] = yp_in && grid_interior(g, x, y + 1, z);
    inn[2][0] = z >= 1 && grid_interior(g, x, y, z - 1);
    inn[2][1] = z + 1 < g.nz && grid_interior(g, x, y, z + 1);
    const double xc[3] = {pc, qc, c.u1};
    const double lc[3] = {c.lx, lyc, c.lz};
    double du = 0.0, dk = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double xp = a == 2 ? xc[2] - up : xc[a];
      const double lp = inp ? lc[a] : 0.0;
#pragma unroll
      for (int sd = 0; sd < 2; sd++) {
        if (!has[a][sd]) continue;
        const double unn = un[a][sd];
        const double rnn = inn[a][sd] ? rn[a][sd] : 0.0;
        const double xnn = a == 2 ? xn[a][sd] - unn : xn[a][sd];
        const double lnn = inn[a][sd] ? ln[a][sd] : 0.0;
        const double dX = sd == 0 ? xp - xnn : xnn - xp;  // X_hi - X_lo
        const double dL = sd == 0 ? lp - lnn : lnn - lp;  // Lam~_hi - Lam~_lo
        const double dU = sd == 0 ? up - unn : unn - up;  // u_hi - u_lo
        const double G = lam * (sd == 0 ? rnn - rp : rp - rnn);
        const double kf = 0.5 * (kp + kn[a][sd]);
        dk += 0.5 * (-lam * dX * dL + G * dU);
        du += sd == 0 ? G * kf : -(G * kf);
      }
    }
    UB[p] = ub0 + du;
    GK[p] = gk0 + dk;
    m = c;
    c = pl;
  }
}

// ubar = u - target; partial sums of ubar^2 per block
__global__ void __launch_bounds__(256)
    k_residual(const double *__restrict__ u, const double *__restrict__ t,
               double *__restrict__ ub, double *__restrict__ partial, long long n) {
  __shared__ double sh[256];
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (p < n) {
    const double e = u[p] - t[p];
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

__global__ void __launch_bounds__(256)
    k_sum(const double *__restrict__ in, double *__restrict__ out, long long n) {
  __shared__ double sh[256];
  double v = 0.0;
  for (long long p = threadIdx.x; p < n; p += blockDim.x) v += in[p];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void __launch_bounds__(256) k_zero(double *__restrict__ x, long long n) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) x[p] = 0.0;
}

}  // namespace tridhip

using namespace tridhip;

namespace {

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

// planes per thread of the gradient gather (k_vadi_grad_col; at 128^3, 16
// planes: 54.5 -> 48 us per step, against 8 planes 50 us)
constexpr int kGradZ = 16;

Grid make_grid(int nx, int ny, int nz) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nxy = (long long)nx * ny;
  g.n = g.nxy * nz;
  return g;
}

SysDesc line_desc(const Grid &g, int axis) {
  SysDesc D;
  memset(&D, 0, sizeof D);
  const long long str[3] = {1, g.nx, g.nxy};
  const int dims[3] = {g.nx, g.ny, g.nz};
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
  D.small_index = D.nsys < (1LL << 31) && !wide_index();
  return D;
}

// Cache policy of the streaming accesses, from the working set of a forward
// and adjoint step: about ten grid arrays (u, kappa, the sweep outputs,
// their adjoints, the gradients).
int vadi_policy(const Grid &g) {
  if ((unsigned long long)g.n * 8 >= (1ULL << 32) || wide_index()) return kStream;
  const int p = cache_policy(10 * (size_t)g.n * sizeof(double));
  return p == kMixed ? kCached : p;  // no coefficient arrays: d and u cached
}

// Generic kappa sweep along `axis` (partitioned Thomas, scalar loads):
// non-inc: out (default d) = A^{-1} d or A^{-T} d; inc: out (default u) =
// u + A^{-1} d.
tridStatus_t vadi_sweep(const double *K, double *d, double *u, double *out, const Grid &g,
                        int axis, double lam, int transpose, int inc, hipStream_t st) {
  const SysDesc D = line_desc(g, axis);
  const int order = axis == 0 ? 0 : 1;
  PartCfg cfg = heuristic(order, D.N, 0, sizeof(double));
  if (cfg.kind != 1) return fail(TRID_STATUS_INVALID_VALUE, "vadi: grid too long along an axis");
  cfg.vec = 0;
  cfg.idx32 = (unsigned long long)g.n * 8 < (1ULL << 32) && !wide_index();
  cfg.cp = cfg.idx32 ? vadi_policy(g) : kStream;
  const CoefKappa<double> cf{K, lam, transpose};
  hipError_t e = inc ? launch_part<double, 1>(cfg, cf, d, u, D, st, out)
                     : launch_part<double, 0>(cfg, cf, d, u, D, st, out);
  return check_hip(e, "vadi sweep launch");
}

tridStatus_t vadi_rhs(const double *K, const double *u, double *r, const Grid &g, double lam,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_vadi_rhs, dim3(blocks_for(g.n)), dim3(256), 0, st, K, u, r, g, lam);
  return check_hip(hipGetLastError(), "vadi rhs launch");
}

// Chunk length of the staged x kernel for this grid (0: not applicable).
int staged_x_rows(const Grid &g, const void *K, const void *in, const void *out) {
  if (!aligned16(K) || !aligned16(in) || !aligned16(out)) return 0;
  for (int R : {4, 8, 16}) {
    if (g.nx % R) continue;
    const int cps = g.nx / R;
    if (cps > kWave || (cps & (cps - 1))) continue;
    if (g.ny % (kWave / cps)) continue;
    return R;
  }
  return 0;
}

template <int MODE>
tridStatus_t launch_vadi_x(int R, const double *K, const double *in, double *out,
                           const Grid &g, double lam, hipStream_t st) {
  const int cps = g.nx / R;
  const int cps_log2 = ilog2(cps);
  const int lines_per_block = 256 / cps;
  const long long nlines = (long long)g.ny * g.nz;
  const unsigned blocks = (unsigned)((nlines + lines_per_block - 1) / lines_per_block);
  const size_t lds = 256 * (size_t)R * sizeof(double);
#define TRID_VX(RR, CC)                                                                  \
  hipLaunchKernelGGL((k_vadi_x<RR, MODE, CC>), dim3(blocks), dim3(256), lds, st, K, in, out, \
                     g.nx, g.ny, g.nz, lam, cps_log2)
#define TRID_VX_R(CC)              \
  if (R == 8) TRID_VX(8, CC);      \
  else if (R == 4) TRID_VX(4, CC); \
  else TRID_VX(16, CC);
  if (vadi_policy(g) == kCached) {
    TRID_VX_R(kCached)
  } else {
    TRID_VX_R(kStream)
  }
#undef TRID_VX_R
#undef TRID_VX
  return check_hip(hipGetLastError(), "vadi x launch");
}

// p = A_x^{-1} D(kappa) u   (r: scratch for the unfused path; may equal p)
tridStatus_t vadi_rhs_x(const double *K, const double *u, double *p, const Grid &g, double lam,
                        hipStream_t st) {
  const int R = staged_x_rows(g, K, u, p);
  if (R) return launch_vadi_x<0>(R, K, u, p, g, lam, st);
  tridStatus_t s = vadi_rhs(K, u, p, g, lam, st);
  if (!s) s = vadi_sweep(K, p, nullptr, nullptr, g, 0, lam, 0, 0, st);
  return s;
}

// out = A_x^{-T} in
tridStatus_t vadi_xT(const double *K, double *in, double *out, const Grid &g, double lam,
                     hipStream_t st) {
  const int R = staged_x_rows(g, K, in, out);
  if (R) return launch_vadi_x<1>(R, K, in, out, g, lam, st);
  return vadi_sweep(K, in, nullptr, out, g, 0, lam, 1, 0, st);
}

// One forward step from src to dst (dst may equal src).  p, q receive the
// x and y sweep solutions (q may equal p: the y sweep then works in place).
tridStatus_t vadi_step(const double *K, const double *src, double *dst, double *p, double *q,
                       const Grid &g, double lam, hipStream_t st) {
  tridStatus_t s = vadi_rhs_x(K, src, p, g, lam, st);
  if (!s) s = vadi_sweep(K, p, nullptr, q == p ? nullptr : q, g, 1, lam, 0, 0, st);
  if (!s)
    s = vadi_sweep(K, q, const_cast<double *>(src), dst == src ? nullptr : dst, g, 2, lam, 0, 1,
                   st);
  return s;
}

}  // namespace

EXTERN_C tridStatus_t tridVadiRhs(const double *kappa, const double *u, double *r, int nx,
                                  int ny, int nz, double lam) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || !kappa || !u || !r)
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  return vadi_rhs(kappa, u, r, make_grid(nx, ny, nz), lam, current_stream());
}

EXTERN_C tridStatus_t tridVadiSweep(const double *kappa, double *d, double *u, int nx, int ny,
                                    int nz, int axis, double lam, int transpose, int inc) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || !kappa || !d || (inc && !u) || axis < 0 || axis > 2 ||
      (inc && transpose))
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  return vadi_sweep(kappa, d, u, nullptr, make_grid(nx, ny, nz), axis, lam, transpose, inc,
                    current_stream());
}

EXTERN_C tridStatus_t tridVadiForward(const double *kappa, double *u, int nx, int ny, int nz,
                                      int steps, double lam, int sync) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || steps < 0 || !kappa || !u)
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  const Grid g = make_grid(nx, ny, nz);
  hipStream_t st = current_stream();
  double *r = static_cast<double *>(workspace((size_t)g.n * sizeof(double), st));
  if (!r) return TRID_STATUS_ALLOC_FAILED;  // reason recorded by workspace()
  for (int s = 0; s < steps; s++) {
    tridStatus_t e = vadi_step(kappa, u, u, r, r, g, lam, st);
    if (e) return e;
  }
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridVadiLossGrad(const double *kappa, const double *u0,
                                       const double *target, int nx, int ny, int nz, int steps,
                                       double lam, int ckpt, double *loss, double *grad_kappa,
                                       double *grad_u0) {
  clear_error();
  if (nx < 1 || ny < 1 || nz < 1 || steps < 1 || ckpt < 1 || !kappa || !u0 || !target ||
      !loss || !grad_kappa || !grad_u0)
    return fail(TRID_STATUS_INVALID_VALUE, "bad vadi arguments");
  ckpt = std::min(ckpt, steps);
  const Grid g = make_grid(nx, ny, nz);
  const long long n = g.n;
  const int nseg = (steps + ckpt - 1) / ckpt;
  // workspace (arrays of n): checkpoints of segments 1..nseg-1, the running
  // u, u_1..u_ckpt of the recorded segment (u_0 is its checkpoint), p and q
  // of each recorded step, Lz, Ly, Lx, the forward scratch r; then partial
  // sums.  Arrays start 256-byte aligned.
  const long long na = (n + 31) / 32 * 32;
  const long long nblk = blocks_for(n);
  const size_t arrays = (size_t)(nseg - 1) + 1 + (size_t)ckpt + 2 * (size_t)ckpt + 3 + 1;
  hipStream_t st = current_stream();
  double *ws =
      static_cast<double *>(workspace((arrays * na + nblk + 1) * sizeof(double), st));
  if (!ws) return TRID_STATUS_ALLOC_FAILED;  // reason recorded by workspace()
  double *cur = ws;
  auto take = [&]() {
    double *p = cur;
    cur += na;
    return p;
  };
  std::vector<double *> ck(nseg);
  ck[0] = const_cast<double *>(u0);  // read only
  for (int s = 1; s < nseg; s++) ck[s] = take();
  double *work = take();
  std::vector<double *> us(ckpt + 1), P(ckpt), Q(ckpt);
  for (int i = 1; i <= ckpt; i++) us[i] = take();
  for (int i = 0; i < ckpt; i++) P[i] = take();
  for (int i = 0; i < ckpt; i++) Q[i] = take();
  double *LZ = take(), *LY = take(), *LX = take(), *r = take();
  double *partial = cur;
  tridStatus_t s = TRID_STATUS_SUCCESS;

  // records steps k0..k1-1 from u_{k0} = us[0]: us[i+1], P[i], Q[i]
  auto record = [&](int seg) {
    const int k0 = seg * ckpt, k1 = std::min(steps, k0 + ckpt);
    us[0] = ck[seg];
    tridStatus_t e = TRID_STATUS_SUCCESS;
    for (int k = k0; k < k1 && !e; k++) {
      const int i = k - k0;
      e = vadi_step(kappa, us[i], us[i + 1], P[i], Q[i], g, lam, st);
    }
    return e;
  };

  // forward: segments 0..nseg-2 in a scratch buffer, each ending in the next
  // checkpoint; the last segment recorded
  for (int sg = 0; sg + 1 < nseg && !s; sg++) {
    const int k0 = sg * ckpt, k1 = k0 + ckpt;
    const double *src = ck[sg];
    for (int k = k0; k < k1 && !s; k++) {
      double *dst = k + 1 == k1 ? ck[sg + 1] : work;
      s = vadi_step(kappa, src, dst, r, r, g, lam, st);
      src = dst;
    }
  }
  if (!s) s = record(nseg - 1);
  if (s) return s;
  const double *uT = us[steps - (nseg - 1) * ckpt];

  // loss and ubar = u_T - target (ubar lives in grad_u0)
  double *ubar = grad_u0;
  hipLaunchKernelGGL(k_residual, dim3((unsigned)nblk), dim3(256), 0, st, uT, target, ubar,
                     partial, n);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, partial, partial + nblk, nblk);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)nblk), dim3(256), 0, st, grad_kappa, n);
  s = check_hip(hipGetLastError(), "loss launch");
  if (s) return s;

  const int grad_tx = (nx + 63) / 64, grad_ty = (ny + 3) / 4;
  const unsigned grad_grid =
      (unsigned)grad_tx * (unsigned)grad_ty * (unsigned)((nz + kGradZ - 1) / kGradZ);
  for (int sg = nseg - 1; sg >= 0 && !s; sg--) {
    const int k0 = sg * ckpt, k1 = std::min(steps, k0 + ckpt);
    if (sg < nseg - 1) s = record(sg);
    for (int k = k1 - 1; k >= k0 && !s; k--) {
      const int i = k - k0;
      s = vadi_sweep(kappa, ubar, nullptr, LZ, g, 2, lam, 1, 0, st);
      if (!s) s = vadi_sweep(kappa, LZ, nullptr, LY, g, 1, lam, 1, 0, st);
      if (!s) s = vadi_xT(kappa, LY, LX, g, lam, st);
      if (s) break;
      hipLaunchKernelGGL(k_vadi_grad_col<kGradZ>, dim3(grad_grid), dim3(64, 4), 0, st, kappa, us[i],
                         us[i + 1], P[i], Q[i], LZ, LY, LX, ubar, grad_kappa, g, lam, grad_tx, grad_ty);
      s = check_hip(hipGetLastError(), "vadi gradient launch");
    }
  }
  if (s) return s;
  double sq = 0.0;
  s = check_hip(hipMemcpyAsync(&sq, partial + nblk, sizeof(double), hipMemcpyDeviceToHost, st),
                "loss copy");
  if (!s) s = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  if (s) return s;
  *loss = 0.5 * sq;
  return TRID_STATUS_SUCCESS;
}
