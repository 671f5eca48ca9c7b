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
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trid_adi.h"
#include "../../include/trid_hip_ext.h"
#include "trid_cc.hpp"
#include "trid_launch.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

constexpr int kRhsPlanes = 8;  // z planes per thread of the RHS kernel

// 7-point right-hand side.  Block (64, 4): one thread per (i, j) column
// segment of kRhsPlanes planes, walking z with the k-1/k/k+1 values in
// registers.  The sum is formed in the reference's order
// (apps/adi/src/adi_orig.c:183-186) without contraction, so it matches the
// reference's rounding.
template <typename T, int CP>
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
    st_stream<CP>(du + idx, v);
    um = uc;
    uc = up;
    idx += nxy;
  }
}

// The 7-point sum in the reference's order (apps/adi/src/adi_orig.c:183-186),
// without contraction.
template <typename T>
__device__ __forceinline__ T rhs_point(T l, T r, T ym, T yp, T zm, T zp, T c,
                                       T lam) {
#pragma clang fp contract(off)
  T s = l + r;
  s = s + ym;
  s = s + yp;
  s = s + zm;
  s = s + zp;
  const T six = T(6) * c;
  s = s - six;
  return lam * s;
}

// Padded wave staging for the fused kernel: piece pc (16 bytes) of the
// wave's span sits at slot pc + pc / P (P = R / E pieces per chunk, one pad
// slot per chunk when P > 1), so lane l's chunk is the P slots from
// l * (P + 1): with an odd slot stride the per-lane chunk accesses of
// consecutive lanes fall in distinct LDS banks (a stride of P slots puts
// lanes l and l + 16 / P in the same bank).
template <typename T, int R>
struct PadStage {
  static constexpr int E = Vec16<T>::n, P = R / E, PAD = P > 1 ? 1 : 0;
  // elements of one wave's buffer
  static constexpr int elems = kWave * (P + PAD) * E;
  __device__ static int slot(int pc) { return pc + (pc / P) * PAD; }
};

template <typename T, int R>
__device__ __forceinline__ void span_stage_pad(const Span<T, R> &t, T *__restrict__ buf,
                                               int lane, T (&x)[R]) {
  using V = typename Vec16<T>::type;
  using S = PadStage<T, R>;
  V *s = reinterpret_cast<V *>(buf);
#pragma unroll
  for (int v = 0; v < S::P; v++) s[S::slot(v * kWave + lane)] = t.v[v];
  __builtin_amdgcn_wave_barrier();
  const V *mine = s + lane * (S::P + S::PAD);
#pragma unroll
  for (int v = 0; v < S::P; v++) {
    const V w = mine[v];
#pragma unroll
    for (int e = 0; e < S::E; e++) x[S::E * v + e] = w[e];
  }
  __builtin_amdgcn_wave_barrier();
}

template <typename T, int R, int CP>
__device__ __forceinline__ void store_staged_pad(T *__restrict__ span, T *__restrict__ buf,
                                                 int lane, const T (&x)[R]) {
  using V = typename Vec16<T>::type;
  using S = PadStage<T, R>;
  V *s = reinterpret_cast<V *>(buf);
  V *mine = s + lane * (S::P + S::PAD);
#pragma unroll
  for (int v = 0; v < S::P; v++) {
    V w;
#pragma unroll
    for (int e = 0; e < S::E; e++) w[e] = x[S::E * v + e];
    mine[v] = w;
  }
  __builtin_amdgcn_wave_barrier();
  V *g = reinterpret_cast<V *>(span);
#pragma unroll
  for (int v = 0; v < S::P; v++) st_stream<CP>(g + v * kWave + lane, s[S::slot(v * kWave + lane)]);
  __builtin_amdgcn_wave_barrier();
}

// Fused right-hand side + x sweep: du <- A_x^{-1} (lambda * Laplacian u).
// One wave owns 64/cps whole x-lines of one z-plane (the host requires
// nx == cps*R, cps <= 64 and ny a multiple of the lines per wave, so a wave
// never straddles planes).  The wave's lines and their four neighbour lines
// (y-1, y+1, z-1, z+1: contiguous spans at -nx, +nx, -nx*ny, +nx*ny) are
// loaded as whole spans; the x-neighbours come from the staging buffer.
// HBM traffic: u once (neighbour spans hit L2 / MALL) and du once.
// xcd = 1: work-group b takes line block (b mod 8) * (B / 8) + b / 8 (B the
// grid rounded down to a multiple of 8), so that each XCD -- work-groups are
// dealt to the 8 XCDs round-robin -- sweeps one contiguous range of planes
// and finds the neighbour planes of its lines in its own L2.
// FR = 1 (a tuner candidate; it ties the in-kernel factors at 256^3 and loses
// at 512^3, where its larger LDS footprint lowers occupancy, since the
// kernel's time is set by its loads and transposes rather than its
// arithmetic): the x sweep runs with precomputed factors F (k_cc_setup<T, R> of
// the interior-line matrix for N = nx and these cps chunks; record of chunk
// k at F + k * nf): the chunk sweep and the in-wave reduced solve touch d
// only, with no reciprocals, and the reduced solve exchanges one value per
// level instead of three.  Lines off the grid's interior have d = 0, so the
// interior factors give them their (zero) solution too.  The work-group
// copies the table into LDS transposed (factor i of chunk k at i * cps + k),
// so a wave's factor reads are consecutive words.  CP: cache policy of the
// du stores.
// YS > 0 (nx = 64 * YS elements per 16-byte piece, so that a line is YS
// piece rows of the span and the wave holds NP / YS >= 2 lines): the y-1 and
// y+1 spans overlap the wave's own span by all but one line, so only their
// outer line is loaded and the rest is taken from the wave's registers.
template <typename T, int R, int CP, int FR, int YS = 0>
__global__ void __launch_bounds__(256)
    k_adi_rhs_x(const T *__restrict__ u, T *__restrict__ du, int nx, int ny,
                int nz, T lam, CoefConst<T> cf, int cps_log2, int xcd,
                const T *__restrict__ F) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int cps = 1 << cps_log2;
  const int k = tid & (cps - 1);
  using Ly = CCLayout<R>;
  // per-wave staging buffers, then (FR) the factor table
  constexpr int SW = PadStage<T, R>::elems;
  const T *Ft = reinterpret_cast<const T *>(smem_raw) + 4 * SW;
  if constexpr (FR) {
    const int nf = Ly::nf(cps_log2);
    T *ft = reinterpret_cast<T *>(smem_raw) + 4 * SW;
    for (int i = tid; i < nf * cps; i += 256) ft[i] = F[(i & (cps - 1)) * nf + (i >> cps_log2)];
    __syncthreads();
  }
  const int lines_per_block = 256 >> cps_log2;
  const long long nlines = (long long)ny * nz;
  const long long nblk = (nlines + lines_per_block - 1) / lines_per_block;
  // one pass per block of lines; a grid smaller than nblk (the persistent
  // launch of FR = 1, which copies the factor table once) strides over them
  for (long long b0 = blockIdx.x; b0 < nblk; b0 += gridDim.x) {
    long long blk = b0;
    if (xcd) {
      const long long nb8 = nblk & ~7LL;
      if (b0 < nb8) blk = (b0 & 7) * (nb8 >> 3) + (b0 >> 3);
    }
    // the wave's first line, its plane and its place in the plane are
    // wave-uniform (scalar registers; the host keeps ny * nz below 2^31); a
    // wave never straddles planes
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long long wline0 = blk * lines_per_block + ((wv * kWave) >> cps_log2);
    const bool wave_active = wline0 < nlines;
    const long long nxy = (long long)nx * ny;
    const unsigned wl = (unsigned)wline0;
    const int kw = (int)(wl / (unsigned)ny);
    const int jw = (int)(wl - (unsigned)kw * (unsigned)ny);  // the wave's first line in its plane
    const bool plane_in = kw >= 1 && kw <= nz - 2;
    T *stage = reinterpret_cast<T *>(smem_raw) + (tid >> 6) * SW;
    const long long wo = wline0 * nx;

    T d[R];
#pragma unroll
    for (int r = 0; r < R; r++) d[r] = T(0);
    if (wave_active && plane_in) {
      constexpr int E = Vec16<T>::n, NP = R / E;
      // all five spans in flight before the first transpose
      Span<T, R> sc, sym, syp, szm, szp;
      span_issue<T, R>(u + wo, lane, sc);
      if constexpr (YS > 0 && YS < NP) {
        // element e of the y-1 span is element e - nx = e - 64*E*YS of the
        // wave's span: piece row v of it is row v - YS of sc (v >= YS); the
        // y+1 span's row v is row v + YS of sc (v < NP - YS)
        using V = typename Vec16<T>::type;
        const V *gm = reinterpret_cast<const V *>(u + wo - nx);
        const V *gp = reinterpret_cast<const V *>(u + wo + nx);
#pragma unroll
        for (int v = 0; v < YS; v++) sym.v[v] = gm[v * kWave + lane];
#pragma unroll
        for (int v = NP - YS; v < NP; v++) syp.v[v] = gp[v * kWave + lane];
#pragma unroll
        for (int v = YS; v < NP; v++) sym.v[v] = sc.v[v - YS];
#pragma unroll
        for (int v = 0; v < NP - YS; v++) syp.v[v] = sc.v[v + YS];
      } else {
        span_issue<T, R>(u + wo - nx, lane, sym);
        span_issue<T, R>(u + wo + nx, lane, syp);
      }
      span_issue<T, R>(u + wo - nxy, lane, szm);
      span_issue<T, R>(u + wo + nxy, lane, szp);
      // the stencil is evaluated in the loaded layout (lane holds elements
      // (v*64 + lane)*E .. +E-1 of the wave's span): the neighbour spans line
      // up element for element, the x-neighbours at piece ends come from the
      // adjacent lanes; only the right-hand side is transposed into chunks
      const int nx_log2 = __builtin_ctz(nx);  // nx = cps * R, a power of two
      Span<T, R> sd;
#pragma unroll
      for (int v = 0; v < NP; v++) {
        const auto cv = sc.v[v];
        T left = lane_read(cv[E - 1], lane > 0 ? lane - 1 : 0);
        T right = lane_read(cv[0], lane < kWave - 1 ? lane + 1 : kWave - 1);
        const T prev_last = v > 0 ? lane_bcast(sc.v[v > 0 ? v - 1 : 0][E - 1], kWave - 1) : T(0);
        const T next_first = v + 1 < NP ? lane_bcast(sc.v[v + 1 < NP ? v + 1 : 0][0], 0) : T(0);
        if (lane == 0) left = prev_last;
        if (lane == kWave - 1) right = next_first;
#pragma unroll
        for (int j = 0; j < E; j++) {
          const int e = (v * kWave + lane) * E + j;
          const int row = e & (nx - 1);
          const int jl_e = jw + (e >> nx_log2);
          const bool in = jl_e >= 1 && jl_e <= ny - 2 && row >= 1 && row + 2 <= nx;
          const T l = j == 0 ? left : cv[j > 0 ? j - 1 : 0];
          const T rr = j == E - 1 ? right : cv[j + 1 < E ? j + 1 : 0];
          sd.v[v][j] = in ? rhs_point(l, rr, sym.v[v][j], syp.v[v][j], szm.v[v][j], szp.v[v][j],
                                      cv[j], lam)
                          : T(0);
        }
      }
      span_stage_pad<T, R>(sd, stage, lane, d);
    }
    T y[R];
    if constexpr (FR) {
      // factor i of this lane's chunk (the empty asm keeps the table reads
      // after the staged right-hand side, as for the coefficients below)
      int kf = k;
      asm volatile("" : "+v"(kf) : "v"(d[0]));
      auto fv = [&](int i) { return Ft[i * cps + kf]; };
      // chunk sweep (d only), the recurrences of k_cc_sweep
      T dp = d[1] * fv(Ly::R1);
      d[1] = dp;
#pragma unroll
      for (int i = 2; i < R; i++) {
        dp = (d[i] - fv(Ly::A + i - 2) * dp) * fv(Ly::RI + i - 2);
        d[i] = dp;
      }
      const T ld = d[R - 1];
#pragma unroll
      for (int i = R - 3; i >= 1; i--) d[i] = d[i] - fv(Ly::CB + i - 1) * d[i + 1];
      const T fd = (d[0] - fv(Ly::C0) * d[1]) * fv(Ly::R0);
      // reduced system over the line's chunks (consecutive lanes)
      T fnd = lane_read(fd, k + 1 < cps ? lane + 1 : lane);
      if (k + 1 >= cps) fnd = T(0);
      T rd = (ld - fv(Ly::LA) * fd - fv(Ly::LC) * fnd) * fv(Ly::INV);
      int j = 0;
      for (int st = 1; st < cps; st <<= 1, j++) {
        T md = lane_read(rd, k >= st ? lane - st : lane);
        T pd = lane_read(rd, k + st < cps ? lane + st : lane);
        if (k < st) md = T(0);
        if (k + st >= cps) pd = T(0);
        rd = (rd - fv(Ly::PCR + 3 * j) * md - fv(Ly::PCR + 3 * j + 1) * pd) *
             fv(Ly::PCR + 3 * j + 2);
      }
      T xm = lane_read(rd, k > 0 ? lane - 1 : lane);
      if (k == 0) xm = T(0);
      const T xf = fd - fv(Ly::FA) * xm - fv(Ly::FC) * rd;
      y[0] = xf;
      y[R - 1] = rd;
#pragma unroll
      for (int i = 1; i < R - 1; i++) y[i] = d[i] - fv(Ly::EA + i - 1) * xf - fv(Ly::EC + i - 1) * rd;
    } else {
      // the interior line's matrix: constant rows, identity rows at the
      // grid's first row (row 0 of chunk 0) and last row (row R-1 of chunk
      // cps-1).  Lines off the interior have d = 0, so this matrix gives them
      // their zero solution too (as the factor form above).  The empty asm
      // ties the coefficients to the staged right-hand side: hoisted above
      // the span loads, the factor arithmetic would hold its results across
      // their latency (163 against 104 VGPRs at R = 8, FP64).
      T ca = cf.a, cb = cf.b, cc = cf.c;
      asm volatile("" : "+v"(ca), "+v"(cb), "+v"(cc) : "v"(d[0]));
      T a[R], b[R], c[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        a[r] = ca;
        b[r] = cb;
        c[r] = cc;
      }
      if (k == 0) {
        a[0] = T(0);
        b[0] = T(1);
        c[0] = T(0);
      }
      if (k == cps - 1) {
        a[R - 1] = T(0);
        b[R - 1] = T(1);
        c[R - 1] = T(0);
      }
      Row<T> f, l;
      chunk_reduce<T, R>(a, b, c, d, f, l);
      T xf, xl;
      reduced_wave<T>(f, l, k, cps, lane, xf, xl);
      chunk_expand<T, R>(a, c, d, xf, xl, y);
    }
    if (wave_active) store_staged_pad<T, R, CP>(du + wo, stage, lane, y);
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
  D.small_index = D.nsys < (1LL << 31) && !wide_index();
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
  int fused_R = 0;          // >0: fused RHS + x sweep with this chunk length
  int fused_cps_log2 = 0;
  int fused_xcd = 0;        // XCD-contiguous block order of the fused kernel
  int fused_cc = 0;         // fused x sweep with precomputed factors
  void *fused_F[3] = {nullptr, nullptr, nullptr};  // their tables, chunk length 4, 8, 16
  int cp = 0;               // cache policy of every stage's streaming accesses
  // y and z sweeps with precomputed factors (trid_cc.hpp): chunk length (0 =
  // the generic kernel), chunks per system and the factor table
  int cc_R[3] = {0, 0, 0};
  int cc_cps_log2[3] = {0, 0, 0};
  int cc_xcd[3] = {0, 0, 0};
  int cc_pair[3] = {0, 0, 0};  // two adjacent systems per lane (16-byte rows)
  void *cc_F[3] = {nullptr, nullptr, nullptr};
  bool tune_pending = false;
  hipGraph_t graph = nullptr;  // tridAdiGraphCapture's graph
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  // tridAdiStep's graph of step_n steps on (step_u, step_du)
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  void *step_u = nullptr, *step_du = nullptr;
  int step_n = 0;
  std::string text;
};

namespace {

template <typename T>
CoefConst<T> coefs(const tridAdiPlan_st *p) {
  const T lam = (T)p->lambda;
  if (p->mode == 0) return CoefConst<T>{(T)-0.5f * lam, (T)1.0f + lam, (T)-0.5f * lam};
  return CoefConst<T>{-lam, (T)1.0f + (T)2.0f * lam, -lam};
}

int fused_slot(int R) { return R == 4 ? 0 : (R == 8 ? 1 : 2); }

template <typename T>
hipError_t launch_fused(tridAdiPlan_st *p, T *u, T *du, hipStream_t st) {
  const CoefConst<T> cf = coefs<T>(p);
  const int lines_per_block = 256 >> p->fused_cps_log2;
  const long long nlines = (long long)p->ny * p->nz;
  const unsigned blocks = (unsigned)((nlines + lines_per_block - 1) / lines_per_block);
  // four padded wave staging buffers
  const int R = p->fused_R;
  size_t lds = 4 * sizeof(T) *
               (size_t)(R == 4 ? PadStage<T, 4>::elems
                               : (R == 8 ? PadStage<T, 8>::elems : PadStage<T, 16>::elems));
  const T *F = static_cast<const T *>(p->fused_F[fused_slot(p->fused_R)]);
  const int fr = p->fused_cc && F;
  if (fr) {
    // the transposed factor table after the staging buffers
    const int cl = p->fused_cps_log2;
    const int nf = p->fused_R == 4 ? CCLayout<4>::nf(cl)
                                   : (p->fused_R == 8 ? CCLayout<8>::nf(cl) : CCLayout<16>::nf(cl));
    lds += (size_t)nf * (1u << cl) * sizeof(T);
  }
  // the factor-table form runs persistent: as many work-groups as are
  // resident at once, each copying the table into LDS once
  // piece rows per line when a line is a whole number (< NP) of them
  const int E = Vec16<T>::n;
  const int ys = p->nx % (kWave * E) == 0 ? p->nx / (kWave * E) : 0;
#define TRID_FUSED_K(RR, CC, FF, YY)                                                          \
  do {                                                                                        \
    auto kern = k_adi_rhs_x<T, RR, CC, FF, YY>;                                               \
    unsigned g = blocks;                                                                      \
    if (FF) g = std::min(g, resident_blocks(reinterpret_cast<const void *>(kern), 256, lds)); \
    (void)launch_kernel(kern, dim3(g), dim3(256), (std::uint32_t)lds, st, u, du, p->nx,      \
                        p->ny, p->nz, (T)p->lambda, cf, p->fused_cps_log2, p->fused_xcd, F); \
  } while (0)
  // the shared-y-span forms (YS = 1, 2, 4 where YS < R / E); everything
  // else loads all five spans
#define TRID_FUSED(RR, CC, FF)                                             \
  do {                                                                     \
    constexpr int NP_ = RR / Vec16<T>::n;                                  \
    if constexpr (NP_ > 1)                                                 \
      if (ys == 1) { TRID_FUSED_K(RR, CC, FF, 1); break; }                 \
    if constexpr (NP_ > 2)                                                 \
      if (ys == 2) { TRID_FUSED_K(RR, CC, FF, (NP_ > 2 ? 2 : 0)); break; } \
    if constexpr (NP_ > 4)                                                 \
      if (ys == 4) { TRID_FUSED_K(RR, CC, FF, (NP_ > 4 ? 4 : 0)); break; } \
    TRID_FUSED_K(RR, CC, FF, 0);                                           \
  } while (0)
#define TRID_FUSED_R(CC, FF)                        \
  if (p->fused_R == 8) TRID_FUSED(8, CC, FF);       \
  else if (p->fused_R == 4) TRID_FUSED(4, CC, FF);  \
  else TRID_FUSED(16, CC, FF);
  if (p->cp == kCached) {
    if (fr) {
      TRID_FUSED_R(kCached, 1)
    } else {
      TRID_FUSED_R(kCached, 0)
    }
  } else {
    if (fr) {
      TRID_FUSED_R(kStream, 1)
    } else {
      TRID_FUSED_R(kStream, 0)
    }
  }
#undef TRID_FUSED_R
#undef TRID_FUSED
#undef TRID_FUSED_K
  return hipGetLastError();
}

// Factor-reuse sweeps: chunk length R in {8, 16, 32} with at most 16 chunks
// per system (work-groups of 64 systems x cps chunks).
bool cc_rows_ok(int N, int R, int *cps_log2) {
  if (R != 8 && R != 16 && R != 32) return false;
  const long long ch = (N + R - 1) / R;
  int cl = 0;
  while ((1LL << cl) < ch) cl++;
  if ((kWave << cl) > cc_max_threads(R, 1)) return false;
  *cps_log2 = cl;
  return true;
}

// Two systems per lane fit chunk lengths up to 32 rows in work-groups within
// the kernel's launch bound.
bool cc_pair_fits(int R, int cps_log2) {
  return R <= 32 && (kWave << cps_log2) <= cc_max_threads(R, 2);
}

// Builds the factor table of direction dir for chunk length R into *F
// (allocated here; waits for the setup kernel).
template <typename T>
tridStatus_t cc_build(tridAdiPlan_st *p, int dir, int R, void **F, int *cps_log2) {
  const int N = p->desc[dir].N;
  if (!cc_rows_ok(N, R, cps_log2)) return fail(TRID_STATUS_INVALID_VALUE, "cc layout");
  const int cps = 1 << *cps_log2;
  size_t nf = 0;
  if (R == 8) nf = CCLayout<8>::nf(*cps_log2);
  else if (R == 16) nf = CCLayout<16>::nf(*cps_log2);
  else nf = CCLayout<32>::nf(*cps_log2);
  tridStatus_t s = check_hip(hipMalloc(F, cps * nf * sizeof(T)), "hipMalloc");
  if (s) return s;
  const CoefConst<T> cf = coefs<T>(p);
  // one wave (cps <= 16)
  hipStream_t st = current_stream();
  if (R == 8)
    hipLaunchKernelGGL((k_cc_setup<T, 8>), dim3(1), dim3(kWave), 0, st, cf, N, *cps_log2,
                       static_cast<T *>(*F));
  else if (R == 16)
    hipLaunchKernelGGL((k_cc_setup<T, 16>), dim3(1), dim3(kWave), 0, st, cf, N, *cps_log2,
                       static_cast<T *>(*F));
  else
    hipLaunchKernelGGL((k_cc_setup<T, 32>), dim3(1), dim3(kWave), 0, st, cf, N, *cps_log2,
                       static_cast<T *>(*F));
  s = check_hip(hipGetLastError(), "factor setup launch");
  if (!s) s = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  if (s) {
    (void)hipFree(*F);
    *F = nullptr;
  }
  return s;
}

// Factor table of the fused x sweep with chunk length R (fused_rows_ok(p, R)
// holds: nx = R * cps, cps <= 64), built once per plan and chunk length into
// p->fused_F (waits for the setup kernel).
template <typename T>
tridStatus_t fused_table(tridAdiPlan_st *p, int R) {
  void *&F = p->fused_F[fused_slot(R)];
  if (F) return TRID_STATUS_SUCCESS;
  const int cl = ilog2(p->nx / R), cps = 1 << cl;
  const size_t nf = R == 4 ? CCLayout<4>::nf(cl) : (R == 8 ? CCLayout<8>::nf(cl) : CCLayout<16>::nf(cl));
  tridStatus_t s = check_hip(hipMalloc(&F, cps * nf * sizeof(T)), "hipMalloc");
  if (s) {
    F = nullptr;
    return s;
  }
  const CoefConst<T> cf = coefs<T>(p);
  hipStream_t st = current_stream();
  T *f = static_cast<T *>(F);
  if (R == 4)
    hipLaunchKernelGGL((k_cc_setup<T, 4>), dim3(1), dim3(kWave), 0, st, cf, p->nx, cl, f);
  else if (R == 8)
    hipLaunchKernelGGL((k_cc_setup<T, 8>), dim3(1), dim3(kWave), 0, st, cf, p->nx, cl, f);
  else
    hipLaunchKernelGGL((k_cc_setup<T, 16>), dim3(1), dim3(kWave), 0, st, cf, p->nx, cl, f);
  s = check_hip(hipGetLastError(), "factor setup launch");
  if (!s) s = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  if (s) {
    (void)hipFree(F);
    F = nullptr;
  }
  return s;
}

tridStatus_t fused_table_any(tridAdiPlan_st *p, int R) {
  return p->dbl ? fused_table<double>(p, R) : fused_table<float>(p, R);
}

// Two adjacent systems per lane need the systems' first other dimension to
// be unit-stride in d (and u), an even count of it, and 16-byte alignment.
template <typename T>
bool cc_pair_ok(const SysDesc &D, const T *d, const T *u, int inc) {
  if (D.nother < 1 || D.ostr[3][0] != 1 || D.osize[0] % 2) return false;
  if (inc && D.ostr[4][0] != 1) return false;
  if (!aligned16(d) || (inc && !aligned16(u))) return false;
  return D.nsys % 2 == 0;
}

template <typename T, int INC>
hipError_t launch_cc_t(int R, int cps_log2, int xcd, int pair, const void *F, const SysDesc &D,
                       int idx32, int cp, T *d, T *u, hipStream_t st) {
  const int cps = 1 << cps_log2;
  const int ns = pair ? 2 : 1;
  const long long per = (long long)kWave * ns;
  const unsigned blocks = (unsigned)((D.nsys + per - 1) / per);
  const size_t lds = 2 * (size_t)ns * kWave * cps * sizeof(T);
  const dim3 grid(blocks), block(kWave * cps);
  const T *f = static_cast<const T *>(F);
#define TRID_CC(RR, II, NN, CC)                                                           \
  (void)launch_kernel(k_cc_sweep<T, INC, RR, II, NN, CC>, grid, block, (std::uint32_t)lds, st, \
                      f, d, u, D, cps_log2, xcd)
#define TRID_CC_R(II, NN, CC)             \
  if (R == 8) TRID_CC(8, II, NN, CC);       \
  else if (R == 16) TRID_CC(16, II, NN, CC); \
  else TRID_CC(32, II, NN, CC);
  // pairs: chunk lengths 8, 16 and 32 (cc_pair_fits); the cached policy
  // with 32-bit offsets only
  if (idx32 && cp == kCached) {
    if (pair) {
      TRID_CC_R(uint32_t, 2, kCached)
    } else {
      TRID_CC_R(uint32_t, 1, kCached)
    }
  } else if (idx32) {
    if (pair) {
      TRID_CC_R(uint32_t, 2, kStream)
    } else {
      TRID_CC_R(uint32_t, 1, kStream)
    }
  } else {
    if (pair) {
      TRID_CC_R(long long, 2, kStream)
    } else {
      TRID_CC_R(long long, 1, kStream)
    }
  }
#undef TRID_CC_R
#undef TRID_CC
  return hipGetLastError();
}

template <typename T, int INC>
hipError_t launch_cc(const tridAdiPlan_st *p, int dir, T *d, T *u, hipStream_t st) {
  const int pair = p->cc_pair[dir] && cc_pair_fits(p->cc_R[dir], p->cc_cps_log2[dir]) &&
                   cc_pair_ok<T>(p->desc[dir], d, u, INC);
  return launch_cc_t<T, INC>(p->cc_R[dir], p->cc_cps_log2[dir], p->cc_xcd[dir], pair,
                             p->cc_F[dir], p->desc[dir], p->cfg[dir].idx32, p->cp, d, u, st);
}

// Stage 4 = fused RHS + x sweep (plans with fused_R > 0 only).
template <typename T>
tridStatus_t stage_t(tridAdiPlan_st *p, T *u, T *du, int stage, hipStream_t st) {
  hipError_t e = hipSuccess;
  const CoefConst<T> cf = coefs<T>(p);
  if (stage == 4) {
    e = launch_fused<T>(p, u, du, st);
  } else if (stage == 0) {
    dim3 grid((p->nx + 63) / 64, (p->ny + 3) / 4, (p->nz + kRhsPlanes - 1) / kRhsPlanes);
    if (p->cp == kCached)
      e = launch_kernel(k_adi_rhs<T, kCached>, grid, dim3(64, 4), 0u, st, (const T *)u, du,
                        p->nx, p->ny, p->nz, (T)p->lambda);
    else
      e = launch_kernel(k_adi_rhs<T, kStream>, grid, dim3(64, 4), 0u, st, (const T *)u, du,
                        p->nx, p->ny, p->nz, (T)p->lambda);
  } else if (stage == 1) {
    PartCfg c = aligned16(du) ? p->cfg[0] : p->cfg_x_unaligned;
    c.cp = p->cp;
    e = launch_part<T, 0>(c, cf, du, du, p->desc[0], st);
  } else if (stage == 2) {
    PartCfg c = p->cfg[1];
    c.cp = p->cp;
    e = p->cc_R[1] ? launch_cc<T, 0>(p, 1, du, du, st)
                   : launch_part<T, 0>(c, cf, du, du, p->desc[1], st);
  } else {
    PartCfg c = p->cfg[2];
    c.cp = p->cp;
    e = p->cc_R[2] ? launch_cc<T, 1>(p, 2, du, u, st)
                   : launch_part<T, 1>(c, cf, du, u, p->desc[2], st);
  }
  return check_hip(e, "ADI kernel launch");
}

tridStatus_t stage_any(tridAdiPlan_st *p, void *u, void *du, int stage, hipStream_t st) {
  if (p->dbl) return stage_t<double>(p, (double *)u, (double *)du, stage, st);
  return stage_t<float>(p, (float *)u, (float *)du, stage, st);
}

bool use_fused(const tridAdiPlan_st *p, const void *u, const void *du) {
  return p->fused_R > 0 && aligned16(u) && aligned16(du);
}

tridStatus_t steps_any(tridAdiPlan_st *p, void *u, void *du, int nsteps, hipStream_t st) {
  const bool fused = use_fused(p, u, du);
  for (int s = 0; s < nsteps; s++) {
    if (fused) {
      for (int g : {4, 2, 3}) {
        tridStatus_t r = stage_any(p, u, du, g, st);
        if (r) return r;
      }
    } else {
      for (int g = 0; g < 4; g++) {
        tridStatus_t r = stage_any(p, u, du, g, st);
        if (r) return r;
      }
    }
  }
  return TRID_STATUS_SUCCESS;
}

// The fused kernel needs nx == cps * R with cps <= 64 a power of two, ny a
// multiple of the lines per wave (64 / cps) and fewer than 2^31 lines (its
// line indices are 32-bit).
bool fused_rows_ok(const tridAdiPlan_st *p, int R) {
  if (R != 4 && R != 8 && R != 16) return false;
  if ((long long)p->ny * p->nz >= (1LL << 31)) return false;
  if (p->nx % R) return false;
  const int cps = p->nx / R;
  if (cps > kWave || (cps & (cps - 1))) return false;
  return p->ny % (kWave / cps) == 0;
}

void set_fused(tridAdiPlan_st *p, int R) {
  p->fused_R = R;
  p->fused_cps_log2 = ilog2(p->nx / R);
}

// Times every candidate configuration of the y and z sweeps (and of the x
// sweep, and the chunk length of the fused RHS + x kernel) on scratch grids
// and keeps the fastest.  Data values do not change
// the kernels' work, so zero-filled scratch is enough.
// Times the candidates of one stage inside the time step: for every timed
// step, prefix() runs the stages before it, the candidate runs between two
// events and suffix() runs the stages after it, so the candidate sees the
// caches as a real step leaves them.  `passes` interleaved rounds of one
// warm-up step and `reps` timed steps; returns the candidate with the lowest
// time of its best round (-1 if none ran).
template <typename Prefix, typename Cand, typename Suffix>
int pick_in_step(int ncand, Prefix prefix, Cand cand, Suffix suffix, hipStream_t st,
                 std::vector<float> *times = nullptr, int reps = 4, int passes = 2) {
  std::vector<hipEvent_t> ev(2 * reps, nullptr);
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      for (auto &f : ev)
        if (f) (void)hipEventDestroy(f);
      (void)hipGetLastError();
      return -1;
    }
  std::vector<float> best(ncand, 1e30f);
  for (int pass = 0; pass < passes; pass++)
    for (int i = 0; i < ncand; i++) {
      prefix();
      if (cand(i) != hipSuccess) {
        (void)hipGetLastError();
        continue;
      }
      suffix();
      for (int r = 0; r < reps; r++) {
        prefix();
        (void)hipEventRecord(ev[2 * r], st);
        (void)cand(i);
        (void)hipEventRecord(ev[2 * r + 1], st);
        suffix();
      }
      if (hipEventSynchronize(ev[2 * reps - 1]) != hipSuccess) break;
      float tot = 0.f;
      for (int r = 0; r < reps; r++) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]);
        tot += ms;
      }
      best[i] = std::min(best[i], tot);
    }
  for (auto &e : ev) (void)hipEventDestroy(e);
  (void)hipGetLastError();
  if (times) *times = best;
  return stable_pick(best);
}

// TRIDHIP_TUNE_LOG=1: the plan tuner prints every candidate's time.
bool tune_log() {
  const char *e = std::getenv("TRIDHIP_TUNE_LOG");
  return e && e[0] == '1';
}

// Tunes the plan on the caller's own arrays (the speed of large power-of-two
// strides depends on the actual addresses): the y sweep, then the z sweep,
// then the chunk length of the fused RHS + x kernel, each timed inside the
// time step.  u and du are saved first and restored at the end.
template <typename T>
void tune_plan_t(tridAdiPlan_st *p, T *u, T *du) {
  const size_t npts = (size_t)p->nx * p->ny * p->nz, bytes = npts * sizeof(T);
  T *save = nullptr;
  if (hipMalloc(&save, 2 * bytes) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  hipStream_t st = current_stream();
  (void)hipMemcpyAsync(save, u, bytes, hipMemcpyDeviceToDevice, st);
  (void)hipMemcpyAsync(save + npts, du, bytes, hipMemcpyDeviceToDevice, st);
  const CoefConst<T> cf = coefs<T>(p);
  const int n[3] = {p->nx, p->ny, p->nz};
  auto stage = [&](int g) { (void)stage_t<T>(p, u, du, g, st); };
  auto x_stage = [&]() {
    if (use_fused(p, u, du)) {
      stage(4);
    } else {
      stage(0);
      stage(1);
    }
  };
  // generic x sweep (the unfused path): timed on its own
  {
    const int vec = (p->nx * sizeof(T)) % 16 == 0;
    std::vector<PartCfg> cands;
    for (PartCfg c : candidates(0, n[0], vec, sizeof(T), true, p->desc[0].nsys)) {
      if (c.vec != 2) continue;  // x keeps the staged layout
      c.idx32 = p->cfg[0].idx32;
      cands.push_back(c);
    }
    auto run = [&](int i) { return launch_part<T, 0>(cands[i], cf, du, du, p->desc[0], st); };
    const int i = pick_fastest((int)cands.size(), [](int) { return true; }, run, st);
    if (i >= 0) p->cfg[0] = cands[i];
  }
  // y and z: generic kernels and the factor-reuse kernel at every chunk
  // length that fits, in both work-group orders
  for (int dir = 1; dir < 3; dir++) {
    struct Choice {
      int cc, R, cl, xcd, pair;
      PartCfg cfg;
      void *F;
    };
    std::vector<Choice> gen, fac;
    for (PartCfg c : candidates(1, n[dir], 0, sizeof(T), true, p->desc[dir].nsys)) {
      c.idx32 = p->cfg[dir].idx32;
      gen.push_back({0, 0, 0, 0, 0, c, nullptr});
    }
    const bool pair_ok = cc_pair_ok<T>(p->desc[dir], du, u, dir == 2);
    if (p->cc_F[dir]) (void)hipFree(p->cc_F[dir]);
    p->cc_F[dir] = nullptr;
    p->cc_R[dir] = 0;
    std::vector<void *> tables;
    for (int R : {16, 32, 8}) {
      int cl = 0;
      void *F = nullptr;
      if (!cc_rows_ok(n[dir], R, &cl)) continue;
      if (cc_build<T>(p, dir, R, &F, &cl) != TRID_STATUS_SUCCESS) continue;
      tables.push_back(F);
      // two lines per lane first: the usual winner
      for (int pr = (pair_ok && cc_pair_fits(R, cl) ? 1 : 0); pr >= 0; pr--) {
        fac.push_back({1, R, cl, 0, pr, p->cfg[dir], F});
        fac.push_back({1, R, cl, 1, pr, p->cfg[dir], F});
      }
    }
    // the pick is the first candidate within 2 % of the fastest, so the
    // usual winner goes first: the partitioned kernel for y, the factor-reuse
    // kernel for z (at 256^3 FP64 66-67 us against 69 us in kernel events)
    std::vector<Choice> ch = dir == 2 ? fac : gen;
    for (const Choice &c : dir == 2 ? gen : fac) ch.push_back(c);
    if (ch.empty()) continue;
    const PartCfg keep = p->cfg[dir];
    auto apply = [&](const Choice &c) {
      if (c.cc) {
        p->cc_R[dir] = c.R;
        p->cc_cps_log2[dir] = c.cl;
        p->cc_xcd[dir] = c.xcd;
        p->cc_pair[dir] = c.pair;
        p->cc_F[dir] = c.F;
        p->cfg[dir] = keep;
      } else {
        p->cc_R[dir] = 0;
        p->cc_F[dir] = nullptr;
        p->cfg[dir] = c.cfg;
      }
    };
    auto prefix = [&]() {
      x_stage();
      if (dir == 2) stage(2);
    };
    auto cand = [&](int i) {
      apply(ch[i]);
      return stage_t<T>(p, u, du, dir == 1 ? 2 : 3, st) == TRID_STATUS_SUCCESS
                 ? hipSuccess
                 : hipErrorLaunchFailure;
    };
    auto suffix = [&]() {
      if (dir == 1) stage(3);
    };
    // the stages around the candidate use the current choices of the
    // other directions; start from the first candidate of this one
    apply(ch[0]);
    std::vector<float> tm;
    const int i = pick_in_step((int)ch.size(), prefix, cand, suffix, st, &tm);
    if (tune_log())
      for (size_t c = 0; c < ch.size(); c++)
        fprintf(stderr, "[tridhip tune] %s sweep %s%s%s: %.4f ms\n", dir == 1 ? "y" : "z",
                ch[c].cc ? ("cc<R" + std::to_string(ch[c].R) + ">").c_str()
                         : cfg_name(ch[c].cfg, sizeof(T)).c_str(),
                ch[c].cc && ch[c].xcd ? ",xcd" : "", ch[c].cc && ch[c].pair ? ",pair" : "",
                tm[c]);
    (void)hipStreamSynchronize(st);
    apply(ch[i >= 0 ? i : 0]);
    for (void *F : tables)
      if (F != p->cc_F[dir]) (void)hipFree(F);
  }
  // chunk length, block order and precomputed factors of the fused RHS + x
  // kernel (the untuned choice first)
  if (p->fused_R) {
    struct FusedChoice {
      int R, xcd, cc;
    };
    std::vector<FusedChoice> ch{{p->fused_R, p->fused_xcd, p->fused_cc}};
    for (int cc : {1, 0})
      for (int R : {4, 8, 16})
        for (int x : {0, 1}) {
          if (!fused_rows_ok(p, R)) continue;
          if (R == p->fused_R && x == p->fused_xcd && cc == p->fused_cc) continue;
          if (cc && fused_table<T>(p, R) != TRID_STATUS_SUCCESS) continue;
          ch.push_back({R, x, cc});
        }
    auto apply = [&](const FusedChoice &c) {
      set_fused(p, c.R);
      p->fused_xcd = c.xcd;
      p->fused_cc = c.cc;
    };
    auto cand = [&](int i) {
      apply(ch[i]);
      return launch_fused<T>(p, u, du, st);
    };
    auto suffix = [&]() {
      stage(2);
      stage(3);
    };
    std::vector<float> tm;
    const int i = pick_in_step((int)ch.size(), []() {}, cand, suffix, st, &tm);
    if (tune_log())
      for (size_t c = 0; c < ch.size(); c++)
        fprintf(stderr, "[tridhip tune] rhs+x fused R%d%s%s: %.4f ms\n", ch[c].R,
                ch[c].cc ? ",cc" : "", ch[c].xcd ? ",xcd" : "", tm[c]);
    apply(ch[i >= 0 ? i : 0]);
  }
  // cache policy of the whole step (the other policy changes what every
  // stage finds in the last-level cache, so whole steps are timed): the
  // footprint rule's choice first, kept unless the other is more than 2 %
  // faster
  if (p->cfg[0].idx32 && forced_cache_policy() < 0) {
    const int pol[2] = {p->cp, p->cp == kCached ? kStream : kCached};
    auto cand = [&](int i) {
      p->cp = pol[i];
      x_stage();
      stage(2);
      stage(3);
      return hipSuccess;
    };
    std::vector<float> tm;
    const int i = pick_in_step(2, []() {}, cand, []() {}, st, &tm);
    if (tune_log())
      for (int c = 0; c < 2; c++)
        fprintf(stderr, "[tridhip tune] step, cache policy %s: %.4f ms\n",
                pol[c] == kCached ? "cached" : "stream", tm[c]);
    p->cp = pol[i >= 0 ? i : 0];
  }
  (void)hipMemcpyAsync(u, save, bytes, hipMemcpyDeviceToDevice, st);
  (void)hipMemcpyAsync(du, save + npts, bytes, hipMemcpyDeviceToDevice, st);
  (void)hipStreamSynchronize(st);
  (void)hipFree(save);
  (void)hipGetLastError();
}

std::string plan_text(const tridAdiPlan_st *p) {
  const size_t ts = p->dbl ? 8 : 4;
  const std::string cached = p->cp == kCached ? ",cached" : "";
  auto part_name = [&](int dir) {
    PartCfg c = p->cfg[dir];
    c.cp = p->cp;
    return cfg_name(c, ts);
  };
  auto sweep_name = [&](int dir) {
    if (!p->cc_R[dir]) return part_name(dir);
    return "cc<R" + std::to_string(p->cc_R[dir]) + ",cps" +
           std::to_string(1 << p->cc_cps_log2[dir]) + (p->cc_xcd[dir] ? ",xcd" : "") +
           (p->cc_pair[dir] ? ",pair" : "") + cached + ">";
  };
  std::string fx = "rhs" + cached + "<";
  if (p->fused_R)
    fx = "rhs+x(fused,R" + std::to_string(p->fused_R) + (p->fused_cc ? ",cc" : "") +
         (p->fused_xcd ? ",xcd" : "") + cached + ")<";
  return fx + (p->dbl ? "double" : "float") + "> x:" + part_name(0) + " y:" + sweep_name(1) +
         " z:" + sweep_name(2);
}

// With autotuning on, a plan tunes at its first use, on that call's arrays,
// and never inside a graph capture: tridAdiGraphCapture tunes before it
// starts capturing.  A plan whose first use is on a stream that a caller is
// capturing stays untuned: tuning replaces the factor tables, whose
// addresses the caller's graph holds.
void ensure_tuned(tridAdiPlan_st *p, void *u, void *du) {
  if (!p->tune_pending) return;
  p->tune_pending = false;
  if (capturing(current_stream())) return;
  if (p->dbl)
    tune_plan_t<double>(p, static_cast<double *>(u), static_cast<double *>(du));
  else
    tune_plan_t<float>(p, static_cast<float *>(u), static_cast<float *>(du));
  p->text = plan_text(p);
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
  const int idx32 = (unsigned long long)nx * ny * nz * ts < (1ULL << 32) && !wide_index();
  // cache policy from the step's working set, u and du
  p->cp = idx32 ? cache_policy(2 * (size_t)nx * ny * nz * ts) : kStream;
  if (p->cp == kMixed) p->cp = kCached;  // no coefficient arrays: d and u cached
  const int n[3] = {nx, ny, nz};
  for (int dir = 0; dir < 3; dir++) {
    p->desc[dir] = grid_desc(nx, ny, nz, dir);
    const int order = dir == 0 ? 0 : 1;
    const int vec = dir == 0 && (nx * ts) % 16 == 0;
    PartCfg c = heuristic(order, n[dir], vec, ts);
    if (order == 1 && c.kind == 1) {
      // y and z sweeps that run without precomputed factors (by default,
      // lines past 512 rows; any line when the plan's sweep is pinned to
      // rows = 0 or the tuner picks a non-factor candidate): the strided
      // batches' layout, most lines per work-group
      const PartCfg s = strided_default(n[dir], ts);
      if (s.kind == 1) c = s;
    }
    if (dir == 0) {
      PartCfg cu = heuristic(order, n[dir], 0, ts);
      cu.idx32 = idx32;
      p->cfg_x_unaligned = cu;
      if (stage_ok(c, true, p->desc[0].nsys, nx)) c.vec = staged_vec(c);
    }
    c.idx32 = idx32;
    p->cfg[dir] = c;
    if (c.kind != 1 || (dir == 0 && p->cfg_x_unaligned.kind != 1)) {
      delete p;
      return fail(TRID_STATUS_INVALID_VALUE, "ADI grid dimensions must be at most 8192");
    }
  }
  // fused RHS + x sweep: whole-chunk lines in one wave, waves within a plane;
  // by default with the x sweep's chunk length, else the first that fits
  if ((nx * ts) % 16 == 0) {
    for (int R : {p->cfg[0].R, 8, 4, 16})
      if (fused_rows_ok(p, R)) {
        set_fused(p, R);
        break;
      }
  }
  // y and z sweeps: factor reuse by default when a chunk layout fits
  for (int dir = 1; dir < 3; dir++) {
    for (int R : {16, 8, 32}) {
      int cl = 0;
      if (!cc_rows_ok(n[dir], R, &cl)) continue;
      void *F = nullptr;
      const tridStatus_t r = p->dbl ? cc_build<double>(p, dir, R, &F, &cl)
                                    : cc_build<float>(p, dir, R, &F, &cl);
      if (r == TRID_STATUS_SUCCESS) {
        p->cc_R[dir] = R;
        p->cc_cps_log2[dir] = cl;
        p->cc_pair[dir] = cc_pair_fits(R, cl);  // used when the arrays allow it (cc_pair_ok)
        p->cc_F[dir] = F;
      }
      break;
    }
  }
  clear_error();
  p->tune_pending = autotune_mode() == 1;
  p->text = plan_text(p);
  *plan = p;
  return TRID_STATUS_SUCCESS;
}

namespace {

// longest step graph tridAdiStep captures (longer runs replay it)
constexpr int kStepGraphMax = 64;

// Drops the plan's graphs (both hold the factor tables and the kernel
// choices of the time they were captured), after any launch of them ended.
void drop_graphs(tridAdiPlan_st *p, bool explicit_graph) {
  p->step_u = p->step_du = nullptr;
  p->step_n = 0;
  if (!p->step_exec && !(explicit_graph && p->exec)) return;
  (void)hipDeviceSynchronize();
  if (p->step_exec) (void)hipGraphExecDestroy(p->step_exec);
  if (p->step_graph) (void)hipGraphDestroy(p->step_graph);
  p->step_exec = nullptr;
  p->step_graph = nullptr;
  if (explicit_graph) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
  }
}

// Captures nsteps steps on (u, du) into a graph on the plan's capture stream.
tridStatus_t capture_steps(tridAdiPlan_st *p, void *u, void *du, int nsteps, hipGraph_t *graph,
                           hipGraphExec_t *exec) {
  if (!p->cap) {
    tridStatus_t s = check_hip(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking),
                               "hipStreamCreate");
    if (s) return s;
  }
  tridStatus_t s = check_hip(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal),
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
  s = check_hip(hipGraphInstantiate(exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  if (s) {
    (void)hipGraphDestroy(g);
    return s;
  }
  *graph = g;
  return TRID_STATUS_SUCCESS;
}

}  // namespace

EXTERN_C tridStatus_t tridAdiPlanDestroy(tridAdiPlan_t p) {
  if (!p) return TRID_STATUS_SUCCESS;
  drop_graphs(p, true);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  for (void *F : p->cc_F)
    if (F) (void)hipFree(F);
  for (void *F : p->fused_F)
    if (F) (void)hipFree(F);
  delete p;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridAdiStep(tridAdiPlan_t p, void *u, void *du, int nsteps,
                                  int sync) {
  clear_error();
  tridStatus_t s = check_args(p, u, du);
  if (s) return s;
  if (nsteps < 0) return fail(TRID_STATUS_INVALID_VALUE, "nsteps < 0");
  ensure_tuned(p, u, du);
  hipStream_t st = current_stream();
  // several steps: from the second consecutive call with these arrays and
  // that count on, launches of a graph of up to kStepGraphMax steps, which
  // removes the launch gaps between the kernels (a one-off call does not
  // pay for a capture); the first call, the remainder and calls on a stream
  // that is itself capturing are launched kernel by kernel
  int rest = nsteps;
  if (nsteps >= 2 && !capturing(st)) {
    const int chunk = std::min(nsteps, kStepGraphMax);
    if (p->step_u == u && p->step_du == du && p->step_n == chunk) {
      if (!p->step_exec) {
        s = capture_steps(p, u, du, chunk, &p->step_graph, &p->step_exec);
        if (s) return s;
      }
      for (; rest >= chunk; rest -= chunk) {
        s = check_hip(hipGraphLaunch(p->step_exec, st), "hipGraphLaunch");
        if (s) return s;
      }
    } else {
      drop_graphs(p, false);
      p->step_u = u;
      p->step_du = du;
      p->step_n = chunk;
    }
  }
  if (rest > 0) s = steps_any(p, u, du, rest, st);
  if (s) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridAdiStage(tridAdiPlan_t p, void *u, void *du, int stage) {
  clear_error();
  tridStatus_t s = check_args(p, u, du);
  if (s) return s;
  if (stage < 0 || stage > 4) return fail(TRID_STATUS_INVALID_VALUE, "stage must be 0..4");
  ensure_tuned(p, u, du);
  if (stage == 4 && !use_fused(p, u, du))
    return fail(TRID_STATUS_INVALID_VALUE, "this plan/layout has no fused RHS + x stage");
  return stage_any(p, u, du, stage, current_stream());
}

EXTERN_C tridStatus_t tridAdiGraphCapture(tridAdiPlan_t p, void *u, void *du,
                                          int nsteps) {
  clear_error();
  tridStatus_t s = check_args(p, u, du);
  if (s) return s;
  if (nsteps < 1) return fail(TRID_STATUS_INVALID_VALUE, "nsteps must be >= 1");
  ensure_tuned(p, u, du);
  if (p->exec) {
    (void)hipDeviceSynchronize();  // a launch of the old graph may still run
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  return capture_steps(p, u, du, nsteps, &p->graph, &p->exec);
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

EXTERN_C tridStatus_t tridAdiPlanSetSweep(tridAdiPlan_t p, int dir, int rows, int pair,
                                          int xcd) {
  if (!p) return fail(TRID_STATUS_INVALID_VALUE, "NULL plan");
  if (dir < 0 || dir > 2) return fail(TRID_STATUS_INVALID_VALUE, "dir must be 0, 1 or 2");
  if ((pair != 0 && pair != 1) || (xcd != 0 && xcd != 1))
    return fail(TRID_STATUS_INVALID_VALUE, "pair and xcd must be 0 or 1");
  if (dir == 0) {
    // the fused right-hand side + x sweep: chunk length and block order
    if (!p->fused_R) return fail(TRID_STATUS_INVALID_VALUE, "the plan has no fused x stage");
    if (!fused_rows_ok(p, rows))
      return fail(TRID_STATUS_INVALID_VALUE, "rows do not fit the fused x stage");
    if (pair) {
      const tridStatus_t s = fused_table_any(p, rows);
      if (s) return s;
    }
    drop_graphs(p, true);
    set_fused(p, rows);
    p->fused_xcd = xcd;
    p->fused_cc = pair;
    p->tune_pending = false;
    p->text = plan_text(p);
    clear_error();
    return TRID_STATUS_SUCCESS;
  }
  void *F = nullptr;
  int cl = 0;
  if (rows != 0) {
    if (!cc_rows_ok(p->desc[dir].N, rows, &cl))
      return fail(TRID_STATUS_INVALID_VALUE, "rows do not fit the line length");
    if (pair && !cc_pair_fits(rows, cl))
      return fail(TRID_STATUS_INVALID_VALUE, "two lines per lane do not fit these rows");
    const tridStatus_t s = p->dbl ? cc_build<double>(p, dir, rows, &F, &cl)
                                  : cc_build<float>(p, dir, rows, &F, &cl);
    if (s) return s;
  } else if (pair || xcd) {
    return fail(TRID_STATUS_INVALID_VALUE, "pair and xcd apply to rows 8, 16 and 32");
  }
  // graphs hold the old factor table and kernel choice; hipFree waits for
  // queued kernels that may still read the old factors
  drop_graphs(p, true);
  if (p->cc_F[dir]) (void)hipFree(p->cc_F[dir]);
  p->cc_F[dir] = F;
  p->cc_R[dir] = rows;
  p->cc_cps_log2[dir] = cl;
  p->cc_pair[dir] = pair;
  p->cc_xcd[dir] = xcd;
  p->tune_pending = false;
  p->text = plan_text(p);
  clear_error();
  return TRID_STATUS_SUCCESS;
}
