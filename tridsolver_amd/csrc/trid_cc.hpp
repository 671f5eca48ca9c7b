// trid_cc.hpp -- partitioned Thomas with precomputed factors ("factor
// reuse") for batches in which every interior system has the same constant
// tridiagonal matrix: the ADI y and z sweeps (constant a, b, c on rows
// interior to the grid, identity rows elsewhere).
//
// In the partitioned solve of trid_part.hpp every multiplier that touches d
// -- the pivot reciprocals and the sub-diagonal entries of the chunk sweep,
// the backward-substitution and expansion multipliers, the first/last row
// coefficients and the (a, c, 1/diagonal) of every reduced-system PCR level
// -- depends only on the matrix, the chunk index and the chunk layout.  For a
// constant-coefficient matrix they are the same for every interior line, so
// k_cc_setup computes them once (per plan) with the very same recurrences,
// and k_cc_sweep only runs the d recurrences: no reciprocals, and one value
// (d) per reduced-system exchange instead of three (a, c, d).  The reference
// keeps precomputed factors for repeated matrices in the same spirit
// (trid_warp_setup / trid_warp_solve, scalar/src/cuda/trid_thomaspcr_small.h:
// 508-605); this is an independent design for the chunked wave64 scheme.
//
// Layout of k_cc_sweep: systems in consecutive lanes (coalesced rows, as
// ORDER 1 of trid_part.hpp) with exactly 64 systems per work-group, so wave w
// holds chunk w of 64 systems and the chunk's factors are wave-uniform
// (loaded once per wave through the scalar cache).  Lines that are not
// interior to the grid are identity systems: their lanes store d unchanged.
#pragma once

#include "trid_part.hpp"

namespace tridhip {

// Largest work-group of k_cc_sweep for chunk length R and NS systems per
// lane (its launch bound): 64 chunk rows per lane (R = 32 with two systems)
// need more than the 128 VGPRs of a 1024-thread work-group.
__host__ __device__ constexpr int cc_max_threads(int R, int NS) {
  return R * NS >= 64 ? 512 : 1024;
}

// Per-chunk factor record (T values) for chunk length R.
template <int R>
struct CCLayout {
  static_assert(R >= 4, "factor layout needs R >= 4");
  static constexpr int R1 = 0;              // 1/b_1
  static constexpr int A = 1;               // a_i, i = 2..R-1
  static constexpr int RI = A + (R - 2);    // 1/(b_i - a_i c'_{i-1}), i = 2..R-1
  static constexpr int CB = RI + (R - 2);   // c'_i of the backward pass, i = 1..R-3
  static constexpr int C0 = CB + (R - 3);   // c_0
  static constexpr int R0 = C0 + 1;         // 1/(first row's diagonal)
  static constexpr int EA = R0 + 1;         // expansion a_i, i = 1..R-2
  static constexpr int EC = EA + (R - 2);   // expansion c_i, i = 1..R-2
  static constexpr int FA = EC + (R - 2);   // first row a, c
  static constexpr int FC = FA + 1;
  static constexpr int LA = FC + 1;         // last row a, c and 1/diagonal
  static constexpr int LC = LA + 1;         //   after eliminating the first
  static constexpr int INV = LC + 1;        //   unknown of the next chunk
  static constexpr int PCR = INV + 1;       // per PCR level: r.a, r.c, 1/bn
  __host__ __device__ static constexpr int nf(int levels) { return PCR + 3 * levels; }
};

// Factors of the interior-line matrix of a constant-coefficient batch (rows
// interior to the grid when 1 <= row <= N-2: a, b, c = cf; identity rows
// otherwise), for chunk length R and cps = 2^cps_log2 chunks.  One
// wave (cps <= 64); thread k writes chunk k's record at F + k * nf.  The recurrences are those of chunk_reduce, eliminate_first
// and pcr_update, applied to the coefficients only.
template <typename T, int R>
__global__ void __launch_bounds__(kWave)
    k_cc_setup(CoefConst<T> cf, int N, int cps_log2, T *__restrict__ F) {
  using Ly = CCLayout<R>;
  __shared__ T sa[kWave], sc[kWave];
  const int cps = 1 << cps_log2;
  const int nf = Ly::nf(cps_log2);
  const int k = threadIdx.x;
  const bool on = k < cps;
  T *out = F + (long long)(on ? k : 0) * nf;
  auto W = [&](int i, T v) {
    if (on) out[i] = v;
  };
  T a[R], b[R], c[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int row = k * R + r;
    const bool in = row >= 1 && row + 2 <= N;
    a[r] = in ? cf.a : T(0);
    b[r] = in ? cf.b : T(1);
    c[r] = in ? cf.c : T(0);
  }
  T ap, cp;
  {
    const T r = rcp(b[1]);
    ap = a[1] * r;
    cp = c[1] * r;
    a[1] = ap;
    c[1] = cp;
    W(Ly::R1, r);
  }
#pragma unroll
  for (int i = 2; i < R; i++) {
    const T r = rcp(b[i] - a[i] * cp);
    W(Ly::A + i - 2, a[i]);
    W(Ly::RI + i - 2, r);
    ap = -a[i] * ap * r;
    cp = c[i] * r;
    a[i] = ap;
    c[i] = cp;
  }
  const T la = a[R - 1], lc = c[R - 1];
#pragma unroll
  for (int i = R - 3; i >= 1; i--) {
    W(Ly::CB + i - 1, c[i]);
    a[i] = a[i] - c[i] * a[i + 1];
    c[i] = -c[i] * c[i + 1];
  }
  const T bf = b[0] - c[0] * a[1];
  const T cff = -c[0] * c[1];
  W(Ly::C0, c[0]);
  const T r0 = rcp(bf);
  W(Ly::R0, r0);
  const T fa = a[0] * r0, fc = cff * r0;
#pragma unroll
  for (int i = 1; i < R - 1; i++) {
    W(Ly::EA + i - 1, a[i]);
    W(Ly::EC + i - 1, c[i]);
  }
  W(Ly::FA, fa);
  W(Ly::FC, fc);
  // reduced system: eliminate the next chunk's first unknown
  if (on) {
    sa[k] = fa;
    sc[k] = fc;
  }
  __syncthreads();
  const T fna = k + 1 < cps ? sa[k + 1] : T(0);
  const T fnc = k + 1 < cps ? sc[k + 1] : T(0);
  T inv = rcp(T(1) - la * fc - lc * fna);
  W(Ly::LA, la);
  W(Ly::LC, lc);
  W(Ly::INV, inv);
  T ra = -la * fa * inv, rc = -lc * fnc * inv;
  int j = 0;
  for (int s = 1; s < cps; s <<= 1, j++) {
    __syncthreads();
    if (on) {
      sa[k] = ra;
      sc[k] = rc;
    }
    __syncthreads();
    const T ma = k >= s && on ? sa[k - s] : T(0), mc = k >= s && on ? sc[k - s] : T(0);
    const T pa = k + s < cps ? sa[k + s] : T(0), pc = k + s < cps ? sc[k + s] : T(0);
    inv = rcp(T(1) - ra * mc - rc * pa);
    W(Ly::PCR + 3 * j, ra);
    W(Ly::PCR + 3 * j + 1, rc);
    W(Ly::PCR + 3 * j + 2, inv);
    ra = -ra * ma * inv;
    rc = -rc * pc * inv;
  }
}

// The solve with precomputed factors F (k_cc_setup of the same N, R, cps).
// blockDim = 64 * cps <= cc_max_threads(R, NS); wave w = chunk w; lane l of work-group b takes NS
// adjacent systems q = NS * (64 * b + l) + s, s < NS.  NS = 2 moves both
// systems' rows as one 16-byte access (the host requires adjacent systems to
// be adjacent elements of d and u, 16-byte aligned, and an even system
// count), so a wave's row access covers 128 systems without gaps.
// Non-Inc: d <- A^{-1} d; Inc: u <- u + A^{-1} d.  CP: cache policy of the
// streaming accesses (kStream / kCached).
template <typename T, int INC, int R, typename I, int NS, int CP>
__global__ void __launch_bounds__(cc_max_threads(R, NS))
    k_cc_sweep(const T *__restrict__ F, T *__restrict__ Dd, T *__restrict__ U, SysDesc D,
               int cps_log2, int xcd) {
  using Ly = CCLayout<R>;
  typedef T V2 __attribute__((ext_vector_type(2)));
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sm = reinterpret_cast<T *>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int cps = 1 << cps_log2;
  const int k = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform chunk index
  const T *__restrict__ Fk = F + k * Ly::nf(cps_log2);
  // xcd: every XCD takes one contiguous range of systems (as k_part)
  long long block = blockIdx.x;
  if (xcd) {
    const unsigned nb8 = gridDim.x & ~7u;
    if (blockIdx.x < nb8) block = (long long)(blockIdx.x & 7u) * (nb8 >> 3) + (blockIdx.x >> 3);
  }
  const long long q0 = (block * kWave + lane) * NS;
  const bool active = q0 < D.nsys;
  const int N = D.N;
  I off[5];
  bool line_in[NS];
  if (active) {
    sys_offsets<I>(D, q0, off, &line_in[0]);
    if (NS == 2) {
      I off1[5];
      sys_offsets<I>(D, q0 + 1, off1, &line_in[NS - 1]);
    }
  } else {
#pragma unroll
    for (int t = 0; t < 5; t++) off[t] = 0;
#pragma unroll
    for (int s = 0; s < NS; s++) line_in[s] = false;
  }
  const int row0 = k * R;
  const I s3 = (I)D.sstr[3], s4 = (I)D.sstr[4];
  // element r of system s: d[s][r]
  T d[NS][R];
  {
    I od = off[3] + (I)row0 * s3;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const bool ok = active && row0 + r < N;
      if (NS == 2) {
        const V2 v = ok ? ld_stream<CP>(reinterpret_cast<const V2 *>(at(Dd, od))) : V2{};
        d[0][r] = v[0];
        d[NS - 1][r] = v[1];
      } else {
        d[0][r] = ok ? ld_stream<CP>(at(Dd, od)) : T(0);
      }
      od += s3;
    }
  }
  T ld[NS], fd[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    // chunk sweep (d only)
    T dp = d[s][1] * Fk[Ly::R1];
    d[s][1] = dp;
#pragma unroll
    for (int i = 2; i < R; i++) {
      dp = (d[s][i] - Fk[Ly::A + i - 2] * dp) * Fk[Ly::RI + i - 2];
      d[s][i] = dp;
    }
    ld[s] = d[s][R - 1];
#pragma unroll
    for (int i = R - 3; i >= 1; i--) d[s][i] = d[s][i] - Fk[Ly::CB + i - 1] * d[s][i + 1];
    fd[s] = (d[s][0] - Fk[Ly::C0] * d[s][1]) * Fk[Ly::R0];
  }
  // reduced systems, d only, through LDS (ping-pong buffers of NS * 64 * cps)
  const int n = NS * cps * kWave;
  const int me = (k * kWave + lane) * NS;
  int par = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) sm[me + s] = fd[s];
  __syncthreads();
  T rd[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const T fnd = k + 1 < cps ? sm[me + NS * kWave + s] : T(0);
    rd[s] = (ld[s] - Fk[Ly::LA] * fd[s] - Fk[Ly::LC] * fnd) * Fk[Ly::INV];
  }
  par = 1;
  int j = 0;
  for (int st = 1; st < cps; st <<= 1, j++) {
    T *b = sm + par * n;
#pragma unroll
    for (int s = 0; s < NS; s++) b[me + s] = rd[s];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const T md = k >= st ? b[me - st * NS * kWave + s] : T(0);
      const T pd = k + st < cps ? b[me + st * NS * kWave + s] : T(0);
      rd[s] = (rd[s] - Fk[Ly::PCR + 3 * j] * md - Fk[Ly::PCR + 3 * j + 1] * pd) *
              Fk[Ly::PCR + 3 * j + 2];
    }
    par ^= 1;
  }
  T xf[NS];
  {
    T *b = sm + par * n;
#pragma unroll
    for (int s = 0; s < NS; s++) b[me + s] = rd[s];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const T xm = k > 0 ? b[me - NS * kWave + s] : T(0);
      xf[s] = fd[s] - Fk[Ly::FA] * xm - Fk[Ly::FC] * rd[s];
    }
  }
  T y[NS][R];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    y[s][0] = xf[s];
    y[s][R - 1] = rd[s];
#pragma unroll
    for (int i = 1; i < R - 1; i++)
      y[s][i] = d[s][i] - Fk[Ly::EA + i - 1] * xf[s] - Fk[Ly::EC + i - 1] * rd[s];
  }
  if (!active) return;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    if (!line_in[s]) {
      // identity system: the solution is d itself (not yet overwritten)
      I od = off[3] + (I)row0 * s3 + (I)s;
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (row0 + r < N) y[s][r] = *at(Dd, od);
        od += s3;
      }
    }
  }
  if (INC) {
    I ou = off[4] + (I)row0 * s4;
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (row0 + r < N) {
        if (NS == 2) {
          V2 *pu = reinterpret_cast<V2 *>(at(U, ou));
          V2 v = ld_stream<CP>(pu);
          v[0] += y[0][r];
          v[1] += y[NS - 1][r];
          st_stream<CP>(pu, v);
        } else {
          T *pu = at(U, ou);
          st_stream<CP>(pu, ld_stream<CP>(pu) + y[0][r]);
        }
      }
      ou += s4;
    }
  } else {
    I od = off[3] + (I)row0 * s3;
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (row0 + r < N) {
        if (NS == 2) {
          V2 v;
          v[0] = y[0][r];
          v[1] = y[NS - 1][r];
          st_stream<CP>(reinterpret_cast<V2 *>(at(Dd, od)), v);
        } else {
          st_stream<CP>(at(Dd, od), y[0][r]);
        }
      }
      od += s3;
    }
  }
}

}  // namespace tridhip
