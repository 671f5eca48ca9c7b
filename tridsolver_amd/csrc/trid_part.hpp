// trid_part.hpp -- the partitioned ("register-blocked") Thomas solver.
//
// Every system is cut into chunks of R consecutive rows; one lane owns one
// chunk of one system and keeps all of it in registers.  Each lane runs a
// modified Thomas sweep over its chunk that expresses every row through the
// chunk's first and last unknowns, which leaves two rows per chunk coupled
// to the neighbouring chunks.  Eliminating each chunk's first unknown from
// those rows leaves one row per chunk (in the chunks' last unknowns), again
// tridiagonal; it is solved by parallel cyclic reduction (PCR) across lanes
// -- with ds_bpermute when all chunks of a system are in one wave, through
// LDS otherwise.  Each lane then recovers its first unknown and its interior
// rows.  Global memory is read once (a, b, c, d) and written once (d, or u
// for the Inc form): no c'/d' scratch traffic.
//
// The reference solves the same systems with the plain Thomas recurrence
// (scalar/src/cuda/trid_strided_multidim.hpp:43-87, c2/d2 in local memory)
// or with a warp-32 Thomas-PCR hybrid (scalar/src/cuda/trid_thomaspcr.hpp);
// this file is an independent wave64 design of the latter idea.
#pragma once

#include "trid_device.hpp"

namespace tridhip {

// Modified Thomas over one chunk.  In: a, b, c, d of R rows (rows past the
// system end are identity rows, a of the system's first row and c of its
// last row are zero).  Out: for i = 1..R-2 the row relations
//   y_i = d[i] - a[i] * y_0 - c[i] * y_{R-1}
// left in a, c, d, and the two reduced rows (normalised, diagonal 1)
//   first: first.a * y_{-1}   + y_0     + first.c * y_{R-1} = first.d
//   last:  last.a  * y_0      + y_{R-1} + last.c  * y_R     = last.d
// where y_{-1} is the previous chunk's last unknown and y_R the next chunk's
// first unknown.
template <typename T, int R>
__device__ __forceinline__ void chunk_reduce(T (&a)[R], const T (&b)[R],
                                             T (&c)[R], T (&d)[R],
                                             Row<T> &first, Row<T> &last) {
  static_assert(R >= 2, "chunks need at least two rows");
  T ap, cp, dp;
  {
    T r = rcp(b[1]);
    ap = a[1] * r;
    cp = c[1] * r;
    dp = d[1] * r;
    a[1] = ap;
    c[1] = cp;
    d[1] = dp;
  }
#pragma unroll
  for (int i = 2; i < R; i++) {
    T r = rcp(b[i] - a[i] * cp);
    ap = -a[i] * ap * r;
    dp = (d[i] - a[i] * dp) * r;
    cp = c[i] * r;
    a[i] = ap;
    c[i] = cp;
    d[i] = dp;
  }
  last.a = a[R - 1];
  last.c = c[R - 1];
  last.d = d[R - 1];
#pragma unroll
  for (int i = R - 3; i >= 1; i--) {
    a[i] = a[i] - c[i] * a[i + 1];
    d[i] = d[i] - c[i] * d[i + 1];
    c[i] = -c[i] * c[i + 1];
  }
  T bf, cf, df;
  if (R >= 3) {
    bf = b[0] - c[0] * a[1];
    cf = -c[0] * c[1];
    df = d[0] - c[0] * d[1];
  } else {
    bf = b[0];
    cf = c[0];
    df = d[0];
  }
  T r = rcp(bf);
  first.a = a[0] * r;
  first.c = cf * r;
  first.d = df * r;
}

template <typename T, int R>
__device__ __forceinline__ void chunk_expand(const T (&a)[R], const T (&c)[R],
                                             const T (&d)[R], T y0, T yl,
                                             T (&y)[R]) {
  y[0] = y0;
  y[R - 1] = yl;
#pragma unroll
  for (int i = 1; i < R - 1; i++) y[i] = d[i] - a[i] * y0 - c[i] * yl;
}

// Row of chunk k in the last-unknown system, given the chunk's own two
// reduced rows f, l and the next chunk's first row fn (zero row if none):
// substituting f_k = f.d - f.a l_{k-1} - f.c l_k and the same for f_{k+1}
// into l's row gives  alpha l_{k-1} + l_k + gamma l_{k+1} = delta.
template <typename T>
__device__ __forceinline__ Row<T> eliminate_first(const Row<T> &f,
                                                  const Row<T> &l,
                                                  const Row<T> &fn) {
  T bn = T(1) - l.a * f.c - l.c * fn.a;
  T inv = rcp(bn);
  Row<T> r;
  r.a = -l.a * f.a * inv;
  r.c = -l.c * fn.c * inv;
  r.d = (l.d - l.a * f.d - l.c * fn.d) * inv;
  return r;
}

// Merges the reduced rows of two adjacent blocks of rows, A = (f, l) and the
// block B = (fb, lb) just after it, into the reduced rows of A followed by B:
// A's last and B's first unknown are eliminated (a 2x2 solve of l's and fb's
// rows), which leaves
//   f: f.a * y_{-1} + y_fA + f.c * y_lB = f.d
//   l: l.a * y_fA + y_lB + l.c * y_{+1} = l.d
// with y_{-1} / y_{+1} the unknowns just outside A / B.  The operation is
// associative, so a segment's rows can be merged in a tree.
template <typename T>
__device__ __forceinline__ void seg_merge(Row<T> &f, Row<T> &l, const Row<T> &fb,
                                          const Row<T> &lb) {
  const T idet = rcp(T(1) - l.c * fb.a);
  const T g = rcp(T(1) - f.c * l.a * idet);
  const T h = rcp(T(1) - lb.a * fb.c * idet);
  Row<T> F, L;
  F.a = f.a * g;
  F.c = f.c * l.c * fb.c * idet * g;
  F.d = (f.d - f.c * (l.d - l.c * fb.d) * idet) * g;
  L.a = lb.a * fb.a * l.a * idet * h;
  L.c = lb.c * h;
  L.d = (lb.d - lb.a * (fb.d - fb.a * l.d) * idet) * h;
  f = F;
  l = L;
}

// Output of the segment-reduction mode of part_block (split-line solves of
// systems longer than one work-group holds).  The work-group's systems are
// segments of L rows of longer systems; each segment's two merged reduced
// rows go to rows 2p and 2p+1 of its system's reduced system (b = 1), at
// element qs * rs_sys + row * rs_row of ra, rb, rc, rd.
//   segment item q -> (system qs, segment p): p_tail >= 0: the launch holds
//   only the last (partial) segment p_tail of every system, qs = q;
//   otherwise ORDER 0: p = q % Pf, qs = q / Pf; ORDER 1: qs = q % nsys,
//   p = q / nsys.
template <typename T>
struct SegOut {
  T *ra = nullptr, *rb = nullptr, *rc = nullptr, *rd = nullptr;
  long long rs_row = 0, rs_sys = 0;
  long long nsys = 0;
  int Pf = 1, P = 1, p_tail = -1;
};

// (system, segment) of segment item q of a SegOut launch.
template <typename T, int ORDER>
__device__ __forceinline__ void seg_item(const SegOut<T> &so, long long q, long long &qs,
                                         int &p) {
  if (so.p_tail >= 0) {
    qs = q;
    p = so.p_tail;
  } else if (ORDER == 0) {
    p = (int)(q % so.Pf);
    qs = q / so.Pf;
  } else {
    qs = q % so.nsys;
    p = (int)(q / so.nsys);
  }
}

template <typename T>
__device__ __forceinline__ Row<T> row_from_lane(const Row<T> &v, int src,
                                                bool valid) {
  int s = valid ? src : 0;
  Row<T> o;
  o.a = lane_read(v.a, s);
  o.c = lane_read(v.c, s);
  o.d = lane_read(v.d, s);
  if (!valid) {
    o.a = T(0);
    o.c = T(0);
    o.d = T(0);
  }
  return o;
}

// Reduced solve for a system whose cps chunks sit in consecutive lanes of
// one wave (cps a power of two <= 64).  k: the lane's chunk index.
// Returns the chunk's first (xf) and last (xl) unknowns.
template <typename T>
__device__ __forceinline__ void reduced_wave(const Row<T> &f, const Row<T> &l,
                                             int k, int cps, int lane, T &xf,
                                             T &xl) {
  Row<T> fn = row_from_lane(f, lane + 1, k + 1 < cps);
  Row<T> r = eliminate_first(f, l, fn);
  for (int s = 1; s < cps; s <<= 1) {
    Row<T> m = row_from_lane(r, lane - s, k >= s);
    Row<T> p = row_from_lane(r, lane + s, k + s < cps);
    r = pcr_update(r, m, p);
  }
  xl = r.d;
  T xm = lane_read(xl, k > 0 ? lane - 1 : lane);
  if (k == 0) xm = T(0);
  xf = f.d - f.a * xm - f.c * xl;
}

// Same through LDS for chunks spread over several waves.  Entry (chunk k,
// system slot s) lives at k * spb + s of each of two ping-pong buffers of
// 3 * cps * spb values (one barrier per exchange).  All threads call it.
template <typename T>
__device__ __forceinline__ void reduced_lds(const Row<T> &f, const Row<T> &l,
                                            int k, int cps, int slot, int spb,
                                            T *sm, T &xf, T &xl) {
  const int n = cps * spb;
  const int me = k * spb + slot;
  int par = 0;
  auto publish = [&](const Row<T> &v) {
    T *b = sm + par * 3 * n;
    b[me] = v.a;
    b[n + me] = v.c;
    b[2 * n + me] = v.d;
  };
  auto fetch = [&](int kk, bool valid) {
    Row<T> o{T(0), T(0), T(0)};
    if (valid) {
      const T *b = sm + par * 3 * n;
      int i = kk * spb + slot;
      o.a = b[i];
      o.c = b[n + i];
      o.d = b[2 * n + i];
    }
    return o;
  };
  publish(f);
  __syncthreads();
  Row<T> fn = fetch(k + 1, k + 1 < cps);
  par ^= 1;
  Row<T> r = eliminate_first(f, l, fn);
  for (int s = 1; s < cps; s <<= 1) {
    publish(r);
    __syncthreads();
    Row<T> m = fetch(k - s, k >= s);
    Row<T> p = fetch(k + s, k + s < cps);
    par ^= 1;
    r = pcr_update(r, m, p);
  }
  xl = r.d;
  {
    T *b = sm + par * 3 * n;
    b[me] = xl;
    __syncthreads();
    T xm = k > 0 ? b[me - spb] : T(0);
    xf = f.d - f.a * xm - f.c * xl;
  }
}

// Same, for chunks spread over several waves with cps <= 64: the chunks'
// two reduced rows go to LDS, every wave then takes whole systems (64 / cps
// systems, chunks in consecutive lanes) and solves them with the in-wave
// PCR of reduced_wave (no barrier per level), and the chunks' unknowns go
// back through LDS -- three barriers in all instead of one per PCR level.
// Entry (chunk k, system slot s) lives at k * (spb + 1) + s of six arrays of
// cps * (spb + 1) values (the padding keeps the wave-local reads, which walk
// k with s fixed, free of bank conflicts).  All threads call it; blockDim =
// cps * spb.
template <typename T>
__device__ __forceinline__ void reduced_lds_w(const Row<T> &f, const Row<T> &l, int k, int cps,
                                              int slot, int spb, T *sm, T &xf, T &xl) {
  const int sp = spb + 1;
  const int n = cps * sp;
  const int me = k * sp + slot;
  sm[me] = f.a;
  sm[n + me] = f.c;
  sm[2 * n + me] = f.d;
  sm[3 * n + me] = l.a;
  sm[4 * n + me] = l.c;
  sm[5 * n + me] = l.d;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int cps_log2 = __builtin_ctz(cps);
  const int k2 = lane & (cps - 1);
  const int s2 = (int)(threadIdx.x >> 6) * (kWave >> cps_log2) + (lane >> cps_log2);
  const int e = k2 * sp + s2;
  const Row<T> f2{sm[e], sm[n + e], sm[2 * n + e]};
  const Row<T> l2{sm[3 * n + e], sm[4 * n + e], sm[5 * n + e]};
  T xf2, xl2;
  reduced_wave<T>(f2, l2, k2, cps, lane, xf2, xl2);
  __syncthreads();  // every entry has been read
  sm[e] = xf2;
  sm[n + e] = xl2;
  __syncthreads();
  xf = sm[me];
  xl = sm[n + me];
}

// A row of the last-unknown system with three right-hand sides: d, and the
// unit couplings p / q to the unknowns just before and just after a wave.
template <typename T>
struct Row3 {
  T a, c, d, p, q;
};

template <typename T>
__device__ __forceinline__ Row3<T> row3_from_lane(const Row3<T> &v, int src, bool valid) {
  const int s = valid ? src : 0;
  Row3<T> o;
  o.a = lane_read(v.a, s);
  o.c = lane_read(v.c, s);
  o.d = lane_read(v.d, s);
  o.p = lane_read(v.p, s);
  o.q = lane_read(v.q, s);
  if (!valid) o = Row3<T>{T(0), T(0), T(0), T(0), T(0)};
  return o;
}

template <typename T>
__device__ __forceinline__ Row3<T> pcr_update3(const Row3<T> &r, const Row3<T> &m,
                                               const Row3<T> &n) {
  const T inv = rcp(T(1) - r.a * m.c - r.c * n.a);
  Row3<T> o;
  o.a = -r.a * m.a * inv;
  o.c = -r.c * n.c * inv;
  o.d = (r.d - r.a * m.d - r.c * n.d) * inv;
  o.p = (r.p - r.a * m.p - r.c * n.p) * inv;
  o.q = (r.q - r.a * m.q - r.c * n.q) * inv;
  return o;
}

// Reduced solve for systems whose cps > 64 chunks sit in consecutive lanes
// of W = cps / 64 whole waves (ORDER 0 and the ragged kernel).  The chunks'
// last unknowns l_k form one tridiagonal system (eliminate_first; the last
// chunk of a wave takes the next wave's first chunk row through LDS), of
// which every wave holds 64 consecutive rows.  Three barriers:
//   1. every wave solves its 64 rows by in-wave PCR with three right-hand
//      sides, so that l_k = D_k - P_k * X_p - Q_k * X_n in terms of the
//      unknowns just outside the wave (X_p = l of the previous wave's last
//      chunk, X_n = l of the next wave's first chunk);
//   2. one lane per system solves for every wave's first and last l: a
//      block-tridiagonal system of 2x2 blocks whose forward sweep is scalar
//      (the coupling blocks have one nonzero column);
//   3. every lane evaluates its l_k and the chunk's first unknown.
// LDS: 12 W values per system slot at sm + slot * 12W.  All threads call it.
template <typename T>
__device__ __forceinline__ void reduced_multi(const Row<T> &f, const Row<T> &l, int k, int cps,
                                              int slot, T *sm, T &xf, T &xl) {
  const int lane = threadIdx.x & (kWave - 1);
  const int W = cps >> 6;
  const int w = k >> 6;
  T *fs = sm + slot * (12 * W);  // [W][3]: first row of each wave's chunk 0
  T *bs = fs + 3 * W;            // [W][6]: D, P, Q of rows 0 and 63; then G, H
  T *zs = bs + 6 * W;            // [W][2]: first and last l of every wave
  if (lane == 0) {
    fs[3 * w] = f.a;
    fs[3 * w + 1] = f.c;
    fs[3 * w + 2] = f.d;
  }
  __syncthreads();
  Row<T> fn = row_from_lane(f, lane + 1, lane + 1 < kWave);
  if (lane == kWave - 1 && w + 1 < W) {
    fn.a = fs[3 * w + 3];
    fn.c = fs[3 * w + 4];
    fn.d = fs[3 * w + 5];
  }
  const Row<T> r = eliminate_first(f, l, fn);
  Row3<T> q;
  q.a = lane == 0 ? T(0) : r.a;
  q.c = lane == kWave - 1 ? T(0) : r.c;
  q.d = r.d;
  q.p = lane == 0 ? r.a : T(0);
  q.q = lane == kWave - 1 ? r.c : T(0);
  for (int s = 1; s < kWave; s <<= 1) {
    const Row3<T> m = row3_from_lane(q, lane - s, lane >= s);
    const Row3<T> n = row3_from_lane(q, lane + s, lane + s < kWave);
    q = pcr_update3(q, m, n);
  }
  if (lane == 0) {
    bs[6 * w] = q.d;
    bs[6 * w + 1] = q.p;
    bs[6 * w + 2] = q.q;
  }
  if (lane == kWave - 1) {
    bs[6 * w + 3] = q.d;
    bs[6 * w + 4] = q.p;
    bs[6 * w + 5] = q.q;
  }
  __syncthreads();
  if (w == 0 && lane == 0) {
    // z_v = (first_v, last_v):  z_v + B_v z_{v-1} + C_v z_{v+1} = (D0, D63)
    // with B_v = [[0, P0], [0, P63]], C_v = [[Q0, 0], [Q63, 0]];
    // z_v = H_v - G_v z_{v+1}, G_v = [[G0, 0], [G1, 0]]
    T g1 = T(0), h1 = T(0);
    for (int v = 0; v < W; v++) {
      T *e = bs + 6 * v;
      const T D0 = e[0], P0 = e[1], Q0 = e[2], D63 = e[3], P63 = e[4], Q63 = e[5];
      const T inv = rcp(T(1) - P0 * g1);
      const T H0 = (D0 - P0 * h1) * inv;
      const T H1 = D63 - P63 * h1 + P63 * g1 * H0;
      const T G0 = Q0 * inv;
      const T G1 = Q63 + P63 * g1 * G0;
      e[0] = G0;
      e[1] = G1;
      e[2] = H0;
      e[3] = H1;
      g1 = G1;
      h1 = H1;
    }
    T xn = T(0);  // first l of the next wave
    for (int v = W - 1; v >= 0; v--) {
      const T *e = bs + 6 * v;
      const T first = e[2] - e[0] * xn;
      zs[2 * v] = first;
      zs[2 * v + 1] = e[3] - e[1] * xn;
      xn = first;
    }
  }
  __syncthreads();
  const T xp = w > 0 ? zs[2 * w - 1] : T(0);
  const T xnx = w + 1 < W ? zs[2 * w + 2] : T(0);
  xl = q.d - q.p * xp - q.q * xnx;
  T xm = lane_read(xl, lane > 0 ? lane - 1 : lane);
  if (lane == 0) xm = xp;
  xf = f.d - f.a * xm - f.c * xl;
}

// Cache policy of the streaming global accesses (every element of a, b, c,
// d and u is touched once per solve):
//   kStream: non-temporal loads and stores, which do not allocate in L2 or
//            in MI355X's 256 MB last-level cache (MALL).  Best when the
//            arrays exceed that cache: a read-4/write-1 solve over 640 MiB
//            runs at about 77 % of peak HBM bandwidth, against 68 % with
//            ordinary accesses.
//   kCached: ordinary accesses.  When the arrays of a solve, or of a chain
//            of solves such as an ADI step (u and du), fit in the MALL, the
//            next kernel (or call) reads what the previous one wrote from
//            there: a copy -> in-place -> add chain over 2 x 128 MiB runs at
//            93-97 % of 8 TB/s cached and at 64-77 % with non-temporal
//            accesses.
//   kMixed:  (coefficient arrays only) non-temporal a, b, c and ordinary d
//            and u: the read-only coefficients do not evict the right-hand
//            side, which a caller's previous kernel wrote and its next one
//            reads, from the last-level cache.
enum { kStream = 0, kCached = 1, kMixed = 2 };

template <int CP = kStream, typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (CP == kCached) return *p;
  else return __builtin_nontemporal_load(p);
}
template <int CP = kStream, typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
  if constexpr (CP == kCached) *p = v;
  else __builtin_nontemporal_store(v, p);
}

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};

// Global access at an element offset held in index type I.  With I =
// uint32_t the byte offset is formed in 32 bits (the host guarantees every
// array fits in 4 GiB), which lets the compiler use the scalar-base +
// 32-bit vector-offset addressing form.
template <typename T, typename I>
__device__ __forceinline__ const T *at(const T *base, I off) {
  if constexpr (sizeof(I) == 4) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                       (uint32_t)(off * (I)sizeof(T)));
  } else {
    return base + off;
  }
}
template <typename T, typename I>
__device__ __forceinline__ T *at(T *base, I off) {
  if constexpr (sizeof(I) == 4) {
    return reinterpret_cast<T *>(reinterpret_cast<char *>(base) +
                                 (uint32_t)(off * (I)sizeof(T)));
  } else {
    return base + off;
  }
}

template <typename T, int R, int CP = kStream>
__device__ __forceinline__ void load_chunk_vec(const T *__restrict__ p,
                                               T (&x)[R]) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  static_assert(R % E == 0, "chunk must be a whole number of 16-byte vectors");
  const V *pv = reinterpret_cast<const V *>(p);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V t = ld_stream<CP>(pv + v);
#pragma unroll
    for (int e = 0; e < E; e++) x[E * v + e] = t[e];
  }
}

template <typename T, int R, int CP = kStream>
__device__ __forceinline__ void store_chunk_vec(T *__restrict__ p,
                                                const T (&x)[R]) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  V *pv = reinterpret_cast<V *>(p);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V t;
#pragma unroll
    for (int e = 0; e < E; e++) t[e] = x[E * v + e];
    st_stream<CP>(pv + v, t);
  }
}

// Wave-staged access for contiguous systems (ORDER 0): the wave's 64*R
// elements of one array form one contiguous span.  Global memory is moved
// in 16-byte pieces with consecutive lanes on consecutive pieces (each
// instruction covers 1 KiB without gaps), and the lane <-> chunk transpose
// goes through a wave-private LDS buffer of 64*R elements.  LDS operations
// of one wave execute in order, so no barrier is needed; wave_barrier only
// keeps the compiler from reordering the LDS accesses.
template <typename T, int R, int CP = kStream>
__device__ __forceinline__ void load_chunk_staged(const T *__restrict__ span,
                                                  T *__restrict__ buf,
                                                  int lane, T (&x)[R]) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  const V *g = reinterpret_cast<const V *>(span);
  V *s = reinterpret_cast<V *>(buf);
  V t[R / E];
#pragma unroll
  for (int v = 0; v < R / E; v++) t[v] = ld_stream<CP>(g + v * kWave + lane);
#pragma unroll
  for (int v = 0; v < R / E; v++) s[v * kWave + lane] = t[v];
  __builtin_amdgcn_wave_barrier();
  const V *mine = reinterpret_cast<const V *>(buf + lane * R);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V w = mine[v];
#pragma unroll
    for (int e = 0; e < E; e++) x[E * v + e] = w[e];
  }
  __builtin_amdgcn_wave_barrier();
}

// add: the span receives its own previous contents (or, with src, those of
// src) plus x.
template <typename T, int R, int CP = kStream>
__device__ __forceinline__ void store_chunk_staged(T *__restrict__ span,
                                                   T *__restrict__ buf,
                                                   int lane, const T (&x)[R],
                                                   bool add,
                                                   const T *src = nullptr) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  V *mine = reinterpret_cast<V *>(buf + lane * R);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V w;
#pragma unroll
    for (int e = 0; e < E; e++) w[e] = x[E * v + e];
    mine[v] = w;
  }
  __builtin_amdgcn_wave_barrier();
  V *g = reinterpret_cast<V *>(span);
  const V *gs = src ? reinterpret_cast<const V *>(src) : g;
  const V *s = reinterpret_cast<const V *>(buf);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V w = s[v * kWave + lane];
    if (add) w += ld_stream<CP>(gs + v * kWave + lane);
    st_stream<CP>(g + v * kWave + lane, w);
  }
  __builtin_amdgcn_wave_barrier();
}

// General wave staging (VEC 4): the wave's chunks cover the contiguous
// element span [start, start + len), len <= 64 R, of an array with a
// 16-byte aligned base; the span may start and end anywhere.  The 16-byte
// pieces that cover it (consecutive lanes on consecutive pieces) go to the
// wave's LDS buffer of 64 R + E elements, so that element start + j sits at
// buf[sh + j], sh = start mod E.  A piece is the aligned 16-byte block that
// holds a span element, so it never reaches past the page of that element.
// Returns sh.
template <typename T, int R, typename I, int CP = kStream>
__device__ __forceinline__ int gspan_load(const T *__restrict__ base, I start, int len,
                                          T *__restrict__ buf, int lane) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  const int sh = (int)(start & (I)(E - 1));
  const V *g = reinterpret_cast<const V *>(at(base, start - (I)sh));
  const int np = (sh + len + E - 1) / E;
  V *s = reinterpret_cast<V *>(buf);
  V t[R / E + 1];
  // every lane loads R/E + 1 pieces; those past the span repeat its last
  // piece (unconditional loads schedule far better than predicated ones)
#pragma unroll
  for (int v = 0; v <= R / E; v++) {
    const int pc = v * kWave + lane;
    t[v] = ld_stream<CP>(g + (pc < np ? pc : np - 1));
  }
#pragma unroll
  for (int v = 0; v <= R / E; v++) {
    const int pc = v * kWave + lane;
    if (pc < np) s[pc] = t[v];
  }
  __builtin_amdgcn_wave_barrier();
  return sh;
}

// The lane's `rows` valid elements from buf[i0 ...] (pad after them).
template <typename T, int R>
__device__ __forceinline__ void gspan_take(const T *__restrict__ buf, int i0, int rows,
                                           T (&x)[R], T pad) {
#pragma unroll
  for (int r = 0; r < R; r++) x[r] = r < rows ? buf[i0 + r] : pad;
  __builtin_amdgcn_wave_barrier();
}

// Writes the span [start, start + len) from buf (element start + j at
// buf[sh + j], sh = start mod E, filled by the lanes before a wave barrier):
// pieces wholly inside the span as 16-byte stores, the elements of the two
// boundary pieces one by one (the rest of those pieces belongs to other
// waves).  add: the span receives its previous contents (or, with src,
// those of src) plus buf.
template <typename T, int R, typename I, int CP = kStream>
__device__ __forceinline__ void gspan_store(T *__restrict__ base, I start, int len,
                                            const T *__restrict__ buf, int lane, bool add,
                                            const T *src = nullptr) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  const int sh = (int)(start & (I)(E - 1));
  V *g = reinterpret_cast<V *>(at(base, start - (I)sh));
  const T *srcb = src ? src : base;
  const V *gs = reinterpret_cast<const V *>(at(srcb, start - (I)sh));
  const int np = (sh + len + E - 1) / E;
  const V *s = reinterpret_cast<const V *>(buf);
#pragma unroll
  for (int v = 0; v <= R / E; v++) {
    const int pc = v * kWave + lane;
    if (pc < np) {
      const int lo = pc * E - sh;  // span index of the piece's first element
      V w = s[pc];
      if (lo >= 0 && lo + E <= len) {
        if (add) w += ld_stream<CP>(gs + pc);
        st_stream<CP>(g + pc, w);
      } else {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int j = lo + e;
          if (j >= 0 && j < len) {
            T val = w[e];
            if (add) val += ld_stream<CP>(at(srcb, start + (I)j));
            st_stream<CP>(at(base, start + (I)j), val);
          }
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// Wave-staged load of 64*R contiguous elements with ordinary (cached) loads:
// a stencil re-reads the span as a neighbour of other lines, so it stays
// cacheable.  Split in two so that a kernel can have the global loads of
// several spans in flight before it transposes them: span_issue moves the
// span into registers (16 bytes per lane per instruction, consecutive lanes
// on consecutive pieces), span_stage hands every lane its own R elements
// through the wave's LDS buffer.  left/right: the elements just before and
// after the lane's chunk within the span (0 outside it).
template <typename T, int R>
struct Span {
  typename Vec16<T>::type v[R / Vec16<T>::n];
};

template <typename T, int R>
__device__ __forceinline__ void span_issue(const T *__restrict__ span, int lane,
                                           Span<T, R> &s) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  const V *g = reinterpret_cast<const V *>(span);
#pragma unroll
  for (int v = 0; v < R / E; v++) s.v[v] = g[v * kWave + lane];
}

template <typename T, int R>
__device__ __forceinline__ void span_stage(const Span<T, R> &t, T *__restrict__ buf,
                                           int lane, T (&x)[R], T *left = nullptr,
                                           T *right = nullptr) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  V *s = reinterpret_cast<V *>(buf);
#pragma unroll
  for (int v = 0; v < R / E; v++) s[v * kWave + lane] = t.v[v];
  __builtin_amdgcn_wave_barrier();
  const V *mine = reinterpret_cast<const V *>(buf + lane * R);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V w = mine[v];
#pragma unroll
    for (int e = 0; e < E; e++) x[E * v + e] = w[e];
  }
  if (left) *left = lane > 0 ? buf[lane * R - 1] : T(0);
  if (right) *right = lane < kWave - 1 ? buf[lane * R + R] : T(0);
  __builtin_amdgcn_wave_barrier();
}

template <typename T, int R>
__device__ __forceinline__ void load_span_cached(const T *__restrict__ span,
                                                 T *__restrict__ buf, int lane,
                                                 T (&x)[R], T *left = nullptr,
                                                 T *right = nullptr) {
  Span<T, R> t;
  span_issue<T, R>(span, lane, t);
  span_stage<T, R>(t, buf, lane, x, left, right);
}

// Coefficient sources of part_block.
//   CoefArrays: a, b, c read from memory with the layout of SysDesc (arrays
//               0, 1, 2) -- the library's strided-batch entry points.
//   CoefConst:  constant a, b, c on rows interior to the grid (the system's
//               line is interior and 1 <= row <= N-2), identity rows
//               (a = c = 0, b = 1) elsewhere -- the ADI sweeps with
//               Dirichlet boundaries.
template <typename T>
struct CoefArrays {
  static constexpr bool kArrays = true;
  static constexpr bool kHook = false;
  const T *__restrict__ A;
  const T *__restrict__ B;
  const T *__restrict__ C;
};
template <typename T>
struct CoefConst {
  static constexpr bool kArrays = false;
  static constexpr bool kHook = false;
  T a, b, c;
};

// Coefficient policies with their own chunk loader (scalar loads, any
// ORDER, VEC 0 only).  load() fills a, b, c of rows row0..row0+R-1 of the
// system at element offset off (stride s) in arrays with the layout of
// SysDesc array 0; rows past N are identity rows.
//   CoefTransposed: the transposed system A^T of tridiag(A, B, C): sub-
//                   diagonal c_{i-1}, diagonal b_i, super-diagonal a_{i+1}.
//   CoefKappa:      the variable-diffusivity ADI operator along a line:
//                   face values kf_{i+1/2} = (K_i + K_{i+1})/2 and, at rows
//                   interior to the grid, a_i = -lam kf_{i-1/2},
//                   c_i = -lam kf_{i+1/2}, b_i = 1 - a_i - c_i (identity rows
//                   elsewhere); with transpose, its transposed operator.
template <typename T>
struct CoefTransposed {
  static constexpr bool kArrays = false;
  static constexpr bool kHook = true;
  const T *__restrict__ A;
  const T *__restrict__ B;
  const T *__restrict__ C;
  template <int R, typename I>
  __device__ __forceinline__ void load(int row0, int N, bool, I off, I s,
                                       T (&a)[R], T (&b)[R], T (&c)[R]) const {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = row0 + r;
      if (row < N) {
        a[r] = row >= 1 ? ld_stream(at(C, off + (I)(row - 1) * s)) : T(0);
        b[r] = ld_stream(at(B, off + (I)row * s));
        c[r] = row + 1 < N ? ld_stream(at(A, off + (I)(row + 1) * s)) : T(0);
      } else {
        a[r] = T(0);
        b[r] = T(1);
        c[r] = T(0);
      }
    }
  }
};

template <typename T>
struct CoefKappa {
  static constexpr bool kArrays = false;
  static constexpr bool kHook = true;
  const T *__restrict__ K;
  T lam;
  int transpose;
  template <int R, typename I>
  __device__ __forceinline__ void load(int row0, int N, bool line_in, I off, I s,
                                       T (&a)[R], T (&b)[R], T (&c)[R]) const {
    // kappa of rows row0-1 .. row0+R (cached loads: neighbours re-read them)
    T k[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++) {
      const int row = row0 - 1 + r;
      k[r] = (row >= 0 && row < N) ? *at(K, off + (I)row * s) : T(0);
    }
    auto inner = [&](int row) { return line_in && row >= 1 && row + 2 <= N; };
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = row0 + r;
      const T am = -lam * (T(0.5) * (k[r + 1] + k[r]));      // -lam kf_{row-1/2}
      const T cp = -lam * (T(0.5) * (k[r + 1] + k[r + 2]));  // -lam kf_{row+1/2}
      const bool in = inner(row);
      b[r] = in ? T(1) - am - cp : T(1);
      if (!transpose) {
        a[r] = in ? am : T(0);
        c[r] = in ? cp : T(0);
      } else {
        // (A^T)_{row,row-1} = c_{row-1} = -lam kf_{row-1/2} if row-1 is an
        // interior row; (A^T)_{row,row+1} = a_{row+1} = -lam kf_{row+1/2}
        a[r] = inner(row - 1) ? am : T(0);
        c[r] = inner(row + 1) ? cp : T(0);
        if (row >= N) b[r] = T(1);
      }
    }
  }
};

// Block-wide partitioned solve of a strided batch.
//   ORDER 0: thread t -> chunk t % cps, system slot t / cps (chunks of a
//            system in consecutive lanes: contiguous systems, solvedim 0).
//   ORDER 1: thread t -> system slot t % spb, chunk t / spb (systems in
//            consecutive lanes: coalesced across systems, solvedim >= 1).
//   VEC:     1: 16-byte vector loads/stores along the system (ORDER 0 with
//            unit stride and 16-byte aligned chunks only; checked on the
//            host).  2: wave-staged access (ORDER 0, cps <= 64, every
//            system exactly cps*R long and the systems of a wave adjacent in
//            memory; checked on the host).  3: the same for cps > 64 (each
//            wave a contiguous span of one system), with the wave-blocked
//            reduced solve (reduced_multi).
//   I:       index type of element offsets (uint32_t or long long).
// blockDim.x == cps * spb, both powers of two.  Non-Inc: d <- A^{-1} d;
// Inc: u <- u + A^{-1} d.
// MODE 1 (segment reduction, CoefArrays only): every system of D is a
// segment of a longer system; instead of solving it, the work-group merges
// each segment's chunk rows into the segment's two reduced rows -- with the
// segment's true couplings a (first row) and c (last row) to its neighbours
// -- and writes them as SegOut describes.  Nothing else is written.
// MODE 2 (segment solve, CoefArrays only): every system of D is a segment
// whose neighbours' boundary unknowns are solved (so.rd holds the reduced
// solution); the couplings to them move to the right-hand side
// (d_first -= a_first * last_{p-1}, d_last -= c_last * first_{p+1}) and the
// segment is solved as an independent system.
template <typename T, int INC, int ORDER, int R, int VEC, typename I,
          typename Coef, int MODE = 0, int CP = kStream>
__device__ __forceinline__ void part_block(const Coef &coef, T *__restrict__ Dd,
                                           T *__restrict__ U, const SysDesc &D,
                                           int cps_log2, int spb_log2,
                                           long long block,
                                           T *__restrict__ Out = nullptr,
                                           SegOut<T> so = SegOut<T>()) {
  static_assert(!Coef::kHook || VEC == 0, "hooked coefficients load scalars");
  static_assert(MODE == 0 || Coef::kArrays, "segment reduction reads coefficient arrays");
  static_assert(CP != kMixed || Coef::kArrays, "the mixed policy is for coefficient arrays");
  // cache policies of the coefficient arrays and of d / u
  constexpr int CPC = CP == kMixed ? kStream : CP;
  constexpr int CPD = CP == kMixed ? kCached : CP;
  constexpr bool STAGED = VEC == 2 || VEC == 3;
  constexpr bool GSTAGE = VEC == 4;
  // per-wave staging buffer (elements)
  constexpr int SW = kWave * R + (GSTAGE ? Vec16<T>::n : 0);
  const int tid = threadIdx.x;
  const int cps = 1 << cps_log2;
  const int spb = 1 << spb_log2;
  int k, slot;
  if (ORDER == 0) {
    k = tid & (cps - 1);
    slot = tid >> cps_log2;
  } else {
    slot = tid & (spb - 1);
    k = tid >> spb_log2;
  }
  const long long q = block * spb + slot;
  const bool active = q < D.nsys;
  const int N = D.N;
  I off[5];
  bool line_in = false;
  if (active) {
    sys_offsets<I>(D, q, off, &line_in);
  } else {
#pragma unroll
    for (int t = 0; t < 5; t++) off[t] = 0;
  }
  const int row0 = k * R;
  const I s0 = (I)D.sstr[0], s1 = (I)D.sstr[1], s2 = (I)D.sstr[2],
          s3 = (I)D.sstr[3], s4 = (I)D.sstr[4];
  T a[R], b[R], c[R], d[R];
  const bool full = active && (row0 + R <= N);

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *stage = nullptr;
  I wave_off[5];
  bool wave_active = false;
  // general staging: the wave's span of wave_len elements starts at
  // wave_off[t] in array t; the lane's chunk is at span index g_idx, with
  // g_rows valid rows
  int wave_len = 0, g_idx = 0, g_rows = 0;
  if (GSTAGE) {
    stage = reinterpret_cast<T *>(smem_raw) + (tid >> 6) * SW;
    const long long q0 = ((long long)__builtin_amdgcn_readfirstlane((int)(q >> 32)) << 32) |
                         (unsigned)__builtin_amdgcn_readfirstlane((int)(q & 0xffffffffLL));
    const int row0w = __builtin_amdgcn_readfirstlane(row0);
    wave_active = q0 < D.nsys;
    if (cps <= kWave) {
      // whole systems, dense: system q at q * N in every array
      const long long nact = D.nsys - q0 < (kWave >> cps_log2) ? D.nsys - q0 : (kWave >> cps_log2);
      wave_len = (int)(nact * N);
      g_idx = (int)(q - q0) * N + row0;
    } else {
      // 64 chunks of one system
      wave_len = N - row0w < kWave * R ? N - row0w : kWave * R;
      g_idx = row0 - row0w;
    }
    g_rows = active ? (N - row0 < R ? N - row0 : R) : 0;
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const I o = off[t] + (I)row0;
      if constexpr (sizeof(I) == 4) {
        wave_off[t] = (I)__builtin_amdgcn_readfirstlane((int)o);
      } else {
        int lo = __builtin_amdgcn_readfirstlane((int)(o & 0xffffffffLL));
        int hi = __builtin_amdgcn_readfirstlane((int)(o >> 32));
        wave_off[t] = (I)(((long long)hi << 32) | (unsigned)lo);
      }
    }
  }
  if (STAGED) {
    // the wave's span starts at the element of its lane 0 (VEC 2: chunk 0
    // of the wave's first system; VEC 3: chunk 64 w of the wave's system,
    // unit stride); the host guarantees that a wave's systems are either
    // all active or all inactive
    stage = reinterpret_cast<T *>(smem_raw) + (tid >> 6) * (kWave * R);
    wave_active = __builtin_amdgcn_readfirstlane((int)active) != 0;
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const I o = VEC == 3 ? off[t] + (I)row0 : off[t];
      if constexpr (sizeof(I) == 4) {
        wave_off[t] = (I)__builtin_amdgcn_readfirstlane((int)o);
      } else {
        int lo = __builtin_amdgcn_readfirstlane((int)(o & 0xffffffffLL));
        int hi = __builtin_amdgcn_readfirstlane((int)(o >> 32));
        wave_off[t] = (I)(((long long)hi << 32) | (unsigned)lo);
      }
    }
  }

  // ---- load ------------------------------------------------------------
  if constexpr (Coef::kHook) {
    if (active) {
      coef.template load<R, I>(row0, N, line_in, off[0], s0, a, b, c);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        a[r] = T(0);
        b[r] = T(1);
        c[r] = T(0);
      }
    }
  } else if constexpr (!Coef::kArrays) {
    // the line's rows (identity rows on lines off the interior), then
    // identity at row 0 and at rows N-1 and past: one select per
    // coefficient and lane, and per row only in the chunks that hold row 0
    // or row N-1
    const T ca = line_in ? coef.a : T(0), cb = line_in ? coef.b : T(1),
            cc = line_in ? coef.c : T(0);
#pragma unroll
    for (int r = 0; r < R; r++) {
      a[r] = ca;
      b[r] = cb;
      c[r] = cc;
    }
    if (row0 == 0) {
      a[0] = T(0);
      b[0] = T(1);
      c[0] = T(0);
    }
    if (row0 + R >= N) {
#pragma unroll
      for (int r = 0; r < R; r++)
        if (row0 + r + 1 >= N) {
          a[r] = T(0);
          b[r] = T(1);
          c[r] = T(0);
        }
    }
  }
  if (GSTAGE) {
    const int lane = tid & (kWave - 1);
    if (wave_active) {
      if constexpr (Coef::kArrays) {
        int sh = gspan_load<T, R, I, CPC>(coef.A, wave_off[0], wave_len, stage, lane);
        gspan_take<T, R>(stage, sh + g_idx, g_rows, a, T(0));
        sh = gspan_load<T, R, I, CPC>(coef.B, wave_off[1], wave_len, stage, lane);
        gspan_take<T, R>(stage, sh + g_idx, g_rows, b, T(1));
        sh = gspan_load<T, R, I, CPC>(coef.C, wave_off[2], wave_len, stage, lane);
        gspan_take<T, R>(stage, sh + g_idx, g_rows, c, T(0));
      }
      const int sh = gspan_load<T, R, I, CPD>(Dd, wave_off[3], wave_len, stage, lane);
      gspan_take<T, R>(stage, sh + g_idx, g_rows, d, T(0));
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        a[r] = T(0);
        b[r] = T(1);
        c[r] = T(0);
        d[r] = T(0);
      }
    }
    if (row0 == 0) a[0] = T(0);
#pragma unroll
    for (int r = 0; r < R; r++)
      if (row0 + r == N - 1) c[r] = T(0);
  } else if (STAGED) {
    // host guarantees: every chunk of an active wave is full and in range
    const int lane = tid & (kWave - 1);
    if (wave_active) {
      if constexpr (Coef::kArrays) {
        load_chunk_staged<T, R, CPC>(at(coef.A, wave_off[0]), stage, lane, a);
        load_chunk_staged<T, R, CPC>(at(coef.B, wave_off[1]), stage, lane, b);
        load_chunk_staged<T, R, CPC>(at(coef.C, wave_off[2]), stage, lane, c);
      }
      load_chunk_staged<T, R, CPD>(at(Dd, wave_off[3]), stage, lane, d);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        a[r] = T(0);
        b[r] = T(1);
        c[r] = T(0);
        d[r] = T(0);
      }
    }
  } else if (full) {
    I oa = off[0] + (I)row0 * s0, ob = off[1] + (I)row0 * s1,
      oc = off[2] + (I)row0 * s2, od = off[3] + (I)row0 * s3;
    // (staged layouts never get here: VEC is 0 or 1)
    if constexpr (VEC == 1) {
      if constexpr (Coef::kArrays) {
        load_chunk_vec<T, R, CPC>(at(coef.A, oa), a);
        load_chunk_vec<T, R, CPC>(at(coef.B, ob), b);
        load_chunk_vec<T, R, CPC>(at(coef.C, oc), c);
      }
      load_chunk_vec<T, R, CPD>(at(Dd, od), d);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        if constexpr (Coef::kArrays) {
          a[r] = ld_stream<CPC>(at(coef.A, oa));
          b[r] = ld_stream<CPC>(at(coef.B, ob));
          c[r] = ld_stream<CPC>(at(coef.C, oc));
        }
        d[r] = ld_stream<CPD>(at(Dd, od));
        oa += s0;
        ob += s1;
        oc += s2;
        od += s3;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = row0 + r;
      if (active && row < N) {
        if constexpr (Coef::kArrays) {
          a[r] = row == 0 ? T(0) : ld_stream<CPC>(at(coef.A, off[0] + (I)row * s0));
          b[r] = ld_stream<CPC>(at(coef.B, off[1] + (I)row * s1));
          c[r] = row == N - 1 ? T(0) : ld_stream<CPC>(at(coef.C, off[2] + (I)row * s2));
        }
        d[r] = ld_stream<CPD>(at(Dd, off[3] + (I)row * s3));
      } else {
        a[r] = T(0);
        b[r] = T(1);
        c[r] = T(0);
        d[r] = T(0);
      }
    }
  }
  if constexpr (Coef::kArrays) {
    // a_0 and c_{N-1} of every system are never used (the partial-chunk
    // path above already masked them)
    if (STAGED || full) {
      if (row0 == 0) a[0] = T(0);
      if (row0 + R == N) c[R - 1] = T(0);
    }
  }
  if constexpr (MODE == 2) {
    if (active && (row0 == 0 || row0 + R == N)) {
      long long qs;
      int p;
      seg_item<T, ORDER>(so, q, qs, p);
      const long long xb = qs * so.rs_sys;
      if (row0 == 0 && p > 0)
        d[0] -= *at(coef.A, off[0]) * so.rd[xb + (2LL * p - 1) * so.rs_row];
      if (row0 + R == N && p + 1 < so.P)
        d[R - 1] -= *at(coef.C, off[2] + (I)(N - 1) * s2) * so.rd[xb + (2LL * p + 2) * so.rs_row];
    }
  }
  if constexpr (MODE == 1) {
    // unit couplings at the segment's ends: the merged rows' outer
    // coefficients are then proportional to the true a and c, which the
    // writer below multiplies in (zero at the ends of the whole system); a
    // partial last chunk ends its system, so its zero c stays
    if (row0 == 0) a[0] = T(1);
    if ((STAGED || full) && row0 + R == N) c[R - 1] = T(1);
  }

  // ---- solve -----------------------------------------------------------
  Row<T> f, l;
  chunk_reduce<T, R>(a, b, c, d, f, l);

  if constexpr (MODE == 1) {
    // merge the segment's chunk rows in a tree; chunk 0 ends with the
    // segment's rows
    if (ORDER == 0 && cps <= kWave) {
      const int lane = tid & (kWave - 1);
      for (int s = 1; s < cps; s <<= 1) {
        const int src = (lane + s) & (kWave - 1);
        Row<T> fb, lb;
        fb.a = lane_read(f.a, src);
        fb.c = lane_read(f.c, src);
        fb.d = lane_read(f.d, src);
        lb.a = lane_read(l.a, src);
        lb.c = lane_read(l.c, src);
        lb.d = lane_read(l.d, src);
        if ((k & (2 * s - 1)) == 0) seg_merge<T>(f, l, fb, lb);
      }
    } else {
      // entry (chunk k, slot) at k * spb + slot of six arrays of cps * spb
      // values; at level s chunk k (k % 2s == 0) reads entry k + s, which
      // was last written at level s/2, and rewrites its own
      T *sm = reinterpret_cast<T *>(smem_raw);
      const int n = cps * spb, e = k * spb + slot;
      auto put = [&]() {
        sm[e] = f.a;
        sm[n + e] = f.c;
        sm[2 * n + e] = f.d;
        sm[3 * n + e] = l.a;
        sm[4 * n + e] = l.c;
        sm[5 * n + e] = l.d;
      };
      put();
      for (int s = 1; s < cps; s <<= 1) {
        __syncthreads();
        if ((k & (2 * s - 1)) == 0) {
          const int o = (k + s) * spb + slot;
          Row<T> fb, lb;
          fb.a = sm[o];
          fb.c = sm[n + o];
          fb.d = sm[2 * n + o];
          lb.a = sm[3 * n + o];
          lb.c = sm[4 * n + o];
          lb.d = sm[5 * n + o];
          seg_merge<T>(f, l, fb, lb);
          put();
        }
      }
    }
    if (k == 0 && active) {
      long long qs;
      int p;
      seg_item<T, ORDER>(so, q, qs, p);
      const T as = p > 0 ? *at(coef.A, off[0]) : T(0);
      const T ce = p + 1 < so.P ? *at(coef.C, off[2] + (I)(N - 1) * s2) : T(0);
      const long long i0 = qs * so.rs_sys + (long long)(2 * p) * so.rs_row;
      const long long i1 = i0 + so.rs_row;
      so.ra[i0] = f.a * as;
      so.rb[i0] = T(1);
      so.rc[i0] = f.c;
      so.rd[i0] = f.d;
      so.ra[i1] = l.a;
      so.rb[i1] = T(1);
      so.rc[i1] = l.c * ce;
      so.rd[i1] = l.d;
    }
    return;
  }

  T xf, xl;
  if (ORDER == 0 && cps <= kWave) {
    reduced_wave<T>(f, l, k, cps, tid & (kWave - 1), xf, xl);
  } else if (ORDER == 0 && (VEC == 3 || VEC == 4)) {
    // wave rows after the wave staging buffers
    reduced_multi<T>(f, l, k, cps, slot,
                     reinterpret_cast<T *>(smem_raw) + (size_t)(blockDim.x >> 6) * SW, xf, xl);
  } else if (cps <= kWave) {
    reduced_lds_w<T>(f, l, k, cps, slot, spb, reinterpret_cast<T *>(smem_raw), xf, xl);
  } else {
    reduced_lds<T>(f, l, k, cps, slot, spb, reinterpret_cast<T *>(smem_raw),
                   xf, xl);
  }

  T y[R];
  chunk_expand<T, R>(a, c, d, xf, xl, y);

  // ---- store -----------------------------------------------------------
  // non-Inc: the solution goes to Out (default d); Inc: Out (default u)
  // receives u + solution
  T *__restrict__ Dst = Out ? Out : Dd;
  T *__restrict__ Udst = Out ? Out : U;
  if (GSTAGE) {
    const int lane = tid & (kWave - 1);
    if (wave_active) {
      const int t = INC ? 4 : 3;
      const int sh = (int)(wave_off[t] & (I)(Vec16<T>::n - 1));
#pragma unroll
      for (int r = 0; r < R; r++)
        if (r < g_rows) stage[sh + g_idx + r] = y[r];
      __builtin_amdgcn_wave_barrier();
      if (INC)
        gspan_store<T, R, I, CPD>(Udst, wave_off[4], wave_len, stage, lane, true,
                             static_cast<const T *>(U));
      else
        gspan_store<T, R, I, CPD>(Dst, wave_off[3], wave_len, stage, lane, false);
    }
  } else if (STAGED) {
    const int lane = tid & (kWave - 1);
    if (wave_active) {
      if (INC)
        store_chunk_staged<T, R, CPD>(at(Udst, wave_off[4]), stage, lane, y, true,
                                 at(static_cast<const T *>(U), wave_off[4]));
      else
        store_chunk_staged<T, R, CPD>(at(Dst, wave_off[3]), stage, lane, y, false);
    }
  } else if (full) {
    if (INC) {
      I ou = off[4] + (I)row0 * s4;
      if constexpr (VEC == 1) {
        T u[R];
        load_chunk_vec<T, R, CPD>(at(U, ou), u);
#pragma unroll
        for (int r = 0; r < R; r++) u[r] += y[r];
        store_chunk_vec<T, R, CPD>(at(Udst, ou), u);
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
          st_stream<CPD>(at(Udst, ou), ld_stream<CPD>(at(U, ou)) + y[r]);
          ou += s4;
        }
      }
    } else {
      I od = off[3] + (I)row0 * s3;
      if constexpr (VEC == 1) {
        store_chunk_vec<T, R, CPD>(at(Dst, od), y);
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
          st_stream<CPD>(at(Dst, od), y[r]);
          od += s3;
        }
      }
    }
  } else if (active) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = row0 + r;
      if (row < N) {
        if (INC) {
          const I ou = off[4] + (I)row * s4;
          st_stream<CPD>(at(Udst, ou), ld_stream<CPD>(at(U, ou)) + y[r]);
        } else {
          st_stream<CPD>(at(Dst, off[3] + (I)row * s3), y[r]);
        }
      }
    }
  }
}

}  // namespace tridhip
