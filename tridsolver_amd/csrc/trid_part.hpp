// trid_part.hpp -- the partitioned ("register-blocked") Thomas solver.
//
// Every system is cut into chunks of R consecutive rows; one lane owns one
// chunk of one system and keeps all of it in registers.  Each lane runs a
// modified Thomas sweep over its chunk that expresses every row through the
// chunk's first and last unknowns, which leaves two rows per chunk coupled
// to the neighbouring chunks.  Those 2*C rows (C = chunks per system) form a
// tridiagonal system again; it is solved by parallel cyclic reduction (PCR),
// across lanes with ds_bpermute when all chunks of a system are in one wave,
// through LDS otherwise.  Each lane then recovers its interior rows from the
// two solved boundary unknowns.  Global memory is read once (a, b, c, d) and
// written once (d, or u for the Inc form): no c'/d' scratch traffic.
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

// PCR over the 2*cps reduced rows of a system whose cps chunks sit in
// consecutive lanes of one wave (cps a power of two <= 64).  k is the lane's
// chunk index, lane its lane id.  Rows are ordered f_0, l_0, f_1, l_1, ...
// On return f.d and l.d hold the chunk's first and last unknowns.
template <typename T>
__device__ __forceinline__ void pcr_wave(Row<T> &f, Row<T> &l, int k, int cps,
                                         int lane) {
  {
    Row<T> lm = row_from_lane(l, lane - 1, k > 0);
    Row<T> fp = row_from_lane(f, lane + 1, k + 1 < cps);
    Row<T> nf = pcr_update(f, lm, l);
    Row<T> nl = pcr_update(l, f, fp);
    f = nf;
    l = nl;
  }
  for (int s = 1; s < cps; s <<= 1) {
    Row<T> fm = row_from_lane(f, lane - s, k >= s);
    Row<T> fp = row_from_lane(f, lane + s, k + s < cps);
    Row<T> lm = row_from_lane(l, lane - s, k >= s);
    Row<T> lp = row_from_lane(l, lane + s, k + s < cps);
    f = pcr_update(f, fm, fp);
    l = pcr_update(l, lm, lp);
  }
}

// Same PCR through LDS for chunks spread over several waves.  Row rho of
// system slot s lives at rho * spb + s.  All threads of the block call it.
template <typename T>
__device__ __forceinline__ void pcr_lds(Row<T> &f, Row<T> &l, int k, int cps,
                                        int slot, int spb, T *sa, T *sc,
                                        T *sd) {
  const int M = 2 * cps;
  const int rf = 2 * k, rl = 2 * k + 1;
  for (int h = 1; h < M; h <<= 1) {
    sa[rf * spb + slot] = f.a;
    sc[rf * spb + slot] = f.c;
    sd[rf * spb + slot] = f.d;
    sa[rl * spb + slot] = l.a;
    sc[rl * spb + slot] = l.c;
    sd[rl * spb + slot] = l.d;
    __syncthreads();
    Row<T> z{T(0), T(0), T(0)};
    Row<T> fm = z, fp = z, lm = z, lp = z;
    if (rf - h >= 0) {
      int i = (rf - h) * spb + slot;
      fm = Row<T>{sa[i], sc[i], sd[i]};
    }
    if (rf + h < M) {
      int i = (rf + h) * spb + slot;
      fp = Row<T>{sa[i], sc[i], sd[i]};
    }
    if (rl - h >= 0) {
      int i = (rl - h) * spb + slot;
      lm = Row<T>{sa[i], sc[i], sd[i]};
    }
    if (rl + h < M) {
      int i = (rl + h) * spb + slot;
      lp = Row<T>{sa[i], sc[i], sd[i]};
    }
    __syncthreads();
    f = pcr_update(f, fm, fp);
    l = pcr_update(l, lm, lp);
  }
}

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
};
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
};

template <typename T, int R>
__device__ __forceinline__ void load_chunk_vec(const T *__restrict__ p,
                                               T (&x)[R]) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  static_assert(R % E == 0, "chunk must be a whole number of 16-byte vectors");
  const V *pv = reinterpret_cast<const V *>(p);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V t = pv[v];
    if constexpr (E == 2) {
      x[2 * v] = t.x;
      x[2 * v + 1] = t.y;
    } else {
      x[4 * v] = t.x;
      x[4 * v + 1] = t.y;
      x[4 * v + 2] = t.z;
      x[4 * v + 3] = t.w;
    }
  }
}

template <typename T, int R>
__device__ __forceinline__ void store_chunk_vec(T *__restrict__ p,
                                                const T (&x)[R]) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  V *pv = reinterpret_cast<V *>(p);
#pragma unroll
  for (int v = 0; v < R / E; v++) {
    V t;
    if constexpr (E == 2) {
      t.x = x[2 * v];
      t.y = x[2 * v + 1];
    } else {
      t.x = x[4 * v];
      t.y = x[4 * v + 1];
      t.z = x[4 * v + 2];
      t.w = x[4 * v + 3];
    }
    pv[v] = t;
  }
}

// Block-wide partitioned solve of a strided batch.
//   ORDER 0: thread t -> chunk t % cps, system slot t / cps (chunks of a
//            system in consecutive lanes: contiguous systems, solvedim 0).
//   ORDER 1: thread t -> system slot t % spb, chunk t / spb (systems in
//            consecutive lanes: coalesced across systems, solvedim >= 1).
//   VEC:     16-byte vector loads/stores along the system (ORDER 0 with unit
//            stride and 16-byte aligned chunks only; checked on the host).
// blockDim.x == cps * spb, both powers of two.
template <typename T, int INC, int ORDER, int R, int VEC>
__device__ __forceinline__ void part_block(const T *__restrict__ A,
                                           const T *__restrict__ B,
                                           const T *__restrict__ C,
                                           T *__restrict__ Dd,
                                           T *__restrict__ U, const SysDesc &D,
                                           int cps_log2, int spb_log2,
                                           long long block) {
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
  long long off[5];
  if (active) {
    sys_offsets(D, q, off);
  } else {
#pragma unroll
    for (int t = 0; t < 5; t++) off[t] = 0;
  }
  const int row0 = k * R;
  T a[R], b[R], c[R], d[R];
  const bool full = active && (row0 + R <= N);
  if (full) {
    const T *pa = A + off[0] + row0 * D.sstr[0];
    const T *pb = B + off[1] + row0 * D.sstr[1];
    const T *pc = C + off[2] + row0 * D.sstr[2];
    const T *pd = Dd + off[3] + row0 * D.sstr[3];
    if (VEC) {
      load_chunk_vec<T, R>(pa, a);
      load_chunk_vec<T, R>(pb, b);
      load_chunk_vec<T, R>(pc, c);
      load_chunk_vec<T, R>(pd, d);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        a[r] = pa[r * D.sstr[0]];
        b[r] = pb[r * D.sstr[1]];
        c[r] = pc[r * D.sstr[2]];
        d[r] = pd[r * D.sstr[3]];
      }
    }
    if (row0 == 0) a[0] = T(0);
    if (row0 + R == N) c[R - 1] = T(0);
  } else {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = row0 + r;
      if (active && row < N) {
        a[r] = row == 0 ? T(0) : A[off[0] + row * D.sstr[0]];
        b[r] = B[off[1] + row * D.sstr[1]];
        c[r] = row == N - 1 ? T(0) : C[off[2] + row * D.sstr[2]];
        d[r] = Dd[off[3] + row * D.sstr[3]];
      } else {
        a[r] = T(0);
        b[r] = T(1);
        c[r] = T(0);
        d[r] = T(0);
      }
    }
  }

  Row<T> f, l;
  chunk_reduce<T, R>(a, b, c, d, f, l);

  if (ORDER == 0 && cps <= kWave) {
    pcr_wave<T>(f, l, k, cps, tid & (kWave - 1));
  } else {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *sm = reinterpret_cast<T *>(smem_raw);
    const int M = 2 * cps * spb;
    pcr_lds<T>(f, l, k, cps, slot, spb, sm, sm + M, sm + 2 * M);
  }

  T y[R];
  chunk_expand<T, R>(a, c, d, f.d, l.d, y);

  if (full) {
    if (INC) {
      T *pu = U + off[4] + row0 * D.sstr[4];
      if (VEC) {
        T u[R];
        load_chunk_vec<T, R>(pu, u);
#pragma unroll
        for (int r = 0; r < R; r++) u[r] += y[r];
        store_chunk_vec<T, R>(pu, u);
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) pu[r * D.sstr[4]] += y[r];
      }
    } else {
      T *pd = Dd + off[3] + row0 * D.sstr[3];
      if (VEC) {
        store_chunk_vec<T, R>(pd, y);
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) pd[r * D.sstr[3]] = y[r];
      }
    }
  } else if (active) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = row0 + r;
      if (row < N) {
        if (INC)
          U[off[4] + row * D.sstr[4]] += y[r];
        else
          Dd[off[3] + row * D.sstr[3]] = y[r];
      }
    }
  }
}

}  // namespace tridhip
