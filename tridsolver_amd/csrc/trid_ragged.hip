// trid_ragged.hip -- ragged batches in one persistent launch
// (include/trid_ragged.h).
//
// Every system is solved with the partitioned Thomas scheme of
// trid_part.hpp (chunks of R = 8 rows per lane, reduced system by PCR), with the number of chunks per system (cps,
// 1..512) chosen per length bin.  A work item is one work-group of 512
// threads processing 512/cps systems of one bin.
// Loads are staged per span group through LDS: the lanes of a system (or of
// one wave's 64 chunks of a longer system) read its contiguous span as
// 16-byte pieces, then each lane takes its chunk from LDS (the transpose of
// the solvedim-0 path, for spans at arbitrary offsets and alignments).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trid_hip.h"
#include "../../include/trid_hip_ext.h"
#include "../../include/trid_ragged.h"
#include "trid_part.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

constexpr int kRagThreads = 512;
constexpr int kRagMaxLen = 4096;  // longest system of the persistent kernel
// systems of kRagMaxLen+1 .. kRagLongLen rows: a second launch of the
// persistent kernel with 512 chunks of 9..16 rows (more registers per lane,
// so it does not share the first launch's instantiation)
constexpr int kRagLongLen = 8192;
constexpr int kRagR2min = 9, kRagR2max = 16;
// cost model of the systems past kRagLongLen rows, measured on MI355X (FP64):
// one strided one-system solve takes ~kSoloUs (27 us for 9000 rows, 42 us
// for 2^20 rows: launches and latency); the one-thread kernel ~kThreadRowUs
// per row of its longest system (0.4-0.8 us: a dependent chain of global
// accesses)
constexpr double kSoloUs = 25.0;
constexpr double kThreadRowUs = 0.5;

// Chunk lengths of the persistent kernel.  A bin is (cps, R): cps a power of
// two (1..512) and R the shortest of 5..8 rows (FP64) with cps * R >= L, so
// that at least 4/5 of a system's lanes hold rows (a power-of-two chunk count
// with one fixed R leaves a quarter idle on average).  The kernel's registers
// are those of the longest, 8 rows (512 chunks of 8 rows cover 4096 rows).
// FP32 rows take half the registers: its bins use 9..16 rows (kRagR2min..
// kRagR2max), half the chunks and reduced-system levels, up to 8192 rows in
// the first launch.
constexpr int kRagRmin = 5, kRagRmax = 8;

struct RagItem {
  int cps_log2;  // chunks per system
  int first;     // first entry of the item in the sorted system list
  int count;     // systems in this item (<= kRagThreads / cps)
  int rows;      // chunk length R of the item's bin (5..8, or 9..16)
};

template <typename T>
struct RagArgs {
  const T *a, *b, *c;
  T *d, *u;
  int inc;
  const int64_t *offsets;
  const int *sys;  // system ids, grouped by bin
  const RagItem *items;
  int nitems;
  unsigned *counter;  // [0] next item, [1] finished work-groups
};

// Span-group staging.  A span group is the G lanes that hold one contiguous
// row range of one system: the whole system when cps <= 64 (G = cps), else
// the 64 chunks of one wave (G = 64), so a group never leaves its wave and
// no work-group barrier is needed.  The group reads its span [p, p + len)
// of an array as the 16-byte aligned pieces that cover it (piece j at group
// lane j mod G; lanes past the span repeat its last piece, which keeps the
// loads unpredicated), into the group's LDS region of G*R + 2E elements, so
// that element p + i sits at reg[sh + i], sh = (address of p / sizeof(T))
// mod E.  A span of G*R elements needs at most G*R/E + 1 pieces (at most
// G*R + 2E - 2 elements), hence ceil(R/E) + 1 rounds of G lanes and the 2E
// spare elements when R is not a multiple of E.  A piece is the aligned
// block that holds a span element, so no load reaches past that element's
// page.  Returns sh.
template <typename T, int R>
__device__ __forceinline__ int rag_load(const T *__restrict__ p, int len, T *__restrict__ reg,
                                        int kg, int G) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  const int sh = (int)((reinterpret_cast<uintptr_t>(p) / sizeof(T)) & (E - 1));
  const V *g = reinterpret_cast<const V *>(p - sh);
  const int np = len > 0 ? (sh + len + E - 1) / E : 0;
  const int last = np > 0 ? np - 1 : 0;
  constexpr int NR = (R + E - 1) / E + 1;
  V t[NR];
#pragma unroll
  for (int j = 0; j < NR; j++) {
    const int pc = j * G + kg;
    t[j] = ld_nt(g + (pc < np ? pc : last));
  }
  V *s = reinterpret_cast<V *>(reg);
#pragma unroll
  for (int j = 0; j < NR; j++) {
    const int pc = j * G + kg;
    if (pc < np) s[pc] = t[j];
  }
  __builtin_amdgcn_wave_barrier();
  return sh;
}

// The lane's rows i0 .. i0+R-1 of the span (pad past len) from reg.
template <typename T, int R>
__device__ __forceinline__ void rag_take(const T *reg, int sh, int i0, int len, T (&x)[R],
                                         T pad) {
#pragma unroll
  for (int r = 0; r < R; r++) x[r] = i0 + r < len ? reg[sh + i0 + r] : pad;
  __builtin_amdgcn_wave_barrier();
}

// Writes the span [p, p + len) from reg (filled by the lanes, then a wave
// barrier): pieces wholly inside the span as 16-byte stores, the elements of
// the boundary pieces one by one (the rest of them belongs to other
// systems).  add: p receives its previous contents plus reg.
template <typename T, int R>
__device__ __forceinline__ void rag_store(T *__restrict__ p, int len, const T *__restrict__ reg,
                                          int kg, int G, bool add) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  const int sh = (int)((reinterpret_cast<uintptr_t>(p) / sizeof(T)) & (E - 1));
  V *g = reinterpret_cast<V *>(p - sh);
  const int np = len > 0 ? (sh + len + E - 1) / E : 0;
  const V *s = reinterpret_cast<const V *>(reg);
  constexpr int NR = (R + E - 1) / E + 1;
#pragma unroll
  for (int j = 0; j < NR; j++) {
    const int pc = j * G + kg;
    if (pc < np) {
      const int lo = pc * E - sh;
      V w = s[pc];
      if (lo >= 0 && lo + E <= len) {
        if (add) w += ld_nt(g + pc);
        st_nt(g + pc, w);
      } else {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int i = lo + e;
          if (i >= 0 && i < len) {
            T v = w[e];
            if (add) v += ld_nt(p + i);
            st_nt(p + i, v);
          }
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
}

template <typename T, int R>
__device__ __forceinline__ void rag_item(const RagArgs<T> &g, const RagItem &it, T *smem) {
  constexpr int E = Vec16<T>::n;
  const int tid = threadIdx.x;
  const int cps = 1 << it.cps_log2;
  const int k = tid & (cps - 1);
  const int slot = tid >> it.cps_log2;
  const bool active = slot < it.count;
  long long off = 0;
  int L = 0;
  if (active) {
    const int id = g.sys[it.first + slot];
    off = g.offsets[id];
    L = (int)(g.offsets[id + 1] - off);
  }
  // the lane's span group: G lanes, rows row_base .. of the system
  const int G = cps < kWave ? cps : kWave;
  const int kg = k & (G - 1);
  const int row_base = (k - kg) * R;
  const int len = L - row_base < 0 ? 0 : (L - row_base < G * R ? L - row_base : G * R);
  T *reg = smem + (tid / G) * (G * R + 2 * E);
  // a group with nothing to load (inactive slot, or a wave past its
  // system's end) points its loads at element 0, which every array has (the
  // plan holds only systems of at least one row)
  const long long p0 = len > 0 ? off + row_base : 0;
  const int i0 = kg * R;  // the lane's first row within the span
  T a[R], b[R], c[R], d[R];
  int sh = rag_load<T, R>(g.a + p0, len, reg, kg, G);
  rag_take<T, R>(reg, sh, i0, len, a, T(0));
  sh = rag_load<T, R>(g.b + p0, len, reg, kg, G);
  rag_take<T, R>(reg, sh, i0, len, b, T(1));
  sh = rag_load<T, R>(g.c + p0, len, reg, kg, G);
  rag_take<T, R>(reg, sh, i0, len, c, T(0));
  sh = rag_load<T, R>(g.d + p0, len, reg, kg, G);
  rag_take<T, R>(reg, sh, i0, len, d, T(0));
  const int row0 = k * R;
  if (row0 == 0) a[0] = T(0);
#pragma unroll
  for (int r = 0; r < R; r++)
    if (row0 + r == L - 1) c[r] = T(0);

  Row<T> f, l;
  chunk_reduce<T, R>(a, b, c, d, f, l);
  T xf, xl;
  if (cps <= kWave) {
    reduced_wave<T>(f, l, k, cps, tid & (kWave - 1), xf, xl);
  } else {
    // the PCR exchange buffers (6 values per thread) reuse the staging
    // region once every wave is done with it
    __syncthreads();
    reduced_lds<T>(f, l, k, cps, slot, kRagThreads >> it.cps_log2, smem, xf, xl);
    __syncthreads();
  }
  T y[R];
  chunk_expand<T, R>(a, c, d, xf, xl, y);
  T *dst = (g.inc ? g.u : g.d) + p0;
  const int shd = (int)((reinterpret_cast<uintptr_t>(dst) / sizeof(T)) & (E - 1));
#pragma unroll
  for (int r = 0; r < R; r++)
    if (i0 + r < len) reg[shd + i0 + r] = y[r];
  __builtin_amdgcn_wave_barrier();
  rag_store<T, R>(dst, len, reg, kg, G, g.inc != 0);
}

// rag_item for the item's chunk length, one of RLO..RHI
template <typename T, int R, int RHI>
__device__ __forceinline__ void rag_dispatch(const RagArgs<T> &g, const RagItem &it, T *smem) {
  if constexpr (R == RHI) {
    rag_item<T, R>(g, it, smem);
  } else {
    if (it.rows == R) rag_item<T, R>(g, it, smem);
    else rag_dispatch<T, R + 1, RHI>(g, it, smem);
  }
}

// Persistent work-groups take items from an atomic counter; the next
// item's index is fetched while the current item is processed.  RLO..RHI:
// the chunk lengths of the launch's items.
template <typename T, int RLO, int RHI>
__global__ void __launch_bounds__(kRagThreads) k_ragged(RagArgs<T> g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  __shared__ int s_next;
  if (threadIdx.x == 0) s_next = (int)atomicAdd(g.counter, 1u);
  __syncthreads();
  int i = s_next;
  while (i < g.nitems) {
    __syncthreads();  // every lane has read s_next
    if (threadIdx.x == 0) s_next = (int)atomicAdd(g.counter, 1u);
    const RagItem it = g.items[i];
    rag_dispatch<T, RLO, RHI>(g, it, smem);
    __syncthreads();  // s_next written; the next item reuses the staging region
    i = s_next;
  }
  // the last work-group to finish resets the counters for the next launch
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned done = atomicAdd(g.counter + 1, 1u);
    if (done == gridDim.x - 1) {
      atomicExch(g.counter, 0u);
      atomicExch(g.counter + 1, 0u);
    }
  }
}

// Systems longer than kRagMaxLen: one thread per system, forward sweep's
// c'/d' in a workspace (ws_off[j] = start of system j's 2*L slots).
template <typename T>
__global__ void __launch_bounds__(64)
    k_ragged_long(RagArgs<T> g, const int *ids, const long long *ws_off, int n,
                  T *ws) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int id = ids[j];
  const long long off = g.offsets[id];
  const int L = (int)(g.offsets[id + 1] - off);
  T *cs = ws + ws_off[j];
  T *ds = cs + L;
  T bb = rcp(g.b[off]);
  T cc = L > 1 ? bb * g.c[off] : T(0);
  T dd = bb * g.d[off];
  cs[0] = cc;
  ds[0] = dd;
  for (int i = 1; i < L; i++) {
    const T aa = g.a[off + i];
    bb = rcp(g.b[off + i] - aa * cc);
    dd = (g.d[off + i] - aa * dd) * bb;
    cc = i < L - 1 ? bb * g.c[off + i] : T(0);
    cs[i] = cc;
    ds[i] = dd;
  }
  for (int i = L - 1; i >= 0; i--) {
    if (i < L - 1) dd = ds[i] - cs[i] * dd;
    if (g.inc)
      g.u[off + i] += dd;
    else
      g.d[off + i] = dd;
  }
}

}  // namespace tridhip

using namespace tridhip;

struct tridRaggedPlan_st {
  int dbl = 1;
  long long nsys = 0;
  long long extent = 0;  // offsets[nsys]: elements every array must hold
  const int64_t *offsets = nullptr;  // caller's device array
  int *sys = nullptr;
  RagItem *items = nullptr;
  int nitems = 0;
  unsigned *counter = nullptr;
  int grid = 0;
  // second launch: systems of kRagMaxLen+1 .. kRagLongLen rows
  int *sys2 = nullptr;
  RagItem *items2 = nullptr;
  int nitems2 = 0;
  int grid2 = 0;
  size_t lds2 = 0;
  int *long_ids = nullptr;
  long long *long_ws_off = nullptr;
  int nlong = 0;
  long long long_ws = 0;  // elements
  // the longest systems, each solved as a one-system strided batch (split-line
  // solve past 8192 rows): (offset, length)
  std::vector<std::pair<long long, int>> solo;
  size_t lds = 0;
  std::string text;
};

namespace {

void free_plan(tridRaggedPlan_st *p) {
  if (!p) return;
  (void)hipFree(p->sys);
  (void)hipFree(p->items);
  (void)hipFree(p->sys2);
  (void)hipFree(p->items2);
  (void)hipFree(p->counter);
  (void)hipFree(p->long_ids);
  (void)hipFree(p->long_ws_off);
  delete p;
}

template <typename T>
tridStatus_t ragged_solve(tridRaggedPlan_st *p, const T *a, const T *b, const T *c,
                          T *d, T *u, int inc, int sync) {
  clear_error();
  if (!p) return fail(TRID_STATUS_INVALID_VALUE, "ragged plan is NULL");
  if ((sizeof(T) == 8) != (p->dbl == 1))
    return fail(TRID_STATUS_INVALID_VALUE, "ragged plan precision differs from the call");
  if (p->nsys == 0) return TRID_STATUS_SUCCESS;
  if (!a || !b || !c || !d || (inc && !u))
    return fail(TRID_STATUS_INVALID_VALUE, "NULL array");
  hipStream_t st = current_stream();
  RagArgs<T> g{a, b, c, d, u, inc ? 1 : 0, p->offsets, p->sys, p->items, p->nitems, p->counter};
  if (p->nitems > 0) {
    // FP64 items have chunks of 5..8 rows, FP32 ones of 9..16
    if constexpr (sizeof(T) == 8)
      hipLaunchKernelGGL((k_ragged<T, kRagRmin, kRagRmax>), dim3(p->grid), dim3(kRagThreads),
                         p->lds, st, g);
    else
      hipLaunchKernelGGL((k_ragged<T, kRagR2min, kRagR2max>), dim3(p->grid), dim3(kRagThreads),
                         p->lds, st, g);
    tridStatus_t s = check_hip(hipGetLastError(), "ragged kernel launch");
    if (s) return s;
  }
  if (p->nitems2 > 0) {
    // stream-ordered after the first launch, whose last work-group has reset
    // the counter
    RagArgs<T> g2 = g;
    g2.sys = p->sys2;
    g2.items = p->items2;
    g2.nitems = p->nitems2;
    hipLaunchKernelGGL((k_ragged<T, kRagR2min, kRagR2max>), dim3(p->grid2), dim3(kRagThreads),
                       p->lds2, st, g2);
    tridStatus_t s = check_hip(hipGetLastError(), "ragged kernel launch");
    if (s) return s;
  }
  if (p->nlong > 0) {
    T *ws = static_cast<T *>(workspace((size_t)p->long_ws * sizeof(T), st));
    if (!ws) return TRID_STATUS_ALLOC_FAILED;  // reason recorded by workspace()
    hipLaunchKernelGGL(k_ragged_long<T>, dim3((p->nlong + 63) / 64), dim3(64), 0, st, g,
                       p->long_ids, p->long_ws_off, p->nlong, ws);
    tridStatus_t s = check_hip(hipGetLastError(), "ragged long-system launch");
    if (s) return s;
  }
  for (const auto &sl : p->solo) {
    int dims[1] = {sl.second}, opts[1] = {0};
    const long long o = sl.first;
    tridStatus_t s;
    if constexpr (sizeof(T) == 8)
      s = inc ? tridDmtsvStridedBatchInc(a + o, b + o, c + o, d + o, u + o, 1, 0, dims, dims,
                                         opts, 0)
              : tridDmtsvStridedBatch(a + o, b + o, c + o, d + o, nullptr, 1, 0, dims, dims,
                                      opts, 0);
    else
      s = inc ? tridSmtsvStridedBatchInc(a + o, b + o, c + o, d + o, u + o, 1, 0, dims, dims,
                                         opts, 0)
              : tridSmtsvStridedBatch(a + o, b + o, c + o, d + o, nullptr, 1, 0, dims, dims,
                                      opts, 0);
    if (s) return s;
  }
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

}  // namespace

EXTERN_C tridStatus_t tridRaggedPlanCreate(tridRaggedPlan_t *plan,
                                           const int64_t *offsets,
                                           long long nsys, int is_double) {
  clear_error();
  if (!plan) return fail(TRID_STATUS_INVALID_VALUE, "plan pointer is NULL");
  *plan = nullptr;
  if (nsys < 0 || nsys > (1LL << 31) - 2)
    return fail(TRID_STATUS_INVALID_VALUE, "nsys out of range");
  if (nsys > 0 && !offsets) return fail(TRID_STATUS_INVALID_VALUE, "offsets is NULL");
  auto *p = new (std::nothrow) tridRaggedPlan_st;
  if (!p) return fail(TRID_STATUS_ALLOC_FAILED, "plan allocation failed");
  p->dbl = is_double ? 1 : 0;
  p->nsys = nsys;
  p->offsets = offsets;
  if (nsys == 0) {
    p->text = "empty";
    *plan = p;
    return TRID_STATUS_SUCCESS;
  }
  std::vector<int64_t> off(nsys + 1);
  tridStatus_t s = check_hip(hipMemcpy(off.data(), offsets, sizeof(int64_t) * (nsys + 1),
                                       hipMemcpyDeviceToHost),
                             "reading offsets");
  if (s) {
    free_plan(p);
    return s;
  }
  if (off[0] < 0) {
    free_plan(p);
    return fail(TRID_STATUS_INVALID_VALUE, "offsets[0] must be >= 0");
  }
  for (long long k = 0; k < nsys; k++)
    if (off[k + 1] < off[k]) {
      free_plan(p);
      return fail(TRID_STATUS_INVALID_VALUE, "offsets must be non-decreasing");
    }
  p->extent = off[nsys];
  // bins (cps = 2^cl >= L / rmax, R = the shortest of rmin..rmax rows with
  // cps*R >= L): FP64 chunks of 5..8 rows; FP32 (half the registers per
  // row) chunks of 9..16 rows, half the chunks and reduced-system levels
  const int rmin = p->dbl ? kRagRmin : kRagR2min, rmax = p->dbl ? kRagRmax : kRagR2max;
  const long long max_len = (long long)kRagThreads * rmax;  // first launch
  auto bin_of = [&](long long L, int &R) {
    const long long ch = (L + rmax - 1) / rmax;
    int cl = 0;
    while ((1LL << cl) < ch) cl++;
    R = (int)std::max<long long>(rmin, (L + (1LL << cl) - 1) >> cl);
    return cl;
  };
  const int nR = rmax - rmin + 1;
  std::vector<std::vector<int>> bins(16 * nR);
  // systems past kRagMaxLen rows: the k longest are solved one by one as
  // strided batches of one system (a few launches each, bandwidth-bound),
  // the rest together by the one-thread-per-system kernel (time set by its
  // longest system); k minimises k * kSoloUs + L_(k+1) * kThreadRowUs
  std::vector<std::pair<long long, long long>> over;  // (length, system)
  for (long long k = 0; k < nsys; k++) {
    const long long L = off[k + 1] - off[k];
    if (L > kRagLongLen) over.push_back({L, k});
  }
  std::sort(over.begin(), over.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
  size_t nsolo = 0;
  {
    double best = 1e300;
    for (size_t k = 0; k <= over.size(); k++) {
      const double t = k * kSoloUs + (k < over.size() ? over[k].first * kThreadRowUs : 0.0);
      if (t < best) {
        best = t;
        nsolo = k;
      }
    }
  }
  std::vector<char> is_solo(nsolo ? nsys : 0, 0);
  for (size_t k = 0; k < nsolo; k++) {
    if (over[k].first > INT32_MAX) {
      free_plan(p);
      return fail(TRID_STATUS_INVALID_VALUE, "system longer than 2^31 - 1 rows");
    }
    is_solo[over[k].second] = 1;
    p->solo.push_back({off[over[k].second], (int)over[k].first});
  }
  std::vector<int> longs;
  std::vector<long long> long_off;
  long long lws = 0;
  constexpr int nR2 = kRagR2max - kRagR2min + 1;
  std::vector<std::vector<int>> bins2(nR2);  // 512 chunks of R = 9..16 rows
  for (long long k = 0; k < nsys; k++) {
    const long long L = off[k + 1] - off[k];
    if (L == 0 || (nsolo && is_solo[k])) continue;
    if (L > kRagLongLen) {
      longs.push_back((int)k);
      long_off.push_back(lws);
      lws += 2 * L;
      continue;
    }
    if (L > max_len) {
      const int R = (int)((L + kRagThreads - 1) / kRagThreads);
      bins2[R - kRagR2min].push_back((int)k);
      continue;
    }
    int R = 0;
    const int cl = bin_of(L, R);
    bins[cl * nR + (R - rmin)].push_back((int)k);
  }
  // items, most expensive first (more chunks per system, then longer chunks)
  std::vector<int> sys;
  std::vector<RagItem> items;
  std::string desc;
  for (int cl = 15; cl >= 0; cl--)
    for (int R = rmax; R >= rmin; R--) {
      std::vector<int> &v = bins[cl * nR + (R - rmin)];
      if (v.empty()) continue;
      const int per = std::max(1, kRagThreads >> cl);
      for (size_t i = 0; i < v.size(); i += per) {
        RagItem it;
        it.cps_log2 = cl;
        it.first = (int)sys.size() + (int)i;
        it.count = (int)std::min<size_t>(per, v.size() - i);
        it.rows = R;
        items.push_back(it);
      }
      sys.insert(sys.end(), v.begin(), v.end());
      char buf[96];
      snprintf(buf, sizeof buf, "R%d/cps%d:%zu ", R, 1 << cl, v.size());
      desc += buf;
    }
  std::vector<int> sys2;
  std::vector<RagItem> items2;
  for (int R = kRagR2max; R >= kRagR2min; R--) {
    const std::vector<int> &v = bins2[R - kRagR2min];
    if (v.empty()) continue;
    for (size_t i = 0; i < v.size(); i++) {
      RagItem it;
      it.cps_log2 = 9;
      it.first = (int)(sys2.size() + i);
      it.count = 1;
      it.rows = R;
      items2.push_back(it);
    }
    sys2.insert(sys2.end(), v.begin(), v.end());
    char buf[96];
    snprintf(buf, sizeof buf, "R%d/cps512:%zu ", R, v.size());
    desc += buf;
  }
  if (!longs.empty()) desc += "long(>8192):" + std::to_string(longs.size()) + " ";
  if (!p->solo.empty()) desc += "solo:" + std::to_string(p->solo.size());
  p->text = desc;
  p->nitems = (int)items.size();
  p->nlong = (int)longs.size();
  p->long_ws = lws;
  const size_t ts = p->dbl ? 8 : 4;
  // span-group regions (at most 512 groups of G*R + 2E elements, G >= 1);
  // the LDS reduced solve reuses them (6 values per thread)
  p->lds = (size_t)kRagThreads * (rmax + 2 * (16 / ts)) * ts;
  p->lds2 = (size_t)kRagThreads * (kRagR2max + 2 * (16 / ts)) * ts;
  p->nitems2 = (int)items2.size();
  auto up = [&](void **dst, const void *src, size_t bytes) -> tridStatus_t {
    if (bytes == 0) return TRID_STATUS_SUCCESS;
    tridStatus_t r = check_hip(hipMalloc(dst, bytes), "hipMalloc");
    if (r) return r;
    return check_hip(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
  };
  s = up((void **)&p->sys, sys.data(), sys.size() * sizeof(int));
  if (!s) s = up((void **)&p->items, items.data(), items.size() * sizeof(RagItem));
  if (!s) s = up((void **)&p->sys2, sys2.data(), sys2.size() * sizeof(int));
  if (!s) s = up((void **)&p->items2, items2.data(), items2.size() * sizeof(RagItem));
  if (!s) s = up((void **)&p->long_ids, longs.data(), longs.size() * sizeof(int));
  if (!s) s = up((void **)&p->long_ws_off, long_off.data(), long_off.size() * sizeof(long long));
  if (!s) {
    unsigned zero[2] = {0, 0};
    s = up((void **)&p->counter, zero, sizeof zero);
  }
  if (s) {
    free_plan(p);
    return s;
  }
  // persistent grid: as many work-groups as fit on the device at once
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  int per_cu = 0, per_cu2 = 0;
  if (p->dbl) {
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k_ragged<double, kRagRmin, kRagRmax>, kRagThreads, p->lds);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu2, k_ragged<double, kRagR2min, kRagR2max>, kRagThreads, p->lds2);
  } else {
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k_ragged<float, kRagR2min, kRagR2max>, kRagThreads, p->lds);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu2, k_ragged<float, kRagR2min, kRagR2max>, kRagThreads, p->lds2);
  }
  per_cu = std::max(1, per_cu);
  per_cu2 = std::max(1, per_cu2);
  p->grid = std::max(1, std::min(p->nitems, cus * per_cu));
  p->grid2 = std::max(1, std::min(p->nitems2, cus * per_cu2));
  *plan = p;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridRaggedPlanDestroy(tridRaggedPlan_t plan) {
  free_plan(plan);
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridDgtsvRaggedBatch(tridRaggedPlan_t plan,
                                           const double *a, const double *b,
                                           const double *c, double *d,
                                           double *u, int inc, int sync) {
  return ragged_solve<double>(plan, a, b, c, d, u, inc, sync);
}

EXTERN_C tridStatus_t tridSgtsvRaggedBatch(tridRaggedPlan_t plan,
                                           const float *a, const float *b,
                                           const float *c, float *d, float *u,
                                           int inc, int sync) {
  return ragged_solve<float>(plan, a, b, c, d, u, inc, sync);
}

EXTERN_C tridStatus_t tridRaggedPlanExtent(tridRaggedPlan_t plan, long long *elements) {
  clear_error();
  if (!plan || !elements) return fail(TRID_STATUS_INVALID_VALUE, "plan or elements is NULL");
  *elements = plan->extent;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C const char *tridRaggedPlanDescribe(tridRaggedPlan_t plan) {
  return plan ? plan->text.c_str() : "";
}
