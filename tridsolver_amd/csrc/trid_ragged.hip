// trid_ragged.hip -- ragged batches in one persistent launch
// (include/trid_ragged.h).
//
// Every system is solved with the partitioned Thomas scheme of
// trid_part.hpp: chunks of R rows per lane (Thomas within the chunk), the
// chunks' reduced system solved across lanes.  The chunk length R and the
// number of chunks per system (cps, 1..512) are chosen per length class.  A
// work item is one work-group of 512 threads processing 512/cps systems of
// one bin.  Loads are staged per span group through LDS: the lanes of a
// system (or of one wave's 64 chunks of a longer system) read its contiguous
// span as 16-byte pieces, then each lane takes its chunk from LDS (the
// transpose of the solvedim-0 path, for spans at arbitrary offsets and
// alignments).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/trid_hip.h"
#include "../../include/trid_hip_ext.h"
#include "../../include/trid_ragged.h"
#include "trid_launch.hpp"
#include "trid_part.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

constexpr int kRagThreads = 512;
// Two row families, each a launch of the persistent kernel with its own
// instantiation (registers are those of the family's longest chunk):
//   family 0: chunks of 5..8 rows   (512 chunks cover 4096 rows)
//   family 1: chunks of 9..16 rows  (512 chunks cover 8192 rows)
// A bin is (cps, R): cps a power of two (1..512) and R the shortest row count
// of the family with cps * R >= L, so that at least 4/5 (family 1: 8/9) of a
// system's lanes hold rows.
constexpr int kRagRlo[2] = {5, 9}, kRagRhi[2] = {8, 16};
constexpr int kRagMaxLen[2] = {kRagThreads * 8, kRagThreads * 16};
// systems past kRagMaxLen[1] rows leave the persistent kernel: the longest
// of them are solved one by one as strided batches of one system, the rest
// by a one-thread-per-system kernel.  Untuned cost model, measured on
// MI355X (FP64): one strided one-system solve ~kSoloUs (27 us for 9000 rows,
// 42 us for 2^20 rows: launches and latency); the one-thread kernel
// ~kThreadRowUs per row of its longest system (a dependent chain of global
// accesses).  The autotuner measures both on the plan's own lengths.
constexpr double kSoloUs = 25.0;
constexpr double kThreadRowUs = 0.5;
// length classes: class c < 10 holds the systems whose family-0 bin has
// 2^c chunks (8 * 2^(c-1) < L <= 8 * 2^c); class 10 holds 4097..8192 rows,
// which only family 1 takes
constexpr int kRagClasses = 11;

struct RagItem {
  int cps_log2;  // chunks per system
  int first;     // first entry of the item in the sorted system list
  int count;     // systems in this item (<= kRagThreads / cps)
  int rows;      // chunk length R of the item's bin (5..8, or 9..16)
  int red;       // reduced solve past 64 chunks: 0 = PCR through LDS (one
                 // barrier per level), 1 = PCR in each wave, then a serial
                 // sweep across the waves (three barriers)
  int pad_;
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
    t[j] = ld_stream(g + (pc < np ? pc : last));
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

// rag_load for a span group of a whole wave (G = 64), straight into LDS:
// global_load_lds moves each lane's 16-byte piece to reg + (round * 64 +
// lane) pieces -- the layout rag_load builds -- without passing through
// registers, so the loads of several arrays can be in flight at once.  The
// caller waits (rag_lds_wait) before reading reg.  Non-temporal, as
// ld_stream.  Returns sh.
template <typename T, int R>
__device__ __forceinline__ int rag_load_lds(const T *__restrict__ p, int len, T *__restrict__ reg,
                                            int lane) {
  using V = typename Vec16<T>::type;
  constexpr int E = Vec16<T>::n;
  const int sh = (int)((reinterpret_cast<uintptr_t>(p) / sizeof(T)) & (E - 1));
  const V *g = reinterpret_cast<const V *>(p - sh);
  const int np = len > 0 ? (sh + len + E - 1) / E : 0;
  constexpr int NR = (R + E - 1) / E + 1;
#pragma unroll
  for (int j = 0; j < NR; j++) {
    const int pc = j * kWave + lane;
    if (pc < np)
      __builtin_amdgcn_global_load_lds(
          (const void *)(g + pc),
          (__attribute__((address_space(3))) void *)(reg + j * kWave * E), 16, 0,
          2 /* non-temporal */);
  }
  return sh;
}

// Waits for the wave's loads into LDS (and every earlier access).
__device__ __forceinline__ void rag_lds_wait() {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
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
        if (add) w += ld_stream(g + pc);
        st_stream(g + pc, w);
      } else {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int i = lo + e;
          if (i >= 0 && i < len) {
            T v = w[e];
            if (add) v += ld_stream(p + i);
            st_stream(p + i, v);
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
  if (G == kWave) {
    // whole-wave span groups (systems of 64 chunks and more): two arrays
    // in flight at a time, straight into LDS (the second array's regions
    // follow the eight waves' first ones)
    T *reg2 = smem + (kRagThreads / kWave) * (kWave * R + 2 * E) + (tid / kWave) * (kWave * R + 2 * E);
    rag_lds_wait();  // the previous item's reads of these regions are done
    int sa = rag_load_lds<T, R>(g.a + p0, len, reg, kg);
    int sb = rag_load_lds<T, R>(g.b + p0, len, reg2, kg);
    rag_lds_wait();
    rag_take<T, R>(reg, sa, i0, len, a, T(0));
    rag_take<T, R>(reg2, sb, i0, len, b, T(1));
    rag_lds_wait();
    sa = rag_load_lds<T, R>(g.c + p0, len, reg, kg);
    sb = rag_load_lds<T, R>(g.d + p0, len, reg2, kg);
    rag_lds_wait();
    rag_take<T, R>(reg, sa, i0, len, c, T(0));
    rag_take<T, R>(reg2, sb, i0, len, d, T(0));
  } else {
    int sh = rag_load<T, R>(g.a + p0, len, reg, kg, G);
    rag_take<T, R>(reg, sh, i0, len, a, T(0));
    sh = rag_load<T, R>(g.b + p0, len, reg, kg, G);
    rag_take<T, R>(reg, sh, i0, len, b, T(1));
    sh = rag_load<T, R>(g.c + p0, len, reg, kg, G);
    rag_take<T, R>(reg, sh, i0, len, c, T(0));
    sh = rag_load<T, R>(g.d + p0, len, reg, kg, G);
    rag_take<T, R>(reg, sh, i0, len, d, T(0));
  }
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
    // the reduced solve's LDS (PCR: 6 values per thread; per wave: 12
    // values per wave of a system) reuses the staging region once every
    // wave is done with it
    __syncthreads();
    if (it.red)
      reduced_multi<T>(f, l, k, cps, slot, smem, xf, xl);
    else
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

// Systems longer than kRagMaxLen[1]: one thread per system, forward sweep's
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

// Synthetic diagonally dominant rows for the autotuner's timing runs.
template <typename T>
__global__ void k_rag_fill(T *a, T *b, T *c, T *d, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] = T(-0.25);
  b[i] = T(1);
  c[i] = T(-0.25);
  d[i] = T((i % 13) - 6) / T(8);
}

}  // namespace tridhip

using namespace tridhip;

namespace {

// One launch of the persistent kernel: the items of one row family.
struct RagLaunch {
  int *sys = nullptr;
  RagItem *items = nullptr;
  int nitems = 0;
  int grid = 0;
  size_t lds = 0;
};

// Row family and reduced solve of one length class.
struct RagChoice {
  int fam = 0;
  int red = 1;
};

}  // namespace

struct tridRaggedPlan_st {
  int dbl = 1;
  long long nsys = 0;
  long long extent = 0;  // offsets[nsys]: elements every array must hold
  const int64_t *offsets = nullptr;  // caller's device array
  unsigned *counter = nullptr;  // both launches' (each resets it at its end)
  RagLaunch launch[2];          // family 0, family 1
  int *long_ids = nullptr;
  long long *long_ws_off = nullptr;
  int nlong = 0;
  long long long_ws = 0;  // elements
  // the longest systems, each solved as a one-system strided batch (split-line
  // solve past 8192 rows): (offset, length)
  std::vector<std::pair<long long, int>> solo;
  std::string text;
};

namespace {

std::mutex g_rag_mu;
RagChoice g_rag_forced{-1, -1};  // tridRaggedForceChoice

void free_launch(RagLaunch &l) {
  (void)hipFree(l.sys);
  (void)hipFree(l.items);
  l = RagLaunch{};
}

void free_plan(tridRaggedPlan_st *p) {
  if (!p) return;
  free_launch(p->launch[0]);
  free_launch(p->launch[1]);
  (void)hipFree(p->counter);
  (void)hipFree(p->long_ids);
  (void)hipFree(p->long_ws_off);
  delete p;
}

int ceil_log2(long long v) {
  int l = 0;
  while ((1LL << l) < v) l++;
  return l;
}

// Bin of a system of L rows in family f: cps = 2^cl, R rows per chunk.
int bin_of(long long L, int f, int &R) {
  const int cl = ceil_log2((L + kRagRhi[f] - 1) / kRagRhi[f]);
  R = (int)std::max<long long>(kRagRlo[f], (L + (1LL << cl) - 1) >> cl);
  return cl;
}

int class_of(long long L) {
  if (L > kRagMaxLen[0]) return 10;
  int R;
  return bin_of(L, 0, R);
}

bool fam_ok(int c, int f) { return f == 1 || c < 10; }

// The reduced-solve choice matters only past 64 chunks per system.
bool red_matters(int c, int f) {
  const long long Lmax = c < 10 ? 8LL << c : kRagMaxLen[1];
  int R;
  return bin_of(Lmax, f, R) > 6;
}

RagChoice default_choice(int c, bool dbl) {
  // measured on MI355X over 200,000 systems of 8..4096 rows: FP64 runs
  // fastest with family 0 up to 4096 rows (family 1 past it) and PCR
  // through LDS (1.26 ms against 1.31-1.35 ms for the other three
  // combinations); FP32 (half the registers per row) with family 1
  // throughout and the per-wave reduced solve (0.68 ms against 0.69-0.78)
  RagChoice ch;
  ch.fam = dbl && c < 10 ? 0 : 1;
  ch.red = dbl ? 0 : 1;
  return ch;
}

// Systems of each row family and the reduced solve of each chunk count,
// from the per-class choices.
void family_lists(const std::vector<std::vector<int>> &cls, const RagChoice *choice, int f,
                  std::vector<int> &ids, int *red) {
  ids.clear();
  for (int i = 0; i < 10; i++) red[i] = 0;
  for (int c = 0; c < kRagClasses; c++) {
    if (choice[c].fam != f) continue;
    ids.insert(ids.end(), cls[c].begin(), cls[c].end());
    // the items of 2^cl chunks take the reduced solve of the class they come
    // from (in family 1, a class's systems take half its chunk count)
    const long long Lmax = c < 10 ? 8LL << c : kRagMaxLen[1];
    int R;
    const int cl = bin_of(Lmax, f, R);
    if (cl > 6) red[cl] = choice[c].red;
  }
  std::sort(ids.begin(), ids.end());
}

// Work items of the systems `ids` (lengths from `off`) in family f, appended
// to sys/items, most expensive first (more chunks, then longer chunks).
// red[cl]: reduced solve of the items with 2^cl chunks.
void build_items(const std::vector<int64_t> &off, const std::vector<int> &ids, int f,
                 const int *red, std::vector<int> &sys, std::vector<RagItem> &items,
                 std::string *desc) {
  const int nR = kRagRhi[f] - kRagRlo[f] + 1;
  std::vector<std::vector<int>> bins(10 * nR);
  for (int id : ids) {
    int R;
    const int cl = bin_of(off[id + 1] - off[id], f, R);
    bins[cl * nR + (R - kRagRlo[f])].push_back(id);
  }
  for (int cl = 9; cl >= 0; cl--)
    for (int R = kRagRhi[f]; R >= kRagRlo[f]; R--) {
      const std::vector<int> &v = bins[cl * nR + (R - kRagRlo[f])];
      if (v.empty()) continue;
      const int per = std::max(1, kRagThreads >> cl);
      for (size_t i = 0; i < v.size(); i += per) {
        RagItem it;
        it.cps_log2 = cl;
        it.first = (int)(sys.size() + i);
        it.count = (int)std::min<size_t>(per, v.size() - i);
        it.rows = R;
        it.red = cl > 6 ? red[cl] : 0;
        it.pad_ = 0;
        items.push_back(it);
      }
      sys.insert(sys.end(), v.begin(), v.end());
      if (desc) {
        char buf[96];
        snprintf(buf, sizeof buf, "R%d/cps%d:%zu ", R, 1 << cl, v.size());
        *desc += buf;
      }
    }
}

template <typename T, int RLO, int RHI>
void occupancy(size_t lds, int &per_cu) {
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ragged<T, RLO, RHI>,
                                                      kRagThreads, lds);
  per_cu = std::max(1, per_cu);
}

size_t family_lds(int f, size_t ts) {
  // span-group regions (at most 512 groups of G*R + 2E elements, G >= 1),
  // or, for whole-wave groups, two sets of eight wave regions of 64R + 2E;
  // the reduced solves reuse them
  const size_t E = 16 / ts, R = kRagRhi[f];
  const size_t groups = (size_t)kRagThreads * (R + 2 * E);
  const size_t waves2 = 2 * (size_t)(kRagThreads / kWave) * (kWave * R + 2 * E);
  return std::max(groups, waves2) * ts;
}

int family_grid(int f, bool dbl, int nitems) {
  int dev = 0, cus = 256, per_cu = 1;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t lds = family_lds(f, dbl ? 8 : 4);
  if (dbl) {
    if (f == 0) occupancy<double, 5, 8>(lds, per_cu);
    else occupancy<double, 9, 16>(lds, per_cu);
  } else {
    if (f == 0) occupancy<float, 5, 8>(lds, per_cu);
    else occupancy<float, 9, 16>(lds, per_cu);
  }
  return std::max(1, std::min(nitems, cus * per_cu));
}

template <typename T>
hipError_t launch_family(int f, const RagLaunch &l, RagArgs<T> g, hipStream_t st) {
  if (l.nitems == 0) return hipSuccess;
  g.sys = l.sys;
  g.items = l.items;
  g.nitems = l.nitems;
  if (f == 0)
    hipLaunchKernelGGL((k_ragged<T, 5, 8>), dim3(l.grid), dim3(kRagThreads), l.lds, st, g);
  else
    hipLaunchKernelGGL((k_ragged<T, 9, 16>), dim3(l.grid), dim3(kRagThreads), l.lds, st, g);
  return hipGetLastError();
}

tridStatus_t upload(void **dst, const void *src, size_t bytes) {
  if (bytes == 0) return TRID_STATUS_SUCCESS;
  tridStatus_t r = check_hip(hipMalloc(dst, bytes), "hipMalloc");
  if (r) return r;
  return check_hip(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
}

tridStatus_t make_launch(RagLaunch &l, int f, bool dbl, const std::vector<int> &sys,
                         const std::vector<RagItem> &items) {
  l.nitems = (int)items.size();
  l.lds = family_lds(f, dbl ? 8 : 4);
  l.grid = family_grid(f, dbl, l.nitems);
  tridStatus_t s = upload((void **)&l.sys, sys.data(), sys.size() * sizeof(int));
  if (!s) s = upload((void **)&l.items, items.data(), items.size() * sizeof(RagItem));
  return s;
}

// ---------------------------------------------------------------------------
// autotuner
//
// For every length class present, the candidate choices (row family, and
// past 64 chunks per system the reduced solve) are timed on a synthetic
// batch of the class's own lengths (cycled until it holds 2^22 unknowns, so
// that the persistent grid runs several rounds of items) and the fastest is
// kept.  For systems past 8192 rows the two cost constants of the
// solo / one-thread split are measured on the plan's longest system.  The
// timing runs use buffers of their own; the caller's arrays are not touched.

struct TuneBufs {
  void *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  int64_t *off = nullptr;
  unsigned *counter = nullptr;
  ~TuneBufs() {
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(c);
    (void)hipFree(d);
    (void)hipFree(off);
    (void)hipFree(counter);
  }
};

template <typename T>
tridStatus_t alloc_synthetic(TuneBufs &tb, long long n, hipStream_t st) {
  const size_t bytes = (size_t)std::max(1LL, n) * sizeof(T);
  tridStatus_t s = check_hip(hipMalloc(&tb.a, bytes), "hipMalloc");
  if (!s) s = check_hip(hipMalloc(&tb.b, bytes), "hipMalloc");
  if (!s) s = check_hip(hipMalloc(&tb.c, bytes), "hipMalloc");
  if (!s) s = check_hip(hipMalloc(&tb.d, bytes), "hipMalloc");
  if (s) return s;
  hipLaunchKernelGGL(k_rag_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     static_cast<T *>(tb.a), static_cast<T *>(tb.b), static_cast<T *>(tb.c),
                     static_cast<T *>(tb.d), n);
  return check_hip(hipGetLastError(), "fill");
}

// Time of one solve with both family launches (best of `passes` x `reps`
// launches after a warm-up), in ms; < 0 on a launch failure.
template <typename T>
float time_launches(const RagLaunch *ls, const RagArgs<T> &g, hipStream_t st, int reps = 3,
                    int passes = 2) {
  auto run = [&]() {
    hipError_t e = launch_family<T>(0, ls[0], g, st);
    if (e == hipSuccess) e = launch_family<T>(1, ls[1], g, st);
    return e;
  };
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1.f;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1.f;
  }
  float best = -1.f;
  if (run() == hipSuccess) {
    for (int pass = 0; pass < passes; pass++) {
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < reps; r++) (void)run();
      (void)hipEventRecord(e1, st);
      if (hipEventSynchronize(e1) != hipSuccess) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= reps;
      if (best < 0.f || ms < best) best = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipGetLastError();
  return best;
}

// The plan autotuner.  The choices are timed in context: a synthetic batch
// with the plan's own lengths (every k-th system of each class when the
// batch holds more than kTuneMaxElems unknowns) is solved with both family
// launches built from a whole choice vector -- per-class runs in isolation
// mispredict the mixed launch (their tails and occupancy differ).  Starting
// from the untuned choices, the classes are visited from the most to the
// fewest unknowns, and each alternative (other row family; past 64 chunks
// the other reduced solve) is kept when the whole solve gets faster by more
// than 2 %.  The timing runs use buffers of their own; the caller's arrays
// are not touched.
constexpr long long kTuneMaxElems = 1LL << 27;

template <typename T>
tridStatus_t tune_plan(const std::vector<int64_t> &off, const std::vector<std::vector<int>> &cls,
                       bool dbl, RagChoice *choice, std::string &log, hipStream_t st) {
  long long total = 0;
  std::vector<long long> unk(kRagClasses, 0);
  for (int c = 0; c < kRagClasses; c++)
    for (int id : cls[c]) unk[c] += off[id + 1] - off[id];
  for (long long u : unk) total += u;
  if (total == 0) return TRID_STATUS_SUCCESS;
  const long long stride = std::max(1LL, (total + kTuneMaxElems - 1) / kTuneMaxElems);
  // the synthetic batch: sampled systems of every class, contiguous
  std::vector<int64_t> soff{0};
  std::vector<std::vector<int>> scls(kRagClasses);
  for (int c = 0; c < kRagClasses; c++)
    for (size_t i = 0; i < cls[c].size(); i += stride) {
      const int id = cls[c][i];
      scls[c].push_back((int)soff.size() - 1);
      soff.push_back(soff.back() + (off[id + 1] - off[id]));
    }
  TuneBufs tb;
  tridStatus_t s = alloc_synthetic<T>(tb, soff.back(), st);
  if (!s) s = upload((void **)&tb.off, soff.data(), soff.size() * sizeof(int64_t));
  unsigned zero[2] = {0, 0};
  if (!s) s = upload((void **)&tb.counter, zero, sizeof zero);
  if (s) return s;
  const RagArgs<T> g{static_cast<const T *>(tb.a), static_cast<const T *>(tb.b),
                     static_cast<const T *>(tb.c), static_cast<T *>(tb.d), nullptr, 0,
                     tb.off, nullptr, nullptr, 0, tb.counter};
  auto eval = [&](const RagChoice *ch, float &ms) -> tridStatus_t {
    RagLaunch ls[2];
    tridStatus_t r = TRID_STATUS_SUCCESS;
    for (int f = 0; f < 2 && !r; f++) {
      std::vector<int> ids, sys;
      std::vector<RagItem> items;
      int red[10];
      family_lists(scls, ch, f, ids, red);
      build_items(soff, ids, f, red, sys, items, nullptr);
      r = make_launch(ls[f], f, dbl, sys, items);
    }
    if (!r) {
      ms = time_launches<T>(ls, g, st);
      if (ms < 0.f) r = fail(TRID_STATUS_EXECUTION_FAILED, "ragged autotuner: launch failed");
    }
    free_launch(ls[0]);
    free_launch(ls[1]);
    return r;
  };
  float best = 0.f;
  s = eval(choice, best);
  if (s) return s;
  char buf[96];
  snprintf(buf, sizeof buf, "sample 1/%lld, untuned %.1fus;", stride, 1000.f * best);
  log += buf;
  std::vector<int> order;
  for (int c = 0; c < kRagClasses; c++)
    if (unk[c] > 0) order.push_back(c);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return unk[x] > unk[y]; });
  for (int c : order) {
    std::vector<RagChoice> alts;
    for (int f = 0; f < 2; f++) {
      if (!fam_ok(c, f)) continue;
      for (int red = 0; red < 2; red++) {
        if (!red_matters(c, f) && red != choice[c].red) continue;
        if (f == choice[c].fam && red == choice[c].red) continue;
        alts.push_back({f, red});
      }
    }
    for (const RagChoice &alt : alts) {
      RagChoice trial[kRagClasses];
      std::copy(choice, choice + kRagClasses, trial);
      trial[c] = alt;
      float ms = 0.f;
      s = eval(trial, ms);
      if (s) return s;
      const bool take = ms < 0.98f * best;
      snprintf(buf, sizeof buf, " c%d:%d%s=%.1fus%s", c, alt.fam, alt.red ? "w" : "p",
               1000.f * ms, take ? "*" : "");
      log += buf;
      if (take) {
        best = ms;
        choice[c] = alt;
      }
    }
  }
  snprintf(buf, sizeof buf, "; tuned %.1fus", 1000.f * best);
  log += buf;
  return TRID_STATUS_SUCCESS;
}

// Measured cost constants of the solo / one-thread split on the plan's
// longest system of Lmax rows: (one strided solve, one-thread kernel per row).
template <typename T>
tridStatus_t tune_long(long long Lmax, double &solo_us, double &row_us, hipStream_t st) {
  TuneBufs tb;
  tridStatus_t s = alloc_synthetic<T>(tb, Lmax, st);
  int64_t off[2] = {0, Lmax};
  if (!s) s = upload((void **)&tb.off, off, sizeof off);
  if (s) return s;
  int id = 0;
  long long wso = 0;
  int *dids = nullptr;
  long long *dwso = nullptr;
  T *ws = nullptr;
  s = upload((void **)&dids, &id, sizeof id);
  if (!s) s = upload((void **)&dwso, &wso, sizeof wso);
  if (!s) s = check_hip(hipMalloc((void **)&ws, 2 * (size_t)Lmax * sizeof(T)), "hipMalloc");
  if (!s) {
    T *a = static_cast<T *>(tb.a), *b = static_cast<T *>(tb.b), *c = static_cast<T *>(tb.c),
      *d = static_cast<T *>(tb.d);
    RagArgs<T> g{a, b, c, d, nullptr, 0, tb.off, nullptr, nullptr, 0, nullptr};
    int dims[1] = {(int)Lmax}, opts[1] = {0};
    std::vector<float> tm;
    auto prep = [&](int) { return true; };
    auto run = [&](int k) -> hipError_t {
      if (k == 0) {
        tridStatus_t r;
        if constexpr (sizeof(T) == 8)
          r = tridDmtsvStridedBatch(a, b, c, d, nullptr, 1, 0, dims, dims, opts, 0);
        else
          r = tridSmtsvStridedBatch(a, b, c, d, nullptr, 1, 0, dims, dims, opts, 0);
        return r ? hipErrorUnknown : hipSuccess;
      }
      hipLaunchKernelGGL(k_ragged_long<T>, dim3(1), dim3(64), 0, st, g, dids, dwso, 1, ws);
      return hipGetLastError();
    };
    if (pick_fastest(2, prep, run, st, &tm, 3, 1) >= 0) {
      solo_us = 1000.0 * tm[0] / 3;
      row_us = 1000.0 * tm[1] / 3 / (double)Lmax;
    }
  }
  (void)hipFree(dids);
  (void)hipFree(dwso);
  (void)hipFree(ws);
  return s;
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
  RagArgs<T> g{a, b, c, d, u, inc ? 1 : 0, p->offsets, nullptr, nullptr, 0, p->counter};
  // the two launches are stream-ordered: the first one's last work-group
  // has reset the counter when the second starts
  for (int f = 0; f < 2; f++) {
    tridStatus_t s = check_hip(launch_family<T>(f, p->launch[f], g, st), "ragged kernel launch");
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

tridStatus_t create_plan(tridRaggedPlan_st *p, const std::vector<int64_t> &off) {
  const long long nsys = p->nsys;
  const bool dbl = p->dbl == 1;
  hipStream_t st = current_stream();
  RagChoice forced;
  {
    std::lock_guard<std::mutex> lk(g_rag_mu);
    forced = g_rag_forced;
  }
  const bool tune = autotune_mode() == 1 && forced.fam < 0 && forced.red < 0 && !capturing(st);
  std::string tlog;

  // systems past kRagMaxLen[1] rows: the k longest are solved one by one as
  // strided batches of one system (a few launches each, bandwidth-bound),
  // the rest together by the one-thread-per-system kernel (time set by its
  // longest system); k minimises k * solo + L_(k+1) * row
  std::vector<std::pair<long long, long long>> over;  // (length, system)
  std::vector<std::vector<int>> cls(kRagClasses);
  for (long long k = 0; k < nsys; k++) {
    const long long L = off[k + 1] - off[k];
    if (L == 0) continue;
    if (L > kRagMaxLen[1]) {
      over.push_back({L, k});
      continue;
    }
    cls[class_of(L)].push_back((int)k);
  }
  std::sort(over.begin(), over.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
  for (const auto &o : over)
    if (o.first > INT32_MAX) return fail(TRID_STATUS_INVALID_VALUE, "system longer than 2^31 - 1 rows");
  double solo_us = kSoloUs, row_us = kThreadRowUs;
  if (tune && !over.empty()) {
    tridStatus_t s = dbl ? tune_long<double>(over[0].first, solo_us, row_us, st)
                         : tune_long<float>(over[0].first, solo_us, row_us, st);
    if (s) return s;
    char buf[96];
    snprintf(buf, sizeof buf, "long:solo=%.1fus,row=%.3fus ", solo_us, row_us);
    tlog += buf;
  }
  size_t nsolo = 0;
  {
    double best = 1e300;
    for (size_t k = 0; k <= over.size(); k++) {
      const double t = k * solo_us + (k < over.size() ? over[k].first * row_us : 0.0);
      if (t < best) {
        best = t;
        nsolo = k;
      }
    }
  }
  std::vector<int> longs;
  std::vector<long long> long_off;
  long long lws = 0;
  for (size_t k = 0; k < over.size(); k++) {
    if (k < nsolo) {
      p->solo.push_back({off[over[k].second], (int)over[k].first});
    } else {
      longs.push_back((int)over[k].second);
      long_off.push_back(lws);
      lws += 2 * over[k].first;
    }
  }
  // keep the one-thread systems in offset order
  {
    std::vector<size_t> perm(longs.size());
    for (size_t i = 0; i < perm.size(); i++) perm[i] = i;
    std::sort(perm.begin(), perm.end(), [&](size_t x, size_t y) { return longs[x] < longs[y]; });
    std::vector<int> l2;
    std::vector<long long> o2;
    lws = 0;
    for (size_t i : perm) {
      l2.push_back(longs[i]);
      o2.push_back(lws);
      lws += 2 * (off[longs[i] + 1] - off[longs[i]]);
    }
    longs.swap(l2);
    long_off.swap(o2);
  }

  // a row family and reduced solve per length class
  RagChoice choice[kRagClasses];
  for (int c = 0; c < kRagClasses; c++) {
    choice[c] = default_choice(c, dbl);
    if (forced.fam >= 0 && fam_ok(c, forced.fam)) choice[c].fam = forced.fam;
    if (forced.red >= 0) choice[c].red = forced.red;
  }
  if (tune) {
    tridStatus_t s = dbl ? tune_plan<double>(off, cls, dbl, choice, tlog, st)
                         : tune_plan<float>(off, cls, dbl, choice, tlog, st);
    if (s) return s;
  }

  // items of each family; a family's reduced-solve choice is per chunk count
  std::string desc;
  for (int f = 0; f < 2; f++) {
    std::vector<int> ids;
    int red[10];
    family_lists(cls, choice, f, ids, red);
    std::vector<int> sys;
    std::vector<RagItem> items;
    build_items(off, ids, f, red, sys, items, &desc);
    tridStatus_t s = make_launch(p->launch[f], f, dbl, sys, items);
    if (s) return s;
  }
  if (!longs.empty()) desc += "long(>8192):" + std::to_string(longs.size()) + " ";
  if (!p->solo.empty()) desc += "solo:" + std::to_string(p->solo.size()) + " ";
  {
    std::string cs = "classes:";
    for (int c = 0; c < kRagClasses; c++)
      if (!cls[c].empty()) {
        char buf[48];
        snprintf(buf, sizeof buf, " c%d=%d%s", c, choice[c].fam, choice[c].red ? "w" : "p");
        cs += buf;
      }
    desc += cs;
  }
  if (tune) desc += " tuned: " + tlog;
  p->text = desc;
  p->nlong = (int)longs.size();
  p->long_ws = lws;
  tridStatus_t s = upload((void **)&p->long_ids, longs.data(), longs.size() * sizeof(int));
  if (!s) s = upload((void **)&p->long_ws_off, long_off.data(), long_off.size() * sizeof(long long));
  if (!s) {
    unsigned zero[2] = {0, 0};
    s = upload((void **)&p->counter, zero, sizeof zero);
  }
  return s;
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
  s = create_plan(p, off);
  if (s) {
    free_plan(p);
    return s;
  }
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

EXTERN_C tridStatus_t tridRaggedForceChoice(int family, int red) {
  clear_error();
  if (family < -1 || family > 1 || red < -1 || red > 1)
    return fail(TRID_STATUS_INVALID_VALUE, "family and red must be -1, 0 or 1");
  std::lock_guard<std::mutex> lk(g_rag_mu);
  g_rag_forced = RagChoice{family, red};
  return TRID_STATUS_SUCCESS;
}
