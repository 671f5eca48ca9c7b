// trid_scalar.hip -- strided-batch scalar tridiagonal solves: kernels, the
// per-call dispatcher and autotuner, and the C entry points of trid_hip.h.
//
// Host dispatch replaces tridMultiDimBatchSolve<REAL,INC>
// (scalar/src/cuda/trid_cuda.cu:191-445): instead of a kernel per solvedim
// and opts value plus cuBLAS transposes for solvedim 0, every layout goes to
// one partitioned-Thomas kernel family (trid_part.hpp) whose lane mapping is
// chosen by the layout class:
//   solvedim 0  (unit stride along the system)  -> ORDER 0, chunks in lanes
//   solvedim >= 1 (systems adjacent in memory)  -> ORDER 1, systems in lanes
// Systems longer than one work-group holds (8192 rows) take the split-line
// path below, which runs the same kernels on segments of the systems; a
// one-thread-per-system Thomas kernel remains as a forced configuration.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/trid_hip.h"
#include "../../include/trid_hip_ext.h"
#include "trid_launch.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

// ---------------------------------------------------------------------------
// kernels

// Fallback: one thread per system, plain Thomas with the forward sweep's
// c' and d' in a global workspace laid out [row][system] (coalesced across
// threads).  Same recurrence as the reference's trid_strided_multidim_kernel
// (scalar/src/cuda/trid_strided_multidim.hpp:43-87) without the N_MAX cap.
template <typename T, int INC>
__global__ void __launch_bounds__(256)
    k_thomas_gm(const T *__restrict__ A, const T *__restrict__ B,
                const T *__restrict__ C, T *__restrict__ Dd,
                T *__restrict__ U, SysDesc D, T *__restrict__ ws) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= D.nsys) return;
  long long off[5];
  sys_offsets<long long>(D, q, off);
  const int N = D.N;
  const long long ns = D.nsys;
  T *cs = ws;
  T *ds = ws + (long long)N * ns;
  const T *pa = A + off[0], *pb = B + off[1], *pc = C + off[2];
  T *pd = Dd + off[3];
  T *pu = INC ? U + off[4] : nullptr;
  const long long sa = D.sstr[0], sb = D.sstr[1], sc = D.sstr[2],
                  sd = D.sstr[3], su = D.sstr[4];
  T bb = rcp(pb[0]);
  T cc = N > 1 ? bb * pc[0] : T(0);
  T dd = bb * pd[0];
  cs[q] = cc;
  ds[q] = dd;
  for (int i = 1; i < N; i++) {
    T aa = pa[i * sa];
    bb = pb[i * sb] - aa * cc;
    dd = pd[i * sd] - aa * dd;
    bb = rcp(bb);
    cc = i < N - 1 ? bb * pc[i * sc] : T(0);
    dd = bb * dd;
    cs[i * ns + q] = cc;
    ds[i * ns + q] = dd;
  }
  if (INC)
    pu[(N - 1) * su] += dd;
  else
    pd[(N - 1) * sd] = dd;
  for (int i = N - 2; i >= 0; i--) {
    dd = ds[i * ns + q] - cs[i * ns + q] * dd;
    if (INC)
      pu[i * su] += dd;
    else
      pd[i * sd] = dd;
  }
}

// Applies the solver's reciprocal to n values (diagnostic entry point).
template <typename T>
__global__ void k_rcp(const T *__restrict__ x, T *__restrict__ y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = rcp(x[i]);
}

// ---------------------------------------------------------------------------
// host side

namespace {

template <typename T, int INC>
hipError_t launch_fallback(const T *a, const T *b, const T *c, T *d, T *u,
                           const SysDesc &D, hipStream_t st, T *ws) {
  const long long blocks = (D.nsys + 255) / 256;
  hipLaunchKernelGGL((k_thomas_gm<T, INC>), dim3((unsigned)blocks), dim3(256),
                     0, st, a, b, c, d, u, D, ws);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// split-line solves: systems longer than one work-group holds
//
// Every system is cut into P segments of L = R * cps rows (the last one may
// be shorter).  The solve takes three steps, each a batched launch over all
// segments of all systems:
//   1. segment reduction (k_seg MODE 1): each work-group reduces its
//      segments (partitioned Thomas on the chunks, then a merge tree over the
//      chunks' rows) to the segment's two rows in its first and last
//      unknowns, coupled to the neighbouring segments -- the reduced system
//      of 2P rows per system;
//   2. the reduced systems are solved by the partitioned kernel (or, when
//      2P is itself too long, by a split-line solve of their own);
//   3. segment solve (k_seg MODE 2): every segment is solved as an
//      independent system, its couplings to the neighbouring segments moved
//      to the right-hand side with their solved boundary unknowns.
// d is only read until step 3 writes the solution (Inc: never written).
// Global traffic: a, b, c, d read twice and the solution written once (9
// values per unknown against 5 for one-pass solves), plus 8 reduced values
// per segment.  The reference solves the same reduced structure across MPI
// ranks (scalar/src/cpu/trid_mpi_cpu.hpp:139-328: modified Thomas per rank,
// reduced system of two rows per rank, then a backward pass per rank); here
// the ranks are segments of one GPU's batch.

inline long long round4(long long v) { return (v + 3) & ~3LL; }

// Workspace (elements) of a split-line solve, including its recursive
// reduced solves.
inline long long split_ws_elems(const PartCfg &cfg, int N, long long nsys, size_t tsize) {
  const long long L = (long long)cfg.R << cfg.cps_log2;
  const long long P = (N + L - 1) / L;
  const long long n2 = round4(2 * P * nsys);
  long long need = 4 * n2;
  if (heuristic(cfg.order, (int)(2 * P), 1, tsize).kind != 1)
    need += split_ws_elems(split_cfg(cfg.order, 1, tsize, nsys), (int)(2 * P), nsys, tsize);
  return need;
}

template <typename T, int INC>
hipError_t split_solve(const PartCfg &cfg, const T *a, const T *b, const T *c, T *d,
                       T *u, const SysDesc &D, hipStream_t st, T *ws) {
  const int order = cfg.order;
  const int L = cfg.R << cfg.cps_log2;
  const int N = D.N;
  const int P = (N + L - 1) / L, Pf = N / L;
  const int Nt = N - Pf * L;
  const long long ns = D.nsys;
  const long long n2 = round4(2LL * P * ns);
  T *ra = ws, *rb = ws + n2, *rc = ws + 2 * n2, *rd = ws + 3 * n2;
  T *next = ws + 4 * n2;

  SegOut<T> so;
  so.ra = ra;
  so.rb = rb;
  so.rc = rc;
  so.rd = rd;
  so.nsys = ns;
  so.P = P;
  so.Pf = Pf > 0 ? Pf : 1;
  so.rs_row = order == 0 ? 1 : ns;
  so.rs_sys = order == 0 ? 2LL * P : 1;

  // the full segments as one batch: the segment index is one more non-solve
  // dimension (fastest for ORDER 0, so that a system's segments are adjacent
  // work items; slowest for ORDER 1, which keeps systems in adjacent lanes)
  SysDesc Ds = D;
  Ds.N = L;
  if (Pf > 0) {
    const int slot = order == 0 ? 0 : D.nother;
    for (int k = D.nother; k > slot; k--) {
      Ds.osize[k] = D.osize[k - 1];
      for (int t = 0; t < 5; t++) Ds.ostr[t][k] = D.ostr[t][k - 1];
    }
    Ds.osize[slot] = Pf;
    for (int t = 0; t < 5; t++) Ds.ostr[t][slot] = (long long)L * D.sstr[t];
    Ds.nother = D.nother + 1;
    Ds.nsys = ns * Pf;
    Ds.small_index = Ds.nsys < (1LL << 31) && !wide_index();
  }
  PartCfg cs = cfg;
  cs.kind = 1;
  cs.xcd = 0;
  if (cs.vec == 2) cs.vec = 1;
  // wave-staged access: the segments of one wave must be adjacent in memory
  if (cs.vec == 1 && order == 0 && (1 << cs.cps_log2) <= kWave &&
      Pf % (kWave >> cs.cps_log2) == 0)
    cs.vec = 2;

  // the last, shorter segment of every system (rows Pf*L .. N-1)
  SysDesc Dt = D;
  Dt.N = Nt;
  PartCfg ct;
  const long long sh[5] = {(long long)Pf * L * D.sstr[0], (long long)Pf * L * D.sstr[1],
                           (long long)Pf * L * D.sstr[2], (long long)Pf * L * D.sstr[3],
                           (long long)Pf * L * D.sstr[4]};
  if (Nt > 0) {
    ct = make_cfg(order, cfg.R, Nt, cfg.vec ? 1 : 0, 1 << (cfg.cps_log2 + cfg.spb_log2));
    ct.idx32 = cfg.idx32;
    if (!cfg_valid(ct, Nt)) return hipErrorInvalidValue;
  }

  SegOut<T> sot = so;  // the tail launch
  sot.p_tail = Pf;

  hipError_t e;
  // 1. segment reduction
  if (Pf > 0) {
    e = launch_seg<T, 1, 0>(cs, CoefArrays<T>{a, b, c}, d, nullptr, Ds, so, st);
    if (e != hipSuccess) return e;
  }
  if (Nt > 0) {
    e = launch_seg<T, 1, 0>(ct, CoefArrays<T>{a + sh[0], b + sh[1], c + sh[2]}, d + sh[3],
                            nullptr, Dt, sot, st);
    if (e != hipSuccess) return e;
  }

  // 2. reduced systems: 2P rows each, laid out like the original order
  SysDesc Dr;
  memset(&Dr, 0, sizeof Dr);
  Dr.N = 2 * P;
  Dr.nsys = ns;
  Dr.nother = 1;
  Dr.osize[0] = (int)ns;
  for (int t = 0; t < 5; t++) {
    Dr.sstr[t] = so.rs_row;
    Dr.ostr[t][0] = so.rs_sys;
  }
  Dr.small_index = !wide_index();
  const int ridx32 = (unsigned long long)n2 * sizeof(T) < (1ULL << 32) && !wide_index();
  // (16-byte loads along the reduced rows for ORDER 0: the workspace arrays
  // are 16-byte aligned and each system's 2P rows start at a multiple of 2P)
  const int rvec = order == 0 && (2 * P * (int)sizeof(T)) % 16 == 0;
  PartCfg cr = heuristic(order, 2 * P, rvec, sizeof(T));
  if (cr.kind == 1) {
    if (stage_ok(cr, order == 0, ns, 2 * P)) cr.vec = staged_vec(cr);
    cr.idx32 = ridx32;
    e = launch_part<T, 0>(cr, CoefArrays<T>{ra, rb, rc}, rd, nullptr, Dr, st);
  } else {
    cr = split_cfg(order, rvec, sizeof(T), ns);
    cr.idx32 = ridx32;
    e = split_solve<T, 0>(cr, ra, rb, rc, rd, nullptr, Dr, st, next);
  }
  if (e != hipSuccess) return e;

  // 3. independent segments, couplings from the reduced solution
  if (Pf > 0) {
    e = launch_seg<T, 2, INC>(cs, CoefArrays<T>{a, b, c}, d, u, Ds, so, st);
    if (e != hipSuccess) return e;
  }
  if (Nt > 0)
    e = launch_seg<T, 2, INC>(ct, CoefArrays<T>{a + sh[0], b + sh[1], c + sh[2]}, d + sh[3],
                              INC ? u + sh[4] : nullptr, Dt, sot, st);
  return e;
}

std::mutex g_tune_mu;
// tuned configuration per (element size, Inc, order, N, systems, layout,
// solve stride, first other stride): strides decide the memory pattern (sd1
// and sd2 of a cube share everything else)
std::map<std::tuple<int, int, int, int, long long, int, long long, long long>, PartCfg> g_tuned;

// Extent (in elements) of an array with the given dims/pads.
long long extent(int ndim, const int *dims, const int *pads) {
  long long e = 0, cum = 1;
  for (int k = 0; k < ndim; k++) {
    if (dims[k] == 0) return 0;
    e += (long long)(dims[k] - 1) * cum;
    cum *= pads[k];
  }
  return e + 1;
}

// Runs one configuration.  Scratch of the fallback and split-line kinds
// comes from the stream's workspace, or, when `fixed` is given, from that
// buffer of `fixed_bytes`, which must hold it.
template <typename T, int INC>
tridStatus_t run_cfg(const PartCfg &cfg, const T *a, const T *b, const T *c,
                     T *d, T *u, const SysDesc &D, hipStream_t st, void *fixed = nullptr,
                     size_t fixed_bytes = 0) {
  hipError_t e;
  if (cfg.kind == 2 || cfg.kind == 4) {
    const size_t bytes =
        cfg.kind == 2 ? 2 * (size_t)D.N * (size_t)D.nsys * sizeof(T)
                      : (size_t)split_ws_elems(cfg, D.N, D.nsys, sizeof(T)) * sizeof(T);
    T *ws;
    if (fixed) {
      if (bytes > fixed_bytes)
        return fail(TRID_STATUS_INVALID_VALUE,
                    "this solve needs " + std::to_string(bytes) + " bytes of scratch; the "
                    "no-allocation entry point holds " + std::to_string(fixed_bytes));
      ws = static_cast<T *>(fixed);
    } else {
      ws = static_cast<T *>(workspace(bytes, st));
      if (!ws) return TRID_STATUS_ALLOC_FAILED;  // reason recorded by workspace()
    }
    if (cfg.kind == 2)
      e = launch_fallback<T, INC>(a, b, c, d, u, D, st, ws);
    else
      e = split_solve<T, INC>(cfg, a, b, c, d, u, D, st, ws);
  } else {
    e = launch_part<T, INC>(cfg, CoefArrays<T>{a, b, c}, d, u, D, st);
  }
  return check_hip(e, "kernel launch");
}

// cp: the cache policy the candidates run with; unless a policy is forced
// (tridForceCachePolicy), the winner is then timed against its twins under
// the other policies, and keeps its own unless a twin is more than 2 %
// faster.
template <typename T, int INC>
PartCfg autotune(int order, int vec, bool gvec, int idx32, bool dense, const T *a,
                 const T *b, const T *c, T *d, T *u, const SysDesc &D,
                 long long d_extent, int cp, hipStream_t st) {
  std::vector<PartCfg> cands =
      candidates(order, D.N, vec, sizeof(T), dense, D.nsys, gvec);
  for (auto &cf : cands) {
    cf.idx32 = idx32;
    cf.cp = idx32 ? cp : kStream;
  }
  PartCfg best = heuristic(order, D.N, vec, sizeof(T));
  if (stage_ok(best, dense, D.nsys, D.N)) best.vec = staged_vec(best);
  else if (gstage_ok(best, gvec, dense)) best.vec = 4;
  best.idx32 = idx32;
  best.cp = idx32 ? cp : kStream;
  if (cands.size() < 2) return best;
  // the candidates run on the caller's own arrays (memory-channel effects of
  // large power-of-two strides depend on the actual addresses), repeatedly
  // and in place as a caller runs them; the written array is saved first and
  // restored at the end
  T *saved = static_cast<T *>(workspace((size_t)d_extent * sizeof(T), st));
  if (!saved) return best;
  T *dst = INC ? u : d;
  const size_t bytes = (size_t)d_extent * sizeof(T);
  if (hipMemcpyAsync(saved, dst, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) return best;
  auto prep = [&](int) { return true; };
  auto run = [&](int i) {
    return launch_part<T, INC>(cands[i], CoefArrays<T>{a, b, c}, d, u, D, st);
  };
  std::vector<float> tm;
  const int i = pick_fastest((int)cands.size(), prep, run, st, &tm);
  if (i >= 0) best = cands[i];
  const char *log = std::getenv("TRIDHIP_TUNE_LOG");
  if (log && log[0] == '1')
    for (size_t c = 0; c < cands.size(); c++)
      fprintf(stderr, "[tridhip tune] N=%d nsys=%lld order %d %s: %.4f ms\n", D.N, D.nsys,
              order, cfg_name(cands[c], sizeof(T)).c_str(), tm[c]);
  if (i >= 0 && idx32 && forced_cache_policy() < 0) {
    // the same layout under the other cache policies
    const PartCfg win = best;
    cands.assign({win});
    for (int p : {kCached, kMixed, kStream})
      if (p != win.cp) {
        PartCfg twin = win;
        twin.cp = p;
        cands.push_back(twin);
      }
    const int j = pick_fastest((int)cands.size(), prep, run, st, &tm);
    if (j >= 0) best = cands[j];
    if (log && log[0] == '1')
      for (size_t c = 0; c < cands.size(); c++)
        fprintf(stderr, "[tridhip tune] N=%d nsys=%lld order %d cache policy: %s %.4f ms\n",
                D.N, D.nsys, order, cfg_name(cands[c], sizeof(T)).c_str(), tm[c]);
  }
  (void)hipMemcpyAsync(dst, saved, bytes, hipMemcpyDeviceToDevice, st);
  return best;
}

// Untuned configuration of a strided batch: the heuristic chunk layout,
// the strided default for solvedim >= 1, wave staging where it applies and
// the chunk length that pads a solvedim-0 system least.
template <typename T>
PartCfg default_cfg(int order, const SysDesc &D, int vec, bool gvec, bool dense) {
  PartCfg cfg = heuristic(order, D.N, vec, sizeof(T));
  if (order == 1 && cfg.kind == 1) {
    const PartCfg s = strided_default(D.N, sizeof(T));
    if (s.kind == 1) cfg = s;
  }
  // staging pays when a chunk is more than one 16-byte vector
  if (cfg.R * sizeof(T) > 16) {
    if (stage_ok(cfg, dense, D.nsys, D.N)) {
      cfg.vec = staged_vec(cfg);
    } else {
      // general staging takes up to 512 threads: longer FP64 chunks if
      // needed (FP32 is faster unstaged than in 16-row chunks)
      PartCfg c16 = make_cfg(order, 16, D.N, cfg.vec, 256);
      if (sizeof(T) == 8 && !gstage_ok(cfg, gvec, dense) && cfg_valid(c16, D.N) &&
          gstage_ok(c16, gvec, dense))
        cfg = c16;
      if (gstage_ok(cfg, gvec, dense)) cfg.vec = 4;
    }
  }
  // solvedim 0 between powers of two: 12-row chunks when they pad less
  return pick_rows(cfg, D.N, gvec, dense, sizeof(T));
}

template <typename T, int INC>
tridStatus_t solve(const T *a, const int *a_pads, const T *b,
                   const int *b_pads, const T *c, const int *c_pads, T *d,
                   const int *d_pads, T *u, const int *u_pads, int ndim,
                   int solvedim, const int *dims, const int *opts, int sync,
                   void *fixed = nullptr, size_t fixed_bytes = 0) {
  (void)opts;  // accepted for compatibility; the kernel choice is automatic
  clear_error();
  if (!dims) return fail(TRID_STATUS_INVALID_VALUE, "dims is NULL");
  if (ndim < 1 || ndim > MAXDIM)
    return fail(TRID_STATUS_INVALID_VALUE, "ndim must be in [1, 8]");
  if (solvedim < 0 || solvedim >= ndim)
    return fail(TRID_STATUS_INVALID_VALUE, "solvedim must be in [0, ndim)");
  const int *pads[5] = {a_pads ? a_pads : dims, b_pads ? b_pads : dims,
                        c_pads ? c_pads : dims, d_pads ? d_pads : dims,
                        u_pads ? u_pads : (d_pads ? d_pads : dims)};
  for (int k = 0; k < ndim; k++) {
    if (dims[k] < 0) return fail(TRID_STATUS_INVALID_VALUE, "negative dims");
    for (int t = 0; t < 5; t++)
      if (pads[t][k] < dims[k])
        return fail(TRID_STATUS_INVALID_VALUE, "pads[i] < dims[i]");
  }
  SysDesc D;
  memset(&D, 0, sizeof D);
  D.N = dims[solvedim];
  D.nsys = 1;
  for (int k = 0; k < ndim; k++)
    if (k != solvedim) {
      D.osize[D.nother++] = dims[k];
      D.nsys *= dims[k];
    }
  if (D.N == 0 || D.nsys == 0) return TRID_STATUS_SUCCESS;
  if (!a || !b || !c || !d)
    return fail(TRID_STATUS_INVALID_VALUE, "a, b, c and d must be non-NULL");
  if (INC && !u) return fail(TRID_STATUS_INVALID_VALUE, "u is NULL in an Inc call");
  for (int t = 0; t < 5; t++) {
    long long cum = 1;
    int o = 0;
    for (int k = 0; k < ndim; k++) {
      if (k == solvedim)
        D.sstr[t] = cum;
      else
        D.ostr[t][o++] = cum;
      cum *= pads[t][k];
    }
  }
  D.small_index = D.nsys < (1LL << 31) && !wide_index();
  // 32-bit offsets when every array (its allocated extent) fits in 4 GiB
  int idx32 = !wide_index();
  size_t footprint = 0, rhs_bytes = 0;  // bytes spanned by all arrays, by d and u
  for (int t = 0; t < 5; t++) {
    if (t == 4 && !INC) break;
    const unsigned long long bytes = (unsigned long long)extent(ndim, dims, pads[t]) * sizeof(T);
    if (bytes >= (1ULL << 32)) idx32 = 0;
    footprint += bytes;
    if (t >= 3) rhs_bytes += bytes;
  }
  const int cp_rule = cache_policy(footprint, rhs_bytes);

  const int order = solvedim == 0 ? 0 : 1;
  // vector loads along the system need unit stride and 16-byte aligned chunks
  constexpr int E = 16 / sizeof(T);
  int vec = order == 0;
  const void *ptrs[5] = {a, b, c, d, u};
  for (int t = 0; t < 5 && vec; t++) {
    if (t == 4 && !INC) break;
    if (!aligned16(ptrs[t])) vec = 0;
    for (int k = 0; k < D.nother; k++)
      if (D.ostr[t][k] % E != 0) vec = 0;
  }

  // general wave staging needs only unit stride and 16-byte aligned bases
  bool gvec = order == 0;
  for (int t = 0; t < 5 && gvec; t++) {
    if (t == 4 && !INC) break;
    if (!aligned16(ptrs[t])) gvec = false;
  }

  // dense: every array is unpadded (system q of a solvedim-0 batch at q*N)
  bool dense = true;
  for (int t = 0; t < 5; t++) {
    if (t == 4 && !INC) break;
    for (int k = 0; k < ndim; k++) dense &= pads[t][k] == dims[k];
  }
  hipStream_t st = current_stream();
  PartCfg cfg;
  bool tuned = false;  // cfg (and its cache policy) from the autotuner
  ForcedConfig fc = forced_config();
  if (fc.kind == 2) {
    cfg.kind = 2;
  } else if (fc.kind == 4) {
    cfg = split_cfg(order, vec, sizeof(T), D.nsys);
    if (fc.rows) cfg.R = fc.rows;
    if (fc.cps) {
      // keep the work-group size of the default shape
      const int thr_log2 = cfg.cps_log2 + cfg.spb_log2;
      cfg.cps_log2 = ilog2(fc.cps);
      cfg.spb_log2 = std::max(0, thr_log2 - cfg.cps_log2);
      if (order == 1) cfg.spb_log2 = std::min(cfg.spb_log2, 6);
    }
    if (fc.spb) cfg.spb_log2 = ilog2(fc.spb);
    cfg.vec = vec && (cfg.R * (int)sizeof(T)) % 16 == 0;
    if (!cfg_valid(cfg, D.N))
      return fail(TRID_STATUS_INVALID_VALUE, "forced configuration does not fit this shape");
  } else if (fc.kind == 1 || fc.kind == 3) {
    int R = fc.rows ? fc.rows : (sizeof(T) == 8 ? 8 : 16);
    cfg = make_cfg(order, R, D.N, vec && (R * (int)sizeof(T)) % 16 == 0,
                   order == 0 ? 256 : 512);
    if (fc.cps) cfg.cps_log2 = ilog2(fc.cps);
    if (fc.spb) cfg.spb_log2 = ilog2(fc.spb);
    if (is_fill_rows(R)) {
      // these chunk lengths exist in the general staged layout only
      if (!gstage_ok(cfg, gvec, dense))
        return fail(TRID_STATUS_INVALID_VALUE,
                    "rows 5, 6, 7 and 12 need the general staged layout");
      cfg.vec = 4;
    }
    if (!cfg_valid(cfg, D.N, sizeof(T)))
      return fail(TRID_STATUS_INVALID_VALUE, "forced configuration does not fit this shape");
    if (fc.stage && stage_ok(cfg, dense, D.nsys, D.N)) cfg.vec = staged_vec(cfg);
    else if (fc.stage && gstage_ok(cfg, gvec, dense)) cfg.vec = 4;
    cfg.xcd = fc.kind == 3;
  } else {
    // tuned configurations are cached per shape; a shape first seen while
    // `st` is being captured takes the untuned default (tuning synchronises
    // and times launches, which a capture cannot hold) and is tuned at its
    // first call outside a capture
    // (a call with fixed scratch never tunes: tuning allocates)
    bool have = false;
    const bool tune = autotune_mode() == 1 && !fixed;
    tuned = tune;
    auto key = std::make_tuple((int)sizeof(T), INC, order, D.N, D.nsys,
                               vec * 4 + (int)gvec * 2 + (int)dense, D.sstr[3],
                               D.nother > 0 ? D.ostr[3][0] : 0LL);
    if (tune) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) {
        cfg = it->second;
        have = true;
      }
    }
    if (!have && tune && !capturing(st)) {
      long long ext = extent(ndim, dims, INC ? pads[4] : pads[3]);
      cfg = autotune<T, INC>(order, vec, gvec, idx32, dense, a, b, c, d, u, D, ext, cp_rule,
                             st);
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tuned[key] = cfg;
      have = true;
    }
    if (!have) {
      cfg = default_cfg<T>(order, D, vec, gvec, dense);
      tuned = false;
    }
  }
  // systems longer than one work-group holds: split-line solve.  Strided
  // systems (ORDER 1) of 2048 rows and more take it too: a work-group that
  // holds whole systems of that length has room for one or two of them,
  // i.e. 8- to 16-byte row runs (3-30 % of HBM bandwidth on MI355X), while
  // the split-line segments keep 64 systems side by side (about 69 % of its
  // 9 words per unknown, i.e. 38 % on the one-pass scale)
  if (fc.kind == 0 && cfg.kind == 1 && order == 1 && D.N >= 2048) cfg.kind = 4;
  if (cfg.kind == 4 && fc.kind != 4) cfg = split_cfg(order, vec, sizeof(T), D.nsys);
  if (cfg.kind == 4 && D.nsys > (1LL << 30))
    return fail(TRID_STATUS_INVALID_VALUE, "split-line solve: too many systems");
  cfg.idx32 = idx32;
  // cache policy: the tuned one, else the footprint rule; a forced policy
  // wins; split-line, fallback and 64-bit-offset solves always stream
  if (!tuned) cfg.cp = cp_rule;
  if (forced_cache_policy() >= 0) cfg.cp = forced_cache_policy();
  if (!idx32 || cfg.kind != 1) cfg.cp = kStream;
  set_last_config(cfg_name(cfg, sizeof(T)));
  // launch events (tridSetLaunchEvents) time single-kernel solves only
  if (cfg.kind != 1) (void)take_launch_events();
  tridStatus_t s = run_cfg<T, INC>(cfg, a, b, c, d, u, D, st, fixed, fixed_bytes);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (sync) return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  return TRID_STATUS_SUCCESS;
}

}  // namespace
}  // namespace tridhip

using tridhip::solve;

EXTERN_C tridStatus_t tridDebugReciprocal(const void *x, void *y, long long n,
                                          int is_double) {
  using namespace tridhip;
  clear_error();
  if (n < 0 || (n > 0 && (!x || !y)))
    return fail(TRID_STATUS_INVALID_VALUE, "tridDebugReciprocal: bad arguments");
  if (n == 0) return TRID_STATUS_SUCCESS;
  hipStream_t st = current_stream();
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (is_double)
    hipLaunchKernelGGL(k_rcp<double>, dim3(blocks), dim3(256), 0, st,
                       static_cast<const double *>(x), static_cast<double *>(y), n);
  else
    hipLaunchKernelGGL(k_rcp<float>, dim3(blocks), dim3(256), 0, st,
                       static_cast<const float *>(x), static_cast<float *>(y), n);
  tridStatus_t s = check_hip(hipGetLastError(), "k_rcp launch");
  if (s) return s;
  return check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
}

EXTERN_C void initTridMultiDimBatchSolve(int ndim, int *dims, int *pads) {
  (void)ndim;
  (void)dims;
  (void)pads;
}

EXTERN_C tridStatus_t tridSmtsvStridedBatch(const float *a, const float *b,
                                            const float *c, float *d, float *u,
                                            int ndim, int solvedim, int *dims,
                                            int *pads, int *opts, int sync) {
  (void)u;
  return solve<float, 0>(a, pads, b, pads, c, pads, d, pads, nullptr, pads,
                         ndim, solvedim, dims, opts, sync);
}

EXTERN_C tridStatus_t tridDmtsvStridedBatch(const double *a, const double *b,
                                            const double *c, double *d,
                                            double *u, int ndim, int solvedim,
                                            int *dims, int *pads, int *opts,
                                            int sync) {
  (void)u;
  return solve<double, 0>(a, pads, b, pads, c, pads, d, pads, nullptr, pads,
                          ndim, solvedim, dims, opts, sync);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchInc(const float *a, const float *b,
                                               const float *c, float *d,
                                               float *u, int ndim,
                                               int solvedim, int *dims,
                                               int *pads, int *opts,
                                               int sync) {
  return solve<float, 1>(a, pads, b, pads, c, pads, d, pads, u, pads, ndim,
                         solvedim, dims, opts, sync);
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchInc(const double *a,
                                               const double *b,
                                               const double *c, double *d,
                                               double *u, int ndim,
                                               int solvedim, int *dims,
                                               int *pads, int *opts,
                                               int sync) {
  return solve<double, 1>(a, pads, b, pads, c, pads, d, pads, u, pads, ndim,
                          solvedim, dims, opts, sync);
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchPadded(
    const double *a, const int *a_pads, const double *b, const int *b_pads,
    const double *c, const int *c_pads, double *d, const int *d_pads,
    double *u, const int *u_pads, int ndim, int solvedim, int *dims, int *opts,
    int sync) {
  (void)u;
  return solve<double, 0>(a, a_pads, b, b_pads, c, c_pads, d, d_pads, nullptr,
                          u_pads, ndim, solvedim, dims, opts, sync);
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchPaddedInc(
    const double *a, const int *a_pads, const double *b, const int *b_pads,
    const double *c, const int *c_pads, double *d, const int *d_pads,
    double *u, const int *u_pads, int ndim, int solvedim, int *dims, int *opts,
    int sync) {
  return solve<double, 1>(a, a_pads, b, b_pads, c, c_pads, d, d_pads, u,
                          u_pads, ndim, solvedim, dims, opts, sync);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchPadded(
    const float *a, const int *a_pads, const float *b, const int *b_pads,
    const float *c, const int *c_pads, float *d, const int *d_pads, float *u,
    const int *u_pads, int ndim, int solvedim, int *dims, int *opts,
    int sync) {
  (void)u;
  return solve<float, 0>(a, a_pads, b, b_pads, c, c_pads, d, d_pads, nullptr,
                         u_pads, ndim, solvedim, dims, opts, sync);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchPaddedInc(
    const float *a, const int *a_pads, const float *b, const int *b_pads,
    const float *c, const int *c_pads, float *d, const int *d_pads, float *u,
    const int *u_pads, int ndim, int solvedim, int *dims, int *opts,
    int sync) {
  return solve<float, 1>(a, a_pads, b, b_pads, c, c_pads, d, d_pads, u,
                         u_pads, ndim, solvedim, dims, opts, sync);
}

// ---------------------------------------------------------------------------
// no-allocation entry point
//
// The scratch of split-line and one-thread-per-system solves comes from this
// file-scope device array instead of the per-stream workspace.  Every device
// holds its own copy at its own address; host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol) but the
// address hipGetSymbolAddress returns for the current device, fetched once
// per device.  Nothing in it needs clearing between calls: every kernel that
// uses it writes each value before reading it.  Calls that overlap on one
// device (two streams, two threads) share the array and race.
static __device__ __align__(256) unsigned char scratch[TRID_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device -> address of `scratch` there

tridStatus_t scratch_address(void **out) {
  int dev = 0;
  tridStatus_t s = tridhip::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (s) return s;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return TRID_STATUS_SUCCESS;
  }
  void *ptr = nullptr;
  s = tridhip::check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                         "hipGetSymbolAddress(scratch)");
  if (s) return s;
  g_scratch_ptr[dev] = ptr;
  *out = ptr;
  return TRID_STATUS_SUCCESS;
}
}  // namespace

EXTERN_C tridStatus_t tridDmtsvStridedBatchNoAlloc(const double *a, const double *b,
                                                   const double *c, double *d, double *u,
                                                   int ndim, int solvedim, int *dims,
                                                   int *pads, int *opts, int sync) {
  (void)u;
  tridhip::clear_error();
  void *ptr = nullptr;
  tridStatus_t s = scratch_address(&ptr);
  if (s) return s;
  return solve<double, 0>(a, pads, b, pads, c, pads, d, pads, nullptr, pads, ndim, solvedim,
                          dims, opts, sync, ptr, TRID_NOALLOC_SCRATCH_BYTES);
}
