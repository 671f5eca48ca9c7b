// trid_launch.hpp -- kernel configuration table and launchers of the
// partitioned Thomas kernels (shared by the strided-batch entry points and
// the ADI driver).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "trid_part.hpp"
#include "trid_runtime.hpp"

namespace tridhip {

template <typename T, int INC, int ORDER, int R, int VEC, typename I,
          typename Coef, int CP>
__global__ void __launch_bounds__(((R >= 16 && sizeof(T) == 8) || VEC == 4) ? 512 : 1024)
    k_part(Coef coef, T *__restrict__ Dd, T *__restrict__ U, SysDesc D,
           int cps_log2, int spb_log2, int xcd, T *__restrict__ Out) {
  // CP: cache policy of the streaming accesses (kStream / kCached)
  // xcd: consecutive hardware blocks go round-robin to the 8 XCDs (each with
  // its own L2); this remap hands every XCD one contiguous range of systems
  // instead (a bijection on the first 8*floor(nblocks/8) blocks, identity on
  // the rest).
  long long block = blockIdx.x;
  if (xcd) {
    const unsigned nb8 = gridDim.x & ~7u;
    if (blockIdx.x < nb8) block = (long long)(blockIdx.x & 7u) * (nb8 >> 3) + (blockIdx.x >> 3);
  }
  part_block<T, INC, ORDER, R, VEC, I, Coef, 0, CP>(coef, Dd, U, D, cps_log2, spb_log2,
                                                    block, Out);
}

// Segment steps of a split-line solve: D's systems are segments of longer
// systems.  MODE 1: their merged reduced rows go to `so`; MODE 2: solved
// with the couplings to the neighbouring segments taken from so.rd.
template <typename T, int MODE, int INC, int ORDER, int R, int VEC, typename I>
__global__ void __launch_bounds__((R >= 16 && sizeof(T) == 8) ? 512 : 1024)
    k_seg(CoefArrays<T> coef, T *__restrict__ Dd, T *__restrict__ U, SysDesc D, int cps_log2,
          int spb_log2, SegOut<T> so) {
  part_block<T, INC, ORDER, R, VEC, I, CoefArrays<T>, MODE>(
      coef, Dd, U, D, cps_log2, spb_log2, blockIdx.x, nullptr, so);
}

struct PartCfg {
  int kind = 1;  // 1 partitioned, 2 fallback, 4 split-line (segments of R*cps rows)
  int R = 8;
  int cps_log2 = 0;
  int spb_log2 = 0;
  int order = 0;
  int vec = 0;
  int idx32 = 1;  // 32-bit element offsets (every array < 4 GiB)
  int xcd = 0;    // XCD-contiguous block order (see k_part)
  int cp = 0;     // cache policy of the streaming accesses: kStream / kCached / kMixed
};

inline bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

inline int ilog2(long long v) {
  int l = 0;
  while ((1LL << l) < v) l++;
  return l;
}

template <typename T, int INC, int ORDER, int R, int VEC, typename Coef>
hipError_t launch_part_t(const PartCfg &cfg, const Coef &coef, T *d, T *u,
                         const SysDesc &D, hipStream_t st, T *out) {
  const int cps = 1 << cfg.cps_log2, spb = 1 << cfg.spb_log2;
  const int threads = cps * spb;
  const long long blocks = (D.nsys + spb - 1) / spb;
  size_t lds = 0;
  // reduced-system exchange: reduced_lds (6 * cps * spb values) or, for
  // ORDER 1 with cps <= 64, reduced_lds_w (6 * cps * (spb + 1))
  if (ORDER == 1 || cps > kWave) lds = 2 * 3 * (size_t)cps * (spb + 1) * sizeof(T);
  if (VEC == 2) lds = (size_t)threads * R * sizeof(T);  // wave staging buffers
  // staging buffers, then 12 values per wave for the reduced solve
  if (VEC == 3) lds = (size_t)threads * R * sizeof(T) + 12 * (size_t)(threads / kWave) * sizeof(T);
  // general staging: per wave 64 R + E elements (E = 16 / sizeof(T)), then
  // 12 values per wave for the reduced solve of systems over several waves
  if (VEC == 4)
    lds = (size_t)(threads / kWave) * (kWave * R + 16 / sizeof(T) + (cps > kWave ? 12 : 0)) *
          sizeof(T);
  // the cached policy exists for 32-bit offsets only: a solve whose arrays
  // fit the last-level cache has arrays far below 4 GiB
  if (cfg.idx32 && cfg.cp == kCached)
    return launch_kernel(k_part<T, INC, ORDER, R, VEC, uint32_t, Coef, kCached>,
                         dim3((unsigned)blocks), dim3(threads), (std::uint32_t)lds, st, coef, d,
                         u, D, cfg.cps_log2, cfg.spb_log2, cfg.xcd, out);
  if constexpr (Coef::kArrays) {
    if (cfg.idx32 && cfg.cp == kMixed)
      return launch_kernel(k_part<T, INC, ORDER, R, VEC, uint32_t, Coef, kMixed>,
                           dim3((unsigned)blocks), dim3(threads), (std::uint32_t)lds, st, coef,
                           d, u, D, cfg.cps_log2, cfg.spb_log2, cfg.xcd, out);
  }
  if (cfg.idx32)
    return launch_kernel(k_part<T, INC, ORDER, R, VEC, uint32_t, Coef, kStream>,
                         dim3((unsigned)blocks), dim3(threads), (std::uint32_t)lds, st, coef, d,
                         u, D, cfg.cps_log2, cfg.spb_log2, cfg.xcd, out);
  return launch_kernel(k_part<T, INC, ORDER, R, VEC, long long, Coef, kStream>,
                       dim3((unsigned)blocks), dim3(threads), (std::uint32_t)lds, st, coef, d, u,
                       D, cfg.cps_log2, cfg.spb_log2, cfg.xcd, out);
}

// Launches the partitioned kernel for any configuration of the table.
template <typename T, int INC, typename Coef>
hipError_t launch_part(const PartCfg &cfg, const Coef &coef, T *d, T *u,
                       const SysDesc &D, hipStream_t st, T *out = nullptr) {
  if constexpr (Coef::kHook) {
    // hooked coefficient loaders take the scalar (VEC 0) paths only
#define TRID_CASE0(ORD, RR)                                                    \
  if (cfg.order == ORD && cfg.R == RR)                                         \
    return launch_part_t<T, INC, ORD, RR, 0>(cfg, coef, d, u, D, st, out);
    TRID_CASE0(0, 4)
    TRID_CASE0(0, 8)
    TRID_CASE0(0, 16)
    TRID_CASE0(1, 4)
    TRID_CASE0(1, 8)
    TRID_CASE0(1, 16)
#undef TRID_CASE0
    return hipErrorInvalidValue;
  } else {
  #define TRID_CASE(ORD, RR, VV)                                                 \
    if (cfg.order == ORD && cfg.R == RR && cfg.vec == VV)                        \
      return launch_part_t<T, INC, ORD, RR, VV>(cfg, coef, d, u, D, st, out);
    TRID_CASE(0, 4, 0)
    TRID_CASE(0, 8, 0)
    TRID_CASE(0, 16, 0)
    TRID_CASE(0, 4, 1)
    TRID_CASE(0, 8, 1)
    TRID_CASE(0, 16, 1)
    TRID_CASE(0, 4, 2)
    TRID_CASE(0, 8, 2)
    TRID_CASE(0, 16, 2)
    TRID_CASE(0, 4, 3)
    TRID_CASE(0, 8, 3)
    TRID_CASE(0, 16, 3)
    if constexpr (Coef::kArrays) {
      TRID_CASE(0, 4, 4)
      TRID_CASE(0, 8, 4)
      TRID_CASE(0, 16, 4)
      // chunk lengths that fill a power-of-two chunk count better
      TRID_CASE(0, 5, 4)
      TRID_CASE(0, 6, 4)
      TRID_CASE(0, 7, 4)
      TRID_CASE(0, 12, 4)
    }
    TRID_CASE(1, 4, 0)
    TRID_CASE(1, 8, 0)
    TRID_CASE(1, 16, 0)
  #undef TRID_CASE
    return hipErrorInvalidValue;
  }
}

template <typename T, int MODE, int INC, int ORDER, int R, int VEC>
hipError_t launch_seg_t(const PartCfg &cfg, const CoefArrays<T> &coef, T *d, T *u,
                        const SysDesc &D, const SegOut<T> &so, hipStream_t st) {
  const int cps = 1 << cfg.cps_log2, spb = 1 << cfg.spb_log2;
  const int threads = cps * spb;
  const long long blocks = (D.nsys + spb - 1) / spb;
  size_t lds = 0;
  // LDS: the merge tree (MODE 1) or the reduced solve (MODE 2) runs in LDS
  // unless all chunks of a segment are in one wave (ORDER 0, cps <= 64)
  if (ORDER == 1 || cps > kWave) lds = 2 * 3 * (size_t)cps * (spb + 1) * sizeof(T);
  if (VEC == 2) lds = (size_t)threads * R * sizeof(T);
  if (cfg.idx32)
    hipLaunchKernelGGL((k_seg<T, MODE, INC, ORDER, R, VEC, uint32_t>), dim3((unsigned)blocks),
                       dim3(threads), lds, st, coef, d, u, D, cfg.cps_log2, cfg.spb_log2, so);
  else
    hipLaunchKernelGGL((k_seg<T, MODE, INC, ORDER, R, VEC, long long>), dim3((unsigned)blocks),
                       dim3(threads), lds, st, coef, d, u, D, cfg.cps_log2, cfg.spb_log2, so);
  return hipGetLastError();
}

// Launches segment step MODE (1 reduce, 2 solve) for a configuration of the
// table.
template <typename T, int MODE, int INC>
hipError_t launch_seg(const PartCfg &cfg, const CoefArrays<T> &coef, T *d, T *u,
                      const SysDesc &D, const SegOut<T> &so, hipStream_t st) {
#define TRID_SEG(ORD, RR, VV)                                                  \
  if (cfg.order == ORD && cfg.R == RR && cfg.vec == VV)                        \
    return launch_seg_t<T, MODE, INC, ORD, RR, VV>(cfg, coef, d, u, D, so, st);
  TRID_SEG(0, 4, 0)
  TRID_SEG(0, 8, 0)
  TRID_SEG(0, 16, 0)
  TRID_SEG(0, 4, 1)
  TRID_SEG(0, 8, 1)
  TRID_SEG(0, 16, 1)
  TRID_SEG(0, 4, 2)
  TRID_SEG(0, 8, 2)
  TRID_SEG(0, 16, 2)
  TRID_SEG(1, 4, 0)
  TRID_SEG(1, 8, 0)
  TRID_SEG(1, 16, 0)
#undef TRID_SEG
  return hipErrorInvalidValue;
}

// Chunk lengths of the general staged layout (VEC 4) besides 4, 8 and 16:
// with a power-of-two chunk count they leave fewer rows idle for lengths
// between powers of two (300 rows: 60 chunks of 5 in 64, against 38 of 8).
inline bool is_fill_rows(int R) { return R == 5 || R == 6 || R == 7 || R == 12; }

// tsize: element size; FP32 16-row chunks of strided systems (ORDER 1,
// unstaged) fit 1024-thread work-groups (79 VGPRs), FP64 ones and staged
// layouts do not.
inline bool cfg_valid(const PartCfg &cfg, int N, size_t tsize = 8) {
  if (cfg.kind == 4) {
    // split-line: the segment shape must be a valid work-group shape
    PartCfg s = cfg;
    s.kind = 1;
    return cfg_valid(s, cfg.R << cfg.cps_log2, tsize);
  }
  if (cfg.kind != 1) return true;
  // chunk lengths 5, 6, 7 and 12 exist for the general staged layout only
  if (cfg.R != 4 && cfg.R != 8 && cfg.R != 16 &&
      !(is_fill_rows(cfg.R) && cfg.vec == 4 && cfg.order == 0))
    return false;
  const long long cps = 1LL << cfg.cps_log2, spb = 1LL << cfg.spb_log2;
  const long long thr = cps * spb;
  if (thr < kWave || thr > 1024) return false;
  const bool r16_wide = tsize == 4 && cfg.order == 1 && cfg.vec == 0;
  if (cfg.R >= 16 && thr > 512 && !r16_wide) return false;  // register budget of R=16
  if (cfg.vec == 4 && thr > 512) return false;  // launch bound of general staging
  if (cps * cfg.R < N) return false;
  return true;
}

// Threads per block preferred for a given chunk count.
inline PartCfg make_cfg(int order, int R, int N, int vec, int threads) {
  PartCfg c;
  c.kind = 1;
  c.order = order;
  c.R = R;
  c.vec = vec;
  long long chunks = (std::max(N, 1) + R - 1) / R;
  c.cps_log2 = ilog2(chunks);
  int cps = 1 << c.cps_log2;
  int spb = std::max(1, threads / cps);
  if (order == 1) spb = std::min(spb, 64);
  c.spb_log2 = ilog2(spb);
  return c;
}

// Wave-staged access (VEC 2 / 3) needs dense arrays (system q at q * N) and
// whole chunks (cps * R == N); with cps < 64 also whole waves of systems
// (nsys a multiple of the systems per wave).
inline bool stage_ok(const PartCfg &c, bool dense, long long nsys, int N) {
  if (c.kind != 1 || c.order != 0 || c.vec != 1 || !dense) return false;
  const long long cps = 1LL << c.cps_log2;
  if (cps * c.R != N) return false;
  return cps >= kWave || nsys % (kWave / cps) == 0;
}

// The staged variant of a configuration that stage_ok accepts: VEC 2 with
// all chunks of a system in one wave, VEC 3 for systems over several waves.
inline int staged_vec(const PartCfg &c) { return (1 << c.cps_log2) > kWave ? 3 : 2; }

// General wave staging (VEC 4: any N, spans of any alignment) needs ORDER 0,
// unit stride and 16-byte aligned array bases (gvec), and either one system
// per wave or more (cps >= 64) or dense arrays (the systems of a wave
// adjacent).
inline bool gstage_ok(const PartCfg &c, bool gvec, bool dense) {
  if (c.kind != 1 || c.order != 0 || c.vec >= 2 || !gvec) return false;
  if ((1 << (c.cps_log2 + c.spb_log2)) > 512) return false;
  return (1 << c.cps_log2) >= kWave || dense;
}

inline std::string cfg_name(const PartCfg &c, size_t tsize) {
  char buf[128];
  if (c.kind == 2) {
    snprintf(buf, sizeof buf, "fallback<%s>", tsize == 8 ? "double" : "float");
  } else if (c.kind == 4) {
    snprintf(buf, sizeof buf, "split<%s,o%d,R%d,L%d,spb%d%s>",
             tsize == 8 ? "double" : "float", c.order, c.R, c.R << c.cps_log2,
             1 << c.spb_log2, c.vec >= 2 ? ",staged" : (c.vec ? ",vec" : ""));
  } else {
    snprintf(buf, sizeof buf, "part<%s,o%d,R%d,cps%d,spb%d%s%s%s>",
             tsize == 8 ? "double" : "float", c.order, c.R, 1 << c.cps_log2,
             1 << c.spb_log2,
             c.vec == 4 ? ",gstaged" : (c.vec >= 2 ? ",staged" : (c.vec ? ",vec" : "")),
             c.xcd ? ",xcd" : "",
             c.cp == kCached ? ",cached" : (c.cp == kMixed ? ",mixed" : ""));
  }
  return buf;
}

inline std::vector<PartCfg> candidates(int order, int N, int vec, size_t tsize,
                                bool dense, long long nsys, bool gvec = false) {
  std::vector<PartCfg> out;
  const int Rs[3] = {4, 8, 16};
  const int thr[3] = {256, 512, 1024};
  auto add = [&](const PartCfg &c) {
    for (auto &o : out)
      if (o.R == c.R && o.cps_log2 == c.cps_log2 && o.spb_log2 == c.spb_log2 &&
          o.vec == c.vec && o.xcd == c.xcd)
        return;
    out.push_back(c);
  };
  auto add_both = [&](PartCfg c) {
    add(c);
    // the XCD-contiguous order only matters with several blocks per XCD
    if ((nsys >> c.spb_log2) >= 64) {
      c.xcd = 1;
      add(c);
    }
  };
  for (int R : Rs)
    for (int t : thr) {
      PartCfg c = make_cfg(order, R, N, vec && (R * (int)tsize) % 16 == 0, t);
      if (!cfg_valid(c, N, tsize)) continue;
      add_both(c);
      if (stage_ok(c, dense, nsys, N)) {
        c.vec = staged_vec(c);
        add_both(c);
      } else if (gstage_ok(c, gvec, dense)) {
        c.vec = 4;
        add_both(c);
      }
    }
  for (int R : {5, 6, 7, 12})
    for (int t : thr) {
      PartCfg c = make_cfg(order, R, N, 0, t);
      if (!gstage_ok(c, gvec, dense)) continue;
      c.vec = 4;
      if (cfg_valid(c, N, tsize)) add_both(c);
    }
  return out;
}

// Default layout of strided (solvedim >= 1) batches with coefficient arrays:
// as many systems side by side as the register budget allows, i.e. the
// longest chunks at the largest work-group (FP64: 16 rows at 512 threads
// past 256 rows; FP32: 8 rows up to 256 rows, else 16, at 1024 threads),
// with the XCD-contiguous order past 512 rows (FP32: past 128).  From tuning
// sweeps on MI355X over [256, N, k] batches of 2^24 unknowns: at N = 300-512
// this moves FP64 from 33 % to 46-59 % of 8 TB/s and FP32 from 31 % to
// 52-58 %; at N = 1500 FP64 from 8 % to 35 %.  Returns kind 0 when the rule
// does not apply (FP64 up to 256 rows keeps the generic choice).
inline PartCfg strided_default(int N, size_t tsize) {
  PartCfg c;
  c.kind = 0;
  if (tsize == 8 && N <= 256) return c;
  const int R = (tsize == 4 && N <= 256) ? 8 : 16;
  PartCfg t = make_cfg(1, R, N, 0, tsize == 8 ? 512 : 1024);
  t.xcd = N > (tsize == 8 ? 512 : 128);
  if (!cfg_valid(t, N, tsize)) return c;
  return t;
}

// Default chunk length of solvedim-0 solves that are not exactly staged:
// the configuration c (FP64: 8 or 16 rows, general staging; FP32: 4 or 8
// rows, 16-byte loads) or general staging with 12-row chunks (FP64 also
// 7-row chunks), whichever pads the system to the fewest rows (cps * R, cps
// the power of two >= N / R); ties keep c.  From tuning sweeps on MI355X
// (N = 70..8000): 12- and 7-row chunks fill the lengths between the powers
// of two that 8-row (or 4-row) chunks fit, and win there by 10-30 %.
inline PartCfg pick_rows(const PartCfg &c, int N, bool gvec, bool dense, size_t tsize) {
  if (c.kind != 1 || c.order != 0 || c.vec == 2 || c.vec == 3) return c;
  PartCfg best = c;
  for (int R : {12, 7}) {
    if (R == 7 && tsize != 8) continue;
    PartCfg t = make_cfg(0, R, N, 0, 256);
    if (!gstage_ok(t, gvec, dense)) continue;
    t.vec = 4;
    if (!cfg_valid(t, N, tsize)) continue;
    if (((long long)t.R << t.cps_log2) >= ((long long)best.R << best.cps_log2)) continue;
    t.xcd = c.xcd;
    t.idx32 = c.idx32;
    best = t;
  }
  return best;
}

// Autotuner timing: `passes` interleaved rounds over the candidates; in each
// round candidate i gets prep(i) (untimed), one warm-up run(i) and `reps`
// timed run(i) launches on st.  Returns the index with the lowest time of
// its best round (-1 if no candidate could be launched).
// The first candidate (in the caller's order) within `tol` of the fastest
// time: a choice that timing noise between near-equal candidates does not
// flip from run to run.  -1 when no candidate ran.
inline int stable_pick(const std::vector<float> &t, float tol = 0.02f) {
  float lo = 1e30f;
  for (float v : t) lo = std::min(lo, v);
  if (!(lo < 1e29f)) return -1;
  for (size_t i = 0; i < t.size(); i++)
    if (t[i] <= lo * (1.0f + tol)) return (int)i;
  return -1;
}

template <typename Prep, typename Run>
int pick_fastest(int ncand, Prep prep, Run run, hipStream_t st,
                 std::vector<float> *times = nullptr, int reps = 5, int passes = 2) {
  std::vector<float> best(ncand, 1e30f);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  for (int pass = 0; pass < passes; pass++) {
    for (int i = 0; i < ncand; i++) {
      if (!prep(i)) continue;
      if (run(i) != hipSuccess) {
        (void)hipGetLastError();
        continue;
      }
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < reps; r++) (void)run(i);
      (void)hipEventRecord(e1, st);
      if (hipEventSynchronize(e1) != hipSuccess) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      best[i] = std::min(best[i], ms);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipGetLastError();
  if (times) *times = best;
  return stable_pick(best);
}

// Split-line configuration for systems longer than one work-group holds:
// segments of L = R * cps rows, from sweeps on MI355X over 2^25 unknowns:
//   ORDER 0: 64 chunks per segment (L = 512 FP64 / 1024 FP32), 256 threads;
//   ORDER 1: up to 64 systems side by side and 512 (FP64) / 256 (FP32)
//            threads -- short segments (L = 64) for wide batches, which beat
//            longer ones despite the longer reduced system.
inline PartCfg split_cfg(int order, int vec, size_t tsize, long long nsys) {
  PartCfg c;
  c.kind = 4;
  c.order = order;
  c.R = tsize == 8 ? 8 : 16;
  c.vec = order == 0 && vec && (c.R * (int)tsize) % 16 == 0;
  if (order == 0) {
    c.cps_log2 = 6;
    c.spb_log2 = 2;
  } else {
    c.spb_log2 = std::min(6, ilog2(std::max(1LL, nsys)));
    c.cps_log2 = std::max(1, (tsize == 8 ? 9 : 8) - c.spb_log2);
    if (c.spb_log2 + c.cps_log2 < 6) c.spb_log2 = 6 - c.cps_log2;
  }
  return c;
}

inline PartCfg heuristic(int order, int N, int vec, size_t tsize) {
  if (order == 1 && N > 512 && N <= 1024) {
    // strided systems of 513..1024 rows: 16-row chunks keep 8 systems per
    // 512-thread group (8-row chunks leave 4), in the XCD-contiguous group
    // order (15 % -> 40-48 % of HBM bandwidth on MI355X)
    PartCfg c = make_cfg(order, 16, N, 0, 512);
    c.xcd = 1;
    if (cfg_valid(c, N)) return c;
  }
  // preferred chunk lengths: FP64 8 rows, FP32 16 rows (register budget);
  // FP32 along contiguous rows with 16-byte loads: 4 rows, one load per
  // lane and array, consecutive lanes on consecutive 16 bytes (fastest at
  // every length measured on MI355X, exact or not)
  const bool r4 = tsize == 4 && order == 0 && vec;
  const int pref[3] = {r4 ? 4 : (tsize == 8 ? 8 : 16), tsize == 8 ? 16 : 8, r4 ? 16 : 4};
  for (int R : pref) {
    int v = vec && (R * (int)tsize) % 16 == 0;
    PartCfg c = make_cfg(order, R, N, v, order == 0 ? 256 : 512);
    if (cfg_valid(c, N)) return c;
    c = make_cfg(order, R, N, v, 1024);
    if (cfg_valid(c, N)) return c;
  }
  for (int R : pref) {
    for (int t : {64, 128, 256, 512, 1024}) {
      int v = vec && (R * (int)tsize) % 16 == 0;
      PartCfg c = make_cfg(order, R, N, v, t);
      if (cfg_valid(c, N)) return c;
    }
  }
  return split_cfg(order, vec, tsize, 1 << 20);
}

}  // namespace tridhip
