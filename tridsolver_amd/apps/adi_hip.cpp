// adi_hip.cpp -- command-line ADI application on tridsolver_amd, for the
// validation scripts written around the reference's ADI apps.
//
// Runs the heat-diffusion ADI time steps of apps/adi/src/adi_cuda.cu:283-330
// (u = 1 on the boundary and 0 inside, lambda = 1, the reference's
// coefficients of apps/adi/src/preproc.cu:70-72) through the library's ADI
// plan (trid_adi.h), then reports as the reference apps do:
//   * "end h_u sum = <sum of u>"                (adi_cuda.cu:88-101, 342);
//   * with -dump 1 (default) the two corner tables of
//     apps/adi/src/print_array.c and the raw array in <argv[0]>.dat
//     (nx*ny*nz values of FP, x fastest), the file apps/adi/src/compare.c
//     reads;
//   * the timing summary of adi_cuda.cu:375-388: -prof 0 the mean time per
//     iteration, -prof 1 (default) the time per element of the whole step
//     and of its four stages (right-hand side, x, y, z sweeps).
// Options (-name value or -name=value): -devid, -nx, -ny, -nz, -iter,
// -optx/-opty/-optz (accepted; the library picks its kernels), -prof, -dump,
// -help.  FPPREC selects the precision at compile time as in the reference
// (0 = float, 1 = double).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trid_adi.h"
#include "../../include/trid_hip_ext.h"

#ifndef FPPREC
#error "define FPPREC (0 = float, 1 = double)"
#endif
#if FPPREC == 0
typedef float FP;
#elif FPPREC == 1
typedef double FP;
#else
#error "FPPREC must be 0 or 1"
#endif

namespace {

struct Options {
  int devid = 0, nx = 256, ny = 256, nz = 256, iter = 10, prof = 1, dump = 1;
};

[[noreturn]] void usage(const char *prog, int code) {
  printf("Usage: %s -nx NX -ny NY -nz NZ -iter ITER [-devid ID] [-prof 0|1] [-dump 0|1]\n",
         prog);
  exit(code);
}

Options parse(int argc, char **argv) {
  Options o;
  struct {
    const char *name;
    int *dst;
  } table[] = {{"devid", &o.devid}, {"nx", &o.nx},     {"ny", &o.ny},     {"nz", &o.nz},
               {"iter", &o.iter},   {"prof", &o.prof}, {"dump", &o.dump}, {"optx", nullptr},
               {"opty", nullptr},   {"optz", nullptr}};
  for (int i = 1; i < argc; i++) {
    const char *arg = argv[i];
    if (arg[0] != '-') usage(argv[0], 1);
    while (*arg == '-') arg++;
    std::string name(arg), value;
    const size_t eq = name.find('=');
    if (eq != std::string::npos) {
      value = name.substr(eq + 1);
      name = name.substr(0, eq);
    }
    if (name == "help" || name == "h") usage(argv[0], 0);
    bool known = false;
    for (auto &t : table) {
      if (name != t.name) continue;
      known = true;
      if (eq == std::string::npos) {
        if (i + 1 >= argc) usage(argv[0], 1);
        value = argv[++i];
      }
      if (t.dst) *t.dst = atoi(value.c_str());
    }
    if (!known) usage(argv[0], 1);
  }
  return o;
}

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    exit(1);
  }
}

void check(tridStatus_t s, const char *what) {
  if (s != TRID_STATUS_SUCCESS) {
    fprintf(stderr, "%s: %s (%s)\n", what, tridGetStatusString(s), tridGetLastError());
    exit(1);
  }
}

// The corner tables of print_array.c: the first min(nx,17) and the last
// min(nx,17) columns of the first min(ny,17) rows of planes k = 0, 1.
void print_corners(const std::vector<FP> &u, int nx, int ny, int nz) {
  for (int side = 0; side < 2; side++) {
    const int i0 = side == 0 ? 0 : std::max(0, nx - 17);
    const int i1 = side == 0 ? std::min(nx, 17) : nx;
    for (int k = 0; k < std::min(nz, 2); k++) {
      printf("k = %i\n", k);
      for (int j = 0; j < std::min(ny, 17); j++) {
        printf(" %d   ", j);
        for (int i = i0; i < i1; i++)
          printf(" %5.5g ", (double)u[(size_t)k * nx * ny + (size_t)j * nx + i]);
        printf("\n");
      }
      printf("\n");
    }
  }
}

}  // namespace

int main(int argc, char **argv) {
  const Options o = parse(argc, argv);
  printf("\nGrid dimensions: %d x %d x %d\n", o.nx, o.ny, o.nz);
  if (o.nx < 1 || o.ny < 1 || o.nz < 1 || o.iter < 0) usage(argv[0], 1);
  check(hipSetDevice(o.devid), "hipSetDevice");

  const size_t n = (size_t)o.nx * o.ny * o.nz;
  std::vector<FP> h_u(n);
  for (int k = 0; k < o.nz; k++)
    for (int j = 0; j < o.ny; j++)
      for (int i = 0; i < o.nx; i++) {
        const bool edge = i == 0 || i == o.nx - 1 || j == 0 || j == o.ny - 1 || k == 0 ||
                          k == o.nz - 1;
        h_u[(size_t)k * o.nx * o.ny + (size_t)j * o.nx + i] = edge ? FP(1) : FP(0);
      }

  FP *d_u = nullptr, *d_du = nullptr;
  check(hipMalloc(&d_u, n * sizeof(FP)), "hipMalloc");
  check(hipMalloc(&d_du, n * sizeof(FP)), "hipMalloc");
  check(hipMemset(d_du, 0, n * sizeof(FP)), "hipMemset");
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  auto elapsed = [&]() {
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    return ms * 1e-3;
  };
  hipStream_t st = static_cast<hipStream_t>(tridGetStream());

  check(hipEventRecord(e0, st), "hipEventRecord");
  check(hipMemcpyAsync(d_u, h_u.data(), n * sizeof(FP), hipMemcpyHostToDevice, st), "copy");
  check(hipEventRecord(e1, st), "hipEventRecord");
  printf("\nCopy u to device: %f (s) \n", elapsed());

  tridAdiPlan_t plan = nullptr;
  check(tridAdiPlanCreate(&plan, o.nx, o.ny, o.nz, FPPREC == 1, 1.0, 0), "tridAdiPlanCreate");

  // stage times of -prof 1: right-hand side, x, y, z sweeps
  double t_stage[4] = {0, 0, 0, 0}, t_total = 0;
  if (o.prof == 1) {
    for (int it = 0; it < o.iter; it++)
      for (int s = 0; s < 4; s++) {
        check(hipEventRecord(e0, st), "hipEventRecord");
        check(tridAdiStage(plan, d_u, d_du, s), "tridAdiStage");
        check(hipEventRecord(e1, st), "hipEventRecord");
        const double t = elapsed();
        t_stage[s] += t;
        t_total += t;
      }
  } else {
    check(hipEventRecord(e0, st), "hipEventRecord");
    if (o.iter > 0) check(tridAdiStep(plan, d_u, d_du, o.iter, 0), "tridAdiStep");
    check(hipEventRecord(e1, st), "hipEventRecord");
    t_total = elapsed();
  }
  printf("\nComputing ADI on GPU: %f (s) \n", t_total);

  check(hipEventRecord(e0, st), "hipEventRecord");
  check(hipMemcpyAsync(h_u.data(), d_u, n * sizeof(FP), hipMemcpyDeviceToHost, st), "copy");
  check(hipEventRecord(e1, st), "hipEventRecord");
  printf("\nCopy u to host: %f (s) \n", elapsed());

  double sum = 0.0;
  for (size_t i = 0; i < n; i++) sum += h_u[i];
  printf("%s sum = %lg\n", "end h_u", sum);

  if (o.dump) {
    print_corners(h_u, o.nx, o.ny, o.nz);
    const std::string path = std::string(argv[0]) + ".dat";
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
      printf("ERROR: File stream could not be opened. Data will not be written to file!\n");
    } else {
      if (fwrite(h_u.data(), sizeof(FP), n, f) != n) printf("ERROR: short write to %s\n", path.c_str());
      fclose(f);
    }
  }

  check(tridAdiPlanDestroy(plan), "tridAdiPlanDestroy");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d_u);
  (void)hipFree(d_du);
  printf("Done.\n");

  const double pts = (double)n;
  if (o.prof == 0) {
    printf("Avg(per iter) \n[total]\n");
    printf("%f\n", o.iter ? t_total / o.iter : 0.0);
  } else if (o.prof == 1) {
    const double it = o.iter ? (double)o.iter : 1.0;
    printf("Time per element averaged on %d iterations: \n[total] \t[prepro] \t[trid_x] "
           "\t[trid_y] \t[trid_z]\n", o.iter);
    printf("%e \t%e \t%e \t%e \t%e\n", (t_total / it) / pts, (t_stage[0] / it) / pts,
           (t_stage[1] / it) / pts, (t_stage[2] / it) / pts, (t_stage[3] / it) / pts);
  }
  return 0;
}
