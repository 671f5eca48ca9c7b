// compare.cpp -- compares two raw ADI output files (nx*ny*nz values of FP,
// x fastest, as adi_hip and the reference's ADI apps write them) and prints
// the reference's summary line
//     SumOfDiff = <sum of (file2 - file1)>; Normalized SumOfDiff = <sum / n>
// with the sum accumulated in FP in file order, as apps/adi/src/compare.c
// does (its output format and arithmetic are kept so that validation
// scripts read both alike).  The grid points where the difference is NaN are
// printed as "i j k" lines before it.
// Usage: compare file1.dat file2.dat -nx NX -ny NY -nz NZ
// FPPREC selects the precision at compile time (0 = float, 1 = double).
// Exit status: 0, or 2 when a file cannot be opened.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

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

bool read_file(const char *path, std::vector<FP> &v) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  if (fread(v.data(), sizeof(FP), v.size(), f) != v.size())
    printf("There was an error while reading the file %s!\n", path);
  fclose(f);
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    printf("Please specify the ADI configuration, e.g.: \n$ e.g. ./compare file1.dat file2.dat "
           "-nx NX -ny NY -nz NZ \n");
    return 0;
  }
  int nx = 256, ny = 256, nz = 256;
  for (int i = 3; i < argc; i++) {
    std::string name(argv[i]), value;
    if (name.empty() || name[0] != '-') continue;
    name.erase(0, name.find_first_not_of('-'));
    const size_t eq = name.find('=');
    if (eq != std::string::npos) {
      value = name.substr(eq + 1);
      name = name.substr(0, eq);
    } else if (i + 1 < argc) {
      value = argv[i + 1];
    }
    if (name == "nx") nx = atoi(value.c_str());
    if (name == "ny") ny = atoi(value.c_str());
    if (name == "nz") nz = atoi(value.c_str());
  }
  printf("\nGrid dimensions: %d x %d x %d\n", nx, ny, nz);
  const size_t n = (size_t)nx * ny * nz;
  std::vector<FP> u1(n, FP(0)), u2(n, FP(0));
  printf("Opening file: %s \n", argv[1]);
  printf("Opening file: %s \n", argv[2]);
  if (!read_file(argv[2], u2) || !read_file(argv[1], u1)) {
    printf("Cannot open the input files.\n");
    return 2;
  }
  FP sum = FP(0);
  size_t ind = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++, ind++) {
        const FP diff = u2[ind] - u1[ind];
        if (std::isnan(diff)) printf("%d %d %d\n", i, j, k);
        sum += diff;
      }
  printf("\nSumOfDiff = %e; Normalized SumOfDiff = %e \n", (double)sum,
         (double)(sum / (FP)(nx * ny * nz)));
  return 0;
}
