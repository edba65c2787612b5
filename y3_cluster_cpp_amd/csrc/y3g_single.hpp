// y3g_single.hpp -- device-wide single-integral cubature (the plain path).
#pragma once

#include <hip/hip_runtime.h>

#include "y3g_internal.hpp"

namespace y3g {

constexpr int kSingleKmax = 16384;  // regions split per step, plain path

struct SingleArgs {
  KernArgs K;                  // tables and points
  const y3g_integral* ig;      // the integral (device)
  double *lo, *hi, *est, *err; // regions: lo/hi [kMaxDim][cap]
  int8_t* ax;
  int cap;
  int* list;                   // regions to evaluate (2 * kmax)
  int* sel;                    // selected parents (kmax)
  double *pest, *perr;         // parents' est / err (kmax)
  double* partial;             // 2 * 256 block partial sums
  double* IE;                  // I, E
  int* hc;                     // 64 error-bin counts
  unsigned long long* hq;      // 64 error-bin masses
  int* blockcnt;               // 2 per compaction tile
  int* blockoff;               // 4 per compaction tile
};

hipError_t run_single(SingleArgs A, const y3g_integral& ig, int precision, int kmax,
                      int init_split, int init_lnm, int eval_grid, hipStream_t s,
                      y3g_result* out, long long* nevals);

}  // namespace y3g
