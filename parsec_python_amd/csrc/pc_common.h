// Shared helpers for the PARSEC hot-path HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace pc {

// Thread-local text of the last failure; the C-ABI returns an int status and
// pc_last_error() hands this string to the caller.
void set_error(const std::string& msg);
const char* last_error();

enum Status : int {
  PC_OK = 0,
  PC_ERR_ARG = 1,      // invalid argument (shape, pointer, size)
  PC_ERR_HIP = 2,      // HIP runtime error
  PC_ERR_SOLVER = 3,   // dense solver failure (not SPD, no convergence)
  PC_ERR_LIMIT = 4,    // size above a documented maximum
};

constexpr int kWave = 64;

}  // namespace pc

#define PC_CHECK_ARG(cond, msg)                                   \
  do {                                                            \
    if (!(cond)) {                                                \
      pc::set_error(std::string(__func__) + ": " + (msg));        \
      return pc::PC_ERR_ARG;                                      \
    }                                                             \
  } while (0)

#define PC_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      pc::set_error(std::string(__func__) + ": " #call " -> " + hipGetErrorString(e_)); \
      return pc::PC_ERR_HIP;                                                           \
    }                                                                                  \
  } while (0)

#define PC_LAUNCH_CHECK() PC_HIP(hipGetLastError())

// Block-wide FP64 sum of one value per thread (blockDim.x multiple of 64,
// <= 1024).  Result valid in thread 0.
__device__ inline double pc_block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) r += sh[w];
  }
  return r;
}
