// Hartree solve: conjugate gradients on A x = rhs, A = -lap/h^2 on the grid,
// FP64, no preconditioner (src/Eigensolvers/pcg.py:7-85; default cg_prec=0,
// src/rsdft_models.py:31).  Stops when r.r <= tol^2 * r0.r0 or after maxit
// iterations.  With alias_first=1 the second search direction is (1+beta) r_1,
// which is what the reference's CPU loop computes because its p and r are the
// same array in the first iteration (pcg.py:45-77).
//
// All scalars stay on the device; each kernel checks the device "done" flag,
// so the host only reads the flag back every `check_every` iterations.
#include "pc_common.h"

int pc_internal_stencil_direct_f64(void* grid, const double* Y, const double* X, double* Out, const double* V,
                                   double scale, double ca, double cb, double cc, double* dot_part, void* stream,
                                   int* out_nblocks);
long long pc_internal_grid_npts(void* grid);
long long pc_internal_grid_nrow(void* grid);

namespace {

constexpr int kBlk = 256;
enum { S_RO = 0, S_TOL1, S_DONE, S_ITS, S_ALPHA, S_BETA, S_MAXIT, S_COUNT = 8 };

__device__ double sum_parts(const double* part, int n, double* red) {
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += kBlk) s += part[b];
  return pc_block_sum(s, red);
}

__global__ __launch_bounds__(kBlk) void cg_init_scalars(const double* part, int n, double tol, int maxit, double* sc) {
  __shared__ double red[4];
  const double ro = sum_parts(part, n, red);
  if (threadIdx.x == 0) {
    sc[S_RO] = ro;
    sc[S_TOL1] = tol * tol * ro;
    sc[S_ITS] = 0.0;
    sc[S_MAXIT] = (double)maxit;
    sc[S_DONE] = (maxit <= 0 || !(ro > tol * tol * ro)) ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(kBlk) void cg_copy_rr(long long n, const double* __restrict__ r, double* __restrict__ p,
                                                   double* __restrict__ part) {
  __shared__ double red[4];
  const long long i = (long long)blockIdx.x * kBlk + threadIdx.x;
  double s = 0.0;
  if (i < n) {
    const double v = r[i];
    p[i] = v;
    s = v * v;
  }
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlk) void cg_alpha(const double* part, int n, double* sc) {
  __shared__ double red[4];
  if (sc[S_DONE] != 0.0) return;
  const double pap = sum_parts(part, n, red);
  if (threadIdx.x == 0) sc[S_ALPHA] = sc[S_RO] / pap;
}

__global__ __launch_bounds__(kBlk) void cg_update_xr(long long n, const double* __restrict__ sc, double* __restrict__ x,
                                                     double* __restrict__ r, const double* __restrict__ p,
                                                     const double* __restrict__ ap, double* __restrict__ part) {
  __shared__ double red[4];
  if (sc[S_DONE] != 0.0) return;
  const double alpha = sc[S_ALPHA];
  const long long i = (long long)blockIdx.x * kBlk + threadIdx.x;
  double s = 0.0;
  if (i < n) {
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * ap[i];
    r[i] = rv;
    s = rv * rv;
  }
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlk) void cg_beta(const double* part, int n, double* sc) {
  __shared__ double red[4];
  if (sc[S_DONE] != 0.0) return;
  const double ro1 = sum_parts(part, n, red);
  if (threadIdx.x == 0) {
    const double its = sc[S_ITS] + 1.0;
    sc[S_BETA] = ro1 / sc[S_RO];
    sc[S_RO] = ro1;
    sc[S_ITS] = its;
    // the next p update still runs for this iteration; done takes effect after it
    sc[S_DONE] = (its >= sc[S_MAXIT] || !(ro1 > sc[S_TOL1])) ? 2.0 : 0.0;
  }
}

__global__ __launch_bounds__(kBlk) void cg_update_p(long long n, double* __restrict__ sc, const double* __restrict__ r,
                                                    double* __restrict__ p, int alias_first) {
  const double done = sc[S_DONE];
  if (done == 1.0) return;
  const double beta = sc[S_BETA];
  const bool first = alias_first && sc[S_ITS] == 1.0;
  const long long i = (long long)blockIdx.x * kBlk + threadIdx.x;
  if (i < n) {
    const double rv = r[i];
    p[i] = rv + beta * (first ? rv : p[i]);
  }
}

__global__ void cg_finish_iter(double* sc) {
  if (threadIdx.x == 0 && sc[S_DONE] == 2.0) sc[S_DONE] = 1.0;
}

}  // namespace

extern "C" {

int pc_pcg_workspace(void* grid, size_t* out_bytes) {
  PC_CHECK_ARG(grid && out_bytes, "null pointer");
  const long long n = pc_internal_grid_npts(grid);
  const long long nb_st = (pc_internal_grid_nrow(grid) + 3) / 4;
  const long long nb_vec = (n + kBlk - 1) / kBlk;
  *out_bytes = (size_t)(3 * n + nb_st + nb_vec + S_COUNT) * sizeof(double);
  return pc::PC_OK;
}

// x (in: warm start, out: solution) and rhs are device FP64 arrays of npts.
// inv_h2 = 1/h^2.  its_out receives the iteration count (host); res_out (host,
// optional, 2 doubles) receives the final r.r and tol^2 * r0.r0.
int pc_pcg_solve(void* grid, double inv_h2, const double* rhs, double* x, int maxit, double tol, int alias_first,
                 int check_every, void* work, size_t work_bytes, int* its_out, double* res_out, void* stream) {
  PC_CHECK_ARG(grid && rhs && x && work && its_out, "null pointer");
  size_t need = 0;
  pc_pcg_workspace(grid, &need);
  PC_CHECK_ARG(work_bytes >= need, "workspace too small");
  if (check_every < 1) check_every = 8;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long n = pc_internal_grid_npts(grid);
  const int nb_vec = (int)((n + kBlk - 1) / kBlk);
  double* r = static_cast<double*>(work);
  double* p = r + n;
  double* ap = p + n;
  const long long nb_st_max = (pc_internal_grid_nrow(grid) + 3) / 4;
  double* part_a = ap + n;
  double* part_b = part_a + nb_st_max;
  double* sc = part_b + nb_vec;
  int nb_st = 0;
  // r = rhs - A x
  int s = pc_internal_stencil_direct_f64(grid, x, rhs, r, x, -inv_h2, 0.0, 0.0, 1.0, nullptr, stream, &nb_st);
  if (s) return s;
  PC_CHECK_ARG(nb_st <= nb_st_max, "partials overflow");
  hipLaunchKernelGGL(cg_copy_rr, dim3(nb_vec), dim3(kBlk), 0, st, n, r, p, part_b);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(cg_init_scalars, dim3(1), dim3(kBlk), 0, st, part_b, nb_vec, tol, maxit, sc);
  PC_LAUNCH_CHECK();
  // pinned readback buffer, allocated once per host thread and kept
  static thread_local double* host = nullptr;
  if (!host) PC_HIP(hipHostMalloc(&host, 4 * sizeof(double), hipHostMallocDefault));
  int status = pc::PC_OK;
  for (int it0 = 0; it0 < maxit; it0 += check_every) {
    for (int t = 0; t < check_every && it0 + t < maxit; ++t) {
      status = pc_internal_stencil_direct_f64(grid, p, nullptr, ap, x, inv_h2, 0.0, 0.0, 0.0, part_a, stream, &nb_st);
      if (status) break;
      hipLaunchKernelGGL(cg_alpha, dim3(1), dim3(kBlk), 0, st, part_a, nb_st, sc);
      hipLaunchKernelGGL(cg_update_xr, dim3(nb_vec), dim3(kBlk), 0, st, n, sc, x, r, p, ap, part_b);
      hipLaunchKernelGGL(cg_beta, dim3(1), dim3(kBlk), 0, st, part_b, nb_vec, sc);
      hipLaunchKernelGGL(cg_update_p, dim3(nb_vec), dim3(kBlk), 0, st, n, sc, r, p, alias_first);
      hipLaunchKernelGGL(cg_finish_iter, dim3(1), dim3(64), 0, st, sc);
      if (hipGetLastError() != hipSuccess) {
        pc::set_error("pc_pcg_solve: kernel launch failed");
        status = pc::PC_ERR_HIP;
        break;
      }
    }
    if (status) break;
    if (hipMemcpyAsync(host, sc, 4 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      pc::set_error("pc_pcg_solve: readback failed");
      status = pc::PC_ERR_HIP;
      break;
    }
    if (host[S_DONE] != 0.0) break;
  }
  if (!status) {
    if (hipMemcpyAsync(host, sc, 4 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      pc::set_error("pc_pcg_solve: readback failed");
      status = pc::PC_ERR_HIP;
    } else {
      *its_out = (int)host[S_ITS];
      if (res_out) {
        res_out[0] = host[S_RO];
        res_out[1] = host[S_TOL1];
      }
    }
  }
  return status;
}

}  // extern "C"
