// FP64 dense k x k stage of Rayleigh-Ritz, on the GPU (rocSOLVER / rocBLAS).
//
// The reference orthonormalises the filtered block with a Householder QR and
// diagonalises G = W^T H W (src/Eigensolvers/chefsi1.py:43-67; GPU twin
// chefsi1_gpu.py:54-66).  With S = Y^T Y and Gy = Y^T (H Y) from the MFMA Gram
// kernels, the same Ritz pairs follow from the Cholesky factor S = R^T R:
//   W = Y R^-1 (orthonormal), G = R^-T Gy R^-1, G = Q diag(lam) Q^T,
//   Ritz vectors = W Q = Y (R^-1 Q) = Y T.
// This file computes T and lam; the caller applies T with one rotation GEMM.
#include "pc_common.h"
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <vector>

namespace {

struct DenseCtx {
  rocblas_handle h;
  hipStream_t stream;
  int kmax;
  double* tmp;      // kmax*kmax
  double* E;        // kmax (syevd workspace)
  double* out;      // kmax + 4 : eigenvalues then 3 info codes (as doubles)
  int* info;        // 4
  double* host_out; // pinned, kmax + 4
  int method;       // 0: divide and conquer (syevd), 1: Jacobi (syevj)
  double* resid;    // syevj: off-diagonal norm at exit
  int* sweeps;      // syevj: sweeps used
};

__global__ void symmetrize(double* G, int k, int ldg) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k * k) return;
  const int i = e % k, j = e / k;
  if (i < j) {
    const double a = G[(long long)j * ldg + i], b = G[(long long)i * ldg + j];
    const double m = 0.5 * (a + b);
    G[(long long)j * ldg + i] = m;
    G[(long long)i * ldg + j] = m;
  }
}

__global__ void zero_info(int* info) {
  if (threadIdx.x < 4) info[threadIdx.x] = 0;
}

__global__ void pack_out(const int* info, double* out, int k) {
  if (threadIdx.x < 3) out[k + threadIdx.x] = (double)info[threadIdx.x];
}

#define PC_RB(call)                                                                          \
  do {                                                                                       \
    rocblas_status s_ = (call);                                                              \
    if (s_ != rocblas_status_success) {                                                      \
      pc::set_error(std::string(__func__) + ": " #call " -> rocblas status " + std::to_string((int)s_)); \
      return pc::PC_ERR_SOLVER;                                                              \
    }                                                                                        \
  } while (0)

}  // namespace

extern "C" {

int pc_dense_create(void* stream, int kmax, void** out_ctx) {
  PC_CHECK_ARG(out_ctx && kmax >= 1, "bad argument");
  DenseCtx* c = new DenseCtx();
  c->stream = static_cast<hipStream_t>(stream);
  c->kmax = kmax;
  if (rocblas_create_handle(&c->h) != rocblas_status_success) {
    delete c;
    pc::set_error("pc_dense_create: rocblas_create_handle failed");
    return pc::PC_ERR_SOLVER;
  }
  rocblas_set_stream(c->h, c->stream);
  rocblas_set_pointer_mode(c->h, rocblas_pointer_mode_host);
  PC_HIP(hipMalloc(&c->tmp, (size_t)kmax * kmax * sizeof(double)));
  PC_HIP(hipMalloc(&c->E, (size_t)kmax * sizeof(double)));
  PC_HIP(hipMalloc(&c->out, (size_t)(kmax + 4) * sizeof(double)));
  PC_HIP(hipMalloc(&c->info, 4 * sizeof(int)));
  PC_HIP(hipMalloc(&c->resid, sizeof(double)));
  PC_HIP(hipMalloc(&c->sweeps, sizeof(int)));
  PC_HIP(hipHostMalloc(&c->host_out, (size_t)(kmax + 4) * sizeof(double), hipHostMallocDefault));
  c->method = 0;
  *out_ctx = c;
  return pc::PC_OK;
}

int pc_dense_destroy(void* ctx) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  if (!c) return pc::PC_OK;
  rocblas_destroy_handle(c->h);
  (void)hipFree(c->tmp); (void)hipFree(c->E); (void)hipFree(c->out); (void)hipFree(c->info);
  (void)hipFree(c->resid); (void)hipFree(c->sweeps);
  (void)hipHostFree(c->host_out);
  delete c;
  return pc::PC_OK;
}

// S (k x k, ld lds) and Gy (k x k, ld ldg) are overwritten.  On return T
// (k x k, ld ldt, device) holds R^-1 Q and lam_host[0..k) the ascending Ritz
// values.  Synchronises the stream once (to read info codes and lam).
int pc_dense_rr(void* ctx, int k, double* S, int lds, double* Gy, int ldg, double* T, int ldt, double* lam_dev,
                double* lam_host) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && S && Gy && T && lam_dev, "null pointer");
  PC_CHECK_ARG(k >= 1 && k <= c->kmax, "k out of range");
  PC_CHECK_ARG(lds >= k && ldg >= k && ldt >= k, "leading dimension too small");
  const double one = 1.0;
  hipLaunchKernelGGL(zero_info, dim3(1), dim3(64), 0, c->stream, c->info);
  PC_LAUNCH_CHECK();
  PC_RB(rocsolver_dpotrf(c->h, rocblas_fill_upper, k, S, lds, c->info));
  PC_RB(rocsolver_dtrtri(c->h, rocblas_fill_upper, rocblas_diagonal_non_unit, k, S, lds, c->info + 1));
  // tmp = Gy R^-1
  PC_RB(rocblas_dtrmm(c->h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                      rocblas_diagonal_non_unit, k, k, &one, S, lds, Gy, ldg, c->tmp, k));
  // Gy = R^-T tmp
  PC_RB(rocblas_dtrmm(c->h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                      rocblas_diagonal_non_unit, k, k, &one, S, lds, c->tmp, k, Gy, ldg));
  hipLaunchKernelGGL(symmetrize, dim3((k * k + 255) / 256), dim3(256), 0, c->stream, Gy, k, ldg);
  PC_LAUNCH_CHECK();
  if (c->method == 1)
    PC_RB(rocsolver_dsyevj(c->h, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, k, Gy, ldg,
                           0.0, c->resid, 100, c->sweeps, c->out, c->info + 2));
  else
    PC_RB(rocsolver_dsyevd(c->h, rocblas_evect_original, rocblas_fill_upper, k, Gy, ldg, c->out, c->E,
                           c->info + 2));
  // T = R^-1 Q
  PC_RB(rocblas_dtrmm(c->h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                      rocblas_diagonal_non_unit, k, k, &one, S, lds, Gy, ldg, T, ldt));
  PC_HIP(hipMemcpyAsync(lam_dev, c->out, k * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(pack_out, dim3(1), dim3(64), 0, c->stream, c->info, c->out, k);
  PC_LAUNCH_CHECK();
  PC_HIP(hipMemcpyAsync(c->host_out, c->out, (k + 3) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PC_HIP(hipStreamSynchronize(c->stream));
  const int i0 = (int)c->host_out[k], i1 = (int)c->host_out[k + 1], i2 = (int)c->host_out[k + 2];
  if (i0 != 0) {
    pc::set_error("pc_dense_rr: overlap matrix not positive definite (potrf info " + std::to_string(i0) + ")");
    return pc::PC_ERR_SOLVER;
  }
  if (i1 != 0 || i2 != 0) {
    pc::set_error("pc_dense_rr: trtri/syevd failed (info " + std::to_string(i1) + "," + std::to_string(i2) + ")");
    return pc::PC_ERR_SOLVER;
  }
  if (lam_host)
    for (int i = 0; i < k; ++i) lam_host[i] = c->host_out[i];
  return pc::PC_OK;
}

// Eigensolver of the projected k x k problem: 0 = divide and conquer
// (rocSOLVER syevd), 1 = cyclic Jacobi (rocSOLVER syevj, tolerance = machine
// precision, at most 100 sweeps).
int pc_dense_set_method(void* ctx, int method) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c, "null context");
  PC_CHECK_ARG(method == 0 || method == 1, "method must be 0 (syevd) or 1 (syevj)");
  c->method = method;
  return pc::PC_OK;
}

// Symmetric eigen-decomposition of a small FP64 matrix in place (eigenvectors
// overwrite A, ascending eigenvalues to w_host).  Used for the Lanczos
// tridiagonal and tests.
int pc_dense_syevd(void* ctx, int k, double* A, int lda, double* w_host) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && A, "null pointer");
  PC_CHECK_ARG(k >= 1 && k <= c->kmax && lda >= k, "bad size");
  hipLaunchKernelGGL(zero_info, dim3(1), dim3(64), 0, c->stream, c->info);
  PC_LAUNCH_CHECK();
  PC_RB(rocsolver_dsyevd(c->h, rocblas_evect_original, rocblas_fill_upper, k, A, lda, c->out, c->E, c->info + 2));
  hipLaunchKernelGGL(pack_out, dim3(1), dim3(64), 0, c->stream, c->info, c->out, k);
  PC_LAUNCH_CHECK();
  PC_HIP(hipMemcpyAsync(c->host_out, c->out, (k + 3) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PC_HIP(hipStreamSynchronize(c->stream));
  if ((int)c->host_out[k + 2] != 0) {
    pc::set_error("pc_dense_syevd: no convergence");
    return pc::PC_ERR_SOLVER;
  }
  if (w_host)
    for (int i = 0; i < k; ++i) w_host[i] = c->host_out[i];
  return pc::PC_OK;
}

}  // extern "C"
