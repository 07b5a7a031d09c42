// Point-wise SCF kernels: density from occupied states, LDA XC + potential
// update + energy/error sums, Hartree right-hand side, deterministic
// reductions, the mixer's multi-dot / linear-combination kernels, type
// conversion and the counter-based random start vectors.
#include "pc_common.h"

namespace {

constexpr int kBlk = 256;

// ---------------------------------------------------------------- density
// rho[p] = sum_{v<nev} occ2[v] * W[v*ld + p]^2 (src/rsdft_solver.py:324).
template <typename T>
__global__ __launch_bounds__(kBlk) void density_kernel(const T* __restrict__ W, long long ld, long long n, int nev,
                                                       const double* __restrict__ occ2, double* __restrict__ rho) {
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  if (p >= n) return;
  double s = 0.0;
  // unrolled so that 16 loads per thread are in flight (the sum keeps its order)
#pragma unroll 16
  for (int v = 0; v < nev; ++v) {
    const double w = (double)W[(long long)v * ld + p];
    s += occ2[v] * (w * w);
  }
  rho[p] = s;
}

// ---------------------------------------------------------------- XC + potential
// LDA, Ceperley-Alder / Perdew-Zunger in Ry (src/V_xc/exc_nspn.py:22-91).
__device__ __forceinline__ void lda_pz(double rho, double& vxc, double& e) {
  vxc = 0.0;
  e = 0.0;
  if (rho > 0.0) {
    const double third = 1.0 / 3.0;
    const double a0 = 0.52106176119784611;  // (4/(9 pi))^(1/3)
    const double twovpia0 = 2.0 / (M_PI * a0);
    const double p75vpi = 0.75 / M_PI;
    const double rs = cbrt(p75vpi / rho);
    double v = -twovpia0 / rs;
    e = 0.75 * rho * v;
    double ec;
    if (rs >= 1.0) {
      const double g = -0.2846, b1 = 1.0529, b2 = 0.3334;
      const double sq = sqrt(rs);
      ec = g / (1.0 + b1 * sq + b2 * rs);
      v += (ec * ec) * (1.0 + 3.5 * b1 * sq * third + 4.0 * b2 * rs * third) / g;
    } else {
      const double c1 = 0.0622, c2 = 0.096, c3 = 0.004, c4 = 0.0232, c5 = 0.0192;
      const double al = log(rs);
      ec = c1 * al - c2 + (c3 * al - c4) * rs;
      v += ec - (c1 + (c3 * al - c5) * rs) * third;
    }
    e += rho * ec;
    vxc = v;
  }
}

// rho_cell in electrons/cell -> rho = rho_cell / h3 (e/bohr^3, written to rho_out),
// vxc [Ry], pot_new = ppot + 0.5 vxc + hpot + hpot0 (src/rsdft_solver.py:328,350-351),
// partial sums per block: [E_xc/h3, |pot_new-pot|^2, |pot_new|^2, sum rho (hpot+hpot0), sum rho vxc]
__global__ __launch_bounds__(kBlk) void scf_post_kernel(long long n, const double* __restrict__ rho_cell, double h3,
                                                        const double* __restrict__ ppot,
                                                        const double* __restrict__ hpot,
                                                        const double* __restrict__ hpot0,
                                                        const double* __restrict__ pot_old, double* __restrict__ rho_out,
                                                        double* __restrict__ vxc_out, double* __restrict__ pot_new,
                                                        double* __restrict__ part, int nblk) {
  __shared__ double red[kBlk / 64];
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  double s[5] = {0, 0, 0, 0, 0};
  if (p < n) {
    const double rho = rho_cell[p] / h3;
    double v, e;
    lda_pz(rho, v, e);
    const double hh = hpot[p], h0 = hpot0[p];
    const double pn = ppot[p] + 0.5 * v + hh + h0;
    const double d = pn - pot_old[p];
    rho_out[p] = rho;
    vxc_out[p] = v;
    pot_new[p] = pn;
    s[0] = e; s[1] = d * d; s[2] = pn * pn; s[3] = rho * (hh + h0); s[4] = rho * v;
  }
  for (int q = 0; q < 5; ++q) {
    const double t = pc_block_sum(s[q], red);
    if (threadIdx.x == 0) part[(long long)q * nblk + blockIdx.x] = t;
  }
}

// XC only (initial potential): vxc and per-block E_xc partials.
__global__ __launch_bounds__(kBlk) void xc_kernel(long long n, const double* __restrict__ rho, double* __restrict__ vxc,
                                                  double* __restrict__ part) {
  __shared__ double red[kBlk / 64];
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  double e = 0.0;
  if (p < n) {
    double v;
    lda_pz(rho[p], v, e);
    vxc[p] = v;
  }
  const double t = pc_block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlk) void hartree_rhs_kernel(long long n, const double* __restrict__ rho_cell,
                                                           const double* __restrict__ rho0, double fac,
                                                           double* __restrict__ rhs) {
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  if (p < n) rhs[p] = fac * (rho_cell[p] - rho0[p]);
}

// out[s] = sum_b part[s*nblk + b], fixed order -> deterministic.
__global__ __launch_bounds__(kBlk) void reduce_sets_kernel(const double* __restrict__ part, int nblk,
                                                           double* __restrict__ out) {
  __shared__ double red[kBlk / 64];
  const double* ps = part + (long long)blockIdx.x * nblk;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kBlk) s += ps[b];
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// ---------------------------------------------------------------- mixer helpers
// part[i*nblk + b] = sum_{p in chunk b} A[i*lda + p] * v[p]
__global__ __launch_bounds__(kBlk) void dots_kernel(long long n, const double* __restrict__ A, long long lda,
                                                    const double* __restrict__ v, long long chunk,
                                                    double* __restrict__ part) {
  __shared__ double red[kBlk / 64];
  const int i = blockIdx.y;
  const long long p0 = (long long)blockIdx.x * chunk;
  const long long p1 = min(n, p0 + chunk);
  const double* a = A + (long long)i * lda;
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kBlk) s += a[p] * v[p];
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) part[(long long)i * gridDim.x + blockIdx.x] = t;
}

struct Coef64 {
  double c[64];
};

// y = alpha*y + beta*x + sum_i coef[i] * A[i*lda + :]   (x may be null)
__global__ __launch_bounds__(kBlk) void lincomb_kernel(long long n, const double* __restrict__ A, long long lda, int m,
                                                       Coef64 cf, double alpha, const double* __restrict__ x,
                                                       double beta, double* __restrict__ y) {
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  if (p >= n) return;
  double s = alpha * y[p];
  if (x) s += beta * x[p];
#pragma unroll 8
  for (int i = 0; i < m; ++i) s += cf.c[i] * A[(long long)i * lda + p];
  y[p] = s;
}

// The same combination with its scalars read from device memory (results of
// pc_dots, so a caller needs no host round trip): coef[i] = cs * cdev[i];
// alpha = 1 / adev[0] (0 when |adev[0]| < atol |adev[0]|) when adev is set;
// beta = bs * bdev[0] when bdev is set.  Same arithmetic as lincomb_kernel.
__global__ __launch_bounds__(kBlk) void lincomb_dev_kernel(long long n, const double* __restrict__ A, long long lda,
                                                           int m, const double* __restrict__ cdev, double cs,
                                                           double alpha, const double* __restrict__ adev, double atol,
                                                           const double* __restrict__ x, double beta,
                                                           const double* __restrict__ bdev, double bs,
                                                           double* __restrict__ y) {
  __shared__ double cf[64];
  if ((int)threadIdx.x < m) cf[threadIdx.x] = cs * cdev[threadIdx.x];
  __syncthreads();
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  if (p >= n) return;
  double al = alpha;
  if (adev) {
    const double v = adev[0];
    al = fabs(v) < atol * fabs(v) ? 0.0 : 1.0 / v;
  }
  const double be = bdev ? bs * bdev[0] : beta;
  double s = al * y[p];
  if (x) s += be * x[p];
#pragma unroll 8
  for (int i = 0; i < m; ++i) s += cf[i] * A[(long long)i * lda + p];
  y[p] = s;
}

// ---------------------------------------------------------------- conversion
template <typename A, typename B>
__global__ __launch_bounds__(kBlk) void convert_kernel(long long n, int ncol, const A* __restrict__ src, long long lds,
                                                       B* __restrict__ dst, long long ldd) {
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  const int c = blockIdx.y;
  if (p < n && c < ncol) dst[(long long)c * ldd + p] = (B)src[(long long)c * lds + p];
}

// ---------------------------------------------------------------- random start vectors
// splitmix64 stream: element idx of stream `seed` is mix(seed + (idx+1)*gamma);
// uniform = (mix >> 11) * 2^-53 in [0,1).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(unsigned long long seed, unsigned long long idx) {
  return (double)(mix64(seed + (idx + 1ULL) * 0x9E3779B97F4A7C15ULL) >> 11) * (1.0 / 9007199254740992.0);
}

// Column c of the block gets elements [c*n, (c+1)*n) of the stream.
template <typename T>
__global__ __launch_bounds__(kBlk) void rand_kernel(long long n, int ncol, long long ld, unsigned long long seed,
                                                    int normal, T* __restrict__ out) {
  const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
  const int c = blockIdx.y;
  if (p >= n) return;
  const unsigned long long idx = (unsigned long long)c * (unsigned long long)n + (unsigned long long)p;
  double v;
  if (!normal) {
    v = uniform01(seed, idx);
  } else {
    const double u1 = uniform01(seed, 2ULL * idx);
    const double u2 = uniform01(seed, 2ULL * idx + 1ULL);
    v = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
  }
  out[(long long)c * ld + p] = (T)v;
}

inline unsigned blocks(long long n) { return (unsigned)((n + kBlk - 1) / kBlk); }

// ---------------------------------------------------------------- Lanczos
// Device-side three-term Lanczos step (the bound of src/Eigensolvers/
// lanczosForChefsi1.py:5-99) without host round trips: alpha and beta stay in
// coef[] on the device and the host reads all of them once.  One workgroup per
// kLzChunk points; every workgroup sums the partials of the previous stage
// itself, in one fixed order, so all hold the same scalar.
constexpr int kLzChunk = 2048;

inline unsigned lz_blocks(long long n) { return (unsigned)((n + kLzChunk - 1) / kLzChunk); }

__device__ __forceinline__ double lz_sum_parts(const double* __restrict__ part, int nb, double* red) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlk) s += part[b];
  const double t = pc_block_sum(s, red);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = t;
  __syncthreads();
  return bc;
}

// part[b] = sum over chunk b of a[p] * c[p]
__global__ __launch_bounds__(kBlk) void lz_dot(long long n, const double* __restrict__ a, const double* __restrict__ c,
                                               double* __restrict__ part) {
  __shared__ double red[kBlk / 64];
  const long long p0 = (long long)blockIdx.x * kLzChunk, p1 = min(n, p0 + kLzChunk);
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kBlk) s += a[p] * c[p];
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// alpha = sum(pa) (v1 . H v1); w -= alpha v1 + beta v0 (beta = coef[2k-1], 0 at
// k = 0); pb = partials of w . w; coef[2k] = alpha
__global__ __launch_bounds__(kBlk) void lz_orth(long long n, const double* __restrict__ v0,
                                                const double* __restrict__ v1, double* __restrict__ w,
                                                double* __restrict__ coef, int k, const double* __restrict__ pa,
                                                double* __restrict__ pb, int nb) {
  __shared__ double red[kBlk / 64];
  const double alpha = lz_sum_parts(pa, nb, red);
  const double beta = k > 0 ? coef[2 * k - 1] : 0.0;
  const long long p0 = (long long)blockIdx.x * kLzChunk, p1 = min(n, p0 + kLzChunk);
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kBlk) {
    const double x = w[p] - alpha * v1[p] - beta * v0[p];
    w[p] = x;
    s += x * x;
  }
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) {
    pb[blockIdx.x] = t;
    if (blockIdx.x == 0) coef[2 * k] = alpha;
  }
}

// nrm = sqrt(sum(pb)); v /= nrm; *slot = nrm (slot may be null)
__global__ __launch_bounds__(kBlk) void lz_norm(long long n, double* __restrict__ v, const double* __restrict__ pb,
                                                int nb, double* __restrict__ slot) {
  __shared__ double red[kBlk / 64];
  const double nrm = sqrt(lz_sum_parts(pb, nb, red));
  const double inv = 1.0 / nrm;
  const long long p0 = (long long)blockIdx.x * kLzChunk, p1 = min(n, p0 + kLzChunk);
  for (long long p = p0 + threadIdx.x; p < p1; p += kBlk) v[p] *= inv;
  if (slot && blockIdx.x == 0 && threadIdx.x == 0) *slot = nrm;
}

}  // namespace

extern "C" {

int pc_density_f64(const double* W, long long ld, long long n, int nev, const double* occ2, double* rho, void* stream) {
  PC_CHECK_ARG(W && occ2 && rho && n >= 1 && nev >= 0 && ld >= n, "bad argument");
  hipLaunchKernelGGL(density_kernel<double>, dim3(blocks(n)), dim3(kBlk), 0, (hipStream_t)stream, W, ld, n, nev,
                     occ2, rho);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_density_f32(const float* W, long long ld, long long n, int nev, const double* occ2, double* rho, void* stream) {
  PC_CHECK_ARG(W && occ2 && rho && n >= 1 && nev >= 0 && ld >= n, "bad argument");
  hipLaunchKernelGGL(density_kernel<float>, dim3(blocks(n)), dim3(kBlk), 0, (hipStream_t)stream, W, ld, n, nev, occ2,
                     rho);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_num_blocks(long long n, int* out) {
  PC_CHECK_ARG(out && n >= 0, "bad argument");
  *out = (int)blocks(n);
  return pc::PC_OK;
}

// part must hold 5*pc_num_blocks(n) doubles; sums5 (device) receives the 5 totals.
int pc_scf_post(long long n, const double* rho_cell, double h3, const double* ppot, const double* hpot,
                const double* hpot0, const double* pot_old, double* rho_out, double* vxc_out, double* pot_new,
                double* part, double* sums5, void* stream) {
  PC_CHECK_ARG(n >= 1 && rho_cell && ppot && hpot && hpot0 && pot_old && rho_out && vxc_out && pot_new && part && sums5,
               "bad argument");
  const int nb = (int)blocks(n);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(scf_post_kernel, dim3(nb), dim3(kBlk), 0, st, n, rho_cell, h3, ppot, hpot, hpot0, pot_old,
                     rho_out, vxc_out, pot_new, part, nb);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_sets_kernel, dim3(5), dim3(kBlk), 0, st, part, nb, sums5);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// vxc for rho [e/bohr^3]; exc_sum (device, 1 double) = sum of per-point
// energy density (multiply by h^3 for E_xc).  part: pc_num_blocks(n) doubles.
int pc_xc(long long n, const double* rho, double* vxc, double* part, double* exc_sum, void* stream) {
  PC_CHECK_ARG(n >= 1 && rho && vxc && part && exc_sum, "bad argument");
  const int nb = (int)blocks(n);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(xc_kernel, dim3(nb), dim3(kBlk), 0, st, n, rho, vxc, part);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_sets_kernel, dim3(1), dim3(kBlk), 0, st, part, nb, exc_sum);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_hartree_rhs(long long n, const double* rho_cell, const double* rho0, double fac, double* rhs, void* stream) {
  PC_CHECK_ARG(n >= 1 && rho_cell && rho0 && rhs, "bad argument");
  hipLaunchKernelGGL(hartree_rhs_kernel, dim3(blocks(n)), dim3(kBlk), 0, (hipStream_t)stream, n, rho_cell, rho0, fac,
                     rhs);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_reduce_sets(const double* part, int nblk, int nsets, double* out, void* stream) {
  PC_CHECK_ARG(part && out && nblk >= 1 && nsets >= 1, "bad argument");
  hipLaunchKernelGGL(reduce_sets_kernel, dim3(nsets), dim3(kBlk), 0, (hipStream_t)stream, part, nblk, out);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// Workspace for pc_dots: m * pc_dots_nchunk(n) doubles.
int pc_dots_nchunk(long long n, int* out) {
  PC_CHECK_ARG(out && n >= 1, "bad argument");
  // one workgroup per 2048 points (8 per thread): a single dot over 605,808
  // points takes 296 workgroups (10 at 65,536 points each took 70 us)
  long long nch = (n + 2047) / 2048;
  if (nch < 1) nch = 1;
  if (nch > 1024) nch = 1024;
  *out = (int)nch;
  return pc::PC_OK;
}

// out[i] = A_i . v for i < m (device out, deterministic).
int pc_dots(long long n, const double* A, long long lda, int m, const double* v, double* part, double* out,
            void* stream) {
  PC_CHECK_ARG(n >= 1 && A && v && part && out && m >= 1 && m <= 65535 && lda >= n, "bad argument");
  int nch;
  pc_dots_nchunk(n, &nch);
  const long long chunk = (n + nch - 1) / nch;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(dots_kernel, dim3(nch, m), dim3(kBlk), 0, st, n, A, lda, v, chunk, part);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_sets_kernel, dim3(m), dim3(kBlk), 0, st, part, nch, out);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// y = alpha*y + beta*x + sum_{i<m} coef[i] A_i   (coef: host array, m <= 64)
int pc_lincomb(long long n, const double* A, long long lda, int m, const double* coef_host, double alpha,
               const double* x, double beta, double* y, void* stream) {
  PC_CHECK_ARG(n >= 1 && y && m >= 0 && m <= 64, "bad argument");
  PC_CHECK_ARG(m == 0 || (A && coef_host && lda >= n), "bad matrix argument");
  Coef64 cf;
  for (int i = 0; i < 64; ++i) cf.c[i] = i < m ? coef_host[i] : 0.0;
  hipLaunchKernelGGL(lincomb_kernel, dim3(blocks(n)), dim3(kBlk), 0, (hipStream_t)stream, n, A, lda, m, cf, alpha, x,
                     beta, y);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// y = alpha*y + beta*x + sum_{i<m} coef[i] A_i with device scalars (see
// lincomb_dev_kernel); coef_dev may be null when m = 0, alpha_recip_dev and
// beta_dev null to use alpha and beta as given.
int pc_lincomb_dev(long long n, const double* A, long long lda, int m, const double* coef_dev, double coef_scale,
                   double alpha, const double* alpha_recip_dev, double alpha_tol, const double* x, double beta,
                   const double* beta_dev, double beta_scale, double* y, void* stream) {
  PC_CHECK_ARG(n >= 1 && y && m >= 0 && m <= 64, "bad argument");
  PC_CHECK_ARG(m == 0 || (A && coef_dev && lda >= n), "bad matrix argument");
  PC_CHECK_ARG(!beta_dev || x, "beta_dev without x");
  hipLaunchKernelGGL(lincomb_dev_kernel, dim3(blocks(n)), dim3(kBlk), 0, (hipStream_t)stream, n, A, lda, m, coef_dev,
                     coef_scale, alpha, alpha_recip_dev, alpha_tol, x, beta, beta_dev, beta_scale, y);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_convert_f64_to_f32(long long n, int ncol, const double* src, long long lds, float* dst, long long ldd,
                          void* stream) {
  PC_CHECK_ARG(n >= 1 && ncol >= 1 && ncol <= 65535 && src && dst, "bad argument");
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks(n), ncol), dim3(kBlk), 0, (hipStream_t)stream, n,
                     ncol, src, lds, dst, ldd);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_convert_f32_to_f64(long long n, int ncol, const float* src, long long lds, double* dst, long long ldd,
                          void* stream) {
  PC_CHECK_ARG(n >= 1 && ncol >= 1 && ncol <= 65535 && src && dst, "bad argument");
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks(n), ncol), dim3(kBlk), 0, (hipStream_t)stream, n,
                     ncol, src, lds, dst, ldd);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// Lanczos workspace: 2 * pc_lanczos_nblocks(n) doubles.
int pc_lanczos_nblocks(long long n, int* out) {
  PC_CHECK_ARG(out && n >= 1, "bad argument");
  *out = (int)lz_blocks(n);
  return pc::PC_OK;
}

// v <- v / ||v|| (device only).
int pc_lanczos_start(long long n, double* v, double* part, void* stream) {
  PC_CHECK_ARG(n >= 1 && v && part, "bad argument");
  const unsigned nb = lz_blocks(n);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(lz_dot, dim3(nb), dim3(kBlk), 0, st, n, v, v, part);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(lz_norm, dim3(nb), dim3(kBlk), 0, st, n, v, part, (int)nb, (double*)nullptr);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// Step k (0-based) after w = H v1: alpha_k = v1 . w, w -= alpha_k v1 + beta_{k-1} v0,
// beta_k = ||w||, w /= beta_k; coef[2k] = alpha_k, coef[2k+1] = beta_k (device).
int pc_lanczos_step(long long n, const double* v0, const double* v1, double* w, double* coef, int k, double* part,
                    void* stream) {
  PC_CHECK_ARG(n >= 1 && v0 && v1 && w && coef && part && k >= 0, "bad argument");
  const unsigned nb = lz_blocks(n);
  hipStream_t st = (hipStream_t)stream;
  double* pa = part;
  double* pb = part + nb;
  hipLaunchKernelGGL(lz_dot, dim3(nb), dim3(kBlk), 0, st, n, v1, w, pa);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(lz_orth, dim3(nb), dim3(kBlk), 0, st, n, v0, v1, w, coef, k, pa, pb, (int)nb);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(lz_norm, dim3(nb), dim3(kBlk), 0, st, n, w, pb, (int)nb, coef + 2 * k + 1);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_rand_f64(long long n, int ncol, long long ld, unsigned long long seed, int normal, double* out, void* stream) {
  PC_CHECK_ARG(n >= 1 && ncol >= 1 && ncol <= 65535 && out && ld >= n, "bad argument");
  hipLaunchKernelGGL(rand_kernel<double>, dim3(blocks(n), ncol), dim3(kBlk), 0, (hipStream_t)stream, n, ncol, ld, seed,
                     normal, out);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_rand_f32(long long n, int ncol, long long ld, unsigned long long seed, int normal, float* out, void* stream) {
  PC_CHECK_ARG(n >= 1 && ncol >= 1 && ncol <= 65535 && out && ld >= n, "bad argument");
  hipLaunchKernelGGL(rand_kernel<float>, dim3(blocks(n), ncol), dim3(kBlk), 0, (hipStream_t)stream, n, ncol, ld, seed,
                     normal, out);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

}  // extern "C"
