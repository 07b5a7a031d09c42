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
#include <algorithm>
#include <type_traits>
#include <vector>

extern "C" {
int pc_gram_f64(const double* X, long long ldx, const double* Y, long long ldy, long long n, int kx, int ky,
                double* S, long long lds, int sym, void* work, size_t work_bytes, void* stream);
int pc_gram_f32(const float* X, long long ldx, const float* Y, long long ldy, long long n, int kx, int ky,
                double* S, long long lds, int sym, void* work, size_t work_bytes, void* stream);
}

namespace {

constexpr int kTriMaxM = 32;
constexpr int kTriClusterFlag = -7;
constexpr int kTriMaxK = 512;

struct TriWs {
  double* x;    // [k*k] eigenvectors of T, element j of vector i at [j * k + i]
  double* lam;  // [k] bisection eigenvalues
  double* par;  // [4] gl, gu, ||T||, pivmin
};

struct DenseCtx {
  rocblas_handle h;
  hipStream_t stream;
  int kmax;
  double* tmp;      // kmax*kmax
  double* E;        // kmax (syevd workspace)
  double* out;      // kmax + 4 : eigenvalues then 3 info codes (as doubles)
  int* info;        // 4
  double* host_out; // pinned, kmax + 4
  int method;       // 0: syevd, 1: Jacobi (syevj), 2: one-launch tridiagonalization + tri_* + Q Z
  double* tau;      // kmax (method 2: reflector scalars)
  double* P;        // method 2: padded copy of G, 64*ceil(kmax/64) x (kmax+1)
  double* resid;    // syevj: off-diagonal norm at exit
  int* sweeps;      // syevj: sweeps used
  int* sync;             // sytrd_lower_mwg: flags and abort word
  double* pub;           // sytrd_lower_mwg: published p and column, [4][320]
  bool mwg_ok;           // false after a multi-workgroup reduction gave up waiting
  int mwg_aborts;        // reductions repeated with the one-workgroup kernel
  int trd_spin;          // polls of one wait before sytrd_lower_mwg gives up
  double* S0;            // kmax*kmax: copy of the overlap matrix (shifted retry)
  double* G0;            // kmax*kmax: copy of the projected matrix (shifted retry)
  int shifted;           // Rayleigh-Ritz steps that needed a shifted Cholesky factor
  TriWs tw;              // tridiagonal eigensolver workspace (kmax <= kTriMaxK)
  double* tri_lam;       // kmax: its eigenvalues before they replace the diagonal
  bool tri_ok;           // false after a cluster exceeded kTriMaxM (stedc from then on)
  int tri_mode;          // R^-1: 0 = trtri + trmm, 1 = triangular solves, 2 = chol_blk_* + GEMMs (k <= 320)
  int blk_fallbacks;     // attempts repeated with potrf after chol_blk_factor found a pivot that is not positive
  // pc_dense_chol_async: the chol_blk_* kernels on a second stream (default
  // priority), ordered by events: ev_s (S ready on the main stream) before
  // them, ev_c after them; the next pc_dense_rr waits for ev_c
  hipStream_t side;
  hipEvent_t ev_s, ev_c;
  int* cinfo;            // their potrf-style info
  bool chol_pending;
  bool chol_made;        // the pending side work includes the factor (else only S)
  int chol_k;
  const double* chol_S;
  int chol_lds;
  double* cR;            // chol_blk_*: factor, inverse (32 nb x 32 nb each) and the diagonal blocks' inverses
  double* cRi;
  double* cD;
  int tri_fallbacks;     // steps repeated with stedc
  int orm_mode;          // Q Z: 1 = orm_tmat + orm_apply (default, k <= kOrmMaxK), 0 = rocSOLVER's ormtr
  double* ormT;          // orm_tmat's triangular factors, kOrmNB^2 per block of reflectors
};

// S += shift * max_i S_ii on the diagonal (one workgroup of 256 threads)
__global__ void add_diag_shift(double* S, int k, int lds, double shift) {
  __shared__ double red[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < k; i += 256) m = fmax(m, S[(long long)i * lds + i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double d = shift * red[0];
  for (int i = threadIdx.x; i < k; i += 256) S[(long long)i * lds + i] += d;
}

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

__global__ void copy_info(const int* src, int* dst) {
  if (threadIdx.x == 0) dst[0] = src[0];
}

__global__ void pack_out(const int* info, double* out, int k) {
  if (threadIdx.x < 4) out[k + threadIdx.x] = (double)info[threadIdx.x];
}

// ---- blocked Cholesky factor and its inverse (k <= 320), FP64 MFMA
//
// S = R^T R with R upper, in 32 x 32 blocks (S padded with the identity to
// 32 nb, nb = ceil(k / 32)), then Ri = R^-1; both in column-major 32 nb x 32 nb
// workspaces.  Replaces rocSOLVER potrf + trtri (about 30 launches at k = 275)
// with three: chol_blk_init (the copy of S), chol_blk_factor (one workgroup,
// right-looking by block columns: the diagonal block factored and inverted by
// one wavefront, then R_jc = R_jj^-T A_jc and A_ab -= R_ja^T R_jb shared out
// over the wavefronts, one 32 x 32 block each) and chol_blk_inverse
// (workgroup j: block column j of R^-1 by block back substitution).  The block
// products run on v_mfma_f64_16x16x4f64 (A[m][k] from lane m + 16 k, B[k][n]
// from lane n + 16 k, D[m][n] in lane n + 16 (m % 4), register m / 4).
constexpr int kCb = 32;            // block size
constexpr int kCbMaxNb = 10;       // k <= 320
constexpr int kCholMaxK = kCb * kCbMaxNb;
constexpr int kCbLW = kCb + 1;     // LDS column stride of a block
typedef double v4d __attribute__((ext_vector_type(4)));

struct CbLane {
  int lm, lk, wm, wn;
  __device__ CbLane() {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    lm = l & 15; lk = l >> 4; wm = w >> 1; wn = w & 1;
  }
  // the element of the wavefront's quarter held in result register r
  __device__ int row(int r) const { return 16 * wm + 4 * r + lk; }
  __device__ int col() const { return 16 * wn + lm; }
};

__device__ __forceinline__ long long cb_blk(int i, int j, int ld) { return (long long)(kCb * j) * ld + kCb * i; }

// acc += op(A) B over K = 32 (A, B column-major blocks; op(A) = A^T when TA)
template <bool TA>
__device__ __forceinline__ v4d cb_mma(v4d acc, const double* A, int lda, const double* B, int ldb, const CbLane& q) {
  const int m = 16 * q.wm + q.lm, n = 16 * q.wn + q.lm;
  double a[8], b[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int kk = 4 * s + q.lk;
    a[s] = TA ? A[m * lda + kk] : A[kk * lda + m];
    b[s] = B[n * ldb + kk];
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
  return acc;
}

// acc += sum_{u < nt} op(A[u]) B[u]; the operands of four terms are loaded
// before their MFMAs, so four block reads are in flight at once
template <bool TA>
__device__ __forceinline__ v4d cb_mma_sum(v4d acc, const double* const* A, int lda, const double* const* B, int ldb,
                                          int nt, const CbLane& q) {
  const int m = 16 * q.wm + q.lm, n = 16 * q.wn + q.lm;
  for (int u0 = 0; u0 < nt; u0 += 4) {
    double a[4][8], b[4][8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = u0 + u < nt;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int kk = 4 * s + q.lk;
        a[u][s] = in ? (TA ? A[u0 + u][m * lda + kk] : A[u0 + u][kk * lda + m]) : 0.0;
        b[u][s] = in ? B[u0 + u][n * ldb + kk] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][s], b[u][s], acc, 0, 0, 0);
  }
  return acc;
}

__device__ __forceinline__ double cb_readlane(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Cholesky factor and inverse of the symmetric 32 x 32 block in W (LDS), by
// wavefront 0: lane t holds column t % 32 in registers and reads row c of R
// from the lanes that hold it.  On success W holds R (upper, zero below) and
// Di = R^-1 (upper, zero below).  Returns the 1-based column of the first
// pivot that is not positive (or not finite), 0 on success.
__device__ __forceinline__ int cb_chol32(double* W, double* Di) {
  const int t = threadIdx.x & 31;
  double w[kCb];
#pragma unroll
  for (int r = 0; r < kCb; ++r) w[r] = W[t * kCbLW + r];
#pragma unroll
  for (int c = 0; c < kCb; ++c) {
    const double d = cb_readlane(w[c], c);
    if (!(d > 0.0) || !(d < __builtin_inf())) return c + 1;
    const double sq = __builtin_sqrt(d);
    const double is = 1.0 / sq;
    w[c] = t == c ? sq : w[c] * is;
    // rows a > c of column t: W(a, t) -= R(c, a) R(c, t) (below the diagonal: unused)
#pragma unroll
    for (int a = c + 1; a < kCb; ++a) w[a] -= cb_readlane(w[c], a) * w[c];
  }
#pragma unroll
  for (int r = 0; r < kCb; ++r) W[t * kCbLW + r] = r <= t ? w[r] : 0.0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // column t of R^-1 by back substitution, R's columns read from LDS
  double x[kCb];
#pragma unroll
  for (int r = 0; r < kCb; ++r) x[r] = r == t ? 1.0 : 0.0;
#pragma unroll
  for (int c = kCb - 1; c >= 0; --c) {
    x[c] = x[c] / W[c * kCbLW + c];
#pragma unroll
    for (int i = 0; i < c; ++i) x[i] -= W[c * kCbLW + i] * x[c];
  }
#pragma unroll
  for (int r = 0; r < kCb; ++r) Di[t * kCbLW + r] = x[r];
  return 0;
}

// acc += op(A) B over K = 32 for a whole 32 x 32 block by one wavefront: four
// 16 x 16 tiles, acc[tm][tn] (A, B column-major; op(A) = A^T when TA)
template <bool TA>
__device__ __forceinline__ void cb_mma_block(v4d (&acc)[2][2], const double* A, int lda, const double* B, int ldb,
                                             double sgn) {
  const int lane = threadIdx.x & 63, lm = lane & 15, lk = lane >> 4;
  double a[2][8], b[2][8];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int kk = 4 * s + lk, m = 16 * t + lm;
      a[t][s] = sgn * (TA ? A[m * lda + kk] : A[kk * lda + m]);
      b[t][s] = B[m * ldb + kk];
    }
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
      for (int tn = 0; tn < 2; ++tn)
        acc[tm][tn] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tm][s], b[tn][s], acc[tm][tn], 0, 0, 0);
}

// R (ld 32 nb) <- the upper blocks of S padded with the identity, one thread
// per element (a launch of its own: in the one-workgroup factor kernel the
// copy's load latencies add up to most of its time)
__global__ void chol_blk_init(const double* __restrict__ S, int lds, int k, double* __restrict__ R) {
  const int nb = (k + kCb - 1) / kCb, KP = kCb * nb;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= KP * KP) return;
  const int r = e % KP, c = e / KP;
  if (r / kCb > c / kCb) return;
  R[(long long)c * KP + r] = (r < k && c < k) ? S[(long long)c * lds + r] : (r == c ? 1.0 : 0.0);
}

// One workgroup of 256 threads, right-looking by block columns: R (ld 32 nb)
// holds the upper blocks of S padded with the identity (chol_blk_init); for each block
// column j the diagonal block is factored and inverted by wavefront 0
// (cb_chol32), then the wavefronts share out the panel blocks
// R_jc = R_jj^-T A_jc (c > j) and afterwards the trailing blocks
// A_ab -= R_ja^T R_jb (j < a <= b), one whole 32 x 32 block per wavefront at a
// time: three dependent phases per block column.  Dinv (nb blocks,
// column-major 32 x 32): the inverses of R's diagonal blocks; info[0]: as
// potrf's (0, or the 1-based column of the first pivot that is not positive).
__global__ __launch_bounds__(256) void chol_blk_factor(const double* __restrict__ S, int lds, int k, double* R,
                                                       double* Dinv, int* info) {
  __shared__ double W[kCb * kCbLW], Di[kCb * kCbLW];
  __shared__ int fail;
  const int nb = (k + kCb - 1) / kCb, KP = kCb * nb;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lm = lane & 15, lk = lane >> 4;
  if (threadIdx.x == 0) {
    fail = 0;
    info[0] = 0;
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    double* rjj = R + cb_blk(j, j, KP);
    for (int e = threadIdx.x; e < kCb * kCb; e += 256) {
      const int a = e & (kCb - 1), b = e / kCb;
      W[b * kCbLW + a] = rjj[(long long)b * KP + a];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int f = cb_chol32(W, Di);
      if (threadIdx.x == 0 && f != 0) {
        fail = 1;
        info[0] = kCb * j + f;
      }
    }
    __syncthreads();
    if (fail) return;
    {
      double* dj = Dinv + (long long)j * kCb * kCb;
      for (int e = threadIdx.x; e < kCb * kCb; e += 256) {
        const int a = e & (kCb - 1), b = e / kCb;
        rjj[(long long)b * KP + a] = W[b * kCbLW + a];
        dj[e] = Di[b * kCbLW + a];
      }
    }
    // panel: R_jc = Dinv_j^T A_jc, one block per wavefront at a time (Dinv_j from LDS)
    for (int c = j + 1 + wv; c < nb; c += 4) {
      v4d acc[2][2] = {{{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}}, {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}}};
      double* rjc = R + cb_blk(j, c, KP);
      cb_mma_block<true>(acc, Di, kCbLW, rjc, KP, 1.0);
      // every lane's loads of A_jc are done before any lane overwrites it
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            rjc[(long long)(16 * tn + lm) * KP + 16 * tm + 4 * r + lk] = acc[tm][tn][r];
    }
    __syncthreads();
    // trailing blocks: A_ab -= R_ja^T R_jb for j < a <= b, one block per wavefront at a time
    const int nt = nb - 1 - j;
    for (int q = wv; q < nt * (nt + 1) / 2; q += 4) {
      // q -> (a, b), a <= b, row by row of the upper triangle of the trailing blocks
      int a = 0, rem = q;
      while (rem >= nt - a) { rem -= nt - a; ++a; }
      const int b = a + rem;
      const int ga = j + 1 + a, gb = j + 1 + b;
      double* aab = R + cb_blk(ga, gb, KP);
      v4d acc[2][2];
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[tm][tn][r] = aab[(long long)(16 * tn + lm) * KP + 16 * tm + 4 * r + lk];
      cb_mma_block<true>(acc, R + cb_blk(j, ga, KP), KP, R + cb_blk(j, gb, KP), KP, -1.0);
#pragma unroll
      for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
          for (int r = 0; r < 4; ++r) aab[(long long)(16 * tn + lm) * KP + 16 * tm + 4 * r + lk] = acc[tm][tn][r];
    }
    __syncthreads();
  }
}

// Workgroup j (nb of them) writes block column j of Ri = R^-1 (ld 32 nb):
// X_jj = Dinv_j, X_ij = -Dinv_i sum_{l=i+1..j} R_il X_lj for i = j-1 .. 0,
// zero below the diagonal blocks.
__global__ __launch_bounds__(256) void chol_blk_inverse(const double* __restrict__ R, const double* __restrict__ Dinv,
                                                        int k, double* Ri) {
  __shared__ double Y[kCb * kCbLW];
  __shared__ const double* pa[kCbMaxNb];
  __shared__ const double* pb[kCbMaxNb];
  const int nb = (k + kCb - 1) / kCb, KP = kCb * nb;
  const int j = blockIdx.x;
  const CbLane q;
  for (int e = threadIdx.x; e < kCb * kCb; e += 256) {
    const int a = e & (kCb - 1), b = e / kCb;
    Ri[cb_blk(j, j, KP) + (long long)b * KP + a] = Dinv[(long long)j * kCb * kCb + e];
    for (int i = j + 1; i < nb; ++i) Ri[cb_blk(i, j, KP) + (long long)b * KP + a] = 0.0;
  }
  for (int i = j - 1; i >= 0; --i) {
    if (threadIdx.x < j - i) {
      const int l = i + 1 + threadIdx.x;
      pa[threadIdx.x] = R + cb_blk(i, l, KP);
      pb[threadIdx.x] = Ri + cb_blk(l, j, KP);
    }
    __syncthreads();   // also: the blocks X_lj written before are visible
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = cb_mma_sum<false>(acc, pa, KP, pb, KP, j - i, q);
#pragma unroll
    for (int r = 0; r < 4; ++r) Y[q.col() * kCbLW + q.row(r)] = acc[r];
    __syncthreads();
    v4d o = {0.0, 0.0, 0.0, 0.0};
    o = cb_mma<false>(o, Dinv + (long long)i * kCb * kCb, kCb, Y, kCbLW, q);
#pragma unroll
    for (int r = 0; r < 4; ++r) Ri[cb_blk(i, j, KP) + (long long)q.col() * KP + q.row(r)] = -o[r];
    __syncthreads();
  }
}

// ---- one-workgroup Householder tridiagonalization (LAPACK dsytrd layout, lower)
//
// For i = 0..n-2 the reflector H(i) = I - tau_i v_i v_i^T (v_i(i+1) = 1,
// v_i(i+2:) stored in P(i+2:, i)) annihilates P(i+2:, i) of the trailing
// matrix; d = diagonal, e = subdiagonal of the tridiagonal Q^T G Q with
// Q = H(0)..H(n-2) -- the output of LAPACK's dsytd2 with uplo = 'L', so
// rocSOLVER's ormtr applies Q.  The whole reduction is one launch (rocSOLVER's
// sytrd spends most of its time launching per-column kernels at this size):
//   pass i (all waves): over the trailing columns c >= i+1 (interleaved over
//     the waves, U columns loaded per batch) apply the pending rank-2 update
//     of step i-1 and accumulate p = P v_i; wave 0 keeps column i+1;
//   serial part (wave 0 alone): w_i = tau_i p - (tau_i/2)(tau_i p.v_i) v_i,
//     column i+1 brought up to date with (v_i, w_i), and the reflector of
//     step i+1 built from it with wave-level sums.
// P is a padded copy of G (ld = 64 * NCH rows, one extra all-zero column), so
// the pass has no per-element conditions: a lane owns rows lane + 64 q of
// every column, rows outside [i+1, n) are padding or stale upper-triangle
// entries that nothing reads, and columns past n map to the zero column.
// (Per-element conditions turn into branches whose joins make the compiler
// wait for every outstanding store.)  P stays L2-resident for one workgroup.
constexpr int kTrdMax = 512;                    // largest n
constexpr int kTrdThreads = 512;
constexpr int kTrdWaves = kTrdThreads / 64;

__device__ __forceinline__ double wave_allsum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;   // the same in every lane (butterfly)
}

template <int NCH>
__global__ __launch_bounds__(kTrdThreads) void sytrd_lower_1wg(const double* __restrict__ G, int ldg, int n,
                                                               double* __restrict__ P, double* __restrict__ d,
                                                               double* __restrict__ e, double* __restrict__ tau) {
  constexpr int LDP = 64 * NCH;      // padded leading dimension of P
  __shared__ double v[kTrdMax + 1], w[kTrdMax + 1], vn[kTrdMax + 1], colsh[kTrdMax];
  __shared__ double red[kTrdWaves][kTrdMax];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int r = tid; r <= kTrdMax; r += kTrdThreads) {
    v[r] = 0.0;
    w[r] = 0.0;
    vn[r] = 0.0;
  }
  // P <- G, zero padding rows and the extra column n
  for (int c = wv; c <= n; c += kTrdWaves)
    for (int r = lane; r < LDP; r += 64)
      P[(long long)c * LDP + r] = (c < n && r < n) ? G[(long long)c * ldg + r] : 0.0;
  __syncthreads();
  double tau_cur = 0.0;   // wave 0: tau of the step being passed
  double col[NCH];        // wave 0: up-to-date column j, rows lane + 64 q
  // Reflector of column j from col[] (wave 0): d_j, e_j, tau_j, vn and P(:, j).
  auto reflector = [&](int j) {
#pragma unroll
    for (int q = 0; q < NCH; ++q) colsh[lane + 64 * q] = col[q];
    __builtin_amdgcn_wave_barrier();
    double* __restrict__ pj = P + (long long)j * LDP;
    const double dj = colsh[j];
    if (j == n - 1) {
      if (lane == 0) {
        pj[j] = dj;
        d[j] = dj;
      }
      return;
    }
    const double alpha = colsh[j + 1];
    double sq = 0.0;
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      const int r = lane + 64 * q;
      sq += (r >= j + 2 && r < n) ? col[q] * col[q] : 0.0;
    }
    const double xnorm2 = wave_allsum(sq);
    double beta = alpha, t = 0.0, scal = 0.0;
    if (xnorm2 > 0.0) {
      const double nrm = sqrt(alpha * alpha + xnorm2);
      beta = alpha >= 0.0 ? -nrm : nrm;
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      const int r = lane + 64 * q;
      if (r >= j && r < n) {
        const double out = r == j ? dj : (r == j + 1 ? beta : col[q] * scal);
        pj[r] = out;
        vn[r] = r == j ? 0.0 : (r == j + 1 ? 1.0 : out);
      }
    }
    if (lane == 0) {
      d[j] = dj;
      e[j] = beta;
      tau[j] = t;
    }
    tau_cur = t;
  };
  if (wv == 0) {
#pragma unroll
    for (int q = 0; q < NCH; ++q) col[q] = P[lane + 64 * q];
    reflector(0);
  }
  __syncthreads();
  for (int i = 0; i + 1 < n; ++i) {
    const double pf = i > 0 ? 1.0 : 0.0;   // a rank-2 update of step i-1 is pending
    const int r0 = i + 1;
    const int qb = r0 >> 6;                // first 64-row chunk holding rows >= i+1
    // ---- pass i over the NQ chunks qb.. (one instantiation per NQ, no
    // per-element conditions inside)
    auto pass = [&](auto nq_c) {
      constexpr int NQ = decltype(nq_c)::value;
      constexpr int U = NQ == 1 ? 24 : 48 / NQ;   // columns per load batch
      const int rb = 64 * qb + lane;
      double pacc[NQ], vr[NQ], wr[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        pacc[q] = 0.0;
        vr[q] = pf * v[rb + 64 * q];
        wr[q] = pf * w[rb + 64 * q];
      }
      for (int c0 = r0 + wv; c0 < n; c0 += kTrdWaves * U) {
        double x[U][NQ];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int c = min(c0 + kTrdWaves * u, n);
          const double* __restrict__ pc = P + (long long)c * LDP + rb;
#pragma unroll
          for (int q = 0; q < NQ; ++q) x[u][q] = pc[64 * q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int c = min(c0 + kTrdWaves * u, n);
          double* __restrict__ pc = P + (long long)c * LDP + rb;
          const double vc = vn[c], pv = pf * v[c], pw = pf * w[c];
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const double y = x[u][q] - (vr[q] * pw + wr[q] * pv);
            pc[64 * q] = y;
            pacc[q] += y * vc;
            x[u][q] = y;
          }
        }
        if (wv == 0 && c0 == r0) {   // wave 0's first column is i+1: keep it
#pragma unroll
          for (int q = 0; q < NQ; ++q) colsh[rb + 64 * q] = x[0][q];
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q) red[wv][rb + 64 * q] = pacc[q];
    };
    switch (NCH - qb) {
#define PC_TRD_PASS(N)                                    \
  case N:                                                 \
    if constexpr (N <= NCH) pass(std::integral_constant<int, N>{}); \
    break;
      PC_TRD_PASS(1) PC_TRD_PASS(2) PC_TRD_PASS(3) PC_TRD_PASS(4)
      PC_TRD_PASS(5) PC_TRD_PASS(6) PC_TRD_PASS(7) PC_TRD_PASS(8)
#undef PC_TRD_PASS
      default: break;
    }
    __syncthreads();
    // ---- serial part (wave 0): w_i, column i+1, reflector i+1
    if (wv == 0) {
      double wq[NCH], vq[NCH], dp = 0.0;
#pragma unroll
      for (int q = 0; q < NCH; ++q) {
        const int rr = lane + 64 * q;
        const bool act = rr >= r0 && rr < n;
        double p = 0.0;
        if (act) {
#pragma unroll
          for (int ww = 0; ww < kTrdWaves; ++ww) p += red[ww][rr];
        }
        vq[q] = act ? vn[rr] : 0.0;
        wq[q] = act ? tau_cur * p : 0.0;
        col[q] = act ? colsh[rr] : 0.0;
        dp += wq[q] * vq[q];
      }
      const double alpha2 = -0.5 * tau_cur * wave_allsum(dp);
#pragma unroll
      for (int q = 0; q < NCH; ++q) {
        const int rr = lane + 64 * q;
        wq[q] += alpha2 * vq[q];
        if (rr >= r0 && rr < n) {
          w[rr] = wq[q];
          v[rr] = vq[q];
        }
      }
      __builtin_amdgcn_wave_barrier();
      // column i+1 (rows >= i+1): v_i(i+1) = 1
      const double w1 = w[r0];
#pragma unroll
      for (int q = 0; q < NCH; ++q) col[q] -= vq[q] * w1 + wq[q];
      reflector(r0);
    }
    __syncthreads();
  }
  if (n == 1 && tid == 0) d[0] = P[0];
}

// ---- multi-workgroup Householder tridiagonalization (same output as above)
//
// kTrdW workgroups of 256 threads; column c of the matrix lives in the LDS of
// workgroup c % kTrdW for the whole reduction, so the trailing updates read and
// write LDS instead of L2.  Step i:
//   wave 0 of every workgroup (redundantly): w_{i-1} from the published
//     p_{i-1}, then column i (published by its owner, updated through step
//     i-2) brought up to date with (v_{i-1}, w_{i-1}) and reflector i built
//     from it -- the same arithmetic in every workgroup, so no broadcast;
//   all waves: each owned column c >= i+1 gets the pending rank-2 update of
//     step i-1 and p_i(c) = column . v_i; p_i(c) and column i+1 are published;
//   a barrier over the workgroups: each writes its step number to its own
//     flag after a release fence, and polls the others' flags.
// Published values, flags and the abort word are agent-scope atomics, so the
// exchange does not depend on where the workgroups run.  A wait that exceeds
// kTrdSpin polls (settable per context, for tests) sets the abort word and
// every workgroup returns; the host then repeats the reduction with the
// one-workgroup kernel.  All kTrdW workgroups must be resident at once: each
// takes one CU's worth of LDS at most, far below the device's CU count, and the
// kernel is the only work on its stream.  Measured at k = 275: 1.27 ms with
// 32 workgroups, 1.36 ms with 16, 3.26 ms for the one-workgroup kernel.
#ifndef PC_TRD_W
#define PC_TRD_W 32
#endif
constexpr int kTrdW = PC_TRD_W;           // workgroups (at most 64: one poll lane each)
constexpr int kTrdMwgMax = 320;           // largest n (LDS: ceil(320 / kTrdW) columns of 320 rows)
constexpr int kTrdMwgMin = 128;           // smaller n use the one-workgroup kernel
constexpr int kTrdSpin = 1 << 21;         // polls before a wait is declared failed
constexpr int kTrdFlagStride = 32;        // ints between two workgroups' flags
constexpr int kTrdSyncInts = kTrdFlagStride * (kTrdW + 1);   // flags, then the abort word
constexpr int kTrdLoc = (kTrdMwgMax + kTrdW - 1) / kTrdW;

__device__ __forceinline__ void pub_store(double* p, double x) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(x),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pub_load(const double* p) {
  return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

template <int NCH>
__global__ __launch_bounds__(256) void sytrd_lower_mwg(const double* __restrict__ G, int ldg, int n,
                                                       double* __restrict__ P, double* __restrict__ d,
                                                       double* __restrict__ e, double* __restrict__ tau,
                                                       int* sync, double* pub, int* info_abort,
                                                       int spin_limit) {
  constexpr int LDP = 64 * NCH;
  __shared__ double loc[kTrdLoc * LDP];                 // owned columns, rows 0..LDP
  __shared__ double vp[LDP + 1], wp[LDP + 1], vc[LDP + 1], colsh[LDP];
  __shared__ double tau_sh;
  __shared__ int abort_sh;
  const int wg = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nloc = (n - wg + kTrdW - 1) / kTrdW;        // columns wg, wg + W, ... < n
  int* flags = sync;
  int* abort_word = sync + kTrdFlagStride * kTrdW;
  double* pbuf = pub;                                   // [2][LDP]: p_i
  double* cbuf = pub + 2 * LDP;                         // [2][LDP]: column i+1
  for (int t = tid; t < kTrdLoc * LDP; t += 256) {
    const int lc = t / LDP, r = t - lc * LDP;
    const int c = wg + kTrdW * lc;
    loc[t] = (lc < nloc && r < n) ? G[(long long)c * ldg + r] : 0.0;
  }
  for (int r = tid; r <= LDP; r += 256) {
    vp[r] = 0.0;
    wp[r] = 0.0;
    vc[r] = 0.0;
  }
  if (tid == 0) abort_sh = 0;
  __syncthreads();
  for (int i = 0; i + 1 < n; ++i) {
    // ---- wave 0: w_{i-1}, column i, reflector i
    if (wv == 0) {
      double x[NCH];
      if (i > 0) {
        const double* pb = pbuf + ((i - 1) & 1) * LDP;
        const double* cb = cbuf + ((i - 1) & 1) * LDP;   // column i, published at step i-1
        double pq[NCH], vq[NCH], dp = 0.0;
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
          const int r = lane + 64 * q;
          pq[q] = (r >= i && r < n) ? pub_load(pb + r) : 0.0;
          x[q] = r < n ? pub_load(cb + r) : 0.0;
        }
        const double t_prev = tau_sh;
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
          const int r = lane + 64 * q;
          vq[q] = vc[r];
          pq[q] *= t_prev;
          dp += pq[q] * vq[q];
        }
        const double alpha2 = -0.5 * t_prev * wave_allsum(dp);
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
          const int r = lane + 64 * q;
          const double wr = pq[q] + alpha2 * vq[q];
          wp[r] = wr;
          vp[r] = vq[q];
        }
        __builtin_amdgcn_wave_barrier();
        const double wi = wp[i], vi = vp[i];
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
          const int r = lane + 64 * q;
          x[q] -= vp[r] * wi + wp[r] * vi;
        }
      } else {
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
          const int r = lane + 64 * q;
          x[q] = r < n ? G[r] : 0.0;
        }
      }
#pragma unroll
      for (int q = 0; q < NCH; ++q) colsh[lane + 64 * q] = x[q];
      __builtin_amdgcn_wave_barrier();
      const double dj = colsh[i], alpha = colsh[i + 1];
      double sq = 0.0;
#pragma unroll
      for (int q = 0; q < NCH; ++q) {
        const int r = lane + 64 * q;
        sq += (r >= i + 2 && r < n) ? x[q] * x[q] : 0.0;
      }
      const double xnorm2 = wave_allsum(sq);
      double beta = alpha, t = 0.0, scal = 0.0;
      if (xnorm2 > 0.0) {
        const double nrm = sqrt(alpha * alpha + xnorm2);
        beta = alpha >= 0.0 ? -nrm : nrm;
        t = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      double* __restrict__ pj = P + (long long)i * LDP;
#pragma unroll
      for (int q = 0; q < NCH; ++q) {
        const int r = lane + 64 * q;
        const double out = r == i ? dj : (r == i + 1 ? beta : x[q] * scal);
        vc[r] = (r <= i || r >= n) ? 0.0 : (r == i + 1 ? 1.0 : out);
        if (wg == 0 && r >= i && r < n) pj[r] = out;
      }
      if (wg == 0 && lane == 0) {
        d[i] = dj;
        e[i] = beta;
        tau[i] = t;
      }
      if (lane == 0) tau_sh = t;
    }
    __syncthreads();
    // ---- all waves: pending update of step i-1 on the owned columns c >= i+1, p_i(c)
    double* pb = pbuf + (i & 1) * LDP;
    double* cb = cbuf + (i & 1) * LDP;
    for (int lc = wv; lc < nloc; lc += 4) {
      const int c = wg + kTrdW * lc;
      if (c < i + 1) continue;
      double* col = loc + lc * LDP;
      const double vpc = vp[c], wpc = wp[c];
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < NCH; ++q) {
        const int r = lane + 64 * q;
        const double y = col[r] - (vp[r] * wpc + wp[r] * vpc);
        col[r] = y;
        acc += y * vc[r];
        if (c == i + 1) pub_store(cb + r, y);
      }
      acc = wave_allsum(acc);
      if (lane == 0) pub_store(pb + c, acc);
    }
    // ---- barrier over the workgroups
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (wv == 0) {
      if (lane == 0) __hip_atomic_store(flags + kTrdFlagStride * wg, i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int spins = 0;
      bool ok = false;
      while (true) {
        const int f = lane < kTrdW ? __hip_atomic_load(flags + kTrdFlagStride * lane, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT) : (i + 1);
        const int ab = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        const bool all = __all(f >= i + 1);
        if (ab != 0) break;
        if (all) {
          ok = true;
          break;
        }
        if (++spins >= spin_limit) {
          if (lane == 0) __hip_atomic_store(abort_word, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      if (lane == 0) abort_sh = ok ? 0 : 1;
    }
    __syncthreads();
    if (abort_sh) {
      if (wg == 0 && tid == 0) *info_abort = 1;
      return;
    }
  }
  // last diagonal entry: column n-1 through step n-2 (wave 0 of workgroup 0)
  if (wg == 0 && wv == 0) {
    const int i = n - 1;
    const double* pb = pbuf + ((i - 1) & 1) * LDP;
    const double* cb = cbuf + ((i - 1) & 1) * LDP;
    const double t_prev = tau_sh;
    double pq[NCH], vq[NCH], dp = 0.0;
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      const int r = lane + 64 * q;
      pq[q] = (r >= i && r < n) ? t_prev * pub_load(pb + r) : 0.0;
      vq[q] = vc[r];
      dp += pq[q] * vq[q];
    }
    const double alpha2 = -0.5 * t_prev * wave_allsum(dp);
    // only row n-1 is needed: x(n-1) - 2 v(n-1) w(n-1), with v(n-1) = 1
    double dn = 0.0;
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      const int r = lane + 64 * q;
      if (r == i) {
        const double wr = pq[q] + alpha2 * vq[q];
        dn = pub_load(cb + r) - 2.0 * vq[q] * wr;
        P[(long long)i * LDP + i] = dn;
        d[i] = dn;
      }
    }
  }
}

// ---- eigenpairs of the symmetric tridiagonal matrix (d, e) -----------------
// Replaces rocSOLVER's stedc (five merge levels, about a hundred launches at
// k = 275) with three launches:
//   tri_bisect:  eigenvalue i by multisection on the Sturm count, one wavefront
//                per eigenvalue: each round its 64 lanes count the eigenvalues
//                below 64 equally spaced shifts of the current interval and the
//                interval shrinks 65-fold, until it is 2 eps |lambda| +
//                1e-2 eps ||T|| wide (about 10 rounds; two interleaved chains
//                of 128 shifts per round measured slower: the step is bound
//                by FP64 issue, not latency);
//   tri_invit:   inverse iteration, one thread per eigenvalue: LU with partial
//                pivoting of T - sigma I, three solves from a fixed start
//                vector, each normalised; equal shifts are separated by
//                10 eps max(|lambda|, 1e-3 ||T||) per repeat;
//   tri_cluster: eigenvalues closer than 1e-3 ||T|| form a cluster (one
//                workgroup each): its vectors are orthonormalised (modified
//                Gram-Schmidt, twice) and rotated to the eigenvectors of their
//                projected m x m matrix (Jacobi), so vectors of close or equal
//                eigenvalues come out orthogonal and exact within the cluster.
// The eigenvectors pass from tri_invit to tri_cluster through a workspace
// interleaved as [j * k + i] (element j of eigenvalue i), so the threads'
// accesses are contiguous; tri_invit's factors stay in its LDS.  A cluster of
// more than kTriMaxM eigenvalues is reported in info (kTriClusterFlag) and the
// caller repeats the step with rocSOLVER's stedc.
// 1/q to about one ulp: the hardware reciprocal and two Newton steps (a
// shorter dependent chain than IEEE division; q is never 0 where it is used)
__device__ __forceinline__ double tri_rcp(double q) {
  double r = __builtin_amdgcn_rcp(q);
  double t = fma(-q, r, 1.0);
  r = fma(r, t, r);
  t = fma(-q, r, 1.0);
  return fma(r, t, r);
}

__device__ __forceinline__ void tri_bounds(const double* d, const double* e, int k, double* red, double* out4) {
  // Gershgorin interval, ||T|| and the pivot threshold (256 threads)
  double gl = 1e300, gu = -1e300, em = 0.0;
  for (int j = threadIdx.x; j < k; j += 256) {
    const double a = j > 0 ? fabs(e[j - 1]) : 0.0, b = j + 1 < k ? fabs(e[j]) : 0.0;
    gl = fmin(gl, d[j] - a - b);
    gu = fmax(gu, d[j] + a + b);
    if (j + 1 < k) em = fmax(em, e[j] * e[j]);
  }
  red[threadIdx.x] = gl;
  red[256 + threadIdx.x] = gu;
  red[512 + threadIdx.x] = em;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + o]);
      red[256 + threadIdx.x] = fmax(red[256 + threadIdx.x], red[256 + threadIdx.x + o]);
      red[512 + threadIdx.x] = fmax(red[512 + threadIdx.x], red[512 + threadIdx.x + o]);
    }
    __syncthreads();
  }
  const double eps = 2.220446049250313e-16;
  const double tn = fmax(fmax(fabs(red[0]), fabs(red[256])), 1e-300);
  out4[0] = red[0] - 2.0 * eps * tn * k;
  out4[1] = red[256] + 2.0 * eps * tn * k;
  out4[2] = tn;
  out4[3] = 2.2250738585072014e-308 * fmax(1.0, red[512]);
}

__global__ __launch_bounds__(256) void tri_bisect(const double* __restrict__ d, const double* __restrict__ e, int k,
                                                  TriWs ws) {
  __shared__ double sd[kTriMaxK], se2[kTriMaxK], red[768];
  __shared__ double par[4];
  for (int j = threadIdx.x; j < k; j += 256) {
    sd[j] = d[j];
    se2[j] = j + 1 < k ? e[j] * e[j] : 0.0;
  }
  double b4[4];
  tri_bounds(d, e, k, red, b4);
  if (threadIdx.x == 0) {
    for (int t = 0; t < 4; ++t) par[t] = b4[t];
    if (blockIdx.x == 0)
      for (int t = 0; t < 4; ++t) ws.par[t] = b4[t];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= k) return;
  const double pivmin = par[3];
  double lo = par[0], hi = par[1];
  for (int rnd = 0; rnd < 16; ++rnd) {
    const double w = hi - lo;
    const double x = lo + w * ((double)(lane + 1) / 65.0);
    double q = sd[0] - x;
    int cnt = q < 0.0;
    for (int j = 1; j < k; ++j) {
      if (fabs(q) < pivmin) q = -pivmin;
      q = sd[j] - x - se2[j - 1] * tri_rcp(q);
      cnt += q < 0.0;
    }
    // counts grow with the shift: the first lane counting more than i
    const unsigned long long b = __ballot(cnt > i);
    const int f = b ? __ffsll((long long)b) - 1 : 64;
    const double nlo = f == 0 ? lo : lo + w * ((double)f / 65.0);
    const double nhi = f == 64 ? hi : lo + w * ((double)(f + 1) / 65.0);
    lo = nlo;
    hi = nhi;
    if (hi - lo <= 4.440892098500626e-16 * fmax(fabs(lo), fabs(hi)) + 2.220446049250313e-18 * par[2] + pivmin)
      break;
  }
  if (lane == 0) ws.lam[i] = 0.5 * (lo + hi);
}

// Inverse iteration, one thread per eigenvalue.  Each thread's LU factors of
// T - sigma I (the reciprocal of U's diagonal, U's first superdiagonal, the
// multipliers and the pivot choices; U's second superdiagonal is nonzero only
// after a row swap and then equals e[j+1]) and its iterate sit in dynamic LDS,
// interleaved over the workgroup's TI threads ([row][thread], 33 bytes per row
// and thread): TI = 16 where 16 threads' rows fit in kInvLdsBytes (k <= 290),
// else 8.  TI is a compile-time constant so every row's LDS address is a base
// register plus an immediate offset; each pass walks the rows in blocks of
// kInvTB whose LDS reads are issued together before the block's dependent
// chain, with a row-by-row tail, and the pivot choices are selects, not
// branches (the threads of a wavefront pivot differently).  The start vector
// is generated inside the first forward solve.  At k = 275 this is 155 us
// against 255 us for the same arithmetic with a run-time thread count and a
// bounds check per row (and 270 us with the factors in global memory): the
// passes are bound by instruction issue and FP64 latency of one wave per CU.
constexpr int kInvLdsBytes = 150 * 1024;
constexpr int kInvTB = 16;
inline int tri_invit_threads(int k) { return (size_t)33 * 16 * k <= (size_t)kInvLdsBytes ? 16 : 8; }
inline size_t tri_invit_lds(int k) { return (size_t)33 * k * tri_invit_threads(k); }

// fixed pseudo-random start value of element j of vector i, in [-0.5, 0.5)
// (no two eigenvalues start alike)
__device__ __forceinline__ double tri_start(int j, int i) {
  unsigned h = (unsigned)j * 0x9E3779B1u ^ (unsigned)(i + 1) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return (double)(h & 0xFFFFFFu) * (1.0 / 16777216.0) - 0.5;
}

template <int TI>
__global__ __launch_bounds__(TI) void tri_invit(const double* __restrict__ d, const double* __restrict__ e, int k,
                                                TriWs ws) {
  constexpr int TB = kInvTB;
  __shared__ double sdd[kTriMaxK], sde[kTriMaxK];
  extern __shared__ double dyn[];
  for (int j = threadIdx.x; j < k; j += TI) {
    sdd[j] = d[j];
    sde[j] = j + 1 < k ? e[j] : 0.0;   // sde[k - 1] = 0
  }
  __syncthreads();
  const int lt = threadIdx.x;
  const int i = blockIdx.x * TI + lt;
  if (i >= k) return;
  // row j of thread lt at [j * TI] of each array
  double* RP = dyn + lt;                                   // 1 / U(j,j), the pivot held away from 0
  double* U2 = dyn + (size_t)k * TI + lt;                  // U(j,j+1)
  double* LM = dyn + (size_t)2 * k * TI + lt;              // multiplier of step j
  double* XV = dyn + (size_t)3 * k * TI + lt;              // iterate
  unsigned char* PV = reinterpret_cast<unsigned char*>(dyn + (size_t)4 * k * TI) + lt;   // rows j, j+1 swapped
  const double eps = 2.220446049250313e-16;
  const double tn = ws.par[2];
  const double tiny = eps * tn;
  // separate equal shifts: the r-th repeat of a value moves up by r deltas
  const double delta = 10.0 * eps * fmax(fabs(ws.lam[i]), 1e-3 * tn);
  int rep = 0;
  for (int t = i; t > 0 && ws.lam[t] - ws.lam[t - 1] < delta; --t) ++rep;
  const double sg = ws.lam[i] + rep * delta;
  auto rcp_piv = [&](double u) { return tri_rcp(fabs(u) > tiny ? u : (u >= 0.0 ? tiny : -tiny)); };
  // LU with partial pivoting of T - sg I: cd, cu are the pending row's (j,j), (j,j+1)
  double cd = sdd[0] - sg, cu = k > 1 ? sde[0] : 0.0;
  auto lu_step = [&](int j, double sub, double dnext, double nu) {
    const double nd = dnext - sg;
    const bool pv = fabs(cd) < fabs(sub);
    const double den = pv ? sub : cd, num = pv ? cd : sub;
    const double m = den != 0.0 ? num * tri_rcp(den) : 0.0;
    RP[j * TI] = rcp_piv(den);
    U2[j * TI] = pv ? nd : cu;
    LM[j * TI] = m;
    PV[j * TI] = pv;
    const double a = pv ? cu : nd, b = pv ? nd : cu;
    cd = a - m * b;
    cu = pv ? -m * nu : nu;
  };
  int j0 = 0;
  for (; j0 + TB <= k - 1; j0 += TB) {
    double rs[TB], rn[TB], ru[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) {
      rs[t] = sde[j0 + t];
      rn[t] = sdd[j0 + t + 1];
      ru[t] = sde[j0 + t + 1];
    }
#pragma unroll
    for (int t = 0; t < TB; ++t) lu_step(j0 + t, rs[t], rn[t], ru[t]);
  }
  for (int j = j0; j < k - 1; ++j) lu_step(j, sde[j], sdd[j + 1], sde[j + 1]);
  // the last row has no superdiagonals
  RP[(k - 1) * TI] = rcp_piv(cd);
  U2[(k - 1) * TI] = 0.0;
  PV[(k - 1) * TI] = 0;
  double sc = 1.0;
  for (int it = 0; it < 3; ++it) {
    const bool first = it == 0;
    // L y = P x (forward, the row swaps as recorded)
    double yc = first ? tri_start(0, i) : XV[0];
    auto fw_step = [&](int j, double m, double yn, bool p) {
      XV[j * TI] = p ? yn : yc;
      const double a = p ? yc : yn, b = p ? yn : yc;
      yc = a - m * b;
    };
    j0 = 0;
    for (; j0 + TB <= k - 1; j0 += TB) {
      double rm[TB], rx[TB];
      bool rp[TB];
#pragma unroll
      for (int t = 0; t < TB; ++t) {
        rm[t] = LM[(j0 + t) * TI];
        rx[t] = first ? tri_start(j0 + t + 1, i) : XV[(j0 + t + 1) * TI];
        rp[t] = PV[(j0 + t) * TI];
      }
#pragma unroll
      for (int t = 0; t < TB; ++t) fw_step(j0 + t, rm[t], rx[t], rp[t]);
    }
    for (int j = j0; j < k - 1; ++j) fw_step(j, LM[j * TI], first ? tri_start(j + 1, i) : XV[(j + 1) * TI], PV[j * TI]);
    XV[(k - 1) * TI] = yc;
    // U x = y (backward); the previous pass's normalisation is folded into this one
    double y1 = 0.0, y2 = 0.0, nrm = 0.0;
    auto bw_step = [&](int j, double rp, double u2, double u3, double x) {
      const double y = (sc * x - u2 * y1 - u3 * y2) * rp;
      XV[j * TI] = y;
      y2 = y1;
      y1 = y;
      nrm += y * y;
    };
    // U(j,j+2) = sde[j + 1] after a swap (row k - 1 never swaps, and sde[k - 1] = 0)
    int jt = k - 1;
    for (; jt - TB + 1 >= 0; jt -= TB) {
      double r1[TB], r2[TB], r3[TB], rx[TB];
#pragma unroll
      for (int t = 0; t < TB; ++t) {
        const int j = jt - t;
        r1[t] = RP[j * TI];
        r2[t] = U2[j * TI];
        r3[t] = PV[j * TI] ? sde[min(j + 1, k - 1)] : 0.0;
        rx[t] = XV[j * TI];
      }
#pragma unroll
      for (int t = 0; t < TB; ++t) bw_step(jt - t, r1[t], r2[t], r3[t], rx[t]);
    }
    for (int j = jt; j >= 0; --j)
      bw_step(j, RP[j * TI], U2[j * TI], PV[j * TI] ? sde[min(j + 1, k - 1)] : 0.0, XV[j * TI]);
    sc = 1.0 / sqrt(nrm);
  }
  for (int j = 0; j < k; ++j) ws.x[(long long)j * k + i] = sc * XV[j * TI];
}

__device__ __forceinline__ double tri_block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Sum over the 256 threads of a workgroup: wavefront reductions, then the
// four partials in a fixed order (the same value in every thread).
__device__ __forceinline__ double tri_wg_sum(double v, double* red4) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (red4[0] + red4[1]) + (red4[2] + red4[3]);
  __syncthreads();
  return r;
}

// Z (k x k, ld k) and the eigenvalues (ascending) into lam_out; info_flag
// receives kTriClusterFlag when a cluster is larger than kTriMaxM.  The first
// workgroup of a cluster of m > 1 close eigenvalues copies the m inverse-
// iteration vectors into LDS (dynamic, kTriMaxM * k doubles), orthonormalises
// them by modified Gram-Schmidt twice (each thread owns the same rows
// throughout, so only the reductions synchronise), forms C = Xc^T T Xc,
// diagonalises it by cyclic Jacobi and writes Xc U in ascending order.
__global__ __launch_bounds__(256) void tri_cluster(const double* __restrict__ d, const double* __restrict__ e, int k,
                                                   TriWs ws, double* Z, double* lam_out, int* info_flag) {
  extern __shared__ double xs[];
  __shared__ double red4[4];
  __shared__ double C[kTriMaxM * kTriMaxM], U[kTriMaxM * kTriMaxM], wv[kTriMaxM];
  __shared__ int order[kTriMaxM];
  const int i = blockIdx.x;
  const double ortol = 1e-3 * ws.par[2];
  if (i > 0 && ws.lam[i] - ws.lam[i - 1] < ortol) return;   // not the first of its cluster
  int j = i;
  while (j + 1 < k && ws.lam[j + 1] - ws.lam[j] < ortol) ++j;
  const int m = j - i + 1;
  if (m == 1) {
    for (int r = threadIdx.x; r < k; r += 256) Z[(long long)i * k + r] = ws.x[(long long)r * k + i];
    if (threadIdx.x == 0) lam_out[i] = ws.lam[i];
    return;
  }
  if (m > kTriMaxM) {
    if (threadIdx.x == 0) *info_flag = kTriClusterFlag;
    return;
  }
  // the cluster's vectors into LDS, column c at xs + c k (global reads in runs of m)
  for (int t = threadIdx.x; t < m * k; t += 256) {
    const int r = t / m, c = t - r * m;
    xs[c * k + r] = ws.x[(long long)r * k + i + c];
  }
  __syncthreads();
  auto X = [&](int r, int c) -> double& { return xs[c * k + r]; };
  // modified Gram-Schmidt, twice
  for (int pass = 0; pass < 2; ++pass) {
    for (int a = 0; a < m; ++a) {
      for (int b = 0; b < a; ++b) {
        double s = 0.0;
        for (int r = threadIdx.x; r < k; r += 256) s += X(r, b) * X(r, a);
        s = tri_wg_sum(s, red4);
        for (int r = threadIdx.x; r < k; r += 256) X(r, a) -= s * X(r, b);
      }
      double s = 0.0;
      for (int r = threadIdx.x; r < k; r += 256) s += X(r, a) * X(r, a);
      s = tri_wg_sum(s, red4);
      const double inv = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
      for (int r = threadIdx.x; r < k; r += 256) X(r, a) *= inv;
    }
  }
  __syncthreads();   // T Xc reads the neighbours' rows
  // C = Xc^T T Xc
  for (int a = 0; a < m; ++a)
    for (int b = 0; b <= a; ++b) {
      double s = 0.0;
      for (int r = threadIdx.x; r < k; r += 256) {
        double tz = d[r] * X(r, b);
        if (r > 0) tz += e[r - 1] * X(r - 1, b);
        if (r + 1 < k) tz += e[r] * X(r + 1, b);
        s += X(r, a) * tz;
      }
      s = tri_wg_sum(s, red4);
      if (threadIdx.x == 0) {
        C[a * kTriMaxM + b] = s;
        C[b * kTriMaxM + a] = s;
      }
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    // cyclic Jacobi on the m x m matrix (nearly diagonal: few sweeps)
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) U[a * kTriMaxM + b] = a == b ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
      double off = 0.0, dia = 0.0;
      for (int a = 0; a < m; ++a) {
        dia += C[a * kTriMaxM + a] * C[a * kTriMaxM + a];
        for (int b = a + 1; b < m; ++b) off += C[a * kTriMaxM + b] * C[a * kTriMaxM + b];
      }
      if (off <= 1e-34 * dia || off == 0.0) break;
      for (int p = 0; p < m - 1; ++p)
        for (int q = p + 1; q < m; ++q) {
          const double apq = C[p * kTriMaxM + q];
          if (apq == 0.0) continue;
          const double app = C[p * kTriMaxM + p], aqq = C[q * kTriMaxM + q];
          const double th = (aqq - app) / (2.0 * apq);
          const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
          const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
          for (int r = 0; r < m; ++r) {   // C <- J^T C J, U <- U J
            const double crp = C[r * kTriMaxM + p], crq = C[r * kTriMaxM + q];
            C[r * kTriMaxM + p] = cs * crp - sn * crq;
            C[r * kTriMaxM + q] = sn * crp + cs * crq;
          }
          for (int r = 0; r < m; ++r) {
            const double cpr = C[p * kTriMaxM + r], cqr = C[q * kTriMaxM + r];
            C[p * kTriMaxM + r] = cs * cpr - sn * cqr;
            C[q * kTriMaxM + r] = sn * cpr + cs * cqr;
          }
          for (int r = 0; r < m; ++r) {
            const double urp = U[r * kTriMaxM + p], urq = U[r * kTriMaxM + q];
            U[r * kTriMaxM + p] = cs * urp - sn * urq;
            U[r * kTriMaxM + q] = sn * urp + cs * urq;
          }
        }
    }
    // ascending order of the cluster's eigenvalues
    for (int a = 0; a < m; ++a) { wv[a] = C[a * kTriMaxM + a]; order[a] = a; }
    for (int a = 1; a < m; ++a) {
      const int o = order[a];
      int b = a - 1;
      while (b >= 0 && wv[order[b]] > wv[o]) { order[b + 1] = order[b]; --b; }
      order[b + 1] = o;
    }
  }
  __syncthreads();
  if (threadIdx.x < m) lam_out[i + threadIdx.x] = wv[order[threadIdx.x]];
  for (int r = threadIdx.x; r < k; r += 256)
    for (int a = 0; a < m; ++a) {
      const int c = order[a];
      double s = 0.0;
      for (int b = 0; b < m; ++b) s += X(r, b) * U[b * kTriMaxM + c];
      Z[(long long)(i + a) * k + r] = s;
    }
}

#define PC_RB(call)                                                                          \
  do {                                                                                       \
    rocblas_status s_ = (call);                                                              \
    if (s_ != rocblas_status_success) {                                                      \
      pc::set_error(std::string(__func__) + ": " #call " -> rocblas status " + std::to_string((int)s_)); \
      return pc::PC_ERR_SOLVER;                                                              \
    }                                                                                        \
  } while (0)

// ---- back-transformation Q Z of the tridiagonal eigenvectors (k <= 320) ----
//
// Q = H(0) H(1) .. H(k-2) from the tridiagonalization above (reflector i:
// v_i(i+1) = 1, v_i(i+2:) in P(i+2:, i), scalar tau_i).  The reflectors go in
// blocks of kOrmNB, B_b = H(i0) .. H(i0 + kOrmNB - 1) = I - V_b T_b V_b^T with
// T_b upper triangular (the compact WY form of LAPACK's dlarft, forward and
// columnwise: T(t,t) = tau_t, T(0:t,t) = -tau_t T(0:t,0:t) V(:,0:t)^T v_t), and
// Q Z = B_0 (B_1 (.. (B_last Z))).  orm_tmat builds every T_b, one workgroup
// per block: the Gram matrix V_b^T V_b on MFMA, then row s of T_b by lane s
// (T(s,t) needs only T(s,q < t) and the Gram matrix, so the rows need no
// barriers).  orm_apply holds kOrmNC columns of Z in LDS per
// workgroup and applies the blocks last to first, W = V_b^T Z, W = T_b W,
// Z -= V_b W, each product on MFMA (v_mfma_f64_16x16x4f64, the lane layout of
// the Cholesky kernels above), with V_b staged through LDS.  Two launches in
// place of rocSOLVER's ormtr, whose blocked path launches about thirty small
// kernels (0.5 ms at k = 275).
constexpr int kOrmNB = 32;
constexpr int kOrmNC = 16;
constexpr int kOrmMaxK = 320;
constexpr int kOrmLdv = kOrmMaxK | 1;   // LDS stride of a staged V column

// Vs[t * kOrmLdv + rr] = v_{i0+t}(i0 + 1 + rr) for rr < k - i0 - 1 (zero
// columns past the last reflector).  A thread loads all kOrmNB entries of its
// row before storing any (the column index is clamped to a reflector that
// exists, the value selected afterwards), so its loads are in flight together.
__device__ __forceinline__ void orm_stage_v(const double* __restrict__ P, int ldp, int k, int i0, int nr,
                                            double* Vs) {
  const int L = k - i0 - 1;
  for (int rr = threadIdx.x; rr < L; rr += blockDim.x) {
    const double* src = P + (long long)i0 * ldp + i0 + 1 + rr;
    double v[kOrmNB];
#pragma unroll
    for (int t = 0; t < kOrmNB; ++t) {
      v[t] = *src;
      if (t + 1 < nr) src += ldp;
    }
#pragma unroll
    for (int t = 0; t < kOrmNB; ++t) Vs[t * kOrmLdv + rr] = t < nr && rr >= t ? (rr > t ? v[t] : 1.0) : 0.0;
  }
}

__device__ __forceinline__ v4d orm_mfma(double a, double b, v4d acc) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}

__global__ __launch_bounds__(256) void orm_tmat(const double* __restrict__ P, int ldp, int k,
                                                const double* __restrict__ tau, double* __restrict__ Tw) {
  __shared__ double Vs[kOrmNB * kOrmLdv];
  __shared__ double g[kOrmNB][kOrmNB + 1];
  __shared__ double Tl[kOrmNB][kOrmNB + 1];
  const int b = blockIdx.x, i0 = b * kOrmNB;
  const int nr = min(kOrmNB, k - 1 - i0);
  const int L = k - i0 - 1;
  orm_stage_v(P, ldp, k, i0, nr, Vs);
  __syncthreads();
  // g = V^T V: waves 0, 1, 2 take the 16 x 16 tiles (0,0), (0,16), (16,16)
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, lm = l & 15, lk = l >> 4;
  if (w < 3) {
    const int s0 = w == 2 ? 16 : 0, t0 = w == 0 ? 0 : 16;
    const double* va = Vs + (s0 + lm) * kOrmLdv;
    const double* vb = Vs + (t0 + lm) * kOrmLdv;
    v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    int rr = 0;
    for (; rr + 8 <= L; rr += 8) {
      acc0 = orm_mfma(va[rr + lk], vb[rr + lk], acc0);
      acc1 = orm_mfma(va[rr + 4 + lk], vb[rr + 4 + lk], acc1);
    }
    for (; rr < L; rr += 4) {
      const int r = rr + lk;
      acc0 = orm_mfma(r < L ? va[r] : 0.0, r < L ? vb[r] : 0.0, acc0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) g[s0 + 4 * r + lk][t0 + lm] = acc0[r] + acc1[r];
  }
  __syncthreads();
  // row s of T_b by lane s, kept in LDS (only lane s reads it back: no
  // barriers between the columns)
  if (threadIdx.x < kOrmNB) {
    const int s = threadIdx.x;
    for (int t = 0; t < kOrmNB; ++t) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int q = s;   // T(s,q) = 0 for q < s
      for (; q + 4 <= t; q += 4) {
        a0 += Tl[s][q] * g[q][t];
        a1 += Tl[s][q + 1] * g[q + 1][t];
        a2 += Tl[s][q + 2] * g[q + 2][t];
        a3 += Tl[s][q + 3] * g[q + 3][t];
      }
      for (; q < t; ++q) a0 += Tl[s][q] * g[q][t];
      const double tt = t < nr ? tau[i0 + t] : 0.0;
      Tl[s][t] = s < t ? -tt * ((a0 + a1) + (a2 + a3)) : (s == t ? tt : 0.0);
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kOrmNB * kOrmNB; idx += blockDim.x)
    Tw[(long long)b * kOrmNB * kOrmNB + idx] = Tl[idx / kOrmNB][idx % kOrmNB];
}

__global__ __launch_bounds__(256) void orm_apply(const double* __restrict__ P, int ldp, int k,
                                                 const double* __restrict__ Tw, double* __restrict__ Z, int ldz) {
  __shared__ double Vs[kOrmNB * kOrmLdv];
  __shared__ double Zs[kOrmMaxK * kOrmNC];            // Zs[r * kOrmNC + c]
  __shared__ double Ts[kOrmNB][kOrmNB + 1];
  __shared__ double Wp[2][kOrmNB][kOrmNC + 1];        // V^T Z over the two halves of the rows
  __shared__ double W2[kOrmNB][kOrmNC + 1];
  const int c0 = blockIdx.x * kOrmNC;
  const int nc = min(kOrmNC, k - c0);
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, lm = l & 15, lk = l >> 4;
  for (int c = 0; c < kOrmNC; ++c)
    for (int r = tid; r < k; r += blockDim.x) Zs[r * kOrmNC + c] = c < nc ? Z[(long long)(c0 + c) * ldz + r] : 0.0;
  const int nbk = (k - 1 + kOrmNB - 1) / kOrmNB;
  for (int b = nbk - 1; b >= 0; --b) {
    const int i0 = b * kOrmNB, nr = min(kOrmNB, k - 1 - i0), L = k - i0 - 1;
    __syncthreads();   // the previous block is done with Vs, Ts and Zs
    orm_stage_v(P, ldp, k, i0, nr, Vs);
    for (int idx = tid; idx < kOrmNB * kOrmNB; idx += blockDim.x)
      Ts[idx / kOrmNB][idx % kOrmNB] = Tw[(long long)b * kOrmNB * kOrmNB + idx];
    __syncthreads();
    {
      // W = V^T Z (rows i0 + 1 .. k - 1): wave w takes rows t0 = 16 (w & 1) of
      // W over half h = w >> 1 of the rows of Z
      const int t0 = 16 * (w & 1), h = w >> 1;
      const int steps = (L + 3) / 4, half = 4 * ((steps + 1) / 2);
      const int rb = h ? half : 0, re = h ? 4 * steps : half;
      const double* va = Vs + (t0 + lm) * kOrmLdv;
      const double* zb = Zs + (i0 + 1) * kOrmNC + lm;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      for (int rr = rb; rr < re; rr += 4) {
        const int r = rr + lk;
        acc = orm_mfma(r < L ? va[r] : 0.0, r < L ? zb[r * kOrmNC] : 0.0, acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) Wp[h][t0 + 4 * r + lk][lm] = acc[r];
    }
    __syncthreads();
    if (w < 2) {
      // W2 = T W (rows t0 = 16 w)
      const int t0 = 16 * w;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q0 = 0; q0 < kOrmNB; q0 += 4) {
        const int q = q0 + lk;
        acc = orm_mfma(Ts[t0 + lm][q], Wp[0][q][lm] + Wp[1][q][lm], acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) W2[t0 + 4 * r + lk][lm] = acc[r];
    }
    __syncthreads();
    // Z -= V W2, tiles of 16 rows over the waves
    for (int rr0 = 16 * w; rr0 < L; rr0 += 64) {
      const int ra = rr0 + lm;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q0 = 0; q0 < kOrmNB; q0 += 4) {
        const int q = q0 + lk;
        acc = orm_mfma(ra < L ? Vs[q * kOrmLdv + ra] : 0.0, W2[q][lm], acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = rr0 + 4 * r + lk;
        if (rr < L) Zs[(i0 + 1 + rr) * kOrmNC + lm] -= acc[r];
      }
    }
  }
  __syncthreads();
  for (int c = 0; c < nc; ++c)
    for (int r = tid; r < k; r += blockDim.x) Z[(long long)(c0 + c) * ldz + r] = Zs[r * kOrmNC + c];
}

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
  PC_HIP(hipMalloc(&c->tau, (size_t)kmax * sizeof(double)));
  c->P = nullptr;
  if (kmax <= kTrdMax)
    PC_HIP(hipMalloc(&c->P, (size_t)(64 * ((kmax + 63) / 64)) * (kmax + 1) * sizeof(double)));
  PC_HIP(hipHostMalloc(&c->host_out, (size_t)(kmax + 4) * sizeof(double), hipHostMallocDefault));
  PC_HIP(hipMalloc(&c->sync, kTrdSyncInts * sizeof(int)));
  PC_HIP(hipMalloc(&c->pub, 4 * kTrdMwgMax * sizeof(double)));
  c->tri_mode = 2;
  c->blk_fallbacks = 0;
  c->orm_mode = 1;
  PC_HIP(hipMalloc(&c->ormT, (size_t)((kOrmMaxK - 1 + kOrmNB - 1) / kOrmNB) * kOrmNB * kOrmNB * sizeof(double)));
  PC_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  PC_HIP(hipEventCreateWithFlags(&c->ev_s, hipEventDisableTiming));
  PC_HIP(hipEventCreateWithFlags(&c->ev_c, hipEventDisableTiming));
  PC_HIP(hipMalloc(&c->cinfo, sizeof(int)));
  c->chol_pending = false;
  c->chol_made = false;
  c->chol_k = 0;
  c->chol_S = nullptr;
  c->chol_lds = 0;
  {
    const int kb = std::min(kmax, kCholMaxK), nbb = (kb + kCb - 1) / kCb, kp = kCb * nbb;
    PC_HIP(hipMalloc(&c->cR, (size_t)kp * kp * sizeof(double)));
    PC_HIP(hipMalloc(&c->cRi, (size_t)kp * kp * sizeof(double)));
    PC_HIP(hipMalloc(&c->cD, (size_t)nbb * kCb * kCb * sizeof(double)));
  }
  c->tri_ok = kmax <= kTriMaxK;
  if (c->tri_ok) {
    // dynamic LDS limits of the tridiagonal kernels, for every k they take (a
    // process-wide attribute: set to the largest need, not this context's):
    // tri_cluster a cluster's vectors, kTriMaxM * k doubles; tri_invit its
    // factors and iterates, at most kInvLdsBytes
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&tri_cluster), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)((size_t)kTriMaxM * kTriMaxK * sizeof(double))) != hipSuccess)
      c->tri_ok = false;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&tri_invit<16>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kInvLdsBytes) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void*>(&tri_invit<8>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kInvLdsBytes) != hipSuccess)
      c->tri_ok = false;
  }
  c->tri_fallbacks = 0;
  c->tri_lam = nullptr;
  c->tw = TriWs{};
  if (c->tri_ok) {
    PC_HIP(hipMalloc(&c->tw.x, (size_t)kmax * kmax * sizeof(double)));
    PC_HIP(hipMalloc(&c->tw.lam, (2 * (size_t)kmax + 4) * sizeof(double)));
    c->tri_lam = c->tw.lam + kmax;
    c->tw.par = c->tri_lam + kmax;
  }
  PC_HIP(hipMalloc(&c->S0, (size_t)kmax * kmax * sizeof(double)));
  PC_HIP(hipMalloc(&c->G0, (size_t)kmax * kmax * sizeof(double)));
  c->shifted = 0;
  c->mwg_ok = true;
  c->mwg_aborts = 0;
  c->trd_spin = kTrdSpin;
  c->method = 2;
  *out_ctx = c;
  return pc::PC_OK;
}

int pc_dense_destroy(void* ctx) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  if (!c) return pc::PC_OK;
  (void)hipStreamSynchronize(c->side);
  (void)hipStreamDestroy(c->side);
  (void)hipEventDestroy(c->ev_s);
  (void)hipEventDestroy(c->ev_c);
  (void)hipFree(c->cinfo);
  (void)hipFree(c->sync);
  (void)hipFree(c->pub);
  rocblas_destroy_handle(c->h);
  (void)hipFree(c->tmp); (void)hipFree(c->E); (void)hipFree(c->out); (void)hipFree(c->info);
  (void)hipFree(c->resid); (void)hipFree(c->sweeps); (void)hipFree(c->tau); (void)hipFree(c->P);
  (void)hipFree(c->S0); (void)hipFree(c->G0);
  (void)hipFree(c->cR); (void)hipFree(c->cRi); (void)hipFree(c->cD); (void)hipFree(c->ormT);
  if (c->tw.x) { (void)hipFree(c->tw.x); (void)hipFree(c->tw.lam); }
  (void)hipHostFree(c->host_out);
  delete c;
  return pc::PC_OK;
}

// S (k x k, ld lds) and Gy (k x k, ld ldg) are overwritten.  On return T
// (k x k, ld ldt, device) holds R^-1 Q and lam_host[0..k) the ascending Ritz
// values.  Synchronises the stream once (to read info codes and lam).
//
// When the overlap matrix is not numerically positive definite (a block whose
// columns are nearly dependent, as the Ritz vectors of a Lanczos basis without
// reorthogonalisation can be), the step is repeated from saved copies of S
// and Gy with S + shift * max_i S_ii I, shift = 1e-13, 1e-11, 1e-9, 1e-7: the
// dependent directions then get large Ritz values and leave the wanted lower
// end of the spectrum, which the reference's Householder QR leaves in place.
int pc_dense_rr(void* ctx, int k, double* S, int lds, double* Gy, int ldg, double* T, int ldt, double* lam_dev,
                double* lam_host) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && S && Gy && T && lam_dev, "null pointer");
  PC_CHECK_ARG(k >= 1 && k <= c->kmax, "k out of range");
  PC_CHECK_ARG(lds >= k && ldg >= k && ldt >= k, "leading dimension too small");
  const double one = 1.0;
  // the blocked Cholesky kernels, until they report a pivot that is not
  // positive: that attempt is repeated with rocSOLVER's potrf (the same
  // shift), which then serves the rest of the call
  bool blk = c->tri_mode == 2 && k <= kCholMaxK;
  // a factor from pc_dense_chol_async: the main stream waits for it before it
  // touches the factor buffers (used or not); used for the first attempt when
  // it was made from this S
  bool pre = false;
  if (c->chol_pending) {
    PC_HIP(hipStreamWaitEvent(c->stream, c->ev_c, 0));
    pre = blk && c->chol_made && c->chol_k == k && c->chol_S == S && c->chol_lds == lds;
    c->chol_pending = false;
  }
  PC_HIP(hipMemcpy2DAsync(c->S0, k * sizeof(double), S, lds * sizeof(double), k * sizeof(double), k,
                          hipMemcpyDeviceToDevice, c->stream));
  PC_HIP(hipMemcpy2DAsync(c->G0, k * sizeof(double), Gy, ldg * sizeof(double), k * sizeof(double), k,
                          hipMemcpyDeviceToDevice, c->stream));
  constexpr int kShiftTries = 4;
  constexpr double kShifts[kShiftTries + 1] = {0.0, 1e-13, 1e-11, 1e-9, 1e-7};
  bool first = true;
  for (int shift_try = 0;; ++shift_try) {
  if (!first) {
    PC_HIP(hipMemcpy2DAsync(S, lds * sizeof(double), c->S0, k * sizeof(double), k * sizeof(double), k,
                            hipMemcpyDeviceToDevice, c->stream));
    PC_HIP(hipMemcpy2DAsync(Gy, ldg * sizeof(double), c->G0, k * sizeof(double), k * sizeof(double), k,
                            hipMemcpyDeviceToDevice, c->stream));
  }
  first = false;
  if (shift_try > 0) {
    hipLaunchKernelGGL(add_diag_shift, dim3(1), dim3(256), 0, c->stream, S, k, lds, kShifts[shift_try]);
    PC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(zero_info, dim3(1), dim3(64), 0, c->stream, c->info);
  PC_LAUNCH_CHECK();
  if (blk) {
    // R^-1 from chol_blk_*, then Gy <- R^-T (Gy R^-1) by two GEMMs
    const int nbk = (k + kCb - 1) / kCb;
    if (pre) {
      hipLaunchKernelGGL(copy_info, dim3(1), dim3(64), 0, c->stream, c->cinfo, c->info);
      PC_LAUNCH_CHECK();
      pre = false;   // a retry factors again on this stream
    } else {
      hipLaunchKernelGGL(chol_blk_init, dim3((kCb * nbk * kCb * nbk + 255) / 256), dim3(256), 0, c->stream, S, lds,
                         k, c->cR);
      PC_LAUNCH_CHECK();
      hipLaunchKernelGGL(chol_blk_factor, dim3(1), dim3(256), 0, c->stream, S, lds, k, c->cR, c->cD, c->info);
      PC_LAUNCH_CHECK();
      hipLaunchKernelGGL(chol_blk_inverse, dim3(nbk), dim3(256), 0, c->stream, c->cR, c->cD, k, c->cRi);
      PC_LAUNCH_CHECK();
    }
    const double zero = 0.0;
    PC_RB(rocblas_dgemm(c->h, rocblas_operation_none, rocblas_operation_none, k, k, k, &one, Gy, ldg, c->cRi,
                        kCb * nbk, &zero, c->tmp, k));
    PC_RB(rocblas_dgemm(c->h, rocblas_operation_transpose, rocblas_operation_none, k, k, k, &one, c->cRi, kCb * nbk,
                        c->tmp, k, &zero, Gy, ldg));
  } else {
  PC_RB(rocsolver_dpotrf(c->h, rocblas_fill_upper, k, S, lds, c->info));
  if (c->tri_mode == 1) {
    // Gy <- R^-T Gy R^-1 by two triangular solves in place
    PC_RB(rocblas_dtrsm(c->h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                        rocblas_diagonal_non_unit, k, k, &one, S, lds, Gy, ldg));
    PC_RB(rocblas_dtrsm(c->h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                        rocblas_diagonal_non_unit, k, k, &one, S, lds, Gy, ldg));
  } else {
    PC_RB(rocsolver_dtrtri(c->h, rocblas_fill_upper, rocblas_diagonal_non_unit, k, S, lds, c->info + 1));
    // tmp = Gy R^-1
    PC_RB(rocblas_dtrmm(c->h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                        rocblas_diagonal_non_unit, k, k, &one, S, lds, Gy, ldg, c->tmp, k));
    // Gy = R^-T tmp
    PC_RB(rocblas_dtrmm(c->h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                        rocblas_diagonal_non_unit, k, k, &one, S, lds, c->tmp, k, Gy, ldg));
  }
  }
  hipLaunchKernelGGL(symmetrize, dim3((k * k + 255) / 256), dim3(256), 0, c->stream, Gy, k, ldg);
  PC_LAUNCH_CHECK();
  for (int attempt = 0;; ++attempt) {
    const double* Q = Gy;   // eigenvectors of G
    int ldq = ldg;
    bool mwg = false;
    bool tri = false;
    if ((c->method == 2 || c->method == 3 || c->method == 4) && c->P && k <= kTrdMax) {
      // G = H T H^T, T = Z diag(lam) Z^T (tri_* kernels or stedc), Q = H Z (ormtr)
      const int nch = (k + 63) / 64;
      mwg = (c->method == 2 || c->method == 4) && c->mwg_ok && k >= kTrdMwgMin && k <= kTrdMwgMax;
      if (mwg) {
        PC_HIP(hipMemsetAsync(c->sync, 0, kTrdSyncInts * sizeof(int), c->stream));
        switch (nch) {
#define PC_TRDM(N)                                                                                            \
  case N:                                                                                                     \
    hipLaunchKernelGGL(sytrd_lower_mwg<N>, dim3(kTrdW), dim3(256), 0, c->stream, Gy, ldg, k, c->P, c->out,   \
                       c->E, c->tau, c->sync, c->pub, c->info + 3, c->trd_spin);                              \
    break;
          PC_TRDM(2) PC_TRDM(3) PC_TRDM(4) PC_TRDM(5)
#undef PC_TRDM
          default: break;
        }
      } else {
        switch (nch) {
#define PC_TRD(N)                                                                                                  \
  case N:                                                                                                          \
    hipLaunchKernelGGL(sytrd_lower_1wg<N>, dim3(1), dim3(kTrdThreads), 0, c->stream, Gy, ldg, k, c->P, c->out,  \
                       c->E, c->tau);                                                                              \
    break;
          PC_TRD(1) PC_TRD(2) PC_TRD(3) PC_TRD(4) PC_TRD(5) PC_TRD(6) PC_TRD(7) PC_TRD(8)
#undef PC_TRD
        }
      }
      PC_LAUNCH_CHECK();
      tri = c->method != 4 && c->tri_ok && k <= kTriMaxK;
      if (tri) {
        hipLaunchKernelGGL(tri_bisect, dim3((k + 3) / 4), dim3(256), 0, c->stream, c->out, c->E, k, c->tw);
        PC_LAUNCH_CHECK();
        {
          const int nt = tri_invit_threads(k);
          if (nt == 16)
            hipLaunchKernelGGL(tri_invit<16>, dim3((k + 15) / 16), dim3(16), tri_invit_lds(k), c->stream, c->out, c->E,
                               k, c->tw);
          else
            hipLaunchKernelGGL(tri_invit<8>, dim3((k + 7) / 8), dim3(8), tri_invit_lds(k), c->stream, c->out, c->E,
                               k, c->tw);
        }
        PC_LAUNCH_CHECK();
        hipLaunchKernelGGL(tri_cluster, dim3(k), dim3(256), (size_t)kTriMaxM * k * sizeof(double), c->stream, c->out,
                           c->E, k, c->tw, c->tmp, c->tri_lam, c->info + 2);
        PC_LAUNCH_CHECK();
        PC_HIP(hipMemcpyAsync(c->out, c->tri_lam, k * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      } else {
        PC_RB(rocsolver_dstedc(c->h, rocblas_evect_tridiagonal, k, c->out, c->E, c->tmp, k, c->info + 2));
      }
      if (c->orm_mode == 1 && k <= kOrmMaxK) {
        if (k > 1) {   // (k = 1: no reflectors, Q = I)
          hipLaunchKernelGGL(orm_tmat, dim3((k - 1 + kOrmNB - 1) / kOrmNB), dim3(256), 0, c->stream, c->P, 64 * nch,
                             k, c->tau, c->ormT);
          PC_LAUNCH_CHECK();
          hipLaunchKernelGGL(orm_apply, dim3((k + kOrmNC - 1) / kOrmNC), dim3(256), 0, c->stream, c->P, 64 * nch, k,
                             c->ormT, c->tmp, k);
          PC_LAUNCH_CHECK();
        }
      } else {
        PC_RB(rocsolver_dormtr(c->h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, k, k, c->P,
                               64 * nch, c->tau, c->tmp, k));
      }
      Q = c->tmp;
      ldq = k;
    } else if (c->method == 1) {
      PC_RB(rocsolver_dsyevj(c->h, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, k, Gy, ldg,
                             0.0, c->resid, 100, c->sweeps, c->out, c->info + 2));
    } else {
      PC_RB(rocsolver_dsyevd(c->h, rocblas_evect_original, rocblas_fill_upper, k, Gy, ldg, c->out, c->E,
                             c->info + 2));
    }
    // T = R^-1 Q
    if (blk) {
      const double zero = 0.0;
      const int kp = kCb * ((k + kCb - 1) / kCb);
      PC_RB(rocblas_dgemm(c->h, rocblas_operation_none, rocblas_operation_none, k, k, k, &one, c->cRi, kp, Q, ldq,
                          &zero, T, ldt));
    } else if (c->tri_mode == 1) {
      PC_HIP(hipMemcpy2DAsync(T, ldt * sizeof(double), Q, ldq * sizeof(double), k * sizeof(double), k,
                              hipMemcpyDeviceToDevice, c->stream));
      PC_RB(rocblas_dtrsm(c->h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                          rocblas_diagonal_non_unit, k, k, &one, S, lds, T, ldt));
    } else {
      PC_RB(rocblas_dtrmm(c->h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                          rocblas_diagonal_non_unit, k, k, &one, S, lds, Q, ldq, T, ldt));
    }
    PC_HIP(hipMemcpyAsync(lam_dev, c->out, k * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(pack_out, dim3(1), dim3(64), 0, c->stream, c->info, c->out, k);
    PC_LAUNCH_CHECK();
    PC_HIP(hipMemcpyAsync(c->host_out, c->out, (k + 4) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PC_HIP(hipStreamSynchronize(c->stream));
    // (a multi-workgroup reduction that gave up leaves a meaningless
    // tridiagonal matrix: that is repeated first, whatever came after it)
    if (tri && !(mwg && c->host_out[k + 3] != 0.0) && (int)c->host_out[k + 2] == kTriClusterFlag &&
        attempt < 2) {
      // a cluster of close eigenvalues larger than the kernels take: repeat
      // the step with rocSOLVER's stedc (G is unchanged) and keep using it
      c->tri_ok = false;
      c->tri_fallbacks += 1;
      PC_HIP(hipMemsetAsync(c->info + 2, 0, sizeof(int), c->stream));
      continue;
    }
    if (mwg && c->host_out[k + 3] != 0.0 && attempt < 2) {
      // a workgroup of the multi-workgroup reduction gave up waiting: repeat
      // with the one-workgroup kernel (G is unchanged) and keep using it
      c->mwg_ok = false;
      c->mwg_aborts += 1;
      PC_HIP(hipMemsetAsync(c->info + 2, 0, 2 * sizeof(int), c->stream));
      continue;
    }
    break;
  }
  const int i0 = (int)c->host_out[k];
  if (i0 != 0 && blk) {
    blk = false;
    c->blk_fallbacks += 1;
    --shift_try;
    continue;
  }
  if (i0 != 0 && shift_try < kShiftTries) continue;
  if (shift_try > 0) c->shifted += 1;
  break;
  }
  const int i0 = (int)c->host_out[k], i1 = (int)c->host_out[k + 1], i2 = (int)c->host_out[k + 2];
  if (i0 != 0) {
    pc::set_error("pc_dense_rr: overlap matrix not positive definite (potrf info " + std::to_string(i0) +
                  ") even with a shift of 1e-7 of its largest diagonal entry");
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

// The blocked Cholesky factor and inverse of S (k x k, ld lds) on the
// context's second stream, after the work already queued on the main stream
// (which produces S): the next pc_dense_rr with this S uses it, so the factor
// overlaps what the caller queues in between (H Y and its Gram matrix).  S
// must stay unchanged until that call.  Without the blocked kernels (another
// triangular mode, k > 320) nothing is queued.
static int side_begin(DenseCtx* c) {
  if (c->chol_pending) PC_HIP(hipStreamWaitEvent(c->stream, c->ev_c, 0));   // one batch of side work at a time
  PC_HIP(hipEventRecord(c->ev_s, c->stream));
  PC_HIP(hipStreamWaitEvent(c->side, c->ev_s, 0));
  return pc::PC_OK;
}

static int side_end(DenseCtx* c, int k, const double* S, int lds) {
  c->chol_made = false;
  if (c->tri_mode == 2 && k <= kCholMaxK) {
    const int nbk = (k + kCb - 1) / kCb;
    hipLaunchKernelGGL(chol_blk_init, dim3((kCb * nbk * kCb * nbk + 255) / 256), dim3(256), 0, c->side, S, lds, k,
                       c->cR);
    PC_LAUNCH_CHECK();
    hipLaunchKernelGGL(chol_blk_factor, dim3(1), dim3(256), 0, c->side, S, lds, k, c->cR, c->cD, c->cinfo);
    PC_LAUNCH_CHECK();
    hipLaunchKernelGGL(chol_blk_inverse, dim3(nbk), dim3(256), 0, c->side, c->cR, c->cD, k, c->cRi);
    PC_LAUNCH_CHECK();
    c->chol_made = true;
  }
  PC_HIP(hipEventRecord(c->ev_c, c->side));
  c->chol_pending = true;
  c->chol_k = k;
  c->chol_S = S;
  c->chol_lds = lds;
  return pc::PC_OK;
}

int pc_dense_chol_async(void* ctx, int k, const double* S, int lds) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && S, "null pointer");
  PC_CHECK_ARG(k >= 1 && k <= c->kmax && lds >= k, "bad size");
  if (c->tri_mode != 2 || k > kCholMaxK) return pc::PC_OK;
  if (side_begin(c) != pc::PC_OK) return pc::PC_ERR_HIP;
  return side_end(c, k, S, lds);
}

// S = Y^T Y (symmetric Gram matrix of the k columns of Y, FP32 when fp32 = 1,
// else FP64; n points, ld ldy) and then its blocked Cholesky factor and
// inverse, both on the context's second stream after the work queued on the
// main stream (which produces Y): the main stream can form H Y and its Gram
// matrix meanwhile.  work (work_bytes from pc_gram_workspace) must not be
// the workspace of the main stream's Gram matrix.  Y, S and work must stay
// unchanged until the next pc_dense_rr, which waits for this work.
int pc_dense_gram_chol_async(void* ctx, int fp32, const void* Y, long long ldy, long long n, int k, double* S,
                             int lds, void* work, size_t work_bytes) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && Y && S && work, "null pointer");
  PC_CHECK_ARG(k >= 1 && k <= c->kmax && lds >= k, "bad size");
  if (side_begin(c) != pc::PC_OK) return pc::PC_ERR_HIP;
  const int st = fp32 ? pc_gram_f32(static_cast<const float*>(Y), ldy, static_cast<const float*>(Y), ldy, n, k, k, S,
                                    lds, 1, work, work_bytes, c->side)
                      : pc_gram_f64(static_cast<const double*>(Y), ldy, static_cast<const double*>(Y), ldy, n, k, k,
                                    S, lds, 1, work, work_bytes, c->side);
  if (st != pc::PC_OK) return st;
  return side_end(c, k, S, lds);
}

// Eigensolver of the projected k x k problem: 0 = divide and conquer
// (rocSOLVER syevd), 1 = cyclic Jacobi (rocSOLVER syevj, tolerance = machine
// precision, at most 100 sweeps), 2 = Householder tridiagonalization in one
// launch (32 workgroups for 128 <= k <= 320, one workgroup otherwise; k > 512
// uses syevd) + the tri_* tridiagonal eigensolver (rocSOLVER stedc if a
// cluster of close eigenvalues exceeds kTriMaxM) + ormtr, 3 = as 2 with one
// workgroup always, 4 = as 2 with stedc always.
int pc_dense_set_method(void* ctx, int method) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c, "null context");
  PC_CHECK_ARG(method >= 0 && method <= 4,
               "method must be 0 (syevd), 1 (syevj), 2 (tridiagonalization + tridiagonal eigensolver), "
               "3 (2, one workgroup) or 4 (tridiagonalization + stedc)");
  c->method = method;
  return pc::PC_OK;
}

// Polls of one inter-workgroup wait before the multi-workgroup
// tridiagonalization gives up (spins >= 1; tests force the fallback with a
// small value).  out_aborts (optional) receives the number of reductions this
// context has repeated with the one-workgroup kernel; spins = 0 only reads it.
int pc_dense_trd_control(void* ctx, int spins, int reset, int* out_aborts) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c, "null context");
  PC_CHECK_ARG(spins >= 0, "spins < 0");
  if (out_aborts) *out_aborts = c->mwg_aborts;
  if (spins > 0) c->trd_spin = spins;
  if (reset) {
    c->mwg_ok = true;
    c->mwg_aborts = 0;
  }
  return pc::PC_OK;
}

// How the Cholesky factor and R^-1 are formed in pc_dense_rr: 0 = rocSOLVER
// potrf + trtri, then three trmm; 1 = potrf, then three trsm in place; 2 =
// chol_blk_factor + chol_blk_inverse and three GEMMs (default; k > 320 uses 0).
int pc_dense_set_triangular(void* ctx, int mode) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c, "null context");
  PC_CHECK_ARG(mode >= 0 && mode <= 2, "mode must be 0 (trtri + trmm), 1 (trsm) or 2 (blocked Cholesky + GEMM)");
  c->tri_mode = mode;
  return pc::PC_OK;
}

int pc_dense_set_backtransform(void* ctx, int mode) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c, "null context");
  PC_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (rocSOLVER ormtr) or 1 (orm_tmat + orm_apply)");
  c->orm_mode = mode;
  return pc::PC_OK;
}

// Steps of this context whose tridiagonal eigenproblem went to stedc because
// a cluster was too large for the tri_* kernels.
int pc_dense_tri_fallbacks(void* ctx, int* out) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && out, "null pointer");
  *out = c->tri_fallbacks;
  return pc::PC_OK;
}

// Number of Rayleigh-Ritz steps of this context that needed the shifted
// Cholesky factor of the overlap matrix.
int pc_dense_shift_count(void* ctx, int* out) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && out, "null pointer");
  *out = c->shifted;
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
