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
#include <type_traits>
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
  int method;       // 0: syevd, 1: Jacobi (syevj), 2: one-launch tridiagonalization + stedc + ormtr
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

__global__ void pack_out(const int* info, double* out, int k) {
  if (threadIdx.x < 4) out[k + threadIdx.x] = (double)info[threadIdx.x];
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
  PC_HIP(hipMalloc(&c->tau, (size_t)kmax * sizeof(double)));
  c->P = nullptr;
  if (kmax <= kTrdMax)
    PC_HIP(hipMalloc(&c->P, (size_t)(64 * ((kmax + 63) / 64)) * (kmax + 1) * sizeof(double)));
  PC_HIP(hipHostMalloc(&c->host_out, (size_t)(kmax + 4) * sizeof(double), hipHostMallocDefault));
  PC_HIP(hipMalloc(&c->sync, kTrdSyncInts * sizeof(int)));
  PC_HIP(hipMalloc(&c->pub, 4 * kTrdMwgMax * sizeof(double)));
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
  (void)hipFree(c->sync);
  (void)hipFree(c->pub);
  rocblas_destroy_handle(c->h);
  (void)hipFree(c->tmp); (void)hipFree(c->E); (void)hipFree(c->out); (void)hipFree(c->info);
  (void)hipFree(c->resid); (void)hipFree(c->sweeps); (void)hipFree(c->tau); (void)hipFree(c->P);
  (void)hipFree(c->S0); (void)hipFree(c->G0);
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
// and Gy with S + shift * max_i S_ii I, shift = 1e-13, 1e-11, 1e-9: the
// dependent directions then get large Ritz values and leave the wanted lower
// end of the spectrum, which the reference's Householder QR leaves in place.
int pc_dense_rr(void* ctx, int k, double* S, int lds, double* Gy, int ldg, double* T, int ldt, double* lam_dev,
                double* lam_host) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c && S && Gy && T && lam_dev, "null pointer");
  PC_CHECK_ARG(k >= 1 && k <= c->kmax, "k out of range");
  PC_CHECK_ARG(lds >= k && ldg >= k && ldt >= k, "leading dimension too small");
  const double one = 1.0;
  PC_HIP(hipMemcpy2DAsync(c->S0, k * sizeof(double), S, lds * sizeof(double), k * sizeof(double), k,
                          hipMemcpyDeviceToDevice, c->stream));
  PC_HIP(hipMemcpy2DAsync(c->G0, k * sizeof(double), Gy, ldg * sizeof(double), k * sizeof(double), k,
                          hipMemcpyDeviceToDevice, c->stream));
  constexpr int kShiftTries = 3;
  for (int shift_try = 0;; ++shift_try) {
  if (shift_try > 0) {
    PC_HIP(hipMemcpy2DAsync(S, lds * sizeof(double), c->S0, k * sizeof(double), k * sizeof(double), k,
                            hipMemcpyDeviceToDevice, c->stream));
    PC_HIP(hipMemcpy2DAsync(Gy, ldg * sizeof(double), c->G0, k * sizeof(double), k * sizeof(double), k,
                            hipMemcpyDeviceToDevice, c->stream));
    const double shift = shift_try == 1 ? 1e-13 : (shift_try == 2 ? 1e-11 : 1e-9);
    hipLaunchKernelGGL(add_diag_shift, dim3(1), dim3(256), 0, c->stream, S, k, lds, shift);
    PC_LAUNCH_CHECK();
  }
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
  for (int attempt = 0;; ++attempt) {
    const double* Q = Gy;   // eigenvectors of G
    int ldq = ldg;
    bool mwg = false;
    if ((c->method == 2 || c->method == 3) && c->P && k <= kTrdMax) {
      // G = H T H^T, T = Z diag(lam) Z^T (stedc), Q = H Z (ormtr)
      const int nch = (k + 63) / 64;
      mwg = c->method == 2 && c->mwg_ok && k >= kTrdMwgMin && k <= kTrdMwgMax;
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
      PC_RB(rocsolver_dstedc(c->h, rocblas_evect_tridiagonal, k, c->out, c->E, c->tmp, k, c->info + 2));
      PC_RB(rocsolver_dormtr(c->h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, k, k, c->P,
                             64 * nch, c->tau, c->tmp, k));
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
    PC_RB(rocblas_dtrmm(c->h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                        rocblas_diagonal_non_unit, k, k, &one, S, lds, Q, ldq, T, ldt));
    PC_HIP(hipMemcpyAsync(lam_dev, c->out, k * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(pack_out, dim3(1), dim3(64), 0, c->stream, c->info, c->out, k);
    PC_LAUNCH_CHECK();
    PC_HIP(hipMemcpyAsync(c->host_out, c->out, (k + 4) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PC_HIP(hipStreamSynchronize(c->stream));
    if (mwg && c->host_out[k + 3] != 0.0 && attempt == 0) {
      // a workgroup of the multi-workgroup reduction gave up waiting: repeat
      // with the one-workgroup kernel (G is unchanged) and keep using it
      c->mwg_ok = false;
      c->mwg_aborts += 1;
      PC_HIP(hipMemsetAsync(c->info + 3, 0, sizeof(int), c->stream));
      continue;
    }
    break;
  }
  const int i0 = (int)c->host_out[k];
  if (i0 != 0 && shift_try < kShiftTries) continue;
  if (shift_try > 0) c->shifted += 1;
  break;
  }
  const int i0 = (int)c->host_out[k], i1 = (int)c->host_out[k + 1], i2 = (int)c->host_out[k + 2];
  if (i0 != 0) {
    pc::set_error("pc_dense_rr: overlap matrix not positive definite (potrf info " + std::to_string(i0) +
                  ") even with a shift of 1e-9 of its largest diagonal entry");
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
// precision, at most 100 sweeps), 2 = Householder tridiagonalization in one
// launch (32 workgroups for 128 <= k <= 320, one workgroup otherwise; k > 512
// uses syevd) + rocSOLVER stedc + ormtr, 3 = as 2 with one workgroup always.
int pc_dense_set_method(void* ctx, int method) {
  DenseCtx* c = static_cast<DenseCtx*>(ctx);
  PC_CHECK_ARG(c, "null context");
  PC_CHECK_ARG(method >= 0 && method <= 3,
               "method must be 0 (syevd), 1 (syevj), 2 (tridiagonalization + stedc) or 3 (2, one workgroup)");
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
