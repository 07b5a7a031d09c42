// Multigrid-preconditioned CG for the Hartree solve (FP64).
//
// Same system and stop rule as pc_pcg_solve (A = -lap/h^2 with the domain's
// finite-difference stencil, Dirichlet zero outside the domain; stop when
// r.r <= tol^2 r0.r0 or after maxit iterations), preconditioned by one
// symmetric multigrid V-cycle per iteration -- the GPU counterpart of the
// reference's optional preconditioned CG (cg_prec, src/Eigensolvers/pcg.py:7-85,
// whose ILU preconditioner does not map to a GPU).  The preconditioner changes
// the iteration count, not the tolerance the solution meets.
//
// V-cycle, cell-centred: level 0 is the domain with its own stencil; level
// l >= 1 is a full cube of ceil(n_{l-1}/2) cells per axis, a cell active when
// any of its 8 children is, with the 7-point Laplacian at spacing 2^l h.
// Restriction = average of the active children, prolongation = injection
// (R = P^T / 8), one damped-Jacobi step before and after the coarse
// correction on every level, and a fixed-degree Chebyshev iteration on the
// coarsest cube (a polynomial in A, so the cycle is a fixed symmetric linear
// operator and CG theory applies).  From a zero start the pre-smoothed value
// is omega b / D, so the pre-smooth, residual and restriction of a cube level
// are one kernel, and the prolongation with the post-smooth is another.
//
// Launches per iteration:
//   fine level   mg_res0 (z-march: t0 = r - c0 A r), mg_restrict0,
//                mg_up0 (z-march: z = post-smooth(c0 r + P e_1), partials of r.z),
//                mg_apply (z-march: p = z + beta p, A p, partials of p.Ap), mg_update;
//   cube levels  mg_down / mg_up per level above the tail;
//   tail         mg_tail: the smallest levels and the coarsest Chebyshev in
//                one workgroup with every level in LDS.
// Arrays of a cube level hold zero on inactive cells, so a neighbour read
// needs no mask.  Every kernel returns at once when the iteration has
// converged, so the host reads the state back only every check_every
// iterations.
#include "pc_common.h"
#include "pc_zmarch.h"
#include <algorithm>
#include <cmath>
#include <vector>

long long pc_internal_grid_npts(void* grid);
int pc_internal_grid_view(void* grid, int* nx, int* ny, int* nz, const int** lo, const int** hi, const int** off);

namespace {

using pcz::ZMap;

constexpr int kB = 256;
constexpr int kNB = 512;          // workgroups of the point-wise launches (= r.r partials)
constexpr int kMaxLev = 8;
constexpr int kCheb = 8;          // Chebyshev steps on the coarsest cube
constexpr int kCoarsest = 512;    // cells of the coarsest cube, at most
constexpr int kTailCells = 512;   // cells of the levels held in LDS by mg_tail, at most (measured: one level beats more)
constexpr int kTailThreads = 1024;
// Damped Jacobi: the cycle stays SPD while omega < 2 / lambda_max(D^-1 A)
// (lambda_max(D^-1 A) of the radius-R stencil: 2.28 at R = 4, 2.37 at R = 6,
// so omega < 0.843 for every supported radius; 7-point: < 2, omega < 1).
// A V-cycle emulation on the Si87H76 sphere (smooth charge, tolerance 1e-5):
// 34 CG iterations at (0.6, 0.8), 29 at (0.8, 0.8), 28 at (0.8, 6/7).
constexpr double kOmega0 = 0.8;         // on the domain's own stencil
constexpr double kOmegaC = 6.0 / 7.0;   // on the 7-point cubes
enum { S_RR0 = 0, S_RZ0, S_RZ1, S_TOL1, S_DONE, S_ITS, S_RES, S_COUNT = 8 };

struct Fine {
  const int2* meta;
  const int* prow;
  long long n;
  int nx, ny, nz;
  double w[8];        // weights of A (w[0] = centre)
  double c0;          // omega0 / w[0]
  long long chunk;
  int nb, nb8;
};

struct Cube {
  int nx, ny, nz;
  long long cells;
  double inv_h2;      // 1 / (2^l h)^2
  const unsigned char* mask;
  double* b;
  double* e;
};

__device__ __forceinline__ double block_sum_bcast(double v, double* red) {
  const double t = pc_block_sum(v, red);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = t;
  __syncthreads();
  return bc;
}

// Sum of part[0..nb) in a fixed order; the same in every workgroup.
__device__ __forceinline__ double sum_parts(const double* part, int nb, double* red) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
  return block_sum_bcast(s, red);
}

__device__ __forceinline__ int chunk_of(int nb8) { return (blockIdx.x & 7) * nb8 + (blockIdx.x >> 3); }

// value of a compact-domain array at cube coordinates (0 outside the domain)
__device__ __forceinline__ double dom_at(const Fine& g, const double* __restrict__ a, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return 0.0;
  const int2 m = g.meta[j + g.ny * k];
  const int lo = m.y & 0xffff, hi = m.y >> 16;
  return (i >= lo && i < hi) ? a[(long long)m.x + i] : 0.0;
}

// ---------------------------------------------------------------- fine level
// r = rhs - A x; partials of r.r (once per solve)
template <int R>
__global__ __launch_bounds__(kB) void mg_init(Fine g, const double* __restrict__ x, const double* __restrict__ rhs,
                                              double* __restrict__ r, double* __restrict__ prr) {
  __shared__ double red[kB / 64];
  const int c = chunk_of(g.nb8);
  if (c >= g.nb) return;
  const long long p0 = (long long)c * g.chunk, p1 = min(p0 + g.chunk, g.n);
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kB) {
    const int row = g.prow[p];
    const int2 m = g.meta[row];
    const int i = (int)(p - m.x), k = row / g.ny, j = row - k * g.ny;
    double acc = g.w[0] * x[p];
#pragma unroll
    for (int d = 1; d <= R; ++d)
      acc += g.w[d] * (dom_at(g, x, i - d, j, k) + dom_at(g, x, i + d, j, k) + dom_at(g, x, i, j - d, k) +
                       dom_at(g, x, i, j + d, k) + dom_at(g, x, i, j, k - d) + dom_at(g, x, i, j, k + d));
    const double rv = rhs[p] - acc;
    r[p] = rv;
    s += rv * rv;
  }
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) prr[c] = t;
}

// t0 = r - c0 A r: the residual after one damped-Jacobi step from zero
template <int R, int SX>
__global__ __launch_bounds__(256) void mg_res0(ZMap zm, double c0, const double* __restrict__ r,
                                               double* __restrict__ t0, const double* sc) {
  __shared__ double tile[pcz::Tile<R, SX>::SIZE];
  if (sc[S_DONE] != 0.0) return;
  const int L = pcz::block_of(zm);
  if (L >= zm.nwg) return;
  pcz::zmarch<R, SX>(
      zm, L, tile, [&](long long idx, int, int, int) { return r[idx]; },
      [&](long long idx, int, int, int, double qc, double acc) { t0[idx] = qc - c0 * acc; });
}

// level-1 right-hand side: average of the domain children of each cell
__global__ __launch_bounds__(kB) void mg_restrict0(Fine g, const double* __restrict__ t0, Cube c1, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const long long q = (long long)blockIdx.x * kB + threadIdx.x;
  if (q >= c1.cells) return;
  const int I = (int)(q % c1.nx), J = (int)((q / c1.nx) % c1.ny), K = (int)(q / ((long long)c1.nx * c1.ny));
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < 8; ++d) s += dom_at(g, t0, 2 * I + (d & 1), 2 * J + ((d >> 1) & 1), 2 * K + (d >> 2));
  c1.b[q] = 0.125 * s;   // zero on inactive cells: none of their children is in the domain
}

// z = post-smooth(c0 r + P e_1) with the domain's stencil; partials of r.z
template <int R, int SX>
__global__ __launch_bounds__(256) void mg_up0(ZMap zm, double c0, const double* __restrict__ r, Cube c1,
                                              double* __restrict__ z, double* __restrict__ prz, const double* sc) {
  __shared__ double tile[pcz::Tile<R, SX>::SIZE];
  __shared__ double red[pcz::kZW];
  if (sc[S_DONE] != 0.0) return;
  const int L = pcz::block_of(zm);
  if (L >= zm.nwg) return;
  double s = 0.0;
  const double* __restrict__ e1 = c1.e;
  pcz::zmarch<R, SX>(
      zm, L, tile,
      [&](long long idx, int i, int j, int k) {
        return c0 * r[idx] + e1[((long long)(k >> 1) * c1.ny + (j >> 1)) * c1.nx + (i >> 1)];
      },
      [&](long long idx, int, int, int, double qc, double acc) {
        const double rv = r[idx];
        const double zn = qc + c0 * (rv - acc);
        z[idx] = zn;
        s += rv * zn;
      });
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) prz[L] = t;
}

// PCG step t: convergence from r_t.r_t, beta = rz_t / rz_{t-1}, p = z + beta p_old
// (at t = 0 the caller passes p_old = z and beta is 0), A p, partials of p.Ap.
// Workgroup 0 records the scalars.
template <int R, int SX>
__global__ __launch_bounds__(256) void mg_apply(ZMap zm, int t, int maxit, double tol, int nb_rr,
                                                const double* __restrict__ z, const double* __restrict__ p_old,
                                                double* __restrict__ p_new, double* __restrict__ ap,
                                                const double* __restrict__ prr, const double* __restrict__ prz,
                                                double* __restrict__ ppq, double* sc) {
  __shared__ double tile[pcz::Tile<R, SX>::SIZE];
  __shared__ double red[pcz::kZW];
  if (t > 0 && sc[S_DONE] != 0.0) return;
  const int L = pcz::block_of(zm);
  if (L >= zm.nwg) return;
  const double rr = sum_parts(prr, nb_rr, red);
  const double rz = sum_parts(prz, zm.nwg, red);
  const double tol1 = t == 0 ? tol * tol * rr : sc[S_TOL1];
  const bool done = t >= maxit || !(rr > tol1);
  const double beta = t > 0 ? rz / sc[S_RZ0 + ((t - 1) & 1)] : 0.0;
  if (L == 0 && threadIdx.x == 0) {
    sc[S_RZ0 + (t & 1)] = rz;
    if (t == 0) {
      sc[S_TOL1] = tol1;
      sc[S_RR0] = rr;
    }
    sc[S_ITS] = (double)t;
    sc[S_RES] = rr;
    sc[S_DONE] = done ? 1.0 : 0.0;
  }
  if (done) return;
  double s = 0.0;
  pcz::zmarch<R, SX>(
      zm, L, tile, [&](long long idx, int, int, int) { return z[idx] + beta * p_old[idx]; },
      [&](long long idx, int, int, int, double qc, double acc) {
        p_new[idx] = qc;
        ap[idx] = acc;
        s += qc * acc;
      });
  const double tt = pc_block_sum(s, red);
  if (threadIdx.x == 0) ppq[L] = tt;
}

// alpha = rz_t / p.Ap; x += alpha p; r -= alpha Ap; partials of r_{t+1}.r_{t+1}
__global__ __launch_bounds__(kB) void mg_update(Fine g, int npq, int t, const double* sc, double* __restrict__ x,
                                                double* __restrict__ r, const double* __restrict__ p,
                                                const double* __restrict__ ap, const double* __restrict__ ppq,
                                                double* __restrict__ prr) {
  __shared__ double red[kB / 64];
  if (sc[S_DONE] != 0.0) return;
  const int c = chunk_of(g.nb8);
  if (c >= g.nb) return;
  const double pq = sum_parts(ppq, npq, red);
  const double alpha = sc[S_RZ0 + (t & 1)] / pq;
  const long long p0 = (long long)c * g.chunk, p1 = min(p0 + g.chunk, g.n);
  double s = 0.0;
  for (long long i = p0 + threadIdx.x; i < p1; i += kB) {
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * ap[i];
    r[i] = rv;
    s += rv * rv;
  }
  const double tt = pc_block_sum(s, red);
  if (threadIdx.x == 0) prr[c] = tt;
}

// ---------------------------------------------------------------- cube levels
__device__ __forceinline__ double cube_at(const Cube& c, const double* __restrict__ a, int i, int j, int k) {
  const bool in = i >= 0 && j >= 0 && k >= 0 && i < c.nx && j < c.ny && k < c.nz;
  return in ? a[((long long)k * c.ny + j) * c.nx + i] : 0.0;
}

// b_{l+1} = R (b_l - A_l (omega b_l / D_l)) on the active cells.  One thread
// per child (8 per coarse cell, adjacent lanes), summed across the 8 lanes.
__global__ __launch_bounds__(kB) void mg_down(Cube f, Cube c, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const long long gid = (long long)blockIdx.x * kB + threadIdx.x;
  const long long q = gid >> 3;
  const int d = (int)(gid & 7);
  double v = 0.0;
  if (q < c.cells) {
    const int I = (int)(q % c.nx), J = (int)((q / c.nx) % c.ny), K = (int)(q / ((long long)c.nx * c.ny));
    const int i = 2 * I + (d & 1), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d >> 2);
    if (i < f.nx && j < f.ny && k < f.nz) {
      const long long fq = ((long long)k * f.ny + j) * f.nx + i;
      const double act = f.mask[fq] ? 1.0 : 0.0;
      const double bc = f.b[fq];
      const double nb = cube_at(f, f.b, i - 1, j, k) + cube_at(f, f.b, i + 1, j, k) + cube_at(f, f.b, i, j - 1, k) +
                        cube_at(f, f.b, i, j + 1, k) + cube_at(f, f.b, i, j, k - 1) + cube_at(f, f.b, i, j, k + 1);
      // t = b - (omega/6) (6 b - sum nb)
      v = act * ((1.0 - kOmegaC) * bc + (kOmegaC / 6.0) * nb);
    }
  }
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  if (d == 0 && q < c.cells) c.b[q] = 0.125 * v;
}

// e_l = post-smooth(omega b_l / D_l + P e_{l+1}) on the active cells
__global__ __launch_bounds__(kB) void mg_up(Cube f, Cube c, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const long long q = (long long)blockIdx.x * kB + threadIdx.x;
  if (q >= f.cells) return;
  const int i = (int)(q % f.nx), j = (int)((q / f.nx) % f.ny), k = (int)(q / ((long long)f.nx * f.ny));
  const double s0 = kOmegaC / (6.0 * f.inv_h2);   // omega / D_l
  auto ev = [&](int a, int b, int cc) -> double {
    const bool in = a >= 0 && b >= 0 && cc >= 0 && a < f.nx && b < f.ny && cc < f.nz;
    if (!in) return 0.0;
    const long long fq = ((long long)cc * f.ny + b) * f.nx + a;
    const long long cq = ((long long)(cc >> 1) * c.ny + (b >> 1)) * c.nx + (a >> 1);
    const double act = f.mask[fq] ? 1.0 : 0.0;
    return act * (s0 * f.b[fq] + c.e[cq]);
  };
  const double e0 = ev(i, j, k);
  const double ae = f.inv_h2 * (6.0 * e0 - ev(i - 1, j, k) - ev(i + 1, j, k) - ev(i, j - 1, k) - ev(i, j + 1, k) -
                                ev(i, j, k - 1) - ev(i, j, k + 1));
  f.e[q] = f.mask[q] ? e0 + s0 * (f.b[q] - ae) : 0.0;
}

// ---------------------------------------------------------------- tail
// Levels T..L (T >= 1) in LDS, one workgroup: restriction down to L, the
// Chebyshev iteration on L, prolongation and post-smoothing back up to T.
// Input b_T, output e_T (global).
struct TailLev {
  int nx, ny, nz, cells, off;   // off: first cell of the level in the LDS arrays
  double inv_h2;
  const unsigned char* mask;
};
struct Tail {
  int nlev;                     // levels T..L = lev[0..nlev)
  TailLev lev[kMaxLev];
  const double* bT;
  double* eT;
  double lmin, lmax;
};

__global__ __launch_bounds__(kTailThreads) void mg_tail(Tail tl, const double* sc) {
  __shared__ double bl[kTailCells], el[kTailCells];
  __shared__ unsigned char ml[kTailCells];
  __shared__ double xs[2][kCoarsest];
  if (sc[S_DONE] != 0.0) return;
  const int tid = threadIdx.x;
  const int ncell = tl.lev[tl.nlev - 1].off + tl.lev[tl.nlev - 1].cells;
  for (int l = 0; l < tl.nlev; ++l)
    for (int q = tid; q < tl.lev[l].cells; q += kTailThreads) ml[tl.lev[l].off + q] = tl.lev[l].mask[q];
  for (int q = tid; q < tl.lev[0].cells; q += kTailThreads) bl[q] = tl.bT[q];
  for (int q = tid; q < ncell; q += kTailThreads) el[q] = 0.0;
  __syncthreads();
  auto at = [&](const double* a, const TailLev& v, int i, int j, int k) -> double {
    const bool in = i >= 0 && j >= 0 && k >= 0 && i < v.nx && j < v.ny && k < v.nz;
    return in ? a[v.off + (k * v.ny + j) * v.nx + i] : 0.0;
  };
  // down: b_{l+1} = R (b_l - A_l omega b_l / D_l)
  for (int l = 0; l + 1 < tl.nlev; ++l) {
    const TailLev& f = tl.lev[l];
    const TailLev& c = tl.lev[l + 1];
    for (int q = tid; q < c.cells; q += kTailThreads) {
      const int I = q % c.nx, J = (q / c.nx) % c.ny, K = q / (c.nx * c.ny);
      double s = 0.0;
      for (int d = 0; d < 8; ++d) {
        const int i = 2 * I + (d & 1), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d >> 2);
        if (i >= f.nx || j >= f.ny || k >= f.nz) continue;
        const int fq = f.off + (k * f.ny + j) * f.nx + i;
        const double nb = at(bl, f, i - 1, j, k) + at(bl, f, i + 1, j, k) + at(bl, f, i, j - 1, k) +
                          at(bl, f, i, j + 1, k) + at(bl, f, i, j, k - 1) + at(bl, f, i, j, k + 1);
        s += ml[fq] ? (1.0 - kOmegaC) * bl[fq] + (kOmegaC / 6.0) * nb : 0.0;
      }
      bl[c.off + q] = 0.125 * s;
    }
    __syncthreads();
  }
  // coarsest: e = p(A) b, Chebyshev iteration on [lmin, lmax] from zero, one
  // barrier per step (x alternates between two LDS buffers)
  {
    const TailLev& c = tl.lev[tl.nlev - 1];
    const bool act = tid < c.cells && ml[c.off + tid];
    const int I = tid % c.nx, J = (tid / c.nx) % c.ny, K = tid / (c.nx * c.ny);
    const double theta = 0.5 * (tl.lmax + tl.lmin), delta = 0.5 * (tl.lmax - tl.lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    const double b = act ? bl[c.off + tid] : 0.0;
    double x = 0.0, d = b / theta;
    auto xat = [&](const double* a, int i, int j, int k) -> double {
      const bool in = i >= 0 && j >= 0 && k >= 0 && i < c.nx && j < c.ny && k < c.nz;
      return in ? a[(k * c.ny + j) * c.nx + i] : 0.0;
    };
    for (int it = 0; it < kCheb; ++it) {
      x += d;
      double* xb = xs[it & 1];
      if (tid < c.cells) xb[tid] = act ? x : 0.0;
      __syncthreads();
      double r = 0.0;
      if (act)
        r = b - c.inv_h2 * (6.0 * x - xat(xb, I - 1, J, K) - xat(xb, I + 1, J, K) - xat(xb, I, J - 1, K) -
                            xat(xb, I, J + 1, K) - xat(xb, I, J, K - 1) - xat(xb, I, J, K + 1));
      const double rho1 = 1.0 / (2.0 * sigma - rho);
      d = act ? rho1 * rho * d + 2.0 * rho1 / delta * r : 0.0;
      rho = rho1;
    }
    if (tid < c.cells) el[c.off + tid] = act ? x + d : 0.0;
    __syncthreads();
  }
  // up: e_l = post-smooth(omega b_l / D_l + P e_{l+1})
  for (int l = tl.nlev - 2; l >= 0; --l) {
    const TailLev& f = tl.lev[l];
    const TailLev& c = tl.lev[l + 1];
    const double s0 = kOmegaC / (6.0 * f.inv_h2);
    auto ev = [&](int a, int b, int cc) -> double {
      const bool in = a >= 0 && b >= 0 && cc >= 0 && a < f.nx && b < f.ny && cc < f.nz;
      if (!in) return 0.0;
      const int fq = f.off + (cc * f.ny + b) * f.nx + a;
      return ml[fq] ? s0 * bl[fq] + el[c.off + ((cc >> 1) * c.ny + (b >> 1)) * c.nx + (a >> 1)] : 0.0;
    };
    for (int q = tid; q < f.cells; q += kTailThreads) {
      const int i = q % f.nx, j = (q / f.nx) % f.ny, k = q / (f.nx * f.ny);
      const double e0 = ev(i, j, k);
      const double ae = f.inv_h2 * (6.0 * e0 - ev(i - 1, j, k) - ev(i + 1, j, k) - ev(i, j - 1, k) -
                                    ev(i, j + 1, k) - ev(i, j, k - 1) - ev(i, j, k + 1));
      const double e = ml[f.off + q] ? e0 + s0 * (bl[f.off + q] - ae) : 0.0;
      if (l == 0) tl.eT[q] = e;
      else el[f.off + q] = e;
    }
    __syncthreads();
  }
  if (tl.nlev == 1)
    for (int q = tid; q < tl.lev[0].cells; q += kTailThreads) tl.eT[q] = el[q];
}

__global__ void mg_reset(double* sc) {
  if (threadIdx.x < S_COUNT) sc[threadIdx.x] = 0.0;
}

// ---------------------------------------------------------------- host
struct MGCtx {
  void* grid;
  Fine g;
  ZMap zm;
  int radius;
  int nlev;                   // cube levels 1..nlev
  int tail;                   // first level held by mg_tail
  Cube lev[kMaxLev + 1];      // lev[1..nlev]
  Tail tl;
  std::vector<void*> allocs;
  double* t0;
  int first_check;            // iterations before the first state readback (0: check_every)
};

size_t work_bytes_for(long long n) {
  return (size_t)(5 * n + 2 * kNB + 2 * pcz::kMaxWG + S_COUNT) * sizeof(double);
}

template <int R, int SX>
struct FineLaunch {
  static void res0(const MGCtx* c, hipStream_t st, const double* r, const double* sc) {
    hipLaunchKernelGGL((mg_res0<R, SX>), dim3(8 * c->zm.nwg8), dim3(256), 0, st, c->zm, c->g.c0, r, c->t0, sc);
  }
  static void up0(const MGCtx* c, hipStream_t st, const double* r, double* z, double* prz, const double* sc) {
    hipLaunchKernelGGL((mg_up0<R, SX>), dim3(8 * c->zm.nwg8), dim3(256), 0, st, c->zm, c->g.c0, r, c->lev[1], z, prz,
                       sc);
  }
  static void apply(const MGCtx* c, hipStream_t st, int t, int maxit, double tol, const double* z,
                    const double* p_old, double* p_new, double* ap, const double* prr, const double* prz, double* ppq,
                    double* sc) {
    hipLaunchKernelGGL((mg_apply<R, SX>), dim3(8 * c->zm.nwg8), dim3(256), 0, st, c->zm, t, maxit, tol, c->g.nb, z,
                       p_old, p_new, ap, prr, prz, ppq, sc);
  }
  static void init(const MGCtx* c, hipStream_t st, const double* x, const double* rhs, double* r, double* prr) {
    hipLaunchKernelGGL((mg_init<R>), dim3(8 * c->g.nb8), dim3(kB), 0, st, c->g, x, rhs, r, prr);
  }
};

struct FineOps {
  void (*res0)(const MGCtx*, hipStream_t, const double*, const double*);
  void (*up0)(const MGCtx*, hipStream_t, const double*, double*, double*, const double*);
  void (*apply)(const MGCtx*, hipStream_t, int, int, double, const double*, const double*, double*, double*,
                const double*, const double*, double*, double*);
  void (*init)(const MGCtx*, hipStream_t, const double*, const double*, double*, double*);
};

template <int R, int SX>
FineOps fine_ops_rs() {
  using F = FineLaunch<R, SX>;
  return {F::res0, F::up0, F::apply, F::init};
}

template <int R>
FineOps fine_ops_r(int sx) {
  if (sx == 1) return fine_ops_rs<R, 1>();
  if (sx == 2) return fine_ops_rs<R, 2>();
  return fine_ops_rs<R, 4>();
}

bool fine_ops(int radius, int sx, FineOps& ops) {
  switch (radius) {
    case 1: ops = fine_ops_r<1>(sx); return true;
    case 2: ops = fine_ops_r<2>(sx); return true;
    case 3: ops = fine_ops_r<3>(sx); return true;
    case 4: ops = fine_ops_r<4>(sx); return true;
    case 6: ops = fine_ops_r<6>(sx); return true;
    default: return false;
  }
}

}  // namespace

extern "C" {

// Builds the cube hierarchy for `grid` (spacing h): masks, right-hand-side and
// correction arrays on every level (a few MB for a 106^3 box).
int pc_mg_create(void* grid, double h, void** out_mg) {
  PC_CHECK_ARG(grid && out_mg && h > 0, "bad argument");
  MGCtx* c = new MGCtx();
  c->first_check = 0;
  c->grid = grid;
  Fine& g = c->g;
  double w8[8];
  pc_internal_grid_points(grid, &g.ny, &g.nz, &c->radius, &g.meta, &g.prow, w8);
  const int* zr;
  pc_internal_grid_zmap(grid, &g.nx, &zr);
  g.n = pc_internal_grid_npts(grid);
  const double inv_h2 = 1.0 / (h * h);
  for (int d = 0; d < 8; ++d) g.w[d] = w8[d] * inv_h2;
  g.w[0] *= 3.0;
  g.c0 = kOmega0 / g.w[0];
  const long long bp = (g.n + kB - 1) / kB;
  g.nb = (int)std::min<long long>(kNB, bp);
  g.chunk = ((bp + g.nb - 1) / g.nb) * kB;
  g.nb = (int)((g.n + g.chunk - 1) / g.chunk);
  g.nb8 = (g.nb + 7) / 8;
  int rz = 0;
  if (pcz::zmap_init(grid, g.w, c->zm, &rz)) {
    delete c;
    PC_CHECK_ARG(false, "nx > 256");
  }
  FineOps ops;
  if (!fine_ops(c->radius, c->zm.sx, ops)) {
    delete c;
    PC_CHECK_ARG(false, "unsupported stencil radius");
  }
  // domain mask on the full cube (host)
  int nx, ny, nz;
  const int *lo_d, *hi_d, *off_d;
  pc_internal_grid_view(grid, &nx, &ny, &nz, &lo_d, &hi_d, &off_d);
  std::vector<int> lo((size_t)ny * nz), hi((size_t)ny * nz);
  PC_HIP(hipMemcpy(lo.data(), lo_d, lo.size() * sizeof(int), hipMemcpyDeviceToHost));
  PC_HIP(hipMemcpy(hi.data(), hi_d, hi.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<unsigned char> prev((size_t)nx * ny * nz, 0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = lo[j + ny * k]; i < hi[j + ny * k]; ++i) prev[((size_t)k * ny + j) * nx + i] = 1;
  int pn[3] = {nx, ny, nz};
  double hl = h;
  c->nlev = 0;
  while (c->nlev < kMaxLev) {
    const int cn[3] = {(pn[0] + 1) / 2, (pn[1] + 1) / 2, (pn[2] + 1) / 2};
    std::vector<unsigned char> m((size_t)cn[0] * cn[1] * cn[2], 0);
    for (int k = 0; k < pn[2]; ++k)
      for (int j = 0; j < pn[1]; ++j)
        for (int i = 0; i < pn[0]; ++i)
          if (prev[((size_t)k * pn[1] + j) * pn[0] + i]) m[((size_t)(k / 2) * cn[1] + j / 2) * cn[0] + i / 2] = 1;
    hl *= 2.0;
    Cube& L = c->lev[++c->nlev];
    L.nx = cn[0]; L.ny = cn[1]; L.nz = cn[2];
    L.cells = (long long)cn[0] * cn[1] * cn[2];
    L.inv_h2 = 1.0 / (hl * hl);
    unsigned char* md = nullptr;
    double *bd = nullptr, *ed = nullptr;
    PC_HIP(hipMalloc(&md, L.cells));
    PC_HIP(hipMalloc(&bd, L.cells * sizeof(double)));
    PC_HIP(hipMalloc(&ed, L.cells * sizeof(double)));
    c->allocs.push_back(md); c->allocs.push_back(bd); c->allocs.push_back(ed);
    PC_HIP(hipMemcpy(md, m.data(), L.cells, hipMemcpyHostToDevice));
    PC_HIP(hipMemset(bd, 0, L.cells * sizeof(double)));
    PC_HIP(hipMemset(ed, 0, L.cells * sizeof(double)));
    L.mask = md; L.b = bd; L.e = ed;
    prev.swap(m);
    pn[0] = cn[0]; pn[1] = cn[1]; pn[2] = cn[2];
    if (L.cells <= kCoarsest) break;
  }
  PC_CHECK_ARG(c->lev[c->nlev].cells <= kCoarsest, "grid too large for the multigrid hierarchy");
  // tail: the deepest run of levels whose cells fit the LDS arrays
  c->tail = c->nlev;
  long long tcells = c->lev[c->nlev].cells;
  while (c->tail > 1 && tcells + c->lev[c->tail - 1].cells <= kTailCells) tcells += c->lev[--c->tail].cells;
  Tail& tl = c->tl;
  tl.nlev = c->nlev - c->tail + 1;
  int off = 0;
  for (int l = 0; l < tl.nlev; ++l) {
    const Cube& L = c->lev[c->tail + l];
    tl.lev[l] = {L.nx, L.ny, L.nz, (int)L.cells, off, L.inv_h2, L.mask};
    off += (int)L.cells;
  }
  tl.bT = c->lev[c->tail].b;
  tl.eT = c->lev[c->tail].e;
  // Chebyshev interval of the coarsest 7-point operator: lambda_max <= 12/h^2
  // (Gershgorin); lambda_min >= that of the full box (Dirichlet on a subset)
  const Cube& C = c->lev[c->nlev];
  auto lam1 = [](int n) { return 2.0 * (1.0 - std::cos(M_PI / (n + 1))); };
  tl.lmin = C.inv_h2 * (lam1(C.nx) + lam1(C.ny) + lam1(C.nz));
  tl.lmax = 12.0 * C.inv_h2;
  PC_HIP(hipMalloc(&c->t0, g.n * sizeof(double)));
  c->allocs.push_back(c->t0);
  *out_mg = c;
  return pc::PC_OK;
}

int pc_mg_destroy(void* mg) {
  MGCtx* c = static_cast<MGCtx*>(mg);
  if (!c) return pc::PC_OK;
  (void)hipDeviceSynchronize();   // queued V-cycles may still read the level arrays
  for (void* p : c->allocs) (void)hipFree(p);
  delete c;
  return pc::PC_OK;
}

int pc_mg_workspace(void* mg, size_t* out_bytes) {
  MGCtx* c = static_cast<MGCtx*>(mg);
  PC_CHECK_ARG(c && out_bytes, "null pointer");
  *out_bytes = work_bytes_for(c->g.n);
  return pc::PC_OK;
}

// Iterations before the first state readback of pc_pcg_solve_mg (0: its
// check_every); the results do not depend on it.
int pc_mg_set_first_check(void* mg, int first) {
  MGCtx* c = static_cast<MGCtx*>(mg);
  PC_CHECK_ARG(c && first >= 0, "bad argument");
  c->first_check = first;
  return pc::PC_OK;
}

// Multigrid-preconditioned CG: same arguments and stop rule as pc_pcg_solve
// (no first-iteration aliasing: this is the textbook preconditioned method).
int pc_pcg_solve_mg(void* mg, const double* rhs, double* x, int maxit, double tol, int check_every, void* work,
                    size_t work_bytes, int* its_out, double* res_out, void* stream) {
  MGCtx* c = static_cast<MGCtx*>(mg);
  PC_CHECK_ARG(c && rhs && x && work && its_out, "null pointer");
  PC_CHECK_ARG(work_bytes >= work_bytes_for(c->g.n), "workspace too small");
  PC_CHECK_ARG(((uintptr_t)work & 7) == 0, "workspace not 8-byte aligned");
  PC_CHECK_ARG(maxit >= 0, "maxit < 0");
  if (check_every < 1) check_every = 8;
  hipStream_t st = static_cast<hipStream_t>(stream);
  FineOps ops;
  fine_ops(c->radius, c->zm.sx, ops);
  const long long n = c->g.n;
  double* r = static_cast<double*>(work);
  double* z = r + n;
  double* p[2] = {z + n, z + 2 * n};
  double* ap = z + 3 * n;
  double* prr[2] = {ap + n, ap + n + kNB};
  double* prz = prr[1] + kNB;
  double* ppq = prz + pcz::kMaxWG;
  double* sc = ppq + pcz::kMaxWG;
  const int T = c->tail;
  auto vcycle = [&](const double* rr, double* zz) {
    ops.res0(c, st, rr, sc);
    hipLaunchKernelGGL(mg_restrict0, dim3((unsigned)((c->lev[1].cells + kB - 1) / kB)), dim3(kB), 0, st, c->g, c->t0,
                       c->lev[1], sc);
    for (int l = 1; l < T; ++l)
      hipLaunchKernelGGL(mg_down, dim3((unsigned)((8 * c->lev[l + 1].cells + kB - 1) / kB)), dim3(kB), 0, st,
                         c->lev[l], c->lev[l + 1], sc);
    hipLaunchKernelGGL(mg_tail, dim3(1), dim3(kTailThreads), 0, st, c->tl, sc);
    for (int l = T - 1; l >= 1; --l)
      hipLaunchKernelGGL(mg_up, dim3((unsigned)((c->lev[l].cells + kB - 1) / kB)), dim3(kB), 0, st, c->lev[l],
                         c->lev[l + 1], sc);
    ops.up0(c, st, rr, zz, prz, sc);
  };
  hipLaunchKernelGGL(mg_reset, dim3(1), dim3(64), 0, st, sc);
  ops.init(c, st, x, rhs, r, prr[0]);
  vcycle(r, z);
  PC_LAUNCH_CHECK();
  static thread_local double* host = nullptr;
  if (!host) PC_HIP(hipHostMalloc(&host, S_COUNT * sizeof(double), hipHostMallocDefault));
  int status = pc::PC_OK;
  // the state is read back after iteration first_check (when set), then every
  // check_every iterations; kernels queued past convergence return at once
  int t = 0;
  int next = c->first_check > 0 ? c->first_check : check_every;
  while (t <= maxit) {
    for (; t < next && t <= maxit; ++t) {
      // at t = 0 the direction is z itself: p_old = z with beta = 0
      const double* p_old = t == 0 ? z : p[t & 1];
      ops.apply(c, st, t, maxit, tol, z, p_old, p[(t + 1) & 1], ap, prr[t & 1], prz, ppq, sc);
      if (t < maxit) {
        hipLaunchKernelGGL(mg_update, dim3(8 * c->g.nb8), dim3(kB), 0, st, c->g, c->zm.nwg, t, sc, x, r,
                           p[(t + 1) & 1], ap, ppq, prr[(t + 1) & 1]);
        vcycle(r, z);
      }
      if (hipGetLastError() != hipSuccess) {
        pc::set_error("pc_pcg_solve_mg: kernel launch failed");
        status = pc::PC_ERR_HIP;
        break;
      }
    }
    if (status) break;
    if (hipMemcpyAsync(host, sc, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      pc::set_error("pc_pcg_solve_mg: readback failed");
      status = pc::PC_ERR_HIP;
      break;
    }
    if (host[S_DONE] != 0.0) break;
    next = t + check_every;
  }
  if (!status) {
    if (host[S_DONE] == 0.0) {
      pc::set_error("pc_pcg_solve_mg: internal error (no final state)");
      return pc::PC_ERR_HIP;
    }
    *its_out = (int)host[S_ITS];
    if (res_out) {
      res_out[0] = host[S_RES];
      res_out[1] = host[S_TOL1];
    }
  }
  return status;
}

}  // extern "C"
