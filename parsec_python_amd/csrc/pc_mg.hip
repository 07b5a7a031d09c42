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
// operator and CG theory applies).  Each level costs two launches: the
// pre-smooth, residual and restriction are one kernel (from a zero start
// the smoothed value is omega b / D, formed on the fly), and the
// prolongation with the post-smooth is another.
#include "pc_common.h"
#include <algorithm>
#include <cmath>
#include <vector>

long long pc_internal_grid_npts(void* grid);
int pc_internal_grid_points(void* grid, int* ny, int* nz, int* radius, const int2** meta, const int** prow,
                            double* w8);
int pc_internal_grid_zmap(void* grid, int* nx, const int** zrange);
int pc_internal_grid_view(void* grid, int* nx, int* ny, int* nz, const int** lo, const int** hi, const int** off);

namespace {

constexpr int kB = 256;
constexpr int kNB = 512;          // workgroups of the point-wise launches (= partials)
constexpr int kMaxLev = 8;
constexpr int kCheb = 16;         // Chebyshev steps on the coarsest cube
constexpr double kOmega0 = 0.6;   // damped Jacobi on the domain's stencil (lambda_max(D^-1 A) < 2.3 for fd8)
constexpr double kOmegaC = 0.8;   // damped Jacobi on the 7-point cubes (lambda_max(D^-1 A) < 2)
enum { S_RR0 = 0, S_RZ0, S_RZ1, S_TOL1, S_DONE, S_ITS, S_RES, S_COUNT = 8 };

struct Fine {
  const int2* meta;
  const int* prow;
  long long n;
  int nx, ny, nz;
  double w[8];        // weights of A (w[0] = centre)
  double c0;          // omega0 / w[0]
  long long chunk;
  int nb;
};

struct Cube {
  int nx, ny, nz;
  long long cells;
  double inv_h2;      // 1 / (2^l h)^2
  const unsigned char* mask;
  double* b;
  double* e;
};

__device__ __forceinline__ void row_of(const Fine& g, long long p, int& row, int2& m, int& i, int& j, int& k) {
  row = g.prow[p];
  m = g.meta[row];
  i = (int)(p - m.x);
  k = row / g.ny;
  j = row - k * g.ny;
}

// value of a compact-domain array at cube coordinates (0 outside the domain)
__device__ __forceinline__ double dom_at(const Fine& g, const double* __restrict__ a, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return 0.0;
  const int2 m = g.meta[j + g.ny * k];
  const int lo = m.y & 0xffff, hi = m.y >> 16;
  return (i >= lo && i < hi) ? a[(long long)m.x + i] : 0.0;
}

__device__ __forceinline__ double block_sum_bcast(double v, double* red) {
  const double t = pc_block_sum(v, red);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = t;
  __syncthreads();
  return bc;
}

__device__ __forceinline__ double sum_parts(const double* part, int nb, double* red) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
  return block_sum_bcast(s, red);
}

// A applied at point p to a field given by f(i, j, k) (FP64), radius R
template <int R, typename F>
__device__ __forceinline__ double apply_A(const Fine& g, int i, int j, int k, double fc, F f) {
  double acc = g.w[0] * fc;
#pragma unroll
  for (int d = 1; d <= R; ++d)
    acc += g.w[d] * (f(i - d, j, k) + f(i + d, j, k) + f(i, j - d, k) + f(i, j + d, k) + f(i, j, k - d) +
                     f(i, j, k + d));
  return acc;
}

// ---------------------------------------------------------------- fine level
// r = rhs - A x; partials of r.r
template <int R>
__global__ __launch_bounds__(kB) void mg_init(Fine g, const double* __restrict__ x, const double* __restrict__ rhs,
                                              double* __restrict__ r, double* __restrict__ prr) {
  __shared__ double red[kB / 64];
  const long long p0 = (long long)blockIdx.x * g.chunk, p1 = min(p0 + g.chunk, g.n);
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kB) {
    int row, i, j, k;
    int2 m;
    row_of(g, p, row, m, i, j, k);
    auto f = [&](int a, int b, int c) { return dom_at(g, x, a, b, c); };
    const double rv = rhs[p] - apply_A<R>(g, i, j, k, x[p], f);
    r[p] = rv;
    s += rv * rv;
  }
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) prr[blockIdx.x] = t;
}

// t0 = r - A (c0 r): the residual after one damped-Jacobi step from zero
template <int R>
__global__ __launch_bounds__(kB) void mg_res0(Fine g, const double* __restrict__ r, double* __restrict__ t0,
                                              const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const long long p = (long long)blockIdx.x * kB + threadIdx.x;
  if (p >= g.n) return;
  int row, i, j, k;
  int2 m;
  row_of(g, p, row, m, i, j, k);
  auto f = [&](int a, int b, int c) { return dom_at(g, r, a, b, c); };
  t0[p] = r[p] - g.c0 * apply_A<R>(g, i, j, k, r[p], f);
}

// level-1 right-hand side: average of the domain children of each cell
__global__ void mg_restrict0(Fine g, const double* __restrict__ t0, Cube c1, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const long long q = (long long)blockIdx.x * kB + threadIdx.x;
  if (q >= c1.cells) return;
  const int I = (int)(q % c1.nx), J = (int)((q / c1.nx) % c1.ny), K = (int)(q / ((long long)c1.nx * c1.ny));
  double s = 0.0;
  if (c1.mask[q]) {
#pragma unroll
    for (int d = 0; d < 8; ++d) s += dom_at(g, t0, 2 * I + (d & 1), 2 * J + ((d >> 1) & 1), 2 * K + (d >> 2));
  }
  c1.b[q] = 0.125 * s;
}

// 7-point value helpers on a cube level
__device__ __forceinline__ double cube_at(const Cube& c, const double* __restrict__ a, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= c.nx || j >= c.ny || k >= c.nz) return 0.0;
  const long long q = ((long long)k * c.ny + j) * c.nx + i;
  return c.mask[q] ? a[q] : 0.0;
}

// b_{l+1} = R (b_l - A_l (omega b_l / D_l)) on the active cells
__global__ void mg_down(Cube f, Cube c, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const long long q = (long long)blockIdx.x * kB + threadIdx.x;
  if (q >= c.cells) return;
  const int I = (int)(q % c.nx), J = (int)((q / c.nx) % c.ny), K = (int)(q / ((long long)c.nx * c.ny));
  double s = 0.0;
  if (c.mask[q]) {
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const int i = 2 * I + (d & 1), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d >> 2);
      if (i >= f.nx || j >= f.ny || k >= f.nz) continue;
      const long long fq = ((long long)k * f.ny + j) * f.nx + i;
      if (!f.mask[fq]) continue;
      const double bc = f.b[fq];
      const double nb = cube_at(f, f.b, i - 1, j, k) + cube_at(f, f.b, i + 1, j, k) + cube_at(f, f.b, i, j - 1, k) +
                        cube_at(f, f.b, i, j + 1, k) + cube_at(f, f.b, i, j, k - 1) + cube_at(f, f.b, i, j, k + 1);
      // t = b - (omega/6) (6 b - sum nb)
      s += (1.0 - kOmegaC) * bc + (kOmegaC / 6.0) * nb;
    }
  }
  c.b[q] = 0.125 * s;
}

// coarsest cube: e = p(A) b, Chebyshev iteration on [lmin, lmax] from zero,
// one workgroup, the level in LDS
__global__ __launch_bounds__(512) void mg_coarsest(Cube c, double lmin, double lmax, const double* sc) {
  __shared__ double xs[512], rs[512], ds[512];
  if (sc[S_DONE] != 0.0) return;
  const int q = threadIdx.x;
  const bool act = q < c.cells && c.mask[q];
  const int I = q % c.nx, J = (q / c.nx) % c.ny, K = q / (c.nx * c.ny);
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  const double b = act ? c.b[q] : 0.0;
  xs[q] = 0.0;
  rs[q] = b;
  ds[q] = b / theta;
  __syncthreads();
  auto at = [&](const double* a, int i, int j, int k) -> double {
    if (i < 0 || j < 0 || k < 0 || i >= c.nx || j >= c.ny || k >= c.nz) return 0.0;
    const int qq = (k * c.ny + j) * c.nx + i;
    return (qq < c.cells && c.mask[qq]) ? a[qq] : 0.0;
  };
  for (int it = 0; it < kCheb; ++it) {
    double x = xs[q] + ds[q];
    __syncthreads();
    xs[q] = act ? x : 0.0;
    __syncthreads();
    // r = b - A x
    double r = 0.0;
    if (act) {
      const double ax = c.inv_h2 * (6.0 * xs[q] - at(xs, I - 1, J, K) - at(xs, I + 1, J, K) - at(xs, I, J - 1, K) -
                                    at(xs, I, J + 1, K) - at(xs, I, J, K - 1) - at(xs, I, J, K + 1));
      r = b - ax;
    }
    const double rho1 = 1.0 / (2.0 * sigma - rho);
    const double d = rho1 * rho * ds[q] + 2.0 * rho1 / delta * r;
    rho = rho1;
    __syncthreads();
    ds[q] = act ? d : 0.0;
    __syncthreads();
  }
  if (q < c.cells) c.e[q] = act ? xs[q] + ds[q] : 0.0;
}

// e_l = post-smooth(omega b_l / D_l + P e_{l+1}) on the active cells
__global__ void mg_up(Cube f, Cube c, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const long long q = (long long)blockIdx.x * kB + threadIdx.x;
  if (q >= f.cells) return;
  if (!f.mask[q]) {
    f.e[q] = 0.0;
    return;
  }
  const int i = (int)(q % f.nx), j = (int)((q / f.nx) % f.ny), k = (int)(q / ((long long)f.nx * f.ny));
  const double s0 = kOmegaC / (6.0 * f.inv_h2);   // omega / D_l
  auto ev = [&](int a, int b, int cc) -> double {
    if (a < 0 || b < 0 || cc < 0 || a >= f.nx || b >= f.ny || cc >= f.nz) return 0.0;
    const long long fq = ((long long)cc * f.ny + b) * f.nx + a;
    if (!f.mask[fq]) return 0.0;
    const long long cq = ((long long)(cc >> 1) * c.ny + (b >> 1)) * c.nx + (a >> 1);
    return s0 * f.b[fq] + c.e[cq];
  };
  const double e0 = ev(i, j, k);
  const double ae = f.inv_h2 * (6.0 * e0 - ev(i - 1, j, k) - ev(i + 1, j, k) - ev(i, j - 1, k) - ev(i, j + 1, k) -
                                ev(i, j, k - 1) - ev(i, j, k + 1));
  f.e[q] = e0 + s0 * (f.b[q] - ae);
}

// z = post-smooth(c0 r + P e_1) with the domain's stencil; partials of r.z
template <int R>
__global__ __launch_bounds__(kB) void mg_up0(Fine g, const double* __restrict__ r, Cube c1, double* __restrict__ z,
                                             double* __restrict__ prz, const double* sc) {
  __shared__ double red[kB / 64];
  if (sc[S_DONE] != 0.0) return;
  const long long p0 = (long long)blockIdx.x * g.chunk, p1 = min(p0 + g.chunk, g.n);
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kB) {
    int row, i, j, k;
    int2 m;
    row_of(g, p, row, m, i, j, k);
    auto zv = [&](int a, int b, int cc) -> double {
      if (a < 0 || b < 0 || cc < 0 || a >= g.nx || b >= g.ny || cc >= g.nz) return 0.0;
      const int2 mm = g.meta[b + g.ny * cc];
      const int lo = mm.y & 0xffff, hi = mm.y >> 16;
      if (a < lo || a >= hi) return 0.0;
      const long long cq = ((long long)(cc >> 1) * c1.ny + (b >> 1)) * c1.nx + (a >> 1);
      return g.c0 * r[(long long)mm.x + a] + c1.e[cq];
    };
    const double zc = zv(i, j, k);
    const double zn = zc + g.c0 * (r[p] - apply_A<R>(g, i, j, k, zc, zv));
    z[p] = zn;
    s += r[p] * zn;
  }
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) prz[blockIdx.x] = t;
}

// PCG step t: convergence from r_t.r_t, beta = rz_t / rz_{t-1}, p = z + beta p_old,
// Ap, partials of p.Ap.  Workgroup 0 records the scalars.
template <int R>
__global__ __launch_bounds__(kB) void mg_apply(Fine g, int t, int maxit, double tol, const double* __restrict__ z,
                                               const double* __restrict__ p_old, double* __restrict__ p_new,
                                               double* __restrict__ ap, const double* __restrict__ prr,
                                               const double* __restrict__ prz, double* __restrict__ ppq, double* sc) {
  __shared__ double red[kB / 64];
  if (t > 0 && sc[S_DONE] != 0.0) return;
  const double rr = sum_parts(prr, g.nb, red);
  const double rz = sum_parts(prz, g.nb, red);
  const double tol1 = t == 0 ? tol * tol * rr : sc[S_TOL1];
  const bool done = t >= maxit || !(rr > tol1);
  const double beta = t > 0 ? rz / sc[S_RZ0 + ((t - 1) & 1)] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
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
  const long long p0 = (long long)blockIdx.x * g.chunk, p1 = min(p0 + g.chunk, g.n);
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kB) {
    int row, i, j, k;
    int2 m;
    row_of(g, p, row, m, i, j, k);
    auto pv = [&](int a, int b, int cc) -> double {
      if (a < 0 || b < 0 || cc < 0 || a >= g.nx || b >= g.ny || cc >= g.nz) return 0.0;
      const int2 mm = g.meta[b + g.ny * cc];
      const int lo = mm.y & 0xffff, hi = mm.y >> 16;
      if (a < lo || a >= hi) return 0.0;
      const long long idx = (long long)mm.x + a;
      return t == 0 ? z[idx] : z[idx] + beta * p_old[idx];
    };
    const double pc = pv(i, j, k);
    const double apv = apply_A<R>(g, i, j, k, pc, pv);
    p_new[p] = pc;
    ap[p] = apv;
    s += pc * apv;
  }
  const double tt = pc_block_sum(s, red);
  if (threadIdx.x == 0) ppq[blockIdx.x] = tt;
}

// alpha = rz_t / p.Ap; x += alpha p; r -= alpha Ap; partials of r_{t+1}.r_{t+1}
__global__ __launch_bounds__(kB) void mg_update(Fine g, int t, const double* sc, double* __restrict__ x,
                                                double* __restrict__ r, const double* __restrict__ p,
                                                const double* __restrict__ ap, const double* __restrict__ ppq,
                                                double* __restrict__ prr) {
  __shared__ double red[kB / 64];
  if (sc[S_DONE] != 0.0) return;
  const double pq = sum_parts(ppq, g.nb, red);
  const double alpha = sc[S_RZ0 + (t & 1)] / pq;
  const long long p0 = (long long)blockIdx.x * g.chunk, p1 = min(p0 + g.chunk, g.n);
  double s = 0.0;
  for (long long i = p0 + threadIdx.x; i < p1; i += kB) {
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * ap[i];
    r[i] = rv;
    s += rv * rv;
  }
  const double tt = pc_block_sum(s, red);
  if (threadIdx.x == 0) prr[blockIdx.x] = tt;
}

__global__ void mg_reset(double* sc) {
  if (threadIdx.x < S_COUNT) sc[threadIdx.x] = 0.0;
}

struct MGCtx {
  void* grid;
  Fine g;
  int radius;
  int nlev;                   // cube levels 1..nlev
  Cube lev[kMaxLev + 1];      // lev[1..nlev]
  std::vector<void*> allocs;
  double lmin, lmax;          // Chebyshev interval on the coarsest cube
  double* t0;
};

}  // namespace

extern "C" {

// Builds the cube hierarchy for `grid` (spacing h): masks, right-hand-side and
// correction arrays on every level (a few MB for a 106^3 box).
int pc_mg_create(void* grid, double h, void** out_mg) {
  PC_CHECK_ARG(grid && out_mg && h > 0, "bad argument");
  MGCtx* c = new MGCtx();
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
    L.mask = md; L.b = bd; L.e = ed;
    prev.swap(m);
    pn[0] = cn[0]; pn[1] = cn[1]; pn[2] = cn[2];
    if (L.cells <= 512) break;
  }
  PC_CHECK_ARG(c->lev[c->nlev].cells <= 512, "grid too large for the multigrid hierarchy");
  // Chebyshev interval of the coarsest 7-point operator: lambda_max <= 12/h^2
  // (Gershgorin); lambda_min >= that of the full box (Dirichlet on a subset)
  const Cube& C = c->lev[c->nlev];
  auto lam1 = [](int n) { return 2.0 * (1.0 - std::cos(M_PI / (n + 1))); };
  c->lmin = C.inv_h2 * (lam1(C.nx) + lam1(C.ny) + lam1(C.nz));
  c->lmax = 12.0 * C.inv_h2;
  PC_HIP(hipMalloc(&c->t0, g.n * sizeof(double)));
  c->allocs.push_back(c->t0);
  *out_mg = c;
  return pc::PC_OK;
}

int pc_mg_destroy(void* mg) {
  MGCtx* c = static_cast<MGCtx*>(mg);
  if (!c) return pc::PC_OK;
  for (void* p : c->allocs) (void)hipFree(p);
  delete c;
  return pc::PC_OK;
}

int pc_mg_workspace(void* mg, size_t* out_bytes) {
  MGCtx* c = static_cast<MGCtx*>(mg);
  PC_CHECK_ARG(c && out_bytes, "null pointer");
  *out_bytes = (size_t)(5 * c->g.n + 3 * kNB + 2 * kNB + S_COUNT) * sizeof(double);
  return pc::PC_OK;
}

// Multigrid-preconditioned CG: same arguments and stop rule as pc_pcg_solve
// (no first-iteration aliasing: this is the textbook preconditioned method).
int pc_pcg_solve_mg(void* mg, const double* rhs, double* x, int maxit, double tol, int check_every, void* work,
                    size_t work_bytes, int* its_out, double* res_out, void* stream) {
  MGCtx* c = static_cast<MGCtx*>(mg);
  PC_CHECK_ARG(c && rhs && x && work && its_out, "null pointer");
  size_t need = 0;
  pc_mg_workspace(mg, &need);
  PC_CHECK_ARG(work_bytes >= need, "workspace too small");
  PC_CHECK_ARG(((uintptr_t)work & 7) == 0, "workspace not 8-byte aligned");
  PC_CHECK_ARG(maxit >= 0, "maxit < 0");
  if (check_every < 1) check_every = 8;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Fine& g = c->g;
  const long long n = g.n;
  double* r = static_cast<double*>(work);
  double* z = r + n;
  double* p[2] = {z + n, z + 2 * n};
  double* ap = z + 3 * n;
  double* prr[2] = {ap + n, ap + n + kNB};
  double* prz = prr[1] + kNB;
  double* ppq = prz + kNB;
  double* sc = ppq + 2 * kNB;
  const int nbp = (int)((n + kB - 1) / kB);
  const int L = c->nlev;
  auto vcycle = [&](const double* rr, double* zz) {
    switch (c->radius) {
#define PC_MG_R(RR)                                                                                       \
  case RR:                                                                                                \
    hipLaunchKernelGGL(mg_res0<RR>, dim3(nbp), dim3(kB), 0, st, g, rr, c->t0, sc);                       \
    break;
      PC_MG_R(1) PC_MG_R(2) PC_MG_R(3) PC_MG_R(4) PC_MG_R(6)
#undef PC_MG_R
    }
    hipLaunchKernelGGL(mg_restrict0, dim3((unsigned)((c->lev[1].cells + kB - 1) / kB)), dim3(kB), 0, st, g, c->t0,
                       c->lev[1], sc);
    for (int l = 1; l < L; ++l)
      hipLaunchKernelGGL(mg_down, dim3((unsigned)((c->lev[l + 1].cells + kB - 1) / kB)), dim3(kB), 0, st,
                         c->lev[l], c->lev[l + 1], sc);
    hipLaunchKernelGGL(mg_coarsest, dim3(1), dim3(512), 0, st, c->lev[L], c->lmin, c->lmax, sc);
    for (int l = L - 1; l >= 1; --l)
      hipLaunchKernelGGL(mg_up, dim3((unsigned)((c->lev[l].cells + kB - 1) / kB)), dim3(kB), 0, st, c->lev[l],
                         c->lev[l + 1], sc);
    switch (c->radius) {
#define PC_MG_R(RR)                                                                                        \
  case RR:                                                                                                 \
    hipLaunchKernelGGL(mg_up0<RR>, dim3(g.nb), dim3(kB), 0, st, g, rr, c->lev[1], zz, prz, sc);          \
    break;
      PC_MG_R(1) PC_MG_R(2) PC_MG_R(3) PC_MG_R(4) PC_MG_R(6)
#undef PC_MG_R
    }
  };
  hipLaunchKernelGGL(mg_reset, dim3(1), dim3(64), 0, st, sc);
  switch (c->radius) {
#define PC_MG_R(RR)                                                                                            \
  case RR:                                                                                                     \
    hipLaunchKernelGGL(mg_init<RR>, dim3(g.nb), dim3(kB), 0, st, g, x, rhs, r, prr[0]);                       \
    break;
    PC_MG_R(1) PC_MG_R(2) PC_MG_R(3) PC_MG_R(4) PC_MG_R(6)
#undef PC_MG_R
    default: PC_CHECK_ARG(false, "unsupported stencil radius");
  }
  vcycle(r, z);
  PC_LAUNCH_CHECK();
  static thread_local double* host = nullptr;
  if (!host) PC_HIP(hipHostMalloc(&host, S_COUNT * sizeof(double), hipHostMallocDefault));
  int status = pc::PC_OK;
  for (int t0 = 0; t0 <= maxit; t0 += check_every) {
    for (int t = t0; t < t0 + check_every && t <= maxit; ++t) {
      switch (c->radius) {
#define PC_MG_R(RR)                                                                                          \
  case RR:                                                                                                   \
    hipLaunchKernelGGL(mg_apply<RR>, dim3(g.nb), dim3(kB), 0, st, g, t, maxit, tol, z, p[t & 1],             \
                       p[(t + 1) & 1], ap, prr[t & 1], prz, ppq, sc);                                        \
    break;
        PC_MG_R(1) PC_MG_R(2) PC_MG_R(3) PC_MG_R(4) PC_MG_R(6)
#undef PC_MG_R
      }
      if (t < maxit) {
        hipLaunchKernelGGL(mg_update, dim3(g.nb), dim3(kB), 0, st, g, t, sc, x, r, p[(t + 1) & 1], ap, ppq,
                           prr[(t + 1) & 1]);
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
