// Ionic setup on the device: superposition of atomic splines at every domain
// point (src/V_ion/pseudoDiag.py:199-278; the reference evaluates the three
// splines point by point in Python, 84 s for C10H8 on its CPU log).
//   pot   += pot_S(d)                         (local pseudopotential, Ha)
//   rho   += max(0, charge(d) * h^3 / (4 pi))  (SAD density, electrons/cell)
//   hpot0 += hartree(d)                       (atomic Hartree potential, Ha)
// with d the distance from grid point (i*h - rad, j*h - rad, k*h - rad) to the
// atom.  Splines follow src/Splines/fsplevalIO.py:2-65: clamp into the knot
// range, locate the bracketing interval, evaluate
//   t1 (z_j t1^2/h_j + c_j) + t2 (z_{j+1} t2^2/h_j + d_j).
// The caller normalises rho to the valence charge (pseudoDiag.py:294-296).
#include "pc_common.h"

long long pc_internal_grid_npts(void* grid);

namespace {

struct GridView {
  int nx, ny, nz;
  const int* lo;
  const int* hi;
  const int* off;
};

}  // namespace

// defined in pc_grid.hip
int pc_internal_grid_view(void* grid, int* nx, int* ny, int* nz, const int** lo, const int** hi, const int** off);

namespace {

__device__ __forceinline__ double spline_eval(const double* __restrict__ xi, const double* __restrict__ z,
                                              const double* __restrict__ c, const double* __restrict__ d, int n,
                                              double x) {
  x = fmin(fmax(x, xi[0]), xi[n - 1]);
  // upper_bound: first index with xi > x
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xi[mid] > x) hi = mid; else lo = mid + 1;
  }
  int j = lo - 1;
  j = j < 0 ? 0 : (j > n - 2 ? n - 2 : j);
  const double t1 = xi[j + 1] - x;
  const double t2 = x - xi[j];
  const double hj = xi[j + 1] - xi[j];
  return t1 * (z[j] * t1 * t1 / hj + c[j]) + t2 * (z[j + 1] * t2 * t2 / hj + d[j]);
}

struct SadArgs {
  GridView g;
  double h, rad, h3;
  int natoms;
  const double* pos;      // [3*natoms]
  const int* species;     // [natoms]
  const int* tab_off;     // [3*nspecies]: charge, pot_S, hartree
  const int* tab_len;     // [3*nspecies]
  const double* xi;
  const double* z;
  const double* c;
  const double* d;
};

__global__ __launch_bounds__(256) void sad_kernel(SadArgs a, double* __restrict__ pot, double* __restrict__ rho,
                                                  double* __restrict__ hpot0) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= a.g.ny * a.g.nz) return;
  const int j = row % a.g.ny, k = row / a.g.ny;
  const int lo = a.g.lo[row], hi = a.g.hi[row], off = a.g.off[row];
  const double y = j * a.h - a.rad;
  const double zc = k * a.h - a.rad;
  const double fac = a.h3 / (4.0 * M_PI);
  for (int i = lo + lane; i < hi; i += 64) {
    const double x = i * a.h - a.rad;
    double sp = 0.0, sr = 0.0, sh = 0.0;
    for (int at = 0; at < a.natoms; ++at) {
      const double dz = zc - a.pos[3 * at + 2];
      const double dy = y - a.pos[3 * at + 1];
      const double dx = x - a.pos[3 * at];
      const double r = sqrt(dz * dz + dy * dy + dx * dx);
      const int s = a.species[at];
      const int oc = a.tab_off[3 * s], nc = a.tab_len[3 * s];
      const int op = a.tab_off[3 * s + 1], np = a.tab_len[3 * s + 1];
      const int oh = a.tab_off[3 * s + 2], nh = a.tab_len[3 * s + 2];
      sp += spline_eval(a.xi + op, a.z + op, a.c + op, a.d + op, np, r);
      const double ch = spline_eval(a.xi + oc, a.z + oc, a.c + oc, a.d + oc, nc, r);
      sr += fmax(0.0, ch * fac);
      sh += spline_eval(a.xi + oh, a.z + oh, a.c + oh, a.d + oh, nh, r);
    }
    const long long p = off + (i - lo);
    pot[p] = sp;
    rho[p] = sr;
    hpot0[p] = sh;
  }
}

// Batched spline evaluation at arbitrary radii: out[i] = S_tab(r[i]).
__global__ __launch_bounds__(256) void spline_batch_kernel(const double* __restrict__ xi, const double* __restrict__ z,
                                                           const double* __restrict__ c, const double* __restrict__ d,
                                                           int n, const double* __restrict__ r, long long m,
                                                           double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < m) out[i] = spline_eval(xi, z, c, d, n, r[i]);
}

}  // namespace

extern "C" {

// pos/species/tab_off/tab_len/xi/z/c/d are device arrays.
int pc_sad_eval(void* grid, double h, double rad, int natoms, const double* pos, const int* species,
                const int* tab_off, const int* tab_len, const double* xi, const double* z, const double* c,
                const double* d, double* pot, double* rho, double* hpot0, void* stream) {
  PC_CHECK_ARG(grid && pos && species && tab_off && tab_len && xi && z && c && d && pot && rho && hpot0,
               "null pointer");
  PC_CHECK_ARG(natoms >= 0, "natoms < 0");
  SadArgs a;
  pc_internal_grid_view(grid, &a.g.nx, &a.g.ny, &a.g.nz, &a.g.lo, &a.g.hi, &a.g.off);
  a.h = h; a.rad = rad; a.h3 = h * h * h;
  a.natoms = natoms; a.pos = pos; a.species = species; a.tab_off = tab_off; a.tab_len = tab_len;
  a.xi = xi; a.z = z; a.c = c; a.d = d;
  const int nrow = a.g.ny * a.g.nz;
  hipLaunchKernelGGL(sad_kernel, dim3((nrow + 3) / 4), dim3(256), 0, (hipStream_t)stream, a, pot, rho, hpot0);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_spline_eval(const double* xi, const double* z, const double* c, const double* d, int n, const double* r,
                   long long m, double* out, void* stream) {
  PC_CHECK_ARG(xi && z && c && d && r && out && n >= 2 && m >= 0, "bad argument");
  if (m == 0) return pc::PC_OK;
  hipLaunchKernelGGL(spline_batch_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, (hipStream_t)stream, xi, z,
                     c, d, n, r, m, out);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

}  // extern "C"
