// Hartree solve: conjugate gradients on A x = rhs, A = -lap/h^2 on the grid,
// FP64, no preconditioner (src/Eigensolvers/pcg.py:7-85; default cg_prec=0,
// src/rsdft_models.py:31).  Stops when r.r <= tol^2 * r0.r0 or after maxit
// iterations.  With alias_first=1 the second search direction is r_1 + beta r_1,
// which is what the reference's CPU loop computes because its p and r are the
// same array in the first iteration (pcg.py:45-77).
//
// Two launches per iteration, each a fixed grid of NB workgroups striding
// over the domain points:
//   cg_apply : p_new = r + beta p_old (formed on the fly at the centre and at
//              every stencil neighbour), writes p_new and A p_new and one
//              partial of p_new . A p_new per workgroup;
//   cg_update: x += alpha p_new, r -= alpha A p_new, one partial of r.r per
//              workgroup.
// Every workgroup sums the NB partials it needs (r.r of the last two
// iterations, p.Ap of this one) itself, in the same fixed order, so all
// workgroups hold bit-identical alpha, beta and convergence decisions without
// a separate reduction launch and without device-scope fences; workgroup 0
// records the scalars for the host.  When the iteration has converged (or hit
// maxit) both kernels return at once, so the host only reads the state back
// every `check_every` iterations.
#include "pc_common.h"
#include "pc_zmarch.h"
#include <algorithm>
#include <mutex>
#include <string>

long long pc_internal_grid_npts(void* grid);

namespace {

constexpr int kBlk = 512;       // threads per workgroup
constexpr int kMaxNB = 512;     // workgroups of the point-wise launches (= r.r partials)
constexpr int kMaxWG = pcz::kMaxWG;   // workgroups of the z-march step (= p.Ap partials)
// scalars for the host (written by workgroup 0); r.r of iteration t in slot S_RO0 + (t & 1)
enum { S_RO0 = 0, S_RO1, S_TOL1, S_DONE, S_ITS, S_RES, S_COUNT = 8 };
enum Mode { M_FIRST = 1, M_ALIAS = 2, M_NORMAL = 3 };

struct PointMap {
  const int2* meta;   // per row {off - lo, lo | hi << 16}
  const int* prow;    // per point: its row j + ny*k
  long long n;
  long long chunk;    // points per workgroup (multiple of kBlk)
  int nb, nb8;        // chunks (workgroups with work) and ceil(nb / 8)
  int ny, nz;
  double w[8];        // weights of -lap/h^2; w[0] is the 3D centre weight
};

// Chunk of this workgroup: consecutive chunks go to the same XCD (workgroup
// ids are dealt round-robin over the 8 XCDs), so the z-neighbour planes a
// chunk reads are mostly in that XCD's L2.
__device__ __forceinline__ int chunk_of(int nb8) { return (blockIdx.x & 7) * nb8 + (blockIdx.x >> 3); }

// Sum of part[0..nb) in a fixed order; the same in every workgroup.
__device__ __forceinline__ double sum_parts(const double* part, int nb, double* red) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
  double t = pc_block_sum(s, red);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = t;
  __syncthreads();
  return bc;
}

template <int MODE>
__device__ __forceinline__ double qv(const double* __restrict__ a, const double* __restrict__ b, double beta,
                                     long long idx) {
  if (MODE == M_FIRST) return a[idx];
  if (MODE == M_ALIAS) {
    const double r = a[idx];
    return r + beta * r;
  }
  return a[idx] + beta * b[idx];
}

// sum_d w_d (6 neighbours of q at distance d) + w_0 q(p), q formed on the fly
template <int R, int MODE>
__device__ __forceinline__ double apply_point(const PointMap& g, const double* __restrict__ a,
                                              const double* __restrict__ b, double beta, long long p, double& qc) {
  const int row = g.prow[p];
  const int2 m = g.meta[row];
  const int i = (int)(p - m.x);
  const int lo = m.y & 0xffff, hi = m.y >> 16;
  const int k = row / g.ny;
  const int j = row - k * g.ny;
  qc = qv<MODE>(a, b, beta, p);
  double acc = g.w[0] * qc;
  auto qrow = [&](int2 mm) -> double {
    const int l2 = mm.y & 0xffff, h2 = mm.y >> 16;
    return (i >= l2 && i < h2) ? qv<MODE>(a, b, beta, (long long)mm.x + i) : 0.0;
  };
#pragma unroll
  for (int d = 1; d <= R; ++d) {
    double nb = ((i - d >= lo) ? qv<MODE>(a, b, beta, p - d) : 0.0) +
                ((i + d < hi) ? qv<MODE>(a, b, beta, p + d) : 0.0);
    if (j - d >= 0) nb += qrow(g.meta[row - d]);
    if (j + d < g.ny) nb += qrow(g.meta[row + d]);
    if (k - d >= 0) nb += qrow(g.meta[row - d * g.ny]);
    if (k + d < g.nz) nb += qrow(g.meta[row + d * g.ny]);
    acc += g.w[d] * nb;
  }
  return acc;
}

// r = rhs - A x; partial r.r -> part_rr
template <int R>
__global__ __launch_bounds__(kBlk) void cg_init(PointMap g, const double* __restrict__ x,
                                                const double* __restrict__ rhs, double* __restrict__ r,
                                                double* __restrict__ part_rr) {
  __shared__ double red[kBlk / 64];
  const int c = chunk_of(g.nb8);
  if (c >= g.nb) return;
  const long long p0 = (long long)c * g.chunk;
  const long long p1 = min(p0 + g.chunk, g.n);
  double s = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += kBlk) {
    double qc;
    const double acc = apply_point<R, M_FIRST>(g, x, nullptr, 0.0, p, qc);
    const double rv = rhs[p] - acc;
    r[p] = rv;
    s += rv * rv;
  }
  const double t = pc_block_sum(s, red);
  if (threadIdx.x == 0) part_rr[c] = t;
}

// ---- iteration step t (t iterations completed so far) ------------------------
// rr partials of r_t.r_t -> convergence decision (t >= maxit or r.r <= tol^2
// r0.r0) and beta = r_t.r_t / r_{t-1}.r_{t-1} in every workgroup; workgroup 0
// records the scalars.  Then p_new = r + beta p_old (formed on the fly at every
// stencil point of the z-march), A p_new and the partials of p_new . A p_new,
// one per workgroup.
using pcz::ZMap;

template <int R, int SX, int MODE>
__global__ __launch_bounds__(256) void cg_apply_z(ZMap g, int t, int maxit, double tol, int nb_rr,
                                                  const double* __restrict__ r, const double* __restrict__ p_old,
                                                  double* __restrict__ p_new, double* __restrict__ ap,
                                                  const double* __restrict__ part_rr, double* __restrict__ part_pq,
                                                  double* sc) {
  __shared__ double tile[pcz::Tile<R, SX>::SIZE];
  __shared__ double red[pcz::kZW];
  if (t > 0 && sc[S_DONE] != 0.0) return;
  const int L = pcz::block_of(g);
  if (L >= g.nwg) return;
  const double rr = sum_parts(part_rr, nb_rr, red);
  const double tol1 = t == 0 ? tol * tol * rr : sc[S_TOL1];
  const bool done = t >= maxit || !(rr > tol1);
  const double beta = t > 0 ? rr / sc[S_RO0 + ((t - 1) & 1)] : 0.0;
  if (L == 0 && threadIdx.x == 0) {
    sc[S_RO0 + (t & 1)] = rr;
    if (t == 0) sc[S_TOL1] = tol1;
    sc[S_ITS] = (double)t;
    sc[S_RES] = rr;
    sc[S_DONE] = done ? 1.0 : 0.0;
  }
  if (done) return;
  double sacc = 0.0;
  pcz::zmarch<R, SX>(
      g, L, tile, [&](long long idx, int, int, int) { return qv<MODE>(r, p_old, beta, idx); },
      [&](long long idx, int, int, int, double qc, double acc) {
        p_new[idx] = qc;
        ap[idx] = acc;
        sacc += qc * acc;
      });
  const double bs = pc_block_sum(sacc, red);
  if (threadIdx.x == 0) part_pq[L] = bs;
}

// alpha = r_t.r_t / p.Ap; x += alpha p; r -= alpha Ap; partials of r_{t+1}.r_{t+1}
__global__ __launch_bounds__(kBlk) void cg_update(long long n, long long chunk, int nb, int nb8, int npq, int t,
                                                  const double* sc,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ ap,
                                                  const double* __restrict__ part_pq,
                                                  double* __restrict__ part_rr_next) {
  __shared__ double red[kBlk / 64];
  if (sc[S_DONE] != 0.0) return;
  const int c = chunk_of(nb8);
  if (c >= nb) return;
  const double pq = sum_parts(part_pq, npq, red);
  const double alpha = sc[S_RO0 + (t & 1)] / pq;
  const long long p0 = (long long)c * chunk;
  const long long p1 = min(p0 + chunk, n);
  double s = 0.0;
  for (long long i = p0 + threadIdx.x; i < p1; i += kBlk) {
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * ap[i];
    r[i] = rv;
    s += rv * rv;
  }
  const double tt = pc_block_sum(s, red);
  if (threadIdx.x == 0) part_rr_next[c] = tt;
}

struct Work {
  double* r;
  double* p[2];
  double* ap;
  double* prr[2];
  double* ppq;
  double* sc;
};

size_t work_bytes_for(long long n) {
  return (size_t)(4 * n + 2 * kMaxNB + kMaxWG + S_COUNT) * sizeof(double);
}

template <int R>
int launch_init(int nb, hipStream_t st, const PointMap& g, const double* x, const double* rhs, const Work& w) {
  hipLaunchKernelGGL((cg_init<R>), dim3(nb), dim3(kBlk), 0, st, g, x, rhs, w.r, w.prr[0]);
  return hipGetLastError() == hipSuccess ? pc::PC_OK : pc::PC_ERR_HIP;
}

template <int R, int SX>
int launch_apply_z(int mode, hipStream_t st, const ZMap& z, int t, int maxit, double tol, int nb_rr, const Work& w) {
  const double* p_old = w.p[t & 1];
  double* p_new = w.p[(t + 1) & 1];
  const double* prr = w.prr[t & 1];
  const dim3 grid(8 * z.nwg8);
  switch (mode) {
    case M_FIRST:
      hipLaunchKernelGGL((cg_apply_z<R, SX, M_FIRST>), grid, dim3(256), 0, st, z, t, maxit, tol, nb_rr, w.r, p_old,
                         p_new, w.ap, prr, w.ppq, w.sc);
      break;
    case M_ALIAS:
      hipLaunchKernelGGL((cg_apply_z<R, SX, M_ALIAS>), grid, dim3(256), 0, st, z, t, maxit, tol, nb_rr, w.r, p_old,
                         p_new, w.ap, prr, w.ppq, w.sc);
      break;
    default:
      hipLaunchKernelGGL((cg_apply_z<R, SX, M_NORMAL>), grid, dim3(256), 0, st, z, t, maxit, tol, nb_rr, w.r, p_old,
                         p_new, w.ap, prr, w.ppq, w.sc);
      break;
  }
  return hipGetLastError() == hipSuccess ? pc::PC_OK : pc::PC_ERR_HIP;
}

struct Launchers {
  int (*init)(int, hipStream_t, const PointMap&, const double*, const double*, const Work&);
  int (*apply)(int, hipStream_t, const ZMap&, int, int, double, int, const Work&);
};

template <int R>
Launchers launchers(int sx) {
  if (sx == 1) return {launch_init<R>, launch_apply_z<R, 1>};
  if (sx == 2) return {launch_init<R>, launch_apply_z<R, 2>};
  return {launch_init<R>, launch_apply_z<R, 4>};
}

// Shared body of the two solve entry points: `work` holds work_bytes_for(n)
// bytes (8-byte aligned, no initial contents needed: every word is written
// before it is read), `host` S_COUNT doubles of host memory for the state
// read back every check_every iterations.
int pcg_solve_impl(void* grid, double inv_h2, const double* rhs, double* x, int maxit, double tol, int alias_first,
                   int check_every, double* work, double* host, int* its_out, double* res_out, hipStream_t st) {
  const long long n = pc_internal_grid_npts(grid);
  if (check_every < 1) check_every = 8;

  PointMap g;
  int radius = 0;
  double w8[8];
  pc_internal_grid_points(grid, &g.ny, &g.nz, &radius, &g.meta, &g.prow, w8);
  g.n = n;
  for (int d = 0; d < 8; ++d) g.w[d] = w8[d] * inv_h2;
  g.w[0] *= 3.0;
  ZMap zm;
  int radius_z = 0;
  const int zbad = pcz::zmap_init(grid, g.w, zm, &radius_z);
  const int sx = zm.sx;
  PC_CHECK_ARG(!zbad, "nx > 256");
  Launchers L;
  switch (radius) {
    case 1: L = launchers<1>(sx); break;
    case 2: L = launchers<2>(sx); break;
    case 3: L = launchers<3>(sx); break;
    case 4: L = launchers<4>(sx); break;
    case 6: L = launchers<6>(sx); break;
    default: PC_CHECK_ARG(false, "unsupported stencil radius");
  }
  // workgroups own contiguous chunks (multiples of kBlk points)
  const long long blocks_of_pts = (n + kBlk - 1) / kBlk;
  const int nb = (int)std::min<long long>(kMaxNB, blocks_of_pts);
  g.chunk = ((blocks_of_pts + nb - 1) / nb) * kBlk;
  const int nb_used = (int)((n + g.chunk - 1) / g.chunk);
  g.nb = nb_used;
  g.nb8 = (nb_used + 7) / 8;
  const int nb_launch = 8 * g.nb8;

  Work w;
  w.r = work;
  w.p[0] = w.r + n;
  w.p[1] = w.p[0] + n;
  w.ap = w.p[1] + n;
  w.prr[0] = w.ap + n;
  w.prr[1] = w.prr[0] + kMaxNB;
  w.ppq = w.prr[1] + kMaxNB;
  w.sc = w.ppq + kMaxWG;

  if (L.init(nb_launch, st, g, x, rhs, w)) {
    pc::set_error("pc_pcg_solve: launch failed");
    return pc::PC_ERR_HIP;
  }
  int status = pc::PC_OK;
  // iteration t: apply(t) decides convergence from r_t; update(t) makes r_{t+1}.
  // apply(maxit) only records the final state.
  for (int t0 = 0; t0 <= maxit; t0 += check_every) {
    for (int t = t0; t < t0 + check_every && t <= maxit; ++t) {
      const int mode = t == 0 ? M_FIRST : ((t == 1 && alias_first) ? M_ALIAS : M_NORMAL);
      status = L.apply(mode, st, zm, t, maxit, tol, nb_used, w);
      if (!status && t < maxit) {
        hipLaunchKernelGGL(cg_update, dim3(nb_launch), dim3(kBlk), 0, st, n, g.chunk, g.nb, g.nb8, zm.nwg, t, w.sc, x,
                           w.r,
                           w.p[(t + 1) & 1], w.ap, w.ppq, w.prr[(t + 1) & 1]);
        if (hipGetLastError() != hipSuccess) status = pc::PC_ERR_HIP;
      }
      if (status) {
        pc::set_error("pc_pcg_solve: kernel launch failed");
        break;
      }
    }
    if (status) break;
    if (hipMemcpyAsync(host, w.sc, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      pc::set_error("pc_pcg_solve: readback failed");
      status = pc::PC_ERR_HIP;
      break;
    }
    if (host[S_DONE] != 0.0) break;
  }
  if (!status) {
    if (host[S_DONE] == 0.0) {
      pc::set_error("pc_pcg_solve: internal error (no final state)");
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

// ---- the solve without allocations ----------------------------------------------
// A file-scope device array replaces the caller workspace, sized for grids of at
// most kNoAllocMaxPts points (128^3).  Every device holds its own copy at its own
// address; host code never uses the array's name as a device address (it is a
// host-side stand-in), only the pointer hipGetSymbolAddress returned for the
// current device.
constexpr long long kNoAllocMaxPts = 1LL << 21;
constexpr size_t kNoAllocDoubles = (size_t)(4 * kNoAllocMaxPts + 2 * kMaxNB + kMaxWG + S_COUNT);

}  // namespace

// 64 MiB + 12 KiB of device memory per device the library runs on, held until
// the process ends.
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_ptr(double** out) {
  int dev = -1;
  PC_HIP(hipGetDevice(&dev));
  PC_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    PC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return pc::PC_OK;
}

}  // namespace

extern "C" {

int pc_pcg_workspace(void* grid, size_t* out_bytes) {
  PC_CHECK_ARG(grid && out_bytes, "null pointer");
  *out_bytes = work_bytes_for(pc_internal_grid_npts(grid));
  return pc::PC_OK;
}

// x (in: warm start, out: solution) and rhs are device FP64 arrays of npts.
// inv_h2 = 1/h^2.  its_out receives the iteration count (host); res_out (host,
// optional, 2 doubles) receives the final r.r and tol^2 * r0.r0.
int pc_pcg_solve(void* grid, double inv_h2, const double* rhs, double* x, int maxit, double tol, int alias_first,
                 int check_every, void* work, size_t work_bytes, int* its_out, double* res_out, void* stream) {
  PC_CHECK_ARG(grid && rhs && x && work && its_out, "null pointer");
  const long long n = pc_internal_grid_npts(grid);
  PC_CHECK_ARG(n > 0, "empty domain");
  PC_CHECK_ARG(work_bytes >= work_bytes_for(n), "workspace too small");
  PC_CHECK_ARG(((uintptr_t)work & 7) == 0, "workspace not 8-byte aligned");
  PC_CHECK_ARG(maxit >= 0, "maxit < 0");
  // pinned readback buffer, allocated once per host thread and kept
  static thread_local double* host = nullptr;
  if (!host) PC_HIP(hipHostMalloc(&host, S_COUNT * sizeof(double), hipHostMallocDefault));
  return pcg_solve_impl(grid, inv_h2, rhs, x, maxit, tol, alias_first, check_every, static_cast<double*>(work), host,
                        its_out, res_out, static_cast<hipStream_t>(stream));
}

// The same solve with no allocation of any kind: the workspace is the library's
// device array `scratch` (grids of at most 2^21 points) and the state is read
// back into a stack buffer.  Calls on one device share that array, so calls
// that overlap on one device (other host threads, other streams) race.
int pc_pcg_solve_noalloc(void* grid, double inv_h2, const double* rhs, double* x, int maxit, double tol,
                         int alias_first, int check_every, int* its_out, double* res_out, void* stream) {
  PC_CHECK_ARG(grid && rhs && x && its_out, "null pointer");
  const long long n = pc_internal_grid_npts(grid);
  PC_CHECK_ARG(n > 0, "empty domain");
  PC_CHECK_ARG(maxit >= 0, "maxit < 0");
  if (n > kNoAllocMaxPts) {
    pc::set_error("pc_pcg_solve_noalloc: grid of " + std::to_string(n) + " points exceeds the maximum of " +
                  std::to_string(kNoAllocMaxPts) + " points of the library's fixed workspace");
    return pc::PC_ERR_LIMIT;
  }
  double* work = nullptr;
  if (int e = scratch_ptr(&work)) return e;
  double host[S_COUNT];
  return pcg_solve_impl(grid, inv_h2, rhs, x, maxit, tol, alias_first, check_every, work, host, its_out, res_out,
                        static_cast<hipStream_t>(stream));
}

int pc_pcg_noalloc_max_points(long long* out) {
  PC_CHECK_ARG(out, "null pointer");
  *out = kNoAllocMaxPts;
  return pc::PC_OK;
}

}  // extern "C"
