// Kleinman-Bylander p-channel projectors in factored form.
//
// The reference stores V_nl = sum_a sum_{m=x,y,z} u_am u_am^T / (Zvalue/h^3) as
// dense outer products in a sparse matrix (src/V_ion/pseudoNL_original.py:
// 187-206).  Here each atom keeps its support points and the three u values,
// and V_nl*Y is applied as
//   gather : C[a, v, m] = sum_{p in supp(a)} u_am(p) * Y[p, v]          (FP64)
//   scatter: Out[p, v] += alpha * sum_{(a,m) covering p} scale_a u_am(p) C[a, v, m]
// C is atom-major, C[(a*nvec + v)*3 + m], so the scatter's 8 vectors of one
// entry read 24 consecutive doubles (two cache lines, not eight).
// The scatter runs over unique support points (no atomics, deterministic).
#include "pc_common.h"
#include <algorithm>
#include <numeric>
#include <vector>

namespace {

struct ProjDev {
  int natoms;
  int nent;
  int nupts;
  int* atom_ptr;   // [natoms+1]
  int* ent_pt;     // [nent]  point index of entry
  double* ent_u;   // [3*nent]
  int* ent_atom;   // [nent]
  double* scale;   // [natoms]
  int* upt;        // [nupts] unique point index
  int* upt_ptr;    // [nupts+1] into upt_ent
  int* upt_ent;    // [nent] entry ids grouped by point
  int* se_atom;    // [nent] atom of the t-th entry in point order
  double* se_u;    // [3*nent] scale[atom] * u of the t-th entry in point order
};

// C[a][v][m] = sum_e u_m[e] * Y[v][pt[e]] (FP64 accumulation).  Workgroup =
// one atom x 16 vectors; each wavefront walks the atom's entries once and
// applies them to 4 vectors, so the entry tables are read once per 4 vectors.
constexpr int kGV = 4;   // vectors per wavefront in the gather
template <typename T>
__global__ __launch_bounds__(256) void proj_gather(ProjDev P, const T* __restrict__ Y, long long ld, int nvec,
                                                   double* __restrict__ C) {
  const int a = blockIdx.x;
  const int v0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * kGV;
  const int lane = threadIdx.x & 63;
  if (v0 >= nvec) return;
  double s[kGV][3];
#pragma unroll
  for (int u = 0; u < kGV; ++u) s[u][0] = s[u][1] = s[u][2] = 0.0;
  const int e0 = P.atom_ptr[a], e1 = P.atom_ptr[a + 1];
  for (int e = e0 + lane; e < e1; e += 64) {
    const long long pt = P.ent_pt[e];
    const double u0 = P.ent_u[3 * e], u1 = P.ent_u[3 * e + 1], u2 = P.ent_u[3 * e + 2];
#pragma unroll
    for (int u = 0; u < kGV; ++u) {
      const int v = min(v0 + u, nvec - 1);
      const double y = (double)Y[(long long)v * ld + pt];
      s[u][0] += u0 * y;
      s[u][1] += u1 * y;
      s[u][2] += u2 * y;
    }
  }
#pragma unroll
  for (int u = 0; u < kGV; ++u)
#pragma unroll
    for (int m = 0; m < 3; ++m)
      for (int o = 32; o > 0; o >>= 1) s[u][m] += __shfl_down(s[u][m], o, 64);
  if (lane == 0) {
#pragma unroll
    for (int u = 0; u < kGV; ++u) {
      if (v0 + u >= nvec) break;
      double* cv = C + ((long long)a * nvec + (v0 + u)) * 3;
      cv[0] = s[u][0]; cv[1] = s[u][1]; cv[2] = s[u][2];
    }
  }
}

// Out[v][upt[q]] += alpha * sum over the entries e at point q of
// scale[atom] * (u[e] . C[atom][v]).  One thread per unique point and 8
// vectors: the point's entry list (atom and scaled u, in point order, so one
// dependent load instead of three) is read once for the 8.
constexpr int kSV = 8;   // vectors per thread in the scatter
template <typename T>
__global__ __launch_bounds__(256) void proj_scatter(ProjDev P, const double* __restrict__ C, T* Out, long long ld,
                                                    int nvec, double alpha) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int v0 = blockIdx.y * kSV;
  if (q >= P.nupts) return;
  const int t0 = P.upt_ptr[q], t1 = P.upt_ptr[q + 1];
  const long long pt = P.upt[q];
  // the read half of the read-modify-write goes out with the entry loads
  T ov[kSV];
#pragma unroll
  for (int u = 0; u < kSV; ++u) ov[u] = Out[(long long)min(v0 + u, nvec - 1) * ld + pt];
  double acc[kSV];
#pragma unroll
  for (int u = 0; u < kSV; ++u) acc[u] = 0.0;
  for (int t = t0; t < t1; ++t) {
    const int a = P.se_atom[t];
    const double u0 = P.se_u[3 * t], u1 = P.se_u[3 * t + 1], u2 = P.se_u[3 * t + 2];
#pragma unroll
    for (int u = 0; u < kSV; ++u) {
      const int v = min(v0 + u, nvec - 1);
      const double* c = C + ((long long)a * nvec + v) * 3;
      acc[u] += u0 * c[0] + u1 * c[1] + u2 * c[2];
    }
  }
#pragma unroll
  for (int u = 0; u < kSV; ++u) {
    if (v0 + u >= nvec) break;
    Out[(long long)(v0 + u) * ld + pt] = (T)((double)ov[u] + alpha * acc[u]);
  }
}

template <typename T>
int gather_impl(void* proj, const T* Y, long long ld, int nvec, double* C, void* stream) {
  const ProjDev* P = static_cast<const ProjDev*>(proj);
  PC_CHECK_ARG(P && Y && C, "null pointer");
  if (P->natoms == 0 || nvec == 0) return pc::PC_OK;
  dim3 grid(P->natoms, (nvec + 4 * kGV - 1) / (4 * kGV));
  hipLaunchKernelGGL(proj_gather<T>, grid, dim3(256), 0, static_cast<hipStream_t>(stream), *P, Y, ld, nvec, C);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

template <typename T>
int scatter_impl(void* proj, const double* C, T* Out, long long ld, int nvec, double alpha, void* stream) {
  const ProjDev* P = static_cast<const ProjDev*>(proj);
  PC_CHECK_ARG(P && C && Out, "null pointer");
  if (P->nupts == 0 || nvec == 0) return pc::PC_OK;
  dim3 grid((P->nupts + 255) / 256, (nvec + kSV - 1) / kSV);
  hipLaunchKernelGGL(proj_scatter<T>, grid, dim3(256), 0, static_cast<hipStream_t>(stream), *P, C, Out, ld, nvec,
                     alpha);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

}  // namespace

extern "C" {

// atom_ptr[natoms+1] groups entries by atom; ent_pt[nent] are point indices in
// [0, npts); ent_u[3*nent] the (u_x,u_y,u_z) values; scale[natoms] = h^3/Zvalue.
int pc_proj_create(int natoms, const int* atom_ptr, const int* ent_pt, const double* ent_u, const double* scale,
                   long long npts, void** out_proj) {
  PC_CHECK_ARG(out_proj && atom_ptr && scale, "null pointer");
  PC_CHECK_ARG(natoms >= 0, "natoms < 0");
  const int nent = atom_ptr[natoms];
  PC_CHECK_ARG(nent >= 0 && (nent == 0 || (ent_pt && ent_u)), "bad entries");
  std::vector<int> ent_atom(nent);
  for (int a = 0; a < natoms; ++a) {
    PC_CHECK_ARG(atom_ptr[a] <= atom_ptr[a + 1], "atom_ptr not monotone");
    for (int e = atom_ptr[a]; e < atom_ptr[a + 1]; ++e) ent_atom[e] = a;
  }
  for (int e = 0; e < nent; ++e) PC_CHECK_ARG(ent_pt[e] >= 0 && ent_pt[e] < npts, "entry point out of range");
  std::vector<int> order(nent);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ent_pt[x] < ent_pt[y]; });
  std::vector<int> upt, upt_ptr;
  for (int t = 0; t < nent; ++t) {
    if (t == 0 || ent_pt[order[t]] != ent_pt[order[t - 1]]) {
      upt.push_back(ent_pt[order[t]]);
      upt_ptr.push_back(t);
    }
  }
  upt_ptr.push_back(nent);
  ProjDev* P = new ProjDev();
  P->natoms = natoms; P->nent = nent; P->nupts = (int)upt.size();
  hipError_t e = hipSuccess;
  auto alloc_copy = [&](void** dst, const void* src, size_t bytes) {
    if (e != hipSuccess) return;
    e = hipMalloc(dst, std::max<size_t>(bytes, 8));
    if (e == hipSuccess && bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  alloc_copy((void**)&P->atom_ptr, atom_ptr, (natoms + 1) * sizeof(int));
  alloc_copy((void**)&P->ent_pt, ent_pt, nent * sizeof(int));
  alloc_copy((void**)&P->ent_u, ent_u, 3 * (size_t)nent * sizeof(double));
  alloc_copy((void**)&P->ent_atom, ent_atom.data(), nent * sizeof(int));
  alloc_copy((void**)&P->scale, scale, natoms * sizeof(double));
  alloc_copy((void**)&P->upt, upt.data(), upt.size() * sizeof(int));
  alloc_copy((void**)&P->upt_ptr, upt_ptr.data(), upt_ptr.size() * sizeof(int));
  alloc_copy((void**)&P->upt_ent, order.data(), nent * sizeof(int));
  std::vector<int> se_atom(nent);
  std::vector<double> se_u(3 * (size_t)nent);
  for (int t = 0; t < nent; ++t) {
    const int e = order[t], a = ent_atom[e];
    se_atom[t] = a;
    for (int m = 0; m < 3; ++m) se_u[3 * (size_t)t + m] = scale[a] * ent_u[3 * (size_t)e + m];
  }
  alloc_copy((void**)&P->se_atom, se_atom.data(), nent * sizeof(int));
  alloc_copy((void**)&P->se_u, se_u.data(), 3 * (size_t)nent * sizeof(double));
  if (e != hipSuccess) {
    pc::set_error(std::string("pc_proj_create: ") + hipGetErrorString(e));
    return pc::PC_ERR_HIP;
  }
  *out_proj = P;
  return pc::PC_OK;
}

int pc_proj_destroy(void* proj) {
  ProjDev* P = static_cast<ProjDev*>(proj);
  if (!P) return pc::PC_OK;
  (void)hipFree(P->atom_ptr); (void)hipFree(P->ent_pt); (void)hipFree(P->ent_u); (void)hipFree(P->ent_atom);
  (void)hipFree(P->scale); (void)hipFree(P->upt); (void)hipFree(P->upt_ptr); (void)hipFree(P->upt_ent);
  (void)hipFree(P->se_atom); (void)hipFree(P->se_u);
  delete P;
  return pc::PC_OK;
}

int pc_proj_ncols(void* proj, int* out_ncols) {
  const ProjDev* P = static_cast<const ProjDev*>(proj);
  PC_CHECK_ARG(P && out_ncols, "null pointer");
  *out_ncols = 3 * P->natoms;
  return pc::PC_OK;
}

int pc_proj_gather_f64(void* proj, const double* Y, long long ld, int nvec, double* C, void* stream) {
  return gather_impl<double>(proj, Y, ld, nvec, C, stream);
}
int pc_proj_gather_f32(void* proj, const float* Y, long long ld, int nvec, double* C, void* stream) {
  return gather_impl<float>(proj, Y, ld, nvec, C, stream);
}
int pc_proj_scatter_f64(void* proj, const double* C, double* Out, long long ld, int nvec, double alpha,
                        void* stream) {
  return scatter_impl<double>(proj, C, Out, ld, nvec, alpha, stream);
}
int pc_proj_scatter_f32(void* proj, const double* C, float* Out, long long ld, int nvec, double alpha,
                        void* stream) {
  return scatter_impl<float>(proj, C, Out, ld, nvec, alpha, stream);
}

}  // extern "C"
