// Rayleigh-Ritz block GEMMs on MFMA (gfx950).
//
//   gram  : S = X^T Y        X: n x kx, Y: n x ky (vector-major, ld), S FP64 kx x ky
//   rotate: Z = X Q          X: n x kx, Q: FP64 kx x ky (column-major), Z: n x ky
//
// The reference forms G = W^T (H W) with a Python double loop of vdot
// (src/Eigensolvers/chefsi1.py:51-53) or cuBLAS (chefsi1_gpu.py:61) and
// rotates with W @ Q (chefsi1.py:67, chefsi1_gpu.py:66).
//
// MFMA 16x16x4 (f64: v_mfma_f64_16x16x4f64, f32: v_mfma_f32_16x16x4f32) lane
// layout on gfx950 (measured with one-hot operands): A[m][k] from lane m + 16k,
// B[k][n] from lane n + 16k; D[m][n] in lane n + 16*(m/4), register m%4 for
// f32, and in lane n + 16*(m%4), register m/4 for f64.  The K index of an operand is only a label
// that A and B must share, so each lane loads 4 consecutive points (one
// 16-byte/32-byte load) and step s of 4 uses point 4*(lane/16)+s.
//
// gram splits the n points into chunks; each workgroup accumulates one output
// tile over one chunk (FP64 MFMA, or FP32 MFMA per 16 points flushed into FP64
// registers) and writes an FP64 partial; a second kernel sums the partials in a
// fixed order (deterministic).
#include "pc_common.h"
#include <algorithm>

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
template <> struct Acc<double> {
  typedef v4d type;
  static __device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output row of register r in lane group lk (= lane/16)
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
};
template <> struct Acc<float> {
  typedef v4f type;
  static __device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// Load 4 consecutive points [p, p+4) of column col; zero beyond n.
template <typename T, bool VEC>
__device__ __forceinline__ void load4(const T* __restrict__ base, long long ld, int col, int ncol, long long p,
                                      long long n, T out[4]) {
  if (col < ncol) {
    const T* src = base + (long long)col * ld + p;
    if (VEC && p + 4 <= n) {
      if constexpr (sizeof(T) == 4) {
        const float4 t = *reinterpret_cast<const float4*>(src);
        out[0] = t.x; out[1] = t.y; out[2] = t.z; out[3] = t.w;
      } else {
        const double2 t0 = *reinterpret_cast<const double2*>(src);
        const double2 t1 = *reinterpret_cast<const double2*>(src + 2);
        out[0] = t0.x; out[1] = t0.y; out[2] = t1.x; out[3] = t1.y;
      }
      return;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) out[s] = (p + s < n) ? src[s] : T(0);
    return;
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) out[s] = T(0);
}

// ---------------------------------------------------------------- gram
// Tile t -> (ti, tj): row-major over ti_n x tj_n, or the upper triangle
// (ti <= tj) row by row when sym.
__device__ __forceinline__ void tile_coords(int t, int ti_n, int tj_n, int sym, int& ti, int& tj) {
  if (!sym) {
    ti = t / tj_n;
    tj = t - ti * tj_n;
    return;
  }
  int i = 0;
  while (t >= ti_n - i) {
    t -= ti_n - i;
    ++i;
  }
  ti = i;
  tj = i + t;
}

// Workgroup = 4 waves, output tile 64x64 (each wave 32x32 = 2x2 MFMA tiles).
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void gram_kernel(const T* __restrict__ X, long long ldx, const T* __restrict__ Y,
                                                   long long ldy, long long n, int kx, int ky, int ti_n, int tj_n,
                                                   int ntile, int sym, int nchunk, long long chunk,
                                                   double* __restrict__ part) {
  typedef typename Acc<T>::type AT;
  const int b = blockIdx.x;
  const int tile = b % ntile;
  const int ch = b / ntile;
  int ti, tj;
  tile_coords(tile, ti_n, tj_n, sym, ti, tj);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wi = w >> 1, wj = w & 1;
  const int i0 = ti * 64 + wi * 32, j0 = tj * 64 + wj * 32;
  const int lm = lane & 15, lk = lane >> 4;
  // FP32 inputs: the MFMA accumulates 16 points in FP32, then the block is
  // added into FP64 accumulators (products of FP32 values are exact in FP64,
  // so the only FP32 rounding is within one 16-term block).
  AT acc[2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      acc[a][c] = AT{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[a][c][r] = 0.0;
    }
  const long long pbeg = (long long)ch * chunk;
  const long long pend = min(n, pbeg + chunk);
  for (long long p = pbeg; p < pend; p += 16) {
    T a[2][4], bb[2][4];
    const long long pl = p + 4 * lk;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      load4<T, VEC>(X, ldx, i0 + 16 * t + lm, kx, pl, pend, a[t]);
      load4<T, VEC>(Y, ldy, j0 + 16 * t + lm, ky, pl, pend, bb[t]);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = Acc<T>::mfma(a[ta][s], bb[tb][s], acc[ta][tb]);
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64[ta][tb][r] += (double)acc[ta][tb][r];
          acc[ta][tb] = AT{0, 0, 0, 0};
        }
    }
  }
  // partial layout: part[(ch*ntile + tile)*4096 + jj*64 + ii], tile-local 64x64 col-major
  double* out = part + ((long long)ch * ntile + tile) * 4096;
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ii = wi * 32 + 16 * ta + Acc<T>::row(lk, r);
        const int jj = wj * 32 + 16 * tb + lm;
        out[jj * 64 + ii] = (sizeof(T) == 4) ? acc64[ta][tb][r] : (double)acc[ta][tb][r];
      }
}

// FP32 form with the operand tiles staged in LDS: per stage of 32 points the
// workgroup loads its 64 X columns and 64 Y columns once (two float4 per thread
// per operand) into one of two LDS buffers, and each wave reads its 32-column
// halves from there, so every operand value comes from L2 once per workgroup
// instead of once per wave.  The next stage's global loads are in flight while
// the MFMAs run; one barrier per stage.  Same MFMA inputs and the same FP64
// flush every 16 points as gram_kernel<float, true>, so the results match it.
// Requires 16-byte aligned columns (ld % 4 == 0) and chunk % 32 == 0.
__global__ __launch_bounds__(256) void gram_lds_f32(const float* __restrict__ X, long long ldx,
                                                    const float* __restrict__ Y, long long ldy, long long n, int kx,
                                                    int ky, int ti_n, int tj_n, int ntile, int sym, int nchunk,
                                                    long long chunk, double* __restrict__ part) {
  constexpr int SP = 32;        // points per stage
  constexpr int LS = SP + 4;    // LDS row stride (floats)
  __shared__ float xs[2][64 * LS], ys[2][64 * LS];
  const int b0 = blockIdx.x;
  const int tile = b0 % ntile;
  const int ch = b0 / ntile;
  int ti, tj;
  tile_coords(tile, ti_n, tj_n, sym, ti, tj);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wi = w >> 1, wj = w & 1;
  const int lm = lane & 15, lk = lane >> 4;
  const int ib = ti * 64, jb = tj * 64;
  // loader: thread -> column lc of the tile, 8-point segment ls of the stage
  const int lc = threadIdx.x >> 2, ls = (threadIdx.x & 3) * 8;
  const bool xin = ib + lc < kx, yin = jb + lc < ky;
  const float* __restrict__ xcol = X + (long long)(xin ? ib + lc : 0) * ldx;
  const float* __restrict__ ycol = Y + (long long)(yin ? jb + lc : 0) * ldy;
  const long long pbeg = (long long)ch * chunk;
  const long long pend = min(n, pbeg + chunk);
  float gx[8], gy[8];
  auto gload = [&](long long p) {
    const long long q = p + ls;
    if (q + 8 <= pend) {
      if (xin) {
        const float4 t0 = *reinterpret_cast<const float4*>(xcol + q);
        const float4 t1 = *reinterpret_cast<const float4*>(xcol + q + 4);
        gx[0] = t0.x; gx[1] = t0.y; gx[2] = t0.z; gx[3] = t0.w; gx[4] = t1.x; gx[5] = t1.y; gx[6] = t1.z; gx[7] = t1.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) gx[e] = 0.0f;
      }
      if (yin) {
        const float4 t0 = *reinterpret_cast<const float4*>(ycol + q);
        const float4 t1 = *reinterpret_cast<const float4*>(ycol + q + 4);
        gy[0] = t0.x; gy[1] = t0.y; gy[2] = t0.z; gy[3] = t0.w; gy[4] = t1.x; gy[5] = t1.y; gy[6] = t1.z; gy[7] = t1.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) gy[e] = 0.0f;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        gx[e] = (xin && q + e < pend) ? xcol[q + e] : 0.0f;
        gy[e] = (yin && q + e < pend) ? ycol[q + e] : 0.0f;
      }
    }
  };
  auto sstore = [&](int bf) {
    float* xd = &xs[bf][lc * LS + ls];
    float* yd = &ys[bf][lc * LS + ls];
    *reinterpret_cast<float4*>(xd) = make_float4(gx[0], gx[1], gx[2], gx[3]);
    *reinterpret_cast<float4*>(xd + 4) = make_float4(gx[4], gx[5], gx[6], gx[7]);
    *reinterpret_cast<float4*>(yd) = make_float4(gy[0], gy[1], gy[2], gy[3]);
    *reinterpret_cast<float4*>(yd + 4) = make_float4(gy[4], gy[5], gy[6], gy[7]);
  };
  v4f acc[2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      acc[a][c] = v4f{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[a][c][r] = 0.0;
    }
  if (pbeg < pend) {
    gload(pbeg);
    sstore(0);
  }
  __syncthreads();
  int bf = 0;
  for (long long p = pbeg; p < pend; p += SP) {
    const bool more = p + SP < pend;
    if (more) gload(p + SP);
#pragma unroll
    for (int g = 0; g < SP / 16; ++g) {
      float a[2][4], bb[2][4];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const float4 av = *reinterpret_cast<const float4*>(&xs[bf][(wi * 32 + 16 * t + lm) * LS + 16 * g + 4 * lk]);
        const float4 bv = *reinterpret_cast<const float4*>(&ys[bf][(wj * 32 + 16 * t + lm) * LS + 16 * g + 4 * lk]);
        a[t][0] = av.x; a[t][1] = av.y; a[t][2] = av.z; a[t][3] = av.w;
        bb[t][0] = bv.x; bb[t][1] = bv.y; bb[t][2] = bv.z; bb[t][3] = bv.w;
      }
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int ta = 0; ta < 2; ++ta)
#pragma unroll
          for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = Acc<float>::mfma(a[ta][s2], bb[tb][s2], acc[ta][tb]);
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64[ta][tb][r] += (double)acc[ta][tb][r];
          acc[ta][tb] = v4f{0, 0, 0, 0};
        }
    }
    if (more) sstore(bf ^ 1);
    __syncthreads();
    bf ^= 1;
  }
  double* out = part + ((long long)ch * ntile + tile) * 4096;
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ii = wi * 32 + 16 * ta + Acc<float>::row(lk, r);
        const int jj = wj * 32 + 16 * tb + lm;
        out[jj * 64 + ii] = acc64[ta][tb][r];
      }
}

// 16 workgroups per 64x64 tile, one element per thread; chunks summed in
// order (deterministic)
__global__ void gram_reduce(const double* __restrict__ part, int ntile, int nchunk, int ti_n, int tj_n, int kx,
                            int ky, double* __restrict__ S, long long lds, int sym) {
  const int tile = blockIdx.x >> 4;
  int ti, tj;
  tile_coords(tile, ti_n, tj_n, sym, ti, tj);
  {
    const int e = (blockIdx.x & 15) * 256 + threadIdx.x;
    const int ii = e & 63, jj = e >> 6;
    const int i = ti * 64 + ii, j = tj * 64 + jj;
    if (i >= kx || j >= ky) return;
    double s = 0.0;
    // unrolled so that 16 loads per thread are in flight (the sum keeps its order)
#pragma unroll 16
    for (int c = 0; c < nchunk; ++c) s += part[((long long)c * ntile + tile) * 4096 + e];
    S[(long long)j * lds + i] = s;
    if (sym && ti != tj) S[(long long)i * lds + j] = s;
  }
}

// ---------------------------------------------------------------- rotate
// Z^T (ky x n) = Q^T (ky x kx) * X^T (kx x n).  Workgroup = 4 waves; wave tile
// = 64 output columns x 32 points (4x2 MFMA tiles); workgroup = 64 cols x 128 pts.
template <typename T>
__global__ __launch_bounds__(256) void rotate_kernel(const T* __restrict__ X, long long ldx,
                                                     const T* __restrict__ Q, int ldq, T* __restrict__ Z,
                                                     long long ldz, long long n, int kx, int ky, int npt_tiles,
                                                     int ncol_tiles, int nblocks, int nb8) {
  typedef typename Acc<T>::type AT;
  const int L = (blockIdx.x & 7) * nb8 + (blockIdx.x >> 3);
  if (L >= nblocks) return;
  const int ct = L % ncol_tiles;
  const int pt = L / ncol_tiles;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lm = lane & 15, lk = lane >> 4;
  const int c0 = ct * 64;
  const long long p0 = (long long)pt * 128 + w * 32;
  AT acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = AT{0, 0, 0, 0};
  for (int k = 0; k < kx; k += 4) {
    const int kk = k + lk;
    T qa[4], xb[2];
#pragma unroll
    for (int ta = 0; ta < 4; ++ta) {
      const int col = c0 + 16 * ta + lm;
      qa[ta] = (kk < kx && col < ky) ? Q[(long long)kk * ldq + col] : T(0);
    }
#pragma unroll
    for (int tb = 0; tb < 2; ++tb) {
      const long long p = p0 + 16 * tb + lm;
      xb[tb] = (kk < kx && p < n) ? X[(long long)kk * ldx + p] : T(0);
    }
#pragma unroll
    for (int ta = 0; ta < 4; ++ta)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = Acc<T>::mfma(qa[ta], xb[tb], acc[ta][tb]);
  }
#pragma unroll
  for (int ta = 0; ta < 4; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = c0 + 16 * ta + Acc<T>::row(lk, r);
        const long long p = p0 + 16 * tb + lm;
        if (col < ky && p < n) Z[(long long)col * ldz + p] = acc[ta][tb][r];
      }
}

// FP32 form with every output column in one workgroup, so X is read from
// memory once: waves (pw, cw) = (point half, column half) of 128 points x
// 32*NCH columns; a wave accumulates 64 points x NCH column tiles.  Its 64
// points form 4 interleaved tiles (tile j = points 4 n + j), so each lane loads
// the 4 consecutive points it feeds to the 4 tiles with one float4 per
// 4-vector step and stores them back with one float4 per output row.  Q^T
// comes through LDS in 16-vector slabs (double-buffered; the next slab is
// fetched into registers while the MFMAs run).  Requires ky <= 32*NCH,
// 16-byte aligned X/Z columns (ld % 4 == 0).
template <int NCH>
__global__ __launch_bounds__(256, 2) void rotate_wide_f32(const float* __restrict__ X, long long ldx,
                                                       const float* __restrict__ Qt, float* __restrict__ Z,
                                                       long long ldz, long long n, int kx, int ky, int nblocks,
                                                       int nb8) {
  constexpr int KS = 16;             // vectors per Q slab
  constexpr int NC = 32 * NCH;       // columns per workgroup
  constexpr int QPT = KS * NC / 256; // slab values per thread
  __shared__ float qs[2][KS * NC];
  const int L = (blockIdx.x & 7) * nb8 + (blockIdx.x >> 3);
  if (L >= nblocks) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pw = w >> 1, cw = w & 1;
  const int lm = lane & 15, lk = lane >> 4;
  const long long pl = (long long)L * 128 + pw * 64 + 4 * lm;   // this lane's 4 points
  const int c0 = cw * 16 * NCH;
  v4f acc[NCH][4];
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[c][j] = v4f{0, 0, 0, 0};
  float qr[QPT];
  auto qfetch = [&](int kb) {
#pragma unroll
    for (int t = 0; t < QPT; ++t) {
      const int e = threadIdx.x + 256 * t;
      const int kk = e / NC, col = e - kk * NC;
      qr[t] = (kb + kk < kx && col < ky) ? Qt[(long long)(kb + kk) * ky + col] : 0.0f;
    }
  };
  auto qstore = [&](int bf) {
#pragma unroll
    for (int t = 0; t < QPT; ++t) qs[bf][threadIdx.x + 256 * t] = qr[t];
  };
  const bool full = pl + 3 < n;
  // X of one slab (4 steps x 4 points per lane), fetched one slab ahead
  float xc[KS / 4][4], xn[KS / 4][4];
  auto xfetch = [&](int kb, float (&xo)[KS / 4][4]) {
#pragma unroll
    for (int st = 0; st < KS / 4; ++st) {
      const int kk = kb + 4 * st + lk;
#pragma unroll
      for (int j = 0; j < 4; ++j) xo[st][j] = 0.0f;
      if (kk < kx) {
        const float* xr = X + (long long)kk * ldx + pl;
        if (full) {
          const float4 t = *reinterpret_cast<const float4*>(xr);
          xo[st][0] = t.x; xo[st][1] = t.y; xo[st][2] = t.z; xo[st][3] = t.w;
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) xo[st][j] = pl + j < n ? xr[j] : 0.0f;
        }
      }
    }
  };
  qfetch(0);
  xfetch(0, xc);
  qstore(0);
  __syncthreads();
  int bf = 0;
  for (int kb = 0; kb < kx; kb += KS) {
    const bool more = kb + KS < kx;
    if (more) {
      qfetch(kb + KS);
      xfetch(kb + KS, xn);
    }
#pragma unroll
    for (int st = 0; st < KS / 4; ++st) {
      const float* qrow = &qs[bf][(4 * st + lk) * NC + c0 + lm];
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const float qa = qrow[16 * c];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[c][j] = Acc<float>::mfma(qa, xc[st][j], acc[c][j]);
      }
    }
    if (more) {
      qstore(bf ^ 1);
#pragma unroll
      for (int st = 0; st < KS / 4; ++st)
#pragma unroll
        for (int j = 0; j < 4; ++j) xc[st][j] = xn[st][j];
    }
    __syncthreads();
    bf ^= 1;
  }
  // lane (lm, lk) register r of tile (c, j): column c0 + 16c + 4lk + r, point pl + j
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = c0 + 16 * c + Acc<float>::row(lk, r);
      if (col >= ky) continue;
      float* zr = Z + (long long)col * ldz + pl;
      if (full) {
        *reinterpret_cast<float4*>(zr) = make_float4(acc[c][0][r], acc[c][1][r], acc[c][2][r], acc[c][3][r]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (pl + j < n) zr[j] = acc[c][j][r];
      }
    }
}

// FP32 rotation on bf16 MFMA with three-way splits: every FP32 operand value
// x is split exactly into bf16 parts x = h + m + l (truncation: each part keeps
// the next 8 significant bits), and Z = X Q takes the six products whose parts
// carry at least 2^-16 of the value (lo*hi, mid*mid, hi*lo, mid*hi, hi*mid,
// hi*hi, small terms first) on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation: on random operands the same accuracy as the FP32 MFMA path
// (the dropped products are below 2^-24 relative) at 6 / 16 of its MFMA cost
// (Si87H76 block, 605,808 x 275 x 275: 784 us against 1086 us).  Q (FP64) is rounded
// to FP32 and split once by split_q_bf3 into slabs of 32 k's:
// Qp[slab][column][part][32] (bf16 bits, columns padded to 32 NCH, zero
// beyond kx / ky).  The 16x16x32 operand layouts are only relied on through
// their shared K order: A (Q^T) lane m + 16 g element e and B (X^T) lane
// n + 16 g element e both carry k = 8 g + e; D[m][n] in lane n + 16 (m / 4),
// register m % 4, as for the other 16 x 16 FP32-output MFMAs.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int kBf3K = 32;   // k's per slab

__device__ __forceinline__ void split3(float x, unsigned& h, unsigned& m, unsigned& l) {
  h = __builtin_bit_cast(unsigned, x) & 0xffff0000u;
  const float r1 = x - __builtin_bit_cast(float, h);
  m = __builtin_bit_cast(unsigned, r1) & 0xffff0000u;
  l = __builtin_bit_cast(unsigned, r1 - __builtin_bit_cast(float, m));   // at most 8 significant bits
}

__global__ void split_q_bf3(const double* __restrict__ Q, int ldq, int kx, int ky, int nc, int nks,
                            unsigned short* __restrict__ Qp) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nks * nc * kBf3K) return;
  const int kk = (int)(t % kBf3K), col = (int)((t / kBf3K) % nc), sl = (int)(t / ((long long)kBf3K * nc));
  const int k = kBf3K * sl + kk;
  const float x = (k < kx && col < ky) ? (float)Q[(long long)col * ldq + k] : 0.0f;
  unsigned h, m, l;
  split3(x, h, m, l);
  unsigned short* o = Qp + ((long long)sl * nc + col) * 3 * kBf3K + kk;
  o[0] = (unsigned short)(h >> 16);
  o[kBf3K] = (unsigned short)(m >> 16);
  o[2 * kBf3K] = (unsigned short)(l >> 16);
}

// Workgroup: 8 waves (pw = point quarter, cw = column half) over 256 points
// and all 32 NCH columns; a wave accumulates 64 points (4 interleaved tiles,
// tile j = points 4 n + j, one float4 of X per vector and lane) x 16 NCH
// columns.  Per slab of 32 k's: the Q parts through LDS (double-buffered),
// X of the next slab loaded into registers while the MFMAs run.  Requires
// 16-byte aligned X/Z columns (ld % 4 == 0).
template <int NCH>
__global__ __launch_bounds__(512) void rotate_bf3(const float* __restrict__ X, long long ldx,
                                                  const unsigned short* __restrict__ Qp, float* __restrict__ Z,
                                                  long long ldz, long long n, int kx, int ky, int nblocks, int nb8) {
  constexpr int NC = 32 * NCH;
  constexpr int SLAB = NC * 3 * kBf3K;   // bf16 values per slab
  __shared__ __attribute__((aligned(16))) unsigned short qs[2][SLAB];
  const int L = (blockIdx.x & 7) * nb8 + (blockIdx.x >> 3);
  if (L >= nblocks) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pw = w >> 1, cw = w & 1;
  const int lm = lane & 15, g = lane >> 4;
  const long long pl = (long long)L * 256 + pw * 64 + 4 * lm;
  const int c0 = cw * 16 * NCH;
  const int nks = (kx + kBf3K - 1) / kBf3K;
  const bool full = pl + 3 < n;
  v4f acc[NCH][4];
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[c][j] = v4f{0, 0, 0, 0};
  auto qload = [&](int sl, int bf) {
    const u32x4* src = reinterpret_cast<const u32x4*>(Qp + (long long)sl * SLAB);
    u32x4* dst = reinterpret_cast<u32x4*>(qs[bf]);
    for (int t = threadIdx.x; t < SLAB / 8; t += 512) dst[t] = src[t];
  };
  float xr[8][4];
  auto xload = [&](int sl) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int kk = kBf3K * sl + 8 * g + e;
#pragma unroll
      for (int j = 0; j < 4; ++j) xr[e][j] = 0.0f;
      if (kk < kx) {
        const float* xp = X + (long long)kk * ldx + pl;
        if (full) {
          const float4 t = *reinterpret_cast<const float4*>(xp);
          xr[e][0] = t.x; xr[e][1] = t.y; xr[e][2] = t.z; xr[e][3] = t.w;
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) xr[e][j] = pl + j < n ? xp[j] : 0.0f;
        }
      }
    }
  };
  qload(0, 0);
  xload(0);
  __syncthreads();
  for (int sl = 0; sl < nks; ++sl) {
    const int bf = sl & 1;
    bf16x8 bh[4], bm[4], bl[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      unsigned ph[4], pm[4], pq[4];
#pragma unroll
      for (int e2 = 0; e2 < 4; ++e2) {
        unsigned h0, m0, l0, h1, m1, l1;
        split3(xr[2 * e2][j], h0, m0, l0);
        split3(xr[2 * e2 + 1][j], h1, m1, l1);
        // the high halves of two floats -> one bf16 pair (element 2 e2 in the low half)
        ph[e2] = __builtin_amdgcn_perm(h1, h0, 0x07060302u);
        pm[e2] = __builtin_amdgcn_perm(m1, m0, 0x07060302u);
        pq[e2] = __builtin_amdgcn_perm(l1, l0, 0x07060302u);
      }
      bh[j] = __builtin_bit_cast(bf16x8, u32x4{ph[0], ph[1], ph[2], ph[3]});
      bm[j] = __builtin_bit_cast(bf16x8, u32x4{pm[0], pm[1], pm[2], pm[3]});
      bl[j] = __builtin_bit_cast(bf16x8, u32x4{pq[0], pq[1], pq[2], pq[3]});
    }
    const bool more = sl + 1 < nks;
    if (more) {
      xload(sl + 1);
      qload(sl + 1, bf ^ 1);
    }
    const unsigned short* qb = qs[bf];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const unsigned short* qa = qb + (c0 + 16 * c + lm) * 3 * kBf3K + 8 * g;
      const bf16x8 ah = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(qa));
      const bf16x8 am = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(qa + kBf3K));
      const bf16x8 al = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(qa + 2 * kBf3K));
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[c][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm[j], acc[c][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[c][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh[j], acc[c][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm[j], acc[c][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[c][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = c0 + 16 * c + Acc<float>::row(g, r);
      if (col >= ky) continue;
      float* zr = Z + (long long)col * ldz + pl;
      if (full) {
        *reinterpret_cast<float4*>(zr) = make_float4(acc[c][0][r], acc[c][1][r], acc[c][2][r], acc[c][3][r]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (pl + j < n) zr[j] = acc[c][j][r];
      }
    }
}

size_t bf3_work_bytes(int kx, int ky) {
  const size_t nks = (size_t)((kx + kBf3K - 1) / kBf3K), nc = 32 * (size_t)((ky + 31) / 32);
  return nks * nc * 3 * kBf3K * sizeof(unsigned short);
}
// Off by default: on random operands it matches the FP32 MFMA rotation, but
// in the Si87H76 SCF, where the rotation matrix R^-1 Q has large entries that
// cancel, the mixed-precision run took 19 SCF iterations instead of 15 with it
// (same energy to 3.6e-7 Ha/atom), so the 32-product sums inside the bf16 MFMA
// lose more than the FP32 path's sequential accumulation there.
bool g_rotate_bf3 = false;   // pc_rotate_set_split_bf16

__global__ void convert_q_f32(const double* __restrict__ Q, int ldq, int kx, int ky, float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kx * ky) return;
  const int i = e % kx, j = e / kx;
  out[(long long)i * ky + j] = (float)Q[(long long)j * ldq + i];  // row-major kx x ky -> ldq' = ky
}
__global__ void convert_q_f64(const double* __restrict__ Q, int ldq, int kx, int ky, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kx * ky) return;
  const int i = e % kx, j = e / kx;
  out[(long long)i * ky + j] = Q[(long long)j * ldq + i];
}

struct GramPlan {
  int ti_n, tj_n, ntile, nchunk;
  long long chunk;
};

GramPlan plan_gram(long long n, int kx, int ky, int sym) {
  GramPlan p;
  p.ti_n = (kx + 63) / 64;
  p.tj_n = (ky + 63) / 64;
  p.ntile = sym ? p.ti_n * (p.ti_n + 1) / 2 : p.ti_n * p.tj_n;
  long long target = std::max<long long>(1, 2048 / p.ntile);
  long long maxch = std::max<long long>(1, (n + 1023) / 1024);
  p.nchunk = (int)std::min(target, maxch);
  long long chunk = (n + p.nchunk - 1) / p.nchunk;
  chunk = (chunk + 31) / 32 * 32;   // whole 32-point stages of gram_lds_f32
  p.chunk = chunk;
  p.nchunk = (int)((n + chunk - 1) / chunk);
  if (p.nchunk < 1) p.nchunk = 1;
  return p;
}

size_t gram_work_bytes(const GramPlan& p) { return (size_t)p.nchunk * p.ntile * 4096 * sizeof(double); }

template <typename T>
int gram_impl(const T* X, long long ldx, const T* Y, long long ldy, long long n, int kx, int ky, double* S,
              long long lds, int sym, void* work, size_t work_bytes, void* stream) {
  PC_CHECK_ARG(X && Y && S && work, "null pointer");
  PC_CHECK_ARG(n >= 1 && kx >= 1 && ky >= 1, "empty operand");
  PC_CHECK_ARG(ldx >= n && ldy >= n && lds >= kx, "leading dimension too small");
  PC_CHECK_ARG(!sym || (kx == ky), "sym requires kx == ky");
  GramPlan p = plan_gram(n, kx, ky, sym);
  PC_CHECK_ARG(work_bytes >= gram_work_bytes(p), "workspace too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(work);
  const bool vec = (ldx % 4 == 0) && (ldy % 4 == 0) &&
                   (reinterpret_cast<uintptr_t>(X) % (4 * sizeof(T)) == 0) &&
                   (reinterpret_cast<uintptr_t>(Y) % (4 * sizeof(T)) == 0);
  dim3 grid(p.ntile * p.nchunk);
  if constexpr (sizeof(T) == 4) {
    if (vec) {
      hipLaunchKernelGGL(gram_lds_f32, grid, dim3(256), 0, st, X, ldx, Y, ldy, n, kx, ky, p.ti_n, p.tj_n, p.ntile,
                         sym, p.nchunk, p.chunk, part);
      PC_LAUNCH_CHECK();
      hipLaunchKernelGGL(gram_reduce, dim3(16 * p.ntile), dim3(256), 0, st, part, p.ntile, p.nchunk, p.ti_n, p.tj_n,
                         kx, ky, S, lds, sym);
      PC_LAUNCH_CHECK();
      return pc::PC_OK;
    }
  }
  if (vec)
    hipLaunchKernelGGL((gram_kernel<T, true>), grid, dim3(256), 0, st, X, ldx, Y, ldy, n, kx, ky, p.ti_n, p.tj_n,
                       p.ntile, sym, p.nchunk, p.chunk, part);
  else
    hipLaunchKernelGGL((gram_kernel<T, false>), grid, dim3(256), 0, st, X, ldx, Y, ldy, n, kx, ky, p.ti_n, p.tj_n,
                       p.ntile, sym, p.nchunk, p.chunk, part);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(gram_reduce, dim3(16 * p.ntile), dim3(256), 0, st, part, p.ntile, p.nchunk, p.ti_n, p.tj_n, kx,
                     ky, S, lds, sym);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

template <typename T>
int rotate_impl(const T* X, long long ldx, const double* Q, int ldq, T* Z, long long ldz, long long n, int kx,
                int ky, void* work, size_t work_bytes, void* stream) {
  PC_CHECK_ARG(X && Q && Z && work, "null pointer");
  PC_CHECK_ARG(X != Z, "output must not alias input");
  PC_CHECK_ARG(n >= 1 && kx >= 1 && ky >= 1, "empty operand");
  PC_CHECK_ARG(ldx >= n && ldz >= n && ldq >= kx, "leading dimension too small");
  PC_CHECK_ARG(work_bytes >= (size_t)kx * ky * sizeof(T), "workspace too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  T* qt = static_cast<T*>(work);
  const int ne = kx * ky;
  auto convert_q = [&]() -> int {
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL(convert_q_f32, dim3((ne + 255) / 256), dim3(256), 0, st, Q, ldq, kx, ky, qt);
    else
      hipLaunchKernelGGL(convert_q_f64, dim3((ne + 255) / 256), dim3(256), 0, st, Q, ldq, kx, ky, qt);
    PC_LAUNCH_CHECK();
    return pc::PC_OK;
  };
  if constexpr (sizeof(T) == 4) {
    const bool vec = (ldx % 4 == 0) && (ldz % 4 == 0) && (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(Z) % 16 == 0);
    const int nch = (ky + 31) / 32;
    if (vec && nch <= 9 && g_rotate_bf3 && work_bytes >= bf3_work_bytes(kx, ky)) {
      unsigned short* qp = static_cast<unsigned short*>(work);
      const int nks = (kx + kBf3K - 1) / kBf3K;
      const long long nq = (long long)nks * 32 * nch * kBf3K;
      hipLaunchKernelGGL(split_q_bf3, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, Q, ldq, kx, ky, 32 * nch,
                         nks, qp);
      PC_LAUNCH_CHECK();
      const int nb = (int)((n + 255) / 256);
      const int nb8 = (nb + 7) / 8;
      switch (nch) {
#define PC_ROTB(N)                                                                                                \
  case N:                                                                                                         \
    hipLaunchKernelGGL(rotate_bf3<N>, dim3(8 * nb8), dim3(512), 0, st, X, ldx, qp, Z, ldz, n, kx, ky, nb, nb8);    \
    break;
        PC_ROTB(1) PC_ROTB(2) PC_ROTB(3) PC_ROTB(4) PC_ROTB(5) PC_ROTB(6) PC_ROTB(7) PC_ROTB(8) PC_ROTB(9)
#undef PC_ROTB
      }
      PC_LAUNCH_CHECK();
      return pc::PC_OK;
    }
    if (vec && nch <= 9) {
      if (convert_q() != pc::PC_OK) return pc::PC_ERR_HIP;
      const int nb = (int)((n + 127) / 128);
      const int nb8 = (nb + 7) / 8;
      switch (nch) {
#define PC_ROT(N)                                                                                                 \
  case N:                                                                                                         \
    hipLaunchKernelGGL(rotate_wide_f32<N>, dim3(8 * nb8), dim3(256), 0, st, X, ldx, qt, Z, ldz, n, kx, ky, nb, nb8); \
    break;
        PC_ROT(1) PC_ROT(2) PC_ROT(3) PC_ROT(4) PC_ROT(5) PC_ROT(6) PC_ROT(7) PC_ROT(8) PC_ROT(9)
#undef PC_ROT
      }
      PC_LAUNCH_CHECK();
      return pc::PC_OK;
    }
  }
  if (convert_q() != pc::PC_OK) return pc::PC_ERR_HIP;
  const int npt = (int)((n + 127) / 128);
  const int nct = (ky + 63) / 64;
  const int nb = npt * nct;
  const int nb8 = (nb + 7) / 8;
  hipLaunchKernelGGL((rotate_kernel<T>), dim3(8 * nb8), dim3(256), 0, st, X, ldx, qt, ky, Z, ldz, n, kx, ky, npt,
                     nct, nb, nb8);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

}  // namespace

extern "C" {

int pc_gram_workspace(long long n, int kx, int ky, int sym, size_t* out_bytes) {
  PC_CHECK_ARG(out_bytes, "null pointer");
  *out_bytes = gram_work_bytes(plan_gram(n, kx, ky, sym));
  return pc::PC_OK;
}

int pc_gram_f64(const double* X, long long ldx, const double* Y, long long ldy, long long n, int kx, int ky,
                double* S, long long lds, int sym, void* work, size_t work_bytes, void* stream) {
  return gram_impl<double>(X, ldx, Y, ldy, n, kx, ky, S, lds, sym, work, work_bytes, stream);
}

int pc_gram_f32(const float* X, long long ldx, const float* Y, long long ldy, long long n, int kx, int ky,
                double* S, long long lds, int sym, void* work, size_t work_bytes, void* stream) {
  return gram_impl<float>(X, ldx, Y, ldy, n, kx, ky, S, lds, sym, work, work_bytes, stream);
}

int pc_rotate_workspace(int kx, int ky, int fp32, size_t* out_bytes) {
  PC_CHECK_ARG(out_bytes && kx >= 1 && ky >= 1, "bad argument");
  const size_t base = (size_t)kx * ky * (fp32 ? sizeof(float) : sizeof(double));
  *out_bytes = fp32 ? std::max(base, bf3_work_bytes(kx, ky)) : base;
  return pc::PC_OK;
}

int pc_rotate_set_split_bf16(int on) {
  g_rotate_bf3 = on != 0;
  return pc::PC_OK;
}

int pc_rotate_f64(const double* X, long long ldx, const double* Q, int ldq, double* Z, long long ldz, long long n,
                  int kx, int ky, void* work, size_t work_bytes, void* stream) {
  return rotate_impl<double>(X, ldx, Q, ldq, Z, ldz, n, kx, ky, work, work_bytes, stream);
}

int pc_rotate_f32(const float* X, long long ldx, const double* Q, int ldq, float* Z, long long ldz, long long n,
                  int kx, int ky, void* work, size_t work_bytes, void* stream) {
  return rotate_impl<float>(X, ldx, Q, ldq, Z, ldz, n, kx, ky, work, work_bytes, stream);
}

}  // extern "C"
