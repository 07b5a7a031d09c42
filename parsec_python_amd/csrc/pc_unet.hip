// Learned density initialiser: the kernels of a small 3D U-Net (forward,
// backward, Adam), BF16 storage with FP32 accumulation on the bf16 MFMA.
//
// Layout: one volume, channel-last: element (voxel v, channel c) at v*ld + c,
// v = (z*H + y)*W + x.  A channel stride ld larger than the channel count lets
// a layer read or write a slice of a concatenated tensor in place.
// Weights: W[co][t][ci] (tap t = (dz+1)*9 + (dy+1)*3 + (dx+1), ci fastest),
// 3x3x3 taps, zero padding 1.
//
// conv3d (implicit GEMM, M = output channels, N = voxels, K = 27 * C_in):
// each wavefront owns 16*MT output channels x 32 voxels and issues
// v_mfma_f32_16x16x16_bf16 on fragments gathered straight from the
// channel-last tensors (4 consecutive channels = one 8-byte load per lane).
// The input gradient is the same kernel on the tap-flipped, channel-transposed
// weights; the weight gradient is a GEMM over voxels per tap.
#include "pc_common.h"
#include <algorithm>
#include <cmath>

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float f) {
  const unsigned u = __float_as_uint(f);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct Vol {
  int D, H, W;
  long long V;
};

__device__ __forceinline__ void tap_offsets(int t, int& dz, int& dy, int& dx) {
  dz = t / 9 - 1;
  dy = (t / 3) % 3 - 1;
  dx = t % 3 - 1;
}

constexpr int kNT = 2;   // 16-voxel tiles per wavefront
constexpr int kWaves = 4;

// mode: 0 = FP32 output with bias, 1 = leaky ReLU -> BF16 output (with bias),
//       2 = FP32 output, no bias
template <int MT, bool CIN1>
__global__ __launch_bounds__(256) void conv_fwd(Vol vol, const unsigned short* __restrict__ X, int ldx, int cin,
                                                const unsigned short* __restrict__ Wt, const float* __restrict__ bias,
                                                int cout, int mode, float slope, void* Y, int ldy) {
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int co_base = blockIdx.y * 64;
  const long long vbase = ((long long)blockIdx.x * kWaves + wv) * (16 * kNT);
  const int ln = lane & 15, lk = lane >> 4;

  // the voxels of this lane's B columns
  int vz[kNT], vy[kNT], vx[kNT];
  bool vin[kNT];
#pragma unroll
  for (int n = 0; n < kNT; ++n) {
    const long long v = vbase + 16 * n + ln;
    vin[n] = v < vol.V;
    const long long vv = vin[n] ? v : 0;
    vx[n] = (int)(vv % vol.W);
    vy[n] = (int)((vv / vol.W) % vol.H);
    vz[n] = (int)(vv / ((long long)vol.W * vol.H));
  }
  f32x4 acc[MT][kNT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int n = 0; n < kNT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};

  if constexpr (!CIN1) {
    const int ncb = cin >> 4;
    for (int t = 0; t < 27; ++t) {
      int dz, dy, dx;
      tap_offsets(t, dz, dy, dx);
      long long src[kNT];
      bool ok[kNT];
#pragma unroll
      for (int n = 0; n < kNT; ++n) {
        const int z = vz[n] + dz, y = vy[n] + dy, x = vx[n] + dx;
        ok[n] = vin[n] && z >= 0 && z < vol.D && y >= 0 && y < vol.H && x >= 0 && x < vol.W;
        src[n] = ok[n] ? ((long long)(z * vol.H + y) * vol.W + x) * ldx : 0;
      }
      for (int cb = 0; cb < ncb; ++cb) {
        const int ci = cb * 16 + 4 * lk;
        bf16x4 b[kNT];
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
          b[n] = ok[n] ? *reinterpret_cast<const bf16x4*>(X + src[n] + ci) : bf16x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const int co = co_base + 16 * m + ln;
          const bf16x4 a = co < cout ? *reinterpret_cast<const bf16x4*>(Wt + ((long long)co * 27 + t) * cin + ci)
                                     : bf16x4{0, 0, 0, 0};
#pragma unroll
          for (int n = 0; n < kNT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[n], acc[m][n], 0, 0, 0);
        }
      }
    }
  } else {
    // one input channel: K = the 27 taps (two k-steps of 16, taps 27..31 zero)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x4 b[kNT];
      short av[MT][4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int t = ks * 16 + 4 * lk + j;
        int dz = 0, dy = 0, dx = 0;
        if (t < 27) tap_offsets(t, dz, dy, dx);
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
          const int z = vz[n] + dz, y = vy[n] + dy, x = vx[n] + dx;
          const bool ok = t < 27 && vin[n] && z >= 0 && z < vol.D && y >= 0 && y < vol.H && x >= 0 && x < vol.W;
          b[n][j] = ok ? (short)X[((long long)(z * vol.H + y) * vol.W + x) * ldx] : (short)0;
        }
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const int co = co_base + 16 * m + ln;
          av[m][j] = (t < 27 && co < cout) ? (short)Wt[(long long)co * 27 + t] : (short)0;
        }
      }
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const bf16x4 a = {av[m][0], av[m][1], av[m][2], av[m][3]};
#pragma unroll
        for (int n = 0; n < kNT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[n], acc[m][n], 0, 0, 0);
      }
    }
  }

  // epilogue: lane holds D[co = 16m + 4*lk + r][voxel = 16n + ln]
#pragma unroll
  for (int n = 0; n < kNT; ++n) {
    if (!vin[n]) continue;
    const long long v = vbase + 16 * n + ln;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int co0 = co_base + 16 * m + 4 * lk;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = co0 + r;
        if (co >= cout) continue;
        float o = acc[m][n][r];
        if (mode != 2) o += bias[co];
        if (mode == 1) {
          o = o > 0.f ? o : slope * o;
          static_cast<unsigned short*>(Y)[v * ldy + co] = f2bf(o);
        } else {
          static_cast<float*>(Y)[v * ldy + co] = o;
        }
      }
    }
  }
}

// Wt[ci][t][co] = W[co][26 - t][ci]: the weights of the input-gradient conv
__global__ void conv_wflip(const unsigned short* __restrict__ W, int cout, int cin, unsigned short* __restrict__ Wt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)cout * cin * 27;
  if (i >= n) return;
  const int ci = (int)(i % cin);
  const int t = (int)((i / cin) % 27);
  const int co = (int)(i / ((long long)cin * 27));
  Wt[((long long)ci * 27 + (26 - t)) * cout + co] = W[i];
}

// Weight gradient for one tap-block: workgroup (tap t, 16 output channels,
// 16*NT input channels, voxel split s).  dW_part[s][co][t][ci].
template <int NTC>
__global__ __launch_bounds__(256) void conv_wgrad(Vol vol, const unsigned short* __restrict__ dZ, int ldz, int cout,
                                                  const unsigned short* __restrict__ X, int ldx, int cin,
                                                  long long vchunk, float* __restrict__ part) {
  __shared__ float red[kWaves][16 * 16 * NTC];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int ln = lane & 15, lk = lane >> 4;
  const int t = blockIdx.x % 27;
  const int cob = blockIdx.x / 27;
  const int cib = blockIdx.y;
  const int s = blockIdx.z;
  int dz, dy, dx;
  tap_offsets(t, dz, dy, dx);
  const int co = cob * 16 + ln;
  const long long v0 = (long long)s * vchunk;
  const long long v1 = min(v0 + vchunk, vol.V);
  f32x4 acc[NTC];
#pragma unroll
  for (int n = 0; n < NTC; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
  // each wavefront takes every 4th group of 16 voxels
  for (long long vb = v0 + 16 * wv; vb < v1; vb += 16 * kWaves) {
    bf16x4 a;
    bf16x4 b[NTC];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long v = vb + 4 * lk + j;
      const bool vok = v < v1;
      a[j] = (vok && co < cout) ? (short)dZ[v * ldz + co] : (short)0;
      long long src = -1;
      if (vok) {
        const int x = (int)(v % vol.W) + dx;
        const int y = (int)((v / vol.W) % vol.H) + dy;
        const int z = (int)(v / ((long long)vol.W * vol.H)) + dz;
        if (z >= 0 && z < vol.D && y >= 0 && y < vol.H && x >= 0 && x < vol.W)
          src = ((long long)(z * vol.H + y) * vol.W + x) * ldx;
      }
#pragma unroll
      for (int n = 0; n < NTC; ++n) {
        const int ci = cib * 16 * NTC + 16 * n + ln;
        b[n][j] = (src >= 0 && ci < cin) ? (short)X[src + ci] : (short)0;
      }
    }
#pragma unroll
    for (int n = 0; n < NTC; ++n) acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[n], acc[n], 0, 0, 0);
  }
  // D[m = co row 4*lk + r][n = ci col ln]
#pragma unroll
  for (int n = 0; n < NTC; ++n)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wv][(4 * lk + r) * (16 * NTC) + 16 * n + ln] = acc[n][r];
  __syncthreads();
  for (int e = threadIdx.x; e < 16 * 16 * NTC; e += 256) {
    const float sum = red[0][e] + red[1][e] + red[2][e] + red[3][e];
    const int m = e / (16 * NTC), c = e % (16 * NTC);
    const int com = cob * 16 + m, cim = cib * 16 * NTC + c;
    if (com < cout && cim < cin) part[(((long long)s * cout + com) * 27 + t) * cin + cim] = sum;
  }
}

// dW[i] = sum_s part[s][i] (fixed order)
__global__ void reduce_splits(const float* __restrict__ part, int nsplit, long long n, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < nsplit; ++k) s += part[(long long)k * n + i];
  out[i] = s;
}

// dZ = dA * (a > 0 ? 1 : slope) (a == nullptr: identity), rounded to BF16;
// per-block partial bias gradients sum(dZ) per channel (FP32 before rounding).
__global__ __launch_bounds__(256) void act_bwd(const float* __restrict__ dA, int lda,
                                               const unsigned short* __restrict__ a, int lda_a, int C, long long V,
                                               float slope, unsigned short* __restrict__ dZ, int ldz,
                                               long long vchunk, float* __restrict__ dbpart) {
  __shared__ float red[256];
  const int c = threadIdx.x % C;             // C divides 256 or C <= 256 handled by the grid-stride below
  const int lanes = (256 / C) * C;
  const int slot = threadIdx.x / C;
  const int nslot = 256 / C;
  float s = 0.f;
  if (threadIdx.x < lanes) {
    const long long v0 = (long long)blockIdx.x * vchunk;
    const long long v1 = min(v0 + vchunk, V);
    for (long long v = v0 + slot; v < v1; v += nslot) {
      float g = dA[v * lda + c];
      if (a) g = bf2f(a[v * lda_a + c]) > 0.f ? g : slope * g;
      dZ[v * ldz + c] = f2bf(g);
      s += g;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float tot = 0.f;
    for (int k = 0; k < nslot; ++k) tot += red[k * C + threadIdx.x];
    dbpart[(long long)blockIdx.x * C + threadIdx.x] = tot;
  }
}

__global__ void pool2_fwd(Vol out, const unsigned short* __restrict__ X, int ldx, int C,
                          unsigned short* __restrict__ Y, int ldy) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= out.V * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int x = (int)(v % out.W), y = (int)((v / out.W) % out.H), z = (int)(v / ((long long)out.W * out.H));
  const int W2 = 2 * out.W, H2 = 2 * out.H;
  float s = 0.f;
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
    s += bf2f(X[((long long)(zz * H2 + yy) * W2 + xx) * ldx + c]);
  }
  Y[v * ldy + c] = f2bf(s * 0.125f);
}

// dX (fine) = dY(coarse)/8 + dSkip (fine, optional)
__global__ void pool2_bwd(Vol fine, const float* __restrict__ dY, int ldy, const float* __restrict__ dS, int lds,
                          int C, float* __restrict__ dX, int ldx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fine.V * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int x = (int)(v % fine.W), y = (int)((v / fine.W) % fine.H), z = (int)(v / ((long long)fine.W * fine.H));
  const int Wc = fine.W / 2, Hc = fine.H / 2;
  float g = dY[((long long)((z / 2) * Hc + y / 2) * Wc + x / 2) * ldy + c] * 0.125f;
  if (dS) g += dS[v * lds + c];
  dX[v * ldx + c] = g;
}

__global__ void up2_fwd(Vol fine, const unsigned short* __restrict__ X, int ldx, int C,
                        unsigned short* __restrict__ Y, int ldy) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fine.V * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int x = (int)(v % fine.W), y = (int)((v / fine.W) % fine.H), z = (int)(v / ((long long)fine.W * fine.H));
  const int Wc = fine.W / 2, Hc = fine.H / 2;
  Y[v * ldy + c] = X[((long long)((z / 2) * Hc + y / 2) * Wc + x / 2) * ldx + c];
}

// dX (coarse) = sum of the 8 fine children of dY
__global__ void up2_bwd(Vol coarse, const float* __restrict__ dY, int ldy, int C, float* __restrict__ dX, int ldx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= coarse.V * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int x = (int)(v % coarse.W), y = (int)((v / coarse.W) % coarse.H);
  const int z = (int)(v / ((long long)coarse.W * coarse.H));
  const int W2 = 2 * coarse.W, H2 = 2 * coarse.H;
  float s = 0.f;
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
    s += dY[((long long)(zz * H2 + yy) * W2 + xx) * ldy + c];
  }
  dX[v * ldx + c] = s;
}

__global__ void cast_bf16(const float* __restrict__ src, long long n, unsigned short* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

// Adam with bias correction; writes the BF16 copy of the updated weights.
__global__ void adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                     long long n, float lr, float b1, float b2, float eps, float wd, float bc1, float bc2,
                     unsigned short* __restrict__ wbf) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wn = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wn;
  if (wbf) wbf[i] = f2bf(wn);
}

// dy = 2 (y - t) / n; per-block partial sums of (y - t)^2 (FP64)
__global__ __launch_bounds__(256) void mse(const float* __restrict__ y, const float* __restrict__ t, long long n,
                                           float* __restrict__ dy, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float d = y[i] - t[i];
    dy[i] = 2.f * d / (float)n;
    s += (double)d * d;
  }
  const double tot = pc_block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void sum_parts_f32(const float* __restrict__ part, int nblk, int C, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[(long long)b * C + c];
  out[c] = s;
}

__global__ void sum_parts_f64(const double* __restrict__ part, int nblk, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[b];
  out[0] = s;
}

Vol make_vol(int D, int H, int W) { return Vol{D, H, W, (long long)D * H * W}; }

inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

constexpr int kWgradSplit = 64;
constexpr int kActChunk = 2048;
constexpr int kMseBlocks = 512;

}  // namespace

extern "C" {

int pc_unet_conv3d_fwd(const void* X, int ldx, int cin, const void* W, const float* bias, int cout, int D, int H,
                       int Wd, int mode, float slope, void* Y, int ldy, void* stream) {
  PC_CHECK_ARG(X && W && Y, "null pointer");
  PC_CHECK_ARG(D > 0 && H > 0 && Wd > 0, "bad volume");
  PC_CHECK_ARG(cin == 1 || (cin % 16 == 0), "cin must be 1 or a multiple of 16");
  PC_CHECK_ARG(cout >= 1, "cout < 1");
  PC_CHECK_ARG(ldx >= cin && ldy >= cout, "channel stride smaller than the channel count");
  PC_CHECK_ARG(cin == 1 || (ldx % 4 == 0), "ldx must be a multiple of 4");
  PC_CHECK_ARG(mode >= 0 && mode <= 2, "bad mode");
  PC_CHECK_ARG(mode == 2 || bias, "bias required");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Vol vol = make_vol(D, H, Wd);
  const long long vox_per_block = 16LL * kNT * kWaves;
  dim3 grid((unsigned)((vol.V + vox_per_block - 1) / vox_per_block), (unsigned)((cout + 63) / 64));
  const int mt = std::min(4, (cout + 15) / 16);
  auto X16 = static_cast<const unsigned short*>(X);
  auto W16 = static_cast<const unsigned short*>(W);
#define PC_CONV_LAUNCH(MT, C1)                                                                                  \
  hipLaunchKernelGGL((conv_fwd<MT, C1>), grid, dim3(256), 0, st, vol, X16, ldx, cin, W16, bias, cout, mode, slope, \
                     Y, ldy)
  if (cin == 1) {
    if (mt == 1) PC_CONV_LAUNCH(1, true);
    else if (mt == 2) PC_CONV_LAUNCH(2, true);
    else if (mt == 3) PC_CONV_LAUNCH(3, true);
    else PC_CONV_LAUNCH(4, true);
  } else {
    if (mt == 1) PC_CONV_LAUNCH(1, false);
    else if (mt == 2) PC_CONV_LAUNCH(2, false);
    else if (mt == 3) PC_CONV_LAUNCH(3, false);
    else PC_CONV_LAUNCH(4, false);
  }
#undef PC_CONV_LAUNCH
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_unet_conv3d_wflip(const void* W, int cout, int cin, void* Wt, void* stream) {
  PC_CHECK_ARG(W && Wt, "null pointer");
  PC_CHECK_ARG(cout >= 1 && cin >= 1, "bad channels");
  const long long n = (long long)cout * cin * 27;
  hipLaunchKernelGGL(conv_wflip, dim3(nblk(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const unsigned short*>(W), cout, cin, static_cast<unsigned short*>(Wt));
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_unet_wgrad_workspace(int cout, int cin, size_t* out_bytes) {
  PC_CHECK_ARG(out_bytes, "null pointer");
  *out_bytes = (size_t)kWgradSplit * cout * 27 * cin * sizeof(float);
  return pc::PC_OK;
}

// dW[co][t][ci] = sum_v dZ[v][co] X[v + off_t][ci]  (FP32 result)
int pc_unet_conv3d_wgrad(const void* dZ, int ldz, int cout, const void* X, int ldx, int cin, int D, int H, int Wd,
                         float* dW, void* work, size_t work_bytes, void* stream) {
  PC_CHECK_ARG(dZ && X && dW && work, "null pointer");
  PC_CHECK_ARG(D > 0 && H > 0 && Wd > 0, "bad volume");
  PC_CHECK_ARG(ldz >= cout && ldx >= cin, "channel stride smaller than the channel count");
  size_t need = 0;
  pc_unet_wgrad_workspace(cout, cin, &need);
  PC_CHECK_ARG(work_bytes >= need, "workspace too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Vol vol = make_vol(D, H, Wd);
  const int nsplit = (int)std::min<long long>(kWgradSplit, (vol.V + 255) / 256);
  const long long vchunk = (vol.V + nsplit - 1) / nsplit;
  const int ntc = cin >= 32 ? 2 : 1;
  dim3 grid(27 * ((cout + 15) / 16), (cin + 16 * ntc - 1) / (16 * ntc), nsplit);
  auto dZ16 = static_cast<const unsigned short*>(dZ);
  auto X16 = static_cast<const unsigned short*>(X);
  float* part = static_cast<float*>(work);
  if (ntc == 2)
    hipLaunchKernelGGL((conv_wgrad<2>), grid, dim3(256), 0, st, vol, dZ16, ldz, cout, X16, ldx, cin, vchunk, part);
  else
    hipLaunchKernelGGL((conv_wgrad<1>), grid, dim3(256), 0, st, vol, dZ16, ldz, cout, X16, ldx, cin, vchunk, part);
  PC_LAUNCH_CHECK();
  const long long n = (long long)cout * 27 * cin;
  hipLaunchKernelGGL(reduce_splits, dim3(nblk(n)), dim3(256), 0, st, part, nsplit, n, dW);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_unet_act_bwd_workspace(int C, long long V, size_t* out_bytes) {
  PC_CHECK_ARG(out_bytes && C >= 1 && V >= 1, "bad argument");
  *out_bytes = (size_t)((V + kActChunk - 1) / kActChunk) * C * sizeof(float);
  return pc::PC_OK;
}

// dZ (BF16) = dA * leaky'(a) (a == NULL: identity); db[c] = sum_v dZ (FP32)
int pc_unet_act_bwd(const float* dA, int lda, const void* a, int lda_a, int C, int D, int H, int Wd, float slope,
                    void* dZ, int ldz, float* db, void* work, size_t work_bytes, void* stream) {
  PC_CHECK_ARG(dA && dZ && db && work, "null pointer");
  PC_CHECK_ARG(C >= 1 && C <= 256, "C must be in [1, 256]");
  const long long V = (long long)D * H * Wd;
  size_t need = 0;
  pc_unet_act_bwd_workspace(C, V, &need);
  PC_CHECK_ARG(work_bytes >= need, "workspace too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = (int)((V + kActChunk - 1) / kActChunk);
  float* part = static_cast<float*>(work);
  hipLaunchKernelGGL(act_bwd, dim3(nb), dim3(256), 0, st, dA, lda, static_cast<const unsigned short*>(a), lda_a, C, V,
                     slope, static_cast<unsigned short*>(dZ), ldz, (long long)kActChunk, part);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_parts_f32, dim3(nblk(C)), dim3(256), 0, st, part, nb, C, db);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// pooled (D/2, H/2, W/2) from X (D, H, W)
int pc_unet_pool2_fwd(const void* X, int ldx, int C, int D, int H, int Wd, void* Y, int ldy, void* stream) {
  PC_CHECK_ARG(X && Y, "null pointer");
  PC_CHECK_ARG(D % 2 == 0 && H % 2 == 0 && Wd % 2 == 0, "even volume required");
  const Vol out = make_vol(D / 2, H / 2, Wd / 2);
  hipLaunchKernelGGL(pool2_fwd, dim3(nblk(out.V * C)), dim3(256), 0, static_cast<hipStream_t>(stream), out,
                     static_cast<const unsigned short*>(X), ldx, C, static_cast<unsigned short*>(Y), ldy);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_unet_pool2_bwd(const float* dY, int ldy, const float* dS, int lds, int C, int D, int H, int Wd, float* dX,
                      int ldx, void* stream) {
  PC_CHECK_ARG(dY && dX, "null pointer");
  PC_CHECK_ARG(D % 2 == 0 && H % 2 == 0 && Wd % 2 == 0, "even volume required");
  const Vol fine = make_vol(D, H, Wd);
  hipLaunchKernelGGL(pool2_bwd, dim3(nblk(fine.V * C)), dim3(256), 0, static_cast<hipStream_t>(stream), fine, dY, ldy,
                     dS, lds, C, dX, ldx);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// fine (D, H, W) from coarse X (D/2, H/2, W/2)
int pc_unet_up2_fwd(const void* X, int ldx, int C, int D, int H, int Wd, void* Y, int ldy, void* stream) {
  PC_CHECK_ARG(X && Y, "null pointer");
  PC_CHECK_ARG(D % 2 == 0 && H % 2 == 0 && Wd % 2 == 0, "even volume required");
  const Vol fine = make_vol(D, H, Wd);
  hipLaunchKernelGGL(up2_fwd, dim3(nblk(fine.V * C)), dim3(256), 0, static_cast<hipStream_t>(stream), fine,
                     static_cast<const unsigned short*>(X), ldx, C, static_cast<unsigned short*>(Y), ldy);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// coarse gradient (D/2, H/2, W/2) from the fine gradient dY (D, H, W)
int pc_unet_up2_bwd(const float* dY, int ldy, int C, int D, int H, int Wd, float* dX, int ldx, void* stream) {
  PC_CHECK_ARG(dY && dX, "null pointer");
  PC_CHECK_ARG(D % 2 == 0 && H % 2 == 0 && Wd % 2 == 0, "even volume required");
  const Vol coarse = make_vol(D / 2, H / 2, Wd / 2);
  hipLaunchKernelGGL(up2_bwd, dim3(nblk(coarse.V * C)), dim3(256), 0, static_cast<hipStream_t>(stream), coarse, dY,
                     ldy, C, dX, ldx);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_unet_cast_bf16(const float* src, long long n, void* dst, void* stream) {
  PC_CHECK_ARG(src && dst && n >= 0, "bad argument");
  if (n == 0) return pc::PC_OK;
  hipLaunchKernelGGL(cast_bf16, dim3(nblk(n)), dim3(256), 0, static_cast<hipStream_t>(stream), src, n,
                     static_cast<unsigned short*>(dst));
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// step >= 1; wbf (optional) receives the BF16 rounding of the updated weights
int pc_unet_adam(float* w, const float* g, float* m, float* v, long long n, double lr, double b1, double b2,
                 double eps, double wd, int step, void* wbf, void* stream) {
  PC_CHECK_ARG(w && g && m && v && n >= 0 && step >= 1, "bad argument");
  if (n == 0) return pc::PC_OK;
  const float bc1 = (float)(1.0 - std::pow(b1, step));
  const float bc2 = (float)(1.0 - std::pow(b2, step));
  hipLaunchKernelGGL(adam, dim3(nblk(n)), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m, v, n, (float)lr,
                     (float)b1, (float)b2, (float)eps, (float)wd, bc1, bc2, static_cast<unsigned short*>(wbf));
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

int pc_unet_mse_workspace(size_t* out_bytes) {
  PC_CHECK_ARG(out_bytes, "null pointer");
  *out_bytes = (size_t)kMseBlocks * sizeof(double);
  return pc::PC_OK;
}

// dy = 2 (y - t)/n; loss_dev[0] (FP64, device) = sum (y - t)^2
int pc_unet_mse(const float* y, const float* t, long long n, float* dy, double* loss_dev, void* work,
                size_t work_bytes, void* stream) {
  PC_CHECK_ARG(y && t && dy && loss_dev && work && n > 0, "bad argument");
  PC_CHECK_ARG(work_bytes >= (size_t)kMseBlocks * sizeof(double), "workspace too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = (int)std::min<long long>(kMseBlocks, (n + 255) / 256);
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(mse, dim3(nb), dim3(256), 0, st, y, t, n, dy, part);
  PC_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_parts_f64, dim3(1), dim3(64), 0, st, part, nb, loss_dev);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

}  // extern "C"
