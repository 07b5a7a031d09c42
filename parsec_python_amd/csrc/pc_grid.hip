// Grid index map and the fused multi-vector finite-difference stencil.
//
// Storage ("compact, vector-major"): the domain is a set of x-rows (j,k) of an
// nx*ny*nz cube, each row holding the contiguous points lo[j,k] <= i < hi[j,k]
// (cube: the whole row; sphere: the chord).  Points are numbered k-major, then
// j, then i, so row (j,k) starts at off[j,k].  Vector v of a block lives at
// [v*ld, v*ld + npts).  Outside the domain values are zero (Dirichlet), which
// is the reference's truncated-diagonal Laplacian (src/Laplacian/fd3d.py:79-100)
// when the domain is the full cube.
//
// One launch computes, for every vector v and domain point p,
//     out = sum_d w_d * (6 neighbours at distance d) + w_0 * y  (the -lap weights,
//           pre-multiplied by the caller's scale)
//         + (a*V[p] + b) * y[p] + c * x[p]
// which is H*y (a=1,b=0,c=0) and every step of the reference's scaled and slim
// Chebyshev filters (src/Eigensolvers/ch_filter.py:66-80, first_filt.py:94-101).
// x may alias out (each point reads x before writing out at the same place).
//
// Kernel shape (CDNA4): a workgroup owns one vector, a slab of TY y-rows and
// all x, and marches along z.  Each thread keeps its points' 2R+1 z-values in
// registers; the current plane of the slab plus R halo rows above/below sits
// in LDS for the x/y differences.  Workgroup ids are remapped so the slabs of
// one vector run on the same XCD (its L2 then serves the halo rows).
#include "pc_common.h"
#include "pc_zmarch.h"
#include <type_traits>
#include <utility>
#include <vector>
#include <algorithm>

namespace {

struct GridDev {
  int nx, ny, nz;
  long long npts;
  int* lo;       // [ny*nz]
  int* hi;       // [ny*nz]
  int* off;      // [ny*nz]
  int* zrange;   // [2*ny]: min z, max z+1 of non-empty rows per y
  int2* meta;    // [ny*nz]: {off - lo, lo | hi << 16}
  int* prow;     // [npts]: row (j + ny*k) of each domain point
  int radius;
  double w[8];   // 1D -d2/dx2 weights, centre first (unscaled)
  int* cc_tab;   // stencil_cc: per slab of 8 rows, the own and halo columns (tile row << 16 | x), -1 padded
  int cc_wmax;   // its wavefronts per workgroup (0: not built)
};

struct StArgs {
  int nx, ny, nz;
  long long npts;
  const int* lo;
  const int* hi;
  const int* off;
  const int* zrange;
  const int2* meta;
  long long ld;
  int nslab;
  int nvec;
  int nblocks;   // real (unpadded) number of workgroups
  int nb8;       // ceil(nblocks/8)
  int nzc, zc;   // z-chunks per slab and planes per chunk (lean and paired kernels)
  double w[8];   // scaled weights; w[0] already includes the factor 3
};

template <typename T>
__device__ __forceinline__ T row_val(const T* __restrict__ y, int lo, int hi, int off, int i) {
  return (i >= lo && i < hi) ? y[off + (i - lo)] : T(0);
}

// LDS-only workgroup barrier: waits for this wave's LDS traffic, not for
// outstanding global loads (the prefetch of the next plane stays in flight).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt/expcnt left at max
  __builtin_amdgcn_s_barrier();
}

template <typename T>
__device__ __forceinline__ T meta_val(const T* __restrict__ y, int2 m, int i) {
  const int lo = m.y & 0xffff, hi = m.y >> 16;
  return (i >= lo && i < hi) ? y[m.x + i] : T(0);
}

// Software-pipelined z-march: while plane z is computed from LDS and the
// register queues, the loads for the next PD planes (queue entry z+R+1.., halo
// rows, V and x) are already in flight (PD register stages, the z loop unrolled
// by PD).  A workgroup handles NV vectors at once: the row metadata, the V
// values and the control flow are shared by them.  Two LDS tiles alternate,
// so one barrier per plane suffices.
template <typename T, int R, int SX, int TYW, int NV, int NB, bool HASX, bool DOT, int PD = 1, int NW = 4>
__global__ __launch_bounds__(64 * NW) void stencil_zmarch(StArgs g, const T* __restrict__ Y, const T* X,
                                                          T* Out, const T* __restrict__ V, T ca, T cb, T cc,
                                                          double* __restrict__ dot_part) {
  constexpr int TY = NW * TYW;
  constexpr int NXP = 64 * SX;
  constexpr int LW = NXP + 2 * R;   // LDS row stride (R zero columns each side)
  constexpr int LR = TY + 2 * R;    // LDS rows
  constexpr int HT = (2 * R + NW - 1) / NW;  // halo rows per wave
  static_assert(PD == 1 || PD == 2, "prefetch depth 1 or 2");
  static_assert(PD == 1 || NB == 2, "two planes in flight need two LDS tiles");
  __shared__ T tile[NB][NV][LR * LW];
  __shared__ double red[NW];

  const int bid = blockIdx.x;
  const int L = (bid & 7) * g.nb8 + (bid >> 3);
  if (L >= g.nblocks) return;
  const int vg = L / g.nslab;
  const int slab = L - vg * g.nslab;
  const int j0 = slab * TY;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ny = g.ny, nz = g.nz;
  const int2* __restrict__ meta = g.meta;

  const T* yv[NV];
  const T* xv[NV];
  T* ov[NV];
  bool act[NV];
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int v = vg * NV + u;
    act[u] = v < g.nvec;
    const long long base = (long long)(act[u] ? v : vg * NV) * g.ld;
    yv[u] = Y + base;
    xv[u] = X ? X + base : nullptr;
    ov[u] = Out + base;
  }
  constexpr bool hasx = HASX;

  int zb = nz, ze = 0;
  for (int t = 0; t < TY; ++t) {
    const int j = j0 + t;
    if (j < ny) {
      zb = min(zb, g.zrange[2 * j]);
      ze = max(ze, g.zrange[2 * j + 1]);
    }
  }
  double dot = 0.0;
  if (zb >= ze) {
    if constexpr (DOT) {
      const double s = pc_block_sum(0.0, red);
      if (threadIdx.x == 0) dot_part[L] = s;
    }
    return;
  }
  for (int t = threadIdx.x; t < NB * NV * LR * LW; t += 64 * NW) (&tile[0][0][0])[t] = T(0);
  __syncthreads();

  auto rowmeta = [&](int j, int z) -> int2 {
    return (j >= 0 && j < ny && z >= 0 && z < nz) ? meta[j + ny * z] : make_int2(0, 0);
  };
  // Row metadata windows: lane l of a window holds meta(j, wb + l), so the
  // per-plane lookups are register reads (readlane) and the addresses of a
  // plane's loads never wait for a metadata load; a window is reloaded once
  // every 64 planes.  Own row: planes z (output) and z + R (queue front).
  struct Win {
    int2 v;
    int wb;
  };
  auto win_load = [&](Win& w, int j, int wb) {
    w.wb = wb;
    w.v = rowmeta(j, wb + lane);
  };
  auto win_at = [&](const Win& w, int z) -> int2 {
    const int l = z - w.wb;
    return make_int2(__builtin_amdgcn_readlane(w.v.x, l), __builtin_amdgcn_readlane(w.v.y, l));
  };
  Win wown[TYW], whalo[HT];
#pragma unroll
  for (int r = 0; r < TYW; ++r) win_load(wown[r], j0 + wv + NW * r, zb - R);
#pragma unroll
  for (int t = 0; t < HT; ++t) {
    const int hr = wv + NW * t;
    const int tr = hr < R ? hr : hr + TY;
    win_load(whalo[t], j0 - R + tr, zb);
  }

  T q[NV][TYW][SX][2 * R + 1];
  T xq[NV][TYW][SX];
  T vq[TYW][SX];
  T hc[NV][HT][SX] = {};
  int2 mcur[TYW];
  // prefetch stages
  T nq[PD][NV][TYW][SX], nxv[PD][NV][TYW][SX];
  T nv[PD][TYW][SX];
  T hn[PD][NV][HT][SX] = {};
  int2 mnext[PD][TYW];

  // prefill planes zb-R .. zb+R-1
#pragma unroll
  for (int r = 0; r < TYW; ++r) {
    const int j = j0 + wv + NW * r;
#pragma unroll
    for (int u = 0; u < NV; ++u)
#pragma unroll
      for (int s = 0; s < SX; ++s) q[u][r][s][0] = T(0);
#pragma unroll
    for (int dz = 1; dz <= 2 * R; ++dz) {
      const int2 m = win_at(wown[r], zb - R - 1 + dz);
#pragma unroll
      for (int u = 0; u < NV; ++u)
#pragma unroll
        for (int s = 0; s < SX; ++s) q[u][r][s][dz] = meta_val(yv[u], m, lane + 64 * s);
    }
  }
  // issue everything plane z needs into stage S
  auto prefetch = [&](auto stage, int z) {
    constexpr int S = decltype(stage)::value;
#pragma unroll
    for (int r = 0; r < TYW; ++r) {
      if (z + R >= wown[r].wb + 64) win_load(wown[r], j0 + wv + NW * r, z);
      const int2 mq = win_at(wown[r], z + R);
      const int2 mo = win_at(wown[r], z);
      mnext[S][r] = mo;
      const int lo = mo.y & 0xffff, hi = mo.y >> 16;
#pragma unroll
      for (int s = 0; s < SX; ++s) {
        const int i = lane + 64 * s;
        const bool in = i >= lo && i < hi;
        nv[S][r][s] = in ? V[mo.x + i] : T(0);
#pragma unroll
        for (int u = 0; u < NV; ++u) {
          nq[S][u][r][s] = meta_val(yv[u], mq, i);
          nxv[S][u][r][s] = (in && hasx) ? xv[u][mo.x + i] : T(0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < HT; ++t) {
      const int hr = wv + NW * t;
      if (hr < 2 * R) {
        const int tr = hr < R ? hr : hr + TY;
        if (z >= whalo[t].wb + 64) win_load(whalo[t], j0 - R + tr, z);
        const int2 mh = win_at(whalo[t], z);
#pragma unroll
        for (int u = 0; u < NV; ++u)
#pragma unroll
          for (int s = 0; s < SX; ++s) hn[S][u][t][s] = meta_val(yv[u], mh, lane + 64 * s);
      }
    }
  };
  // plane z from stage S: commit, refill the stage with plane z+PD, compute
  auto plane = [&](auto stage, int z) {
    constexpr int S = decltype(stage)::value;
    const int b = NB == 2 ? ((z - zb) & 1) : 0;
#pragma unroll
    for (int r = 0; r < TYW; ++r) {
      mcur[r] = mnext[S][r];
#pragma unroll
      for (int s = 0; s < SX; ++s) {
        vq[r][s] = nv[S][r][s];
#pragma unroll
        for (int u = 0; u < NV; ++u) {
#pragma unroll
          for (int d = 0; d < 2 * R; ++d) q[u][r][s][d] = q[u][r][s][d + 1];
          q[u][r][s][2 * R] = nq[S][u][r][s];
          xq[u][r][s] = nxv[S][u][r][s];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NV; ++u)
#pragma unroll
      for (int t = 0; t < HT; ++t)
#pragma unroll
        for (int s = 0; s < SX; ++s) hc[u][t][s] = hn[S][u][t][s];
    // issue the loads for plane z + PD
    if (z + PD < ze) prefetch(stage, z + PD);
    // plane z into LDS
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      T* tb = tile[b][u];
#pragma unroll
      for (int r = 0; r < TYW; ++r) {
        const int tr = wv + NW * r + R;
#pragma unroll
        for (int s = 0; s < SX; ++s) tb[tr * LW + R + lane + 64 * s] = q[u][r][s][R];
      }
#pragma unroll
      for (int t = 0; t < HT; ++t) {
        const int hr = wv + NW * t;
        if (hr < 2 * R) {
          const int tr = hr < R ? hr : hr + TY;
#pragma unroll
          for (int s = 0; s < SX; ++s) tb[tr * LW + R + lane + 64 * s] = hc[u][t][s];
        }
      }
    }
    lds_barrier();
    // outputs of plane z
#pragma unroll
    for (int r = 0; r < TYW; ++r) {
      const int2 mo = mcur[r];
      const int lo = mo.y & 0xffff, hi = mo.y >> 16;
      const int tr = wv + NW * r + R;
#pragma unroll
      for (int s = 0; s < SX; ++s) {
        const int i = lane + 64 * s;
        // computed on every lane (values outside the domain are zeros), stored inside it
        const bool in = i >= lo && i < hi;
        const T va = ca * vq[r][s] + cb;
#pragma unroll
        for (int u = 0; u < NV; ++u) {
          const T yc = q[u][r][s][R];
          T acc = T(g.w[0]) * yc;
          const T* trow = tile[b][u] + tr * LW + R + i;
#pragma unroll
          for (int d = 1; d <= R; ++d) {
            const T nb = trow[-d] + trow[d] + trow[-d * LW] + trow[d * LW] + q[u][r][s][R - d] +
                         q[u][r][s][R + d];
            acc += T(g.w[d]) * nb;
          }
          T o = acc + va * yc;
          if constexpr (HASX) o += cc * xq[u][r][s];
          if (in && act[u]) {
            ov[u][mo.x + i] = o;
            if constexpr (DOT) dot += (double)o * (double)yc;
          }
        }
      }
    }
    if (NB == 1) lds_barrier();   // one tile: the next plane overwrites it
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, PD - 1>;
  prefetch(S0{}, zb);
  if constexpr (PD == 2) {
    if (zb + 1 < ze) prefetch(S1{}, zb + 1);
    for (int z = zb; z < ze; z += 2) {
      plane(S0{}, z);
      if (z + 1 < ze) plane(S1{}, z + 1);
    }
  } else {
    for (int z = zb; z < ze; ++z) plane(S0{}, z);
  }
  if constexpr (DOT) {
    const double s = pc_block_sum(dot, red);
    if (threadIdx.x == 0) dot_part[L] = s;
  }
}


// ---- the lean z-march (one vector per workgroup, no partial sums) ------------
// The same computation as stencil_zmarch<T,R,SX,1,1,2,HASX,false,1,NW>, written
// to spend fewer vector instructions per plane:
//   * the z-queue is a ring of QL registers per x-chunk, indexed by the plane's
//     phase; the plane loop is unrolled QL times so the ring turns by renaming
//     instead of register moves (slot (z - zb + R) mod QL holds plane z);
//   * every global access is a raw buffer access on a per-vector resource whose
//     range is the vector's npts points: a lane outside its row's chord gets an
//     out-of-range offset, so its load returns 0 and its store is dropped, with
//     no branches around the memory instructions;
//   * X and V are read at unmasked offsets (their values only reach stored
//     lanes); X, V and the halo rows of plane z+2 go out before plane z is
//     computed, and the queue front of plane z-R+QL right after plane z's last
//     use of slot z-R (QL-2R-1 planes ahead of its use).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pc_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T pc_bload(__amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
}
template <typename T>
__device__ __forceinline__ void pc_bstore(__amdgpu_buffer_rsrc_t r, int off, T v) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
  } else {
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, 0, 0);
  }
}

template <int... I, class F>
__device__ __forceinline__ void pc_unroll_seq(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}

#ifndef PC_LEAN_NS
#define PC_LEAN_NS 3
#endif
#ifndef PC_LEAN_PF
#define PC_LEAN_PF 3
#endif
#ifndef PC_LEAN_WAVES
#define PC_LEAN_WAVES 0
#endif
// queue length: 2R+1 planes plus PC_LEAN_PF fronts in flight, rounded up to a
// multiple of 2*NS (tile parity and register sets repeat within the block)
constexpr int pc_lean_ql(int R) {
  return (2 * R + 1 + PC_LEAN_PF + 2 * PC_LEAN_NS - 1) / (2 * PC_LEAN_NS) * (2 * PC_LEAN_NS);
}

// Workgroup: TY y-rows of one vector; wave w owns row w % TY and the CPW
// x-chunks of 64 points starting at chunk (w / TY) * CPW, so a workgroup has
// TY * SX / CPW waves and each wave keeps only CPW chunks of queue registers.
// ZC: the slab's z-range is split into chunks of g.zc planes (few vectors);
// a separate instantiation, so the unchunked kernel keeps its registers.
template <typename T, int R, int SX, int TY, int CPW, bool HASX, bool ZC>
__global__ __launch_bounds__(64 * TY * SX / CPW)
#if PC_LEAN_WAVES > 0
__attribute__((amdgpu_waves_per_eu(PC_LEAN_WAVES, PC_LEAN_WAVES)))
#endif
void stencil_lean(StArgs g, const T* Y, const T* X, T* Out, const T* V, T ca, T cb, T cc) {
  constexpr int NWV = TY * SX / CPW;          // waves per workgroup
  constexpr int NXP = 64 * SX;
  constexpr int LW = NXP + 2 * R;
  constexpr int LR = TY + 2 * R;
  constexpr int NHI = 2 * R * (SX / CPW);     // halo items: (halo row, chunk group)
  constexpr int HT = (NHI + NWV - 1) / NWV;   // halo items per wave
  constexpr int NS = PC_LEAN_NS;
  constexpr int QL = pc_lean_ql(R);
  constexpr int ES = (int)sizeof(T);
  constexpr int OOB = 0x7ffffff0;     // beyond every range (npts * ES < 2^31 is checked on the host)
  static_assert(SX % CPW == 0, "chunks per wave must divide the chunks per row");
  __shared__ T tile[2][LR * LW];
  // row metadata of the slab's LR rows for the planes the march touches (the
  // launch sizes it for the grid's nz), read with uniform addresses every plane
  extern __shared__ int2 mtab[];

  const int bid = blockIdx.x;
  const int L = (bid & 7) * g.nb8 + (bid >> 3);
  if (L >= g.nblocks) return;
  // workgroup L: vector v, z-chunk zci, slab of TY rows (slab fastest)
  const int nvz = ZC ? g.nslab * g.nzc : g.nslab;
  const int v = L / nvz;
  const int zci = ZC ? (L - v * nvz) / g.nslab : 0;
  const int slab = L - v * nvz - zci * g.nslab;
  const int j0 = slab * TY;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ny = g.ny, nz = g.nz;

  int zb = nz, ze = 0;
  for (int t = 0; t < TY; ++t) {
    const int j = j0 + t;
    if (j < ny) {
      zb = min(zb, g.zrange[2 * j]);
      ze = max(ze, g.zrange[2 * j + 1]);
    }
  }
  // this workgroup's chunk of the slab's z-range; planes past its end are
  // computed by the march's last block but not stored (the next chunk's
  // workgroup stores them; see the store range below)
  if constexpr (ZC) {
    zb = max(zb, zci * g.zc);
    ze = min(ze, (zci + 1) * g.zc);
  }
  if (zb >= ze) return;
  const int zlo = zb - R;
  // planes zb-R .. zb-R + QL*(ceil((ze-zb)/QL) + 1) - 1: the launch sizes the
  // table for nz + 2 QL planes
  const int zm = (ze - zb + QL - 1) / QL * QL + QL;

  const unsigned nbytes = (unsigned)(g.npts * ES);
  const long long base = (long long)v * g.ld;
  const __amdgpu_buffer_rsrc_t ry = pc_rsrc(Y + base, nbytes);
  const __amdgpu_buffer_rsrc_t rx = pc_rsrc(HASX ? X + base : Y + base, nbytes);
  // Points are stored plane by plane, so every point of planes >= ze lies at
  // or past the first point of plane ze: a chunk's store range ends there and
  // the planes its march computes past its end are dropped by the range check.
  unsigned sbytes = nbytes;
  if constexpr (ZC) sbytes = (unsigned)((ze < nz ? (long long)g.off[ny * ze] : g.npts) * ES);
  const __amdgpu_buffer_rsrc_t ro = pc_rsrc(Out + base, sbytes);
  const __amdgpu_buffer_rsrc_t rv = pc_rsrc(V, nbytes);

  for (int t = threadIdx.x; t < 2 * LR * LW; t += 64 * NWV) (&tile[0][0])[t] = T(0);
  for (int e = threadIdx.x; e < LR * zm; e += 64 * NWV) {
    const int t = e / zm, z = zlo + (e - t * zm);
    const int j = j0 - R + t;
    mtab[e] = (j >= 0 && j < ny && z >= 0 && z < nz) ? g.meta[j + ny * z] : make_int2(0, 0);
  }
  __syncthreads();

  // metadata of slab row t (0 .. LR-1, own rows at R .. R+TY-1) at plane z, as scalars
  auto mrow = [&](int t, int z) -> int2 {
    const int2 m = mtab[t * zm + (z - zlo)];
    return make_int2(__builtin_amdgcn_readfirstlane(m.x), __builtin_amdgcn_readfirstlane(m.y));
  };
  // byte offset of x-chunk c in a row with metadata m: masked (out of range
  // outside the chord) or plain (any in-range or out-of-range address)
  auto off_masked = [&](int2 m, int c) -> int {
    const int i = lane + 64 * c;
    const int lo = m.y & 0xffff, len = (m.y >> 16) - lo;
    return (unsigned)(i - lo) < (unsigned)len ? (m.x + i) * ES : OOB;
  };
  auto off_plain = [&](int2 m, int c) -> int { return (m.x + lane + 64 * c) * ES; };

  const int town = R + wv % TY;               // own row in the tile
  const int c0 = (wv / TY) * CPW;              // first own chunk
  int th[HT], tc[HT];                          // halo items: tile row and first chunk
#pragma unroll
  for (int u = 0; u < HT; ++u) {
    const int k = wv + NWV * u;
    const int hr = k % (2 * R);
    th[u] = hr < R ? hr : hr + TY;
    tc[u] = (k / (2 * R)) * CPW;
  }
  auto has_halo = [&](int u) { return wv + NWV * u < NHI; };

  T q[CPW][QL];
  T xq[NS][CPW], vq[NS][CPW], hq[NS][HT][CPW];
  int so[NS][CPW];
  // prefill: planes zb-R .. zb-R+QL-1 into slots 0 .. QL-1 (slot of plane p: (p - zb + R) mod QL)
#pragma unroll
  for (int d = 0; d < QL; ++d) {
    const int2 m = mrow(town, zb - R + d);
#pragma unroll
    for (int c = 0; c < CPW; ++c) q[c][d] = pc_bload<T>(ry, off_masked(m, c0 + c));
  }
  // X, V, halo rows and store offsets of plane z into set S
  auto issue_plane = [&](auto set, int z) {
    constexpr int S = decltype(set)::value;
    const int2 mo = mrow(town, z);
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const int po = off_plain(mo, c0 + c);
      so[S][c] = off_masked(mo, c0 + c);
      vq[S][c] = pc_bload<T>(rv, po);
      if constexpr (HASX) xq[S][c] = pc_bload<T>(rx, po);
    }
#pragma unroll
    for (int u = 0; u < HT; ++u) {
      if (has_halo(u)) {
        const int2 mh = mrow(th[u], z);
#pragma unroll
        for (int c = 0; c < CPW; ++c) hq[S][u][c] = pc_bload<T>(ry, off_masked(mh, tc[u] + c));
      }
    }
  };
  pc_unroll_seq(std::make_integer_sequence<int, NS - 1>{}, [&](auto k) {
    constexpr int K = decltype(k)::value;
    issue_plane(std::integral_constant<int, K>{}, zb + K);
  });

  const T w0 = T(g.w[0]);
  T wd[R + 1];
#pragma unroll
  for (int d = 1; d <= R; ++d) wd[d] = T(g.w[d]);

  // Planes run in whole blocks of QL, past ze up to the next multiple of QL:
  // there the slab's rows are empty, so those planes' stores are all out of
  // range, and the loads they issue read zeros or stay inside the vectors.
  // No branch separates a load from its use, which keeps the compiler from
  // sinking the prefetches to their uses.
  auto plane = [&](auto phase, int z) {
    constexpr int P = decltype(phase)::value;
    constexpr int S = P % NS;      // (z - zb) mod NS: the register set of plane z
    constexpr int B = P & 1;       // LDS tile of this plane
    issue_plane(std::integral_constant<int, (P + NS - 1) % NS>{}, z + NS - 1);
    __builtin_amdgcn_sched_barrier(0);
    T* tb = tile[B];
#pragma unroll
    for (int c = 0; c < CPW; ++c) tb[town * LW + R + lane + 64 * (c0 + c)] = q[c][(P + R) % QL];
#pragma unroll
    for (int u = 0; u < HT; ++u) {
      if (has_halo(u)) {
#pragma unroll
        for (int c = 0; c < CPW; ++c) tb[th[u] * LW + R + lane + 64 * (tc[u] + c)] = hq[S][u][c];
      }
    }
    lds_barrier();
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const T* trow = tb + town * LW + R + lane + 64 * (c0 + c);
      const T yc = q[c][(P + R) % QL];
      T acc = w0 * yc;
#pragma unroll
      for (int d = 1; d <= R; ++d) {
        const T nb = trow[-d] + trow[d] + trow[-d * LW] + trow[d * LW] + q[c][(P + R - d) % QL] +
                     q[c][(P + R + d) % QL];
        acc += wd[d] * nb;
      }
      T o = acc + (ca * vq[S][c] + cb) * yc;
      if constexpr (HASX) o += cc * xq[S][c];
      pc_bstore<T>(ro, so[S][c], o);
    }
    // slot P (plane z-R) is free: plane z-R+QL, the queue front of plane z-2R+QL
    {
      const int2 m = mrow(town, z - R + QL);
#pragma unroll
      for (int c = 0; c < CPW; ++c) q[c][P] = pc_bload<T>(ry, off_masked(m, c0 + c));
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  for (int z0 = zb; z0 < ze; z0 += QL) {
    pc_unroll_seq(std::make_integer_sequence<int, QL>{}, [&](auto ph) {
      constexpr int P = decltype(ph)::value;
      plane(ph, z0 + P);
    });
  }
}

#ifndef PC_LEAN_TY
#define PC_LEAN_TY 8
#endif

// z-chunks for the lean and paired kernels: with few vectors the slab x vector
// workgroups do not fill the chip (256 CUs x 32 wavefronts), so each slab's
// z-range is split until the wavefronts would fill it once, keeping at least 8
// planes per chunk (a chunk re-reads 2R queue planes of its neighbour and runs
// a variant with a few more registers).  Measured (FP32, radius 4): H2O sphere
// (69,264 points, k 23, 8 waves per workgroup) 30 -> 20 us; C10H8 sphere
// (k 49, 16 waves per workgroup: 9,408 wavefronts already) 134 us unchunked,
// 159-164 us when chunked to 1024-2048 workgroups.
#ifndef PC_LEAN_ZCHUNK_WAVES
#define PC_LEAN_ZCHUNK_WAVES 8192
#endif
#ifndef PC_LEAN_FORCE_ZC
#define PC_LEAN_FORCE_ZC 0
#endif
inline void pc_zchunks(StArgs& a, int nz, int waves_per_wg) {
  constexpr long long kWaves = PC_LEAN_ZCHUNK_WAVES;
  const long long base = (long long)a.nslab * a.nvec * waves_per_wg;
  int nzc = 1;
  if (base < kWaves) nzc = (int)std::min<long long>((kWaves + base - 1) / base, std::max(1, nz / 8));
  a.zc = (nz + nzc - 1) / nzc;
  if (PC_LEAN_FORCE_ZC > 0) a.zc = std::min(nz, PC_LEAN_FORCE_ZC);
  a.nzc = (nz + a.zc - 1) / a.zc;
}

template <typename T, int R, int SX>
int launch_lean(const GridDev* gd, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld, double scale,
                double ca, double cb, double cc, hipStream_t st) {
  constexpr int TY = PC_LEAN_TY;
  constexpr int CPW = SX > 2 ? SX / 2 : 1;     // at most 16 waves (1024 threads) per workgroup
  constexpr int QL = pc_lean_ql(R);
  constexpr int WPG = TY * SX / CPW;
  StArgs a;
  a.nx = gd->nx; a.ny = gd->ny; a.nz = gd->nz; a.npts = gd->npts;
  a.lo = gd->lo; a.hi = gd->hi; a.off = gd->off; a.zrange = gd->zrange; a.meta = gd->meta;
  a.ld = ld;
  a.nslab = (gd->ny + TY - 1) / TY;
  a.nvec = nvec;
  pc_zchunks(a, gd->nz, WPG);
  a.nblocks = a.nslab * a.nzc * nvec;
  a.nb8 = (a.nblocks + 7) / 8;
  for (int d = 0; d < 8; ++d) a.w[d] = 0.0;
  a.w[0] = 3.0 * gd->w[0] * scale;
  for (int d = 1; d <= R; ++d) a.w[d] = gd->w[d] * scale;
  const dim3 grid(8 * a.nb8), block(64 * TY * SX / CPW);
  const size_t mbytes = sizeof(int2) * (size_t)(TY + 2 * R) * (size_t)(gd->nz + 2 * QL);
  if (a.nzc > 1) {
    if (X)
      hipLaunchKernelGGL((stencil_lean<T, R, SX, TY, CPW, true, true>), grid, block, mbytes, st, a, Y, X, Out, V,
                         T(ca), T(cb), T(cc));
    else
      hipLaunchKernelGGL((stencil_lean<T, R, SX, TY, CPW, false, true>), grid, block, mbytes, st, a, Y, X, Out, V,
                         T(ca), T(cb), T(cc));
  } else if (X) {
    hipLaunchKernelGGL((stencil_lean<T, R, SX, TY, CPW, true, false>), grid, block, mbytes, st, a, Y, X, Out, V,
                       T(ca), T(cb), T(cc));
  } else {
    hipLaunchKernelGGL((stencil_lean<T, R, SX, TY, CPW, false, false>), grid, block, mbytes, st, a, Y, X, Out, V,
                       T(ca), T(cb), T(cc));
  }
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// ---- the compacted-column z-march (one vector per workgroup) ----------------
// stencil_lean gives each wavefront a fixed 64-point x-chunk of one row, so on a
// sphere about half of its lanes hold no point.  Here the lanes of a workgroup
// are dealt the slab's actual columns: the host lists, per slab of TY rows,
// the columns (tile row, x) that any plane of the slab's march touches, row
// after row (own rows first, then the 2R halo rows' columns for the halo
// loads), and a workgroup has as many wavefronts as the larger list needs.
// Each lane keeps the z-queue of its own column, loads one halo column per
// plane, and reads its row metadata per plane from the LDS table (per lane:
// the rows of one wavefront differ).  The arithmetic and its order are those
// of stencil_lean, so the results are the same bits.
constexpr int kCcTY = 8;

template <typename T, int R, int NXP, bool HASX>
__global__ __launch_bounds__(1024) void stencil_cc(StArgs g, const int* __restrict__ ctab, int wmax, const T* Y,
                                                   const T* X, T* Out, const T* V, T ca, T cb, T cc) {
  constexpr int TY = kCcTY;
  constexpr int LW = NXP + 2 * R;
  constexpr int LR = TY + 2 * R;
  constexpr int NS = PC_LEAN_NS;
  constexpr int QL = pc_lean_ql(R);
  constexpr int ES = (int)sizeof(T);
  constexpr int OOB = 0x7ffffff0;
  __shared__ T tile[2][LR * LW];
  extern __shared__ int2 mtab[];

  const int bid = blockIdx.x;
  const int L = (bid & 7) * g.nb8 + (bid >> 3);
  if (L >= g.nblocks) return;
  const int v = L / g.nslab;
  const int slab = L - v * g.nslab;
  const int j0 = slab * TY;
  const int ny = g.ny, nz = g.nz;
  int zb = nz, ze = 0;
  for (int t = 0; t < TY; ++t) {
    const int j = j0 + t;
    if (j < ny) {
      zb = min(zb, g.zrange[2 * j]);
      ze = max(ze, g.zrange[2 * j + 1]);
    }
  }
  if (zb >= ze) return;
  const int zlo = zb - R;
  const int zm = (ze - zb + QL - 1) / QL * QL + QL;

  const unsigned nbytes = (unsigned)(g.npts * ES);
  const long long base = (long long)v * g.ld;
  const __amdgpu_buffer_rsrc_t ry = pc_rsrc(Y + base, nbytes);
  const __amdgpu_buffer_rsrc_t rx = pc_rsrc(HASX ? X + base : Y + base, nbytes);
  const __amdgpu_buffer_rsrc_t ro = pc_rsrc(Out + base, nbytes);
  const __amdgpu_buffer_rsrc_t rv = pc_rsrc(V, nbytes);

  // this lane's own column (tile row tr, x) and halo column (tile row th, x)
  const int* cs = ctab + (long long)slab * 2 * wmax * 64;
  const int co = cs[threadIdx.x], ch = cs[wmax * 64 + threadIdx.x];
  const bool own = co >= 0, hal = ch >= 0;
  const int tr = own ? (co >> 16) : R, xo = own ? (co & 0xffff) : 0;
  const int th = hal ? (ch >> 16) : 0, xh = hal ? (ch & 0xffff) : 0;

  for (int t = threadIdx.x; t < 2 * LR * LW; t += blockDim.x) (&tile[0][0])[t] = T(0);
  for (int e = threadIdx.x; e < LR * zm; e += blockDim.x) {
    const int t = e / zm, z = zlo + (e - t * zm);
    const int j = j0 - R + t;
    mtab[e] = (j >= 0 && j < ny && z >= 0 && z < nz) ? g.meta[j + ny * z] : make_int2(0, 0);
  }
  __syncthreads();

  auto mrow = [&](int t, int z) -> int2 { return mtab[t * zm + (z - zlo)]; };
  auto off_masked = [&](int2 m, int x, bool ok) -> int {
    const int lo = m.y & 0xffff, len = (m.y >> 16) - lo;
    return (ok && (unsigned)(x - lo) < (unsigned)len) ? (m.x + x) * ES : OOB;
  };
  auto off_plain = [&](int2 m, int x, bool ok) -> int { return ok ? (m.x + x) * ES : OOB; };

  T q[QL];
  T xq[NS], vq[NS], hq[NS];
  int so[NS];
#pragma unroll
  for (int d = 0; d < QL; ++d) q[d] = pc_bload<T>(ry, off_masked(mrow(tr, zb - R + d), xo, own));
  auto issue_plane = [&](auto set, int z) {
    constexpr int S = decltype(set)::value;
    const int2 mo = mrow(tr, z);
    const int po = off_plain(mo, xo, own);
    so[S] = off_masked(mo, xo, own);
    vq[S] = pc_bload<T>(rv, po);
    if constexpr (HASX) xq[S] = pc_bload<T>(rx, po);
    hq[S] = pc_bload<T>(ry, off_masked(mrow(th, z), xh, hal));
  };
  pc_unroll_seq(std::make_integer_sequence<int, NS - 1>{}, [&](auto k) {
    constexpr int K = decltype(k)::value;
    issue_plane(std::integral_constant<int, K>{}, zb + K);
  });

  const T w0 = T(g.w[0]);
  T wd[R + 1];
#pragma unroll
  for (int d = 1; d <= R; ++d) wd[d] = T(g.w[d]);

  auto plane = [&](auto phase, int z) {
    constexpr int P = decltype(phase)::value;
    constexpr int S = P % NS;
    constexpr int B = P & 1;
    issue_plane(std::integral_constant<int, (P + NS - 1) % NS>{}, z + NS - 1);
    __builtin_amdgcn_sched_barrier(0);
    T* tb = tile[B];
    if (own) tb[tr * LW + R + xo] = q[(P + R) % QL];
    if (hal) tb[th * LW + R + xh] = hq[S];
    lds_barrier();
    const T* trow = tb + tr * LW + R + xo;
    const T yc = q[(P + R) % QL];
    T acc = w0 * yc;
#pragma unroll
    for (int d = 1; d <= R; ++d) {
      const T nb = trow[-d] + trow[d] + trow[-d * LW] + trow[d * LW] + q[(P + R - d) % QL] + q[(P + R + d) % QL];
      acc += wd[d] * nb;
    }
    T o = acc + (ca * vq[S] + cb) * yc;
    if constexpr (HASX) o += cc * xq[S];
    pc_bstore<T>(ro, so[S], o);
    q[P] = pc_bload<T>(ry, off_masked(mrow(tr, z - R + QL), xo, own));
    __builtin_amdgcn_sched_barrier(0);
  };
  for (int z0 = zb; z0 < ze; z0 += QL) {
    pc_unroll_seq(std::make_integer_sequence<int, QL>{}, [&](auto ph) {
      constexpr int P = decltype(ph)::value;
      plane(ph, z0 + P);
    });
  }
}

// pc_stencil_set_variant: 0 = stencil_lean, 1 = stencil_cc, 2 = stencil_cc for
// FP32 launches of fewer than 128 vectors that stencil_lean would not split
// into z-chunks (the default).  Measured (FP32 step, radius 4, same bits):
// C10H8 sphere (nx 92, 49 vectors) 118 us against 133 us; Si87H76 sphere
// (nx 106, 275 vectors) 824 against 787 us; H2O sphere (nx 52, 23 vectors,
// z-chunked by stencil_lean) 30 against 18 us.
int g_stencil_cc = 2;

template <typename T, int R>
int launch_cc(const GridDev* gd, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld, double scale,
              double ca, double cb, double cc, hipStream_t st) {
  constexpr int TY = kCcTY;
  constexpr int QL = pc_lean_ql(R);
  StArgs a;
  a.nx = gd->nx; a.ny = gd->ny; a.nz = gd->nz; a.npts = gd->npts;
  a.lo = gd->lo; a.hi = gd->hi; a.off = gd->off; a.zrange = gd->zrange; a.meta = gd->meta;
  a.ld = ld;
  a.nslab = (gd->ny + TY - 1) / TY;
  a.nvec = nvec;
  a.nzc = 1; a.zc = gd->nz;
  a.nblocks = a.nslab * nvec;
  a.nb8 = (a.nblocks + 7) / 8;
  for (int d = 0; d < 8; ++d) a.w[d] = 0.0;
  a.w[0] = 3.0 * gd->w[0] * scale;
  for (int d = 1; d <= R; ++d) a.w[d] = gd->w[d] * scale;
  const dim3 grid(8 * a.nb8), block(64 * gd->cc_wmax);
  const size_t mbytes = sizeof(int2) * (size_t)(TY + 2 * R) * (size_t)(gd->nz + 2 * QL);
  if (gd->nx <= 64) {
    if (X) hipLaunchKernelGGL((stencil_cc<T, R, 64, true>), grid, block, mbytes, st, a, gd->cc_tab, gd->cc_wmax, Y, X, Out, V, T(ca), T(cb), T(cc));
    else hipLaunchKernelGGL((stencil_cc<T, R, 64, false>), grid, block, mbytes, st, a, gd->cc_tab, gd->cc_wmax, Y, X, Out, V, T(ca), T(cb), T(cc));
  } else {
    if (X) hipLaunchKernelGGL((stencil_cc<T, R, 128, true>), grid, block, mbytes, st, a, gd->cc_tab, gd->cc_wmax, Y, X, Out, V, T(ca), T(cb), T(cc));
    else hipLaunchKernelGGL((stencil_cc<T, R, 128, false>), grid, block, mbytes, st, a, gd->cc_tab, gd->cc_wmax, Y, X, Out, V, T(ca), T(cb), T(cc));
  }
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// ---- the paired z-march (FP32, rows of 65..128 points) ---------------------
// stencil_lean with two adjacent x points per lane: lane l holds x = 2l, 2l+1,
// so one wavefront covers a whole row and every Y / X / V access and the store
// move 8 bytes per lane (64-bit buffer accesses) instead of 4.  The row data
// starts at any element, so a pair straddles the chord ends: a pair with at
// least one point in the chord is loaded whole and its outside point is set to
// zero at first use (queue entries: when the plane enters the stencil as z+R;
// halo rows: before they go to LDS), and a store writes the pair when both
// points are in the chord, else the one point that is (a 32-bit store).  The
// two pairs that would straddle the ends of the array (flat start -1 or
// npts-1) are loaded from 0 or npts-2 and shifted into place, so every 64-bit
// access lies wholly inside the vector.
typedef float pc_f2 __attribute__((ext_vector_type(2)));
typedef unsigned pc_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ pc_f2 pc_bload2(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(pc_f2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

template <int R, int TY, bool HASX>
__global__ __launch_bounds__(64 * TY)
void stencil_pair(StArgs g, const float* Y, const float* X, float* Out, const float* V, float ca, float cb, float cc) {
  using T = float;
  constexpr int NWV = TY;                     // one wave per row
  constexpr int LW = 128 + 2 * R;
  constexpr int LR = TY + 2 * R;
  constexpr int NHI = 2 * R;                  // halo rows
  constexpr int HT = (NHI + NWV - 1) / NWV;
  constexpr int NS = PC_LEAN_NS;
  constexpr int QL = pc_lean_ql(R);
  constexpr int ES = 4;
  constexpr int OOB = 0x7ffffff0;
  __shared__ __align__(16) T tile[2][LR * LW];
  extern __shared__ int2 mtab[];

  const int bid = blockIdx.x;
  const int L = (bid & 7) * g.nb8 + (bid >> 3);
  if (L >= g.nblocks) return;
  // workgroup L: vector v, z-chunk zci, slab of TY rows (slab fastest)
  const int nvz = g.nslab * g.nzc;
  const int v = L / nvz;
  const int zci = (L - v * nvz) / g.nslab;
  const int slab = L - v * nvz - zci * g.nslab;
  const int j0 = slab * TY;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ny = g.ny, nz = g.nz;

  int zb = nz, ze = 0;
  for (int t = 0; t < TY; ++t) {
    const int j = j0 + t;
    if (j < ny) {
      zb = min(zb, g.zrange[2 * j]);
      ze = max(ze, g.zrange[2 * j + 1]);
    }
  }
  zb = max(zb, zci * g.zc);
  ze = min(ze, (zci + 1) * g.zc);
  if (zb >= ze) return;
  const int zlo = zb - R;
  const int zm = (ze - zb + QL - 1) / QL * QL + QL;

  const unsigned nbytes = (unsigned)(g.npts * ES);
  const long long base = (long long)v * g.ld;
  const __amdgpu_buffer_rsrc_t ry = pc_rsrc(Y + base, nbytes);
  const __amdgpu_buffer_rsrc_t rx = pc_rsrc(HASX ? X + base : Y + base, nbytes);
  const __amdgpu_buffer_rsrc_t ro = pc_rsrc(Out + base, nbytes);
  const __amdgpu_buffer_rsrc_t rv = pc_rsrc(V, nbytes);

  for (int t = threadIdx.x; t < 2 * LR * LW; t += 64 * NWV) (&tile[0][0])[t] = T(0);
  for (int e = threadIdx.x; e < LR * zm; e += 64 * NWV) {
    const int t = e / zm, z = zlo + (e - t * zm);
    const int j = j0 - R + t;
    mtab[e] = (j >= 0 && j < ny && z >= 0 && z < nz) ? g.meta[j + ny * z] : make_int2(0, 0);
  }
  __syncthreads();

  auto mrow = [&](int t, int z) -> int2 {
    const int2 m = mtab[t * zm + (z - zlo)];
    return make_int2(__builtin_amdgcn_readfirstlane(m.x), __builtin_amdgcn_readfirstlane(m.y));
  };
  const int i0 = 2 * lane;
  const int last2 = (int)g.npts - 2;           // the host checks npts >= 2
  // flat start of the pair (m.x + i0) clamped to [0, npts-2]
  auto pair_start = [&](int2 m) -> int { return min(max(m.x + i0, 0), last2); };
  // byte offset of the pair: out of range unless a point is in the chord
  auto off_pair = [&](int2 m) -> int {
    const int lo = m.y & 0xffff, len = (m.y >> 16) - lo;
    const bool any = (unsigned)(i0 - lo) < (unsigned)len || (unsigned)(i0 + 1 - lo) < (unsigned)len;
    return any ? pair_start(m) * ES : OOB;
  };
  // the pair loaded from pair_start(m), shifted back to x = i0, i0+1, with the
  // points outside the chord set to zero
  auto fix_pair = [&](int2 m, pc_f2 a) -> pc_f2 {
    const int start = m.x + i0, st = pair_start(m);
    const int lo = m.y & 0xffff, len = (m.y >> 16) - lo;
    T x0 = start > st ? a.y : (start < st ? T(0) : a.x);
    T x1 = start < st ? a.x : (start > st ? T(0) : a.y);
    pc_f2 r;
    r.x = (unsigned)(i0 - lo) < (unsigned)len ? x0 : T(0);
    r.y = (unsigned)(i0 + 1 - lo) < (unsigned)len ? x1 : T(0);
    return r;
  };

  const int town = R + wv;
  int th[HT];
#pragma unroll
  for (int u = 0; u < HT; ++u) {
    const int k = wv + NWV * u;
    th[u] = k < R ? k : k + TY;
  }
  auto has_halo = [&](int u) { return wv + NWV * u < NHI; };

  pc_f2 q[QL];
  pc_f2 xq[NS], vq[NS], hq[NS][HT];
  // prefill planes zb-R .. zb-R+QL-1; the 2R planes zb-R .. zb+R-1 are used
  // before the march masks anything, so they are masked here
#pragma unroll
  for (int d = 0; d < QL; ++d) q[d] = pc_bload2(ry, off_pair(mrow(town, zb - R + d)));
#pragma unroll
  for (int d = 0; d < 2 * R; ++d) q[d] = fix_pair(mrow(town, zb - R + d), q[d]);

  auto issue_plane = [&](auto set, int z) {
    constexpr int S = decltype(set)::value;
    const int po = pair_start(mrow(town, z)) * ES;
    vq[S] = pc_bload2(rv, po);
    if constexpr (HASX) xq[S] = pc_bload2(rx, po);
#pragma unroll
    for (int u = 0; u < HT; ++u)
      if (has_halo(u)) hq[S][u] = pc_bload2(ry, off_pair(mrow(th[u], z)));
  };
  pc_unroll_seq(std::make_integer_sequence<int, NS - 1>{}, [&](auto k) {
    constexpr int K = decltype(k)::value;
    issue_plane(std::integral_constant<int, K>{}, zb + K);
  });

  const T w0 = T(g.w[0]);
  T wd[R + 1];
#pragma unroll
  for (int d = 1; d <= R; ++d) wd[d] = T(g.w[d]);

  auto plane = [&](auto phase, int z) {
    constexpr int P = decltype(phase)::value;
    constexpr int S = P % NS;
    constexpr int B = P & 1;
    issue_plane(std::integral_constant<int, (P + NS - 1) % NS>{}, z + NS - 1);
    __builtin_amdgcn_sched_barrier(0);
    // plane z+R enters the stencil: shift and zero its points outside the chord
    q[(P + 2 * R) % QL] = fix_pair(mrow(town, z + R), q[(P + 2 * R) % QL]);
    T* tb = tile[B];
    {
      const pc_f2 c = q[(P + R) % QL];
      tb[town * LW + R + i0] = c.x;
      tb[town * LW + R + i0 + 1] = c.y;
    }
#pragma unroll
    for (int u = 0; u < HT; ++u) {
      if (has_halo(u)) {
        const pc_f2 h = fix_pair(mrow(th[u], z), hq[S][u]);
        tb[th[u] * LW + R + i0] = h.x;
        tb[th[u] * LW + R + i0 + 1] = h.y;
      }
    }
    lds_barrier();
    const int2 mo = mrow(town, z);
    const T* trow = tb + town * LW + R + i0;
    T xs[2 * R + 2];                           // own row at x = i0-R .. i0+R+1
#pragma unroll
    for (int d = 0; d < 2 * R + 2; ++d) xs[d] = trow[d - R];
    const pc_f2 yc = q[(P + R) % QL];
    pc_f2 acc = w0 * yc;
#pragma unroll
    for (int d = 1; d <= R; ++d) {
      // the summation order of stencil_lean: x-, x+, y-, y+, z-, z+
      const pc_f2 zm = q[(P + R - d) % QL], zp = q[(P + R + d) % QL];
      pc_f2 nb;
      nb.x = xs[R - d] + xs[R + d] + trow[-d * LW] + trow[d * LW] + zm.x + zp.x;
      nb.y = xs[R + 1 - d] + xs[R + 1 + d] + trow[-d * LW + 1] + trow[d * LW + 1] + zm.y + zp.y;
      acc += wd[d] * nb;
    }
    pc_f2 o = acc + (ca * fix_pair(mo, vq[S]) + cb) * yc;
    if constexpr (HASX) o += cc * fix_pair(mo, xq[S]);
    {
      const int lo = mo.y & 0xffff, len = (mo.y >> 16) - lo;
      const bool inz = z < ze;
      const bool in0 = inz && (unsigned)(i0 - lo) < (unsigned)len;
      const bool in1 = inz && (unsigned)(i0 + 1 - lo) < (unsigned)len;
      const int o2 = (in0 && in1) ? (mo.x + i0) * ES : OOB;
      const int o1 = (in0 != in1) ? (mo.x + i0 + (in0 ? 0 : 1)) * ES : OOB;
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(pc_u2, o), ro, o2, 0, 0);
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, in0 ? o.x : o.y), ro, o1, 0, 0);
    }
    q[P] = pc_bload2(ry, off_pair(mrow(town, z - R + QL)));
    __builtin_amdgcn_sched_barrier(0);
  };
  for (int z0 = zb; z0 < ze; z0 += QL) {
    pc_unroll_seq(std::make_integer_sequence<int, QL>{}, [&](auto ph) {
      constexpr int P = decltype(ph)::value;
      plane(ph, z0 + P);
    });
  }
}

#ifndef PC_STENCIL_PAIR
#define PC_STENCIL_PAIR 1
#endif

template <int R>
int launch_pair(const GridDev* gd, const float* Y, const float* X, float* Out, const float* V, int nvec, long long ld,
                double scale, double ca, double cb, double cc, hipStream_t st) {
  constexpr int TY = PC_LEAN_TY;
  constexpr int QL = pc_lean_ql(R);
  constexpr int WPG = TY;
  StArgs a;
  a.nx = gd->nx; a.ny = gd->ny; a.nz = gd->nz; a.npts = gd->npts;
  a.lo = gd->lo; a.hi = gd->hi; a.off = gd->off; a.zrange = gd->zrange; a.meta = gd->meta;
  a.ld = ld;
  a.nslab = (gd->ny + TY - 1) / TY;
  a.nvec = nvec;
  pc_zchunks(a, gd->nz, WPG);
  a.nblocks = a.nslab * a.nzc * nvec;
  a.nb8 = (a.nblocks + 7) / 8;
  for (int d = 0; d < 8; ++d) a.w[d] = 0.0;
  a.w[0] = 3.0 * gd->w[0] * scale;
  for (int d = 1; d <= R; ++d) a.w[d] = gd->w[d] * scale;
  const dim3 grid(8 * a.nb8), block(64 * TY);
  const size_t mbytes = sizeof(int2) * (size_t)(TY + 2 * R) * (size_t)(gd->nz + 2 * QL);
  if (X)
    hipLaunchKernelGGL((stencil_pair<R, TY, true>), grid, block, mbytes, st, a, Y, X, Out, V, float(ca), float(cb),
                       float(cc));
  else
    hipLaunchKernelGGL((stencil_pair<R, TY, false>), grid, block, mbytes, st, a, Y, X, Out, V, float(ca), float(cb),
                       float(cc));
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

template <typename T, int R, int SX, int TYW, int NV, int NB = 2, int NW = 4>
constexpr int zmarch_lds_bytes() {
  return NB * NV * (NW * TYW + 2 * R) * (64 * SX + 2 * R) * (int)sizeof(T);
}

template <typename T, int R, int SX, int TYW, int NV, int NB = 2, int PD = 1, int NW = 4>
int launch_cfg(const GridDev* gd, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld, double scale,
               double ca, double cb, double cc, double* dot_part, hipStream_t st) {
  constexpr int TY = NW * TYW;
  StArgs a;
  a.nx = gd->nx; a.ny = gd->ny; a.nz = gd->nz; a.npts = gd->npts;
  a.lo = gd->lo; a.hi = gd->hi; a.off = gd->off; a.zrange = gd->zrange; a.meta = gd->meta;
  a.ld = ld;
  a.nslab = (gd->ny + TY - 1) / TY;
  a.nvec = nvec;
  a.nzc = 1;
  a.zc = gd->nz;
  a.nblocks = a.nslab * ((nvec + NV - 1) / NV);
  a.nb8 = (a.nblocks + 7) / 8;
  for (int d = 0; d < 8; ++d) a.w[d] = 0.0;
  a.w[0] = 3.0 * gd->w[0] * scale;
  for (int d = 1; d <= R; ++d) a.w[d] = gd->w[d] * scale;
  const dim3 grid(8 * a.nb8), block(64 * NW);
  if (dot_part) {
    if constexpr (NV == 1) {
      if (X)
        hipLaunchKernelGGL((stencil_zmarch<T, R, SX, TYW, NV, NB, true, true, PD, NW>), grid, block, 0, st, a, Y, X,
                           Out, V, T(ca), T(cb), T(cc), dot_part);
      else
        hipLaunchKernelGGL((stencil_zmarch<T, R, SX, TYW, NV, NB, false, true, PD, NW>), grid, block, 0, st, a, Y, X,
                           Out, V, T(ca), T(cb), T(cc), dot_part);
    } else {
      pc::set_error("stencil: partial sums need one vector per workgroup");
      return pc::PC_ERR_ARG;
    }
  } else if (X) {
    hipLaunchKernelGGL((stencil_zmarch<T, R, SX, TYW, NV, NB, true, false, PD, NW>), grid, block, 0, st, a, Y, X, Out,
                       V, T(ca), T(cb), T(cc), dot_part);
  } else {
    hipLaunchKernelGGL((stencil_zmarch<T, R, SX, TYW, NV, NB, false, false, PD, NW>), grid, block, 0, st, a, Y, X, Out,
                       V, T(ca), T(cb), T(cc), dot_part);
  }
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

#ifndef PC_STENCIL_NW
#define PC_STENCIL_NW 8
#endif
#ifndef PC_STENCIL_NV
#define PC_STENCIL_NV 1
#endif

// One y-row per wavefront (TYW = 1), PC_STENCIL_NW = 8 wavefronts (y-rows) and
// PC_STENCIL_NV = 1 vector per workgroup where the LDS tiles fit in 56 KB, else
// 4 wavefronts.  Measured on Si87H76 (one recurrence step, 275 vectors):
//   4 rows: FP32 0.95 ms with 2 vectors, 0.96 ms with 1, 1.04 ms with 4;
//           FP64 1.72 ms with 2, 1.79 ms with 1;
//   8 rows: FP32 0.91 ms with 2 vectors (92 VGPRs, 5 waves per SIMD), 0.84 ms
//           with 1 (60 VGPRs, 8 waves per SIMD); FP64 1.65 ms with 1;
//   16 rows, 1 vector: FP32 0.84 ms, FP64 1.66 ms.
// A prefetch depth of two planes (PD = 2) was slower (1.22 ms FP32, 4 vectors):
// the registers it takes cost more occupancy than the extra plane gains.
template <typename T, int R, int SX>
int launch_r_sx(const GridDev* gd, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld, double scale,
                double ca, double cb, double cc, double* dot_part, hipStream_t st) {
  constexpr int kLds = 56 * 1024;
  constexpr int NW = PC_STENCIL_NW;
  if (dot_part)   // per-vector partial sums need one vector per workgroup
    return launch_cfg<T, R, SX, 1, 1, 2, 1, 4>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
#if PC_STENCIL_PAIR
  // Measured on MI355X (one FP32 recurrence step, lean vs paired, identical
  // results): Si87H76 sphere nx 106, 275 vectors, radius 4: 801 vs 881 us;
  // radius 3: 716 vs 880 us; C10H8 sphere nx 92, 49 vectors: 133 vs 178 us;
  // sphere nx 75, radius 1: 46 vs 72 us; 128^3 sphere, 256 vectors, radius 6:
  // 1539 vs 1430 us.  The paired kernel needs 89 VGPRs at radius 4 (5 waves
  // per SIMD against 8), which costs more than the wider accesses gain,
  // except at radius 6, where the lean kernel's queue is long too.
  if constexpr (SX == 2 && sizeof(T) == 4 && R >= 6)
    if (gd->npts >= 2)
      return launch_pair<R>(gd, reinterpret_cast<const float*>(Y), reinterpret_cast<const float*>(X),
                          reinterpret_cast<float*>(Out), reinterpret_cast<const float*>(V), nvec, ld, scale, ca, cb,
                          cc, st);
#endif
  if constexpr (SX <= 2 && R == 4) {
    bool use_cc = g_stencil_cc == 1;
    if (g_stencil_cc == 2 && sizeof(T) == 4 && nvec < 128) {
      constexpr int CPW = SX > 2 ? SX / 2 : 1;
      const long long lean_waves = (long long)((gd->ny + PC_LEAN_TY - 1) / PC_LEAN_TY) * nvec * (PC_LEAN_TY * SX / CPW);
      use_cc = lean_waves >= PC_LEAN_ZCHUNK_WAVES;
    }
    if (use_cc && gd->cc_wmax > 0) return launch_cc<T, R>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, st);
  }
#ifndef PC_STENCIL_NO_LEAN
  if constexpr (SX <= 2) return launch_lean<T, R, SX>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, st);
#endif
  if constexpr (PC_STENCIL_NV >= 2 && zmarch_lds_bytes<T, R, SX, 1, 2, 2, NW>() <= kLds)
    return launch_cfg<T, R, SX, 1, 2, 2, 1, NW>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
  else if constexpr (zmarch_lds_bytes<T, R, SX, 1, 1, 2, NW>() <= kLds)
    return launch_cfg<T, R, SX, 1, 1, 2, 1, NW>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
  else
    return launch_cfg<T, R, SX, 1, 1, 2, 1, 4>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
}

template <typename T, int R>
int launch_r(const GridDev* gd, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld, double scale,
             double ca, double cb, double cc, double* dot_part, hipStream_t st) {
  if (gd->nx <= 64) return launch_r_sx<T, R, 1>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
  if (gd->nx <= 128) return launch_r_sx<T, R, 2>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
  return launch_r_sx<T, R, 4>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
}

template <typename T>
int stencil_dispatch(void* grid, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld, double scale,
                     double ca, double cb, double cc, double* dot_part, void* stream) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  PC_CHECK_ARG(gd != nullptr, "null grid");
  PC_CHECK_ARG(Y && Out && V, "null operand");
  PC_CHECK_ARG(Y != Out, "output must not alias the stencil input");
  PC_CHECK_ARG(nvec >= 1, "nvec < 1");
  PC_CHECK_ARG(ld >= gd->npts, "ld < npts");
  PC_CHECK_ARG(gd->npts * (long long)sizeof(T) < 0x7ffffff0LL, "vector larger than 2 GiB");
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (gd->radius) {
    case 1: return launch_r<T, 1>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 2: return launch_r<T, 2>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 3: return launch_r<T, 3>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 4: return launch_r<T, 4>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 6: return launch_r<T, 6>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    default:
      pc::set_error("stencil: unsupported radius");
      return pc::PC_ERR_ARG;
  }
}

// Direct form for few vectors (Lanczos, CG): one wave per x-row, neighbours
// read through the caches.  Used where the z-march has too few workgroups.
template <typename T, int R>
__global__ __launch_bounds__(256) void stencil_direct(StArgs g, const T* __restrict__ Y, const T* X, T* Out,
                                                      const T* __restrict__ V, T ca, T cb, T cc,
                                                      double* __restrict__ dot_part) {
  __shared__ double red[4];
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int v = blockIdx.y;
  const int nrow = g.ny * g.nz;
  const T* __restrict__ yv = Y + (long long)v * g.ld;
  const T* xv = X ? X + (long long)v * g.ld : nullptr;
  T* ov = Out + (long long)v * g.ld;
  double dot = 0.0;
  if (row < nrow) {
    const int j = row % g.ny, k = row / g.ny;
    const int lo = g.lo[row], hi = g.hi[row], off = g.off[row];
    for (int i = lo + lane; i < hi; i += 64) {
      const long long p = off + (i - lo);
      const T yc = yv[p];
      T acc = T(g.w[0]) * yc;
#pragma unroll
      for (int d = 1; d <= R; ++d) {
        T nb = row_val(yv, lo, hi, off, i - d) + row_val(yv, lo, hi, off, i + d);
        if (j - d >= 0) {
          const int r2 = row - d;
          nb += row_val(yv, g.lo[r2], g.hi[r2], g.off[r2], i);
        }
        if (j + d < g.ny) {
          const int r2 = row + d;
          nb += row_val(yv, g.lo[r2], g.hi[r2], g.off[r2], i);
        }
        if (k - d >= 0) {
          const int r2 = row - d * g.ny;
          nb += row_val(yv, g.lo[r2], g.hi[r2], g.off[r2], i);
        }
        if (k + d < g.nz) {
          const int r2 = row + d * g.ny;
          nb += row_val(yv, g.lo[r2], g.hi[r2], g.off[r2], i);
        }
        acc += T(g.w[d]) * nb;
      }
      T o = acc + (ca * V[p] + cb) * yc;
      if (xv) o += cc * xv[p];
      ov[p] = o;
      dot += (double)o * (double)yc;
    }
  }
  if (dot_part) {
    const double s = pc_block_sum(dot, red);
    if (threadIdx.x == 0) dot_part[(long long)v * gridDim.x + blockIdx.x] = s;
  }
}

template <typename T, int R>
int launch_direct_r(const GridDev* gd, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld,
                    double scale, double ca, double cb, double cc, double* dot_part, hipStream_t st) {
  StArgs a;
  a.nx = gd->nx; a.ny = gd->ny; a.nz = gd->nz; a.npts = gd->npts;
  a.lo = gd->lo; a.hi = gd->hi; a.off = gd->off; a.zrange = gd->zrange; a.meta = gd->meta;
  a.ld = ld; a.nslab = 0; a.nvec = nvec; a.nblocks = 0; a.nb8 = 0; a.nzc = 1; a.zc = gd->nz;
  for (int d = 0; d < 8; ++d) a.w[d] = 0.0;
  a.w[0] = 3.0 * gd->w[0] * scale;
  for (int d = 1; d <= R; ++d) a.w[d] = gd->w[d] * scale;
  const int nrow = gd->ny * gd->nz;
  dim3 grid((nrow + 3) / 4, nvec);
  hipLaunchKernelGGL((stencil_direct<T, R>), grid, dim3(256), 0, st, a, Y, X, Out, V, T(ca), T(cb), T(cc), dot_part);
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// One FP64 vector (the Lanczos bounds): the z-march of the CG step over
// z-chunks (pc_zmarch.h), so each point's neighbours come from registers and
// LDS instead of dependent row-table loads.  Out may alias X.
template <int R, int SX>
__global__ __launch_bounds__(256) void stencil_single(pcz::ZMap zm, const double* __restrict__ Y, const double* X,
                                                      double* Out, const double* __restrict__ V, double ca,
                                                      double cb, double cc) {
  __shared__ double tile[pcz::Tile<R, SX>::SIZE];
  const int L = pcz::block_of(zm);
  if (L >= zm.nwg) return;
  pcz::zmarch<R, SX>(
      zm, L, tile, [&](long long idx, int, int, int) { return Y[idx]; },
      [&](long long idx, int, int, int, double qc, double acc) {
        double o = acc + (ca * V[idx] + cb) * qc;
        if (X) o += cc * X[idx];
        Out[idx] = o;
      });
}

template <int R>
int launch_single_r(const pcz::ZMap& zm, const double* Y, const double* X, double* Out, const double* V, double ca,
                    double cb, double cc, hipStream_t st) {
  const dim3 grid(8 * zm.nwg8), block(256);
  switch (zm.sx) {
    case 1: hipLaunchKernelGGL((stencil_single<R, 1>), grid, block, 0, st, zm, Y, X, Out, V, ca, cb, cc); break;
    case 2: hipLaunchKernelGGL((stencil_single<R, 2>), grid, block, 0, st, zm, Y, X, Out, V, ca, cb, cc); break;
    default: hipLaunchKernelGGL((stencil_single<R, 4>), grid, block, 0, st, zm, Y, X, Out, V, ca, cb, cc); break;
  }
  PC_LAUNCH_CHECK();
  return pc::PC_OK;
}

// Measured on Si87H76 (605,808 points, radius 4): one FP64 vector 17 us, 46 us
// with stencil_direct (same summation order, so the same bits).
int single_dispatch(void* grid, const double* Y, const double* X, double* Out, const double* V, double scale,
                    double ca, double cb, double cc, hipStream_t st, bool* done) {
  *done = false;
  const GridDev* gd = static_cast<const GridDev*>(grid);
  double w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  w[0] = 3.0 * gd->w[0] * scale;
  for (int d = 1; d <= gd->radius && d < 8; ++d) w[d] = gd->w[d] * scale;
  pcz::ZMap zm;
  int radius = 0;
  if (pcz::zmap_init(grid, w, zm, &radius) != 0) return pc::PC_OK;   // nx > 256: direct form
  int rc = pc::PC_OK;
  switch (radius) {
    case 1: rc = launch_single_r<1>(zm, Y, X, Out, V, ca, cb, cc, st); break;
    case 2: rc = launch_single_r<2>(zm, Y, X, Out, V, ca, cb, cc, st); break;
    case 3: rc = launch_single_r<3>(zm, Y, X, Out, V, ca, cb, cc, st); break;
    case 4: rc = launch_single_r<4>(zm, Y, X, Out, V, ca, cb, cc, st); break;
    case 6: rc = launch_single_r<6>(zm, Y, X, Out, V, ca, cb, cc, st); break;
    default: return pc::PC_OK;
  }
  *done = rc == pc::PC_OK;
  return rc;
}

template <typename T>
int direct_dispatch(void* grid, const T* Y, const T* X, T* Out, const T* V, int nvec, long long ld, double scale,
                    double ca, double cb, double cc, double* dot_part, void* stream) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  PC_CHECK_ARG(gd != nullptr, "null grid");
  PC_CHECK_ARG(Y && Out && V, "null operand");
  PC_CHECK_ARG(Y != Out, "output must not alias the stencil input");
  PC_CHECK_ARG(nvec >= 1 && nvec <= 65535, "nvec out of range");
  PC_CHECK_ARG(ld >= gd->npts, "ld < npts");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if constexpr (std::is_same<T, double>::value) {
    if (nvec == 1 && dot_part == nullptr) {
      bool done = false;
      const int rc = single_dispatch(grid, Y, X, Out, V, scale, ca, cb, cc, st, &done);
      if (rc != pc::PC_OK || done) return rc;
    }
  }
  switch (gd->radius) {
    case 1: return launch_direct_r<T, 1>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 2: return launch_direct_r<T, 2>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 3: return launch_direct_r<T, 3>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 4: return launch_direct_r<T, 4>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    case 6: return launch_direct_r<T, 6>(gd, Y, X, Out, V, nvec, ld, scale, ca, cb, cc, dot_part, st);
    default:
      pc::set_error("stencil: unsupported radius");
      return pc::PC_ERR_ARG;
  }
}

}  // namespace

// Used by the CG solver in pc_pcg.hip.
int pc_internal_stencil_direct_f64(void* grid, const double* Y, const double* X, double* Out, const double* V,
                                   double scale, double ca, double cb, double cc, double* dot_part, void* stream,
                                   int* out_nblocks) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  if (out_nblocks) *out_nblocks = (gd->ny * gd->nz + 3) / 4;
  return direct_dispatch<double>(grid, Y, X, Out, V, 1, gd->npts, scale, ca, cb, cc, dot_part, stream);
}

long long pc_internal_grid_npts(void* grid) { return static_cast<const GridDev*>(grid)->npts; }
int pc_internal_grid_view(void* grid, int* nx, int* ny, int* nz, const int** lo, const int** hi, const int** off) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  *nx = gd->nx; *ny = gd->ny; *nz = gd->nz;
  *lo = gd->lo; *hi = gd->hi; *off = gd->off;
  return 0;
}
// Point map for single-vector kernels (the Hartree CG): row metadata
// {off-lo, lo|hi<<16}, the row of every point, and the unscaled weights.
int pc_internal_grid_points(void* grid, int* ny, int* nz, int* radius, const int2** meta, const int** prow,
                            double* w8) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  *ny = gd->ny; *nz = gd->nz; *radius = gd->radius;
  *meta = gd->meta; *prow = gd->prow;
  for (int d = 0; d < 8; ++d) w8[d] = gd->w[d];
  return 0;
}
// z-extent per y (min z, max z + 1 of the non-empty rows) and nx, for z-marching kernels
int pc_internal_grid_zmap(void* grid, int* nx, const int** zrange) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  *nx = gd->nx;
  *zrange = gd->zrange;
  return 0;
}
long long pc_internal_grid_nrow(void* grid) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  return (long long)gd->ny * gd->nz;
}

extern "C" {

int pc_stencil_direct_f64(void* grid, const double* Y, const double* X, double* Out, const double* V, int nvec,
                          long long ld, double scale, double a, double b, double c, double* dot_part, void* stream) {
  return direct_dispatch<double>(grid, Y, X, Out, V, nvec, ld, scale, a, b, c, dot_part, stream);
}

int pc_stencil_direct_f32(void* grid, const float* Y, const float* X, float* Out, const float* V, int nvec,
                          long long ld, double scale, double a, double b, double c, double* dot_part, void* stream) {
  return direct_dispatch<float>(grid, Y, X, Out, V, nvec, ld, scale, a, b, c, dot_part, stream);
}

int pc_grid_create(int nx, int ny, int nz, const int* row_lo, const int* row_hi, int radius, const double* weights,
                   void** out_grid, long long* out_npts) {
  PC_CHECK_ARG(out_grid && row_lo && row_hi && weights, "null pointer");
  PC_CHECK_ARG(nx >= 1 && ny >= 1 && nz >= 1, "bad dims");
  PC_CHECK_ARG(nx <= 256, "nx > 256 not supported (x rows are held by one workgroup)");
  PC_CHECK_ARG(radius == 1 || radius == 2 || radius == 3 || radius == 4 || radius == 6,
               "radius must be 1,2,3,4 or 6");
  const int nrow = ny * nz;
  std::vector<int> off(nrow), zr(2 * ny);
  for (int j = 0; j < ny; ++j) { zr[2 * j] = nz; zr[2 * j + 1] = 0; }
  long long cnt = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const int r = j + ny * k;
      PC_CHECK_ARG(row_lo[r] >= 0 && row_hi[r] <= nx, "row extent outside [0,nx]");
      off[r] = (int)cnt;
      const int len = std::max(0, row_hi[r] - row_lo[r]);
      if (len > 0) {
        zr[2 * j] = std::min(zr[2 * j], k);
        zr[2 * j + 1] = std::max(zr[2 * j + 1], k + 1);
      }
      cnt += len;
    }
  PC_CHECK_ARG(cnt < (1LL << 31), "too many points");
  // normalise empty rows to lo=hi=0
  std::vector<int> lo(row_lo, row_lo + nrow), hi(row_hi, row_hi + nrow);
  for (int r = 0; r < nrow; ++r)
    if (hi[r] <= lo[r]) { lo[r] = 0; hi[r] = 0; }
  std::vector<int2> meta(nrow);
  for (int r = 0; r < nrow; ++r) meta[r] = make_int2(off[r] - lo[r], lo[r] | (hi[r] << 16));
  GridDev* g = new GridDev();
  g->nx = nx; g->ny = ny; g->nz = nz; g->npts = cnt; g->radius = radius;
  g->cc_tab = nullptr;
  g->cc_wmax = 0;
  // stencil_cc column lists: per slab of kCcTY rows, each own row's and halo
  // row's x-range over the planes of the slab's march (an interval on a convex
  // domain), as (tile row << 16 | x)
  std::vector<int> cc;
  int ccw = 0;
  if (nx <= 128) {
    const int TY = kCcTY, nslab = (ny + TY - 1) / TY, LR = TY + 2 * radius;
    std::vector<std::vector<int>> own(nslab), hal(nslab);
    for (int sl = 0; sl < nslab; ++sl) {
      int zb = nz, ze = 0;
      for (int t = 0; t < TY; ++t) {
        const int j = sl * TY + t;
        if (j < ny) { zb = std::min(zb, zr[2 * j]); ze = std::max(ze, zr[2 * j + 1]); }
      }
      for (int t = 0; t < LR; ++t) {
        const int j = sl * TY - radius + t;
        if (j < 0 || j >= ny || zb >= ze) continue;
        int xl = nx, xh = 0;
        for (int k = zb; k < ze; ++k) {
          const int r = j + ny * k;
          if (hi[r] > lo[r]) { xl = std::min(xl, lo[r]); xh = std::max(xh, hi[r]); }
        }
        const bool is_own = t >= radius && t < radius + TY;
        for (int x = xl; x < xh; ++x) (is_own ? own[sl] : hal[sl]).push_back((t << 16) | x);
      }
      ccw = std::max(ccw, (int)std::max((own[sl].size() + 63) / 64, (hal[sl].size() + 63) / 64));
    }
    if (ccw >= 1 && ccw <= 16) {
      cc.assign((size_t)nslab * 2 * ccw * 64, -1);
      for (int sl = 0; sl < nslab; ++sl) {
        std::copy(own[sl].begin(), own[sl].end(), cc.begin() + (size_t)sl * 2 * ccw * 64);
        std::copy(hal[sl].begin(), hal[sl].end(), cc.begin() + (size_t)sl * 2 * ccw * 64 + ccw * 64);
      }
    } else {
      ccw = 0;
    }
  }
  for (int d = 0; d < 8; ++d) g->w[d] = d <= radius ? weights[d] : 0.0;
  hipError_t e = hipSuccess;
  e = hipMalloc(&g->lo, nrow * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&g->hi, nrow * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&g->off, nrow * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&g->zrange, 2 * ny * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&g->meta, nrow * sizeof(int2));
  std::vector<int> prow((size_t)std::max(cnt, 1LL));
  for (int r = 0; r < nrow; ++r)
    for (int t = 0; t < hi[r] - lo[r]; ++t) prow[(size_t)off[r] + t] = r;
  if (e == hipSuccess) e = hipMalloc(&g->prow, prow.size() * sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(g->prow, prow.data(), prow.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->meta, meta.data(), nrow * sizeof(int2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->lo, lo.data(), nrow * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->hi, hi.data(), nrow * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->off, off.data(), nrow * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->zrange, zr.data(), 2 * ny * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && ccw > 0) {
    e = hipMalloc(&g->cc_tab, cc.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(g->cc_tab, cc.data(), cc.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) g->cc_wmax = ccw;
  }
  if (e != hipSuccess) {
    pc::set_error(std::string("pc_grid_create: ") + hipGetErrorString(e));
    return pc::PC_ERR_HIP;
  }
  *out_grid = g;
  if (out_npts) *out_npts = cnt;
  return pc::PC_OK;
}

int pc_grid_destroy(void* grid) {
  GridDev* g = static_cast<GridDev*>(grid);
  if (!g) return pc::PC_OK;
  (void)hipFree(g->lo); (void)hipFree(g->hi); (void)hipFree(g->off); (void)hipFree(g->zrange);
  (void)hipFree(g->meta); (void)hipFree(g->prow);
  if (g->cc_tab) (void)hipFree(g->cc_tab);
  delete g;
  return pc::PC_OK;
}

// Stencil kernel for radius-4 grids of at most 128 points per row
// (process-wide): 0 = stencil_lean, 1 = stencil_cc, 2 = stencil_cc where it
// measured faster (the default; see g_stencil_cc).
int pc_stencil_set_variant(int variant) {
  PC_CHECK_ARG(variant >= 0 && variant <= 2, "variant must be 0, 1 or 2");
  g_stencil_cc = variant;
  return pc::PC_OK;
}

int pc_stencil_f64(void* grid, const double* Y, const double* X, double* Out, const double* V, int nvec,
                   long long ld, double scale, double a, double b, double c, double* dot_part, void* stream) {
  return stencil_dispatch<double>(grid, Y, X, Out, V, nvec, ld, scale, a, b, c, dot_part, stream);
}

int pc_stencil_f32(void* grid, const float* Y, const float* X, float* Out, const float* V, int nvec,
                   long long ld, double scale, double a, double b, double c, double* dot_part, void* stream) {
  return stencil_dispatch<float>(grid, Y, X, Out, V, nvec, ld, scale, a, b, c, dot_part, stream);
}

// Number of per-workgroup partial sums pc_stencil_* writes when dot_part is
// given (padded layout: entry L of the logical workgroup order, vector-major).
int pc_stencil_num_partials(void* grid, int dtype_bytes, int nvec, int* out_count, int* out_per_vector) {
  const GridDev* gd = static_cast<const GridDev*>(grid);
  PC_CHECK_ARG(gd && out_count && out_per_vector, "null pointer");
  PC_CHECK_ARG(dtype_bytes == 4 || dtype_bytes == 8, "dtype_bytes must be 4 or 8");
  // with partial sums: one vector and one y-row per wavefront, 4 rows per workgroup
  const int nslab = (gd->ny + 3) / 4;
  *out_per_vector = nslab;
  *out_count = nslab * nvec;
  return pc::PC_OK;
}

}  // extern "C"
