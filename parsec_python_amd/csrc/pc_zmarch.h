// FP64 z-march for one application of the domain's stencil to a field that is
// formed on the fly, shared by the CG step and the multigrid fine level.
//
// A workgroup of 256 threads owns 4 y-rows (one per wavefront) of the sphere-
// masked grid and a chunk of zc z-planes of them, and marches z: each lane
// keeps its points' 2R+1 z-values in registers, the current plane of the 4
// rows plus R halo rows on each side sits in LDS (two tiles, one barrier per
// plane) for the x/y differences, and the loads for plane z+1 are issued
// before plane z is computed.  Values outside the domain are zero (Dirichlet).
//
//   src(idx, i, j, z) -> value of the field at domain point idx = (i, j, z)
//   out(idx, i, j, z, qc, acc) receives the centre value qc and
//       acc = sum_d w_d (6 neighbours at distance d) + w_0 qc
//       for every domain point of the workgroup's block, once.
#pragma once
#include "pc_common.h"

int pc_internal_grid_points(void* grid, int* ny, int* nz, int* radius, const int2** meta, const int** prow,
                            double* w8);
int pc_internal_grid_zmap(void* grid, int* nx, const int** zrange);

namespace pcz {

constexpr int kZW = 4;          // y-rows per workgroup (one per wavefront)
constexpr int kMaxWG = 2048;    // workgroups of one march (= partials)

struct ZMap {
  const int2* meta;    // per row {off - lo, lo | hi << 16}
  const int* zrange;   // [2*ny]: z extent of each y-row column
  int nx, ny, nz;
  int nslab, nzc, zc;  // 4-row slabs, z-chunks per slab, planes per chunk
  int nwg, nwg8;       // nslab * nzc workgroups, ceil(nwg / 8)
  int sx;              // 64-point x-strips per lane row (1, 2 or 4)
  double w[8];         // weights; w[0] is the 3D centre weight
};

// Fills the map for `grid` with weights w[0..7] (w[0] already the 3D centre
// weight).  zc = 6 planes per chunk (measured on a 106^3 sphere: 38 us per CG
// iteration at 6 planes, 41 us at 2), raised until the workgroups fit kMaxWG.
inline int zmap_init(void* grid, const double* w, ZMap& zm, int* radius) {
  int ny, nz;
  const int* prow;
  double w8[8];
  pc_internal_grid_points(grid, &ny, &nz, radius, &zm.meta, &prow, w8);
  zm.ny = ny;
  zm.nz = nz;
  pc_internal_grid_zmap(grid, &zm.nx, &zm.zrange);
  for (int d = 0; d < 8; ++d) zm.w[d] = w[d];
  zm.nslab = (zm.ny + kZW - 1) / kZW;
  zm.zc = 6;
  while ((long long)zm.nslab * ((zm.nz + zm.zc - 1) / zm.zc) > kMaxWG) ++zm.zc;
  zm.nzc = (zm.nz + zm.zc - 1) / zm.zc;
  zm.nwg = zm.nslab * zm.nzc;
  zm.nwg8 = (zm.nwg + 7) / 8;
  zm.sx = zm.nx <= 64 ? 1 : (zm.nx <= 128 ? 2 : 4);
  return zm.nx <= 256 ? 0 : 1;
}

// Block of launch-order workgroup b: consecutive blocks go to the same XCD
// (workgroup ids are dealt round-robin over the 8 XCDs), so the z-neighbour
// planes a block reads are mostly in that XCD's L2.
__device__ __forceinline__ int block_of(const ZMap& g) { return (blockIdx.x & 7) * g.nwg8 + (blockIdx.x >> 3); }

__device__ __forceinline__ int2 zrow(const ZMap& g, int j, int z) {
  return (j >= 0 && j < g.ny && z >= 0 && z < g.nz) ? g.meta[j + g.ny * z] : make_int2(0, 0);
}

template <int R, int SX>
struct Tile {
  static constexpr int LW = 64 * SX + 2 * R;
  static constexpr int LR = kZW + 2 * R;
  static constexpr int SIZE = 2 * LR * LW;   // doubles
};

// The march of block L (0 <= L < g.nwg); `tile` holds Tile<R, SX>::SIZE doubles
// of LDS.  Every thread of the workgroup must call it.
template <int R, int SX, class Src, class Out>
__device__ __forceinline__ void zmarch(const ZMap& g, int L, double* tile, Src src, Out out) {
  constexpr int TY = kZW;
  constexpr int LW = Tile<R, SX>::LW;
  constexpr int LR = Tile<R, SX>::LR;
  constexpr int HT = (2 * R + kZW - 1) / kZW;
  const int slab = L % g.nslab;
  const int zci = L / g.nslab;
  const int j0 = slab * TY;
  int zb = g.nz, ze = 0;
  for (int q = 0; q < TY; ++q) {
    const int j = j0 + q;
    if (j < g.ny) {
      zb = min(zb, g.zrange[2 * j]);
      ze = max(ze, g.zrange[2 * j + 1]);
    }
  }
  const int z0 = zb + zci * g.zc;
  const int z1 = min(ze, z0 + g.zc);
  if (z0 >= z1) return;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = j0 + wv;
  auto load = [&](int2 m, int i, int jj, int z) -> double {
    const int lo = m.y & 0xffff, hi = m.y >> 16;
    return (i >= lo && i < hi) ? src((long long)m.x + i, i, jj, z) : 0.0;
  };
  // Row metadata windows: lane l holds meta(row, wb + l), so a plane's
  // lookups are register reads and its loads never wait for a metadata load;
  // a window is reloaded once every 64 planes.  The own row serves planes z
  // (output) and z + R (queue front).
  struct Win {
    int2 v;
    int wb;
  };
  auto win_load = [&](Win& w, int jj, int wb) {
    w.wb = wb;
    w.v = zrow(g, jj, wb + lane);
  };
  auto win_at = [&](const Win& w, int z) -> int2 {
    const int l = z - w.wb;
    return make_int2(__builtin_amdgcn_readlane(w.v.x, l), __builtin_amdgcn_readlane(w.v.y, l));
  };
  Win wown, whalo[HT];
  win_load(wown, j, z0 - R);
#pragma unroll
  for (int h = 0; h < HT; ++h) {
    const int hr = wv + kZW * h;
    const int tr = hr < R ? hr : hr + TY;
    win_load(whalo[h], j0 - R + tr, z0);
  }
  for (int e = threadIdx.x; e < 2 * LR * LW; e += 256) tile[e] = 0.0;
  __syncthreads();
  double q[SX][2 * R + 1];
  double nq[SX], hc[HT][SX], hn[HT][SX];
  int2 mcur, mnext;
#pragma unroll
  for (int s = 0; s < SX; ++s) {
    q[s][0] = 0.0;
#pragma unroll
    for (int h = 0; h < HT; ++h) hc[h][s] = hn[h][s] = 0.0;
  }
#pragma unroll
  for (int dz = 1; dz <= 2 * R; ++dz) {
    const int zz = z0 - R - 1 + dz;
    const int2 m = win_at(wown, zz);
#pragma unroll
    for (int s = 0; s < SX; ++s) q[s][dz] = load(m, lane + 64 * s, j, zz);
  }
  auto prefetch = [&](int z) {
    if (z + R >= wown.wb + 64) win_load(wown, j, z);
    const int2 mq = win_at(wown, z + R);
    mnext = win_at(wown, z);
#pragma unroll
    for (int s = 0; s < SX; ++s) nq[s] = load(mq, lane + 64 * s, j, z + R);
#pragma unroll
    for (int h = 0; h < HT; ++h) {
      const int hr = wv + kZW * h;
      if (hr < 2 * R) {
        const int tr = hr < R ? hr : hr + TY;
        const int jh = j0 - R + tr;
        if (z >= whalo[h].wb + 64) win_load(whalo[h], jh, z);
        const int2 mh = win_at(whalo[h], z);
#pragma unroll
        for (int s = 0; s < SX; ++s) hn[h][s] = load(mh, lane + 64 * s, jh, z);
      }
    }
  };
  prefetch(z0);
  for (int z = z0; z < z1; ++z) {
    double* tb = tile + ((z - z0) & 1) * (LR * LW);
    mcur = mnext;
#pragma unroll
    for (int s = 0; s < SX; ++s) {
#pragma unroll
      for (int d = 0; d < 2 * R; ++d) q[s][d] = q[s][d + 1];
      q[s][2 * R] = nq[s];
#pragma unroll
      for (int h = 0; h < HT; ++h) hc[h][s] = hn[h][s];
    }
    if (z + 1 < z1) prefetch(z + 1);
#pragma unroll
    for (int s = 0; s < SX; ++s) tb[(wv + R) * LW + R + lane + 64 * s] = q[s][R];
#pragma unroll
    for (int h = 0; h < HT; ++h) {
      const int hr = wv + kZW * h;
      if (hr < 2 * R) {
        const int tr = hr < R ? hr : hr + TY;
#pragma unroll
        for (int s = 0; s < SX; ++s) tb[tr * LW + R + lane + 64 * s] = hc[h][s];
      }
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);   // LDS only: the prefetch stays in flight
    __builtin_amdgcn_s_barrier();
    const int lo = mcur.y & 0xffff, hi = mcur.y >> 16;
#pragma unroll
    for (int s = 0; s < SX; ++s) {
      const int i = lane + 64 * s;
      const double* trow = tb + (wv + R) * LW + R + i;
      double acc = g.w[0] * q[s][R];
#pragma unroll
      for (int d = 1; d <= R; ++d) {
        const double nb = trow[-d] + trow[d] + trow[-d * LW] + trow[d * LW] + q[s][R - d] + q[s][R + d];
        acc += g.w[d] * nb;
      }
      if (i >= lo && i < hi) out((long long)mcur.x + i, i, j, z, q[s][R], acc);
    }
  }
}

}  // namespace pcz
