// stencil.h -- the stream+collide tile bodies shared by the per-domain
// kernels (kernels2d.hip, kernels3d.hip) and the batched kernels (batch.hip).
//
// The arithmetic is that of mrDeformableSolver2DKernel_Shrd
// (src/core/FsiCoupling2D.cu:82-255) and mrDeformableSolver3DKernel_ShrdM
// (src/core/FsiCoupling3D.cu:83-334) for cells without a moving-wall link.
#pragma once

#include <type_traits>

#include "ff_internal.h"
#include "hermite.h"

namespace ff {

// XCD-aware tile order.  Workgroups are dispatched to the 8 XCDs round-robin
// by block index and every XCD has its own L2, so block p runs tile
// (p % 8) * ceil(n / 8) + p / 8: each XCD works through one contiguous run of
// tiles (whole rows), and the halo rows and columns its tiles share hit in
// its L2 instead of being fetched once per XCD.  Grids are rounded up to a
// multiple of 8 blocks; -1: no tile for this block.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_grid(int n) { return (n + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ int xcd_tile(int p, int n) {
  const int per = (n + kXcds - 1) / kXcds;
  const int t = (p % kXcds) * per + p / kXcds;
  return t < n ? t : -1;
}

// 2D tile shapes: TX x (TY * CPT) cells per 256-thread block, each thread
// updating CPT cells of one column (rows ly, ly + TY, ...).  Taller tiles
// re-read and re-reconstruct less halo per cell at the price of LDS.
struct Tile2 {
  int tx, ty, cpt;
};
constexpr Tile2 kTiles2[] = {{32, 8, 1}, {32, 8, 2}, {64, 4, 2}, {32, 8, 4}, {64, 4, 4}};
constexpr int kNumTiles2 = 5;

// Stage the reconstructed populations of the tile and its one-cell halo in
// LDS (clamped or periodic halo; binary16 when the reference's half tile is
// reproduced), then pull, collide and store every updated cell.  FAST picks
// recon2_fast / collide2_fast.
template <typename S, bool STAGE, bool PERIODIC, int TX, int TY, int CPT, bool FAST = false>
__device__ __forceinline__ void bulk2d_tile(const S* __restrict__ in, S* __restrict__ out,
                                            const uint8_t* __restrict__ flags, int nx, int ny, size_t plane,
                                            float omega, int bx, int by) {
  constexpr int NT = TX * TY, RY = TY * CPT;
  constexpr int HX = TX + 2, HY = RY + 2, HN = HX * HY;
  using T = typename std::conditional<STAGE, __half, float>::type;
  constexpr int PER = (HN + NT - 1) / NT;  // halo cells per thread
  __shared__ T sp[9][HN];
  const int x0 = bx * TX, y0 = by * RY;
  // every load of the tile is issued before the first reconstruction, so the
  // block waits for memory once
  float m[PER][6];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int k = threadIdx.x + r * NT;
    if (k < HN) {
      const int hx = k % HX, hy = k / HX;
      int gx = x0 + hx - 1, gy = y0 + hy - 1;
      if (PERIODIC) {
        gx = ((gx % nx) + nx) % nx;
        gy = ((gy % ny) + ny) % ny;
      } else {
        gx = min(nx - 1, max(0, gx));
        gy = min(ny - 1, max(0, gy));
      }
      const size_t c = (size_t)gy * nx + gx;
#pragma unroll
      for (int q = 0; q < 6; q++) m[r][q] = Store<S>::template load<2>(in + q * plane + c, q, plane);
    }
  }
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int k = threadIdx.x + r * NT;
    if (k < HN) {
      float f[9];
      if constexpr (FAST) recon2_fast(m[r][0], m[r][1], m[r][2], m[r][3], m[r][5], m[r][4], f);
      else recon2_all(m[r][0], m[r][1], m[r][2], m[r][3], m[r][5], m[r][4], f);
#pragma unroll
      for (int i = 0; i < 9; i++) {
        if constexpr (STAGE) sp[i][k] = to_half_rn(f[i]);
        else sp[i][k] = f[i];
      }
    }
  }
  __syncthreads();
  const int lx = threadIdx.x % TX, ly0 = threadIdx.x / TX;
  const int x = x0 + lx;
  if (x >= nx) return;
#pragma unroll
  for (int r = 0; r < CPT; r++) {
    const int ly = ly0 + r * TY;
    const int y = y0 + ly;
    if (y >= ny) break;
    const size_t cell = (size_t)y * nx + x;
    if (!is_updated(flags[cell])) continue;
    float p[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int idx = (ly + 1 - D2::cy(i)) * HX + (lx + 1 - D2::cx(i));
      float v;
      if constexpr (STAGE) v = __half2float(sp[i][idx]);
      else v = sp[i][idx];
      p[i] = fmaxf(0.0f, v);
    }
    float o[6];
    if constexpr (FAST) collide2_fast(p, omega, o);
    else collide2(p, omega, o);
#pragma unroll
    for (int q = 0; q < 6; q++) Store<S>::template store<2>(out + q * plane + cell, q, o[q], plane);
  }
}

// 3D box tile of TX x TY x TZ cells (one per thread) with a one-cell halo in
// LDS (FP32 as the reference's tile, FsiCoupling3D.cu:94).  The reference
// keeps the flag bounce-back value in 3D (FsiCoupling3D.cu:209-222), so
// upstream flags are staged too.  FAST selects recon3_fast / collide3_fast.
template <typename S, bool PERIODIC, int TX, int TY, int TZ, bool FAST = false>
__device__ __forceinline__ void bulk3d_tile(const S* __restrict__ in, S* __restrict__ out,
                                            const uint8_t* __restrict__ flags, int nx, int ny, int nz, size_t plane,
                                            float omega, int bx, int by, int bz) {
  constexpr int kTX3 = TX, kTY3 = TY, kTZ3 = TZ;
  constexpr int HX = kTX3 + 2, HY = kTY3 + 2, HZ = kTZ3 + 2, HN = HX * HY * HZ;
  constexpr int NT = kTX3 * kTY3 * kTZ3;
  constexpr int PER = (HN + NT - 1) / NT;  // halo cells per thread
  constexpr int NW = (NT + 63) / 64;
  __shared__ float sp[27][HN];
  __shared__ uint8_t sf[HN];
  __shared__ int swall[NW];  // per wave: a staged cell is a bounce-back wall
  const int x0 = bx * kTX3, y0 = by * kTY3, z0 = bz * kTZ3;
  // all loads first, then the reconstructions (one wait for memory)
  float m[PER][10];
  uint8_t fl[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int k = threadIdx.x + r * NT;
    if (k < HN) {
      const int hx = k % HX, hy = (k / HX) % HY, hz = k / (HX * HY);
      const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
      const int wx = ((gx % nx) + nx) % nx, wy = ((gy % ny) + ny) % ny, wz = ((gz % nz) + nz) % nz;
      int sx, sy, sz;
      if (PERIODIC) {
        sx = wx; sy = wy; sz = wz;
      } else {
        sx = min(nx - 1, max(0, gx));
        sy = min(ny - 1, max(0, gy));
        sz = min(nz - 1, max(0, gz));
      }
      const size_t c = ((size_t)sz * ny + sy) * nx + sx;
#pragma unroll
      for (int q = 0; q < 10; q++) m[r][q] = Store<S>::template load<3>(in + q * plane + c, q, plane);
      fl[r] = flags[((size_t)wz * ny + wy) * nx + wx];
    }
  }
  bool wall = false;
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int k = threadIdx.x + r * NT;
    if (k < HN) {
      float f[27];
      if constexpr (FAST) recon3_fast(m[r], f);
      else recon3_all(m[r], f);
#pragma unroll
      for (int i = 0; i < 27; i++) sp[i][k] = f[i];
      sf[k] = fl[r];
      wall = wall || is_wall3(fl[r]);
    }
  }
  const unsigned long long vote = __ballot(wall);
  if ((threadIdx.x & 63) == 0) swall[threadIdx.x >> 6] = vote != 0ull;
  __syncthreads();
  // without a wall anywhere in the tile and halo every population is pulled,
  // and the upstream flags need not be read (a block-uniform branch)
  int walls = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) walls |= swall[w];
  const int lx = threadIdx.x % kTX3, ly = (threadIdx.x / kTX3) % kTY3, lz = threadIdx.x / (kTX3 * kTY3);
  const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
  if (x >= nx || y >= ny || z >= nz) return;
  const int ctr = ((lz + 1) * HY + (ly + 1)) * HX + (lx + 1);
  if (!is_updated(sf[ctr])) return;
  float p[27];
  if (walls) {
#pragma unroll
    for (int i = 0; i < 27; i++) {
      const int up = ctr - (D3::cz(i) * HY + D3::cy(i)) * HX - D3::cx(i);
      if (is_wall3(sf[up])) p[i] = fmaxf(0.0f, sp[D3::inv(i)][ctr]);
      else p[i] = fmaxf(0.0f, sp[i][up]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 27; i++) p[i] = fmaxf(0.0f, sp[i][ctr - (D3::cz(i) * HY + D3::cy(i)) * HX - D3::cx(i)]);
  }
  float o[10];
  if constexpr (FAST) collide3_fast(p, omega, o);
  else collide3(p, omega, o);
  const size_t cell = ((size_t)z * ny + y) * nx + x;
#pragma unroll
  for (int q = 0; q < 10; q++) Store<S>::template store<3>(out + q * plane + cell, q, o[q], plane);
}

// 3D ring variant: a TX x TY column of cells per block marching along z over
// ZC planes.  Reconstructed populations of three planes (z-1, z, z+1, each
// with a one-cell xy halo) rotate through LDS, so each halo cell is loaded
// and reconstructed once per column instead of once per box (1.41x instead
// of 2.53x for 16x8).  The 27 directions fall into three classes of nine by
// c_z, and each class is kept only while a pull or the centre cell's own
// bounce-back needs it.  Plane z+1 is reconstructed before plane z is
// computed; its c_z = -1 class goes to LDS at once (plane z pulls it), its
// c_z = +1 and c_z = 0 classes wait in registers until plane z is done:
//   c_z = +1 (pulled from the plane below): planes z-1, z -> 2 slots
//   c_z =  0 (pulled within the plane):      plane z       -> 1 slot
//   c_z = -1 (pulled from the plane above):  planes z, z+1 -> 2 slots
// 45 instead of 81 population planes in LDS (16x8: 33 KB, four blocks per
// CU).  The next plane's moments are fetched into registers before the
// current plane is computed, hiding their latency.  Arithmetic per cell is
// that of bulk3d_tile.
struct Tile3 {
  int tx, ty, tz;  // tz: planes per block
  int kind;        // 0 box (bulk3d_tile), 1 ring (bulk3d_ring), 2 column (bulk3d_col)
  int nt;          // block size
  int pf;          // column: planes of moments in flight (bulk3d_col's PF)
  int dma;         // column: the planes in flight are an LDS ring filled by LDS DMA (FP32 storage)
};
constexpr Tile3 kTiles3[] = {
    {8, 8, 4, 0, 256, 0, 0},   {16, 4, 4, 0, 256, 0, 0},  {16, 8, 32, 1, 128, 0, 0}, {16, 8, 16, 1, 128, 0, 0},
    {16, 8, 8, 1, 128, 0, 0},  {16, 16, 32, 1, 256, 0, 0}, {16, 16, 16, 1, 256, 0, 0}, {16, 12, 8, 2, 256, 1, 0},
    {16, 26, 8, 2, 512, 1, 0}, {64, 4, 16, 2, 448, 1, 0}, {16, 12, 8, 2, 256, 2, 1},  {16, 8, 8, 2, 192, 3, 1},
    {32, 8, 8, 2, 384, 2, 1},  {16, 12, 16, 2, 256, 3, 1}, {64, 4, 32, 2, 448, 1, 0}, {32, 12, 16, 2, 512, 1, 0}};
constexpr int kNumTiles3 = 16;

// The shapes of kTiles3 by index (kept in step with the table);
// F(FAST, KIND, TX, TY, ZC, NT, PF, DMA) instantiates one.
#define FF_TILES3(F, FASTV, tile)                              \
  switch (tile) {                                              \
    case 1: F(FASTV, 0, 16, 4, 4, 256, 1, false); break;       \
    case 2: F(FASTV, 1, 16, 8, 32, 128, 1, false); break;      \
    case 3: F(FASTV, 1, 16, 8, 16, 128, 1, false); break;      \
    case 4: F(FASTV, 1, 16, 8, 8, 128, 1, false); break;       \
    case 5: F(FASTV, 1, 16, 16, 32, 256, 1, false); break;     \
    case 6: F(FASTV, 1, 16, 16, 16, 256, 1, false); break;     \
    case 7: F(FASTV, 2, 16, 12, 8, 256, 1, false); break;      \
    case 8: F(FASTV, 2, 16, 26, 8, 512, 1, false); break;      \
    case 9: F(FASTV, 2, 64, 4, 16, 448, 1, false); break;      \
    case 10: F(FASTV, 2, 16, 12, 8, 256, 2, true); break;      \
    case 11: F(FASTV, 2, 16, 8, 8, 192, 3, true); break;       \
    case 12: F(FASTV, 2, 32, 8, 8, 384, 2, true); break;       \
    case 13: F(FASTV, 2, 16, 12, 16, 256, 3, true); break;     \
    case 14: F(FASTV, 2, 64, 4, 32, 448, 1, false); break;     \
    case 15: F(FASTV, 2, 32, 12, 16, 512, 1, false); break;    \
    default: F(FASTV, 0, 8, 8, 4, 256, 1, false); break;       \
  }

// position of direction i among the nine directions sharing its c_z
__host__ __device__ constexpr int zclass_index(int i) {
  int n = 0;
  for (int j = 0; j < i; j++) n += D3::cz(j) == D3::cz(i) ? 1 : 0;
  return n;
}

template <typename S, bool PERIODIC, int TX, int TY, int ZC, bool FAST = false>
__device__ __forceinline__ void bulk3d_ring(const S* __restrict__ in, S* __restrict__ out,
                                            const uint8_t* __restrict__ flags, int nx, int ny, int nz, size_t plane,
                                            float omega, int bx, int by, int bz) {
  constexpr int NT = TX * TY, HX = TX + 2, HY = TY + 2, HA = HX * HY;
  constexpr int PER = (HA + NT - 1) / NT;  // halo cells per thread per plane
  constexpr int NW = (NT + 63) / 64;
  // rel = z - z0 + 1 numbers the planes from the lower halo plane
  __shared__ float sup[2][9][HA];  // c_z = +1 class, slot rel % 2
  __shared__ float szc[9][HA];     // c_z = 0 class of the current plane
  __shared__ float sdn[2][9][HA];  // c_z = -1 class, slot rel % 2
  __shared__ uint8_t sf[3][HA];    // flags, slot rel % 3
  __shared__ int swall[3][NW];     // per flag slot and wave: a staged cell is a wall
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
  const int zend = min(z0 + ZC, nz);
  float mreg[PER][10];
  uint8_t freg[PER];
  float hold[PER][18];  // c_z = +1 (0..8) and c_z = 0 (9..17) of the plane in flight
  // in-plane offsets of this thread's halo cells, the same on every plane:
  // the pull source (clamped or periodic) and the flag (always wrapped)
  int off_src[PER], off_flag[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int k = threadIdx.x + r * NT;
    const int gx = x0 + k % HX - 1, gy = y0 + k / HX - 1;
    const int wx = ((gx % nx) + nx) % nx, wy = ((gy % ny) + ny) % ny;
    const int sx = PERIODIC ? wx : min(nx - 1, max(0, gx));
    const int sy = PERIODIC ? wy : min(ny - 1, max(0, gy));
    off_src[r] = sy * nx + sx;
    off_flag[r] = wy * nx + wx;
  }
  // fetch the moments and flags of plane gz (may be -1 or nz: halo) into registers
  auto fetch = [&](int gz) {
    const int wz = ((gz % nz) + nz) % nz;
    const int sz = PERIODIC ? wz : min(nz - 1, max(0, gz));
    const size_t zs = (size_t)sz * ny * nx, zf = (size_t)wz * ny * nx;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int k = threadIdx.x + r * NT;
      if (k < HA) {
        const size_t c = zs + off_src[r];
#pragma unroll
        for (int q = 0; q < 10; q++) mreg[r][q] = Store<S>::template load<3>(in + q * plane + c, q, plane);
        freg[r] = flags[zf + off_flag[r]];
      }
    }
  };
  // reconstruct the fetched plane (number rel): c_z = -1 class and flags to
  // LDS, the other two classes to `hold`
  auto put_dn = [&](int rel) {
    const int s3 = rel % 3, s2 = rel % 2;
    bool wall = false;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int k = threadIdx.x + r * NT;
      if (k < HA) {
        float f[27];
        if constexpr (FAST) recon3_fast(mreg[r], f);
        else recon3_all(mreg[r], f);
#pragma unroll
        for (int i = 0; i < 27; i++) {
          const int cz = D3::cz(i), j = zclass_index(i);
          if (cz < 0) sdn[s2][j][k] = f[i];
          else hold[r][cz > 0 ? j : 9 + j] = f[i];
        }
        sf[s3][k] = freg[r];
        wall = wall || is_wall3(freg[r]);
      }
    }
    const unsigned long long vote = __ballot(wall);
    if ((threadIdx.x & 63) == 0) swall[s3][threadIdx.x >> 6] = vote != 0ull;
  };
  // the held classes of plane rel to LDS
  auto put_rest = [&](int rel) {
    const int s2 = rel % 2;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int k = threadIdx.x + r * NT;
      if (k < HA) {
#pragma unroll
        for (int j = 0; j < 9; j++) {
          sup[s2][j][k] = hold[r][j];
          szc[j][k] = hold[r][9 + j];
        }
      }
    }
  };
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int x = x0 + lx, y = y0 + ly;
  const bool inside = x < nx && y < ny;
  const int ctr = (ly + 1) * HX + (lx + 1);
  fetch(z0 - 1);
  put_dn(0);
  put_rest(0);
  fetch(z0);
  put_dn(1);
  put_rest(1);
  fetch(z0 + 1);
  for (int z = z0; z < zend; z++) {
    const int rel = z - z0 + 1;
    const int s_m = (rel - 1) % 3, s_c = rel % 3, s_p = (rel + 1) % 3, s2 = rel % 2;
    put_dn(rel + 1);  // plane z+1 (fetched during the previous iteration)
    __syncthreads();
    if (z + 2 <= zend) fetch(z + 2);  // overlaps the compute below
    // a wall in any of the three planes' staged cells (block-uniform)
    int walls = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) walls |= swall[0][w] | swall[1][w] | swall[2][w];
    if (inside && is_updated(sf[s_c][ctr])) {
      float p[27];
      if (walls) {
#pragma unroll
        for (int i = 0; i < 27; i++) {
          const int cz = D3::cz(i), j = zclass_index(i), up = ctr - D3::cy(i) * HX - D3::cx(i);
          const int sl = cz > 0 ? s_m : (cz < 0 ? s_p : s_c);
          const int ib = D3::inv(i), jb = zclass_index(ib);
          float v;
          if (is_wall3(sf[sl][up])) {
            // bounce-back: the centre's own population in direction inv(i), plane z
            v = cz > 0 ? sdn[s2][jb][ctr] : (cz < 0 ? sup[s2][jb][ctr] : szc[jb][ctr]);
          } else {
            v = cz > 0 ? sup[s2 ^ 1][j][up] : (cz < 0 ? sdn[s2 ^ 1][j][up] : szc[j][up]);
          }
          p[i] = fmaxf(0.0f, v);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 27; i++) {
          const int cz = D3::cz(i), j = zclass_index(i), up = ctr - D3::cy(i) * HX - D3::cx(i);
          p[i] = fmaxf(0.0f, cz > 0 ? sup[s2 ^ 1][j][up] : (cz < 0 ? sdn[s2 ^ 1][j][up] : szc[j][up]));
        }
      }
      float o[10];
      if constexpr (FAST) collide3_fast(p, omega, o);
      else collide3(p, omega, o);
      const size_t cell = ((size_t)z * ny + y) * nx + x;
#pragma unroll
      for (int q = 0; q < 10; q++) Store<S>::template store<3>(out + q * plane + cell, q, o[q], plane);
    }
    __syncthreads();  // plane z-1's c_z = +1 slot and plane z's c_z = 0 slot are free
    put_rest(rel + 1);
  }
}

// 3D column variant with one staged cell per thread.  A block owns a TX x TY
// column of cells and marches it along z over the planes [z0, zend).  Its
// first TX*TY threads own one interior cell each (they stage it and compute
// it); the next 2*(TX+2) + 2*TY threads own the cells of the one-cell xy halo
// ring (they only stage them).  Every staged cell is loaded and reconstructed
// once per plane, and no thread holds more than one staged cell's moments and
// populations in registers.  LDS keeps the population classes by c_z:
//   c_z = +1 (pulled from plane z-1): planes z-1 and z  -> 2 slots
//   c_z =  0 (pulled within plane z): plane z           -> 1 slot
//   c_z = -1 (pulled from plane z+1): plane z+1         -> 1 slot
// 36 population planes of (TX+2)(TY+2) floats.  Per plane: reconstruct plane
// z+1 in registers, barrier, store its c_z = -1 class and the held classes
// of plane z, fetch plane z+2 into registers, barrier, compute plane z.
// The bounce-back of a wall needs the centre's own c_z = -1 class of plane z,
// which the slot no longer holds: blocks with a wall in reach (a block-uniform
// vote) reconstruct it again from the centre's moments (the same arithmetic,
// so the same values).  FAST selects recon3_fast / collide3_fast.
template <int TX, int TY>
struct ColTile {
  static constexpr int HX = TX + 2, HY = TY + 2, HA = HX * HY, NI = TX * TY;
  static constexpr int NT = (HA + 63) / 64 * 64;  // block size
};

// PF planes of moments are in flight while a plane is computed: in registers,
// or (DMA, FP32 storage) in an LDS ring of PF planes filled by LDS DMA loads
// (global_load_lds: the load writes LDS without passing through VGPRs, so
// the planes in flight cost no registers).
template <typename S, bool PERIODIC, bool FAST, int TX, int TY, int PF = 1, bool DMA = false>
__device__ __forceinline__ void bulk3d_col(const S* __restrict__ in, S* __restrict__ out,
                                           const uint8_t* __restrict__ flags, int nx, int ny, int nz, size_t plane,
                                           float omega, int bx, int by, int z0, int zend) {
  using CT = ColTile<TX, TY>;
  constexpr int HX = CT::HX, HY = CT::HY, HA = CT::HA, NI = CT::NI, NW = CT::NT / 64;
  constexpr bool USE_DMA = DMA && std::is_same<S, float>::value;
  static_assert(PF >= 1 && PF <= 4, "planes in flight");
  // 16-bit storage loads through registers: one plane in flight for a DMA shape
  constexpr int NP = (DMA && !USE_DMA) ? 1 : PF;
  __shared__ float ring[USE_DMA ? NP : 1][10][USE_DMA ? CT::NT : 1];  // moments of the planes in flight
  __shared__ float sup[2][9][HA];  // c_z = +1 class, slot rel & 1
  __shared__ float szc[9][HA];     // c_z = 0 class of plane z
  __shared__ float sdn[9][HA];     // c_z = -1 class of plane z + 1
  __shared__ uint8_t sf[3][HA];    // flags, slot rel % 3
  __shared__ int swall[3][NW];     // per flag slot and wave: a staged cell is a wall
  const int t = threadIdx.x;
  __syncthreads();  // a block may run several columns in turn: the last one's reads are done
  // staged cell (hx, hy) of this thread in the (TX+2) x (TY+2) halo frame
  int hx, hy;
  if (t < NI) {
    hx = t % TX + 1;
    hy = t / TX + 1;
  } else {
    const int h = t - NI;
    if (h < HX) {
      hx = h; hy = 0;
    } else if (h < 2 * HX) {
      hx = h - HX; hy = HY - 1;
    } else if (h < 2 * HX + TY) {
      hx = 0; hy = 1 + (h - 2 * HX);
    } else if (h < 2 * HX + 2 * TY) {
      hx = HX - 1; hy = 1 + (h - 2 * HX - TY);
    } else {
      hx = 0; hy = 0;  // idle lane of the last wave
    }
  }
  const bool staged = t < HA;
  const int s = hy * HX + hx;
  const int x0 = bx * TX, y0 = by * TY;
  // in-plane offsets of the staged cell: pull source (clamped or periodic) and flag (wrapped)
  int off_src, off_flag;
  {
    const int gx = x0 + hx - 1, gy = y0 + hy - 1;
    const int wx = ((gx % nx) + nx) % nx, wy = ((gy % ny) + ny) % ny;
    const int sx = PERIODIC ? wx : min(nx - 1, max(0, gx));
    const int sy = PERIODIC ? wy : min(ny - 1, max(0, gy));
    off_src = sy * nx + sx;
    off_flag = wy * nx + wx;
  }
  const size_t nxy = (size_t)nx * ny;
  // NP planes of moments in flight: slot k holds plane z0 + 1 + j for j = k (mod NP)
  float m[USE_DMA ? 1 : NP][10];
  uint8_t fl[NP];
  auto fetch_regs = [&](int gz, float* mm, uint8_t& f) {
    const int wz = ((gz % nz) + nz) % nz;
    const int sz = PERIODIC ? wz : min(nz - 1, max(0, gz));
    if (staged) {
      const size_t c = (size_t)sz * nxy + off_src;
#pragma unroll
      for (int q = 0; q < 10; q++) mm[q] = Store<S>::template load<3>(in + q * plane + c, q, plane);
      f = flags[(size_t)wz * nxy + off_flag];
    } else {
      f = 0;
    }
  };
  auto fetch = [&](int gz, int k) {
    if constexpr (USE_DMA) {
      // every lane loads (idle lanes a valid cell): lane l of wave w lands in ring[k][q][64 w + l]
      const int wz = ((gz % nz) + nz) % nz;
      const int sz = PERIODIC ? wz : min(nz - 1, max(0, gz));
      const float* src = reinterpret_cast<const float*>(in) + (size_t)sz * nxy + off_src;
      const int wbase = t & ~63;
#pragma unroll
      for (int q = 0; q < 10; q++)
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(src + q * plane),
                                         static_cast<void*>(&ring[k][q][wbase]), 4, 0, 0);
      fl[k] = staged ? flags[(size_t)wz * nxy + off_flag] : (uint8_t)0;
    } else {
      fetch_regs(gz, m[k], fl[k]);
    }
  };
  // the moments of the plane in slot k, in registers
  auto moments_of = [&](int k, float* mm) {
    if constexpr (USE_DMA) {
#pragma unroll
      for (int q = 0; q < 10; q++) mm[q] = ring[k][q][t];
    } else {
#pragma unroll
      for (int q = 0; q < 10; q++) mm[q] = m[k][q];
    }
  };
  auto recon = [&](const float* mm, float* f) {
    if constexpr (FAST) recon3_fast(mm, f);
    else recon3_all(mm, f);
  };
  auto vote = [&](int slot, uint8_t f) {
    const unsigned long long v = __ballot(staged && is_wall3(f));
    if ((t & 63) == 0) swall[slot][t >> 6] = v != 0ull;
  };
  float hold[18];  // c_z = +1 (0..8) and c_z = 0 (9..17) classes of the plane in flight
  // with DMA the loop's planes are requested first, so they load during the prologue
  if constexpr (USE_DMA) {
#pragma unroll
    for (int k = 0; k < NP; k++)
      if (z0 + 1 + k <= zend) fetch(z0 + 1 + k, k);
  }
  // prologue: plane z0 - 1 (rel 0) gives its c_z = +1 class, plane z0 (rel 1) is held
  {
    float mp[10], f[27];
    uint8_t fp;
    fetch_regs(z0 - 1, mp, fp);
    recon(mp, f);
    if (staged) {
#pragma unroll
      for (int i = 0; i < 27; i++)
        if (D3::cz(i) > 0) sup[0][zclass_index(i)][s] = f[i];
      sf[0][s] = fp;
    }
    vote(0, fp);
    fetch_regs(z0, mp, fp);
    recon(mp, f);
#pragma unroll
    for (int i = 0; i < 27; i++) {
      const int cz = D3::cz(i), j = zclass_index(i);
      if (cz > 0) hold[j] = f[i];
      else if (cz == 0) hold[9 + j] = f[i];
    }
    if (staged) sf[1][s] = fp;
    vote(1, fp);
  }
  if constexpr (!USE_DMA) {
#pragma unroll
    for (int k = 0; k < NP; k++)
      if (z0 + 1 + k <= zend) fetch(z0 + 1 + k, k);
  }
  const int lx = t % TX, ly = t / TX;
  const int x = x0 + lx, y = y0 + ly;
  const bool inside = t < NI && x < nx && y < ny;
  // Each plane's results are stored at the start of the next plane's step,
  // after its loads are issued.  Loads and stores share one wait counter, and
  // waiting for a load with a store of another kind pending waits for all of
  // them: stored at the end of compute(z), they had to be written back before
  // plane z + 1 could use its moments at the top of the next step.
  float po[10];
  size_t pcell = 0;
  bool pend = false;
  auto flush = [&]() {
    if (pend) {
#pragma unroll
      for (int q = 0; q < 10; q++) Store<S>::template store<3>(out + q * plane + pcell, q, po[q], plane);
    }
    pend = false;
  };
  // one plane: K is the register slot holding plane z + 1 (a constant in each
  // unrolled copy of the loop body)
  auto plane_step = [&](int z, auto KC) {
    constexpr int K = decltype(KC)::value;
    const int rel = z - z0 + 1;
    const int s_m = (rel + 2) % 3, s_c = rel % 3, s_p = (rel + 1) % 3;
    float f[27];
    const uint8_t fz1 = fl[K];
    float mm[10];
    moments_of(K, mm);
    // register slot K is free again: the loads of plane z + 1 + NP are issued
    // first, so they stay in flight over this whole step (both barriers, the
    // reconstruction and compute(z)); the launch is bound by the loads a CU
    // has in flight (issued after the first barrier instead: 116 -> 107 us
    // per 256x128x128 launch of the 64x4x32 column)
    if constexpr (!USE_DMA)
      if (z + 1 + NP <= zend) fetch(z + 1 + NP, K);
    flush();  // plane z - 1
    __syncthreads();  // compute(z - 1) is done with the slots rewritten below
    // plane z + 1 is reconstructed after the barrier, so that its 27 values
    // go to LDS or to the held registers as they are produced (before the
    // barrier they were all live across it, beside the held classes)
    if (staged) {
#pragma unroll
      for (int j = 0; j < 9; j++) {
        sup[rel & 1][j][s] = hold[j];
        szc[j][s] = hold[9 + j];
      }
      sf[s_p][s] = fz1;
    }
    recon(mm, f);  // plane z + 1
    if (staged) {
#pragma unroll
      for (int i = 0; i < 27; i++) {
        const int cz = D3::cz(i), j = zclass_index(i);
        if (cz < 0) sdn[j][s] = f[i];
      }
    }
    vote(s_p, fz1);
#pragma unroll
    for (int i = 0; i < 27; i++) {
      const int cz = D3::cz(i), j = zclass_index(i);
      if (cz > 0) hold[j] = f[i];
      else if (cz == 0) hold[9 + j] = f[i];
    }
    if constexpr (USE_DMA)
      if (z + 1 + NP <= zend) fetch(z + 1 + NP, K);  // in flight while planes z (and on) are computed
    __syncthreads();
    int walls = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) walls |= swall[0][w] | swall[1][w] | swall[2][w];
    if (inside && is_updated(sf[s_c][s])) {
      const int su = (rel - 1) & 1;  // c_z = +1 slot of plane z - 1
      float p[27];
      if (walls) {
        // the centre's own c_z = -1 class of plane z, reconstructed again
        float own[27];
        {
          float mc[10];
          const size_t c = (size_t)z * nxy + (size_t)y * nx + x;
#pragma unroll
          for (int q = 0; q < 10; q++) mc[q] = Store<S>::template load<3>(in + q * plane + c, q, plane);
          recon(mc, own);
        }
#pragma unroll
        for (int i = 0; i < 27; i++) {
          const int cz = D3::cz(i), j = zclass_index(i), up = s - D3::cy(i) * HX - D3::cx(i);
          const int sl = cz > 0 ? s_m : (cz < 0 ? s_p : s_c);
          const int ib = D3::inv(i), jb = zclass_index(ib);
          float v;
          if (is_wall3(sf[sl][up])) {
            // bounce-back: the centre's own population in direction inv(i), plane z
            v = cz > 0 ? own[ib] : (cz < 0 ? sup[rel & 1][jb][s] : szc[jb][s]);
          } else {
            v = cz > 0 ? sup[su][j][up] : (cz < 0 ? sdn[j][up] : szc[j][up]);
          }
          p[i] = fmaxf(0.0f, v);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 27; i++) {
          const int cz = D3::cz(i), j = zclass_index(i), up = s - D3::cy(i) * HX - D3::cx(i);
          p[i] = fmaxf(0.0f, cz > 0 ? sup[su][j][up] : (cz < 0 ? sdn[j][up] : szc[j][up]));
        }
      }
      if constexpr (FAST) collide3_fast(p, omega, po);
      else collide3(p, omega, po);
      pcell = (size_t)z * nxy + (size_t)y * nx + x;
      pend = true;
    }
  };
  for (int z = z0; z < zend; z += NP) {
    plane_step(z, std::integral_constant<int, 0>{});
    if constexpr (NP > 1)
      if (z + 1 < zend) plane_step(z + 1, std::integral_constant<int, (NP > 1 ? 1 : 0)>{});
    if constexpr (NP > 2)
      if (z + 2 < zend) plane_step(z + 2, std::integral_constant<int, (NP > 2 ? 2 : 0)>{});
    if constexpr (NP > 3)
      if (z + 3 < zend) plane_step(z + 3, std::integral_constant<int, (NP > 3 ? 3 : 0)>{});
  }
  flush();  // the last plane
}

}  // namespace ff
