// links.hip -- moving-boundary coupling: sparse link detection, boundary-cell
// update with moving-wall bounce-back, and deterministic force reduction.
//
// The reference clears and fills a dense N*Q table of packed (depth, element)
// keys every step (cudaMemsetAsync at src/core/FsiCoupling2D.cu:376 and
// mrSolver2DComputeOccupancyFlagKernel_v3 at :31-80; 3D at FsiCoupling3D.cu:
// 31-81, 440) and re-tests each keyed link inside the fused kernel
// (:183-215).  Here the keys live in an open-addressing table keyed by cell,
// sized to the boundary, so the per-step traffic scales with the number of
// boundary links instead of N*Q.  The winning element of a link is the same:
// min over elements of depthIndexCompact (SolidGeometryProxy_Device.cuh:79-85),
// tested over the same candidate cells (element bounding box +-1 cell).
//
// Forces: the reference scatters each link force to the element's vertices
// with float atomicAdd (SolidGeometryProxy_Device.cuh:184-194, 401-421), which
// is order-nondeterministic.  Here every link force is stored, then each
// element sums its own links in a fixed enumeration order, bodies sum their
// elements with a fixed tree and vertices sum their incident elements in CSR
// order, all in FP64.  Results are independent of scheduling.
//
// The predicates below must take the same discrete decisions as the reference,
// so this unit is compiled with -ffp-contract=off and uses IEEE division and
// square root.

#include <cfloat>

#include "ff_internal.h"
#include "hermite.h"

namespace ff {

__device__ __forceinline__ unsigned mix32(unsigned h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ int ht_insert(int cell, int* ht_cell, int cap, int* list, int* counters) {
  unsigned h = mix32((unsigned)cell) & (unsigned)(cap - 1);
  for (int probe = 0; probe < cap; probe++) {
    const int prev = atomicCAS(&ht_cell[h], -1, cell);
    if (prev == -1) {
      const int idx = atomicAdd(&counters[0], 1);
      list[idx] = (int)h;
      return (int)h;
    }
    if (prev == cell) return (int)h;
    h = (h + 1) & (unsigned)(cap - 1);
  }
  atomicOr(&counters[1], 1);
  return -1;
}

__device__ __forceinline__ int ht_find(int cell, const int* ht_cell, int cap) {
  unsigned h = mix32((unsigned)cell) & (unsigned)(cap - 1);
  for (int probe = 0; probe < cap; probe++) {
    const int c = ht_cell[h];
    if (c == cell) return (int)h;
    if (c == -1) return -1;
    h = (h + 1) & (unsigned)(cap - 1);
  }
  return -1;
}

// depthIndexCompact (SolidGeometryProxy_Device.cuh:79-85)
__device__ __forceinline__ uint32_t depth_key(float depth, int elem) {
  return (((uint32_t)(depth * 32768.0f)) << 16) | (uint32_t)(uint16_t)elem;
}

// line_intersect (SolidGeometryProxy_Device.cuh:89-124): link p1->p2 against
// segment p3-p4; returns depth = 1 - t in (0,1) or -1.  Only the segment's
// x-extent is tested (:117), so exactly vertical segments never hit.
__device__ __forceinline__ float seg_link(float x1, float y1, float x2, float y2, float x3, float y3, float x4,
                                          float y4, float& xo, float& yo) {
  const float a1 = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
  const float b1 = (x4 - x2) * (y3 - y4) - (y4 - y2) * (x3 - x4);
  if (fabsf(a1) <= 0.0f) return -1.0f;
  const float t1 = b1 / a1;
  if (t1 <= 0.0f || t1 >= 1.0f) return -1.0f;
  xo = t1 * x1 + (1.0f - t1) * x2;
  if ((x3 - xo) * (x4 - xo) < 0.0f) {
    yo = t1 * y1 + (1.0f - t1) * y2;
    return (float)(1.0 - (double)t1);
  }
  return -1.0f;
}

struct Hit2 {
  float depth, b0, b1, hx, hy;
};

// intersectAtIndex<2> (:211-227); barycentric from glm::distance (:220-221)
__device__ __forceinline__ bool hit_elem2(const float* v, const uint32_t* e, int elem, float p1x, float p1y, float p2x,
                                          float p2y, Hit2& h) {
  const uint32_t s = e[2 * elem], t = e[2 * elem + 1];
  const float p3x = v[2 * s], p3y = v[2 * s + 1], p4x = v[2 * t], p4y = v[2 * t + 1];
  float xo = 0.f, yo = 0.f;
  const float d = seg_link(p1x, p1y, p2x, p2y, p3x, p3y, p4x, p4y, xo, yo);
  if (!(d > 0.0f && d < 1.0f)) return false;
  h.depth = d;
  h.hx = xo;
  h.hy = yo;
  const float ax = p4x - xo, ay = p4y - yo, bx = p4x - p3x, by = p4y - p3y;
  h.b0 = sqrtf(ax * ax + ay * ay) / sqrtf(bx * bx + by * by);
  h.b1 = 1.0f - h.b0;
  return true;
}

struct Hit3 {
  float depth, b[3], hit[3];
};

// line_triangle_intersect (:306-341) with NumericUtils 1e-6 tolerances;
// glm::normalize = v * (1/sqrt(dot(v,v))).
__device__ __forceinline__ bool hit_elem3(const float* v, const uint32_t* e, int elem, const float* p1, const float* p2,
                                          Hit3& h) {
  float tv[3][3];
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int k = 0; k < 3; k++) tv[j][k] = v[e[3 * elem + j] * 3 + k];
  const float e1[3] = {tv[1][0] - tv[0][0], tv[1][1] - tv[0][1], tv[1][2] - tv[0][2]};
  const float e2[3] = {tv[2][0] - tv[0][0], tv[2][1] - tv[0][1], tv[2][2] - tv[0][2]};
  const float c[3] = {e1[1] * e2[2] - e2[1] * e1[2], e1[2] * e2[0] - e2[2] * e1[0], e1[0] * e2[1] - e2[0] * e1[1]};
  const float cs = 1.0f / sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const float n[3] = {c[0] * cs, c[1] * cs, c[2] * cs};
  const float dd[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const float ds = 1.0f / sqrtf(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
  const float d[3] = {dd[0] * ds, dd[1] * ds, dd[2] * ds};
  if (fabsf(n[0] * d[0] + n[1] * d[1] + n[2] * d[2]) <= 1e-6f) return false;
  const float a[3] = {tv[0][0] - p1[0], tv[0][1] - p1[1], tv[0][2] - p1[2]};
  const float t = (n[0] * a[0] + n[1] * a[1] + n[2] * a[2]) / (n[0] * dd[0] + n[1] * dd[1] + n[2] * dd[2]);
  if (t < 0.0f || t > 1.0f) return false;
  const float ip[3] = {p1[0] + t * dd[0], p1[1] + t * dd[1], p1[2] + t * dd[2]};
  const float w[3] = {ip[0] - tv[0][0], ip[1] - tv[0][1], ip[2] - tv[0][2]};
  const float d00 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  const float d01 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
  const float d11 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  const float d20 = w[0] * e1[0] + w[1] * e1[1] + w[2] * e1[2];
  const float d21 = w[0] * e2[0] + w[1] * e2[1] + w[2] * e2[2];
  const float den = d00 * d11 - d01 * d01;
  const float u0 = (d11 * d20 - d01 * d21) / den;
  const float v0 = (d00 * d21 - d01 * d20) / den;
  if (u0 >= 0.0f - 1e-6f && v0 >= 0.0f - 1e-6f && u0 + v0 <= 1.0f + 1e-6f) {
    h.depth = t;
    h.b[0] = 1.0f - u0 - v0;
    h.b[1] = u0;
    h.b[2] = v0;
    h.hit[0] = ip[0];
    h.hit[1] = ip[1];
    h.hit[2] = ip[2];
    return true;
  }
  return false;
}

// Candidate range of getBoundingBox (max starts at FLT_MIN, :198-209) / dx, +-1.
template <int DIM>
__device__ __forceinline__ void elem_range(const float* v, const uint32_t* e, int elem, float dx, int* lo, int* hi) {
#pragma unroll
  for (int a = 0; a < DIM; a++) {
    float mn = FLT_MAX, mx = FLT_MIN;
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      const float c = v[DIM * e[DIM * elem + j] + a];
      mn = fminf(mn, c);
      mx = fmaxf(mx, c);
    }
    lo[a] = (int)(mn / dx) - 1;
    hi[a] = (int)(mx / dx) + 1;
  }
}

// ---------------------------------------------------------------- reset
__global__ void __launch_bounds__(1024) k_links_reset(int* ht_cell, uint32_t* ht_key, const int* list, int* counters,
                                                      int kdir) {
  const int n = counters[0];
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int slot = list[t];
    ht_cell[slot] = -1;
    for (int j = 0; j < kdir; j++) ht_key[(size_t)slot * kdir + j] = 0xFFFFFFFFu;
  }
  __syncthreads();
  if (threadIdx.x == 0) counters[0] = 0;
}

int launch_links_reset(Domain& d) {
  k_links_reset<<<1, 1024, 0, d.stream>>>(d.ht_cell, d.ht_key, d.ht_list, d.counters, d.q - 1);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- build
template <int DIM>
__global__ void __launch_bounds__(256)
    k_links_build(const float* __restrict__ v, const uint32_t* __restrict__ e, int ne, const uint8_t* __restrict__ flags,
                  int nx, int ny, int nz, float dx, int* ht_cell, uint32_t* ht_key, int cap, int* list, int* counters) {
  constexpr int K = DIM == 2 ? 8 : 26;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int elem = (int)(t / K);
  const int l = (int)(t % K) + 1;
  if (elem >= ne) return;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  elem_range<DIM>(v, e, elem, dx, lo, hi);
  const int cx = DIM == 2 ? D2::cx(l) : D3::cx(l);
  const int cy = DIM == 2 ? D2::cy(l) : D3::cy(l);
  const int cz = DIM == 2 ? 0 : D3::cz(l);
  const int zlo = DIM == 2 ? 0 : max(lo[2], 0), zhi = DIM == 2 ? 0 : min(hi[2], nz - 1);
  for (int k = zlo; k <= zhi; k++)
    for (int j = max(lo[1], 0); j <= min(hi[1], ny - 1); j++)
      for (int i = max(lo[0], 0); i <= min(hi[0], nx - 1); i++) {
        const int cell = (k * ny + j) * nx + i;
        if (!is_updated(flags[cell])) continue;
        bool hit;
        float depth;
        if constexpr (DIM == 2) {
          Hit2 h;
          hit = hit_elem2(v, e, elem, (float)i * dx, (float)j * dx, ((float)i - (float)cx) * dx,
                          ((float)j - (float)cy) * dx, h);
          depth = h.depth;
        } else {
          Hit3 h;
          const float p1[3] = {(float)i * dx, (float)j * dx, (float)k * dx};
          const float p2[3] = {((float)i - (float)cx) * dx, ((float)j - (float)cy) * dx, ((float)k - (float)cz) * dx};
          hit = hit_elem3(v, e, elem, p1, p2, h);
          depth = h.depth;
        }
        if (!hit) continue;
        const int slot = ht_insert(cell, ht_cell, cap, list, counters);
        if (slot < 0) return;
        atomicMin(&ht_key[(size_t)slot * K + (l - 1)], depth_key(depth, elem));
      }
}

int launch_links_build(Domain& d, const float* verts, const uint32_t* elems) {
  const int K = d.q - 1;
  const long long nt = (long long)d.ne * K;
  const int b = 256;
  const unsigned g = (unsigned)((nt + b - 1) / b);
  if (g == 0) return FF_OK;
  if (d.dim == 2)
    k_links_build<2><<<g, b, 0, d.stream>>>(verts, elems, d.ne, d.flags, d.nx, d.ny, 1, d.cl, d.ht_cell, d.ht_key,
                                            d.ht_cap, d.ht_list, d.counters);
  else
    k_links_build<3><<<g, b, 0, d.stream>>>(verts, elems, d.ne, d.flags, d.nx, d.ny, d.nz, d.cl, d.ht_cell, d.ht_key,
                                            d.ht_cap, d.ht_list, d.counters);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- boundary cells, 2D
// Full update of a cell that owns at least one keyed link: the per-direction
// logic of mrDeformableSolver2DKernel_Shrd (FsiCoupling2D.cu:154-224) with the
// moving-wall value from CalculateDistributionD2Q9AtIndex and the momentum
// exchange force (:193-214).  Overwrites the bulk kernel's value for the cell.
template <typename S, bool STAGE, bool PERIODIC>
__device__ __forceinline__ void staged_pops2(const S* in, size_t plane, size_t c, float* f) {
  float m[6];
#pragma unroll
  for (int q = 0; q < 6; q++) m[q] = Store<S>::template load<2>(in + q * plane + c, q);
  recon2_all(m[0], m[1], m[2], m[3], m[5], m[4], f);
  if constexpr (STAGE) {
#pragma unroll
    for (int i = 0; i < 9; i++) f[i] = __half2float(to_half_rn(f[i]));
  }
}

template <typename S, bool STAGE, bool PERIODIC>
__global__ void __launch_bounds__(256)
    k_bcells2d(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
               size_t plane, float omega, float cl, float ct, float cf, const float* __restrict__ v,
               const float* __restrict__ vel, const uint32_t* __restrict__ e, const int* __restrict__ ht_cell,
               const uint32_t* __restrict__ ht_key, float* __restrict__ ht_f, float* __restrict__ ht_b,
               float* __restrict__ ht_hit, const int* __restrict__ list, const int* __restrict__ counters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= counters[0]) return;
  const int slot = list[t];
  const int cell = ht_cell[slot];
  const int x = cell % nx, y = cell / nx;
  float mc[6];
#pragma unroll
  for (int q = 0; q < 6; q++) mc[q] = Store<S>::template load<2>(in + q * plane + cell, q);
  float fc[9];
  staged_pops2<S, STAGE, PERIODIC>(in, plane, (size_t)cell, fc);
  float p[9];
  for (int i = 0; i < 9; i++) {
    const int dx = D2::cx(i), dy = D2::cy(i);
    const int x1 = x - dx, y1 = y - dy;
    const bool notCross = x1 >= 0 && x1 < nx && y1 >= 0 && y1 < ny;
    const int xw = (x1 + nx) % nx, yw = (y1 + ny) % ny;
    bool blocked = false;
    const size_t li = (size_t)slot * 8 + (i - 1);
    if (i > 0) ht_b[li * 2] = -1.0f;
    if (i > 0 && notCross && !is_wall2(flags[(size_t)yw * nx + xw])) {
      const uint32_t key = ht_key[li];
      if (key != 0xFFFFFFFFu) {
        const int elem = (int)(key & 0xFFFFu);
        Hit2 h;
        if (hit_elem2(v, e, elem, (float)x * cl, (float)y * cl, (float)xw * cl, (float)yw * cl, h)) {
          blocked = true;
          const uint32_t v1 = e[2 * elem], v2 = e[2 * elem + 1];
          float ux = vel[2 * v1] * h.b0 + vel[2 * v2] * h.b1;
          float uy = vel[2 * v1 + 1] * h.b0 + vel[2 * v2 + 1] * h.b1;
          ux = ux * (ct / cl);
          uy = uy * (ct / cl);
          const float pixx = mc[3] + ux * ux - mc[1] * mc[1];
          const float piyy = mc[4] + uy * uy - mc[2] * mc[2];
          const float pixy = mc[5] + uy * ux - mc[1] * mc[2];
          const float fin = fmaxf(0.0f, fc[D2::inv(i)]);
          const float fout = recon2_at(mc[0], ux, uy, pixx, pixy, piyy, i);
          p[i] = fout;
          const float fx = (fin * ((float)(-dx) - ux) - fout * ((float)dx - ux)) * cf;
          const float fy = (fin * ((float)(-dy) - uy) - fout * ((float)dy - uy)) * cf;
          ht_f[li * 2] = fx;
          ht_f[li * 2 + 1] = fy;
          ht_b[li * 2] = h.b0;
          ht_b[li * 2 + 1] = h.b1;
          ht_hit[li * 2] = h.hx;
          ht_hit[li * 2 + 1] = h.hy;
        }
      }
    }
    if (!blocked) {
      int sx, sy;
      if (PERIODIC) {
        sx = xw;
        sy = yw;
      } else {
        sx = min(nx - 1, max(0, x1));
        sy = min(ny - 1, max(0, y1));
      }
      const size_t sc = (size_t)sy * nx + sx;
      float fn[9];
      if (sc == (size_t)cell) {
#pragma unroll
        for (int k = 0; k < 9; k++) fn[k] = fc[k];
      } else {
        staged_pops2<S, STAGE, PERIODIC>(in, plane, sc, fn);
      }
      float val = fn[0];
#pragma unroll
      for (int k = 1; k < 9; k++)
        if (k == i) val = fn[k];
      p[i] = fmaxf(0.0f, val);
    }
  }
  if (!is_updated(flags[cell])) return;
  float o[6];
  collide2(p, omega, o);
#pragma unroll
  for (int q = 0; q < 6; q++) Store<S>::template store<2>(out + q * plane + cell, q, o[q]);
}

// ---------------------------------------------------------------- boundary cells, 3D
template <typename S>
__device__ __forceinline__ void pops3(const S* in, size_t plane, size_t c, float* f) {
  float m[10];
#pragma unroll
  for (int q = 0; q < 10; q++) m[q] = Store<S>::template load<3>(in + q * plane + c, q);
  recon3_all(m, f);
}

// Per-direction logic of mrDeformableSolver3DKernel_ShrdM (FsiCoupling3D.cu:194-285).
template <typename S, bool PERIODIC>
__global__ void __launch_bounds__(128)
    k_bcells3d(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
               int nz, size_t plane, float omega, float cl, float ct, float cf, const float* __restrict__ v,
               const float* __restrict__ vel, const uint32_t* __restrict__ e, const int* __restrict__ ht_cell,
               const uint32_t* __restrict__ ht_key, float* __restrict__ ht_f, float* __restrict__ ht_b,
               float* __restrict__ ht_hit, const int* __restrict__ list, const int* __restrict__ counters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= counters[0]) return;
  const int slot = list[t];
  const int cell = ht_cell[slot];
  const int x = cell % nx, y = (cell / nx) % ny, z = cell / (nx * ny);
  float mc[10];
#pragma unroll
  for (int q = 0; q < 10; q++) mc[q] = Store<S>::template load<3>(in + q * plane + cell, q);
  float fc[27];
  recon3_all(mc, fc);
  float p[27];
  for (int i = 0; i < 27; i++) {
    const int dx = D3::cx(i), dy = D3::cy(i), dz = D3::cz(i);
    const int x1 = x - dx, y1 = y - dy, z1 = z - dz;
    const bool notCross = x1 >= 0 && x1 < nx && y1 >= 0 && y1 < ny && z1 >= 0 && z1 < nz;
    const int xw = (x1 + nx) % nx, yw = (y1 + ny) % ny, zw = (z1 + nz) % nz;
    const size_t li = (size_t)slot * 26 + (i - 1);
    if (i > 0) ht_b[li * 3] = -1.0f;
    if (is_wall3(flags[((size_t)zw * ny + yw) * nx + xw])) {
      p[i] = fmaxf(0.0f, fc[D3::inv(i)]);
      continue;
    }
    bool blocked = false;
    if (i > 0 && notCross) {
      const uint32_t key = ht_key[li];
      if (key != 0xFFFFFFFFu) {
        const int elem = (int)(key & 0xFFFFu);
        const float p1[3] = {(float)x * cl, (float)y * cl, (float)z * cl};
        const float p2[3] = {(float)xw * cl, (float)yw * cl, (float)zw * cl};
        Hit3 h;
        if (hit_elem3(v, e, elem, p1, p2, h)) {
          blocked = true;
          const uint32_t v1 = e[3 * elem], v2 = e[3 * elem + 1], v3 = e[3 * elem + 2];
          float u[3];
#pragma unroll
          for (int c = 0; c < 3; c++)
            u[c] = ((vel[3 * v1 + c] * h.b[0] + vel[3 * v2 + c] * h.b[1] + vel[3 * v3 + c] * h.b[2]) * ct) / cl;
          float ms[10];
          ms[0] = mc[0];
          ms[1] = u[0];
          ms[2] = u[1];
          ms[3] = u[2];
          ms[4] = mc[4] + u[0] * u[0] - mc[1] * mc[1];
          ms[5] = mc[5] + u[1] * u[1] - mc[2] * mc[2];
          ms[6] = mc[6] + u[2] * u[2] - mc[3] * mc[3];
          ms[7] = mc[7] + u[0] * u[1] - mc[1] * mc[2];
          ms[8] = mc[8] + u[1] * u[2] - mc[2] * mc[3];
          ms[9] = mc[9] + u[0] * u[2] - mc[1] * mc[3];
          const float fout = recon3_at(ms, i);
          p[i] = fout;
          const float fin = fmaxf(0.0f, fc[D3::inv(i)]);
          const float cc[3] = {(float)dx, (float)dy, (float)dz};
#pragma unroll
          for (int c = 0; c < 3; c++) {
            ht_f[li * 3 + c] = (fin * (-cc[c] - u[c]) - fout * (cc[c] - u[c])) * cf;
            ht_b[li * 3 + c] = h.b[c];
            ht_hit[li * 3 + c] = h.hit[c];
          }
        }
      }
    }
    if (!blocked) {
      int sx, sy, sz;
      if (PERIODIC) {
        sx = xw; sy = yw; sz = zw;
      } else {
        sx = min(nx - 1, max(0, x1));
        sy = min(ny - 1, max(0, y1));
        sz = min(nz - 1, max(0, z1));
      }
      const size_t sc = ((size_t)sz * ny + sy) * nx + sx;
      float val;
      if (sc == (size_t)cell) {
        val = fc[i];
      } else {
        float fn[27];
        pops3<S>(in, plane, sc, fn);
        val = fn[0];
#pragma unroll
        for (int k = 1; k < 27; k++)
          if (k == i) val = fn[k];
      }
      p[i] = fmaxf(0.0f, val);
    }
  }
  if (!is_updated(flags[cell])) return;
  float o[10];
  collide3(p, omega, o);
#pragma unroll
  for (int q = 0; q < 10; q++) Store<S>::template store<3>(out + q * plane + cell, q, o[q]);
}

template <typename S>
static int bcells_typed(Domain& d, const void* in, void* out, const float* verts, const float* vels,
                        const uint32_t* elems) {
  const bool stage = d.options & FF_OPT_STAGE_FP16;
  const bool per = d.options & FF_OPT_PERIODIC;
  const S* i = (const S*)in;
  S* o = (S*)out;
  if (d.dim == 2) {
    const int b = 256, g = (d.ht_cap + b - 1) / b;
#define FF_BC2(ST, PE)                                                                                            \
  k_bcells2d<S, ST, PE><<<g, b, 0, d.stream>>>(i, o, d.flags, d.nx, d.ny, d.plane, d.omega, d.cl, d.ct, d.cf,    \
                                               verts, vels, elems, d.ht_cell, d.ht_key, d.ht_f, d.ht_b, d.ht_hit, \
                                               d.ht_list, d.counters)
    if (stage && per) FF_BC2(true, true);
    else if (stage) FF_BC2(true, false);
    else if (per) FF_BC2(false, true);
    else FF_BC2(false, false);
#undef FF_BC2
  } else {
    const int b = 128, g = (d.ht_cap + b - 1) / b;
#define FF_BC3(PE)                                                                                                 \
  k_bcells3d<S, PE><<<g, b, 0, d.stream>>>(i, o, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega, d.cl, d.ct, d.cf,    \
                                           verts, vels, elems, d.ht_cell, d.ht_key, d.ht_f, d.ht_b, d.ht_hit,     \
                                           d.ht_list, d.counters)
    if (per) FF_BC3(true);
    else FF_BC3(false);
#undef FF_BC3
  }
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_boundary_cells(Domain& d, const void* in, void* out, const float* verts, const float* vels,
                          const uint32_t* elems) {
  switch (d.storage) {
    case FF_STORAGE_F32: return bcells_typed<float>(d, in, out, verts, vels, elems);
    case FF_STORAGE_F16: return bcells_typed<__half>(d, in, out, verts, vels, elems);
    default: return bcells_typed<__hip_bfloat16>(d, in, out, verts, vels, elems);
  }
}

// ---------------------------------------------------------------- forces
// Per element: re-enumerate the element's candidate links in a fixed order,
// keep those whose winning key names this element and that were applied, and
// sum (FP64) the vertex shares F*b_j (the products addForceAt forms in FP32),
// the total force and the torque about the body reference point.
// elem_acc layout per element: DIM*DIM vertex shares, DIM force, (2D: 1, 3D: 3) torque.
template <int DIM>
__global__ void __launch_bounds__(256)
    k_elem_forces(const float* __restrict__ v, const uint32_t* __restrict__ e, const int* __restrict__ elem_body,
                  const double* __restrict__ body_ref, int ne, const uint8_t* __restrict__ flags, int nx, int ny,
                  int nz, float dx, const int* __restrict__ ht_cell, const uint32_t* __restrict__ ht_key,
                  const float* __restrict__ ht_f, const float* __restrict__ ht_b, const float* __restrict__ ht_hit,
                  int cap, double* __restrict__ elem_acc) {
  constexpr int K = DIM == 2 ? 8 : 26;
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = DIM * DIM + DIM + NT;
  const int elem = blockIdx.x * blockDim.x + threadIdx.x;
  if (elem >= ne) return;
  double acc[STRIDE];
#pragma unroll
  for (int a = 0; a < STRIDE; a++) acc[a] = 0.0;
  const int b = elem_body ? elem_body[elem] : 0;
  double ref[3] = {body_ref[b * DIM], body_ref[b * DIM + 1], DIM == 3 ? body_ref[b * DIM + DIM - 1] : 0.0};
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  elem_range<DIM>(v, e, elem, dx, lo, hi);
  const int zlo = DIM == 2 ? 0 : max(lo[2], 0), zhi = DIM == 2 ? 0 : min(hi[2], nz - 1);
  const uint32_t tag = (uint32_t)(uint16_t)elem;
  for (int k = zlo; k <= zhi; k++)
    for (int j = max(lo[1], 0); j <= min(hi[1], ny - 1); j++)
      for (int i = max(lo[0], 0); i <= min(hi[0], nx - 1); i++) {
        const int cell = (k * ny + j) * nx + i;
        if (!is_updated(flags[cell])) continue;
        const int slot = ht_find(cell, ht_cell, cap);
        if (slot < 0) continue;
        for (int l = 0; l < K; l++) {
          const size_t li = (size_t)slot * K + l;
          const uint32_t key = ht_key[li];
          if (key == 0xFFFFFFFFu || (key & 0xFFFFu) != tag) continue;
          if (ht_b[li * DIM] < 0.0f) continue;
          float F[DIM];
#pragma unroll
          for (int c = 0; c < DIM; c++) F[c] = ht_f[li * DIM + c];
#pragma unroll
          for (int jv = 0; jv < DIM; jv++) {
            const float bj = ht_b[li * DIM + jv];
#pragma unroll
            for (int c = 0; c < DIM; c++) acc[jv * DIM + c] += (double)(F[c] * bj);
          }
#pragma unroll
          for (int c = 0; c < DIM; c++) acc[DIM * DIM + c] += (double)F[c];
          double r[3];
#pragma unroll
          for (int c = 0; c < DIM; c++) r[c] = (double)ht_hit[li * DIM + c] - ref[c];
          if constexpr (DIM == 2) {
            acc[DIM * DIM + DIM] += r[0] * (double)F[1] - r[1] * (double)F[0];
          } else {
            acc[DIM * DIM + DIM + 0] += r[1] * (double)F[2] - r[2] * (double)F[1];
            acc[DIM * DIM + DIM + 1] += r[2] * (double)F[0] - r[0] * (double)F[2];
            acc[DIM * DIM + DIM + 2] += r[0] * (double)F[1] - r[1] * (double)F[0];
          }
        }
      }
#pragma unroll
  for (int a = 0; a < STRIDE; a++) elem_acc[(size_t)elem * STRIDE + a] = acc[a];
}

// Per body: fixed-order strided sums then a fixed LDS tree.
template <int DIM>
__global__ void __launch_bounds__(256) k_body_wrench(const double* __restrict__ elem_acc, const int* __restrict__ ptr,
                                                     const int* __restrict__ idx, double* __restrict__ wrench,
                                                     double* __restrict__ hist, int hist_slot) {
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = DIM * DIM + DIM + NT;
  constexpr int NW = DIM + NT;
  __shared__ double s[NW][256];
  const int b = blockIdx.x;
  double a[NW];
#pragma unroll
  for (int c = 0; c < NW; c++) a[c] = 0.0;
  for (int k = ptr[b] + threadIdx.x; k < ptr[b + 1]; k += 256) {
    const double* ea = elem_acc + (size_t)idx[k] * STRIDE + DIM * DIM;
#pragma unroll
    for (int c = 0; c < NW; c++) a[c] += ea[c];
  }
#pragma unroll
  for (int c = 0; c < NW; c++) s[c][threadIdx.x] = a[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < NW; c++) s[c][threadIdx.x] += s[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < NW) {
    wrench[b * NW + threadIdx.x] = s[threadIdx.x][0];
    if (hist) hist[((size_t)hist_slot * gridDim.x + b) * NW + threadIdx.x] = s[threadIdx.x][0];
  }
}

// Per vertex: sum incident element shares in CSR order; add to the float
// accumulator(s) (the reference accumulates into proxy.forces).
template <int DIM>
__global__ void __launch_bounds__(256) k_vertex_forces(const double* __restrict__ elem_acc, const int* __restrict__ ptr,
                                                       const int* __restrict__ inc, int nv, float* __restrict__ acc1,
                                                       float* __restrict__ acc2) {
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = DIM * DIM + DIM + NT;
  const int vtx = blockIdx.x * blockDim.x + threadIdx.x;
  if (vtx >= nv) return;
  double f[DIM];
#pragma unroll
  for (int c = 0; c < DIM; c++) f[c] = 0.0;
  for (int k = ptr[vtx]; k < ptr[vtx + 1]; k++) {
    const int code = inc[k];
    const int elem = code >> 2, end = code & 3;
#pragma unroll
    for (int c = 0; c < DIM; c++) f[c] += elem_acc[(size_t)elem * STRIDE + end * DIM + c];
  }
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    if (acc1) acc1[vtx * DIM + c] += (float)f[c];
    if (acc2) acc2[vtx * DIM + c] += (float)f[c];
  }
}

int launch_forces(Domain& d, const float* verts, const uint32_t* elems, float* proxy_forces) {
  if (d.ne == 0) return FF_OK;
  const int b = 256;
  const int ge = (d.ne + b - 1) / b;
  double* hist = nullptr;
  int hslot = 0;
  if (d.hist_cap > 0) {
    hist = d.wrench_hist;
    hslot = (int)(d.hist_count % d.hist_cap);
    d.hist_count++;
  }
  if (d.dim == 2) {
    k_elem_forces<2><<<ge, b, 0, d.stream>>>(verts, elems, d.elem_body, d.body_ref, d.ne, d.flags, d.nx, d.ny, 1,
                                             d.cl, d.ht_cell, d.ht_key, d.ht_f, d.ht_b, d.ht_hit, d.ht_cap, d.elem_acc);
    k_body_wrench<2><<<d.n_bodies, 256, 0, d.stream>>>(d.elem_acc, d.body_elem_ptr, d.body_elem_idx, d.wrench, hist,
                                                       hslot);
    k_vertex_forces<2><<<(d.nv + b - 1) / b, b, 0, d.stream>>>(d.elem_acc, d.vert_inc_ptr, d.vert_inc_idx, d.nv,
                                                               d.vforce, proxy_forces);
  } else {
    k_elem_forces<3><<<ge, b, 0, d.stream>>>(verts, elems, d.elem_body, d.body_ref, d.ne, d.flags, d.nx, d.ny, d.nz,
                                             d.cl, d.ht_cell, d.ht_key, d.ht_f, d.ht_b, d.ht_hit, d.ht_cap, d.elem_acc);
    k_body_wrench<3><<<d.n_bodies, 256, 0, d.stream>>>(d.elem_acc, d.body_elem_ptr, d.body_elem_idx, d.wrench, hist,
                                                       hslot);
    k_vertex_forces<3><<<(d.nv + b - 1) / b, b, 0, d.stream>>>(d.elem_acc, d.vert_inc_ptr, d.vert_inc_idx, d.nv,
                                                               d.vforce, proxy_forces);
  }
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- wave kinematics
// Lateral travelling wave along the body axis, evaluated in FP64:
//   y = y_rest + A s sin(theta), v_y = -2 pi f A s cos(theta),
//   theta = 2 pi (s L/lambda - f t) + phi.
template <int DIM>
__global__ void k_wave(const float* __restrict__ rest, const float* __restrict__ s, const int* __restrict__ vbody,
                       const double* __restrict__ par, int nv, double t, float* __restrict__ v,
                       float* __restrict__ vel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nv) return;
  const double* p = par + 4 * vbody[k];
  const double A = p[0], f = p[1], phi = p[2], kl = p[3];
  const double sk = (double)s[k];
  const double two_pi = 6.283185307179586;
  const double th = two_pi * (sk * kl - f * t) + phi;
  double sn, cs;
  sincos(th, &sn, &cs);
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    v[k * DIM + c] = rest[k * DIM + c];
    vel[k * DIM + c] = 0.0f;
  }
  v[k * DIM + 1] = (float)((double)rest[k * DIM + 1] + A * sk * sn);
  vel[k * DIM + 1] = (float)(-two_pi * f * A * sk * cs);
}

int launch_wave(Domain& d) {
  const int b = 256, g = (d.nv + b - 1) / b;
  const double t = (double)d.step_index * d.dt;
  if (d.dim == 2)
    k_wave<2><<<g, b, 0, d.stream>>>(d.rest_verts, d.s_coord, d.vert_body, d.wave_params, d.nv, t, d.verts, d.vels);
  else
    k_wave<3><<<g, b, 0, d.stream>>>(d.rest_verts, d.s_coord, d.vert_body, d.wave_params, d.nv, t, d.verts, d.vels);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- fillSolid
// resetSolidFlagKernel + updateSolidFlagKernel (FsiCoupling2D.cu:318-354):
// each updated cell becomes SolidDynamic when its centre is inside the
// polygon (even-odd crossing test of isInside, SolidGeometryProxy_Device.cuh:126-144),
// Fluid otherwise.
__global__ void k_fill_solid2d(uint8_t* flags, int nx, int ny, float dx, const float* __restrict__ v,
                               const uint32_t* __restrict__ e, int ne) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const size_t c = (size_t)y * nx + x;
  const uint8_t fl = flags[c];
  if (fl != kFluid && fl != kSolidDynamic) return;
  const float px = ((float)x + 0.5f) * dx, py = ((float)y + 0.5f) * dx;
  bool in = false;
  for (int i = 0; i < ne; i++) {
    const uint32_t a = e[2 * i], b = e[2 * i + 1];
    const float sx = v[2 * a], sy = v[2 * a + 1], tx = v[2 * b], ty = v[2 * b + 1];
    if ((sy < py && ty >= py) || (sy >= py && ty < py)) in ^= (sx + (py - sy) * (tx - sx) / (ty - sy) > px);
  }
  flags[c] = in ? kSolidDynamic : kFluid;
}

int launch_fill_solid(Domain& d) {
  dim3 b(256), g((d.nx + 255) / 256, d.ny);
  k_fill_solid2d<<<g, b, 0, d.stream>>>(d.flags, d.nx, d.ny, d.cl, d.verts, d.elems, d.ne);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

}  // namespace ff
