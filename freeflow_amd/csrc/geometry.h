// geometry.h -- link/element predicates and the generation-tagged link table.
//
// Predicates restate SolidGeometryProxy_Device.cuh (line_intersect :89-124,
// intersectAtIndex<2> :211-227, line_triangle_intersect :306-341,
// getBoundingBox :198-209, :424-440, depthIndexCompact :79-85).  They take the
// same discrete decisions as the reference only without FMA contraction and
// with IEEE division/square root, which is how every unit including this file
// is compiled.
//
// Link table: open addressing keyed by cell.  Entries carry the step's
// generation in their upper 32 bits, so entries of earlier steps read as
// empty and the table never needs clearing:
//   cell word : (gen << 32) | cell
//   key word  : (gen << 32) | ~depthIndexCompact   (atomicMax keeps the
//               current generation and, within it, the smallest key)
#pragma once

#include <cfloat>
#include <stdint.h>

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

// Generation and tag of the step being run (see LinkArgs).
__device__ __forceinline__ unsigned step_gen(const LinkArgs& la) { return la.dev_state ? la.st->gen : la.gen; }
__device__ __forceinline__ unsigned step_tag(const LinkArgs& la) { return la.dev_state ? la.st->tag : la.tag; }

// Two link tables alternate by tag parity: the next step's links are detected
// into one while the current step's are read from the other.
__device__ __forceinline__ unsigned long long* table_cells(const LinkArgs& la, unsigned tag) {
  return la.cells + (size_t)(tag & 1u) * la.cap;
}
template <int DIM>
__device__ __forceinline__ unsigned long long* table_keys(const LinkArgs& la, unsigned tag) {
  return la.keys + (size_t)(tag & 1u) * la.cap * (DIM == 2 ? 8 : 26);
}

// Slot of `cell` under generation `gen`, entering it if absent.  The cells of
// a generation are found again by scanning the slots for its tag (no list, so
// no counter that every detecting lane would serialise on).
__device__ __forceinline__ int lt_claim(unsigned long long* cells, int cap, unsigned gen, int cell, int* overflow) {
  const unsigned long long want = ((unsigned long long)gen << 32) | (unsigned)cell;
  unsigned h = mix32((unsigned)cell) & (unsigned)(cap - 1);
  for (int probe = 0; probe < cap; probe++) {
    unsigned long long cur = cells[h];
    while ((unsigned)(cur >> 32) != gen) {  // stale or never used: try to claim
      const unsigned long long prev = atomicCAS(&cells[h], cur, want);
      if (prev == cur) return (int)h;
      cur = prev;
    }
    if (cur == want) return (int)h;
    h = (h + 1) & (unsigned)(cap - 1);
  }
  atomicOr(overflow, 1);
  return -1;
}

__device__ __forceinline__ int lt_find(const unsigned long long* cells, int cap, unsigned gen, int cell) {
  const unsigned long long want = ((unsigned long long)gen << 32) | (unsigned)cell;
  unsigned h = mix32((unsigned)cell) & (unsigned)(cap - 1);
  for (int probe = 0; probe < cap; probe++) {
    const unsigned long long cur = cells[h];
    if (cur == want) return (int)h;
    if ((unsigned)(cur >> 32) != gen) return -1;
    h = (h + 1) & (unsigned)(cap - 1);
  }
  return -1;
}

__device__ __forceinline__ void lt_offer_key(unsigned long long* keys, size_t idx, unsigned gen, uint32_t key) {
  atomicMax(&keys[idx], ((unsigned long long)gen << 32) | (unsigned)(~key));
}
// 0xFFFFFFFF when the link has no key in this generation.
__device__ __forceinline__ uint32_t lt_key(const unsigned long long* keys, size_t idx, unsigned gen) {
  const unsigned long long v = keys[idx];
  return (unsigned)(v >> 32) == gen ? ~(uint32_t)v : 0xFFFFFFFFu;
}

// depthIndexCompact (:79-85)
__device__ __forceinline__ uint32_t depth_key(float depth, int elem) {
  return (((uint32_t)(depth * 32768.0f)) << 16) | (uint32_t)(uint16_t)elem;
}

// line_intersect (:89-124): link p1->p2 against segment p3-p4; depth = 1 - t
// in (0,1) or -1.  Only the segment's x-extent is tested (:117), so exactly
// vertical segments never hit.
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

struct Hit {
  float depth;
  float b[3];
  float hit[3];
};

// intersectAtIndex<2>: barycentrics from glm::distance (:220-221).
// Segment endpoints (p3, p4) given by value.
__device__ __forceinline__ bool hit_seg2(float p3x, float p3y, float p4x, float p4y, float p1x, float p1y, float p2x,
                                         float p2y, Hit& h) {
  float xo = 0.f, yo = 0.f;
  const float d = seg_link(p1x, p1y, p2x, p2y, p3x, p3y, p4x, p4y, xo, yo);
  if (!(d > 0.0f && d < 1.0f)) return false;
  h.depth = d;
  h.hit[0] = xo;
  h.hit[1] = yo;
  h.hit[2] = 0.0f;
  const float ax = p4x - xo, ay = p4y - yo, bx = p4x - p3x, by = p4y - p3y;
  h.b[0] = sqrtf(ax * ax + ay * ay) / sqrtf(bx * bx + by * by);
  h.b[1] = 1.0f - h.b[0];
  h.b[2] = 0.0f;
  return true;
}

__device__ __forceinline__ bool hit_elem2(const float* v, const uint32_t* e, int elem, float p1x, float p1y, float p2x,
                                          float p2y, Hit& h) {
  const uint32_t s = e[2 * elem], t = e[2 * elem + 1];
  return hit_seg2(v[2 * s], v[2 * s + 1], v[2 * t], v[2 * t + 1], p1x, p1y, p2x, p2y, h);
}

// line_triangle_intersect (:306-341) with the NumericUtils 1e-6 tolerances;
// glm::normalize = v * (1 / sqrt(dot(v, v))).  The triangle's own quantities
// (edges, unit normal, Gram terms) are formed once per element.
struct Tri3 {
  float v0[3], e1[3], e2[3], n[3];
  float d00, d01, d11, den;
};

__device__ __forceinline__ Tri3 tri3(const float* v, const uint32_t* e, int elem) {
  float tv[3][3];
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int k = 0; k < 3; k++) tv[j][k] = v[e[3 * elem + j] * 3 + k];
  Tri3 T;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    T.v0[k] = tv[0][k];
    T.e1[k] = tv[1][k] - tv[0][k];
    T.e2[k] = tv[2][k] - tv[0][k];
  }
  const float c[3] = {T.e1[1] * T.e2[2] - T.e2[1] * T.e1[2], T.e1[2] * T.e2[0] - T.e2[2] * T.e1[0],
                      T.e1[0] * T.e2[1] - T.e2[0] * T.e1[1]};
  const float cs = 1.0f / sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
#pragma unroll
  for (int k = 0; k < 3; k++) T.n[k] = c[k] * cs;
  T.d00 = T.e1[0] * T.e1[0] + T.e1[1] * T.e1[1] + T.e1[2] * T.e1[2];
  T.d01 = T.e1[0] * T.e2[0] + T.e1[1] * T.e2[1] + T.e1[2] * T.e2[2];
  T.d11 = T.e2[0] * T.e2[0] + T.e2[1] * T.e2[1] + T.e2[2] * T.e2[2];
  T.den = T.d00 * T.d11 - T.d01 * T.d01;
  return T;
}

// The plane parameter of the link p1 -> p1 + dd is num / den with
// num = n . (v0 - p1) and den = n . dd.
__device__ __forceinline__ float tri3_num(const Tri3& T, const float* p1) {
  const float a[3] = {T.v0[0] - p1[0], T.v0[1] - p1[1], T.v0[2] - p1[2]};
  return T.n[0] * a[0] + T.n[1] * a[1] + T.n[2] * a[2];
}
__device__ __forceinline__ float tri3_den(const Tri3& T, const float* dd) {
  return T.n[0] * dd[0] + T.n[1] * dd[1] + T.n[2] * dd[2];
}

// The rest of the test for a link whose plane parameter t passed 0 <= t <= 1
// (or is NaN): parallel-link rejection, then the barycentric inside test.
__device__ __forceinline__ bool tri3_finish(const Tri3& T, const float* p1, const float* dd, float t, Hit& h) {
  const float ds = 1.0f / sqrtf(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
  const float d[3] = {dd[0] * ds, dd[1] * ds, dd[2] * ds};
  if (fabsf(T.n[0] * d[0] + T.n[1] * d[1] + T.n[2] * d[2]) <= 1e-6f) return false;
  const float ip[3] = {p1[0] + t * dd[0], p1[1] + t * dd[1], p1[2] + t * dd[2]};
  const float w[3] = {ip[0] - T.v0[0], ip[1] - T.v0[1], ip[2] - T.v0[2]};
  const float d20 = w[0] * T.e1[0] + w[1] * T.e1[1] + w[2] * T.e1[2];
  const float d21 = w[0] * T.e2[0] + w[1] * T.e2[1] + w[2] * T.e2[2];
  const float u0 = (T.d11 * d20 - T.d01 * d21) / T.den;
  const float v0 = (T.d00 * d21 - T.d01 * d20) / T.den;
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

// Link p1 -> p2 against a triangle: every test of the reference, the plane
// parameter first (a conjunction of the same predicates on the same values).
__device__ __forceinline__ bool hit_tri3(const Tri3& T, const float* p1, const float* p2, Hit& h) {
  const float dd[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const float t = tri3_num(T, p1) / tri3_den(T, dd);
  if (t < 0.0f || t > 1.0f) return false;
  return tri3_finish(T, p1, dd, t, h);
}

__device__ __forceinline__ bool hit_elem3(const float* v, const uint32_t* e, int elem, const float* p1, const float* p2,
                                          Hit& h) {
  return hit_tri3(tri3(v, e, elem), p1, p2, h);
}

// Candidate cells of an element: getBoundingBox (max starts at FLT_MIN) / dx,
// widened by one cell, clipped to the grid.  lo/hi inclusive.
template <int DIM>
__device__ __forceinline__ void elem_window(const float* v, const uint32_t* e, int elem, float dx, const int* n, int* lo,
                                            int* hi) {
#pragma unroll
  for (int a = 0; a < DIM; a++) {
    float mn = FLT_MAX, mx = FLT_MIN;
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      const float c = v[DIM * e[DIM * elem + j] + a];
      mn = fminf(mn, c);
      mx = fmaxf(mx, c);
    }
    lo[a] = max((int)(mn / dx) - 1, 0);
    hi[a] = min((int)(mx / dx) + 1, n[a] - 1);
  }
}

}  // namespace ff

#include "ff_internal.h"

namespace ff {

// Lattice vectors of D3Q27 packed two bits per direction (c + 1), for
// direction indices only known at run time.
constexpr unsigned long long d3_pack(int axis) {
  unsigned long long p = 0;
  for (int i = 0; i < 27; i++) {
    const int c = axis == 0 ? D3::cx(i) : (axis == 1 ? D3::cy(i) : D3::cz(i));
    p |= (unsigned long long)(c + 1) << (2 * i);
  }
  return p;
}
__device__ __forceinline__ int d3_c(unsigned long long packed, int i) { return (int)((packed >> (2 * i)) & 3ull) - 1; }

// Lanes that detect one element's links in a batch (many domains: whole
// elements per lane group need fewer blocks for the same work).
template <int DIM>
constexpr int kDetectLanes = DIM == 2 ? 32 : 256;

// 2D link detection: a group of LPE lanes per element walks the element's
// candidate window once for every lattice direction, one (window cell,
// direction) task per lane (the loops of mrSolver2DComputeOccupancyFlagKernel_v3,
// FsiCoupling2D.cu:44-74), and offers depthIndexCompact keys for every hit of
// an updated cell.  `block` indexes groups of blockDim / LPE elements; the
// links are entered under `tag` for the vertex positions `v`.
template <int LPE>
__device__ __forceinline__ void build_links_block2(const LinkArgs& la, long long block, unsigned tag, const float* v) {
  const unsigned gen = tag;
  constexpr int K = 8;
  const long long e = block * (long long)(blockDim.x / LPE) + threadIdx.x / LPE;
  if (e >= (long long)la.ne) return;
  const int elem = (int)e, lane = (int)(threadIdx.x % LPE);
  const int n[3] = {la.nx, la.ny, la.nz};
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  elem_window<2>(v, la.e, elem, la.cl, n, lo, hi);
  const int wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1;
  if (wx <= 0 || wy <= 0) return;
  const int ntask = wx * wy * K;
  const uint32_t sa = la.e[2 * elem], sb = la.e[2 * elem + 1];
  const float p3x = v[2 * sa], p3y = v[2 * sa + 1], p4x = v[2 * sb], p4y = v[2 * sb + 1];
  unsigned long long* cells = table_cells(la, gen);
  unsigned long long* keys = table_keys<2>(la, gen);
  for (int t = lane; t < ntask; t += LPE) {
    const int k = t / K;
    const int l = t - k * K + 1;
    const int i = lo[0] + k % wx, j = lo[1] + k / wx;
    const int cell = j * la.nx + i;
    Hit h;
    bool hit = false;
    if (is_updated(la.flags[cell]))
      hit = hit_seg2(p3x, p3y, p4x, p4y, (float)i * la.cl, (float)j * la.cl, ((float)i - (float)D2::cx(l)) * la.cl,
                     ((float)j - (float)D2::cy(l)) * la.cl, h);
    const int slot = hit ? lt_claim(cells, la.cap, gen, cell, &la.st->overflow) : -1;
    if (slot >= 0) lt_offer_key(keys, (size_t)slot * K + (l - 1), gen, depth_key(h.depth, elem));
  }
}

// Link detection with one 16-lane group per (element, direction) pair, each
// lane one window cell: the shortest latency for a domain stepped on its own
// (the same predicates, hits and keys as build_links_block2 and
// build_links_block3).  `block` indexes
// groups of blockDim / 16 pairs.
template <int DIM>
__device__ __forceinline__ void build_links_pairs(const LinkArgs& la, long long block, unsigned tag, const float* v) {
  const unsigned gen = tag;
  constexpr int K = DIM == 2 ? 8 : 26;
  const int lane = threadIdx.x & 15;
  const long long pair = block * (blockDim.x >> 4) + (threadIdx.x >> 4);
  if (pair >= (long long)la.ne * K) return;
  const int elem = (int)(pair / K);
  const int l = (int)(pair % K) + 1;
  const int n[3] = {la.nx, la.ny, la.nz};
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  elem_window<DIM>(v, la.e, elem, la.cl, n, lo, hi);
  const int wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1, wz = DIM == 3 ? hi[2] - lo[2] + 1 : 1;
  if (wx <= 0 || wy <= 0 || wz <= 0) return;
  const int ncell = wx * wy * wz;
  const int cx = DIM == 2 ? D2::cx(l) : d3_c(d3_pack(0), l);
  const int cy = DIM == 2 ? D2::cy(l) : d3_c(d3_pack(1), l);
  const int cz = DIM == 2 ? 0 : d3_c(d3_pack(2), l);
  unsigned long long* cells = table_cells(la, gen);
  unsigned long long* keys = table_keys<DIM>(la, gen);
  Tri3 T;
  float p3x = 0.f, p3y = 0.f, p4x = 0.f, p4y = 0.f;
  if constexpr (DIM == 3) {
    T = tri3(v, la.e, elem);
  } else {
    const uint32_t sa = la.e[2 * elem], sb = la.e[2 * elem + 1];
    p3x = v[2 * sa];
    p3y = v[2 * sa + 1];
    p4x = v[2 * sb];
    p4y = v[2 * sb + 1];
  }
  for (int k = lane; k < ncell; k += 16) {
    const int i = lo[0] + k % wx;
    const int j = lo[1] + (k / wx) % wy;
    const int z = DIM == 3 ? lo[2] + k / (wx * wy) : 0;
    const int cell = (z * la.ny + j) * la.nx + i;
    Hit h;
    bool hit = false;
    if (is_updated(la.flags[cell])) {
      if constexpr (DIM == 2) {
        hit = hit_seg2(p3x, p3y, p4x, p4y, (float)i * la.cl, (float)j * la.cl, ((float)i - (float)cx) * la.cl,
                       ((float)j - (float)cy) * la.cl, h);
      } else {
        const float p1[3] = {(float)i * la.cl, (float)j * la.cl, (float)z * la.cl};
        const float p2[3] = {((float)i - (float)cx) * la.cl, ((float)j - (float)cy) * la.cl,
                             ((float)z - (float)cz) * la.cl};
        hit = hit_tri3(T, p1, p2, h);
      }
    }
    const int slot = hit ? lt_claim(cells, la.cap, gen, cell, &la.st->overflow) : -1;
    if (slot >= 0) lt_offer_key(keys, (size_t)slot * K + (l - 1), gen, depth_key(h.depth, elem));
  }
}

// Survivor capacity of build_links_block3's LDS lists (an int and a byte
// each).
constexpr int kDet3Cap = 2048;

// Superset test, without a division, of "num / den rounds into [0, 1] or is
// NaN" (the plane-parameter test of line_triangle_intersect): num = 0 gives
// +-0 or NaN; a zero den otherwise gives +-inf; equal signs need
// |num| <= |den| (1 + 2^-24) for the quotient to round to at most 1, which
// fl(|den| * 1.0001) bounds from above (rounding is monotone); opposite signs
// round to -0 only when the quotient underflows, below 1e-30 by far.
__device__ __forceinline__ bool tri3_t_may_pass(float num, float den) {
  if (num == 0.0f) return true;
  if (den == 0.0f) return false;
  const float an = fabsf(num), ad = fabsf(den);
  return (num > 0.0f) == (den > 0.0f) ? an <= ad * 1.0001f : an < ad * 1e-30f;
}

// 3D link detection with one 256-thread block per element, cell-major: each
// thread takes one cell of the element's window per pass (window coordinates
// formed once per cell), forms num = n . (v0 - p1), and only cells within
// 2 cl of the triangle's plane go on (no link of length <= sqrt(3) cl from a
// farther cell reaches it; the margin is far above the rounding of the link
// ends for grids up to millions of cells per axis).  For each of the 26
// directions (unrolled, lattice vectors as constants) links whose plane
// parameter t = num / (n . dd) may lie in [0, 1] (tri3_t_may_pass, no
// division) are gathered in LDS; once the lists may fill, and at the end, the
// block finishes them (t itself, parallel test, barycentrics, table entry).
// Same predicates, hits and keys as build_links_pairs.
__device__ __forceinline__ void build_links_block3(const LinkArgs& la, long long block, unsigned tag, const float* v) {
  constexpr int K = 26;
  constexpr int NT = 256;
  __shared__ int s_n;
  __shared__ int s_cell[kDet3Cap];
  __shared__ unsigned char s_dir[kDet3Cap];
  if (block >= (long long)la.ne) return;  // block-uniform
  const unsigned gen = tag;
  const int elem = (int)block;
  const int n[3] = {la.nx, la.ny, la.nz};
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  elem_window<3>(v, la.e, elem, la.cl, n, lo, hi);
  const int wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1, wz = hi[2] - lo[2] + 1;
  if (wx <= 0 || wy <= 0 || wz <= 0) return;  // block-uniform
  const int ncell = wx * wy * wz;
  const Tri3 T = tri3(v, la.e, elem);
  const float far = 2.0f * la.cl;
  unsigned long long* cells = table_cells(la, gen);
  unsigned long long* keys = table_keys<3>(la, gen);
  const int lane = (int)(threadIdx.x & 63u);
  // finishes the ns gathered links; called block-uniformly, with ns read by
  // every thread between two barriers
  auto flush = [&](int ns) {
    for (int q = (int)threadIdx.x; q < ns; q += NT) {
      const int k = s_cell[q], l = s_dir[q];
      const int i = lo[0] + k % wx, j = lo[1] + (k / wx) % wy, z = lo[2] + k / (wx * wy);
      const float p1[3] = {(float)i * la.cl, (float)j * la.cl, (float)z * la.cl};
      const float p2[3] = {((float)i - (float)d3_c(d3_pack(0), l)) * la.cl,
                           ((float)j - (float)d3_c(d3_pack(1), l)) * la.cl,
                           ((float)z - (float)d3_c(d3_pack(2), l)) * la.cl};
      const float dd[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
      const float t = tri3_num(T, p1) / tri3_den(T, dd);
      Hit h;
      if (!(t < 0.0f || t > 1.0f) && tri3_finish(T, p1, dd, t, h)) {
        const int slot = lt_claim(cells, la.cap, gen, (z * la.ny + j) * la.nx + i, &la.st->overflow);
        if (slot >= 0) lt_offer_key(keys, (size_t)slot * K + (l - 1), gen, depth_key(h.depth, elem));
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
  };
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int c0 = 0; c0 < ncell; c0 += NT) {
    const int k = c0 + (int)threadIdx.x;
    bool near = false;
    float fi = 0.0f, fj = 0.0f, fz = 0.0f, num = 0.0f;
    if (k < ncell) {
      const int i = lo[0] + k % wx, j = lo[1] + (k / wx) % wy, z = lo[2] + k / (wx * wy);
      if (is_updated(la.flags[(z * la.ny + j) * la.nx + i])) {
        fi = (float)i;
        fj = (float)j;
        fz = (float)z;
        const float p1[3] = {fi * la.cl, fj * la.cl, fz * la.cl};
        num = tri3_num(T, p1);
        near = fabsf(num) <= far;
      }
    }
    const float p1[3] = {fi * la.cl, fj * la.cl, fz * la.cl};
#pragma unroll
    for (int l = 1; l <= K; l++) {
      bool keep = false;
      if (near) {
        const float p2[3] = {(fi - (float)d3_c(d3_pack(0), l)) * la.cl, (fj - (float)d3_c(d3_pack(1), l)) * la.cl,
                             (fz - (float)d3_c(d3_pack(2), l)) * la.cl};
        const float dd[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        keep = tri3_t_may_pass(num, tri3_den(T, dd));
      }
      const unsigned long long m = __ballot(keep);
      if (m) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&s_n, __popcll(m));
        base = __shfl(base, 0, 64);
        if (keep) {
          const int idx = base + __popcll(m & ((1ull << lane) - 1ull));
          s_cell[idx] = k;
          s_dir[idx] = (unsigned char)l;
        }
      }
      // at most 4 * NT entries join between checks: finish the lists while
      // that still fits
      if (l % 4 == 0 || l == K) {
        __syncthreads();
        const int ns = s_n;
        __syncthreads();  // read by all before any wave adds again
        if (ns > kDet3Cap - 4 * NT) flush(ns);
      }
    }
  }
  __syncthreads();
  const int ns = s_n;
  if (ns > 0) flush(ns);
}

// Detection block of a batch (blockDim = kDetectLanes<3> = 256): 2D packs
// blockDim / kDetectLanes<2> elements per block (32 lanes per segment: about
// 5 tasks each, fewer waves in flight for the same work than 64), 3D takes one
// element per block.
static_assert(kDetectLanes<3> == 256, "build_links_block3 assumes 256-thread blocks");
template <int DIM>
__device__ __forceinline__ void detect_block(const LinkArgs& la, long long block, unsigned tag, const float* v) {
  if constexpr (DIM == 2)
    build_links_block2<kDetectLanes<2>>(la, block, tag, v);
  else
    build_links_block3(la, block, tag, v);
}

}  // namespace ff
