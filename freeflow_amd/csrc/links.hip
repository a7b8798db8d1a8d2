// links.hip -- moving-boundary coupling: link detection, boundary-cell update
// with moving-wall bounce-back, deterministic force reduction and the
// per-step finalisation (outlet update, next-step kinematics).
//
// The reference clears and fills a dense N*Q table of packed (depth, element)
// keys every step (cudaMemsetAsync at src/core/FsiCoupling2D.cu:376,
// mrSolver2DComputeOccupancyFlagKernel_v3 at :31-80; 3D at FsiCoupling3D.cu:
// 31-81, 440) and re-tests each keyed link inside the fused kernel (:183-215).
// Here the keys live in a generation-tagged open-addressing table sized to
// the boundary (geometry.h), so per-step work scales with the boundary.  The
// winning element of a link is the same: the minimum depthIndexCompact key
// over the same candidate cells.
//
// Forces: the reference scatters every link force to the element's vertices
// with float atomicAdd (SolidGeometryProxy_Device.cuh:184-194, 401-421), whose
// order is nondeterministic.  Here each link force is stored with its link;
// each element then sums its own links in a fixed order (one block per
// element), and one block sums elements into bodies and vertices in CSR
// order, all in FP64.  Results do not depend on scheduling.
//
// A step with a boundary is four launches: stream-collide + link detection
// (kernels2d.hip / kernels3d.hip), boundary cells (k_bcells), element sums +
// outlet (k_elem_sums) and the one-block finish (k_finish).  Per-step values
// (generation, counters, vertex buffer parity) live in StepState on the
// device, so the launches take identical arguments every step.

#include "ff_internal.h"
#include "geometry.h"
#include "hermite.h"

namespace ff {

// ---------------------------------------------------------------- build (standalone)
template <int DIM>
__global__ void __launch_bounds__(256) k_links_build(LinkArgs la) {
  build_links_block<DIM>(la, blockIdx.x, la.st->gen);
}

int launch_links_build(Domain& d, const LinkArgs& la) {
  const int K = d.q - 1;
  const long long pairs = (long long)d.ne * K;
  const unsigned g = (unsigned)((pairs + 15) / 16);
  if (g == 0) return FF_OK;
  if (d.dim == 2) k_links_build<2><<<g, 256, 0, d.stream>>>(la);
  else k_links_build<3><<<g, 256, 0, d.stream>>>(la);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- boundary cells
// Full update of every cell that owns a keyed link, one lane per lattice
// direction: the per-direction logic of mrDeformableSolver2DKernel_Shrd
// (FsiCoupling2D.cu:154-224) / mrDeformableSolver3DKernel_ShrdM
// (FsiCoupling3D.cu:194-285), moving-wall value from the single-population
// reconstruction, momentum-exchange force (FsiCoupling2D.cu:193-214,
// FsiCoupling3D.cu:243-271).  Overwrites the bulk kernel's value for the cell.
template <int DIM, typename S, bool STAGE>
__device__ __forceinline__ void cell_pops(const S* in, size_t plane, size_t c, float* f) {
  if constexpr (DIM == 2) {
    float m[6];
#pragma unroll
    for (int q = 0; q < 6; q++) m[q] = Store<S>::template load<2>(in + q * plane + c, q);
    recon2_all(m[0], m[1], m[2], m[3], m[5], m[4], f);
    if constexpr (STAGE) {
#pragma unroll
      for (int i = 0; i < 9; i++) f[i] = __half2float(to_half_rn(f[i]));
    }
  } else {
    float m[10];
#pragma unroll
    for (int q = 0; q < 10; q++) m[q] = Store<S>::template load<3>(in + q * plane + c, q);
    recon3_all(m, f);
  }
}

template <int DIM, typename S, bool STAGE, bool PERIODIC>
__global__ void __launch_bounds__(256) k_bcells(const S* __restrict__ in, S* __restrict__ out, LinkArgs la) {
  constexpr int Q = DIM == 2 ? 9 : 27;
  constexpr int K = Q - 1;
  constexpr int NM = DIM == 2 ? 6 : 10;
  constexpr int LPC = DIM == 2 ? 16 : 32;  // lanes per cell
  constexpr int CPB = 256 / LPC;           // cells per block
  __shared__ float sp[CPB][Q];
  const unsigned gen = la.st->gen;
  const int count = la.st->count[gen & 1];
  const int* list = la.list + (size_t)(gen & 1) * la.cap;
  const float* v = la.v[gen & 1];
  const float* u = la.u[gen & 1];
  const int lc = threadIdx.x / LPC, dir = threadIdx.x % LPC;
  for (int base = blockIdx.x * CPB; base < count; base += gridDim.x * CPB) {
    const int t = base + lc;
    int cell = 0;
    if (t < count && dir < Q) {
      const int slot = list[t];
      cell = (int)(uint32_t)la.cells[slot];
      const int x = cell % la.nx, y = (cell / la.nx) % la.ny, z = DIM == 3 ? cell / (la.nx * la.ny) : 0;
      const int dx = DIM == 2 ? D2::cx(dir) : D3::cx(dir);
      const int dy = DIM == 2 ? D2::cy(dir) : D3::cy(dir);
      const int dz = DIM == 2 ? 0 : D3::cz(dir);
      const int inv = DIM == 2 ? D2::inv(dir) : D3::inv(dir);
      const int x1 = x - dx, y1 = y - dy, z1 = z - dz;
      const bool notCross = x1 >= 0 && x1 < la.nx && y1 >= 0 && y1 < la.ny && z1 >= 0 && z1 < la.nz;
      const int xw = (x1 + la.nx) % la.nx, yw = (y1 + la.ny) % la.ny, zw = (z1 + la.nz) % la.nz;
      int sx, sy, sz;  // pull source: clamped halo (reference) or periodic
      if (PERIODIC) {
        sx = xw; sy = yw; sz = zw;
      } else {
        sx = min(la.nx - 1, max(0, x1));
        sy = min(la.ny - 1, max(0, y1));
        sz = min(la.nz - 1, max(0, z1));
      }
      const size_t sc = ((size_t)sz * la.ny + sy) * la.nx + sx;
      const size_t li = (size_t)slot * K + (dir - 1);
      // every load that depends only on the cell index is issued up front
      const uint8_t fup = la.flags[((size_t)zw * la.ny + yw) * la.nx + xw];
      const uint32_t key = dir > 0 ? lt_key(la.keys, li, gen) : 0xFFFFFFFFu;
      float mc[NM];
#pragma unroll
      for (int q = 0; q < NM; q++) mc[q] = Store<S>::template load<DIM>(in + q * la.plane + cell, q);
      float fn[Q];
      cell_pops<DIM, S, STAGE>(in, la.plane, sc, fn);
      float fc[Q];
      cell_pops<DIM, S, STAGE>(in, la.plane, (size_t)cell, fc);
      float pop = 0.0f;
      bool done = false;
      if (DIM == 3 && is_wall3(fup)) {  // 3D keeps flag bounce-back (FsiCoupling3D.cu:209-222)
        pop = fmaxf(0.0f, fc[inv]);
        done = true;
      }
      if (dir > 0) la.lb[li * DIM] = -1.0f;
      if (!done && dir > 0 && notCross && (DIM == 3 || !is_wall2(fup)) && key != 0xFFFFFFFFu) {
        const int elem = (int)(key & 0xFFFFu);
        Hit h;
        bool hit;
        if constexpr (DIM == 2) {
          hit = hit_elem2(v, la.e, elem, (float)x * la.cl, (float)y * la.cl, (float)xw * la.cl, (float)yw * la.cl, h);
        } else {
          const float p1[3] = {(float)x * la.cl, (float)y * la.cl, (float)z * la.cl};
          const float p2[3] = {(float)xw * la.cl, (float)yw * la.cl, (float)zw * la.cl};
          hit = hit_elem3(v, la.e, elem, p1, p2, h);
        }
        if (hit) {
          done = true;
          const float fin = fmaxf(0.0f, fc[inv]);
          float uw[3] = {0.f, 0.f, 0.f};
          float fout;
          if constexpr (DIM == 2) {
            // getVelocityAt<2> (:173-180), then `ux *= ct / cl` (FsiCoupling2D.cu:200-201)
            const uint32_t v1 = la.e[2 * elem], v2 = la.e[2 * elem + 1];
            uw[0] = u[2 * v1] * h.b[0] + u[2 * v2] * h.b[1];
            uw[1] = u[2 * v1 + 1] * h.b[0] + u[2 * v2 + 1] * h.b[1];
            uw[0] = uw[0] * (la.ct / la.cl);
            uw[1] = uw[1] * (la.ct / la.cl);
            const float pixx = mc[3] + uw[0] * uw[0] - mc[1] * mc[1];
            const float piyy = mc[4] + uw[1] * uw[1] - mc[2] * mc[2];
            const float pixy = mc[5] + uw[1] * uw[0] - mc[1] * mc[2];
            fout = recon2_at(mc[0], uw[0], uw[1], pixx, pixy, piyy, dir);
          } else {
            // getVelocityAt<3> (:385-398), then `u * ct / cl` (FsiCoupling3D.cu:244)
            const uint32_t v1 = la.e[3 * elem], v2 = la.e[3 * elem + 1], v3 = la.e[3 * elem + 2];
#pragma unroll
            for (int c = 0; c < 3; c++)
              uw[c] = ((u[3 * v1 + c] * h.b[0] + u[3 * v2 + c] * h.b[1] + u[3 * v3 + c] * h.b[2]) * la.ct) / la.cl;
            float ms[10];
            ms[0] = mc[0];
            ms[1] = uw[0];
            ms[2] = uw[1];
            ms[3] = uw[2];
            ms[4] = mc[4] + uw[0] * uw[0] - mc[1] * mc[1];
            ms[5] = mc[5] + uw[1] * uw[1] - mc[2] * mc[2];
            ms[6] = mc[6] + uw[2] * uw[2] - mc[3] * mc[3];
            ms[7] = mc[7] + uw[0] * uw[1] - mc[1] * mc[2];
            ms[8] = mc[8] + uw[1] * uw[2] - mc[2] * mc[3];
            ms[9] = mc[9] + uw[0] * uw[2] - mc[1] * mc[3];
            fout = recon3_at(ms, dir);
          }
          pop = fout;
          const float cc[3] = {(float)dx, (float)dy, (float)dz};
#pragma unroll
          for (int c = 0; c < DIM; c++) {
            la.lf[li * DIM + c] = (fin * (-cc[c] - uw[c]) - fout * (cc[c] - uw[c])) * la.cf;
            la.lb[li * DIM + c] = h.b[c];
            la.lh[li * DIM + c] = h.hit[c];
          }
        }
      }
      if (!done) pop = fmaxf(0.0f, fn[dir]);
      sp[lc][dir] = pop;
    }
    __syncthreads();
    if (t < count && dir == 0 && is_updated(la.flags[cell])) {
      float p[Q];
#pragma unroll
      for (int i = 0; i < Q; i++) p[i] = sp[lc][i];
      float o[NM];
      if constexpr (DIM == 2) collide2(p, la.omega, o);
      else collide3(p, la.omega, o);
#pragma unroll
      for (int q = 0; q < NM; q++) Store<S>::template store<DIM>(out + q * la.plane + cell, q, o[q]);
    }
    __syncthreads();
  }
}

template <int DIM, typename S>
static void bcells_launch(Domain& d, const void* in, void* out, const LinkArgs& la) {
  constexpr int CPB = DIM == 2 ? 16 : 8;
  int blocks = (d.lt_cap + CPB - 1) / CPB;
  blocks = blocks > 512 ? 512 : blocks;
  const bool stage = DIM == 2 && (d.options & FF_OPT_STAGE_FP16);
  const bool per = d.options & FF_OPT_PERIODIC;
  const S* i = (const S*)in;
  S* o = (S*)out;
  if (stage && per) k_bcells<DIM, S, true, true><<<blocks, 256, 0, d.stream>>>(i, o, la);
  else if (stage) k_bcells<DIM, S, true, false><<<blocks, 256, 0, d.stream>>>(i, o, la);
  else if (per) k_bcells<DIM, S, false, true><<<blocks, 256, 0, d.stream>>>(i, o, la);
  else k_bcells<DIM, S, false, false><<<blocks, 256, 0, d.stream>>>(i, o, la);
}

int launch_boundary_cells(Domain& d, const void* in, void* out, const LinkArgs& la) {
  if (d.dim == 2) {
    if (d.storage == FF_STORAGE_F32) bcells_launch<2, float>(d, in, out, la);
    else if (d.storage == FF_STORAGE_F16) bcells_launch<2, __half>(d, in, out, la);
    else bcells_launch<2, __hip_bfloat16>(d, in, out, la);
  } else {
    if (d.storage == FF_STORAGE_F32) bcells_launch<3, float>(d, in, out, la);
    else if (d.storage == FF_STORAGE_F16) bcells_launch<3, __half>(d, in, out, la);
    else bcells_launch<3, __hip_bfloat16>(d, in, out, la);
  }
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- finalize
struct FinalArgs {
  const int* elem_body;
  const double* body_ref;
  int n_bodies;
  double* elem_acc;
  double* wrench;
  double* hist;
  int hist_cap;
  const int* bptr;
  const int* bidx;
  const int* vptr;
  const int* vinc;
  float* vforce;
  float* proxy_forces;
  int wave;
  const float* rest;
  const float* s;
  const int* vbody;
  const double* wpar;
  double dt;
  float* vn[2];
  float* un[2];
  void* buf;
  const OutletCell* outlets;
  int n_outlets;
  int outlet_blocks;
};

// Travelling wave along the body axis, FP64:
//   y = y_rest + A s sin(theta), v_y = -2 pi f A s cos(theta),
//   theta = 2 pi (s L/lambda - f t) + phi.
template <int DIM>
__device__ __forceinline__ void wave_vertex(const float* rest, const float* s, const int* vbody, const double* par,
                                            int k, double t, float* v, float* vel) {
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

// Element blocks: each sums its own links (fixed order, FP64) into elem_acc.
// The next blocks update the outlet cells of the freshly written buffer
// (mrApplyOutlet2DKernel / mrApplyOutlet3DKernel); the last ones evaluate the
// next step's wave kinematics.
template <int DIM, typename S, int FB>
__global__ void __launch_bounds__(FB) k_elem_sums(LinkArgs la, FinalArgs fa) {
  constexpr int K = DIM == 2 ? 8 : 26;
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = DIM * DIM + DIM + NT;
  constexpr int CH = 256;  // window cells resolved per chunk
  if ((int)blockIdx.x >= la.ne + fa.outlet_blocks) {
    // kinematics of the next step (its vertex buffer is not read in this step)
    const int k = ((int)blockIdx.x - la.ne - fa.outlet_blocks) * FB + threadIdx.x;
    if (fa.wave && k < la.nv) {
      const unsigned g = la.st->gen;
      wave_vertex<DIM>(fa.rest, fa.s, fa.vbody, fa.wpar, k, (double)g * fa.dt, fa.vn[(g + 1) & 1], fa.un[(g + 1) & 1]);
    }
    return;
  }
  if ((int)blockIdx.x >= la.ne) {
    const int t = ((int)blockIdx.x - la.ne) * FB + threadIdx.x;
    if (t >= fa.n_outlets) return;
    S* buf = (S*)fa.buf;
    const OutletCell c = fa.outlets[t];
    if (la.flags[c.nb] != kFluid) return;
    const float up = Store<S>::template load<DIM>(buf + c.comp * la.plane + c.nb, c.comp);
    if ((c.sign > 0 && up > 0.0f) || (c.sign < 0 && up < 0.0f))
      Store<S>::template store<DIM>(buf + c.comp * la.plane + c.cell, c.comp, up);
    Store<S>::template store<DIM>(buf + c.pcomp * la.plane + c.cell, c.pcomp, up * up);
    return;
  }
  __shared__ int s_slot[CH];
  __shared__ double s_red[STRIDE][FB];
  const unsigned gen = la.st->gen;
  const int elem = blockIdx.x;
  const float* v = la.v[gen & 1];
  const int n[3] = {la.nx, la.ny, la.nz};
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  elem_window<DIM>(v, la.e, elem, la.cl, n, lo, hi);
  const int wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1, wz = DIM == 3 ? hi[2] - lo[2] + 1 : 1;
  const int ncell = (wx > 0 && wy > 0 && wz > 0) ? wx * wy * wz : 0;
  const int b = fa.elem_body ? fa.elem_body[elem] : 0;
  double ref[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < DIM; c++) ref[c] = fa.body_ref[b * DIM + c];
  double acc[STRIDE];
#pragma unroll
  for (int a = 0; a < STRIDE; a++) acc[a] = 0.0;
  const uint32_t tag = (uint32_t)(uint16_t)elem;
  for (int c0 = 0; c0 < ncell; c0 += CH) {
    const int chunk = min(CH, ncell - c0);
    for (int k = threadIdx.x; k < chunk; k += FB) {
      const int kk = c0 + k;
      const int i = lo[0] + kk % wx, j = lo[1] + (kk / wx) % wy, z = DIM == 3 ? lo[2] + kk / (wx * wy) : 0;
      const int cell = (z * la.ny + j) * la.nx + i;
      s_slot[k] = is_updated(la.flags[cell]) ? lt_find(la.cells, la.cap, gen, cell) : -1;
    }
    __syncthreads();
    for (int task = threadIdx.x; task < K * chunk; task += FB) {
      const int k = task % chunk, l = task / chunk;
      const int slot = s_slot[k];
      if (slot < 0) continue;
      const size_t li = (size_t)slot * K + l;
      const uint32_t key = lt_key(la.keys, li, gen);
      if (key == 0xFFFFFFFFu || (key & 0xFFFFu) != tag) continue;
      if (la.lb[li * DIM] < 0.0f) continue;
      float F[DIM];
#pragma unroll
      for (int c = 0; c < DIM; c++) F[c] = la.lf[li * DIM + c];
#pragma unroll
      for (int jv = 0; jv < DIM; jv++) {
        const float bj = la.lb[li * DIM + jv];
#pragma unroll
        for (int c = 0; c < DIM; c++) acc[jv * DIM + c] += (double)(F[c] * bj);  // addForceAt products
      }
#pragma unroll
      for (int c = 0; c < DIM; c++) acc[DIM * DIM + c] += (double)F[c];
      double r[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < DIM; c++) r[c] = (double)la.lh[li * DIM + c] - ref[c];
      if constexpr (DIM == 2) {
        acc[DIM * DIM + DIM] += r[0] * (double)F[1] - r[1] * (double)F[0];
      } else {
        acc[DIM * DIM + DIM + 0] += r[1] * (double)F[2] - r[2] * (double)F[1];
        acc[DIM * DIM + DIM + 1] += r[2] * (double)F[0] - r[0] * (double)F[2];
        acc[DIM * DIM + DIM + 2] += r[0] * (double)F[1] - r[1] * (double)F[0];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < STRIDE; a++) s_red[a][threadIdx.x] = acc[a];
  __syncthreads();
  for (int w = FB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int a = 0; a < STRIDE; a++) s_red[a][threadIdx.x] += s_red[a][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < STRIDE) fa.elem_acc[(size_t)elem * STRIDE + threadIdx.x] = s_red[threadIdx.x][0];
}

// Vertex blocks sum their incident element shares in CSR order (the
// per-vertex forces addForceAt accumulates); the last block folds element
// sums into body wrenches with a fixed tree and starts the next step's link
// table generation.  Only that block reads or writes the generation.
template <int DIM, int FB>
__global__ void __launch_bounds__(FB) k_finish(LinkArgs la, FinalArgs fa) {
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = DIM * DIM + DIM + NT;
  constexpr int NW = DIM + NT;
  if ((int)blockIdx.x < (int)gridDim.x - 1) {
    const int vtx = blockIdx.x * FB + threadIdx.x;
    if (vtx >= la.nv) return;
    double f[DIM];
#pragma unroll
    for (int c = 0; c < DIM; c++) f[c] = 0.0;
    for (int k = fa.vptr[vtx]; k < fa.vptr[vtx + 1]; k++) {
      const int code = fa.vinc[k];
      const int e = code >> 2, end = code & 3;
#pragma unroll
      for (int c = 0; c < DIM; c++) f[c] += fa.elem_acc[(size_t)e * STRIDE + end * DIM + c];
    }
#pragma unroll
    for (int c = 0; c < DIM; c++) {
      fa.vforce[vtx * DIM + c] += (float)f[c];
      if (fa.proxy_forces) fa.proxy_forces[vtx * DIM + c] += (float)f[c];
    }
    return;
  }
  __shared__ double s_red[NW][FB];
  const unsigned gen = la.st->gen;
  for (int bd = 0; bd < fa.n_bodies; bd++) {
    double a[NW];
#pragma unroll
    for (int c = 0; c < NW; c++) a[c] = 0.0;
    for (int k = fa.bptr[bd] + threadIdx.x; k < fa.bptr[bd + 1]; k += FB) {
      const double* ea = fa.elem_acc + (size_t)fa.bidx[k] * STRIDE + DIM * DIM;
#pragma unroll
      for (int c = 0; c < NW; c++) a[c] += ea[c];
    }
#pragma unroll
    for (int c = 0; c < NW; c++) s_red[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int w = FB / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
#pragma unroll
        for (int c = 0; c < NW; c++) s_red[c][threadIdx.x] += s_red[c][threadIdx.x + w];
      __syncthreads();
    }
    if ((int)threadIdx.x < NW) {
      fa.wrench[bd * NW + threadIdx.x] = s_red[threadIdx.x][0];
      if (fa.hist)
        fa.hist[((size_t)((gen - 1) % (unsigned)fa.hist_cap) * fa.n_bodies + bd) * NW + threadIdx.x] =
            s_red[threadIdx.x][0];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    la.st->count[(gen + 1) & 1] = 0;
    la.st->gen = gen + 1;
  }
}

int launch_finalize(Domain& d, const LinkArgs& la, void* out, float* vn0, float* vn1, float* un0, float* un1,
                    float* proxy_forces) {
  FinalArgs fa;
  fa.elem_body = d.elem_body;
  fa.body_ref = d.body_ref;
  fa.n_bodies = d.n_bodies;
  fa.elem_acc = d.elem_acc;
  fa.wrench = d.wrench;
  fa.hist = d.hist_cap > 0 ? d.wrench_hist : nullptr;
  fa.hist_cap = d.hist_cap > 0 ? d.hist_cap : 1;
  fa.bptr = d.body_elem_ptr;
  fa.bidx = d.body_elem_idx;
  fa.vptr = d.vert_inc_ptr;
  fa.vinc = d.vert_inc_idx;
  fa.vforce = d.vforce;
  fa.proxy_forces = proxy_forces;
  fa.wave = d.wave ? 1 : 0;
  fa.rest = d.rest_verts;
  fa.s = d.s_coord;
  fa.vbody = d.vert_body;
  fa.wpar = d.wave_params;
  fa.dt = d.dt;
  fa.vn[0] = vn0;
  fa.vn[1] = vn1;
  fa.un[0] = un0;
  fa.un[1] = un1;
  fa.buf = out;
  fa.outlets = d.outlets;
  fa.n_outlets = d.n_outlets;
  constexpr int FB2 = 128, FB3 = 256;
  const int fb = d.dim == 2 ? FB2 : FB3;
  fa.outlet_blocks = (d.n_outlets + fb - 1) / fb;
  const int wave_blocks = fa.wave ? (d.nv + fb - 1) / fb : 0;
  const unsigned grid = (unsigned)(d.ne + fa.outlet_blocks + wave_blocks);
#define FF_FIN(DIM, S, FB) k_elem_sums<DIM, S, FB><<<grid, FB, 0, d.stream>>>(la, fa)
  if (d.dim == 2) {
    if (d.storage == FF_STORAGE_F32) FF_FIN(2, float, FB2);
    else if (d.storage == FF_STORAGE_F16) FF_FIN(2, __half, FB2);
    else FF_FIN(2, __hip_bfloat16, FB2);
  } else {
    if (d.storage == FF_STORAGE_F32) FF_FIN(3, float, FB3);
    else if (d.storage == FF_STORAGE_F16) FF_FIN(3, __half, FB3);
    else FF_FIN(3, __hip_bfloat16, FB3);
  }
#undef FF_FIN
  const unsigned fgrid = (unsigned)((d.nv + 255) / 256 + 1);
  if (d.dim == 2) k_finish<2, 256><<<fgrid, 256, 0, d.stream>>>(la, fa);
  else k_finish<3, 256><<<fgrid, 256, 0, d.stream>>>(la, fa);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// Per-vertex forces of the most recent step from the element sums.
template <int DIM>
__global__ void k_step_vertex_forces(const double* __restrict__ elem_acc, const int* __restrict__ vptr,
                                     const int* __restrict__ vinc, int nv, float* __restrict__ out) {
  constexpr int STRIDE = DIM * DIM + DIM + (DIM == 2 ? 1 : 3);
  const int vtx = blockIdx.x * blockDim.x + threadIdx.x;
  if (vtx >= nv) return;
  double f[DIM];
#pragma unroll
  for (int c = 0; c < DIM; c++) f[c] = 0.0;
  for (int k = vptr[vtx]; k < vptr[vtx + 1]; k++) {
    const int e = vinc[k] >> 2, end = vinc[k] & 3;
#pragma unroll
    for (int c = 0; c < DIM; c++) f[c] += elem_acc[(size_t)e * STRIDE + end * DIM + c];
  }
#pragma unroll
  for (int c = 0; c < DIM; c++) out[vtx * DIM + c] = (float)f[c];
}

int launch_step_vertex_forces(Domain& d, float* out) {
  const int b = 256, g = (d.nv + b - 1) / b;
  if (g == 0) return FF_OK;
  if (d.dim == 2) k_step_vertex_forces<2><<<g, b, 0, d.stream>>>(d.elem_acc, d.vert_inc_ptr, d.vert_inc_idx, d.nv, out);
  else k_step_vertex_forces<3><<<g, b, 0, d.stream>>>(d.elem_acc, d.vert_inc_ptr, d.vert_inc_idx, d.nv, out);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// Kinematics of the current step (used when the wave is (re)set).
template <int DIM>
__global__ void k_wave_now(const float* rest, const float* s, const int* vbody, const double* par, int nv, double t,
                           float* v, float* vel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nv) wave_vertex<DIM>(rest, s, vbody, par, k, t, v, vel);
}

int launch_wave_now(Domain& d) {
  const int b = 256, g = (d.nv + b - 1) / b;
  const double t = (double)(d.gen_host - 1) * d.dt;
  float* v = d.vbuf[d.gen_host & 1];
  float* u = d.ubuf[d.gen_host & 1];
  if (d.dim == 2)
    k_wave_now<2><<<g, b, 0, d.stream>>>(d.rest_verts, d.s_coord, d.vert_body, d.wave_params, d.nv, t, v, u);
  else
    k_wave_now<3><<<g, b, 0, d.stream>>>(d.rest_verts, d.s_coord, d.vert_body, d.wave_params, d.nv, t, v, u);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- fillSolid
// resetSolidFlagKernel + updateSolidFlagKernel (FsiCoupling2D.cu:318-354):
// each updated cell becomes SolidDynamic when its centre is inside the
// polygon (even-odd crossing test of isInside, SolidGeometryProxy_Device.cuh:
// 126-144), Fluid otherwise.
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

int launch_fill_solid(Domain& d, const float* verts) {
  dim3 b(256), g((d.nx + 255) / 256, d.ny);
  k_fill_solid2d<<<g, b, 0, d.stream>>>(d.flags, d.nx, d.ny, d.cl, verts, d.elems, d.ne);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

}  // namespace ff
