// links_body.h -- device bodies of the boundary-cell, element-sum and finish
// phases, shared by the per-domain kernels (links.hip) and the batched
// kernels (batch.hip).  Each body takes its block index within the domain's
// part of the grid explicitly.
#pragma once

#include "ff_internal.h"
#include "geometry.h"
#include "hermite.h"

namespace ff {

// Virtual lanes of the per-body wrench sum (finish_last_body).
constexpr int kSumLanes = 256;


// ---------------------------------------------------------------- boundary cells
// First barycentric of a (slot, direction) whose link was not applied this
// step.  Applied 3D links can carry a slightly negative barycentric (the
// triangle test accepts u + v <= 1 + 1e-6), so the marker is far outside.
constexpr float kLinkUnused = -1.0e30f;

// Full update of every cell that owns a keyed link, one lane per lattice
// direction: the per-direction logic of mrDeformableSolver2DKernel_Shrd
// (FsiCoupling2D.cu:154-224) / mrDeformableSolver3DKernel_ShrdM
// (FsiCoupling3D.cu:194-285), moving-wall value from the single-population
// reconstruction, momentum-exchange force (FsiCoupling2D.cu:193-214,
// FsiCoupling3D.cu:243-271).  Overwrites the bulk kernel's value for the cell.
// Population i of a cell from its moments m: the reference's rounding
// (recon2_pop / recon3_pop, with the FP16 tile staging in 2D when STAGE),
// or under FAST the reference's all-FP32 single-population form in 2D.
template <int DIM, bool STAGE, bool FAST>
__device__ __forceinline__ float cell_pop(const float* m, int i) {
  if constexpr (DIM == 2) {
    if constexpr (FAST) {
      return recon2_at(m[0], m[1], m[2], m[3], m[5], m[4], i);
    } else {
      const float f = recon2_pop(m[0], m[1], m[2], m[3], m[5], m[4], i);
      if constexpr (STAGE) return __half2float(to_half_rn(f));
      return f;
    }
  } else {
    return recon3_pop(m, i);
  }
}

// The result is written to the cell's side-buffer row (la.sb), not to the
// output moments: the stream-collide tiles of the same launch also write the
// cell, and the element-sum launch copies the side buffer over them.
// Each block scans la.bc_span (<= NT) link-table slots at a time (grid
// stride nblk * bc_span), compacts the occupied ones in LDS and updates those
// cells CPB at a time.
template <int DIM, typename S, bool STAGE, bool PERIODIC, int NT, bool FAST = false>
__device__ __forceinline__ void bcells_body(const S* __restrict__ in, const LinkArgs& la, int blk, int nblk) {
  constexpr int Q = DIM == 2 ? 9 : 27;
  constexpr int K = Q - 1;
  constexpr int NM = DIM == 2 ? 6 : 10;
  constexpr int LPC = DIM == 2 ? 16 : 32;  // lanes per cell
  constexpr int CPB = NT / LPC;            // cells per block
  constexpr int NW = NT / 64;
  __shared__ float sp[CPB][Q];
  __shared__ int s_slot[NT], s_cell[NT], s_wn[NW], s_upd[CPB];
  const unsigned gen = step_gen(la);
  const unsigned tag = step_tag(la);
  const float* v = la.v[gen & 1];
  const float* u = la.u[gen & 1];
  const int lc = threadIdx.x / LPC, dir = threadIdx.x % LPC;
  const unsigned long long* cells = table_cells(la, tag);
  const unsigned long long* keys = table_keys<DIM>(la, tag);
  const int span = la.bc_span;
  for (int base = blk * span; base < la.cap; base += nblk * span) {
    // occupied slots of this chunk, in slot order
    const int s = base + (int)threadIdx.x;
    const bool scan = (int)threadIdx.x < span && s < la.cap;
    const unsigned long long ent = scan ? cells[s] : 0ull;
    const bool occ = scan && (unsigned)(ent >> 32) == tag;
    const unsigned long long m = __ballot(occ);
    const int w = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63u) == 0) s_wn[w] = (int)__popcll(m);
    __syncthreads();
    int off = 0, count = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
      off += i < w ? s_wn[i] : 0;
      count += s_wn[i];
    }
    if (occ) {
      const int r = off + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      s_slot[r] = s;
      s_cell[r] = (int)(uint32_t)ent;
    }
    __syncthreads();
    for (int c0 = 0; c0 < count; c0 += CPB) {
      const int t = c0 + lc;
      int cell = 0;
      if (t < count && dir < Q) {
        const int slot = s_slot[t];
        cell = s_cell[t];
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
        const uint32_t key = dir > 0 ? lt_key(keys, li, tag) : 0xFFFFFFFFu;
        float mc[NM], mn[NM];
#pragma unroll
        for (int q = 0; q < NM; q++) {
          mc[q] = Store<S>::template load<DIM>(in + q * la.plane + cell, q, la.plane);
          mn[q] = Store<S>::template load<DIM>(in + q * la.plane + sc, q, la.plane);
        }
        // the lane needs two populations: the pulled one and the cell's own opposite one
        const float fn_dir = cell_pop<DIM, STAGE, FAST>(mn, dir);
        const float fc_inv = cell_pop<DIM, STAGE, FAST>(mc, inv);
        // direction 0 pulls from the cell itself: its flag is the cell's own,
        // kept for the collision after the barrier (a load there would come
        // after this block's link stores and wait for their write-back)
        if (dir == 0) s_upd[lc] = is_updated(fup) ? 1 : 0;
        float pop = 0.0f;
        bool done = false;
        bool linked = false;
        if (DIM == 3 && is_wall3(fup)) {  // 3D keeps flag bounce-back (FsiCoupling3D.cu:209-222)
          pop = fmaxf(0.0f, fc_inv);
          done = true;
        }
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
            const float fin = fmaxf(0.0f, fc_inv);
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
            linked = true;
            const float cc[3] = {(float)dx, (float)dy, (float)dz};
#pragma unroll
            for (int c = 0; c < DIM; c++) {
              la.lf[li * DIM + c] = (fin * (-cc[c] - uw[c]) - fout * (cc[c] - uw[c])) * la.cf;
              la.lb[li * DIM + c] = h.b[c];
              la.lh[li * DIM + c] = h.hit[c];
            }
          }
        }
        // after every load of the lane: a store before the hit test's loads
        // would make their wait include its write-back
        if (dir > 0 && !linked) la.lb[li * DIM] = kLinkUnused;
        if (!done) pop = fmaxf(0.0f, fn_dir);
        sp[lc][dir] = pop;
      }
      __syncthreads();
      if (t < count && dir == 0 && s_upd[lc]) {
        float p[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) p[i] = sp[lc][i];
        float o[NM];
        if constexpr (DIM == 2 && FAST) collide2_fast(p, la.omega, o);
        else if constexpr (DIM == 2) collide2(p, la.omega, o);
        else if constexpr (FAST) collide3_fast(p, la.omega, o);
        else collide3(p, la.omega, o);
        float* row = la.sb + (size_t)s_slot[t] * NM;
#pragma unroll
        for (int q = 0; q < NM; q++) row[q] = o[q];
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- finalize
struct FinalArgs {
  const int* elem_body;
  const double* body_ref;
  int n_bodies;
  double* elem_acc;
  double* wrench;
  double* hist;
  double* power;       // per body, latest step
  double* power_hist;  // ring of hist_cap steps (with hist)
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
  int wave_fast;  // FP32 kinematics (fast math)
  double dt;
  float* vn[2];
  float* un[2];
  void* buf;
  const OutletCell* outlets;
  int n_outlets;
  int outlet_blocks;
  int copy_blocks;  // copy-back blocks of the sums launch
  int kin_blocks;   // stream launch extras: kinematics blocks, then fin_blocks
  int fin_blocks;   // of a deferred finish (vertex blocks + 1) of step fin_gen
  unsigned fin_gen;
};

// Block size of the element-sum launch (2D and 3D) and elements per block
// (one per wave).
constexpr int kSumBlock = 256, kElemPerBlock = kSumBlock / 64;
__host__ __device__ inline int elem_blocks(int ne) { return (ne + kElemPerBlock - 1) / kElemPerBlock; }

// Host: the FinalArgs of a domain (outlets written into `out`, next-step
// kinematics into vn/un, per-vertex forces optionally added to proxy_forces).
FinalArgs make_final_args(Domain& d, void* out, float* vn0, float* vn1, float* un0, float* un1, float* proxy_forces);
// Host: grid sizes of the boundary-cell and copy-back parts of a step.
int boundary_cell_blocks(const Domain& d, int span);
int copyback_blocks(const Domain& d, int fb, int per_thread);
int detect_blocks(const Domain& d, int nt, bool batched);

// Travelling wave along the body axis:
//   y = y_rest + A s sin(theta), v_y = -2 pi f A s cos(theta),
//   theta = 2 pi (s L/lambda - f t) + phi,  t = g dt at step g.
// par holds per body the four FP64 parameters (A, f, phi, L/lambda)
// followed by eight FP32 values derived from them on the host
// (wave_fp32_params); with fast32 the phase is reduced in cycles in FP32:
// f t mod 1 comes from g times (f dt) held as a float pair hi + lo, the
// product's rounding error recovered with an FMA, so the phase error stays
// about 1e-7 cycle for any g below 2^24.
template <int DIM>
__device__ __forceinline__ void wave_vertex(const float* rest, const float* s, const int* vbody, const double* par,
                                            int fast32, int k, unsigned g, double dt, float* v, float* vel) {
  // every input is read before the first store: a read issued after a store
  // (which may alias it, for the compiler) waits for that store's write-back
  const int b = vbody[k];
  float r[DIM];
#pragma unroll
  for (int c = 0; c < DIM; c++) r[c] = rest[k * DIM + c];
  const float sk32 = s[k];
  float y, vy;
  if (fast32) {
    const float* q = (const float*)(par + kWaveParStride * b + 4);
    const float A = q[0], kl = q[1], phic = q[2], hi = q[3], lo = q[4], vamp = q[5];
    const float gf = (float)g;
    const float pr = gf * hi;
    const float er = fmaf(gf, hi, -pr);
    const float tc = (pr - floorf(pr)) + fmaf(gf, lo, er);  // f t mod 1 (up to a whole cycle)
    float cyc = fmaf(sk32, kl, -tc) + phic;
    cyc -= rintf(cyc);
    float sn, cs;
    sincospif(2.0f * cyc, &sn, &cs);
    const float as = A * sk32;
    y = fmaf(as, sn, r[1]);
    vy = -vamp * sk32 * cs;
  } else {
    const double* p = par + kWaveParStride * b;
    const double A = p[0], f = p[1], phi = p[2], kl = p[3];
    const double sk = (double)sk32;
    const double t = (double)g * dt;
    const double two_pi = 6.283185307179586;
    const double th = two_pi * (sk * kl - f * t) + phi;
    double sn, cs;
    sincos(th, &sn, &cs);
    y = (float)((double)r[1] + A * sk * sn);
    vy = (float)(-two_pi * f * A * sk * cs);
  }
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    v[k * DIM + c] = c == 1 ? y : r[c];
    vel[k * DIM + c] = c == 1 ? vy : 0.0f;
  }
}

// Next-step kinematics for one block of vertices (stream launch): the vertex
// buffers of the next step are not read during this step.
// FAST kernels take the FP32 kinematics (FF_OPT_FAST_MATH always sets wave_fast).
template <int DIM, int FB, bool FAST = false>
__device__ __forceinline__ void wave_block_body(const LinkArgs& la, const FinalArgs& fa, int blk) {
  const int k = blk * FB + threadIdx.x;
  if (k < la.nv) {
    const unsigned g = step_gen(la);
    wave_vertex<DIM>(fa.rest, fa.s, fa.vbody, fa.wpar, FAST ? 1 : fa.wave_fast, k, g, fa.dt, fa.vn[(g + 1) & 1],
                     fa.un[(g + 1) & 1]);
  }
}

// Detection of the next step's links (tag + 1, the next step's vertex
// positions) into the other link table.
template <int DIM, bool BATCHED>
__device__ __forceinline__ void detect_next_body(const LinkArgs& la, int blk) {
  const unsigned gen = step_gen(la);
  if constexpr (BATCHED)
    detect_block<DIM>(la, (long long)blk, step_tag(la) + 1u, la.v[(gen + 1u) & 1u]);
  else
    build_links_pairs<DIM>(la, (long long)blk, step_tag(la) + 1u, la.v[(gen + 1u) & 1u]);
}

// Second launch of a boundary step, kSumBlock threads per block.  The first
// ceil(ne / kElemPerBlock) blocks sum links per element, one element per
// wave: each sums its own links (fixed order, FP64) into elem_acc.  Then
// outlet blocks: outlet cells of the freshly written buffer
// (mrApplyOutlet2DKernel / mrApplyOutlet3DKernel; a neighbour that is a
// boundary cell is read from its side-buffer row, which the copy-back of this
// launch writes).  Then copy-back blocks: boundary-cell rows into the output.
// Any further blocks detect the next step's links.
// Sum of one element's links by one wave (lane = lane index, sl = 64 ints of
// LDS for the wave): window cells are resolved 64 at a time, then every lane
// takes (cell, direction) tasks; a fixed shuffle tree folds the FP64 sums.
// PAIR: a lane's 2D tasks two at a time (a domain stepped on its own, where
// the launch's latency is the slowest wave's); the sums do not depend on it.
template <int DIM, bool PAIR>
__device__ __forceinline__ void elem_sum_wave(const LinkArgs& la, const FinalArgs& fa, int elem, int lane, int* sl) {
  constexpr int K = DIM == 2 ? 8 : 26;
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = elem_stride_v<DIM>;
  const unsigned gen = step_gen(la);
  const unsigned tag = step_tag(la);
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
  const uint32_t etag = (uint32_t)(uint16_t)elem;
  const unsigned long long* cells = table_cells(la, tag);
  const unsigned long long* keys = table_keys<DIM>(la, tag);
  for (int c0 = 0; c0 < ncell; c0 += 64) {
    const int chunk = min(64, ncell - c0);
    if (lane < chunk) {
      const int kk = c0 + lane;
      const int i = lo[0] + kk % wx, j = lo[1] + (kk / wx) % wy, z = DIM == 3 ? lo[2] + kk / (wx * wy) : 0;
      const int cell = (z * la.ny + j) * la.nx + i;
      // the flag and the table probe are independent loads: issue both
      const int found = lt_find(cells, la.cap, tag, cell);
      sl[lane] = is_updated(la.flags[cell]) ? found : -1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // 2D: every load of a task issued before any test, and with PAIR a
    // lane's tasks two at a time (the loads of both overlap), added in task
    // order, so the FP64 sums are those of a one-at-a-time loop.  3D: one task at a time,
    // and a link's force and geometry are loaded only once its key names
    // this element (a 3D window's cells carry mostly other elements' links;
    // loading them all doubled the batched 3D sums launch).
    constexpr int U = DIM == 2 && PAIR ? 2 : 1;
    constexpr bool EARLY = DIM == 2;
    for (int t0 = lane; t0 < K * chunk; t0 += 64 * U) {
      bool use[U];
      uint32_t key[U];
      size_t lis[U];
      float F[U][DIM], B[U][DIM], H[U][DIM];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int task = t0 + 64 * u;
        const int k = task % chunk, l = task / chunk;
        const int slot = task < K * chunk ? sl[k] : -1;
        use[u] = slot >= 0;
        key[u] = 0xFFFFFFFFu;
        lis[u] = 0;
        if (use[u]) {  // every load of the link lies inside the slot's rows
          const size_t li = (size_t)slot * K + l;
          lis[u] = li;
          key[u] = lt_key(keys, li, tag);
          if constexpr (EARLY) {
#pragma unroll
            for (int c = 0; c < DIM; c++) {
              F[u][c] = la.lf[li * DIM + c];
              B[u][c] = la.lb[li * DIM + c];
              H[u][c] = la.lh[li * DIM + c];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (!use[u] || key[u] == 0xFFFFFFFFu || (key[u] & 0xFFFFu) != etag) continue;
        if constexpr (!EARLY) {
          const size_t li = lis[u];
          if (la.lb[li * DIM] == kLinkUnused) continue;
#pragma unroll
          for (int c = 0; c < DIM; c++) {
            F[u][c] = la.lf[li * DIM + c];
            B[u][c] = la.lb[li * DIM + c];
            H[u][c] = la.lh[li * DIM + c];
          }
        }
        if (B[u][0] == kLinkUnused) continue;
#pragma unroll
        for (int jv = 0; jv < DIM; jv++) {
          const float bj = B[u][jv];
#pragma unroll
          for (int c = 0; c < DIM; c++) acc[jv * DIM + c] += (double)(F[u][c] * bj);  // addForceAt products
        }
#pragma unroll
        for (int c = 0; c < DIM; c++) acc[DIM * DIM + c] += (double)F[u][c];
        double r[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < DIM; c++) r[c] = (double)H[u][c] - ref[c];
        if constexpr (DIM == 2) {
          acc[DIM * DIM + DIM] += r[0] * (double)F[u][1] - r[1] * (double)F[u][0];
        } else {
          acc[DIM * DIM + DIM + 0] += r[1] * (double)F[u][2] - r[2] * (double)F[u][1];
          acc[DIM * DIM + DIM + 1] += r[2] * (double)F[u][0] - r[0] * (double)F[u][2];
          acc[DIM * DIM + DIM + 2] += r[0] * (double)F[u][1] - r[1] * (double)F[u][0];
        }
      }
    }
    // the next chunk overwrites sl[] only after every lane's reads above
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // fixed shuffle tree over the wave's 64 lanes
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int a = 0; a < STRIDE; a++) acc[a] += __shfl_down(acc[a], off, 64);
  }
  if (lane == 0) {
    // the element's power: its vertex shares against the velocities the step used
    const float* u = la.u[gen & 1];
    double pw = 0.0;
#pragma unroll
    for (int jv = 0; jv < DIM; jv++) {
      const size_t vid = la.e[(size_t)elem * DIM + jv];
#pragma unroll
      for (int c = 0; c < DIM; c++) pw -= acc[jv * DIM + c] * (double)u[vid * DIM + c];
    }
    acc[STRIDE - 1] = pw;
#pragma unroll
    for (int a = 0; a < STRIDE; a++) fa.elem_acc[(size_t)elem * STRIDE + a] = acc[a];
  }
}


template <int DIM, typename S, int FB, bool BATCHED>
__device__ __forceinline__ void elem_sums_body(const LinkArgs& la, const FinalArgs& fa, int blk, void* out) {
  static_assert(FB == kSumBlock, "element sums run one element per wave of a kSumBlock block");
  constexpr int K = DIM == 2 ? 8 : 26;
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int NM = DIM == 2 ? 6 : 10;
  constexpr int STRIDE = elem_stride_v<DIM>;
  const int neb = elem_blocks(la.ne);
  if (blk >= neb + fa.outlet_blocks + fa.copy_blocks) {
    detect_next_body<DIM, BATCHED>(la, blk - neb - fa.outlet_blocks - fa.copy_blocks);
    return;
  }
  if (blk >= neb + fa.outlet_blocks) {
    const unsigned tag = step_tag(la);
    const int b = blk - neb - fa.outlet_blocks;
    const int nb = fa.copy_blocks;
    const unsigned long long* cells = table_cells(la, tag);
    S* o = (S*)out;
    for (int slot = b * FB + threadIdx.x; slot < la.cap; slot += nb * FB) {
      const unsigned long long ent = cells[slot];
      if ((unsigned)(ent >> 32) != tag) continue;
      const int cell = (int)(uint32_t)ent;
      if (!is_updated(la.flags[cell])) continue;
      // the whole row is loaded before the first store: a load issued after a
      // store (which may alias it, for the compiler) waits for that store's
      // write-back
      const float* row = la.sb + (size_t)slot * NM;
      float r[NM];
#pragma unroll
      for (int q = 0; q < NM; q++) r[q] = row[q];
#pragma unroll
      for (int q = 0; q < NM; q++) Store<S>::template store<DIM>(o + q * la.plane + cell, q, r[q], la.plane);
    }
    return;
  }
  if (blk >= neb) {
    const int t = (blk - neb) * FB + threadIdx.x;
    if (t >= fa.n_outlets) return;
    S* buf = (S*)out;
    const OutletCell c = fa.outlets[t];
    if (la.flags[c.nb] != kFluid) return;
    const unsigned tag = step_tag(la);
    const int nslot = la.ne > 0 ? lt_find(table_cells(la, tag), la.cap, tag, c.nb) : -1;
    const float up = nslot >= 0 ? la.sb[(size_t)nslot * NM + c.comp]
                                : Store<S>::template load<DIM>(buf + c.comp * la.plane + c.nb, c.comp, la.plane);
    if ((c.sign > 0 && up > 0.0f) || (c.sign < 0 && up < 0.0f))
      Store<S>::template store<DIM>(buf + c.comp * la.plane + c.cell, c.comp, up, la.plane);
    Store<S>::template store<DIM>(buf + c.pcomp * la.plane + c.cell, c.pcomp, up * up, la.plane);
    return;
  }
  __shared__ int s_slot[kElemPerBlock][64];
  const int wv = threadIdx.x >> 6;
  const int elem = blk * kElemPerBlock + wv;
  if (elem < la.ne) elem_sum_wave<DIM, !BATCHED>(la, fa, elem, threadIdx.x & 63, s_slot[wv]);
}

// Vertex blocks sum their incident element shares in CSR order (the
// per-vertex forces addForceAt accumulates); the last block folds element
// sums into body wrenches with a fixed tree (ring slot of step `gen`) and,
// for batched domains, advances gen and tag.
template <int DIM, int FB>
__device__ __forceinline__ void finish_vertex_body(const LinkArgs& la, const FinalArgs& fa, int blk) {
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = elem_stride_v<DIM>;
  const int vtx = blk * FB + threadIdx.x;
  if (vtx >= la.nv) return;
  // the forces being added to are read first, beside the element shares: an
  // interleaved read-add-write chain makes every read wait for the previous
  // write's write-back (the two arrays may alias, for the compiler)
  float fv[DIM], fp[DIM];
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    fv[c] = fa.vforce[vtx * DIM + c];
    fp[c] = fa.proxy_forces ? fa.proxy_forces[vtx * DIM + c] : 0.0f;
  }
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
  for (int c = 0; c < DIM; c++) fa.vforce[vtx * DIM + c] = fv[c] + (float)f[c];
  if (fa.proxy_forces) {
#pragma unroll
    for (int c = 0; c < DIM; c++) fa.proxy_forces[vtx * DIM + c] = fp[c] + (float)f[c];
  }
}

template <int DIM, int FB>
__device__ __forceinline__ void finish_last_body(const LinkArgs& la, const FinalArgs& fa, unsigned gen, bool advance) {
  constexpr int NT = DIM == 2 ? 1 : 3;
  constexpr int STRIDE = elem_stride_v<DIM>;
  constexpr int NW = DIM + NT;
  // The sums run over kSumLanes virtual lanes (lane v: elements v, v + kSumLanes,
  // ... of the body, then a fixed tree), whatever the block size FB, so a
  // body's wrench is the same FP64 number from every launch shape.  Channels:
  // the wrench (NW), then the power.
  constexpr int V = kSumLanes, NC = NW + 1;
  __shared__ double s_red[NC][V];
  for (int bd = 0; bd < fa.n_bodies; bd++) {
    for (int v = threadIdx.x; v < V; v += FB) {
      double a[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) a[c] = 0.0;
      for (int k = fa.bptr[bd] + v; k < fa.bptr[bd + 1]; k += V) {
        const double* ea = fa.elem_acc + (size_t)fa.bidx[k] * STRIDE + DIM * DIM;
#pragma unroll
        for (int c = 0; c < NC; c++) a[c] += ea[c];
      }
#pragma unroll
      for (int c = 0; c < NC; c++) s_red[c][v] = a[c];
    }
    __syncthreads();
    for (int w = V / 2; w > 0; w >>= 1) {
      for (int v = threadIdx.x; v < w; v += FB)
#pragma unroll
        for (int c = 0; c < NC; c++) s_red[c][v] += s_red[c][v + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      fa.power[bd] = s_red[NW][0];
      if (fa.power_hist) fa.power_hist[(size_t)((gen - 1) % (unsigned)fa.hist_cap) * fa.n_bodies + bd] = s_red[NW][0];
    }
    if ((int)threadIdx.x < NW) {
      fa.wrench[bd * NW + threadIdx.x] = s_red[threadIdx.x][0];
      if (fa.hist)
        fa.hist[((size_t)((gen - 1) % (unsigned)fa.hist_cap) * fa.n_bodies + bd) * NW + threadIdx.x] =
            s_red[threadIdx.x][0];
    }
    __syncthreads();
  }
  if (advance && threadIdx.x == 0) {  // batched domains: the device keeps gen and tag
    la.st->gen = gen + 1;
    la.st->tag = la.st->tag + 1;
  }
}

// Extra blocks of the stream launch of a boundary step: next-step
// kinematics (kin_blocks), then, when the previous step's finish was
// deferred to this launch, its vertex blocks and wrench block (fin_blocks).
// The previous step's element sums are complete (earlier launch) and this
// step's are not started (later launch).
template <int DIM, int NT, bool FAST = false>
__device__ __forceinline__ void stream_extra_body(const LinkArgs& la, const FinalArgs& fa, int blk) {
  if (blk < fa.kin_blocks) {
    wave_block_body<DIM, NT, FAST>(la, fa, blk);
    return;
  }
  blk -= fa.kin_blocks;
  if (blk < fa.fin_blocks - 1) finish_vertex_body<DIM, NT>(la, fa, blk);
  else finish_last_body<DIM, NT>(la, fa, fa.fin_gen, false);
}

}  // namespace ff
