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
#include "links_body.h"

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

// ---------------------------------------------------------------- per-domain kernels
template <int DIM, typename S, bool STAGE, bool PERIODIC>
__global__ void __launch_bounds__(256) k_bcells(const S* __restrict__ in, S* __restrict__ out, LinkArgs la) {
  bcells_body<DIM, S, STAGE, PERIODIC>(in, out, la, blockIdx.x, gridDim.x);
}

template <int DIM, typename S, int FB>
__global__ void __launch_bounds__(FB) k_elem_sums(LinkArgs la, FinalArgs fa) {
  elem_sums_body<DIM, S, FB>(la, fa, blockIdx.x, fa.buf);
}

// Vertex blocks, then one last block for the body wrenches and the next
// generation (only that block reads or writes the generation).
template <int DIM, int FB>
__global__ void __launch_bounds__(FB) k_finish(LinkArgs la, FinalArgs fa) {
  if ((int)blockIdx.x < (int)gridDim.x - 1) finish_vertex_body<DIM, FB>(la, fa, blockIdx.x);
  else finish_last_body<DIM, FB>(la, fa);
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

FinalArgs make_final_args(Domain& d, void* out, float* vn0, float* vn1, float* un0, float* un1, float* proxy_forces) {
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
  return fa;
}

int launch_finalize(Domain& d, const LinkArgs& la, void* out, float* vn0, float* vn1, float* un0, float* un1,
                    float* proxy_forces) {
  constexpr int FB2 = 128, FB3 = 256;
  const int fb = d.dim == 2 ? FB2 : FB3;
  const FinalArgs fa = make_final_args(d, out, vn0, vn1, un0, un1, proxy_forces);
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
