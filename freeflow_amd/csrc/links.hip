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
// each element then sums its own links in a fixed order (one wave per
// element), and the finish phase sums elements into vertices (CSR order) and
// bodies, all in FP64.  Results do not depend on scheduling.
//
// A step with a boundary is two launches (three when the links must be
// detected first):
//   [detect]  links of this step under the current tag (first step, after a
//             host update of the boundary, proxy steps)
//   stream    boundary cells (into side-buffer rows) + next-step wave
//             kinematics + the previous step's deferred finish (vertex
//             forces, body wrench) + stream-collide tiles (kernels2d.hip /
//             kernels3d.hip)
//   sums      per-element force sums, outlet cells, boundary-cell copy-back,
//             and detection of the next step's links under the next tag (the
//             next step's positions are already known)
// k_step_finish runs the finish on its own only before a reader needs it or
// for proxy steps.  A domain stepped on its own passes generation and tag by
// value (the host keeps them); batch members keep them in StepState on the
// device so that graph replays take identical arguments every step.

#include <algorithm>

#include "ff_internal.h"
#include "geometry.h"
#include "links_body.h"

namespace ff {

// ---------------------------------------------------------------- detect
// Links of the current step under the current tag and vertex buffer.
template <int DIM>
__global__ void __launch_bounds__(256) k_links_build(LinkArgs la) {
  build_links_pairs<DIM>(la, blockIdx.x, step_tag(la), la.v[step_gen(la) & 1]);
}

// Links of the current step of a domain stepped on its own (a batch detects
// its pending members with kb_detect).
int launch_detect(Domain& d, const LinkArgs& la) {
  const unsigned g = (unsigned)detect_blocks(d, 256, false);
  if (g > 0) {
    if (d.dim == 2) k_links_build<2><<<g, 256, 0, d.stream>>>(la);
    else k_links_build<3><<<g, 256, 0, d.stream>>>(la);
  }
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- sums, finish
// Element sums, outlets, copy-back and next-step detection.
template <int DIM, typename S, int FB>
__global__ void __launch_bounds__(FB) k_step_sums(LinkArgs la, FinalArgs fa, void* out) {
  elem_sums_body<DIM, S, FB, false>(la, fa, blockIdx.x, out);
}

// Vertex blocks, then one last block for the body wrenches and the advance
// of gen and tag.  (A separate launch: folding it into the element-sum launch
// would need device-scope release/acquire between blocks on different XCDs,
// i.e. L2 write-back and invalidation, which costs more than the launch.)
template <int DIM, int FB>
__global__ void __launch_bounds__(FB) k_step_finish(LinkArgs la, FinalArgs fa, int nvb) {
  const int b = blockIdx.x;
  if (b < nvb) finish_vertex_body<DIM, FB>(la, fa, b);
  else finish_last_body<DIM, FB>(la, fa, fa.fin_gen, false);
}

// Detection blocks of `nt` threads: batched, one element per kDetectLanes
// lanes; otherwise one (element, direction) pair per 16 lanes.
int detect_blocks(const Domain& d, int nt, bool batched) {
  if (!batched) return (int)(((long long)d.ne * (d.q - 1) + nt / 16 - 1) / (nt / 16));
  const long long per = std::max(1, nt / (d.dim == 2 ? kDetectLanes<2> : kDetectLanes<3>));
  return (int)(((long long)d.ne + per - 1) / per);
}

// Copy-back blocks of the sums launch and boundary-cell blocks of the stream
// launch: both scan the link table's slots (grid stride), so the block
// counts follow its capacity.
int copyback_blocks(const Domain& d, int fb, int per_thread) {
  const long long span = (long long)fb * per_thread;
  return (int)std::min<long long>(256, std::max<long long>(1, ((long long)d.lt_cap + span - 1) / span));
}

// A domain stepped on its own spreads its boundary cells over about 256
// blocks (one per CU, few cells each: the launch's latency is the slowest
// block's), scanning 64..(smallest block size) slots per pass.
int bc_span(const Domain& d) {
  const int nt_min = d.dim == 2 ? 256 : 128;
  int s = 64;
  while (s < nt_min && (long long)s * 256 < d.lt_cap) s *= 2;
  return s;
}

int boundary_cell_blocks(const Domain& d, int span) {
  return (int)std::min<long long>(512, std::max<long long>(1, ((long long)d.lt_cap + span - 1) / span));
}

int launch_step_sums(Domain& d, const LinkArgs& la, void* out, bool detect_next) {
  constexpr int FB = kSumBlock;
  const FinalArgs fa = make_final_args(d, out, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], nullptr);
  const unsigned grid =
      (unsigned)(elem_blocks(d.ne) + fa.outlet_blocks + fa.copy_blocks + (detect_next ? detect_blocks(d, kSumBlock, false) : 0));
#define FF_SUM(DIM, S) k_step_sums<DIM, S, FB><<<grid, FB, 0, d.stream>>>(la, fa, out)
  if (d.dim == 2) {
    if (d.storage == FF_STORAGE_F32) FF_SUM(2, float);
    else if (d.storage == FF_STORAGE_F16) FF_SUM(2, __half);
    else if (d.storage == FF_STORAGE_F16U) FF_SUM(2, h16u);
    else FF_SUM(2, __hip_bfloat16);
  } else {
    if (d.storage == FF_STORAGE_F32) FF_SUM(3, float);
    else if (d.storage == FF_STORAGE_F16) FF_SUM(3, __half);
    else if (d.storage == FF_STORAGE_F16U) FF_SUM(3, h16u);
    else FF_SUM(3, __hip_bfloat16);
  }
#undef FF_SUM
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_step_finish(Domain& d, const LinkArgs& la, unsigned gen, float* proxy_forces) {
  FinalArgs fa = make_final_args(d, nullptr, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], proxy_forces);
  fa.fin_gen = gen;
  const int nvb = (d.nv + 255) / 256;
  const unsigned grid = (unsigned)(nvb + 1);
  if (d.dim == 2) k_step_finish<2, 256><<<grid, 256, 0, d.stream>>>(la, fa, nvb);
  else k_step_finish<3, 256><<<grid, 256, 0, d.stream>>>(la, fa, nvb);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// Run the finish phase of the last step if it was deferred (to the next
// step's stream launch) and something is about to read its results.
int flush_finish(Domain& d) {
  if (!d.finish_pending) return FF_OK;
  d.finish_pending = false;
  return launch_step_finish(d, link_args(d, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], d.elems), d.finish_gen,
                            nullptr);
}

FinalArgs make_final_args(Domain& d, void* out, float* vn0, float* vn1, float* un0, float* un1, float* proxy_forces) {
  FinalArgs fa;
  fa.elem_body = d.elem_body;
  fa.body_ref = d.body_ref;
  fa.n_bodies = d.n_bodies;
  fa.elem_acc = d.elem_acc;
  fa.wrench = d.wrench;
  fa.hist = d.hist_cap > 0 ? d.wrench_hist : nullptr;
  fa.power = d.power;
  fa.power_hist = d.hist_cap > 0 ? d.power_hist : nullptr;
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
  fa.wave_fast = (d.options & FF_OPT_FAST_MATH) ? 1 : 0;
  fa.dt = d.dt;
  fa.vn[0] = vn0;
  fa.vn[1] = vn1;
  fa.un[0] = un0;
  fa.un[1] = un1;
  fa.buf = out;
  fa.outlets = d.outlets;
  fa.n_outlets = d.n_outlets;
  const int fb = kSumBlock;
  fa.outlet_blocks = (d.n_outlets + fb - 1) / fb;
  fa.copy_blocks = (d.has_boundary && d.ne > 0) ? copyback_blocks(d, fb, 1) : 0;
  fa.kin_blocks = 0;
  fa.fin_blocks = 0;
  fa.fin_gen = 0;
  return fa;
}

// Per-vertex forces of the most recent step from the element sums.
template <int DIM>
__global__ void k_step_vertex_forces(const double* __restrict__ elem_acc, const int* __restrict__ vptr,
                                     const int* __restrict__ vinc, int nv, float* __restrict__ out) {
  constexpr int STRIDE = elem_stride_v<DIM>;
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
__global__ void k_wave_now(const float* rest, const float* s, const int* vbody, const double* par, int fast32, int nv,
                           unsigned gstep, double dt, float* v, float* vel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nv) wave_vertex<DIM>(rest, s, vbody, par, fast32, k, gstep, dt, v, vel);
}

// Per-body wave parameters as the device reads them: the four FP64 values,
// then A, L/lambda, phi / 2 pi mod 1, f dt mod 1 as a float pair hi + lo,
// and 2 pi f A, in FP32 (g (f dt) mod 1 = g (f dt mod 1) mod 1 for whole g).
void wave_device_params(const double* p4, double dt, int n_bodies, double* out) {
  const double two_pi = 6.283185307179586;
  for (int b = 0; b < n_bodies; b++) {
    const double* p = p4 + 4 * b;
    double* o = out + kWaveParStride * b;
    const double A = p[0], f = p[1], phi = p[2], kl = p[3];
    o[0] = A; o[1] = f; o[2] = phi; o[3] = kl;
    float* q = (float*)(o + 4);
    double phic = phi / two_pi;
    phic -= std::floor(phic);
    double fdt = f * dt;
    fdt -= std::floor(fdt);
    const float hi = (float)fdt;
    q[0] = (float)A;
    q[1] = (float)kl;
    q[2] = (float)phic;
    q[3] = hi;
    q[4] = (float)(fdt - (double)hi);
    q[5] = (float)(two_pi * f * A);
    q[6] = q[7] = 0.0f;
  }
}

int launch_wave_now(Domain& d) {
  const int b = 256, g = (d.nv + b - 1) / b;
  const unsigned gstep = (unsigned)(d.gen_host - 1);
  const int fast32 = (d.options & FF_OPT_FAST_MATH) ? 1 : 0;
  float* v = d.vbuf[d.gen_host & 1];
  float* u = d.ubuf[d.gen_host & 1];
  if (d.dim == 2)
    k_wave_now<2><<<g, b, 0, d.stream>>>(d.rest_verts, d.s_coord, d.vert_body, d.wave_params, fast32, d.nv, gstep,
                                         d.dt, v, u);
  else
    k_wave_now<3><<<g, b, 0, d.stream>>>(d.rest_verts, d.s_coord, d.vert_body, d.wave_params, fast32, d.nv, gstep,
                                         d.dt, v, u);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- fillSolid
// resetSolidFlagKernel + updateSolidFlagKernel (FsiCoupling2D.cu:318-354):
// an updated cell becomes SolidDynamic when its centre ((x + 0.5) dx,
// (y + 0.5) dx) is inside the boundary by isInside's even-odd ray cast
// (SolidGeometryProxy_Device.cuh:126-144), Fluid otherwise.
//
// Organised by rows instead of cells: a segment that straddles row y's centre
// height (the reference's half-open test) crosses it at xc, computed with the
// reference's expression, and toggles exactly the cells whose centre lies left
// of xc.  Centres are non-decreasing in x, so those cells are a prefix
// x < k of the row; the segment marks bit k - 1 of the row (XOR), and a cell is
// inside when the marks at or right of it have odd parity.  O(segments x rows
// spanned + cells) instead of O(cells x segments), with the same decisions.
__global__ void __launch_bounds__(256) k_fill_clear(uint32_t* __restrict__ bits, size_t nwords) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nwords; i += (size_t)gridDim.x * 256) bits[i] = 0u;
}

// One wave per segment, its lanes over the rows the segment may straddle.
__global__ void __launch_bounds__(256) k_fill_mark(uint32_t* __restrict__ bits, int W, int nx, int ny, float dx,
                                                   const float* __restrict__ v, const uint32_t* __restrict__ e,
                                                   int ne) {
  const int seg = (int)((blockIdx.x * 256u + threadIdx.x) >> 6);
  const int lane = (int)(threadIdx.x & 63u);
  if (seg >= ne) return;
  const uint32_t a = e[2 * seg], b = e[2 * seg + 1];
  const float sx = v[2 * a], sy = v[2 * a + 1], tx = v[2 * b], ty = v[2 * b + 1];
  // rows whose centre height can lie in (min(sy, ty), max(sy, ty)], one row of margin each side
  const float r0 = floorf(fminf(sy, ty) / dx - 0.5f) - 1.0f;
  const float r1 = floorf(fmaxf(sy, ty) / dx - 0.5f) + 1.0f;
  const float f0 = fmaxf(r0, 0.0f), f1 = fminf(r1, (float)(ny - 1));
  if (!(f0 <= f1)) return;  // outside the grid's rows (or not a number)
  const int y0 = (int)f0, y1 = (int)f1;
  for (int y = y0 + lane; y <= y1; y += 64) {
    const float py = ((float)y + 0.5f) * dx;
    if ((sy < py && ty >= py) || (sy >= py && ty < py)) {
      const float xc = sx + (py - sy) * (tx - sx) / (ty - sy);
      // k = number of cells with centre < xc: an estimate, then the exact test
      float fk = ceilf(xc / dx - 0.5f);
      fk = fminf(fmaxf(fk, 0.0f), (float)nx);
      int k = (int)fk;
      while (k > 0 && ((float)(k - 1) + 0.5f) * dx >= xc) k--;
      while (k < nx && ((float)k + 0.5f) * dx < xc) k++;
      if (k > 0) atomicXor(&bits[(size_t)y * W + ((k - 1) >> 5)], 1u << ((k - 1) & 31));
    }
  }
}

// One block per row: suffix parity of the row's marks, right to left in
// chunks of 256 words, then every cell's flag.
__global__ void __launch_bounds__(256) k_fill_apply(uint8_t* __restrict__ flags, const uint32_t* __restrict__ bits,
                                                    int W, int nx) {
  __shared__ uint32_t s_word[256];
  __shared__ uint32_t s_right[256];  // parity of the row's marks in the words right of word i
  __shared__ uint32_t s_wave[4];
  const int y = blockIdx.x;
  const uint32_t* row = bits + (size_t)y * W;
  uint8_t* frow = flags + (size_t)y * nx;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  uint32_t carry = 0;  // parity of the chunks already done (to the right)
  for (int base = (W - 1) / 256 * 256; base >= 0; base -= 256) {
    const int w = base + t;
    const uint32_t word = w < W ? row[w] : 0u;
    const unsigned long long bal = __ballot(__popc(word) & 1u);
    const uint32_t in_wave = lane == 63 ? 0u : (uint32_t)(__popcll(bal >> (lane + 1)) & 1);
    if (lane == 0) s_wave[wv] = (uint32_t)(__popcll(bal) & 1);
    __syncthreads();
    uint32_t above = 0;
    for (int k = wv + 1; k < 4; k++) above ^= s_wave[k];
    s_word[t] = word;
    s_right[t] = carry ^ above ^ in_wave;
    const uint32_t chunk = s_wave[0] ^ s_wave[1] ^ s_wave[2] ^ s_wave[3];
    __syncthreads();
    const int x0 = base * 32;
    for (int i = t; i < 256 * 32; i += 256) {
      const int x = x0 + i;
      if (x >= nx) break;
      const int wl = i >> 5, bit = i & 31;
      const uint32_t inside = s_right[wl] ^ ((uint32_t)__popc(s_word[wl] >> bit) & 1u);
      const uint8_t f = frow[x];
      if (f == kFluid || f == kSolidDynamic) frow[x] = inside ? kSolidDynamic : kFluid;
    }
    carry ^= chunk;
    __syncthreads();
  }
}

int launch_fill_solid(Domain& d, const float* verts) {
  const int W = (d.nx + 31) / 32;
  const size_t nwords = (size_t)d.ny * W;
  k_fill_clear<<<(unsigned)std::min<size_t>((nwords + 255) / 256, 4096), 256, 0, d.stream>>>(d.fill_bits, nwords);
  if (d.ne > 0)
    k_fill_mark<<<(unsigned)((d.ne + 3) / 4), 256, 0, d.stream>>>(d.fill_bits, W, d.nx, d.ny, d.cl, verts, d.elems,
                                                                   d.ne);
  k_fill_apply<<<(unsigned)d.ny, 256, 0, d.stream>>>(d.flags, d.fill_bits, W, d.nx);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

}  // namespace ff
