// kernels2d.hip -- D2Q9 stream+collide over structure-of-arrays moments, the
// outlet update, storage conversion and the FP64 mass reduction.
//
// Replaces mrDeformableSolver2DKernel_Shrd and mrApplyOutlet2DKernel
// (src/core/FsiCoupling2D.cu:82-316).  The bulk kernel treats every updated
// cell as if no moving-wall link existed; the few cells that own a
// moving-wall link are recomputed afterwards by the boundary-cell kernel
// (links.hip), so the bulk pass stays a pure stencil with one byte of flag
// traffic per cell.  With a boundary present the same launch also runs the
// step's link detection in extra workgroups.
//
// In 2D a wall-upstream population is overwritten by the ordinary pull in the
// reference (FsiCoupling2D.cu:168-223: `is_blocked` stays false on the wall
// branch), so the bulk kernel never needs neighbour flags.

#include <type_traits>

#include "ff_internal.h"
#include "geometry.h"
#include "hermite.h"

namespace ff {

constexpr int kTX2 = 32, kTY2 = 8;

// One 32x8 tile: stage the reconstructed populations of the tile and its
// one-cell halo in LDS (clamped or periodic halo), then pull, collide and
// store every updated cell.  `bx, by` select the tile.
template <typename S, bool STAGE, bool PERIODIC>
__device__ __forceinline__ void bulk2d_tile(const S* __restrict__ in, S* __restrict__ out,
                                            const uint8_t* __restrict__ flags, int nx, int ny, size_t plane,
                                            float omega, int bx, int by) {
  constexpr int HX = kTX2 + 2, HY = kTY2 + 2, HN = HX * HY;
  using T = typename std::conditional<STAGE, __half, float>::type;
  __shared__ T sp[9][HN];
  const int x0 = bx * kTX2, y0 = by * kTY2;
  for (int k = threadIdx.x; k < HN; k += kTX2 * kTY2) {
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
    float m[6];
#pragma unroll
    for (int q = 0; q < 6; q++) m[q] = Store<S>::template load<2>(in + q * plane + c, q);
    float f[9];
    recon2_all(m[0], m[1], m[2], m[3], m[5], m[4], f);
#pragma unroll
    for (int i = 0; i < 9; i++) {
      if constexpr (STAGE) sp[i][k] = to_half_rn(f[i]);
      else sp[i][k] = f[i];
    }
  }
  __syncthreads();
  const int lx = threadIdx.x % kTX2, ly = threadIdx.x / kTX2;
  const int x = x0 + lx, y = y0 + ly;
  if (x >= nx || y >= ny) return;
  const size_t cell = (size_t)y * nx + x;
  if (!is_updated(flags[cell])) return;
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
  collide2(p, omega, o);
#pragma unroll
  for (int q = 0; q < 6; q++) Store<S>::template store<2>(out + q * plane + cell, q, o[q]);
}

template <typename S, bool STAGE, bool PERIODIC>
__global__ void __launch_bounds__(kTX2* kTY2)
    k_bulk2d(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
             size_t plane, float omega) {
  bulk2d_tile<S, STAGE, PERIODIC>(in, out, flags, nx, ny, plane, omega, blockIdx.x, blockIdx.y);
}

// Link detection and stream-collide tiles in one launch: the first `nbuild`
// blocks detect the step's moving-wall links (dispatched first: they carry the
// longest dependent chains), the rest are tiles.  The two parts are
// independent (boundary cells are recomputed afterwards).
template <typename S, bool STAGE, bool PERIODIC>
__global__ void __launch_bounds__(kTX2* kTY2)
    k_bulk2d_links(const S* __restrict__ in, S* __restrict__ out, int tiles_x, int nbuild, LinkArgs la) {
  if ((int)blockIdx.x < nbuild) {
    build_links_block<2>(la, (long long)blockIdx.x, la.st->gen);
  } else {
    const int t = (int)blockIdx.x - nbuild;
    bulk2d_tile<S, STAGE, PERIODIC>(in, out, la.flags, la.nx, la.ny, la.plane, la.omega, t % tiles_x, t / tiles_x);
  }
}

template <typename S>
static int bulk2d_typed(Domain& d, const void* in, void* out, const LinkArgs* la) {
  const int tx = (d.nx + kTX2 - 1) / kTX2, ty = (d.ny + kTY2 - 1) / kTY2;
  const bool stage = d.options & FF_OPT_STAGE_FP16;
  const bool per = d.options & FF_OPT_PERIODIC;
  const S* i = (const S*)in;
  S* o = (S*)out;
  if (!la) {
    dim3 grid(tx, ty), block(kTX2 * kTY2);
#define FF_B(ST, PE) k_bulk2d<S, ST, PE><<<grid, block, 0, d.stream>>>(i, o, d.flags, d.nx, d.ny, d.plane, d.omega)
    if (stage && per) FF_B(true, true);
    else if (stage) FF_B(true, false);
    else if (per) FF_B(false, true);
    else FF_B(false, false);
#undef FF_B
  } else {
    const int nbulk = tx * ty;
    const long long pairs = (long long)d.ne * 8;
    const int nbuild = (int)((pairs + 15) / 16);
    const unsigned grid = (unsigned)(nbulk + nbuild);
#define FF_B(ST, PE) k_bulk2d_links<S, ST, PE><<<grid, kTX2 * kTY2, 0, d.stream>>>(i, o, tx, nbuild, *la)
    if (stage && per) FF_B(true, true);
    else if (stage) FF_B(true, false);
    else if (per) FF_B(false, true);
    else FF_B(false, false);
#undef FF_B
  }
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_bulk2d(Domain& d, const void* in, void* out) {
  switch (d.storage) {
    case FF_STORAGE_F32: return bulk2d_typed<float>(d, in, out, nullptr);
    case FF_STORAGE_F16: return bulk2d_typed<__half>(d, in, out, nullptr);
    default: return bulk2d_typed<__hip_bfloat16>(d, in, out, nullptr);
  }
}

int launch_bulk2d_links(Domain& d, const void* in, void* out, const LinkArgs& la) {
  switch (d.storage) {
    case FF_STORAGE_F32: return bulk2d_typed<float>(d, in, out, &la);
    case FF_STORAGE_F16: return bulk2d_typed<__half>(d, in, out, &la);
    default: return bulk2d_typed<__hip_bfloat16>(d, in, out, &la);
  }
}

// ---------------------------------------------------------------- outlet
// In place on the freshly written buffer (mrApplyOutlet2DKernel /
// mrApplyOutlet3DKernel, FsiCoupling2D.cu:257-316, FsiCoupling3D.cu:336-420):
// copy the outward normal velocity of the fluid neighbour when it flows out
// and always set the normal Pi to its square.  Only listed outlet cells run.
template <typename S, int DIM>
__global__ void k_outlet(S* buf, const OutletCell* __restrict__ oc, int n, size_t plane,
                         const uint8_t* __restrict__ flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const OutletCell c = oc[t];
  if (flags[c.nb] != kFluid) return;
  const float up = Store<S>::template load<DIM>(buf + c.comp * plane + c.nb, c.comp);
  if ((c.sign > 0 && up > 0.0f) || (c.sign < 0 && up < 0.0f))
    Store<S>::template store<DIM>(buf + c.comp * plane + c.cell, c.comp, up);
  Store<S>::template store<DIM>(buf + c.pcomp * plane + c.cell, c.pcomp, up * up);
}

int launch_outlet(Domain& d, void* buf) {
  if (d.n_outlets == 0) return FF_OK;
  const int b = 256, g = (d.n_outlets + b - 1) / b;
#define FF_OUT(S)                                                                                       \
  if (d.dim == 2) k_outlet<S, 2><<<g, b, 0, d.stream>>>((S*)buf, d.outlets, d.n_outlets, d.plane, d.flags); \
  else k_outlet<S, 3><<<g, b, 0, d.stream>>>((S*)buf, d.outlets, d.n_outlets, d.plane, d.flags);
  if (d.storage == FF_STORAGE_F32) { FF_OUT(float) }
  else if (d.storage == FF_STORAGE_F16) { FF_OUT(__half) }
  else { FF_OUT(__hip_bfloat16) }
#undef FF_OUT
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- conversion
template <typename S, int DIM>
__global__ void k_aos_to_soa(const float* __restrict__ aos, S* __restrict__ soa, long long n, size_t plane) {
  constexpr int NM = DIM == 2 ? 6 : 10;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int q = 0; q < NM; q++) Store<S>::template store<DIM>(soa + q * plane + c, q, aos[c * NM + q]);
}
template <typename S, int DIM>
__global__ void k_soa_to_aos(const S* __restrict__ soa, float* __restrict__ aos, long long n, size_t plane) {
  constexpr int NM = DIM == 2 ? 6 : 10;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int q = 0; q < NM; q++) aos[c * NM + q] = Store<S>::template load<DIM>(soa + q * plane + c, q);
}

int launch_convert_in(Domain& d, const float* aos, void* buf) {
  const int b = 256;
  const unsigned g = (unsigned)((d.n + b - 1) / b);
#define FF_CV(S)                                                                          \
  if (d.dim == 2) k_aos_to_soa<S, 2><<<g, b, 0, d.stream>>>(aos, (S*)buf, d.n, d.plane); \
  else k_aos_to_soa<S, 3><<<g, b, 0, d.stream>>>(aos, (S*)buf, d.n, d.plane);
  if (d.storage == FF_STORAGE_F32) { FF_CV(float) }
  else if (d.storage == FF_STORAGE_F16) { FF_CV(__half) }
  else { FF_CV(__hip_bfloat16) }
#undef FF_CV
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_convert_out(Domain& d, const void* buf, float* aos) {
  const int b = 256;
  const unsigned g = (unsigned)((d.n + b - 1) / b);
#define FF_CV(S)                                                                                \
  if (d.dim == 2) k_soa_to_aos<S, 2><<<g, b, 0, d.stream>>>((const S*)buf, aos, d.n, d.plane); \
  else k_soa_to_aos<S, 3><<<g, b, 0, d.stream>>>((const S*)buf, aos, d.n, d.plane);
  if (d.storage == FF_STORAGE_F32) { FF_CV(float) }
  else if (d.storage == FF_STORAGE_F16) { FF_CV(__half) }
  else { FF_CV(__hip_bfloat16) }
#undef FF_CV
  FF_HIP(hipGetLastError());
  return FF_OK;
}

// ---------------------------------------------------------------- mass
// Deterministic FP64 sum of rho over updated cells: fixed per-block tree,
// then one block folds the partials in index order.
constexpr int kRedBlocks = 1024;

template <typename S, int DIM>
__global__ void __launch_bounds__(256) k_mass_partial(const S* __restrict__ rho, const uint8_t* __restrict__ flags,
                                                      long long n, double* __restrict__ partial) {
  __shared__ double s[256];
  double acc = 0.0;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256)
    if (is_updated(flags[c])) acc += (double)Store<S>::template load<DIM>(rho + c, 0);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) k_fold(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

int launch_mass(Domain& d, double* out) {
  const void* buf = d.mom[d.cur];
#define FF_MS(S)                                                                                              \
  if (d.dim == 2) k_mass_partial<S, 2><<<kRedBlocks, 256, 0, d.stream>>>((const S*)buf, d.flags, d.n, d.red); \
  else k_mass_partial<S, 3><<<kRedBlocks, 256, 0, d.stream>>>((const S*)buf, d.flags, d.n, d.red);
  if (d.storage == FF_STORAGE_F32) { FF_MS(float) }
  else if (d.storage == FF_STORAGE_F16) { FF_MS(__half) }
  else { FF_MS(__hip_bfloat16) }
#undef FF_MS
  k_fold<<<1, 256, 0, d.stream>>>(d.red, kRedBlocks, out);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

}  // namespace ff
