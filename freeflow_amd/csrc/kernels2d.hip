// kernels2d.hip -- D2Q9 stream+collide over structure-of-arrays moments, the
// outlet update, storage conversion and the FP64 mass reduction.
//
// Replaces mrDeformableSolver2DKernel_Shrd and mrApplyOutlet2DKernel
// (src/core/FsiCoupling2D.cu:82-316).  The bulk kernel treats every updated
// cell as if no moving-wall link existed; the few cells that own a
// moving-wall link are recomputed afterwards by the boundary-cell kernel
// (links.hip), so the bulk pass stays a pure stencil with one byte of flag
// traffic per cell.  With a boundary present the same launch also runs the
// boundary cells and the next step's wave kinematics in extra workgroups.
//
// In 2D a wall-upstream population is overwritten by the ordinary pull in the
// reference (FsiCoupling2D.cu:168-223: `is_blocked` stays false on the wall
// branch), so the bulk kernel never needs neighbour flags.

#include "ff_internal.h"
#include "geometry.h"
#include "links_body.h"
#include "stencil.h"

namespace ff {

template <typename S, bool STAGE, bool PERIODIC, int TX, int TY, int CPT, bool FAST>
__global__ void __launch_bounds__(256)
    k_bulk2d(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
             size_t plane, float omega, int tiles_x, int ntiles) {
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t < 0) return;
  bulk2d_tile<S, STAGE, PERIODIC, TX, TY, CPT, FAST>(in, out, flags, nx, ny, plane, omega, t % tiles_x, t / tiles_x);
}

// Stream launch of a step with a boundary: the first `nbc` blocks update the
// cells owning a moving-wall link (into the side buffer), the next `nwave`
// blocks evaluate the next step's wave kinematics, the rest are the
// stream-collide tiles (which also write the boundary cells; the copy-back of
// the sums launch overwrites them).  nbc + nwave is a multiple of 8, so the
// tiles keep the XCD-aware order.
template <typename S, bool STAGE, bool PERIODIC, int TX, int TY, int CPT, bool FAST>
__global__ void __launch_bounds__(256)
    k_step2d(const S* __restrict__ in, S* __restrict__ out, int tiles_x, int ntiles, int nbc, int nwave, LinkArgs la,
             FinalArgs fa) {
  const int b = blockIdx.x;
  if (b < nbc) {
    bcells_body<2, S, STAGE, PERIODIC, 256, FAST>(in, la, b, nbc);
  } else if (b < nbc + nwave) {
    stream_extra_body<2, 256, FAST>(la, fa, b - nbc);
  } else {
    const int t = xcd_tile(b - nbc - nwave, ntiles);
    if (t < 0) return;
    bulk2d_tile<S, STAGE, PERIODIC, TX, TY, CPT, FAST>(in, out, la.flags, la.nx, la.ny, la.plane, la.omega,
                                                       t % tiles_x, t / tiles_x);
  }
}

// Step launch parameters beyond the tiles (null: plain stream-collide).
struct StreamExtra {
  const LinkArgs* la;
  const FinalArgs* fa;
  int nbc, nwave;
};

template <typename S, bool STAGE, bool PERIODIC, int TX, int TY, int CPT, bool FAST>
static void bulk2d_shape(Domain& d, const S* i, S* o, const StreamExtra* x) {
  const int tx = (d.nx + TX - 1) / TX, ty = (d.ny + TY * CPT - 1) / (TY * CPT);
  const int nt = tx * ty;
  if (!x) {
    launch_timed(k_bulk2d<S, STAGE, PERIODIC, TX, TY, CPT, FAST>, dim3(xcd_grid(nt)), dim3(256), d.stream, d.tev0,
                 d.tev1, i, o, d.flags, d.nx, d.ny, d.plane, d.omega, tx, nt);
  } else {
    const int nbc = xcd_grid(x->nbc + x->nwave) - x->nwave;  // extra boundary-cell blocks share the grid stride
    const unsigned grid = (unsigned)(nbc + x->nwave + xcd_grid(nt));
    launch_timed(k_step2d<S, STAGE, PERIODIC, TX, TY, CPT, FAST>, dim3(grid), dim3(256), d.stream, d.tev0, d.tev1, i,
                 o, tx, nt, nbc, x->nwave, *x->la, *x->fa);
  }
}

template <typename S, bool STAGE, bool PERIODIC, bool FAST>
static void bulk2d_variant(Domain& d, const S* i, S* o, const StreamExtra* x) {
  switch (d.tile2) {
    case 1: bulk2d_shape<S, STAGE, PERIODIC, 32, 8, 2, FAST>(d, i, o, x); break;
    case 2: bulk2d_shape<S, STAGE, PERIODIC, 64, 4, 2, FAST>(d, i, o, x); break;
    case 3: bulk2d_shape<S, STAGE, PERIODIC, 32, 8, 4, FAST>(d, i, o, x); break;
    case 4: bulk2d_shape<S, STAGE, PERIODIC, 64, 4, 4, FAST>(d, i, o, x); break;
    default: bulk2d_shape<S, STAGE, PERIODIC, 32, 8, 1, FAST>(d, i, o, x); break;
  }
}

// FF_OPT_FAST_MATH applies to unstaged tiles (the staged tile reproduces the
// reference's half-precision rounding of the reference's arithmetic).
template <typename S>
static int bulk2d_typed(Domain& d, const void* in, void* out, const StreamExtra* x) {
  const bool stage = d.options & FF_OPT_STAGE_FP16;
  const bool per = d.options & FF_OPT_PERIODIC;
  const bool fast = d.options & FF_OPT_FAST_MATH;
  const S* i = (const S*)in;
  S* o = (S*)out;
  if (stage && per) bulk2d_variant<S, true, true, false>(d, i, o, x);
  else if (stage) bulk2d_variant<S, true, false, false>(d, i, o, x);
  else if (per && fast) bulk2d_variant<S, false, true, true>(d, i, o, x);
  else if (per) bulk2d_variant<S, false, true, false>(d, i, o, x);
  else if (fast) bulk2d_variant<S, false, false, true>(d, i, o, x);
  else bulk2d_variant<S, false, false, false>(d, i, o, x);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

static int bulk2d_dispatch(Domain& d, const void* in, void* out, const StreamExtra* x) {
  switch (d.storage) {
    case FF_STORAGE_F32: return bulk2d_typed<float>(d, in, out, x);
    case FF_STORAGE_F16: return bulk2d_typed<__half>(d, in, out, x);
    case FF_STORAGE_F16U: return bulk2d_typed<h16u>(d, in, out, x);
    default: return bulk2d_typed<__hip_bfloat16>(d, in, out, x);
  }
}

int launch_bulk2d(Domain& d, const void* in, void* out) { return bulk2d_dispatch(d, in, out, nullptr); }

int launch_step_stream3d(Domain& d, const void* in, void* out, const LinkArgs& la, FinalArgs fa, bool wave);

int launch_step_stream(Domain& d, const void* in, void* out, const LinkArgs& la, bool wave) {
  FinalArgs fa = make_final_args(d, out, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], nullptr);
  fa.fin_gen = d.finish_gen;
  if (d.dim == 3) return launch_step_stream3d(d, in, out, la, fa, wave);
  fa.kin_blocks = wave ? (d.nv + 255) / 256 : 0;
  fa.fin_blocks = d.finish_pending ? (d.nv + 255) / 256 + 1 : 0;
  const StreamExtra x{&la, &fa, boundary_cell_blocks(d, la.bc_span), fa.kin_blocks + fa.fin_blocks};
  return bulk2d_dispatch(d, in, out, &x);
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
  const float up = Store<S>::template load<DIM>(buf + c.comp * plane + c.nb, c.comp, plane);
  if ((c.sign > 0 && up > 0.0f) || (c.sign < 0 && up < 0.0f))
    Store<S>::template store<DIM>(buf + c.comp * plane + c.cell, c.comp, up, plane);
  Store<S>::template store<DIM>(buf + c.pcomp * plane + c.cell, c.pcomp, up * up, plane);
}

int launch_outlet(Domain& d, void* buf) {
  if (d.n_outlets == 0) return FF_OK;
  const int b = 256, g = (d.n_outlets + b - 1) / b;
#define FF_OUT(S)                                                                                       \
  if (d.dim == 2) k_outlet<S, 2><<<g, b, 0, d.stream>>>((S*)buf, d.outlets, d.n_outlets, d.plane, d.flags); \
  else k_outlet<S, 3><<<g, b, 0, d.stream>>>((S*)buf, d.outlets, d.n_outlets, d.plane, d.flags);
  if (d.storage == FF_STORAGE_F32) { FF_OUT(float) }
  else if (d.storage == FF_STORAGE_F16) { FF_OUT(__half) }
  else if (d.storage == FF_STORAGE_F16U) { FF_OUT(h16u) }
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
  for (int q = 0; q < NM; q++) Store<S>::template store<DIM>(soa + q * plane + c, q, aos[c * NM + q], plane);
}
template <typename S, int DIM>
__global__ void k_soa_to_aos(const S* __restrict__ soa, float* __restrict__ aos, long long n, size_t plane) {
  constexpr int NM = DIM == 2 ? 6 : 10;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int q = 0; q < NM; q++) aos[c * NM + q] = Store<S>::template load<DIM>(soa + q * plane + c, q, plane);
}

// ---------------------------------------------------------------- initial state
// initialize_state_kernel_2d/3d (LbmInitializer2D.cu:32-95,
// LbmInitializer3D.cu:33-113): rho = 1, the inlet velocity of Inlet* cells,
// Pi from the equilibrium populations; same arithmetic, on the device.
struct InitBC {
  float v[6];
};

template <int DIM>
__device__ __forceinline__ void init_moments(uint8_t fl, const InitBC& bc, float* m) {
  const float cs2 = 1.0f / 3.0f;
  const float rho = 1.0f;
  float ux = 0.0f, uy = 0.0f, uz = 0.0f;
  if constexpr (DIM == 2) {
    constexpr float ex2[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    constexpr float ey2[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    constexpr float w2[9] = {4.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f,
                             1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};
    if (fl == kInletLeft) ux = bc.v[0];
    else if (fl == kInletDown) uy = bc.v[1];
    else if (fl == kInletRight) ux = -bc.v[2];
    else if (fl == kInletUp) uy = -bc.v[3];
    const float u_sq = ux * ux + uy * uy;
    float feq[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const float cu = ex2[k] * ux + ey2[k] * uy;
      feq[k] = w2[k] * rho * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * u_sq);
    }
    const float inv = 1 / rho;
    const float pixx = feq[1] + feq[3] + feq[5] + feq[6] + feq[7] + feq[8];
    const float piyy = feq[2] + feq[4] + feq[5] + feq[6] + feq[7] + feq[8];
    const float pixy = feq[5] - feq[6] + feq[7] - feq[8];
    m[0] = rho;
    m[1] = ux;
    m[2] = uy;
    m[3] = (float)((double)(pixx * inv) - 1.0 * (double)cs2);
    m[4] = (float)((double)(piyy * inv) - 1.0 * (double)cs2);
    m[5] = pixy * inv;
  } else {
    if (fl == kInletLeft) ux = bc.v[0];
    else if (fl == kInletDown) uz = bc.v[1];
    else if (fl == kInletRight) ux = -bc.v[2];
    else if (fl == kInletUp) uz = -bc.v[3];
    else if (fl == kInletFront) uy = bc.v[4];
    else if (fl == kInletBack) uy = -bc.v[5];
    const float u_sq = ux * ux + uy * uy + uz * uz;
    float f[27];
#pragma unroll
    for (int k = 0; k < 27; k++) {
      const float cu = (float)D3::cx(k) * ux + (float)D3::cy(k) * uy + (float)D3::cz(k) * uz;
      const float w = k == 0 ? 8.0f / 27.0f : (k < 7 ? 2.0f / 27.0f : (k < 19 ? 1.0f / 54.0f : 1.0f / 216.0f));
      f[k] = w * rho * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * u_sq);
    }
    const float pixx = f[1] + f[2] + f[7] + f[8] + f[9] + f[10] + f[13] + f[14] + f[15] + f[16] + f[19] + f[20] +
                       f[21] + f[22] + f[23] + f[24] + f[25] + f[26];
    const float piyy = f[3] + f[4] + f[7] + f[8] + f[11] + f[12] + f[13] + f[14] + f[17] + f[18] + f[19] + f[20] +
                       f[21] + f[22] + f[23] + f[24] + f[25] + f[26];
    const float pizz = f[5] + f[6] + f[9] + f[10] + f[11] + f[12] + f[15] + f[16] + f[17] + f[18] + f[19] + f[20] +
                       f[21] + f[22] + f[23] + f[24] + f[25] + f[26];
    const float pixy = (f[7] + f[8] + f[19] + f[20] + f[21] + f[22]) - (f[13] + f[14] + f[23] + f[24] + f[25] + f[26]);
    const float piyz = (f[11] + f[12] + f[19] + f[20] + f[25] + f[26]) - (f[17] + f[18] + f[21] + f[22] + f[23] + f[24]);
    const float pixz = (f[9] + f[10] + f[19] + f[20] + f[23] + f[24]) - (f[15] + f[16] + f[21] + f[22] + f[25] + f[26]);
    const float inv = 1 / rho;
    m[0] = rho;
    m[1] = ux;
    m[2] = uy;
    m[3] = uz;
    m[4] = (float)((double)(pixx * inv) - 1.0 * (double)cs2);
    m[5] = (float)((double)(piyy * inv) - 1.0 * (double)cs2);
    m[6] = (float)((double)(pizz * inv) - 1.0 * (double)cs2);
    m[7] = pixy * inv;
    m[8] = piyz * inv;
    m[9] = pixz * inv;
  }
}

template <typename S, int DIM>
__global__ void k_init_state(const uint8_t* __restrict__ flags, InitBC bc, S* __restrict__ a, S* __restrict__ b,
                             long long n, size_t plane) {
  constexpr int NM = DIM == 2 ? 6 : 10;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float m[NM];
  init_moments<DIM>(flags[c], bc, m);
#pragma unroll
  for (int q = 0; q < NM; q++) {
    Store<S>::template store<DIM>(a + q * plane + c, q, m[q], plane);
    Store<S>::template store<DIM>(b + q * plane + c, q, m[q], plane);
  }
}

// Both moment buffers of d from its flags and the inlet velocities bc[6].
int launch_init_state(Domain& d, const float* bc6) {
  InitBC bc;
  for (int k = 0; k < 6; k++) bc.v[k] = bc6[k];
  const int b = 256;
  const unsigned g = (unsigned)((d.n + b - 1) / b);
#define FF_INIT(S)                                                                                                \
  if (d.dim == 2) k_init_state<S, 2><<<g, b, 0, d.stream>>>(d.flags, bc, (S*)d.mom[0], (S*)d.mom[1], d.n, d.plane); \
  else k_init_state<S, 3><<<g, b, 0, d.stream>>>(d.flags, bc, (S*)d.mom[0], (S*)d.mom[1], d.n, d.plane);
  if (d.storage == FF_STORAGE_F32) { FF_INIT(float) }
  else if (d.storage == FF_STORAGE_F16) { FF_INIT(__half) }
  else if (d.storage == FF_STORAGE_F16U) { FF_INIT(h16u) }
  else { FF_INIT(__hip_bfloat16) }
#undef FF_INIT
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_convert_in(Domain& d, const float* aos, void* buf) {
  const int b = 256;
  const unsigned g = (unsigned)((d.n + b - 1) / b);
#define FF_CV(S)                                                                          \
  if (d.dim == 2) k_aos_to_soa<S, 2><<<g, b, 0, d.stream>>>(aos, (S*)buf, d.n, d.plane); \
  else k_aos_to_soa<S, 3><<<g, b, 0, d.stream>>>(aos, (S*)buf, d.n, d.plane);
  if (d.storage == FF_STORAGE_F32) { FF_CV(float) }
  else if (d.storage == FF_STORAGE_F16) { FF_CV(__half) }
  else if (d.storage == FF_STORAGE_F16U) { FF_CV(h16u) }
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
  else if (d.storage == FF_STORAGE_F16U) { FF_CV(h16u) }
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
    if (is_updated(flags[c])) acc += (double)Store<S>::template load<DIM>(rho + c, 0, 0);
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
  else if (d.storage == FF_STORAGE_F16U) { FF_MS(h16u) }
  else { FF_MS(__hip_bfloat16) }
#undef FF_MS
  k_fold<<<1, 256, 0, d.stream>>>(d.red, kRedBlocks, out);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

}  // namespace ff
