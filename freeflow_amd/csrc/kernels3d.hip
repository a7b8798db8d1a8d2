// kernels3d.hip -- D3Q27 stream+collide over structure-of-arrays moments.
//
// Replaces mrDeformableSolver3DKernel_ShrdM (src/core/FsiCoupling3D.cu:83-334)
// for every updated cell without a moving-wall link (those are recomputed by
// the boundary-cell kernel in links.hip).  Unlike 2D, the reference keeps the
// flag bounce-back value in 3D (FsiCoupling3D.cu:209-222), so the upstream
// flags are staged next to the populations.  The body-force field of the
// reference is identically zero (LbmFlowField3D.cpp:35-36) and is not read.

#include <algorithm>

#include "ff_internal.h"
#include "links_body.h"
#include "stencil.h"

namespace ff {

// One tile per block (XCD-aware order).  KIND 0: boxes of TX x TY x ZC cells;
// 1: TX x TY columns marching ZC planes (bulk3d_ring); 2: TX x TY columns
// marching ZC planes with one staged cell per thread (bulk3d_col: FAST picks
// its arithmetic, PF and DMA how the next planes' moments are fetched).
template <typename S, bool PERIODIC, bool FAST, int KIND, int TX, int TY, int ZC, int PF, bool DMA>
__device__ __forceinline__ void tile3d(const S* in, S* out, const uint8_t* flags, int nx, int ny, int nz, size_t plane,
                                       float omega, int t, int tiles_x, int tiles_y) {
  const int bx = t % tiles_x, by = (t / tiles_x) % tiles_y, bz = t / (tiles_x * tiles_y);
  if constexpr (KIND == 2)
    bulk3d_col<S, PERIODIC, FAST, TX, TY, PF, DMA>(in, out, flags, nx, ny, nz, plane, omega, bx, by, bz * ZC,
                                                   min(bz * ZC + ZC, nz));
  else if constexpr (KIND == 1)
    bulk3d_ring<S, PERIODIC, TX, TY, ZC, FAST>(in, out, flags, nx, ny, nz, plane, omega, bx, by, bz);
  else
    bulk3d_tile<S, PERIODIC, TX, TY, ZC, FAST>(in, out, flags, nx, ny, nz, plane, omega, bx, by, bz);
}

// Staged columns are built for four waves per SIMD (at most 128 VGPRs); with
// an LDS ring of moments the LDS allows two or three.
template <typename S, bool PERIODIC, bool FAST, int KIND, int TX, int TY, int ZC, int NT, int PF, bool DMA>
__global__ void __launch_bounds__(NT, (KIND == 2 && !DMA ? 4 : 2))
    k_bulk3d(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
             int nz, size_t plane, float omega, int tiles_x, int tiles_y, int ntiles) {
  static_assert(KIND != 2 || NT == ColTile<TX, TY>::NT, "block size of a staged column");
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t < 0) return;
  tile3d<S, PERIODIC, FAST, KIND, TX, TY, ZC, PF, DMA>(in, out, flags, nx, ny, nz, plane, omega, t, tiles_x, tiles_y);
}

template <typename S, bool PERIODIC, bool FAST, int KIND, int TX, int TY, int ZC, int NT, int PF, bool DMA>
static void bulk3d_shape(Domain& d, const S* i, S* o) {
  const int tx = (d.nx + TX - 1) / TX, ty = (d.ny + TY - 1) / TY, tz = (d.nz + ZC - 1) / ZC;
  const int nt = tx * ty * tz;
  launch_timed(k_bulk3d<S, PERIODIC, FAST, KIND, TX, TY, ZC, NT, PF, DMA>, dim3(xcd_grid(nt)), dim3(NT), d.stream,
               d.tev0, d.tev1, i, o, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega, tx, ty, nt);
}

template <typename S, bool PERIODIC, bool FAST>
static void bulk3d_variant(Domain& d, const S* i, S* o) {
#define FF_B3(FA, K, X, Y, Z, N, P, DM) bulk3d_shape<S, PERIODIC, FA, K, X, Y, Z, N, P, DM>(d, i, o)
  FF_TILES3(FF_B3, FAST, d.tile3)
#undef FF_B3
}

// Stream launch of a 3D step with a boundary (see k_step2d): boundary-cell
// blocks, wave blocks, then the tiles of the chosen shape; NT = block size.
template <typename S, bool PERIODIC, bool FAST, int KIND, int TX, int TY, int ZC, int NT, int PF, bool DMA>
__global__ void __launch_bounds__(NT)
    k_step3d(const S* __restrict__ in, S* __restrict__ out, int tiles_x, int tiles_y, int ntiles, int nbc, int nwave,
             LinkArgs la, FinalArgs fa) {
  static_assert(KIND != 2 || NT == ColTile<TX, TY>::NT, "block size of a staged column");
  const int b = blockIdx.x;
  if (b < nbc) {
    bcells_body<3, S, false, PERIODIC, NT, FAST>(in, la, b, nbc);
  } else if (b < nbc + nwave) {
    stream_extra_body<3, NT, FAST>(la, fa, b - nbc);
  } else {
    const int t = xcd_tile(b - nbc - nwave, ntiles);
    if (t < 0) return;
    tile3d<S, PERIODIC, FAST, KIND, TX, TY, ZC, PF, DMA>(in, out, la.flags, la.nx, la.ny, la.nz, la.plane, la.omega,
                                                         t, tiles_x, tiles_y);
  }
}

template <typename S, bool PERIODIC, bool FAST, int KIND, int TX, int TY, int ZC, int NT, int PF, bool DMA>
static void step3d_shape(Domain& d, const S* i, S* o, const LinkArgs& la, FinalArgs fa, bool wave) {
  const int tx = (d.nx + TX - 1) / TX, ty = (d.ny + TY - 1) / TY, tz = (d.nz + ZC - 1) / ZC;
  const int nt = tx * ty * tz;
  fa.kin_blocks = wave ? (d.nv + NT - 1) / NT : 0;
  fa.fin_blocks = d.finish_pending ? (d.nv + NT - 1) / NT + 1 : 0;
  const int nw = fa.kin_blocks + fa.fin_blocks;
  const int nbc = xcd_grid(boundary_cell_blocks(d, la.bc_span) + nw) - nw;
  const unsigned grid = (unsigned)(nbc + nw + xcd_grid(nt));
  launch_timed(k_step3d<S, PERIODIC, FAST, KIND, TX, TY, ZC, NT, PF, DMA>, dim3(grid), dim3(NT), d.stream, d.tev0,
               d.tev1, i, o, tx, ty, nt, nbc, nw, la, fa);
}

template <typename S, bool PERIODIC, bool FAST>
static void step3d_variant(Domain& d, const S* i, S* o, const LinkArgs& la, const FinalArgs& fa, bool wave) {
#define FF_S3(FA, K, X, Y, Z, N, P, DM) step3d_shape<S, PERIODIC, FA, K, X, Y, Z, N, P, DM>(d, i, o, la, fa, wave)
  FF_TILES3(FF_S3, FAST, d.tile3)
#undef FF_S3
}

template <typename S>
static int step3d_typed(Domain& d, const void* in, void* out, const LinkArgs& la, const FinalArgs& fa, bool wave) {
  const bool per = d.options & FF_OPT_PERIODIC, fast = d.options & FF_OPT_FAST_MATH;
  if (per && fast) step3d_variant<S, true, true>(d, (const S*)in, (S*)out, la, fa, wave);
  else if (per) step3d_variant<S, true, false>(d, (const S*)in, (S*)out, la, fa, wave);
  else if (fast) step3d_variant<S, false, true>(d, (const S*)in, (S*)out, la, fa, wave);
  else step3d_variant<S, false, false>(d, (const S*)in, (S*)out, la, fa, wave);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_step_stream3d(Domain& d, const void* in, void* out, const LinkArgs& la, FinalArgs fa, bool wave) {
  switch (d.storage) {
    case FF_STORAGE_F32: return step3d_typed<float>(d, in, out, la, fa, wave);
    case FF_STORAGE_F16: return step3d_typed<__half>(d, in, out, la, fa, wave);
    case FF_STORAGE_F16U: return step3d_typed<h16u>(d, in, out, la, fa, wave);
    default: return step3d_typed<__hip_bfloat16>(d, in, out, la, fa, wave);
  }
}

template <typename S>
static int bulk3d_typed(Domain& d, const void* in, void* out) {
  const bool per = d.options & FF_OPT_PERIODIC, fast = d.options & FF_OPT_FAST_MATH;
  if (per && fast) bulk3d_variant<S, true, true>(d, (const S*)in, (S*)out);
  else if (per) bulk3d_variant<S, true, false>(d, (const S*)in, (S*)out);
  else if (fast) bulk3d_variant<S, false, true>(d, (const S*)in, (S*)out);
  else bulk3d_variant<S, false, false>(d, (const S*)in, (S*)out);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_bulk3d(Domain& d, const void* in, void* out) {
  switch (d.storage) {
    case FF_STORAGE_F32: return bulk3d_typed<float>(d, in, out);
    case FF_STORAGE_F16: return bulk3d_typed<__half>(d, in, out);
    case FF_STORAGE_F16U: return bulk3d_typed<h16u>(d, in, out);
    default: return bulk3d_typed<__hip_bfloat16>(d, in, out);
  }
}

}  // namespace ff
