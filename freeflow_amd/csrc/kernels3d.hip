// kernels3d.hip -- D3Q27 stream+collide over structure-of-arrays moments.
//
// Replaces mrDeformableSolver3DKernel_ShrdM (src/core/FsiCoupling3D.cu:83-334)
// for every updated cell without a moving-wall link (those are recomputed by
// the boundary-cell kernel in links.hip).  Unlike 2D, the reference keeps the
// flag bounce-back value in 3D (FsiCoupling3D.cu:209-222), so the upstream
// flags are staged next to the populations.  The body-force field of the
// reference is identically zero (LbmFlowField3D.cpp:35-36) and is not read.

#include "ff_internal.h"
#include "stencil.h"

namespace ff {

template <typename S, bool PERIODIC>
__global__ void __launch_bounds__(kTX3* kTY3* kTZ3)
    k_bulk3d(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
             int nz, size_t plane, float omega) {
  bulk3d_tile<S, PERIODIC>(in, out, flags, nx, ny, nz, plane, omega, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <typename S, bool PERIODIC, int TX, int TY, int ZC>
__global__ void __launch_bounds__(TX* TY)
    k_bulk3d_ring(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
                  int nz, size_t plane, float omega) {
  bulk3d_ring<S, PERIODIC, TX, TY, ZC>(in, out, flags, nx, ny, nz, plane, omega, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <typename S, bool PERIODIC>
static void bulk3d_variant(Domain& d, const S* i, S* o) {
  const Tile3 t = kTiles3[d.tile3];
  dim3 grid((d.nx + t.tx - 1) / t.tx, (d.ny + t.ty - 1) / t.ty, (d.nz + t.tz - 1) / t.tz);
  switch (d.tile3) {
    case 1: k_bulk3d_ring<S, PERIODIC, 16, 8, 32><<<grid, 128, 0, d.stream>>>(i, o, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega); break;
    case 2: k_bulk3d_ring<S, PERIODIC, 8, 8, 32><<<grid, 64, 0, d.stream>>>(i, o, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega); break;
    case 3: k_bulk3d_ring<S, PERIODIC, 16, 16, 32><<<grid, 256, 0, d.stream>>>(i, o, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega); break;
    default: k_bulk3d<S, PERIODIC><<<grid, kTX3 * kTY3 * kTZ3, 0, d.stream>>>(i, o, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega); break;
  }
}

template <typename S>
static int bulk3d_typed(Domain& d, const void* in, void* out) {
  if (d.options & FF_OPT_PERIODIC) bulk3d_variant<S, true>(d, (const S*)in, (S*)out);
  else bulk3d_variant<S, false>(d, (const S*)in, (S*)out);
  FF_HIP(hipGetLastError());
  return FF_OK;
}

int launch_bulk3d(Domain& d, const void* in, void* out) {
  switch (d.storage) {
    case FF_STORAGE_F32: return bulk3d_typed<float>(d, in, out);
    case FF_STORAGE_F16: return bulk3d_typed<__half>(d, in, out);
    default: return bulk3d_typed<__hip_bfloat16>(d, in, out);
  }
}

}  // namespace ff
