// kernels3d.hip -- D3Q27 stream+collide over structure-of-arrays moments.
//
// Replaces mrDeformableSolver3DKernel_ShrdM (src/core/FsiCoupling3D.cu:83-334)
// for every updated cell without a moving-wall link (those are recomputed by
// the boundary-cell kernel in links.hip).  Unlike 2D, the reference keeps the
// flag bounce-back value in 3D (FsiCoupling3D.cu:209-222), so the upstream
// flags are staged next to the populations.  The body-force field of the
// reference is identically zero (LbmFlowField3D.cpp:35-36) and is not read.

#include "ff_internal.h"
#include "hermite.h"

namespace ff {

constexpr int kTX3 = 8, kTY3 = 8, kTZ3 = 4;

template <typename S, bool PERIODIC>
__global__ void __launch_bounds__(kTX3* kTY3* kTZ3)
    k_bulk3d(const S* __restrict__ in, S* __restrict__ out, const uint8_t* __restrict__ flags, int nx, int ny,
             int nz, size_t plane, float omega) {
  constexpr int HX = kTX3 + 2, HY = kTY3 + 2, HZ = kTZ3 + 2, HN = HX * HY * HZ;
  constexpr int NT = kTX3 * kTY3 * kTZ3;
  __shared__ float sp[27][HN];
  __shared__ uint8_t sf[HN];
  const int x0 = blockIdx.x * kTX3, y0 = blockIdx.y * kTY3, z0 = blockIdx.z * kTZ3;
  for (int k = threadIdx.x; k < HN; k += NT) {
    const int hx = k % HX, hy = (k / HX) % HY, hz = k / (HX * HY);
    const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
    const int wx = ((gx % nx) + nx) % nx, wy = ((gy % ny) + ny) % ny, wz = ((gz % nz) + nz) % nz;
    int sx, sy, sz;
    if (PERIODIC) {
      sx = wx; sy = wy; sz = wz;
    } else {
      sx = min(nx - 1, max(0, gx));
      sy = min(ny - 1, max(0, gy));
      sz = min(nz - 1, max(0, gz));
    }
    const size_t c = ((size_t)sz * ny + sy) * nx + sx;
    float m[10];
#pragma unroll
    for (int q = 0; q < 10; q++) m[q] = Store<S>::template load<3>(in + q * plane + c, q);
    float f[27];
    recon3_all(m, f);
#pragma unroll
    for (int i = 0; i < 27; i++) sp[i][k] = f[i];
    sf[k] = flags[((size_t)wz * ny + wy) * nx + wx];
  }
  __syncthreads();
  const int lx = threadIdx.x % kTX3, ly = (threadIdx.x / kTX3) % kTY3, lz = threadIdx.x / (kTX3 * kTY3);
  const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
  if (x >= nx || y >= ny || z >= nz) return;
  const int ctr = ((lz + 1) * HY + (ly + 1)) * HX + (lx + 1);
  if (!is_updated(sf[ctr])) return;
  float p[27];
#pragma unroll
  for (int i = 0; i < 27; i++) {
    const int up = ctr - (D3::cz(i) * HY + D3::cy(i)) * HX - D3::cx(i);
    if (is_wall3(sf[up])) p[i] = fmaxf(0.0f, sp[D3::inv(i)][ctr]);
    else p[i] = fmaxf(0.0f, sp[i][up]);
  }
  float o[10];
  collide3(p, omega, o);
  const size_t cell = ((size_t)z * ny + y) * nx + x;
#pragma unroll
  for (int q = 0; q < 10; q++) Store<S>::template store<3>(out + q * plane + cell, q, o[q]);
}

template <typename S>
static int bulk3d_typed(Domain& d, const void* in, void* out) {
  dim3 grid((d.nx + kTX3 - 1) / kTX3, (d.ny + kTY3 - 1) / kTY3, (d.nz + kTZ3 - 1) / kTZ3);
  dim3 block(kTX3 * kTY3 * kTZ3);
  if (d.options & FF_OPT_PERIODIC)
    k_bulk3d<S, true><<<grid, block, 0, d.stream>>>((const S*)in, (S*)out, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega);
  else
    k_bulk3d<S, false><<<grid, block, 0, d.stream>>>((const S*)in, (S*)out, d.flags, d.nx, d.ny, d.nz, d.plane, d.omega);
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
