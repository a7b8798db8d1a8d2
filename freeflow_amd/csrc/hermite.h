// hermite.h -- device-side HOME-LBM algebra (D2Q9 / D3Q27).
//
// The moment-encoded scheme of the reference keeps only rho, u and the
// normalised second moment Pi per cell; populations are rebuilt on the fly
// with a third-order Hermite expansion and relaxed in moment space.  The
// arithmetic below is written in the same operation order as the reference
// (src/lbm/LbmUtilsFunc.cu) and every translation unit that includes it is
// compiled with -ffp-contract=off, so that the FP32 results match the
// reference rounding, including its FP16 tile staging and the double
// promotions of CalculateDistributionD2Q9ALL.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace ff {

// Lattice direction ordering is fixed by the reference's tables
// (LbmConstants.hpp:24-27, 48-50): it indexes the link keys, so it is kept.
struct D2 {
  static constexpr int Q = 9;
  __host__ __device__ static constexpr int cx(int i) {
    return i == 1 || i == 5 || i == 8 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0);
  }
  __host__ __device__ static constexpr int cy(int i) {
    return i == 2 || i == 5 || i == 6 ? 1 : (i == 4 || i == 7 || i == 8 ? -1 : 0);
  }
  __host__ __device__ static constexpr int inv(int i) {
    return i == 0 ? 0 : (i <= 4 ? ((i + 1) % 4) + 1 : ((i - 5 + 2) % 4) + 5);
  }
};

struct D3 {
  static constexpr int Q = 27;
  // direction i = 0..26 in the reference order
  __host__ __device__ static constexpr int cx(int i) {
    constexpr int t[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    return t[i];
  }
  __host__ __device__ static constexpr int cy(int i) {
    constexpr int t[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    return t[i];
  }
  __host__ __device__ static constexpr int cz(int i) {
    constexpr int t[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
    return t[i];
  }
  // opposite directions come in adjacent pairs (1,2), (3,4), ...
  __host__ __device__ static constexpr int inv(int i) { return i == 0 ? 0 : (i & 1 ? i + 1 : i - 1); }
};

// Flag sets as bit masks over the flag value (one shift and test, no branch chain).
// FsiCoupling2D.cu:171-176: Solid 3, Wall 5, WallLeft 6, WallRight 7, WallDown 10, WallUp 11
__device__ __forceinline__ bool is_wall2(uint8_t f) { return f < 12 && ((0xCE8u >> f) & 1u); }
// FsiCoupling3D.cu:209-216: Solid 3, Wall 5, WallLeft..WallUp 6-11
__device__ __forceinline__ bool is_wall3(uint8_t f) { return f < 12 && ((0xFE8u >> f) & 1u); }
// Fluid 1, SolidDynamic 4
__device__ __forceinline__ bool is_updated(uint8_t f) { return f < 5 && ((0x12u >> f) & 1u); }

// Materialise an FP32 value in a register.  Without it the backend may fuse the
// preceding multiply into a float->half conversion (v_fma_mix*), rounding the
// exact product once to FP16 where the reference rounds to FP32 first.
__device__ __forceinline__ float opaque(float x) {
  __asm__ volatile("" : "+v"(x));
  return x;
}
__device__ __forceinline__ __half to_half_rn(float x) { return __float2half_rn(opaque(x)); }

// -------------------------------------------------------------------- 2D
#define FF_W0_2 (4.0f / 9.0f)
#define FF_W1_2 (1.0f / 9.0f)
#define FF_W5_2 (1.0f / 36.0f)

// All nine populations from (rho, u, Pi) -- CalculateDistributionD2Q9ALL.
// Argument order (rho, ux, uy, pixx, pixy, piyy) as in the reference.
__host__ __device__ __forceinline__ void recon2_all(float rho, float ux, float uy, float pixx, float pixy,
                                                    float piyy, float* f) {
  const float m1x = 3.0f * rho * ux;
  const float m1y = 3.0f * rho * uy;
  const float m2xx = 3.0f * rho * pixx;
  const float m2yy = 3.0f * rho * piyy;
  const float m2xy = 9.0f * rho * pixy;
  const float qxxy = 9.0f * (pixx * uy + 2.0f * pixy * ux - 2.0f * rho * ux * ux * uy);
  const float qxyy = 9.0f * (piyy * ux + 2.0f * pixy * uy - 2.0f * rho * uy * uy * ux);
  const float bulk = rho - 0.5f * (m2xx + m2yy);
  f[0] = FF_W0_2 * bulk;
  f[1] = (float)((double)FF_W1_2 * ((double)(bulk + m1x + 1.5f * m2xx) - 0.5 * (double)qxyy));
  f[2] = (float)((double)FF_W1_2 * ((double)(bulk + m1y + 1.5f * m2yy) - 0.5 * (double)qxxy));
  f[3] = (float)((double)FF_W1_2 * ((double)(bulk - m1x + 1.5f * m2xx) + 0.5 * (double)qxyy));
  f[4] = (float)((double)FF_W1_2 * ((double)(bulk - m1y + 1.5f * m2yy) + 0.5 * (double)qxxy));
  const float dbase = bulk + 1.5f * (m2xx + m2yy);
  f[5] = FF_W5_2 * (dbase + (m1x + m1y) + m2xy + (qxxy + qxyy));
  f[6] = FF_W5_2 * (dbase - m1x + m1y - m2xy + qxxy - qxyy);
  f[7] = FF_W5_2 * (dbase - m1x - m1y + m2xy - qxxy - qxyy);
  f[8] = FF_W5_2 * (dbase + m1x - m1y - m2xy - qxxy + qxyy);
}

// Population i alone, rounded exactly as recon2_all rounds f[i]: the same
// operations in the same order, with the direction's signs (+-1, so the
// products are exact) selected per lane instead of one expression per
// direction.  The boundary-cell update needs only the pulled population and
// the cell's own opposite one, not all nine.
__host__ __device__ __forceinline__ float recon2_pop(float rho, float ux, float uy, float pixx, float pixy, float piyy,
                                                     int i) {
  const float m1x = 3.0f * rho * ux;
  const float m1y = 3.0f * rho * uy;
  const float m2xx = 3.0f * rho * pixx;
  const float m2yy = 3.0f * rho * piyy;
  const float m2xy = 9.0f * rho * pixy;
  const float qxxy = 9.0f * (pixx * uy + 2.0f * pixy * ux - 2.0f * rho * ux * ux * uy);
  const float qxyy = 9.0f * (piyy * ux + 2.0f * pixy * uy - 2.0f * rho * uy * uy * ux);
  const float bulk = rho - 0.5f * (m2xx + m2yy);
  const float sx = (float)D2::cx(i), sy = (float)D2::cy(i);
  if (i == 0) return FF_W0_2 * bulk;
  if (i <= 4) {  // axis: f = w ((bulk + s m1 + 1.5 m2) - 0.5 s q), the bracket in FP64
    const bool ax = sx != 0.0f;
    const float sg = ax ? sx : sy;
    const float m1 = ax ? m1x : m1y, m2 = ax ? m2xx : m2yy, q = ax ? qxyy : qxxy;
    return (float)((double)FF_W1_2 * ((double)(bulk + sg * m1 + 1.5f * m2) - 0.5 * (double)(sg * q)));
  }
  const float dbase = bulk + 1.5f * (m2xx + m2yy);
  const float g = i == 5 ? dbase + (m1x + m1y) + m2xy + (qxxy + qxyy)
                         : dbase + sx * m1x + sy * m1y + (sx * sy) * m2xy + sy * qxxy + sx * qxyy;
  return FF_W5_2 * g;
}

// One population -- CalculateDistributionD2Q9AtIndex (all-float axis terms).
__device__ __forceinline__ float recon2_at(float rho, float ux, float uy, float pixx, float pixy, float piyy, int i) {
  const float m1x = 3.0f * rho * ux;
  const float m1y = 3.0f * rho * uy;
  const float m2xx = 3.0f * rho * pixx;
  const float m2yy = 3.0f * rho * piyy;
  const float m2xy = 9.0f * rho * pixy;
  const float qxxy = 9.0f * (pixx * uy + 2.0f * pixy * ux - 2.0f * rho * ux * ux * uy);
  const float qxyy = 9.0f * (piyy * ux + 2.0f * pixy * uy - 2.0f * rho * uy * uy * ux);
  const float bulk = rho - 0.5f * (m2xx + m2yy);
  const float dbase = bulk + 1.5f * (m2xx + m2yy);
  float s, w;
  switch (i) {
    case 0: s = bulk; w = FF_W0_2; break;
    case 1: s = bulk + m1x + 1.5f * m2xx - 0.5f * qxyy; w = FF_W1_2; break;
    case 2: s = bulk + m1y + 1.5f * m2yy - 0.5f * qxxy; w = FF_W1_2; break;
    case 3: s = bulk - m1x + 1.5f * m2xx + 0.5f * qxyy; w = FF_W1_2; break;
    case 4: s = bulk - m1y + 1.5f * m2yy + 0.5f * qxxy; w = FF_W1_2; break;
    case 5: s = dbase + (m1x + m1y) + m2xy + (qxxy + qxyy); w = FF_W5_2; break;
    case 6: s = dbase - m1x + m1y - m2xy + qxxy - qxyy; w = FF_W5_2; break;
    case 7: s = dbase - m1x - m1y + m2xy - qxxy - qxyy; w = FF_W5_2; break;
    default: s = dbase + m1x - m1y - m2xy - qxxy + qxyy; w = FF_W5_2; break;
  }
  return w * s;
}

// Moments of the pulled populations, relaxation and normalisation -- the tail
// of mrDeformableSolver2DKernel_Shrd (FsiCoupling2D.cu:226-252) with the
// body force identically zero (:227-228).  out = [rho, ux, uy, Pxx, Pyy, Pxy].
__device__ __forceinline__ void collide2(const float* p, float omega, float* out) {
  const float rho = p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8];
  const float inv = 1.0f / rho;
  const float ux = (p[1] - p[3] + p[5] - p[6] - p[7] + p[8]) * inv;
  const float uy = (p[2] - p[4] + p[5] + p[6] - p[7] - p[8]) * inv;
  const float pxx = p[1] + p[3] + p[5] + p[6] + p[7] + p[8];
  const float pyy = p[2] + p[4] + p[5] + p[6] + p[7] + p[8];
  const float pxy = p[5] - p[6] + p[7] - p[8];
  // LbmUtilsFuncGpu2D::Collision (LbmUtilsFunc.cu:512-543), F = 0
  const float omw = 1.0f - omega;
  const float pr = rho / 3.0f;
  const float ru2 = rho * ux * ux;
  const float rv2 = rho * uy * uy;
  const float dev = (pxx - pyy) * 0.5f;
  const float be = pr + 0.5f * (ru2 + rv2);
  const float an = 0.5f * omega * (ru2 - rv2);
  const float nxx = be + dev * omw + an;
  const float nyy = be - dev * omw - an;
  const float nxy = pxy * omw + (omega * rho * ux * uy);
  out[0] = rho;
  out[1] = ux;
  out[2] = uy;
  out[3] = nxx * inv - (1.0f / 3.0f);
  out[4] = nyy * inv - (1.0f / 3.0f);
  out[5] = nxy * inv;
}

// ------------------------------------------------------------- 2D, fast form
// recon2_all / collide2 for instruction count rather than the reference's
// rounding (no double promotions): opposite directions share an even part
// E and differ in the sign of the odd part O, f_i = w (E + O),
// f_inv(i) = w (E - O); moments from the pair sums and differences; FMA.
// The 2D quirk of the reference is kept: the Q terms use Pi without rho.
__device__ __forceinline__ void recon2_fast(float rho, float ux, float uy, float pixx, float pixy, float piyy,
                                            float* f) {
  const float r3 = 3.0f * rho;
  const float m1x = r3 * ux, m1y = r3 * uy, m2xx = r3 * pixx, m2yy = r3 * piyy, m2xy = (3.0f * r3) * pixy;
  // qxxy = 9 (Pxx uy + 2 Pxy ux - 2 rho ux^2 uy), qxyy = 9 (Pyy ux + 2 Pxy uy - 2 rho uy^2 ux)
  const float t2 = -2.0f * rho * ux * uy;
  const float qxxy = 9.0f * fmaf(pixx, uy, fmaf(2.0f * pixy, ux, t2 * ux));
  const float qxyy = 9.0f * fmaf(piyy, ux, fmaf(2.0f * pixy, uy, t2 * uy));
  const float bulk = fmaf(-0.5f, m2xx + m2yy, rho);
  f[0] = FF_W0_2 * bulk;
  const float ex = fmaf(1.5f, m2xx, bulk), ey = fmaf(1.5f, m2yy, bulk);
  const float ox = FF_W1_2 * fmaf(-0.5f, qxyy, m1x), oy = FF_W1_2 * fmaf(-0.5f, qxxy, m1y);
  f[1] = fmaf(FF_W1_2, ex, ox);
  f[3] = fmaf(FF_W1_2, ex, -ox);
  f[2] = fmaf(FF_W1_2, ey, oy);
  f[4] = fmaf(FF_W1_2, ey, -oy);
  const float dbase = fmaf(1.5f, m2xx + m2yy, bulk);
  const float e5 = dbase + m2xy, e6 = dbase - m2xy;
  const float o5 = FF_W5_2 * ((m1x + m1y) + (qxxy + qxyy)), o6 = FF_W5_2 * ((m1y - m1x) + (qxxy - qxyy));
  f[5] = fmaf(FF_W5_2, e5, o5);
  f[7] = fmaf(FF_W5_2, e5, -o5);
  f[6] = fmaf(FF_W5_2, e6, o6);
  f[8] = fmaf(FF_W5_2, e6, -o6);
}

__device__ __forceinline__ void collide2_fast(const float* p, float omega, float* out) {
  const float s13 = p[1] + p[3], d13 = p[1] - p[3], s24 = p[2] + p[4], d24 = p[2] - p[4];
  const float s57 = p[5] + p[7], d57 = p[5] - p[7], s68 = p[6] + p[8], d68 = p[6] - p[8];
  const float sd = s57 + s68;
  const float rho = (p[0] + (s13 + s24)) + sd;
  const float jx = d13 + (d57 - d68), jy = d24 + (d57 + d68);
  const float pxy = s57 - s68;
  const float inv = __builtin_amdgcn_rcpf(rho);  // v_rcp_f32, 1 ulp
  const float ux = jx * inv, uy = jy * inv;
  // LbmUtilsFuncGpu2D::Collision (LbmUtilsFunc.cu:512-543), F = 0
  const float omw = 1.0f - omega;
  const float ru2 = jx * ux, rv2 = jy * uy;
  const float dev = 0.5f * (s13 - s24);  // (Pxx - Pyy) / 2 of the raw sums
  const float be = fmaf(rho, 1.0f / 3.0f, 0.5f * (ru2 + rv2));
  const float an = (0.5f * omega) * (ru2 - rv2);
  const float nxx = fmaf(dev, omw, be + an);
  const float nyy = fmaf(-dev, omw, be - an);
  const float nxy = fmaf(pxy, omw, (omega * ux) * jy);
  out[0] = rho;
  out[1] = ux;
  out[2] = uy;
  out[3] = fmaf(nxx, inv, -1.0f / 3.0f);
  out[4] = fmaf(nyy, inv, -1.0f / 3.0f);
  out[5] = nxy * inv;
}

// -------------------------------------------------------------------- 3D
#define FF_W0_3 (8.0f / 27.0f)
#define FF_W1_3 (2.0f / 27.0f)
#define FF_W2_3 (1.0f / 54.0f)
#define FF_W3_3 (1.0f / 216.0f)

// CalculateDistributionD3Q27All; m = stored order
// [rho, ux, uy, uz, pxx, pyy, pzz, pxy, pyz, pxz].
__host__ __device__ __forceinline__ void recon3_all(const float* m, float* f) {
  const float rho = m[0], ux = m[1], uy = m[2], uz = m[3];
  const float pixx = m[4], piyy = m[5], pizz = m[6], pixy = m[7], piyz = m[8], pixz = m[9];
  const float m1_x = 3.0f * rho * ux, m1_y = 3.0f * rho * uy, m1_z = 3.0f * rho * uz;
  const float m2_xx = 3.0f * rho * pixx, m2_yy = 3.0f * rho * piyy, m2_zz = 3.0f * rho * pizz;
  const float m2_xy = 9.0f * rho * pixy, m2_yz = 9.0f * rho * piyz, m2_xz = 9.0f * rho * pixz;
#define FF_Q3(Paa, ub, Pab, ua) (9.0f * (rho * (Paa) * (ub) + 2.0f * rho * (Pab) * (ua) - 2.0f * rho * (ua) * (ua) * (ub)))
  const float Q_xxy = FF_Q3(pixx, uy, pixy, ux);
  const float Q_xyy = FF_Q3(piyy, ux, pixy, uy);
  const float Q_xxz = FF_Q3(pixx, uz, pixz, ux);
  const float Q_xzz = FF_Q3(pizz, ux, pixz, uz);
  const float Q_yyz = FF_Q3(piyy, uz, piyz, uy);
  const float Q_yzz = FF_Q3(pizz, uy, piyz, uz);
#undef FF_Q3
  const float Q_xyz = 27.0f * (rho * (pixy * uz + piyz * ux + pixz * uy - 2.0f * ux * uy * uz));
  const float bulk = rho - 0.5f * (m2_xx + m2_yy + m2_zz);
  f[0] = FF_W0_3 * bulk;
  const float abx = bulk + 1.5f * m2_xx;
  f[1] = FF_W1_3 * (abx + m1_x - 0.5f * (Q_xyy + Q_xzz));
  f[2] = FF_W1_3 * (abx - m1_x + 0.5f * (Q_xyy + Q_xzz));
  const float aby = bulk + 1.5f * m2_yy;
  f[3] = FF_W1_3 * (aby + m1_y - 0.5f * (Q_xxy + Q_yzz));
  f[4] = FF_W1_3 * (aby - m1_y + 0.5f * (Q_xxy + Q_yzz));
  const float abz = bulk + 1.5f * m2_zz;
  f[5] = FF_W1_3 * (abz + m1_z - 0.5f * (Q_xxz + Q_yyz));
  f[6] = FF_W1_3 * (abz - m1_z + 0.5f * (Q_xxz + Q_yyz));
  const float ebxy = rho + m2_xx + m2_yy - 0.5f * m2_zz;
  const float eqxy = Q_xxy + Q_xyy;
  const float eoxy = 0.5f * (Q_xzz + Q_yzz);
  f[7] = FF_W2_3 * (ebxy + (m1_x + m1_y) + m2_xy + eqxy - eoxy);
  f[8] = FF_W2_3 * (ebxy - (m1_x + m1_y) + m2_xy - eqxy + eoxy);
  f[13] = FF_W2_3 * (ebxy + (m1_x - m1_y) - m2_xy - (Q_xxy - Q_xyy) - 0.5f * (Q_xzz - Q_yzz));
  f[14] = FF_W2_3 * (ebxy - (m1_x - m1_y) - m2_xy + (Q_xxy - Q_xyy) + 0.5f * (Q_xzz - Q_yzz));
  const float ebxz = rho + m2_xx + m2_zz - 0.5f * m2_yy;
  const float eqxz = Q_xxz + Q_xzz;
  const float eoxz = 0.5f * (Q_xyy + Q_yyz);
  f[9] = FF_W2_3 * (ebxz + (m1_x + m1_z) + m2_xz + eqxz - eoxz);
  f[10] = FF_W2_3 * (ebxz - (m1_x + m1_z) + m2_xz - eqxz + eoxz);
  f[15] = FF_W2_3 * (ebxz + (m1_x - m1_z) - m2_xz - (Q_xxz - Q_xzz) - 0.5f * (Q_xyy - Q_yyz));
  f[16] = FF_W2_3 * (ebxz - (m1_x - m1_z) - m2_xz + (Q_xxz - Q_xzz) + 0.5f * (Q_xyy - Q_yyz));
  const float ebyz = rho + m2_yy + m2_zz - 0.5f * m2_xx;
  const float eqyz = Q_yyz + Q_yzz;
  const float eoyz = 0.5f * (Q_xxy + Q_xxz);
  f[11] = FF_W2_3 * (ebyz + (m1_y + m1_z) + m2_yz + eqyz - eoyz);
  f[12] = FF_W2_3 * (ebyz - (m1_y + m1_z) + m2_yz - eqyz + eoyz);
  f[17] = FF_W2_3 * (ebyz + (m1_y - m1_z) - m2_yz - (Q_yyz - Q_yzz) - 0.5f * (Q_xxy - Q_xxz));
  f[18] = FF_W2_3 * (ebyz - (m1_y - m1_z) - m2_yz + (Q_yyz - Q_yzz) + 0.5f * (Q_xxy - Q_xxz));
  const float cM1 = m1_x + m1_y + m1_z;
  const float cS = m2_xx + m2_yy + m2_zz;
  const float cO = m2_xy + m2_yz + m2_xz;
  const float cQ = Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz + Q_yzz;
  const float cc = rho + cS;
  f[19] = FF_W3_3 * (cc + cM1 + cO + cQ + Q_xyz);
  f[20] = FF_W3_3 * (cc - cM1 + cO - cQ - Q_xyz);
  f[21] = FF_W3_3 * (cc + (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) + (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) - Q_xyz);
  f[22] = FF_W3_3 * (cc - (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) - (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) + Q_xyz);
  f[23] = FF_W3_3 * (cc + (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) + (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) - Q_xyz);
  f[24] = FF_W3_3 * (cc - (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) - (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) + Q_xyz);
  f[25] = FF_W3_3 * (cc + (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) + (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) - Q_xyz);
  f[26] = FF_W3_3 * (cc - (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) - (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) + Q_xyz);
}

// Population i alone, rounded exactly as recon3_all rounds f[i] (same
// operations and order; the direction's +-1 signs selected per lane).
__host__ __device__ __forceinline__ float recon3_pop(const float* m, int i) {
  const float rho = m[0], ux = m[1], uy = m[2], uz = m[3];
  const float pixx = m[4], piyy = m[5], pizz = m[6], pixy = m[7], piyz = m[8], pixz = m[9];
  const float m1_x = 3.0f * rho * ux, m1_y = 3.0f * rho * uy, m1_z = 3.0f * rho * uz;
  const float m2_xx = 3.0f * rho * pixx, m2_yy = 3.0f * rho * piyy, m2_zz = 3.0f * rho * pizz;
  const float m2_xy = 9.0f * rho * pixy, m2_yz = 9.0f * rho * piyz, m2_xz = 9.0f * rho * pixz;
#define FF_Q3(Paa, ub, Pab, ua) (9.0f * (rho * (Paa) * (ub) + 2.0f * rho * (Pab) * (ua) - 2.0f * rho * (ua) * (ua) * (ub)))
  const float Q_xxy = FF_Q3(pixx, uy, pixy, ux);
  const float Q_xyy = FF_Q3(piyy, ux, pixy, uy);
  const float Q_xxz = FF_Q3(pixx, uz, pixz, ux);
  const float Q_xzz = FF_Q3(pizz, ux, pixz, uz);
  const float Q_yyz = FF_Q3(piyy, uz, piyz, uy);
  const float Q_yzz = FF_Q3(pizz, uy, piyz, uz);
#undef FF_Q3
  const float bulk = rho - 0.5f * (m2_xx + m2_yy + m2_zz);
  const int ix = D3::cx(i), iy = D3::cy(i), iz = D3::cz(i);
  const float cx = (float)ix, cy = (float)iy, cz = (float)iz;
  const int n = (ix != 0) + (iy != 0) + (iz != 0);
  if (n == 0) return FF_W0_3 * bulk;
  if (n == 1) {  // axis a; b, c the other two axes in x, y, z order
    const float s = cx + cy + cz;
    const float m1 = ix ? m1_x : (iy ? m1_y : m1_z);
    const float m2 = ix ? m2_xx : (iy ? m2_yy : m2_zz);
    const float qb = ix ? Q_xyy : (iy ? Q_xxy : Q_xxz);
    const float qc = ix ? Q_xzz : (iy ? Q_yzz : Q_yyz);
    const float ab = bulk + 1.5f * m2;
    return FF_W1_3 * (ab + s * m1 - s * (0.5f * (qb + qc)));
  }
  if (n == 2) {  // plane (a, b), a before b; c the axis left out
    const bool xy = iz == 0, xz = iy == 0;
    const float sa = ix ? cx : cy, sb = iz ? cz : cy, t = sa * sb;
    const float m1a = ix ? m1_x : m1_y, m1b = iz ? m1_z : m1_y;
    const float m2aa = ix ? m2_xx : m2_yy, m2bb = iz ? m2_zz : m2_yy;
    const float m2cc = xy ? m2_zz : (xz ? m2_yy : m2_xx);
    const float m2ab = xy ? m2_xy : (xz ? m2_xz : m2_yz);
    const float qaab = xy ? Q_xxy : (xz ? Q_xxz : Q_yyz);
    const float qabb = xy ? Q_xyy : (xz ? Q_xzz : Q_yzz);
    const float qacc = xy ? Q_xzz : (xz ? Q_xyy : Q_xxy);
    const float qbcc = xy ? Q_yzz : (xz ? Q_yyz : Q_xxz);
    const float eb = rho + m2aa + m2bb - 0.5f * m2cc;
    return FF_W2_3 * (eb + sa * (m1a + t * m1b) + t * m2ab + sb * (qaab + t * qabb) - sa * (0.5f * (qacc + t * qbcc)));
  }
  const float Q_xyz = 27.0f * (rho * (pixy * uz + piyz * ux + pixz * uy - 2.0f * ux * uy * uz));
  const float cc = rho + (m2_xx + m2_yy + m2_zz);
  const float M = cx * m1_x + cy * m1_y + cz * m1_z;
  const float O = (cx * cy) * m2_xy + (cy * cz) * m2_yz + (cx * cz) * m2_xz;
  const float Qs = cy * Q_xxy + cz * Q_xxz + cx * Q_xyy + cz * Q_yyz + cx * Q_xzz + cy * Q_yzz;
  return FF_W3_3 * (cc + M + O + Qs + (cx * cy * cz) * Q_xyz);
}

// CalculateDistributionD3Q27AtIndex (LbmUtilsFunc.cu:11-210).
__device__ __forceinline__ float recon3_at(const float* m, int i) {
  const float rho = m[0], ux = m[1], uy = m[2], uz = m[3];
  const float pixx = m[4], piyy = m[5], pizz = m[6], pixy = m[7], piyz = m[8], pixz = m[9];
  const float m1_x = 3.0f * rho * ux, m1_y = 3.0f * rho * uy, m1_z = 3.0f * rho * uz;
  const float m2_xx = 3.0f * rho * pixx, m2_yy = 3.0f * rho * piyy, m2_zz = 3.0f * rho * pizz;
  const float m2_xy = 9.0f * rho * pixy, m2_yz = 9.0f * rho * piyz, m2_xz = 9.0f * rho * pixz;
#define FF_Q3(Paa, ub, Pab, ua) (9.0f * (rho * (Paa) * (ub) + 2.0f * rho * (Pab) * (ua) - 2.0f * rho * (ua) * (ua) * (ub)))
  const float Q_xxy = FF_Q3(pixx, uy, pixy, ux);
  const float Q_xyy = FF_Q3(piyy, ux, pixy, uy);
  const float Q_xxz = FF_Q3(pixx, uz, pixz, ux);
  const float Q_xzz = FF_Q3(pizz, ux, pixz, uz);
  const float Q_yyz = FF_Q3(piyy, uz, piyz, uy);
  const float Q_yzz = FF_Q3(pizz, uy, piyz, uz);
#undef FF_Q3
  const float Q_xyz = 27.0f * (rho * (pixy * uz + piyz * ux + pixz * uy - 2.0f * ux * uy * uz));
  const float b1 = rho - 0.5f * (m2_xx + m2_yy + m2_zz);
  const float b2 = rho + (m2_xx + m2_yy + m2_zz);
  const float h = 0.5f, t = 1.5f;
  float s, w;
  switch (i) {
    case 0: s = b1; w = FF_W0_3; break;
    case 1: s = b1 + m1_x + t * m2_xx - h * (Q_xyy + Q_xzz); w = FF_W1_3; break;
    case 2: s = b1 - m1_x + t * m2_xx + h * (Q_xyy + Q_xzz); w = FF_W1_3; break;
    case 3: s = b1 + m1_y + t * m2_yy - h * (Q_xxy + Q_yzz); w = FF_W1_3; break;
    case 4: s = b1 - m1_y + t * m2_yy + h * (Q_xxy + Q_yzz); w = FF_W1_3; break;
    case 5: s = b1 + m1_z + t * m2_zz - h * (Q_xxz + Q_yyz); w = FF_W1_3; break;
    case 6: s = b1 - m1_z + t * m2_zz + h * (Q_xxz + Q_yyz); w = FF_W1_3; break;
    case 7: s = rho + (m1_x + m1_y) + (m2_xx + m2_yy - h * m2_zz) + m2_xy + (Q_xxy + Q_xyy) - h * (Q_xzz + Q_yzz); w = FF_W2_3; break;
    case 8: s = rho - (m1_x + m1_y) + (m2_xx + m2_yy - h * m2_zz) + m2_xy - (Q_xxy + Q_xyy) + h * (Q_xzz + Q_yzz); w = FF_W2_3; break;
    case 9: s = rho + (m1_x + m1_z) + (m2_xx + m2_zz - h * m2_yy) + m2_xz + (Q_xxz + Q_xzz) - h * (Q_xyy + Q_yyz); w = FF_W2_3; break;
    case 10: s = rho - (m1_x + m1_z) + (m2_xx + m2_zz - h * m2_yy) + m2_xz - (Q_xxz + Q_xzz) + h * (Q_xyy + Q_yyz); w = FF_W2_3; break;
    case 11: s = rho + (m1_y + m1_z) + (m2_yy + m2_zz - h * m2_xx) + m2_yz + (Q_yyz + Q_yzz) - h * (Q_xxy + Q_xxz); w = FF_W2_3; break;
    case 12: s = rho - (m1_y + m1_z) + (m2_yy + m2_zz - h * m2_xx) + m2_yz - (Q_yyz + Q_yzz) + h * (Q_xxy + Q_xxz); w = FF_W2_3; break;
    case 13: s = rho + (m1_x - m1_y) + (m2_xx + m2_yy - h * m2_zz) - m2_xy - (Q_xxy - Q_xyy) - h * (Q_xzz - Q_yzz); w = FF_W2_3; break;
    case 14: s = rho - (m1_x - m1_y) + (m2_xx + m2_yy - h * m2_zz) - m2_xy + (Q_xxy - Q_xyy) + h * (Q_xzz - Q_yzz); w = FF_W2_3; break;
    case 15: s = rho + (m1_x - m1_z) + (m2_xx + m2_zz - h * m2_yy) - m2_xz - (Q_xxz - Q_xzz) - h * (Q_xyy - Q_yyz); w = FF_W2_3; break;
    case 16: s = rho - (m1_x - m1_z) + (m2_xx + m2_zz - h * m2_yy) - m2_xz + (Q_xxz - Q_xzz) + h * (Q_xyy - Q_yyz); w = FF_W2_3; break;
    case 17: s = rho + (m1_y - m1_z) + (m2_yy + m2_zz - h * m2_xx) - m2_yz - (Q_yyz - Q_yzz) - h * (Q_xxy - Q_xxz); w = FF_W2_3; break;
    case 18: s = rho - (m1_y - m1_z) + (m2_yy + m2_zz - h * m2_xx) - m2_yz + (Q_yyz - Q_yzz) + h * (Q_xxy - Q_xxz); w = FF_W2_3; break;
    case 19: s = b2 + (m1_x + m1_y + m1_z) + (m2_xy + m2_yz + m2_xz) + (Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz + Q_yzz) + Q_xyz; w = FF_W3_3; break;
    case 20: s = b2 - (m1_x + m1_y + m1_z) + (m2_xy + m2_yz + m2_xz) - (Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz + Q_yzz) - Q_xyz; w = FF_W3_3; break;
    case 21: s = b2 + (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) + (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) - Q_xyz; w = FF_W3_3; break;
    case 22: s = b2 - (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) - (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) + Q_xyz; w = FF_W3_3; break;
    case 23: s = b2 + (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) + (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) - Q_xyz; w = FF_W3_3; break;
    case 24: s = b2 - (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) - (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) + Q_xyz; w = FF_W3_3; break;
    case 25: s = b2 + (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) + (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) - Q_xyz; w = FF_W3_3; break;
    default: s = b2 - (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) - (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) + Q_xyz; w = FF_W3_3; break;
  }
  return w * s;
}

// Moments of the pulled populations + LbmUtilsFuncGpu3D::Collision
// (LbmUtilsFunc.cu:337-414) + normalisation (FsiCoupling3D.cu:291-320), F = 0.
__device__ __forceinline__ void collide3(const float* p, float omega, float* out) {
  float rho = p[0];
#pragma unroll
  for (int i = 1; i < 27; i++) rho = rho + p[i];
  const float inv = 1.0f / rho;
  const float sxp = p[1] + p[7] + p[9] + p[13] + p[15] + p[19] + p[21] + p[23] + p[26];
  const float sxm = p[2] + p[8] + p[10] + p[14] + p[16] + p[20] + p[22] + p[24] + p[25];
  const float syp = p[3] + p[7] + p[11] + p[14] + p[17] + p[19] + p[21] + p[24] + p[25];
  const float sym = p[4] + p[8] + p[12] + p[13] + p[18] + p[20] + p[22] + p[23] + p[26];
  const float szp = p[5] + p[9] + p[11] + p[16] + p[18] + p[19] + p[22] + p[23] + p[25];
  const float szm = p[6] + p[10] + p[12] + p[15] + p[17] + p[20] + p[21] + p[24] + p[26];
  const float ux = (sxp - sxm) * inv;
  const float uy = (syp - sym) * inv;
  const float uz = (szp - szm) * inv;
  float pixx = p[1] + p[2] + p[7] + p[8] + p[9] + p[10] + p[13] + p[14] + p[15] + p[16] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
  float piyy = p[3] + p[4] + p[7] + p[8] + p[11] + p[12] + p[13] + p[14] + p[17] + p[18] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
  float pizz = p[5] + p[6] + p[9] + p[10] + p[11] + p[12] + p[15] + p[16] + p[17] + p[18] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
  float pixy = (p[7] + p[8] + p[19] + p[20] + p[21] + p[22]) - (p[13] + p[14] + p[23] + p[24] + p[25] + p[26]);
  float piyz = (p[11] + p[12] + p[19] + p[20] + p[25] + p[26]) - (p[17] + p[18] + p[21] + p[22] + p[23] + p[24]);
  float pixz = (p[9] + p[10] + p[19] + p[20] + p[23] + p[24]) - (p[15] + p[16] + p[21] + p[22] + p[25] + p[26]);
  const float third = 0.33333333f;
  const float omw = 1.0f - omega;
  const float r3 = rho * third;
  const float rux = rho * ux * ux, ruy = rho * uy * uy, ruz = rho * uz * uz;
  const float wr = omega * rho;
  pixy = omw * pixy + wr * (ux * uy);
  pixz = omw * pixz + wr * (ux * uz);
  piyz = omw * piyz + wr * (uy * uz);
  const float dxx = (2.0f * pixx - piyy - pizz) * third;
  const float dyy = (2.0f * piyy - pixx - pizz) * third;
  const float dzz = (2.0f * pizz - pixx - piyy) * third;
  const float en = (rux + ruy + ruz) * third;
  const float wt = omega * third;
  pixx = r3 + en + omw * dxx + wt * (2.0f * rux - ruy - ruz);
  piyy = r3 + en + omw * dyy + wt * (2.0f * ruy - rux - ruz);
  pizz = r3 + en + omw * dzz + wt * (2.0f * ruz - rux - ruy);
  out[0] = rho;
  out[1] = ux;
  out[2] = uy;
  out[3] = uz;
  out[4] = pixx * inv - (1.0f / 3.0f);
  out[5] = piyy * inv - (1.0f / 3.0f);
  out[6] = pizz * inv - (1.0f / 3.0f);
  out[7] = pixy * inv;
  out[8] = piyz * inv;
  out[9] = pixz * inv;
}

// ------------------------------------------------------------- 3D, fast form
// The same reconstruction and relaxation written for instruction count rather
// than the reference's rounding: opposite directions share their even part E
// (rho and second-order terms) and differ in the sign of their odd part O
// (first- and third-order terms), f_i = w (E + O), f_inv(i) = w (E - O); the
// moments of the pulled populations come from the 13 pair sums and
// differences; products and sums are contracted to FMA.  Algebraically equal
// to recon3_all / collide3 term by term (every grouping below is written out
// in the comments); only the FP32 rounding differs.
__device__ __forceinline__ void recon3_fast(const float* m, float* f) {
  const float rho = m[0], ux = m[1], uy = m[2], uz = m[3];
  const float pxx = m[4], pyy = m[5], pzz = m[6], pxy = m[7], pyz = m[8], pxz = m[9];
  const float r3 = 3.0f * rho, r9 = 9.0f * rho;
  const float m1x = r3 * ux, m1y = r3 * uy, m1z = r3 * uz;
  const float m2xx = r3 * pxx, m2yy = r3 * pyy, m2zz = r3 * pzz;
  const float m2xy = r9 * pxy, m2yz = r9 * pyz, m2xz = r9 * pxz;
  // Q_aab = 9 rho (P_aa u_b + 2 P_ab u_a - 2 u_a^2 u_b) = 9 rho (u_b D_a + 2 P_ab u_a), D_a = P_aa - 2 u_a^2
  const float Dx = fmaf(-2.0f * ux, ux, pxx), Dy = fmaf(-2.0f * uy, uy, pyy), Dz = fmaf(-2.0f * uz, uz, pzz);
  const float txy = 2.0f * pxy, tyz = 2.0f * pyz, txz = 2.0f * pxz;
  const float Qxxy = r9 * fmaf(uy, Dx, txy * ux);
  const float Qxyy = r9 * fmaf(ux, Dy, txy * uy);
  const float Qxxz = r9 * fmaf(uz, Dx, txz * ux);
  const float Qxzz = r9 * fmaf(ux, Dz, txz * uz);
  const float Qyyz = r9 * fmaf(uz, Dy, tyz * uy);
  const float Qyzz = r9 * fmaf(uy, Dz, tyz * uz);
  // Q_xyz = 27 rho (P_xy u_z + P_yz u_x + P_xz u_y - 2 u_x u_y u_z)
  const float Qxyz = (3.0f * r9) * fmaf(pxy, uz, fmaf(pyz, ux, fmaf(pxz, uy, (-2.0f * ux) * (uy * uz))));
  const float cS = m2xx + m2yy + m2zz;
  const float bulk = fmaf(-0.5f, cS, rho);  // rho - (m2xx + m2yy + m2zz) / 2
  const float cc = rho + cS;
  f[0] = FF_W0_3 * bulk;
  // axes: E = bulk + 3/2 m2_aa, O = m1_a - (Q_abb + Q_acc) / 2
  const float S1 = Qxyy + Qxzz, S2 = Qxxy + Qyzz, S3 = Qxxz + Qyyz;
  {
    const float wx = FF_W1_3 * fmaf(-0.5f, S1, m1x), wy = FF_W1_3 * fmaf(-0.5f, S2, m1y),
                wz = FF_W1_3 * fmaf(-0.5f, S3, m1z);
    const float ex = fmaf(1.5f, m2xx, bulk), ey = fmaf(1.5f, m2yy, bulk), ez = fmaf(1.5f, m2zz, bulk);
    f[1] = fmaf(FF_W1_3, ex, wx);
    f[2] = fmaf(FF_W1_3, ex, -wx);
    f[3] = fmaf(FF_W1_3, ey, wy);
    f[4] = fmaf(FF_W1_3, ey, -wy);
    f[5] = fmaf(FF_W1_3, ez, wz);
    f[6] = fmaf(FF_W1_3, ez, -wz);
  }
  // edges in the ab plane: E = rho + m2_aa + m2_bb - m2_cc / 2 +- m2_ab (= cc - 3/2 m2_cc +- m2_ab),
  // O = H_a +- H_b with H_a = m1_a + Q_abb - Q_acc / 2, H_b = m1_b + Q_aab - Q_bcc / 2
  auto edge = [&](float eb, float m2ab, float ha, float hb, float* fp, float* fpn, float* fm, float* fmn) {
    const float op = FF_W2_3 * (ha + hb), om = FF_W2_3 * (ha - hb);
    const float ep = eb + m2ab, em = eb - m2ab;
    *fp = fmaf(FF_W2_3, ep, op);
    *fpn = fmaf(FF_W2_3, ep, -op);
    *fm = fmaf(FF_W2_3, em, om);
    *fmn = fmaf(FF_W2_3, em, -om);
  };
  edge(fmaf(-1.5f, m2zz, cc), m2xy, fmaf(-0.5f, Qxzz, m1x + Qxyy), fmaf(-0.5f, Qyzz, m1y + Qxxy), &f[7], &f[8],
       &f[13], &f[14]);
  edge(fmaf(-1.5f, m2yy, cc), m2xz, fmaf(-0.5f, Qxyy, m1x + Qxzz), fmaf(-0.5f, Qyyz, m1z + Qxxz), &f[9], &f[10],
       &f[15], &f[16]);
  edge(fmaf(-1.5f, m2xx, cc), m2yz, fmaf(-0.5f, Qxxy, m1y + Qyzz), fmaf(-0.5f, Qxxz, m1z + Qyyz), &f[11], &f[12],
       &f[17], &f[18]);
  // corners (sx, sy, sz): E = cc + sx sy m2_xy + sy sz m2_yz + sx sz m2_xz,
  // O = sx G_x + sy G_y + sz G_z + sx sy sz Q_xyz with G_a = m1_a + (sum of Q_a.. of axis a)
  const float Gx = m1x + S1, Gy = m1y + S2, Gz = m1z + S3;
  const float a = cc + m2xy, b = cc - m2xy, c = m2yz + m2xz, d = m2yz - m2xz;
  const float gp = Gx + Gy, gm = Gx - Gy, g1 = Gz + Qxyz, g2 = Gz - Qxyz;
  auto corner = [&](float e, float o, float* fp, float* fn) {
    const float wo = FF_W3_3 * o;
    *fp = fmaf(FF_W3_3, e, wo);
    *fn = fmaf(FF_W3_3, e, -wo);
  };
  corner(a + c, gp + g1, &f[19], &f[20]);  // (+,+,+)
  corner(a - c, gp - g1, &f[21], &f[22]);  // (+,+,-)
  corner(b - d, gm + g2, &f[23], &f[24]);  // (+,-,+)
  corner(b + d, g2 - gm, &f[25], &f[26]);  // (-,+,+)
}

// collide3 in the fast form: moments from pair sums s and differences d of
// opposite populations, then the same relaxation (F = 0) and normalisation.
__device__ __forceinline__ void collide3_fast(const float* p, float omega, float* out) {
  float s[13], d[13];  // pair k = directions (2k+1, 2k+2)
#pragma unroll
  for (int k = 0; k < 13; k++) {
    s[k] = p[2 * k + 1] + p[2 * k + 2];
    d[k] = p[2 * k + 1] - p[2 * k + 2];
  }
  // pairs: 0 x, 1 y, 2 z, 3 (+x+y), 4 (+x+z), 5 (+y+z), 6 (+x-y), 7 (+x-z), 8 (+y-z),
  //        9 (+,+,+), 10 (+,+,-), 11 (+,-,+), 12 (-,+,+)
  const float exy = s[3] + s[6], exz = s[4] + s[7], eyz = s[5] + s[8];
  const float c1 = s[9] + s[10], c2 = s[11] + s[12], c3 = s[9] - s[10], c4 = s[11] - s[12];
  const float C = c1 + c2;
  const float rho = ((p[0] + C) + (s[0] + s[1] + s[2])) + (exy + exz + eyz);
  const float pxx = s[0] + exy + exz + C;  // raw sums of c_a c_b f
  const float pyy = s[1] + exy + eyz + C;
  const float pzz = s[2] + exz + eyz + C;
  const float pxy = (s[3] - s[6]) + (c1 - c2);
  const float pyz = (s[5] - s[8]) + (c3 - c4);
  const float pxz = (s[4] - s[7]) + (c3 + c4);
  const float e1 = d[9] + d[10], e2 = d[11] - d[12], e3 = d[9] - d[10], e4 = d[11] + d[12];
  const float jx = (d[0] + (d[3] + d[6])) + ((d[4] + d[7]) + (e1 + e2));
  const float jy = (d[1] + (d[3] - d[6])) + ((d[5] + d[8]) + (e1 - e2));
  const float jz = (d[2] + (d[4] - d[7])) + ((d[5] - d[8]) + (e3 + e4));
  const float inv = __builtin_amdgcn_rcpf(rho);  // v_rcp_f32, 1 ulp
  const float ux = jx * inv, uy = jy * inv, uz = jz * inv;
  // Collision (LbmUtilsFunc.cu:337-414), F = 0: trace at rho/3 + rho|u|^2/3 per axis,
  // deviatoric and off-diagonal parts relaxed by (1 - omega) towards rho u_a u_b
  const float third = 1.0f / 3.0f, omw = 1.0f - omega;
  const float rux = jx * ux, ruy = jy * uy, ruz = jz * uz;
  const float t3 = (pxx + pyy + pzz) * third, r3 = (rux + ruy + ruz) * third;
  const float base = fmaf(rho, third, r3);
  const float nxx = fmaf(omw, pxx - t3, fmaf(omega, rux - r3, base));
  const float nyy = fmaf(omw, pyy - t3, fmaf(omega, ruy - r3, base));
  const float nzz = fmaf(omw, pzz - t3, fmaf(omega, ruz - r3, base));
  const float wux = omega * ux, wuy = omega * uy;
  const float nxy = fmaf(omw, pxy, wux * jy);
  const float nxz = fmaf(omw, pxz, wux * jz);
  const float nyz = fmaf(omw, pyz, wuy * jz);
  out[0] = rho;
  out[1] = ux;
  out[2] = uy;
  out[3] = uz;
  out[4] = fmaf(nxx, inv, -third);
  out[5] = fmaf(nyy, inv, -third);
  out[6] = fmaf(nzz, inv, -third);
  out[7] = nxy * inv;
  out[8] = nyz * inv;
  out[9] = nxz * inv;
}

// -------------------------------------------------------------------- storage
// 16-bit storage: slot 0 holds rho-1, the velocity slots u, the rest Pi, each
// scaled by a power of two (exact) before rounding.
template <int DIM>
__device__ __forceinline__ float slot_scale(int k) { return k == 0 ? 256.0f : (k <= DIM ? 16.0f : 256.0f); }
template <int DIM>
__device__ __forceinline__ float slot_inv_scale(int k) { return k == 0 ? (1.0f / 256.0f) : (k <= DIM ? (1.0f / 16.0f) : (1.0f / 256.0f)); }

// Element of the split-velocity 16-bit storage (FF_STORAGE_F16U): every plane
// holds binary16 values; rho - 1 and Pi are stored as in FF_STORAGE_F16, and
// each velocity component u_a (scaled by 2^4) as a pair hi + lo of binary16
// values, hi in the component's own plane and lo = e - hi in an extra plane
// NM - 1 planes further on (planes NM .. NM + DIM - 1).  e - hi is exact in
// FP32 and so is hi + lo (22 significant bits), so a stored velocity keeps
// about FP32 precision: the velocity rounding dominates the FP16 storage
// error, rho - 1 and Pi rounding much less.
struct h16u {
  unsigned short bits;
};

// Moment loads and stores by storage type: p points at the cell's value in
// plane k; `plane` is the plane stride in elements (used by the split
// velocity of h16u).
template <typename S> struct Store;
template <> struct Store<float> {
  template <int DIM> __device__ static __forceinline__ float load(const float* p, int k, size_t plane) { return *p; }
  template <int DIM> __device__ static __forceinline__ void store(float* p, int k, float v, size_t plane) { *p = v; }
};
template <> struct Store<__half> {
  template <int DIM> __device__ static __forceinline__ float load(const __half* p, int k, size_t plane) {
    float v = __half2float(*p) * slot_inv_scale<DIM>(k);
    return k == 0 ? v + 1.0f : v;
  }
  template <int DIM> __device__ static __forceinline__ void store(__half* p, int k, float v, size_t plane) {
    float e = (k == 0 ? v - 1.0f : v) * slot_scale<DIM>(k);
    *p = to_half_rn(e);
  }
};
template <> struct Store<__hip_bfloat16> {
  template <int DIM> __device__ static __forceinline__ float load(const __hip_bfloat16* p, int k, size_t plane) {
    float v = __bfloat162float(*p) * slot_inv_scale<DIM>(k);
    return k == 0 ? v + 1.0f : v;
  }
  template <int DIM> __device__ static __forceinline__ void store(__hip_bfloat16* p, int k, float v, size_t plane) {
    float e = (k == 0 ? v - 1.0f : v) * slot_scale<DIM>(k);
    *p = __float2bfloat16(opaque(e));
  }
};
template <> struct Store<h16u> {
  template <int DIM> __device__ static constexpr size_t lo_offset(size_t plane) {
    return (size_t)((DIM == 2 ? 6 : 10) - 1) * plane;
  }
  template <int DIM> __device__ static __forceinline__ float load(const h16u* p, int k, size_t plane) {
    const float hi = __half2float(__ushort_as_half(p->bits));
    if (k >= 1 && k <= DIM) {
      const float lo = __half2float(__ushort_as_half(p[lo_offset<DIM>(plane)].bits));
      return (hi + lo) * (1.0f / 16.0f);
    }
    const float v = hi * slot_inv_scale<DIM>(k);
    return k == 0 ? v + 1.0f : v;
  }
  template <int DIM> __device__ static __forceinline__ void store(h16u* p, int k, float v, size_t plane) {
    const float e = (k == 0 ? v - 1.0f : v) * slot_scale<DIM>(k);
    const __half h = to_half_rn(e);
    p->bits = __half_as_ushort(h);
    if (k >= 1 && k <= DIM) p[lo_offset<DIM>(plane)].bits = __half_as_ushort(to_half_rn(e - __half2float(h)));
  }
};

}  // namespace ff
