// lbm_oracle.cpp -- CPU restatement of FreeFlow's HOME-LBM fluid step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is compiled
// into oracle/_build/liboracle.so and loaded solely by tests/, by
// __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg.
// Nothing in the product (freeflow_amd/) includes, links or calls it.
//
// It restates, operation by operation, the reference CUDA path (paths are
// relative to the FreeFlow source tree):
//   lattice tables ............ include/FSI_Simulator/lbm/LbmConstants.hpp:20-59
//   node flags ................ include/FSI_Simulator/lbm/LbmDataTypes.cuh:17-56
//   3D reconstruction/collision src/lbm/LbmUtilsFunc.cu:11-414
//   2D reconstruction/collision src/lbm/LbmUtilsFunc.cu:422-543
//   2D fused step + outlet .... src/core/FsiCoupling2D.cu:31-316, 358-432
//   3D fused step + outlet .... src/core/FsiCoupling3D.cu:31-420, 424-484
//   geometry predicates ....... include/FSI_Simulator/core/SolidGeometryProxy_Device.cuh:79-440
//   tolerance helpers ......... include/FSI_Simulator/utils/NumericUtils.cuh:23-94
//   initial state ............. src/lbm/LbmInitializer2D.cu:32-149, LbmInitializer3D.cu:33-165
//
// The float instantiation (R = float) keeps every double-promotion site of the
// reference (e.g. `- 0.5 * Q_xyy` in CalculateDistributionD2Q9ALL) so that it
// reproduces the reference's FP32 rounding; the double instantiation (R =
// double) is the FP64 oracle used for error bars.  Geometry (vertex positions,
// link/segment predicates, barycentrics) is always evaluated in float, as the
// reference's proxy is float, so both instantiations take the same discrete
// link decisions.
//
// Compile with -ffp-contract=off: the reference's predicates and the FP16 tile
// rounding are discontinuous, so the oracle must not fuse multiplies and adds.
//
// Deviations from the reference that are switches (reference behaviour is the
// default):
//   stage_fp16 (2D only, default on): the reference stages reconstructed
//     populations in an FP16 LDS tile (FsiCoupling2D.cu:89,120-128), so every
//     pulled population is rounded to binary16.  Off = pure FP32 pulls.
//   periodic (default off): pulls wrap around the domain instead of using the
//     reference's clamped halo (FsiCoupling2D.cu:104-105).
//   storage (default FP32): after each step the stored moments are quantised
//     to FP16/BF16 with the encoding of the product's 16-bit storage modes
//     (3: FP16 with each velocity component kept as a hi + lo FP16 pair).
// Reproduced quirk: in 2D a population whose upstream node is a wall is
// bounce-back-assigned and then overwritten by the ordinary pull
// (FsiCoupling2D.cu:168-223), so 2D walls act through their frozen initial
// moments; in 3D the bounce-back value is kept (FsiCoupling3D.cu:209-284).
// Not reproduced: the out-of-bounds LDS write at FsiCoupling2D.cu:129 (it
// writes past the tile and never feeds a result).
// Force accumulation is sequential in (cell, direction) order in double; the
// reference uses order-nondeterministic float atomicAdd
// (SolidGeometryProxy_Device.cuh:190-193, :415-420).

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Node flags (LbmDataTypes.cuh:17-56).  Values are the enum's ordinal.
enum Flag : uint8_t {
  Invalid = 0, Fluid, FluidRest, Solid, SolidDynamic, Wall,
  WallLeft, WallRight, WallFront, WallBack, WallDown, WallUp, WallCorner,
  InletLeft, InletRight, InletUp, InletDown, InletFront, InletBack,
  OutletLeft, OutletRight, OutletUp, OutletDown, OutletFront, OutletBack,
  Receptor, Deactivated, Smoke
};

// ---------------------------------------------------------------------------
// Lattice tables (LbmConstants.hpp:24-27, 48-54).
static const float ex2[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const float ey2[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int inv2[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

static const float ex3[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
static const float ey3[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
static const float ez3[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
static const int inv3[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

template <class R> struct W2 {
  static R w(int i) {
    const R a = R(4) / R(9), b = R(1) / R(9), c = R(1) / R(36);
    return i == 0 ? a : (i < 5 ? b : c);
  }
  static R cs2() { return R(1) / R(3); }
};
template <class R> struct W3 {
  static R w(int i) {
    if (i == 0) return R(8) / R(27);
    if (i < 7) return R(2) / R(27);
    if (i < 19) return R(1) / R(54);
    return R(1) / R(216);
  }
  static R cs2() { return R(1) / R(3); }
};

// ---------------------------------------------------------------------------
// IEEE-754 binary16 round-to-nearest-even (the CUDA `half` store in the 2D
// LDS tile).  Written from the format definition.
static inline uint16_t f2h(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) {  // inf / nan
    return (uint16_t)(sign | (ax > 0x7F800000u ? 0x7E00u : 0x7C00u));
  }
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to inf
  int e = (int)(ax >> 23) - 127;
  uint32_t mant = ax & 0x7FFFFFu;
  if (e >= -14) {  // normal half
    uint32_t hm = mant >> 13;
    uint32_t rem = mant & 0x1FFFu;
    uint32_t h = ((uint32_t)(e + 15) << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1;  // carry may bump exponent: correct
    return (uint16_t)(sign | h);
  }
  // subnormal half: value = m * 2^-24
  if (e < -25) return (uint16_t)sign;  // below half of smallest subnormal -> 0 (e=-25 handled below)
  uint32_t full = mant | 0x800000u;   // 24-bit significand, value = full * 2^(e-23)
  int shift = (-14 - e) + 13;         // bits to drop to get units of 2^-24
  uint32_t hm = full >> shift;
  uint32_t rem = full & ((1u << shift) - 1u);
  uint32_t half = 1u << (shift - 1);
  if (rem > half || (rem == half && (hm & 1u))) hm += 1;
  return (uint16_t)(sign | hm);
}
static inline float h2f(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1Fu;
  uint32_t m = h & 0x3FFu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {
      float v = (float)m * 5.9604644775390625e-8f;  // m * 2^-24, exact
      std::memcpy(&x, &v, 4);
      x |= sign;
    }
  } else if (e == 31) {
    x = sign | 0x7F800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float r;
  std::memcpy(&r, &x, 4);
  return r;
}
// bfloat16 round-to-nearest-even (product's BF16 storage mode).
static inline uint16_t f2bf(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  if ((x & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x >> 16) | 0x40u);
  uint32_t lsb = (x >> 16) & 1u;
  x += 0x7FFFu + lsb;
  return (uint16_t)(x >> 16);
}
static inline float bf2f(uint16_t b) {
  uint32_t x = (uint32_t)b << 16;
  float r;
  std::memcpy(&r, &x, 4);
  return r;
}

// ---------------------------------------------------------------------------
// 2D Hermite reconstruction, all 9 populations (LbmUtilsFunc.cu:476-510).
// Arguments in the reference's parameter order (rho, ux, uy, pixx, pixy, piyy).
// Note the reference's 2D Q terms omit rho (:488-490) and the axis terms carry
// a double `0.5 * Q` (:498-501); both are kept.
template <class R>
static void recon2d_all(R rho, R ux, R uy, R pixx, R pixy, R piyy, R* f) {
  R m1_x = R(3) * rho * ux;
  R m1_y = R(3) * rho * uy;
  R m2_xx = R(3) * rho * pixx;
  R m2_yy = R(3) * rho * piyy;
  R m2_xy = R(9) * rho * pixy;
  auto calc_Q = [&](R Pi_aa, R u_b, R Pi_ab, R u_a) {
    return R(9) * (Pi_aa * u_b + R(2) * Pi_ab * u_a - R(2) * rho * u_a * u_a * u_b);
  };
  R Q_xxy = calc_Q(pixx, uy, pixy, ux);
  R Q_xyy = calc_Q(piyy, ux, pixy, uy);
  R bulk = rho - R(0.5) * (m2_xx + m2_yy);
  f[0] = W2<R>::w(0) * bulk;
  f[1] = (R)((double)W2<R>::w(1) * ((double)(bulk + m1_x + R(1.5) * m2_xx) - 0.5 * (double)Q_xyy));
  f[2] = (R)((double)W2<R>::w(2) * ((double)(bulk + m1_y + R(1.5) * m2_yy) - 0.5 * (double)Q_xxy));
  f[3] = (R)((double)W2<R>::w(3) * ((double)(bulk - m1_x + R(1.5) * m2_xx) + 0.5 * (double)Q_xyy));
  f[4] = (R)((double)W2<R>::w(4) * ((double)(bulk - m1_y + R(1.5) * m2_yy) + 0.5 * (double)Q_xxy));
  R diag_base = bulk + R(1.5) * (m2_xx + m2_yy);
  f[5] = W2<R>::w(5) * (diag_base + (m1_x + m1_y) + m2_xy + (Q_xxy + Q_xyy));
  f[6] = W2<R>::w(6) * (diag_base - m1_x + m1_y - m2_xy + Q_xxy - Q_xyy);
  f[7] = W2<R>::w(7) * (diag_base - m1_x - m1_y + m2_xy - Q_xxy - Q_xyy);
  f[8] = W2<R>::w(8) * (diag_base + m1_x - m1_y - m2_xy - Q_xxy + Q_xyy);
}

// Single population (LbmUtilsFunc.cu:422-474); used for the moving-wall value.
template <class R>
static R recon2d_at(R rho, R ux, R uy, R pixx, R pixy, R piyy, int i) {
  R m1_x = R(3) * rho * ux;
  R m1_y = R(3) * rho * uy;
  R m2_xx = R(3) * rho * pixx;
  R m2_yy = R(3) * rho * piyy;
  R m2_xy = R(9) * rho * pixy;
  auto calc_Q = [&](R Pi_aa, R u_b, R Pi_ab, R u_a) {
    return R(9) * (Pi_aa * u_b + R(2) * Pi_ab * u_a - R(2) * rho * u_a * u_a * u_b);
  };
  R Q_xxy = calc_Q(pixx, uy, pixy, ux);
  R Q_xyy = calc_Q(piyy, ux, pixy, uy);
  R bulk = rho - R(0.5) * (m2_xx + m2_yy);
  R diag_base = bulk + R(1.5) * (m2_xx + m2_yy);
  R w = W2<R>::w(i);
  R sum = R(0);
  switch (i) {
    case 0: sum = bulk; break;
    case 1: sum = bulk + m1_x + R(1.5) * m2_xx - R(0.5) * Q_xyy; break;
    case 2: sum = bulk + m1_y + R(1.5) * m2_yy - R(0.5) * Q_xxy; break;
    case 3: sum = bulk - m1_x + R(1.5) * m2_xx + R(0.5) * Q_xyy; break;
    case 4: sum = bulk - m1_y + R(1.5) * m2_yy + R(0.5) * Q_xxy; break;
    case 5: sum = diag_base + (m1_x + m1_y) + m2_xy + (Q_xxy + Q_xyy); break;
    case 6: sum = diag_base - m1_x + m1_y - m2_xy + Q_xxy - Q_xyy; break;
    case 7: sum = diag_base - m1_x - m1_y + m2_xy - Q_xxy - Q_xyy; break;
    case 8: sum = diag_base + m1_x - m1_y - m2_xy - Q_xxy + Q_xyy; break;
  }
  return w * sum;
}

// 2D moment-space collision on raw second moments (LbmUtilsFunc.cu:512-543).
template <class R>
static void collide2d(R rho, R ux, R uy, R Fx, R Fy, R omega, R& pixx, R& piyy, R& pixy) {
  R one_minus_w = R(1) - omega;
  R p = rho / R(3);
  R ru2 = rho * ux * ux;
  R rv2 = rho * uy * uy;
  R dev = (pixx - piyy) * R(0.5);
  R bulk_energy = p + R(0.5) * (ru2 + rv2);
  R aniso_eq = R(0.5) * omega * (ru2 - rv2);
  R force_factor = R(0.5) * one_minus_w;
  R f_term_x = Fx * ux + force_factor * (Fx * ux - Fy * uy);
  R f_term_y = Fy * uy + force_factor * (Fy * uy - Fx * ux);
  R pi_xx_new = bulk_energy + dev * one_minus_w + aniso_eq + f_term_x;
  R pi_yy_new = bulk_energy - dev * one_minus_w - aniso_eq + f_term_y;
  R pi_xy_new = pixy * one_minus_w + (omega * rho * ux * uy) + (R(1) - R(0.5) * omega) * (Fy * ux + Fx * uy);
  pixx = pi_xx_new;
  piyy = pi_yy_new;
  pixy = pi_xy_new;
}

// ---------------------------------------------------------------------------
// 3D reconstruction, all 27 (LbmUtilsFunc.cu:212-335).
template <class R>
static void recon3d_all(R rho, R ux, R uy, R uz, R pixx, R piyy, R pizz, R pixy, R piyz, R pixz, R* f) {
  R m1_x = R(3) * rho * ux, m1_y = R(3) * rho * uy, m1_z = R(3) * rho * uz;
  R m2_xx = R(3) * rho * pixx, m2_yy = R(3) * rho * piyy, m2_zz = R(3) * rho * pizz;
  R m2_xy = R(9) * rho * pixy, m2_yz = R(9) * rho * piyz, m2_xz = R(9) * rho * pixz;
  auto calc_Q = [&](R Pi_aa, R u_b, R Pi_ab, R u_a) {
    return R(9) * (rho * Pi_aa * u_b + R(2) * rho * Pi_ab * u_a - R(2) * rho * u_a * u_a * u_b);
  };
  R Q_xxy = calc_Q(pixx, uy, pixy, ux);
  R Q_xyy = calc_Q(piyy, ux, pixy, uy);
  R Q_xxz = calc_Q(pixx, uz, pixz, ux);
  R Q_xzz = calc_Q(pizz, ux, pixz, uz);
  R Q_yyz = calc_Q(piyy, uz, piyz, uy);
  R Q_yzz = calc_Q(pizz, uy, piyz, uz);
  R Q_xyz = R(27) * (rho * (pixy * uz + piyz * ux + pixz * uy - R(2) * ux * uy * uz));
  R bulk = rho - R(0.5) * (m2_xx + m2_yy + m2_zz);
  R w0 = W3<R>::w(0), w1 = W3<R>::w(1), w2 = W3<R>::w(7), w3 = W3<R>::w(19);
  f[0] = w0 * bulk;
  R axis_base_x = bulk + R(1.5) * m2_xx;
  f[1] = w1 * (axis_base_x + m1_x - R(0.5) * (Q_xyy + Q_xzz));
  f[2] = w1 * (axis_base_x - m1_x + R(0.5) * (Q_xyy + Q_xzz));
  R axis_base_y = bulk + R(1.5) * m2_yy;
  f[3] = w1 * (axis_base_y + m1_y - R(0.5) * (Q_xxy + Q_yzz));
  f[4] = w1 * (axis_base_y - m1_y + R(0.5) * (Q_xxy + Q_yzz));
  R axis_base_z = bulk + R(1.5) * m2_zz;
  f[5] = w1 * (axis_base_z + m1_z - R(0.5) * (Q_xxz + Q_yyz));
  f[6] = w1 * (axis_base_z - m1_z + R(0.5) * (Q_xxz + Q_yyz));
  R edge_base_xy = rho + m2_xx + m2_yy - R(0.5) * m2_zz;
  R edge_Q_xy = Q_xxy + Q_xyy;
  R edge_Q_orth_xy = R(0.5) * (Q_xzz + Q_yzz);
  f[7] = w2 * (edge_base_xy + (m1_x + m1_y) + m2_xy + edge_Q_xy - edge_Q_orth_xy);
  f[8] = w2 * (edge_base_xy - (m1_x + m1_y) + m2_xy - edge_Q_xy + edge_Q_orth_xy);
  f[13] = w2 * (edge_base_xy + (m1_x - m1_y) - m2_xy - (Q_xxy - Q_xyy) - R(0.5) * (Q_xzz - Q_yzz));
  f[14] = w2 * (edge_base_xy - (m1_x - m1_y) - m2_xy + (Q_xxy - Q_xyy) + R(0.5) * (Q_xzz - Q_yzz));
  R edge_base_xz = rho + m2_xx + m2_zz - R(0.5) * m2_yy;
  R edge_Q_xz = Q_xxz + Q_xzz;
  R edge_Q_orth_xz = R(0.5) * (Q_xyy + Q_yyz);
  f[9] = w2 * (edge_base_xz + (m1_x + m1_z) + m2_xz + edge_Q_xz - edge_Q_orth_xz);
  f[10] = w2 * (edge_base_xz - (m1_x + m1_z) + m2_xz - edge_Q_xz + edge_Q_orth_xz);
  f[15] = w2 * (edge_base_xz + (m1_x - m1_z) - m2_xz - (Q_xxz - Q_xzz) - R(0.5) * (Q_xyy - Q_yyz));
  f[16] = w2 * (edge_base_xz - (m1_x - m1_z) - m2_xz + (Q_xxz - Q_xzz) + R(0.5) * (Q_xyy - Q_yyz));
  R edge_base_yz = rho + m2_yy + m2_zz - R(0.5) * m2_xx;
  R edge_Q_yz = Q_yyz + Q_yzz;
  R edge_Q_orth_yz = R(0.5) * (Q_xxy + Q_xxz);
  f[11] = w2 * (edge_base_yz + (m1_y + m1_z) + m2_yz + edge_Q_yz - edge_Q_orth_yz);
  f[12] = w2 * (edge_base_yz - (m1_y + m1_z) + m2_yz - edge_Q_yz + edge_Q_orth_yz);
  f[17] = w2 * (edge_base_yz + (m1_y - m1_z) - m2_yz - (Q_yyz - Q_yzz) - R(0.5) * (Q_xxy - Q_xxz));
  f[18] = w2 * (edge_base_yz - (m1_y - m1_z) - m2_yz + (Q_yyz - Q_yzz) + R(0.5) * (Q_xxy - Q_xxz));
  R c_M1 = m1_x + m1_y + m1_z;
  R c_S = m2_xx + m2_yy + m2_zz;
  R c_O = m2_xy + m2_yz + m2_xz;
  R c_Q = Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz + Q_yzz;
  R corner_common = rho + c_S;
  f[19] = w3 * (corner_common + c_M1 + c_O + c_Q + Q_xyz);
  f[20] = w3 * (corner_common - c_M1 + c_O - c_Q - Q_xyz);
  f[21] = w3 * (corner_common + (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) + (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) - Q_xyz);
  f[22] = w3 * (corner_common - (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) - (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) + Q_xyz);
  f[23] = w3 * (corner_common + (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) + (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) - Q_xyz);
  f[24] = w3 * (corner_common - (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) - (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) + Q_xyz);
  f[25] = w3 * (corner_common + (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) + (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) - Q_xyz);
  f[26] = w3 * (corner_common - (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) - (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) + Q_xyz);
}

// 3D single population (LbmUtilsFunc.cu:11-210).
template <class R>
static R recon3d_at(R rho, R ux, R uy, R uz, R pixx, R piyy, R pizz, R pixy, R piyz, R pixz, int i) {
  R m1_x = R(3) * rho * ux, m1_y = R(3) * rho * uy, m1_z = R(3) * rho * uz;
  R m2_xx = R(3) * rho * pixx, m2_yy = R(3) * rho * piyy, m2_zz = R(3) * rho * pizz;
  R m2_xy = R(9) * rho * pixy, m2_yz = R(9) * rho * piyz, m2_xz = R(9) * rho * pixz;
  auto calc_Q_mixed = [&](R Pi_aa, R u_b, R Pi_ab, R u_a) {
    return R(9) * (rho * Pi_aa * u_b + R(2) * rho * Pi_ab * u_a - R(2) * rho * u_a * u_a * u_b);
  };
  R Q_xxy = calc_Q_mixed(pixx, uy, pixy, ux);
  R Q_xyy = calc_Q_mixed(piyy, ux, pixy, uy);
  R Q_xxz = calc_Q_mixed(pixx, uz, pixz, ux);
  R Q_xzz = calc_Q_mixed(pizz, ux, pixz, uz);
  R Q_yyz = calc_Q_mixed(piyy, uz, piyz, uy);
  R Q_yzz = calc_Q_mixed(pizz, uy, piyz, uz);
  R Q_xyz = R(27) * (rho * (pixy * uz + piyz * ux + pixz * uy - R(2) * ux * uy * uz));
  R bulk_part = rho - R(0.5) * (m2_xx + m2_yy + m2_zz);
  R bulk_part2 = rho + (m2_xx + m2_yy + m2_zz);
  R w = W3<R>::w(i);
  R sum = R(0);
  const R h = R(0.5), t = R(1.5);
  switch (i) {
    case 0: sum = bulk_part; break;
    case 1: sum = bulk_part + m1_x + t * m2_xx - h * (Q_xyy + Q_xzz); break;
    case 2: sum = bulk_part - m1_x + t * m2_xx + h * (Q_xyy + Q_xzz); break;
    case 3: sum = bulk_part + m1_y + t * m2_yy - h * (Q_xxy + Q_yzz); break;
    case 4: sum = bulk_part - m1_y + t * m2_yy + h * (Q_xxy + Q_yzz); break;
    case 5: sum = bulk_part + m1_z + t * m2_zz - h * (Q_xxz + Q_yyz); break;
    case 6: sum = bulk_part - m1_z + t * m2_zz + h * (Q_xxz + Q_yyz); break;
    case 7: sum = rho + (m1_x + m1_y) + (m2_xx + m2_yy - h * m2_zz) + m2_xy + (Q_xxy + Q_xyy) - h * (Q_xzz + Q_yzz); break;
    case 8: sum = rho - (m1_x + m1_y) + (m2_xx + m2_yy - h * m2_zz) + m2_xy - (Q_xxy + Q_xyy) + h * (Q_xzz + Q_yzz); break;
    case 9: sum = rho + (m1_x + m1_z) + (m2_xx + m2_zz - h * m2_yy) + m2_xz + (Q_xxz + Q_xzz) - h * (Q_xyy + Q_yyz); break;
    case 10: sum = rho - (m1_x + m1_z) + (m2_xx + m2_zz - h * m2_yy) + m2_xz - (Q_xxz + Q_xzz) + h * (Q_xyy + Q_yyz); break;
    case 11: sum = rho + (m1_y + m1_z) + (m2_yy + m2_zz - h * m2_xx) + m2_yz + (Q_yyz + Q_yzz) - h * (Q_xxy + Q_xxz); break;
    case 12: sum = rho - (m1_y + m1_z) + (m2_yy + m2_zz - h * m2_xx) + m2_yz - (Q_yyz + Q_yzz) + h * (Q_xxy + Q_xxz); break;
    case 13: sum = rho + (m1_x - m1_y) + (m2_xx + m2_yy - h * m2_zz) - m2_xy - (Q_xxy - Q_xyy) - h * (Q_xzz - Q_yzz); break;
    case 14: sum = rho - (m1_x - m1_y) + (m2_xx + m2_yy - h * m2_zz) - m2_xy + (Q_xxy - Q_xyy) + h * (Q_xzz - Q_yzz); break;
    case 15: sum = rho + (m1_x - m1_z) + (m2_xx + m2_zz - h * m2_yy) - m2_xz - (Q_xxz - Q_xzz) - h * (Q_xyy - Q_yyz); break;
    case 16: sum = rho - (m1_x - m1_z) + (m2_xx + m2_zz - h * m2_yy) - m2_xz + (Q_xxz - Q_xzz) + h * (Q_xyy - Q_yyz); break;
    case 17: sum = rho + (m1_y - m1_z) + (m2_yy + m2_zz - h * m2_xx) - m2_yz - (Q_yyz - Q_yzz) - h * (Q_xxy - Q_xxz); break;
    case 18: sum = rho - (m1_y - m1_z) + (m2_yy + m2_zz - h * m2_xx) - m2_yz + (Q_yyz - Q_yzz) + h * (Q_xxy - Q_xxz); break;
    case 19: sum = bulk_part2 + (m1_x + m1_y + m1_z) + (m2_xy + m2_yz + m2_xz) + (Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz + Q_yzz) + Q_xyz; break;
    case 20: sum = bulk_part2 - (m1_x + m1_y + m1_z) + (m2_xy + m2_yz + m2_xz) - (Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz + Q_yzz) - Q_xyz; break;
    case 21: sum = bulk_part2 + (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) + (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) - Q_xyz; break;
    case 22: sum = bulk_part2 - (m1_x + m1_y - m1_z) + (m2_xy - m2_yz - m2_xz) - (Q_xxy - Q_xxz + Q_xyy - Q_yyz + Q_xzz + Q_yzz) + Q_xyz; break;
    case 23: sum = bulk_part2 + (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) + (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) - Q_xyz; break;
    case 24: sum = bulk_part2 - (m1_x - m1_y + m1_z) + (-m2_xy - m2_yz + m2_xz) - (-Q_xxy + Q_xxz + Q_xyy + Q_yyz + Q_xzz - Q_yzz) + Q_xyz; break;
    case 25: sum = bulk_part2 + (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) + (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) - Q_xyz; break;
    case 26: sum = bulk_part2 - (-m1_x + m1_y + m1_z) + (-m2_xy + m2_yz - m2_xz) - (Q_xxy + Q_xxz - Q_xyy + Q_yyz - Q_xzz + Q_yzz) + Q_xyz; break;
  }
  return w * sum;
}

// 3D collision (LbmUtilsFunc.cu:337-414).
template <class R>
static void collide3d(R rho, R ux, R uy, R uz, R Fx, R Fy, R Fz, R omega,
                      R& pixx, R& piyy, R& pizz, R& pixy, R& piyz, R& pixz) {
  const R third = std::is_same<R, float>::value ? R(0.33333333f) : R(1) / R(3);
  const R two_third = std::is_same<R, float>::value ? R(0.6666667f) : R(2) / R(3);
  R one_minus_w = R(1) - omega;
  R dev_force_coeff = one_minus_w * R(0.5);
  R off_diag_force_coeff = R(1) - R(0.5) * omega;
  R r_3 = rho * third;
  R ruu_x = rho * ux * ux;
  R ruu_y = rho * uy * uy;
  R ruu_z = rho * uz * uz;
  R w_rho = omega * rho;
  pixy = one_minus_w * pixy + w_rho * (ux * uy) + off_diag_force_coeff * (Fx * uy + Fy * ux);
  pixz = one_minus_w * pixz + w_rho * (ux * uz) + off_diag_force_coeff * (Fx * uz + Fz * ux);
  piyz = one_minus_w * piyz + w_rho * (uy * uz) + off_diag_force_coeff * (Fy * uz + Fz * uy);
  R s_xx = pixx, s_yy = piyy, s_zz = pizz;
  R dev_xx = (R(2) * s_xx - s_yy - s_zz) * third;
  R dev_yy = (R(2) * s_yy - s_xx - s_zz) * third;
  R dev_zz = (R(2) * s_zz - s_xx - s_yy) * third;
  R fu_x = Fx * ux, fu_y = Fy * uy, fu_z = Fz * uz;
  R f_dev_common = (dev_force_coeff * two_third);
  R energy = (ruu_x + ruu_y + ruu_z) * third;
  R relax_u2_x = (omega * third) * (R(2) * ruu_x - ruu_y - ruu_z);
  R force_x = fu_x + f_dev_common * (R(2) * fu_x - fu_y - fu_z);
  pixx = r_3 + energy + one_minus_w * dev_xx + relax_u2_x + force_x;
  R relax_u2_y = (omega * third) * (R(2) * ruu_y - ruu_x - ruu_z);
  R force_y = fu_y + f_dev_common * (R(2) * fu_y - fu_x - fu_z);
  piyy = r_3 + energy + one_minus_w * dev_yy + relax_u2_y + force_y;
  R relax_u2_z = (omega * third) * (R(2) * ruu_z - ruu_x - ruu_y);
  R force_z = fu_z + f_dev_common * (R(2) * fu_z - fu_x - fu_y);
  pizz = r_3 + energy + one_minus_w * dev_zz + relax_u2_z + force_z;
}

// ---------------------------------------------------------------------------
// Geometry (SolidGeometryProxy_Device.cuh).  Always float.
struct Hit {
  float hit[3];
  float depth;
  int elem;
  float bary[3];
};

// depthIndexCompact (:79-85)
static inline uint32_t depth_index_compact(float depth, int index) {
  uint32_t ret = 0;
  ret |= ((uint32_t)(depth * (float)(1 << 15))) << 16;
  ret |= (uint16_t)(index);
  return ret;
}

// line_intersect (:89-124).  Only the x-extent of the segment is tested
// (:117), so exactly vertical segments never report a hit.
static inline float line_intersect(float x1, float y1, float x2, float y2, float x3, float y3, float x4, float y4,
                                   float& xo, float& yo) {
  float a1 = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
  float b1 = (x4 - x2) * (y3 - y4) - (y4 - y2) * (x3 - x4);
  if ((double)std::fabs(a1) <= 0.) return -1.0f;
  float t1 = b1 / a1;
  if ((double)t1 <= 0.0 || (double)t1 >= 1.0) return -1.0f;
  xo = t1 * x1 + (1.0f - t1) * x2;
  if ((double)((x3 - xo) * (x4 - xo)) < 0.0) {
    yo = t1 * y1 + (1.0f - t1) * y2;
    return (float)(1.0 - (double)t1);
  }
  return -1.0f;
}

struct Geo2 {
  const float* v;
  const float* vel;
  const uint32_t* e;
  int nv, ne;
};

// intersectAtIndex<2> (:211-227); glm::distance = sqrt(dot(d,d)) (:220)
static inline void intersect_at_index2(const Geo2& g, float p1x, float p1y, float p2x, float p2y, Hit& info) {
  int idx = info.elem;
  uint32_t s = g.e[2 * idx], t = g.e[2 * idx + 1];
  float p3x = g.v[2 * s], p3y = g.v[2 * s + 1];
  float p4x = g.v[2 * t], p4y = g.v[2 * t + 1];
  float d = line_intersect(p1x, p1y, p2x, p2y, p3x, p3y, p4x, p4y, info.hit[0], info.hit[1]);
  if (d > 0 && d < 1) {
    info.elem = idx;
    info.depth = d;
    float dx1 = p4x - info.hit[0], dy1 = p4y - info.hit[1];
    float dx2 = p4x - p3x, dy2 = p4y - p3y;
    float l1 = std::sqrt(dx1 * dx1 + dy1 * dy1);
    float l2 = std::sqrt(dx2 * dx2 + dy2 * dy2);
    info.bary[0] = l1 / l2;
    info.bary[1] = 1.0f - info.bary[0];
  } else {
    info.elem = -1;
  }
}

// getBoundingBox<2> (:198-209): max starts at FLT_MIN (smallest positive normal).
static inline void bbox2(const Geo2& g, int idx, float mn[2], float mx[2]) {
  mn[0] = mn[1] = FLT_MAX;
  mx[0] = mx[1] = FLT_MIN;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      float c = g.v[2 * g.e[2 * idx + j] + i];
      mn[i] = std::min(mn[i], c);
      mx[i] = std::max(mx[i], c);
    }
}

// isInside<2> (:126-144): even-odd ray cast.
static inline bool is_inside2(const Geo2& g, float px, float py) {
  bool in = false;
  for (int i = 0; i < g.ne; i++) {
    uint32_t a = g.e[2 * i], b = g.e[2 * i + 1];
    float sx = g.v[2 * a], sy = g.v[2 * a + 1];
    float tx = g.v[2 * b], ty = g.v[2 * b + 1];
    if ((sy < py && ty >= py) || (sy >= py && ty < py)) {
      in ^= (sx + (py - sy) * (tx - sx) / (ty - sy) > px);
    }
  }
  return in;
}

struct Geo3 {
  const float* v;
  const float* vel;
  const uint32_t* e;
  int nv, ne;
};

static inline void sub3(const float a[3], const float b[3], float o[3]) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const float a[3], const float b[3], float o[3]) {
  o[0] = a[1] * b[2] - b[1] * a[2];
  o[1] = a[2] * b[0] - b[2] * a[0];
  o[2] = a[0] * b[1] - b[0] * a[1];
}
// glm::normalize(v) = v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt.
static inline void normalize3(const float a[3], float o[3]) {
  float s = 1.0f / std::sqrt(dot3(a, a));
  o[0] = a[0] * s; o[1] = a[1] * s; o[2] = a[2] * s;
}

// barycentric_coordinates (:291-304)
static inline void barycentric(const float p[3], const float v0[3], const float v1[3], const float v2[3], float& u, float& v) {
  float e1[3], e2[3], w[3];
  sub3(v1, v0, e1);
  sub3(v2, v0, e2);
  sub3(p, v0, w);
  float d00 = dot3(e1, e1), d01 = dot3(e1, e2), d11 = dot3(e2, e2), d20 = dot3(w, e1), d21 = dot3(w, e2);
  float den = d00 * d11 - d01 * d01;
  u = (d11 * d20 - d01 * d21) / den;
  v = (d00 * d21 - d01 * d20) / den;
}

// line_triangle_intersect (:306-341) with the 1e-6 tolerances of NumericUtils.cuh.
static inline bool line_triangle_intersect(const float p1[3], const float p2[3], const float tv[3][3], Hit& info) {
  const float eps = 1e-6f;
  float e1[3], e2[3], c[3], n[3], dd[3], d[3];
  sub3(tv[1], tv[0], e1);
  sub3(tv[2], tv[0], e2);
  cross3(e1, e2, c);
  normalize3(c, n);
  sub3(p2, p1, dd);
  normalize3(dd, d);
  if (std::fabs(dot3(n, d)) <= eps) {
    info.elem = -1;
    return false;
  }
  float a[3];
  sub3(tv[0], p1, a);
  float t = dot3(n, a) / dot3(n, dd);
  if ((double)t < 0. || (double)t > 1.) {
    info.elem = -1;
    return false;
  }
  float ip[3] = {p1[0] + t * dd[0], p1[1] + t * dd[1], p1[2] + t * dd[2]};
  float u0, v0;
  barycentric(ip, tv[0], tv[1], tv[2], u0, v0);
  if (u0 >= 0.0f - eps && v0 >= 0.0f - eps && u0 + v0 <= 1.0f + eps) {
    info.hit[0] = ip[0]; info.hit[1] = ip[1]; info.hit[2] = ip[2];
    info.depth = t;
    info.bary[0] = 1.0f - u0 - v0;
    info.bary[1] = u0;
    info.bary[2] = v0;
    return true;
  }
  info.elem = -1;
  return false;
}

static inline void intersect_at_index3(const Geo3& g, const float p1[3], const float p2[3], Hit& info) {
  int idx = info.elem;
  if (idx < 0 || idx >= g.ne) { info.elem = -1; return; }
  float tv[3][3];
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) tv[j][k] = g.v[g.e[3 * idx + j] * 3 + k];
  line_triangle_intersect(p1, p2, tv, info);
}

static inline void bbox3(const Geo3& g, int idx, float mn[3], float mx[3]) {
  mn[0] = mn[1] = mn[2] = FLT_MAX;
  mx[0] = mx[1] = mx[2] = FLT_MIN;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float c = g.v[g.e[3 * idx + i] * 3 + j];
      mn[j] = std::min(mn[j], c);
      mx[j] = std::max(mx[j], c);
    }
}

// ---------------------------------------------------------------------------
// Occupancy / hit tables (FsiCoupling2D.cu:31-80, FsiCoupling3D.cu:31-81):
// dense N*Q table of min(depthIndexCompact) per link; 0xFFFFFFFF = no hit.
static void hit_table2d(const Geo2& g, int nx, int ny, float dx, std::vector<uint32_t>& hit) {
  hit.assign((size_t)nx * ny * 9, 0xFFFFFFFFu);
  for (int x = 0; x < g.ne; x++) {
    float mn[2], mx[2];
    bbox2(g, x, mn, mx);
    mn[0] /= dx; mn[1] /= dx; mx[0] /= dx; mx[1] /= dx;
    for (int i = int(mn[0]) - 1; i <= int(mx[0]) + 1; i++) {
      if (i < 0 || i >= nx) continue;
      for (int j = int(mn[1]) - 1; j <= int(mx[1]) + 1; j++) {
        if (j < 0 || j >= ny) continue;
        size_t ind = (size_t)j * nx + i;
        for (int l = 1; l <= 8; l++) {
          Hit info;
          float p1x = (float)i * dx, p1y = (float)j * dx;
          float p2x = ((float)i - ex2[l]) * dx, p2y = ((float)j - ey2[l]) * dx;
          info.elem = x;
          info.depth = 3.0f;
          intersect_at_index2(g, p1x, p1y, p2x, p2y, info);
          if (info.elem != -1) {
            uint32_t key = depth_index_compact(info.depth, x);
            uint32_t& slot = hit[ind * 9 + l];
            if (key < slot) slot = key;
          }
        }
      }
    }
  }
}

static void hit_table3d(const Geo3& g, int nx, int ny, int nz, float dx, std::vector<uint32_t>& hit) {
  hit.assign((size_t)nx * ny * nz * 27, 0xFFFFFFFFu);
  for (int x = 0; x < g.ne; x++) {
    float mn[3], mx[3];
    bbox3(g, x, mn, mx);
    for (int c = 0; c < 3; c++) { mn[c] /= dx; mx[c] /= dx; }
    for (int i = int(mn[0]) - 1; i <= int(mx[0]) + 1; i++) {
      if (i < 0 || i >= nx) continue;
      for (int j = int(mn[1]) - 1; j <= int(mx[1]) + 1; j++) {
        if (j < 0 || j >= ny) continue;
        for (int k = int(mn[2]) - 1; k <= int(mx[2]) + 1; k++) {
          if (k < 0 || k >= nz) continue;
          size_t ind = ((size_t)k * ny + j) * nx + i;
          for (int l = 1; l <= 26; l++) {
            Hit info;
            float p1[3] = {(float)i * dx, (float)j * dx, (float)k * dx};
            float p2[3] = {((float)i - ex3[l]) * dx, ((float)j - ey3[l]) * dx, ((float)k - ez3[l]) * dx};
            info.elem = x;
            intersect_at_index3(g, p1, p2, info);
            if (info.elem != -1) {
              uint32_t key = depth_index_compact(info.depth, x);
              uint32_t& slot = hit[ind * 27 + l];
              if (key < slot) slot = key;
            }
          }
        }
      }
    }
  }
}

static inline bool is_wall2(uint8_t f) {  // FsiCoupling2D.cu:171-176
  return f == Solid || f == WallLeft || f == WallRight || f == WallUp || f == WallDown || f == Wall;
}
static inline bool is_wall3(uint8_t f) {  // FsiCoupling3D.cu:209-216
  return f == Solid || f == Wall || f == WallRight || f == WallUp || f == WallDown || f == WallFront || f == WallBack ||
         f == WallLeft;
}

// Storage quantisation matching the product's 16-bit storage modes (applied
// when a moment is stored: by set_moments, the step and the outlet update):
// slot 0 holds rho-1, slots 1..D hold u, the rest hold Pi; each is scaled by a
// power of two before rounding to the 16-bit format.
static const double kScaleRho = 256.0, kScaleU = 16.0, kScalePi = 256.0;
static inline double quantize(double v, int slot, int dim, int storage) {
  if (storage == 0) return v;
  double s, off = 0.0;
  if (slot == 0) { s = kScaleRho; off = 1.0; }
  else if (slot <= dim) s = kScaleU;
  else s = kScalePi;
  float enc = (float)((v - off) * s);
  float dec;
  if (storage == 3 && slot >= 1 && slot <= dim) {
    // split velocity: hi + lo binary16 pair, e - hi and hi + lo exact in FP32
    const float hi = h2f(f2h(enc));
    dec = hi + h2f(f2h(enc - hi));
  } else {
    dec = storage == 2 ? bf2f(f2bf(enc)) : h2f(f2h(enc));
  }
  return (double)dec / s + off;
}

// ---------------------------------------------------------------------------
struct Domain {
  int dim, nx, ny, nz;
  int fp64, stage_fp16, periodic, storage;
  double nu = 0.001, cl = 1.0, ct = 0.001, cf = 1.0;
  size_t n;
  int nm;
  std::vector<uint8_t> flags;
  std::vector<double> buf[2];  // AoS moments, values exactly representable in R
  int cur = 0;
  // boundary
  std::vector<float> verts, vels;
  std::vector<uint32_t> elems;
  int nv = 0, ne = 0;
  std::vector<double> vforce;  // per-vertex accumulated force (dim comps)
  std::vector<uint32_t> hit;
  long long steps_done = 0;
};

// Per-link force record gathered during a step, reduced sequentially after.
struct LinkF {
  int elem;
  float bary[3];
  double F[3];
};

template <class R>
static void step2d(Domain& d) {
  const int nx = d.nx, ny = d.ny;
  const size_t N = d.n;
  const std::vector<double>& cur = d.buf[d.cur];
  std::vector<double>& nxt = d.buf[d.cur ^ 1];
  Geo2 g{d.verts.data(), d.vels.data(), d.elems.data(), d.nv, d.ne};
  const float vis_f = (float)d.nu;
  const float cl = (float)d.cl, ct = (float)d.ct, cf = (float)d.cf;
  // hit table (FsiCoupling2D.cu:376-386)
  if (d.ne > 0) hit_table2d(g, nx, ny, cl, d.hit);
  // Omega (FsiCoupling2D.cu:236): computed in double from the float viscosity.
  const R Omega = std::is_same<R, float>::value ? (R)(float)(1.0 / ((double)vis_f * 3.0 + 0.5))
                                                  : (R)(1.0 / (d.nu * 3.0 + 0.5));
  // staged populations of the current buffer for every cell
  std::vector<R> pops(N * 9);
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < (long long)N; c++) {
    const double* m = &cur[(size_t)c * 6];
    R f[9];
    recon2d_all<R>((R)m[0], (R)m[1], (R)m[2], (R)m[3], (R)m[5], (R)m[4], f);
    for (int i = 0; i < 9; i++) {
      R v = f[i];
      if (d.stage_fp16) v = (R)h2f(f2h((float)v));
      pops[(size_t)c * 9 + i] = v;
    }
  }
  std::vector<std::vector<LinkF>> links_per_row(ny);
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < ny; y++) {
    for (int x = 0; x < nx; x++) {
      size_t curind = (size_t)y * nx + x;
      uint8_t fl = d.flags[curind];
      if (!(fl == Fluid || fl == SolidDynamic)) continue;
      const double* m = &cur[curind * 6];
      R ux_cur = (R)m[1], uy_cur = (R)m[2], pixx_cur = (R)m[3], piyy_cur = (R)m[4], pixy_cur = (R)m[5];
      R pop[9];
      for (int i = 0; i < 9; i++) {
        int dx = (int)ex2[i], dy = (int)ey2[i];
        int x1 = x - dx, y1 = y - dy;
        bool notCross = x1 >= 0 && x1 < nx && y1 >= 0 && y1 < ny;
        int xw = (x1 + nx) % nx, yw = (y1 + ny) % ny;
        size_t ind_back = (size_t)yw * nx + xw;
        bool is_blocked = false;
        if (is_wall2(d.flags[ind_back])) {
          pop[i] = std::max(R(0), pops[curind * 9 + inv2[i]]);
        } else if (notCross && d.ne > 0 && d.hit[9 * curind + i] != 0xFFFFFFFFu) {
          Hit info;
          info.elem = (int)(d.hit[9 * curind + i] & 0x0000FFFFu);
          intersect_at_index2(g, (float)x * cl, (float)y * cl, (float)xw * cl, (float)yw * cl, info);
          is_blocked = info.elem != -1;
          if (is_blocked) {
            R rhoVar = (R)m[0];
            uint32_t v1 = g.e[2 * info.elem], v2 = g.e[2 * info.elem + 1];
            float vx = g.vel[2 * v1] * info.bary[0] + g.vel[2 * v2] * info.bary[1];
            float vy = g.vel[2 * v1 + 1] * info.bary[0] + g.vel[2 * v2 + 1] * info.bary[1];
            float s = ct / cl;
            R ux = (R)(vx * s), uy = (R)(vy * s);
            if (!std::is_same<R, float>::value) {
              ux = (R)vx * ((R)d.ct / (R)d.cl);
              uy = (R)vy * ((R)d.ct / (R)d.cl);
            }
            R pixx = pixx_cur + ux * ux - ux_cur * ux_cur;
            R piyy = piyy_cur + uy * uy - uy_cur * uy_cur;
            R pixy = pixy_cur + uy * ux - ux_cur * uy_cur;
            R fin = std::max(R(0), pops[curind * 9 + inv2[i]]);
            R fout = recon2d_at<R>(rhoVar, ux, uy, pixx, pixy, piyy, i);
            pop[i] = fout;
            R cfR = std::is_same<R, float>::value ? (R)cf : (R)d.cf;
            R fx = (fin * (R(-dx) - ux) - fout * (R(dx) - ux)) * cfR;
            R fy = (fin * (R(-dy) - uy) - fout * (R(dy) - uy)) * cfR;
            LinkF lf;
            lf.elem = info.elem;
            lf.bary[0] = info.bary[0];
            lf.bary[1] = info.bary[1];
            lf.bary[2] = 0.f;
            lf.F[0] = (double)fx;
            lf.F[1] = (double)fy;
            lf.F[2] = 0.0;
            links_per_row[y].push_back(lf);
          }
        }
        // FsiCoupling2D.cu:218-223: `is_blocked` is never set on the wall branch,
        // so a wall-upstream population is overwritten by the pull below; in 2D
        // the wall flag only suppresses the moving-wall test for that link.
        if (!is_blocked) {
          int sx, sy;
          if (d.periodic) { sx = xw; sy = yw; }
          else { sx = std::min(nx - 1, std::max(0, x1)); sy = std::min(ny - 1, std::max(0, y1)); }
          pop[i] = std::max(R(0), pops[((size_t)sy * nx + sx) * 9 + i]);
        }
      }
      R rhoVar = pop[0] + pop[1] + pop[2] + pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
      R FX = 0, FY = 0;
      R invRho = R(1) / rhoVar;
      R ux = (R)(((double)(pop[1] - pop[3] + pop[5] - pop[6] - pop[7] + pop[8]) + 0.5 * (double)FX) * (double)invRho);
      R uy = (R)(((double)(pop[2] - pop[4] + pop[5] + pop[6] - pop[7] - pop[8]) + 0.5 * (double)FY) * (double)invRho);
      R pixx = pop[1] + pop[3] + pop[5] + pop[6] + pop[7] + pop[8];
      R piyy = pop[2] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
      R pixy = pop[5] - pop[6] + pop[7] - pop[8];
      collide2d<R>(rhoVar, ux, uy, FX, FY, Omega, pixx, piyy, pixy);
      R cs2 = W2<R>::cs2();
      if (std::is_same<R, float>::value) cs2 = (R)(1.0f / 3.0f);
      pixx = (R)((double)(pixx * invRho) - 1.0 * (double)cs2);
      piyy = (R)((double)(piyy * invRho) - 1.0 * (double)cs2);
      pixy = pixy * invRho;
      double* o = &nxt[curind * 6];
      const double ov[6] = {(double)rhoVar, (double)(ux + FX * invRho / R(2)), (double)(uy + FY * invRho / R(2)),
                            (double)pixx, (double)piyy, (double)pixy};
      for (int k = 0; k < 6; k++) o[k] = quantize(ov[k], k, 2, d.storage);
    }
  }
  // force accumulation, sequential (row, x, direction) order
  for (int y = 0; y < ny; y++)
    for (const LinkF& lf : links_per_row[y]) {
      uint32_t v1 = g.e[2 * lf.elem], v2 = g.e[2 * lf.elem + 1];
      for (int c = 0; c < 2; c++) {
        if (d.fp64) {
          d.vforce[2 * v1 + c] += lf.F[c] * (double)lf.bary[0];
          d.vforce[2 * v2 + c] += lf.F[c] * (double)lf.bary[1];
        } else {
          d.vforce[2 * v1 + c] += (double)((float)lf.F[c] * lf.bary[0]);
          d.vforce[2 * v2 + c] += (double)((float)lf.F[c] * lf.bary[1]);
        }
      }
    }
  // swap + outlet on the new buffer (FsiCoupling2D.cu:419-427, 257-316)
  d.cur ^= 1;
  std::vector<double>& mm = d.buf[d.cur];
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      size_t c = (size_t)y * nx + x;
      uint8_t fl = d.flags[c];
      long long nb = -1;
      int comp = -1, pcomp = -1, sign = 0;
      if (fl == OutletUp && y - 1 >= 0) { nb = (long long)(y - 1) * nx + x; comp = 2; pcomp = 4; sign = +1; }
      else if (fl == OutletDown && y + 1 < ny) { nb = (long long)(y + 1) * nx + x; comp = 2; pcomp = 4; sign = -1; }
      else if (fl == OutletLeft && x + 1 < nx) { nb = (long long)y * nx + x + 1; comp = 1; pcomp = 3; sign = -1; }
      else if (fl == OutletRight && x - 1 >= 0) { nb = (long long)y * nx + x - 1; comp = 1; pcomp = 3; sign = +1; }
      if (nb < 0) continue;
      if (d.flags[nb] != Fluid) continue;
      R up = (R)mm[(size_t)nb * 6 + comp];
      if ((sign > 0 && up > 0) || (sign < 0 && up < 0)) mm[c * 6 + comp] = (double)up;
      mm[c * 6 + pcomp] = quantize((double)(up * up), pcomp, 2, d.storage);
    }
  d.steps_done++;
}

template <class R>
static void step3d(Domain& d) {
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const size_t N = d.n;
  const std::vector<double>& cur = d.buf[d.cur];
  std::vector<double>& nxt = d.buf[d.cur ^ 1];
  Geo3 g{d.verts.data(), d.vels.data(), d.elems.data(), d.nv, d.ne};
  const float vis_f = (float)d.nu;
  const float cl = (float)d.cl, ct = (float)d.ct, cf = (float)d.cf;
  if (d.ne > 0) hit_table3d(g, nx, ny, nz, cl, d.hit);
  const R Omega = std::is_same<R, float>::value ? (R)(float)(1.0 / ((double)vis_f * 3.0 + 0.5))
                                                  : (R)(1.0 / (d.nu * 3.0 + 0.5));
  std::vector<R> pops(N * 27);
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < (long long)N; c++) {
    const double* m = &cur[(size_t)c * 10];
    recon3d_all<R>((R)m[0], (R)m[1], (R)m[2], (R)m[3], (R)m[4], (R)m[5], (R)m[6], (R)m[7], (R)m[8], (R)m[9],
                   &pops[(size_t)c * 27]);
  }
  std::vector<std::vector<LinkF>> links_per_plane(nz);
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t curind = ((size_t)z * ny + y) * nx + x;
        uint8_t fl = d.flags[curind];
        if (!(fl == Fluid || fl == SolidDynamic)) continue;
        const double* m = &cur[curind * 10];
        R rho_cur = (R)m[0], ux_cur = (R)m[1], uy_cur = (R)m[2], uz_cur = (R)m[3];
        R pixx_cur = (R)m[4], piyy_cur = (R)m[5], pizz_cur = (R)m[6], pixy_cur = (R)m[7], piyz_cur = (R)m[8],
          pixz_cur = (R)m[9];
        R pop[27];
        for (int i = 0; i < 27; i++) {
          int dx = (int)ex3[i], dy = (int)ey3[i], dz = (int)ez3[i];
          int x1 = x - dx, y1 = y - dy, z1 = z - dz;
          bool notCross = x1 >= 0 && x1 < nx && y1 >= 0 && y1 < ny && z1 >= 0 && z1 < nz;
          int xw = (x1 + nx) % nx, yw = (y1 + ny) % ny, zw = (z1 + nz) % nz;
          size_t ind_back = ((size_t)zw * ny + yw) * nx + xw;
          if (is_wall3(d.flags[ind_back])) {
            pop[i] = std::max(R(0), pops[curind * 27 + inv3[i]]);
          } else {
            bool intersect = notCross && d.ne > 0 && d.hit[27 * curind + i] != 0xFFFFFFFFu;
            if (intersect) {
              float p1[3] = {(float)x * cl, (float)y * cl, (float)z * cl};
              float p2[3] = {(float)xw * cl, (float)yw * cl, (float)zw * cl};
              Hit info;
              info.elem = (int)(d.hit[27 * curind + i] & 0x0000FFFFu);
              intersect_at_index3(g, p1, p2, info);
              if (info.elem != -1) {
                uint32_t v1 = g.e[3 * info.elem], v2 = g.e[3 * info.elem + 1], v3 = g.e[3 * info.elem + 2];
                float u3[3];
                for (int c = 0; c < 3; c++)
                  u3[c] = g.vel[3 * v1 + c] * info.bary[0] + g.vel[3 * v2 + c] * info.bary[1] + g.vel[3 * v3 + c] * info.bary[2];
                R ux, uy, uz;
                if (std::is_same<R, float>::value) {
                  ux = (R)(u3[0] * ct / cl); uy = (R)(u3[1] * ct / cl); uz = (R)(u3[2] * ct / cl);
                } else {
                  ux = (R)u3[0] * (R)d.ct / (R)d.cl; uy = (R)u3[1] * (R)d.ct / (R)d.cl; uz = (R)u3[2] * (R)d.ct / (R)d.cl;
                }
                R pixx = pixx_cur + ux * ux - ux_cur * ux_cur;
                R piyy = piyy_cur + uy * uy - uy_cur * uy_cur;
                R pizz = pizz_cur + uz * uz - uz_cur * uz_cur;
                R pixy = pixy_cur + ux * uy - ux_cur * uy_cur;
                R piyz = piyz_cur + uy * uz - uy_cur * uz_cur;
                R pixz = pixz_cur + ux * uz - ux_cur * uz_cur;
                pop[i] = recon3d_at<R>(rho_cur, ux, uy, uz, pixx, piyy, pizz, pixy, piyz, pixz, i);
                R fin = std::max(R(0), pops[curind * 27 + inv3[i]]);
                R fout = pop[i];
                R cfR = std::is_same<R, float>::value ? (R)cf : (R)d.cf;
                LinkF lf;
                lf.elem = info.elem;
                for (int c = 0; c < 3; c++) lf.bary[c] = info.bary[c];
                lf.F[0] = (double)((fin * (R(-dx) - ux) - fout * (R(dx) - ux)) * cfR);
                lf.F[1] = (double)((fin * (R(-dy) - uy) - fout * (R(dy) - uy)) * cfR);
                lf.F[2] = (double)((fin * (R(-dz) - uz) - fout * (R(dz) - uz)) * cfR);
                links_per_plane[z].push_back(lf);
              } else {
                intersect = false;
              }
            }
            if (!intersect) {
              int sx, sy, sz;
              if (d.periodic) { sx = xw; sy = yw; sz = zw; }
              else {
                sx = std::min(nx - 1, std::max(0, x1));
                sy = std::min(ny - 1, std::max(0, y1));
                sz = std::min(nz - 1, std::max(0, z1));
              }
              pop[i] = std::max(R(0), pops[(((size_t)sz * ny + sy) * nx + sx) * 27 + i]);
            }
          }
        }
        R FX = 0, FY = 0, FZ = 0;  // body-force field is always zero (LbmFlowField3D.cpp:35-36)
        R rhoVar = 0;
        for (int i = 0; i < 27; i++) rhoVar += pop[i];
        R invRho = R(1) / rhoVar;
        const R* p = pop;
        R sxp = p[1] + p[7] + p[9] + p[13] + p[15] + p[19] + p[21] + p[23] + p[26];
        R sxm = p[2] + p[8] + p[10] + p[14] + p[16] + p[20] + p[22] + p[24] + p[25];
        R syp = p[3] + p[7] + p[11] + p[14] + p[17] + p[19] + p[21] + p[24] + p[25];
        R sym = p[4] + p[8] + p[12] + p[13] + p[18] + p[20] + p[22] + p[23] + p[26];
        R szp = p[5] + p[9] + p[11] + p[16] + p[18] + p[19] + p[22] + p[23] + p[25];
        R szm = p[6] + p[10] + p[12] + p[15] + p[17] + p[20] + p[21] + p[24] + p[26];
        R ux = (R)(((double)(sxp - sxm) + 0.5 * (double)FX) * (double)invRho);
        R uy = (R)(((double)(syp - sym) + 0.5 * (double)FY) * (double)invRho);
        R uz = (R)(((double)(szp - szm) + 0.5 * (double)FZ) * (double)invRho);
        R pixx = p[1] + p[2] + p[7] + p[8] + p[9] + p[10] + p[13] + p[14] + p[15] + p[16] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        R piyy = p[3] + p[4] + p[7] + p[8] + p[11] + p[12] + p[13] + p[14] + p[17] + p[18] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        R pizz = p[5] + p[6] + p[9] + p[10] + p[11] + p[12] + p[15] + p[16] + p[17] + p[18] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        R pixy = (p[7] + p[8] + p[19] + p[20] + p[21] + p[22]) - (p[13] + p[14] + p[23] + p[24] + p[25] + p[26]);
        R piyz = (p[11] + p[12] + p[19] + p[20] + p[25] + p[26]) - (p[17] + p[18] + p[21] + p[22] + p[23] + p[24]);
        R pixz = (p[9] + p[10] + p[19] + p[20] + p[23] + p[24]) - (p[15] + p[16] + p[21] + p[22] + p[25] + p[26]);
        collide3d<R>(rhoVar, ux, uy, uz, FX, FY, FZ, Omega, pixx, piyy, pizz, pixy, piyz, pixz);
        R cs2 = std::is_same<R, float>::value ? (R)(1.0f / 3.0f) : W3<R>::cs2();
        double* o = &nxt[curind * 10];
        const double ov[10] = {(double)rhoVar, (double)(ux + FX * invRho / R(2)), (double)(uy + FY * invRho / R(2)),
                               (double)(uz + FZ * invRho / R(2)), (double)(pixx * invRho - cs2),
                               (double)(piyy * invRho - cs2), (double)(pizz * invRho - cs2), (double)(pixy * invRho),
                               (double)(piyz * invRho), (double)(pixz * invRho)};
        for (int k = 0; k < 10; k++) o[k] = quantize(ov[k], k, 3, d.storage);
      }
  for (int z = 0; z < nz; z++)
    for (const LinkF& lf : links_per_plane[z]) {
      for (int k = 0; k < 3; k++) {
        uint32_t v = g.e[3 * lf.elem + k];
        for (int c = 0; c < 3; c++)
          d.vforce[3 * v + c] += d.fp64 ? lf.F[c] * (double)lf.bary[k] : (double)((float)lf.F[c] * lf.bary[k]);
      }
    }
  d.cur ^= 1;
  std::vector<double>& mm = d.buf[d.cur];
  const size_t sxy = (size_t)nx * ny;
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t c = ((size_t)z * ny + y) * nx + x;
        uint8_t fl = d.flags[c];
        long long nb = -1;
        int comp = -1, pcomp = -1, sign = 0;
        if (fl == OutletUp && z - 1 >= 0) { nb = (long long)(c - sxy); comp = 3; pcomp = 6; sign = +1; }
        else if (fl == OutletDown && z + 1 < nz) { nb = (long long)(c + sxy); comp = 3; pcomp = 6; sign = -1; }
        else if (fl == OutletLeft && x + 1 < nx) { nb = (long long)(c + 1); comp = 1; pcomp = 4; sign = -1; }
        else if (fl == OutletRight && x - 1 >= 0) { nb = (long long)(c - 1); comp = 1; pcomp = 4; sign = +1; }
        else if (fl == OutletFront && y + 1 < ny) { nb = (long long)(c + nx); comp = 2; pcomp = 5; sign = -1; }
        else if (fl == OutletBack && y - 1 >= 0) { nb = (long long)(c - nx); comp = 2; pcomp = 5; sign = +1; }
        if (nb < 0) continue;
        if (d.flags[nb] != Fluid) continue;
        R up = (R)mm[(size_t)nb * 10 + comp];
        if ((sign > 0 && up > 0) || (sign < 0 && up < 0)) mm[c * 10 + comp] = (double)up;
        mm[c * 10 + pcomp] = quantize((double)(up * up), pcomp, 3, d.storage);
      }
  d.steps_done++;
}

// Initial state (LbmInitializer2D.cu:32-95, LbmInitializer3D.cu:33-113):
// rho=1, u from the inlet flag, Pi from the moments of f_eq.  Written to both buffers.
template <class R>
static void init_reference(Domain& d, const float* bc, int nbc) {
  float bcv[6] = {0, 0, 0, 0, 0, 0};
  int need = d.dim == 2 ? 4 : 6;
  if (nbc == need)
    for (int k = 0; k < need; k++) bcv[k] = bc[k];
  for (size_t c = 0; c < d.n; c++) {
    uint8_t fl = d.flags[c];
    R ux = 0, uy = 0, uz = 0, rho = 1;
    if (d.dim == 2) {
      if (fl == InletLeft) ux = (R)bcv[0];
      else if (fl == InletDown) uy = (R)bcv[1];
      else if (fl == InletRight) ux = -(R)bcv[2];
      else if (fl == InletUp) uy = -(R)bcv[3];
      R u_sq = ux * ux + uy * uy;
      R feq[9];
      for (int k = 0; k < 9; k++) {
        R cu = (R)ex2[k] * ux + (R)ey2[k] * uy;
        feq[k] = W2<R>::w(k) * rho * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * u_sq);
      }
      R invRho = R(1) / rho;
      R pixx = feq[1] + feq[3] + feq[5] + feq[6] + feq[7] + feq[8];
      R piyy = feq[2] + feq[4] + feq[5] + feq[6] + feq[7] + feq[8];
      R pixy = feq[5] - feq[6] + feq[7] - feq[8];
      R cs2 = std::is_same<R, float>::value ? (R)(1.0f / 3.0f) : W2<R>::cs2();
      pixx = (R)((double)(pixx * invRho) - 1.0 * (double)cs2);
      piyy = (R)((double)(piyy * invRho) - 1.0 * (double)cs2);
      pixy = pixy * invRho;
      double mv[6] = {(double)rho, (double)ux, (double)uy, (double)pixx, (double)piyy, (double)pixy};
      for (int k = 0; k < 6; k++) d.buf[0][c * 6 + k] = d.buf[1][c * 6 + k] = mv[k];
    } else {
      if (fl == InletLeft) ux = (R)bcv[0];
      else if (fl == InletDown) uz = (R)bcv[1];
      else if (fl == InletRight) ux = -(R)bcv[2];
      else if (fl == InletUp) uz = -(R)bcv[3];
      else if (fl == InletFront) uy = (R)bcv[4];
      else if (fl == InletBack) uy = -(R)bcv[5];
      R u_sq = ux * ux + uy * uy + uz * uz;
      R feq[27];
      for (int k = 0; k < 27; k++) {
        R cu = (R)ex3[k] * ux + (R)ey3[k] * uy + (R)ez3[k] * uz;
        feq[k] = W3<R>::w(k) * rho * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * u_sq);
      }
      const R* p = feq;
      R pixx = p[1] + p[2] + p[7] + p[8] + p[9] + p[10] + p[13] + p[14] + p[15] + p[16] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
      R piyy = p[3] + p[4] + p[7] + p[8] + p[11] + p[12] + p[13] + p[14] + p[17] + p[18] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
      R pizz = p[5] + p[6] + p[9] + p[10] + p[11] + p[12] + p[15] + p[16] + p[17] + p[18] + p[19] + p[20] + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
      R pixy = (p[7] + p[8] + p[19] + p[20] + p[21] + p[22]) - (p[13] + p[14] + p[23] + p[24] + p[25] + p[26]);
      R piyz = (p[11] + p[12] + p[19] + p[20] + p[25] + p[26]) - (p[17] + p[18] + p[21] + p[22] + p[23] + p[24]);
      R pixz = (p[9] + p[10] + p[19] + p[20] + p[23] + p[24]) - (p[15] + p[16] + p[21] + p[22] + p[25] + p[26]);
      R invRho = R(1) / rho;
      R cs2 = std::is_same<R, float>::value ? (R)(1.0f / 3.0f) : W3<R>::cs2();
      double mv[10] = {(double)rho, (double)ux, (double)uy, (double)uz,
                       (double)(R)((double)(pixx * invRho) - 1.0 * (double)cs2),
                       (double)(R)((double)(piyy * invRho) - 1.0 * (double)cs2),
                       (double)(R)((double)(pizz * invRho) - 1.0 * (double)cs2),
                       (double)(pixy * invRho), (double)(piyz * invRho), (double)(pixz * invRho)};
      for (int k = 0; k < 10; k++) d.buf[0][c * 10 + k] = d.buf[1][c * 10 + k] = mv[k];
    }
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

void* orc_create(int dim, int nx, int ny, int nz, int fp64, int stage_fp16, int periodic) {
  if (dim != 2 && dim != 3) return nullptr;
  if (nx <= 0 || ny <= 0 || (dim == 3 && nz <= 0)) return nullptr;
  Domain* d = new Domain();
  d->dim = dim;
  d->nx = nx;
  d->ny = ny;
  d->nz = dim == 3 ? nz : 1;
  d->fp64 = fp64;
  d->stage_fp16 = dim == 2 ? stage_fp16 : 0;
  d->periodic = periodic;
  d->storage = 0;
  d->n = (size_t)nx * ny * d->nz;
  d->nm = dim == 2 ? 6 : 10;
  d->flags.assign(d->n, Fluid);
  d->buf[0].assign(d->n * d->nm, 0.0);
  d->buf[1].assign(d->n * d->nm, 0.0);
  return d;
}

void orc_destroy(void* h) { delete (Domain*)h; }

void orc_set_params(void* h, double nu, double cl, double ct, double cf) {
  Domain* d = (Domain*)h;
  d->nu = nu;
  d->cl = cl;
  d->ct = ct;
  d->cf = cf;
}

void orc_set_storage(void* h, int storage) { ((Domain*)h)->storage = storage; }

void orc_set_flags(void* h, const uint8_t* flags) {
  Domain* d = (Domain*)h;
  std::memcpy(d->flags.data(), flags, d->n);
}

// Moments in AoS order (m[nm*cell + k]).  which: 0 current, 1 next, -1 both.
// Values are rounded to float for the FP32 instantiation.
void orc_set_moments(void* h, const double* aos, int which) {
  Domain* d = (Domain*)h;
  for (int b = 0; b < 2; b++) {
    int target = b == 0 ? d->cur : (d->cur ^ 1);
    if (which >= 0 && which != b) continue;
    for (size_t i = 0; i < d->n * d->nm; i++) {
      double v = d->fp64 ? aos[i] : (double)(float)aos[i];
      d->buf[target][i] = quantize(v, (int)(i % d->nm), d->dim, d->storage);
    }
  }
}

void orc_get_moments(void* h, double* aos, int which) {
  Domain* d = (Domain*)h;
  int target = which == 1 ? (d->cur ^ 1) : d->cur;
  std::memcpy(aos, d->buf[target].data(), d->n * d->nm * sizeof(double));
}

void orc_init_reference(void* h, const float* bc, int nbc) {
  Domain* d = (Domain*)h;
  d->cur = 0;
  if (d->fp64) init_reference<double>(*d, bc, nbc);
  else init_reference<float>(*d, bc, nbc);
}

void orc_set_boundary(void* h, int nv, const float* verts, const float* vels, int ne, const uint32_t* elems) {
  Domain* d = (Domain*)h;
  int D = d->dim;
  d->nv = nv;
  d->ne = ne;
  d->verts.assign(verts, verts + (size_t)nv * D);
  d->vels.assign(vels, vels + (size_t)nv * D);
  d->elems.assign(elems, elems + (size_t)ne * D);
  if ((int)d->vforce.size() != nv * D) d->vforce.assign((size_t)nv * D, 0.0);
}

// Replace vertex positions/velocities, keeping topology (a moving body).
void orc_update_boundary(void* h, const float* verts, const float* vels) {
  Domain* d = (Domain*)h;
  std::memcpy(d->verts.data(), verts, d->verts.size() * sizeof(float));
  std::memcpy(d->vels.data(), vels, d->vels.size() * sizeof(float));
}

void orc_get_vertex_forces(void* h, double* out) {
  Domain* d = (Domain*)h;
  std::memcpy(out, d->vforce.data(), d->vforce.size() * sizeof(double));
}
void orc_zero_vertex_forces(void* h) {
  Domain* d = (Domain*)h;
  std::fill(d->vforce.begin(), d->vforce.end(), 0.0);
}

void orc_step(void* h, int nsteps) {
  Domain* d = (Domain*)h;
  for (int s = 0; s < nsteps; s++) {
    if (d->dim == 2) {
      if (d->fp64) step2d<double>(*d);
      else step2d<float>(*d);
    } else {
      if (d->fp64) step3d<double>(*d);
      else step3d<float>(*d);
    }
  }
}

void orc_get_hit_table(void* h, uint32_t* out) {
  Domain* d = (Domain*)h;
  Geo2 g2{d->verts.data(), d->vels.data(), d->elems.data(), d->nv, d->ne};
  Geo3 g3{d->verts.data(), d->vels.data(), d->elems.data(), d->nv, d->ne};
  std::vector<uint32_t> t;
  if (d->dim == 2) hit_table2d(g2, d->nx, d->ny, (float)d->cl, t);
  else hit_table3d(g3, d->nx, d->ny, d->nz, (float)d->cl, t);
  std::memcpy(out, t.data(), t.size() * sizeof(uint32_t));
}

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// ---- unit-level entry points (stored moment order) ------------------------
void orc_recon2d_all_f(const float* m, float* f) { recon2d_all<float>(m[0], m[1], m[2], m[3], m[5], m[4], f); }
void orc_recon2d_all_d(const double* m, double* f) { recon2d_all<double>(m[0], m[1], m[2], m[3], m[5], m[4], f); }
float orc_recon2d_at_f(const float* m, int i) { return recon2d_at<float>(m[0], m[1], m[2], m[3], m[5], m[4], i); }
double orc_recon2d_at_d(const double* m, int i) { return recon2d_at<double>(m[0], m[1], m[2], m[3], m[5], m[4], i); }
void orc_collide2d_f(float rho, float ux, float uy, float omega, float* pi) {  // pi = raw [xx, yy, xy]
  collide2d<float>(rho, ux, uy, 0.f, 0.f, omega, pi[0], pi[1], pi[2]);
}
void orc_collide2d_d(double rho, double ux, double uy, double omega, double* pi) {
  collide2d<double>(rho, ux, uy, 0.0, 0.0, omega, pi[0], pi[1], pi[2]);
}
void orc_recon3d_all_f(const float* m, float* f) {
  recon3d_all<float>(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], f);
}
void orc_recon3d_all_d(const double* m, double* f) {
  recon3d_all<double>(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], f);
}
float orc_recon3d_at_f(const float* m, int i) {
  return recon3d_at<float>(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], i);
}
double orc_recon3d_at_d(const double* m, int i) {
  return recon3d_at<double>(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], i);
}
void orc_collide3d_f(float rho, float ux, float uy, float uz, float omega, float* pi) {  // raw [xx,yy,zz,xy,yz,xz]
  collide3d<float>(rho, ux, uy, uz, 0.f, 0.f, 0.f, omega, pi[0], pi[1], pi[2], pi[3], pi[4], pi[5]);
}
void orc_collide3d_d(double rho, double ux, double uy, double uz, double omega, double* pi) {
  collide3d<double>(rho, ux, uy, uz, 0.0, 0.0, 0.0, omega, pi[0], pi[1], pi[2], pi[3], pi[4], pi[5]);
}
float orc_line_intersect(const float* p, float* hit) {  // p = x1,y1,x2,y2,x3,y3,x4,y4
  return line_intersect(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], hit[0], hit[1]);
}
// Returns elem (-1 = miss); out = depth, bary0, bary1, bary2, hit xyz
int orc_intersect3(const float* p1, const float* p2, const float* tri, float* out) {
  float tv[3][3];
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) tv[j][k] = tri[3 * j + k];
  Hit info;
  info.elem = 0;
  bool ok = line_triangle_intersect(p1, p2, tv, info);
  if (!ok) return -1;
  out[0] = info.depth;
  out[1] = info.bary[0];
  out[2] = info.bary[1];
  out[3] = info.bary[2];
  out[4] = info.hit[0];
  out[5] = info.hit[1];
  out[6] = info.hit[2];
  return 0;
}
uint32_t orc_depth_index_compact(float depth, int index) { return depth_index_compact(depth, index); }
uint16_t orc_f2h(float f) { return f2h(f); }
float orc_h2f(uint16_t h) { return h2f(h); }
uint16_t orc_f2bf(float f) { return f2bf(f); }
float orc_bf2f(uint16_t b) { return bf2f(b); }
double orc_quantize(double v, int slot, int dim, int storage) { return quantize(v, slot, dim, storage); }

// fillSolidFlags2D restatement (FsiCoupling2D.cu:318-354): Fluid/SolidDynamic
// cells become Fluid, then Fluid cells whose centre is inside the polygon
// become SolidDynamic.
void orc_fill_solid2d(int nx, int ny, float dxl, int nv, const float* verts, int ne, const uint32_t* elems,
                      uint8_t* flags) {
  Geo2 g{verts, verts, elems, nv, ne};
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      size_t c = (size_t)y * nx + x;
      if (flags[c] == Fluid || flags[c] == SolidDynamic) flags[c] = Fluid;
    }
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++) {
      size_t c = (size_t)y * nx + x;
      if (flags[c] == Fluid) {
        float px = ((float)x + 0.5f) * dxl, py = ((float)y + 0.5f) * dxl;
        if (is_inside2(g, px, py)) flags[c] = SolidDynamic;
      }
    }
}

}  // extern "C"
