// ff_internal.h -- domain state and helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/freeflow_lbm.h"

namespace ff {

// Node flags: ordinals of fsi::lbm::LbmNodeFlag (LbmDataTypes.cuh:17-56).
enum Flag : uint8_t {
  kInvalid = 0, kFluid, kFluidRest, kSolid, kSolidDynamic, kWall,
  kWallLeft, kWallRight, kWallFront, kWallBack, kWallDown, kWallUp, kWallCorner,
  kInletLeft, kInletRight, kInletUp, kInletDown, kInletFront, kInletBack,
  kOutletLeft, kOutletRight, kOutletUp, kOutletDown, kOutletFront, kOutletBack,
  kReceptor, kDeactivated, kSmoke
};

// Thread-local error string behind ff_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define FF_HIP(call)                                  \
  do {                                                \
    hipError_t _e = (call);                           \
    if (_e != hipSuccess) return ::ff::hip_fail(_e, #call); \
  } while (0)

// One outlet cell: in-place update of the new buffer after a step.
struct OutletCell {
  int cell;   // outlet cell index
  int nb;     // neighbour index (must be Fluid to act)
  int8_t comp;   // velocity slot
  int8_t pcomp;  // Pi slot
  int8_t sign;   // +1: copy when u > 0, -1: copy when u < 0
  int8_t pad;
};

// Library-held state of one fluid domain (the role of LbmFlowField2D/3D).
struct Domain {
  int device = 0;
  hipStream_t stream = nullptr;
  int dim = 2, nx = 1, ny = 1, nz = 1;
  long long n = 0;   // cells
  int nm = 6;        // moments per cell
  int q = 9;         // lattice directions
  int storage = FF_STORAGE_F32;
  unsigned options = 0;
  double nu = 0, dx = 1, dt = 1, rho_f = 1;
  float nu_f = 0, cl = 1, ct = 1, cf = 1, omega = 1;
  size_t plane = 0;  // elements between moment planes
  void* mom[2] = {nullptr, nullptr};
  int cur = 0;
  uint8_t* flags = nullptr;
  std::vector<uint8_t> h_flags;
  // outlets
  OutletCell* outlets = nullptr;
  int n_outlets = 0;
  // boundary geometry (device, interleaved like the reference proxy)
  int nv = 0, ne = 0, n_bodies = 0;
  float* verts = nullptr;
  float* vels = nullptr;
  uint32_t* elems = nullptr;
  int* elem_body = nullptr;
  float* vforce = nullptr;       // accumulated per-vertex force
  // CSR: body -> elements, vertex -> (element, end)
  int* body_elem_ptr = nullptr;
  int* body_elem_idx = nullptr;
  int* vert_inc_ptr = nullptr;
  int* vert_inc_idx = nullptr;   // element*4 + end (end < dim)
  double* body_ref = nullptr;    // n_bodies * dim
  double* elem_acc = nullptr;    // per element: dim (force) + dim*dim? see links.hip
  double* wrench = nullptr;      // n_bodies * (2D:3, 3D:6), latest step
  double* wrench_hist = nullptr; // ring
  int hist_cap = 0;
  long long hist_count = 0;      // steps recorded in the ring since last set
  // link hash table (sparse replacement of the dense N*Q hit table)
  int ht_cap = 0;                // power of two
  int* ht_cell = nullptr;        // cap, -1 = empty
  uint32_t* ht_key = nullptr;    // cap * (q-1), 0xFFFFFFFF = no hit
  float* ht_f = nullptr;         // cap * (q-1) * dim  link force
  float* ht_b = nullptr;         // cap * (q-1) * dim  barycentrics (b[0] < 0: link not applied)
  float* ht_hit = nullptr;       // cap * (q-1) * dim  hit point
  int* ht_list = nullptr;        // cap occupied slots
  int* counters = nullptr;       // [0] occupied count, [1] overflow flag
  // wave kinematics
  bool wave = false;
  float* rest_verts = nullptr;
  float* s_coord = nullptr;
  double* wave_params = nullptr;  // 4 per body
  int* vert_body = nullptr;
  long long step_index = 0;       // steps taken since creation (wave time)
  // scratch
  double* red = nullptr;          // reduction scratch
  bool has_boundary = false;
};

// plane stride: moments planes aligned to 256 elements
inline size_t plane_stride(long long n) { return (size_t)((n + 255) / 256 * 256); }
inline size_t storage_bytes(int storage) { return storage == FF_STORAGE_F32 ? 4 : 2; }

// launchers (defined in the .hip units)
int launch_bulk2d(Domain& d, const void* in, void* out);
int launch_bulk3d(Domain& d, const void* in, void* out);
int launch_outlet(Domain& d, void* buf);
int launch_links_reset(Domain& d);
int launch_links_build(Domain& d, const float* verts, const uint32_t* elems);
int launch_boundary_cells(Domain& d, const void* in, void* out, const float* verts, const float* vels,
                          const uint32_t* elems);
int launch_forces(Domain& d, const float* verts, const uint32_t* elems, float* vforce_out);
int launch_wave(Domain& d);
int launch_init_state(Domain& d, const float* bc6);
int launch_convert_in(Domain& d, const float* aos_dev, void* buf);
int launch_convert_out(Domain& d, const void* buf, float* aos_dev);
int launch_mass(Domain& d, double* out_dev);
int launch_fill_solid(Domain& d);

}  // namespace ff
