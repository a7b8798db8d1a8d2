// ff_internal.h -- domain state and helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/freeflow_lbm.h"

namespace ff {

struct Batch;

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

#define FF_HIP(call)                                        \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return ::ff::hip_fail(_e, #call); \
  } while (0)

// Per-element FP64 record of a step (elem_acc): DIM x DIM vertex shares of
// the force (vertex j, component c at j * DIM + c), the force, the torque
// (1 in 2D, 3 in 3D) about the body reference point, and the element's power
// -sum_j share_j . u_j with the step's vertex velocities.
template <int DIM>
constexpr int elem_stride_v = DIM * DIM + DIM + (DIM == 2 ? 1 : 3) + 1;

// One outlet cell: in-place update of the new buffer after a step.
struct OutletCell {
  int cell;      // outlet cell index
  int nb;        // neighbour index (must be Fluid to act)
  int8_t comp;   // velocity slot
  int8_t pcomp;  // Pi slot
  int8_t sign;   // +1: copy when u > 0, -1: copy when u < 0
  int8_t pad;
};

// Step state kept on the device so a step's launches need no host-side
// arguments that change from step to step (they can be replayed as a graph).
struct StepState {
  unsigned gen;       // batched domains: steps since the boundary was set + 1
  unsigned tag;       // batched domains: link-table tag of the current step's links
  int overflow;       // sticky: link table full
  int pad;
};

// Everything the link kernels read, passed by value.
struct LinkArgs {
  const float* v[2];   // vertex positions by generation parity
  const float* u[2];   // vertex velocities by generation parity
  const uint32_t* e;
  int ne, nv;
  unsigned long long* cells;  // 2 * cap: two tables, by tag parity
  unsigned long long* keys;   // 2 * cap * (Q-1), by tag parity
  float* lf;                  // cap * (Q-1) * DIM  link force
  float* lb;                  // cap * (Q-1) * DIM  barycentrics, lb[0] < 0: link not applied
  float* lh;                  // cap * (Q-1) * DIM  hit point
  float* sb;                  // cap * NM  updated moments of each boundary cell (copied back after the step)
  int cap;
  int bc_span;  // link-table slots a boundary-cell block scans per pass (<= its block size)
  StepState* st;
  // Step generation (vertex-buffer parity gen & 1, wrench ring slot) and
  // link-table tag of the step: passed by value for a domain stepped on its
  // own (the host keeps them), read from st for batched domains (graph
  // replays advance them on the device).
  unsigned gen, tag;
  int dev_state;
  const uint8_t* flags;
  int nx, ny, nz;
  float cl, ct, cf, omega;
  size_t plane;
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
  int tile2 = 0;     // 2D stream-collide tile shape (index into kTiles2)
  int tile3 = 0;     // 3D stream-collide tile shape (index into kTiles3)
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
  void* bnd_mem = nullptr;   // one allocation holding the ff_set_boundary arrays below
  void* core_mem = nullptr;  // one allocation holding the moments, flags and reduction scratch
  float* vbuf[2] = {nullptr, nullptr};  // positions by generation parity
  float* ubuf[2] = {nullptr, nullptr};  // velocities by generation parity
  uint32_t* elems = nullptr;
  int* elem_body = nullptr;
  float* vforce = nullptr;        // accumulated per-vertex force
  int* body_elem_ptr = nullptr;   // CSR body -> elements
  int* body_elem_idx = nullptr;
  int* vert_inc_ptr = nullptr;    // CSR vertex -> (element << 2 | end)
  int* vert_inc_idx = nullptr;
  double* body_ref = nullptr;     // n_bodies * dim
  double* elem_acc = nullptr;     // per element: dim*dim vertex shares, dim force, torque
  double* wrench = nullptr;       // n_bodies * (2D:3, 3D:6), latest step
  double* wrench_hist = nullptr;  // ring of hist_cap steps
  double* power = nullptr;        // n_bodies: latest step's sum over the body of -F_v . u_v
  double* power_hist = nullptr;   // ring of hist_cap steps
  int hist_cap = 0;
  long long hist_count = 0;       // steps recorded since the ring was set
  // link table (sparse replacement of the dense N*Q hit table)
  int lt_cap = 0;
  unsigned long long* lt_cells = nullptr;
  unsigned long long* lt_keys = nullptr;
  float* lk_f = nullptr;
  float* lk_b = nullptr;
  float* lk_h = nullptr;
  float* lk_sb = nullptr;         // boundary-cell results (lt_cap * nm)
  uint32_t* fill_bits = nullptr;  // 2D fill-solid row marks (ny * ceil(nx / 32) words)
  StepState* st = nullptr;
  unsigned gen_host = 1;          // generation of the next step (st->gen while batched)
  unsigned tag_host = 1;          // link-table tag of the next step (st->tag while batched)
  bool finish_pending = false;    // the last step's vertex forces / wrench are not yet summed
  unsigned finish_gen = 0;        // generation of that step
  // the links of the next step were detected by the previous step's finish
  // launch; `stale` when the boundary moved since (detect again, new tag)
  bool predetected = false;
  bool stale = false;
  // wave kinematics
  bool wave = false;
  float* rest_verts = nullptr;
  float* s_coord = nullptr;
  double* wave_params = nullptr;  // kWaveParStride per body
  std::vector<double> wave_host;  // host staging of wave_params
  int* vert_body = nullptr;
  // scratch
  double* red = nullptr;          // reduction scratch
  bool has_boundary = false;
  // batch membership: while set, `stream` is the batch's stream and the
  // domain's own stream is parked in `own_stream`
  Batch* batch = nullptr;
  hipStream_t own_stream = nullptr;
  // set only while ff_time_bulk_kernel times a launch (see launch_timed)
  hipEvent_t tev0 = nullptr, tev1 = nullptr;
};

// HIP events owned by a scope: destroyed on every return, error returns
// included; a domain's or batch's timing slots (tev0/tev1) registered with
// `clears` are reset to null with them.
struct EventSet {
  std::vector<hipEvent_t> ev;
  hipEvent_t* clears[2] = {nullptr, nullptr};
  hipError_t create(size_t n) {
    ev.assign(n, nullptr);
    for (auto& e : ev) {
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  }
  hipEvent_t operator[](size_t i) const { return ev[i]; }
  ~EventSet() {
    for (auto* c : clears)
      if (c) *c = nullptr;
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// Batch bookkeeping used by the per-domain entry points (batch.hip).
void batch_mark_dirty(Batch* b);
int batch_detach(Domain& d);

// plane stride: moment planes aligned to 256 elements
inline size_t plane_stride(long long n) { return (size_t)((n + 255) / 256 * 256); }
inline size_t storage_bytes(int storage) { return storage == FF_STORAGE_F32 ? 4 : 2; }
// planes per moment buffer: the moments, plus the velocities' low halves in FF_STORAGE_F16U
inline size_t storage_planes(int storage, int dim) {
  const size_t nm = dim == 2 ? 6 : 10;
  return storage == FF_STORAGE_F16U ? nm + (size_t)dim : nm;
}

LinkArgs link_args(Domain& d, const float* v0, const float* v1, const float* u0, const float* u1, const uint32_t* e);

// launchers (defined in the .hip units)
int launch_bulk2d(Domain& d, const void* in, void* out);
int launch_bulk3d(Domain& d, const void* in, void* out);
// boundary step: [detect] -> stream (tiles + boundary cells + next kinematics)
// -> element sums + outlets + boundary-cell copy-back + next step's detection
// -> finish (vertex forces, wrench)
int launch_detect(Domain& d, const LinkArgs& la);
int launch_step_stream(Domain& d, const void* in, void* out, const LinkArgs& la, bool wave);
int launch_step_sums(Domain& d, const LinkArgs& la, void* out, bool detect_next);
int launch_step_finish(Domain& d, const LinkArgs& la, unsigned gen, float* proxy_forces);
int flush_finish(Domain& d);
int launch_outlet(Domain& d, void* buf);
int launch_init_state(Domain& d, const float* bc6);
// Launch of `kern` on `stream`.  With timing events (ev0 non-null) the
// events take the kernel's own start and end (hipExtLaunchKernel), so a timed
// launch measures the kernel alone, as a kernel trace does, without the
// event packets around it.
template <typename... P, typename... A>
inline void launch_timed(void (*kern)(P...), dim3 grid, dim3 block, hipStream_t stream, hipEvent_t ev0,
                         hipEvent_t ev1, A... args) {
  if (ev0)
    hipExtLaunchKernelGGL(kern, grid, block, 0, stream, ev0, ev1, 0, (P)args...);
  else
    hipLaunchKernelGGL(kern, grid, block, 0, stream, (P)args...);
}

int launch_wave_now(Domain& d);
// Wave parameters on the device: per body 4 FP64 values (A, f, phi, L/lambda)
// and 8 FP32 values derived from them for the fast-math kinematics.
constexpr int kWaveParStride = 8;  // doubles per body
void wave_device_params(const double* p4, double dt, int n_bodies, double* out);
int bc_span(const Domain& d);  // boundary-cell scan span of a domain stepped on its own
int launch_step_vertex_forces(Domain& d, float* out_dev);
int launch_convert_in(Domain& d, const float* aos_dev, void* buf);
int launch_convert_out(Domain& d, const void* buf, float* aos_dev);
int launch_mass(Domain& d, double* out_dev);
int launch_fill_solid(Domain& d, const float* verts);

}  // namespace ff
