/*
 * freeflow_lbm.h -- C ABI of the MI355X (gfx950) HOME-LBM fluid step with
 * moving-boundary coupling.
 *
 * This is the seam the reference's host code calls into.  In FreeFlow the
 * fluid side is entered through
 *   fsi::lbm::fsi_coupling::solveLbmAndFsiStep2D(LbmFlowField2D&,
 *       const SolidGeometryProxy_Device<2>&, const SimulationParameters2D&,
 *       cudaStream_t)                       (include/FSI_Simulator/core/FsiCoupling.hpp:21-25)
 *   ... solveLbmAndFsiStep3D(...)          (FsiCoupling.hpp:34-38)
 *   fillSolidFlags2D(...)                   (FsiCoupling.hpp:27-31)
 * with the flow field owning two moment buffers, the node flags and a dense
 * hit table (include/FSI_Simulator/lbm/LbmFlowField2D.hpp:51-59).  Here the
 * flow field is an opaque ff_domain owned by the library; every function
 * takes plain pointers and sizes, returns an int status (FF_OK = 0) and
 * records a message retrievable with ff_last_error().
 *
 * Threading: calls on one domain must be serialised by the caller; different
 * domains may be driven from different threads.  ff_step* calls are
 * asynchronous on the domain's HIP stream; getters synchronise.
 *
 * Moments use the reference's per-cell order
 *   2D: [rho, ux, uy, Pi_xx, Pi_yy, Pi_xy]
 *   3D: [rho, ux, uy, uz, Pi_xx, Pi_yy, Pi_zz, Pi_xy, Pi_yz, Pi_xz]
 * (src/core/FsiCoupling2D.cu:247-252, FsiCoupling3D.cu:310-320) and the
 * reference's cell order y*nx+x / (z*ny+y)*nx+x; on the device they are kept
 * as structure-of-arrays planes.
 */
#ifndef FREEFLOW_LBM_H
#define FREEFLOW_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FF_OK 0
#define FF_ERR_INVALID 1  /* bad argument or size */
#define FF_ERR_HIP 2      /* HIP runtime error (message has the HIP string) */
#define FF_ERR_NOMEM 3    /* device allocation failed */
#define FF_ERR_CAPACITY 4 /* a fixed-capacity device table overflowed */
#define FF_ERR_STATE 5    /* call made in the wrong state (e.g. no boundary set) */
#define FF_ERR_NODEVICE 6 /* no usable gfx950 device */

/* moment storage formats */
#define FF_STORAGE_F32 0
#define FF_STORAGE_F16 1  /* binary16, rho-1 and Pi scaled by 2^8, u by 2^4 */
#define FF_STORAGE_BF16 2 /* bfloat16, same encoding */
#define FF_STORAGE_F16U 3 /* binary16 as FF_STORAGE_F16, each velocity component
                             kept as a pair hi + lo of binary16 values (about
                             FP32 precision; lo in DIM extra planes after the
                             NM moment planes): 2 B x (NM + DIM) per cell */

/* option bits */
#define FF_OPT_STAGE_FP16 0x1u /* 2D: round pulled populations to binary16 as the
                                  reference's half LDS tile does
                                  (src/core/FsiCoupling2D.cu:89,120-128) */
#define FF_OPT_PERIODIC 0x2u   /* pull across the domain edges periodically instead of
                                  the reference's clamped halo (FsiCoupling2D.cu:104-105) */
#define FF_OPT_FAST_MATH 0x4u  /* the same update in a reassociated FP32 form with FMA
                                  (pair sums of opposite populations) instead of the
                                  reference's operation order (src/lbm/LbmUtilsFunc.cu:
                                  212-543, FsiCoupling2D.cu:154-252, FsiCoupling3D.cu:
                                  291-320): 2D tiles without FF_OPT_STAGE_FP16 and every
                                  3D tile, with their boundary cells, and FP32 wave
                                  kinematics; FP64 stays only in the
                                  per-body force and power sums.  Moments agree with the
                                  reference order to <= 1e-5 relative L2 after 100 steps,
                                  not bit for bit */

/* Node flag values: the ordinals of fsi::lbm::LbmNodeFlag
   (include/FSI_Simulator/lbm/LbmDataTypes.cuh:17-56). */
#define FF_FLAG_INVALID 0
#define FF_FLAG_FLUID 1
#define FF_FLAG_SOLID 3
#define FF_FLAG_SOLID_DYNAMIC 4
#define FF_FLAG_WALL 5

typedef struct ff_domain_desc {
  int dim;               /* 2 (D2Q9) or 3 (D3Q27) */
  int nx, ny, nz;        /* nz ignored for dim 2 */
  double viscosity;      /* lattice viscosity: SimulationParameters*::fluid_viscosity */
  double dx;             /* fluid_dx (cl) */
  double dt;             /* dt (ct) */
  double fluid_density;  /* fluid_density, used for cf = rho dx^D dx / dt^2
                            (include/FSI_Simulator/utils/Config.hpp:96-100, 154-158) */
  int storage;           /* FF_STORAGE_* */
  unsigned options;      /* FF_OPT_* */
  int device;            /* HIP device ordinal */
} ff_domain_desc;

typedef struct ff_domain ff_domain;

/* Replaces LbmFlowField2D/3D construction (src/lbm/LbmFlowField2D.cpp:17-40,
   LbmFlowField3D.cpp:18-42).  All cells start as Fluid with zero moments. */
int ff_domain_create(const ff_domain_desc* desc, ff_domain** out);
int ff_domain_destroy(ff_domain* d);

/* Node flags, one LbmNodeFlag byte per cell (the flag upload of
   LbmInitializer2D.cu:106-128). */
int ff_set_flags(ff_domain* d, const uint8_t* flags);
int ff_get_flags(ff_domain* d, uint8_t* flags);

/* Equilibrium initial state from the flags and the inlet velocities
   (initialize_state_kernel_2d/3d, LbmInitializer2D.cu:32-95,
   LbmInitializer3D.cu:33-113): writes both ping-pong buffers.  n_bc is 4
   (2D) or 6 (3D); any other count means zeros, as in the reference. */
int ff_init_state(ff_domain* d, const float* bc_velocities, int n_bc);

/* Host AoS moments in/out.  which: 0 = current, 1 = the other ping-pong
   buffer, -1 = both (set only).  Outlet cells keep a history per buffer, as
   in the reference, so a faithful restart sets both buffers. */
int ff_set_moments(ff_domain* d, const float* aos, int which);
int ff_get_moments(ff_domain* d, float* aos, int which);

/* Boundary geometry: the data of SolidGeometryProxy_Device<2/3>
   (include/FSI_Simulator/core/SolidGeometryProxy_Device.cuh:29-73):
   nv vertices (dim floats each), their velocities, ne elements (segments:
   2 vertex ids; triangles: 3).  elem_body (may be NULL) assigns each
   element to one of n_bodies bodies for the per-body wrench.  At most
   65536 elements (the reference's 16-bit element field of the hit key,
   SolidGeometryProxy_Device.cuh:79-85).  Vertex coordinates must be finite
   and within 2^30 cells of the origin.  Copies the arrays.  A rejected call
   (FF_ERR_INVALID) keeps the previous boundary. */
int ff_set_boundary(ff_domain* d, int nv, const float* verts, const float* vels, int ne, const uint32_t* elems,
                    const int32_t* elem_body, int n_bodies);
/* New vertex positions/velocities for the next step (same topology). */
int ff_update_boundary(ff_domain* d, const float* verts, const float* vels);
/* Reference point per body for the torque (default: the vertex centroid at
   ff_set_boundary time). */
int ff_set_body_ref(ff_domain* d, const double* ref);

/* Prescribed travelling-wave kinematics evaluated on the device each step:
   vertex k moves laterally (along y in 2D) by
     A_b * s_k * sin(2 pi (s_k * L_b / lambda_b - f_b t) + phi_b)
   from its rest position, where s_k in [0,1] is the body-axis coordinate
   given per vertex and t = step * dt.  params holds 4 doubles per body:
   {A (m), f (Hz), phi (rad), L/lambda}.  Replaces ff_update_boundary. */
int ff_set_wave(ff_domain* d, const float* rest_verts, const float* s_coord, const double* params);
/* New wave parameters (n_bodies x 4 doubles, as ff_set_wave) for the wave set
   by ff_set_wave, e.g. an RL action: ordered on the domain's stream without
   waiting for it (the host array may be reused on return).  No counterpart
   in the reference (its actions go through applyKinematicControl). */
int ff_set_wave_params(ff_domain* d, const double* params);

/* One or more fluid steps with the library-held boundary: the work of
   solveLbmAndFsiStep2D/3D (FsiCoupling2D.cu:358-432, FsiCoupling3D.cu:424-484):
   link detection, stream + collide with moving-wall bounce-back and
   momentum-exchange forces, ping-pong swap, outlet update. */
int ff_step(ff_domain* d, int nsteps);

/* Direct replacement of the proxy call: device pointers with the layout of
   SolidGeometryProxy_Device (interleaved float vertices/velocities, uint32
   elements).  The step's per-vertex forces are added to d_forces, as
   addForceAt does (SolidGeometryProxy_Device.cuh:184-194, :401-421).  The
   pointers must stay valid until the step has executed on the stream. */
int ff_step_proxy(ff_domain* d, const float* d_vertices, const float* d_velocities, const uint32_t* d_elements,
                  float* d_forces, int num_vertices, int num_elements);

/* Per-vertex hydrodynamic forces accumulated over steps (nv*dim floats);
   reset != 0 zeroes the accumulator after reading. */
int ff_get_vertex_forces(ff_domain* d, float* out, int reset);
/* Per-vertex forces of the most recent step only (nv*dim floats): what the
   reference's VBD path scatters to the volume each step
   (scatter_forces_kernel2d, src/fsi/FsiCouplingKernels.cu:77-91). */
int ff_get_step_vertex_forces(ff_domain* d, float* out);
/* ff_get_step_vertex_forces without any device allocation: the forces are
   staged in a static per-GPU array of 1 048 576 floats (4 MiB per GPU, held
   until the process ends) instead of a hipMalloc/hipFree per call; a domain
   with nv * dim above that is rejected with FF_ERR_CAPACITY.  No counterpart
   in the reference: for callers that cannot allocate device memory (a
   per-step RL readback loop, code under stream capture, allocation-free
   serving loops).  Calls on one GPU share the array; calls that overlap on
   one GPU (other threads or streams) race. */
int ff_get_step_vertex_forces_noalloc(ff_domain* d, float* out);
/* Per-body wrench of the most recent step, FP64: 2D {Fx, Fy, Tz},
   3D {Fx, Fy, Fz, Tx, Ty, Tz} per body, torque about the body reference. */
int ff_get_body_wrench(ff_domain* d, double* out);
/* Per-step wrench history of the last ff_step call (ring of up to
   `capacity` steps, set by ff_set_wrench_history); writes min(count, steps)
   records, most recent last. Returns the number written in *n_out. */
int ff_set_wrench_history(ff_domain* d, int capacity);
int ff_get_wrench_history(ff_domain* d, double* out, int count, int* n_out);

/* Per-body power of the latest step, n_bodies doubles: the sum over the
   body's vertices of -F_v . u_v, F_v the step's hydrodynamic vertex force
   (the FP64 vertex shares of ff_get_step_vertex_forces) and u_v the vertex
   velocities the step used; accumulated in FP64 on the device.  Replaces the
   host-side power = sum(-force * velocity) of the reference's RL loop
   (rl/env.py:175-178), which copies every vertex force and velocity to the
   host each simulation step. */
int ff_get_body_power(ff_domain* d, double* out);
/* The per-body power of the last `count` steps (oldest first) from the ring
   that ff_set_wrench_history sizes: *n_out rows of n_bodies doubles. */
int ff_get_power_history(ff_domain* d, double* out, int count, int* n_out);
/* Current vertex positions/velocities held by the library (after wave
   kinematics this is the state used by the most recent step). */
int ff_get_boundary(ff_domain* d, float* verts, float* vels);

/* Global FP64 sum of rho over Fluid/SolidDynamic cells of the current buffer. */
int ff_mass(ff_domain* d, double* out);

/* fillSolidFlags2D (FsiCoupling2D.cu:318-354, 434-464): Fluid/SolidDynamic
   cells become Fluid, then Fluid cells whose centre lies inside the boundary
   polygon become SolidDynamic.  2D only. */
int ff_fill_solid(ff_domain* d);

int ff_sync(ff_domain* d);
/* Average duration (ms) of the step's stream launch (the stream-collide
   kernel; with a boundary, the fused boundary-cell and kinematics blocks too)
   over `iters` launches on the domain's stream, each timed by HIP events
   attached to the launch itself (hipExtLaunchKernel: the kernel's own start
   and end); does not advance the simulation state.  Every launch reads the
   same buffer and writes the other, so caches hold what a repeated launch
   leaves there, not what stepping leaves (ff_step_timed times real steps). */
int ff_time_bulk_kernel(ff_domain* d, int iters, double* avg_ms);
/* Wall time (ms) of `nsteps` ff_step-equivalent steps measured with HIP
   events on the domain stream (state advances). */
int ff_time_steps(ff_domain* d, int nsteps, double* total_ms);
/* `nsteps` ff_step steps (the state advances), timed: *total_ms is the wall
   time from HIP events on the domain stream, *stream_ms the sum over the
   steps of the stream launch's own duration (events attached to each launch,
   so the launch is measured in the cache state real stepping leaves, each
   launch reading what the previous one wrote).  Synchronises. */
int ff_step_timed(ff_domain* d, int nsteps, double* total_ms, double* stream_ms);

/* Device bytes ff_domain_create + ff_set_boundary would allocate for this
   descriptor and boundary (verts/elems may be NULL when ne == 0): the figure
   an admission controller charges against the GPU's memory.  Laid out by the
   same plans the two calls allocate with, each allocation rounded up to
   2 MiB, the per-body arrays sized for ne bodies (an upper bound).  A wrench
   history ring (ff_set_wrench_history) is extra: capacity x bodies x 3 (2D)
   or 6 (3D) doubles.  No device work. */
int ff_domain_footprint(const ff_domain_desc* desc, int nv, int ne, const float* verts, const uint32_t* elems,
                        unsigned long long* bytes);

/* Stream-collide tile shape.  2D (0..4): 32x8, 32x16, 64x8, 32x32, 64x16
   cells per 256-thread block.  3D (0..15): 0-1 8x8x4 and 16x4x4 boxes; 2-6
   columns that rotate reconstructed populations through LDS (16x8 marching
   32, 16 or 8 planes, 16x16 marching 32 or 16); 7-15 columns with one
   staged cell per thread (7: 16x12 x 8 planes, 8: 16x26 x 8, 9: 64x4 x 16;
   10-13 keep the next planes' moments in an LDS ring filled by LDS DMA with
   FP32 storage: 16x12 x 8 with 2 planes in flight, 16x8 x 8 with 3, 32x8 x 8
   with 2, 16x12 x 16 with 3; 14: 64x4 x 32, 15: 32x12 x 16).
   0 is the default; results do not depend on it (in either arithmetic form).
   No counterpart in the reference (its tile is fixed, CudaCommon.cuh:44-51). */
int ff_set_tile(ff_domain* d, int tile);
/* Time every tile shape on this domain (three interleaved rounds, each
   shape's best round; the state is left unchanged) and keep the fastest;
   writes its index to *tile. */
int ff_autotune(ff_domain* d, int* tile);

/* ---- batches: many domains of one kind stepped together ----------------
   No counterpart in the reference, which steps one flow field per Simulator
   (src/core/Simulator2D.cpp:88-130).  A batch holds domains with the same
   dim, storage, options and device.  One ff_batch_step step of every member
   is one replay of a captured HIP graph (runs of 8 or more steps replay an
   8-step graph; the first 4 steps after a membership change are launched
   directly); results of each member equal those of ff_step on it alone.  While a domain is in a batch, its own ff_step /
   ff_step_proxy refuse to run; every other per-domain call (getters,
   ff_set_wave, ff_update_boundary, ...) keeps working and is ordered with
   the batch's steps.  Changing a member's buffers (ff_set_boundary,
   ff_set_flags, ...) or the membership rebuilds the graph at the next step. */
typedef struct ff_batch ff_batch;
int ff_batch_create(int dim, int storage, unsigned options, int device, ff_batch** out);
/* Detaches remaining members (they stay valid domains). */
int ff_batch_destroy(ff_batch* b);
int ff_batch_add(ff_batch* b, ff_domain* d);
int ff_batch_remove(ff_batch* b, ff_domain* d);
int ff_batch_size(ff_batch* b, int* n);
/* Asynchronous on the batch stream. */
int ff_batch_step(ff_batch* b, int nsteps);
/* Waits for the batch stream; reports a link-table overflow of any member. */
int ff_batch_sync(ff_batch* b);
/* Fix the tile shape of the batch (index as for ff_set_tile; -1: choose it
   by timing the candidates on the members, the default). */
int ff_batch_set_tile(ff_batch* b, int tile);
/* Tile index in use and its cell extent. */
int ff_batch_get_tile(ff_batch* b, int* tile, int* tile_x, int* tile_y);
/* Run nsteps steps and return their GPU time in ms (HIP events). */
int ff_batch_time_steps(ff_batch* b, int nsteps, double* total_ms);
/* Average duration (ms, HIP events attached to each launch) of the batched
   stream-collide kernel over `iters` launches, tiles only (no boundary-cell
   or kinematics blocks); the state is left unchanged. */
int ff_batch_time_bulk(ff_batch* b, int iters, double* avg_ms);

const char* ff_last_error(void);
int ff_device_count(int* n);
/* Library build identification string. */
const char* ff_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FREEFLOW_LBM_H */
