// capi.hip -- the extern "C" boundary: domain lifecycle, uploads, the step
// driver and getters.  Host-side orchestration of the kernels in
// kernels2d.hip, kernels3d.hip and links.hip.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "ff_internal.h"
#include "stencil.h"

struct ff_domain : public ff::Domain {};

namespace ff {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? FF_ERR_NOMEM : FF_ERR_HIP;
}

template <class T>
static int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return FF_OK;
}
template <class T>
static void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Streams of destroyed domains, kept per device for the next domains created
// there (a server creates and destroys a domain per episode, and creating a
// HIP stream costs more than a step of a small domain).  A stream is idle
// when it is returned.
static std::mutex g_stream_mu;
static std::vector<std::pair<int, hipStream_t>> g_streams;
constexpr size_t kStreamPool = 64;

static hipError_t acquire_stream(int device, hipStream_t* s) {
  {
    std::lock_guard<std::mutex> lock(g_stream_mu);
    for (size_t i = g_streams.size(); i-- > 0;)
      if (g_streams[i].first == device) {
        *s = g_streams[i].second;
        g_streams.erase(g_streams.begin() + (long)i);
        return hipSuccess;
      }
  }
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

static void release_stream(int device, hipStream_t s) {
  if (!s) return;
  if (hipStreamSynchronize(s) == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_stream_mu);
    if (g_streams.size() < kStreamPool) {
      g_streams.push_back({device, s});
      return;
    }
  }
  (void)hipStreamDestroy(s);
}

// Arrays carved out of one device allocation (one hipMalloc and one hipFree
// for a group of arrays that live and die together), each 256-byte aligned.
struct Carve {
  std::vector<std::function<void(unsigned char*)>> setters;  // each sets one array from the base
  size_t total = 0;
  template <class T>
  size_t add(T** p, size_t count) {
    const size_t off = total;
    setters.push_back([p, off](unsigned char* base) { *p = reinterpret_cast<T*>(base + off); });
    total += (std::max<size_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
    return off;
  }
  int commit(void** base) {
    *base = nullptr;
    hipError_t e = hipMalloc(base, total);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    for (auto& set : setters) set(static_cast<unsigned char*>(*base));
    return FF_OK;
  }
};

static void free_boundary(Domain& d) {
  // the arrays set up by ff_set_boundary share one allocation
  dfree(d.bnd_mem);
  for (int b = 0; b < 2; b++) d.vbuf[b] = d.ubuf[b] = nullptr;
  d.elems = nullptr; d.elem_body = nullptr; d.vforce = nullptr;
  d.body_elem_ptr = d.body_elem_idx = d.vert_inc_ptr = d.vert_inc_idx = nullptr;
  d.body_ref = d.elem_acc = d.wrench = nullptr;
  d.lt_cells = d.lt_keys = nullptr; d.lk_f = d.lk_b = d.lk_h = d.lk_sb = nullptr;
  d.st = nullptr; d.vert_body = nullptr;
  d.rest_verts = nullptr; d.s_coord = nullptr; d.wave_params = nullptr;
  dfree(d.wrench_hist);
  dfree(d.power_hist);
  d.power = nullptr;
  d.nv = d.ne = d.n_bodies = 0;
  d.gen_host = 1;
  d.tag_host = 1;
  d.finish_pending = false;
  d.has_boundary = false;
  d.wave = false;
  d.hist_cap = 0;
  d.hist_count = 0;
  d.predetected = false;
  d.stale = false;
}

static int elem_stride(int dim) { return dim == 2 ? elem_stride_v<2> : elem_stride_v<3>; }
static int wrench_width(int dim) { return dim == 2 ? 3 : 6; }

// Outlet list from the flags (mrApplyOutlet2DKernel / 3D branch table).
static void build_outlets(Domain& d, std::vector<OutletCell>& out) {
  out.clear();
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  for (long long c = 0; c < d.n; c++) {
    const uint8_t f = d.h_flags[c];
    if (f < kOutletLeft || f > kOutletBack) continue;
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((long long)nx * ny));
    OutletCell o{};
    o.cell = (int)c;
    int nbx = x, nby = y, nbz = z;
    if (d.dim == 2) {
      if (f == kOutletUp) { nby = y - 1; o.comp = 2; o.pcomp = 4; o.sign = 1; }
      else if (f == kOutletDown) { nby = y + 1; o.comp = 2; o.pcomp = 4; o.sign = -1; }
      else if (f == kOutletLeft) { nbx = x + 1; o.comp = 1; o.pcomp = 3; o.sign = -1; }
      else if (f == kOutletRight) { nbx = x - 1; o.comp = 1; o.pcomp = 3; o.sign = 1; }
      else continue;  // Front/Back have no 2D branch
    } else {
      if (f == kOutletUp) { nbz = z - 1; o.comp = 3; o.pcomp = 6; o.sign = 1; }
      else if (f == kOutletDown) { nbz = z + 1; o.comp = 3; o.pcomp = 6; o.sign = -1; }
      else if (f == kOutletLeft) { nbx = x + 1; o.comp = 1; o.pcomp = 4; o.sign = -1; }
      else if (f == kOutletRight) { nbx = x - 1; o.comp = 1; o.pcomp = 4; o.sign = 1; }
      else if (f == kOutletFront) { nby = y + 1; o.comp = 2; o.pcomp = 5; o.sign = -1; }
      else { nby = y - 1; o.comp = 2; o.pcomp = 5; o.sign = 1; }
    }
    // the reference reads out of the grid here; such cells are skipped
    if (nbx < 0 || nbx >= nx || nby < 0 || nby >= ny || nbz < 0 || nbz >= nz) continue;
    o.nb = (int)(((long long)nbz * ny + nby) * nx + nbx);
    out.push_back(o);
  }
}

static int upload_aos(Domain& d, const float* aos, int which) {
  float* tmp = nullptr;
  const size_t cnt = (size_t)d.n * d.nm;
  int rc = dalloc(&tmp, cnt);
  if (rc) return rc;
  hipError_t e = hipMemcpy(tmp, aos, cnt * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) { dfree(tmp); return hip_fail(e, "hipMemcpy"); }
  for (int b = 0; b < 2; b++) {
    if (which >= 0 && which != b) continue;
    void* target = d.mom[b == 0 ? d.cur : (d.cur ^ 1)];
    rc = launch_convert_in(d, tmp, target);
    if (rc) { dfree(tmp); return rc; }
  }
  e = hipStreamSynchronize(d.stream);
  dfree(tmp);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return FF_OK;
}

static unsigned next_pow2(unsigned long long v) {
  unsigned p = 1024;
  while (p < v && p < (1u << 30)) p <<= 1;
  return p;
}

// Descriptor checks that need no device: dimension, grid, storage, dx.
static int check_grid(const ff_domain_desc* desc) {
  if (desc->dim != 2 && desc->dim != 3) return fail(FF_ERR_INVALID, "dim must be 2 or 3");
  if (desc->nx <= 0 || desc->ny <= 0 || (desc->dim == 3 && desc->nz <= 0))
    return fail(FF_ERR_INVALID, "grid sizes must be positive");
  const long long n = (long long)desc->nx * desc->ny * (desc->dim == 3 ? desc->nz : 1);
  if (n >= (1ll << 31)) return fail(FF_ERR_INVALID, "grid larger than 2^31 cells");
  if (desc->storage < 0 || desc->storage > 3) return fail(FF_ERR_INVALID, "unknown storage format");
  if (!(desc->dx > 0.0)) return fail(FF_ERR_INVALID, "dx must be positive");
  return FF_OK;
}

// Vertex coordinates must be finite and within 2^30 cells of the origin: the
// link-table sizing converts them to 64-bit cell indices.
static int check_verts(int D, double dx, int nv, const float* verts) {
  for (long long k = 0; k < (long long)nv * D; k++)
    if (!(std::fabs(verts[k] / dx) <= (double)(1 << 30)))
      return fail(FF_ERR_INVALID, "vertex coordinate not finite or beyond 2^30 cells");
  return FF_OK;
}

// Link-table sizing: the distinct cells within two cells of some element's
// bounding box in the shape given (a link cell lies within sqrt(3) cells of
// an element it crosses; the extra cell, and the factor 2 the table size
// applies on top, leave room for motion and deformation).  Counted on a
// bitmap over the mesh's cell box; a box above 2^28 cells falls back to the
// sum of the element boxes.  A table that still fills up reports
// FF_ERR_CAPACITY.
static unsigned long long link_estimate(int D, double dx, int ne, const float* verts, const uint32_t* elems) {
  if (ne <= 0) return 0;
  std::vector<long long> lo((size_t)ne * D), hi((size_t)ne * D);
  long long glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
  unsigned long long sum = 0;
  for (int k = 0; k < ne; k++) {
    unsigned long long cells = 1;
    for (int c = 0; c < D; c++) {
      double mn = 1e300, mx = -1e300;
      for (int j = 0; j < D; j++) {
        const double x = verts[(size_t)elems[(size_t)k * D + j] * D + c] / dx;
        mn = std::min(mn, x);
        mx = std::max(mx, x);
      }
      const long long a = (long long)std::floor(mn) - 2, b = (long long)std::floor(mx) + 2;
      lo[(size_t)k * D + c] = a;
      hi[(size_t)k * D + c] = b;
      glo[c] = k ? std::min(glo[c], a) : a;
      ghi[c] = k ? std::max(ghi[c], b) : b;
      const unsigned long long w = (unsigned long long)(b - a + 1);
      cells = cells > (1ull << 40) / w ? (1ull << 40) : cells * w;  // saturates; the sum stays below 2^57
    }
    sum += cells;
  }
  unsigned long long box = 1;
  for (int c = 0; c < D; c++) box *= (unsigned long long)(ghi[c] - glo[c] + 1);
  if (box > (1ull << 28)) return sum;
  const long long ex = ghi[0] - glo[0] + 1, ey = ghi[1] - glo[1] + 1;
  std::vector<uint64_t> bits((box + 63) / 64, 0);
  unsigned long long count = 0;
  for (int k = 0; k < ne; k++) {
    const long long* l = &lo[(size_t)k * D];
    const long long* h = &hi[(size_t)k * D];
    const long long z0 = D == 3 ? l[2] : 0, z1 = D == 3 ? h[2] : 0, gz = D == 3 ? glo[2] : 0;
    for (long long z = z0; z <= z1; z++)
      for (long long y = l[1]; y <= h[1]; y++)
        for (long long x = l[0]; x <= h[0]; x++) {
          const unsigned long long idx = (unsigned long long)(((z - gz) * ey + (y - glo[1])) * ex + (x - glo[0]));
          const uint64_t m = 1ull << (idx & 63);
          if (!(bits[idx >> 6] & m)) {
            bits[idx >> 6] |= m;
            count++;
          }
        }
  }
  return count;
}

// The core allocation of a domain: both moment buffers, the flags and the
// reduction scratch.  Returns the offset of the flags.
static size_t plan_core(Carve& cv, Domain& d, unsigned char** mom0, unsigned char** mom1) {
  const size_t bytes = d.plane * storage_planes(d.storage, d.dim) * storage_bytes(d.storage);
  cv.add(mom0, bytes);
  cv.add(mom1, bytes);
  const size_t flags_off = cv.add(&d.flags, (size_t)d.n);
  cv.add(&d.red, 1024);
  return flags_off;
}

// The one allocation behind ff_set_boundary: region 1 is initialised from
// the host, region 2 cleared, region 3 written by the kernels before any read.
struct BoundaryPlan {
  size_t o_v[2], o_u[2], o_el, o_eb, o_vb, o_bp, o_bi, o_vp, o_vi, o_rf, o_st;
  size_t init_end, zero_end;
};
static BoundaryPlan plan_boundary(Carve& cv, Domain& d, int nv, int ne, int n_bodies, size_t cap) {
  const int D = d.dim;
  const size_t K = (size_t)(d.q - 1);
  const size_t vcount = (size_t)std::max(nv, 1) * D;
  BoundaryPlan p;
  for (int b = 0; b < 2; b++) {
    p.o_v[b] = cv.add(&d.vbuf[b], vcount);
    p.o_u[b] = cv.add(&d.ubuf[b], vcount);
  }
  p.o_el = cv.add(&d.elems, (size_t)std::max(ne, 1) * D);
  p.o_eb = cv.add(&d.elem_body, (size_t)std::max(ne, 1));
  p.o_vb = cv.add(&d.vert_body, (size_t)std::max(nv, 1));
  p.o_bp = cv.add(&d.body_elem_ptr, (size_t)n_bodies + 1);
  p.o_bi = cv.add(&d.body_elem_idx, (size_t)std::max(ne, 1));
  p.o_vp = cv.add(&d.vert_inc_ptr, (size_t)nv + 1);
  p.o_vi = cv.add(&d.vert_inc_idx, std::max<size_t>((size_t)ne * D, 1));
  p.o_rf = cv.add(&d.body_ref, (size_t)n_bodies * D);
  p.o_st = cv.add(&d.st, 1);
  p.init_end = cv.total;
  // region 2 (generation 0 marks every table entry empty; steps run from generation 1)
  cv.add(&d.vforce, vcount);
  cv.add(&d.wrench, (size_t)n_bodies * wrench_width(D));
  cv.add(&d.power, (size_t)n_bodies);
  cv.add(&d.lt_cells, 2 * cap);  // two tables (cells and keys), alternating by tag parity
  cv.add(&d.lt_keys, 2 * cap * K);
  p.zero_end = cv.total;
  // region 3
  cv.add(&d.elem_acc, (size_t)std::max(ne, 1) * elem_stride(D));
  cv.add(&d.lk_f, cap * K * D);
  cv.add(&d.lk_b, cap * K * D);
  cv.add(&d.lk_h, cap * K * D);
  cv.add(&d.lk_sb, cap * d.nm);
  cv.add(&d.fill_bits, D == 2 ? (size_t)d.ny * (size_t)((d.nx + 31) / 32) : 1);
  // wave kinematics inputs (ff_set_wave fills them when it is used)
  cv.add(&d.rest_verts, vcount);
  cv.add(&d.s_coord, (size_t)std::max(nv, 1));
  cv.add(&d.wave_params, (size_t)n_bodies * kWaveParStride);
  return p;
}

// hipMalloc hands out whole pages of the device allocator: each of a
// domain's allocations is counted rounded up to 2 MiB.
static unsigned long long alloc_rounded(size_t bytes) { return (bytes + (2u << 20) - 1) / (2u << 20) * (2ull << 20); }

// Device bytes a domain allocates: its core allocation and, with a boundary,
// the ff_set_boundary allocation, laid out by the same plans the allocating
// calls use, with n_bodies bodies (the per-body arrays; ne bounds it).  A
// wrench history ring adds its own allocation: capacity x bodies x 3 (2D) or
// 6 (3D) doubles.
static unsigned long long footprint(int dim, long long n, int storage, int nv, int ne, int n_bodies,
                                    unsigned long long est) {
  Domain d;
  d.dim = dim;
  d.nm = dim == 2 ? 6 : 10;
  d.q = dim == 2 ? 9 : 27;
  d.storage = storage;
  d.n = n;
  d.plane = plane_stride(n);
  unsigned char *m0 = nullptr, *m1 = nullptr;
  Carve core;
  plan_core(core, d, &m0, &m1);
  unsigned long long b = alloc_rounded(core.total);
  if (nv > 0 || ne > 0) {
    Carve bc;
    plan_boundary(bc, d, nv, ne, std::max(n_bodies, 1), (size_t)next_pow2(est * 2 + 1024));
    b += alloc_rounded(bc.total);
  }
  return b;
}

LinkArgs link_args(Domain& d, const float* v0, const float* v1, const float* u0, const float* u1, const uint32_t* e) {
  LinkArgs la;
  la.v[0] = v0;
  la.v[1] = v1;
  la.u[0] = u0;
  la.u[1] = u1;
  la.e = e;
  la.ne = d.ne;
  la.nv = d.nv;
  la.cells = d.lt_cells;
  la.keys = d.lt_keys;
  la.lf = d.lk_f;
  la.lb = d.lk_b;
  la.lh = d.lk_h;
  la.sb = d.lk_sb;
  la.cap = d.lt_cap;
  la.bc_span = bc_span(d);
  la.st = d.st;
  la.gen = d.gen_host;
  la.tag = d.tag_host;
  la.dev_state = 0;
  la.flags = d.flags;
  la.nx = d.nx;
  la.ny = d.ny;
  la.nz = d.nz;
  la.cl = d.cl;
  la.ct = d.ct;
  la.cf = d.cf;
  la.omega = d.omega;
  la.plane = d.plane;
  return la;
}

// The links detected ahead by the last step no longer describe the boundary.
static void links_moved(Domain& d) {
  if (d.predetected) d.stale = true;
  d.predetected = false;
}

// One step.  Without a boundary: [stream-collide] -> [outlet].  With one:
// [detect (first step or after the boundary moved)] -> [stream: tiles +
// boundary cells + next wave kinematics] -> [sums: element forces, outlets,
// boundary-cell copy-back, next step's links] -> [finish: vertex forces,
// wrench], the finish deferred into the next step's stream launch (or run
// before anything reads its results).  `pv` != null selects caller-owned
// proxy arrays (positions not known ahead: detected every step; finish run
// at once).
static int one_step(Domain& d, const float* pv, const float* pu, const uint32_t* pe, float* proxy_forces) {
  void* in = d.mom[d.cur];
  void* out = d.mom[d.cur ^ 1];
  int rc;
  if (!(d.has_boundary && d.ne > 0)) {
    if ((rc = d.dim == 2 ? launch_bulk2d(d, in, out) : launch_bulk3d(d, in, out))) return rc;
    d.cur ^= 1;
    return launch_outlet(d, d.mom[d.cur]);
  }
  const bool proxy = pv != nullptr;
  if (proxy) links_moved(d);
  // links detected ahead for a boundary moved since: a fresh tag drops them
  if (!d.predetected && d.stale) d.tag_host++;
  const LinkArgs la = proxy ? link_args(d, pv, pv, pu, pu, pe) : link_args(d, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], d.elems);
  if (!d.predetected) {
    if ((rc = launch_detect(d, la))) return rc;
    d.stale = false;
  }
  // the stream launch also runs the previous step's deferred finish
  if ((rc = launch_step_stream(d, in, out, la, d.wave && !proxy))) return rc;
  d.finish_pending = false;
  if ((rc = launch_step_sums(d, la, out, !proxy))) return rc;
  if (proxy) {
    // caller-owned force arrays are filled by the step itself
    if ((rc = launch_step_finish(d, la, d.gen_host, proxy_forces))) return rc;
  } else {
    d.finish_pending = true;
    d.finish_gen = d.gen_host;
  }
  d.predetected = !proxy;
  d.cur ^= 1;
  d.gen_host++;
  d.tag_host++;
  if (d.hist_cap > 0) d.hist_count++;
  return FF_OK;
}

static int check_overflow(Domain& d) {
  if (!d.has_boundary) return FF_OK;
  StepState st;
  FF_HIP(hipMemcpy(&st, d.st, sizeof(st), hipMemcpyDeviceToHost));
  if (st.overflow) {
    int z = 0;
    FF_HIP(hipMemcpy(&d.st->overflow, &z, sizeof(int), hipMemcpyHostToDevice));
    return fail(FF_ERR_CAPACITY, "boundary link table overflow (capacity " + std::to_string(d.lt_cap) + " cells)");
  }
  return FF_OK;
}

}  // namespace ff

using namespace ff;

#define FF_CHECK_DOMAIN(d) \
  if (!(d)) return fail(FF_ERR_INVALID, "null domain")

extern "C" {

const char* ff_last_error(void) { return g_error.c_str(); }
const char* ff_version(void) { return "freeflow_lbm gfx950 1.0"; }

int ff_device_count(int* n) {
  if (!n) return fail(FF_ERR_INVALID, "null output");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) { *n = 0; return hip_fail(e, "hipGetDeviceCount"); }
  *n = c;
  return FF_OK;
}

int ff_domain_create(const ff_domain_desc* desc, ff_domain** out) {
  if (!desc || !out) return fail(FF_ERR_INVALID, "null argument");
  *out = nullptr;
  if (int rc = check_grid(desc)) return rc;
  const long long nz = desc->dim == 3 ? desc->nz : 1;
  const long long n = (long long)desc->nx * desc->ny * nz;
  if (!(desc->viscosity > 0.0)) return fail(FF_ERR_INVALID, "viscosity must be positive");
  if (!(desc->dt > 0.0)) return fail(FF_ERR_INVALID, "dt must be positive");
  if (desc->options & ~(FF_OPT_STAGE_FP16 | FF_OPT_PERIODIC | FF_OPT_FAST_MATH))
    return fail(FF_ERR_INVALID, "unknown option bits");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FF_ERR_NODEVICE, "no HIP device available");
  if (desc->device < 0 || desc->device >= ndev) return fail(FF_ERR_INVALID, "device ordinal out of range");
  FF_HIP(hipSetDevice(desc->device));
  hipDeviceProp_t prop;
  FF_HIP(hipGetDeviceProperties(&prop, desc->device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(FF_ERR_NODEVICE, std::string("library built for gfx950, device is ") + prop.gcnArchName);

  ff_domain* dd = new ff_domain();
  Domain& d = *dd;
  d.device = desc->device;
  d.dim = desc->dim;
  d.nx = desc->nx;
  d.ny = desc->ny;
  d.nz = (int)nz;
  d.n = n;
  d.nm = d.dim == 2 ? 6 : 10;
  d.q = d.dim == 2 ? 9 : 27;
  d.storage = desc->storage;
  d.options = desc->options;
  if (d.dim == 3) d.options &= ~FF_OPT_STAGE_FP16;  // the 3D tile is FP32 in the reference
  d.nu = desc->viscosity;
  d.dx = desc->dx;
  d.dt = desc->dt;
  d.rho_f = desc->fluid_density;
  // FsiCoupling2D.cu:368-371 and :236: float viscosity, omega from double arithmetic
  d.nu_f = (float)d.nu;
  d.omega = (float)(1.0 / ((double)d.nu_f * 3.0 + 0.5));
  d.cl = (float)d.dx;
  d.ct = (float)d.dt;
  // SimulationParameters2D/3D::getCf (Config.hpp:96-100, 154-158): cm is float
  const float cm = d.dim == 2 ? (float)(d.rho_f * d.dx * d.dx) : (float)(d.rho_f * d.dx * d.dx * d.dx);
  d.cf = (float)((double)cm * d.dx / (d.dt * d.dt));
  d.plane = plane_stride(n);
  d.h_flags.assign((size_t)n, kFluid);
  hipError_t e = acquire_stream(d.device, &d.stream);
  if (e != hipSuccess) { delete dd; return hip_fail(e, "hipStreamCreate"); }
  // one allocation: both moment buffers (cleared), the flags (Fluid) and the
  // reduction scratch
  Carve cv;
  unsigned char* mom_bytes[2] = {nullptr, nullptr};
  const size_t flags_off = plan_core(cv, d, &mom_bytes[0], &mom_bytes[1]);
  int rc = cv.commit(&d.core_mem);
  if (!rc) {
    d.mom[0] = mom_bytes[0];
    d.mom[1] = mom_bytes[1];
    unsigned char* base = (unsigned char*)d.core_mem;
    if ((e = hipMemsetAsync(base, 0, flags_off, d.stream)) != hipSuccess) rc = hip_fail(e, "hipMemset(moments)");
    else if ((e = hipMemsetAsync(d.flags, kFluid, (size_t)n, d.stream)) != hipSuccess) rc = hip_fail(e, "hipMemset(flags)");
    else if ((e = hipStreamSynchronize(d.stream)) != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
  }
  if (rc) { ff_domain_destroy(dd); return rc; }
  *out = dd;
  return FF_OK;
}

int ff_domain_destroy(ff_domain* dd) {
  if (!dd) return FF_OK;
  Domain& d = *dd;
  (void)hipSetDevice(d.device);
  (void)batch_detach(d);
  if (d.stream) (void)hipStreamSynchronize(d.stream);
  free_boundary(d);
  dfree(d.core_mem);  // moments, flags and reduction scratch
  d.mom[0] = d.mom[1] = nullptr;
  d.flags = nullptr;
  d.red = nullptr;
  dfree(d.outlets);
  release_stream(d.device, d.stream);
  delete dd;
  return FF_OK;
}

int ff_set_flags(ff_domain* dd, const uint8_t* flags) {
  FF_CHECK_DOMAIN(dd);
  if (!flags) return fail(FF_ERR_INVALID, "null flags");
  Domain& d = *dd;
  FF_HIP(hipSetDevice(d.device));
  std::memcpy(d.h_flags.data(), flags, (size_t)d.n);
  FF_HIP(hipStreamSynchronize(d.stream));
  batch_mark_dirty(d.batch);
  links_moved(d);
  FF_HIP(hipMemcpy(d.flags, flags, (size_t)d.n, hipMemcpyHostToDevice));
  std::vector<OutletCell> oc;
  build_outlets(d, oc);
  dfree(d.outlets);
  d.n_outlets = (int)oc.size();
  if (d.n_outlets) {
    int rc = dalloc(&d.outlets, oc.size());
    if (rc) return rc;
    FF_HIP(hipMemcpy(d.outlets, oc.data(), oc.size() * sizeof(OutletCell), hipMemcpyHostToDevice));
  }
  return FF_OK;
}

int ff_get_flags(ff_domain* dd, uint8_t* flags) {
  FF_CHECK_DOMAIN(dd);
  if (!flags) return fail(FF_ERR_INVALID, "null flags");
  Domain& d = *dd;
  FF_HIP(hipSetDevice(d.device));
  FF_HIP(hipStreamSynchronize(d.stream));
  FF_HIP(hipMemcpy(flags, d.flags, (size_t)d.n, hipMemcpyDeviceToHost));
  return FF_OK;
}

int ff_init_state(ff_domain* dd, const float* bc, int n_bc) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  FF_HIP(hipSetDevice(d.device));
  float b[6] = {0, 0, 0, 0, 0, 0};
  const int need = d.dim == 2 ? 4 : 6;
  if (bc && n_bc == need) for (int k = 0; k < need; k++) b[k] = bc[k];
  FF_HIP(hipStreamSynchronize(d.stream));
  d.cur = 0;
  batch_mark_dirty(d.batch);
  int rc = launch_init_state(d, b);
  if (rc) return rc;
  FF_HIP(hipStreamSynchronize(d.stream));
  return FF_OK;
}

int ff_set_moments(ff_domain* dd, const float* aos, int which) {
  FF_CHECK_DOMAIN(dd);
  if (!aos) return fail(FF_ERR_INVALID, "null moments");
  if (which < -1 || which > 1) return fail(FF_ERR_INVALID, "which must be -1, 0 or 1");
  FF_HIP(hipSetDevice(dd->device));
  FF_HIP(hipStreamSynchronize(dd->stream));
  return upload_aos(*dd, aos, which);
}

int ff_get_moments(ff_domain* dd, float* aos, int which) {
  FF_CHECK_DOMAIN(dd);
  if (!aos) return fail(FF_ERR_INVALID, "null output");
  if (which < 0 || which > 1) return fail(FF_ERR_INVALID, "which must be 0 or 1");
  Domain& d = *dd;
  FF_HIP(hipSetDevice(d.device));
  float* tmp = nullptr;
  const size_t cnt = (size_t)d.n * d.nm;
  int rc = dalloc(&tmp, cnt);
  if (rc) return rc;
  rc = launch_convert_out(d, d.mom[which == 0 ? d.cur : (d.cur ^ 1)], tmp);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(d.stream);
    if (e == hipSuccess) e = hipMemcpy(aos, tmp, cnt * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "download moments");
  }
  dfree(tmp);
  if (!rc) rc = check_overflow(d);
  return rc;
}

int ff_set_boundary(ff_domain* dd, int nv, const float* verts, const float* vels, int ne, const uint32_t* elems,
                    const int32_t* elem_body, int n_bodies) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  const int D = d.dim;
  if (nv < 0 || ne < 0) return fail(FF_ERR_INVALID, "negative sizes");
  if (ne > 65536) return fail(FF_ERR_INVALID, "at most 65536 boundary elements (16-bit element field of the link key)");
  if ((nv > 0 && (!verts || !vels)) || (ne > 0 && !elems)) return fail(FF_ERR_INVALID, "null boundary array");
  if (n_bodies <= 0) n_bodies = 1;
  for (long long k = 0; k < (long long)ne * D; k++)
    if (elems[k] >= (uint32_t)nv) return fail(FF_ERR_INVALID, "element references a vertex out of range");
  if (int rc = check_verts(D, d.dx, nv, verts)) return rc;
  std::vector<int> eb(ne, 0);
  if (elem_body)
    for (int k = 0; k < ne; k++) {
      if (elem_body[k] < 0 || elem_body[k] >= n_bodies) return fail(FF_ERR_INVALID, "elem_body out of range");
      eb[k] = elem_body[k];
    }
  FF_HIP(hipSetDevice(d.device));
  FF_HIP(hipStreamSynchronize(d.stream));
  batch_mark_dirty(d.batch);
  free_boundary(d);
  d.nv = nv;
  d.ne = ne;
  d.n_bodies = n_bodies;
  int rc = FF_OK;
  // host-side tables first, then one allocation, one upload and one clear
  // CSR body -> elements and vertex -> (element, end), in index order
  std::vector<int> bptr(n_bodies + 1, 0), bidx(ne);
  for (int k = 0; k < ne; k++) bptr[eb[k] + 1]++;
  for (int b = 0; b < n_bodies; b++) bptr[b + 1] += bptr[b];
  {
    std::vector<int> fill(bptr.begin(), bptr.end() - 1);
    for (int k = 0; k < ne; k++) bidx[fill[eb[k]]++] = k;
  }
  std::vector<int> vptr(nv + 1, 0), vinc((size_t)ne * D);
  for (long long k = 0; k < (long long)ne * D; k++) vptr[elems[k] + 1]++;
  for (int v = 0; v < nv; v++) vptr[v + 1] += vptr[v];
  {
    std::vector<int> fill(vptr.begin(), vptr.end() - 1);
    for (int k = 0; k < ne; k++)
      for (int j = 0; j < D; j++) vinc[fill[elems[(size_t)k * D + j]]++] = k * 4 + j;
  }
  // body reference: centroid of the body's vertices; vertex -> body
  std::vector<double> ref((size_t)n_bodies * D, 0.0);
  std::vector<int> vb(nv, -1);
  {
    for (int k = 0; k < ne; k++)
      for (int j = 0; j < D; j++) if (vb[elems[(size_t)k * D + j]] < 0) vb[elems[(size_t)k * D + j]] = eb[k];
    std::vector<int> cnt(n_bodies, 0);
    for (int v = 0; v < nv; v++)
      if (vb[v] >= 0) {
        cnt[vb[v]]++;
        for (int c = 0; c < D; c++) ref[(size_t)vb[v] * D + c] += verts[(size_t)v * D + c];
      }
    for (int b = 0; b < n_bodies; b++)
      if (cnt[b]) for (int c = 0; c < D; c++) ref[(size_t)b * D + c] /= cnt[b];
    for (int v = 0; v < nv; v++) if (vb[v] < 0) vb[v] = 0;
  }
  // link table capacity: distinct candidate cells near the mesh, with margin
  const unsigned long long est = link_estimate(D, d.dx, ne, verts, elems);
  d.lt_cap = (int)next_pow2(est * 2 + 1024);
  Carve cv;
  const BoundaryPlan bp = plan_boundary(cv, d, nv, ne, n_bodies, (size_t)d.lt_cap);
  const size_t init_end = bp.init_end, zero_end = bp.zero_end;
  if ((rc = cv.commit(&d.bnd_mem))) return rc;
  std::vector<unsigned char> host(init_end, 0);
  auto put = [&](size_t off, const void* src, size_t bytes) {
    if (bytes) std::memcpy(host.data() + off, src, bytes);
  };
  for (int b = 0; b < 2; b++) {
    put(bp.o_v[b], verts, (size_t)nv * D * sizeof(float));
    put(bp.o_u[b], vels, (size_t)nv * D * sizeof(float));
  }
  put(bp.o_el, elems, (size_t)ne * D * sizeof(uint32_t));
  put(bp.o_eb, eb.data(), (size_t)ne * sizeof(int));
  put(bp.o_vb, vb.data(), (size_t)nv * sizeof(int));
  put(bp.o_bp, bptr.data(), bptr.size() * sizeof(int));
  put(bp.o_bi, bidx.data(), bidx.size() * sizeof(int));
  put(bp.o_vp, vptr.data(), vptr.size() * sizeof(int));
  put(bp.o_vi, vinc.data(), vinc.size() * sizeof(int));
  put(bp.o_rf, ref.data(), ref.size() * sizeof(double));
  StepState st0{};
  st0.gen = 1;
  st0.tag = 1;
  put(bp.o_st, &st0, sizeof(st0));
  unsigned char* base = (unsigned char*)d.bnd_mem;
  FF_HIP(hipMemcpyAsync(base, host.data(), init_end, hipMemcpyHostToDevice, d.stream));
  FF_HIP(hipMemsetAsync(base + init_end, 0, zero_end - init_end, d.stream));
  FF_HIP(hipStreamSynchronize(d.stream));
  d.gen_host = 1;
  d.tag_host = 1;
  d.has_boundary = true;
  return FF_OK;
}

int ff_update_boundary(ff_domain* dd, const float* verts, const float* vels) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!verts || !vels) return fail(FF_ERR_INVALID, "null boundary array");
  if (int rc = check_verts(d.dim, d.dx, d.nv, verts)) return rc;
  FF_HIP(hipSetDevice(d.device));
  const size_t bytes = (size_t)d.nv * d.dim * sizeof(float);
  for (int b = 0; b < 2; b++) {
    FF_HIP(hipMemcpyAsync(d.vbuf[b], verts, bytes, hipMemcpyHostToDevice, d.stream));
    FF_HIP(hipMemcpyAsync(d.ubuf[b], vels, bytes, hipMemcpyHostToDevice, d.stream));
  }
  FF_HIP(hipStreamSynchronize(d.stream));
  if (d.wave) batch_mark_dirty(d.batch);
  d.wave = false;
  links_moved(d);
  return FF_OK;
}

int ff_set_body_ref(ff_domain* dd, const double* ref) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!ref) return fail(FF_ERR_INVALID, "null reference points");
  FF_HIP(hipSetDevice(d.device));
  FF_HIP(hipStreamSynchronize(d.stream));
  FF_HIP(hipMemcpy(d.body_ref, ref, (size_t)d.n_bodies * d.dim * sizeof(double), hipMemcpyHostToDevice));
  return FF_OK;
}

int ff_set_wave(ff_domain* dd, const float* rest, const float* s, const double* params) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!rest || !s || !params) return fail(FF_ERR_INVALID, "null wave array");
  FF_HIP(hipSetDevice(d.device));
  FF_HIP(hipStreamSynchronize(d.stream));
  int rc;
  // the arrays live in the boundary allocation (ff_set_boundary)
  FF_HIP(hipMemcpyAsync(d.rest_verts, rest, (size_t)d.nv * d.dim * sizeof(float), hipMemcpyHostToDevice, d.stream));
  FF_HIP(hipMemcpyAsync(d.s_coord, s, (size_t)d.nv * sizeof(float), hipMemcpyHostToDevice, d.stream));
  d.wave_host.resize((size_t)d.n_bodies * kWaveParStride);
  wave_device_params(params, d.dt, d.n_bodies, d.wave_host.data());
  FF_HIP(hipMemcpyAsync(d.wave_params, d.wave_host.data(), d.wave_host.size() * sizeof(double),
                        hipMemcpyHostToDevice, d.stream));
  if (!d.wave) batch_mark_dirty(d.batch);
  d.wave = true;
  links_moved(d);
  // kinematics of the next step to run; later steps are produced by the
  // finalize kernel of the step before
  if ((rc = launch_wave_now(d))) return rc;
  FF_HIP(hipStreamSynchronize(d.stream));
  return FF_OK;
}

int ff_set_wave_params(ff_domain* dd, const double* params) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.wave) return fail(FF_ERR_STATE, "no wave set (ff_set_wave)");
  if (!params) return fail(FF_ERR_INVALID, "null wave parameters");
  FF_HIP(hipSetDevice(d.device));
  // pageable source: the copy is staged before hipMemcpyAsync returns
  d.wave_host.resize((size_t)d.n_bodies * kWaveParStride);
  wave_device_params(params, d.dt, d.n_bodies, d.wave_host.data());
  FF_HIP(hipMemcpyAsync(d.wave_params, d.wave_host.data(), d.wave_host.size() * sizeof(double),
                        hipMemcpyHostToDevice, d.stream));
  links_moved(d);
  return launch_wave_now(d);
}

int ff_step(ff_domain* dd, int nsteps) {
  FF_CHECK_DOMAIN(dd);
  if (nsteps < 0) return fail(FF_ERR_INVALID, "negative step count");
  Domain& d = *dd;
  if (d.batch) return fail(FF_ERR_STATE, "domain belongs to a batch: step it with ff_batch_step");
  FF_HIP(hipSetDevice(d.device));
  for (int s = 0; s < nsteps; s++) {
    int rc = one_step(d, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  return FF_OK;
}

int ff_step_proxy(ff_domain* dd, const float* dv, const float* dvel, const uint32_t* de, float* dforces, int nv,
                  int ne) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "ff_set_boundary must describe the proxy topology first");
  if (nv != d.nv || ne != d.ne) return fail(FF_ERR_INVALID, "proxy sizes differ from the boundary topology");
  if (!dv || !dvel || !de) return fail(FF_ERR_INVALID, "null proxy pointer");
  if (d.batch) return fail(FF_ERR_STATE, "domain belongs to a batch: step it with ff_batch_step");
  FF_HIP(hipSetDevice(d.device));
  return one_step(d, dv, dvel, de, dforces);
}

int ff_get_vertex_forces(ff_domain* dd, float* out, int reset) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!out) return fail(FF_ERR_INVALID, "null output");
  FF_HIP(hipSetDevice(d.device));
  if (int frc = flush_finish(d)) return frc;  // the last step's deferred finish
  FF_HIP(hipStreamSynchronize(d.stream));
  const size_t bytes = (size_t)d.nv * d.dim * sizeof(float);
  FF_HIP(hipMemcpy(out, d.vforce, bytes, hipMemcpyDeviceToHost));
  if (reset) {
    FF_HIP(hipMemsetAsync(d.vforce, 0, bytes, d.stream));
    FF_HIP(hipStreamSynchronize(d.stream));
  }
  return check_overflow(d);
}

int ff_get_step_vertex_forces(ff_domain* dd, float* out) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!out) return fail(FF_ERR_INVALID, "null output");
  FF_HIP(hipSetDevice(d.device));
  const size_t cnt = (size_t)d.nv * d.dim;
  if (d.gen_host <= 1 || d.ne == 0) {  // no step has run on this boundary
    std::memset(out, 0, cnt * sizeof(float));
    return FF_OK;
  }
  float* tmp = nullptr;
  int rc = dalloc(&tmp, cnt);
  if (rc) return rc;
  rc = launch_step_vertex_forces(d, tmp);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(d.stream);
    if (e == hipSuccess) e = hipMemcpy(out, tmp, cnt * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "step vertex forces");
  }
  dfree(tmp);
  return rc;
}

// ---------------------------------------------------------------- no-allocation readback
// Staging of ff_get_step_vertex_forces_noalloc: a file-scope device array,
// one copy per GPU, held until the process ends.  Its name in host code is
// not a device address: the address is resolved with hipGetSymbolAddress,
// once per device with that device current, and only that pointer is handed
// to kernels and copies.  Calls on one GPU share the array; calls that
// overlap on one GPU (other threads or streams) race.
constexpr size_t kNoAllocFloats = size_t(1) << 20;  // 4 MiB: nv * dim <= 1 048 576
static __device__ __align__(256) float scratch[kNoAllocFloats];
constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxScratchDevices] = {};

// The calling thread's current device must be `device`.
static int scratch_for_device(int device, float** out) {
  if (device < 0 || device >= kMaxScratchDevices) return fail(FF_ERR_INVALID, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(FF_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[device] = (float*)p;
  }
  *out = g_scratch_ptr[device];
  return FF_OK;
}

int ff_get_step_vertex_forces_noalloc(ff_domain* dd, float* out) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!out) return fail(FF_ERR_INVALID, "null output");
  const size_t cnt = (size_t)d.nv * d.dim;
  if (cnt > kNoAllocFloats)
    return fail(FF_ERR_CAPACITY, "nv * dim = " + std::to_string(cnt) + " exceeds the no-allocation scratch of " +
                                     std::to_string(kNoAllocFloats) + " floats (use ff_get_step_vertex_forces)");
  FF_HIP(hipSetDevice(d.device));
  if (d.gen_host <= 1 || d.ne == 0) {  // no step has run on this boundary
    std::memset(out, 0, cnt * sizeof(float));
    return FF_OK;
  }
  float* ptr = nullptr;
  if (int rc = scratch_for_device(d.device, &ptr)) return rc;
  // every vertex's nv * dim values are written by the kernel before the copy
  if (int rc = launch_step_vertex_forces(d, ptr)) return rc;
  FF_HIP(hipMemcpyAsync(out, ptr, cnt * sizeof(float), hipMemcpyDeviceToHost, d.stream));
  FF_HIP(hipStreamSynchronize(d.stream));
  return FF_OK;
}

int ff_get_body_wrench(ff_domain* dd, double* out) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!out) return fail(FF_ERR_INVALID, "null output");
  FF_HIP(hipSetDevice(d.device));
  if (int frc = flush_finish(d)) return frc;  // the last step's deferred finish
  FF_HIP(hipStreamSynchronize(d.stream));
  FF_HIP(hipMemcpy(out, d.wrench, (size_t)d.n_bodies * wrench_width(d.dim) * sizeof(double), hipMemcpyDeviceToHost));
  return check_overflow(d);
}

int ff_set_wrench_history(ff_domain* dd, int capacity) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (capacity < 0) return fail(FF_ERR_INVALID, "negative capacity");
  FF_HIP(hipSetDevice(d.device));
  if (int frc = flush_finish(d)) return frc;  // the last step's deferred finish
  FF_HIP(hipStreamSynchronize(d.stream));
  batch_mark_dirty(d.batch);
  dfree(d.wrench_hist);
  dfree(d.power_hist);
  d.hist_cap = 0;
  d.hist_count = 0;
  if (capacity == 0) return FF_OK;
  int rc = dalloc(&d.wrench_hist, (size_t)capacity * d.n_bodies * wrench_width(d.dim));
  if (!rc) rc = dalloc(&d.power_hist, (size_t)capacity * d.n_bodies);
  if (rc) {
    dfree(d.wrench_hist);
    dfree(d.power_hist);
    return rc;
  }
  d.hist_cap = capacity;
  return FF_OK;
}

int ff_get_wrench_history(ff_domain* dd, double* out, int count, int* n_out) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!out || !n_out || count < 0) return fail(FF_ERR_INVALID, "bad output");
  FF_HIP(hipSetDevice(d.device));
  if (int frc = flush_finish(d)) return frc;  // the last step's deferred finish
  FF_HIP(hipStreamSynchronize(d.stream));
  const long long avail = std::min<long long>(d.hist_count, d.hist_cap);
  const int n = (int)std::min<long long>(avail, count);
  const size_t rec = (size_t)d.n_bodies * wrench_width(d.dim);
  std::vector<double> ring((size_t)d.hist_cap * rec);
  if (d.hist_cap) FF_HIP(hipMemcpy(ring.data(), d.wrench_hist, ring.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < n; k++) {
    // the step with generation g wrote slot (g - 1) % cap; the latest step is gen_host - 1
    const long long g = (long long)d.gen_host - n + k;
    std::memcpy(out + (size_t)k * rec, ring.data() + (size_t)((g - 1) % d.hist_cap) * rec, rec * sizeof(double));
  }
  *n_out = n;
  return FF_OK;
}

int ff_get_body_power(ff_domain* dd, double* out) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  if (!out) return fail(FF_ERR_INVALID, "null output");
  FF_HIP(hipSetDevice(d.device));
  if (int frc = flush_finish(d)) return frc;  // the last step's deferred finish
  FF_HIP(hipStreamSynchronize(d.stream));
  FF_HIP(hipMemcpy(out, d.power, (size_t)d.n_bodies * sizeof(double), hipMemcpyDeviceToHost));
  return check_overflow(d);
}

int ff_get_power_history(ff_domain* dd, double* out, int count, int* n_out) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!out || !n_out || count < 0) return fail(FF_ERR_INVALID, "bad output");
  FF_HIP(hipSetDevice(d.device));
  if (int frc = flush_finish(d)) return frc;  // the last step's deferred finish
  FF_HIP(hipStreamSynchronize(d.stream));
  const long long avail = std::min<long long>(d.hist_count, d.hist_cap);
  const int n = (int)std::min<long long>(avail, count);
  const size_t rec = (size_t)d.n_bodies;
  std::vector<double> ring((size_t)d.hist_cap * rec);
  if (d.hist_cap) FF_HIP(hipMemcpy(ring.data(), d.power_hist, ring.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < n; k++) {
    // the step with generation g wrote slot (g - 1) % cap; the latest step is gen_host - 1
    const long long g = (long long)d.gen_host - n + k;
    std::memcpy(out + (size_t)k * rec, ring.data() + (size_t)((g - 1) % d.hist_cap) * rec, rec * sizeof(double));
  }
  *n_out = n;
  return FF_OK;
}

int ff_get_boundary(ff_domain* dd, float* verts, float* vels) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  FF_HIP(hipSetDevice(d.device));
  FF_HIP(hipStreamSynchronize(d.stream));
  const size_t bytes = (size_t)d.nv * d.dim * sizeof(float);
  // the buffer of the most recent step (both buffers agree before the first)
  const int b = d.gen_host > 1 ? (int)((d.gen_host - 1) & 1) : (int)(d.gen_host & 1);
  if (verts) FF_HIP(hipMemcpy(verts, d.vbuf[b], bytes, hipMemcpyDeviceToHost));
  if (vels) FF_HIP(hipMemcpy(vels, d.ubuf[b], bytes, hipMemcpyDeviceToHost));
  return FF_OK;
}

int ff_mass(ff_domain* dd, double* out) {
  FF_CHECK_DOMAIN(dd);
  if (!out) return fail(FF_ERR_INVALID, "null output");
  Domain& d = *dd;
  FF_HIP(hipSetDevice(d.device));
  double* dev = nullptr;
  int rc = dalloc(&dev, 1);
  if (rc) return rc;
  rc = launch_mass(d, dev);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(d.stream);
    if (e == hipSuccess) e = hipMemcpy(out, dev, sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "mass");
  }
  dfree(dev);
  return rc;
}

int ff_fill_solid(ff_domain* dd) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (d.dim != 2) return fail(FF_ERR_INVALID, "fill_solid is 2D only");
  if (!d.has_boundary) return fail(FF_ERR_STATE, "no boundary set");
  FF_HIP(hipSetDevice(d.device));
  int rc = launch_fill_solid(d, d.vbuf[d.gen_host & 1]);
  if (rc) return rc;
  FF_HIP(hipStreamSynchronize(d.stream));
  FF_HIP(hipMemcpy(d.h_flags.data(), d.flags, (size_t)d.n, hipMemcpyDeviceToHost));
  return FF_OK;
}

int ff_sync(ff_domain* dd) {
  FF_CHECK_DOMAIN(dd);
  FF_HIP(hipSetDevice(dd->device));
  FF_HIP(hipStreamSynchronize(dd->stream));
  return check_overflow(*dd);
}

int ff_time_bulk_kernel(ff_domain* dd, int iters, double* avg_ms) {
  FF_CHECK_DOMAIN(dd);
  if (iters <= 0 || !avg_ms) return fail(FF_ERR_INVALID, "bad arguments");
  Domain& d = *dd;
  if (d.batch) return fail(FF_ERR_STATE, "domain belongs to a batch");
  FF_HIP(hipSetDevice(d.device));
  if (int frc = flush_finish(d)) return frc;  // the last step's deferred finish
  EventSet ev;
  ev.clears[0] = &d.tev0;
  ev.clears[1] = &d.tev1;
  FF_HIP(ev.create(2 * (size_t)iters));
  void* in = d.mom[d.cur];
  void* out = d.mom[d.cur ^ 1];
  // the stream launch of a step (with a boundary it also runs the boundary
  // cells and the next wave kinematics; all of it is idempotent within a step)
  const bool links = d.has_boundary && d.ne > 0;
  const LinkArgs la = links ? link_args(d, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], d.elems) : LinkArgs{};
  int rc = FF_OK;
  for (int k = 0; k < iters && !rc; k++) {
    d.tev0 = ev[2 * k];  // the launch's own start and end (launch_timed)
    d.tev1 = ev[2 * k + 1];
    if (links) rc = launch_step_stream(d, in, out, la, d.wave);
    else rc = d.dim == 2 ? launch_bulk2d(d, in, out) : launch_bulk3d(d, in, out);
  }
  d.tev0 = d.tev1 = nullptr;
  FF_HIP(hipGetLastError());
  FF_HIP(hipStreamSynchronize(d.stream));
  double tot = 0.0;
  for (int k = 0; k < iters; k++) {
    float ms = 0.f;
    FF_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
    tot += ms;
  }
  *avg_ms = tot / iters;
  return rc;
}

int ff_time_steps(ff_domain* dd, int nsteps, double* total_ms) {
  FF_CHECK_DOMAIN(dd);
  if (nsteps < 0 || !total_ms) return fail(FF_ERR_INVALID, "bad arguments");
  Domain& d = *dd;
  if (d.batch) return fail(FF_ERR_STATE, "domain belongs to a batch: time it with ff_batch_time_steps");
  FF_HIP(hipSetDevice(d.device));
  EventSet ev;
  FF_HIP(ev.create(2));
  FF_HIP(hipEventRecord(ev[0], d.stream));
  int rc = FF_OK;
  for (int s = 0; s < nsteps && !rc; s++) rc = one_step(d, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = flush_finish(d);  // the timed work includes the last step's finish
  FF_HIP(hipEventRecord(ev[1], d.stream));
  FF_HIP(hipEventSynchronize(ev[1]));
  float ms = 0.f;
  FF_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  *total_ms = ms;
  if (rc) return rc;
  return check_overflow(d);
}

int ff_step_timed(ff_domain* dd, int nsteps, double* total_ms, double* stream_ms) {
  FF_CHECK_DOMAIN(dd);
  if (nsteps < 0 || !total_ms || !stream_ms) return fail(FF_ERR_INVALID, "bad arguments");
  Domain& d = *dd;
  if (d.batch) return fail(FF_ERR_STATE, "domain belongs to a batch: time it with ff_batch_time_steps");
  FF_HIP(hipSetDevice(d.device));
  EventSet ev;
  ev.clears[0] = &d.tev0;
  ev.clears[1] = &d.tev1;
  FF_HIP(ev.create(2 * (size_t)nsteps + 2));
  const size_t ek = 2 * (size_t)nsteps;
  FF_HIP(hipEventRecord(ev[ek], d.stream));
  int rc = FF_OK;
  for (int s = 0; s < nsteps && !rc; s++) {
    d.tev0 = ev[2 * (size_t)s];  // the step's stream launch, its own start and end (launch_timed)
    d.tev1 = ev[2 * (size_t)s + 1];
    rc = one_step(d, nullptr, nullptr, nullptr, nullptr);
    d.tev0 = d.tev1 = nullptr;
  }
  if (!rc) rc = flush_finish(d);  // the timed work includes the last step's finish
  FF_HIP(hipEventRecord(ev[ek + 1], d.stream));
  FF_HIP(hipEventSynchronize(ev[ek + 1]));
  if (rc) return rc;
  float ms = 0.f;
  FF_HIP(hipEventElapsedTime(&ms, ev[ek], ev[ek + 1]));
  *total_ms = ms;
  double sum = 0.0;
  for (int s = 0; s < nsteps; s++) {
    float k = 0.f;
    FF_HIP(hipEventElapsedTime(&k, ev[2 * (size_t)s], ev[2 * (size_t)s + 1]));
    sum += k;
  }
  *stream_ms = sum;
  return check_overflow(d);
}

int ff_domain_footprint(const ff_domain_desc* desc, int nv, int ne, const float* verts, const uint32_t* elems,
                        unsigned long long* bytes) {
  if (!desc || !bytes) return fail(FF_ERR_INVALID, "null argument");
  if (int rc = check_grid(desc)) return rc;
  if (nv < 0 || ne < 0 || (ne > 0 && (!verts || !elems))) return fail(FF_ERR_INVALID, "bad boundary arguments");
  const long long n = (long long)desc->nx * desc->ny * (desc->dim == 3 ? desc->nz : 1);
  for (long long k = 0; k < (long long)ne * desc->dim; k++)
    if (elems[k] >= (uint32_t)nv) return fail(FF_ERR_INVALID, "element references a vertex out of range");
  if (ne > 0)
    if (int rc = check_verts(desc->dim, desc->dx, nv, verts)) return rc;
  const unsigned long long est = ne ? link_estimate(desc->dim, desc->dx, ne, verts, elems) : 0;
  *bytes = footprint(desc->dim, n, desc->storage, nv, ne, ne, est);
  return FF_OK;
}

int ff_set_tile(ff_domain* dd, int tile) {
  FF_CHECK_DOMAIN(dd);
  const int n = dd->dim == 2 ? kNumTiles2 : kNumTiles3;
  if (tile < 0 || tile >= n) return fail(FF_ERR_INVALID, "tile index out of range");
  if (dd->dim == 2) dd->tile2 = tile;
  else dd->tile3 = tile;
  batch_mark_dirty(dd->batch);
  return FF_OK;
}

int ff_autotune(ff_domain* dd, int* tile) {
  FF_CHECK_DOMAIN(dd);
  Domain& d = *dd;
  if (d.batch) return fail(FF_ERR_STATE, "domain belongs to a batch: its tile is tuned with the batch");
  FF_HIP(hipSetDevice(d.device));
  EventSet ev;
  FF_HIP(ev.create(2));
  const hipEvent_t e0 = ev[0], e1 = ev[1];
  // the stream-collide launch rewrites only cells every step rewrites, so
  // timing it leaves the state unchanged
  void* in = d.mom[d.cur];
  void* out = d.mom[d.cur ^ 1];
  int rc = FF_OK;
  int& cur = d.dim == 2 ? d.tile2 : d.tile3;
  const int n = d.dim == 2 ? kNumTiles2 : kNumTiles3;
  const int saved = cur;
  // an error return leaves the tile the domain had
  struct Restore {
    int& t;
    int v;
    bool keep = false;
    ~Restore() {
      if (!keep) t = v;
    }
  } restore{cur, saved};
  auto launch = [&]() { return d.dim == 2 ? launch_bulk2d(d, in, out) : launch_bulk3d(d, in, out); };
  // rounds over every candidate (clock and cache drift hit all alike); each
  // candidate keeps its fastest round
  std::vector<double> best(n, 1e30);
  for (int round = 0; round < 3 && !rc; round++) {
    for (int t = 0; t < n && !rc; t++) {
      cur = t;
      for (int w = 0; w < 3 && !rc; w++) rc = launch();
      if (rc) break;
      FF_HIP(hipEventRecord(e0, d.stream));
      for (int k = 0; k < 20 && !rc; k++) rc = launch();
      FF_HIP(hipEventRecord(e1, d.stream));
      FF_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FF_HIP(hipEventElapsedTime(&ms, e0, e1));
      best[t] = std::min(best[t], (double)ms);
    }
  }
  const int best_t = (int)(std::min_element(best.begin(), best.end()) - best.begin());
  cur = rc ? saved : best_t;
  restore.keep = true;
  if (tile) *tile = cur;
  return rc;
}

}  // extern "C"
