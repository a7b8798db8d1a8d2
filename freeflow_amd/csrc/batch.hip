// batch.hip -- many fluid domains of uneven size stepped together.
//
// A batch holds domains of one kind (dimension, storage format, options).
// One step of every member is one replay of a HIP graph with three kernel
// nodes (runs of 8+ steps replay an 8-step graph), the batched forms of the
// per-domain step (kernels2d/3d.hip, links.hip):
//
//   kb_step      boundary cells (into side-buffer rows), next-step wave
//                kinematics and the stream-collide tiles of every domain
//   kb_sums      per-element FP64 force sums, outlet cells, the
//                boundary-cell copy-back and the next step's link detection
//                of every domain
//   kb_finish    per-vertex forces; per domain one block for the body
//                wrench, the generation/tag advance and the step counter
//
// Members whose links were not detected ahead (new boundary, host update,
// new wave parameters) are detected before the replay by kb_detect, one
// launch for all of them.
//
// Each block finds its domain and its block index within that domain in a
// block map (int2 {domain, local block}) built on the host; per-domain
// arguments live in a device array of DomDev records.  The ping-pong parity
// of each domain is its own device step counter, so the graph is replayed
// unchanged every step; it is rebuilt only when membership or a member's
// buffers change.  The arithmetic per cell, link and element is that of the
// per-domain kernels, so a domain's results are the same alone or in a batch.
//
// The stream-collide tile shape is chosen per batch by timing the
// candidates of kTiles2 / kTiles3 on the batch itself (the bulk kernel only rewrites
// cells every step rewrites, so timing it leaves the state unchanged).

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ff_internal.h"
#include "geometry.h"
#include "links_body.h"
#include "stencil.h"

namespace ff {

struct DomDev {
  LinkArgs la;
  FinalArgs fa;
  void* mom[2];      // mom[0] = the domain's current buffer when the batch was built
  unsigned* step;    // steps since the batch was built (parity selects the buffer)
  int tiles_x, tiles_y;
  int nbc;           // boundary-cell blocks of this domain
  int nvb;           // vertex blocks of the finish phase
  int ndet;          // next-step detection blocks of the element-sum phase
  int boundary;
  int* bflag;        // the batch's overflow flag (set when this domain's table overflowed)
};

struct Batch {
  int dim = 2, storage = 0, device = 0;
  unsigned options = 0;
  hipStream_t stream = nullptr;
  std::vector<Domain*> mem;
  bool dirty = true;
  int tile = -1;        // fixed tile shape (index into kTiles2 / kTiles3), -1 = autotune
  int tuned_tile = -1;  // result of the last tile search
  double tuned_cells = 0.0;  // member cells when the search ran (a change by 2x searches again)
  DomDev* d_doms = nullptr;
  size_t doms_cap = 0;
  int2* d_map = nullptr;
  size_t map_cap = 0;
  unsigned* d_steps = nullptr;
  size_t steps_cap = 0;
  int n_bc = 0, n_wave = 0, n_bulk = 0, n_es = 0, n_fin = 0;
  int* d_flag = nullptr;  // set by a replay whose detection overflowed a member's link table
  // pending detections (members whose links were not detected ahead): pinned
  // staging and its device copy, reused or regrown only once the copy and the
  // launches that read it (pend_ev) have completed
  int2* h_pend = nullptr;
  size_t h_pend_cap = 0;
  int2* d_pend = nullptr;
  size_t d_pend_cap = 0;
  hipEvent_t pend_ev = nullptr;
  std::vector<int2> bumps, dets;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // kGraphSteps steps in one graph: long runs pay one graph launch per
  // kGraphSteps steps
  hipGraph_t graph_n = nullptr;
  hipGraphExec_t exec_n = nullptr;
  hipEvent_t tev0 = nullptr, tev1 = nullptr;  // set only while ff_batch_time_bulk times a launch
  // pinned staging of the DomDev records and block maps (reused once the
  // copies recorded by tab_ev have completed)
  unsigned char* h_tab = nullptr;
  size_t h_tab_cap = 0;
  hipEvent_t tab_ev = nullptr;
  int direct = 0;  // steps launched without a graph since the last rebuild
  // Generation of the device tables (d_doms, d_map, d_steps): build_tables
  // bumps it; a graph records the generation it was captured against and is
  // never launched against another (replay checks, drop_graph forgets).
  unsigned long long tab_gen = 0;
  unsigned long long exec_gen = 0, exec_n_gen = 0;
};

// A rebuilt batch launches its first kDirectSteps steps directly and
// captures a graph only after that: a membership that changes every few
// steps (a server admitting and retiring episodes) skips the capture.
constexpr int kDirectSteps = 4;

constexpr int kGraphSteps = 8;

// ---------------------------------------------------------------- kernels
// A pointer read from the domain table lies in device global memory, but
// being read from memory (not passed as a kernel argument) the compiler
// cannot tell, and addresses it generically (flat instructions, which also
// count against the LDS wait counter).  Passed through an integer into the
// global address space, every access derived from it is a global one.
template <typename T>
__device__ __forceinline__ T* dev_global(T* p) {
  return (T*)(__attribute__((address_space(1))) T*)reinterpret_cast<unsigned long long>(p);
}

// Stream launch: [boundary-cell blocks of every domain][wave blocks][tiles].
template <typename S, bool STAGE, bool PERIODIC, int TX, int TY, int CPT, bool FAST>
__global__ void __launch_bounds__(256) kb_step2d(const DomDev* __restrict__ doms, const int2* __restrict__ map,
                                                 int n_bc, int n_wave) {
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;  // padding of the XCD-aware tile order
  const DomDev* D = doms + m.x;
  const unsigned p = *D->step & 1u;
  const int b = blockIdx.x;
  if (b < n_bc) {
    bcells_body<2, S, STAGE, PERIODIC, 256, FAST>((const S*)D->mom[p], D->la, m.y, D->nbc);
  } else if (b < n_bc + n_wave) {
    wave_block_body<2, 256, FAST>(D->la, D->fa, m.y);
  } else {
    bulk2d_tile<S, STAGE, PERIODIC, TX, TY, CPT, FAST>(dev_global((const S*)D->mom[p]), dev_global((S*)D->mom[p ^ 1u]),
                                                       dev_global(D->la.flags), D->la.nx, D->la.ny, D->la.plane,
                                                       D->la.omega, m.y % D->tiles_x, m.y / D->tiles_x);
  }
}

// RING = 0: TX x TY x ZC box tiles; otherwise TX x TY columns marching ZC planes; NT = block size.
template <typename S, bool PERIODIC, bool FAST, int KIND, int TX, int TY, int ZC, int NT, int PF, bool DMA>
__global__ void __launch_bounds__(NT) kb_step3d(const DomDev* __restrict__ doms, const int2* __restrict__ map,
                                                int n_bc, int n_wave) {
  static_assert(KIND != 2 || NT == ColTile<TX, TY>::NT, "block size of a staged column");
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;  // padding of the XCD-aware tile order
  const DomDev* D = doms + m.x;
  const unsigned p = *D->step & 1u;
  const int b = blockIdx.x;
  if (b < n_bc) {
    bcells_body<3, S, false, PERIODIC, NT, FAST>((const S*)D->mom[p], D->la, m.y, D->nbc);
  } else if (b < n_bc + n_wave) {
    wave_block_body<3, NT, FAST>(D->la, D->fa, m.y);
  } else {
    const int tx = D->tiles_x, ty = D->tiles_y;
    const int bx = m.y % tx, by = (m.y / tx) % ty, bz = m.y / (tx * ty);
    const S* in = dev_global((const S*)D->mom[p]);
    S* out = dev_global((S*)D->mom[p ^ 1u]);
    const LinkArgs& la = D->la;
    const uint8_t* flags = dev_global(la.flags);
    if constexpr (KIND == 2)
      bulk3d_col<S, PERIODIC, FAST, TX, TY, PF, DMA>(in, out, flags, la.nx, la.ny, la.nz, la.plane, la.omega, bx,
                                                     by, bz * ZC, min(bz * ZC + ZC, la.nz));
    else if constexpr (KIND == 1)
      bulk3d_ring<S, PERIODIC, TX, TY, ZC, FAST>(in, out, flags, la.nx, la.ny, la.nz, la.plane, la.omega, bx, by,
                                                 bz);
    else
      bulk3d_tile<S, PERIODIC, TX, TY, ZC, FAST>(in, out, flags, la.nx, la.ny, la.nz, la.plane, la.omega, bx, by,
                                                 bz);
  }
}

template <int DIM, typename S, int FB>
__global__ void __launch_bounds__(FB) kb_sums(const DomDev* __restrict__ doms, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const DomDev* D = doms + m.x;
  elem_sums_body<DIM, S, FB, true>(D->la, D->fa, m.y, D->mom[(*D->step & 1u) ^ 1u]);
}

// Per domain: vertex blocks, then one block that folds the wrench, advances
// gen/tag and the domain's step counter.
template <int DIM, int FB>
__global__ void __launch_bounds__(FB) kb_finish(const DomDev* __restrict__ doms, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const DomDev* D = doms + m.x;
  if (m.y < D->nvb) {
    finish_vertex_body<DIM, FB>(D->la, D->fa, m.y);
    return;
  }
  if (D->boundary) {
    finish_last_body<DIM, FB>(D->la, D->fa, D->la.st->gen, true);
    if (threadIdx.x == 0 && D->la.st->overflow) *D->bflag = 1;
  }
  if (threadIdx.x == 0) *D->step = *D->step + 1u;
}

// Pending detections: tag bumps of members whose earlier detection went
// stale, then the detection blocks of every pending member (one launch for
// the whole batch instead of one per domain).
__global__ void kb_tag_bump(const DomDev* __restrict__ doms, const int2* __restrict__ list, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    StepState* st = doms[list[i].x].la.st;
    st->tag = st->tag + 1u;
  }
}

template <int DIM>
__global__ void __launch_bounds__(kSumBlock) kb_detect(const DomDev* __restrict__ doms, const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const LinkArgs& la = doms[m.x].la;
  detect_block<DIM>(la, (long long)m.y, step_tag(la), la.v[step_gen(la) & 1u]);
}

// ---------------------------------------------------------------- launch helpers
// Block size of the stream launch for a tile index.
static int stream_threads(int dim, int tile) { return dim == 2 ? 256 : kTiles3[tile].nt; }

template <typename S, bool STAGE, bool PERIODIC, int TX, int TY, int CPT, bool FAST>
static void step2d_batch_shape(Batch& b, const int2* map, int nblocks, int n_bc, int n_wave) {
  launch_timed(kb_step2d<S, STAGE, PERIODIC, TX, TY, CPT, FAST>, dim3(nblocks), dim3(256), b.stream, b.tev0, b.tev1,
               (const DomDev*)b.d_doms, map, n_bc, n_wave);
}

template <typename S, bool STAGE, bool PERIODIC, bool FAST>
static void step2d_batch_variant(Batch& b, int tile, const int2* map, int nblocks, int n_bc, int n_wave) {
  switch (tile) {
    case 1: step2d_batch_shape<S, STAGE, PERIODIC, 32, 8, 2, FAST>(b, map, nblocks, n_bc, n_wave); break;
    case 2: step2d_batch_shape<S, STAGE, PERIODIC, 64, 4, 2, FAST>(b, map, nblocks, n_bc, n_wave); break;
    case 3: step2d_batch_shape<S, STAGE, PERIODIC, 32, 8, 4, FAST>(b, map, nblocks, n_bc, n_wave); break;
    case 4: step2d_batch_shape<S, STAGE, PERIODIC, 64, 4, 4, FAST>(b, map, nblocks, n_bc, n_wave); break;
    default: step2d_batch_shape<S, STAGE, PERIODIC, 32, 8, 1, FAST>(b, map, nblocks, n_bc, n_wave); break;
  }
}

template <typename S, bool PERIODIC, bool FAST, int KIND, int TX, int TY, int ZC, int NT, int PF, bool DMA>
static void step3d_batch_shape(Batch& b, const int2* map, int nblocks, int n_bc, int n_wave) {
  launch_timed(kb_step3d<S, PERIODIC, FAST, KIND, TX, TY, ZC, NT, PF, DMA>, dim3(nblocks), dim3(NT), b.stream, b.tev0,
               b.tev1, (const DomDev*)b.d_doms, map, n_bc, n_wave);
}

template <typename S, bool PERIODIC, bool FAST>
static void step3d_batch_variant(Batch& b, int tile, const int2* map, int nblocks, int n_bc, int n_wave) {
#define FF_SB3(FA, K, X, Y, Z, N, P, DM) \
  step3d_batch_shape<S, PERIODIC, FA, K, X, Y, Z, N, P, DM>(b, map, nblocks, n_bc, n_wave)
  FF_TILES3(FF_SB3, FAST, tile)
#undef FF_SB3
}

// The stream launch over `map` (n_bc boundary-cell blocks, n_wave wave blocks, then tiles).
template <typename S>
static void launch_stream_batch(Batch& b, int tile, const int2* map, int nblocks, int n_bc, int n_wave) {
  if (nblocks <= 0) return;
  const bool stage = b.options & FF_OPT_STAGE_FP16, per = b.options & FF_OPT_PERIODIC;
  const bool fast = b.options & FF_OPT_FAST_MATH;
  if (b.dim == 2) {
    // Fast math applies to unstaged tiles only.
    if (stage && per) step2d_batch_variant<S, true, true, false>(b, tile, map, nblocks, n_bc, n_wave);
    else if (stage) step2d_batch_variant<S, true, false, false>(b, tile, map, nblocks, n_bc, n_wave);
    else if (per && fast) step2d_batch_variant<S, false, true, true>(b, tile, map, nblocks, n_bc, n_wave);
    else if (per) step2d_batch_variant<S, false, true, false>(b, tile, map, nblocks, n_bc, n_wave);
    else if (fast) step2d_batch_variant<S, false, false, true>(b, tile, map, nblocks, n_bc, n_wave);
    else step2d_batch_variant<S, false, false, false>(b, tile, map, nblocks, n_bc, n_wave);
  } else {
    if (per && fast) step3d_batch_variant<S, true, true>(b, tile, map, nblocks, n_bc, n_wave);
    else if (per) step3d_batch_variant<S, true, false>(b, tile, map, nblocks, n_bc, n_wave);
    else if (fast) step3d_batch_variant<S, false, true>(b, tile, map, nblocks, n_bc, n_wave);
    else step3d_batch_variant<S, false, false>(b, tile, map, nblocks, n_bc, n_wave);
  }
}

template <int DIM, typename S>
static void launch_tail(Batch& b, const int2* map_sums, const int2* map_fin) {
  constexpr int FB = kSumBlock;
  if (b.n_es > 0) kb_sums<DIM, S, FB><<<b.n_es, FB, 0, b.stream>>>(b.d_doms, map_sums);
  if (b.n_fin > 0) kb_finish<DIM, 256><<<b.n_fin, 256, 0, b.stream>>>(b.d_doms, map_fin);
}

template <typename S>
static void launch_step_typed(Batch& b, int tile) {
  const int n_stream = b.n_bc + b.n_wave + b.n_bulk;
  const int2* map_stream = b.d_map;
  const int2* map_sums = map_stream + n_stream;
  const int2* map_fin = map_sums + b.n_es;
  launch_stream_batch<S>(b, tile, map_stream, n_stream, b.n_bc, b.n_wave);
  if (b.dim == 2) launch_tail<2, S>(b, map_sums, map_fin);
  else launch_tail<3, S>(b, map_sums, map_fin);
}

static void launch_step(Batch& b, int tile) {
  if (b.storage == FF_STORAGE_F32) launch_step_typed<float>(b, tile);
  else if (b.storage == FF_STORAGE_F16) launch_step_typed<__half>(b, tile);
  else if (b.storage == FF_STORAGE_F16U) launch_step_typed<h16u>(b, tile);
  else launch_step_typed<__hip_bfloat16>(b, tile);
}

// Tiles only (the map's tile range), for timing.
static void launch_bulk_only(Batch& b, int tile, const int2* map, int nblocks) {
  if (b.storage == FF_STORAGE_F32) launch_stream_batch<float>(b, tile, map, nblocks, 0, 0);
  else if (b.storage == FF_STORAGE_F16) launch_stream_batch<__half>(b, tile, map, nblocks, 0, 0);
  else if (b.storage == FF_STORAGE_F16U) launch_stream_batch<h16u>(b, tile, map, nblocks, 0, 0);
  else launch_stream_batch<__hip_bfloat16>(b, tile, map, nblocks, 0, 0);
}

// ---------------------------------------------------------------- host side
template <class T>
static int grow(T** p, size_t* cap, size_t need) {
  if (need <= *cap && *p) return FF_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  size_t n = std::max<size_t>(need, 16);
  n += n / 2;
  FF_HIP(hipMalloc((void**)p, n * sizeof(T)));
  *cap = n;
  return FF_OK;
}

static void tiles_of(const Batch& b, const Domain& d, int tile, int* tx, int* ty, int* nt) {
  if (b.dim == 2) {
    const Tile2 t = kTiles2[tile];
    *tx = (d.nx + t.tx - 1) / t.tx;
    *ty = (d.ny + t.ty * t.cpt - 1) / (t.ty * t.cpt);
    *nt = *tx * *ty;
  } else {
    const Tile3 t = kTiles3[tile];
    *tx = (d.nx + t.tx - 1) / t.tx;
    *ty = (d.ny + t.ty - 1) / t.ty;
    *nt = *tx * *ty * ((d.nz + t.tz - 1) / t.tz);
  }
}

// Records and block maps for the current members and `tile`; uploads them.
// Map layout: stream launch [boundary cells][wave][tiles], sums launch,
// finish launch.
static int build_tables(Batch& b, int tile) {
  const int n = (int)b.mem.size();
  std::vector<DomDev> doms(n);
  std::vector<int2> bc, wave, bulk, es, fin;
  const int K = b.dim == 2 ? 8 : 26;
  const int fb = kSumBlock;
  const int nt = stream_threads(b.dim, tile);
  int rc = grow(&b.d_steps, &b.steps_cap, (size_t)std::max(n, 1));
  if (rc) return rc;
  for (int i = 0; i < n; i++) {
    Domain& d = *b.mem[i];
    DomDev& r = doms[i];
    std::memset(&r, 0, sizeof(r));
    const bool bnd = d.has_boundary && d.ne > 0;
    r.la = link_args(d, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], d.elems);
    r.la.dev_state = 1;  // graph replays advance gen and tag on the device
    r.fa = make_final_args(d, nullptr, d.vbuf[0], d.vbuf[1], d.ubuf[0], d.ubuf[1], nullptr);
    r.mom[0] = d.mom[d.cur];
    r.mom[1] = d.mom[d.cur ^ 1];
    r.step = b.d_steps + i;
    r.boundary = bnd ? 1 : 0;
    r.bflag = b.d_flag;
    int tx, ty, ntl;
    tiles_of(b, d, tile, &tx, &ty, &ntl);
    r.tiles_x = tx;
    r.tiles_y = ty;
    for (int t = 0; t < ntl; t++) bulk.push_back(make_int2(i, t));
    if (bnd) {
      r.la.bc_span = nt;  // whole blocks per pass: many domains fill the chip
      r.nbc = boundary_cell_blocks(d, nt);
      for (int t = 0; t < r.nbc; t++) bc.push_back(make_int2(i, t));
      if (d.wave) {
        const int nw = (d.nv + nt - 1) / nt;
        for (int t = 0; t < nw; t++) wave.push_back(make_int2(i, t));
      }
      r.nvb = (d.nv + 255) / 256;
      r.ndet = detect_blocks(d, kSumBlock, true);
      r.fa.copy_blocks = copyback_blocks(d, kSumBlock, 4);  // fewer blocks: many domains share the launch
    } else {
      r.la.ne = 0;
      r.fa.copy_blocks = 0;
    }
    const int nes = (bnd ? elem_blocks(d.ne) : 0) + r.fa.outlet_blocks + r.fa.copy_blocks + r.ndet;
    for (int t = 0; t < nes; t++) es.push_back(make_int2(i, t));
    for (int t = 0; t <= r.nvb; t++) fin.push_back(make_int2(i, t));
  }
  // XCD-aware order of the tiles (see xcd_tile): the tile range starts at a
  // multiple of 8 blocks (the wave range is padded) and block p of the range
  // takes tile (p % 8) * ceil(n / 8) + p / 8 of the concatenated list.
  while ((bc.size() + wave.size()) % kXcds) wave.push_back(make_int2(-1, -1));
  {
    const size_t nt = bulk.size(), per = (nt + kXcds - 1) / kXcds;
    std::vector<int2> order(per * kXcds, make_int2(-1, -1));
    for (size_t p = 0; p < order.size(); p++) {
      const size_t t = (p % kXcds) * per + p / kXcds;
      if (t < nt) order[p] = bulk[t];
    }
    bulk.swap(order);
  }
  b.n_bc = (int)bc.size();
  b.n_wave = (int)wave.size();
  b.n_bulk = (int)bulk.size();
  b.n_es = (int)es.size();
  b.n_fin = (int)fin.size();
  (void)fb;
  std::vector<int2> map;
  map.reserve(bc.size() + wave.size() + bulk.size() + es.size() + fin.size());
  for (auto* v : {&bc, &wave, &bulk, &es, &fin}) map.insert(map.end(), v->begin(), v->end());
  if ((size_t)std::max(n, 1) > b.doms_cap || std::max<size_t>(map.size(), 1) > b.map_cap)
    FF_HIP(hipStreamSynchronize(b.stream));  // the old tables may still be read
  if ((rc = grow(&b.d_doms, &b.doms_cap, (size_t)std::max(n, 1)))) return rc;
  if ((rc = grow(&b.d_map, &b.map_cap, std::max<size_t>(map.size(), 1)))) return rc;
  // stage both tables in pinned memory and copy them on the batch stream
  // (ordered before every later launch, no host wait)
  const size_t nd = (size_t)n * sizeof(DomDev), nm = map.size() * sizeof(int2);
  if (!b.tab_ev) FF_HIP(hipEventCreateWithFlags(&b.tab_ev, hipEventDisableTiming));
  FF_HIP(hipEventSynchronize(b.tab_ev));  // the last copies out of h_tab have completed
  if (nd + nm > b.h_tab_cap) {
    if (b.h_tab) (void)hipHostFree(b.h_tab);
    b.h_tab = nullptr;
    b.h_tab_cap = 0;
    const size_t cap = nd + nm + (nd + nm) / 2 + 4096;
    FF_HIP(hipHostMalloc((void**)&b.h_tab, cap, hipHostMallocDefault));
    b.h_tab_cap = cap;
  }
  if (nd) std::memcpy(b.h_tab, doms.data(), nd);
  if (nm) std::memcpy(b.h_tab + nd, map.data(), nm);
  if (nd) FF_HIP(hipMemcpyAsync(b.d_doms, b.h_tab, nd, hipMemcpyHostToDevice, b.stream));
  if (nm) FF_HIP(hipMemcpyAsync(b.d_map, b.h_tab + nd, nm, hipMemcpyHostToDevice, b.stream));
  FF_HIP(hipMemsetAsync(b.d_steps, 0, std::max(n, 1) * sizeof(unsigned), b.stream));
  FF_HIP(hipEventRecord(b.tab_ev, b.stream));
  b.tab_gen++;
  return FF_OK;
}

static double member_cells(const Batch& b) {
  double c = 0.0;
  for (const Domain* d : b.mem) c += (double)d->nx * d->ny * d->nz;
  return c;
}

// Candidate shapes of the batch's tile search: every 2D shape; in 3D the box
// and the register-prefetch staged columns (the ring columns and the LDS-DMA
// columns were never the fastest on a batch, and every candidate costs a
// table build and a few launches of the whole batch while a server waits).
constexpr int kBatchCand3[] = {0, 7, 8, 9, 14, 15};

// Time the stream-collide tiles of every member for each candidate shape
// and keep the fastest.
static int autotune(Batch& b) {
  if (b.mem.empty()) {
    b.tuned_tile = 0;
    return FF_OK;
  }
  const int ncand = b.dim == 2 ? kNumTiles2 : (int)(sizeof(kBatchCand3) / sizeof(kBatchCand3[0]));
  EventSet ev;
  FF_HIP(ev.create(2));
  const hipEvent_t e0 = ev[0], e1 = ev[1];
  double best = 1e30;
  int best_t = 0;
  int rc = FF_OK;
  for (int c = 0; c < ncand && !rc; c++) {
    const int t = b.dim == 2 ? c : kBatchCand3[c];
    if ((rc = build_tables(b, t))) break;
    const int2* map = b.d_map + b.n_bc + b.n_wave;
    // long launches need fewer repetitions for a stable time
    const bool big = b.n_bulk > 20000;
    const int warm = big ? 1 : 3, iters = big ? 3 : 10;
    for (int w = 0; w < warm; w++) launch_bulk_only(b, t, map, b.n_bulk);
    FF_HIP(hipEventRecord(e0, b.stream));
    for (int k = 0; k < iters; k++) launch_bulk_only(b, t, map, b.n_bulk);
    FF_HIP(hipEventRecord(e1, b.stream));
    FF_HIP(hipEventSynchronize(e1));
    FF_HIP(hipGetLastError());
    float ms = 0.f;
    FF_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      best_t = t;
    }
  }
  b.tuned_tile = best_t;
  b.tuned_cells = member_cells(b);
  return rc;
}

static void drop_graph(Batch& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  if (b.exec_n) (void)hipGraphExecDestroy(b.exec_n);
  if (b.graph_n) (void)hipGraphDestroy(b.graph_n);
  b.exec = nullptr;
  b.graph = nullptr;
  b.exec_n = nullptr;
  b.graph_n = nullptr;
  b.exec_gen = b.exec_n_gen = 0;
}

static int current_tile(const Batch& b) { return b.tile >= 0 ? b.tile : std::max(b.tuned_tile, 0); }

// Captures `steps` batched steps into one instantiated graph.
static int capture_steps(Batch& b, int steps, hipGraph_t* graph, hipGraphExec_t* exec) {
  FF_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < steps; k++) launch_step(b, current_tile(b));
  hipError_t le = hipGetLastError();
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(b.stream, &g);
  if (le != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(le, "batched step launch");
  }
  if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
  *graph = g;
  FF_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return FF_OK;
}

// Runs `nsteps` steps: directly while the batch has taken fewer than
// kDirectSteps steps since its rebuild, otherwise from graphs captured on
// demand: kGraphSteps at a time (captured on the first long run), then one
// at a time.
static int replay(Batch& b, int nsteps) {
  int s = 0;
  if (!b.exec) {
    if (b.direct + nsteps <= kDirectSteps) {
      for (; s < nsteps; s++) launch_step(b, current_tile(b));
      FF_HIP(hipGetLastError());
      b.direct += nsteps;
      return FF_OK;
    }
    const int rc = capture_steps(b, 1, &b.graph, &b.exec);
    if (rc) return rc;
    b.exec_gen = b.tab_gen;
  }
  if (nsteps >= kGraphSteps && !b.exec_n) {
    const int rc = capture_steps(b, kGraphSteps, &b.graph_n, &b.exec_n);
    if (rc) return rc;
    b.exec_n_gen = b.tab_gen;
  }
  // a graph's kernel arguments point into the tables it was captured with
  if (b.exec_gen != b.tab_gen || (b.exec_n && b.exec_n_gen != b.tab_gen))
    return fail(FF_ERR_STATE, "internal: batch graph captured against replaced tables");
  for (; s + kGraphSteps <= nsteps; s += kGraphSteps) FF_HIP(hipGraphLaunch(b.exec_n, b.stream));
  for (; s < nsteps; s++) FF_HIP(hipGraphLaunch(b.exec, b.stream));
  return FF_OK;
}


static int rebuild(Batch& b) {
  FF_HIP(hipSetDevice(b.device));
  FF_HIP(hipStreamSynchronize(b.stream));
  drop_graph(b);
  int rc;
  // the best tile depends on how many blocks the launch holds: search again
  // when the member cells changed by more than 2x since the last search
  const double cells = member_cells(b);
  const bool retune = b.tuned_tile < 0 || cells > 2.0 * b.tuned_cells || 2.0 * cells < b.tuned_cells;
  if (b.tile < 0 && retune && (rc = autotune(b))) return rc;
  if ((rc = build_tables(b, current_tile(b)))) return rc;
  b.direct = 0;  // the graphs are captured on demand (replay)
  b.dirty = false;
  return FF_OK;
}

void batch_mark_dirty(Batch* b) {
  if (b) b->dirty = true;
}

static int detach(Batch& b, Domain& d) {
  auto it = std::find(b.mem.begin(), b.mem.end(), &d);
  if (it == b.mem.end()) return fail(FF_ERR_INVALID, "domain is not in this batch");
  FF_HIP(hipSetDevice(b.device));
  FF_HIP(hipStreamSynchronize(b.stream));
  b.mem.erase(it);
  d.stream = d.own_stream;
  d.own_stream = nullptr;
  d.batch = nullptr;
  b.dirty = true;
  return FF_OK;
}

int batch_detach(Domain& d) { return d.batch ? detach(*d.batch, d) : FF_OK; }

// Members whose links were not detected ahead (new boundary, or a boundary
// moved by the host since the last step) detect them before the replay, all
// in one launch.  The block maps index b.mem in the order of the last
// build_tables (the caller rebuilds a dirty batch first).
static int detect_pending(Batch& b) {
  b.bumps.clear();
  b.dets.clear();
  for (size_t i = 0; i < b.mem.size(); i++) {
    Domain& d = *b.mem[i];
    if (!(d.has_boundary && d.ne > 0) || d.predetected) continue;
    if (d.stale) {
      b.bumps.push_back(make_int2((int)i, 0));
      d.tag_host++;  // mirror of the device bump
    }
    const int nd = detect_blocks(d, kSumBlock, true);
    for (int t = 0; t < nd; t++) b.dets.push_back(make_int2((int)i, t));
    d.stale = false;
    d.predetected = true;
  }
  const size_t nb = b.bumps.size(), nd = b.dets.size(), n = nb + nd;
  if (n == 0) return FF_OK;
  if (!b.pend_ev) FF_HIP(hipEventCreateWithFlags(&b.pend_ev, hipEventDisableTiming));
  FF_HIP(hipEventSynchronize(b.pend_ev));  // the last use of h_pend and d_pend has completed
  if (n > b.h_pend_cap) {
    if (b.h_pend) (void)hipHostFree(b.h_pend);
    b.h_pend = nullptr;
    b.h_pend_cap = 0;
    const size_t cap = n + n / 2 + 256;
    FF_HIP(hipHostMalloc((void**)&b.h_pend, cap * sizeof(int2), hipHostMallocDefault));
    b.h_pend_cap = cap;
  }
  int rc = grow(&b.d_pend, &b.d_pend_cap, n);
  if (rc) return rc;
  std::memcpy(b.h_pend, b.bumps.data(), nb * sizeof(int2));
  std::memcpy(b.h_pend + nb, b.dets.data(), nd * sizeof(int2));
  FF_HIP(hipMemcpyAsync(b.d_pend, b.h_pend, n * sizeof(int2), hipMemcpyHostToDevice, b.stream));
  if (nb) kb_tag_bump<<<(unsigned)((nb + 255) / 256), 256, 0, b.stream>>>(b.d_doms, b.d_pend, (int)nb);
  if (nd) {
    if (b.dim == 2) kb_detect<2><<<(unsigned)nd, kSumBlock, 0, b.stream>>>(b.d_doms, b.d_pend + nb);
    else kb_detect<3><<<(unsigned)nd, kSumBlock, 0, b.stream>>>(b.d_doms, b.d_pend + nb);
  }
  FF_HIP(hipGetLastError());
  FF_HIP(hipEventRecord(b.pend_ev, b.stream));
  return FF_OK;
}

// Host mirrors of the device-side step state after `nsteps` replays.
static void advance_mirrors(Batch& b, int nsteps) {
  for (Domain* dp : b.mem) {
    Domain& d = *dp;
    if (nsteps & 1) d.cur ^= 1;
    if (d.has_boundary && d.ne > 0) {
      d.gen_host += (unsigned)nsteps;
      d.tag_host += (unsigned)nsteps;
      if (d.hist_cap > 0) d.hist_count += nsteps;
      d.predetected = true;
    }
  }
}

}  // namespace ff

using namespace ff;

struct ff_batch : public ff::Batch {};

#define FF_CHECK_BATCH(b) \
  if (!(b)) return fail(FF_ERR_INVALID, "null batch")

extern "C" {

int ff_batch_create(int dim, int storage, unsigned options, int device, ff_batch** out) {
  if (!out) return fail(FF_ERR_INVALID, "null output");
  *out = nullptr;
  if (dim != 2 && dim != 3) return fail(FF_ERR_INVALID, "dim must be 2 or 3");
  if (storage < 0 || storage > 3) return fail(FF_ERR_INVALID, "unknown storage format");
  if (options & ~(FF_OPT_STAGE_FP16 | FF_OPT_PERIODIC | FF_OPT_FAST_MATH))
    return fail(FF_ERR_INVALID, "unknown option bits");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FF_ERR_NODEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(FF_ERR_INVALID, "device ordinal out of range");
  FF_HIP(hipSetDevice(device));
  ff_batch* b = new ff_batch();
  b->dim = dim;
  b->storage = storage;
  b->options = dim == 3 ? (options & ~FF_OPT_STAGE_FP16) : options;
  b->device = device;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return hip_fail(e, "hipStreamCreate");
  }
  e = hipMalloc((void**)&b->d_flag, 256);
  if (e == hipSuccess) e = hipMemset(b->d_flag, 0, 256);
  if (e != hipSuccess) {
    if (b->d_flag) (void)hipFree(b->d_flag);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return hip_fail(e, "batch flag");
  }
  *out = b;
  return FF_OK;
}

int ff_batch_destroy(ff_batch* b) {
  if (!b) return FF_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  while (!b->mem.empty()) (void)detach(*b, *b->mem.back());
  drop_graph(*b);
  if (b->d_doms) (void)hipFree(b->d_doms);
  if (b->d_map) (void)hipFree(b->d_map);
  if (b->d_steps) (void)hipFree(b->d_steps);
  if (b->d_flag) (void)hipFree(b->d_flag);
  if (b->d_pend) (void)hipFree(b->d_pend);
  if (b->h_pend) (void)hipHostFree(b->h_pend);
  if (b->pend_ev) (void)hipEventDestroy(b->pend_ev);
  if (b->h_tab) (void)hipHostFree(b->h_tab);
  if (b->tab_ev) (void)hipEventDestroy(b->tab_ev);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return FF_OK;
}

int ff_batch_add(ff_batch* b, ff_domain* dom) {
  FF_CHECK_BATCH(b);
  if (!dom) return fail(FF_ERR_INVALID, "null domain");
  Domain& d = *reinterpret_cast<Domain*>(dom);
  if (d.batch) return fail(FF_ERR_STATE, "domain already belongs to a batch");
  if (d.dim != b->dim || d.storage != b->storage || d.options != b->options || d.device != b->device)
    return fail(FF_ERR_INVALID, "domain kind (dim, storage, options, device) differs from the batch");
  FF_HIP(hipSetDevice(b->device));
  if (int rc = flush_finish(d)) return rc;
  FF_HIP(hipStreamSynchronize(d.stream));
  FF_HIP(hipStreamSynchronize(b->stream));
  if (d.has_boundary) {
    // batched steps read and advance gen and tag on the device
    StepState st;
    FF_HIP(hipMemcpy(&st, d.st, sizeof(st), hipMemcpyDeviceToHost));
    st.gen = d.gen_host;
    st.tag = d.tag_host;
    FF_HIP(hipMemcpy(d.st, &st, sizeof(st), hipMemcpyHostToDevice));
  }
  d.own_stream = d.stream;
  d.stream = b->stream;
  d.batch = b;
  b->mem.push_back(&d);
  b->dirty = true;
  return FF_OK;
}

int ff_batch_remove(ff_batch* b, ff_domain* dom) {
  FF_CHECK_BATCH(b);
  if (!dom) return fail(FF_ERR_INVALID, "null domain");
  return detach(*b, *reinterpret_cast<Domain*>(dom));
}

int ff_batch_size(ff_batch* b, int* n) {
  FF_CHECK_BATCH(b);
  if (!n) return fail(FF_ERR_INVALID, "null output");
  *n = (int)b->mem.size();
  return FF_OK;
}

int ff_batch_step(ff_batch* b, int nsteps) {
  FF_CHECK_BATCH(b);
  if (nsteps < 0) return fail(FF_ERR_INVALID, "negative step count");
  if (b->mem.empty() || nsteps == 0) return FF_OK;
  FF_HIP(hipSetDevice(b->device));
  int rc;
  if (b->dirty && (rc = rebuild(*b))) return rc;
  if ((rc = detect_pending(*b))) return rc;
  if ((rc = replay(*b, nsteps))) return rc;
  advance_mirrors(*b, nsteps);
  return FF_OK;
}

int ff_batch_sync(ff_batch* b) {
  FF_CHECK_BATCH(b);
  FF_HIP(hipSetDevice(b->device));
  FF_HIP(hipStreamSynchronize(b->stream));
  int flag = 0;  // sticky, like a member's own overflow flag
  FF_HIP(hipMemcpy(&flag, b->d_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) return fail(FF_ERR_CAPACITY, "boundary link table overflow in a batched domain");
  return FF_OK;
}

int ff_batch_set_tile(ff_batch* b, int tile) {
  FF_CHECK_BATCH(b);
  if (tile < -1 || tile >= (b->dim == 2 ? kNumTiles2 : kNumTiles3)) return fail(FF_ERR_INVALID, "tile index out of range");
  b->tile = tile;
  if (tile < 0) b->tuned_tile = -1;
  b->dirty = true;
  return FF_OK;
}

int ff_batch_get_tile(ff_batch* b, int* tile, int* tx, int* ty) {
  FF_CHECK_BATCH(b);
  if (!tile) return fail(FF_ERR_INVALID, "null output");
  FF_HIP(hipSetDevice(b->device));
  int rc;
  if (b->dirty && !b->mem.empty() && (rc = rebuild(*b))) return rc;
  const int t = current_tile(*b);
  *tile = t;
  if (tx) *tx = b->dim == 2 ? kTiles2[t].tx : kTiles3[t].tx;
  if (ty) *ty = b->dim == 2 ? kTiles2[t].ty * kTiles2[t].cpt : kTiles3[t].ty;
  return FF_OK;
}

int ff_batch_time_steps(ff_batch* b, int nsteps, double* total_ms) {
  FF_CHECK_BATCH(b);
  if (nsteps < 0 || !total_ms) return fail(FF_ERR_INVALID, "bad arguments");
  FF_HIP(hipSetDevice(b->device));
  int rc;
  if (!b->mem.empty() && b->dirty && (rc = rebuild(*b))) return rc;
  EventSet ev;
  FF_HIP(ev.create(2));
  if (!b->mem.empty() && (rc = detect_pending(*b))) return rc;
  FF_HIP(hipEventRecord(ev[0], b->stream));
  if (!b->mem.empty() && (rc = replay(*b, nsteps))) return rc;
  FF_HIP(hipEventRecord(ev[1], b->stream));
  FF_HIP(hipEventSynchronize(ev[1]));
  float ms = 0.f;
  FF_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  if (!b->mem.empty()) advance_mirrors(*b, nsteps);
  *total_ms = ms;
  return ff_batch_sync(b);
}

int ff_batch_time_bulk(ff_batch* b, int iters, double* avg_ms) {
  FF_CHECK_BATCH(b);
  if (iters <= 0 || !avg_ms) return fail(FF_ERR_INVALID, "bad arguments");
  FF_HIP(hipSetDevice(b->device));
  int rc;
  if (b->mem.empty()) return fail(FF_ERR_STATE, "empty batch");
  if (b->dirty && (rc = rebuild(*b))) return rc;
  // the stream-collide tiles of every member (the link-detection blocks of a
  // step are excluded); they rewrite only cells every step rewrites
  const int tile = current_tile(*b);
  const int2* map = b->d_map + b->n_bc + b->n_wave;
  EventSet ev;
  ev.clears[0] = &b->tev0;
  ev.clears[1] = &b->tev1;
  FF_HIP(ev.create(2 * (size_t)iters));
  for (int k = 0; k < iters; k++) {
    b->tev0 = ev[2 * k];  // the launch's own start and end (launch_timed)
    b->tev1 = ev[2 * k + 1];
    launch_bulk_only(*b, tile, map, b->n_bulk);
  }
  b->tev0 = b->tev1 = nullptr;
  FF_HIP(hipGetLastError());
  FF_HIP(hipStreamSynchronize(b->stream));
  double tot = 0.0;
  for (int k = 0; k < iters; k++) {
    float ms = 0.f;
    FF_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
    tot += ms;
  }
  *avg_ms = tot / iters;
  return FF_OK;
}

}  // extern "C"
