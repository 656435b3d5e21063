// bh_tracking.h -- the photon tracking loop of the MI355X IMC engine.
//
// One segment = one iteration of `while (active)` in transport_photon
// (src/history_based_transport.h:55-140): opacity fetch, RNG draw, distances
// to scatter / face / census, implicit-capture decay, move, then the event.
// segment() advances a live photon by exactly one segment and reports
// whether its history ended, so the same code drives
//   * k_transport_history    -- one work-item per photon, loop to the end;
//   * k_transport_persistent -- a grid sized to the chip whose lanes refill
//     from a global queue whenever enough of a wave's lanes are idle.
//
// Every random number of a segment comes from ONE Threefry evaluation site:
// the draws an effective scatter needs (direction mu and phi, the
// re-emission test, the group walk) are taken at the start of the photon's
// next segment, right before that segment's distance draw.  That is the
// reference's draw order (a scatter never ends a history, so its draws
// always precede the next distance draw), and a single inlined generator
// instead of five keeps the kernel small enough for a useful occupancy.
//
// Precision forms:
//   FP64 (Real = double): statement-for-statement replay of the reference;
//   FP32 tracking (Real = float): position (in the frame of the photon's
//     cell: offsets from its low corner, so an ulp is 2^-24 of a cell width,
//     not of the domain), direction, path lengths and the decay exponent in
//     FP32; energy weights, tallies, the re-emission test and the group walk
//     stay FP64.  Each history carries running bounds on its FP32 position
//     error (across and along the ray), life error and the relative error
//     of its weight.  A decision the bounds cannot separate -- which of scatter / face / census
//     comes first, which face, whether E/E0 falls below the cutoff -- stops
//     the history before that segment is applied: the deposits of the
//     segments already decided are tallied, and k_transport_replay re-runs
//     the history in FP64 from its initial state (still in place), tallying
//     from that segment on.  Every event therefore matches the FP64 form; the
//     tallies differ only by FP32 rounding of the decided segments.
#pragma once

#include "bh_kernels.h"

// Occupancy target of the tracking kernels (waves per SIMD).  The FP64 form
// needs ~200 VGPRs unconstrained (2 waves); asking for 3 costs a few spilled
// registers and buys a third wave to hide the opacity/face gather latency.
#ifndef BH_TRACK_WAVES_PER_EU
#define BH_TRACK_WAVES_PER_EU 3
#endif
#if BH_TRACK_WAVES_PER_EU > 0
#define BH_TRACK_KERNEL \
  __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BH_TRACK_WAVES_PER_EU)))
#else
#define BH_TRACK_KERNEL __global__ __launch_bounds__(256)
#endif

namespace bh {

// The persistent kernel's work queue: chunk c of the launch's photons is
// handed out by counter c % 8, so the waves claim through eight counters,
// each on a 128-byte line of its own (atomics from all waves on one line
// serialise), while the photons in flight still form one window moving
// through the stores in order.  A wave starts at the counter of its XCD.
constexpr uint32_t kQueueParts = 8;
struct PersistentQueue {
  struct alignas(128) Counter {
    unsigned long long next;  // chunks this counter handed out
  } part[kQueueParts];
};

// This wave's XCD (hardware register XCC_ID), 0-7 on MI355X.
__device__ __forceinline__ uint32_t xcc_id() {
#if __HIP_DEVICE_COMPILE__
  uint32_t v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return v & (kQueueParts - 1u);
#else
  return 0;
#endif
}

// Before each persistent launch: the counters zeroed.
__global__ void k_queue_init(PersistentQueue* q) {
  if (threadIdx.x < kQueueParts) q->part[threadIdx.x].next = 0;
}

// Claims the next chunk of photon indices from the counter of the wave's
// XCD, or from the others in turn once it is used up.  Returns the chunk's
// first index and sets its end, or returns n when every counter is used up.
__device__ __forceinline__ uint64_t claim_chunk(PersistentQueue* q, uint64_t n, uint64_t chunk,
                                                uint64_t& end) {
  const uint32_t xcc = xcc_id();
#pragma unroll 1
  for (uint32_t k = 0; k < kQueueParts; ++k) {
    const uint32_t part = (xcc + k) & (kQueueParts - 1u);
    const uint64_t c = atomicAdd(&q->part[part].next, 1ull);
    const uint64_t base = (c * kQueueParts + part) * chunk;
    if (base < n) {
      end = base + chunk < n ? base + chunk : n;
      return base;
    }
  }
  return n;
}

template <typename R>
struct Live {
  R x, y, z, u, v, w, life;
  double E, E0;
  double t_abs, t_track;
  Rng rng;
  uint32_t cell, group, sc, segs;
  uint32_t first_seg;  // FP64 replay: deposits of earlier segments are muted
  uint32_t frame;      // FP32 form: the cell whose low corner x, y, z are from
  // FP32 form, error bounds against exact arithmetic from the same start:
  // position error perpendicular to the ray (eP) and along it (eT), life
  // (eL), all in cm, and the relative error of the weight (eE)
  float eP, eT, eL, eE;
  double E_handoff;    // FP64 replay: the FP32 weight when FP32 stopped
  uint32_t ijk[3];     // lattice meshes: the cell's (i, j, k)
  uint8_t desc;
  uint8_t stype;  // source type (census photons keep it)
  bool pending;  // an effective scatter awaits its draws
};

// FP32 rounding constants of the error bounds: kU is the unit roundoff (one
// correctly rounded operation), kDir bounds the error of each FP32
// direction component, kE covers a few ulps (logf, expm1f, expf).
constexpr float kU = 0x1p-24f;
constexpr float kDir = 0x1p-22f;
constexpr float kE = 0x1p-21f;

// The data one segment reads about its cell.  kLat = 0: the 64-byte
// CellGeo record and the per-group opacities; kLat = 1 (lattice meshes):
// the 32-byte CellRec and the axis node arrays, indexed by the photon's
// (i, j, k); kLat = 2 (a batch of problems that are each a lattice): the
// CellRec and the axes of the cell's problem.  node(k) is the face
// coordinate k = 2 * axis + (high side).  All views return the same numbers
// for the same cell.
template <int kLat>
struct CellView;

template <>
struct CellView<0> {
  const CellGeo* geo;
  const double2* sig;
  double f;
  __device__ __forceinline__ CellView(const DevCells& cells, uint32_t c, const uint32_t*)
      : geo(&cells.geo[c]), sig(&cells.sigma[(size_t)c * cells.G]), f(geo->fleck) {}
  __device__ __forceinline__ double2 sigma(uint32_t g) const { return sig[g]; }
  __device__ __forceinline__ double node(uint32_t k) const { return geo->nodes[k]; }
  __device__ __forceinline__ uint32_t bc_bits() const { return geo->bc_bits; }
};

// One of three values by a run-time index, as selects: indexing an array
// with a run-time value would put the array (and what holds it) in scratch
// memory instead of registers.
template <typename T>
__device__ __forceinline__ T pick3(T a, T b, T c, uint32_t i) {
  return i == 0 ? a : (i == 1 ? b : c);
}

template <>
struct CellView<1> {
  const double* ax0;
  const double* ax1;
  const double* ax2;
  uint32_t i0, i1, i2;
  double sa, ss, f;
  uint32_t bits;
  __device__ __forceinline__ CellView(const DevCells& c, uint32_t cell, const uint32_t* ijk)
      : ax0(c.axis[0]), ax1(c.axis[1]), ax2(c.axis[2]), i0(ijk[0]), i1(ijk[1]), i2(ijk[2]) {
    const CellRec r = c.rec[cell];
    sa = r.sigma_a;
    ss = r.sigma_s;
    f = r.fleck;
    bits = r.bc_bits;
  }
  __device__ __forceinline__ double2 sigma(uint32_t) const { return make_double2(sa, ss); }
  __device__ __forceinline__ double node(uint32_t k) const {
    const uint32_t a = k >> 1;
    return pick3(ax0, ax1, ax2, a)[pick3(i0, i1, i2, a) + (k & 1u)];
  }
  __device__ __forceinline__ uint32_t bc_bits() const { return bits; }
};

template <>
struct CellView<2> {
  const double* ax;      // every problem's axes
  uint32_t i0, i1, i2;   // the cell's low nodes in ax
  uint32_t nx, nxy;      // index strides of the cell's problem
  double sa, ss, f;
  uint32_t bits;
  __device__ __forceinline__ CellView(const DevCells& c, uint32_t cell, const uint32_t* ijk)
      : ax(c.axis[0]) {
    const CellRec r = c.rec[cell];
    sa = r.sigma_a;
    ss = r.sigma_s;
    f = r.fleck;
    bits = r.bc_bits;
    const LatInst li = c.lat[r.pad];
    i0 = li.ox + ijk[0];
    i1 = li.oy + ijk[1];
    i2 = li.oz + ijk[2];
    nx = li.nx;
    nxy = li.nx * li.ny;
  }
  __device__ __forceinline__ double2 sigma(uint32_t) const { return make_double2(sa, ss); }
  __device__ __forceinline__ double node(uint32_t k) const {
    return ax[pick3(i0, i1, i2, k >> 1) + (k & 1u)];
  }
  __device__ __forceinline__ uint32_t bc_bits() const { return bits; }
};

// (i, j, k) of cell c on a lattice numbered x fastest (kLat = 2: within the
// cell's problem).
template <int kLat = 1>
__device__ __forceinline__ void lattice_ijk(const DevCells& cells, uint32_t c, uint32_t ijk[3]) {
  uint32_t nx = cells.dims[0], ny = cells.dims[1];
  if constexpr (kLat == 2) {
    const LatInst li = cells.lat[cells.cell_inst[c]];
    c -= li.first;
    nx = li.nx;
    ny = li.ny;
  }
  const uint32_t nxy = nx * ny;
  ijk[2] = c / nxy;
  const uint32_t r = c - ijk[2] * nxy;
  ijk[1] = r / nx;
  ijk[0] = r - ijk[1] * nx;
}

template <typename R, int kLat = 0>
__device__ __forceinline__ void load_live(const DevCells& cells, const PhotonSoA& W, uint64_t i,
                                          uint64_t ctr_hi, Live<R>& L) {
  L.cell = W.cell[i];
  if constexpr (kLat) lattice_ijk<kLat>(cells, L.cell, L.ijk);
  if constexpr (sizeof(R) == 4) {
    const CellView<kLat> cv(cells, L.cell, L.ijk);
    L.frame = L.cell;
    L.x = (float)(W.x[i] - cv.node(0));
    L.y = (float)(W.y[i] - cv.node(2));
    L.z = (float)(W.z[i] - cv.node(4));
    L.life = (float)W.life[i];
    L.eP = kU * (fabsf(L.x) + fabsf(L.y) + fabsf(L.z));
    L.eT = L.eP;
    L.eL = kU * fabsf(L.life);
    L.eE = 0.0f;
  } else {
    L.x = W.x[i];
    L.y = W.y[i];
    L.z = W.z[i];
    L.life = W.life[i];
  }
  L.u = (R)W.u[i];
  L.v = (R)W.v[i];
  L.w = (R)W.w[i];
  L.E = W.E[i];
  L.E0 = W.E0[i];
  L.rng.ctr = W.ctr[i];
  L.rng.stream = W.stream[i];
  L.rng.ctr_hi = photon_ctr_hi(cells, L.cell, ctr_hi);
  L.group = W.group[i];
  L.desc = W.desc[i];
  L.stype = W.stype[i];
  L.sc = 0;  // surface_cross starts at 0 for every photon (:43)
  L.segs = 0;
  L.t_abs = 0.0;
  L.t_track = 0.0;
  L.first_seg = 0;
  L.pending = false;
}

template <typename R, int kLat = 0>
__device__ __forceinline__ void store_live(const DevCells& cells, const PhotonSoA& W, uint64_t i,
                                           const Live<R>& L) {
  if constexpr (sizeof(R) == 4) {
    // lattice meshes have no PROCESSOR faces, so the frame is the cell
    const double x0 = kLat ? cells.axis[0][L.ijk[0]] : cells.geo[L.frame].nodes[0];
    const double y0 = kLat ? cells.axis[1][L.ijk[1]] : cells.geo[L.frame].nodes[2];
    const double z0 = kLat ? cells.axis[2][L.ijk[2]] : cells.geo[L.frame].nodes[4];
    W.x[i] = x0 + (double)L.x;
    W.y[i] = y0 + (double)L.y;
    W.z[i] = z0 + (double)L.z;
  } else {
    W.x[i] = L.x;
    W.y[i] = L.y;
    W.z[i] = L.z;
  }
  W.u[i] = (double)L.u;
  W.v[i] = (double)L.v;
  W.w[i] = (double)L.w;
  W.life[i] = (double)L.life;
  W.E[i] = L.E;
  W.ctr[i] = L.rng.ctr;
  W.cell[i] = L.cell;
  W.group[i] = L.group;
  W.desc[i] = L.desc;
}

// At most one tally flush per segment: the cell visit the photon leaves or
// ends in (Flush, bh_kernels.h).  The kernel hands it to the tally sink at a
// wave-uniform point.  t_abs holds the implicit-capture deposit only; a kill
// weight travels separately in Flush::kill.
template <typename R>
__device__ __forceinline__ void flush(Flush& fo, const Live<R>& L) {
  fo.has = true;
  fo.c = L.cell;
  fo.cont = L.t_abs;
  fo.track = L.t_track;
}

// Boundary event shared by both precisions (:102-131).  Returns true when
// the history ends.
// Moving to another cell of this mesh: FP64 positions are global; FP32
// positions move to the new cell's frame (one FP32 rounding of the shift on
// the crossing axis; zero shift on the other axes of a rectilinear mesh).
// A PROCESSOR face ends the history with the neighbour's index, which may
// not be a cell of this mesh, so there the frame stays with the old cell.
// The neighbour across face sc (ELEMENT or PROCESSOR); on a lattice the
// neighbour is the next index along the face's axis.
template <int kLat, typename R>
__device__ __forceinline__ uint32_t neighbour(const DevCells& cells, const CellView<kLat>& cv,
                                              const Live<R>& L) {
  if constexpr (kLat) {
    const uint32_t a = L.sc >> 1;
    uint32_t stride;
    if constexpr (kLat == 2)
      stride = (a == 0) ? 1u : (a == 1) ? cv.nx : cv.nxy;
    else
      stride = (a == 0) ? 1u : (a == 1) ? cells.dims[0] : cells.dims[0] * cells.dims[1];
    return (L.sc & 1u) ? L.cell + stride : L.cell - stride;
  } else {
    return cells.next[(size_t)L.cell * 6 + L.sc];
  }
}

// Lattice: (i, j, k) moves one step along the crossing face's axis.  The
// components are updated by selects (no run-time array index).
template <typename R>
__device__ __forceinline__ uint32_t lattice_step(Live<R>& L, uint32_t* from) {
  const uint32_t a = L.sc >> 1;
  const uint32_t i0 = pick3(L.ijk[0], L.ijk[1], L.ijk[2], a);
  const uint32_t i1 = (L.sc & 1u) ? i0 + 1u : i0 - 1u;
  L.ijk[0] = a == 0 ? i1 : L.ijk[0];
  L.ijk[1] = a == 1 ? i1 : L.ijk[1];
  L.ijk[2] = a == 2 ? i1 : L.ijk[2];
  *from = i0;
  return i1;
}

template <int kLat>
__device__ __forceinline__ void enter_cell(const DevCells& cells, const CellView<kLat>&,
                                           uint32_t nc, Live<double>& L) {
  if constexpr (kLat) {
    uint32_t i0;
    (void)lattice_step(L, &i0);
  }
  L.cell = nc;
}
template <int kLat>
__device__ __forceinline__ void enter_cell(const DevCells& cells, const CellView<kLat>& from,
                                           uint32_t nc, Live<float>& L) {
  static_assert(kLat != 2, "FP32 tracking runs on the general or the single-lattice records");
  float sx, sy, sz;
  if constexpr (kLat) {  // only the crossing axis moves
    const uint32_t a = L.sc >> 1;
    uint32_t i0;
    const uint32_t i1 = lattice_step(L, &i0);
    const double* ax = pick3(cells.axis[0], cells.axis[1], cells.axis[2], a);
    const float s = (float)(ax[i1] - ax[i0]);
    sx = (a == 0) ? s : 0.0f;
    sy = (a == 1) ? s : 0.0f;
    sz = (a == 2) ? s : 0.0f;
  } else {
    const CellGeo* to = &cells.geo[nc];
    sx = (float)(to->nodes[0] - from.node(0));
    sy = (float)(to->nodes[2] - from.node(2));
    sz = (float)(to->nodes[4] - from.node(4));
  }
  L.x -= sx;
  L.y -= sy;
  L.z -= sz;
  const float r = 2.0f * kU * (fabsf(sx) + fabsf(sy) + fabsf(sz));
  L.eP += r;
  L.eT += r;
  L.cell = nc;
  L.frame = nc;
}

// A reflection mirrors both the exact and the FP32 continuation about the
// face; the two mirror planes differ by the FP32 rounding of the face
// offset, which is all a reflection adds to the error.
template <int kLat>
__device__ __forceinline__ void reflect_error(const CellView<kLat>&, uint32_t, Live<double>&) {}
template <int kLat>
__device__ __forceinline__ void reflect_error(const CellView<kLat>& cv, uint32_t axis,
                                              Live<float>& L) {
  const float w = (float)(cv.node(2 * axis + 1) - cv.node(2 * axis));
  L.eP += 2.0f * kU * w;
  L.eT += 2.0f * kU * w;
}

template <int kLat, typename R>
__device__ __forceinline__ bool boundary_event(const DevCells& cells, const CellView<kLat>& cv,
                                               Live<R>& L, Flush& fo) {
  const int bc = face_bc(cv.bc_bits(), L.sc);
  if (bc == kElement) {
    flush(fo, L);
    enter_cell(cells, cv, neighbour<kLat>(cells, cv, L), L);
    L.desc = kBound;
    L.t_abs = 0.0;
    L.t_track = 0.0;
    return false;
  }
  if (bc == kProcessor) {  // (never on a lattice mesh)
    const uint32_t nc = neighbour<kLat>(cells, cv, L);
    L.desc = kPass;
    flush(fo, L);
    L.cell = nc;
    return true;
  }
  if (bc == kVacuum || bc == kSource) {
    L.desc = kExit;
    flush(fo, L);
    return true;
  }
  // REFLECT: Photon::reflect (src/photon.h:150-155) negates the component of
  // axis sc/2.  Written as three value selects: a branch chain lets the
  // compiler turn it into one pointer-selected member access, which forces
  // the whole photon state out of registers into scratch memory.
  const uint32_t axis = L.sc >> 1;
  L.u = (axis == 0) ? -L.u : L.u;
  L.v = (axis == 1) ? -L.v : L.v;
  L.w = (axis == 2) ? -L.w : L.w;
  reflect_error(cv, axis, L);
  L.desc = kBound;
  return false;
}


// ---- Wave-cooperative draws ------------------------------------------------
// A segment takes one draw (the distance), or five after an effective
// scatter (mu, phi, the re-emission test, the group, the distance; the
// group draw only when the test re-samples, so the fifth may go unused).
// Drawn lane by lane, a wave runs the Threefry block up to five times per
// segment with most lanes idle after the first.  Instead, before the
// segments run, the wave computes every draw its lanes may need: draw k of
// lane o is Threefry(counter_o + k), the same value the lane's own k-th
// call would return, and the D <= 5 * 64 draws are spread over all 64 lanes
// in ceil(D / 64) passes.  Results go to a per-wave LDS block; the segment
// reads them in order and advances its counter by the draws it uses.
constexpr uint32_t kMaxDraws = 5;

// (sized for the source draws of a wave of new photons, seven each)
struct DrawLds {
  uint16_t job[kWave * kSourceDraws];  // (owner lane << 3) | k
  double res[kWave * kSourceDraws];
};

__device__ __forceinline__ DrawLds* draw_lds() {
  __shared__ DrawLds d[256 / kWave];
  return &d[threadIdx.x / kWave];
}

struct WaveDraws {
  const double* res;
  uint32_t next;  // index in res of the lane's next draw
  __device__ __forceinline__ double take(Rng& rng) {
    rng.ctr += 1u;
    return res[next++];
  }
};

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, kWave);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

// Every lane of the wave at the same program point (full execution mask:
// the owners' counters are read with cross-lane permutes).  `want` lanes
// are about to run a segment.
template <typename R>
__device__ __forceinline__ void wave_draws(bool want, const Live<R>& L, WaveDraws& dr) {
  const uint32_t K = want ? (L.pending ? kMaxDraws : 1u) : 0u;
  const uint64_t m5 = __ballot(K == kMaxDraws), m1 = __ballot(K == 1u);
  const uint32_t off = kMaxDraws * lane_rank(m5) + lane_rank(m1);
  const uint32_t D = kMaxDraws * (uint32_t)__popcll(m5) + (uint32_t)__popcll(m1);
  DrawLds* s = draw_lds();
  dr.res = s->res;
  dr.next = off;
  if (D == 0) return;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint32_t k = 0; k < kMaxDraws; ++k)
    if (k < K) s->job[off + k] = (uint16_t)((lane << 3) | k);
  wave_lds_sync();
  for (uint32_t base = 0; base < D; base += kWave) {
    const uint32_t j = base + lane;
    const uint32_t d = (j < D) ? (uint32_t)s->job[j] : (lane << 3);
    const int o = (int)(d >> 3);
    const uint64_t ctr = shfl_u64(L.rng.ctr, o) + (d & 7u);
    const uint64_t stream = shfl_u64(L.rng.stream, o);
    const uint64_t hi = shfl_u64(L.rng.ctr_hi, o);
    if (j < D) {
      uint64_t o0, o1;
      Threefry::block(ctr, hi, stream, 0ull, o0, o1);
      s->res[j] = u01(o0);
    }
  }
  wave_lds_sync();
}

// New photons made in a transport kernel: lane o with `gen` needs draws
// 0 .. nd-1 of RNG(its seed, its stream) -- the draws the source kernel
// would take -- and the wave computes all of them across its 64 lanes, as
// wave_draws does for segments.  Every lane at the same program point.
__device__ __forceinline__ void wave_gen_draws(bool gen, int nd, uint64_t stream, uint64_t hi,
                                               WaveDraws& dr) {
  const uint32_t K = gen ? (uint32_t)nd : 0u;
  const uint64_t m7 = __ballot(K == 7u), m6 = __ballot(K == 6u);
  const uint32_t off = 7u * lane_rank(m7) + 6u * lane_rank(m6);
  const uint32_t D = 7u * (uint32_t)__popcll(m7) + 6u * (uint32_t)__popcll(m6);
  DrawLds* s = draw_lds();
  dr.res = s->res;
  dr.next = off;
  if (D == 0) return;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint32_t k = 0; k < (uint32_t)kSourceDraws; ++k)
    if (k < K) s->job[off + k] = (uint16_t)((lane << 3) | k);
  wave_lds_sync();
  for (uint32_t base = 0; base < D; base += kWave) {
    const uint32_t j = base + lane;
    const uint32_t d = (j < D) ? (uint32_t)s->job[j] : (lane << 3);
    const int o = (int)(d >> 3);
    const uint64_t st = shfl_u64(stream, o);
    const uint64_t h = shfl_u64(hi, o);
    if (j < D) {
      uint64_t o0, o1;
      Threefry::block((uint64_t)(d & 7u), h, st, 0ull, o0, o1);  // counter = draw index
      s->res[j] = u01(o0);
    }
  }
  wave_lds_sync();
}

// A new photon's initial state in the registers of a tracking kernel, as
// load_live would read it from the store.
template <typename R, int kLat>
__device__ __forceinline__ void live_from_source(const DevCells& cells, const SourcePhoton& o,
                                                 uint64_t hi, Live<R>& L) {
  L.cell = o.cell;
  if constexpr (kLat) lattice_ijk<kLat>(cells, L.cell, L.ijk);
  if constexpr (sizeof(R) == 4) {
    const CellView<kLat> cv(cells, L.cell, L.ijk);
    L.frame = L.cell;
    L.x = (float)(o.x - cv.node(0));
    L.y = (float)(o.y - cv.node(2));
    L.z = (float)(o.z - cv.node(4));
    L.life = (float)o.life;
    L.eP = kU * (fabsf(L.x) + fabsf(L.y) + fabsf(L.z));
    L.eT = L.eP;
    L.eL = kU * fabsf(L.life);
    L.eE = 0.0f;
  } else {
    L.x = o.x;
    L.y = o.y;
    L.z = o.z;
    L.life = o.life;
  }
  L.u = (R)o.u;
  L.v = (R)o.v;
  L.w = (R)o.w;
  L.E = o.E;
  L.E0 = o.E;
  L.rng.ctr = o.ctr;
  L.rng.stream = o.stream;
  L.rng.ctr_hi = hi;
  L.group = o.group;
  L.desc = kPass;
  L.stype = o.stype;
  L.sc = 0;
  L.segs = 0;
  L.t_abs = 0.0;
  L.t_track = 0.0;
  L.first_seg = 0;
  L.pending = false;
}

// New photon p made by one lane alone (the FP64 replay): its draws one by
// one, then the source's sampling.
template <typename R, int kLat>
__device__ __forceinline__ void gen_photon_lane(const DevCells& cells, const SourceDesc& sd,
                                                uint64_t n_new, double dt, uint64_t p,
                                                Live<R>& L) {
  const uint32_t c = hinted_cell(sd, cells.n_cells, n_new, p);
  const SourceInst sp = source_record(cells, c, sd.one, sd.inst);
  Rng rng{0ull, source_stream(sp, sd.off, p), sp.ctr_hi};
  const SourcePhoton o = make_source_photon(cells, sd.sc, sd.face, sd.off, p, c, sp, dt,
                                            0, [&]() { return rng.next(); });
  live_from_source<R, kLat>(cells, o, sp.ctr_hi, L);
}

// Draw states of the lane-by-lane form of a segment's draws.
enum : int { kDrawMu = 0, kDrawPhi, kDrawTest, kDrawGroup, kDrawDistance, kDrawDone };

// What one FP64 segment tells the adjoint (filled when kInfo is set).
struct SegInfo {
  uint32_t c;            // cell of the segment
  double sa, f, sig;     // sigma_a, Fleck factor, sigma_eff = (1 - f) sa + ss
  double d, E_in, absorbed, kill_E;
  bool flight;           // a flight was sampled (sigma_eff > 0)
  bool scattered;        // ... and ended in a collision (d == d_s)
  bool killed;
  uint8_t reemit;        // re-emission test drawn this segment: 0 no, 1 kept, 2 re-sampled
  double r_sa, r_f, r_sig;  // the opacities that test used
};

// The draws of an FP64 segment taken lane by lane, around ONE Threefry site
// (a draw per loop trip keeps the replay, adjoint and event kernels small):
// after an effective scatter (L.pending) the new direction, the re-emission
// test and, when it re-samples, the group (src/history_based_transport.h:
// 95-100); then, with `distance`, the distance to scatter (:62-63), which is
// returned (1e100 when there is none).  sg / total_sigma_s follow a group
// change.
template <bool kInfo, int kLat>
__device__ __forceinline__ double lane_draws(const DevCells& cells, const CellView<kLat>& cv,
                                             double f, Live<double>& L, double2& sg,
                                             double& total_sigma_s, SegInfo* si, bool distance) {
  const uint32_t G = cells.G;
  double d_s = 1.0e100;
  int st = L.pending ? kDrawMu : (distance ? kDrawDistance : kDrawDone);
  if (st == kDrawDistance && !(total_sigma_s > 0.0)) st = kDrawDone;
  double mu = 0.0;
  while (st != kDrawDone) {
    const double r = L.rng.next();  // the loop's only Threefry site
    if (st == kDrawMu) {  // get_uniform_angle (src/sampling_functions.h:56-71)
      mu = r * 2.0 - 1.0;
      st = kDrawPhi;
    } else if (st == kDrawPhi) {
      const double phi = r * 2.0 * kPi;
      const double sin_theta = sqrt(1.0 - mu * mu);
      double sp, cp;
      sincos(phi, &sp, &cp);
      L.u = sin_theta * cp;
      L.v = sin_theta * sp;
      L.w = mu;
      st = kDrawTest;
    } else if (st == kDrawTest) {  // re-emission test (:97)
      const bool resample = r > (sg.y / total_sigma_s);
      if constexpr (kInfo) {
        si->reemit = resample ? 2 : 1;
        si->r_sa = sg.x;
        si->r_f = f;
        si->r_sig = total_sigma_s;
      }
      st = resample ? kDrawGroup
                    : ((distance && total_sigma_s > 0.0) ? kDrawDistance : kDrawDone);
    } else if (st == kDrawGroup) {  // sample_emission_group walk (:98)
      double cdf = r;
      int g = -1;
      const double norm = 1.0 / (cv.sigma(0).x * (double)G);
      while (cdf > 0) {
        g++;
        if ((uint32_t)g >= G) {
          g = (int)G - 1;
          break;
        }
        cdf -= cv.sigma((uint32_t)g).x * norm;
      }
      L.group = (uint32_t)g;
      sg = cv.sigma((uint32_t)g);
      total_sigma_s = (1.0 - f) * sg.x + sg.y;
      st = (distance && total_sigma_s > 0.0) ? kDrawDistance : kDrawDone;
    } else {  // distance to scatter (:62-63)
      d_s = -log(r) / total_sigma_s;
      st = kDrawDone;
    }
  }
  return d_s;
}

// ---- FP64: the reference's arithmetic, operation for operation ----------
// pre: the cell's view loaded before the wave's draws, so that the load's
// latency hides behind them (the draws' LDS fences keep the compiler from
// moving it there itself); null: loaded here.
template <bool kTrack, int kLat = 0, bool kInfo = false>
__device__ __forceinline__ bool segment(const DevCells& cells, Live<double>& L, Flush& fo,
                                        SegInfo* si = nullptr, WaveDraws* dr = nullptr,
                                        const CellView<kLat>* pre = nullptr) {
  if constexpr (kInfo) {
    si->reemit = 0;
    si->flight = false;
    si->killed = false;
    si->absorbed = 0.0;
    si->kill_E = 0.0;
    si->d = 0.0;
    si->c = L.cell;
  }
  if (++L.segs > kMaxSegments) {
    flush(fo, L);
    L.desc = kStuck;
    return true;
  }
  const uint32_t G = cells.G;
  const CellView<kLat> cv = pre ? *pre : CellView<kLat>(cells, L.cell, L.ijk);
  const double f = cv.f;
  double2 sg = cv.sigma(L.group);
  double total_sigma_s = (1.0 - f) * sg.x + sg.y;

  // Draws in the reference's order: after an effective scatter the new
  // direction (mu, phi), the re-emission test and, when it re-samples, the
  // group; then every photon's distance to scatter.  With the wave's draws
  // the code is straight-line, so the distance draw and its log run once for
  // the whole wave; lane by lane (replay, adjoint) it is a loop around one
  // Threefry site, which keeps those kernels small.
  double d_s = 1.0e100;
  if (dr) {  // the wave drew them beforehand
    auto draw = [&]() -> double { return dr->take(L.rng); };
    if (L.pending) {  // src/history_based_transport.h:95-100
      // get_uniform_angle (src/sampling_functions.h:56-71)
      const double mu = draw() * 2.0 - 1.0;
      const double phi = draw() * 2.0 * kPi;
      const double sin_theta = sqrt(1.0 - mu * mu);
      double sp, cp;
      sincos(phi, &sp, &cp);
      L.u = sin_theta * cp;
      L.v = sin_theta * sp;
      L.w = mu;
      const bool resample = draw() > (sg.y / total_sigma_s);  // re-emission test (:97)
      if constexpr (kInfo) {
        si->reemit = resample ? 2 : 1;
        si->r_sa = sg.x;
        si->r_f = f;
        si->r_sig = total_sigma_s;
      }
      if (resample) {  // sample_emission_group walk (:98)
        double cdf = draw();
        int g = -1;
        const double norm = 1.0 / (cv.sigma(0).x * (double)G);
        while (cdf > 0) {
          g++;
          if ((uint32_t)g >= G) {
            g = (int)G - 1;
            break;
          }
          cdf -= cv.sigma((uint32_t)g).x * norm;
        }
        L.group = (uint32_t)g;
        sg = cv.sigma((uint32_t)g);
        total_sigma_s = (1.0 - f) * sg.x + sg.y;
      }
    }
    if (total_sigma_s > 0.0) d_s = -log(draw()) / total_sigma_s;  // distance to scatter (:62-63)
  } else {  // lane by lane: one Threefry site, a draw per loop trip
    d_s = lane_draws<kInfo, kLat>(cells, cv, f, L, sg, total_sigma_s, si, true);
  }
  L.pending = false;
  const double sigma_a = sg.x;

  // Cell::get_distance_to_boundary (src/cell.h:116-132), sgn(a) = (0 < a);
  // the face coordinate is chosen by select, not by a runtime index
  double d_b = 1.0e16;
  {
    const bool px = 0.0 < L.u, py = 0.0 < L.v, pz = 0.0 < L.w;
    const double nx = cv.node(px ? 1u : 0u);
    const double ny = cv.node(py ? 3u : 2u);
    const double nz = cv.node(pz ? 5u : 4u);
    const double dx = (nx - L.x) / L.u;
    if (dx < d_b) { d_b = dx; L.sc = px ? 1u : 0u; }
    const double dy = (ny - L.y) / L.v;
    if (dy < d_b) { d_b = dy; L.sc = py ? 3u : 2u; }
    const double dz = (nz - L.z) / L.w;
    if (dz < d_b) { d_b = dz; L.sc = pz ? 5u : 4u; }
  }
  const double d_c = L.life;
  const double m_bc = (d_c < d_b) ? d_c : d_b;  // std::min(d_b, d_c)
  const double d = (m_bc < d_s) ? m_bc : d_s;    // std::min(d_s, .)

  const double absorbed = L.E * (1.0 - exp(-sigma_a * f * d));
  if constexpr (kInfo) {
    si->c = L.cell;
    si->sa = sigma_a;
    si->f = f;
    si->sig = total_sigma_s;
    si->d = d;
    si->E_in = L.E;
    si->absorbed = absorbed;
    si->flight = total_sigma_s > 0.0;
    si->scattered = (d == d_s);
  }
  L.t_abs += absorbed;
  if constexpr (kTrack) L.t_track += absorbed / (sigma_a * f);
  L.E = L.E - absorbed;

  L.x += L.u * d;  // Photon::move (src/photon.h:107-114)
  L.y += L.v * d;
  L.z += L.w * d;
  L.life -= d;

  if (L.E / L.E0 < kCutoff) {  // kill, deposit the remainder (:85-91)
    flush(fo, L);
    fo.kill = L.E;
    L.desc = kKilled;
    if constexpr (kInfo) {
      si->killed = true;
      si->kill_E = L.E;
    }
    return true;
  }
  if (d == d_s) {  // effective scatter (:95-100); draws taken next segment
    L.pending = true;
    L.desc = kScatter;
    return false;
  }
  if (d == d_b) return boundary_event<kLat>(cells, cv, L, fo);
  if (d == d_c) {  // census (:133-138)
    L.desc = kCensus;
    flush(fo, L);
    return true;
  }
  return false;
}

// ---- FP32 tracking ------------------------------------------------------
// The segment of the FP64 form with FP32 geometry, plus the error bounds
// that decide whether FP32 may take each decision.  The position error e
// (FP32 minus exact, from the same start) is bounded by its components
// perpendicular to the ray (eP) and along it (eT):
//   * a move by d adds the chosen distance's error along the ray, and the
//     FP32 direction's error (kDir per unit length) across it;
//   * a face hit puts the photon on the face: the along-ray error becomes
//     the face distance's evaluation error plus the perpendicular error seen
//     through the face (eP / |u_axis|), while eP is unchanged -- parallel
//     rays keep their separation, so errors do not compound per crossing;
//   * a new direction (scatter) re-splits |e| <= eP + eT;
//   * every FP32 update adds its rounding (kU per term).
// Distance errors: to a face (eP + kU w + kDir |d|) / |u_axis| + eT + 4 kU |d|
// (node offset, subtraction, reciprocal, product); to scatter
// kE d_s + 2 kU / sigma (u rounded to FP32 moves log u by kU); to census eL.
// A decision is ambiguous when the gap between the chosen candidate and any
// other is within twice the sum of their bounds.  A segment whose deposit
// could be off by more than fp32_tol (relative) is treated the same way, so
// every FP32 deposit -- and with it every cell's tally -- is within fp32_tol
// of FP64.
template <bool kTrack, int kLat = 0>
__device__ __forceinline__ bool segment(const DevCells& cells, Live<float>& L, Flush& fo,
                                        SegInfo* = nullptr, WaveDraws* dr = nullptr,
                                        const CellView<kLat>* pre = nullptr) {
  if (++L.segs > kMaxSegments) {
    flush(fo, L);
    L.desc = kStuck;
    return true;
  }
  const uint32_t G = cells.G;
  const CellView<kLat> cv = pre ? *pre : CellView<kLat>(cells, L.cell, L.ijk);
  const double f = cv.f;
  double2 sg = cv.sigma(L.group);
  double total_sigma_s = (1.0 - f) * sg.x + sg.y;

  // the FP64 form's draws, in its order (straight-line: the distance draw
  // and its logf run once for the whole wave)
  float d_s = 1.0e30f, e_s = 0.0f;
  if (dr) {  // the wave drew them beforehand
    auto draw = [&]() -> double { return dr->take(L.rng); };
    if (L.pending) {
      const double mu = draw() * 2.0 - 1.0;
      // sin(theta) from the FP64 cosine: in FP32, mu can round to +-1 and
      // leave a direction with two exactly-zero components, whose face
      // distances are -inf.  In FP64 sqrt(1 - mu^2) >= 2^-25.5 is never 0.
      // phi = 2 pi r, split as 4r = q + f with q = rint(4r) and |f| <= 1/2
      // (exact in FP64): phi = q pi/2 + f pi/2, so the FP32 argument of
      // sincospif is below pi/4 and rounds by at most 2^-25, and the quarter
      // turns are exact swaps and sign changes
      const double r4 = draw() * 4.0;
      const double q = rint(r4);
      float s0, c0;
      sincospif(0.5f * (float)(r4 - q), &s0, &c0);
      const int qi = (int)q & 3;
      const float cp = (qi == 0) ? c0 : (qi == 1) ? -s0 : (qi == 2) ? -c0 : s0;
      const float sp = (qi == 0) ? s0 : (qi == 1) ? c0 : (qi == 2) ? -s0 : -c0;
      const float s_t = (float)sqrt(1.0 - mu * mu);
      L.u = s_t * cp;
      L.v = s_t * sp;
      L.w = (float)mu;
      const float e = L.eP + L.eT;  // |e| seen from the new direction
      L.eP = e;
      L.eT = e;
      if (draw() > (sg.y / total_sigma_s)) {  // re-emission: new group
        double cdf = draw();
        int g = -1;
        const double norm = 1.0 / (cv.sigma(0).x * (double)G);
        while (cdf > 0) {
          g++;
          if ((uint32_t)g >= G) {
            g = (int)G - 1;
            break;
          }
          cdf -= cv.sigma((uint32_t)g).x * norm;
        }
        L.group = (uint32_t)g;
        sg = cv.sigma((uint32_t)g);
        total_sigma_s = (1.0 - f) * sg.x + sg.y;
      }
    }
    if (total_sigma_s > 0.0) {
      float uf = (float)draw();
      uf = (uf < 1.0f) ? uf : 0x1.fffffep-1f;  // (0,1) stays open after rounding
      const float sig = (float)total_sigma_s;
      d_s = -logf(uf) / sig;
      e_s = kE * d_s + 2.0f * kU / sig;
    }
  } else {  // lane by lane: one Threefry site, a draw per loop trip
    int st = L.pending ? kDrawMu : kDrawDistance;
    if (st == kDrawDistance && !(total_sigma_s > 0.0)) st = kDrawDone;
    double mu = 0.0;
    while (st != kDrawDone) {
      const double r = L.rng.next();  // the loop's only Threefry site
      if (st == kDrawMu) {
        mu = r * 2.0 - 1.0;
        st = kDrawPhi;
      } else if (st == kDrawPhi) {
        // sin(theta) from the FP64 cosine: in FP32, mu can round to +-1 and
        // leave a direction with two exactly-zero components, whose face
        // distances are -inf.  In FP64 sqrt(1 - mu^2) >= 2^-25.5 is never 0.
        // phi = 2 pi r, split as 4r = q + f with q = rint(4r) and |f| <= 1/2
        // (exact in FP64): phi = q pi/2 + f pi/2, so the FP32 argument of
        // sincospif is below pi/4 and rounds by at most 2^-25, and the quarter
        // turns are exact swaps and sign changes
        const double r4 = r * 4.0;
        const double q = rint(r4);
        float s0, c0;
        sincospif(0.5f * (float)(r4 - q), &s0, &c0);
        const int qi = (int)q & 3;
        const float cp = (qi == 0) ? c0 : (qi == 1) ? -s0 : (qi == 2) ? -c0 : s0;
        const float sp = (qi == 0) ? s0 : (qi == 1) ? c0 : (qi == 2) ? -s0 : -c0;
        const float s_t = (float)sqrt(1.0 - mu * mu);
        L.u = s_t * cp;
        L.v = s_t * sp;
        L.w = (float)mu;
        const float e = L.eP + L.eT;  // |e| seen from the new direction
        L.eP = e;
        L.eT = e;
        st = kDrawTest;
      } else if (st == kDrawTest) {
        st = (r > (sg.y / total_sigma_s)) ? kDrawGroup : kDrawDistance;
      } else if (st == kDrawGroup) {
        double cdf = r;
        int g = -1;
        const double norm = 1.0 / (cv.sigma(0).x * (double)G);
        while (cdf > 0) {
          g++;
          if ((uint32_t)g >= G) {
            g = (int)G - 1;
            break;
          }
          cdf -= cv.sigma((uint32_t)g).x * norm;
        }
        L.group = (uint32_t)g;
        sg = cv.sigma((uint32_t)g);
        total_sigma_s = (1.0 - f) * sg.x + sg.y;
        st = (total_sigma_s > 0.0) ? kDrawDistance : kDrawDone;
      } else {
        float uf = (float)r;
        uf = (uf < 1.0f) ? uf : 0x1.fffffep-1f;  // (0,1) stays open after rounding
        const float sig = (float)total_sigma_s;
        d_s = -logf(uf) / sig;
        e_s = kE * d_s + 2.0f * kU / sig;
        st = kDrawDone;
      }
    }
  }
  L.pending = false;
  const double saf = sg.x * f;

  // faces in the cell's frame: low faces at 0, high faces at the width;
  // the strict-< axis order of the reference picks the face, and the
  // runner-up axis is kept for the face ambiguity test
  float d_b = 1.0e16f, e_b = 0.0f, d_b2 = 1.0e16f, e_b2 = 0.0f;
  {
    const bool px = 0.0f < L.u, py = 0.0f < L.v, pz = 0.0f < L.w;
    const float wx = (float)(cv.node(1) - cv.node(0));
    const float wy = (float)(cv.node(3) - cv.node(2));
    const float wz = (float)(cv.node(5) - cv.node(4));
    const float ix = 1.0f / L.u, iy = 1.0f / L.v, iz = 1.0f / L.w;
    const float dx = ((px ? wx : 0.0f) - L.x) * ix;
    const float dy = ((py ? wy : 0.0f) - L.y) * iy;
    const float dz = ((pz ? wz : 0.0f) - L.z) * iz;
    const float ex = (L.eP + kU * wx + kDir * fabsf(dx)) * fabsf(ix) + L.eT + 4.0f * kU * fabsf(dx);
    const float ey = (L.eP + kU * wy + kDir * fabsf(dy)) * fabsf(iy) + L.eT + 4.0f * kU * fabsf(dy);
    const float ez = (L.eP + kU * wz + kDir * fabsf(dz)) * fabsf(iz) + L.eT + 4.0f * kU * fabsf(dz);
    if (dx < d_b) { d_b2 = d_b; e_b2 = e_b; d_b = dx; e_b = ex; L.sc = px ? 1u : 0u; }
    else if (dx < d_b2) { d_b2 = dx; e_b2 = ex; }
    if (dy < d_b) { d_b2 = d_b; e_b2 = e_b; d_b = dy; e_b = ey; L.sc = py ? 3u : 2u; }
    else if (dy < d_b2) { d_b2 = dy; e_b2 = ey; }
    if (dz < d_b) { d_b2 = d_b; e_b2 = e_b; d_b = dz; e_b = ez; L.sc = pz ? 5u : 4u; }
    else if (dz < d_b2) { d_b2 = dz; e_b2 = ez; }
  }
  const float d_c = L.life, e_c = L.eL;
  const float m_bc = (d_c < d_b) ? d_c : d_b;
  const float d = (m_bc < d_s) ? m_bc : d_s;
  const bool is_s = (d == d_s);
  const bool is_b = !is_s && (d == d_b);
  const bool is_c = !is_s && !is_b;
  const float e_d = is_s ? e_s : (is_b ? e_b : e_c);
  bool amb = (!is_s && (d_s - d) <= 2.0f * (e_s + e_d)) ||
             (!is_b && (d_b - d) <= 2.0f * (e_b + e_d)) ||
             (!is_c && (d_c - d) <= 2.0f * (e_c + e_d)) ||
             (is_b && d_b2 < 1.0e16f && (d_b2 - d_b) <= 2.0f * (e_b + e_b2));

  // implicit capture with the weight in FP64: the absorbed fraction from
  // -expm1f below ln 2, the surviving fraction from expf above, so neither
  // small nor large optical depths lose digits to cancellation
  const float x = (float)saf * d;
  double absorbed, E_new;
  float eE = L.eE + (float)saf * e_d;
  if (x < 0.69314718f) {
    absorbed = L.E * (double)(-expm1f(-x));
    E_new = L.E - absorbed;
    eE += kE * fabsf(x);
  } else {
    E_new = L.E * (double)expf(-x);
    absorbed = L.E - E_new;
    eE += kE * (x + 1.0f);
  }
  const double ratio = E_new / L.E0;
  amb = amb || (fabs(ratio - kCutoff) <= 2.0 * (double)eE * ratio);
  // accuracy cap: the deposit's relative error is bounded by the distance's
  // relative bound plus the weight's; past fp32_tol the history goes to FP64
  // (also catches d = 0 and faces met at grazing incidence)
  // (kE / 2 covers expm1f / expf and the FP32 products of the exponent)
  amb = amb || !(e_d + (eE + 0.5f * kE) * fabsf(d) <= cells.fp32_tol * fabsf(d));
  if (amb) {  // stop before this segment; FP64 replays from here on
    flush(fo, L);
    L.first_seg = L.segs;
    L.E_handoff = L.E;
    L.desc = kReplay;
    return true;
  }
  L.t_abs += absorbed;
  if constexpr (kTrack) L.t_track += absorbed / saf;
  L.E = E_new;
  L.eE = eE;

  L.x += L.u * d;
  L.y += L.v * d;
  L.z += L.w * d;
  L.life -= d;
  {
    const float ad = fabsf(d);
    const float rnd = kU * (fabsf(L.x) + fabsf(L.y) + fabsf(L.z) + 2.0f * ad);
    L.eT = (is_b ? e_d - L.eT : L.eT + e_d) + rnd;
    L.eP += kDir * ad + rnd;
    L.eL += e_d + kU * fabsf(L.life);
  }

  if (ratio < kCutoff) {
    flush(fo, L);
    fo.kill = L.E;
    L.desc = kKilled;
    return true;
  }
  if (is_s) {
    L.pending = true;
    L.desc = kScatter;
    return false;
  }
  if (is_b) return boundary_event<kLat>(cells, cv, L, fo);
  L.desc = kCensus;
  flush(fo, L);
  return true;
}

// Wave-level sums, one atomic per wave and counter: valid for any block
// size (the tracking kernels run with 64, 128 or 256 work-items).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// segs per lane; flushes and stuck already summed over the wave.
__device__ __forceinline__ void add_wave_counters(DevCounters* ctrs, uint32_t segs,
                                                  unsigned long long flushes,
                                                  unsigned long long stuck) {
  const unsigned long long s = wave_sum_u64(segs);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicAdd(&ctrs->segments, s);
    atomicAdd(&ctrs->flushes, flushes);
    if (stuck) atomicAdd(&ctrs->stuck, stuck);
  }
}

__device__ __forceinline__ void add_counters(DevCounters* ctrs, unsigned long long segs,
                                             unsigned long long flushes,
                                             unsigned long long stuck) {
  const unsigned long long s = wave_sum_u64(segs);
  const unsigned long long f = wave_sum_u64(flushes);
  const unsigned long long k = wave_sum_u64(stuck);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicAdd(&ctrs->segments, s);
    atomicAdd(&ctrs->flushes, f);
    if (k) atomicAdd(&ctrs->stuck, k);
  }
}

template <class Tally>
__device__ __forceinline__ void bind_lds(Tally& tally, uint32_t n_cells) {
  extern __shared__ double s_tally[];
  if constexpr (Tally::kLds) {
    tally.s_abs = s_tally;
    tally.s_track = s_tally + n_cells;
  }
}

// One work-item per photon: the reference's kernel shape
// (src/history_based_transport.h:263-275) over SoA state.  The segment loop
// runs while any lane of the wave is active, so tally pushes stay
// wave-uniform.
// Wave-aggregated append of the lanes with `want` set to the replay list:
// one atomic per wave on the list length.  Called at wave-uniform points.
__device__ __forceinline__ void replay_append(bool want, uint64_t idx, uint32_t first_seg,
                                              double E, const ReplayList& rl,
                                              DevCounters* ctrs) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(&ctrs->replays, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, kWave);
  if (want) {
    const uint64_t slot = base + lane_rank(m);
    rl.idx[slot] = idx;
    rl.first_seg[slot] = first_seg;
    rl.E[slot] = E;
  }
}

// End of a history inside a tracking kernel: store the photon, or -- FP32
// form, ambiguous decision -- leave its initial state in place for the
// replay.  Returns the segments this kernel applied.
template <typename R, int kLat>
__device__ __forceinline__ uint32_t end_history(const DevCells& cells, const PhotonSoA& W,
                                                uint64_t i, const Live<R>& L, bool& replay,
                                                bool in_place) {
  if constexpr (sizeof(R) == 4) {
    if (L.desc == kReplay) {
      replay = true;
      return L.segs - 1;
    }
  }
  if (in_place) store_live<R, kLat>(cells, W, i, L);
  return L.segs;
}

// The photon stores' 32 field pointers are needed only at refills, history
// ends and census writes.  Held in scalar registers for the whole kernel
// they spill, and every iteration pays the reloads.  The transport kernels
// therefore take the stores as their first two arguments (W at offset 0 of
// the kernel-argument segment, the census output after it) and read them
// there where they are used, through a pointer the optimiser cannot see
// through: scalar loads at the use, nothing held in between.
using KernargPtr = const __attribute__((address_space(4))) char*;
constexpr uint32_t kCensusArgOffset =
    (uint32_t)((sizeof(PhotonSoA) + alignof(CensusOut) - 1) / alignof(CensusOut) *
               alignof(CensusOut));

// The kernels that make new photons take the SourceDesc third, after the
// census output.
constexpr uint32_t kSourceArgOffset =
    (uint32_t)((kCensusArgOffset + sizeof(CensusOut) + alignof(SourceDesc) - 1) /
               alignof(SourceDesc) * alignof(SourceDesc));

// (The host pass parses these device functions too; it has no kernel
// argument segment, and never runs them.)
__device__ __forceinline__ PhotonSoA arg_stores() {
#if __HIP_DEVICE_COMPILE__
  KernargPtr k = (KernargPtr)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return *(const __attribute__((address_space(4))) PhotonSoA*)k;
#else
  return PhotonSoA{};
#endif
}
__device__ __forceinline__ CensusOut arg_census() {
#if __HIP_DEVICE_COMPILE__
  KernargPtr k = (KernargPtr)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return *(const __attribute__((address_space(4))) CensusOut*)(k + kCensusArgOffset);
#else
  return CensusOut{};
#endif
}

__device__ __forceinline__ SourceDesc arg_source() {
#if __HIP_DEVICE_COMPILE__
  KernargPtr k = (KernargPtr)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return *(const __attribute__((address_space(4))) SourceDesc*)(k + kSourceArgOffset);
#else
  return SourceDesc{};
#endif
}

// The view must show the by-value arguments.  Every work-item of a launch
// reaches the same answer; a launch that fails it does nothing, and the
// host reports the count.
__device__ __forceinline__ bool args_ok(const PhotonSoA& W, const CensusOut& co,
                                        DevCounters* ctrs) {
  const PhotonSoA w = arg_stores();
  const CensusOut c = arg_census();
  const bool ok = w.x == W.x && w.key == W.key && c.C.x == co.C.x && c.C.key == co.C.key &&
                  c.mark == co.mark && c.acc == co.acc && c.rec == co.rec &&
                  c.slots == co.slots && c.rec_cap == co.rec_cap && c.on == co.on;
  if (!ok && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctrs->bad_args, 1ull);
  return ok;
}
__device__ __forceinline__ bool args_ok(const PhotonSoA& W, const CensusOut& co,
                                        const SourceDesc& sd, DevCounters* ctrs) {
  const SourceDesc v = arg_source();
  const bool ok = v.off == sd.off && v.hint == sd.hint && v.sc == sd.sc && v.on == sd.on &&
                  v.n_new == sd.n_new;
  if (!ok && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctrs->bad_args, 1ull);
  return ok && args_ok(W, co, ctrs);
}

// A launch's counts: the kernel argument's, or the captured step's.
struct CensusRun {
  uint64_t n_cen, n_new;
  double life;
};
__device__ __forceinline__ CensusRun census_run(const CensusOut& co, const StepParams* P) {
  return P ? CensusRun{P->n_cen, P->n_new, P->census_life}
           : CensusRun{co.n_cen, co.n_new, co.life};
}

// What a lane counts of the histories it ends (census photons go to C), and
// the wave's chunk of census slots (wave-uniform).
// Everything but the exit energy is wave-uniform (counted from ballots), so
// it stays in scalar registers.
struct CensusLocal {
  double exit_E = 0.0;           // per lane
  uint32_t n_exit = 0, n_killed = 0, n_other = 0;
  uint32_t n_cen = 0;            // census photons the wave wrote
  uint64_t cbase = 0;            // first slot of the wave's current chunk
  uint32_t cfill = kCenChunk;    // slots used in it (kCenChunk: no chunk held)
};

// Every lane at the same program point.  `ended`: the lane ended a history
// (not handed to the FP64 replay) this iteration, from physical slot i of W.
// Census photons go to the next slots of the wave's chunk of C (a new chunk,
// one atomic, when it fills), with the reference's census life, and their
// transport index is marked for the census order.  kDense: the wave takes
// exactly the slots its census photons fill, one atomic per call (kernels
// launched many times per step, which must leave no holes).  kLiveType:
// the source type comes from the lane (kernels that make new photons);
// otherwise from the photon's slot of W, so that the lane does not carry it
// through the history.
template <typename R, int kLat, bool kDense = false, bool kLiveType = false>
__device__ __forceinline__ void census_append(bool ended, const DevCells& cells, uint64_t i,
                                              const Live<R>& L, const CensusRun& cr,
                                              CensusLocal& loc) {
  const bool cen = ended && L.desc == kCensus;
  // post_process_photons (src/post_process_functions.h:32-109).  Counted
  // with selects: an if-chain over the three counters becomes a
  // pointer-selected increment that puts them in scratch memory.
  const bool ex = ended && L.desc == kExit, kl = ended && L.desc == kKilled;
  loc.exit_E += ex ? (double)L.E : 0.0;
  if (cells.cell_inst) {  // several problems: each instance's exit energy
    const CensusOut co = arg_census();
    uint64_t pending = co.inst_exit ? __ballot(ex) : 0ull;
    if (pending) {
      // one atomic per instance among this iteration's exits (a wave's
      // photons mostly share an instance), not one per photon: per-photon
      // atomics on a few hundred addresses serialise
      const uint32_t inst = ex ? cells.cell_inst[L.cell] : 0u;
      const int lane = (int)(threadIdx.x & (kWave - 1));
      while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t li = (uint32_t)__shfl((int)inst, leader, kWave);
        const bool mine = ex && inst == li;
        const double v = wave_sum_f64(mine ? (double)L.E : 0.0);
        if (lane == leader) atomicAdd(&co.inst_exit[li], v);
        pending &= ~__ballot(mine);
      }
    }
  }
  loc.n_exit += (uint32_t)__popcll(__ballot(ex));
  loc.n_killed += (uint32_t)__popcll(__ballot(kl));
  loc.n_other += (uint32_t)__popcll(__ballot(ended && !cen && !ex && !kl));
  const uint64_t m = __ballot(cen);
  if (m == 0) return;
  const CensusOut co = arg_census();
  const PhotonSoA W = arg_stores();
  const uint32_t cnt = (uint32_t)__popcll(m);
  const uint32_t fill = kDense ? 0u : loc.cfill;
  const bool take = kDense || fill + cnt > kCenChunk;  // wave-uniform
  uint64_t fresh = 0;
  if (take) {
    const int leader = __ffsll((long long)m) - 1;
    if ((int)(threadIdx.x & (kWave - 1)) == leader)
      fresh = atomicAdd(&co.acc->n_reserved, (unsigned long long)(kDense ? cnt : kCenChunk));
    fresh = shfl_u64(fresh, leader);
    fresh = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(fresh >> 32)) << 32) |
            __builtin_amdgcn_readfirstlane((uint32_t)fresh);  // wave-uniform
  }
  if (cen) {
    const uint32_t q = fill + lane_rank(m);
    const uint64_t s = kDense ? fresh + q
                              : (q < kCenChunk ? loc.cbase + q : fresh + (q - kCenChunk));
    const uint64_t t = (i >= cr.n_cen) ? i - cr.n_cen : cr.n_new + W.key[i];
    if (s < co.slots) {
      store_live<R, kLat>(cells, co.C, s, L);
      co.C.life[s] = cr.life;
      co.C.E0[s] = L.E0;
      co.C.stream[s] = L.rng.stream;
      co.C.stype[s] = kLiveType ? L.stype : W.stype[i];
      co.C.key[s] = (uint32_t)t;
      // the photon's mark at its transport index gives its rank in the
      // census order after the launch; every key is set by its producer
      if (t < cr.n_cen + cr.n_new) co.mark[t] = 1;
    } else {
      atomicAdd(&co.acc->overflow, 1ull);
    }
  }
  loc.n_cen += cnt;
  if constexpr (!kDense) {
    if (take) {
      loc.cbase = fresh;
      loc.cfill = fill + cnt - kCenChunk;
    } else {
      loc.cfill = fill + cnt;
    }
  }
}

static_assert(kCenChunk == kWave, "a chunk's unused tail is marked one slot per lane");

// End of a launch: the wave's exit energy and counts into the launch totals,
// its last chunk's unused tail marked and recorded as holes.
__device__ __forceinline__ void census_flush(const CensusLocal& loc) {
  const CensusOut co = arg_census();
  double e = loc.exit_E;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) e += __shfl_xor(e, d, kWave);
  const unsigned long long nx = loc.n_exit, nk = loc.n_killed, no = loc.n_other;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const bool partial = loc.cfill < kCenChunk;  // wave-uniform
  if (partial && lane >= loc.cfill && loc.cbase + lane < co.slots)
    co.C.key[loc.cbase + lane] = kHole;  // the unused tail of the last chunk
  if (lane == 0) {
    if (nx) {
      atomicAdd(&co.acc->exit_E, e);
      atomicAdd(&co.acc->n_exit, nx);
    }
    if (nk) atomicAdd(&co.acc->n_killed, nk);
    if (no) atomicAdd(&co.acc->n_other, no);
    if (loc.n_cen) atomicAdd(&co.acc->n_census, (unsigned long long)loc.n_cen);
    if (partial) {
      const unsigned long long r = atomicAdd(&co.acc->n_rec, 1ull);
      if (r < co.rec_cap)
        co.rec[r] = HoleRec{loc.cbase + loc.cfill, (uint64_t)(kCenChunk - loc.cfill)};
      else
        atomicAdd(&co.acc->overflow, 1ull);
    }
  }
}

template <class Tally, typename R, int kLat>
BH_TRACK_KERNEL void k_transport_history(const PhotonSoA W, const CensusOut co,
                                         DevCells cells, uint64_t n, uint64_t ctr_hi,
                                         Tally tally, DevCounters* ctrs, ReplayList rl) {
  if (!args_ok(W, co, ctrs)) return;
  CensusLocal loc;
  const CensusRun cr = census_run(co, nullptr);
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool active = i < n;
  unsigned long long flushes = 0, segs = 0, stuck = 0;
  Live<R> L;
  L.rng = Rng{0, 0, 0};
  L.pending = false;
  if (active) load_live<R, kLat>(cells, W, i, ctr_hi, L);
  while (__ballot(active) != 0) {
    Flush fo;
    fo.has = false;
    fo.kill = 0.0;
    bool replay = false, ended = false;
    // the cell's record is loaded before the draws (cell 0 for an idle lane)
    const CellView<kLat> cv(cells, active ? L.cell : 0u, L.ijk);
    WaveDraws dr;
    wave_draws(active, L, dr);
    if (active && segment<Tally::kTrack, kLat>(cells, L, fo, nullptr, &dr, &cv)) {
      stuck = (L.desc == kStuck);
      segs = end_history<R, kLat>(cells, arg_stores(), i, L, replay, !co.on);
      active = false;
      ended = true;
    }
    if constexpr (sizeof(R) == 4) replay_append(replay, i, L.first_seg, L.E_handoff, rl, ctrs);
    if (co.on) census_append<R, kLat>(ended && !replay, cells, i, L, cr, loc);
    flushes += fo.has;
    tally.push(fo);
  }
  tally.finish(cells.n_cells);
  add_counters(ctrs, segs, flushes, stuck);
  if (co.on) census_flush(loc);
}

// Persistent form: each wave keeps 64 live photons and refills idle lanes
// from a chunk of photon indices it owns; a chunk of `chunk` indices is
// claimed with one atomic on the global queue head, so the queue sees one
// atomic per chunk histories instead of one per refill.  The loop ends
// when the queue is drained and no lane holds a photon, so every wave exits.
// With sd.on, the slots from the census count on are the step's new
// photons, made here when a lane takes them (the source's draws computed by
// the whole wave, then the source's sampling) instead of being read.
// The generating instances (kGen) hold the source sampling's temporaries on
// top of the tracking state: at 3 waves per SIMD they spilled 216-248 B per
// lane; at 2 they keep everything in 205-210 VGPRs and measured 12% faster.
template <class Tally, typename R, int kLat, bool kGen>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(kGen ? 2 : BH_TRACK_WAVES_PER_EU)))
void k_transport_persistent(const PhotonSoA W, const CensusOut co,
                                            const SourceDesc sd, DevCells cells, uint64_t n,
                                            uint64_t ctr_hi, Tally tally, DevCounters* ctrs,
                                            PersistentQueue* q, int refill, uint32_t chunk,
                                            ReplayList rl, const StepParams* P) {
  if constexpr (kGen) {
    if (!args_ok(W, co, sd, ctrs)) return;
  } else {
    if (!args_ok(W, co, ctrs)) return;
  }
  const uint64_t kChunk = chunk;
  if (P) n = P->n;  // a step captured as a graph
  const CensusRun cr = census_run(co, P);
  const uint64_t first_gen = (kGen && sd.on) ? cr.n_cen : ~0ull;
  CensusLocal loc;
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const int lane = threadIdx.x & (kWave - 1);
  Live<R> L;
  L.rng = Rng{0, 0, 0};
  L.pending = false;
  bool active = false, gen = false;
  uint64_t idx = 0;
  uint64_t chunk_next = 0, chunk_end = 0;  // wave-uniform
  bool drained = false;
  // segments per lane; flushes and stuck photons counted per wave (scalar)
  uint32_t segs = 0;
  unsigned long long flushes = 0, stuck = 0;
  while (true) {
    uint64_t idle = __ballot(!active);
    int n_idle = __popcll(idle);
    if (!drained && (n_idle >= refill || n_idle == kWave)) {
      // hand out indices from the wave's chunk, claiming chunks as needed
      while (n_idle > 0 && !drained) {
        if (chunk_next >= chunk_end) {
          uint64_t base = 0, end = 0;
          if (lane == 0) base = claim_chunk(q, n, kChunk, end);
          base = shfl_u64(base, 0);
          end = shfl_u64(end, 0);
          if (base >= n) {
            drained = true;
            break;
          }
          chunk_next = base;
          chunk_end = end;
        }
        const uint64_t avail = chunk_end - chunk_next;
        const uint64_t rank = lane_rank(idle);
        const bool take = !active && rank < avail;
        if (take) {
          idx = chunk_next + rank;
          gen = kGen && idx >= first_gen;
          if (!gen) load_live<R, kLat>(cells, arg_stores(), idx, ctr_hi, L);
          active = true;
        }
        const uint64_t taken = (avail < (uint64_t)n_idle) ? avail : (uint64_t)n_idle;
        chunk_next += taken;
        idle = __ballot(!active);
        n_idle = __popcll(idle);
      }
      if (kGen && __ballot(gen) != 0) {  // the new photons this refill took
        const SourceDesc s = arg_source();
        uint64_t p = 0, stream = 0, hi = 0;
        uint32_t c = 0;
        int nd = 0;
        SourceInst sp{};
        if (gen) {
          p = idx - first_gen;
          c = hinted_cell(s, cells.n_cells, cr.n_new, p);
          sp = source_record(cells, c, s.one, s.inst);
          nd = source_draw_count(s.sc, s.off, p, c, 0);
          stream = source_stream(sp, s.off, p);
          hi = sp.ctr_hi;
        }
        WaveDraws gd;
        wave_gen_draws(gen, nd, stream, hi, gd);
        if (gen) {
          uint32_t j = gd.next;
          const SourcePhoton o = make_source_photon(cells, s.sc, s.face, s.off, p, c, sp,
                                                    P ? P->dt : s.dt, 0,
                                                    [&]() { return gd.res[j++]; });
          live_from_source<R, kLat>(cells, o, hi, L);
        }
        gen = false;
      }
    }
    if (__ballot(active) == 0) break;
    Flush fo;
    fo.has = false;
    fo.kill = 0.0;
    bool replay = false, ended = false;
    // the cell's record is loaded before the draws (cell 0 for an idle lane)
    const CellView<kLat> cv(cells, active ? L.cell : 0u, L.ijk);
    WaveDraws dr;
    wave_draws(active, L, dr);
    if (active && segment<Tally::kTrack, kLat>(cells, L, fo, nullptr, &dr, &cv)) {
      segs += (uint32_t)end_history<R, kLat>(cells, arg_stores(), idx, L, replay, !co.on);
      active = false;
      ended = true;
    }
    stuck += (unsigned long long)__popcll(__ballot(ended && L.desc == kStuck));
    if constexpr (sizeof(R) == 4) replay_append(replay, idx, L.first_seg, L.E_handoff, rl, ctrs);
    if (co.on) census_append<R, kLat, false, kGen>(ended && !replay, cells, idx, L, cr, loc);
    flushes += (unsigned long long)__popcll(__ballot(fo.has));
    tally.push(fo);
  }
  tally.finish(cells.n_cells);
  add_wave_counters(ctrs, segs, flushes, stuck);
  if (co.on) census_flush(loc);
}

// ---- Event-split form -------------------------------------------------------
// The reference's event-based algorithm (src/event_based_transport.h:
// 262-305: precompute, distances, state update, per-event processing, then
// compaction of the active list) on the device, with the history form's
// exact per-photon arithmetic and draw order, so both give the same events:
//   k_ev_flight   one segment of every active photon: opacities, its
//                 distance draw, distances to scatter / face / census,
//                 implicit-capture deposit, move, then the event -- a face
//                 crossing or reflection, census, exit and kill are handled
//                 here; an effective scatter is queued;
//   k_ev_scatter  every queued scatter at full width: new direction, the
//                 re-emission test and the group walk (4-5 draws);
// and each pass writes the photons still alive into the next active list
// with one atomic per wave (ballot, then each lane's rank among the set
// lanes): the live photons stay packed, so every wave of the next pass is
// full.  Between passes a photon's open cell visit (its accumulated
// deposits), segment count and last face live in EventState.
struct EventState {
  double* t_abs;
  double* t_track;
  uint32_t* segs;
  uint8_t* sc;
  uint32_t* list[2];           // active photon slots (ping-pong)
  uint32_t* scatter;           // slots whose flight ended in an effective scatter
  unsigned long long* count;   // [0], [1]: active counts; [2]: scatter count
};

// Wave-aggregated append of the lanes with `want` set: one atomic per wave.
// Every lane of the wave at the same program point.
__device__ __forceinline__ void list_append(bool want, uint32_t v, uint32_t* __restrict__ list,
                                            unsigned long long* n) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(n, (unsigned long long)__popcll(m));
  base = shfl_u64(base, leader);
  if (want) list[base + lane_rank(m)] = v;
}

// Active list = every photon; open visits empty.
__global__ void k_ev_init(EventState ev, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    ev.count[0] = n;
    ev.count[1] = 0;
    ev.count[2] = 0;
  }
  if (i >= n) return;
  ev.list[0][i] = (uint32_t)i;
  ev.t_abs[i] = 0.0;
  ev.t_track[i] = 0.0;
  ev.segs[i] = 0;
  ev.sc[i] = 0;
}

// Before a pass pair: the lists it fills start empty.
__global__ void k_ev_clear(EventState ev, int nxt) {
  ev.count[nxt] = 0;
  ev.count[2] = 0;
}

template <class Tally, int kLat>
BH_TRACK_KERNEL void k_ev_flight(const PhotonSoA W, const CensusOut co, DevCells cells,
                                 EventState ev, int cur, uint64_t ctr_hi, Tally tally,
                                 DevCounters* ctrs) {
  if (!args_ok(W, co, ctrs)) return;
  const CensusRun cr = census_run(co, nullptr);
  CensusLocal loc;
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const uint64_t n = ev.count[cur];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long flushes = 0, segs = 0, stuck = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint64_t j = base + threadIdx.x;
    const bool active = j < n;
    const uint32_t i = active ? ev.list[cur][j] : 0u;
    Live<double> L;
    L.rng = Rng{0, 0, 0};
    L.pending = false;
    if (active) {
      load_live<double, kLat>(cells, arg_stores(), i, ctr_hi, L);
      L.t_abs = ev.t_abs[i];
      L.t_track = ev.t_track[i];
      L.segs = ev.segs[i];
      L.sc = ev.sc[i];
    }
    Flush fo;
    fo.has = false;
    fo.kill = 0.0;
    bool ended = false;
    const CellView<kLat> cv(cells, active ? L.cell : 0u, L.ijk);
    if (active && segment<Tally::kTrack, kLat>(cells, L, fo, nullptr, nullptr, &cv)) {
      ended = true;
      if (!co.on) store_live<double, kLat>(cells, arg_stores(), i, L);
    }
    const bool alive = active && !ended;
    if (alive) {
      store_live<double, kLat>(cells, arg_stores(), i, L);
      ev.t_abs[i] = L.t_abs;
      ev.t_track[i] = L.t_track;
      ev.segs[i] = L.segs;
      ev.sc[i] = (uint8_t)L.sc;
    }
    segs += (unsigned long long)wave_sum_u64(ended ? L.segs : 0u);
    stuck += (unsigned long long)__popcll(__ballot(ended && L.desc == kStuck));
    list_append(alive && L.pending, i, ev.scatter, &ev.count[2]);
    list_append(alive && !L.pending, i, ev.list[cur ^ 1], &ev.count[cur ^ 1]);
    if (co.on) census_append<double, kLat, true>(ended, cells, i, L, cr, loc);
    flushes += (unsigned long long)__popcll(__ballot(fo.has));
    tally.push(fo);
  }
  tally.finish(cells.n_cells);
  if ((threadIdx.x & (kWave - 1)) == 0) {  // segs already wave sums
    atomicAdd(&ctrs->segments, segs);
    atomicAdd(&ctrs->flushes, flushes);
    if (stuck) atomicAdd(&ctrs->stuck, stuck);
  }
  if (co.on) census_flush(loc);
}

template <int kLat>
__global__ __launch_bounds__(256) void k_ev_scatter(PhotonSoA W, DevCells cells, EventState ev,
                                                    int cur, uint64_t ctr_hi) {
  const uint64_t n = ev.count[2];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint64_t j = base + threadIdx.x;
    const bool active = j < n;
    uint32_t i = 0;
    if (active) {
      i = ev.scatter[j];
      Live<double> L;
      L.cell = W.cell[i];
      L.group = W.group[i];
      if constexpr (kLat) lattice_ijk<kLat>(cells, L.cell, L.ijk);
      L.rng.ctr = W.ctr[i];
      L.rng.stream = W.stream[i];
      L.rng.ctr_hi = photon_ctr_hi(cells, L.cell, ctr_hi);
      L.pending = true;
      const CellView<kLat> cv(cells, L.cell, L.ijk);
      double2 sg = cv.sigma(L.group);
      double total_sigma_s = (1.0 - cv.f) * sg.x + sg.y;
      (void)lane_draws<false, kLat>(cells, cv, cv.f, L, sg, total_sigma_s, nullptr, false);
      W.u[i] = L.u;
      W.v[i] = L.v;
      W.w[i] = L.w;
      W.group[i] = L.group;
      W.ctr[i] = L.rng.ctr;
    }
    list_append(active, i, ev.list[cur ^ 1], &ev.count[cur ^ 1]);
  }
}

// FP64 replay of the histories the FP32 form stopped: each starts again
// from its initial state, and the deposits and flushes of the segments
// before first_seg -- already tallied in FP32 -- are muted; the
// accumulators restart at first_seg, whose partial cell visit the FP32 pass
// flushed.  From first_seg on, deposits and the final weight are scaled by
// s = (FP32 weight) / (FP64 weight) at that point, a factor within rounding
// of 1: FP32 deposits + s * (FP64 deposits + final weight) then equal the
// initial weight, so conservation holds to round-off, while every event
// (the cutoff test included) still uses the unscaled FP64 weight.  The list
// length is read on the device, so no host round trip separates the two
// launches.  Loop bounds are block-uniform.
// kGen: the new photons of the step were made in the transport kernel (the
// source description is on), so new ones are made again here; a separate
// instance keeps that path's registers out of the plain replay.
template <class Tally, bool kGen>
BH_TRACK_KERNEL void k_transport_replay(const PhotonSoA W, const CensusOut co,
                                        const SourceDesc sd, DevCells cells, ReplayList rl,
                                        uint64_t ctr_hi, Tally tally, DevCounters* ctrs,
                                        const StepParams* P) {
  if constexpr (kGen) {
    if (!args_ok(W, co, sd, ctrs)) return;
  } else {
    if (!args_ok(W, co, ctrs)) return;
  }
  const CensusRun cr = census_run(co, P);
  const uint64_t first_gen = (kGen && sd.on) ? cr.n_cen : ~0ull;
  CensusLocal loc;
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const uint64_t n = ctrs->replays;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long flushes = 0, segs = 0, stuck = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint64_t j = base + threadIdx.x;
    bool active = j < n;
    uint64_t i = 0;
    Live<double> L;
    if (active) {
      i = rl.idx[j];
      if (kGen && i >= first_gen)  // a new photon: made again from its draws
        gen_photon_lane<double, false>(cells, arg_source(), cr.n_new, P ? P->dt : sd.dt,
                                       i - first_gen, L);
      else
        load_live(cells, arg_stores(), i, ctr_hi, L);
      L.first_seg = rl.first_seg[j];
      L.E_handoff = rl.E[j];
    }
    while (__ballot(active) != 0) {
      Flush fo;
      fo.has = false;
      fo.kill = 0.0;
      bool ended = false;
      if (active) {
        if (L.segs + 1 == L.first_seg) {
          L.t_abs = 0.0;
          L.t_track = 0.0;
          L.E_handoff /= L.E;  // from here on: the weight scale s
        }
        const bool end = segment<Tally::kTrack>(cells, L, fo);
        if (L.segs < L.first_seg) {
          fo.has = false;
          fo.kill = 0.0;
        } else {
          fo.cont *= L.E_handoff;
          fo.track *= L.E_handoff;
          fo.kill *= L.E_handoff;
        }
        if (end) {
          stuck += (L.desc == kStuck);
          segs += L.segs + 1 - L.first_seg;
          L.E *= L.E_handoff;
          if (!co.on) store_live(cells, arg_stores(), i, L);
          active = false;
          ended = true;
        }
      }
      if (co.on) census_append<double, false, false, kGen>(ended, cells, i, L, cr, loc);
      flushes += fo.has;
      tally.push(fo);
    }
  }
  tally.finish(cells.n_cells);
  add_counters(ctrs, segs, flushes, stuck);
  if (co.on) census_flush(loc);
}

// Path-replay adjoint of the absorbed-energy tally.  For a loss L with
// g[c] = dL/d abs_E[c], the likelihood-ratio (differential operator)
// derivative of the expected tallies, source photons held fixed, is
//   dL/dtheta = sum over histories of
//     sum over decisions e:  d log p_e / dtheta * R_e
//   + sum over segments j:   dtau_j / dtheta * (g_c E_out_j - R_after_j),
// where R_e is the g-weighted score of the history from decision e on (the
// score of the segment a flight length decides included), R_after_j the
// g-weighted score after segment j (a cutoff deposit is the weight left
// after its segment, so it counts as "after"), and tau_j = sa f d the
// segment's optical depth.  The decisions that depend on a cell's sigma_a
// and Fleck factor f (sigma_eff = (1 - f) sa + ss):
//   flight:   collision at d: d log p = dsigma_eff (1/sigma_eff - d);
//             no collision within d: d log p = -dsigma_eff d;
//   re-emission test: group kept with probability ss / sigma_eff,
//             re-sampled with (1 - f) sa / sigma_eff;
// dsigma_eff/dsa = 1 - f, dsigma_eff/df = -sa, dtau/dsa = f d,
// dtau/df = sa d.  (Group-independent sigma_a: the group walk's
// probabilities then do not depend on the parameters.  The cutoff test is a
// discontinuity and contributes nothing.)
//
// No trajectory is stored: pass 1 runs each history from its initial state
// (its Random123 counter) to get its total R; pass 2 replays it with the
// same draws and, knowing the score accumulated so far, forms every R_e as
// total minus accumulated.  Per-cell gradients are summed over a cell visit
// and flushed through the tally sink: sink "abs" = dL/dsigma_a, sink
// "track" = dL/dfleck.
template <class Sink>
BH_TRACK_KERNEL void k_adjoint(DevCells cells, PhotonSoA W, uint64_t n, uint64_t ctr_hi,
                               const double* g, Sink sink, DevCounters* ctrs) {
  bind_lds(sink, cells.n_cells);
  sink.init(cells.n_cells);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long segs = 0, flushes = 0, stuck = 0;
  Live<double> L;
  SegInfo si;
  // pass 1: total g-weighted score of the history
  bool active = i < n;
  double R = 0.0;
  if (active) load_live(cells, W, i, ctr_hi, L);
  while (active) {
    Flush fo;
    fo.has = false;
    const bool end = segment<false, false, true>(cells, L, fo, &si);
    R += g[si.c] * (si.absorbed + si.kill_E);
    if (end) {
      stuck = (L.desc == kStuck);
      active = false;
    }
  }
  // pass 2: replay and accumulate the gradient
  active = i < n;
  double A = 0.0, gs = 0.0, gf = 0.0;
  if (active) load_live(cells, W, i, ctr_hi, L);
  while (__ballot(active) != 0) {
    Flush out;
    out.has = false;
    out.kill = 0.0;
    if (active) {
      Flush fo;
      fo.has = false;
      const bool end = segment<false, false, true>(cells, L, fo, &si);
      const double gc = g[si.c];
      if (si.reemit) {  // decided before this segment's flight
        const double Rr = R - A;
        const double a1 = 1.0 - si.r_f;
        if (si.reemit == 2) {  // re-sampled: p = (1 - f) sa / sigma_eff
          gs += (1.0 / si.r_sa - a1 / si.r_sig) * Rr;
          gf += (-1.0 / a1 + si.r_sa / si.r_sig) * Rr;
        } else {  // kept: p = ss / sigma_eff
          gs += (-a1 / si.r_sig) * Rr;
          gf += (si.r_sa / si.r_sig) * Rr;
        }
      }
      const double s_own = gc * si.absorbed;
      const double Rj = R - A;
      const double Rafter = Rj - s_own;
      if (si.flight) {
        const double w = si.scattered ? (1.0 / si.sig - si.d) : -si.d;
        gs += (1.0 - si.f) * w * Rj;
        gf += -si.sa * w * Rj;
      }
      const double att = gc * (si.E_in - si.absorbed) - Rafter;
      gs += si.f * si.d * att;
      gf += si.sa * si.d * att;
      A += s_own + gc * si.kill_E;
      if (end || L.cell != si.c) {
        out.has = true;
        out.c = si.c;
        out.cont = gs;
        out.track = gf;
        gs = 0.0;
        gf = 0.0;
      }
      if (end) {
        segs = L.segs;
        active = false;
      }
    }
    flushes += out.has;
    sink.push(out);
  }
  sink.finish(cells.n_cells);
  add_counters(ctrs, segs, flushes, stuck);
}

}  // namespace bh
