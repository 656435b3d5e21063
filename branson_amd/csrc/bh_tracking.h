// bh_tracking.h -- the photon tracking loop of the MI355X IMC engine.
//
// One segment = one iteration of `while (active)` in transport_photon
// (src/history_based_transport.h:55-140): opacity fetch, RNG draw, distances
// to scatter / face / census, implicit-capture decay, move, then the event.
// segment() advances a live photon by exactly one segment and reports
// whether its history ended, so the same code drives
//   * k_transport_history    -- one work-item per photon, loop to the end;
//   * k_transport_persistent -- a grid sized to the chip whose lanes refill
//     from a global queue whenever enough of a wave's lanes are idle, so
//     long and short histories no longer leave most lanes of a wave waiting.
// Precision forms:
//   FP64 (Real = double): statement-for-statement replay of the reference;
//   FP32 tracking (Real = float): position, direction, path lengths and the
//     decay exponent in FP32; energy weights, tallies and the re-emission
//     test stay FP64 so energy conservation is unchanged.
#pragma once

#include "bh_kernels.h"

namespace bh {

struct PersistentQueue {
  unsigned long long next;
  unsigned long long pad[7];
};

template <typename R>
struct Live {
  R x, y, z, u, v, w, life;
  double E, E0;
  double t_abs, t_track;
  Rng rng;
  uint32_t cell, group, sc, segs;
  uint8_t desc;
};

template <typename R>
__device__ __forceinline__ void load_live(const PhotonSoA& W, uint64_t i, uint64_t ctr_hi,
                                          Live<R>& L) {
  L.x = (R)W.x[i];
  L.y = (R)W.y[i];
  L.z = (R)W.z[i];
  L.u = (R)W.u[i];
  L.v = (R)W.v[i];
  L.w = (R)W.w[i];
  L.life = (R)W.life[i];
  L.E = W.E[i];
  L.E0 = W.E0[i];
  L.rng.ctr = W.ctr[i];
  L.rng.stream = W.stream[i];
  L.rng.ctr_hi = ctr_hi;
  L.cell = W.cell[i];
  L.group = W.group[i];
  L.desc = W.desc[i];
  L.sc = 0;  // surface_cross starts at 0 for every photon (:43)
  L.segs = 0;
  L.t_abs = 0.0;
  L.t_track = 0.0;
}

template <typename R>
__device__ __forceinline__ void store_live(const PhotonSoA& W, uint64_t i, const Live<R>& L) {
  W.x[i] = (double)L.x;
  W.y[i] = (double)L.y;
  W.z[i] = (double)L.z;
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

template <class Tally, typename R>
__device__ __forceinline__ void flush(Tally& tally, Live<R>& L, uint32_t& flushes) {
  tally.add(L.cell, L.t_abs, L.t_track);
  ++flushes;
}

// Boundary event shared by both precisions (:102-131).  Returns true when
// the history ends.
template <class Tally, typename R>
__device__ __forceinline__ bool boundary_event(const DevCells& cells, const CellGeo& geo,
                                               Live<R>& L, Tally& tally, uint32_t& flushes) {
  const int bc = face_bc(geo.bc_bits, L.sc);
  if (bc == kElement) {
    flush(tally, L, flushes);
    L.cell = cells.next[(size_t)L.cell * 6 + L.sc];
    L.desc = kBound;
    L.t_abs = 0.0;
    L.t_track = 0.0;
    return false;
  }
  if (bc == kProcessor) {
    const uint32_t nc = cells.next[(size_t)L.cell * 6 + L.sc];
    L.desc = kPass;
    flush(tally, L, flushes);
    L.cell = nc;
    return true;
  }
  if (bc == kVacuum || bc == kSource) {
    L.desc = kExit;
    flush(tally, L, flushes);
    return true;
  }
  // REFLECT: Photon::reflect (src/photon.h:150-155)
  if (L.sc < 2) L.u = -L.u;
  else if (L.sc < 4) L.v = -L.v;
  else L.w = -L.w;
  L.desc = kBound;
  return false;
}

// ---- FP64: the reference's arithmetic, operation for operation ----------
template <class Tally>
__device__ __forceinline__ bool segment(const DevCells& cells, Live<double>& L, Tally& tally,
                                        uint32_t& flushes) {
  if (++L.segs > kMaxSegments) {
    flush(tally, L, flushes);
    L.desc = kStuck;
    return true;
  }
  const uint32_t G = cells.G;
  const double2 sg = cells.sigma[(size_t)L.cell * G + L.group];
  const double sigma_a = sg.x, sigma_s = sg.y;
  const CellGeo& geo = cells.geo[L.cell];
  const double f = geo.fleck;
  const double total_sigma_s = (1.0 - f) * sigma_a + sigma_s;

  const double d_s = (total_sigma_s > 0.0) ? -log(L.rng.next()) / total_sigma_s : 1.0e100;

  // Cell::get_distance_to_boundary (src/cell.h:116-132), sgn(a) = (0 < a)
  double d_b = 1.0e16;
  {
    const uint32_t ix = (0.0 < L.u) ? 1u : 0u;
    const double dx = (geo.nodes[ix] - L.x) / L.u;
    if (dx < d_b) { d_b = dx; L.sc = ix; }
    const uint32_t iy = (0.0 < L.v) ? 3u : 2u;
    const double dy = (geo.nodes[iy] - L.y) / L.v;
    if (dy < d_b) { d_b = dy; L.sc = iy; }
    const uint32_t iz = (0.0 < L.w) ? 5u : 4u;
    const double dz = (geo.nodes[iz] - L.z) / L.w;
    if (dz < d_b) { d_b = dz; L.sc = iz; }
  }
  const double d_c = L.life;
  const double m_bc = (d_c < d_b) ? d_c : d_b;  // std::min(d_b, d_c)
  const double d = (m_bc < d_s) ? m_bc : d_s;    // std::min(d_s, .)

  const double absorbed = L.E * (1.0 - exp(-sigma_a * f * d));
  L.t_abs += absorbed;
  L.t_track += absorbed / (sigma_a * f);
  L.E = L.E - absorbed;

  L.x += L.u * d;  // Photon::move (src/photon.h:107-114)
  L.y += L.v * d;
  L.z += L.w * d;
  L.life -= d;

  if (L.E / L.E0 < kCutoff) {  // kill, deposit the remainder (:85-91)
    L.t_abs += L.E;
    flush(tally, L, flushes);
    L.desc = kKilled;
    return true;
  }
  if (d == d_s) {  // effective scatter (:95-100)
    uniform_angle(L.rng, L.u, L.v, L.w);
    if (L.rng.next() > (sigma_s / ((1.0 - f) * sigma_a + sigma_s)))
      L.group = emission_group(&cells.sigma[(size_t)L.cell * G], G, L.rng);
    L.desc = kScatter;
    return false;
  }
  if (d == d_b) return boundary_event(cells, geo, L, tally, flushes);
  if (d == d_c) {  // census (:133-138)
    L.desc = kCensus;
    flush(tally, L, flushes);
    return true;
  }
  return false;
}

// ---- FP32 tracking ------------------------------------------------------
template <class Tally>
__device__ __forceinline__ bool segment(const DevCells& cells, Live<float>& L, Tally& tally,
                                        uint32_t& flushes) {
  if (++L.segs > kMaxSegments) {
    flush(tally, L, flushes);
    L.desc = kStuck;
    return true;
  }
  const uint32_t G = cells.G;
  const double2 sg = cells.sigma[(size_t)L.cell * G + L.group];
  const double sigma_a = sg.x, sigma_s = sg.y;
  const CellGeo& geo = cells.geo[L.cell];
  const double f = geo.fleck;
  const double saf = sigma_a * f;
  const float tot = (float)((1.0 - f) * sigma_a + sigma_s);

  float d_s = 1.0e30f;
  if (tot > 0.0f) {
    float uf = (float)L.rng.next();
    uf = (uf < 1.0f) ? uf : 0x1.fffffep-1f;  // (0,1) stays open after rounding
    d_s = -__logf(uf) / tot;
  }
  float d_b = 1.0e16f;
  {
    const uint32_t ix = (0.0f < L.u) ? 1u : 0u;
    const float dx = ((float)geo.nodes[ix] - L.x) / L.u;
    if (dx < d_b) { d_b = dx; L.sc = ix; }
    const uint32_t iy = (0.0f < L.v) ? 3u : 2u;
    const float dy = ((float)geo.nodes[iy] - L.y) / L.v;
    if (dy < d_b) { d_b = dy; L.sc = iy; }
    const uint32_t iz = (0.0f < L.w) ? 5u : 4u;
    const float dz = ((float)geo.nodes[iz] - L.z) / L.w;
    if (dz < d_b) { d_b = dz; L.sc = iz; }
  }
  const float d_c = L.life;
  const float m_bc = (d_c < d_b) ? d_c : d_b;
  const float d = (m_bc < d_s) ? m_bc : d_s;

  // 1 - exp(-x) as -expm1(-x): no cancellation for the small optical depths
  // of a single segment; the weight and tallies stay FP64.
  const float frac = -expm1f(-(float)saf * d);
  const double absorbed = L.E * (double)frac;
  L.t_abs += absorbed;
  L.t_track += absorbed / saf;
  L.E = L.E - absorbed;

  L.x += L.u * d;
  L.y += L.v * d;
  L.z += L.w * d;
  L.life -= d;

  if (L.E / L.E0 < kCutoff) {
    L.t_abs += L.E;
    flush(tally, L, flushes);
    L.desc = kKilled;
    return true;
  }
  if (d == d_s) {
    // sin(theta) from the FP64 cosine: in FP32, mu can round to +-1 and
    // leave a direction with two exactly-zero components, whose face
    // distances are -inf.  In FP64 sqrt(1 - mu^2) >= 2^-25.5 is never zero.
    const double mu_d = L.rng.next() * 2.0 - 1.0;
    const float phi = (float)(L.rng.next() * 2.0 * kPi);
    const float st = (float)sqrt(1.0 - mu_d * mu_d);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    L.u = st * cp;
    L.v = st * sp;
    L.w = (float)mu_d;
    if (L.rng.next() > (sigma_s / ((1.0 - f) * sigma_a + sigma_s)))
      L.group = emission_group(&cells.sigma[(size_t)L.cell * G], G, L.rng);
    L.desc = kScatter;
    return false;
  }
  if (d == d_b) return boundary_event(cells, geo, L, tally, flushes);
  if (d == d_c) {
    L.desc = kCensus;
    flush(tally, L, flushes);
    return true;
  }
  return false;
}

__device__ __forceinline__ void add_counters(DevCounters* ctrs, unsigned long long segs,
                                             unsigned long long flushes,
                                             unsigned long long stuck) {
  __shared__ unsigned long long lds[4];
  const unsigned long long s = block_sum<unsigned long long, 256>(segs, lds);
  const unsigned long long f = block_sum<unsigned long long, 256>(flushes, lds);
  const unsigned long long k = block_sum<unsigned long long, 256>(stuck, lds);
  if (threadIdx.x == 0) {
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
// (src/history_based_transport.h:263-275) over SoA state.
template <class Tally, typename R>
__global__ __launch_bounds__(256) void k_transport_history(DevCells cells, PhotonSoA W,
                                                           uint64_t n, uint64_t ctr_hi,
                                                           Tally tally, DevCounters* ctrs) {
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t flushes = 0;
  unsigned long long segs = 0, stuck = 0;
  if (i < n) {
    Live<R> L;
    load_live(W, i, ctr_hi, L);
    while (!segment(cells, L, tally, flushes)) {
    }
    segs = L.segs;
    stuck = (L.desc == kStuck);
    store_live(W, i, L);
  }
  tally.finish(cells.n_cells);
  add_counters(ctrs, segs, flushes, stuck);
}

// Persistent form: each wave keeps 64 live photons; when at least
// `refill` lanes are idle (or all are), the wave claims that many photon
// indices with one atomic on the queue head and the idle lanes load them in
// order (ballot + mbcnt).  The loop ends when the queue is drained and no
// lane holds a photon, so every wave exits.
template <class Tally, typename R>
__global__ __launch_bounds__(256) void k_transport_persistent(DevCells cells, PhotonSoA W,
                                                              uint64_t n, uint64_t ctr_hi,
                                                              Tally tally, DevCounters* ctrs,
                                                              PersistentQueue* q,
                                                              int refill) {
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const int lane = threadIdx.x & (kWave - 1);
  Live<R> L;
  bool active = false;
  bool drained = false;
  uint64_t idx = 0;
  uint32_t flushes = 0;
  unsigned long long segs = 0, stuck = 0;
  while (true) {
    const uint64_t idle = __ballot(!active);
    const int n_idle = __popcll(idle);
    if (!drained && (n_idle >= refill || n_idle == kWave)) {
      const int leader = __ffsll((long long)idle) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(&q->next, (unsigned long long)n_idle);
      base = __shfl(base, leader, kWave);
      if (base >= n) drained = true;
      if (!active) {
        const uint64_t my = base + lane_rank(idle);
        if (my < n) {
          idx = my;
          load_live(W, idx, ctr_hi, L);
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;
    if (active && segment(cells, L, tally, flushes)) {
      segs += L.segs;
      stuck += (L.desc == kStuck);
      store_live(W, idx, L);
      active = false;
    }
  }
  tally.finish(cells.n_cells);
  add_counters(ctrs, segs, flushes, stuck);
}

}  // namespace bh
