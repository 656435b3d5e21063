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
//   FP32 tracking (Real = float): position, direction, path lengths and the
//     decay exponent in FP32; energy weights, tallies and the re-emission
//     test stay FP64 so energy conservation is unchanged.
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
  bool pending;  // an effective scatter awaits its draws
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
  L.pending = false;
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
template <typename R>
__device__ __forceinline__ bool boundary_event(const DevCells& cells, uint32_t bc_bits,
                                               Live<R>& L, Flush& fo) {
  const int bc = face_bc(bc_bits, L.sc);
  if (bc == kElement) {
    flush(fo, L);
    L.cell = cells.next[(size_t)L.cell * 6 + L.sc];
    L.desc = kBound;
    L.t_abs = 0.0;
    L.t_track = 0.0;
    return false;
  }
  if (bc == kProcessor) {
    const uint32_t nc = cells.next[(size_t)L.cell * 6 + L.sc];
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
  L.desc = kBound;
  return false;
}

// Draw states of a segment.
enum : int { kDrawMu = 0, kDrawPhi, kDrawTest, kDrawGroup, kDrawDistance, kDrawDone };

// ---- FP64: the reference's arithmetic, operation for operation ----------
template <bool kTrack>
__device__ __forceinline__ bool segment(const DevCells& cells, Live<double>& L, Flush& fo) {
  if (++L.segs > kMaxSegments) {
    flush(fo, L);
    L.desc = kStuck;
    return true;
  }
  const uint32_t G = cells.G;
  const CellGeo* geo = &cells.geo[L.cell];
  const double f = geo->fleck;
  double2 sg = cells.sigma[(size_t)L.cell * G + L.group];
  double total_sigma_s = (1.0 - f) * sg.x + sg.y;

  double d_s = 1.0e100;
  int st = L.pending ? kDrawMu : kDrawDistance;
  if (st == kDrawDistance && !(total_sigma_s > 0.0)) st = kDrawDone;
  double mu = 0.0;
  while (st != kDrawDone) {
    const double r = L.rng.next();  // the segment's only Threefry site
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
      st = (r > (sg.y / total_sigma_s)) ? kDrawGroup : kDrawDistance;
    } else if (st == kDrawGroup) {  // sample_emission_group walk (:98)
      const double2* sa = &cells.sigma[(size_t)L.cell * G];
      double cdf = r;
      int g = -1;
      const double norm = 1.0 / (sa[0].x * (double)G);
      while (cdf > 0) {
        g++;
        if ((uint32_t)g >= G) {
          g = (int)G - 1;
          break;
        }
        cdf -= sa[g].x * norm;
      }
      L.group = (uint32_t)g;
      sg = sa[g];
      total_sigma_s = (1.0 - f) * sg.x + sg.y;
      st = (total_sigma_s > 0.0) ? kDrawDistance : kDrawDone;
    } else {  // distance to scatter (:62-63)
      d_s = -log(r) / total_sigma_s;
      st = kDrawDone;
    }
  }
  L.pending = false;
  const double sigma_a = sg.x;

  // Cell::get_distance_to_boundary (src/cell.h:116-132), sgn(a) = (0 < a);
  // the face coordinate is chosen by select, not by a runtime index
  double d_b = 1.0e16;
  {
    const bool px = 0.0 < L.u, py = 0.0 < L.v, pz = 0.0 < L.w;
    const double nx = px ? geo->nodes[1] : geo->nodes[0];
    const double ny = py ? geo->nodes[3] : geo->nodes[2];
    const double nz = pz ? geo->nodes[5] : geo->nodes[4];
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
    return true;
  }
  if (d == d_s) {  // effective scatter (:95-100); draws taken next segment
    L.pending = true;
    L.desc = kScatter;
    return false;
  }
  if (d == d_b) return boundary_event(cells, geo->bc_bits, L, fo);
  if (d == d_c) {  // census (:133-138)
    L.desc = kCensus;
    flush(fo, L);
    return true;
  }
  return false;
}

// ---- FP32 tracking ------------------------------------------------------
template <bool kTrack>
__device__ __forceinline__ bool segment(const DevCells& cells, Live<float>& L, Flush& fo) {
  if (++L.segs > kMaxSegments) {
    flush(fo, L);
    L.desc = kStuck;
    return true;
  }
  const uint32_t G = cells.G;
  const CellGeo* geo = &cells.geo[L.cell];
  const double f = geo->fleck;
  double2 sg = cells.sigma[(size_t)L.cell * G + L.group];
  double total_sigma_s = (1.0 - f) * sg.x + sg.y;

  float d_s = 1.0e30f;
  int st = L.pending ? kDrawMu : kDrawDistance;
  if (st == kDrawDistance && !(total_sigma_s > 0.0)) st = kDrawDone;
  double mu = 0.0;
  while (st != kDrawDone) {
    const double r = L.rng.next();
    if (st == kDrawMu) {
      mu = r * 2.0 - 1.0;
      st = kDrawPhi;
    } else if (st == kDrawPhi) {
      // sin(theta) from the FP64 cosine: in FP32, mu can round to +-1 and
      // leave a direction with two exactly-zero components, whose face
      // distances are -inf.  In FP64 sqrt(1 - mu^2) >= 2^-25.5 is never 0.
      const float phi = (float)(r * 2.0 * kPi);
      const float s_t = (float)sqrt(1.0 - mu * mu);
      float sp, cp;
      sincosf(phi, &sp, &cp);
      L.u = s_t * cp;
      L.v = s_t * sp;
      L.w = (float)mu;
      st = kDrawTest;
    } else if (st == kDrawTest) {
      st = (r > (sg.y / total_sigma_s)) ? kDrawGroup : kDrawDistance;
    } else if (st == kDrawGroup) {
      const double2* sa = &cells.sigma[(size_t)L.cell * G];
      double cdf = r;
      int g = -1;
      const double norm = 1.0 / (sa[0].x * (double)G);
      while (cdf > 0) {
        g++;
        if ((uint32_t)g >= G) {
          g = (int)G - 1;
          break;
        }
        cdf -= sa[g].x * norm;
      }
      L.group = (uint32_t)g;
      sg = sa[g];
      total_sigma_s = (1.0 - f) * sg.x + sg.y;
      st = (total_sigma_s > 0.0) ? kDrawDistance : kDrawDone;
    } else {
      float uf = (float)r;
      uf = (uf < 1.0f) ? uf : 0x1.fffffep-1f;  // (0,1) stays open after rounding
      d_s = -__logf(uf) / (float)total_sigma_s;
      st = kDrawDone;
    }
  }
  L.pending = false;
  const double saf = sg.x * f;

  float d_b = 1.0e16f;
  {
    const bool px = 0.0f < L.u, py = 0.0f < L.v, pz = 0.0f < L.w;
    const float nx = (float)(px ? geo->nodes[1] : geo->nodes[0]);
    const float ny = (float)(py ? geo->nodes[3] : geo->nodes[2]);
    const float nz = (float)(pz ? geo->nodes[5] : geo->nodes[4]);
    const float dx = (nx - L.x) / L.u;
    if (dx < d_b) { d_b = dx; L.sc = px ? 1u : 0u; }
    const float dy = (ny - L.y) / L.v;
    if (dy < d_b) { d_b = dy; L.sc = py ? 3u : 2u; }
    const float dz = (nz - L.z) / L.w;
    if (dz < d_b) { d_b = dz; L.sc = pz ? 5u : 4u; }
  }
  const float d_c = L.life;
  const float m_bc = (d_c < d_b) ? d_c : d_b;
  const float d = (m_bc < d_s) ? m_bc : d_s;

  // 1 - exp(-x) as -expm1(-x): no cancellation for the small optical depths
  // of a single segment; the weight and tallies stay FP64.
  const float frac = -expm1f(-(float)saf * d);
  const double absorbed = L.E * (double)frac;
  L.t_abs += absorbed;
  if constexpr (kTrack) L.t_track += absorbed / saf;
  L.E = L.E - absorbed;

  L.x += L.u * d;
  L.y += L.v * d;
  L.z += L.w * d;
  L.life -= d;

  if (L.E / L.E0 < kCutoff) {
    flush(fo, L);
    fo.kill = L.E;
    L.desc = kKilled;
    return true;
  }
  if (d == d_s) {
    L.pending = true;
    L.desc = kScatter;
    return false;
  }
  if (d == d_b) return boundary_event(cells, geo->bc_bits, L, fo);
  if (d == d_c) {
    L.desc = kCensus;
    flush(fo, L);
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
// (src/history_based_transport.h:263-275) over SoA state.  The segment loop
// runs while any lane of the wave is active, so tally pushes stay
// wave-uniform.
template <class Tally, typename R>
BH_TRACK_KERNEL void k_transport_history(DevCells cells, PhotonSoA W, uint64_t n,
                                         uint64_t ctr_hi, Tally tally, DevCounters* ctrs) {
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool active = i < n;
  unsigned long long flushes = 0, segs = 0, stuck = 0;
  Live<R> L;
  if (active) load_live(W, i, ctr_hi, L);
  while (__ballot(active) != 0) {
    Flush fo;
    fo.has = false;
    fo.kill = 0.0;
    if (active && segment<Tally::kTrack>(cells, L, fo)) {
      segs = L.segs;
      stuck = (L.desc == kStuck);
      store_live(W, i, L);
      active = false;
    }
    flushes += fo.has;
    tally.push(fo);
  }
  tally.finish(cells.n_cells);
  add_counters(ctrs, segs, flushes, stuck);
}

// Persistent form: each wave keeps 64 live photons and refills idle lanes
// from a chunk of photon indices it owns; a chunk of kChunk indices is
// claimed with one atomic on the global queue head, so the queue sees one
// atomic per kChunk histories instead of one per refill.  The loop ends
// when the queue is drained and no lane holds a photon, so every wave exits.
template <class Tally, typename R>
BH_TRACK_KERNEL void k_transport_persistent(DevCells cells, PhotonSoA W, uint64_t n,
                                            uint64_t ctr_hi, Tally tally, DevCounters* ctrs,
                                            PersistentQueue* q, int refill) {
  constexpr uint64_t kChunk = 256;
  bind_lds(tally, cells.n_cells);
  tally.init(cells.n_cells);
  const int lane = threadIdx.x & (kWave - 1);
  Live<R> L;
  bool active = false;
  uint64_t idx = 0;
  uint64_t chunk_next = 0, chunk_end = 0;  // wave-uniform
  bool drained = false;
  unsigned long long flushes = 0, segs = 0, stuck = 0;
  while (true) {
    uint64_t idle = __ballot(!active);
    int n_idle = __popcll(idle);
    if (!drained && (n_idle >= refill || n_idle == kWave)) {
      // hand out indices from the wave's chunk, claiming chunks as needed
      while (n_idle > 0 && !drained) {
        if (chunk_next >= chunk_end) {
          unsigned long long base = 0;
          if (lane == 0) base = atomicAdd(&q->next, (unsigned long long)kChunk);
          base = __shfl(base, 0, kWave);
          if (base >= n) {
            drained = true;
            break;
          }
          chunk_next = base;
          chunk_end = (base + kChunk < n) ? base + kChunk : n;
        }
        const uint64_t avail = chunk_end - chunk_next;
        const uint64_t rank = lane_rank(idle);
        const bool take = !active && rank < avail;
        if (take) {
          idx = chunk_next + rank;
          load_live(W, idx, ctr_hi, L);
          active = true;
        }
        const uint64_t taken = (avail < (uint64_t)n_idle) ? avail : (uint64_t)n_idle;
        chunk_next += taken;
        idle = __ballot(!active);
        n_idle = __popcll(idle);
      }
    }
    if (__ballot(active) == 0) break;
    Flush fo;
    fo.has = false;
    fo.kill = 0.0;
    if (active && segment<Tally::kTrack>(cells, L, fo)) {
      segs += L.segs;
      stuck += (L.desc == kStuck);
      store_live(W, idx, L);
      active = false;
    }
    flushes += fo.has;
    tally.push(fo);
  }
  tally.finish(cells.n_cells);
  add_counters(ctrs, segs, flushes, stuck);
}

}  // namespace bh
