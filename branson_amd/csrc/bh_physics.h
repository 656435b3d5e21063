// bh_physics.h -- a time step's physics on the device, so that temperatures,
// opacities, source energies and tallies never leave HBM between steps:
//   Mesh::calculate_photon_energy (src/mesh.h:237-323) -> k_phys_cells +
//     k_phys_partials + k_phys_etot + k_count_check (k_phys_totals serially),
//   Mesh::update_temperature (src/mesh.h:327-423) -> k_update_temperature +
//     k_report_partials.
// Every formula keeps the reference's FP64 operation order (the build uses
// -ffp-contract=off).  Sums run in the reference's cell order: the photon
// counts int(n_user * E_c / E_tot) depend on the last bit of E_tot, so E_tot
// must give the counts of the reference's serial sum: it is summed in
// parallel with an error bound, and serially (exactly as the reference's
// loop) only when a count could depend on the difference.
#pragma once

#include "bh_kernels.h"

namespace bh {

// Per-cell material (Region, src/region.h:27-76) and state, device arrays.
struct MatDev {
  const double *opA, *opB, *opC, *opS, *cv, *rho;
  double* T_e;        // electron temperature, updated every step
  const double* T_r;  // radiation temperature of the initial census
  const double* T_s;  // temperature of the cell's SOURCE face
  const int8_t* face;
};

// Step totals the host reads back (one per engine, in device memory).  The
// first seven are the sums of calculate_photon_energy's loop, in the order
// k_phys_partials produces them.
struct PhysTotals {
  double tot_emission, tot_census, tot_source;  // after the small-energy pass
  double total_photon_E;                        // this rank's photon energy
  double pre_mat;                               // sum T cV V rho before the step
  double e_sum, e_sq;                           // sum e, sum e^2 of e = em + src + cen
  double absorbed, post_mat;                    // after update_temperature
  double pre_census;                            // step 1: initial census energy
  double global_E;                              // E_tot of the photon counts
  double reassign_E;                            // several ranks: the all-reduced
                                                // step total the reassignment tests
  double E_lo, E_hi;                            // bounds on the serial-order E_tot
  uint32_t ambiguous;                           // a count could differ within them
  uint32_t pad;
};

// ---- deterministic parallel sums ------------------------------------------
// Block b sums a fixed range of cells, each work-item a fixed stride of it,
// then a fixed butterfly within the wave and the four waves in order; the
// block partials are summed the same way by one block.  The result does not
// depend on scheduling, so a run is reproducible, but it is not the
// reference's serial rounding (see k_count_check).
constexpr int kSumNT = 256;
constexpr int kSumMaxBlocks = 1024;

template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K]) {
  __shared__ double lds[K][kSumNT / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v[k] += __shfl_xor(v[k], d, kWave);
    if (lane == 0) lds[k][wave] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((lds[k][0] + lds[k][1]) + lds[k][2]) + lds[k][3];
}

__device__ __forceinline__ void sum_range(uint32_t n, uint32_t& lo, uint32_t& hi) {
  const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
  lo = blockIdx.x * per;
  hi = lo + per < n ? lo + per : n;
}

// The seven sums of calculate_photon_energy's loop, one partial per block.
__global__ __launch_bounds__(kSumNT) void k_phys_partials(const double* __restrict__ em,
                                                          const double* __restrict__ src,
                                                          const double* __restrict__ cen,
                                                          const double* __restrict__ mat,
                                                          uint32_t n, double* __restrict__ part) {
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  uint32_t lo, hi;
  sum_range(n, lo, hi);
  for (uint32_t c = lo + threadIdx.x; c < hi; c += kSumNT) {
    const double e = em[c], s = src[c], ce = cen[c];
    v[0] += e;
    v[1] += ce;
    v[2] += s;
    v[3] += s + ce + e;
    v[4] += mat[c];
    const double all = e + s + ce;
    v[5] += all;
    v[6] += all * all;
  }
  block_sums<7>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < 7; ++k) part[blockIdx.x * 8 + k] = v[k];
}

// absorbed energy and post-step material energy, one partial per block
__global__ __launch_bounds__(kSumNT) void k_report_partials(const double* __restrict__ abs_E,
                                                            const double* __restrict__ mat,
                                                            uint32_t n, double* __restrict__ part) {
  double v[2] = {0, 0};
  uint32_t lo, hi;
  sum_range(n, lo, hi);
  for (uint32_t c = lo + threadIdx.x; c < hi; c += kSumNT) {
    v[0] += abs_E[c];
    v[1] += mat[c];
  }
  block_sums<2>(v);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 8] = v[0];
    part[blockIdx.x * 8 + 1] = v[1];
  }
}

// Step 1: energy of the initial census photons, n_c copies of E_c / n_c per
// cell (each cell's copies as one product), one partial per block.
__global__ __launch_bounds__(kSumNT) void k_census_partials(const double* __restrict__ cen,
                                                            const uint64_t* __restrict__ counts,
                                                            uint32_t n, double* __restrict__ part) {
  double v[1] = {0};
  uint32_t lo, hi;
  sum_range(n, lo, hi);
  for (uint32_t c = lo + threadIdx.x; c < hi; c += kSumNT) {
    const uint64_t k = counts[c];
    if (k > 0) v[0] += (double)k * (cen[c] / (double)k);
  }
  block_sums<1>(v);
  if (threadIdx.x == 0) part[blockIdx.x * 8] = v[0];
}

// Deterministic block partials of E[0, *n) (a count read on the device),
// one per block of a fixed grid of kSumMaxBlocks: the split depends on n
// only.  k_final_sums<1> ends the sum.
__global__ __launch_bounds__(kSumNT) void k_rank_energy_partials(
    const double* __restrict__ E, const unsigned long long* __restrict__ n_ptr,
    double* __restrict__ part) {
  double v[1] = {0.0};
  uint32_t lo, hi;
  sum_range((uint32_t)*n_ptr, lo, hi);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kSumNT) v[0] += E[i];
  block_sums<1>(v);
  if (threadIdx.x == 0) part[blockIdx.x * 8] = v[0];
}

// One block: dst[k] = sum over the nb block partials of quantity k.
template <int K>
__global__ __launch_bounds__(kSumNT) void k_final_sums(const double* __restrict__ part,
                                                       uint32_t nb, double* __restrict__ dst) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (uint32_t b = threadIdx.x; b < nb; b += kSumNT)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += part[b * 8 + k];
  block_sums<K>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) dst[k] = v[k];
}

// One rank: E_tot is this rank's total, and the serial loop's value lies
// within E_par +- delta, delta = 2 (n + 64) u E_par (u = 2^-53; the summands
// are the same non-negative doubles on both sides: the serial sum is within
// (n - 1) u sum x of the exact sum, the parallel one within (log2 n + 4) u).
__global__ void k_phys_etot(uint32_t n, PhysTotals* __restrict__ out,
                            SourceInst* __restrict__ inst2) {
  const double E = out->total_photon_E;
  const double delta = 2.0 * ((double)n + 64.0) * 0x1p-53 * E;
  out->global_E = E;
  out->E_lo = E - delta;
  out->E_hi = E + delta;
  out->ambiguous = 0u;
  inst2[0].total_E = E;
  inst2[1].total_E = E;
}

// Per cell: does int(n_user * E_c / E_tot) (the photon count, monotone in
// E_tot) take the same value at both ends of [E_lo, E_hi]?  If it does for
// every energy of every cell, the counts are those of the reference's serial
// E_tot, whatever its last bits; if not, the engine computes E_tot serially.
__global__ __launch_bounds__(256) void k_count_check(const double* __restrict__ em,
                                                     const double* __restrict__ src,
                                                     const double* __restrict__ cen,
                                                     uint32_t n, uint64_t n_user, int first_step,
                                                     PhysTotals* __restrict__ out) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double lo = out->E_lo, hi = out->E_hi;
  const double e3[3] = {em[c], src[c], first_step ? cen[c] : 0.0};
  bool amb = false;
  for (int k = 0; k < 3; ++k)
    if (e3[k] > 0.0) amb |= photons_for(n_user, e3[k], lo) != photons_for(n_user, e3[k], hi);
  if (amb) atomicOr(&out->ambiguous, 1u);
}

__device__ __forceinline__ double cell_volume(const CellGeo& g) {
  return (g.nodes[1] - g.nodes[0]) * (g.nodes[3] - g.nodes[2]) * (g.nodes[5] - g.nodes[4]);
}

// Cell::get_face_area (src/cell.h:83-100); -1 for "no face".
__device__ __forceinline__ double face_area(const CellGeo& g, int face) {
  if (face == 0 || face == 1) return (g.nodes[3] - g.nodes[2]) * (g.nodes[5] - g.nodes[4]);
  if (face == 2 || face == 3) return (g.nodes[1] - g.nodes[0]) * (g.nodes[5] - g.nodes[4]);
  if (face == 4 || face == 5) return (g.nodes[1] - g.nodes[0]) * (g.nodes[3] - g.nodes[2]);
  return -1.0;
}

// Per cell: opacities and Fleck factor into the tracking records (every
// group, Cell::set_op_a / set_op_s, src/cell.h:260-275), and the emission,
// initial-census and face-source energies (src/mesh.h:257-281).
// With P (a step captured as a graph), dt is the step's, read on the device.
__global__ __launch_bounds__(256) void k_phys_cells(CellGeo* __restrict__ geo,
                                                    double2* __restrict__ sigma,
                                                    CellRec* __restrict__ rec, uint32_t G,
                                                    uint32_t n, MatDev m, double dt, double rf,
                                                    int first_step, double* __restrict__ em,
                                                    double* __restrict__ src,
                                                    double* __restrict__ cen,
                                                    double* __restrict__ mat,
                                                    const StepParams* __restrict__ P) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (P) dt = P->dt;
  const CellGeo g = geo[i];
  const double vol = cell_volume(g);
  const double cV = m.cv[i], T = m.T_e[i], Tr = m.T_r[i], Ts = m.T_s[i], rho = m.rho[i];
  const double op_a = m.opA[i] + m.opB[i] * pow(T, m.opC[i]);  // src/region.h:44-46
  const double op_s = m.opS[i];
  const double f = 1.0 / (1.0 + dt * op_a * kC * (4.0 * kA * pow(T, 3.0) / (cV * rho)));
  for (uint32_t k = 0; k < G; ++k) sigma[(size_t)i * G + k] = make_double2(op_a, op_s);
  geo[i].fleck = f;
  if (rec) {
    rec[i].sigma_a = op_a;
    rec[i].sigma_s = op_s;
    rec[i].fleck = f;
  }
  em[i] = rf * dt * vol * f * op_a * kA * kC * pow(T, 4.0);
  cen[i] = first_step ? rf * vol * kA * pow(Tr, 4.0) : 0.0;
  src[i] = rf * 0.25 * kA * kC * face_area(g, m.face[i]) * pow(Ts, 4.0) * dt;
  mat[i] = T * cV * vol * rho;
}

// Serial cell-order sums, the reference loop's exact rounding (the fallback
// when k_count_check finds a count that E_tot's last bits could change, and
// the path for more than one rank).  Every lane of one wave walks the cells
// in order with the same (uniform) index: scalar loads, one dependent chain
// of FP64 adds per sum.
//
// Pass 1 (pass = 0): the sums of calculate_photon_energy's loop; E_tot for
// the photon counts with one rank (its own total, the all-reduce of one
// value); with several, the all-reduced step total of the small-energy
// reassignment (n_ranks equal values summed in rank order), after which
// k_phys_rank_sums sets E_tot and k_phys_reassign applies this rank's
// share.  Pass 2 (pass = 1) recomputes this rank's step totals only.
// only_if_ambiguous: run only when k_count_check flagged a count (the
// serial fallback inside a step that no host round trip splits).
__global__ __launch_bounds__(kWave) void k_phys_totals(const double* __restrict__ em,
                                                       const double* __restrict__ src,
                                                       const double* __restrict__ cen,
                                                       const double* __restrict__ mat,
                                                       uint32_t n, uint32_t n_ranks, int pass,
                                                       PhysTotals* __restrict__ out,
                                                       SourceInst* __restrict__ inst2,
                                                       int only_if_ambiguous) {
  if (only_if_ambiguous && out->ambiguous == 0u) return;
  double t_em = 0.0, t_cen = 0.0, t_src = 0.0, total = 0.0, pre_mat = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll 8
  for (uint32_t c = 0; c < n; ++c) {
    const double e = em[c], s = src[c], ce = cen[c];
    t_em += e;
    t_cen += ce;
    t_src += s;
    total += s + ce + e;
    if (pass == 0) {
      pre_mat += mat[c];
      const double all = e + s + ce;
      e1 += all;
      e2 += all * all;
    }
  }
  if (threadIdx.x != 0) return;
  out->tot_emission = t_em;
  out->tot_census = t_cen;
  out->tot_source = t_src;
  out->total_photon_E = total;
  if (pass == 0) {
    const double local3 = t_em + t_cen + t_src;
    double g = local3;
    for (uint32_t r = 1; r < n_ranks; ++r) g += local3;
    out->reassign_E = g;
    out->pre_mat = pre_mat;
    out->e_sum = e1;
    out->e_sq = e2;
    if (n_ranks == 1) {
      out->global_E = total;
      inst2[0].total_E = total;
      inst2[1].total_E = total;
    }
  }
}

// Small-energy reassignment (src/mesh.h:291-307) of one cell's energies as
// rank r of n_ranks sees them: an energy the all-rank source would give no
// photon of its own (int(n_user * (E / reassign_E)) == 0) goes whole to
// rank cell % n_ranks, and is zero on the others.
__device__ __forceinline__ void reassign_cell(uint32_t c, uint32_t r, uint32_t n_ranks,
                                              double rf, double nu, double g, int first_step,
                                              double& e, double& s, double& ce) {
  const bool mine = (c % n_ranks) == r;
  if (first_step && ce > 0.0 && (int)(nu * (ce / g)) == 0) ce = mine ? ce / rf : 0.0;
  if (e > 0.0 && (int)(nu * (e / g)) == 0) e = mine ? e / rf : 0.0;
  if (s > 0.0 && (int)(nu * (s / g)) == 0) s = mine ? s / rf : 0.0;
}

// More than one rank, before k_phys_reassign: E_tot of the photon counts is
// the all-reduce of every rank's total_photon_E after its own reassignment
// (src/replicated_driver.h:57-59), summed in rank order.  Lane r of the one
// wave computes rank r's total in the reference's cell order from the
// shared pre-reassignment energies.
constexpr uint32_t kMaxRanks = 1024;
__global__ __launch_bounds__(kWave) void k_phys_rank_sums(const double* __restrict__ em,
                                                          const double* __restrict__ src,
                                                          const double* __restrict__ cen,
                                                          uint32_t n, uint64_t n_user,
                                                          uint32_t n_ranks, double rf,
                                                          int first_step,
                                                          PhysTotals* __restrict__ out,
                                                          SourceInst* __restrict__ inst2) {
  __shared__ double rank_total[kMaxRanks];
  const double g = out->reassign_E;
  const double nu = (double)(uint32_t)n_user;
  for (uint32_t r = threadIdx.x; r < n_ranks && r < kMaxRanks; r += kWave) {
    double t = 0.0;
    for (uint32_t c = 0; c < n; ++c) {
      double e = em[c], s = src[c], ce = cen[c];
      reassign_cell(c, r, n_ranks, rf, nu, g, first_step, e, s, ce);
      t += s + ce + e;
    }
    rank_total[r] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double all = 0.0;
  for (uint32_t r = 0; r < n_ranks && r < kMaxRanks; ++r) all += rank_total[r];
  out->global_E = all;
  inst2[0].total_E = all;
  inst2[1].total_E = all;
}

// This rank's share of the reassignment, in place.
__global__ __launch_bounds__(256) void k_phys_reassign(double* __restrict__ em,
                                                       double* __restrict__ src,
                                                       double* __restrict__ cen, uint32_t n,
                                                       uint64_t n_user, uint32_t n_ranks,
                                                       uint32_t rank, double rf, int first_step,
                                                       const PhysTotals* __restrict__ tot) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double e = em[c], s = src[c], ce = cen[c];
  reassign_cell(c, rank, n_ranks, rf, (double)(uint32_t)n_user, tot->reassign_E, first_step, e,
                s, ce);
  cen[c] = ce;
  em[c] = e;
  src[c] = s;
}

// A step captured as a graph: the step's photon counts from the source
// scan's total, within the stores' capacity.
__global__ void k_step_counts(StepParams* __restrict__ P, const uint64_t* __restrict__ n_new) {
  const uint64_t room = P->cap > P->n_cen ? P->cap - P->n_cen : 0;
  const uint64_t k = *n_new;
  P->n_new = k < room ? k : room;
  P->n = P->n_cen + P->n_new;
  P->overflow = k > room ? 1u : 0u;
}

// ---- several independent problems (instances) ------------------------------
// Per-instance step totals (the batched workload of many problems, which has
// no counterpart in the reference): one wave per instance sums its cells in
// a fixed order (lane-strided partial sums, then a fixed butterfly), so a
// run is reproducible.  An instance's photon energy total_E is the sum of
// its cells' source + census + emission energies; its photon counts
// int(n_user E_c / total_E) follow from it.
struct InstReport {  // bh_instance_report, field for field
  double total_E, emission_E, source_E, pre_census_E, post_census_E;
  double absorbed_E, exit_E, pre_mat_E, post_mat_E;
  uint64_t n_new;
};


// After the physics of the step: energies, material energy and total_E per
// instance; total_E goes into the instance's source records (census, new).
__global__ __launch_bounds__(kWave) void k_inst_sums(const double* __restrict__ em,
                                                     const double* __restrict__ src,
                                                     const double* __restrict__ cen,
                                                     const double* __restrict__ mat,
                                                     const uint32_t* __restrict__ first_cell,
                                                     uint32_t K, InstReport* __restrict__ rep,
                                                     SourceInst* __restrict__ rec) {
  const uint32_t k = blockIdx.x;
  if (k >= K) return;
  double e = 0.0, s = 0.0, m = 0.0, t = 0.0;
  for (uint32_t i = first_cell[k] + threadIdx.x; i < first_cell[k + 1]; i += kWave) {
    const double ei = em[i], si = src[i], ci = cen[i];
    e += ei;
    s += si;
    m += mat[i];
    t += si + ci + ei;
  }
  e = wave_sum_f64(e);
  s = wave_sum_f64(s);
  m = wave_sum_f64(m);
  t = wave_sum_f64(t);
  if (threadIdx.x != 0) return;
  InstReport& r = rep[k];
  r.total_E = t;
  r.emission_E = e;
  r.source_E = s;
  r.pre_census_E = 0.0;  // the census photons' energy, summed once they exist
  r.pre_mat_E = m;
  r.post_census_E = 0.0;
  r.exit_E = 0.0;
  rec[k].total_E = t;
  rec[K + k].total_E = t;
}

// Census energy per instance: the photons of W[0, *n) (or [0, n)) summed
// into out[instance] (post = 1: post_census_E, else pre_census_E).  Each
// wave walks a contiguous range of the census, which is in census order, so
// its photons stay in one instance for long runs: the wave keeps a per-lane
// partial for the current instance and adds it with one atomic when the
// instance changes; an iteration whose photons span instances adds lane by
// lane.
__global__ __launch_bounds__(256) void k_inst_census(const uint32_t* __restrict__ cell,
                                                     const double* __restrict__ E, uint64_t n,
                                                     const unsigned long long* __restrict__ n_ptr,
                                                     const uint64_t* __restrict__ n_ptr64,
                                                     const uint32_t* __restrict__ cell_inst,
                                                     InstReport* __restrict__ rep, int post) {
  if (n_ptr) n = *n_ptr < n ? *n_ptr : n;
  if (n_ptr64) n = *n_ptr64 < n ? *n_ptr64 : n;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  // this wave's range, in whole iterations of 64 photons
  const uint64_t per = (n + waves * kWave - 1) / (waves * kWave) * kWave;
  const uint64_t lo = wave * per;
  const uint64_t hi = lo + per < n ? lo + per : n;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  double* dst = post ? &rep[0].post_census_E : &rep[0].pre_census_E;
  const size_t stride_d = sizeof(InstReport) / sizeof(double);
  uint32_t cur = 0xFFFFFFFFu;  // wave-uniform: the instance `part` belongs to
  double part = 0.0;
  for (uint64_t b = lo; b < hi; b += kWave) {
    const uint64_t i = b + (uint64_t)lane;
    const bool in = i < hi;
    const uint32_t k = in ? cell_inst[cell[i]] : 0u;
    const double e = in ? E[i] : 0.0;
    const uint32_t k0 = (uint32_t)__shfl((int)k, 0, kWave);
    const bool same = __ballot(in && k != k0) == 0;
    if (same && k0 == cur) {
      part += e;
    } else if (same) {
      if (cur != 0xFFFFFFFFu) {
        const double w = wave_sum_f64(part);
        if (lane == 0 && w != 0.0) atomicAdd(dst + cur * stride_d, w);
      }
      cur = k0;
      part = e;
    } else if (in) {
      atomicAdd(dst + k * stride_d, e);
    }
  }
  if (cur != 0xFFFFFFFFu) {
    const double w = wave_sum_f64(part);
    if (lane == 0 && w != 0.0) atomicAdd(dst + cur * stride_d, w);
  }
}

// After update_temperature: absorbed and material energy, new photons and
// exit energy per instance.
__global__ __launch_bounds__(kWave) void k_inst_report(const double* __restrict__ abs_E,
                                                       const double* __restrict__ mat,
                                                       const uint64_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ first_cell,
                                                       const double* __restrict__ inst_exit,
                                                       uint32_t K, InstReport* __restrict__ rep) {
  const uint32_t k = blockIdx.x;
  if (k >= K) return;
  double a = 0.0, m = 0.0;
  unsigned long long c = 0;
  for (uint32_t i = first_cell[k] + threadIdx.x; i < first_cell[k + 1]; i += kWave) {
    a += abs_E[i];
    m += mat[i];
    c += counts[i];
  }
  a = wave_sum_f64(a);
  m = wave_sum_f64(m);
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, kWave);
  if (threadIdx.x != 0) return;
  rep[k].absorbed_E = a;
  rep[k].post_mat_E = m;
  rep[k].n_new = c;
  rep[k].exit_E = inst_exit[k];
}

// Mesh::update_temperature, one rank (src/mesh.h:327-423): new electron
// temperature from the step's absorbed and emitted energy, the diagnostic
// radiation temperature from the track tally, and T cV V rho per cell.
__global__ __launch_bounds__(256) void k_update_temperature(
    const CellGeo* __restrict__ geo, MatDev m, const double* __restrict__ abs_E,
    const double* __restrict__ track_E, const double* __restrict__ em, double dt, uint32_t n,
    double* __restrict__ T_r_out, double* __restrict__ mat, const StepParams* __restrict__ P) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (P) dt = P->dt;
  const double cV = m.cv[i], rho = m.rho[i], vol = cell_volume(geo[i]), T = m.T_e[i];
  const double T_new = T + (abs_E[i] - em[i]) / (cV * vol * rho);
  T_r_out[i] = pow(track_E[i] / (vol * dt * kA * kC), 0.25);
  m.T_e[i] = T_new;
  mat[i] = T_new * cV * vol * rho;
}

}  // namespace bh
