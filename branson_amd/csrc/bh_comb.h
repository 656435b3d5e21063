// bh_comb.h -- census population control on the device: comb_photons
// (src/census_functions.h:48-93).  Present upstream but never called there;
// here it is an entry point on the resident census (bh_comb_census).
//
// The reference walks the census in order with one RNG, per photon:
//   p_kill = 1 - E / comb_E, r = next draw;
//   keep if r > p_kill or the cell has one photon left (count == 1), else
//   count--;
// and then adds (cell_E - kept * comb_E) / kept to every survivor of a cell.
// In parallel:
//   * draw i is RNG(seed, stream) at counter ctr0 + i (counter-based);
//   * "one photon left" happens exactly for the last photon of a cell whose
//     earlier photons were all combed away;
//   * the per-cell sums (census energy, corrected energy) run in census order
//     over the cell's photons, after a stable radix sort of the census by cell;
//   * comb_E = E_census / max comes from a deterministic parallel sum whose
//     bound covers the reference's serial sum, and every decision is checked
//     at both ends of the bound (serial sum otherwise).
#pragma once

#include "bh_kernels.h"
#include "bh_physics.h"

namespace bh {

struct CombParams {
  double comb_E;          // E_census / max_census_photons
  double comb_lo, comb_hi;  // comb_E at the ends of E_census's bound
  uint64_t ctr0, stream, ctr_hi;
};

// Deterministic block partials of the census energy (k_final_sums<1> ends
// the sum).
__global__ __launch_bounds__(kSumNT) void k_energy_partials(const double* __restrict__ E,
                                                            uint32_t n,
                                                            double* __restrict__ part) {
  double v[1] = {0.0};
  uint32_t lo, hi;
  sum_range(n, lo, hi);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kSumNT) v[0] += E[i];
  block_sums<1>(v);
  if (threadIdx.x == 0) part[blockIdx.x * 8] = v[0];
}

// The reference's get_photon_list_E (src/census_functions.h:40-46): one
// serial chain over the census in order (uniform loop, scalar loads).
__global__ __launch_bounds__(kWave) void k_serial_energy(const double* __restrict__ E,
                                                         uint32_t n, double* __restrict__ out) {
  double s = 0.0;
#pragma unroll 8
  for (uint32_t i = 0; i < n; ++i) s += E[i];
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_iota_u32(uint32_t* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

__global__ void k_cell_hist(const uint32_t* __restrict__ cell, uint32_t n,
                            unsigned long long* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&counts[cell[i]], 1ull);
}

// One work-item per cell: its photons (sorted[off[c] .. off[c] + cnt[c]),
// census order) combed as the reference's loop combs them.  keep[i] and the
// cell's corrected survivor energy are written; ambiguous is set when a
// decision differs at the two ends of comb_E's bound.
__global__ __launch_bounds__(256) void k_comb_cells(const uint32_t* __restrict__ sorted,
                                                    const unsigned long long* __restrict__ off,
                                                    const unsigned long long* __restrict__ cnt,
                                                    uint32_t n_cells,
                                                    const double* __restrict__ E, CombParams cp,
                                                    uint8_t* __restrict__ keep,
                                                    double* __restrict__ cell_E_new,
                                                    uint32_t* __restrict__ ambiguous,
                                                    unsigned long long* __restrict__ n_kept) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const uint64_t m = cnt[c], base = off[c];
  if (m == 0) return;
  double cell_E = 0.0;
  for (uint64_t j = 0; j < m; ++j) cell_E += E[sorted[base + j]];
  uint32_t kept = 0;
  double corrected = 0.0;
  bool amb = false;
  for (uint64_t j = 0; j < m; ++j) {
    const uint32_t i = sorted[base + j];
    const double e = E[i];
    const double p_kill = 1.0 - e / cp.comb_E;
    uint64_t o0, o1;
    Threefry::block(cp.ctr0 + i, cp.ctr_hi, cp.stream, 0ull, o0, o1);
    const double r = u01(o0);
    amb |= (r > 1.0 - e / cp.comb_lo) != (r > 1.0 - e / cp.comb_hi);
    // count == 1: the last photon, every earlier one of the cell combed away
    const bool k = (r > p_kill) || (j + 1 == m && kept == 0);
    keep[i] = k ? 1 : 0;
    if (k) {
      corrected += cp.comb_E;
      ++kept;
    }
  }
  cell_E_new[c] = cp.comb_E + (cell_E - corrected) / (double)kept;
  atomicAdd(n_kept, (unsigned long long)kept);
  if (amb) atomicOr(ambiguous, 1u);
}

// Survivors take their cell's corrected energy and stay CENSUS; the rest
// become KILLED, so the census compaction drops them.
__global__ void k_comb_commit(PhotonSoA W, uint32_t n, const uint8_t* __restrict__ keep,
                              const double* __restrict__ cell_E_new) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (keep[i]) {
    W.E[i] = cell_E_new[W.cell[i]];
    W.desc[i] = kCensus;
  } else {
    W.desc[i] = kKilled;
  }
}

}  // namespace bh
