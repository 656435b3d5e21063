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
// serial chain over the census in census order (perm: rank -> slot;
// uniform loop, scalar loads).
__global__ __launch_bounds__(kWave) void k_serial_energy(const double* __restrict__ E,
                                                         const uint32_t* __restrict__ perm,
                                                         uint32_t n, double* __restrict__ out) {
  double s = 0.0;
#pragma unroll 8
  for (uint32_t j = 0; j < n; ++j) s += E[perm[j]];
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_iota_u32(uint32_t* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

// Sort keys in census order: the cell of the photon of rank j.
__global__ void k_cells_by_rank(const uint32_t* __restrict__ cell,
                                const uint32_t* __restrict__ perm, uint32_t n,
                                uint32_t* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = cell[perm[j]];
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
// sorted: census ranks grouped by cell, each cell's in census order; perm:
// rank -> slot of W; keep is indexed by rank.
__global__ __launch_bounds__(256) void k_comb_cells(const uint32_t* __restrict__ sorted,
                                                    const uint32_t* __restrict__ perm,
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
  for (uint64_t j = 0; j < m; ++j) cell_E += E[perm[sorted[base + j]]];
  uint32_t kept = 0;
  double corrected = 0.0;
  bool amb = false;
  for (uint64_t j = 0; j < m; ++j) {
    const uint32_t i = sorted[base + j];  // census rank
    const double e = E[perm[i]];
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

// keep[j] as u64 flags for the survivors' exclusive scan.
__global__ void k_keep_flags(const uint8_t* __restrict__ keep, uint32_t n,
                             uint64_t* __restrict__ flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) flag[j] = keep[j];
}

// Survivors, in census order, to C[pos[j]] with their cell's corrected
// energy and their new census rank as key (life and the rest unchanged).
__global__ void k_comb_compact(PhotonSoA W, const uint32_t* __restrict__ perm, uint32_t n,
                               const uint8_t* __restrict__ keep,
                               const uint64_t* __restrict__ pos,
                               const double* __restrict__ cell_E_new, PhotonSoA C) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !keep[j]) return;
  const uint32_t s = perm[j];
  const uint64_t o = pos[j];
  copy_photon(W, s, C, o, cell_E_new[W.cell[s]], (uint32_t)o);
}

}  // namespace bh
