// bh_kernels.h -- kernels of the MI355X IMC engine (source, tracking, census,
// scans) and the device-side data layout they share.  Included once by
// bh_engine.hip.
#pragma once

#include "bh_device.h"

namespace bh {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Photon state in HBM: structure of arrays, one 8/4/1-byte stream per field,
// so every per-photon load and store of a wavefront is one coalesced
// transaction.  98 bytes per photon.
// ---------------------------------------------------------------------------
struct PhotonSoA {
  double *x, *y, *z;  // position (cm)
  double *u, *v, *w;  // direction cosines
  double *E, *E0, *life;
  uint64_t *ctr, *stream;
  uint32_t *cell, *group;
  uint8_t *desc, *stype;
  // resident census: the photon's rank in the reference's census order (the
  // census is stored in the order the transport kernels finish histories)
  uint32_t* key;
};

// Census output of a transport launch: each history that ends as census is
// written by the kernel itself to a slot of C and marked at its transport
// index t (mark[t] = 1) -- the reference's census order is the order of t,
// so a census photon's rank is the count of marks below its t (marks packed
// 64 to a word, a scan of the word counts, a popcount within the word) and
// the census energy is summed by rank -- and exits,
// kills and other endings are counted on the fly.  With `on` false the
// photons stay in place in W (bh_transport_photons).  W = [census (keys =
// census ranks) | new photons], transport index t = n_new + key for a census
// photon, phys - n_cen for a new one (src/replicated_driver.h:73-75: new
// photons first).
//
// Slots: a wave takes kCenChunk slots of C at a time with one atomic on
// n_reserved and fills them in order, so every chunk but a wave's last is
// full.  The unused tail of each last chunk is a run of holes (key kHole)
// recorded in rec; after the launch the census photons beyond n_census are
// moved into the holes below it, leaving C[0, n_census) dense.
constexpr uint32_t kCenChunk = 64;
constexpr uint32_t kHole = 0xFFFFFFFFu;

struct HoleRec {
  uint64_t start, len;
};

struct CensusAcc {  // zeroed before each launch
  unsigned long long n_reserved;  // slots of C handed out
  unsigned long long n_census;    // census photons written
  unsigned long long n_rec;       // hole records
  unsigned long long n_holes;     // n_reserved - n_census, set after the launch
  unsigned long long n_exit, n_killed, n_other;
  unsigned long long overflow;    // a capacity bound was hit (never, by sizing)
  double exit_E;
  unsigned long long n_words;     // census marks' words, set after the launch
  double pad[6];
};

struct CensusOut {
  PhotonSoA C;
  uint8_t* mark;    // [n] 1 at each census photon's transport index, else 0
  CensusAcc* acc;
  HoleRec* rec;     // [rec_cap]: one per wave that took a chunk
  double life;      // c * next_dt given to census photons
  uint64_t n_cen, n_new;
  uint64_t slots, rec_cap;  // capacity of C, of rec
  int on;
  // several independent problems (else null): exit energy per instance,
  // added by every history that leaves through a vacuum or source face
  double* inst_exit;
};

struct DevCells {
  const CellGeo* geo;     // [n_cells]
  const double2* sigma;   // [n_cells * G] (sigma_a, sigma_s)
  const uint32_t* next;   // [n_cells * 6]
  uint32_t n_cells;
  uint32_t G;
  float fp32_tol;         // FP32 tracking: largest relative error bound of a deposit
  // several independent problems in one mesh (else null): the instance of
  // each cell and each instance's RNG counter-high word (seed << 32)
  const uint32_t* cell_inst;
  const uint64_t* inst_ctr_hi;
  // lattice meshes (else rec is null): per-cell records, axis node arrays
  // (nx+1, ny+1, nz+1 entries) and the lattice shape
  const CellRec* rec;
  const double* axis[3];
  uint32_t dims[3];
  // many problems, each a lattice (else null): per-instance shape and axis
  // offsets; CellRec::pad holds the cell's instance, axis[0] the axes of
  // every instance back to back
  const struct LatInst* lat;
};

// One problem of a batch whose problems are lattices: its first cell, its
// shape, and where its x, y and z nodes start in the concatenated axes.
struct LatInst {
  uint32_t first, nx, ny, nz;
  uint32_t ox, oy, oz, pad;
};

// The RNG counter-high word of a photon in cell c.
__device__ __forceinline__ uint64_t photon_ctr_hi(const DevCells& cells, uint32_t c,
                                                  uint64_t ctr_hi) {
  return cells.cell_inst ? cells.inst_ctr_hi[cells.cell_inst[c]] : ctr_hi;
}

// Source parameters of one problem instance (a single problem is one
// instance starting at cell 0).
struct SourceInst {
  uint64_t n_user;       // photons requested for the instance
  double total_E;        // the instance's total photon energy
  uint64_t stream_base;  // 1e13 cycle + n_user rank, or n_user rank (initial census)
  uint64_t ctr_hi;       // seed << 32
  uint32_t first_cell;   // the instance's cells are [first_cell, next instance's)
  uint32_t pad;
};

// Per-step values a step captured as a HIP graph reads on the device (the
// graph copies them in from pinned host memory when it starts).
struct StepParams {
  uint64_t n_cen;       // resident census photons, stored first in W
  uint64_t n_new;       // the step's new photons, stored after them
  uint64_t n;           // n_cen + n_new
  double census_life;   // c * next_dt
  double dt;
  SourceInst fresh;     // one problem's source parameters
  // a resident step (physics on the device) captured as a graph: the photon
  // counts are only known on the device, so n_new and n are set there, from
  // the source counts, within the capacity the stores were sized for
  uint64_t cap;         // photons W and C hold (n is clamped to it)
  uint32_t overflow;    // set when the counts exceeded cap (never, by sizing)
  uint32_t pad;
};

// New photons made where they are transported (the persistent kernel and
// the FP64 replay) instead of being written to HBM by a source kernel and
// read straight back: where photon p of the step's new photons comes from.
// Photon slots from first_gen on are new photon p = slot - first_gen.
// Per cell, from the source counts: its emission photons and the energy of
// each of its emission and face-source photons (initial census: its census
// photons and their energy), which a photon then reads instead of dividing.
struct SourceCell {
  double e_em;    // emission_E / n_em (initial census: census_E / n)
  double e_src;   // source_E / n_src
  uint32_t n_em;  // emission (initial census: census) photons
  uint32_t pad;
};

struct SourceDesc {
  const SourceCell* sc;      // per-cell photon counts and energies
  const int8_t* face;        // per-cell SOURCE face
  const uint64_t* off;       // exclusive prefix of the per-cell photon counts
  const uint32_t* hint;      // hint[j]: the cell of photon j * kHintStride
  const SourceInst* inst;    // device records (one problem: inst[0]); null: `one`
  SourceInst one;
  double dt;                 // (a captured step reads StepParams::dt)
  uint64_t n_new;            // hint covers photons [0, n_new)
  int on;                    // 0: every photon is read from the store
  int pad;
};
constexpr uint32_t kHintStride = 256;

struct DevCounters {
  unsigned long long segments;
  unsigned long long flushes;
  unsigned long long stuck;
  unsigned long long replays;  // histories handed from FP32 tracking to FP64 replay
  unsigned long long bad_args;  // launches whose kernel-argument view was wrong (never)
};

// Histories the FP32-tracking kernels hand over for FP64 replay: the photon
// index (its initial state is still in place), the first segment whose
// deposits the replay tallies, and the FP32 weight at that point.
struct ReplayList {
  uint64_t* idx;
  uint32_t* first_seg;
  double* E;
};

struct CensusPartial {  // a census-writing launch's totals, read back
  double census_E;
  double exit_E;
  unsigned long long n_census;
  unsigned long long n_exit;
  unsigned long long n_killed;
  unsigned long long n_other;
  unsigned long long overflow;
};

// ---------------------------------------------------------------------------
// Block scan helpers (blockDim.x == 256: four wave64s).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const T o = __shfl_up(v, d, kWave);
    if (lane >= d) v += o;
  }
  return v;
}

template <typename T, int NT>
__device__ __forceinline__ T block_excl_scan(T v, T* lds /*[NT/64]*/, T& total) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const T incl = wave_incl_scan(v);
  if (lane == kWave - 1) lds[wave] = incl;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const T c = lds[w];
    if (w < wave) off += c;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return off + incl - v;
}

template <typename T, int NT>
__device__ __forceinline__ T block_sum(T v, T* lds) {
  T tot;
  block_excl_scan<T, NT>(v, lds, tot);
  return tot;
}

// Exclusive prefix over a 64-bit flag ballot: position of this lane's flag
// among the set flags of its wave.
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  const uint32_t lo = (uint32_t)mask, hi = (uint32_t)(mask >> 32);
  return __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
}

// Sum over the 64 lanes of a wave, in a fixed butterfly order; every lane
// gets the sum.  All lanes at the same program point.
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// ---------------------------------------------------------------------------
// Generic exclusive scan of uint64 counts: reduce per tile, scan the tile
// sums in one block, then scan each tile with its offset.
// ---------------------------------------------------------------------------
constexpr int kScanNT = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanNT * kScanItems;

// Three-pass exclusive scan of n items.  n is the argument, or *n_ptr when
// n_ptr is set (sizes that only the device knows, e.g. inside a captured
// graph); the tile loops are grid-stride, so any grid works.
// A length read on the device never exceeds the bound the buffers were
// sized for.
__device__ __forceinline__ uint64_t scan_n(uint64_t n, const uint64_t* n_ptr) {
  return n_ptr ? (*n_ptr < n ? *n_ptr : n) : n;
}

// kNz: an item counts 1 when its bits are nonzero (a census energy stored
// by transport index), else its value.
template <bool kNz>
__device__ __forceinline__ uint64_t scan_item(uint64_t v) {
  return kNz ? (uint64_t)(v != 0) : v;
}

template <bool kNz>
__global__ __launch_bounds__(kScanNT) void k_scan_reduce(const uint64_t* __restrict__ in,
                                                         uint64_t n, const uint64_t* n_ptr,
                                                         uint64_t* __restrict__ tile_sum,
                                                         double* __restrict__ tile_E) {
  __shared__ uint64_t lds[kScanNT / kWave];
  __shared__ double lds_e[kScanNT / kWave];
  n = scan_n(n, n_ptr);
  const uint64_t n_tiles = (n + kScanTile - 1) / kScanTile;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const uint64_t base = t * kScanTile;
    uint64_t s = 0;
    double e = 0.0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      const uint64_t i = base + (uint64_t)it * kScanNT + threadIdx.x;
      if (i < n) {
        const uint64_t v = in[i];
        s += scan_item<kNz>(v);
        if (kNz) e += __longlong_as_double((long long)v);
      }
    }
    s = block_sum<uint64_t, kScanNT>(s, lds);
    if (kNz && tile_E) e = block_sum<double, kScanNT>(e, lds_e);
    if (threadIdx.x == 0) {
      tile_sum[t] = s;
      if (kNz && tile_E) tile_E[t] = e;
    }
  }
}

// One block of 256 threads scans the tile sums in place (exclusive) and
// writes the grand total to *total; with tile_E, also the sum of the tiles'
// energies to *E_total (a fixed tree per 256 tiles, chained in tile order:
// deterministic for a given n).
__global__ __launch_bounds__(kScanNT) void k_scan_top(uint64_t* __restrict__ tile_sum,
                                                      uint64_t n, const uint64_t* n_ptr,
                                                      uint64_t* __restrict__ total,
                                                      const double* __restrict__ tile_E,
                                                      double* __restrict__ E_total) {
  __shared__ uint64_t lds[kScanNT / kWave];
  __shared__ double lds_e[kScanNT / kWave];
  n = scan_n(n, n_ptr);
  const uint64_t n_tiles = (n + kScanTile - 1) / kScanTile;
  uint64_t carry = 0;
  double e_carry = 0.0;
  for (uint64_t b = 0; b < n_tiles; b += kScanNT) {
    const uint64_t i = b + threadIdx.x;
    const uint64_t v = (i < n_tiles) ? tile_sum[i] : 0;
    uint64_t tot;
    const uint64_t ex = block_excl_scan<uint64_t, kScanNT>(v, lds, tot);
    if (i < n_tiles) tile_sum[i] = carry + ex;
    carry += tot;
    if (tile_E) e_carry += block_sum<double, kScanNT>((i < n_tiles) ? tile_E[i] : 0.0, lds_e);
  }
  if (threadIdx.x == 0) {
    *total = carry;
    if (tile_E) *E_total = e_carry;
  }
}

template <bool kNz>
__global__ __launch_bounds__(kScanNT) void k_scan_down(const uint64_t* __restrict__ in,
                                                       uint64_t n, const uint64_t* n_ptr,
                                                       const uint64_t* __restrict__ tile_off,
                                                       uint64_t* __restrict__ out) {
  __shared__ uint64_t lds[kScanNT / kWave];
  n = scan_n(n, n_ptr);
  const uint64_t n_tiles = (n + kScanTile - 1) / kScanTile;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const uint64_t base = t * kScanTile;
    uint64_t carry = tile_off[t];
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      const uint64_t i = base + (uint64_t)it * kScanNT + threadIdx.x;
      const uint64_t v = (i < n) ? scan_item<kNz>(in[i]) : 0;
      uint64_t tot;
      const uint64_t ex = block_excl_scan<uint64_t, kScanNT>(v, lds, tot);
      if (i < n) out[i] = carry + ex;
      carry += tot;
    }
  }
}

__global__ void k_fill_f64(double* __restrict__ p, uint64_t n, double v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Up to kZeroSegs arrays of 64-bit words cleared by one launch (blockIdx.y
// picks the array, a grid-stride loop over it in x): the tallies, counters,
// census marks and totals a step clears before its transport launch, which
// were five fill launches back to back.
constexpr uint32_t kZeroSegs = 6;
struct ZeroSegs {
  unsigned long long* p[kZeroSegs];
  uint64_t n[kZeroSegs];
  uint32_t count;
};

__global__ __launch_bounds__(256) void k_zero_segs(const ZeroSegs z) {
  if (blockIdx.y >= z.count) return;
  unsigned long long* __restrict__ p = z.p[blockIdx.y];
  const uint64_t n = z.n[blockIdx.y];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Source (src/source.h).
// Photons per cell = max(1, int(n_user * E / E_tot)) when E > 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t photons_for(uint64_t n_user, double E, double total_E) {
  if (!(E > 0.0)) return 0u;
  uint32_t n = (uint32_t)(int)((double)n_user * E / total_E);
  return n ? n : 1u;
}

// counts[c] = photons of cell c: emission + face source (mode 0) or initial
// census (mode 1, emission_E then holds the census energy).
__global__ void k_source_counts(const double* __restrict__ emission_E,
                                const double* __restrict__ source_E, uint32_t n_cells,
                                SourceInst one, const SourceInst* __restrict__ inst,
                                const uint32_t* __restrict__ cell_inst, int mode,
                                uint64_t* __restrict__ counts, SourceCell* __restrict__ sc) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const SourceInst sp = inst ? inst[cell_inst ? cell_inst[c] : 0u] : one;
  const uint32_t n_em = photons_for(sp.n_user, emission_E[c], sp.total_E);
  SourceCell o{0.0, 0.0, n_em, 0u};
  if (n_em) o.e_em = emission_E[c] / (double)n_em;  // a photon's E (src/source.h:169, 265)
  uint64_t n = n_em;
  if (mode == 0) {
    const uint32_t n_src = photons_for(sp.n_user, source_E[c], sp.total_E);
    if (n_src) o.e_src = source_E[c] / (double)n_src;  // (src/source.h:280)
    n += n_src;
  }
  counts[c] = n;
  sc[c] = o;
}

// One work-item per new photon p: locate its cell in the exclusive prefix
// `off` (last cell with off[c] <= p), then sample it.  Stream number is the
// instance's stream_base + its photon index within the instance:
// 1e13*cycle + n_user*rank + i for new photons (src/source.h:221-222, 267,
// 283), n_user*rank + i for the initial census (src/source.h:144, 171).
// Output goes to slot out_base + p of W.
// With `P` set (a whole step captured as a graph), n_new, out_base and dt
// are read on the device.  Grid-stride loop.
__device__ __forceinline__ void source_one(const DevCells& cells,
                                           const SourceCell* __restrict__ sc,
                                           const int8_t* __restrict__ face_of,
                                           const uint64_t* __restrict__ off, uint64_t p,
                                           uint32_t c, const SourceInst& one,
                                           const SourceInst* __restrict__ inst, double dt,
                                           int mode, const PhotonSoA& W, uint64_t out_base);

// hint[j] = the cell of new photon j * kHintStride, for j < n_hint (the
// count read on the device when n_ptr is set): the cell of any photon p lies
// in [hint[p / kHintStride], hint[p / kHintStride + 1]].
__global__ void k_source_hints(const uint64_t* __restrict__ off, uint32_t n_cells,
                               uint64_t n_new, const uint64_t* __restrict__ n_ptr,
                               uint32_t* __restrict__ hint);

// upper_bound(off[lo, hi), p) - 1: the cell whose photons include p.
__device__ __forceinline__ uint32_t cell_of_photon(const uint64_t* __restrict__ off, uint64_t p,
                                                   uint32_t lo, uint32_t hi) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= p)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

__global__ void k_source_hints(const uint64_t* __restrict__ off, uint32_t n_cells,
                               uint64_t n_new, const uint64_t* __restrict__ n_ptr,
                               uint32_t* __restrict__ hint) {
  if (n_ptr) n_new = *n_ptr < n_new ? *n_ptr : n_new;
  const uint64_t n_hint = (n_new + kHintStride - 1) / kHintStride;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_hint; j += stride)
    hint[j] = cell_of_photon(off, j * kHintStride, 0u, n_cells);
}

// The cell of new photon p of n_new, from the hint array: one or two search
// steps where a cell holds many photons.
__device__ __forceinline__ uint32_t hinted_cell(const SourceDesc& sd, uint32_t n_cells,
                                                uint64_t n_new, uint64_t p) {
  const uint64_t j = p / kHintStride;
  const uint64_t n_hint = (n_new + kHintStride - 1) / kHintStride;
  const uint32_t lo = sd.hint[j];
  const uint32_t hi = (j + 1 < n_hint) ? sd.hint[j + 1] + 1u : n_cells;
  return cell_of_photon(sd.off, p, lo, hi);
}

// A block's 256 consecutive photons fall in a few consecutive cells.  With
// `hint` (k_source_hints over the same offsets and count, kHintStride = the
// block size), the block's first photon p0 is a multiple of kHintStride, so
// hint[p0 / kHintStride] is its cell and hint[p0 / kHintStride + 1] bounds
// the cell of its last: one scalar load pair per block, then each photon
// searches that range (one or two steps where a cell holds many photons).
// Without hints (the initial census), two waves find the cells of the
// block's first and last photon over the whole offset array (~18 dependent
// loads at 64^3 cells, whose latency the rest of the block waits out).
// n_ptr: the count is read on the device, bounded by n_new (the room in W).
__global__ __launch_bounds__(256) void k_source_fill(
    DevCells cells, const SourceCell* __restrict__ sc, const int8_t* __restrict__ face_of,
    const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ hint, uint64_t n_new, SourceInst one,
    const SourceInst* __restrict__ inst, double dt, int mode, PhotonSoA W, uint64_t out_base,
    const StepParams* __restrict__ P, const uint64_t* __restrict__ n_ptr) {
  __shared__ uint32_t s_first, s_last;
  if (P) {
    out_base = P->n_cen;
    n_new = P->n_new;
    dt = P->dt;
  }
  if (n_ptr) n_new = *n_ptr < n_new ? *n_ptr : n_new;
  const uint64_t n_hint = (n_new + kHintStride - 1) / kHintStride;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p0 = (uint64_t)blockIdx.x * blockDim.x; p0 < n_new; p0 += stride) {
    uint32_t c_first, c_end;
    if (hint) {
      const uint64_t j = p0 / kHintStride;
      c_first = hint[j];
      c_end = (j + 1 < n_hint) ? hint[j + 1] + 1u : cells.n_cells;
    } else {
      const uint64_t p_last = (p0 + blockDim.x < n_new ? p0 + blockDim.x : n_new) - 1;
      if (threadIdx.x == 0) s_first = cell_of_photon(off, p0, 0u, cells.n_cells);
      if (threadIdx.x == kWave) s_last = cell_of_photon(off, p_last, 0u, cells.n_cells);
      __syncthreads();
      c_first = s_first;
      c_end = s_last + 1u;
      __syncthreads();
    }
    const uint64_t p = p0 + threadIdx.x;
    if (p < n_new) {
      const uint32_t c = cell_of_photon(off, p, c_first, c_end);
      source_one(cells, sc, face_of, off, p, c, one, inst, dt, mode, W, out_base);
    }
  }
}

// Photon p of the step's source, from its draws of RNG(seed, stream) at
// counters 0, 1, ... (seven for an emission photon, six for a face-source or
// initial-census one), taken in order from draw(): the sampling of
// set_initial_census_photon / set_emission_photon /
// set_boundary_source_photon (src/source.h:27-81) in the reference's order.
// Shared by the source kernel (draw() runs the generator) and the transport
// kernels that make new photons where they are transported (draw() reads the
// wave's draws); a draw functor rather than an array of draws, which a
// branch-dependent index would put in scratch memory.
struct SourcePhoton {
  double x, y, z, u, v, w, E, life;
  uint64_t ctr, stream;
  uint32_t cell, group;
  uint8_t stype;
};

// The stream and source record of new photon p in cell c.
__device__ __forceinline__ SourceInst source_record(const DevCells& cells, uint32_t c,
                                                    const SourceInst& one,
                                                    const SourceInst* __restrict__ inst) {
  return inst ? inst[cells.cell_inst ? cells.cell_inst[c] : 0u] : one;
}

__device__ __forceinline__ uint64_t source_stream(const SourceInst& sp,
                                                  const uint64_t* __restrict__ off, uint64_t p) {
  return sp.stream_base + (p - off[sp.first_cell]);
}

template <class Draw>
__device__ __forceinline__ SourcePhoton make_source_photon(
    const DevCells& cells, const SourceCell* __restrict__ sc, const int8_t* __restrict__ face_of,
    const uint64_t* __restrict__ off, uint64_t p, uint32_t c, const SourceInst& sp, double dt,
    int mode, Draw&& draw) {
  const uint64_t local = p - off[c];
  const CellGeo& g = cells.geo[c];
  const double n0 = g.nodes[0], n1 = g.nodes[1], n2 = g.nodes[2], n3 = g.nodes[3],
               n4 = g.nodes[4], n5 = g.nodes[5];
  SourcePhoton o;
  const SourceCell k = sc[c];
  if (mode == 1 || local < k.n_em) {  // set_initial_census_photon / set_emission_photon
    o.E = k.e_em;
    o.x = n0 + draw() * (n1 - n0);  // get_uniform_position_in_cell
    o.y = n2 + draw() * (n3 - n2);
    o.z = n4 + draw() * (n5 - n4);
    const double mu = draw() * 2.0 - 1.0;  // get_uniform_angle
    const double phi = draw() * 2.0 * kPi;
    const double sin_theta = sqrt(1.0 - mu * mu);
    o.u = sin_theta * cos(phi);
    o.v = sin_theta * sin(phi);
    o.w = mu;
    if (mode == 1) {
      o.life = kC * dt;
      o.stype = 0;
      o.ctr = 6;
    } else {
      o.life = draw() * kC * dt;
      o.stype = 2;
      o.ctr = 7;
    }
  } else {  // set_boundary_source_photon
    o.E = k.e_src;
    const int face = face_of[c];
    // get_uniform_position_on_face (src/sampling_functions.h:32-53)
    const double a = draw();
    const double b = draw();
    if (face == 0 || face == 1) {
      o.x = (face == 0) ? n0 : n1;
      o.y = n2 + a * (n3 - n2);
      o.z = n4 + b * (n5 - n4);
    } else if (face == 2 || face == 3) {
      o.x = n0 + a * (n1 - n0);
      o.y = (face == 2) ? n2 : n3;
      o.z = n4 + b * (n5 - n4);
    } else {
      o.x = n0 + a * (n1 - n0);
      o.y = n2 + b * (n3 - n2);
      o.z = (face == 4) ? n4 : n5;
    }
    // get_source_angle_on_face (src/sampling_functions.h:94-122): inward
    // sign on x faces only, as the reference does
    const double theta = acos(sqrt(draw()));
    const double phi = draw() * 2.0 * kPi;
    const double sign = (face % 2) ? -1.0 : 1.0;
    const double ct = cos(theta), st = sin(theta);
    if (face == 0 || face == 1) {
      o.u = ct * sign;
      o.v = st * sin(phi);
      o.w = st * cos(phi);
    } else if (face == 2 || face == 3) {
      o.u = st * sin(phi);
      o.v = ct;
      o.w = st * cos(phi);
    } else {
      o.u = st * cos(phi);
      o.v = st * sin(phi);
      o.w = ct;
    }
    o.life = draw() * kC * dt;
    o.stype = 1;
    o.ctr = 6;
  }
  o.group = (uint32_t)floor(draw() * (double)cells.G);
  o.stream = source_stream(sp, off, p);
  o.cell = c;
  return o;
}

constexpr int kSourceDraws = 7;  // the most draws a source photon takes

// Draws photon p in cell c takes: 7 for emission, 6 for a face-source or an
// initial-census photon.
__device__ __forceinline__ int source_draw_count(const SourceCell* __restrict__ sc,
                                                 const uint64_t* __restrict__ off, uint64_t p,
                                                 uint32_t c, int mode) {
  if (mode == 1) return 6;
  return (p - off[c] < sc[c].n_em) ? 7 : 6;
}

__device__ __forceinline__ void source_one(const DevCells& cells,
                                           const SourceCell* __restrict__ sc,
                                           const int8_t* __restrict__ face_of,
                                           const uint64_t* __restrict__ off, uint64_t p,
                                           uint32_t c, const SourceInst& one,
                                           const SourceInst* __restrict__ inst, double dt,
                                           int mode, const PhotonSoA& W, uint64_t out_base) {
  const SourceInst sp = source_record(cells, c, one, inst);
  Rng rng{0ull, source_stream(sp, off, p), sp.ctr_hi};
  const SourcePhoton o = make_source_photon(cells, sc, face_of, off, p, c, sp,
                                            dt, mode, [&]() { return rng.next(); });
  const uint64_t i = out_base + p;
  // an initial census (mode 1) is in the reference's order by p; new photons
  // carry p too
  W.key[i] = (uint32_t)p;
  W.x[i] = o.x;
  W.y[i] = o.y;
  W.z[i] = o.z;
  W.u[i] = o.u;
  W.v[i] = o.v;
  W.w[i] = o.w;
  W.E[i] = o.E;
  W.E0[i] = o.E;
  W.life[i] = o.life;
  W.ctr[i] = o.ctr;
  W.stream[i] = o.stream;
  W.cell[i] = o.cell;
  W.group[i] = o.group;
  W.desc[i] = kPass;
  W.stype[i] = o.stype;
}

// ---------------------------------------------------------------------------
// Tally sinks.  A flush adds one photon's accumulated (absorbed, track)
// energy of one cell visit; push() is called by every lane of a wave at the
// same program point, with has = false for lanes that did not flush.
//
// GlobalTally: a 128-entry ring of flush records per wave in LDS.  Lanes
// append their records (ballot + mbcnt); whenever 64 are queued the whole
// wave drains them with one full-wave global_atomic_add_f64 per tally array.
// On gfx950 vmcnt counts atomics in order with loads, so an atomic issued by
// a few divergent lanes puts a full atomic round trip in front of the next
// opacity fetch; batching turns ~10 such stalls into one and fills all 64
// lanes of every atomic instruction.
//
// LdsTally: block-private tallies of the whole mesh in LDS (ds_add_f64),
// added to HBM once per block -- for meshes small enough that every block
// can hold them, where global atomics would serialise on a few hot cells.
// ---------------------------------------------------------------------------
struct alignas(8) FlushRec {
  double a;
  double t;
  uint32_t c;
  uint32_t pad;
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A flush: one cell visit of one photon.  cont = implicit-capture deposit,
// track = sum of deposit/(sigma_a f) over the visit, kill = the weight
// deposited when the photon is killed (0 otherwise).  The reference adds the
// kill weight to the visit's absorbed energy (src/history_based_transport.h:
// 85-88), so abs = cont + kill.
struct Flush {
  bool has;
  uint32_t c;
  double cont, track, kill;
};

// kTrack = true: two tallies per flush, abs_E += cont + kill and
// track_E += track (any opacities).
// kTrack = false: for cells whose absorption opacity is the same in every
// group (Branson's faux multigroup, Cell::set_op_a) the track tally of a
// visit is cont/(sigma_a f), so only cont is tallied per flush (into abs_E)
// and kill weights separately (into track_E, rare); k_tally_finalize turns
// them into abs = cont + kill, track = cont/(sigma_a f).  Half the atomics
// and no per-segment division.
template <bool kTrackT>
struct GlobalTally {
  static constexpr bool kLds = false;
  static constexpr bool kTrack = kTrackT;
  static constexpr uint32_t kRing = 128;  // power of two, >= 2 * 64
  double* abs_E;
  double* track_E;
  double* s_abs;  // unused
  double* s_track;
  uint32_t head, count;  // wave-uniform ring state

  __device__ __forceinline__ static FlushRec* ring() {
    __shared__ FlushRec r[256 / kWave][kRing];
    return r[threadIdx.x / kWave];
  }
  __device__ __forceinline__ void init(uint32_t) {
    head = 0;
    count = 0;
  }
  __device__ __forceinline__ void drain_one(const FlushRec& e) {
    if constexpr (kTrack) {
      unsafeAtomicAdd(&abs_E[e.c], e.a);
      unsafeAtomicAdd(&track_E[e.c], e.t);
    } else {
      unsafeAtomicAdd(&abs_E[e.c], e.a);
      if (e.t != 0.0) unsafeAtomicAdd(&track_E[e.c], e.t);
    }
  }
  __device__ __forceinline__ void push(const Flush& f) {
    const uint64_t m = __ballot(f.has);
    if (m == 0) return;
    FlushRec* r = ring();
    if (f.has) {
      const uint32_t pos = (head + count + lane_rank(m)) & (kRing - 1);
      r[pos].a = kTrack ? f.cont + f.kill : f.cont;
      r[pos].t = kTrack ? f.track : f.kill;
      r[pos].c = f.c;
    }
    count += (uint32_t)__popcll(m);
    if (count >= kWave) {
      wave_lds_sync();
      drain_one(r[(head + (threadIdx.x & (kWave - 1))) & (kRing - 1)]);
      head = (head + kWave) & (kRing - 1);
      count -= kWave;
      wave_lds_sync();
    }
  }
  __device__ __forceinline__ void finish(uint32_t) {
    wave_lds_sync();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if (lane < count) drain_one(ring()[(head + lane) & (kRing - 1)]);
  }
};

// HashTally: the sink of GlobalTally<false> (cont per flush into abs_E, kill
// weights into track_E, then k_tally_finalize) with the deposits of many
// flushes combined before they reach HBM.  Scattered FP64 atomics run at
// ~23e9/s on this device, a tenth of plain stores, and far slower when many
// of them hit the same few cache lines -- a source face in a thick wall
// puts nearly every flush of a step into one layer of cells.  The photons of
// one wave come from consecutive source indices (born in the same few
// cells), so their flushes keep hitting a small neighbourhood.  Each wave
// owns an open-addressing table of kSlots cells in LDS: a flush claims its
// cell's slot (ds_cmpst) and adds its deposit and kill weight
// (ds_add_f64); the wave writes every occupied slot out with global atomics
// when the table could pass half full, and at exit.  Key 0 is empty; cell c
// has key c + 1.  At most kSlots / 2 + 64 slots are occupied during an
// insertion, so a probe always finds its key or an empty slot.
struct HashTally {
  static constexpr bool kLds = false;
  static constexpr bool kTrack = false;
  static constexpr uint32_t kSlots = 256;  // per wave, power of two
  static constexpr uint32_t kLimit = kSlots / 2;
  double* abs_E;
  double* track_E;
  double* s_abs;  // unused
  double* s_track;
  uint32_t used;  // wave-uniform: occupied slots

  struct Table {
    uint32_t key[kSlots];
    double cont[kSlots];
    double kill[kSlots];
  };
  __device__ __forceinline__ static Table* table() {
    __shared__ Table t[256 / kWave];
    return &t[threadIdx.x / kWave];
  }
  __device__ __forceinline__ static uint32_t slot_of(uint32_t key) {
    return (key * 2654435761u) >> 24;  // Fibonacci hashing to 8 bits
  }
  static_assert(kSlots == 256, "slot_of yields 8 bits");

  __device__ __forceinline__ void init(uint32_t) {
    Table* t = table();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    for (uint32_t s = lane; s < kSlots; s += kWave) {
      t->key[s] = 0u;
      t->cont[s] = 0.0;
      t->kill[s] = 0.0;
    }
    used = 0;
    wave_lds_sync();
  }

  __device__ __forceinline__ void evict() {
    wave_lds_sync();
    Table* t = table();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    for (uint32_t s = lane; s < kSlots; s += kWave) {
      const uint32_t k = t->key[s];
      if (k != 0u) {
        unsafeAtomicAdd(&abs_E[k - 1u], t->cont[s]);
        const double kl = t->kill[s];
        if (kl != 0.0) unsafeAtomicAdd(&track_E[k - 1u], kl);
        t->key[s] = 0u;
        t->cont[s] = 0.0;
        t->kill[s] = 0.0;
      }
    }
    used = 0;
    wave_lds_sync();
  }

  // Every lane at the same program point; lanes with has = false only take
  // part in the wave-uniform bookkeeping.
  __device__ __forceinline__ void push(const Flush& f) {
    const uint64_t m = __ballot(f.has);
    if (m == 0) return;
    if (used + (uint32_t)__popcll(m) > kLimit) evict();
    Table* t = table();
    bool fresh = false;
    if (f.has) {
      const uint32_t key = f.c + 1u;
      uint32_t s = slot_of(key);
      for (;;) {
        const uint32_t old = atomicCAS(&t->key[s], 0u, key);
        if (old == 0u || old == key) {
          fresh = old == 0u;
          unsafeAtomicAdd(&t->cont[s], f.cont);
          if (f.kill != 0.0) unsafeAtomicAdd(&t->kill[s], f.kill);
          break;
        }
        s = (s + 1u) & (kSlots - 1u);
      }
    }
    used += (uint32_t)__popcll(__ballot(fresh));
  }
  __device__ __forceinline__ void finish(uint32_t) { evict(); }
};

// abs = cont + kill, track = cont / (sigma_a f) after a GlobalTally<false>
// or HashTally transport (sigma_a of group 0: every group holds the same
// value).
__global__ void k_tally_finalize(const CellGeo* __restrict__ geo,
                                 const double2* __restrict__ sigma, uint32_t G,
                                 uint32_t n_cells, double* __restrict__ abs_E,
                                 double* __restrict__ track_E) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const double cont = abs_E[c], kill = track_E[c];
  abs_E[c] = cont + kill;
  track_E[c] = cont / (sigma[(size_t)c * G].x * geo[c].fleck);
}

struct LdsTally {
  static constexpr bool kLds = true;
  static constexpr bool kTrack = true;
  double* abs_E;  // global destination
  double* track_E;
  double* s_abs;  // dynamic LDS, n_cells each
  double* s_track;
  uint32_t head, count;  // unused
  __device__ __forceinline__ void init(uint32_t n_cells) {
    for (uint32_t c = threadIdx.x; c < n_cells; c += blockDim.x) {
      s_abs[c] = 0.0;
      s_track[c] = 0.0;
    }
    __syncthreads();
  }
  __device__ __forceinline__ void push(const Flush& f) {
    if (f.has) {
      __hip_atomic_fetch_add(&s_abs[f.c], f.cont + f.kill, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&s_track[f.c], f.track, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __device__ __forceinline__ void finish(uint32_t n_cells) {
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < n_cells; c += blockDim.x) {
      const double a = s_abs[c], t = s_track[c];
      if (a != 0.0 || t != 0.0) {
        unsafeAtomicAdd(&abs_E[c], a);
        unsafeAtomicAdd(&track_E[c], t);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// AOS <-> SoA for the reference's 120-byte Photon.
// ---------------------------------------------------------------------------
struct AosPhoton {
  uint32_t cell, group, source_type;
  uint8_t descriptor[4];
  double pos[3];
  double angle[3];
  double E, E0, life_dx;
  uint64_t rng[4];
};
static_assert(sizeof(AosPhoton) == 120, "reference Photon is 120 bytes");

__global__ void k_aos_to_soa(const AosPhoton* __restrict__ a, uint64_t n, PhotonSoA W,
                             uint64_t base) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const AosPhoton p = a[i];
  const uint64_t o = base + i;
  W.key[o] = (uint32_t)i;  // an imported census is in the reference's order
  W.x[o] = p.pos[0];
  W.y[o] = p.pos[1];
  W.z[o] = p.pos[2];
  W.u[o] = p.angle[0];
  W.v[o] = p.angle[1];
  W.w[o] = p.angle[2];
  W.E[o] = p.E;
  W.E0[o] = p.E0;
  W.life[o] = p.life_dx;
  W.ctr[o] = p.rng[0];
  W.stream[o] = p.rng[2];
  W.cell[o] = p.cell;
  W.group[o] = p.group;
  W.desc[o] = p.descriptor[0];
  W.stype[o] = (uint8_t)p.source_type;
}

// by_key: photon s goes to a[key_off + key[s]] (a resident census, back in
// the reference's order), else to a[i].
__global__ void k_soa_to_aos(PhotonSoA W, uint64_t base, uint64_t n, uint64_t ctr_hi,
                             DevCells cells, AosPhoton* __restrict__ a, int by_key,
                             uint64_t key_off) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = base + i;
  if (by_key && W.key[s] >= n) return;  // keys of a census are a permutation of [0, n)
  const uint64_t o = by_key ? key_off + W.key[s] : i;
  AosPhoton p;
  p.cell = W.cell[s];
  // per-instance seed; a photon that left through a PROCESSOR face may name
  // a cell of another mesh, which keeps the run's word
  if (p.cell < cells.n_cells) ctr_hi = photon_ctr_hi(cells, p.cell, ctr_hi);
  p.group = W.group[s];
  p.source_type = W.stype[s];
  p.descriptor[0] = W.desc[s];
  p.descriptor[1] = p.descriptor[2] = p.descriptor[3] = 0;
  p.pos[0] = W.x[s];
  p.pos[1] = W.y[s];
  p.pos[2] = W.z[s];
  p.angle[0] = W.u[s];
  p.angle[1] = W.v[s];
  p.angle[2] = W.w[s];
  p.E = W.E[s];
  p.E0 = W.E0[s];
  p.life_dx = W.life[s];
  p.rng[0] = W.ctr[s];
  p.rng[1] = ctr_hi;
  p.rng[2] = W.stream[s];
  p.rng[3] = 0;
  a[o] = p;
}

// Copies photon `from` of W to slot `to` of C with energy E and key k.
__device__ __forceinline__ void copy_photon(const PhotonSoA& W, uint64_t from, const PhotonSoA& C,
                                            uint64_t to, double E, uint32_t k) {
  C.x[to] = W.x[from];
  C.y[to] = W.y[from];
  C.z[to] = W.z[from];
  C.u[to] = W.u[from];
  C.v[to] = W.v[from];
  C.w[to] = W.w[from];
  C.E[to] = E;
  C.E0[to] = W.E0[from];
  C.life[to] = W.life[from];
  C.ctr[to] = W.ctr[from];
  C.stream[to] = W.stream[from];
  C.cell[to] = W.cell[from];
  C.group[to] = W.group[from];
  C.desc[to] = W.desc[from];
  C.stype[to] = W.stype[from];
  C.key[to] = k;
}

// The launch's census totals in the readback record (census_E already
// there).
__global__ void k_census_pack(const CensusAcc* __restrict__ acc, CensusPartial* __restrict__ out) {
  out->exit_E = acc->exit_E;
  out->n_census = acc->n_census;
  out->n_exit = acc->n_exit;
  out->n_killed = acc->n_killed;
  out->n_other = acc->n_other;
  out->overflow = acc->overflow;
}

// Hole fill, 1: per record, how many of its holes lie below n_census (those
// get a photon from above), and the hole count.
__global__ void k_hole_heads(const HoleRec* __restrict__ rec, CensusAcc* __restrict__ acc,
                             uint64_t rec_cap, uint64_t* __restrict__ head_len) {
  const uint64_t nr = acc->n_rec < rec_cap ? acc->n_rec : rec_cap;
  const uint64_t nc = acc->n_census;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) acc->n_holes = acc->n_reserved - nc;
  for (uint64_t r = g; r < nr; r += stride) {
    const uint64_t a = rec[r].start, b = a + rec[r].len;
    head_len[r] = a < nc ? (b < nc ? b : nc) - a : 0;
  }
}

// Hole fill, 2: which slots of [n_census, n_reserved) hold a photon.
__global__ void k_hole_tail(const uint32_t* __restrict__ key, const CensusAcc* __restrict__ acc,
                            uint64_t bound, uint64_t* __restrict__ live) {
  const uint64_t nc = acc->n_census;
  const uint64_t nh = acc->n_holes < bound ? acc->n_holes : bound;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nh; p += stride)
    live[p] = key[nc + p] != kHole;
}

// Hole fill, 3: the j-th photon above n_census moves to the j-th hole below
// it (holes ordered by record, then by slot).  Order in C is free: keys
// carry the census order.
__global__ void k_hole_fill(PhotonSoA C, const CensusAcc* __restrict__ acc, uint64_t bound,
                            const uint64_t* __restrict__ live, const uint64_t* __restrict__ rank,
                            const HoleRec* __restrict__ rec, const uint64_t* __restrict__ hole_off,
                            uint64_t rec_cap) {
  const uint64_t nc = acc->n_census;
  const uint64_t nh = acc->n_holes < bound ? acc->n_holes : bound;
  const uint64_t nr = acc->n_rec < rec_cap ? acc->n_rec : rec_cap;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nh; p += stride) {
    if (!live[p]) continue;
    const uint64_t j = rank[p];
    uint64_t lo = 0, hi = nr;  // last record with hole_off <= j
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (hole_off[mid] <= j)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo == 0) continue;
    const uint64_t r = lo - 1, to = rec[r].start + (j - hole_off[r]), from = nc + p;
    if (to < nc) copy_photon(C, from, C, to, C.E[from], C.key[from]);
  }
}

// perm[key[s]] = s: the resident census in the reference's order.
__global__ void k_census_perm(const uint32_t* __restrict__ key, uint64_t n,
                              uint32_t* __restrict__ perm) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && key[s] < n) perm[key[s]] = (uint32_t)s;
}

// Census marks by transport index, 64 to a word: each word's bit mask and
// its count (their exclusive scan gives each word's first rank), and the
// word count into acc->n_words.  The marks are zero up to a whole word past
// the launch's photons.
__global__ void k_mark_words(const uint8_t* __restrict__ mark, uint64_t cap,
                             const uint64_t* __restrict__ n_ptr, uint64_t* __restrict__ words,
                             uint64_t* __restrict__ counts, CensusAcc* __restrict__ acc) {
  uint64_t n = cap;
  if (n_ptr) n = *n_ptr < cap ? *n_ptr : cap;
  const uint64_t nw = (n + 63) / 64;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (first == 0) acc->n_words = nw;
  for (uint64_t w = first; w < nw; w += stride) {
    const uint4* q = reinterpret_cast<const uint4*>(mark + w * 64);
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 v = q[k];
      const uint32_t x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {  // four 0/1 bytes to four bits
        const uint32_t b = (x[j] & 1u) | ((x[j] >> 7) & 2u) | ((x[j] >> 14) & 4u) |
                           ((x[j] >> 21) & 8u);
        m |= (uint64_t)b << (16 * k + 4 * j);
      }
    }
    words[w] = m;
    counts[w] = (uint64_t)__popcll(m);
  }
}

// Each census photon's key, its transport index t, becomes its rank in the
// reference's census order: the census photons with smaller t, counted from
// the mark words.  Its energy goes to position rank of E_rank, so that the
// census energy can be summed in census order.
__global__ void k_census_rekey(uint32_t* __restrict__ key, const double* __restrict__ E,
                               const uint64_t* __restrict__ words,
                               const uint64_t* __restrict__ word_rank,
                               const unsigned long long* __restrict__ count, uint64_t cap,
                               double* __restrict__ E_rank) {
  const uint64_t n = *count;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
    const uint32_t t = key[s];
    if (t < cap) {
      const uint64_t w = t >> 6, below = (1ull << (t & 63u)) - 1ull;
      const uint64_t r = word_rank[w] + (uint64_t)__popcll(words[w] & below);
      key[s] = (uint32_t)r;
      E_rank[r] = E[s];
    } else {
      key[s] = 0xFFFFFFFFu;
    }
  }
}

// ---------------------------------------------------------------------------
// RNG probes: device draws and raw Threefry words.
// ---------------------------------------------------------------------------
__global__ void k_rng_draws(uint64_t seed_hi, uint64_t stream, uint64_t ctr0, uint32_t n,
                            double* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rng r{ctr0 + i, stream, seed_hi};
  out[i] = r.next();
}

// Streaming copy: 16-byte loads and stores, grid-stride.
__global__ __launch_bounds__(256) void k_copy_probe(const double2* __restrict__ src,
                                                     double2* __restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

__global__ void k_threefry(const uint64_t* __restrict__ in, uint64_t n,
                           uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t o0, o1;
  Threefry::block(in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3], o0, o1);
  out[2 * i] = o0;
  out[2 * i + 1] = o1;
}

}  // namespace bh
