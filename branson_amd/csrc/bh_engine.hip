// bh_engine.hip -- handle, device memory and the C-ABI of the MI355X IMC
// engine (include/branson_hip.h).  Kernels live in bh_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/branson_hip.h"
#include "bh_kernels.h"
#include "bh_physics.h"
#include "bh_comb.h"
#include "bh_tracking.h"

#include <hipcub/hipcub.hpp>

using namespace bh;

// Device memory of bh_transport_photons_noalloc: the photons' AOS staging
// and SoA state, the tallies, counters, queue and FP32 replay lists of one
// call, carved from this array instead of allocated.  Every GPU the library
// runs on holds its own copy until the process ends.  In host code the name
// is only a host-side stand-in: the device address comes from
// hipGetSymbolAddress, once per device (scratch_on_device), and only that
// address is ever passed to a kernel or a copy.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on `device`, which must be current.
// Resolved on the first call for a device and kept.
hipError_t scratch_on_device(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}


constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }
inline uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Field sizes of PhotonSoA, in declaration order.
constexpr int kFields = 16;
constexpr size_t kFieldBytes[kFields] = {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 4, 4, 1, 1, 4};

// One slab of device memory holding every field of `cap` photons.
struct PhotonStore {
  void* slab = nullptr;
  uint64_t cap = 0;
  PhotonSoA soa{};

  static size_t bytes_for(uint64_t cap) {
    size_t b = 0;
    for (size_t fb : kFieldBytes) b += align_up(fb * cap);
    return b;
  }
  static PhotonSoA carve(void* slab, uint64_t cap) {
    char* p = static_cast<char*>(slab);
    void* f[kFields];
    for (int k = 0; k < kFields; ++k) {
      f[k] = p;
      p += align_up(kFieldBytes[k] * cap);
    }
    PhotonSoA s;
    s.x = (double*)f[0];
    s.y = (double*)f[1];
    s.z = (double*)f[2];
    s.u = (double*)f[3];
    s.v = (double*)f[4];
    s.w = (double*)f[5];
    s.E = (double*)f[6];
    s.E0 = (double*)f[7];
    s.life = (double*)f[8];
    s.ctr = (uint64_t*)f[9];
    s.stream = (uint64_t*)f[10];
    s.cell = (uint32_t*)f[11];
    s.group = (uint32_t*)f[12];
    s.desc = (uint8_t*)f[13];
    s.stype = (uint8_t*)f[14];
    s.key = (uint32_t*)f[15];
    return s;
  }
  static void* field(const PhotonSoA& s, int k) {
    void* f[kFields] = {s.x, s.y, s.z, s.u, s.v, s.w, s.E, s.E0, s.life,
                        s.ctr, s.stream, s.cell, s.group, s.desc, s.stype, s.key};
    return f[k];
  }
};

}  // namespace

template <typename T>
struct DArr {
  T* p = nullptr;
  size_t cap = 0;
};

struct bh_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  bh_options opt{};
  std::string err;

  // mesh
  uint32_t n_cells = 0, G = 0;
  DArr<CellGeo> geo;
  DArr<double2> sigma;
  DArr<uint32_t> next;
  bool has_mesh = false;
  bool gray_sigma_a = false;     // one tally atomic per flush (see with_tally)
  bool uniform_sigma_a = false;  // sigma_a of every cell equal in all groups
  bool lattice = false;          // mesh is a lattice: CellRec + axis arrays
  bool inst_lattice = false;     // every instance a lattice: CellRec + per-instance axes
  int last_lattice = 0;          // records of the last launch (bh_step_stats.lattice)
  double src_neff_frac = 1.0;    // effective source cells / n_cells (tally choice)
  int last_tally = 0;            // tally form of the last launch (bh_step_stats.tally)
  uint64_t n_allocs = 0;         // device allocations of step calls since the last stats report

  // device-resident step physics (bh_set_material / bh_step_resident)
  bool has_material = false;
  DArr<double> mat;      // [9 n]: opA opB opC opS cv rho T_e T_r T_s
  DArr<double> mat_e;    // [n] T cV V rho of the cell
  DArr<double> tr_diag;  // [n] diagnostic radiation temperature
  double mat_rf = 1.0;
  uint32_t mat_ranks = 1, mat_rank = 0;
  PhysTotals* d_phys = nullptr;
  SourceInst* d_inst2 = nullptr;  // [census, new] source records, E_tot set on the device
  DArr<double> sum_part;          // block partials of the deterministic sums
  // comb (bh_comb_census): sort buffers, keep flags, per-cell energies
  DArr<uint32_t> comb_keys, comb_idx, comb_sorted, comb_perm, comb_rank_cell;
  DArr<uint8_t> comb_keep, comb_tmp;
  DArr<double> comb_cellE;
  DArr<unsigned long long> comb_ctr;
  DArr<uint64_t> cflag;       // keep flags of the comb
  DArr<uint64_t> cmark;       // census marks by transport index, one byte each
  DArr<uint64_t> mwords, mcounts;  // the marks 64 to a word: bit masks and counts
  DArr<double> erank;         // census energies by rank (census order)
  DArr<HoleRec> hrec;         // census slots a wave took and left unused
  DArr<uint64_t> hole_len, hole_off, tail_live, tail_rank;  // hole fill
  CensusAcc* d_cacc = nullptr;  // census totals of a census-writing launch
  uint64_t serial_etot_steps = 0; // steps whose E_tot needed the serial sum
  uint32_t dims[3] = {0, 0, 0};
  DArr<CellRec> rec;
  DArr<LatInst> lat;             // per-instance lattice shapes (inst_lattice)
  DArr<double> axes;             // x nodes, then y, then z

  // per-cell source energies
  DArr<double> em, src, cen;
  DArr<int8_t> face;
  bool has_energy = false, has_census_E = false;
  // the step's new photons: exclusive prefix of the per-cell counts and the
  // cell hints; with the persistent algorithm they are made inside the
  // transport kernel (src_deferred) rather than written to W
  DArr<uint64_t> soff;
  DArr<uint32_t> src_hint;
  SourceDesc sd{};
  bool src_deferred = false;

  // photons: W = [census | new], C = next census / scratch
  PhotonStore W, C;
  uint64_t n_cen = 0, n_new = 0;
  uint64_t ctr_hi = 0;

  // tallies
  DArr<double> abs_E, track_E;

  // scratch
  DArr<uint64_t> counts, off, tiles;
  DArr<SourceCell> src_cell;      // per-cell source photon counts and energies
  DArr<double> tile_E;        // per-tile census energies of the rank scan
  DArr<AosPhoton> aos;
  DArr<double> adj_g, grad_sa, grad_f;  // adjoint: dL/d abs_E in, gradients out

  // several independent problems in one mesh (n_inst = 0: one problem)
  uint32_t n_inst = 0;
  std::vector<uint32_t> inst_first;     // n_inst + 1 cell bounds
  std::vector<uint64_t> inst_hi;        // seed << 32 per instance
  std::vector<uint32_t> cell_inst_host;
  DArr<uint32_t> cell_inst;
  DArr<uint64_t> inst_ctr_hi;
  DArr<SourceInst> inst_src;            // [2 * n_inst]: census, then new photons
  DArr<uint32_t> inst_first_dev;        // n_inst + 1 cell bounds on the device
  DArr<InstReport> inst_rep;            // resident steps: per-instance totals
  DArr<double> inst_exit;               // resident steps: exit energy per instance
  SourceInst* h_rec = nullptr;          // pinned: the step's source records
  InstReport* h_inst_rep = nullptr;     // pinned: per-instance totals read back
  size_t h_inst_cap = 0;                // instances the two pinned buffers hold
  // event-split form: open cell visits, segment counts, last faces and the
  // active / scatter lists, one entry per photon slot
  DArr<double> ev_abs, ev_track;
  DArr<uint32_t> ev_segs, ev_list0, ev_list1, ev_scatter;
  DArr<uint8_t> ev_sc;
  unsigned long long* d_ev_count = nullptr;
  DArr<uint64_t> rep_idx;    // FP32 form: histories handed to the FP64 replay
  DArr<uint32_t> rep_first;
  DArr<double> rep_E;
  uint64_t* d_total = nullptr;      // scan totals
  CensusPartial* d_ctot = nullptr;  // census totals
  DevCounters* d_ctrs = nullptr;
  PersistentQueue* d_queue = nullptr;
  void* h_pinned = nullptr;  // small readbacks

  // the device phase of a step captured as HIP graphs (bh_step_graph), one
  // per photon-store parity (W and C swap every step)
  struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uintptr_t> key;  // the buffers and sizes it was captured with
  };
  StepGraph graphs[2];
  StepGraph resident_graphs[2];    // bh_step_resident(_instances) steps after the first
  StepParams* d_params = nullptr;  // what the graph's kernels read
  StepParams* h_params = nullptr;  // pinned; the graph's first node copies it in
  double* h_tally = nullptr;       // pinned: abs_E then track_E
  size_t h_tally_cap = 0;
  std::vector<double> em_host, src_host;  // the step's energies, for exact counts

  hipEvent_t ev[9] = {};          // 6, 7: the source phase (run_source); 8: its count read
  bool src_ms_pending = false;    // last_source_ms not yet read from ev[6], ev[7]
  uint64_t last_n_new = 0;        // new photons of the last run_source (its fill's grid)
  bool replay_timed = false;  // ev[3] marks the start of the last FP64 replay
  bool inst_exit_on = false;  // bh_transport adds exit energy per instance
  double last_source_ms = 0.0;
  int n_cu = 256;
  size_t lds_budget = 48 * 1024;  // dynamic LDS a tally may use per block

  // this device's address of `scratch` (bh_transport_photons_noalloc);
  // null when it could not be resolved, with the reason
  void* scratch = nullptr;
  std::string scratch_err;
};

namespace {

int fail(bh_handle* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  return code;
}

#define HIPCHK(h, call)                                                             \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail((h), BH_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,               \
                  hipGetErrorString(e_), __FILE__, __LINE__);                       \
  } while (0)

#define TRY(expr)            \
  do {                       \
    int r_ = (expr);         \
    if (r_ != BH_OK) return r_; \
  } while (0)

// Set-up, export and probe calls allocate as set-up, not as step work:
// their allocations are left out of bh_step_stats.n_allocs.
struct SetupAllocs {
  bh_handle* h;
  uint64_t n;
  explicit SetupAllocs(bh_handle* hh) : h(hh), n(hh ? hh->n_allocs : 0) {}
  ~SetupAllocs() {
    if (h) h->n_allocs = n;
  }
};

uint64_t take_allocs(bh_handle* h) {
  const uint64_t n = h->n_allocs;
  h->n_allocs = 0;
  return n;
}

// Makes a hold at least n elements (contents are not kept).  A growing
// array grows by half again, so one that follows a growing census is
// reallocated rarely: every hipMalloc / hipFree synchronises the device.
template <typename T>
int ensure(bh_handle* h, DArr<T>& a, size_t n) {
  if (n <= a.cap && a.p) return BH_OK;
  const size_t want = std::max<size_t>({n, a.cap + a.cap / 2, 1});
  if (a.p) HIPCHK(h, hipFree(a.p));
  a.p = nullptr;
  a.cap = 0;
  ++h->n_allocs;
  if (hipMalloc(&a.p, want * sizeof(T)) != hipSuccess) {
    a.p = nullptr;
    return fail(h, BH_ERR_ALLOC, "device allocation of %zu bytes failed", want * sizeof(T));
  }
  a.cap = want;
  return BH_OK;
}

template <typename T>
void release(DArr<T>& a) {
  if (a.p) (void)hipFree(a.p);
  a.p = nullptr;
  a.cap = 0;
}

// Grows a photon store to hold `need`, keeping the first `keep` photons.
int grow_store(bh_handle* h, PhotonStore& s, uint64_t need, uint64_t keep) {
  if (need <= s.cap && s.slab) return BH_OK;
  // grow by half again: a census that grows step after step reallocates
  // rarely (each reallocation is a device-wide synchronising hipMalloc)
  uint64_t cap = std::max<uint64_t>(need + need / 2, 1024);
  void* slab = nullptr;
  ++h->n_allocs;
  if (hipMalloc(&slab, PhotonStore::bytes_for(cap)) != hipSuccess)
    return fail(h, BH_ERR_ALLOC, "device allocation for %llu photons failed",
                (unsigned long long)cap);
  PhotonSoA ns = PhotonStore::carve(slab, cap);
  if (keep && s.slab) {
    for (int k = 0; k < kFields; ++k)
      HIPCHK(h, hipMemcpyAsync(PhotonStore::field(ns, k), PhotonStore::field(s.soa, k),
                               keep * kFieldBytes[k], hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  if (s.slab) HIPCHK(h, hipFree(s.slab));
  s.slab = slab;
  s.cap = cap;
  s.soa = ns;
  return BH_OK;
}

// Arrays indexed by photon are sized from the photon stores' capacity
// (bh_options.reserve_photons sizes the stores up front), not from the
// step's count, which grows with the census from step to step.
uint64_t photon_cap(const bh_handle* h, uint64_t n) {
  return std::max<uint64_t>({n, h->W.cap, h->C.cap, 1});
}

DevCells dev_cells(const bh_handle* h) {
  const double tol = h->opt.fp32_tol > 0.0 ? h->opt.fp32_tol : 1e-6;
  const bool inst = h->n_inst > 0;
  DevCells d{h->geo.p,        h->sigma.p, h->next.p,
             h->n_cells,      h->G,       (float)tol,
             inst ? h->cell_inst.p : nullptr, inst ? h->inst_ctr_hi.p : nullptr,
             nullptr,         {nullptr, nullptr, nullptr}, {0, 0, 0}};
  if (h->lattice) {
    d.rec = h->rec.p;
    d.axis[0] = h->axes.p;
    d.axis[1] = h->axes.p + h->dims[0] + 1;
    d.axis[2] = d.axis[1] + h->dims[1] + 1;
    for (int a = 0; a < 3; ++a) d.dims[a] = h->dims[a];
  } else if (h->inst_lattice && inst) {
    d.rec = h->rec.p;
    d.axis[0] = h->axes.p;
    d.lat = h->lat.p;
  }
  return d;
}

inline unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)div_up(n, block); }

// Exclusive scan of n uint64 values on the device; the total lands in
// h->d_total[slot].
// Exclusive scan of in[0, n) into out, total into d_total[slot].  With
// n_ptr the count is read on the device and n only bounds it (buffers and
// grid are sized for n).
// nonzero: an item counts 1 when its bits are nonzero (census energies by
// transport index); E_total then also receives the items' sum as doubles,
// summed in a fixed order.
int scan_u64(bh_handle* h, const uint64_t* in, uint64_t n, uint64_t* out, int slot,
             const uint64_t* n_ptr = nullptr, bool nonzero = false,
             double* E_total = nullptr) {
  const uint64_t n_tiles = std::max<uint64_t>(div_up(n, kScanTile), 1);
  // the single-block middle pass loops over the tile sums (of the count read
  // on the device where there is one: n may be a capacity); 2^20 tiles hold
  // every 32-bit photon index
  if (n_tiles > ((uint64_t)1 << 20))
    return fail(h, BH_ERR_ARG, "scan of %llu elements exceeds the supported size",
                (unsigned long long)n);
  TRY(ensure(h, h->tiles, n_tiles));
  double* tile_E = nullptr;
  if (nonzero && E_total) {
    TRY(ensure(h, h->tile_E, n_tiles));
    tile_E = h->tile_E.p;
  }
  const unsigned grid = (unsigned)std::min<uint64_t>(n_tiles, 4096);
  if (nonzero)
    k_scan_reduce<true><<<grid, kScanNT, 0, h->stream>>>(in, n, n_ptr, h->tiles.p, tile_E);
  else
    k_scan_reduce<false><<<grid, kScanNT, 0, h->stream>>>(in, n, n_ptr, h->tiles.p, nullptr);
  k_scan_top<<<1, kScanNT, 0, h->stream>>>(h->tiles.p, n, n_ptr, h->d_total + slot, tile_E,
                                          tile_E ? E_total : nullptr);
  if (nonzero)
    k_scan_down<true><<<grid, kScanNT, 0, h->stream>>>(in, n, n_ptr, h->tiles.p, out);
  else
    k_scan_down<false><<<grid, kScanNT, 0, h->stream>>>(in, n, n_ptr, h->tiles.p, out);
  HIPCHK(h, hipGetLastError());
  return BH_OK;
}

int read_total(bh_handle* h, int slot, uint64_t* v) {
  uint64_t* hp = static_cast<uint64_t*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp, h->d_total + slot, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  *v = *hp;
  return BH_OK;
}

// A lattice: cells numbered x fastest on a tensor-product grid whose node
// coordinates agree bit for bit with three axis arrays, ELEMENT faces whose
// neighbour is the lattice neighbour, no PROCESSOR face, and opacities equal
// in every group of a cell.  Every mesh built from a Branson deck is one.
bool detect_lattice(const bh_cells* c, uint32_t dims[3], std::vector<double>& axes,
                    uint32_t base = 0) {
  const uint32_t n = c->n_cells, G = c->n_groups;
  const double* nd = c->nodes;
  auto at = [&](uint32_t cell, int k) { return nd[(size_t)cell * 6 + k]; };
  uint32_t nx = 1;
  while (nx < n && at(nx, 2) == at(0, 2) && at(nx, 4) == at(0, 4) && at(nx, 0) == at(nx - 1, 1))
    ++nx;
  uint32_t ny = 1;
  while ((uint64_t)(ny + 1) * nx <= n && at(ny * nx, 4) == at(0, 4) &&
         at(ny * nx, 2) == at((ny - 1) * nx, 3))
    ++ny;
  const uint64_t nxy = (uint64_t)nx * ny;
  if (n % nxy) return false;
  const uint32_t nz = (uint32_t)(n / nxy);
  axes.assign((size_t)nx + ny + nz + 3, 0.0);
  double* X = axes.data();
  double* Y = X + nx + 1;
  double* Z = Y + ny + 1;
  for (uint32_t i = 0; i < nx; ++i) X[i] = at(i, 0);
  X[nx] = at(nx - 1, 1);
  for (uint32_t j = 0; j < ny; ++j) Y[j] = at(j * nx, 2);
  Y[ny] = at((ny - 1) * nx, 3);
  for (uint32_t k = 0; k < nz; ++k) Z[k] = at((uint32_t)(k * nxy), 4);
  Z[nz] = at((uint32_t)((nz - 1) * nxy), 5);
  const uint64_t stride[3] = {1, nx, nxy};
  const uint32_t dim[3] = {nx, ny, nz};
  for (uint32_t k = 0, cell = 0; k < nz; ++k)
    for (uint32_t j = 0; j < ny; ++j)
      for (uint32_t i = 0; i < nx; ++i, ++cell) {
        const double want[6] = {X[i], X[i + 1], Y[j], Y[j + 1], Z[k], Z[k + 1]};
        const uint32_t idx[3] = {i, j, k};
        for (int f = 0; f < 6; ++f) {
          if (std::memcmp(&want[f], &nd[(size_t)cell * 6 + f], sizeof(double)) != 0) return false;
          const int32_t bc = c->bc[(size_t)cell * 6 + f];
          if (bc == BH_PROCESSOR) return false;
          if (bc == BH_ELEMENT) {
            const int a = f >> 1;
            const bool hi = f & 1;
            if (hi ? idx[a] + 1 >= dim[a] : idx[a] == 0) return false;
            const uint64_t nb = hi ? cell + stride[a] : cell - stride[a];
            if (c->next[(size_t)cell * 6 + f] != nb + base) return false;
          }
        }
        for (uint32_t g = 1; g < G; ++g)
          if (c->sigma_a[(size_t)cell * G + g] != c->sigma_a[(size_t)cell * G] ||
              c->sigma_s[(size_t)cell * G + g] != c->sigma_s[(size_t)cell * G])
            return false;
      }
  dims[0] = nx;
  dims[1] = ny;
  dims[2] = nz;
  return true;
}

// Many problems whose meshes are each a lattice (cells of instance k are
// [first[k], first[k+1]), numbered x fastest within it, ELEMENT faces only
// inside it): the 32-byte records with each cell's instance in CellRec::pad,
// every instance's axes back to back, and the per-instance shapes.  Leaves
// h->inst_lattice false (tracking on the general records) when any
// instance is not a lattice.  c: the whole mesh in host memory.
int build_inst_lattice(bh_handle* h, const bh_cells* c) {
  h->inst_lattice = false;
  if (h->opt.general_mesh != 0 || h->lattice || h->n_inst == 0) return BH_OK;
  const uint32_t K = h->n_inst, G = c->n_groups;
  std::vector<double> axes;
  std::vector<LatInst> lat(K);
  std::vector<CellRec> rec(c->n_cells);
  for (uint32_t k = 0; k < K; ++k) {
    const uint32_t f = h->inst_first[k], e = h->inst_first[k + 1];
    bh_cells sub = *c;
    sub.n_cells = e - f;
    sub.nodes = c->nodes + (size_t)f * 6;
    sub.bc = c->bc + (size_t)f * 6;
    sub.next = c->next + (size_t)f * 6;
    sub.sigma_a = c->sigma_a + (size_t)f * G;
    sub.sigma_s = c->sigma_s + (size_t)f * G;
    sub.fleck = c->fleck + f;
    uint32_t dims[3];
    std::vector<double> ax;
    if (!detect_lattice(&sub, dims, ax, f)) return BH_OK;
    const uint32_t o = (uint32_t)axes.size();
    lat[k] = LatInst{f, dims[0], dims[1], dims[2], o, o + dims[0] + 1,
                     o + dims[0] + dims[1] + 2, 0u};
    axes.insert(axes.end(), ax.begin(), ax.end());
    for (uint32_t i = f; i < e; ++i) {
      uint32_t bits = 0;
      for (int q = 0; q < 6; ++q) bits |= (uint32_t)c->bc[(size_t)i * 6 + q] << (3 * q);
      rec[i] = CellRec{c->sigma_a[(size_t)i * G], c->sigma_s[(size_t)i * G], c->fleck[i], bits,
                       k};
    }
  }
  TRY(ensure(h, h->rec, c->n_cells));
  TRY(ensure(h, h->axes, axes.size()));
  TRY(ensure(h, h->lat, K));
  HIPCHK(h, hipMemcpyAsync(h->rec.p, rec.data(), rec.size() * sizeof(CellRec),
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->axes.p, axes.data(), axes.size() * sizeof(double),
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->lat.p, lat.data(), K * sizeof(LatInst), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->inst_lattice = true;
  return BH_OK;
}

int check_cells(bh_handle* h, const bh_cells* c) {
  if (!c || !c->nodes || !c->bc || !c->next || !c->sigma_a || !c->sigma_s || !c->fleck)
    return fail(h, BH_ERR_ARG, "bh_cells has a NULL array");
  if (c->n_cells == 0 || c->n_groups == 0)
    return fail(h, BH_ERR_ARG, "bh_cells needs n_cells > 0 and n_groups > 0");
  for (size_t i = 0; i < (size_t)c->n_cells * 6; ++i) {
    if (c->bc[i] < BH_REFLECT || c->bc[i] > BH_PROCESSOR)
      return fail(h, BH_ERR_ARG, "cell %zu face %zu: bc %d out of range", i / 6, i % 6,
                  c->bc[i]);
    if (c->bc[i] == BH_ELEMENT && c->next[i] >= c->n_cells)
      return fail(h, BH_ERR_ARG, "cell %zu face %zu: neighbour %u out of range", i / 6, i % 6,
                  c->next[i]);
  }
  return BH_OK;
}

// The wave hash table pays where flushes concentrate on few cells (many
// photons per cell of the source distribution: a face source, a hot spot),
// and costs ~8% where they spread (uniform emission).  Auto mode decides by
// the effective number of source cells (sum e)^2 / sum e^2 of the per-cell
// source energies, as a fraction of the mesh.
constexpr double kHashTallyNeff = 0.25;

bool use_hash_tally(const bh_handle* h) {
  if (h->opt.tally_mode == 4) return true;
  if (h->opt.tally_mode != 0 || h->n_cells >= 0xFFFFFFFFu) return false;
  return h->src_neff_frac < kHashTallyNeff;
}

// Runs `launch` with the tally sink chosen for this mesh: block-private LDS
// tallies when both arrays fit, else global atomics -- one per flush when
// every cell's absorption opacity is group-independent (then
// k_tally_finalize completes the tallies; combined per wave in an LDS hash
// table where the source is concentrated), two otherwise.
// Where a transport launch puts its results, and the RNG counter-high word
// of its photons.  handle_sinks(): the handle's own tallies, counters, queue
// and replay lists.  additive: the tally arrays already hold values that
// the launch must add to, so no tally form that rewrites them afterwards
// (k_tally_finalize) may be chosen.
struct Sinks {
  double* abs_E;
  double* track_E;
  DevCounters* ctrs;
  PersistentQueue* queue;
  ReplayList rl;  // FP32 form; null members: the handle's lists, sized at launch
  uint64_t ctr_hi;
  bool additive;
};

Sinks handle_sinks(const bh_handle* h, uint64_t ctr_hi) {
  return Sinks{h->abs_E.p, h->track_E.p, h->d_ctrs, h->d_queue,
               ReplayList{nullptr, nullptr, nullptr}, ctr_hi, false};
}

template <class Launch>
int with_tally(bh_handle* h, bool lds, const Sinks& s, Launch&& launch) {
  if (lds) {
    h->last_tally = 2;
    LdsTally t{s.abs_E, s.track_E, nullptr, nullptr, 0, 0};
    return launch(t, (size_t)2 * h->n_cells * sizeof(double));
  }
  if (h->gray_sigma_a && !s.additive && use_hash_tally(h)) {
    h->last_tally = 4;
    HashTally t{s.abs_E, s.track_E, nullptr, nullptr, 0};
    TRY(launch(t, (size_t)0));
    k_tally_finalize<<<grid_for(h->n_cells, 256), 256, 0, h->stream>>>(
        h->geo.p, h->sigma.p, h->G, h->n_cells, s.abs_E, s.track_E);
    HIPCHK(h, hipGetLastError());
    return BH_OK;
  }
  if (h->gray_sigma_a && !s.additive) {
    h->last_tally = 1;
    GlobalTally<false> t{s.abs_E, s.track_E, nullptr, nullptr, 0, 0};
    TRY(launch(t, (size_t)0));
    k_tally_finalize<<<grid_for(h->n_cells, 256), 256, 0, h->stream>>>(
        h->geo.p, h->sigma.p, h->G, h->n_cells, s.abs_E, s.track_E);
    HIPCHK(h, hipGetLastError());
    return BH_OK;
  }
  h->last_tally = 3;
  GlobalTally<true> t{s.abs_E, s.track_E, nullptr, nullptr, 0, 0};
  return launch(t, (size_t)0);
}

bool use_lds_tally(const bh_handle* h, bool persistent) {
  const size_t bytes = (size_t)2 * h->n_cells * sizeof(double);
  if (h->opt.tally_mode == 1 || h->opt.tally_mode == 3 || h->opt.tally_mode == 4) return false;
  if (h->opt.tally_mode == 2) return bytes <= h->lds_budget;
  return bytes <= (persistent ? h->lds_budget : std::min<size_t>(h->lds_budget, 16 * 1024));
}

// FP64 replay of the histories the FP32 kernel stopped (list length on the
// device, in h->d_ctrs->replays); same tally sink as the kernel before it.
template <class T>
int launch_replay(bh_handle* h, const DevCells& cells, const PhotonSoA& W, const ReplayList& rl,
                  T t, size_t shm, const CensusOut& co, const StepParams* P, const Sinks& s,
                  const SourceDesc& sd) {
  // the replay's share of the transport time (not inside a captured graph)
  h->replay_timed = P == nullptr;
  if (h->replay_timed) HIPCHK(h, hipEventRecord(h->ev[3], h->stream));
  if (sd.on)
    k_transport_replay<T, true><<<(unsigned)h->n_cu * 4, 256, shm, h->stream>>>(
        W, co, sd, cells, rl, s.ctr_hi, t, s.ctrs, P);
  else
    k_transport_replay<T, false><<<(unsigned)h->n_cu * 4, 256, shm, h->stream>>>(
        W, co, sd, cells, rl, s.ctr_hi, t, s.ctrs, P);
  HIPCHK(h, hipGetLastError());
  return BH_OK;
}

// Event-split transport of W[0, n): flight and scatter passes over the
// compacted list of live photons until it is empty.  The host reads the live
// count every few pass pairs (the counts only fall, so the last one read
// bounds the grids of the pairs in between).
int launch_event(bh_handle* h, const PhotonSoA& W, uint64_t n, const CensusOut& co,
                 const Sinks& s, bool lds) {
  const uint64_t cap = std::max<uint64_t>({n, h->W.cap, h->C.cap, 1});
  if (cap >= (uint64_t)0xFFFFFFFFu)
    return fail(h, BH_ERR_CAPACITY, "the event-split lists index photons with 32 bits");
  TRY(ensure(h, h->ev_abs, cap));
  TRY(ensure(h, h->ev_track, cap));
  TRY(ensure(h, h->ev_segs, cap));
  TRY(ensure(h, h->ev_sc, cap));
  TRY(ensure(h, h->ev_list0, cap));
  TRY(ensure(h, h->ev_list1, cap));
  TRY(ensure(h, h->ev_scatter, cap));
  const EventState ev{h->ev_abs.p,   h->ev_track.p, h->ev_segs.p, h->ev_sc.p,
                      {h->ev_list0.p, h->ev_list1.p}, h->ev_scatter.p, h->d_ev_count};
  const DevCells cells = dev_cells(h);
  k_ev_init<<<grid_for(n, 256), 256, 0, h->stream>>>(ev, n);
  HIPCHK(h, hipGetLastError());
  const unsigned max_grid = (unsigned)h->n_cu * 16;
  uint64_t live = n;
  int cur = 0;
  unsigned long long* hp = reinterpret_cast<unsigned long long*>(
      static_cast<char*>(h->h_pinned) + 3584);
  return with_tally(h, lds, s, [&](auto t, size_t shm) -> int {
    using T = decltype(t);
    while (live > 0) {
      for (int k = 0; k < 8; ++k) {
        const unsigned grid = std::min<unsigned>(grid_for(live, 256), max_grid);
        k_ev_clear<<<1, 1, 0, h->stream>>>(ev, cur ^ 1);
        if (h->lattice)
          k_ev_flight<T, true><<<grid, 256, shm, h->stream>>>(W, co, cells, ev, cur, s.ctr_hi, t,
                                                               s.ctrs);
        else
          k_ev_flight<T, false><<<grid, 256, shm, h->stream>>>(W, co, cells, ev, cur, s.ctr_hi,
                                                                t, s.ctrs);
        if (h->lattice)
          k_ev_scatter<true><<<grid, 256, 0, h->stream>>>(W, cells, ev, cur, s.ctr_hi);
        else
          k_ev_scatter<false><<<grid, 256, 0, h->stream>>>(W, cells, ev, cur, s.ctr_hi);
        HIPCHK(h, hipGetLastError());
        cur ^= 1;
      }
      HIPCHK(h, hipMemcpyAsync(hp, &h->d_ev_count[cur], sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      live = *hp;
    }
    return BH_OK;
  });
}

// Default persistent grid: the kernel's resident blocks per CU, at most this.
constexpr int kMaxBlocksPerCu = 8;

// Transport of W[0, n) with the configured kernel; counters in h->d_ctrs.
// With P (a step being captured as a graph) the persistent kernel reads n on
// the device and n only sizes the grid and buffers.  co: census output
// (co.on = 0: photons stay in place in W).
// sd.on: the photons from the census count on are made inside the
// persistent kernel (and again in the FP64 replay); no other form makes them.
int launch_transport(bh_handle* h, const PhotonSoA& W, uint64_t n, const StepParams* P,
                     const CensusOut& co, const Sinks& s, const SourceDesc& sd = SourceDesc{}) {
  if (n == 0) return BH_OK;
  const DevCells cells = dev_cells(h);
  const int block = h->opt.block_size;
  const bool persistent = h->opt.algorithm == BH_ALGO_PERSISTENT;
  const bool lds = use_lds_tally(h, persistent);
  const bool fp32 = h->opt.precision == BH_PRECISION_FP32_TRACK;
  const bool lat = h->lattice;
  const uint64_t ctr_hi = s.ctr_hi;
  DevCounters* const ctrs = s.ctrs;
  ReplayList rl = s.rl;
  if (fp32 && !rl.idx) {
    // room for every photon of the launch, sized with the photon stores so
    // that a growing census does not reallocate it every step
    const size_t cap = std::max<size_t>(n, std::max(h->W.cap, h->C.cap));
    TRY(ensure(h, h->rep_idx, cap));
    TRY(ensure(h, h->rep_first, cap));
    TRY(ensure(h, h->rep_E, cap));
    rl = ReplayList{h->rep_idx.p, h->rep_first.p, h->rep_E.p};
  }
  if (sd.on && !persistent)
    return fail(h, BH_ERR_STATE, "only the persistent kernel makes new photons itself");
  h->last_lattice = lat ? 1 : 0;
  if (h->opt.algorithm == BH_ALGO_EVENT) {
    if (P) return fail(h, BH_ERR_STATE, "the event-split algorithm is not captured as a graph");
    return launch_event(h, W, n, co, s, lds);
  }
  if (!persistent) {
    return with_tally(h, lds, s, [&](auto t, size_t shm) -> int {
      using T = decltype(t);
      if (fp32) {
        if (lat)
          k_transport_history<T, float, true><<<grid_for(n, block), block, shm, h->stream>>>(
              W, co, cells, n, ctr_hi, t, ctrs, rl);
        else
          k_transport_history<T, float, false><<<grid_for(n, block), block, shm, h->stream>>>(
              W, co, cells, n, ctr_hi, t, ctrs, rl);
        HIPCHK(h, hipGetLastError());
        return launch_replay(h, cells, W, rl, t, shm, co, P, s, sd);
      }
      if (lat)
        k_transport_history<T, double, true><<<grid_for(n, block), block, shm, h->stream>>>(
            W, co, cells, n, ctr_hi, t, ctrs, rl);
      else
        k_transport_history<T, double, false><<<grid_for(n, block), block, shm, h->stream>>>(
            W, co, cells, n, ctr_hi, t, ctrs, rl);
      HIPCHK(h, hipGetLastError());
      return BH_OK;
    });
  }
  const int refill = h->opt.refill_threshold > 0 ? std::min(h->opt.refill_threshold, kWave) : 16;
  const uint32_t chunk = h->opt.batch_photons > 0 ? (uint32_t)h->opt.batch_photons : 256u;
  // By default the grid is the blocks that stay resident together: a
  // persistent block beyond them starts only after the queue has drained,
  // and exits at once (8 per CU measured 1-3% slower than the 3 an FP64
  // kernel keeps resident).
  auto grid_of = [&](const void* kernel, size_t shm) -> unsigned {
    int bpc = h->opt.blocks_per_cu;
    if (bpc <= 0) {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, shm) != hipSuccess ||
          nb <= 0)
        nb = kMaxBlocksPerCu;
      bpc = std::min(nb, kMaxBlocksPerCu);
    }
    return (unsigned)std::min<uint64_t>((uint64_t)h->n_cu * bpc,
                                        std::max<uint64_t>(div_up(n, block), 1));
  };
  k_queue_init<<<1, 64, 0, h->stream>>>(s.queue);
  return with_tally(h, lds, s, [&](auto t, size_t shm) -> int {
    using T = decltype(t);
    // one instantiation per (precision, records, in-kernel photon making)
    auto go = [&](auto* k) {
      k<<<grid_of(reinterpret_cast<const void*>(k), shm), block, shm, h->stream>>>(
          W, co, sd, cells, n, ctr_hi, t, ctrs, s.queue, refill, chunk, rl, P);
    };
    const bool gen = sd.on != 0;
    if (fp32) {
      if (lat && gen) go(&k_transport_persistent<T, float, true, true>);
      else if (lat) go(&k_transport_persistent<T, float, true, false>);
      else if (gen) go(&k_transport_persistent<T, float, false, true>);
      else go(&k_transport_persistent<T, float, false, false>);
      HIPCHK(h, hipGetLastError());
      return launch_replay(h, cells, W, rl, t, shm, co, P, s, sd);
    }
    if (lat && gen) go(&k_transport_persistent<T, double, 1, true>);
    else if (lat) go(&k_transport_persistent<T, double, 1, false>);
    else if (gen) go(&k_transport_persistent<T, double, 0, true>);
    else if (cells.lat) {  // a batch of problems that are each a lattice
      h->last_lattice = 2;
      go(&k_transport_persistent<T, double, 2, false>);
    } else {
      go(&k_transport_persistent<T, double, 0, false>);
    }
    HIPCHK(h, hipGetLastError());
    return BH_OK;
  });
}

// Arrays to clear before a launch, cleared together by one k_zero_segs.
struct ZeroList {
  ZeroSegs z{};
  uint64_t max_words = 0;
};

int zero_launch(bh_handle* h, ZeroList& zl) {
  if (zl.z.count) {
    const unsigned gx =
        (unsigned)std::min<uint64_t>(std::max<uint64_t>(div_up(zl.max_words, 256), 1), 2048);
    k_zero_segs<<<dim3(gx, zl.z.count), 256, 0, h->stream>>>(zl.z);
    HIPCHK(h, hipGetLastError());
  }
  zl = ZeroList{};
  return BH_OK;
}

int zero_add(bh_handle* h, ZeroList& zl, void* p, size_t bytes) {
  if (zl.z.count == kZeroSegs) TRY(zero_launch(h, zl));
  const uint64_t w = bytes / 8;  // every array cleared here is 8-byte words
  zl.z.p[zl.z.count] = static_cast<unsigned long long*>(p);
  zl.z.n[zl.z.count] = w;
  ++zl.z.count;
  zl.max_words = std::max(zl.max_words, w);
  return BH_OK;
}

// The tallies and counters of a launch cleared: now, or (pending) with the
// census buffers by census_prepare.
int zero_tallies_and_counters(bh_handle* h, ZeroList* pending = nullptr) {
  ZeroList own;
  ZeroList& zl = pending ? *pending : own;
  TRY(zero_add(h, zl, h->abs_E.p, (size_t)h->n_cells * sizeof(double)));
  TRY(zero_add(h, zl, h->track_E.p, (size_t)h->n_cells * sizeof(double)));
  static_assert(sizeof(DevCounters) % 8 == 0, "counters are 64-bit words");
  TRY(zero_add(h, zl, h->d_ctrs, sizeof(DevCounters)));
  return pending ? BH_OK : zero_launch(h, zl);
}

// Capacities of a census-writing launch over at most n photons.  Every
// wave that writes census photons leaves at most kCenChunk - 1 unused slots
// (its last chunk), and holds one hole record.  A persistent or replay grid
// has a fixed number of waves whatever n; a history launch has one wave per
// 64 photons, each needing at most one chunk.
struct CensusSizes {
  uint64_t slots, rec_cap, hole_cap;
};
CensusSizes census_sizes(const bh_handle* h, uint64_t n) {
  const uint64_t block = (uint64_t)h->opt.block_size;
  const bool fp32 = h->opt.precision == BH_PRECISION_FP32_TRACK;
  const uint64_t w_rep = fp32 ? (uint64_t)h->n_cu * 4 * 256 / kWave : 0;
  CensusSizes z;
  if (h->opt.algorithm == BH_ALGO_EVENT) {  // dense census writes: no holes
    z.rec_cap = 1;
    z.slots = std::max<uint64_t>(n, 1);
    z.hole_cap = kCenChunk;
    return z;
  }
  if (h->opt.algorithm == BH_ALGO_PERSISTENT) {
    // the default grid is the resident blocks, at most kMaxBlocksPerCu
    const uint64_t bpc = h->opt.blocks_per_cu > 0 ? (uint64_t)h->opt.blocks_per_cu
                                                  : (uint64_t)kMaxBlocksPerCu;
    z.rec_cap = (uint64_t)h->n_cu * bpc * block / kWave + w_rep;
    z.slots = n + kCenChunk * z.rec_cap;
  } else {
    const uint64_t w_main = div_up(std::max<uint64_t>(n, 1), block) * block / kWave;
    z.rec_cap = w_main + w_rep;
    z.slots = kCenChunk * w_main + (fp32 ? n + kCenChunk * w_rep : 0);
  }
  z.hole_cap = kCenChunk * z.rec_cap;
  return z;
}

// The census marks of launches of up to pc photons: one byte per photon
// and a whole word (64) of zero marks past the last.
inline uint64_t mark_words(uint64_t pc) { return div_up(std::max<uint64_t>(pc, 1), 64) + 1; }

int ensure_marks(bh_handle* h, uint64_t pc) {
  const uint64_t mw = mark_words(pc);
  TRY(ensure(h, h->cmark, mw * 8));
  TRY(ensure(h, h->mwords, mw));
  TRY(ensure(h, h->mcounts, mw));
  TRY(ensure(h, h->off, mw));
  TRY(ensure(h, h->erank, std::max<uint64_t>(pc, 1)));
  return BH_OK;
}

// Buffers and the output record of a census-writing launch over the n
// photons of W = [census | new]: C has room for all of them and the holes
// the waves leave, the marks by transport index are cleared, the totals
// zeroed.  Fill kernels, no memset, so the same code records into a graph
// (P: sizes read on the device; cap bounds n).
int census_prepare(bh_handle* h, uint64_t cap, double next_dt, CensusOut& co,
                   ZeroList* pending = nullptr) {
  // census keys are 32-bit transport indices and ranks, 0xFFFFFFFF marks a hole
  if (cap >= (uint64_t)kHole)
    return fail(h, BH_ERR_CAPACITY, "%llu photons in one launch: at most %u are supported",
                (unsigned long long)cap, kHole - 1u);
  const CensusSizes z = census_sizes(h, cap);
  TRY(grow_store(h, h->C, z.slots, 0));
  const uint64_t pc = photon_cap(h, cap);
  TRY(ensure_marks(h, pc));
  TRY(ensure(h, h->tiles, std::max<uint64_t>(div_up(pc, kScanTile), 1)));
  TRY(ensure(h, h->tile_E, std::max<uint64_t>(div_up(pc, kScanTile), 1)));
  TRY(ensure(h, h->sum_part, (size_t)8 * kSumMaxBlocks));
  TRY(ensure(h, h->hrec, z.rec_cap));
  const uint64_t mw = mark_words(cap);  // the marks of this launch and a zero word
  ZeroList own;
  ZeroList& zl = pending ? *pending : own;
  TRY(zero_add(h, zl, h->cmark.p, (size_t)mw * 64));
  static_assert(sizeof(CensusAcc) % 8 == 0, "census totals are 64-bit words");
  TRY(zero_add(h, zl, h->d_cacc, sizeof(CensusAcc)));
  TRY(zero_launch(h, zl));
  co = CensusOut{h->C.soa, reinterpret_cast<uint8_t*>(h->cmark.p), h->d_cacc, h->hrec.p,
                 kC * next_dt, h->n_cen, h->n_new, h->C.cap, z.rec_cap, 1};
  return BH_OK;
}

// After the launch: the holes below n_census filled from the photons above
// it, each census photon's rank in the reference's order (the count of marks
// below its transport index), the census energy (summed by rank in a fixed
// split) and the totals, into h->d_ctot.
int census_finish(bh_handle* h, uint64_t cap, const uint64_t* n_ptr) {
  const CensusSizes z = census_sizes(h, cap);
  TRY(ensure(h, h->hole_len, z.rec_cap));
  TRY(ensure(h, h->hole_off, z.rec_cap));
  TRY(ensure(h, h->tail_live, z.hole_cap));
  TRY(ensure(h, h->tail_rank, z.hole_cap));
  CensusAcc* acc = h->d_cacc;
  const unsigned g_rec = std::min<unsigned>(grid_for(z.rec_cap, 256), 1024);
  const unsigned g_hole = std::min<unsigned>(grid_for(z.hole_cap, 256), 2048);
  k_hole_heads<<<g_rec, 256, 0, h->stream>>>(h->hrec.p, acc, z.rec_cap, h->hole_len.p);
  TRY(scan_u64(h, h->hole_len.p, z.rec_cap, h->hole_off.p, 3,
               reinterpret_cast<const uint64_t*>(&acc->n_rec)));
  k_hole_tail<<<g_hole, 256, 0, h->stream>>>(h->C.soa.key, acc, z.hole_cap, h->tail_live.p);
  TRY(scan_u64(h, h->tail_live.p, z.hole_cap, h->tail_rank.p, 3,
               reinterpret_cast<const uint64_t*>(&acc->n_holes)));
  k_hole_fill<<<g_hole, 256, 0, h->stream>>>(h->C.soa, acc, z.hole_cap, h->tail_live.p,
                                             h->tail_rank.p, h->hrec.p, h->hole_off.p, z.rec_cap);
  HIPCHK(h, hipGetLastError());
  const uint64_t n = std::max<uint64_t>(cap, 1);
  const uint64_t nw = div_up(n, 64);
  TRY(ensure_marks(h, n));
  // ranks: the marks 64 to a word, each word's first rank by a scan of
  // their counts, then each census photon's rank from its word; the census
  // energy summed in census order (by rank, a fixed split of the count)
  k_mark_words<<<std::min<unsigned>(grid_for(nw, 256), 4096), 256, 0, h->stream>>>(
      reinterpret_cast<const uint8_t*>(h->cmark.p), n, n_ptr, h->mwords.p, h->mcounts.p, acc);
  TRY(scan_u64(h, h->mcounts.p, nw, h->off.p, 2,
               reinterpret_cast<const uint64_t*>(&acc->n_words)));
  const unsigned long long* cnt = &acc->n_census;
  k_census_rekey<<<std::min<unsigned>(grid_for(n, 256), 8192), 256, 0, h->stream>>>(
      h->C.soa.key, h->C.soa.E, h->mwords.p, h->off.p, cnt, n, h->erank.p);
  k_rank_energy_partials<<<kSumMaxBlocks, kSumNT, 0, h->stream>>>(h->erank.p, cnt,
                                                                  h->sum_part.p);
  k_final_sums<1><<<1, kSumNT, 0, h->stream>>>(h->sum_part.p, kSumMaxBlocks,
                                              &h->d_ctot->census_E);
  k_census_pack<<<1, 1, 0, h->stream>>>(acc, h->d_ctot);
  HIPCHK(h, hipGetLastError());
  return BH_OK;
}

// Writes the step's new photons to W behind the census (src_deferred: they
// were left to be made inside the transport kernel).  For everything that
// reads them from W: the history and event-split forms and the exports.
// Whether the step's new photons are made inside the persistent transport
// kernel (bh_options.source_mode 1).  Auto keeps the source kernel: making
// them in the transport kernel saves their 97-byte HBM round trip but costs
// that kernel 30-34 spilled VGPRs at its three waves per SIMD, and measured
// slower on every workload (hohlraum 8.05 against 6.72 ms per step, cube
// 10.2 against 9.7 ms).
bool source_in_transport(const bh_handle* h, uint64_t) {
  return h->opt.algorithm == BH_ALGO_PERSISTENT && h->opt.source_mode == 1;
}

int materialize_source(bh_handle* h, uint64_t n_new) {
  if (!h->src_deferred) return BH_OK;
  if (n_new == 0) {
    h->src_deferred = false;
    return BH_OK;
  }
  TRY(grow_store(h, h->W, h->n_cen + n_new, h->n_cen));
  const SourceDesc& d = h->sd;
  k_source_fill<<<grid_for(n_new, 256), 256, 0, h->stream>>>(
      dev_cells(h), d.sc, d.face, d.off, d.hint, n_new, d.one, d.inst, d.dt, 0, h->W.soa,
      h->n_cen, nullptr, nullptr);
  HIPCHK(h, hipGetLastError());
  h->src_deferred = false;
  return BH_OK;
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

namespace {
// The photon stores for launches of up to `n` photons, and every buffer a
// step sizes from the stores' capacity (census marks and ranks, scan tiles,
// hole records, source hints, the FP32 replay list, the event-split lists),
// allocated together at creation: steps within the reservation then
// allocate nothing.  Each hipMalloc / hipFree synchronises the device, and
// multi-GB ones inside a step cost far more than its kernels.
int reserve_buffers(bh_handle* h, uint64_t n) {
  TRY(grow_store(h, h->W, n, 0));
  TRY(grow_store(h, h->C, census_sizes(h, n).slots, 0));
  const uint64_t pc = photon_cap(h, n);
  const CensusSizes z = census_sizes(h, pc);
  const uint64_t tiles = std::max<uint64_t>(div_up(pc, kScanTile), 1);
  TRY(ensure_marks(h, pc));
  TRY(ensure(h, h->tiles, tiles));
  TRY(ensure(h, h->tile_E, tiles));
  TRY(ensure(h, h->sum_part, (size_t)8 * kSumMaxBlocks));
  TRY(ensure(h, h->hrec, z.rec_cap));
  TRY(ensure(h, h->hole_len, z.rec_cap));
  TRY(ensure(h, h->hole_off, z.rec_cap));
  TRY(ensure(h, h->tail_live, z.hole_cap));
  TRY(ensure(h, h->tail_rank, z.hole_cap));
  TRY(ensure(h, h->src_hint, div_up(pc, kHintStride) + 1));
  if (h->opt.precision == BH_PRECISION_FP32_TRACK) {
    TRY(ensure(h, h->rep_idx, pc));
    TRY(ensure(h, h->rep_first, pc));
    TRY(ensure(h, h->rep_E, pc));
  }
  if (h->opt.algorithm == BH_ALGO_EVENT) {
    TRY(ensure(h, h->ev_abs, pc));
    TRY(ensure(h, h->ev_track, pc));
    TRY(ensure(h, h->ev_segs, pc));
    TRY(ensure(h, h->ev_sc, pc));
    TRY(ensure(h, h->ev_list0, pc));
    TRY(ensure(h, h->ev_list1, pc));
    TRY(ensure(h, h->ev_scatter, pc));
  }
  return BH_OK;
}
}  // namespace

int bh_abi_version(void) { return BH_ABI_VERSION; }

int bh_create(int device, const bh_options* opts, bh_handle** out) {
  if (!out) return BH_ERR_ARG;
  *out = nullptr;
  bh_handle* h = new (std::nothrow) bh_handle();
  if (!h) return BH_ERR_ALLOC;
  if (opts) h->opt = *opts;
  if (h->opt.block_size == 0) h->opt.block_size = 256;
  if (h->opt.block_size != 64 && h->opt.block_size != 128 && h->opt.block_size != 256) {
    int r = fail(h, BH_ERR_ARG, "block_size %d not supported (64, 128 or 256)",
                 h->opt.block_size);
    *out = h;
    return r;
  }
  if (h->opt.batch_photons < 0 || h->opt.batch_photons > (1 << 20)) {
    int r = fail(h, BH_ERR_ARG, "batch_photons %d out of range (0..2^20)",
                 h->opt.batch_photons);
    *out = h;
    return r;
  }
  if (h->opt.algorithm == BH_ALGO_EVENT && h->opt.precision != BH_PRECISION_FP64) {
    int r = fail(h, BH_ERR_ARG, "the event-split algorithm runs FP64 only");
    *out = h;
    return r;
  }
  if (h->opt.precision < 0 || h->opt.precision > 1 || h->opt.algorithm < 0 ||
      h->opt.algorithm > 2 || h->opt.tally_mode < 0 || h->opt.tally_mode > 4 ||
      !(h->opt.fp32_tol >= 0.0 && h->opt.fp32_tol < 1.0) || h->opt.general_mesh < 0 ||
      h->opt.general_mesh > 1 || h->opt.source_mode < 0 || h->opt.source_mode > 2) {
    int r = fail(h, BH_ERR_ARG, "invalid bh_options");
    *out = h;
    return r;
  }
  h->device = device;
  *out = h;
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0)
    return fail(h, BH_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= n_dev)
    return fail(h, BH_ERR_ARG, "device %d out of range (%d devices)", device, n_dev);
  HIPCHK(h, hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(h, hipGetDeviceProperties(&prop, device));
  h->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  // leave room for the kernels' static LDS (block scans, counters)
  if (prop.sharedMemPerBlock > 4096)
    h->lds_budget = std::min<size_t>(48 * 1024, prop.sharedMemPerBlock - 4096);
  HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  for (auto& e : h->ev) HIPCHK(h, hipEventCreate(&e));
  HIPCHK(h, hipMalloc(&h->d_total, 4 * sizeof(uint64_t)));
  HIPCHK(h, hipMalloc(&h->d_ctot, sizeof(CensusPartial)));
  HIPCHK(h, hipMalloc(&h->d_cacc, sizeof(CensusAcc)));
  HIPCHK(h, hipMalloc(&h->d_ctrs, sizeof(DevCounters)));
  HIPCHK(h, hipMalloc(&h->d_queue, sizeof(PersistentQueue)));
  HIPCHK(h, hipMalloc(&h->d_ev_count, 4 * sizeof(unsigned long long)));
  HIPCHK(h, hipMalloc(&h->d_params, sizeof(StepParams)));
  HIPCHK(h, hipMalloc(&h->d_phys, sizeof(PhysTotals)));
  HIPCHK(h, hipMalloc(&h->d_inst2, 2 * sizeof(SourceInst)));
  HIPCHK(h, hipHostMalloc(&h->h_pinned, 4096, hipHostMallocDefault));
  {  // the no-allocation entry point's device memory on this device
    const hipError_t e = scratch_on_device(device, &h->scratch);
    if (e != hipSuccess) {
      h->scratch = nullptr;
      h->scratch_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    }
  }
  if (h->opt.reserve_photons) TRY(reserve_buffers(h, h->opt.reserve_photons));
  h->n_allocs = 0;
  return BH_OK;
}

void bh_destroy(bh_handle* h) {
  if (!h) return;
  if (h->stream) {
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
  }
  release(h->geo);
  release(h->sigma);
  release(h->next);
  release(h->em);
  release(h->src);
  release(h->cen);
  release(h->face);
  release(h->abs_E);
  release(h->track_E);
  release(h->counts);
  release(h->src_cell);
  release(h->off);
  release(h->tiles);
  release(h->aos);
  release(h->rec);
  release(h->axes);
  release(h->cell_inst);
  release(h->inst_ctr_hi);
  release(h->inst_src);
  release(h->adj_g);
  release(h->grad_sa);
  release(h->grad_f);
  release(h->rep_idx);
  release(h->rep_first);
  release(h->rep_E);
  release(h->mat);
  release(h->mat_e);
  release(h->tr_diag);
  release(h->sum_part);
  release(h->comb_keys);
  release(h->comb_idx);
  release(h->comb_sorted);
  release(h->comb_perm);
  release(h->comb_rank_cell);
  release(h->comb_keep);
  release(h->comb_tmp);
  release(h->comb_cellE);
  release(h->comb_ctr);
  release(h->cflag);
  release(h->lat);
  release(h->cmark);
  release(h->mwords);
  release(h->mcounts);
  release(h->erank);
  release(h->tile_E);
  release(h->hrec);
  release(h->hole_len);
  release(h->hole_off);
  release(h->tail_live);
  release(h->tail_rank);
  void* ptrs[] = {h->W.slab, h->C.slab, h->d_total, h->d_ctot, h->d_ctrs, h->d_queue,
                  h->d_phys, h->d_inst2, h->d_cacc};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& g : h->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  for (auto& g : h->resident_graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  release(h->soff);
  release(h->src_hint);
  release(h->ev_abs);
  release(h->ev_track);
  release(h->ev_segs);
  release(h->ev_list0);
  release(h->ev_list1);
  release(h->ev_scatter);
  release(h->ev_sc);
  if (h->d_ev_count) (void)hipFree(h->d_ev_count);
  release(h->inst_first_dev);
  release(h->inst_rep);
  release(h->inst_exit);
  if (h->h_rec) (void)hipHostFree(h->h_rec);
  if (h->h_inst_rep) (void)hipHostFree(h->h_inst_rep);
  if (h->d_params) (void)hipFree(h->d_params);
  if (h->h_params) (void)hipHostFree(h->h_params);
  if (h->h_tally) (void)hipHostFree(h->h_tally);
  if (h->h_pinned) (void)hipHostFree(h->h_pinned);
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

const char* bh_last_error(const bh_handle* h) {
  if (!h) return "null handle";
  return h->err.c_str();
}

int bh_sync(bh_handle* h) {
  if (!h || !h->stream) return BH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return BH_OK;
}

int bh_set_mesh(bh_handle* h, const bh_cells* c) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream) return BH_ERR_ARG;
  TRY(check_cells(h, c));
  if (h->has_mesh) TRY(materialize_source(h, h->n_new));  // they read the old cells
  const uint32_t n = c->n_cells, G = c->n_groups;
  std::vector<CellGeo> geo(n);
  for (uint32_t i = 0; i < n; ++i) {
    CellGeo& g = geo[i];
    std::memcpy(g.nodes, c->nodes + (size_t)i * 6, 6 * sizeof(double));
    g.fleck = c->fleck[i];
    uint32_t bits = 0;
    for (int f = 0; f < 6; ++f) bits |= (uint32_t)c->bc[(size_t)i * 6 + f] << (3 * f);
    g.bc_bits = bits;
    g.source_face_plus1 = 0;
    for (int f = 0; f < 6; ++f)
      if (c->bc[(size_t)i * 6 + f] == BH_SOURCE) {
        g.source_face_plus1 = (uint32_t)f + 1;
        break;
      }
  }
  std::vector<double2> sig((size_t)n * G);
  for (size_t k = 0; k < (size_t)n * G; ++k) sig[k] = make_double2(c->sigma_a[k], c->sigma_s[k]);
  HIPCHK(h, hipSetDevice(h->device));
  TRY(ensure(h, h->geo, n));
  TRY(ensure(h, h->next, (size_t)n * 6));
  TRY(ensure(h, h->sigma, (size_t)n * G));
  TRY(ensure(h, h->abs_E, n));
  TRY(ensure(h, h->track_E, n));
  TRY(ensure(h, h->counts, n));  // the source's per-cell counts and offsets
  TRY(ensure(h, h->src_cell, n));
  TRY(ensure(h, h->soff, n));
  TRY(ensure(h, h->tiles, std::max<uint64_t>(div_up(n, kScanTile), 1)));
  HIPCHK(h, hipMemcpyAsync(h->geo.p, geo.data(), n * sizeof(CellGeo), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->next.p, c->next, (size_t)n * 6 * sizeof(uint32_t),
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->sigma.p, sig.data(), sig.size() * sizeof(double2),
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->has_mesh && (n != h->n_cells || G != h->G)) {
    h->has_energy = false;  // energies were sized for the old mesh
    h->n_cen = 0;           // and a resident census indexes the old cells
    h->n_new = 0;
    h->n_inst = 0;          // as did the instance table
  }
  h->n_cells = n;
  h->G = G;
  h->has_mesh = true;
  std::vector<double> axes;
  uint32_t dims[3];
  h->lattice = false;
  if (h->opt.general_mesh == 0 && detect_lattice(c, dims, axes)) {
    std::vector<CellRec> rec(n);
    for (uint32_t i = 0; i < n; ++i)
      rec[i] = CellRec{c->sigma_a[(size_t)i * G], c->sigma_s[(size_t)i * G], c->fleck[i],
                       geo[i].bc_bits, 0u};
    TRY(ensure(h, h->rec, n));
    TRY(ensure(h, h->axes, axes.size()));
    HIPCHK(h, hipMemcpyAsync(h->rec.p, rec.data(), n * sizeof(CellRec), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(h->axes.p, axes.data(), axes.size() * sizeof(double),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int a = 0; a < 3; ++a) h->dims[a] = dims[a];
    h->lattice = true;
  }
  h->uniform_sigma_a = true;
  for (uint32_t i = 0; i < n && h->uniform_sigma_a; ++i)
    for (uint32_t g = 1; g < G; ++g)
      if (c->sigma_a[(size_t)i * G + g] != c->sigma_a[(size_t)i * G]) {
        h->uniform_sigma_a = false;
        break;
      }
  // tally_mode 3 forces two tallies per flush
  h->gray_sigma_a = h->uniform_sigma_a && h->opt.tally_mode != 3;
  // a batch of problems kept across a new mesh of the same size
  TRY(build_inst_lattice(h, c));
  return BH_OK;
}

int bh_set_source_energy(bh_handle* h, const double* emission_E, const double* source_E,
                         const double* census_E, const int8_t* source_face) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "bh_set_source_energy before bh_set_mesh");
  // photons of an earlier bh_source still to be made read these energies
  TRY(materialize_source(h, h->n_new));
  if (!emission_E || !source_E || !source_face)
    return fail(h, BH_ERR_ARG, "emission_E, source_E and source_face are required");
  const uint32_t n = h->n_cells;
  for (uint32_t c = 0; c < n; ++c)
    if (source_E[c] > 0.0 && (source_face[c] < 0 || source_face[c] > 5))
      return fail(h, BH_ERR_ARG, "cell %u has source energy but no source face", c);
  HIPCHK(h, hipSetDevice(h->device));
  TRY(ensure(h, h->em, n));
  TRY(ensure(h, h->src, n));
  TRY(ensure(h, h->cen, n));
  TRY(ensure(h, h->face, n));
  HIPCHK(h, hipMemcpyAsync(h->em.p, emission_E, n * sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->src.p, source_E, n * sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
  if (census_E)
    HIPCHK(h, hipMemcpyAsync(h->cen.p, census_E, n * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
  HIPCHK(h, hipMemcpyAsync(h->face.p, source_face, n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->em_host.assign(emission_E, emission_E + n);
  h->src_host.assign(source_E, source_E + n);
  {  // photons are drawn in proportion to energy: (sum e)^2 / sum e^2 cells
    double s1 = 0.0, s2 = 0.0;
    for (uint32_t c = 0; c < n; ++c) {
      const double e = emission_E[c] + source_E[c] + (census_E ? census_E[c] : 0.0);
      s1 += e;
      s2 += e * e;
    }
    h->src_neff_frac = s2 > 0.0 ? (s1 * s1 / s2) / (double)n : 1.0;
  }
  h->has_energy = true;
  h->has_census_E = census_E != nullptr;
  return BH_OK;
}

// The step's source for one problem (inst_src null: `census` / `fresh`
// describe it) or for every instance (inst_src[0, n_inst) census,
// [n_inst, 2 n_inst) new photons, on the device).
// icen / inew: source parameter records on the device (per instance, or one
// record whose E_tot the device computed), else `census` / `fresh` are used.
// census_part: step 1 of a resident step, block partials of the initial
// census energy go there (census_nb blocks).
static int run_source(bh_handle* h, const SourceInst& census, const SourceInst& fresh,
                      const SourceInst* icen, const SourceInst* inew, double dt,
                      uint64_t* n_new_out, double* census_part = nullptr,
                      uint32_t census_nb = 0) {
  const uint32_t n = h->n_cells;
  const DevCells cells = dev_cells(h);
  TRY(ensure(h, h->counts, n));
  TRY(ensure(h, h->src_cell, n));
  TRY(ensure(h, h->soff, n));
  const uint32_t* cinst = h->n_inst ? h->cell_inst.p : nullptr;
  HIPCHK(h, hipEventRecord(h->ev[6], h->stream));
  if (h->has_census_E) {  // step 1: initial census replaces the resident one
    k_source_counts<<<grid_for(n, 256), 256, 0, h->stream>>>(h->cen.p, nullptr, n, census, icen,
                                                              cinst, 1, h->counts.p, h->src_cell.p);
    if (census_part) {  // initial census energy, block partials
      k_census_partials<<<census_nb, kSumNT, 0, h->stream>>>(h->cen.p, h->counts.p, n,
                                                             census_part);
      HIPCHK(h, hipGetLastError());
    }
    TRY(scan_u64(h, h->counts.p, n, h->soff.p, 0));
    uint64_t n_c = 0;
    TRY(read_total(h, 0, &n_c));
    TRY(grow_store(h, h->W, n_c, 0));
    if (n_c)
      k_source_fill<<<grid_for(n_c, 256), 256, 0, h->stream>>>(
          cells, h->src_cell.p, h->face.p, h->soff.p, nullptr, n_c, census, icen, dt, 1,
          h->W.soa, 0, nullptr, nullptr);
    HIPCHK(h, hipGetLastError());
    h->n_cen = n_c;
    h->has_census_E = false;
  }
  k_source_counts<<<grid_for(n, 256), 256, 0, h->stream>>>(h->em.p, h->src.p, n, fresh, inew,
                                                            cinst, 0, h->counts.p, h->src_cell.p);
  TRY(scan_u64(h, h->counts.p, n, h->soff.p, 1));
  // The new photons: made inside the persistent transport kernel, or written
  // to W behind the census here.  Written here, the hints and the fill are
  // launched before the host reads the count: they read it on the device,
  // bounded by the room W has, so the GPU works while the host waits for the
  // count.  A count beyond that room grows W and writes them again.
  const bool write_now = !source_in_transport(h, 0);
  const uint64_t room = (write_now && h->W.cap > h->n_cen) ? h->W.cap - h->n_cen : 0;
  if (room) TRY(ensure(h, h->src_hint, div_up(room, kHintStride) + 1));
  uint64_t* hp = static_cast<uint64_t*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp, h->d_total + 1, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipEventRecord(h->ev[8], h->stream));  // the count is on the host
  if (room) {
    // one 256-photon tile per block measured 8% faster than a grid-stride
    // grid of 16 blocks per CU: the grid covers the last step's count, or
    // the photons asked for, and an eighth more (the kernel's grid-stride
    // loop takes any photons beyond)
    const uint64_t guess = h->last_n_new ? h->last_n_new : (inew ? room : fresh.n_user);
    const uint64_t est = guess + guess / 8 + 1;
    k_source_hints<<<std::min<unsigned>(grid_for(div_up(room, kHintStride), 256), 1024), 256, 0,
                     h->stream>>>(h->soff.p, n, room, h->d_total + 1, h->src_hint.p);
    k_source_fill<<<grid_for(std::min(est, room), 256), 256, 0, h->stream>>>(
        cells, h->src_cell.p, h->face.p, h->soff.p, h->src_hint.p, room, fresh, inew, dt, 0,
        h->W.soa, h->n_cen, nullptr, h->d_total + 1);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipEventSynchronize(h->ev[8]));
  const uint64_t n_new = *hp;
  const bool again = n_new > room;  // (or the photons are made in the transport kernel)
  if (again && room) HIPCHK(h, hipStreamSynchronize(h->stream));  // the short fill is done
  if (again) TRY(ensure(h, h->src_hint, div_up(n_new, kHintStride) + 1));
  h->sd = SourceDesc{h->src_cell.p, h->face.p, h->soff.p, h->src_hint.p, inew, fresh, dt,
                     n_new, 1, 0};
  h->src_deferred = n_new > 0 && (again || !write_now);
  if (again) {
    k_source_hints<<<std::min<unsigned>(grid_for(div_up(n_new, kHintStride), 256), 1024), 256, 0,
                     h->stream>>>(h->soff.p, n, n_new, nullptr, h->src_hint.p);
    HIPCHK(h, hipGetLastError());
    if (write_now) TRY(materialize_source(h, n_new));
  }
  // the phase's time is read once its end has passed (source_ms), so the
  // host goes on to queue the transport while the fill runs
  HIPCHK(h, hipEventRecord(h->ev[7], h->stream));
  h->src_ms_pending = true;
  h->n_new = n_new;
  h->last_n_new = n_new;
  if (n_new_out) *n_new_out = n_new;
  return BH_OK;
}

// Duration of the last run_source, from its events once the second has
// completed.
double source_ms(bh_handle* h) {
  if (h->src_ms_pending) {
    h->src_ms_pending = false;
    h->last_source_ms = hipEventSynchronize(h->ev[7]) == hipSuccess
                            ? elapsed_ms(h->ev[6], h->ev[7]) : 0.0;
  }
  return h->last_source_ms;
}

int bh_source(bh_handle* h, uint64_t n_user_photons, double total_E, uint32_t seed,
              uint32_t cycle, uint32_t rank, double dt, uint64_t* n_new_out) {
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh || !h->has_energy)
    return fail(h, BH_ERR_STATE, "bh_source needs bh_set_mesh and bh_set_source_energy");
  if (h->n_inst)
    return fail(h, BH_ERR_STATE, "the mesh holds %u instances: use bh_source_instances",
                h->n_inst);
  if (!(total_E > 0.0)) return fail(h, BH_ERR_ARG, "total_E must be > 0");
  HIPCHK(h, hipSetDevice(h->device));
  const uint64_t ctr_hi = (uint64_t)seed << 32;
  if (h->n_cen && ctr_hi != h->ctr_hi)
    return fail(h, BH_ERR_ARG, "seed differs from the resident census photons' seed");
  h->ctr_hi = ctr_hi;
  const uint64_t rank_off = n_user_photons * (uint64_t)rank;
  const SourceInst census{n_user_photons, total_E, rank_off, ctr_hi, 0, 0};
  const SourceInst fresh{n_user_photons, total_E, 10000000000000ull * (uint64_t)cycle + rank_off,
                         ctr_hi, 0, 0};
  return run_source(h, census, fresh, nullptr, nullptr, dt, n_new_out);
}

int bh_set_instances(bh_handle* h, uint32_t n_instances, const uint32_t* first_cell,
                     const uint32_t* seeds) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "bh_set_instances before bh_set_mesh");
  HIPCHK(h, hipSetDevice(h->device));
  if (n_instances == 0) {  // back to one problem
    h->n_inst = 0;
    h->inst_lattice = false;
    h->n_cen = 0;
    h->n_new = 0;
    return BH_OK;
  }
  if (!first_cell || !seeds) return BH_ERR_ARG;
  if (first_cell[0] != 0 || first_cell[n_instances] != h->n_cells)
    return fail(h, BH_ERR_ARG, "instance cell ranges must start at 0 and end at n_cells (%u)",
                h->n_cells);
  for (uint32_t k = 0; k < n_instances; ++k)
    if (first_cell[k + 1] <= first_cell[k])
      return fail(h, BH_ERR_ARG, "instance %u has no cells", k);
  h->inst_first.assign(first_cell, first_cell + n_instances + 1);
  h->inst_hi.resize(n_instances);
  for (uint32_t k = 0; k < n_instances; ++k) h->inst_hi[k] = (uint64_t)seeds[k] << 32;
  h->cell_inst_host.resize(h->n_cells);
  for (uint32_t k = 0; k < n_instances; ++k)
    for (uint32_t c = first_cell[k]; c < first_cell[k + 1]; ++c) h->cell_inst_host[c] = k;
  TRY(ensure(h, h->cell_inst, h->n_cells));
  TRY(ensure(h, h->inst_ctr_hi, n_instances));
  TRY(ensure(h, h->inst_src, 2 * (size_t)n_instances));
  TRY(ensure(h, h->inst_first_dev, (size_t)n_instances + 1));
  TRY(ensure(h, h->inst_rep, n_instances));
  TRY(ensure(h, h->inst_exit, n_instances));
  if (h->h_inst_cap < n_instances) {
    if (h->h_rec) HIPCHK(h, hipHostFree(h->h_rec));
    if (h->h_inst_rep) HIPCHK(h, hipHostFree(h->h_inst_rep));
    h->h_rec = nullptr;
    h->h_inst_rep = nullptr;
    h->h_inst_cap = 0;
    HIPCHK(h, hipHostMalloc(&h->h_rec, 2 * (size_t)n_instances * sizeof(SourceInst),
                            hipHostMallocDefault));
    HIPCHK(h, hipHostMalloc(&h->h_inst_rep, (size_t)n_instances * sizeof(InstReport),
                            hipHostMallocDefault));
    h->h_inst_cap = n_instances;
  }
  HIPCHK(h, hipMemcpyAsync(h->inst_first_dev.p, first_cell,
                           ((size_t)n_instances + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->cell_inst.p, h->cell_inst_host.data(),
                           h->n_cells * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->inst_ctr_hi.p, h->inst_hi.data(), n_instances * sizeof(uint64_t),
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->n_inst = n_instances;
  h->n_cen = 0;  // a resident census belongs to the previous arrangement
  h->n_new = 0;
  if (!h->lattice && h->opt.general_mesh == 0) {
    // the mesh as the device holds it, for the per-instance lattice test
    const uint32_t n = h->n_cells, G = h->G;
    std::vector<CellGeo> geo(n);
    std::vector<double2> sig((size_t)n * G);
    std::vector<uint32_t> next((size_t)n * 6);
    HIPCHK(h, hipMemcpyAsync(geo.data(), h->geo.p, n * sizeof(CellGeo), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(sig.data(), h->sigma.p, sig.size() * sizeof(double2),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(next.data(), h->next.p, next.size() * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::vector<double> nodes((size_t)n * 6), sa((size_t)n * G), ss((size_t)n * G), f(n);
    std::vector<int32_t> bc((size_t)n * 6);
    for (uint32_t i = 0; i < n; ++i) {
      for (int q = 0; q < 6; ++q) {
        nodes[(size_t)i * 6 + q] = geo[i].nodes[q];
        bc[(size_t)i * 6 + q] = (int32_t)((geo[i].bc_bits >> (3 * q)) & 7u);
      }
      f[i] = geo[i].fleck;
    }
    for (size_t k = 0; k < sig.size(); ++k) {
      sa[k] = sig[k].x;
      ss[k] = sig[k].y;
    }
    const bh_cells view{n, G, nodes.data(), bc.data(), next.data(), sa.data(), ss.data(),
                        f.data()};
    TRY(build_inst_lattice(h, &view));
  }
  return BH_OK;
}

int bh_source_instances(bh_handle* h, const uint64_t* n_user_photons, const double* total_E,
                        uint32_t cycle, uint32_t rank, double dt, uint64_t* n_new_out) {
  if (!h || !h->stream || !n_user_photons || !total_E) return BH_ERR_ARG;
  if (!h->has_mesh || !h->has_energy)
    return fail(h, BH_ERR_STATE,
                "bh_source_instances needs bh_set_mesh and bh_set_source_energy");
  if (!h->n_inst) return fail(h, BH_ERR_STATE, "no instances: call bh_set_instances");
  HIPCHK(h, hipSetDevice(h->device));
  const uint32_t K = h->n_inst;
  std::vector<SourceInst> p(2 * (size_t)K);
  for (uint32_t k = 0; k < K; ++k) {
    if (!(total_E[k] > 0.0)) return fail(h, BH_ERR_ARG, "instance %u: total_E must be > 0", k);
    const uint64_t rank_off = n_user_photons[k] * (uint64_t)rank;
    p[k] = SourceInst{n_user_photons[k], total_E[k], rank_off, h->inst_hi[k], h->inst_first[k], 0};
    p[K + k] = SourceInst{n_user_photons[k], total_E[k],
                          10000000000000ull * (uint64_t)cycle + rank_off, h->inst_hi[k],
                          h->inst_first[k], 0};
  }
  HIPCHK(h, hipMemcpyAsync(h->inst_src.p, p.data(), p.size() * sizeof(SourceInst),
                           hipMemcpyHostToDevice, h->stream));
  const SourceInst unused{};
  return run_source(h, unused, unused, h->inst_src.p, h->inst_src.p + h->n_inst, dt, n_new_out);
}

int bh_source_from_energy(bh_handle* h, const double* emission_E, const double* source_E,
                          const double* census_E, const int8_t* source_face,
                          uint64_t n_user_photons, double total_E, uint32_t seed,
                          uint32_t cycle, uint32_t rank, double dt, uint64_t* n_new) {
  TRY(bh_set_source_energy(h, emission_E, source_E, census_E, source_face));
  return bh_source(h, n_user_photons, total_E, seed, cycle, rank, dt, n_new);
}

int bh_transport(bh_handle* h, double next_dt, double* abs_E, double* track_E,
                 bh_step_stats* stats) {
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "bh_transport before bh_set_mesh");
  HIPCHK(h, hipSetDevice(h->device));
  const uint64_t n = h->n_cen + h->n_new;
  ZeroList zl;
  TRY(zero_tallies_and_counters(h, &zl));
  // the kernels write the census to C themselves (post_process_photons,
  // src/post_process_functions.h:32-109); the reference's census order is
  // restored from the transport indices afterwards
  CensusOut co;
  TRY(census_prepare(h, n, next_dt, co, &zl));
  co.inst_exit = (h->inst_exit_on && h->n_inst) ? h->inst_exit.p : nullptr;
  h->replay_timed = false;
  // new photons not yet in W are made inside the persistent kernel
  if (h->src_deferred && h->opt.algorithm != BH_ALGO_PERSISTENT)
    TRY(materialize_source(h, h->n_new));
  SourceDesc sd = h->sd;
  sd.on = h->src_deferred ? 1 : 0;
  h->src_deferred = false;
  HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
  TRY(launch_transport(h, h->W.soa, n, nullptr, co, handle_sinks(h, h->ctr_hi), sd));
  HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
  TRY(census_finish(h, n, nullptr));
  HIPCHK(h, hipEventRecord(h->ev[2], h->stream));

  // small readback: census totals + counters
  char* hp = static_cast<char*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp, h->d_ctot, sizeof(CensusPartial), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(hp + 256, h->d_ctrs, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->stream));
  if (abs_E)
    HIPCHK(h, hipMemcpyAsync(abs_E, h->abs_E.p, h->n_cells * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
  if (track_E)
    HIPCHK(h, hipMemcpyAsync(track_E, h->track_E.p, h->n_cells * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  CensusPartial tot;
  DevCounters ctrs;
  std::memcpy(&tot, hp, sizeof tot);
  std::memcpy(&ctrs, hp + 256, sizeof ctrs);
  if (ctrs.bad_args) {  // a transport kernel's view of its own arguments disagreed
    h->n_cen = 0;
    h->n_new = 0;
    return fail(h, BH_ERR_DEVICE, "transport kernel argument check failed; nothing transported");
  }
  if (tot.overflow) {  // excluded by the sizing of C and the hole records
    h->n_cen = 0;
    h->n_new = 0;
    return fail(h, BH_ERR_DEVICE, "census output exceeded its capacity (%llu slots)",
                (unsigned long long)h->C.cap);
  }

  std::swap(h->W, h->C);
  h->n_cen = tot.n_census;
  h->n_new = 0;
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->census_E = tot.census_E;
    stats->exit_E = tot.exit_E;
    stats->n_transported = n;
    stats->n_census = tot.n_census;
    stats->n_exit = tot.n_exit;
    stats->n_killed = tot.n_killed;
    stats->n_segments = ctrs.segments;
    stats->n_flushes = ctrs.flushes;
    stats->n_replayed = ctrs.replays;
    stats->lattice = h->last_lattice;
    stats->tally = h->last_tally;
    stats->n_allocs = take_allocs(h);
    stats->transport_ms = elapsed_ms(h->ev[0], h->ev[1]);
    stats->census_ms = elapsed_ms(h->ev[1], h->ev[2]);
    stats->source_ms = source_ms(h);
    stats->replay_ms = h->replay_timed ? elapsed_ms(h->ev[3], h->ev[1]) : 0.0;
  }
  if (ctrs.stuck)
    return fail(h, BH_ERR_DEVICE, "%llu photon(s) exceeded %u segments and were stopped",
                ctrs.stuck, kMaxSegments);
  if (tot.n_other)
    return fail(h, BH_ERR_STATE, "%llu photon(s) ended without EXIT/CENSUS/KILLED",
                tot.n_other);
  return BH_OK;
}

// ---------------------------------------------------------------------------
// A step's device phase as one HIP graph launch.
// ---------------------------------------------------------------------------
// photons_for of the source kernels, on the host: the same IEEE operations,
// so the host knows the step's photon count exactly before the launch.
static inline uint64_t host_photons_for(uint64_t n_user, double E, double total_E) {
  if (!(E > 0.0)) return 0u;
  const uint32_t n = (uint32_t)(int)((double)n_user * E / total_E);
  return n ? n : 1u;
}

// Everything a captured step bakes in: buffer addresses, sizes and the
// options that choose its kernels.  A change means capturing again.
static std::vector<uintptr_t> step_graph_key(const bh_handle* h) {
  auto u = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
  return {u(h->W.slab),      u(h->C.slab),    (uintptr_t)h->W.cap,  (uintptr_t)h->C.cap,
          u(h->geo.p),       u(h->sigma.p),   u(h->next.p),         u(h->em.p),
          u(h->src.p),       u(h->face.p),    u(h->counts.p),       u(h->src_cell.p),
          u(h->off.p),
          u(h->soff.p),      u(h->src_hint.p),
          u(h->tiles.p),     u(h->abs_E.p),
          u(h->track_E.p),   u(h->rep_idx.p), u(h->rep_first.p),    u(h->rep_E.p),
          u(h->h_tally),     (uintptr_t)h->n_cells, (uintptr_t)h->G,
          u(h->rec.p),       u(h->axes.p),    (uintptr_t)h->lattice,
          u(h->lat.p),       (uintptr_t)h->inst_lattice,
          u(h->cmark.p),     u(h->mwords.p),  u(h->mcounts.p),      u(h->erank.p),
          u(h->d_cacc),      u(h->sum_part.p),
          u(h->hrec.p),      u(h->hole_len.p), u(h->hole_off.p),   u(h->tail_live.p),
          u(h->tail_rank.p), u(h->tile_E.p),
          (uintptr_t)h->gray_sigma_a, (uintptr_t)use_hash_tally(h), (uintptr_t)h->ctr_hi};
}

// Records source + transport + census + readback on h->stream (being
// captured).  Sizes come from h->d_params at run time; `cap` bounds them.
static int record_step(bh_handle* h, uint64_t cap, bool gen) {
  const uint32_t nc = h->n_cells;
  StepParams* P = h->d_params;
  HIPCHK(h, hipMemcpyAsync(P, h->h_params, sizeof(StepParams), hipMemcpyHostToDevice,
                           h->stream));
  const SourceInst unused{};
  k_source_counts<<<grid_for(nc, 256), 256, 0, h->stream>>>(h->em.p, h->src.p, nc, unused,
                                                             &P->fresh, nullptr, 0, h->counts.p, h->src_cell.p);
  TRY(scan_u64(h, h->counts.p, nc, h->soff.p, 1));
  // the new photons: made inside the transport kernel, or written to W
  k_source_hints<<<std::min<unsigned>(grid_for(div_up(cap, kHintStride), 256), 1024), 256, 0,
                   h->stream>>>(h->soff.p, nc, cap, &P->n_new, h->src_hint.p);
  if (!gen)
    k_source_fill<<<(unsigned)h->n_cu * 16, 256, 0, h->stream>>>(
        dev_cells(h), h->src_cell.p, h->face.p, h->soff.p, h->src_hint.p, 0, unused,
        &P->fresh, 0.0, 0, h->W.soa, 0, P, nullptr);
  HIPCHK(h, hipGetLastError());
  const SourceDesc sd{h->src_cell.p, h->face.p, h->soff.p, h->src_hint.p, &P->fresh,
                      unused,  0.0,      cap,       gen ? 1 : 0, 0};
  ZeroList zl;
  TRY(zero_tallies_and_counters(h, &zl));
  CensusOut co;  // n_cen, n_new and the census life come from P on the device
  TRY(census_prepare(h, cap, 0.0, co, &zl));
  TRY(launch_transport(h, h->W.soa, cap, P, co, handle_sinks(h, h->ctr_hi), sd));
  TRY(census_finish(h, cap, &P->n));
  char* hp = static_cast<char*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp, h->d_ctot, sizeof(CensusPartial), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(hp + 256, h->d_ctrs, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(hp + 512, h->d_total + 1, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->h_tally, h->abs_E.p, nc * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->h_tally + nc, h->track_E.p, nc * sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  return BH_OK;
}

int bh_step_graph(bh_handle* h, uint64_t n_user_photons, double total_E, uint32_t seed,
                  uint32_t cycle, uint32_t rank, double dt, double next_dt, double* abs_E,
                  double* track_E, bh_step_stats* stats) {
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh || !h->has_energy)
    return fail(h, BH_ERR_STATE, "bh_step_graph needs bh_set_mesh and bh_set_source_energy");
  if (h->opt.algorithm != BH_ALGO_PERSISTENT)
    return fail(h, BH_ERR_STATE, "bh_step_graph needs the persistent algorithm");
  if (h->n_inst) return fail(h, BH_ERR_STATE, "bh_step_graph runs one problem (no instances)");
  if (!(total_E > 0.0)) return fail(h, BH_ERR_ARG, "total_E must be > 0");
  if (h->has_census_E) {  // the first step makes the initial census: eager
    TRY(bh_source(h, n_user_photons, total_E, seed, cycle, rank, dt, nullptr));
    return bh_transport(h, next_dt, abs_E, track_E, stats);
  }
  HIPCHK(h, hipSetDevice(h->device));
  const uint64_t ctr_hi = (uint64_t)seed << 32;
  if (h->n_cen && ctr_hi != h->ctr_hi)
    return fail(h, BH_ERR_ARG, "seed differs from the resident census photons' seed");
  h->ctr_hi = ctr_hi;
  const uint32_t nc = h->n_cells;
  uint64_t n_new = 0;
  for (uint32_t c = 0; c < nc; ++c)
    n_new += host_photons_for(n_user_photons, h->em_host[c], total_E) +
             host_photons_for(n_user_photons, h->src_host[c], total_E);
  const uint64_t n = h->n_cen + n_new;
  // every buffer the graph touches is sized (and allocated) before capture;
  // both stores hold cap photons plus the census slots the waves may leave
  // unused (a constant of the launch shape), so cap stays put from step to
  // step while the stores swap
  const uint64_t slack = census_sizes(h, 0).slots;
  TRY(grow_store(h, h->W, n + slack, h->n_cen));
  TRY(grow_store(h, h->C, n + slack, 0));
  const uint64_t cap = std::min(h->W.cap, h->C.cap) - slack;
  const CensusSizes z = census_sizes(h, cap);
  const uint64_t pc = photon_cap(h, cap);  // as census_prepare sizes them
  TRY(ensure(h, h->counts, nc));
  TRY(ensure(h, h->soff, nc));
  TRY(ensure(h, h->src_hint, div_up(std::max<uint64_t>(cap, 1), kHintStride) + 1));
  TRY(ensure_marks(h, pc));
  TRY(ensure(h, h->hrec, z.rec_cap));
  TRY(ensure(h, h->hole_len, z.rec_cap));
  TRY(ensure(h, h->hole_off, z.rec_cap));
  TRY(ensure(h, h->tail_live, z.hole_cap));
  TRY(ensure(h, h->tail_rank, z.hole_cap));
  TRY(ensure(h, h->sum_part, (size_t)8 * kSumMaxBlocks));
  const uint64_t scan_max = std::max<uint64_t>(std::max<uint64_t>(nc, pc),
                                               std::max(z.hole_cap, z.rec_cap));
  TRY(ensure(h, h->tiles, std::max<uint64_t>(div_up(scan_max, kScanTile), 1)));
  TRY(ensure(h, h->tile_E, std::max<uint64_t>(div_up(scan_max, kScanTile), 1)));
  if (h->opt.precision == BH_PRECISION_FP32_TRACK) {  // as launch_transport sizes them
    const uint64_t rep = std::max(h->W.cap, h->C.cap);
    TRY(ensure(h, h->rep_idx, rep));
    TRY(ensure(h, h->rep_first, rep));
    TRY(ensure(h, h->rep_E, rep));
  }
  if (!h->h_params) HIPCHK(h, hipHostMalloc(&h->h_params, sizeof(StepParams), hipHostMallocDefault));
  if (h->h_tally_cap < 2 * (size_t)nc) {
    if (h->h_tally) HIPCHK(h, hipHostFree(h->h_tally));
    h->h_tally = nullptr;
    h->h_tally_cap = 0;
    HIPCHK(h, hipHostMalloc(&h->h_tally, 2 * (size_t)nc * sizeof(double), hipHostMallocDefault));
    h->h_tally_cap = 2 * (size_t)nc;
  }
  const uint64_t rank_off = n_user_photons * (uint64_t)rank;
  *h->h_params = StepParams{h->n_cen, n_new, n, kC * next_dt, dt,
                            SourceInst{n_user_photons, total_E,
                                       10000000000000ull * (uint64_t)cycle + rank_off, ctr_hi,
                                       0, 0}};
  // one graph per parity of the photon stores (they swap every step)
  bh_handle::StepGraph& g = h->graphs[h->W.slab < h->C.slab ? 0 : 1];
  const bool gen = source_in_transport(h, n_new);
  std::vector<uintptr_t> key = step_graph_key(h);
  key.push_back((uintptr_t)gen);
  if (!g.exec || g.key != key) {
    if (g.exec) HIPCHK(h, hipGraphExecDestroy(g.exec));
    if (g.graph) HIPCHK(h, hipGraphDestroy(g.graph));
    g.exec = nullptr;
    g.graph = nullptr;
    g.key.clear();
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_step(h, cap, gen);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc != BH_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess)
      return fail(h, BH_ERR_DEVICE, "capturing the step failed: %s", hipGetErrorString(e));
    g.graph = graph;
    HIPCHK(h, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    g.key = key;
  }
  HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
  HIPCHK(h, hipGraphLaunch(g.exec, h->stream));
  HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  CensusPartial tot;
  DevCounters ctrs;
  uint64_t dev_new = 0;
  const char* hp = static_cast<const char*>(h->h_pinned);
  std::memcpy(&tot, hp, sizeof tot);
  std::memcpy(&ctrs, hp + 256, sizeof ctrs);
  if (ctrs.bad_args) {  // a transport kernel's view of its own arguments disagreed
    h->n_cen = 0;
    h->n_new = 0;
    return fail(h, BH_ERR_DEVICE, "transport kernel argument check failed; nothing transported");
  }
  std::memcpy(&dev_new, hp + 512, sizeof dev_new);
  if (tot.overflow) {  // excluded by the sizing of C and the hole records
    h->n_cen = 0;
    h->n_new = 0;
    return fail(h, BH_ERR_DEVICE, "census output exceeded its capacity (%llu slots)",
                (unsigned long long)h->C.cap);
  }
  if (dev_new != n_new) {  // the host and device counts are the same IEEE arithmetic
    h->n_cen = 0;
    h->n_new = 0;
    return fail(h, BH_ERR_DEVICE, "step graph: device made %llu photons, host expected %llu",
                (unsigned long long)dev_new, (unsigned long long)n_new);
  }
  if (abs_E) std::memcpy(abs_E, h->h_tally, nc * sizeof(double));
  if (track_E) std::memcpy(track_E, h->h_tally + nc, nc * sizeof(double));
  std::swap(h->W, h->C);
  h->n_cen = tot.n_census;
  h->n_new = 0;
  h->last_source_ms = 0.0;
  h->src_ms_pending = false;
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->census_E = tot.census_E;
    stats->exit_E = tot.exit_E;
    stats->n_transported = n;
    stats->n_census = tot.n_census;
    stats->n_exit = tot.n_exit;
    stats->n_killed = tot.n_killed;
    stats->n_segments = ctrs.segments;
    stats->n_flushes = ctrs.flushes;
    stats->n_replayed = ctrs.replays;
    stats->lattice = h->last_lattice;
    stats->tally = h->last_tally;
    stats->n_allocs = take_allocs(h);
    stats->transport_ms = elapsed_ms(h->ev[0], h->ev[1]);  // the whole graph
  }
  if (ctrs.stuck)
    return fail(h, BH_ERR_DEVICE, "%llu photon(s) exceeded %u segments and were stopped",
                ctrs.stuck, kMaxSegments);
  if (tot.n_other)
    return fail(h, BH_ERR_STATE, "%llu photon(s) ended without EXIT/CENSUS/KILLED",
                tot.n_other);
  return BH_OK;
}

int bh_get_tallies(bh_handle* h, double* abs_E, double* track_E) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "no mesh");
  if (abs_E)
    HIPCHK(h, hipMemcpyAsync(abs_E, h->abs_E.p, h->n_cells * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
  if (track_E)
    HIPCHK(h, hipMemcpyAsync(track_E, h->track_E.p, h->n_cells * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return BH_OK;
}

int bh_census_size(bh_handle* h, uint64_t* n) {
  if (!h || !n) return BH_ERR_ARG;
  *n = h->n_cen;
  return BH_OK;
}

int bh_census_export(bh_handle* h, bh_photon* out, uint64_t cap, uint64_t* n) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || !n) return BH_ERR_ARG;
  *n = h->n_cen;
  if (h->n_cen == 0) return BH_OK;
  if (!out || cap < h->n_cen)
    return fail(h, BH_ERR_CAPACITY, "census has %llu photons, buffer holds %llu",
                (unsigned long long)h->n_cen, (unsigned long long)cap);
  HIPCHK(h, hipSetDevice(h->device));
  TRY(ensure(h, h->aos, h->n_cen));
  k_soa_to_aos<<<grid_for(h->n_cen, 256), 256, 0, h->stream>>>(h->W.soa, 0, h->n_cen,
                                                                h->ctr_hi, dev_cells(h),
                                                                h->aos.p, 1, 0);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, h->aos.p, h->n_cen * sizeof(AosPhoton), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return BH_OK;
}

int bh_photons_export(bh_handle* h, bh_photon* out, uint64_t cap, uint64_t* n) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || !n) return BH_ERR_ARG;
  const uint64_t total = h->n_cen + h->n_new;
  *n = total;
  if (total == 0) return BH_OK;
  if (!out || cap < total)
    return fail(h, BH_ERR_CAPACITY, "%llu resident photons, buffer holds %llu",
                (unsigned long long)total, (unsigned long long)cap);
  HIPCHK(h, hipSetDevice(h->device));
  TRY(materialize_source(h, h->n_new));
  TRY(ensure(h, h->aos, total));
  // new photons live at W[n_cen, n_cen + n_new), census at W[0, n_cen)
  if (h->n_new)
    k_soa_to_aos<<<grid_for(h->n_new, 256), 256, 0, h->stream>>>(h->W.soa, h->n_cen, h->n_new,
                                                                  h->ctr_hi, dev_cells(h),
                                                                  h->aos.p, 0, 0);
  if (h->n_cen)
    k_soa_to_aos<<<grid_for(h->n_cen, 256), 256, 0, h->stream>>>(h->W.soa, 0, h->n_cen,
                                                                  h->ctr_hi, dev_cells(h),
                                                                  h->aos.p, 1, h->n_new);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, h->aos.p, total * sizeof(AosPhoton), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return BH_OK;
}

// Checks the caller's AOS photons against the mesh and returns the RNG
// counter-high word they share (photon 0's; per instance with instances).
// A host loop only: no allocation.
static int check_aos(bh_handle* h, const bh_photon* in, uint64_t n, uint64_t* hi_out) {
  const uint64_t hi = n ? in[0].rng[1] : 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (h->has_mesh && in[i].cell >= h->n_cells)
      return fail(h, BH_ERR_ARG, "photon %llu: cell %u out of range",
                  (unsigned long long)i, in[i].cell);
    if (h->has_mesh && in[i].group >= h->G)
      return fail(h, BH_ERR_ARG, "photon %llu: group %u out of range",
                  (unsigned long long)i, in[i].group);
    // the device keeps one counter-high word per run (or per instance)
    const uint64_t want = h->n_inst ? h->inst_hi[h->cell_inst_host[in[i].cell]] : hi;
    if (in[i].rng[1] != want || in[i].rng[3] != 0)
      return fail(h, BH_ERR_ARG,
                  h->n_inst ? "photon %llu: RNG counter-high/key-high words differ from its "
                              "instance's seed"
                            : "photon %llu: RNG counter-high/key-high words differ from "
                              "photon 0 (one seed per run is required)",
                  (unsigned long long)i);
  }
  *hi_out = hi;
  return BH_OK;
}

// Uploads AOS photons into store dst (after its first `keep`).  *hi_out
// receives their counter-high word; the handle's own (that of the resident
// census) is left as it is.
static int upload_aos(bh_handle* h, const bh_photon* in, uint64_t n, PhotonStore& dst,
                      uint64_t keep, uint64_t* hi_out) {
  *hi_out = h->ctr_hi;
  if (n == 0) return BH_OK;
  uint64_t hi = 0;
  TRY(check_aos(h, in, n, &hi));
  TRY(ensure(h, h->aos, n));
  TRY(grow_store(h, dst, n, keep));
  HIPCHK(h, hipMemcpyAsync(h->aos.p, in, n * sizeof(AosPhoton), hipMemcpyHostToDevice,
                           h->stream));
  k_aos_to_soa<<<grid_for(n, 256), 256, 0, h->stream>>>(h->aos.p, n, dst.soa, 0);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  *hi_out = hi;
  return BH_OK;
}

int bh_census_import(bh_handle* h, const bh_photon* in, uint64_t n) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || (n && !in)) return BH_ERR_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  uint64_t hi = 0;
  TRY(upload_aos(h, in, n, h->W, 0, &hi));
  h->ctr_hi = hi;
  h->n_cen = n;
  h->n_new = 0;
  h->src_deferred = false;
  return BH_OK;
}

// Per-photon outcome totals of transported AOS photons (the
// post_process_photons counts, src/post_process_functions.h:32-59).
static void outcome_stats(const bh_photon* photons, uint64_t n, bh_step_stats* stats) {
  for (uint64_t i = 0; i < n; ++i) {
    const uint8_t d = photons[i].descriptor[0];
    if (d == BH_CENSUS) {
      stats->n_census++;
      stats->census_E += photons[i].E;
    } else if (d == BH_EXIT) {
      stats->n_exit++;
      stats->exit_E += photons[i].E;
    } else if (d == BH_KILLED) {
      stats->n_killed++;
    }
  }
}

int bh_transport_photons(bh_handle* h, bh_photon* photons, uint64_t n, double* abs_E,
                         double* track_E, bh_step_stats* stats) {
  if (!h || !h->stream || (n && !photons)) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "bh_transport_photons before bh_set_mesh");
  HIPCHK(h, hipSetDevice(h->device));
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (n == 0) return BH_OK;
  // the photons' counter-high word goes to the kernels; the resident
  // census keeps its own
  uint64_t hi = 0;
  TRY(upload_aos(h, photons, n, h->C, 0, &hi));
  TRY(zero_tallies_and_counters(h));
  // tally choice from the photons' cells: S evenly spaced samples, effective
  // number of cells S (S - 1) / sum c (c - 1) (the collision estimator: N for
  // photons spread evenly over N cells)
  const double saved_neff = h->src_neff_frac;
  {
    const uint64_t S = std::min<uint64_t>(n, 65536);
    std::vector<uint32_t> cnt(h->n_cells, 0u);
    for (uint64_t k = 0; k < S; ++k) {
      const uint32_t c = photons[k * n / S].cell;
      if (c < h->n_cells) ++cnt[c];
    }
    double coll = 0.0;
    for (const uint32_t v : cnt) coll += (double)v * (double)(v > 0 ? v - 1 : 0);
    h->src_neff_frac = coll > 0.0 ? (double)S * (double)(S - 1) / coll / (double)h->n_cells : 1.0;
  }
  HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
  CensusOut in_place{};  // the caller's photons come back in place, in order
  const int rc_t = launch_transport(h, h->C.soa, n, nullptr, in_place, handle_sinks(h, hi));
  h->src_neff_frac = saved_neff;
  TRY(rc_t);
  HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
  k_soa_to_aos<<<grid_for(n, 256), 256, 0, h->stream>>>(h->C.soa, 0, n, hi, dev_cells(h),
                                                        h->aos.p, 0, 0);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(photons, h->aos.p, n * sizeof(AosPhoton), hipMemcpyDeviceToHost,
                           h->stream));
  std::vector<double> a(h->n_cells), t(h->n_cells);
  HIPCHK(h, hipMemcpyAsync(a.data(), h->abs_E.p, h->n_cells * sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(t.data(), h->track_E.p, h->n_cells * sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  char* hp = static_cast<char*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp + 256, h->d_ctrs, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  DevCounters ctrs;
  std::memcpy(&ctrs, hp + 256, sizeof ctrs);
  // (the resident census was not touched, so it stays)
  if (ctrs.bad_args)  // a transport kernel's view of its own arguments disagreed
    return fail(h, BH_ERR_DEVICE, "transport kernel argument check failed; nothing transported");
  for (uint32_t c = 0; c < h->n_cells; ++c) {
    if (abs_E) abs_E[c] += a[c];
    if (track_E) track_E[c] += t[c];
  }
  if (stats) {
    stats->n_transported = n;
    stats->n_segments = ctrs.segments;
    stats->n_flushes = ctrs.flushes;
    stats->n_replayed = ctrs.replays;
    stats->lattice = h->last_lattice;
    stats->tally = h->last_tally;
    stats->n_allocs = take_allocs(h);
    stats->transport_ms = elapsed_ms(h->ev[0], h->ev[1]);
    outcome_stats(photons, n, stats);
  }
  if (ctrs.stuck)
    return fail(h, BH_ERR_DEVICE, "%llu photon(s) exceeded %u segments and were stopped",
                ctrs.stuck, kMaxSegments);
  return BH_OK;
}

// The scratch layout of one bh_transport_photons_noalloc call; bytes = 0
// when it does not fit.
struct NoallocLayout {
  size_t ctrs, queue, abs_E, track_E, aos, soa, rep_idx, rep_first, rep_E, bytes;
};
static NoallocLayout noalloc_layout(uint64_t n, uint32_t n_cells, bool fp32) {
  NoallocLayout L{};
  size_t o = 0;
  auto take = [&](size_t b) {
    const size_t at = o;
    o += align_up(b);
    return at;
  };
  L.ctrs = take(sizeof(DevCounters));
  L.queue = take(sizeof(PersistentQueue));
  L.abs_E = take((size_t)n_cells * sizeof(double));
  L.track_E = take((size_t)n_cells * sizeof(double));
  L.aos = take((size_t)n * sizeof(AosPhoton));
  L.soa = take(PhotonStore::bytes_for(n));
  if (fp32) {
    L.rep_idx = take((size_t)n * sizeof(uint64_t));
    L.rep_first = take((size_t)n * sizeof(uint32_t));
    L.rep_E = take((size_t)n * sizeof(double));
  }
  L.bytes = o;
  return L;
}

int bh_transport_photons_noalloc(bh_handle* h, bh_photon* photons, uint64_t n, double* abs_E,
                                 double* track_E, bh_step_stats* stats) {
  if (!h || !h->stream || (n && !photons)) return BH_ERR_ARG;
  if (!h->has_mesh)
    return fail(h, BH_ERR_STATE, "bh_transport_photons_noalloc before bh_set_mesh");
  if (!h->scratch)
    return fail(h, BH_ERR_DEVICE, "no scratch memory on device %d: %s", h->device,
                h->scratch_err.c_str());
  if (h->opt.algorithm == BH_ALGO_EVENT)
    return fail(h, BH_ERR_STATE,
                "bh_transport_photons_noalloc runs the history or persistent algorithm");
  HIPCHK(h, hipSetDevice(h->device));
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (n == 0) return BH_OK;
  const bool fp32 = h->opt.precision == BH_PRECISION_FP32_TRACK;
  const NoallocLayout L = noalloc_layout(n, h->n_cells, fp32);
  if (L.bytes > kScratchBytes) {
    const NoallocLayout one = noalloc_layout(1, h->n_cells, fp32);
    const NoallocLayout two = noalloc_layout(2, h->n_cells, fp32);
    const size_t per = two.bytes > one.bytes ? two.bytes - one.bytes : 1;
    return fail(h, BH_ERR_CAPACITY,
                "%llu photons on %u cells need %zu bytes of device scratch; it holds %zu "
                "(about %llu photons on this mesh)",
                (unsigned long long)n, h->n_cells, L.bytes, kScratchBytes,
                (unsigned long long)(kScratchBytes > one.bytes ? (kScratchBytes - one.bytes) / per
                                                               : 0));
  }
  uint64_t hi = 0;
  TRY(check_aos(h, photons, n, &hi));
  char* base = static_cast<char*>(h->scratch);
  Sinks sk;
  sk.abs_E = reinterpret_cast<double*>(base + L.abs_E);
  sk.track_E = reinterpret_cast<double*>(base + L.track_E);
  sk.ctrs = reinterpret_cast<DevCounters*>(base + L.ctrs);
  sk.queue = reinterpret_cast<PersistentQueue*>(base + L.queue);
  sk.rl = fp32 ? ReplayList{reinterpret_cast<uint64_t*>(base + L.rep_idx),
                            reinterpret_cast<uint32_t*>(base + L.rep_first),
                            reinterpret_cast<double*>(base + L.rep_E)}
               : ReplayList{nullptr, nullptr, nullptr};
  sk.ctr_hi = hi;
  sk.additive = true;  // the caller's tallies are added to on the device
  AosPhoton* aos = reinterpret_cast<AosPhoton*>(base + L.aos);
  const PhotonSoA soa = PhotonStore::carve(base + L.soa, n);
  const uint32_t nc = h->n_cells;
  // the previous call's contents are still there: clear what starts at zero
  k_fill_u64<<<1, 64, 0, h->stream>>>(reinterpret_cast<unsigned long long*>(sk.ctrs),
                                      sizeof(DevCounters) / 8);
  if (abs_E)
    HIPCHK(h, hipMemcpyAsync(sk.abs_E, abs_E, nc * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
  else
    k_fill_f64<<<grid_for(nc, 256), 256, 0, h->stream>>>(sk.abs_E, nc, 0.0);
  if (track_E)
    HIPCHK(h, hipMemcpyAsync(sk.track_E, track_E, nc * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
  else
    k_fill_f64<<<grid_for(nc, 256), 256, 0, h->stream>>>(sk.track_E, nc, 0.0);
  HIPCHK(h, hipMemcpyAsync(aos, photons, n * sizeof(AosPhoton), hipMemcpyHostToDevice,
                           h->stream));
  k_aos_to_soa<<<grid_for(n, 256), 256, 0, h->stream>>>(aos, n, soa, 0);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
  CensusOut in_place{};  // the caller's photons come back in place, in order
  TRY(launch_transport(h, soa, n, nullptr, in_place, sk));
  HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
  k_soa_to_aos<<<grid_for(n, 256), 256, 0, h->stream>>>(soa, 0, n, hi, dev_cells(h), aos, 0, 0);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(photons, aos, n * sizeof(AosPhoton), hipMemcpyDeviceToHost,
                           h->stream));
  if (abs_E)
    HIPCHK(h, hipMemcpyAsync(abs_E, sk.abs_E, nc * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  if (track_E)
    HIPCHK(h, hipMemcpyAsync(track_E, sk.track_E, nc * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  char* hp = static_cast<char*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp + 256, sk.ctrs, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  DevCounters ctrs;
  std::memcpy(&ctrs, hp + 256, sizeof ctrs);
  if (ctrs.bad_args)  // a transport kernel's view of its own arguments disagreed
    return fail(h, BH_ERR_DEVICE, "transport kernel argument check failed; nothing transported");
  if (stats) {
    stats->n_transported = n;
    stats->n_segments = ctrs.segments;
    stats->n_flushes = ctrs.flushes;
    stats->n_replayed = ctrs.replays;
    stats->lattice = h->last_lattice;
    stats->tally = h->last_tally;
    stats->n_allocs = take_allocs(h);
    stats->transport_ms = elapsed_ms(h->ev[0], h->ev[1]);
    outcome_stats(photons, n, stats);
  }
  if (ctrs.stuck)
    return fail(h, BH_ERR_DEVICE, "%llu photon(s) exceeded %u segments and were stopped",
                ctrs.stuck, kMaxSegments);
  return BH_OK;
}

int bh_adjoint_photons(bh_handle* h, const bh_photon* photons, uint64_t n,
                       const double* abs_adjoint, double* d_sigma_a, double* d_fleck,
                       bh_step_stats* stats) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || (n && !photons) || !abs_adjoint || !d_sigma_a || !d_fleck)
    return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "bh_adjoint_photons before bh_set_mesh");
  if (!h->uniform_sigma_a)
    return fail(h, BH_ERR_ARG,
                "the adjoint needs sigma_a equal in every group of a cell (the group "
                "walk's probabilities must not depend on it)");
  HIPCHK(h, hipSetDevice(h->device));
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (n == 0) return BH_OK;
  uint64_t hi = 0;  // the photons' counter-high word; the resident census keeps its own
  TRY(upload_aos(h, photons, n, h->C, 0, &hi));
  const uint32_t nc = h->n_cells;
  TRY(ensure(h, h->adj_g, nc));
  TRY(ensure(h, h->grad_sa, nc));
  TRY(ensure(h, h->grad_f, nc));
  HIPCHK(h, hipMemcpyAsync(h->adj_g.p, abs_adjoint, nc * sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
  k_fill_f64<<<grid_for(nc, 256), 256, 0, h->stream>>>(h->grad_sa.p, nc, 0.0);
  k_fill_f64<<<grid_for(nc, 256), 256, 0, h->stream>>>(h->grad_f.p, nc, 0.0);
  k_fill_u64<<<1, 64, 0, h->stream>>>(reinterpret_cast<unsigned long long*>(h->d_ctrs),
                                      sizeof(DevCounters) / 8);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
  GlobalTally<true> sink{h->grad_sa.p, h->grad_f.p, nullptr, nullptr, 0, 0};
  k_adjoint<GlobalTally<true>><<<grid_for(n, 256), 256, 0, h->stream>>>(
      dev_cells(h), h->C.soa, n, hi, h->adj_g.p, sink, h->d_ctrs);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
  std::vector<double> a(nc), f(nc);
  HIPCHK(h, hipMemcpyAsync(a.data(), h->grad_sa.p, nc * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(f.data(), h->grad_f.p, nc * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  char* hp = static_cast<char*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp + 256, h->d_ctrs, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  DevCounters ctrs;
  std::memcpy(&ctrs, hp + 256, sizeof ctrs);
  // (the resident census was not touched, so it stays)
  if (ctrs.bad_args)  // a transport kernel's view of its own arguments disagreed
    return fail(h, BH_ERR_DEVICE, "transport kernel argument check failed; nothing transported");
  for (uint32_t c = 0; c < nc; ++c) {
    d_sigma_a[c] += a[c];
    d_fleck[c] += f[c];
  }
  if (stats) {
    stats->n_transported = n;
    stats->n_segments = ctrs.segments;
    stats->n_flushes = ctrs.flushes;
    stats->transport_ms = elapsed_ms(h->ev[0], h->ev[1]);
  }
  if (ctrs.stuck)
    return fail(h, BH_ERR_DEVICE, "%llu photon(s) exceeded %u segments and were stopped",
                ctrs.stuck, kMaxSegments);
  return BH_OK;
}

// ---- census population control ------------------------------------------------
int bh_comb_census(bh_handle* h, int64_t max_census_photons, uint32_t seed, uint64_t stream,
                   uint64_t ctr0, bh_comb_report* rep) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "bh_comb_census before bh_set_mesh");
  if (max_census_photons <= 0) return fail(h, BH_ERR_ARG, "max_census_photons must be > 0");
  if (h->n_new) return fail(h, BH_ERR_STATE, "new photons are waiting for transport");
  if (h->n_cen >= 0xFFFFFFFFull) return fail(h, BH_ERR_ARG, "census too large to comb");
  if (rep) std::memset(rep, 0, sizeof *rep);
  const uint32_t n = (uint32_t)h->n_cen, nc = h->n_cells;
  if (rep) rep->n_before = n;
  if (n == 0) return BH_OK;
  HIPCHK(h, hipSetDevice(h->device));
  const uint64_t ctr_hi = (uint64_t)seed << 32;
  // E_census: deterministic parallel sum, with the bound on the serial loop's
  const uint32_t nb = std::max<uint32_t>(1u, std::min<uint32_t>(kSumMaxBlocks, div_up(n, 4096)));
  TRY(ensure(h, h->sum_part, (size_t)8 * nb));
  k_energy_partials<<<nb, kSumNT, 0, h->stream>>>(h->W.soa.E, n, h->sum_part.p);
  k_final_sums<1><<<1, kSumNT, 0, h->stream>>>(h->sum_part.p, nb, &h->d_phys->global_E);
  HIPCHK(h, hipGetLastError());
  char* hp = static_cast<char*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp + 3072, &h->d_phys->global_E, sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  // stable sort of the census by cell: each cell's photons in census order
  TRY(ensure(h, h->comb_keys, n));
  TRY(ensure(h, h->comb_rank_cell, n));
  TRY(ensure(h, h->comb_idx, n));
  TRY(ensure(h, h->comb_sorted, n));
  TRY(ensure(h, h->comb_keep, n));
  TRY(ensure(h, h->comb_cellE, nc));
  TRY(ensure(h, h->comb_ctr, 2));
  TRY(ensure(h, h->counts, nc));
  TRY(ensure(h, h->off, nc));
  // the census in the reference's order: perm[rank] = slot of W
  TRY(ensure(h, h->comb_perm, n));
  k_iota_u32<<<grid_for(n, 256), 256, 0, h->stream>>>(h->comb_perm.p, n);  // in bounds always
  k_census_perm<<<grid_for(n, 256), 256, 0, h->stream>>>(h->W.soa.key, n, h->comb_perm.p);
  k_iota_u32<<<grid_for(n, 256), 256, 0, h->stream>>>(h->comb_idx.p, n);
  k_cells_by_rank<<<grid_for(n, 256), 256, 0, h->stream>>>(h->W.soa.cell, h->comb_perm.p, n,
                                                           h->comb_rank_cell.p);
  int bits = 1;
  while (bits < 32 && (1ull << bits) < nc) ++bits;
  size_t tmp_bytes = 0;
  HIPCHK(h, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, h->comb_rank_cell.p,
                                               h->comb_keys.p, h->comb_idx.p, h->comb_sorted.p,
                                               (int)n, 0, bits, h->stream));
  TRY(ensure(h, h->comb_tmp, tmp_bytes));
  HIPCHK(h, hipcub::DeviceRadixSort::SortPairs(h->comb_tmp.p, tmp_bytes, h->comb_rank_cell.p,
                                               h->comb_keys.p, h->comb_idx.p, h->comb_sorted.p,
                                               (int)n, 0, bits, h->stream));
  HIPCHK(h, hipMemsetAsync(h->counts.p, 0, nc * sizeof(uint64_t), h->stream));
  k_cell_hist<<<grid_for(n, 256), 256, 0, h->stream>>>(
      h->W.soa.cell, n, reinterpret_cast<unsigned long long*>(h->counts.p));
  HIPCHK(h, hipGetLastError());
  TRY(scan_u64(h, h->counts.p, nc, h->off.p, 0));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double E_par;
  std::memcpy(&E_par, hp + 3072, sizeof E_par);
  const double delta = 2.0 * ((double)n + 64.0) * 0x1p-53 * E_par;
  const double m = (double)max_census_photons;
  CombParams cp{E_par / m, (E_par - delta) / m, (E_par + delta) / m, ctr0, stream, ctr_hi};
  auto comb = [&](uint32_t* amb, uint64_t* kept) -> int {
    // comb_ctr: [0] ambiguous flag (low 32 bits), [1] survivors
    HIPCHK(h, hipMemsetAsync(h->comb_ctr.p, 0, 2 * sizeof(unsigned long long), h->stream));
    k_comb_cells<<<grid_for(nc, 256), 256, 0, h->stream>>>(
        h->comb_sorted.p, h->comb_perm.p, reinterpret_cast<const unsigned long long*>(h->off.p),
        reinterpret_cast<const unsigned long long*>(h->counts.p), nc, h->W.soa.E, cp,
        h->comb_keep.p, h->comb_cellE.p, reinterpret_cast<uint32_t*>(h->comb_ctr.p),
        h->comb_ctr.p + 1);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(hp + 3072, h->comb_ctr.p, 2 * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    uint64_t v[2];
    std::memcpy(v, hp + 3072, sizeof v);
    *amb = (uint32_t)v[0];
    *kept = v[1];
    return BH_OK;
  };
  uint32_t amb = 0;
  uint64_t kept = 0;
  TRY(comb(&amb, &kept));
  const bool serial = amb != 0;
  if (serial) {  // the reference's serial census energy, then comb again
    k_serial_energy<<<1, kWave, 0, h->stream>>>(h->W.soa.E, h->comb_perm.p, n,
                                                &h->d_phys->global_E);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(hp + 3072, &h->d_phys->global_E, sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double E_ser;
    std::memcpy(&E_ser, hp + 3072, sizeof E_ser);
    cp.comb_E = cp.comb_lo = cp.comb_hi = E_ser / m;
    E_par = E_ser;
    TRY(comb(&amb, &kept));
  }
  // survivors in census order into C: positions by exclusive scan of the
  // keep flags, new census ranks as keys
  TRY(ensure(h, h->cflag, n));
  TRY(ensure(h, h->off, std::max<uint64_t>(n, nc)));
  TRY(grow_store(h, h->C, n, 0));
  k_keep_flags<<<grid_for(n, 256), 256, 0, h->stream>>>(h->comb_keep.p, n, h->cflag.p);
  TRY(scan_u64(h, h->cflag.p, n, h->off.p, 2));
  k_comb_compact<<<grid_for(n, 256), 256, 0, h->stream>>>(h->W.soa, h->comb_perm.p, n,
                                                          h->comb_keep.p, h->off.p,
                                                          h->comb_cellE.p, h->C.soa);
  HIPCHK(h, hipGetLastError());
  // census energy of the survivors (deterministic sum over C)
  k_energy_partials<<<nb, kSumNT, 0, h->stream>>>(h->C.soa.E, (uint32_t)kept, h->sum_part.p);
  k_final_sums<1><<<1, kSumNT, 0, h->stream>>>(h->sum_part.p, nb, &h->d_phys->pre_census);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(hp + 3072, &h->d_phys->pre_census, sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double census_after;
  std::memcpy(&census_after, hp + 3072, sizeof census_after);
  std::swap(h->W, h->C);
  h->n_cen = kept;
  if (rep) {
    rep->n_after = kept;
    rep->comb_E = cp.comb_E;
    rep->census_E_before = E_par;
    rep->census_E_after = census_after;
    rep->comb_serial = serial ? 1 : 0;
  }
  return BH_OK;
}

// ---- device-resident step physics -------------------------------------------
static MatDev mat_dev(bh_handle* h) {
  const size_t n = h->n_cells;
  double* b = h->mat.p;
  return MatDev{b,         b + n,     b + 2 * n, b + 3 * n, b + 4 * n,
                b + 5 * n, b + 6 * n, b + 7 * n, b + 8 * n, h->face.p};
}

int bh_set_material(bh_handle* h, const bh_material* m, double replicated_factor,
                    uint32_t n_ranks, uint32_t rank) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || !m) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "bh_set_material before bh_set_mesh");
  if (h->n_inst && (n_ranks != 1 || replicated_factor != 1.0))
    return fail(h, BH_ERR_ARG, "instances run as one rank (replicated_factor 1)");
  if (n_ranks == 0 || rank >= n_ranks || !(replicated_factor > 0.0))
    return fail(h, BH_ERR_ARG, "need n_ranks >= 1, rank < n_ranks, replicated_factor > 0");
  if (n_ranks > kMaxRanks) return fail(h, BH_ERR_ARG, "at most %u ranks", kMaxRanks);
  const double* cols[9] = {m->opA, m->opB, m->opC, m->opS, m->cv, m->rho, m->T_e, m->T_r, m->T_s};
  for (const double* c : cols)
    if (!c) return fail(h, BH_ERR_ARG, "every bh_material array is required");
  if (!m->source_face) return fail(h, BH_ERR_ARG, "every bh_material array is required");
  const uint32_t n = h->n_cells;
  for (uint32_t c = 0; c < n; ++c) {
    if (m->source_face[c] < -1 || m->source_face[c] > 5)
      return fail(h, BH_ERR_ARG, "cell %u: source_face must be -1..5", c);
    if (!(m->cv[c] > 0.0) || !(m->rho[c] > 0.0))
      return fail(h, BH_ERR_ARG, "cell %u: cv and rho must be > 0", c);
  }
  HIPCHK(h, hipSetDevice(h->device));
  TRY(ensure(h, h->mat, (size_t)9 * n));
  TRY(ensure(h, h->mat_e, n));
  TRY(ensure(h, h->tr_diag, n));
  TRY(ensure(h, h->em, n));
  TRY(ensure(h, h->src, n));
  TRY(ensure(h, h->cen, n));
  TRY(ensure(h, h->face, n));
  if (!h->h_params)
    HIPCHK(h, hipHostMalloc(&h->h_params, sizeof(StepParams), hipHostMallocDefault));
  for (int k = 0; k < 9; ++k)
    HIPCHK(h, hipMemcpyAsync(h->mat.p + (size_t)k * n, cols[k], n * sizeof(double),
                             hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->face.p, m->source_face, n, hipMemcpyHostToDevice, h->stream));
  k_fill_f64<<<grid_for(n, 256), 256, 0, h->stream>>>(h->tr_diag.p, n, 0.0);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->mat_rf = replicated_factor;
  h->mat_ranks = n_ranks;
  h->mat_rank = rank;
  h->has_material = true;
  return BH_OK;
}

}  // extern "C"

namespace {

// Blocks of the deterministic cell sums of a resident step.
uint32_t resident_blocks(const bh_handle* h) {
  return std::max<uint32_t>(1u, std::min<uint32_t>(kSumMaxBlocks, div_up(h->n_cells, 4096)));
}

// Mesh::calculate_photon_energy on the device, with no host round trip:
// per-cell opacities, Fleck factors and source energies, the step totals,
// and E_tot into the source records `rec` (one problem: [census, new];
// instances: [K census | K new]).  One problem and one rank: E_tot summed in
// parallel with a bound, the reference's serial cell-order sum run on the
// device only when a photon count could depend on the difference; several
// ranks: the serial sums, every rank's reassigned total and this rank's
// share; instances: each instance's own totals (k_inst_sums).  P: a step
// captured as a graph (dt read on the device).
int resident_physics(bh_handle* h, bool first, uint64_t n_user, double dt, const StepParams* P,
                     SourceInst* rec, uint32_t nb) {
  const uint32_t n = h->n_cells;
  const MatDev md = mat_dev(h);
  k_phys_cells<<<grid_for(n, 256), 256, 0, h->stream>>>(
      h->geo.p, h->sigma.p, (h->lattice || h->inst_lattice) ? h->rec.p : nullptr, h->G, n, md,
      dt, h->mat_rf,
      first ? 1 : 0, h->em.p, h->src.p, h->cen.p, h->mat_e.p, P);
  k_fill_u64<<<1, 64, 0, h->stream>>>(reinterpret_cast<unsigned long long*>(h->d_phys),
                                      sizeof(PhysTotals) / 8);
  HIPCHK(h, hipGetLastError());
  if (h->n_inst) {
    const uint32_t K = h->n_inst;
    k_fill_f64<<<grid_for(K, 256), 256, 0, h->stream>>>(h->inst_exit.p, K, 0.0);
    k_inst_sums<<<K, kWave, 0, h->stream>>>(h->em.p, h->src.p, h->cen.p, h->mat_e.p,
                                            h->inst_first_dev.p, K, h->inst_rep.p, rec);
    HIPCHK(h, hipGetLastError());
    return BH_OK;
  }
  if (h->mat_ranks == 1) {
    // E_tot in parallel with a bound on the serial loop's value, a check that
    // no photon count depends on where in the bound it lies, and the serial
    // sums only when one does
    k_phys_partials<<<nb, kSumNT, 0, h->stream>>>(h->em.p, h->src.p, h->cen.p, h->mat_e.p, n,
                                                  h->sum_part.p);
    k_final_sums<7><<<1, kSumNT, 0, h->stream>>>(h->sum_part.p, nb, &h->d_phys->tot_emission);
    k_phys_etot<<<1, 1, 0, h->stream>>>(n, h->d_phys, rec);
    k_count_check<<<grid_for(n, 256), 256, 0, h->stream>>>(h->em.p, h->src.p, h->cen.p, n,
                                                           n_user, first ? 1 : 0, h->d_phys);
    k_phys_totals<<<1, kWave, 0, h->stream>>>(h->em.p, h->src.p, h->cen.p, h->mat_e.p, n, 1, 0,
                                              h->d_phys, rec, 1);
  } else {  // the reference's serial sums, then this rank's share of the cells
    k_phys_totals<<<1, kWave, 0, h->stream>>>(h->em.p, h->src.p, h->cen.p, h->mat_e.p, n,
                                              h->mat_ranks, 0, h->d_phys, rec, 0);
    k_phys_rank_sums<<<1, kWave, 0, h->stream>>>(h->em.p, h->src.p, h->cen.p, n, n_user,
                                                 h->mat_ranks, h->mat_rf, first ? 1 : 0,
                                                 h->d_phys, rec);
    k_phys_reassign<<<grid_for(n, 256), 256, 0, h->stream>>>(
        h->em.p, h->src.p, h->cen.p, n, n_user, h->mat_ranks, h->mat_rank, h->mat_rf,
        first ? 1 : 0, h->d_phys);
    k_phys_totals<<<1, kWave, 0, h->stream>>>(h->em.p, h->src.p, h->cen.p, h->mat_e.p, n,
                                              h->mat_ranks, 1, h->d_phys, rec, 0);
  }
  HIPCHK(h, hipGetLastError());
  return BH_OK;
}

// Mesh::update_temperature on the device and the step's report sums; with
// instances, each instance's census energy (photons C[0, *n_census)) and
// totals.
int resident_finish(bh_handle* h, double dt, const StepParams* P, uint32_t nb,
                    const PhotonSoA& census, uint64_t census_cap,
                    const unsigned long long* n_census) {
  const uint32_t n = h->n_cells;
  k_update_temperature<<<grid_for(n, 256), 256, 0, h->stream>>>(
      h->geo.p, mat_dev(h), h->abs_E.p, h->track_E.p, h->em.p, dt, n, h->tr_diag.p,
      h->mat_e.p, P);
  k_report_partials<<<nb, kSumNT, 0, h->stream>>>(h->abs_E.p, h->mat_e.p, n, h->sum_part.p);
  k_final_sums<2><<<1, kSumNT, 0, h->stream>>>(h->sum_part.p, nb, &h->d_phys->absorbed);
  if (h->n_inst) {
    const uint32_t K = h->n_inst;
    k_inst_census<<<std::min<unsigned>(grid_for(std::max<uint64_t>(census_cap, 1), 256), 4096),
                    256, 0, h->stream>>>(census.cell, census.E, census_cap, n_census, nullptr,
                                         h->cell_inst.p, h->inst_rep.p, 1);
    k_inst_report<<<K, kWave, 0, h->stream>>>(h->abs_E.p, h->mat_e.p, h->counts.p,
                                              h->inst_first_dev.p, h->inst_exit.p, K,
                                              h->inst_rep.p);
  }
  HIPCHK(h, hipGetLastError());
  return BH_OK;
}

// The per-step source records: census then new photons, per instance.
void resident_records(const bh_handle* h, const uint64_t* n_user, uint32_t step, uint32_t rank,
                      SourceInst* out) {
  const uint32_t K = h->n_inst ? h->n_inst : 1u;
  const uint64_t hi1 = h->ctr_hi;
  for (uint32_t k = 0; k < K; ++k) {
    const uint64_t nu = n_user[k];
    const uint64_t rank_off = nu * (uint64_t)rank;
    const uint64_t hi = h->n_inst ? h->inst_hi[k] : hi1;
    const uint32_t first = h->n_inst ? h->inst_first[k] : 0u;
    out[k] = SourceInst{nu, 0.0, rank_off, hi, first, 0};
    out[K + k] = SourceInst{nu, 0.0, 10000000000000ull * (uint64_t)step + rank_off, hi, first, 0};
  }
}

void fill_report(const bh_handle* h, const PhysTotals& pt, const bh_step_stats& st,
                 uint64_t n_transported, uint64_t n_new, bool first, bh_resident_report* rep) {
  if (!rep) return;
  std::memset(rep, 0, sizeof *rep);
  rep->total_E = pt.global_E;
  rep->emission_E = pt.tot_emission;
  rep->source_E = pt.tot_source;
  rep->pre_census_E = first ? pt.pre_census : 0.0;
  rep->pre_mat_E = pt.pre_mat;
  rep->post_mat_E = pt.post_mat;
  rep->absorbed_E = pt.absorbed;
  rep->stats = st;
  rep->stats.n_transported = n_transported;
  rep->n_new = n_new;
  rep->etot_serial = (h->mat_ranks > 1 || pt.ambiguous) ? 1 : 0;
}

// Instances: the totals over instances and each instance's own.
void fill_instance_reports(const bh_handle* h, const InstReport* r, bh_resident_report* tot,
                           bh_instance_report* per) {
  const uint32_t K = h->n_inst;
  double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t n_new = 0;
  for (uint32_t k = 0; k < K; ++k) {
    if (per) std::memcpy(&per[k], &r[k], sizeof(bh_instance_report));
    s[0] += r[k].total_E;
    s[1] += r[k].emission_E;
    s[2] += r[k].source_E;
    s[3] += r[k].pre_census_E;
    s[5] += r[k].absorbed_E;
    s[7] += r[k].pre_mat_E;
    s[8] += r[k].post_mat_E;
    n_new += r[k].n_new;
  }
  if (tot) {
    tot->total_E = s[0];
    tot->emission_E = s[1];
    tot->source_E = s[2];
    tot->pre_census_E = s[3];
    tot->absorbed_E = s[5];
    tot->pre_mat_E = s[7];
    tot->post_mat_E = s[8];
    tot->n_new = n_new;
    tot->etot_serial = 0;
  }
}
static_assert(sizeof(InstReport) == sizeof(bh_instance_report),
              "InstReport mirrors bh_instance_report");

// One resident step, launch by launch (the first step, which makes the
// initial census, and the history algorithm): host synchronisations only
// where the photon stores may have to grow.
int resident_step_eager(bh_handle* h, const uint64_t* n_user, uint32_t step, uint32_t rank,
                        double dt, double next_dt, bh_resident_report* rep,
                        bh_instance_report* irep) {
  const bool first = step == 1;
  const uint32_t K = h->n_inst;
  const uint32_t nb = resident_blocks(h);
  TRY(ensure(h, h->sum_part, (size_t)8 * nb));
  SourceInst* rec = nullptr;
  if (K) {
    resident_records(h, n_user, step, rank, h->h_rec);
    HIPCHK(h, hipMemcpyAsync(h->inst_src.p, h->h_rec, 2 * (size_t)K * sizeof(SourceInst),
                             hipMemcpyHostToDevice, h->stream));
    rec = h->inst_src.p;
  } else {
    SourceInst* host2 = reinterpret_cast<SourceInst*>(static_cast<char*>(h->h_pinned) + 1024);
    resident_records(h, n_user, step, rank, host2);
    HIPCHK(h, hipMemcpyAsync(h->d_inst2, host2, 2 * sizeof(SourceInst), hipMemcpyHostToDevice,
                             h->stream));
    rec = h->d_inst2;
  }
  TRY(resident_physics(h, first, n_user[0], dt, nullptr, rec, nb));
  h->has_energy = true;
  h->has_census_E = first;
  uint64_t n_new = 0;
  const SourceInst unused{};
  TRY(run_source(h, unused, unused, rec, rec + (K ? K : 1), dt, &n_new,
                 (first && !K) ? h->sum_part.p : nullptr, nb));
  if (first && !K) {
    k_final_sums<1><<<1, kSumNT, 0, h->stream>>>(h->sum_part.p, nb, &h->d_phys->pre_census);
    HIPCHK(h, hipGetLastError());
  }
  char* hp = static_cast<char*>(h->h_pinned);
  PhysTotals pt;
  HIPCHK(h, hipMemcpyAsync(hp + 2048, h->d_phys, sizeof pt, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::memcpy(&pt, hp + 2048, sizeof pt);
  if (!K) {
    if (!(pt.global_E > 0.0)) return fail(h, BH_ERR_STATE, "no photon energy this step");
    h->src_neff_frac = pt.e_sq > 0.0 ? (pt.e_sum * pt.e_sum / pt.e_sq) / (double)h->n_cells : 1.0;
    if (pt.ambiguous || h->mat_ranks > 1) ++h->serial_etot_steps;
  } else if (h->n_cen) {  // each instance's census photons before transport
    k_inst_census<<<std::min<unsigned>(grid_for(h->n_cen, 256), 4096), 256, 0, h->stream>>>(
        h->W.soa.cell, h->W.soa.E, h->n_cen, nullptr, nullptr, h->cell_inst.p, h->inst_rep.p,
        0);
    HIPCHK(h, hipGetLastError());
  }
  const uint64_t n_transported = h->n_cen + n_new;
  bh_step_stats st;
  h->inst_exit_on = K > 0;
  const int rc = bh_transport(h, next_dt, nullptr, nullptr, &st);
  h->inst_exit_on = false;
  TRY(rc);
  // after the transport the new census is W[0, n_cen)
  TRY(resident_finish(h, dt, nullptr, nb, h->W.soa, h->n_cen, nullptr));
  HIPCHK(h, hipMemcpyAsync(hp + 2048, h->d_phys, sizeof pt, hipMemcpyDeviceToHost, h->stream));
  if (K)
    HIPCHK(h, hipMemcpyAsync(h->h_inst_rep, h->inst_rep.p, (size_t)K * sizeof(InstReport),
                             hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::memcpy(&pt, hp + 2048, sizeof pt);
  fill_report(h, pt, st, n_transported, n_new, first, rep);
  if (K) fill_instance_reports(h, h->h_inst_rep, rep, irep);
  return BH_OK;
}

// Everything a captured resident step bakes in.
std::vector<uintptr_t> resident_graph_key(const bh_handle* h, uint64_t n_user_total) {
  std::vector<uintptr_t> key = step_graph_key(h);
  auto u = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
  const uintptr_t more[] = {u(h->mat.p),      u(h->mat_e.p),       u(h->tr_diag.p),
                            u(h->d_phys),     u(h->d_inst2),       u(h->cen.p),
                            u(h->inst_src.p), u(h->inst_rep.p),    u(h->inst_exit.p),
                            u(h->inst_first_dev.p), u(h->cell_inst.p), u(h->h_rec),
                            u(h->h_inst_rep), (uintptr_t)h->n_inst, (uintptr_t)h->mat_ranks,
                            (uintptr_t)h->mat_rank, (uintptr_t)n_user_total};
  key.insert(key.end(), std::begin(more), std::end(more));
  uint64_t rf_bits;
  std::memcpy(&rf_bits, &h->mat_rf, sizeof rf_bits);
  key.push_back((uintptr_t)rf_bits);
  return key;
}

// Records a resident step after the first: the step's values come in from
// pinned memory, every count stays on the device, and one copy at the end
// takes the report out.  cap bounds the photons of the step.
int record_resident(bh_handle* h, uint64_t cap, uint64_t n_user0, bool gen) {
  const uint32_t nc = h->n_cells, K = h->n_inst;
  const uint32_t nb = resident_blocks(h);
  StepParams* P = h->d_params;
  HIPCHK(h, hipMemcpyAsync(P, h->h_params, sizeof(StepParams), hipMemcpyHostToDevice,
                           h->stream));
  SourceInst* rec = K ? h->inst_src.p : h->d_inst2;
  const size_t n_rec = 2 * (size_t)(K ? K : 1);
  HIPCHK(h, hipMemcpyAsync(rec, h->h_rec, n_rec * sizeof(SourceInst), hipMemcpyHostToDevice,
                           h->stream));
  TRY(resident_physics(h, false, n_user0, 0.0, P, rec, nb));
  // make_photons: counts, offsets, the step's counts into P, then the photons
  const SourceInst unused{};
  const uint32_t* cinst = K ? h->cell_inst.p : nullptr;
  SourceInst* inew = rec + (K ? K : 1);
  k_source_counts<<<grid_for(nc, 256), 256, 0, h->stream>>>(h->em.p, h->src.p, nc, unused, inew,
                                                             cinst, 0, h->counts.p, h->src_cell.p);
  TRY(scan_u64(h, h->counts.p, nc, h->soff.p, 1));
  k_step_counts<<<1, 1, 0, h->stream>>>(P, h->d_total + 1);
  // the new photons: made inside the transport kernel, or written to W
  k_source_hints<<<std::min<unsigned>(grid_for(div_up(cap, kHintStride), 256), 1024), 256, 0,
                   h->stream>>>(h->soff.p, nc, cap, &P->n_new, h->src_hint.p);
  if (!gen)
    k_source_fill<<<(unsigned)h->n_cu * 16, 256, 0, h->stream>>>(
        dev_cells(h), h->src_cell.p, h->face.p, h->soff.p, h->src_hint.p, 0, unused, inew,
        0.0, 0, h->W.soa, 0, P, nullptr);
  HIPCHK(h, hipGetLastError());
  const SourceDesc sd{h->src_cell.p, h->face.p, h->soff.p, h->src_hint.p, inew,
                      unused,  0.0,      cap,       gen ? 1 : 0, 0};
  if (K) {  // each instance's census photons before transport
    k_inst_census<<<std::min<unsigned>(grid_for(std::max<uint64_t>(cap, 1), 256), 4096), 256, 0,
                    h->stream>>>(h->W.soa.cell, h->W.soa.E, cap, nullptr, &P->n_cen,
                                 h->cell_inst.p, h->inst_rep.p, 0);
    HIPCHK(h, hipGetLastError());
  }
  ZeroList zl;
  TRY(zero_tallies_and_counters(h, &zl));
  CensusOut co;
  TRY(census_prepare(h, cap, 0.0, co, &zl));
  co.inst_exit = K ? h->inst_exit.p : nullptr;
  TRY(launch_transport(h, h->W.soa, cap, P, co, handle_sinks(h, h->ctr_hi), sd));
  TRY(census_finish(h, cap, &P->n));
  TRY(resident_finish(h, 0.0, P, nb, h->C.soa, cap, &h->d_cacc->n_census));
  char* hp = static_cast<char*>(h->h_pinned);
  HIPCHK(h, hipMemcpyAsync(hp, h->d_ctot, sizeof(CensusPartial), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(hp + 256, h->d_ctrs, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(hp + 2048, h->d_phys, sizeof(PhysTotals), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(hp + 3072, P, sizeof(StepParams), hipMemcpyDeviceToHost,
                           h->stream));
  if (K)
    HIPCHK(h, hipMemcpyAsync(h->h_inst_rep, h->inst_rep.p, (size_t)K * sizeof(InstReport),
                             hipMemcpyDeviceToHost, h->stream));
  return BH_OK;
}

// A resident step after the first as one HIP graph launch: no host round
// trip between its phases.  The photon stores are sized beforehand for the
// most photons the step can make (the census plus n_user plus two per cell:
// every cell's count is below n_user E_c / E_tot + 1 per source kind).
int resident_step_graph(bh_handle* h, const uint64_t* n_user, uint32_t step, uint32_t rank,
                        double dt, double next_dt, bh_resident_report* rep,
                        bh_instance_report* irep) {
  const uint32_t nc = h->n_cells, K = h->n_inst;
  uint64_t n_user_total = 0;
  for (uint32_t k = 0; k < (K ? K : 1u); ++k) n_user_total += n_user[k];
  const uint64_t bound = h->n_cen + n_user_total + 2 * (uint64_t)nc;
  if (bound >= (uint64_t)kHole)
    return fail(h, BH_ERR_CAPACITY, "up to %llu photons in one step: at most %u are supported",
                (unsigned long long)bound, kHole - 1u);
  const uint64_t slack = census_sizes(h, 0).slots;
  TRY(grow_store(h, h->W, bound + slack, h->n_cen));
  TRY(grow_store(h, h->C, bound + slack, 0));
  const uint64_t cap = std::min(h->W.cap, h->C.cap) - slack;
  const CensusSizes z = census_sizes(h, cap);
  const uint64_t pc = photon_cap(h, cap);
  const uint32_t nb = resident_blocks(h);
  TRY(ensure(h, h->counts, nc));
  TRY(ensure(h, h->soff, nc));
  TRY(ensure(h, h->src_hint, div_up(std::max<uint64_t>(cap, 1), kHintStride) + 1));
  TRY(ensure_marks(h, pc));
  TRY(ensure(h, h->hrec, z.rec_cap));
  TRY(ensure(h, h->hole_len, z.rec_cap));
  TRY(ensure(h, h->hole_off, z.rec_cap));
  TRY(ensure(h, h->tail_live, z.hole_cap));
  TRY(ensure(h, h->tail_rank, z.hole_cap));
  TRY(ensure(h, h->sum_part, (size_t)8 * std::max<uint32_t>(nb, kSumMaxBlocks)));
  const uint64_t scan_max = std::max<uint64_t>(std::max<uint64_t>(nc, pc),
                                               std::max(z.hole_cap, z.rec_cap));
  TRY(ensure(h, h->tiles, std::max<uint64_t>(div_up(scan_max, kScanTile), 1)));
  TRY(ensure(h, h->tile_E, std::max<uint64_t>(div_up(scan_max, kScanTile), 1)));
  if (h->opt.precision == BH_PRECISION_FP32_TRACK) {
    const uint64_t r = std::max(h->W.cap, h->C.cap);
    TRY(ensure(h, h->rep_idx, r));
    TRY(ensure(h, h->rep_first, r));
    TRY(ensure(h, h->rep_E, r));
  }
  if (!K) {  // one problem: the two source records live in pinned memory too
    if (h->h_inst_cap < 1) {
      HIPCHK(h, hipHostMalloc(&h->h_rec, 2 * sizeof(SourceInst), hipHostMallocDefault));
      HIPCHK(h, hipHostMalloc(&h->h_inst_rep, sizeof(InstReport), hipHostMallocDefault));
      h->h_inst_cap = 1;
    }
  }
  resident_records(h, n_user, step, rank, h->h_rec);
  *h->h_params = StepParams{h->n_cen, 0, 0, kC * next_dt, dt, SourceInst{}, cap, 0u, 0u};
  bh_handle::StepGraph& g = h->resident_graphs[h->W.slab < h->C.slab ? 0 : 1];
  // (the step's n_new is not known here: its n_user stands in for it)
  const bool gen = source_in_transport(h, n_user_total);
  std::vector<uintptr_t> key = resident_graph_key(h, n_user_total);
  key.push_back((uintptr_t)gen);
  if (!g.exec || g.key != key) {
    if (g.exec) HIPCHK(h, hipGraphExecDestroy(g.exec));
    if (g.graph) HIPCHK(h, hipGraphDestroy(g.graph));
    g.exec = nullptr;
    g.graph = nullptr;
    g.key.clear();
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_resident(h, cap, n_user[0], gen);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc != BH_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess)
      return fail(h, BH_ERR_DEVICE, "capturing the resident step failed: %s",
                  hipGetErrorString(e));
    g.graph = graph;
    HIPCHK(h, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    g.key = key;
  }
  HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
  HIPCHK(h, hipGraphLaunch(g.exec, h->stream));
  HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const char* hp = static_cast<const char*>(h->h_pinned);
  CensusPartial tot;
  DevCounters ctrs;
  PhysTotals pt;
  StepParams Pd;
  std::memcpy(&tot, hp, sizeof tot);
  std::memcpy(&ctrs, hp + 256, sizeof ctrs);
  std::memcpy(&pt, hp + 2048, sizeof pt);
  std::memcpy(&Pd, hp + 3072, sizeof Pd);
  if (ctrs.bad_args || tot.overflow || Pd.overflow) {
    h->n_cen = 0;
    h->n_new = 0;
    return fail(h, BH_ERR_DEVICE,
                ctrs.bad_args ? "transport kernel argument check failed; nothing transported"
                              : "resident step: photon counts exceeded the stores' capacity");
  }
  if (!K && !(pt.global_E > 0.0)) return fail(h, BH_ERR_STATE, "no photon energy this step");
  std::swap(h->W, h->C);
  h->n_cen = tot.n_census;
  h->n_new = 0;
  h->last_source_ms = 0.0;
  h->src_ms_pending = false;
  if (!K && (pt.ambiguous || h->mat_ranks > 1)) ++h->serial_etot_steps;
  bh_step_stats st;
  std::memset(&st, 0, sizeof st);
  st.census_E = tot.census_E;
  st.exit_E = tot.exit_E;
  st.n_transported = Pd.n;
  st.n_census = tot.n_census;
  st.n_exit = tot.n_exit;
  st.n_killed = tot.n_killed;
  st.n_segments = ctrs.segments;
  st.n_flushes = ctrs.flushes;
  st.n_replayed = ctrs.replays;
  st.lattice = h->last_lattice;
  st.tally = h->last_tally;
  st.n_allocs = take_allocs(h);
  st.transport_ms = elapsed_ms(h->ev[0], h->ev[1]);  // the whole graph
  fill_report(h, pt, st, Pd.n, Pd.n_new, false, rep);
  if (K) fill_instance_reports(h, h->h_inst_rep, rep, irep);
  if (rep) rep->graph = 1;
  if (ctrs.stuck)
    return fail(h, BH_ERR_DEVICE, "%llu photon(s) exceeded %u segments and were stopped",
                ctrs.stuck, kMaxSegments);
  if (tot.n_other)
    return fail(h, BH_ERR_STATE, "%llu photon(s) ended without EXIT/CENSUS/KILLED",
                tot.n_other);
  return BH_OK;
}

int resident_checks(bh_handle* h, uint32_t step, double dt) {
  if (!h->has_mesh || !h->has_material)
    return fail(h, BH_ERR_STATE, "a resident step needs bh_set_mesh and bh_set_material");
  if (step == 0) return fail(h, BH_ERR_ARG, "step counts from 1");
  if (!(dt > 0.0)) return fail(h, BH_ERR_ARG, "dt must be > 0");
  return BH_OK;
}

// Steps after the first run as one graph (persistent algorithm).
bool resident_uses_graph(const bh_handle* h, uint32_t step) {
  return step > 1 && h->opt.algorithm == BH_ALGO_PERSISTENT && !h->has_census_E;
}

}  // namespace

extern "C" {

int bh_step_resident(bh_handle* h, uint64_t n_user_photons, uint32_t seed, uint32_t step,
                     uint32_t rank, double dt, double next_dt, bh_resident_report* rep) {
  if (!h || !h->stream) return BH_ERR_ARG;
  TRY(resident_checks(h, step, dt));
  if (h->n_inst)
    return fail(h, BH_ERR_STATE, "the mesh holds %u instances: use bh_step_resident_instances",
                h->n_inst);
  if (n_user_photons == 0 || n_user_photons > 0xFFFFFFFFull)
    return fail(h, BH_ERR_ARG, "n_user_photons must be in [1, 2^32)");
  HIPCHK(h, hipSetDevice(h->device));
  // the step's physics fills every group alike (Cell::set_op_a / set_op_s)
  h->uniform_sigma_a = true;
  h->gray_sigma_a = h->opt.tally_mode != 3;
  const uint64_t ctr_hi = (uint64_t)seed << 32;
  if (h->n_cen && step > 1 && ctr_hi != h->ctr_hi)
    return fail(h, BH_ERR_ARG, "seed differs from the resident census photons' seed");
  h->ctr_hi = ctr_hi;
  if (resident_uses_graph(h, step))
    return resident_step_graph(h, &n_user_photons, step, rank, dt, next_dt, rep, nullptr);
  return resident_step_eager(h, &n_user_photons, step, rank, dt, next_dt, rep, nullptr);
}

int bh_step_resident_instances(bh_handle* h, const uint64_t* n_user_photons, uint32_t step,
                               uint32_t rank, double dt, double next_dt,
                               bh_resident_report* total, bh_instance_report* per_instance) {
  if (!h || !h->stream || !n_user_photons) return BH_ERR_ARG;
  TRY(resident_checks(h, step, dt));
  if (!h->n_inst) return fail(h, BH_ERR_STATE, "no instances: call bh_set_instances");
  for (uint32_t k = 0; k < h->n_inst; ++k)
    if (n_user_photons[k] == 0 || n_user_photons[k] > 0xFFFFFFFFull)
      return fail(h, BH_ERR_ARG, "instance %u: n_user_photons must be in [1, 2^32)", k);
  HIPCHK(h, hipSetDevice(h->device));
  h->uniform_sigma_a = true;
  h->gray_sigma_a = h->opt.tally_mode != 3;
  if (resident_uses_graph(h, step))
    return resident_step_graph(h, n_user_photons, step, rank, dt, next_dt, total, per_instance);
  return resident_step_eager(h, n_user_photons, step, rank, dt, next_dt, total, per_instance);
}

int bh_get_temperature(bh_handle* h, double* T_e, double* T_r) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_material) return fail(h, BH_ERR_STATE, "no material (bh_set_material)");
  const size_t n = h->n_cells;
  if (T_e)
    HIPCHK(h, hipMemcpyAsync(T_e, h->mat.p + 6 * n, n * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  if (T_r)
    HIPCHK(h, hipMemcpyAsync(T_r, h->tr_diag.p, n * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return BH_OK;
}

int bh_get_physics(bh_handle* h, double* sigma_a, double* sigma_s, double* fleck,
                   double* emission_E, double* source_E) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream) return BH_ERR_ARG;
  if (!h->has_mesh) return fail(h, BH_ERR_STATE, "no mesh");
  const size_t n = h->n_cells, G = h->G;
  std::vector<double2> sig;
  std::vector<CellGeo> geo;
  if (sigma_a || sigma_s) {
    sig.resize(n * G);
    HIPCHK(h, hipMemcpyAsync(sig.data(), h->sigma.p, n * G * sizeof(double2),
                             hipMemcpyDeviceToHost, h->stream));
  }
  if (fleck) {
    geo.resize(n);
    HIPCHK(h, hipMemcpyAsync(geo.data(), h->geo.p, n * sizeof(CellGeo), hipMemcpyDeviceToHost,
                             h->stream));
  }
  if (emission_E && h->em.p)
    HIPCHK(h, hipMemcpyAsync(emission_E, h->em.p, n * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  if (source_E && h->src.p)
    HIPCHK(h, hipMemcpyAsync(source_E, h->src.p, n * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < sig.size(); ++k) {
    if (sigma_a) sigma_a[k] = sig[k].x;
    if (sigma_s) sigma_s[k] = sig[k].y;
  }
  for (size_t c = 0; c < geo.size(); ++c) fleck[c] = geo[c].fleck;
  return BH_OK;
}

int bh_rng_draws(bh_handle* h, uint32_t seed, uint64_t stream, uint64_t ctr0, uint32_t n,
                 double* out) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || (n && !out)) return BH_ERR_ARG;
  if (n == 0) return BH_OK;
  if (ctr0 + n < ctr0)
    return fail(h, BH_ERR_ARG, "counter range wraps the 64-bit low word");
  HIPCHK(h, hipSetDevice(h->device));
  double* d = nullptr;
  HIPCHK(h, hipMalloc(&d, n * sizeof(double)));
  k_rng_draws<<<grid_for(n, 256), 256, 0, h->stream>>>((uint64_t)seed << 32, stream, ctr0, n,
                                                        d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(h, BH_ERR_DEVICE, "bh_rng_draws: %s", hipGetErrorString(e));
  return BH_OK;
}

int bh_copy_probe(bh_handle* h, uint64_t n_bytes, int reps, double* gbps) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || !gbps || reps <= 0 || n_bytes < 16) return BH_ERR_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  const uint64_t n = n_bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  HIPCHK(h, hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(h, BH_ERR_ALLOC, "bh_copy_probe: allocation failed");
  }
  const unsigned grid = (unsigned)std::min<uint64_t>(div_up(n, 256), (uint64_t)h->n_cu * 16);
  hipError_t e = hipMemsetAsync(a, 0, n * 16, h->stream);
  if (e == hipSuccess) {
    k_copy_probe<<<grid, 256, 0, h->stream>>>(a, b, n);  // warm-up
    e = hipEventRecord(h->ev[4], h->stream);
  }
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    k_copy_probe<<<grid, 256, 0, h->stream>>>(a, b, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(h->ev[5], h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  const double ms = elapsed_ms(h->ev[4], h->ev[5]) / reps;
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(h, BH_ERR_DEVICE, "bh_copy_probe: %s", hipGetErrorString(e));
  *gbps = ms > 0 ? 2.0 * (double)(n * 16) / (ms * 1e-3) / 1e9 : 0.0;
  return BH_OK;
}

int bh_threefry(bh_handle* h, const uint64_t* in, uint64_t n, uint64_t* out) {
  SetupAllocs setup_allocs(h);
  if (!h || !h->stream || (n && (!in || !out))) return BH_ERR_ARG;
  if (n == 0) return BH_OK;
  HIPCHK(h, hipSetDevice(h->device));
  uint64_t *di = nullptr, *dout = nullptr;
  HIPCHK(h, hipMalloc(&di, n * 4 * sizeof(uint64_t)));
  if (hipMalloc(&dout, n * 2 * sizeof(uint64_t)) != hipSuccess) {
    (void)hipFree(di);
    return fail(h, BH_ERR_ALLOC, "bh_threefry: allocation failed");
  }
  hipError_t e = hipMemcpyAsync(di, in, n * 4 * sizeof(uint64_t), hipMemcpyHostToDevice,
                                h->stream);
  if (e == hipSuccess) {
    k_threefry<<<grid_for(n, 256), 256, 0, h->stream>>>(di, n, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, dout, n * 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(di);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(h, BH_ERR_DEVICE, "bh_threefry: %s", hipGetErrorString(e));
  return BH_OK;
}

}  // extern "C"
