// bh_device.h -- device primitives of the MI355X IMC engine: counter-based
// RNG, source samplers, cell geometry.  Header-only, gfx950 device code.
//
// Floating point must replay the reference's IEEE double arithmetic, so the
// translation units that include this header are compiled with
// -ffp-contract=off (no fused multiply-add is formed from a*b+c).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bh {

// Physical constants (src/constants.h:16-24).
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kC = 299.792458;  // cm / shake
constexpr double kA = 0.01372;     // GJ / cm^3 / keV^4
constexpr double kCutoff = 0.01;   // E/E0 below which a photon is killed

enum : int { kReflect = 0, kVacuum = 1, kElement = 2, kSource = 3, kProcessor = 4 };
enum : uint8_t { kExit = 0, kPass = 1, kCensus = 2, kScatter = 3, kKilled = 4, kBound = 5,
                 kReplay = 254, kStuck = 255 };
// kReplay: internal to one transport call -- an FP32-tracked history whose
// next decision was too close to call; it is replayed in FP64 before the
// call returns, so the descriptor never reaches the census.

// A photon whose history exceeds this many segments is stopped and flagged
// (kStuck) so that no wave can spin forever on malformed input.
constexpr uint32_t kMaxSegments = 1u << 26;

// ---------------------------------------------------------------------------
// Threefry2x64-20 (Salmon et al., SC'11; vendored by the reference as
// src/random123/threefry.h:190-297).  Rotation table R_64x2 and the Skein key
// schedule parity are the published constants.  A 64-bit word lives in two
// VGPRs; a rotate by a constant N is two v_alignbit_b32 funnel shifts (none
// for N = 32, a register swap), an add is one v_lshl_add_u64.
// ---------------------------------------------------------------------------
template <unsigned N>
__device__ __forceinline__ uint64_t rotl64c(uint64_t x) {
  static_assert(N > 0 && N < 64, "rotation");
  const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  uint32_t nh, nl;
  if constexpr (N == 32) {
    nh = lo;
    nl = hi;
  } else if constexpr (N < 32) {  // alignbit(a, b, s) = ({a, b} >> s)[31:0]
    nh = __builtin_amdgcn_alignbit(hi, lo, 32u - N);
    nl = __builtin_amdgcn_alignbit(lo, hi, 32u - N);
  } else {
    nh = __builtin_amdgcn_alignbit(lo, hi, 64u - N);
    nl = __builtin_amdgcn_alignbit(hi, lo, 64u - N);
  }
  return ((uint64_t)nh << 32) | nl;
}

struct Threefry {
  static constexpr uint64_t kParity = 0x1BD11BDAA9FC1A22ULL;

  template <unsigned N>
  __device__ __forceinline__ static void mix(uint64_t& x0, uint64_t& x1) {
    x0 += x1;
    x1 = rotl64c<N>(x1);
    x1 ^= x0;
  }
  // four rounds with rotation constants R[4q .. 4q+3] then key injection s
  template <unsigned A, unsigned B, unsigned C, unsigned D>
  __device__ __forceinline__ static void quad(uint64_t& x0, uint64_t& x1, uint64_t ka,
                                              uint64_t kb, uint64_t s) {
    mix<A>(x0, x1);
    mix<B>(x0, x1);
    mix<C>(x0, x1);
    mix<D>(x0, x1);
    x0 += ka;
    x1 += kb + s;
  }

  // Full two-word output of Threefry2x64 with 20 rounds.
  __device__ __forceinline__ static void block(uint64_t c0, uint64_t c1, uint64_t k0,
                                               uint64_t k1, uint64_t& o0, uint64_t& o1) {
    const uint64_t k2 = kParity ^ k0 ^ k1;
    uint64_t x0 = c0 + k0;
    uint64_t x1 = c1 + k1;
    // R_64x2 = {16, 42, 12, 31, 16, 32, 24, 21}; injection s uses
    // ks[s % 3] and ks[(s + 1) % 3] + s
    quad<16, 42, 12, 31>(x0, x1, k1, k2, 1);
    quad<16, 32, 24, 21>(x0, x1, k2, k0, 2);
    quad<16, 42, 12, 31>(x0, x1, k0, k1, 3);
    quad<16, 32, 24, 21>(x0, x1, k1, k2, 4);
    quad<16, 42, 12, 31>(x0, x1, k2, k0, 5);
    o0 = x0;
    o1 = x1;
  }
};

// r123::u01fixedpt<double, uint64_t> (src/RNG.h:202-238): odd multiples of
// 2^-53 in (0, 1).  The 53-bit integer converts to double exactly.
__device__ __forceinline__ double u01(uint64_t x) {
  return (double)((x >> 11) | 1ull) * 0x1.0p-53;
}

// Per-photon RNG state as the device keeps it: the low counter word and the
// stream (key low word).  The counter's high word is seed << 32 for every
// photon of a run and the key's high word is 0 (src/RNG.h:318-330), so both
// live in kernel arguments.  A carry out of the low word would take 2^64
// draws and is not represented.
struct Rng {
  uint64_t ctr;
  uint64_t stream;
  uint64_t ctr_hi;

  __device__ __forceinline__ double next() {  // _ran, src/RNG.h:262-285
    uint64_t o0, o1;
    Threefry::block(ctr, ctr_hi, stream, 0ull, o0, o1);
    ctr += 1u;
    return u01(o0);
  }
};

// ---------------------------------------------------------------------------
// Samplers (src/sampling_functions.h).  Operation order follows the
// reference exactly; see the oracle for the same formulas on the host.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pos_in_cell(const double* nd, Rng& rng, double& x, double& y,
                                            double& z) {  // :22-30
  x = nd[0] + rng.next() * (nd[1] - nd[0]);
  y = nd[2] + rng.next() * (nd[3] - nd[2]);
  z = nd[4] + rng.next() * (nd[5] - nd[4]);
}

__device__ __forceinline__ void pos_on_face(const double* nd, int face, Rng& rng, double& x,
                                            double& y, double& z) {  // :32-53
  if (face == 0 || face == 1) {
    x = (face == 0) ? nd[0] : nd[1];
    y = nd[2] + rng.next() * (nd[3] - nd[2]);
    z = nd[4] + rng.next() * (nd[5] - nd[4]);
  } else if (face == 2 || face == 3) {
    x = nd[0] + rng.next() * (nd[1] - nd[0]);
    y = (face == 2) ? nd[2] : nd[3];
    z = nd[4] + rng.next() * (nd[5] - nd[4]);
  } else {
    x = nd[0] + rng.next() * (nd[1] - nd[0]);
    y = nd[2] + rng.next() * (nd[3] - nd[2]);
    z = (face == 4) ? nd[4] : nd[5];
  }
}

// get_uniform_angle (:56-71)
__device__ __forceinline__ void uniform_angle(Rng& rng, double& u, double& v, double& w) {
  const double mu = rng.next() * 2.0 - 1.0;
  const double phi = rng.next() * 2.0 * kPi;
  const double sin_theta = sqrt(1.0 - mu * mu);
  u = sin_theta * cos(phi);
  v = sin_theta * sin(phi);
  w = mu;
}

// get_source_angle_on_face (:94-122); inward sign on x faces only, as the
// reference does.
__device__ __forceinline__ void source_angle(int face, Rng& rng, double& u, double& v,
                                             double& w) {
  const double theta = acos(sqrt(rng.next()));
  const double phi = rng.next() * 2.0 * kPi;
  const double sign = (face % 2) ? -1.0 : 1.0;
  const double ct = cos(theta), st = sin(theta);
  if (face == 0 || face == 1) {
    u = ct * sign;
    v = st * sin(phi);
    w = st * cos(phi);
  } else if (face == 2 || face == 3) {
    u = st * sin(phi);
    v = ct;
    w = st * cos(phi);
  } else {
    u = st * cos(phi);
    v = st * sin(phi);
    w = ct;
  }
}

// sample_emission_group(RNG&, const Cell&) (:125-138): subtraction walk over
// the absorption CDF normalised by sigma_a[0]*G.  sa points at the cell's
// group-contiguous (sigma_a, sigma_s) pairs.  A walk that round-off carries
// past the last group returns the last group.
__device__ __forceinline__ uint32_t emission_group(const double2* sa, uint32_t G, Rng& rng) {
  double cdf = rng.next();
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
  return (uint32_t)g;
}

// ---------------------------------------------------------------------------
// Cell geometry record: one 64-byte line per cell holding everything the
// tracking loop reads every segment except the opacities.
// ---------------------------------------------------------------------------
struct alignas(64) CellGeo {
  double nodes[6];  // x_lo x_hi y_lo y_hi z_lo z_hi
  double fleck;
  uint32_t bc_bits;  // 3 bits per face, face f at bits [3f, 3f+3)
  uint32_t source_face_plus1;  // first SOURCE face + 1, 0 when none
};
static_assert(sizeof(CellGeo) == 64, "CellGeo must be one 64-byte record");

// Lattice meshes (a rectilinear grid numbered x fastest, opacities equal in
// every group -- every mesh Branson builds from a deck): the node
// coordinates live in three small axis arrays and a cell's per-segment data
// in one 32-byte record, half a cache line, so the cell data of a 64^3 mesh
// fits in one XCD's L2.
struct alignas(32) CellRec {
  double sigma_a;
  double sigma_s;
  double fleck;
  uint32_t bc_bits;  // as CellGeo
  uint32_t pad;
};
static_assert(sizeof(CellRec) == 32, "CellRec must be a 32-byte record");

__device__ __forceinline__ int face_bc(uint32_t bits, uint32_t face) {
  return (int)((bits >> (3u * face)) & 7u);
}

}  // namespace bh
