/*
 * branson_oracle.c -- CPU restatement of Branson's replicated-mode IMC step.
 *
 * TEST INFRASTRUCTURE ONLY (see branson_oracle.h): the correctness oracle and
 * the timed CPU baseline.  Never part of the shipped library.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Contraction
 * must stay off so every a*b+c rounds twice, as the reference's x86-64 build
 * without -march does.
 */
#include "branson_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/constants.h:16-24 */
#define ORC_PI 3.1415926535897932384626433832795
#define ORC_C 299.792458
#define ORC_A 0.01372
#define ORC_CUTOFF 0.01

/* ======================================================================== */
/* Threefry2x64-20.  Algorithm of Salmon et al., "Parallel random numbers:   */
/* as easy as 1, 2, 3" (SC'11): Threefish-style add/rotate/xor rounds with a */
/* key schedule injected every 4 rounds.  The reference vendors Random123   */
/* (src/random123/threefry.h:190-297, rotations :86-93, parity :171).       */
/* ======================================================================== */
static const unsigned orc_tf_rot[8] = {16, 42, 12, 31, 16, 32, 24, 21};
#define ORC_TF_PARITY 0x1BD11BDAA9FC1A22ULL

static inline uint64_t orc_rotl(uint64_t x, unsigned n) {
  return (x << n) | (x >> (64u - n));
}

void orc_threefry2x64_20(const uint64_t ctr[2], const uint64_t key[2],
                         uint64_t out[2]) {
  uint64_t ks[3];
  ks[0] = key[0];
  ks[1] = key[1];
  ks[2] = ORC_TF_PARITY ^ key[0] ^ key[1];
  uint64_t x0 = ctr[0] + ks[0];
  uint64_t x1 = ctr[1] + ks[1];
  for (unsigned r = 0; r < 20; ++r) {
    x0 += x1;
    x1 = orc_rotl(x1, orc_tf_rot[r & 7u]);
    x1 ^= x0;
    if ((r & 3u) == 3u) {
      const unsigned s = (r + 1u) >> 2; /* injection number 1..5 */
      x0 += ks[s % 3u];
      x1 += ks[(s + 1u) % 3u] + (uint64_t)s;
    }
  }
  out[0] = x0;
  out[1] = x1;
}

/* r123::u01fixedpt<double, uint64_t> (src/RNG.h:202-238): the top 53 bits,
 * forced odd, times 2^-53.  Output in (0,1), never 0, 0.5 or 1. */
double orc_u01(uint64_t x) {
  return (double)((x >> 11) | 1ULL) * 0x1.0p-53;
}

/* _ran (src/RNG.h:262-285): output word 0 of Threefry(ctr, key), then the
 * 128-bit counter is incremented (src/random123/array.h:133-165). */
static inline double orc_rng_next(uint64_t rng[4]) {
  const uint64_t ctr[2] = {rng[0], rng[1]};
  const uint64_t key[2] = {rng[2], rng[3]};
  uint64_t out[2];
  orc_threefry2x64_20(ctr, key, out);
  rng[0] += 1u;
  if (rng[0] == 0u) rng[1] += 1u;
  return orc_u01(out[0]);
}

/* RNG(seed, stream) (src/RNG.h:318-330) */
static inline void orc_rng_init(uint64_t rng[4], uint32_t seed,
                                uint64_t stream) {
  rng[0] = 0;
  rng[1] = (uint64_t)seed << 32;
  rng[2] = stream;
  rng[3] = 0;
}

void orc_rng_draws(uint32_t seed, uint64_t stream, uint64_t ctr0, uint32_t n,
                   double *out) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  rng[0] = ctr0;
  for (uint32_t i = 0; i < n; ++i) out[i] = orc_rng_next(rng);
}

/* ======================================================================== */
/* Samplers (src/sampling_functions.h)                                      */
/* ======================================================================== */

/* get_uniform_position_in_cell (:22-30) */
static inline void orc_pos_in_cell(const double *nd, uint64_t rng[4],
                                   double pos[3]) {
  pos[0] = nd[0] + orc_rng_next(rng) * (nd[1] - nd[0]);
  pos[1] = nd[2] + orc_rng_next(rng) * (nd[3] - nd[2]);
  pos[2] = nd[4] + orc_rng_next(rng) * (nd[5] - nd[4]);
}

/* get_uniform_position_on_face (:32-53) */
static inline void orc_pos_on_face(const double *nd, int face, uint64_t rng[4],
                                   double pos[3]) {
  if (face == 0 || face == 1) {
    pos[0] = (face == 0) ? nd[0] : nd[1];
    pos[1] = nd[2] + orc_rng_next(rng) * (nd[3] - nd[2]);
    pos[2] = nd[4] + orc_rng_next(rng) * (nd[5] - nd[4]);
  } else if (face == 2 || face == 3) {
    pos[0] = nd[0] + orc_rng_next(rng) * (nd[1] - nd[0]);
    pos[1] = (face == 2) ? nd[2] : nd[3];
    pos[2] = nd[4] + orc_rng_next(rng) * (nd[5] - nd[4]);
  } else {
    pos[0] = nd[0] + orc_rng_next(rng) * (nd[1] - nd[0]);
    pos[1] = nd[2] + orc_rng_next(rng) * (nd[3] - nd[2]);
    pos[2] = (face == 4) ? nd[4] : nd[5];
  }
}

/* get_uniform_angle (:56-71): mu = 2u-1, phi = 2 pi u */
static inline void orc_uniform_angle(uint64_t rng[4], double ang[3]) {
  const double mu = orc_rng_next(rng) * 2.0 - 1.0;
  const double phi = orc_rng_next(rng) * 2.0 * ORC_PI;
  const double sin_theta = sqrt(1.0 - mu * mu);
  ang[0] = sin_theta * cos(phi);
  ang[1] = sin_theta * sin(phi);
  ang[2] = mu;
}

/* get_source_angle_on_face (:94-122).  The reference applies the inward sign
 * only on x faces; y/z faces keep +cos(theta).  Restated as is. */
static inline void orc_source_angle(int face, uint64_t rng[4], double ang[3]) {
  const double theta = acos(sqrt(orc_rng_next(rng)));
  const double phi = orc_rng_next(rng) * 2.0 * ORC_PI;
  const double sign = (face % 2) ? -1.0 : 1.0;
  if (face == 0 || face == 1) {
    ang[0] = cos(theta) * sign;
    ang[1] = sin(theta) * sin(phi);
    ang[2] = sin(theta) * cos(phi);
  } else if (face == 2 || face == 3) {
    ang[0] = sin(theta) * sin(phi);
    ang[1] = cos(theta);
    ang[2] = sin(theta) * cos(phi);
  } else {
    ang[0] = sin(theta) * cos(phi);
    ang[1] = sin(theta) * sin(phi);
    ang[2] = cos(theta);
  }
}

/* sample_emission_group(RNG&, const Cell&) (:125-138): subtraction walk over
 * the normalised absorption CDF.  The reference reads past the last group
 * when round-off leaves cdf > 0 after G subtractions; that undefined read is
 * replaced here by the last group. */
static inline uint32_t orc_emission_group(const double *sa, uint32_t G,
                                          uint64_t rng[4]) {
  double cdf = orc_rng_next(rng);
  int g = -1;
  const double norm = 1.0 / (sa[0] * (double)G);
  while (cdf > 0) {
    g++;
    if ((uint32_t)g >= G) {
      g = (int)G - 1;
      break;
    }
    cdf -= sa[g] * norm;
  }
  return (uint32_t)g;
}

void orc_sample_position_in_cell(const double nodes[6], uint32_t seed,
                                 uint64_t stream, double pos[3],
                                 uint64_t *ctr_after) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  orc_pos_in_cell(nodes, rng, pos);
  *ctr_after = rng[0];
}

void orc_sample_position_on_face(const double nodes[6], int face,
                                 uint32_t seed, uint64_t stream,
                                 double pos[3], uint64_t *ctr_after) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  orc_pos_on_face(nodes, face, rng, pos);
  *ctr_after = rng[0];
}

void orc_sample_uniform_angle(uint32_t seed, uint64_t stream, double ang[3],
                              uint64_t *ctr_after) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  orc_uniform_angle(rng, ang);
  *ctr_after = rng[0];
}

void orc_sample_source_angle(int face, uint32_t seed, uint64_t stream,
                             double ang[3], uint64_t *ctr_after) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  orc_source_angle(face, rng, ang);
  *ctr_after = rng[0];
}

int orc_sample_emission_group(const double *sigma_a_cell, uint32_t n_groups,
                              uint32_t seed, uint64_t stream,
                              uint64_t *ctr_after) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  const uint32_t g = orc_emission_group(sigma_a_cell, n_groups, rng);
  *ctr_after = rng[0];
  return (int)g;
}

/* Cell::get_distance_to_boundary (src/cell.h:116-132) with sgn = (0 < v)
 * (src/cell.h:23-25): per axis only the face the direction points at; the
 * strict minimum from 1e16 sets the crossing face. */
static inline double orc_dist_boundary(const double *nd, const double pos[3],
                                       const double ang[3], uint32_t *sc) {
  double min_dist = 1.0e16;
  for (uint32_t i = 0; i < 3; ++i) {
    const uint32_t idx = 2u * i + (uint32_t)(0.0 < ang[i]);
    const double dist = (nd[idx] - pos[i]) / ang[i];
    if (dist < min_dist) {
      min_dist = dist;
      *sc = idx;
    }
  }
  return min_dist;
}

double orc_distance_to_boundary(const double nodes[6], const double pos[3],
                                const double ang[3], uint32_t *surface_cross) {
  return orc_dist_boundary(nodes, pos, ang, surface_cross);
}

/* ======================================================================== */
/* Source (src/source.h)                                                    */
/* ======================================================================== */

/* photons per cell: int(n_user * E / E_tot), at least one (src/source.h:
 * 229-244, 259-279).  n_user is uint64 converted to double first. */
static inline uint32_t orc_count(uint64_t n_user, double E, double total_E) {
  if (!(E > 0.0)) return 0;
  uint32_t n = (uint32_t)(int)((double)n_user * E / total_E);
  if (n == 0) n = 1;
  return n;
}

uint64_t orc_source_counts(uint32_t n_cells, const double *emission_E,
                           const double *source_E, uint64_t n_user,
                           double total_E, uint32_t *counts) {
  uint64_t total = 0;
  for (uint32_t c = 0; c < n_cells; ++c) {
    const uint32_t ne = orc_count(n_user, emission_E[c], total_E);
    const uint32_t ns = source_E ? orc_count(n_user, source_E[c], total_E) : 0;
    if (counts) {
      counts[2 * c] = ne;
      counts[2 * c + 1] = ns;
    }
    total += (uint64_t)ne + ns;
  }
  return total;
}

/* set_emission_photon (src/source.h:27-43): 7 draws */
static void orc_emission_photon(orc_photon *p, const orc_cells *cells,
                                uint32_t c, double E, double dt, uint32_t seed,
                                uint64_t stream) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  p->source_type = 2;
  orc_pos_in_cell(&cells->nodes[6 * (size_t)c], rng, p->pos);
  orc_uniform_angle(rng, p->angle);
  p->E = E;
  p->E0 = E;
  p->life_dx = orc_rng_next(rng) * ORC_C * dt;
  p->cell = c;
  p->group = (uint32_t)floor(orc_rng_next(rng) * (double)cells->n_groups);
  memcpy(p->rng, rng, sizeof rng);
  memset(p->descriptor, 0, 4);
  p->descriptor[0] = ORC_PASS;
}

/* set_boundary_source_photon (src/source.h:46-62): 6 draws */
static void orc_face_photon(orc_photon *p, const orc_cells *cells, uint32_t c,
                            double E, double dt, uint32_t seed,
                            uint64_t stream, int face) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  p->source_type = 1;
  orc_pos_on_face(&cells->nodes[6 * (size_t)c], face, rng, p->pos);
  orc_source_angle(face, rng, p->angle);
  p->E = E;
  p->E0 = E;
  p->life_dx = orc_rng_next(rng) * ORC_C * dt;
  p->cell = c;
  p->group = (uint32_t)floor(orc_rng_next(rng) * (double)cells->n_groups);
  memcpy(p->rng, rng, sizeof rng);
  memset(p->descriptor, 0, 4);
  p->descriptor[0] = ORC_PASS;
}

/* set_initial_census_photon (src/source.h:65-81): 6 draws, life = c*dt */
static void orc_census_photon(orc_photon *p, const orc_cells *cells,
                              uint32_t c, double E, double dt, uint32_t seed,
                              uint64_t stream) {
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  p->source_type = 0;
  orc_pos_in_cell(&cells->nodes[6 * (size_t)c], rng, p->pos);
  orc_uniform_angle(rng, p->angle);
  p->E = E;
  p->E0 = E;
  p->life_dx = ORC_C * dt;
  p->cell = c;
  p->group = (uint32_t)floor(orc_rng_next(rng) * (double)cells->n_groups);
  memcpy(p->rng, rng, sizeof rng);
  memset(p->descriptor, 0, 4);
  p->descriptor[0] = ORC_PASS;
}

/* make_photons (src/source.h:207-291).  Stream of the i-th new photon:
 * 1e13*cycle + n_user*rank + i (:221-222). */
uint64_t orc_make_photons(const orc_cells *cells, const double *emission_E,
                          const double *source_E, const int8_t *source_face,
                          uint64_t n_user, double total_E, uint32_t seed,
                          uint32_t cycle, uint32_t rank, double dt,
                          orc_photon *out, uint64_t cap) {
  const uint64_t cycle_off = 10000000000000ULL * (uint64_t)cycle;
  const uint64_t rank_off = n_user * (uint64_t)rank;
  uint64_t ith = 0;
  for (uint32_t c = 0; c < cells->n_cells; ++c) {
    const uint32_t ne = orc_count(n_user, emission_E[c], total_E);
    if (ne) {
      const double E = emission_E[c] / ne;
      for (uint32_t p = 0; p < ne; ++p) {
        if (ith < cap)
          orc_emission_photon(&out[ith], cells, c, E, dt, seed,
                              cycle_off + rank_off + ith);
        ith++;
      }
    }
    const uint32_t ns = orc_count(n_user, source_E[c], total_E);
    if (ns) {
      const double E = source_E[c] / ns;
      for (uint32_t p = 0; p < ns; ++p) {
        if (ith < cap)
          orc_face_photon(&out[ith], cells, c, E, dt, seed,
                          cycle_off + rank_off + ith, source_face[c]);
        ith++;
      }
    }
  }
  return ith;
}

/* make_initial_census_photons (src/source.h:134-180): streams n_user*rank+i */
uint64_t orc_make_initial_census(const orc_cells *cells,
                                 const double *census_E, uint64_t n_user,
                                 double total_E, uint32_t seed, uint32_t rank,
                                 double dt, orc_photon *out, uint64_t cap) {
  const uint64_t rank_off = n_user * (uint64_t)rank;
  uint64_t ith = 0;
  for (uint32_t c = 0; c < cells->n_cells; ++c) {
    const uint32_t n = orc_count(n_user, census_E[c], total_E);
    if (!n) continue;
    const double E = census_E[c] / n;
    for (uint32_t p = 0; p < n; ++p) {
      if (ith < cap)
        orc_census_photon(&out[ith], cells, c, E, dt, seed, rank_off + ith);
      ith++;
    }
  }
  return ith;
}

/* ======================================================================== */
/* Transport (src/history_based_transport.h:32-141)                         */
/* ======================================================================== */
static uint64_t orc_transport_photon(const orc_cells *cells, orc_photon *p,
                                     double *abs_E, double *track_E) {
  const uint32_t G = cells->n_groups;
  uint64_t *rng = p->rng;
  uint32_t sc = 0;
  uint32_t c = p->cell;
  double t_abs = 0.0, t_track = 0.0;
  uint64_t segs = 0;
  int active = 1;
  while (active) {
    ++segs;
    const size_t cg = (size_t)c * G + p->group;
    const double sigma_s = cells->sigma_s[cg];
    const double sigma_a = cells->sigma_a[cg];
    const double f = cells->fleck[c];
    const double total_sigma_s = (1.0 - f) * sigma_a + sigma_s;

    const double d_s =
        (total_sigma_s > 0.0) ? -log(orc_rng_next(rng)) / total_sigma_s
                              : 1.0e100;
    const double d_b =
        orc_dist_boundary(&cells->nodes[6 * (size_t)c], p->pos, p->angle, &sc);
    const double d_c = p->life_dx;
    /* std::min(a, b) == (b < a) ? b : a */
    const double m_bc = (d_c < d_b) ? d_c : d_b;
    const double d = (m_bc < d_s) ? m_bc : d_s;

    const double absorbed = p->E * (1.0 - exp(-sigma_a * f * d));
    t_abs += absorbed;
    t_track += absorbed / (sigma_a * f);
    p->E = p->E - absorbed;

    /* Photon::move (src/photon.h:107-114) */
    p->pos[0] += p->angle[0] * d;
    p->pos[1] += p->angle[1] * d;
    p->pos[2] += p->angle[2] * d;
    p->life_dx -= d;

    if (p->E / p->E0 < ORC_CUTOFF) { /* below_cutoff, src/photon.h:46-49 */
      t_abs += p->E;
      abs_E[c] += t_abs;
      track_E[c] += t_track;
      active = 0;
      p->descriptor[0] = ORC_KILLED;
    } else if (d == d_s) {
      orc_uniform_angle(rng, p->angle);
      if (orc_rng_next(rng) > (sigma_s / ((1.0 - f) * sigma_a + sigma_s)))
        p->group = orc_emission_group(&cells->sigma_a[(size_t)c * G], G, rng);
      p->descriptor[0] = ORC_SCATTER;
    } else if (d == d_b) {
      const int32_t bc = cells->bc[6 * (size_t)c + sc];
      if (bc == ORC_ELEMENT) {
        abs_E[c] += t_abs;
        track_E[c] += t_track;
        c = cells->next[6 * (size_t)c + sc];
        p->cell = c;
        p->descriptor[0] = ORC_BOUND;
        t_abs = 0.0;
        t_track = 0.0;
      } else if (bc == ORC_PROCESSOR) {
        active = 0;
        p->cell = cells->next[6 * (size_t)c + sc];
        p->descriptor[0] = ORC_PASS;
        abs_E[c] += t_abs;
        track_E[c] += t_track;
      } else if (bc == ORC_VACUUM || bc == ORC_SOURCE) {
        active = 0;
        p->descriptor[0] = ORC_EXIT;
        abs_E[c] += t_abs;
        track_E[c] += t_track;
      } else { /* REFLECT: Photon::reflect, src/photon.h:150-155 */
        p->angle[sc / 2] = -p->angle[sc / 2];
        p->descriptor[0] = ORC_BOUND;
      }
    } else if (d == d_c) {
      active = 0;
      p->descriptor[0] = ORC_CENSUS;
      abs_E[c] += t_abs;
      track_E[c] += t_track;
    }
  }
  return segs;
}

/* history_cpu_transport_photons (src/history_based_transport.h:279-308) */
uint64_t orc_transport(const orc_cells *cells, orc_photon *photons,
                       uint64_t n, double *abs_E, double *track_E,
                       int n_threads) {
  const uint32_t nc = cells->n_cells;
  uint64_t segs = 0;
  if (n_threads <= 1) {
    for (uint64_t i = 0; i < n; ++i)
      segs += orc_transport_photon(cells, &photons[i], abs_E, track_E);
    return segs;
  }
  double *ta = (double *)calloc((size_t)n_threads * nc, sizeof(double));
  double *tt = (double *)calloc((size_t)n_threads * nc, sizeof(double));
#pragma omp parallel num_threads(n_threads) reduction(+ : segs)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    double *my_a = ta + (size_t)t * nc;
    double *my_t = tt + (size_t)t * nc;
#pragma omp for schedule(guided)
    for (int64_t i = 0; i < (int64_t)n; ++i)
      segs += orc_transport_photon(cells, &photons[i], my_a, my_t);
  }
  /* merge in thread order (:298-302) */
  for (uint32_t c = 0; c < nc; ++c)
    for (int t = 0; t < n_threads; ++t) {
      abs_E[c] += ta[(size_t)t * nc + c];
      track_E[c] += tt[(size_t)t * nc + c];
    }
  free(ta);
  free(tt);
  return segs;
}

/* ======================================================================== */
/* Differential-operator (likelihood-ratio) tally, forward mode             */
/* ======================================================================== */
/* Derivative of the expected absorbed-energy tally along a direction of the
 * per-cell parameters: sigma_a(c) + t v_sa(c), fleck(c) + t v_f(c), t -> 0,
 * with the source photons held fixed.  It is the classic differential
 * operator estimator of Monte Carlo perturbation theory: every score s of a
 * history contributes
 *     s * S + ds/dt|path fixed,
 * where S sums d log p / dt over the random decisions made before (and
 * including) the one the score depends on, and ds/dt|path fixed is the
 * derivative through the implicit-capture weights.  The decisions that
 * depend on the parameters in a cell with sigma_eff = (1 - f) sa + ss:
 *   flight: collision at d, density sigma_eff exp(-sigma_eff d):
 *           d log p = dsigma_eff (1/sigma_eff - d); no collision within d
 *           (face or census first): d log p = -dsigma_eff d;
 *   re-emission test after a collision: keep the group with probability
 *           ss / sigma_eff, re-sample with (1 - f) sa / sigma_eff.
 * dsigma_eff = (1 - f) v_sa - sa v_f; the optical depth of a segment
 * tau = sa f d has dtau = (f v_sa + sa v_f) d.  Group-independent sigma_a
 * only: then the group walk draws each group with a probability that does
 * not depend on the parameters.  The cutoff test is a discontinuity of the
 * estimator and contributes nothing here.
 *
 * The photons are transported in place with exactly the events of
 * orc_transport (the same statements in the same order, plus the derivative
 * bookkeeping).  abs_E and dabs_E are accumulated into. */
static uint64_t orc_dos_photon(const orc_cells *cells, orc_photon *p,
                               const double *v_sa, const double *v_f,
                               double *abs_E, double *dabs_E) {
  const uint32_t G = cells->n_groups;
  uint64_t *rng = p->rng;
  uint32_t sc = 0;
  uint32_t c = p->cell;
  uint64_t segs = 0;
  double S = 0.0;   /* score function: sum of d log p / dt so far */
  double dlE = 0.0; /* d log(weight) / dt through implicit capture  */
  int active = 1;
  while (active) {
    ++segs;
    const size_t cg = (size_t)c * G + p->group;
    const double sigma_s = cells->sigma_s[cg];
    const double sigma_a = cells->sigma_a[cg];
    const double f = cells->fleck[c];
    const double total_sigma_s = (1.0 - f) * sigma_a + sigma_s;
    const double dsig = (1.0 - f) * v_sa[c] - sigma_a * v_f[c];

    const double d_s =
        (total_sigma_s > 0.0) ? -log(orc_rng_next(rng)) / total_sigma_s
                              : 1.0e100;
    const double d_b =
        orc_dist_boundary(&cells->nodes[6 * (size_t)c], p->pos, p->angle, &sc);
    const double d_c = p->life_dx;
    const double m_bc = (d_c < d_b) ? d_c : d_b;
    const double d = (m_bc < d_s) ? m_bc : d_s;

    if (total_sigma_s > 0.0)
      S += dsig * ((d == d_s) ? (1.0 / total_sigma_s - d) : -d);
    const double dtau = (f * v_sa[c] + sigma_a * v_f[c]) * d;

    const double absorbed = p->E * (1.0 - exp(-sigma_a * f * d));
    p->E = p->E - absorbed;
    abs_E[c] += absorbed;
    dabs_E[c] += absorbed * (S + dlE) + p->E * dtau;
    dlE -= dtau;

    p->pos[0] += p->angle[0] * d;
    p->pos[1] += p->angle[1] * d;
    p->pos[2] += p->angle[2] * d;
    p->life_dx -= d;

    if (p->E / p->E0 < ORC_CUTOFF) {
      abs_E[c] += p->E;
      dabs_E[c] += p->E * (S + dlE);
      active = 0;
      p->descriptor[0] = ORC_KILLED;
    } else if (d == d_s) {
      orc_uniform_angle(rng, p->angle);
      if (orc_rng_next(rng) > (sigma_s / ((1.0 - f) * sigma_a + sigma_s))) {
        /* (1 - f) sa = sigma_eff - ss, whose derivative is dsig */
        S += dsig / ((1.0 - f) * sigma_a) - dsig / total_sigma_s;
        p->group = orc_emission_group(&cells->sigma_a[(size_t)c * G], G, rng);
      } else {
        S -= dsig / total_sigma_s;
      }
      p->descriptor[0] = ORC_SCATTER;
    } else if (d == d_b) {
      const int32_t bc = cells->bc[6 * (size_t)c + sc];
      if (bc == ORC_ELEMENT) {
        c = cells->next[6 * (size_t)c + sc];
        p->cell = c;
        p->descriptor[0] = ORC_BOUND;
      } else if (bc == ORC_PROCESSOR) {
        active = 0;
        p->cell = cells->next[6 * (size_t)c + sc];
        p->descriptor[0] = ORC_PASS;
      } else if (bc == ORC_VACUUM || bc == ORC_SOURCE) {
        active = 0;
        p->descriptor[0] = ORC_EXIT;
      } else {
        p->angle[sc / 2] = -p->angle[sc / 2];
        p->descriptor[0] = ORC_BOUND;
      }
    } else if (d == d_c) {
      active = 0;
      p->descriptor[0] = ORC_CENSUS;
    }
  }
  return segs;
}

uint64_t orc_dos_transport(const orc_cells *cells, orc_photon *photons,
                           uint64_t n, const double *v_sigma_a,
                           const double *v_fleck, double *abs_E,
                           double *dabs_E, int n_threads) {
  const uint32_t nc = cells->n_cells;
  uint64_t segs = 0;
  if (n_threads <= 1) {
    for (uint64_t i = 0; i < n; ++i)
      segs += orc_dos_photon(cells, &photons[i], v_sigma_a, v_fleck, abs_E,
                             dabs_E);
    return segs;
  }
  double *ta = (double *)calloc((size_t)n_threads * nc, sizeof(double));
  double *td = (double *)calloc((size_t)n_threads * nc, sizeof(double));
#pragma omp parallel num_threads(n_threads) reduction(+ : segs)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
#pragma omp for schedule(guided)
    for (int64_t i = 0; i < (int64_t)n; ++i)
      segs += orc_dos_photon(cells, &photons[i], v_sigma_a, v_fleck,
                             ta + (size_t)t * nc, td + (size_t)t * nc);
  }
  for (uint32_t c = 0; c < nc; ++c)
    for (int t = 0; t < n_threads; ++t) {
      abs_E[c] += ta[(size_t)t * nc + c];
      dabs_E[c] += td[(size_t)t * nc + c];
    }
  free(ta);
  free(td);
  return segs;
}

/* post_process_photons (src/post_process_functions.h:32-59): census keeps
 * transport order and gets life_dx = c * next_dt. */
/* comb_photons (src/census_functions.h:48-93).  The reference's maps keyed
 * by cell are arrays here; the loops and their order are the reference's. */
uint64_t orc_comb_photons(orc_photon *photons, uint64_t n, uint32_t n_cells,
                          int64_t max_census_photons, uint32_t seed, uint64_t stream,
                          uint64_t ctr0) {
  uint32_t *count = (uint32_t *)calloc(n_cells ? n_cells : 1, sizeof(uint32_t));
  double *cell_E = (double *)calloc(n_cells ? n_cells : 1, sizeof(double));
  double *corrected = (double *)calloc(n_cells ? n_cells : 1, sizeof(double));
  double census_total_E = 0.0; /* get_photon_list_E (:40-46) */
  for (uint64_t i = 0; i < n; ++i) census_total_E += photons[i].E;
  const double comb_E = census_total_E / (double)max_census_photons; /* :56-58 */
  for (uint64_t i = 0; i < n; ++i) {                                  /* :60-63 */
    count[photons[i].cell]++;
    cell_E[photons[i].cell] += photons[i].E;
  }
  uint64_t rng[4];
  orc_rng_init(rng, seed, stream);
  rng[0] = ctr0;
  uint64_t kept = 0;
  for (uint64_t i = 0; i < n; ++i) { /* :65-79 */
    const uint32_t c = photons[i].cell;
    const double p_kill = 1.0 - photons[i].E / comb_E;
    const double rand_check = orc_rng_next(rng);
    if (rand_check > p_kill || count[c] == 1) {
      photons[i].E = comb_E;
      corrected[c] += comb_E;
      photons[kept++] = photons[i];
    } else {
      count[c]--;
    }
  }
  for (uint64_t j = 0; j < kept; ++j) { /* :81-90 */
    const uint32_t c = photons[j].cell;
    photons[j].E = photons[j].E + (cell_E[c] - corrected[c]) / (double)count[c];
  }
  free(count);
  free(cell_E);
  free(corrected);
  return kept;
}

uint64_t orc_post_process(double next_dt, orc_photon *photons, uint64_t n,
                          orc_photon *census_out, uint64_t cap,
                          orc_census_stats *st) {
  uint64_t n_complete = 0;
  for (uint64_t i = 0; i < n; ++i) {
    orc_photon *p = &photons[i];
    switch (p->descriptor[0]) {
    case ORC_KILLED:
      st->n_killed++;
      n_complete++;
      break;
    case ORC_EXIT:
      st->exit_E += p->E;
      st->n_exit++;
      n_complete++;
      break;
    case ORC_CENSUS:
      p->life_dx = ORC_C * next_dt;
      if (st->n_census < cap) census_out[st->n_census] = *p;
      st->n_census++;
      st->census_E += p->E;
      n_complete++;
      break;
    default:
      break;
    }
  }
  return n_complete;
}

/* ======================================================================== */
/* Deck -> mesh, step physics, time stepping and the replicated driver.     */
/* ======================================================================== */
struct orc_sim {
  /* deck */
  uint32_t n_regions;
  double *reg; /* [n_regions][8] cv rho opA opB opC opS T_e T_r */
  uint64_t n_user;
  uint32_t seed, G;
  /* mesh */
  uint32_t n_cells;
  double *nodes;
  int32_t *bc;
  uint32_t *next;
  uint32_t *region;
  double *T_e, *T_r_cell, *T_s, *cv, *rho;
  /* physics of the current step */
  double *sigma_a, *sigma_s, *fleck;
  double *emission_E, *source_E, *census_E;
  int8_t *source_face;
  double total_photon_E;
  int prepared;
  /* tallies and diagnostics */
  double *abs_E, *track_E, *T_r_diag;
  /* census list */
  orc_photon *census;
  uint64_t n_census;
  /* replicated rank of this simulation (src/mesh.h:80-87) and the all-rank
   * photon energy the source divides by (src/replicated_driver.h:57-59) */
  uint32_t rank, n_ranks;
  double global_photon_E;
  /* IMC_State (src/imc_state.h:42-45, 112-134, 292-296) */
  double time, dt, t_stop, t_mult, dt_max;
  uint32_t step;
  double pre_mat_E, emission_tot, source_tot, pre_census_E;
};

static double orc_next_dt(const orc_sim *s) {
  double next_dt;
  if (s->dt * s->t_mult < s->dt_max)
    next_dt = s->dt * s->t_mult;
  else
    next_dt = s->dt_max;
  if (s->time + next_dt > s->t_stop) next_dt = s->t_stop - s->time;
  return next_dt;
}

static double orc_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Proto_Mesh constructor (src/proto_mesh.h:81-218), one rank. */
orc_sim *orc_sim_create(const orc_deck *d) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof *s);
  if (!s) return NULL;
  s->n_regions = d->n_regions;
  s->reg = (double *)malloc(sizeof(double) * 8 * d->n_regions);
  for (uint32_t r = 0; r < d->n_regions; ++r) {
    double *q = &s->reg[8 * r];
    q[0] = d->cv[r];
    q[1] = d->rho[r];
    q[2] = d->opA[r];
    q[3] = d->opB[r];
    q[4] = d->opC[r];
    q[5] = d->opS[r];
    q[6] = d->T_e[r];
    q[7] = d->T_r[r];
  }
  s->n_user = d->n_photons;
  s->seed = d->seed;
  s->G = d->n_groups ? d->n_groups : 1;

  uint32_t ng[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    for (uint32_t k = 0; k < d->n_div[a]; ++k) ng[a] += d->div_cells[a][k];
  const uint32_t n = ng[0] * ng[1] * ng[2];
  s->n_cells = n;
  s->nodes = (double *)malloc(sizeof(double) * 6 * n);
  s->bc = (int32_t *)malloc(sizeof(int32_t) * 6 * n);
  s->next = (uint32_t *)malloc(sizeof(uint32_t) * 6 * n);
  s->region = (uint32_t *)malloc(sizeof(uint32_t) * n);
  s->T_e = (double *)malloc(sizeof(double) * n);
  s->T_r_cell = (double *)malloc(sizeof(double) * n);
  s->T_s = (double *)malloc(sizeof(double) * n);
  s->cv = (double *)malloc(sizeof(double) * n);
  s->rho = (double *)malloc(sizeof(double) * n);
  s->sigma_a = (double *)calloc((size_t)n * s->G, sizeof(double));
  s->sigma_s = (double *)calloc((size_t)n * s->G, sizeof(double));
  s->fleck = (double *)calloc(n, sizeof(double));
  s->emission_E = (double *)calloc(n, sizeof(double));
  s->source_E = (double *)calloc(n, sizeof(double));
  s->census_E = (double *)calloc(n, sizeof(double));
  s->source_face = (int8_t *)calloc(n, 1);
  s->abs_E = (double *)calloc(n, sizeof(double));
  s->track_E = (double *)calloc(n, sizeof(double));
  s->T_r_diag = (double *)calloc(n, sizeof(double));

  const uint32_t ngx = ng[0], ngy = ng[1], ngz = ng[2];
  uint32_t gc = 0, g_i = 0, g_j = 0, g_k = 0;
  for (uint32_t izd = 0; izd < d->n_div[2]; ++izd) {
    const double dz = (d->div_end[2][izd] - d->div_start[2][izd]) /
                      (double)d->div_cells[2][izd];
    const uint32_t nz = d->div_cells[2][izd];
    const double z0 = d->div_start[2][izd];
    for (uint32_t k = 0; k < nz; ++k) {
      g_j = 0;
      for (uint32_t iyd = 0; iyd < d->n_div[1]; ++iyd) {
        const double dy = (d->div_end[1][iyd] - d->div_start[1][iyd]) /
                          (double)d->div_cells[1][iyd];
        const uint32_t ny = d->div_cells[1][iyd];
        const double y0 = d->div_start[1][iyd];
        for (uint32_t j = 0; j < ny; ++j) {
          g_i = 0;
          for (uint32_t ixd = 0; ixd < d->n_div[0]; ++ixd) {
            const double dx = (d->div_end[0][ixd] - d->div_start[0][ixd]) /
                              (double)d->div_cells[0][ixd];
            const uint32_t nx = d->div_cells[0][ixd];
            const double x0 = d->div_start[0][ixd];
            for (uint32_t i = 0; i < nx; ++i) {
              double *nd = &s->nodes[6 * (size_t)gc];
              /* division ends snap to the next division's start (:130-145) */
              const double xe = (i == nx - 1 && ixd != d->n_div[0] - 1)
                                    ? d->div_start[0][ixd + 1]
                                    : x0 + (i + 1) * dx;
              const double ye = (j == ny - 1 && iyd != d->n_div[1] - 1)
                                    ? d->div_start[1][iyd + 1]
                                    : y0 + (j + 1) * dy;
              const double ze = (k == nz - 1 && izd != d->n_div[2] - 1)
                                    ? d->div_start[2][izd + 1]
                                    : z0 + (k + 1) * dz;
              nd[0] = x0 + i * dx;
              nd[1] = xe;
              nd[2] = y0 + j * dy;
              nd[3] = ye;
              nd[4] = z0 + k * dz;
              nd[5] = ze;
              const uint32_t r =
                  d->div_region[((size_t)izd * d->n_div[1] + iyd) *
                                    d->n_div[0] +
                                ixd];
              s->region[gc] = r;
              int32_t *b = &s->bc[6 * (size_t)gc];
              uint32_t *nx6 = &s->next[6 * (size_t)gc];
              /* neighbours and boundary conditions (:156-202) */
              if (g_i < ngx - 1) { nx6[1] = gc + 1; b[1] = ORC_ELEMENT; }
              else { nx6[1] = gc; b[1] = d->bc[1]; }
              if (g_i > 0) { nx6[0] = gc - 1; b[0] = ORC_ELEMENT; }
              else { nx6[0] = gc; b[0] = d->bc[0]; }
              if (g_j < ngy - 1) { nx6[3] = gc + ngx; b[3] = ORC_ELEMENT; }
              else { nx6[3] = gc; b[3] = d->bc[3]; }
              if (g_j > 0) { nx6[2] = gc - ngx; b[2] = ORC_ELEMENT; }
              else { nx6[2] = gc; b[2] = d->bc[2]; }
              if (g_k < ngz - 1) { nx6[5] = gc + ngx * ngy; b[5] = ORC_ELEMENT; }
              else { nx6[5] = gc; b[5] = d->bc[5]; }
              if (g_k > 0) { nx6[4] = gc - ngx * ngy; b[4] = ORC_ELEMENT; }
              else { nx6[4] = gc; b[4] = d->bc[4]; }
              /* Mesh::initialize_physical_properties (src/mesh.h:426-439) */
              const double *q = &s->reg[8 * r];
              s->cv[gc] = q[0];
              s->rho[gc] = q[1];
              s->T_e[gc] = q[6];
              s->T_r_cell[gc] = q[7];
              int face = -1;
              for (int f = 0; f < 6; ++f)
                if (b[f] == ORC_SOURCE) { face = f; break; }
              s->source_face[gc] = (int8_t)face;
              s->T_s[gc] = (face != -1) ? d->T_source : 0.0;
              gc++;
              g_i++;
            }
          }
          g_j++;
        }
      }
      g_k++;
    }
  }
  s->time = d->t_start;
  s->dt = d->dt_start;
  s->t_stop = d->t_stop;
  s->t_mult = d->t_mult;
  s->dt_max = d->dt_max;
  s->step = 1;
  s->rank = 0;
  s->n_ranks = 1;
  return s;
}

void orc_sim_set_rank(orc_sim *s, uint32_t rank, uint32_t n_ranks) {
  s->n_ranks = n_ranks ? n_ranks : 1;
  s->rank = rank < s->n_ranks ? rank : 0;
  s->prepared = 0;
}

void orc_sim_destroy(orc_sim *s) {
  if (!s) return;
  free(s->reg); free(s->nodes); free(s->bc); free(s->next); free(s->region);
  free(s->T_e); free(s->T_r_cell); free(s->T_s); free(s->cv); free(s->rho);
  free(s->sigma_a); free(s->sigma_s); free(s->fleck);
  free(s->emission_E); free(s->source_E); free(s->census_E);
  free(s->source_face); free(s->abs_E); free(s->track_E); free(s->T_r_diag);
  free(s->census);
  free(s);
}

uint32_t orc_sim_n_cells(const orc_sim *s) { return s->n_cells; }

int orc_sim_finished(const orc_sim *s) {
  return fabs(s->time - s->t_stop) < 1.0e-8; /* src/imc_state.h:128-134 */
}

static double orc_face_area(const double *nd, int face) {
  if (face == 0 || face == 1) return (nd[3] - nd[2]) * (nd[5] - nd[4]);
  if (face == 2 || face == 3) return (nd[1] - nd[0]) * (nd[5] - nd[4]);
  if (face == 4 || face == 5) return (nd[1] - nd[0]) * (nd[3] - nd[2]);
  return -1.0;
}

/* One cell's three energies after the replicated small-energy reassignment
 * (src/mesh.h:291-307) as rank r of n_ranks sees them: an energy that the
 * all-rank source would give no photon of its own, int(n_user * (E /
 * global_source_E)) == 0, goes whole (E / replicated_factor) to rank
 * i % n_ranks and is zero on every other rank. */
static void orc_reassign(uint32_t i, uint32_t r, uint32_t n_ranks, double rf,
                         uint32_t n_user, double global_source_E, uint32_t step,
                         double *em, double *src, double *cen) {
  const int mine = (i % n_ranks) == r;
  if (step == 1 && *cen > 0.0 && (int)(n_user * (*cen / global_source_E)) == 0)
    *cen = mine ? *cen / rf : 0.0;
  if (*em > 0.0 && (int)(n_user * (*em / global_source_E)) == 0)
    *em = mine ? *em / rf : 0.0;
  if (*src > 0.0 && (int)(n_user * (*src / global_source_E)) == 0)
    *src = mine ? *src / rf : 0.0;
}

/* Mesh::calculate_photon_energy (src/mesh.h:237-323) for rank `rank` of
 * `n_ranks` (replicated_factor = 1 / n_ranks, src/mesh.h:87).  The two
 * MPI_Allreduce(SUM) calls -- of the three step totals inside
 * calculate_photon_energy (src/mesh.h:292-294) and of total_photon_E in the
 * driver (src/replicated_driver.h:57-59) -- are modelled as sums taken in
 * rank order.  Every rank holds the same mesh, so the first sums n_ranks
 * copies of this rank's value; for the second, every rank's reassigned
 * total is computed here from the same per-cell energies.  One rank: both
 * are this rank's own values and the reassignment keeps every energy. */
int orc_sim_prepare(orc_sim *s) {
  const double dt = s->dt;
  const uint32_t step = s->step;
  const double rf = 1.0 / (double)s->n_ranks;
  const uint32_t n_user = (uint32_t)s->n_user; /* n_user_photons is uint32_t */
  double tot_census = 0.0, tot_emission = 0.0, tot_source = 0.0;
  double pre_mat = 0.0;
  s->total_photon_E = 0.0;
  for (uint32_t i = 0; i < s->n_cells; ++i) {
    const double *nd = &s->nodes[6 * (size_t)i];
    const double vol = (nd[1] - nd[0]) * (nd[3] - nd[2]) * (nd[5] - nd[4]);
    const double cV = s->cv[i];
    const double T = s->T_e[i];
    const double Tr = s->T_r_cell[i];
    const double Ts = s->T_s[i];
    const double rho = s->rho[i];
    const double *q = &s->reg[8 * s->region[i]];
    const double op_a = q[2] + q[3] * pow(T, q[4]); /* src/region.h:44-46 */
    const double op_s = q[5];
    const double f =
        1.0 / (1.0 + dt * op_a * ORC_C * (4.0 * ORC_A * pow(T, 3) / (cV * rho)));
    for (uint32_t g = 0; g < s->G; ++g) {
      s->sigma_a[(size_t)i * s->G + g] = op_a;
      s->sigma_s[(size_t)i * s->G + g] = op_s;
    }
    s->fleck[i] = f;
    s->emission_E[i] = rf * dt * vol * f * op_a * ORC_A * ORC_C * pow(T, 4);
    if (step > 1)
      s->census_E[i] = 0.0;
    else
      s->census_E[i] = rf * vol * ORC_A * pow(Tr, 4);
    s->source_E[i] = rf * 0.25 * ORC_A * ORC_C *
                     orc_face_area(nd, s->source_face[i]) * pow(Ts, 4) * dt;
    pre_mat += T * cV * vol * rho;
    tot_emission += s->emission_E[i];
    tot_census += s->census_E[i];
    tot_source += s->source_E[i];
    s->total_photon_E += s->source_E[i] + s->census_E[i] + s->emission_E[i];
  }
  /* src/mesh.h:289-315 (replicated is set for every replicated run) */
  const double local3 = tot_emission + tot_census + tot_source;
  double global_source_E = local3;
  for (uint32_t r = 1; r < s->n_ranks; ++r) global_source_E += local3;
  /* every rank's total after its own reassignment, summed in rank order;
   * this rank's is recomputed below exactly as the reference's loop does */
  double global_photon_E = 0.0;
  for (uint32_t r = 0; r < s->n_ranks && s->n_ranks > 1; ++r) {
    double t = 0.0;
    for (uint32_t i = 0; i < s->n_cells; ++i) {
      double em = s->emission_E[i], src = s->source_E[i], cen = s->census_E[i];
      orc_reassign(i, r, s->n_ranks, rf, n_user, global_source_E, step, &em, &src, &cen);
      t += src + cen + em;
    }
    global_photon_E += t;
  }
  tot_census = tot_emission = tot_source = 0.0;
  s->total_photon_E = 0.0;
  for (uint32_t i = 0; i < s->n_cells; ++i) {
    orc_reassign(i, s->rank, s->n_ranks, rf, n_user, global_source_E, step,
                 &s->emission_E[i], &s->source_E[i], &s->census_E[i]);
    tot_emission += s->emission_E[i];
    tot_census += s->census_E[i];
    tot_source += s->source_E[i];
    s->total_photon_E += s->source_E[i] + s->census_E[i] + s->emission_E[i];
  }
  s->global_photon_E = (s->n_ranks > 1) ? global_photon_E : s->total_photon_E;
  s->pre_mat_E = pre_mat;
  s->emission_tot = tot_emission;
  s->source_tot = tot_source;
  if (step == 1) s->pre_census_E = tot_census;
  s->prepared = 1;
  return 0;
}

static double orc_list_E(const orc_photon *p, uint64_t n) {
  double t = 0.0;
  for (uint64_t i = 0; i < n; ++i) t += p[i].E;
  return t;
}

/* imc_replicated_driver loop body (src/replicated_driver.h:47-121) */
int orc_sim_step(orc_sim *s, int n_threads, orc_step_report *rep) {
  if (orc_sim_finished(s)) return -1;
  memset(rep, 0, sizeof *rep);
  if (!s->prepared) orc_sim_prepare(s);
  s->prepared = 0;
  const double global_E = s->global_photon_E;
  const double dt = s->dt;
  orc_cells cells = {s->n_cells, s->G, s->nodes, s->bc, s->next,
                     s->sigma_a, s->sigma_s, s->fleck};

  s->pre_census_E = orc_list_E(s->census, s->n_census);
  const double ts = orc_seconds();
  if (s->step == 1) {
    const uint64_t nc = orc_make_initial_census(&cells, s->census_E, s->n_user,
                                                global_E, s->seed, s->rank, dt,
                                                NULL, 0);
    free(s->census);
    s->census = (orc_photon *)malloc(sizeof(orc_photon) * (nc ? nc : 1));
    orc_make_initial_census(&cells, s->census_E, s->n_user, global_E, s->seed,
                            s->rank, dt, s->census, nc);
    s->n_census = nc;
  }
  s->pre_census_E = orc_list_E(s->census, s->n_census);

  const uint64_t n_new = orc_make_photons(&cells, s->emission_E, s->source_E,
                                          s->source_face, s->n_user, global_E,
                                          s->seed, s->step, s->rank, dt, NULL, 0);
  const uint64_t n_all = n_new + s->n_census;
  orc_photon *all = (orc_photon *)malloc(sizeof(orc_photon) * (n_all ? n_all : 1));
  orc_make_photons(&cells, s->emission_E, s->source_E, s->source_face,
                   s->n_user, global_E, s->seed, s->step, s->rank, dt, all, n_new);
  if (s->n_census)
    memcpy(all + n_new, s->census, sizeof(orc_photon) * s->n_census);
  rep->transported = n_all;
  rep->source_seconds = orc_seconds() - ts;

  /* replicated_transport (src/replicated_transport.h:50-157) */
  const double next_dt = orc_next_dt(s);
  memset(s->abs_E, 0, sizeof(double) * s->n_cells);
  memset(s->track_E, 0, sizeof(double) * s->n_cells);
  const double t0 = orc_seconds();
  rep->segments =
      orc_transport(&cells, all, n_all, s->abs_E, s->track_E, n_threads);
  orc_census_stats st;
  memset(&st, 0, sizeof st);
  free(s->census);
  s->census = (orc_photon *)malloc(sizeof(orc_photon) * (n_all ? n_all : 1));
  orc_post_process(next_dt, all, n_all, s->census, n_all, &st);
  rep->transport_seconds = orc_seconds() - t0;
  s->n_census = st.n_census;
  free(all);

  /* Mesh::update_temperature (src/mesh.h:327-423) */
  double total_abs = 0.0, post_mat = 0.0;
  for (uint32_t i = 0; i < s->n_cells; ++i) {
    const double *q = &s->reg[8 * s->region[i]];
    const double cV = q[0], rho = q[1];
    const double *nd = &s->nodes[6 * (size_t)i];
    const double vol = (nd[1] - nd[0]) * (nd[3] - nd[2]) * (nd[5] - nd[4]);
    const double T = s->T_e[i];
    const double T_new = T + (s->abs_E[i] - s->emission_E[i]) / (cV * vol * rho);
    s->T_r_diag[i] = pow(s->track_E[i] / (vol * dt * ORC_A * ORC_C), 0.25);
    s->T_e[i] = T_new;
    total_abs += s->abs_E[i];
    post_mat += T_new * cV * vol * rho;
  }

  rep->step = s->step;
  rep->time = s->time;
  rep->dt = dt;
  rep->next_dt = next_dt;
  rep->census_size = st.n_census;
  rep->emission_E = s->emission_tot;
  rep->source_E = s->source_tot;
  rep->absorbed_E = total_abs;
  rep->exit_E = st.exit_E;
  rep->pre_census_E = s->pre_census_E;
  rep->post_census_E = st.census_E;
  rep->pre_mat_E = s->pre_mat_E;
  rep->post_mat_E = post_mat;
  /* src/imc_state.h:254-258 */
  rep->rad_conservation = (total_abs + st.census_E + st.exit_E) -
                          (s->pre_census_E + s->emission_tot + s->source_tot);
  rep->mat_conservation = post_mat - (s->pre_mat_E + total_abs - s->emission_tot);

  /* IMC_State::next_time_step (src/imc_state.h:292-296) */
  s->time += s->dt;
  s->dt = orc_next_dt(s);
  s->step++;
  return 0;
}

void orc_sim_get_mesh(const orc_sim *s, double *nodes, int32_t *bc,
                      uint32_t *next, uint32_t *region) {
  const size_t n = s->n_cells;
  if (nodes) memcpy(nodes, s->nodes, sizeof(double) * 6 * n);
  if (bc) memcpy(bc, s->bc, sizeof(int32_t) * 6 * n);
  if (next) memcpy(next, s->next, sizeof(uint32_t) * 6 * n);
  if (region) memcpy(region, s->region, sizeof(uint32_t) * n);
}

void orc_sim_get_physics(const orc_sim *s, double *sigma_a, double *sigma_s,
                         double *fleck, double *emission_E, double *source_E,
                         double *census_E, int8_t *source_face,
                         double *total_E) {
  const size_t n = s->n_cells;
  if (sigma_a) memcpy(sigma_a, s->sigma_a, sizeof(double) * n * s->G);
  if (sigma_s) memcpy(sigma_s, s->sigma_s, sizeof(double) * n * s->G);
  if (fleck) memcpy(fleck, s->fleck, sizeof(double) * n);
  if (emission_E) memcpy(emission_E, s->emission_E, sizeof(double) * n);
  if (source_E) memcpy(source_E, s->source_E, sizeof(double) * n);
  if (census_E) memcpy(census_E, s->census_E, sizeof(double) * n);
  if (source_face) memcpy(source_face, s->source_face, n);
  if (total_E) *total_E = s->global_photon_E;
}

void orc_sim_get_tallies(const orc_sim *s, double *abs_E, double *track_E) {
  if (abs_E) memcpy(abs_E, s->abs_E, sizeof(double) * s->n_cells);
  if (track_E) memcpy(track_E, s->track_E, sizeof(double) * s->n_cells);
}

void orc_sim_get_temperature(const orc_sim *s, double *T_e, double *T_r) {
  if (T_e) memcpy(T_e, s->T_e, sizeof(double) * s->n_cells);
  if (T_r) memcpy(T_r, s->T_r_diag, sizeof(double) * s->n_cells);
}

uint64_t orc_sim_census_size(const orc_sim *s) { return s->n_census; }

void orc_sim_clock(const orc_sim *s, double *time, double *dt, double *next_dt,
                   uint32_t *step) {
  if (time) *time = s->time;
  if (dt) *dt = s->dt;
  if (next_dt) *next_dt = orc_next_dt(s);
  if (step) *step = s->step;
}

void orc_sim_get_census(const orc_sim *s, orc_photon *out) {
  if (s->n_census) memcpy(out, s->census, sizeof(orc_photon) * s->n_census);
}
