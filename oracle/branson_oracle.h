/*
 * branson_oracle.h -- CPU restatement of Branson's replicated-mode IMC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only to check results against (or to time the CPU).  The product (branson_amd, libbranson_hip.so)
 * never links, loads or calls anything in this directory.
 *
 * Every function below restates one piece of the reference algorithm and
 * cites the reference file:line it follows (paths relative to the reference
 * checkout, e.g. src/history_based_transport.h).  Floating point is IEEE
 * double with contraction disabled (-ffp-contract=off); transcendental calls
 * go to the host libm exactly as the reference's do.
 *
 * Pinning (what fixes each function independently of this file; the tests
 * named are in tests/):
 *   - orc_threefry2x64_20: known-answer vectors dumped from the reference's own
 *     vendored Random123 header compiled in place (oracle/_ref), plus the
 *     published all-zero Random123 vector (test_oracle_rng.py).
 *   - orc_u01: the closed form documented at src/RNG.h:189-201 (odd multiples
 *     of 2^-53; extremes 2^-53 and 1-2^-53).
 *   - orc_distance_to_boundary / move / reflect: closed forms of
 *     src/test/test_cell.cc:79-133 and src/test/test_photon.cc:69-126.
 *   - samplers and make_photons: source photons hand traced from the pinned
 *     draws (positions, directions, lives, groups, draw counts, streams),
 *     statistical means, unit norms, cell containment; the group walk as
 *     inversion of its CDF (test_oracle_pins.py).
 *   - source counts: the closed form max(1, int(n*E/E_tot)) (src/source.h).
 *   - transport: effective scattering in an infinite medium (scatter rate
 *     sigma_eff = (1-f) sigma_a + sigma_s, re-sampling fraction
 *     (1-f) sigma_a / sigma_eff, weight exp(-sigma_a f L), tallies, every
 *     draw accounted for, isotropy), hand-traced single-scatter histories,
 *     the kill deposit, and the cosine-law beam through a pure absorber
 *     (2 E3(tau)).
 *   - post_process: census order, census life and totals on a hand-built
 *     list.
 *   - step physics: the textbook implicit Monte Carlo forms (Fleck &
 *     Cummings 1971) over three steps of a two-region deck with
 *     temperature-dependent opacities; the material energy balance and the
 *     track-length radiation temperature.
 *   - rank share: n ranks' energies against the one-rank energies (every
 *     cell's energy summed over ranks; small energies whole on rank
 *     cell % n); every rank's E_tot equal to the one-rank total.
 *   - whole steps: per-step photon counts, census sizes and energies printed
 *     by the reference executable for the 100-cell Marshak slab (10 steps,
 *     seed 777) and the 40^3 cube (1 step, seed 1234); conservation to
 *     round-off.
 */
#ifndef BRANSON_ORACLE_H
#define BRANSON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/constants.h:26-28 */
enum { ORC_REFLECT = 0, ORC_VACUUM, ORC_ELEMENT, ORC_SOURCE, ORC_PROCESSOR };
enum { ORC_EXIT = 0, ORC_PASS, ORC_CENSUS, ORC_SCATTER, ORC_KILLED, ORC_BOUND };

/* The reference's 120-byte AOS Photon (src/photon.h:172-181). */
typedef struct orc_photon {
  uint32_t cell;
  uint32_t group;
  uint32_t source_type;
  uint8_t descriptor[4];
  double pos[3];
  double angle[3];
  double E;
  double E0;
  double life_dx;
  uint64_t rng[4]; /* ctr lo, ctr hi (= seed<<32), key lo (= stream), key hi */
} orc_photon;

/* Read-only cell data during transport (src/cell.h:319-336), cell-major. */
typedef struct orc_cells {
  uint32_t n_cells;
  uint32_t n_groups;
  const double *nodes;   /* [n_cells][6] x_lo x_hi y_lo y_hi z_lo z_hi */
  const int32_t *bc;     /* [n_cells][6] */
  const uint32_t *next;  /* [n_cells][6] global neighbour index         */
  const double *sigma_a; /* [n_cells][n_groups] */
  const double *sigma_s; /* [n_cells][n_groups] */
  const double *fleck;   /* [n_cells] */
} orc_cells;

/* ---- RNG (src/RNG.h, src/random123/threefry.h) ------------------------- */
void orc_threefry2x64_20(const uint64_t ctr[2], const uint64_t key[2],
                         uint64_t out[2]);
double orc_u01(uint64_t x);
/* n draws of RNG(seed, stream) starting at counter value ctr0 */
void orc_rng_draws(uint32_t seed, uint64_t stream, uint64_t ctr0, uint32_t n,
                   double *out);

/* ---- samplers (src/sampling_functions.h) -------------------------------- */
/* Each sampler starts a fresh RNG(seed, stream) and returns the counter after
 * use in *ctr_after so draw accounting can be checked. */
void orc_sample_position_in_cell(const double nodes[6], uint32_t seed,
                                 uint64_t stream, double pos[3],
                                 uint64_t *ctr_after);
void orc_sample_position_on_face(const double nodes[6], int face,
                                 uint32_t seed, uint64_t stream,
                                 double pos[3], uint64_t *ctr_after);
void orc_sample_uniform_angle(uint32_t seed, uint64_t stream, double ang[3],
                              uint64_t *ctr_after);
void orc_sample_source_angle(int face, uint32_t seed, uint64_t stream,
                             double ang[3], uint64_t *ctr_after);
int orc_sample_emission_group(const double *sigma_a_cell, uint32_t n_groups,
                              uint32_t seed, uint64_t stream,
                              uint64_t *ctr_after);
double orc_distance_to_boundary(const double nodes[6], const double pos[3],
                                const double ang[3], uint32_t *surface_cross);

/* ---- source (src/source.h) ---------------------------------------------- */
/* Photon counts per cell: counts[2*c] emission (or census), counts[2*c+1]
 * face source.  Returns the total. */
uint64_t orc_source_counts(uint32_t n_cells, const double *emission_E,
                           const double *source_E, uint64_t n_user,
                           double total_E, uint32_t *counts);
/* make_photons (src/source.h:207-291): fills out[0..total). */
uint64_t orc_make_photons(const orc_cells *cells, const double *emission_E,
                          const double *source_E, const int8_t *source_face,
                          uint64_t n_user, double total_E, uint32_t seed,
                          uint32_t cycle, uint32_t rank, double dt,
                          orc_photon *out, uint64_t cap);
/* make_initial_census_photons (src/source.h:134-180). */
uint64_t orc_make_initial_census(const orc_cells *cells,
                                 const double *census_E, uint64_t n_user,
                                 double total_E, uint32_t seed, uint32_t rank,
                                 double dt, orc_photon *out, uint64_t cap);

/* ---- transport (src/history_based_transport.h) -------------------------- */
/* transport_photon for every photon; n_threads > 1 uses OpenMP guided
 * scheduling with per-thread tallies merged in thread order (:279-308).
 * abs_E/track_E are accumulated into (not cleared).  Returns segments. */
uint64_t orc_transport(const orc_cells *cells, orc_photon *photons,
                       uint64_t n, double *abs_E, double *track_E,
                       int n_threads);

/* Differential-operator (likelihood-ratio) derivative of the absorbed-energy
 * tally along the direction (v_sigma_a, v_fleck) of the per-cell parameters,
 * source photons held fixed; forward mode, same events as orc_transport.
 * dabs_E[c] and abs_E[c] are accumulated into.  Group-independent sigma_a
 * only.  Returns segments. */
uint64_t orc_dos_transport(const orc_cells *cells, orc_photon *photons,
                           uint64_t n, const double *v_sigma_a,
                           const double *v_fleck, double *abs_E,
                           double *dabs_E, int n_threads);

/* post_process_photons (src/post_process_functions.h:32-59). */
typedef struct orc_census_stats {
  double census_E;
  double exit_E;
  uint64_t n_census;
  uint64_t n_exit;
  uint64_t n_killed;
} orc_census_stats;
/* comb_photons (src/census_functions.h:48-93), one rank: combs the census
 * (in census order) down toward max_census_photons photons of energy
 * E_census / max_census_photons, with per-cell energy correction.  The
 * reference takes an RNG*; here it is RNG(seed, stream) from counter ctr0,
 * one draw per census photon in order.  Survivors are written to the front
 * of `photons` in census order; returns their number. */
uint64_t orc_comb_photons(orc_photon *photons, uint64_t n, uint32_t n_cells,
                          int64_t max_census_photons, uint32_t seed, uint64_t stream,
                          uint64_t ctr0);

uint64_t orc_post_process(double next_dt, orc_photon *photons, uint64_t n,
                          orc_photon *census_out, uint64_t cap,
                          orc_census_stats *stats);

/* ---- deck, mesh and step physics (src/input.h, src/proto_mesh.h,
 *      src/mesh.h, src/imc_state.h, src/replicated_driver.h) -------------- */
typedef struct orc_deck {
  double t_start, t_stop, dt_start, t_mult, dt_max;
  uint64_t n_photons;
  uint32_t seed;
  uint32_t n_groups;
  uint32_t n_div[3];
  const double *div_start[3];
  const double *div_end[3];
  const uint32_t *div_cells[3];
  const uint32_t *div_region; /* [iz][iy][ix] -> index into region arrays */
  uint32_t n_regions;
  const double *cv, *rho, *opA, *opB, *opC, *opS, *T_e, *T_r;
  int32_t bc[6];
  double T_source;
} orc_deck;

typedef struct orc_sim orc_sim;

typedef struct orc_step_report {
  uint32_t step;
  double time, dt, next_dt;
  uint64_t transported;
  uint64_t census_size;
  uint64_t segments;
  double emission_E, source_E, absorbed_E, exit_E;
  double pre_census_E, post_census_E, pre_mat_E, post_mat_E;
  double rad_conservation, mat_conservation;
  double transport_seconds; /* transport + post-processing into the census */
  double source_seconds;    /* initial census + make_photons + join         */
} orc_step_report;

orc_sim *orc_sim_create(const orc_deck *deck);
void orc_sim_destroy(orc_sim *sim);
/* Runs the simulation as rank `rank` of `n_ranks` replicated ranks
 * (default 0 of 1): replicated_factor 1/n_ranks, the small-energy
 * reassignment and the all-rank E_tot of src/mesh.h:287-315 and
 * src/replicated_driver.h:57-59, and the rank's stream offsets.  A step
 * transports this rank's share only; its temperature update then uses this
 * rank's tallies, not the all-rank sum the reference reduces first. */
void orc_sim_set_rank(orc_sim *sim, uint32_t rank, uint32_t n_ranks);
uint32_t orc_sim_n_cells(const orc_sim *sim);
int orc_sim_finished(const orc_sim *sim);
/* one replicated time step (src/replicated_driver.h:47-121) */
int orc_sim_step(orc_sim *sim, int n_threads, orc_step_report *rep);
/* mesh export: cell-major arrays sized n_cells*6 / n_cells*n_groups */
void orc_sim_get_mesh(const orc_sim *sim, double *nodes, int32_t *bc,
                      uint32_t *next, uint32_t *region);
/* state after the last calculate_photon_energy: per-cell opacities, Fleck
 * factor and the three source energies of the current step; total_E is the
 * all-rank photon energy the source divides by */
void orc_sim_get_physics(const orc_sim *sim, double *sigma_a, double *sigma_s,
                         double *fleck, double *emission_E, double *source_E,
                         double *census_E, int8_t *source_face,
                         double *total_E);
void orc_sim_get_tallies(const orc_sim *sim, double *abs_E, double *track_E);
void orc_sim_get_temperature(const orc_sim *sim, double *T_e, double *T_r);
uint64_t orc_sim_census_size(const orc_sim *sim);
void orc_sim_get_census(const orc_sim *sim, orc_photon *out);
/* clock of the coming step: time, dt, the census life step next_dt
 * (IMC_State::get_next_dt, src/imc_state.h:112-125) and the step number */
void orc_sim_clock(const orc_sim *sim, double *time, double *dt, double *next_dt,
                   uint32_t *step);
/* calculate_photon_energy only (src/mesh.h:237-323) for the coming step */
int orc_sim_prepare(orc_sim *sim);

#ifdef __cplusplus
}
#endif
#endif
