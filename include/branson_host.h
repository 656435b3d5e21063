/*
 * branson_host.h -- C-ABI of the host side: Branson's replicated-mode time
 * loop over the MI355X engine of branson_hip.h.
 *
 * A bh_driver owns what the reference's main() builds for one rank
 * (src/main.cc:71-131): the parsed XML deck (Input, src/input.h), the mesh
 * with its per-cell material state (Mesh, src/mesh.h; numbering of
 * src/proto_mesh.h:106-218), the time-step bookkeeping (IMC_State,
 * src/imc_state.h) and an engine handle.  bh_driver_step is one iteration of
 * imc_replicated_driver (src/replicated_driver.h:47-121); the three phase
 * calls split it so a caller can time the device phase alone.
 *
 * Same conventions as branson_hip.h: BH_OK or a negative BH_ERR_* code,
 * message via bh_driver_last_error, never aborts, not thread safe.
 */
#ifndef BRANSON_HOST_H
#define BRANSON_HOST_H

#include <stdint.h>

#include "branson_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bh_driver_config {
  bh_options engine;     /* engine options (precision, algorithm, ...)       */
  int32_t device;        /* HIP device                                        */
  uint32_t n_groups;     /* the reference's compile-time BRANSON_N_GROUPS; 0 = 1 */
  uint32_t rank;         /* replicated rank of this process (stream offsets)  */
  uint32_t n_ranks;      /* ranks sharing the source energy; 0 = 1            */
  uint64_t n_photons;    /* override of the deck's <photons>; 0 = deck value  */
  double t_stop;         /* override of the deck's <t_stop>; 0 = deck value   */
  int32_t device_physics; /* 1: each step's physics (calculate_photon_energy,
                             update_temperature) runs on the device and the
                             material state stays there (bh_step_resident);
                             0: host physics, per-step uploads             */
  int32_t reserved;      /* set to 0                                          */
} bh_driver_config;

/* Per-step diagnostics: IMC_State::print_conservation's quantities
 * (src/imc_state.h:174-289) plus counters and timings. */
typedef struct bh_driver_report {
  uint32_t step;
  double time, dt, next_dt;
  uint64_t transported, census_size, n_segments, n_flushes, n_exit, n_killed;
  double emission_E, source_E, absorbed_E, exit_E;
  double pre_census_E, post_census_E, pre_mat_E, post_mat_E;
  double rad_conservation, mat_conservation;
  double source_ms, transport_ms, census_ms; /* device (HIP events)          */
  double prepare_ms, device_phase_ms, finish_ms; /* host wall clock per phase */
  uint64_t n_replayed;   /* FP32 tracking: histories finished in FP64       */
  int32_t lattice;       /* 1: tracked on the lattice records               */
  int32_t tally;         /* tally form (bh_step_stats.tally)                */
  double replay_ms;      /* FP32 tracking: FP64 replay part of transport_ms */
  uint64_t n_allocs;     /* device allocations in the step's device phase
                            (bh_step_stats.n_allocs)                       */
} bh_driver_report;

typedef struct bh_driver bh_driver;

/* Parses the deck (Input, src/input.h:77-637 rules: unknown boundary types
 * and division/region-map mismatches are errors) and builds the mesh.  The
 * engine is created on the first device call. */
int bh_driver_create(const char *xml_path, const bh_driver_config *cfg, bh_driver **out);
void bh_driver_destroy(bh_driver *d);
const char *bh_driver_last_error(const bh_driver *d);

int bh_driver_finished(const bh_driver *d); /* IMC_State::finished */
uint32_t bh_driver_n_cells(const bh_driver *d);
uint64_t bh_driver_n_photons(const bh_driver *d);
uint32_t bh_driver_seed(const bh_driver *d);

/* Host only: Mesh::calculate_photon_energy for the coming step, without
 * touching the device (bh_driver_prepare runs it too). */
int bh_driver_calculate_physics(bh_driver *d);

/* One step in three phases:
 *  prepare:   Mesh::calculate_photon_energy, upload of cells and energies;
 *  transport: device source + transport + census (inputs resident in HBM);
 *  finish:    tally copy-out, Mesh::update_temperature, conservation,
 *             IMC_State::next_time_step. */
int bh_driver_prepare(bh_driver *d);
int bh_driver_transport(bh_driver *d);
int bh_driver_finish(bh_driver *d, bh_driver_report *rep);
int bh_driver_step(bh_driver *d, bh_driver_report *rep);

/* State accessors (n_cells entries; NULL skips an array). */
int bh_driver_get_tallies(const bh_driver *d, double *abs_E, double *track_E);
int bh_driver_get_temperature(const bh_driver *d, double *T_e, double *T_r);
int bh_driver_get_mesh(const bh_driver *d, double *nodes, int32_t *bc, uint32_t *next,
                       uint32_t *region);
/* Cell data and source energies of the step last prepared (with
 * device_physics: of the step last transported; census_E is then not kept
 * and comes back as 0, total_E is that step's E_tot). */
int bh_driver_get_physics(const bh_driver *d, double *sigma_a, double *sigma_s, double *fleck,
                          double *emission_E, double *source_E, double *census_E,
                          int8_t *source_face, double *total_E);
/* The engine handle the driver uses (owned by the driver; created on first
 * use, NULL when no device is available). */
bh_handle *bh_driver_engine(bh_driver *d);

#ifdef __cplusplus
}
#endif
#endif
