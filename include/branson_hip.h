/*
 * branson_hip.h -- C-ABI of the MI355X (gfx950) replicated-mode IMC
 * transport engine.
 *
 * The engine replaces the transport phase of Branson's replicated driver,
 * src/replicated_driver.h:66-88 (source sampling, join with the census,
 * replicated_transport, post-processing into the census), with hand-written
 * HIP kernels.  Plain pointers and sizes only; every entry point returns
 * BH_OK (0) or a negative BH_ERR_* code and never aborts the process; the
 * message of the last failure is available from bh_last_error().
 *
 * Ownership: the handle owns all device buffers; the photon census stays
 * device-resident between steps.  The caller owns every host array it
 * passes.  A handle is not thread safe: one handle per host thread / rank /
 * device.  Every call is synchronous at return unless stated otherwise.
 *
 * Array layouts are cell-major, matching a walk over Branson's
 * std::vector<Cell> (src/cell.h:319-336): nodes[c*6 + face],
 * bc[c*6 + face], next[c*6 + face], sigma_a[c*n_groups + g].
 * Face order is X_NEG, X_POS, Y_NEG, Y_POS, Z_NEG, Z_POS (src/constants.h:27).
 */
#ifndef BRANSON_HIP_H
#define BRANSON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_ABI_VERSION 4

/* return codes */
#define BH_OK 0
#define BH_ERR_ARG (-1)      /* invalid argument or size                   */
#define BH_ERR_DEVICE (-2)   /* HIP runtime / device failure               */
#define BH_ERR_ALLOC (-3)    /* device or host allocation failed           */
#define BH_ERR_STATE (-4)    /* call out of order (e.g. no mesh yet)       */
#define BH_ERR_CAPACITY (-5) /* caller buffer too small                    */

/* Constants::bc_type and Constants::event_type (src/constants.h:26-28) */
enum { BH_REFLECT = 0, BH_VACUUM = 1, BH_ELEMENT = 2, BH_SOURCE = 3, BH_PROCESSOR = 4 };
enum { BH_EXIT = 0, BH_PASS = 1, BH_CENSUS = 2, BH_SCATTER = 3, BH_KILLED = 4, BH_BOUND = 5 };

/* arithmetic of the tracking loop */
enum {
  BH_PRECISION_FP64 = 0,      /* everything FP64: replays the reference     */
  BH_PRECISION_FP32_TRACK = 1 /* FP32 position/direction/path, FP64 energy and
                                 tallies; a history whose next event FP32
                                 cannot decide is finished in FP64, so every
                                 event matches the FP64 form               */
};
/* loop organisation */
enum {
  BH_ALGO_HISTORY = 0,    /* one work-item per photon, loop to completion  */
  BH_ALGO_PERSISTENT = 1, /* persistent waves refill finished lanes        */
  BH_ALGO_EVENT = 2       /* event-split passes over a compacted list of live
                             photons (src/event_based_transport.h:262-305);
                             FP64 only, same events as the other forms     */
};

typedef struct bh_options {
  int32_t precision;     /* BH_PRECISION_*                                  */
  int32_t algorithm;     /* BH_ALGO_*                                       */
  int32_t block_size;    /* tracking work-items per block: 64, 128 or 256;
                            0 = default (256)                              */
  int32_t blocks_per_cu; /* persistent kernel grid per CU, 0 = the blocks  */
                         /* that stay resident together (at most 8)        */
  int32_t tally_mode;    /* 0 auto (block-private LDS where the mesh fits;
                            else, for group-independent opacities, mode 4
                            when the source energy sits in under a quarter
                            of the cells by (sum e)^2 / sum e^2, mode 1
                            otherwise), 1 global atomics, 2 block-private LDS,
                            3 global atomics with the track tally summed per
                            segment even for group-independent opacities,
                            4 global atomics of deposits combined per wave
                            in an LDS table (group-independent opacities;
                            otherwise as 1)                               */
  int32_t refill_threshold; /* persistent: idle lanes of a wave that trigger
                               a refill (1..64), 0 = default (16)          */
  uint64_t reserve_photons; /* photons per launch that the photon stores and
                               every step buffer are sized for at bh_create
                               (steps within it allocate nothing); 0: grow
                               on demand                                  */
  double fp32_tol;          /* FP32 tracking: largest relative error bound a
                               single deposit (and the photon weight) may
                               carry; a history that would exceed it is
                               finished in FP64.  Bounds every cell's tally
                               error against FP64.  0 = default (1e-6)     */
  int32_t batch_photons;    /* persistent: photons a wave claims at a time
                               (1..2^20), 0 = default (256)                */
  int32_t step_graph;       /* host driver: run each step's device phase as
                               one HIP graph (bh_step_graph), 0 = no      */
  int32_t general_mesh;     /* 0: a mesh that is a lattice (rectilinear,
                               numbered x fastest, opacities equal in every
                               group -- any deck-built mesh) is tracked
                               from compact per-cell records and axis
                               arrays; 1: always the general cell records */
  int32_t source_mode;      /* where the persistent algorithm's new photons
                               are made: 0 auto (= 2, measured faster), 1
                               inside the transport kernel as each is taken
                               (no HBM round trip of their state), 2 by the
                               source kernel into the photon store        */
} bh_options;

/* Read-only cell data for one step (src/cell.h:319-336). */
typedef struct bh_cells {
  uint32_t n_cells;
  uint32_t n_groups;
  const double *nodes;   /* [n_cells*6]  x_lo x_hi y_lo y_hi z_lo z_hi      */
  const int32_t *bc;     /* [n_cells*6]  BH_REFLECT .. BH_PROCESSOR         */
  const uint32_t *next;  /* [n_cells*6]  neighbour global index             */
  const double *sigma_a; /* [n_cells*n_groups] absorption opacity (1/cm)    */
  const double *sigma_s; /* [n_cells*n_groups] scattering opacity (1/cm)    */
  const double *fleck;   /* [n_cells] Fleck factor                          */
} bh_cells;

/* Diagnostics of one transport call (the IMC_State setters of
 * src/replicated_transport.h:151-155 plus counters the reference lacks). */
typedef struct bh_step_stats {
  double census_E;        /* post-transport census energy                 */
  double exit_E;          /* energy that left through VACUUM/SOURCE faces */
  uint64_t n_transported; /* photons transported (new + census)           */
  uint64_t n_census;      /* new census size                              */
  uint64_t n_exit;
  uint64_t n_killed;
  uint64_t n_segments;    /* iterations of the tracking loop              */
  uint64_t n_flushes;     /* tally flushes (cell exits and terminations)  */
  double transport_ms;    /* device time of the tracking kernel(s)        */
  double census_ms;       /* device time of census order and totals       */
  double source_ms;       /* device time of the last bh_source call       */
  uint64_t n_replayed;    /* FP32 tracking: histories whose next decision
                             the FP32 error bound could not separate,
                             finished in FP64 from that segment on        */
  int32_t lattice;        /* 1: tracked on the lattice records (see
                             bh_options.general_mesh), 2: on per-instance
                             lattice records (bh_set_instances, every
                             instance a lattice; FP64 persistent waves),
                             0: general records                          */
  int32_t tally;          /* tally form of the last transport: 1 one global
                             atomic per flush, 2 block-private LDS, 3 two
                             global atomics per flush, 4 wave hash table   */
  double replay_ms;       /* FP32 tracking: device time of the FP64 replay,
                             included in transport_ms (0 in FP64)          */
  uint64_t n_allocs;      /* device allocations made by the step's calls
                             (bh_source*, this transport) since the last
                             stats report; each synchronises the device.
                             0 for steps within bh_options.reserve_photons */
} bh_step_stats;

/* A photon in the reference's 120-byte AOS layout (src/photon.h:172-181). */
typedef struct bh_photon {
  uint32_t cell;
  uint32_t group;
  uint32_t source_type;
  uint8_t descriptor[4];
  double pos[3];
  double angle[3];
  double E;
  double E0;
  double life_dx;
  uint64_t rng[4]; /* counter lo, counter hi (seed<<32), key lo (stream), key hi (0) */
} bh_photon;

typedef struct bh_handle bh_handle;

/* Library version / ABI check.  No counterpart in the reference. */
int bh_abi_version(void);

/* Creates a handle on HIP device `device` (the rank -> device mapping of
 * GPU_Setup, src/gpu_setup.h:68-85, is the caller's).  Replaces the
 * per-step GPU_Setup construction at src/replicated_driver.h:64.
 * opts may be NULL (defaults). */
int bh_create(int device, const bh_options *opts, bh_handle **out);
void bh_destroy(bh_handle *h);
const char *bh_last_error(const bh_handle *h);
/* Waits for all work of the handle. */
int bh_sync(bh_handle *h);

/* Uploads the cells for the coming step (opacities and Fleck factor change
 * every step, src/mesh.h:265-270).  Replaces GPU_Setup's device copy of the
 * cell vector (src/gpu_setup.h:19-90). */
int bh_set_mesh(bh_handle *h, const bh_cells *cells);

/* Uploads the per-cell source energies of the step computed by
 * Mesh::calculate_photon_energy (src/mesh.h:272-280): emission, face source
 * and (first step only, else NULL) initial census energy, plus each cell's
 * first SOURCE face or -1 (Cell::get_source_face, src/cell.h:69-76). */
int bh_set_source_energy(bh_handle *h, const double *emission_E,
                         const double *source_E, const double *census_E,
                         const int8_t *source_face);

/* Samples the step's photons on the device: the initial census when census
 * energy was uploaded (make_initial_census_photons, src/source.h:134-180),
 * then emission and face-source photons (make_photons, src/source.h:
 * 207-291), joined ahead of the resident census (join_photon_arrays,
 * src/census_functions.h:20-29).  cycle is the step number; total_E the
 * all-rank photon energy (src/replicated_driver.h:57-59).
 * *n_new (may be NULL) receives the number of new photons.  The call returns
 * once that count is known; the photons are written on the handle's stream,
 * ahead of whatever the handle does next. */
int bh_source(bh_handle *h, uint64_t n_user_photons, double total_E,
              uint32_t seed, uint32_t cycle, uint32_t rank, double dt,
              uint64_t *n_new);

/* Convenience form of the two calls above with the argument list of the
 * reference's source phase. */
int bh_source_from_energy(bh_handle *h, const double *emission_E,
                          const double *source_E, const double *census_E,
                          const int8_t *source_face, uint64_t n_user_photons,
                          double total_E, uint32_t seed, uint32_t cycle,
                          uint32_t rank, double dt, uint64_t *n_new);

/* Several independent problems in one engine (no counterpart in the
 * reference; the batched workload of many problems of uneven size).  The
 * mesh of bh_set_mesh is the concatenation of the instances' cells:
 * instance k owns cells [first_cell[k], first_cell[k+1]) with
 * first_cell[0] = 0 and first_cell[n_instances] = n_cells, and neighbour
 * indices are global.  Each instance draws from its own seed; source,
 * transport and census of all instances then run in the same launches, so
 * short and long problems share the GPU.  Clears the resident census;
 * n_instances = 0 returns to a single problem.  Stays in force across
 * bh_set_mesh calls with the same number of cells. */
int bh_set_instances(bh_handle *h, uint32_t n_instances, const uint32_t *first_cell,
                     const uint32_t *seeds);
/* bh_source for every instance at once: instance k samples with
 * n_user_photons[k] and total_E[k], streams numbered within the instance
 * exactly as bh_source numbers them for a lone problem.  The per-cell
 * energies come from bh_set_source_energy. */
int bh_source_instances(bh_handle *h, const uint64_t *n_user_photons, const double *total_E,
                        uint32_t cycle, uint32_t rank, double dt, uint64_t *n_new);

/* Transports every resident photon (new + census), tallies per-cell absorbed
 * and track-length energy, and rebuilds the census in transport order with
 * life_dx = c * next_dt.  Replaces replicated_transport
 * (src/replicated_transport.h:33-158).  abs_E / track_E (n_cells each) are
 * overwritten when non-NULL; stats may be NULL. */
int bh_transport(bh_handle *h, double next_dt, double *abs_E, double *track_E,
                 bh_step_stats *stats);

/* One step's device phase -- bh_source followed by bh_transport, same
 * arguments and results -- as a single HIP graph launch (no counterpart in
 * the reference: each time step captured as one graph).
 * Every photon count stays on the device: the host computes the step's
 * photon count exactly (the source's own arithmetic) and the graph's first
 * node copies the step's values in, so no host round trip splits the step.
 * The graph is captured on first use and replayed; it is captured again
 * when a buffer is reallocated or the mesh changes.  Persistent algorithm,
 * one problem; the first step (initial census) runs as bh_source +
 * bh_transport.  stats->transport_ms is the whole graph's device time. */
int bh_step_graph(bh_handle *h, uint64_t n_user_photons, double total_E, uint32_t seed,
                  uint32_t cycle, uint32_t rank, double dt, double next_dt, double *abs_E,
                  double *track_E, bh_step_stats *stats);

/* Copies the last step's tallies out (n_cells doubles each; either may be
 * NULL).  Lets a caller keep the copy out of a timed transport call. */
int bh_get_tallies(bh_handle *h, double *abs_E, double *track_E);

/* Number of resident census photons. */
int bh_census_size(bh_handle *h, uint64_t *n);
/* Copies the census out in the reference's 120-byte layout.  *n receives the
 * census size; fails with BH_ERR_CAPACITY when cap is smaller. */
int bh_census_export(bh_handle *h, bh_photon *out, uint64_t cap, uint64_t *n);
/* Copies every resident photon out in the order transport will see them --
 * the step's new photons first, then the census (src/replicated_driver.h:
 * 73-75) -- in the reference's 120-byte layout.  Between bh_source and
 * bh_transport this is the joined all_photons array. */
int bh_photons_export(bh_handle *h, bh_photon *out, uint64_t cap, uint64_t *n);
/* Replaces the resident census (e.g. a census made by the reference). */
int bh_census_import(bh_handle *h, const bh_photon *in, uint64_t n);

/* Drop-in for gpu_transport_photons (src/history_based_transport.h:
 * 348-413): transports the caller's AOS photons in place against the cells
 * of the last bh_set_mesh and accumulates into abs_E/track_E (n_cells each,
 * added to, not overwritten).  Does not touch the resident census. */
int bh_transport_photons(bh_handle *h, bh_photon *photons, uint64_t n,
                         double *abs_E, double *track_E, bh_step_stats *stats);

/* bh_transport_photons without any device allocation (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments and results as bh_transport_photons, except that the
 * tallies always use a form that adds to abs_E / track_E on the device
 * (events identical, tallies equal to round-off).  All of the call's device
 * memory (AOS staging, SoA state, tallies, counters, replay lists) comes
 * from one static 64 MiB array per GPU, held until the process ends: about
 * 270,000 photons per call on a 40^3 mesh; a call that needs more fails with
 * BH_ERR_CAPACITY and the limit in bh_last_error.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (two handles on different
 * threads or streams) race and must be serialised by the caller. */
int bh_transport_photons_noalloc(bh_handle *h, bh_photon *photons, uint64_t n,
                                 double *abs_E, double *track_E,
                                 bh_step_stats *stats);

/* Path-replay adjoint of the absorbed-energy tally (no counterpart in the
 * reference).  For photons in their state BEFORE transport (the input of
 * bh_transport_photons), the cells of the last bh_set_mesh and
 * abs_adjoint[c] = dL/d abs_E[c] of a loss L on the absorbed-energy tally,
 * accumulates (adds) into d_sigma_a[c] and d_fleck[c] the derivative of the
 * expected L with respect to the cell's absorption opacity (all groups
 * together) and Fleck factor, source photons held fixed: the
 * differential-operator (likelihood-ratio) estimator, computed by replaying
 * each history from its Random123 counter, with no trajectory stored.  FP64.
 * Requires sigma_a equal in every group of a cell.  n_cells doubles each. */
int bh_adjoint_photons(bh_handle *h, const bh_photon *photons, uint64_t n,
                       const double *abs_adjoint, double *d_sigma_a, double *d_fleck,
                       bh_step_stats *stats);

/* ---- Census population control ---------------------------------------- */
typedef struct bh_comb_report {
  uint64_t n_before, n_after;  /* census photons before and after          */
  double comb_E;               /* E_census / max_census_photons            */
  double census_E_before, census_E_after;
  int32_t comb_serial;         /* 1: E_census summed serially (a decision
                                  depended on its last bits)               */
  int32_t pad;
} bh_comb_report;

/* comb_photons (src/census_functions.h:48-93) on the resident census: each
 * photon survives with probability E / comb_E (comb_E = E_census /
 * max_census_photons), survivors take energy comb_E, a cell always keeps its
 * last photon when the others are combed away, and each cell's survivors
 * share the cell's energy correction, so every cell's census energy is
 * conserved.  The reference passes an RNG*; here the draws are those of
 * RNG(seed, stream) from counter ctr0, one per census photon in census
 * order.  Survivors keep their census order.  (The reference defines the
 * function but never calls it.) */
int bh_comb_census(bh_handle *h, int64_t max_census_photons, uint32_t seed, uint64_t stream,
                   uint64_t ctr0, bh_comb_report *rep);

/* ---- Device-resident step physics ------------------------------------ */
/* Per-cell material and initial state for step physics on the device: the
 * cell's Region (src/region.h:27-76: opacity law op_a = opA + opB T^opC,
 * scattering opS, heat capacity cv, density rho), its initial electron and
 * radiation temperatures, the temperature of its SOURCE face and that face
 * (Cell::get_source_face, src/cell.h:69-76; -1 for none).  n_cells entries
 * each, for the cells of the last bh_set_mesh. */
typedef struct bh_material {
  const double *opA, *opB, *opC, *opS;
  const double *cv, *rho;
  const double *T_e, *T_r, *T_s;
  const int8_t *source_face;
} bh_material;

/* What one resident step reports (IMC_State::print_conservation quantities,
 * src/imc_state.h:174-289, less the pre-census energy after step 1, which is
 * the previous step's post_census_E). */
typedef struct bh_resident_report {
  double total_E;         /* E_tot of the photon counts (all ranks)        */
  double emission_E, source_E;
  double pre_census_E;    /* step 1: the initial census energy             */
  double pre_mat_E, post_mat_E, absorbed_E;
  bh_step_stats stats;    /* transport and census of the step             */
  uint64_t n_new;         /* new photons sourced this step                 */
  int32_t etot_serial;    /* 1: E_tot was summed serially in cell order (more
                             than one rank, or a photon count that a
                             parallel sum's last bits could have changed);
                             0: parallel sum, every count proven equal     */
  int32_t graph;          /* 1: the step ran as one HIP graph launch       */
} bh_resident_report;

/* Uploads the material (replaces any earlier one).  replicated_factor,
 * n_ranks and rank are the replicated-mode share of this engine
 * (Mesh::calculate_photon_energy, src/mesh.h:237-323). */
int bh_set_material(bh_handle *h, const bh_material *m, double replicated_factor,
                    uint32_t n_ranks, uint32_t rank);

/* One replicated time step with its physics on the device
 * (src/replicated_driver.h:55-101): Mesh::calculate_photon_energy
 * (src/mesh.h:237-323, sums in cell order as the reference's loop), source
 * (initial census at step 1), transport, census and Mesh::update_temperature
 * (src/mesh.h:327-423).  Temperatures, opacities, energies and tallies stay
 * on the device; only the report crosses to the host.  `step` is the
 * reference's cycle number (1 = first step: initial census).  With the
 * persistent algorithm every step after the first is one HIP graph launch
 * with no host round trip inside it (the photon stores are sized before it
 * for the most photons the step can make). */
int bh_step_resident(bh_handle *h, uint64_t n_user_photons, uint32_t seed, uint32_t step,
                     uint32_t rank, double dt, double next_dt, bh_resident_report *rep);

/* Several independent problems (bh_set_instances) stepped with their physics
 * on the device (no counterpart in the reference: the batched workload).
 * The material of bh_set_material holds every instance's cells (one rank,
 * replicated_factor 1); instance k sources n_user_photons[k] photons with
 * its own seed, and its photon energy E_tot is the sum of its cells'
 * energies, summed in a fixed order.  total (may be NULL) receives the
 * sums over instances and the step's transport stats; per_instance (may be
 * NULL) one record per instance.  Steps after the first run as one HIP
 * graph, as bh_step_resident's. */
typedef struct bh_instance_report {
  double total_E, emission_E, source_E;
  double pre_census_E, post_census_E; /* census photons' energy before/after  */
  double absorbed_E, exit_E;
  double pre_mat_E, post_mat_E;
  uint64_t n_new;                     /* photons the instance sourced          */
} bh_instance_report;
int bh_step_resident_instances(bh_handle *h, const uint64_t *n_user_photons, uint32_t step,
                               uint32_t rank, double dt, double next_dt,
                               bh_resident_report *total, bh_instance_report *per_instance);

/* Device-resident state after bh_step_resident: electron temperature and
 * the diagnostic radiation temperature (T_r = (track_E / (V dt a c))^1/4),
 * and the opacities and Fleck factor of the last step.  Any pointer may be
 * NULL; n_cells entries each (sigma: n_cells * n_groups). */
int bh_get_temperature(bh_handle *h, double *T_e, double *T_r);
int bh_get_physics(bh_handle *h, double *sigma_a, double *sigma_s, double *fleck,
                   double *emission_E, double *source_E);

/* Counter-based RNG as the device computes it: n draws of RNG(seed, stream)
 * from counter ctr0 (src/RNG.h:262-342).  Runs on the device. */
int bh_rng_draws(bh_handle *h, uint32_t seed, uint64_t stream, uint64_t ctr0,
                 uint32_t n, double *out);
/* Bandwidth probe (no counterpart in the reference): copies n_bytes between
 * two device buffers with a plain streaming kernel (k_copy_probe), reps times,
 * and returns the mean rate in GB/s (read + write bytes / kernel time).
 * Used to confirm the HBM peak on a machine and to calibrate hardware byte
 * counters against a kernel of known traffic. */
int bh_copy_probe(bh_handle *h, uint64_t n_bytes, int reps, double *gbps);

/* Raw Threefry2x64-20 words on the device for n (ctr, key) pairs
 * (src/random123/threefry.h:190-297): in[4*i..4*i+3] = ctr0 ctr1 key0 key1,
 * out[2*i..2*i+1]. */
int bh_threefry(bh_handle *h, const uint64_t *in, uint64_t n, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
