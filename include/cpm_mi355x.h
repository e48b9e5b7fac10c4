/*
 * cpm_mi355x: C-ABI of the MI355X (gfx950) engine for CorePotts' data-parallel
 * Metropolis sweeps (CheckerboardSweepCPM, LotteryCPM and TiledCheckerboardCPM), their Philox4x32-10
 * semantic RNG, copy-attempt energy evaluation and volume/boundary tracker updates.
 *
 * Plain pointers and sizes only. Every function returns 0 on success and a
 * nonzero CPM_E_* code otherwise; the message is available from cpm_last_error().
 * No C++ exception crosses this boundary. A context owns its device memory and
 * one HIP stream; calls on one context are not thread-safe, distinct contexts are.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   cpm_create          init_scientific(state, proposal_relation, components, alg; seed)
 *                       lib/CorePotts/src/algorithms/checkerboard.jl:197-219,
 *                       lib/CorePotts/src/algorithms/lottery.jl:94-115 (workspace
 *                       allocation and realized colouring/conflict graph :159-195 / :46-92)
 *   cpm_upload_state    Adapt.adapt(backend, compile_scientific_state(...)) host->device copy
 *                       lib/CorePotts/src/components/scientific_trackers.jl:210-228,
 *                       lib/CorePotts/src/sciml/interface.jl:607-631
 *   cpm_step            SciMLBase.step!(integrator, n) -> perform_scientific_mcs!
 *                       lib/CorePotts/src/algorithms/sequential.jl:599-609,
 *                       checkerboard.jl:564-575, lottery.jl:328-395, tiled_checkerboard_device.jl:
 *                       743-806 (asynchronous, no host sync)
 *   cpm_report          current_mcs_report(integrator) -> ScientificMCSReport (12 x UInt64)
 *                       sequential.jl:236-251,611-623; checkerboard.jl:577-586; lottery.jl:397-415
 *   cpm_download_state  logical_state(integrator) / logical_snapshot(plan, state)
 *                       sequential.jl:630-632; lib/CorePotts/src/execution/contracts.jl:420-431
 *   cpm_sync            synchronize_observation!(plan)  contracts.jl:366-370
 *   cpm_metrics         ExecutionMetrics counters       contracts.jl:306-316,372-396
 *   cpm_set_seed/mcs    integrator.seed / integrator.mcs fields (sequential.jl:216-234)
 *   cpm_schedule        workspace.color_count/color_order (checkerboard.jl:1-15),
 *                       workspace.round_count/round_order (lottery.jl:2-17)
 *   cpm_ensemble_seed   ensemble_seed(EnsembleSeedDerivationV1(), master, trajectory, rerun)
 *                       lib/CorePotts/src/sciml/interface.jl:1009-1026
 *   cpm_*_mechanics     mechanical state columns and _initialize_mechanics!
 *                       lib/CorePotts/src/components/scientific_mechanics.jl:63-103; the
 *                       half-interval OU steps around every colour pass / lottery round
 *                       (checkerboard.jl:521-558, lottery.jl:353-391) run inside cpm_step
 *   cpm_batch_*         no single counterpart: many independent integrators (the reference's
 *                       EnsembleProblem trajectories, interface.jl:1059-1084) advanced by one
 *                       fused launch per colour pass.
 *
 * Data layout follows the reference's compiled state: lattice arrays are column-major
 * (x fastest), site s (1-based) has 0-based index s-1; owner tags 1 = cell, 2 = medium;
 * per-cell arrays have `capacity` entries where entry c-1 belongs to cell id c.
 */
#ifndef CPM_MI355X_H
#define CPM_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPM_ABI_VERSION 1
#define CPM_MAX_DIRS 128
#define CPM_MAX_TYPES 16
#define CPM_MAX_MEDIA 16

enum cpm_status {
    CPM_OK = 0,
    CPM_E_INVALID = 1,      /* invalid argument or model */
    CPM_E_UNSUPPORTED = 2,  /* valid in the reference but outside this engine */
    CPM_E_DEVICE = 3,       /* HIP runtime error */
    CPM_E_NOMEM = 4,        /* allocation failed or budget exceeded */
    CPM_E_STATE = 5         /* call not valid in the current state */
};

enum cpm_face { CPM_FACE_PERIODIC = 0, CPM_FACE_CLOSED = 1, CPM_FACE_FIXED = 2 };
enum cpm_term { CPM_TERM_VOLUME = 0, CPM_TERM_CONTACT = 1, CPM_TERM_BOUNDARY = 2 };
enum cpm_algorithm { CPM_ALG_CHECKERBOARD = 1, CPM_ALG_LOTTERY = 2, CPM_ALG_TILED = 3 };
/* Mechanical components (lib/CorePotts/src/components/scientific_components.jl:130-240,
 * 575-660) and their initialization (MechanicalInitialization, :130-134). */
enum cpm_mechanics { CPM_MECH_PRESSURE = 1, CPM_MECH_TENSION = 2 };
enum cpm_mech_init {
    CPM_MECH_INIT_CONSTITUTIVE = 1,  /* ConstitutiveMeanInitialization */
    CPM_MECH_INIT_STATIONARY = 2,    /* StationaryMechanicalInitialization */
    CPM_MECH_INIT_PRESERVE = 3       /* PreserveMechanicalInitialization */
};
#define CPM_MAX_MECHANICS 2

/* StaticCartesianRelation (lib/CorePotts/src/spatial/cartesian.jl:252-260), already in
 * canonical direction order; weights hold the values of the relation's eltype. */
typedef struct cpm_relation_desc {
    int32_t count;
    int32_t offsets[CPM_MAX_DIRS][3];
    double weights[CPM_MAX_DIRS];
} cpm_relation_desc;

/* FluctuatingVolumePressure (state column volume_pressure, observable finite volume,
 * target_volume / volume_strength) or FluctuatingSurfaceTension (state column
 * surface_tension, observable boundary measure, target_boundary / boundary_strength),
 * Float32. scientific_components.jl:152-182, 576-616. */
typedef struct cpm_mechanics_desc {
    int32_t kind;                  /* CPM_MECH_* */
    int32_t initialization;        /* CPM_MECH_INIT_* */
    uint32_t instance_id;          /* RNG operation of the component's draws (1..65535) */
    float eta;                     /* relaxation rate, finite and positive */
    int32_t noise_is_temperature;  /* AlgorithmTemperatureNoise: the algorithm temperature */
    float noise;                   /* FixedMechanicalNoise value (non-negative) otherwise */
} cpm_mechanics_desc;

/* Everything init_scientific receives besides the state arrays. */
typedef struct cpm_model_desc {
    int32_t abi_version;           /* CPM_ABI_VERSION */
    int32_t algorithm;             /* CPM_ALG_* */
    int32_t ndim;                  /* 2 or 3 */
    int32_t dims[3];
    int32_t face_kind[3][2];       /* CPM_FACE_* per axis, [0] negative [1] positive face */
    uint32_t face_medium[3][2];    /* medium owner id of a fixed-exterior face */
    const uint8_t *mutable_mask;   /* NULL: every site mutable; else N bytes (1 = mutable) */
    const uint8_t *immutable_tags; /* obstacle owner tags/ids, read where mask == 0 */
    const uint32_t *immutable_ids;

    cpm_relation_desc proposal;    /* ProposalRole relation */
    cpm_relation_desc contact;     /* UnorderedContactHamiltonian relation */
    cpm_relation_desc surface;     /* BoundaryMeasureTracker relation (BoundaryEdgeCount) */

    int32_t n_energies;            /* energies in ScientificComponentSet tuple order */
    int32_t energy_order[3];       /* CPM_TERM_* */
    int32_t volume_f64;            /* arithmetic type of each term: 0 Float32, 1 Float64 */
    int32_t contact_f64;
    int32_t boundary_f64;

    int32_t n_types;               /* contact matrix J (row-major, 1-based type ids) */
    double J[CPM_MAX_TYPES * CPM_MAX_TYPES];
    int32_t n_table;               /* MediumTypeTable: medium owner id -> type id */
    uint32_t table_owner[CPM_MAX_MEDIA];
    uint32_t table_type[CPM_MAX_MEDIA];
    int32_t n_media;               /* state medium_ids (order of the medium volume tracker) */
    uint32_t medium_ids[CPM_MAX_MEDIA];

    float temperature;             /* Float32 algorithm temperature */
    uint32_t capacity;             /* cell slots */
    uint64_t seed;
    double retire_target_volume;   /* retirement defaults of the property columns */
    double retire_volume_strength;
    int64_t retire_target_boundary;
    double retire_boundary_strength;
    int32_t n_mechanics;           /* ScientificComponentSet.mechanics, tuple order */
    cpm_mechanics_desc mechanics[CPM_MAX_MECHANICS];
    /* TiledCheckerboardCPM (tiled_checkerboard.jl:1-60): tile edge per axis (0 = the
     * reference default, 4 in 2D and 2 in 3D) and switching interval (0 = nothing: the
     * site count of the largest tile) */
    int32_t tile_size[3];
    int32_t switching_interval;
} cpm_model_desc;

typedef struct cpm_ctx cpm_ctx;

/* Message of the last failing call on ctx (or of the last failing cpm_create when ctx is NULL). */
const char *cpm_last_error(const cpm_ctx *ctx);
int32_t cpm_abi_version(void);

/* Compile schedules (realized colouring / conflict graph) and allocate device state. */
int32_t cpm_create(const cpm_model_desc *desc, cpm_ctx **out);
int32_t cpm_destroy(cpm_ctx *ctx);

/* Host -> device copy of the compiled state (arrays sized N, capacity or n_media). */
int32_t cpm_upload_state(cpm_ctx *ctx, const uint8_t *tags, const uint32_t *ids,
                         const int32_t *finite_volumes, const int32_t *medium_volumes,
                         const int64_t *boundary_measures, const uint8_t *active,
                         const uint32_t *cell_types, const double *target_volume,
                         const double *volume_strength, const int64_t *target_boundary,
                         const double *boundary_strength);
/* Device -> host copy (synchronizes). Any output pointer may be NULL. */
int32_t cpm_download_state(cpm_ctx *ctx, uint8_t *tags, uint32_t *ids, int32_t *finite_volumes,
                           int32_t *medium_volumes, int64_t *boundary_measures, uint8_t *active,
                           uint32_t *cell_types, double *target_volume, double *volume_strength,
                           int64_t *target_boundary, double *boundary_strength);

/* Mechanical state columns (capacity Float32 each; NULL for a component the model lacks).
 * The property columns volume_pressure and surface_tension of the compiled state. */
int32_t cpm_upload_mechanics(cpm_ctx *ctx, const float *volume_pressure, const float *surface_tension);
/* Device -> host copy (synchronizes). */
int32_t cpm_download_mechanics(cpm_ctx *ctx, float *volume_pressure, float *surface_tension);
/* _initialize_mechanics! (lib/CorePotts/src/components/scientific_mechanics.jl:63-103) on the
 * uploaded state: init_scientific with initialize_mechanics = true. */
int32_t cpm_initialize_mechanics(cpm_ctx *ctx);

int32_t cpm_set_seed(cpm_ctx *ctx, uint64_t seed);
int32_t cpm_set_mcs(cpm_ctx *ctx, uint64_t mcs);
int32_t cpm_get_mcs(const cpm_ctx *ctx, uint64_t *mcs);

/* Enqueue n MCS on the context stream; returns without synchronizing. */
int32_t cpm_step(cpm_ctx *ctx, uint64_t n);
/* Report of the last completed MCS (one launch + one synchronization). Returns
 * CPM_E_STATE when no MCS has run since the state was set. */
int32_t cpm_report(cpm_ctx *ctx, uint64_t out[12]);
int32_t cpm_sync(cpm_ctx *ctx);

/* FP64 sum of the accepted copies' delta_h (the Float32 energy changes the acceptance used)
 * over the last completed MCS, from the same counters as cpm_report (one launch + one
 * synchronization when not cached). Each accepted copy changes the global energy by its
 * delta_h (lib/CorePotts/test/test_scientific_hamiltonians.jl:47-80), so summed over MCS this
 * tracks the energy change for drift checks against cpm_total_energy. CPM_E_UNSUPPORTED for
 * TiledCheckerboardCPM (tile-local views). (No single counterpart in the reference: the
 * per-MCS energy bookkeeping of its evaluation, scientific_inner_loop.jl:257-269.) */
int32_t cpm_report_energy(cpm_ctx *ctx, double *accepted_delta_h);
/* Global energy of the current state in FP64: the model's terms summed as the reference's
 * global_energy methods define them (scientific_components.jl:97-107 volume, 315-340 contact,
 * 525-538 boundary; volume and boundary measures from the trackers), as
 * RealisticScaleRunner's _global_energy adds them. Synchronizes. The first call allocates
 * the context's partial-sum buffer (8 KiB), kept until cpm_destroy. */
int32_t cpm_total_energy(cpm_ctx *ctx, double *energy);
/* cpm_total_energy without any allocation. No counterpart in the reference: this entry point
 * is for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). Same kernels and result as cpm_total_energy; its
 * partial sums live in the library's static device array of kEnergyParts + 1 = 1025 doubles
 * (8 KiB per device, held until the process ends), whose address cpm_create resolved. The
 * reduction's grid is fixed, so no lattice exceeds that maximum. Every call on one device
 * shares the array: calls that overlap on one device (other threads, contexts or streams)
 * race. CPM_E_STATE when the context's device is not the current device. Synchronizes. */
int32_t cpm_total_energy_noalloc(cpm_ctx *ctx, double *energy);

/* Colour (checkerboard), round (lottery) or tile-colour (tiled) count, and the order used
 * by the last MCS. order may be NULL; it needs round_count entries. Synchronizes. */
int32_t cpm_schedule(cpm_ctx *ctx, int32_t *round_count, uint32_t *order);
/* Realized colouring of a model, computed on the host as cpm_create computes it (no device
 * needed): the greedy colour of every site in ascending site order (checkerboard.jl:57-90) or
 * the halo colour of its tile (tiled), 0 for immutable sites; N bytes, and the colour count.
 * Not for lottery. Errors as cpm_create. */
int32_t cpm_model_colors(const cpm_model_desc *desc, uint8_t *colors, int32_t *count);
/* TiledCheckerboardCPM layout of a model, computed on the host (no device needed):
 * tiled_layout(algorithm, dims; halo_radius, periodic) (tiled_checkerboard.jl:166-188) with
 * the halo of _tiled_required_halo (:258-296) and the device batch schedule
 * (tiled_checkerboard_device.jl:28-121). grid and reach per axis, halo radius, the halo
 * colour of every tile (grid[0]*grid[1]*grid[2] entries, first axis fastest; may be NULL),
 * colour count, switching interval, batches per colour and the positions of the box a lane
 * stages in LDS (0: the kernel that reads the lattice in global memory). Errors as cpm_create. */
int32_t cpm_tiled_layout(const cpm_model_desc *desc, int32_t grid[3], int32_t reach[3], int32_t *halo,
                         uint16_t *tile_colors, int32_t *color_count, int32_t *interval, int32_t *batches,
                         int32_t *staged);
/* Process-wide choice of the TiledCheckerboardCPM batch kernel for models compiled after the
 * call: nonzero (default) stages each tile's box and halo in LDS where it holds at most 64
 * positions, zero always reads the lattice in global memory. Both give identical results. No
 * reference counterpart (an implementation choice, like cpm_set_row_kernels). */
int32_t cpm_set_tiled_staging(int32_t enable);
/* Realized colour of every site (0 for immutable sites), N bytes: the site's own colour
 * (checkerboard) or the halo colour of its tile (tiled). Not for lottery. */
int32_t cpm_site_colors(const cpm_ctx *ctx, uint8_t *colors);
/* Kernel configuration of the context's lattice: tile shape id (bits 0-3), kernels for
 * proposal relations of other than 2 * ndim directions (bit 4), row propose kernel (bit 5;
 * with bit 6: an x side of 2 mod 4), lottery kernel for periodic obstacle-free lattices
 * (bit 6 without bit 5), 256 for TiledCheckerboardCPM;
 * bits 8-11: the commit kernel's specialised evaluation (0: general; 1-3 2D with 0 / 4 / 8
 * contact directions, 4-6 3D with 0 / 6 / 26).
 * (No counterpart in the reference: kernel selection of this engine.) */
int32_t cpm_kernel_config(const cpm_ctx *ctx, int32_t *config);

/* ExecutionMetrics mirror: [launches, host_synchronizations, h2d_transfers,
 * d2h_transfers, device_bytes_allocated, graph_launches]. */
int32_t cpm_metrics(const cpm_ctx *ctx, uint64_t out[6]);

/* Like cpm_step, but attaches a HIP event pair to every kernel launch on the context stream
 * (the pair brackets that kernel's execution) and synchronizes at the end. out[0..7] =
 * propose-kernel total ms (TiledCheckerboardCPM: batch kernel), propose launches,
 * commit-kernel total ms (tiled: reconciliation), commit launches, MCS-order kernels' total ms
 * (the plan and end launches around each chunk of up to 64 MCS, or the begin launch of every MCS
 * where a TiledCheckerboardCPM replica runs), their launches, wall ms of the whole run, MCS run. Mechanics kernels are not in any of the
 * kernel totals. (No counterpart in the reference: the measurement hook behind the per-kernel
 * roofline.) */
int32_t cpm_step_timed(cpm_ctx *ctx, uint64_t n, double out[8]);

/* Tile shapes of the propose kernels (the autotuner's search space): out = {dim, halo,
 * TX, TY, TZ} of shape id 0..15; CPM_E_INVALID for ids without a shape. (No counterpart in
 * the reference: launch geometry of this engine.) */
int32_t cpm_tile_shape_info(int32_t shape, int32_t out[5]);
/* Process-wide preferred tile shape for replicas of (algorithm, ndim, halo) created after
 * the call; shape -1 restores the default. The shape's dim and halo must match. Results do
 * not depend on the shape. (No counterpart in the reference.) */
int32_t cpm_set_tile_shape(int32_t algorithm, int32_t ndim, int32_t halo, int32_t shape);
/* Process-wide switch (default on) of the row propose kernels, used for replicas created
 * after the call whose model is a 2D periodic obstacle-free checkerboard with the canonical
 * von Neumann proposal relation and a 2 x 2 colour pattern. Results do not depend on it.
 * (No counterpart in the reference: kernel selection of this engine.) */
int32_t cpm_set_row_kernels(int32_t enable);

/* Device-to-device copy of `bytes` bytes (a multiple of 16) with a plain 16-byte-per-lane
 * kernel, launched once untimed and `reps` times timed; *out_ms = mean launch time. Errors
 * are reported through cpm_last_error(NULL). (No counterpart in the reference: the
 * calibration kernel of known traffic for hardware counters and the achievable-bandwidth
 * figure beside the roofline.) */
int32_t cpm_copy_probe(uint64_t bytes, int32_t reps, double *out_ms);

/* Capture chunks of `chunk` MCS into a HIP graph that cpm_step replays when n is a
 * multiple of chunk (0 disables graphs). */
int32_t cpm_set_graph_chunk(cpm_ctx *ctx, uint32_t chunk);
/* Capture the HIP graph that the next cpm_step calls replay, performing no MCS, so that
 * capture and instantiation stay out of a caller's timed loop. *ready = 1 when the next
 * cpm_step of at least one chunk starts with that graph, 0 when graphs are off or the claim
 * buffers' parity makes it start with one MCS of plain launches. (No counterpart in the
 * reference: launch management of this engine.) */
int32_t cpm_prepare_graph(cpm_ctx *ctx, int32_t *ready);

/* Philox-derived replica seed (sciml/interface.jl:1009-1026). */
uint64_t cpm_ensemble_seed(uint64_t master_seed, uint32_t trajectory, uint32_t rerun);

/* ---------------- batched replicas and served jobs ----------------
 * A batch holds many independent replicas (possibly of different sizes, temperatures,
 * contact matrices and algorithms of one kind) and advances every replica that has not
 * reached its target MCS with one fused launch per colour pass. Replicas can be added
 * and removed between steps (continuous batching); each replica's trajectory is
 * identical to a single-context run with the same model, state and seed. */
typedef struct cpm_batch cpm_batch;

int32_t cpm_batch_create(int32_t algorithm, uint64_t hbm_budget_bytes, cpm_batch **out);
int32_t cpm_batch_destroy(cpm_batch *b);
const char *cpm_batch_last_error(const cpm_batch *b);
/* Device bytes a replica of this model would hold (for admission by budget). */
int32_t cpm_replica_bytes(const cpm_model_desc *desc, uint64_t *bytes);
/* Add a replica; fails with CPM_E_NOMEM when the budget would be exceeded. */
int32_t cpm_batch_add(cpm_batch *b, const cpm_model_desc *desc, int32_t *slot);
int32_t cpm_batch_remove(cpm_batch *b, int32_t slot);
int32_t cpm_batch_upload_state(cpm_batch *b, int32_t slot, const uint8_t *tags,
                               const uint32_t *ids, const int32_t *finite_volumes,
                               const int32_t *medium_volumes, const int64_t *boundary_measures,
                               const uint8_t *active, const uint32_t *cell_types,
                               const double *target_volume, const double *volume_strength,
                               const int64_t *target_boundary, const double *boundary_strength);
int32_t cpm_batch_download_state(cpm_batch *b, int32_t slot, uint8_t *tags, uint32_t *ids,
                                 int32_t *finite_volumes, int32_t *medium_volumes,
                                 int64_t *boundary_measures, uint8_t *active,
                                 uint32_t *cell_types, double *target_volume,
                                 double *volume_strength, int64_t *target_boundary,
                                 double *boundary_strength);
int32_t cpm_batch_upload_mechanics(cpm_batch *b, int32_t slot, const float *volume_pressure,
                                   const float *surface_tension);
int32_t cpm_batch_download_mechanics(cpm_batch *b, int32_t slot, float *volume_pressure,
                                     float *surface_tension);
int32_t cpm_batch_initialize_mechanics(cpm_batch *b, int32_t slot);
/* Set a replica's MCS counter (checkpoint restore; integrator.mcs, sequential.jl:216-234);
 * its target becomes the same value, so it stays idle until the next cpm_batch_set_target. */
int32_t cpm_batch_set_mcs(cpm_batch *b, int32_t slot, uint64_t mcs);
/* The replica advances until its MCS counter reaches target_mcs. */
int32_t cpm_batch_set_target(cpm_batch *b, int32_t slot, uint64_t target_mcs);

/* Device-resident snapshots. A snapshot holds one replica's state at its last completed MCS in
 * device memory: owner words, cell records (trackers, property columns, mechanical state),
 * medium volumes, the MCS counter and the seed, which is everything exact continuation needs
 * (persistence/scientific.jl:464-468: state, seed, MCS). Capture and restore are device-to-device
 * copies with no host transfer, so a served job can be pre-empted into HBM at about 4 bytes per
 * site plus 64 per cell and resumed at once. The canonical checkpoint of
 * persistence/scientific.jl:250-292 is built on the host from cpm_snapshot_download. (The
 * reference keeps checkpoints on the host only; this is the engine's device-side form.) */
typedef struct cpm_snapshot cpm_snapshot;
/* Capture the context's replica (synchronizes). *out is freed with cpm_snapshot_destroy. */
int32_t cpm_snapshot_capture(cpm_ctx *ctx, cpm_snapshot **out);
int32_t cpm_batch_snapshot_capture(cpm_batch *b, int32_t slot, cpm_snapshot **out);
/* Copy a snapshot into a replica with the same lattice, capacity, media and cell types, and set
 * its seed and MCS counter to the snapshot's (a batch replica then waits for a target). */
int32_t cpm_snapshot_restore(cpm_ctx *ctx, const cpm_snapshot *snapshot);
int32_t cpm_batch_snapshot_restore(cpm_batch *b, int32_t slot, const cpm_snapshot *snapshot);
/* The snapshot's state in cpm_download_state's arrays and order (null pointers skip arrays), and
 * its mechanical state as cpm_download_mechanics. Errors are reported by cpm_last_error(NULL). */
int32_t cpm_snapshot_download(const cpm_snapshot *snapshot, uint8_t *tags, uint32_t *ids,
                              int32_t *finite_volumes, int32_t *medium_volumes,
                              int64_t *boundary_measures, uint8_t *active, uint32_t *cell_types,
                              double *target_volume, double *volume_strength,
                              int64_t *target_boundary, double *boundary_strength);
int32_t cpm_snapshot_download_mechanics(const cpm_snapshot *snapshot, float *volume_pressure,
                                        float *surface_tension);
/* MCS counter, seed and device bytes of a snapshot. */
int32_t cpm_snapshot_info(const cpm_snapshot *snapshot, uint64_t *mcs, uint64_t *seed, uint64_t *bytes);
void cpm_snapshot_destroy(cpm_snapshot *snapshot);
/* Enqueue up to n MCS for every replica below its target (asynchronous). */
int32_t cpm_batch_step(cpm_batch *b, uint64_t n);
/* Current MCS of every slot (synchronizes); out needs cpm_batch_capacity entries,
 * UINT64_MAX marks free slots. */
int32_t cpm_batch_progress(cpm_batch *b, uint64_t *out, int32_t capacity);
int32_t cpm_batch_capacity(const cpm_batch *b);
int32_t cpm_batch_report(cpm_batch *b, int32_t slot, uint64_t out[12]);
int32_t cpm_batch_sync(cpm_batch *b);
int32_t cpm_batch_set_graph_chunk(cpm_batch *b, uint32_t chunk);
/* Resident device bytes of the batch (replica state + shared scratch). */
int32_t cpm_batch_bytes(const cpm_batch *b, uint64_t *bytes);

#ifdef __cplusplus
}
#endif

#endif
