/*
 * CPU oracle for the CorePotts Metropolis hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C restatement of the reference algorithms, written so the parity
 * tests can check the HIP engine against it. Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it; the product library never links it.
 *
 * Reference files restated here (paths relative to the reference checkout):
 *   lib/CorePotts/src/rng/semantic.jl                      Philox4x32-10, address packing,
 *                                                          uniform_open01, bounded_uint
 *   lib/CorePotts/src/spatial/cartesian.jl:415-525         coordinates and neighbour realization
 *   lib/CorePotts/src/topology/topology.jl:195-209         idx_to_coord
 *   lib/CorePotts/src/protocols/scientific.jl:241-273      copy attempt construction
 *   lib/CorePotts/src/protocols/scientific.jl:372-385      conventional Metropolis acceptance
 *   lib/CorePotts/src/components/scientific_components.jl  volume, contact and boundary terms
 *   lib/CorePotts/src/components/scientific_trackers.jl    staging, commit, retirement
 *   lib/CorePotts/src/components/scientific_inner_loop.jl  energy fold order
 *   lib/CorePotts/src/algorithms/checkerboard.jl           colouring + checkerboard MCS
 *   lib/CorePotts/src/algorithms/lottery.jl                lottery MCS
 *   lib/CorePotts/src/algorithms/sequential.jl:474-596     sequential MCS
 *   lib/CorePotts/src/algorithms/tiled_checkerboard.jl     tiled layout, halo colouring, colour
 *   lib/CorePotts/src/algorithms/tiled_checkerboard_device.jl  order, tile-local proposals and
 *                                                          reconciliation (host recount or
 *                                                          snapshot plus additive deltas)
 *   lib/CorePotts/src/sciml/interface.jl:1009-1026         ensemble_seed
 *   lib/CorePotts/src/components/scientific_mechanics.jl   mechanical OU transition, init,
 *   lib/CorePotts/src/components/scientific_components.jl  advance; mechanical work
 *     :152-240, 576-660;  rng/semantic.jl:323-331          normal_box_muller(Float32)
 *
 * Pinning (see tests/test_oracle_*.py): Philox known answers and the frozen address
 * vector of lib/CorePotts/test/test_rng_semantics.jl:15-37; the literal conflict-claim
 * fixtures of lib/CorePotts/test/test_checkerboard_algorithms.jl:180-196; and the
 * 262,144-replica transition counts archived under
 * design/evidence/phase-13/empirical/{cpu,rocm,metal}/row{3,4,9,10}-*.toml.
 * Lottery has no golden vector in the reference; it is pinned by the reference's
 * statistical tests (lib/CorePotts/test/test_lottery_algorithms.jl:78-181), by
 * invariants, and by brute-force checks on tiny inputs.
 * Parity unpinned: Julia's Base.exp(::Float32) is not available here. The oracle
 * evaluates exp in double precision and rounds once to float (a correctly rounded
 * exp except at exact ties); Julia's result is within 1 ulp of that value. The same
 * holds for the Float32 exp, log and cos of the mechanics (OU factor, Box-Muller); the
 * mechanics are pinned by the reference's closed-form OU value, its one-MCS clock test
 * and its kernel-moment test (lib/CorePotts/test/test_scientific_mechanics.jl).
 */
#ifndef CPM_ORACLE_H
#define CPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_DIRS 128
#define ORC_MAX_TYPES 16
#define ORC_MAX_MEDIA 16

enum { ORC_FACE_PERIODIC = 0, ORC_FACE_CLOSED = 1, ORC_FACE_FIXED = 2 };
enum { ORC_TERM_VOLUME = 0, ORC_TERM_CONTACT = 1, ORC_TERM_BOUNDARY = 2 };
enum { ORC_ALG_SEQUENTIAL = 0, ORC_ALG_CHECKERBOARD = 1, ORC_ALG_LOTTERY = 2, ORC_ALG_TILED = 3 };
enum { ORC_MECH_PRESSURE = 1, ORC_MECH_TENSION = 2 };
enum { ORC_MECH_INIT_CONSTITUTIVE = 1, ORC_MECH_INIT_STATIONARY = 2, ORC_MECH_INIT_PRESERVE = 3 };

typedef struct orc_relation {
    int32_t n;                     /* direction count, canonical order */
    int32_t off[ORC_MAX_DIRS][3];  /* offsets (x, y, z); z = 0 in 2D */
    double weight[ORC_MAX_DIRS];   /* relation weights (values of the relation's eltype) */
} orc_relation;

typedef struct orc_model {
    int32_t ndim;
    int32_t dims[3];
    int32_t face_kind[3][2];        /* [axis][0 = negative, 1 = positive] */
    uint32_t face_medium[3][2];     /* medium owner id of a fixed-exterior face */
    const uint8_t *mutable_mask;    /* NULL: every site mutable */
    const uint8_t *imm_tags;        /* obstacle owners (read only where mask == 0) */
    const uint32_t *imm_ids;

    orc_relation proposal;
    orc_relation contact;           /* contact energy relation (if present) */
    orc_relation surface;           /* boundary tracker relation (always present) */

    int32_t n_energies;
    int32_t energy_order[3];        /* ORC_TERM_* in ScientificComponentSet tuple order */
    int32_t volume_f64;             /* arithmetic type of each term: 0 Float32, 1 Float64 */
    int32_t contact_f64;
    int32_t boundary_f64;

    int32_t n_types;                /* J is n_types x n_types, row-major, 1-based types */
    double J[ORC_MAX_TYPES * ORC_MAX_TYPES];
    int32_t n_table;                /* MediumTypeTable (sorted by owner id) */
    uint32_t table_owner[ORC_MAX_MEDIA];
    uint32_t table_type[ORC_MAX_MEDIA];
    int32_t n_media;                /* state medium_ids (medium volume index order) */
    uint32_t medium_ids[ORC_MAX_MEDIA];

    float temperature;              /* algorithm temperature (Float32) */
    int32_t capacity;
    /* retirement defaults of the property columns */
    double retire_target_volume;
    double retire_volume_strength;
    int64_t retire_target_boundary;
    double retire_boundary_strength;
    /* mechanical components in ScientificComponentSet.mechanics order (Float32) */
    int32_t n_mech;
    int32_t mech_kind[2];           /* ORC_MECH_* */
    int32_t mech_init[2];           /* ORC_MECH_INIT_* */
    uint32_t mech_instance[2];      /* RNG operation of the component's draws */
    float mech_eta[2];
    int32_t mech_noise_is_temperature[2];
    float mech_noise[2];
    /* TiledCheckerboardCPM: tile extent per axis (0: the default 4 x 4 / 2 x 2 x 2) and
     * switching interval (0: the largest tile's mutable site count) */
    int32_t tile_size[3];
    int32_t switching_interval;
} orc_model;

typedef struct orc_state {
    uint8_t *tags;                  /* N, column-major (x fastest) */
    uint32_t *ids;                  /* N */
    int32_t *finite_volumes;        /* capacity */
    int32_t *medium_volumes;        /* n_media */
    int64_t *boundary_measures;     /* capacity */
    uint8_t *active;                /* capacity */
    uint32_t *cell_types;           /* capacity */
    double *target_volume;          /* capacity */
    double *volume_strength;        /* capacity */
    int64_t *target_boundary;       /* capacity */
    double *boundary_strength;      /* capacity */
    float *volume_pressure;         /* capacity (FluctuatingVolumePressure state) */
    float *surface_tension;         /* capacity (FluctuatingSurfaceTension state) */
} orc_state;

/* ---- semantic RNG (rng/semantic.jl) ---- */
void orc_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);
void orc_counter_key(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                     uint32_t operation, uint32_t entity_kind, uint32_t entity,
                     uint64_t generation, uint32_t invocation, uint32_t draw,
                     uint64_t out[3]);
void orc_rng_words(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                   uint32_t operation, uint32_t entity_kind, uint32_t entity,
                   uint64_t generation, uint32_t invocation, uint32_t draw, uint32_t out[4]);
float orc_uniform_open01_f32(uint32_t word1);
uint32_t orc_bounded_uint(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                          uint32_t operation, uint32_t entity_kind, uint32_t entity,
                          uint32_t draw, uint32_t bound);
uint64_t orc_ensemble_seed(uint64_t master_seed, uint32_t trajectory, uint32_t rerun);
/* normal_box_muller(Float32) at (stream, mcs, subround, operation, CellEntity, entity,
 * generation 0, invocation, draw 0) */
float orc_normal_box_muller(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                            uint32_t operation, uint32_t entity, uint32_t invocation);
/* _mechanical_ou_transition in Float32 */
float orc_ou_transition(float q, float displacement, float strength, float eta, float noise,
                        float interval, float standard_normal);

/* Threads of the colour-pass loops of the checkerboard and lottery sweeps (0: every
 * processor); results do not depend on it. */
void orc_set_threads(int32_t threads);
int32_t orc_threads(void);

/* ---- schedules ---- */
/* Greedy realized colouring. colors[N] receives 1-based colours (0 for immutable
 * sites); returns the colour count or a negative error. */
int32_t orc_realized_coloring(const orc_model *m, uint8_t *colors, int32_t *max_degree);
/* Realized conflict degree per site (0 for immutable sites); returns max degree. */
int32_t orc_conflict_degrees(const orc_model *m, uint8_t *degrees);

/* ---- simulation handle ---- */
typedef struct orc_sim orc_sim;
orc_sim *orc_create(const orc_model *m, int32_t algorithm, uint64_t seed);
void orc_destroy(orc_sim *s);
/* _initialize_mechanics! and one _advance_mechanics! call (mcs, subround, phase, interval) */
void orc_initialize_mechanics(orc_sim *s);
void orc_advance_mechanics(orc_sim *s, uint64_t mcs, uint32_t subround, uint32_t phase, float interval);
/* TiledCheckerboardCPM: reconciliation by recount from ownership (the host reference,
 * tiled_checkerboard.jl:371-380) when mode = 1, by snapshot plus additive per-tile deltas
 * (the device path, tiled_checkerboard_device.jl:688-705) when mode = 0 (default). */
void orc_set_tiled_reconcile(orc_sim *s, int32_t mode);
/* Tile layout: grid[3], reach[3], halo; tile_colors[prod(grid)] (1-based) and the colour count. */
int32_t orc_tiled_layout(const orc_sim *s, int32_t grid[3], int32_t reach[3], int32_t *halo,
                         uint16_t *tile_colors);

/* delta_h and mechanical work of one proposal (recipient site, 1-based direction) */
int32_t orc_copy_evaluation(const orc_sim *s, uint32_t recipient, int32_t direction, float out[2]);
/* the energy fold of one proposal in the terms' own types, before evaluate_copy rounds it to
 * Float32 (the per-term energy_change sum the reference's local/global identity compares) */
int32_t orc_copy_energy_fold(const orc_sim *s, uint32_t recipient, int32_t direction, double *out);
orc_state *orc_state_of(orc_sim *s);     /* mutable view of the simulation state */
void orc_set_seed(orc_sim *s, uint64_t seed);
void orc_set_mcs(orc_sim *s, uint64_t mcs);
uint64_t orc_mcs(const orc_sim *s);
int32_t orc_round_count(const orc_sim *s); /* colours (checkerboard) or rounds (lottery) */
int32_t orc_step(orc_sim *s, uint64_t n);
int32_t orc_report(const orc_sim *s, uint64_t out[12]);
int32_t orc_last_order(const orc_sim *s, uint32_t *out); /* last colour/round order */
/* Lottery: activated opportunities of every mutable site (ascending site order) in the last
 * MCS, the reference's workspace.accounting[1:N] (lottery.jl:240-262); returns the count
 * of mutable sites, or -1 for other algorithms. */
int64_t orc_last_activations(const orc_sim *s, uint32_t *out);
/* Copy the logical state in / out (arrays as in orc_state, sized by the model). */
int32_t orc_load_state(orc_sim *s, const orc_state *src);
int32_t orc_store_state(const orc_sim *s, orc_state *dst);
/* Rebuild volumes / boundary measures from ownership and compare with trackers.
 * Returns 0 when every tracker conforms. */
int32_t orc_tracker_errors(const orc_sim *s);

/* Direct claim-protocol check on literal attempts (checkerboard.jl:281-369).
 * losing/gaining are cell ids (0 = medium). Outputs selected[n], dispositions[n],
 * max_priority[cap+1], min_identity[cap+1] (1-based cell index). */
void orc_claims(int32_t n, const uint32_t *recipient, const uint32_t *losing,
                const uint32_t *gaining, const uint32_t *priority, int32_t capacity,
                uint8_t *selected, uint8_t *dispositions, uint32_t *max_priority,
                uint32_t *min_identity);

/* One-MCS transition sampling over independent seeded replicas
 * (integration/transition/ProductionTransitionSampler.jl:97-127). Each replica
 * restarts from the loaded state with mcs 0 and seed seed_base + r. The final
 * lattice is mapped to a catalog index: sum_i label(site_i) * n_labels^(i-1),
 * where label(site) is the position of its owner word in labels[]
 * (owner word = +cell id, -medium id). counts[n_labels^N] accumulate. */
int32_t orc_sample_transitions(orc_sim *s, uint64_t seed_base, uint64_t replicas,
                               const int64_t *labels, int32_t n_labels, uint64_t *counts);

/* Global energies (FP64 accumulations) for statistical checks. */
double orc_total_energy(const orc_sim *s);
/* FP64 sums of the accepted copies' delta_h (the Float32 values the acceptance used):
 * out[0] over the last MCS, out[1] since the state was loaded. */
void orc_accepted_energy(const orc_sim *s, double out[2]);

#ifdef __cplusplus
}
#endif

#endif
