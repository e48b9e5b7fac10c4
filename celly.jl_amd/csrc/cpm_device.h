// Device-side data structures and primitives of the MI355X CPM engine.
//
// Owner encoding on the device: one int32 word per site, id | type << kTypeShift for a cell
// (positive), -id for a medium; 0 never names an owner and marks an absent neighbour inside
// LDS tiles. A cell's type is fixed while it owns sites, so OwnerRef equality
// (lib/CorePotts/src/state/logical.jl:9-21) is still word equality, and the contact energy reads
// a neighbour cell's type from its word instead of from a per-cell table.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdint.h>

#include "../../include/cpm_mi355x.h"

namespace cpm {

constexpr int kTypeShift = 26;                       // cell words: id in bits 0-25, type in 26-30
constexpr uint32_t kIdMask = (1u << kTypeShift) - 1u;  // also the largest cell capacity
__host__ __device__ __forceinline__ uint32_t word_cell(int32_t w) { return (uint32_t)w & kIdMask; }
__host__ __device__ __forceinline__ uint32_t word_type(int32_t w) { return (uint32_t)w >> kTypeShift; }
// per-cell slot of an owner word: the cell id, 0 (the spare slot) for a medium
__host__ __device__ __forceinline__ uint32_t cell_slot(int32_t w) { return w > 0 ? word_cell(w) : 0u; }

// Canonical direction order of the Moore shells (cartesian.jl:324-328: squared length, then
// the offset tuple); their first 2 * dim directions are the first (von Neumann) shell. The
// commit kernel's specialised evaluations use these neighbourhoods with compile-time offsets.
constexpr int8_t kMoore2[8][2] = {{-1, 0}, {0, -1}, {0, 1}, {1, 0}, {-1, -1}, {-1, 1}, {1, -1}, {1, 1}};
constexpr int8_t kMoore3[26][3] = {
    {-1, 0, 0},  {0, -1, 0},  {0, 0, -1}, {0, 0, 1},  {0, 1, 0},   {1, 0, 0},  {-1, -1, 0}, {-1, 0, -1}, {-1, 0, 1},
    {-1, 1, 0},  {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1},   {1, -1, 0}, {1, 0, -1},  {1, 0, 1},   {1, 1, 0},
    {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}, {1, -1, 1}, {1, 1, -1}, {1, 1, 1}};
template <int DIM>
__host__ __device__ constexpr int moore_offset(int k, int axis) {
    return DIM == 2 ? (axis < 2 ? kMoore2[k][axis] : 0) : kMoore3[k][axis];
}

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kCommitTiles = 4;    // tile segments per commit workgroup (claim check, evaluation)
constexpr int kMaxRounds = 256;    // colour / round order capacity per replica
constexpr int kSlotFields = 10;    // per-tile accounting fields (8-9: FP64 sum of the accepted
                                   // copies' delta_h of the tile's pass)
constexpr uint32_t kEnergyParts = 1024;  // partial sums of the global-energy reduction

// Semantic RNG stream ids (lib/CorePotts/src/rng/semantic.jl:37-58).
enum : uint32_t {
    kStreamRecipient = 3,
    kStreamDirection = 4,
    kStreamAcceptance = 5,
    kStreamLotteryActivation = 6,
    kStreamLotteryPriority = 7,
    kStreamCheckerboardOrder = 8,
    kStreamCheckerboardPriority = 17,
    kStreamAuxEvolution = 9,
    kStreamAuxInitialization = 18,
    kStreamTiledOrder = 19,
    kStreamTiledProposal = 20,
};
enum : uint32_t { kGlobalEntity = 0, kSiteEntity = 1, kCellEntity = 2 };

// Division by a runtime-uniform divisor via a precomputed reciprocal: q = umulhi(n, m)
// with m = floor(2^32 / d) undershoots the quotient by at most one.
struct FastDiv {
    uint32_t d;
    uint32_t m;
};

// Bounded draw constants: rejection threshold (2^32 - bound) mod bound and divisor.
struct Bound {
    uint32_t threshold;
    FastDiv div;
};

// Relation offsets compiled for the device (canonical direction order kept).
struct RelDev {
    int32_t n;
    int8_t off[CPM_MAX_DIRS][3];
    float wf[CPM_MAX_DIRS];
    double wd[CPM_MAX_DIRS];
};

// Static description of one replica (lattice, model, schedule and buffers).
// Per-cell state in one 64-byte record, so a copy attempt reads each of its two cells
// with one cache line: trackers (volume, boundary measure), type, activity and the
// property columns of the volume and boundary energies. The host converts to and from the
// reference's per-cell arrays at upload and download.
struct __attribute__((aligned(64))) CellRec {
    int32_t vol;       // finite volume
    uint32_t ctype;    // cell type (0 when inactive)
    int64_t bnd;       // boundary measure
    double tv;         // target_volume
    double vs;         // volume_strength
    int64_t tb;        // target_boundary
    double bs;         // boundary_strength
    float qp;          // volume_pressure (FluctuatingVolumePressure state)
    float qt;          // surface_tension (FluctuatingSurfaceTension state)
    uint8_t active;
    uint8_t pad[7];
};
static_assert(sizeof(CellRec) == 64, "one cache line per cell");
static_assert(offsetof(CellRec, bnd) == 8 && offsetof(CellRec, tv) == 16 && offsetof(CellRec, vs) == 24 &&
                  offsetof(CellRec, tb) == 32 && offsetof(CellRec, bs) == 40 && offsetof(CellRec, qp) == 48 &&
                  offsetof(CellRec, qt) == 52 && offsetof(CellRec, active) == 56,
              "load_cell reads the fields at these offsets");

struct RepDev {
    int32_t ndim;
    int32_t dims[3];
    uint32_t nsites;
    int32_t face_kind[3][2];
    int32_t face_word[3][2];     // owner word of fixed faces
    int32_t simple;              // every face periodic, no obstacles, dims >= halo
    int32_t ntiles[3];
    uint32_t tile_base;          // first tile of this replica inside its group
    uint32_t tile_count;
    int32_t algorithm;
    int32_t rounds;              // colours (checkerboard) or lottery rounds
    int32_t uniform_degree;      // lottery: every site has rounds-1 conflict neighbours
    float temperature;
    uint32_t key_lo, key_hi;     // Philox key = seed ^ mix64(generation domain)
    int32_t pat[3];              // periodic colour pattern periods (0: per-site array)
    int32_t pat_aligned;         // every period divides the tile side: class offsets = residues
    uint8_t pattern[64];
    uint8_t pat_coord[64][3];    // residue coordinates of each pattern index
    // aligned patterns whose residue classes have power-of-two extents inside a tile:
    // colour c owns residues col_res[col_first[c] .. + col_count[c]), each class holds
    // 2^cls_shift tile sites laid out 2^cls_sx along x by 2^cls_sy along y
    int32_t pat_fast;
    int32_t cls_shift, cls_sx, cls_sy;
    uint8_t col_first[65], col_count[65];
    uint8_t col_res[64];
    int32_t prop_lds[8];         // first proposal offsets as LDS position deltas of the tile shape
    int32_t halo;                // largest |offset| component over the relations
    int32_t surf_lin[32];        // first surface / contact offsets as linear site deltas, valid
    int32_t cont_lin[32];        // for sites at least `halo` away from every face
    Bound prop_bound;            // direction draw over the proposal relation
    Bound round_bound;           // lottery activation draw over the round count
    FastDiv div_x, div_xy;       // site -> coordinates
    int32_t n_energies;
    int32_t energy_order[3];
    int32_t volume_f64, contact_f64, boundary_f64;
    int32_t has_contact;
    // Specialised evaluation of the commit kernel (0: the general one): 1-3 2D with the von
    // Neumann surface relation and no / von Neumann / Moore contact relation, 4-6 the same in
    // 3D (relations in canonical order, kMoore2 / kMoore3); periodic obstacle-free lattices,
    // Float32 terms, no mechanics.
    int32_t fast_eval;
    int32_t n_types;
    float Jf[CPM_MAX_TYPES * CPM_MAX_TYPES];
    double Jd[CPM_MAX_TYPES * CPM_MAX_TYPES];
    int32_t n_table;
    int32_t table_word[CPM_MAX_MEDIA];
    uint32_t table_type[CPM_MAX_MEDIA];
    int32_t n_media;
    int32_t medium_word[CPM_MAX_MEDIA];
    double retire_tv, retire_vs, retire_bs;
    int64_t retire_tb;
    uint32_t claim_base;         // first claim slot of this replica (cell id 0)
    RelDev prop, cont, surf, conf;
    // device buffers
    int32_t *lattice;
    const uint8_t *mask;         // nullptr: all mutable
    const uint16_t *colour_list; // colourings without pat_fast: tile positions by tile and colour
    const uint32_t *colour_off;  // [tile][rounds + 1] ranges into colour_list
    const uint8_t *site_degree;  // lottery, nullptr: uniform degree
    uint32_t capacity;           // cell slots (ids 1..capacity)
    // mechanics (FluctuatingVolumePressure / FluctuatingSurfaceTension, tuple order)
    int32_t n_mech;
    int32_t mech_kind[2];        // CPM_MECH_*
    int32_t mech_init[2];        // CPM_MECH_INIT_*
    uint32_t mech_instance[2];   // RNG operation
    float mech_eta[2];
    float mech_noise[2];         // noise scale (the temperature for AlgorithmTemperatureNoise)
    // half interval of the OU steps around a pass: by colour (checkerboard,
    // size / (2 * mutable sites)) or entry 0 (lottery, 1 / (2 * rounds)), Float32
    float mech_half[kMaxRounds + 1];
    CellRec *cells;              // cap + 1, index = cell id (entry 0: spare read for media)
    int32_t *medium_vol;         // n_media
    // TiledCheckerboardCPM: halo-coloured tiles (1-based ids, first axis fastest), the
    // mutable sites of each tile (x fastest inside a tile), the tiles of each colour, the
    // per-cell tracker changes of one colour batch and the counters of the current MCS
    int32_t tiled_interval, tiled_batches;
    int32_t tile_edge[3], tile_grid[3];
    int32_t tiled_halo;             // radius of the declared accesses
    int32_t tiled_stage[3];         // staged box per tile: edge + 2 halo (0 when not staged)
    const uint32_t *tile_sites;     // mutable sites (1-based), grouped by tile
    const uint32_t *tile_offsets;   // tiles + 1 ranges into tile_sites
    const uint32_t *color_tiles;    // tiles (1-based) grouped by colour, ascending
    const uint32_t *color_offsets;  // colours + 1 ranges into color_tiles
    int32_t *tile_dvol;             // cap + 1 volume changes (summed over the batch's tiles)
    int32_t *tile_dbnd;             // cap + 1 boundary-measure changes
    int32_t *tile_spill;            // touched-cell entries past the LDS list, per tile of a colour
    int32_t tile_spill_stride;      // words per tile: 2 * (2 * interval - LDS entries), or 0
    unsigned long long *tiled_counts;  // realized, same owner, boundary null, rejected, accepted
};

// Per-replica dynamic state, updated on the device by the begin-MCS kernel (one MCS at a time)
// or by the plan and end kernels around a planned chunk of MCS.
struct RepDyn {
    uint64_t mcs;      // MCS being performed (after begin) / last completed
    uint64_t target;   // replica advances while mcs < target
    uint32_t active;   // performing an MCS in the current step
    uint32_t in_use;
    uint32_t order[kMaxRounds];
};

// What one replica does in the k-th MCS of a planned chunk: the MCS number (1-based, as in the
// RNG addresses), whether it performs it (in use and below its target) and that MCS's colour or
// round order. A plan depends only on the replica's MCS counter, target and key at the start of
// the chunk, so one launch writes the plans of a whole chunk; the pass kernels of its k-th MCS
// read plan k instead of a per-MCS state update.
constexpr int kPlanMcs = 64;  // MCS one plan launch covers
struct McsPlan {
    uint64_t mcs;
    uint32_t active;
    uint32_t pad;
    uint16_t order[kMaxRounds];
};

// One actionable copy attempt staged for the claim-resolution / commit kernel.
struct Entry {
    uint32_t site;      // recipient (1-based linear site id)
    uint32_t priority;  // claim priority (Philox word 1)
    int32_t losing;     // owner word of the recipient
    int32_t gaining;    // owner word of the donor
};

// Four packed int32 lanes: one 16-byte global load or store.
typedef int32_t v4i __attribute__((ext_vector_type(4)));

// Per-tile launch data: replica slot, tile origin and the replica's first claim word.
struct __align__(16) TileInfo {
    uint32_t rep;
    int32_t x0, y0, z0;
    uint32_t claim_base;
    uint32_t pad[3];
};

// Stretch count written by a propose kernel for a tile whose replica does not run this
// pass (inactive replica, or fewer colours / rounds than the group's maximum).
constexpr uint32_t kIdleTile = 0xffffffffu;

struct PassArgs {
    const RepDev *reps;
    RepDyn *dyn;
    const TileInfo *tiles;      // group tile -> replica slot and origin
    unsigned long long *claims; // two buffers of total_claims each
    Entry *seg;                 // per-tile segments of seg_cap entries
    uint32_t *seg_count;        // [tile][kWaves] stretch lengths (kIdleTile: not running)
    uint32_t *slots;            // [pass][tile][kSlotFields]
    uint32_t ntiles;
    uint32_t seg_cap;
    uint32_t pass;              // ordinal inside the MCS (0-based)
    uint32_t parity;            // claim buffer used by this pass
    uint32_t total_claims;
    uint32_t fast_eval;         // RepDev::fast_eval shared by every replica of the group (0: mixed)
    const McsPlan *plan;        // [slot][kPlanMcs]
    uint32_t plan_k;            // MCS of the chunk this pass belongs to
};

// Device buffers are reached through pointers stored in RepDev; casting them to the
// global address space lets the compiler emit global_* (not flat_*) memory operations,
// which do not wait on the LDS counter.
#define CPM_GLOBAL __attribute__((address_space(1)))
#define CPM_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ CPM_GLOBAL T *gp(T *p) {
    return (CPM_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ const CPM_GLOBAL T *gp(const T *p) {
    return (const CPM_GLOBAL T *)p;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (rng/semantic.jl:136-187) and the v1 address packing (:249-261)
// ---------------------------------------------------------------------------
struct Words {
    uint32_t w0, w1, w2, w3;
};

// a ^ b ^ c in one VALU instruction: v_bitop3_b32 with truth table 0x96 (odd parity of the
// three input bits, whatever their order), which the compiler does not form for these XORs.
// All three operands are vector registers: the round keys are loop-invariant and the compiler
// hoists their copies out of the draw loops. (With the key as a scalar operand the compiler
// put round keys in the scalar registers that the multiplies' carry-outs write, and results
// were wrong.)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Rounds 1-9: two 32 x 32 -> 64-bit multiplies and two three-input XORs. Round 0 is left to
// the compiler (the counter words other than the entity are often uniform there).
__device__ __forceinline__ Words philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                          uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9e3779b9u;
            k1 += 0xbb67ae85u;
        }
        const uint64_t p0 = (uint64_t)0xd2511f53u * c0;
        const uint64_t p1 = (uint64_t)0xcd9e8d57u * c2;
        uint32_t n0, n2;
        if (r == 0) {
            n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
            n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        } else {
            n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
            n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        }
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    return Words{c0, c1, c2, c3};
}

// counter0 = mcs | subround << 48 | invocation << 56
// counter1 = entity | operation << 32 | stream << 44 | kind << 52 | draw << 54
__device__ __forceinline__ Words rng_words(uint32_t key_lo, uint32_t key_hi, uint32_t stream,
                                           uint64_t mcs, uint32_t subround, uint32_t operation,
                                           uint32_t kind, uint32_t entity, uint32_t invocation,
                                           uint32_t draw) {
    const uint32_t c0 = (uint32_t)mcs;
    const uint32_t c1 = ((uint32_t)(mcs >> 32) & 0xffffu) | ((subround & 0xffu) << 16) |
                        ((invocation & 0xffu) << 24);
    const uint32_t c2 = entity;
    const uint32_t c3 = operation | (stream << 12) | (kind << 20) | (draw << 22);
    return philox10(c0, c1, c2, c3, key_lo, key_hi);
}

__device__ __forceinline__ uint32_t fast_mod(uint32_t n, FastDiv f) {
    if (f.d == 1) return 0;
    const uint32_t q = __umulhi(n, f.m);
    uint32_t r = n - q * f.d;
    return r >= f.d ? r - f.d : r;
}
__device__ __forceinline__ uint32_t fast_div(uint32_t n, FastDiv f) {
    if (f.d == 1) return n;
    uint32_t q = __umulhi(n, f.m);
    return n - q * f.d >= f.d ? q + 1 : q;
}

// bounded_uint with host-precomputed threshold and divisor (rng/semantic.jl:298-310).
__device__ __forceinline__ uint32_t bounded_draw(uint32_t key_lo, uint32_t key_hi, uint32_t stream,
                                                 uint64_t mcs, uint32_t subround, uint32_t operation,
                                                 uint32_t kind, uint32_t entity, uint32_t draw,
                                                 Bound b) {
    uint32_t invocation = 0;
    while (true) {
        const Words w = rng_words(key_lo, key_hi, stream, mcs, subround, operation, kind,
                                  entity, invocation, draw);
        if (w.w0 >= b.threshold || invocation == 255u) return fast_mod(w.w0, b.div);
        ++invocation;
    }
}

// bounded_uint (rng/semantic.jl:298-310): reject below (2^32 - bound) mod bound,
// retrying with invocation + 1.
__device__ __forceinline__ uint32_t bounded_uint(uint32_t key_lo, uint32_t key_hi,
                                                 uint32_t stream, uint64_t mcs,
                                                 uint32_t subround, uint32_t operation,
                                                 uint32_t kind, uint32_t entity, uint32_t draw,
                                                 uint32_t bound) {
    const uint32_t threshold = (0u - bound) % bound;
    uint32_t invocation = 0;
    while (true) {
        const Words w = rng_words(key_lo, key_hi, stream, mcs, subround, operation, kind,
                                  entity, invocation, draw);
        if (w.w0 >= threshold || invocation == 255u) return w.w0 % bound;
        ++invocation;
    }
}

// uniform_open01(Float32) (rng/semantic.jl:276-280), all Float32 operations.
__device__ __forceinline__ float uniform_open01(uint32_t w0) {
    const float bits = (float)(w0 >> 8);
    return __fmul_rn(__fadd_rn(bits, 0.5f), 0x1.0p-24f);
}

// Fisher-Yates schedule permutation (checkerboard.jl:226-242, lottery.jl:124-140).
template <class T>
__device__ inline void schedule_order(uint32_t key_lo, uint32_t key_hi, uint32_t stream,
                                      uint32_t operation, uint64_t mcs, int count,
                                      T *order) {
    constexpr int kRegs = 16;
    if (count <= kRegs) {
        // the permutation in registers (swaps by selects), stored once: no store-to-load
        // round trips through global memory between the swaps
        uint32_t p[kRegs];
#pragma unroll
        for (int i = 0; i < kRegs; ++i) p[i] = (uint32_t)(i + 1);
        for (int c = count; c >= 2; --c) {
            const uint32_t draw = (uint32_t)(count - c);
            const int sel = (int)bounded_uint(key_lo, key_hi, stream, mcs, 0, operation, kGlobalEntity, 0, draw,
                                              (uint32_t)c);
            uint32_t vc = 0, vs = 0;
#pragma unroll
            for (int i = 0; i < kRegs; ++i) {
                vc = i == c - 1 ? p[i] : vc;
                vs = i == sel ? p[i] : vs;
            }
#pragma unroll
            for (int i = 0; i < kRegs; ++i) p[i] = i == c - 1 ? vs : i == sel ? vc : p[i];
        }
#pragma unroll
        for (int i = 0; i < kRegs; ++i)
            if (i < count) order[i] = (T)p[i];
        return;
    }
    for (int i = 0; i < count; ++i) order[i] = (T)(i + 1);
    for (int c = count; c >= 2; --c) {
        const uint32_t draw = (uint32_t)(count - c);
        const uint32_t sel = bounded_uint(key_lo, key_hi, stream, mcs, 0, operation,
                                          kGlobalEntity, 0, draw, (uint32_t)c) + 1;
        const T t = order[c - 1];
        order[c - 1] = order[sel - 1];
        order[sel - 1] = t;
    }
}

// XCD-aware workgroup -> tile map. Workgroups are dispatched round-robin over the
// eight XCDs; this gives each XCD one contiguous band of tiles so that halo rows
// shared by neighbouring tiles stay in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t wg, uint32_t n) {
    constexpr uint32_t kXcd = 8;
    const uint32_t q = n / kXcd, rem = n % kXcd;
    const uint32_t xcd = wg % kXcd, k = wg / kXcd;
    return xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
}

}  // namespace cpm
