/*
 * CPU oracle for the CorePotts Metropolis hot path (TEST INFRASTRUCTURE ONLY).
 * See cpm_oracle.h for the list of reference files restated here and how each
 * function is pinned. Every routine follows the reference's data layout
 * (tags/ids SoA ownership, 1-based site and cell ids) and its evaluation order.
 */
#include "cpm_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Semantic RNG: rng/semantic.jl:136-310                                      */
/* ------------------------------------------------------------------------- */

#define PHILOX_M0 0xd2511f53u
#define PHILOX_M1 0xcd9e8d57u
#define PHILOX_W0 0x9e3779b9u
#define PHILOX_W1 0xbb67ae85u
#define GENERATION_DOMAIN 0xd2b74407b1ce6e93ull

/* rng/semantic.jl:144-187: ten rounds, key bumped between rounds. */
void orc_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = counter[0], c1 = counter[1], c2 = counter[2], c3 = counter[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* rng/semantic.jl:191-195: SplitMix64 finalizer. */
static uint64_t rng_mix64(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

/* rng/semantic.jl:249-261: injective packing of an address. */
void orc_counter_key(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                     uint32_t operation, uint32_t entity_kind, uint32_t entity,
                     uint64_t generation, uint32_t invocation, uint32_t draw,
                     uint64_t out[3]) {
    uint64_t c0 = mcs | ((uint64_t)(subround & 0xff) << 48) | ((uint64_t)(invocation & 0xff) << 56);
    uint64_t c1 = (uint64_t)entity | ((uint64_t)operation << 32) | ((uint64_t)stream << 44) |
                  ((uint64_t)entity_kind << 52) | ((uint64_t)draw << 54);
    out[0] = c0;
    out[1] = c1;
    out[2] = seed ^ rng_mix64(generation ^ GENERATION_DOMAIN);
}

void orc_rng_words(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                   uint32_t operation, uint32_t entity_kind, uint32_t entity,
                   uint64_t generation, uint32_t invocation, uint32_t draw, uint32_t out[4]) {
    uint64_t ck[3];
    orc_counter_key(seed, stream, mcs, subround, operation, entity_kind, entity, generation,
                    invocation, draw, ck);
    uint32_t counter[4] = {(uint32_t)ck[0], (uint32_t)(ck[0] >> 32), (uint32_t)ck[1],
                           (uint32_t)(ck[1] >> 32)};
    uint32_t key[2] = {(uint32_t)ck[2], (uint32_t)(ck[2] >> 32)};
    orc_philox4x32_10(counter, key, out);
}

/* rng/semantic.jl:276-280: (Float32(w >> 8) + 0.5f0) * 2^-24, all in Float32. */
float orc_uniform_open01_f32(uint32_t word1) {
    volatile float bits = (float)(word1 >> 8);
    volatile float shifted = bits + 0.5f;
    return shifted * 0x1.0p-24f;
}

/* rng/semantic.jl:298-310: rejection with invocation retries. */
uint32_t orc_bounded_uint(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                          uint32_t operation, uint32_t entity_kind, uint32_t entity,
                          uint32_t draw, uint32_t bound) {
    uint32_t threshold = (uint32_t)(0u - bound) % bound;
    for (uint32_t invocation = 0; invocation < 256; ++invocation) {
        uint32_t w[4];
        orc_rng_words(seed, stream, mcs, subround, operation, entity_kind, entity, 0, invocation,
                      draw, w);
        if (w[0] >= threshold) return w[0] % bound;
    }
    fprintf(stderr, "oracle: bounded_uint exhausted the invocation domain\n");
    abort();
}

/* sciml/interface.jl:1009-1026: EnsembleStream = 16, op 1, EnsembleEntity = 3. */
uint64_t orc_ensemble_seed(uint64_t master_seed, uint32_t trajectory, uint32_t rerun) {
    uint32_t w[4];
    orc_rng_words(master_seed, 16, 0, 0, 1, 3, trajectory, 0, rerun - 1, 0, w);
    return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}

enum {
    S_RECIPIENT = 3, S_DIRECTION = 4, S_ACCEPTANCE = 5, S_LOTTERY_ACTIVATION = 6,
    S_LOTTERY_PRIORITY = 7, S_CHECKERBOARD_ORDER = 8, S_CHECKERBOARD_PRIORITY = 17,
    S_AUX_EVOLUTION = 9, S_AUX_INITIALIZATION = 18, S_TILED_ORDER = 19, S_TILED_PROPOSAL = 20
};
enum { K_GLOBAL = 0, K_SITE = 1, K_CELL = 2 };

/* ------------------------------------------------------------------------- */
/* Domain and neighbour realization: spatial/cartesian.jl:415-525             */
/* ------------------------------------------------------------------------- */

enum { NB_MUTABLE = 1, NB_FIXED = 2, NB_EXTERIOR = 3, NB_ABSENT = 4, NB_INVALID = 5 };
enum { ROLE_PROPOSAL = 0, ROLE_CONTACT = 1, ROLE_SURFACE = 2 };

typedef struct {
    int32_t kind;
    uint32_t site;  /* 1-based */
    int64_t owner;  /* owner word for FIXED / EXTERIOR: +cell, -medium */
} nbr_t;

static inline int64_t owner_word(uint8_t tag, uint32_t id) {
    return tag == 1 ? (int64_t)id : -(int64_t)id;
}

static inline int64_t nsites(const orc_model *m) {
    int64_t n = 1;
    for (int d = 0; d < m->ndim; ++d) n *= m->dims[d];
    return n;
}

static inline int site_mutable(const orc_model *m, uint32_t site) {
    return m->mutable_mask == NULL || m->mutable_mask[site - 1] != 0;
}

/* topology.jl:195-209 then cartesian.jl:494-525. */
static nbr_t realize(const orc_model *m, const orc_relation *rel, int role, uint32_t site,
                     int dir) {
    uint32_t idx0 = site - 1;
    int32_t c[3] = {0, 0, 0};
    uint32_t W = (uint32_t)m->dims[0];
    if (m->ndim == 2) {
        c[0] = (int32_t)(idx0 % W);
        c[1] = (int32_t)(idx0 / W);
    } else {
        uint32_t H = (uint32_t)m->dims[1];
        uint32_t stride = W * H;
        uint32_t rem = idx0 % stride;
        c[2] = (int32_t)(idx0 / stride);
        c[0] = (int32_t)(rem % W);
        c[1] = (int32_t)(rem / W);
    }
    int closed = 0, fixed = 0, invalid = 0;
    int64_t fixed_owner = 0;
    for (int d = 0; d < m->ndim; ++d) {
        int32_t v = c[d] + rel->off[dir][d];
        int face;
        if (v < 0) face = 0;
        else if (v >= m->dims[d]) face = 1;
        else { c[d] = v; continue; }
        int kind = m->face_kind[d][face];
        if (kind == ORC_FACE_PERIODIC) {
            while (v < 0) v += m->dims[d];
            while (v >= m->dims[d]) v -= m->dims[d];
            c[d] = v;
        } else if (kind == ORC_FACE_CLOSED) {
            invalid |= fixed;
            closed = 1;
            c[d] = v;
        } else {
            int64_t fo = -(int64_t)m->face_medium[d][face];
            int incompatible = fixed && (fixed_owner != fo);
            invalid |= closed | incompatible;
            fixed = 1;
            fixed_owner = fo;
            c[d] = v;
        }
    }
    nbr_t r = {NB_ABSENT, 0, 0};
    if (invalid) { r.kind = NB_INVALID; return r; }
    if (closed) return r;
    int allows_fixed = role != ROLE_PROPOSAL;
    if (fixed) {
        if (!allows_fixed) return r;
        r.kind = NB_EXTERIOR;
        r.owner = fixed_owner;
        return r;
    }
    int64_t lin = c[0];
    if (m->ndim >= 2) lin += (int64_t)c[1] * m->dims[0];
    if (m->ndim == 3) lin += (int64_t)c[2] * m->dims[0] * m->dims[1];
    uint32_t nsite = (uint32_t)(lin + 1);
    if (site_mutable(m, nsite)) {
        r.kind = NB_MUTABLE;
        r.site = nsite;
    } else if (allows_fixed) {
        r.kind = NB_FIXED;
        r.site = nsite;
        r.owner = owner_word(m->imm_tags[nsite - 1], m->imm_ids[nsite - 1]);
    }
    return r;
}

/* ------------------------------------------------------------------------- */
/* Simulation handle                                                          */
/* ------------------------------------------------------------------------- */

struct orc_sim {
    orc_model m;
    int32_t algorithm;
    uint64_t seed;
    uint64_t mcs;
    int64_t n;           /* lattice sites */
    int64_t n_mutable;
    uint32_t *mutable_sites; /* ascending 1-based */
    orc_state st;
    /* schedules */
    int32_t color_count;     /* colours (checkerboard) or rounds (lottery) */
    uint32_t *color_sites;   /* sites grouped by colour, ascending inside a colour */
    uint32_t *color_offsets; /* 0-based offsets, color_count + 1 */
    uint32_t *csr_offsets;   /* lottery: per mutable position, 0-based */
    uint32_t *csr_nbrs;      /* lottery: neighbour positions (0-based in mutable_sites) */
    uint32_t *order;         /* last colour / round order (1-based values) */
    /* report of the last MCS */
    uint64_t report[12];
    int has_report;
    /* per-attempt scratch */
    int64_t *att_losing, *att_gaining;
    uint32_t *att_recipient, *att_priority;
    uint8_t *att_outcome, *att_selected, *att_disp;
    int64_t *att_dl, *att_dg;
    float *att_dh;           /* delta_h of evaluated attempts */
    uint32_t *cell_max_priority, *cell_min_identity;
    uint32_t (*tickets)[4];
    uint32_t *activations;   /* lottery: activated opportunities of each mutable position in the
                                last MCS (lottery.jl accounting[1:N]) */
    /* TiledCheckerboardCPM layout (tiled_checkerboard.jl:87-196) */
    int32_t tile_size[3], tile_grid[3], tile_reach[3], tile_halo;
    int64_t ntiles;
    uint16_t *tile_colors;   /* per tile, 1-based */
    uint32_t *tile_sites;    /* mutable sites of each tile, x fastest inside a tile */
    uint32_t *tile_offsets;  /* ntiles + 1 */
    uint32_t *color_tiles;   /* tiles grouped by colour, ascending */
    uint32_t *color_tile_offsets; /* color_count + 1 */
    int32_t tiled_interval, tiled_batches, tiled_recount;
    int64_t *acc_dvol, *acc_dbnd;
    /* FP64 sums of the accepted copies' Float32 delta_h: over the last MCS and since the
       state was loaded (drift checks against the global energy) */
    double dh_last_mcs, dh_total;
};

enum { OUT_ACTIONABLE = 1, OUT_SAME_OWNER = 2, OUT_BOUNDARY_NULL = 3, OUT_IMMUTABLE = 4 };
enum { D_PENDING = 0, D_CONFLICT = 1, D_CONSTRAINT = 2, D_REJECTED = 3, D_ACCEPTED = 4 };

static inline int64_t owner_at(const orc_sim *s, uint32_t site) {
    return owner_word(s->st.tags[site - 1], s->st.ids[site - 1]);
}

static uint32_t medium_type(const orc_model *m, int64_t owner) {
    uint32_t id = (uint32_t)(-owner);
    for (int i = 0; i < m->n_table; ++i)
        if (m->table_owner[i] == id) return m->table_type[i];
    fprintf(stderr, "oracle: medium owner %u has no declared type\n", id);
    abort();
}

static uint32_t owner_type(const orc_sim *s, int64_t owner) {
    if (owner > 0) return s->st.cell_types[owner - 1];
    return medium_type(&s->m, owner);
}

static int medium_index(const orc_model *m, int64_t owner) {
    uint32_t id = (uint32_t)(-owner);
    for (int i = 0; i < m->n_media; ++i)
        if (m->medium_ids[i] == id) return i;
    fprintf(stderr, "oracle: undeclared medium owner %u\n", id);
    abort();
}

/* cartesian.jl / checkerboard.jl:31-54: union of realized mutable neighbours. */
static int collect_conflict_neighbours(const orc_model *m, uint32_t site, uint32_t *out) {
    const orc_relation *rels[3] = {&m->proposal, &m->surface, &m->contact};
    int roles[3] = {ROLE_PROPOSAL, ROLE_SURFACE, ROLE_CONTACT};
    int nrel = 2;
    for (int t = 0; t < m->n_energies; ++t)
        if (m->energy_order[t] == ORC_TERM_CONTACT) nrel = 3;
    int count = 0;
    for (int r = 0; r < nrel; ++r) {
        for (int d = 0; d < rels[r]->n; ++d) {
            nbr_t nb = realize(m, rels[r], roles[r], site, d);
            if (nb.kind != NB_MUTABLE || nb.site == site) continue;
            int dup = 0;
            for (int k = 0; k < count; ++k) dup |= out[k] == nb.site;
            if (!dup) out[count++] = nb.site;
        }
    }
    /* sort ascending (the CSR order of the reference; order does not change results) */
    for (int i = 1; i < count; ++i) {
        uint32_t v = out[i];
        int j = i - 1;
        while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; --j; }
        out[j + 1] = v;
    }
    return count;
}

int32_t orc_conflict_degrees(const orc_model *m, uint8_t *degrees) {
    int64_t n = nsites(m);
    uint32_t nb[3 * ORC_MAX_DIRS];
    int32_t maxdeg = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t site = (uint32_t)(i + 1);
        if (!site_mutable(m, site)) { degrees[i] = 0; continue; }
        int c = collect_conflict_neighbours(m, site, nb);
        degrees[i] = (uint8_t)c;
        if (c > maxdeg) maxdeg = c;
    }
    return maxdeg;
}

/* checkerboard.jl:57-90: greedy smallest-free colour in ascending site order. */
int32_t orc_realized_coloring(const orc_model *m, uint8_t *colors, int32_t *max_degree) {
    int64_t n = nsites(m);
    uint32_t nb[3 * ORC_MAX_DIRS];
    int forbidden[3 * ORC_MAX_DIRS + 2];
    int32_t maxcolor = 0, maxdeg = 0;
    memset(colors, 0, (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t site = (uint32_t)(i + 1);
        if (!site_mutable(m, site)) continue;
        int c = collect_conflict_neighbours(m, site, nb);
        if (c > maxdeg) maxdeg = c;
        for (int k = 0; k <= c + 1; ++k) forbidden[k] = 0;
        for (int k = 0; k < c; ++k) {
            int col = colors[nb[k] - 1];
            if (col > 0 && col <= c + 1) forbidden[col] = 1;
        }
        int color = 1;
        while (forbidden[color]) ++color;
        if (color > 255) return -1;
        colors[i] = (uint8_t)color;
        if (color > maxcolor) maxcolor = color;
    }
    if (max_degree) *max_degree = maxdeg;
    return maxcolor;
}

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

static int32_t relation_radius(const orc_relation *r) {
    int32_t h = 0;
    for (int k = 0; k < r->n; ++k)
        for (int a = 0; a < 3; ++a) {
            int32_t v = r->off[k][a] < 0 ? -r->off[k][a] : r->off[k][a];
            if (v > h) h = v;
        }
    return h;
}

/* tiled_checkerboard.jl:87-196 and tiled_checkerboard_device.jl:28-121: halo radius of the
 * declared accesses, tile grid, greedy halo colouring of the tiles in linear order (two tiles
 * conflict when their separation, periodic where the axis is, is within the reach on every
 * axis), the mutable sites of every tile, the switching interval and batch count. Only the
 * tiles within the reach of a tile can conflict with it, so the greedy scan visits those. */
static int tiled_layout(orc_sim *s) {
    const orc_model *m = &s->m;
    int32_t halo = relation_radius(&m->proposal);
    int32_t rs = relation_radius(&m->surface);
    if (rs > halo) halo = rs;
    for (int t = 0; t < m->n_energies; ++t) {
        int32_t r = 0;
        if (m->energy_order[t] == ORC_TERM_CONTACT) r = relation_radius(&m->contact);
        else if (m->energy_order[t] == ORC_TERM_BOUNDARY) r = rs;
        else continue;
        if (r < 1) r = 1; /* dependency_radius = 1 (scientific_inner_loop.jl:321-332) */
        if (r > halo) halo = r;
    }
    s->tile_halo = halo;
    int64_t ntiles = 1;
    int periodic[3];
    for (int a = 0; a < 3; ++a) {
        int32_t ts = a < m->ndim ? (m->tile_size[a] > 0 ? m->tile_size[a] : (m->ndim == 2 ? 4 : 2)) : 1;
        s->tile_size[a] = ts;
        s->tile_grid[a] = a < m->ndim ? (m->dims[a] + ts - 1) / ts : 1;
        s->tile_reach[a] = a < m->ndim ? (halo + ts - 1) / ts : 0;
        periodic[a] = a < m->ndim && m->face_kind[a][0] == ORC_FACE_PERIODIC;
        ntiles *= s->tile_grid[a];
    }
    s->ntiles = ntiles;
    const int32_t gx = s->tile_grid[0], gy = s->tile_grid[1];
    s->tile_colors = (uint16_t *)xcalloc((size_t)ntiles, 2);
    int32_t maxc = 0;
    uint8_t *taken = (uint8_t *)xcalloc(65538, 1);
    for (int64_t t = 0; t < ntiles; ++t) {
        int32_t c[3] = {(int32_t)(t % gx), (int32_t)((t / gx) % gy), (int32_t)(t / ((int64_t)gx * gy))};
        int32_t lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = -s->tile_reach[a]; hi[a] = s->tile_reach[a]; }
        int32_t marked[4096];
        int nm = 0;
        for (int32_t dz = lo[2]; dz <= hi[2]; ++dz)
            for (int32_t dy = lo[1]; dy <= hi[1]; ++dy)
                for (int32_t dx = lo[0]; dx <= hi[0]; ++dx) {
                    int32_t d[3] = {dx, dy, dz}, q[3];
                    int ok = 1;
                    for (int a = 0; a < 3 && ok; ++a) {
                        int32_t v = c[a] + d[a], n = s->tile_grid[a];
                        if (periodic[a]) { v %= n; if (v < 0) v += n; }
                        else if (v < 0 || v >= n) ok = 0;
                        q[a] = v;
                    }
                    if (!ok) continue;
                    int64_t u = q[0] + (int64_t)gx * (q[1] + (int64_t)gy * q[2]);
                    if (u >= t) continue;
                    int32_t col = s->tile_colors[u];
                    if (!taken[col] && nm < 4096) { taken[col] = 1; marked[nm++] = col; }
                }
        int32_t col = 1;
        while (taken[col]) ++col;
        for (int k = 0; k < nm; ++k) taken[marked[k]] = 0;
        if (col > 65535) { free(taken); return 1; }
        s->tile_colors[t] = (uint16_t)col;
        if (col > maxc) maxc = col;
    }
    free(taken);
    if (maxc > 255) return 1;
    s->color_count = maxc;
    s->color_tile_offsets = (uint32_t *)xcalloc((size_t)maxc + 1, 4);
    s->color_tiles = (uint32_t *)xcalloc((size_t)ntiles, 4);
    uint32_t pos = 0;
    for (int32_t col = 1; col <= maxc; ++col) {
        s->color_tile_offsets[col - 1] = pos;
        for (int64_t t = 0; t < ntiles; ++t)
            if (s->tile_colors[t] == col) s->color_tiles[pos++] = (uint32_t)(t + 1);
    }
    s->color_tile_offsets[maxc] = pos;
    s->tile_offsets = (uint32_t *)xcalloc((size_t)ntiles + 1, 4);
    s->tile_sites = (uint32_t *)xcalloc((size_t)s->n_mutable + 1, 4);
    uint32_t at = 0, maxsites = 0;
    const int32_t X = m->dims[0], Y = m->dims[1], Z = m->ndim == 3 ? m->dims[2] : 1;
    for (int64_t t = 0; t < ntiles; ++t) {
        s->tile_offsets[t] = at;
        int32_t c[3] = {(int32_t)(t % gx), (int32_t)((t / gx) % gy), (int32_t)(t / ((int64_t)gx * gy))};
        int32_t z1 = c[2] * s->tile_size[2] + s->tile_size[2]; if (z1 > Z) z1 = Z;
        int32_t y1 = c[1] * s->tile_size[1] + s->tile_size[1]; if (y1 > Y) y1 = Y;
        int32_t x1 = c[0] * s->tile_size[0] + s->tile_size[0]; if (x1 > X) x1 = X;
        for (int32_t z = c[2] * s->tile_size[2]; z < z1; ++z)
            for (int32_t y = c[1] * s->tile_size[1]; y < y1; ++y)
                for (int32_t x = c[0] * s->tile_size[0]; x < x1; ++x) {
                    uint32_t site = 1u + (uint32_t)x + (uint32_t)X * ((uint32_t)y + (uint32_t)Y * (uint32_t)z);
                    if (site_mutable(m, site)) s->tile_sites[at++] = site;
                }
        if (at - s->tile_offsets[t] > maxsites) maxsites = at - s->tile_offsets[t];
    }
    s->tile_offsets[ntiles] = at;
    if (maxsites == 0) return 1;
    s->tiled_interval = m->switching_interval > 0 && (uint32_t)m->switching_interval < maxsites
                            ? m->switching_interval : (int32_t)maxsites;
    s->tiled_batches = ((int32_t)maxsites + s->tiled_interval - 1) / s->tiled_interval;
    if (s->color_count * s->tiled_batches > 255) return 1;
    s->acc_dvol = (int64_t *)xcalloc((size_t)m->capacity, 8);
    s->acc_dbnd = (int64_t *)xcalloc((size_t)m->capacity, 8);
    return 0;
}

void orc_set_tiled_reconcile(orc_sim *s, int32_t mode) { s->tiled_recount = mode; }

int32_t orc_tiled_layout(const orc_sim *s, int32_t grid[3], int32_t reach[3], int32_t *halo,
                         uint16_t *tile_colors) {
    for (int a = 0; a < 3; ++a) {
        grid[a] = s->tile_grid[a];
        reach[a] = s->tile_reach[a];
    }
    *halo = s->tile_halo;
    if (tile_colors) memcpy(tile_colors, s->tile_colors, (size_t)s->ntiles * 2);
    return s->color_count;
}

orc_sim *orc_create(const orc_model *model, int32_t algorithm, uint64_t seed) {
    orc_sim *s = (orc_sim *)xcalloc(1, sizeof(orc_sim));
    s->m = *model;
    s->algorithm = algorithm;
    s->seed = seed;
    s->n = nsites(model);
    int64_t cap = model->capacity;
    /* own copies of the domain arrays */
    if (model->mutable_mask) {
        uint8_t *mask = (uint8_t *)xcalloc((size_t)s->n, 1);
        uint8_t *tags = (uint8_t *)xcalloc((size_t)s->n, 1);
        uint32_t *ids = (uint32_t *)xcalloc((size_t)s->n, 4);
        memcpy(mask, model->mutable_mask, (size_t)s->n);
        memcpy(tags, model->imm_tags, (size_t)s->n);
        memcpy(ids, model->imm_ids, (size_t)s->n * 4);
        s->m.mutable_mask = mask;
        s->m.imm_tags = tags;
        s->m.imm_ids = ids;
    }
    s->mutable_sites = (uint32_t *)xcalloc((size_t)s->n, 4);
    for (int64_t i = 0; i < s->n; ++i)
        if (site_mutable(&s->m, (uint32_t)(i + 1))) s->mutable_sites[s->n_mutable++] = (uint32_t)(i + 1);
    s->st.tags = (uint8_t *)xcalloc((size_t)s->n, 1);
    s->st.ids = (uint32_t *)xcalloc((size_t)s->n, 4);
    s->st.finite_volumes = (int32_t *)xcalloc((size_t)cap, 4);
    s->st.medium_volumes = (int32_t *)xcalloc(ORC_MAX_MEDIA, 4);
    s->st.boundary_measures = (int64_t *)xcalloc((size_t)cap, 8);
    s->st.active = (uint8_t *)xcalloc((size_t)cap, 1);
    s->st.cell_types = (uint32_t *)xcalloc((size_t)cap, 4);
    s->st.volume_pressure = (float *)xcalloc((size_t)cap, 4);
    s->st.surface_tension = (float *)xcalloc((size_t)cap, 4);
    s->st.target_volume = (double *)xcalloc((size_t)cap, 8);
    s->st.volume_strength = (double *)xcalloc((size_t)cap, 8);
    s->st.target_boundary = (int64_t *)xcalloc((size_t)cap, 8);
    s->st.boundary_strength = (double *)xcalloc((size_t)cap, 8);
    int64_t nm = s->n_mutable;
    s->att_losing = (int64_t *)xcalloc((size_t)nm, 8);
    s->att_gaining = (int64_t *)xcalloc((size_t)nm, 8);
    s->att_recipient = (uint32_t *)xcalloc((size_t)nm, 4);
    s->att_priority = (uint32_t *)xcalloc((size_t)nm, 4);
    s->att_outcome = (uint8_t *)xcalloc((size_t)nm, 1);
    s->att_selected = (uint8_t *)xcalloc((size_t)nm, 1);
    s->att_disp = (uint8_t *)xcalloc((size_t)nm, 1);
    s->att_dl = (int64_t *)xcalloc((size_t)nm, 8);
    s->att_dg = (int64_t *)xcalloc((size_t)nm, 8);
    s->att_dh = (float *)xcalloc((size_t)nm, 4);
    s->cell_max_priority = (uint32_t *)xcalloc((size_t)cap + 1, 4);
    s->cell_min_identity = (uint32_t *)xcalloc((size_t)cap + 1, 4);
    s->order = (uint32_t *)xcalloc(256, 4);

    if (algorithm == ORC_ALG_CHECKERBOARD) {
        uint8_t *colors = (uint8_t *)xcalloc((size_t)s->n, 1);
        int32_t count = orc_realized_coloring(&s->m, colors, NULL);
        if (count < 0) { free(colors); orc_destroy(s); return NULL; }
        s->color_count = count;
        s->color_offsets = (uint32_t *)xcalloc((size_t)count + 1, 4);
        s->color_sites = (uint32_t *)xcalloc((size_t)nm, 4);
        uint32_t pos = 0;
        for (int c = 1; c <= count; ++c) {
            s->color_offsets[c - 1] = pos;
            for (int64_t i = 0; i < s->n; ++i)
                if (colors[i] == c) s->color_sites[pos++] = (uint32_t)(i + 1);
        }
        s->color_offsets[count] = pos;
        free(colors);
    } else if (algorithm == ORC_ALG_LOTTERY) {
        /* lottery.jl:46-92 + checkerboard.jl:31-54: CSR over mutable positions */
        uint32_t nb[3 * ORC_MAX_DIRS];
        uint32_t *pos_of = (uint32_t *)xcalloc((size_t)s->n, 4);
        for (int64_t i = 0; i < nm; ++i) pos_of[s->mutable_sites[i] - 1] = (uint32_t)i;
        s->csr_offsets = (uint32_t *)xcalloc((size_t)nm + 1, 4);
        size_t capn = (size_t)nm * 8 + 16, used = 0;
        s->csr_nbrs = (uint32_t *)xcalloc(capn, 4);
        int32_t maxdeg = 0;
        for (int64_t i = 0; i < nm; ++i) {
            int c = collect_conflict_neighbours(&s->m, s->mutable_sites[i], nb);
            if (used + (size_t)c > capn) {
                capn = capn * 2 + (size_t)c;
                s->csr_nbrs = (uint32_t *)realloc(s->csr_nbrs, capn * 4);
            }
            s->csr_offsets[i] = (uint32_t)used;
            uint32_t tmp[3 * ORC_MAX_DIRS];
            for (int k = 0; k < c; ++k) tmp[k] = pos_of[nb[k] - 1];
            /* sort positions ascending (sort!(unique!(...)) in the reference) */
            for (int a = 1; a < c; ++a) {
                uint32_t v = tmp[a];
                int b = a - 1;
                while (b >= 0 && tmp[b] > v) { tmp[b + 1] = tmp[b]; --b; }
                tmp[b + 1] = v;
            }
            for (int k = 0; k < c; ++k) s->csr_nbrs[used++] = tmp[k];
            if (c > maxdeg) maxdeg = c;
        }
        s->csr_offsets[nm] = (uint32_t)used;
        free(pos_of);
        s->color_count = maxdeg + 1;
        if (s->color_count > 255) { orc_destroy(s); return NULL; }
        s->tickets = (uint32_t (*)[4])xcalloc((size_t)nm, 16);
        s->activations = (uint32_t *)xcalloc((size_t)nm, sizeof(uint32_t));
    } else if (algorithm == ORC_ALG_TILED) {
        if (tiled_layout(s) != 0) { orc_destroy(s); return NULL; }
    } else {
        s->color_count = 1;
    }
    return s;
}

void orc_destroy(orc_sim *s) {
    if (!s) return;
    if (s->m.mutable_mask) {
        free((void *)s->m.mutable_mask);
        free((void *)s->m.imm_tags);
        free((void *)s->m.imm_ids);
    }
    free(s->mutable_sites);
    free(s->st.tags); free(s->st.ids); free(s->st.finite_volumes); free(s->st.medium_volumes);
    free(s->st.boundary_measures); free(s->st.active); free(s->st.cell_types);
    free(s->st.target_volume); free(s->st.volume_strength); free(s->st.target_boundary);
    free(s->st.boundary_strength);
    free(s->st.volume_pressure); free(s->st.surface_tension);
    free(s->att_losing); free(s->att_gaining); free(s->att_recipient); free(s->att_priority);
    free(s->att_outcome); free(s->att_selected); free(s->att_disp); free(s->att_dl); free(s->att_dh);
    free(s->att_dg); free(s->cell_max_priority); free(s->cell_min_identity); free(s->order);
    free(s->color_sites); free(s->color_offsets); free(s->csr_offsets); free(s->csr_nbrs);
    free(s->tickets);
    free(s->activations);
    free(s->tile_colors); free(s->tile_sites); free(s->tile_offsets); free(s->color_tiles);
    free(s->color_tile_offsets); free(s->acc_dvol); free(s->acc_dbnd);
    free(s);
}

orc_state *orc_state_of(orc_sim *s) { return &s->st; }
void orc_set_seed(orc_sim *s, uint64_t seed) { s->seed = seed; }
void orc_set_mcs(orc_sim *s, uint64_t mcs) { s->mcs = mcs; s->has_report = 0; }
uint64_t orc_mcs(const orc_sim *s) { return s->mcs; }
int32_t orc_round_count(const orc_sim *s) { return s->color_count; }

int32_t orc_last_order(const orc_sim *s, uint32_t *out) {
    for (int i = 0; i < s->color_count; ++i) out[i] = s->order[i];
    return s->color_count;
}

int64_t orc_last_activations(const orc_sim *s, uint32_t *out) {
    if (!s->activations) return -1;
    memcpy(out, s->activations, (size_t)s->n_mutable * sizeof(uint32_t));
    return s->n_mutable;
}

static void copy_state(const orc_model *m, int64_t n, const orc_state *src, orc_state *dst) {
    size_t cap = (size_t)m->capacity;
    memcpy(dst->tags, src->tags, (size_t)n);
    memcpy(dst->ids, src->ids, (size_t)n * 4);
    memcpy(dst->finite_volumes, src->finite_volumes, cap * 4);
    memcpy(dst->medium_volumes, src->medium_volumes, (size_t)m->n_media * 4);
    memcpy(dst->boundary_measures, src->boundary_measures, cap * 8);
    memcpy(dst->active, src->active, cap);
    memcpy(dst->cell_types, src->cell_types, cap * 4);
    memcpy(dst->target_volume, src->target_volume, cap * 8);
    memcpy(dst->volume_strength, src->volume_strength, cap * 8);
    memcpy(dst->target_boundary, src->target_boundary, cap * 8);
    memcpy(dst->boundary_strength, src->boundary_strength, cap * 8);
    if (src->volume_pressure && dst->volume_pressure)
        memcpy(dst->volume_pressure, src->volume_pressure, cap * 4);
    if (src->surface_tension && dst->surface_tension)
        memcpy(dst->surface_tension, src->surface_tension, cap * 4);
}

int32_t orc_load_state(orc_sim *s, const orc_state *src) {
    s->dh_last_mcs = s->dh_total = 0.0;
    copy_state(&s->m, s->n, src, &s->st);
    return 0;
}

int32_t orc_store_state(const orc_sim *s, orc_state *dst) {
    copy_state(&s->m, s->n, &s->st, dst);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Attempt construction: protocols/scientific.jl:241-267                       */
/* ------------------------------------------------------------------------- */

static int construct_attempt(const orc_sim *s, uint32_t recipient, int direction,
                             int64_t *losing, int64_t *gaining, uint32_t *donor) {
    *losing = owner_at(s, recipient);
    *gaining = *losing;
    *donor = recipient;
    if (!site_mutable(&s->m, recipient)) return OUT_IMMUTABLE;
    nbr_t nb = realize(&s->m, &s->m.proposal, ROLE_PROPOSAL, recipient, direction - 1);
    if (nb.site == 0) return OUT_BOUNDARY_NULL;
    *donor = nb.site;
    *gaining = owner_at(s, nb.site);
    if (*gaining == *losing) return OUT_SAME_OWNER;
    return OUT_ACTIONABLE;
}

/* ------------------------------------------------------------------------- */
/* Energy terms                                                               */
/* ------------------------------------------------------------------------- */

/* scientific_components.jl:466-481 with BoundaryEdgeCount weights (Int64 1). */
static void boundary_change(const orc_sim *s, uint32_t recipient, int64_t losing,
                            int64_t gaining, int64_t *dl, int64_t *dg) {
    int64_t l = 0, g = 0;
    for (int d = 0; d < s->m.surface.n; ++d) {
        nbr_t nb = realize(&s->m, &s->m.surface, ROLE_SURFACE, recipient, d);
        if (nb.kind == NB_ABSENT || nb.kind == NB_INVALID) continue;
        int64_t o = nb.kind == NB_MUTABLE ? owner_at(s, nb.site) : nb.owner;
        l += (o == losing) ? 1 : -1;
        g += (o == gaining) ? -1 : 1;
    }
    *dl = l;
    *dg = g;
}

/* scientific_components.jl:90-95 */
static float vol_change_f32(float strength, float target, int32_t volume, int32_t change) {
    float d0 = (float)volume - target;
    float old = strength * (d0 * d0);
    int32_t updated = volume + change;
    if (updated == 0) return -old;
    float d1 = (float)updated - target;
    return strength * (d1 * d1) - old;
}
static double vol_change_f64(double strength, double target, int32_t volume, int32_t change) {
    double d0 = (double)volume - target;
    double old = strength * (d0 * d0);
    int32_t updated = volume + change;
    if (updated == 0) return -old;
    double d1 = (double)updated - target;
    return strength * (d1 * d1) - old;
}

/* scientific_trackers.jl:123-142 */
static double volume_term(const orc_sim *s, int64_t losing, int64_t gaining) {
    const orc_state *st = &s->st;
    if (s->m.volume_f64) {
        double delta = 0.0;
        if (losing > 0) {
            int64_t i = losing - 1;
            delta += vol_change_f64(st->volume_strength[i], st->target_volume[i],
                                    st->finite_volumes[i], -1);
        }
        if (gaining > 0) {
            int64_t i = gaining - 1;
            delta += vol_change_f64(st->volume_strength[i], st->target_volume[i],
                                    st->finite_volumes[i], 1);
        }
        return delta;
    }
    float delta = 0.0f;
    if (losing > 0) {
        int64_t i = losing - 1;
        delta += vol_change_f32((float)st->volume_strength[i], (float)st->target_volume[i],
                                st->finite_volumes[i], -1);
    }
    if (gaining > 0) {
        int64_t i = gaining - 1;
        delta += vol_change_f32((float)st->volume_strength[i], (float)st->target_volume[i],
                                st->finite_volumes[i], 1);
    }
    return (double)delta;
}

/* scientific_components.jl:289-308 */
static double contact_term(const orc_sim *s, uint32_t recipient, int64_t losing, int64_t gaining) {
    const orc_model *m = &s->m;
    uint32_t lt = owner_type(s, losing), gt = owner_type(s, gaining);
    int nt = m->n_types;
    if (lt < 1 || lt > (uint32_t)nt || gt < 1 || gt > (uint32_t)nt) {
        fprintf(stderr, "oracle: owner type outside the contact matrix\n");
        abort();
    }
    double dd = 0.0;
    float df = 0.0f;
    for (int d = 0; d < m->contact.n; ++d) {
        nbr_t nb = realize(m, &m->contact, ROLE_CONTACT, recipient, d);
        if (nb.kind == NB_ABSENT || nb.kind == NB_INVALID) continue;
        int64_t o = nb.kind == NB_MUTABLE ? owner_at(s, nb.site) : nb.owner;
        uint32_t ntp = owner_type(s, o);
        if (ntp < 1 || ntp > (uint32_t)nt) {
            fprintf(stderr, "oracle: owner type outside the contact matrix\n");
            abort();
        }
        double jl = m->J[(lt - 1) * nt + (ntp - 1)];
        double jg = m->J[(gt - 1) * nt + (ntp - 1)];
        if (m->contact_f64) {
            double w = m->contact.weight[d];
            if (losing != o) dd -= w * jl;
            if (gaining != o) dd += w * jg;
        } else {
            float w = (float)m->contact.weight[d];
            if (losing != o) df -= w * (float)jl;
            if (gaining != o) df += w * (float)jg;
        }
    }
    return m->contact_f64 ? dd : (double)df;
}

/* scientific_components.jl:540-545 */
static double bnd_owner_f64(double strength, double target, double measure, int32_t volume,
                            double bchange, int32_t vchange) {
    double d0 = measure - target;
    double old = strength * (d0 * d0);
    if (vchange == -1 && volume == 1) return -old;
    double d1 = (measure + bchange) - target;
    return strength * (d1 * d1) - old;
}
static float bnd_owner_f32(float strength, float target, float measure, int32_t volume,
                           float bchange, int32_t vchange) {
    float d0 = measure - target;
    float old = strength * (d0 * d0);
    if (vchange == -1 && volume == 1) return -old;
    float d1 = (measure + bchange) - target;
    return strength * (d1 * d1) - old;
}

/* scientific_trackers.jl:144-172 */
static double boundary_term(const orc_sim *s, int64_t losing, int64_t gaining, int64_t dl,
                            int64_t dg) {
    const orc_state *st = &s->st;
    if (s->m.boundary_f64) {
        double r = 0.0;
        if (losing > 0) {
            int64_t i = losing - 1;
            r += bnd_owner_f64(st->boundary_strength[i], (double)st->target_boundary[i],
                               (double)st->boundary_measures[i], st->finite_volumes[i],
                               (double)dl, -1);
        }
        if (gaining > 0) {
            int64_t i = gaining - 1;
            r += bnd_owner_f64(st->boundary_strength[i], (double)st->target_boundary[i],
                               (double)st->boundary_measures[i], st->finite_volumes[i],
                               (double)dg, 1);
        }
        return r;
    }
    float r = 0.0f;
    if (losing > 0) {
        int64_t i = losing - 1;
        r += bnd_owner_f32((float)st->boundary_strength[i], (float)st->target_boundary[i],
                           (float)st->boundary_measures[i], st->finite_volumes[i], (float)dl,
                           -1);
    }
    if (gaining > 0) {
        int64_t i = gaining - 1;
        r += bnd_owner_f32((float)st->boundary_strength[i], (float)st->target_boundary[i],
                           (float)st->boundary_measures[i], st->finite_volumes[i], (float)dg,
                           1);
    }
    return (double)r;
}

/* scientific_inner_loop.jl:213-217: ((0 + E1) + E2) + ... with promotion, before evaluate_copy
 * rounds the fold to the algorithm's number type (:257-259). */
static double energy_fold(const orc_sim *s, uint32_t recipient, int64_t losing, int64_t gaining,
                          int64_t dl, int64_t dg) {
    double acc = 0.0;
    int acc64 = 0;
    for (int t = 0; t < s->m.n_energies; ++t) {
        int term = s->m.energy_order[t];
        double v;
        int v64;
        if (term == ORC_TERM_VOLUME) {
            v = volume_term(s, losing, gaining);
            v64 = s->m.volume_f64;
        } else if (term == ORC_TERM_CONTACT) {
            v = contact_term(s, recipient, losing, gaining);
            v64 = s->m.contact_f64;
        } else {
            v = boundary_term(s, losing, gaining, dl, dg);
            v64 = s->m.boundary_f64;
        }
        if (!acc64 && !v64) {
            acc = (double)((float)acc + (float)v);
        } else {
            acc = acc + v;
            acc64 = 1;
        }
    }
    return acc;
}

/* evaluate_copy's delta_h: the fold rounded once to Float32 (scientific_inner_loop.jl:257-259). */
static float delta_h(const orc_sim *s, uint32_t recipient, int64_t losing, int64_t gaining,
                     int64_t dl, int64_t dg) {
    return (float)energy_fold(s, recipient, losing, gaining, dl, dg);
}

/* ------------------------------------------------------------------------- */
/* Mechanics: scientific_mechanics.jl, scientific_components.jl:220-231,646-657 */
/* ------------------------------------------------------------------------- */

/* rng/semantic.jl:323-331: lanes 1 and 2 of one Philox call; log and cos in double
 * precision, rounded once to Float32. */
float orc_normal_box_muller(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t subround,
                            uint32_t operation, uint32_t entity, uint32_t invocation) {
    uint32_t w[4];
    orc_rng_words(seed, stream, mcs, subround, operation, K_CELL, entity, 0, invocation, 0, w);
    float u1 = orc_uniform_open01_f32(w[0]);
    float u2 = orc_uniform_open01_f32(w[1]);
    float radius = sqrtf(-2.0f * (float)log((double)u1));
    float angle = (2.0f * 3.14159274101257324f) * u2; /* 2 * Float32(pi) */
    return radius * (float)cos((double)angle);
}

/* scientific_mechanics.jl:29-35, every operation in Float32 */
float orc_ou_transition(float q, float displacement, float strength, float eta, float noise,
                        float interval, float standard_normal) {
    float alpha = (float)exp((double)(-eta * interval));
    float two_s = 2.0f * strength;
    float mean = (1.0f - alpha) * (two_s * displacement);
    float rest = 1.0f - alpha * alpha;
    float variance = (two_s * noise) * (rest > 0.0f ? rest : 0.0f);
    return (alpha * q + mean) + sqrtf(variance) * standard_normal;
}

static float mech_noise(const orc_sim *s, int k) {
    return s->m.mech_noise_is_temperature[k] ? s->m.temperature : s->m.mech_noise[k];
}

static float *mech_column(orc_sim *s, int k) {
    return s->m.mech_kind[k] == ORC_MECH_PRESSURE ? s->st.volume_pressure : s->st.surface_tension;
}

/* _mechanical_mean_inputs (scientific_mechanics.jl:47-56) */
static void mech_inputs(const orc_sim *s, int k, int64_t i, float *displacement, float *strength) {
    const orc_state *st = &s->st;
    if (s->m.mech_kind[k] == ORC_MECH_PRESSURE) {
        *displacement = (float)st->finite_volumes[i] - (float)st->target_volume[i];
        *strength = (float)st->volume_strength[i];
    } else {
        *displacement = (float)st->boundary_measures[i] - (float)st->target_boundary[i];
        *strength = (float)st->boundary_strength[i];
    }
}

/* scientific_mechanics.jl:63-103 */
void orc_initialize_mechanics(orc_sim *s) {
    for (int k = 0; k < s->m.n_mech; ++k) {
        float *q = mech_column(s, k);
        for (int64_t i = 0; i < s->m.capacity; ++i) {
            if (!s->st.active[i]) {
                q[i] = 0.0f;
                continue;
            }
            if (s->m.mech_init[k] == ORC_MECH_INIT_PRESERVE) continue;
            float displacement, strength;
            mech_inputs(s, k, i, &displacement, &strength);
            float mean = (2.0f * strength) * displacement;
            if (s->m.mech_init[k] == ORC_MECH_INIT_CONSTITUTIVE) {
                q[i] = mean;
            } else {
                float normal = orc_normal_box_muller(s->seed, S_AUX_INITIALIZATION, 0, 0,
                                                     s->m.mech_instance[k], (uint32_t)(i + 1), 0);
                q[i] = mean + sqrtf((2.0f * strength) * mech_noise(s, k)) * normal;
            }
        }
    }
}

/* scientific_mechanics.jl:135-168 */
void orc_advance_mechanics(orc_sim *s, uint64_t mcs, uint32_t subround, uint32_t phase, float interval) {
    for (int k = 0; k < s->m.n_mech; ++k) {
        float *q = mech_column(s, k);
        for (int64_t i = 0; i < s->m.capacity; ++i) {
            if (!s->st.active[i]) {
                q[i] = 0.0f;
                continue;
            }
            float displacement, strength;
            mech_inputs(s, k, i, &displacement, &strength);
            float normal = orc_normal_box_muller(s->seed, S_AUX_EVOLUTION, mcs, subround,
                                                 s->m.mech_instance[k], (uint32_t)(i + 1), phase);
            q[i] = orc_ou_transition(q[i], displacement, strength, s->m.mech_eta[k], mech_noise(s, k),
                                     interval, normal);
        }
    }
}

/* mechanical_work of every component folded in tuple order (scientific_inner_loop.jl:238-242) */
static float mechanical_work(const orc_sim *s, int64_t losing, int64_t gaining, int64_t dl, int64_t dg) {
    float total = 0.0f;
    for (int k = 0; k < s->m.n_mech; ++k) {
        float r = 0.0f;
        if (s->m.mech_kind[k] == ORC_MECH_PRESSURE) {
            if (losing > 0) r -= s->st.volume_pressure[losing - 1];
            if (gaining > 0) r += s->st.volume_pressure[gaining - 1];
        } else {
            if (losing > 0) r += s->st.surface_tension[losing - 1] * (float)dl;
            if (gaining > 0) r += s->st.surface_tension[gaining - 1] * (float)dg;
        }
        total += r;
    }
    return total;
}

/* protocols/scientific.jl:372-385 with zero drive, kinetic and yield terms.
 * exp is evaluated in double precision and rounded once to float. */
static float acceptance_probability(float dh, float mech, float temperature) {
    volatile float d = dh + mech; /* delta_h + mechanical_work */
    if (temperature == 0.0f) {
        float eff = d + 0.0f;
        return eff <= 0.0f ? 1.0f : 0.0f;
    }
    volatile float neg = -(d + 0.0f);
    volatile float q = neg / temperature;
    volatile float lr = (q + 0.0f) + 0.0f;
    if (lr >= 0.0f) return 1.0f;
    return (float)exp((double)lr);
}

/* checkerboard.jl:410-440 and scientific_trackers.jl:341-373 */
static void commit(orc_sim *s, uint32_t recipient, int64_t losing, int64_t gaining, int64_t dl,
                   int64_t dg) {
    orc_state *st = &s->st;
    st->tags[recipient - 1] = gaining > 0 ? 1 : 2;
    st->ids[recipient - 1] = (uint32_t)(gaining > 0 ? gaining : -gaining);
    if (losing > 0) {
        int64_t i = losing - 1;
        st->finite_volumes[i] -= 1;
        st->boundary_measures[i] += dl;
        if (st->finite_volumes[i] == 0) {
            st->active[i] = 0;
            st->cell_types[i] = 0;
            st->target_volume[i] = s->m.retire_target_volume;
            st->volume_strength[i] = s->m.retire_volume_strength;
            st->target_boundary[i] = s->m.retire_target_boundary;
            st->boundary_strength[i] = s->m.retire_boundary_strength;
            st->volume_pressure[i] = 0.0f;
            st->surface_tension[i] = 0.0f;
        }
    } else {
        st->medium_volumes[medium_index(&s->m, losing)] -= 1;
    }
    if (gaining > 0) {
        int64_t i = gaining - 1;
        st->finite_volumes[i] += 1;
        st->boundary_measures[i] += dg;
    } else {
        st->medium_volumes[medium_index(&s->m, gaining)] += 1;
    }
}

/* ------------------------------------------------------------------------- */
/* Claim protocol: checkerboard.jl:273-369                                    */
/* ------------------------------------------------------------------------- */

/* Loops over the sites or attempts of one colour run on several threads when they are long
 * enough. Everything they do is independent per element or an order-free maximum / minimum
 * (claims), and the sums are formed afterwards in element order, so the result is the same for
 * any thread count. */
#define ORC_PAR_MIN 16384

static void atomic_max_u32(uint32_t *p, uint32_t v) {
    uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}
static void atomic_min_u32(uint32_t *p, uint32_t v) {
    uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

void orc_set_threads(int32_t threads) {
#ifdef _OPENMP
    omp_set_num_threads(threads > 0 ? threads : omp_get_num_procs());
#else
    (void)threads;
#endif
}

int32_t orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void claims_run(int32_t n, const uint32_t *idx, const uint8_t *use,
                       const uint32_t *recipient, const int64_t *losing, const int64_t *gaining,
                       const uint32_t *priority, uint32_t *maxp, uint32_t *minid,
                       uint8_t *selected, uint8_t *disp) {
    /* idx maps 0..n-1 to attempt slots; use[] marks attempts that take part */
#pragma omp parallel for if (n >= ORC_PAR_MIN) schedule(static)
    for (int32_t k = 0; k < n; ++k) {
        uint32_t a = idx ? idx[k] : (uint32_t)k;
        if (!use[a]) continue;
        if (losing[a] > 0) atomic_max_u32(&maxp[losing[a]], priority[a]);
        if (gaining[a] > 0) atomic_max_u32(&maxp[gaining[a]], priority[a]);
    }
#pragma omp parallel for if (n >= ORC_PAR_MIN) schedule(static)
    for (int32_t k = 0; k < n; ++k) {
        uint32_t a = idx ? idx[k] : (uint32_t)k;
        if (!use[a]) continue;
        if (losing[a] > 0 && maxp[losing[a]] == priority[a]) atomic_min_u32(&minid[losing[a]], recipient[a]);
        if (gaining[a] > 0 && maxp[gaining[a]] == priority[a]) atomic_min_u32(&minid[gaining[a]], recipient[a]);
    }
#pragma omp parallel for if (n >= ORC_PAR_MIN) schedule(static)
    for (int32_t k = 0; k < n; ++k) {
        uint32_t a = idx ? idx[k] : (uint32_t)k;
        if (!use[a]) continue;
        int wl = losing[a] <= 0 || (maxp[losing[a]] == priority[a] && minid[losing[a]] == recipient[a]);
        int wg = gaining[a] <= 0 || (maxp[gaining[a]] == priority[a] && minid[gaining[a]] == recipient[a]);
        int wins = wl && wg;
        selected[a] = wins ? 1 : 0;
        if (!wins) disp[a] = D_CONFLICT;
    }
}

void orc_claims(int32_t n, const uint32_t *recipient, const uint32_t *losing,
                const uint32_t *gaining, const uint32_t *priority, int32_t capacity,
                uint8_t *selected, uint8_t *dispositions, uint32_t *max_priority,
                uint32_t *min_identity) {
    int64_t *l = (int64_t *)xcalloc((size_t)n, 8);
    int64_t *g = (int64_t *)xcalloc((size_t)n, 8);
    uint8_t *use = (uint8_t *)xcalloc((size_t)n, 1);
    for (int32_t i = 0; i < n; ++i) {
        l[i] = losing[i];
        g[i] = gaining[i];
        use[i] = 1;
        dispositions[i] = D_PENDING;
    }
    for (int32_t c = 0; c <= capacity; ++c) {
        max_priority[c] = 0;
        min_identity[c] = 0xffffffffu;
    }
    claims_run(n, NULL, use, recipient, l, g, priority, max_priority, min_identity, selected,
               dispositions);
    free(l); free(g); free(use);
}

/* ------------------------------------------------------------------------- */
/* MCS drivers                                                                */
/* ------------------------------------------------------------------------- */

/* rng/semantic.jl:427-444 / checkerboard.jl:226-242 / lottery.jl:124-140 */
static void fisher_yates(uint64_t seed, uint32_t stream, uint64_t mcs, uint32_t operation,
                         int32_t count, uint32_t *order) {
    for (int32_t i = 0; i < count; ++i) order[i] = (uint32_t)(i + 1);
    for (int32_t c = count; c >= 2; --c) {
        uint32_t draw = (uint32_t)(count - c);
        uint32_t sel = orc_bounded_uint(seed, stream, mcs, 0, operation, K_GLOBAL, 0, draw,
                                        (uint32_t)c) + 1;
        uint32_t t = order[c - 1];
        order[c - 1] = order[sel - 1];
        order[sel - 1] = t;
    }
}

static void evaluate_selected(orc_sim *s, uint32_t a, uint64_t mcs, uint32_t subround) {
    uint32_t rec = s->att_recipient[a];
    int64_t l = s->att_losing[a], g = s->att_gaining[a];
    int64_t dl, dg;
    boundary_change(s, rec, l, g, &dl, &dg);
    s->att_dl[a] = dl;
    s->att_dg[a] = dg;
    float dh = delta_h(s, rec, l, g, dl, dg);
    float p = acceptance_probability(dh, mechanical_work(s, l, g, dl, dg), s->m.temperature);
    uint32_t w[4];
    orc_rng_words(s->seed, S_ACCEPTANCE, mcs, subround, 0, K_SITE, rec, 0, 0, 0, w);
    float u = orc_uniform_open01_f32(w[0]);
    s->att_disp[a] = u < p ? D_ACCEPTED : D_REJECTED;
    s->att_dh[a] = dh;
}

/* checkerboard.jl:521-586 */
static void checkerboard_mcs(orc_sim *s) {
    uint64_t mcs = s->mcs + 1;
    int32_t K = s->m.proposal.n;
    fisher_yates(s->seed, S_CHECKERBOARD_ORDER, mcs, 0, s->color_count, s->order);
    uint64_t realized = 0, same = 0, boundary = 0, immutable = 0, conflicts = 0, rejected = 0,
             accepted = 0;
    for (int32_t ordinal = 0; ordinal < s->color_count; ++ordinal) {
        int32_t color = (int32_t)s->order[ordinal];
        uint32_t first = s->color_offsets[color - 1], stop = s->color_offsets[color];
        /* half interval of the colour (scientific_mechanics.jl:175-178) */
        float half = (float)(stop - first) / (2.0f * (float)s->n_mutable);
        orc_advance_mechanics(s, mcs, (uint32_t)ordinal, 0, half);
        for (int64_t c = 0; c <= s->m.capacity; ++c) {
            s->cell_max_priority[c] = 0;
            s->cell_min_identity[c] = 0xffffffffu;
        }
        /* candidates: checkerboard.jl:244-271 */
#pragma omp parallel for if (stop - first >= ORC_PAR_MIN) schedule(static)
        for (uint32_t k = first; k < stop; ++k) {
            uint32_t site = s->color_sites[k];
            uint32_t dir = orc_bounded_uint(s->seed, S_DIRECTION, mcs, 0, 0, K_SITE, site, 0,
                                            (uint32_t)K) + 1;
            uint32_t donor;
            int out = construct_attempt(s, site, (int)dir, &s->att_losing[k],
                                        &s->att_gaining[k], &donor);
            s->att_recipient[k] = site;
            s->att_outcome[k] = (uint8_t)out;
            uint32_t w[4];
            orc_rng_words(s->seed, S_CHECKERBOARD_PRIORITY, mcs, 0, 0, K_SITE, site, 0, 0, 0, w);
            s->att_priority[k] = w[0];
            s->att_selected[k] = out == OUT_ACTIONABLE;
            s->att_disp[k] = D_PENDING;
        }
        /* claims over the colour's slot range: checkerboard.jl:281-369 */
        {
            int32_t n = (int32_t)(stop - first);
            uint8_t *use = (uint8_t *)xcalloc((size_t)n, 1);
            for (int32_t k = 0; k < n; ++k) use[k] = s->att_outcome[first + k] == OUT_ACTIONABLE;
            claims_run(n, NULL, use, s->att_recipient + first, s->att_losing + first,
                       s->att_gaining + first, s->att_priority + first, s->cell_max_priority,
                       s->cell_min_identity, s->att_selected + first, s->att_disp + first);
            free(use);
        }
        /* evaluate every selected attempt against the common snapshot, then commit */
#pragma omp parallel for if (stop - first >= ORC_PAR_MIN) schedule(dynamic, 1024)
        for (uint32_t k = first; k < stop; ++k)
            if (s->att_selected[k]) evaluate_selected(s, k, mcs, 0);
        for (uint32_t k = first; k < stop; ++k)
            if (s->att_disp[k] == D_ACCEPTED) {
                commit(s, s->att_recipient[k], s->att_losing[k], s->att_gaining[k], s->att_dl[k],
                       s->att_dg[k]);
                s->dh_last_mcs += (double)s->att_dh[k];
            }
        orc_advance_mechanics(s, mcs, (uint32_t)ordinal, 1, half);
    }
    /* checkerboard.jl:459-505 */
    for (int64_t k = 0; k < s->n_mutable; ++k) {
        switch (s->att_outcome[k]) {
        case OUT_ACTIONABLE:
            realized++;
            conflicts += s->att_disp[k] == D_CONFLICT;
            rejected += s->att_disp[k] == D_REJECTED;
            accepted += s->att_disp[k] == D_ACCEPTED;
            break;
        case OUT_SAME_OWNER: same++; break;
        case OUT_BOUNDARY_NULL: boundary++; break;
        default: immutable++; break;
        }
    }
    uint64_t *r = s->report;
    r[0] = mcs; r[1] = (uint64_t)s->color_count; r[2] = (uint64_t)s->n_mutable;
    r[3] = (uint64_t)s->n_mutable; r[4] = realized; r[5] = same; r[6] = boundary;
    r[7] = immutable; r[8] = conflicts; r[9] = 0; r[10] = rejected; r[11] = accepted;
    s->has_report = 1;
    s->mcs = mcs;
}

static int ticket_better(const uint32_t *left, const uint32_t *right, uint32_t left_site,
                         uint32_t right_site) {
    for (int i = 0; i < 4; ++i)
        if (left[i] != right[i]) return left[i] > right[i];
    return left_site < right_site;
}

/* lottery.jl:328-415 */
static void lottery_mcs(orc_sim *s) {
    uint64_t mcs = s->mcs + 1;
    int32_t R = s->color_count;
    int32_t K = s->m.proposal.n;
    int64_t nm = s->n_mutable;
    fisher_yates(s->seed, S_LOTTERY_ACTIVATION, mcs, 1, R, s->order);
    uint64_t acc[9] = {0};
    uint8_t *use = (uint8_t *)xcalloc((size_t)nm, 1);
    memset(s->activations, 0, (size_t)nm * sizeof(uint32_t));
    float half = 1.0f / (2.0f * (float)R); /* lottery.jl:353 */
    for (int32_t ordinal = 0; ordinal < R; ++ordinal) {
        uint32_t round = s->order[ordinal];
        uint32_t sub = round - 1;
        orc_advance_mechanics(s, mcs, (uint32_t)ordinal, 0, half);
        for (int64_t c = 0; c <= s->m.capacity; ++c) {
            s->cell_max_priority[c] = 0;
            s->cell_min_identity[c] = 0xffffffffu;
        }
#pragma omp parallel for if (nm >= ORC_PAR_MIN) schedule(static)
        for (int64_t i = 0; i < nm; ++i)
            orc_rng_words(s->seed, S_LOTTERY_PRIORITY, mcs, sub, 0, K_SITE, s->mutable_sites[i], 0,
                          0, 0, s->tickets[i]);
        uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
#pragma omp parallel for if (nm >= ORC_PAR_MIN) schedule(static) reduction(+ : a0, a1, a2, a3, a4)
        for (int64_t i = 0; i < nm; ++i) {
            uint32_t site = s->mutable_sites[i];
            uint32_t f = s->csr_offsets[i], e = s->csr_offsets[i + 1];
            uint32_t degree = e - f;
            int eligible = orc_bounded_uint(s->seed, S_LOTTERY_ACTIVATION, mcs, sub, 0, K_SITE, site,
                                            0, (uint32_t)R) < degree + 1;
            int wins = 1;
            for (uint32_t k = f; k < e; ++k) {
                uint32_t p = s->csr_nbrs[k];
                if (ticket_better(s->tickets[p], s->tickets[i], s->mutable_sites[p], site)) wins = 0;
            }
            s->att_selected[i] = 0;
            s->att_disp[i] = D_PENDING;
            use[i] = 0;
            if (eligible && wins) {
                a0++;
                s->activations[i]++;
                uint32_t dir = orc_bounded_uint(s->seed, S_DIRECTION, mcs, sub, 0, K_SITE, site, 0,
                                                (uint32_t)K) + 1;
                uint32_t donor;
                int out = construct_attempt(s, site, (int)dir, &s->att_losing[i],
                                            &s->att_gaining[i], &donor);
                s->att_recipient[i] = site;
                s->att_outcome[i] = (uint8_t)out;
                s->att_priority[i] = s->tickets[i][0];
                if (out == OUT_ACTIONABLE) { a1++; s->att_selected[i] = 1; use[i] = 1; }
                else if (out == OUT_SAME_OWNER) a2++;
                else if (out == OUT_BOUNDARY_NULL) a3++;
                else a4++;
            }
        }
        acc[0] += a0; acc[1] += a1; acc[2] += a2; acc[3] += a3; acc[4] += a4;
        claims_run((int32_t)nm, NULL, use, s->att_recipient, s->att_losing, s->att_gaining,
                   s->att_priority, s->cell_max_priority, s->cell_min_identity, s->att_selected,
                   s->att_disp);
        for (int64_t i = 0; i < nm; ++i)
            if (use[i] && !s->att_selected[i]) acc[5]++;
#pragma omp parallel for if (nm >= ORC_PAR_MIN) schedule(dynamic, 1024)
        for (int64_t i = 0; i < nm; ++i)
            if (s->att_selected[i]) evaluate_selected(s, (uint32_t)i, mcs, sub);
        for (int64_t i = 0; i < nm; ++i)
            if (s->att_selected[i]) {
                if (s->att_disp[i] == D_ACCEPTED) {
                    acc[8]++;
                    commit(s, s->att_recipient[i], s->att_losing[i], s->att_gaining[i], s->att_dl[i],
                           s->att_dg[i]);
                    s->dh_last_mcs += (double)s->att_dh[i];
                } else {
                    acc[7]++;
                }
            }
        orc_advance_mechanics(s, mcs, (uint32_t)ordinal, 1, half);
    }
    free(use);
    uint64_t *r = s->report;
    r[0] = mcs; r[1] = (uint64_t)R; r[2] = (uint64_t)nm * (uint64_t)R; r[3] = acc[0];
    r[4] = acc[1]; r[5] = acc[2]; r[6] = acc[3]; r[7] = acc[4]; r[8] = acc[5]; r[9] = acc[6];
    r[10] = acc[7]; r[11] = acc[8];
    s->has_report = 1;
    s->mcs = mcs;
}

/* Recount of finite volumes and boundary measures from ownership (rebuild_derived_state!). */
static void recount_trackers(orc_sim *s) {
    int64_t cap = s->m.capacity;
    for (int64_t c = 0; c < cap; ++c) {
        s->st.finite_volumes[c] = 0;
        s->st.boundary_measures[c] = 0;
    }
    for (int64_t i = 0; i < s->n; ++i) {
        int64_t o = owner_at(s, (uint32_t)(i + 1));
        if (o > 0) s->st.finite_volumes[o - 1]++;
    }
    for (int64_t k = 0; k < s->n_mutable; ++k) {
        uint32_t site = s->mutable_sites[k];
        int64_t o = owner_at(s, site);
        if (o <= 0) continue;
        for (int d = 0; d < s->m.surface.n; ++d) {
            nbr_t nb = realize(&s->m, &s->m.surface, ROLE_SURFACE, site, d);
            if (nb.kind == NB_ABSENT || nb.kind == NB_INVALID) continue;
            int64_t no = nb.kind == NB_MUTABLE ? owner_at(s, nb.site) : nb.owner;
            if (no != o) s->st.boundary_measures[o - 1]++;
        }
    }
}

/* TiledCheckerboardCPM MCS (tiled_checkerboard_device.jl:407-586,688-705,743-806; the host
 * reference tiled_checkerboard.jl:382-463). Tiles of one colour run from one snapshot: a
 * tile's proposals are sequential and see the snapshot trackers plus the tile's own changes;
 * ownership is written at once (tiles of a colour lie beyond each other's halo). After the
 * colour's batch the trackers become the snapshot plus every tile's changes (or a recount)
 * and cells left with zero volume retire. */
static void tiled_mcs(orc_sim *s) {
    uint64_t mcs = s->mcs + 1;
    int32_t C = s->color_count, B = s->tiled_batches, interval = s->tiled_interval;
    uint32_t K = (uint32_t)s->m.proposal.n;
    int64_t cap = s->m.capacity;
    orc_state *st = &s->st;
    fisher_yates(s->seed, S_TILED_ORDER, mcs, 0, C, s->order);
    uint64_t realized = 0, same = 0, boundary = 0, immutable = 0, rejected = 0, accepted = 0;
    uint32_t subround = 0;
    int64_t *t_cell = (int64_t *)xcalloc((size_t)interval * 2 + 2, 8);
    int64_t *t_dvol = (int64_t *)xcalloc((size_t)interval * 2 + 2, 8);
    int64_t *t_dbnd = (int64_t *)xcalloc((size_t)interval * 2 + 2, 8);
    for (int32_t ordinal = 0; ordinal < C; ++ordinal) {
        uint32_t color = s->order[ordinal];
        uint32_t first_tile = s->color_tile_offsets[color - 1], stop_tile = s->color_tile_offsets[color];
        for (int32_t batch = 1; batch <= B; ++batch) {
            subround++;
            uint32_t batch_start = (uint32_t)(batch - 1) * (uint32_t)interval + 1;
            memset(s->acc_dvol, 0, (size_t)cap * 8);
            memset(s->acc_dbnd, 0, (size_t)cap * 8);
            for (uint32_t k = first_tile; k < stop_tile; ++k) {
                uint32_t tile = s->color_tiles[k];
                uint32_t first = s->tile_offsets[tile - 1];
                uint32_t count = s->tile_offsets[tile] - first;
                if (batch_start > count) continue;
                uint32_t stop = batch_start + (uint32_t)interval - 1;
                if (stop > count) stop = count;
                int touched = 0;
                for (uint32_t lp = batch_start; lp <= stop; ++lp) {
                    uint32_t sel = orc_bounded_uint(s->seed, S_TILED_PROPOSAL, mcs, subround, lp, K_SITE,
                                                    tile, 0, count);
                    uint32_t rec = s->tile_sites[first + sel];
                    uint32_t dir = orc_bounded_uint(s->seed, S_TILED_PROPOSAL, mcs, subround, lp, K_SITE,
                                                    tile, 1, K) + 1;
                    int64_t l, g;
                    uint32_t donor;
                    int out = construct_attempt(s, rec, (int)dir, &l, &g, &donor);
                    if (out == OUT_SAME_OWNER) { same++; continue; }
                    if (out == OUT_BOUNDARY_NULL) { boundary++; continue; }
                    if (out == OUT_IMMUTABLE) { immutable++; continue; }
                    realized++;
                    int64_t dl, dg;
                    boundary_change(s, rec, l, g, &dl, &dg);
                    float dh = delta_h(s, rec, l, g, dl, dg);
                    float p = acceptance_probability(dh, 0.0f, s->m.temperature);
                    uint32_t w[4];
                    orc_rng_words(s->seed, S_TILED_PROPOSAL, mcs, subround, lp, K_SITE, tile, 0, 0, 2, w);
                    if (!(orc_uniform_open01_f32(w[0]) < p)) { rejected++; continue; }
                    accepted++;
                    s->dh_last_mcs += (double)dh;
                    st->tags[rec - 1] = g > 0 ? 1 : 2;
                    st->ids[rec - 1] = (uint32_t)(g > 0 ? g : -g);
                    int64_t cells[2] = {l, g}, dv[2] = {-1, 1}, db[2] = {dl, dg};
                    for (int e = 0; e < 2; ++e) {
                        if (cells[e] <= 0) {
                            st->medium_volumes[medium_index(&s->m, cells[e])] += (int32_t)dv[e];
                            continue;
                        }
                        int64_t i = cells[e] - 1;
                        /* the tile's view: snapshot trackers plus its own changes */
                        st->finite_volumes[i] += (int32_t)dv[e];
                        st->boundary_measures[i] += db[e];
                        int j = 0;
                        while (j < touched && t_cell[j] != i) ++j;
                        if (j == touched) { t_cell[touched] = i; t_dvol[touched] = 0; t_dbnd[touched] = 0; ++touched; }
                        t_dvol[j] += dv[e];
                        t_dbnd[j] += db[e];
                    }
                }
                /* back to the snapshot for the next tile; the changes join the colour's sum */
                for (int j = 0; j < touched; ++j) {
                    st->finite_volumes[t_cell[j]] -= (int32_t)t_dvol[j];
                    st->boundary_measures[t_cell[j]] -= t_dbnd[j];
                    s->acc_dvol[t_cell[j]] += t_dvol[j];
                    s->acc_dbnd[t_cell[j]] += t_dbnd[j];
                }
            }
            if (s->tiled_recount) {
                recount_trackers(s);
            } else {
                for (int64_t c = 0; c < cap; ++c) {
                    st->finite_volumes[c] += (int32_t)s->acc_dvol[c];
                    st->boundary_measures[c] += s->acc_dbnd[c];
                }
            }
            for (int64_t c = 0; c < cap; ++c)
                if (st->finite_volumes[c] == 0 && st->active[c]) {
                    st->active[c] = 0;
                    st->cell_types[c] = 0;
                    st->target_volume[c] = s->m.retire_target_volume;
                    st->volume_strength[c] = s->m.retire_volume_strength;
                    st->target_boundary[c] = s->m.retire_target_boundary;
                    st->boundary_strength[c] = s->m.retire_boundary_strength;
                }
        }
    }
    free(t_cell); free(t_dvol); free(t_dbnd);
    uint64_t *r = s->report;
    r[0] = mcs; r[1] = subround; r[2] = (uint64_t)s->n_mutable; r[3] = (uint64_t)s->n_mutable;
    r[4] = realized; r[5] = same; r[6] = boundary; r[7] = immutable; r[8] = 0; r[9] = 0;
    r[10] = rejected; r[11] = accepted;
    s->has_report = 1;
    s->mcs = mcs;
}

/* sequential.jl:488-567 */
static void sequential_mcs(orc_sim *s) {
    uint64_t mcs = s->mcs + 1;
    int64_t nm = s->n_mutable;
    uint32_t K = (uint32_t)s->m.proposal.n;
    uint64_t realized = 0, same = 0, boundary = 0, immutable = 0, rejected = 0, accepted = 0;
    orc_advance_mechanics(s, mcs, 0, 0, 0.5f); /* sequential.jl:578-579 */
    for (int64_t raw = 1; raw <= nm; ++raw) {
        uint32_t id = (uint32_t)raw;
        uint32_t mi = orc_bounded_uint(s->seed, S_RECIPIENT, mcs, 0, 0, K_GLOBAL, id, 0,
                                       (uint32_t)nm) + 1;
        uint32_t rec = s->mutable_sites[mi - 1];
        uint32_t dir = orc_bounded_uint(s->seed, S_DIRECTION, mcs, 0, 0, K_GLOBAL, id, 0, K) + 1;
        int64_t l, g;
        uint32_t donor;
        int out = construct_attempt(s, rec, (int)dir, &l, &g, &donor);
        if (out == OUT_SAME_OWNER) { same++; continue; }
        if (out == OUT_BOUNDARY_NULL) { boundary++; continue; }
        if (out == OUT_IMMUTABLE) { immutable++; continue; }
        realized++;
        int64_t dl, dg;
        boundary_change(s, rec, l, g, &dl, &dg);
        float dh = delta_h(s, rec, l, g, dl, dg);
        float p = acceptance_probability(dh, mechanical_work(s, l, g, dl, dg), s->m.temperature);
        uint32_t w[4];
        orc_rng_words(s->seed, S_ACCEPTANCE, mcs, 0, 0, K_GLOBAL, id, 0, 0, 0, w);
        float u = orc_uniform_open01_f32(w[0]);
        if (u < p) {
            commit(s, rec, l, g, dl, dg);
            s->dh_last_mcs += (double)dh;
            accepted++;
        } else {
            rejected++;
        }
    }
    orc_advance_mechanics(s, mcs, 0, 1, 0.5f);
    uint64_t *r = s->report;
    r[0] = mcs; r[1] = 1; r[2] = (uint64_t)nm; r[3] = (uint64_t)nm; r[4] = realized;
    r[5] = same; r[6] = boundary; r[7] = immutable; r[8] = 0; r[9] = 0; r[10] = rejected;
    r[11] = accepted;
    s->has_report = 1;
    s->mcs = mcs;
}

/* evaluate_copy + acceptance_inputs for one proposal (recipient, 1-based direction):
 * out = {delta_h, mechanical_work}; returns the attempt outcome (1 = actionable). */
int32_t orc_copy_evaluation(const orc_sim *s, uint32_t recipient, int32_t direction, float out[2]) {
    int64_t l, g, dl = 0, dg = 0;
    uint32_t donor;
    int outcome = construct_attempt(s, recipient, direction, &l, &g, &donor);
    out[0] = out[1] = 0.0f;
    if (outcome != OUT_ACTIONABLE) return outcome;
    boundary_change(s, recipient, l, g, &dl, &dg);
    out[0] = delta_h(s, recipient, l, g, dl, dg);
    out[1] = mechanical_work(s, l, g, dl, dg);
    return outcome;
}

int32_t orc_copy_energy_fold(const orc_sim *s, uint32_t recipient, int32_t direction, double *out) {
    int64_t l, g, dl = 0, dg = 0;
    uint32_t donor;
    int outcome = construct_attempt(s, recipient, direction, &l, &g, &donor);
    *out = 0.0;
    if (outcome != OUT_ACTIONABLE) return outcome;
    boundary_change(s, recipient, l, g, &dl, &dg);
    *out = energy_fold(s, recipient, l, g, dl, dg);
    return outcome;
}

int32_t orc_step(orc_sim *s, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        s->dh_last_mcs = 0.0;
        if (s->algorithm == ORC_ALG_CHECKERBOARD) checkerboard_mcs(s);
        else if (s->algorithm == ORC_ALG_LOTTERY) lottery_mcs(s);
        else if (s->algorithm == ORC_ALG_TILED) tiled_mcs(s);
        else sequential_mcs(s);
        s->dh_total += s->dh_last_mcs;
    }
    return 0;
}

void orc_accepted_energy(const orc_sim *s, double out[2]) {
    out[0] = s->dh_last_mcs;
    out[1] = s->dh_total;
}

int32_t orc_report(const orc_sim *s, uint64_t out[12]) {
    if (!s->has_report) return 1;
    memcpy(out, s->report, sizeof(s->report));
    return 0;
}

/* scientific_trackers.jl:399-423: rebuild trackers from ownership. */
int32_t orc_tracker_errors(const orc_sim *s) {
    int64_t cap = s->m.capacity;
    int32_t *fv = (int32_t *)xcalloc((size_t)cap, 4);
    int32_t mv[ORC_MAX_MEDIA] = {0};
    int64_t *bm = (int64_t *)xcalloc((size_t)cap, 8);
    int32_t errors = 0;
    for (int64_t i = 0; i < s->n; ++i) {
        int64_t o = owner_at(s, (uint32_t)(i + 1));
        if (o > 0) {
            if (o > cap) { errors |= 8; continue; }
            fv[o - 1]++;
        } else {
            mv[medium_index(&s->m, o)]++;
        }
    }
    for (int64_t k = 0; k < s->n_mutable; ++k) {
        uint32_t site = s->mutable_sites[k];
        int64_t o = owner_at(s, site);
        if (o <= 0) continue;
        for (int d = 0; d < s->m.surface.n; ++d) {
            nbr_t nb = realize(&s->m, &s->m.surface, ROLE_SURFACE, site, d);
            if (nb.kind == NB_ABSENT || nb.kind == NB_INVALID) continue;
            int64_t no = nb.kind == NB_MUTABLE ? owner_at(s, nb.site) : nb.owner;
            if (no != o) bm[o - 1]++;
        }
    }
    for (int64_t c = 0; c < cap; ++c) {
        if (fv[c] != s->st.finite_volumes[c]) errors |= 1;
        if (s->st.active[c] && bm[c] != s->st.boundary_measures[c]) errors |= 4;
    }
    for (int i = 0; i < s->m.n_media; ++i)
        if (mv[i] != s->st.medium_volumes[i]) errors |= 2;
    free(fv);
    free(bm);
    return errors;
}

int32_t orc_sample_transitions(orc_sim *s, uint64_t seed_base, uint64_t replicas,
                               const int64_t *labels, int32_t n_labels, uint64_t *counts) {
    orc_state pristine;
    int64_t cap = s->m.capacity;
    pristine.tags = (uint8_t *)xcalloc((size_t)s->n, 1);
    pristine.ids = (uint32_t *)xcalloc((size_t)s->n, 4);
    pristine.finite_volumes = (int32_t *)xcalloc((size_t)cap, 4);
    pristine.medium_volumes = (int32_t *)xcalloc(ORC_MAX_MEDIA, 4);
    pristine.boundary_measures = (int64_t *)xcalloc((size_t)cap, 8);
    pristine.active = (uint8_t *)xcalloc((size_t)cap, 1);
    pristine.cell_types = (uint32_t *)xcalloc((size_t)cap, 4);
    pristine.target_volume = (double *)xcalloc((size_t)cap, 8);
    pristine.volume_strength = (double *)xcalloc((size_t)cap, 8);
    pristine.target_boundary = (int64_t *)xcalloc((size_t)cap, 8);
    pristine.boundary_strength = (double *)xcalloc((size_t)cap, 8);
    orc_store_state(s, &pristine);
    int32_t status = 0;
    for (uint64_t r = 0; r < replicas; ++r) {
        orc_load_state(s, &pristine);
        s->seed = seed_base + r;
        s->mcs = 0;
        orc_step(s, 1);
        uint64_t id = 0, scale = 1;
        for (int64_t i = 0; i < s->n; ++i) {
            int64_t o = owner_at(s, (uint32_t)(i + 1));
            int32_t lab = -1;
            for (int32_t k = 0; k < n_labels; ++k)
                if (labels[k] == o) lab = k;
            if (lab < 0) { status = 1; break; }
            id += (uint64_t)lab * scale;
            scale *= (uint64_t)n_labels;
        }
        counts[id]++;
    }
    orc_load_state(s, &pristine);
    free(pristine.tags); free(pristine.ids); free(pristine.finite_volumes);
    free(pristine.medium_volumes); free(pristine.boundary_measures); free(pristine.active);
    free(pristine.cell_types); free(pristine.target_volume); free(pristine.volume_strength);
    free(pristine.target_boundary); free(pristine.boundary_strength);
    return status;
}

/* Global energies (scientific_components.jl:97-107, 315-340, 525-538) in FP64. */
double orc_total_energy(const orc_sim *s) {
    const orc_model *m = &s->m;
    double total = 0.0;
    int has_vol = 0, has_con = 0, has_bnd = 0;
    for (int t = 0; t < m->n_energies; ++t) {
        has_vol |= m->energy_order[t] == ORC_TERM_VOLUME;
        has_con |= m->energy_order[t] == ORC_TERM_CONTACT;
        has_bnd |= m->energy_order[t] == ORC_TERM_BOUNDARY;
    }
    for (int64_t c = 0; c < m->capacity; ++c) {
        if (!s->st.active[c]) continue;
        if (has_vol) {
            double d = (double)s->st.finite_volumes[c] - s->st.target_volume[c];
            total += s->st.volume_strength[c] * d * d;
        }
        if (has_bnd) {
            double d = (double)s->st.boundary_measures[c] - (double)s->st.target_boundary[c];
            total += s->st.boundary_strength[c] * d * d;
        }
    }
    if (has_con) {
        int nt = m->n_types;
        for (int64_t k = 0; k < s->n_mutable; ++k) {
            uint32_t site = s->mutable_sites[k];
            int64_t o = owner_at(s, site);
            uint32_t ot = owner_type(s, o);
            for (int d = 0; d < m->contact.n; ++d) {
                nbr_t nb = realize(m, &m->contact, ROLE_CONTACT, site, d);
                if (nb.kind == NB_ABSENT || nb.kind == NB_INVALID) continue;
                if (nb.kind == NB_MUTABLE && nb.site < site) continue; /* each mutable pair once */
                int64_t no = nb.kind == NB_MUTABLE ? owner_at(s, nb.site) : nb.owner;
                if (no == o) continue;
                uint32_t nt2 = owner_type(s, no);
                total += m->contact.weight[d] * m->J[(ot - 1) * nt + (nt2 - 1)];
            }
        }
    }
    return total;
}
