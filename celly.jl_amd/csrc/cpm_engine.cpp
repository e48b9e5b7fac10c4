// Host side of the MI355X CPM engine: schedule compilation (realized greedy colouring,
// conflict degrees), replica groups and their device arenas, stepping with optional
// HIP-graph chunks, state transfer and the C-ABI declared in include/cpm_mi355x.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cpm_mi355x.h"
#include "cpm_device.h"

extern "C" int cpm_internal_tile_config(int id, int *dim, int *halo, int *tx, int *ty, int *tz);
extern "C" hipError_t cpm_internal_launch_pass(int config, int algorithm, const cpm::PassArgs *a,
                                               hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_mech_advance(const cpm::RepDev *reps, const cpm::McsPlan *plan, int slot,
                                                       uint32_t capacity, uint32_t pass, uint32_t phase,
                                                       uint32_t plan_k, hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_mech_init(const cpm::RepDev *reps, int slot, uint32_t capacity,
                                                    hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_propose(int config, int algorithm, const cpm::PassArgs *a,
                                                  hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_commit(const cpm::PassArgs *a, hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_tiled_batch(const cpm::RepDev *reps, const cpm::RepDyn *dyn, int slot,
                                                      uint32_t ordinal, uint32_t batch, uint32_t max_tiles,
                                                      uint32_t interval, uint32_t staged, hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_tiled_reconcile(const cpm::RepDev *reps, const cpm::RepDyn *dyn, int slot,
                                                          uint32_t capacity, hipStream_t stream);
extern "C" void cpm_internal_time_next(hipEvent_t start, hipEvent_t stop);
extern "C" hipError_t cpm_internal_launch_begin(const cpm::RepDev *reps, cpm::RepDyn *dyn, cpm::McsPlan *plan,
                                                int nslots, hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_plan(const cpm::RepDev *reps, const cpm::RepDyn *dyn, cpm::McsPlan *plan,
                                               int nslots, int K, hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_plan_end(const cpm::RepDev *reps, cpm::RepDyn *dyn,
                                                   const cpm::McsPlan *plan, int nslots, int K, hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_set_dyn(cpm::RepDyn *dyn, int slot, uint64_t mcs,
                                                  uint64_t target, uint32_t in_use, uint32_t mask,
                                                  hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_copy_probe(const void *src, void *dst, size_t n4,
                                                     hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_energy(const cpm::RepDev *reps, int slot, double *buf,
                                                 hipStream_t stream);
extern "C" hipError_t cpm_internal_launch_report(const cpm::RepDev *reps, const uint32_t *slots,
                                                 uint32_t ntiles, int slot,
                                                 unsigned long long *out, hipStream_t stream);

namespace {

thread_local std::string g_create_error;

struct Failure {
    int32_t code;
    std::string message;
};

[[noreturn]] void fail(int32_t code, const std::string &msg) { throw Failure{code, msg}; }

void hip_check(hipError_t err, const char *what) {
    if (err != hipSuccess)
        fail(CPM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(err));
}

// rng/semantic.jl:191-195
uint64_t mix64(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

// Host Philox for cpm_ensemble_seed (rng/semantic.jl:144-187).
void philox_host(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9e3779b9u;
            k1 += 0xbb67ae85u;
        }
        const uint64_t p0 = (uint64_t)0xd2511f53u * c[0];
        const uint64_t p1 = (uint64_t)0xcd9e8d57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
    }
}

constexpr uint64_t kGenerationDomain = 0xd2b74407b1ce6e93ull;

cpm::FastDiv make_div(uint32_t d) {
    cpm::FastDiv f;
    f.d = d;
    f.m = d <= 1 ? 0u : (uint32_t)((1ull << 32) / d);
    return f;
}

cpm::Bound make_bound(uint32_t bound) {
    cpm::Bound b;
    b.threshold = (uint32_t)(0u - bound) % bound;
    b.div = make_div(bound);
    return b;
}

template <typename T>
T *dev_alloc(size_t count, uint64_t &bytes) {
    if (count == 0) count = 1;
    void *p = nullptr;
    hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    bytes += count * sizeof(T);
    return static_cast<T *>(p);
}

void dev_free(void *p) {
    if (p) (void)hipFree(p);
}

// Device owner words and cell records (host copies) -> the per-site and per-cell arrays of
// cpm_download_state (null outputs are skipped).
void unpack_state(const int32_t *words, size_t n, const cpm::CellRec *cells, size_t cap, uint8_t *tags,
                  uint32_t *ids, int32_t *fv, int64_t *bm, uint8_t *active, uint32_t *ctype, double *tv,
                  double *vs, int64_t *tb, double *bs) {
    if (words)
        for (size_t i = 0; i < n; ++i) {
            const int32_t w = words[i];
            if (tags) tags[i] = w > 0 ? 1 : 2;
            if (ids) ids[i] = w > 0 ? cpm::word_cell(w) : (uint32_t)-w;
        }
    if (cells)
        for (size_t c = 0; c < cap; ++c) {
            const cpm::CellRec &q = cells[c + 1];
            if (fv) fv[c] = q.vol;
            if (bm) bm[c] = q.bnd;
            if (active) active[c] = q.active;
            if (ctype) ctype[c] = q.ctype;
            if (tv) tv[c] = q.tv;
            if (vs) vs[c] = q.vs;
            if (tb) tb[c] = q.tb;
            if (bs) bs[c] = q.bs;
        }
}

// ---------------------------------------------------------------------------
// Replica: one lattice with its compiled schedule and device state
// ---------------------------------------------------------------------------
struct Replica {
    cpm_model_desc desc{};
    std::vector<uint8_t> mask, imm_tags;
    std::vector<uint32_t> imm_ids;
    bool has_mask = false;
    int64_t n = 0, nmut = 0;
    int halo = 1;
    int config = 0;
    int rounds = 0;
    std::vector<int32_t> conf_off;  // flattened union offsets (x,y,z)
    std::vector<uint8_t> colors;    // checkerboard realized colouring
    std::vector<uint8_t> degrees;   // lottery conflict degrees
    bool uniform_degree = true;
    bool per_site_color = false;
    // checkerboard colourings without a fast periodic pattern: per tile, the tile positions
    // of each colour (clist) and, per tile and colour, their range (coff, rounds + 1 per tile)
    std::vector<uint16_t> clist;
    std::vector<uint32_t> coff;
    // TiledCheckerboardCPM layout: tile edge, grid and reach per axis, halo radius, halo
    // colour of every tile (1-based), mutable sites by tile, tiles by colour, the switching
    // interval, batches per colour and the most tiles of any colour
    int tile_edge[3] = {1, 1, 1}, tile_grid[3] = {1, 1, 1}, tile_reach[3] = {0, 0, 0};
    int tiled_halo = 0, tiled_interval = 0, tiled_batches = 0;
    int tiled_stage[3] = {0, 0, 0};  // staged box per tile (0: the global-memory kernel)
    uint32_t tiled_staged = 0;       // positions of the staged box
    uint32_t tiled_max_tiles = 0;
    std::vector<uint16_t> tile_colors;
    std::vector<uint32_t> tile_sites, tile_offsets, color_tiles, color_offsets;
    cpm::RepDev dev{};
    // device buffers
    int32_t *d_lattice = nullptr;
    uint8_t *d_mask = nullptr, *d_color = nullptr, *d_degree = nullptr;
    uint16_t *d_clist = nullptr;
    uint32_t *d_coff = nullptr;
    int32_t *d_mvol = nullptr;
    cpm::CellRec *d_cells = nullptr;
    uint32_t *d_tile_sites = nullptr, *d_tile_offsets = nullptr, *d_color_tiles = nullptr, *d_color_offsets = nullptr;
    int32_t *d_tile_dvol = nullptr, *d_tile_dbnd = nullptr, *d_tile_spill = nullptr;
    unsigned long long *d_tiled_counts = nullptr;
    uint64_t bytes = 0;
    uint64_t estimate = 0;      // estimate_replica_bytes at admission
    uint64_t base_mcs = 0;      // MCS counter when the state was last set
    bool report_cached = false;  // cached_report valid while mcs == cached_mcs
    uint64_t cached_mcs = 0;
    uint64_t cached_report[12] = {};
    double cached_dh = 0.0;      // FP64 sum of the accepted copies' delta_h of that MCS (NaN: tiled)
    int group = -1;

    ~Replica() {
        dev_free(d_lattice); dev_free(d_mask); dev_free(d_color); dev_free(d_degree);
        dev_free(d_clist); dev_free(d_coff);
        dev_free(d_cells); dev_free(d_mvol);
        dev_free(d_tile_sites); dev_free(d_tile_offsets); dev_free(d_color_tiles); dev_free(d_color_offsets);
        dev_free(d_tile_dvol); dev_free(d_tile_dbnd); dev_free(d_tile_spill); dev_free(d_tiled_counts);
    }
};

int max_abs_offset(const cpm_relation_desc &r, int ndim) {
    int h = 0;
    for (int k = 0; k < r.count; ++k)
        for (int a = 0; a < ndim; ++a) h = std::max(h, std::abs(r.offsets[k][a]));
    return h;
}

bool has_term(const cpm_model_desc &d, int term) {
    for (int t = 0; t < d.n_energies; ++t)
        if (d.energy_order[t] == term) return true;
    return false;
}

// Process-wide tile-shape preference per (algorithm, dimension, halo); -1 = default shape.
std::atomic<int> g_shape_pref[3][4][3];

// Process-wide switch of the row kernels (2D periodic von Neumann checkerboards with a
// 2 x 2 colour pattern); on by default.
std::atomic<int> g_row_kernels{1};

// Process-wide switch of the staged TiledCheckerboardCPM kernel (tile box and halo in LDS);
// on by default, read when a model is compiled.
std::atomic<int> g_tiled_staging{1};

int default_shape(int ndim, int halo) { return ndim == 2 ? (halo == 1 ? 0 : 1) : (halo == 1 ? 2 : 3); }

int tile_shape_for(int algorithm, int ndim, int halo) {
    if (algorithm >= 1 && algorithm <= 2 && ndim >= 2 && ndim <= 3 && halo >= 1 && halo <= 2) {
        const int p = g_shape_pref[algorithm][ndim][halo].load();
        if (p > 0) return p - 1;  // stored as shape + 1 (0 = default)
    }
    return default_shape(ndim, halo);
}

void validate_relation(const cpm_relation_desc &r, int ndim, const char *name) {
    if (r.count < 1 || r.count > CPM_MAX_DIRS)
        fail(CPM_E_INVALID, std::string(name) + " relation needs 1..128 directions");
    for (int k = 0; k < r.count; ++k) {
        bool nonzero = false;
        for (int a = 0; a < 3; ++a) {
            if (a >= ndim && r.offsets[k][a] != 0)
                fail(CPM_E_INVALID, std::string(name) + " relation offset has a component beyond ndim");
            nonzero |= r.offsets[k][a] != 0;
        }
        if (!nonzero) fail(CPM_E_INVALID, std::string(name) + " relation offsets must be nonzero");
        if (!(r.weights[k] >= 0.0) || !std::isfinite(r.weights[k]))
            fail(CPM_E_INVALID, std::string(name) + " relation weights must be finite and non-negative");
    }
}

// Host realization of a neighbour position (cartesian.jl:445-525).
// Returns 0-based linear index of the realized site, or -1 if absent/invalid/exterior.
struct HostGeom {
    int ndim;
    int dims[3];
    int face_kind[3][2];
    uint32_t face_medium[3][2];
    const uint8_t *mask;
    bool simple;
};

int64_t realize_site(const HostGeom &g, const int c0[3], const int32_t *off) {
    int c[3] = {c0[0] + off[0], c0[1] + off[1], c0[2] + off[2]};
    bool closed = false, fixed = false, invalid = false;
    uint32_t fixed_medium = 0;
    for (int a = 0; a < g.ndim; ++a) {
        const int n = g.dims[a];
        int v = c[a];
        int face;
        if (v < 0) face = 0;
        else if (v >= n) face = 1;
        else continue;
        const int kind = g.face_kind[a][face];
        if (kind == CPM_FACE_PERIODIC) {
            v %= n;
            if (v < 0) v += n;
            c[a] = v;
        } else if (kind == CPM_FACE_CLOSED) {
            invalid |= fixed;
            closed = true;
        } else {
            const uint32_t m = g.face_medium[a][face];
            const bool incompatible = fixed && fixed_medium != m;
            invalid |= closed | incompatible;
            fixed = true;
            fixed_medium = m;
        }
    }
    if (invalid || closed || fixed) return -1;
    return (int64_t)c[0] + (int64_t)g.dims[0] * ((int64_t)c[1] + (int64_t)g.dims[1] * c[2]);
}

// Realized mutable conflict neighbours (distinct, excluding the site itself).
int conflict_neighbours(const HostGeom &g, const std::vector<int32_t> &offs, int64_t i,
                        int64_t *out) {
    int c[3];
    c[0] = (int)(i % g.dims[0]);
    c[1] = (int)((i / g.dims[0]) % g.dims[1]);
    c[2] = (int)(i / ((int64_t)g.dims[0] * g.dims[1]));
    const int k = (int)offs.size() / 3;
    int count = 0;
    for (int d = 0; d < k; ++d) {
        const int64_t j = realize_site(g, c, &offs[3 * d]);
        if (j < 0 || j == i) continue;
        if (g.mask && !g.mask[j]) continue;
        bool dup = false;
        for (int q = 0; q < count; ++q) dup |= out[q] == j;
        if (!dup) out[count++] = j;
    }
    return count;
}

// Group configuration of TiledCheckerboardCPM replicas (outside the tile configurations 0..127).
constexpr int kTiledConfig = 256;
// Touched-cell entries the tiled kernels keep in LDS per tile (as cpm::kListLds).
constexpr uint32_t kTiledListLds = 8;

// Largest per-axis offset of a relation (its radius).
int relation_radius(const cpm_relation_desc &rel, int ndim) { return max_abs_offset(rel, ndim); }

// TiledCheckerboardCPM layout (tiled_checkerboard.jl:87-196; tiled_checkerboard_device.jl:
// 28-121): halo radius of the declared accesses; tile grid and reach; greedy halo colouring
// of the tiles in linear order, first axis fastest (two tiles conflict when their tile
// separation, periodic where the axis is, lies within the reach on every axis; only tiles
// within the reach can conflict, so the scan visits those); the mutable sites of every tile
// (x fastest inside a tile); the tiles of each colour; switching interval and batches.
void compile_tiled(Replica &r, const HostGeom &g) {
    const cpm_model_desc &d = r.desc;
    const int nd = d.ndim;
    int halo = std::max(relation_radius(d.proposal, nd), relation_radius(d.surface, nd));
    for (int t = 0; t < d.n_energies; ++t) {
        int rad;
        if (d.energy_order[t] == CPM_TERM_CONTACT) rad = relation_radius(d.contact, nd);
        else if (d.energy_order[t] == CPM_TERM_BOUNDARY) rad = relation_radius(d.surface, nd);
        else continue;
        halo = std::max(halo, std::max(rad, 1));  // dependency radius at least 1
    }
    r.tiled_halo = halo;
    bool periodic[3] = {false, false, false};
    uint64_t ntiles = 1;
    for (int a = 0; a < 3; ++a) {
        const int edge = a < nd ? (d.tile_size[a] > 0 ? d.tile_size[a] : (nd == 2 ? 4 : 2)) : 1;
        r.tile_edge[a] = edge;
        r.tile_grid[a] = a < nd ? (g.dims[a] + edge - 1) / edge : 1;
        r.tile_reach[a] = a < nd ? (halo + edge - 1) / edge : 0;
        periodic[a] = a < nd && g.face_kind[a][0] == CPM_FACE_PERIODIC;
        ntiles *= (uint64_t)r.tile_grid[a];
    }
    if (ntiles > 0x7fffffffull) fail(CPM_E_UNSUPPORTED, "too many tiles");
    const int gx = r.tile_grid[0], gy = r.tile_grid[1];
    auto tile_coords = [&](uint64_t t, int c[3]) {
        c[0] = (int)(t % (uint64_t)gx);
        c[1] = (int)((t / (uint64_t)gx) % (uint64_t)gy);
        c[2] = (int)(t / ((uint64_t)gx * (uint64_t)gy));
    };
    // visits the tiles within `reach + extra` tiles of tile t (wrapped on periodic axes)
    auto around = [&](uint64_t t, int extra, auto &&fn) {
        int c[3];
        tile_coords(t, c);
        const int rx = r.tile_reach[0] + extra, ry = r.tile_reach[1] + extra, rz = nd == 3 ? r.tile_reach[2] + extra : 0;
        for (int dz = -rz; dz <= rz; ++dz)
            for (int dy = -ry; dy <= ry; ++dy)
                for (int dx = -rx; dx <= rx; ++dx) {
                    const int off[3] = {dx, dy, dz};
                    int q[3];
                    bool ok = true;
                    for (int a = 0; a < 3 && ok; ++a) {
                        int v = c[a] + off[a];
                        const int n = r.tile_grid[a];
                        if (periodic[a]) v = ((v % n) + n) % n;
                        else if (v < 0 || v >= n) ok = false;
                        q[a] = v;
                    }
                    if (ok) fn((uint64_t)q[0] + (uint64_t)gx * ((uint64_t)q[1] + (uint64_t)gy * (uint64_t)q[2]));
                }
    };
    r.tile_colors.assign((size_t)ntiles, 0);
    int maxc = 0;
    std::vector<uint64_t> taken(8, 0);  // colours 1..511 as bits
    for (uint64_t t = 0; t < ntiles; ++t) {
        std::fill(taken.begin(), taken.end(), 0ull);
        around(t, 0, [&](uint64_t u) {
            if (u < t) {
                const int col = r.tile_colors[(size_t)u];
                taken[(size_t)col >> 6] |= 1ull << (col & 63);
            }
        });
        int col = 1;
        while (col < 511 && ((taken[(size_t)col >> 6] >> (col & 63)) & 1ull)) ++col;
        if (col > 255) fail(CPM_E_UNSUPPORTED, "tile colouring exceeds 255 colours");
        r.tile_colors[(size_t)t] = (uint16_t)col;
        maxc = std::max(maxc, col);
    }
    // Tiles of one colour run concurrently and read the lattice within `halo` of their own
    // sites, so no two may lie within `halo` of each other on every axis. The colouring
    // guarantees this unless a periodic axis ends in a partial tile, whose seam can bring
    // tiles closer than their tile separation; such layouts are checked site by site.
    bool seam = false;
    for (int a = 0; a < nd; ++a) seam = seam || (periodic[a] && g.dims[a] % r.tile_edge[a] != 0);
    if (seam) {
        auto separation = [&](int a, int ta, int tb) {
            const int n = g.dims[a], e = r.tile_edge[a];
            const int a0 = ta * e, a1 = std::min(a0 + e, n) - 1, b0 = tb * e, b1 = std::min(b0 + e, n) - 1;
            if (ta == tb) return 0;
            const int lo0 = ta < tb ? a0 : b0, lo1 = ta < tb ? a1 : b1, hi0 = ta < tb ? b0 : a0, hi1 = ta < tb ? b1 : a1;
            int sep = hi0 - lo1;
            if (periodic[a]) sep = std::min(sep, lo0 + n - hi1);
            return sep;
        };
        for (uint64_t t = 0; t < ntiles; ++t) {
            int c[3];
            tile_coords(t, c);
            around(t, 1, [&](uint64_t u) {
                if (u == t || r.tile_colors[(size_t)u] != r.tile_colors[(size_t)t]) return;
                int q[3];
                tile_coords(u, q);
                bool near = true;
                for (int a = 0; a < nd; ++a) near = near && separation(a, c[a], q[a]) <= halo;
                if (near)
                    fail(CPM_E_UNSUPPORTED, "tiles of one colour meet within the halo across a periodic seam; "
                                            "choose a tile size that divides the periodic axes");
            });
        }
    }
    r.rounds = maxc;
    r.color_offsets.assign((size_t)maxc + 1, 0);
    r.color_tiles.assign((size_t)ntiles, 0);
    {
        std::vector<uint32_t> count((size_t)maxc + 2, 0);
        for (uint64_t t = 0; t < ntiles; ++t) count[r.tile_colors[(size_t)t]]++;
        uint32_t at = 0;
        for (int col = 1; col <= maxc; ++col) {
            r.color_offsets[(size_t)col - 1] = at;
            at += count[(size_t)col];
            r.tiled_max_tiles = std::max(r.tiled_max_tiles, count[(size_t)col]);
        }
        r.color_offsets[(size_t)maxc] = at;
        std::vector<uint32_t> fill(r.color_offsets.begin(), r.color_offsets.end());
        for (uint64_t t = 0; t < ntiles; ++t) r.color_tiles[fill[r.tile_colors[(size_t)t] - 1]++] = (uint32_t)(t + 1);
    }
    r.tile_offsets.assign((size_t)ntiles + 1, 0);
    r.tile_sites.clear();
    r.tile_sites.reserve((size_t)r.nmut);
    uint32_t maxsites = 0;
    const int X = g.dims[0], Y = g.dims[1], Z = g.dims[2];
    for (uint64_t t = 0; t < ntiles; ++t) {
        r.tile_offsets[(size_t)t] = (uint32_t)r.tile_sites.size();
        int c[3];
        tile_coords(t, c);
        const int x0 = c[0] * r.tile_edge[0], y0 = c[1] * r.tile_edge[1], z0 = c[2] * r.tile_edge[2];
        const int x1 = std::min(x0 + r.tile_edge[0], X), y1 = std::min(y0 + r.tile_edge[1], Y),
                  z1 = std::min(z0 + r.tile_edge[2], Z);
        for (int z = z0; z < z1; ++z)
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    const int64_t i = x + (int64_t)X * (y + (int64_t)Y * z);
                    if (!g.mask || g.mask[i]) r.tile_sites.push_back((uint32_t)(i + 1));
                }
        maxsites = std::max(maxsites, (uint32_t)r.tile_sites.size() - r.tile_offsets[(size_t)t]);
    }
    r.tile_offsets[(size_t)ntiles] = (uint32_t)r.tile_sites.size();
    if (maxsites == 0) fail(CPM_E_INVALID, "a tiled domain requires a mutable site");
    // the RNG operation field holds the 1-based proposal index in 12 bits
    if (maxsites > 4095) fail(CPM_E_UNSUPPORTED, "tiles of more than 4095 mutable sites");
    r.tiled_interval = d.switching_interval > 0 && (uint32_t)d.switching_interval < maxsites
                           ? d.switching_interval : (int)maxsites;
    if (r.tiled_interval > 64) fail(CPM_E_UNSUPPORTED, "switching intervals above 64 proposals per tile");
    r.tiled_batches = (int)((maxsites + (uint32_t)r.tiled_interval - 1) / (uint32_t)r.tiled_interval);
    // staged kernel: box and halo of at most 64 positions, no site staged twice (a periodic axis
    // at least as long as the staged extent), LDS within 64 KiB per workgroup
    {
        uint32_t positions = 1;
        bool ok = g_tiled_staging.load() != 0;
        for (int a = 0; a < 3; ++a) {
            const int ext = a < nd ? r.tile_edge[a] + 2 * halo : 1;
            r.tiled_stage[a] = ext;
            positions *= (uint32_t)ext;
            if (a < nd && periodic[a] && ext > g.dims[a]) ok = false;
        }
        const uint64_t lds = 3584 + (uint64_t)(positions + 2 * kTiledListLds) * 64 * 4;
        ok = ok && positions <= 64 && lds <= 65536;
        r.tiled_staged = ok ? positions : 0;
        if (!ok) r.tiled_stage[0] = r.tiled_stage[1] = r.tiled_stage[2] = 0;
    }
    if ((int64_t)maxc * r.tiled_batches > 255) fail(CPM_E_UNSUPPORTED, "tile colours times batches exceed 255 subrounds");
    // realized colour of every site: the colour of its tile (0 for immutable sites)
    r.colors.assign((size_t)r.n, 0);
    for (uint64_t t = 0; t < ntiles; ++t)
        for (uint32_t k = r.tile_offsets[(size_t)t]; k < r.tile_offsets[(size_t)t + 1]; ++k)
            r.colors[r.tile_sites[k] - 1] = (uint8_t)r.tile_colors[(size_t)t];
}

void compile_replica(Replica &r) {
    const cpm_model_desc &d = r.desc;
    if (d.abi_version != CPM_ABI_VERSION) fail(CPM_E_INVALID, "model descriptor ABI version mismatch");
    if (d.algorithm != CPM_ALG_CHECKERBOARD && d.algorithm != CPM_ALG_LOTTERY && d.algorithm != CPM_ALG_TILED)
        fail(CPM_E_UNSUPPORTED, "only CheckerboardSweepCPM, LotteryCPM and TiledCheckerboardCPM run on this engine");
    if (d.ndim != 2 && d.ndim != 3) fail(CPM_E_INVALID, "Cartesian domains are two- or three-dimensional");
    r.desc.dims[2] = d.ndim == 2 ? 1 : d.dims[2];
    uint64_t n = 1;
    for (int a = 0; a < 3; ++a) {
        if (r.desc.dims[a] <= 0) fail(CPM_E_INVALID, "Cartesian dimensions must be positive");
        n *= (uint64_t)r.desc.dims[a];
    }
    if (n > 0xffffffffull - 1) fail(CPM_E_INVALID, "site count must fit the UInt32 site representation");
    r.n = (int64_t)n;
    for (int a = 0; a < d.ndim; ++a) {
        const int k0 = d.face_kind[a][0], k1 = d.face_kind[a][1];
        if (k0 < 0 || k0 > 2 || k1 < 0 || k1 > 2) fail(CPM_E_INVALID, "unknown boundary kind");
        if ((k0 == CPM_FACE_PERIODIC) != (k1 == CPM_FACE_PERIODIC))
            fail(CPM_E_INVALID, "periodic boundary faces must occur as a paired axis");
    }
    validate_relation(d.proposal, d.ndim, "proposal");
    validate_relation(d.surface, d.ndim, "surface");
    const bool contact = has_term(d, CPM_TERM_CONTACT);
    if (contact) validate_relation(d.contact, d.ndim, "contact");
    if (d.n_energies < 0 || d.n_energies > 3) fail(CPM_E_INVALID, "at most three energy terms");
    for (int t = 0; t < d.n_energies; ++t)
        if (d.energy_order[t] < 0 || d.energy_order[t] > 2) fail(CPM_E_INVALID, "unknown energy term");
    if (contact && (d.n_types < 1 || d.n_types > CPM_MAX_TYPES))
        fail(CPM_E_INVALID, "contact matrix size must be 1..16");
    if (d.n_table < 0 || d.n_table > CPM_MAX_MEDIA || d.n_media < 0 || d.n_media > CPM_MAX_MEDIA)
        fail(CPM_E_INVALID, "at most 16 media");
    if (!(d.temperature >= 0.0f) || !std::isfinite(d.temperature))
        fail(CPM_E_INVALID, "temperature must be finite and non-negative");
    if (d.capacity < 1 || d.capacity > 0x7ffffffeu) fail(CPM_E_INVALID, "capacity must be 1..2^31-2");
    if (d.capacity > cpm::kIdMask)
        fail(CPM_E_UNSUPPORTED, "cell capacity beyond 2^26 - 1 does not fit the owner word (id and type bits)");
    if (d.algorithm == CPM_ALG_TILED) {
        // tiled_checkerboard.jl:27-60 and the tiled admission of init_scientific
        for (int a = 0; a < 3; ++a)
            if (d.tile_size[a] < 0 || (a < d.ndim && d.tile_size[a] > 0xffff) || (a >= d.ndim && d.tile_size[a] > 1))
                fail(CPM_E_INVALID, "tile size needs one positive edge per lattice dimension");
        if (d.switching_interval < 0) fail(CPM_E_INVALID, "switching interval must be positive");
        if (d.n_mechanics > 0) fail(CPM_E_UNSUPPORTED, "TiledCheckerboardCPM does not carry mechanics");
    }
    // mechanics (scientific_components.jl:162-182, 592-616; component checks of init_scientific)
    if (d.n_mechanics < 0 || d.n_mechanics > CPM_MAX_MECHANICS) fail(CPM_E_INVALID, "at most two mechanical components");
    for (int k = 0; k < d.n_mechanics; ++k) {
        const cpm_mechanics_desc &m = d.mechanics[k];
        if (m.kind != CPM_MECH_PRESSURE && m.kind != CPM_MECH_TENSION) fail(CPM_E_INVALID, "unknown mechanical component");
        if (m.initialization < CPM_MECH_INIT_CONSTITUTIVE || m.initialization > CPM_MECH_INIT_PRESERVE)
            fail(CPM_E_INVALID, "unknown mechanical initialization");
        if (m.instance_id < 1 || m.instance_id > 0xffffu)
            fail(CPM_E_INVALID, "mechanical component instance ID must be positive and fit UInt16");
        if (!(m.eta > 0.0f) || !std::isfinite(m.eta))
            fail(CPM_E_INVALID, "mechanical relaxation rate must be finite and positive");
        if (!m.noise_is_temperature && (!(m.noise >= 0.0f) || !std::isfinite(m.noise)))
            fail(CPM_E_INVALID, "mechanical noise must be finite and non-negative");
        for (int j = 0; j < k; ++j) {
            if (d.mechanics[j].instance_id == m.instance_id)
                fail(CPM_E_INVALID, "mechanical components need distinct instance IDs");
            if (d.mechanics[j].kind == m.kind)
                fail(CPM_E_INVALID, "mechanical components need distinct state properties");
        }
    }

    if (contact) {
        // every medium a contact relation can see needs a declared type (scientific_components.jl:35-40)
        auto typed = [&](uint32_t medium) {
            for (int i = 0; i < d.n_table; ++i)
                if (d.table_owner[i] == medium) return d.table_type[i] >= 1 && (int)d.table_type[i] <= d.n_types;
            return false;
        };
        for (int i = 0; i < d.n_media; ++i)
            if (!typed(d.medium_ids[i])) fail(CPM_E_INVALID, "medium owner has no declared type");
        for (int a = 0; a < d.ndim; ++a)
            for (int f = 0; f < 2; ++f)
                if (d.face_kind[a][f] == CPM_FACE_FIXED && !typed(d.face_medium[a][f]))
                    fail(CPM_E_INVALID, "fixed-exterior owner has no declared type");
        if (d.mutable_mask)
            for (int64_t i = 0; i < r.n; ++i)
                if (!d.mutable_mask[i] && !typed(d.immutable_ids[i]))
                    fail(CPM_E_INVALID, "obstacle owner has no declared type");
    }
    r.halo = std::max(max_abs_offset(d.proposal, d.ndim), max_abs_offset(d.surface, d.ndim));
    if (contact) r.halo = std::max(r.halo, max_abs_offset(d.contact, d.ndim));
    if (r.halo > 2) fail(CPM_E_UNSUPPORTED, "relations reaching beyond distance 2 are not supported");
    r.halo = std::max(r.halo, 1);

    r.has_mask = false;
    if (d.mutable_mask) {
        r.mask.assign(d.mutable_mask, d.mutable_mask + r.n);
        r.imm_tags.assign(d.immutable_tags, d.immutable_tags + r.n);
        r.imm_ids.assign(d.immutable_ids, d.immutable_ids + r.n);
        for (int64_t i = 0; i < r.n; ++i)
            if (!r.mask[i]) r.has_mask = true;
        if (!r.has_mask) {
            r.mask.clear();
            r.imm_tags.clear();
            r.imm_ids.clear();
        }
    }
    r.desc.mutable_mask = nullptr;
    r.desc.immutable_tags = nullptr;
    r.desc.immutable_ids = nullptr;
    r.nmut = 0;
    for (int64_t i = 0; i < r.n; ++i) r.nmut += r.has_mask ? (r.mask[i] != 0) : 1;
    if (r.nmut < 1) fail(CPM_E_INVALID, "a Cartesian ownership domain requires a mutable site");

    bool all_periodic = true;
    for (int a = 0; a < d.ndim; ++a)
        all_periodic = all_periodic && d.face_kind[a][0] == CPM_FACE_PERIODIC;
    bool big = true;
    for (int a = 0; a < d.ndim; ++a) big = big && r.desc.dims[a] >= r.halo;
    const bool simple = all_periodic && !r.has_mask && big;

    // conflict relations: proposal, boundary tracker, contact (checkerboard.jl:150-157)
    r.conf_off.clear();
    auto add_rel = [&](const cpm_relation_desc &rel) {
        for (int k = 0; k < rel.count; ++k) {
            bool dup = false;
            for (size_t q = 0; q < r.conf_off.size(); q += 3)
                dup |= r.conf_off[q] == rel.offsets[k][0] && r.conf_off[q + 1] == rel.offsets[k][1] &&
                       r.conf_off[q + 2] == rel.offsets[k][2];
            if (!dup)
                for (int a = 0; a < 3; ++a) r.conf_off.push_back(rel.offsets[k][a]);
        }
    };
    add_rel(d.proposal);
    add_rel(d.surface);
    if (contact) add_rel(d.contact);
    if (r.conf_off.size() / 3 > CPM_MAX_DIRS) fail(CPM_E_UNSUPPORTED, "conflict relation too large");

    HostGeom g{};
    g.ndim = d.ndim;
    for (int a = 0; a < 3; ++a) {
        g.dims[a] = r.desc.dims[a];
        g.face_kind[a][0] = d.face_kind[a][0];
        g.face_kind[a][1] = d.face_kind[a][1];
        g.face_medium[a][0] = d.face_medium[a][0];
        g.face_medium[a][1] = d.face_medium[a][1];
    }
    g.mask = r.has_mask ? r.mask.data() : nullptr;
    g.simple = simple;

    std::vector<int64_t> nb(r.conf_off.size() / 3 + 1);
    int pattern[3] = {0, 0, 0};
    std::vector<uint8_t> pattern_table;
    // Sites at least `halo` from every face of an obstacle-free lattice reach their conflict
    // neighbours by linear deltas; distinct offsets within the halo are distinct sites there
    // (dims >= 2 halo + 1), so no de-duplication is needed. Other sites use realize_site.
    const int kconf = (int)r.conf_off.size() / 3;
    std::vector<int64_t> lin((size_t)kconf);
    for (int q = 0; q < kconf; ++q)
        lin[q] = r.conf_off[3 * q] + (int64_t)g.dims[0] * (r.conf_off[3 * q + 1] + (int64_t)g.dims[1] * r.conf_off[3 * q + 2]);
    const int hh = r.halo;
    bool wide = !g.mask;
    for (int a = 0; a < d.ndim; ++a) wide = wide && g.dims[a] >= 2 * hh + 1;
    // Layers along the slowest axis: rows in 2D, planes in 3D.
    const int layer_axis = d.ndim - 1;
    const int n_layers = g.dims[layer_axis];
    const int64_t layer_size = r.n / n_layers;
    // visits the sites of layers [l0, l1) in ascending order with their coordinates;
    // fn(i, x, y, z, interior)
    auto for_layers = [&](int l0, int l1, auto &&fn) {
        const int z0 = layer_axis == 2 ? l0 : 0, z1 = layer_axis == 2 ? l1 : g.dims[2];
        const int y0 = layer_axis == 1 ? l0 : 0, y1 = layer_axis == 1 ? l1 : g.dims[1];
        for (int z = z0; z < z1; ++z)
            for (int y = y0; y < y1; ++y) {
                const bool iny = wide && y >= hh && y < g.dims[1] - hh &&
                                 (d.ndim == 2 || (z >= hh && z < g.dims[2] - hh));
                int64_t i = (int64_t)g.dims[0] * (y + (int64_t)g.dims[1] * z);
                for (int x = 0; x < g.dims[0]; ++x, ++i)
                    fn(i, x, y, z, iny && x >= hh && x < g.dims[0] - hh);
            }
    };
    auto for_sites = [&](auto &&fn) { for_layers(0, n_layers, fn); };
    if (d.algorithm == CPM_ALG_CHECKERBOARD) {
        // checkerboard.jl:57-90: greedy smallest free colour, ascending site order
        r.colors.assign((size_t)r.n, 0);
        int maxcolor = 0;
        std::vector<char> forbidden;
        uint8_t *colors = r.colors.data();
        // the smallest colour no conflict neighbour holds; neighbours of higher index are
        // still uncoloured (0) when a site is visited in ascending order
        auto colour_site = [&](int64_t i, int, int, int, bool interior) {
            if (g.mask && !g.mask[i]) return;
            int col;
            if (kconf + 1 <= 63) {
                // colours never exceed degree + 1 <= 63: a bit mask of the taken colours
                uint64_t taken = 1;  // colour 0 (uncoloured) is never chosen
                if (interior) {
                    for (int q = 0; q < kconf; ++q) taken |= 1ull << colors[i + lin[q]];
                } else {
                    const int c = conflict_neighbours(g, r.conf_off, i, nb.data());
                    for (int q = 0; q < c; ++q) taken |= 1ull << colors[nb[q]];
                }
                col = __builtin_ctzll(~taken);
            } else {
                const int c = conflict_neighbours(g, r.conf_off, i, nb.data());
                forbidden.assign((size_t)c + 2, 0);
                for (int q = 0; q < c; ++q) {
                    const int cq = colors[nb[q]];
                    if (cq > 0 && cq <= c + 1) forbidden[cq] = 1;
                }
                col = 1;
                while (forbidden[col]) ++col;
            }
            if (col > 255) fail(CPM_E_UNSUPPORTED, "realized colouring exceeds 255 colours");
            colors[i] = (uint8_t)col;
            maxcolor = std::max(maxcolor, col);
        };
        // Every layer at least `hh` from both ends of the layer axis is coloured by one
        // deterministic function of the `hh` layers before it and of its own obstacles (its
        // conflict neighbours of lower index lie in those layers or earlier in its own layer).
        // When the state (the last `hh` layers) repeats exactly, each following obstacle-free
        // layer equals the layer one period before it if that one is obstacle-free too; such
        // layers are copied, the others coloured site by site. The last `hh` layers see layer
        // 0 onward across a periodic face and are always coloured site by site. The result
        // equals the site-by-site greedy colouring. source[m] is the layer m repeats (or -1).
        std::vector<int> source((size_t)n_layers, -1);
        std::vector<char> layer_free((size_t)n_layers, 1);
        if (g.mask)
            for (int l = 0; l < n_layers; ++l)
                for (int64_t k = 0; k < layer_size && layer_free[(size_t)l]; ++k)
                    layer_free[(size_t)l] = g.mask[(int64_t)l * layer_size + k] != 0;
        {
            std::unordered_map<uint64_t, int> seen;
            const size_t state_bytes = (size_t)hh * (size_t)layer_size;
            for (int l = 0; l < n_layers; ++l) {
                if (source[(size_t)l] < 0) for_layers(l, l + 1, colour_site);
                if (l + 1 < hh || l + 1 >= n_layers - hh) continue;
                const uint8_t *state = colors + (int64_t)(l + 1 - hh) * layer_size;
                uint64_t hsh = 0x9e3779b97f4a7c15ull;
                for (size_t k = 0; k < state_bytes; k += 8) {
                    uint64_t w = 0;
                    std::memcpy(&w, state + k, std::min<size_t>(8, state_bytes - k));
                    hsh = mix64(hsh ^ w);
                }
                auto it = seen.find(hsh);
                if (it == seen.end()) {
                    seen.emplace(hsh, l);
                    continue;
                }
                if (std::memcmp(colors + (int64_t)(it->second + 1 - hh) * layer_size, state, state_bytes) != 0) continue;
                const int period = l - it->second;
                int m = l + 1;
                for (; m < n_layers - hh && layer_free[(size_t)m] && layer_free[(size_t)(m - period)]; ++m) {
                    std::memcpy(colors + (int64_t)m * layer_size, colors + (int64_t)(m - period) * layer_size,
                                (size_t)layer_size);
                    source[(size_t)m] = m - period;
                }
                l = m - 1;  // continues with the first layer not copied
            }
        }
        r.rounds = maxcolor;
        // periodic pattern detection: colour == f(x mod px, y mod py, z mod pz). Each
        // candidate is first probed on the sites within 2 of a face (where wrap seams break
        // periodicity), then verified on every site. A repeated layer matches when the layer
        // it repeats does and the pattern's period along the layer axis divides the
        // repetition period, so the verification skips such layers.
        auto matches = [&](int px, int py, int pz, std::vector<uint8_t> &table, bool probe) {
            const int pl = layer_axis == 1 ? py : pz;
            auto repeated = [&](int l) {
                const int src = source[(size_t)l];
                return !probe && src >= 0 && (l - src) % pl == 0;
            };
            int rz = 0;
            for (int z = 0; z < g.dims[2]; ++z, rz = (rz + 1 == pz ? 0 : rz + 1)) {
                const bool ez = d.ndim == 3 && (z < 2 || z >= g.dims[2] - 2);
                if (layer_axis == 2 && repeated(z)) continue;
                int ry = 0;
                for (int y = 0; y < g.dims[1]; ++y, ry = (ry + 1 == py ? 0 : ry + 1)) {
                    if (layer_axis == 1 && repeated(y)) continue;
                    const bool ey = ez || y < 2 || y >= g.dims[1] - 2;
                    const int64_t row = (int64_t)g.dims[0] * (y + (int64_t)g.dims[1] * z);
                    const int base = px * (ry + py * rz);
                    int rx = 0;
                    for (int x = 0; x < g.dims[0]; ++x, rx = (rx + 1 == px ? 0 : rx + 1)) {
                        if (probe && !ey && x >= 2 && x < g.dims[0] - 2) {
                            // interior of the row: skip to its last two sites
                            const int skip = g.dims[0] - 2 - x;
                            x += skip - 1;
                            rx = (rx + skip - 1) % px;
                            continue;
                        }
                        const int64_t i = row + x;
                        if (g.mask && !g.mask[i]) continue;
                        uint8_t &t = table[(size_t)(base + rx)];
                        if (t == 0) t = colors[i];
                        else if (t != colors[i]) return false;
                    }
                }
            }
            return true;
        };
        bool found = false;
        for (int prod = 1; prod <= 64 && !found; ++prod) {
            for (int px = 1; px <= 8 && !found; ++px)
                for (int py = 1; py <= 8 && !found; ++py)
                    for (int pz = 1; pz <= (d.ndim == 3 ? 8 : 1) && !found; ++pz) {
                        if (px * py * pz != prod) continue;
                        std::vector<uint8_t> table((size_t)prod, 0);
                        if (!matches(px, py, pz, table, true)) continue;
                        std::fill(table.begin(), table.end(), 0);
                        if (matches(px, py, pz, table, false)) {
                            found = true;
                            pattern[0] = px;
                            pattern[1] = py;
                            pattern[2] = pz;
                            pattern_table = table;
                        }
                    }
        }
        r.per_site_color = !found;
    } else if (d.algorithm == CPM_ALG_TILED) {
        compile_tiled(r, g);
    } else {
        // lottery.jl:46-92: rounds = maximum realized conflict degree + 1
        int maxdeg = 0;
        if (all_periodic && wide) {
            // distinct offsets within the halo reach distinct sites from every site
            maxdeg = kconf;
            r.uniform_degree = true;
            r.degrees.clear();
        } else {
            r.degrees.assign((size_t)r.n, 0);
            for_sites([&](int64_t i, int, int, int, bool interior) {
                if (g.mask && !g.mask[i]) return;
                const int c = interior ? kconf : conflict_neighbours(g, r.conf_off, i, nb.data());
                r.degrees[i] = (uint8_t)std::min(c, 255);
                maxdeg = std::max(maxdeg, c);
            });
            r.uniform_degree = true;
            for (int64_t i = 0; i < r.n; ++i) {
                if (g.mask && !g.mask[i]) continue;
                if (r.degrees[i] != maxdeg) {
                    r.uniform_degree = false;
                    break;
                }
            }
        }
        if (maxdeg + 1 > 255) fail(CPM_E_UNSUPPORTED, "lottery conflict degree exceeds the subround domain");
        r.rounds = maxdeg + 1;
    }

    // tile shape by dimension and halo (process-wide preference, else the default); bit 4
    // selects the kernels for proposal relations other than 2 * ndim directions (the
    // fixed-count kernels keep the offsets in registers)
    const int shape = tile_shape_for(d.algorithm, d.ndim, r.halo);
    r.config = d.algorithm == CPM_ALG_TILED ? kTiledConfig : shape | (d.proposal.count == 2 * d.ndim ? 0 : 16);
    int cdim, chalo, tx, ty, tz;
    cpm_internal_tile_config(shape, &cdim, &chalo, &tx, &ty, &tz);

    cpm::RepDev &v = r.dev;
    std::memset(&v, 0, sizeof(v));
    v.ndim = d.ndim;
    for (int a = 0; a < 3; ++a) v.dims[a] = r.desc.dims[a];
    v.nsites = (uint32_t)r.n;
    for (int a = 0; a < 3; ++a)
        for (int f = 0; f < 2; ++f) {
            v.face_kind[a][f] = a < d.ndim ? d.face_kind[a][f] : CPM_FACE_PERIODIC;
            v.face_word[a][f] = -(int32_t)d.face_medium[a][f];
        }
    v.simple = simple ? 1 : 0;
    v.ntiles[0] = (v.dims[0] + tx - 1) / tx;
    v.ntiles[1] = (v.dims[1] + ty - 1) / ty;
    v.ntiles[2] = (v.dims[2] + tz - 1) / tz;
    v.algorithm = d.algorithm;
    v.rounds = r.rounds;
    v.uniform_degree = r.uniform_degree ? 1 : 0;
    v.temperature = d.temperature;
    const uint64_t key = d.seed ^ mix64(0ull ^ kGenerationDomain);
    v.key_lo = (uint32_t)key;
    v.key_hi = (uint32_t)(key >> 32);
    if (d.algorithm == CPM_ALG_CHECKERBOARD && !r.per_site_color) {
        for (int a = 0; a < 3; ++a) v.pat[a] = pattern[a];
        std::memcpy(v.pattern, pattern_table.data(), pattern_table.size());
    } else {
        v.pat[0] = v.pat[1] = v.pat[2] = 1;
    }
    for (int i = 0; i < v.pat[0] * v.pat[1] * v.pat[2]; ++i) {
        v.pat_coord[i][0] = (uint8_t)(i % v.pat[0]);
        v.pat_coord[i][1] = (uint8_t)((i / v.pat[0]) % v.pat[1]);
        v.pat_coord[i][2] = (uint8_t)(i / (v.pat[0] * v.pat[1]));
    }
    v.pat_aligned = (tx % v.pat[0] == 0 && ty % v.pat[1] == 0 && tz % v.pat[2] == 0) ? 1 : 0;
    v.pat_fast = 0;
    if (d.algorithm == CPM_ALG_CHECKERBOARD && !r.per_site_color && v.pat_aligned) {
        const int nx = tx / v.pat[0], ny = ty / v.pat[1], nz = tz / v.pat[2];
        auto pow2 = [](int q) { return q > 0 && (q & (q - 1)) == 0; };
        auto log2i = [](int q) {
            int l = 0;
            while ((1 << l) < q) ++l;
            return l;
        };
        if (pow2(nx) && pow2(ny) && pow2(nz) && r.rounds <= 64) {
            v.pat_fast = 1;
            v.cls_sx = log2i(nx);
            v.cls_sy = log2i(ny);
            v.cls_shift = log2i(nx * ny * nz);
            const int np = v.pat[0] * v.pat[1] * v.pat[2];
            int at = 0;
            for (int c = 1; c <= r.rounds; ++c) {
                v.col_first[c] = (uint8_t)at;
                for (int q = 0; q < np; ++q)
                    if (v.pattern[q] == c) v.col_res[at++] = (uint8_t)q;
                v.col_count[c] = (uint8_t)(at - v.col_first[c]);
            }
        }
    }
    // row kernel: 2D, periodic and obstacle-free, canonical von Neumann proposal, one
    // residue class of a 2 x 2 pattern per colour, even x side (rows of whole 16-byte groups,
    // or of 8-byte halves when the side is 2 mod 4)
    if (g_row_kernels.load() && d.algorithm == CPM_ALG_CHECKERBOARD && simple && d.ndim == 2 && v.pat_fast &&
        (r.config & 16) == 0 && v.pat[0] == 2 && v.pat[1] == 2 && v.pat[2] == 1 && r.rounds == 4 &&
        r.desc.dims[0] % 2 == 0 &&
        (shape == 0 || shape == 1 || (shape >= 4 && shape <= 7) || shape == 12 || shape == 13)) {
        static const int vn[4][2] = {{-1, 0}, {0, -1}, {0, 1}, {1, 0}};
        bool ok = d.proposal.count == 4;
        for (int k = 0; ok && k < 4; ++k)
            ok = d.proposal.offsets[k][0] == vn[k][0] && d.proposal.offsets[k][1] == vn[k][1];
        for (int c = 1; ok && c <= 4; ++c) ok = v.col_count[c] == 1;
        if (ok) r.config |= r.desc.dims[0] % 4 == 0 ? 32 : 32 | 64;
    }
    // 3D row kernel: periodic and obstacle-free, canonical von Neumann proposal (six
    // directions), one residue class of a 2 x 2 x 2 pattern per colour, rows of whole 16-byte
    // groups
    if (g_row_kernels.load() && d.algorithm == CPM_ALG_CHECKERBOARD && simple && d.ndim == 3 && v.pat_fast &&
        (r.config & 16) == 0 && v.pat[0] == 2 && v.pat[1] == 2 && v.pat[2] == 2 && r.rounds == 8 &&
        r.desc.dims[0] % 4 == 0 &&
        (shape == 2 || shape == 8 || shape == 9)) {
        static const int vn3[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};
        bool ok = d.proposal.count == 6;
        for (int k = 0; ok && k < 6; ++k)
            for (int a = 0; a < 3; ++a) ok = ok && d.proposal.offsets[k][a] == vn3[k][a];
        for (int c = 1; ok && c <= 8; ++c) ok = v.col_count[c] == 1;
        if (ok) r.config |= 32;
    }
    // lottery on a periodic obstacle-free lattice with at least 2 halo + 1 sites per axis:
    // winners follow from the tickets alone
    if (d.algorithm == CPM_ALG_LOTTERY && all_periodic && wide) r.config |= 64;
    // plain lottery whose conflict shell is the whole radius-1 Moore shell (8 positions in 2D,
    // 26 in 3D): the winner test reads the tickets at compile-time offsets
    if (d.algorithm == CPM_ALG_LOTTERY && (r.config & 64) && (r.config & 16) == 0) {
        const int want = d.ndim == 2 ? 8 : 26;
        bool moore = (int)(r.conf_off.size() / 3) == want;
        for (size_t q = 0; moore && q < r.conf_off.size(); q += 3) {
            const int64_t ox = r.conf_off[q], oy = r.conf_off[q + 1], oz = r.conf_off[q + 2];
            moore = ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1 &&
                    (d.ndim == 3 || oz == 0) && (ox | oy | oz) != 0;
        }
        if (moore) r.config |= 128;  // offsets are distinct, so the count fixes the set
    }
    v.capacity = d.capacity;
    v.tiled_interval = r.tiled_interval;
    v.tiled_batches = r.tiled_batches;
    for (int a = 0; a < 3; ++a) {
        v.tile_edge[a] = r.tile_edge[a];
        v.tile_grid[a] = r.tile_grid[a];
        v.tiled_stage[a] = r.tiled_stage[a];
    }
    v.tiled_halo = r.tiled_halo;
    v.n_mech = d.n_mechanics;
    for (int k = 0; k < d.n_mechanics; ++k) {
        const cpm_mechanics_desc &m = d.mechanics[k];
        v.mech_kind[k] = m.kind;
        v.mech_init[k] = m.initialization;
        v.mech_instance[k] = m.instance_id;
        v.mech_eta[k] = m.eta;
        v.mech_noise[k] = m.noise_is_temperature ? d.temperature : m.noise;
    }
    if (d.n_mechanics > 0) {
        // OU half intervals in Float32: colour size / (2 * mutable sites) per colour
        // (scientific_mechanics.jl:170-182), or 1 / (2 * rounds) (lottery.jl:353)
        if (d.algorithm == CPM_ALG_LOTTERY) {
            v.mech_half[0] = 1.0f / (2.0f * (float)r.rounds);
        } else {
            if (r.rounds > cpm::kMaxRounds) fail(CPM_E_UNSUPPORTED, "too many colours for the mechanical clock");
            std::vector<uint64_t> size((size_t)r.rounds + 1, 0);
            uint64_t mutable_sites = 0;
            for (int64_t i = 0; i < r.n; ++i)
                if (r.colors[(size_t)i] > 0) {
                    size[r.colors[(size_t)i]]++;
                    mutable_sites++;
                }
            for (int c = 1; c <= r.rounds; ++c)
                v.mech_half[c] = (float)size[c] / (2.0f * (float)mutable_sites);
        }
    }
    v.halo = r.halo;
    {
        const int64_t X = r.desc.dims[0], XY = (int64_t)r.desc.dims[0] * r.desc.dims[1];
        for (int k = 0; k < 32 && k < d.surface.count; ++k)
            v.surf_lin[k] = (int32_t)(d.surface.offsets[k][0] + X * d.surface.offsets[k][1] + XY * d.surface.offsets[k][2]);
        if (contact)
            for (int k = 0; k < 32 && k < d.contact.count; ++k)
                v.cont_lin[k] = (int32_t)(d.contact.offsets[k][0] + X * d.contact.offsets[k][1] + XY * d.contact.offsets[k][2]);
    }
    r.clist.clear();
    r.coff.clear();
    if (d.algorithm == CPM_ALG_CHECKERBOARD && !v.pat_fast) {
        const int C = r.rounds;
        const uint32_t ntx = (uint32_t)v.ntiles[0], nty = (uint32_t)v.ntiles[1], ntz = (uint32_t)v.ntiles[2];
        const uint64_t ntiles = (uint64_t)ntx * nty * ntz;
        r.clist.resize((size_t)r.n);
        r.coff.assign((size_t)ntiles * (size_t)(C + 1), 0);
        std::vector<uint32_t> cnt((size_t)C + 1);
        uint32_t at = 0;
        for (uint64_t t = 0; t < ntiles; ++t) {
            const int tx0 = (int)(t % ntx) * tx, ty0 = (int)((t / ntx) % nty) * ty, tz0 = (int)(t / ((uint64_t)ntx * nty)) * tz;
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (int pass = 0; pass < 2; ++pass) {
                // pass 0 counts the sites of each colour, pass 1 writes them in position order
                uint32_t *off = r.coff.data() + t * (uint64_t)(C + 1);
                if (pass == 1) {
                    uint32_t run = at;
                    for (int c = 1; c <= C; ++c) {
                        off[c - 1] = run;
                        run += cnt[c];
                        cnt[c] = off[c - 1];
                    }
                    off[C] = run;
                }
                for (int iz = 0; iz < tz; ++iz)
                    for (int iy = 0; iy < ty; ++iy)
                        for (int ix = 0; ix < tx; ++ix) {
                            const int gx = tx0 + ix, gy = ty0 + iy, gz = tz0 + iz;
                            if (gx >= v.dims[0] || gy >= v.dims[1] || gz >= v.dims[2]) continue;
                            const int64_t site = gx + (int64_t)v.dims[0] * (gy + (int64_t)v.dims[1] * gz);
                            const int c = r.colors[(size_t)site];
                            if (c == 0) continue;  // immutable site
                            if (pass == 0) ++cnt[c];
                            else r.clist[cnt[c]++] = (uint16_t)(ix + tx * (iy + ty * iz));
                        }
                if (pass == 1) at = off[C];
            }
        }
        r.clist.resize(at);
    }
    {
        // LDS layout of the tile shape: rows of tx + 8 positions (4-position x pads), ty + 2 halo rows
        const int lx = tx + 8, ly = ty + 2 * chalo;
        for (int k = 0; k < 8 && k < d.proposal.count; ++k)
            v.prop_lds[k] = d.proposal.offsets[k][0] +
                            lx * (d.proposal.offsets[k][1] + ly * d.proposal.offsets[k][2]);
    }
    v.prop_bound = make_bound((uint32_t)d.proposal.count);
    v.round_bound = make_bound((uint32_t)std::max(r.rounds, 1));
    v.div_x = make_div((uint32_t)v.dims[0]);
    v.div_xy = make_div((uint32_t)v.dims[0] * (uint32_t)v.dims[1]);
    v.n_energies = d.n_energies;
    for (int t = 0; t < 3; ++t) v.energy_order[t] = d.energy_order[t];
    v.volume_f64 = d.volume_f64 ? 1 : 0;
    v.contact_f64 = d.contact_f64 ? 1 : 0;
    v.boundary_f64 = d.boundary_f64 ? 1 : 0;
    v.has_contact = contact ? 1 : 0;
    v.fast_eval = 0;
    {
        const int ns = d.surface.count, nc = contact ? d.contact.count : 0;
        const bool f32 = !(has_term(d, CPM_TERM_VOLUME) && d.volume_f64) && !(contact && d.contact_f64) &&
                         !(has_term(d, CPM_TERM_BOUNDARY) && d.boundary_f64);
        // relations that are the canonical first shell / Moore shell in canonical order
        auto canonical = [&](const cpm_relation_desc &rel) {
            for (int k = 0; k < rel.count; ++k)
                for (int a = 0; a < 3; ++a) {
                    const int want = d.ndim == 2 ? cpm::moore_offset<2>(k, a) : cpm::moore_offset<3>(k, a);
                    if (rel.offsets[k][a] != want) return false;
                }
            return true;
        };
        const int shell = 2 * d.ndim, moore = d.ndim == 2 ? 8 : 26;
        const bool rels = ns == shell && canonical(d.surface) &&
                          (nc == 0 || ((nc == shell || nc == moore) && canonical(d.contact)));
        if (v.simple && d.n_mechanics == 0 && f32 && rels) {
            if (d.ndim == 2 && ns == 4) v.fast_eval = nc == 0 ? 1 : nc == 4 ? 2 : nc == 8 ? 3 : 0;
            if (d.ndim == 3 && ns == 6) v.fast_eval = nc == 0 ? 4 : nc == 6 ? 5 : nc == 26 ? 6 : 0;
        }
    }
    v.n_types = d.n_types;
    for (int i = 0; i < CPM_MAX_TYPES * CPM_MAX_TYPES; ++i) {
        v.Jd[i] = d.J[i];
        v.Jf[i] = (float)d.J[i];
    }
    v.n_table = d.n_table;
    for (int i = 0; i < d.n_table; ++i) {
        v.table_word[i] = -(int32_t)d.table_owner[i];
        v.table_type[i] = d.table_type[i];
    }
    v.n_media = d.n_media;
    for (int i = 0; i < d.n_media; ++i) v.medium_word[i] = -(int32_t)d.medium_ids[i];
    v.retire_tv = d.retire_target_volume;
    v.retire_vs = d.retire_volume_strength;
    v.retire_bs = d.retire_boundary_strength;
    v.retire_tb = d.retire_target_boundary;
    auto fill_rel = [](cpm::RelDev &o, const cpm_relation_desc &rel) {
        o.n = rel.count;
        for (int k = 0; k < rel.count; ++k) {
            for (int a = 0; a < 3; ++a) o.off[k][a] = (int8_t)rel.offsets[k][a];
            o.wd[k] = rel.weights[k];
            o.wf[k] = (float)rel.weights[k];
        }
    };
    fill_rel(v.prop, d.proposal);
    fill_rel(v.surf, d.surface);
    if (contact) fill_rel(v.cont, d.contact);
    v.conf.n = (int32_t)(r.conf_off.size() / 3);
    for (int k = 0; k < v.conf.n; ++k)
        for (int a = 0; a < 3; ++a) v.conf.off[k][a] = (int8_t)r.conf_off[3 * k + a];
}

// Device buffers of a replica, filled in stream order (no host synchronization): copy
// sources are host vectors the replica keeps for its lifetime.
void allocate_replica(Replica &r, hipStream_t stream) {
    const size_t cap1 = (size_t)r.desc.capacity + 1;
    r.d_lattice = dev_alloc<int32_t>((size_t)r.n, r.bytes);
    if (r.has_mask) {
        r.d_mask = dev_alloc<uint8_t>((size_t)r.n, r.bytes);
        hip_check(hipMemcpyAsync(r.d_mask, r.mask.data(), (size_t)r.n, hipMemcpyHostToDevice, stream), "copy mask");
    }
    if (!r.coff.empty()) {
        r.d_clist = dev_alloc<uint16_t>(std::max<size_t>(r.clist.size(), 1), r.bytes);
        r.d_coff = dev_alloc<uint32_t>(r.coff.size(), r.bytes);
        if (!r.clist.empty())
            hip_check(hipMemcpyAsync(r.d_clist, r.clist.data(), r.clist.size() * 2, hipMemcpyHostToDevice, stream), "copy colour lists");
        hip_check(hipMemcpyAsync(r.d_coff, r.coff.data(), r.coff.size() * 4, hipMemcpyHostToDevice, stream), "copy colour ranges");
    }
    if (r.desc.algorithm == CPM_ALG_LOTTERY && !r.uniform_degree) {
        r.d_degree = dev_alloc<uint8_t>((size_t)r.n, r.bytes);
        hip_check(hipMemcpyAsync(r.d_degree, r.degrees.data(), (size_t)r.n, hipMemcpyHostToDevice, stream), "copy degrees");
    }
    r.d_cells = dev_alloc<cpm::CellRec>(cap1, r.bytes);
    r.d_mvol = dev_alloc<int32_t>(CPM_MAX_MEDIA, r.bytes);
    if (r.desc.algorithm == CPM_ALG_TILED) {
        auto upload = [&](uint32_t *&dst, const std::vector<uint32_t> &src) {
            dst = dev_alloc<uint32_t>(std::max<size_t>(src.size(), 1), r.bytes);
            if (!src.empty())
                hip_check(hipMemcpyAsync(dst, src.data(), src.size() * 4, hipMemcpyHostToDevice, stream), "copy tiles");
        };
        upload(r.d_tile_sites, r.tile_sites);
        upload(r.d_tile_offsets, r.tile_offsets);
        upload(r.d_color_tiles, r.color_tiles);
        upload(r.d_color_offsets, r.color_offsets);
        r.d_tile_dvol = dev_alloc<int32_t>(cap1, r.bytes);
        r.d_tile_dbnd = dev_alloc<int32_t>(cap1, r.bytes);
        r.d_tiled_counts = dev_alloc<unsigned long long>(8, r.bytes);
        // touched-cell entries beyond the kernels' LDS list (at most two per proposal)
        const uint32_t entries = 2 * (uint32_t)r.tiled_interval;
        r.dev.tile_spill_stride = entries > kTiledListLds ? (int32_t)(2 * (entries - kTiledListLds)) : 0;
        r.d_tile_spill = dev_alloc<int32_t>(std::max<size_t>((size_t)r.tiled_max_tiles * r.dev.tile_spill_stride, 1),
                                            r.bytes);
        hip_check(hipMemsetAsync(r.d_tile_dvol, 0, cap1 * 4, stream), "memset");
        hip_check(hipMemsetAsync(r.d_tile_dbnd, 0, cap1 * 4, stream), "memset");
        hip_check(hipMemsetAsync(r.d_tiled_counts, 0, 8 * 8, stream), "memset");
    }
    hip_check(hipMemsetAsync(r.d_lattice, 0, (size_t)r.n * 4, stream), "memset");
    hip_check(hipMemsetAsync(r.d_cells, 0, cap1 * sizeof(cpm::CellRec), stream), "memset");
    hip_check(hipMemsetAsync(r.d_mvol, 0, CPM_MAX_MEDIA * 4, stream), "memset");
    cpm::RepDev &v = r.dev;
    v.lattice = r.d_lattice;
    v.mask = r.d_mask;
    v.colour_list = r.d_clist;
    v.colour_off = r.d_coff;
    v.site_degree = r.d_degree;
    v.cells = r.d_cells;
    v.medium_vol = r.d_mvol;
    v.tile_sites = r.d_tile_sites;
    v.tile_offsets = r.d_tile_offsets;
    v.color_tiles = r.d_color_tiles;
    v.color_offsets = r.d_color_offsets;
    v.tile_dvol = r.d_tile_dvol;
    v.tile_dbnd = r.d_tile_dbnd;
    v.tile_spill = r.d_tile_spill;
    v.tiled_counts = r.d_tiled_counts;
}

// Upper bound on the device bytes a replica adds to a batch: its own buffers plus its share
// of the group arenas (claim entries per tile site, tile table, accounting slots for at most
// (2 halo + 1)^ndim colour passes, two claim words per cell) and its descriptor slot.
uint64_t estimate_replica_bytes(const cpm_model_desc &d) {
    const int ndim = d.ndim == 3 ? 3 : 2;
    int dims[3] = {std::max(d.dims[0], 1), std::max(d.dims[1], 1), ndim == 3 ? std::max(d.dims[2], 1) : 1};
    uint64_t n = (uint64_t)dims[0] * (uint64_t)dims[1] * (uint64_t)dims[2];
    int halo = std::max(max_abs_offset(d.proposal, ndim), max_abs_offset(d.surface, ndim));
    if (has_term(d, CPM_TERM_CONTACT)) halo = std::max(halo, max_abs_offset(d.contact, ndim));
    halo = std::min(std::max(halo, 1), 2);
    int cdim, chalo, t[3];
    cpm_internal_tile_config(tile_shape_for(d.algorithm, ndim, halo), &cdim, &chalo, &t[0],
                             &t[1], &t[2]);
    uint64_t tiles = 1;
    for (int a = 0; a < 3; ++a) tiles *= (uint64_t)((dims[a] + t[a] - 1) / t[a]);
    const uint64_t tile_sites = (uint64_t)t[0] * t[1] * t[2];
    uint64_t rounds = 1;
    for (int a = 0; a < ndim; ++a) rounds *= (uint64_t)(2 * halo + 1);
    const uint64_t cap1 = (uint64_t)d.capacity + 1;
    const uint64_t own = n * 4 + (d.mutable_mask ? n : 0) + 2 * n /* colour lists or degree table */ +
                         tiles * (rounds + 1) * 4 /* colour ranges */ +
                         cap1 * sizeof(cpm::CellRec) + CPM_MAX_MEDIA * 4 +
                         (d.algorithm == CPM_ALG_TILED ? 3 * n * 4 + 16 + cap1 * 8 + 64 : 0) /* tile lists */;
    const uint64_t arenas = tiles * (tile_sites * sizeof(cpm::Entry) + 4 * cpm::kWaves + sizeof(cpm::TileInfo) +
                                     rounds * cpm::kSlotFields * 4) +
                            2 * cap1 * 8;
    return own + arenas + sizeof(cpm::RepDev) + sizeof(cpm::RepDyn);
}

// ---------------------------------------------------------------------------
// Groups: replicas sharing one kernel instantiation and one set of arenas
// ---------------------------------------------------------------------------
struct Group {
    int config = 0;
    bool tiled() const { return config == kTiledConfig; }
    int tx = 0, ty = 0, tz = 0;
    std::vector<int> slots;
    uint32_t ntiles = 0, total_claims = 0, seg_cap = 0;
    int max_rounds = 0;
    int fast_eval = 0;  // the commit evaluation every replica of the group takes (0: general)
    cpm::TileInfo *d_tiles = nullptr;
    unsigned long long *d_claims = nullptr;
    cpm::Entry *d_seg = nullptr;
    uint32_t *d_seg_count = nullptr;
    uint32_t *d_slots = nullptr;
    size_t cap_tiles = 0, cap_claims = 0, cap_slots = 0, cap_seg = 0;
    uint64_t bytes = 0;
    uint64_t pass_counter = 0;
    bool dirty = true;
    bool dirty_arena_fresh = true;  // no accounting written since the last rebuild

    void release() {
        dev_free(d_tiles); dev_free(d_claims); dev_free(d_seg); dev_free(d_seg_count);
        dev_free(d_slots);
        d_tiles = nullptr; d_claims = nullptr; d_seg = nullptr; d_seg_count = nullptr;
        d_slots = nullptr;
        cap_tiles = cap_claims = cap_slots = cap_seg = 0;
        bytes = 0;
    }
};

}  // namespace

// In-HBM snapshot of one replica at a completed MCS: its owner words, cell records (trackers,
// property columns, mechanical state) and medium volumes, copied on the device, with the MCS
// counter and seed. The snapshot owns its device buffers.
struct cpm_snapshot {
    int ndim = 0;
    int64_t dims[3] = {1, 1, 1};
    size_t n = 0;
    uint32_t capacity = 0;
    int32_t n_media = 0, n_types = 0;
    uint64_t mcs = 0, seed = 0;
    int32_t *d_lattice = nullptr;
    cpm::CellRec *d_cells = nullptr;
    int32_t *d_mvol = nullptr;
    uint64_t bytes = 0;
    ~cpm_snapshot() {
        if (d_lattice) (void)hipFree(d_lattice);
        if (d_cells) (void)hipFree(d_cells);
        if (d_mvol) (void)hipFree(d_mvol);
    }
};

namespace {

struct Engine {
    int algorithm = CPM_ALG_CHECKERBOARD;
    hipStream_t stream = nullptr;
    std::vector<Replica *> reps;
    cpm::RepDev *d_reps = nullptr;
    cpm::RepDyn *d_dyn = nullptr;
    cpm::McsPlan *d_plan = nullptr;  // [slot][kPlanMcs]
    int slot_cap = 0;
    std::vector<Group> groups;
    unsigned long long *d_report = nullptr;
    unsigned long long *h_report = nullptr;
    unsigned long long *d_many = nullptr;   // batched report counters (cache_reports)
    unsigned long long *h_many = nullptr;
    size_t many_cap = 0;
    uint64_t metrics[6] = {0, 0, 0, 0, 0, 0};
    uint64_t budget = 0;
    uint32_t graph_chunk = 0;
    uint32_t graph_mcs = 0;  // MCS per captured graph
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    bool graph_valid = false;
    std::string err;

    ~Engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (Replica *r : reps) delete r;
        for (Group &g : groups) g.release();
        dev_free(d_reps);
        dev_free(d_dyn);
        dev_free(d_plan);
        dev_free(d_report);
        if (h_report) (void)hipHostFree(h_report);
        dev_free(d_many);
        if (h_many) (void)hipHostFree(h_many);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_valid = false;
    }

    void init(int alg) {
        algorithm = alg;
        hip_check(hipStreamCreate(&stream), "hipStreamCreate");
        hip_check(hipMalloc(&d_report, 16 * sizeof(unsigned long long)), "hipMalloc");
        hip_check(hipHostMalloc((void **)&h_report, 16 * sizeof(unsigned long long), hipHostMallocDefault),
                  "hipHostMalloc");
    }

    uint64_t bytes() const {
        uint64_t b = (uint64_t)slot_cap * (sizeof(cpm::RepDev) + sizeof(cpm::RepDyn) + cpm::kPlanMcs * sizeof(cpm::McsPlan));
        for (const Replica *r : reps)
            if (r) b += r->bytes;
        for (const Group &g : groups) b += g.bytes;
        return b;
    }

    void ensure_slots(int need) {
        if (need <= slot_cap) return;
        int cap = std::max(need, std::max(8, slot_cap * 2));
        cpm::RepDev *nr = nullptr;
        cpm::RepDyn *nd = nullptr;
        hip_check(hipMalloc(&nr, (size_t)cap * sizeof(cpm::RepDev)), "hipMalloc reps");
        hip_check(hipMalloc(&nd, (size_t)cap * sizeof(cpm::RepDyn)), "hipMalloc dyn");
        hip_check(hipMemsetAsync(nr, 0, (size_t)cap * sizeof(cpm::RepDev), stream), "memset");
        hip_check(hipMemsetAsync(nd, 0, (size_t)cap * sizeof(cpm::RepDyn), stream), "memset");
        if (slot_cap) {
            hip_check(hipMemcpyAsync(nr, d_reps, (size_t)slot_cap * sizeof(cpm::RepDev),
                                     hipMemcpyDeviceToDevice, stream), "copy reps");
            hip_check(hipMemcpyAsync(nd, d_dyn, (size_t)slot_cap * sizeof(cpm::RepDyn),
                                     hipMemcpyDeviceToDevice, stream), "copy dyn");
        }
        hip_check(hipStreamSynchronize(stream), "sync");
        dev_free(d_reps);
        dev_free(d_dyn);
        dev_free(d_plan);
        d_reps = nr;
        d_dyn = nd;
        // plans are written at the start of every chunk: nothing to carry over
        d_plan = nullptr;
        hip_check(hipMalloc(&d_plan, (size_t)cap * cpm::kPlanMcs * sizeof(cpm::McsPlan)), "hipMalloc plans");
        slot_cap = cap;
        reps.resize((size_t)cap, nullptr);
        drop_graph();
    }

    // Stream-ordered descriptor upload: kernels enqueued earlier still see the old slot.
    void upload_rep(int slot) {
        hip_check(hipMemcpyAsync(d_reps + slot, &reps[slot]->dev, sizeof(cpm::RepDev), hipMemcpyHostToDevice,
                                 stream),
                  "upload replica");
        metrics[2]++;
    }

    void set_dyn(int slot, uint64_t mcs, uint64_t target, uint32_t in_use, uint32_t mask) {
        hip_check(cpm_internal_launch_set_dyn(d_dyn, slot, mcs, target, in_use, mask, stream), "set_dyn");
        metrics[0]++;
    }

    int add(const cpm_model_desc &desc) {
        if (desc.algorithm != algorithm) fail(CPM_E_INVALID, "replica algorithm differs from the batch");
        Replica *r = new Replica();
        try {
            r->desc = desc;
            compile_replica(*r);
            // admission: resident replicas' estimates plus this one stay within the budget
            r->estimate = estimate_replica_bytes(desc);
            uint64_t resident = 0;
            for (const Replica *q : reps)
                if (q) resident += q->estimate;
            if (budget && resident + r->estimate > budget) fail(CPM_E_NOMEM, "HBM budget exceeded");
            allocate_replica(*r, stream);
        } catch (...) {
            delete r;
            throw;
        }
        int slot = -1;
        for (int s = 0; s < slot_cap; ++s)
            if (!reps[s]) {
                slot = s;
                break;
            }
        if (slot < 0) {
            slot = slot_cap;
            ensure_slots(slot_cap + 1);
        }
        reps[slot] = r;
        int gi = -1;
        for (size_t k = 0; k < groups.size(); ++k)
            if (groups[k].config == r->config) gi = (int)k;
        if (gi < 0) {
            Group g;
            g.config = r->config;
            int dim, halo;
            if (!g.tiled() && !cpm_internal_tile_config(g.config, &dim, &halo, &g.tx, &g.ty, &g.tz))
                fail(CPM_E_STATE, "unknown kernel configuration");
            groups.push_back(g);
            gi = (int)groups.size() - 1;
        }
        r->group = gi;
        groups[gi].slots.push_back(slot);
        groups[gi].dirty = true;
        upload_rep(slot);
        set_dyn(slot, 0, 0, 1, 1 | 2 | 4 | 8);
        metrics[4] = bytes();
        drop_graph();
        return slot;
    }

    // Keep the last-MCS report of every replica of g before its accounting arena moves.
    // One report launch per replica whose last-MCS report is not cached yet, one copy
    // back and one synchronization for the whole group.
    void cache_reports(Group &g) {
        if (g.dirty_arena_fresh || g.slots.empty()) return;
        std::vector<cpm::RepDyn> y((size_t)slot_cap);
        hip_check(hipMemcpyAsync(y.data(), d_dyn, (size_t)slot_cap * sizeof(cpm::RepDyn), hipMemcpyDeviceToHost,
                                 stream), "read progress");
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        metrics[3]++;
        std::vector<int> todo;
        for (int s : g.slots) {
            const Replica *r = reps[s];
            if (!r) continue;
            const uint64_t m = y[(size_t)s].mcs;
            if ((r->report_cached && r->cached_mcs == m) || m <= r->base_mcs) continue;
            todo.push_back(s);
        }
        if (todo.empty()) return;
        if (todo.size() * 8 > many_cap) {
            dev_free(d_many);
            if (h_many) (void)hipHostFree(h_many);
            d_many = nullptr;
            h_many = nullptr;
            many_cap = todo.size() * 8 * 2;
            hip_check(hipMalloc(&d_many, many_cap * sizeof(unsigned long long)), "hipMalloc");
            hip_check(hipHostMalloc((void **)&h_many, many_cap * sizeof(unsigned long long), hipHostMallocDefault),
                      "hipHostMalloc");
        }
        for (size_t k = 0; k < todo.size(); ++k) {
            hip_check(cpm_internal_launch_report(d_reps, g.d_slots, g.ntiles, todo[k], d_many + 8 * k, stream),
                      "report");
            metrics[0]++;
        }
        hip_check(hipMemcpyAsync(h_many, d_many, todo.size() * 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                 stream), "copy reports");
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        metrics[3]++;
        for (size_t k = 0; k < todo.size(); ++k) {
            uint64_t out[12];
            fill_report(reps[todo[k]], y[(size_t)todo[k]].mcs, h_many + 8 * k, out);
        }
    }

    void remove(int slot) {
        Replica *r = check_slot(slot);
        hip_check(hipStreamSynchronize(stream), "sync");
        Group &g = groups[r->group];
        g.slots.erase(std::remove(g.slots.begin(), g.slots.end(), slot), g.slots.end());
        g.dirty = true;
        set_dyn(slot, 0, 0, 0, 1 | 2 | 4 | 8);
        hip_check(hipStreamSynchronize(stream), "sync");
        delete r;
        reps[slot] = nullptr;
        drop_graph();
    }

    Replica *check_slot(int slot) {
        if (slot < 0 || slot >= slot_cap || !reps[slot]) fail(CPM_E_INVALID, "no replica in this slot");
        return reps[slot];
    }

    void rebuild(Group &g) {
        if (g.tiled()) {
            // TiledCheckerboardCPM replicas keep their own layout and counters: no arenas
            g.ntiles = (uint32_t)g.slots.size();
            g.max_rounds = 0;
            std::sort(g.slots.begin(), g.slots.end());
            for (int s : g.slots) upload_rep(s);
            g.dirty = false;
            g.dirty_arena_fresh = false;
            metrics[4] = bytes();
            drop_graph();
            return;
        }
        cache_reports(g);
        std::sort(g.slots.begin(), g.slots.end());
        g.fast_eval = g.slots.empty() ? 0 : reps[g.slots[0]]->dev.fast_eval;
        for (int s : g.slots)
            if (reps[s]->dev.fast_eval != g.fast_eval) g.fast_eval = 0;
        uint32_t tiles = 0, claims = 0;
        int maxr = 1;
        std::vector<cpm::TileInfo> tile_info;
        for (int s : g.slots) {
            Replica *r = reps[s];
            cpm::RepDev &v = r->dev;
            v.tile_base = tiles;
            v.tile_count = (uint32_t)v.ntiles[0] * (uint32_t)v.ntiles[1] * (uint32_t)v.ntiles[2];
            v.claim_base = claims;
            tiles += v.tile_count;
            claims += r->desc.capacity + 1;
            maxr = std::max(maxr, r->rounds);
            for (uint32_t t = 0; t < v.tile_count; ++t) {
                const uint32_t tx = t % (uint32_t)v.ntiles[0];
                const uint32_t ty = (t / (uint32_t)v.ntiles[0]) % (uint32_t)v.ntiles[1];
                const uint32_t tz = t / ((uint32_t)v.ntiles[0] * (uint32_t)v.ntiles[1]);
                tile_info.push_back(cpm::TileInfo{(uint32_t)s, (int32_t)(tx * g.tx), (int32_t)(ty * g.ty),
                                                  (int32_t)(tz * g.tz), v.claim_base, {0, 0, 0}});
            }
        }
        g.ntiles = tiles;
        g.total_claims = std::max(claims, 1u);
        g.seg_cap = (uint32_t)(g.tx * g.ty * g.tz);
        g.max_rounds = maxr;
        const size_t need_tiles = std::max<size_t>(tiles, 1);
        const size_t need_slots = need_tiles * (size_t)maxr * cpm::kSlotFields;
        const size_t need_seg = need_tiles * g.seg_cap;
        hip_check(hipStreamSynchronize(stream), "sync");
        // grow to fit; shrink once the claim-entry arena (the large one) is less than half used
        if (need_tiles > g.cap_tiles || 2 * (size_t)g.total_claims > g.cap_claims || need_slots > g.cap_slots ||
            need_seg > g.cap_seg || 2 * need_seg < g.cap_seg) {
            g.release();
            g.d_tiles = dev_alloc<cpm::TileInfo>(need_tiles, g.bytes);
            g.d_seg_count = dev_alloc<uint32_t>(need_tiles * cpm::kWaves, g.bytes);
            g.d_claims = dev_alloc<unsigned long long>(2 * (size_t)g.total_claims, g.bytes);
            g.d_slots = dev_alloc<uint32_t>(need_slots, g.bytes);
            g.d_seg = dev_alloc<cpm::Entry>(need_seg, g.bytes);
            g.cap_tiles = need_tiles;
            g.cap_claims = 2 * (size_t)g.total_claims;
            g.cap_slots = need_slots;
            g.cap_seg = need_seg;
        }
        hip_check(hipMemsetAsync(g.d_claims, 0, 2 * (size_t)g.total_claims * 8, stream), "memset claims");
        hip_check(hipMemsetAsync(g.d_slots, 0, need_slots * 4, stream), "memset slots");
        hip_check(hipMemsetAsync(g.d_seg_count, 0, need_tiles * cpm::kWaves * 4, stream), "memset seg");
        if (tiles)
            hip_check(hipMemcpy(g.d_tiles, tile_info.data(), tiles * sizeof(cpm::TileInfo), hipMemcpyHostToDevice),
                      "tile map");
        for (int s : g.slots) upload_rep(s);
        g.pass_counter = 0;
        g.dirty = false;
        g.dirty_arena_fresh = true;
        metrics[4] = bytes();
        drop_graph();
    }

    void rebuild_dirty() {
        for (Group &g : groups)
            if (g.dirty) rebuild(g);
    }

    cpm::PassArgs args(const Group &g, uint32_t pass, uint32_t parity, uint32_t plan_k) const {
        cpm::PassArgs a{};
        a.plan = d_plan;
        a.plan_k = plan_k;
        a.reps = d_reps;
        a.dyn = d_dyn;
        a.tiles = g.d_tiles;
        a.claims = g.d_claims;
        a.seg = g.d_seg;
        a.seg_count = g.d_seg_count;
        a.slots = g.d_slots;
        a.ntiles = g.ntiles;
        a.seg_cap = g.seg_cap;
        a.pass = pass;
        a.parity = parity;
        a.total_claims = g.total_claims;
        a.fast_eval = (uint32_t)g.fast_eval;
        return a;
    }

    // Optional per-kernel timing: events[2i], events[2i + 1] bracket the execution of one
    // kernel (attached to its launch), kinds[i] its kind (0 propose or tiled batch, 1 commit
    // or tiled reconciliation, 2 begin).
    std::vector<hipEvent_t> *timing_events = nullptr;
    std::vector<int> *timing_kinds = nullptr;

    void mark(int kind) {
        if (!timing_events) return;
        hipEvent_t ev[2];
        for (hipEvent_t &e : ev) {
            hip_check(hipEventCreate(&e), "hipEventCreate");
            timing_events->push_back(e);
        }
        timing_kinds->push_back(kind);
        cpm_internal_time_next(ev[0], ev[1]);
    }

    // MCS are planned in chunks (plan kernel, the chunk's passes, end kernel) unless a
    // TiledCheckerboardCPM group needs its per-MCS counter reset by the begin kernel.
    bool chunk_plans() const {
        for (const Group &g : groups)
            if (g.ntiles && g.tiled()) return false;
        return true;
    }

    // The colour passes of every non-tiled group for the plan_k-th MCS of the current chunk.
    void enqueue_passes(uint32_t plan_k) {
        for (Group &g : groups) {
            if (g.ntiles == 0) continue;
            if (g.tiled()) {
                enqueue_tiled(g);
                continue;
            }
            for (int p = 0; p < g.max_rounds; ++p) {
                const cpm::PassArgs a = args(g, (uint32_t)p, (uint32_t)(g.pass_counter & 1), plan_k);
                // mechanics: half-interval OU steps before and after the pass
                // (checkerboard.jl:522-557, lottery.jl:353-391); timed with the adjacent kernel
                if (timing_events) {
                    advance_mechanics(g, (uint32_t)p, 0, plan_k);
                    mark(0);
                    hip_check(cpm_internal_launch_propose(g.config, algorithm, &a, stream), "propose");
                    mark(1);
                    hip_check(cpm_internal_launch_commit(&a, stream), "commit");
                    advance_mechanics(g, (uint32_t)p, 1, plan_k);
                } else {
                    advance_mechanics(g, (uint32_t)p, 0, plan_k);
                    hip_check(cpm_internal_launch_pass(g.config, algorithm, &a, stream), "pass kernels");
                    advance_mechanics(g, (uint32_t)p, 1, plan_k);
                }
                metrics[0] += 2;  // propose, commit
                g.pass_counter++;
            }
        }
    }

    // One MCS for every active replica: begin (mcs + 1, order) + the passes.
    void enqueue_mcs() {
        mark(2);
        hip_check(cpm_internal_launch_begin(d_reps, d_dyn, d_plan, slot_cap, stream), "begin_mcs");
        metrics[0]++;
        enqueue_passes(0);
    }

    // n MCS: chunks of at most kPlanMcs MCS, each one plan launch, its passes and one end launch
    // (the counters advance at the end), or one begin launch per MCS where a tiled group runs.
    void enqueue_steps(uint64_t n) {
        if (!chunk_plans()) {
            for (uint64_t i = 0; i < n; ++i) enqueue_mcs();
            return;
        }
        while (n > 0) {
            const int K = (int)std::min<uint64_t>(n, (uint64_t)cpm::kPlanMcs);
            mark(2);
            hip_check(cpm_internal_launch_plan(d_reps, d_dyn, d_plan, slot_cap, K, stream), "plan");
            metrics[0]++;
            for (int k = 0; k < K; ++k) enqueue_passes((uint32_t)k);
            mark(2);
            hip_check(cpm_internal_launch_plan_end(d_reps, d_dyn, d_plan, slot_cap, K, stream), "plan end");
            metrics[0]++;
            n -= (uint64_t)K;
        }
    }

    // TiledCheckerboardCPM MCS of every replica of g: per colour in the MCS's order and per
    // batch of the switching interval, the tiles' proposals, then the reconciliation
    // (tiled_checkerboard_device.jl:743-806).
    void enqueue_tiled(const Group &g) {
        for (int s : g.slots) {
            const Replica *r = reps[s];
            for (int o = 0; o < r->rounds; ++o)
                for (int b = 0; b < r->tiled_batches; ++b) {
                    mark(0);
                    hip_check(cpm_internal_launch_tiled_batch(d_reps, d_dyn, s, (uint32_t)o, (uint32_t)b,
                                                              r->tiled_max_tiles, (uint32_t)r->tiled_interval,
                                                              r->tiled_staged, stream),
                              "tiled batch");
                    mark(1);
                    hip_check(cpm_internal_launch_tiled_reconcile(d_reps, d_dyn, s, r->desc.capacity, stream),
                              "tiled reconcile");
                    metrics[0] += 2;
                }
        }
    }

    void step_timed(uint64_t n, double out[8]) {
        rebuild_dirty();
        for (Group &g : groups) g.dirty_arena_fresh = false;
        std::vector<hipEvent_t> events;
        std::vector<int> kinds;
        hipEvent_t wall[2] = {nullptr, nullptr};
        auto release = [&]() {
            cpm_internal_time_next(nullptr, nullptr);
            for (hipEvent_t e : events) (void)hipEventDestroy(e);
            for (hipEvent_t e : wall)
                if (e) (void)hipEventDestroy(e);
        };
        timing_events = &events;
        timing_kinds = &kinds;
        try {
            for (hipEvent_t &e : wall) hip_check(hipEventCreate(&e), "hipEventCreate");
            hip_check(hipEventRecord(wall[0], stream), "hipEventRecord");
            enqueue_steps(n);
            hip_check(hipEventRecord(wall[1], stream), "hipEventRecord");
        } catch (...) {
            timing_events = nullptr;
            timing_kinds = nullptr;
            (void)hipStreamSynchronize(stream);
            release();
            throw;
        }
        timing_events = nullptr;
        timing_kinds = nullptr;
        try {
            hip_check(hipStreamSynchronize(stream), "sync");
            metrics[1]++;
            for (int k = 0; k < 8; ++k) out[k] = 0.0;
            for (size_t i = 0; i < kinds.size(); ++i) {
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, events[2 * i], events[2 * i + 1]), "hipEventElapsedTime");
                const int kind = kinds[i];
                const int slot = kind == 0 ? 0 : kind == 1 ? 2 : 4;
                out[slot] += ms;
                out[slot + 1] += 1.0;
            }
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, wall[0], wall[1]), "hipEventElapsedTime");
            out[6] = ms;
            out[7] = (double)n;
        } catch (...) {
            release();
            throw;
        }
        release();
    }

    bool graph_parity_ok(uint32_t mcs) const {
        for (const Group &g : groups) {
            if (g.ntiles == 0) continue;
            if ((g.pass_counter & 1) != 0) return false;
            if (((uint64_t)mcs * (uint64_t)g.max_rounds) % 2 != 0) return false;
        }
        return true;
    }

    void capture(uint32_t mcs) {
        drop_graph();
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        std::vector<uint64_t> saved;
        for (Group &g : groups) saved.push_back(g.pass_counter);
        try {
            enqueue_steps(mcs);
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(stream, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &graph), "end capture");
        hip_check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "graph instantiate");
        for (size_t k = 0; k < groups.size(); ++k) groups[k].pass_counter = saved[k];
        graph_mcs = mcs;
        graph_valid = true;
    }

    // Captures the graph that the next steps replay, performing no MCS, so that capture and
    // instantiation stay out of a caller's timed loop. False when graphs are off or the claim
    // buffers' parity would make the next step start with a plain MCS.
    bool prepare_graph() {
        rebuild_dirty();
        if (graph_chunk == 0) return false;
        uint32_t chunk = graph_chunk;
        bool even = true;
        for (const Group &g : groups)
            if (g.ntiles && ((uint64_t)chunk * g.max_rounds) % 2) even = false;
        if (!even) chunk *= 2;
        if (!graph_parity_ok(chunk)) return false;
        if (!graph_valid || graph_mcs != chunk) capture(chunk);
        return true;
    }

    void step(uint64_t n) {
        rebuild_dirty();
        for (Group &g : groups) g.dirty_arena_fresh = false;
        while (n > 0) {
            if (graph_chunk > 0) {
                uint32_t chunk = graph_chunk;
                bool even = true;
                for (const Group &g : groups)
                    if (g.ntiles && ((uint64_t)chunk * g.max_rounds) % 2) even = false;
                if (!even) chunk *= 2;
                if (n >= chunk && graph_parity_ok(chunk)) {
                    if (!graph_valid || graph_mcs != chunk) capture(chunk);
                    hip_check(hipGraphLaunch(graph_exec, stream), "graph launch");
                    metrics[5]++;
                    for (Group &g : groups) g.pass_counter += (uint64_t)chunk * g.max_rounds;
                    n -= chunk;
                    continue;
                }
                if (n >= chunk) {
                    // one MCS of plain launches brings the claim-buffer parity back for the graph
                    enqueue_steps(1);
                    n -= 1;
                    continue;
                }
            }
            // the rest (or every MCS without graphs): planned chunks of plain launches
            enqueue_steps(n);
            n = 0;
        }
    }

    void sync() {
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        metrics[1]++;
    }

    uint64_t read_mcs(int slot) {
        cpm::RepDyn y;
        hip_check(hipStreamSynchronize(stream), "sync");
        hip_check(hipMemcpy(&y, d_dyn + slot, sizeof(uint64_t) * 2 + sizeof(uint32_t) * 2, hipMemcpyDeviceToHost),
                  "read mcs");
        metrics[1]++;
        metrics[3]++;
        return y.mcs;
    }

    void report(int slot, uint64_t out[12]) {
        Replica *r = check_slot(slot);
        const uint64_t mcs = read_mcs(slot);
        if (r->report_cached && r->cached_mcs == mcs) {
            std::memcpy(out, r->cached_report, sizeof(r->cached_report));
            return;
        }
        if (mcs <= r->base_mcs) fail(CPM_E_STATE, "no MCS has completed since the state was set");
        const Group &g = groups[r->group];
        if (g.dirty_arena_fresh) fail(CPM_E_STATE, "replica accounting was discarded by a membership change");
        if (g.tiled()) {
            // the replica's MCS counters, zeroed by the begin kernel of each MCS
            hip_check(hipMemcpyAsync(h_report, r->d_tiled_counts, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                     stream), "copy report");
        } else {
            hip_check(cpm_internal_launch_report(d_reps, g.d_slots, g.ntiles, slot, d_report, stream), "report");
            metrics[0]++;
            hip_check(hipMemcpyAsync(h_report, d_report, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                     stream), "copy report");
        }
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        metrics[3]++;
        if (g.tiled()) fill_tiled_report(r, mcs, h_report, out);
        else fill_report(r, mcs, h_report, out);
    }

    // TiledCheckerboardCPM report (tiled_checkerboard_device.jl:790-806): internal rounds
    // colours x batches, every mutable site a candidate and activated, no conflicts.
    void fill_tiled_report(Replica *r, uint64_t mcs, const unsigned long long *s, uint64_t out[12]) {
        out[0] = mcs;
        out[1] = (uint64_t)r->rounds * (uint64_t)r->tiled_batches;
        out[2] = out[3] = (uint64_t)r->nmut;
        out[4] = s[0];
        out[5] = s[1];
        out[6] = s[2];
        out[7] = 0;
        out[8] = 0;
        out[9] = 0;
        out[10] = s[3];
        out[11] = s[4];
        std::memcpy(r->cached_report, out, sizeof(r->cached_report));
        r->cached_mcs = mcs;
        r->report_cached = true;
    }

    // The 12-field report (sequential.jl:236-251) of replica r from its 7 summed counters.
    void fill_report(Replica *r, uint64_t mcs, const unsigned long long *s, uint64_t out[12]) {
        const uint64_t rounds = (uint64_t)r->rounds;
        out[0] = mcs;
        out[1] = rounds;
        out[2] = algorithm == CPM_ALG_LOTTERY ? (uint64_t)r->nmut * rounds : s[0];
        out[3] = s[0];
        out[4] = s[3];
        out[5] = s[1];
        out[6] = s[2];
        out[7] = 0;
        out[8] = s[4];
        out[9] = 0;
        out[10] = s[5];
        out[11] = s[6];
        std::memcpy(r->cached_report, out, sizeof(r->cached_report));
        double dh;
        std::memcpy(&dh, &s[7], sizeof(dh));
        r->cached_dh = dh;
        r->cached_mcs = mcs;
        r->report_cached = true;
    }

    // FP64 sum of the accepted copies' delta_h over the replica's last MCS (report counters).
    double report_energy(int slot) {
        uint64_t out[12];
        report(slot, out);
        Replica *r = reps[slot];
        if (groups[r->group].tiled())
            fail(CPM_E_UNSUPPORTED, "TiledCheckerboardCPM evaluates against tile-local views; no energy report");
        return r->cached_dh;
    }

    // Global FP64 energy of the replica's current state; buf holds kEnergyParts + 1 doubles of
    // device memory (partial sums, then the total). One launch pair, one copy, one sync.
    double total_energy(int slot, double *buf) {
        check_slot(slot);
        hip_check(cpm_internal_launch_energy(d_reps, slot, buf, stream), "energy");
        metrics[0] += 2;
        hip_check(hipMemcpyAsync(h_report, buf + cpm::kEnergyParts, sizeof(double), hipMemcpyDeviceToHost, stream),
                  "copy energy");
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        metrics[3]++;
        double e;
        std::memcpy(&e, h_report, sizeof(e));
        return e;
    }

    void upload_state(int slot, const uint8_t *tags, const uint32_t *ids, const int32_t *fv,
                      const int32_t *mv, const int64_t *bm, const uint8_t *active, const uint32_t *ctype,
                      const double *tv, const double *vs, const int64_t *tb, const double *bs) {
        Replica *r = check_slot(slot);
        if (!tags || !ids || !fv || !bm || !active || !ctype)
            fail(CPM_E_INVALID, "ownership, volume, boundary, active and type arrays are required");
        const size_t n = (size_t)r->n, cap = r->desc.capacity, cap1 = cap + 1;
        std::vector<int32_t> words(n);
        for (size_t i = 0; i < n; ++i) {
            const uint32_t id = ids[i];
            if (id == 0 || id > 0x7fffffffu) fail(CPM_E_INVALID, "owner identity must be positive and fit Int32");
            if (tags[i] == 1) {
                if (id > cap) fail(CPM_E_INVALID, "cell owner exceeds the capacity");
                // the cell's word carries its type (cpm_device.h: id | type << kTypeShift)
                if (ctype[id - 1] > 31u)
                    fail(CPM_E_UNSUPPORTED, "a cell type beyond 31 does not fit the owner word");
                words[i] = (int32_t)(id | (ctype[id - 1] << cpm::kTypeShift));
            } else if (tags[i] == 2) {
                words[i] = -(int32_t)id;
            } else {
                fail(CPM_E_INVALID, "owner tag must identify a cell or medium domain");
            }
        }
        // per-cell arrays -> 64-byte cell records (entry 0 is the spare read for media)
        std::vector<cpm::CellRec> cells(cap1);
        std::memset(cells.data(), 0, cap1 * sizeof(cpm::CellRec));
        for (size_t c = 0; c < cap; ++c) {
            cpm::CellRec &q = cells[c + 1];
            q.vol = fv[c];
            q.bnd = bm[c];
            q.ctype = ctype[c];
            q.active = active[c];
            q.tv = tv ? tv[c] : 0.0;
            q.vs = vs ? vs[c] : 0.0;
            q.tb = tb ? tb[c] : 0;
            q.bs = bs ? bs[c] : 0.0;
        }
        if (has_term(r->desc, CPM_TERM_CONTACT))
            for (size_t c = 0; c < cap; ++c)
                if (active[c] && (ctype[c] < 1 || ctype[c] > (uint32_t)r->desc.n_types))
                    fail(CPM_E_INVALID, "an owner type lies outside the contact interaction matrix");
        std::vector<int32_t> media(CPM_MAX_MEDIA, 0);
        for (int i = 0; i < r->desc.n_media; ++i) media[i] = mv ? mv[i] : 0;
        hip_check(hipStreamSynchronize(stream), "sync");
        hip_check(hipMemcpy(r->d_lattice, words.data(), n * 4, hipMemcpyHostToDevice), "upload lattice");
        hip_check(hipMemcpy(r->d_cells, cells.data(), cap1 * sizeof(cpm::CellRec), hipMemcpyHostToDevice), "upload");
        hip_check(hipMemcpy(r->d_mvol, media.data(), CPM_MAX_MEDIA * 4, hipMemcpyHostToDevice), "upload");
        metrics[2]++;
        r->base_mcs = read_mcs(slot);
        r->report_cached = false;
    }

    void download_state(int slot, uint8_t *tags, uint32_t *ids, int32_t *fv, int32_t *mv, int64_t *bm,
                        uint8_t *active, uint32_t *ctype, double *tv, double *vs, int64_t *tb, double *bs) {
        Replica *r = check_slot(slot);
        const size_t n = (size_t)r->n, cap = r->desc.capacity, cap1 = cap + 1;
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        metrics[3]++;
        if (tags || ids) {
            std::vector<int32_t> words(n);
            hip_check(hipMemcpy(words.data(), r->d_lattice, n * 4, hipMemcpyDeviceToHost), "download lattice");
            unpack_state(words.data(), n, nullptr, 0, tags, ids, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr, nullptr);
        }
        if (fv || bm || active || ctype || tv || vs || tb || bs) {
            // 64-byte cell records -> per-cell arrays
            std::vector<cpm::CellRec> cells(cap1);
            hip_check(hipMemcpy(cells.data(), r->d_cells, cap1 * sizeof(cpm::CellRec), hipMemcpyDeviceToHost),
                      "download");
            unpack_state(nullptr, 0, cells.data(), cap, nullptr, nullptr, fv, bm, active, ctype, tv, vs, tb, bs);
        }
        if (mv && r->desc.n_media > 0)
            hip_check(hipMemcpy(mv, r->d_mvol, (size_t)r->desc.n_media * 4, hipMemcpyDeviceToHost), "download");
    }

    // The Float32 mechanical state fields of the cell records (strided copies, one per column).
    void upload_mechanics(int slot, const float *qp, const float *qt) {
        Replica *r = check_slot(slot);
        const size_t cap = r->desc.capacity;
        hip_check(hipStreamSynchronize(stream), "sync");
        if (qp)
            hip_check(hipMemcpy2D(&r->d_cells[1].qp, sizeof(cpm::CellRec), qp, sizeof(float), sizeof(float), cap,
                                  hipMemcpyHostToDevice), "upload mechanics");
        if (qt)
            hip_check(hipMemcpy2D(&r->d_cells[1].qt, sizeof(cpm::CellRec), qt, sizeof(float), sizeof(float), cap,
                                  hipMemcpyHostToDevice), "upload mechanics");
        metrics[2]++;
    }

    void download_mechanics(int slot, float *qp, float *qt) {
        Replica *r = check_slot(slot);
        const size_t cap = r->desc.capacity;
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        if (qp)
            hip_check(hipMemcpy2D(qp, sizeof(float), &r->d_cells[1].qp, sizeof(cpm::CellRec), sizeof(float), cap,
                                  hipMemcpyDeviceToHost), "download mechanics");
        if (qt)
            hip_check(hipMemcpy2D(qt, sizeof(float), &r->d_cells[1].qt, sizeof(cpm::CellRec), sizeof(float), cap,
                                  hipMemcpyDeviceToHost), "download mechanics");
        metrics[3]++;
    }

    void initialize_mechanics(int slot) {
        Replica *r = check_slot(slot);
        if (r->desc.n_mechanics == 0) return;
        hip_check(cpm_internal_launch_mech_init(d_reps, slot, r->desc.capacity, stream), "mechanics init");
        metrics[0]++;
    }

    // Half-interval OU steps of the replicas of a group that carry mechanics.
    void advance_mechanics(const Group &g, uint32_t pass, uint32_t phase, uint32_t plan_k) {
        for (int s : g.slots) {
            const Replica *r = reps[s];
            if (!r || r->desc.n_mechanics == 0) continue;
            hip_check(cpm_internal_launch_mech_advance(d_reps, d_plan, s, r->desc.capacity, pass, phase, plan_k,
                                                       stream),
                      "mechanics");
            metrics[0]++;
        }
    }

    void set_seed(int slot, uint64_t seed) {
        Replica *r = check_slot(slot);
        r->desc.seed = seed;
        const uint64_t key = seed ^ mix64(kGenerationDomain);
        r->dev.key_lo = (uint32_t)key;
        r->dev.key_hi = (uint32_t)(key >> 32);
        upload_rep(slot);
    }

    // Device-to-device copies of the replica's state into a new snapshot at MCS `mcs`.
    cpm_snapshot *capture_snapshot(int slot, uint64_t mcs) {
        Replica *r = check_slot(slot);
        std::unique_ptr<cpm_snapshot> sn(new cpm_snapshot());
        const size_t n = (size_t)r->n, cap1 = (size_t)r->desc.capacity + 1;
        sn->ndim = r->desc.ndim;
        for (int a = 0; a < 3; ++a) sn->dims[a] = r->desc.dims[a];
        sn->n = n;
        sn->capacity = r->desc.capacity;
        sn->n_media = r->desc.n_media;
        sn->n_types = r->desc.n_types;
        sn->mcs = mcs;
        sn->seed = r->desc.seed;
        sn->d_lattice = dev_alloc<int32_t>(n, sn->bytes);
        sn->d_cells = dev_alloc<cpm::CellRec>(cap1, sn->bytes);
        sn->d_mvol = dev_alloc<int32_t>(CPM_MAX_MEDIA, sn->bytes);
        hip_check(hipMemcpyAsync(sn->d_lattice, r->d_lattice, n * 4, hipMemcpyDeviceToDevice, stream), "snapshot");
        hip_check(hipMemcpyAsync(sn->d_cells, r->d_cells, cap1 * sizeof(cpm::CellRec), hipMemcpyDeviceToDevice,
                                 stream), "snapshot");
        hip_check(hipMemcpyAsync(sn->d_mvol, r->d_mvol, CPM_MAX_MEDIA * 4, hipMemcpyDeviceToDevice, stream),
                  "snapshot");
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        return sn.release();
    }

    // The snapshot's state copied into the replica (device to device) with its seed; the
    // caller sets the MCS counter. The replica must have the snapshot's lattice, capacity,
    // media and cell types.
    void restore_snapshot(int slot, const cpm_snapshot &sn) {
        Replica *r = check_slot(slot);
        bool same = sn.ndim == r->desc.ndim && sn.capacity == r->desc.capacity && sn.n_media == r->desc.n_media &&
                    sn.n_types == r->desc.n_types && sn.n == (size_t)r->n;
        for (int a = 0; a < 3; ++a) same = same && sn.dims[a] == r->desc.dims[a];
        if (!same) fail(CPM_E_INVALID, "the snapshot's lattice, capacity, media or cell types differ from the replica's");
        const size_t cap1 = (size_t)r->desc.capacity + 1;
        hip_check(hipStreamSynchronize(stream), "sync");
        hip_check(hipMemcpyAsync(r->d_lattice, sn.d_lattice, sn.n * 4, hipMemcpyDeviceToDevice, stream), "restore");
        hip_check(hipMemcpyAsync(r->d_cells, sn.d_cells, cap1 * sizeof(cpm::CellRec), hipMemcpyDeviceToDevice,
                                 stream), "restore");
        hip_check(hipMemcpyAsync(r->d_mvol, sn.d_mvol, CPM_MAX_MEDIA * 4, hipMemcpyDeviceToDevice, stream),
                  "restore");
        hip_check(hipStreamSynchronize(stream), "sync");
        metrics[1]++;
        if (sn.seed != r->desc.seed) set_seed(slot, sn.seed);
        r->base_mcs = sn.mcs;
        r->report_cached = false;
    }
};

template <typename F>
int32_t guarded(std::string *err, F &&fn) {
    try {
        fn();
        return CPM_OK;
    } catch (const Failure &f) {
        if (err) *err = f.message;
        return f.code;
    } catch (const std::bad_alloc &) {
        if (err) *err = "host allocation failed";
        return CPM_E_NOMEM;
    } catch (...) {
        if (err) *err = "unexpected failure";
        return CPM_E_INVALID;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Static scratch of cpm_total_energy_noalloc: the partial sums and the total of one global
// energy reduction (kEnergyParts + 1 doubles = 8 KiB on every device the library runs on,
// held until the process ends). Host code never passes `scratch` itself to a kernel (on the
// host it only names the symbol): cpm_create resolves its device address once per device
// with hipGetSymbolAddress, with that device current, and every use goes through that address.
// Every call on one device shares the array; each reduction overwrites all of it, so nothing
// needs clearing, and calls that overlap on one device race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[cpm::kEnergyParts + 1];
constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mutex;
static double *g_scratch_address[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, resolved on the first call per device.
static double *resolve_scratch(int *device) {
    hip_check(hipGetDevice(device), "hipGetDevice");
    if (*device < 0 || *device >= kMaxScratchDevices) fail(CPM_E_UNSUPPORTED, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[*device]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_address[*device] = static_cast<double *>(p);
    }
    return g_scratch_address[*device];
}

struct cpm_ctx {
    Engine engine;
    int slot = 0;
    int device = -1;             // device current at cpm_create
    double *scratch_ptr = nullptr;  // device address of `scratch` on that device
    uint64_t host_mcs = 0;
    double *d_energy = nullptr;  // cpm_total_energy's partial sums, allocated by its first call
    ~cpm_ctx() {
        if (d_energy) (void)hipFree(d_energy);
    }
};

struct cpm_batch {
    Engine engine;
};

extern "C" {

const char *cpm_last_error(const cpm_ctx *ctx) {
    return ctx ? ctx->engine.err.c_str() : g_create_error.c_str();
}

int32_t cpm_abi_version(void) { return CPM_ABI_VERSION; }

int32_t cpm_create(const cpm_model_desc *desc, cpm_ctx **out) {
    if (!desc || !out) {
        g_create_error = "null argument";
        return CPM_E_INVALID;
    }
    cpm_ctx *ctx = nullptr;
    const int32_t rc = guarded(&g_create_error, [&] {
        ctx = new cpm_ctx();
        ctx->scratch_ptr = resolve_scratch(&ctx->device);
        ctx->engine.init(desc->algorithm);
        ctx->slot = ctx->engine.add(*desc);
        ctx->engine.rebuild_dirty();
        hip_check(hipStreamSynchronize(ctx->engine.stream), "sync");
    });
    if (rc != CPM_OK) {
        delete ctx;
        *out = nullptr;
        return rc;
    }
    *out = ctx;
    return CPM_OK;
}

int32_t cpm_destroy(cpm_ctx *ctx) {
    delete ctx;
    return CPM_OK;
}

int32_t cpm_upload_state(cpm_ctx *ctx, const uint8_t *tags, const uint32_t *ids,
                         const int32_t *finite_volumes, const int32_t *medium_volumes,
                         const int64_t *boundary_measures, const uint8_t *active,
                         const uint32_t *cell_types, const double *target_volume,
                         const double *volume_strength, const int64_t *target_boundary,
                         const double *boundary_strength) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        ctx->engine.upload_state(ctx->slot, tags, ids, finite_volumes, medium_volumes, boundary_measures,
                                 active, cell_types, target_volume, volume_strength, target_boundary,
                                 boundary_strength);
    });
}

int32_t cpm_download_state(cpm_ctx *ctx, uint8_t *tags, uint32_t *ids, int32_t *finite_volumes,
                           int32_t *medium_volumes, int64_t *boundary_measures, uint8_t *active,
                           uint32_t *cell_types, double *target_volume, double *volume_strength,
                           int64_t *target_boundary, double *boundary_strength) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        ctx->engine.download_state(ctx->slot, tags, ids, finite_volumes, medium_volumes,
                                   boundary_measures, active, cell_types, target_volume,
                                   volume_strength, target_boundary, boundary_strength);
    });
}

int32_t cpm_upload_mechanics(cpm_ctx *ctx, const float *volume_pressure, const float *surface_tension) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] { ctx->engine.upload_mechanics(ctx->slot, volume_pressure, surface_tension); });
}

int32_t cpm_download_mechanics(cpm_ctx *ctx, float *volume_pressure, float *surface_tension) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err,
                   [&] { ctx->engine.download_mechanics(ctx->slot, volume_pressure, surface_tension); });
}

int32_t cpm_initialize_mechanics(cpm_ctx *ctx) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] { ctx->engine.initialize_mechanics(ctx->slot); });
}

int32_t cpm_set_seed(cpm_ctx *ctx, uint64_t seed) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] { ctx->engine.set_seed(ctx->slot, seed); });
}

int32_t cpm_set_mcs(cpm_ctx *ctx, uint64_t mcs) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        ctx->engine.set_dyn(ctx->slot, mcs, mcs, 1, 1 | 2 | 8);
        ctx->host_mcs = mcs;
        Replica *r = ctx->engine.check_slot(ctx->slot);
        r->base_mcs = mcs;
        r->report_cached = false;
    });
}

int32_t cpm_get_mcs(const cpm_ctx *ctx, uint64_t *mcs) {
    if (!ctx || !mcs) return CPM_E_INVALID;
    *mcs = ctx->host_mcs;
    return CPM_OK;
}

int32_t cpm_step(cpm_ctx *ctx, uint64_t n) {
    if (!ctx) return CPM_E_INVALID;
    if (n == 0) return CPM_OK;
    return guarded(&ctx->engine.err, [&] {
        Engine &e = ctx->engine;
        const uint64_t target = ctx->host_mcs + n;
        e.set_dyn(ctx->slot, 0, target, 1, 2);
        e.step(n);
        ctx->host_mcs = target;
    });
}

int32_t cpm_report(cpm_ctx *ctx, uint64_t out[12]) {
    if (!ctx || !out) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] { ctx->engine.report(ctx->slot, out); });
}

int32_t cpm_report_energy(cpm_ctx *ctx, double *accepted_delta_h) {
    if (!ctx || !accepted_delta_h) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] { *accepted_delta_h = ctx->engine.report_energy(ctx->slot); });
}

int32_t cpm_total_energy(cpm_ctx *ctx, double *energy) {
    if (!ctx || !energy) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        if (!ctx->d_energy)
            hip_check(hipMalloc(&ctx->d_energy, (cpm::kEnergyParts + 1) * sizeof(double)), "hipMalloc energy");
        *energy = ctx->engine.total_energy(ctx->slot, ctx->d_energy);
    });
}

int32_t cpm_total_energy_noalloc(cpm_ctx *ctx, double *energy) {
    if (!ctx || !energy) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        static_assert(sizeof(scratch) == (cpm::kEnergyParts + 1) * sizeof(double),
                      "the reduction's fixed grid writes kEnergyParts partial sums and one total");
        int current = -1;
        hip_check(hipGetDevice(&current), "hipGetDevice");
        if (current != ctx->device)
            fail(CPM_E_STATE, "the context's device is not current (its scratch address belongs to its device)");
        if (!ctx->scratch_ptr) fail(CPM_E_STATE, "scratch address was not resolved at cpm_create");
        *energy = ctx->engine.total_energy(ctx->slot, ctx->scratch_ptr);
    });
}

int32_t cpm_sync(cpm_ctx *ctx) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] { ctx->engine.sync(); });
}

int32_t cpm_schedule(cpm_ctx *ctx, int32_t *round_count, uint32_t *order) {
    if (!ctx) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        Engine &e = ctx->engine;
        Replica *r = e.check_slot(ctx->slot);
        if (round_count) *round_count = r->rounds;
        if (order) {
            cpm::RepDyn y;
            hip_check(hipStreamSynchronize(e.stream), "sync");
            hip_check(hipMemcpy(&y, e.d_dyn + ctx->slot, sizeof(y), hipMemcpyDeviceToHost), "copy order");
            e.metrics[1]++;
            e.metrics[3]++;
            std::memcpy(order, y.order, (size_t)r->rounds * 4);
        }
    });
}

int32_t cpm_site_colors(const cpm_ctx *ctx, uint8_t *colors) {
    if (!ctx || !colors) return CPM_E_INVALID;
    const Replica *r = ctx->engine.reps[ctx->slot];
    if (r->desc.algorithm == CPM_ALG_LOTTERY) return CPM_E_STATE;
    std::memcpy(colors, r->colors.data(), (size_t)r->n);
    return CPM_OK;
}

int32_t cpm_model_colors(const cpm_model_desc *desc, uint8_t *colors, int32_t *count) {
    if (!desc) {
        g_create_error = "null argument";
        return CPM_E_INVALID;
    }
    return guarded(&g_create_error, [&] {
        if (desc->algorithm == CPM_ALG_LOTTERY) fail(CPM_E_INVALID, "LotteryCPM has no realized colouring");
        Replica r;
        r.desc = *desc;
        compile_replica(r);  // host work only
        if (colors) std::memcpy(colors, r.colors.data(), (size_t)r.n);
        if (count) *count = r.rounds;
    });
}

int32_t cpm_tiled_layout(const cpm_model_desc *desc, int32_t grid[3], int32_t reach[3], int32_t *halo,
                         uint16_t *tile_colors, int32_t *color_count, int32_t *interval, int32_t *batches,
                         int32_t *staged) {
    if (!desc) {
        g_create_error = "null argument";
        return CPM_E_INVALID;
    }
    return guarded(&g_create_error, [&] {
        if (desc->algorithm != CPM_ALG_TILED) fail(CPM_E_INVALID, "not a TiledCheckerboardCPM model");
        Replica r;
        r.desc = *desc;
        compile_replica(r);  // host work only
        for (int a = 0; a < 3; ++a) {
            if (grid) grid[a] = r.tile_grid[a];
            if (reach) reach[a] = r.tile_reach[a];
        }
        if (halo) *halo = r.tiled_halo;
        if (tile_colors) std::memcpy(tile_colors, r.tile_colors.data(), r.tile_colors.size() * 2);
        if (color_count) *color_count = r.rounds;
        if (interval) *interval = r.tiled_interval;
        if (staged) *staged = (int32_t)r.tiled_staged;
        if (batches) *batches = r.tiled_batches;
    });
}

int32_t cpm_kernel_config(const cpm_ctx *ctx, int32_t *config) {
    if (!ctx || !config) return CPM_E_INVALID;
    const Replica *r = ctx->engine.reps[ctx->slot];
    *config = r->config | (r->dev.fast_eval << 8);
    return CPM_OK;
}

int32_t cpm_metrics(const cpm_ctx *ctx, uint64_t out[6]) {
    if (!ctx || !out) return CPM_E_INVALID;
    std::memcpy(out, ctx->engine.metrics, sizeof(ctx->engine.metrics));
    return CPM_OK;
}

int32_t cpm_step_timed(cpm_ctx *ctx, uint64_t n, double out[8]) {
    if (!ctx || !out) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        Engine &e = ctx->engine;
        const uint64_t target = ctx->host_mcs + n;
        e.set_dyn(ctx->slot, 0, target, 1, 2);
        e.step_timed(n, out);
        ctx->host_mcs = target;
    });
}

int32_t cpm_set_graph_chunk(cpm_ctx *ctx, uint32_t chunk) {
    if (!ctx) return CPM_E_INVALID;
    ctx->engine.graph_chunk = chunk;
    ctx->engine.drop_graph();
    return CPM_OK;
}

int32_t cpm_prepare_graph(cpm_ctx *ctx, int32_t *ready) {
    if (!ctx || !ready) return CPM_E_INVALID;
    *ready = 0;
    return guarded(&ctx->engine.err, [&] { *ready = ctx->engine.prepare_graph() ? 1 : 0; });
}

uint64_t cpm_ensemble_seed(uint64_t master_seed, uint32_t trajectory, uint32_t rerun) {
    // EnsembleStream = 16, operation 1, EnsembleEntity = 3, invocation = rerun - 1
    const uint64_t c0 = (uint64_t)((rerun - 1) & 0xffu) << 56;
    const uint64_t c1 = (uint64_t)trajectory | (1ull << 32) | (16ull << 44) | (3ull << 52);
    const uint64_t key = master_seed ^ mix64(kGenerationDomain);
    uint32_t c[4] = {(uint32_t)c0, (uint32_t)(c0 >> 32), (uint32_t)c1, (uint32_t)(c1 >> 32)};
    philox_host(c, (uint32_t)key, (uint32_t)(key >> 32));
    return (uint64_t)c[0] | ((uint64_t)c[1] << 32);
}

int32_t cpm_tile_shape_info(int32_t shape, int32_t out[5]) {
    if (!out || shape < 0 || shape > 15) return CPM_E_INVALID;
    int dim, halo, tx, ty, tz;
    if (!cpm_internal_tile_config(shape, &dim, &halo, &tx, &ty, &tz)) return CPM_E_INVALID;
    out[0] = dim;
    out[1] = halo;
    out[2] = tx;
    out[3] = ty;
    out[4] = tz;
    return CPM_OK;
}

int32_t cpm_set_tile_shape(int32_t algorithm, int32_t ndim, int32_t halo, int32_t shape) {
    if (algorithm < 1 || algorithm > 2 || ndim < 2 || ndim > 3 || halo < 1 || halo > 2) return CPM_E_INVALID;
    if (shape >= 0) {
        int dim, h, tx, ty, tz;
        if (shape > 15 || !cpm_internal_tile_config(shape, &dim, &h, &tx, &ty, &tz) || dim != ndim || h != halo)
            return CPM_E_INVALID;
    }
    g_shape_pref[algorithm][ndim][halo].store(shape >= 0 ? shape + 1 : 0);
    return CPM_OK;
}

int32_t cpm_set_tiled_staging(int32_t enable) {
    g_tiled_staging.store(enable ? 1 : 0);
    return CPM_OK;
}

int32_t cpm_set_row_kernels(int32_t enable) {
    g_row_kernels.store(enable ? 1 : 0);
    return CPM_OK;
}

int32_t cpm_copy_probe(uint64_t bytes, int32_t reps, double *out_ms) {
    if (!out_ms || reps < 1 || bytes < 16 || bytes % 16) {
        g_create_error = "cpm_copy_probe: bytes must be a positive multiple of 16, reps >= 1";
        return CPM_E_INVALID;
    }
    void *src = nullptr, *dst = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    const int32_t rc = guarded(&g_create_error, [&] {
        hip_check(hipMalloc(&src, bytes), "copy probe alloc");
        hip_check(hipMalloc(&dst, bytes), "copy probe alloc");
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        hip_check(hipEventCreate(&ev[0]), "event");
        hip_check(hipEventCreate(&ev[1]), "event");
        hip_check(hipMemsetAsync(src, 0x5a, bytes, stream), "memset");
        // one untimed launch, then `reps` timed launches
        hip_check(cpm_internal_launch_copy_probe(src, dst, bytes / 16, stream), "copy probe");
        hip_check(hipEventRecord(ev[0], stream), "event");
        for (int32_t i = 0; i < reps; ++i)
            hip_check(cpm_internal_launch_copy_probe(src, dst, bytes / 16, stream), "copy probe");
        hip_check(hipEventRecord(ev[1], stream), "event");
        hip_check(hipEventSynchronize(ev[1]), "sync");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ev[0], ev[1]), "elapsed");
        *out_ms = (double)ms / reps;
    });
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return rc;
}

// ---------------- batch ----------------
int32_t cpm_batch_create(int32_t algorithm, uint64_t hbm_budget_bytes, cpm_batch **out) {
    if (!out) return CPM_E_INVALID;
    if (algorithm != CPM_ALG_CHECKERBOARD && algorithm != CPM_ALG_LOTTERY && algorithm != CPM_ALG_TILED) {
        g_create_error = "unknown algorithm";
        return CPM_E_INVALID;
    }
    cpm_batch *b = nullptr;
    const int32_t rc = guarded(&g_create_error, [&] {
        b = new cpm_batch();
        b->engine.init(algorithm);
        b->engine.budget = hbm_budget_bytes;
    });
    if (rc != CPM_OK) {
        delete b;
        *out = nullptr;
        return rc;
    }
    *out = b;
    return CPM_OK;
}

int32_t cpm_batch_destroy(cpm_batch *b) {
    delete b;
    return CPM_OK;
}

const char *cpm_batch_last_error(const cpm_batch *b) {
    return b ? b->engine.err.c_str() : g_create_error.c_str();
}

int32_t cpm_replica_bytes(const cpm_model_desc *desc, uint64_t *bytes) {
    if (!desc || !bytes) return CPM_E_INVALID;
    *bytes = estimate_replica_bytes(*desc);
    return CPM_OK;
}

int32_t cpm_batch_add(cpm_batch *b, const cpm_model_desc *desc, int32_t *slot) {
    if (!b || !desc || !slot) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { *slot = b->engine.add(*desc); });
}

int32_t cpm_batch_remove(cpm_batch *b, int32_t slot) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { b->engine.remove(slot); });
}

int32_t cpm_batch_upload_state(cpm_batch *b, int32_t slot, const uint8_t *tags,
                               const uint32_t *ids, const int32_t *finite_volumes,
                               const int32_t *medium_volumes, const int64_t *boundary_measures,
                               const uint8_t *active, const uint32_t *cell_types,
                               const double *target_volume, const double *volume_strength,
                               const int64_t *target_boundary, const double *boundary_strength) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] {
        b->engine.upload_state(slot, tags, ids, finite_volumes, medium_volumes, boundary_measures, active,
                               cell_types, target_volume, volume_strength, target_boundary,
                               boundary_strength);
    });
}

int32_t cpm_batch_download_state(cpm_batch *b, int32_t slot, uint8_t *tags, uint32_t *ids,
                                 int32_t *finite_volumes, int32_t *medium_volumes,
                                 int64_t *boundary_measures, uint8_t *active,
                                 uint32_t *cell_types, double *target_volume,
                                 double *volume_strength, int64_t *target_boundary,
                                 double *boundary_strength) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] {
        b->engine.download_state(slot, tags, ids, finite_volumes, medium_volumes, boundary_measures,
                                 active, cell_types, target_volume, volume_strength, target_boundary,
                                 boundary_strength);
    });
}

int32_t cpm_batch_set_target(cpm_batch *b, int32_t slot, uint64_t target_mcs) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] {
        b->engine.check_slot(slot);
        b->engine.set_dyn(slot, 0, target_mcs, 1, 2);
    });
}

int32_t cpm_batch_upload_mechanics(cpm_batch *b, int32_t slot, const float *volume_pressure,
                                   const float *surface_tension) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { b->engine.upload_mechanics(slot, volume_pressure, surface_tension); });
}

int32_t cpm_batch_download_mechanics(cpm_batch *b, int32_t slot, float *volume_pressure, float *surface_tension) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { b->engine.download_mechanics(slot, volume_pressure, surface_tension); });
}

int32_t cpm_batch_initialize_mechanics(cpm_batch *b, int32_t slot) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { b->engine.initialize_mechanics(slot); });
}

int32_t cpm_batch_set_mcs(cpm_batch *b, int32_t slot, uint64_t mcs) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] {
        Replica *r = b->engine.check_slot(slot);
        b->engine.set_dyn(slot, mcs, mcs, 1, 1 | 2 | 8);
        r->base_mcs = mcs;
        r->report_cached = false;
    });
}

int32_t cpm_snapshot_capture(cpm_ctx *ctx, cpm_snapshot **out) {
    if (!ctx || !out) return CPM_E_INVALID;
    *out = nullptr;
    return guarded(&ctx->engine.err, [&] { *out = ctx->engine.capture_snapshot(ctx->slot, ctx->host_mcs); });
}

int32_t cpm_batch_snapshot_capture(cpm_batch *b, int32_t slot, cpm_snapshot **out) {
    if (!b || !out) return CPM_E_INVALID;
    *out = nullptr;
    return guarded(&b->engine.err, [&] {
        b->engine.check_slot(slot);
        *out = b->engine.capture_snapshot(slot, b->engine.read_mcs(slot));
    });
}

int32_t cpm_snapshot_restore(cpm_ctx *ctx, const cpm_snapshot *snapshot) {
    if (!ctx || !snapshot) return CPM_E_INVALID;
    return guarded(&ctx->engine.err, [&] {
        ctx->engine.restore_snapshot(ctx->slot, *snapshot);
        ctx->engine.set_dyn(ctx->slot, snapshot->mcs, snapshot->mcs, 1, 1 | 2 | 8);
        ctx->host_mcs = snapshot->mcs;
    });
}

int32_t cpm_batch_snapshot_restore(cpm_batch *b, int32_t slot, const cpm_snapshot *snapshot) {
    if (!b || !snapshot) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] {
        b->engine.restore_snapshot(slot, *snapshot);
        b->engine.set_dyn(slot, snapshot->mcs, snapshot->mcs, 1, 1 | 2 | 8);
    });
}

int32_t cpm_snapshot_download(const cpm_snapshot *sn, uint8_t *tags, uint32_t *ids, int32_t *finite_volumes,
                              int32_t *medium_volumes, int64_t *boundary_measures, uint8_t *active,
                              uint32_t *cell_types, double *target_volume, double *volume_strength,
                              int64_t *target_boundary, double *boundary_strength) {
    if (!sn) {
        g_create_error = "null snapshot";
        return CPM_E_INVALID;
    }
    return guarded(&g_create_error, [&] {
        const size_t cap = sn->capacity;
        if (tags || ids) {
            std::vector<int32_t> words(sn->n);
            hip_check(hipMemcpy(words.data(), sn->d_lattice, sn->n * 4, hipMemcpyDeviceToHost), "snapshot download");
            unpack_state(words.data(), sn->n, nullptr, 0, tags, ids, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr, nullptr);
        }
        if (finite_volumes || boundary_measures || active || cell_types || target_volume || volume_strength ||
            target_boundary || boundary_strength) {
            std::vector<cpm::CellRec> cells(cap + 1);
            hip_check(hipMemcpy(cells.data(), sn->d_cells, (cap + 1) * sizeof(cpm::CellRec), hipMemcpyDeviceToHost),
                      "snapshot download");
            unpack_state(nullptr, 0, cells.data(), cap, nullptr, nullptr, finite_volumes, boundary_measures, active,
                         cell_types, target_volume, volume_strength, target_boundary, boundary_strength);
        }
        if (medium_volumes && sn->n_media > 0)
            hip_check(hipMemcpy(medium_volumes, sn->d_mvol, (size_t)sn->n_media * 4, hipMemcpyDeviceToHost),
                      "snapshot download");
    });
}

int32_t cpm_snapshot_download_mechanics(const cpm_snapshot *sn, float *volume_pressure, float *surface_tension) {
    if (!sn) {
        g_create_error = "null snapshot";
        return CPM_E_INVALID;
    }
    return guarded(&g_create_error, [&] {
        const size_t cap = sn->capacity;
        if (volume_pressure)
            hip_check(hipMemcpy2D(volume_pressure, sizeof(float), &sn->d_cells[1].qp, sizeof(cpm::CellRec), sizeof(float),
                                  cap, hipMemcpyDeviceToHost), "snapshot mechanics");
        if (surface_tension)
            hip_check(hipMemcpy2D(surface_tension, sizeof(float), &sn->d_cells[1].qt, sizeof(cpm::CellRec), sizeof(float),
                                  cap, hipMemcpyDeviceToHost), "snapshot mechanics");
    });
}

int32_t cpm_snapshot_info(const cpm_snapshot *sn, uint64_t *mcs, uint64_t *seed, uint64_t *bytes) {
    if (!sn) return CPM_E_INVALID;
    if (mcs) *mcs = sn->mcs;
    if (seed) *seed = sn->seed;
    if (bytes) *bytes = sn->bytes;
    return CPM_OK;
}

void cpm_snapshot_destroy(cpm_snapshot *sn) { delete sn; }

int32_t cpm_batch_step(cpm_batch *b, uint64_t n) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { b->engine.step(n); });
}

int32_t cpm_batch_progress(cpm_batch *b, uint64_t *out, int32_t capacity) {
    if (!b || !out) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] {
        Engine &e = b->engine;
        std::vector<cpm::RepDyn> y((size_t)e.slot_cap);
        hip_check(hipStreamSynchronize(e.stream), "sync");
        if (e.slot_cap)
            hip_check(hipMemcpy(y.data(), e.d_dyn, (size_t)e.slot_cap * sizeof(cpm::RepDyn),
                                hipMemcpyDeviceToHost), "progress");
        e.metrics[1]++;
        e.metrics[3]++;
        for (int s = 0; s < capacity; ++s)
            out[s] = (s < e.slot_cap && e.reps[s]) ? y[s].mcs : UINT64_MAX;
    });
}

int32_t cpm_batch_capacity(const cpm_batch *b) { return b ? b->engine.slot_cap : 0; }

int32_t cpm_batch_report(cpm_batch *b, int32_t slot, uint64_t out[12]) {
    if (!b || !out) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { b->engine.report(slot, out); });
}

int32_t cpm_batch_sync(cpm_batch *b) {
    if (!b) return CPM_E_INVALID;
    return guarded(&b->engine.err, [&] { b->engine.sync(); });
}

int32_t cpm_batch_set_graph_chunk(cpm_batch *b, uint32_t chunk) {
    if (!b) return CPM_E_INVALID;
    b->engine.graph_chunk = chunk;
    b->engine.drop_graph();
    return CPM_OK;
}

int32_t cpm_batch_bytes(const cpm_batch *b, uint64_t *bytes) {
    if (!b || !bytes) return CPM_E_INVALID;
    *bytes = b->engine.bytes();
    return CPM_OK;
}

}  // extern "C"
