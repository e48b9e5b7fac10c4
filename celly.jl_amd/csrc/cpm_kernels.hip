// HIP kernels of the MI355X CPM engine (gfx950).
//
// One colour pass (checkerboard) or lottery round is two launches:
//   propose  : LDS-staged lattice tile + halo (or, for 2D periodic von Neumann
//              checkerboards with a 2 x 2 colour pattern, rows read straight into
//              registers); schedule filter (colour or lottery local-maximum); interior
//              sites whose proposal neighbours all share the recipient's owner are
//              settled as same-owner no-ops without RNG; the rest are compacted into an
//              LDS work list (wave ballot) and get direction / priority Philox draws;
//              actionable attempts publish a 64-bit claim key (priority << 32 |
//              ~recipient) to both cells with one atomic max and are staged in a
//              per-tile segment.
//   commit   : per-tile segment; an attempt wins iff it holds the key of both of its
//              cells (equivalent to the reference's max-priority then min-identity
//              claims, checkerboard.jl:281-369); winners evaluate the energy change
//              against the pass snapshot, draw acceptance and commit at once (winners
//              touch disjoint cells and non-adjacent sites, so evaluation and commit
//              need no separation).
#include <hip/hip_runtime.h>

#include "cpm_device.h"

#include <hip/hip_ext.h>

// Kernel-exact timing: an event pair set by cpm_internal_time_next is attached to the next
// launch made here (hipExtLaunchKernelGGL), so the pair brackets that kernel's own execution
// on its stream rather than the queue gaps around it.
static thread_local hipEvent_t t_time_start = nullptr, t_time_stop = nullptr;

#define CPM_LAUNCH(K, GRID, BLOCK, SHMEM, STREAM, ...)                                               \
    do {                                                                                            \
        if (t_time_start) {                                                                         \
            hipExtLaunchKernelGGL(K, GRID, BLOCK, SHMEM, STREAM, t_time_start, t_time_stop, 0u, __VA_ARGS__); \
            t_time_start = t_time_stop = nullptr;                                                   \
        } else {                                                                                    \
            hipLaunchKernelGGL(K, GRID, BLOCK, SHMEM, STREAM, __VA_ARGS__);                         \
        }                                                                                           \
    } while (0)

extern "C" void cpm_internal_time_next(hipEvent_t start, hipEvent_t stop) {
    t_time_start = start;
    t_time_stop = stop;
}

namespace cpm {

template <int DIM, int H, int TX, int TY, int TZ>
struct Geom {
    static constexpr int PX = 4;  // x padding (>= halo) keeping LDS rows 16-byte aligned
    static constexpr int HZ = DIM == 3 ? H : 0;
    static constexpr int LX = TX + 2 * PX;
    static constexpr int LY = TY + 2 * H;
    static constexpr int LZ = TZ + 2 * HZ;
    static constexpr int LN = LX * LY * LZ;
    static constexpr int TN = TX * TY * TZ;
    static constexpr int LV = LX / 4;           // int4 per LDS row
    static constexpr int NV = LV * LY * LZ;     // int4 per tile
    static_assert(DIM == 3 || TZ == 1, "2D tiles have TZ == 1");
    static_assert(TX % 4 == 0 && H <= PX, "tile rows must be whole int4");
    static_assert(LN < 65536, "LDS positions must fit 16 bits");
    static __device__ __forceinline__ int pos(int ix, int iy, int iz) {
        return (ix + PX) + LX * ((iy + H) + LY * (iz + HZ));
    }
};

__device__ __forceinline__ int wrap_axis(int g, int n) {
    if (g < 0) g += n;
    if (g >= n) g -= n;
    if (g < 0 || g >= n) {
        g %= n;
        if (g < 0) g += n;
    }
    return g;
}

// Realize one lattice position for contact/surface roles and the proposal role at
// once (lib/CorePotts/src/spatial/cartesian.jl:445-525). Returns the owner word seen
// by contact/surface relations (0 for absent or invalid) and whether the position is
// a mutable site (the only donors a proposal relation realizes).
struct Realized {
    int32_t word;   // owner seen by contact/surface relations, 0 when absent or invalid
    uint32_t site;  // 1-based site when the position is inside the domain, else 0
    uint8_t mut;    // mutable site (the only donors a proposal relation realizes)
};

template <class D>
__device__ inline Realized realize_position(const D &d, int gx, int gy, int gz) {
    bool closed = false, fixed = false, invalid = false;
    int32_t fixed_word = 0;
    auto axis = [&](int a, int &v) {
        const int n = d.dims[a];
        int face;
        if (v < 0) face = 0;
        else if (v >= n) face = 1;
        else return;
        const int kind = d.face_kind[a][face];
        if (kind == CPM_FACE_PERIODIC) {
            v = wrap_axis(v, n);
        } else if (kind == CPM_FACE_CLOSED) {
            invalid |= fixed;
            closed = true;
        } else {
            const int32_t w = d.face_word[a][face];
            const bool incompatible = fixed && fixed_word != w;
            invalid |= closed | incompatible;
            fixed = true;
            fixed_word = w;
        }
    };
    axis(0, gx);
    axis(1, gy);
    if (d.ndim == 3) axis(2, gz);
    Realized r{0, 0, 0};
    if (invalid || closed) return r;
    if (fixed) {
        r.word = fixed_word;
        return r;
    }
    const uint32_t lin = (uint32_t)gx + (uint32_t)d.dims[0] * ((uint32_t)gy + (uint32_t)d.dims[1] * (uint32_t)gz);
    r.site = lin + 1;
    r.mut = d.mask ? gp(d.mask)[lin] : (uint8_t)1;
    r.word = gp(d.lattice)[lin];
    return r;
}

template <int DIM, int H, int TX, int TY, int TZ, int NT = kBlock>
__device__ inline void load_tile(const CPM_CONST RepDev &d, int x0, int y0, int z0, int32_t *s_own,
                                 uint8_t *s_mut, uint32_t *s_sid) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const bool inner_yz = y0 >= H && y0 + TY + H <= Y && (DIM == 2 || (z0 >= G::HZ && z0 + TZ + G::HZ <= Z));
    const CPM_GLOBAL int32_t *lat = gp(d.lattice);
    if (d.simple && (X & 3) == 0 && x0 >= G::PX && x0 + TX + G::PX <= X && inner_yz) {
        // interior tile of a periodic, obstacle-free lattice: 16-byte loads, all in flight
        constexpr int NL = (G::NV + NT - 1) / NT;
        v4i v[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const int jj = (int)threadIdx.x + j * NT;
            if (jj < G::NV) {
                const int row = jj / G::LV, c4 = jj % G::LV;
                const int gx = x0 - G::PX + 4 * c4, gy = y0 - H + row % G::LY, gz = z0 - G::HZ + row / G::LY;
                v[j] = *(const CPM_GLOBAL v4i *)(lat + (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz));
            }
        }
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const int jj = (int)threadIdx.x + j * NT;
            if (jj < G::NV) {
                const int row = jj / G::LV, c4 = jj % G::LV;
                *reinterpret_cast<v4i *>(s_own + row * G::LX + 4 * c4) = v[j];
                if (s_sid) {
                    const int gx = x0 - G::PX + 4 * c4, gy = y0 - H + row % G::LY, gz = z0 - G::HZ + row / G::LY;
                    const uint32_t lin = (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz);
                    for (int q = 0; q < 4; ++q) s_sid[row * G::LX + 4 * c4 + q] = lin + 1 + q;
                }
            }
        }
        if (s_sid)  // lottery kernels read the mutable flags of every position
            for (int i = threadIdx.x; i < G::LN; i += NT) s_mut[i] = 1;
        return;
    }
    constexpr int NL = (G::LN + NT - 1) / NT;
    const bool interior = x0 >= G::PX && x0 + TX + G::PX <= X && inner_yz;
    const CPM_GLOBAL uint8_t *mask = d.mask ? gp(d.mask) : nullptr;
    int32_t w[NL];
    // every load of the tile is issued before the first LDS store
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const int i = (int)threadIdx.x + j * NT;
        w[j] = 0;
        if (i < G::LN) {
            const int lx = i % G::LX;
            const int ly = (i / G::LX) % G::LY;
            const int lz = i / (G::LX * G::LY);
            int gx = x0 - G::PX + lx, gy = y0 - H + ly, gz = z0 - G::HZ + lz;
            if (interior || d.simple) {
                if (!interior) {
                    gx = wrap_axis(gx, X);
                    gy = wrap_axis(gy, Y);
                    if (DIM == 3) gz = wrap_axis(gz, Z);
                }
                const uint32_t lin = (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz);
                w[j] = lat[lin];
                if (mask || s_sid) {
                    s_mut[i] = mask ? mask[lin] : (uint8_t)1;
                    if (s_sid) s_sid[i] = lin + 1;
                }
            } else {
                const Realized rz = realize_position(d, gx, gy, gz);
                w[j] = rz.word;
                s_mut[i] = rz.mut;
                if (s_sid) s_sid[i] = rz.site;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const int i = (int)threadIdx.x + j * NT;
        if (i < G::LN) s_own[i] = w[j];
    }
}

// Append lanes with `flag` set to an LDS list of 16-bit tile indices.
__device__ __forceinline__ void ballot_push(bool flag, uint16_t value, uint32_t *counter,
                                            uint16_t *list) {
    const uint64_t m = __ballot(flag);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (flag) list[base + __popcll(m & __lanemask_lt())] = value;
}

// Block-wide sum of four counters (one LDS round).
__device__ __forceinline__ void block_sum4(uint32_t v[4], uint32_t (*s_red)[kBlock / 64]) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) s_red[k][wave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += s_red[k][q];
        v[k] = t;
    }
}

// ---------------------------------------------------------------------------
// Energy terms (scientific_components.jl, scientific_trackers.jl)
// ---------------------------------------------------------------------------

// A bounded-draw constant read from the descriptor in the constant address space.
__device__ __forceinline__ Bound const_bound(const CPM_CONST Bound &b) {
    return Bound{b.threshold, FastDiv{b.div.d, b.div.m}};
}

// Cell type or declared medium type, clamped into 1..n_types (the host rejects models
// and states that could reach an undeclared type).
template <class D>
__device__ __forceinline__ uint32_t owner_type(const D &d, int32_t w) {
    uint32_t t = 1;
    if (w > 0) {
        t = word_type(w);
    } else {
        for (int i = 0; i < d.n_table; ++i)
            if (d.table_word[i] == w) t = d.table_type[i];
    }
    return min(max(t, 1u), (uint32_t)d.n_types);
}

template <class D>
__device__ __forceinline__ int medium_slot(const D &d, int32_t w) {
    for (int i = 0; i < d.n_media; ++i)
        if (d.medium_word[i] == w) return i;
    return 0;
}

// scientific_components.jl:90-95
__device__ __forceinline__ float vol_change_f32(float s, float t, int32_t v, int32_t c) {
    const float d0 = __fsub_rn((float)v, t);
    const float old = __fmul_rn(s, __fmul_rn(d0, d0));
    const int32_t u = v + c;
    if (u == 0) return -old;
    const float d1 = __fsub_rn((float)u, t);
    return __fsub_rn(__fmul_rn(s, __fmul_rn(d1, d1)), old);
}
__device__ __forceinline__ double vol_change_f64(double s, double t, int32_t v, int32_t c) {
    const double d0 = __dsub_rn((double)v, t);
    const double old = __dmul_rn(s, __dmul_rn(d0, d0));
    const int32_t u = v + c;
    if (u == 0) return -old;
    const double d1 = __dsub_rn((double)u, t);
    return __dsub_rn(__dmul_rn(s, __dmul_rn(d1, d1)), old);
}
// scientific_components.jl:540-545
__device__ __forceinline__ float bnd_change_f32(float s, float t, float m, int32_t v, float bc,
                                                int32_t vc) {
    const float d0 = __fsub_rn(m, t);
    const float old = __fmul_rn(s, __fmul_rn(d0, d0));
    if (vc == -1 && v == 1) return -old;
    const float d1 = __fsub_rn(__fadd_rn(m, bc), t);
    return __fsub_rn(__fmul_rn(s, __fmul_rn(d1, d1)), old);
}
__device__ __forceinline__ double bnd_change_f64(double s, double t, double m, int32_t v,
                                                 double bc, int32_t vc) {
    const double d0 = __dsub_rn(m, t);
    const double old = __dmul_rn(s, __dmul_rn(d0, d0));
    if (vc == -1 && v == 1) return -old;
    const double d1 = __dsub_rn(__dadd_rn(m, bc), t);
    return __dsub_rn(__dmul_rn(s, __dmul_rn(d1, d1)), old);
}

// ---------------------------------------------------------------------------
// Propose side
// ---------------------------------------------------------------------------

// Claims of one propose tile are first combined in a workgroup table (cell -> maximum key;
// open addressing with linear probing, a full neighbourhood falls back to the global atomic),
// then flushed with one global atomicMax per distinct cell: the maximum over a cell's keys is
// the same, with fewer atomics at the memory side. One 64-bit key per cell replaces the
// max-priority and min-identity claims (checkerboard.jl:281-335).
constexpr int kClaimBits = 7;
constexpr int kClaimSlots = 1 << kClaimBits;
constexpr int kClaimProbes = 2;

__device__ __forceinline__ void claim_local(uint32_t *s_ckey, unsigned long long *s_cval, unsigned long long *claims,
                                            int32_t cell, unsigned long long key) {
    uint32_t slot = ((uint32_t)cell * 0x9E3779B1u) >> (32 - kClaimBits);
    for (int probe = 0; probe < kClaimProbes; ++probe) {
        const uint32_t old = atomicCAS(&s_ckey[slot], 0u, (uint32_t)cell);
        if (old == 0u || old == (uint32_t)cell) {
            atomicMax(&s_cval[slot], key);
            return;
        }
        slot = (slot + 1) & (kClaimSlots - 1);
    }
    atomicMax(&claims[(uint32_t)cell], key);
}


// Appends the lanes with `flag` set to a wave-private LDS list. `count` is the wave's
// running length (uniform across the wave), so no atomics are needed.
__device__ __forceinline__ void wave_append(bool flag, uint16_t value, uint32_t &count, uint16_t *list) {
    const uint64_t m = __ballot(flag);
    if (flag) list[count + __popcll(m & __lanemask_lt())] = value;
    count += (uint32_t)__popcll(m);
}

// Entry e of the concatenation of the kWaves wave-private lists (each of capacity Q) whose
// lengths have exclusive prefix sums p1 <= p2 <= p3.
template <int Q>
__device__ __forceinline__ uint32_t list_slot(uint32_t e, uint32_t p1, uint32_t p2, uint32_t p3) {
    const uint32_t sg = (uint32_t)(e >= p1) + (uint32_t)(e >= p2) + (uint32_t)(e >= p3);
    const uint32_t base = sg == 0 ? 0u : sg == 1 ? p1 : sg == 2 ? p2 : p3;
    return sg * (uint32_t)Q + (e - base);
}

// Four per-tile counters packed as two 16-bit pairs (a tile holds at most 2048 sites),
// summed over the workgroup into s_sum (zeroed before the first barrier).
__device__ __forceinline__ void block_add_packed(uint32_t lo, uint32_t hi, uint32_t *s_sum) {
    lo = __ockl_wfred_add_u32(lo);
    hi = __ockl_wfred_add_u32(hi);
    if (__lane_id() == 0) {
        atomicAdd(&s_sum[0], lo);
        atomicAdd(&s_sum[1], hi);
    }
}

// True when every proposal neighbour of pos is a mutable site owned by `own`: any
// direction draw would then yield a same-owner no-op. KF > 0: the relation has KF
// directions held in registers (po); KF == 0: K directions in LDS (s_ploff).
template <int KF, bool SIMPLE>
__device__ __forceinline__ bool interior_site(const int32_t *s_own, const uint8_t *s_mut, int pos,
                                              int32_t own, const int32_t *po, int K,
                                              const int32_t *s_ploff) {
    bool interior = true;
    if constexpr (KF > 0) {
#pragma unroll
        for (int k = 0; k < KF; ++k) {
            const int q = pos + po[k];
            interior &= s_own[q] == own;
            if (!SIMPLE) interior &= s_mut[q] != 0;
        }
    } else {
        for (int k = 0; k < K; ++k) {
            const int q = pos + s_ploff[k];
            interior &= s_own[q] == own;
            if (!SIMPLE) interior &= s_mut[q] != 0;
        }
    }
    return interior;
}

// Offset of proposal direction dir (0-based), from the LDS copy of the offsets (a
// register select chain would be turned into an indexed scratch load).
template <int KF>
__device__ __forceinline__ int direction_offset(uint32_t dir, const int32_t *, const int32_t *s_ploff) {
    return s_ploff[dir];
}

// ---------------------------------------------------------------------------
// Checkerboard colour pass, propose + claim (checkerboard.jl:244-304)
// ---------------------------------------------------------------------------
template <int DIM, int H, int TX, int TY, int TZ, int KF, int NT>
__global__ __launch_bounds__(NT) void cb_propose_kernel(PassArgs a) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    constexpr int NW = NT / 64;        // waves per workgroup
    constexpr int Q = G::TN / NW;      // capacity of a wave-private list
    static_assert(G::TN <= 4096, "packed 16-bit tile counters");
    __shared__ __align__(16) int32_t s_own[G::LN];
    __shared__ uint8_t s_mut[G::LN];
    __shared__ uint16_t s_work[G::TN];
    __shared__ uint16_t s_act[G::TN];
    __shared__ uint8_t s_dir[G::TN];
    __shared__ int32_t s_ploff[KF > 0 ? KF : CPM_MAX_DIRS];
    __shared__ uint32_t s_cls[64];
    __shared__ uint32_t s_len[2][NW];
    __shared__ uint32_t s_sum[2];
    __shared__ uint32_t s_ckey[kClaimSlots];
    __shared__ unsigned long long s_cval[kClaimSlots];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const TileInfo ti = a.tiles[tile];
    const uint32_t r = ti.rep;
    // descriptor and MCS state are not written by this kernel (constant address space:
    // scalar loads that stay valid across the kernel's own stores)
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)a.reps)[r];
    const CPM_CONST McsPlan &y = ((const CPM_CONST McsPlan *)a.plan)[(size_t)r * kPlanMcs + a.plan_k];
    if (!y.active || a.pass >= (uint32_t)d.rounds) {
        if (threadIdx.x < kWaves) a.seg_count[tile * kWaves + threadIdx.x] = kIdleTile;
        return;
    }
    const uint64_t mcs = y.mcs;
    const uint32_t colour = y.order[a.pass];

    const int x0 = ti.x0, y0 = ti.y0, z0 = ti.z0;
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const int K = d.prop.n;
    const bool simple = d.simple;
    const bool fast = d.pat_fast;
    const int wave = (int)(threadIdx.x >> 6);
    // colourings without a fast pattern: this tile's sites of the colour, from the lists
    uint32_t lbeg = 0, lcnt = 0;
    if (!fast) {
        const CPM_GLOBAL uint32_t *off = gp(d.colour_off) + (size_t)(tile - d.tile_base) * (uint32_t)(d.rounds + 1);
        lbeg = off[colour - 1];
        lcnt = off[colour] - lbeg;
        if (lcnt == 0) {
            // nothing of this colour here: no attempts, no tile load
            if (threadIdx.x < kWaves) a.seg_count[tile * kWaves + threadIdx.x] = 0;
            if (threadIdx.x == 0) {
                uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
                slot[0] = slot[1] = slot[2] = slot[3] = 0;
            }
            return;
        }
    }

    if (threadIdx.x < 2) s_sum[threadIdx.x] = 0;
    for (int i = (int)threadIdx.x; i < kClaimSlots; i += NT) {
        s_ckey[i] = 0u;
        s_cval[i] = 0ull;
    }
    if constexpr (KF == 0) {
        for (int k = threadIdx.x; k < K; k += NT)
            s_ploff[k] = d.prop.off[k][0] + G::LX * (d.prop.off[k][1] + G::LY * d.prop.off[k][2]);
    } else {
        if (threadIdx.x < KF) s_ploff[threadIdx.x] = d.prop_lds[threadIdx.x];
    }
    int32_t po[KF > 0 ? KF : 1];
#pragma unroll
    for (int k = 0; k < (KF > 0 ? KF : 1); ++k) po[k] = d.prop_lds[k];
    int ccount = 0;
    if (fast) {
        // residue classes of this colour: packed residue coordinates
        ccount = d.col_count[colour];
        if ((int)threadIdx.x < ccount) {
            const int rr = d.col_res[d.col_first[colour] + threadIdx.x];
            s_cls[threadIdx.x] = (uint32_t)d.pat_coord[rr][0] | ((uint32_t)d.pat_coord[rr][1] << 8) |
                                 ((uint32_t)d.pat_coord[rr][2] << 16);
        }
    }
    load_tile<DIM, H, TX, TY, TZ, NT>(d, x0, y0, z0, s_own, s_mut, nullptr);
    __syncthreads();

    // 1. colour sites; interior sites (every proposal neighbour a mutable site of the
    //    same owner) are same-owner no-ops whatever the direction draw
    uint32_t c01 = 0, c23 = 0;  // candidates | same owner << 16, boundary null | realized << 16
    uint32_t nw = 0;
    uint16_t *my_work = s_work + wave * Q;
    auto visit = [&](int i, int ix, int iy, int iz) __attribute__((always_inline)) {
        bool need = false;
        if (i >= 0) {
            const int pos = G::pos(ix, iy, iz);
            if (simple || s_mut[pos]) {
                c01 += 1u;
                const int32_t own = s_own[pos];
                const bool interior = simple ? interior_site<KF, true>(s_own, s_mut, pos, own, po, K, s_ploff)
                                             : interior_site<KF, false>(s_own, s_mut, pos, own, po, K, s_ploff);
                if (interior) c01 += 1u << 16;
                else need = true;
            }
        }
        wave_append(need, (uint16_t)(i < 0 ? 0 : i), nw, my_work);
    };
    if (fast) {
        // the colour's residue classes as one flat index space: class k = j >> shift
        const int shift = d.cls_shift, sx = d.cls_sx, sy = d.cls_sy;
        const int px = d.pat[0], py = d.pat[1], pz = d.pat[2];
        const int total = ccount << shift;
        for (int j0 = 0; j0 < total; j0 += NT) {
            const int j = j0 + (int)threadIdx.x;
            int i = -1, ix = 0, iy = 0, iz = 0;
            if (j < total) {
                const uint32_t c = s_cls[j >> shift];
                const int w = j & ((1 << shift) - 1);
                const int kx = w & ((1 << sx) - 1);
                const int ky = (w >> sx) & ((1 << sy) - 1);
                const int kz = w >> (sx + sy);
                ix = (int)(c & 0xff) + __mul24(px, kx);
                iy = (int)((c >> 8) & 0xff) + __mul24(py, ky);
                iz = (int)(c >> 16) + __mul24(pz, kz);
                if (x0 + ix < X && y0 + iy < Y && z0 + iz < Z) i = ix + TX * (iy + TY * iz);
            }
            visit(i, ix, iy, iz);
        }
    } else {
        // the tile's sites of this colour (ascending tile positions, all inside the lattice)
        const CPM_GLOBAL uint16_t *list = gp(d.colour_list) + lbeg;
        for (uint32_t j0 = 0; j0 < lcnt; j0 += NT) {
            const uint32_t j = j0 + threadIdx.x;
            int i = -1, ix = 0, iy = 0, iz = 0;
            if (j < lcnt) {
                i = list[j];
                ix = i % TX;
                iy = (i / TX) % TY;
                iz = i / (TX * TY);
            }
            visit(i, ix, iy, iz);
        }
    }
    if (__lane_id() == 0) s_len[0][wave] = nw;
    __syncthreads();

    // 2. direction draws of the boundary sites (ProposalDirectionStream), attempt outcome
    const uint32_t p1 = s_len[0][0], p2 = p1 + (NW > 1 ? s_len[0][NW > 1 ? 1 : 0] : 0u);
    const uint32_t p3 = p2 + (NW > 2 ? s_len[0][NW > 2 ? 2 : 0] : 0u);
    const uint32_t nwork = p3 + (NW > 3 ? s_len[0][NW > 3 ? 3 : 0] : 0u);
    uint32_t na = 0;
    for (uint32_t e0 = 0; e0 < nwork; e0 += NT) {
        const uint32_t e = e0 + threadIdx.x;
        bool act = false;
        int i = 0;
        uint32_t dir = 0;
        if (e < nwork) {
            i = s_work[list_slot<Q>(e, p1, p2, p3)];
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int pos = G::pos(ix, iy, iz);
            const uint32_t site = 1u + (uint32_t)(x0 + ix) +
                                  (uint32_t)X * ((uint32_t)(y0 + iy) + (uint32_t)Y * (uint32_t)(z0 + iz));
            dir = bounded_draw(d.key_lo, d.key_hi, kStreamDirection, mcs, 0, 0, kSiteEntity, site, 0,
                               const_bound(d.prop_bound));
            const int q = pos + direction_offset<KF>(dir, po, s_ploff);
            if (!simple && !s_mut[q]) {
                c23 += 1u;
            } else if (s_own[q] == s_own[pos]) {
                c01 += 1u << 16;
            } else {
                c23 += 1u << 16;
                act = true;
            }
        }
        const uint64_t m = __ballot(act);
        if (act) {
            const uint32_t at = wave * Q + na + __popcll(m & __lanemask_lt());
            s_act[at] = (uint16_t)i;
            s_dir[at] = (uint8_t)dir;
        }
        na += (uint32_t)__popcll(m);
    }
    block_add_packed(c01, c23, s_sum);
    if (__lane_id() == 0) s_len[1][wave] = na;
    __syncthreads();

    // 3. priority draws and claims of the actionable attempts; the tile segment keeps the
    //    wave-private layout of the lists (stretch w = entries [w * Q, w * Q + length w))
    const uint32_t q1 = s_len[1][0], q2 = q1 + (NW > 1 ? s_len[1][NW > 1 ? 1 : 0] : 0u);
    const uint32_t q3 = q2 + (NW > 2 ? s_len[1][NW > 2 ? 2 : 0] : 0u);
    const uint32_t nact = q3 + (NW > 3 ? s_len[1][NW > 3 ? 3 : 0] : 0u);
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims + d.claim_base;
    CPM_GLOBAL v4i *seg = (CPM_GLOBAL v4i *)(a.seg + (size_t)tile * a.seg_cap);
    for (uint32_t e = threadIdx.x; e < nact; e += NT) {
        const uint32_t at = list_slot<Q>(e, q1, q2, q3);
        const int i = s_act[at];
        const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
        const int pos = G::pos(ix, iy, iz);
        const uint32_t site = 1u + (uint32_t)(x0 + ix) +
                              (uint32_t)X * ((uint32_t)(y0 + iy) + (uint32_t)Y * (uint32_t)(z0 + iz));
        const uint32_t prio = rng_words(d.key_lo, d.key_hi, kStreamCheckerboardPriority, mcs, 0, 0,
                                        kSiteEntity, site, 0, 0).w0;
        const int32_t losing = s_own[pos];
        const int32_t gaining = s_own[pos + direction_offset<KF>(s_dir[at], po, s_ploff)];
        // media are never claimed
        const unsigned long long key = ((unsigned long long)prio << 32) | (uint32_t)(~site);
        if (losing > 0) claim_local(s_ckey, s_cval, claims, (int32_t)word_cell(losing), key);
        if (gaining > 0) claim_local(s_ckey, s_cval, claims, (int32_t)word_cell(gaining), key);
        seg[at] = v4i{(int32_t)site, (int32_t)prio, losing, gaining};
    }

    if (threadIdx.x < kWaves)
        a.seg_count[tile * kWaves + threadIdx.x] = (int)threadIdx.x < NW ? s_len[1][threadIdx.x % NW] : 0u;
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        const uint32_t lo = s_sum[0], hi = s_sum[1];
        slot[0] = lo & 0xffffu;
        slot[1] = lo >> 16;
        slot[2] = hi & 0xffffu;
        slot[3] = hi >> 16;
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < kClaimSlots; i += NT) {
        const uint32_t c = s_ckey[i];
        if (c != 0u) atomicMax(&claims[c], s_cval[i]);
    }
}

// ---------------------------------------------------------------------------
// Checkerboard colour pass on a 2D periodic obstacle-free lattice whose proposal relation
// is the von Neumann shell in canonical order (-1,0), (0,-1), (0,1), (1,0) and whose
// colouring is a 2 x 2 pattern, one residue class per colour (checkerboard.jl:244-304).
// No LDS tile and no barrier before the end: every wave reads its rows of the tile
// straight into registers (16-byte loads of the colour row and the rows above and below,
// plus the one x neighbour outside the 16 bytes), settles interior sites, and compacts
// boundary sites and then actionable attempts in wave-private LDS lists so that the
// Philox draws run on full waves. Outputs (claims, segment stretches, counters) are those
// of cb_propose_kernel with NT = 256. An x side of 2 mod 4 (HALF) puts every other row 8
// bytes into a 16-byte group, and the last group of a row holds two sites: its loads are
// 8-byte halves, the upper half of the last group is not read, and the x + 1 neighbour of
// the row's last site is the row's first word.
// ---------------------------------------------------------------------------
template <int TX, int TY>
struct RowGeom {
    static constexpr int LPR = TX / 4;               // lanes per lattice row
    static constexpr int RPI = 64 / LPR;             // rows per wave instruction
    static constexpr int BAND = TY / kWaves;         // tile rows of one wave
    static constexpr int CROWS = BAND / 2;           // rows of the colour in the band
    static constexpr int ITERS = (CROWS + RPI - 1) / RPI;  // row strips of one wave
    static constexpr int CAP = 2 * 64;               // colour sites of one strip
    static constexpr int Q = TX * TY / kWaves;       // segment stretch capacity
    static_assert(TX % 4 == 0 && 64 % LPR == 0 && TY % (2 * kWaves) == 0, "row tile shape");
    static_assert(CAP * ITERS <= Q, "stretch holds every colour site of the wave");
};

// One row strip of a wave in registers: 16 bytes of the colour row and of the rows below
// and above, and the x neighbour outside the 16 bytes (x - 1 for RX == 0, x + 4 else).
struct RowStrip {
    v4i mid, dn, up;
    int32_t side;
    uint32_t site0;  // 1-based site of the first of the 16 bytes
    bool valid;
    bool full;       // all four sites lie in the row (false only for the last group of a HALF row)
};

// Four words at p: one 16-byte load, or (HALF: p is 8-byte aligned) two 8-byte loads, the
// upper one only when the group is whole.
template <bool HALF>
__device__ __forceinline__ v4i row_group(const CPM_GLOBAL int32_t *p, bool full) {
    if constexpr (!HALF) {
        return *(const CPM_GLOBAL v4i *)p;
    } else {
        typedef int32_t v2i __attribute__((ext_vector_type(2)));
        const v2i lo = *(const CPM_GLOBAL v2i *)p;
        v2i hi = v2i{0, 0};
        if (full) hi = *(const CPM_GLOBAL v2i *)(p + 2);
        return v4i{lo.x, lo.y, hi.x, hi.y};
    }
}

template <int RX, int TX, int TY, bool HALF>
__device__ __forceinline__ RowStrip row_load(const CPM_GLOBAL int32_t *lat, int X, int Y, int x0, int y0, int ry,
                                             int wave, int it) {
    using R = RowGeom<TX, TY>;
    const int lane = (int)__lane_id();
    const int k = it * R::RPI + lane / R::LPR;
    RowStrip st;
    const int gx = x0 + 4 * (lane % R::LPR);
    const int gy = y0 + wave * R::BAND + 2 * k + ry;
    // a tile at the upper x or y edge may extend past the lattice
    st.valid = k < R::CROWS && gx < X && gy < Y;
    st.full = !HALF || gx + 4 <= X;
    st.mid = st.dn = st.up = v4i{0, 0, 0, 0};
    st.side = 0;
    st.site0 = 1u + (uint32_t)gx + (uint32_t)gy * (uint32_t)X;
    if (st.valid) {
        const int gym = gy == 0 ? Y - 1 : gy - 1;
        const int gyp = gy + 1 == Y ? 0 : gy + 1;
        const uint32_t row = (uint32_t)gy * (uint32_t)X;
        st.mid = row_group<HALF>(lat + row + (uint32_t)gx, st.full);
        st.dn = row_group<HALF>(lat + (uint32_t)gym * (uint32_t)X + (uint32_t)gx, st.full);
        st.up = row_group<HALF>(lat + (uint32_t)gyp * (uint32_t)X + (uint32_t)gx, st.full);
        const int sx = RX == 0 ? (gx == 0 ? X - 1 : gx - 1) : (gx + 4 >= X ? 0 : gx + 4);
        st.side = lat[row + (uint32_t)sx];
        // the last group of a HALF row: its second site (x = X - 1) has the row's first word at x + 1
        if (HALF && RX == 1 && !st.full) st.mid.z = st.side;
    }
    return st;
}

// Colour sites of one strip (x residues RX and RX + 2 of each 16 bytes): interior sites are
// counted as same-owner no-ops, boundary sites appended to the wave's candidate list.
template <int RX>
__device__ __forceinline__ void row_collect(const RowStrip &st, uint32_t &ncand, uint32_t &nsites,
                                            uint32_t &nsame, uint32_t *s_site, int32_t *s_own,
                                            int32_t (*s_nb)[2 * 64]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = RX + 2 * h;
        const bool valid = h == 0 ? st.valid : st.valid && st.full;
        nsites += (uint32_t)__popcll(__ballot(valid));
        const v4i mid = st.mid, dn = st.dn, up = st.up;
        const int32_t own = p == 0 ? mid.x : p == 1 ? mid.y : p == 2 ? mid.z : mid.w;
        const int32_t left = p == 0 ? st.side : p == 1 ? mid.x : p == 2 ? mid.y : mid.z;
        const int32_t right = p == 0 ? mid.y : p == 1 ? mid.z : p == 2 ? mid.w : st.side;
        const int32_t below = p == 0 ? dn.x : p == 1 ? dn.y : p == 2 ? dn.z : dn.w;
        const int32_t above = p == 0 ? up.x : p == 1 ? up.y : p == 2 ? up.z : up.w;
        const bool interior = own == left && own == below && own == above && own == right;
        const bool boundary = valid && !interior;
        nsame += (uint32_t)__popcll(__ballot(valid && interior));
        const uint64_t m = __ballot(boundary);
        if (boundary) {
            const uint32_t e = ncand + __popcll(m & __lanemask_lt());
            s_site[e] = st.site0 + (uint32_t)p;
            s_own[e] = own;
            s_nb[0][e] = left;
            s_nb[1][e] = below;
            s_nb[2][e] = above;
            s_nb[3][e] = right;
        }
        ncand += (uint32_t)__popcll(m);
    }
}

// LDS of one row-kernel tile: each wave's boundary sites of the current strip (site, owner,
// the four von Neumann neighbour words), then the workgroup's actionable attempts of that
// strip, pooled over the four waves.
struct RowLds {
    uint32_t site[kWaves][2 * 64];
    int32_t own[kWaves][2 * 64];
    int32_t nb[kWaves][4][2 * 64];
    uint32_t cnt[kWaves];
    uint32_t act_site[kWaves * 2 * 64];
    int32_t act_own[kWaves * 2 * 64];
    int32_t act_gain[kWaves * 2 * 64];
    uint32_t nact[2];  // double-buffered by strip parity
};

// One row-kernel tile. Every wave issues all its strips' loads first. Then strip by strip:
// each wave compacts its boundary sites (interior sites are same-owner no-ops without a draw);
// after a barrier the workgroup draws the directions (ProposalDirectionStream) over the four
// waves' pooled boundary sites in full 64-lane chunks and compacts the actionable attempts;
// after a second barrier it draws their priorities (CheckerboardPriorityStream), claims both
// cells in the LDS table and writes the tile's segment (stretch 0). Pooling over the waves
// fills the Philox chunks: a strip holds about 40 boundary sites and 13 actionable attempts
// per wave, so four waves need 3 + 1 wave-wide Philox calls instead of 4 + 4.
template <int RX, int TX, int TY, bool HALF>
__device__ __forceinline__ void row_tile(const PassArgs &a, const CPM_CONST RepDev &d, const TileInfo &ti,
                                         uint32_t tile, uint64_t mcs, int ry, int wave, RowLds &L, uint32_t *s_ckey,
                                         unsigned long long *s_cval, uint32_t &nsites, uint32_t &nsame,
                                         uint32_t &nact) {
    using R = RowGeom<TX, TY>;
    const CPM_GLOBAL int32_t *lat = gp(d.lattice);
    const int X = d.dims[0], Y = d.dims[1];
    const int lane = (int)__lane_id();
    const uint32_t tid = threadIdx.x;
    RowStrip st[R::ITERS];
#pragma unroll
    for (int it = 0; it < R::ITERS; ++it) st[it] = row_load<RX, TX, TY, HALF>(lat, X, Y, ti.x0, ti.y0, ry, wave, it);
    __syncthreads();  // the claim table and the first strip's counter are clear (loads in flight)
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims + d.claim_base;
    CPM_GLOBAL v4i *seg = (CPM_GLOBAL v4i *)(a.seg + (size_t)tile * a.seg_cap);
#pragma unroll
    for (int it = 0; it < R::ITERS; ++it) {
        uint32_t ncand = 0;
        row_collect<RX>(st[it], ncand, nsites, nsame, L.site[wave], L.own[wave], L.nb[wave]);
        if (lane == 0) L.cnt[wave] = ncand;
        __syncthreads();
        // the next strip's counter: its last readers (the previous strip) passed the barrier above
        if (tid == 0) L.nact[(it + 1) & 1] = 0u;
        const uint32_t p1 = L.cnt[0], p2 = p1 + L.cnt[1], p3 = p2 + L.cnt[2], total = p3 + L.cnt[3];
        // direction draws over the pooled boundary sites; four directions: bounded_uint(4)
        // never rejects ((2^32 - 4) mod 4 == 0)
        for (uint32_t i0 = (uint32_t)wave * 64u; i0 < total; i0 += kBlock) {
            const uint32_t i = i0 + (uint32_t)lane;
            bool act = false;
            uint32_t site = 0;
            int32_t own = 0, gain = 0;
            if (i < total) {
                const uint32_t sg = (uint32_t)(i >= p1) + (uint32_t)(i >= p2) + (uint32_t)(i >= p3);
                const uint32_t e = i - (sg == 0 ? 0u : sg == 1 ? p1 : sg == 2 ? p2 : p3);
                site = L.site[sg][e];
                own = L.own[sg][e];
                const uint32_t dir = rng_words(d.key_lo, d.key_hi, kStreamDirection, mcs, 0, 0, kSiteEntity, site,
                                               0, 0).w0 & 3u;
                gain = L.nb[sg][dir][e];
                act = gain != own;
            }
            const uint64_t valid = __ballot(i < total);
            const uint64_t m = __ballot(act);
            nsame += (uint32_t)__popcll(valid & ~m);
            uint32_t base = 0;
            if (lane == 0 && m) base = atomicAdd(&L.nact[it & 1], (uint32_t)__popcll(m));
            base = __shfl(base, 0);
            if (act) {
                const uint32_t at = base + __popcll(m & __lanemask_lt());
                L.act_site[at] = site;
                L.act_own[at] = own;
                L.act_gain[at] = gain;
            }
        }
        __syncthreads();
        const uint32_t na = L.nact[it & 1];
        // priority draws, claims and the segment entries
        for (uint32_t e = tid; e < na; e += kBlock) {
            const uint32_t site = L.act_site[e];
            const int32_t losing = L.act_own[e];
            const int32_t gaining = L.act_gain[e];
            const uint32_t prio = rng_words(d.key_lo, d.key_hi, kStreamCheckerboardPriority, mcs, 0, 0,
                                            kSiteEntity, site, 0, 0).w0;
            const unsigned long long key = ((unsigned long long)prio << 32) | (uint32_t)(~site);
            if (losing > 0) claim_local(s_ckey, s_cval, claims, (int32_t)word_cell(losing), key);
            if (gaining > 0) claim_local(s_ckey, s_cval, claims, (int32_t)word_cell(gaining), key);
            seg[nact + e] = v4i{(int32_t)site, (int32_t)prio, losing, gaining};
        }
        nact += na;
    }
}

template <int TX, int TY, bool HALF>
__global__ __launch_bounds__(kBlock) void cb_propose_rows_kernel(PassArgs a) {
    __shared__ RowLds L;
    __shared__ uint32_t s_cnt[kWaves][2];
    __shared__ uint32_t s_ckey[kClaimSlots];
    __shared__ unsigned long long s_cval[kClaimSlots];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const TileInfo ti = a.tiles[tile];
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)a.reps)[ti.rep];
    const CPM_CONST McsPlan &y = ((const CPM_CONST McsPlan *)a.plan)[(size_t)ti.rep * kPlanMcs + a.plan_k];
    if (!y.active || a.pass >= (uint32_t)d.rounds) {
        if (threadIdx.x < kWaves) a.seg_count[tile * kWaves + threadIdx.x] = kIdleTile;
        return;
    }
    for (int i = (int)threadIdx.x; i < kClaimSlots; i += kBlock) {
        s_ckey[i] = 0u;
        s_cval[i] = 0ull;
    }
    if (threadIdx.x == 0) L.nact[0] = 0u;
    const uint64_t mcs = y.mcs;
    const uint32_t colour = y.order[a.pass];
    const int res = d.col_res[d.col_first[colour]];
    const int rx = d.pat_coord[res][0], ry = d.pat_coord[res][1];
    const int wave = (int)(threadIdx.x >> 6);
    uint32_t nsites = 0, nsame = 0, nact = 0;
    if (rx == 0)
        row_tile<0, TX, TY, HALF>(a, d, ti, tile, mcs, ry, wave, L, s_ckey, s_cval, nsites, nsame, nact);
    else
        row_tile<1, TX, TY, HALF>(a, d, ti, tile, mcs, ry, wave, L, s_ckey, s_cval, nsites, nsame, nact);
    if (__lane_id() == 0) {
        s_cnt[wave][0] = nsites;
        s_cnt[wave][1] = nsame;
    }
    __syncthreads();
    if (threadIdx.x < kWaves)  // the whole segment is stretch 0
        a.seg_count[tile * kWaves + threadIdx.x] = threadIdx.x == 0 ? nact : 0u;
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        uint32_t t0 = 0, t1 = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            t0 += s_cnt[w][0];
            t1 += s_cnt[w][1];
        }
        slot[0] = t0;    // candidates
        slot[1] = t1;    // same-owner no-ops
        slot[2] = 0;     // boundary no-ops: none on a periodic obstacle-free lattice
        slot[3] = nact;  // realized (actionable) attempts
    }
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims + d.claim_base;
    for (int i = (int)threadIdx.x; i < kClaimSlots; i += kBlock) {
        const uint32_t c = s_ckey[i];
        if (c != 0u) atomicMax(&claims[c], s_cval[i]);
    }
}

// ---------------------------------------------------------------------------
// 3D row propose kernel: 3D periodic obstacle-free checkerboards whose proposal relation is
// the canonical von Neumann shell (6 directions) and whose colouring is a 2 x 2 x 2 pattern
// with one residue class per colour (checkerboard.jl:244-310). A colour's sites lie in every
// other row of every other plane; a wave reads its colour rows and their four neighbour rows
// (y - 1, y + 1, z - 1, z + 1) straight into registers.
// ---------------------------------------------------------------------------
template <int TX, int TY, int TZ>
struct RowGeom3 {
    static constexpr int LPR = TX / 4;                      // lanes per lattice row
    static constexpr int RPI = 64 / LPR;                    // rows per wave instruction
    static constexpr int CRY = TY / 2, CRZ = TZ / 2;        // colour rows along y and z
    static constexpr int CROWS = CRY * CRZ / kWaves;        // colour rows of one wave
    static constexpr int ITERS = (CROWS + RPI - 1) / RPI;   // row strips of one wave
    static constexpr int CAP = 2 * 64;                      // colour sites of one strip
    static constexpr int Q = TX * TY * TZ / kWaves;         // segment stretch capacity
    static_assert(TX % 4 == 0 && 64 % LPR == 0 && TY % 2 == 0 && TZ % 2 == 0 && (CRY * CRZ) % kWaves == 0,
                  "3D row tile shape");
    static_assert(CAP * ITERS <= Q, "stretch holds every colour site of the wave");
};

// One row strip of a wave: 16 bytes of the colour row and of its four neighbour rows, and
// the x neighbour outside the 16 bytes (x - 1 for RX == 0, x + 4 else).
struct RowStrip3 {
    v4i mid, dn, up, bk, fr;
    int32_t side;
    uint32_t site0;  // 1-based site of the first of the 16 bytes
    bool valid;
};

template <int RX, int TX, int TY, int TZ>
__device__ __forceinline__ RowStrip3 row_load3(const CPM_GLOBAL int32_t *lat, int X, int Y, int Z, const TileInfo &ti,
                                               int ry, int rz, int wave, int it) {
    using R = RowGeom3<TX, TY, TZ>;
    const int lane = (int)__lane_id();
    const int k = it * R::RPI + lane / R::LPR;
    RowStrip3 st;
    const int j = wave * R::CROWS + (k < R::CROWS ? k : 0);  // colour row of the tile
    const int gx = ti.x0 + 4 * (lane % R::LPR);
    const int gy = ti.y0 + 2 * (j % R::CRY) + ry;
    const int gz = ti.z0 + 2 * (j / R::CRY) + rz;
    // a tile at an upper edge may extend past the lattice (whose X is a multiple of 4)
    st.valid = k < R::CROWS && gx < X && gy < Y && gz < Z;
    const uint32_t XY = (uint32_t)X * (uint32_t)Y;
    st.mid = st.dn = st.up = st.bk = st.fr = v4i{0, 0, 0, 0};
    st.side = 0;
    st.site0 = 1u + (uint32_t)gx + (uint32_t)gy * (uint32_t)X + (uint32_t)gz * XY;
    if (st.valid) {
        const int gym = gy == 0 ? Y - 1 : gy - 1, gyp = gy + 1 == Y ? 0 : gy + 1;
        const int gzm = gz == 0 ? Z - 1 : gz - 1, gzp = gz + 1 == Z ? 0 : gz + 1;
        const uint32_t plane = (uint32_t)gz * XY, row = plane + (uint32_t)gy * (uint32_t)X;
        st.mid = *(const CPM_GLOBAL v4i *)(lat + row + (uint32_t)gx);
        st.dn = *(const CPM_GLOBAL v4i *)(lat + plane + (uint32_t)gym * (uint32_t)X + (uint32_t)gx);
        st.up = *(const CPM_GLOBAL v4i *)(lat + plane + (uint32_t)gyp * (uint32_t)X + (uint32_t)gx);
        st.bk = *(const CPM_GLOBAL v4i *)(lat + (uint32_t)gzm * XY + (uint32_t)gy * (uint32_t)X + (uint32_t)gx);
        st.fr = *(const CPM_GLOBAL v4i *)(lat + (uint32_t)gzp * XY + (uint32_t)gy * (uint32_t)X + (uint32_t)gx);
        const int sx = RX == 0 ? (gx == 0 ? X - 1 : gx - 1) : (gx + 4 == X ? 0 : gx + 4);
        st.side = lat[row + (uint32_t)sx];
    }
    return st;
}

// Colour sites of one strip: interior sites (all six neighbours share the owner) are
// same-owner no-ops; boundary sites are appended with their neighbours in the canonical
// direction order (-x, -y, -z, +z, +y, +x).
template <int RX>
__device__ __forceinline__ void row_collect3(const RowStrip3 &st, uint32_t &ncand, uint32_t &nsites, uint32_t &nsame,
                                             uint32_t *s_site, int32_t *s_own, int32_t (*s_nb)[2 * 64]) {
    nsites += 2u * (uint32_t)__popcll(__ballot(st.valid));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = RX + 2 * h;
        const v4i mid = st.mid, dn = st.dn, up = st.up, bk = st.bk, fr = st.fr;
        const int32_t own = p == 0 ? mid.x : p == 1 ? mid.y : p == 2 ? mid.z : mid.w;
        const int32_t left = p == 0 ? st.side : p == 1 ? mid.x : p == 2 ? mid.y : mid.z;
        const int32_t right = p == 0 ? mid.y : p == 1 ? mid.z : p == 2 ? mid.w : st.side;
        const int32_t below = p == 0 ? dn.x : p == 1 ? dn.y : p == 2 ? dn.z : dn.w;
        const int32_t above = p == 0 ? up.x : p == 1 ? up.y : p == 2 ? up.z : up.w;
        const int32_t back = p == 0 ? bk.x : p == 1 ? bk.y : p == 2 ? bk.z : bk.w;
        const int32_t front = p == 0 ? fr.x : p == 1 ? fr.y : p == 2 ? fr.z : fr.w;
        const bool interior = own == left && own == below && own == above && own == right && own == back &&
                              own == front;
        const bool boundary = st.valid && !interior;
        nsame += (uint32_t)__popcll(__ballot(st.valid && interior));
        const uint64_t m = __ballot(boundary);
        if (boundary) {
            const uint32_t e = ncand + __popcll(m & __lanemask_lt());
            s_site[e] = st.site0 + (uint32_t)p;
            s_own[e] = own;
            s_nb[0][e] = left;
            s_nb[1][e] = below;
            s_nb[2][e] = back;
            s_nb[3][e] = front;
            s_nb[4][e] = above;
            s_nb[5][e] = right;
        }
        ncand += (uint32_t)__popcll(m);
    }
}

template <int RX, int TX, int TY, int TZ>
__device__ __forceinline__ void row_wave3(const PassArgs &a, const CPM_CONST RepDev &d, const TileInfo &ti,
                                          uint32_t tile, uint64_t mcs, int ry, int rz, int wave, uint32_t *s_site,
                                          int32_t *s_own, int32_t (*s_nb)[2 * 64], uint32_t *s_ckey,
                                          unsigned long long *s_cval, uint32_t &nsites, uint32_t &nsame,
                                          uint32_t &nact) {
    using R = RowGeom3<TX, TY, TZ>;
    const CPM_GLOBAL int32_t *lat = gp(d.lattice);
    const int lane = (int)__lane_id();
    RowStrip3 st[R::ITERS];
#pragma unroll
    for (int it = 0; it < R::ITERS; ++it)
        st[it] = row_load3<RX, TX, TY, TZ>(lat, d.dims[0], d.dims[1], d.dims[2], ti, ry, rz, wave, it);
    __syncthreads();  // the claim table is clear (the strips' loads are in flight)
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims + d.claim_base;
    CPM_GLOBAL v4i *seg = (CPM_GLOBAL v4i *)(a.seg + (size_t)tile * a.seg_cap) + wave * R::Q;
    const Bound dirs = const_bound(d.prop_bound);
#pragma unroll
    for (int it = 0; it < R::ITERS; ++it) {
        uint32_t ncand = 0;
        row_collect3<RX>(st[it], ncand, nsites, nsame, s_site, s_own, s_nb);
        // direction draws (bounded_uint(6) may retry); actionable attempts compacted in place
        uint32_t na = 0;
        for (uint32_t e0 = 0; e0 < ncand; e0 += 64) {
            const uint32_t e = e0 + (uint32_t)lane;
            bool act = false;
            uint32_t site = 0;
            int32_t own = 0, gain = 0;
            if (e < ncand) {
                site = s_site[e];
                own = s_own[e];
                const uint32_t dir =
                    bounded_draw(d.key_lo, d.key_hi, kStreamDirection, mcs, 0, 0, kSiteEntity, site, 0, dirs);
                gain = s_nb[dir][e];
                act = gain != own;
            }
            const uint64_t valid = __ballot(e < ncand);
            const uint64_t m = __ballot(act);
            nsame += (uint32_t)__popcll(valid & ~m);
            if (act) {
                const uint32_t at = na + __popcll(m & __lanemask_lt());
                s_site[at] = site;
                s_own[at] = own;
                s_nb[0][at] = gain;
            }
            na += (uint32_t)__popcll(m);
        }
        // priority draws and claims (CheckerboardPriorityStream); stretch `wave` of the segment
        for (uint32_t e = (uint32_t)lane; e < na; e += 64) {
            const uint32_t site = s_site[e];
            const int32_t losing = s_own[e];
            const int32_t gaining = s_nb[0][e];
            const uint32_t prio = rng_words(d.key_lo, d.key_hi, kStreamCheckerboardPriority, mcs, 0, 0,
                                            kSiteEntity, site, 0, 0).w0;
            const unsigned long long key = ((unsigned long long)prio << 32) | (uint32_t)(~site);
            if (losing > 0) claim_local(s_ckey, s_cval, claims, (int32_t)word_cell(losing), key);
            if (gaining > 0) claim_local(s_ckey, s_cval, claims, (int32_t)word_cell(gaining), key);
            seg[nact + e] = v4i{(int32_t)site, (int32_t)prio, losing, gaining};
        }
        nact += na;
    }
}

template <int TX, int TY, int TZ>
__global__ __launch_bounds__(kBlock) void cb_propose_rows3_kernel(PassArgs a) {
    __shared__ uint32_t s_site[kWaves][2 * 64];
    __shared__ int32_t s_own[kWaves][2 * 64];
    __shared__ int32_t s_nb[kWaves][6][2 * 64];
    __shared__ uint32_t s_cnt[kWaves][3];
    __shared__ uint32_t s_ckey[kClaimSlots];
    __shared__ unsigned long long s_cval[kClaimSlots];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const TileInfo ti = a.tiles[tile];
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)a.reps)[ti.rep];
    const CPM_CONST McsPlan &y = ((const CPM_CONST McsPlan *)a.plan)[(size_t)ti.rep * kPlanMcs + a.plan_k];
    if (!y.active || a.pass >= (uint32_t)d.rounds) {
        if (threadIdx.x < kWaves) a.seg_count[tile * kWaves + threadIdx.x] = kIdleTile;
        return;
    }
    const uint64_t mcs = y.mcs;
    const uint32_t colour = y.order[a.pass];
    const int res = d.col_res[d.col_first[colour]];
    const int rx = d.pat_coord[res][0], ry = d.pat_coord[res][1], rz = d.pat_coord[res][2];
    const int wave = (int)(threadIdx.x >> 6);
    for (int i = (int)threadIdx.x; i < kClaimSlots; i += kBlock) {
        s_ckey[i] = 0u;
        s_cval[i] = 0ull;
    }
    uint32_t nsites = 0, nsame = 0, nact = 0;
    if (rx == 0)
        row_wave3<0, TX, TY, TZ>(a, d, ti, tile, mcs, ry, rz, wave, s_site[wave], s_own[wave], s_nb[wave], s_ckey,
                                 s_cval, nsites, nsame, nact);
    else
        row_wave3<1, TX, TY, TZ>(a, d, ti, tile, mcs, ry, rz, wave, s_site[wave], s_own[wave], s_nb[wave], s_ckey,
                                 s_cval, nsites, nsame, nact);
    if (__lane_id() == 0) {
        a.seg_count[tile * kWaves + wave] = nact;
        s_cnt[wave][0] = nsites;
        s_cnt[wave][1] = nsame;
        s_cnt[wave][2] = nact;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        uint32_t t0 = 0, t1 = 0, t3 = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            t0 += s_cnt[w][0];
            t1 += s_cnt[w][1];
            t3 += s_cnt[w][2];
        }
        slot[0] = t0;   // candidates
        slot[1] = t1;   // same-owner no-ops
        slot[2] = 0;    // boundary no-ops: none on a periodic obstacle-free lattice
        slot[3] = t3;   // realized (actionable) attempts
    }
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims + d.claim_base;
    for (int i = (int)threadIdx.x; i < kClaimSlots; i += kBlock) {
        const uint32_t c = s_ckey[i];
        if (c != 0u) atomicMax(&claims[c], s_cval[i]);
    }
}

// ---------------------------------------------------------------------------
// Lottery round, propose + claim (lottery.jl:153-256)
// ---------------------------------------------------------------------------
// PLAIN: periodic, obstacle-free and at least 2 halo + 1 sites along every axis, so every
// conflict neighbour position of a tile site is a mutable site other than itself.
// MOORE (with PLAIN, halo 1): the conflict shell is the radius-1 Moore shell, so the winner
// test is the maximum of the 8 (26) neighbour tickets at compile-time LDS offsets.
template <int DIM, int H, int TX, int TY, int TZ, int KF, int NT, bool PLAIN, bool MOORE = false>
__global__ __launch_bounds__(NT) void lot_propose_kernel(PassArgs a) {
    static_assert(!MOORE || (PLAIN && H == 1), "Moore winner test needs a plain lattice and halo 1");
    using G = Geom<DIM, H, TX, TY, TZ>;
    constexpr int NW = NT / 64;
    constexpr int Q = G::TN / NW;
    static_assert(G::TN <= 4096, "packed 16-bit tile counters");
    __shared__ __align__(16) int32_t s_own[G::LN];
    // plain lattices stage neither mutability nor identity (smaller footprint, more
    // workgroups per CU)
    __shared__ uint8_t s_mut[PLAIN ? 4 : G::LN];
    __shared__ uint32_t s_sid[PLAIN ? 1 : G::LN];
    __shared__ __align__(16) uint32_t s_tk[G::LN];
    __shared__ uint16_t s_work[G::TN];
    __shared__ int32_t s_ploff[KF > 0 ? KF : CPM_MAX_DIRS];
    __shared__ int32_t s_cloff[CPM_MAX_DIRS];
    __shared__ uint32_t s_len[NW];
    __shared__ uint32_t s_sum[2];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const TileInfo ti = a.tiles[tile];
    const uint32_t r = ti.rep;
    // descriptor and MCS state are not written by this kernel (constant address space:
    // scalar loads that stay valid across the kernel's own stores)
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)a.reps)[r];
    const CPM_CONST McsPlan &y = ((const CPM_CONST McsPlan *)a.plan)[(size_t)r * kPlanMcs + a.plan_k];
    if (!y.active || a.pass >= (uint32_t)d.rounds) {
        if (threadIdx.x < kWaves) a.seg_count[tile * kWaves + threadIdx.x] = kIdleTile;
        return;
    }
    const uint64_t mcs = y.mcs;
    const uint32_t sub = y.order[a.pass] - 1;

    const int x0 = ti.x0, y0 = ti.y0, z0 = ti.z0;
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const int K = d.prop.n;
    const int KC = d.conf.n;
    const int wave = (int)(threadIdx.x >> 6);

    if (threadIdx.x < 2) s_sum[threadIdx.x] = 0;
    if constexpr (KF == 0) {
        for (int k = threadIdx.x; k < K; k += NT)
            s_ploff[k] = d.prop.off[k][0] + G::LX * (d.prop.off[k][1] + G::LY * d.prop.off[k][2]);
    } else {
        if (threadIdx.x < KF) s_ploff[threadIdx.x] = d.prop_lds[threadIdx.x];
    }
    for (int k = threadIdx.x; k < KC; k += NT)
        s_cloff[k] = d.conf.off[k][0] + G::LX * (d.conf.off[k][1] + G::LY * d.conf.off[k][2]);
    int32_t po[KF > 0 ? KF : 1];
#pragma unroll
    for (int k = 0; k < (KF > 0 ? KF : 1); ++k) po[k] = d.prop_lds[k];
    // site id of an LDS position of a plain lattice (every position a mutable site; the
    // mutability and identity arrays are not staged)
    auto sid = [&](int pos) __attribute__((always_inline)) {
        const int lx = pos % G::LX, ly = (pos / G::LX) % G::LY, lz = pos / (G::LX * G::LY);
        int gx = x0 - G::PX + lx, gy = y0 - H + ly, gz = z0 - G::HZ + lz;
        gx += gx < 0 ? X : 0;
        gx -= gx >= X ? X : 0;
        gy += gy < 0 ? Y : 0;
        gy -= gy >= Y ? Y : 0;
        if (DIM == 3) {
            gz += gz < 0 ? Z : 0;
            gz -= gz >= Z ? Z : 0;
        }
        return 1u + (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz);
    };
    load_tile<DIM, H, TX, TY, TZ, NT>(d, x0, y0, z0, s_own, s_mut, PLAIN ? nullptr : s_sid);
    __syncthreads();
    // ticket word 1 of every mutable position (LotteryPriorityStream, subround = round - 1)
    if constexpr (PLAIN) {
        auto ticket = [&](uint32_t site) __attribute__((always_inline)) {
            return rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority, mcs, sub, 0, kSiteEntity, site, 0, 0).w0;
        };
        // The tile itself: on a tile that lies inside the lattice the ids follow from the tile
        // coordinates (shifts, no wrap); a tile reaching past the upper faces wraps them.
        const bool inside = x0 + TX <= X && y0 + TY <= Y && z0 + TZ <= Z;
        for (int i = threadIdx.x; i < G::TN; i += NT) {
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int pos = G::pos(ix, iy, iz);
            const uint32_t site =
                inside ? 1u + (uint32_t)(x0 + ix) + (uint32_t)X * ((uint32_t)(y0 + iy) + (uint32_t)Y * (uint32_t)(z0 + iz))
                       : sid(pos);
            s_tk[pos] = ticket(site);
        }
        // The halo shell of the window (TX + 2H) x (TY + 2H) (x (TZ + 2H)): whole halo planes
        // (3D), then per tile plane the H rows above and below and the H columns either side.
        // The x pads beyond the halo are never read.
        constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
        constexpr int PL = G::HZ > 0 ? 2 * G::HZ * WX * WY : 0;
        constexpr int RG = 2 * H * WX + 2 * H * TY;
        constexpr int HN = PL + TZ * RG;
        for (int h = threadIdx.x; h < HN; h += NT) {
            int wx, wy, wz;
            if (PL > 0 && h < PL) {
                const int p = h / (WX * WY), r = h % (WX * WY);
                wz = p < G::HZ ? p : p + TZ;
                wy = r / WX;
                wx = r % WX;
            } else {
                const int q = h - PL, r = q % RG;
                wz = G::HZ + q / RG;
                if (r < 2 * H * WX) {
                    const int row = r / WX;
                    wx = r % WX;
                    wy = row < H ? row : row + TY;
                } else {
                    const int c = (r - 2 * H * WX) / TY;
                    wy = H + (r - 2 * H * WX) % TY;
                    wx = c < H ? c : c + TX;
                }
            }
            const int pos = (G::PX - H + wx) + G::LX * (wy + G::LY * wz);
            s_tk[pos] = ticket(sid(pos));
        }
    } else {
        for (int i = threadIdx.x; i < G::LN; i += 2 * NT) {
            const int i2 = i + NT;
            const bool m1 = s_mut[i] != 0, m2 = i2 < G::LN && s_mut[i2] != 0;
            const uint32_t t1 = rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority, mcs, sub, 0, kSiteEntity,
                                          s_sid[i], 0, 0).w0;
            const uint32_t t2 = rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority, mcs, sub, 0, kSiteEntity,
                                          i2 < G::LN ? s_sid[i2] : 0u, 0, 0).w0;
            s_tk[i] = m1 ? t1 : 0u;
            if (i2 < G::LN) s_tk[i2] = m2 ? t2 : 0u;
        }
    }
    __syncthreads();

    uint32_t c01 = 0, c23 = 0;  // activated | same owner << 16, boundary null | realized << 16
    uint32_t nw = 0;
    uint16_t *my_work = s_work + wave * Q;
    // Full-ticket comparison against every tying conflict neighbour (rare): words 2-4 of the
    // tickets, then the site id (lottery.jl:153-185).
    auto tie_wins = [&](int pos, uint32_t site, uint32_t mine) __attribute__((always_inline)) {
        bool wins = true;
        const Words b4 = rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority, mcs, sub, 0, kSiteEntity, site, 0, 0);
        for (int k = 0; k < KC && wins; ++k) {
            const int q = pos + s_cloff[k];
            if (!PLAIN && !s_mut[q]) continue;
            const uint32_t nsite = PLAIN ? sid(q) : s_sid[q];
            if (nsite == site || s_tk[q] != mine) continue;
            const Words a4 = rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority, mcs, sub, 0, kSiteEntity, nsite, 0, 0);
            bool better;
            if (a4.w1 != b4.w1) better = a4.w1 > b4.w1;
            else if (a4.w2 != b4.w2) better = a4.w2 > b4.w2;
            else if (a4.w3 != b4.w3) better = a4.w3 > b4.w3;
            else better = nsite < site;
            if (better) wins = false;
        }
        return wins;
    };
    auto eligible_site = [&](uint32_t site) __attribute__((always_inline)) {
        if (d.uniform_degree) return true;
        const uint32_t degree = gp(d.site_degree)[site - 1];
        return bounded_draw(d.key_lo, d.key_hi, kStreamLotteryActivation, mcs, sub, 0, kSiteEntity, site, 0,
                            const_bound(d.round_bound)) < degree + 1;
    };
    if constexpr (MOORE && DIM == 2) {
        // Four consecutive sites of a row per lane: the three ticket rows over them in three
        // 16-byte reads plus the columns either side. The neighbour maximum of site j is
        // max(ud[j - 1], ud[j], ud[j + 1], mid[j - 1], mid[j + 1]) with ud[c] the larger of the
        // tickets above and below column c.
        static_assert(TX % 4 == 0 && G::PX % 4 == 0 && G::LX % 4 == 0, "quads are 16-byte aligned");
        constexpr int QX = TX / 4, NQ = QX * TY;
        for (int q0 = 0; q0 < NQ; q0 += NT) {
            const int qd = q0 + (int)threadIdx.x;
            uint32_t win4 = 0u;
            int ix0 = 0, iy = 0;
            if (qd < NQ) {
                ix0 = 4 * (qd % QX);
                iy = qd / QX;
                const int gx0 = x0 + ix0, gy = y0 + iy;
                const int pos0 = G::pos(ix0, iy, 0);
                const v4i up = *(const v4i *)&s_tk[pos0 - G::LX];
                const v4i md = *(const v4i *)&s_tk[pos0];
                const v4i dn = *(const v4i *)&s_tk[pos0 + G::LX];
                uint32_t ud[6], mid[6];
                ud[0] = max(s_tk[pos0 - G::LX - 1], s_tk[pos0 + G::LX - 1]);
                ud[5] = max(s_tk[pos0 - G::LX + 4], s_tk[pos0 + G::LX + 4]);
                mid[0] = s_tk[pos0 - 1];
                mid[5] = s_tk[pos0 + 4];
                ud[1] = max((uint32_t)up.x, (uint32_t)dn.x);
                ud[2] = max((uint32_t)up.y, (uint32_t)dn.y);
                ud[3] = max((uint32_t)up.z, (uint32_t)dn.z);
                ud[4] = max((uint32_t)up.w, (uint32_t)dn.w);
                mid[1] = (uint32_t)md.x;
                mid[2] = (uint32_t)md.y;
                mid[3] = (uint32_t)md.z;
                mid[4] = (uint32_t)md.w;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (gx0 + j >= X || gy >= Y) continue;
                    const uint32_t mine = mid[j + 1];
                    const uint32_t mx = max(max(max(ud[j], ud[j + 1]), max(ud[j + 2], mid[j])), mid[j + 2]);
                    bool wins = !(mx > mine);
                    const uint32_t site = 1u + (uint32_t)(gx0 + j) + (uint32_t)X * (uint32_t)gy;
                    if (wins && mx == mine) wins = tie_wins(pos0 + j, site, mine);
                    if (wins && eligible_site(site)) win4 |= 1u << j;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                wave_append((win4 >> j) & 1u, (uint16_t)(iy * TX + ix0 + j), nw, my_work);
        }
    } else {
        for (int i0 = 0; i0 < G::TN; i0 += NT) {
            const int i = i0 + (int)threadIdx.x;
            bool need = false;
            if (i < G::TN) {
                const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
                const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
                if (gx < X && gy < Y && gz < Z) {
                    const int pos = G::pos(ix, iy, iz);
                    if (PLAIN || s_mut[pos]) {
                        // a tile site lies inside the lattice: its id needs no wrap
                        const uint32_t site = PLAIN ? 1u + (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz)
                                                    : s_sid[pos];
                        const uint32_t mine = s_tk[pos];
                        // strict lexicographic ticket maximum over the conflict neighbours
                        // (lottery.jl:153-185): word 1 decides unless it ties, which needs the
                        // full tickets (rare)
                        bool beaten = false, tied = false;
                        if constexpr (MOORE) {
                            uint32_t mx = 0u;
#pragma unroll
                            for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
                                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                                    for (int dx = -1; dx <= 1; ++dx)
                                        if (dx != 0 || dy != 0 || dz != 0)
                                            mx = max(mx, s_tk[pos + dx + G::LX * (dy + G::LY * dz)]);
                            beaten = mx > mine;
                            tied = mx == mine;  // a neighbour ties and none is higher
                        } else {
#pragma unroll 8
                            for (int k = 0; k < KC; ++k) {
                                const int q = pos + s_cloff[k];
                                const bool other_site = PLAIN || (s_mut[q] && s_sid[q] != site);
                                const uint32_t other = s_tk[q];
                                beaten |= other_site && other > mine;
                                tied |= other_site && other == mine;
                            }
                        }
                        bool wins = !beaten;
                        if (wins && tied) wins = tie_wins(pos, site, mine);
                        // activated: counted, classified and drawn for in the dense pass below
                        if (wins) need = eligible_site(site);
                    }
                }
            }
            wave_append(need, (uint16_t)(i < G::TN ? i : 0), nw, my_work);
        }
    }
    if (__lane_id() == 0) s_len[wave] = nw;
    __syncthreads();

    const uint32_t p1 = s_len[0], p2 = p1 + (NW > 1 ? s_len[NW > 1 ? 1 : 0] : 0u);
    const uint32_t p3 = p2 + (NW > 2 ? s_len[NW > 2 ? 2 : 0] : 0u);
    const uint32_t nwork = p3 + (NW > 3 ? s_len[NW > 3 ? 3 : 0] : 0u);
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims;
    CPM_GLOBAL v4i *seg = (CPM_GLOBAL v4i *)(a.seg + (size_t)tile * a.seg_cap);
    uint32_t na = 0;  // actionable attempts of this wave (entries e with the same position)
    for (uint32_t e0 = 0; e0 < nwork; e0 += NT) {
        const uint32_t e = e0 + threadIdx.x;
        bool act = false;
        v4i v = v4i{0, 0, 0, 0};
        int pos = 0;
        bool drawn = false;
        if (e < nwork) {
            const int i = s_work[list_slot<Q>(e, p1, p2, p3)];
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            pos = G::pos(ix, iy, iz);
            c01 += 1u;  // activated
            // a winner whose proposal neighbours all share its owner is a same-owner no-op
            // whatever direction it would draw
            if (interior_site<KF, PLAIN>(s_own, s_mut, pos, s_own[pos], po, K, s_ploff)) c01 += 1u << 16;
            else drawn = true;
        }
        if (drawn) {
            const int i = s_work[list_slot<Q>(e, p1, p2, p3)];
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const uint32_t site = PLAIN ? 1u + (uint32_t)(x0 + ix) +
                                              (uint32_t)X * ((uint32_t)(y0 + iy) + (uint32_t)Y * (uint32_t)(z0 + iz))
                                        : s_sid[pos];
            const uint32_t dir = bounded_draw(d.key_lo, d.key_hi, kStreamDirection, mcs, sub, 0,
                                              kSiteEntity, site, 0, const_bound(d.prop_bound));
            const int q = pos + direction_offset<KF>(dir, po, s_ploff);
            if (!PLAIN && !s_mut[q]) {
                c23 += 1u;
            } else {
                const int32_t own = s_own[pos], gain = s_own[q];
                if (gain == own) {
                    c01 += 1u << 16;
                } else {
                    c23 += 1u << 16;
                    const uint32_t prio = s_tk[pos];
                    const unsigned long long key = ((unsigned long long)prio << 32) | (uint32_t)(~site);
                    if (own > 0) atomicMax(&claims[d.claim_base + word_cell(own)], key);
                    if (gain > 0) atomicMax(&claims[d.claim_base + word_cell(gain)], key);
                    v = v4i{(int32_t)site, (int32_t)prio, own, gain};
                    act = true;
                }
            }
        }
        // wave-private stretches of the tile segment: wave w owns entries [w * Q, (w + 1) * Q)
        const uint64_t m = __ballot(act);
        if (act) seg[wave * Q + na + __popcll(m & __lanemask_lt())] = v;
        na += (uint32_t)__popcll(m);
    }
    block_add_packed(c01, c23, s_sum);
    if (__lane_id() == 0) a.seg_count[tile * kWaves + wave] = na;
    if ((int)threadIdx.x >= NW && threadIdx.x < kWaves) a.seg_count[tile * kWaves + threadIdx.x] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        const uint32_t lo = s_sum[0], hi = s_sum[1];
        slot[0] = lo & 0xffffu;
        slot[1] = lo >> 16;
        slot[2] = hi & 0xffffu;
        slot[3] = hi >> 16;
    }
}

// ---------------------------------------------------------------------------
// Claim resolution, evaluation, acceptance and commit (checkerboard.jl:337-457)
// ---------------------------------------------------------------------------

// Owner word seen by a contact/surface relation at site (x,y,z) + offset; 0 when the
// neighbour is absent or invalid (cartesian.jl:494-525).
template <class D, class O>
__device__ __forceinline__ int32_t neighbour_word(const D &d, int x, int y, int z, const O *o) {
    if (d.simple) {
        const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
        int nx = x + o[0], ny = y + o[1], nz = z + o[2];
        nx += nx < 0 ? X : 0;
        nx -= nx >= X ? X : 0;
        ny += ny < 0 ? Y : 0;
        ny -= ny >= Y ? Y : 0;
        nz += nz < 0 ? Z : 0;
        nz -= nz >= Z ? Z : 0;
        return gp(d.lattice)[(uint32_t)nx + (uint32_t)X * ((uint32_t)ny + (uint32_t)Y * (uint32_t)nz)];
    }
    return realize_position(d, x + o[0], y + o[1], z + o[2]).word;
}

// Evaluation of a copy attempt (checkerboard.jl:371-408; scientific_inner_loop.jl:213-269;
// protocols/scientific.jl:372-385). Checkerboard and lottery winners touch disjoint cells and
// non-adjacent sites, so other winners' commits never change what an evaluation reads.
//
// Memory latency dominates this function, so every load that depends only on the entry
// (both cells' trackers and parameters, up to kNb surface and kNb contact neighbour words)
// is issued before any is used, then one batch of neighbour cell types; the arithmetic
// keeps the reference's operation order.
constexpr int kNb = 8;

// The fields of a cell record that evaluation reads, from three 16-byte loads of its line.
struct CellHot {
    int32_t vol;
    uint32_t ctype;
    int64_t bnd;
    double tv, vs;
    int64_t tb;
    double bs;
};

__device__ __forceinline__ CellHot load_cell(const CPM_GLOBAL CellRec *cells, uint32_t i) {
    const CPM_GLOBAL v4i *p = (const CPM_GLOBAL v4i *)(cells + i);
    const v4i a = p[0], b = p[1], c = p[2];
    CellHot h;
    h.vol = a.x;
    h.ctype = (uint32_t)a.y;
    h.bnd = (int64_t)(((uint64_t)(uint32_t)a.w << 32) | (uint32_t)a.z);
    h.tv = __hiloint2double(b.y, b.x);
    h.vs = __hiloint2double(b.w, b.z);
    h.tb = (int64_t)(((uint64_t)(uint32_t)c.y << 32) | (uint32_t)c.x);
    h.bs = __hiloint2double(c.w, c.z);
    return h;
}

// Tracker values of a winner's cells at the pass snapshot and the boundary changes of its
// copy: the commit writes snapshot + change (no other winner touches these cells).
struct Staged {
    int32_t dl, dg;    // boundary-measure changes of the losing and gaining cells
    int32_t vl, vg;    // snapshot volumes (cell owners only)
    int64_t bl, bg;    // snapshot boundary measures (cell owners only)
};

// Changes added to the snapshot trackers of an attempt's two cells: a tile's own earlier
// copies in a TiledCheckerboardCPM batch (the tile-local view), zero for the other algorithms.
struct View {
    int32_t dvl, dvg;  // volume changes of the losing and gaining cells
    int32_t dbl, dbg;  // boundary-measure changes
};

// evaluate_copy of one actionable attempt (losing owner l, gaining owner g) from its gathered
// inputs: both cells' records and mechanical states, the first kNb surface and contact
// neighbour words with the contact neighbours' types, `more(role, k)` for the words of
// directions k >= kNb (role 0 surface, 1 contact) and `jf` / `jd` for the contact matrix.
// Trackers are the records plus `view`: delta_h folded over the energy terms, plus the
// mechanical work (scientific_inner_loop.jl:213-269). Fills the tracker values and boundary
// changes a commit writes.
template <class D, class More, class JF, class JD>
__device__ __forceinline__ float copy_energy_from(const D &d, int32_t l, int32_t g, const View &view,
                                                  const CellHot &cl, const CellHot &cg, float qpl, float qtl,
                                                  float qpg, float qtg, const int32_t (&sw)[kNb],
                                                  const int32_t (&cw)[kNb], const uint32_t (&ct)[kNb], More &&more,
                                                  JF &&jf, JD &&jd, Staged *st, float *dh_out = nullptr) {
    const int32_t vl0 = cl.vol + view.dvl, vg0 = cg.vol + view.dvg;
    const double vs_l = cl.vs, tv_l = cl.tv, vs_g = cg.vs, tv_g = cg.tv;
    const double bs_l = cl.bs, bs_g = cg.bs;
    const int64_t tb_l = cl.tb, tb_g = cg.tb, bm_l = cl.bnd + view.dbl, bm_g = cg.bnd + view.dbg;
    const int ns = d.surf.n, nc = d.has_contact ? d.cont.n : 0;
    const int32_t vl = l > 0 ? vl0 : 0, vg = g > 0 ? vg0 : 0;

    // boundary tracker deltas (BoundaryEdgeCount, scientific_components.jl:466-481)
    int32_t dl = 0, dg = 0;
    auto surface_word = [&](int32_t w) __attribute__((always_inline)) {
        if (w == 0) return;
        dl += (w == l) ? 1 : -1;
        dg += (w == g) ? -1 : 1;
    };
#pragma unroll
    for (int u = 0; u < kNb; ++u) surface_word(sw[u]);
    for (int k = kNb; k < ns; ++k) surface_word(more(0, k));
    st->dl = dl;
    st->dg = dg;
    st->vl = vl;
    st->vg = vg;
    st->bl = bm_l;
    st->bg = bm_g;

    double acc = 0.0;
    bool acc64 = false;
    for (int t = 0; t < d.n_energies; ++t) {
        const int term = d.energy_order[t];
        double v;
        bool v64;
        if (term == CPM_TERM_VOLUME) {
            v64 = d.volume_f64;
            if (v64) {
                double delta = 0.0;
                if (l > 0) delta = __dadd_rn(delta, vol_change_f64(vs_l, tv_l, vl, -1));
                if (g > 0) delta = __dadd_rn(delta, vol_change_f64(vs_g, tv_g, vg, 1));
                v = delta;
            } else {
                float delta = 0.0f;
                if (l > 0) delta = __fadd_rn(delta, vol_change_f32((float)vs_l, (float)tv_l, vl, -1));
                if (g > 0) delta = __fadd_rn(delta, vol_change_f32((float)vs_g, (float)tv_g, vg, 1));
                v = (double)delta;
            }
        } else if (term == CPM_TERM_CONTACT) {
            // scientific_components.jl:289-308, directions in canonical order
            v64 = d.contact_f64;
            const uint32_t nt = (uint32_t)d.n_types;
            const uint32_t lt = l > 0 ? min(max(cl.ctype, 1u), nt) : owner_type(d, l);
            const uint32_t gt = g > 0 ? min(max(cg.ctype, 1u), nt) : owner_type(d, g);
            double dd = 0.0;
            float df = 0.0f;
            auto contact_term = [&](int k, int32_t w, uint32_t tw) __attribute__((always_inline)) {
                if (w == 0) return;
                uint32_t wt;
                if (w == l) wt = lt;
                else if (w == g) wt = gt;
                else if (w > 0) wt = min(max(tw, 1u), nt);
                else wt = owner_type(d, w);
                const uint32_t il = (lt - 1) * nt + (wt - 1), ig = (gt - 1) * nt + (wt - 1);
                if (v64) {
                    const double wk = d.cont.wd[k];
                    if (l != w) dd = __dsub_rn(dd, __dmul_rn(wk, jd(il)));
                    if (g != w) dd = __dadd_rn(dd, __dmul_rn(wk, jd(ig)));
                } else {
                    const float wk = d.cont.wf[k];
                    if (l != w) df = __fsub_rn(df, __fmul_rn(wk, jf(il)));
                    if (g != w) df = __fadd_rn(df, __fmul_rn(wk, jf(ig)));
                }
            };
#pragma unroll
            for (int u = 0; u < kNb; ++u)
                if (u < nc) contact_term(u, cw[u], ct[u]);
            // directions beyond the first kNb in groups of kTail: the group's neighbour words,
            // then their types, are all in flight before the terms fold in canonical order
            constexpr int kTail = 4;
            for (int k0 = kNb; k0 < nc; k0 += kTail) {
                int32_t gw[kTail];
                uint32_t gt[kTail];
#pragma unroll
                for (int u = 0; u < kTail; ++u) gw[u] = k0 + u < nc ? more(1, k0 + u) : 0;
#pragma unroll
                for (int u = 0; u < kTail; ++u) gt[u] = gw[u] > 0 ? word_type(gw[u]) : 0u;
#pragma unroll
                for (int u = 0; u < kTail; ++u)
                    if (k0 + u < nc) contact_term(k0 + u, gw[u], gt[u]);
            }
            v = v64 ? dd : (double)df;
        } else {
            // scientific_trackers.jl:144-172
            v64 = d.boundary_f64;
            if (v64) {
                double r = 0.0;
                if (l > 0) r = __dadd_rn(r, bnd_change_f64(bs_l, (double)tb_l, (double)bm_l, vl, (double)dl, -1));
                if (g > 0) r = __dadd_rn(r, bnd_change_f64(bs_g, (double)tb_g, (double)bm_g, vg, (double)dg, 1));
                v = r;
            } else {
                float r = 0.0f;
                if (l > 0) r = __fadd_rn(r, bnd_change_f32((float)bs_l, (float)tb_l, (float)bm_l, vl, (float)dl, -1));
                if (g > 0) r = __fadd_rn(r, bnd_change_f32((float)bs_g, (float)tb_g, (float)bm_g, vg, (float)dg, 1));
                v = (double)r;
            }
        }
        if (!acc64 && !v64) {
            acc = (double)__fadd_rn((float)acc, (float)v);
        } else {
            acc = __dadd_rn(acc, v);
            acc64 = true;
        }
    }
    const float dh = (float)acc;
    if (dh_out) *dh_out = dh;
    // mechanical work folded in tuple order (scientific_inner_loop.jl:238-242;
    // scientific_components.jl:220-231, 646-657), then added to delta_h
    float mech = 0.0f;
    for (int k = 0; k < d.n_mech; ++k) {
        float r = 0.0f;
        if (d.mech_kind[k] == CPM_MECH_PRESSURE) {
            if (l > 0) r = __fsub_rn(r, qpl);
            if (g > 0) r = __fadd_rn(r, qpg);
        } else {
            if (l > 0) r = __fadd_rn(r, __fmul_rn(qtl, (float)dl));
            if (g > 0) r = __fadd_rn(r, __fmul_rn(qtg, (float)dg));
        }
        mech = __fadd_rn(mech, r);
    }
    return __fadd_rn(dh, mech);
}

// evaluate_copy of one actionable attempt (recipient `site`, losing owner l, gaining owner g)
// against the trackers in the cell records plus `view`, reading the lattice in global memory.
template <class D>
__device__ __forceinline__ float copy_energy(const D &d, uint32_t site, int32_t l, int32_t g, const View &view,
                                             Staged *st, float *dh_out = nullptr) {
    const uint32_t idx0 = site - 1;
    const FastDiv div_xy{d.div_xy.d, d.div_xy.m}, div_x{d.div_x.d, d.div_x.m};
    const uint32_t zq = fast_div(idx0, div_xy);
    const uint32_t rem = idx0 - zq * d.div_xy.d;
    const uint32_t yq = fast_div(rem, div_x);
    const int x = (int)(rem - yq * d.div_x.d);
    const int y = (int)yq;
    const int z = (int)zq;
    // slot 0 of every per-cell array is a spare, read in place of a medium owner
    const uint32_t li = cell_slot(l), gi = cell_slot(g);
    const CellHot cl = load_cell(gp(d.cells), li), cg = load_cell(gp(d.cells), gi);
    // mechanical state of both cells (volume_pressure, surface_tension)
    float qpl = 0.0f, qtl = 0.0f, qpg = 0.0f, qtg = 0.0f;
    if (d.n_mech > 0) {
        const v4i ml = ((const CPM_GLOBAL v4i *)(gp(d.cells) + li))[3];
        const v4i mg = ((const CPM_GLOBAL v4i *)(gp(d.cells) + gi))[3];
        qpl = __int_as_float(ml.x);
        qtl = __int_as_float(ml.y);
        qpg = __int_as_float(mg.x);
        qtg = __int_as_float(mg.y);
    }
    const int ns = d.surf.n, nc = d.has_contact ? d.cont.n : 0;
    // sites at least `halo` from every face of a periodic, obstacle-free lattice reach
    // their neighbours by a linear delta
    const int h = d.halo;
    const bool inner = d.simple && x >= h && x < d.dims[0] - h && y >= h && y < d.dims[1] - h &&
                       (d.ndim == 2 || (z >= h && z < d.dims[2] - h));
    const CPM_GLOBAL int32_t *lat = gp(d.lattice);
    auto nb = [&](const auto *o, int32_t lin) __attribute__((always_inline)) {
        return inner ? lat[(int32_t)idx0 + lin] : neighbour_word(d, x, y, z, o);
    };
    // neighbour words: one unconditional load per direction (lanes off the fast path load
    // their own site), so every load is in flight at once; lanes within `halo` of a face
    // or on a lattice with walls or obstacles then realize their neighbours one by one
    int32_t sw[kNb], cw[kNb];
#pragma unroll
    for (int u = 0; u < kNb; ++u) sw[u] = u < ns ? lat[(int32_t)idx0 + (inner ? d.surf_lin[u] : 0)] : 0;
#pragma unroll
    for (int u = 0; u < kNb; ++u) cw[u] = u < nc ? lat[(int32_t)idx0 + (inner ? d.cont_lin[u] : 0)] : 0;
    if (!inner) {
#pragma unroll
        for (int u = 0; u < kNb; ++u)
            if (u < ns) sw[u] = neighbour_word(d, x, y, z, d.surf.off[u]);
#pragma unroll
        for (int u = 0; u < kNb; ++u)
            if (u < nc) cw[u] = neighbour_word(d, x, y, z, d.cont.off[u]);
    }
    uint32_t ct[kNb];
#pragma unroll
    for (int u = 0; u < kNb; ++u) ct[u] = cw[u] > 0 ? word_type(cw[u]) : 0u;
    auto more = [&](int role, int k) __attribute__((always_inline)) {
        if (role == 0) return k < 32 ? nb(d.surf.off[k], d.surf_lin[k]) : neighbour_word(d, x, y, z, d.surf.off[k]);
        return k < 32 ? nb(d.cont.off[k], d.cont_lin[k]) : neighbour_word(d, x, y, z, d.cont.off[k]);
    };
    return copy_energy_from(
        d, l, g, view, cl, cg, qpl, qtl, qpg, qtg, sw, cw, ct, more,
        [&](uint32_t i) __attribute__((always_inline)) { return d.Jf[i]; },
        [&](uint32_t i) __attribute__((always_inline)) { return d.Jd[i]; }, st, dh_out);
}

// evaluate_copy of one actionable attempt on the common model shape (RepDev::fast_eval): a
// periodic obstacle-free lattice, Float32 energy terms in tuple order, no mechanics, NS surface
// directions that are the first of the NC contact directions (NC = 0: no contact term), all
// within one site. The same arithmetic as copy_energy_from, with every loop fixed at compile
// time: both cell records and all NW neighbour words are loaded at once, then all NC
// neighbour types, so an evaluation waits on two memory levels whatever NC is.
// sJ: the replica's Float32 contact matrix staged in LDS, or nullptr to read it from the
// descriptor.
typedef const __attribute__((address_space(3))) float LdsFloat;
template <int DIM, int NS, int NC, class D>
__device__ __forceinline__ float fast_copy_energy(const D &d, uint32_t site, int32_t l, int32_t g, Staged *st,
                                                  LdsFloat *sJ) {
    constexpr int NW = NC > 0 ? NC : NS;
    const uint32_t idx0 = site - 1;
    const FastDiv div_xy{d.div_xy.d, d.div_xy.m}, div_x{d.div_x.d, d.div_x.m};
    const uint32_t zq = DIM == 3 ? fast_div(idx0, div_xy) : 0u;
    const uint32_t rem = idx0 - zq * d.div_xy.d;
    const uint32_t yq = fast_div(rem, div_x);
    const int x = (int)(rem - yq * d.div_x.d), y = (int)yq, z = (int)zq;
    const uint32_t li = cell_slot(l), gi = cell_slot(g);
    const CellHot cl = load_cell(gp(d.cells), li), cg = load_cell(gp(d.cells), gi);
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const bool inner = x >= 1 && x < X - 1 && y >= 1 && y < Y - 1 && (DIM == 2 || (z >= 1 && z < Z - 1));
    const CPM_GLOBAL int32_t *lat = gp(d.lattice);
    const int32_t sy = X, sz = X * Y;
    int32_t w[NW];
    if (inner) {
#pragma unroll
        for (int k = 0; k < NW; ++k)
            w[k] = lat[(int32_t)idx0 + moore_offset<DIM>(k, 0) + sy * moore_offset<DIM>(k, 1) +
                       sz * moore_offset<DIM>(k, 2)];
    } else {
        // periodic wrap of each axis (the relations reach one site)
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            int nx = x + moore_offset<DIM>(k, 0), ny = y + moore_offset<DIM>(k, 1), nz = z + moore_offset<DIM>(k, 2);
            nx = nx < 0 ? X - 1 : nx == X ? 0 : nx;
            ny = ny < 0 ? Y - 1 : ny == Y ? 0 : ny;
            nz = nz < 0 ? Z - 1 : nz == Z ? 0 : nz;
            w[k] = lat[(uint32_t)nx + (uint32_t)sy * (uint32_t)ny + (uint32_t)sz * (uint32_t)nz];
        }
    }
    uint32_t ct[NC > 0 ? NC : 1];
#pragma unroll
    for (int k = 0; k < NC; ++k) ct[k] = w[k] > 0 ? word_type(w[k]) : 0u;

    const int32_t vl = l > 0 ? cl.vol : 0, vg = g > 0 ? cg.vol : 0;
    // boundary tracker deltas (BoundaryEdgeCount, scientific_components.jl:466-481)
    int32_t dl = 0, dg = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        dl += (w[k] == l) ? 1 : -1;
        dg += (w[k] == g) ? -1 : 1;
    }
    st->dl = dl;
    st->dg = dg;
    st->vl = vl;
    st->vg = vg;
    st->bl = cl.bnd;
    st->bg = cg.bnd;
    float acc = 0.0f;
    for (int t = 0; t < d.n_energies; ++t) {
        const int term = d.energy_order[t];
        float v;
        if (term == CPM_TERM_VOLUME) {
            float delta = 0.0f;
            if (l > 0) delta = __fadd_rn(delta, vol_change_f32((float)cl.vs, (float)cl.tv, vl, -1));
            if (g > 0) delta = __fadd_rn(delta, vol_change_f32((float)cg.vs, (float)cg.tv, vg, 1));
            v = delta;
        } else if (term == CPM_TERM_CONTACT) {
            // scientific_components.jl:289-308, directions in canonical order
            const uint32_t nt = (uint32_t)d.n_types;
            const uint32_t lt = l > 0 ? min(max(cl.ctype, 1u), nt) : owner_type(d, l);
            const uint32_t gt = g > 0 ? min(max(cg.ctype, 1u), nt) : owner_type(d, g);
            float df = 0.0f;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const int32_t o = w[k];
                uint32_t wt;
                if (o == l) wt = lt;
                else if (o == g) wt = gt;
                else if (o > 0) wt = min(max(ct[k], 1u), nt);
                else wt = owner_type(d, o);
                const float wk = d.cont.wf[k];
                const uint32_t il = (lt - 1) * nt + (wt - 1), ig = (gt - 1) * nt + (wt - 1);
                const float jl = sJ ? sJ[il] : d.Jf[il], jg = sJ ? sJ[ig] : d.Jf[ig];
                if (l != o) df = __fsub_rn(df, __fmul_rn(wk, jl));
                if (g != o) df = __fadd_rn(df, __fmul_rn(wk, jg));
            }
            v = df;
        } else {
            // scientific_trackers.jl:144-172
            float r = 0.0f;
            if (l > 0) r = __fadd_rn(r, bnd_change_f32((float)cl.bs, (float)cl.tb, (float)cl.bnd, vl, (float)dl, -1));
            if (g > 0) r = __fadd_rn(r, bnd_change_f32((float)cg.bs, (float)cg.tb, (float)cg.bnd, vg, (float)dg, 1));
            v = r;
        }
        acc = __fadd_rn(acc, v);
    }
    return acc;
}

// ConventionalMetropolis acceptance probability of an energy change at temperature T
// (protocols/scientific.jl:372-385); exp of the Float32 log-ratio in double precision,
// rounded once.
__device__ __forceinline__ float metropolis_p(float T, float dd) {
    if (T == 0.0f) return __fadd_rn(dd, 0.0f) <= 0.0f ? 1.0f : 0.0f;
    float lr = __fdiv_rn(-__fadd_rn(dd, 0.0f), T);
    lr = __fadd_rn(__fadd_rn(lr, 0.0f), 0.0f);
    return lr >= 0.0f ? 1.0f : (float)exp((double)lr);
}

// evaluate_copy + ConventionalMetropolis + acceptance draw of one winning attempt against
// the pass snapshot (checkerboard.jl:371-408).
// *dh receives the energy change (delta_h, without the mechanical work).
// FAST: the group's RepDev::fast_eval (every replica of the launch shares it), 0 for the
// general evaluation. Each kind is its own kernel instantiation, so the registers of the
// largest evaluation do not limit the occupancy of the others.
template <int FAST, class D>
__device__ inline bool evaluate_winner(const D &d, const Entry &e, uint64_t mcs, uint32_t sub, Staged *st,
                                       float *dh, LdsFloat *sJ) {
    float total;
    if constexpr (FAST == 0) {
        const View none{0, 0, 0, 0};
        total = copy_energy(d, e.site, e.losing, e.gaining, none, st, dh);
    } else {
        constexpr int DIM = FAST <= 3 ? 2 : 3;
        constexpr int NS = 2 * DIM;
        constexpr int NC = FAST == 1 || FAST == 4 ? 0 : FAST == 2 ? 4 : FAST == 5 ? 6 : FAST == 3 ? 8 : 26;
        total = fast_copy_energy<DIM, NS, NC>(d, e.site, e.losing, e.gaining, st, sJ);
        *dh = total;
    }
    const float p = metropolis_p(d.temperature, total);
    const Words w = rng_words(d.key_lo, d.key_hi, kStreamAcceptance, mcs, sub, 0, kSiteEntity, e.site, 0, 0);
    return uniform_open01(w.w0) < p;
}

// Commit of one accepted winning attempt: ownership, volumes, boundary measures,
// retirement, medium volumes (checkerboard.jl:410-440).
template <class D>
__device__ inline void commit_attempt(const D &d, const Entry &e, const Staged &st) {
    const int32_t l = e.losing, g = e.gaining;
    gp(d.lattice)[e.site - 1] = g;
    CPM_GLOBAL CellRec *cells = gp(d.cells);
    if (l > 0) {
        CPM_GLOBAL CellRec &c = cells[word_cell(l)];
        const int32_t nv = st.vl - 1;
        c.vol = nv;
        c.bnd = st.bl + st.dl;
        if (nv == 0) {
            // retirement; the cell owns no site, so no owner word holds its type any more
            c.active = 0;
            c.ctype = 0;
            c.tv = d.retire_tv;
            c.vs = d.retire_vs;
            c.tb = d.retire_tb;
            c.bs = d.retire_bs;
            c.qp = 0.0f;
            c.qt = 0.0f;
        }
    } else {
        atomicAdd(&d.medium_vol[medium_slot(d, l)], -1);
    }
    if (g > 0) {
        CPM_GLOBAL CellRec &c = cells[word_cell(g)];
        c.vol = st.vg + 1;
        c.bnd = st.bg + st.dg;
    } else {
        atomicAdd(&d.medium_vol[medium_slot(d, g)], 1);
    }
}


// Claim check, evaluation, acceptance and commit of one pass (checkerboard.jl:337-457) in
// one workgroup per region of CT tile segments (4, or 2 for tiles of 4096 sites and more, so
// that a region holds a similar number of sites). An attempt wins iff its key is the
// maximum on both of its cells. Entries are read in batches of kCheckBatch chunks whose
// entry loads, then claim loads, are all in flight together; winners are compacted into
// LDS and evaluated against the pass snapshot in full chunks, and commit at once (winners
// touch disjoint cells and non-adjacent sites). The workgroup also clears a slice of the
// next pass's claim buffer. The tiles of a replica are contiguous, so a region whose
// first and last tile share a replica belongs to that replica alone (one round per chunk);
// otherwise one round per tile. The replica index is uniform in every round, so its
// descriptor is read through the scalar cache.
constexpr int kCheckBatch = 4;
template <int NWV, int CT, int FAST>
__global__ __launch_bounds__(64 * NWV) void commit_kernel(PassArgs a) {
    constexpr int NT = 64 * NWV;
    constexpr int NS = CT * kWaves;
    constexpr int WCAP = kCheckBatch * NT;
    static_assert((NS & (NS - 1)) == 0 && NS <= 64, "stretch search halves a power of two");
    __shared__ uint32_t s_base[NS + 1];
    __shared__ uint32_t s_cbase[CT];
    __shared__ uint32_t s_cnt[CT][3];  // conflicts, rejections, acceptances
    __shared__ double s_dh[CT];        // accepted copies' delta_h
    __shared__ float s_J[FAST ? CPM_MAX_TYPES * CPM_MAX_TYPES : 1];  // contact matrix of a one-replica region
    __shared__ v4i s_win[WCAP];
    __shared__ uint32_t s_nw;
    const uint32_t nwg = gridDim.x;
    const uint32_t b = xcd_tile(blockIdx.x, nwg);
    const uint32_t lane = __lane_id();
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    {
        unsigned long long *next = a.claims + (size_t)(a.parity ^ 1u) * a.total_claims;
        const uint32_t per = (a.total_claims + nwg - 1) / nwg;
        const uint32_t lo = b * per;
        const uint32_t hi = min(lo + per, a.total_claims);
        for (uint32_t j = lo + tid; j < hi; j += NT) next[j] = 0ull;
    }
    const uint32_t t0 = b * CT;
    const uint32_t nt = min((uint32_t)CT, a.ntiles - t0);
    // stretch counts (idle tiles count zero) and their inclusive scan
    uint32_t n = 0;
    bool live = false;
    if (tid < (uint32_t)NS && lane / kWaves < nt) {
        const uint32_t c = a.seg_count[t0 * kWaves + lane];
        live = c != kIdleTile;
        n = live ? c : 0u;
    }
#pragma unroll
    for (int o = 1; o < NS; o <<= 1) {
        const uint32_t t = __shfl_up(n, o);
        if ((int)lane >= o) n += t;
    }
    if (tid < (uint32_t)NS) s_base[tid + 1] = n;
    if (tid == 0) {
        s_base[0] = 0;
        s_nw = 0;
    }
    const uint64_t live_tiles = __ballot(live && lane % kWaves == 0);  // wave 0: stretch 0 carries the flag
    if (tid < nt) s_cbase[tid] = a.tiles[t0 + tid].claim_base;
    if (tid < (uint32_t)CT * 3) s_cnt[tid / 3][tid % 3] = 0;
    if (tid < (uint32_t)CT) s_dh[tid] = 0.0;
    __syncthreads();
    const uint32_t total = s_base[NS];
    const uint32_t Q = a.seg_cap / kWaves;
    const unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims;
    const CPM_GLOBAL v4i *region = (const CPM_GLOBAL v4i *)(a.seg + (size_t)t0 * a.seg_cap);
    const uint32_t r_first = a.tiles[t0].rep;
    const bool one = a.tiles[t0 + nt - 1].rep == r_first;
    if (FAST && one) {
        // staged while the claim check runs; read after its barrier
        const CPM_CONST RepDev &d0 = ((const CPM_CONST RepDev *)a.reps)[r_first];
        const uint32_t nj = (uint32_t)(d0.n_types * d0.n_types);
        for (uint32_t i = tid; i < nj; i += NT) s_J[i] = d0.Jf[i];
    }
    LdsFloat *sJ = FAST && one ? (LdsFloat *)s_J : nullptr;
    for (uint32_t blk = 0; blk < total; blk += WCAP) {
        // entries of the batch, then their claims
        v4i v[kCheckBatch];
        uint32_t kt[kCheckBatch];
        // branch-free: lanes past the end read entry 0 and cell slot 0, so all loads of
        // the batch are in flight together
#pragma unroll
        for (int j = 0; j < kCheckBatch; ++j) {
            const uint32_t e = min(blk + (uint32_t)j * NT + tid, total - 1);
            int k = 0;
#pragma unroll
            for (int step = NS / 2; step > 0; step >>= 1)
                if (e >= s_base[k + step]) k += step;
            kt[j] = (uint32_t)k / kWaves;
            v[j] = region[kt[j] * a.seg_cap + (uint32_t)(k % kWaves) * Q + (e - s_base[k])];
        }
        unsigned long long cl[kCheckBatch], cg[kCheckBatch];
#pragma unroll
        for (int j = 0; j < kCheckBatch; ++j) {
            const uint32_t cb = s_cbase[kt[j]];
            cl[j] = claims[cb + cell_slot(v[j].z)];
            cg[j] = claims[cb + cell_slot(v[j].w)];
        }
#pragma unroll
        for (int j = 0; j < kCheckBatch; ++j) {
            const uint32_t e = blk + (uint32_t)j * NT + tid;
            bool win = false;
            if (e < total) {
                const unsigned long long key =
                    ((unsigned long long)(uint32_t)v[j].y << 32) | (uint32_t)(~(uint32_t)v[j].x);
                win = (v[j].z <= 0 || cl[j] == key) && (v[j].w <= 0 || cg[j] == key);
                if (!win) atomicAdd(&s_cnt[kt[j]][0], 1u);
            }
            const uint64_t m = __ballot(win);
            if (m) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&s_nw, (uint32_t)__popcll(m));
                base = __shfl(base, 0);
                if (win) s_win[base + __popcll(m & __lanemask_lt())] = v4i{v[j].x, (int32_t)kt[j], v[j].z, v[j].w};
            }
        }
        __syncthreads();
        const uint32_t nw = s_nw;
        // winners in full chunks (wave w takes chunks w, w + NWV, ...): energy change,
        // acceptance and commit
        for (uint32_t w0 = wave * 64; w0 < nw; w0 += NT) {
            const uint32_t w = w0 + lane;
            v4i rec = v4i{0, -1, 0, 0};
            if (w < nw) rec = s_win[w];
            const int kw = rec.y;  // -1 for lanes without a winner
            for (uint32_t k = 0; k < nt; ++k) {
                const uint32_t r = one ? r_first : a.tiles[t0 + k].rep;
                if (w < nw && (one || kw == (int)k)) {
                    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)a.reps)[r];
                    const CPM_CONST McsPlan &y = ((const CPM_CONST McsPlan *)a.plan)[(size_t)r * kPlanMcs + a.plan_k];
                    const uint32_t sub = d.algorithm == CPM_ALG_LOTTERY ? y.order[a.pass] - 1u : 0u;
                    const Entry ent{(uint32_t)rec.x, 0u, rec.z, rec.w};
                    Staged st;
                    float dh;
                    const bool accept = evaluate_winner<FAST>(d, ent, y.mcs, sub, &st, &dh, sJ);
                    if (accept) {
                        commit_attempt(d, ent, st);
                        atomicAdd(&s_dh[kw], (double)dh);
                    }
                    atomicAdd(&s_cnt[kw][accept ? 2 : 1], 1u);
                }
                if (one) break;
            }
        }
        __syncthreads();
        if (tid == 0) s_nw = 0;
        __syncthreads();
    }
    if (tid < nt && ((live_tiles >> (lane * kWaves)) & 1u)) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + t0 + lane) * kSlotFields;
        slot[4] = s_cnt[lane][0];
        slot[5] = s_cnt[lane][1];
        slot[6] = s_cnt[lane][2];
        *(double *)(slot + 8) = s_dh[lane];
    }
}

// ---------------------------------------------------------------------------
// Mechanics (lib/CorePotts/src/components/scientific_mechanics.jl)
// ---------------------------------------------------------------------------

// Correctly rounded Float32 square root: the double-precision root rounded once (the
// double has more than 2 * 24 + 2 significand bits, so the second rounding is exact for
// square roots). The single-precision hardware root is accurate to 1 ulp only.
__device__ __forceinline__ float sqrt_rn(float x) { return (float)sqrt((double)x); }

// normal_box_muller(Float32) (rng/semantic.jl:323-331): lanes 1 and 2 of one Philox call;
// log and cos are evaluated in double precision and rounded once to Float32.
__device__ inline float normal_box_muller(uint32_t key_lo, uint32_t key_hi, uint32_t stream, uint64_t mcs,
                                          uint32_t subround, uint32_t operation, uint32_t entity,
                                          uint32_t invocation) {
    const Words w = rng_words(key_lo, key_hi, stream, mcs, subround, operation, kCellEntity, entity, invocation, 0);
    const float u1 = uniform_open01(w.w0), u2 = uniform_open01(w.w1);
    const float radius = sqrt_rn(__fmul_rn(-2.0f, (float)log((double)u1)));
    const float angle = __fmul_rn(__fmul_rn(2.0f, 3.14159274101257324f), u2);  // 2 Float32(pi)
    return __fmul_rn(radius, (float)cos((double)angle));
}

// _mechanical_ou_transition (scientific_mechanics.jl:29-35), Float32 in the reference's
// operation order; exp in double precision rounded once.
__device__ inline float ou_transition(float q, float disp, float strength, float eta, float noise, float interval,
                                      float normal) {
    const float alpha = (float)exp((double)__fmul_rn(-eta, interval));
    const float two_s = __fmul_rn(2.0f, strength);
    const float mean = __fmul_rn(__fsub_rn(1.0f, alpha), __fmul_rn(two_s, disp));
    const float var = __fmul_rn(__fmul_rn(two_s, noise), fmaxf(0.0f, __fsub_rn(1.0f, __fmul_rn(alpha, alpha))));
    return __fadd_rn(__fadd_rn(__fmul_rn(alpha, q), mean), __fmul_rn(sqrt_rn(var), normal));
}

// Observable minus target, and strength, of mechanical component k for one cell
// (_mechanical_mean_inputs, scientific_mechanics.jl:47-56): finite volume against
// target_volume / volume_strength, or boundary measure against target_boundary /
// boundary_strength, converted to Float32.
__device__ inline void mech_inputs(const CPM_CONST RepDev &d, int k, const CPM_GLOBAL CellRec *c, float *disp, float *strength) {
    if (d.mech_kind[k] == CPM_MECH_PRESSURE) {
        *disp = __fsub_rn((float)c->vol, (float)c->tv);
        *strength = (float)c->vs;
    } else {
        *disp = __fsub_rn((float)c->bnd, (float)c->tb);
        *strength = (float)c->bs;
    }
}

// One half-interval OU step of every mechanical state of every cell of one replica,
// before (phase 0) or after (phase 1) pass `pass` (_advance_checkerboard_mechanics!,
// scientific_mechanics.jl:170-196; lottery.jl:353-391 through _advance_mechanics!).
__global__ __launch_bounds__(kBlock) void mech_advance_kernel(const RepDev *reps, const McsPlan *plan, int slot,
                                                              uint32_t pass, uint32_t phase, uint32_t plan_k) {
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)reps)[slot];
    const CPM_CONST McsPlan &y = ((const CPM_CONST McsPlan *)plan)[(size_t)slot * kPlanMcs + plan_k];
    if (!y.active || pass >= (uint32_t)d.rounds) return;
    const uint32_t cell = 1u + blockIdx.x * kBlock + threadIdx.x;
    if (cell > d.capacity) return;
    const float half = d.algorithm == CPM_ALG_LOTTERY ? d.mech_half[0] : d.mech_half[y.order[pass]];
    CPM_GLOBAL CellRec *c = gp(d.cells) + cell;
    const bool active = c->active != 0;
    for (int k = 0; k < d.n_mech; ++k) {
        const bool pressure = d.mech_kind[k] == CPM_MECH_PRESSURE;
        float q = 0.0f;
        if (active) {
            float disp, strength;
            mech_inputs(d, k, c, &disp, &strength);
            const float normal = normal_box_muller(d.key_lo, d.key_hi, kStreamAuxEvolution, y.mcs, pass,
                                                   d.mech_instance[k], cell, phase);
            q = ou_transition(pressure ? c->qp : c->qt, disp, strength, d.mech_eta[k], d.mech_noise[k], half,
                              normal);
        }
        if (pressure) c->qp = q;
        else c->qt = q;
    }
}

// _initialize_mechanics! (scientific_mechanics.jl:63-103): inactive cells hold zero;
// constitutive mean 2 strength displacement, or the stationary draw around it.
__global__ __launch_bounds__(kBlock) void mech_init_kernel(const RepDev *reps, int slot) {
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)reps)[slot];
    const uint32_t cell = 1u + blockIdx.x * kBlock + threadIdx.x;
    if (cell > d.capacity) return;
    CPM_GLOBAL CellRec *c = gp(d.cells) + cell;
    const bool active = c->active != 0;
    for (int k = 0; k < d.n_mech; ++k) {
        const bool pressure = d.mech_kind[k] == CPM_MECH_PRESSURE;
        float q = pressure ? c->qp : c->qt;
        if (!active) {
            q = 0.0f;
        } else if (d.mech_init[k] != CPM_MECH_INIT_PRESERVE) {
            float disp, strength;
            mech_inputs(d, k, c, &disp, &strength);
            const float mean = __fmul_rn(__fmul_rn(2.0f, strength), disp);
            q = mean;
            if (d.mech_init[k] == CPM_MECH_INIT_STATIONARY) {
                const float normal = normal_box_muller(d.key_lo, d.key_hi, kStreamAuxInitialization, 0, 0,
                                                       d.mech_instance[k], cell, 0);
                q = __fadd_rn(mean, __fmul_rn(sqrt_rn(__fmul_rn(__fmul_rn(2.0f, strength), d.mech_noise[k])),
                                              normal));
            }
        }
        if (pressure) c->qp = q;
        else c->qt = q;
    }
}

// ---------------------------------------------------------------------------
// TiledCheckerboardCPM (tiled_checkerboard_device.jl:407-586, 688-705, 743-806)
// ---------------------------------------------------------------------------
// One thread per tile of the colour: the tile's proposals of one batch run in order, each
// against the batch snapshot of the trackers plus the tile's own earlier changes, and
// accepted copies write ownership at once (tiles of one colour lie beyond each other's
// halo, so no tile reads what another writes). The tile's changes are kept per touched
// cell in LDS, one column per thread, and summed into per-cell accumulators at the end;
// the reconcile kernel adds them to the trackers after the batch.
constexpr int kTiledBlock = 64;

// Packed (volume change << 16) + boundary change of one touched cell; both stay far inside
// 16 bits (at most 64 copies of at most 128 surface directions each).
__device__ __forceinline__ int32_t tiled_boundary_part(int32_t packed) { return (int32_t)(int16_t)(packed & 0xffff); }
__device__ __forceinline__ int32_t tiled_volume_part(int32_t packed) {
    return (packed - tiled_boundary_part(packed)) >> 16;
}

// The touched cells of one tile and their packed changes: one LDS column per lane (stride
// kTiledBlock), at most two entries per proposal.
constexpr uint32_t kListLds = 8;  // touched cells kept in LDS; the rest spill to global memory

struct TileList {
    int32_t *cell, *chg;              // LDS columns (stride kTiledBlock)
    CPM_GLOBAL int32_t *spill;        // (cell, change) pairs of entries kListLds onward
    uint32_t touched;

    __device__ __forceinline__ int find(int32_t c) const {
        const uint32_t nl = min(touched, kListLds);
        for (uint32_t j = 0; j < nl; ++j)
            if (cell[j * kTiledBlock] == c) return (int)j;
        for (uint32_t j = kListLds; j < touched; ++j)
            if (spill[2 * (j - kListLds)] == c) return (int)j;
        return -1;
    }
    __device__ __forceinline__ int32_t change(int j) const {
        return (uint32_t)j < kListLds ? chg[j * kTiledBlock] : spill[2 * ((uint32_t)j - kListLds) + 1];
    }
    __device__ __forceinline__ void add(int32_t c, int32_t dv, int32_t db) {
        int j = find(c);
        const int32_t q = dv * 65536 + db;
        if (j < 0) {
            j = (int)touched++;
            if ((uint32_t)j < kListLds) {
                cell[j * kTiledBlock] = c;
                chg[j * kTiledBlock] = q;
            } else {
                spill[2 * ((uint32_t)j - kListLds)] = c;
                spill[2 * ((uint32_t)j - kListLds) + 1] = q;
            }
        } else if ((uint32_t)j < kListLds) {
            chg[j * kTiledBlock] += q;
        } else {
            spill[2 * ((uint32_t)j - kListLds) + 1] += q;
        }
    }
    // the tile-local view of the attempt's cells (media carry no trackers)
    __device__ __forceinline__ View view(int32_t l, int32_t g) const {
        View v{0, 0, 0, 0};
        if (l > 0) {
            const int j = find(l);
            if (j >= 0) {
                const int32_t q = change(j);
                v.dvl = tiled_volume_part(q);
                v.dbl = tiled_boundary_part(q);
            }
        }
        if (g > 0) {
            const int j = find(g);
            if (j >= 0) {
                const int32_t q = change(j);
                v.dvg = tiled_volume_part(q);
                v.dbg = tiled_boundary_part(q);
            }
        }
        return v;
    }
    // the tile's changes join the colour batch's per-cell sums
    template <class D>
    __device__ __forceinline__ void flush(const D &d) const {
        for (uint32_t j = 0; j < touched; ++j) {
            const int32_t c = j < kListLds ? cell[j * kTiledBlock] : spill[2 * (j - kListLds)];
            const int32_t q = change((int)j);
            const int32_t dv = tiled_volume_part(q), db = tiled_boundary_part(q);
            if (dv != 0) atomicAdd(&d.tile_dvol[word_cell(c)], dv);
            if (db != 0) atomicAdd(&d.tile_dbnd[word_cell(c)], db);
        }
    }
};

// Wave sums of the per-lane proposal counters into the replica's MCS counters.
template <class D>
__device__ __forceinline__ void tiled_count(const D &d, uint32_t (&v)[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i)
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
    if (threadIdx.x == 0)
        for (int i = 0; i < 5; ++i)
            if (v[i]) atomicAdd(&d.tiled_counts[i], (unsigned long long)v[i]);
}

// Ownership change and tracker bookkeeping of an accepted tile copy.
template <class D>
__device__ __forceinline__ void tiled_commit(const D &d, TileList &list, int32_t l, int32_t g, const Staged &st) {
    if (l > 0) list.add(l, -1, st.dl);
    else atomicAdd(&d.medium_vol[medium_slot(d, l)], -1);
    if (g > 0) list.add(g, 1, st.dg);
    else atomicAdd(&d.medium_vol[medium_slot(d, g)], 1);
}

// Any domain: recipients from the tile's site list, the lattice read in global memory.
__global__ __launch_bounds__(kTiledBlock) void tiled_batch_kernel(const RepDev *reps, const RepDyn *dyn, int slot,
                                                                  uint32_t ordinal, uint32_t batch) {
    extern __shared__ int32_t s_touched[];  // kListLds cell ids, then as many packed changes
    // descriptor and MCS state are not written by this kernel: constant address space, so
    // they stay in scalar loads however the lattice is written
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)reps)[slot];
    const CPM_CONST RepDyn &y = ((const CPM_CONST RepDyn *)dyn)[slot];
    if (!y.active) return;
    const uint32_t color = y.order[ordinal];
    const uint32_t t0 = d.color_offsets[color - 1], t1 = d.color_offsets[color];
    const uint32_t k = t0 + blockIdx.x * kTiledBlock + threadIdx.x;
    const uint32_t interval = (uint32_t)d.tiled_interval;
    uint32_t v[5] = {0, 0, 0, 0, 0};  // realized, same owner, boundary null, rejected, accepted
    const bool have = k < t1;
    const uint32_t tile = have ? d.color_tiles[k] : 1u;
    const uint32_t first = d.tile_offsets[tile - 1], count = d.tile_offsets[tile] - first;
    const uint32_t start = batch * interval + 1;
    const uint32_t stop = have ? min(start + interval - 1, count) : 0u;
    const uint32_t sub = ordinal * (uint32_t)d.tiled_batches + batch + 1;
    TileList list{s_touched + threadIdx.x, s_touched + kListLds * kTiledBlock + threadIdx.x,
                  gp(d.tile_spill) + (size_t)(have ? k - t0 : 0u) * (uint32_t)d.tile_spill_stride, 0};
    CPM_GLOBAL int32_t *lat = gp(d.lattice);
    const uint32_t X = (uint32_t)d.dims[0], XY = X * (uint32_t)d.dims[1];
    // Each lane advances through its proposals to the next realized one (no-ops are counted on
    // the way), then the wave evaluates all lanes' realized proposals together: evaluation
    // passes follow the most realized proposals of any lane, not the proposal count. A lane's
    // proposals keep their order, each seeing the lane's earlier copies.
    uint32_t lp = start, rec = 0;
    int32_t l = 0, g = 0;
    bool pending = false;
    while (true) {
        while (!pending && lp <= stop) {
            const uint32_t sel = bounded_uint(d.key_lo, d.key_hi, kStreamTiledProposal, y.mcs, sub, lp, kSiteEntity,
                                              tile, 0, count);
            rec = d.tile_sites[first + sel];
            const uint32_t dir = bounded_uint(d.key_lo, d.key_hi, kStreamTiledProposal, y.mcs, sub, lp, kSiteEntity,
                                              tile, 1, (uint32_t)d.prop.n);
            const uint32_t idx0 = rec - 1;
            const int z = (int)(idx0 / XY), yy = (int)((idx0 % XY) / X), x = (int)(idx0 % X);
            const Realized donor = realize_position(d, x + d.prop.off[dir][0], yy + d.prop.off[dir][1],
                                                    z + d.prop.off[dir][2]);
            l = lat[idx0];
            g = donor.word;
            if (donor.site == 0 || !donor.mut) {
                ++v[2];
                ++lp;
            } else if (g == l) {
                ++v[1];
                ++lp;
            } else {
                pending = true;
            }
        }
        if (__ballot(pending) == 0) break;
        if (pending) {
            ++v[0];
            Staged st;
            const float p = metropolis_p(d.temperature, copy_energy(d, rec, l, g, list.view(l, g), &st));
            const Words w = rng_words(d.key_lo, d.key_hi, kStreamTiledProposal, y.mcs, sub, lp, kSiteEntity, tile, 0, 2);
            if (uniform_open01(w.w0) < p) {
                ++v[4];
                lat[rec - 1] = g;
                tiled_commit(d, list, l, g, st);
            } else {
                ++v[3];
            }
            pending = false;
            ++lp;
        }
    }
    if (have) list.flush(d);
    tiled_count(d, v);
}

// LDS bytes of the staged kernel before the per-lane columns: the contact matrix (Float32
// and Float64) and the proposal directions as offsets inside a staged box.
constexpr int kTiledStagedHead = CPM_MAX_TYPES * CPM_MAX_TYPES * (4 + 8) + CPM_MAX_DIRS * 4;

// Tiles whose box plus halo holds at most 64 positions, none of them the same site twice: the
// lane stages its box and halo in LDS once (owner words; a bit mask of the positions a
// proposal may copy from), reads every neighbour there and writes its own copies there and
// to the lattice. Global reads per realized proposal: the two cell records and the contact
// neighbours' types. Recipients of an obstacle-free tile come from the box arithmetic.
__global__ __launch_bounds__(kTiledBlock) void tiled_staged_kernel(const RepDev *reps, const RepDyn *dyn, int slot,
                                                                   uint32_t ordinal, uint32_t batch) {
    extern __shared__ int32_t s_mem[];
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)reps)[slot];
    const CPM_CONST RepDyn &y = ((const CPM_CONST RepDyn *)dyn)[slot];
    if (!y.active) return;
    float *s_jf = (float *)s_mem;
    double *s_jd = (double *)(s_mem + CPM_MAX_TYPES * CPM_MAX_TYPES);
    int32_t *s_pd = s_mem + CPM_MAX_TYPES * CPM_MAX_TYPES * 3;
    int32_t *s_lane = s_mem + kTiledStagedHead / 4;
    const int SX = d.tiled_stage[0], SY = d.tiled_stage[1], SZ = d.tiled_stage[2];
    const int H = d.tiled_halo, HZ = d.ndim == 3 ? H : 0;
    for (int i = threadIdx.x; i < d.n_types * d.n_types; i += kTiledBlock) {
        s_jf[i] = d.Jf[i];
        s_jd[i] = d.Jd[i];
    }
    for (int i = threadIdx.x; i < d.prop.n; i += kTiledBlock)
        s_pd[i] = d.prop.off[i][0] + SX * (d.prop.off[i][1] + SY * d.prop.off[i][2]);
    __syncthreads();
    const uint32_t color = y.order[ordinal];
    const uint32_t t0 = d.color_offsets[color - 1], t1 = d.color_offsets[color];
    const uint32_t k = t0 + blockIdx.x * kTiledBlock + threadIdx.x;
    const uint32_t interval = (uint32_t)d.tiled_interval;
    const uint32_t S = (uint32_t)(SX * SY * SZ);
    uint32_t v[5] = {0, 0, 0, 0, 0};  // realized, same owner, boundary null, rejected, accepted
    const bool have = k < t1;
    const uint32_t tile = have ? d.color_tiles[k] : 1u;
    const uint32_t first = d.tile_offsets[tile - 1], count = d.tile_offsets[tile] - first;
    const uint32_t start = batch * interval + 1;
    const uint32_t stop = have ? min(start + interval - 1, count) : 0u;
    const uint32_t sub = ordinal * (uint32_t)d.tiled_batches + batch + 1;
    const int X = d.dims[0], Y = d.dims[1];
    const uint32_t tc = tile - 1, gx = (uint32_t)d.tile_grid[0], gy = (uint32_t)d.tile_grid[1];
    const int ox = (int)(tc % gx) * d.tile_edge[0], oy = (int)((tc / gx) % gy) * d.tile_edge[1],
              oz = (int)(tc / (gx * gy)) * d.tile_edge[2];
    const int bx = min(d.tile_edge[0], X - ox), by = min(d.tile_edge[1], Y - oy);
    int32_t *s_w = s_lane + threadIdx.x;
    TileList list{s_lane + S * kTiledBlock + threadIdx.x, s_lane + (S + kListLds) * kTiledBlock + threadIdx.x,
                  gp(d.tile_spill) + (size_t)(have ? k - t0 : 0u) * (uint32_t)d.tile_spill_stride, 0};
    // stage the box and its halo: eight positions' loads in flight at a time
    uint64_t donors = 0;
    if (have) {
        for (uint32_t p0 = 0; p0 < S; p0 += 8) {
            int32_t w[8];
            bool m[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t p = p0 + j;
                w[j] = 0;
                m[j] = false;
                if (p < S) {
                    const int sx = (int)(p % (uint32_t)SX), sy = (int)((p / (uint32_t)SX) % (uint32_t)SY),
                              sz = (int)(p / (uint32_t)(SX * SY));
                    const Realized rz = realize_position(d, ox - H + sx, oy - H + sy, oz - HZ + sz);
                    w[j] = rz.word;
                    m[j] = rz.site != 0 && rz.mut;
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (p0 + j < S) {
                    s_w[(p0 + j) * kTiledBlock] = w[j];
                    if (m[j]) donors |= 1ull << (p0 + j);
                }
        }
    }
    const int ns = d.surf.n, nc = d.has_contact ? d.cont.n : 0;
    auto delta = [&](const auto *o) __attribute__((always_inline)) { return o[0] + SX * (o[1] + SY * o[2]); };
    CPM_GLOBAL int32_t *lat = gp(d.lattice);
    // Each lane advances through its proposals to the next realized one (no-ops are counted on
    // the way), then the wave evaluates all lanes' realized proposals together: evaluation
    // passes follow the most realized proposals of any lane, not the proposal count. A lane's
    // proposals keep their order, each seeing the lane's earlier copies.
    uint32_t lp = start, rec = 0;
    int pr = 0;
    int32_t l = 0, g = 0;
    bool pending = false;
    while (true) {
        while (!pending && lp <= stop) {
            const uint32_t sel = bounded_uint(d.key_lo, d.key_hi, kStreamTiledProposal, y.mcs, sub, lp, kSiteEntity,
                                              tile, 0, count);
            int lx, ly, lz;
            if (d.mask == nullptr) {
                // the tile's sites in order, first axis fastest
                lx = (int)(sel % (uint32_t)bx);
                ly = (int)((sel / (uint32_t)bx) % (uint32_t)by);
                lz = (int)(sel / (uint32_t)(bx * by));
                rec = 1u + (uint32_t)(ox + lx) + (uint32_t)X * ((uint32_t)(oy + ly) + (uint32_t)Y * (uint32_t)(oz + lz));
            } else {
                rec = d.tile_sites[first + sel];
                const uint32_t i0 = rec - 1, XY = (uint32_t)X * (uint32_t)Y;
                lz = (int)(i0 / XY) - oz;
                ly = (int)((i0 % XY) / (uint32_t)X) - oy;
                lx = (int)(i0 % (uint32_t)X) - ox;
            }
            const uint32_t dir = bounded_uint(d.key_lo, d.key_hi, kStreamTiledProposal, y.mcs, sub, lp, kSiteEntity,
                                              tile, 1, (uint32_t)d.prop.n);
            pr = (lx + H) + SX * ((ly + H) + SY * (lz + HZ));
            const int pd = pr + s_pd[dir];
            l = s_w[pr * kTiledBlock];
            g = s_w[pd * kTiledBlock];
            if (!((donors >> pd) & 1ull)) {
                ++v[2];
                ++lp;
            } else if (g == l) {
                ++v[1];
                ++lp;
            } else {
                pending = true;
            }
        }
        if (__ballot(pending) == 0) break;
        if (pending) {
            ++v[0];
            const uint32_t li = cell_slot(l), gi = cell_slot(g);
            const CellHot cl = load_cell(gp(d.cells), li), cg = load_cell(gp(d.cells), gi);
            int32_t sw[kNb], cw[kNb];
            uint32_t ct[kNb];
#pragma unroll
            for (int u = 0; u < kNb; ++u) sw[u] = u < ns ? s_w[(pr + delta(d.surf.off[u])) * kTiledBlock] : 0;
#pragma unroll
            for (int u = 0; u < kNb; ++u) cw[u] = u < nc ? s_w[(pr + delta(d.cont.off[u])) * kTiledBlock] : 0;
#pragma unroll
            for (int u = 0; u < kNb; ++u) ct[u] = cw[u] > 0 ? word_type(cw[u]) : 0u;
            auto more = [&](int role, int q) __attribute__((always_inline)) {
                return s_w[(pr + delta(role == 0 ? d.surf.off[q] : d.cont.off[q])) * kTiledBlock];
            };
            Staged st;
            const float dd = copy_energy_from(
                d, l, g, list.view(l, g), cl, cg, 0.0f, 0.0f, 0.0f, 0.0f, sw, cw, ct, more,
                [&](uint32_t i) __attribute__((always_inline)) { return s_jf[i]; },
                [&](uint32_t i) __attribute__((always_inline)) { return s_jd[i]; }, &st);
            const float p = metropolis_p(d.temperature, dd);
            const Words w = rng_words(d.key_lo, d.key_hi, kStreamTiledProposal, y.mcs, sub, lp, kSiteEntity, tile, 0, 2);
            if (uniform_open01(w.w0) < p) {
                ++v[4];
                lat[rec - 1] = g;
                s_w[pr * kTiledBlock] = g;
                tiled_commit(d, list, l, g, st);
            } else {
                ++v[3];
            }
            pending = false;
            ++lp;
        }
    }
    if (have) list.flush(d);
    tiled_count(d, v);
}

// After each colour batch: trackers = snapshot + the summed tile changes, accumulators
// cleared, and active cells left with zero volume retire (tiled_checkerboard_device.jl:
// 688-705; _retire_zero_volume_cells!).
__global__ __launch_bounds__(kBlock) void tiled_reconcile_kernel(const RepDev *reps, const RepDyn *dyn, int slot) {
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)reps)[slot];
    if (!((const CPM_CONST RepDyn *)dyn)[slot].active) return;
    const uint32_t cell = 1u + blockIdx.x * kBlock + threadIdx.x;
    if (cell > d.capacity) return;
    CPM_GLOBAL CellRec *c = gp(d.cells) + cell;
    const int32_t dv = d.tile_dvol[cell], db = d.tile_dbnd[cell];
    int32_t vol = c->vol;
    if ((dv | db) != 0) {
        d.tile_dvol[cell] = 0;
        d.tile_dbnd[cell] = 0;
        vol += dv;
        c->vol = vol;
        c->bnd = c->bnd + db;
    }
    if (vol == 0 && c->active) {
        c->active = 0;
        c->ctype = 0;
        c->tv = d.retire_tv;
        c->vs = d.retire_vs;
        c->tb = d.retire_tb;
        c->bs = d.retire_bs;
        c->qp = 0.0f;
        c->qt = 0.0f;
    }
}

// Starts one MCS for every replica below its target: mcs += 1 and the colour / round
// order of that MCS (checkerboard.jl:226-242, lottery.jl:124-140).
// The colour / round order of MCS m of replica d (checkerboard.jl:226-242, lottery.jl:124-140,
// tiled_checkerboard_device.jl:743-760).
template <class T>
__device__ inline void mcs_order(const RepDev &d, uint64_t m, T *order) {
    if (d.algorithm == CPM_ALG_LOTTERY) schedule_order(d.key_lo, d.key_hi, kStreamLotteryActivation, 1, m, d.rounds, order);
    else if (d.algorithm == CPM_ALG_TILED) schedule_order(d.key_lo, d.key_hi, kStreamTiledOrder, 0, m, d.rounds, order);
    else schedule_order(d.key_lo, d.key_hi, kStreamCheckerboardOrder, 0, m, d.rounds, order);
}

// Starts one MCS for every replica below its target: mcs += 1 and the colour / round order of
// that MCS, also as plan 0 for the pass kernels; TiledCheckerboardCPM counters are reset.
__global__ void begin_mcs_kernel(const RepDev *reps, RepDyn *dyn, McsPlan *plan, int nslots) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nslots) return;
    RepDyn &y = dyn[r];
    McsPlan &p = plan[(size_t)r * kPlanMcs];
    if (!y.in_use || y.mcs >= y.target) {
        y.active = 0;
        p.active = 0;
        return;
    }
    const uint64_t m = y.mcs + 1;
    y.mcs = m;
    y.active = 1;
    p.mcs = m;
    p.active = 1;
    const RepDev &d = reps[r];
    mcs_order(d, m, y.order);
    for (int i = 0; i < d.rounds; ++i) p.order[i] = (uint16_t)y.order[i];
    if (d.algorithm == CPM_ALG_TILED)
        for (int i = 0; i < 5; ++i) d.tiled_counts[i] = 0;
}

// Plans of a chunk of K MCS (K <= kPlanMcs), one thread per (replica, MCS): the k-th MCS is
// mcs + k + 1 and runs while the replica is in use and below its target. The counters stay
// untouched until plan_end_kernel.
__global__ void plan_kernel(const RepDev *reps, const RepDyn *dyn, McsPlan *plan, int nslots, int K) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots * K) return;
    const int r = i / K, k = i % K;
    const RepDyn &y = dyn[r];
    McsPlan &p = plan[(size_t)r * kPlanMcs + k];
    const bool run = y.in_use && y.mcs + (uint64_t)k < y.target;
    const uint64_t m = y.mcs + (uint64_t)k + 1;
    p.mcs = m;
    p.active = run ? 1u : 0u;
    if (run) mcs_order(reps[r], m, p.order);
}

// End of a planned chunk of K MCS: every replica's counter advances by the MCS it performed,
// and its last performed MCS's order becomes the current one.
__global__ void plan_end_kernel(const RepDev *reps, RepDyn *dyn, const McsPlan *plan, int nslots, int K) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nslots) return;
    RepDyn &y = dyn[r];
    const uint64_t done = y.in_use && y.target > y.mcs ? min((uint64_t)K, y.target - y.mcs) : 0;
    y.active = done > 0 ? 1u : 0u;
    if (done == 0) return;
    const McsPlan &p = plan[(size_t)r * kPlanMcs + (done - 1)];
    y.mcs += done;
    for (int i = 0; i < reps[r].rounds; ++i) y.order[i] = p.order[i];
}

// Sets the host-controlled fields of one replica's dynamic state.
__global__ void set_dyn_kernel(RepDyn *dyn, int slot, uint64_t mcs, uint64_t target,
                               uint32_t in_use, uint32_t mask) {
    RepDyn &y = dyn[slot];
    if (mask & 1u) y.mcs = mcs;
    if (mask & 2u) y.target = target;
    if (mask & 4u) y.in_use = in_use;
    if (mask & 8u) y.active = 0;
}

// Sums the accounting slots of one replica over the passes of its last MCS.
__global__ __launch_bounds__(kBlock) void report_kernel(const RepDev *reps, const uint32_t *slots,
                                                        uint32_t ntiles, int slot,
                                                        unsigned long long *out) {
    __shared__ unsigned long long s_acc[7][kBlock];
    __shared__ double s_dh[kBlock];
    const RepDev &d = reps[slot];
    unsigned long long acc[7] = {0, 0, 0, 0, 0, 0, 0};
    double dh = 0.0;
    const uint32_t per_pass = d.tile_count;
    const uint32_t total = per_pass * (uint32_t)d.rounds;
    for (uint32_t i = threadIdx.x; i < total; i += kBlock) {
        const uint32_t p = i / per_pass, t = d.tile_base + i % per_pass;
        const uint32_t *s = slots + ((size_t)p * ntiles + t) * kSlotFields;
        for (int k = 0; k < 7; ++k) acc[k] += s[k];
        dh += *(const double *)(s + 8);
    }
    for (int k = 0; k < 7; ++k) s_acc[k][threadIdx.x] = acc[k];
    s_dh[threadIdx.x] = dh;
    __syncthreads();
    if (threadIdx.x < 7) {
        unsigned long long t = 0;
        for (int j = 0; j < kBlock; ++j) t += s_acc[threadIdx.x][j];
        out[threadIdx.x] = t;
    } else if (threadIdx.x == 7) {
        double t = 0.0;  // fixed order: the result does not depend on scheduling
        for (int j = 0; j < kBlock; ++j) t += s_dh[j];
        out[7] = (unsigned long long)__double_as_longlong(t);
    }
}

// ---------------------------------------------------------------------------
// Global energy in FP64 (scientific_components.jl:97-107, 315-340, 525-538): over active
// cells lambda_V (V - V*)^2 + lambda_S (S - S*)^2 from the trackers, and over mutable sites
// and realized contact directions w J[type(owner), type(neighbour)] where the owners differ,
// a pair of mutable sites once (the direction not after its opposite). A fixed grid of
// kEnergyParts workgroups writes one partial each; energy_sum_kernel adds them in a fixed
// order, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void energy_partials_kernel(const RepDev *reps, int slot, double *partials) {
    __shared__ double s_part[kBlock];
    __shared__ uint8_t s_keep[CPM_MAX_DIRS];  // direction k is not after its opposite
    const CPM_CONST RepDev &d = ((const CPM_CONST RepDev *)reps)[slot];
    if ((int)threadIdx.x < d.cont.n) {
        const int k = (int)threadIdx.x;
        int opp = k;  // index of the opposite offset (k itself when the relation lacks it)
        for (int e = 0; e < d.cont.n; ++e)
            if (d.cont.off[e][0] == -d.cont.off[k][0] && d.cont.off[e][1] == -d.cont.off[k][1] &&
                d.cont.off[e][2] == -d.cont.off[k][2])
                opp = e;
        s_keep[k] = k <= opp ? 1 : 0;
    }
    __syncthreads();
    bool has_vol = false, has_con = false, has_bnd = false;
    for (int t = 0; t < d.n_energies; ++t) {
        has_vol |= d.energy_order[t] == CPM_TERM_VOLUME;
        has_con |= d.energy_order[t] == CPM_TERM_CONTACT;
        has_bnd |= d.energy_order[t] == CPM_TERM_BOUNDARY;
    }
    const uint32_t first = blockIdx.x * kBlock + threadIdx.x, stride = gridDim.x * kBlock;
    double acc = 0.0;
    if (has_vol || has_bnd) {
        for (uint32_t c = 1 + first; c <= d.capacity; c += stride) {
            const CellRec &r = d.cells[c];
            if (!r.active) continue;
            if (has_vol) {
                const double dv = (double)r.vol - r.tv;
                acc += r.vs * (dv * dv);
            }
            if (has_bnd) {
                const double db = (double)r.bnd - (double)r.tb;
                acc += r.bs * (db * db);
            }
        }
    }
    if (has_con && d.simple) {
        // periodic lattice without obstacles: every position is a mutable site, neighbours wrap
        // once; eight neighbour loads in flight per step
        const int nc = d.cont.n;
        const uint32_t nt = (uint32_t)d.n_types;
        const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
        const CPM_GLOBAL int32_t *lat = gp(d.lattice);
        for (uint32_t i = first; i < d.nsites; i += stride) {
            const int x = (int)(i % (uint32_t)X);
            const int y = (int)((i / (uint32_t)X) % (uint32_t)Y);
            const int z = (int)(i / ((uint32_t)X * (uint32_t)Y));
            const int32_t o = lat[i];
            const uint32_t ot = owner_type(d, o);
            for (int k0 = 0; k0 < nc; k0 += 8) {
                int32_t w[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = k0 + j;
                    w[j] = o;
                    if (k < nc) {
                        int gx = x + d.cont.off[k][0], gy = y + d.cont.off[k][1], gz = z + d.cont.off[k][2];
                        gx += gx < 0 ? X : 0;
                        gx -= gx >= X ? X : 0;
                        gy += gy < 0 ? Y : 0;
                        gy -= gy >= Y ? Y : 0;
                        gz += gz < 0 ? Z : 0;
                        gz -= gz >= Z ? Z : 0;
                        w[j] = lat[(uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz)];
                    }
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = k0 + j;
                    if (k < nc && w[j] != 0 && w[j] != o && s_keep[k])
                        acc += d.cont.wd[k] * d.Jd[(ot - 1) * nt + (owner_type(d, w[j]) - 1)];
                }
            }
        }
    } else if (has_con) {
        const int nc = d.cont.n;
        const uint32_t nt = (uint32_t)d.n_types;
        for (uint32_t i = first; i < d.nsites; i += stride) {
            if (d.mask && !d.mask[i]) continue;
            const int x = (int)(i % (uint32_t)d.dims[0]);
            const int y = (int)((i / (uint32_t)d.dims[0]) % (uint32_t)d.dims[1]);
            const int z = (int)(i / ((uint32_t)d.dims[0] * (uint32_t)d.dims[1]));
            const int32_t o = d.lattice[i];
            const uint32_t ot = owner_type(d, o);
            for (int k = 0; k < nc; ++k) {
                const Realized nb = realize_position(d, x + d.cont.off[k][0], y + d.cont.off[k][1],
                                                     z + d.cont.off[k][2]);
                if (nb.word == 0 || nb.word == o) continue;
                if (nb.site && nb.mut && !s_keep[k]) continue;  // the pair is counted from the other site
                acc += d.cont.wd[k] * d.Jd[(ot - 1) * nt + (owner_type(d, nb.word) - 1)];
            }
        }
    }
    s_part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int j = 0; j < kBlock; ++j) t += s_part[j];
        partials[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void energy_sum_kernel(const double *partials, uint32_t n, double *out) {
    __shared__ double s_part[kBlock];
    double t = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) t += partials[i];
    s_part[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int j = 0; j < kBlock; ++j) sum += s_part[j];
        out[0] = sum;
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// generic == 0: the proposal relation has 2 * DIM directions (offsets in registers);
// plain: lottery winners are decided from the tickets alone (see lot_propose_kernel); 2: and
// the conflict shell is the radius-1 Moore shell
template <int DIM, int H, int TX, int TY, int TZ, int NT = kBlock>
static hipError_t launch_propose_t(int algorithm, int generic, int plain, const PassArgs &a, hipStream_t stream) {
    constexpr int KF = 2 * DIM;
    if (algorithm == CPM_ALG_LOTTERY) {
        if (generic && plain)
            CPM_LAUNCH((lot_propose_kernel<DIM, H, TX, TY, TZ, 0, NT, true>), dim3(a.ntiles), dim3(NT), 0,
                               stream, a);
        else if (generic)
            CPM_LAUNCH((lot_propose_kernel<DIM, H, TX, TY, TZ, 0, NT, false>), dim3(a.ntiles), dim3(NT), 0,
                               stream, a);
        else if (plain == 2 && H == 1)
            CPM_LAUNCH((lot_propose_kernel<DIM, H, TX, TY, TZ, KF, NT, true, H == 1>), dim3(a.ntiles), dim3(NT), 0,
                               stream, a);
        else if (plain)
            CPM_LAUNCH((lot_propose_kernel<DIM, H, TX, TY, TZ, KF, NT, true>), dim3(a.ntiles), dim3(NT), 0,
                               stream, a);
        else
            CPM_LAUNCH((lot_propose_kernel<DIM, H, TX, TY, TZ, KF, NT, false>), dim3(a.ntiles), dim3(NT), 0,
                               stream, a);
    } else {
        if (generic)
            CPM_LAUNCH((cb_propose_kernel<DIM, H, TX, TY, TZ, 0, NT>), dim3(a.ntiles), dim3(NT), 0,
                               stream, a);
        else
            CPM_LAUNCH((cb_propose_kernel<DIM, H, TX, TY, TZ, KF, NT>), dim3(a.ntiles), dim3(NT), 0,
                               stream, a);
    }
    return hipGetLastError();
}

constexpr int kCommitWaves = 2;  // claim check on both waves, winners evaluated in full chunks
template <int CT>
static void launch_commit_ct(const PassArgs &a, hipStream_t stream) {
    const uint32_t wgs = (a.ntiles + CT - 1) / CT;
    switch (a.fast_eval) {
#define CPM_COMMIT_CASE(F)                                                                              \
    case F:                                                                                            \
        CPM_LAUNCH((commit_kernel<kCommitWaves, CT, F>), dim3(wgs), dim3(64 * kCommitWaves), 0, stream, a); \
        break;
        CPM_COMMIT_CASE(1)
        CPM_COMMIT_CASE(2)
        CPM_COMMIT_CASE(3)
        CPM_COMMIT_CASE(4)
        CPM_COMMIT_CASE(5)
        CPM_COMMIT_CASE(6)
#undef CPM_COMMIT_CASE
    default:
        CPM_LAUNCH((commit_kernel<kCommitWaves, CT, 0>), dim3(wgs), dim3(64 * kCommitWaves), 0, stream, a);
    }
}

static hipError_t launch_commit(const PassArgs &a, hipStream_t stream) {
    if (a.seg_cap >= 4096) launch_commit_ct<2>(a, stream);
    else if (a.fast_eval == 6) launch_commit_ct<8>(a, stream);  // 26 neighbours: two waves per SIMD fit
    else launch_commit_ct<kCommitTiles>(a, stream);
    return hipGetLastError();
}

}  // namespace cpm

// Tile configurations: id & 15 -> (dim, halo, TX, TY, TZ); bit 4 of id selects the kernels
// for proposal relations whose direction count differs from 2 * dim; bit 5 the row kernel
// of 2D periodic von Neumann checkerboards with a 2 x 2 colour pattern (with bit 6: an x
// side of 2 mod 4, rows of 8-byte halves); bit 6 without bit 5 the lottery kernel for
// periodic obstacle-free lattices of at least 2 halo + 1 sites per axis; bit 7, with bit 6,
// its winner test over the radius-1 Moore conflict shell.
extern "C" int cpm_internal_tile_config(int id, int *dim, int *halo, int *tx, int *ty, int *tz) {
    if (id < 0 || id > 255) return 0;
    id &= 15;
    static const int table[][5] = {
        {2, 1, 128, 16, 1}, {2, 2, 128, 16, 1}, {3, 1, 32, 8, 8}, {3, 2, 32, 8, 8},
        {2, 1, 64, 32, 1},  {2, 1, 256, 8, 1},  {2, 1, 64, 16, 1}, {2, 1, 32, 32, 1},
        {3, 1, 64, 8, 4},   {3, 1, 32, 16, 4},
        // one-wave workgroups
        {2, 1, 32, 16, 1},  {2, 1, 64, 8, 1},
        // 4096-site tiles (two row strips per wave in the row kernel)
        {2, 1, 256, 16, 1}, {2, 1, 128, 32, 1},
    };
    const int n = (int)(sizeof(table) / sizeof(table[0]));
    if (id < 0 || id >= n) return 0;
    *dim = table[id][0];
    *halo = table[id][1];
    *tx = table[id][2];
    *ty = table[id][3];
    *tz = table[id][4];
    return 1;
}

namespace cpm {
template <int TX, int TY>
static void launch_rows(bool half, const PassArgs *a, hipStream_t stream) {
    if (half)
        CPM_LAUNCH((cb_propose_rows_kernel<TX, TY, true>), dim3(a->ntiles), dim3(kBlock), 0, stream, *a);
    else
        CPM_LAUNCH((cb_propose_rows_kernel<TX, TY, false>), dim3(a->ntiles), dim3(kBlock), 0, stream, *a);
}
}  // namespace cpm

extern "C" hipError_t cpm_internal_launch_propose(int config, int algorithm, const cpm::PassArgs *a,
                                                  hipStream_t stream) {
    using namespace cpm;
    if (config < 0 || config > 255) return hipErrorInvalidValue;
    if (config & 32) {
        // row kernels: checkerboard, periodic von Neumann proposal, 2 x 2 (x 2) colour pattern
        if (algorithm != CPM_ALG_CHECKERBOARD || (config & 16) || (config & 128)) return hipErrorInvalidValue;
        const bool half = (config & 64) != 0;  // 2D rows of 8-byte halves (x side 2 mod 4)
        switch (config & 15) {
        case 0:
        case 1: launch_rows<128, 16>(half, a, stream); break;
        case 4: launch_rows<64, 32>(half, a, stream); break;
        case 5: launch_rows<256, 8>(half, a, stream); break;
        case 6: launch_rows<64, 16>(half, a, stream); break;
        case 7: launch_rows<32, 32>(half, a, stream); break;
        case 12: launch_rows<256, 16>(half, a, stream); break;
        case 13: launch_rows<128, 32>(half, a, stream); break;
        // 3D: 2 x 2 x 2 colour pattern, six von Neumann directions
        case 2: CPM_LAUNCH((cb_propose_rows3_kernel<32, 8, 8>), dim3(a->ntiles), dim3(kBlock), 0, stream, *a); break;
        case 8: CPM_LAUNCH((cb_propose_rows3_kernel<64, 8, 4>), dim3(a->ntiles), dim3(kBlock), 0, stream, *a); break;
        case 9: CPM_LAUNCH((cb_propose_rows3_kernel<32, 16, 4>), dim3(a->ntiles), dim3(kBlock), 0, stream, *a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const int g = (config >> 4) & 1;
    const int plain = (config >> 6) & 1 ? 1 + ((config >> 7) & 1) : 0;  // 2: Moore conflict shell
    if (plain && algorithm != CPM_ALG_LOTTERY) return hipErrorInvalidValue;
    if ((config & 128) && (g || !plain)) return hipErrorInvalidValue;
    switch (config & 15) {
    case 0: return launch_propose_t<2, 1, 128, 16, 1>(algorithm, g, plain, *a, stream);
    case 1: return launch_propose_t<2, 2, 128, 16, 1>(algorithm, g, plain, *a, stream);
    case 2: return launch_propose_t<3, 1, 32, 8, 8>(algorithm, g, plain, *a, stream);
    case 3: return launch_propose_t<3, 2, 32, 8, 8>(algorithm, g, plain, *a, stream);
    case 4: return launch_propose_t<2, 1, 64, 32, 1>(algorithm, g, plain, *a, stream);
    case 5: return launch_propose_t<2, 1, 256, 8, 1>(algorithm, g, plain, *a, stream);
    case 6: return launch_propose_t<2, 1, 64, 16, 1>(algorithm, g, plain, *a, stream);
    case 7: return launch_propose_t<2, 1, 32, 32, 1>(algorithm, g, plain, *a, stream);
    case 8: return launch_propose_t<3, 1, 64, 8, 4>(algorithm, g, plain, *a, stream);
    case 9: return launch_propose_t<3, 1, 32, 16, 4>(algorithm, g, plain, *a, stream);
    case 10: return launch_propose_t<2, 1, 32, 16, 1, 64>(algorithm, g, plain, *a, stream);
    case 11: return launch_propose_t<2, 1, 64, 8, 1, 64>(algorithm, g, plain, *a, stream);
    case 12: return launch_propose_t<2, 1, 256, 16, 1>(algorithm, g, plain, *a, stream);
    case 13: return launch_propose_t<2, 1, 128, 32, 1>(algorithm, g, plain, *a, stream);
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t cpm_internal_launch_commit(const cpm::PassArgs *a, hipStream_t stream) {
    return cpm::launch_commit(*a, stream);
}

extern "C" hipError_t cpm_internal_launch_pass(int config, int algorithm, const cpm::PassArgs *a,
                                               hipStream_t stream) {
    hipError_t err = cpm_internal_launch_propose(config, algorithm, a, stream);
    if (err != hipSuccess) return err;
    return cpm::launch_commit(*a, stream);
}

extern "C" hipError_t cpm_internal_launch_begin(const cpm::RepDev *reps, cpm::RepDyn *dyn, cpm::McsPlan *plan,
                                                int nslots, hipStream_t stream) {
    const int blocks = (nslots + 63) / 64;
    CPM_LAUNCH(cpm::begin_mcs_kernel, dim3(blocks), dim3(64), 0, stream, reps, dyn, plan, nslots);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_plan(const cpm::RepDev *reps, const cpm::RepDyn *dyn, cpm::McsPlan *plan,
                                               int nslots, int K, hipStream_t stream) {
    if (K < 1 || K > cpm::kPlanMcs) return hipErrorInvalidValue;
    const int blocks = (nslots * K + 63) / 64;
    CPM_LAUNCH(cpm::plan_kernel, dim3(blocks), dim3(64), 0, stream, reps, dyn, plan, nslots, K);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_plan_end(const cpm::RepDev *reps, cpm::RepDyn *dyn,
                                                   const cpm::McsPlan *plan, int nslots, int K, hipStream_t stream) {
    const int blocks = (nslots + 63) / 64;
    CPM_LAUNCH(cpm::plan_end_kernel, dim3(blocks), dim3(64), 0, stream, reps, dyn, plan, nslots, K);
    return hipGetLastError();
}

// staged = positions of a tile's staged box (0: the kernel reading the lattice in global memory)
extern "C" hipError_t cpm_internal_launch_tiled_batch(const cpm::RepDev *reps, const cpm::RepDyn *dyn, int slot,
                                                      uint32_t ordinal, uint32_t batch, uint32_t max_tiles,
                                                      uint32_t interval, uint32_t staged, hipStream_t stream) {
    const uint32_t blocks = (max_tiles + cpm::kTiledBlock - 1) / cpm::kTiledBlock;
    if (blocks == 0) return hipSuccess;
    (void)interval;
    if (staged) {
        const size_t lds = cpm::kTiledStagedHead + (size_t)(staged + 2 * cpm::kListLds) * cpm::kTiledBlock * sizeof(int32_t);
        CPM_LAUNCH(cpm::tiled_staged_kernel, dim3(blocks), dim3(cpm::kTiledBlock), lds, stream, reps, dyn,
                           slot, ordinal, batch);
    } else {
        const size_t lds = (size_t)2 * cpm::kListLds * cpm::kTiledBlock * sizeof(int32_t);
        CPM_LAUNCH(cpm::tiled_batch_kernel, dim3(blocks), dim3(cpm::kTiledBlock), lds, stream, reps, dyn,
                           slot, ordinal, batch);
    }
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_tiled_reconcile(const cpm::RepDev *reps, const cpm::RepDyn *dyn, int slot,
                                                          uint32_t capacity, hipStream_t stream) {
    const uint32_t blocks = (capacity + cpm::kBlock - 1) / cpm::kBlock;
    if (blocks == 0) return hipSuccess;
    CPM_LAUNCH(cpm::tiled_reconcile_kernel, dim3(blocks), dim3(cpm::kBlock), 0, stream, reps, dyn, slot);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_mech_advance(const cpm::RepDev *reps, const cpm::McsPlan *plan, int slot,
                                                       uint32_t capacity, uint32_t pass, uint32_t phase,
                                                       uint32_t plan_k, hipStream_t stream) {
    const uint32_t blocks = (capacity + cpm::kBlock - 1) / cpm::kBlock;
    if (blocks == 0) return hipSuccess;
    CPM_LAUNCH(cpm::mech_advance_kernel, dim3(blocks), dim3(cpm::kBlock), 0, stream, reps, plan, slot, pass,
                       phase, plan_k);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_mech_init(const cpm::RepDev *reps, int slot, uint32_t capacity,
                                                    hipStream_t stream) {
    const uint32_t blocks = (capacity + cpm::kBlock - 1) / cpm::kBlock;
    if (blocks == 0) return hipSuccess;
    CPM_LAUNCH(cpm::mech_init_kernel, dim3(blocks), dim3(cpm::kBlock), 0, stream, reps, slot);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_set_dyn(cpm::RepDyn *dyn, int slot, uint64_t mcs,
                                                  uint64_t target, uint32_t in_use, uint32_t mask,
                                                  hipStream_t stream) {
    CPM_LAUNCH(cpm::set_dyn_kernel, dim3(1), dim3(1), 0, stream, dyn, slot, mcs, target,
                       in_use, mask);
    return hipGetLastError();
}

namespace cpm {
// Plain 16-byte-per-lane device copy: the calibration kernel of known traffic
// (reads and writes n4 * 16 bytes) for hardware traffic counters and bandwidth probes.
__global__ __launch_bounds__(kBlock) void copy_probe_kernel(const v4i *__restrict__ src,
                                                            v4i *__restrict__ dst, size_t n4) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += stride)
        gp(dst)[i] = gp(src)[i];
}
}  // namespace cpm

extern "C" hipError_t cpm_internal_launch_copy_probe(const void *src, void *dst, size_t n4,
                                                     hipStream_t stream) {
    const size_t want = (n4 + cpm::kBlock - 1) / cpm::kBlock;
    const unsigned blocks = (unsigned)(want < 8192 ? (want ? want : 1) : 8192);
    CPM_LAUNCH(cpm::copy_probe_kernel, dim3(blocks), dim3(cpm::kBlock), 0, stream,
                       (const cpm::v4i *)src, (cpm::v4i *)dst, n4);
    return hipGetLastError();
}

// Global FP64 energy of replica `slot` into buf[kEnergyParts] (buf[0 .. kEnergyParts) hold the
// partial sums, buf[kEnergyParts] the total).
extern "C" hipError_t cpm_internal_launch_energy(const cpm::RepDev *reps, int slot, double *buf,
                                                 hipStream_t stream) {
    hipLaunchKernelGGL(cpm::energy_partials_kernel, dim3(cpm::kEnergyParts), dim3(cpm::kBlock), 0, stream, reps,
                       slot, buf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cpm::energy_sum_kernel, dim3(1), dim3(cpm::kBlock), 0, stream, (const double *)buf,
                       cpm::kEnergyParts, buf + cpm::kEnergyParts);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_report(const cpm::RepDev *reps, const uint32_t *slots,
                                                 uint32_t ntiles, int slot,
                                                 unsigned long long *out, hipStream_t stream) {
    CPM_LAUNCH(cpm::report_kernel, dim3(1), dim3(cpm::kBlock), 0, stream, reps, slots,
                       ntiles, slot, out);
    return hipGetLastError();
}
