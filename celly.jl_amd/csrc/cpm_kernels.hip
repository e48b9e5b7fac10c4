// HIP kernels of the MI355X CPM engine (gfx950).
//
// One colour pass (checkerboard) or lottery round is two launches:
//   propose  : LDS-staged lattice tile + halo; schedule filter (colour or lottery
//              local-maximum); interior sites whose proposal neighbours all share
//              the recipient's owner are settled as same-owner no-ops without RNG;
//              the rest are compacted into an LDS work list (wave ballot) and get
//              direction / priority Philox draws; actionable attempts publish a 64-bit
//              claim key (priority << 32 | ~recipient) to both cells with one atomic
//              max and are staged in a per-tile segment.
//   commit   : per-tile segment; an attempt wins iff it holds the key of both of its
//              cells (equivalent to the reference's max-priority then min-identity
//              claims, checkerboard.jl:281-369); winners evaluate the energy change
//              against the pass snapshot, draw acceptance and commit at once (winners
//              touch disjoint cells and non-adjacent sites, so evaluation and commit
//              need no separation).
#include <hip/hip_runtime.h>

#include "cpm_device.h"

namespace cpm {

template <int DIM, int H, int TX, int TY, int TZ>
struct Geom {
    static constexpr int HZ = DIM == 3 ? H : 0;
    static constexpr int LX = TX + 2 * H;
    static constexpr int LY = TY + 2 * H;
    static constexpr int LZ = TZ + 2 * HZ;
    static constexpr int LN = LX * LY * LZ;
    static constexpr int TN = TX * TY * TZ;
    static_assert(DIM == 3 || TZ == 1, "2D tiles have TZ == 1");
    static_assert(LN < 65536, "LDS positions must fit 16 bits");
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
__device__ inline int32_t realize_position(const RepDev &d, int gx, int gy, int gz,
                                           uint8_t *mutable_out, uint32_t *site_out) {
    int c[3] = {gx, gy, gz};
    bool closed = false, fixed = false, invalid = false;
    int32_t fixed_word = 0;
    for (int a = 0; a < d.ndim; ++a) {
        const int n = d.dims[a];
        int v = c[a];
        int face;
        if (v < 0) face = 0;
        else if (v >= n) face = 1;
        else continue;
        const int kind = d.face_kind[a][face];
        if (kind == CPM_FACE_PERIODIC) {
            c[a] = wrap_axis(v, n);
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
    }
    *site_out = 0;
    *mutable_out = 0;
    if (invalid || closed) return 0;
    if (fixed) return fixed_word;
    const uint32_t lin = (uint32_t)c[0] + (uint32_t)d.dims[0] *
                                               ((uint32_t)c[1] + (uint32_t)d.dims[1] * (uint32_t)c[2]);
    *site_out = lin + 1;
    *mutable_out = d.mask ? d.mask[lin] : (uint8_t)1;
    return d.lattice[lin];
}

template <int DIM, int H, int TX, int TY, int TZ>
__device__ inline void load_tile(const RepDev &d, int x0, int y0, int z0, int32_t *s_own,
                                 uint8_t *s_mut, uint32_t *s_sid) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const bool interior = x0 >= H && x0 + TX + H <= X && y0 >= H && y0 + TY + H <= Y &&
                          (DIM == 2 || (z0 >= G::HZ && z0 + TZ + G::HZ <= Z));
    const int32_t *lat = d.lattice;
    const uint8_t *mask = d.mask;
    for (int i = threadIdx.x; i < G::LN; i += kBlock) {
        const int lx = i % G::LX;
        const int ly = (i / G::LX) % G::LY;
        const int lz = i / (G::LX * G::LY);
        int gx = x0 - H + lx, gy = y0 - H + ly, gz = z0 - G::HZ + lz;
        int32_t w;
        uint8_t mu;
        uint32_t sid;
        if (interior || d.simple) {
            if (!interior) {
                gx = wrap_axis(gx, X);
                gy = wrap_axis(gy, Y);
                if (DIM == 3) gz = wrap_axis(gz, Z);
            }
            const uint32_t lin = (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz);
            w = lat[lin];
            mu = mask ? mask[lin] : (uint8_t)1;
            sid = lin + 1;
        } else {
            w = realize_position(d, gx, gy, gz, &mu, &sid);
        }
        s_own[i] = w;
        s_mut[i] = mu;
        if (s_sid) s_sid[i] = sid;
    }
}

// Append lanes with `flag` set to an LDS list; returns nothing, writes value.
__device__ __forceinline__ void ballot_push(bool flag, uint16_t value, uint32_t *counter,
                                            uint16_t *list) {
    const uint64_t m = __ballot(flag);
    if (m == 0) return;
    const uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (lane == (uint32_t)__ffsll((unsigned long long)m) - 1) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, __ffsll((unsigned long long)m) - 1);
    if (flag) list[base + __popcll(m & __lanemask_lt())] = value;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *s_red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) s_red[wave] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int k = 0; k < kBlock / 64; ++k) t += s_red[k];
    return t;
}

// Stage one actionable attempt: publish the claim key and append to the tile segment.
__device__ __forceinline__ void stage_attempt(bool act, const Entry &ent, const RepDev &d,
                                              unsigned long long *claims, uint32_t *s_nseg,
                                              Entry *seg) {
    if (act) {
        const unsigned long long key = ((unsigned long long)ent.priority << 32) | (uint32_t)(~ent.site);
        if (ent.losing > 0) atomicMax(&claims[d.claim_base + (uint32_t)ent.losing], key);
        if (ent.gaining > 0) atomicMax(&claims[d.claim_base + (uint32_t)ent.gaining], key);
    }
    const uint64_t m = __ballot(act);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(s_nseg, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (act) {
        const uint32_t idx = base + (uint32_t)__popcll(m & __lanemask_lt());
        uint4 v;
        v.x = ent.site;
        v.y = ent.priority;
        v.z = (uint32_t)ent.losing;
        v.w = (uint32_t)ent.gaining;
        reinterpret_cast<uint4 *>(seg)[idx] = v;
    }
}

// ---------------------------------------------------------------------------
// Checkerboard colour pass, propose + claim (checkerboard.jl:244-304)
// ---------------------------------------------------------------------------
template <int DIM, int H, int TX, int TY, int TZ>
__global__ __launch_bounds__(kBlock) void cb_propose_kernel(PassArgs a) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    __shared__ int32_t s_own[G::LN];
    __shared__ uint8_t s_mut[G::LN];
    __shared__ uint16_t s_work[G::TN];
    __shared__ int32_t s_ploff[CPM_MAX_DIRS];
    __shared__ uint8_t s_res[TX + TY + TZ];
    __shared__ uint32_t s_nwork, s_nseg;
    __shared__ uint32_t s_red[kBlock / 64];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const uint32_t r = a.tile_rep[tile];
    const RepDev &d = a.reps[r];
    const RepDyn &y = a.dyn[r];
    if (!y.active || a.pass >= (uint32_t)d.rounds) return;
    const uint64_t mcs = y.mcs;
    const uint32_t colour = y.order[a.pass];

    const uint32_t lt = tile - d.tile_base;
    const int ntx = d.ntiles[0], nty = d.ntiles[1];
    const int x0 = (int)(lt % ntx) * TX;
    const int y0 = (int)((lt / ntx) % nty) * TY;
    const int z0 = (int)(lt / (ntx * nty)) * TZ;
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const int K = d.prop.n;
    const int px = d.pat[0], py = d.pat[1], pz = d.pat[2];
    const bool by_pattern = d.site_color == nullptr;

    if (threadIdx.x == 0) {
        s_nwork = 0;
        s_nseg = 0;
    }
    for (int k = threadIdx.x; k < K; k += kBlock)
        s_ploff[k] = d.prop.off[k][0] + G::LX * (d.prop.off[k][1] + G::LY * d.prop.off[k][2]);
    if (by_pattern) {
        for (int i = threadIdx.x; i < TX + TY + TZ; i += kBlock) {
            int v;
            if (i < TX) v = (x0 + i) % px;
            else if (i < TX + TY) v = (y0 + i - TX) % py;
            else v = (z0 + i - TX - TY) % pz;
            s_res[i] = (uint8_t)v;
        }
    }
    load_tile<DIM, H, TX, TY, TZ>(d, x0, y0, z0, s_own, s_mut, nullptr);
    __syncthreads();

    uint32_t c_cand = 0, c_same = 0, c_bnd = 0, c_real = 0;
    for (int i0 = 0; i0 < G::TN; i0 += kBlock) {
        const int i = i0 + (int)threadIdx.x;
        bool need = false;
        if (i < G::TN) {
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
            if (gx < X && gy < Y && gz < Z) {
                const int pos = (ix + H) + G::LX * ((iy + H) + G::LY * (iz + G::HZ));
                if (s_mut[pos]) {
                    uint32_t c;
                    if (by_pattern)
                        c = d.pattern[s_res[ix] + px * (s_res[TX + iy] + py * s_res[TX + TY + iz])];
                    else
                        c = d.site_color[(uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz)];
                    if (c == colour) {
                        ++c_cand;
                        const int32_t own = s_own[pos];
                        bool interior = true;
                        for (int k = 0; k < K; ++k) {
                            const int q = pos + s_ploff[k];
                            if (!s_mut[q] || s_own[q] != own) {
                                interior = false;
                                break;
                            }
                        }
                        if (interior) ++c_same;
                        else need = true;
                    }
                }
            }
        }
        ballot_push(need, (uint16_t)i, &s_nwork, s_work);
    }
    __syncthreads();

    const uint32_t nwork = s_nwork;
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims;
    Entry *seg = a.seg + (size_t)tile * a.seg_cap;
    for (uint32_t e0 = 0; e0 < nwork; e0 += kBlock) {
        const uint32_t e = e0 + threadIdx.x;
        bool act = false;
        Entry ent{0, 0, 0, 0};
        if (e < nwork) {
            const int i = s_work[e];
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int pos = (ix + H) + G::LX * ((iy + H) + G::LY * (iz + G::HZ));
            const uint32_t site = 1u + (uint32_t)(x0 + ix) +
                                  (uint32_t)X * ((uint32_t)(y0 + iy) + (uint32_t)Y * (uint32_t)(z0 + iz));
            const uint32_t dir = bounded_uint(d.key_lo, d.key_hi, kStreamDirection, mcs, 0, 0,
                                              kSiteEntity, site, 0, (uint32_t)K);
            const int q = pos + s_ploff[dir];
            if (!s_mut[q]) {
                ++c_bnd;
            } else {
                const int32_t own = s_own[pos], gain = s_own[q];
                if (gain == own) {
                    ++c_same;
                } else {
                    ++c_real;
                    const Words w = rng_words(d.key_lo, d.key_hi, kStreamCheckerboardPriority, mcs,
                                              0, 0, kSiteEntity, site, 0, 0);
                    ent = Entry{site, w.w0, own, gain};
                    act = true;
                }
            }
        }
        stage_attempt(act, ent, d, claims, &s_nseg, seg);
    }

    c_cand = block_sum(c_cand, s_red);
    c_same = block_sum(c_same, s_red);
    c_bnd = block_sum(c_bnd, s_red);
    c_real = block_sum(c_real, s_red);
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        slot[0] = c_cand;
        slot[1] = c_same;
        slot[2] = c_bnd;
        slot[3] = c_real;
        a.seg_count[tile] = s_nseg;
    }
}

// ---------------------------------------------------------------------------
// Lottery round, propose + claim (lottery.jl:153-256)
// ---------------------------------------------------------------------------
template <int DIM, int H, int TX, int TY, int TZ>
__global__ __launch_bounds__(kBlock) void lot_propose_kernel(PassArgs a) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    __shared__ int32_t s_own[G::LN];
    __shared__ uint8_t s_mut[G::LN];
    __shared__ uint32_t s_sid[G::LN];
    __shared__ uint32_t s_tk[G::LN];
    __shared__ uint16_t s_work[G::TN];
    __shared__ int32_t s_ploff[CPM_MAX_DIRS];
    __shared__ int32_t s_cloff[CPM_MAX_DIRS];
    __shared__ uint32_t s_nwork, s_nseg;
    __shared__ uint32_t s_red[kBlock / 64];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const uint32_t r = a.tile_rep[tile];
    const RepDev &d = a.reps[r];
    const RepDyn &y = a.dyn[r];
    if (!y.active || a.pass >= (uint32_t)d.rounds) return;
    const uint64_t mcs = y.mcs;
    const uint32_t sub = y.order[a.pass] - 1;
    const uint32_t R = (uint32_t)d.rounds;

    const uint32_t lt = tile - d.tile_base;
    const int ntx = d.ntiles[0], nty = d.ntiles[1];
    const int x0 = (int)(lt % ntx) * TX;
    const int y0 = (int)((lt / ntx) % nty) * TY;
    const int z0 = (int)(lt / (ntx * nty)) * TZ;
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const int K = d.prop.n;
    const int KC = d.conf.n;

    if (threadIdx.x == 0) {
        s_nwork = 0;
        s_nseg = 0;
    }
    for (int k = threadIdx.x; k < K; k += kBlock)
        s_ploff[k] = d.prop.off[k][0] + G::LX * (d.prop.off[k][1] + G::LY * d.prop.off[k][2]);
    for (int k = threadIdx.x; k < KC; k += kBlock)
        s_cloff[k] = d.conf.off[k][0] + G::LX * (d.conf.off[k][1] + G::LY * d.conf.off[k][2]);
    load_tile<DIM, H, TX, TY, TZ>(d, x0, y0, z0, s_own, s_mut, s_sid);
    __syncthreads();
    // ticket word 1 of every mutable position (LotteryPriorityStream, subround = round - 1)
    for (int i = threadIdx.x; i < G::LN; i += kBlock) {
        uint32_t t = 0;
        if (s_mut[i])
            t = rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority, mcs, sub, 0, kSiteEntity,
                          s_sid[i], 0, 0).w0;
        s_tk[i] = t;
    }
    __syncthreads();

    uint32_t c_act = 0, c_same = 0, c_bnd = 0, c_real = 0;
    for (int i0 = 0; i0 < G::TN; i0 += kBlock) {
        const int i = i0 + (int)threadIdx.x;
        bool need = false;
        if (i < G::TN) {
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
            if (gx < X && gy < Y && gz < Z) {
                const int pos = (ix + H) + G::LX * ((iy + H) + G::LY * (iz + G::HZ));
                if (s_mut[pos]) {
                    const uint32_t site = s_sid[pos];
                    const uint32_t mine = s_tk[pos];
                    bool wins = true;
                    for (int k = 0; k < KC && wins; ++k) {
                        const int q = pos + s_cloff[k];
                        if (!s_mut[q]) continue;
                        const uint32_t nsite = s_sid[q];
                        if (nsite == site) continue;
                        const uint32_t other = s_tk[q];
                        if (other > mine) {
                            wins = false;
                        } else if (other == mine) {
                            // full lexicographic ticket comparison (lottery.jl:168-176)
                            const Words a4 = rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority,
                                                       mcs, sub, 0, kSiteEntity, nsite, 0, 0);
                            const Words b4 = rng_words(d.key_lo, d.key_hi, kStreamLotteryPriority,
                                                       mcs, sub, 0, kSiteEntity, site, 0, 0);
                            bool better;
                            if (a4.w1 != b4.w1) better = a4.w1 > b4.w1;
                            else if (a4.w2 != b4.w2) better = a4.w2 > b4.w2;
                            else if (a4.w3 != b4.w3) better = a4.w3 > b4.w3;
                            else better = nsite < site;
                            if (better) wins = false;
                        }
                    }
                    if (wins) {
                        bool eligible = true;
                        if (!d.uniform_degree) {
                            const uint32_t degree = d.site_degree[site - 1];
                            eligible = bounded_uint(d.key_lo, d.key_hi, kStreamLotteryActivation, mcs,
                                                    sub, 0, kSiteEntity, site, 0, R) < degree + 1;
                        }
                        if (eligible) {
                            ++c_act;
                            const int32_t own = s_own[pos];
                            bool interior = true;
                            for (int k = 0; k < K; ++k) {
                                const int q = pos + s_ploff[k];
                                if (!s_mut[q] || s_own[q] != own) {
                                    interior = false;
                                    break;
                                }
                            }
                            if (interior) ++c_same;
                            else need = true;
                        }
                    }
                }
            }
        }
        ballot_push(need, (uint16_t)i, &s_nwork, s_work);
    }
    __syncthreads();

    const uint32_t nwork = s_nwork;
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims;
    Entry *seg = a.seg + (size_t)tile * a.seg_cap;
    for (uint32_t e0 = 0; e0 < nwork; e0 += kBlock) {
        const uint32_t e = e0 + threadIdx.x;
        bool act = false;
        Entry ent{0, 0, 0, 0};
        if (e < nwork) {
            const int i = s_work[e];
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int pos = (ix + H) + G::LX * ((iy + H) + G::LY * (iz + G::HZ));
            const uint32_t site = s_sid[pos];
            const uint32_t dir = bounded_uint(d.key_lo, d.key_hi, kStreamDirection, mcs, sub, 0,
                                              kSiteEntity, site, 0, (uint32_t)K);
            const int q = pos + s_ploff[dir];
            if (!s_mut[q]) {
                ++c_bnd;
            } else {
                const int32_t own = s_own[pos], gain = s_own[q];
                if (gain == own) {
                    ++c_same;
                } else {
                    ++c_real;
                    ent = Entry{site, s_tk[pos], own, gain};
                    act = true;
                }
            }
        }
        stage_attempt(act, ent, d, claims, &s_nseg, seg);
    }

    c_act = block_sum(c_act, s_red);
    c_same = block_sum(c_same, s_red);
    c_bnd = block_sum(c_bnd, s_red);
    c_real = block_sum(c_real, s_red);
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        slot[0] = c_act;
        slot[1] = c_same;
        slot[2] = c_bnd;
        slot[3] = c_real;
        a.seg_count[tile] = s_nseg;
    }
}

// ---------------------------------------------------------------------------
// Claim resolution, energy change, acceptance and commit
// ---------------------------------------------------------------------------

// Owner word seen by a contact/surface relation at site (x,y,z) + offset; false when
// the neighbour is absent or invalid (cartesian.jl:494-525).
__device__ __forceinline__ bool neighbour_owner(const RepDev &d, int x, int y, int z,
                                                const int8_t *o, int32_t *word) {
    if (d.simple) {
        const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
        int nx = x + o[0], ny = y + o[1], nz = z + o[2];
        nx += nx < 0 ? X : 0;
        nx -= nx >= X ? X : 0;
        ny += ny < 0 ? Y : 0;
        ny -= ny >= Y ? Y : 0;
        nz += nz < 0 ? Z : 0;
        nz -= nz >= Z ? Z : 0;
        *word = d.lattice[(uint32_t)nx + (uint32_t)X * ((uint32_t)ny + (uint32_t)Y * (uint32_t)nz)];
        return true;
    }
    uint8_t mu;
    uint32_t sid;
    const int32_t w = realize_position(d, x + o[0], y + o[1], z + o[2], &mu, &sid);
    *word = w;
    return w != 0;
}

// Cell type or declared medium type, clamped into 1..n_types (the host rejects models
// and states that could reach an undeclared type).
__device__ __forceinline__ uint32_t owner_type(const RepDev &d, int32_t w) {
    uint32_t t = 1;
    if (w > 0) {
        t = d.ctype[w];
    } else {
        for (int i = 0; i < d.n_table; ++i)
            if (d.table_word[i] == w) t = d.table_type[i];
    }
    return min(max(t, 1u), (uint32_t)d.n_types);
}

__device__ __forceinline__ int medium_slot(const RepDev &d, int32_t w) {
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

struct EvalCounts {
    uint32_t conflicts = 0, rejected = 0, accepted = 0;
};

__device__ inline void resolve_entry(const RepDev &d, uint64_t mcs, uint32_t sub,
                                     const unsigned long long *claims, const Entry &e,
                                     EvalCounts &cnt) {
    const unsigned long long key = ((unsigned long long)e.priority << 32) | (uint32_t)(~e.site);
    const int32_t l = e.losing, g = e.gaining;
    bool win = true;
    if (l > 0) win = win && claims[d.claim_base + (uint32_t)l] == key;
    if (g > 0) win = win && claims[d.claim_base + (uint32_t)g] == key;
    if (!win) {
        ++cnt.conflicts;
        return;
    }
    const uint32_t idx0 = e.site - 1;
    const uint32_t X = (uint32_t)d.dims[0], Y = (uint32_t)d.dims[1];
    const int x = (int)(idx0 % X);
    const int y = (int)((idx0 / X) % Y);
    const int z = (int)(idx0 / (X * Y));

    // boundary tracker deltas (BoundaryEdgeCount, scientific_components.jl:466-481)
    int32_t dl = 0, dg = 0;
    for (int k = 0; k < d.surf.n; ++k) {
        int32_t w;
        if (!neighbour_owner(d, x, y, z, d.surf.off[k], &w)) continue;
        dl += (w == l) ? 1 : -1;
        dg += (w == g) ? -1 : 1;
    }

    // energy fold in component order (scientific_inner_loop.jl:213-217,257-269)
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
                if (l > 0) delta = __dadd_rn(delta, vol_change_f64(d.vs[l], d.tv[l], d.vol[l], -1));
                if (g > 0) delta = __dadd_rn(delta, vol_change_f64(d.vs[g], d.tv[g], d.vol[g], 1));
                v = delta;
            } else {
                float delta = 0.0f;
                if (l > 0)
                    delta = __fadd_rn(delta, vol_change_f32((float)d.vs[l], (float)d.tv[l], d.vol[l], -1));
                if (g > 0)
                    delta = __fadd_rn(delta, vol_change_f32((float)d.vs[g], (float)d.tv[g], d.vol[g], 1));
                v = (double)delta;
            }
        } else if (term == CPM_TERM_CONTACT) {
            v64 = d.contact_f64;
            const uint32_t nt = (uint32_t)d.n_types;
            const uint32_t lt = owner_type(d, l), gt = owner_type(d, g);
            double dd = 0.0;
            float df = 0.0f;
            for (int k = 0; k < d.cont.n; ++k) {
                int32_t w;
                if (!neighbour_owner(d, x, y, z, d.cont.off[k], &w)) continue;
                const uint32_t wt = owner_type(d, w);
                const uint32_t il = (lt - 1) * nt + (wt - 1), ig = (gt - 1) * nt + (wt - 1);
                if (v64) {
                    const double wk = d.cont.wd[k];
                    if (l != w) dd = __dsub_rn(dd, __dmul_rn(wk, d.Jd[il]));
                    if (g != w) dd = __dadd_rn(dd, __dmul_rn(wk, d.Jd[ig]));
                } else {
                    const float wk = d.cont.wf[k];
                    if (l != w) df = __fsub_rn(df, __fmul_rn(wk, d.Jf[il]));
                    if (g != w) df = __fadd_rn(df, __fmul_rn(wk, d.Jf[ig]));
                }
            }
            v = v64 ? dd : (double)df;
        } else {
            v64 = d.boundary_f64;
            if (v64) {
                double r = 0.0;
                if (l > 0)
                    r = __dadd_rn(r, bnd_change_f64(d.bs[l], (double)d.tb[l], (double)d.bnd[l], d.vol[l],
                                                    (double)dl, -1));
                if (g > 0)
                    r = __dadd_rn(r, bnd_change_f64(d.bs[g], (double)d.tb[g], (double)d.bnd[g], d.vol[g],
                                                    (double)dg, 1));
                v = r;
            } else {
                float r = 0.0f;
                if (l > 0)
                    r = __fadd_rn(r, bnd_change_f32((float)d.bs[l], (float)d.tb[l], (float)d.bnd[l],
                                                    d.vol[l], (float)dl, -1));
                if (g > 0)
                    r = __fadd_rn(r, bnd_change_f32((float)d.bs[g], (float)d.tb[g], (float)d.bnd[g],
                                                    d.vol[g], (float)dg, 1));
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

    // ConventionalMetropolis (protocols/scientific.jl:372-385), zero drive/kinetic/yield
    const float T = d.temperature;
    const float dd = __fadd_rn(dh, 0.0f);
    float p;
    if (T == 0.0f) {
        p = __fadd_rn(dd, 0.0f) <= 0.0f ? 1.0f : 0.0f;
    } else {
        float lr = __fdiv_rn(-__fadd_rn(dd, 0.0f), T);
        lr = __fadd_rn(__fadd_rn(lr, 0.0f), 0.0f);
        p = lr >= 0.0f ? 1.0f : (float)exp((double)lr);
    }
    const Words w = rng_words(d.key_lo, d.key_hi, kStreamAcceptance, mcs, sub, 0, kSiteEntity, e.site, 0, 0);
    if (!(uniform_open01(w.w0) < p)) {
        ++cnt.rejected;
        return;
    }
    ++cnt.accepted;
    // commit (checkerboard.jl:410-440)
    d.lattice[idx0] = g;
    if (l > 0) {
        const int32_t nv = d.vol[l] - 1;
        d.vol[l] = nv;
        d.bnd[l] += dl;
        if (nv == 0) {
            d.active[l] = 0;
            d.ctype[l] = 0;
            d.tv[l] = d.retire_tv;
            d.vs[l] = d.retire_vs;
            d.tb[l] = d.retire_tb;
            d.bs[l] = d.retire_bs;
        }
    } else {
        atomicAdd(&d.medium_vol[medium_slot(d, l)], -1);
    }
    if (g > 0) {
        d.vol[g] += 1;
        d.bnd[g] += dg;
    } else {
        atomicAdd(&d.medium_vol[medium_slot(d, g)], 1);
    }
}

__global__ __launch_bounds__(kBlock) void commit_kernel(PassArgs a) {
    __shared__ uint32_t s_red[kBlock / 64];
    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    // clear the claim buffer of the next pass (its previous reader finished earlier)
    {
        unsigned long long *next = a.claims + (size_t)(a.parity ^ 1u) * a.total_claims;
        const uint32_t per = (a.total_claims + gridDim.x - 1) / gridDim.x;
        const uint32_t lo = blockIdx.x * per;
        const uint32_t hi = min(lo + per, a.total_claims);
        for (uint32_t j = lo + threadIdx.x; j < hi; j += kBlock) next[j] = 0ull;
    }
    const uint32_t r = a.tile_rep[tile];
    const RepDev &d = a.reps[r];
    const RepDyn &y = a.dyn[r];
    if (!y.active || a.pass >= (uint32_t)d.rounds) return;
    const uint64_t mcs = y.mcs;
    const uint32_t sub = d.algorithm == CPM_ALG_LOTTERY ? y.order[a.pass] - 1 : 0;
    const unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims;
    const uint32_t n = a.seg_count[tile];
    const Entry *seg = a.seg + (size_t)tile * a.seg_cap;
    EvalCounts cnt;
    for (uint32_t e = threadIdx.x; e < n; e += kBlock) {
        const uint4 v = reinterpret_cast<const uint4 *>(seg)[e];
        const Entry ent{v.x, v.y, (int32_t)v.z, (int32_t)v.w};
        resolve_entry(d, mcs, sub, claims, ent, cnt);
    }
    const uint32_t c0 = block_sum(cnt.conflicts, s_red);
    const uint32_t c1 = block_sum(cnt.rejected, s_red);
    const uint32_t c2 = block_sum(cnt.accepted, s_red);
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        slot[4] = c0;
        slot[5] = c1;
        slot[6] = c2;
    }
}

// Starts one MCS for every replica below its target: mcs += 1 and the colour / round
// order of that MCS (checkerboard.jl:226-242, lottery.jl:124-140).
__global__ void begin_mcs_kernel(const RepDev *reps, RepDyn *dyn, int nslots) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nslots) return;
    RepDyn &y = dyn[r];
    if (!y.in_use || y.mcs >= y.target) {
        y.active = 0;
        return;
    }
    const uint64_t m = y.mcs + 1;
    y.mcs = m;
    y.active = 1;
    const RepDev &d = reps[r];
    if (d.algorithm == CPM_ALG_LOTTERY)
        schedule_order(d.key_lo, d.key_hi, kStreamLotteryActivation, 1, m, d.rounds, y.order);
    else
        schedule_order(d.key_lo, d.key_hi, kStreamCheckerboardOrder, 0, m, d.rounds, y.order);
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
    const RepDev &d = reps[slot];
    unsigned long long acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t per_pass = d.tile_count;
    const uint32_t total = per_pass * (uint32_t)d.rounds;
    for (uint32_t i = threadIdx.x; i < total; i += kBlock) {
        const uint32_t p = i / per_pass, t = d.tile_base + i % per_pass;
        const uint32_t *s = slots + ((size_t)p * ntiles + t) * kSlotFields;
        for (int k = 0; k < 7; ++k) acc[k] += s[k];
    }
    for (int k = 0; k < 7; ++k) s_acc[k][threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x < 7) {
        unsigned long long t = 0;
        for (int j = 0; j < kBlock; ++j) t += s_acc[threadIdx.x][j];
        out[threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int DIM, int H, int TX, int TY, int TZ>
static hipError_t launch_pass_t(int algorithm, const PassArgs &a, hipStream_t stream) {
    if (algorithm == CPM_ALG_LOTTERY)
        hipLaunchKernelGGL((lot_propose_kernel<DIM, H, TX, TY, TZ>), dim3(a.ntiles), dim3(kBlock), 0,
                           stream, a);
    else
        hipLaunchKernelGGL((cb_propose_kernel<DIM, H, TX, TY, TZ>), dim3(a.ntiles), dim3(kBlock), 0,
                           stream, a);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(commit_kernel, dim3(a.ntiles), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

}  // namespace cpm

// Tile configurations: id -> (dim, halo, TX, TY, TZ).
extern "C" int cpm_internal_tile_config(int id, int *dim, int *halo, int *tx, int *ty, int *tz) {
    static const int table[][5] = {
        {2, 1, 128, 16, 1}, {2, 2, 128, 16, 1}, {3, 1, 32, 8, 8}, {3, 2, 32, 8, 8},
        {2, 1, 64, 32, 1},  {2, 1, 256, 8, 1},  {2, 1, 64, 16, 1}, {2, 1, 32, 32, 1},
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

extern "C" hipError_t cpm_internal_launch_pass(int config, int algorithm, const cpm::PassArgs *a,
                                               hipStream_t stream) {
    using namespace cpm;
    switch (config) {
    case 0: return launch_pass_t<2, 1, 128, 16, 1>(algorithm, *a, stream);
    case 1: return launch_pass_t<2, 2, 128, 16, 1>(algorithm, *a, stream);
    case 2: return launch_pass_t<3, 1, 32, 8, 8>(algorithm, *a, stream);
    case 3: return launch_pass_t<3, 2, 32, 8, 8>(algorithm, *a, stream);
    case 4: return launch_pass_t<2, 1, 64, 32, 1>(algorithm, *a, stream);
    case 5: return launch_pass_t<2, 1, 256, 8, 1>(algorithm, *a, stream);
    case 6: return launch_pass_t<2, 1, 64, 16, 1>(algorithm, *a, stream);
    case 7: return launch_pass_t<2, 1, 32, 32, 1>(algorithm, *a, stream);
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t cpm_internal_launch_begin(const cpm::RepDev *reps, cpm::RepDyn *dyn,
                                                int nslots, hipStream_t stream) {
    const int blocks = (nslots + 63) / 64;
    hipLaunchKernelGGL(cpm::begin_mcs_kernel, dim3(blocks), dim3(64), 0, stream, reps, dyn, nslots);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_set_dyn(cpm::RepDyn *dyn, int slot, uint64_t mcs,
                                                  uint64_t target, uint32_t in_use, uint32_t mask,
                                                  hipStream_t stream) {
    hipLaunchKernelGGL(cpm::set_dyn_kernel, dim3(1), dim3(1), 0, stream, dyn, slot, mcs, target,
                       in_use, mask);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_report(const cpm::RepDev *reps, const uint32_t *slots,
                                                 uint32_t ntiles, int slot,
                                                 unsigned long long *out, hipStream_t stream) {
    hipLaunchKernelGGL(cpm::report_kernel, dim3(1), dim3(cpm::kBlock), 0, stream, reps, slots,
                       ntiles, slot, out);
    return hipGetLastError();
}
