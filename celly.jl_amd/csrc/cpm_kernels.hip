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

__device__ inline Realized realize_position(const RepDev &d, int gx, int gy, int gz) {
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

template <int DIM, int H, int TX, int TY, int TZ>
__device__ inline void load_tile(const RepDev &d, int x0, int y0, int z0, int32_t *s_own,
                                 uint8_t *s_mut, uint32_t *s_sid) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    const bool inner_yz = y0 >= H && y0 + TY + H <= Y && (DIM == 2 || (z0 >= G::HZ && z0 + TZ + G::HZ <= Z));
    const CPM_GLOBAL int32_t *lat = gp(d.lattice);
    if (d.simple && (X & 3) == 0 && x0 >= G::PX && x0 + TX + G::PX <= X && inner_yz) {
        // interior tile of a periodic, obstacle-free lattice: 16-byte loads, all in flight
        constexpr int NL = (G::NV + kBlock - 1) / kBlock;
        v4i v[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const int jj = (int)threadIdx.x + j * kBlock;
            if (jj < G::NV) {
                const int row = jj / G::LV, c4 = jj % G::LV;
                const int gx = x0 - G::PX + 4 * c4, gy = y0 - H + row % G::LY, gz = z0 - G::HZ + row / G::LY;
                v[j] = *(const CPM_GLOBAL v4i *)(lat + (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz));
            }
        }
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const int jj = (int)threadIdx.x + j * kBlock;
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
            for (int i = threadIdx.x; i < G::LN; i += kBlock) s_mut[i] = 1;
        return;
    }
    constexpr int NL = (G::LN + kBlock - 1) / kBlock;
    const bool interior = x0 >= G::PX && x0 + TX + G::PX <= X && inner_yz;
    const CPM_GLOBAL uint8_t *mask = d.mask ? gp(d.mask) : nullptr;
    int32_t w[NL];
    // every load of the tile is issued before the first LDS store
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const int i = (int)threadIdx.x + j * kBlock;
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
        const int i = (int)threadIdx.x + j * kBlock;
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

// Cell type or declared medium type, clamped into 1..n_types (the host rejects models
// and states that could reach an undeclared type).
__device__ __forceinline__ uint32_t owner_type(const RepDev &d, int32_t w) {
    uint32_t t = 1;
    if (w > 0) {
        t = gp(d.ctype)[w];
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

// ---------------------------------------------------------------------------
// Propose side
// ---------------------------------------------------------------------------

// Publish the claim key of an actionable attempt on both cells (one 64-bit atomic max
// per cell replaces the max-priority and min-identity claims, checkerboard.jl:281-335)
// and append the attempt to the tile segment.
__device__ __forceinline__ void stage_claim(bool act, const Entry &ent, const RepDev &d,
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
        uint4 v;
        v.x = ent.site;
        v.y = ent.priority;
        v.z = (uint32_t)ent.losing;
        v.w = (uint32_t)ent.gaining;
        reinterpret_cast<uint4 *>(seg)[base + __popcll(m & __lanemask_lt())] = v;
    }
}

__device__ __forceinline__ int ilog2_exact(int v) { return 31 - __clz(v); }

// Enumerates the tile sites of one colour. Pattern colourings visit only the residue
// classes of that colour; per-site colourings scan the tile. fn(i, ix, iy, iz) is called
// by every lane of the workgroup in lock-step (i < 0 for idle lanes).
template <int DIM, int TX, int TY, int TZ, typename F>
__device__ inline void for_colour_sites(const RepDev &d, const uint8_t *s_pattern, uint32_t colour,
                                        int x0, int y0, int z0, F &&fn) {
    const int X = d.dims[0], Y = d.dims[1], Z = d.dims[2];
    if (d.site_color == nullptr) {
        const int px = d.pat[0], py = d.pat[1], pz = d.pat[2];
        const int np = px * py * pz;
        // periods dividing the tile sides give power-of-two class extents: shifts, no division
        const bool pow2 = (TX % px == 0) && (TY % py == 0) && (TZ % pz == 0) && ((TX / px) & (TX / px - 1)) == 0 &&
                          ((TY / py) & (TY / py - 1)) == 0 && ((TZ / pz) & (TZ / pz - 1)) == 0;
        for (int r = 0; r < np; ++r) {
            if (s_pattern[r] != colour) continue;
            const int rx = d.pat_coord[r][0], ry = d.pat_coord[r][1], rz = d.pat_coord[r][2];
            int fx = rx, fy = ry, fz = rz;
            if (!d.pat_aligned) {
                fx = (rx - x0 % px + px) % px;
                fy = (ry - y0 % py + py) % py;
                fz = (rz - z0 % pz + pz) % pz;
            }
            const int nx = fx < TX ? (TX - fx + px - 1) / px : 0;
            const int ny = fy < TY ? (TY - fy + py - 1) / py : 0;
            const int nz = fz < TZ ? (TZ - fz + pz - 1) / pz : 0;
            const int n = nx * ny * nz;
            const int sx = pow2 ? ilog2_exact(max(nx, 1)) : 0, sy = pow2 ? ilog2_exact(max(ny, 1)) : 0;
            for (int j0 = 0; j0 < n; j0 += kBlock) {
                const int j = j0 + (int)threadIdx.x;
                int i = -1, ix = 0, iy = 0, iz = 0;
                if (j < n) {
                    int kx, ky, kz;
                    if (pow2) {
                        kx = j & (nx - 1);
                        ky = (j >> sx) & (ny - 1);
                        kz = j >> (sx + sy);
                    } else {
                        kx = j % nx;
                        ky = (j / nx) % ny;
                        kz = j / (nx * ny);
                    }
                    ix = fx + px * kx;
                    iy = fy + py * ky;
                    iz = fz + pz * kz;
                    if (x0 + ix < X && y0 + iy < Y && z0 + iz < Z) i = ix + TX * (iy + TY * iz);
                }
                fn(i, ix, iy, iz);
            }
        }
    } else {
        constexpr int TN = TX * TY * TZ;
        for (int i0 = 0; i0 < TN; i0 += kBlock) {
            int i = i0 + (int)threadIdx.x;
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            if (i < TN) {
                const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
                if (gx >= X || gy >= Y || gz >= Z ||
                    gp(d.site_color)[(uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz)] != colour)
                    i = -1;
            } else {
                i = -1;
            }
            fn(i, ix, iy, iz);
        }
    }
}

// Up to kRegDirs proposal offsets kept in (scalar) registers for the interior test.
constexpr int kRegDirs = 8;

template <int LX, int LY>
__device__ __forceinline__ void reg_offsets(const RelDev &rel, int32_t (&po)[kRegDirs]) {
#pragma unroll
    for (int k = 0; k < kRegDirs; ++k)
        po[k] = k < rel.n ? rel.off[k][0] + LX * (rel.off[k][1] + LY * rel.off[k][2]) : 0;
}

// True when every proposal neighbour of pos is a mutable site owned by `own`: any
// direction draw would then yield a same-owner no-op. Branch-free, independent LDS reads.
template <bool SIMPLE>
__device__ __forceinline__ bool interior_site(const int32_t *s_own, const uint8_t *s_mut, int pos,
                                              int32_t own, const int32_t (&po)[kRegDirs], int K,
                                              const int32_t *s_ploff) {
    bool interior = true;
    if (K <= kRegDirs) {
#pragma unroll
        for (int k = 0; k < kRegDirs; ++k) {
            if (k < K) {
                const int q = pos + po[k];
                interior &= s_own[q] == own;
                if (!SIMPLE) interior &= s_mut[q] != 0;
            }
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

// ---------------------------------------------------------------------------
// Checkerboard colour pass, propose + claim (checkerboard.jl:244-304)
// ---------------------------------------------------------------------------
template <int DIM, int H, int TX, int TY, int TZ>
__global__ __launch_bounds__(kBlock) void cb_propose_kernel(PassArgs a) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    __shared__ __align__(16) int32_t s_own[G::LN];
    __shared__ uint8_t s_mut[G::LN];
    __shared__ uint16_t s_work[G::TN];
    __shared__ uint16_t s_act[G::TN];
    __shared__ uint8_t s_dir[G::TN];
    __shared__ int32_t s_ploff[CPM_MAX_DIRS];
    __shared__ uint8_t s_pattern[64];
    __shared__ uint32_t s_nwork, s_nact, s_nseg;
    __shared__ uint32_t s_red[4][kBlock / 64];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const TileInfo ti = a.tiles[tile];
    const uint32_t r = ti.rep;
    const RepDev &d = a.reps[r];
    const RepDyn &y = a.dyn[r];
    if (!y.active || a.pass >= (uint32_t)d.rounds) return;
    const uint64_t mcs = y.mcs;
    const uint32_t colour = y.order[a.pass];

    const int x0 = ti.x0, y0 = ti.y0, z0 = ti.z0;
    const int X = d.dims[0], Y = d.dims[1];
    const int K = d.prop.n;
    const bool simple = d.simple;

    if (threadIdx.x == 0) {
        s_nwork = 0;
        s_nact = 0;
        s_nseg = 0;
    }
    for (int k = threadIdx.x; k < K; k += kBlock)
        s_ploff[k] = d.prop.off[k][0] + G::LX * (d.prop.off[k][1] + G::LY * d.prop.off[k][2]);
    if (threadIdx.x < 64) s_pattern[threadIdx.x] = d.pattern[threadIdx.x];
    int32_t po[kRegDirs];
    reg_offsets<G::LX, G::LY>(d.prop, po);
    load_tile<DIM, H, TX, TY, TZ>(d, x0, y0, z0, s_own, s_mut, nullptr);
    __syncthreads();

    // 1. colour sites; interior sites (every proposal neighbour a mutable site of the
    //    same owner) are same-owner no-ops whatever the direction draw
    uint32_t cnt[4] = {0, 0, 0, 0};  // candidates, same owner, boundary null, realized
    for_colour_sites<DIM, TX, TY, TZ>(d, s_pattern, colour, x0, y0, z0, [&](int i, int ix, int iy, int iz) {
        bool need = false;
        if (i >= 0) {
            const int pos = G::pos(ix, iy, iz);
            if (simple || s_mut[pos]) {
                ++cnt[0];
                const int32_t own = s_own[pos];
                const bool interior = simple ? interior_site<true>(s_own, s_mut, pos, own, po, K, s_ploff)
                                             : interior_site<false>(s_own, s_mut, pos, own, po, K, s_ploff);
                if (interior) ++cnt[1];
                else need = true;
            }
        }
        ballot_push(need, (uint16_t)(i < 0 ? 0 : i), &s_nwork, s_work);
    });
    __syncthreads();

    // 2. direction draws of the boundary sites (ProposalDirectionStream), attempt outcome
    const uint32_t nwork = s_nwork;
    for (uint32_t e0 = 0; e0 < nwork; e0 += kBlock) {
        const uint32_t e = e0 + threadIdx.x;
        bool act = false;
        int i = 0;
        if (e < nwork) {
            i = s_work[e];
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int pos = G::pos(ix, iy, iz);
            const uint32_t site = 1u + (uint32_t)(x0 + ix) +
                                  (uint32_t)X * ((uint32_t)(y0 + iy) + (uint32_t)Y * (uint32_t)(z0 + iz));
            const uint32_t dir = bounded_draw(d.key_lo, d.key_hi, kStreamDirection, mcs, 0, 0,
                                              kSiteEntity, site, 0, d.prop_bound);
            int off = 0;
            if (K <= kRegDirs) {
#pragma unroll
                for (int k = 0; k < kRegDirs; ++k) off = dir == (uint32_t)k ? po[k] : off;
            } else {
                off = s_ploff[dir];
            }
            const int q = pos + off;
            if (!simple && !s_mut[q]) {
                ++cnt[2];
            } else if (s_own[q] == s_own[pos]) {
                ++cnt[1];
            } else {
                ++cnt[3];
                act = true;
                i |= (int)(dir << 16);
            }
        }
        const uint64_t m = __ballot(act);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(&s_nact, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (act) {
                const uint32_t idx = base + __popcll(m & __lanemask_lt());
                s_act[idx] = (uint16_t)(i & 0xffff);
                s_dir[idx] = (uint8_t)(i >> 16);
            }
        }
    }
    __syncthreads();

    // 3. priority draws and claims of the actionable attempts (compacted)
    const uint32_t nact = s_nact;
    unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims;
    Entry *seg = a.seg + (size_t)tile * a.seg_cap;
    for (uint32_t e0 = 0; e0 < nact; e0 += kBlock) {
        const uint32_t e = e0 + threadIdx.x;
        bool act = false;
        Entry ent{0, 0, 0, 0};
        if (e < nact) {
            const int i = s_act[e];
            const int dir = s_dir[e];
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int pos = G::pos(ix, iy, iz);
            const uint32_t site = 1u + (uint32_t)(x0 + ix) +
                                  (uint32_t)X * ((uint32_t)(y0 + iy) + (uint32_t)Y * (uint32_t)(z0 + iz));
            const Words w = rng_words(d.key_lo, d.key_hi, kStreamCheckerboardPriority, mcs, 0, 0,
                                      kSiteEntity, site, 0, 0);
            ent = Entry{site, w.w0, s_own[pos], s_own[pos + s_ploff[dir]]};
            act = true;
        }
        stage_claim(act, ent, d, claims, &s_nseg, seg);
    }

    block_sum4(cnt, s_red);
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        slot[0] = cnt[0];
        slot[1] = cnt[1];
        slot[2] = cnt[2];
        slot[3] = cnt[3];
        a.seg_count[tile] = s_nseg;
    }
}

// ---------------------------------------------------------------------------
// Lottery round, propose + claim (lottery.jl:153-256)
// ---------------------------------------------------------------------------
template <int DIM, int H, int TX, int TY, int TZ>
__global__ __launch_bounds__(kBlock) void lot_propose_kernel(PassArgs a) {
    using G = Geom<DIM, H, TX, TY, TZ>;
    __shared__ __align__(16) int32_t s_own[G::LN];
    __shared__ uint8_t s_mut[G::LN];
    __shared__ uint32_t s_sid[G::LN];
    __shared__ uint32_t s_tk[G::LN];
    __shared__ uint16_t s_work[G::TN];
    __shared__ int32_t s_ploff[CPM_MAX_DIRS];
    __shared__ int32_t s_cloff[CPM_MAX_DIRS];
    __shared__ uint32_t s_nwork, s_nseg;
    __shared__ uint32_t s_red[4][kBlock / 64];

    const uint32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const TileInfo ti = a.tiles[tile];
    const uint32_t r = ti.rep;
    const RepDev &d = a.reps[r];
    const RepDyn &y = a.dyn[r];
    if (!y.active || a.pass >= (uint32_t)d.rounds) return;
    const uint64_t mcs = y.mcs;
    const uint32_t sub = y.order[a.pass] - 1;

    const int x0 = ti.x0, y0 = ti.y0, z0 = ti.z0;
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
    int32_t po[kRegDirs];
    reg_offsets<G::LX, G::LY>(d.prop, po);
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

    uint32_t cnt[4] = {0, 0, 0, 0};  // activated, same owner, boundary null, realized
    for (int i0 = 0; i0 < G::TN; i0 += kBlock) {
        const int i = i0 + (int)threadIdx.x;
        bool need = false;
        if (i < G::TN) {
            const int ix = i % TX, iy = (i / TX) % TY, iz = i / (TX * TY);
            const int gx = x0 + ix, gy = y0 + iy, gz = z0 + iz;
            if (gx < X && gy < Y && gz < Z) {
                const int pos = G::pos(ix, iy, iz);
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
                            const uint32_t degree = gp(d.site_degree)[site - 1];
                            eligible = bounded_draw(d.key_lo, d.key_hi, kStreamLotteryActivation, mcs,
                                                    sub, 0, kSiteEntity, site, 0, d.round_bound) < degree + 1;
                        }
                        if (eligible) {
                            ++cnt[0];
                            const int32_t own = s_own[pos];
                            if (interior_site<false>(s_own, s_mut, pos, own, po, K, s_ploff)) ++cnt[1];
                            else need = true;
                        }
                    }
                }
            }
        }
        ballot_push(need, (uint16_t)(i < G::TN ? i : 0), &s_nwork, s_work);
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
            const int pos = G::pos(ix, iy, iz);
            const uint32_t site = s_sid[pos];
            const uint32_t dir = bounded_draw(d.key_lo, d.key_hi, kStreamDirection, mcs, sub, 0,
                                              kSiteEntity, site, 0, d.prop_bound);
            const int q = pos + s_ploff[dir];
            if (!s_mut[q]) {
                ++cnt[2];
            } else {
                const int32_t own = s_own[pos], gain = s_own[q];
                if (gain == own) {
                    ++cnt[1];
                } else {
                    ++cnt[3];
                    ent = Entry{site, s_tk[pos], own, gain};
                    act = true;
                }
            }
        }
        stage_claim(act, ent, d, claims, &s_nseg, seg);
    }

    block_sum4(cnt, s_red);
    if (threadIdx.x == 0) {
        uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
        slot[0] = cnt[0];
        slot[1] = cnt[1];
        slot[2] = cnt[2];
        slot[3] = cnt[3];
        a.seg_count[tile] = s_nseg;
    }
}

// ---------------------------------------------------------------------------
// Claim resolution, evaluation, acceptance and commit (checkerboard.jl:337-457)
// ---------------------------------------------------------------------------

// Owner word seen by a contact/surface relation at site (x,y,z) + offset; 0 when the
// neighbour is absent or invalid (cartesian.jl:494-525).
__device__ __forceinline__ int32_t neighbour_word(const RepDev &d, int x, int y, int z, const int8_t *o) {
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

// evaluate_copy + ConventionalMetropolis + acceptance draw of one winning attempt against
// the pass snapshot (checkerboard.jl:371-408; scientific_inner_loop.jl:213-269;
// protocols/scientific.jl:372-385). Winners touch disjoint cells and non-adjacent sites,
// so other winners' commits never change what this evaluation reads.
__device__ inline bool evaluate_winner(const RepDev &d, const Entry &e, uint64_t mcs, uint32_t sub,
                                       int32_t *dl_out, int32_t *dg_out) {
    const int32_t l = e.losing, g = e.gaining;
    const uint32_t idx0 = e.site - 1;
    const uint32_t zq = fast_div(idx0, d.div_xy);
    const uint32_t rem = idx0 - zq * d.div_xy.d;
    const uint32_t yq = fast_div(rem, d.div_x);
    const int x = (int)(rem - yq * d.div_x.d);
    const int y = (int)yq;
    const int z = (int)zq;
    const int32_t vl = l > 0 ? gp(d.vol)[l] : 0;
    const int32_t vg = g > 0 ? gp(d.vol)[g] : 0;
    // boundary tracker deltas (BoundaryEdgeCount, scientific_components.jl:466-481)
    int32_t dl = 0, dg = 0;
    for (int k = 0; k < d.surf.n; ++k) {
        const int32_t w = neighbour_word(d, x, y, z, d.surf.off[k]);
        if (w == 0) continue;
        dl += (w == l) ? 1 : -1;
        dg += (w == g) ? -1 : 1;
    }
    *dl_out = dl;
    *dg_out = dg;
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
                if (l > 0) delta = __dadd_rn(delta, vol_change_f64(gp(d.vs)[l], gp(d.tv)[l], vl, -1));
                if (g > 0) delta = __dadd_rn(delta, vol_change_f64(gp(d.vs)[g], gp(d.tv)[g], vg, 1));
                v = delta;
            } else {
                float delta = 0.0f;
                if (l > 0) delta = __fadd_rn(delta, vol_change_f32((float)gp(d.vs)[l], (float)gp(d.tv)[l], vl, -1));
                if (g > 0) delta = __fadd_rn(delta, vol_change_f32((float)gp(d.vs)[g], (float)gp(d.tv)[g], vg, 1));
                v = (double)delta;
            }
        } else if (term == CPM_TERM_CONTACT) {
            // scientific_components.jl:289-308
            v64 = d.contact_f64;
            const uint32_t nt = (uint32_t)d.n_types;
            const uint32_t lt = owner_type(d, l), gt = owner_type(d, g);
            double dd = 0.0;
            float df = 0.0f;
            const int nc = d.cont.n;
            constexpr int B = 4;
            for (int k0 = 0; k0 < nc; k0 += B) {
                int32_t wb[B];
                uint32_t tb[B];
#pragma unroll
                for (int u = 0; u < B; ++u) wb[u] = k0 + u < nc ? neighbour_word(d, x, y, z, d.cont.off[k0 + u]) : 0;
#pragma unroll
                for (int u = 0; u < B; ++u) tb[u] = gp(d.ctype)[wb[u] > 0 ? wb[u] : 0];  // slot 0 is a spare
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const int32_t w = wb[u];
                    if (w == 0) continue;
                    const int k = k0 + u;
                    uint32_t wt;
                    if (w == l) wt = lt;
                    else if (w == g) wt = gt;
                    else if (w > 0) wt = min(max(tb[u], 1u), nt);
                    else wt = owner_type(d, w);
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
            }
            v = v64 ? dd : (double)df;
        } else {
            // scientific_trackers.jl:144-172
            v64 = d.boundary_f64;
            if (v64) {
                double r = 0.0;
                if (l > 0)
                    r = __dadd_rn(r, bnd_change_f64(gp(d.bs)[l], (double)gp(d.tb)[l], (double)gp(d.bnd)[l], vl, (double)dl, -1));
                if (g > 0)
                    r = __dadd_rn(r, bnd_change_f64(gp(d.bs)[g], (double)gp(d.tb)[g], (double)gp(d.bnd)[g], vg, (double)dg, 1));
                v = r;
            } else {
                float r = 0.0f;
                if (l > 0)
                    r = __fadd_rn(r, bnd_change_f32((float)gp(d.bs)[l], (float)gp(d.tb)[l], (float)gp(d.bnd)[l], vl, (float)dl, -1));
                if (g > 0)
                    r = __fadd_rn(r, bnd_change_f32((float)gp(d.bs)[g], (float)gp(d.tb)[g], (float)gp(d.bnd)[g], vg, (float)dg, 1));
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
    return uniform_open01(w.w0) < p;
}

// Commit of one accepted winning attempt: ownership, volumes, boundary measures,
// retirement, medium volumes (checkerboard.jl:410-440).
__device__ inline void commit_attempt(const RepDev &d, const Entry &e, int32_t dl, int32_t dg) {
    const int32_t l = e.losing, g = e.gaining;
    gp(d.lattice)[e.site - 1] = g;
    if (l > 0) {
        const int32_t nv = gp(d.vol)[l] - 1;
        gp(d.vol)[l] = nv;
        gp(d.bnd)[l] += dl;
        if (nv == 0) {
            gp(d.active)[l] = 0;
            gp(d.ctype)[l] = 0;
            gp(d.tv)[l] = d.retire_tv;
            gp(d.vs)[l] = d.retire_vs;
            gp(d.tb)[l] = d.retire_tb;
            gp(d.bs)[l] = d.retire_bs;
        }
    } else {
        atomicAdd(&d.medium_vol[medium_slot(d, l)], -1);
    }
    if (g > 0) {
        gp(d.vol)[g] += 1;
        gp(d.bnd)[g] += dg;
    } else {
        atomicAdd(&d.medium_vol[medium_slot(d, g)], 1);
    }
}

constexpr int kCommitTiles = 8;     // tile segments per commit workgroup
constexpr int kCommitChunk = 1024;  // staged attempts per claim-check round

__global__ __launch_bounds__(kBlock) void commit_kernel(PassArgs a) {
    __shared__ uint32_t s_base[kCommitTiles + 1];
    __shared__ uint32_t s_cnt[kCommitTiles][3];
    __shared__ uint32_t s_win[kCommitChunk];
    __shared__ uint32_t s_nwin;
    const uint32_t nwg = gridDim.x;
    const uint32_t b = xcd_tile(blockIdx.x, nwg);
    // clear the claim buffer of the next pass (its previous reader finished earlier)
    {
        unsigned long long *next = a.claims + (size_t)(a.parity ^ 1u) * a.total_claims;
        const uint32_t per = (a.total_claims + nwg - 1) / nwg;
        const uint32_t lo = b * per;
        const uint32_t hi = min(lo + per, a.total_claims);
        for (uint32_t j = lo + threadIdx.x; j < hi; j += kBlock) next[j] = 0ull;
    }
    if (threadIdx.x < kCommitTiles) {
        const int k = threadIdx.x;
        const uint32_t tile = b * kCommitTiles + k;
        uint32_t n = 0;
        if (tile < a.ntiles) {
            const uint32_t r = a.tiles[tile].rep;
            if (a.dyn[r].active && a.pass < (uint32_t)a.reps[r].rounds) n = a.seg_count[tile];
        }
        s_base[k + 1] = n;
        s_cnt[k][0] = s_cnt[k][1] = s_cnt[k][2] = 0;
    }
    if (threadIdx.x == 0) s_nwin = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base[0] = 0;
        for (int k = 1; k <= kCommitTiles; ++k) s_base[k] += s_base[k - 1];
    }
    __syncthreads();
    const uint32_t total = s_base[kCommitTiles];
    const unsigned long long *claims = a.claims + (size_t)a.parity * a.total_claims;
    for (uint32_t c0 = 0; c0 < total; c0 += kCommitChunk) {
        const uint32_t c1 = min(c0 + kCommitChunk, total);
        // A. claim check of the staged attempts; winners are compacted into LDS
        for (uint32_t e0 = c0; e0 < c1; e0 += kBlock) {
            const uint32_t e = e0 + threadIdx.x;
            bool win = false;
            uint32_t code = 0;
            if (e < c1) {
                int k = 0;
                while (k + 1 < kCommitTiles && e >= s_base[k + 1]) ++k;
                const uint32_t tile = b * kCommitTiles + k;
                const uint32_t i = e - s_base[k];
                const uint4 v = reinterpret_cast<const uint4 *>(a.seg)[(size_t)tile * a.seg_cap + i];
                const RepDev &d = a.reps[a.tiles[tile].rep];
                const unsigned long long key = ((unsigned long long)v.y << 32) | (uint32_t)(~v.x);
                win = true;
                if ((int32_t)v.z > 0) win = claims[d.claim_base + v.z] == key;
                if ((int32_t)v.w > 0) win = win && claims[d.claim_base + v.w] == key;
                if (!win) atomicAdd(&s_cnt[k][0], 1u);
                code = ((uint32_t)k << 16) | i;
            }
            const uint64_t m = __ballot(win);
            if (m) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                uint32_t base = 0;
                if ((int)__lane_id() == leader) base = atomicAdd(&s_nwin, (uint32_t)__popcll(m));
                base = __shfl(base, leader);
                if (win) s_win[base + __popcll(m & __lanemask_lt())] = code;
            }
        }
        __syncthreads();
        // B. winners: energy change, acceptance and commit
        const uint32_t nwin = s_nwin;
        // winners are packed densely into as few waves as possible
        for (uint32_t e = threadIdx.x; e < nwin; e += kBlock) {
            const uint32_t code = s_win[e];
            const int k = (int)(code >> 16);
            const uint32_t tile = b * kCommitTiles + k;
            const uint4 v = reinterpret_cast<const uint4 *>(a.seg)[(size_t)tile * a.seg_cap + (code & 0xffffu)];
            const Entry ent{v.x, v.y, (int32_t)v.z, (int32_t)v.w};
            const uint32_t r = a.tiles[tile].rep;
            const RepDev &d = a.reps[r];
            const RepDyn &y = a.dyn[r];
            const uint32_t sub = d.algorithm == CPM_ALG_LOTTERY ? y.order[a.pass] - 1 : 0;
            int32_t dl, dg;
            if (evaluate_winner(d, ent, y.mcs, sub, &dl, &dg)) {
                atomicAdd(&s_cnt[k][2], 1u);
                commit_attempt(d, ent, dl, dg);
            } else {
                atomicAdd(&s_cnt[k][1], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) s_nwin = 0;
        __syncthreads();
    }
    if (threadIdx.x < kCommitTiles) {
        const int k = threadIdx.x;
        const uint32_t tile = b * kCommitTiles + k;
        if (tile < a.ntiles) {
            const uint32_t r = a.tiles[tile].rep;
            if (a.dyn[r].active && a.pass < (uint32_t)a.reps[r].rounds) {
                uint32_t *slot = a.slots + ((size_t)a.pass * a.ntiles + tile) * kSlotFields;
                slot[4] = s_cnt[k][0];
                slot[5] = s_cnt[k][1];
                slot[6] = s_cnt[k][2];
            }
        }
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
static hipError_t launch_propose_t(int algorithm, const PassArgs &a, hipStream_t stream) {
    if (algorithm == CPM_ALG_LOTTERY)
        hipLaunchKernelGGL((lot_propose_kernel<DIM, H, TX, TY, TZ>), dim3(a.ntiles), dim3(kBlock), 0,
                           stream, a);
    else
        hipLaunchKernelGGL((cb_propose_kernel<DIM, H, TX, TY, TZ>), dim3(a.ntiles), dim3(kBlock), 0,
                           stream, a);
    return hipGetLastError();
}

static hipError_t launch_commit(const PassArgs &a, hipStream_t stream) {
    const uint32_t wgs = (a.ntiles + kCommitTiles - 1) / kCommitTiles;
    hipLaunchKernelGGL(commit_kernel, dim3(wgs), dim3(kBlock), 0, stream, a);
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

extern "C" hipError_t cpm_internal_launch_propose(int config, int algorithm, const cpm::PassArgs *a,
                                                  hipStream_t stream) {
    using namespace cpm;
    switch (config) {
    case 0: return launch_propose_t<2, 1, 128, 16, 1>(algorithm, *a, stream);
    case 1: return launch_propose_t<2, 2, 128, 16, 1>(algorithm, *a, stream);
    case 2: return launch_propose_t<3, 1, 32, 8, 8>(algorithm, *a, stream);
    case 3: return launch_propose_t<3, 2, 32, 8, 8>(algorithm, *a, stream);
    case 4: return launch_propose_t<2, 1, 64, 32, 1>(algorithm, *a, stream);
    case 5: return launch_propose_t<2, 1, 256, 8, 1>(algorithm, *a, stream);
    case 6: return launch_propose_t<2, 1, 64, 16, 1>(algorithm, *a, stream);
    case 7: return launch_propose_t<2, 1, 32, 32, 1>(algorithm, *a, stream);
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
    hipLaunchKernelGGL(cpm::copy_probe_kernel, dim3(blocks), dim3(cpm::kBlock), 0, stream,
                       (const cpm::v4i *)src, (cpm::v4i *)dst, n4);
    return hipGetLastError();
}

extern "C" hipError_t cpm_internal_launch_report(const cpm::RepDev *reps, const uint32_t *slots,
                                                 uint32_t ntiles, int slot,
                                                 unsigned long long *out, hipStream_t stream) {
    hipLaunchKernelGGL(cpm::report_kernel, dim3(1), dim3(cpm::kBlock), 0, stream, reps, slots,
                       ntiles, slot, out);
    return hipGetLastError();
}
