// Pair-interaction kernels: the device form of _apply_binary! /
// apply_binary! (src/core.jl:94-129) for the registered operators.
//
// Work decomposition.  A tile is up to T targets of one cell (T = 8..64).
// One 64-lane wave handles a tile: lane = (target t, slice s), t = lane % T,
// s = lane / T, with S = 64 / T slices.  The candidates of a target are the
// cells key + key_diff of the reference; with di innermost, each of the 3
// (2D) or 9 (3D) (dj, dk) rows is the contiguous cell range
// [key + L1*dj + L1*L2*dk - 1, ... + 1], clamped to [0, key_max) exactly as
// the reference's linear bounds test (src/core.jl:98), so rows wrap and
// small grids visit aliased cells as often as the reference does.  Since
// the particles are cell-sorted, each row is one contiguous particle range.
// The wave streams it through LDS in chunks of 64 (one coalesced load per
// lane).  Each chunk is handled in two phases:
//   1. mask: every lane tests the chunk entries jj = s (mod S) and sets bit
//      jj of a 64-bit mask for each accepted candidate -- or, after a
//      density pass on the same cell list, loads that pass's mask;
//   2. interact: every lane walks the set bits of its mask, so the pair
//      arithmetic runs once per accepted pair, not once per candidate.
// At the end the S slices of a target are combined with wave shuffles.
//
// Predicate (src/core.jl:104-105, src/algebra.jl:49-60): accept q != p when
// RN(sqrt((d1*d1 + d2*d2) + d3*d3)) <= h.  sqrt is correctly rounded and
// monotonic, so this equals s <= s_thr with s_thr = max{s : RN(sqrt(s)) <= h}
// (computed on the host).  The FP64 kernels evaluate s without FMA
// (-ffp-contract=off), i.e. bit-exactly as the reference.  The mixed
// kernels evaluate s in FP32 on tile-relative coordinates and fall back to
// the exact FP64 test inside a guard band around s_thr, so their neighbour
// sets (and masks) are the reference's too.
#include "sph_internal.h"

namespace sph {

#define COL(c, f, i) ((c).cols[(long long)(f) * (c).cap + (i)])

// Guard band of the mixed predicate, relative to s_thr.  The FP32 squared
// distance of tile-relative coordinates (|coords| <= 2*sqrt(3)*h) is within
// ~2e-6 * h^2 of the exact one; the band is 5x wider.
static constexpr double BAND = 1.0e-5;

__device__ inline void count_pairs(const PairCtx& c, unsigned n)
{
    // wave reduction of the per-lane counts, one atomic per wave
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) n += __shfl_xor(n, o);
    if (__lane_id() == 0 && n) atomicAdd(c.pair_counter, (unsigned long long)n);
}

// Compensated (2Sum) addition of a chunk partial into (s, e).
__device__ inline void two_sum_add(float& s, float& e, float x)
{
    const float t = s + x;
    const float bp = t - s;
    e += (s - (t - bp)) + (x - bp);
    s = t;
}

// Bits jj with jj % S == 0 (S = 1, 2, 4, 8).
__device__ inline unsigned long long slice_pattern(int S)
{
    return ~0ull / ((1ull << S) - 1ull);
}

// Tiles come from a work queue (one atomic per tile); every wave leaves
// the loop once the queue passes the tile count.
__device__ inline int next_tile(const PairCtx& c)
{
    int t = 0;
    if (__lane_id() == 0) t = atomicAdd(c.queue, 1);
    return __shfl(t, 0);
}

__device__ inline long long mask_base(const int4& td)
{
    return (long long)(unsigned)td.z | ((long long)td.w << 32);
}

// ================================================================ FP64 ====

// LDS staging of one 64-candidate chunk (FP64): xy, (z, rho), (vx, vy),
// (vz, aux) where aux is P/rho^2 (force) or P (force with rho0).
struct Lds64 {
    double2 xy[64], zr[64], v01[64], v2a[64];
};

// Candidate positions are staged with the periodic image shift of their
// range added (zero for reference systems, where x + 0.0 == x).
__device__ inline void stage_pos64(const PairCtx& c, int q, int l, const Seg& g, Lds64& s)
{
    s.xy[l] = make_double2(COL(c, C_X, q) + g.sx, COL(c, C_Y, q) + g.sy);
    s.zr[l].x = COL(c, C_Z, q) + g.sz;
}

// rho += m W (OUT = C_RHO) or rho0 += m W (OUT = C_RHO0).
template <int OUT>
struct Density64 {
    static constexpr bool INTERACT = true;
    struct Tgt {
        double x, y, z;
    };
    struct Acc {
        double w;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, Lds64& s)
    {
        stage_pos64(c, q, l, g, s);
    }
    static __device__ void init(Acc& a) { a.w = 0.0; }
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt&, const Lds64&,
                                    int, int, int, double, double, double, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double x = r * d.g.inv_h;
        const double u = 1.0 - x;
        const double u2 = u * u;
        a.w += u2 * u2 * (1.0 + 4.0 * x);
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) a.w += __shfl_xor(a.w, o);
        if (!write) return;
        double v = c.overwrite ? 0.0 : COL(c, OUT, p);
        v = v + d.ph.kW * a.w;
        if (c.self) v = v + d.ph.kW;
        COL(c, OUT, p) = v;
    }
};

// internal_force! (examples/collapse_dry.jl:135-141) + Monaghan viscosity.
// With u = 1 - r/h and rDW = C_dW u^3 / h^(d+2), kdW = m C_dW / h^(d+2):
//   pressure  a += -kdW u^3 (P_p/rho_p^2 + P_q/rho_q^2) x_pq
//   viscosity a += -kdW u^3 Pi x_pq,  Pi = avk (v.x)/(rhobar (r^2 + eps)), v.x < 0
//   laminar   a += 2 kdW mu / rho0^2 u^3 v_pq
struct Force64 {
    static constexpr bool INTERACT = true;
    struct Tgt {
        double x, y, z, vx, vy, vz, pr, rho;
        bool fluid;
    };
    struct Acc {
        double ax, ay, az, lx, ly, lz;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
        t.vx = COL(c, C_VX, p); t.vy = COL(c, C_VY, p); t.vz = COL(c, C_VZ, p);
        t.pr = c.prho2[p];
        t.rho = COL(c, C_RHO, p);
        t.fluid = COL(c, C_TYPE, p) == 0.0;
    }
    static __device__ bool active(const Tgt& t) { return t.fluid; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, Lds64& s)
    {
        s.xy[l] = make_double2(COL(c, C_X, q) + g.sx, COL(c, C_Y, q) + g.sy);
        s.zr[l] = make_double2(COL(c, C_Z, q) + g.sz, COL(c, C_RHO, q));
        s.v01[l] = make_double2(COL(c, C_VX, q), COL(c, C_VY, q));
        s.v2a[l] = make_double2(COL(c, C_VZ, q), c.prho2[q]);
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = a.lx = a.ly = a.lz = 0.0; }
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds64& s, int jj, int, int, double dx, double dy,
                                    double dz, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double x = r * d.g.inv_h;
        const double u = 1.0 - x;
        const double u3 = u * u * u;
        const double2 v01 = s.v01[jj], v2a = s.v2a[jj];
        double A = t.pr + v2a.y;
        const double dvx = t.vx - v01.x, dvy = t.vy - v01.y, dvz = t.vz - v2a.x;
        if (d.ph.avk != 0.0) {
            double vx = dvx * dx + dvy * dy;
            vx = vx + dvz * dz;
            if (vx < 0.0) {
                const double rhobar = 0.5 * (t.rho + s.zr[jj].y);
                A = A + d.ph.avk * vx / (rhobar * (r * r + d.ph.eps));
            }
        }
        const double f = u3 * A;
        a.ax += f * dx; a.ay += f * dy; a.az += f * dz;
        if (d.ph.mu != 0.0) { a.lx += u3 * dvx; a.ly += u3 * dvy; a.lz += u3 * dvz; }
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) {
            a.ax += __shfl_xor(a.ax, o); a.ay += __shfl_xor(a.ay, o); a.az += __shfl_xor(a.az, o);
            a.lx += __shfl_xor(a.lx, o); a.ly += __shfl_xor(a.ly, o); a.lz += __shfl_xor(a.lz, o);
        }
        if (!write) return;
        const double kp = -d.ph.kdW;
        const double kl = 2.0 * d.ph.kdW * d.ph.mu / (d.ph.rho0 * d.ph.rho0);
        double ax = c.overwrite ? 0.0 : COL(c, C_AX, p);
        double ay = c.overwrite ? 0.0 : COL(c, C_AY, p);
        double az = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        ax = ax + kp * a.ax; ay = ay + kp * a.ay; az = az + kp * a.az;
        if (d.ph.mu != 0.0) { ax = ax + kl * a.lx; ay = ay + kl * a.ly; az = az + kl * a.lz; }
        COL(c, C_AX, p) = ax; COL(c, C_AY, p) = ay; COL(c, C_AZ, p) = az;
    }
};

// internal_force! of tests/test_collision_2d.jl:75-78 (P/rho0^2, all targets).
struct ForceRho064 {
    static constexpr bool INTERACT = true;
    struct Tgt {
        double x, y, z, P;
    };
    struct Acc {
        double ax, ay, az;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p); t.P = COL(c, C_P, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, Lds64& s)
    {
        stage_pos64(c, q, l, g, s);
        s.v2a[l].y = COL(c, C_P, q);
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = 0.0; }
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds64& s, int jj, int, int, double dx, double dy,
                                    double dz, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double u = 1.0 - r * d.g.inv_h;
        const double f = u * u * u * (t.P + s.v2a[jj].y);
        a.ax += f * dx; a.ay += f * dy; a.az += f * dz;
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) {
            a.ax += __shfl_xor(a.ax, o); a.ay += __shfl_xor(a.ay, o); a.az += __shfl_xor(a.az, o);
        }
        if (!write) return;
        const double k = -d.ph.kdW / (d.ph.rho0 * d.ph.rho0);
        const double ax = c.overwrite ? 0.0 : COL(c, C_AX, p);
        const double ay = c.overwrite ? 0.0 : COL(c, C_AY, p);
        const double az = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        COL(c, C_AX, p) = ax + k * a.ax;
        COL(c, C_AY, p) = ay + k * a.ay;
        COL(c, C_AZ, p) = az + k * a.az;
    }
};

// balance_of_mass! (examples/collapse_dry.jl:112-115):
//   Drho += kdW u^3 (x_pq.v_pq + 2 nu (rho_p - rho_q))
struct Continuity64 {
    static constexpr bool INTERACT = true;
    struct Tgt {
        double x, y, z, vx, vy, vz, rho;
    };
    struct Acc {
        double d;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
        t.vx = COL(c, C_VX, p); t.vy = COL(c, C_VY, p); t.vz = COL(c, C_VZ, p);
        t.rho = COL(c, C_RHO, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, Lds64& s)
    {
        s.xy[l] = make_double2(COL(c, C_X, q) + g.sx, COL(c, C_Y, q) + g.sy);
        s.zr[l] = make_double2(COL(c, C_Z, q) + g.sz, COL(c, C_RHO, q));
        s.v01[l] = make_double2(COL(c, C_VX, q), COL(c, C_VY, q));
        s.v2a[l].x = COL(c, C_VZ, q);
    }
    static __device__ void init(Acc& a) { a.d = 0.0; }
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds64& s, int jj, int, int, double dx, double dy,
                                    double dz, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double u = 1.0 - r * d.g.inv_h;
        const double2 v01 = s.v01[jj];
        double xv = dx * (t.vx - v01.x) + dy * (t.vy - v01.y);
        xv = xv + dz * (t.vz - s.v2a[jj].x);
        a.d += u * u * u * (xv + 2.0 * d.ph.nu * (t.rho - s.zr[jj].y));
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) a.d += __shfl_xor(a.d, o);
        if (!write) return;
        const double v = c.overwrite ? 0.0 : COL(c, C_DRHO, p);
        COL(c, C_DRHO, p) = v + d.ph.kdW * a.d;
    }
};

// Accepted-pair count (COUNT) or emission of (refidx p, refidx q) (EMIT).
template <bool EMIT>
struct Pairs64 {
    static constexpr bool INTERACT = EMIT;
    struct Tgt {
        double x, y, z;
    };
    struct Acc {
        int dummy;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, Lds64& s)
    {
        stage_pos64(c, q, l, g, s);
    }
    static __device__ void init(Acc&) {}
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const Lds64&,
                                    int, int p, int q, double, double, double, double, Acc&)
    {
        const unsigned long long k = atomicAdd(c.emit_counter, 1ull);
        if ((long long)k < c.pair_cap) {
            c.pair_out[2 * k] = c.refidx[p];
            c.pair_out[2 * k + 1] = c.refidx[q];
        }
    }
    static __device__ void finish(const PairCtx&, const SysDesc&, int, const Tgt&, Acc&, int,
                                  bool) {}
};

template <int ROWS, bool FLAT, class Op>
__device__ inline void tile64(const PairCtx& c, int tile, bool load_masks, bool store_masks,
                              Lds64& lds)
{
    const int4 td = c.tiles[tile];
    const int cell = td.x, first = td.y;
    const SysDesc& d = c.sys[c.cell_sys[cell]];
    if (c.only_flagged && !d.pass_on) return;
    const int lc = cell - d.cell_base;
    const long long mbase = mask_base(td);
    const int cend = c.cell_start[cell + 1];
    const int lane = threadIdx.x;
    const int T = c.T, S = 64 / T;
    const int t = lane & (T - 1), s = lane / T;
    const int p = first + t;
    const bool valid = p < cend;
    typename Op::Tgt tg;
    Op::load(c, valid ? p : first, tg);
    const bool act = valid && Op::active(tg);
    if (!__any(act)) return;
    const unsigned long long sbits = slice_pattern(S) << s;
    const double s_thr = d.g.s_thr;
    typename Op::Acc acc;
    Op::init(acc);
    unsigned npairs = 0;
    int chunk = 0;
    for (int row = 0; row < ROWS; ++row) {
        Seg segs[2];
        const int nseg = row_segments<ROWS>(d, c.cell_start, lc, row, segs);
        for (int sg = 0; sg < nseg; ++sg) {
            const Seg seg = segs[sg];
            for (int b = seg.j0; b < seg.j1; b += 64, ++chunk) {
                const int nb = min(64, seg.j1 - b);
                if (lane < nb) Op::stage(c, b + lane, lane, seg, lds);
                __syncthreads();
                unsigned long long m = 0ull;
                if (load_masks) {
                    if (act) m = c.masks[mbase + (long long)chunk * T + t] & sbits;
                } else {
                    for (int j = 0; j < nb; j += S) {
                        const int jj = j + s;
                        if (jj < nb) {
                            const double2 xy = lds.xy[jj];
                            const double dx = tg.x - xy.x, dy = tg.y - xy.y;
                            double ss = dx * dx + dy * dy;
                            if (!FLAT) {
                                const double dz = tg.z - lds.zr[jj].x;
                                ss = ss + dz * dz;
                            }
                            if (act && ss <= s_thr && b + jj != p) m |= 1ull << jj;
                        }
                    }
                    if (store_masks) {
                        unsigned long long full = m;
                        for (int o = T; o < 64; o <<= 1) full |= __shfl_xor(full, o);
                        if (s == 0) c.masks[mbase + (long long)chunk * T + t] = full;
                    }
                }
                npairs += (unsigned)__popcll(m);
                if (Op::INTERACT) {
                    while (__any(m != 0ull)) {
                        if (m) {
                            const int jj = __ffsll((long long)m) - 1;
                            m &= m - 1ull;
                            const double2 xy = lds.xy[jj];
                            const double dx = tg.x - xy.x, dy = tg.y - xy.y;
                            const double dz = FLAT ? 0.0 : tg.z - lds.zr[jj].x;
                            double ss = dx * dx + dy * dy;
                            if (!FLAT) ss = ss + dz * dz;
                            Op::interact(c, d, tg, lds, jj, p, b + jj, dx, dy, dz, ss, acc);
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    if (c.pair_counter) count_pairs(c, npairs);
    Op::finish(c, d, p, tg, acc, T, act && s == 0);
}

template <int ROWS, bool FLAT, class Op>
__global__ __launch_bounds__(64) void k_pair64(PairCtx c)
{
    __shared__ Lds64 lds;
    const int ntiles = c.meta[META_NTILES];
    const bool ok = c.masks != nullptr && c.meta[META_MASK_OK] != 0;
    const bool load_masks = ok && c.mask_mode == MASK_LOAD;
    const bool store_masks = ok && c.mask_mode == MASK_TEST_STORE;
    for (int tile = next_tile(c); tile < ntiles; tile = next_tile(c))
        tile64<ROWS, FLAT, Op>(c, tile, load_masks, store_masks, lds);
}

// =============================================================== mixed ====
// FP32 pair arithmetic on coordinates relative to the tile origin (the FP64
// position of the tile's first target), converted while staging.  Per-lane
// partial sums are plain FP32 within a 64-candidate chunk and folded into a
// compensated (sum, error) pair after every chunk.

// LDS staging (FP32): (x, y, z, P/rho^2), (vx, vy, vz, rho).
struct Lds32 {
    float4 a[64], v[64];
};

struct Density32 {
    static constexpr bool INTERACT = true;
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        float w;
    };
    struct Acc {
        float s, e;
    };
    static __device__ void load(const PairCtx& c, int p, double ox, double oy, double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, double ox,
                                 double oy, double oz, Lds32& s)
    {
        s.a[l] = make_float4((float)((COL(c, C_X, q) + g.sx) - ox),
                             (float)((COL(c, C_Y, q) + g.sy) - oy),
                             (float)((COL(c, C_Z, q) + g.sz) - oz), 0.0f);
    }
    static __device__ void init(Acc& a) { a.s = a.e = 0.0f; }
    static __device__ void zero(Part& q) { q.w = 0.0f; }
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt&, const Lds32&,
                                    int, int, int, float, float, float, float ss, Part& q)
    {
        const float r = __builtin_sqrtf(ss);
        const float x = r * (float)d.g.inv_h;
        const float u = 1.0f - x;
        const float u2 = u * u;
        q.w = fmaf(u2 * u2, fmaf(4.0f, x, 1.0f), q.w);
    }
    static __device__ void fold(Acc& a, const Part& q) { two_sum_add(a.s, a.e, q.w); }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int out_col)
    {
        double w = (double)a.s + (double)a.e;
        for (int o = T; o < 64; o <<= 1) w += __shfl_xor(w, o);
        if (!write) return;
        double v = c.overwrite ? 0.0 : COL(c, out_col, p);
        v = v + d.ph.kW * w;
        if (c.self) v = v + d.ph.kW;
        COL(c, out_col, p) = v;
    }
};

struct Force32 {
    static constexpr bool INTERACT = true;
    struct Tgt {
        float x, y, z, vx, vy, vz, pr, rho;
        bool fluid;
    };
    struct Part {
        float ax, ay, az, lx, ly, lz;
    };
    struct Acc {
        float sx, sy, sz, ex, ey, ez, lx, ly, lz, fx, fy, fz;
    };
    static __device__ void load(const PairCtx& c, int p, double ox, double oy, double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
        t.vx = (float)COL(c, C_VX, p); t.vy = (float)COL(c, C_VY, p); t.vz = (float)COL(c, C_VZ, p);
        t.pr = (float)c.prho2[p];
        t.rho = (float)COL(c, C_RHO, p);
        t.fluid = COL(c, C_TYPE, p) == 0.0;
    }
    static __device__ bool active(const Tgt& t) { return t.fluid; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, double ox,
                                 double oy, double oz, Lds32& s)
    {
        s.a[l] = make_float4((float)((COL(c, C_X, q) + g.sx) - ox),
                             (float)((COL(c, C_Y, q) + g.sy) - oy),
                             (float)((COL(c, C_Z, q) + g.sz) - oz), (float)c.prho2[q]);
        s.v[l] = make_float4((float)COL(c, C_VX, q), (float)COL(c, C_VY, q),
                             (float)COL(c, C_VZ, q), (float)COL(c, C_RHO, q));
    }
    static __device__ void init(Acc& a)
    {
        a.sx = a.sy = a.sz = a.ex = a.ey = a.ez = 0.0f;
        a.lx = a.ly = a.lz = a.fx = a.fy = a.fz = 0.0f;
    }
    static __device__ void zero(Part& q) { q.ax = q.ay = q.az = q.lx = q.ly = q.lz = 0.0f; }
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds32& s, int jj, int, int, float dx, float dy,
                                    float dz, float ss, Part& q)
    {
        const float r = __builtin_sqrtf(ss);
        const float u = 1.0f - r * (float)d.g.inv_h;
        const float u3 = u * u * u;
        const float4 a = s.a[jj], v = s.v[jj];
        float A = t.pr + a.w;
        const float dvx = t.vx - v.x, dvy = t.vy - v.y, dvz = t.vz - v.z;
        if (d.ph.avk != 0.0) {
            const float vx = fmaf(dvz, dz, fmaf(dvy, dy, dvx * dx));
            if (vx < 0.0f) {
                const float rhobar = 0.5f * (t.rho + v.w);
                const float den = rhobar * (ss + (float)d.ph.eps);
                A = fmaf((float)d.ph.avk * vx, __builtin_amdgcn_rcpf(den), A);
            }
        }
        const float f = u3 * A;
        q.ax = fmaf(f, dx, q.ax); q.ay = fmaf(f, dy, q.ay); q.az = fmaf(f, dz, q.az);
        if (d.ph.mu != 0.0) {
            q.lx = fmaf(u3, dvx, q.lx); q.ly = fmaf(u3, dvy, q.ly); q.lz = fmaf(u3, dvz, q.lz);
        }
    }
    static __device__ void fold(Acc& a, const Part& q)
    {
        two_sum_add(a.sx, a.ex, q.ax);
        two_sum_add(a.sy, a.ey, q.ay);
        two_sum_add(a.sz, a.ez, q.az);
        two_sum_add(a.lx, a.fx, q.lx);
        two_sum_add(a.ly, a.fy, q.ly);
        two_sum_add(a.lz, a.fz, q.lz);
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        double ax = (double)a.sx + (double)a.ex, ay = (double)a.sy + (double)a.ey,
               az = (double)a.sz + (double)a.ez;
        double lx = (double)a.lx + (double)a.fx, ly = (double)a.ly + (double)a.fy,
               lz = (double)a.lz + (double)a.fz;
        for (int o = T; o < 64; o <<= 1) {
            ax += __shfl_xor(ax, o); ay += __shfl_xor(ay, o); az += __shfl_xor(az, o);
            lx += __shfl_xor(lx, o); ly += __shfl_xor(ly, o); lz += __shfl_xor(lz, o);
        }
        if (!write) return;
        const double kp = -d.ph.kdW;
        const double kl = 2.0 * d.ph.kdW * d.ph.mu / (d.ph.rho0 * d.ph.rho0);
        double ox = c.overwrite ? 0.0 : COL(c, C_AX, p);
        double oy = c.overwrite ? 0.0 : COL(c, C_AY, p);
        double oz = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        ox = ox + kp * ax; oy = oy + kp * ay; oz = oz + kp * az;
        if (d.ph.mu != 0.0) { ox = ox + kl * lx; oy = oy + kl * ly; oz = oz + kl * lz; }
        COL(c, C_AX, p) = ox; COL(c, C_AY, p) = oy; COL(c, C_AZ, p) = oz;
    }
};

template <bool EMIT>
struct Pairs32 {
    static constexpr bool INTERACT = EMIT;
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        int dummy;
    };
    struct Acc {
        int dummy;
    };
    static __device__ void load(const PairCtx& c, int p, double ox, double oy, double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    static __device__ void stage(const PairCtx& c, int q, int l, const Seg& g, double ox,
                                 double oy, double oz, Lds32& s)
    {
        s.a[l] = make_float4((float)((COL(c, C_X, q) + g.sx) - ox),
                             (float)((COL(c, C_Y, q) + g.sy) - oy),
                             (float)((COL(c, C_Z, q) + g.sz) - oz), 0.0f);
    }
    static __device__ void init(Acc&) {}
    static __device__ void zero(Part&) {}
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const Lds32&,
                                    int, int p, int q, float, float, float, float, Part&)
    {
        const unsigned long long k = atomicAdd(c.emit_counter, 1ull);
        if ((long long)k < c.pair_cap) {
            c.pair_out[2 * k] = c.refidx[p];
            c.pair_out[2 * k + 1] = c.refidx[q];
        }
    }
    static __device__ void fold(Acc&, const Part&) {}
    static __device__ void finish(const PairCtx&, const SysDesc&, int, const Tgt&, const Acc&,
                                  int, bool, int) {}
};

template <int ROWS, bool FLAT, class Op>
__device__ inline void tile32(const PairCtx& c, int tile, int out_col, bool load_masks,
                              bool store_masks, Lds32& lds)
{
    const int4 td = c.tiles[tile];
    const int cell = td.x, first = td.y;
    const SysDesc& d = c.sys[c.cell_sys[cell]];
    if (c.only_flagged && !d.pass_on) return;
    const int lc = cell - d.cell_base;
    const long long mbase = mask_base(td);
    const int cend = c.cell_start[cell + 1];
    const int lane = threadIdx.x;
    const int T = c.T, S = 64 / T;
    const int t = lane & (T - 1), s = lane / T;
    const int p = first + t;
    const bool valid = p < cend;
    const int pp = valid ? p : first;
    const double ox = COL(c, C_X, first), oy = COL(c, C_Y, first), oz = COL(c, C_Z, first);
    typename Op::Tgt tg;
    Op::load(c, pp, ox, oy, oz, tg);
    const bool act = valid && Op::active(tg);
    if (!__any(act)) return;
    const double px = COL(c, C_X, pp), py = COL(c, C_Y, pp), pz = COL(c, C_Z, pp);
    const double s_thr = d.g.s_thr;
    const float thr_lo = (float)(s_thr * (1.0 - BAND));
    const float thr_hi = (float)(s_thr * (1.0 + BAND));
    const unsigned long long sbits = slice_pattern(S) << s;
    typename Op::Acc acc;
    Op::init(acc);
    unsigned npairs = 0;
    int chunk = 0;
    for (int row = 0; row < ROWS; ++row) {
        Seg segs[2];
        const int nseg = row_segments<ROWS>(d, c.cell_start, lc, row, segs);
        for (int sg = 0; sg < nseg; ++sg) {
            const Seg seg = segs[sg];
            for (int b = seg.j0; b < seg.j1; b += 64, ++chunk) {
                const int nb = min(64, seg.j1 - b);
                if (lane < nb) Op::stage(c, b + lane, lane, seg, ox, oy, oz, lds);
                __syncthreads();
                unsigned long long m = 0ull;
                if (load_masks) {
                    if (act) m = c.masks[mbase + (long long)chunk * T + t] & sbits;
                } else {
                    for (int j = 0; j < nb; j += S) {
                        const int jj = j + s;
                        const bool in = jj < nb;
                        const int jr = in ? jj : 0;
                        const float4 a = lds.a[jr];
                        const float dx = tg.x - a.x, dy = tg.y - a.y;
                        const float dz = FLAT ? 0.0f : tg.z - a.z;
                        const float ss =
                            FLAT ? fmaf(dy, dy, dx * dx) : fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        const int q = b + jr;
                        bool ok = ss <= thr_lo;
                        const bool band = in && act && ss > thr_lo && ss <= thr_hi;
                        if (__any(band)) {
                            if (band) {
                                // exact FP64 predicate of the reference
                                const double e1 = px - (COL(c, C_X, q) + seg.sx);
                                const double e2 = py - (COL(c, C_Y, q) + seg.sy);
                                double es = e1 * e1 + e2 * e2;
                                if (!FLAT) {
                                    const double e3 = pz - (COL(c, C_Z, q) + seg.sz);
                                    es = es + e3 * e3;
                                }
                                ok = es <= s_thr;
                            }
                        }
                        if (in && act && ok && q != p) m |= 1ull << jj;
                    }
                    if (store_masks) {
                        unsigned long long full = m;
                        for (int o = T; o < 64; o <<= 1) full |= __shfl_xor(full, o);
                        if (s == 0) c.masks[mbase + (long long)chunk * T + t] = full;
                    }
                }
                npairs += (unsigned)__popcll(m);
                if (Op::INTERACT) {
                    typename Op::Part part;
                    Op::zero(part);
                    while (__any(m != 0ull)) {
                        if (m) {
                            const int jj = __ffsll((long long)m) - 1;
                            m &= m - 1ull;
                            const float4 a = lds.a[jj];
                            const float dx = tg.x - a.x, dy = tg.y - a.y;
                            const float dz = FLAT ? 0.0f : tg.z - a.z;
                            const float ss =
                                FLAT ? fmaf(dy, dy, dx * dx) : fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                            Op::interact(c, d, tg, lds, jj, p, b + jj, dx, dy, dz, ss, part);
                        }
                    }
                    Op::fold(acc, part);
                }
                __syncthreads();
            }
        }
    }
    if (c.pair_counter) count_pairs(c, npairs);
    Op::finish(c, d, p, tg, acc, T, act && s == 0, out_col);
}

template <int ROWS, bool FLAT, class Op>
__global__ __launch_bounds__(64) void k_pair32(PairCtx c, int out_col)
{
    __shared__ Lds32 lds;
    const int ntiles = c.meta[META_NTILES];
    const bool ok = c.masks != nullptr && c.meta[META_MASK_OK] != 0;
    const bool load_masks = ok && c.mask_mode == MASK_LOAD;
    const bool store_masks = ok && c.mask_mode == MASK_TEST_STORE;
    for (int tile = next_tile(c); tile < ntiles; tile = next_tile(c))
        tile32<ROWS, FLAT, Op>(c, tile, out_col, load_masks, store_masks, lds);
}

// ========================================================= diagnostics ====
// Candidates tested per pass: for every particle, the particles of the 27
// (9) cells it visits (src/core.jl:96-102).  out[0] over all targets,
// out[1] over fluid targets (the force pass).
template <int ROWS>
__global__ __launch_bounds__(256) void k_candidates(PairCtx c, const int* __restrict__ key,
                                                    unsigned long long* out)
{
    const int n = c.meta[META_N];
    const int p = blockIdx.x * 256 + threadIdx.x;
    unsigned long long all = 0, fluid = 0;
    if (p < n) {
        const int cell = key[p];
        const SysDesc& d = c.sys[c.cell_sys[cell]];
        for (int row = 0; row < ROWS; ++row) {
            Seg segs[2];
            const int ns = row_segments<ROWS>(d, c.cell_start, cell - d.cell_base, row, segs);
            for (int q = 0; q < ns; ++q) all += (unsigned long long)(segs[q].j1 - segs[q].j0);
        }
        if (COL(c, C_TYPE, p) == 0.0) fluid = all;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        all += __shfl_xor(all, o);
        fluid += __shfl_xor(fluid, o);
    }
    if (__lane_id() == 0) {
        atomicAdd(&out[0], all);
        atomicAdd(&out[1], fluid);
    }
}

void launch_candidates(int rows, const PairCtx& c, const int* key, unsigned long long* out,
                       long long n, hipStream_t st)
{
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (nb == 0) return;
    if (rows == 3) k_candidates<3><<<nb, 256, 0, st>>>(c, key, out);
    else k_candidates<9><<<nb, 256, 0, st>>>(c, key, out);
}

// ============================================================ dispatch ====

template <int ROWS, bool FLAT>
static void launch_rows(int op, int precision, const PairCtx& c, int grid, hipStream_t st)
{
    if (precision == SPH_FP64) {
        switch (op) {
        case P_DENSITY: k_pair64<ROWS, FLAT, Density64<C_RHO>><<<grid, 64, 0, st>>>(c); break;
        case P_DENSITY_RHO0: k_pair64<ROWS, FLAT, Density64<C_RHO0>><<<grid, 64, 0, st>>>(c); break;
        case P_FORCE: k_pair64<ROWS, FLAT, Force64><<<grid, 64, 0, st>>>(c); break;
        case P_FORCE_RHO0: k_pair64<ROWS, FLAT, ForceRho064><<<grid, 64, 0, st>>>(c); break;
        case P_CONTINUITY: k_pair64<ROWS, FLAT, Continuity64><<<grid, 64, 0, st>>>(c); break;
        case P_COUNT: k_pair64<ROWS, FLAT, Pairs64<false>><<<grid, 64, 0, st>>>(c); break;
        case P_EMIT: k_pair64<ROWS, FLAT, Pairs64<true>><<<grid, 64, 0, st>>>(c); break;
        }
    } else {
        switch (op) {
        case P_DENSITY: k_pair32<ROWS, FLAT, Density32><<<grid, 64, 0, st>>>(c, C_RHO); break;
        case P_DENSITY_RHO0: k_pair32<ROWS, FLAT, Density32><<<grid, 64, 0, st>>>(c, C_RHO0); break;
        case P_FORCE: k_pair32<ROWS, FLAT, Force32><<<grid, 64, 0, st>>>(c, 0); break;
        case P_COUNT: k_pair32<ROWS, FLAT, Pairs32<false>><<<grid, 64, 0, st>>>(c, 0); break;
        case P_EMIT: k_pair32<ROWS, FLAT, Pairs32<true>><<<grid, 64, 0, st>>>(c, 0); break;
        // operators without a mixed form run in FP64
        case P_FORCE_RHO0: k_pair64<ROWS, FLAT, ForceRho064><<<grid, 64, 0, st>>>(c); break;
        case P_CONTINUITY: k_pair64<ROWS, FLAT, Continuity64><<<grid, 64, 0, st>>>(c); break;
        }
    }
}

void launch_pair(int op, int precision, int rows, int flat, const PairCtx& c, int grid,
                 hipStream_t st)
{
    if (grid <= 0) return;
    if (rows == 3) {
        if (flat) launch_rows<3, true>(op, precision, c, grid, st);
        else launch_rows<3, false>(op, precision, c, grid, st);
    } else {
        launch_rows<9, false>(op, precision, c, grid, st);
    }
}

}  // namespace sph
