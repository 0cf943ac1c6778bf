// Mixed-precision pair kernels: FP32 pair arithmetic on tile-relative
// coordinates with FP64 accumulators; operators, tile loop, dispatch.
#include "pair_kernels.h"

namespace sph {

// =============================================================== mixed ====
// FP32 pair arithmetic on coordinates relative to the tile origin (the FP64
// position of the tile's first target), converted while staging.  Per-lane
// partial sums are plain FP32 within a run (the pairs of one target in one
// 64-candidate chunk) and added to an FP64 accumulator after every chunk
// (two instructions per field, where a 2Sum in FP32 takes six).

// The hardware square root (v_sqrt_f32, within 1 ulp) instead of the
// correctly rounded sequence the compiler emits for sqrtf (about 15
// instructions).  The pair values need FP32 accuracy, not correct rounding:
// which pairs interact is decided by the predicate, not by r.
__device__ inline float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

// LDS ring of kept candidates (FP32), N slots: (x, y, z, |x|^2) -- (x, y, z,
// P/rho^2) for the force pass -- and, for the force pass, (vx, vy, vz,
// rho * 0.5 / avk); the other operators stage their own fourth component.
template <int N>
struct Lds32Pos {
    float4 pos[N];
};

template <int N>
struct Lds32Full {
    float4 pos[N], v[N];
};

// Tile-relative FP32 position of a fetched candidate and its |.|^2.
template <bool FLAT>
__device__ inline float4 rel_pos32(const Raw& r, double ox, double oy, double oz)
{
    const float x = (float)(r.x - ox);
    const float y = (float)(r.y - oy);
    const float z = FLAT ? 0.0f : (float)(r.z - oz);
    return make_float4(x, y, z, fmaf(z, z, fmaf(y, y, x * x)));
}

struct Density32 {
    static constexpr bool INTERACT = true;
    static constexpr bool USES_Q = false;
    template <int N>
    using Lds = Lds32Pos<N>;
    static constexpr bool STAGE_Q2 = true;  // pos.w = |x|^2
    static constexpr int WAVES = 6;  // a few spilled registers outside the loops
    static constexpr int WAVES_KEPT = 6;
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        float w;
    };
    struct Acc {
        double s;
    };
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <bool FLAT, class L>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, L& s)
    {
        s.pos[l] = rel_pos32<FLAT>(r, ox, oy, oz);
    }
    static __device__ void init(Acc& a) { a.s = 0.0; }
    static __device__ void zero(Part& q) { q.w = 0.0f; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt&, const L&,
                                    int, int, int, float, float, float, float ss, Part& q)
    {
        const float r = sqrt_hw(ss);
        const float ih = (float)d.g.inv_h;
        const float u = fmaf(-r, ih, 1.0f);
        const float u2 = u * u;
        q.w = fmaf(u2 * u2, fmaf(4.0f * ih, r, 1.0f), q.w);
    }
    static __device__ void fold(Acc& a, const Part& q, const SysDesc&) { a.s += (double)q.w; }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int out_col)
    {
        double w = a.s;
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
    static constexpr bool USES_Q = false;
    template <int N>
    using Lds = Lds32Full<N>;
    static constexpr bool STAGE_Q2 = false;  // pos.w = P/rho^2
    static constexpr int WAVES = 5;  // spills a few registers; LDS fits five blocks too
    static constexpr int WAVES_KEPT = 6;
    // rho is staged as rho * 0.5 / avk (target and candidates), so the
    // viscous term's (avk vx) / (0.5 (rho_p + rho_q) (r^2 + eps)) is
    // vx / ((rh_p + rh_q) (r^2 + eps))
    struct Tgt {
        float x, y, z, vx, vy, vz, pr, rh;
        bool fluid;
    };
    static __device__ double rho_scale(const SysDesc& d)
    {
        return d.ph.avk != 0.0 ? 0.5 / d.ph.avk : 0.0;
    }
    struct Part {
        float ax, ay, az, lx, ly, lz;
    };
    struct Acc {
        double ax, ay, az, lx, ly, lz;
    };
    static __device__ void load(const PairCtx& c, const SysDesc& d, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
        t.vx = (float)COL(c, C_VX, p); t.vy = (float)COL(c, C_VY, p); t.vz = (float)COL(c, C_VZ, p);
        t.pr = (float)c.prho2[p];
        t.rh = (float)(COL(c, C_RHO, p) * rho_scale(d));
        t.fluid = COL(c, C_TYPE, p) == 0.0;
    }
    static __device__ bool active(const Tgt& t) { return t.fluid; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_RHO, q);
        r.a1 = COL(c, C_VX, q); r.a2 = COL(c, C_VY, q); r.a3 = COL(c, C_VZ, q);
        r.a4 = c.prho2[q];
        return r;
    }
    template <bool FLAT, class L>
    static __device__ void put(const Raw& r, int l, const SysDesc& d, double ox, double oy,
                                    double oz, L& s)
    {
        float4 a = rel_pos32<FLAT>(r, ox, oy, oz);
        a.w = (float)r.a4;
        s.pos[l] = a;
        s.v[l] = make_float4((float)r.a1, (float)r.a2, (float)r.a3, (float)(r.a0 * rho_scale(d)));
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = a.lx = a.ly = a.lz = 0.0; }
    static __device__ void zero(Part& q) { q.ax = q.ay = q.az = q.lx = q.ly = q.lz = 0.0f; }
    // variant bit 0: Monaghan viscosity (avk != 0); bit 1: laminar (mu != 0)
    static constexpr int NVAR = 4;
    static __device__ int variant(const SysDesc& d)
    {
        return (d.ph.avk != 0.0 ? 1 : 0) | (d.ph.mu != 0.0 ? 2 : 0);
    }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const L& s, int jj, int, int, float dx, float dy,
                                    float dz, float ss, Part& q)
    {
        const float4 v = s.v[jj];
        const float r = sqrt_hw(ss);
        const float u = fmaf(-r, (float)d.g.inv_h, 1.0f);
        const float u3 = u * u * u;
        float A = t.pr + s.pos[jj].w;
        const float dvx = t.vx - v.x, dvy = t.vy - v.y, dvz = t.vz - v.z;
        if (V & 1) {
            // evaluated for every pair and selected, so the unrolled pair loop
            // stays one basic block
            const float vx = fmaf(dvz, dz, fmaf(dvy, dy, dvx * dx));
            const float den = (t.rh + v.w) * (ss + (float)d.ph.eps);
            const float Av = fmaf(vx, __builtin_amdgcn_rcpf(den), A);
            A = vx < 0.0f ? Av : A;
        }
        const float f = u3 * A;
        q.ax = fmaf(f, dx, q.ax); q.ay = fmaf(f, dy, q.ay); q.az = fmaf(f, dz, q.az);
        if (V & 2) {
            q.lx = fmaf(u3, dvx, q.lx); q.ly = fmaf(u3, dvy, q.ly); q.lz = fmaf(u3, dvz, q.lz);
        }
    }
    // the laminar sums only when the term is on (a uniform branch)
    static __device__ void fold(Acc& a, const Part& q, const SysDesc& d)
    {
        a.ax += (double)q.ax; a.ay += (double)q.ay; a.az += (double)q.az;
        if (d.ph.mu != 0.0) { a.lx += (double)q.lx; a.ly += (double)q.ly; a.lz += (double)q.lz; }
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        double ax = a.ax, ay = a.ay, az = a.az, lx = a.lx, ly = a.ly, lz = a.lz;
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

// internal_force! of tests/test_collision_2d.jl:75-78 in the mixed form:
// a += -kdW/rho0^2 u^3 (P_p + P_q) x_pq for every target, P staged in FP32.
struct ForceRho032 {
    static constexpr bool INTERACT = true;
    static constexpr bool USES_Q = false;
    template <int N>
    using Lds = Lds32Pos<N>;
    static constexpr bool STAGE_Q2 = false;  // pos.w = P
    static constexpr int WAVES = 6;
    static constexpr int WAVES_KEPT = 6;
    struct Tgt {
        float x, y, z, P;
    };
    struct Part {
        float ax, ay, az;
    };
    struct Acc {
        double ax, ay, az;
    };
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
        t.P = (float)COL(c, C_P, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_P, q);
        return r;
    }
    template <bool FLAT, class L>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, L& s)
    {
        float4 a = rel_pos32<FLAT>(r, ox, oy, oz);
        a.w = (float)r.a0;
        s.pos[l] = a;
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = 0.0; }
    static __device__ void zero(Part& q) { q.ax = q.ay = q.az = 0.0f; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const L& s, int jj, int, int, float dx, float dy,
                                    float dz, float ss, Part& q)
    {
        const float r = sqrt_hw(ss);
        const float u = fmaf(-r, (float)d.g.inv_h, 1.0f);
        const float f = u * u * u * (t.P + s.pos[jj].w);
        q.ax = fmaf(f, dx, q.ax); q.ay = fmaf(f, dy, q.ay); q.az = fmaf(f, dz, q.az);
    }
    static __device__ void fold(Acc& a, const Part& q, const SysDesc&)
    {
        a.ax += (double)q.ax; a.ay += (double)q.ay; a.az += (double)q.az;
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        double ax = a.ax, ay = a.ay, az = a.az;
        for (int o = T; o < 64; o <<= 1) {
            ax += __shfl_xor(ax, o); ay += __shfl_xor(ay, o); az += __shfl_xor(az, o);
        }
        if (!write) return;
        const double k = -d.ph.kdW / (d.ph.rho0 * d.ph.rho0);
        const double bx = c.overwrite ? 0.0 : COL(c, C_AX, p);
        const double by = c.overwrite ? 0.0 : COL(c, C_AY, p);
        const double bz = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        COL(c, C_AX, p) = bx + k * ax;
        COL(c, C_AY, p) = by + k * ay;
        COL(c, C_AZ, p) = bz + k * az;
    }
};

// balance_of_mass! (examples/collapse_dry.jl:112-115) in the mixed form:
// Drho += kdW u^3 (x_pq.v_pq + 2 nu (rho_p - rho_q)).  The densities are
// staged as rho - rho0 in FP32, so their difference keeps its precision.
struct Continuity32 {
    static constexpr bool INTERACT = true;
    static constexpr bool USES_Q = false;
    template <int N>
    using Lds = Lds32Full<N>;
    static constexpr bool STAGE_Q2 = false;  // pos.w = rho - rho0
    static constexpr int WAVES = 5;
    static constexpr int WAVES_KEPT = 5;
    struct Tgt {
        float x, y, z, vx, vy, vz, dr;
    };
    struct Part {
        float d;
    };
    struct Acc {
        double s;
    };
    static __device__ void load(const PairCtx& c, const SysDesc& d, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
        t.vx = (float)COL(c, C_VX, p); t.vy = (float)COL(c, C_VY, p); t.vz = (float)COL(c, C_VZ, p);
        t.dr = (float)(COL(c, C_RHO, p) - d.ph.rho0);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_RHO, q);
        r.a1 = COL(c, C_VX, q); r.a2 = COL(c, C_VY, q); r.a3 = COL(c, C_VZ, q);
        return r;
    }
    template <bool FLAT, class L>
    static __device__ void put(const Raw& r, int l, const SysDesc& d, double ox, double oy,
                               double oz, L& s)
    {
        float4 a = rel_pos32<FLAT>(r, ox, oy, oz);
        a.w = (float)(r.a0 - d.ph.rho0);
        s.pos[l] = a;
        s.v[l] = make_float4((float)r.a1, (float)r.a2, (float)r.a3, 0.0f);
    }
    static __device__ void init(Acc& a) { a.s = 0.0; }
    static __device__ void zero(Part& q) { q.d = 0.0f; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const L& s, int jj, int, int, float dx, float dy,
                                    float dz, float ss, Part& q)
    {
        const float r = sqrt_hw(ss);
        const float u = fmaf(-r, (float)d.g.inv_h, 1.0f);
        const float4 v = s.v[jj];
        const float xv = fmaf(dz, t.vz - v.z, fmaf(dy, t.vy - v.y, dx * (t.vx - v.x)));
        const float dd = fmaf(2.0f * (float)d.ph.nu, t.dr - s.pos[jj].w, xv);
        q.d = fmaf(u * u * u, dd, q.d);
    }
    static __device__ void fold(Acc& a, const Part& q, const SysDesc&) { a.s += (double)q.d; }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        double v = a.s;
        for (int o = T; o < 64; o <<= 1) v += __shfl_xor(v, o);
        if (!write) return;
        const double b = c.overwrite ? 0.0 : COL(c, C_DRHO, p);
        COL(c, C_DRHO, p) = b + d.ph.kdW * v;
    }
};

template <bool EMIT>
struct Pairs32 {
    static constexpr bool INTERACT = EMIT;
    static constexpr bool USES_Q = true;  // interact reads the candidate index
    template <int N>
    using Lds = Lds32Pos<N>;
    static constexpr bool STAGE_Q2 = true;
    static constexpr int WAVES = 5;
    static constexpr int WAVES_KEPT = 5;
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        int dummy;
    };
    struct Acc {
        int dummy;
    };
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <bool FLAT, class L>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, L& s)
    {
        s.pos[l] = rel_pos32<FLAT>(r, ox, oy, oz);
    }
    static __device__ void init(Acc&) {}
    static __device__ void zero(Part&) {}
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const L&,
                                    int, int p, int q, float, float, float, float, Part&)
    {
        const unsigned long long k = atomicAdd(c.emit_counter, 1ull);
        if ((long long)k < c.pair_cap) {
            c.pair_out[2 * k] = c.refidx[p];
            c.pair_out[2 * k + 1] = c.refidx[q];
        }
    }
    static __device__ void fold(Acc&, const Part&, const SysDesc&) {}
    static __device__ void finish(const PairCtx&, const SysDesc&, int, const Tgt&, const Acc&,
                                  int, bool, int) {}
};

struct Digest32 {
    static constexpr bool INTERACT = true;
    static constexpr bool USES_Q = true;  // interact reads the candidate index
    template <int N>
    using Lds = Lds32Pos<N>;
    static constexpr bool STAGE_Q2 = true;
    static constexpr int WAVES = 5;
    static constexpr int WAVES_KEPT = 5;
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        unsigned long long h, n;
    };
    using Acc = Part;
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <bool FLAT, class L>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, L& s)
    {
        s.pos[l] = rel_pos32<FLAT>(r, ox, oy, oz);
    }
    static __device__ void init(Acc& a) { a.h = a.n = 0; }
    static __device__ void zero(Part& a) { a.h = a.n = 0; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const L&,
                                    int, int, int q, float, float, float, float, Part& a)
    {
        a.h += mix64((unsigned long long)c.refidx[q] + 1ull);
        a.n += 1;
    }
    static __device__ void fold(Acc& a, const Part& q, const SysDesc&) { a.h += q.h; a.n += q.n; }
    static __device__ void finish(const PairCtx& c, const SysDesc&, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        digest_write(c, p, a.h, a.n, T, write);
    }
};

// A lane's run: cnt accepted slots of w (lowest first) against target ht.
template <int V, bool FLAT, int T, class Op, class L>
__device__ inline void walk32(const PairCtx& c, const SysDesc& d, const typename Op::Tgt& ht,
                              const L& lds, const int* q, unsigned long long w,
                              int cnt, int rb, int hp, typename Op::Part& part)
{
    if constexpr (V < Op::NVAR) {
#pragma unroll 2
        for (int k = 0; k < cnt; ++k) {
            const int b = __builtin_ctzll(w);  // w != 0: cnt <= its popcount
            w &= w - 1;
            const int jj = rb + slot_of_bit<T>(b);
            const float4 a = lds.pos[jj];
            const float dx = ht.x - a.x, dy = ht.y - a.y;
            const float dz = FLAT ? 0.0f : ht.z - a.z;
            const float ss = FLAT ? fmaf(dy, dy, dx * dx) : fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            Op::template interact<V>(c, d, ht, lds, jj, hp, Op::USES_Q ? q[jj] : 0, dx, dy, dz, ss,
                                     part);
        }
    }
}

template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op, bool KEPT>
__device__ inline void tile32(const PairCtx& c, int tile, int out_col, bool load_masks,
                              bool store_masks, bool kept_ok,
                              typename Op::template Lds<ring_slots<KEPT>()>& lds,
                              Cands<ROWS, PERIODIC, cands_q<Op, KEPT>()>& L,
                              Bal<typename Op::Part>& bal, unsigned& npairs,
                              unsigned long long& ntests)
{
    const int4 td = ld_const(c.tiles + tile);
    const int cell = td.x, first = td.y;
    const SysDesc d = ld_const(c.sys + ld_const(c.cell_sys + cell));
    if (c.only_flagged && !d.pass_on) return;
    const int lc = cell - d.cell_base;
    const long long mbase = mask_base(td);
    const int cend = ld_const(c.cell_start + cell + 1);
    const int nvalid = min(T, cend - first);
    const int lane = threadIdx.x;
    const int t = lane & (T - 1), s = lane / T;
    const int p = first + t;
    const bool valid = p < cend;
    const int pp = valid ? p : first;
    // tile origin (positions are not written by the pair kernels)
    const double ox = ld_const(&COL(c, C_X, first)), oy = ld_const(&COL(c, C_Y, first)),
                 oz = ld_const(&COL(c, C_Z, first));
    typename Op::Tgt tg;
    Op::load(c, d, pp, ox, oy, oz, tg);
    const bool act = valid && Op::active(tg);
    if (!__any(act)) return;
    const double s_thr = d.g.s_thr;
    // expanded test: |q|^2 - 2 p.q <= s_thr (1 -+ BAND) - |p|^2 -- below the
    // lower threshold the pair is in, above the upper one out, and in
    // between the exact FP64 test decides
    const float tz = FLAT ? 0.0f : tg.z;
    const float p2 = fmaf(tz, tz, fmaf(tg.y, tg.y, tg.x * tg.x));
    const float thr_in = (float)(s_thr * (1.0 - BAND)) - p2;
    const float thr_out = (float)(s_thr * (1.0 + BAND)) - p2;
    const float mx = -2.0f * tg.x, my = -2.0f * tg.y, mz = -2.0f * tz;
    // a pass that loads the masks takes the kept candidates from the list the
    // testing pass wrote; otherwise it streams the candidate list and culls
    // it again, keeping everything within the guard band of the box
    constexpr bool use_kept = KEPT;  // (the kernel took the tiles whose masks and lists are loaded)
    constexpr int RS = ring_slots<KEPT>();
    const bool store_kept = !KEPT && store_masks && kept_ok;
    const long long lbase = mbase / T * 64;
    const int nk = use_kept ? ld_const(c.kept_n + tile) : 0;
    const float cull_thr = (float)(s_thr * (1.0 + BAND));
    int total = 0, nr = 0;
    Box<float> box{};
    // (periodic systems need the ranges' image shifts in either case)
    if (use_kept && PERIODIC) nr = build_cands<ROWS, PERIODIC>(d, c.cell_start, lc, L, total);
    if (!use_kept) {
        // RN(x - ox) is monotonic in x: the box of the relative positions
        // is the relative box
        const Box<double> b = ld_const(reinterpret_cast<const Box<double>*>(c.boxes) + tile);
        box.x0 = (float)(b.x0 - ox); box.x1 = (float)(b.x1 - ox);
        box.y0 = (float)(b.y0 - oy); box.y1 = (float)(b.y1 - oy);
        box.z0 = (float)(b.z0 - oz); box.z1 = (float)(b.z1 - oz);
        nr = build_cands<ROWS, PERIODIC>(d, c.cell_start, lc, L, total);
    }
    __syncthreads();
    typename Op::Acc acc;
    Op::init(acc);
    int cur = 0, fill = 0, head = 0, nself = 0, chunk = 0;
    const int nstep = use_kept ? (nk + 63) / 64 : (total + 63) / 64;
    unsigned long long kword = 0;  // the chunk's mask word, loaded with its list entries
    for (int step = 0; step < nstep; ++step) {
        bool last;
        if (use_kept) {
            // one chunk of the kept list per step
            if (s == 0 && act) kword = c.masks[mbase + (long long)step * T + t];
            const int k = head + lane;
            if (k < nk) {
                const int q = c.kept[lbase + k];
                Seg sh{};
                if constexpr (PERIODIC) sh = range_shift(L, c.kept_rg[lbase + k]);
                Op::template put<FLAT>(Op::template fetch<PERIODIC>(c, q, sh), k & (RS - 1), d, ox,
                                       oy, oz, lds);
                if constexpr (cands_q<Op, KEPT>() > 1) L.q[k & (RS - 1)] = q;
            }
            fill = min(nk, head + 64);
            last = true;
        } else {
            const int base = step * 64;
            Raw raw;
            int rq = 0, rrg = 0;
            cur = advance_range(L, nr, cur, base);
            fetch_slot<ROWS, PERIODIC, Op>(c, L, total, cur, base + lane, raw, rq, rrg);
            // keep the candidates within reach of the box, and the tile's targets
            const float4 a = rel_pos32<FLAT>(raw, ox, oy, oz);
            const bool keep = base + lane < total &&
                              (box_dist2<FLAT>(box, a.x, a.y, a.z) <= cull_thr ||
                               (unsigned)(rq - first) < (unsigned)nvalid);
            const int pos = compact(L, fill, nself, keep, rq, rrg, first);
            if (pos >= 0) {
                Op::template put<FLAT>(raw, pos & (RS - 1), d, ox, oy, oz, lds);
                if (store_kept) {
                    c.kept[lbase + pos] = rq;
                    if constexpr (PERIODIC) c.kept_rg[lbase + pos] = (unsigned char)rrg;
                }
            }
            last = base + 64 >= total;
        }
        __syncthreads();
        while (fill - head >= 64 || (last && fill > head)) {
            if (fill - head < 64) {
                // the last chunk: pad its free slots
                const int k = head + lane;
                if (k >= fill) {
                    Op::template put<FLAT>(pad_raw(), k & (RS - 1), d, ox, oy, oz, lds);
                    if constexpr (cands_q<Op, KEPT>() > 1) L.q[k & (RS - 1)] = first;
                    if constexpr (PERIODIC) L.rg[k & (RS - 1)] = 0;
                }
                __syncthreads();
            }
            const int rb = head & (RS - 64);  // the chunk's part of the ring
            // accepted-slot word of target t (owner lanes publish it)
            unsigned long long word = 0;
            if (use_kept) {
                word = kword;
            } else if (load_masks) {
                if (s == 0 && act) word = c.masks[mbase + (long long)chunk * T + t];
            } else {
                // 32-bit words, slots shifted in from the highest bit down
                MaskT<T> m = 0, unsure = 0;
#pragma unroll
                for (int h = 0; h < (T + 31) / 32; ++h) {
                    unsigned min = 0, mout = 0;
#pragma unroll 4
                    for (int j = (T < 32 ? T : 32) - 1; j >= 0; --j) {
                        const float4 a = lds.pos[rb + s + (64 / T) * (32 * h + j)];
                        const float q2 =
                            Op::STAGE_Q2 ? a.w : fmaf(a.z, a.z, fmaf(a.y, a.y, a.x * a.x));
                        float e = fmaf(mx, a.x, q2);
                        e = fmaf(my, a.y, e);
                        if (!FLAT) e = fmaf(mz, a.z, e);
                        min = shift_in_le(min, e, thr_in);
                        mout = shift_in_le(mout, e, thr_out);
                    }
                    m |= (MaskT<T>)min << (32 * h);
                    unsure |= (MaskT<T>)mout << (32 * h);
                }
                unsure &= ~m;
                if (!act) m = unsure = 0;
                // the guard band: exact FP64 predicate of the reference (rare)
                while (__any(unsure != 0)) {
                    if (unsure) {
                        const int j = lowest_bit<T>(unsure);
                        unsure &= unsure - 1;
                        const int jj = rb + s + (64 / T) * j;
                        const int q = L.q[jj];
                        Seg sh{};
                        if constexpr (PERIODIC) sh = range_shift(L, L.rg[jj]);
                        const double e1 =
                            COL(c, C_X, pp) - shifted<PERIODIC>(COL(c, C_X, q), sh.sx);
                        const double e2 =
                            COL(c, C_Y, pp) - shifted<PERIODIC>(COL(c, C_Y, q), sh.sy);
                        double es = e1 * e1 + e2 * e2;
                        if (!FLAT) {
                            const double e3 =
                                COL(c, C_Z, pp) - shifted<PERIODIC>(COL(c, C_Z, q), sh.sz);
                            es = es + e3 * e3;
                        }
                        if (es <= s_thr) m |= MaskT<T>(1) << j;
                    }
                }
                drop_self<ROWS, PERIODIC, T>(L, nself, head, t, s, m);
                word = (unsigned long long)m << (T == 64 ? 0 : s * T);
#pragma unroll
                for (int o = T; o < 64; o <<= 1) word |= __shfl_xor(word, o);
                if (s != 0) word = 0;
            }
            if (s == 0) bal.w[t] = word;
            bool own;
            int tt, r0, cnt, frun, nrun;
            const int tot = split_runs(lane, __popcll(word), bal, own, tt, r0, cnt, frun, nrun);
            if (own) {
                // every target's pairs on its owner lane
                typename Op::Part part;
                Op::zero(part);
                const int n = __popcll(word);
                if (Op::INTERACT && n > 0) {
                    const int hp = first + t;
                    switch (Op::variant(d)) {
                    case 0: walk32<0, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    case 1: walk32<1, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    case 2: walk32<2, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    default: walk32<3, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    }
                }
                if (n > 0) Op::fold(acc, part, d);
            } else if (tot > 0) {
                typename Op::Part part;
                Op::zero(part);
                // the run's target record from its owner lane (every lane
                // takes part in the shuffle)
                const typename Op::Tgt ht = shfl_value(tg, tt < 0 ? 0 : tt);
                if (Op::INTERACT && cnt > 0) {
                    const unsigned long long w = drop_low_bits(bal.w[tt], r0);
                    const int hp = first + tt;
                    // the operator's uniform switches are taken outside the loop
                    switch (Op::variant(d)) {
                    case 0: walk32<0, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 1: walk32<1, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 2: walk32<2, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    default: walk32<3, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    }
                }
                bal.part[lane] = part;
                __syncthreads();
                for (int k = 0; k < nrun; ++k) Op::fold(acc, bal.part[frun + k], d);
            }
            npairs += (unsigned)__popcll(word);
            if (store_masks && s == 0) c.masks[mbase + (long long)chunk * T + t] = word;
            ++chunk;
            head += 64;
            __syncthreads();
        }
    }
    // distance tests: the box test of every candidate, the pair test of
    // every kept candidate for every target
    if (!load_masks) ntests += (unsigned long long)total + (unsigned long long)fill * nvalid;
    if (store_kept && lane == 0) c.kept_n[tile] = fill;
    Op::finish(c, d, p, tg, acc, T, act && s == 0, out_col);
}

// KEPT: the kernel of a pass that loads the masks and the kept-candidate
// lists (the force pass after a density pass); it holds no candidate test,
// stream or cull, so it needs fewer registers.  The other kernel does
// everything else.  A loading pass launches both: whichever does not match
// the device's mask state leaves at once.
template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op, bool KEPT>
__global__ __launch_bounds__(64, KEPT ? Op::WAVES_KEPT : Op::WAVES) void k_pair32(PairCtx c,
                                                                               int out_col)
{
    __shared__ typename Op::template Lds<ring_slots<KEPT>()> lds;
    __shared__ Cands<ROWS, PERIODIC, cands_q<Op, KEPT>()> L;
    __shared__ Bal<typename Op::Part> bal;
    const int ntiles = ld_const(c.meta + META_NTILES);
    const bool ok = c.masks != nullptr && ld_const(c.meta + META_MASK_OK) != 0;
    const bool load_masks = ok && c.mask_mode == MASK_LOAD;
    const bool store_masks = ok && c.mask_mode == MASK_TEST_STORE;
    const bool kept_ok = ok && c.kept != nullptr && ld_const(c.meta + META_KEPT_OK) != 0;
    if ((load_masks && kept_ok) != KEPT) return;
    TileCursor tc;
    tile_cursor_init(tc);
    // accepted pairs of all the wave's tiles, one atomic per wave
    unsigned npairs = 0;
    unsigned long long ntests = 0;
    for (int tile = next_tile(c, ntiles, tc); tile < ntiles; tile = next_tile(c, ntiles, tc))
        tile32<ROWS, FLAT, PERIODIC, T, Op, KEPT>(c, tile, out_col, load_masks, store_masks,
                                                  kept_ok, lds, L, bal, npairs, ntests);
    if (c.pair_counter) count_pairs(c, npairs);
    count_tests(c, ntests);
}


// ============================================================ dispatch ====

template <int ROWS, bool FLAT, bool PER, int T, bool KEPT>
static void launch_tile_k(int op, const PairCtx& c, int grid, hipStream_t st)
{
    switch (op) {
    case P_DENSITY: k_pair32<ROWS, FLAT, PER, T, Density32, KEPT><<<grid, 64, 0, st>>>(c, C_RHO); break;
    case P_DENSITY_RHO0: k_pair32<ROWS, FLAT, PER, T, Density32, KEPT><<<grid, 64, 0, st>>>(c, C_RHO0); break;
    case P_FORCE: k_pair32<ROWS, FLAT, PER, T, Force32, KEPT><<<grid, 64, 0, st>>>(c, 0); break;
    case P_COUNT: k_pair32<ROWS, FLAT, PER, T, Pairs32<false>, KEPT><<<grid, 64, 0, st>>>(c, 0); break;
    case P_EMIT: k_pair32<ROWS, FLAT, PER, T, Pairs32<true>, KEPT><<<grid, 64, 0, st>>>(c, 0); break;
    case P_DIGEST: k_pair32<ROWS, FLAT, PER, T, Digest32, KEPT><<<grid, 64, 0, st>>>(c, 0); break;
    case P_FORCE_RHO0: k_pair32<ROWS, FLAT, PER, T, ForceRho032, KEPT><<<grid, 64, 0, st>>>(c, 0); break;
    case P_CONTINUITY: k_pair32<ROWS, FLAT, PER, T, Continuity32, KEPT><<<grid, 64, 0, st>>>(c, 0); break;
    }
}

// A pass that loads the masks launches the kernel for the kept-list path and
// the general one; the device's mask state picks which of them works.  The
// density, count, emit and digest passes never load kept lists.
template <int ROWS, bool FLAT, bool PER, int T>
static void launch_tile(int op, const PairCtx& c, int grid, hipStream_t st)
{
    const bool loads = c.mask_mode == MASK_LOAD && op != P_DENSITY && op != P_DENSITY_RHO0 &&
                       op != P_COUNT && op != P_EMIT && op != P_DIGEST;
    if (loads) launch_tile_k<ROWS, FLAT, PER, T, true>(op, c, grid, st);
    launch_tile_k<ROWS, FLAT, PER, T, false>(op, c, grid, st);
}

template <int ROWS, bool FLAT, bool PER>
static void launch_rows(int op, const PairCtx& c, int grid, hipStream_t st)
{
    switch (c.T) {
    case 8: launch_tile<ROWS, FLAT, PER, 8>(op, c, grid, st); break;
    case 16: launch_tile<ROWS, FLAT, PER, 16>(op, c, grid, st); break;
    case 32: launch_tile<ROWS, FLAT, PER, 32>(op, c, grid, st); break;
    default: launch_tile<ROWS, FLAT, PER, 64>(op, c, grid, st); break;
    }
}

void launch_pair32(int op, int rows, int flat, int periodic, const PairCtx& c, int grid, hipStream_t st)
{
    if (rows == 3) {
        if (flat) {
            if (periodic) launch_rows<3, true, true>(op, c, grid, st);
            else launch_rows<3, true, false>(op, c, grid, st);
        } else {
            if (periodic) launch_rows<3, false, true>(op, c, grid, st);
            else launch_rows<3, false, false>(op, c, grid, st);
        }
    } else {
        if (periodic) launch_rows<9, false, true>(op, c, grid, st);
        else launch_rows<9, false, false>(op, c, grid, st);
    }
}

}  // namespace sph
