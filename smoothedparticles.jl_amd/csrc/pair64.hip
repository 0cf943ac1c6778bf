// FP64 pair kernels (the plain drop-in): operators, tile loop, dispatch.
#include "pair_kernels.h"

namespace sph {

// ================================================================ FP64 ====

// LDS ring of kept candidates (FP64): positions only, or xy, (z, rho),
// (vx, vy), (vz, aux) where aux is P/rho^2 (force) or P (force with rho0).
template <int N>
struct Lds64Pos {
    double2 xy[N];
    double z[N];
};

template <int N>
struct Lds64Full {
    double2 xy[N], zr[N], v01[N], v2a[N];
};

template <int N>
__device__ inline double lds_z(const Lds64Pos<N>& s, int j)
{
    return s.z[j];
}
template <int N>
__device__ inline double lds_z(const Lds64Full<N>& s, int j)
{
    return s.zr[j].x;
}

// rho += m W (OUT = C_RHO) or rho0 += m W (OUT = C_RHO0).
template <int OUT>
struct Density64 {
    static constexpr bool INTERACT = true;
    static constexpr bool USES_Q = false;
    static constexpr int WAVES = 6;  // occupancy the registers are fitted to (80 VGPRs)
    static constexpr int WAVES_KEPT = 5;
    template <int N>
    using Lds = Lds64Pos<N>;
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
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <class L>
    static __device__ void put(const Raw& r, int l, L& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.z[l] = r.z;
    }
    static __device__ void init(Acc& a) { a.w = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { a.w = 0.0; }
    static __device__ void fold(Acc& a, const Part& q) { a.w += q.w; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt&, const L&,
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
    static constexpr bool USES_Q = false;
    static constexpr int WAVES = 3;  // occupancy the registers are fitted to
    static constexpr int WAVES_KEPT = 4;  // 125 VGPRs; LDS 7.9 KB with the one-chunk ring
    template <int N>
    using Lds = Lds64Full<N>;
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
    template <class L>
    static __device__ void put(const Raw& r, int l, L& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.zr[l] = make_double2(r.z, r.a0);
        s.v01[l] = make_double2(r.a1, r.a2);
        s.v2a[l] = make_double2(r.a3, r.a4);
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = a.lx = a.ly = a.lz = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { init(a); }
    static __device__ void fold(Acc& a, const Part& q)
    {
        a.ax += q.ax; a.ay += q.ay; a.az += q.az; a.lx += q.lx; a.ly += q.ly; a.lz += q.lz;
    }
    // variant bit 0: Monaghan viscosity (avk != 0); bit 1: laminar (mu != 0)
    static constexpr int NVAR = 4;
    static __device__ int variant(const SysDesc& d)
    {
        return (d.ph.avk != 0.0 ? 1 : 0) | (d.ph.mu != 0.0 ? 2 : 0);
    }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const L& s, int jj, int, int, double dx, double dy,
                                    double dz, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double x = r * d.g.inv_h;
        const double u = 1.0 - x;
        const double u3 = u * u * u;
        const double2 v01 = s.v01[jj], v2a = s.v2a[jj];
        double A = t.pr + v2a.y;
        const double dvx = t.vx - v01.x, dvy = t.vy - v01.y, dvz = t.vz - v2a.x;
        if (V & 1) {
            double vx = dvx * dx + dvy * dy;
            vx = vx + dvz * dz;
            if (vx < 0.0) {
                const double rhobar = 0.5 * (t.rho + s.zr[jj].y);
                A = A + d.ph.avk * vx / (rhobar * (r * r + d.ph.eps));
            }
        }
        const double f = u3 * A;
        a.ax += f * dx; a.ay += f * dy; a.az += f * dz;
        if (V & 2) { a.lx += u3 * dvx; a.ly += u3 * dvy; a.lz += u3 * dvz; }
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
    static constexpr bool USES_Q = false;
    static constexpr int WAVES = 4;  // occupancy the registers are fitted to
    static constexpr int WAVES_KEPT = 4;
    template <int N>
    using Lds = Lds64Full<N>;
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
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_P, q);
        return r;
    }
    template <class L>
    static __device__ void put(const Raw& r, int l, L& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.zr[l].x = r.z;
        s.v2a[l].y = r.a0;
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { init(a); }
    static __device__ void fold(Acc& a, const Part& q) { a.ax += q.ax; a.ay += q.ay; a.az += q.az; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const L& s, int jj, int, int, double dx, double dy,
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
    static constexpr bool USES_Q = false;
    static constexpr int WAVES = 4;  // occupancy the registers are fitted to
    static constexpr int WAVES_KEPT = 4;
    template <int N>
    using Lds = Lds64Full<N>;
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
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_RHO, q);
        r.a1 = COL(c, C_VX, q); r.a2 = COL(c, C_VY, q); r.a3 = COL(c, C_VZ, q);
        return r;
    }
    template <class L>
    static __device__ void put(const Raw& r, int l, L& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.zr[l] = make_double2(r.z, r.a0);
        s.v01[l] = make_double2(r.a1, r.a2);
        s.v2a[l].x = r.a3;
    }
    static __device__ void init(Acc& a) { a.d = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { a.d = 0.0; }
    static __device__ void fold(Acc& a, const Part& q) { a.d += q.d; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const L& s, int jj, int, int, double dx, double dy,
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
    static constexpr bool USES_Q = true;  // interact reads the candidate index
    static constexpr int WAVES = 5;  // occupancy the registers are fitted to
    static constexpr int WAVES_KEPT = 5;
    template <int N>
    using Lds = Lds64Pos<N>;
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
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <class L>
    static __device__ void put(const Raw& r, int l, L& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.z[l] = r.z;
    }
    static __device__ void init(Acc&) {}
    using Part = Acc;
    static __device__ void zero(Part&) {}
    static __device__ void fold(Acc&, const Part&) {}
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const L&,
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

struct Digest64 {
    static constexpr bool INTERACT = true;
    static constexpr bool USES_Q = true;  // interact reads the candidate index
    static constexpr int WAVES = 5;
    static constexpr int WAVES_KEPT = 5;
    template <int N>
    using Lds = Lds64Pos<N>;
    struct Tgt {
        double x, y, z;
    };
    struct Acc {
        unsigned long long h, n;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <class L>
    static __device__ void put(const Raw& r, int l, L& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.z[l] = r.z;
    }
    static __device__ void init(Acc& a) { a.h = a.n = 0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { a.h = a.n = 0; }
    static __device__ void fold(Acc& a, const Part& q) { a.h += q.h; a.n += q.n; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V, class L>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const L&,
                                    int, int, int q, double, double, double, double, Acc& a)
    {
        a.h += mix64((unsigned long long)c.refidx[q] + 1ull);
        a.n += 1;
    }
    static __device__ void finish(const PairCtx& c, const SysDesc&, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        digest_write(c, p, a.h, a.n, T, write);
    }
};

// A lane's run: cnt accepted slots of w (lowest first) against target ht.
template <int V, bool FLAT, int T, class Op, class L>
__device__ inline void walk64(const PairCtx& c, const SysDesc& d, const typename Op::Tgt& ht,
                              const L& lds, const int* q, unsigned long long w,
                              int cnt, int rb, int hp, typename Op::Part& part)
{
    if constexpr (V < Op::NVAR) {
        for (int k = 0; k < cnt; ++k) {
            const int b = __builtin_ctzll(w);  // w != 0: cnt <= its popcount
            w &= w - 1;
            const int jj = rb + slot_of_bit<T>(b);
            const double2 xy = lds.xy[jj];
            const double dx = ht.x - xy.x, dy = ht.y - xy.y;
            const double dz = FLAT ? 0.0 : ht.z - lds_z(lds, jj);
            double ss = dx * dx + dy * dy;
            if (!FLAT) ss = ss + dz * dz;
            Op::template interact<V>(c, d, ht, lds, jj, hp, Op::USES_Q ? q[jj] : 0, dx, dy, dz, ss,
                                     part);
        }
    }
}

template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op, bool KEPT>
__device__ inline void tile64(const PairCtx& c, int tile, bool load_masks, bool store_masks,
                              bool kept_ok,
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
    typename Op::Tgt tg;
    Op::load(c, valid ? p : first, tg);
    const bool act = valid && Op::active(tg);
    if (!__any(act)) return;
    const double s_thr = d.g.s_thr;
    // a pass that loads the masks takes the kept candidates from the list the
    // testing pass wrote; otherwise it streams the candidate list and culls
    // it again
    constexpr bool use_kept = KEPT;  // (the kernel took the tiles whose masks and lists are loaded)
    constexpr int RS = ring_slots<KEPT>();
    const bool store_kept = !KEPT && store_masks && kept_ok;
    const long long lbase = mbase / T * 64;
    const int nk = use_kept ? ld_const(c.kept_n + tile) : 0;
    int total = 0, nr = 0;
    Box<double> box{};
    if (!use_kept) box = ld_const(reinterpret_cast<const Box<double>*>(c.boxes) + tile);
    // (periodic systems need the ranges' image shifts in either case)
    if (!use_kept || PERIODIC) nr = build_cands<ROWS, PERIODIC>(d, c.cell_start, lc, L, total);
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
                Op::put(Op::template fetch<PERIODIC>(c, q, sh), k & (RS - 1), lds);
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
            const bool keep = base + lane < total &&
                              (box_dist2<FLAT>(box, raw.x, raw.y, raw.z) <= s_thr ||
                               (unsigned)(rq - first) < (unsigned)nvalid);
            const int pos = compact(L, fill, nself, keep, rq, rrg, first);
            if (pos >= 0) {
                Op::put(raw, pos & (RS - 1), lds);
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
                    Op::put(pad_raw(), k & (RS - 1), lds);
                    if constexpr (cands_q<Op, KEPT>() > 1) L.q[k & (RS - 1)] = first;
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
                MaskT<T> m = 0;
#pragma unroll
                for (int h = 0; h < (T + 31) / 32; ++h) {
                    unsigned mh = 0;
#pragma unroll 4
                    for (int j = (T < 32 ? T : 32) - 1; j >= 0; --j) {
                        const int jj = rb + s + (64 / T) * (32 * h + j);
                        const double2 xy = lds.xy[jj];
                        const double dx = tg.x - xy.x, dy = tg.y - xy.y;
                        double ss = dx * dx + dy * dy;
                        if (!FLAT) {
                            const double dz = tg.z - lds_z(lds, jj);
                            ss = ss + dz * dz;
                        }
                        mh = shift_in_le(mh, ss, s_thr);
                    }
                    m |= (MaskT<T>)mh << (32 * h);
                }
                if (!act) m = 0;
                drop_self<ROWS, PERIODIC, T>(L, nself, head, t, s, m);
                word = (unsigned long long)m << (T == 64 ? 0 : s * T);
#pragma unroll
                for (int o = T; o < 64; o <<= 1) word |= __shfl_xor(word, o);
                if (s != 0) word = 0;
                if (store_masks && s == 0) c.masks[mbase + (long long)chunk * T + t] = word;
            }
            const int pop = __popcll(word);
            npairs += (unsigned)pop;
            if (s == 0) bal.w[t] = word;
            bool own;
            int tt, r0, cnt, frun, nrun;
            const int tot = split_runs(lane, pop, bal, own, tt, r0, cnt, frun, nrun);
            if (own) {
                // every target's pairs on its owner lane
                typename Op::Part part;
                Op::zero(part);
                const int n = __popcll(word);
                if (Op::INTERACT && n > 0) {
                    const int hp = first + t;
                    switch (Op::variant(d)) {
                    case 0: walk64<0, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    case 1: walk64<1, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    case 2: walk64<2, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    default: walk64<3, FLAT, T, Op>(c, d, tg, lds, L.q, word, n, rb, hp, part); break;
                    }
                }
                if (n > 0) Op::fold(acc, part);
            } else if (tot > 0) {
                typename Op::Part part;
                Op::zero(part);
                const typename Op::Tgt ht = shfl_value(tg, tt < 0 ? 0 : tt);
                if (Op::INTERACT && cnt > 0) {
                    const unsigned long long w = drop_low_bits(bal.w[tt], r0);
                    const int hp = first + tt;
                    // the operator's uniform switches are taken outside the loop
                    switch (Op::variant(d)) {
                    case 0: walk64<0, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 1: walk64<1, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 2: walk64<2, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    default: walk64<3, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    }
                }
                bal.part[lane] = part;
                __syncthreads();
                for (int k = 0; k < nrun; ++k) Op::fold(acc, bal.part[frun + k]);
            }
            ++chunk;
            head += 64;
            __syncthreads();
        }
    }
    // distance tests: the box test of every candidate, the pair test of
    // every kept candidate for every target
    if (!load_masks) ntests += (unsigned long long)total + (unsigned long long)fill * nvalid;
    if (store_kept && lane == 0) c.kept_n[tile] = fill;
    Op::finish(c, d, p, tg, acc, T, act && s == 0);
}

// KEPT: as k_pair32, the kernel of the loading pass with kept lists.
template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op, bool KEPT>
__global__ __launch_bounds__(64, KEPT ? Op::WAVES_KEPT : Op::WAVES) void k_pair64(PairCtx c)
{
    __shared__ typename Op::template Lds<ring_slots<KEPT>()> lds;
    __shared__ Cands<ROWS, PERIODIC, cands_q<Op, KEPT>()> L;
    __shared__ Bal<typename Op::Part> bal;
    const int ntiles = ld_const(c.meta + META_NTILES);
    const bool ok = c.masks != nullptr && ld_const(c.meta + META_MASK_OK) != 0;
    const bool kept_ok = ok && c.kept != nullptr && ld_const(c.meta + META_KEPT_OK) != 0;
    // masks stored by a mixed pass number the candidates its FP32 cull kept;
    // without that pass's kept lists this kernel's FP64 cull would keep
    // others, so it tests instead
    const bool load_masks = ok && c.mask_mode == MASK_LOAD && (kept_ok || !c.masks_mixed);
    const bool store_masks = ok && c.mask_mode == MASK_TEST_STORE;
    if ((load_masks && kept_ok) != KEPT) return;
    TileCursor tc;
    tile_cursor_init(tc);
    // accepted pairs of all the wave's tiles, one atomic per wave
    unsigned npairs = 0;
    unsigned long long ntests = 0;
    for (int tile = next_tile(c, ntiles, tc); tile < ntiles; tile = next_tile(c, ntiles, tc))
        tile64<ROWS, FLAT, PERIODIC, T, Op, KEPT>(c, tile, load_masks, store_masks, kept_ok, lds, L,
                                                  bal, npairs, ntests);
    if (c.pair_counter) count_pairs(c, npairs);
    count_tests(c, ntests);
}


// ============================================================ dispatch ====

template <int ROWS, bool FLAT, bool PER, int T, bool KEPT>
static void launch_tile_k(int op, const PairCtx& c, int grid, hipStream_t st)
{
    switch (op) {
    case P_DENSITY: k_pair64<ROWS, FLAT, PER, T, Density64<C_RHO>, KEPT><<<grid, 64, 0, st>>>(c); break;
    case P_DENSITY_RHO0: k_pair64<ROWS, FLAT, PER, T, Density64<C_RHO0>, KEPT><<<grid, 64, 0, st>>>(c); break;
    case P_FORCE: k_pair64<ROWS, FLAT, PER, T, Force64, KEPT><<<grid, 64, 0, st>>>(c); break;
    case P_FORCE_RHO0: k_pair64<ROWS, FLAT, PER, T, ForceRho064, KEPT><<<grid, 64, 0, st>>>(c); break;
    case P_CONTINUITY: k_pair64<ROWS, FLAT, PER, T, Continuity64, KEPT><<<grid, 64, 0, st>>>(c); break;
    case P_COUNT: k_pair64<ROWS, FLAT, PER, T, Pairs64<false>, KEPT><<<grid, 64, 0, st>>>(c); break;
    case P_EMIT: k_pair64<ROWS, FLAT, PER, T, Pairs64<true>, KEPT><<<grid, 64, 0, st>>>(c); break;
    case P_DIGEST: k_pair64<ROWS, FLAT, PER, T, Digest64, KEPT><<<grid, 64, 0, st>>>(c); break;
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

void launch_pair64(int op, int rows, int flat, int periodic, const PairCtx& c, int grid, hipStream_t st)
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
