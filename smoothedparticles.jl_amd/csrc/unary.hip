// Unary operators (apply_unary!, src/core.jl:138-142), global reductions
// and the field transfers of ParticleField (src/structs.jl:118-125).
//
// All kernels run over the cell-sorted device order and read the particle
// count from meta[META_N], so they can be enqueued without a host round
// trip.  The arithmetic of every unary operator is the reference's FP64
// expression, evaluated in the same order, so these are bit-exact.
#include "sph_internal.h"

namespace sph {

#define UCOL(cols, cap, f, i) ((cols)[(long long)(f) * (cap) + (i)])

// ------------------------------------------------------------ operators --
__global__ __launch_bounds__(256) void k_unary(int op, const int* meta, double* cols,
                                               long long cap, const SysDesc* __restrict__ sysd,
                                               const int* __restrict__ sysid, double* prho2)
{
    const int n = meta[META_N];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Phys& ph = sysd[sysid[i]].ph;
    switch (op) {
    case SPH_OP_RESET_RHO:  // tests/test_collision_2d.jl:84-86
        UCOL(cols, cap, C_RHO, i) = 0.0;
        break;
    case SPH_OP_RESET_A:  // tests/test_collision_2d.jl:80-82
        UCOL(cols, cap, C_AX, i) = 0.0;
        UCOL(cols, cap, C_AY, i) = 0.0;
        UCOL(cols, cap, C_AZ, i) = 0.0;
        break;
    case SPH_OP_FIND_PRESSURE: {  // examples/collapse_dry.jl:126
        const double rho = UCOL(cols, cap, C_RHO, i);
        const double P = (ph.c * ph.c) * (rho - ph.rho0);
        UCOL(cols, cap, C_P, i) = P;
        if (prho2) prho2[i] = P / (rho * rho);
        break;
    }
    case SPH_OP_FIND_PRESSURE_RHO0: {  // tests/test_collision_2d.jl:71-73
        const double rho = UCOL(cols, cap, C_RHO, i);
        const double P = (ph.c * ph.c) * (rho - UCOL(cols, cap, C_RHO0, i));
        UCOL(cols, cap, C_P, i) = P;
        if (prho2) prho2[i] = P / (rho * rho);
        break;
    }
    case SPH_OP_FIND_PRESSURE_CONTINUITY: {  // examples/collapse_dry.jl:123-127
        double rho = UCOL(cols, cap, C_RHO, i);
        rho = rho + UCOL(cols, cap, C_DRHO, i) * ph.dt;
        UCOL(cols, cap, C_RHO, i) = rho;
        UCOL(cols, cap, C_DRHO, i) = 0.0;
        const double P = (ph.c * ph.c) * (rho - ph.rho0);
        UCOL(cols, cap, C_P, i) = P;
        if (prho2) prho2[i] = P / (rho * rho);
        break;
    }
    case SPH_OP_ACCELERATE: {  // examples/collapse_dry.jl:155-159
        if (UCOL(cols, cap, C_TYPE, i) != 0.0) break;
        const double k = ph.coeff * ph.dt;
        for (int d = 0; d < 3; ++d)
            UCOL(cols, cap, C_VX + d, i) =
                UCOL(cols, cap, C_VX + d, i) + k * (UCOL(cols, cap, C_AX + d, i) + ph.g[d]);
        break;
    }
    case SPH_OP_MOVE: {  // examples/collapse_dry.jl:148-153
        // p.Dv = VEC0 for every particle, then the fluid drift
        UCOL(cols, cap, C_AX, i) = 0.0;
        UCOL(cols, cap, C_AY, i) = 0.0;
        UCOL(cols, cap, C_AZ, i) = 0.0;
        if (UCOL(cols, cap, C_TYPE, i) != 0.0) break;
        const double k = ph.coeff * ph.dt;
        for (int d = 0; d < 3; ++d)
            UCOL(cols, cap, C_X + d, i) = UCOL(cols, cap, C_X + d, i) + k * UCOL(cols, cap, C_VX + d, i);
        break;
    }
    case SPH_OP_DRIFT: {  // tests/test_collision_2d.jl:88-90: p.x += dt*p.v
        const double k = ph.coeff * ph.dt;
        for (int d = 0; d < 3; ++d)
            UCOL(cols, cap, C_X + d, i) = UCOL(cols, cap, C_X + d, i) + k * UCOL(cols, cap, C_VX + d, i);
        break;
    }
    case SPH_OP_KICK: {  // tests/test_collision_2d.jl:92-94: p.v += 0.5*dt*p.a
        const double k = ph.coeff * ph.dt;
        for (int d = 0; d < 3; ++d)
            UCOL(cols, cap, C_VX + d, i) = UCOL(cols, cap, C_VX + d, i) + k * UCOL(cols, cap, C_AX + d, i);
        break;
    }
    default:
        break;
    }
}

// P/rho^2 per particle, consumed by the force pass
// (examples/collapse_dry.jl:138: p.P/p.rho^2 + q.P/q.rho^2).
__global__ __launch_bounds__(256) void k_prho2(const int* meta, const double* cols, long long cap,
                                               double* prho2)
{
    const int n = meta[META_N];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double rho = UCOL(cols, cap, C_RHO, i);
    prho2[i] = UCOL(cols, cap, C_P, i) / (rho * rho);
}

// Fused kick and drift of the symplectic Euler step (fluid only):
// v += dt (a + g); x += dt v  (examples/collapse_dry.jl:148-159 with a
// full step instead of two half steps).
__global__ __launch_bounds__(256) void k_kick_drift(const int* meta, double* cols, long long cap,
                                                    const SysDesc* __restrict__ sysd,
                                                    const int* __restrict__ sysid)
{
    const int n = meta[META_N];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (UCOL(cols, cap, C_TYPE, i) != 0.0) return;
    const Phys& ph = sysd[sysid[i]].ph;
    const double k = ph.coeff * ph.dt;
    for (int d = 0; d < 3; ++d) {
        const double v = UCOL(cols, cap, C_VX + d, i) + k * (UCOL(cols, cap, C_AX + d, i) + ph.g[d]);
        UCOL(cols, cap, C_VX + d, i) = v;
        UCOL(cols, cap, C_X + d, i) = UCOL(cols, cap, C_X + d, i) + k * v;
    }
}

// The unary stages of the time loop of examples/collapse_dry.jl:202-211,
// with the half step k = 0.5*dt of its move! and accelerate! (:148-159):
//   CD_KICK_MOVE  accelerate! then move!: fluid v += k (a + g); a = 0 for
//                 every particle; fluid x += k v
//   CD_MOVE       move!: a = 0; fluid x += k v
//   CD_KICK       accelerate!: fluid v += k (a + g)
// The continuity find_pressure! between them is SPH_OP_FIND_PRESSURE_CONTINUITY.
__global__ __launch_bounds__(256) void k_cd_unary(int stage, const int* meta, double* cols,
                                                  long long cap, const SysDesc* __restrict__ sysd,
                                                  const int* __restrict__ sysid)
{
    const int n = meta[META_N];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Phys& ph = sysd[sysid[i]].ph;
    const double k = 0.5 * ph.dt;
    const bool fluid = UCOL(cols, cap, C_TYPE, i) == 0.0;
    double v[3];
    for (int d = 0; d < 3; ++d) {
        v[d] = UCOL(cols, cap, C_VX + d, i);
        if (fluid && stage != CD_MOVE) {
            v[d] = v[d] + k * (UCOL(cols, cap, C_AX + d, i) + ph.g[d]);
            UCOL(cols, cap, C_VX + d, i) = v[d];
        }
    }
    if (stage == CD_KICK) return;
    for (int d = 0; d < 3; ++d) {
        UCOL(cols, cap, C_AX + d, i) = 0.0;
        if (fluid) UCOL(cols, cap, C_X + d, i) = UCOL(cols, cap, C_X + d, i) + k * v[d];
    }
}

// ----------------------------------------------------------- reductions --
// Deterministic two-level tree: per-block partial sums in a fixed order,
// then one block sums the partials in a fixed order.

__device__ inline double block_sum(double v)
{
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

static constexpr int RED_ITEMS = 8;

__device__ inline double energy_of(int what, const double* cols, long long cap, const Phys& ph,
                                   int i)
{
    const double v0 = UCOL(cols, cap, C_VX, i), v1 = UCOL(cols, cap, C_VY, i),
                 v2 = UCOL(cols, cap, C_VZ, i);
    double vv = v0 * v0 + v1 * v1;
    vv = vv + v2 * v2;
    const double rho = UCOL(cols, cap, C_RHO, i);
    if (what == SPH_REDUCE_ENERGY) {  // examples/collapse_dry.jl:166-171
        const double x0 = UCOL(cols, cap, C_X, i), x1 = UCOL(cols, cap, C_Y, i),
                     x2 = UCOL(cols, cap, C_Z, i);
        double gx = ph.g[0] * x0 + ph.g[1] * x1;
        gx = gx + ph.g[2] * x2;
        const double kinetic = 0.5 * ph.m * vv;
        const double potential = -ph.m * gx;
        const double internal =
            ph.m * (ph.c * ph.c) * (log(fabs(rho / ph.rho0)) + ph.rho0 / rho - 1.0);
        return kinetic + potential + internal;
    }
    if (what == SPH_REDUCE_ENERGY_RHO0) {
        // kinetic and potential of examples/collapse_dry.jl:166-171; internal
        // energy of P = c^2 (rho - rho0_p)
        const double x0 = UCOL(cols, cap, C_X, i), x1 = UCOL(cols, cap, C_Y, i),
                     x2 = UCOL(cols, cap, C_Z, i);
        double gx = ph.g[0] * x0 + ph.g[1] * x1;
        gx = gx + ph.g[2] * x2;
        const double r0 = UCOL(cols, cap, C_RHO0, i);
        const double kinetic = 0.5 * ph.m * vv;
        const double potential = -ph.m * gx;
        const double internal = ph.m * (ph.c * ph.c) * (log(fabs(rho / r0)) + r0 / rho - 1.0);
        return kinetic + potential + internal;
    }
    // tests/test_collision_2d.jl:96-100
    const double d = rho - UCOL(cols, cap, C_RHO0, i);
    return 0.5 * ph.m * vv + 0.5 * ph.m * (ph.c * ph.c) * (d * d) / (ph.rho0 * ph.rho0);
}

// Sum over the particles [off, off + n) of one system.
__global__ __launch_bounds__(256) void k_reduce_part(int what, int off, int n,
                                                     const double* cols, long long cap,
                                                     const SysDesc* __restrict__ sysd, int sys,
                                                     double* part)
{
    const Phys& ph = sysd[sys].ph;
    const int base = off + blockIdx.x * 256 * RED_ITEMS;
    const int end = off + n;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int it = 0; it < RED_ITEMS; ++it) {
        const int i = base + it * 256 + threadIdx.x;
        if (i >= end) break;
        if (what == SPH_REDUCE_MOMENTUM) {
            a0 += ph.m * UCOL(cols, cap, C_VX, i);
            a1 += ph.m * UCOL(cols, cap, C_VY, i);
            a2 += ph.m * UCOL(cols, cap, C_VZ, i);
        } else {
            a0 += energy_of(what, cols, cap, ph, i);
        }
    }
    const double s0 = block_sum(a0);
    const double s1 = what == SPH_REDUCE_MOMENTUM ? block_sum(a1) : 0.0;
    const double s2 = what == SPH_REDUCE_MOMENTUM ? block_sum(a2) : 0.0;
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s0;
        part[3 * blockIdx.x + 1] = s1;
        part[3 * blockIdx.x + 2] = s2;
    }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* part, int nb, double* out)
{
    double a[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int d = 0; d < 3; ++d) a[d] += part[3 * b + d];
    for (int d = 0; d < 3; ++d) {
        const double s = block_sum(a[d]);
        if (threadIdx.x == 0) out[d] = s;
    }
}

// ----------------------------------------------------------- probe sums --
// SmoothedParticles.sum(sys, func, x) (src/core.jl:240-258) for a batch of
// probe points, with the interpolation closures of
// examples/cavity_flow.jl:166-172.  One wave per probe: the wave visits the
// cells in key_diff order (di outermost, src/structs.jl:70-82) and each cell
// in its stored (descending reference index) order, as one concatenated list
// taken 64 candidates at a time.  The lanes test their candidates in
// parallel and append the accepted ones, in order, to a ring; the terms of
// 64 ring entries at a time are formed in parallel and added in ring order,
// each read with readlane, so every sum is accumulated in the reference's
// order, in FP64 without contraction: the results are bit-exact.

// find_key (src/structs.jl:97-106): -1 when Int64(floor(x/h)) throws, i.e.
// for a NaN or infinite coordinate or one outside the Int64 range.
__device__ inline bool probe_floor(double v, double h, long long* out)
{
    const double f = floor(v / h);
    if (!(f >= -9223372036854775808.0 && f < 9223372036854775808.0)) return false;
    *out = (long long)f;
    return true;
}

__device__ inline long long probe_key(const Geo& g, double x, double y, double z)
{
    long long fi, fj, fk;
    if (!probe_floor(x, g.h, &fi) || !probe_floor(y, g.h, &fj) || !probe_floor(z, g.h, &fk))
        return -1;
    const long long i = 1 + fi - g.kp[0], j = 1 + fj - g.kp[1], k = 1 + fk - g.kp[2];
    return i + (long long)g.L1 * (j - 1) + (long long)g.L1 * g.L2 * (k - 1);
}

// wendland2 / wendland3 (src/kernels.jl:108-115, :156-163) by dimension.
__device__ inline double probe_w(const Geo& g, double r)
{
    const double x = r / g.h;
    if (x > 1.0) return 0.0;
    const double u = 1.0 - x, u2 = u * u, u4 = u2 * u2;
    if (g.dim == 2) return 2.228169203286535 * u4 * (1.0 + 4.0 * x) / (g.h * g.h);
    return 3.3422538049298023 * u4 * (1.0 + 4.0 * x) / (g.h * g.h * g.h);
}

// The double in lane j of a wave (j wave-uniform).
__device__ inline double readlane_f64(double v, int j)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Term policies of the ordered sums.  terms() forms the NA terms of one
// accepted candidate e at squared distance s; begin() reads what the sum
// needs of its own point when the point is a particle.
//
// ProbeTerms: the interpolation closures of examples/cavity_flow.jl:166-172,
// (phi m) W(h, r) and ((phi m) v_k) W(h, r), phi = Float64(type == FLUID)
// with fluid_only, else 1.
struct ProbeTerms {
    static constexpr int NA = 4;
    double m;
    int fluid_only;
    __device__ void begin(const double*, long long, int) {}
    __device__ void terms(const Geo& g, const double* __restrict__ cols, long long cap, int e,
                          double s, double* t) const
    {
        const double r = sqrt(s);
        const double phi = fluid_only ? (UCOL(cols, cap, C_TYPE, e) == 0.0 ? 1.0 : 0.0) : 1.0;
        const double w = phi * m;
        const double k = probe_w(g, r);
        t[0] = w * k;
        t[1] = (w * UCOL(cols, cap, C_VX, e)) * k;
        t[2] = (w * UCOL(cols, cap, C_VY, e)) * k;
        t[3] = (w * UCOL(cols, cap, C_VZ, e)) * k;
    }
};

// LJWallTerms: LJ_potential(p, q, r) of examples/collapse_symplectic.jl:144-151,
// m E_wall (0.5 s^2 - 0.25 s^4 - 0.25) with s = dr_wall / (r + eps) when q is
// a wall particle, p a fluid particle and r < dr_wall, else 0.  s^4 is
// formed as (s^2)^2.
struct LJWallTerms {
    static constexpr int NA = 1;
    double m, dr_wall, E_wall, eps;
    bool p_fluid;
    __device__ void begin(const double* __restrict__ cols, long long cap, int k)
    {
        p_fluid = UCOL(cols, cap, C_TYPE, k) == 0.0;
    }
    __device__ void terms(const Geo&, const double* __restrict__ cols, long long cap, int e,
                          double s, double* t) const
    {
        const double r = sqrt(s);
        t[0] = 0.0;
        if (p_fluid && UCOL(cols, cap, C_TYPE, e) == 1.0 && r < dr_wall) {
            const double q = dr_wall / (r + eps);
            const double q2 = q * q;
            t[0] = (m * E_wall) * ((0.5 * q2 - 0.25 * (q2 * q2)) - 0.25);
        }
    }
};

// The terms of the accepted candidates in ring slots [done, done + cnt) of
// one point, cnt <= 64, added to a[] in slot order.
template <class Pol>
__device__ inline void sum_flush(const Geo& g, const double* __restrict__ cols, long long cap,
                                 const Pol& pol, const int* qe, const double* qs, int done,
                                 int cnt, int lane, double* a)
{
    double t[Pol::NA];
    for (int j = 0; j < Pol::NA; ++j) t[j] = 0.0;
    if (lane < cnt) {
        const int slot = (done + lane) & 127;
        pol.terms(g, cols, cap, qe[slot], qs[slot], t);
    }
    for (int j = 0; j < cnt; ++j)
        for (int v = 0; v < Pol::NA; ++v) a[v] = a[v] + readlane_f64(t[v], j);
}

// SmoothedParticles.sum in the reference's order, one point per block of one
// wave: point i is probes[3i..3i+2] (the point form, src/core.jl:240-258;
// out[NA i + v]) or, with probes == nullptr, the particle at device slot
// off + i (the particle form, :271-291, q == p included; out[NA ref + v] at
// the particle's reference index).  Candidates whose r > h is false are kept
// (a NaN r is kept, as in the reference); the test is evaluated as s > s_thr
// (s_thr = max{s : RN(sqrt(s)) <= h}, and a NaN s passes both).  Accepted
// candidates are appended in order to a 128-slot ring and their terms formed
// 64 at a time.
template <class Pol>
__global__ __launch_bounds__(64) void k_ordered_sum(Geo g, int cell_base,
                                                    const int* __restrict__ cell_start,
                                                    const double* __restrict__ cols, long long cap,
                                                    const double* __restrict__ probes,
                                                    const int* __restrict__ ref, int off, int np,
                                                    Pol pol, double* __restrict__ out)
{
    __shared__ int qe[128];
    __shared__ double qs[128];
    const int i = blockIdx.x;
    const int lane = (int)threadIdx.x;
    if (i >= np) return;  // block-uniform
    double x, y, z;
    if (probes) {
        x = probes[3 * i]; y = probes[3 * i + 1]; z = probes[3 * i + 2];
    } else {
        const int k = off + i;
        x = UCOL(cols, cap, C_X, k); y = UCOL(cols, cap, C_Y, k); z = UCOL(cols, cap, C_Z, k);
        pol.begin(cols, cap, k);
    }
    const long long key = probe_key(g, x, y, z);
    const long long L1 = g.L1, L12 = (long long)g.L1 * g.L2;
    const int kr = g.L3 == 1 ? 0 : 1;
    // Lane t < ncell holds the range of the t-th key_diff cell (di outermost,
    // dk innermost, src/structs.jl:70-82) if it passes the bounds test of
    // src/core.jl:244, else an empty range; a wave prefix sum over the range
    // lengths concatenates them in that order.
    const int ncell = kr ? 27 : 9, nd = kr ? 3 : 1;
    int my_e0 = 0, my_cnt = 0;
    if (lane < ncell) {
        const int di = lane / (3 * nd) - 1, dj = (lane / nd) % 3 - 1, dk = kr ? lane % 3 - 1 : 0;
        const long long nk = key + (di + L1 * dj + L12 * dk);
        if (nk >= 1 && nk <= g.key_max) {
            const int c = cell_base + (int)(nk - 1);
            my_e0 = cell_start[c];
            my_cnt = cell_start[c + 1] - my_e0;
        }
    }
    int incl = my_cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    double a[Pol::NA];
    for (int v = 0; v < Pol::NA; ++v) a[v] = 0.0;
    int fill = 0, done = 0;  // ring positions, uniform
    int t0c = 0;             // first cell holding candidate `base`, uniform
    for (int base = 0; base < total; base += 64) {
        const int idx = base + lane;
        // The cells covering candidates [base, base + 64), walked with
        // uniform indices: t0c ends below ncell because base < total.
        while (__builtin_amdgcn_readlane(incl, t0c) <= base) ++t0c;
        int e = 0;
        for (int u = t0c; u < ncell; ++u) {
            const int hi = __builtin_amdgcn_readlane(incl, u);
            const int lo = hi - __builtin_amdgcn_readlane(my_cnt, u);
            if (lo >= base + 64) break;
            if (idx >= lo && idx < hi) e = __builtin_amdgcn_readlane(my_e0, u) + (idx - lo);
        }
        bool take = false;
        double s = 0.0;
        if (idx < total) {
            const double d1 = x - UCOL(cols, cap, C_X, e);
            const double d2 = y - UCOL(cols, cap, C_Y, e);
            const double d3 = z - UCOL(cols, cap, C_Z, e);
            s = d1 * d1 + d2 * d2;
            s = s + d3 * d3;
            take = !(s > g.s_thr);  // src/core.jl:252
        }
        const unsigned long long acc = __ballot(take);
        if (take) {
            const int slot = (fill + (int)__popcll(acc & ((1ull << lane) - 1))) & 127;
            qe[slot] = e;
            qs[slot] = s;
        }
        fill += (int)__popcll(acc);
        __syncthreads();
        if (fill - done >= 64) {
            sum_flush(g, cols, cap, pol, qe, qs, done, 64, lane, a);
            done += 64;
            __syncthreads();  // the slots read here are rewritten by the next appends
        }
    }
    sum_flush(g, cols, cap, pol, qe, qs, done, fill - done, lane, a);
    if (lane == 0) {
        const long long o = probes ? (long long)i : (long long)ref[off + i];
        for (int v = 0; v < Pol::NA; ++v) out[Pol::NA * o + v] = a[v];
    }
}

// ------------------------------------------------------------ field I/O --
// Host order is the system's reference order: host slot refidx[k] <->
// device slot k, for k in the system's range [off, off + n).
__global__ __launch_bounds__(256) void k_field_out(int off, int n, const int* __restrict__ ref,
                                                   const double* cols, long long cap, int col0,
                                                   int ncomp, double* out)
{
    const int k = off + blockIdx.x * 256 + threadIdx.x;
    if (k >= off + n) return;
    const long long r = ref[k];
    for (int d = 0; d < ncomp; ++d) out[ncomp * r + d] = UCOL(cols, cap, col0 + d, k);
}

// find_key of each particle at the last build, 1-based within its system.
__global__ __launch_bounds__(256) void k_key_out(int off, int n, const int* __restrict__ ref,
                                                 const int* __restrict__ key, int cell_base,
                                                 double* out)
{
    const int k = off + blockIdx.x * 256 + threadIdx.x;
    if (k >= off + n) return;
    out[ref[k]] = (double)(key[k] - cell_base + 1);
}

__global__ __launch_bounds__(256) void k_field_in(int off, int n, const int* __restrict__ ref,
                                                  double* cols, long long cap, int col0, int ncomp,
                                                  const double* in)
{
    const int k = off + blockIdx.x * 256 + threadIdx.x;
    if (k >= off + n) return;
    const long long r = ref[k];
    for (int d = 0; d < ncomp; ++d) UCOL(cols, cap, col0 + d, k) = in[ncomp * r + d];
}

// Append m particles at device slots [n0, n0+m) with reference indices
// n0.. ; `in` holds x[m][3], v[m][3], rho[m], P[m], type[m] back to back.
__global__ __launch_bounds__(256) void k_append(int n0, int m, int ref0, int sys, const double* in,
                                                int has_v, int has_rho, int has_P, int has_type,
                                                double* cols, int* icols, long long cap)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int k = n0 + i;
    const double* x = in;
    const double* v = in + 3ll * m;
    const double* rho = in + 6ll * m;
    const double* P = in + 7ll * m;
    const double* type = in + 8ll * m;
    for (int f = 0; f < NCOL; ++f) UCOL(cols, cap, f, k) = 0.0;
    for (int d = 0; d < 3; ++d) {
        UCOL(cols, cap, C_X + d, k) = x[3ll * i + d];
        if (has_v) UCOL(cols, cap, C_VX + d, k) = v[3ll * i + d];
    }
    if (has_rho) UCOL(cols, cap, C_RHO, k) = rho[i];
    if (has_P) UCOL(cols, cap, C_P, k) = P[i];
    if (has_type) UCOL(cols, cap, C_TYPE, k) = type[i];
    icols[I_REF * cap + k] = ref0 + i;
    icols[I_KEY * cap + k] = 0;
    icols[I_SYS * cap + k] = sys;
}

__global__ void k_set_meta(int* meta, int idx, int v) { meta[idx] = v; }

__global__ void k_set_int(int* p, int v) { *p = v; }

// old slot -> new slot of the system column after retired systems are
// compacted out of the table (map has one entry per old slot).
__global__ __launch_bounds__(256) void k_remap_sys(const int* meta, int* sysid, const int* map)
{
    const int n = meta[META_N];
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) sysid[k] = map[sysid[k]];
}

__global__ __launch_bounds__(256) void k_fill_int(int* p, int v, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ launchers --
static inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

static inline const int* sysid_col(sph_system* s) { return s->icols[s->cur] + I_SYS * s->cap; }

void enqueue_unary(sph_system* s, int op, double* prho2)
{
    if (s->n <= 0) return;
    k_unary<<<nblk(s->n), 256, 0, s->stream>>>(op, s->meta, s->cols[s->cur], s->cap, s->sysd,
                                              sysid_col(s), prho2);
}

void enqueue_prho2(sph_system* s)
{
    if (s->n <= 0) return;
    k_prho2<<<nblk(s->n), 256, 0, s->stream>>>(s->meta, s->cols[s->cur], s->cap, s->prho2);
}

void enqueue_cd_unary(sph_system* s, int stage)
{
    if (s->n <= 0) return;
    k_cd_unary<<<nblk(s->n), 256, 0, s->stream>>>(stage, s->meta, s->cols[s->cur], s->cap, s->sysd,
                                                 sysid_col(s));
}

void enqueue_kick_drift(sph_system* s)
{
    if (s->n <= 0) return;
    k_kick_drift<<<nblk(s->n), 256, 0, s->stream>>>(s->meta, s->cols[s->cur], s->cap, s->sysd,
                                                   sysid_col(s));
}

long long reduce_partials_needed(long long n) { return (n + 256 * RED_ITEMS - 1) / (256 * RED_ITEMS) + 1; }

void enqueue_reduce(sph_system* s, int sys, long long off, long long n, int what, double* out_dev)
{
    const int nb = (int)((n + 256 * RED_ITEMS - 1) / (256 * RED_ITEMS));
    if (nb > 0)
        k_reduce_part<<<nb, 256, 0, s->stream>>>(what, (int)off, (int)n, s->cols[s->cur], s->cap,
                                                 s->sysd, sys, s->red_part);
    k_reduce_final<<<1, 256, 0, s->stream>>>(s->red_part, nb, out_dev);
}

void enqueue_probe_sum(sph_system* s, int sys, const double* probes_dev, long long np, double m,
                       int fluid_only, double* out_dev)
{
    if (np <= 0) return;
    const SysDesc& d = s->sys_host[sys];
    ProbeTerms pol{m, fluid_only};
    k_ordered_sum<ProbeTerms><<<(unsigned)np, 64, 0, s->stream>>>(
        d.g, d.cell_base, s->cell_start, s->cols[s->cur], s->cap, probes_dev, nullptr, 0, (int)np,
        pol, out_dev);
}

// SmoothedParticles.sum(sys, LJ_potential, p) for every particle of system
// `sys` (device slots [off, off + n)); out_dev[refidx] receives the sums.
void enqueue_particle_sum_lj(sph_system* s, int sys, long long off, long long n,
                             const double* fp, double* out_dev)
{
    if (n <= 0) return;
    const SysDesc& d = s->sys_host[sys];
    LJWallTerms pol{fp[0], fp[1], fp[2], fp[3], false};
    k_ordered_sum<LJWallTerms><<<(unsigned)n, 64, 0, s->stream>>>(
        d.g, d.cell_base, s->cell_start, s->cols[s->cur], s->cap, nullptr,
        s->icols[s->cur] + I_REF * s->cap, (int)off, (int)n, pol, out_dev);
}

void enqueue_field_out(sph_system* s, long long off, long long n, int col0, int ncomp,
                       double* out_dev)
{
    if (n <= 0) return;
    k_field_out<<<nblk(n), 256, 0, s->stream>>>((int)off, (int)n, s->icols[s->cur] + I_REF * s->cap,
                                               s->cols[s->cur], s->cap, col0, ncomp, out_dev);
}

void enqueue_key_out(sph_system* s, long long off, long long n, int cell_base, double* out_dev)
{
    if (n <= 0) return;
    k_key_out<<<nblk(n), 256, 0, s->stream>>>((int)off, (int)n, s->icols[s->cur] + I_REF * s->cap,
                                             s->icols[s->cur] + I_KEY * s->cap, cell_base, out_dev);
}

void enqueue_field_in(sph_system* s, long long off, long long n, int col0, int ncomp,
                      const double* in_dev)
{
    if (n <= 0) return;
    k_field_in<<<nblk(n), 256, 0, s->stream>>>((int)off, (int)n, s->icols[s->cur] + I_REF * s->cap,
                                              s->cols[s->cur], s->cap, col0, ncomp, in_dev);
}

void enqueue_append(sph_system* s, int n0, int m, int ref0, int sys, const double* in_dev,
                    int has_v, int has_rho, int has_P, int has_type)
{
    if (m <= 0) return;
    k_append<<<nblk(m), 256, 0, s->stream>>>(n0, m, ref0, sys, in_dev, has_v, has_rho, has_P,
                                            has_type, s->cols[s->cur], s->icols[s->cur], s->cap);
    k_set_meta<<<1, 1, 0, s->stream>>>(s->meta, META_N, n0 + m);
}

void enqueue_set_int(sph_system* s, int* p, int v) { k_set_int<<<1, 1, 0, s->stream>>>(p, v); }

void enqueue_set_meta(sph_system* s, int idx, int v) { k_set_meta<<<1, 1, 0, s->stream>>>(s->meta, idx, v); }

void enqueue_fill_int(sph_system* s, int* p, int v, long long n)
{
    if (n > 0) k_fill_int<<<nblk(n), 256, 0, s->stream>>>(p, v, n);
}

void enqueue_remap_sys(sph_system* s, const int* map_dev)
{
    if (s->n <= 0) return;
    k_remap_sys<<<nblk(s->n), 256, 0, s->stream>>>(s->meta, s->icols[s->cur] + I_SYS * s->cap,
                                                  map_dev);
}

}  // namespace sph
