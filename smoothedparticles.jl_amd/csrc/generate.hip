// Device-side initial-state generation: generate_particles!(sys, grid,
// shape, constructor) (src/grids.jl:253-258) for square (2D) and cubic (3D)
// lattices, written straight into the device particle arrays.  The lattice
// points, their order (i outermost, k innermost) and the inside tests are
// the FP64 expressions of the host generator (lattice.cpp), which restates
//
//   covering(Squaregrid/CubicGrid, s)   src/grids.jl:52-66, 128-144
//   BoundaryLayer(s, grid, width)       src/geometry.jl:198-234
//   Specification(s, f) (lid removal)   src/geometry.jl:178-189
//   Box inside test                     src/geometry.jl:24-30
//
// so a device-generated state has the bits of the host-generated one.  The
// "constructor" is a fixed form: type, velocity 0, and either constant rho
// and P or the hydrostatic pressure of examples/collapse_dry.jl:97-100.
// An optional seeded jitter moves the first dim coordinates by a uniform
// amount; draw t of the splitmix64 stream (seed s) is mix(s + (t + 1) gamma),
// so each particle draws its own values in parallel and the result equals
// the host's sequential sph_jitter.
//
// One launch flags every lattice index of the covering range, a scan gives
// the kept points their order, one host sync reads the count (the arrays
// may have to grow), and one launch writes the particles.
#include <vector>

#include "scan_kernels.h"
#include "sph_internal.h"
#include "sph_setup.h"

namespace sph {

void lattice_range(int dim, double dr, const double* lo, const double* hi, int64_t* rlo,
                   int64_t* rhi);
std::vector<double> lattice_ball_offsets(int dim, double dr, double w);



__device__ inline void lat_point(const LatticeGeo& g, long long t, double* x)
{
    const long long i = g.r0[0] + t / (g.ext[1] * g.ext[2]);
    const long long j = g.r0[1] + (t / g.ext[2]) % g.ext[1];
    const long long k = g.r0[2] + t % g.ext[2];
    x[0] = (double)i * g.dr;
    x[1] = (double)j * g.dr;
    x[2] = g.dim == 2 ? 0.0 : (double)k * g.dr;
}

// src/geometry.jl:24-30
__device__ inline bool lat_inside(const double* lo, const double* hi, const double* x)
{
    return lo[0] <= x[0] && x[0] <= hi[0] && lo[1] <= x[1] && x[1] <= hi[1] && lo[2] <= x[2] &&
           x[2] <= hi[2];
}

__device__ inline bool lat_keep(const LatticeGeo& g, const double* __restrict__ ball, long long t)
{
    double x[3];
    lat_point(g, t, x);
    if (g.kind == SPH_LATTICE_BOX) return lat_inside(g.lo, g.hi, x);
    // Specification first, then the layer test: outside the tank and a
    // lattice offset of the ball lands inside it
    if (g.lid_axis >= 0 && !(x[g.lid_axis] < g.lid)) return false;
    if (lat_inside(g.lo, g.hi, x)) return false;
    for (int b = 0; b < g.nball; ++b) {
        const double y[3] = {x[0] + ball[3 * b], x[1] + ball[3 * b + 1], x[2] + ball[3 * b + 2]};
        if (lat_inside(g.lo, g.hi, y)) return true;
    }
    return false;
}

__global__ __launch_bounds__(256) void k_lat_flags(LatticeGeo g, const double* __restrict__ ball,
                                                   long long n, unsigned* flags)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) flags[t] = lat_keep(g, ball, t) ? 1u : 0u;
}

struct SrcFlags {
    using V = unsigned;
    const unsigned* f;
    __device__ V operator()(int i) const { return f[i]; }
};

__device__ inline unsigned long long lat_mix(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct LatticeInit {
    double type, rho, P, jitter;
    unsigned long long seed;
    int hydro_axis;
    double rho0, g, H, c;
};

__global__ __launch_bounds__(256) void k_lat_write(LatticeGeo g, LatticeInit in,
                                                   const unsigned* __restrict__ flags,
                                                   const unsigned* __restrict__ pos, long long n,
                                                   int n0, int ref0, int sys, double* cols,
                                                   int* icols, long long cap)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n || !flags[t]) return;
    const long long m = pos[t];  // order among the kept points
    const long long k = n0 + m;
    double x[3];
    lat_point(g, t, x);
    if (in.jitter != 0.0) {
        for (int d = 0; d < g.dim; ++d) {
            const unsigned long long draw = (unsigned long long)(m * g.dim + d) + 1ull;
            const unsigned long long z = lat_mix(in.seed + draw * 0x9e3779b97f4a7c15ULL);
            const double u = (double)(z >> 11) * 0x1.0p-53;  // [0, 1)
            x[d] += (2.0 * u - 1.0) * in.jitter;
        }
    }
    for (int f = 0; f < NCOL; ++f) cols[(long long)f * cap + k] = 0.0;
    for (int d = 0; d < 3; ++d) cols[(long long)(C_X + d) * cap + k] = x[d];
    double rho = in.rho, P = in.P;
    if (in.hydro_axis >= 0) {  // examples/collapse_dry.jl:97-100
        P = (in.rho0 * in.g) * (x[in.hydro_axis] - in.H);
        rho = in.rho0 + P / (in.c * in.c);
    }
    cols[(long long)C_RHO * cap + k] = rho;
    cols[(long long)C_P * cap + k] = P;
    cols[(long long)C_TYPE * cap + k] = in.type;
    icols[I_REF * cap + k] = ref0 + (int)m;
    icols[I_KEY * cap + k] = 0;
    icols[I_SYS * cap + k] = sys;
}

// Plan of a lattice: geometry, covering range, ball offsets.  Returns the
// number of lattice indices to flag, 0 for an empty covering, -1 for an
// invalid specification and -2 for a covering above 2^31 indices.
long long lattice_plan(const sph_lattice_spec* spec, LatticeGeo* g, std::vector<double>* ball)
{
    if (!spec || (spec->dim != 2 && spec->dim != 3) || !(spec->dr > 0.0) ||
        (spec->kind != SPH_LATTICE_BOX && spec->kind != SPH_LATTICE_WALL_LAYER) ||
        (spec->kind == SPH_LATTICE_WALL_LAYER && !(spec->width > 0.0)) ||
        spec->lid_axis > 2 || spec->hydro_axis > 2)
        return -1;
    *g = LatticeGeo{};
    g->dim = spec->dim;
    g->kind = spec->kind;
    g->dr = spec->dr;
    g->lid_axis = spec->kind == SPH_LATTICE_WALL_LAYER ? spec->lid_axis : -1;
    g->lid = spec->lid;
    double blo[3], bhi[3];  // the covering's box: the shape's bounding box
    for (int d = 0; d < 3; ++d) {
        g->lo[d] = spec->lo[d];
        g->hi[d] = spec->hi[d];
        blo[d] = spec->lo[d];
        bhi[d] = spec->hi[d];
        // boundarybox(BoundaryLayer) = tank grown by width; a 2D layer keeps
        // the tank's x3 extent (src/geometry.jl:221-234)
        if (spec->kind == SPH_LATTICE_WALL_LAYER && !(d == 2 && spec->dim == 2)) {
            blo[d] = spec->lo[d] - spec->width;
            bhi[d] = spec->hi[d] + spec->width;
        }
    }
    int64_t rlo[3], rhi[3];
    lattice_range(g->dim, g->dr, blo, bhi, rlo, rhi);
    long long nlat = 1;
    for (int d = 0; d < 3; ++d) {
        g->r0[d] = rlo[d];
        g->ext[d] = rhi[d] - rlo[d] + 1;
        if (g->ext[d] <= 0) return 0;
        if (nlat > ((1ll << 31) - 1) / g->ext[d]) return -2;
        nlat *= g->ext[d];
    }
    ball->clear();
    if (spec->kind == SPH_LATTICE_WALL_LAYER)
        *ball = lattice_ball_offsets(g->dim, g->dr, spec->width);
    g->nball = (int)(ball->size() / 3);
    return nlat;
}

// Flag the kept lattice indices and scan them: pos[t] is the order of kept
// index t, pos[nlat] the count.  part needs scan_partials_needed(nlat).
void enqueue_lattice_flags(sph_system* s, const LatticeGeo& g, const double* ball, long long nlat,
                           unsigned* flags, unsigned* pos, unsigned* part)
{
    const unsigned nb = (unsigned)((nlat + 255) / 256);
    k_lat_flags<<<nb, 256, 0, s->stream>>>(g, ball, nlat, flags);
    scan(s, SrcFlags{flags}, (int)nlat, pos, part);
}

void enqueue_lattice_write(sph_system* s, const LatticeGeo& g, const sph_lattice_spec* spec,
                           const unsigned* flags, const unsigned* pos, long long nlat, int slot,
                           int ref0)
{
    const unsigned nb = (unsigned)((nlat + 255) / 256);
    LatticeInit in{spec->type, spec->rho, spec->P, spec->jitter, spec->seed, spec->hydro_axis,
                   spec->rho0, spec->g, spec->H, spec->c};
    k_lat_write<<<nb, 256, 0, s->stream>>>(g, in, flags, pos, nlat, (int)s->n, ref0, slot,
                                           s->cols[s->cur], s->icols[s->cur], s->cap);
}

}  // namespace sph
