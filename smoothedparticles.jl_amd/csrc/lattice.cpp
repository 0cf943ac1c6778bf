// Host-side initial-state generation: square (2D) and cubic (3D) lattice
// coverings of a box and of a wall layer around a tank, with the FP64
// expressions of the reference so that particle positions and their order
// match what generate_particles! produces there.
//
//   covering(Squaregrid/CubicGrid, s)   src/grids.jl:52-66, 128-144
//   BoundaryLayer(s, grid, width)       src/geometry.jl:198-234
//   Specification(s, f) (lid removal)   src/geometry.jl:178-189
//   Ball / Box inside tests             src/geometry.jl:24-30, 254-256
//
// This is set-up code: it runs once on the host, never on the hot path.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "sph_setup.h"

namespace {

struct Box3 {
    double lo[3], hi[3];
};

// src/geometry.jl:24-30 -- inclusive box test.
inline bool inside_box(const Box3& b, const double* x)
{
    return b.lo[0] <= x[0] && x[0] <= b.hi[0] && b.lo[1] <= x[1] && x[1] <= b.hi[1] &&
           b.lo[2] <= x[2] && x[2] <= b.hi[2];
}

// Lattice index range of a covering: floor(min/dr) .. ceil(max/dr) per axis
// (src/grids.jl:55-58, 131-136).  2D lattices have a single k = 0 layer.
struct Range3 {
    int64_t lo[3], hi[3];
};

Range3 covering_range(int dim, double dr, const Box3& bb)
{
    Range3 r{};
    for (int d = 0; d < 3; ++d) {
        if (d == 2 && dim == 2) {
            r.lo[d] = r.hi[d] = 0;
            continue;
        }
        r.lo[d] = (int64_t)std::floor(bb.lo[d] / dr);
        r.hi[d] = (int64_t)std::ceil(bb.hi[d] / dr);
    }
    return r;
}

inline void lattice_point(int dim, double dr, int64_t i, int64_t j, int64_t k, double* x)
{
    x[0] = (double)i * dr;
    x[1] = (double)j * dr;
    x[2] = dim == 2 ? 0.0 : (double)k * dr;
}

// covering(grid, Ball(0,0,0,w)) -- the offsets of a BoundaryLayer
// (src/geometry.jl:203-205, 254-256).
std::vector<double> ball_offsets(int dim, double dr, double w)
{
    Box3 bb{{-w, -w, -w}, {w, w, w}};
    Range3 r = covering_range(dim, dr, bb);
    std::vector<double> out;
    for (int64_t i = r.lo[0]; i <= r.hi[0]; ++i)
        for (int64_t j = r.lo[1]; j <= r.hi[1]; ++j)
            for (int64_t k = r.lo[2]; k <= r.hi[2]; ++k) {
                double x[3];
                lattice_point(dim, dr, i, j, k, x);
                double a = (x[0] - 0.0) * (x[0] - 0.0);
                double b = (x[1] - 0.0) * (x[1] - 0.0);
                double c = (x[2] - 0.0) * (x[2] - 0.0);
                double s = a + b;
                s = s + c;
                if (s <= w * w) out.insert(out.end(), x, x + 3);
            }
    return out;
}

}  // namespace

namespace sph {
// For the device generator (generate.hip): the lattice index range of the
// covering of a box, and the offsets of covering(grid, Ball(0, 0, 0, w)).
void lattice_range(int dim, double dr, const double* lo, const double* hi, int64_t* rlo,
                   int64_t* rhi)
{
    Box3 b{{lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
    const Range3 r = covering_range(dim, dr, b);
    for (int d = 0; d < 3; ++d) {
        rlo[d] = r.lo[d];
        rhi[d] = r.hi[d];
    }
}
std::vector<double> lattice_ball_offsets(int dim, double dr, double w)
{
    return ball_offsets(dim, dr, w);
}
}  // namespace sph

extern "C" int64_t sph_lattice_box(int dim, double dr, const double* bmin, const double* bmax,
                                   double* out, int64_t cap)
{
    if ((dim != 2 && dim != 3) || !(dr > 0.0)) return -1;
    Box3 b{{bmin[0], bmin[1], bmin[2]}, {bmax[0], bmax[1], bmax[2]}};
    Range3 r = covering_range(dim, dr, b);
    int64_t n = 0;
    for (int64_t i = r.lo[0]; i <= r.hi[0]; ++i)
        for (int64_t j = r.lo[1]; j <= r.hi[1]; ++j)
            for (int64_t k = r.lo[2]; k <= r.hi[2]; ++k) {
                double x[3];
                lattice_point(dim, dr, i, j, k, x);
                if (!inside_box(b, x)) continue;
                if (n < cap) std::memcpy(out + 3 * n, x, sizeof(x));
                ++n;
            }
    return n;
}

extern "C" int64_t sph_lattice_wall_layer(int dim, double dr, const double* tmin,
                                          const double* tmax, double width, int lid_axis,
                                          double lid, double* out, int64_t cap)
{
    if ((dim != 2 && dim != 3) || !(dr > 0.0) || !(width > 0.0)) return -1;
    Box3 tank{{tmin[0], tmin[1], tmin[2]}, {tmax[0], tmax[1], tmax[2]}};
    // boundarybox(BoundaryLayer) = tank grown by width; a 2D layer keeps
    // the tank's x3 extent (src/geometry.jl:221-234).
    Box3 bb = tank;
    for (int d = 0; d < 3; ++d) {
        if (d == 2 && dim == 2) continue;
        bb.lo[d] = tank.lo[d] - width;
        bb.hi[d] = tank.hi[d] + width;
    }
    std::vector<double> dxs = ball_offsets(dim, dr, width);
    size_t nd = dxs.size() / 3;
    Range3 r = covering_range(dim, dr, bb);
    int64_t n = 0;
    for (int64_t i = r.lo[0]; i <= r.hi[0]; ++i)
        for (int64_t j = r.lo[1]; j <= r.hi[1]; ++j)
            for (int64_t k = r.lo[2]; k <= r.hi[2]; ++k) {
                double x[3];
                lattice_point(dim, dr, i, j, k, x);
                // Specification: f(x) first, then the layer test.
                if (lid_axis >= 0 && !(x[lid_axis] < lid)) continue;
                if (inside_box(tank, x)) continue;
                bool hit = false;
                for (size_t t = 0; t < nd && !hit; ++t) {
                    double y[3] = {x[0] + dxs[3 * t], x[1] + dxs[3 * t + 1],
                                   x[2] + dxs[3 * t + 2]};
                    hit = inside_box(tank, y);
                }
                if (!hit) continue;
                if (n < cap) std::memcpy(out + 3 * n, x, sizeof(x));
                ++n;
            }
    return n;
}

// splitmix64: a small, well-known 64-bit generator (Steele, Lea, Flood 2014).
static inline uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

extern "C" void sph_jitter(double* x, int64_t n, int dim, double amplitude, uint64_t seed)
{
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < dim; ++d) {
            double u = (double)(splitmix64(s) >> 11) * 0x1.0p-53;  // [0,1)
            x[3 * i + d] += (2.0 * u - 1.0) * amplitude;
        }
}
