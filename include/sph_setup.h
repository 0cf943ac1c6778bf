/*
 * sph_setup.h -- host-side initial-state generation (set-up code, not on the
 * hot path).  Square (2D) / cubic (3D) lattices with the FP64 expressions of
 * the reference, so that positions and their order equal what
 * generate_particles! produces:
 *
 *   covering(Squaregrid, s) / covering(CubicGrid, s)  src/grids.jl:52-66, 128-144
 *   BoundaryLayer(s, grid, width)                     src/geometry.jl:198-234
 *   Specification(s, x -> x[axis] < lid)              src/geometry.jl:178-189
 */
#ifndef SPH_SETUP_H
#define SPH_SETUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lattice points i*dr (j*dr, k*dr) inside the inclusive box [bmin, bmax],
 * i outermost.  Writes at most cap points ([n][3]) and returns the total
 * count (negative on invalid arguments). */
int64_t sph_lattice_box(int dim, double dr, const double* bmin, const double* bmax,
                        double* out, int64_t cap);

/* Lattice points of Specification(BoundaryLayer(tank, grid, width),
 * x -> x[lid_axis] < lid): outside the tank and within `width` of it along a
 * lattice offset.  lid_axis < 0 keeps the lid.  Same return convention. */
int64_t sph_lattice_wall_layer(int dim, double dr, const double* tank_min,
                               const double* tank_max, double width, int lid_axis,
                               double lid, double* out, int64_t cap);

/* Adds a seeded uniform jitter in [-amplitude, amplitude) to the first dim
 * coordinates of n points (splitmix64 stream, one draw per coordinate). */
void sph_jitter(double* x, int64_t n, int dim, double amplitude, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* SPH_SETUP_H */
