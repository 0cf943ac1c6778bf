/*
 * sph_setup.h -- initial-state generation (set-up code, not on the hot
 * path), on the host and on the device.  Square (2D) / cubic (3D) lattices with the FP64 expressions of
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

#include "sph.h"
#include "sph_batch.h"

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

/* Device-side generate_particles!(sys, grid, shape, constructor)
 * (src/grids.jl:253-258): the lattice points of a covering are found and
 * written into the device particle arrays without a host copy, in the order
 * and with the FP64 positions of sph_lattice_box / sph_lattice_wall_layer
 * (bit for bit).  The constructor is a fixed form: type, zero velocity, and
 * either constant rho and P or, with hydro_axis >= 0, the hydrostatic state
 * of examples/collapse_dry.jl:97-100:
 *   P = (rho0 * g) * (x[hydro_axis] - H),  rho = rho0 + P / (c * c).
 * jitter != 0 adds the seeded uniform jitter of sph_jitter (same draws, the
 * first dim coordinates of the new particles in their order). */
#define SPH_LATTICE_BOX 0         /* points inside the inclusive box [lo, hi] */
#define SPH_LATTICE_WALL_LAYER 1  /* Specification(BoundaryLayer(tank [lo, hi], grid, width),
                                     x -> x[lid_axis] < lid); lid_axis < 0 keeps the lid */
typedef struct sph_lattice_spec {
    int dim;            /* 2 (square lattice) or 3 (cubic) */
    int kind;           /* SPH_LATTICE_BOX or SPH_LATTICE_WALL_LAYER */
    double dr;          /* lattice spacing */
    double lo[3], hi[3];
    double width;       /* wall layer width */
    int lid_axis;       /* wall layer: Specification axis (< 0: none) */
    int hydro_axis;     /* < 0: constant rho and P below */
    double lid;
    double type, rho, P;
    double rho0, g, H, c;  /* hydrostatic state */
    double jitter;      /* amplitude (0: none) */
    uint64_t seed;
} sph_lattice_spec;

/* Appends to the system (sph.h handle) or to system `id` of a batch (the last
 * system holding particles); *n_added (may be NULL) receives the count. */
int sph_generate_lattice(sph_system* s, const sph_lattice_spec* spec, int64_t* n_added);
int sph_batch_generate_lattice(sph_batch* b, int id, const sph_lattice_spec* spec,
                               int64_t* n_added);

#ifdef __cplusplus
}
#endif

#endif /* SPH_SETUP_H */
