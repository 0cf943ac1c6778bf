/*
 * sph_batch.h -- many independent particle systems in one set of device
 * arrays, stepped together by one launch per stage.
 *
 * No counterpart in SmoothedParticles.jl, which steps one ParticleSystem
 * (src/structs.jl:57-92) at a time.  Each system of a batch behaves
 * exactly as that system alone would behave in a handle of sph.h: its own
 * cell geometry, parameters, reference particle order and removal of
 * particles that leave its bounding box.  Systems may also be periodic
 * (also not in the reference).  Systems join and leave a running batch
 * between calls (continuous batching); a retired system's device rows are
 * reclaimed at the next cell build.
 *
 * Conventions are those of sph.h: 0 or a negative SPH_E* code, messages in
 * sph_last_error(), synchronous calls, one host thread per handle.  A batch
 * handle is an sph_system handle: every system has the same dimension (2D
 * or 3D) and precision mode.
 */
#ifndef SPH_BATCH_H
#define SPH_BATCH_H

#include "sph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sph_system sph_batch;

typedef struct sph_system_desc {
    double h;            /* neighbour radius */
    double bbox_min[3];  /* bounding box (the domain's, src/structs.jl:63) */
    double bbox_max[3];
    int periodic;        /* 0: reference box semantics; 1: periodic in x, y (and z when
                            bbox_min[2] != bbox_max[2]) with minimum-image pairs; needs
                            box >= 3 h per periodic axis */
    int pad;
    sph_params params;   /* the system's parameters for sph_batch_step */
} sph_system_desc;

int sph_batch_create(sph_batch** out, int precision);
int sph_batch_destroy(sph_batch* b);

/* Adds a system with n particles (x [n][3]; v, rho, P, type may be NULL)
 * and returns its id, stable for the life of the batch. */
int sph_batch_add(sph_batch* b, const sph_system_desc* desc, int64_t n, const double* x,
                  const double* v, const double* rho, const double* P, const double* type,
                  int* id);
/* Removes a system from the batch; its id is no longer valid. */
int sph_batch_retire(sph_batch* b, int id);

/* Live systems and their particles in total. */
int sph_batch_size(sph_batch* b, int* nsys, int64_t* particles);
int sph_batch_count(sph_batch* b, int id, int64_t* n);
/* Field transfer of one system in its reference order (as sph_get_field). */
int sph_batch_get_field(sph_batch* b, int id, int field, double* host, int64_t n);
int sph_batch_set_field(sph_batch* b, int id, int field, const double* host, int64_t n);
/* Reduction over one system with its own parameters (as sph_reduce). */
int sph_batch_reduce(sph_batch* b, int id, int what, double* out);

/* Cell lists of every system (create_cell_list!, src/core.jl:51-90). */
int sph_batch_build_cells(sph_batch* b);
/* Neighbour pairs of one system under the current cell lists, as
 * sph_neighbour_pairs. */
int sph_batch_neighbour_pairs(sph_batch* b, int id, int64_t* pairs, int64_t cap, int64_t* n);
/* Asynchronous frame download of one system, as sph_snapshot_begin. */
int sph_batch_snapshot_begin(sph_batch* b, int id, const int* fields, int nfields,
                             sph_snapshot** out);
/* Neighbour-set digest of one system under the current cell lists, as
 * sph_neighbour_digest (n must equal the system's particle count). */
int sph_batch_neighbour_digest(sph_batch* b, int id, int64_t* counts, uint64_t* hashes,
                               int64_t n);
/* Probe sums of one system under the current cell lists, as sph_probe_sum
 * with the mass of the system's parameters (SmoothedParticles.sum at points,
 * src/core.jl:240-258).  Periodic systems: SPH_EUNSUP. */
int sph_batch_probe_sum(sph_batch* b, int id, int flags, const double* x, int64_t np,
                        double* out);

/* find_rho0! at t = 0 (tests/test_collision_2d.jl:67-69,123-126) for the
 * systems added since the last call: rho0_p = sum_q m W(h, r_pq) + m W(h, 0).
 * Running systems keep theirs. */
int sph_batch_init_rho0(sph_batch* b);

/* nsteps fused steps of a scheme of sph.h for every system, each with its
 * own parameters.  *pairs as sph_step, summed over all systems. */
int sph_batch_step(sph_batch* b, int scheme, int nsteps, int64_t* pairs);

/* Tile size and its autotuning as sph_set_tile / sph_autotune. */
int sph_batch_set_tile(sph_batch* b, int targets_per_tile);
int sph_batch_autotune(sph_batch* b, int* best);

/* Device memory held by the handle, in bytes (admission control). */
int sph_batch_footprint(sph_batch* b, int64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* SPH_BATCH_H */
