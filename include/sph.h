/*
 * sph.h -- C ABI of the MI355X (gfx950) cell-list neighbour search and
 * pair-interaction library.  This is the drop-in boundary for the hot path
 * of SmoothedParticles.jl v0.2.0:
 *
 *   ParticleSystem(T, domain, h)      src/structs.jl:57-92   -> sph_create
 *   push!(sys.particles, ...)         src/grids.jl:253-258   -> sph_append
 *   ParticleField(sys, :sym)          src/structs.jl:118-125 -> sph_get_field / sph_set_field
 *   create_cell_list!(sys)            src/core.jl:51-90      -> sph_build_cells
 *   apply_binary!(sys, op!)           src/core.jl:125-129    -> sph_apply_binary
 *   apply_unary!(sys, op!)            src/core.jl:138-142    -> sph_apply_unary
 *   apply!(sys, op!; self)            src/core.jl:151-161    -> sph_apply
 *   sum(p -> energy(p), particles)    tests/test_collision_2d.jl:137,
 *                                     examples/collapse_dry.jl:173-187 -> sph_reduce
 *   SmoothedParticles.sum(sys, f, x)  src/core.jl:240-258    -> sph_probe_sum
 *
 * Conventions
 *   - Every function returns 0 on success and a negative SPH_E* code on
 *     failure; sph_last_error() then holds a message (per host thread).
 *   - Host pointers are borrowed for the duration of the call only.  Device
 *     memory is owned by the system handle.
 *   - Calls are synchronous: they return after the device work finished,
 *     as the reference's threaded loops do.
 *   - A handle is not thread-safe; one host thread drives it at a time.
 *   - Particle order seen through get/set/append is the reference's order of
 *     sys.particles, including its swap-with-end removal of particles that
 *     left the domain (src/core.jl:63-81).  The device keeps its own
 *     cell-sorted order internally.
 *   - Host layout of vector fields is [n][3] doubles (RealVector).
 */
#ifndef SPH_H
#define SPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sph_system sph_system;

/* error codes */
#define SPH_OK 0
#define SPH_EINVAL -1   /* invalid argument */
#define SPH_EHIP -2     /* HIP runtime error */
#define SPH_ESTATE -3   /* call not valid in the current state */
#define SPH_ENOMEM -4   /* device allocation failed */
#define SPH_EUNSUP -5   /* unsupported operator or configuration */

/* precision modes */
#define SPH_FP64 0   /* plain FP64 everywhere (the drop-in) */
#define SPH_MIXED 1  /* FP32 relative-coordinate pair arithmetic with FP64
                        accumulation; FP64 positions, state and integrator */

/* fields (sph_get_field / sph_set_field) */
#define SPH_FIELD_X 0     /* position  [n][3] */
#define SPH_FIELD_V 1     /* velocity  [n][3] */
#define SPH_FIELD_A 2     /* acceleration (Dv) [n][3] */
#define SPH_FIELD_RHO 3   /* density   [n] */
#define SPH_FIELD_P 4     /* pressure  [n] */
#define SPH_FIELD_DRHO 5  /* density rate [n] */
#define SPH_FIELD_RHO0 6  /* per-particle reference density [n] */
#define SPH_FIELD_TYPE 7  /* particle type, FLUID = 0.0 [n] */
#define SPH_FIELD_KEY 8   /* read only: find_key of the particle at the last cell-list
                             build (1-based, src/structs.jl:97-106), as a double [n];
                             periodic systems: their own cell key */

/* Binary operators, registered by the name of the reference operator they
 * replace.  Each runs over the neighbour set of src/core.jl:94-112. */
#define SPH_OP_DENSITY_SUM 1         /* find_rho!  tests/test_collision_2d.jl:63-65:
                                        rho += m*W(h,r) */
#define SPH_OP_DENSITY_SUM_RHO0 2    /* find_rho0! tests/test_collision_2d.jl:67-69 */
#define SPH_OP_INTERNAL_FORCE 3      /* internal_force! examples/collapse_dry.jl:135-141:
                                        fluid targets, pressure gradient with
                                        P/rho^2, laminar mu term, plus Monaghan
                                        artificial viscosity alpha (not in the
                                        reference; alpha = 0 disables it) */
#define SPH_OP_PRESSURE_FORCE_RHO0 4 /* internal_force! tests/test_collision_2d.jl:75-78 */
#define SPH_OP_BALANCE_OF_MASS 5     /* balance_of_mass! examples/collapse_dry.jl:112-115 */

/* Unary operators */
#define SPH_OP_RESET_RHO 16          /* reset_rho!  tests/test_collision_2d.jl:84-86 */
#define SPH_OP_RESET_A 17            /* reset_a!    tests/test_collision_2d.jl:80-82 */
#define SPH_OP_FIND_PRESSURE 18      /* P = c^2 (rho - rho0)  examples/collapse_dry.jl:126 */
#define SPH_OP_FIND_PRESSURE_RHO0 19 /* P = c^2 (rho - rho0_p) tests/test_collision_2d.jl:71-73 */
#define SPH_OP_FIND_PRESSURE_CONTINUITY 20 /* find_pressure! examples/collapse_dry.jl:123-127 */
#define SPH_OP_ACCELERATE 21         /* accelerate! fluid: v += (coeff*dt)(a + g)
                                        examples/collapse_dry.jl:155-159 */
#define SPH_OP_MOVE 22               /* move! examples/collapse_dry.jl:148-153: a = 0 for
                                        every particle, then fluid: x += (coeff*dt) v */
#define SPH_OP_DRIFT 23              /* move! tests/test_collision_2d.jl:88-90: every
                                        particle x += (coeff*dt) v; a is kept */
#define SPH_OP_KICK 24               /* accelerate! tests/test_collision_2d.jl:92-94:
                                        every particle v += (coeff*dt) a (no gravity) */

/* Reductions (sph_reduce) */
#define SPH_REDUCE_ENERGY 0           /* sum of energy(p), examples/collapse_dry.jl:166-171 (1 value) */
#define SPH_REDUCE_ENERGY_COLLISION 1 /* tests/test_collision_2d.jl:96-100 (1 value) */
#define SPH_REDUCE_MOMENTUM 2         /* sum of m v (3 values) */
#define SPH_REDUCE_ENERGY_RHO0 3      /* energy with the per-particle reference density:
                                         internal m c^2 (ln(rho/rho0_p) + rho0_p/rho - 1),
                                         the energy of P = c^2 (rho - rho0_p) (1 value) */

/* Fused step schemes (sph_step) */
#define SPH_SCHEME_WCSPH_SE 0       /* build cells; rho = sum m W (self); P = c^2 (rho - rho0);
                                       a = internal force; v += dt (a + g); x += dt v (fluid) */
#define SPH_SCHEME_WCSPH_SE_RHO0 1  /* the same with P = c^2 (rho - rho0_p), rho0_p the
                                       per-particle field set by find_rho0! at t = 0
                                       (tests/test_collision_2d.jl:67-73,123-126) */
#define SPH_SCHEME_COLLAPSE_DRY 2   /* the time loop of examples/collapse_dry.jl:202-211:
                                       accelerate!, move!, create_cell_list!,
                                       balance_of_mass!, find_pressure! (continuity),
                                       move!, create_cell_list!, internal_force!,
                                       accelerate!, with the half steps 0.5*dt of its
                                       move!/accelerate! (:148-159); params->coeff is
                                       not used.  The caller builds the cells and runs
                                       internal_force! once before the first step
                                       (:198-200).  The pair counts of sph_step_pairs
                                       are those of balance_of_mass! and internal_force! */

typedef struct sph_params {
    double m;      /* particle mass */
    double rho0;   /* reference density */
    double c;      /* speed of sound */
    double alpha;  /* artificial viscosity coefficient (0 = off) */
    double hs;     /* smoothing length of the viscosity term */
    double mu;     /* laminar viscosity (0 = off) */
    double nu;     /* density diffusion of balance_of_mass! */
    double g[3];   /* gravity */
    double dt;     /* time step */
    double coeff;  /* multiplier of dt in MOVE / ACCELERATE */
} sph_params;

/* ParticleSystem(T, domain, h), src/structs.jl:57-92.  The domain is given
 * by its bounding box, which is all the reference keeps of it (:63).
 * h > 0 is the neighbour radius and cell size.  The system is 2D when
 * floor(max_z/h) == floor(min_z/h), as in the reference (:70). */
int sph_create(sph_system** out, double h, const double bbox_min[3],
               const double bbox_max[3], int precision);
int sph_destroy(sph_system* s);

/* push! of n particles at the end of sys.particles (generate_particles!,
 * src/grids.jl:253-258).  x is [n][3]; v may be NULL (zero); rho, P, type
 * may be NULL (zero).  Other fields start at zero. */
int sph_append(sph_system* s, int64_t n, const double* x, const double* v,
               const double* rho, const double* P, const double* type);

/* length(sys.particles) */
int sph_count(sph_system* s, int64_t* n);

/* ParticleField get/set in reference order; n must equal sph_count. */
int sph_set_field(sph_system* s, int field, const double* host, int64_t n);
int sph_get_field(sph_system* s, int field, double* host, int64_t n);

/* create_cell_list!(sys), src/core.jl:51-90: removes particles outside the
 * bounding box (swap-with-end, as the reference) and builds the cell list.
 * *n_out (may be NULL) receives the new particle count. */
int sph_build_cells(sph_system* s, int64_t* n_out);

/* apply_binary!(sys, op) and, if self != 0, op(p, p, 0.0) afterwards
 * (src/core.jl:151-157).  Requires a cell list built since the last change
 * of positions or particle count. */
int sph_apply_binary(sph_system* s, int op, const sph_params* params, int self);
/* apply_unary!(sys, op) */
int sph_apply_unary(sph_system* s, int op, const sph_params* params);
/* apply!(sys, op; self): dispatches on the operator id. */
int sph_apply(sph_system* s, int op, const sph_params* params, int self);

/* Global reductions; out receives 1 or 3 doubles. */
int sph_reduce(sph_system* s, int what, const sph_params* params, double* out);

/* SmoothedParticles.sum(sys, func, x) (src/core.jl:240-258) at np probe
 * points x[np][3] (host), with the interpolation closures of
 * examples/cavity_flow.jl:166-172.  For probe i, out[4i] receives the sum of
 * (phi*m)*W(h,r) and out[4i+1+k] the sum of ((phi*m)*v_k)*W(h,r) over the
 * particles the reference visits from x_i (its find_key cell and the
 * key_diff neighbours, in that order) whose distance r does not exceed h;
 * W is wendland2 in 2D and wendland3 in 3D, m = params->m, and phi is
 * Float64(type == FLUID) with SPH_PROBE_FLUID_ONLY, else 1.  The velocity
 * interpolant is out[4i+1+k] / out[4i].  Sums are FP64 in the reference's
 * order in either precision mode.  As in the reference, a probe with a NaN
 * coordinate gets key -1 and a NaN distance is not rejected.  Needs a
 * current cell list; periodic systems: SPH_EUNSUP. */
#define SPH_PROBE_FLUID_ONLY 1
int sph_probe_sum(sph_system* s, const sph_params* params, int flags, const double* x,
                  int64_t np, double* out);

/* SmoothedParticles.sum(sys, func, p) (src/core.jl:271-291) for every
 * particle p: out[p] (reference order, n = sph_count) receives the sum of
 * func(p, q, r) over the particles q the reference visits from p (the cells
 * find_key(p.x) + key_diff in that order, each in its stored order) whose
 * r = norm(p.x - q.x) does not exceed h -- q == p included (r = 0), unlike
 * apply_binary!.  Sums are FP64 in the reference's order in either precision
 * mode.  func is a registered function:
 *   SPH_SUM_LJ_WALL  LJ_potential, examples/collapse_symplectic.jl:144-151;
 *                    fparams = {m, dr_wall, E_wall, eps}: for q a wall
 *                    particle (type 1), p fluid and r < dr_wall,
 *                    m E_wall (0.5 s^2 - 0.25 s^4 - 0.25), s = dr_wall/(r+eps),
 *                    with s^4 formed as (s^2)^2; else 0.
 * Needs a current cell list; periodic systems: SPH_EUNSUP. */
#define SPH_SUM_LJ_WALL 1
int sph_particle_sum(sph_system* s, int func, const double* fparams, int nfparams, double* out,
                     int64_t n);

/* sph_probe_sum without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * probe points and the sums are staged in a static device array of
 * SPH_PROBE_NOALLOC_MAX * 7 doubles (3.5 MiB) that each GPU the library runs
 * on holds until the process ends; np above SPH_PROBE_NOALLOC_MAX is
 * rejected with SPH_EINVAL.  The call allocates nothing.  Calls that overlap
 * on one GPU (several host threads) share the array and race.  Results are
 * the bits of sph_probe_sum. */
#define SPH_PROBE_NOALLOC_MAX 65536
int sph_probe_sum_noalloc(sph_system* s, const sph_params* params, int flags, const double* x,
                          int64_t np, double* out);

/* save_frame! (src/IO.jl:53-75) without stopping the steps: the fields
 * are gathered into reference order on the system's stream (so they are the
 * state at the time of the call) and copied to pinned host memory on a
 * second stream while later steps run.  sph_snapshot_wait blocks until the
 * copy is done; sph_snapshot_field then gives field k's n or 3n doubles
 * ([n][3] for vector fields), valid until sph_snapshot_free.  Up to
 * SPH_SNAPSHOT_MAX_FIELDS fields; free every snapshot before destroying its
 * handle. */
#define SPH_SNAPSHOT_MAX_FIELDS 9
typedef struct sph_snapshot sph_snapshot;
int sph_snapshot_begin(sph_system* s, const int* fields, int nfields, sph_snapshot** out);
int sph_snapshot_wait(sph_snapshot* snap, int64_t* n);
int sph_snapshot_field(sph_snapshot* snap, int field, const double** data);
int sph_snapshot_free(sph_snapshot* snap);

/* Neighbour pairs of the current cell list as reference particle indices
 * (p, q), p != q, dist <= h: pairs[2k], pairs[2k+1].  The order is
 * unspecified.  *n receives the total; at most cap pairs are written. */
int sph_neighbour_pairs(sph_system* s, int64_t* pairs, int64_t cap, int64_t* n);

/* Diagnostics (no counterpart in the reference): an order-independent
 * digest of every particle's neighbour set under the current cell list, for
 * comparing the sets of _apply_binary! (src/core.jl:94-112) at sizes where
 * listing the pairs is too large.  For particle p (reference order, n must
 * equal sph_count) counts[p] receives the number of accepted q and
 * hashes[p] the sum modulo 2^64 of mix64(q + 1) over them, mix64 being the
 * splitmix64 finaliser:
 *   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9; z = (z ^ (z >> 27)) *
 *   0x94d049bb133111eb; return z ^ (z >> 31).
 * Either output may be NULL. */
int sph_neighbour_digest(sph_system* s, int64_t* counts, uint64_t* hashes, int64_t n);

/* nsteps fused steps of a scheme, device-resident, one host sync at the end.
 * SPH_SCHEME_COLLAPSE_DRY is the time loop of examples/collapse_dry.jl:202-211.
 * SPH_SCHEME_WCSPH_SE(_RHO0) is a symplectic Euler step assembled from the
 * reference's operators (create_cell_list!, reset_rho!, find_rho! with self,
 * find_pressure!, reset_a!, internal_force!, then a full-step fluid kick and
 * drift); no reference program runs exactly this loop -- it is the
 * density-summation scheme of the dam-break and Taylor-Green workloads.
 * *pairs (may be NULL) receives the accepted ordered pairs summed over the
 * first pair pass of every step (density sum, or balance_of_mass!). */
int sph_step(sph_system* s, int scheme, const sph_params* params, int nsteps,
             int64_t* pairs);

/* Accepted ordered pairs of the last sph_step call, summed over its steps:
 * density passes (all targets) and force passes (fluid targets only). */
int sph_step_pairs(sph_system* s, int64_t* density_pairs, int64_t* force_pairs);

/* Diagnostics (no counterpart in the reference): distance tests the density
 * passes of the last sph_step call performed, summed over its steps -- one
 * bounding-box test per candidate per tile of targets, and one pair test per
 * candidate the box test kept, per target of the tile.  The reference tests
 * every candidate of the cells key + key_diff (sph_candidates). */
int sph_step_tests(sph_system* s, int64_t* density_tests);

/* Cell geometry as the reference computes it (src/structs.jl:63-82):
 * key_phase[3], key_lim[3], key_max, key_diff[27] (9 used in 2D). */
int sph_key_geometry(sph_system* s, int64_t* key_phase, int64_t* key_lim,
                     int64_t* key_max, int64_t* key_diff, int* n_diff);

/* Targets per tile of the pair kernels (8, 16, 32 or 64; 0 = automatic: the
 * smallest power of two >= the mean occupancy of the non-empty cells, at
 * most 64, chosen at each sph_build_cells).  The summation order of the
 * pair sums, hence their last bits, depends on the tile size.
 * sph_autotune: in SPH_MIXED mode it times the density and force passes for
 * each tile size on the current cell list and keeps the fastest; in SPH_FP64
 * mode it keeps the automatic choice, so the plain FP64 path gives the same
 * bits for the same input on every run.  *best receives the tile size. */
int sph_set_tile(sph_system* s, int targets_per_tile);
int sph_autotune(sph_system* s, const sph_params* params, int* best);

/* Options (no counterpart in the reference).  SPH_OPTION_PAIR_MASKS (default
 * 1): a density pass records, per target and 64-candidate chunk, the set of
 * accepted candidates, and later pair passes on the same cell list read it
 * instead of re-testing candidates.  0 makes every pass test.
 * SPH_OPTION_KEPT_LISTS (default 1): the density pass also records each
 * tile's candidates that survived the bounding-box cull, and the passes that
 * read the masks gather those instead of streaming and culling the whole
 * candidate list again.  Either setting gives the same bits. */
#define SPH_OPTION_PAIR_MASKS 0
#define SPH_OPTION_KEPT_LISTS 1
/* SPH_OPTION_GRAPH (default 0): sph_step / sph_batch_step capture the
 * kernels of one or two steps into a hipGraph once per call and replay it
 * for the call's steps, instead of enqueueing every kernel of every step
 * (about 25 launches per step).  Same bits; kernel profiling is off while
 * the graph replays. */
#define SPH_OPTION_GRAPH 2
int sph_set_option(sph_system* s, int option, int value);

/* Diagnostics (no counterpart in the reference): candidate particles the
 * pair kernels test per pass under the current cell list, i.e. for every
 * particle the particles of the cells key + key_diff it visits
 * (src/core.jl:96-102); *all over all targets, *fluid over fluid targets. */
int sph_candidates(sph_system* s, int64_t* all, int64_t* fluid);

/* Kernel timing with HIP events on the system's stream.  Kernel ids:
 * 0 cell build, 1 density pass, 2 force pass, 3 unary/integrator. */
#define SPH_KERNEL_BUILD 0
#define SPH_KERNEL_DENSITY 1
#define SPH_KERNEL_FORCE 2
#define SPH_KERNEL_UNARY 3
int sph_profile(sph_system* s, int enable);
int sph_profile_read(sph_system* s, int kernel, double* total_ms, int64_t* launches);

/* Message of the last failing call on this host thread. */
const char* sph_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SPH_H */
