/*
 * sph_oracle.c -- FP64 CPU restatement of SmoothedParticles.jl v0.2.0's
 * cell-list neighbour search and pair-interaction loop, plus the WCSPH
 * operators, integrator and energy that sit on that path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library
 * is compared against (tests/, __graft_entry__.smoke) and the CPU-baseline
 * leg of bench.py.  The shipped library never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference
 * semantics that matter for bit-exact neighbour sets:
 *   - cell size is the ParticleSystem `h` argument (called R here);
 *   - keys are linear, 1-based, with a linear bounds test only
 *     (src/core.jl:98), so rows wrap;
 *   - pair predicate: r = sqrt((d1*d1 + d2*d2) + d3*d3) in FP64 without
 *     FMA, rejected when r > R or p is q (src/core.jl:104-105,
 *     src/algebra.jl:49-60);
 *   - each cell holds particle indices in descending order
 *     (src/core.jl:26-41), visited in key_diff order, di outermost
 *     (src/structs.jl:70-82).
 *
 * How each part is pinned independently of this file (tests/test_oracle*.py):
 *   key geometry ............ closed-form integer tables (golden JSON)
 *   cell list / pairs ........ O(N^2) brute force, lattice neighbour counts,
 *                              exact-tie and duplicate-position cases
 *   neighbour digest ......... brute-force pair list; mix64 against the
 *                              splitmix64 finaliser's published outputs
 *   Wendland W / rDW ......... closed forms, quadrature normalisation,
 *                              rD = D/r (tests/test_kernels.jl:20-43)
 *   density sums (rho, rho0).. lattice sums computed independently (fsum)
 *   pressure forces + AV ..... two-particle closed forms (P/rho^2 and the
 *                              rho0^2 form), momentum antisymmetry
 *   continuity ............... two-particle closed form
 *   EOS (three forms) ........ closed forms
 *   move!/drift/kick ......... closed forms; free fall in closed form for the
 *                              symplectic Euler step and for the leapfrog of
 *                              the collapse_dry time loop
 *   energy (three forms) ..... closed-form energy of hand-built states and the
 *                              energy-drift bound of tests/test_collision_2d.jl
 *   momentum ................. exact sum (math.fsum)
 *   periodic systems ......... numpy minimum-image brute force
 *                              (tests/test_oracle_periodic.py)
 *   probe sums ............... brute force over all particles, an exact visit
 *                              count, the density sum at particle positions,
 *                              closed forms, the NaN-key rule
 *                              (tests/test_oracle_probe.py)
 *   assemble_matrix visits ... all-pairs brute force, aliased-grid
 *                              multiplicities (tests/test_oracle_assemble.py)
 * Parity unpinned at the ulp level: the @fastmath kernels (src/kernels.jl)
 * and Julia's libm log in the energy; those are compared by tolerance.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct orc_params {
    double m;      /* particle mass */
    double rho0;   /* reference density */
    double c;      /* speed of sound */
    double alpha;  /* artificial-viscosity alpha */
    double hs;     /* smoothing length used by the viscosity term */
    double mu;     /* laminar viscosity (examples/collapse_dry.jl:139) */
    double nu;     /* density diffusion (examples/collapse_dry.jl:114) */
    double g[3];   /* gravity */
    double dt;     /* time step */
    double coeff;  /* multiplier of dt in move/accelerate (0.5 = half step) */
} orc_params;

enum { F_X = 0, F_V, F_A, F_RHO, F_P, F_DRHO, F_RHO0, F_TYPE, F_ID };

typedef struct orc_sys {
    double h;
    double bmin[3], bmax[3];
    int64_t key_phase[3], key_lim[3], key_max;
    int64_t key_diff[27];
    int n_diff;
    int dim;
    int64_t n, cap;
    double *x, *v, *a;                       /* [cap*3] */
    double *rho, *P, *Drho, *rho0, *type;    /* [cap]   */
    int64_t *id;                             /* stable insertion index */
    int64_t *cell_start;                     /* [key_max+1], 0-based cells */
    int64_t *cell_idx;                       /* [cap], descending per cell */
    int cells_valid;
    /* periodic systems (not in the reference): minimum-image pairs in a box
     * periodic along x, y (and z unless bmin[2] == bmax[2]) */
    int periodic, axes;
    double box[3], cell[3];
} orc_sys;

/* ---------------------------------------------------------------- keys -- */

/* src/structs.jl:63-82 -- key_phase, key_lim, key_max and key_diff. */
static void key_geometry(orc_sys* s)
{
    for (int d = 0; d < 3; ++d) {
        s->key_phase[d] = (int64_t)floor(s->bmin[d] / s->h);
        s->key_lim[d] = (int64_t)floor(s->bmax[d] / s->h) - s->key_phase[d] + 1;
    }
    s->key_max = s->key_lim[0] * s->key_lim[1] * s->key_lim[2];
    s->n_diff = 0;
    if (s->key_lim[2] == 1) {
        s->dim = 2;
        for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj)
                s->key_diff[s->n_diff++] = di + s->key_lim[0] * dj;
    } else {
        s->dim = 3;
        for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj)
                for (int dk = -1; dk <= 1; ++dk)
                    s->key_diff[s->n_diff++] =
                        di + s->key_lim[0] * (dj + s->key_lim[1] * dk);
    }
}

/* Int64(floor(v)) of Julia: throws (-> caught, key -1) when v is not finite
 * or outside the Int64 range. */
static int jl_floor_int(double v, int64_t* out)
{
    double f = floor(v);
    if (!(f >= -9223372036854775808.0 && f < 9223372036854775808.0)) return 0;
    *out = (int64_t)f;
    return 1;
}

/* src/structs.jl:97-106 -- 1-based linear key, -1 for NaN/Inf. */
static int64_t find_key(const orc_sys* s, const double* x)
{
    int64_t f[3];
    for (int d = 0; d < 3; ++d)
        if (!jl_floor_int(x[d] / s->h, &f[d])) return -1;
    int64_t i = 1 + f[0] - s->key_phase[0];
    int64_t j = 1 + f[1] - s->key_phase[1];
    int64_t k = 1 + f[2] - s->key_phase[2];
    return i + s->key_lim[0] * (j - 1) + s->key_lim[0] * s->key_lim[1] * (k - 1);
}

/* src/geometry.jl:24-30 -- inclusive Box test (false for NaN). */
static int is_inside_box(const orc_sys* s, const double* x)
{
    return s->bmin[0] <= x[0] && x[0] <= s->bmax[0] &&
           s->bmin[1] <= x[1] && x[1] <= s->bmax[1] &&
           s->bmin[2] <= x[2] && x[2] <= s->bmax[2];
}

/* Periodic systems: cell of a position wrapped into the box,
 * floor((x - bmin) / cell) clamped to [0, L-1] on each periodic axis. */
static void periodic_cell(const orc_sys* s, const double* x, int64_t* c)
{
    for (int d = 0; d < 3; ++d) {
        c[d] = 0;
        if (d >= s->axes) continue;
        int64_t v = (int64_t)floor((x[d] - s->bmin[d]) / s->cell[d]);
        c[d] = v < 0 ? 0 : (v >= s->key_lim[d] ? s->key_lim[d] - 1 : v);
    }
}

/* x - box * floor((x - bmin) / box) on each periodic axis. */
static void periodic_wrap(const orc_sys* s, double* x)
{
    for (int d = 0; d < s->axes; ++d)
        x[d] = x[d] - s->box[d] * floor((x[d] - s->bmin[d]) / s->box[d]);
}

static int is_finite3(const double* x)
{
    return isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]);
}

/* Cells particle p visits (0-based) with the image shift added to the
 * candidates' coordinates.  Reference systems: key + key_diff with the
 * linear bounds test of src/core.jl:96-99, no shift.  Periodic systems: the
 * 3^dim neighbouring cells, wrapped per axis, shifted by -+box across the
 * seam (every periodic axis has L >= 3 cells of size >= h, so each
 * neighbouring cell appears once and holds the minimum image). */
static int neighbour_cells(const orc_sys* s, const double* xp, int64_t* cells,
                           double (*shift)[3])
{
    int n = 0;
    if (!s->periodic) {
        int64_t key = find_key(s, xp);
        for (int t = 0; t < s->n_diff; ++t) {
            int64_t nk = key + s->key_diff[t];
            if (1 <= nk && nk <= s->key_max) {
                cells[n] = nk - 1;
                shift[n][0] = shift[n][1] = shift[n][2] = 0.0;
                ++n;
            }
        }
        return n;
    }
    int64_t c[3];
    periodic_cell(s, xp, c);
    int kr = s->axes == 3 ? 1 : 0;
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -kr; dk <= kr; ++dk) {
                int64_t idx[3] = {c[0] + di, c[1] + dj, c[2] + dk};
                for (int d = 0; d < 3; ++d) {
                    shift[n][d] = 0.0;
                    if (idx[d] < 0) { idx[d] += s->key_lim[d]; shift[n][d] = -s->box[d]; }
                    else if (idx[d] >= s->key_lim[d]) { idx[d] -= s->key_lim[d]; shift[n][d] = s->box[d]; }
                }
                cells[n] = idx[0] + s->key_lim[0] * (idx[1] + s->key_lim[1] * idx[2]);
                ++n;
            }
    return n;
}

/* src/core.jl:8-10, src/algebra.jl:49-60 -- FP64 distance, no FMA. */
static double dist3(const double* p, const double* q)
{
    double d1 = p[0] - q[0], d2 = p[1] - q[1], d3 = p[2] - q[2];
    double s = d1 * d1 + d2 * d2;
    s = s + d3 * d3;
    return sqrt(s);
}

/* ------------------------------------------------------------- kernels -- */
/* src/kernels.jl:108-115 */
double orc_wendland2(double h, double r)
{
    double x = r / h;
    if (x > 1.0) return 0.0;
    double u = 1.0 - x, u2 = u * u, u4 = u2 * u2;
    return 2.228169203286535 * u4 * (1.0 + 4.0 * x) / (h * h);
}
/* src/kernels.jl:140-147 */
double orc_rDwendland2(double h, double r)
{
    double x = r / h;
    if (x > 1.0) return 0.0;
    double u = 1.0 - x, h2 = h * h;
    return -44.563384065730695 * (u * u * u) / (h2 * h2);
}
/* src/kernels.jl:156-163 */
double orc_wendland3(double h, double r)
{
    double x = r / h;
    if (x > 1.0) return 0.0;
    double u = 1.0 - x, u2 = u * u, u4 = u2 * u2;
    return 3.3422538049298023 * u4 * (1.0 + 4.0 * x) / (h * h * h);
}
/* src/kernels.jl:188-195 */
double orc_rDwendland3(double h, double r)
{
    double x = r / h;
    if (x > 1.0) return 0.0;
    double u = 1.0 - x, h2 = h * h;
    return -66.84507609859604 * (u * u * u) / (h2 * h2 * h);
}
/* src/kernels.jl:124-131, :172-179 (only used by the kernel tests) */
double orc_Dwendland2(double h, double r)
{
    double x = r / h;
    if (x > 1.0) return 0.0;
    double u = 1.0 - x;
    return -44.563384065730695 * x * (u * u * u) / (h * h * h);
}
double orc_Dwendland3(double h, double r)
{
    double x = r / h;
    if (x > 1.0) return 0.0;
    double u = 1.0 - x, h2 = h * h;
    return -66.84507609859604 * x * (u * u * u) / (h2 * h2);
}

static double W(const orc_sys* s, double r)
{
    return s->dim == 2 ? orc_wendland2(s->h, r) : orc_wendland3(s->h, r);
}
static double rDW(const orc_sys* s, double r)
{
    return s->dim == 2 ? orc_rDwendland2(s->h, r) : orc_rDwendland3(s->h, r);
}

/* ----------------------------------------------------------- lifecycle -- */

static void free_arrays(orc_sys* s)
{
    free(s->x); free(s->v); free(s->a);
    free(s->rho); free(s->P); free(s->Drho); free(s->rho0); free(s->type);
    free(s->id); free(s->cell_idx);
    s->x = s->v = s->a = s->rho = s->P = s->Drho = s->rho0 = s->type = NULL;
    s->id = NULL; s->cell_idx = NULL;
}

/* src/structs.jl:57-92 -- ParticleSystem(T, domain, h) with domain = Box. */
orc_sys* orc_create(double h, const double* bmin, const double* bmax)
{
    if (!(h > 0.0)) return NULL;
    orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->h = h;
    memcpy(s->bmin, bmin, 3 * sizeof(double));
    memcpy(s->bmax, bmax, 3 * sizeof(double));
    key_geometry(s);
    s->cell_start = (int64_t*)calloc((size_t)s->key_max + 1, sizeof(int64_t));
    return s;
}

/* A periodic system (not in the reference) on the box [bmin, bmax): each
 * periodic axis is split into L = floor(box/h) cells (reduced while
 * box/L < h), L >= 3.  Returns NULL when the box is too small. */
orc_sys* orc_create_periodic(double h, const double* bmin, const double* bmax)
{
    if (!(h > 0.0)) return NULL;
    orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->h = h;
    memcpy(s->bmin, bmin, 3 * sizeof(double));
    memcpy(s->bmax, bmax, 3 * sizeof(double));
    s->periodic = 1;
    s->axes = bmin[2] == bmax[2] ? 2 : 3;
    for (int d = 0; d < 3; ++d) {
        s->key_phase[d] = 0;
        if (d >= s->axes) {
            s->key_lim[d] = 1;
            s->cell[d] = 1.0;
            continue;
        }
        s->box[d] = bmax[d] - bmin[d];
        double L = floor(s->box[d] / h);
        while (L >= 1.0 && s->box[d] / L < h) L -= 1.0;
        if (!(L >= 3.0)) {
            free(s);
            return NULL;
        }
        s->key_lim[d] = (int64_t)L;
        s->cell[d] = s->box[d] / L;
    }
    s->key_max = s->key_lim[0] * s->key_lim[1] * s->key_lim[2];
    s->dim = s->key_lim[2] == 1 ? 2 : 3;
    s->n_diff = s->axes == 3 ? 27 : 9;
    s->cell_start = (int64_t*)calloc((size_t)s->key_max + 1, sizeof(int64_t));
    return s;
}

void orc_destroy(orc_sys* s)
{
    if (!s) return;
    free_arrays(s);
    free(s->cell_start);
    free(s);
}

/* Resize to n particles; new particles are zero-initialised and get
 * id = index. */
int orc_resize(orc_sys* s, int64_t n)
{
    free_arrays(s);
    s->n = s->cap = n;
    size_t c = n > 0 ? (size_t)n : 1;
    s->x = calloc(3 * c, sizeof(double));
    s->v = calloc(3 * c, sizeof(double));
    s->a = calloc(3 * c, sizeof(double));
    s->rho = calloc(c, sizeof(double));
    s->P = calloc(c, sizeof(double));
    s->Drho = calloc(c, sizeof(double));
    s->rho0 = calloc(c, sizeof(double));
    s->type = calloc(c, sizeof(double));
    s->id = calloc(c, sizeof(int64_t));
    s->cell_idx = calloc(c, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) s->id[i] = i;
    s->cells_valid = 0;
    return 0;
}

int64_t orc_n(const orc_sys* s) { return s->n; }
int orc_dim(const orc_sys* s) { return s->dim; }

void* orc_field(orc_sys* s, int f)
{
    switch (f) {
    case F_X: return s->x;
    case F_V: return s->v;
    case F_A: return s->a;
    case F_RHO: return s->rho;
    case F_P: return s->P;
    case F_DRHO: return s->Drho;
    case F_RHO0: return s->rho0;
    case F_TYPE: return s->type;
    case F_ID: return s->id;
    }
    return NULL;
}

void orc_key_geometry(const orc_sys* s, int64_t* phase3, int64_t* lim3,
                      int64_t* key_max, int64_t* diff27, int* n_diff)
{
    memcpy(phase3, s->key_phase, 3 * sizeof(int64_t));
    memcpy(lim3, s->key_lim, 3 * sizeof(int64_t));
    *key_max = s->key_max;
    memcpy(diff27, s->key_diff, (size_t)s->n_diff * sizeof(int64_t));
    *n_diff = s->n_diff;
}

int64_t orc_find_key(const orc_sys* s, const double* x) { return find_key(s, x); }

/* ----------------------------------------------------------- cell list -- */

static void copy_particle(orc_sys* s, int64_t dst, int64_t src)
{
    if (dst == src) return;
    memcpy(s->x + 3 * dst, s->x + 3 * src, 3 * sizeof(double));
    memcpy(s->v + 3 * dst, s->v + 3 * src, 3 * sizeof(double));
    memcpy(s->a + 3 * dst, s->a + 3 * src, 3 * sizeof(double));
    s->rho[dst] = s->rho[src];
    s->P[dst] = s->P[src];
    s->Drho[dst] = s->Drho[src];
    s->rho0[dst] = s->rho0[src];
    s->type[dst] = s->type[src];
    s->id[dst] = s->id[src];
}

/* src/core.jl:51-90 -- create_cell_list!: remove particles outside the
 * bounding box by swap-with-end (descending removal indices,
 * src/core.jl:63-81), then insert every particle into its cell so that
 * each cell lists indices in descending order (src/core.jl:26-41,84-89).
 * Returns the new particle count. */
int64_t orc_create_cell_list(orc_sys* s)
{
    int64_t n = s->n;
    /* periodic systems wrap every finite position into the box and
     * remove only non-finite ones */
    if (s->periodic)
        for (int64_t i = 0; i < n; ++i)
            if (is_finite3(s->x + 3 * i)) periodic_wrap(s, s->x + 3 * i);
    /* removal: collect indices outside the box in descending order */
    int64_t m = 0;
    for (int64_t i = n - 1; i >= 0; --i) {
        int keep = s->periodic ? is_finite3(s->x + 3 * i) : is_inside_box(s, s->x + 3 * i);
        if (!keep) s->cell_idx[m++] = i;
    }
    for (int64_t i = 0; i < m; ++i)   /* particles[r_i] = particles[end+1-i] */
        copy_particle(s, s->cell_idx[i], n - 1 - i);
    n -= m;
    s->n = n;

    /* insertion: counts per cell, then fill each cell in descending order */
    int64_t km = s->key_max;
    memset(s->cell_start, 0, (size_t)(km + 1) * sizeof(int64_t));
    int64_t* key = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        if (s->periodic) {
            int64_t c[3];
            periodic_cell(s, s->x + 3 * i, c);
            key[i] = c[0] + s->key_lim[0] * (c[1] + s->key_lim[1] * c[2]);
        } else {
            key[i] = find_key(s, s->x + 3 * i) - 1;   /* 0-based */
        }
        s->cell_start[key[i] + 1]++;
    }
    for (int64_t c = 0; c < km; ++c) s->cell_start[c + 1] += s->cell_start[c];
    int64_t* fill = (int64_t*)malloc((size_t)(km > 0 ? km : 1) * sizeof(int64_t));
    memcpy(fill, s->cell_start, (size_t)km * sizeof(int64_t));
    for (int64_t i = n - 1; i >= 0; --i) s->cell_idx[fill[key[i]]++] = i;
    free(fill);
    free(key);
    s->cells_valid = 1;
    return n;
}

/* Visit every accepted ordered pair of particle p in reference order:
 * src/core.jl:94-112 (_apply_binary!).  The body sees q, r and xq, the
 * candidate's position with the image shift added (the position itself for
 * reference systems). */
#define FOR_EACH_NEIGHBOUR(s, p, ...)                                          \
    do {                                                                       \
        int64_t cells_[27];                                                    \
        double sh_[27][3];                                                     \
        const double* xp_ = (s)->x + 3 * (p);                                  \
        int nc_ = neighbour_cells((s), xp_, cells_, sh_);                      \
        for (int t_ = 0; t_ < nc_; ++t_) {                                     \
            int64_t c_ = cells_[t_];                                           \
            for (int64_t e_ = (s)->cell_start[c_];                             \
                 e_ < (s)->cell_start[c_ + 1]; ++e_) {                         \
                int64_t q = (s)->cell_idx[e_];                                 \
                const double* xs_ = (s)->x + 3 * q;                            \
                double xq[3] = {xs_[0] + sh_[t_][0], xs_[1] + sh_[t_][1],      \
                                xs_[2] + sh_[t_][2]};                          \
                double r = dist3(xp_, xq);                                     \
                if (r > (s)->h || (p) == q) continue;                          \
                __VA_ARGS__                                                    \
            }                                                                  \
        }                                                                      \
    } while (0)

/* Neighbour pairs (p, q) as particle indices in reference visiting order.
 * Writes at most `cap` pairs into out[2*k], out[2*k+1]; returns the total. */
int64_t orc_neighbour_pairs(const orc_sys* s, int64_t* out, int64_t cap)
{
    if (!s->cells_valid) return -1;
    int64_t cnt = 0;
    for (int64_t p = 0; p < s->n; ++p) {
        FOR_EACH_NEIGHBOUR(s, p, {
            if (cnt < cap) { out[2 * cnt] = p; out[2 * cnt + 1] = q; }
            ++cnt;
        });
    }
    return cnt;
}

/* The (p, q) entries assemble_matrix pushes (src/core.jl:196-225): the same
 * cell loop as the pair sweep but without the p == q exclusion (:213-216),
 * each visit once, p ascending.  Writes at most `cap` entries; returns the
 * total, or -1 without a current cell list. */
int64_t orc_matrix_entries(const orc_sys* s, int64_t* out, int64_t cap)
{
    if (!s->cells_valid || s->periodic) return -1;
    int64_t cnt = 0;
    for (int64_t p = 0; p < s->n; ++p) {
        const double* xp = s->x + 3 * p;
        int64_t key = find_key(s, xp);
        for (int t = 0; t < s->n_diff; ++t) {
            int64_t nk = key + s->key_diff[t];
            if (!(1 <= nk && nk <= s->key_max)) continue;
            for (int64_t e = s->cell_start[nk - 1]; e < s->cell_start[nk]; ++e) {
                int64_t q = s->cell_idx[e];
                if (dist3(xp, s->x + 3 * q) > s->h) continue;
                if (cnt < cap) { out[2 * cnt] = p; out[2 * cnt + 1] = q; }
                ++cnt;
            }
        }
    }
    return cnt;
}

/* Number of accepted pairs per particle (reference visiting semantics). */
int64_t orc_neighbour_counts(const orc_sys* s, int64_t* counts)
{
    if (!s->cells_valid) return -1;
    int64_t tot = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tot)
    for (int64_t p = 0; p < s->n; ++p) {
        int64_t c = 0;
        FOR_EACH_NEIGHBOUR(s, p, { (void)q; (void)r; (void)xq; ++c; });
        if (counts) counts[p] = c;
        tot += c;
    }
    return tot;
}

/* Order-independent digest of each particle's neighbour set, for comparing
 * sets at sizes where listing the pairs is too large: counts[p] is the
 * number of accepted q, hash[p] the sum modulo 2^64 of mix(q + 1) over them,
 * mix being the splitmix64 finaliser (Steele, Lea and Flood 2014, constants
 * as published by Vigna).  A repeated visit (aliased small grids) counts
 * twice, so equal digests mean equal multisets with overwhelming
 * probability.  Same neighbour loop as _apply_binary! (src/core.jl:94-112). */
uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

int64_t orc_neighbour_digest(const orc_sys* s, int64_t* counts, uint64_t* hash)
{
    if (!s->cells_valid) return -1;
    int64_t tot = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tot)
    for (int64_t p = 0; p < s->n; ++p) {
        int64_t c = 0;
        uint64_t hh = 0;
        FOR_EACH_NEIGHBOUR(s, p, { (void)r; (void)xq; ++c; hh += orc_mix64((uint64_t)q + 1); });
        counts[p] = c;
        hash[p] = hh;
        tot += c;
    }
    return tot;
}

/* SmoothedParticles.sum(sys, func, x), src/core.jl:240-258, at np probe
 * points x[np][3], with the interpolation closures of
 * examples/cavity_flow.jl:166-172: out[4i] = sum of (phi*m)*W(h,r) and
 * out[4i+1+k] = sum of ((phi*m)*v_k)*W(h,r), phi = Float64(type == FLUID)
 * when fluid_only, else 1.  Cells in key_diff order behind the linear bounds
 * test (:243-245), each cell in its stored order (:246-250), r = norm(x - q.x)
 * (:251), rejected only when r > h (:252), so a NaN r is summed.  There is
 * no p == q exclusion.  Reference (non-periodic) systems only; returns -1
 * without a current cell list or for a periodic system. */
int orc_probe_sum(const orc_sys* s, double m, int fluid_only, const double* x, int64_t np,
                  double* out)
{
    if (!s->cells_valid || s->periodic) return -1;
    for (int64_t i = 0; i < np; ++i) {
        const double* xp = x + 3 * i;
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        int64_t key = find_key(s, xp);
        for (int t = 0; t < s->n_diff; ++t) {
            int64_t nk = key + s->key_diff[t];
            if (!(1 <= nk && nk <= s->key_max)) continue;
            for (int64_t e = s->cell_start[nk - 1]; e < s->cell_start[nk]; ++e) {
                int64_t q = s->cell_idx[e];
                double r = dist3(xp, s->x + 3 * q);
                if (r > s->h) continue;
                double w = (fluid_only ? (s->type[q] == 0.0 ? 1.0 : 0.0) : 1.0) * m;
                double k = W(s, r);
                a[0] = a[0] + w * k;
                for (int d = 0; d < 3; ++d) a[1 + d] = a[1 + d] + (w * s->v[3 * q + d]) * k;
            }
        }
        for (int d = 0; d < 4; ++d) out[4 * i + d] = a[d];
    }
    return 0;
}

/* SmoothedParticles.sum(sys, func, p), src/core.jl:271-291, for every
 * particle p, with func = LJ_potential of examples/collapse_symplectic.jl:
 * 144-151: out[p] = sum over the q visited from find_key(p.x) + key_diff
 * (in that order, each cell in its stored order) with r = norm(p.x - q.x)
 * not above h -- q == p included -- of m E_wall (0.5 s^2 - 0.25 s^4 - 0.25),
 * s = dr_wall/(r + eps), when q is a wall particle (type 1), p a fluid
 * particle and r < dr_wall, else 0.  s^4 is (s^2)^2 here; Julia's s^4 is
 * not restated at the ulp (parity unpinned there).  Returns -1 without a
 * current cell list or for a periodic system. */
int orc_particle_sum_lj(const orc_sys* s, double m, double dr_wall, double E_wall, double eps,
                        double* out)
{
    if (!s->cells_valid || s->periodic) return -1;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t p = 0; p < s->n; ++p) {
        const double* xp = s->x + 3 * p;
        double acc = 0.0;
        int64_t key = find_key(s, xp);
        for (int t = 0; t < s->n_diff; ++t) {
            int64_t nk = key + s->key_diff[t];
            if (!(1 <= nk && nk <= s->key_max)) continue;
            for (int64_t e = s->cell_start[nk - 1]; e < s->cell_start[nk]; ++e) {
                int64_t q = s->cell_idx[e];
                double r = dist3(xp, s->x + 3 * q);
                if (r > s->h) continue;
                double term = 0.0;
                if (s->type[q] == 1.0 && s->type[p] == 0.0 && r < dr_wall) {
                    double sq = dr_wall / (r + eps);
                    double s2 = sq * sq;
                    term = (m * E_wall) * ((0.5 * s2 - 0.25 * (s2 * s2)) - 0.25);
                }
                acc = acc + term;
            }
        }
        out[p] = acc;
    }
    return 0;
}

/* O(N^2) brute force with the same FP64 predicate and no cells: all ordered
 * pairs p != q with dist <= h, p ascending then q ascending. */
int64_t orc_brute_pairs(const double* x, int64_t n, double h, int64_t* out, int64_t cap)
{
    int64_t cnt = 0;
    for (int64_t p = 0; p < n; ++p)
        for (int64_t q = 0; q < n; ++q) {
            if (p == q) continue;
            double r = dist3(x + 3 * p, x + 3 * q);
            if (r > h) continue;
            if (cnt < cap) { out[2 * cnt] = p; out[2 * cnt + 1] = q; }
            ++cnt;
        }
    return cnt;
}

/* --------------------------------------------------------- binary ops -- */

/* tests/test_collision_2d.jl:63-65 (find_rho!) with apply!(...; self=true)
 * (src/core.jl:151-161): rho_p += m*W(h,r) over neighbours, then the self
 * term m*W(h,0) added last. */
int64_t orc_density_sum(orc_sys* s, const orc_params* pr, int self)
{
    double m = pr->m;
    int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cnt)
    for (int64_t p = 0; p < s->n; ++p) {
        double acc = s->rho[p];
        FOR_EACH_NEIGHBOUR(s, p, { (void)q; (void)xq; acc = acc + m * W(s, r); ++cnt; });
        if (self) acc = acc + m * W(s, 0.0);
        s->rho[p] = acc;
    }
    return cnt;
}

/* Same pair sum written to the per-particle rho0 field
 * (tests/test_collision_2d.jl:67-69, find_rho0!). */
void orc_density_sum_rho0(orc_sys* s, const orc_params* pr, int self)
{
    double m = pr->m;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t p = 0; p < s->n; ++p) {
        double acc = s->rho0[p];
        FOR_EACH_NEIGHBOUR(s, p, { (void)q; (void)xq; acc = acc + m * W(s, r); });
        if (self) acc = acc + m * W(s, 0.0);
        s->rho0[p] = acc;
    }
}

/* examples/collapse_dry.jl:135-141 (internal_force!: pressure gradient and
 * laminar viscosity, fluid targets only) plus the artificial viscosity of
 * Monaghan (1992), which the reference does not contain:
 *   Pi = -alpha*c*hs*(v_pq.x_pq) / (rhobar*(r^2 + 0.01*hs^2)) if v_pq.x_pq < 0
 *   Dv += -m*Pi*rDW*x_pq. */
void orc_internal_force(orc_sys* s, const orc_params* pr)
{
    double m = pr->m, mu = pr->mu, rho0 = pr->rho0;
    double avk = -pr->alpha * pr->c * pr->hs, eps = 0.01 * pr->hs * pr->hs;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t p = 0; p < s->n; ++p) {
        if (s->type[p] != 0.0) continue;
        const double* xp = s->x + 3 * p;
        const double* vp = s->v + 3 * p;
        double Ap = s->P[p] / (s->rho[p] * s->rho[p]);
        double a0 = s->a[3 * p], a1 = s->a[3 * p + 1], a2 = s->a[3 * p + 2];
        FOR_EACH_NEIGHBOUR(s, p, {
            const double* vq = s->v + 3 * q;
            double dk = rDW(s, r);
            double ker = m * dk;
            double A = Ap + s->P[q] / (s->rho[q] * s->rho[q]);
            double x0 = xp[0] - xq[0], x1 = xp[1] - xq[1], x2 = xp[2] - xq[2];
            double v0 = vp[0] - vq[0], v1 = vp[1] - vq[1], v2 = vp[2] - vq[2];
            double sp = -ker * A;
            a0 = a0 + sp * x0; a1 = a1 + sp * x1; a2 = a2 + sp * x2;
            if (mu != 0.0) {
                double sl = 2.0 * ker * mu / (rho0 * rho0);
                a0 = a0 + sl * v0; a1 = a1 + sl * v1; a2 = a2 + sl * v2;
            }
            if (avk != 0.0) {
                double vx = v0 * x0 + v1 * x1;
                vx = vx + v2 * x2;
                if (vx < 0.0) {
                    double rhobar = 0.5 * (s->rho[p] + s->rho[q]);
                    double Pi = avk * vx / (rhobar * (r * r + eps));
                    double sv = -m * Pi * dk;
                    a0 = a0 + sv * x0; a1 = a1 + sv * x1; a2 = a2 + sv * x2;
                }
            }
        });
        s->a[3 * p] = a0; s->a[3 * p + 1] = a1; s->a[3 * p + 2] = a2;
    }
}

/* tests/test_collision_2d.jl:75-78 -- pressure force with rho0^2, all
 * particles as targets. */
void orc_pressure_force_rho0(orc_sys* s, const orc_params* pr)
{
    double m = pr->m, r02 = pr->rho0 * pr->rho0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t p = 0; p < s->n; ++p) {
        const double* xp = s->x + 3 * p;
        double a0 = s->a[3 * p], a1 = s->a[3 * p + 1], a2 = s->a[3 * p + 2];
        FOR_EACH_NEIGHBOUR(s, p, {
            double ker = m * rDW(s, r);
            double sp = -ker * (s->P[p] / r02 + s->P[q] / r02);
            a0 = a0 + sp * (xp[0] - xq[0]);
            a1 = a1 + sp * (xp[1] - xq[1]);
            a2 = a2 + sp * (xp[2] - xq[2]);
        });
        s->a[3 * p] = a0; s->a[3 * p + 1] = a1; s->a[3 * p + 2] = a2;
    }
}

/* examples/collapse_dry.jl:112-115 (balance_of_mass!): continuity equation
 * with density diffusion nu. */
int64_t orc_balance_of_mass(orc_sys* s, const orc_params* pr)
{
    double m = pr->m, nu = pr->nu;
    int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cnt)
    for (int64_t p = 0; p < s->n; ++p) {
        const double* xp = s->x + 3 * p;
        const double* vp = s->v + 3 * p;
        double acc = s->Drho[p];
        FOR_EACH_NEIGHBOUR(s, p, {
            const double* vq = s->v + 3 * q;
            double ker = m * rDW(s, r);
            double d = (xp[0] - xq[0]) * (vp[0] - vq[0]) +
                       (xp[1] - xq[1]) * (vp[1] - vq[1]);
            d = d + (xp[2] - xq[2]) * (vp[2] - vq[2]);
            acc = acc + ker * (d + 2.0 * nu * (s->rho[p] - s->rho[q]));
            ++cnt;
        });
        s->Drho[p] = acc;
    }
    return cnt;
}

/* ---------------------------------------------------------- unary ops -- */

/* tests/test_collision_2d.jl:84-86 */
void orc_reset_rho(orc_sys* s) { for (int64_t i = 0; i < s->n; ++i) s->rho[i] = 0.0; }
/* tests/test_collision_2d.jl:80-82 */
void orc_reset_a(orc_sys* s) { memset(s->a, 0, (size_t)(3 * s->n) * sizeof(double)); }

/* tests/test_collision_2d.jl:71-73 / examples/collapse_dry.jl:126 --
 * linear EOS P = c^2*(rho - rho0). use_field selects the per-particle
 * rho0 field (test_collision_2d form) instead of the constant. */
void orc_eos(orc_sys* s, const orc_params* pr, int use_field)
{
    double c2 = pr->c * pr->c;
    for (int64_t i = 0; i < s->n; ++i)
        s->P[i] = c2 * (s->rho[i] - (use_field ? s->rho0[i] : pr->rho0));
}

/* examples/collapse_dry.jl:123-127 -- continuity form of find_pressure! */
void orc_eos_continuity(orc_sys* s, const orc_params* pr)
{
    double c2 = pr->c * pr->c;
    for (int64_t i = 0; i < s->n; ++i) {
        s->rho[i] = s->rho[i] + s->Drho[i] * pr->dt;
        s->Drho[i] = 0.0;
        s->P[i] = c2 * (s->rho[i] - pr->rho0);
    }
}

/* examples/collapse_dry.jl:155-159 (accelerate!): fluid only,
 * v += (coeff*dt)*(a + g). */
void orc_accelerate(orc_sys* s, const orc_params* pr)
{
    double k = pr->coeff * pr->dt;
    for (int64_t i = 0; i < s->n; ++i) {
        if (s->type[i] != 0.0) continue;
        for (int d = 0; d < 3; ++d)
            s->v[3 * i + d] = s->v[3 * i + d] + k * (s->a[3 * i + d] + pr->g[d]);
    }
}

/* examples/collapse_dry.jl:148-153 (move!): p.Dv = VEC0 for every
 * particle, then fluid only x += (coeff*dt)*v. */
void orc_move(orc_sys* s, const orc_params* pr)
{
    double k = pr->coeff * pr->dt;
    for (int64_t i = 0; i < s->n; ++i) {
        s->a[3 * i] = 0.0; s->a[3 * i + 1] = 0.0; s->a[3 * i + 2] = 0.0;
        if (s->type[i] != 0.0) continue;
        for (int d = 0; d < 3; ++d)
            s->x[3 * i + d] = s->x[3 * i + d] + k * s->v[3 * i + d];
    }
}

/* tests/test_collision_2d.jl:88-90 (move!): every particle x += (coeff*dt)*v;
 * the acceleration is kept. */
void orc_drift(orc_sys* s, const orc_params* pr)
{
    double k = pr->coeff * pr->dt;
    for (int64_t i = 0; i < s->n; ++i)
        for (int d = 0; d < 3; ++d)
            s->x[3 * i + d] = s->x[3 * i + d] + k * s->v[3 * i + d];
}

/* tests/test_collision_2d.jl:92-94 (accelerate!): every particle
 * v += (coeff*dt)*a, no gravity. */
void orc_kick(orc_sys* s, const orc_params* pr)
{
    double k = pr->coeff * pr->dt;
    for (int64_t i = 0; i < s->n; ++i)
        for (int d = 0; d < 3; ++d)
            s->v[3 * i + d] = s->v[3 * i + d] + k * s->a[3 * i + d];
}

/* The drift of the fused symplectic Euler step: the fluid drift of
 * examples/collapse_dry.jl:150-152 without its reset of Dv, so the step
 * leaves the last force in a. */
static void drift_fluid(orc_sys* s, const orc_params* pr)
{
    double k = pr->coeff * pr->dt;
    for (int64_t i = 0; i < s->n; ++i) {
        if (s->type[i] != 0.0) continue;
        for (int d = 0; d < 3; ++d)
            s->x[3 * i + d] = s->x[3 * i + d] + k * s->v[3 * i + d];
    }
}

/* ---------------------------------------------------------- reductions -- */

/* examples/collapse_dry.jl:166-171 (energy), summed over all particles. */
static double energy_wcsph(const orc_sys* s, const orc_params* pr, int64_t i)
{
    const double* v = s->v + 3 * i;
    const double* x = s->x + 3 * i;
    double m = pr->m;
    double vv = v[0] * v[0] + v[1] * v[1];
    vv = vv + v[2] * v[2];
    double gx = pr->g[0] * x[0] + pr->g[1] * x[1];
    gx = gx + pr->g[2] * x[2];
    double kinetic = 0.5 * m * vv;
    double potential = -m * gx;
    double internal = m * (pr->c * pr->c) *
                      (log(fabs(s->rho[i] / pr->rho0)) + pr->rho0 / s->rho[i] - 1.0);
    return kinetic + potential + internal;
}

/* Energy whose internal part matches P = c^2 (rho - rho0_p) with the
 * per-particle reference density of find_rho0! (tests/test_collision_2d.jl:
 * 67-73): m c^2 (ln(rho/rho0_p) + rho0_p/rho - 1); kinetic and potential
 * as examples/collapse_dry.jl:166-171. */
static double energy_wcsph_rho0(const orc_sys* s, const orc_params* pr, int64_t i)
{
    const double* v = s->v + 3 * i;
    const double* x = s->x + 3 * i;
    double m = pr->m;
    double vv = v[0] * v[0] + v[1] * v[1];
    vv = vv + v[2] * v[2];
    double gx = pr->g[0] * x[0] + pr->g[1] * x[1];
    gx = gx + pr->g[2] * x[2];
    double kinetic = 0.5 * m * vv;
    double potential = -m * gx;
    double r0 = s->rho0[i];
    double internal = m * (pr->c * pr->c) *
                      (log(fabs(s->rho[i] / r0)) + r0 / s->rho[i] - 1.0);
    return kinetic + potential + internal;
}

/* tests/test_collision_2d.jl:96-100 */
static double energy_collision(const orc_sys* s, const orc_params* pr, int64_t i)
{
    const double* v = s->v + 3 * i;
    double m = pr->m;
    double vv = v[0] * v[0] + v[1] * v[1];
    vv = vv + v[2] * v[2];
    double d = s->rho[i] - s->rho0[i];
    return 0.5 * m * vv + 0.5 * m * (pr->c * pr->c) * (d * d) / (pr->rho0 * pr->rho0);
}

static double pairwise(const orc_sys* s, const orc_params* pr, int kind,
                       int64_t lo, int64_t hi)
{
    if (hi - lo <= 1024) {
        double acc = 0.0;
        for (int64_t i = lo; i < hi; ++i)
            acc += kind == 0 ? energy_wcsph(s, pr, i)
                             : (kind == 1 ? energy_collision(s, pr, i) : energy_wcsph_rho0(s, pr, i));
        return acc;
    }
    int64_t mid = lo + (hi - lo) / 2;
    return pairwise(s, pr, kind, lo, mid) + pairwise(s, pr, kind, mid, hi);
}

/* kind 0: WCSPH energy (collapse_dry.jl:166-171); kind 1: collision energy
 * (test_collision_2d.jl:96-100); kind 2: WCSPH energy with the per-particle
 * rho0.  Pairwise summation as Base.sum. */
double orc_energy(const orc_sys* s, const orc_params* pr, int kind)
{
    return pairwise(s, pr, kind, 0, s->n);
}

void orc_momentum(const orc_sys* s, const orc_params* pr, double* out3)
{
    double a[3] = {0, 0, 0};
    for (int64_t i = 0; i < s->n; ++i)
        for (int d = 0; d < 3; ++d) a[d] += pr->m * s->v[3 * i + d];
    memcpy(out3, a, sizeof(a));
}

/* --------------------------------------------------------------- steps -- */

/* One explicit WCSPH step with density summation and symplectic Euler
 * (the step used by the dam-break and Taylor-Green workloads):
 *   create_cell_list!; reset_rho!; find_rho!(self=true); find_pressure!;
 *   reset_a!; internal_force!; accelerate! (coeff 1); move! (coeff 1).
 * Built from the operators above in the order of
 * tests/test_collision_2d.jl:106-116 with a full kick and drift.
 * use_rho0 selects P = c^2 (rho - rho0_p) with the per-particle reference
 * density (tests/test_collision_2d.jl:71-73) instead of the constant rho0.
 * Returns the accepted ordered pairs of the density passes summed over the
 * steps. */
int64_t orc_step_wcsph(orc_sys* s, const orc_params* pr, int nsteps, int use_rho0)
{
    orc_params p1 = *pr;
    p1.coeff = 1.0;
    int64_t pairs = 0;
    for (int k = 0; k < nsteps; ++k) {
        orc_create_cell_list(s);
        orc_reset_rho(s);
        pairs += orc_density_sum(s, &p1, 1);
        orc_eos(s, &p1, use_rho0);
        orc_reset_a(s);
        orc_internal_force(s, &p1);
        orc_accelerate(s, &p1);
        drift_fluid(s, &p1);
    }
    return pairs;
}

/* The time loop of examples/collapse_dry.jl:202-211, nsteps iterations:
 *   accelerate!; move!; create_cell_list!; balance_of_mass!;
 *   find_pressure! (continuity form); move!; create_cell_list!;
 *   internal_force!; accelerate!
 * with the half steps 0.5*dt of accelerate! and move! (:148-159).  The
 * caller runs the initial create_cell_list! and internal_force! (:198-200).
 * Returns the accepted ordered pairs of the balance_of_mass! passes summed
 * over the steps. */
int64_t orc_step_collapse_dry(orc_sys* s, const orc_params* pr, int nsteps)
{
    orc_params ph = *pr;
    ph.coeff = 0.5;
    int64_t pairs = 0;
    for (int k = 0; k < nsteps; ++k) {
        orc_accelerate(s, &ph);
        orc_move(s, &ph);
        orc_create_cell_list(s);
        pairs += orc_balance_of_mass(s, &ph);
        orc_eos_continuity(s, &ph);
        orc_move(s, &ph);
        orc_create_cell_list(s);
        orc_internal_force(s, &ph);
        orc_accelerate(s, &ph);
    }
    return pairs;
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
