// C ABI of libsph_hip (include/sph.h): system lifetime, device memory,
// operator dispatch by registered id, synchronisation and error reporting.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "sph_internal.h"

using namespace sph;

namespace sph {
void enqueue_build(sph_system* s);
void enqueue_tiles(sph_system* s);
void enqueue_unary(sph_system* s, int op, const Phys& ph, double* prho2);
void enqueue_prho2(sph_system* s);
void enqueue_kick_drift(sph_system* s, const Phys& ph);
void enqueue_reduce(sph_system* s, int what, const Phys& ph, double* out_dev);
void enqueue_field_out(sph_system* s, int col0, int ncomp, double* out_dev);
void enqueue_field_in(sph_system* s, int col0, int ncomp, const double* in_dev);
void enqueue_append(sph_system* s, int n0, int m, const double* in_dev, int has_v, int has_rho,
                    int has_P, int has_type);
long long scan_partials_needed(long long n);
long long reduce_partials_needed(long long n);
void reset_queue(sph_system* s);
}  // namespace sph

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPC(expr)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(SPH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

// Largest s with RN(sqrt(s)) <= h: the squared-distance form of the
// reference's `r > h` rejection (src/core.jl:105).
double sqrt_threshold(double h)
{
    double s = h * h;
    while (std::sqrt(s) > h) s = std::nextafter(s, 0.0);
    while (std::sqrt(std::nextafter(s, INFINITY)) <= h) s = std::nextafter(s, INFINITY);
    return s;
}

// Wendland quintic constants of src/kernels.jl:108-195.
Phys make_phys(const sph_system* s, const sph_params* p)
{
    Phys ph{};
    if (p) {
        ph.m = p->m; ph.rho0 = p->rho0; ph.c = p->c; ph.alpha = p->alpha; ph.hs = p->hs;
        ph.mu = p->mu; ph.nu = p->nu;
        ph.g[0] = p->g[0]; ph.g[1] = p->g[1]; ph.g[2] = p->g[2];
        ph.dt = p->dt; ph.coeff = p->coeff;
    }
    const double h = s->geo.h;
    if (s->geo.dim == 2) {
        ph.kW = ph.m * (2.228169203286535 / (h * h));
        ph.kdW = ph.m * (-44.563384065730695 / ((h * h) * (h * h)));
    } else {
        ph.kW = ph.m * (3.3422538049298023 / (h * h * h));
        ph.kdW = ph.m * (-66.84507609859604 / ((h * h) * (h * h) * h));
    }
    ph.avk = -ph.alpha * ph.c * ph.hs;
    ph.eps = 0.01 * ph.hs * ph.hs;
    return ph;
}

template <class T>
int dalloc(T** p, long long n)
{
    *p = nullptr;
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) return fail(SPH_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

template <class T>
void dfree(T*& p)
{
    if (p) hipFree((void*)p);
    p = nullptr;
}

void free_particle_arrays(sph_system* s)
{
    for (int b = 0; b < 2; ++b) {
        dfree(s->cols[b]);
        dfree(s->icols[b]);
    }
    dfree(s->slot); dfree(s->tmp_idx); dfree(s->tmp_key); dfree(s->tmp_ref);
    dfree(s->perm); dfree(s->inv); dfree(s->tiles); dfree(s->prho2); dfree(s->stage);
    dfree(s->red_part);
}

// Grow the particle arrays to hold `need` particles, keeping the contents
// of the current buffer.
int ensure_capacity(sph_system* s, long long need)
{
    if (need <= s->cap) return 0;
    if (need >= (1ll << 31) - 64) return fail(SPH_EINVAL, "particle count exceeds 2^31");
    long long ncap = need;
    if (ncap < 2 * s->cap) ncap = 2 * s->cap;
    if (ncap < 1024) ncap = 1024;
    if (ncap >= (1ll << 31) - 64) ncap = need;
    double* nc[2] = {nullptr, nullptr};
    int* ni[2] = {nullptr, nullptr};
    int rc;
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(&nc[b], NCOL * ncap))) return rc;
        if ((rc = dalloc(&ni[b], NICOL * ncap))) return rc;
    }
    if (s->cap > 0 && s->n > 0) {
        for (int f = 0; f < NCOL; ++f)
            HIPC(hipMemcpyAsync(nc[s->cur] + f * ncap, s->cols[s->cur] + f * s->cap,
                                sizeof(double) * s->n, hipMemcpyDeviceToDevice, s->stream));
        for (int f = 0; f < NICOL; ++f)
            HIPC(hipMemcpyAsync(ni[s->cur] + f * ncap, s->icols[s->cur] + f * s->cap,
                                sizeof(int) * s->n, hipMemcpyDeviceToDevice, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
    }
    free_particle_arrays(s);
    s->cols[0] = nc[0]; s->cols[1] = nc[1];
    s->icols[0] = ni[0]; s->icols[1] = ni[1];
    s->cap = ncap;
    s->tile_cap = ncap / 8 + std::min<long long>(ncap, s->geo.key_max) + 2;
    s->red_part_n = 3 * reduce_partials_needed(ncap);
    if ((rc = dalloc(&s->slot, ncap)) || (rc = dalloc(&s->tmp_idx, ncap)) ||
        (rc = dalloc(&s->tmp_key, ncap)) || (rc = dalloc(&s->tmp_ref, ncap)) ||
        (rc = dalloc(&s->perm, ncap)) || (rc = dalloc(&s->inv, ncap)) ||
        (rc = dalloc(&s->tiles, s->tile_cap)) || (rc = dalloc(&s->prho2, ncap)) ||
        (rc = dalloc(&s->stage, 9 * ncap)) || (rc = dalloc(&s->red_part, s->red_part_n)))
        return rc;
    return 0;
}

int sync_meta(sph_system* s)
{
    HIPC(hipMemcpyAsync(s->meta_host, s->meta, sizeof(int) * META_SLOTS, hipMemcpyDeviceToHost,
                        s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    s->n = s->meta_host[META_N];
    return 0;
}

void prof_begin(sph_system* s, int k, hipEvent_t* ev)
{
    if (!s->prof) return;
    for (int i = 0; i < 2; ++i) {
        if (s->ev_pool.empty()) {
            hipEvent_t e;
            hipEventCreate(&e);
            s->ev_pool.push_back(e);
        }
        ev[i] = s->ev_pool.back();
        s->ev_pool.pop_back();
    }
    hipEventRecord(ev[0], s->stream);
    (void)k;
}

void prof_end(sph_system* s, int k, hipEvent_t* ev)
{
    if (!s->prof) return;
    hipEventRecord(ev[1], s->stream);
    s->ev_open.push_back({k, {ev[0], ev[1]}});
}

void prof_collect(sph_system* s)
{
    if (s->ev_open.empty()) return;
    hipStreamSynchronize(s->stream);
    for (auto& e : s->ev_open) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e.second.first, e.second.second);
        s->prof_ms[e.first] += ms;
        s->prof_n[e.first] += 1;
        s->ev_pool.push_back(e.second.first);
        s->ev_pool.push_back(e.second.second);
    }
    s->ev_open.clear();
}

int choose_tile(double occupancy)
{
    int T = 8;
    while (T < 64 && T < occupancy) T <<= 1;
    return T;
}

PairCtx make_ctx(sph_system* s, const Phys& ph)
{
    PairCtx c{};
    c.g = s->geo;
    c.ph = ph;
    c.T = s->tile_T;
    c.cap = (int)s->cap;
    c.tiles = s->tiles;
    c.cell_start = s->cell_start;
    c.meta = s->meta;
    c.queue = s->meta + META_QUEUE;
    c.cols = s->cols[s->cur];
    c.refidx = s->icols[s->cur] + I_REF * s->cap;
    c.prho2 = s->prho2;
    c.masks = s->masks;
    c.mask_mode = MASK_TEST;
    return c;
}

// Grow the pair-mask buffer to what the current tile table needs and mark
// it usable; called at host sync points (the fused step tests instead of
// loading while the buffer is too small).
int ensure_masks(sph_system* s)
{
    if (!s->use_masks) return 0;
    unsigned long long need = 0;
    HIPC(hipMemcpyAsync(&need, s->mask_total, sizeof(need), hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    if ((long long)need > s->mask_cap) {
        dfree(s->masks);
        s->mask_cap = 0;
        const long long cap = (long long)need + (long long)need / 4 + 4096;
        if (dalloc(&s->masks, cap)) {
            s->masks = nullptr;  // no masks: the kernels keep testing
            g_err.clear();
            return 0;
        }
        s->mask_cap = cap;
    }
    const int ok = 1;
    HIPC(hipMemcpyAsync(s->meta + META_MASK_OK, &ok, sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

// Grid of the pair kernels: the tiles are taken from a work queue, so a
// grid that fills the chip is enough.
int pair_grid(const sph_system* s, long long ntiles)
{
    long long g = std::min<long long>(ntiles, 256 * 32);
    return (int)std::max<long long>(g, 1);
}

int run_pair(sph_system* s, int op, const Phys& ph, int self, int overwrite,
             unsigned long long* counter, long long ntiles, int prof_kernel, int mask_mode)
{
    PairCtx c = make_ctx(s, ph);
    c.self = self;
    c.overwrite = overwrite;
    c.pair_counter = counter;
    c.mask_mode = s->use_masks ? mask_mode : MASK_TEST;
    hipEvent_t ev[2];
    prof_begin(s, prof_kernel, ev);
    reset_queue(s);
    launch_pair(op, s->precision, c, pair_grid(s, ntiles), s->stream);
    prof_end(s, prof_kernel, ev);
    HIPC(hipGetLastError());
    return 0;
}

bool is_binary(int op) { return op >= SPH_OP_DENSITY_SUM && op <= SPH_OP_BALANCE_OF_MASS; }
bool is_unary(int op) { return op >= SPH_OP_RESET_RHO && op <= SPH_OP_MOVE; }

int col_of_field(int field, int* ncomp)
{
    *ncomp = 1;
    switch (field) {
    case SPH_FIELD_X: *ncomp = 3; return C_X;
    case SPH_FIELD_V: *ncomp = 3; return C_VX;
    case SPH_FIELD_A: *ncomp = 3; return C_AX;
    case SPH_FIELD_RHO: return C_RHO;
    case SPH_FIELD_P: return C_P;
    case SPH_FIELD_DRHO: return C_DRHO;
    case SPH_FIELD_RHO0: return C_RHO0;
    case SPH_FIELD_TYPE: return C_TYPE;
    }
    return -1;
}

}  // namespace

namespace sph {
__global__ void k_reset_queue(int* meta) { meta[META_QUEUE] = 0; }
void reset_queue(sph_system* s) { k_reset_queue<<<1, 1, 0, s->stream>>>(s->meta); }
}  // namespace sph

extern "C" {

const char* sph_last_error(void) { return g_err.c_str(); }

int sph_create(sph_system** out, double h, const double* bmin, const double* bmax, int precision)
{
    if (!out || !bmin || !bmax) return fail(SPH_EINVAL, "null argument");
    *out = nullptr;
    if (!(h > 0.0) || !std::isfinite(h))
        return fail(SPH_EINVAL, "invalid ParticleSystem declaration! (h must be a positive float)");
    if (precision != SPH_FP64 && precision != SPH_MIXED)
        return fail(SPH_EINVAL, "unknown precision mode");
    Geo g{};
    g.h = h;
    g.inv_h = 1.0 / h;
    g.s_thr = sqrt_threshold(h);
    long long lim[3];
    for (int d = 0; d < 3; ++d) {
        if (!std::isfinite(bmin[d]) || !std::isfinite(bmax[d]) || bmin[d] > bmax[d])
            return fail(SPH_EINVAL, "invalid bounding box");
        g.bmin[d] = bmin[d];
        g.bmax[d] = bmax[d];
        // src/structs.jl:66-67
        const double fmin = std::floor(bmin[d] / h), fmax = std::floor(bmax[d] / h);
        if (std::fabs(fmin) > 1e15 || std::fabs(fmax) > 1e15)
            return fail(SPH_EINVAL, "bounding box too large for h");
        g.kp[d] = (long long)fmin;
        lim[d] = (long long)fmax - g.kp[d] + 1;
    }
    const long long km = lim[0] * lim[1] * lim[2];
    if (km <= 0 || km > (1ll << 30))
        return fail(SPH_EINVAL, "cell grid too large (key_max > 2^30)");
    g.L1 = (int)lim[0];
    g.L2 = (int)lim[1];
    g.L3 = (int)lim[2];
    g.key_max = (int)km;
    g.dim = lim[2] == 1 ? 2 : 3;  // src/structs.jl:70
    g.flat = bmin[2] == bmax[2];

    sph_system* s = new sph_system();
    std::memset(s->prof_ms, 0, sizeof(s->prof_ms));
    std::memset(s->prof_n, 0, sizeof(s->prof_n));
    s->geo = g;
    s->precision = precision;
    s->tile_T = 32;
    s->tile_auto = 1;
    s->n = 0;
    s->cap = 0;
    s->cells_valid = false;
    s->cur = 0;
    s->cols[0] = s->cols[1] = nullptr;
    s->icols[0] = s->icols[1] = nullptr;
    s->slot = s->tmp_idx = s->tmp_key = s->tmp_ref = s->perm = s->inv = nullptr;
    s->tiles = nullptr;
    s->prho2 = s->stage = s->red_part = nullptr;
    s->cell_chunks = nullptr;
    s->mask_start = s->scan_part64 = s->masks = s->mask_total = nullptr;
    s->mask_cap = 0;
    s->masks_valid = false;
    s->use_masks = true;
    s->prof = false;
    s->step_pairs[0] = s->step_pairs[1] = 0;
    int rc;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return fail(SPH_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    s->scan_part_n = scan_partials_needed(km + 2);
    if ((rc = dalloc(&s->meta, META_SLOTS)) || (rc = dalloc(&s->pair_counter, 4)) ||
        (rc = dalloc(&s->cnt, km + 1)) || (rc = dalloc(&s->cell_start, km + 2)) ||
        (rc = dalloc(&s->tile_start, km + 2)) || (rc = dalloc(&s->scan_part, s->scan_part_n)) ||
        (rc = dalloc(&s->red_out, 4)) || (rc = dalloc(&s->cell_chunks, km + 1)) ||
        (rc = dalloc(&s->mask_start, km + 2)) || (rc = dalloc(&s->scan_part64, s->scan_part_n)) ||
        (rc = dalloc(&s->mask_total, 1))) {
        sph_destroy(s);
        return rc;
    }
    e = hipHostMalloc((void**)&s->meta_host, sizeof(int) * META_SLOTS, hipHostMallocDefault);
    if (e != hipSuccess) {
        s->meta_host = nullptr;
        sph_destroy(s);
        return fail(SPH_EHIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    std::memset(s->meta_host, 0, sizeof(int) * META_SLOTS);
    e = hipMemsetAsync(s->meta, 0, sizeof(int) * META_SLOTS, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        sph_destroy(s);
        return fail(SPH_EHIP, std::string("init: ") + hipGetErrorString(e));
    }
    *out = s;
    return 0;
}

int sph_destroy(sph_system* s)
{
    if (!s) return 0;
    if (s->stream) hipStreamSynchronize(s->stream);
    free_particle_arrays(s);
    dfree(s->meta); dfree(s->pair_counter); dfree(s->cnt); dfree(s->cell_start);
    dfree(s->tile_start); dfree(s->scan_part); dfree(s->red_out);
    dfree(s->cell_chunks); dfree(s->mask_start); dfree(s->scan_part64); dfree(s->masks);
    dfree(s->mask_total);
    if (s->meta_host) hipHostFree(s->meta_host);
    for (auto e : s->ev_pool) hipEventDestroy(e);
    for (auto& e : s->ev_open) {
        hipEventDestroy(e.second.first);
        hipEventDestroy(e.second.second);
    }
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

int sph_append(sph_system* s, int64_t n, const double* x, const double* v, const double* rho,
               const double* P, const double* type)
{
    if (!s || n < 0 || (n > 0 && !x)) return fail(SPH_EINVAL, "invalid append arguments");
    if (n == 0) return 0;
    int rc;
    if ((rc = ensure_capacity(s, s->n + n))) return rc;
    // pack x, v, rho, P, type back to back into the staging buffer
    double* st = s->stage;
    HIPC(hipMemcpyAsync(st, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    if (v) HIPC(hipMemcpyAsync(st + 3 * n, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    if (rho) HIPC(hipMemcpyAsync(st + 6 * n, rho, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    if (P) HIPC(hipMemcpyAsync(st + 7 * n, P, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    if (type) HIPC(hipMemcpyAsync(st + 8 * n, type, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    enqueue_append(s, (int)s->n, (int)n, st, v != nullptr, rho != nullptr, P != nullptr,
                   type != nullptr);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    s->n += n;
    s->cells_valid = s->masks_valid = false;
    return 0;
}

int sph_count(sph_system* s, int64_t* n)
{
    if (!s || !n) return fail(SPH_EINVAL, "null argument");
    *n = s->n;
    return 0;
}

int sph_set_field(sph_system* s, int field, const double* host, int64_t n)
{
    if (!s || (!host && n > 0)) return fail(SPH_EINVAL, "null argument");
    int nc;
    const int col = col_of_field(field, &nc);
    if (col < 0) return fail(SPH_EINVAL, "unknown field");
    if (n != s->n) return fail(SPH_EINVAL, "field length does not match the particle count");
    if (n == 0) return 0;
    HIPC(hipMemcpyAsync(s->stage, host, sizeof(double) * nc * n, hipMemcpyHostToDevice, s->stream));
    enqueue_field_in(s, col, nc, s->stage);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    if (field == SPH_FIELD_X) s->cells_valid = s->masks_valid = false;
    return 0;
}

int sph_get_field(sph_system* s, int field, double* host, int64_t n)
{
    if (!s || (!host && n > 0)) return fail(SPH_EINVAL, "null argument");
    int nc;
    const int col = col_of_field(field, &nc);
    if (col < 0) return fail(SPH_EINVAL, "unknown field");
    if (n != s->n) return fail(SPH_EINVAL, "field length does not match the particle count");
    if (n == 0) return 0;
    enqueue_field_out(s, col, nc, s->stage);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(host, s->stage, sizeof(double) * nc * n, hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

int sph_build_cells(sph_system* s, int64_t* n_out)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    int rc;
    if (s->cap == 0 && (rc = ensure_capacity(s, 1))) return rc;
    hipEvent_t ev[2];
    prof_begin(s, SPH_KERNEL_BUILD, ev);
    enqueue_build(s);
    prof_end(s, SPH_KERNEL_BUILD, ev);
    HIPC(hipGetLastError());
    if ((rc = sync_meta(s))) return rc;
    if (s->tile_auto && s->meta_host[META_NCELLS] > 0) {
        const int T = choose_tile((double)s->n / s->meta_host[META_NCELLS]);
        if (T != s->tile_T) {
            s->tile_T = T;
            enqueue_tiles(s);
            HIPC(hipGetLastError());
            if ((rc = sync_meta(s))) return rc;
        }
    }
    if ((rc = ensure_masks(s))) return rc;
    s->cells_valid = true;
    s->masks_valid = false;
    if (n_out) *n_out = s->n;
    return 0;
}

int sph_apply_binary(sph_system* s, int op, const sph_params* params, int self)
{
    if (!s || !params) return fail(SPH_EINVAL, "null argument");
    if (!is_binary(op)) return fail(SPH_EUNSUP, "not a registered binary operator");
    if (!s->cells_valid)
        return fail(SPH_ESTATE, "cell list is not current: call sph_build_cells after moving or adding particles");
    if (s->n == 0) return 0;
    const Phys ph = make_phys(s, params);
    int pop = P_DENSITY, kid = SPH_KERNEL_DENSITY;
    switch (op) {
    case SPH_OP_DENSITY_SUM: pop = P_DENSITY; break;
    case SPH_OP_DENSITY_SUM_RHO0: pop = P_DENSITY_RHO0; break;
    case SPH_OP_INTERNAL_FORCE:
        pop = P_FORCE; kid = SPH_KERNEL_FORCE;
        enqueue_prho2(s);
        break;
    case SPH_OP_PRESSURE_FORCE_RHO0: pop = P_FORCE_RHO0; kid = SPH_KERNEL_FORCE; break;
    case SPH_OP_BALANCE_OF_MASS: pop = P_CONTINUITY; kid = SPH_KERNEL_FORCE; break;
    }
    // self = true adds op(p, p, 0.0) (src/core.jl:155-157): m W(h, 0) for the
    // density sums; for the other operators x_pp = v_pp = 0 and rho_p - rho_p
    // = 0, so the term is zero.
    const int dens = pop == P_DENSITY || pop == P_DENSITY_RHO0;
    // A density pass records the pair masks of this cell list; the passes
    // after it read them.
    const int mode = dens ? MASK_TEST_STORE : (s->masks_valid ? MASK_LOAD : MASK_TEST);
    int rc = run_pair(s, pop, ph, dens && self, 0, nullptr, s->meta_host[META_NTILES], kid, mode);
    if (rc) return rc;
    HIPC(hipStreamSynchronize(s->stream));
    if (dens) s->masks_valid = s->use_masks && s->masks != nullptr;
    return 0;
}

int sph_apply_unary(sph_system* s, int op, const sph_params* params)
{
    if (!s || !params) return fail(SPH_EINVAL, "null argument");
    if (!is_unary(op)) return fail(SPH_EUNSUP, "not a registered unary operator");
    const Phys ph = make_phys(s, params);
    hipEvent_t ev[2];
    prof_begin(s, SPH_KERNEL_UNARY, ev);
    enqueue_unary(s, op, ph, nullptr);
    prof_end(s, SPH_KERNEL_UNARY, ev);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    if (op == SPH_OP_MOVE) s->cells_valid = s->masks_valid = false;
    return 0;
}

int sph_apply(sph_system* s, int op, const sph_params* params, int self)
{
    if (is_binary(op)) return sph_apply_binary(s, op, params, self);
    if (is_unary(op)) return sph_apply_unary(s, op, params);
    return fail(SPH_EUNSUP, "unknown operator id");
}

int sph_reduce(sph_system* s, int what, const sph_params* params, double* out)
{
    if (!s || !params || !out) return fail(SPH_EINVAL, "null argument");
    if (what != SPH_REDUCE_ENERGY && what != SPH_REDUCE_ENERGY_COLLISION &&
        what != SPH_REDUCE_MOMENTUM && what != SPH_REDUCE_ENERGY_RHO0)
        return fail(SPH_EUNSUP, "unknown reduction");
    const Phys ph = make_phys(s, params);
    double res[3] = {0.0, 0.0, 0.0};
    if (s->n > 0) {
        enqueue_reduce(s, what, ph, s->red_out);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(res, s->red_out, sizeof(res), hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
    }
    const int nv = what == SPH_REDUCE_MOMENTUM ? 3 : 1;
    for (int i = 0; i < nv; ++i) out[i] = res[i];
    return 0;
}

int sph_neighbour_pairs(sph_system* s, int64_t* pairs, int64_t cap, int64_t* n)
{
    if (!s || !n) return fail(SPH_EINVAL, "null argument");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    *n = 0;
    if (s->n == 0) return 0;
    sph_params zero{};
    const Phys ph = make_phys(s, &zero);
    unsigned long long total = 0;
    HIPC(hipMemsetAsync(s->pair_counter, 0, sizeof(unsigned long long), s->stream));
    int rc = run_pair(s, P_COUNT, ph, 0, 0, s->pair_counter, s->meta_host[META_NTILES],
                      SPH_KERNEL_DENSITY, MASK_TEST);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(&total, s->pair_counter, sizeof(total), hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    *n = (int64_t)total;
    if (!pairs || cap <= 0 || total == 0) return 0;
    const long long ncopy = std::min<long long>(cap, (long long)total);
    long long* dpairs = nullptr;
    if ((rc = dalloc(&dpairs, 2 * ncopy))) return rc;
    HIPC(hipMemsetAsync(s->pair_counter, 0, sizeof(unsigned long long), s->stream));
    PairCtx c = make_ctx(s, ph);
    c.emit_counter = s->pair_counter;
    c.pair_out = dpairs;
    c.pair_cap = ncopy;
    reset_queue(s);
    launch_pair(P_EMIT, s->precision, c, pair_grid(s, s->meta_host[META_NTILES]), s->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(pairs, dpairs, sizeof(long long) * 2 * ncopy, hipMemcpyDeviceToHost,
                           s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    hipFree(dpairs);
    if (e != hipSuccess) return fail(SPH_EHIP, std::string("pair emission: ") + hipGetErrorString(e));
    return 0;
}

int sph_step(sph_system* s, int scheme, const sph_params* params, int nsteps, int64_t* pairs)
{
    if (!s || !params) return fail(SPH_EINVAL, "null argument");
    if (scheme != SPH_SCHEME_WCSPH_SE && scheme != SPH_SCHEME_WCSPH_SE_RHO0)
        return fail(SPH_EUNSUP, "unknown scheme");
    const int eos = scheme == SPH_SCHEME_WCSPH_SE ? SPH_OP_FIND_PRESSURE : SPH_OP_FIND_PRESSURE_RHO0;
    if (nsteps < 0) return fail(SPH_EINVAL, "negative step count");
    int rc;
    if (s->cap == 0 && (rc = ensure_capacity(s, 1))) return rc;
    sph_params p1 = *params;
    p1.coeff = 1.0;
    const Phys ph = make_phys(s, &p1);
    // Tiles are taken from a work queue: the grid needs no tile count.
    const long long tiles_bound = s->tile_cap;
    HIPC(hipMemsetAsync(s->pair_counter, 0, 2 * sizeof(unsigned long long), s->stream));
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t ev[2];
        prof_begin(s, SPH_KERNEL_BUILD, ev);
        enqueue_build(s);
        prof_end(s, SPH_KERNEL_BUILD, ev);
        if ((rc = run_pair(s, P_DENSITY, ph, 1, 1, s->pair_counter, tiles_bound,
                           SPH_KERNEL_DENSITY, MASK_TEST_STORE)))
            return rc;
        prof_begin(s, SPH_KERNEL_UNARY, ev);
        enqueue_unary(s, eos, ph, s->prho2);
        prof_end(s, SPH_KERNEL_UNARY, ev);
        if ((rc = run_pair(s, P_FORCE, ph, 0, 1, s->pair_counter + 1, tiles_bound,
                           SPH_KERNEL_FORCE, MASK_LOAD)))
            return rc;
        prof_begin(s, SPH_KERNEL_UNARY, ev);
        enqueue_kick_drift(s, ph);
        prof_end(s, SPH_KERNEL_UNARY, ev);
        HIPC(hipGetLastError());
    }
    HIPC(hipMemcpyAsync(s->step_pairs, s->pair_counter, 2 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s->stream));
    if ((rc = sync_meta(s))) return rc;
    s->cells_valid = s->masks_valid = false;  // positions moved after the last build
    if ((rc = ensure_masks(s))) return rc;    // size the masks for the next call
    if (pairs) *pairs = (int64_t)s->step_pairs[0];
    return 0;
}

int sph_step_pairs(sph_system* s, int64_t* density_pairs, int64_t* force_pairs)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    if (density_pairs) *density_pairs = (int64_t)s->step_pairs[0];
    if (force_pairs) *force_pairs = (int64_t)s->step_pairs[1];
    return 0;
}

int sph_key_geometry(sph_system* s, int64_t* key_phase, int64_t* key_lim, int64_t* key_max,
                     int64_t* key_diff, int* n_diff)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    const Geo& g = s->geo;
    const long long lim[3] = {g.L1, g.L2, g.L3};
    for (int d = 0; d < 3; ++d) {
        if (key_phase) key_phase[d] = g.kp[d];
        if (key_lim) key_lim[d] = lim[d];
    }
    if (key_max) *key_max = g.key_max;
    // src/structs.jl:70-82, di outermost
    int nd = 0;
    long long diff[27];
    if (g.dim == 2) {
        for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj) diff[nd++] = di + lim[0] * dj;
    } else {
        for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj)
                for (int dk = -1; dk <= 1; ++dk) diff[nd++] = di + lim[0] * (dj + lim[1] * dk);
    }
    if (key_diff)
        for (int i = 0; i < nd; ++i) key_diff[i] = diff[i];
    if (n_diff) *n_diff = nd;
    return 0;
}

int sph_set_tile(sph_system* s, int T)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    if (T == 0) {
        s->tile_auto = 1;
        return 0;
    }
    if (T != 8 && T != 16 && T != 32 && T != 64) return fail(SPH_EINVAL, "tile must be 8, 16, 32 or 64");
    s->tile_auto = 0;
    if (T != s->tile_T) {
        s->tile_T = T;
        s->masks_valid = false;
        if (s->cells_valid) {
            enqueue_tiles(s);
            HIPC(hipGetLastError());
            int rc = sync_meta(s);
            if (!rc) rc = ensure_masks(s);
            if (rc) return rc;
        }
    }
    return 0;
}

int sph_autotune(sph_system* s, const sph_params* params, int* best)
{
    if (!s || !params) return fail(SPH_EINVAL, "null argument");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    // Time the density and force passes for every tile size.  A zero mass
    // makes both passes add exactly zero to rho and a, so the state is left
    // as it was while the full pair arithmetic still runs.
    sph_params p0 = *params;
    p0.m = 0.0;
    const Phys ph = make_phys(s, &p0);
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    float best_ms = 1e30f;
    int best_T = s->tile_T;
    const int cand[4] = {8, 16, 32, 64};
    enqueue_prho2(s);
    for (int i = 0; i < 4; ++i) {
        int rc = sph_set_tile(s, cand[i]);
        if (rc) return rc;
        float tot = 0.0f;
        for (int rep = 0; rep < 3; ++rep) {
            HIPC(hipEventRecord(e0, s->stream));
            if ((rc = run_pair(s, P_DENSITY, ph, 0, 0, nullptr, s->meta_host[META_NTILES],
                               SPH_KERNEL_DENSITY, MASK_TEST_STORE)))
                return rc;
            if ((rc = run_pair(s, P_FORCE, ph, 0, 0, nullptr, s->meta_host[META_NTILES],
                               SPH_KERNEL_FORCE, MASK_LOAD)))
                return rc;
            HIPC(hipEventRecord(e1, s->stream));
            HIPC(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPC(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) tot += ms;
        }
        if (tot < best_ms) {
            best_ms = tot;
            best_T = cand[i];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    int rc = sph_set_tile(s, best_T);
    if (rc) return rc;
    s->masks_valid = false;  // the timing passes ran with a zero mass
    if (best) *best = best_T;
    return 0;
}

int sph_set_option(sph_system* s, int option, int value)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    switch (option) {
    case SPH_OPTION_PAIR_MASKS:
        s->use_masks = value != 0;
        s->masks_valid = false;
        if (s->use_masks && s->cells_valid) return ensure_masks(s);
        return 0;
    }
    return fail(SPH_EINVAL, "unknown option");
}

int sph_candidates(sph_system* s, int64_t* all, int64_t* fluid)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    unsigned long long res[2] = {0, 0};
    if (s->n > 0) {
        sph_params zero{};
        PairCtx c = make_ctx(s, make_phys(s, &zero));
        HIPC(hipMemsetAsync(s->pair_counter, 0, 2 * sizeof(unsigned long long), s->stream));
        launch_candidates(c, s->icols[s->cur] + I_KEY * s->cap, s->pair_counter, s->n, s->stream);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(res, s->pair_counter, sizeof(res), hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
    }
    if (all) *all = (int64_t)res[0];
    if (fluid) *fluid = (int64_t)res[1];
    return 0;
}

int sph_profile(sph_system* s, int enable)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    prof_collect(s);
    s->prof = enable != 0;
    std::memset(s->prof_ms, 0, sizeof(s->prof_ms));
    std::memset(s->prof_n, 0, sizeof(s->prof_n));
    return 0;
}

int sph_profile_read(sph_system* s, int kernel, double* total_ms, int64_t* launches)
{
    if (!s || kernel < 0 || kernel > 3) return fail(SPH_EINVAL, "invalid argument");
    prof_collect(s);
    if (total_ms) *total_ms = s->prof_ms[kernel];
    if (launches) *launches = s->prof_n[kernel];
    return 0;
}

}  // extern "C"
