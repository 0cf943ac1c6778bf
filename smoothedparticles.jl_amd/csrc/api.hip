// C ABI of libsph_hip (include/sph.h, include/sph_batch.h): handle lifetime,
// device memory, the systems of a batch, operator dispatch by registered id,
// synchronisation and error reporting.
//
// A handle is a batch of systems.  The single-system functions of sph.h act
// on slot 0 of a handle made by sph_create; sph_batch.h adds and retires
// systems of a handle made by sph_batch_create.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "sph_batch.h"
#include "sph_internal.h"
#include "sph_setup.h"

using namespace sph;

namespace sph {
void enqueue_build(sph_system* s, unsigned skip_cols = 0);
void enqueue_tiles(sph_system* s);
void enqueue_unary(sph_system* s, int op, double* prho2);
void enqueue_prho2(sph_system* s);
void enqueue_kick_drift(sph_system* s);
void enqueue_cd_unary(sph_system* s, int stage);
void enqueue_reduce(sph_system* s, int sys, long long off, long long n, int what, double* out_dev);
void enqueue_probe_sum(sph_system* s, int sys, const double* probes_dev, long long np, double m,
                       int fluid_only, double* out_dev);
void enqueue_field_out(sph_system* s, long long off, long long n, int col0, int ncomp,
                       double* out_dev);
void enqueue_key_out(sph_system* s, long long off, long long n, int cell_base, double* out_dev);
void enqueue_particle_sum_lj(sph_system* s, int sys, long long off, long long n,
                             const double* fp, double* out_dev);
void enqueue_field_in(sph_system* s, long long off, long long n, int col0, int ncomp,
                      const double* in_dev);
void enqueue_append(sph_system* s, int n0, int m, int ref0, int sys, const double* in_dev,
                    int has_v, int has_rho, int has_P, int has_type);
void enqueue_set_int(sph_system* s, int* p, int v);
void enqueue_set_meta(sph_system* s, int idx, int v);
long long lattice_plan(const sph_lattice_spec* spec, LatticeGeo* g, std::vector<double>* ball);
void enqueue_lattice_flags(sph_system* s, const LatticeGeo& g, const double* ball, long long nlat,
                           unsigned* flags, unsigned* pos, unsigned* part);
void enqueue_lattice_write(sph_system* s, const LatticeGeo& g, const sph_lattice_spec* spec,
                           const unsigned* flags, const unsigned* pos, long long nlat, int slot,
                           int ref0);
void enqueue_fill_int(sph_system* s, int* p, int v, long long n);
void enqueue_remap_sys(sph_system* s, const int* map_dev);
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

// Cell geometry of one system.  Reference systems follow src/structs.jl:63-82.
// Periodic systems (not in the reference) split each periodic axis of the
// box into L = floor(box/h) >= 3 cells of length box/L >= h, so the 3-cell
// stencil holds every minimum-image neighbour exactly once.
int make_geo(double h, const double* bmin, const double* bmax, int periodic, Geo* out)
{
    Geo g{};
    if (!(h > 0.0) || !std::isfinite(h))
        return fail(SPH_EINVAL, "invalid ParticleSystem declaration! (h must be a positive float)");
    g.h = h;
    g.inv_h = 1.0 / h;
    g.s_thr = sqrt_threshold(h);
    for (int d = 0; d < 3; ++d) {
        if (!std::isfinite(bmin[d]) || !std::isfinite(bmax[d]) || bmin[d] > bmax[d])
            return fail(SPH_EINVAL, "invalid bounding box");
        g.bmin[d] = bmin[d];
        g.bmax[d] = bmax[d];
    }
    g.flat = bmin[2] == bmax[2];
    long long lim[3];
    if (!periodic) {
        for (int d = 0; d < 3; ++d) {
            // src/structs.jl:66-67
            const double fmin = std::floor(bmin[d] / h), fmax = std::floor(bmax[d] / h);
            if (std::fabs(fmin) > 1e15 || std::fabs(fmax) > 1e15)
                return fail(SPH_EINVAL, "bounding box too large for h");
            g.kp[d] = (long long)fmin;
            lim[d] = (long long)fmax - g.kp[d] + 1;
        }
    } else {
        const int axes = g.flat ? 2 : 3;
        for (int d = 0; d < 3; ++d) {
            g.kp[d] = 0;
            if (d >= axes) {
                lim[d] = 1;
                g.box[d] = 0.0;
                g.cell[d] = 1.0;
                continue;
            }
            const double box = bmax[d] - bmin[d];
            double L = std::floor(box / h);
            while (L >= 1.0 && box / L < h) L -= 1.0;
            if (!(L >= 3.0) || L > 1e9)
                return fail(SPH_EINVAL, "a periodic box needs at least 3 h along every periodic axis");
            lim[d] = (long long)L;
            g.box[d] = box;
            g.cell[d] = box / L;
        }
        g.periodic = 1;
    }
    const long long km = lim[0] * lim[1] * lim[2];
    if (km <= 0 || km > (1ll << 30)) return fail(SPH_EINVAL, "cell grid too large (key_max > 2^30)");
    g.L1 = (int)lim[0];
    g.L2 = (int)lim[1];
    g.L3 = (int)lim[2];
    g.key_max = (int)km;
    g.dim = lim[2] == 1 ? 2 : 3;  // src/structs.jl:70
    *out = g;
    return 0;
}

// Wendland quintic constants of src/kernels.jl:108-195.
Phys make_phys(const Geo& g, const sph_params* p)
{
    Phys ph{};
    if (p) {
        ph.m = p->m; ph.rho0 = p->rho0; ph.c = p->c; ph.alpha = p->alpha; ph.hs = p->hs;
        ph.mu = p->mu; ph.nu = p->nu;
        ph.g[0] = p->g[0]; ph.g[1] = p->g[1]; ph.g[2] = p->g[2];
        ph.dt = p->dt; ph.coeff = p->coeff;
    }
    const double h = g.h;
    if (g.dim == 2) {
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

// Device allocation with the handle's byte count kept current.
template <class T>
int dalloc(sph_system* s, T** p, long long n)
{
    *p = nullptr;
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(SPH_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    s->bytes += (long long)sizeof(T) * n;
    return 0;
}

template <class T>
void dfree(sph_system* s, T*& p, long long n)
{
    if (p) {
        (void)hipFree((void*)p);
        s->bytes -= (long long)sizeof(T) * (n > 0 ? n : 1);
    }
    p = nullptr;
}

int nsys(const sph_system* s) { return (int)s->sys_host.size(); }

void free_particle_arrays(sph_system* s)
{
    const long long cap = s->cap;
    for (int b = 0; b < 2; ++b) {
        dfree(s, s->cols[b], NCOL * cap);
        dfree(s, s->icols[b], NICOL * cap);
    }
    dfree(s, s->slot, cap); dfree(s, s->tmp_idx, cap); dfree(s, s->tmp_key, cap);
    dfree(s, s->tmp_ref, cap); dfree(s, s->perm, cap); dfree(s, s->inv, cap);
    dfree(s, s->tiles, s->tile_cap); dfree(s, s->kept_n, s->tile_cap);
    dfree(s, s->boxes, 6 * s->tile_cap);
    dfree(s, s->prho2, cap); dfree(s, s->stage, 9 * cap);
    dfree(s, s->red_part, s->red_part_n);
}

// Tiles of a cell list: sum over cells of ceil(count/T) <= n/8 + non-empty cells.
long long tile_bound(const sph_system* s)
{
    return s->cap / 8 + std::min<long long>(s->cap, std::max<long long>(s->cells_cap, 1)) + 2;
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
    int rc = 0;
    for (int b = 0; b < 2 && !rc; ++b) {
        if (!(rc = dalloc(s, &nc[b], NCOL * ncap))) rc = dalloc(s, &ni[b], NICOL * ncap);
    }
    if (rc) {
        for (int b = 0; b < 2; ++b) {
            dfree(s, nc[b], NCOL * ncap);
            dfree(s, ni[b], NICOL * ncap);
        }
        return rc;
    }
    if (s->cap > 0 && s->n > 0) {
        for (int f = 0; f < NCOL; ++f)
            HIPC(hipMemcpyAsync(nc[s->cur] + f * ncap, s->cols[s->cur] + f * s->cap,
                                sizeof(double) * s->n, hipMemcpyDeviceToDevice, s->stream));
        for (int f = 0; f < NICOL; ++f)
            HIPC(hipMemcpyAsync(ni[s->cur] + f * ncap, s->icols[s->cur] + f * s->cap,
                                sizeof(int) * s->n, hipMemcpyDeviceToDevice, s->stream));
    }
    HIPC(hipStreamSynchronize(s->stream));
    free_particle_arrays(s);
    s->cols[0] = nc[0]; s->cols[1] = nc[1];
    s->icols[0] = ni[0]; s->icols[1] = ni[1];
    s->cap = ncap;
    s->tile_cap = tile_bound(s);
    s->red_part_n = 3 * reduce_partials_needed(ncap);
    if ((rc = dalloc(s, &s->slot, ncap)) || (rc = dalloc(s, &s->tmp_idx, ncap)) ||
        (rc = dalloc(s, &s->tmp_key, ncap)) || (rc = dalloc(s, &s->tmp_ref, ncap)) ||
        (rc = dalloc(s, &s->perm, ncap)) || (rc = dalloc(s, &s->inv, ncap)) ||
        (rc = dalloc(s, &s->tiles, s->tile_cap)) || (rc = dalloc(s, &s->kept_n, s->tile_cap)) ||
        (rc = dalloc(s, &s->boxes, 6 * s->tile_cap)) || (rc = dalloc(s, &s->prho2, ncap)) ||
        (rc = dalloc(s, &s->stage, 9 * ncap)) || (rc = dalloc(s, &s->red_part, s->red_part_n)))
        return rc;
    return 0;
}

void free_cell_arrays(sph_system* s)
{
    const long long c = s->cells_cap;
    dfree(s, s->cnt, c); dfree(s, s->cell_start, c + 1); dfree(s, s->tile_start, c + 1);
    dfree(s, s->cell_chunks, c); dfree(s, s->mask_start, c + 1); dfree(s, s->cell_sys, c);
    dfree(s, s->cell_order, c);
    dfree(s, s->scan_part, s->scan_part_n); dfree(s, s->scan_part64, s->scan_part_n);
}

// Per-cell arrays for `buckets` buckets (the cells of every system, one
// removal bucket per system and one bucket for retired systems); the
// cell -> system map is refilled from the host table.
int ensure_cells(sph_system* s, long long buckets)
{
    if (buckets <= s->cells_cap) return 0;
    const long long ncap = std::max<long long>(buckets, s->cells_cap + s->cells_cap / 2);
    if (ncap >= (1ll << 31) - 64) return fail(SPH_EINVAL, "too many cells in the batch");
    free_cell_arrays(s);
    s->cells_cap = ncap;
    s->scan_part_n = scan_partials_needed(ncap + 2);
    int rc;
    if ((rc = dalloc(s, &s->cnt, ncap)) || (rc = dalloc(s, &s->cell_start, ncap + 1)) ||
        (rc = dalloc(s, &s->tile_start, ncap + 1)) || (rc = dalloc(s, &s->cell_chunks, ncap)) ||
        (rc = dalloc(s, &s->mask_start, ncap + 1)) || (rc = dalloc(s, &s->cell_sys, ncap)) ||
        (rc = dalloc(s, &s->cell_order, ncap)) || (rc = dalloc(s, &s->scan_part, s->scan_part_n)) ||
        (rc = dalloc(s, &s->scan_part64, s->scan_part_n))) {
        s->cells_cap = 0;
        return rc;
    }
    if (s->cap > 0) {  // the tile bound grows with the cell count
        dfree(s, s->tiles, s->tile_cap);
        dfree(s, s->kept_n, s->tile_cap);
        dfree(s, s->boxes, 6 * s->tile_cap);
        s->tile_cap = tile_bound(s);
        if ((rc = dalloc(s, &s->tiles, s->tile_cap)) || (rc = dalloc(s, &s->kept_n, s->tile_cap)) ||
            (rc = dalloc(s, &s->boxes, 6 * s->tile_cap)))
            return rc;
    }
    for (int i = 0; i < nsys(s); ++i)
        enqueue_fill_int(s, s->cell_sys + s->sys_host[i].cell_base, i, s->sys_host[i].g.key_max);
    HIPC(hipGetLastError());
    return 0;
}

// Copy the host table of systems to the device, growing the device table
// (and the per-system counters, restarted from the host mirror) if needed.
int upload_sysd(sph_system* s)
{
    const int n = nsys(s);
    if (n > s->sysd_cap) {
        const long long cap = std::max<long long>(16, 2 * (long long)n);
        SysDesc* nd = nullptr;
        int* nc = nullptr;
        int* nh = nullptr;
        int rc;
        if ((rc = dalloc(s, &nd, cap))) return rc;
        if ((rc = dalloc(s, &nc, SYSCNT * cap))) {
            dfree(s, nd, cap);
            return rc;
        }
        hipError_t e = hipHostMalloc((void**)&nh, sizeof(int) * SYSCNT * cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            dfree(s, nd, cap);
            dfree(s, nc, SYSCNT * cap);
            return fail(SPH_EHIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        }
        std::memset(nh, 0, sizeof(int) * SYSCNT * cap);
        if (s->sys_cnt_cap > 0) {
            // keep the device counters of the existing systems
            HIPC(hipMemcpyAsync(nc, s->sys_cnt, sizeof(int) * SYSCNT * s->sys_cnt_cap,
                                hipMemcpyDeviceToDevice, s->stream));
            std::memcpy(nh, s->sys_cnt_host, sizeof(int) * SYSCNT * s->sys_cnt_cap);
        }
        HIPC(hipStreamSynchronize(s->stream));
        dfree(s, s->sysd, s->sysd_cap);
        dfree(s, s->sys_cnt, SYSCNT * s->sys_cnt_cap);
        if (s->sys_cnt_host) (void)hipHostFree(s->sys_cnt_host);
        s->sysd = nd;
        s->sys_cnt = nc;
        s->sys_cnt_host = nh;
        s->sysd_cap = cap;
        s->sys_cnt_cap = cap;
    }
    if (n > 0)
        HIPC(hipMemcpyAsync(s->sysd, s->sys_host.data(), sizeof(SysDesc) * n,
                            hipMemcpyHostToDevice, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

// Host mirrors of the device counters after a build.
int sync_meta(sph_system* s)
{
    HIPC(hipMemcpyAsync(s->meta_host, s->meta, sizeof(int) * META_SLOTS, hipMemcpyDeviceToHost,
                        s->stream));
    if (nsys(s) > 0)
        HIPC(hipMemcpyAsync(s->sys_cnt_host, s->sys_cnt, sizeof(int) * SYSCNT * nsys(s),
                            hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    s->n = s->meta_host[META_N];
    for (int i = 0; i < nsys(s); ++i) {
        s->sys_n[i] = s->sys_cnt_host[SYSCNT * i + SC_N];
        s->sys_off[i] = s->sys_cnt_host[SYSCNT * i + SC_OFF];
    }
    return 0;
}

void prof_begin(sph_system* s, hipEvent_t* ev)
{
    if (!s->prof) return;
    for (int i = 0; i < 2; ++i) {
        if (s->ev_pool.empty()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            s->ev_pool.push_back(e);
        }
        ev[i] = s->ev_pool.back();
        s->ev_pool.pop_back();
    }
    (void)hipEventRecord(ev[0], s->stream);
}

void prof_end(sph_system* s, int k, hipEvent_t* ev)
{
    if (!s->prof) return;
    (void)hipEventRecord(ev[1], s->stream);
    s->ev_open.push_back({k, {ev[0], ev[1]}});
}

void prof_collect(sph_system* s)
{
    if (s->ev_open.empty()) return;
    (void)hipStreamSynchronize(s->stream);
    for (auto& e : s->ev_open) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e.second.first, e.second.second);
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

PairCtx make_ctx(sph_system* s)
{
    PairCtx c{};
    c.sys = s->sysd;
    c.cell_sys = s->cell_sys;
    c.T = s->tile_T;
    c.cap = (int)s->cap;
    c.tiles = s->tiles;
    c.cell_start = s->cell_start;
    c.meta = s->meta;
    c.queue = s->queue;
    c.qgrab = s->rows == 9 ? 1 : 4;
    c.cols = s->cols[s->cur];
    c.refidx = s->icols[s->cur] + I_REF * s->cap;
    c.prho2 = s->prho2;
    c.masks = s->masks;
    c.kept = s->use_kept ? s->kept : nullptr;
    c.kept_rg = s->kept_rg;
    c.kept_n = s->kept_n;
    c.boxes = s->boxes;
    c.mask_mode = MASK_TEST;
    c.masks_mixed = s->precision == SPH_MIXED;
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
        dfree(s, s->masks, s->mask_cap);
        s->mask_cap = 0;
        const long long cap = (long long)need + (long long)need / 4 + 4096;
        if (dalloc(s, &s->masks, cap)) {
            s->masks = nullptr;  // no masks: the kernels keep testing
            g_err.clear();
            return 0;
        }
        s->mask_cap = cap;
    }
    // kept-candidate lists: 64 entries per chunk of a tile (T mask words)
    // (with the candidates' ranges when a system is periodic)
    const long long kneed = (long long)(need / s->tile_T * 64);
    if (kneed > s->kept_cap || (s->periodic && !s->kept_rg)) {
        dfree(s, s->kept, s->kept_cap);
        dfree(s, s->kept_rg, s->kept_cap);
        s->kept_cap = 0;
        const long long cap = kneed + kneed / 4 + 4096;
        if (dalloc(s, &s->kept, cap) || (s->periodic && dalloc(s, &s->kept_rg, cap))) {
            dfree(s, s->kept, cap);  // no lists: the loading passes stream and cull
            s->kept = nullptr;
            s->kept_rg = nullptr;
            g_err.clear();
        } else {
            s->kept_cap = cap;
        }
    }
    const int ok[2] = {1, s->kept ? 1 : 0};
    HIPC(hipMemcpyAsync(s->meta + META_MASK_OK, ok, sizeof(ok), hipMemcpyHostToDevice, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

// Grid of the pair kernels: the tiles are taken from a work queue, so a
// grid that fills the chip is enough.
int pair_grid(long long ntiles)
{
    const long long g = std::min<long long>(ntiles, 256 * 32);
    return (int)std::max<long long>(g, 1);
}

int run_pair(sph_system* s, int op, int self, int overwrite, unsigned long long* counter,
             long long ntiles, int prof_kernel, int mask_mode, int only_flagged = 0,
             unsigned long long* tests = nullptr)
{
    PairCtx c = make_ctx(s);
    c.self = self;
    c.overwrite = overwrite;
    c.pair_counter = counter;
    c.test_counter = tests;
    c.mask_mode = s->use_masks ? mask_mode : MASK_TEST;
    c.only_flagged = only_flagged;
    hipEvent_t ev[2];
    prof_begin(s, ev);
    reset_queue(s);
    launch_pair(op, s->precision, s->rows, s->flat, s->periodic, c, pair_grid(ntiles), s->stream);
    prof_end(s, prof_kernel, ev);
    HIPC(hipGetLastError());
    return 0;
}

bool is_binary(int op) { return op >= SPH_OP_DENSITY_SUM && op <= SPH_OP_BALANCE_OF_MASS; }
bool is_unary(int op) { return op >= SPH_OP_RESET_RHO && op <= SPH_OP_KICK; }

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

int destroy(sph_system* s);

// A handle with its stream, counters and a first particle allocation, but
// no system yet.
int new_handle(int precision, sph_system** out)
{
    *out = nullptr;
    if (precision != SPH_FP64 && precision != SPH_MIXED)
        return fail(SPH_EINVAL, "unknown precision mode");
    sph_system* s = new sph_system();  // value-initialised: pointers null, counts 0
    s->precision = precision;
    s->tile_T = 32;
    s->tile_auto = 1;
    s->rows = 3;
    s->flat = 1;
    s->use_masks = true;
    s->use_kept = true;
    int rc = 0;
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        s->stream = nullptr;
        rc = fail(SPH_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (!rc && !(rc = dalloc(s, &s->meta, META_SLOTS)) && !(rc = dalloc(s, &s->pair_counter, 4)) &&
        !(rc = dalloc(s, &s->queue, NQUEUE * QSTRIDE)) &&
        !(rc = dalloc(s, &s->red_out, 4)) && !(rc = dalloc(s, &s->mask_total, 1))) {
        e = hipHostMalloc((void**)&s->meta_host, sizeof(int) * META_SLOTS, hipHostMallocDefault);
        if (e != hipSuccess) {
            s->meta_host = nullptr;
            rc = fail(SPH_EHIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        }
    }
    if (!rc) {
        std::memset(s->meta_host, 0, sizeof(int) * META_SLOTS);
        e = hipMemsetAsync(s->meta, 0, sizeof(int) * META_SLOTS, s->stream);
        if (e == hipSuccess)
            e = hipMemsetAsync(s->mask_total, 0, sizeof(unsigned long long), s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(SPH_EHIP, std::string("init: ") + hipGetErrorString(e));
    }
    if (!rc) rc = ensure_capacity(s, 1);
    if (!rc) rc = upload_sysd(s);
    if (rc) {
        const std::string msg = g_err;
        destroy(s);
        g_err = msg;
        return rc;
    }
    *out = s;
    return 0;
}

int destroy(sph_system* s)
{
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_particle_arrays(s);
    free_cell_arrays(s);
    dfree(s, s->meta, META_SLOTS); dfree(s, s->pair_counter, 4); dfree(s, s->red_out, 4);
    dfree(s, s->queue, NQUEUE * QSTRIDE);
    dfree(s, s->masks, s->mask_cap); dfree(s, s->mask_total, 1);
    dfree(s, s->kept, s->kept_cap); dfree(s, s->kept_rg, s->kept_cap);
    dfree(s, s->sysd, s->sysd_cap); dfree(s, s->sys_cnt, SYSCNT * s->sys_cnt_cap);
    if (s->meta_host) (void)hipHostFree(s->meta_host);
    if (s->sys_cnt_host) (void)hipHostFree(s->sys_cnt_host);
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
    for (auto& e : s->ev_open) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    if (s->copy_stream) {
        (void)hipStreamSynchronize(s->copy_stream);
        (void)hipStreamDestroy(s->copy_stream);
    }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

// Add a system at the end of the table; its particles follow.
int add_system(sph_system* s, const Geo& g, const Phys& ph, int* slot_out)
{
    const int rows = g.dim == 2 ? 3 : 9;
    if (nsys(s) > 0 && rows != s->rows)
        return fail(SPH_EINVAL, "all systems of a batch must have the same dimension");
    if (s->cells_total + g.key_max >= (1ll << 30))
        return fail(SPH_EINVAL, "too many cells in the batch");
    SysDesc d{};
    d.g = g;
    d.ph = ph;
    d.cell_base = (int)s->cells_total;
    d.live = 1;
    d.pass_on = 0;
    s->sys_host.push_back(d);
    s->sys_n.push_back(0);
    s->sys_off.push_back(s->n);
    s->cells_total += g.key_max;
    const int slot = nsys(s) - 1;
    s->id_of_slot.push_back((int)s->slot_of_id.size());
    s->slot_of_id.push_back(slot);
    s->rho0_pending.push_back(1);
    int rc = upload_sysd(s);
    if (!rc) rc = ensure_cells(s, s->cells_total + nsys(s) + 1);
    if (rc) {  // leave the table as it was
        s->slot_of_id.back() = -1;
        s->sys_host.pop_back();
        s->sys_n.pop_back();
        s->sys_off.pop_back();
        s->id_of_slot.pop_back();
        s->rho0_pending.pop_back();
        s->cells_total -= g.key_max;
        return rc;
    }
    s->rows = rows;
    s->flat = slot == 0 ? g.flat : (s->flat && g.flat);
    s->periodic = slot == 0 ? g.periodic : (s->periodic || g.periodic);
    enqueue_fill_int(s, s->cell_sys + d.cell_base, slot, g.key_max);
    enqueue_set_int(s, s->sys_cnt + SYSCNT * slot + SC_N, 0);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    s->cells_valid = s->masks_valid = false;
    *slot_out = slot;
    return 0;
}

// Append n particles to system `slot`.  The device rows [off, off + n_sys)
// of a system stay contiguous: appends go to the end of the arrays, which
// is the end of the system when it is the last one holding particles.
int append_particles(sph_system* s, int slot, int64_t n, const double* x, const double* v,
                     const double* rho, const double* P, const double* type)
{
    if (n < 0 || (n > 0 && !x)) return fail(SPH_EINVAL, "invalid append arguments");
    if (n == 0) return 0;
    if (s->sys_n[slot] > 0 && s->sys_off[slot] + s->sys_n[slot] != s->n)
        return fail(SPH_EINVAL, "particles can only be appended to the last system of a batch");
    if ((long long)n > (1ll << 31) - 64 - s->n) return fail(SPH_EINVAL, "particle count exceeds 2^31");
    int rc;
    if ((rc = ensure_capacity(s, s->n + n))) return rc;
    // x, v, rho, P, type back to back in the staging buffer
    double* st = s->stage;
    HIPC(hipMemcpyAsync(st, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    if (v) HIPC(hipMemcpyAsync(st + 3 * n, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    if (rho) HIPC(hipMemcpyAsync(st + 6 * n, rho, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    if (P) HIPC(hipMemcpyAsync(st + 7 * n, P, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    if (type) HIPC(hipMemcpyAsync(st + 8 * n, type, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    enqueue_append(s, (int)s->n, (int)n, (int)s->sys_n[slot], slot, st, v != nullptr,
                   rho != nullptr, P != nullptr, type != nullptr);
    if (s->sys_n[slot] == 0) s->sys_off[slot] = s->n;
    s->sys_n[slot] += n;
    enqueue_set_int(s, s->sys_cnt + SYSCNT * slot + SC_N, (int)s->sys_n[slot]);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    s->n += n;
    s->cells_valid = s->masks_valid = false;
    return 0;
}

// Device-side generate_particles! of a lattice (sph_setup.h) into system
// `slot`: plan on the host, flag and scan on the device, one sync for the
// count, then the particles are written where sph_append would put them.
int generate_lattice(sph_system* s, int slot, const sph_lattice_spec* spec, int64_t* n_added)
{
    if (n_added) *n_added = 0;
    if (s->sys_n[slot] > 0 && s->sys_off[slot] + s->sys_n[slot] != s->n)
        return fail(SPH_EINVAL, "particles can only be appended to the last system of a batch");
    LatticeGeo g;
    std::vector<double> ball;
    const long long nlat = lattice_plan(spec, &g, &ball);
    if (nlat == -1) return fail(SPH_EINVAL, "invalid lattice specification");
    if (nlat < 0) return fail(SPH_EINVAL, "lattice covering above 2^31 points");
    if (nlat == 0) return 0;
    unsigned *flags = nullptr, *pos = nullptr, *part = nullptr;
    double* dball = nullptr;
    const long long nparts = scan_partials_needed(nlat);
    int rc = dalloc(s, &flags, nlat);
    if (!rc) rc = dalloc(s, &pos, nlat + 1);
    if (!rc) rc = dalloc(s, &part, nparts);
    if (!rc && !ball.empty()) rc = dalloc(s, &dball, (long long)ball.size());
    unsigned total = 0;
    hipError_t e = hipSuccess;
    if (!rc && dball)
        e = hipMemcpyAsync(dball, ball.data(), sizeof(double) * ball.size(), hipMemcpyHostToDevice,
                           s->stream);
    if (!rc && e == hipSuccess) {
        enqueue_lattice_flags(s, g, dball, nlat, flags, pos, part);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(&total, pos + nlat, sizeof(unsigned), hipMemcpyDeviceToHost,
                               s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    }
    if (!rc && e != hipSuccess) rc = fail(SPH_EHIP, std::string("lattice: ") + hipGetErrorString(e));
    if (!rc && (long long)total > (1ll << 31) - 64 - s->n)
        rc = fail(SPH_EINVAL, "particle count exceeds 2^31");
    if (!rc && total > 0) rc = ensure_capacity(s, s->n + total);
    if (!rc && total > 0) {
        enqueue_lattice_write(s, g, spec, flags, pos, nlat, slot, (int)s->sys_n[slot]);
        enqueue_set_meta(s, META_N, (int)(s->n + total));
        if (s->sys_n[slot] == 0) s->sys_off[slot] = s->n;
        s->sys_n[slot] += total;
        enqueue_set_int(s, s->sys_cnt + SYSCNT * slot + SC_N, (int)s->sys_n[slot]);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(SPH_EHIP, std::string("lattice: ") + hipGetErrorString(e));
        s->n += total;
        s->cells_valid = s->masks_valid = false;
        if (!rc && n_added) *n_added = total;
    }
    dfree(s, flags, nlat);
    dfree(s, pos, nlat + 1);
    dfree(s, part, nparts);
    dfree(s, dball, (long long)ball.size());
    return rc;
}

int build_cells(sph_system* s)
{
    int rc;
    if (nsys(s) == 0) {  // nothing to sort: no cell arrays yet
        s->n = 0;
        s->cells_valid = true;
        return 0;
    }
    hipEvent_t ev[2];
    prof_begin(s, ev);
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
    s->retired_since_build = false;
    return 0;
}

// Single-system calls carry their parameters: they become slot 0's.  Every
// caller synchronises the stream before returning, after which the host
// table may change again.
int set_params0(sph_system* s, const sph_params* p)
{
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    s->sys_host[0].ph = make_phys(s->sys_host[0].g, p);
    HIPC(hipMemcpyAsync(s->sysd, s->sys_host.data(), sizeof(SysDesc), hipMemcpyHostToDevice,
                        s->stream));
    return 0;
}

// The time loop of examples/collapse_dry.jl:202-211 (see k_cd_unary):
// accelerate! + move! (half steps); build; balance_of_mass!; continuity
// find_pressure!; move!; build; P/rho^2 in the new order; internal_force!;
// accelerate!.
int cd_step(sph_system* s, long long tiles_bound)
{
    int rc;
    hipEvent_t ev[2];
    prof_begin(s, ev);
    enqueue_cd_unary(s, CD_KICK_MOVE);
    prof_end(s, SPH_KERNEL_UNARY, ev);
    prof_begin(s, ev);
    enqueue_build(s);
    prof_end(s, SPH_KERNEL_BUILD, ev);
    if ((rc = run_pair(s, P_CONTINUITY, 0, 0, s->pair_counter, tiles_bound, SPH_KERNEL_DENSITY,
                       MASK_TEST, 0, s->pair_counter + 2)))
        return rc;
    prof_begin(s, ev);
    enqueue_unary(s, SPH_OP_FIND_PRESSURE_CONTINUITY, nullptr);
    enqueue_cd_unary(s, CD_MOVE);
    prof_end(s, SPH_KERNEL_UNARY, ev);
    prof_begin(s, ev);
    enqueue_build(s);
    prof_end(s, SPH_KERNEL_BUILD, ev);
    enqueue_prho2(s);  // the build permuted the particles
    if ((rc = run_pair(s, P_FORCE, 0, 0, s->pair_counter + 1, tiles_bound, SPH_KERNEL_FORCE,
                       MASK_TEST)))
        return rc;
    prof_begin(s, ev);
    enqueue_cd_unary(s, CD_KICK);
    prof_end(s, SPH_KERNEL_UNARY, ev);
    HIPC(hipGetLastError());
    return 0;
}

// One step of a density-summation scheme: build; density pass (stores the
// masks); EOS with P/rho^2; force pass (loads them); kick and drift.
int ds_step(sph_system* s, int eos, long long tiles_bound)
{
    int rc;
    hipEvent_t ev[2];
    prof_begin(s, ev);
    // the density pass overwrites rho of every particle in a cell and the
    // EOS then P of every particle, before anything reads them: the build
    // need not carry the two columns over
    enqueue_build(s, (1u << C_RHO) | (1u << C_P));
    prof_end(s, SPH_KERNEL_BUILD, ev);
    if ((rc = run_pair(s, P_DENSITY, 1, 1, s->pair_counter, tiles_bound, SPH_KERNEL_DENSITY,
                       MASK_TEST_STORE, 0, s->pair_counter + 2)))
        return rc;
    prof_begin(s, ev);
    enqueue_unary(s, eos, s->prho2);
    prof_end(s, SPH_KERNEL_UNARY, ev);
    if ((rc = run_pair(s, P_FORCE, 0, 1, s->pair_counter + 1, tiles_bound, SPH_KERNEL_FORCE,
                       MASK_LOAD)))
        return rc;
    prof_begin(s, ev);
    enqueue_kick_drift(s);
    prof_end(s, SPH_KERNEL_UNARY, ev);
    HIPC(hipGetLastError());
    return 0;
}

int one_step(sph_system* s, int scheme, int eos, long long tiles_bound)
{
    return scheme == SPH_SCHEME_COLLAPSE_DRY ? cd_step(s, tiles_bound)
                                             : ds_step(s, eos, tiles_bound);
}

// With SPH_OPTION_GRAPH the steps of a call are replayed from a hipGraph:
// every kernel of a step reads its sizes from device counters, and nothing
// the host enqueues changes while the call runs, so one capture serves the
// whole call.  A build swaps the two column buffers, so the graph holds
// steps enough for an even number of builds (two density-summation steps,
// one collapse_dry step) and leaves the buffers where it found them.  The
// pair counters are cleared before the replays, outside the graph.
int run_graph(sph_system* s, int scheme, int eos, long long tiles_bound, int per, int reps)
{
    const bool prof = s->prof;
    s->prof = false;  // no event records inside the capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = 0;
    hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        s->prof = prof;
        return fail(SPH_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    for (int k = 0; k < per && !rc; ++k) rc = one_step(s, scheme, eos, tiles_bound);
    e = hipStreamEndCapture(s->stream, &graph);
    s->prof = prof;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = hipGraphLaunch(exec, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(SPH_EHIP, std::string("step graph: ") + hipGetErrorString(e));
    return 0;
}

int run_steps(sph_system* s, int scheme, int nsteps)
{
    if (scheme != SPH_SCHEME_WCSPH_SE && scheme != SPH_SCHEME_WCSPH_SE_RHO0 &&
        scheme != SPH_SCHEME_COLLAPSE_DRY)
        return fail(SPH_EUNSUP, "unknown scheme");
    if (nsteps < 0) return fail(SPH_EINVAL, "negative step count");
    const int eos = scheme == SPH_SCHEME_WCSPH_SE ? SPH_OP_FIND_PRESSURE : SPH_OP_FIND_PRESSURE_RHO0;
    int rc;
    if (nsys(s) == 0) {
        s->step_pairs[0] = s->step_pairs[1] = s->step_pairs[2] = 0;
        return 0;
    }
    // Tiles are taken from a work queue: the grid needs no tile count.
    const long long tiles_bound = s->tile_cap;
    HIPC(hipMemsetAsync(s->pair_counter, 0, 3 * sizeof(unsigned long long), s->stream));
    int k = 0;
    const int per = scheme == SPH_SCHEME_COLLAPSE_DRY ? 1 : 2;
    if (s->use_graph && nsteps / per >= 2) {
        if ((rc = run_graph(s, scheme, eos, tiles_bound, per, nsteps / per))) return rc;
        k = nsteps / per * per;
    }
    for (; k < nsteps; ++k)
        if ((rc = one_step(s, scheme, eos, tiles_bound))) return rc;
    HIPC(hipMemcpyAsync(s->step_pairs, s->pair_counter, 3 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s->stream));
    if ((rc = sync_meta(s))) return rc;
    if (nsteps > 0) {
        s->retired_since_build = false;
        s->cells_valid = s->masks_valid = false;  // positions moved after the last build
    }
    return ensure_masks(s);  // size the masks for the next call
}

int slot_of(sph_system* s, int id, bool live_only = true)
{
    if (id < 0 || id >= (int)s->slot_of_id.size() || s->slot_of_id[id] < 0) return -1;
    const int slot = s->slot_of_id[id];
    if (live_only && !s->sys_host[slot].live) return -1;
    return slot;
}

// Drop retired systems from the table once their particles have left the
// cells (a build ran since they were retired): the live systems get new
// slots and cell ranges, and the system column of the particles is
// remapped.  The particle order is unchanged.
int compact(sph_system* s)
{
    const int n0 = nsys(s);
    bool any_dead = false;
    for (int i = 0; i < n0; ++i) any_dead |= !s->sys_host[i].live;
    if (!any_dead || s->retired_since_build) return 0;
    int rc;
    if ((rc = sync_meta(s))) return rc;
    std::vector<int> map(n0, 0);
    std::vector<SysDesc> nh;
    std::vector<long long> nn, no;
    std::vector<int> nid;
    std::vector<char> npend;
    long long cells = 0;
    for (int i = 0; i < n0; ++i) {
        if (!s->sys_host[i].live) continue;
        map[i] = (int)nh.size();
        SysDesc d = s->sys_host[i];
        d.cell_base = (int)cells;
        cells += d.g.key_max;
        nh.push_back(d);
        nn.push_back(s->sys_n[i]);
        no.push_back(s->sys_off[i]);
        nid.push_back(s->id_of_slot[i]);
        npend.push_back(s->rho0_pending[i]);
    }
    int* dmap = nullptr;
    if ((rc = dalloc(s, &dmap, n0))) return rc;
    hipError_t e = hipMemcpyAsync(dmap, map.data(), sizeof(int) * n0, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        enqueue_remap_sys(s, dmap);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(s, dmap, n0);
    if (e != hipSuccess) return fail(SPH_EHIP, std::string("compaction: ") + hipGetErrorString(e));
    for (int i = 0; i < n0; ++i)
        s->slot_of_id[s->id_of_slot[i]] = s->sys_host[i].live ? map[i] : -1;
    s->sys_host = nh;
    s->sys_n = nn;
    s->sys_off = no;
    s->id_of_slot = nid;
    s->rho0_pending = npend;
    s->cells_total = cells;
    s->flat = 1;
    s->periodic = 0;
    for (const SysDesc& d : s->sys_host) {
        s->flat = s->flat && d.g.flat;
        s->periodic = s->periodic || d.g.periodic;
    }
    if ((rc = upload_sysd(s))) return rc;
    for (int i = 0; i < nsys(s); ++i) {
        int* c = s->sys_cnt_host + SYSCNT * i;
        c[SC_NTOT] = c[SC_N] = (int)s->sys_n[i];
        c[SC_NREM] = 0;
        c[SC_OFF] = (int)s->sys_off[i];
        c[SC_REFBASE] = 0;
    }
    if (nsys(s) > 0)
        HIPC(hipMemcpyAsync(s->sys_cnt, s->sys_cnt_host, sizeof(int) * SYSCNT * nsys(s),
                            hipMemcpyHostToDevice, s->stream));
    for (int i = 0; i < nsys(s); ++i)
        enqueue_fill_int(s, s->cell_sys + s->sys_host[i].cell_base, i, s->sys_host[i].g.key_max);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    s->cells_valid = s->masks_valid = false;
    return 0;
}

int field_io(sph_system* s, int slot, int field, double* host, int64_t n, bool in)
{
    if (field == SPH_FIELD_KEY) {  // read only, from the last build
        if (in) return fail(SPH_EINVAL, "the cell-key field is read only");
        if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
        if (n != s->sys_n[slot]) return fail(SPH_EINVAL, "field length does not match the particle count");
        if (n == 0) return 0;
        enqueue_key_out(s, s->sys_off[slot], n, s->sys_host[slot].cell_base, s->stage);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(host, s->stage, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
        return 0;
    }
    int nc;
    const int col = col_of_field(field, &nc);
    if (col < 0) return fail(SPH_EINVAL, "unknown field");
    if (n != s->sys_n[slot]) return fail(SPH_EINVAL, "field length does not match the particle count");
    if (n == 0) return 0;
    const long long off = s->sys_off[slot];
    if (in) {
        HIPC(hipMemcpyAsync(s->stage, host, sizeof(double) * nc * n, hipMemcpyHostToDevice, s->stream));
        enqueue_field_in(s, off, n, col, nc, s->stage);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(s->stream));
        if (field == SPH_FIELD_X) s->cells_valid = s->masks_valid = false;
    } else {
        enqueue_field_out(s, off, n, col, nc, s->stage);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(host, s->stage, sizeof(double) * nc * n, hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
    }
    return 0;
}

int reduce_slot(sph_system* s, int slot, int what, double* out)
{
    if (what != SPH_REDUCE_ENERGY && what != SPH_REDUCE_ENERGY_COLLISION &&
        what != SPH_REDUCE_MOMENTUM && what != SPH_REDUCE_ENERGY_RHO0)
        return fail(SPH_EUNSUP, "unknown reduction");
    double res[3] = {0.0, 0.0, 0.0};
    if (s->sys_n[slot] > 0) {
        enqueue_reduce(s, slot, s->sys_off[slot], s->sys_n[slot], what, s->red_out);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(res, s->red_out, sizeof(res), hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
    }
    const int nv = what == SPH_REDUCE_MOMENTUM ? 3 : 1;
    for (int i = 0; i < nv; ++i) out[i] = res[i];
    return 0;
}

// Probe sums of system `slot` (SmoothedParticles.sum at points,
// src/core.jl:240-258).  The probe and result staging is allocated per call:
// the call runs at frame cadence, not in the step loop.
int probe_sum(sph_system* s, int slot, double m, int flags, const double* x, long long np,
              double* out)
{
    if (np < 0 || np > (1LL << 28)) return fail(SPH_EINVAL, "probe count out of range");
    if (flags & ~SPH_PROBE_FLUID_ONLY) return fail(SPH_EINVAL, "unknown probe flags");
    if (s->sys_host[slot].g.periodic)
        return fail(SPH_EUNSUP, "probe sums of periodic systems are not supported");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    if (np == 0) return 0;
    double* buf = nullptr;
    HIPC(hipMalloc((void**)&buf, sizeof(double) * 7 * (size_t)np));
    double* pin = buf;
    double* pout = buf + 3 * np;
    hipError_t e = hipMemcpyAsync(pin, x, sizeof(double) * 3 * np, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        enqueue_probe_sum(s, slot, pin, np, m, (flags & SPH_PROBE_FLUID_ONLY) ? 1 : 0, pout);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, pout, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    const hipError_t ef = hipFree(buf);
    if (e != hipSuccess) return fail(SPH_EHIP, std::string("probe sum: ") + hipGetErrorString(e));
    if (ef != hipSuccess) return fail(SPH_EHIP, std::string("hipFree: ") + hipGetErrorString(ef));
    return 0;
}

}  // namespace

// An asynchronous frame download (sph_snapshot_begin): the fields are
// gathered into reference order on the system's stream, then copied to
// pinned host memory on a second stream while the system's stream goes on
// with the next steps.
struct sph_snapshot {
    sph_system* s;
    long long n;
    int nf;
    int field[SPH_SNAPSHOT_MAX_FIELDS];
    long long off[SPH_SNAPSHOT_MAX_FIELDS];
    double* dev;
    double* host;
    long long total;
    hipEvent_t gathered, done;
};

namespace {

int snapshot_free(sph_snapshot* q)
{
    if (!q) return 0;
    if (q->done) (void)hipEventSynchronize(q->done);
    if (q->dev) (void)hipFree(q->dev);
    if (q->host) (void)hipHostFree(q->host);
    if (q->gathered) (void)hipEventDestroy(q->gathered);
    if (q->done) (void)hipEventDestroy(q->done);
    delete q;
    return 0;
}

int snapshot_begin(sph_system* s, int slot, const int* fields, int nf, sph_snapshot** out)
{
    *out = nullptr;
    if (nf < 1 || nf > SPH_SNAPSHOT_MAX_FIELDS) return fail(SPH_EINVAL, "1 to 9 fields per snapshot");
    sph_snapshot* q = new sph_snapshot();
    q->s = s;
    q->n = s->sys_n[slot];
    q->nf = nf;
    long long total = 0;
    for (int k = 0; k < nf; ++k) {
        int nc;
        if (fields[k] == SPH_FIELD_KEY || col_of_field(fields[k], &nc) < 0) {
            delete q;
            return fail(SPH_EINVAL, "unknown snapshot field");
        }
        q->field[k] = fields[k];
        q->off[k] = total;
        total += nc * q->n;
    }
    q->total = total;
    hipError_t e = hipSuccess;
    if (!s->copy_stream) e = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&q->gathered, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&q->done, hipEventDisableTiming);
    if (e == hipSuccess && total > 0) e = hipMalloc((void**)&q->dev, sizeof(double) * total);
    if (e == hipSuccess && total > 0) e = hipHostMalloc((void**)&q->host, sizeof(double) * total);
    if (e == hipSuccess) {
        for (int k = 0; k < nf; ++k) {
            int nc;
            const int col = col_of_field(q->field[k], &nc);
            enqueue_field_out(s, s->sys_off[slot], q->n, col, nc, q->dev + q->off[k]);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(q->gathered, s->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s->copy_stream, q->gathered, 0);
    if (e == hipSuccess && total > 0)
        e = hipMemcpyAsync(q->host, q->dev, sizeof(double) * total, hipMemcpyDeviceToHost,
                           s->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(q->done, s->copy_stream);
    if (e != hipSuccess) {
        const std::string msg = std::string("snapshot: ") + hipGetErrorString(e);
        snapshot_free(q);
        return fail(SPH_EHIP, msg);
    }
    *out = q;
    return 0;
}

// SmoothedParticles.sum(sys, func, p) for every particle p of system `slot`
// (src/core.jl:271-291) with a registered func; out[reference index].  The
// sums are staged in the field-transfer buffer.
int particle_sum(sph_system* s, int slot, int func, const double* fp, int nfp, double* out,
                 int64_t n)
{
    if (func != SPH_SUM_LJ_WALL) return fail(SPH_EUNSUP, "unknown particle-sum function");
    if (nfp != 4 || !fp) return fail(SPH_EINVAL, "SPH_SUM_LJ_WALL takes 4 parameters");
    if (s->sys_host[slot].g.periodic)
        return fail(SPH_EUNSUP, "particle sums of periodic systems are not supported");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    if (n != s->sys_n[slot]) return fail(SPH_EINVAL, "n must equal the particle count");
    if (n == 0) return 0;
    enqueue_particle_sum_lj(s, slot, s->sys_off[slot], n, fp, s->stage);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, s->stage, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

// Probe sums without device allocation (sph_probe_sum_noalloc).  The probe
// points and the results are staged in `scratch`, a file-scope device array
// of SPH_PROBE_NOALLOC_MAX * 7 doubles that every GPU holds until the process
// ends.  Its device address is taken once per device with
// hipGetSymbolAddress (the array's name is a host-side stand-in) and kept in
// scratch_ptr; calls that overlap on one device share the array and race.
static __device__ __align__(256) double scratch[7 * SPH_PROBE_NOALLOC_MAX];

static constexpr int MAX_DEVICES = 64;
static double* scratch_ptr[MAX_DEVICES];
static std::mutex scratch_mu;

int scratch_of(int device, double** out)
{
    if (device < 0 || device >= MAX_DEVICES) return fail(SPH_EINVAL, "device id out of range");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SPH_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = scratch_ptr[device];
    return 0;
}

int probe_sum_noalloc(sph_system* s, double m, int flags, const double* x, long long np,
                      double* out)
{
    if (np < 0) return fail(SPH_EINVAL, "probe count out of range");
    if (np > SPH_PROBE_NOALLOC_MAX)
        return fail(SPH_EINVAL, "sph_probe_sum_noalloc takes at most SPH_PROBE_NOALLOC_MAX (" +
                                    std::to_string(SPH_PROBE_NOALLOC_MAX) + ") probes per call");
    if (flags & ~SPH_PROBE_FLUID_ONLY) return fail(SPH_EINVAL, "unknown probe flags");
    if (s->sys_host[0].g.periodic)
        return fail(SPH_EUNSUP, "probe sums of periodic systems are not supported");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    if (np == 0) return 0;
    HIPC(hipSetDevice(s->device));
    double* buf = nullptr;
    int rc = scratch_of(s->device, &buf);
    if (rc) return rc;
    double* pin = buf;
    double* pout = buf + 3 * np;
    HIPC(hipMemcpyAsync(pin, x, sizeof(double) * 3 * np, hipMemcpyHostToDevice, s->stream));
    enqueue_probe_sum(s, 0, pin, np, m, (flags & SPH_PROBE_FLUID_ONLY) ? 1 : 0, pout);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, pout, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, s->stream));
    HIPC(hipStreamSynchronize(s->stream));
    return 0;
}

// Neighbour pairs of the current cell lists; with only_slot >= 0, of that
// system only (the pair kernels skip the tiles of the others).
int neighbour_pairs(sph_system* s, int only_slot, int64_t* pairs, int64_t cap, int64_t* n)
{
    *n = 0;
    if (s->n == 0) return 0;
    int rc;
    const int flagged = only_slot >= 0;
    if (flagged) {
        for (int i = 0; i < nsys(s); ++i) s->sys_host[i].pass_on = i == only_slot;
        if ((rc = upload_sysd(s))) return rc;
    }
    unsigned long long total = 0;
    long long* dpairs = nullptr;
    hipError_t e = hipMemsetAsync(s->pair_counter, 0, sizeof(unsigned long long), s->stream);
    rc = e == hipSuccess ? 0 : fail(SPH_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    if (!rc)
        rc = run_pair(s, P_COUNT, 0, 0, s->pair_counter, s->meta_host[META_NTILES],
                      SPH_KERNEL_DENSITY, MASK_TEST, flagged);
    if (!rc) {
        e = hipMemcpyAsync(&total, s->pair_counter, sizeof(total), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(SPH_EHIP, std::string("pair count: ") + hipGetErrorString(e));
    }
    *n = (int64_t)total;
    if (!rc && pairs && cap > 0 && total > 0) {
        const long long ncopy = std::min<long long>(cap, (long long)total);
        rc = dalloc(s, &dpairs, 2 * ncopy);
        if (!rc) {
            PairCtx c = make_ctx(s);
            c.emit_counter = s->pair_counter;
            c.pair_out = dpairs;
            c.pair_cap = ncopy;
            c.only_flagged = flagged;
            e = hipMemsetAsync(s->pair_counter, 0, sizeof(unsigned long long), s->stream);
            if (e == hipSuccess) {
                reset_queue(s);
                launch_pair(P_EMIT, s->precision, s->rows, s->flat, s->periodic, c,
                            pair_grid(s->meta_host[META_NTILES]), s->stream);
                e = hipGetLastError();
            }
            if (e == hipSuccess)
                e = hipMemcpyAsync(pairs, dpairs, sizeof(long long) * 2 * ncopy,
                                   hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            dfree(s, dpairs, 2 * ncopy);
            if (e != hipSuccess) rc = fail(SPH_EHIP, std::string("pair emission: ") + hipGetErrorString(e));
        }
    }
    if (flagged) {
        for (int i = 0; i < nsys(s); ++i) s->sys_host[i].pass_on = 0;
        const std::string msg = g_err;
        const int rc2 = upload_sysd(s);
        if (rc) g_err = msg;
        else rc = rc2;
    }
    return rc;
}

// Per-particle neighbour count and hash of the current cell lists (of
// system `slot`), in that system's reference order.
int neighbour_digest(sph_system* s, int slot, int64_t* counts, uint64_t* hashes, int64_t n)
{
    if ((long long)n != s->sys_n[slot]) return fail(SPH_EINVAL, "n must equal the particle count");
    if (n == 0) return 0;
    int rc;
    const int flagged = nsys(s) > 1;
    if (flagged) {
        for (int i = 0; i < nsys(s); ++i) s->sys_host[i].pass_on = i == slot;
        if ((rc = upload_sysd(s))) return rc;
    }
    long long* dout = nullptr;
    rc = dalloc(s, &dout, 2 * (long long)n);
    if (!rc) {
        PairCtx c = make_ctx(s);
        c.pair_out = dout;
        c.pair_cap = 2 * (long long)n;
        c.only_flagged = flagged;
        hipError_t e = hipMemsetAsync(dout, 0, sizeof(long long) * 2 * n, s->stream);
        if (e == hipSuccess) {
            reset_queue(s);
            launch_pair(P_DIGEST, s->precision, s->rows, s->flat, s->periodic, c,
                        pair_grid(s->meta_host[META_NTILES]), s->stream);
            e = hipGetLastError();
        }
        std::vector<long long> host;
        if (e == hipSuccess) {
            host.resize(2 * (size_t)n);
            e = hipMemcpyAsync(host.data(), dout, sizeof(long long) * 2 * n,
                               hipMemcpyDeviceToHost, s->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        dfree(s, dout, 2 * (long long)n);
        if (e != hipSuccess) {
            rc = fail(SPH_EHIP, std::string("neighbour digest: ") + hipGetErrorString(e));
        } else {
            for (int64_t i = 0; i < n; ++i) {
                if (counts) counts[i] = host[2 * i];
                if (hashes) hashes[i] = (uint64_t)host[2 * i + 1];
            }
        }
    }
    if (flagged) {
        for (int i = 0; i < nsys(s); ++i) s->sys_host[i].pass_on = 0;
        const std::string msg = g_err;
        const int rc2 = upload_sysd(s);
        if (rc) g_err = msg;
        else rc = rc2;
    }
    return rc;
}

// Mixed precision: times the density and force passes for every tile size on
// the current cell lists with a zero mass (both passes then add exactly zero
// to rho and a, so the state is left as it was) and keeps the fastest.  The
// mixed sums are plain FP32 per run before the FP64 fold, so their bits depend on T
// as well; the mixed path is stated by tolerance.  FP64: the occupancy rule.
int autotune(sph_system* s, int* best)
{
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    if (s->precision == SPH_FP64) {
        // The FP64 sums are added per run and per slice of a tile, so their
        // bits depend on the tile size: the plain FP64 path keeps the size
        // the occupancy rule gives, never one picked by timing, and the same
        // input gives the same bits on every run.
        int rc = sph_set_tile(s, 0);
        if (!rc && s->meta_host[META_NCELLS] > 0) {
            const int T = choose_tile((double)s->n / s->meta_host[META_NCELLS]);
            rc = sph_set_tile(s, T);
            if (!rc) rc = sph_set_tile(s, 0);
        }
        if (rc) return rc;
        if (best) *best = s->tile_T;
        return 0;
    }
    std::vector<Phys> saved;
    for (auto& d : s->sys_host) {
        saved.push_back(d.ph);
        d.ph.m = 0.0;
        d.ph.kW = 0.0;
        d.ph.kdW = 0.0;
    }
    int rc = upload_sysd(s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && hipEventCreate(&e0) != hipSuccess) rc = fail(SPH_EHIP, "hipEventCreate");
    if (!rc && hipEventCreate(&e1) != hipSuccess) rc = fail(SPH_EHIP, "hipEventCreate");
    float best_ms = 1e30f;
    int best_T = s->tile_T;
    const int cand[4] = {8, 16, 32, 64};
    if (!rc) enqueue_prho2(s);
    for (int i = 0; i < 4 && !rc; ++i) {
        if ((rc = sph_set_tile(s, cand[i]))) break;
        float tot = 0.0f;
        for (int rep = 0; rep < 3 && !rc; ++rep) {
            (void)hipEventRecord(e0, s->stream);
            if ((rc = run_pair(s, P_DENSITY, 0, 0, nullptr, s->meta_host[META_NTILES],
                               SPH_KERNEL_DENSITY, MASK_TEST_STORE)))
                break;
            if ((rc = run_pair(s, P_FORCE, 0, 0, nullptr, s->meta_host[META_NTILES],
                               SPH_KERNEL_FORCE, MASK_LOAD)))
                break;
            (void)hipEventRecord(e1, s->stream);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(SPH_EHIP, "hipEventSynchronize");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) tot += ms;
        }
        if (!rc && tot < best_ms) {
            best_ms = tot;
            best_T = cand[i];
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (size_t i = 0; i < s->sys_host.size(); ++i) s->sys_host[i].ph = saved[i];
    const std::string msg = g_err;
    const int rc2 = upload_sysd(s);
    if (rc) {
        g_err = msg;
        return rc;
    }
    if (rc2) return rc2;
    if ((rc = sph_set_tile(s, best_T))) return rc;
    s->masks_valid = false;  // the timing passes ran with a zero mass
    if (best) *best = best_T;
    return 0;
}

}  // namespace

namespace sph {
__global__ void k_reset_queue(int* queue)
{
    if (threadIdx.x < NQUEUE) queue[threadIdx.x * QSTRIDE] = 0;
}
void reset_queue(sph_system* s) { k_reset_queue<<<1, 64, 0, s->stream>>>(s->queue); }
}  // namespace sph

extern "C" {

const char* sph_last_error(void) { return g_err.c_str(); }

// ================================================================ sph.h ==

int sph_create(sph_system** out, double h, const double* bmin, const double* bmax, int precision)
{
    if (!out || !bmin || !bmax) return fail(SPH_EINVAL, "null argument");
    *out = nullptr;
    Geo g;
    int rc = make_geo(h, bmin, bmax, 0, &g);
    if (rc) return rc;
    sph_system* s;
    if ((rc = new_handle(precision, &s))) return rc;
    int slot;
    if ((rc = add_system(s, g, make_phys(g, nullptr), &slot))) {
        const std::string msg = g_err;
        destroy(s);
        g_err = msg;
        return rc;
    }
    *out = s;
    return 0;
}

int sph_destroy(sph_system* s)
{
    if (!s) return 0;
    return destroy(s);
}

int sph_append(sph_system* s, int64_t n, const double* x, const double* v, const double* rho,
               const double* P, const double* type)
{
    if (!s) return fail(SPH_EINVAL, "invalid append arguments");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return append_particles(s, 0, n, x, v, rho, P, type);
}

int sph_count(sph_system* s, int64_t* n)
{
    if (!s || !n) return fail(SPH_EINVAL, "null argument");
    *n = nsys(s) > 0 ? s->sys_n[0] : 0;
    return 0;
}

int sph_set_field(sph_system* s, int field, const double* host, int64_t n)
{
    if (!s || (!host && n > 0)) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return field_io(s, 0, field, const_cast<double*>(host), n, true);
}

int sph_get_field(sph_system* s, int field, double* host, int64_t n)
{
    if (!s || (!host && n > 0)) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return field_io(s, 0, field, host, n, false);
}

int sph_build_cells(sph_system* s, int64_t* n_out)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    int rc = build_cells(s);
    if (rc) return rc;
    if (n_out) *n_out = nsys(s) > 0 ? s->sys_n[0] : 0;
    return 0;
}

int sph_apply_binary(sph_system* s, int op, const sph_params* params, int self)
{
    if (!s || !params) return fail(SPH_EINVAL, "null argument");
    if (!is_binary(op)) return fail(SPH_EUNSUP, "not a registered binary operator");
    if (!s->cells_valid)
        return fail(SPH_ESTATE, "cell list is not current: call sph_build_cells after moving or adding particles");
    if (s->n == 0) return 0;
    int rc = set_params0(s, params);
    if (rc) return rc;
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
    rc = run_pair(s, pop, dens && self, 0, nullptr, s->meta_host[META_NTILES], kid, mode);
    if (rc) return rc;
    HIPC(hipStreamSynchronize(s->stream));
    if (dens) s->masks_valid = s->use_masks && s->masks != nullptr;
    return 0;
}

int sph_apply_unary(sph_system* s, int op, const sph_params* params)
{
    if (!s || !params) return fail(SPH_EINVAL, "null argument");
    if (!is_unary(op)) return fail(SPH_EUNSUP, "not a registered unary operator");
    int rc = set_params0(s, params);
    if (rc) return rc;
    hipEvent_t ev[2];
    prof_begin(s, ev);
    enqueue_unary(s, op, nullptr);
    prof_end(s, SPH_KERNEL_UNARY, ev);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
    if (op == SPH_OP_MOVE || op == SPH_OP_DRIFT) s->cells_valid = s->masks_valid = false;
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
    int rc = set_params0(s, params);
    if (rc) return rc;
    return reduce_slot(s, 0, what, out);
}

int sph_probe_sum(sph_system* s, const sph_params* params, int flags, const double* x,
                  int64_t np, double* out)
{
    if (!s || !params || (np > 0 && (!x || !out))) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return probe_sum(s, 0, params->m, flags, x, np, out);
}

int sph_probe_sum_noalloc(sph_system* s, const sph_params* params, int flags, const double* x,
                          int64_t np, double* out)
{
    if (!s || !params || (np > 0 && (!x || !out))) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return probe_sum_noalloc(s, params->m, flags, x, np, out);
}

int sph_snapshot_begin(sph_system* s, const int* fields, int nfields, sph_snapshot** out)
{
    if (!s || !fields || !out) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return snapshot_begin(s, 0, fields, nfields, out);
}

int sph_batch_snapshot_begin(sph_batch* b, int id, const int* fields, int nfields,
                             sph_snapshot** out)
{
    if (!b || !fields || !out) return fail(SPH_EINVAL, "null argument");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return snapshot_begin(b, slot, fields, nfields, out);
}

int sph_snapshot_wait(sph_snapshot* q, int64_t* n)
{
    if (!q) return fail(SPH_EINVAL, "null snapshot");
    HIPC(hipEventSynchronize(q->done));
    if (n) *n = q->n;
    return 0;
}

int sph_snapshot_field(sph_snapshot* q, int field, const double** data)
{
    if (!q || !data) return fail(SPH_EINVAL, "null argument");
    HIPC(hipEventSynchronize(q->done));
    for (int k = 0; k < q->nf; ++k)
        if (q->field[k] == field) {
            *data = q->host ? q->host + q->off[k] : nullptr;
            return 0;
        }
    return fail(SPH_EINVAL, "field not in the snapshot");
}

int sph_snapshot_free(sph_snapshot* q) { return snapshot_free(q); }

int sph_generate_lattice(sph_system* s, const sph_lattice_spec* spec, int64_t* n_added)
{
    if (!s || !spec) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return generate_lattice(s, 0, spec, n_added);
}

int sph_batch_generate_lattice(sph_batch* b, int id, const sph_lattice_spec* spec,
                               int64_t* n_added)
{
    if (!b || !spec) return fail(SPH_EINVAL, "null argument");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return generate_lattice(b, slot, spec, n_added);
}

int sph_particle_sum(sph_system* s, int func, const double* fparams, int nfparams, double* out,
                     int64_t n)
{
    if (!s || (n > 0 && !out)) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    return particle_sum(s, 0, func, fparams, nfparams, out, n);
}

int sph_neighbour_pairs(sph_system* s, int64_t* pairs, int64_t cap, int64_t* n)
{
    if (!s || !n) return fail(SPH_EINVAL, "null argument");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    return neighbour_pairs(s, -1, pairs, cap, n);
}

int sph_neighbour_digest(sph_system* s, int64_t* counts, uint64_t* hashes, int64_t n)
{
    if (!s || (n > 0 && !counts && !hashes)) return fail(SPH_EINVAL, "null argument");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    if (!s->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    return neighbour_digest(s, 0, counts, hashes, n);
}

int sph_step(sph_system* s, int scheme, const sph_params* params, int nsteps, int64_t* pairs)
{
    if (!s || !params) return fail(SPH_EINVAL, "null argument");
    sph_params p1 = *params;
    p1.coeff = 1.0;
    int rc = set_params0(s, &p1);
    if (rc) return rc;
    if ((rc = run_steps(s, scheme, nsteps))) return rc;
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

int sph_step_tests(sph_system* s, int64_t* density_tests)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    if (density_tests) *density_tests = (int64_t)s->step_pairs[2];
    return 0;
}

int sph_key_geometry(sph_system* s, int64_t* key_phase, int64_t* key_lim, int64_t* key_max,
                     int64_t* key_diff, int* n_diff)
{
    if (!s) return fail(SPH_EINVAL, "null system");
    if (nsys(s) == 0) return fail(SPH_ESTATE, "the handle holds no system");
    const Geo& g = s->sys_host[0].g;
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
    int rc = set_params0(s, params);
    if (rc) return rc;
    return autotune(s, best);
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
    case SPH_OPTION_KEPT_LISTS:
        s->use_kept = value != 0;
        s->masks_valid = false;
        return 0;
    case SPH_OPTION_GRAPH:
        s->use_graph = value != 0;
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
        PairCtx c = make_ctx(s);
        HIPC(hipMemsetAsync(s->pair_counter, 0, 2 * sizeof(unsigned long long), s->stream));
        launch_candidates(s->rows, c, s->icols[s->cur] + I_KEY * s->cap, s->pair_counter, s->n,
                          s->stream);
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

// ========================================================== sph_batch.h ==

int sph_batch_create(sph_batch** out, int precision)
{
    if (!out) return fail(SPH_EINVAL, "null argument");
    return new_handle(precision, out);
}

int sph_batch_destroy(sph_batch* b) { return sph_destroy(b); }

int sph_batch_add(sph_batch* b, const sph_system_desc* desc, int64_t n, const double* x,
                  const double* v, const double* rho, const double* P, const double* type,
                  int* id)
{
    if (!b || !desc || !id) return fail(SPH_EINVAL, "null argument");
    if (n < 0 || (n > 0 && !x)) return fail(SPH_EINVAL, "invalid particle arguments");
    Geo g;
    int rc = make_geo(desc->h, desc->bbox_min, desc->bbox_max, desc->periodic, &g);
    if (rc) return rc;
    if ((rc = compact(b))) return rc;
    sph_params p = desc->params;
    p.coeff = 1.0;  // the fused step advances by dt
    int slot;
    if ((rc = add_system(b, g, make_phys(g, &p), &slot))) return rc;
    if ((rc = append_particles(b, slot, n, x, v, rho, P, type))) {
        // the system never held particles: retire it
        const std::string msg = g_err;
        b->sys_host[slot].live = 0;
        b->slot_of_id[b->id_of_slot[slot]] = -1;
        (void)upload_sysd(b);
        g_err = msg;
        return rc;
    }
    *id = b->id_of_slot[slot];
    return 0;
}

int sph_batch_retire(sph_batch* b, int id)
{
    if (!b) return fail(SPH_EINVAL, "null batch");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    // The device counters keep the system's particle count until the next
    // build moves its particles out of the cells (the removal replay of
    // the other systems relies on the per-system ranges).
    b->sys_host[slot].live = 0;
    b->retired_since_build = true;
    b->cells_valid = b->masks_valid = false;
    return upload_sysd(b);
}

int sph_batch_size(sph_batch* b, int* nsys_live, int64_t* particles)
{
    if (!b) return fail(SPH_EINVAL, "null batch");
    int live = 0;
    long long np = 0;
    for (int i = 0; i < nsys(b); ++i)
        if (b->sys_host[i].live) {
            ++live;
            np += b->sys_n[i];
        }
    if (nsys_live) *nsys_live = live;
    if (particles) *particles = np;
    return 0;
}

int sph_batch_count(sph_batch* b, int id, int64_t* n)
{
    if (!b || !n) return fail(SPH_EINVAL, "null argument");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    *n = b->sys_n[slot];
    return 0;
}

int sph_batch_get_field(sph_batch* b, int id, int field, double* host, int64_t n)
{
    if (!b || (!host && n > 0)) return fail(SPH_EINVAL, "null argument");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return field_io(b, slot, field, host, n, false);
}

int sph_batch_set_field(sph_batch* b, int id, int field, const double* host, int64_t n)
{
    if (!b || (!host && n > 0)) return fail(SPH_EINVAL, "null argument");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return field_io(b, slot, field, const_cast<double*>(host), n, true);
}

int sph_batch_reduce(sph_batch* b, int id, int what, double* out)
{
    if (!b || !out) return fail(SPH_EINVAL, "null argument");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return reduce_slot(b, slot, what, out);
}

int sph_batch_probe_sum(sph_batch* b, int id, int flags, const double* x, int64_t np,
                        double* out)
{
    if (!b || (np > 0 && (!x || !out))) return fail(SPH_EINVAL, "null argument");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return probe_sum(b, slot, b->sys_host[slot].ph.m, flags, x, np, out);
}

int sph_batch_build_cells(sph_batch* b)
{
    if (!b) return fail(SPH_EINVAL, "null batch");
    int rc = compact(b);
    if (rc) return rc;
    return build_cells(b);
}

int sph_batch_neighbour_pairs(sph_batch* b, int id, int64_t* pairs, int64_t cap, int64_t* n)
{
    if (!b || !n) return fail(SPH_EINVAL, "null argument");
    if (!b->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return neighbour_pairs(b, slot, pairs, cap, n);
}

int sph_batch_neighbour_digest(sph_batch* b, int id, int64_t* counts, uint64_t* hashes,
                               int64_t n)
{
    if (!b || (n > 0 && !counts && !hashes)) return fail(SPH_EINVAL, "null argument");
    if (!b->cells_valid) return fail(SPH_ESTATE, "cell list is not current");
    const int slot = slot_of(b, id);
    if (slot < 0) return fail(SPH_EINVAL, "unknown system id");
    return neighbour_digest(b, slot, counts, hashes, n);
}

int sph_batch_init_rho0(sph_batch* b)
{
    if (!b) return fail(SPH_EINVAL, "null batch");
    int rc = compact(b);
    if (rc) return rc;
    bool any = false;
    for (int i = 0; i < nsys(b); ++i) {
        b->sys_host[i].pass_on = b->sys_host[i].live && b->rho0_pending[i];
        any |= b->sys_host[i].pass_on != 0;
    }
    if (!any) return 0;
    if ((rc = upload_sysd(b)) == 0 && (rc = build_cells(b)) == 0) {
        // rho0_p = sum_q m W + m W(h, 0), written (not accumulated), for the
        // flagged systems only
        rc = run_pair(b, P_DENSITY_RHO0, 1, 1, nullptr, b->meta_host[META_NTILES],
                      SPH_KERNEL_DENSITY, MASK_TEST, 1);
        if (!rc && hipStreamSynchronize(b->stream) != hipSuccess)
            rc = fail(SPH_EHIP, "reference density pass failed");
    }
    for (int i = 0; i < nsys(b); ++i) {
        if (!rc && b->sys_host[i].pass_on) b->rho0_pending[i] = 0;
        b->sys_host[i].pass_on = 0;
    }
    const std::string msg = g_err;
    const int rc2 = upload_sysd(b);
    if (rc) {
        g_err = msg;
        return rc;
    }
    return rc2;
}

int sph_batch_step(sph_batch* b, int scheme, int nsteps, int64_t* pairs)
{
    if (!b) return fail(SPH_EINVAL, "null batch");
    int rc = compact(b);
    if (rc) return rc;
    if ((rc = run_steps(b, scheme, nsteps))) return rc;
    if (pairs) *pairs = (int64_t)b->step_pairs[0];
    return 0;
}

int sph_batch_set_tile(sph_batch* b, int targets_per_tile) { return sph_set_tile(b, targets_per_tile); }

int sph_batch_autotune(sph_batch* b, int* best)
{
    if (!b) return fail(SPH_EINVAL, "null batch");
    int rc = b->cells_valid ? 0 : sph_batch_build_cells(b);
    if (rc) return rc;
    return autotune(b, best);
}

int sph_batch_footprint(sph_batch* b, int64_t* bytes)
{
    if (!b || !bytes) return fail(SPH_EINVAL, "null argument");
    *bytes = b->bytes;
    return 0;
}

}  // extern "C"
