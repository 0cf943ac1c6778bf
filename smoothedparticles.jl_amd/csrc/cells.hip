// Cell-list builder: the device form of create_cell_list! (src/core.jl:51-90)
// for every system of a batch at once.
//
// Stages, all on the handle's stream, no host round trip:
//   1. k_keys     per particle: FP64 find_key (src/structs.jl:97-106) and the
//                 inclusive bounding-box test (src/geometry.jl:24-30); a
//                 particle outside goes to its system's removal bucket.
//                 Periodic systems wrap positions into the box instead.
//                 Global key = the system's cell_base + local key.  Per-bucket
//                 counts by wave-aggregated atomics (one atomic per distinct
//                 key in a wave; the input is the previous sorted order, so a
//                 wave usually touches one or two cells).
//   2. scan       exclusive prefix sum of the counts -> cell_start
//   3. k_scatter  bucket every particle
//   4. k_rank     order each bucket by descending reference index, the order
//                 add_index! keeps (src/core.jl:26-41); deterministic
//                 whatever order the atomics handed out (the bucket of
//                 retired systems keeps its scatter order, never used)
//   5. k_gather   permute every column into the new order
//   6. removal    each system's removal bucket holds its removed particles in
//                 descending reference index; the reference's swap-with-end
//                 renumbering (src/core.jl:72-81) is replayed per system
//   7. tiles      a scan of ceil(count/T) per cell, the 64-candidate chunks a
//                 target of each cell walks, a scan of pair-mask words, and
//                 the tile table of the pair kernels
//
// This is a counting sort on the global cell key (one radix digit) followed
// by an in-bucket rank.
#include "scan_kernels.h"
#include "sph_internal.h"

namespace sph {


__global__ void k_fill_u32(unsigned* p, unsigned v, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- prepare --
// Per-system reference-index base of the inverse map and the count at the
// start of the build.
// One block of 256 threads; systems in rounds of 256, an exclusive scan of
// the counts per round.
__global__ __launch_bounds__(256) void k_sys_prep(int nsys, int* sc, int* meta)
{
    __shared__ int part[4];
    int carry = 0;
    for (int s0 = 0; s0 < nsys; s0 += 256) {
        const int s = s0 + (int)threadIdx.x;
        const int n = s < nsys ? sc[SYSCNT * s + SC_N] : 0;
        int incl = n;
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        if (lane == 63) part[w] = incl;
        __syncthreads();
        int off = carry;
        for (int k = 0; k < w; ++k) off += part[k];
        const int round = part[0] + part[1] + part[2] + part[3];
        if (s < nsys) {
            sc[SYSCNT * s + SC_NTOT] = n;
            sc[SYSCNT * s + SC_REFBASE] = off + incl - n;
        }
        carry += round;
        __syncthreads();
    }
    if (threadIdx.x == 0) meta[META_NTOT] = meta[META_N];
}

// ------------------------------------------------------------------ keys --
// Bucket key of particle i of system s: its cell, the system's removal
// bucket (outside the box, or not finite in a periodic system) or the dead
// bucket (retired system).  Periodic systems wrap the position (written
// back).
__device__ inline int key_of(const SysDesc& d, int s, int nsys, int cells_total, double* x,
                             double* y, double* z, int i)
{
    const Geo& g = d.g;
    double px = x[i], py = y[i], pz = z[i];
    if (!d.live) return cells_total + nsys;
    if (g.periodic) {
        // positions live in [bmin, bmin + box) on the periodic axes
        if (!(isfinite(px) && isfinite(py) && isfinite(pz))) return cells_total + s;
        px = px - g.box[0] * floor((px - g.bmin[0]) / g.box[0]);
        py = py - g.box[1] * floor((py - g.bmin[1]) / g.box[1]);
        x[i] = px;
        y[i] = py;
        long long ci = (long long)floor((px - g.bmin[0]) / g.cell[0]);
        long long cj = (long long)floor((py - g.bmin[1]) / g.cell[1]);
        long long ck = 0;
        if (g.L3 > 1) {
            pz = pz - g.box[2] * floor((pz - g.bmin[2]) / g.box[2]);
            z[i] = pz;
            ck = (long long)floor((pz - g.bmin[2]) / g.cell[2]);
        }
        ci = ci < 0 ? 0 : (ci >= g.L1 ? g.L1 - 1 : ci);
        cj = cj < 0 ? 0 : (cj >= g.L2 ? g.L2 - 1 : cj);
        ck = ck < 0 ? 0 : (ck >= g.L3 ? g.L3 - 1 : ck);
        return d.cell_base + (int)(ci + (long long)g.L1 * (cj + (long long)g.L2 * ck));
    }
    const bool inside = g.bmin[0] <= px && px <= g.bmax[0] && g.bmin[1] <= py &&
                        py <= g.bmax[1] && g.bmin[2] <= pz && pz <= g.bmax[2];
    if (!inside) return cells_total + s;
    long long ci = (long long)floor(px / g.h) - g.kp[0];
    long long cj = (long long)floor(py / g.h) - g.kp[1];
    long long ck = (long long)floor(pz / g.h) - g.kp[2];
    // In range by monotonicity of the correctly rounded division; the clamp
    // only guards memory.
    ci = ci < 0 ? 0 : (ci >= g.L1 ? g.L1 - 1 : ci);
    cj = cj < 0 ? 0 : (cj >= g.L2 ? g.L2 - 1 : cj);
    ck = ck < 0 ? 0 : (ck >= g.L3 ? g.L3 - 1 : ck);
    return d.cell_base + (int)(ci + (long long)g.L1 * (cj + (long long)g.L2 * ck));
}

__global__ __launch_bounds__(256) void k_keys(const SysDesc* __restrict__ sysd, int nsys,
                                              int cells_total, double* x, double* y, double* z,
                                              const int* __restrict__ sysid, const int* meta,
                                              int* __restrict__ key, int* __restrict__ slot,
                                              unsigned* __restrict__ cnt)
{
    const int n = meta[META_N];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int s = i < n ? sysid[i] : 0;
    // a wave of one system reads its descriptor once, through the scalar
    // cache (the usual case: systems are contiguous)
    const int s0 = __builtin_amdgcn_readfirstlane(s);
    int k = -1;
    if (__all(i >= n || s == s0)) {
        if (i < n) k = key_of(sysd[s0], s0, nsys, cells_total, x, y, z, i);
    } else if (i < n) {
        k = key_of(sysd[s], s, nsys, cells_total, x, y, z, i);
    }
    if (i < n) key[i] = k;
    // One atomic per distinct key among the wave's lanes.
    const int lane = __lane_id();
    unsigned long long active = __ballot(i < n);
    unsigned myslot = 0;
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int lk = __shfl(k, leader);
        const unsigned long long peers = __ballot(k == lk) & active;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&cnt[lk], (unsigned)__popcll(peers));
        base = __shfl(base, leader);
        if ((peers >> lane) & 1ull)
            myslot = base + (unsigned)__popcll(peers & ((1ull << lane) - 1ull));
        active &= ~peers;
    }
    if (i < n) slot[i] = (int)myslot;
}

// ------------------------------------------------------------------ scan --
// (kernels in scan_kernels.h)
struct SrcCount {
    using V = unsigned;
    const unsigned* c;
    __device__ V operator()(int i) const { return c[i]; }
};
// Tiles of the cell at brick rank i: ceil(count / T).
struct SrcTiles {
    using V = unsigned;
    const int* cs;
    const int* order;
    int ncells;
    int T;
    __device__ V operator()(int i) const
    {
        if (i >= ncells) return 0u;
        const int c = order[i];
        const int n = cs[c + 1] - cs[c];
        return (unsigned)((n + T - 1) / T);
    }
};
// Mask words of the cell at brick rank i: tiles * chunks per target * T.
struct SrcMaskWords {
    using V = unsigned long long;
    const int* tile_start;
    const int* chunks;
    const int* order;
    int ncells;
    int T;
    __device__ V operator()(int i) const
    {
        if (i >= ncells) return 0ull;
        const long long nt = tile_start[i + 1] - tile_start[i];
        return (V)(nt * chunks[order[i]] * T);
    }
};

long long scan_partials_needed(long long n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// -------------------------------------------------- scatter, rank, gather --
__global__ __launch_bounds__(256) void k_scatter(const int* meta, const int* __restrict__ key,
                                                 const int* __restrict__ slot,
                                                 const int* __restrict__ ref,
                                                 const int* __restrict__ cell_start,
                                                 int* tmp_idx, int* tmp_key, int* tmp_ref)
{
    const int n = meta[META_NTOT];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int k = key[i];
    const int pos = cell_start[k] + slot[i];
    tmp_idx[pos] = i;
    tmp_key[pos] = k;
    tmp_ref[pos] = ref[i];
}

// Rank of each particle in its bucket: the number of bucket members with a
// larger reference index (ties -- particles of different retired systems --
// by scatter position), so the ranks are a permutation of the bucket.
//   - the bucket of retired systems (dead_key) keeps its scatter order: its
//     order is never used, and it can hold whole retired systems;
//   - a block whose threads all lie in one bucket of more than RANK_LDS_MIN
//     particles (a removal bucket after many particles left the box) counts
//     against the bucket staged through LDS, 1024 indices at a time;
//   - otherwise the buckets the block's threads lie in (a few cells of tens
//     to hundreds of particles) are staged through LDS when they fit in
//     RANK_SPAN entries, and each thread counts against its bucket there;
//   - else each thread scans its bucket in global memory.
static constexpr int RANK_LDS_MIN = 512;
static constexpr int RANK_SPAN = 2048;

__global__ __launch_bounds__(256) void k_rank(const int* meta, const int* __restrict__ cell_start,
                                              const int* __restrict__ tmp_idx,
                                              const int* __restrict__ tmp_key,
                                              const int* __restrict__ tmp_ref, int* perm,
                                              int dead_key)
{
    __shared__ __align__(16) int tile[RANK_SPAN];
    const int n = meta[META_NTOT];
    const int b0 = blockIdx.x * 256;
    if (b0 >= n) return;  // whole block past the end (uniform)
    const int u = b0 + threadIdx.x;
    const bool in = u < n;
    const int c0 = tmp_key[b0];
    const int c1 = tmp_key[min(n - 1, b0 + 255)];
    const int s0 = cell_start[c0], e0 = cell_start[c0 + 1];
    if (c0 == c1 && c0 != dead_key && e0 - s0 > RANK_LDS_MIN) {
        const int r = in ? tmp_ref[u] : 0;
        int rank = 0;
        for (int base = s0; base < e0; base += 1024) {
            const int m = min(1024, e0 - base);
            __syncthreads();
            for (int j = threadIdx.x; j < m; j += 256) tile[j] = tmp_ref[base + j];
            __syncthreads();
            for (int j = 0; j < m; ++j) {
                const int rv = tile[j];
                rank += rv > r || (rv == r && base + j < u);
            }
        }
        if (in) perm[s0 + rank] = tmp_idx[u];
        return;
    }
    // the span of the buckets of the block's first and last thread holds
    // every bucket in between (scatter order is bucket order)
    const int span0 = s0, span1 = cell_start[c1 + 1];
    if (c1 != dead_key && span1 - span0 <= RANK_SPAN) {
        for (int j = threadIdx.x; j < span1 - span0; j += 256) tile[j] = tmp_ref[span0 + j];
        __syncthreads();
        if (!in) return;
        // a live bucket (a cell or a system's removal bucket) holds one
        // system's particles, whose reference indices are distinct: the
        // rank is the count of larger ones, read four at a time
        const int c = tmp_key[u];
        const int s = cell_start[c], e = cell_start[c + 1];
        const int r = tile[u - span0];
        int rank = 0, v = s - span0;
        const int ve = e - span0;
        for (; v < ve && (v & 3); ++v) rank += tile[v] > r;
        for (; v + 4 <= ve; v += 4) {
            const int4 q = *reinterpret_cast<const int4*>(&tile[v]);
            rank += (q.x > r) + (q.y > r) + (q.z > r) + (q.w > r);
        }
        for (; v < ve; ++v) rank += tile[v] > r;
        perm[s + rank] = tmp_idx[u];
        return;
    }
    if (!in) return;
    const int c = tmp_key[u];
    if (c == dead_key) {
        perm[u] = tmp_idx[u];
        return;
    }
    const int s = cell_start[c], e = cell_start[c + 1];
    const int r = tmp_ref[u];
    int rank = 0;
#pragma unroll 8
    for (int v = s; v < e; ++v) {
        const int rv = tmp_ref[v];
        rank += rv > r || (rv == r && v < u);
    }
    perm[s + rank] = tmp_idx[u];
}

__global__ __launch_bounds__(256) void k_gather(const int* meta, const int* __restrict__ perm,
                                                const double* __restrict__ src,
                                                double* __restrict__ dst,
                                                const int* __restrict__ isrc,
                                                int* __restrict__ idst, long long cap,
                                                unsigned skip)
{
    const int n = meta[META_NTOT];
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const long long j = perm[k];
#pragma unroll
    for (int f = 0; f < NCOL; ++f)
        if (!((skip >> f) & 1u)) dst[f * cap + k] = src[f * cap + j];
#pragma unroll
    for (int f = 0; f < NICOL; ++f) idst[f * cap + k] = isrc[f * cap + j];
}

// Totals of the build, per system and overall.
__global__ void k_finalize(const SysDesc* __restrict__ sysd, int nsys, int cells_total, int* sc,
                           int* meta, const int* cell_start)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) {
        const int na = cell_start[cells_total];
        meta[META_NREM] = meta[META_NTOT] - na;
        meta[META_N] = na;
        meta[META_NCELLS] = 0;
    }
    if (s >= nsys) return;
    const SysDesc& d = sysd[s];
    const int off = cell_start[d.cell_base];
    const int end = cell_start[d.cell_base + d.g.key_max];
    sc[SYSCNT * s + SC_OFF] = off;
    sc[SYSCNT * s + SC_N] = d.live ? end - off : 0;
    sc[SYSCNT * s + SC_NREM] = cell_start[cells_total + s + 1] - cell_start[cells_total + s];
}

// Removal renumbering (src/core.jl:72-81): with the removed reference
// indices r_1 > r_2 > ... > r_m of a system, particles[r_i] =
// particles[N+1-i] in turn.  inv maps (system reference base + reference
// position) to the sorted slot that holds it.
__global__ __launch_bounds__(256) void k_inv(const int* meta, const int* __restrict__ ref,
                                             const int* __restrict__ sysid, const int* sc,
                                             int* inv)
{
    if (meta[META_NREM] == 0) return;
    const int n = meta[META_NTOT];
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) inv[sc[SYSCNT * sysid[k] + SC_REFBASE] + ref[k]] = k;
}

__global__ void k_remove_seq(const int* meta, int nsys, int cells_total, const int* sc,
                             const int* cell_start, int* ref, int* inv)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys || meta[META_NREM] == 0) return;
    const int m = sc[SYSCNT * s + SC_NREM];
    if (m == 0) return;
    const int first = cell_start[cells_total + s];
    const int ntot = sc[SYSCNT * s + SC_NTOT];
    const int rb = sc[SYSCNT * s + SC_REFBASE];
    for (int i = 0; i < m; ++i) {
        const int r = ref[first + i];  // descending
        const int src = ntot - 1 - i;  // reference position N+1-i (1-based)
        const int slot = inv[rb + src];
        inv[rb + r] = slot;
        ref[slot] = r;
    }
}

// Rank of local cell lc of a system in brick order: the grid is cut into
// bricks of BX x BY x BZ cells (BZ = 1 in 2D), the bricks are numbered x
// fastest, then y, then z, and the cells of a brick x fastest; the bricks
// on the far faces are cut to the grid, so the ranks are 0 .. key_max - 1.
// The tile table follows this order: each XCD walks a contiguous share of
// it, and the tiles it has in flight then cover a compact block of cells
// whose neighbour cells (the candidates) are few enough to stay in its L2,
// where a row-major order spreads them over whole rows of the grid.
constexpr int BRICK_X = 4, BRICK_Y = 4, BRICK_Z = 4, BRICK_2D = 8;

__device__ inline int brick_rank(const Geo& g, int lc)
{
    const int bx = g.L3 > 1 ? BRICK_X : BRICK_2D;
    const int by = g.L3 > 1 ? BRICK_Y : BRICK_2D;
    const int bz = g.L3 > 1 ? BRICK_Z : 1;
    const int ix = lc % g.L1, r = lc / g.L1;
    const int iy = r % g.L2, iz = r / g.L2;
    const int kx = ix / bx, ky = iy / by, kz = iz / bz;
    const int wx = min(bx, g.L1 - kx * bx);  // extents of this brick
    const int hy = min(by, g.L2 - ky * by);
    const int hz = min(bz, g.L3 - kz * bz);
    return kz * bz * g.L1 * g.L2 + ky * by * g.L1 * hz + kx * bx * hy * hz +
           (ix - kx * bx) + wx * ((iy - ky * by) + hy * (iz - kz * bz));
}

// 64-candidate chunks one target of global cell c walks: the pair kernels
// concatenate the cell's stencil ranges and cut the list into chunks of 64.
// Also places c at its brick rank in order[].
template <int ROWS>
__global__ __launch_bounds__(256) void k_cell_chunks(const SysDesc* __restrict__ sysd,
                                                     const int* __restrict__ cell_sys,
                                                     int ncells, const int* __restrict__ cell_start,
                                                     int* chunks, int* order)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    const SysDesc& d = sysd[cell_sys[c]];
    const int lc = c - d.cell_base;
    order[d.cell_base + brick_rank(d.g, lc)] = c;
    int n = 0;
    if (cell_start[c + 1] > cell_start[c]) {
        for (int row = 0; row < ROWS; ++row) {
            Seg seg[2];
            const int ns = row_segments<ROWS>(d, cell_start, lc, row, seg);
            for (int q = 0; q < 2; ++q)
                if (q < ns) n += seg[q].j1 - seg[q].j0;
        }
    }
    chunks[c] = (n + 63) / 64;
}

// Tile table, in brick order: tile k of the cell c at rank r covers targets
// cell_start[c] + k*T ...; its pair masks start at mask_start[r] +
// k * chunks[c] * T, its kept-candidate list at 64 / T times that.
__global__ __launch_bounds__(256) void k_tiles(const int* __restrict__ cell_start,
                                               const int* __restrict__ tile_start,
                                               const int* __restrict__ chunks,
                                               const int* __restrict__ order,
                                               const unsigned long long* __restrict__ mask_start,
                                               int ncells, int T, long long mask_cap,
                                               long long kept_cap, int4* tiles, int* meta,
                                               unsigned long long* mask_total)
{
    __shared__ int with_tiles;
    if (threadIdx.x == 0) with_tiles = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        meta[META_NTILES] = tile_start[ncells];
        const unsigned long long need = mask_start[ncells];
        *mask_total = need;
        meta[META_MASK_OK] = (long long)need <= mask_cap ? 1 : 0;
        // 64 kept-candidate entries per chunk of a tile, T mask words
        meta[META_KEPT_OK] = (long long)(need / T * 64) <= kept_cap ? 1 : 0;
    }
    __syncthreads();
    // cells with tiles: counted per block, one global atomic per block (a
    // single address serialises its atomics)
    int n = 0;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < ncells; r += gridDim.x * 256) {
        const int t0 = tile_start[r], nt = tile_start[r + 1] - t0;
        n += nt > 0;
        const int c = order[r];
        const int cs = cell_start[c];
        const long long words = (long long)chunks[c] * T;
        for (int t = 0; t < nt; ++t) {
            const unsigned long long base = mask_start[r] + (unsigned long long)(t * words);
            tiles[t0 + t] = make_int4(c, cs + t * T, (int)(unsigned)(base & 0xffffffffull),
                                      (int)(base >> 32));
        }
    }
    if (n) atomicAdd(&with_tiles, n);
    __syncthreads();
    if (threadIdx.x == 0 && with_tiles) atomicAdd(&meta[META_NCELLS], with_tiles);
}

// Bounding box of each tile's targets (absolute FP64), read by the pair
// kernels to cull candidates.  A wave covers 64 / T tiles at a time, one
// lane per target.  The coordinates go through LDS; of the T lanes of a
// tile, G = T / 8 lanes per bound (c < 6: min x, max x, min y, ...) reduce
// eight coordinates each and combine in log2(G) shuffle steps.  Targets
// past the cell's end take the first target's place; NaN coordinates are
// ignored (fmin/fmax).
template <bool FLAT>
__global__ __launch_bounds__(256) void k_tile_boxes(const int4* __restrict__ tiles,
                                                    const int* __restrict__ cell_start,
                                                    const int* meta, const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z, int T,
                                                    double* boxes)
{
    __shared__ double xs[4][3][64];
    const int ntiles = meta[META_NTILES];
    const int lane = threadIdx.x & 63;
    double(*w)[64] = xs[threadIdx.x >> 6];
    const int per_wave = 64 / T, G = T / 8;
    const int tw = lane / T, lt = lane % T;  // tile in the wave, target in the tile
    const int c = lt / G, j = lt % G, axis = c >> 1;
    const bool lo = (c & 1) == 0;
    const long long per_block = 4ll * per_wave, step = (long long)gridDim.x * per_block;
    // the block's four waves step together (block barriers)
    for (long long b = (long long)blockIdx.x * per_block; b < ntiles; b += step) {
        const long long i = b + (threadIdx.x >> 6) * per_wave + tw;
        if (i < ntiles) {
            const int4 td = tiles[i];
            const int p = td.y + lt;
            const int pp = p < cell_start[td.x + 1] ? p : td.y;
            w[0][lane] = x[pp];
            w[1][lane] = y[pp];
            w[2][lane] = FLAT ? 0.0 : z[pp];
        }
        __syncthreads();
        double v = 0.0;
        if (c < 6) {
            const int k0 = tw * T + 8 * j;
            v = w[axis][k0];
            for (int k = 1; k < 8; ++k) {
                const double u = w[axis][k0 + k];
                v = lo ? fmin(v, u) : fmax(v, u);
            }
        }
        for (int o = 1; o < G; o <<= 1) {
            const double u = __shfl_xor(v, o);
            v = lo ? fmin(v, u) : fmax(v, u);
        }
        if (i < ntiles && c < 6 && j == 0) boxes[6 * i + c] = v;
        __syncthreads();
    }
}

// ------------------------------------------------------------ host side --
static inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

void enqueue_tiles(sph_system* s)
{
    const int ncells = (int)s->cells_total;
    const unsigned nbc = nblocks(ncells + 1, 256);
    if (s->rows == 3)
        k_cell_chunks<3><<<nbc, 256, 0, s->stream>>>(s->sysd, s->cell_sys, ncells, s->cell_start,
                                                     s->cell_chunks, s->cell_order);
    else
        k_cell_chunks<9><<<nbc, 256, 0, s->stream>>>(s->sysd, s->cell_sys, ncells, s->cell_start,
                                                     s->cell_chunks, s->cell_order);
    scan(s, SrcTiles{s->cell_start, s->cell_order, ncells, s->tile_T}, ncells + 1, s->tile_start,
         s->scan_part);
    scan(s, SrcMaskWords{s->tile_start, s->cell_chunks, s->cell_order, ncells, s->tile_T},
         ncells + 1, s->mask_start, s->scan_part64);
    k_tiles<<<nbc < 256 ? nbc : 256, 256, 0, s->stream>>>(s->cell_start, s->tile_start,
                                        s->cell_chunks, s->cell_order,
                                        s->mask_start, ncells, s->tile_T,
                                        s->masks ? s->mask_cap : 0,
                                        s->kept && (s->kept_rg || !s->periodic) ? s->kept_cap : 0,
                                        s->tiles, s->meta,
                                        s->mask_total);
    const double* c = s->cols[s->cur];
    const long long cap = s->cap;
    const unsigned nbt = 2048;  // waves loop over the tiles
    auto boxes = s->flat ? k_tile_boxes<true> : k_tile_boxes<false>;
    boxes<<<nbt, 256, 0, s->stream>>>(s->tiles, s->cell_start, s->meta, c + C_X * cap,
                                      c + C_Y * cap, c + C_Z * cap, s->tile_T, s->boxes);
}

// skip_cols: columns (bit C_*) the caller rewrites for every particle before
// reading them; the gather leaves their values in the new order undefined.
void enqueue_build(sph_system* s, unsigned skip_cols)
{
    hipStream_t st = s->stream;
    const long long cap = s->cap;
    const int nsys = (int)s->sys_host.size();
    const int ncells = (int)s->cells_total;
    const long long nbuckets = s->cells_total + nsys + 1;
    const unsigned nb = nblocks(s->n > 0 ? s->n : 1, 256);
    double* c = s->cols[s->cur];
    int* ic = s->icols[s->cur];
    k_sys_prep<<<1, 256, 0, st>>>(nsys, s->sys_cnt, s->meta);
    k_fill_u32<<<nblocks(nbuckets, 256), 256, 0, st>>>(s->cnt, 0u, nbuckets);
    k_keys<<<nb, 256, 0, st>>>(s->sysd, nsys, ncells, c + C_X * cap, c + C_Y * cap,
                               c + C_Z * cap, ic + I_SYS * cap, s->meta, ic + I_KEY * cap,
                               s->slot, s->cnt);
    scan(s, SrcCount{s->cnt}, (int)nbuckets, s->cell_start, s->scan_part);
    k_scatter<<<nb, 256, 0, st>>>(s->meta, ic + I_KEY * cap, s->slot, ic + I_REF * cap,
                                  s->cell_start, s->tmp_idx, s->tmp_key, s->tmp_ref);
    k_rank<<<nb, 256, 0, st>>>(s->meta, s->cell_start, s->tmp_idx, s->tmp_key, s->tmp_ref,
                               s->perm, ncells + nsys);
    k_gather<<<nb, 256, 0, st>>>(s->meta, s->perm, c, s->cols[s->cur ^ 1], ic,
                                 s->icols[s->cur ^ 1], cap, skip_cols);
    s->cur ^= 1;
    int* ref = s->icols[s->cur] + I_REF * cap;
    int* sysid = s->icols[s->cur] + I_SYS * cap;
    k_finalize<<<nblocks(nsys, 64), 64, 0, st>>>(s->sysd, nsys, ncells, s->sys_cnt, s->meta,
                                                  s->cell_start);
    k_inv<<<nb, 256, 0, st>>>(s->meta, ref, sysid, s->sys_cnt, s->inv);
    k_remove_seq<<<nblocks(nsys, 64), 64, 0, st>>>(s->meta, nsys, ncells, s->sys_cnt,
                                                    s->cell_start, ref, s->inv);
    enqueue_tiles(s);
}

}  // namespace sph
