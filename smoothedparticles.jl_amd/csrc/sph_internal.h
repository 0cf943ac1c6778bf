// Internal definitions shared by the HIP translation units of libsph_hip.
//
// One handle (struct sph_system) holds a batch of independent particle
// systems packed into one set of device arrays; the single-system API of
// sph.h is a batch of one.
//
// Device layout:
//   cols[buf]   NCOL FP64 columns of `cap` entries each, in cell-sorted order:
//               x y z | vx vy vz | ax ay az | rho P Drho rho0 type
//   icols[buf]  refidx (index of the particle in its system's reference
//               order, src/core.jl), key (global cell key) and sys (system id)
//   Two buffers: the cell build gathers buffer `cur` into `cur^1` and flips.
//   sysd        one SysDesc per system: cell geometry, parameters, the
//               system's range of global cell keys
//   meta        small int32 counters read by kernels (particle count, tiles)
//
// Global cell keys: system s owns keys [cell_base_s, cell_base_s + key_max_s);
// after all cells come one removal bucket per system and one bucket for
// particles of retired systems.  Sorting by global key therefore keeps every
// system's particles contiguous and cell-sorted.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "sph.h"

namespace sph {

enum Col {
    C_X = 0, C_Y, C_Z, C_VX, C_VY, C_VZ, C_AX, C_AY, C_AZ,
    C_RHO, C_P, C_DRHO, C_RHO0, C_TYPE, NCOL
};
enum ICol { I_REF = 0, I_KEY, I_SYS, NICOL };

enum Meta {
    META_N = 0,       // active particle count (all systems)
    META_NTOT = 1,    // count at the start of the last build
    META_NREM = 2,    // particles removed by the last build
    META_NTILES = 3,  // tiles of the pair kernels
    META_NCELLS = 4,  // non-empty cells
    META_QUEUE = 5,   // (unused; the pair kernels' tile queues are separate)
    META_MASK_OK = 6, // the pair-mask buffer holds this cell list's masks
    META_KEPT_OK = 7, // the kept-candidate buffer holds this cell list's lists
    META_SLOTS = 8
};

// Per-system counters on the device (SYSCNT_* ints per system).
enum SysCnt { SC_NTOT = 0, SC_N = 1, SC_NREM = 2, SC_OFF = 3, SC_REFBASE = 4, SYSCNT = 5 };

// Cell geometry of the reference (src/structs.jl:63-82) plus the exact
// squared-distance threshold of the pair predicate.
struct Geo {
    double h;        // neighbour radius
    double inv_h;
    double s_thr;    // largest s with RN(sqrt(s)) <= h (src/core.jl:105)
    double bmin[3], bmax[3];
    long long kp[3]; // key_phase
    int L1, L2, L3;  // key_lim
    int key_max;
    int dim;         // 2 when key_lim[3] == 1
    int flat;        // bmin[2] == bmax[2]: every particle has the same z
    int periodic;    // 1: periodic in x and y (and z in 3D), minimum image
    int pad;
    double box[3];   // periodic box lengths (bmax - bmin)
    double cell[3];  // periodic cell lengths box / L (>= h)
};

// Parameters of the pair and unary kernels (derived from sph_params).
struct Phys {
    double m, rho0, c, alpha, hs, mu, nu;
    double g[3];
    double dt, coeff;
    double kW;    // m * C_W / h^d         (W = C_W (1-x)^4 (1+4x) / h^d)
    double kdW;   // m * C_dW / h^(d+2)    (rDW = C_dW (1-x)^3 / h^(d+2))
    double avk;   // -alpha * c * hs
    double eps;   // 0.01 * hs^2
};

// One system of a batch.
struct SysDesc {
    Geo g;
    Phys ph;
    int cell_base;   // first global cell key
    int live;        // 0 once retired: its particles go to the dead bucket
    int pass_on;     // 1: passes with only_flagged set write this system
    int pad;
};

// Pair-mask modes of the pair kernels.  A mask is the 64-bit set of
// accepted candidates of one target in one 64-candidate chunk; the density
// pass records them and the passes after it on the same cell list read
// them instead of re-testing the candidates.
enum MaskMode { MASK_TEST = 0, MASK_TEST_STORE = 1, MASK_LOAD = 2 };

struct PairCtx {
    const SysDesc* sys;       // device array, one per system
    const int* cell_sys;      // global cell -> system
    int T;                    // targets per tile (8..64, power of two)
    int cap;
    const int4* tiles;        // {global cell, first target, mask base lo, mask base hi}
    unsigned long long* masks;
    int* kept;                // kept candidates of each tile (64 per mask chunk)
    unsigned char* kept_rg;   // their candidate ranges (periodic systems: image shift)
    int* kept_n;              // kept candidates per tile
    const double* boxes;      // bounding box of each tile's targets (6 doubles)
    int mask_mode;
    int masks_mixed;          // the masks were stored by a mixed-precision pass
    const int* cell_start;
    const int* meta;
    int* queue;               // NQUEUE tile-queue heads, QSTRIDE ints apart
    int qgrab;                // tiles taken per queue atomic
    double* cols;             // current column buffer
    const int* refidx;
    const double* prho2;      // P/rho^2 per particle (force pass)
    unsigned long long* pair_counter;
    unsigned long long* test_counter;  // distance tests of a testing pass (may be null)
    unsigned long long* emit_counter;
    long long* pair_out;      // pair emission (test hook)
    long long pair_cap;
    int self;                 // add the self term (density)
    int overwrite;            // write instead of accumulate
    int only_flagged;         // skip systems whose pass_on is 0
};

// Tile queues of the pair kernels: the tiles are split into NQUEUE
// contiguous ranges with one head counter each (on its own 128-byte line);
// workgroup b starts on queue b % NQUEUE (the XCD it runs on) and takes
// PairCtx::qgrab tiles per atomic, then helps the other queues.
static constexpr int NQUEUE = 8;
static constexpr int QSTRIDE = 32;

// Binary operator ids inside the library.
enum PairOp {
    P_DENSITY = 0,       // rho  += m W
    P_DENSITY_RHO0 = 1,  // rho0 += m W
    P_FORCE = 2,         // internal force + AV + laminar
    P_FORCE_RHO0 = 3,    // pressure force with rho0^2
    P_CONTINUITY = 4,    // balance_of_mass!
    P_COUNT = 5,         // accepted pair count
    P_EMIT = 6,          // pair emission
    P_DIGEST = 7,        // per-target neighbour count and hash
    NPAIROP
};

// A lattice covering for the device generator (generate.hip).
struct LatticeGeo {
    int dim, kind, lid_axis, nball;
    double dr, lid;
    double lo[3], hi[3];          // the box (SPH_LATTICE_BOX) or the tank
    long long r0[3], ext[3];      // covering index range: first index and extent per axis
};

// Unary stages of the collapse_dry time loop (k_cd_unary).
enum CdStage { CD_KICK_MOVE = 0, CD_MOVE = 1, CD_KICK = 2 };

// One contiguous candidate range of a stencil row, with the periodic image
// shift added to the candidates' coordinates.
struct Seg {
    int j0, j1;
    double sx, sy, sz;
};

// Load through the constant address space: for tables a kernel only reads
// (descriptors, cell starts, tiles), so that wave-uniform loads become
// scalar loads the compiler may repeat instead of keeping the values live in
// vector registers.
#define SPH_CONST_AS __attribute__((address_space(4)))
template <class T>
__device__ inline T ld_const(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "ld_const copies 32-bit words");
    T v;
    const SPH_CONST_AS unsigned* src = (const SPH_CONST_AS unsigned*)p;
    unsigned* dst = reinterpret_cast<unsigned*>(&v);
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) dst[i] = src[i];
    return v;
}

// x + s for periodic systems (s: image shift); x itself otherwise, where
// the shift is zero and x + 0.0 would only turn -0.0 into +0.0.
template <bool PERIODIC>
__device__ inline double shifted(double x, double s)
{
    if constexpr (PERIODIC) return x + s;
    else return x;
}

// The stencil rows of local cell `lc` of system `d`.
//
// Reference systems: row r = (dj, dk) holds the cells
// key + L1*dj + L1*L2*dk + di for di = -1..1 (src/structs.jl:70-82), which
// are consecutive, so after the linear bounds test 1 <= key <= key_max of
// src/core.jl:98 a row is one contiguous particle range; rows wrap and
// small grids alias exactly as in the reference.
//
// Periodic systems (not in the reference): rows wrap per axis with the
// minimum-image shift; a row splits into at most two ranges at the x seam.
// Every periodic axis has at least 3 cells, so no cell is visited twice.
// Returns the number of ranges (0..2) written to seg.
// PERIODIC = false compiles the reference rows only (a batch without
// periodic systems).
template <int ROWS, bool PERIODIC = true>
__device__ inline int row_segments(const SysDesc& d, const int* cell_start, int lc, int row,
                                   Seg* seg)
{
    const Geo& g = d.g;
    const int dj = row % 3 - 1;
    const int dk = ROWS == 3 ? 0 : row / 3 - 1;
    const int base = d.cell_base;
    if (!PERIODIC || !g.periodic) {
        const long long mid = (long long)lc + (long long)g.L1 * dj + (long long)g.L1 * g.L2 * dk;
        long long lo = mid - 1, hi = mid + 1;
        if (lo < 0) lo = 0;
        if (hi > g.key_max - 1) hi = g.key_max - 1;
        if (lo > hi) return 0;
        seg[0].j0 = ld_const(cell_start + base + lo);
        seg[0].j1 = ld_const(cell_start + base + hi + 1);
        seg[0].sx = seg[0].sy = seg[0].sz = 0.0;
        return seg[0].j0 < seg[0].j1 ? 1 : 0;
    }
    const int ci = lc % g.L1;
    const int cj = (lc / g.L1) % g.L2;
    const int ck = lc / (g.L1 * g.L2);
    int j = cj + dj, k = ck + dk;
    double sy = 0.0, sz = 0.0;
    if (j < 0) { j += g.L2; sy = -g.box[1]; }
    else if (j >= g.L2) { j -= g.L2; sy = g.box[1]; }
    if (ROWS == 9) {
        if (k < 0) { k += g.L3; sz = -g.box[2]; }
        else if (k >= g.L3) { k -= g.L3; sz = g.box[2]; }
    }
    const int rowbase = base + g.L1 * (j + g.L2 * k);
    // L1 >= 3: at most one of the two seams is crossed.  The outputs are
    // written with fixed indices (selects), so they stay in registers.
    const bool left = ci == 0, right = ci == g.L1 - 1;
    const int lo = left ? 0 : ci - 1;
    const int hi = right ? g.L1 - 1 : ci + 1;
    Seg m, q;
    m.j0 = ld_const(cell_start + rowbase + lo);
    m.j1 = ld_const(cell_start + rowbase + hi + 1);
    m.sx = 0.0; m.sy = sy; m.sz = sz;
    q.j0 = q.j1 = 0;
    q.sx = 0.0; q.sy = sy; q.sz = sz;
    if (left) {  // cell L1-1 seen through the left seam
        q.j0 = ld_const(cell_start + rowbase + g.L1 - 1);
        q.j1 = ld_const(cell_start + rowbase + g.L1);
        q.sx = -g.box[0];
    } else if (right) {  // cell 0 seen through the right seam
        q.j0 = ld_const(cell_start + rowbase);
        q.j1 = ld_const(cell_start + rowbase + 1);
        q.sx = g.box[0];
    }
    const bool has_m = m.j0 < m.j1, has_q = q.j0 < q.j1;
    seg[0] = has_m ? m : q;
    seg[1] = q;
    return (int)has_m + (int)has_q;
}

// Kernel launchers (implemented in the .hip files).
void launch_pair(int op, int precision, int rows, int flat, int periodic, const PairCtx& c,
                 int grid, hipStream_t st);
void launch_candidates(int rows, const PairCtx& c, const int* key, unsigned long long* out,
                       long long n, hipStream_t st);

}  // namespace sph

// The opaque handle of sph.h / sph_batch.h: a batch of systems.
struct sph_system {
    int precision;
    int tile_T;          // current targets per tile
    int tile_auto;       // choose tile_T from occupancy at each build
    int rows;            // 3 (2D) or 9 (3D): the same for every system of a batch
    int flat;            // every system is flat 2D
    int periodic;        // some system is periodic
    long long n;         // host mirror of meta[META_N]
    long long cap;
    bool cells_valid;
    hipStream_t stream;
    hipStream_t copy_stream;  // asynchronous frame downloads (created on first use)
    int device;          // the device the handle's memory and stream live on

    // systems
    std::vector<sph::SysDesc> sys_host;
    std::vector<long long> sys_n;      // particles per system (host mirror)
    std::vector<long long> sys_off;    // first particle of each system (valid with cells)
    sph::SysDesc* sysd;                // device copy
    long long sysd_cap;
    int* cell_sys;                     // global cell -> system
    long long cells_total;             // sum of key_max of all systems
    long long cells_cap;               // allocation of the per-cell arrays
    int* sys_cnt;                      // device: SYSCNT ints per system
    int* sys_cnt_host;                 // pinned mirror
    long long sys_cnt_cap;

    double* cols[2];
    int* icols[2];
    int cur;

    int* meta;           // device counters
    int* meta_host;      // pinned mirror
    int* queue;          // NQUEUE * QSTRIDE ints: tile-queue heads
    unsigned long long* pair_counter;

    // build scratch
    unsigned* cnt;       // buckets
    int* cell_start;     // buckets + 1
    int* tile_start;     // cells + 1, by brick rank (cell_order)
    unsigned* scan_part; // scan partials
    long long scan_part_n;
    int* slot;           // cap
    int* tmp_idx;        // cap
    int* tmp_key;        // cap
    int* tmp_ref;        // cap
    int* perm;           // cap
    int* inv;            // cap
    int4* tiles;         // tile bound
    long long tile_cap;
    int* cell_chunks;    // cells: 64-candidate chunks of a target of the cell
    int* cell_order;     // cells: global cell at each brick rank (tile-table order)
    unsigned long long* mask_start;  // cells + 1, by brick rank: first mask word
    unsigned long long* scan_part64;
    unsigned long long* masks;       // pair masks of the current cell list
    int* kept;                       // kept-candidate lists of the tiles (64 per chunk)
    long long kept_cap;              // entries allocated
    unsigned char* kept_rg;          // kept_cap entries when a system is periodic, else null
    int* kept_n;                     // tile_cap: kept candidates per tile
    double* boxes;                   // 6 * tile_cap: bounding boxes of the tiles' targets
    long long mask_cap;              // words allocated
    unsigned long long* mask_total;  // device: words needed (1 value)
    bool masks_valid;                // masks match the cell list and positions
    bool use_masks;                  // option: reuse masks across passes
    bool use_kept;                   // option: kept-candidate lists with the masks
    bool use_graph;                  // option: replay the steps of a call from a hipGraph
    double* prho2;       // cap
    double* stage;       // 9 * cap (field IO)
    double* red_part;    // reduction partials
    long long red_part_n;
    double* red_out;     // 3 reduction results
    unsigned long long step_pairs[3];  // density pairs, force pairs, density tests  // pair counts of the last sph_step

    // batch bookkeeping (host only)
    std::vector<int> slot_of_id;       // stable system id -> slot (-1 once compacted out)
    std::vector<int> id_of_slot;       // slot -> stable id
    std::vector<char> rho0_pending;    // slot: reference density not yet initialised
    bool retired_since_build;          // retired particles are still in their cells
    long long bytes;                   // device bytes held by the handle

    // profiling
    bool prof;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open;
    double prof_ms[4];
    long long prof_n[4];
};
