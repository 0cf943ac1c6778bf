// Internal definitions shared by the HIP translation units of libsph_hip.
//
// Device layout (one system):
//   cols[buf]   NCOL FP64 columns of `cap` entries each, in cell-sorted order:
//               x y z | vx vy vz | ax ay az | rho P Drho rho0 type
//   icols[buf]  refidx (reference index of the particle, src/core.jl order)
//               and key (0-based linear cell key, key_max = removed)
//   Two buffers: the cell build gathers buffer `cur` into `cur^1` and flips.
//   meta        small int32 counters read by kernels (particle count, tiles)
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
enum ICol { I_REF = 0, I_KEY, NICOL };

enum Meta {
    META_N = 0,       // active particle count
    META_NTOT = 1,    // count at the start of the last build
    META_NREM = 2,    // particles removed by the last build
    META_NTILES = 3,  // tiles of the pair kernels
    META_NCELLS = 4,  // non-empty cells
    META_QUEUE = 5,   // work-queue head of the pair kernels
    META_MASK_OK = 6, // the pair-mask buffer holds this cell list's masks
    META_SLOTS = 8
};

// Cell geometry of the reference (src/structs.jl:63-82) plus the exact
// squared-distance threshold of the pair predicate.
struct Geo {
    double h;        // neighbour radius = cell size
    double inv_h;
    double s_thr;    // largest s with RN(sqrt(s)) <= h (src/core.jl:105)
    double bmin[3], bmax[3];
    long long kp[3]; // key_phase
    int L1, L2, L3;  // key_lim
    int key_max;
    int dim;         // 2 when key_lim[3] == 1
    int flat;        // bmin[2] == bmax[2]: every particle has the same z
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

// Pair-mask modes of the pair kernels.  A mask is the 64-bit set of
// accepted candidates of one target in one 64-candidate chunk; the density
// pass records them and the passes after it on the same cell list read
// them instead of re-testing the candidates.
enum MaskMode { MASK_TEST = 0, MASK_TEST_STORE = 1, MASK_LOAD = 2 };

struct PairCtx {
    Geo g;
    Phys ph;
    int T;                    // targets per tile (8..64, power of two)
    int cap;
    const int4* tiles;        // {cell, first target, mask base lo, mask base hi}
    unsigned long long* masks;
    int mask_mode;
    const int* cell_start;
    const int* meta;
    int* queue;               // &meta[META_QUEUE]
    double* cols;             // current column buffer
    const int* refidx;
    const double* prho2;      // P/rho^2 per particle (force pass)
    unsigned long long* pair_counter;
    unsigned long long* emit_counter;
    long long* pair_out;      // pair emission (test hook)
    long long pair_cap;
    int self;                 // add the self term (density)
    int overwrite;            // write instead of accumulate
};

// Binary operator ids inside the library.
enum PairOp {
    P_DENSITY = 0,       // rho  += m W
    P_DENSITY_RHO0 = 1,  // rho0 += m W
    P_FORCE = 2,         // internal force + AV + laminar
    P_FORCE_RHO0 = 3,    // pressure force with rho0^2
    P_CONTINUITY = 4,    // balance_of_mass!
    P_COUNT = 5,         // accepted pair count
    P_EMIT = 6,          // pair emission
    NPAIROP
};

// Row r of the 3 (2D) or 9 (3D) stencil rows of `cell`: the cells
// key + L1*dj + L1*L2*dk + di for di = -1..1 (src/structs.jl:70-82) are
// consecutive, so a row is one contiguous particle range [j0, j1) after the
// linear bounds test 1 <= key <= key_max of src/core.jl:98.
template <int ROWS>
__device__ inline bool row_range(const Geo& g, const int* cell_start, int cell, int row, int& j0,
                                 int& j1)
{
    const int dj = row % 3 - 1;
    const int dk = ROWS == 3 ? 0 : row / 3 - 1;
    const long long mid = (long long)cell + (long long)g.L1 * dj + (long long)g.L1 * g.L2 * dk;
    long long lo = mid - 1, hi = mid + 1;
    if (lo < 0) lo = 0;
    if (hi > g.key_max - 1) hi = g.key_max - 1;
    if (lo > hi) return false;
    j0 = cell_start[lo];
    j1 = cell_start[hi + 1];
    return j0 < j1;
}

// Kernel launchers (implemented in the .hip files).
void launch_pair(int op, int precision, const PairCtx& c, int grid, hipStream_t st);
void launch_candidates(const PairCtx& c, const int* key, unsigned long long* out, long long n,
                       hipStream_t st);

}  // namespace sph

// The opaque handle of sph.h.
struct sph_system {
    sph::Geo geo;
    int precision;
    int tile_T;          // current targets per tile
    int tile_auto;       // choose tile_T from occupancy at each build
    long long n;         // host mirror of meta[META_N]
    long long cap;
    bool cells_valid;
    hipStream_t stream;

    double* cols[2];
    int* icols[2];
    int cur;

    int* meta;           // device counters
    int* meta_host;      // pinned mirror
    unsigned long long* pair_counter;

    // build scratch
    unsigned* cnt;       // key_max + 1
    int* cell_start;     // key_max + 2
    int* tile_start;     // key_max + 2
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
    int* cell_chunks;    // key_max: 64-candidate chunks of a target of the cell
    unsigned long long* mask_start;  // key_max + 2: first mask word of each cell
    unsigned long long* scan_part64;
    unsigned long long* masks;       // pair masks of the current cell list
    long long mask_cap;              // words allocated
    unsigned long long* mask_total;  // device: words needed (1 value)
    bool masks_valid;                // masks match the cell list and positions
    bool use_masks;                  // option: reuse masks across passes
    double* prho2;       // cap
    double* stage;       // 3 * cap (field IO)
    double* red_part;    // reduction partials
    long long red_part_n;
    double* red_out;     // 3 reduction results
    unsigned long long step_pairs[2];  // pair counts of the last sph_step

    // profiling
    bool prof;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open;
    double prof_ms[4];
    long long prof_n[4];
};
