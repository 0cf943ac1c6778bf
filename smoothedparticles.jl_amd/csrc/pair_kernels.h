// Pair-interaction kernels (shared device code of pair64.hip and pair32.hip):
// the device form of _apply_binary! /
// apply_binary! (src/core.jl:94-129) for the registered operators.
//
// Work decomposition.  A tile is up to T targets of one cell (T = 8..64, a
// template parameter).  One 64-lane wave handles a tile: lane = (target t,
// slice s), t = lane % T, s = lane / T, with S = 64 / T slices.
//
// Candidates.  The candidates of a target are the cells key + key_diff of
// the reference; with di innermost, each of the 3 (2D) or 9 (3D) (dj, dk)
// rows is the contiguous cell range [key + L1*dj + L1*L2*dk - 1, ... + 1],
// clamped to [0, key_max) exactly as the reference's linear bounds test
// (src/core.jl:98), so rows wrap and small grids visit aliased cells as
// often as the reference does.  Since the particles are cell-sorted, each
// row is one contiguous particle range (periodic systems: at most two, with
// an image shift each).  All targets of a tile share the cell, hence the
// ranges: the tile's candidate list is the concatenation of its ranges.
//
// Culling.  Only the candidates within reach of the tile -- squared
// distance to the bounding box of its targets not above the predicate's
// threshold -- can be accepted by any target, about half of the list in 3D.
// The list is streamed in 64-candidate steps; each step appends its kept
// candidates (ballot + lane rank) to a ring of 128 LDS slots, and every 64
// kept candidates form a chunk.  The kept order is the list order, and the
// box and the cull test are the same in every pass, so all passes see the
// same chunks.  The testing pass also records each tile's kept candidates
// (and their ranges, for the image shifts of periodic systems); a pass that
// loads its masks reads that list instead of streaming and culling again,
// one chunk per step, so its ring holds a single chunk (64 slots).
//
// Each chunk is handled in two phases:
//   1. mask: lane (t, s) tests the chunk slots s + S*j (j < T) against
//      target t and sets bit j of its T-bit mask for slot s + S*j -- or,
//      after a density pass on the same cell list, loads that pass's mask.
//      Interleaving the slices keeps the accepted counts of the lanes
//      alike (neighbouring slots are neighbouring particles);
//   2. interact: the chunk's accepted pairs are cut into runs spread over
//      all 64 lanes (balanced pairs, below), so the pair arithmetic runs
//      once per accepted pair, not once per candidate.
// At the end the S slices of a target are combined with wave shuffles.
// The masks of target t for a chunk form one 64-bit word (the T-bit mask
// of slice s in bits [s*T, s*T + T)), stored at
// mask_base(tile) + chunk * T + t.
//
// Predicate (src/core.jl:104-105, src/algebra.jl:49-60): accept q != p when
// RN(sqrt((d1*d1 + d2*d2) + d3*d3)) <= h.  sqrt is correctly rounded and
// monotonic, so this equals s <= s_thr with s_thr = max{s : RN(sqrt(s)) <= h}
// (computed on the host).  The FP64 kernels evaluate s without FMA
// (-ffp-contract=off), i.e. bit-exactly as the reference.  The mixed
// kernels evaluate |q|^2 - 2 p.q + |p|^2 in FP32 on tile-relative
// coordinates (three FMAs per candidate): below s_thr (1 - BAND) the pair is
// accepted, above s_thr (1 + BAND) rejected, and in between the exact FP64
// test decides, so their neighbour sets (and masks) are the reference's
// too.  The box test never drops a candidate that the predicate accepts:
// in FP64 each box distance component is a correctly rounded difference no
// larger than the target's, and the mixed test uses the same guard band.
// The identity exclusion p != q is applied once per chunk to the slots that
// hold p: the tile's own targets are always kept, so each occurrence of the
// tile's cell in the list puts them in consecutive slots.
#pragma once

#include <type_traits>

#include "sph_internal.h"

namespace sph {

#define COL(c, f, i) ((c).cols[(long long)(f) * (c).cap + (i)])

// Guard band of the mixed predicate, relative to s_thr.  Tile-relative
// coordinates are below 4 h in magnitude; the FP32 expanded form
// |q|^2 - 2 p.q + |p|^2 is then within ~1e-5 * h^2 of the exact squared
// distance and the direct form within ~1e-6 * h^2; the band is 10x wider.
static constexpr double BAND = 1.0e-4;

__device__ inline void count_pairs(const PairCtx& c, unsigned n)
{
    // wave reduction of the per-lane counts, one atomic per wave
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) n += __shfl_xor(n, o);
    if (__lane_id() == 0 && n) atomicAdd(c.pair_counter, (unsigned long long)n);
}

// The wave's distance tests (the same value in every lane).
__device__ inline void count_tests(const PairCtx& c, unsigned long long n)
{
    if (c.test_counter && __lane_id() == 0 && n) atomicAdd(c.test_counter, n);
}

// Tile queues.  The tile range [0, ntiles) is cut into NQUEUE contiguous
// parts, one per queue head.  A wave starts on queue blockIdx % NQUEUE (the
// XCD it runs on, so neighbouring tiles share that XCD's L2), takes qgrab
// tiles per atomic, and when its queue is empty moves on to the next ones.
// With one tile per grab the tiles in flight on an XCD are consecutive and
// share their neighbour cells in the L2 (3D: 9.6 instead of 10.6 ms per
// dam-break step); the short 2D tiles take four per grab, which spares
// atomics.
// Every wave leaves once all queues are past their ends.  The state is
// wave-uniform (readfirstlane), so the tile record, the system id and the
// system's parameters are scalar loads held in SGPRs.
struct TileCursor {
    int q0, k, cur, end;
};

__device__ inline void tile_cursor_init(TileCursor& tc)
{
    tc.q0 = (int)(blockIdx.x % NQUEUE);
    tc.k = 0;
    tc.cur = tc.end = 0;
}

__device__ inline int next_tile(const PairCtx& c, int ntiles, TileCursor& tc)
{
    if (tc.cur < tc.end) return tc.cur++;
    while (tc.k < NQUEUE) {
        const int q = (tc.q0 + tc.k) % NQUEUE;
        const int lo = (int)((long long)ntiles * q / NQUEUE);
        const int hi = (int)((long long)ntiles * (q + 1) / NQUEUE);
        int start = 0;
        if (__lane_id() == 0) start = atomicAdd(&c.queue[q * QSTRIDE], c.qgrab);
        start = __builtin_amdgcn_readfirstlane(start);
        if (start < hi - lo) {
            tc.cur = lo + start;
            tc.end = min(hi, lo + start + c.qgrab);
            return tc.cur++;
        }
        ++tc.k;
    }
    return ntiles;
}

__device__ inline long long mask_base(const int4& td)
{
    return (long long)(unsigned)td.z | ((long long)td.w << 32);
}

// m * 2 + (e <= thr): the predicate shifted into a mask, as one compare
// into VCC and one add with carry (the compiler would spend a select, a
// shift and an OR per bit).  NaN compares false, as e <= thr does.
__device__ inline unsigned shift_in_le(unsigned m, float e, float thr)
{
    unsigned r;
    asm("v_cmp_ge_f32_e32 vcc, %2, %3\n\tv_addc_co_u32_e32 %0, vcc, %1, %1, vcc"
        : "=v"(r)
        : "v"(m), "v"(thr), "v"(e)
        : "vcc");
    return r;
}

__device__ inline unsigned shift_in_le(unsigned m, double e, double thr)
{
    unsigned r;
    asm("v_cmp_ge_f64_e32 vcc, %2, %3\n\tv_addc_co_u32_e32 %0, vcc, %1, %1, vcc"
        : "=v"(r)
        : "v"(m), "v"(thr), "v"(e)
        : "vcc");
    return r;
}

// Per-lane mask of T slots.
template <int T>
using MaskT = typename std::conditional<(T > 32), unsigned long long, unsigned>::type;

template <int T>
__device__ inline int lowest_bit(MaskT<T> m)
{
    if constexpr (T > 32) return __ffsll((long long)m) - 1;
    else return __ffs(m) - 1;
}

// ------------------------------------------------------------ candidates --
// The candidate ranges of a tile: its cell's stencil rows, concatenated.
// pref[i] candidates precede range i; a chunk's slot l holds candidate
// base + l, whose range i satisfies pref[i] <= base + l < pref[i + 1].
// The kept candidates are staged in a ring of at most 128 slots: q (and rg,
// for periodic systems) of each slot, and the kept position of the first
// target at each occurrence of the tile's cell.
template <int ROWS, bool PER, int NQ = 128>
struct Cands {
    static constexpr int MAXR = 2 * ROWS;
    static constexpr int NSH = PER ? MAXR : 1;
    int j0[MAXR];
    int pref[MAXR + 1];
    double sx[NSH], sy[NSH], sz[NSH];  // image shift (periodic systems)
    int q[NQ];                         // particle index of each ring slot (1: not kept,
                                       // in kernels whose operator does not need it)
    int rg[PER ? 128 : 1];             // range of each ring slot
    int self0[MAXR];
};

// Uniform work (every lane computes the same values); lane 0 writes.
// Returns the number of ranges; *total receives the candidate count.
template <int ROWS, bool PERIODIC, int NQ>
__device__ inline int build_cands(const SysDesc& d, const int* cell_start, int lc,
                                  Cands<ROWS, PERIODIC, NQ>& L, int& total)
{
    const bool w = __lane_id() == 0;
    int n = 0, tot = 0;
    for (int row = 0; row < ROWS; ++row) {
        Seg segs[2];
        const int ns = row_segments<ROWS, PERIODIC>(d, cell_start, lc, row, segs);
        for (int k = 0; k < 2; ++k) {
            if (k < ns) {
                if (w) {
                    L.j0[n] = segs[k].j0;
                    L.pref[n] = tot;
                    if constexpr (PERIODIC) {
                        L.sx[n] = segs[k].sx;
                        L.sy[n] = segs[k].sy;
                        L.sz[n] = segs[k].sz;
                    }
                }
                tot += segs[k].j1 - segs[k].j0;
                ++n;
            }
        }
    }
    if (w) L.pref[n] = tot;
    total = tot;
    return n;
}

// Slot of chunk `base` that holds candidate g (g < total): its range and
// particle index.  Ranges are scanned from `cur`, the range of slot 0.
template <int ROWS, bool PER, int NQ>
__device__ inline int slot_range(const Cands<ROWS, PER, NQ>& L, int cur, int g)
{
    int i = cur;
    while (L.pref[i + 1] <= g) ++i;
    return i;
}

// Range of the next chunk's slot 0.
template <int ROWS, bool PER, int NQ>
__device__ inline int advance_range(const Cands<ROWS, PER, NQ>& L, int nr, int cur, int next_base)
{
    while (cur < nr && L.pref[cur + 1] <= next_base) ++cur;
    return cur;
}

template <int ROWS, int NQ>
__device__ inline Seg range_shift(const Cands<ROWS, true, NQ>& L, int i)
{
    Seg g;
    g.j0 = g.j1 = 0;
    g.sx = L.sx[i];
    g.sy = L.sy[i];
    g.sz = L.sz[i];
    return g;
}

// Lanes below this one with their bit set in m.
__device__ inline int lane_rank(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Append the lanes' kept candidates to the ring (list order).  fill: kept so
// far; nself: occurrences of the first target so far.  Returns the lane's
// kept position (its ring slot is pos & 127), -1 if its candidate is not
// kept.
template <int ROWS, bool PER, int NQ>
__device__ inline int compact(Cands<ROWS, PER, NQ>& L, int& fill, int& nself, bool keep, int q,
                              int rg, int first)
{
    const unsigned long long km = __ballot(keep);
    const unsigned long long sm = __ballot(keep && q == first);
    const int pos = fill + lane_rank(km);
    if (keep) {
        const int r = pos & 127;
        if constexpr (NQ > 1) L.q[r] = q;
        if constexpr (PER) L.rg[r] = rg;
        if (q == first) L.self0[nself + lane_rank(sm)] = pos;
    }
    fill += __popcll(km);
    nself += __popcll(sm);
    return keep ? pos : -1;
}

// Clear the bits of the chunk's slots (kept positions head ... head + 63)
// that hold the target itself (the reference skips p === q,
// src/core.jl:105): target t sits t slots after the first target at every
// occurrence of the tile's cell.
template <int ROWS, bool PER, int T, int NQ>
__device__ inline void drop_self(const Cands<ROWS, PER, NQ>& L, int nself, int head, int t, int s,
                                 MaskT<T>& m)
{
    constexpr int S = 64 / T;
    for (int k = 0; k < nself; ++k) {
        const int l = L.self0[k] + t - head;
        if (l >= 0 && l < 64 && l % S == s) m &= ~(MaskT<T>(1) << (l / S));
    }
}

// Bounding box of a tile's targets, computed with the tile table in absolute
// FP64 coordinates; the mixed kernels convert it to FP32 relative to the
// tile's first target.
template <class V>
struct Box {
    V x0, x1, y0, y1, z0, z1;
};

// Squared distance of (x, y, z) to the box (0 inside), evaluated in the
// predicate's operation order.
template <bool FLAT>
__device__ inline double box_dist2(const Box<double>& b, double x, double y, double z)
{
    const double dx = fmax(fmax(b.x0 - x, x - b.x1), 0.0);
    const double dy = fmax(fmax(b.y0 - y, y - b.y1), 0.0);
    double ss = dx * dx + dy * dy;
    if (!FLAT) {
        const double dz = fmax(fmax(b.z0 - z, z - b.z1), 0.0);
        ss = ss + dz * dz;
    }
    return ss;
}

template <bool FLAT>
__device__ inline float box_dist2(const Box<float>& b, float x, float y, float z)
{
    const float dx = fmaxf(fmaxf(b.x0 - x, x - b.x1), 0.0f);
    const float dy = fmaxf(fmaxf(b.y0 - y, y - b.y1), 0.0f);
    const float dz = FLAT ? 0.0f : fmaxf(fmaxf(b.z0 - z, z - b.z1), 0.0f);
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// ------------------------------------------------------ balanced pairs --
// The accepted pairs of a chunk are spread over all 64 lanes instead of
// being walked by the lane that owns the target: the per-lane counts of a
// chunk differ a lot (targets near the chunk's candidates accept many, far
// ones none), and an owner-walked loop runs as long as the busiest lane.
// Per chunk:
//   - the owner lane of target t (slice 0, lane t) publishes the target's
//     64-bit word of accepted slots in LDS;
//   - the pairs are cut into runs of at most B pairs, one target per run,
//     with B a quota from the mean up for which the runs fit in 64 lanes
//     (or, when every target is one run, each owner lane walks its own);
//     lane l takes run l;
//   - the run split is wave-uniform work (DPP scans, ballots, a table
//     lookup), and a lane finds its run from a packed mark and a prefix
//     maximum (split_runs);
//   - a lane reads its target's data from the owner lane (shuffles), skips
//     to its run with a popcount descent and walks the run;
//   - its partial sums go to LDS, and the owner folds the runs of its target
//     in run order, so the result does not depend on timing.
// The words are exact (the mixed test settles its guard band before), so
// the runs -- and the summation order -- are the same whether a pass tests
// the candidates or loads the density pass's words.
template <class Part>
struct Bal {
    unsigned long long w[64];  // accepted slots of each target
    int mark[64];              // run starts: (target + 1, its pop) packed, else 0
    Part part[64];             // each lane's partial sums of the chunk
};

// Bit b of a target's word is slot (b % T) * (64 / T) + b / T of the chunk.
template <int T>
__device__ inline int slot_of_bit(int b)
{
    return (b % T) * (64 / T) + b / T;
}

// w without its n lowest set bits (0 <= n < popcount(w)): the n-th set bit
// is found in the 32-bit half that holds it, halving the field each step.
__device__ inline unsigned long long drop_low_bits(unsigned long long w, int n)
{
    if (n <= 0) return w;
    const unsigned lo = (unsigned)w;
    const int clo = __popc(lo);
    const bool up = n >= clo;
    unsigned x = up ? (unsigned)(w >> 32) : lo;
    n = up ? n - clo : n;
    int pos = up ? 32 : 0;
    int c = __popc(x & 0xffffu);
    if (n >= c) { n -= c; x >>= 16; pos += 16; }
    c = __popc(x & 0xffu);
    if (n >= c) { n -= c; x >>= 8; pos += 8; }
    c = __popc(x & 0xfu);
    if (n >= c) { n -= c; x >>= 4; pos += 4; }
    c = __popc(x & 0x3u);
    if (n >= c) { n -= c; x >>= 2; pos += 2; }
    if (n >= (int)(x & 1u)) pos += 1;
    return w & (~0ull << pos);
}

// Largest v over the wave, 0 <= v < 128, bit by bit from the top: one
// compare per bit, the result is uniform.
__device__ inline int wave_max7(int v)
{
    int r = 0;
#pragma unroll
    for (int b = 6; b >= 0; --b) {
        const int c = r | (1 << b);
        if (__ballot(v >= c) != 0ull) r = c;
    }
    return r;
}

// Copy of v as held by lane src.
template <class V>
__device__ inline V shfl_value(const V& v, int src)
{
    static_assert(sizeof(V) % 4 == 0, "shuffled per 32-bit word");
    V r;
    const int* a = reinterpret_cast<const int*>(&v);
    int* b = reinterpret_cast<int*>(&r);
#pragma unroll
    for (unsigned i = 0; i < sizeof(V) / 4; ++i) b[i] = __shfl(a[i], src);
    return r;
}

// Inclusive prefix sum of v over the wave's lanes, in six DPP steps: the
// row shifts by 1, 2, 4 and 8 lanes scan each row of 16 lanes, then the
// broadcast of lane 15 carries row 0's total into row 1 and row 2's into
// row 3, and the broadcast of lane 31 carries rows 0-1's total into rows 2
// and 3.  Lanes a step does not write add the 0 given as the old value.
// Every lane of the wave must be active.
__device__ inline int wave_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Sum of v over the wave (uniform).
__device__ inline int wave_total(int v) { return __builtin_amdgcn_readlane(wave_scan(v), 63); }

// Inclusive prefix maximum of v >= 0 over the wave's lanes: the steps of
// wave_scan with max for the sum.
__device__ inline int wave_max_scan(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true));   // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true));   // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true));   // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true));   // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return v;
}

// ceil(pop / B) for 0 <= pop <= 64 and a uniform 1 <= B <= 65, branch-free:
// floor((pop + B - 1/2) / B), a value at least 1 / (2B) away from every
// integer, computed as pop * (1/B) + (B - 1/2) / B with one FMA.  The two
// factors are rounded to FP32 at compile time and read from a table with a
// scalar load (B is uniform); the FMA's error stays below 3e-5 / B, so the
// truncation gives the exact quotient.
struct RunDiv {
    float r, c;
};

#define SPH_RUN_DIV(b) {1.0f / (float)(b), ((float)(b) - 0.5f) / (float)(b)}
#define SPH_RUN_DIV8(b)                                                                        \
    SPH_RUN_DIV(b), SPH_RUN_DIV(b + 1), SPH_RUN_DIV(b + 2), SPH_RUN_DIV(b + 3),                \
        SPH_RUN_DIV(b + 4), SPH_RUN_DIV(b + 5), SPH_RUN_DIV(b + 6), SPH_RUN_DIV(b + 7)
static __constant__ RunDiv kRunDiv[72] = {{0.0f, 0.0f},     SPH_RUN_DIV(1),     SPH_RUN_DIV(2),
                                          SPH_RUN_DIV(3),   SPH_RUN_DIV(4),     SPH_RUN_DIV(5),
                                          SPH_RUN_DIV(6),   SPH_RUN_DIV(7),     SPH_RUN_DIV8(8),
                                          SPH_RUN_DIV8(16), SPH_RUN_DIV8(24),   SPH_RUN_DIV8(32),
                                          SPH_RUN_DIV8(40), SPH_RUN_DIV8(48),   SPH_RUN_DIV8(56),
                                          SPH_RUN_DIV8(64)};
#undef SPH_RUN_DIV8
#undef SPH_RUN_DIV

__device__ inline RunDiv run_div(int B) { return kRunDiv[B]; }

__device__ inline int runs_of(int pop, RunDiv d) { return (int)fmaf((float)pop, d.r, d.c); }

// Runs of the chunk.  pop: accepted pairs of the lane's target (owner lanes
// only, 0 elsewhere).  Returns 0 when no target has pairs, else a positive
// value.  When the run length is the largest pop, every target is one run
// and the runs need no moving: own is set and each owner lane walks its own
// target (tt, r0 and cnt are then unused).  That is always so when all 64
// lanes hold a target with pairs, since a shorter run length splits the
// largest and needs 65 runs.  Otherwise sets the lane's run (target tt,
// first rank r0, count cnt) and, for owner lanes, the first run and the
// number of runs of their target.  The wave sums are DPP scans and the
// comparisons with the largest pop are ballots (pop > B on some lane), so
// the split moves no data through LDS until the runs are assigned.
template <class Part>
__device__ inline int split_runs(int lane, int pop, Bal<Part>& bal, bool& own, int& tt, int& r0,
                                 int& cnt, int& first_run, int& nruns)
{
    const unsigned long long with_pairs = __ballot(pop > 0);
    own = false;
    tt = -1;
    r0 = cnt = first_run = nruns = 0;
    if (with_pairs == 0ull) return 0;
    if (with_pairs == ~0ull) {
        own = true;
        return 1;
    }
    const int tot = wave_total(pop);
    // The run length B: the mean (rounded up) or one more if the runs then
    // fit in 64 lanes; else at most two bisection steps between that and
    // ceil(tot / (64 - nz)) (nz: lanes with pairs), which always fits since
    // ceil(pop / B) <= pop / B + 1.  The number of runs does not grow with
    // B; at B >= max(pop) every target is one run.  A full search for the
    // smallest B costs more tests than it saves pairs on sparse chunks.
    int B = (tot + 63) >> 6;  // <= 64: pop <= 64 on each lane
    const int B0 = B;
    // the runs at B and at B + 1 in one scan (16-bit halves: each sum <= 4096,
    // so the low half never carries into the high one); its prefixes serve the
    // run assignment below when B ends at one of the two
    const int n0 = runs_of(pop, run_div(B)), n1 = runs_of(pop, run_div(B + 1));
    const int scan01 = wave_scan(n0 | (n1 << 16));
    const int both = __builtin_amdgcn_readlane(scan01, 63);
    if ((both & 0xffff) > 64 && __ballot(pop > B) != 0ull) {
        ++B;
        if ((both >> 16) > 64 && __ballot(pop > B) != 0ull) {
            const int nz = __popcll(with_pairs);  // < 64: not every lane has pairs
            int lo = B + 1, hi = min((tot + (64 - nz) - 1) / (64 - nz), wave_max7(pop));
            for (int it = 0; it < 2 && lo < hi; ++it) {
                const int mid = (lo + hi) >> 1;
                if (wave_total(runs_of(pop, run_div(mid))) > 64)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            B = hi;
        }
    }
    if (__ballot(pop > B) == 0ull) {  // B >= max(pop)
        own = true;
        return tot;
    }
    int incl;
    if (B == B0) {
        nruns = n0;
        incl = scan01 & 0xffff;
    } else if (B == B0 + 1) {
        nruns = n1;
        incl = scan01 >> 16;
    } else {
        nruns = runs_of(pop, run_div(B));
        incl = wave_scan(nruns);
    }
    first_run = incl - nruns;
    const int used = __builtin_amdgcn_readlane(incl, 63);  // <= 64
    // The lane where a target's runs start gets a mark (target + 1, pop);
    // a lane's run belongs to the last mark at or below it.  The mark, with
    // its lane index added, goes through a prefix maximum: the target
    // numbers grow with the start lanes, so the maximum is that last mark.
    bal.mark[lane] = 0;
    __syncthreads();
    if (nruns > 0 && first_run < 64) bal.mark[first_run] = ((lane + 1) << 16) | (pop << 8);
    __syncthreads();
    int mk = bal.mark[lane];
    if (mk != 0) mk |= lane;
    mk = wave_max_scan(mk);
    if (lane < used) {
        tt = (mk >> 16) - 1;
        r0 = (lane - (mk & 0xff)) * B;
        cnt = min(B, ((mk >> 8) & 0xff) - r0);
    }
    return tot;
}

// Slots of the LDS ring of staged candidates.  A testing pass appends the
// kept candidates of each streamed step behind those not yet worked, so it
// needs two chunks' room; the kept-list kernel stages one chunk per step and
// works it before the next, so one chunk's room does (less LDS, more waves).
template <bool KEPT>
constexpr int ring_slots()
{
    return KEPT ? 64 : 128;
}

// Entries of the candidate-index ring: the kept-list kernel of an operator
// that does not read the index keeps none (LDS for one more wave per SIMD).
template <class Op, bool KEPT>
constexpr int cands_q()
{
    return KEPT && !Op::USES_Q ? 1 : 128;
}

// ------------------------------------------------------------- staging --
// A candidate's fields as loaded from the columns (the image shift already
// added to the position).  A kernel streams its candidates 64 at a time:
// each lane loads one into registers, and the kept ones are converted and
// written to the LDS ring in the same step (the other waves on the SIMD
// cover the loads' latency).  Each operator fills (fetch) and writes (put)
// only the fields it uses.
struct Raw {
    double x, y, z, a0, a1, a2, a3, a4;
};

// Slots past the end of the list: far away, rejected by every test.
__device__ inline Raw pad_raw()
{
    Raw r{};
    r.x = r.y = r.z = 1.0e30;
    return r;
}

template <bool P>
__device__ inline void fetch_pos(const PairCtx& c, int q, const Seg& g, Raw& r)
{
    r.x = shifted<P>(COL(c, C_X, q), g.sx);
    r.y = shifted<P>(COL(c, C_Y, q), g.sy);
    r.z = shifted<P>(COL(c, C_Z, q), g.sz);
}

// Fetch candidate g of the list (a pad past its end) into registers; cur is
// the range of the chunk's first slot.  rq, rg: its particle index and range.
template <int ROWS, bool PERIODIC, class Op, int NQ>
__device__ inline void fetch_slot(const PairCtx& c, const Cands<ROWS, PERIODIC, NQ>& L, int total,
                                  int cur, int g, Raw& raw, int& rq, int& rg)
{
    if (g < total) {
        const int i = slot_range(L, cur, g);
        const int q = L.j0[i] + (g - L.pref[i]);
        Seg sh{};
        if constexpr (PERIODIC) sh = range_shift(L, i);
        raw = Op::template fetch<PERIODIC>(c, q, sh);
        rq = q;
        rg = i;
    } else {
        raw = pad_raw();
        rq = 0;
        rg = 0;
    }
}

// Neighbour-set digest of each target: the number of accepted q and the sum
// modulo 2^64 of mix64(refidx q + 1), written to pair_out[2 refidx p] and
// pair_out[2 refidx p + 1].  Integer sums, so the result does not depend on
// the order in which the pairs are met.
__device__ inline unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__device__ inline void digest_write(const PairCtx& c, int p, unsigned long long h,
                                    unsigned long long n, int T, bool write)
{
    for (int o = T; o < 64; o <<= 1) {
        h += __shfl_xor(h, o);
        n += __shfl_xor(n, o);
    }
    if (!write) return;
    const long long r = c.refidx[p];
    c.pair_out[2 * r] = (long long)n;
    c.pair_out[2 * r + 1] = (long long)h;
}

}  // namespace sph
