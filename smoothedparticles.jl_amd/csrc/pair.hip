// Pair-interaction kernels: the device form of _apply_binary! /
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
// loads its masks reads that list instead of streaming and culling again.
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

// Compensated (2Sum) addition of a chunk partial into (s, e).
__device__ inline void two_sum_add(float& s, float& e, float x)
{
    const float t = s + x;
    const float bp = t - s;
    e += (s - (t - bp)) + (x - bp);
    s = t;
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
// The kept candidates are staged in a ring of 128 slots: q (and rg, for
// periodic systems) of each slot, and the kept position of the first
// target at each occurrence of the tile's cell.
template <int ROWS, bool PER>
struct Cands {
    static constexpr int MAXR = 2 * ROWS;
    static constexpr int NSH = PER ? MAXR : 1;
    int j0[MAXR];
    int pref[MAXR + 1];
    double sx[NSH], sy[NSH], sz[NSH];  // image shift (periodic systems)
    int q[128];                        // particle index of each ring slot
    int rg[PER ? 128 : 1];             // range of each ring slot
    int self0[MAXR];
};

// Uniform work (every lane computes the same values); lane 0 writes.
// Returns the number of ranges; *total receives the candidate count.
template <int ROWS, bool PERIODIC>
__device__ inline int build_cands(const SysDesc& d, const int* cell_start, int lc,
                                  Cands<ROWS, PERIODIC>& L, int& total)
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
template <int ROWS, bool PER>
__device__ inline int slot_range(const Cands<ROWS, PER>& L, int cur, int g)
{
    int i = cur;
    while (L.pref[i + 1] <= g) ++i;
    return i;
}

// Range of the next chunk's slot 0.
template <int ROWS, bool PER>
__device__ inline int advance_range(const Cands<ROWS, PER>& L, int nr, int cur, int next_base)
{
    while (cur < nr && L.pref[cur + 1] <= next_base) ++cur;
    return cur;
}

template <int ROWS>
__device__ inline Seg range_shift(const Cands<ROWS, true>& L, int i)
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
template <int ROWS, bool PER>
__device__ inline int compact(Cands<ROWS, PER>& L, int& fill, int& nself, bool keep, int q, int rg,
                              int first)
{
    const unsigned long long km = __ballot(keep);
    const unsigned long long sm = __ballot(keep && q == first);
    const int pos = fill + lane_rank(km);
    if (keep) {
        const int r = pos & 127;
        L.q[r] = q;
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
template <int ROWS, bool PER, int T>
__device__ inline void drop_self(const Cands<ROWS, PER>& L, int nself, int head, int t, int s,
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
// being walked by the lane that owns the target: with ~15% acceptance the
// per-lane counts of a chunk differ a lot (targets near the chunk's
// particles accept many, far ones none), and an owner-walked loop runs as
// long as the busiest lane.  Per chunk:
//   - the owner lane of target t (slice 0, lane t) publishes the target's
//     64-bit word of accepted slots in LDS;
//   - the pairs are cut into runs of at most B pairs, one target per run,
//     with B the smallest quota (from the mean up) for which the runs fit
//     in 64 lanes; lane l takes run l;
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
    int lane_t[64];            // target of each lane's run (-1: none)
    Part part[64];             // each lane's partial sums of the chunk
};

// Bit b of a target's word is slot (b % T) * (64 / T) + b / T of the chunk.
template <int T>
__device__ inline int slot_of_bit(int b)
{
    return (b % T) * (64 / T) + b / T;
}

// w without its n lowest set bits (0 <= n < popcount(w)).
__device__ inline unsigned long long drop_low_bits(unsigned long long w, int n)
{
    if (n <= 0) return w;
    unsigned long long x = w;
    int pos = 0;
    int c = __popc((unsigned)x);
    if (n >= c) { n -= c; x >>= 32; pos += 32; }
    c = __popc((unsigned)x & 0xffffu);
    if (n >= c) { n -= c; x >>= 16; pos += 16; }
    c = __popc((unsigned)x & 0xffu);
    if (n >= c) { n -= c; x >>= 8; pos += 8; }
    c = __popc((unsigned)x & 0xfu);
    if (n >= c) { n -= c; x >>= 4; pos += 4; }
    c = __popc((unsigned)x & 0x3u);
    if (n >= c) { n -= c; x >>= 2; pos += 2; }
    c = (int)((unsigned)x & 1u);
    if (n >= c) pos += 1;
    return w & (~0ull << pos);
}

__device__ inline int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
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

// Wave sum of v, 0 <= v < 128, from bit-sliced ballots: no cross-lane
// data movement through LDS, the result is uniform.
__device__ inline int wave_sum7(int v)
{
    int sum = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b) sum += __popcll(__ballot((v >> b) & 1)) << b;
    return sum;
}

// Sum of v over the lanes below this one, 0 <= v < 128.
__device__ inline int lanes_below_sum7(int v)
{
    int sum = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b) sum += lane_rank(__ballot((v >> b) & 1)) << b;
    return sum;
}

// ceil(65536 / d) for a uniform 1 <= d <= 64: a reciprocal estimate, then
// corrected to the exact value.
__device__ inline int run_magic(int d)
{
    int m = (int)(65536.0f * __builtin_amdgcn_rcpf((float)d)) + 1;
    while ((m - 1) * d >= 65536) --m;
    while (m * d < 65536) ++m;
    return m;
}

// ceil(pop / B) for 0 <= pop <= 64: with m = ceil(65536 / B) and B <= 64,
// floor(n m / 65536) = floor(n / B) for n < 1024.
__device__ inline int runs_of(int pop, int B, int m)
{
    if (pop <= 0) return 0;
    return B <= 64 ? ((pop + B - 1) * m) >> 16 : (pop + B - 1) / B;
}

// Runs of the chunk.  pop: accepted pairs of the lane's target (owner lanes
// only, 0 elsewhere).  Returns the total; sets the lane's run (target tt,
// first rank r0, count cnt) and, for owner lanes, the first run and the
// number of runs of their target.
template <class Part>
__device__ inline int split_runs(int lane, int pop, Bal<Part>& bal, int& tt, int& r0, int& cnt,
                                 int& first_run, int& nruns)
{
    const int tot = wave_sum7(pop);
    tt = -1;
    r0 = cnt = first_run = nruns = 0;
    if (tot == 0) return 0;
    int B = (tot + 63) >> 6;  // <= 64: pop <= 64 on each lane
    int m = run_magic(B);
    int need = wave_sum7(runs_of(pop, B, m));
    if (need > 64) {
        ++B;
        m = B <= 64 ? run_magic(B) : 0;
        need = wave_sum7(runs_of(pop, B, m));
        if (need > 64) {
            const int nz = __popcll(__ballot(pop > 0));
            B = nz < 64 ? (tot + (64 - nz) - 1) / (64 - nz) : wave_max(pop);
            m = B <= 64 ? run_magic(B) : 0;
        }
    }
    nruns = runs_of(pop, B, m);
    first_run = lanes_below_sum7(nruns);
    bal.lane_t[lane] = -1;
    __syncthreads();
    for (int k = 0; k < nruns; ++k) bal.lane_t[first_run + k] = lane;
    __syncthreads();
    tt = bal.lane_t[lane];
    const int src = tt < 0 ? 0 : tt;
    const int st = __shfl(first_run, src);
    const int pt = __shfl(pop, src);
    if (tt >= 0) {
        r0 = (lane - st) * B;
        cnt = min(B, pt - r0);
    }
    return tot;
}

// ------------------------------------------------------------- staging --
// A candidate's fields as loaded from the columns (the image shift already
// added to the position).  The kernels load the next chunk's candidates into
// registers before working on the current chunk and write them to LDS at
// the next iteration, so the loads' latency overlaps the tests and pairs.
// Each operator fills (fetch) and writes (put) only the fields it uses.
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
template <int ROWS, bool PERIODIC, class Op>
__device__ inline void fetch_slot(const PairCtx& c, const Cands<ROWS, PERIODIC>& L, int total,
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

// ================================================================ FP64 ====

// LDS ring of kept candidates (FP64): positions only, or xy, (z, rho),
// (vx, vy), (vz, aux) where aux is P/rho^2 (force) or P (force with rho0).
struct Lds64Pos {
    double2 xy[128];
    double z[128];
};

struct Lds64Full {
    double2 xy[128], zr[128], v01[128], v2a[128];
};

__device__ inline double lds_z(const Lds64Pos& s, int j) { return s.z[j]; }
__device__ inline double lds_z(const Lds64Full& s, int j) { return s.zr[j].x; }

// rho += m W (OUT = C_RHO) or rho0 += m W (OUT = C_RHO0).
template <int OUT>
struct Density64 {
    static constexpr bool INTERACT = true;
    static constexpr int WAVES = 5;  // occupancy the registers are fitted to
    using Lds = Lds64Pos;
    struct Tgt {
        double x, y, z;
    };
    struct Acc {
        double w;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    static __device__ void put(const Raw& r, int l, Lds& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.z[l] = r.z;
    }
    static __device__ void init(Acc& a) { a.w = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { a.w = 0.0; }
    static __device__ void fold(Acc& a, const Part& q) { a.w += q.w; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt&, const Lds&,
                                    int, int, int, double, double, double, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double x = r * d.g.inv_h;
        const double u = 1.0 - x;
        const double u2 = u * u;
        a.w += u2 * u2 * (1.0 + 4.0 * x);
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) a.w += __shfl_xor(a.w, o);
        if (!write) return;
        double v = c.overwrite ? 0.0 : COL(c, OUT, p);
        v = v + d.ph.kW * a.w;
        if (c.self) v = v + d.ph.kW;
        COL(c, OUT, p) = v;
    }
};

// internal_force! (examples/collapse_dry.jl:135-141) + Monaghan viscosity.
// With u = 1 - r/h and rDW = C_dW u^3 / h^(d+2), kdW = m C_dW / h^(d+2):
//   pressure  a += -kdW u^3 (P_p/rho_p^2 + P_q/rho_q^2) x_pq
//   viscosity a += -kdW u^3 Pi x_pq,  Pi = avk (v.x)/(rhobar (r^2 + eps)), v.x < 0
//   laminar   a += 2 kdW mu / rho0^2 u^3 v_pq
struct Force64 {
    static constexpr bool INTERACT = true;
    static constexpr int WAVES = 3;  // occupancy the registers are fitted to
    using Lds = Lds64Full;
    struct Tgt {
        double x, y, z, vx, vy, vz, pr, rho;
        bool fluid;
    };
    struct Acc {
        double ax, ay, az, lx, ly, lz;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
        t.vx = COL(c, C_VX, p); t.vy = COL(c, C_VY, p); t.vz = COL(c, C_VZ, p);
        t.pr = c.prho2[p];
        t.rho = COL(c, C_RHO, p);
        t.fluid = COL(c, C_TYPE, p) == 0.0;
    }
    static __device__ bool active(const Tgt& t) { return t.fluid; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_RHO, q);
        r.a1 = COL(c, C_VX, q); r.a2 = COL(c, C_VY, q); r.a3 = COL(c, C_VZ, q);
        r.a4 = c.prho2[q];
        return r;
    }
    static __device__ void put(const Raw& r, int l, Lds& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.zr[l] = make_double2(r.z, r.a0);
        s.v01[l] = make_double2(r.a1, r.a2);
        s.v2a[l] = make_double2(r.a3, r.a4);
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = a.lx = a.ly = a.lz = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { init(a); }
    static __device__ void fold(Acc& a, const Part& q)
    {
        a.ax += q.ax; a.ay += q.ay; a.az += q.az; a.lx += q.lx; a.ly += q.ly; a.lz += q.lz;
    }
    // variant bit 0: Monaghan viscosity (avk != 0); bit 1: laminar (mu != 0)
    static constexpr int NVAR = 4;
    static __device__ int variant(const SysDesc& d)
    {
        return (d.ph.avk != 0.0 ? 1 : 0) | (d.ph.mu != 0.0 ? 2 : 0);
    }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds& s, int jj, int, int, double dx, double dy,
                                    double dz, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double x = r * d.g.inv_h;
        const double u = 1.0 - x;
        const double u3 = u * u * u;
        const double2 v01 = s.v01[jj], v2a = s.v2a[jj];
        double A = t.pr + v2a.y;
        const double dvx = t.vx - v01.x, dvy = t.vy - v01.y, dvz = t.vz - v2a.x;
        if (V & 1) {
            double vx = dvx * dx + dvy * dy;
            vx = vx + dvz * dz;
            if (vx < 0.0) {
                const double rhobar = 0.5 * (t.rho + s.zr[jj].y);
                A = A + d.ph.avk * vx / (rhobar * (r * r + d.ph.eps));
            }
        }
        const double f = u3 * A;
        a.ax += f * dx; a.ay += f * dy; a.az += f * dz;
        if (V & 2) { a.lx += u3 * dvx; a.ly += u3 * dvy; a.lz += u3 * dvz; }
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) {
            a.ax += __shfl_xor(a.ax, o); a.ay += __shfl_xor(a.ay, o); a.az += __shfl_xor(a.az, o);
            a.lx += __shfl_xor(a.lx, o); a.ly += __shfl_xor(a.ly, o); a.lz += __shfl_xor(a.lz, o);
        }
        if (!write) return;
        const double kp = -d.ph.kdW;
        const double kl = 2.0 * d.ph.kdW * d.ph.mu / (d.ph.rho0 * d.ph.rho0);
        double ax = c.overwrite ? 0.0 : COL(c, C_AX, p);
        double ay = c.overwrite ? 0.0 : COL(c, C_AY, p);
        double az = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        ax = ax + kp * a.ax; ay = ay + kp * a.ay; az = az + kp * a.az;
        if (d.ph.mu != 0.0) { ax = ax + kl * a.lx; ay = ay + kl * a.ly; az = az + kl * a.lz; }
        COL(c, C_AX, p) = ax; COL(c, C_AY, p) = ay; COL(c, C_AZ, p) = az;
    }
};

// internal_force! of tests/test_collision_2d.jl:75-78 (P/rho0^2, all targets).
struct ForceRho064 {
    static constexpr bool INTERACT = true;
    static constexpr int WAVES = 4;  // occupancy the registers are fitted to
    using Lds = Lds64Full;
    struct Tgt {
        double x, y, z, P;
    };
    struct Acc {
        double ax, ay, az;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p); t.P = COL(c, C_P, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_P, q);
        return r;
    }
    static __device__ void put(const Raw& r, int l, Lds& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.zr[l].x = r.z;
        s.v2a[l].y = r.a0;
    }
    static __device__ void init(Acc& a) { a.ax = a.ay = a.az = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { init(a); }
    static __device__ void fold(Acc& a, const Part& q) { a.ax += q.ax; a.ay += q.ay; a.az += q.az; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds& s, int jj, int, int, double dx, double dy,
                                    double dz, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double u = 1.0 - r * d.g.inv_h;
        const double f = u * u * u * (t.P + s.v2a[jj].y);
        a.ax += f * dx; a.ay += f * dy; a.az += f * dz;
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) {
            a.ax += __shfl_xor(a.ax, o); a.ay += __shfl_xor(a.ay, o); a.az += __shfl_xor(a.az, o);
        }
        if (!write) return;
        const double k = -d.ph.kdW / (d.ph.rho0 * d.ph.rho0);
        const double ax = c.overwrite ? 0.0 : COL(c, C_AX, p);
        const double ay = c.overwrite ? 0.0 : COL(c, C_AY, p);
        const double az = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        COL(c, C_AX, p) = ax + k * a.ax;
        COL(c, C_AY, p) = ay + k * a.ay;
        COL(c, C_AZ, p) = az + k * a.az;
    }
};

// balance_of_mass! (examples/collapse_dry.jl:112-115):
//   Drho += kdW u^3 (x_pq.v_pq + 2 nu (rho_p - rho_q))
struct Continuity64 {
    static constexpr bool INTERACT = true;
    static constexpr int WAVES = 4;  // occupancy the registers are fitted to
    using Lds = Lds64Full;
    struct Tgt {
        double x, y, z, vx, vy, vz, rho;
    };
    struct Acc {
        double d;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
        t.vx = COL(c, C_VX, p); t.vy = COL(c, C_VY, p); t.vz = COL(c, C_VZ, p);
        t.rho = COL(c, C_RHO, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_RHO, q);
        r.a1 = COL(c, C_VX, q); r.a2 = COL(c, C_VY, q); r.a3 = COL(c, C_VZ, q);
        return r;
    }
    static __device__ void put(const Raw& r, int l, Lds& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.zr[l] = make_double2(r.z, r.a0);
        s.v01[l] = make_double2(r.a1, r.a2);
        s.v2a[l].x = r.a3;
    }
    static __device__ void init(Acc& a) { a.d = 0.0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { a.d = 0.0; }
    static __device__ void fold(Acc& a, const Part& q) { a.d += q.d; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds& s, int jj, int, int, double dx, double dy,
                                    double dz, double ss, Acc& a)
    {
        const double r = sqrt(ss);
        const double u = 1.0 - r * d.g.inv_h;
        const double2 v01 = s.v01[jj];
        double xv = dx * (t.vx - v01.x) + dy * (t.vy - v01.y);
        xv = xv + dz * (t.vz - s.v2a[jj].x);
        a.d += u * u * u * (xv + 2.0 * d.ph.nu * (t.rho - s.zr[jj].y));
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        for (int o = T; o < 64; o <<= 1) a.d += __shfl_xor(a.d, o);
        if (!write) return;
        const double v = c.overwrite ? 0.0 : COL(c, C_DRHO, p);
        COL(c, C_DRHO, p) = v + d.ph.kdW * a.d;
    }
};

// Accepted-pair count (COUNT) or emission of (refidx p, refidx q) (EMIT).
template <bool EMIT>
struct Pairs64 {
    static constexpr bool INTERACT = EMIT;
    static constexpr int WAVES = 5;  // occupancy the registers are fitted to
    using Lds = Lds64Pos;
    struct Tgt {
        double x, y, z;
    };
    struct Acc {
        int dummy;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    static __device__ void put(const Raw& r, int l, Lds& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.z[l] = r.z;
    }
    static __device__ void init(Acc&) {}
    using Part = Acc;
    static __device__ void zero(Part&) {}
    static __device__ void fold(Acc&, const Part&) {}
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const Lds&,
                                    int, int p, int q, double, double, double, double, Acc&)
    {
        const unsigned long long k = atomicAdd(c.emit_counter, 1ull);
        if ((long long)k < c.pair_cap) {
            c.pair_out[2 * k] = c.refidx[p];
            c.pair_out[2 * k + 1] = c.refidx[q];
        }
    }
    static __device__ void finish(const PairCtx&, const SysDesc&, int, const Tgt&, Acc&, int,
                                  bool) {}
};

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

struct Digest64 {
    static constexpr bool INTERACT = true;
    static constexpr int WAVES = 5;
    using Lds = Lds64Pos;
    struct Tgt {
        double x, y, z;
    };
    struct Acc {
        unsigned long long h, n;
    };
    static __device__ void load(const PairCtx& c, int p, Tgt& t)
    {
        t.x = COL(c, C_X, p); t.y = COL(c, C_Y, p); t.z = COL(c, C_Z, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    static __device__ void put(const Raw& r, int l, Lds& s)
    {
        s.xy[l] = make_double2(r.x, r.y);
        s.z[l] = r.z;
    }
    static __device__ void init(Acc& a) { a.h = a.n = 0; }
    using Part = Acc;
    static __device__ void zero(Part& a) { a.h = a.n = 0; }
    static __device__ void fold(Acc& a, const Part& q) { a.h += q.h; a.n += q.n; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const Lds&,
                                    int, int, int q, double, double, double, double, Acc& a)
    {
        a.h += mix64((unsigned long long)c.refidx[q] + 1ull);
        a.n += 1;
    }
    static __device__ void finish(const PairCtx& c, const SysDesc&, int p, const Tgt&, Acc& a,
                                  int T, bool write)
    {
        digest_write(c, p, a.h, a.n, T, write);
    }
};

// A lane's run: cnt accepted slots of w (lowest first) against target ht.
template <int V, bool FLAT, int T, class Op>
__device__ inline void walk64(const PairCtx& c, const SysDesc& d, const typename Op::Tgt& ht,
                              const typename Op::Lds& lds, const int* q, unsigned long long w,
                              int cnt, int rb, int hp, typename Op::Part& part)
{
    if constexpr (V < Op::NVAR) {
        for (int k = 0; k < cnt; ++k) {
            const int b = __builtin_ctzll(w);  // w != 0: cnt <= its popcount
            w &= w - 1;
            const int jj = rb + slot_of_bit<T>(b);
            const double2 xy = lds.xy[jj];
            const double dx = ht.x - xy.x, dy = ht.y - xy.y;
            const double dz = FLAT ? 0.0 : ht.z - lds_z(lds, jj);
            double ss = dx * dx + dy * dy;
            if (!FLAT) ss = ss + dz * dz;
            Op::template interact<V>(c, d, ht, lds, jj, hp, q[jj], dx, dy, dz, ss, part);
        }
    }
}

template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op>
__device__ inline void tile64(const PairCtx& c, int tile, bool load_masks, bool store_masks,
                              bool kept_ok,
                              typename Op::Lds& lds, Cands<ROWS, PERIODIC>& L,
                              Bal<typename Op::Part>& bal, unsigned& npairs,
                              unsigned long long& ntests)
{
    const int4 td = ld_const(c.tiles + tile);
    const int cell = td.x, first = td.y;
    const SysDesc d = ld_const(c.sys + ld_const(c.cell_sys + cell));
    if (c.only_flagged && !d.pass_on) return;
    const int lc = cell - d.cell_base;
    const long long mbase = mask_base(td);
    const int cend = ld_const(c.cell_start + cell + 1);
    const int nvalid = min(T, cend - first);
    const int lane = threadIdx.x;
    const int t = lane & (T - 1), s = lane / T;
    const int p = first + t;
    const bool valid = p < cend;
    typename Op::Tgt tg;
    Op::load(c, valid ? p : first, tg);
    const bool act = valid && Op::active(tg);
    if (!__any(act)) return;
    const double s_thr = d.g.s_thr;
    // a pass that loads the masks takes the kept candidates from the list the
    // testing pass wrote; otherwise it streams the candidate list and culls
    // it again
    const bool use_kept = load_masks && kept_ok;
    const bool store_kept = store_masks && kept_ok;
    const long long lbase = mbase / T * 64;
    const int nk = use_kept ? ld_const(c.kept_n + tile) : 0;
    int total = 0, nr = 0;
    Box<double> box{};
    if (!use_kept) box = ld_const(reinterpret_cast<const Box<double>*>(c.boxes) + tile);
    // (periodic systems need the ranges' image shifts in either case)
    if (!use_kept || PERIODIC) nr = build_cands<ROWS, PERIODIC>(d, c.cell_start, lc, L, total);
    __syncthreads();
    typename Op::Acc acc;
    Op::init(acc);
    int cur = 0, fill = 0, head = 0, nself = 0, chunk = 0;
    const int nstep = use_kept ? (nk + 63) / 64 : (total + 63) / 64;
    unsigned long long kword = 0;  // the chunk's mask word, loaded with its list entries
    for (int step = 0; step < nstep; ++step) {
        bool last;
        if (use_kept) {
            // one chunk of the kept list per step
            if (s == 0 && act) kword = c.masks[mbase + (long long)step * T + t];
            const int k = head + lane;
            if (k < nk) {
                const int q = c.kept[lbase + k];
                Seg sh{};
                if constexpr (PERIODIC) sh = range_shift(L, c.kept_rg[lbase + k]);
                Op::put(Op::template fetch<PERIODIC>(c, q, sh), k & 127, lds);
                L.q[k & 127] = q;
            }
            fill = min(nk, head + 64);
            last = true;
        } else {
            const int base = step * 64;
            Raw raw;
            int rq = 0, rrg = 0;
            cur = advance_range(L, nr, cur, base);
            fetch_slot<ROWS, PERIODIC, Op>(c, L, total, cur, base + lane, raw, rq, rrg);
            // keep the candidates within reach of the box, and the tile's targets
            const bool keep = base + lane < total &&
                              (box_dist2<FLAT>(box, raw.x, raw.y, raw.z) <= s_thr ||
                               (unsigned)(rq - first) < (unsigned)nvalid);
            const int pos = compact(L, fill, nself, keep, rq, rrg, first);
            if (pos >= 0) {
                Op::put(raw, pos & 127, lds);
                if (store_kept) {
                    c.kept[lbase + pos] = rq;
                    if constexpr (PERIODIC) c.kept_rg[lbase + pos] = (unsigned char)rrg;
                }
            }
            last = base + 64 >= total;
        }
        __syncthreads();
        while (fill - head >= 64 || (last && fill > head)) {
            if (fill - head < 64) {
                // the last chunk: pad its free slots
                const int k = head + lane;
                if (k >= fill) {
                    Op::put(pad_raw(), k & 127, lds);
                    L.q[k & 127] = first;
                }
                __syncthreads();
            }
            const int rb = head & 64;  // ring half of the chunk
            // accepted-slot word of target t (owner lanes publish it)
            unsigned long long word = 0;
            if (use_kept) {
                word = kword;
            } else if (load_masks) {
                if (s == 0 && act) word = c.masks[mbase + (long long)chunk * T + t];
            } else {
                // 32-bit words, slots shifted in from the highest bit down
                MaskT<T> m = 0;
#pragma unroll
                for (int h = 0; h < (T + 31) / 32; ++h) {
                    unsigned mh = 0;
#pragma unroll 4
                    for (int j = (T < 32 ? T : 32) - 1; j >= 0; --j) {
                        const int jj = rb + s + (64 / T) * (32 * h + j);
                        const double2 xy = lds.xy[jj];
                        const double dx = tg.x - xy.x, dy = tg.y - xy.y;
                        double ss = dx * dx + dy * dy;
                        if (!FLAT) {
                            const double dz = tg.z - lds_z(lds, jj);
                            ss = ss + dz * dz;
                        }
                        mh = shift_in_le(mh, ss, s_thr);
                    }
                    m |= (MaskT<T>)mh << (32 * h);
                }
                if (!act) m = 0;
                drop_self<ROWS, PERIODIC, T>(L, nself, head, t, s, m);
                word = (unsigned long long)m << (T == 64 ? 0 : s * T);
#pragma unroll
                for (int o = T; o < 64; o <<= 1) word |= __shfl_xor(word, o);
                if (s != 0) word = 0;
                if (store_masks && s == 0) c.masks[mbase + (long long)chunk * T + t] = word;
            }
            const int pop = __popcll(word);
            npairs += (unsigned)pop;
            if (s == 0) bal.w[t] = word;
            int tt, r0, cnt, frun, nrun;
            const int tot = split_runs(lane, pop, bal, tt, r0, cnt, frun, nrun);
            if (tot > 0) {
                typename Op::Part part;
                Op::zero(part);
                const typename Op::Tgt ht = shfl_value(tg, tt < 0 ? 0 : tt);
                if (Op::INTERACT && cnt > 0) {
                    const unsigned long long w = drop_low_bits(bal.w[tt], r0);
                    const int hp = first + tt;
                    // the operator's uniform switches are taken outside the loop
                    switch (Op::variant(d)) {
                    case 0: walk64<0, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 1: walk64<1, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 2: walk64<2, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    default: walk64<3, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    }
                }
                bal.part[lane] = part;
                __syncthreads();
                for (int k = 0; k < nrun; ++k) Op::fold(acc, bal.part[frun + k]);
            }
            ++chunk;
            head += 64;
            __syncthreads();
        }
    }
    // distance tests: the box test of every candidate, the pair test of
    // every kept candidate for every target
    if (!load_masks) ntests += (unsigned long long)total + (unsigned long long)fill * nvalid;
    if (store_kept && lane == 0) c.kept_n[tile] = fill;
    Op::finish(c, d, p, tg, acc, T, act && s == 0);
}

template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op>
__global__ __launch_bounds__(64, Op::WAVES) void k_pair64(PairCtx c)
{
    __shared__ typename Op::Lds lds;
    __shared__ Cands<ROWS, PERIODIC> L;
    __shared__ Bal<typename Op::Part> bal;
    const int ntiles = ld_const(c.meta + META_NTILES);
    const bool ok = c.masks != nullptr && ld_const(c.meta + META_MASK_OK) != 0;
    const bool kept_ok = ok && c.kept != nullptr && ld_const(c.meta + META_KEPT_OK) != 0;
    // masks stored by a mixed pass number the candidates its FP32 cull kept;
    // without that pass's kept lists this kernel's FP64 cull would keep
    // others, so it tests instead
    const bool load_masks = ok && c.mask_mode == MASK_LOAD && (kept_ok || !c.masks_mixed);
    const bool store_masks = ok && c.mask_mode == MASK_TEST_STORE;
    TileCursor tc;
    tile_cursor_init(tc);
    // accepted pairs of all the wave's tiles, one atomic per wave
    unsigned npairs = 0;
    unsigned long long ntests = 0;
    for (int tile = next_tile(c, ntiles, tc); tile < ntiles; tile = next_tile(c, ntiles, tc))
        tile64<ROWS, FLAT, PERIODIC, T, Op>(c, tile, load_masks, store_masks, kept_ok, lds, L, bal,
                                            npairs, ntests);
    if (c.pair_counter) count_pairs(c, npairs);
    count_tests(c, ntests);
}

// =============================================================== mixed ====
// FP32 pair arithmetic on coordinates relative to the tile origin (the FP64
// position of the tile's first target), converted while staging.  Per-lane
// partial sums are plain FP32 within a 64-candidate chunk and folded into a
// compensated (sum, error) pair after every chunk.

// The hardware square root (v_sqrt_f32, within 1 ulp) instead of the
// correctly rounded sequence the compiler emits for sqrtf (about 15
// instructions).  The pair values need FP32 accuracy, not correct rounding:
// which pairs interact is decided by the predicate, not by r.
__device__ inline float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

// LDS ring of kept candidates (FP32): (x, y, z, |x|^2) -- (x, y, z,
// P/rho^2) for the force pass -- and, for the force pass, (vx, vy, vz, rho).
struct Lds32Pos {
    float4 pos[128];
};

struct Lds32Full {
    float4 pos[128], v[128];
};

// Tile-relative FP32 position of a fetched candidate and its |.|^2.
template <bool FLAT>
__device__ inline float4 rel_pos32(const Raw& r, double ox, double oy, double oz)
{
    const float x = (float)(r.x - ox);
    const float y = (float)(r.y - oy);
    const float z = FLAT ? 0.0f : (float)(r.z - oz);
    return make_float4(x, y, z, fmaf(z, z, fmaf(y, y, x * x)));
}

struct Density32 {
    static constexpr bool INTERACT = true;
    using Lds = Lds32Pos;
    static constexpr bool STAGE_Q2 = true;  // pos.w = |x|^2
    static constexpr int WAVES = 6;  // a few spilled registers outside the loops
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        float w;
    };
    struct Acc {
        float s, e;
    };
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <bool FLAT>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, Lds& s)
    {
        s.pos[l] = rel_pos32<FLAT>(r, ox, oy, oz);
    }
    static __device__ void init(Acc& a) { a.s = a.e = 0.0f; }
    static __device__ void zero(Part& q) { q.w = 0.0f; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt&, const Lds&,
                                    int, int, int, float, float, float, float ss, Part& q)
    {
        const float r = sqrt_hw(ss);
        const float x = r * (float)d.g.inv_h;
        const float u = 1.0f - x;
        const float u2 = u * u;
        q.w = fmaf(u2 * u2, fmaf(4.0f, x, 1.0f), q.w);
    }
    static __device__ void fold(Acc& a, const Part& q) { two_sum_add(a.s, a.e, q.w); }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int out_col)
    {
        double w = (double)a.s + (double)a.e;
        for (int o = T; o < 64; o <<= 1) w += __shfl_xor(w, o);
        if (!write) return;
        double v = c.overwrite ? 0.0 : COL(c, out_col, p);
        v = v + d.ph.kW * w;
        if (c.self) v = v + d.ph.kW;
        COL(c, out_col, p) = v;
    }
};

struct Force32 {
    static constexpr bool INTERACT = true;
    using Lds = Lds32Full;
    static constexpr bool STAGE_Q2 = false;  // pos.w = P/rho^2
    static constexpr int WAVES = 5;  // spills a few registers; LDS fits five blocks too
    struct Tgt {
        float x, y, z, vx, vy, vz, pr, rho;
        bool fluid;
    };
    struct Part {
        float ax, ay, az, lx, ly, lz;
    };
    struct Acc {
        float sx, sy, sz, ex, ey, ez, lx, ly, lz, fx, fy, fz;
    };
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
        t.vx = (float)COL(c, C_VX, p); t.vy = (float)COL(c, C_VY, p); t.vz = (float)COL(c, C_VZ, p);
        t.pr = (float)c.prho2[p];
        t.rho = (float)COL(c, C_RHO, p);
        t.fluid = COL(c, C_TYPE, p) == 0.0;
    }
    static __device__ bool active(const Tgt& t) { return t.fluid; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_RHO, q);
        r.a1 = COL(c, C_VX, q); r.a2 = COL(c, C_VY, q); r.a3 = COL(c, C_VZ, q);
        r.a4 = c.prho2[q];
        return r;
    }
    template <bool FLAT>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, Lds& s)
    {
        float4 a = rel_pos32<FLAT>(r, ox, oy, oz);
        a.w = (float)r.a4;
        s.pos[l] = a;
        s.v[l] = make_float4((float)r.a1, (float)r.a2, (float)r.a3, (float)r.a0);
    }
    static __device__ void init(Acc& a)
    {
        a.sx = a.sy = a.sz = a.ex = a.ey = a.ez = 0.0f;
        a.lx = a.ly = a.lz = a.fx = a.fy = a.fz = 0.0f;
    }
    static __device__ void zero(Part& q) { q.ax = q.ay = q.az = q.lx = q.ly = q.lz = 0.0f; }
    // variant bit 0: Monaghan viscosity (avk != 0); bit 1: laminar (mu != 0)
    static constexpr int NVAR = 4;
    static __device__ int variant(const SysDesc& d)
    {
        return (d.ph.avk != 0.0 ? 1 : 0) | (d.ph.mu != 0.0 ? 2 : 0);
    }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds& s, int jj, int, int, float dx, float dy,
                                    float dz, float ss, Part& q)
    {
        const float r = sqrt_hw(ss);
        const float u = 1.0f - r * (float)d.g.inv_h;
        const float u3 = u * u * u;
        const float4 v = s.v[jj];
        float A = t.pr + s.pos[jj].w;
        const float dvx = t.vx - v.x, dvy = t.vy - v.y, dvz = t.vz - v.z;
        if (V & 1) {
            // evaluated for every pair and selected, so the unrolled pair loop
            // stays one basic block
            const float vx = fmaf(dvz, dz, fmaf(dvy, dy, dvx * dx));
            const float rhobar = 0.5f * (t.rho + v.w);
            const float den = rhobar * (ss + (float)d.ph.eps);
            const float Av = fmaf((float)d.ph.avk * vx, __builtin_amdgcn_rcpf(den), A);
            A = vx < 0.0f ? Av : A;
        }
        const float f = u3 * A;
        q.ax = fmaf(f, dx, q.ax); q.ay = fmaf(f, dy, q.ay); q.az = fmaf(f, dz, q.az);
        if (V & 2) {
            q.lx = fmaf(u3, dvx, q.lx); q.ly = fmaf(u3, dvy, q.ly); q.lz = fmaf(u3, dvz, q.lz);
        }
    }
    static __device__ void fold(Acc& a, const Part& q)
    {
        two_sum_add(a.sx, a.ex, q.ax);
        two_sum_add(a.sy, a.ey, q.ay);
        two_sum_add(a.sz, a.ez, q.az);
        two_sum_add(a.lx, a.fx, q.lx);
        two_sum_add(a.ly, a.fy, q.ly);
        two_sum_add(a.lz, a.fz, q.lz);
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        double ax = (double)a.sx + (double)a.ex, ay = (double)a.sy + (double)a.ey,
               az = (double)a.sz + (double)a.ez;
        double lx = (double)a.lx + (double)a.fx, ly = (double)a.ly + (double)a.fy,
               lz = (double)a.lz + (double)a.fz;
        for (int o = T; o < 64; o <<= 1) {
            ax += __shfl_xor(ax, o); ay += __shfl_xor(ay, o); az += __shfl_xor(az, o);
            lx += __shfl_xor(lx, o); ly += __shfl_xor(ly, o); lz += __shfl_xor(lz, o);
        }
        if (!write) return;
        const double kp = -d.ph.kdW;
        const double kl = 2.0 * d.ph.kdW * d.ph.mu / (d.ph.rho0 * d.ph.rho0);
        double ox = c.overwrite ? 0.0 : COL(c, C_AX, p);
        double oy = c.overwrite ? 0.0 : COL(c, C_AY, p);
        double oz = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        ox = ox + kp * ax; oy = oy + kp * ay; oz = oz + kp * az;
        if (d.ph.mu != 0.0) { ox = ox + kl * lx; oy = oy + kl * ly; oz = oz + kl * lz; }
        COL(c, C_AX, p) = ox; COL(c, C_AY, p) = oy; COL(c, C_AZ, p) = oz;
    }
};

// internal_force! of tests/test_collision_2d.jl:75-78 in the mixed form:
// a += -kdW/rho0^2 u^3 (P_p + P_q) x_pq for every target, P staged in FP32.
struct ForceRho032 {
    static constexpr bool INTERACT = true;
    using Lds = Lds32Pos;
    static constexpr bool STAGE_Q2 = false;  // pos.w = P
    static constexpr int WAVES = 6;
    struct Tgt {
        float x, y, z, P;
    };
    struct Part {
        float ax, ay, az;
    };
    struct Acc {
        float sx, sy, sz, ex, ey, ez;
    };
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
        t.P = (float)COL(c, C_P, p);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_P, q);
        return r;
    }
    template <bool FLAT>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, Lds& s)
    {
        float4 a = rel_pos32<FLAT>(r, ox, oy, oz);
        a.w = (float)r.a0;
        s.pos[l] = a;
    }
    static __device__ void init(Acc& a) { a.sx = a.sy = a.sz = a.ex = a.ey = a.ez = 0.0f; }
    static __device__ void zero(Part& q) { q.ax = q.ay = q.az = 0.0f; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds& s, int jj, int, int, float dx, float dy,
                                    float dz, float ss, Part& q)
    {
        const float r = sqrt_hw(ss);
        const float u = 1.0f - r * (float)d.g.inv_h;
        const float f = u * u * u * (t.P + s.pos[jj].w);
        q.ax = fmaf(f, dx, q.ax); q.ay = fmaf(f, dy, q.ay); q.az = fmaf(f, dz, q.az);
    }
    static __device__ void fold(Acc& a, const Part& q)
    {
        two_sum_add(a.sx, a.ex, q.ax);
        two_sum_add(a.sy, a.ey, q.ay);
        two_sum_add(a.sz, a.ez, q.az);
    }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        double ax = (double)a.sx + (double)a.ex, ay = (double)a.sy + (double)a.ey,
               az = (double)a.sz + (double)a.ez;
        for (int o = T; o < 64; o <<= 1) {
            ax += __shfl_xor(ax, o); ay += __shfl_xor(ay, o); az += __shfl_xor(az, o);
        }
        if (!write) return;
        const double k = -d.ph.kdW / (d.ph.rho0 * d.ph.rho0);
        const double bx = c.overwrite ? 0.0 : COL(c, C_AX, p);
        const double by = c.overwrite ? 0.0 : COL(c, C_AY, p);
        const double bz = c.overwrite ? 0.0 : COL(c, C_AZ, p);
        COL(c, C_AX, p) = bx + k * ax;
        COL(c, C_AY, p) = by + k * ay;
        COL(c, C_AZ, p) = bz + k * az;
    }
};

// balance_of_mass! (examples/collapse_dry.jl:112-115) in the mixed form:
// Drho += kdW u^3 (x_pq.v_pq + 2 nu (rho_p - rho_q)).  The densities are
// staged as rho - rho0 in FP32, so their difference keeps its precision.
struct Continuity32 {
    static constexpr bool INTERACT = true;
    using Lds = Lds32Full;
    static constexpr bool STAGE_Q2 = false;  // pos.w = rho - rho0
    static constexpr int WAVES = 5;
    struct Tgt {
        float x, y, z, vx, vy, vz, dr;
    };
    struct Part {
        float d;
    };
    struct Acc {
        float s, e;
    };
    static __device__ void load(const PairCtx& c, const SysDesc& d, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
        t.vx = (float)COL(c, C_VX, p); t.vy = (float)COL(c, C_VY, p); t.vz = (float)COL(c, C_VZ, p);
        t.dr = (float)(COL(c, C_RHO, p) - d.ph.rho0);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        r.a0 = COL(c, C_RHO, q);
        r.a1 = COL(c, C_VX, q); r.a2 = COL(c, C_VY, q); r.a3 = COL(c, C_VZ, q);
        return r;
    }
    template <bool FLAT>
    static __device__ void put(const Raw& r, int l, const SysDesc& d, double ox, double oy,
                               double oz, Lds& s)
    {
        float4 a = rel_pos32<FLAT>(r, ox, oy, oz);
        a.w = (float)(r.a0 - d.ph.rho0);
        s.pos[l] = a;
        s.v[l] = make_float4((float)r.a1, (float)r.a2, (float)r.a3, 0.0f);
    }
    static __device__ void init(Acc& a) { a.s = a.e = 0.0f; }
    static __device__ void zero(Part& q) { q.d = 0.0f; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx&, const SysDesc& d, const Tgt& t,
                                    const Lds& s, int jj, int, int, float dx, float dy,
                                    float dz, float ss, Part& q)
    {
        const float r = sqrt_hw(ss);
        const float u = 1.0f - r * (float)d.g.inv_h;
        const float4 v = s.v[jj];
        const float xv = fmaf(dz, t.vz - v.z, fmaf(dy, t.vy - v.y, dx * (t.vx - v.x)));
        const float dd = fmaf(2.0f * (float)d.ph.nu, t.dr - s.pos[jj].w, xv);
        q.d = fmaf(u * u * u, dd, q.d);
    }
    static __device__ void fold(Acc& a, const Part& q) { two_sum_add(a.s, a.e, q.d); }
    static __device__ void finish(const PairCtx& c, const SysDesc& d, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        double v = (double)a.s + (double)a.e;
        for (int o = T; o < 64; o <<= 1) v += __shfl_xor(v, o);
        if (!write) return;
        const double b = c.overwrite ? 0.0 : COL(c, C_DRHO, p);
        COL(c, C_DRHO, p) = b + d.ph.kdW * v;
    }
};

template <bool EMIT>
struct Pairs32 {
    static constexpr bool INTERACT = EMIT;
    using Lds = Lds32Pos;
    static constexpr bool STAGE_Q2 = true;
    static constexpr int WAVES = 5;
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        int dummy;
    };
    struct Acc {
        int dummy;
    };
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <bool FLAT>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, Lds& s)
    {
        s.pos[l] = rel_pos32<FLAT>(r, ox, oy, oz);
    }
    static __device__ void init(Acc&) {}
    static __device__ void zero(Part&) {}
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const Lds&,
                                    int, int p, int q, float, float, float, float, Part&)
    {
        const unsigned long long k = atomicAdd(c.emit_counter, 1ull);
        if ((long long)k < c.pair_cap) {
            c.pair_out[2 * k] = c.refidx[p];
            c.pair_out[2 * k + 1] = c.refidx[q];
        }
    }
    static __device__ void fold(Acc&, const Part&) {}
    static __device__ void finish(const PairCtx&, const SysDesc&, int, const Tgt&, const Acc&,
                                  int, bool, int) {}
};

struct Digest32 {
    static constexpr bool INTERACT = true;
    using Lds = Lds32Pos;
    static constexpr bool STAGE_Q2 = true;
    static constexpr int WAVES = 5;
    struct Tgt {
        float x, y, z;
    };
    struct Part {
        unsigned long long h, n;
    };
    using Acc = Part;
    static __device__ void load(const PairCtx& c, const SysDesc&, int p, double ox, double oy,
                                double oz, Tgt& t)
    {
        t.x = (float)(COL(c, C_X, p) - ox);
        t.y = (float)(COL(c, C_Y, p) - oy);
        t.z = (float)(COL(c, C_Z, p) - oz);
    }
    static __device__ bool active(const Tgt&) { return true; }
    template <bool P>
    static __device__ Raw fetch(const PairCtx& c, int q, const Seg& g)
    {
        Raw r{};
        fetch_pos<P>(c, q, g, r);
        return r;
    }
    template <bool FLAT>
    static __device__ void put(const Raw& r, int l, const SysDesc&, double ox, double oy,
                                    double oz, Lds& s)
    {
        s.pos[l] = rel_pos32<FLAT>(r, ox, oy, oz);
    }
    static __device__ void init(Acc& a) { a.h = a.n = 0; }
    static __device__ void zero(Part& a) { a.h = a.n = 0; }
    static constexpr int NVAR = 1;
    static __device__ int variant(const SysDesc&) { return 0; }
    template <int V>
    static __device__ void interact(const PairCtx& c, const SysDesc&, const Tgt&, const Lds&,
                                    int, int, int q, float, float, float, float, Part& a)
    {
        a.h += mix64((unsigned long long)c.refidx[q] + 1ull);
        a.n += 1;
    }
    static __device__ void fold(Acc& a, const Part& q) { a.h += q.h; a.n += q.n; }
    static __device__ void finish(const PairCtx& c, const SysDesc&, int p, const Tgt&,
                                  const Acc& a, int T, bool write, int)
    {
        digest_write(c, p, a.h, a.n, T, write);
    }
};

// A lane's run: cnt accepted slots of w (lowest first) against target ht.
template <int V, bool FLAT, int T, class Op>
__device__ inline void walk32(const PairCtx& c, const SysDesc& d, const typename Op::Tgt& ht,
                              const typename Op::Lds& lds, const int* q, unsigned long long w,
                              int cnt, int rb, int hp, typename Op::Part& part)
{
    if constexpr (V < Op::NVAR) {
#pragma unroll 2
        for (int k = 0; k < cnt; ++k) {
            const int b = __builtin_ctzll(w);  // w != 0: cnt <= its popcount
            w &= w - 1;
            const int jj = rb + slot_of_bit<T>(b);
            const float4 a = lds.pos[jj];
            const float dx = ht.x - a.x, dy = ht.y - a.y;
            const float dz = FLAT ? 0.0f : ht.z - a.z;
            const float ss = FLAT ? fmaf(dy, dy, dx * dx) : fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            Op::template interact<V>(c, d, ht, lds, jj, hp, q[jj], dx, dy, dz, ss, part);
        }
    }
}

template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op>
__device__ inline void tile32(const PairCtx& c, int tile, int out_col, bool load_masks,
                              bool store_masks, bool kept_ok, typename Op::Lds& lds, Cands<ROWS, PERIODIC>& L,
                              Bal<typename Op::Part>& bal, unsigned& npairs,
                              unsigned long long& ntests)
{
    const int4 td = ld_const(c.tiles + tile);
    const int cell = td.x, first = td.y;
    const SysDesc d = ld_const(c.sys + ld_const(c.cell_sys + cell));
    if (c.only_flagged && !d.pass_on) return;
    const int lc = cell - d.cell_base;
    const long long mbase = mask_base(td);
    const int cend = ld_const(c.cell_start + cell + 1);
    const int nvalid = min(T, cend - first);
    const int lane = threadIdx.x;
    const int t = lane & (T - 1), s = lane / T;
    const int p = first + t;
    const bool valid = p < cend;
    const int pp = valid ? p : first;
    // tile origin (positions are not written by the pair kernels)
    const double ox = ld_const(&COL(c, C_X, first)), oy = ld_const(&COL(c, C_Y, first)),
                 oz = ld_const(&COL(c, C_Z, first));
    typename Op::Tgt tg;
    Op::load(c, d, pp, ox, oy, oz, tg);
    const bool act = valid && Op::active(tg);
    if (!__any(act)) return;
    const double s_thr = d.g.s_thr;
    // expanded test: |q|^2 - 2 p.q <= s_thr (1 -+ BAND) - |p|^2 -- below the
    // lower threshold the pair is in, above the upper one out, and in
    // between the exact FP64 test decides
    const float tz = FLAT ? 0.0f : tg.z;
    const float p2 = fmaf(tz, tz, fmaf(tg.y, tg.y, tg.x * tg.x));
    const float thr_in = (float)(s_thr * (1.0 - BAND)) - p2;
    const float thr_out = (float)(s_thr * (1.0 + BAND)) - p2;
    const float mx = -2.0f * tg.x, my = -2.0f * tg.y, mz = -2.0f * tz;
    // a pass that loads the masks takes the kept candidates from the list the
    // testing pass wrote; otherwise it streams the candidate list and culls
    // it again, keeping everything within the guard band of the box
    const bool use_kept = load_masks && kept_ok;
    const bool store_kept = store_masks && kept_ok;
    const long long lbase = mbase / T * 64;
    const int nk = use_kept ? ld_const(c.kept_n + tile) : 0;
    const float cull_thr = (float)(s_thr * (1.0 + BAND));
    int total = 0, nr = 0;
    Box<float> box{};
    // (periodic systems need the ranges' image shifts in either case)
    if (use_kept && PERIODIC) nr = build_cands<ROWS, PERIODIC>(d, c.cell_start, lc, L, total);
    if (!use_kept) {
        // RN(x - ox) is monotonic in x: the box of the relative positions
        // is the relative box
        const Box<double> b = ld_const(reinterpret_cast<const Box<double>*>(c.boxes) + tile);
        box.x0 = (float)(b.x0 - ox); box.x1 = (float)(b.x1 - ox);
        box.y0 = (float)(b.y0 - oy); box.y1 = (float)(b.y1 - oy);
        box.z0 = (float)(b.z0 - oz); box.z1 = (float)(b.z1 - oz);
        nr = build_cands<ROWS, PERIODIC>(d, c.cell_start, lc, L, total);
    }
    __syncthreads();
    typename Op::Acc acc;
    Op::init(acc);
    int cur = 0, fill = 0, head = 0, nself = 0, chunk = 0;
    const int nstep = use_kept ? (nk + 63) / 64 : (total + 63) / 64;
    unsigned long long kword = 0;  // the chunk's mask word, loaded with its list entries
    for (int step = 0; step < nstep; ++step) {
        bool last;
        if (use_kept) {
            // one chunk of the kept list per step
            if (s == 0 && act) kword = c.masks[mbase + (long long)step * T + t];
            const int k = head + lane;
            if (k < nk) {
                const int q = c.kept[lbase + k];
                Seg sh{};
                if constexpr (PERIODIC) sh = range_shift(L, c.kept_rg[lbase + k]);
                Op::template put<FLAT>(Op::template fetch<PERIODIC>(c, q, sh), k & 127, d, ox,
                                       oy, oz, lds);
                L.q[k & 127] = q;
            }
            fill = min(nk, head + 64);
            last = true;
        } else {
            const int base = step * 64;
            Raw raw;
            int rq = 0, rrg = 0;
            cur = advance_range(L, nr, cur, base);
            fetch_slot<ROWS, PERIODIC, Op>(c, L, total, cur, base + lane, raw, rq, rrg);
            // keep the candidates within reach of the box, and the tile's targets
            const float4 a = rel_pos32<FLAT>(raw, ox, oy, oz);
            const bool keep = base + lane < total &&
                              (box_dist2<FLAT>(box, a.x, a.y, a.z) <= cull_thr ||
                               (unsigned)(rq - first) < (unsigned)nvalid);
            const int pos = compact(L, fill, nself, keep, rq, rrg, first);
            if (pos >= 0) {
                Op::template put<FLAT>(raw, pos & 127, d, ox, oy, oz, lds);
                if (store_kept) {
                    c.kept[lbase + pos] = rq;
                    if constexpr (PERIODIC) c.kept_rg[lbase + pos] = (unsigned char)rrg;
                }
            }
            last = base + 64 >= total;
        }
        __syncthreads();
        while (fill - head >= 64 || (last && fill > head)) {
            if (fill - head < 64) {
                // the last chunk: pad its free slots
                const int k = head + lane;
                if (k >= fill) {
                    Op::template put<FLAT>(pad_raw(), k & 127, d, ox, oy, oz, lds);
                    L.q[k & 127] = first;
                    if constexpr (PERIODIC) L.rg[k & 127] = 0;
                }
                __syncthreads();
            }
            const int rb = head & 64;  // ring half of the chunk
            // accepted-slot word of target t (owner lanes publish it)
            unsigned long long word = 0;
            if (use_kept) {
                word = kword;
            } else if (load_masks) {
                if (s == 0 && act) word = c.masks[mbase + (long long)chunk * T + t];
            } else {
                // 32-bit words, slots shifted in from the highest bit down
                MaskT<T> m = 0, unsure = 0;
#pragma unroll
                for (int h = 0; h < (T + 31) / 32; ++h) {
                    unsigned min = 0, mout = 0;
#pragma unroll 4
                    for (int j = (T < 32 ? T : 32) - 1; j >= 0; --j) {
                        const float4 a = lds.pos[rb + s + (64 / T) * (32 * h + j)];
                        const float q2 =
                            Op::STAGE_Q2 ? a.w : fmaf(a.z, a.z, fmaf(a.y, a.y, a.x * a.x));
                        float e = fmaf(mx, a.x, q2);
                        e = fmaf(my, a.y, e);
                        if (!FLAT) e = fmaf(mz, a.z, e);
                        min = shift_in_le(min, e, thr_in);
                        mout = shift_in_le(mout, e, thr_out);
                    }
                    m |= (MaskT<T>)min << (32 * h);
                    unsure |= (MaskT<T>)mout << (32 * h);
                }
                unsure &= ~m;
                if (!act) m = unsure = 0;
                // the guard band: exact FP64 predicate of the reference (rare)
                while (__any(unsure != 0)) {
                    if (unsure) {
                        const int j = lowest_bit<T>(unsure);
                        unsure &= unsure - 1;
                        const int jj = rb + s + (64 / T) * j;
                        const int q = L.q[jj];
                        Seg sh{};
                        if constexpr (PERIODIC) sh = range_shift(L, L.rg[jj]);
                        const double e1 =
                            COL(c, C_X, pp) - shifted<PERIODIC>(COL(c, C_X, q), sh.sx);
                        const double e2 =
                            COL(c, C_Y, pp) - shifted<PERIODIC>(COL(c, C_Y, q), sh.sy);
                        double es = e1 * e1 + e2 * e2;
                        if (!FLAT) {
                            const double e3 =
                                COL(c, C_Z, pp) - shifted<PERIODIC>(COL(c, C_Z, q), sh.sz);
                            es = es + e3 * e3;
                        }
                        if (es <= s_thr) m |= MaskT<T>(1) << j;
                    }
                }
                drop_self<ROWS, PERIODIC, T>(L, nself, head, t, s, m);
                word = (unsigned long long)m << (T == 64 ? 0 : s * T);
#pragma unroll
                for (int o = T; o < 64; o <<= 1) word |= __shfl_xor(word, o);
                if (s != 0) word = 0;
            }
            if (s == 0) bal.w[t] = word;
            int tt, r0, cnt, frun, nrun;
            const int tot = split_runs(lane, __popcll(word), bal, tt, r0, cnt, frun, nrun);
            if (tot > 0) {
                typename Op::Part part;
                Op::zero(part);
                // the run's target record from its owner lane (every lane
                // takes part in the shuffle)
                const typename Op::Tgt ht = shfl_value(tg, tt < 0 ? 0 : tt);
                if (Op::INTERACT && cnt > 0) {
                    const unsigned long long w = drop_low_bits(bal.w[tt], r0);
                    const int hp = first + tt;
                    // the operator's uniform switches are taken outside the loop
                    switch (Op::variant(d)) {
                    case 0: walk32<0, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 1: walk32<1, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    case 2: walk32<2, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    default: walk32<3, FLAT, T, Op>(c, d, ht, lds, L.q, w, cnt, rb, hp, part); break;
                    }
                }
                bal.part[lane] = part;
                __syncthreads();
                for (int k = 0; k < nrun; ++k) Op::fold(acc, bal.part[frun + k]);
            }
            npairs += (unsigned)__popcll(word);
            if (store_masks && s == 0) c.masks[mbase + (long long)chunk * T + t] = word;
            ++chunk;
            head += 64;
            __syncthreads();
        }
    }
    // distance tests: the box test of every candidate, the pair test of
    // every kept candidate for every target
    if (!load_masks) ntests += (unsigned long long)total + (unsigned long long)fill * nvalid;
    if (store_kept && lane == 0) c.kept_n[tile] = fill;
    Op::finish(c, d, p, tg, acc, T, act && s == 0, out_col);
}

template <int ROWS, bool FLAT, bool PERIODIC, int T, class Op>
__global__ __launch_bounds__(64, Op::WAVES) void k_pair32(PairCtx c, int out_col)
{
    __shared__ typename Op::Lds lds;
    __shared__ Cands<ROWS, PERIODIC> L;
    __shared__ Bal<typename Op::Part> bal;
    const int ntiles = ld_const(c.meta + META_NTILES);
    const bool ok = c.masks != nullptr && ld_const(c.meta + META_MASK_OK) != 0;
    const bool load_masks = ok && c.mask_mode == MASK_LOAD;
    const bool store_masks = ok && c.mask_mode == MASK_TEST_STORE;
    const bool kept_ok = ok && c.kept != nullptr && ld_const(c.meta + META_KEPT_OK) != 0;
    TileCursor tc;
    tile_cursor_init(tc);
    // accepted pairs of all the wave's tiles, one atomic per wave
    unsigned npairs = 0;
    unsigned long long ntests = 0;
    for (int tile = next_tile(c, ntiles, tc); tile < ntiles; tile = next_tile(c, ntiles, tc))
        tile32<ROWS, FLAT, PERIODIC, T, Op>(c, tile, out_col, load_masks, store_masks, kept_ok, lds,
                                            L, bal, npairs, ntests);
    if (c.pair_counter) count_pairs(c, npairs);
    count_tests(c, ntests);
}

// ========================================================= diagnostics ====
// Candidates tested per pass: for every particle, the particles of the 27
// (9) cells it visits (src/core.jl:96-102).  out[0] over all targets,
// out[1] over fluid targets (the force pass).
template <int ROWS>
__global__ __launch_bounds__(256) void k_candidates(PairCtx c, const int* __restrict__ key,
                                                    unsigned long long* out)
{
    const int n = c.meta[META_N];
    const int p = blockIdx.x * 256 + threadIdx.x;
    unsigned long long all = 0, fluid = 0;
    if (p < n) {
        const int cell = key[p];
        const SysDesc& d = c.sys[c.cell_sys[cell]];
        for (int row = 0; row < ROWS; ++row) {
            Seg segs[2];
            const int ns = row_segments<ROWS>(d, c.cell_start, cell - d.cell_base, row, segs);
            for (int q = 0; q < 2; ++q)
                if (q < ns) all += (unsigned long long)(segs[q].j1 - segs[q].j0);
        }
        if (COL(c, C_TYPE, p) == 0.0) fluid = all;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        all += __shfl_xor(all, o);
        fluid += __shfl_xor(fluid, o);
    }
    if (__lane_id() == 0) {
        atomicAdd(&out[0], all);
        atomicAdd(&out[1], fluid);
    }
}

void launch_candidates(int rows, const PairCtx& c, const int* key, unsigned long long* out,
                       long long n, hipStream_t st)
{
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (nb == 0) return;
    if (rows == 3) k_candidates<3><<<nb, 256, 0, st>>>(c, key, out);
    else k_candidates<9><<<nb, 256, 0, st>>>(c, key, out);
}

// ============================================================ dispatch ====

template <int ROWS, bool FLAT, bool PER, int T>
static void launch_tile(int op, int precision, const PairCtx& c, int grid, hipStream_t st)
{
    if (precision == SPH_FP64) {
        switch (op) {
        case P_DENSITY: k_pair64<ROWS, FLAT, PER, T, Density64<C_RHO>><<<grid, 64, 0, st>>>(c); break;
        case P_DENSITY_RHO0: k_pair64<ROWS, FLAT, PER, T, Density64<C_RHO0>><<<grid, 64, 0, st>>>(c); break;
        case P_FORCE: k_pair64<ROWS, FLAT, PER, T, Force64><<<grid, 64, 0, st>>>(c); break;
        case P_FORCE_RHO0: k_pair64<ROWS, FLAT, PER, T, ForceRho064><<<grid, 64, 0, st>>>(c); break;
        case P_CONTINUITY: k_pair64<ROWS, FLAT, PER, T, Continuity64><<<grid, 64, 0, st>>>(c); break;
        case P_COUNT: k_pair64<ROWS, FLAT, PER, T, Pairs64<false>><<<grid, 64, 0, st>>>(c); break;
        case P_EMIT: k_pair64<ROWS, FLAT, PER, T, Pairs64<true>><<<grid, 64, 0, st>>>(c); break;
        case P_DIGEST: k_pair64<ROWS, FLAT, PER, T, Digest64><<<grid, 64, 0, st>>>(c); break;
        }
    } else {
        switch (op) {
        case P_DENSITY: k_pair32<ROWS, FLAT, PER, T, Density32><<<grid, 64, 0, st>>>(c, C_RHO); break;
        case P_DENSITY_RHO0: k_pair32<ROWS, FLAT, PER, T, Density32><<<grid, 64, 0, st>>>(c, C_RHO0); break;
        case P_FORCE: k_pair32<ROWS, FLAT, PER, T, Force32><<<grid, 64, 0, st>>>(c, 0); break;
        case P_COUNT: k_pair32<ROWS, FLAT, PER, T, Pairs32<false>><<<grid, 64, 0, st>>>(c, 0); break;
        case P_EMIT: k_pair32<ROWS, FLAT, PER, T, Pairs32<true>><<<grid, 64, 0, st>>>(c, 0); break;
        case P_DIGEST: k_pair32<ROWS, FLAT, PER, T, Digest32><<<grid, 64, 0, st>>>(c, 0); break;
        case P_FORCE_RHO0: k_pair32<ROWS, FLAT, PER, T, ForceRho032><<<grid, 64, 0, st>>>(c, 0); break;
        case P_CONTINUITY: k_pair32<ROWS, FLAT, PER, T, Continuity32><<<grid, 64, 0, st>>>(c, 0); break;
        }
    }
}

template <int ROWS, bool FLAT, bool PER>
static void launch_rows(int op, int precision, const PairCtx& c, int grid, hipStream_t st)
{
    switch (c.T) {
    case 8: launch_tile<ROWS, FLAT, PER, 8>(op, precision, c, grid, st); break;
    case 16: launch_tile<ROWS, FLAT, PER, 16>(op, precision, c, grid, st); break;
    case 32: launch_tile<ROWS, FLAT, PER, 32>(op, precision, c, grid, st); break;
    default: launch_tile<ROWS, FLAT, PER, 64>(op, precision, c, grid, st); break;
    }
}

void launch_pair(int op, int precision, int rows, int flat, int periodic, const PairCtx& c,
                 int grid, hipStream_t st)
{
    if (grid <= 0) return;
    if (rows == 3) {
        if (flat) {
            if (periodic) launch_rows<3, true, true>(op, precision, c, grid, st);
            else launch_rows<3, true, false>(op, precision, c, grid, st);
        } else {
            if (periodic) launch_rows<3, false, true>(op, precision, c, grid, st);
            else launch_rows<3, false, false>(op, precision, c, grid, st);
        }
    } else {
        if (periodic) launch_rows<9, false, true>(op, precision, c, grid, st);
        else launch_rows<9, false, false>(op, precision, c, grid, st);
    }
}

}  // namespace sph
