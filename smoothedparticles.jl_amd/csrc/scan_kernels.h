// Exclusive device-wide scan shared by the cell-list build (cells.hip) and
// the device lattice generator (generate.hip).
#pragma once

#include "sph_internal.h"

namespace sph {

static constexpr int SCAN_ITEMS = 16;
static constexpr int SCAN_TILE = 256 * SCAN_ITEMS;

// Exclusive device-wide scan, three launches: per-block sums, one block
// scanning the block sums, per-block scans plus offset.  V is unsigned or
// unsigned long long.
template <class V>
__device__ inline V wave_incl_scan(V v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const V t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Inclusive scan over a 256-thread block; `total` gets the block sum.
template <class V>
__device__ inline V block_incl_scan(V v, V& total)
{
    __shared__ V wsum[4];
    const V w = wave_incl_scan(v);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[wid] = w;
    __syncthreads();
    V off = 0;
    for (int i = 0; i < wid; ++i) off += wsum[i];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return w + off;
}

template <class Src>
__global__ __launch_bounds__(256) void k_scan_reduce(Src src, int n, typename Src::V* part)
{
    using V = typename Src::V;
    const int base = blockIdx.x * SCAN_TILE;
    V s = 0;
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int i = base + it * 256 + threadIdx.x;
        if (i < n) s += src(i);
    }
    V total;
    block_incl_scan(s, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

template <class V>
__global__ __launch_bounds__(256) void k_scan_partials(V* part, int nb)
{
    V carry = 0;
    for (int b = 0; b < nb; b += 256) {
        const int i = b + threadIdx.x;
        const V v = i < nb ? part[i] : (V)0;
        V total;
        const V inc = block_incl_scan(v, total);
        if (i < nb) part[i] = carry + inc - v;
        carry += total;
    }
}

// out[i] = exclusive prefix for i < n, out[n] = total.
template <class Src, class O>
__global__ __launch_bounds__(256) void k_scan_down(Src src, int n, const typename Src::V* part,
                                                   O* out)
{
    using V = typename Src::V;
    const int base = blockIdx.x * SCAN_TILE;
    V carry = part[blockIdx.x];
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int i = base + it * 256 + threadIdx.x;
        const V v = i < n ? src(i) : (V)0;
        V total;
        const V inc = block_incl_scan(v, total);
        if (i < n) out[i] = (O)(carry + inc - v);
        if (i == n - 1) out[n] = (O)(carry + inc);
        carry += total;
    }
}

template <class Src, class O>
static void scan(sph_system* s, Src src, int n, O* out, typename Src::V* part)
{
    const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    k_scan_reduce<<<nb, 256, 0, s->stream>>>(src, n, part);
    k_scan_partials<<<1, 256, 0, s->stream>>>(part, nb);
    k_scan_down<<<nb, 256, 0, s->stream>>>(src, n, part, out);
}


}  // namespace sph
