// Pair-pass dispatch by precision, and the candidate-count diagnostic.
// The kernels live in pair64.hip (FP64) and pair32.hip (mixed); their shared
// device code, and the description of the decomposition, in pair_kernels.h.
#include "pair_kernels.h"

namespace sph {

void launch_pair64(int op, int rows, int flat, int periodic, const PairCtx& c, int grid,
                   hipStream_t st);
void launch_pair32(int op, int rows, int flat, int periodic, const PairCtx& c, int grid,
                   hipStream_t st);

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

void launch_pair(int op, int precision, int rows, int flat, int periodic, const PairCtx& c,
                 int grid, hipStream_t st)
{
    if (grid <= 0) return;
    if (precision == SPH_FP64) launch_pair64(op, rows, flat, periodic, c, grid, st);
    else launch_pair32(op, rows, flat, periodic, c, grid, st);
}

}  // namespace sph
