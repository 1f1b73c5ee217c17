// Closed-form max-log LLRs of one axis of a uniform separable lattice, from
// per-cell tables.  Shared by the demapper (ck_llr.hip) and the decoder's
// fused frame load (ck_ldpc_dev.h), so both give the same LLRs bit for bit.
//
// An axis has L = 2^kb levels, ascending and equally spaced.  Its half-grid
// points h_g, g = 0 .. 2L-2, are the levels (even g: h_g = l_{g/2}) and the
// midpoints of neighbouring levels (odd g); they cut the real line into 2L
// cells, w = 0 .. 2L-1, cell w = 2s + side holding the x whose nearest level
// is s, below it (side 0) or at or above it (side 1).  Within a cell the
// max-log LLR of bit b is linear in x: the nearest level s against the
// nearest level j of the opposite bit, which is the nearer to x of the
// nearest levels of opposite bit below and above s (the opposite-level
// tables of ck_axis) and, the levels being equally spaced, the same one for
// every x of the cell (ties aside, where both give the same metric):
//
//     m_s - m_j = ((x - l_j)^2 - (x - l_s)^2) / sigma^2
//               = (x - (l_s + l_j) / 2) * 2 (l_s - l_j) / sigma^2,
//     LLR_b = (x - mid) * coef * inv_s2,  coef = +-2 (l_s - l_j)
//
// with + when s's bit b is 0.  A cell's (mid, coef) per bit and the
// half-grid points are staged in LDS; per symbol and axis the kernel finds
// the cell from a float32 estimate of the half-grid coordinate, corrected by
// one exact float64 comparison with the nearest half-grid point, then takes
// each bit's LLR in float64 (x - mid carries the cancellation near a
// decision boundary; coef and inv_s2 are far from zero) and rounds once to
// float32.  A bit constant over the axis (no opposite level) has coef 0.
#pragma once

#include <hip/hip_runtime.h>

#include "ck_tail.h"

namespace ck {

// Axes of up to this many bits use the table form (64 levels: 4096-QAM);
// the cell tables of wider axes would not fit a workgroup's LDS.
constexpr int kClosedTableBits = 6;

struct ClosedCell {
    double mid;   // (l_s + l_j) / 2
    double coef;  // +-2 (l_s - l_j)
};

template <int KB>
struct __attribute__((aligned(16))) ClosedAxis {
    static constexpr int L = 1 << KB;
    ClosedCell cell[2 * L][KB];  // cell w, bit b
    double half[2 * L - 1];      // half-grid points
    int pos[KB];                 // position of bit b in the symbol's label
    float a, c;                  // half-grid coordinate of x: x * a + c
};

// Stages one axis's tables from its ck_axis (every lane of the workgroup
// takes part; the caller's barrier follows).
template <int KB>
__device__ inline void stage_closed(ClosedAxis<KB>& t, const ck_axis& ax) {
    constexpr int L = 1 << KB;
    for (int i = static_cast<int>(threadIdx.x); i < 2 * L * KB; i += static_cast<int>(blockDim.x)) {
        const int w = i / KB;
        const int b = i - w * KB;
        const int s = w >> 1;
        const int jl = ax.lo[b * L + s];
        const int jh = ax.hi[b * L + s];
        int j;
        if (jl < 0 && jh < 0) {
            j = s;  // the bit is constant over the axis
        } else if (jl < 0) {
            j = jh;
        } else if (jh < 0) {
            j = jl;
        } else {
            // jl nearer for every x of the cell when (l_jl + l_jh) / 2 lies
            // at or above the cell (index units: jl + jh >= 2s below l_s,
            // > 2s from l_s up).
            const int twice_mid = jl + jh;
            j = ((w & 1) ? twice_mid > 2 * s : twice_mid >= 2 * s) ? jl : jh;
        }
        const double ls = ax.levels[s];
        const double lj = ax.levels[j];
        const bool one = (ax.labels[s] >> (KB - 1 - b)) & 1;
        t.cell[w][b].mid = 0.5 * (ls + lj);
        t.cell[w][b].coef = one ? -2.0 * (ls - lj) : 2.0 * (ls - lj);
    }
    for (int g = static_cast<int>(threadIdx.x); g < 2 * L - 1; g += static_cast<int>(blockDim.x))
        t.half[g] = (g & 1) ? 0.5 * (ax.levels[g >> 1] + ax.levels[(g >> 1) + 1])
                            : ax.levels[g >> 1];
    if (threadIdx.x < static_cast<unsigned>(KB)) t.pos[threadIdx.x] = ax.pos[threadIdx.x];
    if (threadIdx.x == 0) {
        const double step = ax.levels[1] - ax.levels[0];
        t.a = static_cast<float>(2.0 / step);
        t.c = static_cast<float>(-2.0 * ax.levels[0] / step);
    }
}

// The kb LLRs of symbol component x into row[pos[b]] (+0.0f added when
// PLUS_ZERO, as the decoder's frame load takes LLRs: -0.0 becomes +0.0).
// The float32 coordinate lands within one half-grid step of x's, so
// g = round(coordinate) is the nearest half-grid point or a neighbour of
// it, and x < h_g in float64 settles the cell: w = g + (x >= h_g).  NaN
// takes cell 1 and gives NaN LLRs.
template <int KB, bool PLUS_ZERO = false>
__device__ __forceinline__ void closed_axis_llrs(const ClosedAxis<KB>& t, float x, double inv_s2,
                                                 float* row) {
    constexpr int L = 1 << KB;
    int g = __float2int_rn(__builtin_fmaf(x, t.a, t.c));
    g = min(max(g, 0), 2 * L - 2);
    const double xd = x;
    const int w = g + (xd < t.half[g] ? 0 : 1);
    const ClosedCell* cw = t.cell[w];
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        const ClosedCell e = cw[b];
        const float v = static_cast<float>(((xd - e.mid) * e.coef) * inv_s2);
        row[t.pos[b]] = PLUS_ZERO ? v + 0.0f : v;
    }
}

}  // namespace ck
