// Bit-LLR demapping on gfx950: three streaming kernels, one symbol per lane.
//
//   separable_closed  max-log LLRs of a uniform separable lattice: per axis,
//                     the cell of x (nearest level and side) from one float64
//                     comparison, then per bit (x - mid) * coef from the
//                     cell's tables (ck_closed.h; axes of up to 6 bits), or
//                     the snapped level against the nearer of the two levels
//                     just outside its run of equal bits (wider axes);
//   separable_levels  the metrics of each axis's levels with the log prior,
//                     max-log or exact log-sum-exp: unimodal Gray forms
//                     (compile-time halves) for Gray labels over equally
//                     spaced levels with a concave log prior, a general form
//                     (argmax per half) otherwise;
//   points            the metrics of all M points (any constellation).
//
// Semantics follow commkit/mapping/_numba_llr.py:35-235 and
// commkit/_cuda/src/llr.cu:100-337 (same routing; a half whose sum would
// lose precision is summed relative to its own maximum).  The arithmetic is
// this library's own, chosen for accuracy against an FP64 evaluation of the
// definition:
//
//   - a difference of two metrics is never formed in float32 from two large
//     metrics: max-log takes it in float64, in factored form,
//       m_a - m_c = (l_a - l_c) * ((x - l_a) + (x - l_c)) * inv_s2
//                   + (logp_a - logp_c),
//     so an LLR near zero (x near a decision boundary) keeps its relative
//     accuracy;
//   - the exact Gray and point forms take every level's (point's) metric in
//     float64 relative to one reference, its exponential as 2^n (exact)
//     times a hardware exp2 of a fraction in [-1/2, 1/2], the half sums in
//     float64, and each LLR as one oriented log(1 + q) of the two half sums
//     (log_ratio); the general form takes exp(m - m_peak) per level with the
//     factored difference and the two log(1 + T) terms of a bit as one logf
//     (log1p_diff).
//
// Layout: per axis, the tables (levels, log prior, labels, opposite levels,
// bit positions) are staged in LDS once per workgroup.  Each lane writes its
// k LLRs into an LDS tile of kThreads rows of stride (k | 1) floats (odd, so
// lane t's row starts in its own bank) and the workgroup stores the tile to
// HBM contiguously, as 16-byte vectors where the tile is whole.
#include "ck_closed.h"
#include "ck_common.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int tile_stride(int k) { return k | 1; }

// Stores `count` staged symbols of k LLRs from the tile to out[base * k ...].
__device__ __forceinline__ void store_tile(const float* tile, float* __restrict__ out,
                                           long long base, int count, int k) {
    const int stride = tile_stride(k);
    float* dst = out + base * k;
    const int total = count * k;
    if (count == kThreads) {  // whole tile: 256 * k floats, 16-byte aligned
        float4* dst4 = reinterpret_cast<float4*>(dst);
        for (int q = threadIdx.x; q < total / 4; q += kThreads) {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = 4 * q + u;
                const int sym = j / k;
                v[u] = tile[sym * stride + (j - sym * k)];
            }
            dst4[q] = make_float4(v[0], v[1], v[2], v[3]);
        }
    } else {
        for (int j = threadIdx.x; j < total; j += kThreads) {
            const int sym = j / k;
            dst[j] = tile[sym * stride + (j - sym * k)];
        }
    }
}

// Stores `count` staged symbols like store_tile, but the LLR of flat index t
// goes to out[(t / period) * period + perm[t % period]] and indices from
// `limit` on are dropped: the bit layout of a code's blocks (probabilistic
// amplitude shaping) applied on the way out.  A tile may span several blocks.
// Lane i handles tile elements i, i + kThreads, ... (K of them at most): it
// gathers all their permutation entries first, so the K loads are in flight
// together, then stores.
template <int K>
__device__ __forceinline__ void store_tile_perm(const float* tile, float* __restrict__ out,
                                                long long base, int count,
                                                const int* __restrict__ perm, long long period,
                                                long long limit) {
    const int stride = tile_stride(K);
    const long long t0 = base * K;
    if (t0 >= limit) return;
    const long long blk0 = t0 / period;
    int total = count * K;
    if (limit - t0 < total) total = static_cast<int>(limit - t0);
    long long off = t0 - blk0 * period + threadIdx.x;
    long long blk = blk0;
    long long dst[K];
#pragma unroll
    for (int u = 0; u < K; ++u) {
        while (off >= period) {  // period may be shorter than a tile
            off -= period;
            ++blk;
        }
        if (threadIdx.x + u * kThreads < total) dst[u] = blk * period + perm[off];
        off += kThreads;
    }
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const int j = threadIdx.x + u * kThreads;
        if (j < total) {
            const int sym = j / K;
            out[dst[u]] = tile[sym * stride + (j - sym * K)];
        }
    }
}

constexpr double kLog2eD = 1.4426950408889634;
constexpr float kLn2 = 0.6931471805599453f;

// LDS copy of one axis with 2^KB levels: float64 levels and log prior for
// the metric differences, float32 copies for locating maxima.
template <int KB>
struct AxisTab {
    static constexpr int L = 1 << KB;
    double lev[L];
    double logp[L];
    double logp2[L];  // log prior in log2 units (exact unimodal Gray form)
    float levf[L];
    float logpf[L];
    int lab[L];
    short lo[KB][L];
    short hi[KB][L];
    int pos[KB];
    // Level-metric form only: nonzero when the labels are the reflected Gray
    // code up to a constant XOR (gray_xor), the levels ascend equally spaced
    // and the log prior is concave over them.  Then the metric
    // -(x - l)^2 inv_s2 + logp(l) is concave in the level index for every x
    // and inv_s2 > 0, and each half's best level is the peak or one of the
    // peak's two nearest levels of opposite bit (gray_opposite).
    int unimodal_gray;
    int gray_xor;
};

// Reflected Gray label of level l.
__host__ __device__ constexpr int gray_of(int l) { return l ^ (l >> 1); }

template <int KB>
__device__ void stage_axis(AxisTab<KB>& t, const ck_axis& a, bool with_opposite) {
    constexpr int L = 1 << KB;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        t.lev[i] = a.levels[i];
        t.logp[i] = a.log_prior[i];
        t.logp2[i] = a.log_prior[i] * kLog2eD;
        t.levf[i] = static_cast<float>(a.levels[i]);
        t.logpf[i] = static_cast<float>(a.log_prior[i]);
        t.lab[i] = a.labels[i];
    }
    if (with_opposite) {
        for (int i = threadIdx.x; i < KB * L; i += blockDim.x) {
            t.lo[i / L][i % L] = static_cast<short>(a.lo[i]);
            t.hi[i / L][i % L] = static_cast<short>(a.hi[i]);
        }
    }
    if (threadIdx.x < KB) t.pos[threadIdx.x] = a.pos[threadIdx.x];
    if (with_opposite) return;
    __syncthreads();
    // One thread per level (L <= blockDim): Gray label up to lab[0], equal
    // spacing, concave log prior around it.
    bool ok = true;
    const int l = threadIdx.x;
    if (l < L) {
        const double step = t.lev[1] - t.lev[0];
        ok = step > 0.0 && (t.lab[l] ^ gray_of(l)) == t.lab[0];
        if (l + 1 < L) ok = ok && fabs((t.lev[l + 1] - t.lev[l]) - step) <= 1e-9 * step;
        if (l >= 1 && l + 1 < L)
            ok = ok && (t.logp[l - 1] + t.logp[l + 1] - 2.0 * t.logp[l]) <= 1e-9;
    }
    const int all = __syncthreads_and(ok ? 1 : 0);
    if (threadIdx.x == 0) {
        t.unimodal_gray = all;
        t.gray_xor = t.lab[0];
    }
}

// Nearest levels below (*lo) and above (*hi) level s whose reflected Gray
// bit sh differs from s's, -1 for none.  That bit changes value entering
// the levels l = 2^sh (mod 2^(sh+1)), so s's run of equal bits starts at
// q 2^(sh+1) - 2^sh (or 0) and the next run at (q + 1) 2^(sh+1) - 2^sh,
// with q = (s + 2^sh) >> (sh + 1).  A constant XOR on the labels keeps the
// runs.
template <int L>
__device__ __forceinline__ void gray_opposite(int s, int sh, int* lo, int* hi) {
    const int h = 1 << sh;
    const int q = (s + h) >> (sh + 1);
    *lo = q == 0 ? -1 : (q << (sh + 1)) - h - 1;
    const int nx = ((q + 1) << (sh + 1)) - h;
    *hi = nx < L ? nx : -1;
}

// m_a - m_c of two levels of one axis, in float64, factored (see the file
// comment): exact up to float64 rounding for the float32 symbol x.
template <int KB>
__device__ __forceinline__ double level_diff(const AxisTab<KB>& t, double x, double inv_s2,
                                             int a, int c) {
    const double la = t.lev[a];
    const double lc = t.lev[c];
    return (la - lc) * ((x - la) + (x - lc)) * inv_s2 + (t.logp[a] - t.logp[c]);
}

// Closed-form max-log LLRs of one axis into row[pos[b]]: the nearest level
// against the nearer of the two levels just outside its run of equal bits.
template <int KB>
__device__ __forceinline__ void axis_closed(const AxisTab<KB>& t, float x, double inv_s2,
                                            float inv_step, float* row) {
    constexpr int L = 1 << KB;
    int s = __float2int_rn((x - t.levf[0]) * inv_step);
    s = min(max(s, 0), L - 1);
    const double xd = x;
    double d = xd - t.lev[s];
    // The float32 estimate can land one level off near a midpoint: keep the
    // nearer of s and its neighbour towards x, compared in float64.
    const int s2 = d > 0.0 ? s + 1 : s - 1;
    if (s2 >= 0 && s2 < L) {
        const double d2 = xd - t.lev[s2];
        if (fabs(d2) < fabs(d)) {
            s = s2;
            d = d2;
        }
    }
    const double ls = t.lev[s];
    const int lab = t.lab[s];
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        const int jl = t.lo[b][s];
        const int jh = t.hi[b][s];
        int j;
        if (jl < 0) {
            j = jh;
        } else if (jh < 0) {
            j = jl;
        } else {
            j = (fabs(xd - t.lev[jh]) < fabs(xd - t.lev[jl])) ? jh : jl;
        }
        j = j < 0 ? s : j;  // a bit constant over the axis: no opposite level
        const double lo = t.lev[j];
        // m_own - m_other, the LLR when the nearest level's bit is 0.
        const double diff = (ls - lo) * (d + (xd - lo)) * inv_s2;
        const float v = static_cast<float>(diff);
        row[t.pos[b]] = ((lab >> (KB - 1 - b)) & 1) ? -v : v;
    }
}

// Below this (in nats) a half's maximum is too far under the peak for its
// exp to stay a normal float32; that half is summed relative to itself.
constexpr double kWeak = -80.0;
constexpr float kLog2e = 1.4426950408889634f;

// m_a - m_c in the factored form of level_diff, with the float64 level and
// log prior of c and x - l_c passed in (taken once per symbol).
__device__ __forceinline__ double diff_to(double la, double pa, double xd, double lc, double pc,
                                          double xc, double inv_s2) {
    return (la - lc) * ((xd - la) + xc) * inv_s2 + (pa - pc);
}

// log(1 + a) - log(1 + b) for a, b >= 0 as one logarithm: oriented so that
// q = (max - min) / (1 + min) >= 0, and log(1 + q) = log(u) + r / u with
// u + r = 1 + q exactly (TwoSum), so the result carries about one rounding
// of logf instead of two of log1pf (ocml's log1pf is a compensated
// double-float routine, several times the cost of logf).
__device__ __forceinline__ float log1p_diff(float a, float b) {
    const bool neg = a < b;
    const float hi = neg ? b : a;
    const float lo = neg ? a : b;
    const float q = (hi - lo) / (1.0f + lo);
    const float u = 1.0f + q;
    const float bv = u - 1.0f;
    const float r = (1.0f - (u - bv)) + (q - bv);
    const float d = logf(u) + r * __builtin_amdgcn_rcpf(u);
    return neg ? -d : d;
}

// Max-log LLRs of one axis whose tables are flagged unimodal_gray (levels
// ascending and equally spaced, log prior concave, labels Gray up to a
// constant XOR).  The float32 metrics locate the peak level pk; for each bit
// the half that holds pk has pk as its best level, and the other half's best
// level a_o is the better, in float64 metric, of pk's two nearest levels of
// opposite bit: LLR = -/+ (m_o - m_pk), sign by the half of pk.
template <int KB>
__device__ __forceinline__ void axis_maxlog_gray(const AxisTab<KB>& t, float x, float inv_s2f,
                                                 double inv_s2, float* row) {
    constexpr int L = 1 << KB;
    float best = -ck::kInf;
    int pk = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const float dx = x - t.levf[l];
        const float m = -(dx * dx) * inv_s2f + t.logpf[l];
        if (m > best) {
            best = m;
            pk = l;
        }
    }
    const double xd = x;
    const double lpk = t.lev[pk];
    const double ppk = t.logp[pk];
    const double xpk = xd - lpk;
    const int gpk = gray_of(pk);
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        const int sh = KB - 1 - b;
        int jl, jh;
        gray_opposite<L>(pk, sh, &jl, &jh);
        double wl = -1.0e300;
        double wh = -1.0e300;
        if (jl >= 0) wl = diff_to(t.lev[jl], t.logp[jl], xd, lpk, ppk, xpk, inv_s2);
        if (jh >= 0) wh = diff_to(t.lev[jh], t.logp[jh], xd, lpk, ppk, xpk, inv_s2);
        const float r = static_cast<float>(-((wh > wl) ? wh : wl));
        row[t.pos[b]] = ((((gpk ^ t.gray_xor) >> sh) & 1) != 0) ? -r : r;
    }
}

// Below this half sum (relative to the reference level's term 1) the half
// is summed again relative to its own best level: its float64 terms would
// lose precision to subnormals.
constexpr double kFarHalf = 0x1p-1000;

// The metric of level l in log2 units: (log p_l - (x - l)^2 / sigma^2) log2(e).
template <int KB>
__device__ __forceinline__ double metric2(const AxisTab<KB>& t, double xd, double s2l, int l) {
    const double d = xd - t.lev[l];
    return fma(-s2l * d, d, t.logp2[l]);
}

// 2^w in float64 for a float64 w: 2^n (exact, by ldexp) times 2^f by the
// hardware exp2 with n = rint(w) and f = w - n in [-1/2, 1/2], so the term
// carries the rounding of exp2 and of f (under 2^-25 absolute) whatever |w|,
// and stays a normal float64 down to 2^-1022.
__device__ __forceinline__ double exp2_split(double w) {
    const double n = __builtin_rint(w);
    const float f = static_cast<float>(w - n);
    return __builtin_ldexp(static_cast<double>(__builtin_amdgcn_exp2f(f)), static_cast<int>(n));
}

// log(x) of a positive float64 from its binary exponent and the float32
// log2 of its mantissa: relative error a few float32 roundings, for any
// normal x.
__device__ __forceinline__ float log_of(double x) {
    int e;
    const double m = frexp(x, &e);
    return (static_cast<float>(e) + __builtin_amdgcn_logf(static_cast<float>(m))) * kLn2;
}

// log(S_a / S_b) for S_a, S_b > 0 given in float64.  For a ratio under 2^64
// it is oriented so that q = |S_a - S_b| / min(S_a, S_b) >= 0 (formed in
// float64, then rounded once) and log(1 + q) = log(u) + r / u with u + r =
// 1 + q exactly (TwoSum): an LLR near zero keeps an absolute error of about
// one float32 rounding of q, and u >= 1 needs no range reduction for the
// hardware log2.  A larger ratio (an LLR above 44) is the difference of the
// two logarithms.
__device__ __forceinline__ float log_ratio(double sa, double sb) {
    const double d = sa - sb;
    const double lo = d < 0.0 ? sa : sb;
    const double hi = d < 0.0 ? sb : sa;
    float v;
    if (lo >= 0x1p-64 * hi) {
        double rl = __builtin_amdgcn_rcp(lo);
        rl = fma(rl, fma(-lo, rl, 1.0), rl);  // one Newton step
        const float q = static_cast<float>(fabs(d) * rl);
        const float u = 1.0f + q;
        const float bv = u - 1.0f;
        const float r = (1.0f - (u - bv)) + (q - bv);
        v = __builtin_amdgcn_logf(u) * kLn2 + r * __builtin_amdgcn_rcpf(u);
    } else {
        v = log_of(hi) - log_of(lo);
    }
    return d < 0.0 ? -v : v;
}

// Exact (log-sum-exp) LLRs of one unimodal Gray axis (conditions as for
// axis_maxlog_gray).  Every level's metric is taken in float64 relative to
// a reference level s (the snapped level of x), in log2 units, w_l =
// M_l - M_s; its term 2^w_l (exp2_split) is added in float64 to one of two
// half sums per bit, the half being the level's Gray bit, a compile-time
// constant of the unrolled loop.  Bit b's LLR is log(S_b0 / S_b1) by
// log_ratio, negated where the label XOR flips the bit.
//
// The half holding s has a sum >= 1 (s's own term).  When the other half's
// sum falls under kFarHalf, every level between s and the peak pk has a
// metric >= M_s (the metric is concave in the level index), so s and pk lie
// in one run of equal bits, the other half's best level a_o is the better of
// s's two nearest levels of opposite bit (gray_opposite), and that half is
// summed again relative to a_o:  log S_oth = (w_o + log2(1 + U)) ln 2 with
// U = sum over the other levels of the half of 2^(M_l - M_o).
//
// A prior far more peaked than the likelihood can put a level 2^996 above s
// (sums over 1e300): then the sums are taken again with s = pk, located in
// float64.
template <int KB>
__device__ __forceinline__ void axis_exact_gray(const AxisTab<KB>& t, float x, double inv_s2,
                                                float inv_step, float* row) {
    constexpr int L = 1 << KB;
    const double xd = x;
    const double s2l = inv_s2 * kLog2eD;
    int s = __float2int_rn((x - t.levf[0]) * inv_step);
    s = min(max(s, 0), L - 1);
    double S0[KB], S1[KB];
    double ms = 0.0;
    for (int pass = 0;; ++pass) {
        ms = metric2(t, xd, s2l, s);
#pragma unroll
        for (int b = 0; b < KB; ++b) {
            S0[b] = 0.0;
            S1[b] = 0.0;
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const double d = xd - t.lev[l];
            const double e = exp2_split(fma(-s2l * d, d, t.logp2[l] - ms));
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                if ((gray_of(l) >> (KB - 1 - b)) & 1)
                    S1[b] += e;
                else
                    S0[b] += e;
            }
            // Four levels at a time: the scheduler would otherwise issue
            // every level's table reads up front and spill.
            if ((l & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
        // Finite sums (also the NaN symbol's second pass) end the loop.
        if (S0[0] + S1[0] < 1.0e300 || pass == 1) break;
        double best = -1.0e300;
#pragma unroll 1
        for (int l = 0; l < L; ++l) {
            const double m = metric2(t, xd, s2l, l);
            if (m > best) {
                best = m;
                s = l;
            }
        }
    }
    const int gs = gray_of(s);
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        const int sh = KB - 1 - b;
        const int own = (gs >> sh) & 1;
        const double s_own = own ? S1[b] : S0[b];
        const double s_oth = own ? S0[b] : S1[b];
        float v;  // log(S_own / S_oth)
        if (s_oth >= kFarHalf) {
            v = log_ratio(s_own, s_oth);
        } else {
            int jl, jh;
            gray_opposite<L>(s, sh, &jl, &jh);
            const double ml = jl >= 0 ? metric2(t, xd, s2l, jl) : -1.0e300;
            const double mh = jh >= 0 ? metric2(t, xd, s2l, jh) : -1.0e300;
            const int a = mh > ml ? jh : jl;
            const double mo = mh > ml ? mh : ml;
            float u = 0.0f;
#pragma unroll 1
            for (int l = 0; l < L; ++l) {
                if (((gray_of(l) >> sh) & 1) != own && l != a)
                    u += __builtin_amdgcn_exp2f(static_cast<float>(metric2(t, xd, s2l, l) - mo));
            }
            // log S_own - log S_oth, with S_oth = 2^(mo - ms) (1 + u)
            const double lo = (ms - mo) * 0.6931471805599453 +
                              static_cast<double>(log_of(s_own) - log1pf(u));
            v = static_cast<float>(lo);
        }
        row[t.pos[b]] = ((own ^ (t.gray_xor >> sh)) & 1) ? -v : v;
    }
}

// LLRs of one axis from the metrics of its levels into row[pos[b]].
// Max-log: the difference of the two halves' maxima, located in float64.
// Exact: LLR = (m_a0 - m_a1) + log(1 + T0) - log(1 + T1), with a_h the
// maximum of half h and T_h = sum over the other levels of half h of
// exp(m_l - m_a_h); the exponentials are taken once per level relative to
// the symbol's peak and divided by the half's own term (or, for a half
// whose maximum is more than 80 nats under the peak, taken relative to that
// maximum directly).
template <int KB>
__device__ __forceinline__ void axis_levels(const AxisTab<KB>& t, float x, float inv_s2f,
                                            double inv_s2, float inv_step, bool exact,
                                            float* row) {
    if constexpr (KB <= 5) {
        if (t.unimodal_gray) {
            if (exact)
                axis_exact_gray<KB>(t, x, inv_s2, inv_step, row);
            else
                axis_maxlog_gray<KB>(t, x, inv_s2f, inv_s2, row);
            return;
        }
    }
    constexpr int L = 1 << KB;
    const double xd = x;
    int a0[KB], a1[KB];
    if (!exact) {
        double p0[KB], p1[KB];
#pragma unroll
        for (int b = 0; b < KB; ++b) {
            p0[b] = -1.0e300;
            p1[b] = -1.0e300;
            a0[b] = 0;
            a1[b] = 0;
        }
#pragma unroll 4
        for (int l = 0; l < L; ++l) {
            const double dx = xd - t.lev[l];
            const double m = -(dx * dx) * inv_s2 + t.logp[l];
            const int lab = t.lab[l];
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                if ((lab >> (KB - 1 - b)) & 1) {
                    if (m > p1[b]) { p1[b] = m; a1[b] = l; }
                } else {
                    if (m > p0[b]) { p0[b] = m; a0[b] = l; }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < KB; ++b)
            row[t.pos[b]] = static_cast<float>(level_diff(t, xd, inv_s2, a0[b], a1[b]));
        return;
    }
    float p0[KB], p1[KB];
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        p0[b] = -ck::kInf;
        p1[b] = -ck::kInf;
        a0[b] = 0;
        a1[b] = 0;
    }
#pragma unroll 4
    for (int l = 0; l < L; ++l) {
        const float dx = x - t.levf[l];
        const float m = -(dx * dx) * inv_s2f + t.logpf[l];
        const int lab = t.lab[l];
#pragma unroll
        for (int b = 0; b < KB; ++b) {
            if ((lab >> (KB - 1 - b)) & 1) {
                if (m > p1[b]) { p1[b] = m; a1[b] = l; }
            } else {
                if (m > p0[b]) { p0[b] = m; a0[b] = l; }
            }
        }
    }
    const int pk = (p0[0] >= p1[0]) ? a0[0] : a1[0];
    double w0[KB], w1[KB];
    float u0[KB], u1[KB];
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        w0[b] = level_diff(t, xd, inv_s2, a0[b], pk);
        w1[b] = level_diff(t, xd, inv_s2, a1[b], pk);
        u0[b] = 0.0f;
        u1[b] = 0.0f;
    }
#pragma unroll 4
    for (int l = 0; l < L; ++l) {
        const float e = exp2f(static_cast<float>(level_diff(t, xd, inv_s2, l, pk)) * kLog2e);
        const int lab = t.lab[l];
#pragma unroll
        for (int b = 0; b < KB; ++b) {
            if ((lab >> (KB - 1 - b)) & 1) {
                u1[b] += (l == a1[b]) ? 0.0f : e;
            } else {
                u0[b] += (l == a0[b]) ? 0.0f : e;
            }
        }
    }
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        float t0, t1;
        if (w0[b] < kWeak) {
            t0 = 0.0f;
            for (int l = 0; l < L; ++l)
                if (!((t.lab[l] >> (KB - 1 - b)) & 1) && l != a0[b])
                    t0 += exp2f(static_cast<float>(level_diff(t, xd, inv_s2, l, a0[b])) * kLog2e);
        } else {
            t0 = u0[b] / exp2f(static_cast<float>(w0[b]) * kLog2e);
        }
        if (w1[b] < kWeak) {
            t1 = 0.0f;
            for (int l = 0; l < L; ++l)
                if (((t.lab[l] >> (KB - 1 - b)) & 1) && l != a1[b])
                    t1 += exp2f(static_cast<float>(level_diff(t, xd, inv_s2, l, a1[b])) * kLog2e);
        } else {
            t1 = u1[b] / exp2f(static_cast<float>(w1[b]) * kLog2e);
        }
        const double d01 = level_diff(t, xd, inv_s2, a0[b], a1[b]);
        row[t.pos[b]] = static_cast<float>(d01 + static_cast<double>(log1p_diff(t0, t1)));
    }
}

// CLOSED: closed form (max-log, uniform); else level metrics.
// inv_rows (optional): 1/sigma^2 per row of row_len symbols, instead of inv_s2.
#define CK_SEPARABLE_PARAMS                                                                      \
    const float2 *__restrict__ x, long long n, ck_axis re, ck_axis im, float inv_s2f,           \
        double inv_s2, int exact, float *__restrict__ out, const double *__restrict__ inv_rows, \
        long long row_len, const int *__restrict__ perm, long long period, long long limit
#define CK_SEPARABLE_ARGS \
    x, n, re, im, inv_s2f, inv_s2, exact, out, inv_rows, row_len, perm, period, limit

// The LDS tables of a separable kernel: the cell tables of the closed form
// (ck_closed.h) for axes of up to kClosedTableBits bits, else the per-level
// tables (level metrics, and the closed form of wider axes).
template <int KB, bool CLOSED>
constexpr bool kCellTables = CLOSED && KB <= ck::kClosedTableBits;

template <int KB, bool CLOSED, bool CELLS = kCellTables<KB, CLOSED>>
struct SeparableLds {
    AxisTab<KB> tab[2];
};

template <int KB, bool CLOSED>
struct SeparableLds<KB, CLOSED, true> {
    ck::ClosedAxis<KB> tab[2];
};

// Both axes' tables into LDS, then a barrier.
template <int KB, bool CLOSED>
__device__ __forceinline__ void stage_separable(SeparableLds<KB, CLOSED>& t, const ck_axis& re,
                                                const ck_axis& im) {
    if constexpr (kCellTables<KB, CLOSED>) {
        ck::stage_closed<KB>(t.tab[0], re);
        ck::stage_closed<KB>(t.tab[1], im);
    } else {
        stage_axis<KB>(t.tab[0], re, CLOSED);
        stage_axis<KB>(t.tab[1], im, CLOSED);
    }
    __syncthreads();
}

// 1 / level step of both axes (per-level tables only: the cell tables carry
// their own coordinate map).
template <int KB, bool CLOSED>
__device__ __forceinline__ float2 level_steps(const SeparableLds<KB, CLOSED>& t) {
    if constexpr (kCellTables<KB, CLOSED>) {
        return make_float2(0.0f, 0.0f);
    } else {
        return make_float2(1.0f / (t.tab[0].levf[1] - t.tab[0].levf[0]),
                           1.0f / (t.tab[1].levf[1] - t.tab[1].levf[0]));
    }
}

// The k LLRs of one symbol v into row: both axes of the staged tables.
template <int KB, bool CLOSED>
__device__ __forceinline__ void separable_symbol(SeparableLds<KB, CLOSED>& t, float2 v,
                                                 float inv_s2f, double inv_s2, float2 inv_step,
                                                 int exact, float* row) {
    if constexpr (kCellTables<KB, CLOSED>) {
        ck::closed_axis_llrs<KB>(t.tab[0], v.x, inv_s2, row);
        ck::closed_axis_llrs<KB>(t.tab[1], v.y, inv_s2, row);
    } else if constexpr (CLOSED) {
        axis_closed<KB>(t.tab[0], v.x, inv_s2, inv_step.x, row);
        axis_closed<KB>(t.tab[1], v.y, inv_s2, inv_step.y, row);
    } else {
        // The table index is opaque to the compiler, so the level tables are
        // read from LDS per symbol (broadcast reads) instead of held in
        // registers across the caller's loop: 2 x 2^KB float64 values per
        // axis would spill.
        int z = 0;
        asm volatile("" : "+v"(z));
        axis_levels<KB>(t.tab[z], v.x, inv_s2f, inv_s2, inv_step.x, exact != 0, row);
        axis_levels<KB>(t.tab[z + 1], v.y, inv_s2f, inv_s2, inv_step.y, exact != 0, row);
    }
}

template <int KB, bool CLOSED>
__device__ __forceinline__ void separable_body(CK_SEPARABLE_PARAMS) {
    constexpr int K = 2 * KB;
    __shared__ SeparableLds<KB, CLOSED> lt;
    __shared__ float tile[kThreads * (K | 1)];
    stage_separable<KB, CLOSED>(lt, re, im);
    const float2 inv_step = level_steps<KB, CLOSED>(lt);
    float* row = tile + threadIdx.x * tile_stride(K);
    for (long long base = static_cast<long long>(blockIdx.x) * kThreads; base < n;
         base += static_cast<long long>(gridDim.x) * kThreads) {
        const long long i = base + threadIdx.x;
        if (i < n) {
            const float2 v = x[i];
            if (inv_rows != nullptr) {
                const long long r = (n <= 0xffffffffLL)
                                        ? static_cast<long long>(static_cast<unsigned>(i) /
                                                                 static_cast<unsigned>(row_len))
                                        : i / row_len;
                inv_s2 = inv_rows[r];
                inv_s2f = static_cast<float>(inv_s2);
            }
            separable_symbol<KB, CLOSED>(lt, v, inv_s2f, inv_s2, inv_step, exact, row);
        }
        __syncthreads();
        const long long left = n - base;
        const int count = left < kThreads ? static_cast<int>(left) : kThreads;
        if (perm != nullptr)
            store_tile_perm<K>(tile, out, base, count, perm, period, limit);
        else
            store_tile(tile, out, base, count, K);
        __syncthreads();
    }
}

template <int KB, bool CLOSED>
__global__ __launch_bounds__(kThreads) void llr_separable_kernel(CK_SEPARABLE_PARAMS) {
    separable_body<KB, CLOSED>(CK_SEPARABLE_ARGS);
}

// Level metrics of axes of up to 16 levels: held to 128 VGPRs, four waves
// per SIMD, which the form of 256-QAM fits without spilling (unconstrained
// it takes 136 and three waves; wider axes would spill under the cap, and
// the closed form needs under 60).
template <int KB, bool CLOSED>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void
llr_separable_kernel_w4(CK_SEPARABLE_PARAMS) {
    separable_body<KB, CLOSED>(CK_SEPARABLE_ARGS);
}

// Up to kMaxSegments requests of one constellation in one launch, passed by
// value: segment j demaps its symbols x[j][0 .. n[j]) into out[j] (the plain
// layout, or the bit layout of perm[j] with period and limit as
// ck_llr_separable_permuted), 1/sigma^2 inv_s2[j] or, when inv_rows[j] is
// set, inv_rows[j][i / row_len[j]].  tile_end[j] is the prefix sum of the
// segments' 256-symbol tiles: a workgroup takes whole tiles of one segment,
// found by a uniform scan.
constexpr int kMaxSegments = 16;

struct Segments {
    const float2* x[kMaxSegments];
    float* out[kMaxSegments];
    const double* inv_rows[kMaxSegments];
    const int* perm[kMaxSegments];
    long long n[kMaxSegments];
    long long row_len[kMaxSegments];
    long long period[kMaxSegments];
    long long limit[kMaxSegments];
    long long tile_end[kMaxSegments];
    double inv_s2[kMaxSegments];
    int count;
};

template <int KB, bool CLOSED>
__device__ __forceinline__ void segments_body(const Segments& sg, const ck_axis& re,
                                              const ck_axis& im, int exact) {
    constexpr int K = 2 * KB;
    __shared__ SeparableLds<KB, CLOSED> lt;
    __shared__ float tile[kThreads * (K | 1)];
    stage_separable<KB, CLOSED>(lt, re, im);
    const float2 inv_step = level_steps<KB, CLOSED>(lt);
    float* row = tile + threadIdx.x * tile_stride(K);
    const long long tiles = sg.tile_end[sg.count - 1];
    for (long long tl = blockIdx.x; tl < tiles; tl += gridDim.x) {
        int j = 0;
        while (j + 1 < sg.count && sg.tile_end[j] <= tl) ++j;
        const long long base = (tl - (j > 0 ? sg.tile_end[j - 1] : 0)) * kThreads;
        const long long n = sg.n[j];
        const long long i = base + threadIdx.x;
        if (i < n) {
            const float2 v = sg.x[j][i];
            double inv_s2 = sg.inv_s2[j];
            const double* rows = sg.inv_rows[j];
            if (rows != nullptr) {
                const long long len = sg.row_len[j];
                inv_s2 = rows[(n <= 0xffffffffLL)
                                  ? static_cast<long long>(static_cast<unsigned>(i) /
                                                           static_cast<unsigned>(len))
                                  : i / len];
            }
            separable_symbol<KB, CLOSED>(lt, v, static_cast<float>(inv_s2), inv_s2, inv_step,
                                         exact, row);
        }
        __syncthreads();
        const int count = n - base < kThreads ? static_cast<int>(n - base) : kThreads;
        if (sg.perm[j] != nullptr)
            store_tile_perm<K>(tile, sg.out[j], base, count, sg.perm[j], sg.period[j],
                               sg.limit[j]);
        else
            store_tile(tile, sg.out[j], base, count, K);
        __syncthreads();
    }
}

template <int KB, bool CLOSED>
__global__ __launch_bounds__(kThreads) void llr_segments_kernel(Segments sg, ck_axis re,
                                                                ck_axis im, int exact) {
    segments_body<KB, CLOSED>(sg, re, im, exact);
}

template <int KB, bool CLOSED>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void
llr_segments_kernel_w4(Segments sg, ck_axis re, ck_axis im, int exact) {
    segments_body<KB, CLOSED>(sg, re, im, exact);
}

// m_a - m_c of two points in float64, factored per component.
__device__ __forceinline__ double point_diff(const double2* pts, const double* logp, double xr,
                                             double xi, double inv_s2, int a, int c) {
    const double2 pa = pts[a];
    const double2 pc = pts[c];
    const double dr = (pa.x - pc.x) * ((xr - pa.x) + (xr - pc.x));
    const double di = (pa.y - pc.y) * ((xi - pa.y) + (xi - pc.y));
    return (dr + di) * inv_s2 + (logp[a] - logp[c]);
}

constexpr int kPointsInLds = 1024;

// Constellations of up to 2^kOrderedBits points are staged in label order
// (point m has label m, when the labels are a permutation of 0 .. M-1, as
// every constellation's are) and the max-log point loop is unrolled, so the
// half of each bit a point belongs to is a compile-time constant.
constexpr int kOrderedBits = 8;

// The LLRs of one symbol (xr, xi) from the metrics of all M = 2^K points,
// into row[b].  ORDERED: the points are in label order and the loops
// unrolled.  Max-log locates each half's maximum in float64 (the maxima
// are the answer) and returns their factored float64 difference; ordered,
// a half's maximum is tracked as one float64 max of a key, the metric with
// its lowest K mantissa bits replaced by the point's label (metrics within
// 2^-(52-K) relative of each other may swap, and either gives the same LLR
// to that precision), the label of the maximum read back from the key.
// Exact takes the peak point (float32 metrics) as reference: every term
// 2^w, w = M_m - M_pk in log2 units (float64, then exp2_split), goes to one
// of the two float64 half sums of each bit; a point's label is the same in
// every lane, read on the scalar unit, and each half sum takes one FMA with
// a scalar-selected weight of 1 or 0 (exact: S + 1 e = S + e, S + 0 e = S).
// (Its point loops stay rolled: unrolled over 128 points the exact form ran
// 1.7 times slower.)
template <int K, bool ORDERED, bool STAGED>
__device__ __forceinline__ void points_symbol(const double2* pts, const double* logp,
                                              const double* logp2, const int* lab_tab,
                                              float2 v, float inv_s2f, double inv_s2,
                                              int exact, float* row) {
    constexpr int M = 1 << K;
    const double s2l = inv_s2 * kLog2eD;
    const double xr = v.x;
    const double xi = v.y;
    // The metric of point m in log2 units relative to a reference metric
    // (unstaged tables carry no log2 copy of the prior).
    const auto metric2 = [&](int m, double ref) {
        const double2 p = pts[m];
        const double dr = xr - p.x;
        const double di = xi - p.y;
        const double lp = STAGED ? logp2[m] : logp[m] * kLog2eD;
        return fma(-s2l, fma(dr, dr, di * di), lp - ref);
    };
    if (!exact) {
        int a0[K], a1[K];
        if constexpr (ORDERED) {
            double k0[K], k1[K];
#pragma unroll
            for (int b = 0; b < K; ++b) {
                k0[b] = -ck::kInf;
                k1[b] = -ck::kInf;
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const double2 p = pts[m];
                const double dr = xr - p.x;
                const double di = xi - p.y;
                const double met = -(dr * dr + di * di) * inv_s2 + logp[m];
                const double key = __longlong_as_double(
                    (__double_as_longlong(met) & ~static_cast<long long>(M - 1)) | m);
#pragma unroll
                for (int b = 0; b < K; ++b) {
                    if ((m >> (K - 1 - b)) & 1)
                        k1[b] = fmax(k1[b], key);
                    else
                        k0[b] = fmax(k0[b], key);
                }
            }
#pragma unroll
            for (int b = 0; b < K; ++b) {  // NaN symbols: keys stay -inf, label 0
                a0[b] = static_cast<int>(__double_as_longlong(k0[b]) & (M - 1));
                a1[b] = static_cast<int>(__double_as_longlong(k1[b]) & (M - 1));
            }
        } else {
            double p0[K], p1[K];
#pragma unroll
            for (int b = 0; b < K; ++b) {
                p0[b] = -1.0e300;
                p1[b] = -1.0e300;
                a0[b] = 0;
                a1[b] = 0;
            }
            for (int m = 0; m < M; ++m) {
                const double2 p = pts[m];
                const double dr = xr - p.x;
                const double di = xi - p.y;
                const double met = -(dr * dr + di * di) * inv_s2 + logp[m];
                const int lab = lab_tab[m];
#pragma unroll
                for (int b = 0; b < K; ++b) {
                    if ((lab >> (K - 1 - b)) & 1) {
                        if (met > p1[b]) { p1[b] = met; a1[b] = m; }
                    } else {
                        if (met > p0[b]) { p0[b] = met; a0[b] = m; }
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < K; ++b)
            row[b] = static_cast<float>(point_diff(pts, logp, xr, xi, inv_s2, a0[b], a1[b]));
        return;
    }
    float best = -ck::kInf;
    int pk = 0;
    for (int m = 0; m < M; ++m) {
        const double2 p = pts[m];
        const float dr = v.x - static_cast<float>(p.x);
        const float di = v.y - static_cast<float>(p.y);
        const float met = -(dr * dr + di * di) * inv_s2f + static_cast<float>(logp[m]);
        if (met > best) {
            best = met;
            pk = m;
        }
    }
    const double ref = metric2(pk, 0.0);
    double S0[K], S1[K];
#pragma unroll
    for (int b = 0; b < K; ++b) {
        S0[b] = 0.0;
        S1[b] = 0.0;
    }
    for (int m = 0; m < M; ++m) {
        const double e = exp2_split(metric2(m, ref));
        const int lab = __builtin_amdgcn_readfirstlane(lab_tab[m]);
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const bool one = (lab >> (K - 1 - b)) & 1;
            S1[b] = fma(one ? 1.0 : 0.0, e, S1[b]);
            S0[b] = fma(one ? 0.0 : 1.0, e, S0[b]);
        }
    }
    const int lpk = lab_tab[pk];
#pragma unroll
    for (int b = 0; b < K; ++b) {
        const int own = (lpk >> (K - 1 - b)) & 1;  // the peak's half: sum >= 1
        const double s_own = own ? S1[b] : S0[b];
        const double s_oth = own ? S0[b] : S1[b];
        float r;  // log(S_own / S_oth)
        if (s_oth >= kFarHalf) {
            r = log_ratio(s_own, s_oth);
        } else {
            // The other half lies more than 1000 octaves under the peak:
            // summed again relative to its own best.
            double mo = -1.0e300;
            for (int m = 0; m < M; ++m)
                if (((lab_tab[m] >> (K - 1 - b)) & 1) != own) mo = fmax(mo, metric2(m, ref));
            double u = 0.0;
            for (int m = 0; m < M; ++m)
                if (((lab_tab[m] >> (K - 1 - b)) & 1) != own)
                    u += exp2_split(metric2(m, ref) - mo);
            r = static_cast<float>(-mo * 0.6931471805599453 +
                                   static_cast<double>(log_of(s_own) - log_of(u)));
        }
        row[b] = own ? -r : r;
    }
}

template <int K>
__global__ __launch_bounds__(kThreads) void llr_points_kernel(
    const float2* __restrict__ x, long long n, const double2* __restrict__ g_pts,
    const double* __restrict__ g_logp, const int* __restrict__ g_lab, int M, float inv_s2f,
    double inv_s2, int exact, float* __restrict__ out) {
    __shared__ float tile[kThreads * (K | 1)];
    extern __shared__ char dyn[];
    // M = 2^K (checked by the host): the tables of up to kPointsInLds
    // points are staged in LDS, decided at compile time so that the
    // compiler reads them as LDS, not through generic pointers.
    constexpr bool staged = (1 << K) <= kPointsInLds;
    static_assert(!(K <= kOrderedBits) || staged, "ordered tables are staged");
    const double2* pts = g_pts;
    const double* logp = g_logp;
    const double* logp2 = nullptr;
    const int* lab_tab = g_lab;
    bool ordered = false;
    if constexpr (staged) {
        double2* s_pts = reinterpret_cast<double2*>(dyn);
        double* s_logp = reinterpret_cast<double*>(s_pts + M);
        double* s_logp2 = s_logp + M;
        int* s_lab = reinterpret_cast<int*>(s_logp2 + M);
        if constexpr (K <= kOrderedBits) {
            // Point i at position lab[i]; the staging is ordered when every
            // position m then holds label m (the labels are a permutation).
            for (int i = threadIdx.x; i < M; i += kThreads) s_lab[i] = -1;
            __syncthreads();
            for (int i = threadIdx.x; i < M; i += kThreads) {
                const int L = g_lab[i];
                if (L >= 0 && L < M) {
                    s_pts[L] = g_pts[i];
                    s_logp[L] = g_logp[i];
                    s_logp2[L] = g_logp[i] * kLog2eD;
                    s_lab[L] = L;
                }
            }
            __syncthreads();
            bool ok = true;
            for (int i = threadIdx.x; i < M; i += kThreads) ok = ok && s_lab[i] == i;
            ordered = __syncthreads_and(ok ? 1 : 0) != 0;
        }
        if (!ordered) {
            for (int i = threadIdx.x; i < M; i += kThreads) {
                s_pts[i] = g_pts[i];
                s_logp[i] = g_logp[i];
                s_logp2[i] = g_logp[i] * kLog2eD;
                s_lab[i] = g_lab[i];
            }
        }
        pts = s_pts;
        logp = s_logp;
        logp2 = s_logp2;
        lab_tab = s_lab;
    }
    __syncthreads();
    float* row = tile + threadIdx.x * tile_stride(K);
    for (long long base = static_cast<long long>(blockIdx.x) * kThreads; base < n;
         base += static_cast<long long>(gridDim.x) * kThreads) {
        const long long i = base + threadIdx.x;
        if (i < n) {
            const float2 v = x[i];
            if constexpr (K <= kOrderedBits) {
                if (ordered)
                    points_symbol<K, true, staged>(pts, logp, logp2, lab_tab, v, inv_s2f, inv_s2,
                                                   exact, row);
                else
                    points_symbol<K, false, staged>(pts, logp, logp2, lab_tab, v, inv_s2f,
                                                    inv_s2, exact, row);
            } else {
                points_symbol<K, false, staged>(pts, logp, logp2, lab_tab, v, inv_s2f, inv_s2,
                                                exact, row);
            }
        }
        __syncthreads();
        const long long left = n - base;
        store_tile(tile, out, base, left < kThreads ? static_cast<int>(left) : kThreads, K);
        __syncthreads();
    }
}

// Workgroups of a grid-stride launch over n symbols: per_cu per CU at most.
// The cell-table closed form streams at close to the copy rate with 8 per
// CU (fewer trips stage the tables; measured 8 against 16: 79 against 87 us
// on 12.3 M symbols of 64-QAM); the level-metric forms keep 16 to hide
// their longer per-symbol latency.
int grid_for(long long n, int per_cu = 16) {
    ck::DeviceInfo info{};
    if (ck::device_info(&info) != CK_OK) info.num_cus = 256;
    const long long tiles = (n + kThreads - 1) / kThreads;
    const long long cap = static_cast<long long>(info.num_cus) * per_cu;
    return static_cast<int>(tiles < cap ? tiles : cap);
}

template <int KB, bool CLOSED>
constexpr int kPerCu = kCellTables<KB, CLOSED> ? 8 : 16;

int32_t check_axis(const ck_axis* a, const char* name, bool closed) {
    CK_REQUIRE(a != nullptr, "%s axis is NULL", name);
    CK_REQUIRE(a->num_bits >= 1 && a->num_bits <= 8, "%s axis: num_bits %d not in [1, 8]", name,
               a->num_bits);
    CK_REQUIRE(a->num_levels == (1 << a->num_bits), "%s axis: num_levels %d != 2^num_bits", name,
               a->num_levels);
    CK_REQUIRE(a->levels && a->log_prior && a->labels && a->pos, "%s axis: NULL table", name);
    CK_REQUIRE(!closed || (a->lo && a->hi), "%s axis: NULL opposite-level table", name);
    return CK_OK;
}

// The kernel instantiation for KB bits per axis (only the chosen wrapper is
// instantiated).
template <int KB, bool CLOSED>
constexpr auto separable_kernel() {
    if constexpr (!CLOSED && KB <= 4)
        return llr_separable_kernel_w4<KB, CLOSED>;
    else
        return llr_separable_kernel<KB, CLOSED>;
}

template <bool CLOSED>
int32_t launch_separable(const void* x, int64_t n, const ck_axis* re, const ck_axis* im,
                         double inv_s2, int exact, float* out, void* stream,
                         const double* inv_rows = nullptr, int64_t row_len = 1,
                         const int32_t* perm = nullptr, int64_t period = 1, int64_t limit = 0) {
    CK_REQUIRE(n >= 0, "n = %lld < 0", static_cast<long long>(n));
    if (int32_t st = check_axis(re, "re", CLOSED)) return st;
    if (int32_t st = check_axis(im, "im", CLOSED)) return st;
    CK_REQUIRE(re->num_bits == im->num_bits,
               "separable axes must carry the same number of bits (%d, %d)", re->num_bits,
               im->num_bits);
    CK_REQUIRE(inv_s2 > 0.0 && inv_s2 < 1e300, "inv_s2 must be positive and finite");
    if (n == 0) return CK_OK;
    CK_REQUIRE(x != nullptr && out != nullptr, "NULL symbols or output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const float2* xs = static_cast<const float2*>(x);
    switch (re->num_bits) {
#define CK_CASE(KB)                                                                           \
    case KB:                                                                                  \
        hipLaunchKernelGGL((separable_kernel<KB, CLOSED>()),                                  \
                           dim3(grid_for(n, kPerCu<KB, CLOSED>)), dim3(kThreads), 0,          \
                           s, xs, static_cast<long long>(n), *re, *im,                        \
                           static_cast<float>(inv_s2), inv_s2, exact, out, inv_rows,          \
                           static_cast<long long>(row_len), perm,                             \
                           static_cast<long long>(period), static_cast<long long>(limit));    \
        break;
        CK_CASE(1) CK_CASE(2) CK_CASE(3) CK_CASE(4) CK_CASE(5) CK_CASE(6) CK_CASE(7) CK_CASE(8)
#undef CK_CASE
    }
    return ck::launched(CLOSED ? "llr_separable_maxlog" : "llr_separable");
}

}  // namespace

extern "C" int32_t ck_llr_separable_maxlog(const void* x, int64_t n, const ck_axis* re,
                                           const ck_axis* im, double inv_s2, float* out,
                                           void* stream) {
    return launch_separable<true>(x, n, re, im, inv_s2, 0, out, stream);
}

extern "C" int32_t ck_llr_separable(const void* x, int64_t n, const ck_axis* re,
                                    const ck_axis* im, double inv_s2, int32_t exact, float* out,
                                    void* stream) {
    CK_REQUIRE(exact == 0 || exact == 1, "exact must be 0 or 1, got %d", exact);
    return launch_separable<false>(x, n, re, im, inv_s2, exact, out, stream);
}

extern "C" int32_t ck_llr_separable_rows(const void* x, int64_t rows, int64_t row_len,
                                         const ck_axis* re, const ck_axis* im,
                                         const double* inv_s2, int32_t closed, int32_t exact,
                                         float* out, void* stream) {
    CK_REQUIRE(rows >= 0 && row_len >= 1, "rows %lld, row_len %lld", static_cast<long long>(rows),
               static_cast<long long>(row_len));
    CK_REQUIRE(closed == 0 || closed == 1, "closed must be 0 or 1, got %d", closed);
    CK_REQUIRE(exact == 0 || exact == 1, "exact must be 0 or 1, got %d", exact);
    CK_REQUIRE(!(closed && exact), "the closed form is max-log only");
    CK_REQUIRE(rows <= INT64_MAX / row_len, "rows x row_len overflows");
    if (rows == 0) return CK_OK;
    CK_REQUIRE(inv_s2 != nullptr, "NULL inv_s2");
    // 1 is a placeholder for the scalar check; the kernel reads inv_s2[row].
    return closed ? launch_separable<true>(x, rows * row_len, re, im, 1.0, 0, out, stream, inv_s2,
                                           row_len)
                  : launch_separable<false>(x, rows * row_len, re, im, 1.0, exact, out, stream,
                                            inv_s2, row_len);
}

extern "C" int32_t ck_llr_separable_permuted(const void* x, int64_t n, const ck_axis* re,
                                             const ck_axis* im, double inv_s2, int32_t closed,
                                             int32_t exact, float* out, const int32_t* perm,
                                             int64_t period, int64_t limit, void* stream) {
    CK_REQUIRE(closed == 0 || closed == 1, "closed must be 0 or 1, got %d", closed);
    CK_REQUIRE(exact == 0 || exact == 1, "exact must be 0 or 1, got %d", exact);
    CK_REQUIRE(!(closed && exact), "the closed form is max-log only");
    CK_REQUIRE(n >= 0 && re != nullptr, "n = %lld or NULL axis", static_cast<long long>(n));
    CK_REQUIRE(period >= 1 && limit >= 0 && limit % period == 0,
               "limit %lld is not a whole number of blocks of %lld", static_cast<long long>(limit),
               static_cast<long long>(period));
    CK_REQUIRE(limit <= n * 2 * re->num_bits, "limit %lld past the %lld LLRs",
               static_cast<long long>(limit), static_cast<long long>(n * 2 * re->num_bits));
    if (n == 0 || limit == 0) return CK_OK;
    CK_REQUIRE(perm != nullptr, "NULL perm");
    // Only the symbols holding LLRs below the limit are demapped.
    const int64_t k = 2 * re->num_bits;
    const int64_t m = (limit + k - 1) / k;
    return closed ? launch_separable<true>(x, m, re, im, inv_s2, 0, out, stream, nullptr, 1, perm,
                                           period, limit)
                  : launch_separable<false>(x, m, re, im, inv_s2, exact, out, stream, nullptr, 1,
                                            perm, period, limit);
}

namespace {

template <int KB, bool CLOSED>
constexpr auto segments_kernel() {
    if constexpr (!CLOSED && KB <= 4)
        return llr_segments_kernel_w4<KB, CLOSED>;
    else
        return llr_segments_kernel<KB, CLOSED>;
}

template <bool CLOSED>
int32_t launch_segments(const Segments& sg, const ck_axis* re, const ck_axis* im, int exact,
                        hipStream_t s) {
    const long long tiles = sg.tile_end[sg.count - 1];
    switch (re->num_bits) {
#define CK_CASE(KB)                                                                        \
    case KB:                                                                               \
        hipLaunchKernelGGL((segments_kernel<KB, CLOSED>()),                                \
                           dim3(grid_for(tiles * kThreads, kPerCu<KB, CLOSED>)),           \
                           dim3(kThreads), 0,                                              \
                           s, sg, *re, *im, exact);                                        \
        break;
        CK_CASE(1) CK_CASE(2) CK_CASE(3) CK_CASE(4) CK_CASE(5) CK_CASE(6) CK_CASE(7) CK_CASE(8)
#undef CK_CASE
    }
    return ck::launched("llr_separable_segments");
}

}  // namespace

extern "C" int32_t ck_llr_separable_segments(const ck_llr_segment* segs, int32_t count,
                                             const ck_axis* re, const ck_axis* im,
                                             int32_t closed, int32_t exact, void* stream) {
    CK_REQUIRE(count >= 0, "count = %d < 0", count);
    CK_REQUIRE(closed == 0 || closed == 1, "closed must be 0 or 1, got %d", closed);
    CK_REQUIRE(exact == 0 || exact == 1, "exact must be 0 or 1, got %d", exact);
    CK_REQUIRE(!(closed && exact), "the closed form is max-log only");
    if (count == 0) return CK_OK;
    CK_REQUIRE(segs != nullptr, "NULL segments");
    if (int32_t st = check_axis(re, "re", closed != 0)) return st;
    if (int32_t st = check_axis(im, "im", closed != 0)) return st;
    CK_REQUIRE(re->num_bits == im->num_bits,
               "separable axes must carry the same number of bits (%d, %d)", re->num_bits,
               im->num_bits);
    const int64_t k = 2 * re->num_bits;
    // Every segment is checked before anything is launched.
    for (int32_t j = 0; j < count; ++j) {
        const ck_llr_segment& g = segs[j];
        CK_REQUIRE(g.n >= 0, "segment %d: n = %lld < 0", j, static_cast<long long>(g.n));
        if (g.inv_rows != nullptr)
            CK_REQUIRE(g.row_len >= 1, "segment %d: row_len %lld < 1", j,
                       static_cast<long long>(g.row_len));
        else
            CK_REQUIRE(g.inv_s2 > 0.0 && g.inv_s2 < 1e300,
                       "segment %d: inv_s2 must be positive and finite", j);
        if (g.perm != nullptr) {
            CK_REQUIRE(g.period >= 1 && g.limit >= 0 && g.limit % g.period == 0,
                       "segment %d: limit %lld is not a whole number of blocks of %lld", j,
                       static_cast<long long>(g.limit), static_cast<long long>(g.period));
            CK_REQUIRE(g.limit <= g.n * k, "segment %d: limit %lld past the %lld LLRs", j,
                       static_cast<long long>(g.limit), static_cast<long long>(g.n * k));
        }
        CK_REQUIRE(g.n == 0 || (g.x != nullptr && g.out != nullptr),
                   "segment %d: NULL symbols or output", j);
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    Segments sg{};
    long long tiles = 0;
    for (int32_t j = 0; j <= count; ++j) {
        if (j == count || sg.count == kMaxSegments) {
            if (sg.count > 0) {
                const int32_t st = closed ? launch_segments<true>(sg, re, im, 0, s)
                                          : launch_segments<false>(sg, re, im, exact, s);
                if (st != CK_OK) return st;
            }
            sg = Segments{};
            tiles = 0;
            if (j == count) break;
        }
        const ck_llr_segment& g = segs[j];
        // With a layout, only the symbols holding LLRs below the limit.
        const long long n = g.perm != nullptr ? (g.limit + k - 1) / k : g.n;
        if (n == 0) continue;
        const int c = sg.count++;
        sg.x[c] = static_cast<const float2*>(g.x);
        sg.out[c] = g.out;
        sg.inv_rows[c] = g.inv_rows;
        sg.perm[c] = g.perm;
        sg.n[c] = n;
        sg.row_len[c] = g.inv_rows != nullptr ? g.row_len : 1;
        sg.period[c] = g.perm != nullptr ? g.period : 1;
        sg.limit[c] = g.perm != nullptr ? g.limit : 0;
        sg.inv_s2[c] = g.inv_rows != nullptr ? 1.0 : g.inv_s2;
        tiles += (n + kThreads - 1) / kThreads;
        sg.tile_end[c] = tiles;
    }
    return CK_OK;
}

extern "C" int32_t ck_llr_points(const void* x, int64_t n, const void* points,
                                 const double* log_prior, const int32_t* labels, int32_t M,
                                 int32_t k, double inv_s2, int32_t exact, float* out,
                                 void* stream) {
    CK_REQUIRE(n >= 0, "n = %lld < 0", static_cast<long long>(n));
    CK_REQUIRE(k >= 1 && k <= 16, "k = %d not in [1, 16]", k);
    CK_REQUIRE(M == (1 << k), "M = %d != 2^k", M);
    CK_REQUIRE(exact == 0 || exact == 1, "exact must be 0 or 1, got %d", exact);
    CK_REQUIRE(inv_s2 > 0.0 && inv_s2 < 1e300, "inv_s2 must be positive and finite");
    if (n == 0) return CK_OK;
    CK_REQUIRE(x && out && points && log_prior && labels, "NULL pointer argument");
    const int grid = grid_for(n);
    const size_t dyn = M <= kPointsInLds ? static_cast<size_t>(M) * 36 : 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const float2* xs = static_cast<const float2*>(x);
    const double2* pts = static_cast<const double2*>(points);
    switch (k) {
#define CK_CASE(K)                                                                           \
    case K:                                                                                  \
        hipLaunchKernelGGL((llr_points_kernel<K>), dim3(grid), dim3(kThreads), dyn, s, xs,   \
                           static_cast<long long>(n), pts, log_prior, labels, M,             \
                           static_cast<float>(inv_s2), inv_s2, exact, out);                  \
        break;
        CK_CASE(1) CK_CASE(2) CK_CASE(3) CK_CASE(4) CK_CASE(5) CK_CASE(6) CK_CASE(7) CK_CASE(8)
        CK_CASE(9) CK_CASE(10) CK_CASE(11) CK_CASE(12) CK_CASE(13) CK_CASE(14) CK_CASE(15)
        CK_CASE(16)
#undef CK_CASE
    }
    return ck::launched("llr_points");
}
