// Constant-composition distribution matching (CCDM) on gfx950.
//
// The matcher of commkit/coding/shaping.py:69-225 maps k uniform bits to a
// block of n constellation points of a fixed composition by a 64-bit range
// coder (commkit/coding/_numba_loops.py:944-1102): at each point the range
// `rng` (kept in [2^31, 2^32) by doubling) is split among the points in
// proportion to their remaining counts, floor(rng * c_i / remaining) each,
// the rounding remainder going to the largest count (lowest index among
// equals).  The matcher follows the input's binary fraction through those
// splits; the dematcher rebuilds the start of the final interval exactly
// (32-bit limbs) and rounds it up to k bits, flagging any sequence that is
// not the image of a payload.
//
// A block is a sequential recursion over its points.  Up to 16 points a
// group of lanes (one per point, a power of two) decodes a block, splitting
// each position's divisions and combining them with DPP moves; larger
// alphabets take one lane per block.  The composition (at most kMaxPoints
// points) comes in the kernel arguments; each lane keeps its counts and the
// current split in registers (up to 16 points; larger alphabets compile to
// a variant whose arrays the compiler keeps in scratch).  floor(a /
// remaining) with a < 2^62 is a float64 estimate from one reciprocal per
// position, corrected exactly with 64-bit integer arithmetic.  The
// dematcher accumulates the interval start in a 128-bit register window
// and writes the limbs it leaves behind to a caller workspace of
// buffer_bits / 32 + 2 limbs per block, limb-major.
#include "ck_common.h"
#include "ck_tail.h"

#include <cstdint>

namespace {

constexpr int kMaxPoints = 64;
constexpr int kLanes = 64;  // blocks per workgroup (one wave)
constexpr int64_t kTop = int64_t{1} << 32;
constexpr int64_t kHalf = int64_t{1} << 31;

struct Composition {
    int64_t count[kMaxPoints];
};

// floor(a / d) for 0 <= a < 2^62, 1 <= d < 2^31, inv ~ 1 / d: the float64
// estimate is within one of the quotient; the remainder fixes it exactly.
__device__ __forceinline__ int64_t div_floor(int64_t a, int64_t d, double inv) {
    int64_t q = static_cast<int64_t>(static_cast<double>(a) * inv);
    int64_t r = a - q * d;
    while (r < 0) {
        --q;
        r += d;
    }
    while (r >= d) {
        ++q;
        r -= d;
    }
    return q;
}

// A lane's remaining counts and the current split, MB >= M entries (the
// unused ones count zero, get no share and are never chosen).  Loops over
// MB unroll, so with MB <= 16 both arrays live in registers.
template <int MB>
struct Coder {
    int cnt[MB];
    int64_t share[MB];

    __device__ __forceinline__ void init(const Composition& comp, int M) {
#pragma unroll
        for (int i = 0; i < MB; ++i) cnt[i] = i < M ? static_cast<int>(comp.count[i]) : 0;
    }

    // floor(rng * c_i / remaining) each, the remainder to the largest count
    // (the lowest index among equals).
    __device__ __forceinline__ void split(int64_t rng, int64_t remaining) {
        const double inv = 1.0 / static_cast<double>(remaining);
        int big = 0;
        int cbig = cnt[0];
        int64_t rest = rng;
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            share[i] = div_floor(rng * cnt[i], remaining, inv);
            rest -= share[i];
            const bool more = cnt[i] > cbig;
            big = more ? i : big;
            cbig = more ? cnt[i] : cbig;
        }
#pragma unroll
        for (int i = 0; i < MB; ++i) share[i] += (i == big) ? rest : 0;
    }

    // The point whose interval holds d: the first s with d < cum_s + share_s;
    // its cumulative start and share.
    __device__ __forceinline__ int find(int64_t d, int64_t* cum_out, int64_t* share_out) const {
        int64_t cum = 0;
        int s = -1;
        int64_t cs = 0;
        int64_t ss = 0;
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            const bool hit = s < 0 && d < cum + share[i];
            s = hit ? i : s;
            cs = hit ? cum : cs;
            ss = hit ? share[i] : ss;
            cum += share[i];
        }
        *cum_out = cs;
        *share_out = ss;
        return s < 0 ? 0 : s;
    }

    __device__ __forceinline__ int64_t start_of(int s) const {
        int64_t cum = 0;
#pragma unroll
        for (int i = 0; i < MB; ++i) cum += (i < s) ? share[i] : 0;
        return cum;
    }

    __device__ __forceinline__ int64_t share_of(int s) const {
        int64_t v = 0;
#pragma unroll
        for (int i = 0; i < MB; ++i) v = (i == s) ? share[i] : v;
        return v;
    }

    __device__ __forceinline__ int count_of(int s) const {
        int v = 0;
#pragma unroll
        for (int i = 0; i < MB; ++i) v = (i == s) ? cnt[i] : v;
        return v;
    }

    __device__ __forceinline__ void take(int s) {
#pragma unroll
        for (int i = 0; i < MB; ++i) cnt[i] -= (i == s) ? 1 : 0;
    }
};

template <int MB>
__global__ __launch_bounds__(kLanes) void ccdm_match_kernel(const signed char* __restrict__ bits,
                                                            long long R, long long k, Composition comp,
                                                            int M, long long n, int* __restrict__ out) {
    const long long b = static_cast<long long>(blockIdx.x) * kLanes + threadIdx.x;
    if (b >= R) return;  // no barriers below: each lane works alone
    Coder<MB> cd;
    cd.init(comp, M);
    const signed char* u = bits + b * k;
    int* o = out + b * n;
    long long pos = 0;
    int64_t d = 0;
    for (int j = 0; j < 32; ++j, ++pos) d = (d << 1) | (pos < k ? u[pos] : 0);
    int64_t rng = kTop;
    int64_t remaining = n;
    for (long long t = 0; t < n; ++t) {
        cd.split(rng, remaining);
        int64_t cum;
        int64_t sh;
        const int s = cd.find(d, &cum, &sh);
        o[t] = s;
        d -= cum;
        rng = sh;
        cd.take(s);
        --remaining;
        for (int g = 0; g < 32 && rng < kHalf; ++g) {  // rng >= 1: at most 31
            rng <<= 1;
            d = (d << 1) | (pos < k ? u[pos] : 0);
            ++pos;
        }
    }
}

// Bit 2^-j (j >= 1) of a block's limbs (limb l at low[l * R]).
__device__ __forceinline__ int limb_bit(const unsigned* low, long long R, long long j) {
    return static_cast<int>((low[((j - 1) >> 5) * R] >> (31 - ((j - 1) & 31))) & 1u);
}

// The start of the interval, low = sum over positions of cum 2^-shift, is
// accumulated in a 128-bit register window over limbs base - 2 .. base + 1
// (hi: the first two, lo: the last two; limbs below zero are not stored).
// Each addition lands on limbs (i - 1, i) with i = (shift - 1) / 32, which
// only grows: the window slides down to base = i - 1 first, writing out its
// first limb, which no later addition reaches except through a carry out
// of all 128 bits (two limbs of ones), taken through memory.
__device__ __forceinline__ void emit(unsigned* low, long long R, long long limbs, long long l,
                                     uint64_t word) {
    if (l >= 0 && l < limbs) low[l * R] = static_cast<unsigned>(word);
}

template <int MB>
__global__ __launch_bounds__(kLanes) void ccdm_dematch_kernel(
    const int* __restrict__ symbols, long long R, long long n, Composition comp, int M,
    long long k, long long buffer_bits, signed char* __restrict__ bits,
    unsigned char* __restrict__ valid, unsigned* __restrict__ ws) {
    const long long b = static_cast<long long>(blockIdx.x) * kLanes + threadIdx.x;
    if (b >= R) return;
    Coder<MB> cd;
    cd.init(comp, M);
    const long long limbs = buffer_bits / 32 + 2;
    unsigned* low = ws + b;  // limb l at low[l * R]
    for (long long l = 0; l < limbs; ++l) low[l * R] = 0u;
    const int* sym = symbols + b * n;
    int64_t rng = kTop;
    long long shift = 32;
    int64_t remaining = n;
    bool ok = true;
    uint64_t hi = 0;  // limbs base - 2, base - 1
    uint64_t lo = 0;  // limbs base, base + 1
    long long base = -1;
    for (long long t = 0; t < n; ++t) {
        const int s = sym[t];
        if (s < 0 || s >= M || cd.count_of(s) == 0) {
            ok = false;
            break;
        }
        cd.split(rng, remaining);
        const int64_t cum = cd.start_of(s);
        const long long i = (shift - 1) >> 5;
        while (base < i - 1) {
            emit(low, R, limbs, base - 2, hi >> 32);
            hi = (hi << 32) | (lo >> 32);
            lo <<= 32;
            ++base;
        }
        // low += cum 2^-shift: v on limbs (i - 1, i) = the window's lo.
        const uint64_t v = static_cast<uint64_t>(cum) << (31 - ((shift - 1) & 31));
        lo += v;
        if (lo < v) {
            hi += 1;
            if (hi == 0) {  // through limbs base - 3, base - 4, ... (none above 2^-1)
                for (long long l = base - 3; l >= 0; --l) {
                    const uint64_t x = static_cast<uint64_t>(low[l * R]) + 1u;
                    low[l * R] = static_cast<unsigned>(x);
                    if ((x >> 32) == 0) break;
                }
            }
        }
        rng = cd.share_of(s);
        cd.take(s);
        --remaining;
        for (int g = 0; g < 32 && rng < kHalf; ++g) {  // rng >= 1: at most 31
            rng <<= 1;
            ++shift;
        }
        if (shift > buffer_bits) {
            ok = false;
            break;
        }
    }
    emit(low, R, limbs, base - 2, hi >> 32);
    emit(low, R, limbs, base - 1, hi);
    emit(low, R, limbs, base, lo >> 32);
    emit(low, R, limbs, base + 1, lo);
    // v = ceil(low 2^k): the first k bits, plus one if any bit follows.
    signed char* out = bits + b * k;
    const long long first = k >> 5;
    bool beyond = (low[first * R] & (0xFFFFFFFFu >> (k & 31))) != 0u;
    for (long long l = first + 1; l < limbs && !beyond; ++l) beyond = low[l * R] != 0u;
    for (long long j = 1; j <= k; ++j) out[j - 1] = static_cast<signed char>(limb_bit(low, R, j));
    bool up = beyond;
    for (long long j = k; up && j > 0; --j) {
        up = out[j - 1] != 0;
        out[j - 1] = static_cast<signed char>(out[j - 1] ^ 1);
    }
    if (ok && up) ok = false;  // v = 2^k: no payload
    if (ok && beyond) {
        // v 2^-k must lie in [low, low + rng 2^-shift): every bit of low after
        // the k-th above its last 32 is one, and 2^w - (those bits) < rng.
        const long long lo_bit = (k + 1 > shift - 31) ? k + 1 : shift - 31;
        for (long long j = k + 1; j < lo_bit && ok; ++j)
            if (!limb_bit(low, R, j)) ok = false;
        if (ok) {
            int64_t f = 0;
            for (long long j = lo_bit; j <= shift; ++j) f = (f << 1) | limb_bit(low, R, j);
            ok = (int64_t{1} << (shift - lo_bit + 1)) - f < rng;
        }
    }
    valid[b] = ok ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Group kernels (up to 16 points): G lanes per block, lane j holds point j's
// count and computes its share, so a position's divisions run side by side;
// the group's sum, argmax, prefix and broadcasts are DPP moves within the
// lane's row of 16 (G divides 16, so a group never straddles a row).  Every
// lane of a group follows the same positions and keeps the same copy of the
// block's range state; lane 0 writes what the block writes.
// ---------------------------------------------------------------------------

// floor(rng * c / d) for rng <= 2^32, c <= d < 2^31: a float64 estimate,
// then one exact 64-bit remainder check each way.
__device__ __forceinline__ int64_t share_of_point(uint64_t rng, uint32_t c, uint32_t d, double inv) {
    const uint64_t a = rng * c;
    uint64_t q = static_cast<uint64_t>(static_cast<double>(rng) * (static_cast<double>(c) * inv));
    const uint64_t qd = q * d;
    if (qd > a) {
        --q;
    } else if (a - qd >= d) {
        ++q;
    }
    return static_cast<int64_t>(q);
}

template <int CTRL>
__device__ __forceinline__ int dpp(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ int64_t dpp64(int64_t v) {
    const int lo = dpp<CTRL>(static_cast<int>(static_cast<uint64_t>(v)));
    const int hi = dpp<CTRL>(static_cast<int>(static_cast<uint64_t>(v) >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(static_cast<unsigned>(hi)) << 32) |
                                static_cast<unsigned>(lo));
}

// The partner of each butterfly step: xor 1, xor 2 (quad permutes), the
// mirror within 8 and within 16 lanes; after the steps up to G every lane of
// the group holds the group's result.
constexpr int kXor1 = 0xB1, kXor2 = 0x4E, kHalfMirror = 0x141, kMirror = 0x140;

template <int G>
__device__ __forceinline__ int64_t group_sum(int64_t v) {
    if constexpr (G >= 2) v += dpp64<kXor1>(v);
    if constexpr (G >= 4) v += dpp64<kXor2>(v);
    if constexpr (G >= 8) v += dpp64<kHalfMirror>(v);
    if constexpr (G >= 16) v += dpp64<kMirror>(v);
    return v;
}

template <int G>
__device__ __forceinline__ int group_max(int v) {
    if constexpr (G >= 2) v = max(v, dpp<kXor1>(v));
    if constexpr (G >= 4) v = max(v, dpp<kXor2>(v));
    if constexpr (G >= 8) v = max(v, dpp<kHalfMirror>(v));
    if constexpr (G >= 16) v = max(v, dpp<kMirror>(v));
    return v;
}

template <int G>
__device__ __forceinline__ int group_min(int v) {
    if constexpr (G >= 2) v = min(v, dpp<kXor1>(v));
    if constexpr (G >= 4) v = min(v, dpp<kXor2>(v));
    if constexpr (G >= 8) v = min(v, dpp<kHalfMirror>(v));
    if constexpr (G >= 16) v = min(v, dpp<kMirror>(v));
    return v;
}

// Exclusive prefix sum over the group's lanes 0..j-1 (row shifts; a lane
// keeps a shifted value only from its own group, j >= the shift).  Each
// move runs with every lane of the group active and the choice comes
// after it: a move under a narrower mask would read its source lane as
// disabled, which returns zero.
template <int G>
__device__ __forceinline__ int64_t group_prefix(int64_t v, int j) {
    int64_t x = v;
    if constexpr (G >= 2) {
        const int64_t y = dpp64<0x111>(x);
        x += y & -static_cast<int64_t>(j >= 1);
    }
    if constexpr (G >= 4) {
        const int64_t y = dpp64<0x112>(x);
        x += y & -static_cast<int64_t>(j >= 2);
    }
    if constexpr (G >= 8) {
        const int64_t y = dpp64<0x114>(x);
        x += y & -static_cast<int64_t>(j >= 4);
    }
    if constexpr (G >= 16) {
        const int64_t y = dpp64<0x118>(x);
        x += y & -static_cast<int64_t>(j >= 8);
    }
    return x - v;
}

// One position of the coder for the group: lane j's share (the remainder
// added at the largest count, lowest index among equals) and its start.
template <int G>
__device__ __forceinline__ void group_split(int cnt, int j, int64_t rng, int64_t remaining,
                                            int64_t* share, int64_t* start) {
    const double inv = 1.0 / static_cast<double>(remaining);
    int64_t sh = share_of_point(static_cast<uint64_t>(rng), static_cast<uint32_t>(cnt),
                                static_cast<uint32_t>(remaining), inv);
    const int64_t rest = rng - group_sum<G>(sh);
    const int cmax = group_max<G>(cnt);
    const int big = group_min<G>(cnt == cmax ? j : G);
    sh += (j == big) ? rest : 0;
    *share = sh;
    *start = group_prefix<G>(sh, j);
}

template <int G>
__global__ __launch_bounds__(kLanes) void ccdm_match_group(const signed char* __restrict__ bits,
                                                           long long R, long long k, Composition comp,
                                                           int M, long long n, int* __restrict__ out) {
    const int j = static_cast<int>(threadIdx.x) & (G - 1);
    const long long b = (static_cast<long long>(blockIdx.x) * kLanes + threadIdx.x) / G;
    if (b >= R) return;  // whole groups leave together
    int cnt = j < M ? static_cast<int>(comp.count[j]) : 0;
    const signed char* u = bits + b * k;
    int* o = out + b * n;
    long long pos = 0;
    int64_t d = 0;
    for (int i = 0; i < 32; ++i, ++pos) d = (d << 1) | (pos < k ? u[pos] : 0);
    int64_t rng = kTop;
    int64_t remaining = n;
    for (long long t = 0; t < n; ++t) {
        int64_t sh;
        int64_t start;
        group_split<G>(cnt, j, rng, remaining, &sh, &start);
        const bool hit = sh > 0 && d >= start && d < start + sh;
        const int s = group_min<G>(hit ? j : G);
        const int64_t cum = group_sum<G>(j == s ? start : 0);
        rng = group_sum<G>(j == s ? sh : 0);
        if (j == 0) o[t] = s;
        d -= cum;
        cnt -= (j == s) ? 1 : 0;
        --remaining;
        for (int g = 0; g < 32 && rng < kHalf; ++g) {  // rng >= 1: at most 31
            rng <<= 1;
            d = (d << 1) | (pos < k ? u[pos] : 0);
            ++pos;
        }
    }
}

template <int G>
__global__ __launch_bounds__(kLanes) void ccdm_dematch_group(
    const int* __restrict__ symbols, long long R, long long n, Composition comp, int M,
    long long k, long long buffer_bits, signed char* __restrict__ bits,
    unsigned char* __restrict__ valid, unsigned* __restrict__ ws) {
    const int j = static_cast<int>(threadIdx.x) & (G - 1);
    const long long b = (static_cast<long long>(blockIdx.x) * kLanes + threadIdx.x) / G;
    if (b >= R) return;
    int cnt = j < M ? static_cast<int>(comp.count[j]) : 0;
    const long long limbs = buffer_bits / 32 + 2;
    unsigned* low = ws + b;  // limb l at low[l * R]
    for (long long l = j; l < limbs; l += G) low[l * R] = 0u;
    const int* sym = symbols + b * n;
    int64_t rng = kTop;
    long long shift = 32;
    int64_t remaining = n;
    bool ok = true;
    uint64_t hi = 0;  // limbs base - 2, base - 1
    uint64_t lo = 0;  // limbs base, base + 1
    long long base = -1;
    // One step of the decoder per received point; false: stop (invalid).
    auto step = [&](const int s) -> bool {
        // The received point is known, so instead of every point's start the
        // group needs three sums of the lanes' shares (all of them, those
        // below s, that of s) and the count of s; they are independent, so
        // their moves overlap.  The remainder of the split goes to the
        // largest count, lowest index among equals (`big`), as in the coder.
        const double inv = 1.0 / static_cast<double>(remaining);
        const int64_t sh = share_of_point(static_cast<uint64_t>(rng), static_cast<uint32_t>(cnt),
                                          static_cast<uint32_t>(remaining), inv);
        const int64_t total = group_sum<G>(sh);
        const int64_t below = group_sum<G>(j < s ? sh : 0);
        const int64_t mine = group_sum<G>(j == s ? sh : 0);
        const int cs = group_max<G>(j == s ? cnt : 0);
        if (!(s >= 0 && s < M) || cs == 0) {  // out of range, or no such point left
            ok = false;
            return false;
        }
        const int cmax = group_max<G>(cnt);
        const int big = group_min<G>(cnt == cmax ? j : G);
        const int64_t rest = rng - total;
        const int64_t cum = below + (big < s ? rest : 0);
        const int64_t sh_s = mine + (big == s ? rest : 0);
        const long long i = (shift - 1) >> 5;
        while (base < i - 1) {
            if (j == 0) emit(low, R, limbs, base - 2, hi >> 32);
            hi = (hi << 32) | (lo >> 32);
            lo <<= 32;
            ++base;
        }
        const uint64_t v = static_cast<uint64_t>(cum) << (31 - ((shift - 1) & 31));
        lo += v;
        if (lo < v) {
            hi += 1;
            if (hi == 0 && j == 0) {
                for (long long l = base - 3; l >= 0; --l) {
                    const uint64_t x = static_cast<uint64_t>(low[l * R]) + 1u;
                    low[l * R] = static_cast<unsigned>(x);
                    if ((x >> 32) == 0) break;
                }
            }
        }
        rng = sh_s;
        cnt -= (j == s) ? 1 : 0;
        --remaining;
        {  // renormalize to [2^31, 2^32) in one shift (as the bounded doubling loop)
            const int z = rng > 0 ? __builtin_clzll(static_cast<uint64_t>(rng)) - 32 : 32;
            const int d = z < 0 ? 0 : (z > 32 ? 32 : z);
            rng <<= d;
            shift += d;
        }
        if (shift > buffer_bits) {
            ok = false;
            return false;
        }
        return true;
    };
    // The points are loaded kAhead steps before their step, into a ring of
    // registers (static indices: the ring is walked by an unrolled loop), so
    // a step never waits for memory.
    constexpr int kAhead = 8;
    int ring[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) ring[u] = u < n ? sym[u] : 0;
    bool go = true;
    for (long long t0 = 0; go && t0 < n; t0 += kAhead) {
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const long long t = t0 + u;
            if (go && t < n) {
                const int s = ring[u];
                ring[u] = t + kAhead < n ? sym[t + kAhead] : 0;
                go = step(s);
            }
        }
    }
    if (j == 0) {
        emit(low, R, limbs, base - 2, hi >> 32);
        emit(low, R, limbs, base - 1, hi);
        emit(low, R, limbs, base, lo >> 32);
        emit(low, R, limbs, base + 1, lo);
    }
    // The limbs are written by lane 0 and read by the whole group below.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // v = ceil(low 2^k).  beyond: any bit after the k-th; adding one flips
    // the trailing ones of the first k bits and the zero before them (at
    // position `last0`, 1-based; 0 when all k are ones: v = 2^k).
    const long long first = k >> 5;
    const bool beyond = (low[first * R] & (0xFFFFFFFFu >> (k & 31))) != 0u ||
                        [&] {
                            for (long long l = first + 1; l < limbs; ++l)
                                if (low[l * R] != 0u) return true;
                            return false;
                        }();
    long long last0 = 0;
    if (beyond) {
        for (long long l = (k - 1) >> 5; l >= 0; --l) {
            unsigned w = low[l * R];
            const long long top = l * 32 + 32;  // bit positions l*32+1 .. top
            if (top > k) w |= 0xFFFFFFFFu >> (k - l * 32);  // ignore bits past k
            if (w != 0xFFFFFFFFu) {
                last0 = l * 32 + 32 - __builtin_ctz(~w);
                break;
            }
        }
    }
    signed char* out = bits + b * k;
    for (long long p = j + 1; p <= k; p += G) {
        const int bit = limb_bit(low, R, p);
        out[p - 1] = static_cast<signed char>(bit ^ ((beyond && p >= last0) ? 1 : 0));
    }
    if (ok && beyond && last0 == 0) ok = false;  // v = 2^k: no payload
    if (ok && beyond) {
        const long long lo_bit = (k + 1 > shift - 31) ? k + 1 : shift - 31;
        for (long long p = k + 1; p < lo_bit && ok; ++p)
            if (!limb_bit(low, R, p)) ok = false;
        if (ok) {
            int64_t f = 0;
            for (long long p = lo_bit; p <= shift; ++p) f = (f << 1) | limb_bit(low, R, p);
            ok = (int64_t{1} << (shift - lo_bit + 1)) - f < rng;
        }
    }
    if (j == 0) valid[b] = ok ? 1 : 0;
}

int32_t check_composition(const int64_t* composition, int32_t num_points, int64_t n,
                          Composition* comp) {
    CK_REQUIRE(composition != nullptr, "composition is NULL");
    CK_REQUIRE(num_points >= 1 && num_points <= kMaxPoints, "%d points, 1 to %d supported",
               num_points, kMaxPoints);
    CK_REQUIRE(n >= 1 && n <= (int64_t{1} << 30), "block of %lld points not in [1, 2^30]",
               static_cast<long long>(n));
    int64_t total = 0;
    for (int i = 0; i < kMaxPoints; ++i) comp->count[i] = 0;
    for (int i = 0; i < num_points; ++i) {
        CK_REQUIRE(composition[i] >= 0, "composition[%d] = %lld is negative", i,
                   static_cast<long long>(composition[i]));
        comp->count[i] = composition[i];
        total += composition[i];
    }
    CK_REQUIRE(total == n, "the composition sums to %lld, not the block length %lld",
               static_cast<long long>(total), static_cast<long long>(n));
    return CK_OK;
}

unsigned blocks_of(int64_t rows) { return static_cast<unsigned>((rows + kLanes - 1) / kLanes); }

// Lanes per block of the group kernels (a power of two, at least the
// points), or 0 for the one-lane kernel (more than 16 points).
int group_of(int32_t num_points) {
    int g = 2;
    while (g < num_points) g *= 2;
    return g <= 16 ? g : 0;
}

unsigned groups_of(int64_t rows, int G) {
    return static_cast<unsigned>((rows * G + kLanes - 1) / kLanes);
}

}  // namespace

extern "C" int32_t ck_ccdm_match(const int8_t* bits, int64_t rows, int64_t k,
                                 const int64_t* composition, int32_t num_points, int64_t n,
                                 int32_t* out, void* stream) {
    Composition comp;
    if (int32_t st = check_composition(composition, num_points, n, &comp)) return st;
    CK_REQUIRE(rows >= 0 && k >= 1, "rows %lld, k %lld", static_cast<long long>(rows),
               static_cast<long long>(k));
    if (rows == 0) return CK_OK;
    CK_REQUIRE(rows <= (int64_t{1} << 31), "%lld rows, at most 2^31", static_cast<long long>(rows));
    CK_REQUIRE(bits != nullptr && out != nullptr, "NULL bits or output");
    const dim3 g(blocks_of(rows)), blk(kLanes);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const signed char* u = reinterpret_cast<const signed char*>(bits);
    const long long R = rows, kk = k, nn = n;
    const int G = group_of(num_points);
    const dim3 gg(groups_of(rows, G));
    if (G == 2)
        hipLaunchKernelGGL(ccdm_match_group<2>, gg, blk, 0, st, u, R, kk, comp, num_points, nn, out);
    else if (G == 4)
        hipLaunchKernelGGL(ccdm_match_group<4>, gg, blk, 0, st, u, R, kk, comp, num_points, nn, out);
    else if (G == 8)
        hipLaunchKernelGGL(ccdm_match_group<8>, gg, blk, 0, st, u, R, kk, comp, num_points, nn, out);
    else if (G == 16)
        hipLaunchKernelGGL(ccdm_match_group<16>, gg, blk, 0, st, u, R, kk, comp, num_points, nn, out);
    else
        hipLaunchKernelGGL(ccdm_match_kernel<kMaxPoints>, g, blk, 0, st, u, R, kk, comp, num_points, nn,
                           out);
    return ck::launched("ccdm match");
}

extern "C" int64_t ck_ccdm_dematch_workspace_bytes(int64_t rows, int64_t buffer_bits) {
    if (rows < 0 || buffer_bits < 1) return -1;
    return rows * (buffer_bits / 32 + 2) * static_cast<int64_t>(sizeof(unsigned));
}

extern "C" int32_t ck_ccdm_dematch(const int32_t* symbols, int64_t rows, int64_t n,
                                   const int64_t* composition, int32_t num_points, int64_t k,
                                   int64_t buffer_bits, int8_t* bits, uint8_t* valid,
                                   void* workspace, void* stream) {
    Composition comp;
    if (int32_t st = check_composition(composition, num_points, n, &comp)) return st;
    CK_REQUIRE(rows >= 0 && k >= 1, "rows %lld, k %lld", static_cast<long long>(rows),
               static_cast<long long>(k));
    CK_REQUIRE(buffer_bits > k, "buffer of %lld bits does not exceed k = %lld",
               static_cast<long long>(buffer_bits), static_cast<long long>(k));
    if (rows == 0) return CK_OK;
    CK_REQUIRE(rows <= (int64_t{1} << 31), "%lld rows, at most 2^31", static_cast<long long>(rows));
    CK_REQUIRE(symbols != nullptr && bits != nullptr && valid != nullptr && workspace != nullptr,
               "NULL symbols, bits, valid or workspace");
    const dim3 g(blocks_of(rows)), blk(kLanes);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    signed char* out = reinterpret_cast<signed char*>(bits);
    unsigned* ws = static_cast<unsigned*>(workspace);
    const long long R = rows, nn = n, kk = k, bb = buffer_bits;
    const int G = group_of(num_points);
    const dim3 gg(groups_of(rows, G));
    if (G == 2)
        hipLaunchKernelGGL(ccdm_dematch_group<2>, gg, blk, 0, st, symbols, R, nn, comp, num_points,
                           kk, bb, out, valid, ws);
    else if (G == 4)
        hipLaunchKernelGGL(ccdm_dematch_group<4>, gg, blk, 0, st, symbols, R, nn, comp, num_points,
                           kk, bb, out, valid, ws);
    else if (G == 8)
        hipLaunchKernelGGL(ccdm_dematch_group<8>, gg, blk, 0, st, symbols, R, nn, comp, num_points,
                           kk, bb, out, valid, ws);
    else if (G == 16)
        hipLaunchKernelGGL(ccdm_dematch_group<16>, gg, blk, 0, st, symbols, R, nn, comp, num_points,
                           kk, bb, out, valid, ws);
    else
        hipLaunchKernelGGL(ccdm_dematch_kernel<kMaxPoints>, g, blk, 0, st, symbols, R, nn, comp,
                           num_points, kk, bb, out, valid, ws);
    return ck::launched("ccdm dematch");
}
