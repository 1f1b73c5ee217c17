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
// A block is a sequential recursion, so one lane decodes one block and the
// launch is parallel over blocks.  The composition (at most kMaxPoints
// points) comes in the kernel arguments; each lane keeps its counts and the
// current split in LDS, point-major so that a wave's lanes touch
// consecutive words.  floor(a / remaining) with a < 2^62 is a float64
// estimate from one reciprocal per point position, corrected exactly with
// 64-bit integer arithmetic.  The dematcher's interval start lives in a
// caller workspace of buffer_bits / 32 + 2 limbs per block, limb-major.
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

// The split of `rng` at this position into share[i * kLanes + lane].
__device__ __forceinline__ void split(const int* cnt, int64_t* share, int M, int lane,
                                      int64_t rng, int64_t remaining) {
    const double inv = 1.0 / static_cast<double>(remaining);
    int big = 0;
    int cbig = cnt[lane];
    int64_t rest = rng;
    for (int i = 0; i < M; ++i) {
        const int c = cnt[i * kLanes + lane];
        const int64_t s = div_floor(rng * c, remaining, inv);
        share[i * kLanes + lane] = s;
        rest -= s;
        if (c > cbig) {
            big = i;
            cbig = c;
        }
    }
    share[big * kLanes + lane] += rest;
}

__global__ __launch_bounds__(kLanes) void ccdm_match_kernel(const signed char* __restrict__ bits,
                                                            long long R, long long k, Composition comp,
                                                            int M, long long n, int* __restrict__ out) {
    __shared__ int cnt[kMaxPoints * kLanes];
    __shared__ int64_t share[kMaxPoints * kLanes];
    const int lane = static_cast<int>(threadIdx.x);
    const long long b = static_cast<long long>(blockIdx.x) * kLanes + lane;
    if (b >= R) return;  // no barriers below: each lane works alone
    for (int i = 0; i < M; ++i) cnt[i * kLanes + lane] = static_cast<int>(comp.count[i]);
    const signed char* u = bits + b * k;
    int* o = out + b * n;
    long long pos = 0;
    int64_t d = 0;
    for (int j = 0; j < 32; ++j, ++pos) d = (d << 1) | (pos < k ? u[pos] : 0);
    int64_t rng = kTop;
    int64_t remaining = n;
    for (long long t = 0; t < n; ++t) {
        split(cnt, share, M, lane, rng, remaining);
        int64_t cum = 0;
        int64_t sh = 0;
        int s = 0;
        for (int i = 0; i < M; ++i) {
            sh = share[i * kLanes + lane];
            if (d < cum + sh) {
                s = i;
                break;
            }
            cum += sh;
        }
        o[t] = s;
        d -= cum;
        rng = sh;
        cnt[s * kLanes + lane] -= 1;
        --remaining;
        while (rng < kHalf) {
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

__global__ __launch_bounds__(kLanes) void ccdm_dematch_kernel(
    const int* __restrict__ symbols, long long R, long long n, Composition comp, int M,
    long long k, long long buffer_bits, signed char* __restrict__ bits,
    unsigned char* __restrict__ valid, unsigned* __restrict__ ws) {
    __shared__ int cnt[kMaxPoints * kLanes];
    __shared__ int64_t share[kMaxPoints * kLanes];
    const int lane = static_cast<int>(threadIdx.x);
    const long long b = static_cast<long long>(blockIdx.x) * kLanes + lane;
    if (b >= R) return;
    for (int i = 0; i < M; ++i) cnt[i * kLanes + lane] = static_cast<int>(comp.count[i]);
    const long long limbs = buffer_bits / 32 + 2;
    unsigned* low = ws + b;  // limb l at low[l * R]
    for (long long l = 0; l < limbs; ++l) low[l * R] = 0u;
    const int* sym = symbols + b * n;
    int64_t rng = kTop;
    long long shift = 32;
    int64_t remaining = n;
    bool ok = true;
    for (long long t = 0; t < n; ++t) {
        const int s = sym[t];
        if (s < 0 || s >= M || cnt[s * kLanes + lane] == 0) {
            ok = false;
            break;
        }
        split(cnt, share, M, lane, rng, remaining);
        int64_t cum = 0;
        for (int i = 0; i < s; ++i) cum += share[i * kLanes + lane];
        // low += cum 2^-shift, the carry running towards 2^-1.
        long long i = (shift - 1) >> 5;
        const uint64_t v = static_cast<uint64_t>(cum) << (31 - ((shift - 1) & 31));
        uint64_t acc = static_cast<uint64_t>(low[i * R]) + (v & 0xFFFFFFFFull);
        low[i * R] = static_cast<unsigned>(acc);
        uint64_t carry = (acc >> 32) + (v >> 32);
        while (carry != 0 && i > 0) {
            --i;
            acc = static_cast<uint64_t>(low[i * R]) + carry;
            low[i * R] = static_cast<unsigned>(acc);
            carry = acc >> 32;
        }
        rng = share[s * kLanes + lane];
        cnt[s * kLanes + lane] -= 1;
        --remaining;
        while (rng < kHalf) {
            rng <<= 1;
            ++shift;
        }
        if (shift > buffer_bits) {
            ok = false;
            break;
        }
    }
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
        const long long lo = (k + 1 > shift - 31) ? k + 1 : shift - 31;
        for (long long j = k + 1; j < lo && ok; ++j)
            if (!limb_bit(low, R, j)) ok = false;
        if (ok) {
            int64_t f = 0;
            for (long long j = lo; j <= shift; ++j) f = (f << 1) | limb_bit(low, R, j);
            ok = (int64_t{1} << (shift - lo + 1)) - f < rng;
        }
    }
    valid[b] = ok ? 1 : 0;
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

}  // namespace

extern "C" int32_t ck_ccdm_match(const int8_t* bits, int64_t rows, int64_t k,
                                 const int64_t* composition, int32_t num_points, int64_t n,
                                 int32_t* out, void* stream) {
    Composition comp;
    if (int32_t st = check_composition(composition, num_points, n, &comp)) return st;
    CK_REQUIRE(rows >= 0 && k >= 1, "rows %lld, k %lld", static_cast<long long>(rows),
               static_cast<long long>(k));
    if (rows == 0) return CK_OK;
    CK_REQUIRE(rows <= (int64_t{1} << 31) * kLanes, "too many rows");
    CK_REQUIRE(bits != nullptr && out != nullptr, "NULL bits or output");
    hipLaunchKernelGGL(ccdm_match_kernel, dim3(blocks_of(rows)), dim3(kLanes), 0,
                       static_cast<hipStream_t>(stream), reinterpret_cast<const signed char*>(bits),
                       static_cast<long long>(rows), static_cast<long long>(k), comp, num_points,
                       static_cast<long long>(n), out);
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
    CK_REQUIRE(rows <= (int64_t{1} << 31) * kLanes, "too many rows");
    CK_REQUIRE(symbols != nullptr && bits != nullptr && valid != nullptr && workspace != nullptr,
               "NULL symbols, bits, valid or workspace");
    hipLaunchKernelGGL(ccdm_dematch_kernel, dim3(blocks_of(rows)), dim3(kLanes), 0,
                       static_cast<hipStream_t>(stream), symbols, static_cast<long long>(rows),
                       static_cast<long long>(n), comp, num_points, static_cast<long long>(k),
                       static_cast<long long>(buffer_bits),
                       reinterpret_cast<signed char*>(bits), valid,
                       static_cast<unsigned*>(workspace));
    return ck::launched("ccdm dematch");
}
