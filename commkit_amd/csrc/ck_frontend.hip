// Device front end on gfx950: systematic LDPC encoding and additive white
// Gaussian noise, so that workloads are built where they are decoded.
//
// Encoding (ck_ldpc_encode): a workgroup of 256 lanes takes kFrames
// codewords.  For each, the information bits are staged in LDS (and copied
// to the codeword), and each wave forms 64 syndrome bits at a time, one
// check per lane (the XOR of the row's information bits, CSR from L2),
// packed by a wave ballot.  The parity is then Hp^-1 s over GF(2) for all
// the group's codewords at once: each lane takes a parity bit, walks the
// syndrome words with the matching column words of Hp^-1 (stored
// transposed, so the lanes' loads are contiguous), and keeps one 64-bit
// accumulator per codeword; the bit is the parity of its popcount.  Reading
// Hp^-1 once per group, not per codeword, divides its L2 traffic by
// kFrames.  A dual-diagonal Hp needs no table: the parity is the prefix XOR
// of the syndrome.
//
// Noise (ck_awgn): Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC 2011) keyed by the 64-bit seed;
// counter (group lo, group hi, stream lo, stream hi), one call per four
// normals.  Uniforms are the top 24 bits plus one half, times 2^-24, so in
// (0, 1); Box-Muller turns each pair into two normals.
#include "ck_common.h"
#include "ck_tail.h"

#include <cstdint>

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kFrames = 16;  // codewords per encoder workgroup

// ---------------------------------------------------------------- encoding

__global__ __launch_bounds__(kThreads) void encode_kernel(
    const signed char* __restrict__ info, long long rows, int k, int m, int n_out, int words,
    const int* __restrict__ row_ptr, const int* __restrict__ cols,
    const unsigned long long* __restrict__ inverse_t, signed char* __restrict__ out) {
    extern __shared__ unsigned long long lds[];
    unsigned long long* syn = lds;                                          // kFrames x words
    signed char* u = reinterpret_cast<signed char*>(lds + kFrames * words);  // k bytes
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const long long first = static_cast<long long>(blockIdx.x) * kFrames;
    const int frames = static_cast<int>(rows - first < kFrames ? rows - first : kFrames);

    for (int f = 0; f < frames; ++f) {
        const signed char* src = info + (first + f) * k;
        signed char* dst = out + (first + f) * n_out;
        for (int i = tid; i < k; i += kThreads) {
            const signed char b = src[i] & 1;
            u[i] = b;
            dst[i] = b;
        }
        __syncthreads();
        for (int w = wave; w < words; w += kWaves) {
            const int r = 64 * w + lane;
            int bit = 0;
            if (r < m)
                for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) bit ^= u[cols[e]];
            const unsigned long long word = __ballot(bit != 0);
            if (lane == 0) syn[f * words + w] = word;
        }
        __syncthreads();  // u is overwritten by the next codeword
    }
    const int parity = n_out - k;
    if (inverse_t == nullptr) {
        // Dual-diagonal Hp: p_j = s_0 ^ ... ^ s_j, a prefix XOR per codeword.
        if (tid < frames) {
            unsigned long long carry = 0;  // all ones if the bits so far have odd parity
            for (int w = 0; w < words; ++w) {
                unsigned long long x = syn[tid * words + w];
                x ^= x << 1;
                x ^= x << 2;
                x ^= x << 4;
                x ^= x << 8;
                x ^= x << 16;
                x ^= x << 32;
                x ^= carry;
                syn[tid * words + w] = x;
                carry = (x >> 63) ? ~0ull : 0ull;
            }
        }
        __syncthreads();
        for (int f = 0; f < frames; ++f) {
            signed char* dst = out + (first + f) * n_out + k;
            for (int j = tid; j < parity; j += kThreads)
                dst[j] = static_cast<signed char>((syn[f * words + (j >> 6)] >> (j & 63)) & 1);
        }
        return;
    }
    for (int j = tid; j < parity; j += kThreads) {
        unsigned long long acc[kFrames];
#pragma unroll
        for (int f = 0; f < kFrames; ++f) acc[f] = 0;
        for (int w = 0; w < words; ++w) {
            const unsigned long long h = inverse_t[static_cast<long long>(w) * m + j];
#pragma unroll
            for (int f = 0; f < kFrames; ++f) acc[f] ^= h & syn[f * words + w];
        }
#pragma unroll
        for (int f = 0; f < kFrames; ++f)
            if (f < frames)
                out[(first + f) * n_out + k + j] = static_cast<signed char>(__popcll(acc[f]) & 1);
    }
}

// ---------------------------------------------------------------- noise

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ float uniform_open(uint32_t x) {
    return (static_cast<float>(x >> 8) + 0.5f) * 5.9604644775390625e-08f;  // 2^-24
}

// Four standard normals of draw `group`.
__device__ __forceinline__ void normals4(unsigned long long group, uint32_t k0, uint32_t k1,
                                         uint32_t s0, uint32_t s1, float z[4]) {
    uint32_t c[4] = {static_cast<uint32_t>(group), static_cast<uint32_t>(group >> 32), s0, s1};
    philox4x32_10(c, k0, k1);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const float r = sqrtf(-2.0f * logf(uniform_open(c[2 * p])));
        float sn, cs;
        sincosf(6.283185307179586f * uniform_open(c[2 * p + 1]), &sn, &cs);
        z[2 * p] = r * cs;
        z[2 * p + 1] = r * sn;
    }
}

// One lane per draw of four normals: 4 real or 2 complex elements.  `out`
// may alias `x` (each element is read, then written, by the same lane).
__global__ __launch_bounds__(kThreads) void awgn_kernel(const float* x, long long rows,
                                                        long long len, int per_element,
                                                        const float* __restrict__ noise_std,
                                                        uint32_t k0, uint32_t k1, uint32_t s0,
                                                        uint32_t s1, float* out) {
    const long long total = rows * len;  // elements (complex pairs count once)
    const int per_draw = 4 / per_element;
    const long long groups = (total + per_draw - 1) / per_draw;
    for (long long g = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x; g < groups;
         g += static_cast<long long>(gridDim.x) * kThreads) {
        float z[4];
        normals4(static_cast<unsigned long long>(g), k0, k1, s0, s1, z);
        for (int i = 0; i < per_draw; ++i) {
            const long long e = g * per_draw + i;
            if (e >= total) break;
            const float sd = noise_std[e / len];
            if (per_element == 2) {
                out[2 * e] = x[2 * e] + sd * z[2 * i];
                out[2 * e + 1] = x[2 * e + 1] + sd * z[2 * i + 1];
            } else {
                out[e] = x[e] + sd * z[i];
            }
        }
    }
}

}  // namespace

extern "C" int32_t ck_ldpc_encode(const ck_ldpc_encoder* enc, const int8_t* info, int64_t rows,
                                  int8_t* words, void* stream) {
    CK_REQUIRE(enc != nullptr, "NULL encoder");
    CK_REQUIRE(rows >= 0, "rows %lld", static_cast<long long>(rows));
    CK_REQUIRE(enc->k >= 1 && enc->m >= 1 && enc->n_out >= enc->k && enc->n_out <= enc->k + enc->m,
               "encoder sizes k %d, m %d, n_out %d", enc->k, enc->m, enc->n_out);
    CK_REQUIRE(enc->words == (enc->m + 63) / 64, "words %d for m %d", enc->words, enc->m);
    if (rows == 0) return CK_OK;
    CK_REQUIRE(info != nullptr && words != nullptr && enc->info_row_ptr != nullptr &&
                   enc->info_cols != nullptr,
               "NULL argument");
    const size_t lds = static_cast<size_t>(kFrames) * enc->words * 8 + static_cast<size_t>(enc->k);
    ck::DeviceInfo dev;
    int32_t st = ck::device_info(&dev);
    if (st != CK_OK) return st;
    CK_REQUIRE(lds <= dev.lds_per_block, "encoder needs %zu B of LDS (k %d, m %d); %zu available",
               lds, enc->k, enc->m, dev.lds_per_block);
    CK_REQUIRE(rows / kFrames < (int64_t{1} << 31), "too many rows");
    if (lds > 65536)
        CK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&encode_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(lds)),
               "encoder LDS");
    const unsigned grid = static_cast<unsigned>((rows + kFrames - 1) / kFrames);
    hipLaunchKernelGGL(encode_kernel, dim3(grid), dim3(kThreads), lds,
                       static_cast<hipStream_t>(stream), reinterpret_cast<const signed char*>(info),
                       static_cast<long long>(rows), enc->k, enc->m, enc->n_out, enc->words,
                       enc->info_row_ptr, enc->info_cols,
                       reinterpret_cast<const unsigned long long*>(enc->inverse_t),
                       reinterpret_cast<signed char*>(words));
    return ck::launched("ldpc encode");
}

extern "C" int32_t ck_awgn(const void* x, int64_t rows, int64_t len, int32_t is_complex,
                           const float* noise_std, uint64_t seed, uint64_t stream_id, void* out,
                           void* stream) {
    CK_REQUIRE(rows >= 0 && len >= 0, "rows %lld, len %lld", static_cast<long long>(rows),
               static_cast<long long>(len));
    if (rows == 0 || len == 0) return CK_OK;
    CK_REQUIRE(x != nullptr && noise_std != nullptr && out != nullptr, "NULL argument");
    const int per_element = is_complex ? 2 : 1;
    const long long total = static_cast<long long>(rows) * len;
    const long long groups = (total * per_element + 3) / 4;
    ck::DeviceInfo dev;
    int32_t st = ck::device_info(&dev);
    if (st != CK_OK) return st;
    const long long want = (groups + kThreads - 1) / kThreads;
    const long long cap = static_cast<long long>(dev.num_cus) * 16;
    const unsigned grid = static_cast<unsigned>(want < cap ? want : cap);
    hipLaunchKernelGGL(awgn_kernel, dim3(grid), dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float*>(x), static_cast<long long>(rows),
                       static_cast<long long>(len), per_element, noise_std,
                       static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32),
                       static_cast<uint32_t>(stream_id), static_cast<uint32_t>(stream_id >> 32),
                       static_cast<float*>(out));
    return ck::launched("awgn");
}
