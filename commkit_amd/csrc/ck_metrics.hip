// Post-FEC scoring on gfx950: bit and block errors (ber, fer) and the GMI
// loss of bit LLRs (gmi, ngmi).
//
// Both are streaming reductions over rows, bound by HBM: a workgroup of 256
// lanes takes a span of one row (for the errors, one block of `block` bits;
// for the GMI, kSpan elements), each lane walks its strided share with
// 16-byte loads where the row allows, the workgroup sums with wave
// reductions and one LDS step, and a single lane adds the span's total to
// the row's result with a 64-bit atomic.  The results are cleared by a small
// kernel first (no memset: the calls may be captured in a graph).
//
// gmi: the reference's per-element term in float32 (the LLRs' precision),
// log2(1 + e^x) = (log1p(e^-|x|) + max(0, x)) / ln 2 with x = -(1 - 2c) L,
// summed in float64 (commkit/metrics/information.py:34-89).
#include "ck_common.h"
#include "ck_tail.h"

#include <cstdint>

namespace {

constexpr int kThreads = 256;
constexpr long long kSpan = 1 << 16;  // GMI elements per workgroup

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Sum over the workgroup; valid in thread 0.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch) {
    v = wave_sum(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) scratch[wave] = v;
    __syncthreads();
    T total = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) total += scratch[w];
    return total;
}

__global__ void clear_i64(long long* a, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void clear_f64(double* a, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// One workgroup per (row, block): its mismatches, and whether it has any
// (block_err may be NULL; the last block of a row may be short).
__global__ __launch_bounds__(kThreads) void bit_errors_kernel(
    const signed char* __restrict__ bits, const signed char* __restrict__ ref, long long n,
    long long block, long long blocks_per_row, unsigned long long* __restrict__ bit_err,
    unsigned long long* __restrict__ block_err) {
    __shared__ long long scratch[kThreads / 64];
    const long long row = blockIdx.x / blocks_per_row;
    const long long blk = blockIdx.x - row * blocks_per_row;
    const long long base = row * n + blk * block;
    const long long len = n - blk * block < block ? n - blk * block : block;
    const signed char* a = bits + base;
    const signed char* b = ref + base;
    long long count = 0;
    const bool vec = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
    long long done = 0;
    if (vec) {
        const long long v16 = len / 16;
        const uint4* a4 = reinterpret_cast<const uint4*>(a);
        const uint4* b4 = reinterpret_cast<const uint4*>(b);
        for (long long i = threadIdx.x; i < v16; i += kThreads) {
            const uint4 x = a4[i];
            const uint4 y = b4[i];
            // Bytes are 0 or 1: the xor marks each differing byte's low bit.
            count += __popc(x.x ^ y.x) + __popc(x.y ^ y.y) + __popc(x.z ^ y.z) +
                     __popc(x.w ^ y.w);
        }
        done = v16 * 16;
    }
    for (long long i = done + threadIdx.x; i < len; i += kThreads) count += (a[i] != b[i]) ? 1 : 0;
    const long long total = block_sum(count, scratch);
    if (threadIdx.x == 0 && total) {
        atomicAdd(&bit_err[row], static_cast<unsigned long long>(total));
        if (block_err != nullptr) atomicAdd(&block_err[row], 1ull);
    }
}

__device__ __forceinline__ double gmi_term(float llr, signed char bit) {
    const float x = bit ? llr : -llr;  // -(1 - 2c) L
    // Divided by ln 2 rounded to float32, as the float32 arrays of the
    // reference are: zero LLRs then give exactly one bit.
    return static_cast<double>((log1pf(expf(-fabsf(x))) + fmaxf(0.0f, x)) / 0.6931471805599453f);
}

__global__ __launch_bounds__(kThreads) void gmi_kernel(const float* __restrict__ llrs,
                                                       const signed char* __restrict__ bits,
                                                       long long n, long long spans_per_row,
                                                       double* __restrict__ loss) {
    __shared__ double scratch[kThreads / 64];
    const long long row = blockIdx.x / spans_per_row;
    const long long span = blockIdx.x - row * spans_per_row;
    const long long lo = span * kSpan;
    const long long hi = lo + kSpan < n ? lo + kSpan : n;
    const float* l = llrs + row * n;
    const signed char* c = bits + row * n;
    double acc = 0.0;
    for (long long i = lo + threadIdx.x; i < hi; i += kThreads) acc += gmi_term(l[i], c[i]);
    const double total = block_sum(acc, scratch);
    if (threadIdx.x == 0) atomicAdd(&loss[row], total / static_cast<double>(n));
}

}  // namespace

extern "C" int32_t ck_bit_errors(const int8_t* bits, const int8_t* reference, int64_t rows,
                                 int64_t n, int64_t block, int64_t* bit_errors,
                                 int64_t* block_errors, void* stream) {
    CK_REQUIRE(rows >= 0 && n >= 0, "rows %lld, n %lld", static_cast<long long>(rows),
               static_cast<long long>(n));
    CK_REQUIRE(block >= 1 && n % block == 0, "block %lld does not divide the row length %lld",
               static_cast<long long>(block), static_cast<long long>(n));
    if (rows == 0) return CK_OK;
    CK_REQUIRE(bit_errors != nullptr, "NULL bit_errors");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const unsigned cg = static_cast<unsigned>((rows + 255) / 256);
    hipLaunchKernelGGL(clear_i64, dim3(cg), dim3(256), 0, st,
                       reinterpret_cast<long long*>(bit_errors), static_cast<long long>(rows));
    if (block_errors != nullptr)
        hipLaunchKernelGGL(clear_i64, dim3(cg), dim3(256), 0, st,
                           reinterpret_cast<long long*>(block_errors), static_cast<long long>(rows));
    if (n == 0) return ck::launched("bit errors");
    CK_REQUIRE(bits != nullptr && reference != nullptr, "NULL bits or reference");
    // Without block errors only the counts matter: long blocks are split
    // into spans of kSpan bits, so that a long row takes many workgroups.
    if (block_errors == nullptr && block > kSpan) block = kSpan;
    const long long per_row = (n + block - 1) / block;
    CK_REQUIRE(rows * per_row < (int64_t{1} << 31), "too many blocks");
    hipLaunchKernelGGL(bit_errors_kernel, dim3(static_cast<unsigned>(rows * per_row)),
                       dim3(kThreads), 0, st, reinterpret_cast<const signed char*>(bits),
                       reinterpret_cast<const signed char*>(reference), static_cast<long long>(n),
                       static_cast<long long>(block), per_row,
                       reinterpret_cast<unsigned long long*>(bit_errors),
                       reinterpret_cast<unsigned long long*>(block_errors));
    return ck::launched("bit errors");
}

extern "C" int32_t ck_gmi_loss(const float* llrs, const int8_t* bits, int64_t rows, int64_t n,
                               double* loss, void* stream) {
    CK_REQUIRE(rows >= 0 && n >= 1, "rows %lld, n %lld", static_cast<long long>(rows),
               static_cast<long long>(n));
    if (rows == 0) return CK_OK;
    CK_REQUIRE(llrs != nullptr && bits != nullptr && loss != nullptr, "NULL argument");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(clear_f64, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0, st,
                       loss, static_cast<long long>(rows));
    const long long spans = (n + kSpan - 1) / kSpan;
    CK_REQUIRE(rows * spans < (int64_t{1} << 31), "too many rows");
    hipLaunchKernelGGL(gmi_kernel, dim3(static_cast<unsigned>(rows * spans)), dim3(kThreads), 0, st,
                       llrs, reinterpret_cast<const signed char*>(bits), static_cast<long long>(n),
                       spans, loss);
    return ck::launched("gmi loss");
}
