// Device building blocks of the LDPC kernels, shared by the per-call
// decoder (ck_ldpc.hip) and the persistent batch decoder (ck_batch.hip):
// constants, the compressed min-sum check state, frame load/store, the QC
// lane mapping, the per-frame syndrome check and the two row updates.
// Arithmetic contract: see ck_ldpc.hip.
#pragma once

#include "ck_common.h"

namespace ckdev {

constexpr int kGenericThreads = 256;
constexpr int kMaxQcThreads = 1024;
constexpr int kMaxFrames = 64;   // frames per QC workgroup
constexpr int kMaxDegree = 32;   // check degree of the compressed min-sum state
constexpr int kMaxSlots = 480;   // nonzero circulants of a QC code the QC kernels take
constexpr int kTabEntries = kMaxSlots + kMaxDegree;  // LDS slot table incl. padding
constexpr size_t kStaticLds = 8 * kTabEntries + 4 * kMaxFrames;  // static LDS of the QC kernels
constexpr float kPhiMin = 1e-7f;
constexpr float kPhiMax = 30.0f;

struct Params {
    float scale;
    float offset;
    int max_iterations;
    int early_stop;
};

// Compressed min-sum state of one check row: R_e = +-(e == idx ? m2 : m1),
// negative where bit e of `signs` is set.  All zeros is R = +0 everywhere,
// the state before the first iteration.
struct alignas(16) CheckState {
    float m1;
    float m2;
    unsigned signs;
    int idx;
};

__device__ __forceinline__ float phi(float x) {
    x = fminf(fmaxf(x, kPhiMin), kPhiMax);
    return -logf(tanhf(0.5f * x));
}

// Offset (floats) of the LDS check states after F APPs of N floats, 16-byte aligned.
__host__ __device__ __forceinline__ size_t state_offset(int F, int N) {
    return (static_cast<size_t>(F) * N + 3) & ~static_cast<size_t>(3);
}

__device__ __forceinline__ float state_msg(const CheckState& st, int e) {
    const float mag = (e == st.idx) ? st.m2 : st.m1;  // m1, m2 >= +0
    return __uint_as_float(__float_as_uint(mag) | ((st.signs << (31 - e)) & 0x80000000u));
}

// Channel LLRs of frames [frame0, frame0 + F) into LDS (zeros for the
// punctured tail and for frames past R).
__device__ inline void load_frames(float* lds, const float* __restrict__ llr, long long frame0, int F,
                            long long R, int n, int N) {
    const bool vec = ((n & 3) == 0) && ((N & 3) == 0) &&
                     ((reinterpret_cast<uintptr_t>(llr) & 15) == 0);
    for (int f = 0; f < F; ++f) {
        const long long frame = frame0 + f;
        float* dst = lds + static_cast<size_t>(f) * N;
        if (frame >= R) {
            for (int v = threadIdx.x; v < N; v += blockDim.x) dst[v] = 0.0f;
            continue;
        }
        const float* src = llr + frame * n;
        if (vec) {
            const float4* s4 = reinterpret_cast<const float4*>(src);
            float4* d4 = reinterpret_cast<float4*>(dst);
            for (int q = threadIdx.x; q < N / 4; q += blockDim.x)
                d4[q] = (4 * q < n) ? s4[q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            for (int v = threadIdx.x; v < N; v += blockDim.x) dst[v] = v < n ? src[v] : 0.0f;
        }
    }
}

// APP, hard info bits, iterations and valid of one frame, written by the
// `lanes` lanes numbered `u` = 0..lanes-1 of that frame.
__device__ inline void store_frame(const ck_ldpc_code& c, const float* a, long long frame, int u,
                            int lanes, int it, bool ok, float* __restrict__ app_out,
                            signed char* __restrict__ bits_out, int* __restrict__ iters_out,
                            unsigned char* __restrict__ valid_out) {
    float* dst = app_out + frame * c.n;
    for (int v = u; v < c.n; v += lanes) dst[v] = a[v];
    if (bits_out != nullptr) {
        signed char* b = bits_out + frame * c.k;
        for (int i = u; i < c.k; i += lanes) {
            const int pos = c.info ? c.info[i] : i;
            b[i] = a[pos] < 0.0f ? 1 : 0;
        }
    }
    if (u == 0) {
        iters_out[frame] = it;
        valid_out[frame] = ok ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Quasi-cyclic kernels
// ---------------------------------------------------------------------------

struct QcLane {
    int f;        // frame within the workgroup
    int r;        // row within the block row
    int p;        // lane within the row's P lanes
    int u;        // lane index within the frame, 0 .. Z*P-1
    long long frame;
    bool live;
};

template <int P>
__device__ __forceinline__ QcLane qc_lane(int Z, int F, long long R) {
    QcLane l;
    const int row_lane = threadIdx.x / P;
    l.p = threadIdx.x - row_lane * P;
    l.f = row_lane / Z;
    l.r = row_lane - l.f * Z;
    l.u = l.r * P + l.p;
    l.frame = static_cast<long long>(blockIdx.x) * F + l.f;
    l.live = (l.f < F) && (l.frame < R);
    return l;
}

// Uniform table loads: through the constant address space, so a
// wave-uniform index becomes a scalar (SMEM) load instead of a vector load
// whose latency every edge of the row would wait on.
#define CK_CONSTANT __attribute__((address_space(4)))

template <class T>
__device__ __forceinline__ T ld_uniform(const T* p, int i) {
    return ((const CK_CONSTANT T*)p)[i];
}

// Stages the slot table in LDS, where every row update reads it.  The code
// holds {block column * Z, shift} per nonzero circulant (S + 32 entries, zero
// padded); lane r of a block row reaches variable jZ + (r + shift) mod Z of
// slot {jZ, shift}, which in bytes from &a[r] is
//   4 (jZ + shift) - (r >= Z - shift ? 4Z : 0),
// so the LDS copy holds {4 (jZ + shift), Z - shift}.  Padding {0, 0} becomes
// {0, Z}: offset 0, the lane's own row, inside the frame.
__device__ inline void stage_slot_table(const ck_ldpc_code& c, int2* tab) {
    const int n = c.num_slots + kMaxDegree;
    const int Z = c.lifting;
    const int2* src = reinterpret_cast<const int2*>(c.qc_slot_packed);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int2 v = src[i];
        tab[i] = make_int2(4 * (v.x + v.y), Z - v.y);
    }
}

// Byte offset from &a[r] of lane r's variable in the staged slot `v`.
__device__ __forceinline__ int qc_off(int2 v, int r, int z4) {
    return v.x + ((r >= v.y) ? -z4 : 0);
}

__device__ __forceinline__ float& lds_at(float* ar, int off) {
    return *reinterpret_cast<float*>(reinterpret_cast<char*>(ar) + off);
}

// `mag` (sign bit clear) with the sign of bit e of `bits`.
__device__ __forceinline__ float with_sign(float mag, unsigned bits, int e) {
    return __uint_as_float(__float_as_uint(mag) | ((bits << (31 - e)) & 0x80000000u));
}

// Whether every check of the lane's frame is satisfied; `test` says whether
// this lane's frame is being checked.  Called by every lane (barriers).
template <int P>
__device__ bool qc_syndrome_ok(const ck_ldpc_code& c, const int2* tab, const float* a,
                               const QcLane& l, bool test, int F, int* s_flag) {
    const int z4 = 4 * c.lifting;
    float* ar = const_cast<float*>(a) + l.r;
    int bad = 0;
    if (test) {
        // Lane p of a row checks the rows of block rows p, p+P, ...
        for (int br0 = 0; br0 < c.base_rows; br0 += P) {
            const int br = min(br0 + l.p, c.base_rows - 1);
            const int s0 = ld_uniform(c.qc_row_ptr, br);
            const int s1 = ld_uniform(c.qc_row_ptr, br + 1);
            unsigned parity = 0u;
            for (int s = s0; s < s1; s += 8) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int2 e = tab[s + u];  // padded: s + u < S + 32
                    v[u] = lds_at(ar, qc_off(e, l.r, z4));
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) parity ^= (s + u < s1 && v[u] < 0.0f) ? 1u : 0u;
            }
            bad |= (br0 + l.p < c.base_rows) ? static_cast<int>(parity) : 0;
        }
    }
    if (F == 1) return !__syncthreads_or(bad);
    if (threadIdx.x < F) s_flag[threadIdx.x] = 0;
    __syncthreads();
    if (bad) s_flag[l.f] = 1;
    __syncthreads();
    const bool ok = (l.f < F) ? (s_flag[l.f] == 0) : true;
    __syncthreads();
    return ok;
}

__device__ __forceinline__ CheckState zero_state() {
    CheckState z;
    z.m1 = 0.0f;
    z.m2 = 0.0f;
    z.signs = 0u;
    z.idx = 0;
    return z;
}

// One min-sum check row of a block row over NT slots of the lane (edges
// e = t*P + p), given its old state; returns the new state.  Branch-free:
// every slot is processed, and only the last four can lie past the row's end
// (their table entries are real or zero padding, so their addresses stay in
// the frame).  UPDATE_APP: layered (APP = Q + R_new); else the flooding check
// phase, which leaves the APP alone.  Q and the LDS address of every edge
// stay in registers between the two passes.
template <int P, int NT, bool UPDATE_APP>
__device__ __forceinline__ CheckState qc_ms_row_n(const int2* tab, float* a, const CheckState st,
                                                  int s0, int dc, int r, int p, int Z,
                                                  const Params& prm) {
    float q[NT];
    float* addr[NT];
    float* ar = a + r;
    const int z4 = 4 * Z;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        addr[t] = &lds_at(ar, qc_off(tab[s0 + t * P + p], r, z4));
        q[t] = *addr[t];
    }
    // Pass 1.  min1 / min2 / idx follow the sequential rule exactly:
    //   idx moves on a strictly smaller |Q| (the first edge attaining min1),
    //   min2 = min(min2, max(min1, |Q|)) = med3(min1, min2, |Q|) as min1 <= min2,
    //   min1 = min(min1, |Q|)
    // (min, max and med3 return one of their operands, so nothing is rounded).
    // Bit e of qneg is Q_e < 0; the sign parity is its popcount.
    float min1 = ck::kInf;
    float min2 = ck::kInf;
    int idx = 0x7fffffff;
    unsigned qneg = 0u;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int e = t * P + p;
        const bool valid = (t < NT - 4) || (e < dc);
        const float qv = q[t] - state_msg(st, e);
        q[t] = qv;
        const float mag = valid ? fabsf(qv) : ck::kInf;
        idx = (mag < min1) ? e : idx;
        min2 = __builtin_amdgcn_fmed3f(min1, min2, mag);
        min1 = fminf(min1, mag);
        qneg |= (valid && qv < 0.0f) ? (1u << e) : 0u;
    }
#pragma unroll
    for (int off = 1; off < P; off <<= 1) {
        const float o1 = __shfl_xor(min1, off);
        const float o2 = __shfl_xor(min2, off);
        const int oi = __shfl_xor(idx, off);
        qneg |= static_cast<unsigned>(__shfl_xor(static_cast<int>(qneg), off));
        if (o1 < min1 || (o1 == min1 && oi < idx)) {
            min2 = fminf(min1, o2);
            min1 = o1;
            idx = oi;
        } else {
            min2 = fminf(min2, o1);
        }
    }
    const float m1 = fmaxf(prm.scale * min1 - prm.offset, 0.0f);
    const float m2 = fmaxf(prm.scale * min2 - prm.offset, 0.0f);
    // Pass 2: R_e is negative where (total parity) ^ (Q_e < 0).
    const unsigned row_mask = (dc >= 32) ? 0xffffffffu : ((1u << dc) - 1u);
    const unsigned signs = ((__popc(qneg) & 1) ? ~qneg : qneg) & row_mask;
    if (UPDATE_APP) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int e = t * P + p;
            if ((t < NT - 4) || (e < dc)) {
                const float mag = (e == idx) ? m2 : m1;
                *addr[t] = q[t] + with_sign(mag, signs, e);
            }
        }
    }
    CheckState out;
    out.m1 = m1;
    out.m2 = m2;
    out.signs = signs;
    out.idx = idx;
    return out;
}

// Chooses the instantiation of qc_ms_row_n by the row's slot count, rounded
// up to a multiple of four (the count is uniform across the workgroup).
template <int P, bool UPDATE_APP, int G = 1>
__device__ __forceinline__ CheckState qc_ms_dispatch(int groups, const int2* tab, float* a,
                                                     const CheckState st, int s0, int dc, int r,
                                                     int p, int Z, const Params& prm) {
    if constexpr (4 * G * P >= kMaxDegree) {
        return qc_ms_row_n<P, 4 * G, UPDATE_APP>(tab, a, st, s0, dc, r, p, Z, prm);
    } else {
        if (groups <= G) return qc_ms_row_n<P, 4 * G, UPDATE_APP>(tab, a, st, s0, dc, r, p, Z, prm);
        return qc_ms_dispatch<P, UPDATE_APP, G + 1>(groups, tab, a, st, s0, dc, r, p, Z, prm);
    }
}

// One min-sum check row `r` of block row `br`.
template <int P, bool UPDATE_APP>
__device__ __forceinline__ CheckState qc_ms_row(const ck_ldpc_code& c, const int2* tab, float* a,
                                                const CheckState st, int br, int r, int p,
                                                const Params& prm) {
    const int s0 = ld_uniform(c.qc_row_ptr, br);
    const int dc = ld_uniform(c.qc_row_ptr, br + 1) - s0;
    const int groups = ((dc + P - 1) / P + 3) >> 2;
    return qc_ms_dispatch<P, UPDATE_APP>(groups, tab, a, st, s0, dc, r, p, c.lifting, prm);
}

// One sum-product check row; Rm is the frame's messages, [slot][Z].  The old
// messages of the row are loaded together up front; Q, phi(|Q|) and the LDS
// address of every edge stay in registers between the passes.
template <bool UPDATE_APP>
__device__ __forceinline__ void qc_sp_row(const ck_ldpc_code& c, const int2* tab, float* a,
                                          float* __restrict__ Rm, int br, int r, bool first) {
    constexpr int MAXT = kMaxDegree;
    const int Z = c.lifting;
    const int s0 = ld_uniform(c.qc_row_ptr, br);
    const int dc = ld_uniform(c.qc_row_ptr, br + 1) - s0;
    float rold[MAXT];
#pragma unroll
    for (int t = 0; t < MAXT; ++t)
        if (t < dc) rold[t] = first ? 0.0f : Rm[static_cast<size_t>(s0 + t) * Z + r];
    float q[MAXT];
    float ph[MAXT];
    int addr[MAXT];
    float total = 0.0f;
    unsigned neg = 0u;
    float* ar = a + r;
    const int z4 = 4 * Z;
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
        if (t < dc) {
            addr[t] = qc_off(tab[s0 + t], r, z4);
            q[t] = lds_at(ar, addr[t]) - rold[t];
            neg ^= (q[t] < 0.0f) ? 1u : 0u;
            ph[t] = phi(fabsf(q[t]));
            total = total + ph[t];
        }
    }
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
        if (t < dc) {
            const float mag = phi(total - ph[t]);
            const float rn = (neg ^ ((q[t] < 0.0f) ? 1u : 0u)) ? -mag : mag;
            if (UPDATE_APP) lds_at(ar, addr[t]) = q[t] + rn;
            Rm[static_cast<size_t>(s0 + t) * Z + r] = rn;
        }
    }
}


}  // namespace ckdev
