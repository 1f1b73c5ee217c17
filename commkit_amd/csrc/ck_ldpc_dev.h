// Device building blocks of the LDPC kernels, shared by the per-call
// decoder (ck_ldpc.hip) and the persistent batch decoder (ck_batch.hip):
// constants, the compressed min-sum check state, frame load/store, the QC
// lane mapping, the per-frame syndrome check and the two row updates.
// Arithmetic contract: see ck_ldpc.hip.
#pragma once

#include "ck_closed.h"
#include "ck_common.h"

namespace ckdev {

constexpr int kGenericThreads = 1024;
constexpr int kMaxQcThreads = 768;  // 12 waves: up to 168 VGPRs per lane
constexpr int kMaxFrames = 64;   // frames per QC workgroup
constexpr int kMaxDegree = 32;   // check degree of the compressed min-sum state
constexpr int kMaxSlots = 480;   // nonzero circulants of a QC code the QC kernels take
constexpr int kTabEntries = kMaxSlots + kMaxDegree;  // LDS slot table incl. padding
constexpr size_t kStaticLds = 8 * kTabEntries + 4 * kMaxFrames;  // static LDS of the QC kernels (slot table, flags)
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

// Sign bitmaps of the syndrome check: per frame and block column j, bit c
// is (APP of variable jZ + (c mod Z)) < 0 for c in [0, 64 ceil((Z + 32) / 64)),
// so the 32 rows from any row below Z read two adjacent words, unwrapped.
__host__ __device__ __forceinline__ int syn_words_per_col(int Z) { return 2 * ((Z + 95) / 64); }
__host__ __device__ __forceinline__ size_t syn_words(int Z, int N) {
    return static_cast<size_t>(N / Z) * syn_words_per_col(Z);
}

__host__ __device__ __forceinline__ size_t align4(size_t x) { return (x + 3) & ~static_cast<size_t>(3); }

// Dynamic LDS of the QC kernels, in floats: F APPs of N floats, then F frames
// of sign bitmaps, then (layered min-sum with the state in LDS) F frames of
// m check states; each part 16-byte aligned.
struct QcLds {
    size_t bits;
    size_t state;
    size_t end;
};
__host__ __device__ __forceinline__ QcLds qc_lds(int F, int N, int Z, int m, bool state) {
    QcLds o;
    o.bits = align4(static_cast<size_t>(F) * N);
    o.state = align4(o.bits + static_cast<size_t>(F) * syn_words(Z, N));
    o.end = state ? o.state + static_cast<size_t>(F) * m * 4 : o.state;
    return o;
}

// The flooding kernel's LDS past the QC layout when a frame's messages fit
// beside it (small codes), in floats: F frames of channel LLRs (N each),
// F frames of messages (min-sum: a 16-byte check state per row; sum-product:
// a float per edge), then the column tables (block column starts, then per
// slot in column order its slot, and per slot its shift, block row and
// position).  end == 0: not used.
struct FloodLds {
    size_t chan;
    size_t msgs;
    size_t tabs;
    size_t end;
};
__host__ __device__ __forceinline__ FloodLds flood_lds(int F, int N, int Z, int m, int E,
                                                      int nb, int slots, bool sp) {
    FloodLds o;
    o.chan = qc_lds(F, N, Z, m, false).end;
    o.msgs = align4(o.chan + static_cast<size_t>(F) * N);
    o.tabs = o.msgs + static_cast<size_t>(F) * (sp ? static_cast<size_t>(align4(E))
                                                 : static_cast<size_t>(m) * 4);
    o.end = o.tabs + static_cast<size_t>(nb + 1) + 4 * static_cast<size_t>(slots);
    return o;
}

__device__ __forceinline__ float state_msg(const CheckState& st, int e) {
    const float mag = (e == st.idx) ? st.m2 : st.m1;  // m1, m2 >= +0
    return __uint_as_float(__float_as_uint(mag) | ((st.signs << (31 - e)) & 0x80000000u));
}

// x + 0.0f per component: -0.0 becomes +0.0, every other value stays (the
// decoders read a Q's sign from its sign bit, and no Q is -0.0 once no
// channel LLR is; the reference's comparisons treat -0.0 and +0.0 alike).
__device__ __forceinline__ float4 plus_zero(float4 v) {
    return make_float4(v.x + 0.0f, v.y + 0.0f, v.z + 0.0f, v.w + 0.0f);
}

// Channel LLRs of frames [frame0, frame0 + F) into LDS (zeros for the
// punctured tail and for frames past R).  The 16-byte path walks the F
// frames as one range of F * N/4 vectors and issues up to kLoadBatch of a
// lane's loads before the first LDS write, so the workgroup waits for a
// couple of HBM round trips instead of one per vector.
constexpr int kLoadBatch = 8;

__device__ inline void load_frames(float* lds, const float* __restrict__ llr, long long frame0, int F,
                            long long R, int n, int N) {
    const bool vec = ((n & 3) == 0) && ((N & 3) == 0) &&
                     ((reinterpret_cast<uintptr_t>(llr) & 15) == 0);
    if (vec) {
        const int n4 = N >> 2;          // vectors per frame in LDS
        const int m4 = n >> 2;          // of them read from the frame
        const int total = F * n4;
        const int step = static_cast<int>(blockDim.x);
        float4* d4 = reinterpret_cast<float4*>(lds);
        // (f, j): frame and vector of this lane's next element q.
        int q = static_cast<int>(threadIdx.x);
        int f = q / n4;
        int j = q - f * n4;
        while (q < total) {
            float4 v[kLoadBatch];
            int at[kLoadBatch];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) {
                at[u] = q;
                const long long frame = frame0 + f;
                const bool load = q < total && j < m4 && frame < R;
                v[u] = load ? reinterpret_cast<const float4*>(llr + frame * n)[j]
                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                v[u] = plus_zero(v[u]);
                q += step;
                j += step;
                while (j >= n4) {
                    j -= n4;
                    ++f;
                }
            }
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (at[u] < total) d4[at[u]] = v[u];
        }
        return;
    }
    for (int f = 0; f < F; ++f) {
        const long long frame = frame0 + f;
        float* dst = lds + static_cast<size_t>(f) * N;
        if (frame >= R) {
            for (int v = threadIdx.x; v < N; v += blockDim.x) dst[v] = 0.0f;
            continue;
        }
        const float* src = llr + frame * n;
        for (int v = threadIdx.x; v < N; v += blockDim.x) dst[v] = v < n ? src[v] + 0.0f : 0.0f;
    }
}

// Channel LLRs of one frame (n of them at src, zeros up to N) into `a`,
// by the `lanes` lanes numbered u = 0..lanes-1 of that frame; 16-byte
// vectors when the row allows, and up to kLoadBatch loads of a lane in
// flight before its LDS writes.
__device__ inline void load_frame(float* a, const float* __restrict__ src, int n, int N, int u,
                                  int lanes) {
    if (((n & 3) == 0) && ((N & 3) == 0) && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
        const int n4 = N >> 2;
        const int m4 = n >> 2;
        const float4* s4 = reinterpret_cast<const float4*>(src);
        float4* d4 = reinterpret_cast<float4*>(a);
        for (int j0 = u; j0 < n4; j0 += kLoadBatch * lanes) {
            float4 v[kLoadBatch];
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                const int j = j0 + k * lanes;
                v[k] = plus_zero(j < m4 ? s4[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            }
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                const int j = j0 + k * lanes;
                if (j < n4) d4[j] = v[k];
            }
        }
        return;
    }
    for (int v0 = u; v0 < N; v0 += kLoadBatch * lanes) {
        float x[kLoadBatch];
#pragma unroll
        for (int k = 0; k < kLoadBatch; ++k) {
            const int v = v0 + k * lanes;
            x[k] = v < n ? src[v] + 0.0f : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kLoadBatch; ++k) {
            const int v = v0 + k * lanes;
            if (v < N) a[v] = x[k];
        }
    }
}

// Demapping fused into the frame load: a frame's channel LLRs computed from
// its received symbols straight into LDS by the closed-form max-log of a
// uniform separable lattice, with the demapper's cell tables and per-axis
// function (ck_closed.h), so the LLRs are bit for bit the demapper's.  One
// symbol holds k = 2 kb consecutive code bits; with this source the decoder
// never reads or writes LLRs in HBM.  The tables are staged in the dynamic
// LDS past the frames' APPs (the sign bitmaps and check states there are
// not used before the load ends).
constexpr int kDemapMaxBits = 4;  // per axis: up to 256-QAM

struct SymbolSource {
    const float2* x;         // (R, n / k) complex64 symbols; NULL: the frames are LLRs
    ck_axis re, im;          // closed-form tables (levels, labels, opposite levels, positions)
    double inv_s2;           // 1 / sigma^2 of every frame, unless inv_rows
    const double* inv_rows;  // (R,) 1 / sigma^2 per frame, or NULL
};

// LDS bytes of both axes' tables for kb bits per axis.
__host__ __device__ constexpr size_t demap_table_bytes(int kb) {
    return kb == 1   ? 2 * sizeof(ck::ClosedAxis<1>)
           : kb == 2 ? 2 * sizeof(ck::ClosedAxis<2>)
           : kb == 3 ? 2 * sizeof(ck::ClosedAxis<3>)
                     : 2 * sizeof(ck::ClosedAxis<4>);
}

// Frames [frame0, frame0 + F) demapped into LDS (zeros for the punctured
// tail and for frames past R); a lane's next kDemapBatch symbols are loaded
// before the first is demapped.  KB bits per axis; `tables` (16-byte
// aligned LDS) takes both axes' tables.
constexpr int kDemapBatch = 4;

template <int KB>
__device__ inline void load_frames_demap_kb(float* lds, const SymbolSource& src, void* tables,
                                            long long frame0, int F, long long R, int n, int N) {
    constexpr int k = 2 * KB;
    ck::ClosedAxis<KB>* tab = static_cast<ck::ClosedAxis<KB>*>(tables);
    ck::stage_closed<KB>(tab[0], src.re);
    ck::stage_closed<KB>(tab[1], src.im);
    __syncthreads();
    const int S = n / k;  // symbols per frame
    const int total = F * S;
    const int step = static_cast<int>(blockDim.x);
    for (int q0 = static_cast<int>(threadIdx.x); q0 < total; q0 += kDemapBatch * step) {
        float2 v[kDemapBatch];
        double inv[kDemapBatch];
#pragma unroll
        for (int u = 0; u < kDemapBatch; ++u) {
            const int q = q0 + u * step;
            const int f = q / S;
            const long long frame = frame0 + f;
            const bool live = q < total && frame < R;
            v[u] = live ? src.x[frame * S + (q - f * S)] : make_float2(0.0f, 0.0f);
            inv[u] = !live ? 1.0 : (src.inv_rows != nullptr ? src.inv_rows[frame] : src.inv_s2);
        }
#pragma unroll
        for (int u = 0; u < kDemapBatch; ++u) {
            const int q = q0 + u * step;
            if (q >= total) break;
            const int f = q / S;
            float* row = lds + static_cast<size_t>(f) * N + static_cast<size_t>(q - f * S) * k;
            if (frame0 + f < R) {
                ck::closed_axis_llrs<KB, true>(tab[0], v[u].x, inv[u], row);
                ck::closed_axis_llrs<KB, true>(tab[1], v[u].y, inv[u], row);
            } else {
#pragma unroll
                for (int b = 0; b < k; ++b) row[b] = 0.0f;
            }
        }
    }
    const int pad = N - n;
    for (int q = static_cast<int>(threadIdx.x); q < F * pad; q += step) {
        const int f = q / pad;
        lds[static_cast<size_t>(f) * N + n + (q - f * pad)] = 0.0f;
    }
}

__device__ inline void load_frames_demap(float* lds, const SymbolSource& src, void* tables,
                                         long long frame0, int F, long long R, int n, int N) {
    switch (src.re.num_bits) {  // a kernel argument: a scalar branch
        case 1: load_frames_demap_kb<1>(lds, src, tables, frame0, F, R, n, N); break;
        case 2: load_frames_demap_kb<2>(lds, src, tables, frame0, F, R, n, N); break;
        case 3: load_frames_demap_kb<3>(lds, src, tables, frame0, F, R, n, N); break;
        default: load_frames_demap_kb<4>(lds, src, tables, frame0, F, R, n, N); break;
    }
}

// APP, hard info bits, iterations and valid of one frame, written by the
// `lanes` lanes numbered `u` = 0..lanes-1 of that frame.
template <bool VEC = true>
__device__ inline void store_frame(const ck_ldpc_code& c, const float* a, long long frame, int u,
                            int lanes, int it, bool ok, float* __restrict__ app_out,
                            signed char* __restrict__ bits_out, int* __restrict__ iters_out,
                            unsigned char* __restrict__ valid_out) {
    // kLoadBatch LDS reads of a lane in flight before its stores.  VEC, with
    // n and k multiples of 4, 16-byte aligned rows and the information bits
    // first (no info table): the APP moves as 16-byte vectors and the hard
    // bits as 4-byte words of four sign bytes (the batched kernel, short of
    // registers, keeps the scalar form).
    float* dst = app_out + frame * c.n;
    const bool vec = VEC && ((c.n & 3) == 0) && ((c.k & 3) == 0) && c.info == nullptr &&
                     ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) &&
                     ((reinterpret_cast<uintptr_t>(a) & 15) == 0) &&
                     (bits_out == nullptr || ((reinterpret_cast<uintptr_t>(bits_out + frame * c.k) & 3) == 0));
    if (vec) {
        const float4* a4 = reinterpret_cast<const float4*>(a);
        float4* d4 = reinterpret_cast<float4*>(dst);
        const int n4 = c.n >> 2;
        for (int v0 = u; v0 < n4; v0 += kLoadBatch * lanes) {
            float4 x[kLoadBatch];
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                const int v = v0 + k * lanes;
                x[k] = v < n4 ? a4[v] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                const int v = v0 + k * lanes;
                if (v < n4) d4[v] = x[k];
            }
        }
        if (bits_out != nullptr) {
            unsigned* b4 = reinterpret_cast<unsigned*>(bits_out + frame * c.k);
            const int k4 = c.k >> 2;
            for (int i0 = u; i0 < k4; i0 += kLoadBatch * lanes) {
                float4 x[kLoadBatch];
#pragma unroll
                for (int k = 0; k < kLoadBatch; ++k) {
                    const int i = i0 + k * lanes;
                    x[k] = a4[i < k4 ? i : 0];
                }
#pragma unroll
                for (int k = 0; k < kLoadBatch; ++k) {
                    const int i = i0 + k * lanes;
                    // byte j = hard decision of bit 4i + j (APP < 0), little-endian
                    const unsigned w = (x[k].x < 0.0f ? 1u : 0u) | (x[k].y < 0.0f ? 0x100u : 0u) |
                                       (x[k].z < 0.0f ? 0x10000u : 0u) |
                                       (x[k].w < 0.0f ? 0x1000000u : 0u);
                    if (i < k4) b4[i] = w;
                }
            }
        }
        if (u == 0) {
            iters_out[frame] = it;
            valid_out[frame] = ok ? 1 : 0;
        }
        return;
    }
    for (int v0 = u; v0 < c.n; v0 += kLoadBatch * lanes) {
        float x[kLoadBatch];
#pragma unroll
        for (int k = 0; k < kLoadBatch; ++k) {
            const int v = v0 + k * lanes;
            x[k] = v < c.n ? a[v] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kLoadBatch; ++k) {
            const int v = v0 + k * lanes;
            if (v < c.n) dst[v] = x[k];
        }
    }
    if (bits_out != nullptr) {
        signed char* b = bits_out + frame * c.k;
        for (int i0 = u; i0 < c.k; i0 += kLoadBatch * lanes) {
            int pos[kLoadBatch];
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                const int i = i0 + k * lanes;
                pos[k] = i < c.k ? (c.info ? c.info[i] : i) : 0;
            }
            bool neg[kLoadBatch];
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) neg[k] = a[pos[k]] < 0.0f;
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                const int i = i0 + k * lanes;
                if (i < c.k) b[i] = neg[k] ? 1 : 0;
            }
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

// The lanes of one check row: P = 2 puts them in lanes i and i + 32 of a
// wave, so each half-wave reads 32 consecutive rows of a circulant (no LDS
// bank conflicts between the two halves' columns); P = 4 puts them in a quad.
// Workgroups are whole waves, so both halves exist.
template <int P>
__device__ __forceinline__ QcLane qc_lane(int Z, int F, long long R) {
    QcLane l;
    int row_lane;
    if constexpr (P == 2) {
        const int lane = static_cast<int>(threadIdx.x & 63);
        row_lane = static_cast<int>((threadIdx.x >> 6) << 5) | (lane & 31);
        l.p = lane >> 5;
    } else {
        row_lane = threadIdx.x / P;
        l.p = threadIdx.x - row_lane * P;
    }
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
//   4 (jZ + shift) - (r >= Z - shift ? 4Z : 0).
// The LDS copy holds x = 4 (jZ + shift) and y = (Z - shift) << 21 | jW, where
// jW = j * syn_words_per_col(Z) locates the column's sign bitmap.  With
// rr = r << 21 | (2^21 - 1), r >= Z - shift is rr >= y (unsigned).  Padding
// {0, 0} becomes x = 0, y = Z << 21: offset 0, the lane's own row.
__device__ inline void stage_slot_table(const ck_ldpc_code& c, int2* tab) {
    const int n = c.num_slots + kMaxDegree;
    const int Z = c.lifting;
    const int W = syn_words_per_col(Z);
    const int2* src = reinterpret_cast<const int2*>(c.qc_slot_packed);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int2 v = src[i];
        const unsigned y = (static_cast<unsigned>(Z - v.y) << 21) |
                           static_cast<unsigned>((v.x / Z) * W);
        tab[i] = make_int2(4 * (v.x + v.y), static_cast<int>(y));
    }
}

__device__ __forceinline__ unsigned row_rank(int r) {
    return (static_cast<unsigned>(r) << 21) | 0x1fffffu;
}

// -4Z, the byte offset back over one block column, as a value the compiler
// cannot fold into the offset arithmetic: the edge address then takes three
// VALU ops (compare, select, three-way add).
__device__ __forceinline__ int neg_col_bytes(int Z) { return __builtin_amdgcn_readfirstlane(-4 * Z); }

// Byte offset from &a[r] of lane r's variable in the staged slot `v`
// (rr = row_rank(r), nz4 = neg_col_bytes(Z)).
__device__ __forceinline__ int qc_off(int2 v, unsigned rr, int nz4) {
    return v.x + ((rr >= static_cast<unsigned>(v.y)) ? nz4 : 0);
}

__device__ __forceinline__ float& lds_at(float* ar, int off) {
    return *reinterpret_cast<float*>(reinterpret_cast<char*>(ar) + off);
}

// The value of lane (this lane ^ off), off = 1 or 2, within a quad: a DPP
// move (one VALU op) instead of an LDS permute.  The partner lane (a lane of
// the same check row) must be active.
__device__ __forceinline__ int lane_xor(int v, int off) {
    return off == 1 ? __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false)   // quad_perm 1,0,3,2
                    : __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
}

// The value of lane J of the quad in every lane of it (quad_perm J,J,J,J);
// the quad's lanes must be active, as for lane_xor.
template <int J>
__device__ __forceinline__ int quad_lane(int v) {
    return __builtin_amdgcn_mov_dpp(v, J * 0x55, 0xF, 0xF, false);
}

// Both values of a two-lane check row (lanes i and i + 32) in each of its
// lanes: .x from the lane with p = 0, .y from p = 1.  v_permlane32_swap
// exchanges the upper half of its first operand with the lower half of its
// second; with v in both, the first result holds the lower half's values in
// every lane and the second the upper half's.
__device__ __forceinline__ int2 row_pair(int v) {
    const auto s = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(v),
                                                    static_cast<unsigned>(v), false, false);
    return make_int2(static_cast<int>(s[0]), static_cast<int>(s[1]));
}

// `mag` (sign bit clear) with the sign of bit e of `bits`.
__device__ __forceinline__ float with_sign(float mag, unsigned bits, int e) {
    return __uint_as_float(__float_as_uint(mag) | ((bits << (31 - e)) & 0x80000000u));
}

// Block rows checked between barriers by the direct syndrome check.
constexpr int kSynGroup = 3;

// Parity (0 or 1) of the hard decisions of row r of block row br of the
// frame whose APP starts at `a`: the XOR of the APPs' sign bits over the
// row's circulants (no APP is -0.0, so a sign bit is APP < 0).  Table
// entries past the row's end are real slots or padding, in the frame.
__device__ __forceinline__ unsigned qc_row_parity(const ck_ldpc_code& c, const int2* tab,
                                                  const float* a, int br, int r) {
    constexpr int U = 8;
    const int s0 = ld_uniform(c.qc_row_ptr, br);
    const int dc = ld_uniform(c.qc_row_ptr, br + 1) - s0;
    float* ar = const_cast<float*>(a) + r;
    const unsigned rr = row_rank(r);
    const int z4 = neg_col_bytes(c.lifting);
    unsigned acc = 0u;
    for (int t = 0; t < dc; t += U) {
        unsigned v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __float_as_uint(lds_at(ar, qc_off(tab[s0 + t + u], rr, z4)));
#pragma unroll
        for (int u = 0; u < U; ++u) acc ^= (t + u < dc) ? v[u] : 0u;
    }
    return acc >> 31;
}

// Whether every check of the lane's frame slot is satisfied.  `bits` is the
// sign-bitmap area of frame slot 0 (slot f at bits + f * syn_words), `app`
// the APP of slot 0 (slot f at app + f * N); `test` (the same on every lane of
// a slot) says whether the slot is checked.  Called by every lane (barriers);
// lanes outside the F slots (lane_ok false) return true.
//
// A sample of rows is checked first (quick rejection).  Otherwise the waves
// write the bitmaps, one ballot of 64 signs per step; then
// each task (slot, block row, 32 rows from r0) XORs, over the block row's
// circulants {j, shift}, the 32 bits from column (r0 + shift) mod Z of
// bitmap j: bit i of the result is the parity of row r0 + i.
//
// prechecked: the caller has already set s_flag of every tested slot to 1
// (a row of the last block row of the iteration has odd hard-decision
// parity: rejected) or 0 (all its rows even).  The other block rows are then
// checked row by row straight from the APP, kSynGroup block rows between
// barriers, last block rows first, until every tested slot is rejected or
// all rows are even (no bitmaps: a frame that passes the last block row is
// near convergence, and a failing row is usually found in the first group).
__device__ __forceinline__ bool qc_syndrome_ok(const ck_ldpc_code& c, const int2* tab, const float* app,
                               unsigned* bits, const QcLane& l, bool lane_ok, bool test, int F,
                               int* s_flag, bool prechecked = false) {
    constexpr int U = 8;  // loads in flight per lane
    const int Z = c.lifting;
    const int N = c.num_variables;
    const int nb = N / Z;
    const int W = syn_words_per_col(Z);
    const int C = W >> 1;
    const int words = nb * W;
    const int K = (Z + 31) >> 5;
    const int per_frame = c.base_rows * K;
    const int checks = F * per_frame;
    if (prechecked) {
        if (lane_ok && l.u == 0 && !test) s_flag[l.f] = 2;
        const float* af = app + static_cast<size_t>(lane_ok ? l.f : 0) * N;
        for (int br = c.base_rows - 2; br >= 0; br -= kSynGroup) {
            if (!__syncthreads_or(lane_ok && l.u == 0 && s_flag[l.f] == 0)) break;
            if (lane_ok && l.p == 0 && s_flag[l.f] == 0) {
                unsigned odd = 0u;
                for (int b = br; b >= 0 && b > br - kSynGroup; --b) odd |= qc_row_parity(c, tab, af, b, l.r);
                if (odd) s_flag[l.f] = 1;
            }
        }
        __syncthreads();
        const bool ok = lane_ok ? (s_flag[l.f] == 0) : true;
        __syncthreads();
        return ok;
    }
    if (lane_ok && l.u == 0) s_flag[l.f] = test ? 0 : 2;
    __syncthreads();
    // Quick rejection: lanes (f, r < 64, p = 0) check row r of block row 0
    // straight from the APP; one unsatisfied row settles the frame, which
    // then skips the full check.  While a frame is far from converged this
    // almost always happens, and the bitmaps are not built at all.
    if (lane_ok && l.p == 0 && l.r < 64 && s_flag[l.f] == 0) {
        const int s0 = ld_uniform(c.qc_row_ptr, 0);
        const int s1 = ld_uniform(c.qc_row_ptr, 1);
        float* ar = const_cast<float*>(app) + static_cast<size_t>(l.f) * N + l.r;
        const unsigned rr = row_rank(l.r);
        const int z4 = neg_col_bytes(Z);
        unsigned parity = 0u;
        for (int s = s0; s < s1; s += U) {
            float v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = lds_at(ar, qc_off(tab[s + u], rr, z4));  // padded table
#pragma unroll
            for (int u = 0; u < U; ++u) parity ^= (s + u < s1 && v[u] < 0.0f) ? 1u : 0u;
        }
        if (parity) s_flag[l.f] = 1;
    }
    if (!__syncthreads_or(lane_ok && l.u == 0 && s_flag[l.f] == 0)) {
        const bool ok = lane_ok ? (s_flag[l.f] == 0) : true;
        __syncthreads();
        return ok;
    }
    // The first check task of this lane; its row range is fetched now so the
    // load completes under the bitmap writes.
    int task = threadIdx.x;
    int s0 = 0;
    int s1 = 0;
    if (task < checks) {
        const int br = (task % per_frame) / K;
        s0 = c.qc_row_ptr[br];
        s1 = c.qc_row_ptr[br + 1];
    }
    {
        // Chunk t = (f nb + j) C + ch (slot f, column j, bits 64 ch ..
        // 64 ch + 63) goes to words 2t, 2t + 1: lane i's bit is the sign of
        // variable (64 ch + i) mod Z of column f nb + j, whose APP starts at
        // app + (f nb + j) Z.  Each wave takes whole columns (one ballot per
        // chunk, kept by lane ch in a pair of registers), then lanes
        // 0..C-1 store the column's C chunks.  Every lane is active here
        // (the ballots count all 64).
        const int lane = threadIdx.x & 63;
        const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
        const int nw = static_cast<int>(blockDim.x >> 6);
        const int cols = F * nb;
        const int d = 64 % Z;  // column step from one chunk to the next
        const int lm = lane % Z;
        for (int fj = wave; fj < cols; fj += nw) {
            const float* src = app + static_cast<size_t>(fj) * Z;
            unsigned lo = 0u;
            unsigned hi = 0u;
            int col = lm;
            for (int ch0 = 0; ch0 < C; ch0 += 4) {
                float v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    v[u] = src[col];  // col < Z: in the column whatever ch0 + u
                    col += d;
                    col -= (col >= Z) ? Z : 0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (ch0 + u < C) {
                        const unsigned long long m = __ballot(v[u] < 0.0f);
                        const bool mine = lane == ch0 + u;
                        lo = mine ? static_cast<unsigned>(m) : lo;
                        hi = mine ? static_cast<unsigned>(m >> 32) : hi;
                    }
                }
            }
            if (lane < C)
                *reinterpret_cast<uint2*>(bits + 2 * (static_cast<size_t>(fj) * C + lane)) =
                    make_uint2(lo, hi);
        }
    }
    __syncthreads();
    for (; task < checks; task += blockDim.x) {
        const int f = task / per_frame;
        const int rem = task - f * per_frame;
        const int br = rem / K;
        const int r0 = (rem - br * K) << 5;
        if (task >= static_cast<int>(blockDim.x)) {
            s0 = c.qc_row_ptr[br];
            s1 = c.qc_row_ptr[br + 1];
        }
        if (s_flag[f] != 0) continue;  // not checked, or already failed
        const unsigned* bm = bits + static_cast<size_t>(f) * words;
        unsigned acc = 0u;
        for (int s = s0; s < s1; s += U) {
            // Entries past the row's end (other rows or padding) stay in bounds.
            unsigned lo[U];
            unsigned hi[U];
            int sh[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned y = static_cast<unsigned>(tab[s + u].y);
                const int d = r0 - static_cast<int>(y >> 21);  // r0 - (Z - shift)
                const int o = d + ((d < 0) ? Z : 0);            // (r0 + shift) mod Z
                const unsigned* w = bm + (y & 0x1fffffu) + (o >> 5);
                lo[u] = w[0];
                hi[u] = w[1];
                sh[u] = o & 31;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (s + u < s1) acc ^= __builtin_amdgcn_alignbit(hi[u], lo[u], static_cast<unsigned>(sh[u]));
        }
        if (Z - r0 < 32) acc &= (1u << (Z - r0)) - 1u;
        if (acc != 0u) s_flag[f] = 1;
    }
    __syncthreads();
    const bool ok = lane_ok ? (s_flag[l.f] == 0) : true;
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

// Slot granularity of the row instantiations: rows are padded to a multiple
// of kSlotStep slots per lane.
template <int P>
constexpr int kSlotStep = 2;

// One min-sum check row of a block row over NT slots of the lane (edges
// e = t*P + p), given its old state; returns the new state.  Branch-free:
// every slot is processed, and only the last kSlotStep can lie past the row's end
// (their table entries are real or zero padding, so their addresses stay in
// the frame).  UPDATE_APP: layered (APP = Q + R_new); else the flooding check
// phase, which leaves the APP alone.  Q and the LDS address of every edge
// stay in registers between the two passes.
//
// The rule is the sequential one of the reference (min1 = the smallest |Q|,
// min2 the second smallest, R_e = sign * (e == argmin ? m2 : m1) with
// sign = parity of the row's Q signs ^ (Q_e < 0)), evaluated so that an edge
// costs few instructions:
//   - pass 1 keeps only min1 and min2 (min and med3 return one of their
//     operands, so nothing is rounded); the argmin is found in pass 2 as an
//     edge whose |Q| equals min1.  With one smallest |Q| that edge is the
//     reference's argmin; with a tie min2 = min1, so m2 = m1 and every edge
//     gets the same message whichever tied edge is called the argmin;
//   - Q_e < 0 is Q_e's sign bit: no Q is -0.0, because the channel LLRs are
//     loaded with -0.0 turned into +0.0 (x + 0.0f) and a - b or a + b is
//     -0.0 only when a is -0.0 (round to nearest), so no APP and no Q ever
//     becomes one.  With one lane per row (P = 1) the sign bits are shifted
//     into a word one alignbit per edge (edge t lands at bit NT - 1 - t) and
//     bit-reversed into edge order once per row;
//   - R_e's sign bit is the row parity's flipped by Q_e's sign bit: the
//     magnitude with the parity's sign bit, XORed with Q_e & 0x80000000.
// PAR: *hard_parity = the parity of the row's hard decisions (the XOR of
// the APPs' sign bits): with UPDATE_APP on the updated APP, the row's
// syndrome bit once no later row of the iteration touches its bits; without
// (flooding) on the APP the row reads, the syndrome bit of the previous
// iteration's result.
template <int P, int NT, bool UPDATE_APP, bool PAR = false>
__device__ __forceinline__ CheckState qc_ms_row_n(const int2* tab, float* a, const CheckState st,
                                                  int s0, int dc, int r, int p, int Z,
                                                  const Params& prm,
                                                  unsigned* hard_parity = nullptr) {
    constexpr int S = kSlotStep<P>;
    float q[NT];
    float* addr[NT];
    float* ar = a + r;
    const int z4 = neg_col_bytes(Z);
    const unsigned rr = row_rank(r);
    unsigned hard = 0u;  // PAR: XOR of the read (flooding) or updated APPs' bit patterns
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        addr[t] = &lds_at(ar, qc_off(tab[s0 + t * P + p], rr, z4));
        q[t] = *addr[t];
        if (PAR && !UPDATE_APP) {
            const bool valid = (t < NT - S) || (t * P + p < dc);
            hard ^= valid ? __float_as_uint(q[t]) : 0u;
        }
    }
    // Pass 1: Q, min1 and min2 (min2 = med3(min1, min2, |Q|) as min1 <= min2),
    // the sign bits of Q.
    float min1 = ck::kInf;
    float min2 = ck::kInf;
    unsigned qneg = 0u;  // P > 1: bit e is Q_e < 0
    unsigned qrev = 0u;  // P = 1: sign bit of Q_t at bit NT - 1 - t
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int e = t * P + p;
        const bool valid = (t < NT - S) || (e < dc);
        const float qv = q[t] - state_msg(st, e);
        q[t] = qv;
        const float mag = valid ? fabsf(qv) : ck::kInf;
        min2 = __builtin_amdgcn_fmed3f(min1, min2, mag);
        min1 = fminf(min1, mag);
        if constexpr (P == 1)
            qrev = __builtin_amdgcn_alignbit(qrev, __float_as_uint(qv), 31u);
        else
            qneg |= (valid && qv < 0.0f) ? (1u << e) : 0u;
    }
    const unsigned row_mask = (dc >= 32) ? 0xffffffffu : ((1u << dc) - 1u);
    if constexpr (P == 1) {
        qneg = (__builtin_bitreverse32(qrev) >> (32 - NT)) & row_mask;
    } else if constexpr (P == 2) {
        // The row's two smallest |Q| from each lane's two smallest.
        const int2 v1 = row_pair(__float_as_int(min1));
        const int2 v2 = row_pair(__float_as_int(min2));
        const int2 vq = row_pair(static_cast<int>(qneg));
        const float x1 = __int_as_float(v1.x);
        const float y1 = __int_as_float(v1.y);
        min1 = fminf(x1, y1);
        min2 = fminf(fmaxf(x1, y1), fminf(__int_as_float(v2.x), __int_as_float(v2.y)));
        qneg = static_cast<unsigned>(vq.x | vq.y);
    } else {
#pragma unroll
        for (int off = 1; off < P; off <<= 1) {
            const float o1 = __int_as_float(lane_xor(__float_as_int(min1), off));
            const float o2 = __int_as_float(lane_xor(__float_as_int(min2), off));
            qneg |= static_cast<unsigned>(lane_xor(static_cast<int>(qneg), off));
            min2 = fminf(fmaxf(min1, o1), fminf(min2, o2));
            min1 = fminf(min1, o1);
        }
    }
    const float m1 = fmaxf(prm.scale * min1 - prm.offset, 0.0f);
    const float m2 = fmaxf(prm.scale * min2 - prm.offset, 0.0f);
    const unsigned parity = static_cast<unsigned>(__popc(qneg) & 1);
    const unsigned signs = (parity ? ~qneg : qneg) & row_mask;
    // Pass 2: the argmin (an edge at min1), and with UPDATE_APP the new APP.
    const unsigned pbit = parity << 31;
    float m1p = __uint_as_float(__float_as_uint(m1) | pbit);
    float m2p = __uint_as_float(__float_as_uint(m2) | pbit);
    // Opaque to the compiler, so that an edge's sign costs one bitop3
    // (mp ^ (Q & 0x80000000)) instead of the parity folded in per edge.
    asm volatile("" : "+v"(m1p), "+v"(m2p));
    int idx = -1;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int e = t * P + p;
        const bool valid = (t < NT - S) || (e < dc);
        const bool at_min = valid && fabsf(q[t]) == min1;
        idx = at_min ? e : idx;
        if (UPDATE_APP && valid) {
            const float mp = at_min ? m2p : m1p;
            const float rn = __uint_as_float(__float_as_uint(mp) ^
                                             (__float_as_uint(q[t]) & 0x80000000u));
            const float an = q[t] + rn;
            *addr[t] = an;
            if (PAR) hard ^= __float_as_uint(an);
        }
    }
    if constexpr (P == 2) {
        const int2 vi = row_pair(idx);
        idx = max(vi.x, vi.y);
        if (PAR) {
            const int2 vh = row_pair(static_cast<int>(hard));
            hard = static_cast<unsigned>(vh.x ^ vh.y);
        }
    } else {
#pragma unroll
        for (int off = 1; off < P; off <<= 1) {
            idx = max(idx, lane_xor(idx, off));
            if (PAR) hard ^= static_cast<unsigned>(lane_xor(static_cast<int>(hard), off));
        }
    }
    if (PAR) *hard_parity = hard >> 31;  // no APP is -0.0: sign bit = (APP < 0)
    CheckState out;
    out.m1 = m1;
    out.m2 = m2;
    out.signs = signs;
    out.idx = idx;
    return out;
}

// Chooses the instantiation of qc_ms_row_n by the row's slot count, rounded
// up to a multiple of kSlotStep (the count is uniform across the workgroup).
template <int P, bool UPDATE_APP, int MAXT, int G = 1, bool PAR = false>
__device__ __forceinline__ CheckState qc_ms_dispatch(int groups, const int2* tab, float* a,
                                                     const CheckState st, int s0, int dc, int r,
                                                     int p, int Z, const Params& prm,
                                                     unsigned* hard_parity = nullptr) {
    constexpr int S = kSlotStep<P>;
    if constexpr (S * G >= MAXT) {
        return qc_ms_row_n<P, S * G, UPDATE_APP, PAR>(tab, a, st, s0, dc, r, p, Z, prm,
                                                      hard_parity);
    } else {
        if (groups <= G)
            return qc_ms_row_n<P, S * G, UPDATE_APP, PAR>(tab, a, st, s0, dc, r, p, Z, prm,
                                                          hard_parity);
        return qc_ms_dispatch<P, UPDATE_APP, MAXT, G + 1, PAR>(groups, tab, a, st, s0, dc, r, p,
                                                               Z, prm, hard_parity);
    }
}

// Slots per lane of the kernels for short rows (P = 1; halved for P = 2):
// fewer registers than rows of up to kMaxDegree edges.  Up to kShortRow
// edges the kernel fits 168 VGPRs (one workgroup of twelve waves per CU);
// up to kShortRowTwo it fits 128 without spilling, so two workgroups of six
// waves share a CU.
constexpr int kShortRow = 24;
constexpr int kShortRowTwo = 22;

// One min-sum check row `r` of block row `br`; MAXT bounds the lane's slots
// (ceil(row degree / P)), which the host checks before choosing the kernel.
template <int P, bool UPDATE_APP, int MAXT = kMaxDegree / P, bool PAR = false>
__device__ __forceinline__ CheckState qc_ms_row(const ck_ldpc_code& c, const int2* tab, float* a,
                                                const CheckState st, int br, int r, int p,
                                                const Params& prm,
                                                unsigned* hard_parity = nullptr) {
    const int s0 = ld_uniform(c.qc_row_ptr, br);
    const int dc = ld_uniform(c.qc_row_ptr, br + 1) - s0;
    const int groups = ((dc + P - 1) / P + kSlotStep<P> - 1) / kSlotStep<P>;
    return qc_ms_dispatch<P, UPDATE_APP, MAXT, 1, PAR>(groups, tab, a, st, s0, dc, r, p,
                                                       c.lifting, prm, hard_parity);
}

// phi(x) = -ln(tanh(x / 2)) = ln((1 + e^-x) / (1 - e^-x)) for x clipped to
// [kPhiMin, kPhiMax], taken in log2 units: phi2(y) = phi(y ln 2) / ln 2 =
// log2((1 + 2^-y) / (1 - 2^-y)), from the hardware exp2, reciprocal and log2
// with no scaling in or out.  A row keeps its phi values and their total in
// these units: |Q| is scaled by log2(e) on the way in and the new message
// by ln 2 on the way out, two multiplies per edge instead of four.  Below
// x = 1/16, 1 - e^-x would lose digits to cancellation and comes from its
// Taylor series (relative error of the four terms below 2e-7 there; written
// with fused multiply-adds, which the file's no-contraction flag would not
// form by itself).  Within about 1e-6 relative of the float64 value where
// phi is above 1e-2; smaller values (x above ~5) carry an absolute error
// near 1e-7.
constexpr float kLog2eF = 1.44269504f;
constexpr float kLn2F = 0.69314718f;

__device__ __forceinline__ float phi2_fast(float y) {
    y = fminf(fmaxf(y, kPhiMin * kLog2eF), kPhiMax * kLog2eF);
    const float t = __builtin_amdgcn_exp2f(-y);  // e^-x, x = y ln 2
    const float x = y * kLn2F;
    const float series =
        x * __builtin_fmaf(-x, __builtin_fmaf(-x, __builtin_fmaf(-x, 0.041666668f, 0.16666667f),
                                              0.5f),
                           1.0f);
    const float den = x < 0.0625f ? series : 1.0f - t;
    return __builtin_amdgcn_logf((1.0f + t) * __builtin_amdgcn_rcpf(den));
}

// One sum-product check row over NT slots of the lane (edges e = t*P + p);
// Rm is the frame's messages, [slot][Z].  Branch-free like qc_ms_row_n: the
// old messages of the lane's edges are loaded together up front (slots past
// the row's end load a real slot of the row and are ignored); Q, phi(|Q|)
// and the LDS address of every edge stay in registers between the passes.
// The phi sum is taken in ascending edge order whatever P is: with P lanes
// every lane gathers the row's phi values of each slot (lane p holds edge
// tP + p) and adds them in edge order, so all lanes hold the one-lane sum
// and every tile decodes bit for bit alike.  Padding edges add +0.
//
// Signs as in qc_ms_row_n: no Q is -0.0, so Q < 0 is Q's sign bit; the
// row's sign parity is the XOR of the valid Q's bit patterns (bit 31), and
// it is folded into the ln 2 that scales each new magnitude, so a new
// message is that product with Q's sign bit XORed in (one bitop3).  The
// messages of the lane's edges sit P Z floats apart from its first edge's
// slot s0 + p; a padding slot (e >= dc, possibly past the last slot) reads
// slot s0 instead.  PAR (one lane per row, flooding): *hard_parity = the
// parity of the hard decisions on the APPs the row reads.
template <int P, int NT, bool UPDATE_APP, bool PAR = false>
__device__ __forceinline__ void qc_sp_row_n(const int2* tab, float* a, float* __restrict__ Rm,
                                            int s0, int dc, int r, int p, int Z, bool first,
                                            unsigned* hard_parity = nullptr) {
    static_assert(!PAR || (P == 1 && !UPDATE_APP), "read-APP parity: one lane per row, flooding");
    constexpr int S = kSlotStep<P>;
    float q[NT];
    float ph[NT];
    float* addr[NT];
    float* ar = a + r;
    const int z4 = neg_col_bytes(Z);
    const unsigned rr = row_rank(r);
    float* rm = Rm + (static_cast<size_t>(s0 + p) * Z + r);  // edge p's message
    const int step = P * Z;
    float rold[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int e = t * P + p;
        const bool valid = (t < NT - S) || (e < dc);
        addr[t] = &lds_at(ar, qc_off(tab[s0 + t * P + p], rr, z4));
        rold[t] = first ? 0.0f : rm[valid ? t * step : -p * Z];  // padding: slot s0
    }
    float total = 0.0f;
    unsigned neg = 0u;   // bit 31: parity of the valid Q's signs
    unsigned hard = 0u;  // PAR: bit 31: parity of the valid APPs' signs
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int e = t * P + p;
        const bool valid = (t < NT - S) || (e < dc);
        const float av = *addr[t];
        if (PAR) hard ^= valid ? __float_as_uint(av) : 0u;
        q[t] = av - rold[t];
        ph[t] = valid ? phi2_fast(fabsf(q[t]) * kLog2eF) : 0.0f;
        neg ^= valid ? __float_as_uint(q[t]) : 0u;
    }
    if constexpr (P == 1) {
#pragma unroll
        for (int t = 0; t < NT; ++t) total = total + ph[t];
    } else if constexpr (P == 2) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int2 v = row_pair(__float_as_int(ph[t]));  // edges 2t, 2t + 1
            total = total + __int_as_float(v.x);
            total = total + __int_as_float(v.y);
        }
        const int2 vn = row_pair(static_cast<int>(neg));
        neg = static_cast<unsigned>(vn.x ^ vn.y);
    } else {
        static_assert(P == 4, "lanes per row: 1, 2 or 4");
#pragma unroll
        for (int t = 0; t < NT; ++t) {  // edges 4t .. 4t + 3
            const int v = __float_as_int(ph[t]);
            total = total + __int_as_float(quad_lane<0>(v));
            total = total + __int_as_float(quad_lane<1>(v));
            total = total + __int_as_float(quad_lane<2>(v));
            total = total + __int_as_float(quad_lane<3>(v));
        }
#pragma unroll
        for (int off = 1; off < P; off <<= 1)
            neg ^= static_cast<unsigned>(lane_xor(static_cast<int>(neg), off));
    }
    // -ln 2 for an odd row parity: the product carries the parity's sign.
    float ln2s = (neg & 0x80000000u) ? -kLn2F : kLn2F;
    asm volatile("" : "+v"(ln2s));  // one select per row, not per edge
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int e = t * P + p;
        if ((t < NT - S) || (e < dc)) {
            const float mag = phi2_fast(total - ph[t]) * ln2s;
            const float rn = __uint_as_float(__float_as_uint(mag) ^
                                             (__float_as_uint(q[t]) & 0x80000000u));
            if (UPDATE_APP) *addr[t] = q[t] + rn;
            rm[t * step] = rn;
        }
    }
    if (PAR) *hard_parity = hard >> 31;  // no APP is -0.0: sign bit = (APP < 0)
}

template <int P, bool UPDATE_APP, int MAXT, int G = 1, bool PAR = false>
__device__ __forceinline__ void qc_sp_dispatch(int groups, const int2* tab, float* a,
                                               float* __restrict__ Rm, int s0, int dc, int r, int p,
                                               int Z, bool first, unsigned* hard_parity) {
    constexpr int S = kSlotStep<P>;
    if constexpr (S * G >= MAXT) {
        qc_sp_row_n<P, S * G, UPDATE_APP, PAR>(tab, a, Rm, s0, dc, r, p, Z, first, hard_parity);
    } else {
        if (groups <= G) {
            qc_sp_row_n<P, S * G, UPDATE_APP, PAR>(tab, a, Rm, s0, dc, r, p, Z, first,
                                                   hard_parity);
            return;
        }
        qc_sp_dispatch<P, UPDATE_APP, MAXT, G + 1, PAR>(groups, tab, a, Rm, s0, dc, r, p, Z, first,
                                                        hard_parity);
    }
}

// One sum-product check row `r` of block row `br`.
template <int P, bool UPDATE_APP, int MAXT = kMaxDegree / P, bool PAR = false>
__device__ __forceinline__ void qc_sp_row(const ck_ldpc_code& c, const int2* tab, float* a,
                                          float* __restrict__ Rm, int br, int r, int p, bool first,
                                          unsigned* hard_parity = nullptr) {
    const int s0 = ld_uniform(c.qc_row_ptr, br);
    const int dc = ld_uniform(c.qc_row_ptr, br + 1) - s0;
    const int groups = ((dc + P - 1) / P + kSlotStep<P> - 1) / kSlotStep<P>;
    qc_sp_dispatch<P, UPDATE_APP, MAXT, 1, PAR>(groups, tab, a, Rm, s0, dc, r, p, c.lifting,
                                                first, hard_parity);
}

}  // namespace ckdev
