// LDPC decoding on gfx950: layered and flooding belief propagation,
// normalized/offset min-sum and sum-product, syndrome early stop.
//
// Arithmetic contract (commkit/coding/_numba_loops.py:31-147,
// commkit/_cuda/src/ldpc.cu:37-159): a check row makes two passes over its
// edges in ascending column order.  Pass 1 forms Q = APP - R_old and keeps
// the sign parity and, for min-sum, min1 (the first edge attaining it is
// `idx`) and min2 (the second order statistic); for sum-product the running
// float32 sum of phi(|Q|) with phi(x) = -log tanh(x/2), x clipped to
// [1e-7, 30].  Pass 2 writes R_new = sign * mag and APP = Q + R_new, with
// mag = max(scale * min - offset, 0) (min2 for edge idx) or
// phi(total - phi(|Q|)).  This translation unit is compiled with
// -ffp-contract=off, so every product and sum rounds on its own, as in the
// reference: min-sum is bit-identical to the CPU restatement, including
// iteration counts and valid flags.  Early stop checks every row's parity of
// the hard decisions (APP < 0) before the first iteration and after each.
//
// Kernels:
//   qc_layered  quasi-cyclic codes.  A workgroup decodes F frames whose
//               APPs sit in LDS; lane (f, r, p) owns row r of the current
//               block row for frame f, with P lanes per row splitting its
//               edges and combining min1/min2/idx/sign with wavefront
//               shuffles (an exact, order-independent reduction on
//               (|Q|, edge) pairs).  Columns come from the circulant shift,
//               (r + s) mod Z, so consecutive lanes touch consecutive LDS
//               words.  Block rows run in schedule order (a layer is a union
//               of whole block rows, whose rows share no variable), with a
//               barrier after each.  Min-sum keeps each check row's state
//               compressed to {m1, m2, sign bits, idx} (16 bytes instead of
//               4 bytes per edge), in LDS or HBM; sum-product keeps every
//               R in HBM, laid out [frame][slot][Z] so lanes over r load
//               consecutive words.
//   qc_flooding the flooding schedule on the same layout: every check uses
//               the previous iteration's APP; then APP = L + sum of R over
//               the variable's checks, in ascending check order.
//   generic     any code through its CSR schedule, one workgroup per frame,
//               rows of a layer strided over the lanes, APP in LDS when it
//               fits (else in the workspace), R in the workspace in CSR edge
//               order.  Also flooding, through the variable-to-edge view.
//
// Frames are independent, and a frame's arithmetic does not depend on its
// workgroup, its slot or its neighbours, so a frame decodes to the same
// bits, LLRs and iteration count alone or in any batch.
#include "ck_common.h"

#include <algorithm>

namespace {

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
    const float mag = (e == st.idx) ? st.m2 : st.m1;
    return ((st.signs >> e) & 1u) ? -mag : mag;
}

// Channel LLRs of frames [frame0, frame0 + F) into LDS (zeros for the
// punctured tail and for frames past R).
__device__ void load_frames(float* lds, const float* __restrict__ llr, long long frame0, int F,
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
__device__ void store_frame(const ck_ldpc_code& c, const float* a, long long frame, int u,
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

// Copies the packed {block column * Z, shift} slot table (S + 32 entries,
// zero padded) into LDS, where every row update reads it.
__device__ void stage_slot_table(const ck_ldpc_code& c, int2* tab) {
    const int n = c.num_slots + kMaxDegree;
    const int2* src = reinterpret_cast<const int2*>(c.qc_slot_packed);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tab[i] = src[i];
}

__device__ __forceinline__ int qc_addr(int jz, int r, int sh, int Z) {
    int col = r + sh;
    col -= (col >= Z) ? Z : 0;
    return jz + col;
}

// Whether every check of the lane's frame is satisfied; `test` says whether
// this lane's frame is being checked.  Called by every lane (barriers).
template <int P>
__device__ bool qc_syndrome_ok(const ck_ldpc_code& c, const int2* tab, const float* a,
                               const QcLane& l, bool test, int F, int* s_flag) {
    const int Z = c.lifting;
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
                    v[u] = a[qc_addr(e.x, l.r, e.y, Z)];
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

// One min-sum check row of a block row, given its old state; returns the new
// state.  UPDATE_APP: layered (APP = Q + R_new); else the flooding check
// phase, which leaves the APP alone.  Q and the LDS address of every edge of
// the lane stay in registers between the two passes (MAXT = 32 / P slots).
template <int P, bool UPDATE_APP>
__device__ __forceinline__ CheckState qc_ms_row(const ck_ldpc_code& c, const int2* tab, float* a,
                                                const CheckState st, int br, int r, int p,
                                                const Params& prm) {
    constexpr int MAXT = kMaxDegree / P;
    constexpr int G = 8;  // slots per group; groups past the row's end are skipped
    const int Z = c.lifting;
    const int s0 = ld_uniform(c.qc_row_ptr, br);
    const int dc = ld_uniform(c.qc_row_ptr, br + 1) - s0;
    const int tmax = (dc + P - 1) / P;
    float q[MAXT];
    int addr[MAXT];
    // Every load of the row first, LDS only (slot table, then APP), so they
    // are in flight together.  Slots past the row's end inside the last group
    // read real or padding table entries: their addresses stay in the frame.
#pragma unroll
    for (int g = 0; g < MAXT; g += G) {
        if (g < tmax) {
#pragma unroll
            for (int t = g; t < g + G && t < MAXT; ++t) {
                const int2 v = tab[s0 + t * P + p];
                addr[t] = qc_addr(v.x, r, v.y, Z);
                q[t] = a[addr[t]];
            }
        }
    }
    // Pass 1.  min1 / min2 / idx follow the sequential rule exactly:
    //   idx moves on a strictly smaller |Q| (the first edge attaining min1),
    //   min2 = min(min2, max(min1, |Q|)), min1 = min(min1, |Q|)
    // (min and max return one of their operands, so no value is rounded).
    // Bit e of qneg is Q_e < 0; the sign parity is its popcount.
    float min1 = ck::kInf;
    float min2 = ck::kInf;
    int idx = 0x7fffffff;
    unsigned qneg = 0u;
#pragma unroll
    for (int g = 0; g < MAXT; g += G) {
        if (g < tmax) {
#pragma unroll
            for (int t = g; t < g + G && t < MAXT; ++t) {
                const int e = t * P + p;
                const bool valid = e < dc;
                const float qv = q[t] - state_msg(st, e);
                q[t] = qv;
                const float mag = valid ? fabsf(qv) : ck::kInf;
                idx = (mag < min1) ? e : idx;
                min2 = fminf(min2, fmaxf(min1, mag));
                min1 = fminf(min1, mag);
                qneg |= (valid && qv < 0.0f) ? (1u << e) : 0u;
            }
        }
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
        for (int g = 0; g < MAXT; g += G) {
            if (g < tmax) {
#pragma unroll
                for (int t = g; t < g + G && t < MAXT; ++t) {
                    const int e = t * P + p;
                    if (e < dc) {
                        const float mag = (e == idx) ? m2 : m1;
                        a[addr[t]] = q[t] + (((signs >> e) & 1u) ? -mag : mag);
                    }
                }
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
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
        if (t < dc) {
            const int2 v = tab[s0 + t];
            addr[t] = qc_addr(v.x, r, v.y, Z);
            q[t] = a[addr[t]] - rold[t];
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
            if (UPDATE_APP) a[addr[t]] = q[t] + rn;
            Rm[static_cast<size_t>(s0 + t) * Z + r] = rn;
        }
    }
}

template <bool SP, int P, bool STATE_LDS>
__global__ __launch_bounds__(kMaxQcThreads) void ldpc_qc_layered(
    ck_ldpc_code c, const float* __restrict__ llr, float* __restrict__ app_out,
    signed char* __restrict__ bits_out, int* __restrict__ iters_out,
    unsigned char* __restrict__ valid_out, void* __restrict__ ws, long long R, int F, Params prm) {
    extern __shared__ float lds[];
    __shared__ int s_flag[kMaxFrames];
    __shared__ int2 s_tab[kTabEntries];
    const int N = c.num_variables;
    const int Z = c.lifting;
    const int mb = c.base_rows;
    const QcLane l = qc_lane<P>(Z, F, R);
    float* a = lds + static_cast<size_t>(l.f < F ? l.f : 0) * N;
    stage_slot_table(c, s_tab);
    load_frames(lds, llr, static_cast<long long>(blockIdx.x) * F, F, R, c.n, N);
    __syncthreads();

    CheckState* st_base = nullptr;
    float* Rm = nullptr;
    if (SP) {
        Rm = static_cast<float*>(ws) + static_cast<size_t>(l.live ? l.frame : 0) * c.num_edges;
    } else if (STATE_LDS) {
        st_base = reinterpret_cast<CheckState*>(lds + state_offset(F, N)) +
                  static_cast<size_t>(l.f < F ? l.f : 0) * c.num_checks;
    } else {
        st_base = static_cast<CheckState*>(ws) +
                  static_cast<size_t>(l.live ? l.frame : 0) * c.num_checks;
    }

    bool done = !l.live;
    bool ok = false;
    int iters = 0;
    if (prm.early_stop) {
        ok = qc_syndrome_ok<P>(c, s_tab, a, l, l.live, F, s_flag);
        done = done || ok;
    }
    int it = 0;
    // The state of the next block row is loaded while the current one is
    // processed; before the first iteration every state is zero.
    CheckState st_next = zero_state();
    while (__syncthreads_or(!done) && it < prm.max_iterations) {
        const bool first = it == 0;
        for (int br = 0; br < mb; ++br) {
            if (!done) {
                if (SP) {
                    qc_sp_row<true>(c, s_tab, a, Rm, br, l.r, first);
                } else {
                    const CheckState st = st_next;
                    if (br + 1 < mb) {
                        st_next = first ? zero_state()
                                        : st_base[static_cast<size_t>(br + 1) * Z + l.r];
                    } else if (mb > 1) {
                        st_next = st_base[l.r];  // block row 0, stored this iteration
                    }
                    const CheckState ns = qc_ms_row<P, true>(c, s_tab, a, st, br, l.r, l.p, prm);
                    if (l.p == 0) st_base[static_cast<size_t>(br) * Z + l.r] = ns;
                    if (mb == 1) st_next = ns;
                }
            }
            __syncthreads();
        }
        ++it;
        if (!done) iters = it;
        if (prm.early_stop) {
            const bool now = qc_syndrome_ok<P>(c, s_tab, a, l, !done, F, s_flag);
            if (!done) {
                ok = now;
                done = now;
            }
        }
    }
    if (!prm.early_stop) ok = qc_syndrome_ok<P>(c, s_tab, a, l, l.live, F, s_flag);
    __syncthreads();
    if (l.live)
        store_frame(c, a, l.frame, l.u, Z * P, iters, ok, app_out, bits_out, iters_out, valid_out);
}

template <bool SP, int P>
__global__ __launch_bounds__(kMaxQcThreads) void ldpc_qc_flooding(
    ck_ldpc_code c, const float* __restrict__ llr, float* __restrict__ app_out,
    signed char* __restrict__ bits_out, int* __restrict__ iters_out,
    unsigned char* __restrict__ valid_out, void* __restrict__ ws, long long R, int F, Params prm) {
    extern __shared__ float lds[];
    __shared__ int s_flag[kMaxFrames];
    __shared__ int2 s_tab[kTabEntries];
    const int N = c.num_variables;
    const int Z = c.lifting;
    const QcLane l = qc_lane<P>(Z, F, R);
    float* a = lds + static_cast<size_t>(l.f < F ? l.f : 0) * N;
    stage_slot_table(c, s_tab);
    load_frames(lds, llr, static_cast<long long>(blockIdx.x) * F, F, R, c.n, N);
    __syncthreads();

    const size_t fr = static_cast<size_t>(l.live ? l.frame : 0);
    CheckState* st_base = SP ? nullptr : static_cast<CheckState*>(ws) + fr * c.num_checks;
    float* Rm = SP ? static_cast<float*>(ws) + fr * c.num_edges : nullptr;
    const float* L = llr + fr * c.n;

    bool done = !l.live;
    bool ok = false;
    int iters = 0;
    if (prm.early_stop) {
        ok = qc_syndrome_ok<P>(c, s_tab, a, l, l.live, F, s_flag);
        done = done || ok;
    }
    int it = 0;
    while (__syncthreads_or(!done) && it < prm.max_iterations) {
        const bool first = it == 0;
        // Check phase: every row reads the previous iteration's APP.
        if (!done) {
            for (int br = 0; br < c.base_rows; ++br) {
                if (SP) {
                    if (l.p == 0) qc_sp_row<false>(c, s_tab, a, Rm, br, l.r, first);
                } else {
                    CheckState* sp = st_base + static_cast<size_t>(br) * Z + l.r;
                    const CheckState st = first ? zero_state() : *sp;
                    const CheckState ns = qc_ms_row<P, false>(c, s_tab, a, st, br, l.r, l.p, prm);
                    if (l.p == 0) *sp = ns;
                }
            }
        }
        __syncthreads();
        // Variable phase: APP = L + the new R of every check, ascending check.
        if (!done) {
            for (int v = l.u; v < N; v += Z * P) {
                const int j = v / Z;
                const int cc = v - j * Z;
                float sum = v < c.n ? L[v] : 0.0f;
                for (int t = c.qc_col_ptr[j]; t < c.qc_col_ptr[j + 1]; ++t) {
                    const int s = c.qc_col_slot[t];
                    int rr = cc - c.qc_slot_shift[s];
                    rr += (rr < 0) ? Z : 0;
                    float rv;
                    if (SP) {
                        rv = Rm[static_cast<size_t>(s) * Z + rr];
                    } else {
                        const CheckState st =
                            st_base[static_cast<size_t>(c.qc_slot_row[s]) * Z + rr];
                        rv = state_msg(st, c.qc_slot_pos[s]);
                    }
                    sum = sum + rv;
                }
                a[v] = sum;
            }
        }
        __syncthreads();
        ++it;
        if (!done) iters = it;
        if (prm.early_stop) {
            const bool now = qc_syndrome_ok<P>(c, s_tab, a, l, !done, F, s_flag);
            if (!done) {
                ok = now;
                done = now;
            }
        }
    }
    if (!prm.early_stop) ok = qc_syndrome_ok<P>(c, s_tab, a, l, l.live, F, s_flag);
    __syncthreads();
    if (l.live)
        store_frame(c, a, l.frame, l.u, Z * P, iters, ok, app_out, bits_out, iters_out, valid_out);
}

// ---------------------------------------------------------------------------
// Generic CSR kernel
// ---------------------------------------------------------------------------

__device__ bool generic_syndrome_ok(const ck_ldpc_code& c, const float* a) {
    int bad = 0;
    for (int row = threadIdx.x; row < c.num_checks; row += blockDim.x) {
        int parity = 0;
        for (int e = c.row_ptr[row]; e < c.row_ptr[row + 1]; ++e) parity ^= a[c.cols[e]] < 0.0f;
        bad |= parity;
    }
    return !__syncthreads_or(bad);
}

template <bool SP, bool UPDATE_APP>
__device__ __forceinline__ void generic_row(const ck_ldpc_code& c, float* a,
                                            float* __restrict__ Rm, int row, bool first,
                                            const Params& prm) {
    const int e0 = c.row_ptr[row];
    const int e1 = c.row_ptr[row + 1];
    unsigned neg = 0u;
    if (SP) {
        float total = 0.0f;
        for (int e = e0; e < e1; ++e) {
            const float q = a[c.cols[e]] - (first ? 0.0f : Rm[e]);
            neg ^= (q < 0.0f) ? 1u : 0u;
            total = total + phi(fabsf(q));
        }
        for (int e = e0; e < e1; ++e) {
            const int col = c.cols[e];
            const float q = a[col] - (first ? 0.0f : Rm[e]);
            const float mag = phi(total - phi(fabsf(q)));
            const float rn = (neg ^ ((q < 0.0f) ? 1u : 0u)) ? -mag : mag;
            if (UPDATE_APP) a[col] = q + rn;
            Rm[e] = rn;
        }
        return;
    }
    float min1 = ck::kInf;
    float min2 = ck::kInf;
    int idx = -1;
    for (int e = e0; e < e1; ++e) {
        const float q = a[c.cols[e]] - (first ? 0.0f : Rm[e]);
        neg ^= (q < 0.0f) ? 1u : 0u;
        const float mag = fabsf(q);
        if (mag < min1) {
            min2 = min1;
            min1 = mag;
            idx = e;
        } else if (mag < min2) {
            min2 = mag;
        }
    }
    const float m1 = fmaxf(prm.scale * min1 - prm.offset, 0.0f);
    const float m2 = fmaxf(prm.scale * min2 - prm.offset, 0.0f);
    for (int e = e0; e < e1; ++e) {
        const int col = c.cols[e];
        const float q = a[col] - (first ? 0.0f : Rm[e]);
        const float mag = (e == idx) ? m2 : m1;
        const float rn = (neg ^ ((q < 0.0f) ? 1u : 0u)) ? -mag : mag;
        if (UPDATE_APP) a[col] = q + rn;
        Rm[e] = rn;
    }
}

template <bool SP, bool FLOOD, bool APP_LDS>
__global__ __launch_bounds__(kGenericThreads) void ldpc_generic(
    ck_ldpc_code c, const float* __restrict__ llr, float* __restrict__ app_out,
    signed char* __restrict__ bits_out, int* __restrict__ iters_out,
    unsigned char* __restrict__ valid_out, float* __restrict__ msg_ws,
    float* __restrict__ app_ws, long long R, Params prm) {
    extern __shared__ float lds[];
    const long long frame = blockIdx.x;
    const int N = c.num_variables;
    float* a = APP_LDS ? lds : app_ws + static_cast<size_t>(frame) * N;
    float* Rm = msg_ws + static_cast<size_t>(frame) * c.num_edges;
    const float* L = llr + frame * c.n;
    for (int v = threadIdx.x; v < N; v += blockDim.x) a[v] = v < c.n ? L[v] : 0.0f;
    __syncthreads();

    int it = 0;
    bool ok = prm.early_stop && generic_syndrome_ok(c, a);
    while (!ok && it < prm.max_iterations) {
        const bool first = it == 0;
        if (FLOOD) {
            for (int row = threadIdx.x; row < c.num_checks; row += blockDim.x)
                generic_row<SP, false>(c, a, Rm, row, first, prm);
            __syncthreads();
            for (int v = threadIdx.x; v < N; v += blockDim.x) {
                float sum = v < c.n ? L[v] : 0.0f;
                for (int t = c.var_ptr[v]; t < c.var_ptr[v + 1]; ++t) sum = sum + Rm[c.var_edges[t]];
                a[v] = sum;
            }
            __syncthreads();
        } else {
            for (int layer = 0; layer < c.num_layers; ++layer) {
                const int end = c.layer_ptr[layer + 1];
                for (int row = c.layer_ptr[layer] + threadIdx.x; row < end; row += blockDim.x)
                    generic_row<SP, true>(c, a, Rm, row, first, prm);
                __syncthreads();
            }
        }
        ++it;
        if (prm.early_stop) ok = generic_syndrome_ok(c, a);
    }
    if (!prm.early_stop) ok = generic_syndrome_ok(c, a);
    __syncthreads();
    store_frame(c, a, frame, threadIdx.x, blockDim.x, it, ok, app_out, bits_out, iters_out,
                valid_out);
}

// ---------------------------------------------------------------------------
// Host: planning, workspace and launch
// ---------------------------------------------------------------------------

struct Plan {
    int kind;          // 1 generic, 2 quasi-cyclic
    bool sp;
    bool flood;
    bool app_lds;      // generic
    int F;             // QC frames per workgroup
    int P;             // QC lanes per row
    bool state_lds;    // QC layered min-sum
    int threads;
    size_t lds_bytes;
    size_t ws_per_frame;  // bytes
};

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int32_t check_code(const ck_ldpc_code* c) {
    CK_REQUIRE(c != nullptr, "code is NULL");
    CK_REQUIRE(c->num_checks >= 1 && c->num_variables > c->num_checks,
               "need 1 <= num_checks < num_variables (%d, %d)", c->num_checks, c->num_variables);
    CK_REQUIRE(c->n >= 1 && c->n <= c->num_variables, "n = %d not in [1, N]", c->n);
    CK_REQUIRE(c->k >= 0 && c->k <= c->num_variables, "k = %d not in [0, N]", c->k);
    CK_REQUIRE(c->num_edges >= 2 * c->num_checks, "num_edges %d < 2 * num_checks", c->num_edges);
    CK_REQUIRE(c->num_layers >= 1 && c->num_layers <= c->num_checks, "num_layers %d out of range",
               c->num_layers);
    CK_REQUIRE(c->row_ptr && c->cols && c->layer_ptr, "NULL CSR table");
    if (c->lifting > 0) {
        CK_REQUIRE(c->base_rows * c->lifting == c->num_checks &&
                       c->base_cols * c->lifting == c->num_variables,
                   "QC view %d x %d x %d does not match m = %d, N = %d", c->base_rows,
                   c->base_cols, c->lifting, c->num_checks, c->num_variables);
        CK_REQUIRE(c->num_slots * c->lifting == c->num_edges, "QC slots do not cover the edges");
        CK_REQUIRE(c->qc_row_ptr && c->qc_slot_col && c->qc_slot_shift && c->qc_row_id &&
                       c->qc_col_ptr && c->qc_col_slot && c->qc_slot_pos && c->qc_slot_row &&
                       c->qc_slot_packed,
                   "NULL QC table");
    }
    return CK_OK;
}

int32_t check_method(const ck_ldpc_method* m) {
    CK_REQUIRE(m != nullptr, "method is NULL");
    CK_REQUIRE(m->rule == 0 || m->rule == 1, "rule %d not 0 (min-sum) or 1 (sum-product)",
               m->rule);
    CK_REQUIRE(m->schedule == 0 || m->schedule == 1, "schedule %d not 0 (layered) or 1 (flooding)",
               m->schedule);
    CK_REQUIRE(m->max_iterations >= 1, "max_iterations %d < 1", m->max_iterations);
    CK_REQUIRE(m->kernel >= 0 && m->kernel <= 2, "kernel %d not in [0, 2]", m->kernel);
    CK_REQUIRE(m->frames_per_group >= 0 && m->frames_per_group <= kMaxFrames,
               "frames_per_group %d not in [0, %d]", m->frames_per_group, kMaxFrames);
    CK_REQUIRE(m->lanes_per_row == 0 || m->lanes_per_row == 1 || m->lanes_per_row == 2 ||
                   m->lanes_per_row == 4,
               "lanes_per_row %d not 0, 1, 2 or 4", m->lanes_per_row);
    CK_REQUIRE(m->state_in_lds >= -1 && m->state_in_lds <= 1, "state_in_lds %d not -1, 0 or 1",
               m->state_in_lds);
    CK_REQUIRE(m->scale > 0.0f && m->scale <= 1.0f, "scale must be in (0, 1]");
    CK_REQUIRE(m->offset >= 0.0f, "offset must be >= 0");
    return CK_OK;
}

int32_t make_plan(const ck_ldpc_code* c, const ck_ldpc_method* m, Plan* plan) {
    if (int32_t st = check_code(c)) return st;
    if (int32_t st = check_method(m)) return st;
    ck::DeviceInfo info{};
    if (int32_t st = ck::device_info(&info)) return st;
    // Dynamic LDS budget: the per-workgroup limit less room for the kernels'
    // static LDS (frame flags, barrier scratch).
    const size_t lds_max =
        info.lds_per_block > kStaticLds + 2048 ? info.lds_per_block - kStaticLds - 1024 : 0;
    Plan p{};
    p.sp = m->rule == 1;
    p.flood = m->schedule == 1;
    const int N = c->num_variables;
    const int Z = c->lifting;
    const size_t app_bytes = static_cast<size_t>(N) * 4;
    bool qc_ok = Z > 0 && Z <= kMaxQcThreads && app_bytes <= lds_max &&
                 c->max_row_degree <= kMaxDegree && c->num_slots <= kMaxSlots;
    if (m->kernel == 2 && !qc_ok)
        return ck::fail(CK_EINVAL,
                        "quasi-cyclic kernel requested but the code has no usable QC view "
                        "(lifting %d, max row degree %d, %d slots, APP %zu bytes, LDS %zu bytes)",
                        Z, c->max_row_degree, c->num_slots, app_bytes, lds_max);
    if (m->kernel == 1) qc_ok = false;
    if (qc_ok) {
        p.kind = 2;
        int P = m->lanes_per_row;
        if (P == 0) P = 1;
        if (p.sp) P = 1;
        if (Z * P > kMaxQcThreads) P = 1;
        int F = m->frames_per_group;
        if (F == 0) F = std::max(1, 256 / (Z * P));
        F = std::min(F, kMaxQcThreads / (Z * P));
        F = std::min<int>(F, static_cast<int>(lds_max / app_bytes));
        F = std::max(F, 1);
        const size_t state_bytes = static_cast<size_t>(c->num_checks) * sizeof(CheckState);
        bool state_lds = false;
        if (!p.sp && !p.flood) {
            const size_t with = state_offset(F, N) * 4 + static_cast<size_t>(F) * state_bytes;
            const size_t without = static_cast<size_t>(F) * app_bytes;
            if (m->state_in_lds == 1) {
                if (with > lds_max)
                    return ck::fail(CK_EINVAL, "check state does not fit in LDS (%zu > %zu bytes)",
                                    with, lds_max);
                state_lds = true;
            } else if (m->state_in_lds == -1) {
                // Keep the state in LDS when it costs no workgroup per CU.
                state_lds = with <= lds_max && (lds_max / with) == (lds_max / without);
            }
        }
        p.F = F;
        p.P = P;
        p.state_lds = state_lds;
        p.threads = static_cast<int>(round_up(static_cast<size_t>(F) * Z * P, 64));
        p.lds_bytes = state_lds ? state_offset(F, N) * 4 + static_cast<size_t>(F) * state_bytes
                                : static_cast<size_t>(F) * app_bytes;
        if (p.sp)
            p.ws_per_frame = static_cast<size_t>(c->num_edges) * 4;
        else
            p.ws_per_frame = state_lds ? 0 : state_bytes;
    } else {
        CK_REQUIRE(!p.flood || (c->var_ptr && c->var_edges),
                   "flooding needs the variable-to-edge view (var_ptr, var_edges)");
        p.kind = 1;
        p.app_lds = app_bytes <= lds_max;
        p.threads = kGenericThreads;
        p.lds_bytes = p.app_lds ? app_bytes : 0;
        p.ws_per_frame = static_cast<size_t>(c->num_edges) * 4 + (p.app_lds ? 0 : app_bytes);
    }
    *plan = p;
    return CK_OK;
}

template <class K>
void allow_lds(K kernel, size_t bytes) {
    if (bytes > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(bytes));
        (void)hipGetLastError();
    }
}

template <bool SP, int P, bool SL>
void launch_qc_layered(const Plan& p, const ck_ldpc_code& c, const float* llr, float* app,
                       signed char* bits, int* iters, unsigned char* valid, void* ws, long long R,
                       const Params& prm, hipStream_t s) {
    auto kernel = ldpc_qc_layered<SP, P, SL>;
    allow_lds(kernel, p.lds_bytes);
    const long long groups = (R + p.F - 1) / p.F;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(groups)), dim3(p.threads), p.lds_bytes,
                       s, c, llr, app, bits, iters, valid, ws, R, p.F, prm);
}

template <bool SP, int P>
void launch_qc_flooding(const Plan& p, const ck_ldpc_code& c, const float* llr, float* app,
                        signed char* bits, int* iters, unsigned char* valid, void* ws,
                        long long R, const Params& prm, hipStream_t s) {
    auto kernel = ldpc_qc_flooding<SP, P>;
    allow_lds(kernel, p.lds_bytes);
    const long long groups = (R + p.F - 1) / p.F;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(groups)), dim3(p.threads), p.lds_bytes,
                       s, c, llr, app, bits, iters, valid, ws, R, p.F, prm);
}

template <bool SP, bool FLOOD, bool APP_LDS>
void launch_generic(const Plan& p, const ck_ldpc_code& c, const float* llr, float* app,
                    signed char* bits, int* iters, unsigned char* valid, void* ws, long long R,
                    const Params& prm, hipStream_t s) {
    auto kernel = ldpc_generic<SP, FLOOD, APP_LDS>;
    allow_lds(kernel, p.lds_bytes);
    float* msg = static_cast<float*>(ws);
    float* app_ws = APP_LDS ? nullptr : msg + static_cast<size_t>(R) * c.num_edges;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(R)), dim3(p.threads), p.lds_bytes, s, c,
                       llr, app, bits, iters, valid, msg, app_ws, R, prm);
}

}  // namespace

extern "C" int64_t ck_ldpc_workspace_bytes(const ck_ldpc_code* code, const ck_ldpc_method* method,
                                           int64_t R) {
    Plan p{};
    if (make_plan(code, method, &p) != CK_OK || R < 0) return -1;
    return static_cast<int64_t>(round_up(p.ws_per_frame * static_cast<size_t>(R), 256));
}

extern "C" int32_t ck_ldpc_decode(const ck_ldpc_code* code, const ck_ldpc_method* method,
                                  const float* llrs, float* app, int8_t* bits,
                                  int32_t* iterations, uint8_t* valid, int64_t R, void* workspace,
                                  void* stream) {
    Plan p{};
    if (int32_t st = make_plan(code, method, &p)) return st;
    CK_REQUIRE(R >= 0, "R = %lld < 0", static_cast<long long>(R));
    if (R == 0) return CK_OK;
    CK_REQUIRE(R <= 0x7fffffffLL, "R = %lld exceeds one launch", static_cast<long long>(R));
    CK_REQUIRE(llrs && app && iterations && valid, "NULL input or output");
    CK_REQUIRE(p.ws_per_frame == 0 || workspace != nullptr, "NULL workspace (%zu bytes per frame)",
               p.ws_per_frame);
    CK_REQUIRE((reinterpret_cast<uintptr_t>(workspace) & 255) == 0, "workspace not 256-byte aligned");
    const ck_ldpc_code& c = *code;
    Params prm{method->scale, method->offset, method->max_iterations, method->early_stop != 0};
    if (p.sp) {
        prm.scale = 1.0f;
        prm.offset = 0.0f;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    signed char* b = reinterpret_cast<signed char*>(bits);
    int* it = iterations;
    unsigned char* v = valid;
    if (p.kind == 2) {
        if (p.flood) {
            if (p.sp) launch_qc_flooding<true, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else if (p.P == 1) launch_qc_flooding<false, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else if (p.P == 2) launch_qc_flooding<false, 2>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else launch_qc_flooding<false, 4>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        } else if (p.sp) {
            launch_qc_layered<true, 1, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        } else if (p.state_lds) {
            if (p.P == 1) launch_qc_layered<false, 1, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else if (p.P == 2) launch_qc_layered<false, 2, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else launch_qc_layered<false, 4, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        } else {
            if (p.P == 1) launch_qc_layered<false, 1, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else if (p.P == 2) launch_qc_layered<false, 2, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else launch_qc_layered<false, 4, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        }
        return ck::launched("ldpc quasi-cyclic decode");
    }
    if (p.flood) {
        if (p.sp) {
            if (p.app_lds) launch_generic<true, true, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else launch_generic<true, true, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        } else {
            if (p.app_lds) launch_generic<false, true, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else launch_generic<false, true, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        }
    } else {
        if (p.sp) {
            if (p.app_lds) launch_generic<true, false, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else launch_generic<true, false, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        } else {
            if (p.app_lds) launch_generic<false, false, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
            else launch_generic<false, false, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s);
        }
    }
    return ck::launched("ldpc generic decode");
}
