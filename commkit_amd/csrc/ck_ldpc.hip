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
// phi(total - phi(|Q|)).  Every kernel evaluates phi the same way, in log2
// units from the hardware exp2 / reciprocal / log2 (phi2_fast, within the
// sum-product tolerance of -logf(tanhf)), and sums a row's phi values in
// ascending edge order, so sum-product, too, decodes bit for bit alike on
// every kernel and tile.  This translation unit is compiled with
// -ffp-contract=off, so every product and sum rounds on its own, as in the
// reference: min-sum is bit-identical to the CPU restatement, including
// iteration counts and valid flags.  Early stop checks every row's parity of
// the hard decisions (APP < 0) before the first iteration and after each.
//
// Kernels:
//   qc_layered  quasi-cyclic codes.  A workgroup decodes F frames whose
//               APPs sit in LDS; lane (f, r, p) owns row r of the current
//               block row for frame f, with P lanes per row splitting its
//               edges and combining min1/min2/idx/sign across the row's
//               lanes (P = 2: the two halves of a wave, one
//               v_permlane32_swap per value; P = 4: DPP quad moves) -- an
//               exact, order-independent reduction on (|Q|, edge) pairs; for
//               sum-product the phi sum and sign parity.  Columns come from
//               the circulant shift, (r + s) mod Z, so consecutive lanes
//               (of a half-wave for P = 2) touch consecutive LDS words.  Block rows run in schedule order (a layer is a union
//               of whole block rows, whose rows share no variable), with a
//               barrier after each.  Min-sum keeps each check row's state
//               compressed to {m1, m2, sign bits, idx} (16 bytes instead of
//               4 bytes per edge), in LDS or HBM; sum-product keeps every
//               R in HBM, laid out [frame][slot][Z] so lanes over r load
//               consecutive words.
//   qc_flooding the flooding schedule on the same layout: every check uses
//               the previous iteration's APP; then APP = L + sum of R over
//               the variable's checks, in ascending check order.  The
//               check phase's reads also give the syndrome of the APP it
//               reads, so the stop rule needs no pass of its own.
//   generic     any code through its CSR schedule, one workgroup per frame,
//               rows of a layer strided over the lanes, APP in LDS when it
//               fits (else in the workspace), R in the workspace in CSR edge
//               order.  Also flooding, through the variable-to-edge view.
//
// Frames are independent, and a frame's arithmetic does not depend on its
// workgroup, its slot or its neighbours, so a frame decodes to the same
// bits, LLRs and iteration count alone or in any batch.
#include "ck_common.h"
#include "ck_ldpc_dev.h"

#include <algorithm>
#include <mutex>

namespace {

using namespace ckdev;

// SHORT (min-sum, short rows): row instantiations up to kShortRow / P slots
// with a register budget of 168 VGPRs (SHORT = 1: one workgroup of up to
// twelve waves per CU), or up to kShortRowTwo slots with 128 (SHORT = 2,
// P = 1: two workgroups of up to six waves, whose barriers and frame
// loads and stores overlap).
template <int P, int SHORT>
constexpr int kRowSlots = SHORT == 2 ? kShortRowTwo / P : (SHORT == 1 ? kShortRow / P : kMaxDegree / P);

template <bool SP, int P, bool STATE_LDS, int SHORT>
__global__ __launch_bounds__(kMaxQcThreads)
__attribute__((amdgpu_waves_per_eu(SHORT == 2 ? 4 : (SHORT == 1 ? 3 : 1)))) void ldpc_qc_layered(
    ck_ldpc_code c, const float* __restrict__ llr, float* __restrict__ app_out,
    signed char* __restrict__ bits_out, int* __restrict__ iters_out,
    unsigned char* __restrict__ valid_out, void* __restrict__ ws, long long R, int F, Params prm,
    SymbolSource sym) {
    // 16-byte aligned: float4 frame loads, the fused demapping's tables
    extern __shared__ __attribute__((aligned(16))) float lds[];
    // s_flag[it & 1]: the syndrome's flags of the check after iteration
    // `it` (min-sum: the last block row's parities, set during the
    // iteration); the other buffer is cleared meanwhile.
    __shared__ int s_flag[2][kMaxFrames];
    __shared__ int2 s_tab[kTabEntries];
    const int N = c.num_variables;
    const int Z = c.lifting;
    const int mb = c.base_rows;
    const QcLane l = qc_lane<P>(Z, F, R);
    float* a = lds + static_cast<size_t>(l.f < F ? l.f : 0) * N;
    const QcLds lay = qc_lds(F, N, Z, c.num_checks, STATE_LDS);
    stage_slot_table(c, s_tab);
    if (sym.x != nullptr) {  // demapping fused into the load; tables past the APPs
        load_frames_demap(lds, sym, lds + lay.bits, static_cast<long long>(blockIdx.x) * F, F, R,
                          c.n, N);
    } else {
        load_frames(lds, llr, static_cast<long long>(blockIdx.x) * F, F, R, c.n, N);
    }
    __syncthreads();

    unsigned* bits = reinterpret_cast<unsigned*>(lds + lay.bits);
    const bool lane_ok = l.f < F;
    CheckState* st_base = nullptr;
    float* Rm = nullptr;
    if (SP) {
        Rm = static_cast<float*>(ws) + static_cast<size_t>(l.live ? l.frame : 0) * c.num_edges;
    } else if (STATE_LDS) {
        st_base = reinterpret_cast<CheckState*>(lds + lay.state) +
                  static_cast<size_t>(l.f < F ? l.f : 0) * c.num_checks;
    } else {
        st_base = static_cast<CheckState*>(ws) +
                  static_cast<size_t>(l.live ? l.frame : 0) * c.num_checks;
    }

    bool done = !l.live;
    bool ok = false;
    int iters = 0;
    int it = 0;
    // The state of the next block row is loaded while the current one is
    // processed; before the first iteration every state is zero.
    CheckState st_next = zero_state();
    for (;;) {
        // The syndrome is checked before the first iteration and after each
        // one with early stop, else once after the last (one call site keeps
        // the kernel's registers down).  After an iteration of min-sum the
        // last block row's hard-decision parities (final: no later row of
        // the iteration touches its bits) have already settled s_flag.
        const bool last = it >= prm.max_iterations;
        int* flag = s_flag[(it + 1) & 1];  // set by the last iteration's last block row
        // One frame whose last block row left an odd parity: not satisfied,
        // and the workgroup runs on; every lane reads the same flag, so the
        // syndrome's pass and the workgroup's OR are skipped alike.
        const bool rejected = F == 1 && !SP && mb > 1 && it > 0 && prm.early_stop && !last &&
                              flag[0] == 1;
        if (rejected) {
            ok = false;
        } else {
            if (prm.early_stop || last) {
                const bool now = qc_syndrome_ok(c, s_tab, lds, bits, l, lane_ok, !done, F, flag,
                                                !SP && it > 0 && mb > 1);
                if (!done) {
                    ok = now;
                    if (prm.early_stop) done = now;
                }
            }
            if (last || !__syncthreads_or(!done)) break;
        }
        const bool first = it == 0;
        // Cleared for the parities of this iteration's last block row: last
        // read before the previous iteration's block rows, whose barriers
        // order those reads before this write and this write before the
        // last block row's.
        int* next = s_flag[it & 1];
        if (!SP && mb > 1 && lane_ok && l.u == 0) next[l.f] = 0;
        for (int br = 0; br < mb; ++br) {
            if (!done) {
                if (SP) {
                    qc_sp_row<P, true>(c, s_tab, a, Rm, br, l.r, l.p, first);
                } else {
                    const CheckState st = st_next;
                    if (br + 1 < mb) {
                        st_next = first ? zero_state()
                                        : st_base[static_cast<size_t>(br + 1) * Z + l.r];
                    } else if (mb > 1) {
                        st_next = st_base[l.r];  // block row 0, stored this iteration
                    }
                    CheckState ns;
                    if (br + 1 < mb || mb == 1) {
                        ns = qc_ms_row<P, true, kRowSlots<P, SHORT>>(c, s_tab, a, st, br, l.r, l.p,
                                                                     prm);
                    } else {
                        unsigned odd = 0u;
                        ns = qc_ms_row<P, true, kRowSlots<P, SHORT>, true>(c, s_tab, a, st, br,
                                                                           l.r, l.p, prm, &odd);
                        if (odd) next[l.f] = 1;
                    }
                    if (l.p == 0) st_base[static_cast<size_t>(br) * Z + l.r] = ns;
                    if (mb == 1) st_next = ns;
                }
            }
            __syncthreads();
        }
        ++it;
        if (!done) iters = it;
    }
    __syncthreads();
    if (l.live)
        store_frame(c, a, l.frame, l.u, Z * P, iters, ok, app_out, bits_out, iters_out, valid_out);
}

// G groups of Z P lanes per frame: in the check phase group g updates block
// rows g, g + G, ... (flooding rows are independent within an iteration:
// each reads the previous iteration's APP and its own state), and the
// variable phase and the frame's store spread over all G Z P lanes.  The
// outputs are the same for every G.  small: the frames' messages (check
// states or sum-product messages), their channel LLRs and the column tables
// live in LDS (flood_lds) instead of HBM and the code's tables, so the
// variable phase's chain of table, state and LLR reads runs from LDS.  The
// syndrome check takes no pass of its own: the check phase reads every
// row's APPs, and the parity of their signs is the row's syndrome bit for
// the APP the previous iteration left (the one the stop decision is about).
// A variable's new APP in the flooding variable phase: l0 plus the messages
// of its checks (column edges t0 .. t1 - 1) in ascending check order.  B
// edges at a time: each edge's chain of table and message reads is issued
// before the first of their sums (with the tables and messages in HBM the
// chains overlap; from LDS, B = 1 measures faster).
constexpr int kVarBatch = 4;

template <bool SP, int B>
__device__ __forceinline__ float var_sum(float sum, int t0, int t1, int cc, int Z,
                                         const int* col_slot, const int* slot_shift,
                                         const int* slot_row, const int* slot_pos,
                                         const CheckState* st_base, const float* Rm) {
    for (int tb = t0; tb < t1; tb += B) {
        float rv[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int t = tb + u < t1 ? tb + u : t0;  // in the column
            const int s = col_slot[t];
            int rr = cc - slot_shift[s];
            rr += (rr < 0) ? Z : 0;
            if (SP) {
                rv[u] = Rm[static_cast<size_t>(s) * Z + rr];
            } else {
                const CheckState st = st_base[static_cast<size_t>(slot_row[s]) * Z + rr];
                rv[u] = state_msg(st, slot_pos[s]);
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (tb + u < t1) sum = sum + rv[u];
    }
    return sum;
}

template <bool SP, int P>
__global__ __launch_bounds__(kMaxQcThreads) void ldpc_qc_flooding(
    ck_ldpc_code c, const float* __restrict__ llr, float* __restrict__ app_out,
    signed char* __restrict__ bits_out, int* __restrict__ iters_out,
    unsigned char* __restrict__ valid_out, void* __restrict__ ws, long long R, int F, int G,
    int small, Params prm, SymbolSource sym) {
    extern __shared__ __attribute__((aligned(16))) float lds[];  // 16-byte check states
    __shared__ int s_flag[2][kMaxFrames];
    __shared__ int2 s_tab[kTabEntries];
    const int N = c.num_variables;
    const int Z = c.lifting;
    // Lanes of (frame f, group g) as those of a virtual frame f G + g.
    QcLane l = qc_lane<P>(Z, F * G, static_cast<long long>(F) * G * gridDim.x);
    const int g = l.f % G;
    l.f = l.f / G;
    l.frame = static_cast<long long>(blockIdx.x) * F + l.f;
    l.live = (l.f < F) && (l.frame < R);
    const int lanes = G * Z * P;      // lanes of one frame
    const int ug = g * Z * P + l.u;  // this lane among them
    float* a = lds + static_cast<size_t>(l.f < F ? l.f : 0) * N;
    stage_slot_table(c, s_tab);
    if (sym.x != nullptr) {  // demapping fused into the load; tables over the unused bitmap area
        load_frames_demap(lds, sym, lds + qc_lds(F, N, Z, c.num_checks, false).bits,
                          static_cast<long long>(blockIdx.x) * F, F, R, c.n, N);
    } else {
        load_frames(lds, llr, static_cast<long long>(blockIdx.x) * F, F, R, c.n, N);
    }
    __syncthreads();

    const bool lane_ok = l.f < F;
    const size_t fr = static_cast<size_t>(l.live ? l.frame : 0);
    const int nb = N / Z;
    const int fl = l.f < F ? l.f : 0;
    CheckState* st_base;
    float* Rm;
    const float* L;  // the frame's channel LLRs; LDS: N of them, -0.0 as +0.0 and zero-padded
    const int* col_ptr = c.qc_col_ptr;
    const int* col_slot = c.qc_col_slot;
    const int* slot_shift = c.qc_slot_shift;
    const int* slot_row = c.qc_slot_row;
    const int* slot_pos = c.qc_slot_pos;
    if (small) {
        const FloodLds fo = flood_lds(F, N, Z, c.num_checks, c.num_edges, nb, c.num_slots, SP);
        float* chan = lds + fo.chan;
        for (int i = threadIdx.x; i < F * N; i += blockDim.x) chan[i] = lds[i];  // the loaded APPs
        int* tabs = reinterpret_cast<int*>(lds + fo.tabs);
        const int S = c.num_slots;
        for (int i = threadIdx.x; i <= nb; i += blockDim.x) tabs[i] = c.qc_col_ptr[i];
        for (int i = threadIdx.x; i < S; i += blockDim.x) {
            tabs[nb + 1 + i] = c.qc_col_slot[i];
            tabs[nb + 1 + S + i] = c.qc_slot_shift[i];
            tabs[nb + 1 + 2 * S + i] = c.qc_slot_row[i];
            tabs[nb + 1 + 3 * S + i] = c.qc_slot_pos[i];
        }
        col_ptr = tabs;
        col_slot = tabs + nb + 1;
        slot_shift = col_slot + S;
        slot_row = slot_shift + S;
        slot_pos = slot_row + S;
        L = chan + static_cast<size_t>(fl) * N;
        st_base = SP ? nullptr : reinterpret_cast<CheckState*>(lds + fo.msgs) +
                                     static_cast<size_t>(fl) * c.num_checks;
        Rm = SP ? lds + fo.msgs + static_cast<size_t>(fl) * align4(c.num_edges) : nullptr;
        __syncthreads();
    } else {
        st_base = SP ? nullptr : static_cast<CheckState*>(ws) + fr * c.num_checks;
        Rm = SP ? static_cast<float*>(ws) + fr * c.num_edges : nullptr;
        L = llr + fr * c.n;
    }

    // s_flag[it & 1][f]: some row of frame f has odd hard-decision parity on
    // the APP the check phase of iteration `it` read.  Two buffers: one is
    // cleared for the next iteration while the other is read, so no barrier
    // separates the reads from the clearing.
    for (int i = threadIdx.x; i < 2 * kMaxFrames; i += blockDim.x)
        s_flag[i / kMaxFrames][i % kMaxFrames] = 0;
    __syncthreads();
    bool done = !l.live;
    bool ok = false;
    int iters = 0;
    int it = 0;
    for (;;) {
        const bool last = it >= prm.max_iterations;
        const bool first = it == 0;
        int* flag = s_flag[it & 1];
        // Check phase: every row reads the previous iteration's APP, so the
        // parities of its hard decisions make the syndrome of that APP, the
        // check before iteration `it` (the APP after `it` iterations).  At
        // the iteration limit the phase runs for the syndrome alone.
        if (!done) {
            for (int br = g; br < c.base_rows; br += G) {
                unsigned odd = 0u;
                if (SP) {
                    if (l.p == 0)
                        qc_sp_row<1, false, kMaxDegree, true>(c, s_tab, a, Rm, br, l.r, 0, first,
                                                              &odd);
                } else {
                    CheckState* sp = st_base + static_cast<size_t>(br) * Z + l.r;
                    const CheckState st = first ? zero_state() : *sp;
                    const CheckState ns = qc_ms_row<P, false, kMaxDegree / P, true>(
                        c, s_tab, a, st, br, l.r, l.p, prm, &odd);
                    if (l.p == 0 && !last) *sp = ns;
                }
                if (odd) flag[fl] = 1;
            }
        }
        __syncthreads();
        const bool clean = flag[fl] == 0;
        if (!done && (prm.early_stop || last)) {
            ok = clean;
            if (prm.early_stop && clean) done = true;
        }
        if (last) done = true;
        // Whether any frame of the workgroup runs on, the same on every lane:
        // a frame that has stopped, or holds no frame, raised no flag.
        bool more = !last;
        if (more && prm.early_stop) {
            more = false;
            for (int f = 0; f < F; ++f) more = more || flag[f] != 0;
        }
        if (!more) break;
        if (lane_ok && ug == 0) s_flag[(it + 1) & 1][fl] = 0;  // read last iteration
        // Variable phase: APP = L + the new R of every check, ascending check.
        if (!done) {
            for (int v = ug; v < N; v += lanes) {
                const int j = v / Z;
                const int cc = v - j * Z;
                // -0.0 as +0.0, as at the load (the LDS copy is the loaded APP)
                const float l0 = small ? L[v] : (v < c.n ? L[v] + 0.0f : 0.0f);
                const int t0 = col_ptr[j];
                const int t1 = col_ptr[j + 1];
                const float sum =
                    small ? var_sum<SP, 1>(l0, t0, t1, cc, Z, col_slot, slot_shift, slot_row,
                                           slot_pos, st_base, Rm)
                          : var_sum<SP, kVarBatch>(l0, t0, t1, cc, Z, col_slot, slot_shift,
                                                   slot_row, slot_pos, st_base, Rm);
                a[v] = sum;
            }
        }
        __syncthreads();
        ++it;
        if (!done) iters = it;
    }
    __syncthreads();
    if (l.live)
        store_frame(c, a, l.frame, ug, lanes, iters, ok, app_out, bits_out, iters_out, valid_out);
}

// ---------------------------------------------------------------------------
// Generic CSR kernel
// ---------------------------------------------------------------------------

// The APP vector of the generic kernel: variables below `split` in LDS, the
// rest in the frame's HBM workspace `glob` (ALL_LDS: all of it in LDS); a
// lane's access goes to one or the other (an LDS and a global instruction,
// each under its lanes' mask), so a frame too long for LDS keeps its
// lowest-numbered variables there (for the systematic codes here, the
// information bits: the columns of highest degree).
template <bool ALL_LDS>
struct GenApp {
    float* lds;
    float* glob;
    int split;
    // The split form names the address spaces, so each access is one LDS or
    // one global instruction rather than a flat access the compiler must
    // resolve at run time.
    using LdsF = __attribute__((address_space(3))) float;
    using GlbF = __attribute__((address_space(1))) float;
    __device__ __forceinline__ float ld(int v) const {
        if constexpr (ALL_LDS) return lds[v];
        float x;
        if (v < split)
            x = ((const LdsF*)lds)[v];
        else
            x = ((const GlbF*)glob)[v];
        return x;
    }
    __device__ __forceinline__ void st(int v, float x) const {
        if constexpr (ALL_LDS) {
            lds[v] = x;
        } else if (v < split) {
            ((LdsF*)lds)[v] = x;
        } else {
            ((GlbF*)glob)[v] = x;
        }
    }
};

// MAXD > 0: rows of at most MAXD edges, whose columns and values are loaded
// as one batch into registers (a row then costs one memory round trip for
// its columns and one for its values, instead of two per edge); MAXD = 0:
// any degree, edge by edge.
template <int MAXD, class App>
__device__ bool generic_syndrome_ok(const ck_ldpc_code& c, const App& A) {
    // Rounds of blockDim rows; the check stops at the first round with an
    // unsatisfied row (the frame is not a codeword), which while decoding
    // is usually the first.
    for (int base = 0; base < c.num_checks; base += blockDim.x) {
        const int row = base + threadIdx.x;
        int parity = 0;
        if (row < c.num_checks) {
            const int e0 = c.row_ptr[row];
            const int e1 = c.row_ptr[row + 1];
            if (MAXD > 0) {
                // unconditional loads from clamped addresses (see row_load_cols)
                constexpr int D = MAXD > 0 ? MAXD : 1;
                const int last = e1 > e0 ? e1 - 1 : 0;
                int col[D];
                float v[D];
#pragma unroll
                for (int i = 0; i < D; ++i) col[i] = c.cols[e0 + i < e1 ? e0 + i : last];
#pragma unroll
                for (int i = 0; i < D; ++i) v[i] = A.ld(col[i]);
#pragma unroll
                for (int i = 0; i < D; ++i) parity ^= (e0 + i < e1 && v[i] < 0.0f) ? 1 : 0;
            } else {
                for (int e = e0; e < e1; ++e) parity ^= A.ld(c.cols[e]) < 0.0f;
            }
        }
        if (__syncthreads_or(parity)) return false;
    }
    return true;
}

// A row of at most MAXD edges with its columns and Q values in registers:
// the arithmetic and its order are those of generic_row below (Q = APP - R
// read once; within a layer no other row touches the row's columns, so the
// second pass's APP equals the first's).  Loading and updating are split so
// that a lane can put the loads of two rows in flight before either update.
// Messages of slot i of the row at Rm[base + i * stride]: the row's CSR edges
// (base = first edge, stride 1; flooding, whose variable update reads them by
// edge id), or slot-major (base = schedule row, stride = m; layered), where
// the lanes of a wave, on consecutive rows, touch consecutive words.
template <int MAXD>
struct RowRegs {
    int base;
    int stride;
    int col[MAXD];
    float q[MAXD];
};

// The loads are unconditional, from clamped in-bounds addresses (the row's
// last edge, or edge / variable 0 for no row), and the padding slots are
// marked afterwards: loads under per-slot conditions would each be waited
// for before the next is issued.
template <int MAXD, bool SLOT_MAJOR>
__device__ __forceinline__ void row_load_cols(const ck_ldpc_code& c, int row, RowRegs<MAXD>& r) {
    if (SLOT_MAJOR && c.slot_cols != nullptr) {  // one coalesced load per slot
        r.base = row >= 0 ? row : 0;
        r.stride = c.num_checks;
#pragma unroll
        for (int i = 0; i < MAXD; ++i) {
            const int v = c.slot_cols[i * c.num_checks + r.base];
            r.col[i] = row >= 0 ? v : -1;
        }
        return;
    }
    const int e0 = row >= 0 ? c.row_ptr[row] : 0;
    const int e1 = row >= 0 ? c.row_ptr[row + 1] : 0;
    r.base = SLOT_MAJOR ? (row >= 0 ? row : 0) : e0;
    r.stride = SLOT_MAJOR ? c.num_checks : 1;
    const int last = e1 > e0 ? e1 - 1 : 0;
#pragma unroll
    for (int i = 0; i < MAXD; ++i) {
        const int v = c.cols[e0 + i < e1 ? e0 + i : last];
        r.col[i] = e0 + i < e1 ? v : -1;
    }
}

template <int MAXD, class App>
__device__ __forceinline__ void row_load_q(const App& A, const float* __restrict__ Rm,
                                           bool first, RowRegs<MAXD>& r) {
    float av[MAXD];
    float rv[MAXD];
#pragma unroll
    for (int i = 0; i < MAXD; ++i) {
        const bool in = r.col[i] >= 0;
        av[i] = A.ld(in ? r.col[i] : 0);
        rv[i] = first ? 0.0f : Rm[in ? r.base + i * r.stride : r.base];
    }
#pragma unroll
    for (int i = 0; i < MAXD; ++i) r.q[i] = av[i] - rv[i];
}

template <bool SP, bool UPDATE_APP, int MAXD, class App>
__device__ __forceinline__ void row_update(const App& A, float* __restrict__ Rm, const Params& prm,
                                           const RowRegs<MAXD>& r) {
    unsigned neg = 0u;
    if (SP) {
        float total = 0.0f;
#pragma unroll
        for (int i = 0; i < MAXD; ++i)
            if (r.col[i] >= 0) {
                neg ^= (r.q[i] < 0.0f) ? 1u : 0u;
                total = total + phi2_fast(fabsf(r.q[i]) * kLog2eF);
            }
#pragma unroll
        for (int i = 0; i < MAXD; ++i)
            if (r.col[i] >= 0) {
                const float mag = phi2_fast(total - phi2_fast(fabsf(r.q[i]) * kLog2eF)) * kLn2F;
                const float rn = (neg ^ ((r.q[i] < 0.0f) ? 1u : 0u)) ? -mag : mag;
                if (UPDATE_APP) A.st(r.col[i], r.q[i] + rn);
                Rm[r.base + i * r.stride] = rn;
            }
        return;
    }
    float min1 = ck::kInf;
    float min2 = ck::kInf;
    int idx = -1;
#pragma unroll
    for (int i = 0; i < MAXD; ++i)
        if (r.col[i] >= 0) {
            neg ^= (r.q[i] < 0.0f) ? 1u : 0u;
            const float mag = fabsf(r.q[i]);
            if (mag < min1) {
                min2 = min1;
                min1 = mag;
                idx = i;
            } else if (mag < min2) {
                min2 = mag;
            }
        }
    const float m1 = fmaxf(prm.scale * min1 - prm.offset, 0.0f);
    const float m2 = fmaxf(prm.scale * min2 - prm.offset, 0.0f);
#pragma unroll
    for (int i = 0; i < MAXD; ++i)
        if (r.col[i] >= 0) {
            const float mag = (i == idx) ? m2 : m1;
            const float rn = (neg ^ ((r.q[i] < 0.0f) ? 1u : 0u)) ? -mag : mag;
            if (UPDATE_APP) A.st(r.col[i], r.q[i] + rn);
            Rm[r.base + i * r.stride] = rn;
        }
}

template <bool SP, bool UPDATE_APP, int MAXD, bool SLOT_MAJOR, class App>
__device__ __forceinline__ void generic_row_reg(const ck_ldpc_code& c, const App& A,
                                                float* __restrict__ Rm, int row, bool first,
                                                const Params& prm) {
    RowRegs<MAXD> r;
    row_load_cols<MAXD, SLOT_MAJOR>(c, row, r);
    row_load_q(A, Rm, first, r);
    row_update<SP, UPDATE_APP>(A, Rm, prm, r);
}

// NR rows of one layer (a negative row: none), all their loads in flight
// before any update.
template <bool SP, bool UPDATE_APP, int MAXD, bool SLOT_MAJOR, int NR, class App>
__device__ __forceinline__ void generic_rows(const ck_ldpc_code& c, const App& A,
                                             float* __restrict__ Rm, const int (&rows)[NR],
                                             bool first, const Params& prm) {
    RowRegs<MAXD> r[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) row_load_cols<MAXD, SLOT_MAJOR>(c, rows[j], r[j]);
#pragma unroll
    for (int j = 0; j < NR; ++j) row_load_q(A, Rm, first, r[j]);
#pragma unroll
    for (int j = 0; j < NR; ++j)
        if (rows[j] >= 0) row_update<SP, UPDATE_APP>(A, Rm, prm, r[j]);
}

template <bool SP, bool UPDATE_APP, int MAXD, bool SLOT_MAJOR, class App>
__device__ __forceinline__ void generic_row(const ck_ldpc_code& c, const App& A,
                                            float* __restrict__ Rm, int row, bool first,
                                            const Params& prm) {
    if (MAXD > 0) {
        generic_row_reg<SP, UPDATE_APP, (MAXD > 0 ? MAXD : 1), SLOT_MAJOR>(c, A, Rm, row, first,
                                                                           prm);
        return;
    }
    const int e0 = c.row_ptr[row];
    const int e1 = c.row_ptr[row + 1];
    unsigned neg = 0u;
    if (SP) {
        float total = 0.0f;
        for (int e = e0; e < e1; ++e) {
            const float q = A.ld(c.cols[e]) - (first ? 0.0f : Rm[e]);
            neg ^= (q < 0.0f) ? 1u : 0u;
            total = total + phi2_fast(fabsf(q) * kLog2eF);
        }
        for (int e = e0; e < e1; ++e) {
            const int col = c.cols[e];
            const float q = A.ld(col) - (first ? 0.0f : Rm[e]);
            const float mag = phi2_fast(total - phi2_fast(fabsf(q) * kLog2eF)) * kLn2F;
            const float rn = (neg ^ ((q < 0.0f) ? 1u : 0u)) ? -mag : mag;
            if (UPDATE_APP) A.st(col, q + rn);
            Rm[e] = rn;
        }
        return;
    }
    float min1 = ck::kInf;
    float min2 = ck::kInf;
    int idx = -1;
    for (int e = e0; e < e1; ++e) {
        const float q = A.ld(c.cols[e]) - (first ? 0.0f : Rm[e]);
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
        const float q = A.ld(col) - (first ? 0.0f : Rm[e]);
        const float mag = (e == idx) ? m2 : m1;
        const float rn = (neg ^ ((q < 0.0f) ? 1u : 0u)) ? -mag : mag;
        if (UPDATE_APP) A.st(col, q + rn);
        Rm[e] = rn;
    }
}

template <bool SP, bool FLOOD, bool APP_LDS, int MAXD>
__global__ __launch_bounds__(kGenericThreads) void ldpc_generic(
    ck_ldpc_code c, const float* __restrict__ llr, float* __restrict__ app_out,
    signed char* __restrict__ bits_out, int* __restrict__ iters_out,
    unsigned char* __restrict__ valid_out, float* __restrict__ msg_ws,
    float* __restrict__ app_ws, long long R, Params prm, int lds_vars, long long msg_stride) {
    extern __shared__ float lds[];
    const long long frame = blockIdx.x;
    const int N = c.num_variables;
    float* glob = APP_LDS ? nullptr : app_ws + static_cast<size_t>(frame) * N;
    const GenApp<APP_LDS> A{lds, glob, APP_LDS ? N : lds_vars};
    float* Rm = msg_ws + static_cast<size_t>(frame) * msg_stride;
    const float* L = llr + frame * c.n;
    for (int v = threadIdx.x; v < N; v += blockDim.x) A.st(v, v < c.n ? L[v] : 0.0f);
    __syncthreads();

    int it = 0;
    bool ok = prm.early_stop && generic_syndrome_ok<MAXD>(c, A);
    while (!ok && it < prm.max_iterations) {
        const bool first = it == 0;
        if (FLOOD) {
            for (int row = threadIdx.x; row < c.num_checks; row += blockDim.x)
                generic_row<SP, false, MAXD, false>(c, A, Rm, row, first, prm);
            __syncthreads();
            for (int v = threadIdx.x; v < N; v += blockDim.x) {
                float sum = v < c.n ? L[v] + 0.0f : 0.0f;  // -0.0 as +0.0, as at the load
                for (int t = c.var_ptr[v]; t < c.var_ptr[v + 1]; ++t) sum = sum + Rm[c.var_edges[t]];
                A.st(v, sum);
            }
            __syncthreads();
        } else {
            for (int layer = 0; layer < c.num_layers; ++layer) {
                const int end = c.layer_ptr[layer + 1];
                if (MAXD > 0 && MAXD <= 16) {
                    // two rows to a lane (four measured slower: 2.3 against
                    // 2.0 ms on DVB-S2 rate 1/2, 64 frames)
                    constexpr int NR = 2;
                    const int step = static_cast<int>(blockDim.x);
                    for (int row = c.layer_ptr[layer] + threadIdx.x; row < end; row += NR * step) {
                        int rows[NR];
#pragma unroll
                        for (int j = 0; j < NR; ++j) rows[j] = row + j * step < end ? row + j * step : -1;
                        generic_rows<SP, true, (MAXD > 0 ? MAXD : 1), true, NR>(c, A, Rm, rows,
                                                                                first, prm);
                    }
                } else {
                    for (int row = c.layer_ptr[layer] + threadIdx.x; row < end; row += blockDim.x)
                        generic_row<SP, true, MAXD, (MAXD > 0)>(c, A, Rm, row, first, prm);
                }
                __syncthreads();
            }
        }
        ++it;
        if (prm.early_stop) ok = generic_syndrome_ok<MAXD>(c, A);
    }
    if (!prm.early_stop) ok = generic_syndrome_ok<MAXD>(c, A);
    if (!APP_LDS) {  // the LDS part back to the workspace, read from there
        for (int v = threadIdx.x; v < A.split; v += blockDim.x) glob[v] = lds[v];
    }
    __syncthreads();
    store_frame(c, APP_LDS ? lds : glob, frame, threadIdx.x, blockDim.x, it, ok, app_out,
                bits_out, iters_out, valid_out);
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
    int short_rows;    // QC layered min-sum, rows of at most kShortRow edges: the
                       // SHORT variant (1: 168 VGPRs, 2: 128 VGPRs), else 0
    int G;             // QC flooding: block rows updated at once per frame
    bool small;        // QC flooding: messages, channel LLRs and column tables in LDS
    int threads;
    size_t lds_bytes;
    size_t ws_per_frame;  // bytes
    size_t msg_floats;    // generic: messages per frame
};

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

namespace ck {

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

}  // namespace ck

namespace {

using ck::check_code;
using ck::check_method;

// Register batch width of the generic kernel for rows of up to d edges (0:
// edge by edge), the same rule launch_generic dispatches on.
int generic_maxd(int d) { return d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : 0; }

// Messages per frame of the generic kernel: slot-major (rows x the batch
// width) for the layered schedule with register batches, one per CSR edge
// otherwise.
size_t generic_msg_floats(const ck_ldpc_code& c, bool flood) {
    const int w = generic_maxd(c.max_row_degree);
    return (!flood && w > 0) ? static_cast<size_t>(w) * c.num_checks
                             : static_cast<size_t>(c.num_edges);
}

int32_t make_plan(const ck_ldpc_code* c, const ck_ldpc_method* m, long long R, Plan* plan) {
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
    // APP plus sign bitmaps of one frame.
    const size_t frame_bytes =
        Z > 0 ? qc_lds(1, N, Z, c->num_checks, false).end * 4 : app_bytes;
    bool qc_ok = Z > 0 && Z <= kMaxQcThreads && frame_bytes <= lds_max &&
                 c->max_row_degree <= kMaxDegree && c->num_slots <= kMaxSlots &&
                 syn_words(Z, N) < (1u << 21);
    if (m->kernel == 2 && !qc_ok)
        return ck::fail(CK_EINVAL,
                        "quasi-cyclic kernel requested but the code has no usable QC view "
                        "(lifting %d, max row degree %d, %d slots, %zu bytes per frame, LDS %zu bytes)",
                        Z, c->max_row_degree, c->num_slots, frame_bytes, lds_max);
    if (m->kernel == 1) qc_ok = false;
    if (qc_ok) {
        p.kind = 2;
        int P = m->lanes_per_row;
        if (P == 0) P = 1;
        if (p.sp && p.flood) P = 1;
        if (Z * P > kMaxQcThreads) P = 1;
        int F = m->frames_per_group;
        if (F == 0) {
            // Frames share a workgroup once there are two workgroups per CU
            // without sharing; more lanes per workgroup, fewer per-row barriers.
            const long long per = R / (2LL * info.num_cus);
            F = static_cast<int>(std::max(1LL, std::min<long long>(per, kMaxFrames)));
        }
        F = std::min(F, kMaxQcThreads / (Z * P));
        while (F > 1 && qc_lds(F, N, Z, c->num_checks, false).end * 4 > lds_max) --F;
        const size_t state_bytes = static_cast<size_t>(c->num_checks) * sizeof(CheckState);
        bool state_lds = false;
        if (!p.sp && !p.flood) {
            const size_t with = qc_lds(F, N, Z, c->num_checks, true).end * 4;
            const size_t without = qc_lds(F, N, Z, c->num_checks, false).end * 4;
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
        p.G = 1;
        p.small = false;
        if (p.flood) {  // as many of a frame's block rows at once as the workgroup holds
            p.G = std::max(1, std::min(c->base_rows, kMaxQcThreads / (F * Z * P)));
            const size_t small_bytes = flood_lds(F, N, Z, c->num_checks, c->num_edges, N / Z,
                                                 c->num_slots, p.sp).end * 4;
            p.small = small_bytes <= lds_max;
        }
        p.threads = static_cast<int>(round_up(static_cast<size_t>(F) * p.G * Z * P, 64));
        p.lds_bytes = qc_lds(F, N, Z, c->num_checks, state_lds).end * 4;
        p.short_rows = 0;
        if (!p.sp && !p.flood && P <= 2 && c->max_row_degree <= kShortRow) {
            // Two workgroups per CU when at most six waves each and LDS allows.
            const bool two = P == 1 && p.threads <= 384 && c->max_row_degree <= kShortRowTwo &&
                             2 * (p.lds_bytes + kStaticLds) <= info.lds_per_block;
            p.short_rows = two ? 2 : 1;
        }
        if (p.small)  // flooding with its messages in LDS: no workspace
            p.ws_per_frame = 0;
        else if (p.sp)
            p.ws_per_frame = static_cast<size_t>(c->num_edges) * 4;
        else
            p.ws_per_frame = state_lds ? 0 : state_bytes;
    } else {
        CK_REQUIRE(!p.flood || (c->var_ptr && c->var_edges),
                   "flooding needs the variable-to-edge view (var_ptr, var_edges)");
        p.kind = 1;
        p.app_lds = app_bytes <= lds_max;
        p.threads = kGenericThreads;
        // too long for LDS: its first lds_max / 4 variables there, the rest in HBM
        p.lds_bytes = p.app_lds ? app_bytes : (lds_max & ~static_cast<size_t>(255));
        p.msg_floats = generic_msg_floats(*c, p.flood);
        p.ws_per_frame = p.msg_floats * 4 + (p.app_lds ? 0 : app_bytes);
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

template <bool SP, int P, bool SL, int SHORT = 0>
void launch_qc_layered(const Plan& p, const ck_ldpc_code& c, const float* llr, float* app,
                       signed char* bits, int* iters, unsigned char* valid, void* ws, long long R,
                       const Params& prm, hipStream_t s, const SymbolSource& sym) {
    auto kernel = ldpc_qc_layered<SP, P, SL, SHORT>;
    allow_lds(kernel, p.lds_bytes);
    const long long groups = (R + p.F - 1) / p.F;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(groups)), dim3(p.threads), p.lds_bytes,
                       s, c, llr, app, bits, iters, valid, ws, R, p.F, prm, sym);
}

template <bool SP, int P>
void launch_qc_flooding(const Plan& p, const ck_ldpc_code& c, const float* llr, float* app,
                        signed char* bits, int* iters, unsigned char* valid, void* ws,
                        long long R, const Params& prm, hipStream_t s, const SymbolSource& sym) {
    auto kernel = ldpc_qc_flooding<SP, P>;
    const long long groups = (R + p.F - 1) / p.F;
    // p.lds_bytes covers the QC layout, and with fused demapping its tables.
    const size_t lds = p.small ? std::max<size_t>(flood_lds(p.F, c.num_variables, c.lifting,
                                                            c.num_checks, c.num_edges,
                                                            c.num_variables / c.lifting,
                                                            c.num_slots, SP).end * 4,
                                                  p.lds_bytes)
                               : p.lds_bytes;
    allow_lds(kernel, lds);
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(groups)), dim3(p.threads), lds, s, c,
                       llr, app, bits, iters, valid, ws, R, p.F, p.G, p.small ? 1 : 0, prm, sym);
}

template <bool SP, bool FLOOD, bool APP_LDS, int MAXD>
void launch_generic_d(const Plan& p, const ck_ldpc_code& c, const float* llr, float* app,
                      signed char* bits, int* iters, unsigned char* valid, void* ws, long long R,
                      const Params& prm, hipStream_t s) {
    auto kernel = ldpc_generic<SP, FLOOD, APP_LDS, MAXD>;
    allow_lds(kernel, p.lds_bytes);
    float* msg = static_cast<float*>(ws);
    float* app_ws = APP_LDS ? nullptr : msg + static_cast<size_t>(R) * p.msg_floats;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(R)), dim3(p.threads), p.lds_bytes, s, c,
                       llr, app, bits, iters, valid, msg, app_ws, R, prm,
                       static_cast<int>(p.lds_bytes / 4), static_cast<long long>(p.msg_floats));
}

// Rows of up to 8, 16 or 32 edges take the register batch; wider rows the
// edge-by-edge loop.
template <bool SP, bool FLOOD, bool APP_LDS>
void launch_generic(const Plan& p, const ck_ldpc_code& c, const float* llr, float* app,
                    signed char* bits, int* iters, unsigned char* valid, void* ws, long long R,
                    const Params& prm, hipStream_t s) {
    const int d = generic_maxd(c.max_row_degree);
    if (d == 8) launch_generic_d<SP, FLOOD, APP_LDS, 8>(p, c, llr, app, bits, iters, valid, ws, R, prm, s);
    else if (d == 16) launch_generic_d<SP, FLOOD, APP_LDS, 16>(p, c, llr, app, bits, iters, valid, ws, R, prm, s);
    else if (d == 32) launch_generic_d<SP, FLOOD, APP_LDS, 32>(p, c, llr, app, bits, iters, valid, ws, R, prm, s);
    else launch_generic_d<SP, FLOOD, APP_LDS, 0>(p, c, llr, app, bits, iters, valid, ws, R, prm, s);
}

}  // namespace

extern "C" int64_t ck_ldpc_workspace_bytes(const ck_ldpc_code* code, const ck_ldpc_method* method,
                                           int64_t R) {
    Plan p{};
    if (R < 0 || make_plan(code, method, R, &p) != CK_OK) return -1;
    return static_cast<int64_t>(round_up(p.ws_per_frame * static_cast<size_t>(R), 256));
}

// The decode of ck_ldpc_decode; with sym.x set (ck_ldpc_decode_symbols) the
// frames come as symbols, demapped in the quasi-cyclic kernel's load.
static int32_t decode_impl(const ck_ldpc_code* code, const ck_ldpc_method* method,
                           const float* llrs, const SymbolSource& sym, float* app, int8_t* bits,
                           int32_t* iterations, uint8_t* valid, int64_t R, void* workspace,
                           void* stream) {
    Plan p{};
    if (int32_t st = make_plan(code, method, R, &p)) return st;
    CK_REQUIRE(R >= 0, "R = %lld < 0", static_cast<long long>(R));
    if (R == 0) return CK_OK;
    CK_REQUIRE(R <= 0x7fffffffLL, "R = %lld exceeds one launch", static_cast<long long>(R));
    CK_REQUIRE((llrs || sym.x) && app && iterations && valid, "NULL input or output");
    CK_REQUIRE(sym.x == nullptr || p.kind == 2,
               "demapping fused into the decode needs a quasi-cyclic kernel (layered or flooding)");
    // The flooding kernel's HBM form reads the channel LLRs from the input
    // in every variable phase; with symbols in there is no such buffer, so
    // only its LDS form (channel LLRs copied to LDS after the load) fuses.
    CK_REQUIRE(sym.x == nullptr || !p.flood || p.small,
               "fused demapping with the flooding schedule needs a code whose channel LLRs, "
               "messages and column tables fit in LDS beside its APPs");
    if (sym.x != nullptr) {
        // The demapping tables go past the frames' APPs, over the sign
        // bitmaps and check states (unused until the load ends).
        ck::DeviceInfo info{};
        if (int32_t st = ck::device_info(&info)) return st;
        const size_t need = qc_lds(p.F, code->num_variables, code->lifting, code->num_checks,
                                   p.state_lds).bits * 4 +
                            demap_table_bytes(sym.re.num_bits);
        CK_REQUIRE(need + kStaticLds + 1024 <= info.lds_per_block,
                   "fused demapping: %zu bytes of LDS per workgroup exceed the device's %zu", need,
                   info.lds_per_block);
        p.lds_bytes = std::max(p.lds_bytes, need);
        if (p.short_rows == 2 && 2 * (p.lds_bytes + kStaticLds) > info.lds_per_block)
            p.short_rows = 1;  // one workgroup per CU
    }
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
            if (p.sp) launch_qc_flooding<true, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 1) launch_qc_flooding<false, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 2) launch_qc_flooding<false, 2>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else launch_qc_flooding<false, 4>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
        } else if (p.sp) {
            if (p.P == 1) launch_qc_layered<true, 1, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 2) launch_qc_layered<true, 2, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else launch_qc_layered<true, 4, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
        } else if (p.state_lds) {
            if (p.short_rows == 2 && p.P == 1) launch_qc_layered<false, 1, true, 2>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.short_rows == 1 && p.P == 1) launch_qc_layered<false, 1, true, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 1) launch_qc_layered<false, 1, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 2 && p.short_rows) launch_qc_layered<false, 2, true, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 2) launch_qc_layered<false, 2, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else launch_qc_layered<false, 4, true>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
        } else {
            if (p.short_rows == 2 && p.P == 1) launch_qc_layered<false, 1, false, 2>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.short_rows == 1 && p.P == 1) launch_qc_layered<false, 1, false, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 1) launch_qc_layered<false, 1, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 2 && p.short_rows) launch_qc_layered<false, 2, false, 1>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else if (p.P == 2) launch_qc_layered<false, 2, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
            else launch_qc_layered<false, 4, false>(p, c, llrs, app, b, it, v, workspace, R, prm, s, sym);
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

// ---------------------------------------------------------------------------
// Decoding without allocation: the workspace is a fixed device array
// ---------------------------------------------------------------------------

namespace {

// Bytes of the static scratch: 64 MiB, held by every device the library runs
// on until the process ends.
constexpr size_t kScratchBytes = static_cast<size_t>(64) << 20;
constexpr int kScratchDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch[kScratchDevices] = {};

// The device address of `scratch` on the current device, resolved once per
// device (the array's name is a host-side stand-in, not a device address).
int32_t scratch_address(void** out) {
    int dev = 0;
    CK_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kScratchDevices)
        return ck::fail(CK_ENOSPACE, "device index %d beyond %d", dev, kScratchDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        CK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (p == nullptr) return ck::fail(CK_EHIP, "hipGetSymbolAddress(scratch) gave NULL");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return CK_OK;
}

}  // namespace


extern "C" int32_t ck_ldpc_decode(const ck_ldpc_code* code, const ck_ldpc_method* method,
                                  const float* llrs, float* app, int8_t* bits,
                                  int32_t* iterations, uint8_t* valid, int64_t R, void* workspace,
                                  void* stream) {
    const SymbolSource none{};
    return decode_impl(code, method, llrs, none, app, bits, iterations, valid, R, workspace,
                       stream);
}

extern "C" int32_t ck_ldpc_decode_symbols(const ck_ldpc_code* code, const ck_ldpc_method* method,
                                          const void* symbols, const ck_axis* re,
                                          const ck_axis* im, double inv_s2,
                                          const double* inv_rows, float* app, int8_t* bits,
                                          int32_t* iterations, uint8_t* valid, int64_t R,
                                          void* workspace, void* stream) {
    CK_REQUIRE(code != nullptr && re != nullptr && im != nullptr, "NULL code or axis");
    CK_REQUIRE(re->num_bits == im->num_bits && re->num_bits >= 1 &&
                   re->num_bits <= kDemapMaxBits,
               "fused demapping: axes of 1..%d bits each, alike (%d, %d)", kDemapMaxBits,
               re->num_bits, im->num_bits);
    for (const ck_axis* a : {re, im})
        CK_REQUIRE(a->num_levels == (1 << a->num_bits) && a->levels && a->labels && a->lo &&
                       a->hi && a->pos,
                   "fused demapping: closed-form axis tables (levels, labels, opposite levels, "
                   "positions) required");
    const int k = 2 * re->num_bits;
    CK_REQUIRE(code->n % k == 0, "fused demapping: n = %d is not a whole number of %d-bit symbols",
               code->n, k);
    CK_REQUIRE(inv_rows != nullptr || (inv_s2 > 0.0 && inv_s2 < 1e300),
               "inv_s2 must be positive and finite");
    CK_REQUIRE(R == 0 || symbols != nullptr, "NULL symbols");
    SymbolSource sym{};
    sym.x = static_cast<const float2*>(symbols);
    sym.re = *re;
    sym.im = *im;
    sym.inv_s2 = inv_s2;
    sym.inv_rows = inv_rows;
    if (R == 0) sym.x = nullptr;
    return decode_impl(code, method, nullptr, sym, app, bits, iterations, valid, R, workspace,
                       stream);
}

extern "C" int32_t ck_ldpc_decode_static(const ck_ldpc_code* code, const ck_ldpc_method* method,
                                         const float* llrs, float* app, int8_t* bits,
                                         int32_t* iterations, uint8_t* valid, int64_t R,
                                         void* stream) {
    CK_REQUIRE(R >= 0, "R = %lld < 0", static_cast<long long>(R));
    Plan p{};
    if (int32_t st = make_plan(code, method, R, &p)) return st;
    if (R == 0) return CK_OK;
    if (p.ws_per_frame == 0)  // the check state lives in LDS: no workspace at all
        return ck_ldpc_decode(code, method, llrs, app, bits, iterations, valid, R, nullptr, stream);
    CK_REQUIRE(p.ws_per_frame <= kScratchBytes,
               "one frame needs %zu bytes of workspace; the static scratch holds %zu",
               p.ws_per_frame, kScratchBytes);
    CK_REQUIRE(llrs && app && iterations && valid, "NULL input or output");
    void* ws = nullptr;
    if (int32_t st = scratch_address(&ws)) return st;
    // Frames in chunks that fit the scratch, one launch after the other on
    // `stream`: a chunk's plan never needs more workspace per frame than the
    // whole batch's (the check state leaves LDS only for more frames per
    // workgroup), which the check below confirms before each launch.
    const long long chunk = static_cast<long long>(kScratchBytes / p.ws_per_frame);
    const ck_ldpc_code& c = *code;
    for (long long r0 = 0; r0 < R; r0 += chunk) {
        const long long rc = std::min<long long>(chunk, R - r0);
        const int64_t need = ck_ldpc_workspace_bytes(code, method, rc);
        CK_REQUIRE(need >= 0 && static_cast<size_t>(need) <= kScratchBytes,
                   "a chunk of %lld frames needs %lld bytes of workspace; the static scratch "
                   "holds %zu", rc, static_cast<long long>(need), kScratchBytes);
        if (int32_t st = ck_ldpc_decode(code, method, llrs + r0 * c.n, app + r0 * c.n,
                                        bits ? bits + r0 * c.k : nullptr, iterations + r0,
                                        valid + r0, rc, need > 0 ? ws : nullptr, stream))
            return st;
    }
    return CK_OK;
}
