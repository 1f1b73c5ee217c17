// Persistent batched QC-LDPC decoding: frames of several codes, uneven in
// length and rate, decoded by one launch with per-frame early exit.
//
// The frames are grouped into classes (one code per class, its frames
// contiguous).  The launch is a persistent grid, one workgroup per CU (or as
// many as LDS allows).  A workgroup works through the classes in order; for
// class c it holds F_c frame slots (the APP of each in LDS, its check state
// in LDS too when it fits, else in a per-slot workspace) and decodes them
// with the layered row updates of
// the per-call QC kernel, lane (f, r, p) owning row r of slot f.  Each slot
// takes its next frame from the class's queue (one atomic counter per class
// in the workspace): at every iteration boundary a slot whose frame has
// finished (syndrome satisfied, or the iteration limit reached) writes that
// frame's outputs and is refilled, so a frame that converges early frees its
// lanes for the next one instead of idling until its neighbours finish.  A
// workgroup moves to the next class once the queue is empty and its slots
// have drained; classes are ordered by cost (largest frames first) by the
// caller, so workgroups finishing early help with the tail.
//
// Every frame's arithmetic is that of the per-call decoder: the same row
// updates in the same order, states starting at zero, the same syndrome
// checks.  Results are identical to decoding each frame alone.
//
// The launch reads only fixed device buffers (the descriptor block written
// once by ck_batch_prepare, the frames' inputs and outputs, the slot
// workspace), so it can be captured in a HIP graph and replayed; the queue
// counters are cleared by a small kernel launched before it (no memset
// node).  Every workgroup exits once every queue is empty and its slots are
// idle: no workgroup waits on another.
#include "ck_batch_int.h"
#include "ck_common.h"
#include "ck_ldpc_dev.h"

#include <algorithm>
#include <climits>
#include <cstring>

namespace {

using namespace ckdev;

using ckbatch::kMaxClasses;
constexpr int kBatchThreads = 768;

struct BatchClass {
    ck_ldpc_code code;
    const float* llrs;
    float* app;
    signed char* bits;
    int* iters;
    unsigned char* valid;
    long long num_frames;
    int F;
    int P;
    int state_lds;  // min-sum check states in LDS (else in the slot workspace)
};

struct alignas(256) BatchDesc {
    unsigned int counters[kMaxClasses];  // next frame of each class
    int num_classes;
    int sum_product;
    Params prm;
    long long slot_bytes;  // workspace per workgroup (all its slots)
    char* slots;           // per-workgroup slot workspace
    BatchClass cls[kMaxClasses];
};

__global__ void batch_reset(BatchDesc* d) {
    if (threadIdx.x < kMaxClasses) d->counters[threadIdx.x] = 0u;
}

// Hands each slot flagged in s_need its class's next frame (or none when the
// queue is empty), loads the frame's LLRs, and with early stop retires at
// once the frames already satisfied (0 iterations), asking again.
template <int P>
__device__ __forceinline__ void refill(const BatchClass& bc, unsigned* counter, const ck_ldpc_code& c,
                       const int2* tab, float* lds, unsigned* bits, float* a, const QcLane& l,
                       bool lane_ok, int F,
                       const Params& prm, int* s_need, int* s_frame, int* s_flag, bool& active,
                       int& my_it, long long& frame) {
    const int Z = c.lifting;
    const int N = c.num_variables;
    const int n = c.n;
    while (true) {
        __syncthreads();
        if (threadIdx.x < F && s_need[threadIdx.x]) {
            const unsigned f = atomicAdd(counter, 1u);
            s_frame[threadIdx.x] = (static_cast<long long>(f) < bc.num_frames) ? static_cast<int>(f) : -1;
        }
        __syncthreads();
        const bool fresh = lane_ok && s_need[l.f] && s_frame[l.f] >= 0;
        if (fresh) {
            frame = s_frame[l.f];
            load_frame(a, bc.llrs + frame * n, n, N, l.u, Z * P);
        }
        __syncthreads();
        if (lane_ok && l.u == 0) s_need[l.f] = 0;
        bool done0 = false;
        if (prm.early_stop) {
            const bool ok = qc_syndrome_ok(c, tab, lds, bits, l, lane_ok, fresh, F, s_flag);
            done0 = fresh && ok;
            if (done0) store_frame<false>(c, a, frame, l.u, Z * P, 0, true, bc.app, bc.bits, bc.iters, bc.valid);
        }
        if (fresh && !done0) {
            active = true;
            my_it = 0;
        }
        if (done0 && l.u == 0) s_need[l.f] = 1;
        if (!__syncthreads_or(done0)) break;
    }
}

template <bool SP, int P, bool STATE_LDS, int MAXT>
__device__ __forceinline__ void run_class(BatchDesc* __restrict__ d, int ci, float* lds, int2* s_tab, int* s_need,
                          int* s_frame, int* s_flag) {
    const BatchClass bc = d->cls[ci];
    const ck_ldpc_code c = bc.code;
    const Params prm = d->prm;
    const int Z = c.lifting;
    const int N = c.num_variables;
    const int mb = c.base_rows;
    const int F = bc.F;
    unsigned* counter = &d->counters[ci];
    const QcLane l = qc_lane<P>(Z, F, LLONG_MAX);
    const bool lane_ok = l.f < F;
    const int slot = lane_ok ? l.f : 0;
    float* a = lds + static_cast<size_t>(slot) * N;
    unsigned* bits = reinterpret_cast<unsigned*>(lds + qc_lds(F, N, Z, c.num_checks, false).bits);
    char* ws = d->slots + static_cast<size_t>(blockIdx.x) * d->slot_bytes;
    CheckState* st_base =
        (STATE_LDS ? reinterpret_cast<CheckState*>(lds + qc_lds(F, N, Z, c.num_checks, true).state)
                   : reinterpret_cast<CheckState*>(ws)) +
        static_cast<size_t>(slot) * c.num_checks;
    float* Rm = reinterpret_cast<float*>(ws) + static_cast<size_t>(slot) * c.num_edges;

    __syncthreads();  // the previous class is done with LDS
    stage_slot_table(c, s_tab);
    if (threadIdx.x < F) s_need[threadIdx.x] = 1;
    bool active = false;
    int my_it = 0;
    long long frame = -1;
    refill<P>(bc, counter, c, s_tab, lds, bits, a, l, lane_ok, F, prm, s_need, s_frame, s_flag, active, my_it,
              frame);
    CheckState st_next = zero_state();
    while (__syncthreads_or(active)) {
        for (int br = 0; br < mb; ++br) {
            if (active) {
                const bool first = my_it == 0;
                if (SP) {
                    qc_sp_row<P, true, MAXT>(c, s_tab, a, Rm, br, l.r, l.p, first);
                } else {
                    const CheckState st = first ? zero_state() : st_next;
                    if (br + 1 < mb) {
                        st_next = first ? zero_state() : st_base[static_cast<size_t>(br + 1) * Z + l.r];
                    } else if (mb > 1) {
                        st_next = st_base[l.r];  // block row 0, stored this iteration
                    }
                    const CheckState ns = qc_ms_row<P, true, MAXT>(c, s_tab, a, st, br, l.r, l.p, prm);
                    if (l.p == 0) st_base[static_cast<size_t>(br) * Z + l.r] = ns;
                    if (mb == 1) st_next = ns;
                }
            }
            __syncthreads();
        }
        if (active) ++my_it;
        const bool check = active && (prm.early_stop || my_it >= prm.max_iterations);
        const bool ok = qc_syndrome_ok(c, s_tab, lds, bits, l, lane_ok, check, F, s_flag);
        const bool fin = active && ((prm.early_stop && ok) || my_it >= prm.max_iterations);
        if (fin) {
            store_frame<false>(c, a, frame, l.u, Z * P, my_it, ok, bc.app, bc.bits, bc.iters, bc.valid);
            active = false;
            if (l.u == 0) s_need[l.f] = 1;
        }
        refill<P>(bc, counter, c, s_tab, lds, bits, a, l, lane_ok, F, prm, s_need, s_frame, s_flag, active,
                  my_it, frame);
    }
}

// ONE: whether some class runs one lane per check row (0: none; 1: all such
// classes have rows of at most kShortRow edges; 2: longer rows).  A lane of
// a one-lane row holds all its slots in registers, which costs the kernel
// registers (spills) that two-lane classes should not pay, so each case is
// its own kernel.  TWO: the slots a lane of a two-lane row holds, kMaxDegree
// / 2 or kTwoShortSlots (rows of at most 28 edges): the longer row
// instantiations alone cost spills (the mixed benchmark's launch: 5.53 ms
// with 16 slots, 5.07 ms with 14).  Everything is inlined: calls would move
// the live registers to scratch memory around every row and syndrome check.
constexpr int kTwoShortSlots = 14;

template <bool SP, int ONE, int TWO = kMaxDegree / 2>
__global__ __launch_bounds__(kBatchThreads) void ldpc_batch(BatchDesc* __restrict__ d) {
    constexpr int kOneLaneSlots = ONE == 1 ? kShortRow : kMaxDegree;
    extern __shared__ float lds[];
    __shared__ int2 s_tab[kTabEntries];
    __shared__ int s_need[kMaxFrames];
    __shared__ int s_frame[kMaxFrames];
    __shared__ int s_flag[kMaxFrames];
    const int nc = d->num_classes;
    for (int ci = 0; ci < nc; ++ci) {
        if (d->cls[ci].P == 2) {
            if (!SP && d->cls[ci].state_lds)
                run_class<SP, 2, true, TWO>(d, ci, lds, s_tab, s_need, s_frame, s_flag);
            else
                run_class<SP, 2, false, TWO>(d, ci, lds, s_tab, s_need, s_frame, s_flag);
        } else if constexpr (ONE != 0) {
            if (!SP && d->cls[ci].state_lds)
                run_class<SP, 1, true, kOneLaneSlots>(d, ci, lds, s_tab, s_need, s_frame, s_flag);
            else
                run_class<SP, 1, false, kOneLaneSlots>(d, ci, lds, s_tab, s_need, s_frame, s_flag);
        }
    }
}

struct Layout {
    int grid;
    size_t lds_bytes;
    size_t slot_bytes;
    int F[kMaxClasses];
    int P[kMaxClasses];
    int S[kMaxClasses];  // check states in LDS
    int variant;         // kernel: one-lane rows (0 none, 1 short, 2 long); 3: none,
                         // and two-lane rows of at most 2 kTwoShortSlots edges
};

int32_t make_layout(const ck_batch_class* classes, int32_t num_classes, const ck_ldpc_method* m,
                    Layout* out) {
    CK_REQUIRE(classes != nullptr, "classes is NULL");
    CK_REQUIRE(num_classes >= 1 && num_classes <= kMaxClasses, "num_classes %d not in [1, %d]",
               num_classes, kMaxClasses);
    if (int32_t st = ck::check_method(m)) return st;
    CK_REQUIRE(m->schedule == 0, "batched decoding runs the layered schedule");
    ck::DeviceInfo info{};
    if (int32_t st = ck::device_info(&info)) return st;
    const size_t lds_max =
        info.lds_per_block > 8 * kTabEntries + 12 * kMaxFrames + 2048
            ? info.lds_per_block - 8 * kTabEntries - 12 * kMaxFrames - 1024
            : 0;
    const bool sp = m->rule == 1;
    Layout L{};
    for (int i = 0; i < num_classes; ++i) {
        const ck_batch_class& bc = classes[i];
        CK_REQUIRE(bc.code != nullptr, "class %d: code is NULL", i);
        if (int32_t st = ck::check_code(bc.code)) return st;
        const ck_ldpc_code& c = *bc.code;
        CK_REQUIRE(c.lifting > 0, "class %d: batched decoding needs a quasi-cyclic code", i);
        CK_REQUIRE(c.max_row_degree <= kMaxDegree, "class %d: row degree %d above %d", i,
                   c.max_row_degree, kMaxDegree);
        CK_REQUIRE(c.num_slots <= kMaxSlots, "class %d: %d circulants above %d", i, c.num_slots,
                   kMaxSlots);
        CK_REQUIRE(bc.num_frames >= 0 && bc.num_frames < (1LL << 31), "class %d: %lld frames", i,
                   static_cast<long long>(bc.num_frames));
        CK_REQUIRE(bc.num_frames == 0 || (bc.llrs && bc.app && bc.iterations && bc.valid),
                   "class %d: NULL input or output", i);
        CK_REQUIRE(syn_words(c.lifting, c.num_variables) < (1u << 21),
                   "class %d: %d block columns of lifting %d are too many", i,
                   c.num_variables / c.lifting, c.lifting);
        // APP plus sign bitmaps of one frame.
        const size_t app_bytes = qc_lds(1, c.num_variables, c.lifting, c.num_checks, false).end * 4;
        int P = bc.lanes_per_row;
        if (P == 0) P = (2 * c.lifting <= kBatchThreads) ? 2 : 1;  // twice the lanes per frame
        CK_REQUIRE(P == 1 || P == 2, "class %d: lanes_per_row %d not 1 or 2", i, P);
        CK_REQUIRE(c.lifting * P <= kBatchThreads, "class %d: lifting %d x %d lanes above %d", i,
                   c.lifting, P, kBatchThreads);
        int F = bc.frames_per_group;
        if (F == 0) F = kBatchThreads / (c.lifting * P);
        F = std::min({F, kBatchThreads / (c.lifting * P), kMaxFrames,
                      static_cast<int>(lds_max / app_bytes)});
        CK_REQUIRE(F >= 1, "class %d: a frame (%zu bytes of APP and bitmaps) does not fit in LDS",
                   i, app_bytes);
        while (F > 1 && qc_lds(F, c.num_variables, c.lifting, c.num_checks, false).end * 4 > lds_max)
            --F;
        // Min-sum check states in LDS when they fit beside the slots' APPs.
        const bool st_lds =
            !sp && qc_lds(F, c.num_variables, c.lifting, c.num_checks, true).end * 4 <= lds_max;
        L.F[i] = F;
        L.P[i] = P;
        L.S[i] = st_lds ? 1 : 0;
        L.lds_bytes = std::max(L.lds_bytes,
                               qc_lds(F, c.num_variables, c.lifting, c.num_checks, st_lds).end * 4);
        const size_t per_slot = sp ? static_cast<size_t>(c.num_edges) * 4
                                   : (st_lds ? 0 : static_cast<size_t>(c.num_checks) * sizeof(CheckState));
        L.slot_bytes = std::max(L.slot_bytes, static_cast<size_t>(F) * per_slot);
    }
    L.slot_bytes = (L.slot_bytes + 255) / 256 * 256;
    L.variant = 0;
    bool two_short = true;
    for (int i = 0; i < num_classes; ++i) {
        if (L.P[i] == 1)
            L.variant = std::max(L.variant, classes[i].code->max_row_degree <= kShortRow ? 1 : 2);
        else
            two_short = two_short && classes[i].code->max_row_degree <= 2 * kTwoShortSlots;
    }
    if (L.variant == 0 && two_short) L.variant = 3;
    // One workgroup of twelve waves per CU: its registers (up to 168 VGPRs
    // a lane) leave no room for a second.
    L.grid = info.num_cus;
    *out = L;
    return CK_OK;
}

}  // namespace

namespace ckbatch {

size_t desc_bytes() { return sizeof(BatchDesc); }

int32_t shape(const ck_batch_class* classes, int32_t num_classes, const ck_ldpc_method* method,
              Shape* out) {
    Layout L{};
    if (int32_t st = make_layout(classes, num_classes, method, &L)) return st;
    out->grid = L.grid;
    out->lds_bytes = L.lds_bytes;
    out->slot_bytes = L.slot_bytes;
    for (int i = 0; i < kMaxClasses; ++i) {
        out->F[i] = L.F[i];
        out->P[i] = L.P[i];
        out->S[i] = L.S[i];
    }
    out->variant = L.variant;
    return CK_OK;
}

void describe(const ck_batch_class* classes, int32_t num_classes, const ck_ldpc_method* method,
              const Shape& L, void* workspace, void* host_desc) {
    BatchDesc& d = *static_cast<BatchDesc*>(host_desc);
    std::memset(&d, 0, sizeof(d));
    d.num_classes = num_classes;
    d.sum_product = method->rule == 1;
    d.prm = Params{method->rule == 1 ? 1.0f : method->scale, method->rule == 1 ? 0.0f : method->offset,
                   method->max_iterations, method->early_stop != 0};
    d.slot_bytes = static_cast<long long>(L.slot_bytes);
    d.slots = static_cast<char*>(workspace) + sizeof(BatchDesc);
    for (int i = 0; i < num_classes; ++i) {
        const ck_batch_class& bc = classes[i];
        BatchClass& o = d.cls[i];
        o.code = *bc.code;
        o.llrs = bc.llrs;
        o.app = bc.app;
        o.bits = reinterpret_cast<signed char*>(bc.bits);
        o.iters = bc.iterations;
        o.valid = bc.valid;
        o.num_frames = bc.num_frames;
        o.F = L.F[i];
        o.P = L.P[i];
        o.state_lds = L.S[i];
    }
}

void allow_lds(size_t lds_bytes) {
    if (lds_bytes > 65536) {
        const void* kernels[] = {
            reinterpret_cast<const void*>(ldpc_batch<false, 0>),
            reinterpret_cast<const void*>(ldpc_batch<false, 1>),
            reinterpret_cast<const void*>(ldpc_batch<false, 2>),
            reinterpret_cast<const void*>(ldpc_batch<false, 0, kTwoShortSlots>),
            reinterpret_cast<const void*>(ldpc_batch<true, 0>),
            reinterpret_cast<const void*>(ldpc_batch<true, 1>),
            reinterpret_cast<const void*>(ldpc_batch<true, 2>),
            reinterpret_cast<const void*>(ldpc_batch<true, 0, kTwoShortSlots>)};
        for (const void* k : kernels)
            (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds_bytes));
        (void)hipGetLastError();
    }
}

int32_t launch(void* workspace, int grid, size_t lds_bytes, bool sum_product, int variant,
               hipStream_t s) {
    CK_REQUIRE(variant >= 0 && variant <= 3, "kernel variant %d", variant);
    BatchDesc* d = static_cast<BatchDesc*>(workspace);
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, s, d);
    if (int32_t st = ck::launched("batch reset")) return st;
    const dim3 g(static_cast<unsigned>(grid)), b(kBatchThreads);
    if (sum_product) {
        if (variant == 0) hipLaunchKernelGGL((ldpc_batch<true, 0>), g, b, lds_bytes, s, d);
        else if (variant == 1) hipLaunchKernelGGL((ldpc_batch<true, 1>), g, b, lds_bytes, s, d);
        else if (variant == 2) hipLaunchKernelGGL((ldpc_batch<true, 2>), g, b, lds_bytes, s, d);
        else hipLaunchKernelGGL((ldpc_batch<true, 0, kTwoShortSlots>), g, b, lds_bytes, s, d);
    } else {
        if (variant == 0) hipLaunchKernelGGL((ldpc_batch<false, 0>), g, b, lds_bytes, s, d);
        else if (variant == 1) hipLaunchKernelGGL((ldpc_batch<false, 1>), g, b, lds_bytes, s, d);
        else if (variant == 2) hipLaunchKernelGGL((ldpc_batch<false, 2>), g, b, lds_bytes, s, d);
        else hipLaunchKernelGGL((ldpc_batch<false, 0, kTwoShortSlots>), g, b, lds_bytes, s, d);
    }
    return ck::launched("batch decode");
}

}  // namespace ckbatch

extern "C" int64_t ck_batch_workspace_bytes(const ck_batch_class* classes, int32_t num_classes,
                                            const ck_ldpc_method* method) {
    Layout L{};
    if (make_layout(classes, num_classes, method, &L) != CK_OK) return -1;
    return static_cast<int64_t>(sizeof(BatchDesc) + L.slot_bytes * static_cast<size_t>(L.grid));
}

extern "C" int32_t ck_batch_prepare(const ck_batch_class* classes, int32_t num_classes,
                                    const ck_ldpc_method* method, void* workspace,
                                    int64_t workspace_bytes, ck_batch_plan* plan) {
    Layout L{};
    if (int32_t st = make_layout(classes, num_classes, method, &L)) return st;
    CK_REQUIRE(plan != nullptr, "plan is NULL");
    const size_t need = sizeof(BatchDesc) + L.slot_bytes * static_cast<size_t>(L.grid);
    CK_REQUIRE(workspace != nullptr && workspace_bytes >= static_cast<int64_t>(need),
               "workspace of %lld bytes, %zu needed", static_cast<long long>(workspace_bytes), need);
    CK_REQUIRE((reinterpret_cast<uintptr_t>(workspace) & 255) == 0, "workspace not 256-byte aligned");
    ckbatch::Shape shp{};
    shp.grid = L.grid;
    shp.lds_bytes = L.lds_bytes;
    shp.slot_bytes = L.slot_bytes;
    for (int i = 0; i < kMaxClasses; ++i) {
        shp.F[i] = L.F[i];
        shp.P[i] = L.P[i];
        shp.S[i] = L.S[i];
    }
    BatchDesc d;
    ckbatch::describe(classes, num_classes, method, shp, workspace, &d);
    CK_HIP(hipMemcpy(workspace, &d, sizeof(d), hipMemcpyHostToDevice), "hipMemcpy(batch descriptor)");
    // Set here, outside any stream capture: the launch itself only enqueues.
    ckbatch::allow_lds(L.lds_bytes);
    plan->workspace = workspace;
    plan->grid = L.grid;
    plan->threads = kBatchThreads;
    plan->lds_bytes = static_cast<int32_t>(L.lds_bytes);
    plan->sum_product = d.sum_product;
    for (int i = 0; i < kMaxClasses; ++i) {
        plan->frames_per_group[i] = i < num_classes ? L.F[i] : 0;
        plan->lanes_per_row[i] = i < num_classes ? L.P[i] : 0;
    }
    plan->kernel_variant = L.variant;
    return CK_OK;
}

extern "C" int32_t ck_batch_launch(const ck_batch_plan* plan, void* stream) {
    CK_REQUIRE(plan != nullptr && plan->workspace != nullptr, "plan is not prepared");
    CK_REQUIRE(plan->grid >= 1 && plan->threads == kBatchThreads, "plan is not prepared");
    return ckbatch::launch(plan->workspace, plan->grid, static_cast<size_t>(plan->lds_bytes),
                           plan->sum_product != 0, plan->kernel_variant,
                           static_cast<hipStream_t>(stream));
}

namespace {
struct GraphHandle {
    hipGraph_t graph;
    hipGraphExec_t exec;
};
}  // namespace

extern "C" int32_t ck_batch_graph_create(const ck_batch_plan* plan, void** graph) {
    CK_REQUIRE(graph != nullptr, "graph is NULL");
    hipStream_t s = nullptr;
    CK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    GraphHandle* h = new GraphHandle{nullptr, nullptr};
    int32_t st = ck::hip_status(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
                                "hipStreamBeginCapture");
    if (st == CK_OK) {
        const int32_t launch = ck_batch_launch(plan, s);
        const int32_t end = ck::hip_status(hipStreamEndCapture(s, &h->graph), "hipStreamEndCapture");
        st = launch != CK_OK ? launch : end;
    }
    if (st == CK_OK)
        st = ck::hip_status(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0),
                            "hipGraphInstantiate");
    (void)hipStreamDestroy(s);
    if (st != CK_OK) {
        if (h->exec) (void)hipGraphExecDestroy(h->exec);
        if (h->graph) (void)hipGraphDestroy(h->graph);
        delete h;
        return st;
    }
    *graph = h;
    return CK_OK;
}

extern "C" int32_t ck_batch_graph_launch(void* graph, void* stream) {
    CK_REQUIRE(graph != nullptr, "graph is NULL");
    GraphHandle* h = static_cast<GraphHandle*>(graph);
    CK_HIP(hipGraphLaunch(h->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
    return CK_OK;
}

extern "C" int32_t ck_batch_graph_destroy(void* graph) {
    if (graph == nullptr) return CK_OK;
    GraphHandle* h = static_cast<GraphHandle*>(graph);
    int32_t st = CK_OK;
    if (h->exec) st = ck::hip_status(hipGraphExecDestroy(h->exec), "hipGraphExecDestroy");
    if (h->graph && st == CK_OK) st = ck::hip_status(hipGraphDestroy(h->graph), "hipGraphDestroy");
    delete h;
    return st;
}
