// Serving engine: continuous batching of decode requests over the batched
// QC-LDPC decoder, with admission against a fixed device-memory budget.
//
// The caller owns the memory: per code a ring of frame buffers (LLRs in,
// APP / bits / iterations / valid out, `capacity` frames), sized from its
// HBM budget, and `max_inflight` launch workspaces.  The engine hands out
// ring space and turns requests into launches:
//
//   admit(code, frames, tag)  reserves `frames` contiguous slots of the
//                             code's ring (CK_EBUSY when the ring, i.e. the
//                             budget, is full; the caller retries later);
//   submit(tag)               the caller has written the request's LLRs;
//   step()                    one launch over the submitted requests not yet
//                             launched, oldest submission first, up to
//                             max_launch_frames frames (at least one
//                             request; 0 = no cap) and 8 classes
//                             (contiguous slots of one code merge into one
//                             class), if a launch workspace is free;
//   poll()                    requests whose launch has completed;
//   release(tag)              the caller has read the results; the slots
//                             return to the ring in admission order.
//
// Requests therefore join the next launch as soon as they arrive instead of
// waiting for a full batch, and a launch in flight never blocks admission
// or submission.  The frame cap bounds how long one launch holds its
// requests: a burst is launched in pieces, so a request arriving behind it
// joins a launch of bounded size instead of waiting for the whole burst.  Each workspace has its own stream, so launches in flight
// overlap on the device (a new launch's workgroups take the CUs that the
// previous one's finished workgroups free); a launch first waits for the
// work enqueued on the caller's stream (the requests' LLRs), and records an
// event that poll() queries without blocking; a launch completes as soon as
// its event has fired, whether or not older launches still run.  A frame decodes the same
// alone or in any launch (the batched decoder's contract), so a request's
// results do not depend on the traffic it was served with.
//
// dry_run: no device work at all (no copies, launches or events); step()
// marks requests launched and poll() completes them.  The admission, ring
// and ordering logic can then be exercised without a GPU.
//
// Not thread-safe: one host thread drives an engine.
#include "ck_batch_int.h"
#include "ck_common.h"

#include <algorithm>
#include <cstring>
#include <deque>
#include <vector>

namespace {

enum State { kAdmitted = 0, kSubmitted = 1, kLaunched = 2, kDone = 3, kReleased = 4, kGap = 5 };

struct Request {
    uint64_t tag;
    int code;
    long long first;   // monotonic slot index of the first frame
    long long frames;
    int state;
    int launch;        // workspace slot of its launch
    long long order;   // submission order
};

struct Ring {
    ck_serve_ring buf;
    long long head = 0;  // first live monotonic index
    long long tail = 0;  // next free monotonic index
    std::deque<Request> live;
};

struct Launch {
    bool busy = false;
    long long seq = 0;
    hipStream_t stream = nullptr;  // own stream: launches in flight overlap on the device
    hipEvent_t done = nullptr;
    void* host_desc = nullptr;     // pinned: read by the descriptor copy
    void* workspace = nullptr;     // device
};

}  // namespace

struct ck_serve {
    std::vector<ck_ldpc_code> codes;
    ck_ldpc_method method{};
    std::vector<Ring> rings;
    std::vector<Launch> launches;
    hipStream_t stream = nullptr;     // the caller's stream, where request LLRs are written
    hipEvent_t inputs = nullptr;      // recorded there at each launch, waited on by the launch
    bool dry = false;
    long long next_seq = 0;
    long long next_order = 0;
    long long max_launch_frames = 0;  // 0: no cap
    int64_t workspace_bytes = 0;
    std::vector<uint64_t> completed;  // tags completed, not yet returned by poll
};

namespace {

Request* find(ck_serve* s, uint64_t tag) {
    for (Ring& r : s->rings)
        for (Request& q : r.live)
            if (q.state != kGap && q.tag == tag) return &q;
    return nullptr;
}

// need_buffers: the rings' device buffers must be set (not for sizing).
int32_t check_config(const ck_serve_config* cfg, bool need_buffers = true) {
    CK_REQUIRE(cfg != nullptr, "config is NULL");
    CK_REQUIRE(cfg->num_codes >= 1 && cfg->num_codes <= ckbatch::kMaxClasses,
               "num_codes %d not in [1, %d]", cfg->num_codes, ckbatch::kMaxClasses);
    CK_REQUIRE(cfg->codes != nullptr && cfg->rings != nullptr && cfg->method != nullptr,
               "NULL codes, rings or method");
    CK_REQUIRE(cfg->max_inflight >= 1 && cfg->max_inflight <= 8, "max_inflight %d not in [1, 8]",
               cfg->max_inflight);
    if (int32_t st = ck::check_method(cfg->method)) return st;
    CK_REQUIRE(cfg->method->schedule == 0, "serving runs the layered schedule");
    CK_REQUIRE(cfg->max_launch_frames >= 0, "max_launch_frames %lld < 0",
               static_cast<long long>(cfg->max_launch_frames));
    for (int i = 0; i < cfg->num_codes; ++i) {
        CK_REQUIRE(cfg->codes[i] != nullptr, "code %d is NULL", i);
        if (int32_t st = ck::check_code(cfg->codes[i])) return st;
        const ck_serve_ring& r = cfg->rings[i];
        CK_REQUIRE(r.capacity >= 1, "ring %d: capacity %lld < 1", i,
                   static_cast<long long>(r.capacity));
        CK_REQUIRE(!need_buffers || cfg->dry_run ||
                       (r.llrs && r.app && r.bits && r.iterations && r.valid),
                   "ring %d: NULL buffer", i);
    }
    return CK_OK;
}

// Launch workspace bound over every subset of the codes: the descriptor
// plus one workgroup's largest slot workspace for each CU.
int32_t workspace_bound(const ck_serve_config* cfg, int64_t* out) {
    size_t slot = 0;
    int grid = 0;
    for (int i = 0; i < cfg->num_codes; ++i) {
        ck_batch_class c{};
        c.code = cfg->codes[i];
        c.num_frames = 0;
        ckbatch::Shape sh{};
        if (int32_t st = ckbatch::shape(&c, 1, cfg->method, &sh)) return st;
        slot = std::max(slot, sh.slot_bytes);
        grid = std::max(grid, sh.grid);
    }
    const size_t desc = (ckbatch::desc_bytes() + 255) / 256 * 256;
    *out = static_cast<int64_t>(desc + slot * static_cast<size_t>(grid));
    return CK_OK;
}

// Completes every launch whose event has fired, oldest first (a short
// launch started after a long one completes without waiting for it).
int32_t reap(ck_serve* s) {
    std::vector<Launch*> busy;
    for (Launch& l : s->launches)
        if (l.busy) busy.push_back(&l);
    std::sort(busy.begin(), busy.end(), [](const Launch* a, const Launch* b) { return a->seq < b->seq; });
    for (Launch* l : busy) {
        if (!s->dry) {
            const hipError_t e = hipEventQuery(l->done);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return ck::hip_status(e, "hipEventQuery");
        }
        const int w = static_cast<int>(l - s->launches.data());
        for (Ring& r : s->rings)
            for (Request& q : r.live)
                if (q.state == kLaunched && q.launch == w) {
                    q.state = kDone;
                    s->completed.push_back(q.tag);
                }
        l->busy = false;
    }
    return CK_OK;
}

}  // namespace

extern "C" int64_t ck_serve_workspace_bytes(const ck_serve_config* config) {
    if (check_config(config, false) != CK_OK) return -1;
    if (config->dry_run) return 0;
    int64_t b = 0;
    if (workspace_bound(config, &b) != CK_OK) return -1;
    return b;
}

extern "C" int32_t ck_serve_create(const ck_serve_config* config, void* const* workspaces,
                                   void* stream, ck_serve** out) {
    if (int32_t st = check_config(config)) return st;
    CK_REQUIRE(out != nullptr, "out is NULL");
    int64_t ws = 0;
    if (!config->dry_run) {
        if (int32_t st = workspace_bound(config, &ws)) return st;
        CK_REQUIRE(workspaces != nullptr, "workspaces is NULL");
    }
    ck_serve* s = new ck_serve();
    s->method = *config->method;
    s->stream = static_cast<hipStream_t>(stream);
    s->dry = config->dry_run != 0;
    s->max_launch_frames = config->max_launch_frames;
    s->workspace_bytes = ws;
    for (int i = 0; i < config->num_codes; ++i) {
        s->codes.push_back(*config->codes[i]);
        Ring r;
        r.buf = config->rings[i];
        s->rings.push_back(r);
    }
    s->launches.resize(static_cast<size_t>(config->max_inflight));
    int32_t st = CK_OK;
    if (!s->dry) {
        size_t lds = 0;
        for (int i = 0; i < config->num_codes && st == CK_OK; ++i) {
            ck_batch_class c{};
            c.code = config->codes[i];
            ckbatch::Shape sh{};
            st = ckbatch::shape(&c, 1, config->method, &sh);
            lds = std::max(lds, sh.lds_bytes);
        }
        if (st == CK_OK) ckbatch::allow_lds(lds);
        for (size_t w = 0; w < s->launches.size() && st == CK_OK; ++w) {
            Launch& l = s->launches[w];
            l.workspace = workspaces[w];
            if (l.workspace == nullptr || (reinterpret_cast<uintptr_t>(l.workspace) & 255) != 0)
                st = ck::fail(CK_EINVAL, "workspace %zu NULL or not 256-byte aligned", w);
            if (st == CK_OK)
                st = ck::hip_status(hipHostMalloc(&l.host_desc, ckbatch::desc_bytes(),
                                                  hipHostMallocDefault),
                                    "hipHostMalloc");
            if (st == CK_OK)
                st = ck::hip_status(hipEventCreateWithFlags(&l.done, hipEventDisableTiming),
                                    "hipEventCreate");
            if (st == CK_OK)  // default priority, like every stream of this library
                st = ck::hip_status(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking),
                                    "hipStreamCreate");
        }
        if (st == CK_OK)
            st = ck::hip_status(hipEventCreateWithFlags(&s->inputs, hipEventDisableTiming),
                                "hipEventCreate");
    }
    if (st != CK_OK) {
        (void)ck_serve_destroy(s);
        return st;
    }
    *out = s;
    return CK_OK;
}

extern "C" int32_t ck_serve_admit(ck_serve* s, int32_t code, int64_t frames, uint64_t tag,
                                  ck_serve_slot* where) {
    CK_REQUIRE(s != nullptr && where != nullptr, "NULL engine or slot");
    CK_REQUIRE(code >= 0 && code < static_cast<int32_t>(s->rings.size()), "code %d out of range",
               code);
    Ring& r = s->rings[static_cast<size_t>(code)];
    CK_REQUIRE(frames >= 1 && frames <= r.buf.capacity,
               "request of %lld frames; the ring of code %d holds %lld",
               static_cast<long long>(frames), code, static_cast<long long>(r.buf.capacity));
    CK_REQUIRE(find(s, tag) == nullptr, "tag %llu is in use", static_cast<unsigned long long>(tag));
    const long long cap = r.buf.capacity;
    long long pos = r.tail % cap;
    long long gap = 0;
    if (pos + frames > cap) {  // a request's frames are contiguous: skip to the ring's start
        gap = cap - pos;
        pos = 0;
    }
    if (r.tail + gap + frames - r.head > cap)
        return ck::fail(CK_EBUSY, "code %d: %lld frames do not fit (%lld of %lld in use)", code,
                        static_cast<long long>(frames), r.tail - r.head, cap);
    if (gap > 0) {
        r.live.push_back(Request{0, code, r.tail, gap, kGap, -1, -1});
        r.tail += gap;
    }
    r.live.push_back(Request{tag, code, r.tail, frames, kAdmitted, -1, -1});
    r.tail += frames;
    const ck_ldpc_code& c = s->codes[static_cast<size_t>(code)];
    where->first = pos;
    where->llrs = r.buf.llrs ? r.buf.llrs + pos * c.n : nullptr;
    where->app = r.buf.app ? r.buf.app + pos * c.n : nullptr;
    where->bits = r.buf.bits ? r.buf.bits + pos * c.k : nullptr;
    where->iterations = r.buf.iterations ? r.buf.iterations + pos : nullptr;
    where->valid = r.buf.valid ? r.buf.valid + pos : nullptr;
    return CK_OK;
}

extern "C" int32_t ck_serve_submit(ck_serve* s, uint64_t tag) {
    CK_REQUIRE(s != nullptr, "NULL engine");
    Request* q = find(s, tag);
    CK_REQUIRE(q != nullptr, "unknown tag %llu", static_cast<unsigned long long>(tag));
    CK_REQUIRE(q->state == kAdmitted, "tag %llu was already submitted",
               static_cast<unsigned long long>(tag));
    q->state = kSubmitted;
    q->order = s->next_order++;
    return CK_OK;
}

extern "C" int32_t ck_serve_step(ck_serve* s, int64_t* launched_frames) {
    CK_REQUIRE(s != nullptr, "NULL engine");
    if (launched_frames) *launched_frames = 0;
    if (int32_t st = reap(s)) return st;
    int w = -1;
    for (size_t i = 0; i < s->launches.size(); ++i)
        if (!s->launches[i].busy) {
            w = static_cast<int>(i);
            break;
        }
    if (w < 0) return CK_OK;  // every workspace in flight: the next step launches
    // The requests of this launch: submitted ones in submission order, up to
    // the frame cap (at least one) and kMaxClasses runs.
    std::vector<Request*> pending;
    for (Ring& r : s->rings)
        for (Request& q : r.live)
            if (q.state == kSubmitted) pending.push_back(&q);
    if (pending.empty()) return CK_OK;
    std::sort(pending.begin(), pending.end(),
              [](const Request* a, const Request* b) { return a->order < b->order; });
    // Runs of chosen requests, contiguous in their ring, become classes.
    struct Run {
        int code;
        long long pos;
        long long frames;
        std::vector<Request*> reqs;
    };
    std::vector<Run> runs;
    long long chosen = 0;
    for (Request* q : pending) {
        if (s->max_launch_frames > 0 && chosen > 0 && chosen + q->frames > s->max_launch_frames)
            break;
        const long long pos = q->first % s->rings[static_cast<size_t>(q->code)].buf.capacity;
        Run* ext = nullptr;
        for (Run& run : runs)
            if (run.code == q->code && run.pos + run.frames == pos) ext = &run;
        if (ext != nullptr) {
            ext->frames += q->frames;
            ext->reqs.push_back(q);
        } else {
            if (runs.size() == static_cast<size_t>(ckbatch::kMaxClasses)) break;
            runs.push_back(Run{q->code, pos, q->frames, {q}});
        }
        chosen += q->frames;
    }
    // Largest frames first (the batched decoder's cost order), then oldest.
    std::stable_sort(runs.begin(), runs.end(), [&](const Run& a, const Run& b) {
        return s->codes[static_cast<size_t>(a.code)].n > s->codes[static_cast<size_t>(b.code)].n;
    });
    ck_batch_class cls[ckbatch::kMaxClasses];
    std::memset(cls, 0, sizeof(cls));
    long long total = 0;
    for (size_t i = 0; i < runs.size(); ++i) {
        const Run& run = runs[i];
        const Ring& r = s->rings[static_cast<size_t>(run.code)];
        const ck_ldpc_code& c = s->codes[static_cast<size_t>(run.code)];
        cls[i].code = &c;
        cls[i].num_frames = run.frames;
        if (!s->dry) {
            cls[i].llrs = r.buf.llrs + run.pos * c.n;
            cls[i].app = r.buf.app + run.pos * c.n;
            cls[i].bits = r.buf.bits + run.pos * c.k;
            cls[i].iterations = r.buf.iterations + run.pos;
            cls[i].valid = r.buf.valid + run.pos;
        }
        total += run.frames;
    }
    Launch& l = s->launches[static_cast<size_t>(w)];
    if (!s->dry) {
        const int n = static_cast<int>(runs.size());
        ckbatch::Shape sh{};
        if (int32_t st = ckbatch::shape(cls, n, &s->method, &sh)) return st;
        const size_t need = ckbatch::desc_bytes() + sh.slot_bytes * static_cast<size_t>(sh.grid);
        CK_REQUIRE(static_cast<int64_t>(need) <= s->workspace_bytes,
                   "launch workspace %zu above the bound %lld", need,
                   static_cast<long long>(s->workspace_bytes));
        ckbatch::describe(cls, n, &s->method, sh, l.workspace, l.host_desc);
        // The launch follows everything enqueued so far on the caller's stream
        // (the requests' LLRs), then runs on its own stream.
        CK_HIP(hipEventRecord(s->inputs, s->stream), "hipEventRecord(inputs)");
        CK_HIP(hipStreamWaitEvent(l.stream, s->inputs, 0), "hipStreamWaitEvent");
        CK_HIP(hipMemcpyAsync(l.workspace, l.host_desc, ckbatch::desc_bytes(),
                              hipMemcpyHostToDevice, l.stream),
               "hipMemcpyAsync(descriptor)");
        if (int32_t st = ckbatch::launch(l.workspace, sh.grid, sh.lds_bytes, s->method.rule == 1,
                                         sh.variant, l.stream))
            return st;
        CK_HIP(hipEventRecord(l.done, l.stream), "hipEventRecord");
    }
    l.busy = true;
    l.seq = s->next_seq++;
    for (const Run& run : runs)
        for (Request* q : run.reqs) {
            q->state = kLaunched;
            q->launch = w;
        }
    if (launched_frames) *launched_frames = total;
    return CK_OK;
}

extern "C" int32_t ck_serve_poll(ck_serve* s, uint64_t* tags, int32_t max_tags, int32_t* count) {
    CK_REQUIRE(s != nullptr && count != nullptr, "NULL engine or count");
    CK_REQUIRE(max_tags >= 0 && (max_tags == 0 || tags != nullptr), "NULL tags");
    if (int32_t st = reap(s)) return st;
    const int32_t n = std::min<int32_t>(max_tags, static_cast<int32_t>(s->completed.size()));
    for (int32_t i = 0; i < n; ++i) tags[i] = s->completed[static_cast<size_t>(i)];
    s->completed.erase(s->completed.begin(), s->completed.begin() + n);
    *count = n;
    return CK_OK;
}

extern "C" int32_t ck_serve_release(ck_serve* s, uint64_t tag) {
    CK_REQUIRE(s != nullptr, "NULL engine");
    Request* q = find(s, tag);
    CK_REQUIRE(q != nullptr, "unknown tag %llu", static_cast<unsigned long long>(tag));
    CK_REQUIRE(q->state == kDone, "tag %llu has not completed",
               static_cast<unsigned long long>(tag));
    q->state = kReleased;
    Ring& r = s->rings[static_cast<size_t>(q->code)];
    while (!r.live.empty() && (r.live.front().state == kReleased || r.live.front().state == kGap)) {
        r.head = r.live.front().first + r.live.front().frames;
        r.live.pop_front();
    }
    if (r.live.empty()) r.head = r.tail;
    return CK_OK;
}

extern "C" int32_t ck_serve_usage(const ck_serve* s, int32_t code, int64_t* in_use,
                                  int64_t* capacity, int32_t* inflight) {
    CK_REQUIRE(s != nullptr, "NULL engine");
    CK_REQUIRE(code >= 0 && code < static_cast<int32_t>(s->rings.size()), "code %d out of range",
               code);
    const Ring& r = s->rings[static_cast<size_t>(code)];
    if (in_use) *in_use = r.tail - r.head;
    if (capacity) *capacity = r.buf.capacity;
    if (inflight) {
        int n = 0;
        for (const Launch& l : s->launches) n += l.busy ? 1 : 0;
        *inflight = n;
    }
    return CK_OK;
}

extern "C" int32_t ck_serve_destroy(ck_serve* s) {
    if (s == nullptr) return CK_OK;
    int32_t st = CK_OK;
    if (!s->dry) {
        for (Launch& l : s->launches) {
            if (l.stream) {
                const int32_t e = ck::hip_status(hipStreamSynchronize(l.stream),
                                                 "hipStreamSynchronize");
                st = st != CK_OK ? st : e;
                (void)hipStreamDestroy(l.stream);
            }
            if (l.done) (void)hipEventDestroy(l.done);
            if (l.host_desc) (void)hipHostFree(l.host_desc);
        }
        if (s->inputs) (void)hipEventDestroy(s->inputs);
    }
    delete s;
    return st;
}
