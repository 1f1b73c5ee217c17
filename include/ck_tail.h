/*
 * ck_tail.h - C ABI of libck_hip.so, the MI355X (gfx950) soft-decision
 * receive tail: bit-LLR demapping and QC-LDPC decoding in hand-written HIP.
 *
 * Conventions (all entry points):
 *   - plain C types only; every pointer named "device" is device memory of
 *     the current HIP device (hipMalloc / torch tensor storage);
 *   - `stream` is a hipStream_t (NULL = the null stream); work is enqueued,
 *     nothing synchronises and nothing is copied device->host;
 *   - the return value is 0 on success, else a ck_status; the message of the
 *     last failure on the calling thread is ck_last_error();
 *   - arguments are validated before any launch; a zero-size problem
 *     returns 0 without launching (reference: commkit/_cuda/llr.py:24-25,
 *     commkit/_cuda/coding.py:54-55).
 *
 * LLR sign convention: LLR = log P(b=0|y) / P(b=1|y); the k bits of a symbol
 * are adjacent in the output, MSB first (commkit/mapping/llr.py:24-95).
 */
#ifndef CK_TAIL_H
#define CK_TAIL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ck_status {
    CK_OK = 0,
    CK_EINVAL = 1,   /* bad argument (shape, size, null pointer, range) */
    CK_EHIP = 2,     /* HIP runtime error (launch, attribute) */
    CK_ENOSPACE = 3, /* problem exceeds a fixed limit of the library */
    CK_EBUSY = 4     /* no room now (serving: the request's ring is full); retry later */
};

/* Library version (major * 10000 + minor * 100 + patch). */
int32_t ck_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* ck_last_error(void);

/* ------------------------------------------------------------------------
 * Demapper
 * ------------------------------------------------------------------------ */

/* One axis of a separable square lattice (Gray QAM, shaped or not); all
 * pointers are device memory.  Reference: the per-axis tables built by
 * commkit/mapping/_demap.py:316-347 (levels, log prior, packed labels,
 * opposite levels, bit positions).  Levels and prior are float64 here (the
 * reference uploads float32): the kernels form metric differences in
 * float64 from them, so an LLR keeps its relative accuracy down to zero. */
typedef struct ck_axis {
    const double* levels;    /* (L,) ascending, equally spaced for max-log */
    const double* log_prior; /* (L,) log marginal prior; zeros = uniform   */
    const int32_t* labels;   /* (L,) label of each level over the axis's
                                num_bits bits, MSB first                  */
    const int32_t* lo;       /* (num_bits, L) nearest level below whose bit
                                differs, -1 for none (closed form only)   */
    const int32_t* hi;       /* (num_bits, L) nearest level above, -1 none */
    const int32_t* pos;      /* (num_bits,) position of each bit in the
                                symbol label                              */
    int32_t num_levels;      /* L = 2^num_bits, 2..256                     */
    int32_t num_bits;        /* 1..8                                       */
} ck_axis;

/* Closed-form max-log LLRs of a uniform separable lattice.
 * x: (n,) complex64 device; out: (n, k) float32 device, k = re.num_bits +
 * im.num_bits.  inv_s2 = 1 / sigma^2 (float64: with a prior, the likelihood
 * and prior terms of a small LLR cancel, and a float32 1/sigma^2 would
 * cost its rounding relative to the larger term).
 * Replaces the kernel `llr_separable_maxlog` (commkit/_cuda/src/llr.cu:
 * 304-337) reached through _cuda.get_kernel("llr_separable_maxlog")
 * (commkit/_cuda/llr.py:150; call site commkit/mapping/_demap.py:330). */
int32_t ck_llr_separable_maxlog(const void* x, int64_t n, const ck_axis* re,
                                const ck_axis* im, double inv_s2, float* out,
                                void* stream);

/* LLRs from the metrics of each axis's levels, with the log prior:
 * exact = 0 max-log, 1 exact log-sum-exp.  Same layout as above.
 * Replaces `llr_separable` (commkit/_cuda/src/llr.cu:225-259; wrapper
 * commkit/_cuda/llr.py:87; call site commkit/mapping/_demap.py:349). */
int32_t ck_llr_separable(const void* x, int64_t n, const ck_axis* re,
                         const ck_axis* im, double inv_s2, int32_t exact,
                         float* out, void* stream);

/* Separable LLRs of `rows` rows of `row_len` symbols, each row with its own
 * noise: inv_s2 is a (rows,) float64 device array of 1/sigma^2 (positive
 * and finite; not checked on the device).  closed = 1: the closed form of
 * ck_llr_separable_maxlog (exact must be 0); closed = 0: the level metrics
 * of ck_llr_separable.  Row r gives the same LLRs, bit for bit, as the
 * scalar entry point called on that row alone with inv_s2[r].
 * No counterpart in the reference, whose compute_llr takes one noise
 * variance per call (commkit/mapping/llr.py:24-95); this serves frames of
 * mixed Es/N0 demapped in one launch. */
int32_t ck_llr_separable_rows(const void* x, int64_t rows, int64_t row_len,
                              const ck_axis* re, const ck_axis* im,
                              const double* inv_s2, int32_t closed, int32_t exact,
                              float* out, void* stream);

/* Separable LLRs written in a code's bit layout: the LLR of flat index t
 * (symbol-major, k adjacent) goes to out[(t / period) * period +
 * perm[t % period]], for t < limit (a whole number of blocks of `period`
 * LLRs; later symbols are not demapped).  perm (period,) int32 device must
 * be a permutation of 0..period-1 (not checked).  closed / exact as for
 * ck_llr_separable_rows.  No counterpart in the reference, which demaps and
 * then gathers (commkit/coding/shaping.py, PAS): probabilistic amplitude
 * shaping's codeword LLRs come out of the demapper in one pass. */
int32_t ck_llr_separable_permuted(const void* x, int64_t n, const ck_axis* re,
                                  const ck_axis* im, double inv_s2, int32_t closed,
                                  int32_t exact, float* out, const int32_t* perm,
                                  int64_t period, int64_t limit, void* stream);

/* One request of ck_llr_separable_segments: the LLRs of x[0 .. n) into
 * out, as ck_llr_separable (perm NULL: (n, k) float32) or as
 * ck_llr_separable_permuted (perm, period, limit: the code's bit layout;
 * only the symbols holding LLRs below limit are demapped), with
 * 1/sigma^2 = inv_s2 or, when inv_rows is set, inv_rows[i / row_len]
 * (float64, device).  All pointers are device pointers. */
typedef struct ck_llr_segment {
    const void* x;           /* (n,) complex64                              */
    float* out;              /* LLRs                                        */
    int64_t n;               /* symbols                                     */
    double inv_s2;           /* 1 / sigma^2 when inv_rows is NULL           */
    const double* inv_rows;  /* (n / row_len,) 1 / sigma^2 per row, or NULL */
    int64_t row_len;         /* symbols per row (with inv_rows)             */
    const int32_t* perm;     /* (period,) bit layout, or NULL               */
    int64_t period;
    int64_t limit;
} ck_llr_segment;

/* The requests segs[0 .. count) of one separable constellation demapped in
 * one launch per 16 of them (closed: the closed form, max-log, uniform;
 * else level metrics, exact or max-log), each exactly as its own call of
 * ck_llr_separable / ck_llr_separable_rows / ck_llr_separable_permuted would
 * give.  segs is a host array, read before the call returns.  No
 * counterpart in the reference: for request loops (a serving loop demapping
 * each admitted request into its ring slot) that would otherwise launch
 * once per request. */
int32_t ck_llr_separable_segments(const ck_llr_segment* segs, int32_t count,
                                  const ck_axis* re, const ck_axis* im, int32_t closed,
                                  int32_t exact, void* stream);

/* LLRs from the metrics of all M points (any constellation).
 * points: (M,) complex128 device; log_prior: (M,) float64 device (zeros =
 * uniform); labels: (M,) int32 packed MSB-first labels; 2 <= M <= 65536,
 * k = log2(M) <= 16.  out: (n, k) float32.
 * Replaces `llr` (commkit/_cuda/src/llr.cu:200-223; wrapper
 * commkit/_cuda/llr.py:36; call site commkit/mapping/_demap.py:210). */
int32_t ck_llr_points(const void* x, int64_t n, const void* points,
                      const double* log_prior, const int32_t* labels, int32_t M,
                      int32_t k, double inv_s2, int32_t exact, float* out,
                      void* stream);

/* ------------------------------------------------------------------------
 * LDPC decoder
 * ------------------------------------------------------------------------ */

/* An LDPC code on the device.  The CSR view is always present; the
 * quasi-cyclic view (lifting > 0) lets the decoder compute columns from
 * circulant shifts instead of reading them.  Reference: LDPCCode._schedule
 * (commkit/coding/ldpc.py:283-296) gives (row_ptr, cols, layer_ptr); the
 * QC view restates LDPCCode.quasi_cyclic (commkit/coding/ldpc.py:183-214). */
typedef struct ck_ldpc_code {
    int32_t num_checks;        /* m                                          */
    int32_t num_variables;     /* N (mother codeword)                        */
    int32_t n;                 /* transmitted bits, N - punctured            */
    int32_t k;                 /* information bits                           */
    int32_t num_edges;         /* E                                          */
    int32_t num_layers;        /* L                                          */
    int32_t max_row_degree;    /* largest check degree                       */
    const int32_t* row_ptr;    /* (m+1) CSR, rows in layer order             */
    const int32_t* cols;       /* (E) ascending within a row                 */
    const int32_t* layer_ptr;  /* (L+1) first row of each layer              */
    const int32_t* var_ptr;    /* (N+1) edges of each variable ...           */
    const int32_t* var_edges;  /* (E) ... as CSR edge ids, ascending check id
                                  (flooding only; may be NULL otherwise)     */
    const int32_t* info;       /* (k) codeword position of each info bit, or
                                  NULL for positions 0..k-1                  */
    int32_t lifting;           /* Z > 0 for a QC view, 0 for none            */
    int32_t base_rows;         /* mb                                         */
    int32_t base_cols;         /* nb                                         */
    int32_t num_slots;         /* S = nonzero base entries                   */
    const int32_t* qc_row_ptr; /* (mb+1) slots of each block row, block rows
                                  in schedule order                          */
    const int32_t* qc_slot_col;   /* (S) block column of each slot           */
    const int32_t* qc_slot_shift; /* (S) circulant shift of each slot        */
    const int32_t* qc_row_id;     /* (mb) natural index of each scheduled
                                     block row                               */
    const int32_t* qc_col_ptr;    /* (nb+1) slots of each block column ...   */
    const int32_t* qc_col_slot;   /* (S) ... in ascending natural block row  */
    const int32_t* qc_slot_pos;   /* (S) position of the slot in its row     */
    const int32_t* qc_slot_row;   /* (S) scheduled block row of each slot    */
    const int32_t* qc_slot_packed;/* (S + 32, 2) {block column * Z, shift}
                                     per slot, then 32 zero entries         */
    const int32_t* slot_cols;     /* (w, m) slot-major columns of the rows in
                                     schedule order, -1 past a row's end,
                                     w = 8, 16 or 32 (the smallest at least
                                     max_row_degree); NULL: the layered
                                     generic kernel reads the CSR            */
} ck_ldpc_code;

/* Decoding method.  rule/scale/offset/max_iterations/early_stop follow
 * MinSum and SumProduct (commkit/coding/ldpc.py:317-371).  schedule 1
 * (flooding) has no counterpart in the reference.  The tile fields are
 * autotuner choices; 0 / -1 let the library choose. */
typedef struct ck_ldpc_method {
    int32_t rule;             /* 0 normalized/offset min-sum, 1 sum-product */
    int32_t schedule;         /* 0 layered, 1 flooding                      */
    float scale;              /* min-sum normalisation (1 for sum-product)  */
    float offset;             /* min-sum offset (0 for sum-product)         */
    int32_t max_iterations;   /* >= 1                                       */
    int32_t early_stop;       /* stop a frame on a zero syndrome            */
    int32_t kernel;           /* 0 auto, 1 generic CSR, 2 quasi-cyclic       */
    int32_t frames_per_group; /* QC: frames per workgroup, 0 = auto         */
    int32_t lanes_per_row;    /* QC min-sum: lanes per check row 1/2/4, 0 = auto */
    int32_t state_in_lds;     /* QC min-sum: check state in LDS 1, HBM 0, -1 auto */
} ck_ldpc_method;

/* Device workspace (bytes) ck_ldpc_decode needs for R frames. */
int64_t ck_ldpc_workspace_bytes(const ck_ldpc_code* code,
                                const ck_ldpc_method* method, int64_t R);

/* Decode R frames.  llrs (R, n) float32 channel LLRs in; app (R, n) float32
 * a-posteriori LLRs out (may alias llrs); bits (R, k) int8 hard decisions of
 * the info bits out (NULL to skip); iterations (R) int32 and valid (R)
 * uint8 out; workspace: ck_ldpc_workspace_bytes() device bytes, 256-byte
 * aligned (NULL when that is 0).  Punctured bits start at LLR 0.
 * Replaces the kernel `ldpc_layered<SUM_PRODUCT>` (commkit/_cuda/src/
 * ldpc.cu:107-159) reached through _cuda.get_kernel("ldpc_layered",
 * sum_product=) (commkit/_cuda/coding.py:11-69; call site
 * commkit/coding/ldpc.py:566-581), including the hard decisions of
 * commkit/coding/ldpc.py:582-587. */
int32_t ck_ldpc_decode(const ck_ldpc_code* code, const ck_ldpc_method* method,
                       const float* llrs, float* app, int8_t* bits,
                       int32_t* iterations, uint8_t* valid, int64_t R,
                       void* workspace, void* stream);

/* ck_ldpc_decode of frames given as channel symbols: symbols (R, n / k)
 * interleaved complex float32 (k = re->num_bits + im->num_bits bits a
 * symbol, both axes of 1..4 bits and alike, uniform levels with their
 * closed-form tables lo/hi), demapped by max-log inside the decoder's frame
 * load, so the LLRs never pass through device memory.  1/sigma^2 is inv_s2
 * for every frame, or inv_rows[r] (R,) float64 device for frame r when
 * inv_rows is not NULL.  Outputs, workspace and results are those of
 * ck_ldpc_decode of the LLRs ck_llr_separable_maxlog (or
 * ck_llr_separable_rows with closed = 1) writes for the same symbols, bit
 * for bit.  Quasi-cyclic codes, layered schedule, or flooding when the
 * code's channel LLRs, messages and column tables fit in LDS beside its
 * APPs (CK_EINVAL otherwise, and for the generic kernel).  No counterpart in the reference, whose receive tail runs
 * compute_llr (commkit/mapping/llr.py:24-95) and then the decoder
 * (commkit/coding/ldpc.py:566-587) as two launches. */
int32_t ck_ldpc_decode_symbols(const ck_ldpc_code* code, const ck_ldpc_method* method,
                               const void* symbols, const ck_axis* re, const ck_axis* im,
                               double inv_s2, const double* inv_rows, float* app,
                               int8_t* bits, int32_t* iterations, uint8_t* valid, int64_t R,
                               void* workspace, void* stream);

/* ck_ldpc_decode without a workspace argument and without allocating: the
 * workspace is a static device array of 64 MiB (each GPU the library runs
 * on holds one until the process ends), and frames are decoded in chunks
 * that fit it, one launch after the other on `stream`; a frame needing more
 * than 64 MiB of workspace is refused (CK_EINVAL).  Nothing in the scratch
 * must start at zero (the first iteration reads no old message).  Calls
 * that overlap on one GPU (other streams, other threads) share the scratch
 * and race: serialise them.  Outputs, validation and results are those of
 * ck_ldpc_decode.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
int32_t ck_ldpc_decode_static(const ck_ldpc_code* code, const ck_ldpc_method* method,
                              const float* llrs, float* app, int8_t* bits,
                              int32_t* iterations, uint8_t* valid, int64_t R, void* stream);

/* ------------------------------------------------------------------------
 * Batched decoding: frames of several QC codes in one persistent launch
 * ------------------------------------------------------------------------ */

/* The frames of one code: inputs and outputs as for ck_ldpc_decode, frames
 * contiguous.  frames_per_group / lanes_per_row: 0 lets the library choose. */
typedef struct ck_batch_class {
    const ck_ldpc_code* code;  /* host struct; its tables stay valid while
                                  the batch is used                           */
    int64_t num_frames;
    const float* llrs;         /* (num_frames, n) device                      */
    float* app;                /* (num_frames, n) device                      */
    int8_t* bits;              /* (num_frames, k) device, or NULL             */
    int32_t* iterations;       /* (num_frames) device                         */
    uint8_t* valid;            /* (num_frames) device                         */
    int32_t frames_per_group;  /* frame slots per workgroup                   */
    int32_t lanes_per_row;     /* 1 or 2                                      */
} ck_batch_class;

/* Launch plan filled by ck_batch_prepare. */
typedef struct ck_batch_plan {
    void* workspace;
    int32_t grid;
    int32_t threads;
    int32_t lds_bytes;
    int32_t sum_product;
    int32_t frames_per_group[8];
    int32_t lanes_per_row[8];
    int32_t kernel_variant;  /* chosen by ck_batch_prepare, read by ck_batch_launch */
} ck_batch_plan;

/* Device workspace for a batch of up to 8 classes (-1 on bad arguments).
 * No counterpart in the reference, which decodes one code per launch in
 * chunks (commkit/coding/ldpc.py:565-581); the batch generalises that loop
 * to uneven codes in one launch. */
int64_t ck_batch_workspace_bytes(const ck_batch_class* classes,
                                 int32_t num_classes,
                                 const ck_ldpc_method* method);

/* Writes the batch's descriptor into the workspace (a synchronous copy) and
 * fills the launch plan.  Layered min-sum or sum-product.  Classes should be
 * ordered by cost, largest frames first. */
int32_t ck_batch_prepare(const ck_batch_class* classes, int32_t num_classes,
                         const ck_ldpc_method* method, void* workspace,
                         int64_t workspace_bytes, ck_batch_plan* plan);

/* Decodes every frame of the prepared batch: a counter-reset kernel, then
 * the persistent decoder.  Reads and writes only the buffers named at
 * prepare time, so it may be captured in a HIP graph. */
int32_t ck_batch_launch(const ck_batch_plan* plan, void* stream);

/* A HIP graph of ck_batch_launch, instantiated; launch it any number of
 * times, then destroy it. */
int32_t ck_batch_graph_create(const ck_batch_plan* plan, void** graph);
int32_t ck_batch_graph_launch(void* graph, void* stream);
int32_t ck_batch_graph_destroy(void* graph);

/* ---------------------------------------------------------------------
 * Serving engine: continuous batching of decode requests over the
 * batched decoder, admission against a device-memory budget.  No
 * counterpart in the reference (it decodes one call at a time,
 * commkit/coding/base.py:317-379); requests are what a caller of
 * `decode` would pass, frames of one code each.
 *
 * The caller owns all memory: per code a ring of `capacity` frame
 * buffers, sized from its budget, and `max_inflight` launch workspaces of
 * ck_serve_workspace_bytes() bytes (256-byte aligned).  One host thread
 * drives an engine.
 * ------------------------------------------------------------------- */

typedef struct ck_serve ck_serve;

typedef struct ck_serve_ring {  /* device buffers of `capacity` frames   */
    float* llrs;                /* (capacity, n) channel LLRs (written by the caller) */
    float* app;                 /* (capacity, n) APP LLRs                */
    int8_t* bits;               /* (capacity, k) hard info bits          */
    int32_t* iterations;        /* (capacity,)                           */
    uint8_t* valid;             /* (capacity,)                           */
    int64_t capacity;
} ck_serve_ring;

typedef struct ck_serve_config {
    const ck_ldpc_code* const* codes;  /* num_codes codes                 */
    const ck_serve_ring* rings;        /* one ring per code               */
    int32_t num_codes;                 /* 1..8                            */
    int32_t max_inflight;              /* launches in flight, 1..8        */
    const ck_ldpc_method* method;      /* layered                         */
    int32_t dry_run;                   /* 1: no device work (host logic only;
                                          buffers may be NULL)            */
    int64_t max_launch_frames;         /* frames one launch takes at most
                                          (whole requests, at least one);
                                          0 = no cap                      */
} ck_serve_config;

typedef struct ck_serve_slot {  /* where an admitted request's frames live */
    float* llrs;
    float* app;
    int8_t* bits;
    int32_t* iterations;
    uint8_t* valid;
    int64_t first;              /* first frame's index in the ring       */
} ck_serve_slot;

/* Bytes of one launch workspace for this configuration (the rings'
 * buffers may still be NULL); -1 on error. */
int64_t ck_serve_workspace_bytes(const ck_serve_config* config);

/* workspaces: max_inflight device pointers (ignored when dry_run). */
int32_t ck_serve_create(const ck_serve_config* config, void* const* workspaces,
                        void* stream, ck_serve** out);

/* Reserves `frames` contiguous ring slots of `code` for request `tag`
 * (unique among live requests); CK_EBUSY when they do not fit now. */
int32_t ck_serve_admit(ck_serve* server, int32_t code, int64_t frames, uint64_t tag,
                       ck_serve_slot* where);

/* The request's LLRs are in place (written on, or ordered before work on,
 * the stream given to ck_serve_create; launches wait for that stream's
 * work, then run on streams of their own). */
int32_t ck_serve_submit(ck_serve* server, uint64_t tag);

/* Launches the submitted requests not yet launched, oldest submission
 * first, up to max_launch_frames frames and 8 contiguous runs, if a
 * workspace is free; *launched_frames = frames launched (0: nothing to
 * launch, or every workspace in flight).  A launch completes when its own
 * work has finished, independently of launches started before it. */
int32_t ck_serve_step(ck_serve* server, int64_t* launched_frames);

/* Up to max_tags requests completed since the last poll (non-blocking). */
int32_t ck_serve_poll(ck_serve* server, uint64_t* tags, int32_t max_tags, int32_t* count);

/* The caller is done with a completed request's outputs; its slots return
 * to the ring (in admission order). */
int32_t ck_serve_release(ck_serve* server, uint64_t tag);

/* Ring slots in use and capacity of `code`, launches in flight. */
int32_t ck_serve_usage(const ck_serve* server, int32_t code, int64_t* in_use,
                       int64_t* capacity, int32_t* inflight);

/* Waits for launches in flight, frees the engine's host resources. */
int32_t ck_serve_destroy(ck_serve* server);

/* ---------------------------------------------------------------------
 * Constant-composition distribution matching (the amplitude matcher of
 * probabilistic amplitude shaping).  `composition` is a HOST array of
 * num_points (1..64) non-negative counts summing to the block length n
 * (<= 2^30); the other arrays are device arrays.
 * ------------------------------------------------------------------- */

/* Matches each row of `bits` (rows, k) int8 0/1 to n point indices,
 * `out` (rows, n) int32: the range coder of ccdm_match
 * (commkit/coding/_numba_loops.py:944-1004; CCDM._match,
 * commkit/coding/shaping.py:178-184). */
int32_t ck_ccdm_match(const int8_t* bits, int64_t rows, int64_t k,
                      const int64_t* composition, int32_t num_points, int64_t n,
                      int32_t* out, void* stream);

/* Device workspace of ck_ccdm_dematch: buffer_bits / 32 + 2 32-bit limbs a
 * row (-1 on bad arguments). */
int64_t ck_ccdm_dematch_workspace_bytes(int64_t rows, int64_t buffer_bits);

/* Dematches each row of `symbols` (rows, n) int32 to `bits` (rows, k) int8
 * and `valid` (rows,) uint8 (0: not a sequence the matcher emits, or its
 * interval needs more than buffer_bits fractional bits): ccdm_dematch
 * (commkit/coding/_numba_loops.py:1010-1102; CCDM._dematch,
 * commkit/coding/shaping.py:186-198). */
int32_t ck_ccdm_dematch(const int32_t* symbols, int64_t rows, int64_t n,
                        const int64_t* composition, int32_t num_points, int64_t k,
                        int64_t buffer_bits, int8_t* bits, uint8_t* valid,
                        void* workspace, void* stream);

/* ---------------------------------------------------------------------
 * Post-FEC scoring.  Device arrays; the outputs are cleared by the call.
 * ------------------------------------------------------------------- */

/* Per row of `bits` and `reference` (rows, n) int8 0/1: bit_errors[r] = the
 * differing bits, block_errors[r] = the blocks of `block` bits (block
 * divides n) with any: ber and fer (commkit/metrics/error_rates.py:30-181).
 * block_errors may be NULL (bit counts only). */
int32_t ck_bit_errors(const int8_t* bits, const int8_t* reference, int64_t rows,
                      int64_t n, int64_t block, int64_t* bit_errors,
                      int64_t* block_errors, void* stream);

/* Per row of `llrs` (rows, n) float32 and the sent `bits` (rows, n) int8:
 * loss[r] = mean over the row of log2(1 + e^{-(1 - 2c) L}) (float64), the
 * term behind gmi = H(X) - k loss and ngmi = 1 - loss
 * (commkit/metrics/information.py:34-89). */
int32_t ck_gmi_loss(const float* llrs, const int8_t* bits, int64_t rows, int64_t n,
                    double* loss, void* stream);

/* ---------------------------------------------------------------------
 * Device front end: LDPC encoding and AWGN.
 * ------------------------------------------------------------------- */

/* A systematic LDPC encoder: codeword = [u, p] with the syndrome of the
 * information part s = H_info u and the parity p = Hp^-1 s over GF(2)
 * (commkit/coding/ldpc.py:445-536 builds the same systematic form). Device
 * tables. */
typedef struct ck_ldpc_encoder {
    int32_t k;                    /* information bits                        */
    int32_t m;                    /* checks = parity bits of the mother code */
    int32_t n_out;                /* bits written per codeword: k + m - punctured */
    int32_t words;                /* 64-bit words of a syndrome, ceil(m / 64) */
    const int32_t* info_row_ptr;  /* (m + 1) CSR of H over the info columns  */
    const int32_t* info_cols;     /* their columns, < k                      */
    const uint64_t* inverse_t;    /* (words, m): bit b of [w][j] is
                                     (Hp^-1)[j][64 w + b]; NULL for a
                                     dual-diagonal Hp (p = prefix XOR of s)  */
} ck_ldpc_encoder;

/* Encode `rows` codewords: info (rows, k) int8 0/1 to words (rows, n_out)
 * int8.  Replaces the encoder of commkit/coding/ldpc.py:445-536 for codes
 * with an invertible or dual-diagonal parity part. */
int32_t ck_ldpc_encode(const ck_ldpc_encoder* enc, const int8_t* info, int64_t rows,
                       int8_t* words, void* stream);

/* out = x + noise_std[row] * n with n standard normal (circular complex of
 * unit variance per component for complex samples), x (rows, len) float32
 * or interleaved complex64 (is_complex), noise_std (rows) float32 device.  Normals
 * come from Philox4x32-10 (key = seed, counter = (group, stream)), four
 * uniforms per call turned into four normals by Box-Muller; element e of
 * the flat array takes normal e of the draw (real), or normals 2e and
 * 2e + 1 (complex), so results depend only on (seed, stream) and the
 * element's index.  out may alias x.  The draws of commkit/_random.py:29-60
 * and commkit/impairments/noise.py:638-672, on the device. */
int32_t ck_awgn(const void* x, int64_t rows, int64_t len, int32_t is_complex,
                const float* noise_std, uint64_t seed, uint64_t stream_id, void* out,
                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CK_TAIL_H */
