// Internals of the batched decoder shared with the serving engine
// (ck_serve.hip): the launch shape of a set of classes, the descriptor a
// launch reads, and the launch itself.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "ck_tail.h"

namespace ckbatch {

constexpr int kMaxClasses = 8;

struct Shape {
    int grid;              // workgroups (one per CU)
    size_t lds_bytes;      // dynamic LDS of a workgroup
    size_t slot_bytes;     // check-state workspace of a workgroup
    int F[kMaxClasses];    // frame slots per workgroup, per class
    int P[kMaxClasses];    // lanes per check row, per class
    int S[kMaxClasses];    // 1: the class's min-sum check states live in LDS
    int variant;           // kernel: one-lane rows (0 none, 1 short, 2 long); 3 none, and
                           // every two-lane class's rows of at most 28 edges
};

// Bytes of the descriptor at the start of a launch's workspace; the
// workgroups' slot workspace follows it (grid * slot_bytes).
size_t desc_bytes();

int32_t shape(const ck_batch_class* classes, int32_t num_classes, const ck_ldpc_method* method,
              Shape* out);

// Writes the descriptor of a launch over `classes` (whose workspace is the
// device buffer `workspace`) into host memory `host_desc` (desc_bytes()).
void describe(const ck_batch_class* classes, int32_t num_classes, const ck_ldpc_method* method,
              const Shape& shape, void* workspace, void* host_desc);

// Raises the kernels' dynamic LDS limit when needed (not under capture).
void allow_lds(size_t lds_bytes);

// Clears the queue counters and launches the decode on `stream`; the
// descriptor must be in `workspace` by then (stream order).
int32_t launch(void* workspace, int grid, size_t lds_bytes, bool sum_product, int one_lane,
               hipStream_t stream);

}  // namespace ckbatch
