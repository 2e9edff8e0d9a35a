// Optional per-kernel-class timing with HIP events on the launching stream.
// Off by default; scirs_gpu_prof_begin()/scirs_gpu_prof_end() switch it on
// and read the totals.  Each scope records an event pair around the launches
// it covers plus the algorithmic work they do (flops and HBM bytes).
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/scirs_gpu.h"   // SCIRS_PROF_NCAT, scirs_prof_stats

enum {
    SCIRS_PROF_PANEL = 0,     // cooperative panel factorisation
    SCIRS_PROF_LASWP = 1,     // row interchanges outside the panel
    SCIRS_PROF_TRSM = 2,      // U12 = L11^-1 A12
    SCIRS_PROF_UPDATE = 3,    // trailing (Schur) update on the matrix cores
    SCIRS_PROF_TRSV = 4,      // substitution after the factorisation
    SCIRS_PROF_RESIDUAL = 5,  // FP64 residual / norms
    SCIRS_PROF_CONVERT = 6,   // FP64 -> FP32 conversion (+ row sums)
    SCIRS_PROF_OTHER = 7,
};

namespace scirs {
bool prof_enabled();
struct ProfScope {
    int cat;
    double flops, bytes;
    hipStream_t s;
    hipEvent_t e0 = nullptr;
    ProfScope(int cat_, double flops_, double bytes_, hipStream_t s_);
    ~ProfScope();
};
} // namespace scirs
