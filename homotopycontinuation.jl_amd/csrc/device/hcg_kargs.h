// Kernel argument block of hcg_track_kernel (shared by the host runtime and the
// hipRTC-compiled device code; plain C types only).
#pragma once
#include "hcgpu_types.h"

// Waves per SIMD the tracker kernels are compiled for (register budget 512 / this).
#ifndef HCG_WAVES_PER_EU
#define HCG_WAVES_PER_EU 1
#endif

typedef struct hcg_kargs {
    int32_t npaths;
    int32_t log_capacity;
    int32_t ngroups;
    int32_t pad_;
    const double* starts;          // npaths * n complex
    double t1, t0;
    const double* start_params;    // nparams complex (PARAMETER)
    const double* target_params;   // ngroups * nparams complex
    const int32_t* group_of_path;  // npaths or NULL
    int32_t* next_path;            // device work-queue head (zeroed before launch)
    const hcg_options* opts;       // device copy
    hcg_path_result* results;
    double* solutions;
    double* last_points;
    double* valuations;
    unsigned long long* counters;  // [0] tracker steps, [1] Newton iterations, [2] counted FP64 flops, [3] learned proposals accepted
    float* log_records;            // step log (HCG_STEP_RECORD_FLOATS per record) or NULL
    int32_t* log_count;
    const float* mlp;              // learned proposer weights (FP32 master copy) or NULL
} hcg_kargs;

// Per-lane totals, added to the launch counters when the lane exits.
typedef struct LaneTotals {
    long long steps, newton, lacc;
    double flops;
} LaneTotals;

// Several systems in one launch: paths [offset[s], offset[s+1]) belong to system s,
// whose inputs and outputs are described by sys[s]; one work queue over all of them.
#define HCG_MAX_SYSTEMS 8
typedef struct hcg_multi_kargs {
    int32_t nsys, total;
    int32_t offset[HCG_MAX_SYSTEMS + 1];
    int32_t* next_path;
    hcg_kargs sys[HCG_MAX_SYSTEMS];
} hcg_multi_kargs;
