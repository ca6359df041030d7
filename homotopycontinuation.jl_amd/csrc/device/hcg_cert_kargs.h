// Kernel argument block of hcg_certify_kernel (shared by the host runtime and the
// hipRTC-compiled device code; plain C types only).
#pragma once
#include "hcgpu_types.h"

typedef struct hcg_cert_kargs {
    int32_t nsol;
    int32_t refine;               // 1: Newton-refine each candidate first
    int32_t real_system;          // 1: F maps real points (at these parameters) to real values
    int32_t extended;             // 1: retry a failed test with a double-double residual enclosure
    const double* candidates;     // nsol * n complex
    const double* params;         // per group: nparams complex (point), or NULL
    const double* params_ival;    // per group: nparams * 4 (re lo, re hi, im lo, im hi), or NULL
    const double* params_ball;    // per group: nparams * 6 (re hi, re lo, re rad, im hi, im lo, im rad), or NULL
    const int32_t* group_of;      // nsol group indices (checked by the host) or NULL (group 0)
    hcg_certificate* out;         // nsol records
    double* box1;                 // nsol * n * 4: the certified box x₁
    double* box0;                 // nsol * n * 4: the inflated box x₀
    double* refined;              // nsol * n complex: the refined candidate x̃
} hcg_cert_kargs;
