// Device-side description of a lane-group program (shared by the host runtime and the
// precompiled lane-group tracker kernel; plain C types only).
//
// A system's tapes become two slot programs (group_sched.h): `ej` evaluates H and its
// Jacobian (the evaluate_and_jacobian! tapes), `e` evaluates H (the evaluate! tapes, also
// run in double-double and on truncated Taylor series).  For a straight-line homotopy each
// program holds G's tape (index 0) and F's tape (index 1); for a parameter homotopy F's only.
#pragma once
#include "hcgpu_types.h"

#define HCG_GROUP_LANES 16

// operand reference kinds (bits 28..31) and index (bits 0..27)
#define HCG_GREF_KIND(r) ((int)((unsigned)(r) >> 28))
#define HCG_GREF_INDEX(r) ((r) & 0x0FFFFFFF)
#define HCG_GK_REG 0
#define HCG_GK_CONST 1
#define HCG_GK_PARAM 2
#define HCG_GK_ZERO 3
#define HCG_GK_INT 4
#define HCG_GK_NONE 5

typedef struct hcg_gprog {
    int32_t nslots, nregs, ntapes, pad_;
    const uint8_t* op;                /* nslots OpType codes */
    const int32_t* arg;               /* nslots * 16 * 4 operand references */
    const int32_t* out;               /* nslots * 16 result registers, -1: none */
    const double* consts;             /* constant table, (re, im) pairs */
    const int32_t* var_reg[2];        /* per tape: n registers the variables are loaded into (-1: unused) */
    const int32_t* u_ref[2];          /* per tape: n output references */
    const int32_t* U_ref[2];          /* per tape: n * n references, row-major (ej only) */
} hcg_gprog;

typedef struct hcg_gsys {
    int32_t n, kind, npar, lds_bytes;  /* lds_bytes: LDS per path (largest program x value size) */
    double gamma_re, gamma_im;
    double fl_eval, fl_ej, fl_dd, fl_t1, fl_t2, fl_t3;  /* counted FP64 flops per evaluation */
    hcg_gprog ej, e;
} hcg_gsys;
