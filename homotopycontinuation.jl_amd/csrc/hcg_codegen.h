// Internal interface of the tape -> HIP code generator.
#pragma once
#include <string>

#include "hcgpu.h"

#define HCG_MAX_N 24

namespace hcg {

// Static FP64 flop counts of the generated functions (one call each).
struct Flops {
    long eval = 0, eval_jac = 0, eval_dd = 0;
    long taylor[4] = {0, 0, 0, 0};
};

bool validate_tape(const hcg_tape* tp, int n, std::string* err);
bool generate_source(const hcg_homotopy_desc* h, std::string* src, Flops* fl, std::string* err);
// Certifier for a square system: point, double-double and interval evaluations plus the
// Krawczyk kernel (device/hcg_certify.h).
bool generate_certify_source(const hcg_system* F, int n, std::string* src, std::string* err);
bool homotopy_flops(const hcg_homotopy_desc* h, long* fl6, std::string* err);
bool generate_multi_source(const hcg_homotopy_desc* const* hs, int nsys, std::string* src, Flops* fls,
                           std::string* err);

}  // namespace hcg
