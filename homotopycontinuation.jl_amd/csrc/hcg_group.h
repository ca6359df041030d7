// Host interface of the lane-group tracker's slot programs (group.cpp); the kernel launcher is
// declared in hcg_group_launch.h.
#pragma once
#include <string>
#include <vector>

#include "device/hcg_group_types.h"
#include "device/hcg_kargs.h"
#include "hcg_codegen.h"
#include "hcgpu.h"

namespace hcg {

// Compile-time system size of the kernel instance that tracks systems of size n (0: none).
int group_nn_for(int n);
// Lanes per group of a kernel instance: 16 up to nn = 16, 32 for nn = 24.
int group_lanes_for(int nn);
int group_mlp_lds_bytes();
int group_block();

// The slot programs of one or more homotopies, uploaded to the device.
struct GroupPrograms {
    int nn = 0;            // kernel instance
    int lds_per_path = 0;  // bytes of LDS per path (largest over the systems)
    int nsys = 0;
    std::vector<hcg_gsys> host;  // with device pointers (host pointers into host_tables without upload)
    std::vector<char> host_tables;
    void* d_tables = nullptr;    // every table of every system, one allocation
    hcg_gsys* d_sys = nullptr;
    // statistics per system: slots and registers of the eval_jac / eval programs, instructions
    std::vector<int> ej_slots, ej_regs, e_slots, e_regs, ej_instr, e_instr;
};

// Schedule the tapes of hs[0..nsys) and, when upload is set, copy the tables to the current
// device.  Returns false with *err.
bool group_build(const hcg_homotopy_desc* const* hs, int nsys, bool upload, GroupPrograms* out, std::string* err);
void group_free(GroupPrograms* g);

}  // namespace hcg
