// Host side of the lane-group tracker: slot programs of each system's tapes, packed into one
// device allocation per program (hcg_gsys records followed by their tables).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "group_sched.h"
#include "hcg_codegen.h"
#include "hcg_group.h"

namespace hcg {
namespace {

// Byte image of every table; pointers are fixed up once the device base is known.
struct Packer {
    std::vector<char> bytes;
    size_t put(const void* p, size_t n) {  // p == nullptr: reserve n zero bytes
        size_t off = (bytes.size() + 15) & ~(size_t)15;
        bytes.resize(off + n, 0);
        if (n && p) std::memcpy(bytes.data() + off, p, n);
        return off;
    }
};

struct ProgOffsets {
    size_t op, arg, out, consts;
    size_t var_reg[2], u_ref[2], U_ref[2];
    bool has_U;
};

ProgOffsets pack(Packer& pk, const GProgram& P, int n, bool with_U) {
    ProgOffsets o;
    std::vector<uint8_t> op = P.op;
    op.resize(std::max<size_t>(op.size(), 1));
    o.op = pk.put(op.data(), op.size());
    o.arg = pk.put(P.arg.data(), P.arg.size() * sizeof(int32_t));
    o.out = pk.put(P.out.data(), P.out.size() * sizeof(int32_t));
    o.consts = pk.put(P.consts.data(), P.consts.size() * sizeof(double));
    o.has_U = with_U;
    for (size_t t = 0; t < 2; t++) {
        o.var_reg[t] = o.u_ref[t] = o.U_ref[t] = 0;
        if (t >= P.io.size()) continue;
        o.var_reg[t] = pk.put(P.io[t].var_reg.data(), P.io[t].var_reg.size() * sizeof(int32_t));
        o.u_ref[t] = pk.put(P.io[t].u_ref.data(), (size_t)n * sizeof(int32_t));
        if (with_U) o.U_ref[t] = pk.put(P.io[t].U_ref.data(), (size_t)n * n * sizeof(int32_t));
    }
    return o;
}

void fix(hcg_gprog* g, const GProgram& P, const ProgOffsets& o, char* base) {
    g->nslots = P.nslots;
    g->nregs = P.nregs;
    g->ntapes = (int32_t)P.io.size();
    g->pad_ = 0;
    g->op = base ? (const uint8_t*)(base + o.op) : nullptr;
    g->arg = base ? (const int32_t*)(base + o.arg) : nullptr;
    g->out = base ? (const int32_t*)(base + o.out) : nullptr;
    g->consts = base ? (const double*)(base + o.consts) : nullptr;
    for (int t = 0; t < 2; t++) {
        const bool have = t < (int)P.io.size() && base;
        g->var_reg[t] = have ? (const int32_t*)(base + o.var_reg[t]) : nullptr;
        g->u_ref[t] = have ? (const int32_t*)(base + o.u_ref[t]) : nullptr;
        g->U_ref[t] = have && o.has_U ? (const int32_t*)(base + o.U_ref[t]) : nullptr;
    }
}

}  // namespace

int group_nn_for(int n) {
    if (n <= 4) return 4;
    if (n <= 8) return 8;
    if (n <= 12) return 12;
    if (n <= 16) return 16;
    if (n <= 24) return 24;  // the library's largest system size
    return 0;
}
int group_lanes_for(int nn) { return nn > 16 ? 32 : HCG_GROUP_LANES; }
int group_mlp_lds_bytes() { return ((HCG_MLP_NPARAMS * 2 + 15) / 16) * 16; }  // BF16 weights, 16-byte aligned
int group_block() { return 256; }

bool group_build(const hcg_homotopy_desc* const* hs, int nsys, bool upload, GroupPrograms* out, std::string* err) {
    *out = GroupPrograms();
    if (nsys < 1 || nsys > HCG_MAX_SYSTEMS) { *err = "between 1 and 8 homotopies per program"; return false; }
    std::vector<GProgram> ej(nsys), e(nsys);
    std::vector<long> fl((size_t)6 * nsys);
    int maxn = 0, lds = 512;  // the learned proposer's hidden layers use 512 bytes of a path's LDS
    for (int s = 0; s < nsys; s++) {
        const hcg_homotopy_desc* h = hs[s];
        if (!h) { *err = "homotopy is NULL"; return false; }
        if (!homotopy_flops(h, &fl[6 * s], err)) return false;
        const int n = h->n;
        if (n > 24) { *err = "the lane-group layout tracks systems of up to 24 variables"; return false; }
        maxn = std::max(maxn, n);
        std::vector<const hcg_tape*> tj, te;
        std::vector<const double*> fx;
        if (h->kind == HCG_STRAIGHT_LINE) {
            tj = {h->start.eval_jac, h->target.eval_jac};
            te = {h->start.eval, h->target.eval};
            fx = {h->start.fixed_params, h->target.fixed_params};
        } else {
            tj = {h->target.eval_jac};
            te = {h->target.eval};
            fx = {h->target.fixed_params};
        }
        if (!schedule_tapes(tj, fx, n, &ej[s], err) || !check_program(tj, fx, ej[s], err)) return false;
        if (!schedule_tapes(te, fx, n, &e[s], err) || !check_program(te, fx, e[s], err)) return false;
        lds = std::max(lds, ej[s].nregs * 16);
        lds = std::max(lds, e[s].nregs * 64);  // Taylor series of order 3: 4 complex per register
        out->ej_slots.push_back(ej[s].nslots);
        out->ej_regs.push_back(ej[s].nregs);
        out->ej_instr.push_back(ej[s].ninstr);
        out->e_slots.push_back(e[s].nslots);
        out->e_regs.push_back(e[s].nregs);
        out->e_instr.push_back(e[s].ninstr);
    }
    out->nn = group_nn_for(maxn);
    out->lds_per_path = (lds + 15) & ~15;
    out->nsys = nsys;
    const size_t lds_total = (size_t)group_mlp_lds_bytes() + (size_t)(group_block() / group_lanes_for(out->nn)) * out->lds_per_path;
    if (lds_total > 160 * 1024) {
        *err = "the slot programs need " + std::to_string(out->lds_per_path) + " bytes of LDS per path, more than a "
               "workgroup of 16 paths can hold";
        return false;
    }
    // pack: the hcg_gsys records first, then every table
    Packer pk;
    const size_t sys_off = pk.put(nullptr, sizeof(hcg_gsys) * nsys);
    std::vector<ProgOffsets> oj(nsys), oe(nsys);
    for (int s = 0; s < nsys; s++) {
        oj[s] = pack(pk, ej[s], hs[s]->n, true);
        oe[s] = pack(pk, e[s], hs[s]->n, false);
    }
    char* dbase = nullptr;
    if (upload) {
        if (hipMalloc((void**)&dbase, pk.bytes.size()) != hipSuccess) { *err = "hipMalloc of the slot programs failed"; return false; }
    } else {  // the tables stay on the host (schedule statistics, CPU checks)
        out->host_tables = pk.bytes;
        dbase = out->host_tables.data();
    }
    out->host.resize(nsys);
    for (int s = 0; s < nsys; s++) {
        const hcg_homotopy_desc* h = hs[s];
        hcg_gsys& g = out->host[s];
        std::memset(&g, 0, sizeof g);
        g.n = h->n;
        g.kind = h->kind;
        g.npar = h->kind == HCG_PARAMETER ? h->target.eval->nparams : 0;
        g.lds_bytes = out->lds_per_path;
        g.gamma_re = h->gamma_re;
        g.gamma_im = h->gamma_im;
        g.fl_eval = (double)fl[6 * s + 0];
        g.fl_ej = (double)fl[6 * s + 1];
        g.fl_dd = (double)fl[6 * s + 2];
        g.fl_t1 = (double)fl[6 * s + 3];
        g.fl_t2 = (double)fl[6 * s + 4];
        g.fl_t3 = (double)fl[6 * s + 5];
        fix(&g.ej, ej[s], oj[s], dbase);
        fix(&g.e, e[s], oe[s], dbase);
    }
    std::memcpy(pk.bytes.data() + sys_off, out->host.data(), sizeof(hcg_gsys) * nsys);
    if (!upload) std::memcpy(out->host_tables.data() + sys_off, out->host.data(), sizeof(hcg_gsys) * nsys);
    if (upload) {
        if (hipMemcpy(dbase, pk.bytes.data(), pk.bytes.size(), hipMemcpyHostToDevice) != hipSuccess) {
            hipFree(dbase);
            *err = "copy of the slot programs failed";
            return false;
        }
        out->d_tables = dbase;
        out->d_sys = (hcg_gsys*)(dbase + sys_off);
    }
    return true;
}

void group_free(GroupPrograms* g) {
    if (g->d_tables) hipFree(g->d_tables);
    g->d_tables = nullptr;
    g->d_sys = nullptr;
}

}  // namespace hcg
