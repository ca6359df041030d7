// Tape -> HIP source generator (host side of hcg_program_create).
//
// Lowers each straight-line program (ModelKit InstructionSequence layout,
// src/model_kit/instruction_sequence.jl:1-27) to straight-line device functions:
//   *_eval      u = F(x)                    (compiled_execute_impl, compiled_system_homotopy.jl:178-243)
//   *_eval_jac  u = F(x), U = dF/dx         (evaluate_and_jacobian!, :277-289)
//   *_eval_dd   F(x) in double-double       (interpreted_system.jl:65-70)
//   *_taylorK   orders K-1 and K of F(x(s)) (execute_taylor!, instruction_interpreter.jl:444-491)
// Taylor code is emitted per coefficient: values carry the highest order that can be
// non-zero (constants 0, parameters 0 or 1, variables K-1), and products only emit
// the terms that are not structurally zero.  The accumulation order of every
// remaining term is the oracle's, so the results agree bit for bit apart from the
// sign of exact zeros.
#include "hcg_codegen.h"

#include <array>
#include <climits>
#include <functional>
#include <map>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <string>
#include <vector>

namespace hcg {

enum {
    OP_STOP = 0, OP_CB, OP_ACOS, OP_ASIN, OP_COS, OP_COSH, OP_EXP, OP_INV, OP_INV_NOT_ZERO, OP_INVSQR, OP_NEG,
    OP_SIN, OP_SINH, OP_SQR, OP_SQRT, OP_TAN, OP_TANH, OP_IDENTITY, OP_ADD, OP_DIV, OP_MUL, OP_SUB, OP_POW_INT,
    OP_POW, OP_ADD3, OP_MUL3, OP_MULADD, OP_MULSUB, OP_SUBMUL, OP_ADD4, OP_MUL4, OP_MULMULADD, OP_MULMULSUB
};

static int arity(int op) {
    switch (op) {
    case OP_STOP: return 0;
    case OP_CB: case OP_INV: case OP_INV_NOT_ZERO: case OP_INVSQR: case OP_NEG: case OP_SQR: case OP_IDENTITY: return 1;
    case OP_ADD: case OP_DIV: case OP_MUL: case OP_SUB: case OP_POW_INT: return 2;
    case OP_ADD3: case OP_MUL3: case OP_MULADD: case OP_MULSUB: case OP_SUBMUL: return 3;
    case OP_ADD4: case OP_MUL4: case OP_MULMULADD: case OP_MULMULSUB: return 4;
    default: return -1;
    }
}

static std::string dlit(double v) {
    if (std::isnan(v)) return "__builtin_nan(\"\")";
    if (std::isinf(v)) return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", v);
    return buf;
}
static std::string clit(double re, double im) { return "mkc(" + dlit(re) + ", " + dlit(im) + ")"; }

bool validate_tape(const hcg_tape* tp, int n, std::string* err) {
    if (!tp || !tp->instr) { *err = "null tape"; return false; }
    if (tp->t_index != 0) { *err = "homotopy tapes (continuation parameter in the tape) are not supported"; return false; }
    if (tp->nvars != n) { *err = "tape variable count differs from the homotopy size"; return false; }
    if (tp->output_dim != n) { *err = "only square systems are supported"; return false; }
    if (tp->nconst < 0 || tp->nparams < 0 || tp->tape_space < tp->nconst + tp->nparams + tp->nvars) { *err = "bad tape ranges"; return false; }
    if (tp->ninstr < 0 || tp->nu < 0 || tp->nU < 0 || (tp->nu > 0 && !tp->u_assign) || (tp->nU > 0 && !tp->U_assign) ||
        (tp->nconst > 0 && !tp->constants)) { *err = "bad tape counts"; return false; }
    if (tp->tape_space > (1 << 22) || tp->ninstr > (1 << 22)) { *err = "tape too large"; return false; }
    if (tp->params_first != tp->nconst + 1 || tp->vars_first != tp->nconst + tp->nparams + 1) { *err = "unexpected tape block layout"; return false; }
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        int ar = arity(op);
        if (ar < 0) { *err = "unsupported SLP op " + std::to_string(op) + " (transcendental ops are not lowered)"; return false; }
        if (op == OP_STOP) break;
        for (int k = 0; k < ar; k++) {
            if (op == OP_POW_INT && k == 1) continue;
            if (I[k] < 1 || I[k] > tp->tape_space) { *err = "tape input index out of range"; return false; }
        }
        if (I[5] < 1 || I[5] > tp->tape_space) { *err = "tape output index out of range"; return false; }
        if (I[5] <= tp->nconst + tp->nparams + tp->nvars) { *err = "instruction overwrites an input slot"; return false; }
    }
    for (int i = 0; i < tp->nu; i++)
        if (tp->u_assign[2 * i] < 1 || tp->u_assign[2 * i] > n || tp->u_assign[2 * i + 1] < 1 || tp->u_assign[2 * i + 1] > tp->tape_space) { *err = "bad u assignment"; return false; }
    for (int i = 0; i < tp->nU; i++)
        if (tp->U_assign[2 * i] < 1 || tp->U_assign[2 * i] > n * n || tp->U_assign[2 * i + 1] < 1 || tp->U_assign[2 * i + 1] > tp->tape_space) { *err = "bad U assignment"; return false; }
    std::vector<char> seen(n + (size_t)n * n, 0);
    for (int i = 0; i < tp->nu; i++)
        if (seen[tp->u_assign[2 * i] - 1]++) { *err = "u entry assigned twice"; return false; }
    for (int i = 0; i < tp->nU; i++)
        if (seen[n + tp->U_assign[2 * i] - 1]++) { *err = "U entry assigned twice"; return false; }
    return true;
}

enum ParamMode { PM_NONE, PM_FIXED, PM_DYNAMIC };

// ---------------------------- ComplexF64 / DD eval ----------------------------
// Instructions that recompute a value an earlier instruction of the same tape computed
// (same operation on the same values, by value numbering).  A tape producer repeats cheap
// products on purpose (rematerialization: shorter live ranges); the generated code passes
// the operands of such an instruction through hcg_opaque so that the compiler does not
// merge it back into the first computation, and its operations are not counted as
// algorithmic flops.
static std::vector<char> recomputed(const hcg_tape* tp) {
    std::vector<long> vn(tp->tape_space + 1);
    for (int sl = 0; sl <= tp->tape_space; sl++) vn[sl] = sl;
    std::map<std::array<long, 5>, long> seen;
    long next = tp->tape_space + 1;
    std::vector<char> dup(tp->ninstr, 0);
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        const int op = I[4];
        if (op == OP_STOP) break;
        std::array<long, 5> key = {op, -1, -1, -1, -1};
        for (int q = 0; q < arity(op); q++) key[q + 1] = (op == OP_POW_INT && q == 1) ? -2L - I[1] : vn[I[q]];
        auto it = seen.find(key);
        if (it != seen.end()) {
            dup[i] = 1;
            vn[I[5]] = it->second;
        } else {
            seen[key] = next;
            vn[I[5]] = next++;
        }
    }
    return dup;
}

// How an output is written: store(kind 'u' or 'U', index, value expression) -> statement.
// The default writes the value; the fused straight-line Jacobian combines it with the
// other system's output instead.
typedef std::function<std::string(char, int, const std::string&)> StoreFn;

static std::string plain_store(char kind, int idx, const std::string& v) {
    return std::string(1, kind) + "[" + std::to_string(idx) + "] = " + v + ";";
}

// Straight-line body of one tape (no signature): inputs x (and pp, pq, pt when pm is
// PM_DYNAMIC), outputs through `store`; local names carry `pfx` so that two bodies can
// share one function.  Returns the counted FP64 flops.
static long gen_eval_body(std::ostringstream& os, const hcg_tape* tp, bool jac, bool dd, int n, ParamMode pm,
                          const double* fixed, const std::string& pfx, const StoreFn& store);

static void gen_eval(std::ostringstream& os, const hcg_tape* tp, const std::string& fname, bool jac, bool dd, int n,
                     ParamMode pm, const double* fixed, long* flops) {
    const char* T = dd ? "cdd" : "cplx";
    os << "__device__ __noinline__ void " << fname << "(const " << T << "* __restrict__ x, ";
    if (pm == PM_DYNAMIC) os << "const double* __restrict__ pp, const double* __restrict__ pq, cplx pt, ";
    os << (dd ? "cdd" : "cplx") << "* __restrict__ u";
    if (jac) os << ", cplx* __restrict__ U";
    os << ") {\n";
    *flops = gen_eval_body(os, tp, jac, dd, n, pm, fixed, "", plain_store);
    os << "}\n";
}

static long gen_eval_body(std::ostringstream& os, const hcg_tape* tp, bool jac, bool dd, int n, ParamMode pm,
                          const double* fixed, const std::string& pfx, const StoreFn& store) {
    const char* T = dd ? "cdd" : "cplx";
    std::vector<std::string> nm(tp->tape_space + 1);
    auto wrap = [&](const std::string& c) { return dd ? "cdd_from(" + c + ")" : c; };
    for (int k = 0; k < tp->nconst; k++)
        nm[k + 1] = wrap(clit(tp->constants[2 * k], tp->constants[2 * k + 1]));
    // parameters and variables are defined right before their first use, so that their live
    // ranges (and the register pressure of long straight-line bodies) stay short
    std::vector<std::string> lazy_def(tp->tape_space + 1);
    for (int k = 0; k < tp->nparams; k++) {
        std::string v = pfx + "P" + std::to_string(k);
        if (pm == PM_FIXED) lazy_def[tp->params_first + k] = "  const " + std::string(T) + " " + v + " = " + wrap(clit(fixed[2 * k], fixed[2 * k + 1])) + ";\n";
        else lazy_def[tp->params_first + k] = "  const " + std::string(T) + " " + v + " = " + wrap("par_lerp(pt, pp, pq, " + std::to_string(k) + ")") + ";\n";
        nm[tp->params_first + k] = v;
    }
    for (int k = 0; k < tp->nvars; k++) {
        std::string v = pfx + "X" + std::to_string(k);
        lazy_def[tp->vars_first + k] = "  const " + std::string(T) + " " + v + " = x[" + std::to_string(k) + "];\n";
        nm[tp->vars_first + k] = v;
    }
    auto need = [&](int slot) {
        if (slot >= 1 && slot <= tp->tape_space && !lazy_def[slot].empty()) { os << lazy_def[slot]; lazy_def[slot].clear(); }
    };
    const char* pre = dd ? "cdd_" : "c_";
    long f = pm == PM_DYNAMIC ? 6L * tp->nparams : 0;  // p(t) = t p + (1 - t) q per parameter
    long dm = dd ? 58 : 6, da = dd ? 22 : 2;
    // every output is stored right after the last instruction writing its slot (short live
    // ranges); outputs read straight from inputs or constants are stored at the end
    std::vector<int> last_write(tp->tape_space + 1, -1);
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        if (I[4] == OP_STOP) break;
        last_write[I[5]] = i;
    }
    struct OutStore { char kind; int idx, slot; };
    std::vector<std::vector<OutStore>> stores_after(tp->ninstr + 1);
    std::vector<int> uset(n, 0), Uset(jac ? n * n : 0, 0);
    for (int i = 0; i < tp->nu; i++) {
        int row = tp->u_assign[2 * i] - 1, slot = tp->u_assign[2 * i + 1];
        int at = last_write[slot] >= 0 ? last_write[slot] : tp->ninstr;
        stores_after[at].push_back({'u', row, slot});
        uset[row] = 1;
    }
    if (jac)
        for (int i = 0; i < tp->nU; i++) {
            int l = tp->U_assign[2 * i] - 1, slot = tp->U_assign[2 * i + 1];
            int at = last_write[slot] >= 0 ? last_write[slot] : tp->ninstr;
            stores_after[at].push_back({'U', l, slot});
            Uset[l] = 1;
        }
    auto flush_stores = [&](int at) {
        for (auto& st : stores_after[at]) {
            need(st.slot);
            os << "  " << store(st.kind, st.idx, nm[st.slot]) << "\n";
        }
    };
    const std::vector<char> dup = recomputed(tp);
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        if (op == OP_STOP) break;
        for (int q = 0; q < arity(op); q++)
            if (!(op == OP_POW_INT && q == 1)) need(I[q]);
        std::string a = nm[I[0]], b = arity(op) > 1 && op != OP_POW_INT ? nm[I[1]] : "", c = arity(op) > 2 ? nm[I[2]] : "",
                    d = arity(op) > 3 ? nm[I[3]] : "";
        std::string out = pfx + "T" + std::to_string(i);
        if (dup[i] && !dd) {
            int k = 0;
            for (std::string* opnd : {&a, &b, &c, &d}) {
                if (opnd->empty() || opnd->compare(0, 4, "mkc(") == 0) continue;
                std::string o = out + "o" + std::to_string(k++);
                os << "  const cplx " << o << " = hcg_opaque(" << *opnd << ");\n";
                *opnd = o;
            }
        }
        const long f_before = f;
        std::string e;
        switch (op) {
        case OP_CB: e = std::string(pre) + "cb(" + a + ")"; f += dm * 2; break;
        case OP_INV: e = std::string(pre) + "inv(" + a + ")"; f += dm + da; break;
        case OP_INV_NOT_ZERO: e = std::string(pre) + "inv_not_zero(" + a + ")"; f += dm + da; break;
        case OP_INVSQR: e = std::string(pre) + "invsqr(" + a + ")"; f += 2 * dm + da; break;
        case OP_NEG: e = std::string(pre) + "neg(" + a + ")"; break;
        case OP_SQR: e = std::string(pre) + "sqr(" + a + ")"; f += dd ? 3 * 11 + 2 * 9 : 5; break;
        case OP_IDENTITY: e = a; break;
        case OP_ADD: e = std::string(pre) + "add(" + a + ", " + b + ")"; f += da; break;
        case OP_SUB: e = std::string(pre) + "sub(" + a + ", " + b + ")"; f += da; break;
        case OP_MUL: e = std::string(pre) + "mul(" + a + ", " + b + ")"; f += dm; break;
        case OP_DIV: e = std::string(pre) + "div(" + a + ", " + b + ")"; f += dm + da + 3; break;
        case OP_POW_INT: e = std::string(pre) + "pow_int(" + a + ", " + std::to_string(I[1]) + ")"; f += dm * 2; break;
        case OP_ADD3: e = std::string(pre) + "add3(" + a + ", " + b + ", " + c + ")"; f += 2 * da; break;
        case OP_MUL3: e = std::string(pre) + "mul3(" + a + ", " + b + ", " + c + ")"; f += 2 * dm; break;
        case OP_MULADD: e = std::string(pre) + "muladd(" + a + ", " + b + ", " + c + ")"; f += dm + da; break;
        case OP_MULSUB: e = std::string(pre) + "mulsub(" + a + ", " + b + ", " + c + ")"; f += dm + da; break;
        case OP_SUBMUL: e = std::string(pre) + "submul(" + a + ", " + b + ", " + c + ")"; f += dm + da; break;
        case OP_ADD4: e = std::string(pre) + "add4(" + a + ", " + b + ", " + c + ", " + d + ")"; f += 3 * da; break;
        case OP_MUL4: e = std::string(pre) + "mul4(" + a + ", " + b + ", " + c + ", " + d + ")"; f += 3 * dm; break;
        case OP_MULMULADD: e = std::string(pre) + "mulmuladd(" + a + ", " + b + ", " + c + ", " + d + ")"; f += 2 * dm + da; break;
        case OP_MULMULSUB: e = std::string(pre) + "mulmulsub(" + a + ", " + b + ", " + c + ", " + d + ")"; f += 2 * dm + da; break;
        }
        if (dup[i]) f = f_before;
        os << "  const " << T << " " << out << " = " << e << ";\n";
        nm[I[5]] = out;
        flush_stores(i);
    }
    flush_stores(tp->ninstr);
    for (int i = 0; i < n; i++)
        if (!uset[i]) os << "  " << store('u', i, dd ? "cdd_from(mkc(0.0, 0.0))" : "mkc(0.0, 0.0)") << "\n";
    if (jac)
        for (int i = 0; i < n * n; i++)
            if (!Uset[i]) os << "  " << store('U', i, "mkc(0.0, 0.0)") << "\n";
    return f;
}

// Output values of a tape that are compile-time constants: the literal for each output
// (u: 0..n-1, U: n..n+n*n-1) whose slot is a tape constant or that is never assigned,
// "" for outputs the tape computes.
static std::vector<std::string> constant_outputs(const hcg_tape* tp, int n) {
    std::vector<std::string> c(n + n * n, "mkc(0.0, 0.0)");
    auto lit = [&](int slot) -> std::string {
        if (slot >= 1 && slot <= tp->nconst) return clit(tp->constants[2 * (slot - 1)], tp->constants[2 * (slot - 1) + 1]);
        return "";
    };
    for (int i = 0; i < tp->nu; i++) c[tp->u_assign[2 * i] - 1] = lit(tp->u_assign[2 * i + 1]);
    for (int i = 0; i < tp->nU; i++) c[n + tp->U_assign[2 * i] - 1] = lit(tp->U_assign[2 * i + 1]);
    return c;
}

// Straight-line homotopy H = ts G + tt F (ts = γ t, tt = 1 - t): value and Jacobian in one
// function, each entry combined exactly as c_muladd(ts, G_l, c_mul(tt, F_l)).  F's body runs
// first; entries whose G value is a compile-time constant are combined where F stores them,
// the others are stored raw and combined where G's body stores its value, so no second
// Jacobian is written to memory.
static long gen_straight_eval_jac(std::ostringstream& os, const hcg_system& G, const hcg_system& F, int n,
                                  ParamMode pmG, ParamMode pmF) {
    os << "__device__ __noinline__ void sysH_eval_jac(const cplx* __restrict__ x, cplx ts, cplx tt, "
          "cplx* __restrict__ u, cplx* __restrict__ U) {\n";
    const std::vector<std::string> gc = constant_outputs(G.eval_jac, n);
    auto pos = [n](char kind, int idx) { return kind == 'u' ? idx : n + idx; };
    long f = gen_eval_body(os, F.eval_jac, true, false, n, pmF, F.fixed_params, "f",
                           [&](char kind, int idx, const std::string& v) {
                               const std::string& g = gc[pos(kind, idx)];
                               if (g.empty()) return plain_store(kind, idx, v);
                               return plain_store(kind, idx, "c_muladd(ts, " + g + ", c_mul(tt, " + v + "))");
                           });
    f += gen_eval_body(os, G.eval_jac, true, false, n, pmG, G.fixed_params, "g",
                       [&](char kind, int idx, const std::string& v) -> std::string {
                           if (!gc[pos(kind, idx)].empty()) return "";
                           std::string dst = std::string(1, kind) + "[" + std::to_string(idx) + "]";
                           return dst + " = c_muladd(ts, " + v + ", c_mul(tt, " + dst + "));";
                       });
    os << "}\n";
    return f;
}

// ------------------------ double-double tape interpreter ------------------------
// The double-double residual runs rarely (extended-precision steps and final
// refinement), so it is interpreted like the reference's InterpretedSystem
// (src/model_kit/interpreted_system.jl:65-70, instruction_interpreter.jl:252-332)
// instead of being unrolled: the tape is re-encoded with its slots compacted by a
// linear scan (operand >= 0: register slot, < 0: constant -k-1) and stored in
// __constant__ tables; each lane keeps its live slots in a private array.
// An interpreter program for one tape: 6 ints per instruction (op, four operands, result
// register), operands >= 0 registers, < 0 constants -e-1, >= PBASE parameter e - PBASE;
// registers assigned by a linear scan over the tape's values (the instructions are SSA after
// validation except for slot reuse, so liveness is tracked per definition).
struct InterpPlan {
    std::vector<int> code;
    std::vector<double> consts;         // (re, im) pairs: the tape's constants, then fixed parameters
    std::vector<std::pair<int, int>> loads;  // (register, input index) of the variables read
    std::vector<int> out_operand;       // per output: its operand, INT32_MIN for a zero output
    int nreg = 0, ninstr = 0;
};

static InterpPlan plan_interp(const hcg_tape* tp, const std::vector<std::pair<int, int>>& outputs, int nout,
                              ParamMode pm, const double* fixed, int PBASE) {
    InterpPlan P;
    for (int k = 0; k < 2 * tp->nconst; k++) P.consts.push_back(tp->constants[k]);
    int ninstr = 0;
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        if (I[4] == OP_STOP) break;
        ninstr++;
    }
    P.ninstr = ninstr;
    std::vector<int> def_of(tp->tape_space + 1, -1);   // current definition id of slot
    std::vector<int> def_last(ninstr + tp->nparams + tp->nvars + 1, -1);
    int ndefs = 0;
    for (int k = 0; k < tp->nparams + tp->nvars; k++) {
        int slot = tp->params_first + k;
        def_of[slot] = ndefs++;
    }
    std::vector<std::array<int, 4>> in_defs(ninstr);
    std::vector<int> out_def(ninstr);
    for (int i = 0; i < ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4], ar = arity(op);
        for (int q = 0; q < 4; q++) {
            in_defs[i][q] = -2;  // -2 unused, -1 constant
            if (q >= ar || (op == OP_POW_INT && q == 1)) continue;
            int s = I[q];
            if (s <= tp->nconst) { in_defs[i][q] = -1; continue; }
            in_defs[i][q] = def_of[s];
            def_last[def_of[s]] = i;
        }
        def_of[I[5]] = ndefs;
        out_def[i] = ndefs++;
    }
    def_last.resize(ndefs, -1);
    // outputs keep their definitions alive to the end
    std::vector<int> out_is_def(nout, -1), out_const(nout, -1);
    for (auto& o : outputs) {
        int idx = o.first, s = o.second;
        if (s <= tp->nconst) out_const[idx] = s - 1;
        else { out_is_def[idx] = def_of[s]; def_last[def_of[s]] = ninstr; }
    }
    std::vector<int> reg(ndefs, -1);
    std::vector<int> free_regs;
    int nreg = 0;
    auto alloc = [&]() { if (!free_regs.empty()) { int r = free_regs.back(); free_regs.pop_back(); return r; } return nreg++; };
    // parameters are not copied into registers: fixed ones become constants, the others are
    // fetched when an instruction reads them (operand PBASE + k)
    std::vector<int> par_operand(tp->nparams, 0);
    for (int k = 0; k < tp->nparams; k++) {
        if (pm == PM_FIXED) {
            P.consts.push_back(fixed[2 * k]);
            P.consts.push_back(fixed[2 * k + 1]);
            par_operand[k] = -(int)(P.consts.size() / 2) ;
        } else {
            par_operand[k] = PBASE + k;
        }
    }
    for (int k = tp->nparams; k < tp->nparams + tp->nvars; k++) {
        if (def_last[k] < 0) continue;  // unused input
        reg[k] = alloc();
        P.loads.push_back({reg[k], k});
    }
    for (int i = 0; i < ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        int ops[4] = {0, 0, 0, 0};
        for (int q = 0; q < 4; q++) {
            int d = in_defs[i][q];
            if (d == -2) ops[q] = (op == OP_POW_INT && q == 1) ? I[1] : 0;
            else if (d == -1) ops[q] = -(I[q] - 1) - 1;
            else if (d < tp->nparams) ops[q] = par_operand[d];
            else ops[q] = reg[d];
        }
        // free inputs whose last use is this instruction
        for (int q = 0; q < 4; q++) {
            int d = in_defs[i][q];
            if (d >= tp->nparams && def_last[d] == i && reg[d] >= 0) {
                bool dup = false;
                for (int r = 0; r < q; r++) if (in_defs[i][r] == d) dup = true;
                if (!dup) free_regs.push_back(reg[d]);
            }
        }
        int od = out_def[i];
        int outr = alloc();
        reg[od] = outr;
        if (def_last[od] < 0) free_regs.push_back(outr);  // dead value: slot reusable right away
        P.code.insert(P.code.end(), {op, ops[0], ops[1], ops[2], ops[3], outr});
    }
    P.out_operand.assign(nout, INT32_MIN);
    for (int r = 0; r < nout; r++) {
        if (out_is_def[r] >= 0) P.out_operand[r] = out_is_def[r] < tp->nparams ? par_operand[out_is_def[r]] : reg[out_is_def[r]];
        else if (out_const[r] >= 0) P.out_operand[r] = -out_const[r] - 1;
    }
    P.nreg = nreg;
    return P;
}

static void emit_tables(std::ostringstream& os, const std::string& fname, const InterpPlan& P) {
    os << "__constant__ int " << fname << "_code[] = {";
    for (size_t k = 0; k < P.code.size(); k++) os << (k ? "," : "") << P.code[k];
    if (P.code.empty()) os << "0";
    os << "};\n__constant__ double " << fname << "_const[] = {";
    for (size_t k = 0; k < P.consts.size(); k++) os << (k ? "," : "") << dlit(P.consts[k]);
    if (P.consts.empty()) os << "0.0";
    os << "};\n";
}

static void gen_dd_interp(std::ostringstream& os, const hcg_tape* tp, const std::string& fname, int n, ParamMode pm,
                          const double* fixed, long* flops) {
    std::vector<std::pair<int, int>> outs;
    for (int i = 0; i < tp->nu; i++) outs.push_back({tp->u_assign[2 * i] - 1, tp->u_assign[2 * i + 1]});
    const InterpPlan P = plan_interp(tp, outs, n, pm, fixed, 1 << 24);
    const std::vector<int>& code = P.code;
    const int ninstr = P.ninstr, nreg = P.nreg;
    const auto& loads = P.loads;
    const auto& out_operand = P.out_operand;
    emit_tables(os, fname, P);
    os << "__device__ __noinline__ void " << fname << "(const cdd* __restrict__ x, ";
    if (pm == PM_DYNAMIC) os << "const double* __restrict__ pp, const double* __restrict__ pq, cplx pt, ";
    os << "cdd* __restrict__ u) {\n";
    os << "  cdd R[" << std::max(nreg, 1) << "];\n";
    os << "  DDPar P; P.p = " << (pm == PM_DYNAMIC ? "pp" : "nullptr") << "; P.q = " << (pm == PM_DYNAMIC ? "pq" : "nullptr")
       << "; P.t = " << (pm == PM_DYNAMIC ? "pt" : "mkc(0.0, 0.0)") << ";\n";
    for (auto& ld : loads) os << "  R[" << ld.first << "] = x[" << (ld.second - tp->nparams) << "];\n";
    os << "  dd_interpret(" << fname << "_code, " << ninstr << ", " << fname << "_const, R, P);\n";
    for (int r = 0; r < n; r++) {
        if (out_operand[r] == INT32_MIN) os << "  u[" << r << "] = cdd_from(mkc(0.0, 0.0));\n";
        else os << "  u[" << r << "] = dd_fetch(" << fname << "_const, R, P, " << out_operand[r] << ");\n";
    }
    os << "}\n";
    long f = pm == PM_DYNAMIC ? 6L * tp->nparams : 0;  // p(t) of every parameter
    const std::vector<char> dup = recomputed(tp);  // code[i] is tape instruction i
    for (int i = 0; i < ninstr; i++) {
        if (dup[i]) continue;
        int op = code[6 * i];
        switch (op) {
        case OP_MUL: f += 58; break;
        case OP_MULADD: case OP_MULSUB: case OP_SUBMUL: f += 80; break;
        case OP_ADD: case OP_SUB: f += 22; break;
        case OP_SQR: f += 51; break;
        default: f += 120; break;
        }
    }
    *flops = f;
}

// ---------------------------- Taylor series ----------------------------
struct SVal {
    int deg;                    // highest possibly non-zero order, -1 = identically zero
    std::vector<std::string> c; // coefficient expressions (valid for 0..deg)
};

// How a row's Taylor coefficients of orders K-1 and K are written: (row, vKm1, vK) -> statements.
typedef std::function<std::string(int, const std::string&, const std::string&)> TaylorStoreFn;

static long gen_taylor_body(std::ostringstream& os, const hcg_tape* tp, int K, int n, ParamMode pm, const double* fixed,
                            const std::string& pfx, const TaylorStoreFn& store);

static void gen_taylor(std::ostringstream& os, const hcg_tape* tp, const std::string& fname, int K, int n, ParamMode pm,
                       const double* fixed, long* flops) {
    os << "__device__ __noinline__ void " << fname << "(";
    for (int j = 0; j < K; j++) os << "const cplx* __restrict__ t" << j << ", ";
    if (pm == PM_DYNAMIC) os << "const double* __restrict__ pp, const double* __restrict__ pq, cplx pt, ";
    // the parameter homotopy needs order K only (and its tracker tail-calls this function)
    const bool km1 = pm != PM_DYNAMIC;
    if (km1) os << "cplx* __restrict__ oKm1, ";
    os << "cplx* __restrict__ oK) {\n";
    *flops = gen_taylor_body(os, tp, K, n, pm, fixed, "", [km1](int row, const std::string& a, const std::string& b) {
        std::string r = km1 ? "oKm1[" + std::to_string(row) + "] = " + a + "; " : std::string();
        return r + "oK[" + std::to_string(row) + "] = " + b + ";";
    });
    os << "}\n";
}

static long gen_taylor_body(std::ostringstream& os, const hcg_tape* tp, int K, int n, ParamMode pm, const double* fixed,
                            const std::string& pfx, const TaylorStoreFn& store) {
    std::vector<SVal> v(tp->tape_space + 1);
    auto zero = [&]() { SVal s; s.deg = -1; return s; };
    for (auto& s : v) s = zero();
    for (int k = 0; k < tp->nconst; k++) {
        SVal s; s.deg = 0; s.c.push_back(clit(tp->constants[2 * k], tp->constants[2 * k + 1]));
        v[k + 1] = s;
    }
    // parameters and variables are defined right before their first use (short live ranges)
    std::vector<std::string> lazy_def(tp->tape_space + 1);
    for (int k = 0; k < tp->nparams; k++) {
        SVal s;
        std::string nm0 = pfx + "P" + std::to_string(k);
        std::string& d = lazy_def[tp->params_first + k];
        if (pm == PM_FIXED) {
            d = "  const cplx " + nm0 + " = " + clit(fixed[2 * k], fixed[2 * k + 1]) + ";\n";
            s.deg = 0; s.c.push_back(nm0);
        } else {
            d = "  const cplx " + nm0 + " = par_lerp(pt, pp, pq, " + std::to_string(k) + ");\n";
            std::string nm1 = pfx + "Q" + std::to_string(k);
            if (K >= 1) d += "  const cplx " + nm1 + " = par_diff(pp, pq, " + std::to_string(k) + ");\n";
            s.deg = K >= 1 ? 1 : 0; s.c.push_back(nm0);
            if (K >= 1) s.c.push_back(nm1);
        }
        v[tp->params_first + k] = s;
    }
    for (int k = 0; k < tp->nvars; k++) {
        SVal s;
        s.deg = K - 1;
        std::string& d = lazy_def[tp->vars_first + k];
        for (int j = 0; j < K; j++) {
            std::string nm = pfx + "X" + std::to_string(k) + "_" + std::to_string(j);
            d += "  const cplx " + nm + " = t" + std::to_string(j) + "[" + std::to_string(k) + "];\n";
            s.c.push_back(nm);
        }
        v[tp->vars_first + k] = s;
    }
    auto need = [&](int slot) {
        if (slot >= 1 && slot <= tp->tape_space && !lazy_def[slot].empty()) { os << lazy_def[slot]; lazy_def[slot].clear(); }
    };
    long f = pm == PM_DYNAMIC ? 8L * tp->nparams : 0;  // p(t) and p - q of every parameter
    auto has = [&](const SVal& s, int j) { return j >= 0 && j <= s.deg; };
    auto pack = [&](const SVal& s) {  // full ser<K> expression
        std::string e = "ser<" + std::to_string(K) + ">{{";
        for (int j = 0; j <= K; j++) {
            if (j) e += ", ";
            e += has(s, j) ? s.c[j] : "mkc(0.0, 0.0)";
        }
        return e + "}}";
    };
    // outputs are stored right after the last instruction writing their slot
    std::vector<int> last_write(tp->tape_space + 1, -1);
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        if (I[4] == OP_STOP) break;
        last_write[I[5]] = i;
    }
    std::vector<std::vector<std::pair<int, int>>> stores_after(tp->ninstr + 1);  // (row, slot)
    std::vector<int> uset(n, 0);
    for (int i = 0; i < tp->nu; i++) {
        int row = tp->u_assign[2 * i] - 1, slot = tp->u_assign[2 * i + 1];
        stores_after[last_write[slot] >= 0 ? last_write[slot] : tp->ninstr].push_back({row, slot});
        uset[row] = 1;
    }
    auto flush_stores = [&](int at) {
        for (auto& rs : stores_after[at]) {
            need(rs.second);
            const SVal& sv = v[rs.second];
            os << "  " << store(rs.first, has(sv, K - 1) ? sv.c[K - 1] : "mkc(0.0, 0.0)", has(sv, K) ? sv.c[K] : "mkc(0.0, 0.0)")
               << "\n";
        }
    };
    const std::vector<char> tdup = recomputed(tp);
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        if (op == OP_STOP) break;
        for (int q = 0; q < arity(op); q++)
            if (!(op == OP_POW_INT && q == 1)) need(I[q]);
        std::string base = pfx + "S" + std::to_string(i);
        // a recomputation (see recomputed()): its operands' coefficients pass through
        // hcg_opaque; the original series are put back before the result is recorded
        std::vector<std::pair<int, SVal>> saved;
        if (tdup[i]) {
            int k = 0;
            for (int q = 0; q < arity(op); q++) {
                if (op == OP_POW_INT && q == 1) continue;
                const int sl = I[q];
                bool done = false;
                for (auto& sv : saved) done |= sv.first == sl;
                if (done) continue;
                saved.push_back({sl, v[sl]});
                SVal o = v[sl];
                for (auto& cn : o.c) {
                    if (cn.compare(0, 4, "mkc(") == 0) continue;
                    std::string nmo = base + "o" + std::to_string(k++);
                    os << "  const cplx " << nmo << " = hcg_opaque(" << cn << ");\n";
                    cn = nmo;
                }
                v[sl] = o;
            }
        }
        const long f_before = f;
        const SVal& a = v[I[0]];
        SVal r;
        auto emit = [&](int k, const std::string& e) {
            std::string nm = base + "_" + std::to_string(k);
            os << "  const cplx " << nm << " = " << e << ";\n";
            if ((int)r.c.size() <= k) r.c.resize(k + 1);
            r.c[k] = nm;
        };
        auto prod_terms = [&](const SVal& x, const SVal& y, int k, std::string acc, bool have_acc) {
            // Σ_{j} x_j y_{k-j}, j ascending, skipping structural zeros; muladd chain on acc
            for (int j = 0; j <= k; j++) {
                if (!has(x, j) || !has(y, k - j)) continue;
                if (!have_acc) { acc = "c_mul(" + x.c[j] + ", " + y.c[k - j] + ")"; have_acc = true; f += 6; }
                else { acc = "c_muladd(" + x.c[j] + ", " + y.c[k - j] + ", " + acc + ")"; f += 8; }
            }
            return std::make_pair(acc, have_acc);
        };
        switch (op) {
        case OP_IDENTITY: r = a; break;
        case OP_NEG:
            r.deg = a.deg;
            for (int k = 0; k <= a.deg; k++) emit(k, "c_neg(" + a.c[k] + ")");
            break;
        case OP_ADD:
        case OP_SUB: {
            const SVal& b = v[I[1]];
            r.deg = std::max(a.deg, b.deg);
            for (int k = 0; k <= r.deg; k++) {
                bool ha = has(a, k), hb = has(b, k);
                if (ha && hb) { emit(k, std::string(op == OP_ADD ? "c_add(" : "c_sub(") + a.c[k] + ", " + b.c[k] + ")"); f += 2; }
                else if (ha) emit(k, a.c[k]);
                else if (hb) emit(k, op == OP_ADD ? b.c[k] : "c_sub(mkc(0.0, 0.0), " + b.c[k] + ")");
                else emit(k, "mkc(0.0, 0.0)");
            }
            break;
        }
        case OP_MUL: {
            const SVal& b = v[I[1]];
            if (a.deg < 0 || b.deg < 0) { r.deg = -1; break; }
            r.deg = std::min(K, a.deg + b.deg);
            for (int k = 0; k <= r.deg; k++) {
                auto pr = prod_terms(a, b, k, "", false);
                emit(k, pr.second ? pr.first : "mkc(0.0, 0.0)");
            }
            break;
        }
        case OP_MULADD: {
            const SVal& b = v[I[1]];
            const SVal& c = v[I[2]];
            int dp = (a.deg < 0 || b.deg < 0) ? -1 : std::min(K, a.deg + b.deg);
            r.deg = std::max(dp, c.deg);
            for (int k = 0; k <= r.deg; k++) {
                bool hc = has(c, k);
                auto pr = prod_terms(a, b, k, hc ? c.c[k] : "", hc);
                emit(k, pr.second ? pr.first : "mkc(0.0, 0.0)");
            }
            break;
        }
        case OP_SQR: {
            if (a.deg < 0) { r.deg = -1; break; }
            r.deg = std::min(K, 2 * a.deg);
            for (int k = 0; k <= r.deg; k++) {
                if (k == 0) { emit(0, "c_sqr(" + a.c[0] + ")"); f += 5; continue; }
                std::string w; bool hw = false;
                for (int j = 0; j <= (k - 1) / 2; j++) {
                    if (!has(a, j) || !has(a, k - j)) continue;
                    if (!hw) { w = "c_mul(" + a.c[j] + ", " + a.c[k - j] + ")"; hw = true; f += 6; }
                    else { w = "c_muladd(" + a.c[j] + ", " + a.c[k - j] + ", " + w + ")"; f += 8; }
                }
                if (k % 2 == 0) {
                    bool hs = has(a, k / 2);
                    std::string sq = hs ? "c_sqr(" + a.c[k / 2] + ")" : "mkc(0.0, 0.0)";
                    if (hw) { emit(k, "c_add(c_scale(2.0, " + w + "), " + sq + ")"); f += 9; }
                    else emit(k, sq);
                } else {
                    if (hw) { emit(k, "c_add(" + w + ", " + w + ")"); f += 2; }
                    else emit(k, "mkc(0.0, 0.0)");
                }
            }
            break;
        }
        default: {
            // remaining ops: full-series templates
            std::string args = pack(a);
            int ar = arity(op);
            std::string fn;
            switch (op) {
            case OP_CB: fn = "s_cb"; f += 12 * (K + 1) * (K + 2); break;
            case OP_INV: case OP_INV_NOT_ZERO: fn = "s_inv"; f += 8 * (K + 1) * (K + 2); break;
            case OP_INVSQR: fn = "s_invsqr"; f += 12 * (K + 1) * (K + 2); break;
            case OP_DIV: fn = "s_div"; f += 8 * (K + 1) * (K + 2); break;
            case OP_POW_INT: fn = "s_pow_int"; f += 16 * (K + 1) * (K + 2); break;
            case OP_ADD3: fn = "s_add3"; f += 4 * (K + 1); break;
            case OP_MUL3: fn = "s_mul3"; f += 8 * (K + 1) * (K + 2); break;
            case OP_MULSUB: fn = "s_mulsub"; f += 4 * (K + 1) * (K + 2); break;
            case OP_SUBMUL: fn = "s_submul"; f += 4 * (K + 1) * (K + 2); break;
            case OP_ADD4: fn = "s_add4"; f += 6 * (K + 1); break;
            case OP_MUL4: fn = "s_mul4"; f += 12 * (K + 1) * (K + 2); break;
            case OP_MULMULADD: fn = "s_mulmuladd"; f += 8 * (K + 1) * (K + 2); break;
            case OP_MULMULSUB: fn = "s_mulmulsub"; f += 8 * (K + 1) * (K + 2); break;
            }
            if (op == OP_POW_INT) args += ", " + std::to_string(I[1]);
            else for (int q = 1; q < ar; q++) args += ", " + pack(v[I[q]]);
            os << "  const ser<" << K << "> " << base << " = " << fn << "<" << K << ">(" << args << ");\n";
            r.deg = K;
            r.c.resize(K + 1);
            for (int k = 0; k <= K; k++) r.c[k] = base + ".c[" + std::to_string(k) + "]";
            break;
        }
        }
        for (auto& sv : saved) v[sv.first] = sv.second;
        if (tdup[i]) f = f_before;
        v[I[5]] = r;
        flush_stores(i);
    }
    flush_stores(tp->ninstr);
    for (int i = 0; i < n; i++)
        if (!uset[i]) os << "  " << store(i, "mkc(0.0, 0.0)", "mkc(0.0, 0.0)") << "\n";
    return f;
}

// Straight-line homotopy, Taylor coefficient of order K of H = γ t G + (1 - t) F along the
// path (x(t) = Σ t_j τ^j):  order 1 (values):  u = γ G(x) - F(x);  order K >= 2:
// u = γ (G_{K-1} + t G_K) + ((1 - t) F_K - F_{K-1}).  Both bodies in one function: F's part
// is stored first and G's part added where G's body stores its row (same operations and
// order as combining two separately stored results).
static long gen_straight_taylor(std::ostringstream& os, const hcg_system& G, const hcg_system& F, int K, int n,
                                ParamMode pmG, ParamMode pmF) {
    long f = 0;
    os << "__device__ __noinline__ void sysH_taylor" << K << "(";
    if (K == 1) os << "const cplx* __restrict__ x, ";
    else
        for (int j = 0; j < K; j++) os << "const cplx* __restrict__ t" << j << ", ";
    os << "cplx g, cplx t, cplx omt, cplx* __restrict__ u) {\n";
    if (K == 1) {
        f += gen_eval_body(os, F.eval, false, false, n, pmF, F.fixed_params, "f",
                           [](char, int i, const std::string& v) { return plain_store('u', i, v); });
        f += gen_eval_body(os, G.eval, false, false, n, pmG, G.fixed_params, "g", [](char, int i, const std::string& v) {
            std::string d = "u[" + std::to_string(i) + "]";
            return d + " = c_sub(c_mul(g, " + v + "), " + d + ");";
        });
    } else {
        f += gen_taylor_body(os, F.eval, K, n, pmF, F.fixed_params, "f", [](int i, const std::string& a, const std::string& b) {
            return "u[" + std::to_string(i) + "] = c_sub(c_mul(omt, " + b + "), " + a + ");";
        });
        f += gen_taylor_body(os, G.eval, K, n, pmG, G.fixed_params, "g", [](int i, const std::string& a, const std::string& b) {
            std::string d = "u[" + std::to_string(i) + "]";
            return d + " = c_add(c_mul(g, c_add(" + a + ", c_mul(t, " + b + "))), " + d + ");";
        });
    }
    os << "}\n";
    return f;
}

// Functions of one system.  A straight-line homotopy's Jacobian and Taylor coefficients come
// from the fused sysH_* functions instead (gen_straight_*), so only its values are generated
// per system.
static void gen_system(std::ostringstream& os, const hcg_system& s, const std::string& prefix, int n, ParamMode pm,
                       bool straight, Flops* fl) {
    const double* fixed = s.fixed_params;
    long f;
    gen_eval(os, s.eval, prefix + "_eval", false, false, n, pm, fixed, &f); fl->eval += f;
    gen_dd_interp(os, s.eval, prefix + "_eval_dd", n, pm, fixed, &f); fl->eval_dd += f;
    if (straight) return;
    gen_eval(os, s.eval_jac, prefix + "_eval_jac", true, false, n, pm, fixed, &f); fl->eval_jac += f;
    for (int K = 1; K <= 3; K++) {
        gen_taylor(os, s.eval, prefix + "_taylor" + std::to_string(K), K, n, pm, fixed, &f);
        fl->taylor[K] += f;
    }
}

// Validates one homotopy and writes its HCG_* macros, system functions and static flop
// counts (everything a tracker body needs) to `os`.
static bool emit_system(std::ostringstream& os, const hcg_homotopy_desc* h, Flops* fl, std::string* err,
                        long* fl6 = nullptr) {
    int n = h->n;
    if (n < 1 || n > HCG_MAX_N) { *err = "system size out of range (1.." + std::to_string(HCG_MAX_N) + ")"; return false; }
    if (h->kind != HCG_STRAIGHT_LINE && h->kind != HCG_PARAMETER) { *err = "unknown homotopy kind"; return false; }
    if (!validate_tape(h->target.eval_jac, n, err) || !validate_tape(h->target.eval, n, err)) return false;
    if (h->target.eval_jac->nU == 0 && n > 0) { *err = "eval_jac tape has no Jacobian assignments"; return false; }
    if (h->target.eval_jac->nparams != h->target.eval->nparams) { *err = "eval/eval_jac parameter counts differ"; return false; }
    if (h->kind == HCG_STRAIGHT_LINE) {
        if (!validate_tape(h->start.eval_jac, n, err) || !validate_tape(h->start.eval, n, err)) return false;
        if (h->start.eval->nparams && !h->start.fixed_params) { *err = "start system parameters must be fixed"; return false; }
        if (h->target.eval->nparams && !h->target.fixed_params) { *err = "target system parameters must be fixed"; return false; }
    } else if (h->target.eval->nparams < 1) {
        *err = "parameter homotopy needs a system with parameters";
        return false;
    }
    *fl = Flops();
    os << "#define HCG_N " << n << "\n#define HCG_KIND " << h->kind << "\n";
    os << "#define HCG_NPAR " << (h->kind == HCG_PARAMETER ? h->target.eval->nparams : 0) << "\n";
    os << "#define HCG_GAMMA_RE " << dlit(h->gamma_re) << "\n#define HCG_GAMMA_IM " << dlit(h->gamma_im) << "\n";
    if (h->kind == HCG_STRAIGHT_LINE) {
        const ParamMode pmG = h->start.eval->nparams ? PM_FIXED : PM_NONE, pmF = h->target.eval->nparams ? PM_FIXED : PM_NONE;
        gen_system(os, h->start, "sysG", n, pmG, true, fl);
        gen_system(os, h->target, "sysF", n, pmF, true, fl);
        fl->eval_jac += gen_straight_eval_jac(os, h->start, h->target, n, pmG, pmF);
        for (int K = 2; K <= 3; K++) fl->taylor[K] += gen_straight_taylor(os, h->start, h->target, K, n, pmG, pmF);
        gen_straight_taylor(os, h->start, h->target, 1, n, pmG, pmF);  // values: counted as fl->eval
    } else {
        gen_system(os, h->target, "sysF", n, PM_DYNAMIC, false, fl);
    }
    // static FP64 flop counts of one homotopy-level call (system functions + combination)
    long comb_eval = 0, comb_ej = 0, comb_dd = 0, comb_t1 = 0, comb_t23 = 0;
    if (h->kind == HCG_STRAIGHT_LINE) {
        comb_eval = 6 + 14L * n;
        comb_ej = 6 + 14L * (n + (long)n * n);
        comb_dd = 6 + 94L * n;
        comb_t1 = 8L * n;
        comb_t23 = 24L * n;
    }
    const long f6[6] = {fl->eval + comb_eval, fl->eval_jac + comb_ej, fl->eval_dd + comb_dd,
                        h->kind == HCG_STRAIGHT_LINE ? fl->eval + comb_t1 : fl->taylor[1], fl->taylor[2] + comb_t23,
                        fl->taylor[3] + comb_t23};
    if (fl6)
        for (int i = 0; i < 6; i++) fl6[i] = f6[i];
    os << "#define HCG_FL_EVAL " << f6[0] << "\n";
    os << "#define HCG_FL_EJ " << f6[1] << "\n";
    os << "#define HCG_FL_DD " << f6[2] << "\n";
    os << "#define HCG_FL_T1 " << f6[3] << "\n";
    os << "#define HCG_FL_T2 " << f6[4] << "\n";
    os << "#define HCG_FL_T3 " << f6[5] << "\n";
    return true;
}

// The counted FP64 flops of one homotopy-level evaluation (eval, eval_jac, eval_dd, taylor 1..3),
// as the generated kernels count them; also validates the homotopy.
bool homotopy_flops(const hcg_homotopy_desc* h, long* fl6, std::string* err) {
    std::ostringstream os;
    Flops fl;
    return emit_system(os, h, &fl, err, fl6);
}

bool generate_source(const hcg_homotopy_desc* h, std::string* src, Flops* fl, std::string* err) {
    std::ostringstream os;
    os << "// generated by hcgpu codegen\n";
    os << "#include \"hcg_arith.h\"\n";
    if (!emit_system(os, h, fl, err)) return false;
    os << "#include \"hcg_tracker.h\"\n";
    *src = os.str();
    return true;
}

// Several homotopies in one program: system s lives in namespace hcs<s> (its functions
// plus its own instance of the tracker body); one persistent kernel takes paths from a
// single work queue over all systems and dispatches each to its system's tracker.
bool generate_multi_source(const hcg_homotopy_desc* const* hs, int nsys, std::string* src, Flops* fls,
                           std::string* err) {
    if (nsys < 1 || nsys > 8) { *err = "between 1 and 8 systems per program"; return false; }
    std::ostringstream os;
    os << "// generated by hcgpu codegen (multi-system program)\n";
    os << "#include \"hcg_arith.h\"\n#include \"hcgpu_types.h\"\n#include \"hcg_kargs.h\"\n#include \"hcg_sched.h\"\n#include \"hcg_mlp_infer.h\"\n";
    for (int s = 0; s < nsys; s++) {
        os << "namespace hcs" << s << " {\n";
        std::string e;
        if (!emit_system(os, hs[s], &fls[s], &e)) { *err = "system " + std::to_string(s) + ": " + e; return false; }
        os << "#include \"hcg_tracker_body.h\"\n";
        for (const char* m : {"HCG_N", "HCG_KIND", "HCG_NPAR", "HCG_GAMMA_RE", "HCG_GAMMA_IM", "HCG_FL_EVAL", "HCG_FL_EJ",
                              "HCG_FL_DD", "HCG_FL_T1", "HCG_FL_T2", "HCG_FL_T3"})
            os << "#undef " << m << "\n";
        os << "}  // namespace hcs" << s << "\n";
    }
    os << "union hcg_lane_union {\n";
    for (int s = 0; s < nsys; s++) os << "    hcs" << s << "::Lane l" << s << ";\n";
    os << "};\n";
    auto each = [&](const std::string& body) {
        std::string out = "            switch (s) {\n";
        for (int s = 0; s < nsys; s++) {
            std::string b = body;
            std::string ns = "hcs" + std::to_string(s), l = "U.l" + std::to_string(s), a = "m.sys[" + std::to_string(s) + "]";
            for (size_t p; (p = b.find("@NS")) != std::string::npos;) b.replace(p, 3, ns);
            for (size_t p; (p = b.find("@L")) != std::string::npos;) b.replace(p, 2, l);
            for (size_t p; (p = b.find("@A")) != std::string::npos;) b.replace(p, 2, a);
            out += "            case " + std::to_string(s) + ": " + b + " break;\n";
        }
        return out + "            }\n";
    };
    os << R"(extern "C" __global__ void __launch_bounds__(64, HCG_WAVES_PER_EU) hcg_track_multi_kernel(hcg_multi_kargs m) {
    __shared__ unsigned short mlp_lds[HCG_MLP_NPARAMS];
    const unsigned short* mw = stage_mlp(m.sys[0].mlp, m.sys[0].opts, mlp_lds);
    hcg_lane_union U;
    LaneTotals T = {0, 0, 0, 0.0};
    const int refill = refill_threshold(m.sys[0].opts);
    int phase = 0, k = -1, s = 0;
    while (true) {
        int ended = 0;
        const unsigned long long fetch = wave_refill_mask(phase == 0, refill);
        if (phase == 0) {
            if (!fetch) continue;
            int kg = wave_fetch(fetch, m.next_path);
            if (kg >= m.total) break;
            s = 0;
            while (s + 1 < m.nsys && kg >= m.offset[s + 1]) s++;
            k = kg - m.offset[s];
            int live = 0;
)" << each("{ @NS::lane_setup(@L, @A, mw); live = @NS::lane_begin(@L, @A, k); if (live) ended = @NS::lane_ended(@L); }")
       << R"(            if (!live) continue;
            phase = 1;
        } else {
)" << each("{ @NS::lane_iter(@L); ended = @NS::lane_ended(@L); }")
       << R"(        }
        if (ended) {
)" << each("@NS::lane_finish(@L, @A, k, T);")
       << R"(            phase = 0;
        }
    }
    atomicAdd(&m.sys[0].counters[0], (unsigned long long)T.steps);
    atomicAdd(&m.sys[0].counters[1], (unsigned long long)T.newton);
    atomicAdd(&m.sys[0].counters[2], (unsigned long long)T.flops);
    atomicAdd(&m.sys[0].counters[3], (unsigned long long)T.lacc);
}
)";
    *src = os.str();
    return true;
}


// ---------------------------- certification ----------------------------
// Interval (ball = false) or double-double ball (ball = true) evaluation of a System tape:
// the u outputs (jac = false) or the U outputs (jac = true), at a box X (IComplex) or at a
// double-double point X, with parameter boxes / balls PB.  An interpreter program (the same
// register planning as the double-double residual) rather than straight-line code: interval
// and ball operations are long, and straight-line bodies of large tapes took tens of minutes
// to compile.  ci_interpret / cb_interpret (hcg_interval.h) run the generic IComplex
// operations; quotients have no ball form (their results are unbounded balls).
static void gen_cert_interp(std::ostringstream& os, const hcg_tape* tp, const std::string& fname, bool jac, int n,
                            ParamMode pm, const double* fixed, bool ball) {
    const int nout = jac ? n * n : n, na = jac ? tp->nU : tp->nu;
    const int32_t* as = jac ? tp->U_assign : tp->u_assign;
    std::vector<std::pair<int, int>> outs;
    for (int i = 0; i < na; i++) outs.push_back({as[2 * i] - 1, as[2 * i + 1]});
    const InterpPlan P = plan_interp(tp, outs, nout, pm, fixed, 1 << 24);
    emit_tables(os, fname, P);
    const std::string T = ball ? "cball" : "icplx", pre = ball ? "cb" : "ci";
    os << "__device__ __noinline__ void " << fname << "(const " << (ball ? "cdd" : "icplx")
       << "* __restrict__ X, const double* __restrict__ PB, " << T << "* __restrict__ out) {\n";
    os << "  " << T << " R[" << std::max(P.nreg, 1) << "];\n";
    for (auto& ld : P.loads) {
        const std::string x = "X[" + std::to_string(ld.second - tp->nparams) + "]";
        os << "  R[" << ld.first << "] = " << (ball ? "cb_ddpt(" + x + ")" : x) << ";\n";
    }
    os << "  " << pre << "_interpret(" << fname << "_code, " << P.ninstr << ", " << fname << "_const, R, PB);\n";
    for (int r = 0; r < nout; r++) {
        if (P.out_operand[r] == INT32_MIN) os << "  out[" << r << "] = " << (ball ? "cb_pt(mkc(0.0, 0.0))" : "ic_zero()") << ";\n";
        else os << "  out[" << r << "] = " << pre << "_fetch(" << fname << "_const, R, PB, " << P.out_operand[r] << ");\n";
    }
    os << "}\n";
}

bool generate_certify_source(const hcg_system* F, int n, std::string* src, std::string* err) {
    if (!F || !F->eval || !F->eval_jac) { *err = "NULL system"; return false; }
    if (n < 1 || n > HCG_MAX_N) { *err = "system size out of range (1.." + std::to_string(HCG_MAX_N) + ")"; return false; }
    if (!validate_tape(F->eval_jac, n, err) || !validate_tape(F->eval, n, err)) return false;
    if (F->eval_jac->nU == 0) { *err = "eval_jac tape has no Jacobian assignments"; return false; }
    if (F->eval_jac->nparams != F->eval->nparams) { *err = "eval/eval_jac parameter counts differ"; return false; }
    const int np = F->eval->nparams;
    const ParamMode pm = np == 0 ? PM_NONE : (F->fixed_params ? PM_FIXED : PM_DYNAMIC);
    const double* fixed = F->fixed_params;
    std::ostringstream os;
    os << "// generated by hcgpu codegen (certifier)\n#include \"hcg_arith.h\"\n#include \"hcg_interval.h\"\n";
    os << "#define HCG_N " << n << "\n#define HCG_NPAR " << (pm == PM_DYNAMIC ? np : 0) << "\n";
    long f;
    gen_eval(os, F->eval_jac, "certF_eval_jac", true, false, n, pm, fixed, &f);
    gen_dd_interp(os, F->eval, "certF_eval_dd", n, pm, fixed, &f);
    gen_cert_interp(os, F->eval, "cert_ieval", false, n, pm, fixed, false);
    gen_cert_interp(os, F->eval_jac, "cert_ijac", true, n, pm, fixed, false);
    gen_cert_interp(os, F->eval, "cert_beval", false, n, pm, fixed, true);
    gen_cert_interp(os, F->eval_jac, "cert_bjac", true, n, pm, fixed, true);
    const std::string pa = pm == PM_DYNAMIC ? "P, P, mkc(0.0, 0.0), " : "";
    os << "HD void cert_eval_jac(const cplx* x, const double* P, cplx* u, cplx* U) { certF_eval_jac(x, " << pa << "u, U); }\n";
    os << "HD void cert_eval_dd(const cdd* x, const double* P, cdd* u) { certF_eval_dd(x, " << pa << "u); }\n";
    os << "#include \"hcg_certify.h\"\n";
    *src = os.str();
    return true;
}

}  // namespace hcg
