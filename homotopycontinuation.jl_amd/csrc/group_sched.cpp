// Slot scheduling and LDS register allocation of tapes for the lane-group tracker
// (see group_sched.h).  Tape semantics follow ModelKit's InstructionSequence
// (src/model_kit/instruction_sequence.jl:1-27): slots 1..nconst hold the constants, then the
// parameters, then the variables; every instruction writes its output slot.
#include "group_sched.h"

#include <algorithm>
#include <cstdio>
#include <map>
#include <queue>
#include <set>

namespace hcg {
namespace {

// A value may be read kMinDistance slots after the slot that writes it.  The interpreter
// finishes a slot (reads, then writes) before it starts the next, so 1 suffices.
constexpr int kMinDistance = 1;

enum { OPS_MAX = 33 };
const int kArity[OPS_MAX] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                             2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4};
constexpr int OP_POW_INT_ = 22;
// OpTypes the device interpreter implements (src/model_kit/operations.jl:183-248)
bool op_supported(int op) {
    switch (op) {
    case 1: case 7: case 8: case 9: case 10: case 13: case 17: case 18: case 19: case 20: case 21: case 22:
    case 24: case 25: case 26: case 27: case 28: case 29: case 30: case 31: case 32:
        return true;
    default:
        return false;
    }
}

// A value of the merged program: an instruction result (>= 0) or a variable (-(v+1) over
// all tapes' variables), or an operand reference that needs no register.
struct Operand {
    int kind;   // 0: instruction value, 1: variable value, 2: plain reference (const/param/zero/int)
    int id;     // instruction index or variable value index
    int32_t ref;
};

struct Instr {
    int op;
    int nargs;
    Operand a[4];
    int tape;
};

struct Lowered {
    std::vector<Instr> ins;
    int nvarvals = 0;
    std::vector<int> var_tape, var_index;  // per variable value: its tape and variable number
    std::vector<std::vector<Operand>> u;    // per tape: per output row
    std::vector<std::vector<Operand>> U;    // per tape: n*n row-major
};

bool lower(const std::vector<const hcg_tape*>& tapes, const std::vector<const double*>& fixed, int n, Lowered* L,
           std::vector<double>* consts, std::string* err) {
    char buf[256];
    for (size_t t = 0; t < tapes.size(); t++) {
        const hcg_tape* tp = tapes[t];
        if (!tp || !tp->instr || tp->ninstr < 0 || tp->tape_space < 1) { *err = "malformed tape"; return false; }
        std::vector<Operand> def(tp->tape_space + 1);
        for (auto& d : def) d = {2, 0, GREF_ZERO};
        const int cbase = (int)(consts->size() / 2);
        for (int k = 0; k < tp->nconst; k++) {
            consts->push_back(tp->constants[2 * k]);
            consts->push_back(tp->constants[2 * k + 1]);
            def[k + 1] = {2, 0, GREF_CONST | (cbase + k)};
        }
        for (int k = 0; k < tp->nparams; k++) {
            int32_t ref;
            if (fixed[t]) {
                ref = GREF_CONST | (int32_t)(consts->size() / 2);
                consts->push_back(fixed[t][2 * k]);
                consts->push_back(fixed[t][2 * k + 1]);
            } else {
                ref = GREF_PARAM | k;
            }
            def[tp->params_first + k] = {2, 0, ref};
        }
        for (int k = 0; k < tp->nvars; k++) {
            def[tp->vars_first + k] = {1, L->nvarvals, 0};
            L->var_tape.push_back((int)t);
            L->var_index.push_back(k);
            L->nvarvals++;
        }
        for (int i = 0; i < tp->ninstr; i++) {
            const int32_t* I = tp->instr + 6 * i;
            const int op = I[4];
            if (op == 0) break;
            if (op < 0 || op >= OPS_MAX || !op_supported(op)) {
                std::snprintf(buf, sizeof buf, "tape %zu instruction %d: OpType %d has no group form", t, i, op);
                *err = buf;
                return false;
            }
            Instr in;
            in.op = op;
            in.nargs = kArity[op];
            in.tape = (int)t;
            for (int k = 0; k < in.nargs; k++) {
                if (op == OP_POW_INT_ && k == 1) { in.a[k] = {2, 0, GREF_INT | (I[1] & 0x0FFFFFFF)}; continue; }
                if (I[k] < 1 || I[k] > tp->tape_space) { *err = "tape operand out of range"; return false; }
                in.a[k] = def[I[k]];
            }
            if (I[5] < 1 || I[5] > tp->tape_space) { *err = "tape output out of range"; return false; }
            def[I[5]] = {0, (int)L->ins.size(), 0};
            L->ins.push_back(in);
        }
        std::vector<Operand> u(n, Operand{2, 0, GREF_ZERO}), U((size_t)n * n, Operand{2, 0, GREF_ZERO});
        for (int k = 0; k < tp->nu; k++) {
            int row = tp->u_assign[2 * k] - 1, slot = tp->u_assign[2 * k + 1];
            if (row < 0 || row >= n || slot < 1 || slot > tp->tape_space) { *err = "bad u assignment"; return false; }
            u[row] = def[slot];
        }
        for (int k = 0; k < tp->nU; k++) {
            int l = tp->U_assign[2 * k] - 1, slot = tp->U_assign[2 * k + 1];
            if (l < 0 || l >= n * n || slot < 1 || slot > tp->tape_space) { *err = "bad U assignment"; return false; }
            const int i = l % n, j = l / n;  // column-major linear index
            U[(size_t)i * n + j] = def[slot];
        }
        L->u.push_back(u);
        L->U.push_back(U);
    }
    return true;
}

}  // namespace

bool schedule_tapes(const std::vector<const hcg_tape*>& tapes, const std::vector<const double*>& fixed, int n,
                    GProgram* P, std::string* err) {
    const int G = kGroupLanes;
    Lowered L;
    *P = GProgram();
    if (!lower(tapes, fixed, n, &L, &P->consts, err)) return false;
    const int N = (int)L.ins.size();
    P->ninstr = N;
    // consumers and longest path to a sink
    std::vector<std::vector<int>> users(N);
    std::vector<int> npred(N, 0);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < L.ins[i].nargs; k++)
            if (L.ins[i].a[k].kind == 0) { users[L.ins[i].a[k].id].push_back(i); npred[i]++; }
    std::vector<int> lp(N, 1);
    for (int i = N - 1; i >= 0; i--)
        for (int u : users[i]) lp[i] = std::max(lp[i], lp[u] + 1);
    // list scheduling: per-op priority queues of eligible instructions
    typedef std::pair<int, int> Key;  // (priority, -index)
    std::vector<std::priority_queue<Key>> ready(OPS_MAX);
    std::map<int, std::vector<int>> waiting;  // earliest slot -> instructions
    std::vector<int> earliest(N, 0), slot_of(N, -1), left = npred;
    for (int i = 0; i < N; i++)
        if (left[i] == 0) ready[L.ins[i].op].push(Key(lp[i], -i));
    int done = 0, s = 0;
    std::vector<std::vector<int>> slots;
    while (done < N) {
        auto w = waiting.find(s);
        if (w != waiting.end()) {
            for (int i : w->second) ready[L.ins[i].op].push(Key(lp[i], -i));
            waiting.erase(w);
        }
        int best = -1;
        for (int op = 0; op < OPS_MAX; op++)
            if (!ready[op].empty() && (best < 0 || ready[op].top() > ready[best].top())) best = op;
        std::vector<int> pick;
        if (best >= 0)
            while (!ready[best].empty() && (int)pick.size() < G) { pick.push_back(-ready[best].top().second); ready[best].pop(); }
        for (int i : pick) {
            slot_of[i] = s;
            for (int u : users[i]) {
                earliest[u] = std::max(earliest[u], s + kMinDistance);
                if (--left[u] == 0) waiting[earliest[u]].push_back(u);
            }
        }
        done += (int)pick.size();
        slots.push_back(pick);
        s++;
    }
    const int S = (int)slots.size();
    P->nslots = S;
    // lifetimes: instruction values [slot, last use]; variables [-1, last use]; outputs to the end
    const int END = S + 1;
    std::vector<int> last_i(N, -1), last_v(L.nvarvals, -2);
    auto use = [&](const Operand& o, int at) {
        if (o.kind == 0) last_i[o.id] = std::max(last_i[o.id], at);
        else if (o.kind == 1) last_v[o.id] = std::max(last_v[o.id], at);
    };
    for (int i = 0; i < N; i++)
        for (int k = 0; k < L.ins[i].nargs; k++) use(L.ins[i].a[k], slot_of[i]);
    for (size_t t = 0; t < L.u.size(); t++) {
        for (auto& o : L.u[t]) use(o, END);
        for (auto& o : L.U[t]) use(o, END);
    }
    // linear scan: a register read for the last time in a slot may take that slot's results,
    // because the interpreter reads every operand of a slot before it writes any result (the
    // 16 lanes of a group are one wave's instruction stream); check_program verifies it
    std::vector<int> reg_i(N, -1), reg_v(L.nvarvals, -1);
    std::set<int> freeregs;
    int nregs = 0;
    std::multimap<int, int> release;  // last-use slot -> register
    auto alloc = [&]() {
        if (!freeregs.empty()) { int r = *freeregs.begin(); freeregs.erase(freeregs.begin()); return r; }
        return nregs++;
    };
    for (int v = 0; v < L.nvarvals; v++)
        if (last_v[v] >= 0) { reg_v[v] = alloc(); release.insert({last_v[v], reg_v[v]}); }
    for (int si = 0; si < S; si++) {
        // values whose last read is before this slot free their registers
        while (!release.empty() && release.begin()->first <= si) {
            freeregs.insert(release.begin()->second);
            release.erase(release.begin());
        }
        for (int i : slots[si]) {
            if (last_i[i] < 0) continue;  // dead value: computed, never stored
            reg_i[i] = alloc();
            release.insert({last_i[i], reg_i[i]});
        }
    }
    P->nregs = nregs;
    auto ref = [&](const Operand& o) -> int32_t {
        if (o.kind == 0) return reg_i[o.id] >= 0 ? (GREF_REG | reg_i[o.id]) : GREF_ZERO;
        if (o.kind == 1) return GREF_REG | reg_v[o.id];
        return o.ref;
    };
    P->op.assign(S, 0);
    P->arg.assign((size_t)S * G * 4, GREF_NONE);
    P->out.assign((size_t)S * G, -1);
    for (int si = 0; si < S; si++) {
        const auto& pick = slots[si];
        if (!pick.empty()) P->op[si] = (uint8_t)L.ins[pick[0]].op;
        for (size_t ln = 0; ln < pick.size(); ln++) {
            const Instr& in = L.ins[pick[ln]];
            for (int k = 0; k < in.nargs; k++) P->arg[((size_t)si * G + ln) * 4 + k] = ref(in.a[k]);
            P->out[(size_t)si * G + ln] = reg_i[pick[ln]];  // -1: result unused
        }
    }
    P->io.resize(tapes.size());
    for (size_t t = 0; t < tapes.size(); t++) {
        GTapeIO& io = P->io[t];
        io.var_reg.assign(tapes[t]->nvars, -1);
        for (int v = 0; v < L.nvarvals; v++)
            if (L.var_tape[v] == (int)t) io.var_reg[L.var_index[v]] = reg_v[v];
        for (auto& o : L.u[t]) io.u_ref.push_back(ref(o));
        for (auto& o : L.U[t]) io.U_ref.push_back(ref(o));
    }
    return true;
}

bool check_program(const std::vector<const hcg_tape*>& tapes, const std::vector<const double*>& fixed,
                   const GProgram& P, std::string* err) {
    const int G = kGroupLanes;
    std::vector<double> consts;
    Lowered L;
    const int n = P.io.empty() ? 0 : (int)P.io[0].u_ref.size();
    if (!lower(tapes, fixed, n, &L, &consts, err)) return false;
    if (consts != P.consts) { *err = "constant table differs"; return false; }
    const int N = (int)L.ins.size();
    if (N != P.ninstr) { *err = "instruction count differs"; return false; }
    // identify each slot entry with a tape instruction: the k-th placed instruction of the
    // program is matched by (op, operand value ids) against the lowered tape, in order of
    // first availability; value ids: instruction i -> i, variable v -> N + v
    std::vector<long> reg(P.nregs, -1);       // value id held by each register
    std::vector<int> reg_slot(P.nregs, -1000);  // slot that wrote it (variables: before slot 0)
    for (size_t t = 0; t < tapes.size(); t++)
        for (size_t k = 0; k < P.io[t].var_reg.size(); k++)
            if (P.io[t].var_reg[k] >= 0) {
                int v = -1;
                for (int q = 0; q < L.nvarvals; q++)
                    if (L.var_tape[q] == (int)t && L.var_index[q] == (int)k) v = q;
                if (v < 0 || P.io[t].var_reg[k] >= P.nregs) { *err = "bad variable register"; return false; }
                reg[P.io[t].var_reg[k]] = N + v;
            }
    // map instruction -> its position by matching in program order: instructions are unique
    // by (op, operands) up to duplicates, which are interchangeable
    // Instructions that repeat an earlier one (same op on the same values, e.g. a monomial
    // recomputed per row) compute the same value: each gets the id of its first occurrence.
    std::multimap<std::vector<long>, int> sig;
    std::vector<long> canon(N);
    auto opid = [&](const Operand& o) -> long {
        if (o.kind == 0) return canon[o.id];
        if (o.kind == 1) return N + o.id;
        return -1000000000L - (long)(uint32_t)o.ref;
    };
    {
        std::map<std::vector<long>, int> first;
        for (int i = 0; i < N; i++) {
            std::vector<long> key = {L.ins[i].op};
            for (int k = 0; k < L.ins[i].nargs; k++) key.push_back(opid(L.ins[i].a[k]));
            auto it = first.find(key);
            canon[i] = it == first.end() ? i : it->second;
            if (it == first.end()) first[key] = i;
            sig.insert({key, i});
        }
    }
    std::vector<int> placed(N, -1);
    int count = 0;
    for (int s = 0; s < P.nslots; s++) {
        std::vector<std::pair<int, long>> writes;
        for (int ln = 0; ln < G; ln++) {
            const int32_t* a = &P.arg[((size_t)s * G + ln) * 4];
            if (a[0] == GREF_NONE) continue;
            const int op = P.op[s];
            std::vector<long> key = {op};
            for (int k = 0; k < kArity[op]; k++) {
                const int32_t r = a[k];
                if (gref_kind(r) == GREF_REG) {
                    const int ri = gref_index(r);
                    if (ri >= P.nregs || reg[ri] < 0) { *err = "read of an unassigned register"; return false; }
                    key.push_back(reg[ri]);
                } else {
                    key.push_back(-1000000000L - (long)(uint32_t)r);
                }
            }
            auto it = sig.find(key);
            if (it == sig.end()) {
                char buf[160];
                std::snprintf(buf, sizeof buf, "slot %d lane %d: no tape instruction with these operand values", s, ln);
                *err = buf;
                return false;
            }
            const int i = it->second;
            sig.erase(it);
            placed[i] = s;
            count++;
            // operands produced at least two slots earlier (the register read holds a value
            // that was written by an earlier slot; its slot is recorded per register)
            for (int k = 0; k < kArity[op]; k++)
                if (gref_kind(a[k]) == GREF_REG && reg_slot[gref_index(a[k])] > s - kMinDistance) {
                    *err = "operand produced too late";
                    return false;
                }
            const int o = P.out[(size_t)s * G + ln];
            if (o >= 0) writes.push_back({o, canon[i]});
        }
        // results of slot s may not overwrite a register that slot s+1 reads (when the
        // interpreter reads a slot's operands before it writes the previous slot's results)
        if (kMinDistance >= 2 && s + 1 < P.nslots)
            for (auto& w : writes)
                for (int ln = 0; ln < G; ln++) {
                    const int32_t* a = &P.arg[((size_t)(s + 1) * G + ln) * 4];
                    if (a[0] == GREF_NONE) continue;
                    for (int k = 0; k < kArity[P.op[s + 1]]; k++)
                        if (gref_kind(a[k]) == GREF_REG && gref_index(a[k]) == w.first) {
                            *err = "a result overwrites an operand of the next slot";
                            return false;
                        }
                }
        for (auto& w : writes) {
            if (w.first >= P.nregs) { *err = "result register out of range"; return false; }
            reg[w.first] = w.second;
            reg_slot[w.first] = s;
        }
    }
    if (count != N) { *err = "not every instruction was placed"; return false; }
    // outputs: each reference must hold the value the tape assigns
    for (size_t t = 0; t < tapes.size(); t++) {
        auto chk = [&](const Operand& o, int32_t r) {
            if (o.kind == 2) return r == o.ref;
            if (gref_kind(r) != GREF_REG || gref_index(r) >= P.nregs) return false;
            return reg[gref_index(r)] == (o.kind == 0 ? canon[o.id] : (long)(N + o.id));
        };
        for (size_t i = 0; i < L.u[t].size(); i++)
            if (!chk(L.u[t][i], P.io[t].u_ref[i])) { *err = "output u does not hold its value"; return false; }
        for (size_t i = 0; i < L.U[t].size(); i++)
            if (!chk(L.U[t][i], P.io[t].U_ref[i])) { *err = "output U does not hold its value"; return false; }
    }
    return true;
}

}  // namespace hcg
