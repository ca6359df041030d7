// Interpreter of lane-group slot programs (hcg_group_types.h).  Lane r of the group executes
// instruction r of every slot; operands come from the group's LDS registers, the constant
// table, or the homotopy's parameters formed at t.  Three value types: ComplexF64 (evaluate!,
// evaluate_and_jacobian!), ComplexDF64 (evaluate! with x in double-double) and truncated
// Taylor series of order K (taylor!), with the operation semantics of
// src/model_kit/operations.jl:183-248 and src/model_kit/taylor.jl:272-920 written with the
// same arithmetic as the oracle's tape interpreter (hcg_arith.h).
#pragma once
#include "hcg_arith.h"
#include "hcg_group_types.h"

struct GEval {
    const hcg_gprog* P;
    char* lds;           // this group's register file
    const double* consts;  // P->consts (g_bind)
    cplx t;              // parameters are formed at t (parameter homotopies)
    const double* p;     // start parameters (or null)
    const double* q;     // target parameters of the path's group (or null)
};

// The register file is addressed through local (address space 3) pointers: ds_read/ds_write.
// Through the generic pointer the accesses compile to flat instructions, whose waits also
// cover every global load in flight (among them the next slot's table entries).
typedef __attribute__((address_space(3))) double g_lds_d;
__device__ __forceinline__ cplx g_ld_c(const char* R, int k) {
    const g_lds_d* d = (const g_lds_d*)R + 2 * k;
    return mkc(d[0], d[1]);
}
__device__ __forceinline__ void g_st_c(char* R, int k, cplx v) {
    g_lds_d* d = (g_lds_d*)R + 2 * k;
    d[0] = v.re;
    d[1] = v.im;
}
__device__ __forceinline__ cdd g_ld_dd(const char* R, int k) {
    const g_lds_d* d = (const g_lds_d*)R + 4 * k;
    return mkcdd(mkdd(d[0], d[1]), mkdd(d[2], d[3]));
}
__device__ __forceinline__ void g_st_dd(char* R, int k, cdd v) {
    g_lds_d* d = (g_lds_d*)R + 4 * k;
    d[0] = v.re.hi;
    d[1] = v.re.lo;
    d[2] = v.im.hi;
    d[3] = v.im.lo;
}
template <int K> __device__ __forceinline__ ser<K> g_ld_s(const char* R, int k) {
    const g_lds_d* d = (const g_lds_d*)R + 2 * (K + 1) * k;
    ser<K> v;
#pragma unroll
    for (int j = 0; j <= K; j++) v.c[j] = mkc(d[2 * j], d[2 * j + 1]);
    return v;
}
template <int K> __device__ __forceinline__ void g_st_s(char* R, int k, const ser<K>& v) {
    g_lds_d* d = (g_lds_d*)R + 2 * (K + 1) * k;
#pragma unroll
    for (int j = 0; j <= K; j++) {
        d[2 * j] = v.c[j].re;
        d[2 * j + 1] = v.c[j].im;
    }
}

// Slot tables, constants and parameters are read through global (address space 1) pointers:
// global loads count only against the vector-memory counter, so waiting for an LDS operand
// does not also wait for them.
typedef __attribute__((address_space(1))) const double g_glob_d;
typedef __attribute__((address_space(1))) const int32_t g_glob_i;
typedef __attribute__((address_space(1))) const uint8_t g_glob_u8;
__device__ __forceinline__ cplx g_gl_c(const double* base, int k) {
    const g_glob_d* d = (const g_glob_d*)base + 2 * k;
    return mkc(d[0], d[1]);
}
__device__ __forceinline__ int g_gl_i(const int32_t* base, int k) { return ((const g_glob_i*)base)[k]; }
// par_lerp (hcg_arith.h) on the values p_k, q_k: the same operations in the same order
__device__ __forceinline__ cplx g_par_lerp(cplx t, cplx pk, cplx qk) {
    cplx rr = c_add(c_scale(t.re, pk), c_scale(1.0 - t.re, qk));
    cplx rc = c_add(c_mul(t, pk), c_mul(mkc(1.0 - t.re, -t.im), qk));
    const bool real_t = t.im == 0.0;
    return mkc(real_t ? rr.re : rc.re, real_t ? rr.im : rc.im);
}

__device__ __forceinline__ int g_int_imm(int32_t ref) { return ((int)((unsigned)ref << 4)) >> 4; }

// The lanes of a wave execute a slot together (LDS operations of one wave complete in order);
// this marks the slot boundary for the compiler, so no LDS access moves across it.
__device__ __forceinline__ void g_slot_end() { __builtin_amdgcn_wave_barrier(); }

// One slot's table entries for lane r.  The interpreter loads slot s+1's entries before it runs
// slot s, so their memory latency overlaps the slot's operand fetches and arithmetic.
// The table pointers are read from the program record once per run: the record is reached
// through a generic pointer, which the compiler cannot tell apart from the LDS the slots write,
// so it would otherwise read them again at every slot.
struct GSlot {
    int op, o;
    int4 a;
};
struct GTables {
    const uint8_t* op;
    const int32_t* arg;
    const int32_t* out;
    int nslots;
};
__device__ __forceinline__ GTables g_tables(const hcg_gprog* P) {
    GTables t;
    t.op = P->op;
    t.arg = P->arg;
    t.out = P->out;
    t.nslots = P->nslots;
    return t;
}
__device__ __forceinline__ GSlot g_slot(const GTables& P, int s, int r) {
    GSlot q;
    if (s < P.nslots && r < HCG_GROUP_LANES) {  // a slot holds 16 instructions: lanes 16..31 of a 32-lane group idle
        q.op = ((const g_glob_u8*)P.op)[s];
        const g_glob_i* ap = (const g_glob_i*)P.arg + 4 * (s * HCG_GROUP_LANES + r);
        q.a = make_int4(ap[0], ap[1], ap[2], ap[3]);
        q.o = g_gl_i(P.out, s * HCG_GROUP_LANES + r);
    } else {
        q.op = 0;
        q.a = make_int4(0, 0, 0, 0);
        q.o = -1;
    }
    return q;
}

// Points the evaluation at program P (and its constant table).
__device__ __forceinline__ void g_bind(GEval& E, const hcg_gprog* P) {
    E.P = P;
    E.consts = P->consts;
}

// ------------------------------- ComplexF64 -------------------------------
__device__ __forceinline__ cplx g_fetch_c(const GEval& E, int32_t ref) {
    const int kind = HCG_GREF_KIND(ref), idx = HCG_GREF_INDEX(ref);
    if (kind == HCG_GK_REG) return g_ld_c(E.lds, idx);
    if (kind == HCG_GK_CONST) return g_gl_c(E.consts, idx);
    if (kind == HCG_GK_PARAM) return g_par_lerp(E.t, g_gl_c(E.p, idx), g_gl_c(E.q, idx));
    return mkc(0.0, 0.0);
}
// The constants and parameter values p_k, q_k a slot entry's operands need, loaded one slot
// ahead (no load for register and immediate operands, nor for lanes without an instruction).
struct GVals {
    cplx p[4], q[4];
};
__device__ __forceinline__ void g_vals(const GEval& E, const GSlot& sl, GVals& v) {
    const int32_t ref[4] = {sl.a.x, sl.a.y, sl.a.z, sl.a.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int kind = HCG_GREF_KIND(ref[k]), idx = HCG_GREF_INDEX(ref[k]);
        if (sl.o >= 0 && kind == HCG_GK_CONST) v.p[k] = g_gl_c(E.consts, idx);
        if (sl.o >= 0 && kind == HCG_GK_PARAM) {
            v.p[k] = g_gl_c(E.p, idx);
            v.q[k] = g_gl_c(E.q, idx);
        }
    }
}
// g_fetch_c with the constant or parameter values already loaded
__device__ __forceinline__ cplx g_fetch_cv(const GEval& E, int32_t ref, cplx pk, cplx qk) {
    const int kind = HCG_GREF_KIND(ref), idx = HCG_GREF_INDEX(ref);
    if (kind == HCG_GK_REG) return g_ld_c(E.lds, idx);
    if (kind == HCG_GK_CONST) return pk;
    if (kind == HCG_GK_PARAM) return g_par_lerp(E.t, pk, qk);
    return mkc(0.0, 0.0);
}
__device__ __forceinline__ cplx g_op_c(int op, cplx a, cplx b, cplx c, cplx d, int e) {
    switch (op) {
    case 1: return c_cb(a);
    case 7: return c_inv(a);
    case 8: return c_inv_not_zero(a);
    case 9: return c_invsqr(a);
    case 10: return c_neg(a);
    case 13: return c_sqr(a);
    case 17: return a;
    case 18: return c_add(a, b);
    case 19: return c_div(a, b);
    case 20: return c_mul(a, b);
    case 21: return c_sub(a, b);
    case 22: return c_pow_int(a, e);
    case 24: return c_add3(a, b, c);
    case 25: return c_mul3(a, b, c);
    case 26: return c_muladd(a, b, c);
    case 27: return c_mulsub(a, b, c);
    case 28: return c_submul(a, b, c);
    case 29: return c_add4(a, b, c, d);
    case 30: return c_mul4(a, b, c, d);
    case 31: return c_mulmuladd(a, b, c, d);
    case 32: return c_mulmulsub(a, b, c, d);
    default: return mkc(HCG_NAN, HCG_NAN);
    }
}
// Loads lane r's variable value into each tape's variable register, runs every slot.
__device__ __forceinline__ void g_run_c(GEval E, int r, int n, cplx x) {
    const hcg_gprog* P = E.P;
    char* R = E.lds;
    if (r < n)
        for (int t = 0; t < P->ntapes; t++) {
            const int vr = g_gl_i(P->var_reg[t], r);
            if (vr >= 0) g_st_c(R, vr, x);
        }
    g_slot_end();
    // slot s runs with its table entries loaded two slots ahead and its constants and
    // parameter values one slot ahead, so no global load latency lies on the slot chain
    const GTables Tb = g_tables(P);
    GSlot next = g_slot(Tb, 0, r);
    GVals nv;
    g_vals(E, next, nv);
    GSlot next2 = g_slot(Tb, 1, r);
    for (int s = 0; s < Tb.nslots; s++) {
        const GSlot cur = next;
        const GVals cv = nv;
        next = next2;
        g_vals(E, next, nv);
        next2 = g_slot(Tb, s + 2, r);
        const int op = cur.op, o = cur.o;
        const int4 a = cur.a;
        if (o >= 0) {
            cplx va = g_fetch_cv(E, a.x, cv.p[0], cv.q[0]);
            cplx vb = op == 22 ? mkc(0.0, 0.0) : g_fetch_cv(E, a.y, cv.p[1], cv.q[1]);
            cplx vc = g_fetch_cv(E, a.z, cv.p[2], cv.q[2]), vd = g_fetch_cv(E, a.w, cv.p[3], cv.q[3]);
            g_st_c(R, o, g_op_c(op, va, vb, vc, vd, g_int_imm(a.y)));
        }
        g_slot_end();
    }
}

// ------------------------------- ComplexDF64 -------------------------------
__device__ __forceinline__ cdd g_fetch_dd(const GEval& E, int32_t ref) {
    const int kind = HCG_GREF_KIND(ref), idx = HCG_GREF_INDEX(ref);
    if (kind == HCG_GK_REG) return g_ld_dd(E.lds, idx);
    if (kind == HCG_GK_CONST) return cdd_from(g_gl_c(E.consts, idx));
    if (kind == HCG_GK_PARAM) return cdd_from(g_par_lerp(E.t, g_gl_c(E.p, idx), g_gl_c(E.q, idx)));
    return mkcdd(mkdd(0.0, 0.0), mkdd(0.0, 0.0));
}
__device__ __forceinline__ cdd g_op_dd(int op, cdd a, cdd b, cdd c, cdd d, int e) {
    switch (op) {
    case 1: return cdd_cb(a);
    case 7: return cdd_inv(a);
    case 8: return cdd_inv_not_zero(a);
    case 9: return cdd_invsqr(a);
    case 10: return cdd_neg(a);
    case 13: return cdd_sqr(a);
    case 17: return a;
    case 18: return cdd_add(a, b);
    case 19: return cdd_div(a, b);
    case 20: return cdd_mul(a, b);
    case 21: return cdd_sub(a, b);
    case 22: return cdd_pow_int(a, e);
    case 24: return cdd_add3(a, b, c);
    case 25: return cdd_mul3(a, b, c);
    case 26: return cdd_muladd(a, b, c);
    case 27: return cdd_mulsub(a, b, c);
    case 28: return cdd_submul(a, b, c);
    case 29: return cdd_add4(a, b, c, d);
    case 30: return cdd_mul4(a, b, c, d);
    case 31: return cdd_mulmuladd(a, b, c, d);
    case 32: return cdd_mulmulsub(a, b, c, d);
    default: return mkcdd(mkdd(HCG_NAN, 0.0), mkdd(HCG_NAN, 0.0));
    }
}
__device__ __forceinline__ void g_run_dd(GEval E, int r, int n, cplx x) {
    const hcg_gprog* P = E.P;
    char* R = E.lds;
    if (r < n)
        for (int t = 0; t < P->ntapes; t++) {
            const int vr = g_gl_i(P->var_reg[t], r);
            if (vr >= 0) g_st_dd(R, vr, cdd_from(x));
        }
    g_slot_end();
    const GTables Tb = g_tables(P);
    GSlot next = g_slot(Tb, 0, r);
    for (int s = 0; s < Tb.nslots; s++) {
        const GSlot cur = next;
        next = g_slot(Tb, s + 1, r);
        const int op = cur.op, o = cur.o;
        const int4 a = cur.a;
        if (o >= 0) {
            cdd va = g_fetch_dd(E, a.x);
            cdd vb = op == 22 ? cdd_from(mkc(0.0, 0.0)) : g_fetch_dd(E, a.y);
            cdd vc = g_fetch_dd(E, a.z), vd = g_fetch_dd(E, a.w);
            g_st_dd(R, o, g_op_dd(op, va, vb, vc, vd, g_int_imm(a.y)));
        }
        g_slot_end();
    }
}

// --------------------------- Taylor series, order K ---------------------------
template <int K> __device__ __forceinline__ ser<K> g_fetch_s(const GEval& E, int32_t ref) {
    const int kind = HCG_GREF_KIND(ref), idx = HCG_GREF_INDEX(ref);
    if (kind == HCG_GK_REG) return g_ld_s<K>(E.lds, idx);
    if (kind == HCG_GK_CONST) return s_const<K>(g_gl_c(E.consts, idx));
    if (kind == HCG_GK_PARAM) {  // p(t + s) = p(t) + s (p - q)
        const cplx pk = g_gl_c(E.p, idx), qk = g_gl_c(E.q, idx);
        ser<K> r = s_const<K>(g_par_lerp(E.t, pk, qk));
        r.c[1] = c_sub(pk, qk);  // par_diff
        return r;
    }
    return s_const<K>(mkc(0.0, 0.0));
}
template <int K> __device__ __forceinline__ ser<K> g_op_s(int op, const ser<K>& a, const ser<K>& b, const ser<K>& c,
                                                          const ser<K>& d, int e) {
    switch (op) {
    case 1: return s_cb<K>(a);
    case 7: return s_inv<K>(a);
    case 8: return s_inv<K>(a);
    case 9: return s_invsqr<K>(a);
    case 10: return s_neg<K>(a);
    case 13: return s_sqr<K>(a);
    case 17: return a;
    case 18: return s_add<K>(a, b);
    case 19: return s_div<K>(a, b);
    case 20: return s_mul<K>(a, b);
    case 21: return s_sub<K>(a, b);
    case 22: return s_pow_int<K>(a, e);
    case 24: return s_add3<K>(a, b, c);
    case 25: return s_mul3<K>(a, b, c);
    case 26: return s_muladd<K>(a, b, c);
    case 27: return s_mulsub<K>(a, b, c);
    case 28: return s_submul<K>(a, b, c);
    case 29: return s_add4<K>(a, b, c, d);
    case 30: return s_mul4<K>(a, b, c, d);
    case 31: return s_mulmuladd<K>(a, b, c, d);
    case 32: return s_mulmulsub<K>(a, b, c, d);
    default: return s_const<K>(mkc(HCG_NAN, HCG_NAN));
    }
}
// x0..x2: lane r's entries of the Taylor vectors of orders 0..K-1 (order K of x is zero)
template <int K> __device__ __forceinline__ void g_run_s(GEval E, int r, int n, cplx x0, cplx x1, cplx x2) {
    const hcg_gprog* P = E.P;
    char* R = E.lds;
    if (r < n) {
        ser<K> xs = s_const<K>(x0);
        if (K >= 2) xs.c[1] = x1;
        if (K >= 3) xs.c[2] = x2;
        for (int t = 0; t < P->ntapes; t++) {
            const int vr = g_gl_i(P->var_reg[t], r);
            if (vr >= 0) g_st_s<K>(R, vr, xs);
        }
    }
    g_slot_end();
    const GTables Tb = g_tables(P);
    GSlot next = g_slot(Tb, 0, r);
    for (int s = 0; s < Tb.nslots; s++) {
        const GSlot cur = next;
        next = g_slot(Tb, s + 1, r);
        const int op = cur.op, o = cur.o;
        const int4 a = cur.a;
        if (o >= 0) {
            ser<K> va = g_fetch_s<K>(E, a.x);
            ser<K> vb = op == 22 ? s_const<K>(mkc(0.0, 0.0)) : g_fetch_s<K>(E, a.y);
            ser<K> vc = g_fetch_s<K>(E, a.z), vd = g_fetch_s<K>(E, a.w);
            g_st_s<K>(R, o, g_op_s<K>(op, va, vb, vc, vd, g_int_imm(a.y)));
        }
        g_slot_end();
    }
}
// output reads after a run
__device__ __forceinline__ cplx g_read_c(const GEval& E, int32_t ref) { return g_fetch_c(E, ref); }
__device__ __forceinline__ cdd g_read_dd(const GEval& E, int32_t ref) { return g_fetch_dd(E, ref); }
template <int K> __device__ __forceinline__ cplx g_read_s(const GEval& E, int32_t ref, int k) {
    return g_fetch_s<K>(E, ref).c[k];
}
