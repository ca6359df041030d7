// FP64 interval arithmetic for the certification kernel: real intervals [lo, hi] and
// rectangular complex intervals, with the semantics of the reference's IntervalArithmetic
// module (src/interval_arithmetic.jl).  Every bound is computed in round-to-nearest and
// then moved one ulp outward (prevfloat / nextfloat, :57-61): a round-to-nearest result is
// within half an ulp of the exact value, so the moved bound encloses it.  The case
// analysis of products and quotients (:118-234) picks the same extremal products as the
// reference; it is written with selects rather than branches.
// Fused forms (muladd) use one FMA per bound, a single rounding.
#pragma once

struct ival { double lo, hi; };
struct icplx { ival re, im; };

// nextfloat / prevfloat of an FP64 value (NaN stays NaN, +Inf stays +Inf for nextfloat)
HD double iv_next(double x) {
    if (x != x || x == HCG_INF) return x;
    if (x == 0.0) return 4.9406564584124654e-324;
    const long long b = __double_as_longlong(x);
    return __longlong_as_double(x > 0.0 ? b + 1 : b - 1);
}
HD double iv_prev(double x) { return -iv_next(-x); }

HD ival mki(double lo, double hi) { ival r; r.lo = lo; r.hi = hi; return r; }
HD ival iv_pt(double x) { return mki(x, x); }
HD ival iv_round(double lo, double hi) { return mki(iv_prev(lo), iv_next(hi)); }  // @round
HD bool iv_iszero(ival a) { return a.lo == 0.0 && a.hi == 0.0; }
HD ival iv_nan() { return mki(HCG_NAN, HCG_NAN); }
HD double iv_mag(ival a) { return jmax(fabs(a.lo), fabs(a.hi)); }
HD bool iv_contains(ival a, double x) { return a.lo <= x && x <= a.hi; }
HD bool iv_interior(ival a, ival b) { return a.lo > b.lo && a.hi < b.hi; }  // isinterior(a, b)
HD double iv_min_fast(double x, double y) { return y < x ? y : x; }
HD double iv_max_fast(double x, double y) { return y > x ? y : x; }

HD ival iv_add(ival a, ival b) { return iv_round(a.lo + b.lo, a.hi + b.hi); }
HD ival iv_add_r(ival a, double b) { return iv_round(a.lo + b, a.hi + b); }
HD ival iv_sub(ival a, ival b) { return iv_round(a.lo - b.hi, a.hi - b.lo); }
HD ival iv_sub_r(ival a, double b) { return iv_round(a.lo - b, a.hi - b); }
HD ival iv_neg(ival a) { return mki(-a.hi, -a.lo); }

// x * a for a real x (:110-117)
HD ival iv_mul_r(double x, ival a) {
    const bool z = iv_iszero(a) || x == 0.0, pos = x >= 0.0;
    const ival r = iv_round((pos ? a.lo : a.hi) * x, (pos ? a.hi : a.lo) * x);
    return z ? mki(0.0, 0.0) : r;
}

// a * b (:118-144): lo = min of the products (l1a*l1b, l2a*l2b), hi = max of (h1a*h1b,
// h2a*h2b); both pairs coincide except when both intervals straddle zero
HD ival iv_mul(ival a, ival b) {
    const bool bpos = b.lo >= 0.0, bneg = !bpos && b.hi <= 0.0, bmix = !bpos && !bneg;
    const bool apos = bmix ? a.lo > 0.0 : a.lo >= 0.0;
    const bool aneg = !apos && (bmix ? a.hi < 0.0 : a.hi <= 0.0);
    const bool amix = !apos && !aneg;
    double l1a, l1b, h1a, h1b;
    if (bpos) {
        l1a = a.lo; l1b = apos ? b.lo : b.hi;
        h1a = a.hi; h1b = aneg ? b.lo : b.hi;
    } else if (bneg) {
        l1a = a.hi; l1b = aneg ? b.hi : b.lo;
        h1a = a.lo; h1b = apos ? b.hi : b.lo;
    } else {
        l1a = apos ? a.hi : a.lo; l1b = apos ? b.lo : b.hi;
        h1a = apos ? a.hi : a.lo; h1b = apos ? b.hi : b.lo;
    }
    const bool two = bmix && amix;
    const double l1 = l1a * l1b, h1 = h1a * h1b;
    const double l2 = two ? a.hi * b.lo : l1, h2 = two ? a.hi * b.hi : h1;
    return iv_round(two ? jmin(l1, l2) : l1, two ? jmax(h1, h2) : h1);
}

// muladd(a, b, c) with interval a, b, c (:146-166) and with a real first factor (:180-194)
HD ival iv_muladd(ival a, ival b, ival c) {
    double lo = iv_min_fast(iv_min_fast(iv_min_fast(fma(a.lo, b.lo, c.lo), fma(a.lo, b.hi, c.lo)),
                                        fma(a.hi, b.lo, c.lo)), fma(a.hi, b.hi, c.lo));
    double hi = iv_max_fast(iv_max_fast(iv_max_fast(fma(a.lo, b.lo, c.hi), fma(a.lo, b.hi, c.hi)),
                                        fma(a.hi, b.lo, c.hi)), fma(a.hi, b.hi, c.hi));
    return mki(iv_prev(lo), iv_next(hi));
}
HD ival iv_muladd_r(double a, ival b, ival c) {
    double lo = iv_min_fast(fma(a, b.lo, c.lo), fma(a, b.hi, c.lo));
    double hi = iv_max_fast(fma(a, b.lo, c.hi), fma(a, b.hi, c.hi));
    return mki(iv_prev(lo), iv_next(hi));
}

// a / b (:206-227): NaN interval when b contains zero (and a is not zero)
HD ival iv_div(ival a, ival b) {
    const bool bpos = b.lo > 0.0, bneg = !bpos && b.hi < 0.0;
    const bool apos = a.lo >= 0.0, aneg = !apos && a.hi <= 0.0;
    double ln, ld, hn, hd;
    if (bpos) {
        ln = a.lo; ld = apos ? b.hi : b.lo;
        hn = a.hi; hd = aneg ? b.hi : b.lo;
    } else {
        ln = a.hi; ld = aneg ? b.lo : b.hi;
        hn = a.lo; hd = apos ? b.lo : b.hi;
    }
    const ival r = iv_round(ln / ld, hn / hd);
    return iv_iszero(a) ? mki(0.0, 0.0) : ((bpos || bneg) ? r : iv_nan());
}
// inv(a) (:196-201)
HD ival iv_inv(ival a) {
    const ival r = iv_round(1.0 / a.hi, 1.0 / a.lo);
    return iv_contains(a, 0.0) ? iv_nan() : r;
}
// sqr(a) (:229-238) with mig(a) = 0 when 0 is in a, else min(|lo|, |hi|) (:72-75)
HD ival iv_sqr(ival a) {
    const bool pos = a.lo >= 0.0, neg = !pos && a.hi <= 0.0;
    const double m = iv_mag(a), g = iv_contains(a, 0.0) ? 0.0 : jmin(fabs(a.lo), fabs(a.hi));
    const double lo = pos ? a.lo * a.lo : (neg ? a.hi * a.hi : g * g);
    const double hi = pos ? a.hi * a.hi : (neg ? a.lo * a.lo : m * m);
    return iv_round(lo, hi);
}

// ---------------- complex intervals (IComplex, :263-409) ----------------
HD icplx mkic(ival re, ival im) { icplx z; z.re = re; z.im = im; return z; }
HD icplx ic_pt(cplx z) { return mkic(iv_pt(z.re), iv_pt(z.im)); }
HD icplx ic_zero() { return mkic(mki(0.0, 0.0), mki(0.0, 0.0)); }
HD icplx ic_ld(const double* P, int k) {  // parameter k of a packed (re lo, re hi, im lo, im hi) array
    return mkic(mki(P[4 * k], P[4 * k + 1]), mki(P[4 * k + 2], P[4 * k + 3]));
}
HD double ic_mag(icplx z) { return jmax(iv_mag(z.re), iv_mag(z.im)); }
HD bool ic_interior(icplx a, icplx b) { return iv_interior(a.re, b.re) && iv_interior(a.im, b.im); }
HD icplx ic_conj(icplx z) { return mkic(z.re, iv_neg(z.im)); }
HD icplx ci_neg(icplx a) { return mkic(iv_neg(a.re), iv_neg(a.im)); }
HD icplx ci_add(icplx a, icplx b) { return mkic(iv_add(a.re, b.re), iv_add(a.im, b.im)); }
HD icplx ci_sub(icplx a, icplx b) { return mkic(iv_sub(a.re, b.re), iv_sub(a.im, b.im)); }
__device__ __noinline__ icplx ci_mul(icplx z, icplx w) {
    return mkic(iv_sub(iv_mul(z.re, w.re), iv_mul(z.im, w.im)), iv_add(iv_mul(z.re, w.im), iv_mul(z.im, w.re)));
}
HD icplx ci_inv(icplx b) {
    const ival d = iv_add(iv_sqr(b.re), iv_sqr(b.im));
    return mkic(iv_div(b.re, d), iv_div(iv_neg(b.im), d));
}
HD icplx ci_div(icplx a, icplx b) {
    const ival d = iv_add(iv_sqr(b.re), iv_sqr(b.im));
    return mkic(iv_div(iv_add(iv_mul(a.re, b.re), iv_mul(a.im, b.im)), d),
                iv_div(iv_sub(iv_mul(a.im, b.re), iv_mul(a.re, b.im)), d));
}
// muladd(z, w, x) for a point or interval z (:372-375)
HD icplx ic_muladd_pt(cplx z, icplx w, icplx x) {
    return mkic(iv_sub(iv_muladd_r(z.re, w.re, x.re), iv_mul_r(z.im, w.im)),
                iv_muladd_r(z.re, w.im, iv_muladd_r(z.im, w.re, x.im)));
}
HD icplx ic_muladd(icplx z, icplx w, icplx x) {
    return mkic(iv_sub(iv_muladd(z.re, w.re, x.re), iv_mul(z.im, w.im)),
                iv_muladd(z.re, w.im, iv_muladd(z.im, w.re, x.im)));
}

// Straight-line-program operations on IComplex values: the generic definitions of
// src/model_kit/operations.jl:183-248 (x*x*x, x*x, x*y + z, ...), which is what the
// reference's interpreter runs for IComplexF64.
HD icplx ci_sqr(icplx a) { return ci_mul(a, a); }
HD icplx ci_cb(icplx a) { return ci_mul(ci_mul(a, a), a); }
HD bool ci_iszero(icplx a) { return iv_iszero(a.re) && iv_iszero(a.im); }
HD icplx ci_inv_not_zero(icplx a) {
    const icplx r = ci_inv(a);
    return ci_iszero(a) ? a : r;
}
HD icplx ci_invsqr(icplx a) { return ci_sqr(ci_inv(a)); }
HD icplx ci_add3(icplx a, icplx b, icplx c) { return ci_add(ci_add(a, b), c); }
HD icplx ci_add4(icplx a, icplx b, icplx c, icplx d) { return ci_add(ci_add(ci_add(a, b), c), d); }
HD icplx ci_mul3(icplx a, icplx b, icplx c) { return ci_mul(ci_mul(a, b), c); }
HD icplx ci_mul4(icplx a, icplx b, icplx c, icplx d) { return ci_mul(ci_mul(ci_mul(a, b), c), d); }
HD icplx ci_muladd(icplx a, icplx b, icplx c) { return ci_add(ci_mul(a, b), c); }
HD icplx ci_mulsub(icplx a, icplx b, icplx c) { return ci_sub(ci_mul(a, b), c); }
HD icplx ci_submul(icplx a, icplx b, icplx c) { return ci_sub(c, ci_mul(a, b)); }
HD icplx ci_mulmuladd(icplx a, icplx b, icplx c, icplx d) { return ci_add(ci_mul(a, b), ci_mul(c, d)); }
HD icplx ci_mulmulsub(icplx a, icplx b, icplx c, icplx d) { return ci_sub(ci_mul(a, b), ci_mul(c, d)); }
// Base.power_by_squaring on IComplex, p >= 0
__device__ __noinline__ icplx ci_pbs(icplx x, int p) {
    if (p == 1) return x;
    if (p == 0) return mkic(mki(1.0, 1.0), mki(0.0, 0.0));
    if (p == 2) return ci_mul(x, x);
    int t = __builtin_ctz((unsigned)p) + 1;
    p >>= t;
    while (--t > 0) x = ci_mul(x, x);
    icplx y = x;
    while (p > 0) {
        t = __builtin_ctz((unsigned)p) + 1;
        p >>= t;
        while (--t >= 0) x = ci_mul(x, x);
        y = ci_mul(y, x);
    }
    return y;
}
HD icplx ci_pow_int(icplx x, int p) { return p > 0 ? ci_pbs(x, p) : ci_inv(ci_pbs(x, -p)); }

// ---------------- double-double balls (extended-precision residual) ----------------
// A real ball (m, r) holds every v with |v - (m.hi + m.lo)| <= r.  Midpoints use
// double-double arithmetic whose rounding errors have published bounds (Joldes, Muller,
// Popescu, "Tight and rigorous error bounds for basic building blocks of double-word
// arithmetic", ACM TOMS 44(2), 2017): the accurate sum (AccurateDWPlusDW) errs by less than
// 3u²/(1-4u) of the exact sum (so by less than 3.1u² of the computed one), the product
// (DWTimesDW1, which dd_mul is) by less than 7u² of the exact product, u = 2^-53.  The
// radius adds 8u² and 16u² of the magnitudes plus 1e-300 (underflow), rounded upward.
struct dball { ddr m; double r; };
struct cball { dball re, im; };
#define BALL_U2 1.2325951644078310e-32  // u² = 2^-106

HD double up_(double x) { return iv_next(x); }
HD double dd_mag(ddr a) { return up_(fabs(a.hi) + fabs(a.lo)); }
HD ddr dd_add_acc(ddr x, ddr y) {
    const ddr s = two_sum(x.hi, y.hi), t = two_sum(x.lo, y.lo);
    const double c = s.lo + t.hi;
    const ddr v = quick_two_sum(s.hi, c);
    const double w = t.lo + v.lo;
    return quick_two_sum(v.hi, w);
}
HD dball mkdb(ddr m, double r) { dball b; b.m = m; b.r = r; return b; }
HD dball db_add(dball a, dball b) {
    const ddr m = dd_add_acc(a.m, b.m);
    const double e = up_(up_(8.0 * BALL_U2 * dd_mag(m)) + 1e-300);
    return mkdb(m, up_(up_(a.r + b.r) + e));
}
HD dball db_neg(dball a) { return mkdb(mkdd(-a.m.hi, -a.m.lo), a.r); }
HD dball db_sub(dball a, dball b) { return db_add(a, db_neg(b)); }
HD dball db_mul(dball a, dball b) {
    const ddr m = dd_mul(a.m, b.m);
    const double ma = dd_mag(a.m), mb = dd_mag(b.m);
    const double e = up_(up_(16.0 * BALL_U2 * up_(ma * mb)) + 1e-300);
    const double r = up_(up_(up_(ma * b.r) + up_(mb * a.r)) + up_(a.r * b.r));
    return mkdb(m, up_(r + e));
}
// the FP64 interval that encloses a ball
HD ival db_to_ival(dball a) {
    const double t = a.m.hi + a.m.lo;
    return mki(iv_prev(iv_prev(t) - a.r), iv_next(iv_next(t) + a.r));
}

HD cball mkcb(dball re, dball im) { cball z; z.re = re; z.im = im; return z; }
HD cball cb_pt(cplx z) { return mkcb(mkdb(mkdd(z.re, 0.0), 0.0), mkdb(mkdd(z.im, 0.0), 0.0)); }
HD cball cb_ddpt(cdd z) { return mkcb(mkdb(z.re, 0.0), mkdb(z.im, 0.0)); }
HD cball cb_ld(const double* B, int k) {  // parameter k of a packed (re hi, re lo, re rad, im hi, im lo, im rad) array
    return mkcb(mkdb(mkdd(B[6 * k], B[6 * k + 1]), B[6 * k + 2]), mkdb(mkdd(B[6 * k + 3], B[6 * k + 4]), B[6 * k + 5]));
}
HD cball cb_nan() { return mkcb(mkdb(mkdd(HCG_NAN, 0.0), HCG_INF), mkdb(mkdd(HCG_NAN, 0.0), HCG_INF)); }
HD icplx cb_to_ic(cball z) { return mkic(db_to_ival(z.re), db_to_ival(z.im)); }
HD cball cb_neg(cball a) { return mkcb(db_neg(a.re), db_neg(a.im)); }
HD cball cb_add(cball a, cball b) { return mkcb(db_add(a.re, b.re), db_add(a.im, b.im)); }
HD cball cb_sub(cball a, cball b) { return mkcb(db_sub(a.re, b.re), db_sub(a.im, b.im)); }
__device__ __noinline__ cball cb_mul(cball a, cball b) {
    return mkcb(db_sub(db_mul(a.re, b.re), db_mul(a.im, b.im)), db_add(db_mul(a.re, b.im), db_mul(a.im, b.re)));
}
HD cball cb_sqr(cball a) { return cb_mul(a, a); }
HD cball cb_cb(cball a) { return cb_mul(cb_mul(a, a), a); }
HD cball cb_add3(cball a, cball b, cball c) { return cb_add(cb_add(a, b), c); }
HD cball cb_add4(cball a, cball b, cball c, cball d) { return cb_add(cb_add(cb_add(a, b), c), d); }
HD cball cb_mul3(cball a, cball b, cball c) { return cb_mul(cb_mul(a, b), c); }
HD cball cb_mul4(cball a, cball b, cball c, cball d) { return cb_mul(cb_mul(cb_mul(a, b), c), d); }
HD cball cb_muladd(cball a, cball b, cball c) { return cb_add(cb_mul(a, b), c); }
HD cball cb_mulsub(cball a, cball b, cball c) { return cb_sub(cb_mul(a, b), c); }
HD cball cb_submul(cball a, cball b, cball c) { return cb_sub(c, cb_mul(a, b)); }
HD cball cb_mulmuladd(cball a, cball b, cball c, cball d) { return cb_add(cb_mul(a, b), cb_mul(c, d)); }
HD cball cb_mulmulsub(cball a, cball b, cball c, cball d) { return cb_sub(cb_mul(a, b), cb_mul(c, d)); }
// x^p for p >= 1 by squaring; quotients have no ball form here (the residual then stays unbounded)
__device__ __noinline__ cball cb_pow_int(cball x, int p) {
    if (p < 1) return cb_nan();
    if (p == 1) return x;
    cball y = x;
    bool have = false;
    while (p > 0) {
        if (p & 1) { y = have ? cb_mul(y, x) : x; have = true; }
        p >>= 1;
        if (p) x = cb_mul(x, x);
    }
    return y;
}

// ---------------- interpreters over the generated tables ----------------
// Operands: >= 0 register, < 0 constant -e-1 (a point), >= CI_PBASE parameter e - CI_PBASE
// (its box / ball).  Op codes are the OpType ordinals (src/model_kit/operations.jl:5-49).
#define CI_PBASE (1 << 24)
HD icplx ci_fetch(const double* K, const icplx* R, const double* IP, int e) {
    if (e >= CI_PBASE) return ic_ld(IP, e - CI_PBASE);
    return e >= 0 ? R[e] : ic_pt(mkc(K[2 * (-e - 1)], K[2 * (-e - 1) + 1]));
}
HD cball cb_fetch(const double* K, const cball* R, const double* BP, int e) {
    if (e >= CI_PBASE) return cb_ld(BP, e - CI_PBASE);
    return e >= 0 ? R[e] : cb_pt(mkc(K[2 * (-e - 1)], K[2 * (-e - 1) + 1]));
}
__device__ __noinline__ void ci_interpret(const int* code, int ninstr, const double* K, icplx* R, const double* IP) {
    for (int i = 0; i < ninstr; i++) {
        const int* I = code + 6 * i;
        const icplx a = ci_fetch(K, R, IP, I[1]);
        icplx r;
        switch (I[0]) {
        case 1: r = ci_cb(a); break;                       // OP_CB
        case 7: r = ci_inv(a); break;                      // OP_INV
        case 8: r = ci_inv_not_zero(a); break;             // OP_INV_NOT_ZERO
        case 9: r = ci_invsqr(a); break;                   // OP_INVSQR
        case 10: r = ci_neg(a); break;                     // OP_NEG
        case 13: r = ci_sqr(a); break;                     // OP_SQR
        case 17: r = a; break;                             // OP_IDENTITY
        case 18: r = ci_add(a, ci_fetch(K, R, IP, I[2])); break;
        case 19: r = ci_div(a, ci_fetch(K, R, IP, I[2])); break;
        case 20: r = ci_mul(a, ci_fetch(K, R, IP, I[2])); break;
        case 21: r = ci_sub(a, ci_fetch(K, R, IP, I[2])); break;
        case 22: r = ci_pow_int(a, I[2]); break;           // OP_POW_INT (exponent inline)
        case 24: r = ci_add3(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3])); break;
        case 25: r = ci_mul3(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3])); break;
        case 26: r = ci_muladd(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3])); break;
        case 27: r = ci_mulsub(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3])); break;
        case 28: r = ci_submul(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3])); break;
        case 29: r = ci_add4(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3]), ci_fetch(K, R, IP, I[4])); break;
        case 30: r = ci_mul4(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3]), ci_fetch(K, R, IP, I[4])); break;
        case 31: r = ci_mulmuladd(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3]), ci_fetch(K, R, IP, I[4])); break;
        case 32: r = ci_mulmulsub(a, ci_fetch(K, R, IP, I[2]), ci_fetch(K, R, IP, I[3]), ci_fetch(K, R, IP, I[4])); break;
        default: r = mkic(iv_nan(), iv_nan());
        }
        R[I[5]] = r;
    }
}
__device__ __noinline__ void cb_interpret(const int* code, int ninstr, const double* K, cball* R, const double* BP) {
    for (int i = 0; i < ninstr; i++) {
        const int* I = code + 6 * i;
        const cball a = cb_fetch(K, R, BP, I[1]);
        cball r;
        switch (I[0]) {
        case 1: r = cb_cb(a); break;
        case 10: r = cb_neg(a); break;
        case 13: r = cb_sqr(a); break;
        case 17: r = a; break;
        case 18: r = cb_add(a, cb_fetch(K, R, BP, I[2])); break;
        case 20: r = cb_mul(a, cb_fetch(K, R, BP, I[2])); break;
        case 21: r = cb_sub(a, cb_fetch(K, R, BP, I[2])); break;
        case 22: r = cb_pow_int(a, I[2]); break;
        case 24: r = cb_add3(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3])); break;
        case 25: r = cb_mul3(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3])); break;
        case 26: r = cb_muladd(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3])); break;
        case 27: r = cb_mulsub(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3])); break;
        case 28: r = cb_submul(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3])); break;
        case 29: r = cb_add4(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3]), cb_fetch(K, R, BP, I[4])); break;
        case 30: r = cb_mul4(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3]), cb_fetch(K, R, BP, I[4])); break;
        case 31: r = cb_mulmuladd(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3]), cb_fetch(K, R, BP, I[4])); break;
        case 32: r = cb_mulmulsub(a, cb_fetch(K, R, BP, I[2]), cb_fetch(K, R, BP, I[3]), cb_fetch(K, R, BP, I[4])); break;
        default: r = cb_nan();  // inv, div: no ball form
        }
        R[I[5]] = r;
    }
}
