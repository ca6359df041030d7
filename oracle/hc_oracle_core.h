/*
 * ORACLE — test infrastructure only. Never linked into the product library.
 *
 * Scalar / complex / double-double arithmetic used by the CPU restatement of
 * HomotopyContinuation.jl 2.20 (hc_oracle.c).  The product HIP kernels implement
 * the same arithmetic contract independently; this file exists so that the
 * oracle can be compiled with `-ffp-contract=off` and still use explicit FMAs
 * exactly where the contract says so.
 *
 * Arithmetic contract (shared by oracle and GPU, each implemented separately):
 *   cmul(a,b)      = (fma(ar,br,-(ai*bi)), fma(ar,bi, ai*br))
 *   cmuladd(a,b,c) = a*b + c with two nested FMAs per component
 *   abs2(a)        = fma(ar,ar, ai*ai)
 *   naive complex division (conj-multiply / abs2), as the reference's
 *   @fastmath divisions (src/linear_algebra.jl:171,288,343).
 */
#ifndef HC_ORACLE_CORE_H
#define HC_ORACLE_CORE_H
#include <math.h>
#include <stdint.h>
#include <float.h>

typedef struct { double re, im; } cplx;
typedef struct { double hi, lo; } dd;
typedef struct { dd re, im; } cdd;

static inline cplx C(double re, double im) { cplx z = {re, im}; return z; }
static inline cplx cadd(cplx a, cplx b) { return C(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return C(a.re - b.re, a.im - b.im); }
static inline cplx cneg(cplx a) { return C(-a.re, -a.im); }
static inline cplx cconj(cplx a) { return C(a.re, -a.im); }
static inline cplx cmul(cplx a, cplx b) { return C(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)); }
static inline cplx cmuladd(cplx a, cplx b, cplx c) {
    return C(fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im)));
}
/* c - a*b */
static inline cplx csubmul(cplx a, cplx b, cplx c) {
    return C(fma(-a.re, b.re, fma(a.im, b.im, c.re)), fma(-a.re, b.im, fma(-a.im, b.re, c.im)));
}
/* a*b - c */
static inline cplx cmulsub(cplx a, cplx b, cplx c) {
    return C(fma(a.re, b.re, fma(-a.im, b.im, -c.re)), fma(a.re, b.im, fma(a.im, b.re, -c.im)));
}
static inline double cabs2(cplx a) { return fma(a.re, a.re, a.im * a.im); }
/* fast_abs, src/utils.jl:207 */
static inline double fast_abs(cplx a) { return sqrt(cabs2(a)); }
static inline cplx cscale(double s, cplx a) { return C(s * a.re, s * a.im); }
static inline cplx cdivr(cplx a, double s) { return C(a.re / s, a.im / s); }
static inline cplx cdiv(cplx a, cplx b) {
    double d = cabs2(b);
    return C(fma(a.re, b.re, a.im * b.im) / d, fma(a.im, b.re, -(a.re * b.im)) / d);
}
static inline cplx cinv(cplx a) { double d = cabs2(a); return C(a.re / d, -a.im / d); }
static inline int cisnan(cplx a) { return isnan(a.re) || isnan(a.im); }

/* Julia's NaN-propagating min/max/clamp (Base.min, Base.max, Base.clamp). */
static inline double jmin(double a, double b) { if (isnan(a) || isnan(b)) return NAN; return a < b ? a : b; }
static inline double jmax(double a, double b) { if (isnan(a) || isnan(b)) return NAN; return a > b ? a : b; }
static inline double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }
/* nanmin/nanmax, src/utils.jl:211-212 */
static inline double nanmin(double a, double b) { return isnan(a) ? b : (isnan(b) ? a : jmin(a, b)); }
static inline double nanmax(double a, double b) { return isnan(a) ? b : (isnan(b) ? a : jmax(a, b)); }
/* Base.FastMath.max_fast: ifelse(y > x, y, x) */
static inline double max_fast(double x, double y) { return y > x ? y : x; }
/* Base.eps(x::Float64) */
static inline double jeps(double x) {
    if (!isfinite(x)) return NAN;
    x = fabs(x);
    if (x >= DBL_MIN) { int e; frexp(x, &e); return ldexp(1.0, e - 1 - 52); }
    return 4.9406564584124654e-324;
}
/* nthroot, src/utils.jl:408-422 */
static inline double nthroot(double x, int N) {
    if (N == 4) return sqrt(sqrt(x));
    if (N == 2) return sqrt(x);
    if (N == 3) return cbrt(x);
    if (N == 1) return x;
    if (N == 0) return 1.0;
    return pow(x, 1.0 / N);
}
/* integer power of a real by repeated squaring (used for s^(m-1), Δs^order) */
static inline double powi(double x, int n) {
    double r = 1.0, b = x; int k = n < 0 ? -n : n;
    while (k) { if (k & 1) r *= b; b *= b; k >>= 1; }
    return n < 0 ? 1.0 / r : r;
}
/* _h(a), src/tracker.jl:517 */
static inline double h_of(double a) { return 2 * a * (sqrt(4 * (a * a) + 1) - 2 * a); }

/* ---------------- double-double, src/DoubleDouble.jl:14-331 ---------------- */
static inline dd DD(double hi, double lo) { dd r = {hi, lo}; return r; }
static inline dd quick_two_sum(double a, double b) { double s = a + b; double e = b - (s - a); return DD(s, e); }
static inline dd two_sum(double a, double b) { double s = a + b; double v = s - a; double e = (a - (s - v)) + (b - v); return DD(s, e); }
static inline dd two_diff(double a, double b) { double s = a - b; double v = s - a; double e = (a - (s - v)) - (b + v); return DD(s, e); }
static inline dd two_prod(double a, double b) { double p = a * b; double e = fma(a, b, -p); return DD(p, e); }
static inline dd dd_from(double x) { return DD(x, isinf(x) ? INFINITY : 0.0); }
static inline dd dd_add_d(dd a, double b) { dd s = two_sum(a.hi, b); double lo = s.lo + a.lo; return quick_two_sum(s.hi, lo); }
static inline dd dd_add(dd a, dd b) { dd s = two_sum(a.hi, b.hi); double lo = s.lo + (a.lo + b.lo); return quick_two_sum(s.hi, lo); }
static inline dd dd_sub_d(dd a, double b) { dd s = two_diff(a.hi, b); double lo = s.lo + a.lo; return quick_two_sum(s.hi, lo); }
static inline dd d_sub_dd(double a, dd b) { dd s = two_diff(a, b.hi); double lo = s.lo - b.lo; return quick_two_sum(s.hi, lo); }
static inline dd dd_sub(dd a, dd b) { dd s = two_diff(a.hi, b.hi); double lo = s.lo + a.lo; lo = lo - b.lo; return quick_two_sum(s.hi, lo); }
static inline dd dd_neg(dd a) { return DD(-a.hi, -a.lo); }
static inline dd dd_mul_d(dd a, double b) { dd p = two_prod(a.hi, b); double p2 = p.lo + a.lo * b; return quick_two_sum(p.hi, p2); }
static inline dd dd_mul(dd a, dd b) { dd p = two_prod(a.hi, b.hi); double p2 = p.lo + (a.hi * b.lo + a.lo * b.hi); return quick_two_sum(p.hi, p2); }
static inline dd dd_div(dd a, dd b) {
    double q1 = a.hi / b.hi;
    dd r = dd_mul_d(b, q1);
    dd s = two_diff(a.hi, r.hi);
    double s2 = s.lo - r.lo; s2 = s2 + a.lo;
    double q2 = (s.hi + s2) / b.hi;
    return quick_two_sum(q1, q2);
}
static inline cdd CDD(dd re, dd im) { cdd z = {re, im}; return z; }
static inline cdd cdd_from(cplx a) { return CDD(dd_from(a.re), dd_from(a.im)); }
static inline cplx cdd_round(cdd a) { return C(a.re.hi, a.im.hi); }
static inline cdd cdd_add(cdd a, cdd b) { return CDD(dd_add(a.re, b.re), dd_add(a.im, b.im)); }
static inline cdd cdd_sub(cdd a, cdd b) { return CDD(dd_sub(a.re, b.re), dd_sub(a.im, b.im)); }
static inline cdd cdd_neg(cdd a) { return CDD(dd_neg(a.re), dd_neg(a.im)); }
/* Base complex `*` applied to DoubleF64 components */
static inline cdd cdd_mul(cdd a, cdd b) {
    return CDD(dd_sub(dd_mul(a.re, b.re), dd_mul(a.im, b.im)), dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re)));
}
/* ComplexF64 * ComplexDF64: Float64 * DoubleF64 dispatches to the mixed method */
static inline cdd cdd_mul_c(cplx a, cdd b) {
    return CDD(dd_sub(dd_mul_d(b.re, a.re), dd_mul_d(b.im, a.im)), dd_add(dd_mul_d(b.im, a.re), dd_mul_d(b.re, a.im)));
}
/* ComplexDF64 - ComplexF64 (mixed method per component) */
static inline cdd cdd_sub_c(cdd a, cplx b) { return CDD(dd_sub_d(a.re, b.re), dd_sub_d(a.im, b.im)); }
/* op_sqr(z::Complex), src/model_kit/operations.jl:199-202 */
static inline cdd cdd_sqr(cdd z) { return CDD(dd_mul(dd_add(z.re, z.im), dd_sub(z.re, z.im)), dd_mul(dd_add(z.re, z.re), z.im)); }
static inline cdd cdd_cb(cdd z) {
    dd a = dd_mul(dd_add(z.re, z.im), dd_sub(z.re, z.im));
    dd b = dd_mul(dd_add(z.re, z.re), z.im);
    return CDD(dd_sub(dd_mul(a, z.re), dd_mul(b, z.im)), dd_add(dd_mul(a, z.im), dd_mul(b, z.re)));
}
static inline cdd cdd_div(cdd a, cdd b) {
    dd d = dd_add(dd_mul(b.re, b.re), dd_mul(b.im, b.im));
    return CDD(dd_div(dd_add(dd_mul(a.re, b.re), dd_mul(a.im, b.im)), d),
               dd_div(dd_sub(dd_mul(a.im, b.re), dd_mul(a.re, b.im)), d));
}
static inline cdd cdd_inv(cdd a) { cdd one = {{1.0, 0.0}, {0.0, 0.0}}; return cdd_div(one, a); }

#endif
