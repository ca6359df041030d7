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
/* Base.abs(::Complex) = hypot(re, im): overflow-free.  Scaling by a power of two is exact, so
 * this is deterministic (IEEE-exact operations only) and within one ulp. */
static inline double chypot(cplx a) {
    double x = fabs(a.re), y = fabs(a.im);
    if (isinf(x) || isinf(y)) return INFINITY;
    if (isnan(x) || isnan(y)) return NAN;
    double m = x > y ? x : y;
    if (m == 0) return 0.0;
    int e;
    frexp(m, &e);
    double xs = ldexp(x, -e), ys = ldexp(y, -e);
    return ldexp(sqrt(fma(xs, xs, ys * ys)), e);
}
static inline cplx cscale(double s, cplx a) { return C(s * a.re, s * a.im); }
static inline cplx cdivr(cplx a, double s) { return C(a.re / s, a.im / s); }
static inline cplx cdiv(cplx a, cplx b) {
    double d = cabs2(b);
    return C(fma(a.re, b.re, a.im * b.im) / d, fma(a.im, b.re, -(a.re * b.im)) / d);
}
static inline cplx cinv(cplx a) { double d = cabs2(a); return C(a.re / d, -a.im / d); }
static inline int cisnan(cplx a) { return isnan(a.re) || isnan(a.im); }


/* Deterministic log / exp / pow / cbrt built from IEEE-exact operations only
 * (+ - * / sqrt, frexp/ldexp), so that the CPU oracle and the GPU kernels produce
 * identical bits; accuracy is about 1 ulp.  log: atanh series on m in [sqrt(1/2), sqrt(2));
 * exp: k*ln2 range reduction + degree-16 Taylor; pow(x,y) = exp(y log x) for x > 0;
 * cbrt: exp(log|x|/3) plus one Newton correction. */
#define HC_LN2_HI 6.93147180369123816490e-01
#define HC_LN2_LO 1.90821492927058770002e-10
static inline double hc_log(double x) {
    if (isnan(x) || x < 0) return NAN;
    if (x == 0) return -INFINITY;
    if (isinf(x)) return x;
    int e;
    double m = frexp(x, &e);               /* x = m 2^e, m in [0.5, 1) */
    if (m < 0.70710678118654752440) { m = 2.0 * m; e -= 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 25.0;
    p = p * s2 + 1.0 / 23.0; p = p * s2 + 1.0 / 21.0; p = p * s2 + 1.0 / 19.0; p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0; p = p * s2 + 1.0 / 13.0; p = p * s2 + 1.0 / 11.0; p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0; p = p * s2 + 1.0 / 5.0; p = p * s2 + 1.0 / 3.0;
    double lm = 2.0 * s + 2.0 * s * (s2 * p);
    double de = (double)e;
    return de * HC_LN2_HI + (de * HC_LN2_LO + lm);
}
static inline double hc_exp(double x) {
    if (isnan(x)) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.2) return 0.0;
    double kf = nearbyint(x * 1.44269504088896340736);
    double r = (x - kf * HC_LN2_HI) - kf * HC_LN2_LO;
    double p = 1.0 / 20922789888000.0;           /* 1/16! */
    const double inv[16] = {1.0 / 1307674368000.0, 1.0 / 87178291200.0, 1.0 / 6227020800.0, 1.0 / 479001600.0,
                            1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0,
                            1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0, 1.0};
    for (int i = 0; i < 16; i++) p = p * r + inv[i];
    int k = (int)kf;
    if (k > 1000) { p = ldexp(p, 1000); k -= 1000; }
    if (k < -1000) { p = ldexp(p, -1000); k += 1000; }
    return ldexp(p, k);
}
static inline double hc_pow(double x, double y) {
    if (y == 0) return 1.0;
    if (x == 1.0) return 1.0;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0) return y > 0 ? 0.0 : INFINITY;
    if (isinf(x)) return y > 0 ? INFINITY : 0.0;
    if (x < 0) return NAN;
    return hc_exp(y * hc_log(x));
}
static inline double hc_cbrt(double x) {
    if (x == 0 || isnan(x) || isinf(x)) return x;
    double a = fabs(x);
    double y = hc_exp(hc_log(a) / 3.0);
    y = y - (y * y * y - a) / (3.0 * (y * y));
    return x < 0 ? -y : y;
}

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
    if (N == 3) return hc_cbrt(x);
    if (N == 1) return x;
    if (N == 0) return 1.0;
    return hc_pow(x, 1.0 / N);
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
