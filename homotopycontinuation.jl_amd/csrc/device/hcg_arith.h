// Device arithmetic for the generated path-tracker kernels (compiled by hipRTC for gfx950).
// Complex ComplexF64 arithmetic, double-double (ComplexDF64) arithmetic and truncated
// Taylor series of order <= 3, with the same FMA contract as the oracle's: FMAs are
// written explicitly, everything else is compiled with -ffp-contract=off.
//   double-double: src/DoubleDouble.jl:14-331
//   SLP op semantics: src/model_kit/operations.jl:183-248
//   Taylor rules: src/model_kit/taylor.jl:272-920
#pragma once

#define HCG_INF __builtin_inf()
#define HCG_NAN __builtin_nan("")
#define HCG_PI 3.14159265358979323846

struct cplx { double re, im; };
struct ddr { double hi, lo; };
struct cdd { ddr re, im; };

#define HD __device__ __forceinline__

HD cplx mkc(double re, double im) { cplx z; z.re = re; z.im = im; return z; }
HD cplx c_add(cplx a, cplx b) { return mkc(a.re + b.re, a.im + b.im); }
HD cplx c_sub(cplx a, cplx b) { return mkc(a.re - b.re, a.im - b.im); }
HD cplx c_neg(cplx a) { return mkc(-a.re, -a.im); }
HD cplx c_conj(cplx a) { return mkc(a.re, -a.im); }
HD cplx c_mul(cplx a, cplx b) { return mkc(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)); }
HD cplx c_muladd(cplx a, cplx b, cplx c) {
    return mkc(fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im)));
}
HD cplx c_submul(cplx a, cplx b, cplx c) {  // c - a*b
    return mkc(fma(-a.re, b.re, fma(a.im, b.im, c.re)), fma(-a.re, b.im, fma(-a.im, b.re, c.im)));
}
HD cplx c_mulsub(cplx a, cplx b, cplx c) {  // a*b - c
    return mkc(fma(a.re, b.re, fma(-a.im, b.im, -c.re)), fma(a.re, b.im, fma(a.im, b.re, -c.im)));
}
HD double c_abs2(cplx a) { return fma(a.re, a.re, a.im * a.im); }
HD double c_fabs(cplx a) { return sqrt(c_abs2(a)); }
// Base.abs(::Complex) = hypot(re, im), overflow-free; power-of-two scaling is exact, so the
// result is the same on every device (IEEE-exact operations only)
HD double c_hypot(cplx a) {
    double x = fabs(a.re), y = fabs(a.im);
    if (isinf(x) || isinf(y)) return HCG_INF;
    if (isnan(x) || isnan(y)) return HCG_NAN;
    double m = x > y ? x : y;
    if (m == 0) return 0.0;
    int e;
    frexp(m, &e);
    double xs = ldexp(x, -e), ys = ldexp(y, -e);
    return ldexp(sqrt(fma(xs, xs, ys * ys)), e);
}
HD cplx c_scale(double s, cplx a) { return mkc(s * a.re, s * a.im); }
HD cplx c_divr(cplx a, double s) { return mkc(a.re / s, a.im / s); }
HD cplx c_div(cplx a, cplx b) {
    double d = c_abs2(b);
    return mkc(fma(a.re, b.re, a.im * b.im) / d, fma(a.im, b.re, -(a.re * b.im)) / d);
}
HD cplx c_inv(cplx a) { double d = c_abs2(a); return mkc(a.re / d, -a.im / d); }
HD bool c_isnan(cplx a) { return isnan(a.re) || isnan(a.im); }
HD cplx c_sqr(cplx z) { return mkc((z.re + z.im) * (z.re - z.im), (z.re + z.re) * z.im); }
HD cplx c_cb(cplx z) {
    double a = (z.re + z.im) * (z.re - z.im), b = (z.re + z.re) * z.im;
    return mkc(fma(a, z.re, -(b * z.im)), fma(a, z.im, b * z.re));
}
HD cplx c_add3(cplx a, cplx b, cplx c) { return c_add(c_add(a, b), c); }
HD cplx c_add4(cplx a, cplx b, cplx c, cplx d) { return c_add(c_add(c_add(a, b), c), d); }
HD cplx c_mul3(cplx a, cplx b, cplx c) { return c_mul(c_mul(a, b), c); }
HD cplx c_mul4(cplx a, cplx b, cplx c, cplx d) { return c_mul(c_mul(c_mul(a, b), c), d); }
HD cplx c_mulmuladd(cplx a, cplx b, cplx c, cplx d) { return c_muladd(a, b, c_mul(c, d)); }
HD cplx c_mulmulsub(cplx a, cplx b, cplx c, cplx d) { return c_submul(c, d, c_mul(a, b)); }
HD cplx c_inv_not_zero(cplx a) { return (a.re == 0.0 && a.im == 0.0) ? a : c_inv(a); }
HD cplx c_invsqr(cplx a) { return c_sqr(c_inv(a)); }
// Base.power_by_squaring, p >= 0
__device__ __noinline__ cplx c_pbs(cplx x, int p) {
    if (p == 1) return x;
    if (p == 0) return mkc(1.0, 0.0);
    if (p == 2) return c_mul(x, x);
    int t = __builtin_ctz((unsigned)p) + 1;
    p >>= t;
    while (--t > 0) x = c_mul(x, x);
    cplx y = x;
    while (p > 0) {
        t = __builtin_ctz((unsigned)p) + 1;
        p >>= t;
        while (--t >= 0) x = c_mul(x, x);
        y = c_mul(y, x);
    }
    return y;
}
HD cplx c_pow_int(cplx x, int p) { return p > 0 ? c_pbs(x, p) : c_inv(c_pbs(x, -p)); }

// ---------------------------- double-double ----------------------------
HD ddr mkdd(double hi, double lo) { ddr r; r.hi = hi; r.lo = lo; return r; }
HD ddr quick_two_sum(double a, double b) { double s = a + b; double e = b - (s - a); return mkdd(s, e); }
HD ddr two_sum(double a, double b) { double s = a + b; double v = s - a; double e = (a - (s - v)) + (b - v); return mkdd(s, e); }
HD ddr two_diff(double a, double b) { double s = a - b; double v = s - a; double e = (a - (s - v)) - (b + v); return mkdd(s, e); }
HD ddr two_prod(double a, double b) { double p = a * b; double e = fma(a, b, -p); return mkdd(p, e); }
HD ddr dd_from(double x) { return mkdd(x, isinf(x) ? HCG_INF : 0.0); }
HD ddr dd_add(ddr a, ddr b) { ddr s = two_sum(a.hi, b.hi); double lo = s.lo + (a.lo + b.lo); return quick_two_sum(s.hi, lo); }
HD ddr dd_sub(ddr a, ddr b) { ddr s = two_diff(a.hi, b.hi); double lo = s.lo + a.lo; lo = lo - b.lo; return quick_two_sum(s.hi, lo); }
HD ddr dd_sub_d(ddr a, double b) { ddr s = two_diff(a.hi, b); double lo = s.lo + a.lo; return quick_two_sum(s.hi, lo); }
HD ddr dd_neg(ddr a) { return mkdd(-a.hi, -a.lo); }
HD ddr dd_mul(ddr a, ddr b) { ddr p = two_prod(a.hi, b.hi); double p2 = p.lo + (a.hi * b.lo + a.lo * b.hi); return quick_two_sum(p.hi, p2); }
HD ddr dd_mul_d(ddr a, double b) { ddr p = two_prod(a.hi, b); double p2 = p.lo + a.lo * b; return quick_two_sum(p.hi, p2); }
HD ddr dd_div(ddr a, ddr b) {
    double q1 = a.hi / b.hi;
    ddr r = dd_mul_d(b, q1);
    ddr s = two_diff(a.hi, r.hi);
    double s2 = s.lo - r.lo; s2 = s2 + a.lo;
    double q2 = (s.hi + s2) / b.hi;
    return quick_two_sum(q1, q2);
}
HD cdd mkcdd(ddr re, ddr im) { cdd z; z.re = re; z.im = im; return z; }
HD cdd cdd_from(cplx a) { return mkcdd(dd_from(a.re), dd_from(a.im)); }
HD cplx cdd_round(cdd a) { return mkc(a.re.hi, a.im.hi); }
HD cdd cdd_add(cdd a, cdd b) { return mkcdd(dd_add(a.re, b.re), dd_add(a.im, b.im)); }
HD cdd cdd_sub(cdd a, cdd b) { return mkcdd(dd_sub(a.re, b.re), dd_sub(a.im, b.im)); }
HD cdd cdd_neg(cdd a) { return mkcdd(dd_neg(a.re), dd_neg(a.im)); }
HD cdd cdd_mul(cdd a, cdd b) {
    return mkcdd(dd_sub(dd_mul(a.re, b.re), dd_mul(a.im, b.im)), dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re)));
}
// ComplexF64 * ComplexDF64 (mixed Float64 * DoubleF64 per component)
HD cdd cdd_mul_c(cplx a, cdd b) {
    return mkcdd(dd_sub(dd_mul_d(b.re, a.re), dd_mul_d(b.im, a.im)), dd_add(dd_mul_d(b.im, a.re), dd_mul_d(b.re, a.im)));
}
HD cdd cdd_sub_c(cdd a, cplx b) { return mkcdd(dd_sub_d(a.re, b.re), dd_sub_d(a.im, b.im)); }
HD cdd cdd_sqr(cdd z) { return mkcdd(dd_mul(dd_add(z.re, z.im), dd_sub(z.re, z.im)), dd_mul(dd_add(z.re, z.re), z.im)); }
HD cdd cdd_cb(cdd z) {
    ddr a = dd_mul(dd_add(z.re, z.im), dd_sub(z.re, z.im));
    ddr b = dd_mul(dd_add(z.re, z.re), z.im);
    return mkcdd(dd_sub(dd_mul(a, z.re), dd_mul(b, z.im)), dd_add(dd_mul(a, z.im), dd_mul(b, z.re)));
}
HD cdd cdd_div(cdd a, cdd b) {
    ddr d = dd_add(dd_mul(b.re, b.re), dd_mul(b.im, b.im));
    return mkcdd(dd_div(dd_add(dd_mul(a.re, b.re), dd_mul(a.im, b.im)), d),
                 dd_div(dd_sub(dd_mul(a.im, b.re), dd_mul(a.re, b.im)), d));
}
HD cdd cdd_inv(cdd a) { return cdd_div(mkcdd(mkdd(1.0, 0.0), mkdd(0.0, 0.0)), a); }
HD cdd cdd_add3(cdd a, cdd b, cdd c) { return cdd_add(cdd_add(a, b), c); }
HD cdd cdd_add4(cdd a, cdd b, cdd c, cdd d) { return cdd_add(cdd_add(cdd_add(a, b), c), d); }
HD cdd cdd_mul3(cdd a, cdd b, cdd c) { return cdd_mul(cdd_mul(a, b), c); }
HD cdd cdd_mul4(cdd a, cdd b, cdd c, cdd d) { return cdd_mul(cdd_mul(cdd_mul(a, b), c), d); }
HD cdd cdd_muladd(cdd a, cdd b, cdd c) { return cdd_add(cdd_mul(a, b), c); }
HD cdd cdd_mulsub(cdd a, cdd b, cdd c) { return cdd_sub(cdd_mul(a, b), c); }
HD cdd cdd_submul(cdd a, cdd b, cdd c) { return cdd_sub(c, cdd_mul(a, b)); }
HD cdd cdd_mulmuladd(cdd a, cdd b, cdd c, cdd d) { return cdd_add(cdd_mul(a, b), cdd_mul(c, d)); }
HD cdd cdd_mulmulsub(cdd a, cdd b, cdd c, cdd d) { return cdd_sub(cdd_mul(a, b), cdd_mul(c, d)); }
HD cdd cdd_inv_not_zero(cdd a) {
    return (a.re.hi == 0.0 && a.re.lo == 0.0 && a.im.hi == 0.0 && a.im.lo == 0.0) ? a : cdd_inv(a);
}
HD cdd cdd_invsqr(cdd a) { return cdd_sqr(cdd_inv(a)); }
__device__ __noinline__ cdd cdd_pbs(cdd x, int p) {
    if (p == 1) return x;
    if (p == 0) return mkcdd(mkdd(1.0, 0.0), mkdd(0.0, 0.0));
    if (p == 2) return cdd_mul(x, x);
    int t = __builtin_ctz((unsigned)p) + 1;
    p >>= t;
    while (--t > 0) x = cdd_mul(x, x);
    cdd y = x;
    while (p > 0) {
        t = __builtin_ctz((unsigned)p) + 1;
        p >>= t;
        while (--t >= 0) x = cdd_mul(x, x);
        y = cdd_mul(y, x);
    }
    return y;
}
HD cdd cdd_pow_int(cdd x, int p) { return p > 0 ? cdd_pbs(x, p) : cdd_inv(cdd_pbs(x, -p)); }

// ---------------------- truncated Taylor series, orders 0..K ----------------------
template <int K> struct ser { cplx c[K + 1]; };

template <int K> HD ser<K> s_const(cplx a) { ser<K> r; r.c[0] = a;
#pragma unroll
    for (int k = 1; k <= K; k++) r.c[k] = mkc(0.0, 0.0);
    return r; }
template <int K> HD ser<K> s_add(const ser<K>& x, const ser<K>& y) { ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) r.c[k] = c_add(x.c[k], y.c[k]);
    return r; }
template <int K> HD ser<K> s_sub(const ser<K>& x, const ser<K>& y) { ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) r.c[k] = c_sub(x.c[k], y.c[k]);
    return r; }
template <int K> HD ser<K> s_neg(const ser<K>& x) { ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) r.c[k] = c_neg(x.c[k]);
    return r; }
template <int K> HD ser<K> s_mul(const ser<K>& x, const ser<K>& y) {
    ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) {
        cplx c = c_mul(x.c[0], y.c[k]);
#pragma unroll
        for (int j = 1; j <= k; j++) c = c_muladd(x.c[j], y.c[k - j], c);
        r.c[k] = c;
    }
    return r;
}
template <int K> HD ser<K> s_sqr(const ser<K>& x) {
    ser<K> r;
    r.c[0] = c_sqr(x.c[0]);
#pragma unroll
    for (int k = 1; k <= K; k++) {
        cplx w = c_mul(x.c[0], x.c[k]);
#pragma unroll
        for (int j = 1; j <= (k - 1) / 2; j++) w = c_muladd(x.c[j], x.c[k - j], w);
        if (k % 2 == 0) r.c[k] = c_add(c_scale(2.0, w), c_sqr(x.c[k / 2]));
        else r.c[k] = c_add(w, w);
    }
    return r;
}
template <int K> HD ser<K> s_cb(const ser<K>& x) { return s_mul<K>(s_sqr<K>(x), x); }
template <int K> HD ser<K> s_div(const ser<K>& x, const ser<K>& y) {
    ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) {
        cplx num = x.c[k];
        if (k > 0) {
            cplx s = c_mul(r.c[0], y.c[k]);
#pragma unroll
            for (int j = 1; j < k; j++) s = c_muladd(r.c[j], y.c[k - j], s);
            num = c_sub(num, s);
        }
        r.c[k] = c_div(num, y.c[0]);
    }
    return r;
}
template <int K> HD ser<K> s_inv(const ser<K>& x) {
    ser<K> r;
    r.c[0] = c_div(mkc(1.0, 0.0), x.c[0]);
#pragma unroll
    for (int k = 1; k <= K; k++) {
        cplx s = c_mul(r.c[0], x.c[k]);
#pragma unroll
        for (int j = 1; j < k; j++) s = c_muladd(r.c[j], x.c[k - j], s);
        r.c[k] = c_div(c_neg(s), x.c[0]);
    }
    return r;
}
template <int K> HD ser<K> s_invsqr(const ser<K>& x) { return s_inv<K>(s_sqr<K>(x)); }
// taylor_op_pow_int with the reference's x0 == 0 -> 0 rule
template <int K> __device__ __noinline__ ser<K> s_pow_int(ser<K> x, int rr) {
    ser<K> w;
    if (x.c[0].re == 0.0 && x.c[0].im == 0.0) {
#pragma unroll
        for (int k = 0; k <= K; k++) w.c[k] = mkc(0.0, 0.0);
        return w;
    }
    w.c[0] = c_pow_int(x.c[0], rr);
    cplx u0inv = c_inv(x.c[0]);
#pragma unroll
    for (int k = 1; k <= K; k++) {
        cplx s = c_mul(w.c[k - 1], c_scale(1.0, x.c[1]));
#pragma unroll
        for (int j = 2; j <= k; j++) s = c_muladd(w.c[k - j], c_scale((double)j, x.c[j]), s);
        s = c_scale((double)rr, s);
        cplx wt;
        if (k > 1) {
            cplx t = c_mul(x.c[k - 1], c_scale(1.0, w.c[1]));
#pragma unroll
            for (int j = 2; j <= k - 1; j++) t = c_muladd(x.c[k - j], c_scale((double)j, w.c[j]), t);
            wt = c_mul(u0inv, c_sub(s, t));
        } else {
            wt = c_mul(u0inv, s);
        }
        w.c[k] = c_divr(wt, (double)k);
    }
    return w;
}
template <int K> HD ser<K> s_muladd(const ser<K>& a, const ser<K>& y, const ser<K>& z) {
    ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) {
        cplx c = z.c[k];
#pragma unroll
        for (int j = 0; j <= k; j++) c = c_muladd(a.c[j], y.c[k - j], c);
        r.c[k] = c;
    }
    return r;
}
template <int K> HD ser<K> s_mulsub(const ser<K>& a, const ser<K>& y, const ser<K>& z) {
    ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) {
        cplx c = c_mulsub(a.c[0], y.c[k], z.c[k]);
#pragma unroll
        for (int j = 1; j <= k; j++) c = c_muladd(a.c[j], y.c[k - j], c);
        r.c[k] = c;
    }
    return r;
}
template <int K> HD ser<K> s_submul(const ser<K>& a, const ser<K>& y, const ser<K>& z) {
    ser<K> r;
#pragma unroll
    for (int k = 0; k <= K; k++) {
        cplx c = z.c[k];
#pragma unroll
        for (int j = 0; j <= k; j++) c = c_submul(a.c[j], y.c[k - j], c);
        r.c[k] = c;
    }
    return r;
}
template <int K> HD ser<K> s_add3(const ser<K>& a, const ser<K>& b, const ser<K>& c) { return s_add<K>(s_add<K>(a, b), c); }
template <int K> HD ser<K> s_mul3(const ser<K>& a, const ser<K>& b, const ser<K>& c) { return s_mul<K>(s_mul<K>(a, b), c); }
template <int K> HD ser<K> s_add4(const ser<K>& a, const ser<K>& b, const ser<K>& c, const ser<K>& d) { return s_add<K>(s_add<K>(a, b), s_add<K>(c, d)); }
template <int K> HD ser<K> s_mul4(const ser<K>& a, const ser<K>& b, const ser<K>& c, const ser<K>& d) { return s_mul<K>(s_mul<K>(a, b), s_mul<K>(c, d)); }
template <int K> HD ser<K> s_mulmuladd(const ser<K>& a, const ser<K>& b, const ser<K>& c, const ser<K>& d) { return s_add<K>(s_mul<K>(a, b), s_mul<K>(c, d)); }
template <int K> HD ser<K> s_mulmulsub(const ser<K>& a, const ser<K>& b, const ser<K>& c, const ser<K>& d) { return s_sub<K>(s_mul<K>(a, b), s_mul<K>(c, d)); }


/* Deterministic log / exp / pow / cbrt built from IEEE-exact operations only
 * (+ - * / sqrt, frexp/ldexp), so that the CPU oracle and the GPU kernels produce
 * identical bits; accuracy is about 1 ulp.  log: atanh series on m in [sqrt(1/2), sqrt(2));
 * exp: k*ln2 range reduction + degree-16 Taylor; pow(x,y) = exp(y log x) for x > 0;
 * cbrt: exp(log|x|/3) plus one Newton correction. */
#define HC_LN2_HI 6.93147180369123816490e-01
#define HC_LN2_LO 1.90821492927058770002e-10
// The *_inl forms are inlined where the caller keeps much state in registers (a call would
// save and restore it); the plain names are out-of-line calls of the same code.
HD double hc_log_inl(double x) {
    if (isnan(x) || x < 0) return HCG_NAN;
    if (x == 0) return -HCG_INF;
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
HD double hc_exp_inl(double x) {
    if (isnan(x)) return x;
    if (x > 709.782712893384) return HCG_INF;
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
HD double hc_pow_inl(double x, double y) {
    if (y == 0) return 1.0;
    if (x == 1.0) return 1.0;
    if (isnan(x) || isnan(y)) return HCG_NAN;
    if (x == 0) return y > 0 ? 0.0 : HCG_INF;
    if (isinf(x)) return y > 0 ? HCG_INF : 0.0;
    if (x < 0) return HCG_NAN;
    return hc_exp_inl(y * hc_log_inl(x));
}
HD double hc_cbrt_inl(double x) {
    if (x == 0 || isnan(x) || isinf(x)) return x;
    double a = fabs(x);
    double y = hc_exp_inl(hc_log_inl(a) / 3.0);
    y = y - (y * y * y - a) / (3.0 * (y * y));
    return x < 0 ? -y : y;
}
__device__ __noinline__ double hc_log(double x) { return hc_log_inl(x); }
__device__ __noinline__ double hc_exp(double x) { return hc_exp_inl(x); }
__device__ __noinline__ double hc_pow(double x, double y) { return hc_pow_inl(x, y); }
__device__ __noinline__ double hc_cbrt(double x) { return hc_cbrt_inl(x); }

// Julia's NaN-propagating min/max/clamp and helpers (src/utils.jl:207-212, 408-422)
HD double jmin(double a, double b) { if (isnan(a) || isnan(b)) return HCG_NAN; return a < b ? a : b; }
HD double jmax(double a, double b) { if (isnan(a) || isnan(b)) return HCG_NAN; return a > b ? a : b; }
HD double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }
HD double nanmin(double a, double b) { return isnan(a) ? b : (isnan(b) ? a : jmin(a, b)); }
HD double nanmax(double a, double b) { return isnan(a) ? b : (isnan(b) ? a : jmax(a, b)); }
HD double max_fast(double x, double y) { return y > x ? y : x; }
HD double jeps(double x) {
    if (!isfinite(x)) return HCG_NAN;
    x = fabs(x);
    if (x >= 2.2250738585072014e-308) { int e; frexp(x, &e); return ldexp(1.0, e - 1 - 52); }
    return 4.9406564584124654e-324;
}
HD double nthroot(double x, int N) {
    if (N == 4) return sqrt(sqrt(x));
    if (N == 2) return sqrt(x);
    if (N == 3) return hc_cbrt(x);
    if (N == 1) return x;
    if (N == 0) return 1.0;
    return hc_pow(x, 1.0 / N);
}
HD double nthroot_inl(double x, int N) {
    if (N == 4) return sqrt(sqrt(x));
    if (N == 2) return sqrt(x);
    if (N == 3) return hc_cbrt_inl(x);
    if (N == 1) return x;
    if (N == 0) return 1.0;
    return hc_pow_inl(x, 1.0 / N);
}
HD double ipow(double x, int n) {
    double r = 1.0, b = x; int k = n < 0 ? -n : n;
    while (k) { if (k & 1) r *= b; b *= b; k >>= 1; }
    return n < 0 ? 1.0 / r : r;
}
HD double h_of(double a) { return 2 * a * (sqrt(4 * (a * a) + 1) - 2 * a); }
#define DEPS 2.220446049250313e-16

// Interpreter for double-double tape evaluation (operand >= 0: register, < 0: constant -k-1)
// Parameters of a parameter homotopy at t: p(t) = t p + (1 - t) q, and the direction
// p - q of its Taylor series (tp!, src/homotopies/parameter_homotopy.jl:66-87).
// Both forms are computed and one is selected (no branch): inlined for every parameter of
// a straight-line body, a branch would split it into hundreds of blocks.
// A copy of v that the compiler cannot identify with v: the generated code passes the
// operands of a deliberately repeated computation through it, so that the repetition is not
// merged back into the first computation (which would keep its result live in between).
HD cplx hcg_opaque(cplx v) {
    asm volatile("" : "+v"(v.re), "+v"(v.im));
    return v;
}
HD cplx par_lerp(cplx t, const double* p, const double* q, int k) {
    cplx pk = mkc(p[2 * k], p[2 * k + 1]), qk = mkc(q[2 * k], q[2 * k + 1]);
    cplx rr = c_add(c_scale(t.re, pk), c_scale(1.0 - t.re, qk));
    cplx rc = c_add(c_mul(t, pk), c_mul(mkc(1.0 - t.re, -t.im), qk));
    const bool real_t = t.im == 0.0;
    return mkc(real_t ? rr.re : rc.re, real_t ? rr.im : rc.im);
}
HD cplx par_diff(const double* p, const double* q, int k) {
    return c_sub(mkc(p[2 * k], p[2 * k + 1]), mkc(q[2 * k], q[2 * k + 1]));
}
// Operands of the double-double interpreter: >= 0 register, < 0 constant -e-1,
// >= DD_PBASE homotopy parameter e - DD_PBASE (formed at t when read).
#define DD_PBASE (1 << 24)
struct DDPar {
    const double* p;
    const double* q;
    cplx t;
};
HD cdd dd_fetch(const double* K, const cdd* R, const DDPar& P, int e) {
    if (e >= DD_PBASE) return cdd_from(par_lerp(P.t, P.p, P.q, e - DD_PBASE));
    return e >= 0 ? R[e] : cdd_from(mkc(K[2 * (-e - 1)], K[2 * (-e - 1) + 1]));
}
__device__ __noinline__ void dd_interpret(const int* code, int ninstr, const double* K, cdd* R, DDPar P) {
    for (int i = 0; i < ninstr; i++) {
        const int* I = code + 6 * i;
        const int op = I[0];
        cdd a = dd_fetch(K, R, P, I[1]), r;
        switch (op) {
        case 1: r = cdd_cb(a); break;                       // OP_CB
        case 7: r = cdd_inv(a); break;                      // OP_INV
        case 8: r = cdd_inv_not_zero(a); break;             // OP_INV_NOT_ZERO
        case 9: r = cdd_invsqr(a); break;                   // OP_INVSQR
        case 10: r = cdd_neg(a); break;                     // OP_NEG
        case 13: r = cdd_sqr(a); break;                     // OP_SQR
        case 17: r = a; break;                              // OP_IDENTITY
        case 18: r = cdd_add(a, dd_fetch(K, R, P, I[2])); break;
        case 19: r = cdd_div(a, dd_fetch(K, R, P, I[2])); break;
        case 20: r = cdd_mul(a, dd_fetch(K, R, P, I[2])); break;
        case 21: r = cdd_sub(a, dd_fetch(K, R, P, I[2])); break;
        case 22: r = cdd_pow_int(a, I[2]); break;           // OP_POW_INT (exponent inline)
        case 24: r = cdd_add3(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3])); break;
        case 25: r = cdd_mul3(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3])); break;
        case 26: r = cdd_muladd(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3])); break;
        case 27: r = cdd_mulsub(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3])); break;
        case 28: r = cdd_submul(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3])); break;
        case 29: r = cdd_add4(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3]), dd_fetch(K, R, P, I[4])); break;
        case 30: r = cdd_mul4(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3]), dd_fetch(K, R, P, I[4])); break;
        case 31: r = cdd_mulmuladd(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3]), dd_fetch(K, R, P, I[4])); break;
        case 32: r = cdd_mulmulsub(a, dd_fetch(K, R, P, I[2]), dd_fetch(K, R, P, I[3]), dd_fetch(K, R, P, I[4])); break;
        default: r = mkcdd(mkdd(HCG_NAN, 0.0), mkdd(HCG_NAN, 0.0));
        }
        R[I[5]] = r;
    }
}
